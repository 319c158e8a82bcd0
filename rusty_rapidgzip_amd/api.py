"""Host-side mirror of rusty_rapidgzip's public API over the MI355X decoder.

Reference surface (crates/rusty-rapidgzip/src/lib.rs):
  read_gz(path, sink, config) -> DecodeStats          lib.rs:366-496
  Config{num_threads, chunk_size_bytes, verbose}      lib.rs:49-100
  Verbosity::{Off, On}                                lib.rs:72-88
  DecodeStats{compressed_bytes, uncompressed_bytes,
              chunks_decoded, speculation_failures}   lib.rs:102-108
  Error::{Io, Gzip(GzipError), Deflate(DeflateError)} lib.rs:110-118

The sink receives the decompressed stream as in-order ``bytes`` chunks, one per
reference chunk (the reference's ``Arc<Vec<u8>>`` per ``Sender::send``).  It may
be any object with ``send`` or ``put`` (crossbeam / ``queue.Queue`` style), a
list (appended to) or a callable.  A sink that raises :class:`SinkClosed` or
returns ``False`` from ``send`` ends the stream like a dropped Receiver.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass, field

from . import _native

# ---------------------------------------------------------------------------
# ctypes mirror of include/rrg_gpu.h
# ---------------------------------------------------------------------------
RRG_OK = 0
RRG_E_EOF, RRG_E_INVALID, RRG_E_BADMAGIC, RRG_E_METHOD, RRG_E_RESERVED = 1, 2, 3, 4, 5
RRG_E_TRUNCATED, RRG_E_CRC, RRG_E_ISIZE, RRG_E_IO, RRG_E_DEVICE, RRG_E_ARG = 6, 7, 8, 9, 10, 11


class _Err(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("expected", ctypes.c_uint32),
        ("got", ctypes.c_uint32),
        ("got64", ctypes.c_uint64),
        ("member", ctypes.c_uint32),
        ("trailer_byte", ctypes.c_uint64),
        ("method", ctypes.c_uint32),
        ("message", ctypes.c_char * 192),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [
        ("compressed_bytes", ctypes.c_uint64),
        ("uncompressed_bytes", ctypes.c_uint64),
        ("chunks_decoded", ctypes.c_uint64),
        ("speculation_failures", ctypes.c_uint64),
    ]


class _Cfg(ctypes.Structure):
    _fields_ = [
        ("chunk_size_bytes", ctypes.c_size_t),
        ("num_threads", ctypes.c_int),
        ("verbose", ctypes.c_int),
        ("segment_bytes", ctypes.c_size_t),
        ("lut_bits", ctypes.c_int),
        ("lanes_per_block", ctypes.c_int),
    ]


class Profile(ctypes.Structure):
    """Stage timings of the last decode on a context (rrg_profile)."""

    _fields_ = [
        ("ms_total", ctypes.c_double),
        ("ms_find", ctypes.c_double),
        ("ms_inflate", ctypes.c_double),
        ("ms_tail", ctypes.c_double),
        ("ms_resolve", ctypes.c_double),
        ("ms_crc", ctypes.c_double),
        ("n_segments", ctypes.c_uint64),
        ("n_markers_tail", ctypes.c_uint64),
        ("n_inflate_launches", ctypes.c_uint64),
        ("n_tail_rounds", ctypes.c_uint64),
        ("compressed_bytes", ctypes.c_uint64),
        ("uncompressed_bytes", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t)

# every symbol include/rrg_gpu.h declares
EXPORTED_SYMBOLS = (
    "rrg_gpu_abi_version",
    "rrg_cfg_default",
    "rrg_gpu_create",
    "rrg_gpu_destroy",
    "rrg_gpu_read_gz",
    "rrg_gpu_decode",
    "rrg_gpu_decode_device",
    "rrg_gpu_last_profile",
)


def _lib() -> ctypes.CDLL:
    lib = _native.gpu_lib()
    if not getattr(lib, "_rrg_typed", False):
        lib.rrg_gpu_abi_version.restype = ctypes.c_int
        lib.rrg_cfg_default.argtypes = [ctypes.POINTER(_Cfg)]
        lib.rrg_gpu_create.restype = ctypes.c_void_p
        lib.rrg_gpu_create.argtypes = [ctypes.c_int]
        lib.rrg_gpu_destroy.argtypes = [ctypes.c_void_p]
        lib.rrg_gpu_read_gz.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(_Cfg), _CB, ctypes.c_void_p,
                                        ctypes.POINTER(_Stats), ctypes.POINTER(_Err)]
        lib.rrg_gpu_decode.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(_Cfg), _CB,
                                       ctypes.c_void_p, ctypes.POINTER(_Stats), ctypes.POINTER(_Err)]
        lib.rrg_gpu_decode_device.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(_Cfg),
            ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
            ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_Stats), ctypes.POINTER(_Err)]
        lib.rrg_gpu_last_profile.argtypes = [ctypes.c_void_p, ctypes.POINTER(Profile)]
        lib._rrg_typed = True
    return lib


# ---------------------------------------------------------------------------
# Public types
# ---------------------------------------------------------------------------
class Verbosity(enum.Enum):
    Off = 0
    On = 1


@dataclass
class Config:
    """lib.rs:49-100.  ``segment_bytes``/``lut_bits``/``lanes_per_block`` tune the GPU decomposition."""

    num_threads: int = 0
    chunk_size_bytes: int = 4 * 1024 * 1024
    verbose: Verbosity = Verbosity.Off
    segment_bytes: int = 0
    lut_bits: int = 9
    lanes_per_block: int = 32

    def _c(self) -> _Cfg:
        c = _Cfg()
        _lib().rrg_cfg_default(ctypes.byref(c))
        c.chunk_size_bytes = int(self.chunk_size_bytes)
        c.num_threads = int(self.num_threads)
        c.verbose = 1 if self.verbose == Verbosity.On else 0
        c.segment_bytes = int(self.segment_bytes)
        c.lut_bits = int(self.lut_bits)
        c.lanes_per_block = int(self.lanes_per_block)
        return c


@dataclass
class DecodeStats:
    """lib.rs:102-108"""

    compressed_bytes: int = 0
    uncompressed_bytes: int = 0
    chunks_decoded: int = 0
    speculation_failures: int = 0


class RapidgzipError(Exception):
    """Error (lib.rs:110-118)."""


class IoError(RapidgzipError):
    """Error::Io"""


class DeviceError(RapidgzipError):
    """HIP runtime failure (no reference counterpart)."""


class DeflateError(RapidgzipError):
    """DeflateError::{UnexpectedEof, Invalid(msg)} (deflate/mod.rs:22-28)."""

    def __init__(self, kind: str, message: str):
        super().__init__(message)
        self.kind = kind  # "UnexpectedEof" | "Invalid"


@dataclass
class _GzipFields:
    expected: int = 0
    got: int = 0
    member: int = 0
    uncompressed: int = 0
    trailer_byte: int = 0


class GzipError(RapidgzipError):
    """GzipError::{BadMagic, UnsupportedMethod, ReservedFlags, Truncated, CrcMismatch, IsizeMismatch} (gzip.rs:25-61)."""

    def __init__(self, kind: str, message: str, **fields):
        super().__init__(message)
        self.kind = kind
        for k, v in fields.items():
            setattr(self, k, v)


class SinkClosed(Exception):
    """Raised by a sink to end the stream (the reference's dropped Receiver)."""


def _raise(rc: int, e: _Err) -> None:
    msg = e.message.decode(errors="replace")
    if rc == RRG_E_EOF:
        raise DeflateError("UnexpectedEof", msg)
    if rc == RRG_E_INVALID:
        raise DeflateError("Invalid", msg)
    if rc == RRG_E_BADMAGIC:
        raise GzipError("BadMagic", msg)
    if rc == RRG_E_METHOD:
        raise GzipError("UnsupportedMethod", msg, method=e.method)
    if rc == RRG_E_RESERVED:
        raise GzipError("ReservedFlags", msg, flags=e.method)
    if rc == RRG_E_TRUNCATED:
        raise GzipError("Truncated", msg)
    if rc == RRG_E_CRC:
        raise GzipError("CrcMismatch", msg, expected=e.expected, got=e.got, member=e.member,
                        uncompressed=e.got64, trailer_byte=e.trailer_byte)
    if rc == RRG_E_ISIZE:
        raise GzipError("IsizeMismatch", msg, expected=e.expected, got=e.got64, member=e.member,
                        trailer_byte=e.trailer_byte)
    if rc == RRG_E_IO:
        raise IoError(msg)
    raise DeviceError(msg or f"rrg_gpu error {rc}")


def _sink_fn(sink):
    if sink is None:
        return lambda b: True
    if isinstance(sink, list):
        return lambda b: (sink.append(b), True)[1]
    for name in ("send", "put"):
        fn = getattr(sink, name, None)
        if fn is not None:
            return lambda b, fn=fn: fn(b) is not False
    if callable(sink):
        return lambda b: sink(b) is not False
    raise TypeError("sink must have send()/put(), be a list, or be callable")


class Decoder:
    """One decoder context on one GPU (one HIP stream, its own device buffers).

    Not thread-safe, like the reference's per-worker native handles
    (libdeflate_ffi.rs:28-35)."""

    def __init__(self, device: int = 0):
        self._lib = _lib()
        self._ctx = self._lib.rrg_gpu_create(int(device))
        if not self._ctx:
            raise DeviceError(f"cannot create a decoder context on HIP device {device}")
        self.device = device

    def close(self) -> None:
        if self._ctx:
            self._lib.rrg_gpu_destroy(self._ctx)
            self._ctx = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def profile(self) -> Profile:
        p = Profile()
        self._lib.rrg_gpu_last_profile(self._ctx, ctypes.byref(p))
        return p

    def _run(self, fn, config: Config | None, sink) -> DecodeStats:
        config = config or Config()
        push = _sink_fn(sink)
        pending: list[BaseException] = []

        def cb(_user, ptr, n):
            try:
                ok = push(ctypes.string_at(ptr, n))
            except SinkClosed:
                return 1
            except BaseException as ex:  # propagate after the native call returns
                pending.append(ex)
                return 1
            return 0 if ok else 1

        ccb = _CB(cb)
        st, er = _Stats(), _Err()
        rc = fn(ctypes.byref(config._c()), ccb, ctypes.byref(st), ctypes.byref(er))
        if pending:
            raise pending[0]
        if rc != RRG_OK:
            _raise(rc, er)
        return DecodeStats(st.compressed_bytes, st.uncompressed_bytes, st.chunks_decoded, st.speculation_failures)

    def read_gz(self, path, sink, config: Config | None = None) -> DecodeStats:
        p = str(path).encode()
        return self._run(lambda c, cb, st, er: self._lib.rrg_gpu_read_gz(self._ctx, p, c, cb, None, st, er),
                         config, sink)

    def decode_bytes(self, data: bytes, sink, config: Config | None = None) -> DecodeStats:
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data) if data else (ctypes.c_uint8 * 1)()
        return self._run(
            lambda c, cb, st, er: self._lib.rrg_gpu_decode(self._ctx, ctypes.addressof(buf), len(data), c, cb, None,
                                                           st, er), config, sink)

    def decode_device(self, d_ptr: int, length: int, host: bytes | None = None, config: Config | None = None,
                      max_chunks: int = 1 << 20):
        """Device-resident decode.  Returns (d_out_ptr, out_len, chunk_ends, stats); the output
        stays in device memory owned by this decoder until its next call."""
        config = config or Config()
        ends = (ctypes.c_uint64 * max_chunks)()
        d_out = ctypes.c_void_p()
        out_len = ctypes.c_uint64()
        n = ctypes.c_size_t()
        st, er = _Stats(), _Err()
        # host copy of the file (framing walks only), passed without copying
        hptr = None
        if host is not None:
            if isinstance(host, bytes):
                hptr = ctypes.cast(ctypes.c_char_p(host), ctypes.c_void_p).value
            else:
                import numpy as np

                arr = np.ascontiguousarray(host, dtype=np.uint8)
                host = arr  # keep alive for the call
                hptr = arr.ctypes.data
        rc = self._lib.rrg_gpu_decode_device(self._ctx, ctypes.c_void_p(d_ptr), length, hptr,
                                             ctypes.byref(config._c()), ctypes.byref(d_out), ctypes.byref(out_len),
                                             ends, max_chunks, ctypes.byref(n), ctypes.byref(st), ctypes.byref(er))
        if rc != RRG_OK:
            _raise(rc, er)
        stats = DecodeStats(st.compressed_bytes, st.uncompressed_bytes, st.chunks_decoded, st.speculation_failures)
        return d_out.value or 0, out_len.value, [ends[i] for i in range(min(n.value, max_chunks))], stats


_default: dict[int, Decoder] = {}


def _decoder(device: int = 0) -> Decoder:
    d = _default.get(device)
    if d is None:
        d = _default[device] = Decoder(device)
    return d


def read_gz(path, sink, config: Config | None = None) -> DecodeStats:
    """Decode the gzip file at ``path`` into ``sink`` (lib.rs:366-496).  Blocks until done."""
    return _decoder().read_gz(path, sink, config)


def decode_bytes(data: bytes, sink, config: Config | None = None) -> DecodeStats:
    """read_gz over an in-memory gzip file."""
    return _decoder().decode_bytes(data, sink, config)


def decode_all(data: bytes, config: Config | None = None) -> bytes:
    """All members of an in-memory gzip file, concatenated (gzip.rs:63-85 decode_all)."""
    out: list[bytes] = []
    decode_bytes(data, out, config)
    return b"".join(out)
