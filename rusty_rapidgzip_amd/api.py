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
        ("device_budget_bytes", ctypes.c_size_t),
        ("launch_graph", ctypes.c_int),
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
        ("n_redecoded", ctypes.c_uint64),
        ("n_inflate_launches", ctypes.c_uint64),
        ("n_tail_rounds", ctypes.c_uint64),
        ("compressed_bytes", ctypes.c_uint64),
        ("uncompressed_bytes", ctypes.c_uint64),
        ("n_grown", ctypes.c_uint64),
        ("ms_expand", ctypes.c_double),
        ("n_tokens", ctypes.c_uint64),
        ("find_bits", ctypes.c_uint64),
        ("ms_first_chunk", ctypes.c_double),
        ("n_slabs", ctypes.c_uint64),
        ("device_bytes_peak", ctypes.c_uint64),
        ("n_graph_launches", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class _BatchItem(ctypes.Structure):
    _fields_ = [
        ("d_file", ctypes.c_void_p),
        ("h_file", ctypes.c_void_p),
        ("len", ctypes.c_size_t),
        ("out_off", ctypes.c_uint64),
        ("out_len", ctypes.c_uint64),
        ("chunk_ends", ctypes.POINTER(ctypes.c_uint64)),
        ("n_chunks", ctypes.c_size_t),
        ("stats", _Stats),
        ("error", _Err),
    ]


class _ServerOpts(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_size_t),
        ("max_batch_bytes", ctypes.c_size_t),
        ("max_batch_items", ctypes.c_int),
        ("cfg", _Cfg),
        ("engines", ctypes.c_int),
    ]


class _Completion(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("tag", ctypes.c_uint64),
        ("d_out", ctypes.c_void_p),
        ("out_len", ctypes.c_uint64),
        ("n_chunks", ctypes.c_uint64),
        ("batch", ctypes.c_uint64),
        ("batch_items", ctypes.c_uint32),
        ("ms_queue", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("stats", _Stats),
        ("error", _Err),
    ]


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
    "rrg_gpu_decode_fd",
    "rrg_gpu_autotune",
    "rrg_gpu_last_profile",
    "rrg_gpu_decode_batch",
    "rrg_server_create",
    "rrg_server_submit",
    "rrg_server_poll",
    "rrg_server_chunk_ends",
    "rrg_server_copy_out",
    "rrg_server_release",
    "rrg_server_pending",
    "rrg_server_destroy",
    "rrg_gpu_debug_fail_alloc",
    "rrg_gpu_decode_device_noalloc",
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
        lib.rrg_gpu_autotune.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(_Cfg), ctypes.c_int,
            ctypes.POINTER(_Cfg), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_Err)]
        lib.rrg_gpu_decode_device.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(_Cfg),
            ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
            ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_Stats), ctypes.POINTER(_Err)]
        lib.rrg_gpu_decode_device_noalloc.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(_Cfg), ctypes.c_void_p,
            ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t,
            ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_Stats), ctypes.POINTER(_Err)]
        lib.rrg_gpu_last_profile.argtypes = [ctypes.c_void_p, ctypes.POINTER(Profile)]
        lib.rrg_gpu_decode_fd.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(_Cfg), _CB, ctypes.c_void_p,
                                          ctypes.POINTER(_Stats), ctypes.POINTER(_Err)]
        lib.rrg_gpu_decode_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(_BatchItem), ctypes.c_size_t,
                                             ctypes.POINTER(_Cfg), ctypes.POINTER(ctypes.c_void_p),
                                             ctypes.POINTER(_Err)]
        lib.rrg_server_create.restype = ctypes.c_void_p
        lib.rrg_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(_ServerOpts)]
        lib.rrg_server_submit.restype = ctypes.c_uint64
        lib.rrg_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64]
        lib.rrg_server_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Completion), ctypes.c_int, ctypes.c_int]
        lib.rrg_server_chunk_ends.restype = ctypes.c_size_t
        lib.rrg_server_chunk_ends.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.c_size_t]
        lib.rrg_server_copy_out.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t]
        lib.rrg_server_release.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        lib.rrg_server_pending.restype = ctypes.c_uint64
        lib.rrg_server_pending.argtypes = [ctypes.c_void_p]
        lib.rrg_server_destroy.argtypes = [ctypes.c_void_p]
        lib.rrg_gpu_debug_fail_alloc.argtypes = [ctypes.c_int]
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
    """lib.rs:49-100.  ``num_threads == 1`` gives the reference's single-worker chunking (chunks
    end at block ends, pipeline.rs:1211-1345 / lib.rs:244-361); the GPU decode is the same for
    any value.  ``segment_bytes``/``lut_bits``/``lanes_per_block`` tune the GPU decomposition
    (``lanes_per_block`` 64: one wavefront per segment, the default; 8/16/32: one lane per
    segment; ``lut_bits`` 0 picks the LUT width for either);
    ``device_budget_bytes`` bounds the device memory of a host-output decode (0 = half the
    device's free memory), which then runs in slabs of whole chunks.  ``launch_graph`` runs the
    tail, resolve and CRC kernels of a batched or served decode as one HIP graph launch."""

    num_threads: int = 0
    chunk_size_bytes: int = 4 * 1024 * 1024
    verbose: Verbosity = Verbosity.Off
    segment_bytes: int = 0
    lut_bits: int = 0
    lanes_per_block: int = 64
    device_budget_bytes: int = 0
    launch_graph: bool = False

    def _c(self) -> _Cfg:
        c = _Cfg()
        _lib().rrg_cfg_default(ctypes.byref(c))
        c.chunk_size_bytes = int(self.chunk_size_bytes)
        c.num_threads = int(self.num_threads)
        c.verbose = 1 if self.verbose == Verbosity.On else 0
        c.segment_bytes = int(self.segment_bytes)
        c.lut_bits = int(self.lut_bits)
        c.lanes_per_block = int(self.lanes_per_block)
        c.device_budget_bytes = int(self.device_budget_bytes)
        c.launch_graph = 1 if self.launch_graph else 0
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


def _error(rc: int, e: _Err) -> RapidgzipError:
    msg = e.message.decode(errors="replace")
    if rc == RRG_E_EOF:
        return DeflateError("UnexpectedEof", msg)
    if rc == RRG_E_INVALID:
        return DeflateError("Invalid", msg)
    if rc == RRG_E_BADMAGIC:
        return GzipError("BadMagic", msg)
    if rc == RRG_E_METHOD:
        return GzipError("UnsupportedMethod", msg, method=e.method)
    if rc == RRG_E_RESERVED:
        return GzipError("ReservedFlags", msg, flags=e.method)
    if rc == RRG_E_TRUNCATED:
        return GzipError("Truncated", msg)
    if rc == RRG_E_CRC:
        return GzipError("CrcMismatch", msg, expected=e.expected, got=e.got, member=e.member,
                         uncompressed=e.got64, trailer_byte=e.trailer_byte)
    if rc == RRG_E_ISIZE:
        return GzipError("IsizeMismatch", msg, expected=e.expected, got=e.got64, member=e.member,
                         trailer_byte=e.trailer_byte)
    if rc == RRG_E_IO:
        return IoError(msg)
    return DeviceError(msg or f"rrg_gpu error {rc}")


def _raise(rc: int, e: _Err) -> None:
    raise _error(rc, e)


def _host_ptr(data):
    """(address, keep-alive object) of a bytes / bytearray / numpy buffer, without copying bytes objects."""
    if data is None:
        return None, None
    if isinstance(data, bytes):
        return ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p).value, data
    import numpy as np

    arr = np.ascontiguousarray(np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytearray, memoryview))
                               else np.asarray(data, dtype=np.uint8))
    return (arr.ctypes.data if arr.size else None), arr


def _stats(st: _Stats) -> DecodeStats:
    return DecodeStats(st.compressed_bytes, st.uncompressed_bytes, st.chunks_decoded, st.speculation_failures)


@dataclass
class BatchResult:
    """One file of a batched decode: its output slice, chunk ends (relative to
    the slice), stats, and the error it ended with (None on success)."""

    d_out: int
    out_len: int
    chunk_ends: list
    stats: DecodeStats
    error: RapidgzipError | None
    data: bytes | None = None


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
        self._ends_buf = None

    def _ends(self, n: int):
        """The chunk-end array the library fills, kept from call to call (a
        fresh ctypes array of 2^20 entries costs half a millisecond to clear)."""
        if self._ends_buf is None or len(self._ends_buf) < n:
            self._ends_buf = (ctypes.c_uint64 * n)()
        return self._ends_buf

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

    def decode_bytes(self, data, sink, config: Config | None = None) -> DecodeStats:
        """read_gz over a gzip file in host memory (bytes, bytearray or a uint8 array; not copied)."""
        ptr, keep = _host_ptr(data)
        n = len(keep) if keep is not None else 0
        if not ptr:
            keep = (ctypes.c_uint8 * 1)()
            ptr = ctypes.addressof(keep)
        return self._run(
            lambda c, cb, st, er: self._lib.rrg_gpu_decode(self._ctx, ptr, n, c, cb, None, st, er), config, sink)

    def decode_bytes_native(self, data, cb_addr: int, user_addr: int, config: Config | None = None) -> DecodeStats:
        """decode_bytes with a native rrg_chunk_cb (its address and user pointer):
        no Python runs per chunk."""
        config = config or Config()
        ptr, keep = _host_ptr(data)
        n = len(keep) if keep is not None else 0
        if not ptr:
            keep = (ctypes.c_uint8 * 1)()
            ptr = ctypes.addressof(keep)
        st, er = _Stats(), _Err()
        rc = self._lib.rrg_gpu_decode(self._ctx, ptr, n, ctypes.byref(config._c()), _CB(cb_addr),
                                      ctypes.c_void_p(user_addr), ctypes.byref(st), ctypes.byref(er))
        if rc != RRG_OK:
            _raise(rc, er)
        return _stats(st)

    def decode_fd(self, fd: int, sink, config: Config | None = None) -> DecodeStats:
        """read_gz over a non-seekable source (pipe, stdin): the streaming path's chunking."""
        return self._run(lambda c, cb, st, er: self._lib.rrg_gpu_decode_fd(self._ctx, int(fd), c, cb, None, st, er),
                         config, sink)

    def decode_device(self, d_ptr: int, length: int, host: bytes | None = None, config: Config | None = None,
                      max_chunks: int = 1 << 20):
        """Device-resident decode.  Returns (d_out_ptr, out_len, chunk_ends, stats); the output
        stays in device memory owned by this decoder until its next call."""
        config = config or Config()
        ends = self._ends(max_chunks)
        d_out = ctypes.c_void_p()
        out_len = ctypes.c_uint64()
        n = ctypes.c_size_t()
        st, er = _Stats(), _Err()
        # host copy of the file (framing walks only), passed without copying
        hptr, _keep = _host_ptr(host)
        rc = self._lib.rrg_gpu_decode_device(self._ctx, ctypes.c_void_p(d_ptr), length, hptr,
                                             ctypes.byref(config._c()), ctypes.byref(d_out), ctypes.byref(out_len),
                                             ends, max_chunks, ctypes.byref(n), ctypes.byref(st), ctypes.byref(er))
        if rc != RRG_OK:
            _raise(rc, er)
        stats = DecodeStats(st.compressed_bytes, st.uncompressed_bytes, st.chunks_decoded, st.speculation_failures)
        return d_out.value or 0, out_len.value, [ends[i] for i in range(min(n.value, max_chunks))], stats

    def decode_device_noalloc(self, d_ptr: int, length: int, d_out: int, out_cap: int, host: bytes | None = None,
                              config: Config | None = None, max_chunks: int = 1 << 16):
        """Device-resident decode that allocates no device memory
        (rrg_gpu_decode_device_noalloc): the output goes to the caller's
        buffer d_out.  Returns (out_len, chunk_ends, stats)."""
        config = config or Config()
        ends = self._ends(max_chunks)
        out_len = ctypes.c_uint64()
        n = ctypes.c_size_t()
        st, er = _Stats(), _Err()
        hptr, _keep = _host_ptr(host)
        rc = self._lib.rrg_gpu_decode_device_noalloc(self._ctx, ctypes.c_void_p(d_ptr), length, hptr,
                                                     ctypes.byref(config._c()), ctypes.c_void_p(d_out), out_cap,
                                                     ctypes.byref(out_len), ends, max_chunks, ctypes.byref(n),
                                                     ctypes.byref(st), ctypes.byref(er))
        if rc != RRG_OK:
            _raise(rc, er)
        stats = DecodeStats(st.compressed_bytes, st.uncompressed_bytes, st.chunks_decoded, st.speculation_failures)
        return out_len.value, [ends[i] for i in range(min(n.value, max_chunks))], stats

    def autotune(self, d_ptr: int, length: int, host: bytes | None = None, config: Config | None = None,
                 reps: int = 2) -> tuple[Config, float]:
        """Pick lanes_per_block / lut_bits / segment_bytes for this GPU by decoding the
        device-resident sample (a complete gzip file) under each combination
        (rrg_gpu_autotune).  Returns (the fastest Config, its time in ms)."""
        config = config or Config()
        best = _Cfg()
        ms = ctypes.c_double()
        er = _Err()
        hptr, _keep = _host_ptr(host)
        rc = self._lib.rrg_gpu_autotune(self._ctx, ctypes.c_void_p(d_ptr), length, hptr, ctypes.byref(config._c()),
                                        int(reps), ctypes.byref(best), ctypes.byref(ms), ctypes.byref(er))
        if rc != RRG_OK:
            _raise(rc, er)
        tuned = Config(num_threads=best.num_threads, chunk_size_bytes=best.chunk_size_bytes,
                       verbose=config.verbose, segment_bytes=best.segment_bytes, lut_bits=best.lut_bits,
                       lanes_per_block=best.lanes_per_block)
        return tuned, ms.value

    def decode_batch(self, files, config: Config | None = None, device_ptrs=None, device_lens=None,
                     copy: bool = True):
        """Decode independent gzip files with one pass of every kernel.

        ``files``: host buffers (bytes / numpy), or None where ``device_ptrs[i]``
        gives the file in device memory (then ``files[i]`` may still be given as
        the host copy, or left None) with its length in ``device_lens[i]``.
        Returns one :class:`BatchResult` per file;
        with ``copy`` the decoded bytes are copied back into ``result.data``.
        Device outputs stay valid until this decoder's next call."""
        config = config or Config()
        n = len(files)
        items = (_BatchItem * max(1, n))()
        keep = []
        for i, f in enumerate(files):
            hp, k = _host_ptr(f)
            keep.append(k)
            items[i].h_file = hp
            dp = device_ptrs[i] if device_ptrs is not None else None
            items[i].d_file = dp or None
            if dp:
                items[i].len = int(device_lens[i])
            elif f is not None:
                items[i].len = len(f) if not hasattr(f, "nbytes") else int(f.nbytes)
            else:
                raise ValueError(f"batch file {i} has neither host bytes nor a device pointer")
        d_out = ctypes.c_void_p()
        er = _Err()
        rc = self._lib.rrg_gpu_decode_batch(self._ctx, items, n, ctypes.byref(config._c()), ctypes.byref(d_out),
                                            ctypes.byref(er))
        if rc != RRG_OK:
            _raise(rc, er)
        base = d_out.value or 0
        out = []
        if copy:
            from . import hipmem
        for i in range(n):
            it = items[i]
            ends = [it.chunk_ends[k] for k in range(it.n_chunks)]
            r = BatchResult(base + it.out_off, it.out_len, ends, _stats(it.stats),
                            _error(it.error.kind, it.error) if it.error.kind else None)
            if copy:
                r.data = hipmem.to_host(r.d_out, r.out_len).tobytes() if r.out_len else b""
            out.append(r)
        return out


@dataclass
class Completion:
    """A finished server request."""

    id: int
    tag: int
    d_out: int
    out_len: int
    n_chunks: int
    batch: int
    batch_items: int
    ms_queue: float
    ms_total: float
    stats: DecodeStats
    error: RapidgzipError | None


class Server:
    """Serving loop on one GPU (rrg_server_*): requests are queued, admitted
    while the device bytes they hold stay under ``hbm_budget_bytes``, and
    decoded in batches formed from whatever is queued, ``engines`` batches at
    a time (each on a decode context and stream of its own; 0 = 2).  Outputs
    stay in device memory until :meth:`release`."""

    def __init__(self, device: int = 0, config: Config | None = None, hbm_budget_bytes: int = 0,
                 max_batch_bytes: int = 0, max_batch_items: int = 0, engines: int = 0):
        self._lib = _lib()
        o = _ServerOpts()
        o.hbm_budget_bytes = int(hbm_budget_bytes)
        o.max_batch_bytes = int(max_batch_bytes)
        o.max_batch_items = int(max_batch_items)
        o.engines = int(engines)
        o.cfg = (config or Config())._c()
        self._srv = self._lib.rrg_server_create(int(device), ctypes.byref(o))
        if not self._srv:
            raise DeviceError(f"cannot start a server on HIP device {device}")

    def submit(self, data, tag: int = 0) -> int:
        hp, _k = _host_ptr(data)
        n = len(data) if not hasattr(data, "nbytes") else int(data.nbytes)
        rid = self._lib.rrg_server_submit(self._srv, hp, n, int(tag))
        if not rid:
            raise IoError("server is shutting down")
        return rid

    def poll(self, max_items: int = 64, timeout_ms: int = -1) -> list:
        buf = (_Completion * max_items)()
        k = self._lib.rrg_server_poll(self._srv, buf, max_items, int(timeout_ms))
        out = []
        for i in range(k):
            c = buf[i]
            out.append(Completion(c.id, c.tag, c.d_out or 0, c.out_len, c.n_chunks, c.batch, c.batch_items,
                                  c.ms_queue, c.ms_total, _stats(c.stats),
                                  _error(c.error.kind, c.error) if c.error.kind else None))
        return out

    def chunk_ends(self, rid: int) -> list:
        n = self._lib.rrg_server_chunk_ends(self._srv, rid, None, 0)
        arr = (ctypes.c_uint64 * max(1, n))()
        self._lib.rrg_server_chunk_ends(self._srv, rid, arr, n)
        return [arr[i] for i in range(n)]

    def result(self, rid: int, out_len: int) -> bytes:
        buf = (ctypes.c_uint8 * max(1, out_len))()
        rc = self._lib.rrg_server_copy_out(self._srv, rid, ctypes.addressof(buf), out_len)
        if rc != RRG_OK:
            raise DeviceError(f"copy of request {rid} failed ({rc})")
        return bytes(buf)[:out_len]

    def release(self, rid: int) -> None:
        self._lib.rrg_server_release(self._srv, rid)

    def pending(self) -> int:
        return self._lib.rrg_server_pending(self._srv)

    def close(self) -> None:
        if self._srv:
            self._lib.rrg_server_destroy(self._srv)
            self._srv = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def debug_fail_alloc(n: int) -> None:
    """Testing hook: the n-th next device allocation of the library fails with
    HIP's out-of-memory error (n <= 0 turns it off)."""
    _lib().rrg_gpu_debug_fail_alloc(int(n))


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
