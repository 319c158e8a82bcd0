"""ctypes wrapper of the CPU oracle (rrg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker / the timed CPU baseline,
never as a product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_rrg.so")

ORC_OK, ORC_E_EOF, ORC_E_INVALID, ORC_E_BADMAGIC, ORC_E_METHOD, ORC_E_RESERVED = 0, 1, 2, 3, 4, 5
ORC_E_TRUNCATED, ORC_E_CRC, ORC_E_ISIZE, ORC_E_IO, ORC_E_NOMEM = 6, 7, 8, 9, 10


def build() -> str:
    """Compile the oracle (see oracle/Makefile for the same recipe)."""
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-o", tmp, os.path.join(HERE, "rrg_oracle.c"), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(r.stderr)
    os.replace(tmp, LIB)
    return LIB


class Error(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("expected", ctypes.c_uint32), ("got32", ctypes.c_uint32),
                ("got64", ctypes.c_uint64), ("member", ctypes.c_uint32), ("trailer_byte", ctypes.c_uint64),
                ("msg", ctypes.c_char_p)]


class Stats(ctypes.Structure):
    _fields_ = [("compressed_bytes", ctypes.c_uint64), ("uncompressed_bytes", ctypes.c_uint64),
                ("chunks_decoded", ctypes.c_uint64), ("speculation_failures", ctypes.c_uint64)]


class MemberBoundary(ctypes.Structure):
    _fields_ = [("out_off", ctypes.c_uint64), ("crc", ctypes.c_uint32), ("isize", ctypes.c_uint32),
                ("trailer_byte", ctypes.c_uint64)]


class Chunk(ctypes.Structure):
    _fields_ = [("cells", ctypes.POINTER(ctypes.c_uint16)), ("n", ctypes.c_size_t), ("cap", ctypes.c_size_t),
                ("start_bit", ctypes.c_uint64), ("end_hint", ctypes.c_uint64), ("actual_end_bit", ctypes.c_uint64),
                ("final_block", ctypes.c_int), ("mb", ctypes.POINTER(MemberBoundary)), ("n_mb", ctypes.c_size_t),
                ("cap_mb", ctypes.c_size_t), ("err", ctypes.c_int), ("msg", ctypes.c_char_p)]


SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t)

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.orc_crc32.argtypes = [ctypes.c_uint32, vp, sz]
        L.orc_crc32.restype = ctypes.c_uint32
        L.orc_crc32_combine.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
        L.orc_crc32_combine.restype = ctypes.c_uint32
        L.orc_parse_header.argtypes = [vp, sz]
        L.orc_parse_header.restype = ctypes.c_long
        L.orc_parse_bgzf_block_size.argtypes = [vp, sz]
        L.orc_parse_bgzf_block_size.restype = ctypes.c_uint32
        L.orc_find_next_dynamic_block.argtypes = [vp, sz, ctypes.c_uint64, ctypes.c_uint64,
                                                  ctypes.POINTER(ctypes.c_uint64)]
        L.orc_strict_header_ok.argtypes = [vp, sz, ctypes.c_uint64]
        L.orc_strict_header_ok.restype = ctypes.c_int
        L.orc_next_lut.argtypes = [ctypes.c_uint32]
        L.orc_next_lut.restype = ctypes.c_uint8
        L.orc_safe_inflate.argtypes = [vp, sz, vp, sz, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        L.orc_decode_chunk.argtypes = [vp, sz, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(Chunk)]
        L.orc_chunk_free.argtypes = [ctypes.POINTER(Chunk)]
        L.orc_resolve_cells.argtypes = [vp, sz, vp, sz, vp]
        L.orc_read_gz.argtypes = [vp, sz, sz, ctypes.c_int, SINK, vp, ctypes.POINTER(Stats), ctypes.POINTER(Error)]
        L.orc_read_gz_stream_serial.argtypes = [vp, sz, SINK, vp, ctypes.POINTER(Stats), ctypes.POINTER(Error)]
        L.orc_chunk_boundaries.argtypes = [vp, sz, sz, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), sz]
        L.orc_chunk_boundaries.restype = sz
        _lib = L
    return _lib


def _buf(data) -> tuple[np.ndarray, int]:
    a = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray, memoryview)) else np.asarray(
        data, dtype=np.uint8)
    a = np.ascontiguousarray(a)
    return a, (a.ctypes.data if a.size else 0)


def crc32(data, crc: int = 0) -> int:
    a, p = _buf(data)
    return lib().orc_crc32(crc, p, a.size)


def crc32_combine(a: int, b: int, len_b: int) -> int:
    return lib().orc_crc32_combine(a, b, len_b)


def parse_header(data) -> int:
    a, p = _buf(data)
    return lib().orc_parse_header(p, a.size)


def bgzf_block_size(data) -> int:
    a, p = _buf(data)
    return lib().orc_parse_bgzf_block_size(p, a.size)


def find_next_dynamic_block(data, from_bit: int, until_bit: int | None = None):
    a, p = _buf(data)
    out = ctypes.c_uint64()
    ub = a.size * 8 if until_bit is None else until_bit
    return out.value if lib().orc_find_next_dynamic_block(p, a.size, from_bit, ub, ctypes.byref(out)) else None


def strict_header_ok(data, pos: int) -> bool:
    """block_finder.rs:196-262: the strict dynamic header parse at bit `pos`."""
    a, p = _buf(data)
    return bool(lib().orc_strict_header_ok(p, a.size, pos))


def next_lut(w: int) -> int:
    return lib().orc_next_lut(w)


def safe_inflate(data, cap: int):
    """(status, bytes, consumed) of the textbook inflate (safe_inflate.rs:617-644)."""
    a, p = _buf(data)
    out = np.empty(max(cap, 1), dtype=np.uint8)
    n, used = ctypes.c_size_t(), ctypes.c_size_t()
    rc = lib().orc_safe_inflate(p, a.size, out.ctypes.data, cap, ctypes.byref(n), ctypes.byref(used))
    return rc, out[:n.value].tobytes(), used.value


def decode_chunk(body, start_bit: int, end_hint: int):
    """Speculative chunk decode: dict(err, cells (np.uint16), actual_end_bit, final, members)."""
    a, p = _buf(body)
    c = Chunk()
    rc = lib().orc_decode_chunk(p, a.size, start_bit, end_hint, ctypes.byref(c))
    cells = np.ctypeslib.as_array(c.cells, shape=(c.n,)).copy() if c.n else np.zeros(0, np.uint16)
    mbs = [(c.mb[i].out_off, c.mb[i].crc, c.mb[i].isize, c.mb[i].trailer_byte) for i in range(c.n_mb)]
    res = dict(err=rc, cells=cells, actual_end_bit=c.actual_end_bit, final=bool(c.final_block), members=mbs)
    lib().orc_chunk_free(ctypes.byref(c))
    return res


def resolve_cells(cells: np.ndarray, tail: bytes):
    cells = np.ascontiguousarray(cells, dtype=np.uint16)
    t, tp = _buf(tail)
    out = np.empty(max(cells.size, 1), dtype=np.uint8)
    rc = lib().orc_resolve_cells(cells.ctypes.data if cells.size else 0, cells.size, tp, t.size, out.ctypes.data)
    return rc, out[:cells.size].tobytes()


def chunk_boundaries(body, chunk_size: int, threads: int = 4) -> list[int]:
    a, p = _buf(body)
    cap = a.size * 8 // (max(chunk_size, 65536) * 8) + 2
    out = (ctypes.c_uint64 * cap)()
    n = lib().orc_chunk_boundaries(p, a.size, chunk_size, threads, out, cap)
    return [out[i] for i in range(n)]


def read_gz(data, chunk_size: int = 4 << 20, threads: int = 4, keep: bool = True, sink_limit: int | None = None):
    """Restated read_gz over in-memory data: (rc, chunks, Stats, Error).

    keep=False discards the bytes (timed baseline); sink_limit closes the sink
    after that many chunks."""
    a, p = _buf(data)
    chunks: list[bytes] = []
    count = [0]

    def cb(_u, ptr, n):
        if sink_limit is not None and count[0] >= sink_limit:
            return 1
        count[0] += 1
        if keep:
            chunks.append(ctypes.string_at(ptr, n))
        return 0

    st, er = Stats(), Error()
    rc = lib().orc_read_gz(p, a.size, chunk_size, threads, SINK(cb), None, ctypes.byref(st), ctypes.byref(er))
    return rc, chunks, st, er


def read_gz_stream_serial(data, sink_limit: int | None = None):
    """The single-worker streaming path for non-seekable inputs (lib.rs:244-361):
    (rc, chunks, Stats, Error).  sink_limit closes the sink after that many chunks."""
    a, p = _buf(data)
    chunks: list[bytes] = []

    def cb(_u, ptr, n):
        if sink_limit is not None and len(chunks) >= sink_limit:
            return 1
        chunks.append(ctypes.string_at(ptr, n))
        return 0

    st, er = Stats(), Error()
    rc = lib().orc_read_gz_stream_serial(p, a.size, SINK(cb), None, ctypes.byref(st), ctypes.byref(er))
    return rc, chunks, st, er
