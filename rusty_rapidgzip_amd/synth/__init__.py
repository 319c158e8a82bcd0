"""Seeded synthetic workloads (text / FASTQ-like / mixed-entropy) and gzip/BGZF writers.

Thin ctypes wrapper over librrg_synth.so; every generator is deterministic in
its seed and independent of the thread count.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .. import _native

SEED_TEXT = 0x7E570001
SEED_FASTQ = 0xFA570002
SEED_MIXED = 0x313D0004

_typed = False


def _lib():
    global _typed
    lib = _native.synth_lib()
    if not _typed:
        u8p = ctypes.c_void_p
        lib.rrg_synth_text.argtypes = [ctypes.c_uint64, u8p, ctypes.c_size_t, ctypes.c_int]
        lib.rrg_synth_fastq.argtypes = [ctypes.c_uint64, u8p, ctypes.c_size_t, ctypes.c_int]
        lib.rrg_synth_mixed.argtypes = [ctypes.c_uint64, u8p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_size_t, ctypes.c_int]
        lib.rrg_synth_mixed.restype = ctypes.c_size_t
        lib.rrg_gzip_bound.argtypes = [ctypes.c_size_t]
        lib.rrg_gzip_bound.restype = ctypes.c_size_t
        lib.rrg_bgzf_bound.argtypes = [ctypes.c_size_t]
        lib.rrg_bgzf_bound.restype = ctypes.c_size_t
        lib.rrg_gzip.argtypes = [u8p, ctypes.c_size_t, ctypes.c_int, u8p, ctypes.c_size_t, ctypes.c_int,
                                 ctypes.c_size_t]
        lib.rrg_gzip.restype = ctypes.c_size_t
        lib.rrg_gzip_segments.argtypes = [u8p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                          u8p, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t]
        lib.rrg_gzip_segments.restype = ctypes.c_size_t
        lib.rrg_bgzf.argtypes = [u8p, ctypes.c_size_t, ctypes.c_int, u8p, ctypes.c_size_t, ctypes.c_int]
        lib.rrg_bgzf.restype = ctypes.c_size_t
        lib.rrg_members_bound.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
        lib.rrg_members_bound.restype = ctypes.c_size_t
        lib.rrg_members.argtypes = [u8p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, u8p, ctypes.c_size_t,
                                    ctypes.c_int]
        lib.rrg_members.restype = ctypes.c_size_t
        lib.rrg_sink_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        lib.rrg_sink_copy.restype = ctypes.c_int
        lib.rrg_sink_count.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        lib.rrg_sink_count.restype = ctypes.c_int
        _typed = True
    return lib


def _threads(n: int | None) -> int:
    if n:
        return n
    return min(16, os.cpu_count() or 1)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def text(n: int, seed: int = SEED_TEXT, threads: int | None = None) -> np.ndarray:
    """Order-2 Markov text over '!'..'`' with Dirichlet(0.3) rows and ~80-char lines."""
    out = np.empty(n, dtype=np.uint8)
    if _lib().rrg_synth_text(seed, _ptr(out), n, _threads(threads)) != 0:
        raise RuntimeError("text generator failed")
    return out


def fastq(n: int, seed: int = SEED_FASTQ, threads: int | None = None) -> np.ndarray:
    """FASTQ-like records: '@r<id>', 150 bp ACGT (+N at ~0.001), '+', 150-char Phred walk."""
    out = np.empty(n, dtype=np.uint8)
    if _lib().rrg_synth_fastq(seed, _ptr(out), n, _threads(threads)) != 0:
        raise RuntimeError("fastq generator failed")
    return out


def mixed(n: int, seed: int = SEED_MIXED, threads: int | None = None):
    """Mixed-entropy segments; returns (raw, levels, segment_ends)."""
    out = np.empty(n, dtype=np.uint8)
    cap = n // (1 << 20) + 16
    lv = np.zeros(cap, dtype=np.int32)
    ends = np.zeros(cap, dtype=np.uint64)
    ns = _lib().rrg_synth_mixed(seed, _ptr(out), n, _ptr(lv), _ptr(ends), cap, _threads(threads))
    return out, lv[:ns].copy(), ends[:ns].copy()


def gzip(raw: np.ndarray, level: int = 6, threads: int | None = None, piece: int = 64 << 20) -> np.ndarray:
    """One gzip member at `level` (pieces primed with the previous 32 KiB, joined by sync flushes)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    cap = _lib().rrg_gzip_bound(raw.size)
    out = np.empty(cap, dtype=np.uint8)
    n = _lib().rrg_gzip(_ptr(raw), raw.size, level, _ptr(out), cap, _threads(threads), piece)
    if n == 0:
        raise RuntimeError("gzip writer failed")
    return out[:n].copy()


def gzip_segments(raw: np.ndarray, levels: np.ndarray, ends: np.ndarray, threads: int | None = None,
                  piece: int = 64 << 20) -> np.ndarray:
    """One gzip member whose compression level switches at segment ends (deflateParams)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    levels = np.ascontiguousarray(levels, dtype=np.int32)
    ends = np.ascontiguousarray(ends, dtype=np.uint64)
    cap = _lib().rrg_gzip_bound(raw.size)
    out = np.empty(cap, dtype=np.uint8)
    n = _lib().rrg_gzip_segments(_ptr(raw), raw.size, _ptr(levels), _ptr(ends), levels.size, _ptr(out), cap,
                                 _threads(threads), piece)
    if n == 0:
        raise RuntimeError("gzip writer failed")
    return out[:n].copy()


def members(raw: np.ndarray, member: int = 1 << 20, level: int = 6, threads: int | None = None) -> np.ndarray:
    """pigz-style multi-member gzip: consecutive `member`-byte pieces, each its own
    member with a plain header (no BGZF extra field)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    cap = _lib().rrg_members_bound(raw.size, member)
    out = np.empty(cap, dtype=np.uint8)
    n = _lib().rrg_members(_ptr(raw), raw.size, member, level, _ptr(out), cap, _threads(threads))
    if n == 0 and raw.size:
        raise RuntimeError("member writer failed")
    return out[:n].copy()


def bgzf(raw: np.ndarray, level: int = 6, threads: int | None = None) -> np.ndarray:
    """BGZF: 65280-byte blocks as gzip members with BC/BSIZE, plus the EOF block."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    cap = _lib().rrg_bgzf_bound(raw.size)
    out = np.empty(cap, dtype=np.uint8)
    n = _lib().rrg_bgzf(_ptr(raw), raw.size, level, _ptr(out), cap, _threads(threads))
    if n == 0:
        raise RuntimeError("bgzf writer failed")
    return out[:n].copy()


class CopySink(ctypes.Structure):
    """State of rrg_sink_copy: a C chunk callback that copies every chunk into
    one reused buffer of `cap` bytes and counts bytes and chunks."""

    _fields_ = [("buf", ctypes.c_void_p), ("cap", ctypes.c_size_t), ("bytes", ctypes.c_uint64),
                ("chunks", ctypes.c_uint64)]

    def __init__(self, cap: int = 8 << 20):
        super().__init__()
        self._mem = np.empty(cap, np.uint8)
        self.buf = self._mem.ctypes.data
        self.cap = cap

    @staticmethod
    def callback(copy: bool = True) -> int:
        """Address of rrg_sink_copy (an rrg_chunk_cb), or with copy=False of
        rrg_sink_count, which only counts."""
        return ctypes.cast(_lib().rrg_sink_copy if copy else _lib().rrg_sink_count, ctypes.c_void_p).value
