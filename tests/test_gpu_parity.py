"""GPU parity: the HIP decoder through the C ABI against the oracle.

Bit-exact bytes, identical chunk boundaries (the reference's Vec stream),
identical DecodeStats, and errors (never faults) on corrupt input.
"""
import gzip as pygzip
import hashlib
import json
import os
import sys
import zlib

import numpy as np
import pytest

import crafted
from conftest import GOLDEN
from oracle import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lcg_payload(n, seed):
    rng = np.random.default_rng(seed & 0xFFFFFFFF)
    return (rng.integers(0, 95, n, dtype=np.uint8) + 32).tobytes()


def gz(data, level=6):
    return pygzip.compress(data, compresslevel=level, mtime=0)


def gpu_decode(decoder, data, chunk, seg=0, lut_bits=8):
    import rusty_rapidgzip_amd as rrg

    out = []
    st = decoder.decode_bytes(data, out, rrg.Config(chunk_size_bytes=chunk, segment_bytes=seg, lut_bits=lut_bits))
    return out, st


def check_same(decoder, data, chunk, seg=0, lut_bits=8, expect=None):
    rc, ochunks, ost, oer = oracle.read_gz(data, chunk_size=chunk, threads=4)
    assert rc == 0, oer.msg
    out, st = gpu_decode(decoder, data, chunk, seg, lut_bits)
    got = b"".join(out)
    want = b"".join(ochunks)
    if expect is not None:
        assert want == expect
    assert len(got) == len(want)
    assert got == want
    assert [len(c) for c in out] == [len(c) for c in ochunks]
    assert st.uncompressed_bytes == ost.uncompressed_bytes
    assert st.chunks_decoded == ost.chunks_decoded
    assert st.speculation_failures == ost.speculation_failures
    return st


def test_corpus_fixtures(decoder):
    sums = json.load(open(os.path.join(GOLDEN, "corpus_sums.json")))
    for name, meta in sorted(sums.items()):
        _say("fixture", name)
        data = open(os.path.join(GOLDEN, "corpus", name), "rb").read()
        for chunk in (65536, 4 << 20):
            out, st = gpu_decode(decoder, data, chunk)
            got = b"".join(out)
            assert hashlib.sha256(got).hexdigest() == meta["sha256"], name
            check_same(decoder, data, chunk)


@pytest.mark.parametrize("seg", [32 << 10, 128 << 10, 0])
@pytest.mark.parametrize("lut_bits", [8, 9, 10])
def test_fastq_small(decoder, seg, lut_bits):
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(6 << 20)
    g = synth.gzip(raw, 6, piece=2 << 20).tobytes()
    _say("fastq", seg, lut_bits, len(g))
    check_same(decoder, g, 1 << 20, seg, lut_bits, expect=raw.tobytes())


@pytest.mark.parametrize("level", [1, 6, 9])
def test_text_levels(decoder, level):
    from rusty_rapidgzip_amd import synth

    raw = synth.text(5 << 20)
    g = synth.gzip(raw, level, piece=2 << 20).tobytes()
    _say("text", level, len(g))
    check_same(decoder, g, 512 << 10, 48 << 10, expect=raw.tobytes())


def test_lcg_levels_matrix(decoder):
    for level in range(1, 10):
        p = lcg_payload(1 << 20, 1000 + level)
        _say("lcg level", level)
        check_same(decoder, gz(p, level), 65536, expect=p)


def test_mixed_entropy(decoder):
    from rusty_rapidgzip_amd import synth

    raw, lv, ends = synth.mixed(24 << 20)
    g = synth.gzip_segments(raw, lv, ends, piece=4 << 20).tobytes()
    _say("mixed", len(g))
    check_same(decoder, g, 1 << 20, 64 << 10, expect=raw.tobytes())


def test_multi_member(decoder):
    parts = [lcg_payload(n, 100 + n) for n in (0, 1, 5000, 300_000, 0, 70_000, 1 << 20)]
    g = b"".join(gz(p, 6) for p in parts)
    check_same(decoder, g, 65536, expect=b"".join(parts))
    tiny = b"".join(gz(bytes([65 + i % 26]) * (i % 7), 6) for i in range(500))
    check_same(decoder, tiny, 65536)


@pytest.mark.parametrize("member_kib", [64, 256, 1024])
def test_pigz_style_members(decoder, member_kib):
    """Many members with plain headers (not BGZF): member ends fall inside
    segments and chunks; chunk boundaries, stats and bytes as the oracle's."""
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(12 << 20, seed=17)
    g = synth.members(raw, member_kib << 10, 6).tobytes()
    _say("pigz members", member_kib, len(g))
    check_same(decoder, g, 1 << 20, expect=raw.tobytes())
    check_same(decoder, g, 1 << 20, seg=16 << 10, expect=raw.tobytes())


def test_nested_gzip_speculation_failures(decoder):
    inner = gz(lcg_payload(2 << 20, 42), 6)
    outer = gz(inner, 1)
    st = check_same(decoder, outer, 65536, expect=inner)
    _say("nested spec failures", st.speculation_failures)


def test_bgzf(decoder):
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(5 << 20)
    bg = synth.bgzf(raw, 6).tobytes()
    check_same(decoder, bg, 1 << 20, expect=raw.tobytes())
    # BGZF followed by a plain gzip member (the non-BGZF remainder path)
    tail = lcg_payload(200_000, 9)
    check_same(decoder, bg[:-28] + gz(tail, 6), 1 << 20, expect=raw.tobytes() + tail)


def test_errors_not_faults(decoder):
    import rusty_rapidgzip_amd as rrg

    p = lcg_payload(400_000, 5)
    good = gz(p, 6)
    bad = bytearray(good)
    bad[-6] ^= 1
    with pytest.raises(rrg.GzipError) as ei:
        gpu_decode(decoder, bytes(bad), 65536)
    assert ei.value.kind == "CrcMismatch"
    bad = bytearray(good)
    bad[-2] ^= 1
    with pytest.raises(rrg.GzipError) as ei:
        gpu_decode(decoder, bytes(bad), 65536)
    assert ei.value.kind == "IsizeMismatch"
    with pytest.raises(rrg.RapidgzipError):
        gpu_decode(decoder, good[: len(good) // 2], 65536)
    with pytest.raises(rrg.RapidgzipError):
        gpu_decode(decoder, good + b"junk", 65536)
    with pytest.raises(rrg.GzipError) as ei:
        gpu_decode(decoder, b"\x1f\x8c" + good[2:], 65536)
    assert ei.value.kind == "BadMagic"
    with pytest.raises(rrg.IoError):
        gpu_decode(decoder, b"", 65536)
    # random corruption inside the deflate body: an error or (if the damage is
    # undetectable by deflate) a CRC error -- never a fault
    rng = np.random.default_rng(11)
    for _ in range(8):
        b = bytearray(good)
        for pos in rng.integers(20, len(b) - 8, 3):
            b[pos] ^= int(rng.integers(1, 256))
        with pytest.raises(rrg.RapidgzipError):
            gpu_decode(decoder, bytes(b), 65536)
    # decoder still healthy afterwards
    check_same(decoder, good, 65536, expect=p)


def test_high_ratio_overflow(decoder):
    # runs expand far beyond a segment's first output region: the lane counts the
    # cells it needs and the segment is regrown once
    parts = []
    rng = np.random.default_rng(21)
    for i in range(24):
        parts.append(bytes([65 + i % 26]) * int(rng.integers(1 << 19, 3 << 20)))
        parts.append(lcg_payload(int(rng.integers(2000, 20000)), i))
    raw = b"".join(parts)
    g = gz(raw, 6)
    _say("high ratio", len(raw), len(g))
    check_same(decoder, g, 1 << 20, seg=32 << 10, expect=raw)
    prof = decoder.profile()
    assert prof.n_inflate_launches <= 3, prof.as_dict()


def test_periodic_copies(decoder):
    # back-references with periods 1, 2, 3, 4 and 8 (the store-only fast path
    # covers 1, 2 and 4), long and short, next to random text
    rng = np.random.default_rng(5)
    parts = []
    for i in range(400):
        period = int(rng.choice([1, 2, 3, 4, 8]))
        unit = lcg_payload(period, 1000 + i)
        parts.append(unit * int(rng.integers(1, 400)))
        parts.append(lcg_payload(int(rng.integers(1, 300)), 2000 + i))
    raw = b"".join(parts)
    for level in (1, 6, 9):
        check_same(decoder, gz(raw, level), 65536, seg=32 << 10, expect=raw)


def test_stored_stretches(decoder):
    # level 0: the whole member is stored blocks; segment targets start at
    # verified stored headers, so the stretch is decoded by many lanes
    raw = np.random.default_rng(9).integers(0, 256, 24 << 20, dtype=np.uint8).tobytes()
    g = gz(raw, 0)
    check_same(decoder, g, 4 << 20, expect=raw)
    assert decoder.profile().n_segments > 100
    # stored stretches between compressed ones
    mixed = raw[: 6 << 20] + lcg_payload(2 << 20, 4) + raw[6 << 20: 9 << 20]
    from rusty_rapidgzip_amd import synth

    levels = np.array([0, 6, 0], dtype=np.int32)
    ends = np.array([6 << 20, 8 << 20, 11 << 20], dtype=np.uint64)
    g2 = synth.gzip_segments(np.frombuffer(mixed, dtype=np.uint8), levels, ends, piece=4 << 20).tobytes()
    check_same(decoder, g2, 1 << 20, expect=mixed)


def test_large_offsets(decoder):
    # compressed offsets beyond 2^31 bits (256 MiB): 64-bit positions end to end
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(1100 << 20)
    g = synth.gzip(raw, 6, piece=64 << 20)
    assert g.size * 8 > (1 << 31) + (8 << 20)
    _say("large", g.size)
    check_same(decoder, g.tobytes(), 4 << 20, expect=raw.tobytes())


def test_device_api(decoder):
    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

    raw = synth.fastq(4 << 20)
    g = synth.gzip(raw, 6, piece=1 << 20).tobytes()
    d = hipmem.DeviceBuffer.from_host(g)
    for host in (g, None):  # with and without the optional host copy
        ptr, n, ends, st = decoder.decode_device(d.ptr, len(g), host, rrg.Config(chunk_size_bytes=1 << 20))
        assert n == raw.size and ends[-1] == n
        rc, ochunks, _, _ = oracle.read_gz(g, chunk_size=1 << 20, threads=4)
        assert np.diff([0] + ends).tolist() == [len(c) for c in ochunks]
        assert hipmem.to_host(ptr, n).tobytes() == raw.tobytes()
    d.free()


def test_sink_closed(decoder):
    import rusty_rapidgzip_amd as rrg

    p = lcg_payload(1 << 20, 77)
    g = gz(p, 6)
    got = []

    def sink(b):
        got.append(b)
        if len(got) == 2:
            raise rrg.SinkClosed()

    st = decoder.decode_bytes(g, sink, rrg.Config(chunk_size_bytes=65536))
    assert st.chunks_decoded == 1  # the third send never happens; the second is not counted
    assert b"".join(got) == p[: sum(len(x) for x in got)]


def test_read_gz_file(decoder, tmp_path):
    import rusty_rapidgzip_amd as rrg

    p = lcg_payload(700_000, 3)
    f = tmp_path / "x.gz"
    f.write_bytes(gz(p, 6))
    out = []
    st = rrg.read_gz(f, out, rrg.Config(chunk_size_bytes=65536))
    assert b"".join(out) == p
    assert st.compressed_bytes == f.stat().st_size
    assert zlib.crc32(b"".join(out)) == zlib.crc32(p)


def test_streaming_pipe(decoder, tmp_path):
    import threading

    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(9 << 20, seed=77)
    g = synth.gzip(raw, 6, piece=4 << 20).tobytes()
    chunk = 1 << 20
    want = crafted.stream_chunk_lengths(g, chunk)
    cfg = rrg.Config(chunk_size_bytes=chunk)

    def feed(fd, data):
        with os.fdopen(fd, "wb") as f:
            f.write(data)

    rfd, wfd = os.pipe()
    t = threading.Thread(target=feed, args=(wfd, g))
    t.start()
    out = []
    st = decoder.decode_fd(rfd, out, cfg)
    t.join()
    os.close(rfd)
    assert b"".join(out) == raw.tobytes()
    assert [len(c) for c in out] == want
    assert st.chunks_decoded == len(want) and st.speculation_failures == 0
    assert st.compressed_bytes == len(g)
    # read_gz on a named pipe takes the same path
    fifo = tmp_path / "in.fifo"
    os.mkfifo(fifo)

    def feed_fifo():
        with open(fifo, "wb") as f:
            f.write(g)

    t = threading.Thread(target=feed_fifo)
    t.start()
    out2 = []
    rrg.read_gz(fifo, out2, cfg)
    t.join()
    assert [len(c) for c in out2] == want
    # an empty file reads no header: Truncated, as the reference's streaming path
    empty = tmp_path / "empty.gz"
    empty.write_bytes(b"")
    with pytest.raises(rrg.GzipError) as ei:
        rrg.read_gz(empty, [], cfg)
    assert ei.value.kind == "Truncated"


def test_autotune_picks_an_equivalent_configuration(decoder):
    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

    raw = synth.fastq(24 << 20, seed=5)
    g = synth.gzip(raw, 6, piece=8 << 20)
    d = hipmem.DeviceBuffer.from_host(g)
    try:
        base = rrg.Config(chunk_size_bytes=1 << 20)
        _say("autotune on", g.size, "bytes")
        tuned, ms = decoder.autotune(d.ptr, g.size, g, base, reps=1)
        assert ms > 0
        assert (tuned.lanes_per_block, tuned.lut_bits) in ((16, 8), (16, 9), (64, 9), (64, 10), (64, 11))
        assert tuned.segment_bytes in (0, 40960, 57344)
        assert tuned.chunk_size_bytes == base.chunk_size_bytes
        ptr, n, ends, st = decoder.decode_device(d.ptr, g.size, g, tuned)
        assert n == raw.size and np.array_equal(hipmem.to_host(ptr, n), raw)
        _, _, ends0, st0 = decoder.decode_device(d.ptr, g.size, g, base)
        assert ends == ends0 and st.speculation_failures == st0.speculation_failures
    finally:
        d.free()
