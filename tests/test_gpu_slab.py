"""GPU: rrg_gpu_decode under a device-memory budget decodes in slabs of whole
chunks (the bounded, streaming form of read_gz; pipeline.rs:272-375,
858-952).  Bytes, chunk boundaries, DecodeStats and errors must equal the
oracle's and the one-slab decode's, whatever the slab boundaries cross: a
member still open, markers reaching into the previous slab, a chunk start the
chain has not reached, BGZF batches."""
import sys

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gz(data, level=6):
    import gzip

    return gzip.compress(data, compresslevel=level, mtime=0)


def _lcg(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 95, n, dtype=np.uint8) + 32).tobytes()


def _decode(decoder, data, chunk, budget, **kw):
    import rusty_rapidgzip_amd as rrg

    out = []
    st = decoder.decode_bytes(data, out, rrg.Config(chunk_size_bytes=chunk, device_budget_bytes=budget, **kw))
    return out, st, decoder.profile()


def _check(decoder, data, chunk, budget, min_slabs, expect=None, threads=16):
    rc, ochunks, ost, oer = oracle.read_gz(data, chunk_size=chunk, threads=threads)
    assert rc == 0, oer.msg
    out, st, prof = _decode(decoder, data, chunk, budget)
    _say("slabs", prof.n_slabs, "peak MiB", prof.device_bytes_peak / 2**20, "first chunk ms", prof.ms_first_chunk,
         "total ms", prof.ms_total)
    assert prof.n_slabs >= min_slabs
    if expect is not None:
        assert b"".join(ochunks) == expect
    assert [len(c) for c in out] == [len(c) for c in ochunks]
    assert b"".join(out) == b"".join(ochunks)
    assert (st.uncompressed_bytes, st.chunks_decoded, st.speculation_failures) == \
        (ost.uncompressed_bytes, ost.chunks_decoded, ost.speculation_failures)
    assert st.compressed_bytes == len(data)
    return out, st, prof


def test_budget_fastq_256m(decoder):
    """A 64 MiB budget on 256 MiB of FASTQ: many slabs, within the budget."""
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(256 << 20, seed=5)
    g = synth.gzip(raw, 6).tobytes()
    _say("fastq 256 MiB", len(g))
    out, st, prof = _check(decoder, g, 1 << 20, 64 << 20, 8, expect=raw.tobytes())
    # the context gave back what earlier calls had grown before the decode
    assert prof.device_bytes_peak <= 64 << 20
    assert prof.ms_first_chunk < prof.ms_total / 4
    # the same chunks as one slab
    out1, st1, prof1 = _decode(decoder, g, 1 << 20, 0)
    assert prof1.n_slabs <= 4
    assert [len(c) for c in out1] == [len(c) for c in out]


@pytest.mark.parametrize("member_kib,min_slabs", [(64, 1), (300, 4), (1024, 4)])
def test_budget_members(decoder, member_kib, min_slabs):
    """Members open across slab boundaries: CRC and ISIZE carried over.  A
    64 KiB member of this text is one final block, which the block finder
    never returns: the file is one chunk, so one slab."""
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(24 << 20, seed=23)
    g = synth.members(raw, member_kib << 10, 6).tobytes()
    _say("members", member_kib, len(g))
    _check(decoder, g, 256 << 10, 8 << 20, min_slabs, expect=raw.tobytes())


def test_budget_plain_members(decoder):
    """Members of every size, including empty ones, at slab boundaries."""
    parts = [_lcg(n, 300 + i) for i, n in enumerate([0, 1, 70_000, 900_000, 0, 333_333, 2_000_000, 5, 1 << 20])]
    g = b"".join(_gz(p, 6) for p in parts) * 3
    _check(decoder, g, 65536, 2 << 20, 4, expect=b"".join(parts) * 3)


def test_budget_speculation_failures(decoder):
    """A gzip stream inside a gzip stream: many finds are false starts, so
    slab boundaries fall where the chain has to re-decode."""
    inner = _gz(_lcg(6 << 20, 42), 6)
    outer = _gz(inner, 1)
    out, st, prof = _check(decoder, outer, 65536, 2 << 20, 4, expect=inner)
    _say("speculation failures", st.speculation_failures)


def test_budget_mixed_entropy(decoder):
    from rusty_rapidgzip_amd import synth

    raw, lv, ends = synth.mixed(24 << 20)
    g = synth.gzip_segments(raw, lv, ends, piece=4 << 20).tobytes()
    _check(decoder, g, 1 << 20, 16 << 20, 3, expect=raw.tobytes())


def test_budget_bgzf(decoder):
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(12 << 20, seed=3)
    bg = synth.bgzf(raw, 6).tobytes()
    _check(decoder, bg, 256 << 10, 4 << 20, 3, expect=raw.tobytes())
    # BGZF followed by a plain gzip member, in the last slab
    tail = _lcg(300_000, 9)
    _check(decoder, bg[:-28] + _gz(tail, 6), 256 << 10, 4 << 20, 3, expect=raw.tobytes() + tail)


def test_budget_errors(decoder):
    """A CRC, an ISIZE and a deflate error in a later slab: the same error
    (kind, member number, trailer offset) after the same chunks as the
    oracle's."""
    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(16 << 20, seed=29)
    g = bytearray(synth.members(raw, 1 << 20, 6).tobytes())
    # the trailer of the 12th member
    ends, pos = [], 0
    while len(ends) < 12:
        hl = oracle.parse_header(bytes(g[pos:pos + 64]))
        n = len(_member_body(g, pos + hl))
        pos += hl + n + 8
        ends.append(pos)
    for flip, kind in ((ends[11] - 6, "CrcMismatch"), (ends[11] - 2, "IsizeMismatch")):
        bad = bytearray(g)
        bad[flip] ^= 0x10
        rc, ochunks, ost, oer = oracle.read_gz(bytes(bad), chunk_size=256 << 10, threads=16)
        assert rc != 0
        got = []
        with pytest.raises(rrg.GzipError) as ei:
            decoder.decode_bytes(bytes(bad), got, rrg.Config(chunk_size_bytes=256 << 10, device_budget_bytes=4 << 20))
        assert ei.value.kind == kind
        assert ei.value.member == oer.member == 11
        assert ei.value.trailer_byte == oer.trailer_byte
        assert [len(c) for c in got] == [len(c) for c in ochunks]
        assert decoder.profile().n_slabs >= 3
    # a broken block in the middle of the stream
    bad = bytearray(g)
    bad[len(g) * 2 // 3] ^= 0xFF
    rc, ochunks, ost, oer = oracle.read_gz(bytes(bad), chunk_size=256 << 10, threads=16)
    got = []
    with pytest.raises(rrg.RapidgzipError):
        decoder.decode_bytes(bytes(bad), got, rrg.Config(chunk_size_bytes=256 << 10, device_budget_bytes=4 << 20))
    assert rc != 0
    assert b"".join(got) == b"".join(ochunks)[:sum(len(c) for c in got)]


def _member_body(g, start):
    import zlib

    d = zlib.decompressobj(-15)
    d.decompress(bytes(g[start:]))
    return bytes(g[start:len(g) - len(d.unused_data)])


def test_budget_sink_closed(decoder):
    """A sink that closes in the third slab ends the decode: Ok, the refused
    chunk's bytes counted, no chunk after it."""
    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(16 << 20, seed=31)
    g = synth.gzip(raw, 6).tobytes()
    got = []

    def sink(b):
        got.append(b)
        if len(got) == 9:
            raise rrg.SinkClosed()

    st = decoder.decode_bytes(g, sink, rrg.Config(chunk_size_bytes=256 << 10, device_budget_bytes=4 << 20))
    assert len(got) == 9 and st.chunks_decoded == 8
    assert st.uncompressed_bytes == sum(len(c) for c in got)
    assert b"".join(got) == raw.tobytes()[:st.uncompressed_bytes]
    # the context is usable afterwards
    _check(decoder, g, 256 << 10, 4 << 20, 3, expect=raw.tobytes())


def _pipe_decode(decoder, data, cfg, on_first=None, hold=None):
    """decode_fd over a pipe fed by a thread.  With `hold`, the feeder writes
    data[:hold], waits (at most 30 s) for the first chunk, then the rest; the
    returned flag says whether the first chunk came before the last byte."""
    import os
    import threading

    first = threading.Event()
    fed_all = threading.Event()

    def feed(fd):
        with os.fdopen(fd, "wb") as f:
            if hold is None:
                f.write(data)
            else:
                f.write(data[:hold])
                f.flush()
                first.wait(30)
                f.write(data[hold:])
        fed_all.set()

    early = []
    out = []

    def sink(b):
        if not out:
            early.append(not fed_all.is_set())
            first.set()
        out.append(b)

    rfd, wfd = os.pipe()
    t = threading.Thread(target=feed, args=(wfd,))
    t.start()
    try:
        st = decoder.decode_fd(rfd, sink, cfg)
    finally:
        first.set()
        t.join()
        os.close(rfd)
    return out, st, bool(early and early[0])


@pytest.mark.parametrize("budget", [0, 4 << 20])
def test_stream_slabs(decoder, budget):
    """The streaming chunking over a pipe, in slabs: chunk lengths as the
    reference's streaming path (each start found from the previous one)."""
    import crafted
    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(24 << 20, seed=41)
    g = synth.gzip(raw, 6, piece=4 << 20).tobytes()
    want = crafted.stream_chunk_lengths(g, 256 << 10)
    out, st, _ = _pipe_decode(decoder, g, rrg.Config(chunk_size_bytes=256 << 10, device_budget_bytes=budget))
    prof = decoder.profile()
    _say("stream slabs", prof.n_slabs, "chunks", len(out), "peak MiB", prof.device_bytes_peak / 2**20)
    assert b"".join(out) == raw.tobytes()
    assert [len(c) for c in out] == want
    assert st.chunks_decoded == len(want) and st.speculation_failures == 0
    assert st.compressed_bytes == len(g)
    if budget:
        assert prof.n_slabs >= 8


def test_stream_first_chunk_before_eof(decoder):
    """A pipe that stops after its first 1.5 MB until a chunk has come out:
    the decode hands out chunks before the input ends."""
    import crafted
    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(16 << 20, seed=43)
    g = synth.gzip(raw, 6, piece=4 << 20).tobytes()
    cfg = rrg.Config(chunk_size_bytes=256 << 10, device_budget_bytes=8 << 20)
    out, st, early = _pipe_decode(decoder, g, cfg, hold=1_500_000)
    assert early, "the first chunk came only after the whole input"
    assert b"".join(out) == raw.tobytes()
    assert [len(c) for c in out] == crafted.stream_chunk_lengths(g, 256 << 10)


def test_stream_members_and_errors(decoder):
    """Members across stream slabs, and a CRC error in a later member: the
    error after the same chunks as the streaming path's."""
    import gzip

    import rusty_rapidgzip_amd as rrg

    parts = [_lcg(n, 500 + i) for i, n in enumerate([700_000, 1, 1_300_000, 0, 900_000, 2_100_000])]
    g = b"".join(gzip.compress(p, 6, mtime=0) for p in parts)
    cfg = rrg.Config(chunk_size_bytes=65536, device_budget_bytes=2 << 20)
    out, st, _ = _pipe_decode(decoder, g, cfg)
    assert b"".join(out) == b"".join(parts)
    assert decoder.profile().n_slabs >= 4
    bad = bytearray(g)
    bad[-6] ^= 1  # the last member's CRC
    got = []
    with pytest.raises(rrg.GzipError) as ei:
        got, st, _ = _pipe_decode(decoder, bytes(bad), cfg)
    assert ei.value.kind == "CrcMismatch" and ei.value.member == len(parts) - 1
