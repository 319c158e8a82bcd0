"""Long expansion items cut into pieces: segments whose tokens expand to
millions of cells (runs, short-period repeats) expand as pieces of about
512 Ki cells on separate wavefronts, each piece a chain segment whose
sources before it are markers.  Bytes, chunk boundaries and stats must equal
the oracle's (pipeline.rs:1043-1078 carries the window the pieces rely on).
"""
import gzip as pygzip
import sys

import numpy as np
import pytest

import rusty_rapidgzip_amd as rrg
from oracle import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gz(data, level=6):
    return pygzip.compress(data, compresslevel=level, mtime=0)


def text(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, 95, n, dtype=np.uint8) + 32).tobytes()


def check(decoder, data, chunk, expect, seg=0, capfd=None, cut=True):
    rc, ochunks, ost, oer = oracle.read_gz(data, chunk_size=chunk, threads=4)
    assert rc == 0, oer.msg
    out = []
    st = decoder.decode_bytes(data, out, rrg.Config(chunk_size_bytes=chunk, segment_bytes=seg,
                                                    verbose=rrg.Verbosity.On))
    got = b"".join(out)
    assert b"".join(ochunks) == expect
    assert len(got) == len(expect)
    assert got == expect
    assert [len(c) for c in out] == [len(c) for c in ochunks]
    assert st.uncompressed_bytes == ost.uncompressed_bytes
    assert st.chunks_decoded == ost.chunks_decoded
    assert st.speculation_failures == ost.speculation_failures
    if capfd is not None:
        err = capfd.readouterr().err
        assert ("long items cut into" in err) == cut, err[-2000:]


def periodic(seed, periods, reps_bytes):
    parts = []
    for i, p in enumerate(periods):
        unit = text(p, seed + i)
        parts.append((unit * (reps_bytes // p + 1))[:reps_bytes])
        parts.append(text(3000 + 97 * i, seed + 100 + i))
    return b"".join(parts)


@pytest.mark.parametrize("level", [1, 6, 9])
def test_periodic_pieces(decoder, level, capfd):
    # periods up to 32000: matches reach across piece cuts into the piece before
    raw = periodic(3, [1, 3, 7, 64, 300, 4093, 31999], 3 << 20)
    g = gz(raw, level)
    _say("periodic pieces", level, len(raw), len(g))
    check(decoder, g, 64 << 10, raw, seg=32 << 10, capfd=capfd)


def test_first_segment_pieces(decoder, capfd):
    # the file's first segment (no markers of its own) is cut too: its later
    # pieces hold markers into the first
    raw = b"Q" * (5 << 20) + text(1 << 20, 8) + (text(250, 9) * 20000)
    g = gz(raw, 6)
    _say("first segment pieces", len(raw), len(g))
    check(decoder, g, 4 << 20, raw, capfd=capfd)


def test_member_pieces(decoder, capfd):
    # later members start with long runs: the member checks read only the
    # first piece of a cut segment
    a = text(200000, 11)
    b = b"R" * (3 << 20) + text(100000, 12)
    c = (text(17, 13) * 300000) + text(5000, 14)
    raw = a + b + c
    g = gz(a, 6) + gz(b, 6) + gz(c, 1)
    _say("member pieces", len(raw), len(g))
    check(decoder, g, 64 << 10, raw, capfd=capfd)


def test_small_items_not_cut(decoder, capfd):
    raw = text(3 << 20, 21)
    check(decoder, gz(raw, 6), 1 << 20, raw, capfd=capfd, cut=False)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_mixed_seeds(decoder, seed):
    # stored random stretches, runs, short-period repeats, distance-32768
    # repeats and text at levels 1/9 in one member: cut pieces and long
    # exact finder walks together
    from rusty_rapidgzip_amd import synth

    raw, lv, ends = synth.mixed(40 << 20, seed=seed, threads=8)
    g = synth.gzip_segments(raw, lv, ends, threads=8, piece=16 << 20).tobytes()
    _say("mixed seed", seed, len(g))
    check(decoder, g, 1 << 20, raw.tobytes(), seg=32 << 10)
