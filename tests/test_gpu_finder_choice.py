"""GPU: the finder's candidate choice on streams whose answer is fixed by
construction (tests/crafted.py, pinned on the CPU by test_finder_choice.py).

A chunk target meets (a) a false candidate whose 64th trial symbol cannot be
decoded (rejected: no speculation failure), (b) the same with the 65th
(accepted: one speculation failure), (c) a true block start at offset +11 of
a window the skip chain jumps over (the chunk starts at the next block).  The
chunk lengths and DecodeStats of the HIP path must be the constructed ones
and the oracle's.  BGZF batches must follow the 64 KiB -> chunk size ramp.
"""
import sys

import pytest

import crafted
from oracle import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _decode(decoder, data, chunk, **kw):
    import rusty_rapidgzip_amd as rrg

    out = []
    st = decoder.decode_bytes(data, out, rrg.Config(chunk_size_bytes=chunk, **kw))
    return out, st


@pytest.mark.parametrize("seg", [0, 8 << 10])
@pytest.mark.parametrize("kind", ["trial64", "trial65", "skip"])
def test_trap_at_chunk_target(decoder, kind, seg):
    f, payload, info = crafted.finder_trap(kind)
    _say("trap", kind, seg, len(f))
    rc, ochunks, ost, oer = oracle.read_gz(f, chunk_size=65536, threads=4)
    assert rc == 0, oer.msg
    out, st = _decode(decoder, f, 65536, segment_bytes=seg)
    assert b"".join(out) == payload
    outs = info["block_out"]
    if kind == "trial65":
        want = [outs[0], sum(outs[1:])]
        want_spec = 1
    else:
        want = [len(c) for c in ochunks]
        want_spec = 0
        # the chunk starts at a true block start after the trap (B2 or B3)
        assert want[0] in (outs[0], outs[0] + outs[1])
        if kind == "skip":
            assert want[0] == outs[0] + outs[1]
    assert [len(c) for c in out] == [len(c) for c in ochunks] == want
    assert st.speculation_failures == ost.speculation_failures == want_spec
    assert st.chunks_decoded == len(want)


@pytest.mark.parametrize("chunk", [256 << 10, 4096])
def test_bgzf_batch_ramp(decoder, chunk):
    sizes = [20000] * 10 + [60000] * 2 + [1000] * 5 + [65000] * 6
    members = [crafted.printable(n, 900 + i) for i, n in enumerate(sizes)]
    f = b"".join(crafted.bgzf_stored_member(m) for m in members)
    want = {256 << 10: [80000, 180000, 325000, 130000],
            4096: [80000, 80000, 100000, 130000, 130000, 130000, 65000]}[chunk]
    out, st = _decode(decoder, f, chunk)
    assert b"".join(out) == b"".join(members)
    assert [len(c) for c in out] == want
    assert st.chunks_decoded == len(want)
