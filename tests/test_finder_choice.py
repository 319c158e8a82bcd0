"""CPU: the block finder's candidate choice and the BGZF batch ramp, pinned by
streams whose answer is fixed by construction rather than by the oracle.

The chunk boundaries and `speculation_failures` of read_gz follow from which
candidate `find_next_dynamic_block` returns (block_finder.rs:109-141): the
first position on the NEXT_LUT skip chain that passes the strict header parse
and a trial decode of TRIAL_SYMBOLS = 64 symbols (block_finder.rs:47-55,
268-301).  BGZF output is cut into batches whose compressed size ramps from
64 KiB to the chunk size (pipeline.rs:1488-1551).  Each test below fails for a
plausible restatement error: a trial of 63 or 65 symbols, a match counted as
two symbols, candidates tried in bit order instead of chain order, a ramp that
does not start at 64 KiB or does not double.
"""
import random
import zlib

import pytest

import crafted
from oracle import oracle


def _gz_body(f):
    return f[oracle.parse_header(f):]


@pytest.mark.parametrize("k", [1, 2, 63, 64])
def test_trial_rejects_undecodable_symbol_within_64(k):
    # the candidate at bit 0 passes the strict parse; symbol k has no code
    bits = crafted.trial_candidate_bits(k, "bad")
    w = crafted.BitWriter()
    w.bits = bits
    data = w.tobytes(pad_bits=512)  # zero bits after it decode as literals
    assert oracle.strict_header_ok(data, 0)
    assert oracle.find_next_dynamic_block(data, 0, 13) is None


@pytest.mark.parametrize("k", [65, 66, 200])
def test_trial_accepts_after_64_symbols(k):
    bits = crafted.trial_candidate_bits(k, "bad")
    w = crafted.BitWriter()
    w.bits = bits
    data = w.tobytes(pad_bits=512)
    assert oracle.find_next_dynamic_block(data, 0, 13) == 0


@pytest.mark.parametrize("k", [1, 30, 64])
def test_trial_end_of_block_accepts(k):
    # end-of-block at symbol k <= 64 ends the trial successfully, whatever follows
    w = crafted.BitWriter()
    w.bits = crafted.trial_candidate_bits(k, "eob") + [1, 1, 1] * 40
    data = w.tobytes(pad_bits=512)
    assert oracle.find_next_dynamic_block(data, 0, 13) == 0


def test_trial_bits_need_ten_available():
    # a LUT probe peeks 10 bits (huffman.rs:357-366): the trial of a candidate
    # whose 64 symbols end within the last bits of the input fails at the end
    w = crafted.BitWriter()
    w.bits = crafted.trial_candidate_bits(65, "bad")[:-3]
    cut = w.tobytes()
    assert oracle.find_next_dynamic_block(cut, 0, 13) is None
    assert oracle.find_next_dynamic_block(w.tobytes(pad_bits=16), 0, 13) == 0


def test_skip_chain_passes_over_header_at_offset_11():
    f, payload, info = crafted.finder_trap("skip")
    assert zlib.decompress(f, 31) == payload
    body = _gz_body(f)
    T, B2, B3 = info["T"], info["B2"], info["B3"]
    assert B2 == T + 11
    # B2 is a real, verifiable block start: a scan entering at it finds it ...
    assert oracle.find_next_dynamic_block(body, B2, B2 + 13) == B2
    # ... but the window at T has no candidate at any shift, so the chain jumps
    # to T + 13 and the reference's find from T is the next block on the chain
    assert crafted.next_lut_brute(int.from_bytes(body, "little") >> T & 0x1FFF) == 13
    chain = crafted.chain_positions(body, T, B3 + 1)
    assert all(p != B2 for p, _ in chain)
    assert (B3, True) in chain
    assert oracle.find_next_dynamic_block(body, T) == B3


def _expected_find(body, T, starts):
    """First true block start among `starts` that the chain from T visits
    (every other candidate on the way is assumed not to verify)."""
    for p, cand in crafted.chain_positions(body, T, max(starts) + 1):
        if cand and p in starts:
            return p
    return None


@pytest.mark.parametrize("kind", ["trial64", "trial65", "skip"])
def test_trap_chunks_and_speculation(kind):
    """read_gz chunk boundaries and speculation failures around a trap at the
    first chunk target (64 KiB chunks)."""
    f, payload, info = crafted.finder_trap(kind)
    body = _gz_body(f)
    T, B2, B3, outs = info["T"], info["B2"], info["B3"], info["block_out"]
    rc, chunks, st, er = oracle.read_gz(f, chunk_size=65536, threads=4)
    assert rc == 0, er.msg
    assert b"".join(chunks) == payload
    if kind == "trial65":
        # the false candidate is the find: chunk 1's guessed start is not
        # where chunk 0 ends (block 2), a speculation failure
        assert oracle.find_next_dynamic_block(body, T) == T
        assert [len(c) for c in chunks] == [outs[0], sum(outs[1:])]
        assert st.speculation_failures == 1
    else:
        want = _expected_find(body, T + (1 if kind == "trial64" else 0), {B2, B3})
        assert oracle.find_next_dynamic_block(body, T) == want
        first = outs[0] + (outs[1] if want == B3 else 0)
        assert [len(c) for c in chunks] == [first, len(payload) - first]
        assert st.speculation_failures == 0
    assert st.chunks_decoded == len(chunks)


def test_finder_on_known_block_starts():
    """On zlib members whose piece blocks start right after each sync flush,
    the find from any position is the first dynamic block start that the
    skip chain from there visits."""
    rng = random.Random(17)
    pieces = [crafted.printable(rng.randint(2000, 12000), 500 + i) for i in range(40)]
    body, payload, starts = crafted.synced_member_starts(pieces, level=6)
    assert zlib.decompress(body, -15) == payload
    bits = int.from_bytes(body, "little")
    dyn = {8 * s for s in starts if (bits >> (8 * s)) & 7 == 0b100}  # BFINAL = 0, BTYPE = 2
    assert len(dyn) >= 30
    last = max(dyn)
    picks = sorted(rng.randrange(0, last) for _ in range(90)) + sorted(dyn)[:10]
    for frm in picks:
        want = _expected_find(body, frm, dyn | {last + 1})
        got = oracle.find_next_dynamic_block(body, frm)
        assert got == want, (frm, got, want)


def test_bgzf_batch_ramp():
    """BGZF batches: members are added until their compressed bytes reach the
    target, which starts at 64 KiB and doubles up to the chunk size."""
    sizes = [20000] * 10 + [60000] * 2 + [1000] * 5 + [65000] * 6  # member = 31 + payload bytes
    members = [crafted.printable(n, 900 + i) for i, n in enumerate(sizes)]
    f = b"".join(crafted.bgzf_stored_member(m) for m in members)
    # by hand, chunk size 256 KiB: targets 64 Ki, 128 Ki, 256 Ki, 256 Ki
    #   batch 1: 4 x 20031 = 80124 >= 65536
    #   batch 2: 6 x 20031 + 60031 = 180217 >= 131072
    #   batch 3: 60031 + 5 x 1031 + 4 x 65031 = 325310 >= 262144
    #   batch 4: the 2 members left
    counts = [4, 7, 10, 2]
    want, k = [], 0
    for c in counts:
        want.append(sum(sizes[k:k + c]))
        k += c
    rc, chunks, st, er = oracle.read_gz(f, chunk_size=256 << 10, threads=4)
    assert rc == 0, er.msg
    assert b"".join(chunks) == b"".join(members)
    assert [len(c) for c in chunks] == want == [80000, 180000, 325000, 130000]
    # a chunk size below 64 KiB is raised to it, so every target is 64 KiB:
    #   4 x 20031; 4 x 20031; 2 x 20031 + 60031 = 100093;
    #   60031 + 5 x 1031 = 65186 < 65536, + 65031; then 2, 2 and 1 x 65031
    rc, chunks, st, er = oracle.read_gz(f, chunk_size=4096, threads=4)
    assert rc == 0
    assert [len(c) for c in chunks] == [80000, 80000, 100000, 130000, 130000, 130000, 65000]
