"""CPU: the finder's per-lane strict header check (csrc/rrg_hdr.h) against the
oracle's strict header parse (block_finder.rs:196-262), bit position by bit
position.  The finder rejects candidates with the per-lane check before its
wave-cooperative verification, so the two must agree everywhere: a position
the check wrongly rejects is a block start the finder would miss.

The streams hold true dynamic headers (zlib levels 1, 6, 9 over text and
FASTQ-like records), random bits, and stream ends cut at every alignment, so
each early exit of the parse (field ranges, precode, overruns, EOB length,
Kraft sums, end of input) is reached.
"""
import os
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("strict_header")
    obj = str(d / "oracle.o")
    exe = str(d / "strict_header_check")
    subprocess.check_call(["gcc", "-O2", "-c", "-o", obj, os.path.join(ROOT, "oracle", "rrg_oracle.c")])
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "strict_header_check.cpp"),
                           obj, "-lpthread"])
    return exe, d


def _run(harness, name, body, first=0, end=None):
    exe, d = harness
    path = str(d / (name + ".bin"))
    with open(path, "wb") as f:
        f.write(body)
    end = len(body) * 8 if end is None else end
    r = subprocess.run([exe, path, str(first), str(end)], capture_output=True, text=True)
    assert r.returncode == 0, (name, r.stdout, r.stderr)
    return int(r.stdout.split("strict ")[1].split()[0])


def _deflate(data, level):
    co = zlib.compressobj(level, zlib.DEFLATED, -15)
    return co.compress(data) + co.flush()


def test_matches_oracle_on_compressed_streams(harness):
    from rusty_rapidgzip_amd import synth

    fq = synth.fastq(3 << 20, seed=5).tobytes()
    text = b"".join(b"line %d of some text with words %d %d\n" % (i, i * 7 % 1000, i * 13 % 777) for i in range(60000))
    rng = np.random.default_rng(3)
    skewed = rng.choice(np.frombuffer(b"ACGTN\n@+", dtype=np.uint8), 1 << 20, p=[.3, .3, .2, .15, .01, .02, .01, .01])
    passes = 0
    for name, raw in (("fastq", fq), ("text", text), ("skewed", skewed.tobytes())):
        for level in (1, 6, 9):
            passes += _run(harness, f"{name}{level}", _deflate(raw, level)[: 1 << 20])
    assert passes > 30  # true headers and strict false positives are among the positions


def test_matches_oracle_on_random_bits_and_cut_ends(harness):
    rng = np.random.default_rng(11)
    _run(harness, "random", rng.integers(0, 256, 1 << 19, dtype=np.uint8).tobytes())
    from rusty_rapidgzip_amd import synth

    body = _deflate(synth.fastq(200_000, seed=2).tobytes(), 6)
    # every header of the stream, with the input cut short inside it
    for cut in range(1, 700, 7):
        _run(harness, f"cut{cut}", body[:cut], 0, cut * 8)
    n = len(body) * 8
    _run(harness, "tail", body, max(0, n - 20000), n)


# code-length code: eight symbols of 3 bits (a complete code)
PRE = [3 if s in (0, 1, 5, 7, 8, 16, 17, 18) else 0 for s in range(19)]
EIGHTS = [(8, 0)] + [(16, 3)] * 41 + [(16, 2)]  # 252 lengths of 8


def _cases():
    """name -> (precode, symbols, hlit, hdist, hclen, the strict parse's verdict)"""
    lit_tail = [(17, 1), (7, 0), (7, 0), (0, 0), (0, 0)]  # 4 zeros, EOB and 257 at 7, 2 zeros
    return {
        "valid": (PRE, EIGHTS + lit_tail + [(1, 0), (1, 0)], 260, 2, 19, True),
        "single_distance_code": (PRE, EIGHTS + lit_tail + [(5, 0)], 260, 1, 19, True),
        "no_distance_codes": (PRE, EIGHTS + lit_tail + [(0, 0), (0, 0)], 260, 2, 19, True),
        "incomplete_distance_code": (PRE, EIGHTS + lit_tail + [(5, 0), (5, 0)], 260, 2, 19, False),
        "oversubscribed_distance_code": (PRE, EIGHTS + lit_tail + [(1, 0), (1, 0), (1, 0)], 260, 3, 19, False),
        # 3 zeros first, then 253 eights, EOB at 7 and 257 at 8
        "repeat_first": (PRE, [(16, 0)] + EIGHTS + [(8, 0), (7, 0), (8, 0), (1, 0), (1, 0)], 258, 2, 19, False),
        # a repeat after a zero run repeats zeros: 6 zeros, 250 eights, 7 7 8 8
        "repeat_after_zeros": (PRE, [(17, 0), (16, 0)] + EIGHTS[:-1] + [(16, 0), (7, 0), (7, 0), (8, 0), (8, 0),
                                                                       (1, 0), (1, 0)], 260, 2, 19, True),
        # a literal/length code of one code (the end-of-block symbol)
        "only_eob": (PRE, [(18, 127), (18, 106), (0, 0), (1, 0), (1, 0), (1, 0)], 257, 2, 19, True),
        "eob_unused": (PRE, EIGHTS + [(8, 0), (8, 0), (8, 0), (0, 0), (0, 0), (8, 0), (1, 0), (1, 0)], 258, 2, 19,
                       False),
        "incomplete_literal_code": (PRE, EIGHTS + [(17, 1), (7, 0), (0, 0), (0, 0), (0, 0), (1, 0), (1, 0)], 260, 2,
                                    19, False),
        "zero_run_overruns": (PRE, EIGHTS + [(17, 1), (7, 0), (7, 0), (18, 0)], 260, 2, 19, False),
        "zero_run_overruns_by_one": (PRE, EIGHTS + lit_tail + [(17, 0)], 260, 2, 19, False),
        "repeat_overruns": (PRE, EIGHTS + lit_tail + [(1, 0), (16, 0)], 260, 2, 19, False),
        "short_precode": ([3 if s in (0, 1, 7, 8, 16, 17, 18) else 0 for s in range(19)], EIGHTS + lit_tail + [(1, 0),
                                                                                                          (1, 0)],
                          260, 2, 19, False),
        "hlit_287": (PRE, EIGHTS + lit_tail + [(0, 0)] * 27 + [(1, 0), (1, 0)], 287, 2, 19, False),
    }


def test_crafted_headers_every_rule(harness):
    """Each rule of the strict parse decides one crafted header; the header
    sits at every bit alignment, and the input also ends inside it and just
    after it (the 10-bit probe rule at the end of input)."""
    from oracle import oracle

    exe, d = harness
    rng = np.random.default_rng(9)
    for name, (pre, seq, hlit, hdist, hclen, want) in _cases().items():
        for lead in range(8):
            w, nh = crafted_header(pre, seq, hlit, hdist, hclen, lead)
            body = w.tobytes() + rng.integers(0, 256, 64, dtype=np.uint8).tobytes()
            assert oracle.strict_header_ok(body, lead) == want, (name, lead)
            _run(harness, f"{name}_{lead}", body)
            for extra in (-2, 0, 1, 2, 3):  # input ends inside or just after the header
                cut = (lead + nh + 7) // 8 + extra
                if cut > 0:
                    _run(harness, f"{name}_{lead}_cut{extra}", body[:cut])


def crafted_header(pre, seq, hlit, hdist, hclen, lead):
    import crafted

    return crafted.dynamic_header(pre, seq, hlit, hdist, hclen, lead=lead)


def test_crafted_valid_header_decodes_with_zlib():
    """The crafted encoder is right: the valid header, two literals and EOB
    make a final block that zlib decodes."""
    import crafted
    from oracle import oracle

    pre, seq, hlit, hdist, hclen, want = _cases()["valid"]
    assert want
    w, nh = crafted.dynamic_header(pre, seq, hlit, hdist, hclen, bfinal=1)
    lit = [8] * 252 + [0] * 4 + [7, 7] + [0, 0]  # the lengths `seq` encodes
    codes = crafted.canonical_codes(lit)
    for sym, n in ((65, 8), (66, 8), (256, 7)):
        w.put_code(codes[sym], n)
    body = w.tobytes()
    assert zlib.decompress(body, -15) == b"AB"
    assert oracle.strict_header_ok(body, 0)
