"""Gzip inputs with a known block structure, shared by the oracle and GPU tests.

zlib ends the current block at every Z_SYNC_FLUSH (the data block, then an
empty stored block), and never inside a piece shorter than its 16 K-symbol
buffer, so the non-final block ends of these members are exactly the piece
ends.  That pins the single-worker chunking (chunks end at block ends)
independently of any decoder.
"""
import random
import struct
import zlib

import numpy as np

from oracle import oracle

HEADER = b"\x1f\x8b\x08\x00\x00\x00\x00\x00\x00\xff"
WINDOW = 32768


def printable(n, seed, lo=32, hi=127):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi, n, dtype=np.uint8).tobytes()


def trailer(data):
    return struct.pack("<II", zlib.crc32(data) & 0xFFFFFFFF, len(data) & 0xFFFFFFFF)


def synced_member(pieces, level=6):
    """One gzip member whose non-final blocks end at the piece ends.

    Returns (member bytes, payload, block ends as payload offsets)."""
    co = zlib.compressobj(level, zlib.DEFLATED, -15)
    body, ends, pos = [], [], 0
    for p in pieces:
        assert len(p) < 16000
        body.append(co.compress(p))
        body.append(co.flush(zlib.Z_SYNC_FLUSH))
        pos += len(p)
        ends += [pos, pos]  # the data block and the empty stored block
    body.append(co.flush(zlib.Z_FINISH))  # an empty final block
    payload = b"".join(pieces)
    return HEADER + b"".join(body) + trailer(payload), payload, ends


def synced_file(seed, n_members=2):
    """Members of synced_member over random pieces of 1-12 KB at levels 1/6/9.
    Returns (file, payload, [(member length, block ends)])."""
    rng = random.Random(seed)
    data, payload, members = b"", b"", []
    for m in range(n_members):
        pieces = [printable(rng.randint(1000, 12000), seed * 1000 + m * 100 + i) for i in range(rng.randint(20, 60))]
        g, p, ends = synced_member(pieces, level=rng.choice([1, 6, 9]))
        data, payload = data + g, payload + p
        members.append((len(p), ends))
    return data, payload, members


def serial_chunks(members, chunk_size, streaming):
    """Chunk lengths of the single-worker paths for members given as
    (member length, non-final block ends): after a non-final block all but the
    last 32 KiB are sent once chunk size + 32 KiB (streaming: more than
    32 KiB) are pending; a member's final block sends the rest
    (pipeline.rs:1254-1287, lib.rs:286-303)."""
    T = max(chunk_size, 65536)
    out = []
    for mlen, ends in members:
        sent = 0
        for e in ends:
            if (e > sent + WINDOW) if streaming else (e >= sent + T + WINDOW):
                out.append(e - WINDOW - sent)
                sent = e - WINDOW
        if mlen > sent:
            out.append(mlen - sent)
    return out


def reach_before_member(a_len, z_len, y_ref, y_new, seed=1):
    """Two members cut from one deflate stream: member 1 holds A, member 2 holds
    Z and then Y, whose first y_ref bytes repeat A's tail.  zlib encodes them
    as back-references reaching across Z into A, that is, before member 2's
    start.  A decoder with a member-local window must reject member 2; a
    chunk that starts at Y's block resolves them from the previous chunk's
    output instead (the bytes are then A's, and member 2's trailer, computed
    over Z + Y, matches).

    Returns (file, payload A + Z + Y, file offset of Y's first block, lengths)."""
    # disjoint alphabets: Z and Y's new bytes never match A by chance
    A = printable(a_len, seed, 33, 80)
    Z = printable(z_len, seed + 1, 80, 127)
    Y = A[-y_ref:] + printable(y_new, seed + 2, 80, 127)
    assert z_len + y_ref <= WINDOW - 262
    co = zlib.compressobj(6, zlib.DEFLATED, -15)
    sa = co.compress(A) + co.flush(zlib.Z_SYNC_FLUSH)
    sz = co.compress(Z) + co.flush(zlib.Z_SYNC_FLUSH)
    sy = co.compress(Y) + co.flush(zlib.Z_SYNC_FLUSH) + co.flush(zlib.Z_FINISH)  # Y's block is non-final
    m1 = HEADER + sa + b"\x03\x00" + trailer(A)  # an empty final fixed block ends member 1
    m2 = HEADER + sz + sy + trailer(Z + Y)
    y_at = len(m1) + len(HEADER) + len(sz)
    return m1 + m2, A + Z + Y, y_at, (len(A), len(Z), len(Y))


CL_ORDER = (16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15)


class BitWriter:
    """Bits LSB-first, as DEFLATE packs them (RFC 1951 3.1.1)."""

    def __init__(self, lead=0):
        self.bits = [0] * lead

    def put(self, value, n):
        self.bits += [(value >> k) & 1 for k in range(n)]

    def put_code(self, code, n):  # Huffman codes go most significant bit first
        self.bits += [(code >> (n - 1 - k)) & 1 for k in range(n)]

    def tobytes(self, pad_bits=0):
        bits = self.bits + [0] * pad_bits
        bits += [0] * (-len(bits) % 8)
        return bytes(sum(b << k for k, b in enumerate(bits[i:i + 8])) for i in range(0, len(bits), 8))


def canonical_codes(lengths):
    """RFC 1951 3.2.2 codes for a list of code lengths (0 = unused)."""
    count = [0] * 16
    for L in lengths:
        if L:
            count[L] += 1
    nxt, code = [0] * 16, 0
    for L in range(1, 16):
        code = (code + count[L - 1]) << 1 if L > 1 else 0
        nxt[L] = code
    out = []
    for L in lengths:
        out.append(nxt[L] if L else None)
        if L:
            nxt[L] += 1
    return out


def dynamic_header(precode, seq, hlit, hdist, hclen=19, lead=0, bfinal=0):
    """A dynamic block header: `precode` = 19 code-length-code lengths in
    symbol order, `seq` = (code-length symbol, extra bits value) pairs.
    Returns (BitWriter, number of header bits after the lead)."""
    w = BitWriter(lead)
    w.put(bfinal, 1)
    w.put(2, 2)
    w.put(hlit - 257, 5)
    w.put(hdist - 1, 5)
    w.put(hclen - 4, 4)
    for k in range(hclen):
        w.put(precode[CL_ORDER[k]], 3)
    codes = canonical_codes(precode)
    for sym, extra in seq:
        w.put_code(codes[sym], precode[sym])
        if sym == 16:
            w.put(extra, 2)
        elif sym == 17:
            w.put(extra, 3)
        elif sym == 18:
            w.put(extra, 7)
    return w, len(w.bits) - lead


# ---------------------------------------------------------------------------
# Streams that pin the block finder's candidate choice
# ---------------------------------------------------------------------------
# A complete precode over all 19 code-length symbols: 13 of 4 bits, 6 of 5.
# The 5-bit codes (the all-ones one among them) go to symbols 10-14, which no
# header here uses, so no header has nine one bits in a row.
PRECODE_19 = [4] * 9 + [5] * 6 + [4] * 4


def length_seq(lengths):
    """(code-length symbol, extra) pairs for a list of code lengths: zero runs
    of 3-10 as symbol 17, of 11-138 as symbol 18, other lengths literally."""
    seq, i = [], 0
    while i < len(lengths):
        if lengths[i] == 0:
            j = i
            while j < len(lengths) and lengths[j] == 0 and j - i < 138:
                j += 1
            run = j - i
            if run >= 11:
                seq.append((18, run - 11))
                i = j
                continue
            if run >= 3:
                seq.append((17, run - 3))
                i = j
                continue
        seq.append((lengths[i], 0))
        i += 1
    return seq


def put_header(w, lit_lens, dist_lens, bfinal=0):
    """Append a dynamic block header (complete precode) to BitWriter w."""
    hdr, _ = dynamic_header(PRECODE_19, length_seq(list(lit_lens) + list(dist_lens)), len(lit_lens), len(dist_lens),
                            bfinal=bfinal)
    w.bits += hdr.bits


# "Byte code": literals 0..254 in 8 bits (code = the byte value), literal 255
# and end-of-block in 9 bits (111111110, 111111111), one 1-bit distance code.
# Both codes pass the strict parse (the literal code is complete, a one-symbol
# distance code is allowed), and any bit string without nine ones in a row can
# be written as a run of its literals.
BYTE_LIT = [8] * 255 + [9, 9]
ONE_DIST = [1]
BYTE_CODES = canonical_codes(BYTE_LIT)


def put_literal(w, b, out):
    w.put_code(BYTE_CODES[b], BYTE_LIT[b])
    out.append(b)


def put_eob(w):
    w.put_code(BYTE_CODES[256], 9)


def put_bits_as_literals(w, bits, out):
    """Write `bits` (stream order) into a byte-code block as literals: each
    group of 8 is the code of the literal of that value, a group of eight
    ones followed by a zero is literal 255.  Zeros complete the last group."""
    bits = list(bits)
    i = 0
    while i < len(bits):
        g = (bits[i:i + 8] + [0] * 8)[:8]
        b = sum(bit << (7 - k) for k, bit in enumerate(g))
        if b != 255:
            put_literal(w, b, out)
            i += 8
        else:
            assert bits[i:i + 9] == [1] * 8 + [0], "nine ones in a row cannot be written"
            put_literal(w, 255, out)
            i += 9


def put_filler(w, nbits, out, seed):
    """Exactly nbits (>= 63) of byte-code literals: printable bytes in 8 bits,
    with nbits mod 8 literals 255 (9 bits each) first."""
    assert nbits >= 63
    nine = nbits % 8
    eight = (nbits - 9 * nine) // 8
    for _ in range(nine):
        put_literal(w, 255, out)
    for b in printable(eight, seed):
        put_literal(w, b, out)


def put_literals(w, data, out):
    for b in data:
        put_literal(w, b, out)


# The code of the trial-decode candidates: literal 'A' = 0, end-of-block = 10,
# length 3 (symbol 257) = 11, and one distance code (0 = distance 1).  The bit
# 1 after a length symbol has no distance code: "no matching code".
TRIAL_LIT = [0] * 258
TRIAL_LIT[65], TRIAL_LIT[256], TRIAL_LIT[257] = 1, 2, 2


def trial_candidate_bits(k, last="bad"):
    """A non-final dynamic header with the trial code, then k - 1 valid
    symbols (literals, every third a length-3 match at distance 1), then at
    symbol k either an undecodable distance ("bad") or end-of-block ("eob")."""
    w = BitWriter()
    put_header(w, TRIAL_LIT, ONE_DIST)
    for i in range(1, k):
        if i % 3 == 0:
            w.put_code(0b11, 2)  # length 3
            w.put_code(0b0, 1)   # distance 1
        else:
            w.put_code(0b0, 1)   # literal 'A'
    if last == "bad":
        w.put_code(0b11, 2)
        w.put_code(0b1, 1)
    else:
        w.put_code(0b10, 2)
    return w.bits


def gzip_member(body_bits, payload):
    w = BitWriter()
    w.bits = list(body_bits)
    return HEADER + w.tobytes() + trailer(payload)


CHUNK_TARGET_BITS = 65536 * 8  # the first chunk target of a 64 KiB chunk size


def finder_trap(kind, seed=11):
    """A gzip member built around its first chunk target T (bit 524288 of the
    deflate body, chunk size 64 KiB).  Block 1 is a byte-code block; what sits
    at T depends on `kind`:

      "trial64" / "trial65": a false candidate written as block-1 literals --
          a dynamic header whose trial symbol 64 / 65 cannot be decoded.  The
          reference's 64-symbol trial rejects the first and accepts the
          second (which then fails as a chunk start: a speculation failure).
      "skip": eleven one bits (a literal ending in 11, then block 1's
          end-of-block), then block 2's header at T + 11.  The 13-bit window
          at T has no candidate at any shift, so the NEXT_LUT chain jumps to
          T + 13 and never tests T + 11; the reference's find is block 3.

    Returns (file, payload, info) with info = dict(B2, B3, T, block_out=[...])
    (block starts as body bit offsets, the output length of each block)."""
    T = CHUNK_TARGET_BITS
    w = BitWriter()
    outs = []
    out = []
    put_header(w, BYTE_LIT, ONE_DIST)
    if kind in ("trial64", "trial65"):
        put_filler(w, T - len(w.bits), out, seed)
        assert len(w.bits) == T
        put_bits_as_literals(w, trial_candidate_bits(64 if kind == "trial64" else 65), out)
        put_literals(w, printable(2000, seed + 1), out)
        put_eob(w)
    else:
        put_filler(w, T - 6 - len(w.bits), out, seed)
        put_literal(w, 35, out)  # code 00100011: ends in two ones at T, T + 1
        put_eob(w)               # nine ones: T + 2 .. T + 10
        assert len(w.bits) == T + 11
    outs.append(out)
    B2 = len(w.bits)
    out = []
    put_header(w, BYTE_LIT, ONE_DIST)
    put_literals(w, printable(3000, seed + 2), out)
    put_eob(w)
    outs.append(out)
    B3 = len(w.bits)
    out = []
    put_header(w, BYTE_LIT, ONE_DIST)
    put_literals(w, printable(1500, seed + 3), out)
    put_eob(w)
    outs.append(out)
    out = []
    put_header(w, BYTE_LIT, ONE_DIST, bfinal=1)
    put_literals(w, printable(700, seed + 4), out)
    put_eob(w)
    outs.append(out)
    payload = b"".join(bytes(o) for o in outs)
    return gzip_member(w.bits, payload), payload, dict(T=T, B2=B2, B3=B3, block_out=[len(o) for o in outs])


def next_lut_brute(w):
    """NEXT_LUT[w] from its definition (block_finder.rs:84-101): the smallest
    shift j in [0, 13) whose zero-padded window passes the 13-bit predicate."""
    def cand(b):
        return (b & 1) == 0 and ((b >> 1) & 3) == 2 and ((b >> 3) & 31) <= 29 and ((b >> 8) & 31) <= 29

    return next((j for j in range(13) if cand(w >> j)), 13)


_NEXT_LUT = []


def chain_positions(body, start, stop):
    """Positions the reference's skip chain visits from `start` (below `stop`),
    with the candidates (jump 0) marked: [(pos, is_candidate)].  A candidate
    that does not verify is left by one bit."""
    if not _NEXT_LUT:
        _NEXT_LUT.extend(next_lut_brute(w) for w in range(1 << 13))
    padded = bytes(body) + b"\0\0\0"
    nbits = len(body) * 8
    out, p = [], start
    while p < stop and p + 13 <= nbits:
        q = p >> 3
        s = _NEXT_LUT[(int.from_bytes(padded[q:q + 3], "little") >> (p & 7)) & 0x1FFF]
        out.append((p, s == 0))
        p += s if s else 1
    return out


def bgzf_stored_member(data):
    """One BGZF member (BC extra field, BSIZE) holding `data` in a stored block:
    31 + len(data) bytes."""
    assert len(data) <= 65535
    size = 18 + 5 + len(data) + 8
    head = b"\x1f\x8b\x08\x04\x00\x00\x00\x00\x00\xff" + struct.pack("<H", 6) + b"BC" + struct.pack("<HH", 2, size - 1)
    block = b"\x01" + struct.pack("<HH", len(data), len(data) ^ 0xFFFF) + data
    return head + block + trailer(data)


def synced_member_starts(pieces, level=6):
    """As synced_member, also returning the byte offset (in the deflate body)
    where each piece's data block starts: zlib writes each piece as one block
    right after the previous sync flush, which ends on a byte boundary."""
    co = zlib.compressobj(level, zlib.DEFLATED, -15)
    body, starts, n = [], [], 0
    for p in pieces:
        assert len(p) < 16000
        starts.append(n)
        a = co.compress(p) + co.flush(zlib.Z_SYNC_FLUSH)
        body.append(a)
        n += len(a)
    body.append(co.flush(zlib.Z_FINISH))
    payload = b"".join(pieces)
    return b"".join(body), payload, starts


def stream_chunk_lengths(data, chunk):
    """Chunk lengths of the reference's streaming path (streaming.rs:278-346):
    each chunk starts at the first verified block at or after the previous
    start + chunk_size, and ends where the decode from its start reaches the
    next chunk's start.  Every start must be a true one (checked here)."""
    hl = oracle.parse_header(data)
    body = data[hl:]
    nbits = len(body) * 8
    cb = max(chunk, 65536) * 8
    B = [0]
    while B[-1] + cb < nbits:
        f = oracle.find_next_dynamic_block(body, B[-1] + cb, nbits)
        if f is None:
            break
        B.append(f)
    lens = []
    for k in range(len(B)):
        end = B[k + 1] if k + 1 < len(B) else (1 << 64) - 1
        r = oracle.decode_chunk(body, B[k], end)
        assert r["err"] == 0
        if k + 1 < len(B):
            assert r["actual_end_bit"] == B[k + 1]  # a true start: the expected lengths are exact
        lens.append(len(r["cells"]))
    return lens
