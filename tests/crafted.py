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
