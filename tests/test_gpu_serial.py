"""GPU: the single-worker paths (Config.num_threads == 1) and member-local
back-references, through the C ABI, against the oracle.

num_threads == 1 selects the reference's serial path (pipeline.rs:1211-1345)
for a regular file and its streaming serial path (lib.rs:244-361) for a
non-seekable one: the bytes are the same, but chunks end at block ends and
errors, stats and a closed sink behave as those paths do.  The oracle's
restatement of both is pinned in test_oracle.py against inputs whose block
ends are known.
"""
import gzip as pygzip
import os
import sys
import threading

import numpy as np
import pytest

import crafted
from oracle import oracle

pytestmark = pytest.mark.gpu

KIND = {1: "UnexpectedEof", 2: "Invalid", 3: "BadMagic", 4: "UnsupportedMethod", 5: "ReservedFlags",
        6: "Truncated", 7: "CrcMismatch", 8: "IsizeMismatch", 9: "Io"}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _kind(e):
    import rusty_rapidgzip_amd as rrg

    return "Io" if isinstance(e, rrg.IoError) else getattr(e, "kind", type(e).__name__)


def _feed(fd, data):
    with os.fdopen(fd, "wb") as f:
        f.write(data)


def gpu(decoder, data, fd=False, sink_limit=None, **cfg):
    """(chunks, stats, error or None) of read_gz on the GPU; fd=True reads the
    data through a pipe (the non-seekable paths)."""
    import rusty_rapidgzip_amd as rrg

    out = []

    def sink(b):
        if sink_limit is not None and len(out) >= sink_limit:
            raise rrg.SinkClosed()
        out.append(bytes(b))

    config = rrg.Config(**cfg)
    try:
        if fd:
            rfd, wfd = os.pipe()
            t = threading.Thread(target=_feed, args=(wfd, data))
            t.start()
            try:
                st = decoder.decode_fd(rfd, sink, config)
            finally:
                t.join()
                os.close(rfd)
        else:
            st = decoder.decode_bytes(data, sink, config)
        return out, st, None
    except rrg.RapidgzipError as e:
        return out, None, e


def want_serial(data, chunk, stream, sink_limit=None):
    if stream:
        return oracle.read_gz_stream_serial(data, sink_limit=sink_limit)
    return oracle.read_gz(data, chunk_size=chunk, threads=1, sink_limit=sink_limit)


def check(name, got, want):
    chunks, st, err = got
    rc, ochunks, ost, oer = want
    assert [len(c) for c in chunks] == [len(c) for c in ochunks], name
    assert chunks == [bytes(c) for c in ochunks], name
    if rc == 0:
        assert err is None, (name, err)
        assert (st.uncompressed_bytes, st.chunks_decoded, st.speculation_failures, st.compressed_bytes) == (
            ost.uncompressed_bytes, ost.chunks_decoded, ost.speculation_failures, ost.compressed_bytes), name
    else:
        assert err is not None, name
        assert _kind(err) == KIND[rc], (name, err, oer.msg)
        if rc == 2:  # the same DeflateError::Invalid message
            assert str(err).endswith(oer.msg.decode()), (name, str(err), oer.msg)
        if rc in (7, 8):
            assert (err.member, err.trailer_byte) == (oer.member, oer.trailer_byte), name


@pytest.mark.parametrize("seg", [0, 4096])
def test_serial_known_block_ends(decoder, seg):
    data, payload, members = crafted.synced_file(5)
    for chunk in (65536, 100_000, 1 << 20):
        _say("serial", "seg", seg, "chunk", chunk)
        got = gpu(decoder, data, num_threads=1, chunk_size_bytes=chunk, segment_bytes=seg)
        check(f"serial {chunk}", got, want_serial(data, chunk, False))
        assert [len(c) for c in got[0]] == crafted.serial_chunks(members, chunk, streaming=False)
    _say("stream serial", "seg", seg)
    got = gpu(decoder, data, fd=True, num_threads=1, segment_bytes=seg)
    check("stream serial", got, want_serial(data, 0, True))
    assert [len(c) for c in got[0]] == crafted.serial_chunks(members, 0, streaming=True)


def test_serial_large_multi_piece(decoder):
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(9 << 20, seed=21)
    g = synth.gzip(raw, 6, piece=4 << 20).tobytes()
    for chunk in (1 << 20, 4 << 20):
        _say("serial fastq9m", chunk)
        got = gpu(decoder, g, num_threads=1, chunk_size_bytes=chunk)
        check(f"fastq {chunk}", got, want_serial(g, chunk, False))
    _say("stream serial fastq9m")
    got = gpu(decoder, g, fd=True, num_threads=1)
    check("fastq stream", got, want_serial(g, 0, True))
    assert b"".join(got[0]) == raw.tobytes()


def _corrupt_set():
    data, payload, _ = crafted.synced_file(6, n_members=2)
    n = len(data)
    cases = [("good", data)]
    b = bytearray(data)
    b[-6] ^= 1
    cases.append(("bad_crc", bytes(b)))
    b = bytearray(data)
    b[-1] ^= 1
    cases.append(("bad_isize", bytes(b)))
    cases += [("cut_body", data[: n // 2]), ("cut_trailer", data[:-3]), ("junk_tail", data + b"junk"),
              ("bad_magic", b"\x1f\x8c" + data[2:]), ("tail_zero", data + b"\0" * 20)]
    rng = np.random.default_rng(17)
    for k in range(6):
        b = bytearray(data)
        for pos in rng.integers(30, n - 8, 2):
            b[pos] ^= int(rng.integers(1, 256))
        cases.append((f"corrupt{k}", bytes(b)))
    return cases


@pytest.mark.parametrize("stream", [False, True])
def test_serial_errors_match_oracle(decoder, stream):
    for name, data in _corrupt_set():
        for seg in (0, 4096):
            _say("serial errors", name, "stream", stream, "seg", seg)
            got = gpu(decoder, data, fd=stream, num_threads=1, chunk_size_bytes=65536, segment_bytes=seg)
            check(f"{name} seg={seg}", got, want_serial(data, 65536, stream))


def test_serial_sink_closed(decoder):
    data, payload, _ = crafted.synced_file(7, n_members=1)
    _say("serial sink closed")
    got = gpu(decoder, data, sink_limit=1, num_threads=1, chunk_size_bytes=65536)
    check("serial closed", got, want_serial(data, 65536, False, sink_limit=1))
    got = gpu(decoder, data, fd=True, sink_limit=2, num_threads=1)
    check("stream closed", got, want_serial(data, 0, True, sink_limit=2))
    assert got[1].chunks_decoded > 2  # the streaming serial path decodes on


@pytest.mark.parametrize("seg", [0, 2048, 4096, 8192])
def test_back_reference_before_member_start(decoder, seg):
    """Member 2's block Y reaches back into member 1 (see crafted.reach_before_member)."""
    f, pay, y_at, (la, lz, ly) = crafted.reach_before_member(80000, 15000, 12000, 3000)
    _say("reach chunk at Y", "seg", seg)
    # speculative path: a chunk starts at Y, its markers resolve from member 1
    for threads in (0, 4):
        got = gpu(decoder, f, num_threads=threads, chunk_size_bytes=65536, segment_bytes=seg)
        rc, ochunks, ost, oer = oracle.read_gz(f, chunk_size=65536, threads=4)
        check("parallel, chunk at Y", got, (rc, ochunks, ost, oer))
        assert b"".join(got[0]) == pay
    # serial paths: a member-local window rejects it
    got = gpu(decoder, f, num_threads=1, chunk_size_bytes=65536, segment_bytes=seg)
    check("serial", got, want_serial(f, 65536, False))
    assert [len(c) for c in got[0]] == [la]
    got = gpu(decoder, f, fd=True, num_threads=1, segment_bytes=seg)
    check("stream serial", got, want_serial(f, 0, True))
    # Y inside the first chunk: a later member of a chunk decodes on its own
    f, pay, y_at, (la, lz, ly) = crafted.reach_before_member(76000, 15000, 12000, 3000)
    _say("reach Y inside a chunk", "seg", seg)
    got = gpu(decoder, f, num_threads=4, chunk_size_bytes=65536, segment_bytes=seg)
    check("parallel, Y in chunk 0", got, oracle.read_gz(f, chunk_size=65536, threads=4))
    got = gpu(decoder, f, num_threads=1, chunk_size_bytes=65536, segment_bytes=seg)
    check("serial, Y in chunk 0", got, want_serial(f, 65536, False))


CONCAT = {
    # tests/concat.rs:84-265 (threads 1, 2, 4 there)
    "hello_world": [(b"Hello", 6), (b"World", 6)],
    "empty_interspersed": [(b"", 6), (b"alpha", 6), (b"", 9), (b"", 1), (b"beta", 6), (b"", 6)],
    "many_small": [(f"member-{i}-payload\n".encode(), 6) for i in range(500)],
}


def test_concat_all_thread_counts(decoder):
    from test_oracle import lcg_payload

    cases = dict(CONCAT)
    cases["mixed_levels"] = [(lcg_payload(1, 1), 6), (lcg_payload(100, 2), 1), (lcg_payload(65_537, 3), 9),
                             (lcg_payload(1, 4), 6), (lcg_payload(300_000, 5), 6), (lcg_payload(7, 6), 1),
                             (lcg_payload(1_000_000, 7), 9)]
    for name, parts in sorted(cases.items()):
        g = b"".join(pygzip.compress(p, compresslevel=lv, mtime=0) for p, lv in parts)
        for threads in (1, 2, 4):
            _say("concat", name, "threads", threads)
            got = gpu(decoder, g, num_threads=threads, chunk_size_bytes=65536)
            check(f"{name} t={threads}", got, oracle.read_gz(g, chunk_size=65536, threads=threads))
        got = gpu(decoder, g, fd=True, num_threads=1)
        check(f"{name} stream serial", got, want_serial(g, 0, True))


def test_serial_batch_equals_alone(decoder):
    import rusty_rapidgzip_amd as rrg

    files = [d for _, d in _corrupt_set()]
    files.append(crafted.reach_before_member(80000, 15000, 12000, 3000)[0])
    cfg = rrg.Config(num_threads=1, chunk_size_bytes=65536)
    _say("serial batch of", len(files))
    res = decoder.decode_batch(files, cfg)
    for i, (data, r) in enumerate(zip(files, res)):
        chunks, st, err = gpu(decoder, data, num_threads=1, chunk_size_bytes=65536)
        got = [r.data[a:b] for a, b in zip([0] + list(r.chunk_ends[:-1]), r.chunk_ends)]
        assert got == chunks, i
        assert (r.error is None) == (err is None), i
        if err is not None:
            assert (_kind(r.error), str(r.error)) == (_kind(err), str(err)), i
