"""GPU: batched decode and the serving loop give every file exactly the result
it gets alone.

"Alone" is the single-file C ABI (rrg_gpu_decode: bytes streamed per chunk,
DecodeStats, error), itself pinned to the oracle by test_gpu_parity.py; good
inputs are also checked against the oracle here.  Each file of a batch, and
each request of a server run, must reproduce the same chunk boundaries, bytes,
stats and error -- including corrupt files mixed in with good ones.
"""
import gzip as pygzip
import json
import os
import sys

import numpy as np
import pytest

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


def _inputs():
    from rusty_rapidgzip_amd import synth

    files = []
    raw = synth.fastq(3 << 20)
    files.append(("fastq3m", synth.gzip(raw, 6, piece=1 << 20).tobytes()))
    files.append(("text_l1", synth.gzip(synth.text(1 << 20), 1).tobytes()))
    files.append(("text_l9", synth.gzip(synth.text(1 << 20, seed=7), 9).tobytes()))
    mraw, lv, ends = synth.mixed(8 << 20)
    files.append(("mixed8m", synth.gzip_segments(mraw, lv, ends, piece=2 << 20).tobytes()))
    parts = [lcg_payload(n, 100 + n) for n in (0, 1, 5000, 300_000, 0, 70_000)]
    files.append(("multi", b"".join(gz(p, 6) for p in parts)))
    files.append(("tiny_members", b"".join(gz(bytes([65 + i % 26]) * (i % 7), 6) for i in range(300))))
    files.append(("nested", gz(gz(lcg_payload(1 << 20, 42), 6), 1)))
    bg = synth.bgzf(synth.fastq(2 << 20, seed=3), 6).tobytes()
    files.append(("bgzf", bg))
    files.append(("bgzf_rest", bg[:-28] + gz(lcg_payload(100_000, 9), 6)))
    good = gz(lcg_payload(300_000, 5), 6)
    b = bytearray(good)
    b[-6] ^= 1
    files.append(("bad_crc", bytes(b)))
    b = bytearray(good)
    b[-2] ^= 1
    files.append(("bad_isize", bytes(b)))
    files.append(("truncated", good[: len(good) // 2]))
    files.append(("junk_tail", good + b"junk"))
    files.append(("bad_magic", b"\x1f\x8c" + good[2:]))
    files.append(("empty", b""))
    rng = np.random.default_rng(11)
    for k in range(3):
        b = bytearray(good)
        for pos in rng.integers(20, len(b) - 8, 3):
            b[pos] ^= int(rng.integers(1, 256))
        files.append((f"corrupt{k}", bytes(b)))
    sums = json.load(open(os.path.join(GOLDEN, "corpus_sums.json")))
    for name in sorted(sums):
        files.append((name, open(os.path.join(GOLDEN, "corpus", name), "rb").read()))
    return files


CHUNK = 256 << 10


def _alone(decoder, data, cfg):
    """(chunks, stats or None, (error type, kind, message) or None)"""
    import rusty_rapidgzip_amd as rrg

    out = []
    try:
        st = decoder.decode_bytes(data, out, cfg)
        return out, st, None
    except rrg.RapidgzipError as e:
        return out, None, (type(e).__name__, getattr(e, "kind", None), str(e))


def _split(data, ends):
    prev, out = 0, []
    for e in ends:
        out.append(data[prev:e])
        prev = e
    return out


@pytest.fixture(scope="module")
def cases(decoder):
    import rusty_rapidgzip_amd as rrg

    cfg = rrg.Config(chunk_size_bytes=CHUNK)
    files = _inputs()
    alone = []
    for name, data in files:
        a = _alone(decoder, data, cfg)
        if a[2] is None:  # good inputs: the oracle agrees
            rc, ochunks, ost, _ = oracle.read_gz(data, chunk_size=CHUNK, threads=4)
            assert rc == 0, name
            assert [bytes(c) for c in ochunks] == a[0], name
            assert ost.speculation_failures == a[1].speculation_failures, name
        alone.append(a)
    return cfg, files, alone


def _check(name, want, chunks, stats, err):
    wchunks, wstats, werr = want
    assert [len(c) for c in chunks] == [len(c) for c in wchunks], name
    assert chunks == wchunks, name
    if werr is None:
        assert err is None, (name, err)
        assert stats.uncompressed_bytes == wstats.uncompressed_bytes, name
        assert stats.chunks_decoded == wstats.chunks_decoded, name
        assert stats.speculation_failures == wstats.speculation_failures, name
    else:
        assert err is not None, name
        assert (type(err).__name__, getattr(err, "kind", None), str(err)) == werr, name


def test_batch_equals_alone(decoder, cases):
    cfg, files, alone = cases
    _say("batch of", len(files))
    res = decoder.decode_batch([d for _, d in files], cfg)
    for (name, _), want, r in zip(files, alone, res):
        _check(name, want, _split(r.data, r.chunk_ends), r.stats, r.error)
    # twice in a row on the same context, in another order, with duplicates
    order = list(range(len(files)))[::-1] + [0, 3, 7]
    res = decoder.decode_batch([files[i][1] for i in order], cfg)
    for i, r in zip(order, res):
        _check(files[i][0], alone[i], _split(r.data, r.chunk_ends), r.stats, r.error)


def test_batch_graph_replays(decoder, cases):
    """Tail, resolve and CRC as one graph launch, batch after batch: the
    instantiated graph is updated in place when the next batch launches the
    same kernels and instantiated anew when it does not (a BGZF-only batch has
    no tail chain), with eager single-file decodes and an eager batch in
    between.  Every batch gives each file its result alone."""
    import rusty_rapidgzip_amd as rrg

    _, files, alone = cases
    cfg = rrg.Config(chunk_size_bytes=CHUNK, launch_graph=True)
    eager = rrg.Config(chunk_size_bytes=CHUNK, launch_graph=False)
    everything = list(range(len(files)))
    bgzf_only = [i for i, (n, _) in enumerate(files) if n == "bgzf"] * 2
    subsets = [everything, bgzf_only, [0, 1, 2], everything[::-1], bgzf_only, [3, 6, 0]]
    for rep in range(2):
        for sub in subsets:
            _say("graph batch", rep, len(sub))
            res = decoder.decode_batch([files[i][1] for i in sub], cfg)
            assert decoder.profile().n_graph_launches == 1
            for i, r in zip(sub, res):
                _check(files[i][0], alone[i], _split(r.data, r.chunk_ends), r.stats, r.error)
            # eager work on the same context in between
            k = sub[-1]
            got = _alone(decoder, files[k][1], cfg)
            assert got[0] == alone[k][0] and got[2] == alone[k][2]
            res = decoder.decode_batch([files[i][1] for i in sub[:3]], eager)
            assert decoder.profile().n_graph_launches == 0
            for i, r in zip(sub[:3], res):
                _check(files[i][0], alone[i], _split(r.data, r.chunk_ends), r.stats, r.error)


def test_batch_device_inputs(decoder, cases):
    from rusty_rapidgzip_amd import hipmem

    cfg, files, alone = cases
    pick = [i for i, (_, d) in enumerate(files) if len(d)][:8]
    bufs = [hipmem.DeviceBuffer.from_host(files[i][1]) for i in pick]
    try:
        # device inputs without host copies: the headers are fetched from the device
        res = decoder.decode_batch([None] * len(pick), cfg, device_ptrs=[b.ptr for b in bufs],
                                   device_lens=[len(files[i][1]) for i in pick])
        for i, r in zip(pick, res):
            _check(files[i][0], alone[i], _split(r.data, r.chunk_ends), r.stats, r.error)
    finally:
        for b in bufs:
            b.free()


def test_server_mix_equals_alone(cases):
    import rusty_rapidgzip_amd as rrg

    cfg, files, alone = cases
    order = list(range(len(files))) * 2
    np.random.default_rng(5).shuffle(order)
    with rrg.Server(0, cfg, hbm_budget_bytes=256 << 20, max_batch_items=9) as srv:
        ids = {}
        for k, i in enumerate(order):
            ids[srv.submit(files[i][1], tag=k)] = i
        done, batches, mixed = 0, set(), 0
        while done < len(order):
            for c in srv.poll(16, timeout_ms=60000):
                i = ids[c.id]
                data = srv.result(c.id, c.out_len) if c.out_len else b""
                ends = srv.chunk_ends(c.id)
                _check(files[i][0], alone[i], _split(data, ends), c.stats, c.error)
                batches.add(c.batch)
                mixed += c.batch_items > 1
                assert c.ms_total >= c.ms_queue >= 0
                srv.release(c.id)
                done += 1
        _say("server batches", len(batches), "requests in shared batches", mixed)
        assert len(batches) >= 2
        assert mixed > 0
        assert srv.pending() == 0


def test_server_admission_waits_for_release(cases):
    """With a budget smaller than one request, each request is admitted only
    once the previous output has been released (the InflightCap analogue)."""
    import rusty_rapidgzip_amd as rrg

    cfg, files, alone = cases
    pick = [0, 1, 4, 7]
    with rrg.Server(0, cfg, hbm_budget_bytes=1) as srv:
        ids = {srv.submit(files[i][1]): i for i in pick}
        seen = 0
        while seen < len(pick):
            got = srv.poll(4, timeout_ms=60000)
            assert len(got) == 1  # nothing else can be admitted while this output is held
            c = got[0]
            assert c.batch_items == 1
            i = ids[c.id]
            data = srv.result(c.id, c.out_len)
            _check(files[i][0], alone[i], _split(data, srv.chunk_ends(c.id)), c.stats, c.error)
            assert srv.poll(4, timeout_ms=200) == []  # held output blocks the next admission
            srv.release(c.id)
            seen += 1


def test_server_release_before_poll(cases):
    """A completion released before it was polled is dropped, not returned."""
    import time

    import rusty_rapidgzip_amd as rrg

    cfg, files, alone = cases
    with rrg.Server(0, cfg) as srv:
        rid = srv.submit(files[0][1])
        deadline = time.time() + 60
        while not srv.chunk_ends(rid) and time.time() < deadline:
            time.sleep(0.01)
        assert srv.chunk_ends(rid), "request did not complete"
        srv.release(rid)
        assert srv.poll(4, timeout_ms=100) == []
        rid2 = srv.submit(files[1][1])
        got = srv.poll(4, timeout_ms=60000)
        assert [c.id for c in got] == [rid2]
        c = got[0]
        _check(files[1][0], alone[1], _split(srv.result(rid2, c.out_len), srv.chunk_ends(rid2)), c.stats, c.error)
        srv.release(rid2)


def test_server_recovers_from_allocation_failure(cases):
    """A batch whose device allocation fails reports an error for its own
    requests only; later requests still decode to their solo results."""
    import rusty_rapidgzip_amd as rrg

    cfg, files, alone = cases
    with rrg.Server(0, cfg) as srv:
        rrg.debug_fail_alloc(3)  # an allocation of the first batch's decode fails
        try:
            rid = srv.submit(files[0][1])
            got = srv.poll(4, timeout_ms=60000)
        finally:
            rrg.debug_fail_alloc(0)
        assert [c.id for c in got] == [rid]
        assert isinstance(got[0].error, rrg.DeviceError), got[0].error
        srv.release(rid)
        for i in (0, 3, 7):
            r = srv.submit(files[i][1])
            c = srv.poll(4, timeout_ms=60000)[0]
            assert c.id == r
            _check(files[i][0], alone[i], _split(srv.result(r, c.out_len) if c.out_len else b"",
                                                 srv.chunk_ends(r)), c.stats, c.error)
            srv.release(r)
