"""GPU: the no-allocation device decode (rrg_gpu_decode_device_noalloc) gives
what rrg_gpu_decode_device gives -- bytes, chunk ends, stats, errors -- on
consecutive calls (the second with a smaller input) through the C ABI."""
import gzip as pygzip
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs():
    from rusty_rapidgzip_amd import synth

    fastq = synth.gzip(synth.fastq(6 << 20, seed=21), 6, piece=2 << 20).tobytes()
    text = synth.gzip(synth.text(2 << 20, seed=5), 9).tobytes()
    rng = np.random.default_rng(8)
    parts = [(rng.integers(0, 40, n, dtype=np.uint8) + 60).tobytes() for n in (70_000, 1, 300_000)]
    multi = b"".join(pygzip.compress(p, 6, mtime=0) for p in parts)
    bgzf = synth.bgzf(synth.fastq(1 << 20, seed=4), 6).tobytes()
    return [("fastq6m", fastq), ("text2m_l9", text), ("multi", multi), ("bgzf1m", bgzf)]


def _plain(decoder, buf, data, cfg):
    from rusty_rapidgzip_amd import hipmem

    d_out, n, ends, st = decoder.decode_device(buf.ptr, len(data), data, cfg)
    return hipmem.to_host(d_out, n).tobytes(), ends, st


def _noalloc(decoder, buf, data, cfg, out):
    from rusty_rapidgzip_amd import hipmem

    n, ends, st = decoder.decode_device_noalloc(buf.ptr, len(data), out.ptr, out.nbytes, data, cfg)
    return hipmem.to_host(out.ptr, n).tobytes(), ends, st


def test_noalloc_equals_device_decode(decoder):
    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem

    cfg = rrg.Config(chunk_size_bytes=1 << 20)
    files = _inputs()
    # largest first, so that every later call runs on a smaller input than the one before
    files.sort(key=lambda f: -len(f[1]))
    out = hipmem.DeviceBuffer(16 << 20)
    try:
        for rep in range(2):
            for name, data in files:
                _say("noalloc", rep, name, len(data))
                buf = hipmem.DeviceBuffer.from_host(data)
                try:
                    want = _plain(decoder, buf, data, cfg)
                    got = _noalloc(decoder, buf, data, cfg, out)
                finally:
                    buf.free()
                assert got[0] == want[0], name
                assert got[1] == want[1], name
                assert got[2] == want[2], name
    finally:
        out.free()


def test_noalloc_errors(decoder):
    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

    cfg = rrg.Config(chunk_size_bytes=1 << 20)
    data = synth.gzip(synth.text(1 << 20, seed=9), 6).tobytes()
    buf = hipmem.DeviceBuffer.from_host(data)
    out = hipmem.DeviceBuffer(4 << 20)
    try:
        # an output buffer smaller than the decompressed size
        with pytest.raises(rrg.DeviceError, match="too small"):
            decoder.decode_device_noalloc(buf.ptr, len(data), out.ptr, 1 << 19, data, cfg)
        # a corrupt trailer: the same error as the allocating decode
        bad = bytearray(data)
        bad[-6] ^= 1
        bad = bytes(bad)
        bb = hipmem.DeviceBuffer.from_host(bad)
        try:
            with pytest.raises(rrg.GzipError) as e1:
                decoder.decode_device(bb.ptr, len(bad), bad, cfg)
            with pytest.raises(rrg.GzipError) as e2:
                decoder.decode_device_noalloc(bb.ptr, len(bad), out.ptr, out.nbytes, bad, cfg)
            assert e1.value.kind == e2.value.kind == "CrcMismatch"
        finally:
            bb.free()
        # and the next call on the same workspace is correct again
        got = _noalloc(decoder, buf, data, cfg, out)
        assert got[0] == synth.text(1 << 20, seed=9).tobytes()
    finally:
        buf.free()
        out.free()
    # an input over the limit is refused before anything runs
    big = hipmem.DeviceBuffer(9 << 20)
    try:
        with pytest.raises(rrg.DeviceError, match="limit"):
            decoder.decode_device_noalloc(big.ptr, 9 << 20, big.ptr, big.nbytes, None, cfg)
    finally:
        big.free()
