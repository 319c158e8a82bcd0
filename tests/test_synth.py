"""The workload generators write valid gzip: the mixed-entropy writer once
switched levels with deflateParams, which crashed inside zlib 1.2.11's deflate
on some seeds (seed 4 at 256 MiB); it now ends the stream with a sync flush
and starts a new one, primed with the 32 KiB before, at each level change."""
import zlib

import pytest

from rusty_rapidgzip_amd import synth


@pytest.mark.parametrize("seed", [4, 5])
def test_mixed_writer_valid(seed):
    raw, lv, ends = synth.mixed(256 << 20, seed=seed, threads=8)
    gz = synth.gzip_segments(raw, lv, ends, threads=8, piece=64 << 20)
    assert (lv == 0).any() and ((lv == 9) | (lv == 1)).any()  # stored and compressed segments
    out = zlib.decompressobj(31).decompress(gz.tobytes())
    assert out == raw.tobytes()
