"""GPU: the command-line drop-in (rapidgzip-gpu, main.rs:12-105 of the
reference's binary) writes exactly what `gzip -dc` writes, from a file path
and from a pipe on stdin, and fails with a message on corrupt input."""
import gzip as pygzip
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cli():
    from rusty_rapidgzip_amd import _native

    return _native.CLI_BIN


def _data():
    from rusty_rapidgzip_amd import synth

    raw = synth.fastq(5 << 20, seed=77).tobytes()
    rng = np.random.default_rng(3)
    parts = [(rng.integers(0, 60, n, dtype=np.uint8) + 40).tobytes() for n in (100_000, 7, 600_000)]
    multi = b"".join(pygzip.compress(p, 6, mtime=0) for p in parts)
    return [("fastq", synth.gzip(np.frombuffer(raw, np.uint8), 6, piece=1 << 20).tobytes(), raw),
            ("multi", multi, b"".join(parts))]


@pytest.mark.parametrize("args", [[], ["-P", "1"], ["--chunk-size", "65536", "-v"]])
def test_cli_file_and_stdin(tmp_path, args):
    for name, gz, raw in _data():
        f = tmp_path / f"{name}.gz"
        f.write_bytes(gz)
        _say("cli", name, args)
        want = subprocess.run(["gzip", "-dc", str(f)], capture_output=True, check=True).stdout
        assert want == raw
        r = subprocess.run([_cli(), *args, str(f)], capture_output=True, timeout=120)
        assert r.returncode == 0, r.stderr.decode()
        assert r.stdout == want
        if "-v" in args:
            assert b"[rapidgzip]" in r.stderr
        # `-` reads stdin; a pipe takes the streaming path
        p = subprocess.Popen([_cli(), *args, "-"], stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                             stderr=subprocess.PIPE)
        out, err = p.communicate(gz, timeout=120)
        assert p.returncode == 0, err.decode()
        assert out == want
        # a redirected regular file on stdin is read as a file
        with open(f, "rb") as fin:
            r = subprocess.run([_cli(), *args, "-"], stdin=fin, capture_output=True, timeout=120)
        assert r.returncode == 0 and r.stdout == want


def test_cli_corrupt_input(tmp_path):
    name, gz, raw = _data()[1]
    bad = bytearray(gz)
    bad[-6] ^= 1  # the last member's CRC
    f = tmp_path / "bad.gz"
    f.write_bytes(bytes(bad))
    r = subprocess.run([_cli(), str(f)], capture_output=True, timeout=120)
    assert r.returncode == 1
    assert b"CRC32 mismatch" in r.stderr
    r = subprocess.run([_cli(), str(tmp_path / "missing.gz")], capture_output=True, timeout=120)
    assert r.returncode == 1 and b"cannot open" in r.stderr
