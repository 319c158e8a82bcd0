"""Benchmark: decompressed GB/s of the MI355X gzip decoder on one GPU.

Workload (the headline configuration): 4 GiB of synthetic FASTQ-like text
(4-line records '@r<id>', 150 bp ACGT with ~0.1% N, '+', 150-char Phred walk),
gzip level 6, one member, read_gz chunk size 4 MiB.  One step = one complete
device-resident decode of the file: block finder, speculative inflate, chain
verification, tail chain, marker resolution, CRC32 + ISIZE checks, output left
in HBM with the reference's chunk boundaries.  The compressed input is already
resident in HBM when the timed region starts.

Prints ONE JSON line.  `cpu_baseline` times the CPU restatement of read_gz
(oracle/, "port") on this host's cores over a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--gib", type=float, default=4.0, help="uncompressed size of the workload")
    ap.add_argument("--chunk-mib", type=int, default=4)
    ap.add_argument("--segment-kib", type=int, default=0, help="0 = automatic")
    ap.add_argument("--lut-bits", type=int, default=9)
    ap.add_argument("--lanes", type=int, default=16)
    ap.add_argument("--cpu-sample-mib", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build targets one GPU; --gpus must be 1"}))
        log("only --gpus 1 is supported: the decoder runs on a single MI355X")
        return 2

    import numpy as np

    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

    threads = min(16, os.cpu_count() or 1)
    n = int(args.gib * (1 << 30))
    t0 = time.time()
    raw = synth.fastq(n, threads=threads)
    log(f"generated {n} bytes of FASTQ-like text in {time.time() - t0:.1f}s")
    t0 = time.time()
    gz = synth.gzip(raw, 6, threads=threads, piece=64 << 20)
    C = int(gz.size)
    log(f"gzip -6: {C} bytes ({n / C:.2f}:1) in {time.time() - t0:.1f}s")
    d_in = hipmem.DeviceBuffer.from_host(gz)
    hipmem.synchronize()
    cfg = rrg.Config(chunk_size_bytes=args.chunk_mib << 20, segment_bytes=args.segment_kib << 10,
                     lut_bits=args.lut_bits, lanes_per_block=args.lanes)
    dec = rrg.Decoder(0)
    host = gz  # numpy array, passed by pointer

    def step():
        return dec.decode_device(d_in.ptr, C, host, cfg)

    # warmup (first call also checks the output bytes against the generator)
    for w in range(max(1, args.warmup)):
        ptr, out_len, ends, st = step()
        if w == 0:
            assert out_len == n, (out_len, n)
            check = hipmem.to_host(ptr, n)
            if not np.array_equal(check, raw):
                raise RuntimeError("GPU output differs from the generated bytes")
            del check
            log(f"warmup decode verified: {out_len} bytes, {len(ends)} chunks, "
                f"{st.speculation_failures} speculation failures")
        p = dec.profile()
        log("warmup", w, {k: round(v, 3) if isinstance(v, float) else v for k, v in p.as_dict().items()})

    lat = []
    prof = []
    hipmem.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        step()
        lat.append((time.perf_counter() - t1) * 1e3)
        prof.append(dec.profile().as_dict())
    hipmem.synchronize()
    elapsed = time.perf_counter() - t_start
    value = n * args.steps / elapsed / 1e9
    ms_per_step = elapsed * 1e3 / args.steps

    def avg(key):
        return sum(p[key] for p in prof) / len(prof)

    ms_inflate = avg("ms_inflate")
    launches = avg("n_inflate_launches")
    segs = avg("n_segments")
    # dominant kernel: speculative inflate.  Algorithmic bytes per launch: the
    # compressed stream read once plus the u16 cells written (C + 2U).
    infl_bytes = C + 2 * n
    infl_ms_launch = ms_inflate / max(1.0, launches)
    achieved = infl_bytes / max(1, launches) / (infl_ms_launch / 1e3) / 1e9

    result = {
        "metric": "decompressed GB/s on 1 MI355X (device-resident read_gz, CRC-checked)",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"fastq-like {args.gib:g} GiB, gzip -6 single member, 4-line records, 150 bp",
            "chunk_size_bytes": args.chunk_mib << 20,
            "compressed_bytes": C,
            "uncompressed_bytes": n,
            "segments": int(segs),
            "lut_bits": args.lut_bits,
            "lanes_per_block": args.lanes,
            "input": "resident in HBM",
        },
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "stages_ms": {k: round(avg(k), 3) for k in ("ms_find", "ms_inflate", "ms_tail", "ms_resolve", "ms_crc")},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_inflate",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
            "end_to_end_frac": round((C + n) * args.steps / elapsed / 1e9 / HBM_PEAK_GBS, 5),
        },
        "cpu_baseline": None,
    }
    d_in.free()

    if not args.no_cpu:
        from oracle import oracle

        sn = min(n, args.cpu_sample_mib << 20)
        sraw = synth.fastq(sn, threads=threads)
        sgz = synth.gzip(sraw, 6, threads=threads, piece=64 << 20)
        cores = threads
        times = []
        for _ in range(3):
            t1 = time.perf_counter()
            rc, _, st, _ = oracle.read_gz(sgz, chunk_size=args.chunk_mib << 20, threads=cores, keep=False)
            times.append(time.perf_counter() - t1)
            if rc != 0 or st.uncompressed_bytes != sn:
                raise RuntimeError("CPU baseline decode failed")
        t_med = sorted(times)[len(times) // 2]
        result["cpu_baseline"] = {
            "value": round(sn / t_med / 1e9, 3),
            "unit": "GB/s",
            "cores": cores,
            "kind": "port",
            "sample": f"{sn >> 20} MiB of the same FASTQ-like generator, gzip -6, same chunking; median of 3",
        }
        log(f"cpu baseline {result['cpu_baseline']}")

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
