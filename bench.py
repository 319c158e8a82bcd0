"""Benchmark: decompressed GB/s of the MI355X gzip decoder on one GPU.

Default workload (the headline configuration): 4 GiB of synthetic FASTQ-like
text (4-line records '@r<id>', 150 bp ACGT with ~0.1% N, '+', 150-char Phred
walk), gzip level 6, one member, read_gz chunk size 4 MiB.  One step = one
complete device-resident decode of the file: block finder, inflate (Huffman
tokens, then expansion), chain verification, tail chain, marker resolution,
CRC32 + ISIZE checks, output left in HBM with the reference's chunk
boundaries.  The compressed input is already resident in HBM when the timed
region starts.

Prints ONE JSON line.  `cpu_baseline` times the CPU restatement of read_gz
(oracle/, "port") on this host's cores over the whole workload, with zlib and
`gzip -dc` single-thread anchors.  `roofline` is for the dominant kernel (the
longer of inflate pass 1, k_tokens_wave, and pass 2, k_expand, by their HIP
event times in this run); `roofline.traffic` is its HBM traffic per launch
measured with rocprofv3 counters (profiles/traffic_fastq4g.json), when
present.  `other_configs` holds the BGZF, mixed-entropy and 64 MiB text
configurations of the same run (each output-verified against its generator).

Other workloads (--workload), each one JSON line of its own:
  bgzf    the same 4 GiB as BGZF (65280-byte members, level 6), one batch
  members the same 4 GiB as pigz-style 1 MiB gzip members (plain headers)
  mixed   2 GiB mixed-entropy stress in one member (stored random segments,
          runs, repeats, text at levels 1/9)
  batch   64 FASTQ files of 64 MiB, one batched call vs one call per file
  serve   request mix through the serving loop (sizes log-uniform, text /
          FASTQ / mixed at levels 1/6/9), closed-loop capacity, then Poisson
          arrivals at fractions of it: throughput and p50/p99 latency
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_fastq4g.json")
STAGES = ("ms_find", "ms_inflate", "ms_expand", "ms_tail", "ms_resolve", "ms_crc")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _pct(v, q):
    import numpy as np

    return round(float(np.percentile(v, q)), 3) if len(v) else None


def _base(metric, value, unit, steps, warmup, ms_per_step, dtype="u8"):
    return {
        "metric": metric,
        "value": round(value, 3),
        "unit": unit,
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
    }


def make_cfg(args):
    import rusty_rapidgzip_amd as rrg

    return rrg.Config(chunk_size_bytes=args.chunk_mib << 20, segment_bytes=args.segment_kib << 10,
                      lut_bits=args.lut_bits, lanes_per_block=args.lanes)


def timed_steps(step, steps):
    from rusty_rapidgzip_amd import hipmem

    lat = []
    hipmem.synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        step()
        lat.append((time.perf_counter() - t1) * 1e3)
    hipmem.synchronize()
    return time.perf_counter() - t_start, lat


# ---------------------------------------------------------------------------
def run_fastq(args, threads):
    import numpy as np

    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

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
    cfg = make_cfg(args)
    dec = rrg.Decoder(0)
    tuned_ms = None
    if args.autotune:
        # untimed: the autotuner's pick replaces the command-line decomposition
        cfg, tuned_ms = dec.autotune(d_in.ptr, C, gz, cfg, reps=1)
        log(f"autotuned: lanes {cfg.lanes_per_block}, lut {cfg.lut_bits}, segment {cfg.segment_bytes} "
            f"({tuned_ms:.1f} ms)")

    def step():
        return dec.decode_device(d_in.ptr, C, gz, cfg)

    for w in range(max(1, args.warmup)):
        ptr, out_len, ends, st = step()
        if w == 0:
            assert out_len == n, (out_len, n)
            if not np.array_equal(hipmem.to_host(ptr, n), raw):
                raise RuntimeError("GPU output differs from the generated bytes")
            log(f"warmup decode verified: {out_len} bytes, {len(ends)} chunks, "
                f"{st.speculation_failures} speculation failures")
        log("warmup", w, {k: round(v, 3) if isinstance(v, float) else v for k, v in dec.profile().as_dict().items()})
    prof = []

    def step_p():
        step()
        prof.append(dec.profile().as_dict())

    elapsed, lat = timed_steps(step_p, args.steps)
    value = n * args.steps / elapsed / 1e9

    def avg(key):
        return sum(p[key] for p in prof) / len(prof)

    # Algorithmic bytes per launch.  Inflate pass 1 (k_tokens_wave, or k_tokens
    # with one lane per segment): the compressed stream read once plus the
    # 32-bit tokens written, C + 4 * tokens.  Pass 2 (k_expand): the tokens read
    # and the u16 cells written, 4 * tokens + 2U.  The roofline object is for
    # the one that takes longer (the dominant kernel).
    ms_inflate = avg("ms_inflate")
    launches = max(1.0, avg("n_inflate_launches"))
    n_tok = avg("n_tokens")
    lut_eff = cfg.lut_bits or (10 if cfg.lanes_per_block >= 64 else 9)
    tok_kernel = "k_tokens_wave" if cfg.lanes_per_block >= 64 else "k_tokens"
    kernels = {
        tok_kernel: ((C + 4 * n_tok) / launches, ms_inflate / launches),
        "k_expand": (4 * n_tok + 2 * n, avg("ms_expand")),
    }
    dom = max(kernels, key=lambda k: kernels[k][1])
    dom_bytes, dom_ms = kernels[dom]
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        t = json.load(open(TRAFFIC_FILE))
        # counters measured on this workload with this launch geometry
        if (t.get("compressed_bytes") == C and t.get("uncompressed_bytes") == n
                and t.get("lut_bits") == lut_eff and t.get("lanes_per_block") == cfg.lanes_per_block
                and not cfg.segment_bytes):
            traffic = t.get("kernels", {}).get(dom, {}).get("bytes_per_launch")
    result = _base("decompressed GB/s on 1 MI355X (device-resident read_gz, CRC-checked)", value, "GB/s",
                   args.steps, args.warmup, elapsed * 1e3 / args.steps)
    result["config"] = {
        "workload": f"fastq-like {args.gib:g} GiB, gzip -6 single member, 4-line records, 150 bp "
                    "(written by zlib in 64 MiB pieces joined by sync flushes)",
        "chunk_size_bytes": args.chunk_mib << 20,
        "compressed_bytes": C,
        "uncompressed_bytes": n,
        "segments": int(avg("n_segments")),
        "tokens": int(n_tok),
        "lut_bits": lut_eff,
        "lanes_per_block": cfg.lanes_per_block,
        "segment_bytes": cfg.segment_bytes or "automatic",
        "autotuned": tuned_ms is not None,
        "input": "resident in HBM",
    }
    result["latency_ms"] = {"p50": _pct(lat, 50), "p99": _pct(lat, 99)}
    result["stages_ms"] = {k: round(avg(k), 3) for k in STAGES}
    result["finder"] = {
        "candidate_bit_offsets_per_s": round(avg("find_bits") / (avg("ms_find") / 1e3), 1),
        "scanned_bits": int(avg("find_bits")),
        "how": "bit offsets from each target to its find (or scan end), per second of the find stage",
    }
    others = {k: round(b / (ms / 1e3) / 1e9, 1) for k, (b, ms) in kernels.items() if k != dom}
    others.update({
        "k_resolve": round(3 * n / (avg("ms_resolve") / 1e3) / 1e9, 1),
        "k_crc_pieces": round(n / (avg("ms_crc") / 1e3) / 1e9, 1),
        "unit": f"GB/s algorithmic ({tok_kernel} C + 4*tokens, k_expand 4*tokens + 2U, resolve 3U, CRC U)",
    })
    result["roofline"] = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": int(dom_bytes),
        "end_to_end_frac": round((C + n) * args.steps / elapsed / 1e9 / HBM_PEAK_GBS, 5),
        "other_kernels": others,
    }
    result["cpu_baseline"] = None
    d_in.free()
    raw_keep = raw
    if not args.no_e2e:
        # the same decode through the host-buffer entry point (rrg_gpu_decode):
        # the file in host memory, decoded in slabs, each chunk copied to the
        # host and handed to a C callback that copies it into one reused
        # buffer; never the value
        from rusty_rapidgzip_amd import synth

        runs = []
        for _ in range(3):
            sink = synth.CopySink()
            t1 = time.perf_counter()
            st_h = dec.decode_bytes_native(gz, synth.CopySink.callback(), ctypes.addressof(sink), cfg)
            dt = time.perf_counter() - t1
            assert sink.bytes == n == st_h.uncompressed_bytes and sink.chunks == st_h.chunks_decoded
            p = dec.profile()
            runs.append((n / dt / 1e9, p.ms_first_chunk, p.n_slabs, p.device_bytes_peak))
        runs.sort()
        med = runs[1]
        counted = []
        for _ in range(3):
            sink = synth.CopySink(4096)
            t1 = time.perf_counter()
            dec.decode_bytes_native(gz, synth.CopySink.callback(copy=False), ctypes.addressof(sink), cfg)
            counted.append(n / (time.perf_counter() - t1) / 1e9)
            assert sink.bytes == n
        result["host_to_host"] = {"value": round(med[0], 3), "unit": "GB/s", "ms_first_chunk": round(med[1], 2),
                                  "slabs": int(med[2]), "device_gib_peak": round(med[3] / 2**30, 2),
                                  "delivery_only": round(sorted(counted)[1], 3),
                                  "how": "rrg_gpu_decode on the file in host memory: slabs decoded while earlier "
                                         "slabs' chunks go to the host (pinned staging) and to a C callback that "
                                         "copies each into one reused 8 MiB buffer; median of 3.  delivery_only: "
                                         "the same with a callback that only counts the bytes"}
        log(f"host to host {result['host_to_host']}")

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args, threads, raw_keep, gz, n)
    result["_raw"] = raw_keep
    return result


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, threads, raw, gz, n):
    """The CPU restatement of read_gz (oracle/, "port") on the whole workload,
    all host threads, median of 5; plus two single-thread anchors on a
    bounded sample: zlib inflate (Python's zlib) and `gzip -dc > /dev/null`."""
    import subprocess
    import tempfile
    import zlib

    from oracle import oracle

    times = []
    for _ in range(5):
        t1 = time.perf_counter()
        rc, _, st, _ = oracle.read_gz(gz, chunk_size=args.chunk_mib << 20, threads=threads, keep=False)
        times.append(time.perf_counter() - t1)
        if rc != 0 or st.uncompressed_bytes != n:
            raise RuntimeError("CPU baseline decode failed")
    t_med = sorted(times)[len(times) // 2]
    # anchors on the first 256 MiB of output (a gzip file of its own)
    from rusty_rapidgzip_amd import synth

    an = min(n, 256 << 20)
    agz = synth.gzip(raw[:an], 6, threads=threads, piece=64 << 20).tobytes()
    t1 = time.perf_counter()
    d = zlib.decompressobj(31)
    out = 0
    for k in range(0, len(agz), 1 << 22):
        out += len(d.decompress(agz[k:k + (1 << 22)]))
    out += len(d.flush())
    z_gbs = an / (time.perf_counter() - t1) / 1e9
    assert out == an
    g_gbs = None
    with tempfile.NamedTemporaryFile(suffix=".gz", delete=False) as f:
        f.write(agz)
        path = f.name
    try:
        t1 = time.perf_counter()
        r = subprocess.run(["gzip", "-dc", path], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        if r.returncode == 0:
            g_gbs = an / (time.perf_counter() - t1) / 1e9
    except OSError:
        pass
    finally:
        os.unlink(path)
    res = {
        "value": round(n / t_med / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the whole workload ({n >> 20} MiB out, gzip -6, same chunking); median of 5 "
                  f"({', '.join(f'{t:.3f}' for t in times)} s)",
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "anchors": {"zlib_1_thread_GBps": round(z_gbs, 3),
                    "gzip_dc_GBps": round(g_gbs, 3) if g_gbs else None,
                    "anchor_sample": f"first {an >> 20} MiB of the workload as its own gzip -6 file"},
    }
    log(f"cpu baseline {res}")
    return res


def _sub(result):
    """The fields of a workload's result kept in the main line's other_configs."""
    keep = ("metric", "value", "unit", "ms_per_step", "config", "latency_ms", "stages_ms", "inflate_launches",
            "cpu_read_gz", "verified")
    return {k: result[k] for k in keep if k in result}


def run_text64m(args, threads):
    """Config 1: 64 MiB of order-2 Markov text, gzip -6, one member: GPU
    decode and the CPU restatement of read_gz (all threads and one)."""
    import numpy as np

    import rusty_rapidgzip_amd as rrg
    from oracle import oracle
    from rusty_rapidgzip_amd import hipmem, synth

    n = 64 << 20
    raw = synth.text(n, threads=threads)
    gz = synth.gzip(raw, 6, threads=threads, piece=64 << 20)
    C = int(gz.size)
    d_in = hipmem.DeviceBuffer.from_host(gz)
    cfg = rrg.Config()
    dec = rrg.Decoder(0)

    def step():
        return dec.decode_device(d_in.ptr, C, gz, cfg)

    ptr, out_len, ends, st = step()
    ok = out_len == n and np.array_equal(hipmem.to_host(ptr, n), raw)
    if not ok:
        raise RuntimeError("text64m: GPU output differs from the generated bytes")
    prof = []

    def step_p():
        step()
        prof.append(dec.profile().as_dict())

    steps = max(3, args.steps)
    elapsed, lat = timed_steps(step_p, steps)
    result = _base("decompressed GB/s on 1 MI355X (device-resident read_gz, text)", n * steps / elapsed / 1e9,
                   "GB/s", steps, 1, elapsed * 1e3 / steps)
    result["config"] = {"workload": "64 MiB order-2 Markov text (64 symbols, Dirichlet 0.3 rows), gzip -6, one member",
                        "compressed_bytes": C, "uncompressed_bytes": n, "chunks": len(ends)}
    result["latency_ms"] = {"p50": _pct(lat, 50), "p99": _pct(lat, 99)}
    result["stages_ms"] = {k: round(sum(p[k] for p in prof) / len(prof), 3) for k in STAGES}
    result["verified"] = True
    cpu = {}
    for th in (() if args.no_cpu else (threads, 1)):
        ts = []
        for _ in range(5):
            t1 = time.perf_counter()
            rc, _, cst, _ = oracle.read_gz(gz, chunk_size=4 << 20, threads=th, keep=False)
            ts.append(time.perf_counter() - t1)
            if rc != 0 or cst.uncompressed_bytes != n:
                raise RuntimeError("text64m: CPU read_gz failed")
        cpu[f"threads_{th}"] = round(n / sorted(ts)[2] / 1e9, 3)
    if cpu:
        result["cpu_read_gz"] = {**cpu, "unit": "GB/s", "kind": "port", "how": "median of 5"}
    d_in.free()
    return result


# ---------------------------------------------------------------------------
def run_bgzf(args, threads, raw=None):
    import numpy as np

    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

    n = int(args.gib * (1 << 30))
    if raw is None:
        raw = synth.fastq(n, threads=threads)
    t0 = time.time()
    bg = synth.bgzf(raw, 6, threads=threads)
    C = int(bg.size)
    log(f"BGZF level 6: {C} bytes in {time.time() - t0:.1f}s")
    d_in = hipmem.DeviceBuffer.from_host(bg)
    cfg = make_cfg(args)
    dec = rrg.Decoder(0)

    def step():
        return dec.decode_device(d_in.ptr, C, bg, cfg)

    for w in range(max(1, args.warmup)):
        ptr, out_len, ends, st = step()
        if w == 0:
            assert out_len == n and np.array_equal(hipmem.to_host(ptr, n), raw)
            log(f"warmup verified: {len(ends)} batches")
    prof = []

    def step_p():
        step()
        prof.append(dec.profile().as_dict())

    elapsed, lat = timed_steps(step_p, args.steps)
    result = _base("decompressed GB/s on 1 MI355X (device-resident read_gz, BGZF)", n * args.steps / elapsed / 1e9,
                   "GB/s", args.steps, args.warmup, elapsed * 1e3 / args.steps)
    result["config"] = {"workload": f"fastq-like {args.gib:g} GiB as BGZF (65280-byte members, level 6)",
                        "compressed_bytes": C, "uncompressed_bytes": n,
                        "members": int(prof[-1]["n_segments"]), "input": "resident in HBM"}
    result["latency_ms"] = {"p50": _pct(lat, 50), "p99": _pct(lat, 99)}
    result["stages_ms"] = {k: round(sum(p[k] for p in prof) / len(prof), 3)
                           for k in STAGES}
    result["verified"] = True
    d_in.free()
    return result


# ---------------------------------------------------------------------------
def run_members(args, threads):
    """pigz-style multi-member gzip (next row 1 of the scope table): the FASTQ
    bytes cut into 1 MiB pieces, each an independent member with a plain header,
    so the file is not BGZF and takes the speculative member path."""
    import numpy as np

    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

    n = int(args.gib * (1 << 30))
    raw = synth.fastq(n, threads=threads)
    t0 = time.time()
    g = synth.members(raw, args.member_kib << 10, 6, threads=threads)
    C = int(g.size)
    log(f"{args.member_kib} KiB members, level 6: {C} bytes in {time.time() - t0:.1f}s")
    d_in = hipmem.DeviceBuffer.from_host(g)
    cfg = make_cfg(args)
    dec = rrg.Decoder(0)

    def step():
        return dec.decode_device(d_in.ptr, C, g, cfg)

    for w in range(max(1, args.warmup)):
        ptr, out_len, ends, st = step()
        if w == 0:
            assert out_len == n and np.array_equal(hipmem.to_host(ptr, n), raw)
            log(f"warmup verified: {len(ends)} chunks, {st.speculation_failures} speculation failures")
    prof = []

    def step_p():
        step()
        prof.append(dec.profile().as_dict())

    elapsed, lat = timed_steps(step_p, args.steps)
    result = _base("decompressed GB/s on 1 MI355X (device-resident read_gz, multi-member)",
                   n * args.steps / elapsed / 1e9, "GB/s", args.steps, args.warmup, elapsed * 1e3 / args.steps)
    result["config"] = {"workload": f"fastq-like {args.gib:g} GiB as {args.member_kib} KiB gzip members "
                                    "(level 6, plain headers)",
                        "compressed_bytes": C, "uncompressed_bytes": n, "members": (n + (args.member_kib << 10) - 1)
                        // (args.member_kib << 10), "segments": int(prof[-1]["n_segments"]),
                        "input": "resident in HBM"}
    result["latency_ms"] = {"p50": _pct(lat, 50), "p99": _pct(lat, 99)}
    result["stages_ms"] = {k: round(sum(p[k] for p in prof) / len(prof), 3)
                           for k in STAGES}
    d_in.free()
    return result


# ---------------------------------------------------------------------------
def run_mixed(args, threads):
    """Mixed-entropy stress in one gzip member: 1-8 MiB segments of random
    bytes (stored), byte runs, period-3..300 repeats and Markov text at levels
    1/9, switched inside the member."""
    import numpy as np

    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

    n = int(args.mixed_gib * (1 << 30))
    raw, lv, ends = synth.mixed(n, threads=threads)
    t0 = time.time()
    gz = synth.gzip_segments(raw, lv, ends, threads=threads, piece=64 << 20)
    C = int(gz.size)
    log(f"mixed {n} -> {C} bytes ({len(ends)} segments) in {time.time() - t0:.1f}s")
    d_in = hipmem.DeviceBuffer.from_host(gz)
    cfg = make_cfg(args)
    dec = rrg.Decoder(0)

    def step():
        return dec.decode_device(d_in.ptr, C, gz, cfg)

    for w in range(max(1, args.warmup)):
        ptr, out_len, ends_out, st = step()
        if w == 0:
            assert out_len == n and np.array_equal(hipmem.to_host(ptr, n), raw)
            log(f"warmup verified: {len(ends_out)} chunks, {st.speculation_failures} speculation failures")
        log(f"warmup {w} {dec.profile().as_dict()}")
    prof = []

    def step_p():
        step()
        prof.append(dec.profile().as_dict())

    elapsed, lat = timed_steps(step_p, args.steps)
    result = _base("decompressed GB/s on 1 MI355X (device-resident read_gz, mixed entropy)",
                   n * args.steps / elapsed / 1e9, "GB/s", args.steps, args.warmup, elapsed * 1e3 / args.steps)
    result["config"] = {"workload": f"mixed-entropy {args.mixed_gib:g} GiB single member: 1-8 MiB segments of "
                                    "random bytes (stored), runs, period-3..300 repeats, Markov text at levels 1/9",
                        "compressed_bytes": C, "uncompressed_bytes": n, "chunk_size_bytes": cfg.chunk_size_bytes,
                        "segments": int(prof[-1]["n_segments"]), "input": "resident in HBM"}
    result["latency_ms"] = {"p50": _pct(lat, 50), "p99": _pct(lat, 99)}
    result["stages_ms"] = {k: round(sum(p[k] for p in prof) / len(prof), 3)
                           for k in STAGES}
    result["inflate_launches"] = int(prof[-1]["n_inflate_launches"])
    result["verified"] = True
    d_in.free()
    return result


# ---------------------------------------------------------------------------
def run_batch(args, threads):
    import numpy as np

    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import hipmem, synth

    nf, fsz = args.batch_files, args.batch_mib << 20
    raws, gzs = [], []
    for i in range(nf):
        r = synth.fastq(fsz, seed=synth.SEED_FASTQ + 1000 + i, threads=threads)
        raws.append(r)
        gzs.append(synth.gzip(r, 6, threads=threads, piece=16 << 20))
    bufs = [hipmem.DeviceBuffer.from_host(g) for g in gzs]
    ptrs, lens = [b.ptr for b in bufs], [int(g.size) for g in gzs]
    cfg = make_cfg(args)
    dec = rrg.Decoder(0)
    U = nf * fsz

    def batched():
        return dec.decode_batch(gzs, cfg, device_ptrs=ptrs, device_lens=lens, copy=False)

    def plain():
        for i in range(nf):
            dec.decode_device(ptrs[i], lens[i], gzs[i], cfg)

    for w in range(max(1, args.warmup)):
        res = batched()
        if w == 0:
            for i in (0, nf - 1):
                assert res[i].error is None and np.array_equal(hipmem.to_host(res[i].d_out, fsz), raws[i])
        plain()
    t_b, lat_b = timed_steps(batched, args.steps)
    t_p, lat_p = timed_steps(plain, args.steps)
    vb, vp = U * args.steps / t_b / 1e9, U * args.steps / t_p / 1e9
    result = _base("decompressed GB/s on 1 MI355X (batched decode of many files)", vb, "GB/s", args.steps,
                   args.warmup, t_b * 1e3 / args.steps)
    result["config"] = {"workload": f"{nf} FASTQ-like files of {args.batch_mib} MiB, gzip -6, one batched call",
                        "compressed_bytes": sum(lens), "uncompressed_bytes": U, "input": "resident in HBM"}
    result["plain_path"] = {"value": round(vp, 3), "unit": "GB/s", "how": "one rrg_gpu_decode_device call per file"}
    result["batched_over_plain"] = round(vb / vp, 2)
    result["latency_ms"] = {"batch_p50": _pct(lat_b, 50), "batch_p99": _pct(lat_b, 99)}
    for b in bufs:
        b.free()
    # tail + resolve + CRC as one graph launch against separate launches, on
    # batches of small files, where launch overheads are a visible share
    small = [synth.gzip(synth.fastq(256 << 10, seed=synth.SEED_FASTQ + 5000 + i, threads=threads), 6,
                        threads=threads).tobytes() for i in range(64)]
    modes = {}
    for name, graph in (("graph", True), ("separate", False)):
        c2 = make_cfg(args)
        c2.launch_graph = graph
        for _ in range(3):
            dec.decode_batch(small, c2, copy=False)
        ts = []
        for _ in range(20):
            t1 = time.perf_counter()
            dec.decode_batch(small, c2, copy=False)
            ts.append((time.perf_counter() - t1) * 1e3)
        p = dec.profile()
        modes[name] = {"ms_per_batch_p50": _pct(ts, 50), "graph_launches": int(p.n_graph_launches),
                       "ms_tail_resolve_crc": round(p.ms_tail + p.ms_resolve + p.ms_crc, 3)}
    result["post_kernels_graph"] = {
        "batch": "64 FASTQ-like files of 256 KiB (gzip -6), host inputs, one batched call",
        **modes,
        "saved_ms_per_batch": round(modes["separate"]["ms_per_batch_p50"] - modes["graph"]["ms_per_batch_p50"], 3),
    }
    return result


# ---------------------------------------------------------------------------
def run_serve(args, threads):
    import numpy as np

    import rusty_rapidgzip_amd as rrg
    from rusty_rapidgzip_amd import synth

    rng = np.random.default_rng(0x5E4E0005)
    lo, hi = 64 << 10, args.serve_max_mib << 20
    pool = []  # (gz bytes, uncompressed size)
    for g in ("text", "fastq", "mixed"):
        for lvl in (1, 6, 9):
            for _ in range(args.serve_pool):
                size = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
                seed = 0x5E4E0005 + 97 * len(pool)
                if g == "text":
                    raw = synth.text(size, seed=seed, threads=threads)
                    gz = synth.gzip(raw, lvl, threads=threads, piece=16 << 20)
                elif g == "fastq":
                    raw = synth.fastq(size, seed=seed, threads=threads)
                    gz = synth.gzip(raw, lvl, threads=threads, piece=16 << 20)
                else:
                    raw, lv, ends = synth.mixed(size, seed=seed, threads=threads)
                    gz = synth.gzip_segments(raw, lv, ends, threads=threads, piece=16 << 20)
                pool.append((gz.tobytes(), int(raw.size)))
    log(f"request pool: {len(pool)} inputs, {sum(u for _, u in pool) >> 20} MiB uncompressed")
    reqs = [pool[i] for i in rng.integers(0, len(pool), args.serve_requests)]
    cfg = rrg.Config(chunk_size_bytes=args.chunk_mib << 20)

    def run(rate, engines=args.serve_engines):
        """rate None: everything submitted at once (closed loop); else Poisson arrivals per second."""
        with rrg.Server(0, cfg, hbm_budget_bytes=args.serve_budget_gib << 30, engines=engines) as srv:
            done, lat, batches, total = 0, [], set(), 0
            arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs))) if rate else np.zeros(len(reqs))
            t0 = time.perf_counter()
            k = 0
            while done < len(reqs):
                now = time.perf_counter() - t0
                while k < len(reqs) and arrivals[k] <= now:
                    srv.submit(reqs[k][0], tag=k)
                    k += 1
                for c in srv.poll(256, timeout_ms=1000 if k >= len(reqs) else 1):
                    if c.error is not None or c.out_len != reqs[c.tag][1]:
                        raise RuntimeError(f"request {c.tag} failed: {c.error}")
                    lat.append(c.ms_total)
                    batches.add(c.batch)
                    total += c.out_len
                    srv.release(c.id)
                    done += 1
            el = time.perf_counter() - t0
        return total / el / 1e9, lat, len(batches), el

    run(None)  # warmup: first allocations
    cap, lat0, nb0, el0 = run(None)
    log(f"closed loop: {cap:.2f} GB/s, {nb0} batches, {el0:.2f}s")
    cap1, lat1, nb1, el1 = run(None, engines=1)
    log(f"closed loop, one engine: {cap1:.2f} GB/s, {nb1} batches, {el1:.2f}s")
    u_mean = sum(u for _, u in reqs) / len(reqs)
    sweeps = []
    for frac in args.serve_loads:
        rate = frac * cap * 1e9 / u_mean
        thr, lat, nb, el = run(rate)
        sweeps.append({"offered_frac": frac, "req_per_s": round(rate, 1), "GB_per_s": round(thr, 3),
                       "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99), "batches": nb})
        log(sweeps[-1])
    result = _base("decompressed GB/s on 1 MI355X (serving loop, request mix)", cap, "GB/s", 1, 1, el0 * 1e3)
    result["config"] = {
        "workload": f"{len(reqs)} requests, U log-uniform on [64 KiB, {args.serve_max_mib} MiB], "
                    "text/FASTQ/mixed x levels 1/6/9, inputs from host memory, outputs left in HBM",
        "pool_inputs": len(pool),
        "hbm_budget_gib": args.serve_budget_gib,
        "engines": args.serve_engines,
    }
    result["latency_ms"] = {"closed_loop_p50": _pct(lat0, 50), "closed_loop_p99": _pct(lat0, 99)}
    result["one_engine"] = {"GB_per_s": round(cap1, 3), "p50_ms": _pct(lat1, 50), "p99_ms": _pct(lat1, 99),
                            "batches": nb1}
    result["poisson"] = sweeps
    return result


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="fastq4g",
                    choices=("fastq4g", "bgzf", "members", "mixed", "text64m", "batch", "serve"))
    ap.add_argument("--member-kib", type=int, default=1024, help="member size of the members workload")
    ap.add_argument("--mixed-gib", type=float, default=2.0, help="uncompressed size of the mixed workload")
    ap.add_argument("--gib", type=float, default=4.0, help="uncompressed size of the workload")
    ap.add_argument("--chunk-mib", type=int, default=4)
    ap.add_argument("--segment-kib", type=int, default=0, help="0 = automatic")
    ap.add_argument("--lut-bits", type=int, default=0, help="0 = automatic (10 with --lanes 64, 9 otherwise)")
    ap.add_argument("--lanes", type=int, default=64,
                    help="64: one wavefront per segment (k_tokens_wave); 8/16/32: one lane per segment (k_tokens)")
    ap.add_argument("--cpu-sample-mib", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e2e", action="store_true", help="skip the host-buffer (PCIe-inclusive) rate")
    ap.add_argument("--no-others", action="store_true",
                    help="fastq4g only: skip the other configurations (BGZF, mixed entropy, text)")
    ap.add_argument("--autotune", action="store_true", help="pick lanes / LUT width / segment size with the autotuner")
    ap.add_argument("--batch-files", type=int, default=64)
    ap.add_argument("--batch-mib", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-max-mib", type=int, default=64)
    ap.add_argument("--serve-pool", type=int, default=2, help="inputs per generator x level cell")
    ap.add_argument("--serve-budget-gib", type=int, default=32)
    ap.add_argument("--serve-loads", type=float, nargs="*", default=[0.5, 0.9])
    ap.add_argument("--serve-engines", type=int, default=2, help="batches the server decodes at once")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build targets one GPU; --gpus must be 1"}))
        log("only --gpus 1 is supported: the decoder runs on a single MI355X")
        return 2
    threads = min(16, os.cpu_count() or 1)
    fn = {"fastq4g": run_fastq, "bgzf": run_bgzf, "members": run_members, "mixed": run_mixed, "text64m": run_text64m,
          "batch": run_batch, "serve": run_serve}[args.workload]
    result = fn(args, threads)
    result.pop("_raw", None)
    if args.workload == "fastq4g" and not args.no_others:
        # the other configurations of the scope table, each output-verified
        # against its generator (one step = one whole decode), each in a
        # process of its own
        others = {}
        for name in ("bgzf", "mixed", "text64m"):
            cmd = [sys.executable, os.path.abspath(__file__), "--workload", name, "--steps", str(args.steps),
                   "--warmup", "1", "--no-e2e", "--lut-bits", str(args.lut_bits), "--lanes", str(args.lanes)]
            if args.no_cpu:
                cmd.append("--no-cpu")
            import subprocess

            r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
            lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
            if r.returncode != 0 or not lines:
                raise RuntimeError(f"{name}: child bench failed with status {r.returncode}")
            others[name] = _sub(json.loads(lines[-1]))
            log(f"{name} {others[name]['value']} GB/s")
        result["other_configs"] = others
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
