"""Benchmark: 2^28 float32 standard-normal samples per step on one MI355X.

Workload (the univariate-sweep configuration, normal(0,1) leg): 4096
Vmt19937_64 engines seeded 0, 4, ..., 16380 (16384 MT19937-64 streams seeded
0..16383), each filling a contiguous block of 65536 floats, i.e. one
`Rand::normal<ArrayXf>(65536, 1, engine_g)` per engine. One step = one launch
of the fill over the whole 1 GiB output, which stays resident in HBM.

Prints one JSON line (rank 0) with Gsamples/s, the HBM-write roofline of the
fill kernel (algorithmic bytes = 4 B written per sample, timed with HIP events
on the launch stream), and the CPU baseline (the oracle restatement of the
reference algorithm, timed on this host's cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gsamples/s + % HBM peak (normal f32, 2^28); p50/p99 batched-request latency"
N_ENGINES = 4096
PER_ENGINE = 65536
N_SAMPLES = N_ENGINES * PER_ENGINE  # 2^28
HBM_PEAK_GBS = 8000.0


def cpu_baseline(max_threads=16):
    """Oracle restatement (C++, lane-emulated AVX2 layout) on the host: one
    engine per thread over the same 4096-engine workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    e = pyoracle.Engines(n=N_ENGINES, seed0=0, stride=4)
    t0 = time.perf_counter()
    pyoracle.fill_normal_f32(e, PER_ENGINE, threads=threads)
    dt = time.perf_counter() - t0
    # single-thread rate on a 64-engine sample of the same workload
    e1 = pyoracle.Engines(n=64, seed0=0, stride=4)
    t1 = time.perf_counter()
    pyoracle.fill_normal_f32(e1, PER_ENGINE, threads=1)
    dt1 = time.perf_counter() - t1
    return {
        "value": N_SAMPLES / dt / 1e9,
        "unit": "Gsamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full workload: 4096 engines x 65536 normals, {threads} threads (one engine per "
                  f"thread at a time); reference itself not buildable (Eigen absent)",
        "single_thread_value": 64 * PER_ENGINE / dt1 / 1e9,
        "single_thread_sample": "64 engines x 65536 normals on 1 thread",
    }


def load_traffic():
    p = os.path.join(ROOT, "profiles", "normal_f32_2p28_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the sampling path runs on a single MI355X "
              "(independent engines shard across GPUs with no collective)", file=sys.stderr)
        sys.exit(2)

    import torch
    import eigenrand_amd as er

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    urng = er.Vmt19937_64(0, n=N_ENGINES, stride=4)
    out = torch.empty(N_SAMPLES, dtype=torch.float32, device="cuda")

    for _ in range(args.warmup):
        er.normal(PER_ENGINE, 1, urng, out=out)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        er.normal(PER_ENGINE, 1, urng, out=out)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = N_SAMPLES / (wall / args.steps) / 1e9
    alg_bytes = 4 * N_SAMPLES
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (engine-generated; no input data)",
        "config": {
            "workload": "normal(0,1) f32, 2^28 samples: 4096 Vmt19937_64 engines seeded 0,4,..,16380 "
                        "x 65536 contiguous samples each",
            "n_samples": N_SAMPLES,
            "engines": N_ENGINES,
            "mt_streams": 4 * N_ENGINES,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(),
            "kernel": "k_fill_fixed<NormalF32<0>>",
            "kernel_ms": round(kernel_ms, 4),
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
