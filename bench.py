#!/usr/bin/env python3
"""Benchmark: posit32 GEMM throughput on the blocked-LU trailing-update shapes.

Workload (one "step"): the four trailing-update GEMMs of a blocked LU at n = 16384 with
nb = 256, i.e. C = C - A*B with M = N = n - k for k in {0, 4096, 8192, 12288}, K = 256,
alpha = -1, beta = 1; A, B, C uniform[-1,1) in FP64 converted to posit32 (seeded, generated
on the device), all resident in HBM before timing.  value = sum(2*M*N*K) / time.

Besides the JSON fields of the run contract the line carries:
  roofline      integer-VALU roofline of the dominant kernel (the M = N = 16384 GEMM):
                achieved = VALU lane-ops per launch / its average duration (HIP events on
                the library stream), where VALU lane-ops per launch = 64 * SQ_INSTS_VALU
                measured with rocprofv3 (profiles/rgemm_valu_counts.json, scaled by MNK);
                peak = CUs * 4 SIMD * 16 lanes * clock; traffic = HBM bytes per launch from
                FETCH_SIZE + WRITE_SIZE in the same file.
  cpu_baseline  the oracle's restatement of the reference CPU Rgemm (OpenMP over columns,
                as libRpack/Rgemm.cpp:168) on a bounded sample of the 16384 shape.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SHAPES = [16384, 12288, 8192, 4096]
K = 256
P_ONE, P_MONE = 0x40000000, 0xC0000000
COUNTS = os.path.join(ROOT, "profiles", "rgemm_valu_counts.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(seconds_target=12.0):
    """Time the oracle Rgemm (port of libRpack/Rgemm.cpp, NN branch) on column samples of
    the M=16384, K=256 trailing shape with all host cores (capped at 16)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from datagen import posit_matrix
    from oracle_lib import Oracle

    O = Oracle()
    cores = min(16, os.cpu_count() or 1)
    M = SHAPES[0]
    A = posit_matrix(O, 501, M, K)
    ncols = 4
    elapsed = 0.0
    while True:
        B = posit_matrix(O, 502, K, ncols)
        C = posit_matrix(O, 503, M, ncols)
        t0 = time.perf_counter()
        O.rgemm("N", "N", M, ncols, K, P_MONE, A, B, P_ONE, C, nthreads=cores)
        elapsed = time.perf_counter() - t0
        if elapsed > seconds_target / 4 or ncols >= 4096:
            break
        ncols = int(ncols * max(2.0, min(16.0, (seconds_target / 4) / max(elapsed, 1e-3))))
    gops = 2.0 * M * ncols * K / elapsed / 1e9
    return {
        "value": round(gops, 4),
        "unit": "Gop/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle Rgemm NN, M={M}, N={ncols} columns, K={K}, alpha=-1, beta=1, "
                  f"{elapsed:.2f} s on {cores} OpenMP threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-shape", type=int, default=0,
                    help="run only the M=N=<value> GEMM (used for per-dispatch profiling)")
    args = ap.parse_args()
    global SHAPES
    if args.only_shape:
        SHAPES = [args.only_shape]
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the posit32 GEMM path is "
              "benchmarked on a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import posit_gpu_amd as pga

    pga.init(0)
    info = pga.device_info()
    log(f"device: {info['name']} {info['arch']} CUs={info['compute_units']} clock={info['clock_khz']} kHz")

    bufs = []
    for i, n in enumerate(SHAPES):
        dA = pga.DeviceArray((n, K))
        dB = pga.DeviceArray((K, n))
        dC = pga.DeviceArray((n, n))
        pga.fill_uniform(dA, 1000 + 3 * i)
        pga.fill_uniform(dB, 1001 + 3 * i)
        pga.fill_uniform(dC, 1002 + 3 * i)
        bufs.append((n, dA, dB, dC))
    pga.synchronize()

    def step(evs=None):
        for s, (n, dA, dB, dC) in enumerate(bufs):
            if evs is not None:
                evs[s].record()
            pga.rgemm_device("N", "N", n, n, K, P_MONE, dA, n, dB, K, P_ONE, dC, n)
        if evs is not None:
            evs[len(bufs)].record()

    for w in range(args.warmup):
        step()
        pga.synchronize()
        log(f"warmup step {w + 1}/{args.warmup} done")

    evs = [[pga.Event() for _ in range(len(bufs) + 1)] for _ in range(args.steps)]
    pga.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(evs[s])
    pga.synchronize()
    wall = time.perf_counter() - t0
    per_shape_ms = np.zeros(len(bufs))
    for s in range(args.steps):
        for i in range(len(bufs)):
            per_shape_ms[i] += evs[s][i].elapsed_ms(evs[s][i + 1])
    per_shape_ms /= args.steps
    log(f"timed {args.steps} steps: wall {wall:.3f} s, per-shape ms {np.round(per_shape_ms, 3).tolist()}")

    ops_step = sum(2.0 * n * n * K for n in SHAPES)
    ms_step = wall / args.steps * 1e3
    gops = ops_step * args.steps / wall / 1e9

    # roofline of the dominant kernel (the 16384 GEMM)
    n0 = SHAPES[0]
    t_dom = per_shape_ms[0] / 1e3
    macs = float(n0) * n0 * K
    peak = info["int32_valu_peak"]
    roof = {"bound": "valu", "achieved": None, "peak": round(peak / 1e12, 3), "unit": "Tlane-op/s",
            "frac": None, "traffic": None, "kernel": "rgemm_tiled NN 64x64x16",
            "avg_ms": round(per_shape_ms[0], 4), "gops": round(2 * macs / t_dom / 1e9, 2)}
    if os.path.exists(COUNTS):
        c = json.load(open(COUNTS))
        lane_ops_per_mac = c["valu_lane_ops_per_mac"]
        achieved = lane_ops_per_mac * macs / t_dom
        roof["achieved"] = round(achieved / 1e12, 3)
        roof["frac"] = round(achieved / peak, 4)
        roof["valu_insts_per_posit_op"] = round(lane_ops_per_mac / 2.0, 2)
        if c.get("hbm_bytes_per_mac") is not None:
            roof["traffic"] = int(c["hbm_bytes_per_mac"] * macs)
        roof["counts_from"] = os.path.relpath(COUNTS, ROOT)

    out = {
        "metric": "posit32 GEMM Gop/s (2MNK/t)",
        "value": round(gops, 2),
        "unit": "Gop/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "posit32 (integer VALU)",
        "data": "synthetic: seeded uniform[-1,1) FP64 -> posit32 RNE, generated on the device",
        "config": {
            "workload": "blocked-LU trailing-update GEMMs at n=16384, nb=256: C -= A*B with "
                        "M=N in {16384,12288,8192,4096}, K=256, alpha=-1, beta=1, column-major",
            "shapes_mnk": [[n, n, K] for n in SHAPES],
            "per_shape_ms": [round(x, 4) for x in per_shape_ms],
            "per_shape_gops": [round(2.0 * n * n * K / (t / 1e3) / 1e9, 2) for n, t in zip(SHAPES, per_shape_ms)],
            "device": info["name"],
        },
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
