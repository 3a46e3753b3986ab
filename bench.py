#!/usr/bin/env python3
"""Benchmark: posit32 GEMM throughput on the blocked-LU trailing-update shapes.

Workload (one "step"): the four trailing-update GEMMs of a blocked LU at n = 16384 with
nb = 256, i.e. C = C - A*B with M = N = n - k for k in {0, 4096, 8192, 12288}, K = 256,
alpha = -1, beta = 1; A, B, C uniform[-1,1) in FP64 converted to posit32 (seeded, generated
on the device), all resident in HBM before timing.  value = sum(2*M*N*K) / time.

Besides the JSON fields of the run contract the line carries:
  roofline      integer-VALU roofline of the dominant kernel (the M = N = 16384 GEMM: its
                two FP64 operand-panel kernels and the 64 x 32 tile kernel):
                achieved = VALU lane-ops per call / its average duration (HIP events on
                the library stream), where VALU lane-ops per call = 64 * SQ_INSTS_VALU
                measured with rocprofv3 (profiles/rgemm_valu_counts.json, scaled by MNK);
                peak = CUs * 4 SIMD * 16 lanes * clock; traffic = HBM bytes per launch from
                FETCH_SIZE + WRITE_SIZE in the same file (scaled by a streaming kernel's
                counts, bench.py --config copycal).
  batched_gemm, batched_factor, serving
                the other parts of the metric, each with its bitwise-mismatch count:
                4096 GEMMs (M, N, K on {16..512}) in one persistent launch from a
                hipGraph against per-call launches; 2048 LU + 2048 lower Cholesky
                factorizations (n in [32, 512]) likewise; a request stream (50% LU / 50%
                Cholesky, n log-uniform [64, 4096], 1-16 RHS, 20% re-solves) served with
                continuous batching against plain per-request launches, as capacity and
                at a fixed Poisson rate (solves/s, p50, p99).  --no-extras leaves them out.
  cpu_baseline  the oracle's restatement of the reference CPU Rgemm (OpenMP over columns,
                as libRpack/Rgemm.cpp:168) on a bounded sample of the 16384 shape, on every
                thread the process is given and on one thread (medians of three), with the
                host's core counts and CPU model.
"""
import argparse
import json
import math
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
METRIC = "posit32 GEMM Gop/s (2MNK/t), %INT VALU peak; batched solves/s + p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_facts():
    """Host cores this process may use, the OpenMP thread count it is given and the CPU
    model (the GPU box assigns one GPU's share of its cores: OMP_NUM_THREADS there)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": avail, "omp_num_threads": omp or None,
            "threads": min(avail, omp) if omp > 0 else avail, "cpu_model": model}


def cpu_baseline(seconds_target=8.0):
    """The oracle's restatement of the reference CPU Rgemm (NN branch of libRpack/Rgemm.cpp,
    OpenMP over columns as at Rgemm.cpp:168) on column samples of the M=16384, K=256
    trailing shape: the median of three runs on every thread this process is given, and
    the median of three single-thread runs."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from datagen import posit_matrix
    from oracle_lib import Oracle

    O = Oracle()
    facts = _cpu_facts()
    threads = facts["threads"]
    M = SHAPES[0]
    A = posit_matrix(O, 501, M, K)

    def run(ncols, nthreads):
        B = posit_matrix(O, 502, K, ncols)
        C = posit_matrix(O, 503, M, ncols)
        t0 = time.perf_counter()
        O.rgemm("N", "N", M, ncols, K, P_MONE, A, B, P_ONE, C, nthreads=nthreads)
        return time.perf_counter() - t0

    def sample(nthreads, target):
        ncols = 2
        while True:  # grow the sample until one run takes about target / 3
            t = run(ncols, nthreads)
            if t > target / 6 or ncols >= 8192:
                break
            ncols = int(ncols * max(2.0, min(16.0, (target / 3) / max(t, 1e-3))))
        ts = sorted(run(ncols, nthreads) for _ in range(3))
        return ncols, ts[1]

    ncols, t_mt = sample(threads, seconds_target)
    ncols1, t_1 = sample(1, seconds_target / 2)
    return {
        "value": round(2.0 * M * ncols * K / t_mt / 1e9, 4),
        "unit": "Gop/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle Rgemm NN, M={M}, K={K}, alpha=-1, beta=1, N={ncols} columns on {threads} OpenMP "
                  f"threads (median of 3: {t_mt:.2f} s) and N={ncols1} columns on 1 thread (median of 3: "
                  f"{t_1:.2f} s)",
        "single_thread": round(2.0 * M * ncols1 * K / t_1 / 1e9, 4),
        **{k: v for k, v in facts.items() if k != "threads"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="default run: leave out the batched and served sub-measurements")
    ap.add_argument("--config", default="gemm",
                    choices=["gemm", "dropin", "lu", "cholesky", "batched", "batched_factor", "blocked", "serve",
                             "copycal"],
                    help="gemm: trailing-update GEMMs (headline); dropin: Rgemm 512^3 (alpha=1, beta=0) "
                         "through the host-pointer RgemmG drop-in and device-resident; lu / cholesky: blocked "
                         "factorization + solve at n=8192; batched: 4096 small GEMMs; batched_factor: "
                         "2048 LU + 2048 Cholesky factorizations of n in [32, 512]; blocked: LU + "
                         "upper Cholesky solves of n in [513, 4096] in one blocked batch; serve: Poisson "
                         "stream of solve requests, served vs plain per-request; copycal: a "
                         "streaming kernel of known size (2 x 1 GiB read, 1 GiB written) that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--n", type=int, default=8192, help="order for --config lu / cholesky")
    ap.add_argument("--requests", type=int, default=400, help="requests for --config serve")
    ap.add_argument("--qps", type=float, default=60.0, help="arrival rate for --config serve")
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
    if args.config == "dropin":
        return bench_dropin(pga, info, args)
    if args.config == "lu":
        return bench_lu(pga, info, args)
    if args.config == "cholesky":
        return bench_cholesky(pga, info, args)
    if args.config == "batched":
        return bench_batched(pga, info, args)
    if args.config == "copycal":
        return bench_copycal(pga, info, args)
    if args.config == "batched_factor":
        return bench_batched_factor(pga, info, args)
    if args.config == "serve":
        return bench_serve(pga, info, args)
    if args.config == "blocked":
        return bench_blocked(pga, info, args)

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
            "frac": None, "traffic": None, "kernel": "rgemm_tiled_panel NN 64x32 + k_panel_a + k_panel_t",
            "avg_ms": round(per_shape_ms[0], 4), "gops": round(2 * macs / t_dom / 1e9, 2)}
    if os.path.exists(COUNTS):
        # rocprofv3 counts of one launch of this kernel at M = N = 16384, K = 256:
        # VALU lane-ops per multiply-add = 64 * SQ_INSTS_VALU / MNK (the per-element work
        # does not depend on the shape); HBM bytes = FETCH_SIZE and WRITE_SIZE scaled by
        # what they read for a streaming kernel of known size.
        c = json.load(open(COUNTS))
        lane_ops_per_mac = c["valu_lane_ops_per_mac"]
        achieved = lane_ops_per_mac * macs / t_dom
        roof["achieved"] = round(achieved / 1e12, 3)
        roof["frac"] = round(achieved / peak, 4)
        roof["valu_insts_per_posit_op"] = round(lane_ops_per_mac / 2.0, 2)
        if c.get("shape_mnk") == [n0, n0, K]:
            roof["traffic"] = int(c["hbm_bytes"])
            roof["algorithmic_bytes"] = int(c["algorithmic_bytes"])
            roof["hbm_GBps"] = round(c["hbm_bytes"] / t_dom / 1e9, 1)
        roof["counts_from"] = os.path.relpath(COUNTS, ROOT)

    out = {
        "metric": METRIC,
        "value": round(gops, 2),
        "unit": "Gop/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "posit32 (VALU: integer and exact FP64 operations)",
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
    for b in bufs:
        for d in b[1:]:
            d.free()
    bufs.clear()
    if not args.no_extras:
        # the other measured parts of the metric: batched GEMMs, batched factorizations
        # (both one persistent launch replayed from a hipGraph) and the served form
        log("batched GEMMs ...")
        out["batched_gemm"] = run_batched_gemm(pga, 3)
        log("batched factorizations ...")
        out["batched_factor"] = run_batched_factor(pga, 2)
        log("serving ...")
        out["serving"] = run_serving(pga, args.requests, args.qps)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def _timed(pga, fn):
    e0, e1 = pga.Event(), pga.Event()
    e0.record()
    fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_ms(e1) / 1e3


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    return Oracle()


def bench_dropin(pga, info, args):
    """Plain Rgemm C = A*B, M = N = K = 512, alpha = 1, beta = 0, through the host-pointer
    drop-in RgemmG (copies in, kernel, copies out, synchronous like libRpack's RgemmG),
    timed wall-clock around each call as main_gemm.cpp:128-140 does; beside it the same GEMM
    on device-resident operands (HIP events).  value is the device-resident rate."""
    n = 512
    dA, dB, dC = pga.DeviceArray((n, n)), pga.DeviceArray((n, n)), pga.DeviceArray((n, n))
    pga.fill_uniform(dA, 5101)
    pga.fill_uniform(dB, 5102)
    pga.fill_uniform(dC, 5103)
    pga.synchronize()
    A, B = dA.to_host(), dB.to_host()
    C = dC.to_host()
    ops = 2.0 * n ** 3
    reps = max(args.steps, 5)
    for _ in range(max(args.warmup, 1)):
        pga.RgemmG("N", "N", n, n, n, P_ONE, A, n, B, n, 0, C, n)
    t0 = time.perf_counter()
    for _ in range(reps):
        pga.RgemmG("N", "N", n, n, n, P_ONE, A, n, B, n, 0, C, n)
    t_e2e = (time.perf_counter() - t0) / reps
    pga.rgemm_device("N", "N", n, n, n, P_ONE, dA, n, dB, n, 0, dC, n)
    pga.synchronize()
    t_dev = _timed(pga, lambda: [pga.rgemm_device("N", "N", n, n, n, P_ONE, dA, n, dB, n, 0, dC, n)
                                 for _ in range(reps)]) / reps
    same = bool(np.array_equal(dC.to_host(), C))
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle's Rgemm (port of Rgemm.cpp NN) on the same operands
        O = _oracle()
        cores = _cpu_facts()["threads"]
        Cc = np.zeros((n, n), np.uint32, order="F")
        t0 = time.perf_counter()
        O.rgemm("N", "N", n, n, n, P_ONE, np.asfortranarray(A), np.asfortranarray(B), 0, Cc, nthreads=cores)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(ops / t_cpu / 1e9, 4), "unit": "Gop/s", "cores": cores, "kind": "port",
               "sample": f"oracle Rgemm NN 512^3, {t_cpu:.2f} s", "bitwise_equal": bool(np.array_equal(Cc, C))}
    out = {
        "metric": "posit32 GEMM Gop/s (2MNK/t)", "value": round(ops / t_dev / 1e9, 2), "unit": "Gop/s",
        "n_gpus": 1, "steps": reps, "ms_per_step": round(t_dev * 1e3, 4), "higher_is_better": True,
        "dtype": "posit32 (VALU: integer and exact FP64 operations)",
        "data": "synthetic: seeded uniform[-1,1) FP64 -> posit32, generated on the device",
        "config": {"workload": "Rgemm NN M=N=K=512, alpha=1, beta=0 (plain drop-in case)",
                   "dropin_ms_per_call": round(t_e2e * 1e3, 4),
                   "dropin_gops_including_copies": round(ops / t_e2e / 1e9, 2),
                   "dropin_equals_device_resident": same, "device": info["name"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_lu(pga, info, args):
    """Rgetrf_opencl (nb=256) + Rgetrs on the device, b = A*ones by the posit GEMV;
    block-size invariance (nb=64 gives the same bits) as the full-size parity check."""
    n, nb = args.n, 256
    dA0 = pga.DeviceArray((n, n))
    pga.fill_uniform(dA0, 4242)
    dA = pga.DeviceArray((n, n))
    dA.copy_from(dA0)
    ones = pga.DeviceArray.from_host(np.full(n, P_ONE, np.uint32))
    db = pga.DeviceArray(n, np.uint32, "C")
    pga.gemv_device("N", n, n, P_ONE, dA0, n, ones, 1, 0, db, 1)
    dx = pga.DeviceArray((n, 1))
    dx.copy_from(db)
    dp = pga.DeviceArray(n, np.int32, "C")
    di = pga.DeviceArray(1, np.int32, "C")
    pga.synchronize()
    t_f = _timed(pga, lambda: pga.getrf_device(n, n, dA, n, dp, di, nb))
    t_s = _timed(pga, lambda: pga.getrs_device("N", n, 1, dA, n, dp, dx, n))
    berr = pga.backward_error_device(n, dA0, n, dx, db)
    info_v = int(di.to_host()[0])
    F1 = dA.to_host()
    dA.copy_from(dA0)
    pga.getrf_device(n, n, dA, n, dp, di, 64)
    pga.synchronize()
    same = bool(np.array_equal(F1, dA.to_host()))
    del F1
    ops = 2.0 / 3.0 * n ** 3
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle's blocked LU (port of Rgetrf_opencl / Rgetrf2) at n_cpu
        O = _oracle()
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from datagen import posit_matrix

        cores = _cpu_facts()["threads"]
        n_cpu = 1536
        Ac = posit_matrix(O, 4243, n_cpu, n_cpu)
        t0 = time.perf_counter()
        O.rgetrf(Ac, nb=64, nthreads=cores)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(2.0 / 3.0 * n_cpu ** 3 / t_cpu / 1e9, 4), "unit": "Gop/s", "cores": cores,
               "kind": "port", "sample": f"oracle Rgetrf n={n_cpu} nb=64, {t_cpu:.2f} s"}
    out = {
        "metric": "posit32 LU Gop/s ((2/3)n^3/t)", "value": round(ops / t_f / 1e9, 2), "unit": "Gop/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "posit32 (VALU: integer and exact FP64 operations)",
        "data": "synthetic: seeded uniform[-1,1) FP64 -> posit32, generated on the device",
        "config": {"workload": f"Rgetrf_opencl n={n} nb={nb} + Rgetrs (b = A*ones)", "factor_s": round(t_f, 4),
                   "solve_s": round(t_s, 4), "info": info_v, "backward_error": berr,
                   "nb256_equals_nb64": same, "device": info["name"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_cholesky(pga, info, args):
    """Rpotrf_opencl (lower, nb=64) + Rpotrs on A = M*M^T/n + I (M uniform[-1,1))."""
    n = args.n
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from datagen import uniform_doubles

    M = uniform_doubles(4300, n * n).reshape(n, n)
    S = M @ M.T
    S /= n
    S[np.diag_indices(n)] += 1.0
    del M
    dS = pga.DeviceArray.from_host(np.asfortranarray(S))
    del S
    dA0 = pga.DeviceArray((n, n))
    pga._check(pga._lib.pga_p32_from_f64_device(dS.ptr, dA0.ptr, n * n, None), "from_f64")
    dS.free()
    dA = pga.DeviceArray((n, n))
    dA.copy_from(dA0)
    ones = pga.DeviceArray.from_host(np.full(n, P_ONE, np.uint32))
    db = pga.DeviceArray(n, np.uint32, "C")
    pga.gemv_device("N", n, n, P_ONE, dA0, n, ones, 1, 0, db, 1)
    dx = pga.DeviceArray((n, 1))
    dx.copy_from(db)
    di = pga.DeviceArray(1, np.int32, "C")
    pga.synchronize()
    t_f = _timed(pga, lambda: pga.potrf_device("L", n, dA, n, di, 0))
    t_s = _timed(pga, lambda: pga.potrs_device("L", n, 1, dA, n, dx, n))
    berr = pga.backward_error_device(n, dA0, n, dx, db)
    info_v = int(di.to_host()[0])
    L1 = np.tril(dA.to_host())
    dA.copy_from(dA0)
    pga.potrf_device("L", n, dA, n, di, 128)
    pga.synchronize()
    same = bool(np.array_equal(L1, np.tril(dA.to_host())))
    del L1
    # the upper factor of the same matrix (Rpotrf_opencl 'U', nb = 64): its own sums per
    # block row, so no block-size invariance to check; its bits are pinned by the tests
    dA.copy_from(dA0)
    pga.synchronize()
    t_fu = _timed(pga, lambda: pga.potrf_device("U", n, dA, n, di, 0))
    info_u = int(di.to_host()[0])
    ops = n ** 3 / 3.0
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle's blocked lower Cholesky (port of Rpotrf_opencl / Rpotrf2) at n_cpu
        O = _oracle()
        cores = _cpu_facts()["threads"]
        n_cpu = 2048
        Mc = uniform_doubles(4301, n_cpu * n_cpu).reshape(n_cpu, n_cpu)
        Sc = Mc @ Mc.T / n_cpu + np.eye(n_cpu)
        Ac = np.asfortranarray(O.from_double_v(Sc.ravel(order="F")).reshape(n_cpu, n_cpu, order="F"))
        t0 = time.perf_counter()
        O.rpotrf("L", Ac, nb=64, nthreads=cores)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(n_cpu ** 3 / 3.0 / t_cpu / 1e9, 4), "unit": "Gop/s", "cores": cores,
               "kind": "port", "sample": f"oracle Rpotrf lower n={n_cpu} nb=64, {t_cpu:.2f} s"}
    out = {
        "metric": "posit32 Cholesky Gop/s ((1/3)n^3/t)", "value": round(ops / t_f / 1e9, 2), "unit": "Gop/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "posit32 (VALU: integer and exact FP64 operations)",
        "data": "synthetic: M*M^T/n + I with M seeded uniform[-1,1), FP64 -> posit32",
        "config": {"workload": f"Rpotrf_opencl lower n={n} nb=64 + Rpotrs", "factor_s": round(t_f, 4),
                   "solve_s": round(t_s, 4), "info": info_v, "backward_error": berr,
                   "lower_nb64_equals_nb128": same, "upper_factor_s": round(t_fu, 4), "upper_info": info_u,
                   "device": info["name"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_batched_gemm(pga, reps):
    """4096 independent GEMMs, M, N, K uniform on multiples of 16 in [16, 512], alpha =
    beta = 1: one persistent launch (autotuned plan) replayed from a hipGraph, against
    per-call launches one after another; every C compared bit for bit."""
    rng = np.random.default_rng(4096)
    count = 4096
    shapes = rng.integers(1, 33, (count, 3)) * 16
    offs = np.zeros((count + 1, 3), np.int64)
    for i, (m, n, k) in enumerate(shapes):
        offs[i + 1] = offs[i] + [m * k, k * n, m * n]
    dA = pga.DeviceArray(int(offs[-1, 0]), np.uint32, "C")
    dB = pga.DeviceArray(int(offs[-1, 1]), np.uint32, "C")
    dC0 = pga.DeviceArray(int(offs[-1, 2]), np.uint32, "C")
    dC = pga.DeviceArray(int(offs[-1, 2]), np.uint32, "C")
    pga.fill_uniform(dA, 41)
    pga.fill_uniform(dB, 42)
    pga.fill_uniform(dC0, 43)
    probs = []
    total_ops = 0.0
    for i, (m, n, k) in enumerate(shapes):
        m, n, k = int(m), int(n), int(k)
        probs.append(dict(m=m, n=n, k=k, alpha=P_ONE, beta=P_ONE, a=dA.ptr + 4 * int(offs[i, 0]), lda=m,
                          b=dB.ptr + 4 * int(offs[i, 1]), ldb=k, c=dC.ptr + 4 * int(offs[i, 2]), ldc=m))
        total_ops += 2.0 * m * n * k
    dC.copy_from(dC0)
    pga.synchronize()
    t0 = time.perf_counter()
    plan = pga.BatchPlan("N", "N", probs, tile=0)
    t_tune = time.perf_counter() - t0
    plan.capture_graph()
    dC.copy_from(dC0)
    plan.launch_graph()
    pga.synchronize()
    got = dC.to_host()
    t_graph = min(_timed(pga, plan.launch_graph) for _ in range(max(1, reps)))

    def per_call():
        for p in probs:
            pga.rgemm_device("N", "N", p["m"], p["n"], p["k"], P_ONE, p["a"], p["lda"], p["b"], p["ldb"], P_ONE,
                             p["c"], p["ldc"])

    dC.copy_from(dC0)
    pga.synchronize()
    t_plain = _timed(pga, per_call)
    want = dC.to_host()
    mism = sum(0 if np.array_equal(got[offs[i, 2]:offs[i + 1, 2]], want[offs[i, 2]:offs[i + 1, 2]]) else 1
               for i in range(count))
    info = plan.info()
    del plan
    for d in (dA, dB, dC0, dC):
        d.free()
    return {"workload": "4096 GEMMs C += A*B, M,N,K uniform on {16,32,...,512}, alpha=beta=1, one persistent "
                        "launch replayed from a hipGraph",
            "value": round(total_ops / t_graph / 1e9, 2), "unit": "Gop/s",
            "gemms_per_s": round(count / t_graph, 1), "graph_s": round(t_graph, 5),
            "per_call_s": round(t_plain, 5), "speedup_vs_per_call": round(t_plain / t_graph, 2),
            "plan": info, "plan_build_s": round(t_tune, 2), "bitwise_mismatches_vs_per_call": mism}


def bench_batched(pga, info, args):
    r = run_batched_gemm(pga, max(1, args.steps))
    out = {"metric": "batched posit32 GEMM Gop/s (sum 2MNK/t)", "value": r["value"], "unit": "Gop/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "posit32 (VALU: integer and exact FP64 operations)",
           "data": "synthetic: seeded uniform[-1,1) posit32, shapes from a seeded generator",
           "config": dict(r, device=info["name"])}
    print(json.dumps(out), flush=True)


def run_batched_factor(pga, reps):
    """2048 LU (Rgetrf_opencl, nb = 64) and 2048 lower Cholesky (Rpotrf_opencl, nb = 64)
    factorizations, n uniform in [32, 512], in one persistent launch replayed from a
    hipGraph, against the per-call device routines one problem after another; factors,
    pivots and info compared bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from datagen import uniform_doubles

    rng = np.random.default_rng(2048)
    count = 2048
    ns = [("getrf", int(n)) for n in rng.integers(32, 513, count)] + \
         [("potrf", int(n)) for n in rng.integers(32, 513, count)]
    items = []
    for i, (kind, n) in enumerate(ns):
        d0 = pga.DeviceArray((n, n))
        if kind == "getrf":
            pga.fill_uniform(d0, 50000 + i)
        else:
            M = uniform_doubles(60000 + i, n * n).reshape(n, n)
            S = M @ M.T
            S /= n
            S[np.diag_indices(n)] += 1.0
            dS = pga.DeviceArray.from_host(np.asfortranarray(S))
            pga._check(pga._lib.pga_p32_from_f64_device(dS.ptr, d0.ptr, n * n, None), "from_f64")
            pga.synchronize()
            dS.free()
        items.append(dict(kind=kind, n=n, orig=d0, a=pga.DeviceArray((n, n)),
                          ipiv=pga.DeviceArray(n, np.int32, "C") if kind == "getrf" else None,
                          info=pga.DeviceArray(1, np.int32, "C")))
    pga.synchronize()
    log(f"inputs ready: {len(items)} matrices")

    def restore():
        for it in items:
            it["a"].copy_from(it["orig"])
        pga.synchronize()

    ops = sum((2.0 / 3.0 if k == "getrf" else 1.0 / 3.0) * n ** 3 for k, n in ns)
    tasks = [dict(kind=it["kind"], n=it["n"], a=it["a"], lda=it["n"], ipiv=it["ipiv"], info=it["info"], uplo="L",
                  nb=64) for it in items]
    t0 = time.perf_counter()
    plan = pga.FactorPlan(tasks)
    default = plan.info()
    tune = plan.autotune()
    log(f"plan tuned: grid {default['grid']} -> {tune['grid']}, task order {tune['schedule']}, "
        f"{tune['tune_ms'] / 1e3:.2f} s")
    plan.capture_graph()
    t_build = time.perf_counter() - t0
    times = []
    for _ in range(max(1, min(reps, 3))):
        restore()
        times.append(_timed(pga, plan.launch_graph))
    t_graph = min(times)
    log(f"graph replay {times}")
    got = [(it["a"].to_host(), None if it["ipiv"] is None else it["ipiv"].to_host(), int(it["info"].to_host()[0]))
           for it in items]
    restore()

    def per_call():
        for it in items:
            if it["kind"] == "getrf":
                pga.getrf_device(it["n"], it["n"], it["a"], it["n"], it["ipiv"], it["info"], 64)
            else:
                pga.potrf_device("L", it["n"], it["a"], it["n"], it["info"], 64)

    t_plain = _timed(pga, per_call)
    log(f"per-call {t_plain:.3f} s")
    mismatches = 0
    for it, (a, piv, inf) in zip(items, got):
        same = np.array_equal(a, it["a"].to_host()) and inf == int(it["info"].to_host()[0])
        if piv is not None:
            same = same and np.array_equal(piv, it["ipiv"].to_host())
        mismatches += 0 if same else 1
    pinfo = plan.info()
    del plan
    for it in items:
        for key in ("orig", "a", "ipiv", "info"):
            if it[key] is not None:
                it[key].free()
    return {"workload": f"{count} Rgetrf_opencl (nb=64) + {count} lower Rpotrf_opencl (nb=64), n uniform "
                        "in [32, 512], one persistent launch replayed from a hipGraph",
            "value": round(len(items) / t_graph, 1), "unit": "factorizations/s",
            "graph_s": round(t_graph, 4), "per_call_s": round(t_plain, 4),
            "speedup_vs_per_call": round(t_plain / t_graph, 2), "gops": round(ops / t_graph / 1e9, 2),
            "plan": dict(pinfo, schedule=tune["schedule"], untuned_grid=default["grid"]),
            "plan_build_s": round(t_build, 3), "bitwise_mismatches_vs_per_call": mismatches}


def bench_batched_factor(pga, info, args):
    r = run_batched_factor(pga, args.steps)
    out = {"metric": "batched posit32 factorizations/s", "value": r["value"], "unit": "factorizations/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "posit32 (VALU: integer and exact FP64 operations)",
           "data": "synthetic: LU inputs uniform[-1,1), Cholesky inputs M*M^T/n + I, posit32",
           "config": dict(r, device=info["name"])}
    print(json.dumps(out), flush=True)


def bench_blocked(pga, info, args):
    """Factor-and-solve tasks too large for one workgroup (LU and upper Cholesky, n
    log-uniform in [513, 4096], 1-16 right-hand sides) in one blocked batch, against the
    per-call device drivers one problem after another; every factor and solution compared
    bit for bit."""
    rng = np.random.default_rng(4096)
    count = max(2, args.requests // 4)
    N = 4096
    G = pga.DeviceArray((N, N))
    pga.fill_uniform(G, 77)
    S = _spd_device(pga, N, 78)
    Gh, Sh = G.to_host(), S.to_host()
    items = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(513), math.log(4096)))))
        kind = "gesv" if i % 2 == 0 else "posv"
        nrhs = int(rng.integers(1, 17))
        a0 = pga.DeviceArray.from_host(np.asfortranarray((Gh if kind == "gesv" else Sh)[:n, :n]))
        b0 = pga.DeviceArray.from_host(np.asfortranarray(Gh[:n, N - nrhs:]))
        items.append(dict(kind=kind, n=n, nrhs=nrhs, a0=a0, b0=b0, a=pga.DeviceArray((n, n)), b=pga.DeviceArray((n, nrhs)),
                          ipiv=pga.DeviceArray(n, np.int32, "C") if kind == "gesv" else None,
                          info=pga.DeviceArray(1, np.int32, "C")))
    pga.synchronize()
    log(f"inputs ready: {count} problems, n median {int(np.median([it['n'] for it in items]))}")

    def restore():
        for it in items:
            it["a"].copy_from(it["a0"])
            it["b"].copy_from(it["b0"])
        pga.synchronize()

    ops = sum((2.0 / 3.0 if it["kind"] == "gesv" else 1.0 / 3.0) * it["n"] ** 3 for it in items)
    tasks = [dict(kind=it["kind"], n=it["n"], a=it["a"], lda=it["n"], ipiv=it["ipiv"], info=it["info"], uplo="U", nb=64,
                  b=it["b"], ldb=it["n"], nrhs=it["nrhs"]) for it in items]
    plan = pga.FactorPlan(tasks)
    times = []
    for _ in range(max(1, min(args.steps, 3))):
        restore()
        times.append(_timed(pga, plan.execute))
    t_plan = min(times)
    log(f"blocked batch {times}")
    got = [(it["a"].to_host(), it["b"].to_host(), int(it["info"].to_host()[0])) for it in items]
    restore()

    def per_call():
        for it in items:
            n = it["n"]
            if it["kind"] == "gesv":
                pga.getrf_device(n, n, it["a"], n, it["ipiv"], it["info"], 64)
                pga.getrs_device("N", n, it["nrhs"], it["a"], n, it["ipiv"], it["b"], n)
            else:
                pga.potrf_device("U", n, it["a"], n, it["info"], 64)
                pga.potrs_device("U", n, it["nrhs"], it["a"], n, it["b"], n)

    t_plain = _timed(pga, per_call)
    log(f"per-call {t_plain:.3f} s")
    mism = sum(0 if (np.array_equal(a, it["a"].to_host()) and np.array_equal(b, it["b"].to_host()) and
                     inf == int(it["info"].to_host()[0])) else 1 for it, (a, b, inf) in zip(items, got))
    out = {
        "metric": "blocked batched posit32 solves/s", "value": round(count / t_plan, 2), "unit": "solves/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "posit32 (VALU: integer and exact FP64 operations)",
        "data": "synthetic: leading blocks of a uniform[-1,1) matrix and of an SPD M*M^T/n + I matrix, posit32",
        "config": {"workload": f"{count} factor-and-solve tasks, alternating Rgetrf_opencl+Rgetrs and upper "
                               "Rpotrf_opencl+Rpotrs (nb 64), n log-uniform in [513, 4096], 1-16 RHS",
                   "plan_s": round(t_plan, 4), "per_call_s": round(t_plain, 4),
                   "speedup_vs_per_call": round(t_plain / t_plan, 2), "gops": round(ops / t_plan / 1e9, 2),
                   "bitwise_mismatches_vs_per_call": mism, "device": info["name"]},
    }
    print(json.dumps(out), flush=True)


def _spd_device(pga, N, seed):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from datagen import uniform_doubles
    M = uniform_doubles(seed, N * N).reshape(N, N)
    S = M @ M.T
    S /= N
    S[np.diag_indices(N)] += 1.0
    del M
    dS = pga.DeviceArray.from_host(np.asfortranarray(S))
    del S
    P = pga.DeviceArray((N, N))
    pga._check(pga._lib.pga_p32_from_f64_device(dS.ptr, P.ptr, N * N, None), "from_f64")
    pga.synchronize()
    dS.free()
    return P


def _serve_requests(pga, count, seed):
    """Requests of the serving workload: 50% LU / 50% SPD Cholesky, n log-uniform in
    [64, 4096], 1..16 right-hand sides uniform[-1,1), 20% re-solves against an earlier
    factor.  Matrices are leading principal blocks of one 4096 x 4096 general matrix
    (uniform[-1,1)) and one SPD matrix M*M^T/4096 + I (M uniform[-1,1)), converted to
    posit32 on the device."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from datagen import uniform_doubles

    N = 4096
    G = pga.DeviceArray((N, N))
    pga.fill_uniform(G, seed)
    Gh = G.to_host()
    M = uniform_doubles(seed + 1, N * N).reshape(N, N)
    S = M @ M.T
    S /= N
    S[np.diag_indices(N)] += 1.0
    del M
    dS = pga.DeviceArray.from_host(np.asfortranarray(S))
    del S
    P = pga.DeviceArray((N, N))
    pga._check(pga._lib.pga_p32_from_f64_device(dS.ptr, P.ptr, N * N, None), "from_f64")
    pga.synchronize()
    dS.free()
    Sh = P.to_host()
    rng = np.random.default_rng(seed + 2)
    reqs, news = [], []
    for i in range(count):
        nrhs = int(rng.integers(1, 17))
        if news and rng.random() < 0.2:
            j = news[int(rng.integers(0, len(news)))]
            n = reqs[j]["n"]
            reqs.append(dict(kind="resolve", n=n, of=j, b=None, nrhs=nrhs))
        else:
            n = int(round(math.exp(rng.uniform(math.log(64), math.log(4096)))))
            kind = "lu" if rng.random() < 0.5 else "chol"
            src = Gh if kind == "lu" else Sh
            reqs.append(dict(kind=kind, n=n, a=np.asfortranarray(src[:n, :n]), nrhs=nrhs))
            news.append(i)
    for r in reqs:
        r["b"] = np.asfortranarray(Gh[: r["n"], 4096 - r["nrhs"]:])  # uniform[-1,1) columns
    return reqs


def _serve_run(pga, reqs, mode, arrivals, batch_max_n=512, batch_slots=2, workers=4):
    """Submit reqs at the given arrival times (seconds from start); return results."""
    budget = 64 << 30
    server = pga.Server(budget, mode, batch_max_n=batch_max_n, batch_slots=batch_slots, workers=workers,
                        chol_uplo="U")
    ids = [None] * len(reqs)
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        wait = arrivals[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        if r["kind"] == "resolve":
            ids[i] = server.submit("resolve", r["n"], b=r["b"], factor=ids[r["of"]])
        else:
            ids[i] = server.submit(r["kind"], r["n"], r["a"], r["b"])
    res = [server.wait(i) for i in ids]
    wall = time.perf_counter() - t0
    st = server.stats()
    server.close()
    lat = np.array([r["t_done"] - r["t_submit"] for r in res])
    first = min(r["t_submit"] for r in res)
    span = max(r["t_done"] for r in res) - first
    ok = sum(1 for r in res if r["status"] == 0 and r["info"] == 0)
    return res, {"solves_per_s": round(len(res) / span, 2), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                 "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2), "wall_s": round(wall, 3), "ok": ok,
                 "batches": st["batches"], "batched_requests": st["batched_requests"],
                 "peak_bytes": st["bytes_peak"], "evictions": st["evictions"]}


def run_serving(pga, requests, qps):
    """Serving: capacity (every request submitted at once) and latency under a Poisson
    arrival stream at a fixed rate, for the served form (continuous batching) and for
    plain per-request launches; every solution compared bit for bit across the four runs."""
    args = argparse.Namespace(requests=requests, qps=qps)
    reqs = _serve_requests(pga, args.requests, 777)
    log(f"{len(reqs)} requests generated")
    zero = [0.0] * len(reqs)
    rng = np.random.default_rng(778)
    poisson = np.cumsum(rng.exponential(1.0 / args.qps, len(reqs))).tolist()
    cap_s, cap_s_st = _serve_run(pga, reqs, "batched", zero)
    log(f"served capacity {cap_s_st}")
    cap_p, cap_p_st = _serve_run(pga, reqs, "plain", zero)
    log(f"plain capacity {cap_p_st}")
    lat_s, lat_s_st = _serve_run(pga, reqs, "batched", poisson)
    log(f"served at {args.qps} qps {lat_s_st}")
    lat_p, lat_p_st = _serve_run(pga, reqs, "plain", poisson)
    log(f"plain at {args.qps} qps {lat_p_st}")
    mism = 0
    for runs in zip(cap_s, cap_p, lat_s, lat_p):
        xs = [r["x"] for r in runs]
        infos = {r["info"] for r in runs}
        if len(infos) != 1 or any(not np.array_equal(xs[0], x) for x in xs[1:]):
            mism += 1
    sizes = [r["n"] for r in reqs]
    return {"workload": f"{len(reqs)} solve requests: 50% LU / 50% Cholesky (upper), n log-uniform "
                        "[64, 4096], 1-16 RHS, 20% re-solves on resident factors",
            "value": cap_s_st["solves_per_s"], "unit": "solves/s",
            "n_median": int(np.median(sizes)), "resolves": sum(1 for r in reqs if r["kind"] == "resolve"),
            "served_capacity": cap_s_st, "plain_capacity": cap_p_st,
            "speedup_vs_plain": round(cap_s_st["solves_per_s"] / cap_p_st["solves_per_s"], 2),
            "qps": args.qps, "served_at_qps": lat_s_st, "plain_at_qps": lat_p_st,
            "speedup_at_qps": round(lat_s_st["solves_per_s"] / lat_p_st["solves_per_s"], 2),
            "bitwise_mismatches_across_runs": mism}


def bench_serve(pga, info, args):
    r = run_serving(pga, args.requests, args.qps)
    out = {"metric": "served posit32 solves/s", "value": r["value"], "unit": "solves/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "posit32 (VALU: integer and exact FP64 operations)",
           "data": "synthetic: general uniform[-1,1) and SPD M*M^T/n + I blocks, RHS uniform[-1,1), posit32",
           "config": dict(r, device=info["name"])}
    print(json.dumps(out), flush=True)


def bench_copycal(pga, info, args):
    """o[i] = a[i] + b[i] over 2^28 posit32 words: 2 GiB fetched, 1 GiB written per launch
    (algorithmic), the reference for scaling the HBM traffic counters."""
    n = 1 << 28
    da, db, do = (pga.DeviceArray(n, np.uint32, "C") for _ in range(3))
    pga.fill_uniform(da, 1)
    pga.fill_uniform(db, 2)
    pga.synchronize()
    ts = []
    for _ in range(max(1, args.steps)):
        ts.append(_timed(pga, lambda: pga._check(
            pga._lib.pga_p32_elementwise_device(0, da.ptr, db.ptr, do.ptr, n, None), "elementwise")))
    t = min(ts)
    out = {"metric": "streaming add GB/s", "value": round(3.0 * 4 * n / t / 1e9, 1), "unit": "GB/s",
           "n_gpus": 1, "higher_is_better": True, "config": {"workload": "posit32 add of two 1 GiB arrays",
                                                             "read_bytes": 8 * n, "write_bytes": 4 * n,
                                                             "ms": round(t * 1e3, 4)}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
