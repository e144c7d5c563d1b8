"""Benchmark on one MI355X: H_AC TFLOP/s at D = 1024 for the JW Fermi-Hubbard chain
(t = 1, U = 4, mu = 2, N = 2, d = 4, w = 6), plus VUMPS outer-loop timings on the same
configuration.

A "step" is one application of the effective Hamiltonian H_AC (effective.jl:90-102) to a
device-resident (D, d, D) tensor with device-resident environments.  Algorithmic work per
step: F_HAC = 4 w d D^3 flop (dense-level two-GEMM form).  `value` is F_HAC / (wall time
per step), inputs resident in HBM, timed with a device synchronisation on both sides.

Also reported:
  roofline      the dominant GEMM of H_AC against the FP64 MFMA peak (HIP events on the
                library's stream);
  vumps         seconds per VUMPS iteration (and its phases) for a few iterations at
                D = 1024 from a random canonical state (disable with --no-vumps);
  time_to_galerkin  with --ttg: the full run to Galerkin < 1e-8 (long);
  cpu_baseline  the FP64 CPU restatement of H_AC (numpy / OpenBLAS) on a bounded sample.
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "VUMPS time-to-Galerkin<1e-8 and H_AC TFLOP/s at D=1024, mixed vs FP64 GPU"
PEAKS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}
# HBM bytes per launch of the dominant H_AC GEMM measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (FETCH_SIZE scaled by 1.995 from a copy kernel of known size); keyed by
# (D, precision, kernel).  profiles/hac_d1024_hubbard_traffic.md holds the runs.
TRAFFIC = {(1024, "fp64", "gemm1_GLx"): 662.3 * 2 ** 20, (1024, "fp64", "gemm2_GR"): 606.7 * 2 ** 20}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(D, d, w, seconds=15.0):
    import numpy as np
    from oracle import imps_oracle as oracle
    rng = np.random.default_rng(0)
    W = oracle.hubbard_mpo()
    GL = rng.standard_normal((D, w, D))
    GR = rng.standard_normal((D, w, D))
    x = rng.standard_normal((D, d, D))
    oracle.hac_apply_dense(GL, W, GR, x)  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        oracle.hac_apply_dense(GL, W, GR, x)
        n += 1
        if time.perf_counter() - t0 > seconds or n >= 20:
            break
    dt = (time.perf_counter() - t0) / n
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 4 * w * d * D ** 3 / dt / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": f"{n} H_AC applications of the numpy FP64 restatement (oracle/imps_oracle.py, "
                      f"einsum -> OpenBLAS, {cores} threads) at D={D}, d={d}, w={w}; {dt * 1e3:.1f} ms each"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--prec", default="fp64", choices=sorted(PREC))
    ap.add_argument("--D", type=int, default=1024)
    ap.add_argument("--no-vumps", action="store_true", help="skip the VUMPS iteration timing")
    ap.add_argument("--vumps-iters", type=int, default=3)
    ap.add_argument("--ttg", action="store_true", help="run VUMPS to Galerkin < 1e-8 (long)")
    ap.add_argument("--ttg-maxiter", type=int, default=2000)
    ap.add_argument("--ttg-seconds", type=float, default=1e9, help="stop the --ttg run after this wall time")
    ap.add_argument("--refine-below", type=float, default=1e-6, help="mixed precision: FP64 below this Galerkin error")
    ap.add_argument("--calib", action="store_true", help="also run a 256 MiB copy kernel (traffic calibration)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); how the path would shard "
              "over several GPUs is described in DESIGN.md and is not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import infinitempsalgorithms_amd as imps
    from infinitempsalgorithms_amd import _lib

    D, d, w, N = args.D, 4, 6, 2
    H = imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=N)
    prec = PREC[args.prec]
    F = 4.0 * w * d * D ** 3
    out = {"metric": METRIC}

    # ---- H_AC throughput on device-resident data
    ctx = imps.Context(0)
    ctx.set_mpo(H)
    rng = np.random.default_rng(7)
    rnd = [np.asfortranarray(rng.standard_normal((D, d, D))) for _ in range(N)]
    rc = [np.asfortranarray(rng.standard_normal((D, D))) for _ in range(N)]
    ctx.set_state(imps.CanonicalIMPS(rnd, rnd, rc, rnd))
    for l in range(N):
        ctx.set_env(l, 0, rng.standard_normal((D, w, D)))
        ctx.set_env(l, 1, rng.standard_normal((D, w, D)))
    n = D * d * D
    x, y = ctypes.c_void_p(), ctypes.c_void_p()
    ctx.call("imps_dev_alloc", n * 8, ctypes.byref(x))
    ctx.call("imps_dev_alloc", n * 8, ctypes.byref(y))
    xh = np.asfortranarray(rng.standard_normal(n))
    ctx.call("imps_memcpy_h2d", x, xh.ctypes.data_as(ctypes.c_void_p), n * 8)
    if args.calib:
        nb = 256 << 20
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        ctx.call("imps_dev_alloc", nb, ctypes.byref(a))
        ctx.call("imps_dev_alloc", nb, ctypes.byref(b))
        for _ in range(3):
            ctx.call("imps_dev_copy", b, a, nb)
        ctx.call("imps_synchronize")
        ctx.call("imps_dev_free", a)
        ctx.call("imps_dev_free", b)
    for _ in range(args.warmup):
        ctx.call("imps_hac_apply_dev", 0, x, y, prec)
    ctx.call("imps_synchronize")
    log(f"bench: timing {args.steps} H_AC applications at D={D} ({args.prec})")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.call("imps_hac_apply_dev", 0, x, y, prec)
    ctx.call("imps_synchronize")
    dt = (time.perf_counter() - t0) / args.steps
    value = F / dt / 1e12
    t = _lib.HacTiming()
    ctx.call("imps_bench_hac", 0, max(10, args.steps // 2), prec, ctypes.byref(t))
    g1 = 2.0 * w * d * D ** 3
    kernels = {
        "gemm1_GLx": {"ms": t.ms_gemm1, "tflops": g1 / (t.ms_gemm1 * 1e-3) / 1e12},
        "wmix": {"ms": t.ms_wmix, "gbs": 2 * w * d * D * D * 8 / (t.ms_wmix * 1e-3) / 1e9},
        "gemm2_GR": {"ms": t.ms_gemm2, "tflops": g1 / (t.ms_gemm2 * 1e-3) / 1e12},
    }
    dom = max(("gemm1_GLx", "gemm2_GR"), key=lambda k: kernels[k]["ms"])
    achieved = kernels[dom]["tflops"]
    peak = PEAKS[args.prec]
    ctx.call("imps_dev_free", x)
    ctx.call("imps_dev_free", y)
    ctx.close()
    out.update({
        "value": value, "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32-on-f64", "bf16x3": "bf16x3-on-f64"}[args.prec],
        "data": "synthetic (random tensors and environments of the benchmark shape)",
        "config": {"workload": "H_AC apply, JW Fermi-Hubbard t=1 U=4 mu=2, N=2, d=4, w=6, D=1024"
                   if D == 1024 else f"H_AC apply, JW Fermi-Hubbard, N=2, d=4, w=6, D={D}",
                   "D": D, "d": d, "w": w, "flop_per_step": F, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": TRAFFIC.get((D, args.prec, dom))},
        "kernels": kernels,
    })

    # ---- VUMPS outer iterations on the same configuration
    if not args.no_vumps or args.ttg:
        log(f"bench: canonicalising a random D={D} state (PCG64 seed 2) on the device")
        ctx = imps.Context(0)
        t0 = time.perf_counter()
        psi, ctx = imps.randomimps(D, H, seed=2, ctx=ctx, tol=1e-12, maxiter=100)
        t_setup = time.perf_counter() - t0
        envs = imps.GPUCache(psi, H, ctx)
        maxit = args.ttg_maxiter if args.ttg else args.vumps_iters
        tstart = time.perf_counter()

        trace = []

        def cb(it, eps, f):
            el = time.perf_counter() - tstart
            trace.append((it, eps, el))
            log(f"bench: VUMPS iter {it} eps {eps:.3e} e/site {f / N:.12f} t {el:.1f}s")
            return args.ttg and el > args.ttg_seconds

        alg = imps.VUMPS(tol=1e-8, maxiter=maxit, prec=prec, refine_below=args.refine_below)
        _, _, info = imps.find_groundstate(psi, H, alg, envs=envs, callback=cb)
        rec = {"setup_seconds": t_setup, "iterations": info.niter, "loop_seconds": info.seconds,
               "seconds_per_iteration": info.seconds / max(info.niter, 1), "itererr": info.itererr,
               "energy_per_site": info.energy / N, "hac_applies": info.hac_applies,
               "hc_applies": info.hc_applies, "phases_seconds": info.timings}
        if args.ttg:
            marks = {}
            for it, eps, el in trace:
                for thr in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8):
                    if eps <= thr and thr not in marks:
                        marks[thr] = {"iter": it, "seconds": el}
            out["time_to_galerkin"] = {args.prec: {**rec, "seconds": info.t_galerkin_reached if info.converged else None,
                                                   "converged": info.converged,
                                                   "prec_switch_iter": info.prec_switch_iter,
                                                   "first_reached": {f"{k:.0e}": v for k, v in marks.items()}}}
        else:
            out["vumps"] = rec
        ctx.close()

    if not args.no_cpu:
        log("bench: CPU baseline (numpy restatement)")
        out["cpu_baseline"] = cpu_baseline(D, d, w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
