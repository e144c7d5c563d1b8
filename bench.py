"""Benchmark: H_AC TFLOP/s at D = 1024 (JW Fermi-Hubbard, N = 2, d = 4, w = 6) on one MI355X,
plus (with --ttg) VUMPS time-to-Galerkin < 1e-8 on the same configuration.

A "step" is one application of the effective Hamiltonian H_AC (effective.jl:90-102) to a
device-resident (D, d, D) tensor with device-resident environments.  Algorithmic work per
step: F_HAC = 4 w d D^3 flop (dense-level two-GEMM form).  Prints one JSON line.

The cpu_baseline leg times the FP64 CPU restatement of the same contraction (numpy /
OpenBLAS on the host cores) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "VUMPS time-to-Galerkin<1e-8 and H_AC TFLOP/s at D=1024, mixed vs FP64 GPU"
PEAKS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}


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
                      f"einsum -> OpenBLAS) at D={D}, d={d}, w={w}; {dt * 1e3:.1f} ms each"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--prec", default="fp64", choices=sorted(PREC))
    ap.add_argument("--D", type=int, default=1024)
    ap.add_argument("--ttg", action="store_true", help="also run VUMPS to Galerkin < 1e-8")
    ap.add_argument("--ttg-maxiter", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU layout is "
              "described in DESIGN.md and is not part of this build", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import numpy as np

    import infinitempsalgorithms_amd as imps
    from infinitempsalgorithms_amd import _lib

    D, d, w, N = args.D, 4, 6, 2
    H = imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=N)
    ctx = imps.Context(0)
    log(f"bench: Hubbard N={N} d={d} w={w} D={D}: canonicalising a random state on the device")
    t0 = time.perf_counter()
    psi, _ = imps.randomimps(D, H, seed=2, ctx=ctx, tol=1e-12, maxiter=50)
    log(f"bench: state ready in {time.perf_counter() - t0:.1f}s; building random device environments")
    rng = np.random.default_rng(7)
    for l in range(N):
        ctx.set_env(l, 0, rng.standard_normal((D, w, D)))
        ctx.set_env(l, 1, rng.standard_normal((D, w, D)))
    prec = PREC[args.prec]
    n = D * d * D
    x = ctypes.c_void_p()
    y = ctypes.c_void_p()
    ctx.call("imps_dev_alloc", n * 8, ctypes.byref(x))
    ctx.call("imps_dev_alloc", n * 8, ctypes.byref(y))
    xh = np.asfortranarray(rng.standard_normal(n))
    ctx.call("imps_memcpy_h2d", x, xh.ctypes.data_as(ctypes.c_void_p), n * 8)
    for _ in range(args.warmup):
        ctx.call("imps_hac_apply_dev", 0, x, y, prec)
    ctx.call("imps_synchronize")
    log(f"bench: timing {args.steps} H_AC applications ({args.prec})")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.call("imps_hac_apply_dev", 0, x, y, prec)
    ctx.call("imps_synchronize")
    dt = (time.perf_counter() - t0) / args.steps
    F = 4.0 * w * d * D ** 3
    value = F / dt / 1e12
    # per-kernel durations on the library's stream (HIP events)
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
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32-on-f64", "bf16x3": "bf16x3-on-f64"}[args.prec],
        "data": "synthetic (random canonical state, random environments)",
        "config": {"workload": "H_AC apply, JW Fermi-Hubbard t=1 U=4 mu=2, N=2, d=4, w=6",
                   "D": D, "flop_per_step": F, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None},
        "kernels": kernels,
    }
    if args.ttg:
        log("bench: VUMPS to Galerkin < 1e-8 (FP64) from the canonicalised random state")
        ctx.call("imps_envs_reset")
        ctx.set_state(psi)
        envs = imps.GPUCache(psi, H, ctx)
        _, _, info = imps.find_groundstate(psi, H, imps.VUMPS(tol=1e-8, maxiter=args.ttg_maxiter), envs=envs)
        out["time_to_galerkin"] = {"fp64": {"seconds": info.t_galerkin_reached if info.converged else None,
                                            "loop_seconds": info.seconds, "niter": info.niter,
                                            "itererr": info.itererr, "energy_per_site": info.energy / N,
                                            "hac_applies": info.hac_applies}}
    if not args.no_cpu:
        log("bench: CPU baseline (numpy restatement)")
        out["cpu_baseline"] = cpu_baseline(D, d, w)
    ctx.call("imps_dev_free", x)
    ctx.call("imps_dev_free", y)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
