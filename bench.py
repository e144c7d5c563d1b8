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
  mixed         (FP64 runs) the same H_AC in the split-BF16 planes mode: wall-clock value,
                per-application cost inside an eigensolve, kernels and roofline;
  vumps         seconds per VUMPS iteration (and its phases) for a few iterations at
                D = 1024 from a random canonical state (disable with --no-vumps);
  time_to_galerkin  (default; --no-ttg skips it) wall seconds of VUMPS from a random canonical
                state (PCG64 seed 2, canonicalised on the device), FP64 against the mixed bulk
                arithmetics with FP64 refinement: full runs to Galerkin < 1e-8 at D = 128 (a bond
                dimension at which all three converge for this model) and the time to Galerkin
                <= 1e-4 at D = 1024 (where the iteration settles into a limit cycle near 2e-5
                instead of converging);
  time_to_galerkin_long  with --ttg: the run at --D to Galerkin < 1e-8 (long);
  idmrg         with --idmrg K: seconds per single-site IDMRG sweep over K sweeps;
  tdvp          with --tdvp K: seconds per imaginary-time TDVP step over K steps;
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
# HBM bytes per launch of the H_AC GEMMs measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (FETCH_SIZE scaled by 2.0, the factor a copy kernel of known size shows); keyed by (D, precision,
# kernel).  profiles/hac_d1024_hubbard_mfma_util_and_traffic.md holds the runs.
TRAFFIC = {(1024, "fp64", "gemm1_GLx"): 737.2 * 2 ** 20, (1024, "fp64", "gemm2_GR"): 539.5 * 2 ** 20,
           (1024, "bf16x3", "gemm1_GLx"): 309.1 * 2 ** 20, (1024, "bf16x3", "gemm2_GR"): 304.1 * 2 ** 20}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def oracle_ttg(D, maxiter):
    """The CPU baseline of the time-to-Galerkin runs: the oracle's find_groundstate (numpy
    restatement of vumps.jl:51-86) on the JW Hubbard chain (t=1, U=4, mu=2, N=2) from
    randomimps(D, [4, 4], seed 2), VUMPS tol 1e-8, on this host's cores."""
    from oracle import imps_oracle as O
    W = O.hubbard_mpo(1.0, 4.0, 2.0)
    psi = O.randomimps(D, [4, 4], 2)
    t0 = time.perf_counter()

    def cb(it, eps, f):
        if it % 10 == 0 or eps < 1e-8:
            log(f"oracle: VUMPS iter {it} eps {eps:.3e} e/site {f / 2:.12f} t {time.perf_counter() - t0:.1f}s")
    _, _, info = O.find_groundstate(psi, [W, W], O.VUMPS(tol=1e-8, maxiter=maxiter), callback=cb)
    return {"D": D, "iterations": info.niter, "converged": bool(info.converged), "itererr": info.losses[-1],
            "seconds": time.perf_counter() - t0, "threads": os.environ.get("OPENBLAS_NUM_THREADS", "default"),
            "kind": "port"}


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
    ap.add_argument("--no-ttg", action="store_true", help="skip the default time-to-Galerkin runs")
    ap.add_argument("--vumps-only", action="store_true", help="only the VUMPS iterations (for kernel traces)")
    ap.add_argument("--sweep", action="store_true",
                    help="batched parameter sweep: 64 TFIM instances, h in U[0.5,1.5], D in {32,64,128,256}")
    ap.add_argument("--sweep-workers", type=int, default=8)
    ap.add_argument("--sweep-mode", default="both", choices=["device", "threads", "both"],
                    help="--sweep: the lockstep device batch, the threaded batch, or both")
    ap.add_argument("--no-sweep", action="store_true", help="skip the parameter sweep of the default run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the fixed-point response of the default run")
    ap.add_argument("--no-realtime", action="store_true", help="skip the real-time TDVP record of the default run")
    ap.add_argument("--gradient", action="store_true",
                    help="energy gradients: TFIM D=128 de/dh and XXZ D=512 de/dDelta against finite differences")
    ap.add_argument("--vumps-iters", type=int, default=3)
    ap.add_argument("--ttg", action="store_true", help="run VUMPS to Galerkin < 1e-8 (long)")
    ap.add_argument("--ttg-maxiter", type=int, default=2000)
    ap.add_argument("--ttg-seconds", type=float, default=1e9, help="stop the --ttg run after this wall time")
    ap.add_argument("--stop-below", type=float, default=0.0,
                    help="VUMPS runs: stop once the Galerkin error is at or below this value")
    ap.add_argument("--state-out", default="", help="VUMPS runs: write the final state to this state file")
    ap.add_argument("--state-in", default="",
                    help="VUMPS runs: start from this state file instead of the random canonical state")
    ap.add_argument("--refine-below", type=float, default=1e-6, help="mixed precision: FP64 below this Galerkin error")
    ap.add_argument("--idmrg", type=int, default=0, help="also time this many single-site IDMRG sweeps")
    ap.add_argument("--tdvp", type=int, default=0, help="also time this many imaginary-time TDVP steps (dt = -0.05)")
    ap.add_argument("--gauge-eig-miniter", type=int, default=10,
                    help="gauge sweeps before the transfer-map eigenvector refresh (RightCanonical eig_miniter)")
    ap.add_argument("--eig-floor", type=float, default=0.0,
                    help="mixed precision: eigensolver tolerance floor (0: library default)")
    ap.add_argument("--calib", action="store_true", help="also run a 256 MiB copy kernel (traffic calibration)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the split-BF16 H_AC line of an FP64 run")
    ap.add_argument("--oracle-ttg", action="store_true",
                    help="CPU only: the oracle's VUMPS run to Galerkin < 1e-8 at --D (JW Hubbard, "
                         "PCG64 seed 2; long: minutes at D=64)")
    ap.add_argument("--gemm-variant", type=int, default=-1,
                    help="run FP64/FP32 GEMM tile variant V wherever valid instead of autotuning (-1)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the library runs on a "
              "single GPU and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    if args.oracle_ttg:
        print(json.dumps({"metric": METRIC, "oracle_time_to_galerkin": oracle_ttg(args.D, args.ttg_maxiter)}),
              flush=True)
        return

    import infinitempsalgorithms_amd as imps
    from infinitempsalgorithms_amd import _lib

    if args.gemm_variant >= 0:
        _lib.check(None, _lib.load().imps_set_gemm_autotune(2 + args.gemm_variant), "imps_set_gemm_autotune")
    D, d, w, N = args.D, 4, 6, 2
    H = imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=N)
    prec = PREC[args.prec]
    F = 4.0 * w * d * D ** 3
    out = {"metric": METRIC}

    if args.sweep or args.gradient:
        if args.sweep:
            out["sweep"] = parameter_sweep(args, imps, prec)
        if args.gradient:
            out["gradient"] = gradients(imps)
        print(json.dumps(out), flush=True)
        return

    # ---- H_AC throughput on device-resident data
    if not args.vumps_only:
        r = hac_throughput(args, imps, _lib, H, args.prec, D, d, w, N, F)
        out.update({
            "value": r["value"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": {"fp64": "f64", "fp32": "f32-on-f64", "bf16x3": "bf16x3-on-f64"}[args.prec],
            "data": "synthetic (random tensors and environments of the benchmark shape)",
            "config": {"workload": "H_AC apply, JW Fermi-Hubbard t=1 U=4 mu=2, N=2, d=4, w=6, D=1024"
                       if D == 1024 else f"H_AC apply, JW Fermi-Hubbard, N=2, d=4, w=6, D={D}",
                       "D": D, "d": d, "w": w, "flop_per_step": F, "parallelism": "single GPU"},
            "roofline": r["roofline"], "kernels": r["kernels"],
            # F_HAC counts the identity levels' products, which run as copies: the rate of the
            # flops actually executed, 4 (w-1) d D^3 per step
            "value_executed_flops": F * (w - 1) / w / (r["ms_per_step"] * 1e-3) / 1e12,
        })
        if r["krylov"]:
            out["krylov"] = r["krylov"]
        if r["hc"]:
            out["hc"] = r["hc"]
        if args.prec == "fp64" and not args.no_mixed:
            # the mixed-precision bulk path of the same H_AC (split-BF16 planes, FP32 accumulation)
            m = hac_throughput(args, imps, _lib, H, "bf16x3", D, d, w, N, F)
            k = m["kernels"]
            out["mixed"] = {"prec": "bf16x3", "value": m["value"], "ms_per_step": m["ms_per_step"],
                            "value_in_eigensolve": k["apply_in_eigensolve"]["tflops"] if "apply_in_eigensolve" in k else None,
                            "speedup_vs_fp64": r["ms_per_step"] / m["ms_per_step"],
                            "roofline": m["roofline"], "kernels": k}

    # ---- VUMPS outer iterations on the same configuration
    if not args.no_vumps or args.ttg or args.vumps_only:
        vumps_iterations(args, imps, H, prec, out, D, N)

    if not args.no_ttg and not args.vumps_only and not args.ttg:
        out["time_to_galerkin"] = time_to_galerkin(imps, H, N)

    if not args.no_sweep and not args.vumps_only:
        # the 64-instance TFIM parameter sweep: lockstep device batch against the threaded batch
        out["sweep"] = parameter_sweep(args, imps, PREC["fp64"])
    if not args.no_adjoint and not args.vumps_only:
        out["adjoint"] = adjoint_check(imps)
    if not args.no_realtime and not args.vumps_only:
        out["tdvp_real_time"] = realtime_tdvp(imps)

    if args.idmrg > 0:
        out["idmrg"] = idmrg_sweeps(args, imps, H, D, N)
    if args.tdvp > 0:
        out["tdvp"] = tdvp_steps(args, imps, H, D, N)

    if not args.no_cpu and not args.vumps_only:
        log("bench: CPU baseline (numpy restatement)")
        out["cpu_baseline"] = cpu_baseline(D, d, w)
    print(json.dumps(out), flush=True)


def hac_throughput(args, imps, _lib, H, pname, D, d, w, N, F):
    """Times H_AC at precision `pname` on random device-resident data; returns the wall-clock
    value, the per-kernel HIP-event timings and the roofline entry of the dominant GEMM."""
    import numpy as np
    prec = PREC[pname]
    ctx = imps.Context(0)
    ctx.set_mpo(H)
    rng = np.random.default_rng(7)
    rnd = [np.asfortranarray(rng.standard_normal((D, d, D))) for _ in range(N)]
    rc = [np.asfortranarray(rng.standard_normal((D, D))) for _ in range(N)]
    ctx.set_state(imps.CanonicalIMPS(rnd, rnd, rc, rnd))
    # random environments with the Schur structure DMRGCache gives them: GL level 0 and GR
    # level w-1 are the identity (H_AC applies those levels as copies)
    for l in range(N):
        GL = rng.standard_normal((D, w, D))
        GR = rng.standard_normal((D, w, D))
        GL[:, 0, :] = np.eye(D)
        GR[:, w - 1, :] = np.eye(D)
        ctx.set_env(l, 0, GL)
        ctx.set_env(l, 1, GR)
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
    log(f"bench: timing {args.steps} H_AC applications at D={D} ({pname})")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.call("imps_hac_apply_dev", 0, x, y, prec)
    ctx.call("imps_synchronize")
    dt = (time.perf_counter() - t0) / args.steps
    t = _lib.HacTiming()
    ctx.call("imps_bench_hac", 0, max(10, args.steps // 2), prec, ctypes.byref(t))
    ctx.call("imps_dev_free", x)
    ctx.call("imps_dev_free", y)
    hc = None
    if pname == "fp64":
        # H_C (effective.jl:72-76) on the same environments: C-sized vectors, F_HC = 4 w D^3
        # (the identity levels included, as in F_HAC); wall clock over device-resident data
        nc = D * D
        xc, yc = ctypes.c_void_p(), ctypes.c_void_p()
        ctx.call("imps_dev_alloc", nc * 8, ctypes.byref(xc))
        ctx.call("imps_dev_alloc", nc * 8, ctypes.byref(yc))
        xch = np.asfortranarray(rng.standard_normal(nc))
        ctx.call("imps_memcpy_h2d", xc, xch.ctypes.data_as(ctypes.c_void_p), nc * 8)
        for _ in range(args.warmup):
            ctx.call("imps_hc_apply_dev", 0, xc, yc, prec)
        ctx.call("imps_synchronize")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.call("imps_hc_apply_dev", 0, xc, yc, prec)
        ctx.call("imps_synchronize")
        dtc = (time.perf_counter() - t0) / args.steps
        # executed flops: the identity levels run as copies, so the GEMMs do 4 (w-1) D^3; the dense
        # form 4 w D^3 (the H_C definition's work) is kept as a separately labelled rate
        fhc = 4.0 * w * D ** 3
        fex = 4.0 * (w - 1) * D ** 3
        hc = {"ms": dtc * 1e3, "flop_executed": fex, "tflops_executed": fex / dtc / 1e12,
              "frac_of_fp64_peak": fex / dtc / 1e12 / PEAKS["fp64"],
              "flop_dense_form": fhc, "tflops_dense_form": fhc / dtc / 1e12}
        ctx.call("imps_dev_free", xc)
        ctx.call("imps_dev_free", yc)
    krylov = None
    if pname == "fp64":
        # the Krylov solvers' Gram-Schmidt sweeps (HBM-bound) on vectors of the H_AC length
        # against a 20-vector basis, operands rotated through three bases
        kt = _lib.KrylovTiming()
        nv = 20
        ctx.call("imps_bench_krylov", n, nv, 20, ctypes.byref(kt))

        def rate(ms, b):
            return {"ms": ms, "bytes": b, "gbs": b / (ms * 1e-3) / 1e9}
        krylov = {"n": n, "basis_vectors": nv,
                  "dot_sweep": rate(kt.ms_multidot, kt.bytes_multidot),
                  "second_pass_sweep": rate(kt.ms_axpy_dots, kt.bytes_axpy),
                  "norm_sweep": rate(kt.ms_axpy_norm, kt.bytes_axpy)}
    ctx.close()
    # per-launch algorithmic flops of the GEMMs as executed: 2 (w-1) d D^3 each, the identity
    # levels being copies (F keeps the full 4 w d D^3 of the H_AC definition)
    planes = t.ms_prepare > 0
    # mix bytes: FP64 in and out per channel, or (planes) FP32 in / bf16 hi+lo out for the
    # levels through the GEMMs plus the FP64 x level in and y level out
    mix_bytes = (2 * w * d * D * D * 8 if not planes
                 else 2 * ((w - 1) * d * D * D * 4 + d * D * D * 8))
    kernels = {
        "gemm1_GLx": {"ms": t.ms_gemm1, "flop": t.flop_gemm1, "tflops": t.flop_gemm1 / (t.ms_gemm1 * 1e-3) / 1e12},
        "wmix": {"ms": t.ms_wmix, "bytes": mix_bytes, "gbs": mix_bytes / (t.ms_wmix * 1e-3) / 1e9},
        "gemm2_GR": {"ms": t.ms_gemm2, "flop": t.flop_gemm2, "tflops": t.flop_gemm2 / (t.ms_gemm2 * 1e-3) / 1e12},
    }
    if planes:
        # split-BF16 planes: the environments are split once per eigensolve; inside a solve an
        # application costs ms_apply (the wall-clock loop above splits them on every call)
        kernels["apply_in_eigensolve"] = {"ms": t.ms_apply, "tflops": F / (t.ms_apply * 1e-3) / 1e12}
        kernels["env_split_per_eigensolve"] = {"ms": t.ms_prepare}
    dom = max(("gemm1_GLx", "gemm2_GR"), key=lambda k: kernels[k]["ms"])
    achieved = kernels[dom]["tflops"]
    roof = {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": PEAKS[pname], "unit": "TFLOP/s",
            "frac": achieved / PEAKS[pname], "traffic": TRAFFIC.get((D, pname, dom))}
    if pname == "bf16x3":
        # three bf16 MFMA products per FP64-equivalent multiply-add, against the dense bf16 peak
        roof.update({"achieved": 3 * achieved, "frac": 3 * achieved / PEAKS[pname],
                     "unit": "TFLOP/s (bf16 MFMA: 3 products per FP64-equivalent flop)"})
    return {"value": F / dt / 1e12, "ms_per_step": dt * 1e3, "kernels": kernels, "roofline": roof, "krylov": krylov,
            "hc": hc}


def time_to_galerkin(imps, H, N):
    """VUMPS from randomimps(D, seed 2) (canonicalised on the device) with each bulk arithmetic:
    full runs to Galerkin < 1e-8 at D = 128 and the time to Galerkin <= 1e-4 at D = 1024.  Wall
    seconds are the library's loop time (info.t_galerkin_reached, or the callback's clock for the
    1e-4 mark); the mixed runs refine in FP64 below 1e-6 or once the reduced arithmetic stalls."""
    out = {}
    for D, target, maxiter in ((128, 1e-8, 600), (1024, 1e-4, 60)):
        key = f"D{D}_to_{target:.0e}"
        rec = {}
        precs = ("fp64", "fp32", "bf16x3") if D <= 128 else ("fp64", "bf16x3")
        for pname in precs:
            ctx = imps.Context(0)
            psi, ctx = imps.randomimps(D, H, seed=2, ctx=ctx, tol=1e-12, maxiter=100)
            envs = imps.GPUCache(psi, H, ctx)
            marks = []
            t0 = time.perf_counter()

            def cb(it, eps, f):
                marks.append((it, eps, time.perf_counter() - t0))
                return eps <= target

            prec = PREC[pname]
            alg = imps.VUMPS(tol=target, maxiter=maxiter, prec=prec, refine_below=1e-6 if prec else 0.0)
            _, _, info = imps.find_groundstate(psi, H, alg, envs=envs, callback=cb)
            hit = next(((it, el) for it, eps, el in marks if eps <= target), None)
            rec[pname] = {"seconds": hit[1] if hit else None, "iterations": hit[0] if hit else info.niter,
                          "reached": hit is not None, "itererr": info.itererr,
                          "energy_per_site": info.energy / N, "prec_switch_iter": info.prec_switch_iter,
                          "hac_applies": info.hac_applies, "phases_seconds": info.timings}
            log(f"bench: time to Galerkin {target:.0e} at D={D} ({pname}): "
                f"{rec[pname]['seconds']} s, {rec[pname]['iterations']} iterations")
            ctx.close()
        for pname in precs[1:]:
            a, b = rec["fp64"]["seconds"], rec[pname]["seconds"]
            rec[pname]["speedup_vs_fp64"] = a / b if a and b else None
            rec[pname]["energy_diff_vs_fp64"] = abs(rec[pname]["energy_per_site"] - rec["fp64"]["energy_per_site"])
        out[key] = rec
    return out


def idmrg_sweeps(args, imps, H, D, N):
    """Single-site IDMRG (idmrg.jl:12-86) from the same random canonical state as the VUMPS
    timing: seconds per sweep (two eigensolves and two MPO-channel pushes per site)."""
    ctx = imps.Context(0)
    psi, ctx = imps.randomimps(D, H, seed=2, ctx=ctx, tol=1e-12, maxiter=100)
    envs = imps.GPUCache(psi, H, ctx)
    marks = []
    t0 = time.perf_counter()

    def cb(it, eps, e):
        marks.append((it, eps, time.perf_counter() - t0))
        log(f"bench: IDMRG sweep {it} drift {eps:.3e} t {marks[-1][2]:.1f}s")

    _, _, info = imps.find_groundstate(psi, H, imps.IDMRG(tol=1e-12, maxiter=args.idmrg), envs=envs, callback=cb)
    ctx.close()
    loop = marks[-1][2] if marks else 0.0
    return {"sweeps": info.niter, "seconds_per_sweep": loop / max(len(marks), 1), "loop_seconds": loop,
            "total_seconds": info.seconds, "drift": info.itererr, "energy_per_site": info.energy / N,
            "hac_applies": info.hac_applies}


def tdvp_steps(args, imps, H, D, N, dt=-0.05):
    """Imaginary-time TDVP steps (tdvp.jl:33-52) from the random canonical state: seconds per
    step (one Lanczos exponential per site for AC and for C, regauge!, gauge fix, recalculate!)."""
    ctx = imps.Context(0)
    psi, ctx = imps.randomimps(D, H, seed=2, ctx=ctx, tol=1e-12, maxiter=100)
    envs = imps.DMRGCache(psi, H, ctx=ctx)
    e0 = imps.expectationvalue(psi, H, envs) / N
    times = []
    for k in range(args.tdvp):
        t0 = time.perf_counter()
        psi, envs = imps.timestep(psi, H, dt, imps.TDVP(), envs)
        times.append(time.perf_counter() - t0)
        log(f"bench: TDVP step {k + 1} dt {dt} t {times[-1]:.2f}s")
    e1 = imps.expectationvalue(psi, H, envs) / N
    ctx.close()
    return {"steps": args.tdvp, "dt": dt, "seconds_per_step": sum(times) / len(times),
            "energy_per_site_before": e0, "energy_per_site_after": e1}


def realtime_tdvp(imps, D=64, steps=10, dt=-0.01j):
    """Real-time TDVP (tdvp.jl:33-52 with dt = -im t, the state promoted to ComplexF64 and run on the
    device as its real embedding of bond dimension 2 D): the XXX chain, two-site cell, from the
    VUMPS(D) ground state, ten steps; seconds per step, energy drift and norm (the reference's
    real-time test asserts drift < 1e-6 and |norm - 1| < 1e-8 at D = 8)."""
    m = imps.models.heisenberg_xxz(J=1.0, Delta=1.0)
    H = imps.SparseIMPO([m.bulk, m.bulk])
    ctx = imps.Context(0)
    g, envs, info = imps.find_groundstate(None, H, imps.VUMPS(D=D, tol=1e-9, maxiter=400), ctx=ctx, seed=3)
    e0 = imps.expectationvalue(g, H, envs) / 2
    psi = g
    times = []
    for k in range(steps):
        t0 = time.perf_counter()
        psi, envs = imps.timestep(psi, H, dt, imps.TDVP(), envs)
        times.append(time.perf_counter() - t0)
    e1 = imps.expectationvalue(psi, H, envs) / 2
    nrm = imps.norm(psi)
    ctx.close()
    log(f"bench: real-time TDVP XXX D={D}: {sum(times) / steps:.3f} s per step, energy drift {abs(e1 - e0):.2e}")
    return {"workload": f"XXX (Delta = 1), two-site cell, VUMPS D={D} ground state, {steps} steps dt = -0.01i",
            "complex_embedding_bond_dimension": 2 * D, "seconds_per_step": sum(times) / steps,
            "first_step_seconds": times[0], "energy_per_site": e0, "energy_drift": abs(e1 - e0),
            "norm_error": abs(nrm - 1.0), "vumps_iterations": info.niter}


def tfim_exact_energy(h, J=1.0, n=200000):
    """Ground-state energy per site of -J sum sx sx - h sum sz (free fermions):
    -(1/pi) int_0^pi sqrt(J^2 + h^2 - 2 J h cos k) dk, midpoint rule."""
    import numpy as np
    k = (np.arange(n) + 0.5) * np.pi / n
    return -float(np.mean(np.sqrt(J * J + h * h - 2 * J * h * np.cos(k))))


def parameter_sweep(args, imps, prec):
    """64 TFIM instances (N=1, d=2, w=3): h_k ~ U[0.5, 1.5] and D_k in {32, 64, 128, 256},
    both drawn from PCG64(3); instance k starts from random_tensors(D_k, [2], seed=k).
    Work = sum_k (H_AC applications of k) x 4 w d D_k^3."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(3))
    hs = rng.uniform(0.5, 1.5, 64)
    Ds = rng.choice([32, 64, 128, 256], 64)
    problems = [(imps.models.tfim_hamiltonian(J=1.0, h=float(h), nsites=1), imps.random_tensors(int(D), [2], seed=k))
                for k, (h, D) in enumerate(zip(hs, Ds))]
    alg = imps.VUMPS(tol=1e-8, maxiter=200, prec=prec, refine_below=args.refine_below,
                     gauge_eig_miniter=args.gauge_eig_miniter)
    out = {}
    energies = {}
    modes = ["device", "threads"] if args.sweep_mode == "both" else [args.sweep_mode]
    for mode in modes:
        log(f"bench: batched sweep of {len(problems)} TFIM instances, mode {mode}")
        stats = {}
        t0 = time.perf_counter()
        res = imps.find_groundstates_batched(problems, alg, nworkers=args.sweep_workers, mode=mode, stats=stats)
        wall = time.perf_counter() - t0
        flop = sum(r["hac_applies"] * 4.0 * 3 * 2 * float(D) ** 3 for r, D in zip(res, Ds))
        err = [abs(r["energy"] - tfim_exact_energy(float(h))) for r, h in zip(res, hs)]
        energies[mode] = [r["energy"] for r in res]
        out[mode] = {"instances": len(res), "workers": args.sweep_workers if mode == "threads" else 1,
                     "prec": args.prec, "wall_seconds": wall, "stats": stats,
                     "converged": sum(r["converged"] for r in res), "total_iterations": sum(r["niter"] for r in res),
                     "max_iterations": max(r["niter"] for r in res),
                     "hac_applies": sum(r["hac_applies"] for r in res), "hac_tflops": flop / wall / 1e12,
                     "max_itererr": max(r["itererr"] for r in res),
                     "energy_error_vs_exact": {"max": max(err), "median": float(np.median(err))},
                     "per_D": {int(D): {"n": int(sum(Ds == D)),
                                        "mean_seconds": float(np.mean([r["seconds"] for r, x in zip(res, Ds) if x == D])),
                                        "mean_iters": float(np.mean([r["niter"] for r, x in zip(res, Ds) if x == D]))}
                               for D in (32, 64, 128, 256)}}
        log(f"bench: sweep {mode}: {wall:.2f}s, {out[mode]['converged']} converged, max |e - exact| {max(err):.1e}")
    if len(modes) == 2:
        out["device_speedup"] = out["threads"]["wall_seconds"] / out["device"]["wall_seconds"]
        out["max_energy_diff_device_threads"] = max(abs(a - b) for a, b in zip(energies["device"],
                                                                                energies["threads"]))
    return out


def adjoint_check(imps):
    """The fixed-point response at TFIM J=1, h=1.3, D=128: d<sigma_z>/dh through the converged state
    (imps_fixedpoint_gradient) against the closed form -e''(h) of the free-fermion energy."""
    import math

    import numpy as np
    from scipy import integrate
    h = 1.3
    H = imps.models.tfim_hamiltonian(J=1.0, h=h, nsites=1)
    ctx = imps.Context(0)
    psi0, ctx = imps.randomimps(128, H, seed=0, ctx=ctx, tol=1e-12, maxiter=100)
    envs = imps.GPUCache(psi0, H, ctx)
    psi, envs, info = imps.find_groundstate(psi0, H, imps.VUMPS(tol=1e-12, maxiter=1000), envs=envs)
    t0 = time.perf_counter()
    _, dW, rinfo = imps.fixedpoint_gradient(psi, H, envs, imps.objective_mpo_1site(imps.models.sigma_z()), tol=1e-10)
    dt = time.perf_counter() - t0
    ctx.close()
    pred = float(np.sum(dW * (imps.models.tfim_hamiltonian(J=1.0, h=h + 1.0, nsites=1).Ws[0] - H.Ws[0])))
    exact = integrate.quad(lambda k: math.sin(k) ** 2 / (1 + h * h - 2 * h * math.cos(k)) ** 1.5, 0, math.pi,
                           epsabs=1e-14, epsrel=1e-14)[0] / math.pi
    log(f"bench: adjoint TFIM D=128: d<sz>/dh {pred:.12f} (closed form {exact:.12f}), {dt:.2f}s")
    return {"workload": "d<sigma_z>/dh through the VUMPS fixed point, TFIM J=1 h=1.3, D=128",
            "itererr": info.itererr, "dsz_dh": pred, "closed_form": exact, "abs_err": abs(pred - exact),
            "minres_iterations": rinfo.iterations, "relres": rinfo.relres, "seconds": dt}


def gradients(imps):
    """de/dh of TFIM (D=128, h=1.3) against -<sz> and central differences of the converged
    energy (step 1e-5), and de/dDelta of XXZ (D=512, Delta drawn from U[0.5,1.5] with PCG64(5))
    against central differences; the gradient contracts imps_energy_gradient with dW/dparameter
    (the MPOs are linear in h and Delta)."""
    import math

    import numpy as np
    from scipy import integrate
    out = {}

    def solve(H, D, seed, tol, psi0=None, maxiter=1000):
        ctx = imps.Context(0)
        if psi0 is None:
            psi0, ctx = imps.randomimps(D, H, seed=seed, ctx=ctx, tol=1e-12, maxiter=100)
        envs = imps.GPUCache(psi0, H, ctx)
        t0 = time.perf_counter()
        psi, envs, info = imps.find_groundstate(psi0, H, imps.VUMPS(tol=tol, maxiter=maxiter), envs=envs)
        log(f"bench: VUMPS D={D}: {info.niter} iterations, eps {info.itererr:.2e}, {time.perf_counter() - t0:.1f}s")
        return psi, envs, info, ctx

    def energy_at(H, D, tol, psi0, maxiter=1000):
        _, _, info, ctx = solve(H, D, 0, tol, psi0=psi0, maxiter=maxiter)
        ctx.close()
        return info.energy

    def grad_along(psi, H, Hp, envs):
        t0 = time.perf_counter()
        G = imps.energy_gradient(psi, H, envs)
        dt = time.perf_counter() - t0
        return sum(float(np.sum(g * (Wp - W))) for g, W, Wp in zip(G, H.Ws, Hp.Ws)), dt

    # TFIM, N=1, D=128
    h, dh = 1.3, 1e-5
    H = imps.models.tfim_hamiltonian(J=1.0, h=h, nsites=1)
    psi, envs, info, ctx = solve(H, 128, 0, 1e-10)
    g, dt = grad_along(psi, H, imps.models.tfim_hamiltonian(J=1.0, h=h + 1.0, nsites=1), envs)
    sz = imps.expectationvalue(psi, (0, imps.models.sigma_z()), envs)
    ep = energy_at(imps.models.tfim_hamiltonian(J=1.0, h=h + dh, nsites=1), 128, 1e-10, psi)
    em = energy_at(imps.models.tfim_hamiltonian(J=1.0, h=h - dh, nsites=1), 128, 1e-10, psi)
    fd = (ep - em) / (2 * dh)
    out["tfim_D128"] = {"h": h, "itererr": info.itererr, "de_dh": g, "minus_sz": -sz, "fd": fd,
                        "abs_err_vs_sz": abs(g + sz), "abs_err_vs_fd": abs(g - fd), "gradient_seconds": dt}
    # the fixed-point response of <sz> (imps_fixedpoint_gradient) against the closed form -e''(h)
    t0 = time.perf_counter()
    _, dW, rinfo = imps.fixedpoint_gradient(psi, H, envs, imps.objective_mpo_1site(imps.models.sigma_z()), tol=1e-10)
    dt = time.perf_counter() - t0
    pred = float(np.sum(dW * (imps.models.tfim_hamiltonian(J=1.0, h=h + 1.0, nsites=1).Ws[0] - H.Ws[0])))
    exact = integrate.quad(lambda k: math.sin(k) ** 2 / (1 + h * h - 2 * h * math.cos(k)) ** 1.5, 0, math.pi,
                           epsabs=1e-14, epsrel=1e-14)[0] / math.pi
    out["tfim_D128_response_sz"] = {"dsz_dh": pred, "closed_form": exact, "abs_err": abs(pred - exact),
                                    "minres_iterations": rinfo.iterations, "relres": rinfo.relres,
                                    "level_solves": rinfo.level_solves, "level_ops": rinfo.level_ops,
                                    "seconds": dt}
    ctx.close()
    # response of the nearest-neighbour <Sz Sz> to Delta, XXZ one-site cell, D=256 (timing)
    sz2 = np.diag([0.5, -0.5])
    Wszsz = np.zeros((3, 2, 3, 2))
    Wszsz[0, :, 0, :] = np.eye(2)
    Wszsz[2, :, 2, :] = np.eye(2)
    Wszsz[0, :, 1, :] = sz2
    Wszsz[1, :, 2, :] = sz2
    H = imps.models.heisenberg_hamiltonian(J=1.0, Delta=0.7, nsites=1)
    psi, envs, info, ctx = solve(H, 256, 2, 1e-10, maxiter=1500)
    t0 = time.perf_counter()
    val, dW, rinfo = imps.fixedpoint_gradient(psi, H, envs, Wszsz, tol=1e-10)
    dt = time.perf_counter() - t0
    pred = float(np.sum(dW * (imps.models.heisenberg_hamiltonian(J=1.0, Delta=1.7, nsites=1).Ws[0] - H.Ws[0])))
    out["xxz_D256_response_szsz"] = {"Delta": 0.7, "itererr": info.itererr, "szsz": val, "dszsz_dDelta": pred,
                                     "minres_iterations": rinfo.iterations, "relres": rinfo.relres,
                                     "level_solves": rinfo.level_solves, "level_ops": rinfo.level_ops,
                                     "seconds": dt}
    log(f"bench: response XXZ D=256: {rinfo.iterations} MINRES iterations, relres {rinfo.relres:.1e}, {dt:.1f}s")
    ctx.close()
    # XXZ, N=2, D=512
    delta = float(np.random.Generator(np.random.PCG64(5)).uniform(0.5, 1.5))
    dd = 1e-5
    H = imps.models.heisenberg_hamiltonian(J=1.0, Delta=delta, nsites=2)
    psi, envs, info, ctx = solve(H, 512, 1, 1e-9, maxiter=300)
    g, dt = grad_along(psi, H, imps.models.heisenberg_hamiltonian(J=1.0, Delta=delta + 1.0, nsites=2), envs)
    ep = energy_at(imps.models.heisenberg_hamiltonian(J=1.0, Delta=delta + dd, nsites=2), 512, 1e-9, psi, maxiter=300)
    em = energy_at(imps.models.heisenberg_hamiltonian(J=1.0, Delta=delta - dd, nsites=2), 512, 1e-9, psi, maxiter=300)
    fd = (ep - em) / (2 * dd)
    out["xxz_D512"] = {"Delta": delta, "itererr": info.itererr, "de_dDelta_cell": g, "fd": fd,
                       "abs_err_vs_fd": abs(g - fd), "gradient_seconds": dt}
    ctx.close()
    return out


def vumps_iterations(args, imps, H, prec, out, D, N):
    ctx = imps.Context(0)
    t0 = time.perf_counter()
    if args.state_in:
        log(f"bench: starting from the state file {args.state_in}")
        ctx.set_mpo(H)
        ctx.resume(args.state_in)
        psi = ctx.get_state()
    else:
        log(f"bench: canonicalising a random D={D} state (PCG64 seed 2) on the device")
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
        return (args.ttg and el > args.ttg_seconds) or eps <= args.stop_below

    alg = imps.VUMPS(tol=1e-8, maxiter=maxit, prec=prec, refine_below=args.refine_below, eig_floor=args.eig_floor,
                     gauge_eig_miniter=args.gauge_eig_miniter)
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
        out["time_to_galerkin_long"] = {args.prec: {**rec, "seconds": info.t_galerkin_reached if info.converged else None,
                                               "converged": info.converged,
                                               "prec_switch_iter": info.prec_switch_iter,
                                               "first_reached": {f"{k:.0e}": v for k, v in marks.items()}}}
    else:
        out["vumps"] = rec
    if args.state_out:
        ctx.checkpoint(args.state_out)
        log(f"bench: state written to {args.state_out}")
    ctx.close()


if __name__ == "__main__":
    main()
