/*
 * imps.h — C ABI of the MI355X (gfx950) VUMPS library (libimps.so).
 *
 * The entry points replace the Julia callables of InfiniteMPSAlgorithms.jl's VUMPS
 * path; each declaration cites the reference interface it stands in for (paths are
 * relative to the reference repository root).  A host language binds them with its
 * FFI (Julia `ccall`, Python `ctypes`).
 *
 * Conventions
 *  - Every call returns an int status: IMPS_OK (0) or a negative IMPS_E_* code;
 *    imps_last_error() describes the last failure on a context.  Nothing throws or
 *    longjmps across the ABI.
 *  - Tensors cross the ABI as column-major float64 arrays (Julia layout):
 *      MPS tensor   A[a,s,b]      at a + D*(s + d*b)            (D, d, D)
 *      bond matrix  C[a,b]        at a + D*b                    (D, D)
 *      MPO tensor   W[wl,u,wr,dd] at wl + w*(u + d*(wr + w*dd)) (w, d, w, d)
 *      environment  G[a,k,b]      at a + D*(k + w*b)            (D, w, D)
 *    GL = leftenv is (bra, w, ket), GR = rightenv is (ket, w, bra).
 *  - Sites are 0-based here (the reference is 1-based); site indices wrap
 *    periodically like PeriodicVector (src/utility.jl:16-33).
 *  - Host pointers are owned by the caller and only read/written during the call;
 *    the context owns all device memory.  Entry points with the suffix `_dev` take
 *    device pointers and run on the context's stream.
 *  - A context is not thread-safe: use one context per host thread.
 *  - Scope: real float64 data, a uniform bond dimension D, square Schur MPO tensors
 *    with the same level count w on every site.
 */
#ifndef IMPS_H
#define IMPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct imps_ctx imps_ctx;

enum {
  IMPS_OK = 0,
  IMPS_E_ARG = -1,        /* invalid argument / dimension mismatch (DimensionMismatch, ArgumentError) */
  IMPS_E_STATE = -2,      /* required model/state/environments not set */
  IMPS_E_HIP = -3,        /* HIP runtime failure */
  IMPS_E_SOLVER = -4,     /* rocSOLVER/rocBLAS failure */
  IMPS_E_UNSUPPORTED = -5, /* configuration outside the supported scope */
  IMPS_E_IO = -6           /* file missing, unreadable, of another kind, truncated or corrupted */
};

/* Arithmetic of the bulk contractions (H_AC, H_C, transfer pushes). */
enum {
  IMPS_PREC_FP64 = 0,  /* FP64 MFMA */
  IMPS_PREC_FP32 = 1,  /* FP32 MFMA on FP64 data, FP32 accumulation */
  IMPS_PREC_BF16X3 = 2 /* split-BF16 three-term products, FP32 accumulation */
};

enum { IMPS_SIDE_LEFT = 0, IMPS_SIDE_RIGHT = 1 };

/* ---- lifecycle ---------------------------------------------------------------- */

/* Creates a context on HIP device `device` with its own stream. */
int imps_create(int device, imps_ctx** out);
int imps_destroy(imps_ctx* ctx);
/* Message of the last failed call on `ctx` (or of the last failed imps_create when
 * ctx is NULL).  Valid until the next call on the same context. */
const char* imps_last_error(const imps_ctx* ctx);

/* ---- model and state ---------------------------------------------------------- */

/* SparseIMPO (src/operators/mpohamiltonian.jl:33-56) as the dense Schur tensors of
 * tompotensors(H) (mpohamiltonian.jl:150): W[site] has shape (w, d, w, d).
 * Levels 0 and w-1 must be the unit levels (W[0,:,0,:] = W[w-1,:,w-1,:] = I). */
int imps_set_mpo(imps_ctx* ctx, int nsites, int w, int d, const double* const* W);
/* SparseIMPO with rectangular Schur tensors and per-site physical dimensions
 * (src/operators/mpohamiltonian.jl:26-31 and 57-66, the level-chain closure; phydims at :88):
 * W[l] has shape (levels[l-1], dims[l], levels[l], dims[l]) (levels[l] = bonddim(H, l), the
 * levels of the bond right of site l; levels[-1] = levels[nsites-1]), with its unit levels at
 * (0, 0) and (levels[l-1]-1, levels[l]-1).  The chain is embedded exactly into the square form
 * (w = max levels, d = max dims) on which every other call runs: bond level i < levels - 1 keeps
 * its index, the closing level goes to w - 1, and physical indices s >= dims[l] form an empty
 * sector (identity on the unit levels, zero elsewhere) that no operation of the library
 * populates.  States, environments and gradients are then exchanged in that square form (states
 * with zeros in the empty sectors: imps_set_state and imps_set_state_raw refuse others with
 * IMPS_E_ARG).  Closure (levels agree bond by bond) holds by construction of the arguments. */
int imps_set_mpo_chain(imps_ctx* ctx, int nsites, const int* dims, const int* levels, const double* const* W);

/* CanonicalIMPS (src/states/canonicalmps.jl:17-40): AL, AR, AC of shape (D, d, D)
 * and C of shape (D, D) per site; C[l] is the bond right of site l. */
int imps_set_state(imps_ctx* ctx, int nsites, int D, const double* const* AL,
                   const double* const* AR, const double* const* C, const double* const* AC);
/* Unit cell, bond and physical dimension of the context's state (IMPS_E_STATE without one); D is
 * the largest bond dimension when they differ from bond to bond. */
int imps_state_dims(imps_ctx* ctx, int* nsites, int* D, int* d);

/* CanonicalIMPS with bond dimensions that differ from bond to bond (canonicalmps.jl:17-40; the
 * per-site Dl/Dr of DMRGCache, envs.jl:327-343): chi[l] is the dimension of the bond right of
 * site l, D = max chi, and every tensor is passed at (D, d, D) (C at (D, D)) with zeros outside
 * its active block (rows chi[l-1], columns chi[l]).  Each site must admit a full-rank tensor
 * (chi[l] <= d chi[l-1], chi[l-1] <= d chi[l]).  The library keeps the zeros: the effective
 * operators, transfer maps and Krylov solvers map such tensors to such tensors, and the QRpos /
 * LQpos factorisations of the gauge steps act on the active blocks only, so VUMPS, IDMRG, TDVP and
 * the observables preserve the profile.  The environments are returned at (D, w, D); their
 * active block is leftenv(l): chi[l-1], rightenv(l): chi[l]. */
int imps_set_state_nonuniform(imps_ctx* ctx, int nsites, int D, const int* chi, const double* const* AL,
                              const double* const* AR, const double* const* C, const double* const* AC);

/* CanonicalIMPS(As) for raw tensors with bond dimensions chi (canonicalmps.jl:151-170):
 * gaugefix!(:LR) from C0 = I (the active block) as imps_set_state_raw. */
int imps_set_state_raw_nonuniform(imps_ctx* ctx, int nsites, int D, const int* chi, const double* const* A,
                                  double tol, int maxiter);

/* bonddim(ψ, l) for every bond (chi[l]: the bond right of site l; all equal for a uniform state). */
int imps_bond_dims(imps_ctx* ctx, int* chi);
int imps_get_state(imps_ctx* ctx, double* const* AL, double* const* AR, double* const* C,
                   double* const* AC);

/* CanonicalIMPS(As) (src/states/canonicalmps.jl:151-170): uploads raw tensors and
 * brings them to the mixed-canonical form on the device with gaugefix!(order = :LR)
 * (src/states/ortho.jl:64-99), tolerance `tol`, at most `maxiter` sweeps. */
int imps_set_state_raw(imps_ctx* ctx, int nsites, int D, const double* const* A, double tol,
                       int maxiter);

/* ---- effective Hamiltonians (src/algorithms/groundstates/effective.jl) -------- */

/* Replaces AC_hamiltonian(site, ψ, H, ψ, envs)(x) (effective.jl:66-68, 90-102):
 * y[b,u,c] = Σ GL[b,i,a] x[a,s,bb] W[i,u,j,s] GR[bb,j,c] with GL = leftenv(site),
 * GR = rightenv(site).  x, y: (D, d, D). */
int imps_hac_apply(imps_ctx* ctx, int site, const double* x, double* y, int prec);
int imps_hac_apply_dev(imps_ctx* ctx, int site, const double* d_x, double* d_y, int prec);
/* No counterpart in the reference: imps_hac_apply_dev for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * calls no hipMalloc/hipFree and no allocating library routine: the H_AC intermediates and the
 * GEMM split-K workspace live in the library's file-scope device scratch of 2^24 doubles (128 MiB,
 * one copy on each GPU the library runs on, held until the process ends), and GEMM tiles are the
 * autotuned ones already known or the fixed heuristic (never newly tuned), so the result equals
 * imps_hac_apply_dev's bit for bit.  Limit: 2 D w d D plus the split-K workspace must fit the scratch
 * (e.g. D = 512 with w d = 24), else IMPS_E_ARG.  prec: FP64 or FP32 (IMPS_E_UNSUPPORTED for
 * split-BF16).  Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * (several contexts or streams at once) race. */
int imps_hac_apply_noalloc(imps_ctx* ctx, int site, const double* d_x, double* d_y, int prec);

/* Replaces C_hamiltonian(site, ...)(x) (effective.jl:52-54, 72-76):
 * y[α,β] = Σ GL[α,w,α'] x[α',β'] GR[β',w,β] with GL = leftenv(site+1),
 * GR = rightenv(site).  x, y: (D, D). */
int imps_hc_apply(imps_ctx* ctx, int site, const double* x, double* y, int prec);
int imps_hc_apply_dev(imps_ctx* ctx, int site, const double* d_x, double* d_y, int prec);

/* ---- transfer matrices and environments --------------------------------------- */

/* TransferMatrix(ψ.AL, ψ.AL; side=:left) * v or TransferMatrix(ψ.AR, ψ.AR;
 * side=:right) * v over the unit cell (src/transfermatrix.jl:139-166), v: (D, D). */
int imps_transfer_apply(imps_ctx* ctx, int side, const double* v, double* out);

/* regularize!(v, lvec, rvec) (src/algorithms/groundstates/envs.jl:28-32):
 * v -= (Σ lvec[a,b] v[b,a]) * rvec, all (n, n), in place on the host array v. */
int imps_regularize(imps_ctx* ctx, int n, double* v, const double* lvec, const double* rvec);

/* DMRGCache(ψ, H; tol, maxiter, krylovdim) (envs.jl:321-518) when no environments
 * exist yet, otherwise recalculate!(envs, ψ, H; tol) (envs.jl:530-546): the previous
 * levels are the GMRES starting vectors. */
int imps_envs_recalculate(imps_ctx* ctx, double tol, int maxiter, int krylovdim);
/* Drops the environments so the next recalculation starts from zero vectors. */
int imps_envs_reset(imps_ctx* ctx);

/* leftenv(envs, site) / rightenv(envs, site) (src/environments.jl:36-39):
 * (D, w, D) arrays.  imps_set_env overrides one environment (for tests/drivers). */
int imps_get_env(imps_ctx* ctx, int site, int side, double* out);
int imps_set_env(imps_ctx* ctx, int site, int side, const double* in);

/* ---- Krylov solvers on device-resident vectors --------------------------------- */

typedef struct imps_krylov_params {
  double tol;    /* absolute residual tolerance */
  int krylovdim; /* Defaults.krylovdim = 30 */
  int maxiter;   /* Defaults.maxiter = 300 (restarts) */
  int eager;     /* 1: test convergence after every expansion */
  int prec;      /* IMPS_PREC_* of the operator applications */
} imps_krylov_params;

typedef struct imps_krylov_info {
  int converged;
  int numiter; /* restarts + 1 */
  int numops;  /* operator applications */
  double normres;
} imps_krylov_info;

/* fixedpoint(AC_hamiltonian(site,...), ψ.AC[site], :SR, Lanczos) (src/utility.jl:
 * 180-186): the lowest eigenpair; the eigenvector (unit norm, sign fixed so that
 * <x0, v> >= 0) is written to `vec` (host, (D, d, D)) when vec != NULL and kept on
 * the device as the pending AC of `site`. */
int imps_eigsolve_ac(imps_ctx* ctx, int site, const imps_krylov_params* p, double* lambda,
                     double* vec, imps_krylov_info* info);
/* Same for C_hamiltonian(site, ...) and ψ.C[site] (vumps.jl:14-15). */
int imps_eigsolve_c(imps_ctx* ctx, int site, const imps_krylov_params* p, double* lambda,
                    double* vec, imps_krylov_info* info);

/* regauge!(AC, C; alg = QRpos()) (src/states/ortho.jl:214-219) of the pending AC and
 * C of `site` (from imps_eigsolve_ac/c): AL = Q(QRpos(AC)) * Q(QRpos(C))^T, kept as
 * the pending AL of `site` and written to `al` (host) when al != NULL. */
int imps_regauge(imps_ctx* ctx, int site, double* al);

/* gauge_step! (src/algorithms/groundstates/vumps.jl:23-32): AL <- pending ALs,
 * gaugefix!(ψ, AL, C[end]; order = :R, tol, maxiter) (ortho.jl:95-99, 166-202),
 * then AC = AL*C.  The pending ALs are those of the last imps_regauge, or the state's own ALs
 * after the state was set.  `sweeps` receives the number of LQ sweeps (may be NULL). */
int imps_gaugefix_right(imps_ctx* ctx, double tol, int maxiter, int* sweeps);

/* calc_galerkin(ψ, H, envs) (vumps.jl:111-124). */
int imps_galerkin(imps_ctx* ctx, double* eps);
/* expectationvalue(ψ, H::SparseIMPO, envs) (src/observables/expval.jl:93-112):
 * unit-cell total energy divided by dot(ψ, ψ) (canonicalmps.jl:262-269). */
int imps_energy(imps_ctx* ctx, double* e);
/* Gradient of the unit-cell energy of expectationvalue(ψ, H, envs) (expval.jl:93-112) with
 * respect to every MPO tensor at fixed ψ and environments: the reverse of the H_AC
 * contraction (effective.jl:90-102),
 *   dW[site][i,u,j,s] = Σ GL[b,i,a] AC[a,s,bb] GR[bb,j,c] AC[b,u,c] / dot(ψ, ψ),
 * i.e. ⟨AC| GL_i ⊗ δW ⊗ GR_j |AC⟩.  At a converged VUMPS state ψ is stationary, so for
 * every non-structural Schur block this is the derivative of the ground-state energy
 * (Hellmann–Feynman); the unit corners (0,0) and (w-1,w-1) are structural.  The reference
 * has no automatic differentiation.  dW: nsites host arrays (w, d, w, d). */
int imps_energy_gradient(imps_ctx* ctx, double* const* dW);

/* ---- fixed-point response (adjoint of the VUMPS fixed point) -------------------------
 * The reference has no derivative of its fixed point; these entry points differentiate it with
 * the reference's own operators: H_AC (src/algorithms/groundstates/effective.jl:90-102), the
 * slice pushes (envs.jl:104-114), regularize! (envs.jl:28-32), linsolve (envs.jl:14-19) and the
 * level order of DMRGCache (envs.jl:321-518).  One-site unit cell, converged state, environments
 * of the Hamiltonian current (imps_envs_recalculate).
 *
 * For an objective given as a Schur MPO Wobj (wo, d, wo, d) (its value per site is
 * expectationvalue(ψ, Wobj), expval.jl:93-112; a one-site observable O is the wo = 2 MPO with O
 * in block (0, 1)) the response of the fixed point to the Hamiltonian tensor is
 *   dO/dW = -2 d/dW <Φ(L)| H[W] |ψ>,  (H_eff + K) L = g = P[H_AC^O(AC)],
 * H_eff the zero-momentum tangent-space Hamiltonian (one-defect environments solved as device
 * linsolves), K the curvature of the uniform-MPS manifold, the system solved by MINRES on the
 * device, and the sandwich the reverse H_AC contraction plus the adjoint environments (bra defect,
 * ket ψ).  The total derivative of <O> is this plus the explicit dependence of Wobj on W (none
 * for a fixed observable).  For the energy itself (Wobj = W) g is the Galerkin residual and the
 * response vanishes: imps_energy_gradient is then the whole derivative. */
typedef struct imps_response_info {
  int iterations;          /* MINRES iterations */
  int converged;           /* the MINRES estimate reached tol */
  double relres;           /* |(H_eff + K) L - g| / |g| of the returned L (one more application) */
  double estimate;         /* MINRES residual estimate relative to |g| */
  long hessian_applies;    /* applications of H_eff + K */
  long level_solves;       /* GMRES solves of defect-environment levels */
  long level_ops;          /* their operator applications */
  long unconverged_solves; /* level solves that stopped at maxiter */
} imps_response_info;

/* value <- <Wobj> per site, dW (w, d, w, d) <- dO/dW as above.  tol: MINRES tolerance relative to
 * |g|; rcond: relative cutoff of the pseudo-inverse of C (Schmidt values below rcond * s_max carry
 * no curvature term).  IMPS_E_UNSUPPORTED for nsites != 1 or w*d > 32. */
int imps_fixedpoint_gradient(imps_ctx* ctx, int wo, int d, const double* Wobj, double tol, int maxiter,
                             double rcond, double* value, double* dW, imps_response_info* info);

/* out <- P (H_eff + K) P B for a host tensor B (D, d, D) (the operator of the response system,
 * P = 1 - AL AL^T); tol: relative tolerance of its level solves. */
int imps_tangent_hessian_apply(imps_ctx* ctx, double rcond, double tol, const double* B, double* out);

/* expectationvalue(ψ, (site,) => O) (expval.jl:33-37), O: (d, d). */
int imps_local_expectation(imps_ctx* ctx, int site, const double* O, double* val);
/* expectationvalue(ψ, site => O) (expval.jl:33-37) for a ComplexF64 d x d operator O (column-major,
 * re and im interleaved), on real and complex states: val[0] = Re <O>, val[1] = Im <O> (the value
 * is real for a Hermitian O).  On a complex state this is the entry point for Hermitian operators
 * with an imaginary part (Sy, currents), which imps_local_expectation refuses. */
int imps_local_expectation_complex(imps_ctx* ctx, int site, const double* O, double* val);

/* ---- the VUMPS driver ------------------------------------------------------------ */

/* finalize(iter, ψ, operator, envs) -> (ψ, envs) of VUMPS/IDMRG (algdefs.jl:34, vumps.jl:75,
 * idmrg.jl:60): called once per iteration with the context, after recalculate! (VUMPS) or the sweep
 * (IDMRG) and before the iteration's error is computed.  It may read or replace the state and the
 * environments through this ABI (imps_get_state / imps_set_state / imps_set_env ...); environments
 * it leaves invalid (imps_set_state drops them) are recalculated at the iteration's tolerance, warm
 * starts lost.  Return IMPS_OK to continue, a positive value to stop after this iteration, or a
 * negative IMPS_E_* code to abort the run with that status.  The unit cell must not change. */
typedef int (*imps_finalize_cb)(int iter, imps_ctx* ctx, void* user);

/* VUMPS(; tol, maxiter, verbosity, alg_gauge, alg_eigsolve, alg_environments, finalize)
 * (src/algorithms/algdefs.jl:26-35) with the Defaults of src/utility.jl:308-364.  Each of
 * alg_eigsolve, alg_gauge and alg_environments is either a DynamicTol (utility.jl:136-171; its
 * tolerance becomes clamp(eps * factor / sqrt(max(iter, 1)), tol_min, tol_max) every iteration) or,
 * with *_dynamic = 0, a fixed-tolerance solver (updatetol is then the identity, utility.jl:151). */
typedef struct imps_vumps_params {
  double tol;             /* Defaults.tol = 1e-10 (Galerkin target) */
  int maxiter;            /* Defaults.maxiter = 300 */
  int verbosity;          /* 0: silent, >0: one line per iteration on stdout */
  double tol_min;         /* alg_eigsolve DynamicTol tol_min = 1e-14 */
  double tol_max;         /* alg_eigsolve DynamicTol tol_max = 1e-4 */
  double eigs_tolfactor;  /* 1e-3 */
  double gauge_tolfactor; /* 1e-6 */
  double envs_tolfactor;  /* 1e-4 */
  int krylovdim;          /* alg_eigsolve Lanczos krylovdim = 30 */
  int eig_maxiter;        /* alg_eigsolve Lanczos maxiter = 300 */
  int gauge_maxiter;      /* alg_gauge maxiter = 300 */
  int envs_maxiter;       /* DMRGCache maxiter = 300 (recalculate! passes only tol) */
  int prec;               /* IMPS_PREC_* of the bulk contractions */
  double refine_below;    /* mixed precision: switch to FP64 once eps < refine_below */
  double eig_floor;       /* mixed precision: eigensolver tolerance floor relative to |lambda| in the
                             reduced-precision phase (0: 1e-6 for FP32, 1e-4 for split-BF16) */
  int eig_dynamic;        /* 1: alg_eigsolve is a DynamicTol (default); 0: fixed tolerance eig_tol */
  int eig_eager;          /* Lanczos eager = 1 */
  double eig_tol;         /* fixed eigensolver tolerance (Defaults.tol = 1e-10) */
  int gauge_dynamic;      /* 1: alg_gauge is a DynamicTol (default) */
  double gauge_tol;       /* fixed gauge tolerance (Defaults.tolgauge = 1e-13) */
  double gauge_tol_min;   /* 1e-14 */
  double gauge_tol_max;   /* 1e-4 */
  int envs_dynamic;       /* 1: alg_environments is a DynamicTol (default) */
  double envs_tol;        /* fixed environment tolerance (Defaults.tol = 1e-10) */
  double envs_tol_min;    /* 1e-14 */
  double envs_tol_max;    /* 1e-4 */
  int envs_krylovdim;     /* DMRGCache krylovdim = 30 */
  imps_finalize_cb finalize; /* NULL: Defaults._finalize (identity) */
  void* finalize_user;
  int refine_eig;         /* mixed precision, once the FP64 refinement starts: 1 (default) the
                             eigensolves keep the reduced-precision operator for Jacobi-Davidson
                             corrections (MINRES) of FP64 residuals, until |H x - theta x| <= tol;
                             0: FP64 Lanczos */
  int gauge_eig_miniter;  /* RightCanonical eig_miniter = 10 (ortho.jl:22-31, 166-202): gauge sweeps
                             before the dominant-eigenvector refresh of C from the mixed transfer map */
} imps_vumps_params;

/* IterativeConvergenceInfo (algdefs.jl:180-194) plus timings. */
typedef struct imps_vumps_info {
  int niter;
  int converged;
  double itererr;     /* last Galerkin error */
  double energy;      /* last expectationvalue(ψ, H, envs) */
  double seconds;     /* wall time of the loop */
  double t_galerkin_reached; /* seconds until eps <= tol first held (-1 if never) */
  int hac_applies;
  int hc_applies;
  int prec_switch_iter; /* iteration at which FP64 refinement started (-1 if none) */
  /* wall seconds spent in each phase of the loop */
  double t_local;    /* localupdate_step!: AC/C eigensolves and regauge! */
  double t_gauge;    /* gauge_step! */
  double t_envs;     /* recalculate! */
  double t_galerkin; /* calc_galerkin */
  double t_energy;   /* expectationvalue */
  int gauge_sweeps;  /* LQ sweeps summed over iterations */
  int gmres_ops;     /* environment-solver operator applications */
  int qr_fallbacks;  /* Cholesky-QR breakdowns answered by Householder QR */
  int refine_fallbacks; /* refined eigensolves that fell back to FP64 Lanczos */
} imps_vumps_info;

/* Per-iteration callback: return nonzero to stop the loop. */
typedef int (*imps_iter_cb)(int iter, double eps, double energy, void* user);

/* Fills *p with the reference defaults. */
int imps_vumps_default_params(imps_vumps_params* p);

/* find_groundstate(ψ₀, H::SparseIMPO, alg::VUMPS, envs; which = :SR)
 * (src/algorithms/groundstates/vumps.jl:51-86) on the context's state and MPO.
 * Builds DMRGCache(ψ₀, H) first when no environments exist.  `losses` (may be NULL)
 * receives up to `nloss` Galerkin errors: the initial one, then one per iteration. */
int imps_vumps_run(imps_ctx* ctx, const imps_vumps_params* p, imps_vumps_info* info,
                   double* losses, int nloss, imps_iter_cb cb, void* user);

/* find_groundstate(ψ₀, H::SparseIMPO, alg::IDMRG, envs) (src/algorithms/groundstates/idmrg.jl:44-86):
 * single-site IDMRG sweeps (_localupdate_sweep_idmrg!, idmrg.jl:12-34: forward H_AC eigensolves
 * with QRpos splits and MPO-channel left pushes, backward eigensolves with LQpos splits and right
 * pushes), stopping when the bond-0 centre matrix moves by at most p->tol; the state is then
 * rebuilt from AR (CanonicalIMPS with the dynamic gauge tolerance) and the environments
 * recalculated.  Uses p's tol, maxiter, verbosity, tol_min/max, eigs/gauge tolerance factors,
 * krylovdim and maxiters; FP64 throughout (prec, refine_below and eig_floor are ignored).
 * `losses` receives the initial Galerkin error, then the centre drift of every sweep; `cb`
 * receives (iter, drift, lowest H_AC eigenvalue of the sweep).  info->energy is
 * expectationvalue(ψ, H, envs) of the rebuilt state. */
int imps_idmrg_run(imps_ctx* ctx, const imps_vumps_params* p, imps_vumps_info* info,
                   double* losses, int nloss, imps_iter_cb cb, void* user);

/* timestep(ψ, H, dt, alg::TDVP, envs) (src/algorithms/timeevo/tdvp.jl:33-52) for real dt
 * (imaginary time, dt = -τ): every site's AC and C become exp(dt H_AC) AC and exp(dt H_C) C
 * (Lanczos exponential with tolerance tol, krylovdim, maxiter: KrylovKit.exponentiate as used by
 * tdvp.jl:14-19), AL follows from regauge! (ortho.jl:214-219), the state is rebuilt as
 * CanonicalIMPS(ALs, C[end]) with (gauge_tol, gauge_maxiter) and the environments are
 * recalculated.  Builds DMRGCache first when no environments exist.  A non-converged exponential
 * is reported on stderr, as the reference warns.  Works on real and complex states. */
int imps_tdvp_timestep(imps_ctx* ctx, double dt, double tol, int krylovdim, int maxiter, double gauge_tol,
                       int gauge_maxiter);

/* The same step for a complex dt = dt_re + i dt_im (tdvp.jl:33-52 with dt = -im t: real-time
 * evolution exp(-i t H)); IMPS_E_STATE when dt_im != 0 and the state is real (time_evolve,
 * tdvp.jl:91-112, promotes a real state to complex first: imps_set_state_complex). */
int imps_tdvp_timestep_complex(imps_ctx* ctx, double dt_re, double dt_im, double tol, int krylovdim, int maxiter,
                               double gauge_tol, int gauge_maxiter);

/* ---- complex states (CanonicalIMPS{ComplexF64}, canonicalmps.jl:17-40).  Tensors are
 * ComplexF64 arrays as Julia stores them: column-major (D, d, D) or (D, D), real and imaginary
 * parts interleaved (2 doubles per entry).  The device holds their exact real embedding (each
 * entry z as the 2 x 2 block [[Re z, -Im z], [Im z, Re z]], bond dimension 2 D), on which the
 * real kernels run unchanged: products, adjoints (as transposes) and the positive-diagonal
 * QR/LQ/Cholesky factors commute with the embedding.  On a context holding a complex state:
 * find_groundstate (imps_vumps_run), the environments (imps_envs_recalculate), imps_energy,
 * imps_galerkin, imps_gaugefix_right, both TDVP steps and imps_local_expectation /
 * imps_correlator with symmetric real operators work as for real states;
 * imps_local_expectation_complex and imps_correlator_complex take any ComplexF64 operators; imps_hac_apply,
 * imps_hc_apply, imps_transfer_apply, imps_get_env and the vectors of imps_eigsolve_ac/c take and
 * return ComplexF64 arrays of the complex bond dimension (eigenvectors of unit norm);
 * imps_state_dims reports the complex bond dimension; the entry points that pass real
 * state-shaped host arrays or need real arithmetic (imps_get_state, set_env, regauge, the device-
 * pointer applies, IDMRG, imps_dot_state (imps_dot_state_complex takes complex states),
 * dominant_env, dense_expectation, the gradients, checkpointing) return IMPS_E_UNSUPPORTED. */

/* CanonicalIMPS{ComplexF64} from canonical tensors (the complex twin of imps_set_state; AC may
 * be NULL: AL C). */
int imps_set_state_complex(imps_ctx* ctx, int nsites, int D, const double* const* AL, const double* const* AR,
                           const double* const* C, const double* const* AC);

/* CanonicalIMPS(As) for complex raw tensors (canonicalmps.jl:151-170; randomimps(ComplexF64, ...),
 * constructors.jl:13-22): gaugefix!(:LR) from C0 = I as imps_set_state_raw. */
int imps_set_state_raw_complex(imps_ctx* ctx, int nsites, int D, const double* const* A, double tol, int maxiter);

/* CanonicalIMPS{ComplexF64} with bond dimensions that differ from bond to bond (the complex twin
 * of imps_set_state_nonuniform: chi[l] complex bond dimensions, D = max chi, tensors at (D, d, D)
 * with zeros outside their active blocks).  The embedding keeps the profile (complex index a is
 * rows 2a and 2a + 1, so the active block of chi complex indices is the first 2 chi rows), and
 * imps_bond_dims reports chi. */
int imps_set_state_complex_nonuniform(imps_ctx* ctx, int nsites, int D, const int* chi, const double* const* AL,
                                      const double* const* AR, const double* const* C, const double* const* AC);

/* CanonicalIMPS(As) for complex raw tensors with bond dimensions chi (canonicalmps.jl:151-170):
 * gaugefix!(:LR) from C0 = I on the active block, as imps_set_state_raw_nonuniform. */
int imps_set_state_raw_complex_nonuniform(imps_ctx* ctx, int nsites, int D, const int* chi, const double* const* A,
                                          double tol, int maxiter);

/* The complex state's AL, AR, C, AC (any pointer array or entry may be NULL); IMPS_E_STATE for a
 * real state. */
int imps_get_state_complex(imps_ctx* ctx, double* const* AL, double* const* AR, double* const* C,
                           double* const* AC);

/* *flag = 1 when the context holds a complex state, 0 for a real one. */
int imps_state_is_complex(imps_ctx* ctx, int* flag);

/* correlator(ψ, O1, O2, i, js) (src/observables/correlators.jl:13-42) on the context's state:
 * out[k] = ⟨ψ|O1[i] O2[js[k]]|ψ⟩ / dot(ψ, ψ) for i < js[0] < js[1] < ..., O1 and O2 d x d
 * column-major on the host.  Every value equals the reference's single-j form; for a range the
 * sites already evaluated are pushed through before the next j. */
int imps_correlator(imps_ctx* ctx, const double* O1, const double* O2, int i, int nj, const int* js,
                    double* out);
/* imps_correlator for ComplexF64 operators O1, O2 (d x d, column-major, re and im interleaved), on
 * real and complex states: out[2k], out[2k+1] = Re, Im <O1_i O2_js[k]> (correlators.jl:13-42 with
 * complex operators, e.g. Sy Sy). */
int imps_correlator_complex(imps_ctx* ctx, const double* O1, const double* O2, int i, int nj, const int* js,
                            double* out);
/* dot(ψ₁, ψ₂) (src/states/canonicalmps.jl:262-269) with ψ₁ the context's state and ψ₂ given by
 * its AL tensors (host, same unit cell and bond dimension): the dominant eigenvalue of
 * TransferMatrix(ψ₂.AL, ψ₁.AL) from the identity (Arnoldi, tolerance 1e-10).  The eigenvalue of
 * a real transfer matrix may be complex, as the reference's is: out[0] receives its real part and
 * out[1] its imaginary part (0 for a real dominant eigenvalue).  IMPS_E_SOLVER when the
 * eigensolver does not converge. */
int imps_dot_state(imps_ctx* ctx, int nsites, const double* const* AL2, double* out);
/* dot(ψ₁, ψ₂) for complex states (canonicalmps.jl:262-269): ψ₁ the context's complex state, AL2
 * the ComplexF64 (D, d, D) AL tensors of ψ₂ (re and im interleaved, the complex bond dimension);
 * out[0], out[1] = Re, Im ⟨ψ₁|ψ₂⟩.  IMPS_E_STATE for a real context state. */
int imps_dot_state_complex(imps_ctx* ctx, int nsites, const double* const* AL2, double* out);
/* dominant_env(ψ; side) and dominant_env(W::DenseIMPO, ψ; side) (src/algorithms/groundstates/
 * envs.jl:71-99, transfer maps 34-68) on the context's state: the dominant (:LM) eigenpair of the
 * unit-cell transfer map, from the left over AL (side 0) or from the right over AR (side 1).
 * W == NULL: identity channel (nw = 0, w = 1).  Otherwise nw DenseIMPO tensors (w, d, w, d)
 * column-major on the host, W[l % nw] acting at site l.  Arnoldi from ones (eager; tol,
 * krylovdim, maxiter as in the reference's keyword arguments).  lambda receives the eigenvalue,
 * env the eigenvector (D, w, D) column-major, Frobenius norm 1, sign fixed by a non-negative
 * entry sum.  IMPS_E_UNSUPPORTED when the dominant eigenvalue is complex (real arithmetic). */
int imps_dominant_env(imps_ctx* ctx, int side, int nw, int w, int d, const double* const* W, double tol,
                      int krylovdim, int maxiter, double* lambda, double* env);
/* expectationvalue(ψ, W::DenseIMPO, DMRGCache(ψ, W)) (src/observables/expval.jl:69-77 with the
 * environments of src/algorithms/groundstates/envs.jl:284-304 and normalize_envs! 559-581) on the
 * context's state: dominant left environment at site 0 and right environment at site N-1
 * (imps_dominant_env), MPO-channel pushes in between, unit-norm right environments and left
 * environments scaled so that <C[i], H_C(i) C[i]> = 1; value = sum over the cell of the
 * single-site sandwiches / dot(ψ, ψ).  W as for imps_dominant_env (nw >= 1).  GL_out / GR_out
 * (optional, NULL or N host pointers, each NULL or D*w*D doubles) receive the environments,
 * (bra, w, ket) for GL and (ket, w, bra) for GR, column-major. */
int imps_dense_expectation(imps_ctx* ctx, int nw, int w, int d, const double* const* W, double tol, int krylovdim,
                           int maxiter, double* value, double* const* GL_out, double* const* GR_out);

/* State and model files (no counterpart in the reference, which has no file format; for
 * checkpoint and resume of long runs and for exchanging initial states).  Little-endian: an
 * 8-byte magic ("IMPSSTA1" state, "IMPSMPO1" MPO), uint32 version 1, uint32 0x01020304, int32
 * nsites, D (or w), d, int32 element type 1 (float64), the column-major payload (per site AL, AR
 * (D d D), C (D D), AC (D d D); or W (w d w d)), then a 64-bit FNV-1a hash of the payload.
 * Writes go to "<path>.tmp", renamed onto path when complete.  Reads check the header, the
 * requested dimensions (IMPS_E_ARG), the length and the hash (IMPS_E_IO) before writing the
 * output.  Errors of these context-free calls are read with imps_last_error(NULL). */
int imps_write_state_file(const char* path, int nsites, int D, int d, const double* const* AL,
                          const double* const* AR, const double* const* C, const double* const* AC);
int imps_state_file_info(const char* path, int* nsites, int* D, int* d);
int imps_read_state_file(const char* path, int nsites, int D, int d, double* const* AL, double* const* AR,
                         double* const* C, double* const* AC);
int imps_write_mpo_file(const char* path, int nsites, int w, int d, const double* const* W);
int imps_mpo_file_info(const char* path, int* nsites, int* w, int* d);
int imps_read_mpo_file(const char* path, int nsites, int w, int d, double* const* W);
/* The context's state to a state file / a state file into the context (its model must have the
 * file's unit cell and physical dimension, as for imps_set_state). */
int imps_checkpoint_state(imps_ctx* ctx, const char* path);
int imps_resume_state(imps_ctx* ctx, const char* path);

/* ---- batched sweeps ------------------------------------------------------------------ */

/* One instance of a parameter sweep: its MPO, raw ψ₀ tensors (canonicalised on the device
 * like CanonicalIMPS(As)) and the results. */
typedef struct imps_batch_instance {
  int nsites, w, d, D;
  const double* const* W; /* nsites tensors (w, d, w, d) */
  const double* const* A; /* nsites raw tensors (D, d, D) */
  /* outputs */
  int status; /* IMPS_OK or the failing call's code */
  int niter;
  int converged;
  double itererr;
  double energy; /* unit-cell energy */
  double seconds;
  int hac_applies; /* H_AC applications of the instance (canonicalisation excluded) */
} imps_batch_instance;

/* Runs find_groundstate (vumps.jl:51-86) on every instance with the parameters `p`,
 * `nworkers` host threads each owning a context (its own HIP stream) on `device`, so the
 * instances' kernels overlap on the GPU.  Instances are taken largest D^3 first. */
int imps_vumps_batch(int device, int ninst, imps_batch_instance* inst, const imps_vumps_params* p,
                     int nworkers, double gauge_tol);

/* The same sweep as one device batch: every instance runs find_groundstate (vumps.jl:51-86) from
 * CanonicalIMPS(As) (canonicalmps.jl:151-170), and all instances advance in lockstep, one host
 * thread and one HIP stream for the whole batch: each Krylov step (Lanczos eigsolve, GMRES
 * linsolve, Arnoldi gauge fixed point) of all active instances is one grouped GEMM launch per
 * contraction stage, one grouped Gram-Schmidt sequence and one readback; every QR/LQ sweep of all
 * instances is one batched Householder launch.  No counterpart in the reference (its sweeps are
 * loops over find_groundstate).  Scope: one-site unit cells with the same w and d (bond
 * dimensions may differ: multiples of 16, d D <= 2048), FP64 or FP32 bulk arithmetic (the FP32
 * phase switches to FP64 eigensolves below refine_below or at a stall), no finalize callback;
 * IMPS_E_UNSUPPORTED otherwise.  The message of a failure is read with imps_last_error(NULL). */
typedef struct imps_batch_stats {
  long rounds;   /* lockstep rounds (each ends with one host synchronisation), all groups */
  long launches; /* kernel launches for the whole batch */
  double seconds;
  int groups;    /* lockstep groups run concurrently (one host thread and stream each) */
  double t_phase[6]; /* wall seconds of the slowest group in: canonicalisation, local update,
                        gauge_step!, recalculate!, calc_galerkin, expectationvalue */
} imps_batch_stats;
int imps_vumps_batch_device(int device, int ninst, imps_batch_instance* inst, const imps_vumps_params* p,
                            double gauge_tol, imps_batch_stats* stats);

/* ---- measurement helpers ----------------------------------------------------------- */

/* Runs `reps` H_AC applications on device-resident random data of the current
 * site-0 shapes and returns the average duration of one application and of each of
 * its kernels (HIP events on the context's stream), and the flops of each GEMM launch as
 * executed (2 (w-1) d D^3 when the environments' structural levels are identities, else
 * 2 w d D^3).  In the split-BF16 mode the GL x bracket includes the split of x into planes,
 * and the split of the environments into planes (once per eigensolve) is timed apart. */
typedef struct imps_hac_timing {
  double ms_apply;  /* whole application */
  double ms_gemm1;  /* GL x  */
  double ms_wmix;   /* W mix */
  double ms_gemm2;  /* (.) GR */
  double flop_gemm1;
  double flop_gemm2;
  double ms_prepare; /* per-eigensolve operand split of the split-BF16 mode (0 otherwise) */
} imps_hac_timing;
int imps_bench_hac(imps_ctx* ctx, int site, int reps, int prec, imps_hac_timing* t);

/* Average times of the Krylov solvers' Gram-Schmidt kernels on synthetic vectors of length n
 * against a basis of nv (1..32) vectors, `reps` launches each with HIP events on the context's
 * stream, rotating over three bases so that the operands come from HBM:
 * X_v . w (the dot sweep), w -= X c followed by X_v . w (the second-pass sweep) and w -= X c
 * followed by w . w (the norm sweep).  Algorithmic bytes: (nv + 1) n 8 for the dot sweep,
 * (nv + 2) n 8 for the others (the basis, and w read and written). */
typedef struct imps_krylov_timing {
  double ms_multidot;
  double ms_axpy_dots;
  double ms_axpy_norm;
  double bytes_multidot;
  double bytes_axpy;
} imps_krylov_timing;
int imps_bench_krylov(imps_ctx* ctx, long n, int nv, int reps, imps_krylov_timing* t);

/* General strided GEMM of the library on device pointers (column-major):
 * C = alpha*op(A)*op(B) + beta*C, op = 'N' or 'T'.  Exposed for tests. */
int imps_gemm_dev(imps_ctx* ctx, char transa, char transb, int m, int n, int k, double alpha,
                  const double* d_a, int lda, const double* d_b, int ldb, double beta, double* d_c,
                  int ldc, int prec);
/* GEMM tile autotuning, process-wide, on by default (no counterpart in the reference, whose
 * contractions go to BLAS): the first FP64/FP32 GEMM of each shape times the library's
 * bit-identical tile variants once and keeps the fastest; 0 switches back to the fixed
 * shape heuristic; 2 + v (v = 0..5) runs tile variant v wherever it is valid and autotunes
 * elsewhere (for tests and measurement).  Results do not depend on the setting.  Returns
 * IMPS_E_ARG for other values. */
int imps_set_gemm_autotune(int on);
/* Graph capture of fixed kernel sequences, process-wide, on by default (no counterpart in the
 * reference): the gauge sweeps (uniform_rightorth!, ortho.jl:166-202) run each LQ sweep as a
 * captured hipGraph replayed from the second sweep on; 0 runs them eagerly (same kernels, same
 * order, bit-identical results; for tests and measurement).  IMPS_E_ARG for other values. */
int imps_set_graph_capture(int on);
/* Device-to-device copy of `bytes` bytes with the library's copy kernel on the
 * context's stream (also the known-size reference for HBM traffic counters). */
int imps_dev_copy(imps_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* Device memory helpers for callers without their own allocator. */
int imps_dev_alloc(imps_ctx* ctx, size_t bytes, void** d_ptr);
int imps_dev_free(imps_ctx* ctx, void* d_ptr);
int imps_memcpy_h2d(imps_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int imps_memcpy_d2h(imps_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
int imps_synchronize(imps_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* IMPS_H */
