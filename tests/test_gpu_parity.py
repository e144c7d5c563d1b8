"""GPU parity of libimps (called through the C ABI) against the FP64 oracle.

Tolerances: the bulk contractions are FP64 on MFMA, so single operator applications
agree with the oracle to relative 1e-12; Krylov/GMRES results agree to their solver
tolerance; gauge-invariant VUMPS results (energy, Galerkin error) to the stated bounds.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def imps():
    import infinitempsalgorithms_amd as m
    return m


@pytest.fixture(scope="module")
def ctx(imps):
    c = imps.Context(0)
    yield c
    c.close()


class DevArray:
    def __init__(self, ctx, a):
        self.ctx, self.a = ctx, np.asfortranarray(a, dtype=np.float64)
        self.p = ctypes.c_void_p()
        ctx.call("imps_dev_alloc", self.a.nbytes, ctypes.byref(self.p))
        ctx.call("imps_memcpy_h2d", self.p, self.a.ctypes.data_as(ctypes.c_void_p), self.a.nbytes)

    def get(self):
        out = np.zeros_like(self.a, order="F")
        self.ctx.call("imps_memcpy_d2h", out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes)
        return out

    def free(self):
        self.ctx.call("imps_dev_free", self.p)


@pytest.mark.parametrize("ta,tb,m,n,k", [
    ("N", "N", 64, 64, 64), ("N", "N", 130, 77, 45), ("T", "N", 33, 129, 70), ("N", "T", 257, 100, 31),
    ("T", "T", 5, 3, 2), ("N", "N", 1, 1, 1), ("N", "N", 640, 512, 96), ("T", "N", 1024, 1024, 4096),
    # 64 x 128 tiles with paired loads, all four operand orientations; K % 16 != 0 falls back
    ("N", "N", 2048, 1536, 64), ("T", "N", 1920, 1664, 80), ("N", "T", 2048, 1536, 48),
    ("T", "T", 2048, 1536, 32), ("N", "N", 2048, 1536, 33),
    # long K on a full-tile grid
    ("N", "N", 2048, 1536, 2304),
])
def test_gemm_matches_numpy(ctx, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C0 = rng.standard_normal((m, n))
    ref = 0.7 * (A if ta == "N" else A.T) @ (B if tb == "N" else B.T) - 0.3 * C0
    dA, dB, dC = DevArray(ctx, A), DevArray(ctx, B), DevArray(ctx, C0)
    ctx.call("imps_gemm_dev", ta.encode(), tb.encode(), m, n, k, 0.7, dA.p, A.shape[0], dB.p, B.shape[0],
             -0.3, dC.p, m, 0)
    got = dC.get()
    for d in (dA, dB, dC):
        d.free()
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("ta,tb,m,n,k", [("N", "N", 2048, 1536, 64), ("T", "N", 1920, 1664, 80), ("N", "T", 640, 512, 96),
                                         ("N", "N", 130, 77, 45), ("N", "N", 5120, 4096, 64)])
def test_gemm_autotune_is_bit_identical(imps, ctx, prec, ta, tb, m, n, k):
    # the autotuner picks among tile variants that sum over K in the same order: the output
    # with autotuning equals the fixed heuristic's bit for bit
    lib = imps._lib.load()
    rng = np.random.default_rng(m + 2 * n + 3 * k + prec)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C0 = rng.standard_normal((m, n))
    outs = []
    # tune, cached choice, heuristic, then each tile variant forced where valid
    for on in (1, 1, 0) + tuple(range(2, 8)):
        assert lib.imps_set_gemm_autotune(on) == 0
        dA, dB, dC = DevArray(ctx, A), DevArray(ctx, B), DevArray(ctx, C0)
        ctx.call("imps_gemm_dev", ta.encode(), tb.encode(), m, n, k, 0.7, dA.p, A.shape[0], dB.p, B.shape[0],
                 -0.3, dC.p, m, prec)
        outs.append(dC.get())
        for d in (dA, dB, dC):
            d.free()
    lib.imps_set_gemm_autotune(1)
    for o in outs:
        assert np.array_equal(o, outs[2])
    ref = 0.7 * (A if ta == "N" else A.T) @ (B if tb == "N" else B.T) - 0.3 * C0
    assert rel(outs[0], ref) < (1e-13 if prec == 0 else 2e-6)


@pytest.mark.parametrize("ta,tb,m,n,k", [("N", "N", 130, 77, 45), ("T", "N", 1024, 1024, 4096), ("N", "T", 640, 512, 96),
                                         ("N", "N", 2048, 1536, 256)])
def test_gemm_fp32_mfma(ctx, ta, tb, m, n, k):
    # FP32 products/accumulation on FP64 data: error ~ 2^-24 sqrt(k) relative
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    ref = (A if ta == "N" else A.T) @ (B if tb == "N" else B.T)
    dA, dB, dC = DevArray(ctx, A), DevArray(ctx, B), DevArray(ctx, np.zeros((m, n)))
    ctx.call("imps_gemm_dev", ta.encode(), tb.encode(), m, n, k, 1.0, dA.p, A.shape[0], dB.p, B.shape[0],
             0.0, dC.p, m, 1)
    got = dC.get()
    for d in (dA, dB, dC):
        d.free()
    err = rel(got, ref)
    assert 1e-9 < err < 2e-6  # genuinely FP32, and as accurate as FP32 allows


@pytest.mark.parametrize("ta,tb,m,n,k", [("N", "N", 130, 77, 45), ("T", "N", 1024, 1024, 4096), ("N", "T", 640, 512, 96)])
def test_gemm_bf16x3_mfma(ctx, ta, tb, m, n, k):
    # a*b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi (bf16 parts, FP32 accumulation): ~2^-16 relative
    rng = np.random.default_rng(m + n + k + 1)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    ref = (A if ta == "N" else A.T) @ (B if tb == "N" else B.T)
    dA, dB, dC = DevArray(ctx, A), DevArray(ctx, B), DevArray(ctx, np.zeros((m, n)))
    ctx.call("imps_gemm_dev", ta.encode(), tb.encode(), m, n, k, 1.0, dA.p, A.shape[0], dB.p, B.shape[0],
             0.0, dC.p, m, 2)
    got = dC.get()
    for d in (dA, dB, dC):
        d.free()
    err = rel(got, ref)
    assert 1e-8 < err < 5e-5


def _setup_envs(imps, ctx, oracle, Wfn, D, N, seed, sym=False, eye_levels=False):
    # eye_levels: GL level 0 and GR level w-1 set to the identity, as DMRGCache pins them (the
    # effective operators then take their copy path for those levels)
    rng = np.random.default_rng(seed)
    W = Wfn()
    H = imps.SparseIMPO([W] * N)
    ctx.set_mpo(H)
    psi = oracle.randomimps(D, [H.d] * N, seed=seed)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    ctx.set_state(st)
    w = H.w
    GLs, GRs = [], []
    for l in range(N):
        GL = rng.standard_normal((D, w, D))
        GR = rng.standard_normal((D, w, D))
        if eye_levels:
            GL[:, 0, :] = np.eye(D)
            GR[:, w - 1, :] = np.eye(D)
        ctx.set_env(l, 0, GL)
        ctx.set_env(l, 1, GR)
        GLs.append(GL)
        GRs.append(GR)
    return H, psi, GLs, GRs


@pytest.mark.parametrize("eye", [False, True])
@pytest.mark.parametrize("model,D", [("tfim", 16), ("xxz", 24), ("hubbard", 20), ("hubbard", 64), ("tfim", 7)])
def test_hac_hc_match_oracle(imps, ctx, oracle, model, D, eye):
    Wfn = {"tfim": oracle.tfim_mpo, "xxz": oracle.xxz_mpo, "hubbard": oracle.hubbard_mpo}[model]
    H, psi, GLs, GRs = _setup_envs(imps, ctx, oracle, Wfn, D, 2, seed=D, eye_levels=eye)
    envs = imps.GPUCache.__new__(imps.GPUCache)
    envs.ctx, envs.ket, envs.operator = ctx, None, H
    rng = np.random.default_rng(1)
    x = rng.standard_normal((D, H.d, D))
    for site in range(2):
        y = imps.AC_hamiltonian(site, None, H, None, envs)(x)
        assert rel(y, oracle.hac_apply(GLs[site], H.Ws[site], GRs[site], x)) < 1e-12
        c = rng.standard_normal((D, D))
        yc = imps.C_hamiltonian(site, None, H, None, envs)(c)
        assert rel(yc, oracle.hc_apply(GLs[(site + 1) % 2], GRs[site], c)) < 1e-12


@pytest.mark.parametrize("eye", [False, True])
def test_hac_full_size_hubbard(imps, ctx, oracle, eye):
    # the D = 1024 Hubbard shape of the benchmark (w = 6, d = 4), against the dense einsum
    D = 1024
    H, psi, GLs, GRs = _setup_envs(imps, ctx, oracle, oracle.hubbard_mpo, D, 1, seed=2, eye_levels=eye)
    envs = imps.GPUCache.__new__(imps.GPUCache)
    envs.ctx, envs.ket, envs.operator = ctx, None, H
    x = np.random.default_rng(3).standard_normal((D, 4, D))
    y = imps.AC_hamiltonian(0, None, H, None, envs)(x)
    assert rel(y, oracle.hac_apply_dense(GLs[0], H.Ws[0], GRs[0], x)) < 1e-12
    # linearity: H(2x + y) = 2 H(x) + H(y)
    z = np.random.default_rng(4).standard_normal((D, 4, D))
    lhs = imps.AC_hamiltonian(0, None, H, None, envs)(2 * x + z)
    assert rel(lhs, 2 * y + imps.AC_hamiltonian(0, None, H, None, envs)(z)) < 1e-12


def test_transfer_and_regularize(imps, ctx, oracle):
    D = 12
    H, psi, _, _ = _setup_envs(imps, ctx, oracle, oracle.tfim_mpo, D, 2, seed=5)
    rng = np.random.default_rng(6)
    v = rng.standard_normal((D, D))
    tl = imps.TransferMatrix(ctx, "left")(v)
    assert rel(tl, oracle.transfer_left(psi.AL, psi.AL, v.reshape(-1, order="F")).reshape(D, D, order="F")) < 1e-12
    tr = imps.TransferMatrix(ctx, "right")(v)
    assert rel(tr, oracle.transfer_right(psi.AR, psi.AR, v.reshape(-1, order="F")).reshape(D, D, order="F")) < 1e-12
    # identity is the fixed point of the AL transfer (test/algorithms/envs.jl:25-28)
    assert rel(imps.TransferMatrix(ctx, "left")(np.eye(D)), np.eye(D)) < 1e-12
    M = np.asfortranarray(rng.standard_normal((D, D)))
    l = rng.standard_normal((D, D))
    ref = oracle.regularize(M.copy(), l, np.eye(D))
    imps.regularize_(M, l, np.eye(D), ctx=ctx)
    assert rel(M, ref) < 1e-13


@pytest.mark.parametrize("model,D,N", [("tfim", 8, 1), ("tfim", 12, 2), ("xxz", 10, 2), ("hubbard", 12, 2)])
def test_envs_energy_galerkin_match_oracle(imps, ctx, oracle, model, D, N):
    Wfn = {"tfim": oracle.tfim_mpo, "xxz": oracle.xxz_mpo, "hubbard": oracle.hubbard_mpo}[model]
    W = Wfn()
    H = imps.SparseIMPO([W] * N)
    psi = oracle.randomimps(D, [H.d] * N, seed=11)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, H, tol=1e-12, ctx=ctx)
    ref = oracle.dmrg_cache(psi, [W] * N, tol=1e-12)
    for l in range(N):
        assert rel(imps.leftenv(envs, l), ref.lefts[l]) < 1e-8
        assert rel(imps.rightenv(envs, l), ref.rights[l]) < 1e-8
    e = imps.expectationvalue(st, H, envs)
    assert abs(e - oracle.expectation_energy(psi, [W] * N, ref)) < 1e-10
    g = imps.calc_galerkin(st, H, envs)
    assert abs(g - oracle.calc_galerkin(psi, [W] * N, ref)) < 1e-9 * max(1.0, g)
    # warm-started recalculation keeps the energy (test/algorithms/api.jl:207-210)
    imps.recalculate_(envs, st, H, tol=1e-12)
    assert abs(imps.expectationvalue(st, H, envs) - e) < 1e-10


def test_local_update_and_gauge_match_oracle(imps, ctx, oracle):
    D, N = 10, 2
    W = oracle.tfim_mpo()
    H = imps.SparseIMPO([W] * N)
    psi = oracle.randomimps(D, [2] * N, seed=21)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, H, tol=1e-12, ctx=ctx)
    ref = oracle.dmrg_cache(psi, [W] * N, tol=1e-12)
    tol = 1e-10
    ALs_ref = []
    for site in range(N):
        lam, AC, info = imps.fixedpoint(imps.AC_hamiltonian(site, st, H, st, envs), psi.AC[site], "SR",
                                        imps.Lanczos(tol=tol))
        lr, ACr, _ = oracle.lanczos_eigsolve(lambda x: oracle.hac_apply(ref.leftenv(site), W, ref.rightenv(site), x),
                                             psi.AC[site], tol)
        assert info.converged and abs(lam - lr) < 1e-9
        assert rel(AC, ACr) < 1e-6
        lam, C, info = imps.fixedpoint(imps.C_hamiltonian(site, st, H, st, envs), psi.C[site], "SR",
                                       imps.Lanczos(tol=tol))
        lcr, Cr, _ = oracle.lanczos_eigsolve(lambda x: oracle.hc_apply(ref.leftenv(site + 1), ref.rightenv(site), x),
                                             psi.C[site], tol)
        assert abs(lam - lcr) < 1e-9 and rel(C, Cr) < 1e-6
        AL = imps.regauge_(envs, site)
        ALr = oracle.regauge(ACr, Cr)
        ALs_ref.append(ALr)
        assert rel(AL, ALr) < 1e-5
    sweeps = ctypes.c_int()
    ctx.call("imps_gaugefix_right", 1e-12, 300, ctypes.byref(sweeps))
    psi_g = ctx.get_state()
    ref_psi, ref_sweeps = oracle.gauge_step(psi, ALs_ref, 1e-12, 300)
    el, er, em = oracle.mixedcanonical_errors(oracle.CanonicalIMPS(psi_g.AL, psi_g.AR, psi_g.C, psi_g.AC))
    assert max(el, er) < 1e-10
    for l in range(N):
        # singular values of C are gauge invariant
        assert rel(np.linalg.svd(psi_g.C[l], compute_uv=False), np.linalg.svd(ref_psi.C[l], compute_uv=False)) < 1e-6


@pytest.mark.parametrize("kdim", [12, 30, 40])
def test_lanczos_basis_sizes_match_oracle(imps, ctx, oracle, kdim):
    # krylovdim <= 31 runs the fused (three-sweep) Gram-Schmidt step, 40 the four-sweep one;
    # small bases restart several times
    D, N = 16, 2
    W = oracle.xxz_mpo(1.0, 0.8)
    H = imps.SparseIMPO([W] * N)
    psi = oracle.randomimps(D, [2] * N, seed=31)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, H, tol=1e-12, ctx=ctx)
    ref = oracle.dmrg_cache(psi, [W] * N, tol=1e-12)
    n = 2 * D * D
    Hm = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        Hm[:, j] = oracle.hac_apply(ref.leftenv(0), W, ref.rightenv(0), e.reshape((D, 2, D), order="F")).reshape(-1, order="F")
    ev = np.linalg.eigvalsh(0.5 * (Hm + Hm.T))
    lam, AC, info = imps.fixedpoint(imps.AC_hamiltonian(0, st, H, st, envs), psi.AC[0], "SR",
                                    imps.Lanczos(tol=1e-11, krylovdim=kdim, maxiter=500))
    assert info.converged and abs(lam - ev[0]) < 1e-9 * max(1.0, abs(ev[0]))
    r = (Hm @ AC.reshape(-1, order="F")) - lam * AC.reshape(-1, order="F")
    assert np.linalg.norm(r) < 1e-9 and abs(np.linalg.norm(AC) - 1.0) < 1e-12


def test_canonicalize_raw_tensors(imps, ctx, oracle):
    H = imps.models.fermi_hubbard(nsites=2)
    As = imps.random_tensors(8, H.phydims(), seed=2)
    psi, _ = imps.canonicalimps(As, H, ctx=ctx)
    ref = oracle.canonical_from_tensors(As)
    el, er, em = oracle.mixedcanonical_errors(oracle.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC))
    assert max(el, er) < 1e-12 and em < 1e-10
    for l in range(2):
        assert rel(psi.C[l], ref.C[l]) < 1e-8
        assert rel(psi.AL[l], ref.AL[l]) < 1e-8


def test_canonicalize_rank_deficient_tensor(imps, ctx, oracle):
    # A[:, :, D-1] = 0 gives the left QRs an exactly zero column: the shifted first Cholesky
    # QR pass keeps it zero, the second pass meets a zero pivot, and the Householder QRpos
    # fallback has to produce the factorisation
    H = imps.models.tfim_hamiltonian(nsites=1)
    As = imps.random_tensors(8, H.phydims(), seed=5)
    As[0][:, :, 7] = 0.0
    psi, _ = imps.canonicalimps(As, H, ctx=ctx)
    ref = oracle.canonical_from_tensors(As)
    el, er, em = oracle.mixedcanonical_errors(oracle.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC))
    assert max(el, er) < 1e-12 and em < 1e-10
    sv = np.linalg.svd(psi.C[0], compute_uv=False)
    assert sv[-1] < 1e-12 * sv[0]
    assert np.abs(sv - np.linalg.svd(ref.C[0], compute_uv=False)).max() < 1e-10 * sv[0]


@pytest.mark.parametrize("N,D,seed", [(1, 16, 0), (2, 12, 0)])
def test_vumps_tfim_matches_oracle(imps, ctx, oracle, N, D, seed):
    # configuration "TFIM J=h=1, N=1, D=16" and the reference's D=12 2-site test
    W = oracle.tfim_mpo()
    H = imps.SparseIMPO([W] * N)
    psi0 = oracle.randomimps(D, [2] * N, seed=seed)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)
    envs = imps.GPUCache(st, H, ctx)
    ctx.call("imps_envs_reset")
    psi, envs, info = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-10), envs=envs)
    _, _, rinfo = oracle.find_groundstate(psi0, [W] * N, oracle.VUMPS(tol=1e-10))
    assert info.converged and info.itererr <= 1e-10
    e = info.energy / N
    assert abs(e + 4 / math.pi) < 1e-6
    assert abs(info.energy - rinfo.energies[-1]) < 1e-9
    assert abs(info.niter - rinfo.niter) <= max(3, rinfo.niter // 10)
    for a, b in zip(info.losses[:5], rinfo.losses[:5]):
        assert abs(a - b) <= 1e-6 * max(b, 1e-12) + 1e-9


def test_vumps_xxz_two_site_trajectory_matches_oracle(imps, ctx, oracle):
    # the C2 comparison (XXZ, N=2, real S+- form) at D=64: per-iteration Galerkin errors
    # within 1e-6 relative and the energy within 1e-9 after 10 iterations
    W = oracle.xxz_mpo(1.0, 1.0)
    H = imps.SparseIMPO([W] * 2)
    psi0 = oracle.randomimps(64, [2, 2], seed=1)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)
    envs = imps.GPUCache(st, H, ctx)
    ctx.call("imps_envs_reset")
    _, _, info = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-14, maxiter=10), envs=envs)
    _, _, rinfo = oracle.find_groundstate(psi0, [W] * 2, oracle.VUMPS(tol=1e-14, maxiter=10))
    assert info.niter == rinfo.niter == 10
    for a, b in zip(info.losses, rinfo.losses):
        assert abs(a - b) <= 1e-6 * b
    assert abs(info.energy - rinfo.energies[-1]) < 1e-9


def test_hac_fp32_against_fp64(imps, ctx, oracle):
    D = 128
    H, psi, GLs, GRs = _setup_envs(imps, ctx, oracle, oracle.hubbard_mpo, D, 2, seed=9)
    envs = imps.GPUCache.__new__(imps.GPUCache)
    envs.ctx, envs.ket, envs.operator = ctx, None, H
    x = np.random.default_rng(5).standard_normal((D, 4, D))
    y64 = oracle.hac_apply(GLs[0], H.Ws[0], GRs[0], x)
    y32 = imps.AC_hamiltonian(0, None, H, None, envs, prec=1)(x)
    assert 1e-9 < rel(y32, y64) < 5e-6


def test_vumps_mixed_precision_refines_to_fp64(imps, ctx, oracle):
    # FP32 bulk contractions, then FP64 refinement: same converged energy as the FP64 run
    W = oracle.tfim_mpo()
    N, D = 2, 12
    H = imps.SparseIMPO([W] * N)
    psi0 = oracle.randomimps(D, [2] * N, seed=0)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)
    envs = imps.GPUCache(st, H, ctx)
    ctx.call("imps_envs_reset")
    _, _, i64 = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-10), envs=envs)
    envs = imps.GPUCache(st, H, ctx)
    ctx.call("imps_envs_reset")
    _, _, imx = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-10, prec=1, refine_below=1e-6), envs=envs)
    assert imx.converged and imx.itererr <= 1e-10
    assert imx.prec_switch_iter > 1
    assert abs(imx.energy - i64.energy) < 1e-9


def test_local_expectation_matches_oracle(imps, ctx, oracle):
    # expectationvalue(psi, (i,) => O) (expval.jl:33-37) on a random canonical state
    W = oracle.tfim_mpo()
    H = imps.SparseIMPO([W] * 2)
    psi = oracle.randomimps(10, [2, 2], seed=31)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, H, ctx=ctx)
    for site, O in ((0, oracle.sigma_z()), (1, oracle.sigma_x())):
        got = imps.expectationvalue(st, (site, O), envs)
        assert abs(got - oracle.local_expectation1(psi, site, O)) < 1e-12


def test_hellmann_feynman_tfim(imps, ctx, oracle):
    # H = -J sum sx sx - h sum sz: de/dh = -<sz> at the variational optimum; central FD of
    # converged energies with step 1e-4 (D = 16, one-site cell)
    def run(h):
        H = imps.models.tfim_hamiltonian(J=1.0, h=h, nsites=1)
        psi0 = oracle.randomimps(16, [2], seed=0)
        st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)
        envs = imps.GPUCache(st, H, ctx)
        ctx.call("imps_envs_reset")
        psi, envs, info = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-11, maxiter=400), envs=envs)
        return psi, envs, info
    h0, dh = 1.3, 1e-4
    psi, envs, info = run(h0)
    sz = imps.expectationvalue(psi, (0, oracle.sigma_z()), envs)
    ep = run(h0 + dh)[2].energy
    em = run(h0 - dh)[2].energy
    fd = (ep - em) / (2 * dh)
    assert abs(fd + sz) < 1e-6
    # gapped point: D = 16 is converged to the closed form e(h) = -(1/pi) int sqrt(1 + h^2 + 2h cos k)
    ks = np.linspace(0, np.pi, 200001)
    e_exact = -np.trapezoid(np.sqrt(1 + h0 ** 2 + 2 * h0 * np.cos(ks)), ks) / np.pi
    assert abs(info.energy - e_exact) < 1e-9


def test_batched_sweep_matches_single_runs(imps, ctx, oracle):
    # several TFIM instances with uneven D run concurrently (one stream per worker); each
    # equals its own single run and the closed form e(h) at these gapped fields
    cases = [(0.6, 8), (1.4, 12), (0.5, 16), (1.5, 10), (0.8, 12)]
    alg = imps.VUMPS(tol=1e-10, maxiter=300)
    problems = []
    for k, (h, D) in enumerate(cases):
        H = imps.models.tfim_hamiltonian(J=1.0, h=h, nsites=1)
        problems.append((H, imps.random_tensors(D, [2], seed=100 + k)))
    res = imps.find_groundstates_batched(problems, alg, nworkers=3)
    ks = np.linspace(0, np.pi, 200001)
    for (h, D), (H, As), r in zip(cases, problems, res):
        assert r["converged"] and r["itererr"] <= 1e-10
        psi0, c2 = imps.canonicalimps(As, H, ctx=ctx, tol=1e-12)
        envs = imps.GPUCache(psi0, H, c2)
        ctx.call("imps_envs_reset")
        _, _, info = imps.find_groundstate(psi0, H, alg, envs=envs)
        assert abs(r["energy"] - info.energy) < 1e-10
        e_exact = -np.trapezoid(np.sqrt(1 + h ** 2 + 2 * h * np.cos(ks)), ks) / np.pi
        assert abs(r["energy"] - e_exact) < 1e-7


def test_hac_bf16x3_and_mixed_vumps(imps, ctx, oracle):
    D = 64
    H, psi, GLs, GRs = _setup_envs(imps, ctx, oracle, oracle.hubbard_mpo, D, 2, seed=19)
    envs = imps.GPUCache.__new__(imps.GPUCache)
    envs.ctx, envs.ket, envs.operator = ctx, None, H
    x = np.random.default_rng(8).standard_normal((D, 4, D))
    y64 = oracle.hac_apply(GLs[1], H.Ws[1], GRs[1], x)
    yb = imps.AC_hamiltonian(1, None, H, None, envs, prec=2)(x)
    assert 1e-8 < rel(yb, y64) < 1e-4
    # split-BF16 bulk phase, then FP64 refinement: converged energy equals the FP64 run's
    W = oracle.tfim_mpo()
    Ht = imps.SparseIMPO([W] * 2)
    psi0 = oracle.randomimps(12, [2, 2], seed=0)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)
    e = imps.GPUCache(st, Ht, ctx)
    ctx.call("imps_envs_reset")
    _, _, i64 = imps.find_groundstate(st, Ht, imps.VUMPS(tol=1e-10), envs=e)
    e = imps.GPUCache(st, Ht, ctx)
    ctx.call("imps_envs_reset")
    _, _, ib = imps.find_groundstate(st, Ht, imps.VUMPS(tol=1e-10, prec=2, refine_below=1e-4), envs=e)
    assert ib.converged and ib.prec_switch_iter > 1
    assert abs(ib.energy - i64.energy) < 1e-9


@pytest.mark.parametrize("model,D", [("hubbard", 128), ("xxz", 256), ("tfim", 128)])
def test_hac_bf16x3_planes_match_oracle(imps, ctx, oracle, model, D):
    # identity structural levels and D % 128 == 0: the split-BF16 planes kernels (split, two
    # plane GEMMs, transposing mix).  Three bf16 products per pair with FP32 accumulation: the
    # relative error sits near 2^-17 (about 4.5e-6 in an emulation of the same arithmetic);
    # a dropped correction term would give ~2^-9, a wrong index O(1).
    Wfn = {"tfim": oracle.tfim_mpo, "xxz": oracle.xxz_mpo, "hubbard": oracle.hubbard_mpo}[model]
    H, psi, GLs, GRs = _setup_envs(imps, ctx, oracle, Wfn, D, 2, seed=D + 1, eye_levels=True)
    envs = imps.GPUCache.__new__(imps.GPUCache)
    envs.ctx, envs.ket, envs.operator = ctx, None, H
    rng = np.random.default_rng(21)
    for site in range(2):
        x = rng.standard_normal((D, H.d, D))
        ref = oracle.hac_apply(GLs[site], H.Ws[site], GRs[site], x)
        yb = imps.AC_hamiltonian(site, None, H, None, envs, prec=2)(x)
        assert 1e-8 < rel(yb, ref) < 2e-5
        # linearity up to the same rounding
        z = rng.standard_normal((D, H.d, D))
        lhs = imps.AC_hamiltonian(site, None, H, None, envs, prec=2)(2 * x + z)
        rhs = 2 * yb + imps.AC_hamiltonian(site, None, H, None, envs, prec=2)(z)
        assert rel(lhs, rhs) < 2e-5


@pytest.mark.parametrize("model,D", [("hubbard", 128), ("xxz", 256), ("tfim", 128)])
def test_hc_bf16x3_planes_match_oracle(imps, ctx, oracle, model, D):
    # H_C on planes: GEMM 1 over the non-identity GL levels, transposing level splits, split-K
    # GEMM 2 and the reduction that adds the identity level of GR (rel ~6e-6 in an emulation)
    Wfn = {"tfim": oracle.tfim_mpo, "xxz": oracle.xxz_mpo, "hubbard": oracle.hubbard_mpo}[model]
    H, psi, GLs, GRs = _setup_envs(imps, ctx, oracle, Wfn, D, 2, seed=D + 3, eye_levels=True)
    envs = imps.GPUCache.__new__(imps.GPUCache)
    envs.ctx, envs.ket, envs.operator = ctx, None, H
    rng = np.random.default_rng(23)
    for site in range(2):
        x = rng.standard_normal((D, D))
        ref = oracle.hc_apply(GLs[(site + 1) % 2], GRs[site], x)
        yb = imps.C_hamiltonian(site, None, H, None, envs, prec=2)(x)
        assert 1e-8 < rel(yb, ref) < 2e-5


def test_hac_bf16x3_planes_full_size_hubbard(imps, ctx, oracle):
    D = 1024
    H, psi, GLs, GRs = _setup_envs(imps, ctx, oracle, oracle.hubbard_mpo, D, 1, seed=2, eye_levels=True)
    envs = imps.GPUCache.__new__(imps.GPUCache)
    envs.ctx, envs.ket, envs.operator = ctx, None, H
    x = np.random.default_rng(3).standard_normal((D, 4, D))
    y = imps.AC_hamiltonian(0, None, H, None, envs, prec=2)(x)
    assert 1e-8 < rel(y, oracle.hac_apply_dense(GLs[0], H.Ws[0], GRs[0], x)) < 2e-5
    c = np.random.default_rng(5).standard_normal((D, D))
    yc = imps.C_hamiltonian(0, None, H, None, envs, prec=2)(c)
    assert 1e-8 < rel(yc, oracle.hc_apply(GLs[0], GRs[0], c)) < 2e-5


def test_vumps_bf16x3_planes_refined_matches_fp64(imps, ctx, oracle):
    # TFIM h = 1.2, D = 128: split-BF16 planes bulk phase, FP64 from Galerkin 1e-4 on; the
    # converged energy equals the plain FP64 run's
    W = oracle.tfim_mpo(1.0, 1.2)
    H = imps.SparseIMPO([W])
    psi0 = oracle.randomimps(128, [2], seed=0)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)
    e = imps.GPUCache(st, H, ctx)
    ctx.call("imps_envs_reset")
    _, _, i64 = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-9, maxiter=300), envs=e)
    e = imps.GPUCache(st, H, ctx)
    ctx.call("imps_envs_reset")
    _, _, ib = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-9, maxiter=300, prec=2, refine_below=1e-4), envs=e)
    assert i64.converged and ib.converged and ib.prec_switch_iter > 1
    assert abs(ib.energy - i64.energy) < 1e-9


@pytest.mark.parametrize("model,D,N,sweeps", [("xxz", 12, 2, 40), ("tfim", 8, 1, 30)])
def test_idmrg_matches_oracle(imps, ctx, oracle, model, D, N, sweeps):
    # single-site IDMRG (idmrg.jl:12-86) through imps_idmrg_run against the oracle's restatement
    # from the same psi0: the per-sweep centre drift and the final energy; XXX also against the
    # reference test's bounds (test/algorithms/idmrg.jl: 2e-4 from 1/4 - ln 2)
    W = {"xxz": lambda: oracle.xxz_mpo(1.0, 1.0), "tfim": lambda: oracle.tfim_mpo(1.0, 1.0)}[model]()
    H = imps.SparseIMPO([W] * N)
    psi0 = oracle.randomimps(D, [H.d] * N, 3)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)
    e = imps.GPUCache(st, H, ctx)
    ctx.call("imps_envs_reset")
    psi, envs, info = imps.find_groundstate(st, H, imps.IDMRG(tol=1e-12, maxiter=sweeps), envs=e)
    ro = oracle.IDMRG(tol=1e-12, maxiter=sweeps)
    psi_o, envs_o, info_o = oracle.find_groundstate_idmrg(psi0, [W] * N, ro)
    assert info.niter == info_o.niter == sweeps
    # the whole trajectory (measured: last drift equal to 2e-10 relative, energy to 1e-15)
    lg, lo = np.array(info.losses), np.array(info_o.losses)
    assert lg.shape == lo.shape
    assert np.all(np.abs(lg - lo) <= 1e-6 * lo + 1e-12)
    eo = oracle.expectation_energy(psi_o, [W] * N, envs_o)
    assert abs(info.energy - eo) < 1e-10 * N
    assert abs(info.energy - imps.expectationvalue(psi, H, envs)) < 1e-12
    assert max(oracle.mixedcanonical_errors(oracle.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC))) < 1e-8
    if model == "xxz":
        assert abs(info.energy / N - (0.25 - math.log(2))) < 2e-4


def test_tdvp_imaginary_time_matches_oracle(imps, ctx, oracle):
    # TDVP time steps (tdvp.jl:33-52) with dt = -0.05 through imps_tdvp_timestep: the first ten
    # steps' energies against the oracle's restatement from the same psi0, then the reference
    # test's protocol (test/algorithms/tdvp.jl:28-35: t = 20, within 2e-3 of 1/4 - ln 2)
    W = oracle.xxz_mpo(1.0, 1.0)
    H = imps.SparseIMPO([W, W])
    psi0 = oracle.randomimps(8, [2, 2], 5)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)
    envs = imps.DMRGCache(st, H, ctx=ctx)
    t10 = -np.arange(0, 0.50001, 0.05)
    psi, envs, hist = imps.time_evolve(st, H, t10, imps.TDVP(), envs)
    _, _, hist_o = oracle.time_evolve(psi0, [W, W], t10)
    assert len(hist) == len(hist_o) == 11
    assert np.allclose(hist, hist_o, rtol=0, atol=1e-10)
    t_span = -np.arange(0, 20.0001, 0.05)
    psi, envs, hist = imps.time_evolve(psi, H, t_span[10:], imps.TDVP(), envs)
    assert abs(hist[-1] / 2 - (0.25 - math.log(2))) < 2e-3
    assert all(hist[i + 1] <= hist[i] + 1e-10 for i in range(len(hist) - 1))
    assert max(oracle.mixedcanonical_errors(oracle.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC))) < 1e-8
    # a complex dt (real time) promotes the state to complex (tdvp.jl:97-100) and keeps its norm
    e_before = imps.expectationvalue(psi, H, envs)
    pc, envs = imps.timestep(psi, H, -0.01j, imps.TDVP(), envs)
    assert pc.is_complex() and abs(imps.norm(pc) - 1) < 1e-8
    assert abs(imps.expectationvalue(pc, H, envs) - e_before) < 1e-6


def test_correlator_and_fidelity_match_oracle(imps, ctx, oracle):
    # correlator (correlators.jl:13-42) and dot/fidelity (canonicalmps.jl:262-282) on the device
    # against the oracle on a random XXZ-cell state, plus the reference test's product states
    W = oracle.xxz_mpo(1.0, 1.0)
    H = imps.SparseIMPO([W, W])
    Z, X = oracle.sigma_z(), oracle.sigma_x()
    psi = oracle.randomimps(16, [2, 2], 8)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, H, ctx=ctx)
    js = list(range(1, 9)) + [20, 57]
    got = imps.correlator(st, Z, X, 0, js, envs)
    ref = oracle.correlator(psi, Z, X, 0, js)
    assert np.max(np.abs(got - ref)) < 1e-12
    assert abs(imps.correlator(st, X, X, 1, 4, envs) - oracle.correlator(psi, X, X, 1, [4])[0]) < 1e-12
    psi2 = oracle.randomimps(16, [2, 2], 9)
    st2 = imps.CanonicalIMPS(psi2.AL, psi2.AR, psi2.C, psi2.AC)
    assert abs(imps.dot(st, st2, envs) - oracle.dot_psi(psi, psi2)) < 1e-10 * abs(oracle.dot_psi(psi, psi2))
    assert abs(imps.fidelity(st, st2, envs) - oracle.fidelity(psi, psi2)) < 1e-10
    assert abs(imps.fidelity(st, st, envs) - 1.0) < 1e-10
    assert abs(imps.infidelity(st, st2, envs) - (1.0 - oracle.fidelity(psi, psi2))) < 1e-10
    # the reference's phase-invariance check (test/states/mps.jl:57-65) in real arithmetic: site 1's
    # AL, AR and AC times -1 makes the dominant transfer eigenvalue negative, the fidelity stays 1
    stp = imps.CanonicalIMPS([-psi.AL[0], psi.AL[1]], [-psi.AR[0], psi.AR[1]], psi.C,
                             [-psi.AC[0], psi.AC[1]])
    assert imps.dot(st, stp, envs) < 0
    assert abs(imps.fidelity(st, stp, envs) - 1.0) < 1e-10
    up = oracle.prodimps([2, 2], [0, 0])
    stu = imps.CanonicalIMPS(up.AL, up.AR, up.C, up.AC)
    envs_u = imps.DMRGCache(stu, H, ctx=ctx)
    assert np.allclose(imps.correlator(stu, Z, Z, 0, range(1, 5), envs_u), 1.0, atol=1e-12)
    assert np.allclose(imps.correlator(stu, X, X, 0, range(1, 4), envs_u), 0.0, atol=1e-12)


def test_dominant_env_matches_oracle(imps, ctx, oracle):
    # dominant_env (envs.jl:71-99) on the device against the oracle: identity channel (closed form
    # lambda = 1, L = R = I / sqrt(D)) and a DenseIMPO channel with a real dominant eigenvalue (a unit
    # level plus a random perturbation, w = 3); a channel whose dominant pair is complex is refused
    W = oracle.xxz_mpo(1.0, 1.0)
    H = imps.SparseIMPO([W, W])
    psi = oracle.randomimps(12, [2, 2], 21)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, H, ctx=ctx)
    for side in ("left", "right"):
        lam, L = imps.dominant_env(st, envs, side=side)
        lam_o, L_o = oracle.dominant_env(psi, side=side)
        assert L.shape == (12, 1, 12)
        assert abs(lam - 1) < 1e-9 and abs(lam - lam_o) < 1e-9
        assert np.allclose(L[:, 0, :], np.eye(12) / np.sqrt(12), atol=1e-8)
        assert np.linalg.norm(L - L_o) < 1e-8
    rng = np.random.default_rng(8)
    Ws = []
    for _ in range(2):
        Wd = 0.3 * rng.standard_normal((3, 2, 3, 2))
        Wd[0, :, 0, :] += np.eye(2)
        Ws.append(np.asfortranarray(Wd))
    for side in ("left", "right"):
        lam, L = imps.dominant_env(st, envs, W=imps.DenseIMPO(Ws), side=side)
        lam_o, L_o = oracle.dominant_env(psi, Ws, side=side)
        assert L.shape == (12, 3, 12)
        assert abs(lam - lam_o) < 1e-9 * abs(lam_o)
        assert abs(np.linalg.norm(L) - 1) < 1e-12
        assert np.linalg.norm(L - L_o) < 1e-7
    th = 0.7
    R = np.array([[np.cos(th), np.sin(th)], [-np.sin(th), np.cos(th)]])
    Wr = np.asfortranarray(np.einsum("ij,us->iujs", R, np.eye(2)))
    with pytest.raises(imps._lib.ImpsError, match="complex"):
        imps.dominant_env(st, envs, W=[Wr, Wr], side="left", maxiter=20)
    # the context still works after the refused call
    lam, _ = imps.dominant_env(st, envs, side="left")
    assert abs(lam - 1) < 1e-9


def test_dense_expectation_matches_oracle(imps, ctx, oracle):
    # expectationvalue(psi, W::DenseIMPO) with DMRGCache(psi, W) (expval.jl:69-77; envs.jl:284-304,
    # 559-581) on the device against the oracle, environments included; for a one-site cell the
    # value is the dominant eigenvalue of the transfer map (AC = AL C and the left eigen-equation)
    W = oracle.xxz_mpo(1.0, 1.0)
    rng = np.random.default_rng(8)
    Ws = []
    for _ in range(2):
        Wd = 0.3 * rng.standard_normal((3, 2, 3, 2))
        Wd[0, :, 0, :] += np.eye(2)
        Ws.append(np.asfortranarray(Wd))
    psi = oracle.randomimps(12, [2, 2], 21)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, imps.SparseIMPO([W, W]), ctx=ctx)
    E, lefts, rights = imps.dense_expectation(st, imps.DenseIMPO(Ws), envs, return_envs=True)
    E_o = oracle.expectation_dense(psi, Ws)
    lefts_o, rights_o = oracle.dense_envs(psi, Ws)
    assert abs(E - E_o) < 1e-9 * abs(E_o)
    for l in range(2):
        assert np.linalg.norm(lefts[l] - lefts_o[l]) < 1e-7 * np.linalg.norm(lefts_o[l])
        assert np.linalg.norm(rights[l] - rights_o[l]) < 1e-7
    assert abs(imps.dense_expectation(st, [2.5 * x for x in Ws], envs) - 2.5 * E) < 1e-8 * abs(E)
    psi1 = oracle.randomimps(8, [2], 31)
    st1 = imps.CanonicalIMPS(psi1.AL, psi1.AR, psi1.C, psi1.AC)
    envs1 = imps.DMRGCache(st1, imps.SparseIMPO([W]), ctx=ctx)
    lam, _ = imps.dominant_env(st1, envs1, W=[Ws[0]])
    assert abs(imps.dense_expectation(st1, [Ws[0]], envs1) - lam) < 1e-9 * abs(lam)
    assert abs(lam - oracle.dominant_env(psi1, [Ws[0]], "left")[0]) < 1e-9 * abs(lam)


def test_checkpoint_and_resume_state(imps, ctx, oracle, tmp_path):
    # imps_checkpoint_state / imps_resume_state: the device state goes to a file and back into
    # another context bit for bit, and gives the same energy there
    W = oracle.xxz_mpo(1.0, 1.0)
    H = imps.SparseIMPO([W, W])
    psi = oracle.randomimps(10, [2, 2], 41)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, H, ctx=ctx)
    e0 = imps.expectationvalue(st, H, envs)
    path = tmp_path / "ck.imps"
    ctx.checkpoint(path)
    c2 = imps.Context(0)
    try:
        c2.set_mpo(H)
        c2.resume(path)
        back = c2.get_state()
        for xs, ys in zip((st.AL, st.AR, st.C, st.AC), (back.AL, back.AR, back.C, back.AC)):
            for x, y in zip(xs, ys):
                assert np.array_equal(x, y)
        envs2 = imps.DMRGCache(back, H, ctx=c2)
        assert abs(imps.expectationvalue(back, H, envs2) - e0) < 1e-12
        assert np.array_equal(imps.load_state(path).AC[1], st.AC[1])
    finally:
        c2.close()
    c3 = imps.Context(0)
    try:
        with pytest.raises(imps._lib.ImpsError, match="MPO not set"):
            c3.resume(path)
        c3.set_mpo(imps.SparseIMPO([W]))
        with pytest.raises(imps._lib.ImpsError, match="unit cell"):
            c3.resume(path)
        with pytest.raises(imps._lib.ImpsError, match="cannot open"):
            c3.resume(tmp_path / "missing.imps")
        # a d = 2 file into a d = 4 model (and the reverse) is refused before any buffer is sized
        c3.set_mpo(imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=2))
        with pytest.raises(imps._lib.ImpsError, match="physical dimension"):
            c3.resume(path)
        hub = oracle.randomimps(6, [4, 4], 42)
        c3.set_state(imps.CanonicalIMPS(hub.AL, hub.AR, hub.C, hub.AC))
        path4 = tmp_path / "ck4.imps"
        c3.checkpoint(path4)
        c3.set_mpo(H)
        with pytest.raises(imps._lib.ImpsError, match="physical dimension"):
            c3.resume(path4)
    finally:
        c3.close()


def test_python_example_sequence(imps, tmp_path):
    # the Python example of the integration notes at D=16: VUMPS, observables, IDMRG and a TDVP
    # step on one GPUCache, then a checkpoint; each call sees the state the previous one left
    H = imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=2)
    psi0, c = imps.randomimps(16, H, seed=2)
    try:
        envs = imps.GPUCache(psi0, H, c)
        psi, envs, info = imps.find_groundstate(psi0, H, imps.VUMPS(tol=1e-6, maxiter=60), envs=envs)
        assert np.isfinite(info.energy) and info.niter >= 1
        Z = np.diag([1.0, -1.0, -1.0, 1.0])
        cz = imps.correlator(psi, Z, Z, 0, range(1, 9), envs)
        assert cz.shape == (8,) and np.all(np.abs(cz) <= 1 + 1e-9)
        f = imps.fidelity(psi, psi0, envs)
        assert 0.0 <= f <= 1.0 + 1e-9
        assert abs(imps.fidelity(psi, psi, envs) - 1) < 1e-9
        lam, L = imps.dominant_env(psi, envs, side="left")
        assert abs(lam - 1) < 1e-8 and np.allclose(L[:, 0, :], np.eye(16) / 4, atol=1e-6)
        e_v = imps.expectationvalue(psi, H, envs)
        psi2, envs, info2 = imps.find_groundstate(psi, H, imps.IDMRG(tol=1e-6, maxiter=20), envs=envs)
        assert np.isfinite(info2.energy)
        psi3, envs = imps.timestep(psi2, H, -0.05, imps.TDVP(), envs)
        e3 = imps.expectationvalue(psi3, H, envs)
        assert np.isfinite(e3) and e3 < e_v + 1e-3 * abs(e_v)
        c.checkpoint(tmp_path / "psi.imps")
        psi4 = imps.load_state(tmp_path / "psi.imps")
        assert np.array_equal(psi4.AC[0], c.get_state().AC[0])
    finally:
        c.close()


def _pad_all_bonds(imps, psi, n, seed=0):
    """Every bond of psi grown by n with zero weight (the reference's _padbond!, test/testhelpers.jl:
    125-146, applied to all bonds): AL gets n orthonormal completing columns, AR n orthonormal
    completing rows, C and AC are zero-extended; AL C = AC = C AR still holds."""
    rng = np.random.default_rng(seed)

    def complete(M):
        R = rng.standard_normal((M.shape[0], n))
        for _ in range(2):
            R -= M @ (M.T @ R)
        return np.hstack([M, np.linalg.qr(R)[0]])
    D, d = psi.AL[0].shape[0], psi.AL[0].shape[1]
    Dn = D + n
    AL, AR, C, AC = [], [], [], []
    for l in range(len(psi.AL)):
        A = np.zeros((Dn, d, D))
        A[:D] = psi.AL[l]
        AL.append(np.asfortranarray(complete(A.reshape(Dn * d, D, order="F")).reshape(Dn, d, Dn, order="F")))
        B = np.zeros((D, d, Dn))
        B[:, :, :D] = psi.AR[l]
        AR.append(np.asfortranarray(complete(B.reshape(D, d * Dn, order="F").T).T.reshape(Dn, d, Dn, order="F")))
        Cp = np.zeros((Dn, Dn))
        Cp[:D, :D] = psi.C[l]
        C.append(np.asfortranarray(Cp))
        X = np.zeros((Dn, d, Dn))
        X[:D, :, :D] = psi.AC[l]
        AC.append(np.asfortranarray(X))
    return imps.CanonicalIMPS(AL, AR, C, AC)


def test_zero_padding_leaves_energy_and_fidelity_unchanged(imps, ctx, oracle):
    # test/algorithms/envs.jl:62-67: the same physical state with zero-weight bond directions
    # (here on every bond; the library keeps one bond dimension) has the same energy to 1e-10,
    # with the environments solved on the device around a singular C
    W = oracle.xxz_mpo(1.0, 1.0)
    H = imps.SparseIMPO([W, W])
    psi = oracle.randomimps(6, [2, 2], 5)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    e0 = imps.expectationvalue(st, H, imps.DMRGCache(st, H, ctx=ctx))
    pad = _pad_all_bonds(imps, st, 2)
    envs = imps.DMRGCache(pad, H, ctx=ctx)
    assert abs(imps.expectationvalue(pad, H, envs) - e0) < 1e-10
    assert abs(e0 - oracle.expectation_energy(psi, [W, W], oracle.dmrg_cache(psi, [W, W]))) < 1e-10
    assert abs(imps.fidelity(pad, pad, envs) - 1) < 1e-10


def test_xxx_ground_state_observables(imps, ctx, oracle):
    # test/algorithms/observables.jl:20-32 and twosite.jl:66-73 on a ground state found on the
    # device: <Sz^2> = 1/4, antiferromagnetic nearest-neighbour correlator, all correlators below
    # 1/4, and the SU(2) relation <Sz Sz> = e0 / 3 within the reference's 5e-3
    H = imps.models.heisenberg_hamiltonian(J=1.0, Delta=1.0, nsites=2)
    psi0, c = imps.randomimps(12, H, seed=1)
    try:
        envs = imps.GPUCache(psi0, H, c)
        psi, envs, info = imps.find_groundstate(psi0, H, imps.VUMPS(tol=1e-9, maxiter=200), envs=envs)
        e0 = 0.25 - np.log(2)
        assert abs(info.energy / 2 - e0) < 2e-4
        Sz = np.diag([0.5, -0.5])
        assert abs(imps.expectationvalue(psi, (0, Sz @ Sz), envs) - 0.25) < 1e-12
        czz = imps.correlator(psi, Sz, Sz, 0, range(1, 6), envs)
        assert czz[0] < 0 and np.all(czz < 0.26)
        assert abs(czz[0] - e0 / 3) < 5e-3
    finally:
        c.close()


def test_energy_gradient_matches_oracle_and_fd(imps, ctx, oracle):
    # reverse H_AC sandwich against the oracle on a random state
    Wx = oracle.xxz_mpo(1.0, 0.8)
    H = imps.SparseIMPO([Wx] * 2)
    psi = oracle.randomimps(16, [2, 2], seed=12)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    envs = imps.DMRGCache(st, H, tol=1e-12, ctx=ctx)
    G = imps.energy_gradient(st, H, envs)
    Gr = oracle.energy_gradient(psi, [Wx] * 2, oracle.dmrg_cache(psi, [Wx] * 2, tol=1e-12))
    for a, b in zip(G, Gr):
        assert rel(a, b) < 1e-9
    # dE/dDelta of the converged XXZ cell energy vs central differences (step 1e-5):
    # Delta enters only the SzSz channel block W[0, :, 3, :] = J Delta Sz
    def run(delta):
        Hd = imps.models.heisenberg_hamiltonian(J=1.0, Delta=delta, nsites=2)
        p0 = oracle.randomimps(16, [2, 2], seed=1)
        s0 = imps.CanonicalIMPS(p0.AL, p0.AR, p0.C, p0.AC)
        e = imps.GPUCache(s0, Hd, ctx)
        ctx.call("imps_envs_reset")
        return imps.find_groundstate(s0, Hd, imps.VUMPS(tol=1e-11, maxiter=500), envs=e), Hd
    d0, dd = 1.3, 1e-5
    (psi_c, envs_c, info), Hd = run(d0)
    Gc = imps.energy_gradient(psi_c, Hd, envs_c)
    sz = 0.5 * oracle.sigma_z()
    grad = sum(np.sum(g[0, :, 3, :] * sz) for g in Gc)
    fd = (run(d0 + dd)[0][2].energy - run(d0 - dd)[0][2].energy) / (2 * dd)
    assert info.itererr < 1e-9
    assert abs(grad - fd) < 1e-6


def test_vumps_finalize_and_solver_objects(imps, ctx, oracle):
    # VUMPS(finalize, alg_eigsolve, alg_gauge, alg_environments) (algdefs.jl:26-35, vumps.jl:51-86):
    # an identity finalize leaves the trajectory unchanged bit for bit and runs once per iteration;
    # a finalize that hands back a sign-regauged copy of the state (the same physical state) makes
    # the library rebuild the environments and still converges to the same energy; fixed-tolerance
    # solver objects converge to the oracle's energy; an exception inside finalize propagates
    W = oracle.tfim_mpo()
    H = imps.SparseIMPO([W] * 2)
    psi0 = oracle.randomimps(12, [2, 2], seed=0)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)

    def run(alg):
        envs = imps.GPUCache(st, H, ctx)
        ctx.call("imps_envs_reset")
        return imps.find_groundstate(st, H, alg, envs=envs)

    _, _, base = run(imps.VUMPS(tol=1e-10))
    calls = []

    def ident(it, psi, HH, envs):
        calls.append(it)
        assert HH is H and len(psi) == 2
        return psi, envs

    _, _, i1 = run(imps.VUMPS(tol=1e-10, finalize=ident))
    assert calls == list(range(1, i1.niter + 1))
    assert i1.losses == base.losses

    def flip(it, psi, HH, envs):
        if it == 3:
            psi = imps.CanonicalIMPS([-psi.AL[0], psi.AL[1]], [-psi.AR[0], psi.AR[1]], psi.C,
                                     [-psi.AC[0], psi.AC[1]])
        return psi, envs

    _, _, i2 = run(imps.VUMPS(tol=1e-10, finalize=flip))
    assert i2.converged and abs(i2.energy - base.energy) < 1e-9
    # the rebuilt environments differ from the warm-started ones at the solver tolerance
    assert abs(i2.losses[3] - base.losses[3]) < 1e-3 * base.losses[3]
    fixed = imps.VUMPS(tol=1e-10, alg_eigsolve=imps.alg_eigsolve(dynamic_tols=False, tol=1e-12),
                       alg_gauge=imps.alg_gauge(dynamic_tols=False, tol=1e-13),
                       alg_environments=imps.alg_environments(dynamic_tols=False, tol=1e-12))
    _, _, i3 = run(fixed)
    _, _, ref = oracle.find_groundstate(psi0, [W] * 2, oracle.VUMPS(tol=1e-10))
    assert i3.converged and abs(i3.energy - ref.energies[-1]) < 1e-9

    def boom(it, psi, HH, envs):
        raise RuntimeError("finalize failed on purpose")

    with pytest.raises(RuntimeError, match="on purpose"):
        run(imps.VUMPS(tol=1e-10, finalize=boom))
    # IDMRG's finalize runs after every sweep
    seen = []
    envs = imps.GPUCache(st, H, ctx)
    ctx.call("imps_envs_reset")
    _, _, ii = imps.find_groundstate(st, H, imps.IDMRG(tol=1e-8, maxiter=5,
                                                         finalize=lambda it, p, h, e: (seen.append(it), (p, e))[1]),
                                     envs=envs)
    assert seen == list(range(1, ii.niter + 1))


def test_graph_captured_gauge_sweeps_match_eager(imps, ctx, oracle):
    # the gauge sweeps run as captured hipGraphs replayed from the second sweep on (many replays per
    # VUMPS iteration, with the eigensolves' eager work and host synchronisations in between): the
    # whole trajectory and the final state equal the eager run's bit for bit, twice in a row
    lib = imps._lib.load()
    H = imps.models.fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=2)
    psi0 = oracle.randomimps(24, [4, 4], seed=3)
    st = imps.CanonicalIMPS(psi0.AL, psi0.AR, psi0.C, psi0.AC)

    def run():
        envs = imps.GPUCache(st, H, ctx)
        ctx.call("imps_envs_reset")
        psi, _, info = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-14, maxiter=12), envs=envs)
        return psi, info

    try:
        assert lib.imps_set_graph_capture(0) == 0
        p0, i0 = run()
        assert lib.imps_set_graph_capture(1) == 0
        for _ in range(2):
            p1, i1 = run()
            assert i1.losses == i0.losses
            assert i1.timings["gauge_sweeps"] == i0.timings["gauge_sweeps"] > 2 * i0.niter
            for xs, ys in zip((p0.AL, p0.AR, p0.C, p0.AC), (p1.AL, p1.AR, p1.C, p1.AC)):
                for x, y in zip(xs, ys):
                    assert np.array_equal(x, y)
    finally:
        lib.imps_set_graph_capture(1)
    assert lib.imps_set_graph_capture(2) == -1
