"""Pins of the complex restatement (oracle/imps_complex.py), CPU only.

The complex oracle is the checker of the library's complex states (real-time TDVP); these
tests fix it against the real restatement on real inputs, against exact gauge covariance,
against a closed-form precession and against the conservation laws the reference's own TDVP
test asserts (test/algorithms/tdvp.jl:15-24)."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def cx():
    from oracle import imps_complex
    return imps_complex


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.mark.parametrize("model,N,D,seed", [("tfim", 1, 6, 3), ("xxz", 2, 5, 4)])
def test_complex_envs_and_energy_match_real_oracle_on_real_states(oracle, cx, model, N, D, seed):
    W = oracle.tfim_mpo(1.0, 0.7) if model == "tfim" else oracle.xxz_mpo(1.0, 0.6)
    Ws = [W] * N
    psi = oracle.randomimps(D, [2] * N, seed)
    renv = oracle.dmrg_cache(psi, Ws, tol=1e-14)
    cenv = cx.dmrg_cache(cx.from_real(psi), Ws)
    for l in range(N):
        assert _rel(cenv.leftenv(l), renv.leftenv(l)) < 1e-9
        assert _rel(cenv.rightenv(l), renv.rightenv(l)) < 1e-9
    e_r = oracle.expectation_energy(psi, Ws, renv)
    e_c = cx.expectation_energy(cx.from_real(psi), Ws, cenv)
    assert abs(e_c.imag) < 1e-13 and abs(e_c.real - e_r) < 1e-11


def test_complex_imaginary_time_step_matches_real_oracle(oracle, cx):
    # real dt on a real state: the complex step (exact exponentials) equals the real restatement
    # (Lanczos exponentials at tol 1e-13) and stays real
    W = oracle.xxz_mpo(1.0, 0.6)
    Ws = [W, W]
    psi = oracle.randomimps(4, [2, 2], 11)
    renv = oracle.dmrg_cache(psi, Ws, tol=1e-14)
    alg = oracle.TDVP(tol=1e-13, gauge_tol=1e-14, gauge_maxiter=2000)
    r2, _ = oracle.tdvp_timestep(psi, Ws, renv, -0.05, alg)
    c2, _ = cx.tdvp_timestep(cx.from_real(psi), Ws, cx.dmrg_cache(cx.from_real(psi), Ws), -0.05)
    for a, b in zip(c2.AL + c2.AR + c2.C + c2.AC, r2.AL + r2.AR + r2.C + r2.AC):
        assert _rel(a, b) < 1e-9
        assert float(np.abs(a.imag).max()) < 1e-10


@pytest.mark.parametrize("dt", [-0.05, -0.04j, -0.03 - 0.02j])
def test_complex_step_is_covariant_under_phase_gauges(oracle, cx, dt):
    # the same state in a complex diagonal gauge: the step's AL, AR, AC, C transform with the
    # same phases (QRpos/LQpos commute with them) and every gauge-invariant value agrees; a
    # misplaced conjugation breaks this
    W = oracle.xxz_mpo(1.0, 0.8)
    Ws = [W, W]
    psi = cx.canonical_from_tensors([a + 0.3j * b for a, b in zip(oracle.random_tensors(4, [2, 2], 5),
                                                                    oracle.random_tensors(4, [2, 2], 6))])
    rng = np.random.default_rng(7)
    ph = [rng.uniform(-math.pi, math.pi, 4) for _ in range(2)]
    psig = cx.phase_gauge(psi, ph)
    e1 = cx.expectation_energy(psi, Ws, cx.dmrg_cache(psi, Ws))
    e2 = cx.expectation_energy(psig, Ws, cx.dmrg_cache(psig, Ws))
    assert abs(e1 - e2) < 1e-12 and abs(e1.imag) < 1e-12
    p1, env1 = cx.tdvp_timestep(psi, Ws, cx.dmrg_cache(psi, Ws), dt)
    p2, env2 = cx.tdvp_timestep(psig, Ws, cx.dmrg_cache(psig, Ws), dt)
    q = cx.phase_gauge(p1, ph)
    for a, b in zip(p2.AL + p2.AR + p2.C + p2.AC, q.AL + q.AR + q.C + q.AC):
        assert _rel(a, b) < 1e-9
    assert abs(cx.expectation_energy(p1, Ws, env1) - cx.expectation_energy(p2, Ws, env2)) < 1e-11


def test_product_state_precession_is_exact(oracle, cx):
    # H = -h sum sz, |+x> at D = 1: the effective Hamiltonian is -h sz, TDVP is exact and
    # (<sx>, <sy>)(t) = (cos 2ht, -sin 2ht) (real time, dt = -0.1 i; the sign of <sy> fixes the
    # sign of i in the exponent)
    h = 0.7
    W = oracle.tfim_mpo(0.0, h)
    a = np.zeros((1, 2, 1), dtype=complex)
    a[0, :, 0] = [1 / math.sqrt(2), 1 / math.sqrt(2)]
    psi = cx.canonical_from_tensors([a])
    env = cx.dmrg_cache(psi, [W])
    sx = oracle.sigma_x()
    sy = np.array([[0, -1j], [1j, 0]])
    for k in range(1, 11):
        psi, env = cx.tdvp_timestep(psi, [W], env, -0.1j)
        ex, ey = cx.larmor(h, 0.1 * k)
        assert abs(cx.local_expectation1(psi, 0, sx) - ex) < 1e-12
        assert abs(cx.local_expectation1(psi, 0, sy) - ey) < 1e-12


def test_real_time_from_ground_state_conserves_energy_and_norm(oracle, cx):
    # test/algorithms/tdvp.jl:15-24: from a VUMPS ground state (XXX, two-site cell), ten real-time
    # steps dt = -0.01 i keep the energy to 1e-6 and the norm to 1e-8
    W = oracle.xxz_mpo(1.0, 1.0)
    Ws = [W, W]
    psi0 = oracle.randomimps(6, [2, 2], 1)
    g, genv, _ = oracle.find_groundstate(psi0, Ws, oracle.VUMPS(tol=1e-9, maxiter=400))
    psi = cx.from_real(g)
    env = cx.dmrg_cache(psi, Ws)
    e0 = cx.expectation_energy(psi, Ws, env).real / 2
    for _ in range(10):
        psi, env = cx.tdvp_timestep(psi, Ws, env, -0.01j)
    e1 = cx.expectation_energy(psi, Ws, env).real / 2
    assert abs(e1 - e0) < 1e-6
    assert abs(np.linalg.norm(psi.AC[0]) - 1) < 1e-8
    # the state is no longer real (a phase per site at least)
    assert float(np.abs(psi.AC[0].imag).max()) > 1e-4


def test_complex_conjugation_symmetry(oracle, cx):
    # H real: stepping conj(psi) with conj(dt) gives conj of stepping psi with dt
    W = oracle.tfim_mpo(1.0, 0.9)
    psi = cx.canonical_from_tensors([oracle.random_tensors(5, [2], 8)[0] + 0.5j * oracle.random_tensors(5, [2], 9)[0]])
    pc = cx.CanonicalIMPS(*[[x.conj() for x in xs] for xs in (psi.AL, psi.AR, psi.C, psi.AC)])
    dt = -0.02 - 0.07j
    p1, _ = cx.tdvp_timestep(psi, [W], cx.dmrg_cache(psi, [W]), dt)
    p2, _ = cx.tdvp_timestep(pc, [W], cx.dmrg_cache(pc, [W]), np.conj(dt))
    for a, b in zip(p1.AL + p1.C, p2.AL + p2.C):
        assert _rel(a.conj(), b) < 1e-10


def _embed(Z):
    # each complex entry as the real block [[re, -im], [im, re]] on interleaved rows and columns
    D1, m, D2 = Z.shape
    R = np.zeros((2 * D1, m, 2 * D2))
    R[0::2, :, 0::2] = Z.real
    R[1::2, :, 0::2] = Z.imag
    R[0::2, :, 1::2] = -Z.imag
    R[1::2, :, 1::2] = Z.real
    return R


def _extract(R, scale=1.0):
    return scale * (0.5 * (R[0::2, :, 0::2] + R[1::2, :, 1::2]) + 0.5j * (R[1::2, :, 0::2] - R[0::2, :, 1::2]))


def test_real_algorithms_on_the_embedding_equal_the_complex_restatement(oracle, cx):
    # the library runs complex states as their real embedding through its real algorithms; the
    # real restatement applied to the embedded tensors reproduces the complex restatement
    # (canonical form, environments' energy and an imaginary-time TDVP step), with C and AC
    # scaled by 1/sqrt(2)
    W = oracle.xxz_mpo(1.0, 0.8)
    Ws = [W, W]
    raw = [a + 0.4j * b for a, b in zip(oracle.random_tensors(4, [2, 2], 21), oracle.random_tensors(4, [2, 2], 1021))]
    ref = cx.canonical_from_tensors(raw)
    emb = oracle.canonical_from_tensors([_embed(a) for a in raw], tol=1e-14, maxiter=2000)
    s2 = math.sqrt(2.0)
    for x, y in zip(emb.AL + emb.AR, ref.AL + ref.AR):
        assert _rel(_extract(x), y) < 1e-12
    for x, y in zip(emb.C, ref.C):
        assert _rel(_extract(x[:, None, :], s2)[:, 0, :], y) < 1e-12
    e_emb = oracle.expectation_energy(emb, Ws, oracle.dmrg_cache(emb, Ws, tol=1e-14))
    e_ref = cx.expectation_energy(ref, Ws, cx.dmrg_cache(ref, Ws))
    assert abs(e_emb - e_ref.real) < 1e-11 and abs(e_ref.imag) < 1e-12
    alg = oracle.TDVP(tol=1e-13, gauge_tol=1e-14, gauge_maxiter=2000)
    e2, _ = oracle.tdvp_timestep(emb, Ws, oracle.dmrg_cache(emb, Ws, tol=1e-14), -0.05, alg)
    r2, _ = cx.tdvp_timestep(ref, Ws, cx.dmrg_cache(ref, Ws), -0.05)
    for x, y in zip(e2.AL, r2.AL):
        assert _rel(_extract(x), y) < 1e-9


def test_hessenberg_eigenvalues_with_a_nearly_defective_pair():
    # an Arnoldi Hessenberg matrix of an embedded complex transfer map (17 x 17, dominant real
    # eigenvalue doubled and split by rounding into a nearly defective pair): the library's
    # host eigenvalue routine must converge and agree with LAPACK
    import ctypes
    import os
    from infinitempsalgorithms_amd import _lib
    L = _lib.load()
    H = np.asfortranarray(np.load(os.path.join(os.path.dirname(__file__), "golden", "hessenberg_paired_spectrum.npy")))
    n = H.shape[0]
    wr, wi = np.zeros(n), np.zeros(n)
    st = L.imps_host_hess_eigvals(n, H.ctypes.data_as(_lib.P_double), n, wr.ctypes.data_as(_lib.P_double),
                                  wi.ctypes.data_as(_lib.P_double))
    assert st == 0
    got = np.sort(np.abs(wr + 1j * wi))
    want = np.sort(np.abs(np.linalg.eigvals(H)))
    assert np.max(np.abs(got - want)) < 1e-7 * want.max()
    lam, u = ctypes.c_double(), np.zeros(n)
    assert L.imps_host_hess_dominant(n, H.ctypes.data_as(_lib.P_double), n, ctypes.byref(lam),
                                     u.ctypes.data_as(_lib.P_double)) in (0, 1)
    assert abs(lam.value - want.max()) < 1e-7 * want.max()


def test_complex_hessenberg_dominant_eigenpair():
    # the complex Arnoldi projection's dominant eigenpair (host routine) against LAPACK, on a
    # random complex Hessenberg matrix and on one whose dominant eigenvalue is nearly real
    import ctypes
    from infinitempsalgorithms_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(5)
    for shift in (0.0, 3.0):
        n = 12
        Hm = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), -1)
        Hm[0, 0] += shift
        Hf = np.asfortranarray(Hm)
        lam = np.zeros(2)
        x = np.zeros(n, dtype=complex)
        last = ctypes.c_double()
        assert L.imps_host_hess_dominant_z(n, Hf.ctypes.data_as(_lib.P_double), n, lam.ctypes.data_as(_lib.P_double),
                                           x.ctypes.data_as(_lib.P_double), ctypes.byref(last)) == 0
        w, V = np.linalg.eig(Hm)
        k = int(np.argmax(np.abs(w)))
        assert abs(complex(lam[0], lam[1]) - w[k]) < 1e-12 * abs(w[k])
        assert np.linalg.norm(Hm @ x - w[k] * x) < 1e-10 and abs(np.linalg.norm(x) - 1) < 1e-12
        assert abs(last.value - abs(x[-1])) < 1e-15


def test_complex_correlator_pins(oracle, cx):
    # (1) real state, real operators: the real restatement; (2) neighbouring sites: the two-site
    # wavefunction theta = AC_i AR_{i+1} contracted with O1 (x) O2 directly; (3) a product state
    # (D = 1): <O1_i O2_j> = <O1><O2> for every j > i
    sx, sz = oracle.sigma_x(), oracle.sigma_z()
    rpsi = oracle.randomimps(4, [2, 2], 8)
    got = cx.correlator(cx.from_real(rpsi), sx, sz, 0, [1, 2, 5])
    want = oracle.correlator(rpsi, sx, sz, 0, [1, 2, 5])
    assert np.max(np.abs(got - want)) < 1e-13
    psi = cx.canonical_from_tensors([a + 0.5j * b for a, b in zip(oracle.random_tensors(3, [2, 2], 12),
                                                                   oracle.random_tensors(3, [2, 2], 13))])
    rng = np.random.default_rng(2)
    O1 = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    O2 = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    theta = np.einsum("asb,btc->astc", psi.AC[0], psi.AR[1])
    brute = np.einsum("auvc,us,vt,astc->", theta.conj(), O1, O2, theta) / np.vdot(theta, theta)
    assert abs(cx.correlator(psi, O1, O2, 0, [1])[0] - brute) < 1e-12
    a = [np.array([0.6, 0.8j]).reshape(1, 2, 1), np.array([0.8, -0.6 + 0.0j]).reshape(1, 2, 1)]
    prod = cx.canonical_from_tensors(a)
    e1 = [cx.local_expectation1(prod, l, O1) for l in (0, 1)]
    e2 = [cx.local_expectation1(prod, l, O2) for l in (0, 1)]
    vals = cx.correlator(prod, O1, O2, 0, [1, 2, 3])
    for j, v in zip([1, 2, 3], vals):
        assert abs(v - e1[0] * e2[j % 2]) < 1e-12
