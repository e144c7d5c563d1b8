"""Pins of the oracle on states whose bond dimension differs from bond to bond (the per-site
Dl/Dr of DMRGCache, envs.jl:327-343; test/algorithms/envs.jl:31-73), CPU only."""
import numpy as np
import pytest


def _profile_tensors(chi, ds, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    N = len(chi)
    return [np.reshape(rng.standard_normal(chi[(l - 1) % N] * ds[l] * chi[l]), (chi[(l - 1) % N], ds[l], chi[l]),
                       order="F") for l in range(N)]


def _tfim_cell_energy_from_local_terms(oracle, psi, J, h):
    # sum over the cell of -J <sx_l sx_{l+1}> - h <sz_l>, from AC/AR contractions only
    N = len(psi)
    sx, sz = oracle.sigma_x(), oracle.sigma_z()
    E = 0.0
    for l in range(N):
        E += -h * oracle.local_expectation1(psi, l, sz)
        E += -J * oracle.correlator(psi, sx, sx, l, [l + 1])[0]
    return E


def test_nonuniform_environments_energy_equals_local_terms(oracle):
    J, h = 1.0, 1.3
    W = oracle.tfim_mpo(J, h)
    psi = oracle.canonical_from_tensors(_profile_tensors([2, 4, 2], [2, 2, 2], 7))
    assert [a.shape for a in psi.AL] == [(2, 2, 2), (2, 2, 4), (4, 2, 2)]
    env = oracle.dmrg_cache(psi, [W] * 3)
    assert [env.leftenv(l).shape[0] for l in range(3)] == [2, 2, 4]
    assert [env.rightenv(l).shape[0] for l in range(3)] == [2, 4, 2]
    E = oracle.expectation_energy(psi, [W] * 3, env)
    assert abs(E - _tfim_cell_energy_from_local_terms(oracle, psi, J, h)) < 1e-11


def test_zero_padded_bond_leaves_the_energy_unchanged(oracle):
    # test/algorithms/envs.jl:56-62: a bond padded with a zero-weight direction is the same state
    W = oracle.tfim_mpo(1.0, 1.3)
    psi = oracle.randomimps(4, [2, 2, 2], 5)
    E0 = oracle.expectation_energy(psi, [W] * 3, oracle.dmrg_cache(psi, [W] * 3))

    def padc(X):
        Y = np.zeros((X.shape[0] + 1, X.shape[1] + 1))
        Y[:X.shape[0], :X.shape[1]] = X
        return Y

    def pad(A, left, right):
        Y = np.zeros((A.shape[0] + left, A.shape[1], A.shape[2] + right))
        Y[:A.shape[0], :, :A.shape[2]] = A
        return Y

    AL = [psi.AL[0], pad(psi.AL[1], 0, 1), pad(psi.AL[2], 1, 0)]
    AR = [psi.AR[0], pad(psi.AR[1], 0, 1), pad(psi.AR[2], 1, 0)]
    AC = [psi.AC[0], pad(psi.AC[1], 0, 1), pad(psi.AC[2], 1, 0)]
    C = [psi.C[0], padc(psi.C[1]), psi.C[2]]
    pp = oracle.CanonicalIMPS(AL, AR, C, AC)
    assert [a.shape[2] for a in pp.AL] == [4, 5, 4]
    E1 = oracle.expectation_energy(pp, [W] * 3, oracle.dmrg_cache(pp, [W] * 3))
    assert abs(E1 - E0) < 1e-10


@pytest.mark.slow
def test_nonuniform_vumps_is_bracketed_by_the_uniform_runs(oracle):
    # the [2, 4, 2] manifold contains the D = 2 states and lies inside the D = 4 ones
    W = oracle.tfim_mpo(1.0, 1.3)
    Ws = [W] * 3
    alg = oracle.VUMPS(tol=1e-10, maxiter=300)
    psi = oracle.canonical_from_tensors(_profile_tensors([2, 4, 2], [2, 2, 2], 7))
    g, genv, _ = oracle.find_groundstate(psi, Ws, alg)
    assert [a.shape[2] for a in g.AL] == [2, 4, 2]
    E = oracle.expectation_energy(g, Ws, genv)
    e2 = []
    for D in (2, 4):
        gg, ee, _ = oracle.find_groundstate(oracle.randomimps(D, [2, 2, 2], 3), Ws, alg)
        e2.append(oracle.expectation_energy(gg, Ws, ee))
    assert e2[0] + 1e-9 >= E >= e2[1] - 1e-9
    assert abs(E - _tfim_cell_energy_from_local_terms(oracle, g, 1.0, 1.3)) < 1e-10


# ---- complex states with a bond profile (the complex restatement, oracle/imps_complex.py)


def _cx():
    from oracle import imps_complex
    return imps_complex


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def test_complex_nonuniform_environments_match_real_oracle(oracle):
    # a real [2, 4, 2] state through the complex restatement: per-bond environment shapes and
    # values equal to the real restatement's, the energy real and equal
    cx = _cx()
    W = oracle.xxz_mpo(1.0, 0.6)
    Ws = [W] * 3
    psi = oracle.canonical_from_tensors(_profile_tensors([2, 4, 2], [2, 2, 2], 13))
    renv = oracle.dmrg_cache(psi, Ws, tol=1e-14)
    cenv = cx.dmrg_cache(cx.from_real(psi), Ws)
    for l in range(3):
        assert cenv.leftenv(l).shape == renv.leftenv(l).shape
        assert cenv.rightenv(l).shape == renv.rightenv(l).shape
        assert _rel(cenv.leftenv(l), renv.leftenv(l)) < 1e-9
        assert _rel(cenv.rightenv(l), renv.rightenv(l)) < 1e-9
    e = cx.expectation_energy(cx.from_real(psi), Ws, cenv)
    assert abs(e.imag) < 1e-13 and abs(e.real - oracle.expectation_energy(psi, Ws, renv)) < 1e-11


def _complex_profile_tensors(chi, ds, seed):
    return [a + 0.4j * b for a, b in zip(_profile_tensors(chi, ds, seed), _profile_tensors(chi, ds, seed + 1000))]


def _embed(Z):
    # each complex entry as the real block [[re, -im], [im, re]] on interleaved rows and columns
    D1, m, D2 = Z.shape
    R = np.zeros((2 * D1, m, 2 * D2))
    R[0::2, :, 0::2] = Z.real
    R[1::2, :, 0::2] = Z.imag
    R[0::2, :, 1::2] = -Z.imag
    R[1::2, :, 1::2] = Z.real
    return R


def _extract(R):
    return 0.5 * (R[0::2, :, 0::2] + R[1::2, :, 1::2]) + 0.5j * (R[1::2, :, 0::2] - R[0::2, :, 1::2])


def test_complex_nonuniform_restatement_equals_the_real_one_on_the_embedding(oracle):
    # complex raw tensors with bond dimensions [2, 4, 3]: the real restatement run on their
    # embedding (profile [4, 8, 6]) gives the complex restatement's canonical form, energy and
    # imaginary-time step
    cx = _cx()
    W = oracle.tfim_mpo(1.0, 0.9)
    Ws = [W] * 3
    raw = _complex_profile_tensors([2, 4, 3], [2, 2, 2], 31)
    ref = cx.canonical_from_tensors(raw)
    assert [a.shape[2] for a in ref.AL] == [2, 4, 3]
    emb = oracle.canonical_from_tensors([_embed(a) for a in raw], tol=1e-14, maxiter=2000)
    for x, y in zip(emb.AL + emb.AR, ref.AL + ref.AR):
        assert _rel(_extract(x), y) < 1e-12
    e_emb = oracle.expectation_energy(emb, Ws, oracle.dmrg_cache(emb, Ws, tol=1e-14))
    e_ref = cx.expectation_energy(ref, Ws, cx.dmrg_cache(ref, Ws))
    assert abs(e_ref.imag) < 1e-12 and abs(e_emb - e_ref.real) < 1e-11
    alg = oracle.TDVP(tol=1e-13, gauge_tol=1e-14, gauge_maxiter=2000)
    e2, _ = oracle.tdvp_timestep(emb, Ws, oracle.dmrg_cache(emb, Ws, tol=1e-14), -0.05, alg)
    r2, _ = cx.tdvp_timestep(ref, Ws, cx.dmrg_cache(ref, Ws), -0.05)
    for x, y in zip(e2.AL, r2.AL):
        assert _rel(_extract(x), y) < 1e-9


@pytest.mark.parametrize("dt", [-0.04j, -0.03 - 0.02j])
def test_complex_nonuniform_step_is_covariant_under_phase_gauges(oracle, dt):
    # per-bond diagonal unitaries of the profile's sizes: the real-time step commutes with them
    cx = _cx()
    W = oracle.xxz_mpo(1.0, 0.8)
    Ws = [W] * 3
    chi = [2, 4, 3]
    psi = cx.canonical_from_tensors(_complex_profile_tensors(chi, [2, 2, 2], 41))
    rng = np.random.default_rng(9)
    ph = [rng.uniform(-np.pi, np.pi, c) for c in chi]
    psig = cx.phase_gauge(psi, ph)
    p1, env1 = cx.tdvp_timestep(psi, Ws, cx.dmrg_cache(psi, Ws), dt)
    p2, env2 = cx.tdvp_timestep(psig, Ws, cx.dmrg_cache(psig, Ws), dt)
    assert [a.shape[2] for a in p1.AL] == chi
    q = cx.phase_gauge(p1, ph)
    for a, b in zip(p2.AL + p2.AR + p2.C + p2.AC, q.AL + q.AR + q.C + q.AC):
        assert _rel(a, b) < 1e-9
    e1 = cx.expectation_energy(p1, Ws, env1)
    assert abs(e1 - cx.expectation_energy(p2, Ws, env2)) < 1e-11
    if dt.real == 0:  # a real-time step keeps the energy to the step's order
        e0 = cx.expectation_energy(psi, Ws, cx.dmrg_cache(psi, Ws))
        assert abs(e1.imag) < 1e-12 and abs(e1 - e0) < 5e-3


def test_complex_zero_padded_bond_leaves_the_energy_unchanged(oracle):
    # test/algorithms/envs.jl:56-62 for a complex state
    cx = _cx()
    W = oracle.tfim_mpo(1.0, 1.3)
    Ws = [W] * 3
    psi = cx.canonical_from_tensors(_complex_profile_tensors([4, 4, 4], [2, 2, 2], 51))
    E0 = cx.expectation_energy(psi, Ws, cx.dmrg_cache(psi, Ws))

    def pad(A, left, right):
        Y = np.zeros((A.shape[0] + left, A.shape[1], A.shape[2] + right), dtype=complex)
        Y[:A.shape[0], :, :A.shape[2]] = A
        return Y

    C1 = np.zeros((5, 5), dtype=complex)
    C1[:4, :4] = psi.C[1]
    pp = cx.CanonicalIMPS([psi.AL[0], pad(psi.AL[1], 0, 1), pad(psi.AL[2], 1, 0)],
                          [psi.AR[0], pad(psi.AR[1], 0, 1), pad(psi.AR[2], 1, 0)],
                          [psi.C[0], C1, psi.C[2]],
                          [psi.AC[0], pad(psi.AC[1], 0, 1), pad(psi.AC[2], 1, 0)])
    E1 = cx.expectation_energy(pp, Ws, cx.dmrg_cache(pp, Ws))
    assert abs(E1 - E0) < 1e-10
