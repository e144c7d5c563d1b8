"""States whose bond dimension differs from bond to bond (test/algorithms/envs.jl:31-73): the
environments, the canonical form, VUMPS, IDMRG and TDVP on the device against the oracle, the
profile kept by every algorithm, and zero-padding invariance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHI = [2, 4, 2]


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def imps():
    import infinitempsalgorithms_amd as m
    return m


def _lib_state(imps, p):
    return imps.CanonicalIMPS([x.copy() for x in p.AL], [x.copy() for x in p.AR], [x.copy() for x in p.C],
                              [x.copy() for x in p.AC])


def _tfim3(oracle, imps):
    W = oracle.tfim_mpo(1.0, 1.3)
    return [W] * 3, imps.SparseIMPO([W] * 3)


def _psi(oracle, imps, seed=7):
    As = imps.random_tensors(CHI, [2, 2, 2], seed)
    return As, oracle.canonical_from_tensors(As)


def test_nonuniform_environments_and_energy_match_oracle(imps, oracle):
    Ws, H = _tfim3(oracle, imps)
    _, psi = _psi(oracle, imps)
    ref = oracle.dmrg_cache(psi, Ws, tol=1e-13)
    print("bonds: environments of a [2, 4, 2] state", flush=True)
    with imps.Context(0) as c:
        st = _lib_state(imps, psi)
        envs = imps.DMRGCache(st, H, tol=1e-13, ctx=c)
        assert [imps.leftenv(envs, l).shape[0] for l in range(3)] == [2, 2, 4]
        assert [imps.rightenv(envs, l).shape[0] for l in range(3)] == [2, 4, 2]
        for l in range(3):
            assert rel(imps.leftenv(envs, l), ref.leftenv(l)) < 1e-8
            assert rel(imps.rightenv(envs, l), ref.rightenv(l)) < 1e-8
        E = imps.expectationvalue(st, H, envs)
        assert abs(E - oracle.expectation_energy(psi, Ws, ref)) < 1e-10
        sz = oracle.sigma_z()
        for l in range(3):
            assert abs(imps.expectationvalue(st, (l, sz), envs) - oracle.local_expectation1(psi, l, sz)) < 1e-12


def test_nonuniform_canonical_form_on_device(imps, oracle):
    Ws, H = _tfim3(oracle, imps)
    As, psi = _psi(oracle, imps, seed=9)
    print("bonds: CanonicalIMPS(As) of raw [2, 4, 2] tensors", flush=True)
    with imps.Context(0) as c:
        got, _ = imps.canonicalimps(As, H, ctx=c, tol=1e-14, maxiter=2000)
        assert got.bonddims() == CHI
        ref = oracle.canonical_from_tensors(As, tol=1e-14, maxiter=2000)
        for a, b in zip(got.AL + got.AR + got.C + got.AC, ref.AL + ref.AR + ref.C + ref.AC):
            assert a.shape == b.shape and rel(a, b) < 1e-9


def test_nonuniform_vumps_idmrg_and_tdvp_keep_the_profile(imps, oracle):
    # envs.jl test: VUMPS and IDMRG from the same [2, 4, 2] state reach the same energy (1e-8)
    # and keep the profile; VUMPS follows the oracle's Galerkin trajectory; one TDVP step equals
    # the oracle's
    Ws, H = _tfim3(oracle, imps)
    _, psi = _psi(oracle, imps, seed=11)
    g_o, env_o, info_o = oracle.find_groundstate(psi, Ws, oracle.VUMPS(tol=1e-10, maxiter=200))
    E_o = oracle.expectation_energy(g_o, Ws, env_o)
    print("bonds: VUMPS, IDMRG and a TDVP step on a [2, 4, 2] state", flush=True)
    with imps.Context(0) as c:
        st = _lib_state(imps, psi)
        g, envs, info = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-10, maxiter=200), envs=imps.GPUCache(st, H, c))
        assert g.bonddims() == CHI and info.converged
        E = imps.expectationvalue(g, H, envs)
        assert abs(E - E_o) < 1e-9
        k = min(6, len(info.losses), len(info_o.losses))
        for a, b in zip(info.losses[:k], info_o.losses[:k]):
            assert abs(a - b) <= 1e-6 * abs(b)
    with imps.Context(0) as c:
        st = _lib_state(imps, psi)
        g2, envs2, info2 = imps.find_groundstate(st, H, imps.IDMRG(tol=1e-10, maxiter=300), envs=imps.GPUCache(st, H, c))
        assert g2.bonddims() == CHI
        assert abs(imps.expectationvalue(g2, H, envs2) - E) < 1e-8
    ref_env = oracle.dmrg_cache(psi, Ws, tol=1e-13)
    p_o, _ = oracle.tdvp_timestep(psi, Ws, ref_env, -0.05, oracle.TDVP(tol=1e-12, gauge_tol=1e-14, gauge_maxiter=2000))
    with imps.Context(0) as c:
        st = _lib_state(imps, psi)
        envs = imps.DMRGCache(st, H, tol=1e-13, ctx=c)
        p, envs = imps.timestep(st, H, -0.05, imps.TDVP(tol=1e-12, gauge_tol=1e-14, gauge_maxiter=2000), envs)
        assert p.bonddims() == CHI
        for a, b in zip(p.AL + p.AR + p.C + p.AC, p_o.AL + p_o.AR + p_o.C + p_o.AC):
            assert a.shape == b.shape and rel(a, b) < 1e-8


def test_zero_padded_bond_leaves_the_energy_unchanged_on_device(imps, oracle):
    # test/algorithms/envs.jl:56-62 on the device
    Ws, H = _tfim3(oracle, imps)
    psi = oracle.randomimps(4, [2, 2, 2], 5)

    def pad(A, left, right):
        Y = np.zeros((A.shape[0] + left, A.shape[1], A.shape[2] + right))
        Y[:A.shape[0], :, :A.shape[2]] = A
        return Y

    C1 = np.zeros((5, 5))
    C1[:4, :4] = psi.C[1]
    pp = imps.CanonicalIMPS([psi.AL[0], pad(psi.AL[1], 0, 1), pad(psi.AL[2], 1, 0)],
                            [psi.AR[0], pad(psi.AR[1], 0, 1), pad(psi.AR[2], 1, 0)],
                            [psi.C[0], C1, psi.C[2]],
                            [psi.AC[0], pad(psi.AC[1], 0, 1), pad(psi.AC[2], 1, 0)])
    print("bonds: zero-padded bond [4, 5, 4]", flush=True)
    with imps.Context(0) as c:
        st = _lib_state(imps, psi)
        E0 = imps.expectationvalue(st, H, imps.DMRGCache(st, H, ctx=c))
    with imps.Context(0) as c:
        E1 = imps.expectationvalue(pp, H, imps.DMRGCache(pp, H, ctx=c))
    assert abs(E1 - E0) < 1e-10


def test_larger_profile_vumps_matches_oracle(imps, oracle):
    # bond dimensions [8, 16, 12] on a three-site XXZ cell: VUMPS energy and Galerkin trajectory
    # against the oracle, the profile kept
    W = oracle.xxz_mpo(1.0, 0.6)
    Ws = [W] * 3
    H = imps.SparseIMPO(Ws)
    chi = [8, 16, 12]
    psi = oracle.canonical_from_tensors(imps.random_tensors(chi, [2, 2, 2], 23))
    g_o, env_o, info_o = oracle.find_groundstate(psi, Ws, oracle.VUMPS(tol=1e-9, maxiter=300))
    print("bonds: VUMPS on a [8, 16, 12] XXZ cell", flush=True)
    with imps.Context(0) as c:
        st = _lib_state(imps, psi)
        g, envs, info = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-9, maxiter=300), envs=imps.GPUCache(st, H, c))
        assert g.bonddims() == chi
        assert abs(imps.expectationvalue(g, H, envs) - oracle.expectation_energy(g_o, Ws, env_o)) < 1e-8
        k = min(4, len(info.losses), len(info_o.losses))
        for a, b in zip(info.losses[:k], info_o.losses[:k]):
            assert abs(a - b) <= 1e-6 * abs(b)


# ---- complex states with a bond profile, against the complex restatement (oracle/imps_complex.py)


def _cx():
    from oracle import imps_complex
    return imps_complex


def _complex_profile_raw(imps, chi, seed):
    re = imps.random_tensors(chi, [2] * len(chi), seed)
    im = imps.random_tensors(chi, [2] * len(chi), seed + 1000)
    return [a + 0.4j * b for a, b in zip(re, im)]


def _lib_state_c(imps, p):
    return imps.CanonicalIMPS([np.array(x) for x in p.AL], [np.array(x) for x in p.AR], [np.array(x) for x in p.C],
                              [np.array(x) for x in p.AC])


def test_complex_nonuniform_canonical_form_and_environments(imps, oracle):
    cx = _cx()
    chi = [2, 4, 3]
    W = oracle.xxz_mpo(1.0, 0.8)
    Ws = [W] * 3
    H = imps.SparseIMPO(Ws)
    raw = _complex_profile_raw(imps, chi, 61)
    ref = cx.canonical_from_tensors(raw)
    print("bonds: complex CanonicalIMPS(As) of raw [2, 4, 3] tensors", flush=True)
    with imps.Context(0) as c:
        got, _ = imps.canonicalimps(raw, H, ctx=c, tol=1e-14, maxiter=2000)
        assert got.is_complex() and got.bonddims() == chi
        for a, b in zip(got.AL + got.AR + got.C + got.AC, ref.AL + ref.AR + ref.C + ref.AC):
            assert a.shape == b.shape and rel(a, b) < 1e-9
    renv = cx.dmrg_cache(ref, Ws)
    print("bonds: environments of a complex [2, 4, 3] state", flush=True)
    with imps.Context(0) as c:
        st = _lib_state_c(imps, ref)
        envs = imps.DMRGCache(st, H, tol=1e-13, ctx=c)
        for l in range(3):
            assert rel(imps.leftenv(envs, l), renv.leftenv(l)) < 1e-8
            assert rel(imps.rightenv(envs, l), renv.rightenv(l)) < 1e-8
        assert abs(imps.expectationvalue(st, H, envs) - cx.expectation_energy(ref, Ws, renv).real) < 1e-10
        back = c.get_state()
        assert back.bonddims() == chi and back.is_complex()
        for a, b in zip(back.AL + back.C, ref.AL + ref.C):
            assert rel(a, b) < 1e-13


@pytest.mark.parametrize("dt", [-0.04j, -0.03 - 0.02j])
def test_complex_nonuniform_tdvp_step_matches_oracle(imps, oracle, dt):
    cx = _cx()
    chi = [2, 4, 3]
    W = oracle.tfim_mpo(1.0, 0.9)
    Ws = [W] * 3
    H = imps.SparseIMPO(Ws)
    psi = cx.canonical_from_tensors(_complex_profile_raw(imps, chi, 71))
    env = cx.dmrg_cache(psi, Ws)
    p2, env2 = cx.tdvp_timestep(psi, Ws, env, dt)
    print(f"bonds: complex TDVP step on a [2, 4, 3] state, dt={dt}", flush=True)
    with imps.Context(0) as c:
        st = _lib_state_c(imps, psi)
        envs = imps.DMRGCache(st, H, tol=1e-13, ctx=c)
        g2, envs = imps.timestep(st, H, dt, imps.TDVP(tol=1e-12, gauge_tol=1e-14, gauge_maxiter=2000), envs)
        assert g2.is_complex() and g2.bonddims() == chi
        for a, b in zip(g2.AL + g2.AR + g2.C + g2.AC, p2.AL + p2.AR + p2.C + p2.AC):
            assert a.shape == b.shape and rel(a, b) < 1e-8
        assert abs(imps.expectationvalue(g2, H, envs) - cx.expectation_energy(p2, Ws, env2).real) < 1e-10


def test_complex_nonuniform_refuses_weight_outside_the_profile(imps, oracle):
    import ctypes

    from infinitempsalgorithms_amd import _lib
    chi = [2, 4, 3]
    H = imps.SparseIMPO([oracle.tfim_mpo(1.0, 0.9)] * 3)
    raw = _complex_profile_raw(imps, chi, 81)
    with imps.Context(0) as c:
        c.set_mpo(H)
        D = max(chi)
        padded = []
        for l, a in enumerate(raw):
            z = np.zeros((D, 2, D), dtype=complex, order="F")
            z[:a.shape[0], :, :a.shape[2]] = a
            padded.append(z)
        padded[1][3, 0, 3] = 1e-3j  # row 3 lies outside site 1's left bond (chi[0] = 2)
        pp, keep = _lib.zptr_array(padded)
        carr = (ctypes.c_int * 3)(*chi)
        with pytest.raises(_lib.ImpsError, match="outside its bond dimensions"):
            c.call("imps_set_state_raw_complex_nonuniform", 3, D, carr, pp, 1e-12, 100)
