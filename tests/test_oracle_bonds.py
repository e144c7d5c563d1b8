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
