"""Rectangular Schur chains and per-site physical dimensions on the device (imps_set_mpo_chain),
through the C ABI, at the reference's own test cases and closed forms:

- the dimerised TFIM chain of test/operators/mpo.jl:94-140 (tensor shapes 2x3, 3x2, 2x2, bond
  levels [3, 2, 2]): ground-state energy per cell -sqrt(4h^2 + J^2) - |h| with VUMPS and IDMRG,
  environment level counts per bond, and equality with the reference's hand-built square form;
- per-site physical dimensions [2, 3, 2] (test/algorithms/envs.jl:77-130): on-site models whose
  ground state is a product state, E = sum_l min eig(h_l), with VUMPS and IDMRG, physical
  dimensions preserved, identity expectation values 1;
- an interacting heterogeneous cell: TFIM on a d = 2 and a d = 3 site whose third state is a
  decoupled sector raised by +10, so that the energy per cell is 2 e_TFIM(h) in closed form;
- DenseIMPO bonds padded with zeros (test/algorithms/envs.jl:66-70): the same value.

The library runs on the chain's exact square embedding; the tests also check that the empty
physical sectors of the device state hold exact zeros after the solvers.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import integrate

pytestmark = pytest.mark.gpu

SX = np.array([[0.0, 1.0], [1.0, 0.0]])
SZ = np.diag([1.0, -1.0])
I2 = np.eye(2)


@pytest.fixture(scope="module")
def imps():
    import infinitempsalgorithms_amd as m
    return m


def tensor(wl, wr, d, blocks):
    W = np.zeros((wl, d, wr, d), order="F")
    for (i, j), B in blocks.items():
        W[i, :, j, :] = B
    return W


def dimer_chain(h, J):
    """test/operators/mpo.jl:97-107: the NN coupling only on bond 1 (levels [3, 2, 2])."""
    W1 = tensor(2, 3, 2, {(0, 0): I2, (1, 2): I2, (0, 2): -h * SZ, (0, 1): -J * SX})
    W2 = tensor(3, 2, 2, {(0, 0): I2, (2, 1): I2, (0, 1): -h * SZ, (1, 1): SX})
    W3 = tensor(2, 2, 2, {(0, 0): I2, (1, 1): I2, (0, 1): -h * SZ})
    return [W1, W2, W3]


def dimer_square(h, J):
    """The reference's uniform 3 x 3 form of the same operator (test/operators/mpo.jl:127-137)."""
    V1 = tensor(3, 3, 2, {(0, 0): I2, (2, 2): I2, (0, 2): -h * SZ, (0, 1): -J * SX})
    V2 = tensor(3, 3, 2, {(0, 0): I2, (2, 2): I2, (0, 2): -h * SZ, (1, 2): SX})
    V3 = tensor(3, 3, 2, {(0, 0): I2, (2, 2): I2, (0, 2): -h * SZ})
    return [V1, V2, V3]


def onsite_chain(hs):
    """Schur chain with on-site terms only (mpohamiltonian(h_l, []), w = 2) per site."""
    return [tensor(2, 2, h.shape[0], {(0, 0): np.eye(h.shape[0]), (1, 1): np.eye(h.shape[0]), (0, 1): h})
            for h in hs]


def padded_state(ctx):
    """The context's state in the library's square form (imps_get_state), AL and AC."""
    N, D, d = ctx.N, ctx.D, ctx.d
    AL = [np.zeros((D, d, D), order="F") for _ in range(N)]
    AR = [np.zeros((D, d, D), order="F") for _ in range(N)]
    C = [np.zeros((D, D), order="F") for _ in range(N)]
    AC = [np.zeros((D, d, D), order="F") for _ in range(N)]
    keep = [ctx_ptrs(x) for x in (AL, AR, C, AC)]
    ctx.call("imps_get_state", *(k for k in keep))
    return AL, AR, AC


def ctx_ptrs(xs):
    arr = (ctypes.POINTER(ctypes.c_double) * len(xs))()
    for i, x in enumerate(xs):
        arr[i] = x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return arr


def assert_empty_sectors_zero(ctx, dims):
    for xs in padded_state(ctx):
        for l, x in enumerate(xs):
            assert not np.any(x[:, dims[l]:, :]), f"site {l}: weight in the empty sector"


def test_rectangular_chain_ground_state_closed_form(imps):
    h, J = 0.3, 1.0
    H = imps.SparseIMPO(dimer_chain(h, J))
    assert [H.bonddim(l) for l in range(3)] == [3, 2, 2] and H.max_bonddim() == 3
    with pytest.raises(ValueError):
        H.bonddim()
    E_exact = -math.sqrt(4 * h * h + J * J) - abs(h)
    for alg in (imps.VUMPS(D=8, tol=1e-12, maxiter=300), imps.IDMRG(D=8, tol=1e-12, maxiter=300)):
        psi, envs, info = imps.find_groundstate(None, H, alg, seed=3)
        try:
            e = imps.expectationvalue(psi, H, envs)
            print(type(alg).__name__, "dimer chain", e, E_exact, info.itererr, flush=True)
            assert abs(e - E_exact) < 1e-8, (e, E_exact)
            # environment levels per bond: leftenv(l) on the bond left of site l
            assert [imps.leftenv(envs, l).shape[1] for l in range(3)] == [2, 3, 2]
            assert [imps.rightenv(envs, l).shape[1] for l in range(3)] == [3, 2, 2]
            dW = imps.energy_gradient(psi, H, envs)
            assert [g.shape for g in dW] == [W.shape for W in H.Ws]
        finally:
            envs.ctx.close()


def test_rectangular_chain_equals_square_form(imps):
    h, J = 0.3, 1.0
    H = imps.SparseIMPO(dimer_chain(h, J))
    Hsq = imps.SparseIMPO(dimer_square(h, J))
    assert Hsq.square and not H.square
    psi, c = imps.randomimps(6, H, seed=5)
    try:
        e = imps.expectationvalue(psi, H, imps.DMRGCache(psi, H, tol=1e-13, ctx=c))
    finally:
        c.close()
    c2 = imps.Context(0)
    try:
        esq = imps.expectationvalue(psi, Hsq, imps.DMRGCache(psi, Hsq, tol=1e-13, ctx=c2))
    finally:
        c2.close()
    assert abs(e - esq) < 1e-10 * max(1.0, abs(esq)), (e, esq)


def test_heterogeneous_physical_dims_closed_forms(imps):
    dims = [2, 3, 2]
    lams = [np.array([-2.0, 1.0]), np.array([-3.0, 0.0, 2.0]), np.array([-2.0, 1.0])]
    H = imps.SparseIMPO(onsite_chain([np.diag(l) for l in lams]))
    assert H.phydims() == dims and H.bonddim() == 2
    E_exact = sum(l.min() for l in lams)
    for alg in (imps.VUMPS(D=4, tol=1e-10, maxiter=200), imps.IDMRG(D=4, tol=1e-10, maxiter=200)):
        psi, envs, info = imps.find_groundstate(None, H, alg, seed=1)
        try:
            e = imps.expectationvalue(psi, H, envs)
            print(type(alg).__name__, "diagonal heterogeneous", e, E_exact, flush=True)
            assert abs(e - E_exact) < 1e-8
            assert [a.shape for a in psi.AL] == [(4, dl, 4) for dl in dims]
            assert_empty_sectors_zero(envs.ctx, dims)
        finally:
            envs.ctx.close()
    # random symmetric on-site terms: a product ground state, E = sum of the lowest eigenvalues
    rng = np.random.default_rng(5)
    hs = []
    for dl in dims:
        a = rng.standard_normal((dl, dl))
        hs.append((a + a.T) / 2)
    H2 = imps.SparseIMPO(onsite_chain(hs))
    psi, envs, info = imps.find_groundstate(None, H2, imps.VUMPS(D=4, tol=1e-10, maxiter=200), seed=2)
    try:
        e = imps.expectationvalue(psi, H2, envs)
        assert abs(e - sum(np.linalg.eigvalsh(h)[0] for h in hs)) < 1e-8
        for l, dl in enumerate(dims):
            assert abs(imps.expectationvalue(psi, (l, np.eye(dl)), envs) - 1.0) < 1e-10
        assert_empty_sectors_zero(envs.ctx, dims)
    finally:
        envs.ctx.close()


def tfim_energy(h):
    """Energy per site of H = -sum sx sx - h sum sz: -(1/pi) int_0^pi sqrt(1 + h^2 - 2 h cos k) dk."""
    return -integrate.quad(lambda k: math.sqrt(1 + h * h - 2 * h * math.cos(k)), 0, math.pi,
                           epsabs=1e-14, epsrel=1e-14)[0] / math.pi


def test_heterogeneous_cell_with_decoupled_sector_matches_tfim(imps):
    """Site 0: a spin (d = 2); site 1: d = 3, the TFIM acting on its first two states and the third
    one raised by +10 and coupled to nothing.  The ground state has no weight in that state, so the
    energy per two-site cell is 2 e_TFIM(h)."""
    h, J = 1.3, 1.0
    X3 = np.zeros((3, 3))
    X3[:2, :2] = SX
    Z3 = np.zeros((3, 3))
    Z3[:2, :2] = SZ
    P = np.diag([0.0, 0.0, 10.0])
    W0 = tensor(3, 3, 2, {(0, 0): I2, (2, 2): I2, (0, 1): -J * SX, (1, 2): SX, (0, 2): -h * SZ})
    W1 = tensor(3, 3, 3, {(0, 0): np.eye(3), (2, 2): np.eye(3), (0, 1): -J * X3, (1, 2): X3, (0, 2): -h * Z3 + P})
    H = imps.SparseIMPO([W0, W1])
    assert H.phydims() == [2, 3] and not H.square
    psi, envs, info = imps.find_groundstate(None, H, imps.VUMPS(D=32, tol=1e-10, maxiter=500), seed=4)
    try:
        e = imps.expectationvalue(psi, H, envs)
        print("decoupled sector", e, 2 * tfim_energy(h), info.itererr, flush=True)
        assert abs(e - 2 * tfim_energy(h)) < 1e-7
        assert abs(imps.expectationvalue(psi, (1, np.diag([0.0, 0.0, 1.0])), envs)) < 1e-10
        assert_empty_sectors_zero(envs.ctx, [2, 3])
    finally:
        envs.ctx.close()


def test_dense_impo_padded_bond_same_value(imps):
    """DenseIMPO bonds of different sizes: padding one bond with zeros leaves the value unchanged
    (test/algorithms/envs.jl:66-70)."""
    rng = np.random.default_rng(8)
    Ws = []
    for _ in range(2):
        Wd = 0.3 * rng.standard_normal((3, 2, 3, 2))
        Wd[0, :, 0, :] += np.eye(2)
        Ws.append(np.asfortranarray(Wd))
    Wp = [np.zeros((3, 2, 4, 2), order="F"), np.zeros((4, 2, 3, 2), order="F")]
    Wp[0][:, :, :3, :] = Ws[0]
    Wp[1][:3, :, :, :] = Ws[1]
    H = imps.models.tfim_hamiltonian(J=1.0, h=0.8, nsites=2)
    psi, c = imps.randomimps(12, H, seed=21)
    try:
        envs = imps.DMRGCache(psi, H, ctx=c)
        E = imps.dense_expectation(psi, imps.DenseIMPO(Ws), envs)
        Ep, lefts, rights = imps.dense_expectation(psi, imps.DenseIMPO(Wp), envs, return_envs=True)
        assert abs(E - Ep) < 1e-10 * abs(E), (E, Ep)
        assert [G.shape[1] for G in lefts] == [3, 4] and [G.shape[1] for G in rights] == [4, 3]
    finally:
        c.close()


def test_state_with_weight_in_an_empty_sector_is_refused(imps):
    H = imps.SparseIMPO(onsite_chain([np.diag([-1.0, 1.0]), np.diag([-1.0, 0.0, 1.0])]))
    c = imps.Context(0)
    try:
        c.set_mpo(H)
        bad = [np.ones((2, 3, 2), order="F"), np.ones((2, 3, 2), order="F")]  # site 0 has d = 2
        pp = ctx_ptrs(bad)
        with pytest.raises(imps._lib.ImpsError):
            c.call("imps_set_state_raw", 2, 2, pp, 1e-12, 50)
    finally:
        c.close()
