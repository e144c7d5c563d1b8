"""The fixed-point response (adjoint of the VUMPS fixed point) on the device, through the C ABI.

Parity: the tangent-space Hessian apply and the full response against the oracle's restatement
(oracle/imps_response.py) on the same converged states and tangent vectors.  At the TFIM
configuration (J = 1, h = 1.3, D = 128) the response of <sigma_z> to h against the closed form
-e''(h) of the free-fermion energy (e'(h) = -<sigma_z>), and against central finite differences of
re-converged VUMPS runs; XXZ (Delta = 0.7, one-site cell) the response of the nearest-neighbour
<Sz Sz> to Delta against central finite differences.
"""
import math
import warnings

import numpy as np
import pytest
from scipy import integrate

pytestmark = pytest.mark.gpu

warnings.filterwarnings("ignore", message="gmres_linsolve not converged")


@pytest.fixture(scope="module")
def imps():
    import infinitempsalgorithms_amd as m
    return m


@pytest.fixture(scope="module")
def resp():
    from oracle import imps_response
    return imps_response


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def szsz_mpo():
    sz = np.diag([0.5, -0.5])
    Wo = np.zeros((3, 2, 3, 2))
    Wo[0, :, 0, :] = np.eye(2)
    Wo[2, :, 2, :] = np.eye(2)
    Wo[0, :, 1, :] = sz
    Wo[1, :, 2, :] = sz
    return Wo


def tfim_energy_d2(h):
    """e''(h) of e(h) = -(1/pi) int_0^pi sqrt(1 + h^2 - 2 h cos k) dk:
    -(1/pi) int_0^pi sin^2 k / (1 + h^2 - 2 h cos k)^(3/2) dk."""
    return -integrate.quad(lambda k: math.sin(k) ** 2 / (1 + h * h - 2 * h * math.cos(k)) ** 1.5, 0, math.pi,
                           epsabs=1e-14, epsrel=1e-14)[0] / math.pi


def _oracle_groundstate(oracle, W, D, seed=3):
    psi, _, info = oracle.find_groundstate(oracle.randomimps(D, [W.shape[1]], seed), [W],
                                           oracle.VUMPS(tol=1e-13, maxiter=1500))
    assert info.converged, info.itererr
    return psi, oracle.dmrg_cache(psi, [W], tol=1e-15)


def _device_state(imps, oracle_psi, W, ctx):
    H = imps.SparseIMPO([W])
    st = imps.CanonicalIMPS(oracle_psi.AL, oracle_psi.AR, oracle_psi.C, oracle_psi.AC)
    envs = imps.DMRGCache(st, H, tol=1e-14, ctx=ctx)
    return H, st, envs


@pytest.mark.parametrize("model", ["tfim", "xxz"])
def test_tangent_hessian_matches_oracle(imps, oracle, resp, model):
    W = oracle.tfim_mpo(1.0, 0.8) if model == "tfim" else oracle.xxz_mpo(1.0, 0.7)
    D = 6
    psi, envs_o = _oracle_groundstate(oracle, W, D)
    ctx = imps.Context(0)
    try:
        H, st, envs = _device_state(imps, psi, W, ctx)
        ts = resp.TangentSpace(psi, W, envs_o)
        rng = np.random.default_rng(11)
        for _ in range(2):
            B = rng.standard_normal((D, 2, D))
            ref = ts.hessian_apply(B)
            got = imps.tangent_hessian_apply(st, envs, np.asfortranarray(B))
            assert rel(got, ref) < 1e-9, rel(got, ref)
            # the output is a tangent vector: AL^T out = 0
            assert np.abs(np.einsum("asc,asb->cb", psi.AL[0], got)).max() < 1e-12 * np.linalg.norm(got) * D
    finally:
        ctx.close()


@pytest.mark.parametrize("case", ["tfim_sx", "xxz_szsz"])
def test_fixedpoint_gradient_matches_oracle(imps, oracle, resp, case):
    if case == "tfim_sx":
        W, D, Wobj = oracle.tfim_mpo(1.0, 0.8), 4, resp.objective_mpo_1site(oracle.sigma_x())
    else:
        W, D, Wobj = oracle.xxz_mpo(1.0, 0.7), 6, szsz_mpo()
    psi, envs_o = _oracle_groundstate(oracle, W, D)
    val_o, dW_o, info_o = resp.fixedpoint_gradient(psi, W, envs_o, Wobj, tol=1e-11)
    ctx = imps.Context(0)
    try:
        H, st, envs = _device_state(imps, psi, W, ctx)
        val, dW, info = imps.fixedpoint_gradient(st, H, envs, Wobj, tol=1e-11)
        assert info.converged and info.relres < 1e-9, info
        assert abs(val - val_o) < 1e-12
        assert rel(dW, dW_o) < 1e-7, rel(dW, dW_o)
        # the energy as objective has no response (its tangent gradient is the Galerkin residual)
        e, dWe, _ = imps.fixedpoint_gradient(st, H, envs, W)
        assert abs(e - oracle.expectation_energy(psi, [W], envs_o)) < 1e-12
        assert np.abs(dWe).max() < 1e-9
    finally:
        ctx.close()


def test_fixedpoint_gradient_refuses_unsupported(imps, oracle):
    W = oracle.tfim_mpo(1.0, 0.8)
    psi = oracle.randomimps(4, [2, 2], seed=0)
    ctx = imps.Context(0)
    try:
        H = imps.SparseIMPO([W, W])
        st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
        envs = imps.DMRGCache(st, H, ctx=ctx)
        with pytest.raises(ValueError):
            imps.fixedpoint_gradient(st, H, envs, imps.objective_mpo_1site(oracle.sigma_x()))
        with pytest.raises(imps._lib.ImpsError):
            ctx.call("imps_fixedpoint_gradient", 2, 2, imps._lib.dptr(np.asfortranarray(
                imps.objective_mpo_1site(oracle.sigma_x()))), 1e-10, 10, 1e-10, None,
                imps._lib.dptr(np.zeros((3, 2, 3, 2), order="F")), None)
    finally:
        ctx.close()


def _solve(imps, H, D, psi0=None, seed=0, tol=1e-12):
    c = imps.Context(0)
    if psi0 is None:
        psi0, c = imps.randomimps(D, H, seed=seed, ctx=c, tol=1e-12, maxiter=100)
    envs = imps.GPUCache(psi0, H, c)
    psi, envs, info = imps.find_groundstate(psi0, H, imps.VUMPS(tol=tol, maxiter=2000), envs=envs)
    assert info.itererr < 10 * tol, info.itererr
    return psi, envs, info, c


def _richardson_fd(obs, p, h1, h2):
    f1 = (obs(p + h1) - obs(p - h1)) / (2 * h1)
    f2 = (obs(p + h2) - obs(p - h2)) / (2 * h2)
    return f2 + (f2 - f1) * h2 ** 2 / (h1 ** 2 - h2 ** 2)


@pytest.mark.timeout(900)
def test_tfim_d128_response_of_sz_matches_closed_form_and_fd(imps, oracle):
    """d<sigma_z>/dh through the fixed point at J = 1, h = 1.3, D = 128: -e''(h) in closed form."""
    h = 1.3
    mk = lambda hh: imps.models.tfim_hamiltonian(J=1.0, h=hh, nsites=1)  # noqa: E731
    H = mk(h)
    psi, envs, info, c = _solve(imps, H, 128)
    Wobj = imps.objective_mpo_1site(oracle.sigma_z())
    try:
        val, dW, rinfo = imps.fixedpoint_gradient(psi, H, envs, Wobj, tol=1e-10)
        print("tfim d128 response", rinfo, flush=True)
        assert rinfo.converged and rinfo.relres < 1e-6, rinfo
        dWdh = mk(h + 1.0).Ws[0] - H.Ws[0]
        pred = float(np.sum(dW * dWdh))
        assert abs(val - imps.expectationvalue(psi, (0, oracle.sigma_z()), envs)) < 1e-12
        print("tfim d128 d<sz>/dh", pred, -tfim_energy_d2(h), flush=True)
        assert abs(pred + tfim_energy_d2(h)) < 1e-8, (pred, -tfim_energy_d2(h))
    finally:
        c.close()

    def obs(hh):
        Hq = mk(hh)
        ps, ev, _, cq = _solve(imps, Hq, 128, psi0=psi)
        try:
            return imps.expectationvalue(ps, (0, oracle.sigma_z()), ev)
        finally:
            cq.close()

    fd = _richardson_fd(obs, h, 2e-3, 1e-3)
    print("tfim d128 fd", fd, flush=True)
    assert abs(pred - fd) < 1e-7, (pred, fd)


@pytest.mark.timeout(900)
def test_xxz_response_of_szsz_matches_fd(imps, oracle):
    """d<Sz_i Sz_i+1>/dDelta through the fixed point, XXZ Delta = 0.7, one-site cell, D = 32."""
    p, D = 0.7, 32
    mk = lambda dl: imps.SparseIMPO([oracle.xxz_mpo(1.0, dl)])  # noqa: E731
    H = mk(p)
    psi, envs, info, c = _solve(imps, H, D, seed=2, tol=1e-11)
    Wobj = szsz_mpo()
    try:
        val, dW, rinfo = imps.fixedpoint_gradient(psi, H, envs, Wobj, tol=1e-10)
        print("xxz response", rinfo, flush=True)
        assert rinfo.converged and rinfo.relres < 1e-6, rinfo
        pred = float(np.sum(dW * (mk(p + 1.0).Ws[0] - H.Ws[0])))
    finally:
        c.close()

    def obs(q):
        Hq = mk(q)
        ps, ev, _, cq = _solve(imps, Hq, D, psi0=psi, tol=1e-11)
        try:
            sz = np.diag([0.5, -0.5])
            return imps.correlator(ps, sz, sz, 0, 1, ev)
        finally:
            cq.close()

    assert abs(val - obs(p)) < 1e-9
    fd = _richardson_fd(obs, p, 2e-3, 1e-3)
    print("xxz d<SzSz>/dDelta", pred, fd, flush=True)
    assert abs(pred - fd) <= 1e-6 * max(1.0, abs(fd)), (pred, fd)
