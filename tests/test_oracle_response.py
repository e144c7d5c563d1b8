"""Pins of the fixed-point response restatement (oracle/imps_response.py) on the CPU.

The tangent-space Hessian is checked against the brute-force second derivative of the energy
density along the curve AL + t B C^-1 (every point canonicalised and its DMRGCache energy
taken), and the full response against central finite differences of converged VUMPS runs.
"""
import warnings

import numpy as np
import pytest

from oracle import imps_oracle as O
from oracle import imps_response as R

warnings.filterwarnings("ignore", message="gmres_linsolve not converged")


def _groundstate(W, D, seed=3, tol=1e-13, psi0=None):
    alg = O.VUMPS(tol=tol, maxiter=1500)
    psi, _, info = O.find_groundstate(psi0 if psi0 is not None else O.randomimps(D, [W.shape[1]], seed), [W], alg)
    assert info.converged, info.itererr
    return psi, O.dmrg_cache(psi, [W], tol=1e-15)


def _energy_of_tensor(A, W):
    psi = O.canonical_from_tensors([A], tol=1e-15, maxiter=4000)
    return O.expectation_energy(psi, [W], O.dmrg_cache(psi, [W], tol=1e-15))


def _szsz_density():
    sz = np.diag([0.5, -0.5])
    Wo = np.zeros((3, 2, 3, 2))
    Wo[0, :, 0, :] = np.eye(2)
    Wo[2, :, 2, :] = np.eye(2)
    Wo[0, :, 1, :] = sz
    Wo[1, :, 2, :] = sz
    return Wo


@pytest.mark.parametrize("model", ["tfim", "xxz"])
def test_hessian_matches_second_derivative_of_energy_density(model):
    W = O.tfim_mpo(1.0, 1.0) if model == "tfim" else O.xxz_mpo(1.0, 1.0)
    D = 4
    psi, envs = _groundstate(W, D)
    ts = R.TangentSpace(psi, W, envs)
    U, S, _ = np.linalg.svd(psi.C[0])
    rng = np.random.default_rng(7)
    # a tangent direction on the Schmidt components above 1e-2 keeps B C^-1 of order one, so that
    # the curve stays in its quadratic regime for the steps below
    X = rng.standard_normal((D, D)) @ np.diag((S >= 1e-2).astype(float)) @ U.T
    B = ts.project(np.einsum("ask,kb->asb", rng.standard_normal((D, 2, D)), X, optimize=True))
    B /= np.linalg.norm(B)
    q = float(np.sum(B * ts.hessian_apply(B)))
    Delta = np.einsum("asb,bc->asc", B, ts.Cinv, optimize=True)
    e0 = _energy_of_tensor(psi.AL[0], W)

    def d2(t):
        return (_energy_of_tensor(psi.AL[0] + t * Delta, W) - 2 * e0
                + _energy_of_tensor(psi.AL[0] - t * Delta, W)) / t ** 2

    t1, t2 = 3e-4, 1e-4
    a1, a2 = d2(t1), d2(t2)
    rich = a2 + (a2 - a1) * t2 ** 2 / (t1 ** 2 - t2 ** 2)
    assert abs(rich - 2 * q) <= 2e-6 * abs(2 * q), (rich, 2 * q)
    # the curvature term is a visible part of the second derivative: without it the check fails
    qh = float(np.sum(B * ts.heff_apply(B)))
    assert abs(rich - 2 * qh) > 50 * abs(rich - 2 * q)


def test_hessian_symmetric_positive_at_minimum():
    W = O.tfim_mpo(1.0, 0.8)
    D = 4
    psi, envs = _groundstate(W, D)
    ts = R.TangentSpace(psi, W, envs)
    AL = psi.AL[0]
    Q, _ = np.linalg.qr(AL.reshape(D * 2, D), mode="complete")
    VL = Q[:, D:].reshape(D, 2, D)
    n = D * D
    H = np.zeros((n, n))
    for k in range(n):
        X = np.zeros(n)
        X[k] = 1.0
        B = np.einsum("ask,kb->asb", VL, X.reshape(D, D))
        H[:, k] = np.einsum("ask,asb->kb", VL, ts.hessian_apply(B)).ravel()
    assert np.abs(H - H.T).max() < 1e-9
    assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() > 0.1


def _fd(obs, p, steps=(1e-3, 3e-4)):
    (h1, h2) = steps
    f1 = (obs(p + h1) - obs(p - h1)) / (2 * h1)
    f2 = (obs(p + h2) - obs(p - h2)) / (2 * h2)
    return f2 + (f2 - f1) * h2 ** 2 / (h1 ** 2 - h2 ** 2)


@pytest.mark.parametrize("case", ["tfim_sx", "xxz_szsz"])
def test_fixedpoint_gradient_matches_vumps_finite_differences(case):
    if case == "tfim_sx":
        mk, p, D, Wobj = (lambda h: O.tfim_mpo(1.0, h)), 0.8, 4, R.objective_mpo_1site(O.sigma_x())
    else:
        mk, p, D, Wobj = (lambda dl: O.xxz_mpo(1.0, dl)), 0.7, 6, _szsz_density()
    W = mk(p)
    psi, envs = _groundstate(W, D)
    val, dW, info = R.fixedpoint_gradient(psi, W, envs, Wobj, tol=1e-11)
    assert info["relres"] < 1e-9
    dWdp = (mk(p + 1e-6) - mk(p - 1e-6)) / 2e-6
    pred = float(np.sum(dW * dWdp))

    def obs(q):
        ps, _ = _groundstate(mk(q), D, psi0=psi)
        eo = O.dmrg_cache(ps, [Wobj], tol=1e-15)
        return O.expectation_energy(ps, [Wobj], eo)

    assert abs(val - obs(p)) < 1e-10
    fd = _fd(obs, p)
    assert abs(pred - fd) <= 1e-7 * max(1.0, abs(fd)), (pred, fd)


def test_energy_objective_has_no_response():
    """With the energy itself as objective the tangent gradient is the Galerkin residual (zero at
    the fixed point), so the response adds nothing to the Hellmann-Feynman derivative."""
    W = O.xxz_mpo(1.0, 0.9)
    psi, envs = _groundstate(W, 4)
    val, dW, _ = R.fixedpoint_gradient(psi, W, envs, W)
    assert abs(val - O.expectation_energy(psi, [W], envs)) < 1e-12
    assert np.abs(dW).max() < 1e-9


def test_response_of_sz_matches_tfim_closed_form():
    """d<sigma_z>/dh through the fixed point of the TFIM (J = 1, h = 1.3, gapped) against the
    closed form -e''(h) of the free-fermion energy e(h) = -(1/pi) int_0^pi sqrt(1 + h^2 - 2 h cos k)
    (e'(h) = -<sigma_z>): -e'' = (1/pi) int_0^pi sin^2 k / (1 + h^2 - 2 h cos k)^(3/2) dk.  A
    dropped curvature term or a wrong sign of an environment term misses it by far more than the
    truncation error at D = 16."""
    import math

    from scipy import integrate

    h, D = 1.3, 16
    W = O.tfim_mpo(1.0, h)
    psi, envs = _groundstate(W, D)
    val, dW, info = R.fixedpoint_gradient(psi, W, envs, R.objective_mpo_1site(O.sigma_z()), tol=1e-11)
    assert info["relres"] < 1e-7, info
    pred = float(np.sum(dW * (O.tfim_mpo(1.0, h + 1.0) - W)))
    exact = integrate.quad(lambda k: math.sin(k) ** 2 / (1 + h * h - 2 * h * math.cos(k)) ** 1.5, 0, math.pi,
                           epsabs=1e-14, epsrel=1e-14)[0] / math.pi
    assert abs(pred - exact) < 1e-9, (pred, exact)
    # without the curvature term the prediction is visibly off
    ts = R.TangentSpace(psi, W, envs)
    _, g = R.objective_gradient(psi, R.objective_mpo_1site(O.sigma_z()))
    Lh, _, _ = R.minres(lambda B: ts.project(ts.heff_apply(ts.project(B))) + B - ts.project(B), g, 1e-11, 500)
    pred_h = float(np.sum(-2.0 * ts.w_sandwich(Lh) * (O.tfim_mpo(1.0, h + 1.0) - W)))
    assert abs(pred_h - exact) > 1e3 * abs(pred - exact)
