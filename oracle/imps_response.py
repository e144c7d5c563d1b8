"""FP64 CPU restatement of the fixed-point response of a VUMPS ground state (numpy).

TEST INFRASTRUCTURE ONLY, like ``imps_oracle``: only ``tests/`` import it.  The product
package never imports it.

InfiniteMPSAlgorithms.jl has no derivative of its fixed point; this module restates the
library's tangent-space linear response with the reference's own operators:
H_AC (src/algorithms/groundstates/effective.jl:90-102), the slice pushes
(src/algorithms/groundstates/envs.jl:104-114), ``regularize!`` (envs.jl:28-32), the GMRES
linsolve of the environments (envs.jl:14-19) and DMRGCache (envs.jl:321-518).

One-site unit cell, real data.  Conventions as in ``imps_oracle``: ``A[a, s, b]``,
``W[wl, u, wr, d]`` (u = bra), ``GL[b, w, a]`` = (bra, w, ket), ``GR[bb, w, c]`` =
(ket, w, bra).

Tangent vectors are tensors ``B`` of the MPS shape with ``AL^T B = 0`` (B = VL X, left gauge);
``P = 1 - AL AL^T`` projects on them.  The state ``AL + t B C^-1`` (a uniform MPS) has energy
density ``e0 + t^2 <B, (H_eff + K) B> + O(t^3)`` at a variational fixed point, where

* ``H_eff B = P[(H_AC - lambda) B + H_AC[LK, GR](AR) + H_AC[GL, RK](AL)]`` is the zero-momentum
  excitation operator (``lambda = <AC, H_AC AC>``);
* ``K B = P[H_AC[GL, RKb](AC) + C^-T H_AC[LKb, GR](AC)]`` is the curvature of the manifold of
  uniform MPS (the two-defect terms of the second-order state, whose second defect is
  ``C^-1 B`` once written between AL on the left and AR on the right).

The one-defect environments are sums over the defect's distance, solved level by level like
DMRGCache; identity diagonal levels use the fixed points of the mixed transfer matrix for the
regularisation (``x -> x - T x + <x, r0> l0``):

    LK   ket B on the left,  then ket AR / bra AL   l0 = C,     r0 = C^T
    RK   ket B on the right, then ket AL / bra AR   l0 = C^T,   r0 = C
    LKb  bra B on the left,  then bra AR / ket AL   l0 = C^T,   r0 = C
    RKb  bra C^-1 B on the right, then AR / AR      l0 = C^T C, r0 = I

For an objective O (expectation value per site of a Schur MPO ``Wobj``) with tangent gradient
``g = P[H_AC^O(AC)]``, the response of the fixed point to the Hamiltonian tensor is

    dO/dW = -2 d/dW <Phi(B_L)| H[W] |psi>,   (H_eff + K) B_L = g,

the sandwich of the reverse H_AC contraction at the defect site plus the adjoint environments
(bra defect, ket psi) on either side of it.

C^-1 is the pseudo-inverse with relative cutoff ``rcond`` (directions of Schmidt values below
it carry no curvature term; their weight in every observable is below the cutoff).

Parity status (tests/test_oracle_response.py):
  hessian_apply          brute-force second derivative of the energy density of
                         AL + t B C^-1 (canonicalisation + DMRGCache at each t), symmetry,
                         positive spectrum at a minimum
  fixedpoint_gradient    central finite differences of converged VUMPS runs (d<sx>/dh for
                         TFIM, d<SzSz>/dDelta for XXZ); the energy as objective has zero
                         response (the Hellmann-Feynman derivative is all of dE/dW)
"""
from __future__ import annotations

import numpy as np

from . import imps_oracle as O


def push_left_bk(L, M, bra, ket):
    """Slice push (envs.jl:104-107) with distinct layers:
    L'[a',b'] = bra[a,u,a'] L[a,b] M[u,s] ket[b,s,b']."""
    Y = np.einsum("ab,bsc->asc", L, ket, optimize=True)
    Y = np.einsum("us,asc->auc", M, Y, optimize=True)
    return np.einsum("aup,auc->pc", bra, Y, optimize=True)


def push_right_bk(R, M, bra, ket):
    """Slice push (envs.jl:110-114) with distinct layers:
    R'[a',b'] = ket[a',s,a] R[a,b] M[u,s] bra[b',u,b]."""
    Y = np.einsum("psa,ab->psb", ket, R, optimize=True)
    Y = np.einsum("us,psb->pub", M, Y, optimize=True)
    return np.einsum("pub,qub->pq", Y, bra, optimize=True)


def _is_eye(M):
    return np.array_equal(M, np.eye(M.shape[0]))


def _level_solve(push, rhs, M, dual, fix, tol):
    """x = rhs + push(x, M).  For an identity M the transfer matrix has the fixed point ``fix``
    with dual ``dual`` (tr(dual fix) = 1): the component tr(dual rhs) fix is dropped and the
    solve runs on x - T x + tr(dual x) fix (regularize!, envs.jl:28-32, with the fixed points
    of a mixed transfer matrix)."""
    if not np.any(M):
        return rhs.copy()
    if _is_eye(M):
        rhs = rhs - np.trace(dual @ rhs) * fix

        def op(X):
            return push(X, M) - np.trace(dual @ X) * fix
    else:
        def op(X):
            return push(X, M)
    nb = np.linalg.norm(rhs)
    if nb == 0.0:
        return np.zeros_like(rhs)
    D = rhs.shape[0]
    x, _ = O.gmres_linsolve(op, rhs, np.zeros_like(rhs), tol * nb, min(D * D, O.Defaults.krylovdim),
                            O.Defaults.maxiter, a0=1.0, a1=-1.0)
    return x


def defect_env_left(G, W, sbra, sket, bra, ket, l0, r0, lo, hi, tol):
    """Levels lo..hi of sum_k [G pushed through (sbra, sket)] pushed k times through (bra, ket),
    all Schur channels (levels below lo are zero)."""
    D, w = G.shape[0], W.shape[0]
    E = np.zeros((D, w, D))
    for j in range(lo, hi + 1):
        rhs = np.zeros((D, D))
        for i in range(j + 1):
            M = W[i, :, j, :]
            if not np.any(M):
                continue
            rhs += push_left_bk(G[:, i, :], M, sbra, sket)
            if lo <= i < j:
                rhs += push_left_bk(E[:, i, :], M, bra, ket)
        # left vectors: fixed point l0, dual r0
        E[:, j, :] = _level_solve(lambda X, M: push_left_bk(X, M, bra, ket), rhs, W[j, :, j, :], r0, l0, tol)
    return E


def defect_env_right(G, W, sbra, sket, bra, ket, l0, r0, lo, hi, tol):
    """Levels hi..lo (descending) of the right-hand analogue of defect_env_left."""
    D, w = G.shape[0], W.shape[0]
    E = np.zeros((D, w, D))
    for i in range(hi, lo - 1, -1):
        rhs = np.zeros((D, D))
        for j in range(i, w):
            M = W[i, :, j, :]
            if not np.any(M):
                continue
            rhs += push_right_bk(G[:, j, :], M, sbra, sket)
            if i < j <= hi:
                rhs += push_right_bk(E[:, j, :], M, bra, ket)
        # right vectors: fixed point r0, dual l0
        E[:, i, :] = _level_solve(lambda X, M: push_right_bk(X, M, bra, ket), rhs, W[i, :, i, :], l0, r0, tol)
    return E


def tangent_project(AL, E):
    """P E = E - AL (AL^T E) (the Galerkin projector of vumps.jl:111-124)."""
    q = np.einsum("asc,asb->cb", AL, E, optimize=True)
    return E - np.einsum("asc,cb->asb", AL, q, optimize=True)


def pinv_cutoff(C, rcond):
    U, S, Vt = np.linalg.svd(C)
    inv = np.where(S > rcond * S[0], 1.0 / np.where(S > 0, S, 1.0), 0.0)
    return (Vt.T * inv) @ U.T


class TangentSpace:
    """The Hessian of the energy density on the tangent space of a converged one-site state."""

    def __init__(self, psi, W, envs, rcond=1e-10, tol=1e-13):
        if len(psi) != 1:
            raise ValueError("fixed-point response: one-site unit cell only")
        self.AL, self.AR, self.C, self.AC = psi.AL[0], psi.AR[0], psi.C[0], psi.AC[0]
        self.W = W
        self.GL, self.GR = envs.lefts[0], envs.rights[0]
        self.D = self.AL.shape[0]
        self.w = W.shape[0]
        self.tol = tol
        self.Cinv = pinv_cutoff(self.C, rcond)
        self.lam = float(np.sum(self.AC * O.hac_apply(self.GL, W, self.GR, self.AC)))

    def project(self, E):
        return tangent_project(self.AL, E)

    def heff_apply(self, B):
        AL, AR, C, W, GL, GR, w = self.AL, self.AR, self.C, self.W, self.GL, self.GR, self.w
        E = O.hac_apply(GL, W, GR, B) - self.lam * B
        LK = defect_env_left(GL, W, AL, B, AL, AR, C, C.T, 1, w - 1, self.tol)
        E += O.hac_apply(LK, W, GR, AR)
        RK = defect_env_right(GR, W, AR, B, AR, AL, C.T, C, 1, w - 1, self.tol)
        E += O.hac_apply(GL, W, RK, AL)
        return self.project(E)

    def curvature_apply(self, B):
        AR, AL, C, AC, W, GL, GR, w = self.AR, self.AL, self.C, self.AC, self.W, self.GL, self.GR, self.w
        CiB = np.einsum("ab,bsc->asc", self.Cinv, B, optimize=True)
        RKb = defect_env_right(GR, W, CiB, AR, AR, AR, C.T @ C, np.eye(self.D), 1, w - 1, self.tol)
        E = O.hac_apply(GL, W, RKb, AC)
        LKb = defect_env_left(GL, W, B, AL, AR, AL, C.T, C, 1, w - 1, self.tol)
        E += np.einsum("ba,bsc->asc", self.Cinv, O.hac_apply(LKb, W, GR, AC), optimize=True)
        return self.project(E)

    def hessian_apply(self, B):
        """P (H_eff + K) P B: projecting the input too keeps the operator symmetric on the whole
        space (the complement of the tangent space is its null space), so that rounding cannot
        leave the Krylov basis of the solver outside the tangent space."""
        B = self.project(B)
        return self.heff_apply(B) + self.curvature_apply(B)

    def system_apply(self, B):
        """The operator of the response solve: P (H_eff + K) P + (1 - P).  The complement of the
        tangent space is the null space of the Hessian; giving it the eigenvalue 1 (where the
        right-hand side has no component) keeps the solution in the tangent space and the operator
        nonsingular.  On the singular form MINRES's recurrence estimate detaches from the true
        residual once rounding components along the null space are picked up (a D = 8 TFIM case:
        estimate 1e-11, residual 5e-3 after 1000 iterations; 23 iterations to 2e-11 here)."""
        return self.hessian_apply(B) + (B - self.project(B))

    def w_sandwich(self, B):
        """d/dW <Phi(B)| H[W] |psi> per site: the defect site, then the adjoint environments
        (bra defect, ket psi) for the MPO tensors right of it (bra AR / ket AR between) and
        left of it (bra AL / ket AL between)."""
        AL, AR, C, AC, W, GL, GR, w = self.AL, self.AR, self.C, self.AC, self.W, self.GL, self.GR, self.w
        D = self.D

        def sand(L, bra, ket, R):
            return np.einsum("bia,buc,asx,xjc->iujs", L, bra, ket, R, optimize=True)

        S = sand(GL, B, AC, GR)
        LB = defect_env_left(GL, W, B, AC, AR, AR, C.T @ C, np.eye(D), 1, w - 2, self.tol)
        S += sand(LB, AR, AR, GR)
        RB = defect_env_right(GR, W, B, AC, AL, AL, np.eye(D), C @ C.T, 1, w - 1, self.tol)
        S += sand(GL, AL, AL, RB)
        return S


def objective_mpo_1site(Oloc):
    """Schur MPO (w = 2) whose energy density is the one-site expectation value <Oloc>."""
    d = Oloc.shape[0]
    Wo = np.zeros((2, d, 2, d))
    Wo[0, :, 0, :] = np.eye(d)
    Wo[1, :, 1, :] = np.eye(d)
    Wo[0, :, 1, :] = Oloc
    return Wo


def objective_gradient(psi, Wobj, tol=1e-14):
    """Value per site of the objective MPO and its tangent gradient P[H_AC^O(AC)]."""
    eo = O.dmrg_cache(psi, [Wobj], tol=tol)
    val = O.expectation_energy(psi, [Wobj], eo)
    g = tangent_project(psi.AL[0], O.hac_apply(eo.lefts[0], Wobj, eo.rights[0], psi.AC[0]))
    return val, g


def minres(op, b, tol, maxiter, atol=1e-14):
    """MINRES (Paige & Saunders 1975) for a symmetric operator, x0 = 0; stops when the
    recurrence's residual estimate is below max(tol |b|, atol) (the operator is only as exact
    as its inner linear solves, so a relative target on a tiny |b| is out of reach).  Returns
    (x, iterations, estimate relative to |b|)."""
    beta1 = float(np.linalg.norm(b))
    x = np.zeros_like(b)
    if beta1 <= atol:
        return x, 0, 0.0
    r1 = b.copy()
    r2 = b.copy()
    y = b.copy()
    w = np.zeros_like(b)
    w2 = np.zeros_like(b)
    oldb, beta, dbar, epsln, phibar, cs, sn = 0.0, beta1, 0.0, 0.0, beta1, -1.0, 0.0
    it = 0
    while it < maxiter:
        it += 1
        v = y / beta
        y = op(v)
        if it >= 2:
            y = y - (beta / oldb) * r1
        alfa = float(np.sum(v * y))
        y = y - (alfa / beta) * r2
        r1, r2 = r2, y
        oldb, beta = beta, float(np.linalg.norm(y))
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = max(np.hypot(gbar, beta), np.finfo(float).eps)
        cs, sn = gbar / gamma, beta / gamma
        phi = cs * phibar
        phibar = sn * phibar
        w1, w2 = w2, w
        w = (v - oldeps * w1 - delta * w2) / gamma
        x = x + phi * w
        if phibar <= max(tol * beta1, atol) or beta == 0.0:
            break
    return x, it, phibar / beta1


def fixedpoint_gradient(psi, W, envs, Wobj, tol=1e-10, maxiter=1000, rcond=1e-10):
    """(O, dO/dW, info) for the objective MPO Wobj at the converged one-site state psi of the
    Hamiltonian W with environments envs."""
    ts = TangentSpace(psi, W, envs, rcond=rcond, tol=min(1e-13, 1e-2 * tol))
    val, g = objective_gradient(psi, Wobj)
    Lam, its, est = minres(ts.system_apply, g, tol, maxiter)
    res = float(np.linalg.norm(ts.hessian_apply(Lam) - g) / max(np.linalg.norm(g), 1e-300))
    dW = -2.0 * ts.w_sandwich(Lam)
    return val, dW, {"iterations": its, "relres": res, "estimate": est}
