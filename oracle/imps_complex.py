"""Complex FP64 CPU restatement of the TDVP step and its environments (numpy, dense).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the library's complex
states; the product package never imports it.

The real restatement (imps_oracle.py) covers Float64 chains; this module restates, for
ComplexF64 chains and small bond dimensions, the pieces of one TDVP step and of the
canonical form, written directly in complex arithmetic (no real embedding), so that it checks
the library's embedded path independently:

* the reference's complex contractions: push_env_left / push_env_right
  (src/algorithms/transfermatrix.jl:5-26: conj on the bra layer) and the slice pushes
  (src/algorithms/groundstates/envs.jl:104-114);
* regularize! without conjugation (envs.jl:28-32) and DMRGCache's level recursion
  (envs.jl:321-518, square Schur tensors, bond dimensions that may differ from bond to bond),
  each linsolve done as a dense solve of
  (1 - T) x = b;
* H_AC / H_C (effective.jl:72-102), the TDVP integrator as the exact exponential
  scipy.linalg.expm(dt H) of the dense effective Hamiltonian (tdvp.jl:14-19: the Lanczos
  exponential approximates this to its tolerance);
* QRpos / LQpos with a real positive diagonal, regauge! (ortho.jl:214-219), uniform_leftorth /
  uniform_rightorth (ortho.jl:125-202) with the dominant eigenvector taken from the explicit
  transfer matrix, CanonicalIMPS(As) (canonicalmps.jl:151-170) and timestep (tdvp.jl:33-52).

Parity status: pinned in tests/test_oracle_complex.py against the real restatement on real
inputs (environments, energies and the imaginary-time step), against gauge covariance under
complex diagonal gauge transformations, against the exact precession of a product state
under a field (the TDVP flow is exact there), against energy and norm conservation of
real-time steps from a ground state (the reference's own test bounds,
test/algorithms/tdvp.jl:15-24) and against complex conjugation symmetry.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.linalg

from . import imps_oracle as ro


def fmat(A, m, n):
    return np.reshape(A, (m, n), order="F")


# ---------------------------------------------------------------- contractions (complex)


def push_left2(L, above, below):
    """transfermatrix.jl:5-7 with distinct layers: L'[a',b'] = conj(below[a,s,a']) L[a,b] above[b,s,b']."""
    Y = np.einsum("ab,bsc->asc", L, above)
    return np.einsum("asp,asc->pc", below.conj(), Y)


def push_right2(R, above, below):
    """transfermatrix.jl:23-26: R'[a',b'] = above[a',s,a] R[a,b] conj(below[b',s,b])."""
    Y = np.einsum("psa,ab->psb", above, R)
    return np.einsum("psb,qsb->pq", Y, below.conj())


def push_slice_left(L, Wl, A):
    """envs.jl:104-107: L'[a',b'] = conj(A[a,u,a']) L[a,b] Wl[u,d] A[b,d,b']."""
    Y = np.einsum("ab,bdc->adc", L, A)
    Y = np.einsum("ud,adc->auc", Wl, Y)
    return np.einsum("aup,auc->pc", A.conj(), Y)


def push_slice_right(R, Wl, A):
    """envs.jl:110-114: R'[a',b'] = A[a',d,a] R[a,b] Wl[u,d] conj(A[b',u,b])."""
    Y = np.einsum("pda,ab->pdb", A, R)
    Y = np.einsum("ud,pdb->pub", Wl, Y)
    return np.einsum("pub,qub->pq", Y, A.conj())


def regularize(v, lvec, rvec):
    """envs.jl:28-32: c = sum lvec[a,b] v[b,a] (no conjugation); v -= c rvec."""
    c = np.sum(lvec * v.T)
    v -= c * rvec
    return v


def dense_solve(op, rhs):
    """linsolve(op, b; a0 = 1, a1 = -1) (envs.jl:14-19) as a dense solve of (1 - op) x = b."""
    D1, D2 = rhs.shape
    n = D1 * D2
    M = np.zeros((n, n), dtype=complex)
    for k in range(n):
        e = np.zeros(n, dtype=complex)
        e[k] = 1.0
        M[:, k] = e - op(fmat(e, D1, D2)).reshape(-1, order="F")
    return fmat(np.linalg.solve(M, rhs.reshape(-1, order="F")), D1, D2)


# ---------------------------------------------------------------- states and gauge


@dataclass
class CanonicalIMPS:
    AL: list
    AR: list
    C: list
    AC: list

    def __len__(self):
        return len(self.AL)


def qrpos(M):
    """QRpos: thin QR with a real positive diagonal of R (Q R = M)."""
    Q, R = np.linalg.qr(M, mode="reduced")
    dg = np.diag(R)
    ph = np.where(np.abs(dg) > 0, dg / np.where(np.abs(dg) > 0, np.abs(dg), 1.0), 1.0)
    return Q * ph[None, :], ph.conj()[:, None] * R


def lqpos(M):
    """LQpos: L Q = M with L lower triangular with a real positive diagonal."""
    Q, R = qrpos(M.conj().T)
    return R.conj().T, Q.conj().T


def mul_AL_C(AL, C):
    return np.einsum("asb,bc->asc", AL, C)


def transfer_right_matrix(above, below):
    """The cell's right transfer map v -> push_right2 through every site (last site first), as
    an explicit (D^2 x D^2) matrix on column-major vectors."""
    D = above[0].shape[0]
    n = D * D
    T = np.zeros((n, n), dtype=complex)
    for k in range(n):
        e = np.zeros(n, dtype=complex)
        e[k] = 1.0
        R = fmat(e, D, D)
        for ab, bl in zip(reversed(above), reversed(below)):
            R = push_right2(R, ab, bl)
        T[:, k] = R.reshape(-1, order="F")
    return T


def transfer_left_matrix(above, below):
    D = above[0].shape[0]
    n = D * D
    T = np.zeros((n, n), dtype=complex)
    for k in range(n):
        e = np.zeros(n, dtype=complex)
        e[k] = 1.0
        L = fmat(e, D, D)
        for ab, bl in zip(above, below):
            L = push_left2(L, ab, bl)
        T[:, k] = L.reshape(-1, order="F")
    return T


def _dominant_vec(T, D):
    w, V = np.linalg.eig(T)
    k = int(np.argmax(np.abs(w)))
    return fmat(V[:, k], D, D)


def uniform_leftorth(A, C0, tol, maxiter, eig_miniter=10):
    """ortho.jl:125-164 (complex): QRpos sweeps C[i-1] A[i] -> AL[i] C[i]; after eig_miniter
    sweeps C[N] is refreshed from the dominant eigenvector of the mixed left transfer map."""
    N = len(A)
    D = C0.shape[0]
    C = [None] * N
    C[N - 1] = C0 / np.linalg.norm(C0)
    AL = [np.zeros_like(a, dtype=complex) for a in A]
    it = 0
    while True:
        if it >= eig_miniter:
            ev = _dominant_vec(transfer_left_matrix(A, AL), D)
            _, C[N - 1] = qrpos(ev)
        C_pre = C[N - 1].copy()
        for i in range(N):
            Dl, d, Dr = A[i].shape
            M = np.einsum("ab,bsc->asc", C[i - 1], A[i])
            Q, R = qrpos(fmat(M, Dl * d, Dr))
            AL[i] = np.reshape(Q, (Dl, d, Q.shape[1]), order="F")
            C[i] = R
        C[N - 1] = C[N - 1] / np.linalg.norm(C[N - 1])
        eps = float(np.linalg.norm(C[N - 1] - C_pre))
        it += 1
        if eps < tol or it > maxiter:
            return AL, C


def uniform_rightorth(A, C0, tol, maxiter, eig_miniter=10):
    """ortho.jl:166-202 (complex): LQpos sweeps A[i] C[i] -> C[i-1] AR[i]; after eig_miniter
    sweeps C[N] is refreshed from the dominant eigenvector of the mixed right transfer map."""
    N = len(A)
    D = C0.shape[0]
    C = [None] * N
    C[N - 1] = C0 / np.linalg.norm(C0)
    AR = [np.zeros_like(a, dtype=complex) for a in A]
    it = 0
    while True:
        if it >= eig_miniter:
            ev = _dominant_vec(transfer_right_matrix(A, AR), D)
            C[N - 1], _ = lqpos(ev)
        C_pre = C[N - 1].copy()
        for i in range(N - 1, -1, -1):
            Dl, d, Dr = A[i].shape
            M = np.einsum("asb,bc->asc", A[i], C[i])
            L, Q = lqpos(fmat(M, Dl, d * Dr))
            AR[i] = np.reshape(Q, (L.shape[1], d, Dr), order="F")
            C[i - 1] = L
        C[N - 1] = C[N - 1] / np.linalg.norm(C[N - 1])
        eps = float(np.linalg.norm(C[N - 1] - C_pre))
        it += 1
        if eps < tol or it > maxiter:
            return AR, C


def canonical_from_tensors(As, tol=1e-14, maxiter=2000):
    """CanonicalIMPS(As) (canonicalmps.jl:151-170): gaugefix!(:LR) from C0 = I."""
    N = len(As)
    D = As[0].shape[0]
    AL, C = uniform_leftorth([np.asarray(a, dtype=complex) for a in As], np.eye(D, dtype=complex), tol, maxiter)
    AR, C = uniform_rightorth(AL, C[N - 1], tol, maxiter)
    return CanonicalIMPS(AL, AR, C, [mul_AL_C(AL[i], C[i]) for i in range(N)])


def regauge(AC, C):
    """ortho.jl:214-219: AL = Q(QRpos(AC)) Q(QRpos(C))^H."""
    Dl, d, Dr = AC.shape
    Q_AC, _ = qrpos(fmat(AC, Dl * d, Dr))
    Q_C, _ = qrpos(C.copy())
    return np.reshape(Q_AC @ Q_C.conj().T, (Dl, d, Dr), order="F")


# ---------------------------------------------------------------- environments


@dataclass
class Envs:
    lefts: list
    rights: list

    def leftenv(self, l):
        return self.lefts[l % len(self.lefts)]

    def rightenv(self, l):
        return self.rights[l % len(self.rights)]


def _left_cyclethrough(lefts, Ws, ALs, i):
    N = len(ALs)
    for site in range(N):
        tgt = np.zeros((ALs[site].shape[2],) * 2, dtype=complex)
        for l in range(i + 1):
            tgt += push_slice_left(lefts[site][:, l, :], Ws[site][l, :, i, :], ALs[site])
        lefts[(site + 1) % N][:, i, :] = tgt


def _right_cyclethrough(rights, Ws, ARs, i):
    N = len(ARs)
    for site in range(N - 1, -1, -1):
        ns = Ws[site].shape[2]
        tgt = np.zeros((ARs[site].shape[0],) * 2, dtype=complex)
        for l in range(i, ns):
            tgt += push_slice_right(rights[site][:, l, :], Ws[site][i, :, l, :], ARs[site])
        rights[(site - 1) % N][:, i, :] = tgt


def dmrg_cache(psi, Ws):
    """DMRGCache(ψ, H::SparseIMPO) (envs.jl:321-518) for square Schur tensors, complex ψ:
    levels in Schur order, identity levels by the regularised solve (ρ = C C^H on the left,
    C^H C on the right), other diagonal levels by (1 - T_W) x = b, each solve dense."""
    N = len(psi)
    w = Ws[0].shape[0]
    # bond dimensions may differ from bond to bond: leftenv(l) lives on the bond left of site l,
    # rightenv(l) on the bond right of it (envs.jl:327-343)
    Dl = [psi.AL[l].shape[0] for l in range(N)]
    Dr = [psi.AR[l].shape[2] for l in range(N)]
    lefts = [np.zeros((Dl[l], w, Dl[l]), dtype=complex) for l in range(N)]
    rights = [np.zeros((Dr[l], w, Dr[l]), dtype=complex) for l in range(N)]
    IL = [np.eye(Dl[l], dtype=complex) for l in range(N)]
    IR = [np.eye(Dr[l], dtype=complex) for l in range(N)]
    for l in range(N):
        lefts[l][:, 0, :] = IL[l]
        rights[l][:, w - 1, :] = IR[l]
    AL, AR, C = psi.AL, psi.AR, psi.C
    for i in range(1, w):
        _left_cyclethrough(lefts, Ws, AL, i)
        rhs = lefts[0][:, i, :].copy()
        if ro.isidentitylevel(Ws, i):
            rho = C[N - 1] @ C[N - 1].conj().T

            def op(X, rho=rho):
                for l in range(N):
                    X = push_left2(X, AL[l], AL[l])
                return regularize(X, rho, IL[0])

            lefts[0][:, i, :] = dense_solve(op, rhs)
            if N > 1:
                _left_cyclethrough(lefts, Ws, AL, i)
            for l in range(N):
                regularize(lefts[l][:, i, :], C[l - 1] @ C[l - 1].conj().T, IL[l])
        else:
            if not ro.isemptylevel(Ws, i):
                def op(X, i=i):
                    for l in range(N):
                        X = push_slice_left(X, Ws[l][i, :, i, :], AL[l])
                    return X

                lefts[0][:, i, :] = dense_solve(op, rhs)
            if N > 1:
                _left_cyclethrough(lefts, Ws, AL, i)
    for i in range(w - 2, -1, -1):
        _right_cyclethrough(rights, Ws, AR, i)
        rhs = rights[N - 1][:, i, :].copy()
        if ro.isidentitylevel(Ws, i):
            rho = C[N - 1].conj().T @ C[N - 1]

            def op(X, rho=rho):
                for l in range(N - 1, -1, -1):
                    X = push_right2(X, AR[l], AR[l])
                return regularize(X, rho, IR[N - 1])

            rights[N - 1][:, i, :] = dense_solve(op, rhs)
            if N > 1:
                _right_cyclethrough(rights, Ws, AR, i)
            for l in range(N):
                regularize(rights[l][:, i, :], C[l].conj().T @ C[l], IR[l])
        else:
            if not ro.isemptylevel(Ws, i):
                def op(X, i=i):
                    for l in range(N - 1, -1, -1):
                        X = push_slice_right(X, Ws[l][i, :, i, :], AR[l])
                    return X

                rights[N - 1][:, i, :] = dense_solve(op, rhs)
            if N > 1:
                _right_cyclethrough(rights, Ws, AR, i)
    return Envs(lefts, rights)


# ---------------------------------------------------------------- effective operators, observables


def hac_apply(GL, W, GR, x):
    """effective.jl:79-102: y[b,u,c] = GL[b,i,a] x[a,s,bb] W[i,u,j,s] GR[bb,j,c]."""
    t1 = np.einsum("bia,asc->bisc", GL, x)
    t2 = np.einsum("iujs,bisc->bujc", W, t1)
    return np.einsum("bujc,cjz->buz", t2, GR)


def hc_apply(GL, GR, x):
    """effective.jl:72-76: y[a,b] = GL[a,w,a'] x[a',b'] GR[b',w,b]."""
    return np.einsum("awq,qwb->ab", np.einsum("awp,pq->awq", GL, x), GR)


def dense_operator(op, shape):
    n = int(np.prod(shape))
    M = np.zeros((n, n), dtype=complex)
    for k in range(n):
        e = np.zeros(n, dtype=complex)
        e[k] = 1.0
        M[:, k] = op(np.reshape(e, shape, order="F")).reshape(-1, order="F")
    return M


def expm_apply(op, dt, x):
    """exp(dt op) x with the dense operator (the exact value exponentiate approximates)."""
    M = dense_operator(op, x.shape)
    return np.reshape(scipy.linalg.expm(dt * M) @ x.reshape(-1, order="F"), x.shape, order="F")


def dot_psi(psi):
    """dot(ψ, ψ) (canonicalmps.jl:262-269): dominant eigenvalue of the AL transfer map."""
    w = np.linalg.eigvals(transfer_left_matrix(psi.AL, psi.AL))
    return w[int(np.argmax(np.abs(w)))]


def expectation_energy(psi, Ws, envs):
    """expectationvalue(ψ, H::SparseIMPO, envs) (expval.jl:93-112), the cell total:
    sum_l <AC_l| GL_l W_l[:, :, end] GR_l[:, end] |AC_l> / dot(ψ, ψ)."""
    E = 0.0
    for l in range(len(psi)):
        W = Ws[l]
        ns = W.shape[2] - 1
        y = hac_apply(envs.leftenv(l), W[:, :, ns:ns + 1, :], envs.rightenv(l)[:, ns:ns + 1, :], psi.AC[l])
        E += np.vdot(psi.AC[l], y)
    return E / dot_psi(psi)


def local_expectation1(psi, site, O):
    """expval.jl:33-37: <AC|O|AC> / dot(ψ, ψ)."""
    AC = psi.AC[site]
    return np.einsum("aub,us,asb->", AC.conj(), O, AC) / dot_psi(psi)


def correlator(psi, O1, O2, i, js):
    """correlators.jl:13-42 (complex): <psi| O1[i] O2[j] |psi> / dot(psi, psi) for increasing j > i;
    V[b, c] = conj(AC[i][a,u,b]) O1[u,s] AC[i][a,s,c], pushed through the AR of the sites between,
    closed with conj(AR[j][b,u,e]) O2[u,s] AR[j][c,s,e]."""
    N = len(psi)
    AC = psi.AC[i % N]
    V = np.einsum("aub,us,asc->bc", AC.conj(), O1, AC)
    ctr, out = i, []
    for j in js:
        for l in range(ctr + 1, j):
            A = psi.AR[l % N]
            V = np.einsum("bc,bud,cue->de", V, A.conj(), A)
        A = psi.AR[j % N]
        out.append(np.einsum("bc,bue,us,cse->", V, A.conj(), O2, A))
        V = np.einsum("bc,bud,cue->de", V, A.conj(), A)
        ctr = j
    return np.array(out) / dot_psi(psi)


def tdvp_timestep(psi, Ws, envs, dt, gauge_tol=1e-14, gauge_maxiter=2000):
    """timestep(ψ, H, dt, alg::TDVP, envs) (tdvp.jl:33-52) for complex dt: AC -> exp(dt H_AC) AC,
    C -> exp(dt H_C) C per site (H_C at site l between leftenv(l+1) and rightenv(l)), AL from
    regauge!, CanonicalIMPS(ALs, C[end]) (right gauge fix, AL kept), new environments."""
    N = len(psi)
    ALs = []
    for loc in range(N):
        GL, GR = envs.leftenv(loc), envs.rightenv(loc)
        tAC = expm_apply(lambda x: hac_apply(GL, Ws[loc], GR, x), dt, psi.AC[loc])
        GLc = envs.leftenv(loc + 1)
        tC = expm_apply(lambda x: hc_apply(GLc, GR, x), dt, psi.C[loc])
        ALs.append(regauge(tAC, tC))
    AR, C = uniform_rightorth(ALs, psi.C[N - 1].copy(), gauge_tol, gauge_maxiter)
    psi2 = CanonicalIMPS(ALs, AR, C, [mul_AL_C(ALs[l], C[l]) for l in range(N)])
    return psi2, dmrg_cache(psi2, Ws)


def from_real(psi):
    """A real canonical state (imps_oracle.CanonicalIMPS) as complex arrays."""
    return CanonicalIMPS(*[[np.asarray(x, dtype=complex) for x in xs] for xs in (psi.AL, psi.AR, psi.C, psi.AC)])


def phase_gauge(psi, phases):
    """The same state in another gauge: every bond b is rotated by the diagonal unitary
    U_b = diag(exp(i phases[b])), AL/AR/AC[l] -> U_{l-1}^H A U_l and C[l] -> U_l^H C U_l (bond l
    is right of site l).  Returns complex arrays."""
    N = len(psi)
    U = [np.exp(1j * np.asarray(p)) for p in phases]

    def rot(A, l):
        return U[(l - 1) % N].conj()[:, None, None] * A * U[l][None, None, :]

    return CanonicalIMPS([rot(psi.AL[l], l) for l in range(N)], [rot(psi.AR[l], l) for l in range(N)],
                         [U[l].conj()[:, None] * psi.C[l] * U[l][None, :] for l in range(N)],
                         [rot(psi.AC[l], l) for l in range(N)])


def larmor(h, t):
    """(<σx>, <σy>)(t) for the product state |+x> under H = -h Σ σz, evolved with exp(-i H t):
    the spin precesses about z, <σx> = cos(2 h t), <σy> = -sin(2 h t)."""
    return math.cos(2.0 * h * t), -math.sin(2.0 * h * t)
