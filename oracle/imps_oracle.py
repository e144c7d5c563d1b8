"""FP64 CPU restatement of InfiniteMPSAlgorithms.jl's VUMPS inner loop (numpy).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP library: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import it.  The product package ``infinitempsalgorithms_amd`` never imports it and has
no CPU fallback through it.

Every function cites the reference file:line it restates (paths relative to the
reference repository root).  Index conventions are the reference's column-major ones,
written as numpy index order (``reshape(..., order="F")`` wherever the reference
reshapes):

* MPS tensor ``A[a, s, b]``  shape ``(Dl, d, Dr)``;
* dense MPO tensor ``W[wl, u, wr, d]`` (``u`` = output/bra, ``d`` = input/ket),
  the ``tompotensor`` layout of a SchurMPOTensor (src/operators/mpohamiltonian.jl:5-13);
* environments ``GL[b, w, a]`` = (bra, w, ket), ``GR[bb, w, c]`` = (ket, w, bra)
  (src/algorithms/groundstates/effective.jl:3-5).

Scope: real Float64 data, square Schur tensors with the same level count on every site
and a uniform bond dimension (the five benchmark configurations).  Rectangular Schur chains
and non-uniform bond profiles are outside this restatement; complex states (real-time TDVP)
are restated densely in imps_complex.py.

KrylovKit (un-vendored dependency of the reference, unpinned in Project.toml:8) is
restated from its published algorithms, parameterised like KrylovKit:

* Lanczos eigsolve: thick restart (keep = div(3*krylovdim + 2*converged, 5)) and the
  eager convergence test |beta * u_K| <= tol after every expansion.  Orthogonalisation
  is classical Gram-Schmidt applied twice (CGS2) against the whole basis;
* GMRES linsolve: restarted GMRES(krylovdim) with Givens rotations; the residual is
  recomputed explicitly at every restart;
* Arnoldi (schursolve / non-Hermitian eigsolve, ``which = :LM``, howmany = 1): eager
  test as above; at krylovdim the factorisation restarts from the dominant Ritz vector.

The returned Lanczos eigenvector has its sign fixed so that <x0, v> >= 0 (KrylovKit
leaves the sign to LAPACK's eigenvector convention).  Gauge-invariant results (energy,
Galerkin error, spectra) do not depend on that sign.

Parity status of each part (how the restatement is pinned, see tests/test_oracle.py):
  models, tompotensor             exact diagonalisation of open chains; free-fermion
                                  closed form (U = 0 Hubbard)
  hac_apply / hc_apply            dense (dD^2)x(dD^2) brute force, Schur-block loop vs
                                  dense einsum, identity MPO => H(x) = x, Hermiticity
  push_env_* / TransferMatrix     sequential pushes, AL transfer eigenvalue 1
  regularize                      traceless projection (api.jl:242-246)
  lanczos / gmres / arnoldi       numpy eigh / solve / eig on random operators
  qrpos / lqpos                   positive diagonal, orthonormality, reconstruction
  gauge fixing                    mixed-canonical errors -> 0, AL.C == C.AR
  DMRGCache / energy              two-site RDM energy (api.jl:510-527), product state
                                  energy -1 (mpo.jl:83-85), recalculate! invariance
  find_groundstate (VUMPS)        TFIM -4/pi, XXX 1/4 - ln 2 closed forms
  push_env_*_mpo                  slice sums and brute-force contractions
  IDMRG                           XXX energy against VUMPS and 1/4 - ln 2
  lanczos_exponentiate            dense scipy expm (both signs of t, with sub-steps)
  TDVP (real dt)                  imaginary-time XXX run to 1/4 - ln 2 (the reference's test)
  correlator / dot / fidelity     product states, a two-site expectation value as the
                                  neighbouring correlator, factorisation at long distance,
                                  self-fidelity 1
  dominant_env                    identity channel: lambda = 1, L = R = I/sqrt(D); product
                                  state closed form; explicit transfer-matrix eigenpair
  dense_envs / expectation_dense  one-site cell value = dominant transfer eigenvalue; doubled
                                  cell; linearity in W; identity MPO = N
  KrylovKit iteration trajectories: parity unpinned (no Julia in this environment).
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------
# Defaults and dynamic tolerances (src/utility.jl:297-365, 136-171)
# ---------------------------------------------------------------------------


class Defaults:
    """src/utility.jl:308-320."""

    D = 64
    maxiter = 300
    tolgauge = 1.0e-13
    tol = 1.0e-10
    verbosity = 0
    krylovdim = 30
    tol_min = 1.0e-14
    tol_max = 1.0e-4
    eigs_tolfactor = 1.0e-3
    gauge_tolfactor = 1.0e-6
    envs_tolfactor = 1.0e-4


def updatetol(tol_factor, it, eps, tol_min=Defaults.tol_min, tol_max=Defaults.tol_max):
    """DynamicTol update: clamp(eps * factor / sqrt(max(it, 1)), tol_min, tol_max).

    src/utility.jl:153-157.
    """
    it = max(int(it), 1)
    return min(max(eps * tol_factor / math.sqrt(it), tol_min), tol_max)


def fmat(A, m, n):
    """Column-major reshape (Julia ``reshape(A, m, n)``)."""
    return np.reshape(A, (m, n), order="F")


# ---------------------------------------------------------------------------
# Models (src/models.jl:82-122) and the Schur MPO layout
# ---------------------------------------------------------------------------


def sigma_x():
    return np.array([[0.0, 1.0], [1.0, 0.0]])


def sigma_z():
    return np.array([[1.0, 0.0], [0.0, -1.0]])


def mpohamiltonian(h1, pairs):
    """Schur level matrix of an on-site + nearest-neighbour Hamiltonian as a dense
    ``(w, d, w, d)`` tensor (src/models.jl:82-99 followed by FiniteMPSAlgorithms'
    ``tompotensor``): level 1 and level n are the unit levels, ``[1, n] = h1``,
    channel k: ``[1, k+1] = coeff_k * a_k`` and ``[k+1, n] = b_k``.  Missing blocks
    (including the interior diagonal) are zero.
    """
    h1 = np.asarray(h1, dtype=np.float64)
    d = h1.shape[0]
    n = len(pairs) + 2
    W = np.zeros((n, d, n, d))
    eye = np.eye(d)
    W[0, :, 0, :] = eye
    W[n - 1, :, n - 1, :] = eye
    W[0, :, n - 1, :] = h1
    for k, (coeff, a, b) in enumerate(pairs):
        W[0, :, k + 1, :] = coeff * np.asarray(a, dtype=np.float64)
        W[k + 1, :, n - 1, :] = np.asarray(b, dtype=np.float64)
    return W


def tfim_mpo(J=1.0, h=1.0):
    """H = -J sum sx sx - h sum sz (src/models.jl:119-122), w = 3."""
    return mpohamiltonian(-h * sigma_z(), [(-J, sigma_x(), sigma_x())])


def xxz_mpo(J=1.0, Delta=1.0, hz=0.0):
    """Real S+- form of H = J sum (SxSx + SySy + Delta SzSz) - hz sum Sz, w = 5.

    The reference's heisenberg_hamiltonian (src/models.jl:107-111) uses the complex Sy;
    SxSx + SySy = (S+S- + S-S+)/2 gives the same operator with real blocks.
    """
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    sm = sp.T.copy()
    sz = 0.5 * sigma_z()
    return mpohamiltonian(-hz * sz, [(J / 2, sp, sm), (J / 2, sm, sp), (J * Delta, sz, sz)])


def hubbard_local_ops():
    """Local fermion operators on (|0>, |up>, |dn>, |updn> = c+_up c+_dn |0>).

    Returns (c_up, c_dn, F, n_up, n_dn) with F = (-1)^n the Jordan-Wigner parity.
    """
    c_up = np.zeros((4, 4))
    c_dn = np.zeros((4, 4))
    c_up[0, 1] = 1.0   # c_up |up> = |0>
    c_up[2, 3] = 1.0   # c_up c+_up c+_dn |0> = c+_dn |0> = |dn>
    c_dn[0, 2] = 1.0   # c_dn |dn> = |0>
    c_dn[1, 3] = -1.0  # c_dn c+_up c+_dn |0> = -c+_up |0> = -|up>
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    n_up = c_up.T @ c_up
    n_dn = c_dn.T @ c_dn
    return c_up, c_dn, F, n_up, n_dn


def hubbard_mpo(t=1.0, U=4.0, mu=2.0):
    """Hermitian Jordan-Wigner Fermi-Hubbard MPO, w = 6.

    H = -t sum_sigma (c+_{i,s} c_{i+1,s} + h.c.) + U sum n_up n_dn - mu sum n.
    The reference's fermi_hubbard (src/models.jl:164-178) keeps only two hopping
    channels and no string operator; this restatement uses the four Hermitian channels
    c+_s F (x) c_s and F c_s (x) c+_s.  mu = U/2 is half filling.
    """
    c_up, c_dn, F, n_up, n_dn = hubbard_local_ops()
    h1 = U * (n_up @ n_dn) - mu * (n_up + n_dn)
    pairs = [
        (-t, c_up.T @ F, c_up),
        (-t, c_dn.T @ F, c_dn),
        (-t, F @ c_up, c_up.T),
        (-t, F @ c_dn, c_dn.T),
    ]
    return mpohamiltonian(h1, pairs)


def isidentitylevel(Ws, i, atol=1e-14):
    """0-based level i; src/operators/mpohamiltonian.jl:100-109 (square tensors)."""
    if i == 0:
        return True
    for W in Ws:
        w = W.shape[0]
        if i == w - 1:
            continue
        blk = W[i, :, i, :]
        if not np.allclose(blk, np.eye(blk.shape[0]), atol=atol, rtol=0):
            return False
    return True


def isemptylevel(Ws, i):
    """0-based level i; src/operators/mpohamiltonian.jl:122-133 (square tensors)."""
    if i == 0:
        return False
    if any(i == W.shape[2] - 1 for W in Ws):
        return False
    for W in Ws:
        dz = not np.any(W[i, :, i, :])
        cz = not np.any(W[0, :, i, :])
        bz = not np.any(W[i, :, W.shape[2] - 1, :])
        if not (dz and cz and bz):
            return False
    return True


def nonzero_blocks(W):
    w = W.shape[0]
    return [(i, j) for i in range(w) for j in range(w) if np.any(W[i, :, j, :])]


# ---------------------------------------------------------------------------
# Effective Hamiltonians (src/algorithms/groundstates/effective.jl)
# ---------------------------------------------------------------------------


def hac_apply(GL, W, GR, x):
    """Schur block loop of H_AC: effective.jl:90-102.

    y[b,u,c] = sum_{W[i,j] != 0} GL[b,i,a] x[a,s,bb] W[i,u,j,s] GR[bb,j,c]
    """
    y = np.zeros((GL.shape[0], x.shape[1], GR.shape[2]))
    for i, j in nonzero_blocks(W):
        Wij = W[i, :, j, :]
        t = np.einsum("ba,asc->bsc", GL[:, i, :], x, optimize=True)
        t = np.einsum("us,bsc->buc", Wij, t, optimize=True)
        y += np.einsum("buc,cz->buz", t, GR[:, j, :], optimize=True)
    return y


def hac_apply_dense(GL, W, GR, x):
    """Dense twin of H_AC: effective.jl:79-85."""
    t1 = np.einsum("bia,asc->bisc", GL, x, optimize=True)
    t2 = np.einsum("iujs,bisc->bujc", W, t1, optimize=True)
    return np.einsum("bujc,cjz->buz", t2, GR, optimize=True)


def hc_apply(GL, GR, x):
    """H_C: effective.jl:72-76, y[a,b] = GL[a,w,a'] x[a',b'] GR[b',w,b]."""
    t = np.einsum("awp,pq->awq", GL, x, optimize=True)
    return np.einsum("awq,qwb->ab", t, GR, optimize=True)


# ---------------------------------------------------------------------------
# Transfer-matrix pushes (src/transfermatrix.jl, src/algorithms/groundstates/envs.jl)
# ---------------------------------------------------------------------------


def push_env_left(L, A):
    """transfermatrix.jl:5-7 (real): L'[a',b'] = A[a,s,a'] L[a,b] A[b,s,b']."""
    Y = np.einsum("ab,bsc->asc", L, A, optimize=True)
    return np.einsum("asp,asc->pc", A, Y, optimize=True)


def push_env_right(R, A):
    """transfermatrix.jl:23-26 (real): R'[a',b'] = A[a',s,a] R[a,b] A[b',s,b]."""
    Y = np.einsum("psa,ab->psb", A, R, optimize=True)
    return np.einsum("psb,qsb->pq", Y, A, optimize=True)


def push_env_left2(L, above, below):
    """Double-layer left push, transfermatrix.jl:54-58:
    L'[a',b'] = below[a,s,a'] L[a,b] above[b,s,b']."""
    Y = np.einsum("ab,bsc->asc", L, above, optimize=True)
    return np.einsum("asp,asc->pc", below, Y, optimize=True)


def push_env_right2(R, above, below):
    """Double-layer right push, transfermatrix.jl:62-66:
    R'[a',b'] = above[a',s,a] R[a,b] below[b',s,b]."""
    Y = np.einsum("ab,qsb->asq", R, below, optimize=True)
    return np.einsum("psa,asq->pq", above, Y, optimize=True)


def push_slice_left(L, Wl, A):
    """envs.jl:104-107: L'[a',b'] = A[a,u,a'] L[a,b] Wl[u,d] A[b,d,b']."""
    Y = np.einsum("ab,bdc->adc", L, A, optimize=True)
    Y = np.einsum("ud,adc->auc", Wl, Y, optimize=True)
    return np.einsum("aup,auc->pc", A, Y, optimize=True)


def push_slice_right(R, Wl, A):
    """envs.jl:110-114: R'[a',b'] = A[a',d,a] R[a,b] Wl[u,d] A[b',u,b]."""
    Y = np.einsum("pda,ab->pdb", A, R, optimize=True)
    Y = np.einsum("ud,pdb->pub", Wl, Y, optimize=True)
    return np.einsum("pub,qub->pq", Y, A, optimize=True)


def transfer_left(above, below, v):
    """TransferMatrix(above, below; side=:left) * v, transfermatrix.jl:145-153."""
    Db, Da = below[0].shape[0], above[0].shape[0]
    L = fmat(v, Db, Da)
    for ab, bl in zip(above, below):
        L = push_env_left2(L, ab, bl)
    return L.reshape(-1, order="F")


def transfer_right(above, below, v):
    """TransferMatrix(above, below; side=:right) * v, transfermatrix.jl:155-162."""
    Da, Db = above[0].shape[0], below[0].shape[0]
    R = fmat(v, Da, Db)
    for ab, bl in zip(reversed(above), reversed(below)):
        R = push_env_right2(R, ab, bl)
    return R.reshape(-1, order="F")


def regularize(v, lvec, rvec):
    """envs.jl:28-32: c = sum lvec[a,b] v[b,a]; v -= c * rvec (in place)."""
    c = float(np.sum(lvec * v.T))
    v -= c * rvec
    return v


# ---------------------------------------------------------------------------
# Orthogonal factorisations with a positive diagonal (FiniteMPSAlgorithms QRpos/LQpos)
# ---------------------------------------------------------------------------


def qrpos(M):
    """Thin QR with R having a non-negative diagonal (``leftorth(M; alg = QRpos())``)."""
    Q, R = np.linalg.qr(M, mode="reduced")
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return Q * s[None, :], R * s[:, None]


def lqpos(M):
    """Thin LQ with L having a non-negative diagonal (``rightorth(M; alg = LQpos())``)."""
    Q, R = qrpos(M.T)
    return R.T, Q.T


# ---------------------------------------------------------------------------
# Krylov solvers (KrylovKit restatements; call sites utility.jl:180-265, envs.jl:14-19)
# ---------------------------------------------------------------------------


def _cgs2(w, V):
    """Classical Gram-Schmidt, two passes, against the columns of V.
    Returns (w_perp, coefficients)."""
    if V.shape[1] == 0:
        return w, np.zeros(0)
    h = V.T @ w
    w = w - V @ h
    h2 = V.T @ w
    w = w - V @ h2
    return w, h + h2


@dataclass
class KrylovInfo:
    converged: int
    normres: float
    numiter: int
    numops: int


def lanczos_eigsolve(op, x0, tol, krylovdim=Defaults.krylovdim, maxiter=Defaults.maxiter,
                     eager=True, which="SR"):
    """Lowest (which='SR') or highest ('LR') eigenpair of a symmetric operator.

    Restates KrylovKit.eigsolve(f, x0, 1, which, Lanczos(...)) as called from
    ``fixedpoint(op, x0, which, ::Lanczos)`` (src/utility.jl:180-186, 379-393).
    """
    shape = x0.shape
    n = x0.size
    kmax = min(krylovdim, n)
    V = np.zeros((n, kmax + 1))
    T = np.zeros((kmax + 1, kmax + 1))
    v = x0.reshape(-1, order="F").astype(np.float64)
    v = v / np.linalg.norm(v)
    V[:, 0] = v
    K = 0          # number of basis vectors with an applied operator
    numops = 0
    numiter = 1
    beta = 0.0

    def expand(k):
        nonlocal beta, numops
        w = op(V[:, k].reshape(shape, order="F")).reshape(-1, order="F")
        numops += 1
        w, h = _cgs2(w, V[:, : k + 1])
        beta = float(np.linalg.norm(w))
        T[: k + 1, k] = h
        T[k, : k + 1] = h
        if beta > 0:
            V[:, k + 1] = w / beta

    expand(0)
    K = 1
    sign = 1.0 if which == "SR" else -1.0
    while True:
        converged = 0
        if K == kmax or beta <= tol or (eager and K >= 1):
            Tk = 0.5 * (T[:K, :K] + T[:K, :K].T)
            theta, U = np.linalg.eigh(Tk)
            p = np.argsort(sign * theta, kind="stable")
            theta, U = theta[p], U[:, p]
            f = beta * U[K - 1, :]
            if abs(f[0]) <= tol:
                converged = 1
                break
        if K < kmax:
            expand(K)
            K += 1
        else:
            if numiter == maxiter:
                break
            keep = (3 * kmax + 2 * converged) // 5
            Vk = V[:, :K] @ U[:, :keep]
            vnext = V[:, K].copy()
            T[:] = 0.0
            T[:keep, :keep] = np.diag(theta[:keep])
            T[keep, :keep] = f[:keep]
            T[:keep, keep] = f[:keep]
            V[:, :keep] = Vk
            V[:, keep] = vnext
            V[:, keep + 1:] = 0.0
            K = keep
            # The coupling column of the retained residual vector is filled by the
            # next expansion (orthogonalisation against the Ritz vectors reproduces f).
            expand(K)
            K += 1
            numiter += 1
    theta0 = float(theta[0])
    x = V[:, :K] @ U[:, 0]
    if float(x @ x0.reshape(-1, order="F")) < 0:
        x = -x
    info = KrylovInfo(converged, float(abs(beta * U[K - 1, 0])), numiter, numops)
    if not converged:
        warnings.warn(f"lanczos_eigsolve not converged: normres={info.normres:.3e} tol={tol:.3e}")
    return theta0, x.reshape(shape, order="F"), info


def _dominant_ritz(H):
    """Dominant (largest modulus) eigenpair of a small real matrix.  Returns (lam, u, last): lam
    the eigenvalue (complex for a complex dominant pair, else float), u the real part of the
    phase-fixed unit eigenvector x, normalised (a vector of the pair's real invariant plane when
    the pair is complex), and last = |x[-1]|, so that the Arnoldi residual is |beta| * last."""
    lam, U = np.linalg.eig(H)
    i = int(np.argmax(np.abs(lam)))
    x = U[:, i] / np.linalg.norm(U[:, i])
    last = float(abs(x[-1]))
    # rotate the phase so the largest entry is real, then take the real part
    k = int(np.argmax(np.abs(x)))
    x = x * (abs(x[k]) / x[k])
    u = np.real(x)
    u = u / np.linalg.norm(u)
    li = lam[i]
    if abs(li.imag) <= 1e-12 * max(1.0, abs(li)):
        return float(np.real(li)), u, last
    return complex(li), u, last


def arnoldi_dominant(op, x0, tol, krylovdim=Defaults.krylovdim, maxiter=Defaults.maxiter,
                     eager=True):
    """Dominant (which = :LM) eigenpair of a real non-symmetric operator.

    Restates KrylovKit.schursolve(op, x0, 1, :LM, Arnoldi(...)) as called from
    ``gauge_fixedpoint`` (src/utility.jl:204-211) and KrylovKit.eigsolve(..., :LM;
    ishermitian=false) as called from ``dot(psi1, psi2)`` (src/states/canonicalmps.jl:262-269).
    For howmany = 1 with a real dominant eigenvalue the first Schur vector is the
    normalised eigenvector.  A complex dominant pair is returned as a complex eigenvalue (as
    KrylovKit's eigsolve returns it) with the real part of the Ritz vector; restarts from that
    vector span the pair's invariant plane.
    """
    shape = x0.shape
    n = x0.size
    kmax = min(krylovdim, n)
    v = x0.reshape(-1, order="F").astype(np.float64)
    numiter = 1
    numops = 0
    while True:
        V = np.zeros((n, kmax + 1))
        H = np.zeros((kmax + 1, kmax))
        V[:, 0] = v / np.linalg.norm(v)
        K = 0
        conv = False
        while True:
            w = op(V[:, K].reshape(shape, order="F")).reshape(-1, order="F")
            numops += 1
            w, h = _cgs2(w, V[:, : K + 1])
            beta = float(np.linalg.norm(w))
            H[: K + 1, K] = h
            H[K + 1, K] = beta
            K += 1
            if beta > 0:
                V[:, K] = w / beta
            if K == kmax or beta <= tol or eager:
                lam, u, last = _dominant_ritz(H[:K, :K])
                res = abs(beta) * last
                if res <= tol:
                    conv = True
                    break
            if K == kmax:
                break
        x = V[:, :K] @ u
        if conv or numiter == maxiter:
            break
        v = x
        numiter += 1
    info = KrylovInfo(int(conv), float(res), numiter, numops)
    if not conv:
        warnings.warn(f"arnoldi_dominant not converged: normres={res:.3e} tol={tol:.3e}")
    return lam, x.reshape(shape, order="F"), info


def gmres_linsolve(op, b, x0, tol, krylovdim=Defaults.krylovdim, maxiter=Defaults.maxiter,
                   a0=1.0, a1=1.0):
    """Solve (a0 + a1 A) x = b with restarted GMRES.

    Restates KrylovKit.linsolve(f, b, x0, GMRES(tol, maxiter, krylovdim), a0, a1) as
    called from ``linsolve`` (src/algorithms/groundstates/envs.jl:14-19).  ``tol`` is an
    absolute bound on the residual norm.
    """
    shape = b.shape
    bv = b.reshape(-1, order="F").astype(np.float64)
    x = x0.reshape(-1, order="F").astype(np.float64).copy()
    n = bv.size
    kmax = min(krylovdim, n)

    def A(v):
        return a0 * v + a1 * op(v.reshape(shape, order="F")).reshape(-1, order="F")

    numops = 1
    r = bv - A(x)
    beta = float(np.linalg.norm(r))
    numiter = 0
    if beta < tol:
        return x.reshape(shape, order="F"), KrylovInfo(1, beta, 0, numops)
    while numiter < maxiter:
        numiter += 1
        V = np.zeros((n, kmax + 1))
        H = np.zeros((kmax + 1, kmax))
        cs = np.zeros(kmax)
        sn = np.zeros(kmax)
        g = np.zeros(kmax + 1)
        g[0] = beta
        V[:, 0] = r / beta
        k = 0
        while True:
            w = A(V[:, k])
            numops += 1
            w, h = _cgs2(w, V[:, : k + 1])
            hn = float(np.linalg.norm(w))
            H[: k + 1, k] = h
            H[k + 1, k] = hn
            if hn > 0:
                V[:, k + 1] = w / hn
            for i in range(k):
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            rho = math.hypot(H[k, k], H[k + 1, k])
            cs[k] = H[k, k] / rho
            sn[k] = H[k + 1, k] / rho
            H[k, k] = rho
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            k += 1
            if abs(g[k]) <= tol or k == kmax:
                break
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
        x = x + V[:, :k] @ y
        r = bv - A(x)
        numops += 1
        beta = float(np.linalg.norm(r))
        if beta < tol:
            return x.reshape(shape, order="F"), KrylovInfo(1, beta, numiter, numops)
    warnings.warn(f"gmres_linsolve not converged: normres={beta:.3e} tol={tol:.3e}")
    return x.reshape(shape, order="F"), KrylovInfo(0, beta, numiter, numops)


# ---------------------------------------------------------------------------
# Mixed-canonical state and gauge fixing (src/states/canonicalmps.jl, ortho.jl)
# ---------------------------------------------------------------------------


@dataclass
class CanonicalIMPS:
    """src/states/canonicalmps.jl:17-40.  ``C[l]`` lives on the bond right of site l
    (0-based sites; C[-1] is the bond left of site 0, i.e. C[N-1])."""

    AL: list
    AR: list
    C: list
    AC: list

    def __len__(self):
        return len(self.AL)

    def copy(self):
        return CanonicalIMPS([a.copy() for a in self.AL], [a.copy() for a in self.AR],
                             [c.copy() for c in self.C], [a.copy() for a in self.AC])


def mul_AL_C(AL, C):
    """canonicalmps.jl:45-48: AC[a,s,c] = AL[a,s,b] C[b,c]."""
    return np.einsum("asb,bc->asc", AL, C, optimize=True)


def mul_C_AR(C, AR):
    """ortho.jl:105-108: AC[x,s,y] = C[x,a] AR[a,s,y]."""
    return np.einsum("xa,asy->xsy", C, AR, optimize=True)


_GAUGE_TOLFACTOR_EIG = 1.0  # _GAUGE_ALG_EIGSOLVE, src/states/ortho.jl:6


def uniform_leftorth(A, C0, tol, maxiter, eig_miniter=10):
    """src/states/ortho.jl:125-164.  Returns (AL, C)."""
    N = len(A)
    D = C0.shape[0]
    C = [None] * N
    C[N - 1] = C0 / np.linalg.norm(C0)
    Awork = [a.copy() for a in A]
    AL = [np.zeros_like(a) for a in A]
    it = 0
    eps = math.inf
    while True:
        if it >= eig_miniter:
            tol_e = updatetol(_GAUGE_TOLFACTOR_EIG, 1, eps ** 2)
            _, evec, _ = arnoldi_dominant(lambda v: transfer_left(Awork, AL, v.reshape(-1, order="F")).reshape(D, D, order="F"),
                                          C[N - 1].copy(), tol_e)
            _, C[N - 1] = qrpos(evec)
        C_pre = C[N - 1].copy()
        for i in range(N):
            Dl, d, Dr = Awork[i].shape
            M = np.einsum("ab,bsc->asc", C[i - 1], Awork[i], optimize=True)
            Q, R = qrpos(fmat(M, Dl * d, Dr))
            AL[i] = np.reshape(Q, (Dl, d, Q.shape[1]), order="F")
            C[i] = R
        C[N - 1] = C[N - 1] / np.linalg.norm(C[N - 1])
        eps = float(np.linalg.norm(C[N - 1] - C_pre))
        it += 1
        if eps < tol:
            return AL, C, it
        if it > maxiter:
            warnings.warn(f"uniform_leftorth: not converged eps={eps:.3e}")
            return AL, C, it


def uniform_rightorth(A, C0, tol, maxiter, eig_miniter=10, AR_init=None):
    """src/states/ortho.jl:166-202.  Returns (AR, C, sweeps)."""
    N = len(A)
    D = C0.shape[0]
    C = [None] * N
    C[N - 1] = C0 / np.linalg.norm(C0)
    Awork = [a.copy() for a in A]
    AR = [np.zeros_like(a) for a in A] if AR_init is None else [a.copy() for a in AR_init]
    it = 0
    eps = math.inf
    while True:
        if it >= eig_miniter:
            tol_e = updatetol(_GAUGE_TOLFACTOR_EIG, 1, eps ** 2)
            _, evec, _ = arnoldi_dominant(lambda v: transfer_right(Awork, AR, v.reshape(-1, order="F")).reshape(D, D, order="F"),
                                          C[N - 1].copy(), tol_e)
            C[N - 1], _ = lqpos(evec)
        C_pre = C[N - 1].copy()
        for i in range(N - 1, -1, -1):
            Dl, d, Dr = Awork[i].shape
            M = np.einsum("asb,bc->asc", Awork[i], C[i], optimize=True)
            L, Q = lqpos(fmat(M, Dl, d * Dr))
            AR[i] = np.reshape(Q, (L.shape[1], d, Dr), order="F")
            C[i - 1] = L
        C[N - 1] = C[N - 1] / np.linalg.norm(C[N - 1])
        eps = float(np.linalg.norm(C[N - 1] - C_pre))
        it += 1
        if eps < tol:
            return AR, C, it
        if it > maxiter:
            warnings.warn(f"uniform_rightorth: not converged eps={eps:.3e}")
            return AR, C, it


def canonical_from_tensors(As, tol=Defaults.tolgauge, maxiter=Defaults.maxiter):
    """CanonicalIMPS(As): src/states/canonicalmps.jl:151-170 (uniform D <= d*D, so
    ``_makefullrank!`` is a no-op).  gaugefix!(:LR) with C0 = I (ortho.jl:77-99)."""
    N = len(As)
    D = As[0].shape[0]
    AL, C, _ = uniform_leftorth(As, np.eye(D), tol, maxiter)
    AR, C, _ = uniform_rightorth(AL, C[N - 1], tol, maxiter)
    AC = [mul_AL_C(AL[i], C[i]) for i in range(N)]
    return CanonicalIMPS(AL, AR, C, AC)


def random_tensors(D, ds, seed):
    """Seeded psi0 raw tensors: N(0,1) entries from numpy PCG64(seed), drawn site by site
    in column-major order (stand-in for randomimps' randn, constructors.jl:13-22)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [np.reshape(rng.standard_normal(D * d * D), (D, d, D), order="F") for d in ds]


def randomimps(D, ds, seed):
    """src/states/constructors.jl:13-22 with the seeded generator above."""
    return canonical_from_tensors(random_tensors(D, ds, seed))


def prodimps(ds, states):
    """src/states/constructors.jl:34-44 (0-based basis indices)."""
    As = []
    for d, s in zip(ds, states):
        A = np.zeros((1, d, 1))
        A[0, s, 0] = 1.0
        As.append(A)
    return canonical_from_tensors(As)


def regauge(AC, C):
    """src/states/ortho.jl:214-219: AL = Q(QRpos(AC)) Q(QRpos(C))^T."""
    Dl, d, Dr = AC.shape
    Q_AC, _ = qrpos(fmat(AC, Dl * d, Dr))
    Q_C, _ = qrpos(C.copy())
    return np.reshape(Q_AC @ Q_C.T, (Dl, d, Dr), order="F")


def mixedcanonical_errors(psi):
    """src/states/canonicalmps.jl:311-332."""
    el = er = em = 0.0
    N = len(psi)
    for l in range(N):
        AL, AR = psi.AL[l], psi.AR[l]
        g = np.einsum("xsa,xsb->ab", AL, AL)
        el = max(el, float(np.linalg.norm(g - np.eye(g.shape[0]))))
        g = np.einsum("asx,bsx->ab", AR, AR)
        er = max(er, float(np.linalg.norm(g - np.eye(g.shape[0]))))
        em = max(em, float(np.linalg.norm(mul_AL_C(AL, psi.C[l]) - mul_C_AR(psi.C[l - 1], AR))))
    return el, er, em


# ---------------------------------------------------------------------------
# Environments: DMRGCache for a Schur MPO (src/algorithms/groundstates/envs.jl:321-546)
# ---------------------------------------------------------------------------


@dataclass
class Envs:
    """DMRGCache (envs.jl:248-253): lefts[l] (D, w, D) on bond l-1, rights[l] on bond l."""

    lefts: list
    rights: list
    stats: dict = field(default_factory=dict)

    def leftenv(self, l):
        """src/environments.jl:36."""
        return self.lefts[l % len(self.lefts)]

    def rightenv(self, l):
        """src/environments.jl:39."""
        return self.rights[l % len(self.rights)]


def _left_cyclethrough(lefts, Ws, ALs, i):
    """envs.jl:121-136 (square Schur tensors): GL[site+1][:, i] = sum_{l<=i} slice pushes;
    at the last site it overwrites lefts[0]."""
    N = len(ALs)
    for site in range(N):
        snext = (site + 1) % N
        tgt = np.zeros((ALs[site].shape[2],) * 2)
        for l in range(i + 1):
            tgt += push_slice_left(lefts[site][:, l, :], Ws[site][l, :, i, :], ALs[site])
        lefts[snext][:, i, :] = tgt


def _left_closing_sweep(lefts, Ws, ALs):
    """envs.jl:143-156."""
    N = len(ALs)
    for site in range(N):
        snext = (site + 1) % N
        w = Ws[site].shape[0]
        ns = Ws[site].shape[2] - 1
        tgt = np.zeros((ALs[site].shape[2],) * 2)
        for l in range(w):
            tgt += push_slice_left(lefts[site][:, l, :], Ws[site][l, :, ns, :], ALs[site])
        lefts[snext][:, ns, :] = tgt


def _right_cyclethrough(rights, Ws, ARs, i):
    """envs.jl:162-177."""
    N = len(ARs)
    for site in range(N - 1, -1, -1):
        sprev = (site - 1) % N
        ns = Ws[site].shape[2]
        tgt = np.zeros((ARs[site].shape[0],) * 2)
        for l in range(i, ns):
            tgt += push_slice_right(rights[site][:, l, :], Ws[site][i, :, l, :], ARs[site])
        rights[sprev][:, i, :] = tgt


def dmrg_cache(psi, Ws, tol=Defaults.tol, maxiter=Defaults.maxiter,
               krylovdim=Defaults.krylovdim, init=None):
    """DMRGCache(psi, H::SparseIMPO; tol, maxiter, krylovdim, init_lefts, init_rights):
    envs.jl:321-518 for square Schur tensors and a uniform bond dimension."""
    N = len(psi)
    w = Ws[0].shape[0]
    # bond dimensions may differ from bond to bond (envs.jl:327-343): lefts[l] lives on the bond
    # left of site l, rights[l] on the bond right of it
    Dl = [psi.AL[l].shape[0] for l in range(N)]
    Dr = [psi.AL[l].shape[2] for l in range(N)]
    lefts = [np.zeros((Dl[l], w, Dl[l])) for l in range(N)]
    rights = [np.zeros((Dr[l], w, Dr[l])) for l in range(N)]
    for l in range(N):
        lefts[l][:, 0, :] = np.eye(Dl[l])
        rights[l][:, w - 1, :] = np.eye(Dr[l])
    I = np.eye(Dl[0])      # bond N-1 = bond left of site 0: the cell's fixed-point bond
    IL = [np.eye(Dl[l]) for l in range(N)]
    IR = [np.eye(Dr[l]) for l in range(N)]
    kdim = min(Dl[0] * Dl[0], krylovdim)
    stats = {"gmres_ops": 0, "solves": 0}

    def solve(op, rhs, x0):
        x, info = gmres_linsolve(op, rhs, x0, tol, kdim, maxiter, a0=1.0, a1=-1.0)
        stats["gmres_ops"] += info.numops
        stats["solves"] += 1
        return x

    AL, AR, C = psi.AL, psi.AR, psi.C

    # ---- left environments: interior levels (envs.jl:364-421) ----
    for i in range(1, w - 1):
        prev = (init.lefts[0][:, i, :].copy() if init is not None and init.lefts[0].shape == lefts[0].shape
                else lefts[0][:, i, :].copy())
        _left_cyclethrough(lefts, Ws, AL, i)
        rhs = lefts[0][:, i, :].copy()
        if isidentitylevel(Ws, i):
            rho = C[N - 1] @ C[N - 1].T

            def op(X, rho=rho):
                for l in range(N):
                    X = push_env_left(X, AL[l])
                return regularize(X, rho, I)

            lefts[0][:, i, :] = solve(op, rhs, prev)
            if N > 1:
                _left_cyclethrough(lefts, Ws, AL, i)
            for l in range(N):
                rho_l = C[l - 1] @ C[l - 1].T
                regularize(lefts[l][:, i, :], rho_l, IL[l])
        else:
            if not isemptylevel(Ws, i):
                def op(X, i=i):
                    for l in range(N):
                        X = push_slice_left(X, Ws[l][i, :, i, :], AL[l])
                    return X

                lefts[0][:, i, :] = solve(op, rhs, prev)
            if N > 1:
                _left_cyclethrough(lefts, Ws, AL, i)

    # ---- left closing level (envs.jl:424-455) ----
    i = w - 1
    prev = (init.lefts[0][:, i, :].copy() if init is not None and init.lefts[0].shape == lefts[0].shape
            else lefts[0][:, i, :].copy())
    _left_closing_sweep(lefts, Ws, AL)
    rhs = lefts[0][:, i, :].copy()
    rho = C[N - 1] @ C[N - 1].T

    def op_close(X):
        for l in range(N):
            X = push_env_left(X, AL[l])
        return regularize(X, rho, I)

    lefts[0][:, i, :] = solve(op_close, rhs, prev)
    if N > 1:
        _left_closing_sweep(lefts, Ws, AL)
    for l in range(N):
        regularize(lefts[l][:, w - 1, :], C[l - 1] @ C[l - 1].T, IL[l])

    # ---- right environments (envs.jl:461-515) ----
    for i in range(w - 2, -1, -1):
        prev = (init.rights[N - 1][:, i, :].copy() if init is not None and init.rights[N - 1].shape == rights[N - 1].shape
                else rights[N - 1][:, i, :].copy())
        _right_cyclethrough(rights, Ws, AR, i)
        rhs = rights[N - 1][:, i, :].copy()
        if isidentitylevel(Ws, i):
            rho = C[N - 1].T @ C[N - 1]

            def op(X, rho=rho):
                for l in range(N - 1, -1, -1):
                    X = push_env_right(X, AR[l])
                return regularize(X, rho, I)

            rights[N - 1][:, i, :] = solve(op, rhs, prev)
            if N > 1:
                _right_cyclethrough(rights, Ws, AR, i)
            for l in range(N):
                regularize(rights[l][:, i, :], C[l].T @ C[l], IR[l])
        else:
            if not isemptylevel(Ws, i):
                def op(X, i=i):
                    for l in range(N - 1, -1, -1):
                        X = push_slice_right(X, Ws[l][i, :, i, :], AR[l])
                    return X

                rights[N - 1][:, i, :] = solve(op, rhs, prev)
            if N > 1:
                _right_cyclethrough(rights, Ws, AR, i)
    return Envs(lefts, rights, stats)


def recalculate(envs, psi, Ws, tol):
    """recalculate! (envs.jl:530-546): rebuild warm-started from the old levels."""
    new = dmrg_cache(psi, Ws, tol=tol, init=envs)
    envs.lefts, envs.rights, envs.stats = new.lefts, new.rights, new.stats
    return envs


# ---------------------------------------------------------------------------
# Observables (src/observables/expval.jl, src/states/canonicalmps.jl:262-269)
# ---------------------------------------------------------------------------


def dot_psi(psi1, psi2, krylovdim=Defaults.krylovdim):
    """dot(psi1, psi2): dominant eigenvalue of TransferMatrix(psi2.AL, psi1.AL),
    canonicalmps.jl:262-269 (Arnoldi, :LM, tol = Defaults.tol); complex when the dominant
    eigenvalue of the real transfer matrix is complex."""
    D1 = psi1.AL[0].shape[0]
    D2 = psi2.AL[0].shape[0]
    v0 = np.eye(D1, D2)
    lam, _, _ = arnoldi_dominant(
        lambda v: transfer_left(psi2.AL, psi1.AL, v.reshape(-1, order="F")).reshape(D1, D2, order="F"),
        v0, Defaults.tol, krylovdim)
    return lam


def contract_mpo_expval(AC, GL, O, GR):
    """expval.jl:60-67."""
    X = np.einsum("bdp,pwq->bdwq", AC, GR, optimize=True)
    Y = np.einsum("bdwq,vuwd->bvuq", X, O, optimize=True)
    Z = np.einsum("bvuq,avb->auq", Y, GL, optimize=True)
    return float(np.einsum("auq,auq->", Z, AC))


def expectation_energy(psi, Ws, envs):
    """expectationvalue(psi, H::SparseIMPO, envs): expval.jl:93-112 (cell total)."""
    E = 0.0
    for l in range(len(psi)):
        W = Ws[l]
        ns = W.shape[2] - 1
        GL = envs.leftenv(l)
        GR = envs.rightenv(l)
        GRe = GR[:, ns:ns + 1, :]
        Wcol = W[:, :, ns:ns + 1, :]
        E += contract_mpo_expval(psi.AC[l], GL, Wcol, GRe)
    return E / dot_psi(psi, psi)


def energy_gradient(psi, Ws, envs):
    """d e_cell / d W[site] at fixed psi and environments: the sandwich
    <AC| GL_i (x) dW (x) GR_j |AC> / dot(psi, psi) (reverse of effective.jl:90-102 applied
    to the energy of expval.jl:93-112).  The reference has no automatic differentiation;
    at a variational fixed point this is the Hellmann-Feynman derivative."""
    nrm = dot_psi(psi, psi)
    out = []
    for l in range(len(psi)):
        AC = psi.AC[l]
        t1 = np.einsum("bia,asc->bisc", envs.leftenv(l), AC, optimize=True)
        t3 = np.einsum("buc,xjc->buxj", AC, envs.rightenv(l), optimize=True)
        out.append(np.einsum("bisx,buxj->iujs", t1, t3, optimize=True) / nrm)
    return out


def local_expectation1(psi, site, O):
    """expval.jl:33-37."""
    AC = psi.AC[site]
    return float(np.einsum("aub,us,asb->", AC, O, AC)) / dot_psi(psi, psi)


def local_expectation2(psi, i, O12):
    """expval.jl:39-56: adjacent sites (i, i+1); O12 is d1 d2 x d1 d2 (Kronecker order)."""
    N = len(psi)
    j = (i + 1) % N
    AC, AR = psi.AC[i], psi.AR[j]
    d1, d2 = AC.shape[1], AR.shape[1]
    T = np.reshape(O12, (d1, d2, d1, d2), order="F")
    E = np.einsum("aux,xvb,uvde,ady,yeb->", AC, AR, T, AC, AR, optimize=True)
    return float(E) / dot_psi(psi, psi)


# ---------------------------------------------------------------------------
# VUMPS (src/algorithms/groundstates/vumps.jl)
# ---------------------------------------------------------------------------


def calc_galerkin(psi, Ws, envs):
    """vumps.jl:111-124."""
    eps = 0.0
    for site in range(len(psi)):
        x = hac_apply(envs.leftenv(site), Ws[site], envs.rightenv(site), psi.AC[site])
        x = x / np.linalg.norm(x)
        AL = psi.AL[site]
        q = np.einsum("asc,asb->cb", AL, x, optimize=True)
        p = np.einsum("asc,cb->asb", AL, q, optimize=True)
        eps = max(eps, float(np.linalg.norm(x - p)))
    return eps


@dataclass
class VUMPS:
    """src/algorithms/algdefs.jl:26-35 (DynamicTol factors from utility.jl:308-320)."""

    D: int = Defaults.D
    tol: float = Defaults.tol
    maxiter: int = Defaults.maxiter
    verbosity: int = Defaults.verbosity
    eigs_tolfactor: float = Defaults.eigs_tolfactor
    gauge_tolfactor: float = Defaults.gauge_tolfactor
    envs_tolfactor: float = Defaults.envs_tolfactor
    krylovdim: int = Defaults.krylovdim
    eig_maxiter: int = Defaults.maxiter
    gauge_maxiter: int = Defaults.maxiter


@dataclass
class IterativeConvergenceInfo:
    """src/algorithms/algdefs.jl:180-194 plus per-iteration energies."""

    niter: int
    converged: bool
    losses: list
    energies: list

    @property
    def itererr(self):
        return self.losses[-1] if self.losses else float("nan")


def localupdate_step(psi, Ws, envs, tol_eig, alg):
    """vumps.jl:7-19."""
    ALs = []
    for site in range(len(psi)):
        GL, GR = envs.leftenv(site), envs.rightenv(site)
        _, AC, _ = lanczos_eigsolve(lambda x: hac_apply(GL, Ws[site], GR, x), psi.AC[site], tol_eig,
                                    alg.krylovdim, alg.eig_maxiter)
        GLc = envs.leftenv(site + 1)
        _, C, _ = lanczos_eigsolve(lambda x: hc_apply(GLc, GR, x), psi.C[site], tol_eig,
                                   alg.krylovdim, alg.eig_maxiter)
        ALs.append(regauge(AC, C))
    return ALs


def gauge_step(psi, ALs, tol, maxiter):
    """vumps.jl:23-32: AL <- ALs; gaugefix!(order = :R) from C0 = C[end]; AC = AL C."""
    N = len(psi)
    C0 = psi.C[N - 1].copy()
    AR, C, sweeps = uniform_rightorth(ALs, C0, tol, maxiter, AR_init=psi.AR)
    AL = [a.copy() for a in ALs]
    AC = [mul_AL_C(AL[l], C[l]) for l in range(N)]
    return CanonicalIMPS(AL, AR, C, AC), sweeps


def find_groundstate(psi0, Ws, alg=None, envs=None, callback=None):
    """find_groundstate(psi0, H::SparseIMPO, alg::VUMPS, envs; which = :SR):
    vumps.jl:51-86."""
    alg = alg or VUMPS()
    if envs is None:
        envs = dmrg_cache(psi0, Ws)
    psi = psi0.copy()
    eps = calc_galerkin(psi, Ws, envs)
    recalculate(envs, psi, Ws, updatetol(alg.envs_tolfactor, 0, eps))
    losses = [eps]
    energies = []
    converged = False
    it = 0
    for it in range(1, alg.maxiter + 1):
        siter = it - 1
        tol_eig = updatetol(alg.eigs_tolfactor, siter, eps)
        ALs = localupdate_step(psi, Ws, envs, tol_eig, alg)
        tol_gauge = updatetol(alg.gauge_tolfactor, siter, eps)
        psi, _ = gauge_step(psi, ALs, tol_gauge, alg.gauge_maxiter)
        tol_envs = updatetol(alg.envs_tolfactor, siter, eps)
        recalculate(envs, psi, Ws, tol_envs)
        eps = calc_galerkin(psi, Ws, envs)
        losses.append(eps)
        f = expectation_energy(psi, Ws, envs)
        energies.append(f)
        if callback is not None:
            callback(it, eps, f)
        if alg.verbosity > 0:
            print(f"VUMPS iter {it:4d} : eps = {eps:.3e} f = {f:.8g}")
        if eps <= alg.tol:
            converged = True
            break
    return psi, envs, IterativeConvergenceInfo(it, converged, losses, energies)


# ---------------------------------------------------------------------------
# Single-site IDMRG (src/algorithms/groundstates/idmrg.jl)
# ---------------------------------------------------------------------------


def push_env_left_mpo(L, W, A):
    """MPO-channel left push, transfermatrix.jl:11-13 (real):
    L'[a',w',b'] = A[a,u,a'] L[a,w,b] W[w,u,w',d] A[b,d,b']."""
    Y = np.einsum("awb,bdc->awdc", L, A, optimize=True)
    Y = np.einsum("wuvd,awdc->auvc", W, Y, optimize=True)
    return np.einsum("aup,auvc->pvc", A, Y, optimize=True)


def push_env_right_mpo(R, W, A):
    """MPO-channel right push, transfermatrix.jl:31-35 (real):
    R'[a',w',b'] = A[a',d,a] R[a,w,b] W[w',u,w,d] A[b',u,b]."""
    Y = np.einsum("pda,awb->pdwb", A, R, optimize=True)
    Y = np.einsum("vuwd,pdwb->pvub", W, Y, optimize=True)
    return np.einsum("pvub,qub->pvq", Y, A, optimize=True)


@dataclass
class IDMRG:
    """src/algorithms/algdefs.jl:72-82 with the Defaults of utility.jl:308-364 (Lanczos eigensolver
    and gauge with DynamicTol; QRpos orthogonal factorisation)."""

    tol: float = Defaults.tol
    maxiter: int = Defaults.maxiter
    verbosity: int = Defaults.verbosity
    eigs_tolfactor: float = Defaults.eigs_tolfactor
    gauge_tolfactor: float = Defaults.gauge_tolfactor
    krylovdim: int = Defaults.krylovdim
    eig_maxiter: int = Defaults.maxiter
    gauge_maxiter: int = Defaults.maxiter


def idmrg_sweep(psi, Ws, envs, tol_eig, alg):
    """_localupdate_sweep_idmrg! (idmrg.jl:12-34): forward sweep (eigsolve, QRpos split into AL, C,
    AC <- AL except at the last site, transfer_leftenv!), backward sweep (eigsolve, LQpos split into
    C, AR, AC <- AR except at the first site, transfer_rightenv!).  transfer_*env! (envs.jl:591-604)
    are the MPO-channel pushes without normalisation.  Returns (psi, envs, C_old, E)."""
    N = len(psi)
    C_old = psi.C[N - 1].copy()
    for pos in range(N):
        GL, GR = envs.leftenv(pos), envs.rightenv(pos)
        _, AC, _ = lanczos_eigsolve(lambda x: hac_apply(GL, Ws[pos], GR, x), psi.AC[pos], tol_eig,
                                    alg.krylovdim, alg.eig_maxiter)
        D1, d, D2 = AC.shape
        Q, R = qrpos(AC.reshape(D1 * d, D2, order="F"))
        psi.AL[pos] = Q.reshape(D1, d, D2, order="F")
        psi.C[pos] = R
        psi.AC[pos] = psi.AL[pos].copy() if pos < N - 1 else AC
        envs.lefts[(pos + 1) % N] = push_env_left_mpo(envs.lefts[pos], Ws[pos], psi.AL[pos])
    E = None
    for pos in range(N - 1, -1, -1):
        GL, GR = envs.leftenv(pos), envs.rightenv(pos)
        E, AC, _ = lanczos_eigsolve(lambda x: hac_apply(GL, Ws[pos], GR, x), psi.AC[pos], tol_eig,
                                    alg.krylovdim, alg.eig_maxiter)
        D1, d, D2 = AC.shape
        L, Q = lqpos(AC.reshape(D1, d * D2, order="F"))
        psi.C[(pos - 1) % N] = L
        psi.AR[pos] = Q.reshape(D1, d, D2, order="F")
        psi.AC[pos] = psi.AR[pos].copy() if pos > 0 else AC
        envs.rights[(pos - 1) % N] = push_env_right_mpo(envs.rights[pos], Ws[pos], psi.AR[pos])
    return psi, envs, C_old, E


def find_groundstate_idmrg(psi0, Ws, alg=None, envs=None, callback=None):
    """find_groundstate(psi0, H::SparseIMPO, alg::IDMRG, envs) (idmrg.jl:44-86): sweeps until
    eps = ||C[bond 0] - C_old|| <= tol, then the state is rebuilt from AR (CanonicalIMPS with the
    dynamic gauge tolerance) and the environments recalculated.  losses = [initial Galerkin
    error, eps per sweep]; energies = the backward sweep's final eigenvalue per sweep."""
    alg = alg or IDMRG()
    if envs is None:
        envs = dmrg_cache(psi0, Ws)
    psi = psi0.copy()
    eps = calc_galerkin(psi, Ws, envs)
    losses, energies = [eps], []
    converged = False
    it = 0
    for it in range(1, alg.maxiter + 1):
        tol_eig = updatetol(alg.eigs_tolfactor, it - 1, eps)
        psi, envs, C_old, E = idmrg_sweep(psi, Ws, envs, tol_eig, alg)
        eps = float(np.linalg.norm(psi.C[len(psi) - 1] - C_old))
        losses.append(eps)
        energies.append(E)
        if callback is not None:
            callback(it, eps, E)
        if eps <= alg.tol:
            converged = True
            break
    tol_gauge = updatetol(alg.gauge_tolfactor, it, eps)
    psi2 = canonical_from_tensors([a.copy() for a in psi.AR], tol=tol_gauge, maxiter=alg.gauge_maxiter)
    recalculate(envs, psi2, Ws, Defaults.tol)
    return psi2, envs, IterativeConvergenceInfo(it, converged, losses, energies)


# ---------------------------------------------------------------------------
# TDVP time step (src/algorithms/timeevo/tdvp.jl), real dt (imaginary time)
# ---------------------------------------------------------------------------


def lanczos_exponentiate(op, t, x0, tol, krylovdim=Defaults.krylovdim, maxiter=Defaults.maxiter):
    """exp(t op) x0 for a symmetric operator and real t: KrylovKit.exponentiate with Lanczos
    (the `integrate` of tdvp.jl:14-19; KrylovKit is not vendored, so this restates its published
    scheme).  A Lanczos basis (full reorthogonalisation) is grown from the current vector; with
    T = U diag(theta) U^T the projected exponential y(tau) = U exp(tau theta) U^T e1 is accepted for
    the whole remaining time when its residual estimate |beta_K y_K(tau)| <= tol; when the basis
    reaches krylovdim first, the largest tau = remaining / 2^j with an accepted estimate is taken,
    the vector advanced to V y(tau) and the basis rebuilt for the rest.  Returns (x, info)."""
    n = x0.size
    kmax = min(krylovdim, n)
    scale = float(np.linalg.norm(x0))
    v = x0.reshape(-1, order="F") / scale
    remaining = float(t)
    steps = 0
    converged = True
    while True:
        V = np.zeros((n, kmax + 1))
        V[:, 0] = v
        T = np.zeros((kmax + 1, kmax + 1))
        beta = 0.0
        K = 0
        accepted = None
        while True:
            w = op(V[:, K].reshape(x0.shape, order="F")).reshape(-1, order="F")
            w, h = _cgs2(w, V[:, :K + 1])
            beta = float(np.linalg.norm(w))
            T[:K + 1, K] = h
            T[K, :K + 1] = h
            K += 1
            if beta > 0:
                V[:, K] = w / beta
            theta, U = np.linalg.eigh(0.5 * (T[:K, :K] + T[:K, :K].T))

            def y_of(tau):
                return U @ (np.exp(tau * theta) * U[0, :])

            y = y_of(remaining)
            if abs(beta * y[K - 1]) <= tol or beta == 0.0:
                accepted = (remaining, y)
                break
            if K == kmax:
                tau = remaining
                for _ in range(60):
                    tau *= 0.5
                    y = y_of(tau)
                    if abs(beta * y[K - 1]) <= tol:
                        accepted = (tau, y)
                        break
                break
        steps += 1
        if accepted is None:
            converged = False
            tau, y = remaining, y_of(remaining)
        else:
            tau, y = accepted
        x = V[:, :K] @ y
        ny = float(np.linalg.norm(x))
        scale *= ny
        v = x / ny
        remaining -= tau
        if remaining == 0.0 or not converged or steps >= maxiter:
            converged = converged and remaining == 0.0
            break
    return (scale * v).reshape(x0.shape, order="F"), {"steps": steps, "converged": converged}


@dataclass
class TDVP:
    """src/algorithms/algdefs.jl:167-173: integrator = alg_expsolve() (Lanczos, utility.jl:343-347),
    alg_gauge = Defaults.alg_gauge() (tolgauge, maxiter)."""

    tol: float = Defaults.tol
    krylovdim: int = Defaults.krylovdim
    maxiter: int = Defaults.maxiter
    gauge_tol: float = Defaults.tolgauge
    gauge_maxiter: int = Defaults.maxiter


def tdvp_timestep(psi, Ws, envs, dt, alg=None):
    """timestep(psi, H, dt, alg::TDVP, envs) (tdvp.jl:33-52) for real dt: every site's AC and C
    evolved with exp(dt H_AC), exp(dt H_C); AL from regauge! (ortho.jl:214-219); the state
    rebuilt as CanonicalIMPS(ALs, C[end]) (canonicalmps.jl:180-195: gaugefix!(:R)); recalculate!."""
    alg = alg or TDVP()
    N = len(psi)
    ALs = []
    for loc in range(N):
        GL, GR = envs.leftenv(loc), envs.rightenv(loc)
        tAC, _ = lanczos_exponentiate(lambda x: hac_apply(GL, Ws[loc], GR, x), dt, psi.AC[loc], alg.tol,
                                      alg.krylovdim, alg.maxiter)
        GLc = envs.leftenv(loc + 1)
        tC, _ = lanczos_exponentiate(lambda x: hc_apply(GLc, GR, x), dt, psi.C[loc], alg.tol,
                                     alg.krylovdim, alg.maxiter)
        ALs.append(regauge(tAC, tC))
    C0 = psi.C[N - 1].copy()
    AR, C, _ = uniform_rightorth(ALs, C0, alg.gauge_tol, alg.gauge_maxiter)
    AL = [a.copy() for a in ALs]
    psi2 = CanonicalIMPS(AL, AR, C, [mul_AL_C(AL[l], C[l]) for l in range(N)])
    recalculate(envs, psi2, Ws, Defaults.tol)
    return psi2, envs


def time_evolve(psi0, Ws, t_span, alg=None, envs=None, observer=None):
    """time_evolve(psi0, H, t_span, alg::TDVP, envs) (tdvp.jl:91-112) for real steps; history holds
    observer(psi, iter, t) or the cell energy at every point of t_span."""
    alg = alg or TDVP()
    if envs is None:
        envs = dmrg_cache(psi0, Ws)
    psi = psi0.copy()
    hist = [expectation_energy(psi, Ws, envs) if observer is None else observer(psi, 0, t_span[0])]
    for it in range(len(t_span) - 1):
        psi, envs = tdvp_timestep(psi, Ws, envs, t_span[it + 1] - t_span[it], alg)
        hist.append(expectation_energy(psi, Ws, envs) if observer is None else observer(psi, it + 1, t_span[it + 1]))
    return psi, envs, hist


# ---------------------------------------------------------------------------
# Observables: correlator, fidelity (src/observables/correlators.jl, src/states/canonicalmps.jl)
# ---------------------------------------------------------------------------


def correlator(psi, O1, O2, i, js):
    """<psi| O1[i] O2[j] |psi> / dot(psi, psi) for increasing j > i (correlators.jl:13-42):
    V[b, c] = AC[i][a,u,b] O1[u,s] AC[i][a,s,c], pushed through AR of the sites between, closed
    with AR[j][b,u,e] O2[u,s] AR[j][c,s,e].  Every value is the reference's single-j form
    correlator(psi, O1, O2, i, j); for a range the sites already evaluated are pushed through
    before the next j."""
    N = len(psi)
    js = list(js)
    if not js or js[0] <= i or any(b <= a for a, b in zip(js, js[1:])):
        raise ValueError("correlator: need i < j1 < j2 < ...")
    AC = psi.AC[i % N]
    V = np.einsum("aub,us,asc->bc", AC, O1, AC, optimize=True)
    ctr = i
    out = []
    for j in js:
        for l in range(ctr + 1, j):
            V = push_env_left2(V, psi.AR[l % N], psi.AR[l % N])
        AR = psi.AR[j % N]
        out.append(float(np.einsum("bc,bue,us,cse->", V, AR, O2, AR, optimize=True)))
        V = push_env_left2(V, AR, AR)
        ctr = j
    return np.array(out) / dot_psi(psi, psi)


def norm_psi(psi):
    """LinearAlgebra.norm(psi) = norm(psi.AC[1]) (canonicalmps.jl:238)."""
    return float(np.linalg.norm(psi.AC[0]))


def fidelity(psi1, psi2, krylovdim=Defaults.krylovdim):
    """fidelity(psi1, psi2) = |dot(psi1, psi2)| / (norm(psi1) norm(psi2)) (canonicalmps.jl:281-282)."""
    return abs(dot_psi(psi1, psi2, krylovdim)) / (norm_psi(psi1) * norm_psi(psi2))


def dominant_env(psi, W=None, side="left", tol=Defaults.tol, krylovdim=Defaults.krylovdim,
                 maxiter=Defaults.maxiter):
    """dominant_env(psi; side) and dominant_env(W::DenseIMPO, psi; side) (envs.jl:71-99 with the
    transfer maps of envs.jl:34-68): the dominant (:LM) eigenpair of the unit-cell transfer map,
    L -> push_env_left(L, AL[l]) over l = 1..N (identity channel) or
    L -> push_env_left(L, W[l], AL[l]) (DenseIMPO channel), and from the right with AR over
    l = N..1; Arnoldi from v0 = ones, eager, as _eigsolve (utility.jl:379-393) calls it.  Returns
    (lambda, L) with L of shape (D, w, D) (w = 1 for the identity channel), Frobenius norm 1.
    The reference leaves the eigenvector's sign to KrylovKit; here it is fixed by
    sum(L) >= 0."""
    N = len(psi)
    D = psi.AL[0].shape[0]
    left = side == "left"
    if side not in ("left", "right"):
        raise ValueError("side must be 'left' or 'right'")
    if W is None:
        w = 1

        def op(L):
            X = L.reshape(D, D, order="F")
            for k in range(N):
                if left:
                    X = push_env_left2(X, psi.AL[k], psi.AL[k])
                else:
                    X = push_env_right2(X, psi.AR[N - 1 - k], psi.AR[N - 1 - k])
            return X.reshape(D, 1, D, order="F")
    else:
        w = W[0].shape[0]

        def op(L):
            X = L
            for k in range(N):
                if left:
                    X = push_env_left_mpo(X, W[k % len(W)], psi.AL[k])
                else:
                    l = N - 1 - k
                    X = push_env_right_mpo(X, W[l % len(W)], psi.AR[l])
            return X
    lam, x, _ = arnoldi_dominant(op, np.ones((D, w, D)), tol, krylovdim, maxiter, eager=True)
    if isinstance(lam, complex):
        raise ValueError("dominant_env: the dominant eigenvalue is complex (real arithmetic only)")
    x = x / np.linalg.norm(x)
    if x.sum() < 0:
        x = -x
    return lam, x


def dense_envs(psi, W, tol=Defaults.tol, krylovdim=Defaults.krylovdim, maxiter=Defaults.maxiter):
    """DMRGCache(psi, W::DenseIMPO) (envs.jl:284-304): lefts[0] and rights[N-1] are the dominant
    environments (dominant_env, left and right), the others follow by MPO-channel pushes through
    AL (left) and AR (right); then normalize_envs! (envs.jl:559-581): for i = 0..N-1 the right
    environment rights[i] gets unit norm and lefts[i+1] is divided by
    lambda_i = <C[i], H_C(i) C[i]> (H_C(i) with lefts[i+1] and rights[i], effective.jl:37-50)."""
    N = len(psi)
    nw = len(W)
    _, L0 = dominant_env(psi, W, "left", tol, krylovdim, maxiter)
    _, R0 = dominant_env(psi, W, "right", tol, krylovdim, maxiter)
    lefts = [None] * N
    rights = [None] * N
    lefts[0] = L0
    for l in range(1, N):
        lefts[l] = push_env_left_mpo(lefts[l - 1], W[(l - 1) % nw], psi.AL[l - 1])
    rights[N - 1] = R0
    for l in range(N - 2, -1, -1):
        rights[l] = push_env_right_mpo(rights[l + 1], W[(l + 1) % nw], psi.AR[l + 1])
    for i in range(N):
        nr = np.linalg.norm(rights[i])
        if nr > 0:
            rights[i] = rights[i] / nr
        j = (i + 1) % N
        lam = float(np.sum(psi.C[i] * hc_apply(lefts[j], rights[i], psi.C[i])))
        if abs(lam) > 0:
            lefts[j] = lefts[j] / lam
    return lefts, rights


def expectation_dense(psi, W, envs=None):
    """expectationvalue(psi, W::DenseIMPO, envs) (expval.jl:69-77): the sum over the unit cell of
    contract_mpo_expval(AC[l], lefts[l], W[l], rights[l]), divided by dot(psi, psi)."""
    lefts, rights = envs if envs is not None else dense_envs(psi, W)
    E = sum(contract_mpo_expval(psi.AC[l], lefts[l], W[l % len(W)], rights[l]) for l in range(len(psi)))
    return E / dot_psi(psi, psi)
