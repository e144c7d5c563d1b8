"""Pins the FP64 CPU restatement (oracle/imps_oracle.py) against closed forms, exact
diagonalisation, brute force and the invariants the reference's own tests assert.

The reference stores no golden-vector files (its tests only assert thresholds), so each
check below names the reference assertion it mirrors or the mathematics that fixes it.
"""
import math

import numpy as np
import pytest


def _open_chain_matrix(W, L):
    """Dense open-chain Hamiltonian from a Schur MPO: boundary vectors e_1 and e_w."""
    w = W.shape[0]
    cur = np.transpose(W[0], (1, 0, 2))
    for _ in range(1, L):
        new = np.einsum("aUD,aubd->bUuDd", cur, W)
        s = new.shape
        cur = new.reshape(s[0], s[1] * s[2], s[3] * s[4])
    return cur[w - 1]


def _kron_all(ms):
    out = np.array([[1.0]])
    for m in ms:
        out = np.kron(out, m)
    return out


# ---------------------------------------------------------------------------- tolerances


def test_updatetol_clamping(oracle):
    # test/algorithms/api.jl:453-463
    assert oracle.updatetol(0.1, 1, 5.0e-4) == pytest.approx(5.0e-5)
    assert oracle.updatetol(0.1, 1, 1.0e-20) == pytest.approx(1.0e-14)
    assert oracle.updatetol(0.1, 1, 1.0e20) == pytest.approx(1.0e-4)
    assert oracle.updatetol(0.1, 0, 5.0e-4) == oracle.updatetol(0.1, 1, 5.0e-4)
    assert oracle.updatetol(1.0, 4, 1e-6) == pytest.approx(5e-7)


# ---------------------------------------------------------------------------- models


def test_tfim_and_xxz_mpo_match_pauli_construction(oracle):
    L = 5
    X, Z = oracle.sigma_x(), oracle.sigma_z()
    I2 = np.eye(2)

    def site(op, i):
        return _kron_all([op if k == i else I2 for k in range(L)])

    J, h = 1.0, 0.7
    H = sum(-J * site(X, i) @ site(X, i + 1) for i in range(L - 1)) + sum(-h * site(Z, i) for i in range(L))
    assert np.abs(_open_chain_matrix(oracle.tfim_mpo(J, h), L) - H).max() < 1e-14

    # XXZ with the complex Sy: SxSx + SySy is real
    Sx, Sz = X / 2, Z / 2
    Sy = np.array([[0, -1j], [1j, 0]]) / 2
    Jx, Dz, hz = 1.0, 0.8, 0.3
    Hc = sum(Jx * (site(Sx, i) @ site(Sx, i + 1) + (_kron_all([Sy if k == i else I2 for k in range(L)])
                                                   @ _kron_all([Sy if k == i + 1 else I2 for k in range(L)])))
             + Jx * Dz * site(Sz, i) @ site(Sz, i + 1) for i in range(L - 1)) - sum(hz * site(Sz, i) for i in range(L))
    assert np.abs(Hc.imag).max() < 1e-15
    assert np.abs(_open_chain_matrix(oracle.xxz_mpo(Jx, Dz, hz), L) - Hc.real).max() < 1e-14


@pytest.mark.parametrize("U", [0.0, 4.0])
def test_hubbard_mpo_matches_fermion_operators(oracle, U):
    L, t, mu = 4, 1.0, U / 2
    c_up, c_dn, F, _, _ = oracle.hubbard_local_ops()
    I4 = np.eye(4)

    def fop(op, i):
        return _kron_all([F] * i + [op] + [I4] * (L - i - 1))

    cu = [fop(c_up, i) for i in range(L)]
    cd = [fop(c_dn, i) for i in range(L)]
    # canonical anticommutation relations of the Jordan-Wigner operators
    for a in cu + cd:
        for b in cu + cd:
            ac = a @ b.T + b.T @ a
            assert np.abs(ac - (np.eye(4 ** L) if a is b else 0)).max() < 1e-14
    H = np.zeros((4 ** L, 4 ** L))
    for i in range(L - 1):
        for c in (cu, cd):
            hop = c[i].T @ c[i + 1]
            H += -t * (hop + hop.T)
    for i in range(L):
        nu, nd = cu[i].T @ cu[i], cd[i].T @ cd[i]
        H += U * nu @ nd - mu * (nu + nd)
    Hm = _open_chain_matrix(oracle.hubbard_mpo(t, U, mu), L)
    assert np.abs(Hm - H).max() < 1e-14
    if U == 0.0:
        ek = -2 * t * np.cos(np.arange(1, L + 1) * np.pi / (L + 1))
        assert np.linalg.eigvalsh(Hm)[0] == pytest.approx(2 * ek[ek < 0].sum(), abs=1e-12)


def test_schur_levels(oracle):
    # test/operators/mpo.jl:50-56: unit corners, identity and empty levels
    for W in (oracle.tfim_mpo(), oracle.xxz_mpo(), oracle.hubbard_mpo()):
        w, d = W.shape[0], W.shape[1]
        assert np.array_equal(W[0, :, 0, :], np.eye(d))
        assert np.array_equal(W[w - 1, :, w - 1, :], np.eye(d))
        assert oracle.isidentitylevel([W], 0) and oracle.isidentitylevel([W], w - 1)
        for i in range(1, w - 1):
            assert not oracle.isidentitylevel([W], i)
            assert not oracle.isemptylevel([W], i)
    assert oracle.tfim_mpo().shape == (3, 2, 3, 2)
    assert oracle.xxz_mpo().shape == (5, 2, 5, 2)
    assert oracle.hubbard_mpo().shape == (6, 4, 6, 4)
    assert len(oracle.nonzero_blocks(oracle.hubbard_mpo())) == 11


# ---------------------------------------------------------------------------- H_AC / H_C


def _rand_env(rng, D, w, sym=True):
    G = rng.standard_normal((D, w, D))
    if sym:
        G = 0.5 * (G + G.transpose(2, 1, 0))
    return G


def test_hac_block_loop_equals_dense_and_brute_force(oracle):
    rng = np.random.default_rng(5)
    D, W = 3, oracle.hubbard_mpo()
    w, d = W.shape[0], W.shape[1]
    GL, GR = _rand_env(rng, D, w, False), _rand_env(rng, D, w, False)
    x = rng.standard_normal((D, d, D))
    y1 = oracle.hac_apply(GL, W, GR, x)
    y2 = oracle.hac_apply_dense(GL, W, GR, x)
    assert np.abs(y1 - y2).max() < 1e-12 * np.abs(y2).max()
    # brute force: explicit (dD^2) x (dD^2) matrix from nested loops
    n = D * d * D
    M = np.zeros((n, n))
    idx = lambda a, s, b: a + D * (s + d * b)  # column-major vec
    for b in range(D):
        for u in range(d):
            for c in range(D):
                for a in range(D):
                    for s in range(d):
                        for bb in range(D):
                            M[idx(b, u, c), idx(a, s, bb)] = sum(
                                GL[b, i, a] * W[i, u, j, s] * GR[bb, j, c] for i in range(w) for j in range(w))
    y3 = (M @ x.reshape(-1, order="F")).reshape(D, d, D, order="F")
    assert np.abs(y1 - y3).max() < 1e-12 * np.abs(y3).max()


def test_hac_identity_mpo_and_hermiticity(oracle):
    rng = np.random.default_rng(24)
    D, d = 4, 2
    # test/algorithms/api.jl:226-231: identity MPO tensor + identity environments => x
    I4 = np.eye(D).reshape(D, 1, D)
    Wid = np.eye(d).reshape(1, d, 1, d)
    x = rng.standard_normal((D, d, D))
    assert np.allclose(oracle.hac_apply(I4, Wid, I4, x), x, atol=1e-14)
    # api.jl:233-240: <x, H y> = <H x, y> with environments from a DMRGCache
    psi = oracle.randomimps(D, [d], seed=23)
    Ws = [oracle.tfim_mpo()]
    envs = oracle.dmrg_cache(psi, Ws)
    GL, GR = envs.leftenv(0), envs.rightenv(0)
    y = rng.standard_normal((D, d, D))
    lhs = np.vdot(x, oracle.hac_apply(GL, Ws[0], GR, y))
    rhs = np.vdot(oracle.hac_apply(GL, Ws[0], GR, x), y)
    assert abs(lhs - rhs) < 1e-9
    c1, c2 = rng.standard_normal((D, D)), rng.standard_normal((D, D))
    GLc = envs.leftenv(1)
    assert abs(np.vdot(c1, oracle.hc_apply(GLc, GR, c2)) - np.vdot(oracle.hc_apply(GLc, GR, c1), c2)) < 1e-9


def test_hc_matches_explicit_sum(oracle):
    rng = np.random.default_rng(3)
    D, w = 5, 4
    GL, GR = _rand_env(rng, D, w, False), _rand_env(rng, D, w, False)
    x = rng.standard_normal((D, D))
    ref = sum(GL[:, k, :] @ x @ GR[:, k, :] for k in range(w))
    assert np.abs(oracle.hc_apply(GL, GR, x) - ref).max() < 1e-13


# ---------------------------------------------------------------------------- primitives


def test_regularize_traceless(oracle):
    # test/algorithms/api.jl:242-246
    rng = np.random.default_rng(1)
    M = rng.standard_normal((3, 3))
    oracle.regularize(M, np.eye(3), np.eye(3) / 3)
    assert abs(np.trace(M)) < 1e-12
    # no conjugation/transposition mix-up: c = sum l[a,b] v[b,a] = tr(l v)
    l, v = rng.standard_normal((4, 4)), rng.standard_normal((4, 4))
    v0 = v.copy()
    oracle.regularize(v, l, np.eye(4))
    assert np.allclose(v, v0 - np.trace(l @ v0) * np.eye(4))


def test_gmres(oracle):
    # test/algorithms/api.jl:248-251: (a0 + a1 A) x = b with A = 2I, a0 = a1 = 1 => b/3
    rng = np.random.default_rng(0)
    b = rng.standard_normal(5)
    x, info = oracle.gmres_linsolve(lambda v: 2 * v, b, np.zeros(5), 1e-12, a0=1.0, a1=1.0)
    assert np.allclose(x, b / 3) and info.converged
    # random well-conditioned operator, restarts exercised (krylovdim 6 < n)
    n = 40
    A = rng.standard_normal((n, n)) / (3 * math.sqrt(n))
    b = rng.standard_normal(n)
    x, info = oracle.gmres_linsolve(lambda v: A @ v, b, np.zeros(n), 1e-12, krylovdim=6, a0=1.0, a1=-1.0)
    assert info.converged and info.numiter > 1
    assert np.allclose(x, np.linalg.solve(np.eye(n) - A, b), atol=1e-10)


def test_lanczos_matches_eigh(oracle):
    rng = np.random.default_rng(2)
    n = 200
    A = rng.standard_normal((n, n))
    A = A + A.T
    x0 = rng.standard_normal(n)
    lam, v, info = oracle.lanczos_eigsolve(lambda x: A @ x, x0, 1e-10, krylovdim=20)
    ev, U = np.linalg.eigh(A)
    assert info.converged and info.numiter > 1  # thick restarts exercised
    assert lam == pytest.approx(ev[0], abs=1e-9)
    assert abs(abs(v @ U[:, 0]) - 1) < 1e-9
    assert v @ x0 >= 0


def test_arnoldi_dominant(oracle):
    rng = np.random.default_rng(4)
    n = 60
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lamd = np.concatenate([[1.5], 0.9 * rng.uniform(-1, 1, n - 1)])
    A = Q @ np.diag(lamd) @ np.linalg.inv(Q) + 0.0
    lam, v, info = oracle.arnoldi_dominant(lambda x: A @ x, rng.standard_normal(n), 1e-12, krylovdim=12)
    assert info.converged
    assert lam == pytest.approx(1.5, abs=1e-10)
    assert np.linalg.norm(A @ v - lam * v) < 1e-9


def test_qrpos_lqpos(oracle):
    rng = np.random.default_rng(6)
    M = rng.standard_normal((12, 5))
    Q, R = oracle.qrpos(M)
    assert np.allclose(Q @ R, M) and np.allclose(Q.T @ Q, np.eye(5))
    assert np.all(np.diag(R) > 0) and np.allclose(np.tril(R, -1), 0)
    L, Qr = oracle.lqpos(M.T)
    assert np.allclose(L @ Qr, M.T) and np.allclose(Qr @ Qr.T, np.eye(5))
    assert np.all(np.diag(L) > 0) and np.allclose(np.triu(L, 1), 0)
    # uniqueness: a sign-flipped input gives the same R and a sign-flipped Q
    Q2, R2 = oracle.qrpos(-M)
    assert np.allclose(R2, R) and np.allclose(Q2, -Q)


# ---------------------------------------------------------------------------- states / gauge


def test_transfer_matrix_and_gauge(oracle):
    psi = oracle.randomimps(5, [2, 3], seed=0)
    el, er, em = oracle.mixedcanonical_errors(psi)
    assert max(el, er, em) < 1e-12
    # test/algorithms/envs.jl:15-22: TransferMatrix == sequential pushes
    L = np.eye(5)
    w = oracle.transfer_left(psi.AL, psi.AL, L.reshape(-1, order="F")).reshape(5, 5, order="F")
    Lseq = oracle.push_env_left(oracle.push_env_left(L, psi.AL[0]), psi.AL[1])
    assert np.linalg.norm(w - Lseq) / np.linalg.norm(Lseq) < 1e-12
    # envs.jl:25-28: AL transfer fixed point eigenvalue 1
    lam = oracle.dot_psi(psi, psi)
    assert abs(lam - 1) < 1e-10
    # api.jl:166-169: regauge!(AC, C) gives AL with AL C = AC
    ALr = oracle.regauge(psi.AC[0], psi.C[0])
    assert np.allclose(oracle.mul_AL_C(ALr, psi.C[0]), psi.AC[0], atol=1e-9)
    assert abs(np.linalg.norm(psi.AC[0]) - 1) < 1e-12


# ---------------------------------------------------------------------------- envs / energy


def test_energy_product_state_and_rdm_crosscheck(oracle):
    Ws = [oracle.tfim_mpo()]
    # test/operators/mpo.jl:83-85: all-up product state, E = -h <sz> = -1
    psi = oracle.prodimps([2], [0])
    envs = oracle.dmrg_cache(psi, Ws)
    assert abs(oracle.expectation_energy(psi, Ws, envs) + 1.0) < 1e-12
    # api.jl:510-527 (generic state): E_cell = sum_l (-<X_l X_l+1> - <Z_l>)
    psi = oracle.randomimps(6, [2, 2], seed=3)
    Ws2 = Ws * 2
    envs = oracle.dmrg_cache(psi, Ws2, tol=1e-13)
    E = oracle.expectation_energy(psi, Ws2, envs)
    X, Z = oracle.sigma_x(), oracle.sigma_z()
    psi_shift = oracle.CanonicalIMPS(psi.AL[1:] + psi.AL[:1], psi.AR[1:] + psi.AR[:1],
                                     psi.C[1:] + psi.C[:1], psi.AC[1:] + psi.AC[:1])
    Erdm = sum(-oracle.local_expectation2(p, 0, np.kron(X, X)) - oracle.local_expectation1(p, 0, Z)
               for p in (psi, psi_shift))
    assert abs(E - Erdm) < 1e-9
    # api.jl:207-210: recalculate! leaves the energy unchanged
    oracle.recalculate(envs, psi, Ws2, 1e-12)
    assert abs(oracle.expectation_energy(psi, Ws2, envs) - E) < 1e-9
    # regularised identity levels are traceless against rho
    w = Ws2[0].shape[0]
    for l in range(2):
        rho = psi.C[l - 1] @ psi.C[l - 1].T
        assert abs(np.sum(rho * envs.lefts[l][:, w - 1, :].T)) < 1e-10
        assert np.array_equal(envs.lefts[l][:, 0, :], np.eye(6))
        assert np.array_equal(envs.rights[l][:, w - 1, :], np.eye(6))


def test_galerkin_zero_on_exact_eigenstate(oracle):
    # diagonal on-site model: the lowest product state is an exact eigen-MPS
    lam = np.array([-2.0, 1.0])
    W = oracle.mpohamiltonian(np.diag(lam), [])
    psi = oracle.prodimps([2], [0])
    envs = oracle.dmrg_cache(psi, [W])
    assert oracle.calc_galerkin(psi, [W], envs) < 1e-14
    assert oracle.expectation_energy(psi, [W], envs) == pytest.approx(-2.0, abs=1e-13)


def test_vumps_tfim_critical_energy(oracle):
    # test/algorithms/vumps.jl:14-18: TFIM J = h = 1, |e + 4/pi| < 1e-6 at D = 12, 2-site cell
    Ws = [oracle.tfim_mpo()] * 2
    psi0 = oracle.randomimps(12, [2, 2], seed=0)
    psi, envs, info = oracle.find_groundstate(psi0, Ws, oracle.VUMPS(tol=1e-10))
    assert info.converged and info.itererr < 1e-9
    assert abs(info.energies[-1] / 2 + 4 / math.pi) < 1e-6
    # api.jl:488: calc_galerkin(psi, H, envs) == info.itererr
    assert oracle.calc_galerkin(psi, Ws, envs) == info.itererr
    # Hellmann-Feynman at h = 1: de/dh = -<sz> = -2/pi (critical: D-limited accuracy)
    assert abs(oracle.local_expectation1(psi, 0, oracle.sigma_z()) - 2 / math.pi) < 1e-3


def test_vumps_xxx_energy(oracle):
    # test/algorithms/vumps.jl:4-12: XXX, |e - (1/4 - ln 2)| < 2e-4, Galerkin < 1e-9
    Ws = [oracle.xxz_mpo()] * 2
    psi0 = oracle.randomimps(12, [2, 2], seed=1)
    psi, envs, info = oracle.find_groundstate(psi0, Ws, oracle.VUMPS(tol=1e-10))
    assert info.itererr < 1e-9
    assert abs(info.energies[-1] / 2 - (0.25 - math.log(2))) < 2e-4


def test_energy_gradient_onsite_block_and_fd(oracle):
    # on-site block of the sandwich contracted with O is the one-site expectation value
    W = oracle.tfim_mpo(1.0, 1.3)
    psi = oracle.randomimps(6, [2], seed=4)
    envs = oracle.dmrg_cache(psi, [W], tol=1e-12)
    G = oracle.energy_gradient(psi, [W], envs)[0]
    w = W.shape[0]
    for O in (oracle.sigma_z(), oracle.sigma_x()):
        assert abs(np.sum(G[0, :, w - 1, :] * O) - oracle.local_expectation1(psi, 0, O)) < 1e-12
    # at the variational optimum the sandwich is the derivative of the converged energy:
    # de/dh with H = -J sum sx sx - h sum sz (central differences, step 1e-5)
    def e_of(h):
        Wh = oracle.tfim_mpo(1.0, h)
        p0 = oracle.randomimps(8, [2], seed=1)
        p, en, info = oracle.find_groundstate(p0, [Wh], oracle.VUMPS(tol=1e-11))
        return p, en, info
    h0, dh = 1.3, 1e-5
    p, en, info = e_of(h0)
    G = oracle.energy_gradient(p, [oracle.tfim_mpo(1.0, h0)], en)[0]
    grad = np.sum(G[0, :, w - 1, :] * (-oracle.sigma_z()))
    fd = (e_of(h0 + dh)[2].energies[-1] - e_of(h0 - dh)[2].energies[-1]) / (2 * dh)
    assert abs(grad - fd) < 1e-6


def test_mpo_pushes_match_slices_and_brute_force(oracle):
    # transfermatrix.jl:11-13 / 31-35 level by level equal the Schur slice pushes
    # (envs.jl:104-114) summed over the incoming level, and an explicit loop on a tiny case;
    # with a one-level identity MPO they reduce to the identity-channel pushes (5-7, 23-26)
    rng = np.random.default_rng(41)
    D, d, w = 5, 3, 4
    A = rng.standard_normal((D, d, D))
    L = rng.standard_normal((D, w, D))
    R = rng.standard_normal((D, w, D))
    W = rng.standard_normal((w, d, w, d))
    Lp = oracle.push_env_left_mpo(L, W, A)
    Rp = oracle.push_env_right_mpo(R, W, A)
    for v in range(w):
        sl = sum(oracle.push_slice_left(L[:, k, :], W[k, :, v, :], A) for k in range(w))
        sr = sum(oracle.push_slice_right(R[:, k, :], W[v, :, k, :], A) for k in range(w))
        assert np.allclose(Lp[:, v, :], sl, atol=1e-12)
        assert np.allclose(Rp[:, v, :], sr, atol=1e-12)
    ref = np.zeros_like(Lp)
    for a, u, ap, k, b, v, s, bp in np.ndindex(D, d, D, w, D, w, d, D):
        ref[ap, v, bp] += A[a, u, ap] * L[a, k, b] * W[k, u, v, s] * A[b, s, bp]
    assert np.allclose(Lp, ref, atol=1e-12)
    I1 = np.eye(d).reshape(1, d, 1, d)
    L1 = L[:, :1, :]
    assert np.allclose(oracle.push_env_left_mpo(L1, I1, A)[:, 0, :], oracle.push_env_left(L1[:, 0, :], A))
    R1 = R[:, :1, :]
    assert np.allclose(oracle.push_env_right_mpo(R1, I1, A)[:, 0, :], oracle.push_env_right(R1[:, 0, :], A))


def test_idmrg_xxx_energy_matches_vumps(oracle):
    # test/algorithms/idmrg.jl:1-20: IDMRG (D = 12) within 2e-4 of 1/4 - ln 2 and of VUMPS
    W = oracle.xxz_mpo(1.0, 1.0)
    exact = 0.25 - math.log(2)
    psi = oracle.randomimps(12, [2, 2], 3)
    psi_i, envs_i, info = oracle.find_groundstate_idmrg(psi, [W, W], oracle.IDMRG(tol=1e-9, maxiter=60))
    ei = oracle.expectation_energy(psi_i, [W, W], envs_i) / 2
    _, _, iv = oracle.find_groundstate(psi, [W, W], oracle.VUMPS(tol=1e-10, maxiter=300))
    ev = iv.energies[-1] / 2
    assert abs(ei - exact) < 2e-4
    assert abs(ei - ev) < 2e-4
    # the sweep keeps the bond dimension and leaves a canonical state
    assert all(a.shape == (12, 2, 12) for a in psi_i.AL)
    assert max(oracle.mixedcanonical_errors(psi_i)) < 1e-8
    # the loss is the bond-0 centre drift, shrinking over the sweeps
    assert info.losses[-1] < 1e-2 * info.losses[1]


def test_lanczos_exponentiate_matches_expm(oracle):
    # exp(t A) x against scipy's dense expm, with one Krylov space (krylovdim 30) and with the
    # sub-stepping that a short space forces (krylovdim 8); real t of both signs
    import scipy.linalg
    rng = np.random.default_rng(3)
    n = 120
    A = rng.standard_normal((n, n))
    A = 0.5 * (A + A.T)
    x = rng.standard_normal(n)
    for t in (-0.05, -1.0, 0.3):
        ref = scipy.linalg.expm(t * A) @ x
        for kd in (30, 8):
            y, info = oracle.lanczos_exponentiate(lambda v: A @ v, t, x, 1e-10, kd, 300)
            assert info["converged"]
            assert np.linalg.norm(y - ref) < 1e-10 * np.linalg.norm(ref)


def test_tdvp_imaginary_time_xxx(oracle):
    # test/algorithms/tdvp.jl:28-35: imaginary time (dt = -0.05 up to t = 20) from a random D = 8
    # state reaches 1/4 - ln 2 within 2e-3; the energy never rises along the way
    W = oracle.xxz_mpo(1.0, 1.0)
    psi0 = oracle.randomimps(8, [2, 2], 5)
    t_span = -np.arange(0, 20.0001, 0.05)
    psi, envs, hist = oracle.time_evolve(psi0, [W, W], t_span)
    assert len(hist) == 401
    assert abs(hist[-1] / 2 - (0.25 - math.log(2))) < 2e-3
    assert all(hist[i + 1] <= hist[i] + 1e-10 for i in range(len(hist) - 1))
    assert max(oracle.mixedcanonical_errors(psi)) < 1e-8


def test_correlator_and_fidelity(oracle):
    # test/algorithms/observables.jl:6-13: |up up> gives <Z1 Zj> = 1 and <X1 Xj> = 0; Neel <Z1 Z2> = -1
    Z = oracle.sigma_z()
    X = oracle.sigma_x()
    up = oracle.prodimps([2, 2], [0, 0])
    assert np.allclose(oracle.correlator(up, Z, Z, 0, range(1, 5)), 1.0, atol=1e-12)
    assert np.allclose(oracle.correlator(up, X, X, 0, range(1, 4)), 0.0, atol=1e-12)
    neel = oracle.prodimps([2, 2], [0, 1])
    assert abs(oracle.correlator(neel, Z, Z, 0, [1])[0] + 1.0) < 1e-12
    # entangled state: the nearest-neighbour value is the two-site expectation value, a range
    # equals its single-j values, and at long distance the correlator factorises
    psi = oracle.randomimps(6, [2, 2], 8)
    c = oracle.correlator(psi, Z, X, 0, range(1, 7))
    # expval.jl:49 reshapes O12 column-major, so kron(A, B) puts B on site i and A on site i+1
    assert abs(c[0] - oracle.local_expectation2(psi, 0, np.kron(X, Z))) < 1e-12
    for k, j in enumerate(range(1, 7)):
        assert abs(c[k] - oracle.correlator(psi, Z, X, 0, [j])[0]) < 1e-12
    far = oracle.correlator(psi, Z, X, 0, [301])[0]
    fact = oracle.local_expectation1(psi, 0, Z) * oracle.local_expectation1(psi, 1, X)
    assert abs(far - fact) < 1e-8
    # fidelity: 1 with itself and under rescaling, < 1 between different states, symmetric
    assert abs(oracle.fidelity(psi, psi) - 1.0) < 1e-10
    psi2 = oracle.randomimps(6, [2, 2], 9)
    f12, f21 = oracle.fidelity(psi, psi2), oracle.fidelity(psi2, psi)
    assert 0.0 <= f12 < 0.999 and abs(f12 - f21) < 1e-10


def _explicit_mixed_transfer(oracle, psi1, psi2):
    # TransferMatrix(psi2.AL, psi1.AL) as an explicit D^2 x D^2 matrix, built column by column
    # from single-site pushes of unit vectors
    D = psi1.AL[0].shape[0]
    T = np.eye(D * D)
    for a1, a2 in zip(psi1.AL, psi2.AL):
        M = np.column_stack([oracle.transfer_left([a2], [a1], e) for e in np.eye(D * D)])
        T = M @ T
    return T


@pytest.mark.parametrize("N,D,s1,s2", [(1, 4, 1003, 2003), (2, 3, 1001, 2001), (2, 6, 8, 9), (1, 5, 40, 41)])
def test_dot_matches_explicit_transfer_matrix(oracle, N, D, s1, s2):
    # dot(psi1, psi2) (canonicalmps.jl:262-269) is the dominant eigenvalue of the mixed transfer
    # matrix; the first two pairs have a complex dominant pair, which dot must return as the
    # complex eigenvalue (fidelity |lambda|), the others a real one
    psi1, psi2 = oracle.randomimps(D, [2] * N, s1), oracle.randomimps(D, [2] * N, s2)
    ev = np.linalg.eigvals(_explicit_mixed_transfer(oracle, psi1, psi2))
    lam = ev[np.argmax(np.abs(ev))]
    got = oracle.dot_psi(psi1, psi2)
    if abs(lam.imag) > 1e-12:
        assert isinstance(got, complex)
        assert abs(got.real - lam.real) < 1e-10 and abs(abs(got.imag) - abs(lam.imag)) < 1e-10
    else:
        assert abs(got - lam.real) < 1e-10
    assert abs(oracle.fidelity(psi1, psi2) - abs(lam)) < 1e-10


def test_arnoldi_complex_dominant_pair(oracle):
    # a real operator whose dominant eigenvalues are the pair 1.2 exp(+-0.7i): Arnoldi returns
    # the complex value, converges, and its vector lies in the pair's invariant plane
    rng = np.random.default_rng(12)
    n = 50
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    B = np.diag(0.8 * rng.uniform(-1, 1, n))
    r, th = 1.2, 0.7
    B[:2, :2] = [[r * np.cos(th), -r * np.sin(th)], [r * np.sin(th), r * np.cos(th)]]
    A = Q @ B @ Q.T
    lam, v, info = oracle.arnoldi_dominant(lambda x: A @ x, rng.standard_normal(n), 1e-12, krylovdim=12)
    assert info.converged and isinstance(lam, complex)
    assert abs(lam.real - r * np.cos(th)) < 1e-10 and abs(abs(lam.imag) - r * np.sin(th)) < 1e-10
    P = Q[:, :2] @ Q[:, :2].T
    assert np.linalg.norm(v - P @ v) < 1e-9 * np.linalg.norm(v)


def test_dominant_env_closed_forms_and_brute_force(oracle):
    # identity channel (envs.jl:79-99): AL is left- and AR right-orthonormal, so the fixed points
    # are lambda = 1 with L = R = I / sqrt(D) (the reference test's |lambda - 1| and norm checks)
    psi = oracle.randomimps(6, [2, 2], 11)
    for side in ("left", "right"):
        lam, L = oracle.dominant_env(psi, side=side)
        assert L.shape == (6, 1, 6)
        assert abs(lam - 1) < 1e-10
        assert np.allclose(L[:, 0, :], np.eye(6) / math.sqrt(6), atol=1e-8)
    # a one-level DenseIMPO holding the identity is the identity channel
    lam, L = oracle.dominant_env(psi, [np.eye(2).reshape(1, 2, 1, 2)] * 2, side="left")
    assert abs(lam - 1) < 1e-10 and np.allclose(L[:, 0, :], np.eye(6) / math.sqrt(6), atol=1e-8)
    # a one-level operator O on a product state |0 0>: lambda = O[0, 0]^2 per two-site cell
    up = oracle.prodimps([2, 2], [0, 0])
    O = np.array([[2.0, 0.3], [0.5, 0.7]])
    lam, L = oracle.dominant_env(up, [O.reshape(1, 2, 1, 2)] * 2, side="right")
    assert abs(lam - 4.0) < 1e-12 and abs(L[0, 0, 0] - 1.0) < 1e-12
    # dense random W (non-symmetric blocks, w = 3): the dominant eigenpair of the explicit
    # transfer matrix built column by column from the unit-cell map (up to the sign convention)
    rng = np.random.default_rng(5)
    psi = oracle.randomimps(4, [2, 2], 12)
    Ws = [np.asfortranarray(rng.standard_normal((3, 2, 3, 2))) for _ in range(2)]
    for side in ("left", "right"):
        lam, L = oracle.dominant_env(psi, Ws, side=side, tol=1e-12)
        n = 4 * 3 * 4
        T = np.zeros((n, n))
        for k in range(n):
            e = np.zeros(n)
            e[k] = 1.0
            X = e.reshape(4, 3, 4, order="F")
            for j in range(2):
                X = (oracle.push_env_left_mpo(X, Ws[j], psi.AL[j]) if side == "left"
                     else oracle.push_env_right_mpo(X, Ws[1 - j], psi.AR[1 - j]))
            T[:, k] = X.reshape(-1, order="F")
        vals, vecs = np.linalg.eig(T)
        i = int(np.argmax(np.abs(vals)))
        assert abs(vals[i].imag) < 1e-12 and abs(vals[i].real - lam) < 1e-9 * abs(lam)
        v = vecs[:, i].real / np.linalg.norm(vecs[:, i].real)
        v = v if v.sum() >= 0 else -v
        assert np.linalg.norm(v - L.reshape(-1, order="F")) < 1e-7


def test_dense_envs_and_expectation_identities(oracle):
    # DMRGCache(psi, DenseIMPO) + expectationvalue (envs.jl:284-304, 559-581; expval.jl:69-77).
    # For a one-site cell, AC = AL C and the left environment's eigen-equation give
    # contract(AC, L, W, R) = Lambda <C, H_C C>, so the normalised value is the dominant eigenvalue
    # Lambda of the cell's transfer map; a doubled cell gives 2 Lambda (this fixes the wrap of the
    # normalisation); the value is linear in a scale of W; the identity MPO gives N.
    rng = np.random.default_rng(8)
    Wd = 0.3 * rng.standard_normal((3, 2, 3, 2))
    Wd[0, :, 0, :] += np.eye(2)
    Wd = np.asfortranarray(Wd)
    psi1 = oracle.randomimps(6, [2], 31)
    lam, _ = oracle.dominant_env(psi1, [Wd], "left")
    E1 = oracle.expectation_dense(psi1, [Wd])
    assert abs(E1 - lam) < 1e-9 * abs(lam)
    psi2 = oracle.canonical_from_tensors([psi1.AL[0], psi1.AL[0]])
    E2 = oracle.expectation_dense(psi2, [Wd, Wd])
    assert abs(E2 - 2 * lam) < 1e-8 * abs(lam)
    # two-site cell with a real dominant eigenvalue (-0.657 against 0.405 next, explicit spectrum);
    # with the state of seed 32 the same channel's dominant pair is complex (-0.350 +- 0.722i),
    # which the real-valued dominant environment refuses
    Ws = [Wd, np.asfortranarray(Wd[:, ::-1, :, :])]
    with pytest.raises(ValueError, match="complex"):
        oracle.expectation_dense(oracle.randomimps(6, [2, 2], 32), Ws)
    psi = oracle.randomimps(6, [2, 2], 38)
    E = oracle.expectation_dense(psi, Ws)
    assert abs(oracle.expectation_dense(psi, [2.5 * x for x in Ws]) - 2.5 * E) < 1e-8 * abs(E)
    assert abs(oracle.expectation_dense(psi, [np.eye(2).reshape(1, 2, 1, 2)] * 2) - 2.0) < 1e-10
    lefts, rights = oracle.dense_envs(psi, Ws)
    for i in range(2):
        assert abs(np.linalg.norm(rights[i]) - 1) < 1e-12
        j = (i + 1) % 2
        lam_i = np.sum(psi.C[i] * oracle.hc_apply(lefts[j], rights[i], psi.C[i]))
        assert abs(lam_i - 1) < 1e-10
