"""CPU tests of the boundary: libimps.so loads, exports every symbol include/imps.h
declares, its host-side Krylov numerics match numpy, and the host models/validation
match the oracle.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "imps.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(imps_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    from infinitempsalgorithms_amd import _lib
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    syms = _header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    from infinitempsalgorithms_amd import _lib
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(syms) == bound, set(syms) ^ bound


def test_default_params_match_reference_defaults(lib):
    from infinitempsalgorithms_amd import _lib
    p = _lib.VumpsParams()
    assert lib.imps_vumps_default_params(ctypes.byref(p)) == 0
    # src/utility.jl:308-320
    assert (p.tol, p.maxiter, p.krylovdim) == (1e-10, 300, 30)
    assert (p.tol_min, p.tol_max) == (1e-14, 1e-4)
    assert (p.eigs_tolfactor, p.gauge_tolfactor, p.envs_tolfactor) == (1e-3, 1e-6, 1e-4)
    # plain FP64, no mixed-precision switch or floor
    assert (p.prec, p.refine_below, p.eig_floor) == (0, 0.0, 0.0)
    # Defaults.alg_eigsolve / alg_gauge / alg_environments (utility.jl:322-364): DynamicTol-wrapped
    # eager Lanczos, (tol = tolgauge, maxiter) and (tol, maxiter); identity finalize
    assert (p.eig_dynamic, p.eig_eager, p.eig_tol) == (1, 1, 1e-10)
    assert (p.gauge_dynamic, p.gauge_tol, p.gauge_tol_min, p.gauge_tol_max) == (1, 1e-13, 1e-14, 1e-4)
    assert (p.envs_dynamic, p.envs_tol, p.envs_tol_min, p.envs_tol_max, p.envs_krylovdim) == (1, 1e-10, 1e-14, 1e-4, 30)
    assert not p.finalize and p.finalize_user is None
    assert p.refine_eig == 1  # mixed runs refine with reduced-precision corrections of FP64 residuals


def test_vumps_solver_objects_map_to_params(lib):
    # VUMPS(alg_eigsolve, alg_gauge, alg_environments) (algdefs.jl:26-35): a DynamicTol carries its
    # own bounds and factor, a bare solver a fixed tolerance (updatetol is then the identity)
    import infinitempsalgorithms_amd as imps
    p = imps.VUMPS(alg_eigsolve=imps.alg_eigsolve(dynamic_tols=False, tol=1e-7, krylovdim=20, eager=False),
                   alg_gauge=imps.alg_gauge(tol_factor=1e-5, tol_min=1e-13, tol_max=1e-3, maxiter=50),
                   alg_environments=imps.SolverTol(tol=1e-11)).params()
    assert (p.eig_dynamic, p.eig_tol, p.krylovdim, p.eig_eager) == (0, 1e-7, 20, 0)
    assert (p.gauge_dynamic, p.gauge_tolfactor, p.gauge_tol_min, p.gauge_tol_max, p.gauge_maxiter) == \
        (1, 1e-5, 1e-13, 1e-3, 50)
    assert (p.envs_dynamic, p.envs_tol) == (0, 1e-11)
    d = imps.VUMPS().params()
    assert (d.eig_dynamic, d.gauge_dynamic, d.envs_dynamic) == (1, 1, 1)
    with pytest.raises(ValueError):
        imps.DynamicTol(imps.Lanczos(), tol_min=1e-3, tol_max=1e-4)


def test_host_symmetric_eig(lib):
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 31):
        A = rng.standard_normal((n, n))
        A = np.asfortranarray(A + A.T)
        ref = np.linalg.eigvalsh(A)
        w = np.zeros(n)
        V = np.zeros((n, n), order="F")
        Ac = A.copy(order="F")
        P = ctypes.POINTER(ctypes.c_double)
        lib.imps_host_sym_eig(n, Ac.ctypes.data_as(P), w.ctypes.data_as(P), V.ctypes.data_as(P))
        assert np.allclose(w, ref, atol=1e-12)
        assert np.allclose(A @ V, V * w, atol=1e-11)
        assert np.allclose(V.T @ V, np.eye(n), atol=1e-12)


def test_host_hessenberg_dominant(lib):
    from scipy.linalg import hessenberg
    rng = np.random.default_rng(1)
    P = ctypes.POINTER(ctypes.c_double)
    for n in (1, 3, 12, 30):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = np.concatenate([[2.0], rng.uniform(-1.5, 1.5, n - 1)])
        A = Q @ np.diag(lam) @ Q.T + 0.1 * rng.standard_normal((n, n)) / np.sqrt(n)
        H = np.asfortranarray(hessenberg(A))
        ev = np.linalg.eigvals(H)
        wr, wi = np.zeros(n), np.zeros(n)
        lib.imps_host_hess_eigvals(n, H.ctypes.data_as(P), n, wr.ctypes.data_as(P), wi.ctypes.data_as(P))
        got = np.sort_complex(wr + 1j * wi)
        assert np.allclose(got, np.sort_complex(ev), atol=1e-9)
        l = ctypes.c_double()
        u = np.zeros(n)
        st = lib.imps_host_hess_dominant(n, H.ctypes.data_as(P), n, ctypes.byref(l), u.ctypes.data_as(P))
        k = np.argmax(np.abs(ev))
        if abs(ev[k].imag) < 1e-12:
            assert st == 0
            assert l.value == pytest.approx(ev[k].real, abs=1e-10)
            assert np.linalg.norm(H @ u - l.value * u) < 1e-9


def test_models_match_oracle(oracle):
    from infinitempsalgorithms_amd import models
    assert np.array_equal(models.tfim_hamiltonian(0.7, 1.3).Ws[0], oracle.tfim_mpo(0.7, 1.3))
    assert np.array_equal(models.heisenberg_hamiltonian(1.0, 0.8, 0.2).Ws[0], oracle.xxz_mpo(1.0, 0.8, 0.2))
    assert np.array_equal(models.fermi_hubbard(1.0, 4.0, 2.0).Ws[0], oracle.hubbard_mpo(1.0, 4.0, 2.0))
    H = models.fermi_hubbard(nsites=2)
    assert len(H) == 2 and H.w == 6 and H.d == 4


def test_sparseimpo_validation():
    from infinitempsalgorithms_amd import SparseIMPO, models
    W = models.tfim_hamiltonian().Ws[0]
    with pytest.raises(ValueError):
        SparseIMPO([])
    with pytest.raises(ValueError):
        SparseIMPO([W, models.heisenberg_hamiltonian().Ws[0]])  # level chain mismatch
    bad = W.copy()
    bad[0, :, 0, :] = 0.0
    with pytest.raises(ValueError):
        SparseIMPO([bad])


def test_random_tensors_match_oracle(oracle):
    from infinitempsalgorithms_amd import random_tensors
    a = random_tensors(6, [2, 2], seed=3)
    b = oracle.random_tensors(6, [2, 2], seed=3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_gemm_autotune_mode_range(lib):
    # 0 heuristic, 1 autotune, 2 + v a forced tile variant (v = 0..5); anything else is refused
    for mode in (-1, 8, 100):
        assert lib.imps_set_gemm_autotune(mode) != 0
    for mode in (0, 7, 1):
        assert lib.imps_set_gemm_autotune(mode) == 0


def test_host_arrays_are_checked_before_the_library_reads_them(oracle):
    # the C interface takes plain pointers, so the Python mirror refuses arrays of the wrong shape
    # before a call could read past them (no GPU needed: the checks come first)
    import numpy as np
    import infinitempsalgorithms_amd as imps
    from infinitempsalgorithms_amd import api
    psi = oracle.randomimps(4, [2, 2], 3)
    st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
    other = oracle.randomimps(5, [2, 2], 4)
    st5 = imps.CanonicalIMPS(other.AL, other.AR, other.C, other.AC)

    class FakeEnvs:
        ket = st
        ctx = None
    with pytest.raises(ValueError, match="same unit cell"):
        api.dot(st, st5, FakeEnvs())
    with pytest.raises(ValueError, match="physical dimension"):
        api.correlator(st, np.eye(3), np.eye(2), 0, [1], FakeEnvs())
    with pytest.raises(ValueError, match="must be 2 x 2"):
        api.expectationvalue(st, (0, np.eye(3)), FakeEnvs())


def test_dense_model_constructors(oracle):
    # tfim / heisenberg_xxz (src/models.jl:135-155): a one-site DenseIMPO of the Schur bulk and
    # the SparseIMPO of the same model; the XXZ bulk has 5 levels (test/operators/mpo.jl:6-10)
    import numpy as np
    from infinitempsalgorithms_amd import models, DenseIMPO, SparseIMPO
    m = models.tfim(J=1.0, h=0.7)
    assert isinstance(m.mpo, DenseIMPO) and len(m.mpo) == 1 and isinstance(m.hamiltonian, SparseIMPO)
    assert np.array_equal(m.bulk, m.hamiltonian.Ws[0]) and m.bulk.shape == (3, 2, 3, 2)
    assert np.allclose(m.bulk, oracle.tfim_mpo(1.0, 0.7))
    x = models.heisenberg_xxz(J=1.0, Delta=0.5)
    assert x.bulk.shape == (5, 2, 5, 2) and len(x.mpo) == 1
    assert np.array_equal(x.bulk, x.hamiltonian.Ws[0])
    # the S+- channels give the same two-site operator as Sx Sx + Sy Sy + Delta Sz Sz
    w = x.bulk
    h2 = sum(np.kron(w[0, :, k, :], w[k, :, 4, :]) for k in range(1, 4))
    sx = np.array([[0, 0.5], [0.5, 0]]); sy = np.array([[0, -0.5j], [0.5j, 0]]); sz = np.diag([0.5, -0.5])
    assert np.allclose(h2, (np.kron(sx, sx) + np.kron(sy, sy) + 0.5 * np.kron(sz, sz)).real)
