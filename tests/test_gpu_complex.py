"""Complex states on the device (CanonicalIMPS{ComplexF64}): state round trips, canonical form,
energies and TDVP steps with complex dt (real time), checked against the complex restatement
(oracle/imps_complex.py), the closed-form precession and the reference's real-time TDVP test
(test/algorithms/tdvp.jl:15-24)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def imps():
    import infinitempsalgorithms_amd as m
    return m


@pytest.fixture(scope="module")
def cx():
    from oracle import imps_complex
    return imps_complex


@pytest.fixture(scope="module")
def ctx(imps):
    c = imps.Context(0)
    yield c
    c.close()


def _complex_raw(oracle, D, ds, seed):
    re = oracle.random_tensors(D, ds, seed)
    im = oracle.random_tensors(D, ds, seed + 1000)
    return [a + 0.4j * b for a, b in zip(re, im)]


def _as_lib(imps, p):
    return imps.CanonicalIMPS([x.copy() for x in p.AL], [x.copy() for x in p.AR], [x.copy() for x in p.C],
                              [x.copy() for x in p.AC])


def test_complex_state_round_trip_and_canonical_form(imps, ctx, oracle, cx):
    W = oracle.xxz_mpo(1.0, 0.8)
    H = imps.SparseIMPO([W, W])
    raw = _complex_raw(oracle, 4, [2, 2], 21)
    ref = cx.canonical_from_tensors(raw)
    print("complex: set_state_raw_complex N=2 D=4", flush=True)
    ctx.set_mpo(H)
    ctx.set_state_raw(raw, 1e-14, 2000)
    got = ctx.get_state()
    assert got.is_complex() and got.D == 4
    for a, b in zip(got.AL + got.AR + got.C + got.AC, ref.AL + ref.AR + ref.C + ref.AC):
        assert rel(a, b) < 1e-9
    # set_state_complex / get_state_complex round trip is exact up to the embedding's averaging
    ctx.set_state(_as_lib(imps, ref))
    back = ctx.get_state()
    for a, b in zip(back.AL + back.AR + back.C + back.AC, ref.AL + ref.AR + ref.C + ref.AC):
        assert rel(a, b) < 1e-15
    flag = __import__("ctypes").c_int()
    ctx.call("imps_state_is_complex", __import__("ctypes").byref(flag))
    assert flag.value == 1


@pytest.mark.parametrize("model,N,D,dt", [("xxz", 2, 4, -0.05j), ("tfim", 1, 6, -0.03 - 0.02j),
                                          ("xxz", 2, 4, -0.05)])
def test_complex_tdvp_step_matches_oracle(imps, oracle, cx, model, N, D, dt):
    W = oracle.xxz_mpo(1.0, 0.8) if model == "xxz" else oracle.tfim_mpo(1.0, 0.9)
    Ws = [W] * N
    psi = cx.canonical_from_tensors(_complex_raw(oracle, D, [2] * N, 31 + N))
    env = cx.dmrg_cache(psi, Ws)
    e0 = cx.expectation_energy(psi, Ws, env)
    p2, env2 = cx.tdvp_timestep(psi, Ws, env, dt)
    e2 = cx.expectation_energy(p2, Ws, env2)
    print(f"complex: TDVP step {model} N={N} D={D} dt={dt}", flush=True)
    H = imps.SparseIMPO(Ws)
    with imps.Context(0) as c:
        st = _as_lib(imps, psi)
        envs = imps.DMRGCache(st, H, ctx=c)
        assert abs(imps.expectationvalue(st, H, envs) - e0.real) < 1e-10
        g2, envs = imps.timestep(st, H, dt, imps.TDVP(tol=1e-12, gauge_tol=1e-14, gauge_maxiter=2000), envs)
        assert g2.is_complex()
        for a, b in zip(g2.AL + g2.AR + g2.C + g2.AC, p2.AL + p2.AR + p2.C + p2.AC):
            assert rel(a, b) < 1e-8
        assert abs(imps.expectationvalue(g2, H, envs) - e2.real) < 1e-10
        assert abs(imps.calc_galerkin(g2, H, envs)) >= 0.0


def test_product_state_precession_on_device(imps, oracle, cx):
    # H = -h sum sz from |+x> (D = 1): TDVP is exact, <sx>(t) = cos 2ht; the states against the
    # complex restatement fix the sign of i in the exponent (<sy>(t) = -sin 2ht there)
    h = 0.7
    W = oracle.tfim_mpo(0.0, h)
    H = imps.SparseIMPO([W])
    a = np.zeros((1, 2, 1), dtype=complex)
    a[0, :, 0] = [1 / math.sqrt(2), 1 / math.sqrt(2)]
    ref = cx.canonical_from_tensors([a])
    renv = cx.dmrg_cache(ref, [W])
    print("complex: precession D=1, 10 real-time steps", flush=True)
    with imps.Context(0) as c:
        st = _as_lib(imps, ref)
        envs = imps.DMRGCache(st, H, ctx=c)
        sx = oracle.sigma_x()
        for k in range(1, 11):
            st, envs = imps.timestep(st, H, -0.1j, imps.TDVP(tol=1e-13), envs)
            ref, renv = cx.tdvp_timestep(ref, [W], renv, -0.1j)
            ex, ey = cx.larmor(h, 0.1 * k)
            assert abs(imps.expectationvalue(st, (0, sx), envs) - ex) < 1e-10
            assert abs(cx.local_expectation1(st, 0, np.array([[0, -1j], [1j, 0]])) - ey) < 1e-10
            # the same <sy> on the device (a complex Hermitian operator)
            assert abs(imps.expectationvalue(st, (0, np.array([[0, -1j], [1j, 0]])), envs) - ey) < 1e-10
            assert rel(st.AC[0], ref.AC[0]) < 1e-10


def test_real_time_from_ground_state_reference_protocol(imps, oracle, cx):
    # test/algorithms/tdvp.jl:15-24 on the device: XXX, two-site cell, VUMPS(D = 8, tol = 1e-9)
    # ground state, t_span = -im (0:0.01:0.1): 11 history points, energy kept to 1e-6 and the
    # norm to 1e-8; the TDVP with explicit gauge parameters gives the same history to 1e-10;
    # every point equals the complex restatement's energy to 1e-9
    from infinitempsalgorithms_amd import models
    m = models.heisenberg_xxz(J=1.0, Delta=1.0)
    H = imps.SparseIMPO([m.bulk, m.bulk])
    print("complex: XXX D=8 ground state, then 10 real-time steps", flush=True)
    with imps.Context(0) as c:
        g, envs, info = imps.find_groundstate(None, H, imps.VUMPS(D=8, maxiter=200, tol=1e-9), ctx=c, seed=3)
        e0 = imps.expectationvalue(g, H, envs) / 2
        tspan = [-1j * 0.01 * k for k in range(11)]
        obs = lambda p, k, t: imps.expectationvalue(p, H, envs) / 2
        pt, envs, hist = imps.time_evolve(g, H, tspan, imps.TDVP(), envs, observer=obs)
        assert len(hist) == 11
        assert abs(hist[-1] - e0) < 1e-6
        assert abs(imps.norm(pt) - 1) < 1e-8
        assert pt.is_complex() and float(np.abs(pt.AC[0].imag).max()) > 1e-4
        with imps.Context(0) as c2:
            envs2 = imps.DMRGCache(g, H, ctx=c2)
            obs2 = lambda p, k, t: imps.expectationvalue(p, H, envs2) / 2
            _, _, hist2 = imps.time_evolve(g, H, tspan, imps.TDVP(gauge_tol=1e-13, gauge_maxiter=300), envs2,
                                           observer=obs2)
        assert np.allclose(hist2, hist, atol=1e-10, rtol=0)
    ref = cx.from_real(g)
    renv = cx.dmrg_cache(ref, [m.bulk, m.bulk])
    hist_o = [cx.expectation_energy(ref, [m.bulk, m.bulk], renv).real / 2]
    for _ in range(10):
        ref, renv = cx.tdvp_timestep(ref, [m.bulk, m.bulk], renv, -0.01j)
        hist_o.append(cx.expectation_energy(ref, [m.bulk, m.bulk], renv).real / 2)
    assert np.allclose(hist, hist_o, atol=1e-9, rtol=0)


def test_complex_vumps_matches_real_ground_state(imps, oracle):
    # find_groundstate from randomimps(ComplexF64, ...): the complex run reaches the real run's
    # ground-state energy (XXZ, two-site cell, D = 8)
    W = oracle.xxz_mpo(1.0, 0.5)
    H = imps.SparseIMPO([W, W])
    print("complex: VUMPS XXZ D=8 from a complex random state", flush=True)
    with imps.Context(0) as c:
        st, _ = imps.randomimps(8, H, seed=4, ctx=c, dtype=complex)
        assert st.is_complex()
        gc, envs, info_c = imps.find_groundstate(st, H, imps.VUMPS(tol=1e-10, maxiter=500), ctx=c)
        assert gc.is_complex() and info_c.converged
    with imps.Context(0) as c:
        sr, _ = imps.randomimps(8, H, seed=4, ctx=c)
        _, _, info_r = imps.find_groundstate(sr, H, imps.VUMPS(tol=1e-10, maxiter=500), ctx=c)
    assert abs(info_c.energy - info_r.energy) < 1e-9


def test_complex_state_refuses_real_only_entry_points(imps, oracle, cx):
    from infinitempsalgorithms_amd import _lib
    W = oracle.tfim_mpo(1.0, 0.9)
    H = imps.SparseIMPO([W])
    psi = cx.canonical_from_tensors(_complex_raw(oracle, 4, [2], 41))
    with imps.Context(0) as c:
        st = _as_lib(imps, psi)
        envs = imps.DMRGCache(st, H, ctx=c)
        with pytest.raises(_lib.ImpsError, match="status -5"):
            imps.regauge_(envs, 0)
        with pytest.raises(_lib.ImpsError, match="status -5"):
            imps.expectationvalue(st, (0, np.array([[0.0, 1.0], [0.0, 0.0]])), envs)
        # the real-state overlap entry refuses a complex context state (imps_dot_state_complex)
        pp, keep = _lib.ptr_array([np.zeros((8, 2, 8), order="F")])
        with pytest.raises(_lib.ImpsError, match="status -5"):
            c.call("imps_dot_state", 1, pp, np.zeros(2).ctypes.data_as(_lib.P_double))
        # symmetric operators are whole on complex states
        sz = oracle.sigma_z()
        assert abs(imps.expectationvalue(st, (0, sz), envs) - cx.local_expectation1(psi, 0, sz).real) < 1e-12


def test_gaugefix_right_of_a_freshly_set_state_keeps_it(imps, oracle, cx):
    # imps_gaugefix_right before any regauge re-fixes the gauge of the state itself (its pending
    # ALs are its ALs): a canonical state comes back unchanged, real and complex, and a TDVP step
    # after it runs
    import ctypes
    W = oracle.xxz_mpo(1.0, 0.8)
    H = imps.SparseIMPO([W, W])
    real = oracle.randomimps(6, [2, 2], 17)
    cplx = cx.canonical_from_tensors(_complex_raw(oracle, 6, [2, 2], 19))
    print("complex: gaugefix_right after set_state", flush=True)
    for psi in (real, cplx):
        with imps.Context(0) as c:
            st = _as_lib(imps, psi)
            envs = imps.DMRGCache(st, H, ctx=c)
            sw = ctypes.c_int()
            c.call("imps_gaugefix_right", 1e-14, 300, ctypes.byref(sw))
            got = c.get_state()
            for a, b in zip(got.AL + got.AR + got.C + got.AC, psi.AL + psi.AR + psi.C + psi.AC):
                assert rel(a, b) < 1e-10
            envs.ket = got
            c.call("imps_envs_recalculate", 1e-10, 300, 30)
            p2, envs = imps.timestep(got, H, -0.01j, imps.TDVP(), envs)
            assert abs(imps.norm(p2) - 1) < 1e-8


def test_complex_tdvp_step_matches_oracle_at_d24(imps, oracle, cx):
    # a larger complex state (XXZ, two-site cell, D = 24, embedded 48): two real-time steps against
    # the dense complex restatement
    W = oracle.xxz_mpo(1.0, 0.7)
    Ws = [W, W]
    psi = cx.canonical_from_tensors(_complex_raw(oracle, 24, [2, 2], 51))
    env = cx.dmrg_cache(psi, Ws)
    refs = []
    p = psi
    for _ in range(2):
        p, env = cx.tdvp_timestep(p, Ws, env, -0.02j)
        refs.append((p, cx.expectation_energy(p, Ws, env).real))
    print("complex: two real-time steps at D=24", flush=True)
    H = imps.SparseIMPO(Ws)
    with imps.Context(0) as c:
        st = _as_lib(imps, psi)
        envs = imps.DMRGCache(st, H, tol=1e-12, ctx=c)
        for ref, e_ref in refs:
            st, envs = imps.timestep(st, H, -0.02j, imps.TDVP(tol=1e-12, gauge_tol=1e-14, gauge_maxiter=2000), envs)
            for a, b in zip(st.AL + st.C, ref.AL + ref.C):
                assert rel(a, b) < 1e-7
            assert abs(imps.expectationvalue(st, H, envs) - e_ref) < 1e-9


def test_complex_primitives_match_oracle(imps, oracle, cx):
    # the boundary primitives on a complex state (ComplexF64 arrays in and out): environments,
    # H_AC, H_C, the transfer matrix and the lowest H_AC eigenpair against the dense complex
    # restatement
    W = oracle.xxz_mpo(1.0, 0.8)
    Ws = [W, W]
    psi = cx.canonical_from_tensors(_complex_raw(oracle, 6, [2, 2], 61))
    env = cx.dmrg_cache(psi, Ws)
    rng = np.random.default_rng(3)
    print("complex: primitives at D=6", flush=True)
    H = imps.SparseIMPO(Ws)
    with imps.Context(0) as c:
        st = _as_lib(imps, psi)
        envs = imps.DMRGCache(st, H, tol=1e-13, ctx=c)
        for l in range(2):
            assert rel(imps.leftenv(envs, l), env.leftenv(l)) < 1e-9
            assert rel(imps.rightenv(envs, l), env.rightenv(l)) < 1e-9
        x = rng.standard_normal((6, 2, 6)) + 1j * rng.standard_normal((6, 2, 6))
        y = imps.AC_hamiltonian(0, st, H, st, envs)(x)
        assert np.iscomplexobj(y) and rel(y, cx.hac_apply(env.leftenv(0), W, env.rightenv(0), x)) < 1e-12
        xc = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
        yc = imps.C_hamiltonian(0, st, H, st, envs)(xc)
        assert rel(yc, cx.hc_apply(env.leftenv(1), env.rightenv(0), xc)) < 1e-12
        v = rng.standard_normal(36) + 1j * rng.standard_normal(36)
        tv = imps.TransferMatrix(envs, "left")(v)
        assert rel(tv, cx.transfer_left_matrix(psi.AL, psi.AL) @ v) < 1e-12
        lam, vec, info = imps.fixedpoint(imps.AC_hamiltonian(0, st, H, st, envs), psi.AC[0], "SR",
                                         imps.Lanczos(tol=1e-12))
        Hd = cx.dense_operator(lambda z: cx.hac_apply(env.leftenv(0), W, env.rightenv(0), z), (6, 2, 6))
        w, V = np.linalg.eigh(0.5 * (Hd + Hd.conj().T))
        assert abs(lam - w[0]) < 1e-9
        vv = vec.reshape(-1, order="F")
        assert abs(np.linalg.norm(vv) - 1) < 1e-10 and abs(abs(np.vdot(V[:, 0], vv)) - 1) < 1e-8



def test_complex_operator_expectation_values_match_oracle(imps, oracle, cx):
    # expectationvalue(ψ, site => O) (expval.jl:33-37) for complex O: Sy, a random Hermitian and a
    # random non-Hermitian operator, on a complex state and on a real one, against the complex
    # restatement (the Hermitian values real to rounding)
    W = oracle.xxz_mpo(1.0, 0.8)
    Ws = [W, W]
    H = imps.SparseIMPO(Ws)
    rng = np.random.default_rng(3)
    G = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    ops = {"sy": np.array([[0, -0.5j], [0.5j, 0]]), "hermitian": G + G.conj().T, "general": G}
    psi = cx.canonical_from_tensors(_complex_raw(oracle, 5, [2, 2], 77))
    rpsi = oracle.randomimps(4, [2, 2], 5)
    print("complex: one-site expectation values of complex operators", flush=True)
    for state, ref in ((psi, psi), (rpsi, cx.from_real(rpsi))):
        with imps.Context(0) as c:
            st = _as_lib(imps, state)
            envs = imps.DMRGCache(st, H, ctx=c)
            for site in (0, 1):
                for name, O in ops.items():
                    want = cx.local_expectation1(ref, site, O)
                    got = imps.expectationvalue(st, (site, O), envs)
                    assert isinstance(got, complex)
                    assert abs(got - want) < 1e-12 * max(1.0, abs(want)), (name, site, got, want)
                    if name != "general":
                        assert abs(got.imag) < 1e-13


def test_complex_operator_correlators_match_oracle(imps, oracle, cx):
    # correlator(ψ, O1, O2, i, js) (correlators.jl:13-42) with complex operators (Sy Sy, random
    # complex ones) on a complex and on a real state, against the complex restatement
    W = oracle.xxz_mpo(1.0, 0.8)
    H = imps.SparseIMPO([W, W])
    rng = np.random.default_rng(11)
    G1 = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    G2 = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    sy = np.array([[0, -0.5j], [0.5j, 0]])
    pairs = [(sy, sy), (G1, G2), (sy, oracle.sigma_z())]
    psi = cx.canonical_from_tensors(_complex_raw(oracle, 5, [2, 2], 91))
    rpsi = oracle.randomimps(4, [2, 2], 6)
    print("complex: correlators of complex operators", flush=True)
    for state, ref in ((psi, psi), (rpsi, cx.from_real(rpsi))):
        with imps.Context(0) as c:
            st = _as_lib(imps, state)
            envs = imps.DMRGCache(st, H, ctx=c)
            for O1, O2 in pairs:
                for i, js in ((0, [1, 2, 5]), (1, [2, 3])):
                    want = cx.correlator(ref, O1, O2, i, js)
                    got = imps.correlator(st, O1, O2, i, js, envs)
                    assert np.iscomplexobj(got)
                    assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, float(np.max(np.abs(want)))), (i, js)
                single = imps.correlator(st, O1, O2, 0, 2, envs)
                assert isinstance(single, complex)
                assert abs(single - cx.correlator(ref, O1, O2, 0, [2])[0]) < 1e-12


def test_complex_dot_and_fidelity(imps, oracle, cx):
    # dot(ψ₁, ψ₂) = <ψ₁|ψ₂> (canonicalmps.jl:262-269) for complex states: 1 for a normalised state
    # with itself, e^{iφ} per unit cell against the same state with one tensor times e^{iφ} (fixes
    # the conjugated layer), the dense transfer matrix's dominant eigenvalue for two random states,
    # and a real state promoted against its complex copy
    W = oracle.xxz_mpo(1.0, 0.8)
    H = imps.SparseIMPO([W, W])
    psi = cx.canonical_from_tensors(_complex_raw(oracle, 4, [2, 2], 101))
    psi2 = cx.canonical_from_tensors(_complex_raw(oracle, 4, [2, 2], 202))
    phi = 0.7
    rot = imps.CanonicalIMPS([psi.AL[0] * np.exp(1j * phi), psi.AL[1].copy()], [x.copy() for x in psi.AR],
                             [x.copy() for x in psi.C], [x.copy() for x in psi.AC])
    w = np.linalg.eigvals(cx.transfer_left_matrix(psi2.AL, psi.AL))
    want = w[int(np.argmax(np.abs(w)))]
    rpsi = oracle.randomimps(4, [2, 2], 9)
    print("complex: dot and fidelity of complex states", flush=True)
    with imps.Context(0) as c:
        st = _as_lib(imps, psi)
        envs = imps.DMRGCache(st, H, ctx=c)
        d11 = imps.dot(st, st, envs)
        assert isinstance(d11, complex) and abs(d11 - 1.0) < 1e-10
        assert abs(imps.dot(st, rot, envs) - np.exp(1j * phi)) < 1e-10
        d12 = imps.dot(st, _as_lib(imps, psi2), envs)
        assert abs(d12 - want) < 1e-9 * max(1.0, abs(want)), (d12, want)
        assert abs(imps.fidelity(st, _as_lib(imps, psi2), envs) - abs(want)) < 1e-9
    with imps.Context(0) as c:
        rs = imps.CanonicalIMPS(rpsi.AL, rpsi.AR, rpsi.C, rpsi.AC)
        envs = imps.DMRGCache(rs, H, ctx=c)
        assert abs(imps.dot(rs, rs.complex(), envs) - 1.0) < 1e-10
