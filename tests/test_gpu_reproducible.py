"""Run-to-run reproducibility on the device: the library has no floating-point atomics; every
reduction (GEMM split-K sums, Gram-Schmidt and trace partials, norms) adds its terms in a fixed
order, the GEMM tile variants the autotuner picks among sum over K in the same order, and the
captured gauge sweeps replay the eager kernel sequence.  So the same inputs give the same bits:
these tests run each algorithm twice in fresh contexts (and once more with the autotuner and graph
capture off) and compare states, Galerkin histories and energies with exact equality."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def imps():
    import infinitempsalgorithms_amd as m
    return m


def _vumps_run(imps, oracle, W, N, D, alg, seed):
    H = imps.SparseIMPO([W] * N)
    psi = oracle.randomimps(D, [W.shape[1]] * N, seed)
    with imps.Context(0) as c:
        st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
        g, envs, info = imps.find_groundstate(st, H, alg, envs=imps.GPUCache(st, H, c))
        return g, list(info.losses), info.energy


def _same_state(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a.AL + a.AR + a.C + a.AC, b.AL + b.AR + b.C + b.AC))


@pytest.mark.parametrize("case", ["fp64-xxz-N2-D32", "split-bf16-hubbard-N1-D128"])
def test_vumps_is_bit_reproducible(imps, oracle, case):
    lib = imps._lib.load()
    if case.startswith("fp64"):
        W, N, D = oracle.xxz_mpo(1.0, 0.7), 2, 32
        alg = imps.VUMPS(tol=1e-14, maxiter=12)
    else:  # the split-BF16 planes phase throughout (refinement only below 1e-12)
        W, N, D = oracle.hubbard_mpo(1.0, 4.0, 2.0), 1, 128
        alg = imps.VUMPS(tol=1e-14, maxiter=8, prec=2, refine_below=1e-12)
    print(f"reproducible: VUMPS {case}", flush=True)
    runs = [_vumps_run(imps, oracle, W, N, D, alg, 5) for _ in range(2)]
    try:
        assert lib.imps_set_gemm_autotune(0) == 0 and lib.imps_set_graph_capture(0) == 0
        runs.append(_vumps_run(imps, oracle, W, N, D, alg, 5))
    finally:
        lib.imps_set_gemm_autotune(1)
        lib.imps_set_graph_capture(1)
    g0, l0, e0 = runs[0]
    assert len(l0) >= 8
    for g, l, e in runs[1:]:
        assert l == l0 and e == e0
        assert _same_state(g, g0)


def test_real_time_tdvp_is_bit_reproducible(imps, oracle):
    W = oracle.xxz_mpo(1.0, 1.0)
    H = imps.SparseIMPO([W, W])
    psi = oracle.randomimps(12, [2, 2], 3)
    print("reproducible: three real-time TDVP steps, XXX D=12", flush=True)
    outs = []
    for _ in range(2):
        with imps.Context(0) as c:
            st = imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC)
            envs = imps.DMRGCache(st, H, ctx=c)
            for _ in range(3):
                st, envs = imps.timestep(st, H, -0.02j, imps.TDVP(), envs)
            outs.append((st, imps.expectationvalue(st, H, envs)))
    assert outs[0][0].is_complex()
    assert _same_state(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]


@pytest.mark.parametrize("mode", ["device", "threads"])
def test_batched_sweep_is_bit_reproducible(imps, mode):
    # the threaded batch deals instances to worker threads dynamically and the device batch
    # groups them; each instance's result is the same bits in every run
    alg = imps.VUMPS(tol=1e-10, maxiter=200)
    problems = [(imps.models.tfim_hamiltonian(J=1.0, h=h, nsites=1), imps.random_tensors(D, [2], seed=40 + k))
                for k, (h, D) in enumerate([(0.7, 16), (1.3, 32), (0.9, 48), (1.1, 16)])]
    print(f"reproducible: batched sweep ({mode}) of 4 instances, twice", flush=True)
    r1 = imps.find_groundstates_batched(problems, alg, mode=mode, nworkers=4)
    r2 = imps.find_groundstates_batched(problems, alg, mode=mode, nworkers=4)
    for a, b in zip(r1, r2):
        assert a["converged"] and a["niter"] == b["niter"]
        assert a["energy"] == b["energy"] and a["itererr"] == b["itererr"]
