"""The lockstep device batch (imps_vumps_batch_device) against single runs, the threaded batch
and closed forms, through the C ABI.

  TFIM (w = 3) and XXZ (w = 5, interior levels with zero diagonal blocks) one-site instances with
  uneven bond dimensions: every instance's energy equals its own single find_groundstate run
  within 1e-10 (the same fixed point; the device batch's QRs are Householder where the single
  runs use Cholesky QR, so the trajectories agree only to rounding), the TFIM energies the
  free-fermion closed form at gapped fields.
"""
import math

import numpy as np
import pytest
from scipy import integrate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def imps():
    import infinitempsalgorithms_amd as m
    return m


def tfim_energy(h):
    return -integrate.quad(lambda k: math.sqrt(1 + h * h - 2 * h * math.cos(k)), 0, math.pi, epsabs=1e-14,
                           epsrel=1e-14)[0] / math.pi


def _single(imps, H, As, alg):
    psi0, c = imps.canonicalimps(As, H, tol=1e-12)
    try:
        envs = imps.GPUCache(psi0, H, c)
        c.call("imps_envs_reset")
        _, _, info = imps.find_groundstate(psi0, H, alg, envs=envs)
        return info
    finally:
        c.close()


def test_device_batch_tfim_matches_single_runs_and_closed_form(imps):
    cases = [(0.6, 16), (1.4, 32), (0.5, 48), (1.5, 16), (0.8, 32), (1.3, 64)]
    alg = imps.VUMPS(tol=1e-10, maxiter=300)
    problems = [(imps.models.tfim_hamiltonian(J=1.0, h=h, nsites=1), imps.random_tensors(D, [2], seed=100 + k))
                for k, (h, D) in enumerate(cases)]
    assert imps.api.device_batch_eligible(problems, alg)
    stats = {}
    res = imps.find_groundstates_batched(problems, alg, mode="device", stats=stats)
    print("device batch", stats, flush=True)
    assert stats["mode"] == "device" and stats["rounds"] > 0
    for (h, D), (H, As), r in zip(cases, problems, res):
        info = _single(imps, H, As, alg)
        print(h, D, r["niter"], info.niter, r["energy"], info.energy, r["itererr"], flush=True)
        assert r["converged"] and r["itererr"] <= 1e-10
        assert abs(r["energy"] - info.energy) < 1e-10, (h, D, r["energy"], info.energy)
        if D >= 32:
            assert abs(r["energy"] - tfim_energy(h)) < 1e-8, (h, D)


def test_device_batch_xxz_matches_threaded_batch(imps, oracle):
    cases = [(0.5, 16), (0.7, 32), (0.9, 16), (0.6, 48)]
    alg = imps.VUMPS(tol=1e-9, maxiter=400)
    problems = [(imps.SparseIMPO([oracle.xxz_mpo(1.0, dl)]), imps.random_tensors(D, [2], seed=7 + k))
                for k, (dl, D) in enumerate(cases)]
    rd = imps.find_groundstates_batched(problems, alg, mode="device")
    rt = imps.find_groundstates_batched(problems, alg, mode="threads", nworkers=4)
    for c, a, b in zip(cases, rd, rt):
        print(c, a["niter"], b["niter"], a["energy"], b["energy"], flush=True)
        assert a["converged"] and b["converged"]
        assert abs(a["energy"] - b["energy"]) < 1e-10, (c, a["energy"], b["energy"])


def test_device_batch_refuses_unsupported(imps, oracle):
    W = oracle.tfim_mpo(1.0, 1.2)
    two_site = [(imps.SparseIMPO([W, W]), imps.random_tensors(16, [2, 2], seed=1))]
    odd_D = [(imps.SparseIMPO([W]), imps.random_tensors(12, [2], seed=1))]
    for probs in (two_site, odd_D):
        assert not imps.api.device_batch_eligible(probs)
        with pytest.raises(imps._lib.ImpsError, match="status -5"):
            imps.find_groundstates_batched(probs, imps.VUMPS(), mode="device")
