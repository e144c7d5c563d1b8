"""Writes c2_xxz_d256_trajectory.json: the oracle's VUMPS trajectory at the XXZ configuration
(Delta = 1, real S+- MPO, N = 2, d = 2, D = 256, psi0 = randomimps from PCG64(1)), 14 iterations
at tol 1e-14 -- the Galerkin error after each iteration (losses[0] is the initial one) and the
energy per site.  Two more runs from psi0's raw tensors perturbed at the rounding level (every
entry times 1 + 1e-15, and times 1 + 1e-15 N(0,1) with PCG64(123)) record how fast the trajectory
of this configuration amplifies a rounding-level difference: the relative split of the Galerkin
errors per iteration and the largest energy difference.

Run from the repository root (takes about 4 minutes per run on 2 OpenBLAS threads):
    python tests/golden/make_c2_trajectory.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import imps_oracle as O  # noqa: E402

ITERS = 14


def run(mode):
    W = O.xxz_mpo(1.0, 1.0)
    As = O.random_tensors(256, [2, 2], 1)
    if mode == "scale":
        As = [a * (1 + 1e-15) for a in As]
    elif mode == "noise":
        rng = np.random.default_rng(123)
        As = [a * (1 + 1e-15 * rng.standard_normal(a.shape)) for a in As]
    psi0 = O.canonical_from_tensors(As)
    _, _, info = O.find_groundstate(psi0, [W, W], O.VUMPS(tol=1e-14, maxiter=ITERS))
    return np.array(info.losses), np.array(info.energies) / 2


def main():
    base_l, base_e = run("base")
    out = {"config": "XXZ Delta=1 (S+- form, w=5), N=2, d=2, D=256, psi0 randomimps PCG64(1), VUMPS tol 1e-14",
           "losses": base_l.tolist(), "energy_per_site": base_e.tolist(), "perturbed": {}}
    for mode in ("scale", "noise"):
        l, e = run(mode)
        out["perturbed"][mode] = {"rel_loss_split": (np.abs(l - base_l) / base_l).tolist(),
                                  "max_abs_energy_diff": float(np.max(np.abs(e - base_e)))}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_xxz_d256_trajectory.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
