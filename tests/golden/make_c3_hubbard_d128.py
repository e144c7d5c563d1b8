"""Writes c3_hubbard_d128_groundstate.json: the oracle's converged VUMPS ground state of the JW
Fermi-Hubbard chain (t = 1, U = 4, mu = 2, N = 2, d = 4, w = 6) at D = 128 from randomimps(128,
[4, 4]) with PCG64(2), tol 1e-8.  On this model VUMPS at a fixed D may settle into a limit cycle
instead of converging, and whether it does depends on rounding along the trajectory; at D = 128
the FP64, FP32-bulk and split-BF16-bulk device runs all converge.  The file holds:
iterations, the final Galerkin error, the energy per site and per-site local densities
<n_up + n_dn>, <n_up n_dn>, <S^z> = <n_up - n_dn> / 2 and <S^x> = <c+_up c_dn + c+_dn c_up> / 2.
The state at this D breaks the spin symmetry; rotations about the y axis are real, so other real
fixed points of the same energy exist and only S^z^2 + S^x^2 is fixed among them.

Run from the repository root (about 16 minutes on 6 OpenBLAS threads):
    python tests/golden/make_c3_hubbard_d128.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import imps_oracle as O  # noqa: E402


def main():
    W = O.hubbard_mpo(1.0, 4.0, 2.0)
    psi0 = O.randomimps(128, [4, 4], 2)
    psi, envs, info = O.find_groundstate(psi0, [W, W], O.VUMPS(tol=1e-8, maxiter=400))
    c_up, c_dn, _, n_up, n_dn = O.hubbard_local_ops()
    sp = c_up.T @ c_dn
    ops = {"density": n_up + n_dn, "double_occupancy": n_up @ n_dn, "sz": 0.5 * (n_up - n_dn),
           "sx": 0.5 * (sp + sp.T)}
    out = {"config": "JW Hubbard t=1 U=4 mu=2, N=2, d=4, w=6, D=128, psi0 randomimps PCG64(2), VUMPS tol 1e-8",
           "niter": info.niter, "converged": bool(info.converged), "itererr": float(info.losses[-1]),
           "energy_per_site": float(info.energies[-1] / 2),
           "local": {k: [float(O.local_expectation1(psi, s, M)) for s in range(2)] for k, M in ops.items()}}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c3_hubbard_d128_groundstate.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
