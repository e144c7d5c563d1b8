"""MI355X (gfx950) VUMPS inner loop of InfiniteMPSAlgorithms.jl behind a C ABI.

The numerics live in ``libimps.so`` (HIP kernels + C++ runtime); this package is the
host-side mirror of the reference's API (``find_groundstate``, ``VUMPS``,
``AC_hamiltonian``, ``C_hamiltonian``, ``DMRGCache``, ``leftenv``/``rightenv``,
``TransferMatrix``, ...).
"""
from . import _lib
from .api import (IDMRG, TDVP, VUMPS, DynamicTol, SolverTol, alg_eigsolve, alg_environments, alg_gauge, correlator, load_mpo, load_state, save_mpo, save_state, dense_expectation, dominant_env, DenseIMPO, dot, fidelity, infidelity, norm, AC_hamiltonian, C_hamiltonian, CanonicalIMPS, Context, DMRGCache,
                  GPUCache, IterativeConvergenceInfo, Lanczos, SparseIMPO, TransferMatrix,
                  calc_galerkin, canonicalimps, energy_gradient, fixedpoint_gradient, objective_mpo_1site,
                  tangent_hessian_apply, expectationvalue, find_groundstate,
                  find_groundstates_batched, fixedpoint,
                  leftenv, random_tensors, randomimps, recalculate_, regauge_, regularize_,
                  rightenv, time_evolve, timestep)
from . import models

__all__ = [
    "DynamicTol", "SolverTol", "alg_eigsolve", "alg_environments", "alg_gauge", "correlator", "load_mpo", "load_state", "save_mpo", "save_state", "dense_expectation", "dominant_env", "DenseIMPO", "dot", "fidelity", "infidelity", "norm", "IDMRG", "TDVP", "time_evolve", "timestep", "VUMPS", "AC_hamiltonian", "C_hamiltonian", "CanonicalIMPS", "Context", "DMRGCache", "GPUCache",
    "IterativeConvergenceInfo", "Lanczos", "SparseIMPO", "TransferMatrix", "calc_galerkin",
    "canonicalimps", "energy_gradient", "fixedpoint_gradient", "objective_mpo_1site", "tangent_hessian_apply", "expectationvalue", "find_groundstate", "find_groundstates_batched", "fixedpoint", "leftenv", "random_tensors",
    "randomimps", "recalculate_", "regauge_", "regularize_", "rightenv", "models",
]
