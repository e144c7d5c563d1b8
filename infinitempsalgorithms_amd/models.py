"""Model constructors (the reference's ``Models`` submodule, src/models.jl:12-180).

Each returns a :class:`~infinitempsalgorithms_amd.api.SparseIMPO` whose site tensors are
the dense Schur tensors ``W[wl, u, wr, d]`` of ``tompotensor`` (u = output index).
Only real (Float64) operators are built: the XXZ chain uses the S+- form and the
Fermi-Hubbard chain uses Hermitian Jordan-Wigner hopping channels.
"""
from __future__ import annotations

import numpy as np

from types import SimpleNamespace

from .api import DenseIMPO, SparseIMPO


def sigma_x():
    """src/models.jl:64."""
    return np.array([[0.0, 1.0], [1.0, 0.0]])


def sigma_z():
    """src/models.jl:66."""
    return np.array([[1.0, 0.0], [0.0, -1.0]])


def Sx():
    return sigma_x() / 2


def Sz():
    return sigma_z() / 2


def mpohamiltonian(h1, pairs):
    """Schur level tensor of an on-site + nearest-neighbour Hamiltonian
    (src/models.jl:82-99): levels 1 and n are unit levels, [1, n] = h1,
    channel k: [1, k+1] = coeff_k a_k, [k+1, n] = b_k.  Returns the dense (w, d, w, d)
    tensor (FiniteMPSAlgorithms' tompotensor layout)."""
    h1 = np.asarray(h1, dtype=np.float64)
    d = h1.shape[0]
    n = len(pairs) + 2
    W = np.zeros((n, d, n, d), order="F")
    I = np.eye(d)
    W[0, :, 0, :] = I
    W[n - 1, :, n - 1, :] = I
    W[0, :, n - 1, :] = h1
    for k, (coeff, a, b) in enumerate(pairs):
        W[0, :, k + 1, :] = coeff * np.asarray(a, dtype=np.float64)
        W[k + 1, :, n - 1, :] = np.asarray(b, dtype=np.float64)
    return W


def bulk_mpo(h1, pairs):
    """Schur bulk tensor of an on-site + nearest-neighbour Hamiltonian (src/models.jl:35-60):
    the cell of unit levels, h1 in the corner and one channel per pair, as the dense
    (w, d, w, d) tensor that tompotensor gives (the same tensor as mpohamiltonian's)."""
    return mpohamiltonian(h1, pairs)


def tfim_hamiltonian(J=1.0, h=1.0, nsites=1):
    """Transverse-field Ising model H = -J sum sx sx - h sum sz (src/models.jl:119-122)."""
    W = mpohamiltonian(-h * sigma_z(), [(-J, sigma_x(), sigma_x())])
    return SparseIMPO([W] * nsites)


def heisenberg_hamiltonian(J=1.0, Delta=1.0, h=0.0, nsites=1):
    """H = J sum (SxSx + SySy + Delta SzSz) - h sum Sz (src/models.jl:107-111), real form:
    SxSx + SySy = (S+S- + S-S+)/2, w = 5."""
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    sm = sp.T.copy()
    W = mpohamiltonian(-h * Sz(), [(J / 2, sp, sm), (J / 2, sm, sp), (J * Delta, Sz(), Sz())])
    return SparseIMPO([W] * nsites)


def hubbard_local_ops():
    """c_up, c_dn, Jordan-Wigner parity F and the number operators on the local basis
    (|0>, |up>, |dn>, |updn> = c+_up c+_dn |0>)."""
    c_up = np.zeros((4, 4))
    c_dn = np.zeros((4, 4))
    c_up[0, 1] = 1.0
    c_up[2, 3] = 1.0
    c_dn[0, 2] = 1.0
    c_dn[1, 3] = -1.0
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    return c_up, c_dn, F, c_up.T @ c_up, c_dn.T @ c_dn


def fermi_hubbard(t=1.0, U=4.0, mu=2.0, nsites=1):
    """H = -t sum_s (c+_{i,s} c_{i+1,s} + h.c.) + U sum n_up n_dn - mu sum n
    (src/models.jl:164-178 with Hermitian Jordan-Wigner channels, w = 6; mu = U/2 is
    half filling)."""
    c_up, c_dn, F, n_up, n_dn = hubbard_local_ops()
    h1 = U * (n_up @ n_dn) - mu * (n_up + n_dn)
    pairs = [(-t, c_up.T @ F, c_up), (-t, c_dn.T @ F, c_dn), (-t, F @ c_up, c_up.T), (-t, F @ c_dn, c_dn.T)]
    return SparseIMPO([mpohamiltonian(h1, pairs)] * nsites)


def tfim(J=1.0, h=1.0):
    """tfim(; J, h) (src/models.jl:150-155): (mpo, bulk, hamiltonian) with mpo the one-site
    periodic DenseIMPO of the Schur bulk, bulk that tensor and hamiltonian the SparseIMPO of
    tfim_hamiltonian (the reference's note: energies come from the SparseIMPO)."""
    W = bulk_mpo(-h * sigma_z(), [(-J, sigma_x(), sigma_x())])
    return SimpleNamespace(mpo=DenseIMPO([W]), bulk=W, hamiltonian=tfim_hamiltonian(J=J, h=h))


def heisenberg_xxz(J=1.0, Delta=1.0, h=0.0):
    """heisenberg_xxz(; J, Delta, h) (src/models.jl:135-139): (mpo, bulk, hamiltonian) as for
    tfim.  The reference's bulk takes the channels Sx Sx, Sy Sy, Sz Sz (complex, through Sy);
    this real library takes the S+- form of the same operator with the same 5 levels."""
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    sm = sp.T.copy()
    W = bulk_mpo(-h * Sz(), [(J / 2, sp, sm), (J / 2, sm, sp), (J * Delta, Sz(), Sz())])
    return SimpleNamespace(mpo=DenseIMPO([W]), bulk=W, hamiltonian=heisenberg_hamiltonian(J=J, Delta=Delta, h=h))
