"""Host-side mirror of InfiniteMPSAlgorithms.jl's VUMPS API on top of libimps.

Names, argument meaning and error behaviour follow the reference (Julia ``!``
functions become a trailing underscore).  All numerics run in the HIP library; the
Python objects only hold host copies of tensors and the device context.

Index conventions are the reference's column-major ones: MPS ``A[a, s, b]``, bond
matrices ``C[a, b]`` (bond right of a site), MPO ``W[wl, u, wr, d]``, environments
``GL[bra, w, ket]`` and ``GR[ket, w, bra]``.  Sites are 0-based.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check, dptr, fortran, ptr_array, zfortran, zptr_array


# Contexts still open when the interpreter exits are closed by an atexit hook, which runs before
# the interpreter tears modules down and before the C runtime's exit handlers (the HIP runtime's and
# a profiler's among them), so no stream, event or pinned buffer of the library is released from a
# finaliser after the runtime has begun to shut down.
_LIVE = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    for c in list(_LIVE):
        try:
            c.close()
        except Exception:
            pass


class Context:
    """One libimps context (a HIP stream plus device-resident state) on `device`.  Close it with
    close() (or use it as a context manager); contexts left open are closed at interpreter exit."""

    cplx = False  # the context holds a complex state
    chi = None    # bond dimensions (bond l right of site l) when they differ from bond to bond

    def __init__(self, device=0):
        L = _lib.load()
        h = ctypes.c_void_p()
        st = L.imps_create(device, ctypes.byref(h))
        if st != 0:
            raise _lib.ImpsError(f"imps_create failed ({st}): {L.imps_last_error(None).decode()}")
        self.h = h
        self.L = L
        self.N = self.w = self.d = self.D = 0
        self.dims, self.levels = [], []
        self.cplx = False  # the context holds a complex state
        self.chi = None    # bond dimensions (bond l right of site l) when they differ from bond to bond
        _LIVE.add(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def close(self):
        if self.h is not None and self.h.value:
            self.L.imps_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        check(self.h, getattr(self.L, name)(self.h, *args), name)

    # ---- model / state
    def set_mpo(self, H: "SparseIMPO"):
        Ws = [fortran(W) for W in H.Ws]
        pp, keep = ptr_array(Ws)
        if H.square:
            self.call("imps_set_mpo", len(Ws), H.w, H.d, pp)
        else:
            dims = (ctypes.c_int * len(Ws))(*H.dims)
            levels = (ctypes.c_int * len(Ws))(*H.levels)
            self.call("imps_set_mpo_chain", len(Ws), dims, levels, pp)
        self.N, self.w, self.d = len(Ws), H.w, H.d
        self.dims, self.levels = list(H.dims), list(H.levels)

    # The library runs on the square embedding of the model (imps_set_mpo_chain): states carry
    # d = max dims physical indices with zeros in the empty sectors, environments w = max levels.
    # These map the reference's per-site shapes to it and back.
    def pad_site(self, l, x):
        """(Dl, dims[l], Dr) -> (D, d, D) with zeros in the empty physical sector and outside the
        site's active bond block (complex arrays stay complex)."""
        x = zfortran(x) if np.iscomplexobj(x) else fortran(x)
        D1, D2 = (self.D, self.D) if self.chi else (x.shape[0], x.shape[2])
        if (not self.d or x.shape[1] == self.d) and x.shape[0] == D1 and x.shape[2] == D2:
            return x
        out = np.zeros((D1, self.d or x.shape[1], D2), dtype=x.dtype, order="F")
        out[:x.shape[0], :x.shape[1], :x.shape[2]] = x
        return out

    def unpad_site(self, l, x):
        dl = self.dims[l % self.N] if self.d else x.shape[1]
        if self.chi:
            N = len(self.chi)
            return np.asfortranarray(x[:self.chi[(l - 1) % N], :dl, :self.chi[l % N]])
        return x if dl == x.shape[1] else np.asfortranarray(x[:, :dl, :])

    def pad_bond(self, l, x):
        """A bond matrix (chi[l], chi[l]) of the bond right of site l as (D, D)."""
        x = zfortran(x) if np.iscomplexobj(x) else fortran(x)
        D = self.D if self.chi else x.shape[0]
        if x.shape == (D, D):
            return x
        out = np.zeros((D, D), dtype=x.dtype, order="F")
        out[:x.shape[0], :x.shape[1]] = x
        return out

    def unpad_bond(self, l, x):
        if not self.chi:
            return x
        c = self.chi[l % len(self.chi)]
        return np.asfortranarray(x[:c, :c])

    def pad_operator(self, O):
        """A one-site operator (dl x dl) as its (d x d) form, zero on the empty sector."""
        O = fortran(O)
        if O.shape[0] == self.d:
            return O
        out = np.zeros((self.d, self.d), order="F")
        out[:O.shape[0], :O.shape[1]] = O
        return out

    def bond_levels(self, bond):
        """Embedded level indices of the bond right of site `bond` (0 .. levels-2, then w-1)."""
        L = self.levels[bond % self.N]
        return list(range(L - 1)) + [self.w - 1]

    def _check_state_shapes(self, what, D, tensors):
        if not self.d:
            return  # without a model the library itself refuses the call
        for l, x in enumerate(tensors):
            if x.shape != (D, self.dims[l % self.N], D):
                raise ValueError(f"{what}: site {l} tensor must be ({D}, {self.dims[l % self.N]}, {D}), got {x.shape}")

    def set_state(self, psi: "CanonicalIMPS"):
        if not psi.uniform_bonds():
            return self._set_state_nonuniform(psi)
        if psi.is_complex():
            return self._set_state_complex(psi)
        self.chi = None
        AL, AR, C, AC = ([fortran(x) for x in xs] for xs in (psi.AL, psi.AR, psi.C, psi.AC))
        D = psi.D
        if self.d:
            if not (len(AL) == len(AR) == len(C) == len(AC)) or any(x.shape != (D, D) for x in C):
                raise ValueError(f"set_state: AL, AR, C, AC per site with every C ({D}, {D})")
            for xs in (AL, AR, AC):
                self._check_state_shapes("set_state", D, xs)
            AL, AR, AC = ([self.pad_site(l, x) for l, x in enumerate(xs)] for xs in (AL, AR, AC))
        keep = [ptr_array(x) for x in (AL, AR, C, AC)]
        self.call("imps_set_state", len(AL), psi.D, *(k[0] for k in keep))
        self.D = psi.D
        self.cplx = False
        self.chi = None

    def _set_state_nonuniform(self, psi: "CanonicalIMPS"):
        # bond dimensions that differ from bond to bond: every tensor at the largest one with
        # zeros outside its active block (imps_set_state_nonuniform; complex states:
        # imps_set_state_complex_nonuniform)
        chi = psi.bonddims()
        cplx = psi.is_complex()
        N = len(psi)
        for l in range(N):
            want = (chi[(l - 1) % N], self.dims[l % self.N] if self.d else psi.AL[l].shape[1], chi[l])
            for xs in (psi.AL, psi.AR, psi.AC):
                if xs[l].shape != want:
                    raise ValueError(f"set_state: site {l} tensor must be {want}, got {xs[l].shape}")
            if psi.C[l].shape != (chi[l], chi[l]):
                raise ValueError(f"set_state: C[{l}] must be ({chi[l]}, {chi[l]})")
        D = max(chi)
        self.D, self.chi = D, list(chi)
        conv = zfortran if cplx else fortran
        AL, AR, AC = ([self.pad_site(l, conv(x)) for l, x in enumerate(xs)] for xs in (psi.AL, psi.AR, psi.AC))
        C = [self.pad_bond(l, conv(x)) for l, x in enumerate(psi.C)]
        keep = [(zptr_array if cplx else ptr_array)(x) for x in (AL, AR, C, AC)]
        carr = (ctypes.c_int * N)(*chi)
        try:
            self.call("imps_set_state_complex_nonuniform" if cplx else "imps_set_state_nonuniform", N, D, carr,
                      *(k[0] for k in keep))
        except Exception:
            self.chi = None
            raise
        self.cplx = cplx

    def _set_state_complex(self, psi: "CanonicalIMPS"):
        # CanonicalIMPS{ComplexF64}: the library holds the exact real embedding (imps_set_state_complex)
        self.chi = None
        AL, AR, C, AC = ([zfortran(x) for x in xs] for xs in (psi.AL, psi.AR, psi.C, psi.AC))
        D = psi.D
        if self.d:
            if not (len(AL) == len(AR) == len(C) == len(AC)) or any(x.shape != (D, D) for x in C):
                raise ValueError(f"set_state: AL, AR, C, AC per site with every C ({D}, {D})")
            for xs in (AL, AR, AC):
                self._check_state_shapes("set_state", D, xs)
            AL, AR, AC = ([self.pad_site(l, x) for l, x in enumerate(xs)] for xs in (AL, AR, AC))
        keep = [zptr_array(x) for x in (AL, AR, C, AC)]
        self.call("imps_set_state_complex", len(AL), D, *(k[0] for k in keep))
        self.D = D
        self.cplx = True

    def set_state_raw(self, As, tol, maxiter):
        N = len(As)
        chi = [a.shape[2] for a in As]
        if any(As[l].shape[0] != chi[(l - 1) % N] for l in range(N)):
            raise ValueError("set_state_raw: bond dimensions of neighbouring tensors differ")
        if len(set(chi)) > 1:
            cplx = any(np.iscomplexobj(a) for a in As)
            D = max(chi)
            self.D, self.chi = D, list(chi)
            Ap = [self.pad_site(l, zfortran(a) if cplx else fortran(a)) for l, a in enumerate(As)]
            pp, keep = (zptr_array if cplx else ptr_array)(Ap)
            carr = (ctypes.c_int * N)(*chi)
            try:
                self.call("imps_set_state_raw_complex_nonuniform" if cplx else "imps_set_state_raw_nonuniform",
                          N, D, carr, pp, float(tol), int(maxiter))
            except Exception:
                self.chi = None
                raise
            self.cplx = cplx
            return
        self.chi = None
        cplx = any(np.iscomplexobj(a) for a in As)
        As = [zfortran(a) if cplx else fortran(a) for a in As]
        D = As[0].shape[0]
        self._check_state_shapes("set_state_raw", D, As)
        As = [self.pad_site(l, a) for l, a in enumerate(As)]
        pp, keep = (zptr_array if cplx else ptr_array)(As)
        self.call("imps_set_state_raw_complex" if cplx else "imps_set_state_raw", len(As), As[0].shape[0], pp,
                  float(tol), int(maxiter))
        self.D = As[0].shape[0]
        self.cplx = cplx

    def checkpoint(self, path):
        """Writes the context's state to a state file (imps_checkpoint_state)."""
        self.call("imps_checkpoint_state", os.fsencode(path))

    def resume(self, path):
        """Loads a state file into the context (imps_resume_state); the model must match."""
        self.call("imps_resume_state", os.fsencode(path))
        n, D, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.call("imps_state_dims", ctypes.byref(n), ctypes.byref(D), ctypes.byref(d))
        self.D = D.value
        self.cplx = False
        self.chi = None

    def get_state(self):
        N, D, d = self.N, self.D, self.d
        dt = np.complex128 if self.cplx else np.float64
        AL = [np.zeros((D, d, D), dtype=dt, order="F") for _ in range(N)]
        AR = [np.zeros((D, d, D), dtype=dt, order="F") for _ in range(N)]
        C = [np.zeros((D, D), dtype=dt, order="F") for _ in range(N)]
        AC = [np.zeros((D, d, D), dtype=dt, order="F") for _ in range(N)]
        keep = [(zptr_array if self.cplx else ptr_array)(x) for x in (AL, AR, C, AC)]
        self.call("imps_get_state_complex" if self.cplx else "imps_get_state", *(k[0] for k in keep))
        AL, AR, AC = ([self.unpad_site(l, x) for l, x in enumerate(xs)] for xs in (AL, AR, AC))
        C = [self.unpad_bond(l, x) for l, x in enumerate(C)]
        return CanonicalIMPS(AL, AR, C, AC)

    def _env_bond(self, site, side):
        # leftenv(site) lives on the bond left of the site, rightenv(site) on the bond right of it
        return site - 1 if side == _lib.IMPS_SIDE_LEFT else site

    def get_env(self, site, side):
        out = np.zeros((self.D, self.w, self.D), dtype=np.complex128 if self.cplx else np.float64, order="F")
        self.call("imps_get_env", int(site), int(side), out.ctypes.data_as(_lib.P_double))
        bond = self._env_bond(int(site), int(side))
        lv = self.bond_levels(bond)
        if self.chi:
            c = self.chi[bond % len(self.chi)]
            out = out[:c, :, :c]
        return out if len(lv) == self.w else np.asfortranarray(out[:, lv, :])

    def set_env(self, site, side, G):
        G = fortran(G)
        lv = self.bond_levels(self._env_bond(int(site), int(side)))
        if G.shape[1] != self.w:
            if G.shape[1] != len(lv):
                raise ValueError(f"set_env: the environment has {len(lv)} levels, got {G.shape[1]}")
            Gp = np.zeros((G.shape[0], self.w, G.shape[2]), order="F")
            Gp[:, lv, :] = G
            G = Gp
        self.call("imps_set_env", int(site), int(side), dptr(G))


# ---------------------------------------------------------------------------- data types


class SparseIMPO:
    """SparseIMPO (src/operators/mpohamiltonian.jl:33-66): per-site Schur tensors, stored
    densely as ``tompotensor`` arrays (w_left, d, w_right, d).  The level counts may differ from
    bond to bond (rectangular Schur tensors) and the physical dimension from site to site, as
    long as the level chain closes (Ws[l] right levels == Ws[l+1] left levels, periodically;
    mpohamiltonian.jl:57-66).  The first and the last level of every tensor are unit levels.
    The device runs on the chain's exact square embedding (imps_set_mpo_chain)."""

    def __init__(self, Ws):
        Ws = [fortran(W) for W in Ws]
        if not Ws:
            raise ValueError("Ws must not be empty")
        N = len(Ws)
        for l, W in enumerate(Ws):
            if W.ndim != 4 or W.shape[1] != W.shape[3] or W.shape[0] < 2 or W.shape[2] < 2:
                raise ValueError(f"Ws[{l}] must have shape (w_left, d, w_right, d) with at least two levels, "
                                 f"got {W.shape}")
        for l in range(N):
            if Ws[l].shape[2] != Ws[(l + 1) % N].shape[0]:
                raise ValueError(f"level chain mismatch: Ws[{l}] right levels {Ws[l].shape[2]} vs "
                                 f"Ws[{(l + 1) % N}] left levels {Ws[(l + 1) % N].shape[0]}")
        for W in Ws:
            wl, d, wr, _ = W.shape
            if not (np.array_equal(W[0, :, 0, :], np.eye(d)) and np.array_equal(W[wl - 1, :, wr - 1, :], np.eye(d))):
                raise ValueError("the first and the last level must be unit levels")
        self.Ws = Ws
        self.dims = [W.shape[1] for W in Ws]
        self.levels = [W.shape[2] for W in Ws]  # the bond right of each site
        self.w, self.d = max(self.levels), max(self.dims)
        self.square = all(W.shape == (self.w, self.d, self.w, self.d) for W in Ws)

    def __len__(self):
        return len(self.Ws)

    def __getitem__(self, i):
        return self.Ws[i % len(self.Ws)]

    def phydims(self):
        """phydims(H) (mpohamiltonian.jl:88)."""
        return list(self.dims)

    def tompotensors(self):
        """tompotensors(H) (mpohamiltonian.jl:150)."""
        return [W.copy(order="F") for W in self.Ws]

    def bonddim(self, l=None):
        """bonddim(H, ℓ): levels of the bond right of site ℓ (0-based); bonddim(H): the uniform
        level count, a ValueError for rectangular chains (mpohamiltonian.jl:68-83)."""
        if l is not None:
            return self.levels[l % len(self.Ws)]
        if not all(W.shape[0] == W.shape[2] == self.levels[0] for W in self.Ws):
            raise ValueError("non-uniform level structure (rectangular Schur chain); use bonddim(l) or max_bonddim()")
        return self.levels[0]

    def max_bonddim(self):
        """max_bonddim(H) (mpohamiltonian.jl:84)."""
        return self.w


@dataclass
class CanonicalIMPS:
    """CanonicalIMPS{Float64} (src/states/canonicalmps.jl:17-40) as host arrays:
    AL[l] C[l] = AC[l] = C[l-1] AR[l]."""

    AL: list
    AR: list
    C: list
    AC: list

    def __len__(self):
        return len(self.AL)

    @property
    def D(self):
        return self.AL[0].shape[0]

    @property
    def d(self):
        return self.AL[0].shape[1]

    def copy(self):
        return CanonicalIMPS(*[[x.copy(order="F") for x in xs] for xs in (self.AL, self.AR, self.C, self.AC)])

    def bonddims(self):
        """bonddim(ψ, l) for every bond l (the bond right of site l)."""
        return [a.shape[2] for a in self.AL]

    def uniform_bonds(self):
        return len(set(self.bonddims())) == 1 and self.AL[0].shape[0] == self.AL[0].shape[2]

    def is_complex(self):
        """scalartype(ψ) <: Complex."""
        return any(np.iscomplexobj(x) for xs in (self.AL, self.AR, self.C, self.AC) for x in xs)

    def complex(self):
        """complex(ψ) (tdvp.jl:97-100 promotes a real chain before real-time steps)."""
        return CanonicalIMPS(*[[np.asfortranarray(x, dtype=np.complex128) for x in xs]
                               for xs in (self.AL, self.AR, self.C, self.AC)])

    def norm(self):
        """norm(ψ) = ‖AC[1]‖ (canonicalmps.jl:238)."""
        return float(np.linalg.norm(self.AC[0]))


def random_tensors(D, phydims, seed, dtype=np.float64):
    """Raw ψ₀ tensors: N(0,1) entries from numpy PCG64(seed), site by site, column-major
    (the role of randn in randomimps, src/states/constructors.jl:13-22); for a complex dtype each
    site draws the real parts, then the imaginary parts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = len(phydims)
    # D: one bond dimension, or one per bond (bond l right of site l)
    chi = [int(D)] * N if np.ndim(D) == 0 else [int(x) for x in D]
    shapes = [(chi[(l - 1) % N], phydims[l], chi[l]) for l in range(N)]
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        out = []
        for sh in shapes:
            re = rng.standard_normal(int(np.prod(sh)))
            im = rng.standard_normal(int(np.prod(sh)))
            out.append(np.reshape(re + 1j * im, sh, order="F"))
        return out
    return [np.reshape(rng.standard_normal(int(np.prod(sh))), sh, order="F") for sh in shapes]


def canonicalimps(As, H: SparseIMPO, ctx: Context | None = None, tol=1e-13, maxiter=300):
    """CanonicalIMPS(As) (canonicalmps.jl:151-170): gaugefix!(order = :LR) from C₀ = I,
    run on the device.  Returns (ψ, ctx)."""
    ctx = ctx or Context()
    ctx.set_mpo(H)
    ctx.set_state_raw(As, tol, maxiter)
    return ctx.get_state(), ctx


def randomimps(D, H: SparseIMPO, seed=0, ctx: Context | None = None, dtype=np.float64, **kw):
    """randomimps(T, phydims(H); D) (constructors.jl:13-22) with the seeded generator; T is
    Float64 or ComplexF64 (dtype=complex); D is one bond dimension or one per bond."""
    return canonicalimps(random_tensors(D, H.phydims(), seed, dtype), H, ctx, **kw)


# ---------------------------------------------------------------------------- environments


class GPUCache:
    """Environments of the Hamiltonian channel (DMRGCache, envs.jl:248-253) resident on
    the device.  leftenv/rightenv copy one environment to the host."""

    def __init__(self, psi: CanonicalIMPS, H: SparseIMPO, ctx: Context | None = None):
        self.ctx = ctx or Context()
        self.operator = H
        self.ket = psi
        self.ctx.set_mpo(H)
        self.ctx.set_state(psi)

    def __len__(self):
        return len(self.ket)


def DMRGCache(psi: CanonicalIMPS, H: SparseIMPO, tol=1e-10, maxiter=300, krylovdim=30, ctx=None):
    """DMRGCache(ψ, H::SparseIMPO; tol, maxiter, krylovdim) (envs.jl:321-518)."""
    envs = GPUCache(psi, H, ctx)
    envs.ctx.call("imps_envs_reset")
    envs.ctx.call("imps_envs_recalculate", float(tol), int(maxiter), int(krylovdim))
    return envs


def recalculate_(envs: GPUCache, psi: CanonicalIMPS, H: SparseIMPO | None = None, tol=1e-10,
                 maxiter=300, krylovdim=30):
    """recalculate!(envs, ψ, H; tol) (envs.jl:530-546): warm-started rebuild."""
    if psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
    envs.ctx.call("imps_envs_recalculate", float(tol), int(maxiter), int(krylovdim))
    return envs


def leftenv(envs: GPUCache, site):
    """leftenv(envs, ℓ) (src/environments.jl:36): (D, w, D), periodic site index."""
    return envs.ctx.get_env(site, _lib.IMPS_SIDE_LEFT)


def rightenv(envs: GPUCache, site):
    """rightenv(envs, ℓ) (src/environments.jl:39)."""
    return envs.ctx.get_env(site, _lib.IMPS_SIDE_RIGHT)


# ---------------------------------------------------------------------------- effective operators


class MPO_AC_Hamiltonian:
    """AC_hamiltonian(site, ψ, H, ψ, envs) (effective.jl:66-68); calling it applies
    effective.jl:90-102 on the device."""

    def __init__(self, site, envs: GPUCache, prec=_lib.IMPS_PREC_FP64):
        self.site, self.envs, self.prec = site, envs, prec

    def __call__(self, x):
        ctx = self.envs.ctx
        x = ctx.pad_site(self.site, zfortran(x) if ctx.cplx else x)
        y = np.zeros_like(x, order="F")
        ctx.call("imps_hac_apply", int(self.site), x.ctypes.data_as(_lib.P_double), y.ctypes.data_as(_lib.P_double),
                 int(self.prec))
        return ctx.unpad_site(self.site, y)


class MPO_C_Hamiltonian:
    """C_hamiltonian(site, ...) (effective.jl:52-54) applying effective.jl:72-76."""

    def __init__(self, site, envs: GPUCache, prec=_lib.IMPS_PREC_FP64):
        self.site, self.envs, self.prec = site, envs, prec

    def __call__(self, x):
        ctx = self.envs.ctx
        x = ctx.pad_bond(self.site, zfortran(x) if ctx.cplx else x)
        y = np.zeros_like(x, order="F")
        ctx.call("imps_hc_apply", int(self.site), x.ctypes.data_as(_lib.P_double), y.ctypes.data_as(_lib.P_double),
                 int(self.prec))
        return ctx.unpad_bond(self.site, y)


def AC_hamiltonian(site, below, operator, above, envs: GPUCache, prec=_lib.IMPS_PREC_FP64):
    return MPO_AC_Hamiltonian(site, envs, prec)


def C_hamiltonian(site, below, operator, above, envs: GPUCache, prec=_lib.IMPS_PREC_FP64):
    return MPO_C_Hamiltonian(site, envs, prec)


@dataclass
class Lanczos:
    """KrylovKit.Lanczos(; tol, krylovdim, maxiter, eager) as used by fixedpoint (utility.jl:180-186)
    and inside alg_eigsolve."""

    tol: float = 1e-10
    krylovdim: int = 30
    maxiter: int = 300
    eager: bool = True


def fixedpoint(op, x0, which="SR", alg: Lanczos | None = None):
    """fixedpoint(op, x₀, :SR, Lanczos) (utility.jl:180-186) for the effective operators:
    the Krylov vectors stay on the device.  Returns (λ, v, info)."""
    if which != "SR":
        raise ValueError("only which = :SR is supported for the effective Hamiltonians")
    alg = alg or Lanczos()
    p = _lib.KrylovParams(alg.tol, alg.krylovdim, alg.maxiter, 1 if alg.eager else 0, int(op.prec))
    info = _lib.KrylovInfo()
    lam = ctypes.c_double()
    ctx = op.envs.ctx
    dt = np.complex128 if ctx.cplx else np.float64
    if isinstance(op, MPO_AC_Hamiltonian):
        v = np.zeros((ctx.D, ctx.d, ctx.D), dtype=dt, order="F")
        ctx.call("imps_eigsolve_ac", int(op.site), ctypes.byref(p), ctypes.byref(lam), v.ctypes.data_as(_lib.P_double),
                 ctypes.byref(info))
        v = ctx.unpad_site(op.site, v)
    elif isinstance(op, MPO_C_Hamiltonian):
        v = np.zeros((ctx.D, ctx.D), dtype=dt, order="F")
        ctx.call("imps_eigsolve_c", int(op.site), ctypes.byref(p), ctypes.byref(lam), v.ctypes.data_as(_lib.P_double),
                 ctypes.byref(info))
        v = ctx.unpad_bond(op.site, v)
    else:
        raise TypeError("fixedpoint: operator must come from AC_hamiltonian or C_hamiltonian")
    return lam.value, v, info


class TransferMatrix:
    """TransferMatrix(ψ.AL, ψ.AL; side = :left) or (ψ.AR, ψ.AR; side = :right)
    (src/transfermatrix.jl:122-191), applied on the device to vec'd (D, D) matrices."""

    def __init__(self, envs_or_ctx, side="left"):
        self.ctx = envs_or_ctx.ctx if isinstance(envs_or_ctx, GPUCache) else envs_or_ctx
        self.side = _lib.IMPS_SIDE_LEFT if side == "left" else _lib.IMPS_SIDE_RIGHT
        D = self.ctx.D
        self.shape = (D * D, D * D)

    def __matmul__(self, v):
        return self(v)

    def __call__(self, v):
        D = self.ctx.D
        conv = zfortran if self.ctx.cplx else fortran
        X = conv(np.reshape(v, (D, D), order="F"))
        out = np.zeros((D, D), dtype=X.dtype, order="F")
        self.ctx.call("imps_transfer_apply", self.side, X.ctypes.data_as(_lib.P_double),
                      out.ctypes.data_as(_lib.P_double))
        return out.reshape(-1, order="F") if np.ndim(v) == 1 else out


def regularize_(v, lvec, rvec, ctx: Context | None = None):
    """regularize!(v, lvec, rvec) (envs.jl:28-32): v -= (Σ lvec[a,b] v[b,a]) rvec, in place."""
    if v.dtype != np.float64 or not v.flags.f_contiguous:
        raise TypeError("regularize_: v must be a Fortran-ordered float64 array")
    n = v.shape[0]
    if v.shape != (n, n) or np.shape(lvec) != (n, n) or np.shape(rvec) != (n, n):
        raise ValueError("regularize_: DimensionMismatch")
    own = ctx is None
    ctx = ctx or Context()
    l, r = fortran(lvec), fortran(rvec)
    ctx.call("imps_regularize", n, dptr(v), dptr(l), dptr(r))
    if own:
        ctx.close()
    return v


def regauge_(envs: GPUCache, site):
    """regauge!(AC, C; alg = QRpos()) (ortho.jl:214-219) of the pending (AC, C) produced by
    fixedpoint on `site`; returns the candidate AL."""
    ctx = envs.ctx
    al = np.zeros((ctx.D, ctx.d, ctx.D), order="F")
    ctx.call("imps_regauge", int(site), dptr(al))
    return ctx.unpad_site(site, al)


def calc_galerkin(psi: CanonicalIMPS, H: SparseIMPO, envs: GPUCache):
    """calc_galerkin(ψ, H, envs) (vumps.jl:111-124)."""
    if psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
    eps = ctypes.c_double()
    envs.ctx.call("imps_galerkin", ctypes.byref(eps))
    return eps.value


def expectationvalue(psi: CanonicalIMPS, op, envs: GPUCache):
    """expectationvalue(ψ, H::SparseIMPO, envs) (expval.jl:93-112) or, for
    op = (site, O), the one-site value (expval.jl:33-37)."""
    if psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
    val = ctypes.c_double()
    if isinstance(op, SparseIMPO):
        envs.ctx.call("imps_energy", ctypes.byref(val))
    else:
        site, O = op
        dl = psi.AL[int(site) % len(psi)].shape[1]
        if np.shape(O) != (dl, dl):
            raise ValueError(f"expectationvalue: the operator at site {site} must be {dl} x {dl}")
        if np.iscomplexobj(O) and np.any(np.imag(O) != 0):
            # a complex operator (Sy, currents): the complex value, as the reference returns
            d = envs.ctx.d or dl
            Z = np.zeros((d, d), dtype=np.complex128, order="F")
            Z[:dl, :dl] = O
            out = np.zeros(2)
            envs.ctx.call("imps_local_expectation_complex", int(site), Z.ctypes.data_as(_lib.P_double),
                          out.ctypes.data_as(_lib.P_double))
            return complex(out[0], out[1])
        O = envs.ctx.pad_operator(fortran(np.real(O)))
        envs.ctx.call("imps_local_expectation", int(site), dptr(O), ctypes.byref(val))
    return val.value


def correlator(psi: CanonicalIMPS, O1, O2, i: int, js, envs: GPUCache):
    """correlator(ψ, O1, O2, i, j | js) (correlators.jl:13-42): ⟨ψ|O1[i] O2[j]|ψ⟩ / dot(ψ, ψ)
    for j > i (0-based sites), computed on the device.  A scalar for an int j, an array for a
    range of increasing js; each value equals the single-j form."""
    if psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
    scalar = isinstance(js, (int, np.integer))
    jl = [int(js)] if scalar else [int(j) for j in js]
    di = psi.AL[int(i) % len(psi)].shape[1]
    if np.shape(O1) != (di, di) or any(np.shape(O2) != (psi.AL[j % len(psi)].shape[1],) * 2 for j in jl):
        raise ValueError("correlator: O1 must match the physical dimension of site i and O2 that of every j")
    jarr = np.array(jl, dtype=np.int32)
    if any(np.iscomplexobj(O) and np.any(np.imag(O) != 0) for O in (O1, O2)):
        # complex operators (Sy ⊗ Sy, currents): complex values (imps_correlator_complex)
        d = envs.ctx.d or di

        def zpad(O):
            Z = np.zeros((d, d), dtype=np.complex128, order="F")
            Z[:np.shape(O)[0], :np.shape(O)[1]] = O
            return Z

        Z1, Z2 = zpad(O1), zpad(O2)
        out = np.zeros(2 * len(jl))
        envs.ctx.call("imps_correlator_complex", Z1.ctypes.data_as(_lib.P_double), Z2.ctypes.data_as(_lib.P_double),
                      int(i), len(jl), jarr.ctypes.data_as(ctypes.c_void_p), dptr(out))
        vals = out[0::2] + 1j * out[1::2]
        return complex(vals[0]) if scalar else vals
    O1, O2 = envs.ctx.pad_operator(fortran(np.real(O1))), envs.ctx.pad_operator(fortran(np.real(O2)))
    out = np.zeros(len(jl))
    envs.ctx.call("imps_correlator", dptr(O1), dptr(O2), int(i), len(jl),
                  jarr.ctypes.data_as(ctypes.c_void_p), dptr(out))
    return float(out[0]) if scalar else out


def dot(psi1: CanonicalIMPS, psi2: CanonicalIMPS, envs: GPUCache):
    """dot(ψ₁, ψ₂) (canonicalmps.jl:262-269): dominant eigenvalue of the mixed AL transfer
    matrix, on the device (ψ₁ becomes the context's state).  A float, or a complex number when the
    dominant eigenvalue is complex (as the reference returns it)."""
    if len(psi2) != len(psi1) or any(a.shape != b.shape for a, b in zip(psi1.AL, psi2.AL)):
        raise ValueError("dot: the library takes states with the same unit cell, bond and physical "
                         "dimensions")
    if psi1.is_complex() or psi2.is_complex():
        # complex states (a real one is promoted, as Julia promotes the scalar type): the complex
        # dominant eigenvalue on the embedding (imps_dot_state_complex), always a complex number
        p1 = psi1 if psi1.is_complex() else psi1.complex()
        if p1 is not envs.ket:
            envs.ctx.set_state(p1)
            envs.ket = p1
        AL2 = [envs.ctx.pad_site(l, np.asarray(a, dtype=np.complex128)) for l, a in enumerate(psi2.AL)]
        pp, keep = zptr_array(AL2)
        val = np.zeros(2)
        envs.ctx.call("imps_dot_state_complex", len(AL2), pp, dptr(val))
        return complex(val[0], val[1])
    if psi1 is not envs.ket:
        envs.ctx.set_state(psi1)
        envs.ket = psi1
    AL2 = [envs.ctx.pad_site(l, a) for l, a in enumerate(psi2.AL)]
    pp, keep = ptr_array(AL2)
    val = np.zeros(2)
    envs.ctx.call("imps_dot_state", len(AL2), pp, dptr(val))
    return complex(val[0], val[1]) if val[1] != 0.0 else float(val[0])


def norm(psi: CanonicalIMPS):
    """norm(ψ) = norm(ψ.AC[1]) (canonicalmps.jl:238)."""
    return float(np.linalg.norm(psi.AC[0]))


def fidelity(psi1: CanonicalIMPS, psi2: CanonicalIMPS, envs: GPUCache):
    """fidelity(ψ₁, ψ₂) = |dot(ψ₁, ψ₂)| / (norm(ψ₁) norm(ψ₂)) (canonicalmps.jl:281-282)."""
    return abs(dot(psi1, psi2, envs)) / (norm(psi1) * norm(psi2))


def infidelity(psi1: CanonicalIMPS, psi2: CanonicalIMPS, envs: GPUCache):
    """infidelity(ψ₁, ψ₂) = 1 − fidelity(ψ₁, ψ₂) (canonicalmps.jl:283-284)."""
    return 1.0 - fidelity(psi1, psi2, envs)


class DenseIMPO:
    """DenseIMPO (src/operators/infinitempo.jl:12-30): per-site dense MPO tensors (wl, d, wr, d),
    periodic over the unit cell; each right bond must match the next site's left bond.  Bond and
    physical dimensions may differ from site to site: the device runs on the tensors zero-padded
    to one square shape (_dense_square)."""

    def __init__(self, Ws):
        Ws = [fortran(W) for W in Ws]
        if not Ws:
            raise ValueError("Ws must not be empty")
        for l, W in enumerate(Ws):
            if W.ndim != 4 or W.shape[1] != W.shape[3]:
                raise ValueError(f"site {l} physical dims: {W.shape}")
            if W.shape[2] != Ws[(l + 1) % len(Ws)].shape[0]:
                raise ValueError(f"MPO bond mismatch at site {l}")
        self.Ws = Ws

    def __len__(self):
        return len(self.Ws)

    def __getitem__(self, i):
        return self.Ws[i % len(self.Ws)]


def _dense_square(Ws, psi, ctx, what):
    """DenseIMPO tensors with per-bond and per-site physical dimensions zero-padded to one square
    shape (w = the largest bond, d = the context's physical dimension).  The padded bond indices
    and physical sectors carry no weight: each transfer map becomes T (+) 0, with T's nonzero
    spectrum and eigenvectors (zeros on the padding), so every value is unchanged."""
    Ws = [fortran(x) for x in Ws]
    for l in range(len(psi)):
        x = Ws[l % len(Ws)]
        if x.ndim != 4 or x.shape[1] != x.shape[3] or x.shape[1] != psi.AL[l].shape[1]:
            raise ValueError(f"{what}: physical dimensions of W and ψ differ at site {l}")
    w = max(max(x.shape[0], x.shape[2]) for x in Ws)
    d = ctx.d or Ws[0].shape[1]
    out = []
    for x in Ws:
        if x.shape == (w, d, w, d):
            out.append(x)
            continue
        P = np.zeros((w, d, w, d), order="F")
        P[:x.shape[0], :x.shape[1], :x.shape[2], :x.shape[3]] = x
        out.append(P)
    return out, w, d


def dominant_env(psi: CanonicalIMPS, envs, W=None, side: str = "left", which: str = "LM",
                 tol: float = 1e-10, krylovdim: int = 30, maxiter: int = 300):
    """dominant_env(ψ; side, which) / dominant_env(W::DenseIMPO, ψ; ...) (envs.jl:71-99) on the
    device: (λ, L) with L of shape (D, w, D) (w = 1 for the identity channel), ‖L‖ = 1.  ``envs``
    is a GPUCache or a Context.  Only ``which = "LM"`` (the reference's default and the only value
    its callers use); the eigenvector's sign is fixed by a non-negative entry sum."""
    if which != "LM":
        raise ValueError("dominant_env: only which='LM' is supported")
    if side not in ("left", "right"):
        raise ValueError("side must be 'left' or 'right'")
    ctx = envs.ctx if hasattr(envs, "ctx") else envs
    if hasattr(envs, "ket"):
        if psi is not envs.ket:
            ctx.set_state(psi)
            envs.ket = psi
    else:
        ctx.set_state(psi)
    D = psi.D
    lam = ctypes.c_double()
    if W is None:
        out = np.zeros((D, 1, D), order="F")
        ctx.call("imps_dominant_env", 0 if side == "left" else 1, 0, 1, psi.d, None, float(tol), int(krylovdim),
                 int(maxiter), ctypes.byref(lam), dptr(out))
        return lam.value, out
    W0 = W.Ws if isinstance(W, DenseIMPO) else list(W)
    Ws, w, d = _dense_square(W0, psi, ctx, "dominant_env")
    pp, keep = ptr_array(Ws)
    out = np.zeros((D, w, D), order="F")
    ctx.call("imps_dominant_env", 0 if side == "left" else 1, len(Ws), w, d, pp, float(tol), int(krylovdim),
             int(maxiter), ctypes.byref(lam), dptr(out))
    # both environments live on the bond left of site 0
    wb = np.shape(W0[0])[0]
    return lam.value, (out if wb == w else np.asfortranarray(out[:, :wb, :]))


def dense_expectation(psi: CanonicalIMPS, W, envs, tol: float = 1e-10, krylovdim: int = 30, maxiter: int = 300,
                      return_envs: bool = False):
    """expectationvalue(ψ, W::DenseIMPO, DMRGCache(ψ, W)) (expval.jl:69-77; envs.jl:284-304,
    559-581) on the device.  ``envs`` is a GPUCache or a Context whose state is ψ's model's.  With
    ``return_envs`` also the DenseIMPO cache's environments (lefts, rights), each (D, w, D)."""
    ctx = envs.ctx if hasattr(envs, "ctx") else envs
    if hasattr(envs, "ket"):
        if psi is not envs.ket:
            ctx.set_state(psi)
            envs.ket = psi
    else:
        ctx.set_state(psi)
    W0 = W.Ws if isinstance(W, DenseIMPO) else list(W)
    Ws, w, d = _dense_square(W0, psi, ctx, "dense_expectation")
    pp, keep = ptr_array(Ws)
    N, D = len(psi), psi.D
    lefts = [np.zeros((D, w, D), order="F") for _ in range(N)]
    rights = [np.zeros((D, w, D), order="F") for _ in range(N)]
    pl, keep_l = ptr_array(lefts)
    pr, keep_r = ptr_array(rights)
    val = ctypes.c_double()
    ctx.call("imps_dense_expectation", len(Ws), w, d, pp, float(tol), int(krylovdim), int(maxiter),
             ctypes.byref(val), pl if return_envs else None, pr if return_envs else None)
    if not return_envs:
        return val.value
    # lefts[l] on the bond left of site l, rights[l] on the bond right of it
    lefts = [np.asfortranarray(G[:, :np.shape(W0[l % len(W0)])[0], :]) for l, G in enumerate(lefts)]
    rights = [np.asfortranarray(G[:, :np.shape(W0[l % len(W0)])[2], :]) for l, G in enumerate(rights)]
    return val.value, lefts, rights


def _file_call(name, *args):
    L = _lib.load()
    st = getattr(L, name)(*args)
    if st != _lib.IMPS_OK:
        msg = L.imps_last_error(None)
        raise _lib.ImpsError(f"{name} failed with status {st}: {msg.decode() if msg else ''}")


def save_state(psi: CanonicalIMPS, path):
    """Writes ψ (AL, AR, C, AC of every site) to a state file (imps_write_state_file); the
    library's own format, for checkpoint and resume (the reference has no file format)."""
    AL, AR, C, AC = ([fortran(x) for x in xs] for xs in (psi.AL, psi.AR, psi.C, psi.AC))
    if any(x.shape != AL[0].shape for x in AL + AR + AC):
        raise ValueError("save_state: the state file holds one physical dimension for every site")
    keep = [ptr_array(x) for x in (AL, AR, C, AC)]
    _file_call("imps_write_state_file", os.fsencode(path), len(AL), psi.D, psi.d, *(k[0] for k in keep))


def load_state(path) -> CanonicalIMPS:
    """Reads a state file written by save_state or Context.checkpoint."""
    n, D, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _file_call("imps_state_file_info", os.fsencode(path), ctypes.byref(n), ctypes.byref(D), ctypes.byref(d))
    N, D, d = n.value, D.value, d.value
    AL = [np.zeros((D, d, D), order="F") for _ in range(N)]
    AR = [np.zeros((D, d, D), order="F") for _ in range(N)]
    AC = [np.zeros((D, d, D), order="F") for _ in range(N)]
    C = [np.zeros((D, D), order="F") for _ in range(N)]
    keep = [ptr_array(x) for x in (AL, AR, C, AC)]
    _file_call("imps_read_state_file", os.fsencode(path), N, D, d, *(k[0] for k in keep))
    return CanonicalIMPS(AL, AR, C, AC)


def save_mpo(H, path):
    """Writes the per-site MPO tensors (w, d, w, d) of a SparseIMPO or DenseIMPO to an MPO file."""
    Ws = [fortran(W) for W in H.Ws]
    w, d = Ws[0].shape[0], Ws[0].shape[1]
    for W in Ws:
        if W.shape != (w, d, w, d):
            raise ValueError(f"save_mpo: every tensor must have shape {(w, d, w, d)}")
    pp, keep = ptr_array(Ws)
    _file_call("imps_write_mpo_file", os.fsencode(path), len(Ws), w, d, pp)


def load_mpo(path, dense: bool = False):
    """Reads an MPO file as a SparseIMPO (or a DenseIMPO with dense=True)."""
    n, w, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _file_call("imps_mpo_file_info", os.fsencode(path), ctypes.byref(n), ctypes.byref(w), ctypes.byref(d))
    Ws = [np.zeros((w.value, d.value, w.value, d.value), order="F") for _ in range(n.value)]
    pp, keep = ptr_array(Ws)
    _file_call("imps_read_mpo_file", os.fsencode(path), n.value, w.value, d.value, pp)
    return DenseIMPO(Ws) if dense else SparseIMPO(Ws)


def energy_gradient(psi: CanonicalIMPS, H: SparseIMPO, envs: GPUCache):
    """d e_cell / d W[site] at fixed ψ and environments (the reverse of the H_AC
    contraction, effective.jl:90-102, applied to the energy of expval.jl:93-112).  At a
    converged VUMPS state this is the derivative of the ground-state energy with respect
    to each non-structural MPO block.  Returns one (w, d, w, d) array per site."""
    if psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
    out = [np.zeros((H.w, H.d, H.w, H.d), order="F") for _ in range(len(H))]
    pp, keep = ptr_array(out)
    envs.ctx.call("imps_energy_gradient", pp)
    if H.square:
        return out
    N = len(H)
    res = []
    for l, G in enumerate(out):
        li = envs.ctx.bond_levels(l - 1)
        lj = envs.ctx.bond_levels(l)
        dl = H.dims[l]
        res.append(np.asfortranarray(G[np.ix_(li, range(dl), lj, range(dl))]))
    return res


def objective_mpo_1site(O):
    """Schur MPO (w = 2) whose value per site is the one-site expectation value <O>: the identity at
    the unit corners and O in block (0, 1)."""
    O = np.asarray(O, dtype=np.float64)
    d = O.shape[0]
    if O.shape != (d, d):
        raise ValueError("objective_mpo_1site: O must be (d, d)")
    Wo = np.zeros((2, d, 2, d), order="F")
    Wo[0, :, 0, :] = np.eye(d)
    Wo[1, :, 1, :] = np.eye(d)
    Wo[0, :, 1, :] = O
    return Wo


@dataclass
class ResponseInfo:
    iterations: int
    converged: bool
    relres: float
    estimate: float
    hessian_applies: int
    level_solves: int
    level_ops: int
    unconverged_solves: int


def fixedpoint_gradient(psi: CanonicalIMPS, H: SparseIMPO, envs: GPUCache, Wobj, tol=1e-10, maxiter=1000,
                        rcond=1e-10):
    """Response of a converged one-site VUMPS fixed point: (value, dO/dW, info) for the objective
    <Wobj> per site (a Schur MPO (wo, d, wo, d); ``objective_mpo_1site(O)`` for a one-site
    observable) with respect to the Hamiltonian tensor W (w, d, w, d).

    The reference has no derivative of its fixed point; this is the tangent-space linear response
    built from its operators (H_AC effective.jl:90-102, the slice pushes envs.jl:104-114,
    regularize! envs.jl:28-32, linsolve envs.jl:14-19): (H_eff + K) L = P[H_AC^O(AC)] solved by
    MINRES on the device, dO/dW = -2 d/dW <Phi(L)|H[W]|psi> by the reverse H_AC contraction and
    the adjoint environments.  The derivative of any scalar f(W) = <Wobj> through the fixed point
    is then sum(dW * dW/dp) for a parameter p."""
    if len(H) != 1:
        raise ValueError("fixedpoint_gradient: one-site unit cell only")
    if psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
        envs.ctx.call("imps_envs_recalculate", 1e-14, 300, 30)
    Wo = fortran(Wobj)
    if Wo.ndim != 4 or Wo.shape[0] != Wo.shape[2] or Wo.shape[1] != H.d or Wo.shape[3] != H.d:
        raise ValueError(f"fixedpoint_gradient: objective must be (wo, {H.d}, wo, {H.d}), got {Wo.shape}")
    dW = np.zeros((H.w, H.d, H.w, H.d), order="F")
    val = ctypes.c_double()
    info = _lib.ResponseInfo()
    envs.ctx.call("imps_fixedpoint_gradient", Wo.shape[0], H.d, dptr(Wo), float(tol), int(maxiter), float(rcond),
                  ctypes.byref(val), dptr(dW), ctypes.byref(info))
    return val.value, dW, ResponseInfo(info.iterations, bool(info.converged), info.relres, info.estimate,
                                       info.hessian_applies, info.level_solves, info.level_ops,
                                       info.unconverged_solves)


def tangent_hessian_apply(psi: CanonicalIMPS, envs: GPUCache, B, rcond=1e-10, tol=1e-13):
    """P (H_eff + K) P B: the operator of the fixed-point response system (one-site cell)."""
    if psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
    Bf = fortran(B)
    if Bf.shape != psi.AL[0].shape:
        raise ValueError(f"tangent_hessian_apply: B must have the MPS shape {psi.AL[0].shape}, got {Bf.shape}")
    out = np.zeros_like(Bf, order="F")
    envs.ctx.call("imps_tangent_hessian_apply", float(rcond), float(tol), dptr(Bf), dptr(out))
    return out


# ---------------------------------------------------------------------------- solver objects


@dataclass
class SolverTol:
    """The NamedTuple (; tol, maxiter) form of alg_gauge and alg_environments (utility.jl:350-364)."""

    tol: float = 1e-10
    maxiter: int = 300


@dataclass
class DynamicTol:
    """DynamicTol(alg; tol_min, tol_max, tol_factor) (utility.jl:136-148): every iteration the inner
    solver's tolerance becomes clamp(eps * tol_factor / sqrt(max(iter, 1)), tol_min, tol_max)."""

    alg: object
    tol_min: float = 1e-6
    tol_max: float = 1e-2
    tol_factor: float = 0.1

    def __post_init__(self):
        if not 0 <= self.tol_min <= self.tol_max:
            raise ValueError("DynamicTol requires 0 ≤ tol_min ≤ tol_max")


def alg_eigsolve(tol=1e-10, maxiter=300, eager=True, krylovdim=30, dynamic_tols=True, tol_min=1e-14,
                 tol_max=1e-4, tol_factor=1e-3):
    """Defaults.alg_eigsolve (utility.jl:322-333), Hermitian case."""
    alg = Lanczos(tol=tol, krylovdim=krylovdim, maxiter=maxiter, eager=eager)
    return DynamicTol(alg, tol_min, tol_max, tol_factor) if dynamic_tols else alg


def alg_gauge(tol=1e-13, maxiter=300, dynamic_tols=True, tol_min=1e-14, tol_max=1e-4, tol_factor=1e-6):
    """Defaults.alg_gauge (utility.jl:350-354)."""
    alg = SolverTol(tol, maxiter)
    return DynamicTol(alg, tol_min, tol_max, tol_factor) if dynamic_tols else alg


def alg_environments(tol=1e-10, maxiter=300, dynamic_tols=True, tol_min=1e-14, tol_max=1e-4, tol_factor=1e-4):
    """Defaults.alg_environments (utility.jl:359-364)."""
    alg = SolverTol(tol, maxiter)
    return DynamicTol(alg, tol_min, tol_max, tol_factor) if dynamic_tols else alg


def _set_solver(p, which, obj):
    """Writes alg_eigsolve / alg_gauge / alg_environments into the parameter struct."""
    dyn = isinstance(obj, DynamicTol)
    inner = obj.alg if dyn else obj
    if which == "eig":
        if not isinstance(inner, Lanczos):
            raise TypeError("alg_eigsolve must be a Lanczos or a DynamicTol(Lanczos)")
        p.eig_dynamic, p.eig_tol, p.eig_maxiter = int(dyn), float(inner.tol), int(inner.maxiter)
        p.krylovdim, p.eig_eager = int(inner.krylovdim), int(bool(inner.eager))
        if dyn:
            p.tol_min, p.tol_max, p.eigs_tolfactor = obj.tol_min, obj.tol_max, obj.tol_factor
    elif which == "gauge":
        p.gauge_dynamic, p.gauge_tol, p.gauge_maxiter = int(dyn), float(inner.tol), int(inner.maxiter)
        if getattr(inner, "eig_miniter", None) is not None:
            p.gauge_eig_miniter = int(inner.eig_miniter)
        if dyn:
            p.gauge_tol_min, p.gauge_tol_max, p.gauge_tolfactor = obj.tol_min, obj.tol_max, obj.tol_factor
    else:
        # recalculate! forwards only tol (vumps.jl:58, 73); maxiter stays the DMRGCache default
        p.envs_dynamic, p.envs_tol = int(dyn), float(inner.tol)
        if dyn:
            p.envs_tol_min, p.envs_tol_max, p.envs_tolfactor = obj.tol_min, obj.tol_max, obj.tol_factor


def _params(alg, prec, refine_below, eig_floor):
    p = _lib.VumpsParams()
    _lib.check(None, _lib.load().imps_vumps_default_params(ctypes.byref(p)), "imps_vumps_default_params")
    p.tol, p.maxiter, p.verbosity = alg.tol, alg.maxiter, alg.verbosity
    p.tol_min, p.tol_max = alg.tol_min, alg.tol_max
    p.eigs_tolfactor, p.gauge_tolfactor, p.envs_tolfactor = alg.eigs_tolfactor, alg.gauge_tolfactor, alg.envs_tolfactor
    p.krylovdim, p.eig_maxiter, p.gauge_maxiter, p.envs_maxiter = (alg.krylovdim, alg.eig_maxiter, alg.gauge_maxiter,
                                                                   alg.envs_maxiter)
    p.prec, p.refine_below, p.eig_floor = prec, refine_below, eig_floor
    p.refine_eig = int(bool(getattr(alg, "refine_eig", True)))
    p.gauge_eig_miniter = int(getattr(alg, "gauge_eig_miniter", 10))
    for which, obj in (("eig", alg.alg_eigsolve), ("gauge", alg.alg_gauge), ("envs", alg.alg_environments)):
        if obj is not None:
            _set_solver(p, which, obj)
    return p


# ---------------------------------------------------------------------------- VUMPS


@dataclass
class VUMPS:
    """VUMPS(; D, tol, maxiter, verbosity, alg_gauge, alg_eigsolve, alg_environments, finalize)
    (algdefs.jl:26-35) with the Defaults of utility.jl:308-364.  The flat fields are the default
    solver objects' parameters; alg_eigsolve / alg_gauge / alg_environments (Lanczos, SolverTol,
    optionally wrapped in DynamicTol) override them when given.  finalize(iter, ψ, H, envs) ->
    (ψ, envs) runs after recalculate! every iteration (vumps.jl:75); a returned state other than
    the one passed in becomes the context's state (its environments are then rebuilt).  `prec`
    selects the arithmetic of the bulk contractions; with a mixed precision the loop refines in
    FP64 once the Galerkin error drops below `refine_below`; `eig_floor` (relative to |λ|) bounds
    the eigensolver tolerance of the reduced-precision phase (0: the library's default);
    `refine_eig` (default) keeps the reduced-precision operator after the switch for Jacobi-Davidson
    corrections of FP64 residuals instead of switching the eigensolves to FP64 Lanczos."""

    D: int = 64
    tol: float = 1e-10
    maxiter: int = 300
    verbosity: int = 0
    tol_min: float = 1e-14
    tol_max: float = 1e-4
    eigs_tolfactor: float = 1e-3
    gauge_tolfactor: float = 1e-6
    envs_tolfactor: float = 1e-4
    krylovdim: int = 30
    eig_maxiter: int = 300
    gauge_maxiter: int = 300
    envs_maxiter: int = 300
    prec: int = _lib.IMPS_PREC_FP64
    refine_below: float = 0.0
    eig_floor: float = 0.0
    refine_eig: bool = True
    gauge_eig_miniter: int = 10
    alg_eigsolve: object = None
    alg_gauge: object = None
    alg_environments: object = None
    finalize: object = None

    def params(self):
        return _params(self, self.prec, self.refine_below, self.eig_floor)


@dataclass
class IDMRG:
    """IDMRG(; D, tol, maxiter, verbosity, alg_gauge, alg_environments, alg_eigsolve, finalize)
    (algdefs.jl:72-82): single-site infinite DMRG with the same solver defaults and objects as VUMPS
    (utility.jl:308-364); finalize runs after every sweep (idmrg.jl:60) and must keep D."""

    D: int = 64
    tol: float = 1e-10
    maxiter: int = 300
    verbosity: int = 0
    tol_min: float = 1e-14
    tol_max: float = 1e-4
    eigs_tolfactor: float = 1e-3
    gauge_tolfactor: float = 1e-6
    envs_tolfactor: float = 1e-4
    krylovdim: int = 30
    eig_maxiter: int = 300
    gauge_maxiter: int = 300
    envs_maxiter: int = 300
    alg_eigsolve: object = None
    alg_gauge: object = None
    alg_environments: object = None
    finalize: object = None

    def params(self):
        return _params(self, _lib.IMPS_PREC_FP64, 0.0, 0.0)


@dataclass
class IterativeConvergenceInfo:
    """IterativeConvergenceInfo (algdefs.jl:180-194) plus the library's timings."""

    niter: int
    converged: bool
    losses: list
    energy: float
    seconds: float
    t_galerkin_reached: float
    hac_applies: int
    hc_applies: int
    prec_switch_iter: int
    energies: list = field(default_factory=list)
    timings: dict = field(default_factory=dict)

    @property
    def itererr(self):
        return self.losses[-1] if self.losses else float("nan")


def find_groundstate(psi0: CanonicalIMPS | None, H: SparseIMPO, alg: VUMPS | IDMRG | None = None,
                     envs: GPUCache | None = None, seed=0, ctx: Context | None = None, callback=None):
    """find_groundstate(ψ₀, H::SparseIMPO, alg::VUMPS, envs) (vumps.jl:51-86), or with
    alg::IDMRG the single-site IDMRG driver (idmrg.jl:44-86).  With psi0 = None a random state
    of bond dimension alg.D is generated (vumps.jl:96-102).  `callback(iter, eps, energy)` is
    called after every iteration (alg.finalize's slot); a truthy return value stops the loop.
    Returns (ψ, envs, info); for IDMRG, info.losses holds the initial Galerkin error and then the
    bond-0 centre drift of each sweep, info.energies the sweeps' lowest H_AC eigenvalues."""
    alg = alg or VUMPS()
    if envs is None:
        if psi0 is None:
            psi0, ctx = randomimps(alg.D, H, seed=seed, ctx=ctx)
        envs = GPUCache(psi0, H, ctx)
    elif psi0 is not None and psi0 is not envs.ket:
        envs.ctx.set_state(psi0)
    p = alg.params()
    info = _lib.VumpsInfo()
    losses = np.zeros(alg.maxiter + 1)
    energies = []

    def _cb(it, eps, f, user):
        energies.append(f)
        if callback is not None and callback(it, eps, f):
            return 1
        return 0

    cb = _lib.ITER_CB(_cb)
    failure = []

    def _fin(it, h, user):
        # alg.finalize(iter, ψ, H, envs) -> (ψ, envs) on the host copy of the current state
        try:
            cur = envs.ctx.get_state()
            envs.ket = cur
            psi2, _ = alg.finalize(it, cur, H, envs)
            if psi2 is not cur:
                envs.ctx.set_state(psi2)
                envs.ket = psi2
            return 0
        except Exception as e:  # reported after the run returns
            failure.append(e)
            return _lib.IMPS_E_ARG

    if alg.finalize is not None:
        fin = _lib.FINALIZE_CB(_fin)
        p.finalize = fin
    entry = "imps_idmrg_run" if isinstance(alg, IDMRG) else "imps_vumps_run"
    try:
        envs.ctx.call(entry, ctypes.byref(p), ctypes.byref(info),
                      losses.ctypes.data_as(_lib.P_double), len(losses), cb, None)
    except _lib.ImpsError:
        if failure:
            raise failure[0]
        raise
    psi = envs.ctx.get_state()
    envs.ket = psi
    nl = info.niter + 1
    res = IterativeConvergenceInfo(info.niter, bool(info.converged), list(losses[:nl]), info.energy,
                                   info.seconds, info.t_galerkin_reached, info.hac_applies,
                                   info.hc_applies, info.prec_switch_iter, energies,
                                   {k: getattr(info, k) for k in ("t_local", "t_gauge", "t_envs", "t_galerkin",
                                                                  "t_energy", "gauge_sweeps", "gmres_ops",
                                                                  "qr_fallbacks", "refine_fallbacks")})
    return psi, envs, res


@dataclass
class TDVP:
    """TDVP(; integrator, alg_gauge, alg_orth) (algdefs.jl:167-173): Lanczos exponential with the
    alg_expsolve defaults (utility.jl:343-347) and the alg_gauge defaults (tolgauge, maxiter)."""

    tol: float = 1e-10
    krylovdim: int = 30
    maxiter: int = 300
    gauge_tol: float = 1e-13
    gauge_maxiter: int = 300


def timestep(psi: CanonicalIMPS, H: SparseIMPO, dt, alg: TDVP | None = None, envs: GPUCache | None = None):
    """timestep(ψ, H, dt, alg::TDVP, envs) (tdvp.jl:33-52): dt is the coefficient of the
    exponent, exp(dt H) — real dt = -τ for imaginary time, dt = -im t for real time.  A real ψ
    with a non-real dt is promoted to complex (the Julia step promotes through exponentiate).
    Returns (ψ′, envs)."""
    alg = alg or TDVP()
    dt = complex(dt)
    if dt.imag != 0.0 and not psi.is_complex():
        psi = psi.complex()
    if envs is None:
        envs = DMRGCache(psi, H)
    elif psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
        envs.ctx.call("imps_envs_reset")
    if dt.imag == 0.0:
        envs.ctx.call("imps_tdvp_timestep", float(dt.real), float(alg.tol), int(alg.krylovdim), int(alg.maxiter),
                      float(alg.gauge_tol), int(alg.gauge_maxiter))
    else:
        envs.ctx.call("imps_tdvp_timestep_complex", float(dt.real), float(dt.imag), float(alg.tol),
                      int(alg.krylovdim), int(alg.maxiter), float(alg.gauge_tol), int(alg.gauge_maxiter))
    psi2 = envs.ctx.get_state()
    envs.ket = psi2
    return psi2, envs


def time_evolve(psi0: CanonicalIMPS, H: SparseIMPO, t_span, alg: TDVP | None = None, envs: GPUCache | None = None,
                observer=None):
    """time_evolve(ψ₀, H, t_span, alg::TDVP, envs; observer) (tdvp.jl:91-112): steps
    dt = t_span[i+1] - t_span[i] (real: imaginary time; -im t: real time, the state promoted to
    complex first); returns (ψ, envs, history) with observer(ψ, iter, t) (default: the cell
    energy) at every point of t_span."""
    alg = alg or TDVP()
    psi = psi0
    if len(t_span) > 1 and complex(t_span[1] - t_span[0]).imag != 0.0 and not psi.is_complex():
        psi = psi.complex()
    if envs is None:
        envs = DMRGCache(psi, H)
    elif psi is not envs.ket:
        envs.ctx.set_state(psi)
        envs.ket = psi
        envs.ctx.call("imps_envs_reset")
        envs.ctx.call("imps_envs_recalculate", 1e-10, 300, 30)

    def obs(p, it, t):
        return expectationvalue(p, H, envs) if observer is None else observer(p, it, t)

    history = [obs(psi, 0, t_span[0])]
    for it in range(len(t_span) - 1):
        psi, envs = timestep(psi, H, t_span[it + 1] - t_span[it], alg, envs)
        history.append(obs(psi, it + 1, t_span[it + 1]))
    return psi, envs, history


def device_batch_eligible(problems, alg: VUMPS | None = None):
    """True when imps_vumps_batch_device takes the sweep: one-site cells with a common MPO shape,
    bond dimensions that are multiples of 16 (d D <= 2048), FP64 or FP32 bulk arithmetic and no
    finalize callback."""
    alg = alg or VUMPS()
    if not problems or alg.prec not in (_lib.IMPS_PREC_FP64, _lib.IMPS_PREC_FP32) or alg.finalize is not None:
        return False
    H0 = problems[0][0]
    for H, As in problems:
        D = np.asarray(As[0]).shape[0]
        if len(H) != 1 or not H.square or H.w != H0.w or H.d != H0.d or D % 16 or D < 16 or H.d * D > 2048:
            return False
    return H0.w * H0.d <= 64 and H0.d <= 8


def find_groundstates_batched(problems, alg: VUMPS | None = None, nworkers=8, device=0, gauge_tol=1e-12,
                              mode="auto", stats=None):
    """Batched parameter sweep: find_groundstate (vumps.jl:51-86) for every
    (H::SparseIMPO, raw tensors As) in `problems`.  mode "device": one lockstep device batch
    (imps_vumps_batch_device: grouped kernels for all instances, one host thread); "threads": the
    instances on `nworkers` host threads, one context and stream each (imps_vumps_batch, any unit
    cell or precision); "auto" takes the device batch when `device_batch_eligible`.  Returns one
    dict per problem with niter, converged, itererr, energy (unit cell), seconds and hac_applies;
    a dict passed as `stats` receives the device batch's rounds, launches and seconds."""
    alg = alg or VUMPS()
    if mode not in ("auto", "device", "threads"):
        raise ValueError("find_groundstates_batched: mode must be 'auto', 'device' or 'threads'")
    use_device = mode == "device" or (mode == "auto" and device_batch_eligible(problems, alg))
    L = _lib.load()
    keep = []
    arr = (_lib.BatchInstance * len(problems))()
    for k, (H, As) in enumerate(problems):
        if not H.square:
            raise ValueError("find_groundstates_batched: the batch takes square Schur chains with one physical "
                             "dimension (use find_groundstate for rectangular or heterogeneous chains)")
        Ws = [fortran(W) for W in H.Ws]
        As = [fortran(a) for a in As]
        pw, kw = ptr_array(Ws)
        pa, ka = ptr_array(As)
        keep.append((Ws, As, kw, ka))
        arr[k] = _lib.BatchInstance(len(Ws), H.w, H.d, As[0].shape[0], pw, pa, 0, 0, 0, 0.0, 0.0, 0.0, 0)
    p = alg.params()
    if use_device:
        bs = _lib.BatchStats()
        st = L.imps_vumps_batch_device(int(device), len(problems), arr, ctypes.byref(p), float(gauge_tol),
                                       ctypes.byref(bs))
        if st != 0:
            msg = L.imps_last_error(None)
            raise _lib.ImpsError(f"imps_vumps_batch_device failed with status {st}: "
                                 f"{msg.decode() if msg else ''}")
        if stats is not None:
            stats.update({"rounds": bs.rounds, "launches": bs.launches, "seconds": bs.seconds, "mode": "device",
                          "groups": bs.groups,
                          "phase_seconds": dict(zip(["canonicalize", "local_update", "gauge", "envs", "galerkin",
                                                     "energy"], list(bs.t_phase)))})
    else:
        st = L.imps_vumps_batch(int(device), len(problems), arr, ctypes.byref(p), int(nworkers), float(gauge_tol))
        if stats is not None:
            stats.update({"mode": "threads"})
    out = [{"status": r.status, "niter": r.niter, "converged": bool(r.converged), "itererr": r.itererr,
            "energy": r.energy, "seconds": r.seconds, "hac_applies": r.hac_applies} for r in arr]
    if st != 0:
        bad = [k for k, r in enumerate(out) if r["status"] != 0]
        raise _lib.ImpsError(f"imps_vumps_batch failed with status {st} on instances {bad}")
    return out
