"""imps_hac_apply_noalloc (H_AC on device pointers with the library's static device scratch, no
allocation) against imps_hac_apply_dev through the C ABI: bit-identical results, two calls in a
row with the second on a smaller problem, both precisions, structural identity levels and not;
the size limit and split-BF16 refused with an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


class Dev:
    def __init__(self, ctx, a):
        self.ctx, self.a = ctx, np.asfortranarray(a, dtype=np.float64)
        self.p = ctypes.c_void_p()
        ctx.call("imps_dev_alloc", self.a.nbytes, ctypes.byref(self.p))
        ctx.call("imps_memcpy_h2d", self.p, self.a.ctypes.data_as(ctypes.c_void_p), self.a.nbytes)

    def get(self):
        out = np.zeros_like(self.a, order="F")
        self.ctx.call("imps_memcpy_d2h", out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes)
        return out

    def free(self):
        self.ctx.call("imps_dev_free", self.p)


def _problem(imps, ctx, oracle, W, N, D, eye, seed):
    rng = np.random.default_rng(seed)
    H = imps.SparseIMPO([W] * N)
    ctx.set_mpo(H)
    psi = oracle.randomimps(D, [H.d] * N, seed=seed)
    ctx.set_state(imps.CanonicalIMPS(psi.AL, psi.AR, psi.C, psi.AC))
    GLs, GRs = [], []
    for l in range(N):
        GL = rng.standard_normal((D, H.w, D))
        GR = rng.standard_normal((D, H.w, D))
        if eye:
            GL[:, 0, :] = np.eye(D)
            GR[:, H.w - 1, :] = np.eye(D)
        ctx.set_env(l, 0, GL)
        ctx.set_env(l, 1, GR)
        GLs.append(GL)
        GRs.append(GR)
    return H, GLs, GRs


@pytest.mark.parametrize("first_noalloc", [True, False])
def test_noalloc_matches_dev_entry_two_calls(imps, oracle, first_noalloc):
    ctx = imps.Context(0)
    try:
        # first call: Hubbard D=256, N=2, identity structural levels; second call: a smaller TFIM
        # problem (D=96, N=1) with general environments on the same context, so the scratch holds
        # the previous call's contents
        cases = [(oracle.hubbard_mpo(), 2, 256, True, 3), (oracle.tfim_mpo(), 1, 96, False, 4)]
        for W, N, D, eye, seed in cases:
            H, GLs, GRs = _problem(imps, ctx, oracle, W, N, D, eye, seed)
            x = np.random.default_rng(seed + 10).standard_normal((D, H.d, D))
            dx, dy_ref, dy_na = Dev(ctx, x), Dev(ctx, np.zeros_like(x)), Dev(ctx, np.full_like(x, np.nan))
            try:
                for prec in (0, 1):
                    for site in range(N):
                        order = ("imps_hac_apply_noalloc", "imps_hac_apply_dev") if first_noalloc else \
                                ("imps_hac_apply_dev", "imps_hac_apply_noalloc")
                        for name in order:
                            ctx.call(name, site, dx.p, (dy_na if name.endswith("noalloc") else dy_ref).p, prec)
                        ctx.call("imps_synchronize")
                        y_ref, y_na = dy_ref.get(), dy_na.get()
                        assert np.array_equal(y_na, y_ref), (D, prec, site)
                        if prec == 0:
                            ref = oracle.hac_apply_dense(GLs[site], H.Ws[site], GRs[site], x)
                            assert rel(y_na, ref) < 1e-12
            finally:
                for a in (dx, dy_ref, dy_na):
                    a.free()
    finally:
        ctx.close()


def test_noalloc_refuses_oversize_and_bf16(imps, oracle):
    ctx = imps.Context(0)
    try:
        # Hubbard D=1024: 2 D w d D = 50.3M doubles > 2^24
        D = 1024
        H = imps.SparseIMPO([oracle.hubbard_mpo()])
        ctx.set_mpo(H)
        rng = np.random.default_rng(1)
        A = [np.asfortranarray(rng.standard_normal((D, 4, D)))]
        C = [np.asfortranarray(np.eye(D))]
        ctx.set_state(imps.CanonicalIMPS(A, A, C, A))
        G = np.zeros((D, H.w, D))
        ctx.set_env(0, 0, G)
        ctx.set_env(0, 1, G)
        dx, dy = Dev(ctx, np.zeros((D, 4, D))), Dev(ctx, np.zeros((D, 4, D)))
        try:
            with pytest.raises(imps._lib.ImpsError, match="scratch"):
                ctx.call("imps_hac_apply_noalloc", 0, dx.p, dy.p, 0)
        finally:
            dx.free()
            dy.free()
        # split-BF16 is refused at a size that fits
        H, _, _ = _problem(imps, ctx, oracle, oracle.tfim_mpo(), 1, 128, True, 5)
        dx, dy = Dev(ctx, np.zeros((128, 2, 128))), Dev(ctx, np.zeros((128, 2, 128)))
        try:
            with pytest.raises(imps._lib.ImpsError, match="FP64 and FP32 only"):
                ctx.call("imps_hac_apply_noalloc", 0, dx.p, dy.p, 2)
            ctx.call("imps_hac_apply_noalloc", 0, dx.p, dy.p, 0)  # the context still works
        finally:
            dx.free()
            dy.free()
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def imps():
    import infinitempsalgorithms_amd as m
    return m
