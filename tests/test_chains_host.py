"""Host-side rules of rectangular Schur chains and per-site physical dimensions (no device):
the level-chain closure and bond level counts of SparseIMPO (src/operators/mpohamiltonian.jl:57-88,
test/operators/mpo.jl:109-113) and the Context's padding of states and environments into the
library's square form."""
import numpy as np
import pytest

import infinitempsalgorithms_amd as imps

SX = np.array([[0.0, 1.0], [1.0, 0.0]])
SZ = np.diag([1.0, -1.0])
I2 = np.eye(2)


def tensor(wl, wr, d, blocks):
    W = np.zeros((wl, d, wr, d), order="F")
    for (i, j), B in blocks.items():
        W[i, :, j, :] = B
    return W


def dimer_chain(h=0.3, J=1.0):
    W1 = tensor(2, 3, 2, {(0, 0): I2, (1, 2): I2, (0, 2): -h * SZ, (0, 1): -J * SX})
    W2 = tensor(3, 2, 2, {(0, 0): I2, (2, 1): I2, (0, 1): -h * SZ, (1, 1): SX})
    W3 = tensor(2, 2, 2, {(0, 0): I2, (1, 1): I2, (0, 1): -h * SZ})
    return [W1, W2, W3]


def test_rectangular_chain_levels_and_closure():
    W1, W2, W3 = dimer_chain()
    H = imps.SparseIMPO([W1, W2, W3])
    assert [H.bonddim(l) for l in range(3)] == [3, 2, 2]
    assert H.max_bonddim() == 3 and H.w == 3 and H.d == 2 and not H.square
    with pytest.raises(ValueError):
        H.bonddim()
    with pytest.raises(ValueError, match="level chain mismatch"):
        imps.SparseIMPO([W1, W1, W3])
    bad = W3.copy(order="F")
    bad[1, :, 1, :] = 2 * I2
    with pytest.raises(ValueError, match="unit levels"):
        imps.SparseIMPO([W1, W2, bad])
    sq = imps.SparseIMPO([tensor(3, 3, 2, {(0, 0): I2, (2, 2): I2})] * 2)
    assert sq.square and sq.bonddim() == 3


def test_heterogeneous_physical_dims():
    Ws = [tensor(2, 2, d, {(0, 0): np.eye(d), (1, 1): np.eye(d), (0, 1): np.diag(np.arange(d, dtype=float))})
          for d in (2, 3, 2)]
    H = imps.SparseIMPO(Ws)
    assert H.phydims() == [2, 3, 2] and H.bonddim() == 2 and H.d == 3 and not H.square
    # the batched sweep takes square chains only
    assert not imps.api.device_batch_eligible([(H, [np.zeros((16, d, 16)) for d in (2, 3, 2)])])
    with pytest.raises(ValueError):
        imps.find_groundstates_batched([(H, [np.zeros((16, d, 16)) for d in (2, 3, 2)])], mode="threads")


def test_context_padding_round_trip():
    # the padding helpers only; no library call
    c = imps.Context.__new__(imps.Context)
    c.N, c.w, c.d, c.D = 3, 3, 3, 4
    c.dims, c.levels = [2, 3, 2], [3, 2, 2]
    rng = np.random.default_rng(0)
    x = np.asfortranarray(rng.standard_normal((4, 2, 4)))
    p = c.pad_site(0, x)
    assert p.shape == (4, 3, 4) and np.array_equal(p[:, :2, :], x) and not p[:, 2:, :].any()
    assert np.array_equal(c.unpad_site(0, p), x)
    assert c.unpad_site(1, np.ones((4, 3, 4))).shape == (4, 3, 4)
    O = c.pad_operator(SX)
    assert O.shape == (3, 3) and np.array_equal(O[:2, :2], SX) and not O[2, :].any() and not O[:, 2].any()
    # a bond profile: tensors padded to the largest bond dimension and back
    c.chi = [2, 4, 3]
    y = np.asfortranarray(rng.standard_normal((3, 2, 2)))  # site 0: bonds chi[2] = 3, chi[0] = 2
    q = c.pad_site(0, y)
    assert q.shape == (4, 3, 4) and np.array_equal(q[:3, :2, :2], y) and np.count_nonzero(q) == y.size
    assert np.array_equal(c.unpad_site(0, q), y)
    Cb = np.asfortranarray(rng.standard_normal((4, 4)))
    assert np.array_equal(c.unpad_bond(1, c.pad_bond(1, Cb)), Cb)
    Cs = np.asfortranarray(rng.standard_normal((2, 2)))
    assert c.pad_bond(0, Cs).shape == (4, 4) and np.array_equal(c.unpad_bond(0, c.pad_bond(0, Cs)), Cs)
    c.chi = None
    # bond levels: the closing level goes to w - 1
    assert c.bond_levels(0) == [0, 1, 2] and c.bond_levels(1) == [0, 2] and c.bond_levels(-1) == [0, 2]
