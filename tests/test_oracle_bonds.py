"""Pins the kd-tree bond build restatement (core/kdtree.h,
core/simulator.cu:825-943) against brute force."""
import numpy as np
import pytest

import oracle as orc


def lattice(nx, ny, nz, h, jitter, seed):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    g = g.transpose(2, 1, 0, 3).reshape(-1, 3).astype(np.float64) * h  # x fastest
    g += rng.uniform(-1, 1, g.shape) * jitter * h
    x4 = np.zeros((g.shape[0], 4))
    x4[:, :3] = g
    return x4


@pytest.mark.parametrize("seed,jit", [(1, 0.05), (2, 0.0), (3, 0.3)])
def test_kdtree_equals_brute(seed, jit):
    x4 = lattice(11, 9, 8, 0.01, jit, seed)
    h = 0.03
    c1, l1, dropped = orc.bonds_kdtree(x4, h, max_bonds=528, leaf_size=100)
    c2, l2, total = orc.bonds_brute(x4, [h], max_bonds=528)
    assert dropped == 0
    assert total == c2.sum()
    assert np.array_equal(c1, c2)
    for p in range(x4.shape[0]):
        assert np.array_equal(l1[p, :c1[p]], l2[p, :c2[p]])
        assert np.all(np.diff(l1[p, :c1[p]]) > 0)  # sorted, no duplicates
        assert p not in l1[p, :c1[p]]


def test_kdtree_random_cloud_small_leaves():
    rng = np.random.default_rng(5)
    x4 = np.zeros((1500, 4))
    x4[:, :3] = rng.uniform(0, 1, (1500, 3))
    for leaf in (1, 7, 100):
        c1, l1, _ = orc.bonds_kdtree(x4, 0.12, max_bonds=528, leaf_size=leaf)
        c2, l2, _ = orc.bonds_brute(x4, [0.12], max_bonds=528)
        assert np.array_equal(c1, c2)
        assert np.array_equal(l1, l2)


def test_strict_horizon():
    # Unjittered lattice: offsets at exactly 3h are excluded by the strict <,
    # offsets just inside are kept; the interior count of an h-lattice with
    # horizon 3h is the number of integer vectors with 0 < |v| < 3: 92.
    h = 0.5  # exactly representable spacing so |v| h is exact for axis offsets
    x4 = lattice(9, 9, 9, h, 0.0, 0)
    c, l, _ = orc.bonds_brute(x4, [3 * h])
    centre = 4 + 9 * 4 + 81 * 4
    assert c[centre] == 92
    v = np.array([[i, j, k] for i in range(-3, 4) for j in range(-3, 4) for k in range(-3, 4)])
    nrm2 = (v ** 2).sum(1)
    assert ((nrm2 > 0) & (nrm2 < 9)).sum() == 92


def test_bodies_do_not_bond():
    x4 = lattice(6, 6, 6, 0.01, 0.05, 9)
    body = (np.arange(x4.shape[0]) % 2).astype(np.int32)
    c, l, _ = orc.bonds_brute(x4, [0.03, 0.02], body=body)
    for p in range(x4.shape[0]):
        qs = l[p, :c[p]]
        assert np.all(body[qs] == body[p])
        d = np.linalg.norm(x4[qs, :3] - x4[p, :3], axis=1)
        assert np.all(d < (0.03 if body[p] == 0 else 0.02))


def test_overflow_reported():
    x4 = lattice(6, 6, 6, 0.01, 0.0, 0)
    c, l, total = orc.bonds_brute(x4, [1.0], max_bonds=16)
    assert total == -1
    c1, l1, dropped = orc.bonds_kdtree(x4, 1.0, max_bonds=16)
    assert dropped == x4.shape[0] * (x4.shape[0] - 1) - 16 * x4.shape[0]
