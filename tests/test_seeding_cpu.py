"""random_seeds_from_mask (gpustreamlines_amd.seeding), DIPY's signature
restated: counts per voxel or in total, every seed inside its voxel,
uniform positions within the voxels, the affine, determinism and the
argument checks.  DIPY is absent, so its own random points are not
compared (parity of the streams unpinned; the distribution is pinned)."""

import numpy as np
import pytest
from scipy import stats

from gpustreamlines_amd.seeding import random_seeds_from_mask


def _mask():
    m = np.zeros((6, 7, 5), bool)
    m[1:4, 2:6, 1:3] = True
    m[5, 0, 4] = True
    return m


def test_per_voxel_counts_and_positions():
    m = _mask()
    s = random_seeds_from_mask(m, np.eye(4), seeds_count=40, seed_count_per_voxel=True, random_seed=3)
    where = np.argwhere(m)
    assert s.shape == (40 * len(where), 3) and s.dtype == np.float64
    vox = np.floor(s + 0.5).astype(int)
    np.testing.assert_array_equal(vox, np.repeat(where, 40, axis=0))  # voxel order, each inside its voxel
    frac = (s - vox + 0.5).ravel()
    assert stats.kstest(frac, "uniform").pvalue > 1e-3  # uniform within the voxel


def test_total_count_spread_evenly():
    m = _mask()
    nvox = int(m.sum())
    for n in (1, nvox - 1, nvox, 5 * nvox + 7):
        s = random_seeds_from_mask(m, np.eye(4), seeds_count=n, seed_count_per_voxel=False, random_seed=n)
        assert len(s) == n
        vox = np.floor(s + 0.5).astype(int)
        assert m[tuple(vox.T)].all()
        counts = np.unique(vox, axis=0, return_counts=True)[1]
        assert counts.max() <= n // nvox + 1 and (len(counts) == nvox) == (n >= nvox)
        if n >= nvox:
            assert counts.min() >= n // nvox


def test_affine_and_determinism():
    m = _mask()
    A = np.array([[2.0, 0, 0, -10], [0, 0, 1.5, 4], [0, -1.0, 0, 7], [0, 0, 0, 1]])
    s = random_seeds_from_mask(m, np.eye(4), seeds_count=3, random_seed=11)
    t = random_seeds_from_mask(m, A, seeds_count=3, random_seed=11)
    np.testing.assert_allclose(t, s @ A[:3, :3].T + A[:3, 3])
    np.testing.assert_array_equal(s, random_seeds_from_mask(m, np.eye(4), seeds_count=3, random_seed=11))
    assert not np.array_equal(s, random_seeds_from_mask(m, np.eye(4), seeds_count=3, random_seed=12))


def test_argument_checks_and_empty_mask():
    with pytest.raises(ValueError):
        random_seeds_from_mask(np.ones((2, 2, 2, 2)), np.eye(4))
    with pytest.raises(ValueError):
        random_seeds_from_mask(_mask(), np.eye(4), seeds_count=0)
    assert random_seeds_from_mask(np.zeros((3, 3, 3)), np.eye(4), seeds_count=5).shape == (0, 3)
