"""Seeds from a region mask without DIPY.

The reference's driver seeds with DIPY's
``dipy.tracking.utils.random_seeds_from_mask(roi, seeds_count=n,
seed_count_per_voxel=False, random_seed=s, affine=np.eye(4))``
(run_gpu_streamlines.py:306-313).  DIPY is not a dependency here, so this
restates that function's documented behaviour under the same signature:
seeds uniformly distributed inside the mask's voxels (voxel index +
U[-0.5, 0.5) on every axis), ``seeds_count`` in each voxel or in total, in
voxel order, mapped through ``affine``.  The random numbers are this
module's own (numpy's PCG64 seeded with ``random_seed``), so the seeds
follow DIPY's distribution but are not DIPY's points.
"""

from __future__ import annotations

import numpy as np


def random_seeds_from_mask(mask, affine, *, seeds_count=1, seed_count_per_voxel=True, random_seed=None):
    """Random seed points inside the voxels of ``mask``.

    ``seed_count_per_voxel`` True: ``seeds_count`` seeds in every voxel of
    the mask.  False: ``seeds_count`` seeds in total, spread evenly over the
    voxels (each voxel gets ``seeds_count // nvoxels`` or one more, the
    voxels with one more chosen at random).  Returns float64 points
    ``[n, 3]`` in the space of ``affine`` (voxel coordinates for the
    identity), ordered by voxel (``np.argwhere`` order).
    """
    mask = np.asarray(mask)
    if mask.ndim != 3:
        raise ValueError("mask must be a 3D array")
    if int(seeds_count) < 1:
        raise ValueError("seeds_count must be a positive integer")
    seeds_count = int(seeds_count)
    where = np.argwhere(mask.astype(bool))
    nvox = len(where)
    if nvox == 0:
        return np.zeros((0, 3), dtype=np.float64)
    rng = np.random.default_rng(random_seed)
    if seed_count_per_voxel:
        per = np.full(nvox, seeds_count, dtype=np.int64)
    else:
        per = np.full(nvox, seeds_count // nvox, dtype=np.int64)
        extra = seeds_count - int(per.sum())
        if extra:
            per[rng.choice(nvox, size=extra, replace=False)] += 1
    vox = np.repeat(where, per, axis=0).astype(np.float64)
    seeds = vox + rng.random(vox.shape) - 0.5
    affine = np.asarray(affine, dtype=np.float64)
    if affine.shape != (4, 4):
        raise ValueError("affine must be a 4x4 matrix")
    return seeds @ affine[:3, :3].T + affine[:3, 3]
