"""Antipodally symmetric unit spheres and gradient tables, without DIPY.

The reference tracks on DIPY's bundled ``repulsion724`` sphere
(``run_gpu_streamlines.py:320``): 362 electrostatically dispersed hemisphere
points plus their antipodes, with the convex-hull edge list (2,166 edges for
724 vertices).  DIPY is absent, so the same construction is restated here:
a Fibonacci hemisphere lattice relaxed by antipodally symmetric Coulomb
repulsion, completed with its antipodes, and the unique edges of the convex
hull.  The vertex order is ``[hemisphere; -hemisphere]``.
"""

from __future__ import annotations

import numpy as np
from scipy.spatial import ConvexHull

GOLDEN_ANGLE = np.pi * (3.0 - np.sqrt(5.0))


def fibonacci_hemisphere(n: int) -> np.ndarray:
    """``n`` quasi-uniform unit vectors with ``z > 0``."""
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (i + 0.5) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * GOLDEN_ANGLE
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def disperse_hemisphere(pts: np.ndarray, iters: int = 60, step: float = 0.2) -> np.ndarray:
    """Relax hemisphere points under Coulomb repulsion from the point set and
    its antipodes (deterministic gradient steps, re-projected on the sphere)."""
    p = pts.copy()
    n = len(p)
    for _ in range(iters):
        allp = np.concatenate([p, -p])
        d = p[:, None, :] - allp[None, :, :]
        r2 = np.einsum("ijk,ijk->ij", d, d)
        r2[np.arange(n), np.arange(n)] = np.inf
        f = (d / (r2[..., None] ** 1.5)).sum(axis=1)
        f -= (f * p).sum(axis=1, keepdims=True) * p  # tangential part
        fn = np.linalg.norm(f, axis=1, keepdims=True)
        scale = step * np.sqrt(4.0 * np.pi / (2 * n)) / max(fn.max(), 1e-30)
        p = p + scale * f
        p /= np.linalg.norm(p, axis=1, keepdims=True)
        p[p[:, 2] < 0] *= -1.0
    return p


def hull_edges(vertices: np.ndarray) -> np.ndarray:
    """Unique undirected edges ``(u < v)`` of the convex hull, sorted."""
    tri = ConvexHull(vertices).simplices
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [0, 2]]])
    e.sort(axis=1)
    e = np.unique(e, axis=0)
    return e.astype(np.int32)


def symmetric_sphere(n_vertices: int = 724, iters: int = 60):
    """``(vertices float32 [n,3], edges int32 [m,2])`` of an antipodal sphere."""
    if n_vertices % 2:
        raise ValueError("an antipodally symmetric sphere has an even vertex count")
    hemi = disperse_hemisphere(fibonacci_hemisphere(n_vertices // 2), iters=iters)
    verts = np.concatenate([hemi, -hemi]).astype(np.float32)
    edges = hull_edges(verts.astype(np.float64))
    return verts, edges


class GradientTable:
    """Minimal gradient table: b-values, unit b-vectors and the b0 mask."""

    def __init__(self, bvals, bvecs, b0_threshold: float = 50.0):
        self.bvals = np.asarray(bvals, dtype=np.float64)
        self.bvecs = np.asarray(bvecs, dtype=np.float64)
        self.b0s_mask = self.bvals <= b0_threshold

    @property
    def gradients(self):
        return self.bvecs * self.bvals[:, None]


def single_shell_gtab(n_b0=10, n_dirs=160, bval=2000.0):
    """``n_b0`` b0 volumes first, then ``n_dirs`` Fibonacci-hemisphere DWIs."""
    dirs = fibonacci_hemisphere(n_dirs)
    bvecs = np.concatenate([np.zeros((n_b0, 3)), dirs])
    bvals = np.concatenate([np.zeros(n_b0), np.full(n_dirs, bval)])
    return GradientTable(bvals, bvecs)


def multi_shell_gtab(n_b0=6, n_per_shell=90, shells=(1000.0, 2000.0, 3000.0)):
    """``n_b0`` b0 volumes, then one rotated Fibonacci hemisphere per shell."""
    bvecs = [np.zeros((n_b0, 3))]
    bvals = [np.zeros(n_b0)]
    for k, b in enumerate(shells):
        d = fibonacci_hemisphere(n_per_shell)
        a = k * GOLDEN_ANGLE / len(shells)
        rot = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
        bvecs.append(d @ rot.T)
        bvals.append(np.full(n_per_shell, b))
    return GradientTable(np.concatenate(bvals), np.concatenate(bvecs))
