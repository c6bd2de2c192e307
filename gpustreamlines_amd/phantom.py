"""Synthetic diffusion phantoms and seeds for the tracking workloads.

The reference's benchmark data (Stanford HARDI, fetched by DIPY) is not
available offline, so the workloads are defined on a seeded multi-tensor
phantom: 1-3 fibre populations per voxel whose orientations follow smooth
random vector fields, Rician noise, a DTI fit for FA, a white-matter mask
``FA > 0.2`` and the FA map as the stopping map.  Everything is a pure
function of its arguments (``numpy.random.default_rng`` seeds).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .sphere import GradientTable, multi_shell_gtab, single_shell_gtab

# Spatial frequencies (cycles per field of view) of the random fields.
DIR_FREQ = 3.0
DENS_FREQ = 4.0
DENS_OFF = -0.5

LAMBDA_PAR = 1.7e-3
LAMBDA_PERP = 0.3e-3
D_TISSUE = 0.8e-3
D_CSF = 3.0e-3


@dataclass
class Phantom:
    dwi: np.ndarray  # float32 [X,Y,Z,T]
    gtab: GradientTable
    fa: np.ndarray  # float32 [X,Y,Z]
    mask: np.ndarray  # bool [X,Y,Z], FA > 0.2


def _smooth_field(rng, shape, ncomp, nmodes=6, max_freq=2.5):
    """Sum of ``nmodes`` random low-frequency plane waves per component."""
    grid = np.stack(np.meshgrid(*[np.linspace(0, 1, s, dtype=np.float32) for s in shape],
                                indexing="ij"), axis=-1)
    out = np.zeros(shape + (ncomp,), dtype=np.float32)
    for c in range(ncomp):
        for _ in range(nmodes):
            k = rng.uniform(-max_freq, max_freq, size=3).astype(np.float32) * 2 * np.pi
            ph = np.float32(rng.uniform(0, 2 * np.pi))
            a = np.float32(rng.normal())
            out[..., c] += a * np.sin(grid @ k + ph)
    return out


def _unit(v):
    n = np.linalg.norm(v, axis=-1, keepdims=True)
    return v / np.maximum(n, 1e-12)


def dti_fa(dwi, gtab, min_signal=1.0, chunk=131072):
    """FA of a log-linear least-squares tensor fit, per voxel (float32)."""
    g = gtab.bvecs
    b = gtab.bvals
    X = np.stack([np.ones_like(b),
                  -b * g[:, 0] ** 2, -b * g[:, 1] ** 2, -b * g[:, 2] ** 2,
                  -2 * b * g[:, 0] * g[:, 1], -2 * b * g[:, 0] * g[:, 2],
                  -2 * b * g[:, 1] * g[:, 2]], axis=1)
    P = np.linalg.pinv(X)
    flat = dwi.reshape(-1, dwi.shape[-1])
    fa = np.empty(flat.shape[0], dtype=np.float32)
    for s in range(0, flat.shape[0], chunk):
        y = np.log(np.maximum(flat[s : s + chunk].astype(np.float64), min_signal))
        p = y @ P.T
        D = np.empty((p.shape[0], 3, 3))
        D[:, 0, 0], D[:, 1, 1], D[:, 2, 2] = p[:, 1], p[:, 2], p[:, 3]
        D[:, 0, 1] = D[:, 1, 0] = p[:, 4]
        D[:, 0, 2] = D[:, 2, 0] = p[:, 5]
        D[:, 1, 2] = D[:, 2, 1] = p[:, 6]
        ev = np.linalg.eigvalsh(D)
        md = ev.mean(axis=1, keepdims=True)
        num = np.sqrt(((ev - md) ** 2).sum(axis=1))
        den = np.sqrt((ev ** 2).sum(axis=1))
        fa[s : s + chunk] = np.sqrt(1.5) * np.where(den > 0, num / np.where(den > 0, den, 1), 0)
    return np.clip(fa, 0.0, 1.0).reshape(dwi.shape[:-1])


def make_phantom(shape=(81, 106, 76), gtab: GradientTable | None = None, S0=1000.0,
                 snr=20.0, seed=0, max_fibers=3) -> Phantom:
    """Multi-tensor phantom of ``shape`` with Rician noise ``sigma = S0/snr``."""
    if gtab is None:
        gtab = single_shell_gtab()
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    # Brain ellipsoid and a smooth white-matter density inside it.
    c = [np.linspace(-1, 1, s, dtype=np.float32) for s in shape]
    xx, yy, zz = np.meshgrid(*c, indexing="ij")
    brain = (xx ** 2 + yy ** 2 + zz ** 2) < 0.92 ** 2
    dens = _smooth_field(rng, shape, 1, nmodes=5, max_freq=DENS_FREQ)[..., 0]
    dens = 1.0 / (1.0 + np.exp(-2.5 * (dens + DENS_OFF)))
    wm = np.where(brain, dens, 0.0).astype(np.float32)
    # Fibre orientations: smooth fields, decorrelated populations.
    dirs = [_unit(_smooth_field(rng, shape, 3, nmodes=4, max_freq=DIR_FREQ)) for _ in range(max_fibers)]
    nf_field = _smooth_field(rng, shape, 2, nmodes=4, max_freq=1.5)
    nfib = 1 + (nf_field[..., 0] > 0.3).astype(np.int32)
    if max_fibers > 2:
        nfib += ((nf_field[..., 1] > 0.9) & (nfib == 2)).astype(np.int32)
    bvecs = gtab.bvecs.astype(np.float32)
    bvals = gtab.bvals.astype(np.float32)
    nt = len(bvals)
    sig = np.empty(shape + (nt,), dtype=np.float32)
    iso_t = np.exp(-bvals * D_TISSUE).astype(np.float32)
    iso_c = np.exp(-bvals * D_CSF).astype(np.float32)
    for x0 in range(shape[0]):
        aniso = np.zeros(shape[1:] + (nt,), dtype=np.float32)
        nf = nfib[x0]
        for k in range(max_fibers):
            cosg = dirs[k][x0] @ bvecs.T  # [Y,Z,T]
            att = np.exp(-bvals * (LAMBDA_PERP + (LAMBDA_PAR - LAMBDA_PERP) * cosg * cosg))
            aniso += np.where((nf > k)[..., None], att, 0.0)
        aniso /= nf[..., None].astype(np.float32)
        w = wm[x0][..., None]
        inside = (1.0 - w) * iso_t + w * aniso
        sig[x0] = np.where(brain[x0][..., None], S0 * inside, 0.3 * S0 * iso_c)
    sigma = np.float32(S0 / snr)
    n1 = rng.standard_normal(sig.shape, dtype=np.float32)
    n2 = rng.standard_normal(sig.shape, dtype=np.float32)
    dwi = np.sqrt((sig + sigma * n1) ** 2 + (sigma * n2) ** 2).astype(np.float32)
    fa = np.where(brain, dti_fa(dwi, gtab), np.float32(0.0)).astype(np.float32)
    return Phantom(dwi=dwi, gtab=gtab, fa=fa, mask=fa > 0.2)


def make_large_phantom(seed=0) -> Phantom:
    """145x174x145 multi-shell phantom (6 b0 + 3x90 DWI, SNR 25)."""
    return make_phantom((145, 174, 145), gtab=multi_shell_gtab(), snr=25.0, seed=seed)


def seeds_in_mask(mask, n, seed=1):
    """``n`` float32 seeds uniform over ``mask`` voxels with U[-0.5,0.5)^3 jitter
    (voxel coordinates, the reference's seed space)."""
    rng = np.random.default_rng(seed)
    idx = np.argwhere(mask)
    if len(idx) == 0:
        raise ValueError("empty seed mask")
    pick = idx[rng.integers(0, len(idx), size=n)]
    jit = rng.uniform(-0.5, 0.5, size=(n, 3))
    return (pick + jit).astype(np.float32)


def mixed_seeds(fa, mask, n, seed=1, low=(0.05, 0.15)):
    """Half the seeds in ``mask``, half in voxels with ``low[0] <= FA <= low[1]``
    (the uneven-length workload: step counts from 1 to several hundred)."""
    a = seeds_in_mask(mask, (n + 1) // 2, seed=seed)
    border = (fa >= low[0]) & (fa <= low[1])
    b = seeds_in_mask(border, n // 2, seed=seed + 1)
    out = np.empty((n, 3), dtype=np.float32)
    out[0::2] = a
    out[1::2] = b
    return out
