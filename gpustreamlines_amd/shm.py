"""Spherical-harmonic model matrices for the tracking getters, without DIPY.

The reference builds its bootstrap inputs from DIPY
(``cuslines/boot_utils.py:11-94``): ``real_sh_descoteaux(legacy=True)``,
``OpdtModel._fit_matrix``, ``CsaOdfModel._fit_matrix``, ``hat`` and
``lcr_matrix``.  DIPY is not a dependency here, so each of those is restated
from its published closed form:

* legacy Descoteaux 2007 real SH basis: column ``j`` of order ``n`` (even) and
  degree ``m`` is ``sqrt(2)*Re(Y_n^|m|)`` for ``m<0``, ``Y_n^0`` for ``m=0``
  and ``sqrt(2)*Im(Y_n^m)`` for ``m>0``; columns are ordered by ``n`` then
  ``m=-n..n`` (``Y`` is the complex harmonic with Condon-Shortley phase,
  polar angle ``theta``, azimuth ``phi``);
* Laplace-Beltrami regularised pseudo-inverse ``smooth_pinv(B, sqrt(l)*L)``
  with ``L = -n(n+1)``;
* Funk-Radon factors ``F = P_n(0)``;
* CSA fit ``(F*L)/(8 pi) * invB`` (Aganj et al. 2010), OPDT fit
  ``delta_b = F*L*invB``, ``delta_q = 4*F*invB`` (Tristan-Vega et al. 2010);
* hat matrix ``H = U U^T`` (thin SVD of ``B``) and leveraged centred
  residuals ``R = (I-H)/sqrt(1-diag(H)) - column means`` (Chung et al. 2006).

DIPY is absent, so agreement with its numbers is not tested directly; the
tests pin each matrix against the closed form above (harmonics of orders 0-4
written out, the fits' inverse relation to the basis, the leveraged
residuals).  The kernels consume whatever matrices are passed, and both the
GPU path and the CPU oracle receive the same arrays.
"""

from __future__ import annotations

import numpy as np
import scipy.special as sps


def cart2sphere(x, y, z):
    """Return ``(r, theta, phi)`` with polar ``theta`` and azimuth ``phi``."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    r = np.sqrt(x * x + y * y + z * z)
    with np.errstate(invalid="ignore", divide="ignore"):
        theta = np.arccos(np.where(r > 0, z / np.where(r > 0, r, 1.0), 1.0))
    phi = np.arctan2(y, x)
    return r, theta, phi


def sph_harm_ind_list(sh_order_max: int, full_basis: bool = False):
    """Degree ``m`` and order ``n`` of every basis column, in basis order."""
    step = 1 if full_basis else 2
    n_range = np.arange(0, sh_order_max + 1, step, dtype=int)
    n_list = np.repeat(n_range, n_range * 2 + 1)
    m_list = np.concatenate([np.arange(-n, n + 1) for n in n_range])
    return m_list, n_list


def real_sh_descoteaux(sh_order_max, theta, phi, full_basis=False, legacy=True):
    """Legacy Descoteaux real SH basis evaluated at ``(theta, phi)``.

    Returns ``(B, m, n)`` with ``B`` of shape ``(len(theta), ncoef)``.
    """
    if not legacy:
        raise NotImplementedError("only the legacy Descoteaux basis is used by the getters")
    m, n = sph_harm_ind_list(sh_order_max, full_basis)
    theta = np.asarray(theta, dtype=np.float64)[:, None]
    phi = np.asarray(phi, dtype=np.float64)[:, None]
    y = sps.sph_harm_y(n[None, :], np.abs(m)[None, :], theta, phi)
    real = np.where(m[None, :] > 0, y.imag, y.real)
    real = real * np.where(m[None, :] == 0, 1.0, np.sqrt(2.0))
    return real, m, n


def smooth_pinv(B, L):
    """Regularised pseudo-inverse of ``B`` with diagonal penalty ``L``."""
    inv = np.linalg.pinv(np.concatenate((B, np.diag(L))))
    return inv[:, : len(B)]


def hat(B):
    """Hat (projection) matrix of the column space of ``B``."""
    U, _, _ = np.linalg.svd(B, full_matrices=False)
    return U @ U.T


def lcr_matrix(H):
    """Leveraged, centred residual matrix of a hat matrix ``H``."""
    H = np.asarray(H, dtype=np.float64)
    if H.ndim != 2 or H.shape[0] != H.shape[1]:
        raise ValueError("H should be a square matrix")
    d = H.diagonal()
    lev = np.sqrt(1.0 - d, where=d < 1, out=np.ones_like(d))[:, None]
    R = (np.eye(len(H)) - H) / lev
    return R - R.mean(0)


def _basis_and_penalty(bvecs_dwi, sh_order_max):
    _, theta, phi = cart2sphere(*np.asarray(bvecs_dwi, dtype=np.float64).T)
    B, m, n = real_sh_descoteaux(sh_order_max, theta, phi)
    L = -n * (n + 1.0)
    F = sps.eval_legendre(n, 0.0)
    return B, L, F


def csa_fit_matrix(bvecs_dwi, sh_order_max=6, sh_lambda=0.006):
    """CSA ODF fit matrix (ncoef x ndwi); coefficient 0 is replaced by
    ``0.5/sqrt(pi)`` when the fit is applied (``cuslines/cuda_c/boot.cu:363-368``)."""
    B, L, F = _basis_and_penalty(bvecs_dwi, sh_order_max)
    invB = smooth_pinv(B, np.sqrt(sh_lambda) * L)
    return (F * L)[:, None] / (8.0 * np.pi) * invB


def opdt_fit_matrices(bvecs_dwi, sh_order_max=6, sh_lambda=0.006):
    """OPDT fit matrices ``(delta_b, delta_q)``, each ncoef x ndwi."""
    B, L, F = _basis_and_penalty(bvecs_dwi, sh_order_max)
    invB = smooth_pinv(B, np.sqrt(sh_lambda) * L)
    delta_b = (F * L)[:, None] * invB
    delta_q = 4.0 * F[:, None] * invB
    return delta_b, delta_q


def hat_and_lcr(bvecs_dwi, sh_order_max=6):
    """``(H, R)`` for residual bootstrap (``cuslines/boot_utils.py:86-94``)."""
    _, theta, phi = cart2sphere(*np.asarray(bvecs_dwi, dtype=np.float64).T)
    B, _, _ = real_sh_descoteaux(sh_order_max, theta, phi)
    H = hat(B)
    return H, lcr_matrix(H)


CSA_C0 = 0.28209479177387814  # 0.5 / sqrt(pi), the CSA zeroth coefficient


def csa_odf(dwi, b0s_mask, fit_matrix, sampling_matrix, min_signal=1.0, chunk=65536):
    """CSA ODF amplitudes on the sphere for every voxel of ``dwi`` (float32).

    Follows the CSA recipe the bootstrap getter applies per voxel
    (``cuslines/cuda_c/boot.cu:283-327,497-534``) in float64 on the host:
    ``max(min_signal)``, divide by the b0 mean, clamp to [0.001, 0.999],
    ``coef = fit . log(-log(x))``, ``coef[0] = 0.5/sqrt(pi)``,
    ``odf = S . coef``.
    """
    shp = dwi.shape
    flat = dwi.reshape(-1, shp[-1])
    out = np.empty((flat.shape[0], sampling_matrix.shape[0]), dtype=np.float32)
    fm = np.asarray(fit_matrix, dtype=np.float64)
    sm = np.asarray(sampling_matrix, dtype=np.float64)
    b0s = np.asarray(b0s_mask, dtype=bool)
    for s in range(0, flat.shape[0], chunk):
        d = np.maximum(flat[s : s + chunk].astype(np.float64), min_signal)
        b0 = d[:, b0s].mean(axis=1, keepdims=True)
        x = np.clip(d[:, ~b0s] / b0, 0.001, 0.999)
        coef = np.log(-np.log(x)) @ fm.T
        coef[:, 0] = CSA_C0
        out[s : s + chunk] = (coef @ sm.T).astype(np.float32)
    return out.reshape(shp[:-1] + (sampling_matrix.shape[0],))
