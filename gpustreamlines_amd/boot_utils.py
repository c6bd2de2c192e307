"""Bootstrap getter inputs for the OPDT and CSA models, without DIPY.

Same recipe and return dictionary as cuslines/boot_utils.py:11-94
(``prepare_opdt``/``prepare_csa``): the sampling matrix is the legacy
Descoteaux basis on the sphere, the fit matrices come from the model's
regularised pseudo-inverse on the DWI directions, and ``H``/``R`` are the hat
and leveraged-centred-residual matrices of the basis on those directions.
``gtab`` needs ``bvecs`` (or ``gradients``) and ``b0s_mask``; ``sphere``
needs ``vertices`` or ``theta``/``phi``.
"""

from __future__ import annotations

import numpy as np

from . import shm


def _sphere_angles(sphere):
    if hasattr(sphere, "theta") and hasattr(sphere, "phi"):
        return np.asarray(sphere.theta), np.asarray(sphere.phi)
    v = np.asarray(sphere.vertices if hasattr(sphere, "vertices") else sphere, dtype=np.float64)
    _, theta, phi = shm.cart2sphere(v[:, 0], v[:, 1], v[:, 2])
    return theta, phi


def _dwi_bvecs(gtab):
    b0 = np.asarray(gtab.b0s_mask, dtype=bool)
    g = np.asarray(gtab.bvecs if hasattr(gtab, "bvecs") else gtab.gradients, dtype=np.float64)
    g = g[~b0]
    n = np.linalg.norm(g, axis=1, keepdims=True)
    return g / np.where(n > 0, n, 1.0), b0


def _common(gtab, sphere, sh_order_max, full_basis):
    if full_basis:
        raise NotImplementedError("the getters use the symmetric (even-order) basis")
    theta, phi = _sphere_angles(sphere)
    sampling_matrix, _, _ = shm.real_sh_descoteaux(sh_order_max, theta, phi)
    bvecs, b0 = _dwi_bvecs(gtab)
    H, R = shm.hat_and_lcr(bvecs, sh_order_max)
    return sampling_matrix, bvecs, b0, H, R


def prepare_opdt(gtab, sphere, sh_order_max=6, full_basis=False, sh_lambda=0.006, min_signal=1):
    """Matrices for ``BootDirectionGetter(model_type='OPDT', ...)``."""
    sampling_matrix, bvecs, b0, H, R = _common(gtab, sphere, sh_order_max, full_basis)
    delta_b, delta_q = shm.opdt_fit_matrices(bvecs, sh_order_max, sh_lambda)
    return dict(model_type="OPDT", min_signal=min_signal, H=H, R=R, delta_b=delta_b,
                delta_q=delta_q, sampling_matrix=sampling_matrix, b0s_mask=b0)


def prepare_csa(gtab, sphere, sh_order_max=6, full_basis=False, sh_lambda=0.006, min_signal=1):
    """Matrices for ``BootDirectionGetter(model_type='CSA', ...)``."""
    sampling_matrix, bvecs, b0, H, R = _common(gtab, sphere, sh_order_max, full_basis)
    fit = shm.csa_fit_matrix(bvecs, sh_order_max, sh_lambda)
    return dict(model_type="CSA", min_signal=min_signal, H=H, R=R, delta_b=fit,
                delta_q=fit, sampling_matrix=sampling_matrix, b0s_mask=b0)
