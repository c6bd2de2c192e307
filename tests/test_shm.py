"""Closed-form checks of the DIPY-free SH/model matrices (the getters'
inputs): basis orthonormality, hat-matrix algebra, residual centring,
Funk-Radon factors and the CSA constant."""

import numpy as np

from gpustreamlines_amd import shm
from gpustreamlines_amd.sphere import fibonacci_hemisphere, single_shell_gtab, symmetric_sphere


def test_basis_is_orthonormal_on_the_sphere():
    n = 40000
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = i * np.pi * (3 - np.sqrt(5))
    v = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    _, th, phi = shm.cart2sphere(*v.T)
    B, m, nn = shm.real_sh_descoteaux(6, th, phi)
    G = (4 * np.pi / n) * B.T @ B
    assert B.shape[1] == 28
    np.testing.assert_allclose(G, np.eye(28), atol=2e-3)


def test_basis_even_order_is_antipodally_symmetric():
    v, _ = symmetric_sphere(724)
    _, th, phi = shm.cart2sphere(*v.astype(np.float64).T)
    B, _, _ = shm.real_sh_descoteaux(6, th, phi)
    np.testing.assert_allclose(B[:362], B[362:], atol=1e-6)


def test_hat_and_lcr_algebra():
    g = single_shell_gtab()
    bv = g.bvecs[~g.b0s_mask]
    H, R = shm.hat_and_lcr(bv, 6)
    np.testing.assert_allclose(H, H.T, atol=1e-10)
    np.testing.assert_allclose(H @ H, H, atol=1e-10)
    assert np.isclose(np.trace(H), 28)
    np.testing.assert_allclose(R.mean(0), 0, atol=1e-12)
    # R annihilates fitted signals: R (B c) = centred zero vector
    _, th, phi = shm.cart2sphere(*bv.T)
    B, _, _ = shm.real_sh_descoteaux(6, th, phi)
    c = np.random.default_rng(0).normal(size=28)
    np.testing.assert_allclose(R @ (B @ c), 0, atol=1e-9)


def test_funk_radon_and_fit_matrices():
    import scipy.special as sps

    assert np.isclose(sps.eval_legendre(2, 0.0), -0.5)
    assert np.isclose(sps.eval_legendre(4, 0.0), 3 / 8)
    g = single_shell_gtab()
    bv = g.bvecs[~g.b0s_mask]
    fit = shm.csa_fit_matrix(bv, 6, 0.006)
    assert fit.shape == (28, 160)
    np.testing.assert_allclose(fit[0], 0, atol=1e-14)  # L_0 = 0: coefficient 0 comes from the constant
    db, dq = shm.opdt_fit_matrices(bv, 6, 0.006)
    np.testing.assert_allclose(db[0], 0, atol=1e-14)
    assert np.isclose(shm.CSA_C0, 0.5 / np.sqrt(np.pi))


def test_csa_odf_of_isotropic_signal_is_flat():
    g = single_shell_gtab()
    bv = g.bvecs[~g.b0s_mask]
    fit = shm.csa_fit_matrix(bv, 6, 0.006)
    v, _ = symmetric_sphere(724)
    _, th, phi = shm.cart2sphere(*v.astype(np.float64).T)
    S, _, _ = shm.real_sh_descoteaux(6, th, phi)
    dwi = np.where(g.b0s_mask, 1000.0, 1000.0 * np.exp(-2000 * 1e-3)).astype(np.float32)[None, None, None]
    odf = shm.csa_odf(dwi, g.b0s_mask, fit, S)[0, 0, 0]
    np.testing.assert_allclose(odf, odf.mean(), rtol=1e-4)


def test_sphere_edges_are_a_triangulation():
    v, e = symmetric_sphere(724)
    assert v.shape == (724, 3) and e.shape == (2166, 2)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1, atol=1e-6)
    np.testing.assert_allclose(v[:362], -v[362:])
    assert e.min() == 0 and e.max() == 723 and np.all(e[:, 0] < e[:, 1])
    deg = np.bincount(e.ravel(), minlength=724)
    assert deg.min() >= 4 and deg.max() <= 8
    assert len(fibonacci_hemisphere(5)) == 5
