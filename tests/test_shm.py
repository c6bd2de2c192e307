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


def _textbook_real_sh_order_le_4(th, ph):
    """The legacy Descoteaux basis of orders 0, 2 and 4 (m = 0 for order 4)
    written out from the textbook complex harmonics with the Condon-Shortley
    phase, Y_2^1 = -sqrt(15/8pi) sin cos e^{i phi},
    Y_2^2 = sqrt(15/32pi) sin^2 e^{2i phi}, Y_4^0 = 3/(16 sqrt(pi))
    (35 c^4 - 30 c^2 + 3): columns n = 0; n = 2, m = -2..2; the n = 4, m = 0
    column."""
    c, s = np.cos(th), np.sin(th)
    r2 = np.sqrt(2.0)
    y00 = np.full_like(th, 0.5 / np.sqrt(np.pi))
    a21 = np.sqrt(15 / (8 * np.pi)) * s * c
    a22 = np.sqrt(15 / (32 * np.pi)) * s * s
    n2 = [
        r2 * a22 * np.cos(2 * ph),  # m = -2: sqrt2 Re Y_2^2
        -r2 * a21 * np.cos(ph),  # m = -1: sqrt2 Re Y_2^1
        np.sqrt(5 / (16 * np.pi)) * (3 * c * c - 1),  # m = 0
        -r2 * a21 * np.sin(ph),  # m = 1: sqrt2 Im Y_2^1
        r2 * a22 * np.sin(2 * ph),  # m = 2: sqrt2 Im Y_2^2
    ]
    y40 = 3 / (16 * np.sqrt(np.pi)) * (35 * c**4 - 30 * c * c + 3)
    return np.stack([y00] + n2, 1), y40


def test_basis_matches_textbook_harmonics_and_column_order():
    """Orders 0-4 against the harmonics written out by hand: a swapped polar
    and azimuth angle, a dropped Condon-Shortley sign, a signed instead of
    absolute degree for m < 0, or another column order fails here."""
    rng = np.random.default_rng(5)
    v = rng.normal(size=(500, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    _, th, ph = shm.cart2sphere(*v.T)
    np.testing.assert_allclose(np.cos(th), v[:, 2], atol=1e-12)
    np.testing.assert_allclose(np.sin(th) * np.cos(ph), v[:, 0], atol=1e-12)
    B, m, n = shm.real_sh_descoteaux(4, th, ph)
    assert B.shape == (500, 15)
    assert list(n) == [0] + [2] * 5 + [4] * 9 and list(m[1:6]) == [-2, -1, 0, 1, 2]
    low, y40 = _textbook_real_sh_order_le_4(th, ph)
    np.testing.assert_allclose(B[:, :6], low, atol=1e-12)
    np.testing.assert_allclose(B[:, 10], y40, atol=1e-12)
    # in Cartesian form: the m = -1 column is -sqrt(15/4pi) x z, m = 1 is -sqrt(15/4pi) y z
    k = np.sqrt(15 / (4 * np.pi))
    np.testing.assert_allclose(B[:, 2], -k * v[:, 0] * v[:, 2], atol=1e-12)
    np.testing.assert_allclose(B[:, 4], -k * v[:, 1] * v[:, 2], atol=1e-12)


def test_fit_matrices_invert_the_basis_without_penalty():
    """With no Laplace-Beltrami penalty the fits reduce to the plain
    pseudo-inverse: CSA fit . B = diag(F L / 8 pi), OPDT delta_b . B =
    diag(F L), delta_q . B = diag(4 F), and smooth_pinv of a zero penalty is
    numpy's pinv.  A transposed operand or a penalty applied to the wrong
    side fails here."""
    g = single_shell_gtab()
    bv = g.bvecs[~g.b0s_mask]
    _, th, ph = shm.cart2sphere(*bv.T)
    B, m, n = shm.real_sh_descoteaux(6, th, ph)
    L = -n * (n + 1.0)
    F = np.array([1, -0.5, 3 / 8, -5 / 16])[n // 2]
    np.testing.assert_allclose(shm.smooth_pinv(B, np.zeros(28)), np.linalg.pinv(B), atol=1e-10)
    fit = shm.csa_fit_matrix(bv, 6, 0.0)
    np.testing.assert_allclose(fit @ B, np.diag(F * L / (8 * np.pi)), atol=1e-10)
    db, dq = shm.opdt_fit_matrices(bv, 6, 0.0)
    np.testing.assert_allclose(db @ B, np.diag(F * L), atol=1e-9)
    np.testing.assert_allclose(dq @ B, np.diag(4 * F), atol=1e-10)
    # the penalty shrinks high orders: a pure order-6 signal is fitted smaller
    # than a pure order-2 one of the same norm
    inv = shm.smooth_pinv(B, np.sqrt(0.006) * L)
    gain = np.diag(inv @ B)
    assert np.all(gain[n == 6] < gain[n == 2].min()) and np.allclose(gain[0], 1, atol=1e-9)


def test_lcr_matrix_gives_centred_leveraged_residuals():
    """R d is the residual bootstrap's vector of leveraged residuals of d,
    centred: r_i = (d_i - x_i . beta) / sqrt(1 - h_i) with beta the least
    squares fit and h_i = x_i (X^T X)^-1 x_i^T the leverage (Chung et al.
    2006), minus the mean of r; the fit and leverages are computed here by
    least squares and a solve, not through the hat matrix."""
    rng = np.random.default_rng(2)
    X = rng.normal(size=(9, 3))
    R = shm.lcr_matrix(shm.hat(X))
    for _ in range(3):
        d = rng.normal(size=9)
        beta = np.linalg.lstsq(X, d, rcond=None)[0]
        h = np.einsum("ij,ij->i", X, np.linalg.solve(X.T @ X, X.T).T)
        r = (d - X @ beta) / np.sqrt(1 - h)
        np.testing.assert_allclose(R @ d, r - r.mean(), atol=1e-12)
