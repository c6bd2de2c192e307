"""The oracle's residual-bootstrap refit (boot.cu:393-545) against closed
forms: an identity bootstrap of a signal generated from known SH
coefficients must return those coefficients (CSA), the OPDT transform's zero
and linear terms, the resampling index order of the 32 lane streams, and the
orientation of the H and R products."""

import numpy as np
import pytest

from oracle.oracle import Oracle, philox_draws

HR, NB0, DN, NV = 40, 2, 6, 50
S0 = 1000.0
C0 = 0.28209479177387814


def _problem(model, H, R, delta_b, delta_q, smat, dwi, rng_seed=7):
    dimt = NB0 + HR
    shape = (3, 3, 3)
    data = np.empty(shape + (dimt,), dtype=np.float32)
    data[..., :NB0] = S0
    data[..., NB0:] = dwi
    rng = np.random.default_rng(3)
    verts = rng.normal(size=(NV, 3))
    verts = (verts / np.linalg.norm(verts, axis=1, keepdims=True)).astype(np.float32)
    edges = np.array([[0, 1]], dtype=np.int32)
    p = dict(dimx=3, dimy=3, dimz=3, dimt=dimt, model_type=model, nverts=NV, nedges=1,
             max_sline_len=501, step_size=0.5, delta_nr=DN, rng_seed=rng_seed, min_signal=1.0)
    b0 = np.zeros(dimt, dtype=np.int32)
    b0[:NB0] = 1
    boot = dict(H=H, R=R, delta_b=delta_b, delta_q=delta_q, sampling_matrix=smat, b0s_mask=b0)
    return Oracle(p, data, np.ones(shape, np.float32), verts, edges, boot=boot)


def _basis(seed=0):
    rng = np.random.default_rng(seed)
    B = rng.normal(size=(HR, DN))
    smat = rng.normal(size=(NV, DN))
    return B, smat


PT = np.array([1, 1, 1], dtype=np.float32)  # a voxel centre: trilinear is exact there


def test_csa_identity_bootstrap_recovers_coefficients():
    """H = I, R = 0: the bootstrap is the signal itself.  With
    x = exp(-exp(B c)) the CSA transform log(-log x) is B c, delta_q = B^+
    returns c, coefficient 0 is replaced by 1/(2 sqrt(pi)) (boot.cu:363-368),
    and the ODF is smat @ c."""
    B, smat = _basis()
    c = np.random.default_rng(1).uniform(-0.4, 0.4, DN)
    z = np.clip(B @ c, -3.0, 1.0)
    c = np.linalg.lstsq(B, z, rcond=None)[0]
    x = np.exp(-np.exp(B @ c))
    assert x.min() > 0.001 and x.max() < 0.999
    o = _problem(1, np.eye(HR), np.zeros((HR, HR)), np.zeros((DN, HR)), np.linalg.pinv(B), smat,
                 (S0 * x).astype(np.float32))
    rv, odf, sig = o.boot_odf(PT)
    assert rv == 0
    np.testing.assert_allclose(sig[NB0:], x, rtol=2e-6)
    np.testing.assert_array_equal(sig[:NB0], 1.0)
    want_c = c.copy()
    want_c[0] = C0
    want = smat @ want_c
    np.testing.assert_allclose(odf, want, atol=2e-4 * np.abs(want).max())


def test_opdt_transform_zero_and_linear_term():
    """OPDT (boot.cu:240-281): coefficients = delta_q @ t(x) - delta_b @ x with
    t(x) = -ln x (1.5 + ln x) x.  At ln x = -1.5, t vanishes and only the
    linear term remains; at ln x = -1 with delta_b = 0, t = x/2."""
    B, smat = _basis(2)
    rng = np.random.default_rng(4)
    db = rng.normal(size=(DN, HR))
    dq = rng.normal(size=(DN, HR))
    x = np.full(HR, np.exp(-1.5))
    o = _problem(0, np.eye(HR), np.zeros((HR, HR)), db, dq, smat, (S0 * x).astype(np.float32))
    rv, odf, sig = o.boot_odf(PT)
    xs = sig[NB0:].astype(np.float64)
    want = smat @ (-(db @ xs))
    np.testing.assert_allclose(odf, want, atol=1e-4 * np.abs(want).max())
    x = np.full(HR, np.exp(-1.0))
    o = _problem(0, np.eye(HR), np.zeros((HR, HR)), np.zeros((DN, HR)), dq, smat, (S0 * x).astype(np.float32))
    rv, odf, sig = o.boot_odf(PT)
    xs = sig[NB0:].astype(np.float64)
    want = smat @ (dq @ (0.5 * xs))
    np.testing.assert_allclose(odf, want, atol=1e-4 * np.abs(want).max())


@pytest.mark.parametrize("slot,phase", [(0, 0), (5, 1)])
def test_bootstrap_resampling_order(slot, phase):
    """H = 0, R = I: element j of the bootstrap is d[perm_j], perm_j the
    (j // 32)-th draw of lane j % 32 modulo HR (boot.cu:479-486), each lane on
    Philox subsequence 32*slot + lane + phase; then min_signal clip and
    division by the b0 mean, exactly."""
    B, smat = _basis()
    dwi = np.random.default_rng(5).uniform(-50, 900, HR).astype(np.float32)
    o = _problem(1, np.zeros((HR, HR)), np.eye(HR), np.zeros((DN, HR)), np.zeros((DN, HR)), smat, dwi)
    rv, odf, sig = o.boot_odf(PT, slot=slot, phase=phase)
    draws = [philox_draws(7, 32 * slot + lane + phase, (HR + 31) // 32) for lane in range(32)]
    perm = np.array([draws[j % 32][j // 32] % HR for j in range(HR)])
    want = np.maximum(dwi[perm], np.float32(1.0)) / np.float32(S0)
    np.testing.assert_array_equal(sig[NB0:], want.astype(np.float32))


def test_hat_and_residual_products_are_row_major():
    """h = H d + (R d)[perm] with H, R indexed [row][column]: a non-symmetric H
    and R = 0 give H @ d, not H.T @ d."""
    B, smat = _basis()
    rng = np.random.default_rng(6)
    H = np.triu(rng.uniform(0, 0.05, (HR, HR)))
    dwi = rng.uniform(100, 900, HR).astype(np.float32)
    o = _problem(1, H, np.zeros((HR, HR)), np.zeros((DN, HR)), np.zeros((DN, HR)), smat, dwi)
    rv, odf, sig = o.boot_odf(PT)
    want = np.maximum(H.astype(np.float32).astype(np.float64) @ dwi, 1.0) / S0
    np.testing.assert_allclose(sig[NB0:], want, rtol=1e-5)
    assert not np.allclose(sig[NB0:], np.maximum(H.T @ dwi, 1.0) / S0, rtol=1e-3)
