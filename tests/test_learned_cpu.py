"""Learned PTT proposal on the CPU: the float64 training objective against
finite differences, the oracle's network and mixture against closed forms,
and the vMF sampler's statistics."""

import numpy as np
import pytest

from gpustreamlines_amd import learned as LP
from oracle import oracle as O


def _data(n=64, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, LP.N_IN))
    a = rng.uniform(-0.4, 0.4, size=(n, 2))
    U = np.column_stack([np.sqrt(1 - (a ** 2).sum(1)), a])
    return X, U


def test_gradient_matches_finite_differences():
    p = LP.init_params(3).astype(np.float64)
    p[LP.W3:LP.B3] *= 10  # make every output matter
    X, U = _data()
    g = LP.grad_np(p, X, U)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(s, e), 8, replace=False) for s, e in
                          [(LP.W1, LP.B1), (LP.B1, LP.W2), (LP.W2, LP.B2), (LP.B2, LP.W3), (LP.W3, LP.B3),
                           (LP.B3, LP.NPARAM)]])
    h = 1e-6
    for i in idx:
        q = p.copy(); q[i] += h
        lp = LP.loss_np(q, X, U)
        q[i] -= 2 * h
        lm = LP.loss_np(q, X, U)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(fd), (i, fd, g[i])


def test_vmf_normaliser_integrates_to_one():
    """The log-normaliser in the objective: a single component's density
    integrates to 1 over the sphere (midpoint rule in cos-angle)."""
    for kap in (0.05, 1.0, 20.0, 300.0):
        t = np.linspace(-1, 1, 200001)
        logc = np.log(kap) - np.log(2 * np.pi) - kap - np.log1p(-np.exp(-2 * kap))
        f = np.exp(logc + kap * t) * 2 * np.pi
        assert abs(np.trapezoid(f, t) - 1) < 1e-6


def test_bf16_round_trip_and_ties():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.5, 0.0, 1e-3], np.float32)
    w = LP.to_bf16(x)
    back = LP.from_bf16(w)
    assert back[0] == 1.0 and back[3] == -3.5 and back[4] == 0.0
    assert back[1] == 1.0            # halfway between 1 and 1+2^-7: ties to even (1.0)
    assert back[2] == np.float32(1.0078125)
    assert abs(back[5] - 1e-3) < 1e-3 * 2 ** -8


def test_oracle_network_matches_float64():
    p = LP.init_params(5)
    w = LP.to_bf16(p)
    X, _ = _data(8, seed=2)
    for x in X.astype(np.float32):
        got = O.mlp_forward(w, x)
        want = LP.forward_np(LP.from_bf16(w), x[None].astype(np.float64))[2][0]
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-5)


def test_mixture_setup_closed_form():
    rng = np.random.default_rng(7)
    o = rng.normal(size=40).astype(np.float32)
    mix = O.mixture_setup(o)
    oc = o.reshape(8, 5).astype(np.float64)
    m = oc[:, :3] / np.linalg.norm(oc[:, :3], axis=1, keepdims=True)
    np.testing.assert_allclose(mix[:, :3], m, atol=1e-6)
    for c in range(8):
        e1, e2 = mix[c, 3:6], mix[c, 6:9]
        B = np.stack([mix[c, :3], e1, e2])
        np.testing.assert_allclose(B @ B.T, np.eye(3), atol=2e-6)   # orthonormal
        assert np.linalg.det(B) > 0                                  # right-handed
    np.testing.assert_allclose(mix[:, 9], np.exp(np.clip(oc[:, 3], -4, 6)), rtol=1e-6)
    w = np.exp(oc[:, 4] - oc[:, 4].max())
    np.testing.assert_allclose(mix[:, 11], np.cumsum(w), rtol=1e-6)


@pytest.mark.parametrize("kap", [0.5, 5.0, 60.0])
def test_vmf_draws_mean_resultant(kap):
    """One component around the tangent: E[cos angle] = coth k - 1/k for the
    3-D vMF, and (k1, k2) = mc * (normal, binormal) parts stay in the disc."""
    o = np.zeros(40, np.float32).reshape(8, 5)
    o[:, 0] = 1.0
    o[:, 3] = np.log(kap)
    o[1:, 4] = -80.0  # only component 0
    mix = O.mixture_setup(o.ravel())
    mc = 4.0
    k1, k2, ok = O.mixture_draws(mix, mc, 0, 17, 40000)
    r2 = (k1.astype(np.float64) ** 2 + k2 ** 2) / mc ** 2
    assert r2.max() <= 1.0 + 1e-6
    cosang = np.where(ok, np.sqrt(np.clip(1 - r2, 0, 1)), -np.sqrt(np.clip(1 - r2, 0, 1)))
    want = 1 / np.tanh(kap) - 1 / kap
    se = cosang.std() / np.sqrt(len(cosang))
    assert abs(cosang.mean() - want) < 5 * se + 1e-4
