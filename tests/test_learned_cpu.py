"""Learned PTT proposal on the CPU: the float64 training objective against
finite differences, bf16 rounding, the oracle's network against float64, its
inputs against their definition, and the supports it is trained to predict
against the plain getter's own measurement."""

import numpy as np

from gpustreamlines_amd import learned as LP
from oracle import oracle as O


def _data(n=64, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, LP.N_IN)), rng.uniform(0, 6, size=(n, LP.N_OUT))


def test_gradient_matches_finite_differences():
    p = LP.init_params(3).astype(np.float64)
    X, T = _data()
    g = LP.grad_np(p, X, T)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(s, e), 8, replace=False) for s, e in
                          [(LP.W1, LP.B1), (LP.B1, LP.W2), (LP.W2, LP.B2), (LP.B2, LP.W3), (LP.W3, LP.B3),
                           (LP.B3, LP.NPARAM)]])
    h = 1e-6
    for i in idx:
        q = p.copy(); q[i] += h
        lp = LP.loss_np(q, X, T)
        q[i] -= 2 * h
        lm = LP.loss_np(q, X, T)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_round_trip_and_ties():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.5, 0.0, 1e-3], np.float32)
    back = LP.from_bf16(LP.to_bf16(x))
    assert back[0] == 1.0 and back[3] == -3.5 and back[4] == 0.0
    assert back[1] == 1.0            # halfway between 1 and 1+2^-7: ties to even
    assert back[2] == np.float32(1.0078125)
    assert abs(back[5] - 1e-3) < 1e-3 * 2 ** -8


def test_oracle_network_matches_float64():
    w = LP.to_bf16(LP.init_params(5))
    X, _ = _data(8, seed=2)
    for x in X.astype(np.float32):
        got = O.mlp_forward(w, x)
        want = LP.forward_np(LP.from_bf16(w), x[None].astype(np.float64))[2][0]
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)


def test_features_and_supports_on_a_uniform_odf(small_workloads):
    """On a constant ODF (every texel 1 inside the volume) every tap is 1:
    each input of a vertex is 2, the ODF at the point is 1, and every measured
    vertex support is 1 + 8 probes = 9 (ptt.cu:156-192)."""
    w = small_workloads["ptt"]
    data = np.ones_like(w.dataf)
    o = O.Oracle(w.params, data, w.stop_map, w.vertices, w.edges, ptt_lut=w.ptt_lut)
    pos = np.array([16.0, 18.0, 15.0], np.float32)
    frame = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    d = np.array([1, 0, 0], np.float32)
    x = o.mlp_features(pos, d, frame)
    np.testing.assert_array_equal(x[:24], 2.0)
    assert x[24] == 1.0 and (x[25:28] == 0).all()
    np.testing.assert_array_equal(x[28:31], d)
    np.testing.assert_array_equal(x[31:], frame)
    np.testing.assert_array_equal(o.vertex_supports(pos, frame), 9.0)


def test_features_follow_the_vertex_curves(small_workloads):
    """Input v is the sum of the ODF taps 4 and 8 probe steps along vertex
    v's curve: a linear ODF ramp along x (same on every vertex channel) makes
    it the ramp at those two points, which the propagator fixes."""
    w = small_workloads["ptt"]
    X = w.params["dimx"]
    ramp = (np.arange(X, dtype=np.float32) / X)[:, None, None, None]
    data = np.broadcast_to(ramp, w.dataf.shape).astype(np.float32).copy()
    o = O.Oracle(w.params, data, w.stop_map, w.vertices, w.edges, ptt_lut=w.ptt_lut)
    pos = np.array([12.0, 18.0, 15.0], np.float32)
    frame = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    x = o.mlp_features(pos, frame[:3], frame)
    mc, ps = np.float32(w.params["max_curvature"]), np.float32(w.params["probe_step_size"])
    import json
    import os

    from oracle.oracle import prepare_propagator, propagate_frame

    disc = np.array(json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                                "reference_tables.json")))["disc2_vert"], np.float64).reshape(24, 2)
    for v in range(24):
        k1, k2 = np.float32(disc[v, 0] * np.float64(mc)), np.float32(disc[v, 1] * np.float64(mc))
        prop = prepare_propagator(k1, k2, np.float32(4) * ps)
        f = frame.copy()
        q = pos.astype(np.float64).copy()
        want = 0.0
        for _ in range(2):
            f, dv = propagate_frame(prop, f)
            q += dv
            # unnormalised texture coordinates put texel i's centre at i + 0.5
            want += (q[0] - 0.5) / X
        assert abs(x[v] - want) < 1e-3, (v, x[v], want)
