"""Unit pins of the oracle's building blocks against the reference's data
tables (golden fixtures) and against closed forms / brute force."""

import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def tables():
    with open(os.path.join(GOLD, "reference_tables.json")) as f:
        return json.load(f)


def apply_network(swap, keys, vals):
    """The reference's warp_sort<.., WSORT_DIR_DEC> rule (cuwsort.cuh:130-166)
    applied layer by layer with its swap table."""
    k = np.array(keys, dtype=np.float32)
    v = np.array(vals, dtype=np.int32)
    for row in swap:
        nk, nv = k.copy(), v.copy()
        for g, src in enumerate(row):
            a, b = k[src], v[src]
            if not (g < src):
                if a < k[g]:
                    nk[g], nv[g] = a, b
            else:
                if a > k[g]:
                    nk[g], nv[g] = a, b
        k, v = nk, nv
    return k, v


def test_sort_network_is_reference_table(tables):
    assert O.sortnet_table() == tables["swap32"]


def test_warp_sort_descending_distinct():
    rng = np.random.default_rng(0)
    for _ in range(50):
        k = rng.permutation(32).astype(np.float32) + rng.random(32).astype(np.float32)
        v = np.arange(32, dtype=np.int32)
        sk, sv = O.warp_sort_dec(k, v)
        assert np.array_equal(sk, np.sort(k)[::-1])
        assert np.array_equal(k[sv], sk)


def test_warp_sort_ties_follow_reference_network(tables):
    rng = np.random.default_rng(1)
    for _ in range(50):
        k = rng.integers(0, 6, 32).astype(np.float32)
        v = rng.permutation(32).astype(np.int32)
        got = O.warp_sort_dec(k, v)
        want = apply_network(tables["swap32"], k, v)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_prefix_sum_exact_on_integers():
    rng = np.random.default_rng(2)
    for n in (1, 5, 31, 32, 33, 64, 100, 724):
        a = rng.integers(0, 1000, n).astype(np.float32)
        assert np.array_equal(O.prefix_sum(a), np.cumsum(a).astype(np.float32))


def blocked_scan(a):
    """Lane structure of prefix_sum_sh_d (utils.cu:81-120) in numpy."""
    a = a.astype(np.float32).copy()
    for j in range(0, len(a), 32):
        if j:
            a[j] = np.float32(a[j] + a[j - 1])
        blk = a[j:j + 32].copy()
        i = 1
        while i < 32:
            old = blk.copy()
            blk[i:] = (blk[i:] + old[:-i]).astype(np.float32)
            i *= 2
        a[j:j + 32] = blk
    return a


def test_prefix_sum_blocked_order():
    rng = np.random.default_rng(3)
    a = rng.random(724).astype(np.float32)
    assert np.array_equal(O.prefix_sum(a), blocked_scan(a))
    np.testing.assert_allclose(O.prefix_sum(a), np.cumsum(a.astype(np.float64)), rtol=1e-5)


def test_disc_tables_match_reference(tables):
    v = np.array(tables["disc2_vert"]).reshape(24, 2)
    f = np.array(tables["disc2_face"]).reshape(31, 3)
    assert f.min() == 0 and f.max() == 23
    assert np.all(np.linalg.norm(v, axis=1) <= 1.0 + 1e-6)
    src = open(os.path.join(os.path.dirname(GOLD), "..", "oracle", "cuslines_oracle.cpp")).read()
    import re
    body = src[src.index("const float DISC_VERT"):src.index("const int DISC_FACE")]
    vals = [float(x) for x in re.findall(r"-?\d+\.\d+", body)]
    assert np.allclose(vals, tables["disc2_vert"], atol=0, rtol=0)
    fbody = src[src.index("const int DISC_FACE"):]
    fbody = fbody[fbody.index("{") + 1:fbody.index("}")]
    assert [int(x) for x in re.findall(r"\d+", fbody)] == tables["disc2_face"]


def test_propagator_straight_line():
    p = O.prepare_propagator(0.0, 0.0, 0.25)
    assert np.array_equal(p, np.array([0.25, 0, 0, 1, 0, 0, 0, 0, 1], dtype=np.float32))
    frame = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], dtype=np.float32)
    f, d = O.propagate_frame(p, frame)
    assert np.allclose(d, [0.25, 0, 0]) and np.allclose(f, frame)


def test_propagator_arc_geometry():
    """Constant curvature k: the chord of an arc of length s is
    (sin(ks)/k, (1-cos(ks))/k) in the (T, N) plane (Frenet)."""
    k1, s = 2.0, 0.25
    p = O.prepare_propagator(k1, 0.0, s)
    k1e, k2e = k1, 1e-4  # |k2| < K_SMALL is raised to K_SMALL
    k = np.hypot(k1e, k2e)
    assert np.isclose(p[0], np.sin(k * s) / k, rtol=1e-6)
    assert np.isclose(p[1], k1e * (1 - np.cos(k * s)) / k ** 2, rtol=1e-5)
    rng = np.random.default_rng(4)
    for _ in range(20):
        a, b = rng.normal(size=3), rng.normal(size=3)
        t = a / np.linalg.norm(a)
        n = b - b.dot(t) * t
        n /= np.linalg.norm(n)
        frame = np.concatenate([t, n, np.cross(t, n)]).astype(np.float32)
        p = O.prepare_propagator(*rng.uniform(-4, 4, 2), 0.25)
        f, d = O.propagate_frame(p, frame)
        F = f.reshape(3, 3).astype(np.float64)
        assert np.allclose(F @ F.T, np.eye(3), atol=2e-6)
        assert np.allclose(np.cross(F[0], F[1]), F[2], atol=2e-6)
        assert np.linalg.norm(d) <= 0.25 + 1e-6


def test_ptt_taps_are_non_negative(small_workloads):
    """The premise of the early stop of rejection-round probes: on the
    prepared (clipped) ODF every tap, inside the volume, in the border cells
    and beyond them (clamp and tile bleed), is >= 0, so a candidate's running
    support never decreases (ptt.cu:156-192 with ALLOW_WEAK_LINK, ptt.cuh:8)."""
    w = small_workloads["ptt"]
    o = O.Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, ptt_lut=w.ptt_lut)
    dims = np.array([w.params[k] for k in ("dimx", "dimy", "dimz")], np.float64)
    rng = np.random.default_rng(4)
    assert w.dataf.min() >= 0
    vals = []
    for _ in range(3000):
        pos = rng.uniform(-3.0, dims + 2.0).astype(np.float32)
        f = rng.normal(size=3)
        f /= np.linalg.norm(f)
        frame = np.zeros(9, np.float32)
        frame[:3] = f
        vals.append(o.interp4(pos, frame))
    vals = np.array(vals)
    assert vals.min() >= 0.0
    assert (vals > 0).mean() > 0.1
