"""Distribution tests of the oracle's two samplers, pinned by what the
mathematics fixes rather than by a second copy of the code:

- prob (generate_streamlines_cuda.cu:141-250): the step draw is the inverse
  CDF of the thresholded, cone-masked ODF at the point, so vertex counts over
  many streams follow that PMF (chi-square), zero-mass vertices are never
  drawn, the returned direction is the vertex flipped into the incoming
  hemisphere, and the bisection + 32-entry window selection equals
  "smallest i with cdf[i] > u" for every draw that does not land exactly on
  a CDF value;
- PTT (ptt.cu:232-410): a candidate passes when first_val + 8 probe taps
  reach PROBE_QUALITY * PMF_THRESHOLD_P = 0.4, so on a uniform ODF of
  amplitude a (support 9a) a = 0.044 yields nothing and a = 0.045 tracks;
  candidates are a face drawn with probability proportional to its value
  (the sum of its 3 vertex supports, 0 if one is cut) and a point uniform in
  that triangle; the initial frame keeps the maximum-support candidate and
  the backward/forward frames are right-handed with tangents -dir / +dir
  (ptt.cu:415-474).
"""

import json
import os

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as orc
from oracle.oracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _disc():
    with open(os.path.join(GOLD, "reference_tables.json")) as f:
        t = json.load(f)
    return (np.asarray(t["disc2_vert"], np.float64).reshape(24, 2),
            np.asarray(t["disc2_face"], np.int64).reshape(31, 3))


# --------------------------------------------------------------------------
# prob
# --------------------------------------------------------------------------
def _prob_setup(w):
    from gpustreamlines_amd import workloads as W

    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges)
    seeds = W.seeds_for(w, 200, placement="mask", seed=31)
    counts, dirs = o.phase1(seeds)
    # a seed with two or more peaks, so the cone holds a spread of mass
    k = int(np.argmax(counts >= 2))
    assert counts[k] >= 2
    pt = seeds[k]
    v = w.vertices.astype(np.float64)
    d = v[dirs[k, 0]] + 0.3 * v[dirs[k, 1]]
    d = (d / np.linalg.norm(d)).astype(np.float32)
    return o, pt, d


def _expected_pmf(o, w, pt, d):
    """The thresholded, cone-masked ODF at pt, normalised (float64)."""
    rv, odf = o.trilinear(pt)
    assert rv == 0
    odf = odf.astype(np.float64)
    mx = odf.max()
    odf[odf < np.float32(0.05 * mx)] = 0.0
    cosang = (w.vertices.astype(np.float64) @ d.astype(np.float64))
    cmax = np.float32(np.cos(np.float64(np.float32(w.params["max_angle"]))))
    # no vertex sits on the cone edge within float rounding
    assert np.min(np.abs(cosang - cmax)) > 1e-5
    odf[cosang < cmax] = 0.0
    assert odf.sum() > 0 and (odf > 0).sum() >= 10
    return odf / odf.sum()


def test_prob_draws_follow_the_masked_pmf(small_workloads):
    w = small_workloads["prob"]
    o, pt, d = _prob_setup(w)
    p = _expected_pmf(o, w, pt, d)
    n = 300_000
    idx, dirs = o.prob_sample(pt, d, n, sub0=1)
    assert (idx >= 0).all()
    counts = np.bincount(idx, minlength=len(p)).astype(np.float64)
    # zero-mass vertices (below 5% of the maximum, or outside the cone) are
    # never drawn; a draw landing exactly on a CDF value at a bisection
    # midpoint (probability ~1e-7 per draw) could pick the entry before the
    # window, so at most one such draw is tolerated
    assert counts[p == 0].sum() <= 1
    nz = p > 0
    exp = p[nz] * counts[nz].sum()
    big = exp >= 5
    obs = np.append(counts[nz][big], counts[nz][~big].sum())
    ex = np.append(exp[big], exp[~big].sum())
    keep = ex > 0
    chi = stats.chisquare(obs[keep], ex[keep] * obs[keep].sum() / ex[keep].sum())
    assert chi.pvalue > 1e-4, chi
    # the returned direction is the drawn vertex, flipped into dir's hemisphere
    v = w.vertices[idx]
    same = np.all(dirs == v, axis=1)
    flip = np.all(dirs == -v, axis=1)
    assert (same | flip).all()
    dot = (v.astype(np.float64) @ d.astype(np.float64))
    assert np.array_equal(same, dot > 0)


def test_prob_selection_is_the_inverse_cdf():
    """prob_select == smallest i with cdf[i] > sel.  Integer-valued PMFs make
    the blocked scan exact, so the CDF is non-decreasing and every sel that
    is not a CDF value -- random ones, one ulp either side of every CDF value,
    the 32-entry window edges -- must give exactly searchsorted(right); an
    exact tie may give an index at or before it.  Real-valued PMFs: the
    scan's rounding can make the CDF dip by an ulp, so there the check is
    against the first crossing, for sels not within a few ulps of a CDF
    value."""
    rng = np.random.default_rng(3)
    for n in (1, 2, 31, 32, 33, 63, 64, 65, 100, 724):
        for trial in range(5):
            pmf = rng.integers(0, 1000, n).astype(np.float32)
            pmf[rng.random(n) < 0.4] = 0.0  # flat runs of the CDF
            if trial == 0:
                pmf[:] = 0.0
                pmf[rng.integers(0, n)] = 1.0
            cdf = orc.prefix_sum(pmf)
            assert np.array_equal(cdf, np.cumsum(pmf.astype(np.float64)).astype(np.float32))
            last = cdf[-1]
            if last == 0:
                continue
            sels = [np.float32(u) * last for u in rng.random(300)]
            for c in np.unique(cdf):
                sels += [np.nextafter(c, np.float32(-np.inf)), np.nextafter(c, np.float32(np.inf)), c]
            ties = set(cdf.tolist())
            for sel in sels:
                sel = np.float32(sel)
                if not (0 < sel < last):
                    continue
                got = orc.prob_select(cdf, sel)
                want = int(np.searchsorted(cdf, sel, side="right"))
                if float(sel) in ties:
                    assert 0 <= got <= want, (n, sel, got, want)
                else:
                    assert got == want, (n, trial, float(sel), got, want)
    for n in (31, 100, 724):
        for trial in range(5):
            pmf = rng.random(n).astype(np.float32)
            pmf[rng.random(n) < 0.4] = 0.0
            cdf = orc.prefix_sum(pmf)
            last = cdf[-1]
            tol = 8 * np.spacing(last)
            for u in rng.random(2000):
                sel = np.float32(u) * last
                if not (0 < sel < last) or np.min(np.abs(cdf - sel)) <= tol:
                    continue
                got = orc.prob_select(cdf, sel)
                assert got == int(np.argmax(cdf > sel)), (n, float(sel), got)


# --------------------------------------------------------------------------
# PTT
# --------------------------------------------------------------------------
def _ptt_problem(w, data):
    stop = np.full(w.stop_map.shape, 0.8, np.float32)
    return Oracle(w.params, data, stop, w.vertices, w.edges, ptt_lut=w.ptt_lut)


@pytest.mark.parametrize("amp,passes", [(0.044, False), (0.045, True)])
def test_ptt_support_threshold_on_uniform_odf(small_workloads, amp, passes):
    w = small_workloads["ptt"]
    shape = w.dataf.shape
    o = _ptt_problem(w, np.full(shape, amp, np.float32))
    c = np.array(shape[:3], np.float32) / 2
    frame = np.array([0, 0, 1, 1, 0, 0, 0, 1, 0], np.float32)
    for sub in (1, 33, 65):
        ok, tr = o.ptt_substep(c, frame, init=False, sub0=sub)
        assert bool(ok) == passes
        ok_i, _ = o.ptt_init_frame(c, frame[:3], sub0=sub)
        assert bool(ok_i) == passes
        if passes:
            # every candidate has support 9a: the lowest lane wins the first round
            assert tr["win"] == 0 and tr["round"] == 0
            np.testing.assert_allclose(tr["sup"], 9 * amp, rtol=1e-5)
            np.testing.assert_allclose(tr["vert_pdf"], 9 * amp, rtol=1e-5)


def _ramp(shape):
    """ODF amplitude 0.06 + 0.01 x (all vertices), >= 0.06: every candidate's
    support is >= 0.54, while vertex supports vary with the curve's reach
    along x."""
    x = np.arange(shape[0], dtype=np.float32)
    return np.broadcast_to((0.06 + 0.01 * x)[:, None, None, None], shape).astype(np.float32).copy()


def test_ptt_candidates_follow_the_face_cdf(small_workloads):
    w = small_workloads["ptt"]
    shape = w.dataf.shape
    o = _ptt_problem(w, _ramp(shape))
    mc = float(w.params["max_curvature"])
    verts, faces = _disc()
    pos = np.array([shape[0] / 2, shape[1] / 2, shape[2] / 2], np.float32)
    # tangent along z, normal along x: k1 bends the probe curves along the ramp
    frame = np.array([0, 0, 1, 1, 0, 0, 0, 1, 0], np.float32)
    ks, vp = [], None
    for call in range(4000):
        ok, tr = o.ptt_substep(pos, frame, init=False, sub0=32 * call + 1)
        assert ok == 1 and tr["round"] == 0 and tr["win"] == 0
        ks.append(tr["k"].astype(np.float64))
        if vp is None:
            vp = tr["vert_pdf"].astype(np.float64)
        else:
            assert np.array_equal(vp, tr["vert_pdf"])  # the vertex phase draws nothing (non-init)
    ks = np.concatenate(ks)
    assert vp.std() > 0.01 * vp.mean(), "vertex supports should vary with the curve"
    fv = vp[faces].sum(axis=1) * (vp[faces] != 0).all(axis=1)
    pf = fv / fv.sum()
    # locate each candidate in the triangulation of the scaled disc
    tri = verts[faces] * mc  # [31, 3, 2]
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    m = np.stack([b - a, c - a], axis=2)  # [31, 2, 2]
    minv = np.linalg.inv(m)
    rel = ks[:, None, :] - a[None]  # [N, 31, 2]
    rs = np.einsum("fij,nfj->nfi", minv, rel)  # r1, r2 per face
    inside = (rs[..., 0] >= -1e-5) & (rs[..., 1] >= -1e-5) & (rs.sum(-1) <= 1 + 1e-5)
    assert inside.any(axis=1).all(), "a candidate lies outside every disc face"
    face = np.argmax(inside, axis=1)
    counts = np.bincount(face, minlength=31)
    chi = stats.chisquare(counts, pf * counts.sum())
    assert chi.pvalue > 1e-4, (chi, counts, np.round(pf * counts.sum()))
    # uniform in the triangle: r1 and r2 each have density 2(1 - x) on [0, 1]
    r = rs[np.arange(len(face)), face]
    for j in range(2):
        ks_ = stats.kstest(np.clip(r[:, j], 0, 1), lambda x: 1 - (1 - x) ** 2)
        assert ks_.pvalue > 1e-4, (j, ks_)


def test_ptt_initial_frames(small_workloads):
    """init_frame_ptt_d (ptt.cu:415-474): the first initial sub-step tries
    +dir; on success its frame (the winning probe's frame after its probe
    curve) is negated in tangent and normal for the backward pass and kept
    for the forward pass.  Both frames are orthonormal and right-handed
    (negating T and N keeps T x N = B)."""
    from gpustreamlines_amd import workloads as W

    w = small_workloads["ptt"]
    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, ptt_lut=w.ptt_lut)
    seeds = W.seeds_for(w, 60, placement="mask", seed=8)
    counts, dirs = o.phase1(seeds)
    plus = minus = 0
    for s in range(len(seeds)):
        for i in range(counts[s]):
            d = w.vertices[dirs[s, i]].astype(np.float32)
            sub = 32 * (s * 8 + i) + 1
            ok, f = o.ptt_init_frame(seeds[s], d, sub0=sub)
            if not ok:
                continue
            back, fwd = f[:9].reshape(3, 3).astype(np.float64), f[9:].reshape(3, 3).astype(np.float64)
            for fr in (back, fwd):
                np.testing.assert_allclose(fr @ fr.T, np.eye(3), atol=2e-5)
                np.testing.assert_allclose(np.cross(fr[0], fr[1]), fr[2], atol=2e-5)  # right-handed
            np.testing.assert_array_equal(fwd[:2], -back[:2])
            np.testing.assert_array_equal(fwd[2], back[2])
            frame = np.zeros(9, np.float32)
            frame[:3] = d
            ok1, tr = o.ptt_substep(seeds[s], frame, init=True, sub0=sub)
            if ok1:  # +dir succeeded: its frame is the forward one
                plus += 1
                np.testing.assert_array_equal(f[9:], tr["frame"])
            else:  # the -dir attempt's frame is the backward one
                minus += 1
    assert plus >= 20


def test_ptt_init_keeps_the_maximum_support(small_workloads):
    from gpustreamlines_amd import workloads as W

    w = small_workloads["ptt"]
    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, ptt_lut=w.ptt_lut)
    seeds = W.seeds_for(w, 40, placement="mask", seed=9)
    counts, dirs = o.phase1(seeds)
    won = 0
    for s in range(len(seeds)):
        if counts[s] == 0:
            continue
        d = w.vertices[dirs[s, 0]].astype(np.float32)
        frame = np.zeros(9, np.float32)
        frame[:3] = d
        ok, tr = o.ptt_substep(seeds[s], frame, init=True, sub0=32 * s + 1)
        sup = tr["sup"].astype(np.float64)
        if ok:
            won += 1
            assert sup[tr["win"]] == sup.max() >= 0.4 - 1e-7
            assert (sup[: tr["win"]] < sup.max()).all()  # lowest lane among equal maxima
            fr = tr["frame"].reshape(3, 3).astype(np.float64)  # the winner's frame after its probe curve
            np.testing.assert_allclose(fr @ fr.T, np.eye(3), atol=2e-5)
        else:
            assert tr["win"] == -1
    assert won >= 10
