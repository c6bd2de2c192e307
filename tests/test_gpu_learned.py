"""Learned PTT proposal on the GPU: bit-exact against the oracle's restatement
for fixed weights (points and proposal counters, envelope misses included),
the training gradient against the float64 objective, a short training run
on pairs the plain getter records, what the trained network gains over an
untrained one, and the tracking statistics of the exactly corrected learned
sampler against the plain getter on the same seeds."""

import numpy as np
import pytest

from conftest import packed_points

pytestmark = pytest.mark.gpu


def _tracker(w):
    from gpustreamlines_amd._device import DeviceTracker

    dt = DeviceTracker(w.params, device=0)
    dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, ptt_lut=w.ptt_lut)
    return dt


@pytest.mark.parametrize("flat", [True, False])
def test_learned_bit_exact_vs_oracle(small_workloads, flat):
    """Both stepping modes: one loop of sub-steps per slot (the learned
    proposal's per-pass state kept across the loop) and per-seed nesting."""
    from gpustreamlines_amd import learned as LP
    from gpustreamlines_amd import workloads as W
    from oracle.oracle import Oracle

    w = small_workloads["ptt"]
    seeds = W.seeds_for(w, 48, placement="mixed")
    wts = LP.to_bf16(LP.init_params(11))
    print(f"[learned] bit-exact, 48 seeds, flat={flat}", flush=True)
    dt = _tracker(w)
    try:
        dt.set_stepping(flat)
        dt.set_proposal(wts)
        g = dt.track(seeds, keep_dirs=True)
        gstats = dt.proposal_counters()
    finally:
        dt.close()
    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, ptt_lut=w.ptt_lut)
    o.set_learned(wts)
    r = o.track(seeds)
    ostats = o.proposal_counters()
    assert gstats.sum() > 0 and gstats[0] > 0
    assert np.array_equal(gstats, ostats), (gstats, ostats)
    assert np.array_equal(g["sline_len"], r["sline_len"])
    assert np.array_equal(g["points"], packed_points(r))


def test_training_gradient_matches_float64(small_workloads):
    from gpustreamlines_amd import learned as LP

    w = small_workloads["ptt"]
    rng = np.random.default_rng(0)
    n = 512
    X = rng.normal(size=(n, LP.N_IN)).astype(np.float32)
    T = rng.uniform(0, 6, size=(n, LP.N_OUT)).astype(np.float32)
    pairs = np.zeros((n, LP.PAIR), np.float32)
    pairs[:, :LP.N_IN] = X
    pairs[:, LP.N_IN:LP.N_IN + LP.N_OUT] = T
    p = LP.init_params(4)
    dt = _tracker(w)
    try:
        pd = dt.alloc(pairs.nbytes)
        pd.upload(pairs)
        tr = LP.Trainer(dt, p, batch=n)
        loss, g = tr.loss_grad(pd.ptr, n)
        tr.free()
        pd.free()
    finally:
        dt.close()
    pr = LP.from_bf16(LP.to_bf16(p)).astype(np.float64)
    want_loss = LP.loss_np(pr, X.astype(np.float64), T.astype(np.float64))
    want = LP.grad_np(pr, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - want_loss) < 1e-4 * max(1.0, abs(want_loss))
    for s, e in [(LP.W1, LP.B1), (LP.B1, LP.W2), (LP.W2, LP.B2), (LP.B2, LP.W3), (LP.W3, LP.B3), (LP.B3, LP.NPARAM)]:
        scale = np.abs(want[s:e]).max()
        np.testing.assert_allclose(g[s:e], want[s:e], atol=2e-3 * scale, rtol=0)


@pytest.fixture(scope="module")
def trained(small_workloads):
    """Pairs recorded by the plain getter on 2,000 seeds, 4 epochs of Adam."""
    from gpustreamlines_amd import learned as LP
    from gpustreamlines_amd import workloads as W

    w = small_workloads["ptt"]
    seeds = W.seeds_for(w, 2000, placement="mask", seed=5)
    print("[learned] recording + training", flush=True)
    dt = _tracker(w)
    try:
        pairs, n = LP.record_pairs(dt, seeds, capacity=400_000, rate=0.25)
        host = pairs.download(np.zeros((n, LP.PAIR), np.float32))
        tr = LP.Trainer(dt, batch=4096)
        l0 = tr.loss_grad(pairs.ptr, 4096)[0]
        hist = tr.fit(pairs, n, epochs=4)
        params = tr.params()
        tr.free()
        pairs.free()
    finally:
        dt.close()
    print(f"[learned] {n} pairs, loss {l0:.4f} -> {hist}", flush=True)
    return dict(n=n, pairs=host, loss0=l0, hist=hist, params=params)


def test_recorded_pairs_are_well_formed(small_workloads, trained):
    """Recorded pairs hold finite inputs, unit frame rows and supports within
    [0, 9] (first tap plus 8 probes of an ODF normalised to 1 per voxel)."""
    P = trained["pairs"]
    assert trained["n"] > 50_000
    assert np.isfinite(P).all()
    f = P[:, 31:40].reshape(-1, 3, 3).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(f, axis=2), 1.0, atol=1e-4)
    T = P[:, 40:64]
    assert T.min() >= 0.0 and T.max() <= 9.0 + 1e-4
def test_training_lowers_loss_and_envelope_misses(small_workloads, trained):
    """What the trained network gains over the method without it (an
    untrained network): the squared error of the predicted supports falls by
    more than 10x, and the face draws whose exact value breaks the envelope
    of the exact correction (F > M g, the only draws where the learned
    sampler can depart from the reference's face law) fall to a small
    fraction of the untrained network's.  Both keep nearly every sub-step in
    the learned rounds: the correction accepts any proposal."""
    from gpustreamlines_amd import learned as LP
    from gpustreamlines_amd import workloads as W

    assert trained["hist"][-1] < 0.1 * trained["loss0"]
    w = small_workloads["ptt"]
    seeds = W.seeds_for(w, 1500, placement="mixed", seed=9)
    miss, share = {}, {}
    dt = _tracker(w)
    try:
        for name, p in (("untrained", LP.init_params()), ("trained", trained["params"])):
            dt.set_proposal(LP.to_bf16(p))
            dt.track(seeds)
            s = dt.proposal_counters()
            share[name] = s[0] / max(1, s[:3].sum())
            miss[name] = s[3] / max(1, s[0] + s[1])
            print(f"[learned] {name}: accepted {s[0]} fell back {s[1]} plain {s[2]} envelope misses {s[3]} "
                  f"({miss[name]:.4f} per learned sub-step)", flush=True)
    finally:
        dt.close()
    assert share["trained"] > 0.95
    assert miss["trained"] < 0.05
    assert miss["trained"] < 0.2 * miss["untrained"]


def _density(r, shape):
    idx = np.clip(np.rint(r["points"]).astype(int), 0, np.array(shape) - 1)
    d = np.zeros(shape)
    np.add.at(d, tuple(idx.T), 1)
    return d.ravel()


@pytest.fixture(scope="module")
def stats_runs(small_workloads, trained):
    """Same seeds: the plain getter, the plain getter under another rng_seed
    (the noise floor of every statistic), and the learned proposal."""
    import dataclasses

    from gpustreamlines_amd import learned as LP
    from gpustreamlines_amd import workloads as W

    w = small_workloads["ptt"]
    seeds = W.seeds_for(w, 30000, placement="mixed", seed=13)
    res = {}
    dt = _tracker(w)
    try:
        res["plain"] = dt.track(seeds)
        dt.set_proposal(LP.to_bf16(trained["params"]))
        res["learned"] = dt.track(seeds)
    finally:
        dt.close()
    dt = _tracker(dataclasses.replace(w, params=dict(w.params, rng_seed=int(w.params["rng_seed"]) + 1)))
    try:
        res["plain_other_seed"] = dt.track(seeds)
    finally:
        dt.close()
    return res


def test_learned_count_and_density_match_plain(small_workloads, stats_runs):
    """Streamline count within 2 % and voxel visit density of all points
    correlated > 0.98 with the plain getter's (the plain getter's own
    correlation under another random seed is printed beside it)."""
    shape = tuple(small_workloads["ptt"].params[k] for k in ("dimx", "dimy", "dimz"))
    a, b, c = stats_runs["plain"], stats_runs["learned"], stats_runs["plain_other_seed"]
    na, nb = len(a["sline_len"]), len(b["sline_len"])
    da = _density(a, shape)
    r_learned = np.corrcoef(da, _density(b, shape))[0, 1]
    r_floor = np.corrcoef(da, _density(c, shape))[0, 1]
    print(f"[learned] streamlines {na} vs {nb}; density r learned {r_learned:.5f}, plain under another seed "
          f"{r_floor:.5f}", flush=True)
    assert abs(na - nb) <= 0.02 * na
    assert r_learned > 0.98


def test_learned_length_distribution_matches_plain(stats_runs):
    """Two-sample KS of the streamline lengths, plain vs learned, against
    the bar of p > 0.05 (plain vs plain under another seed printed beside)."""
    from scipy.stats import ks_2samp

    a, b, c = stats_runs["plain"], stats_runs["learned"], stats_runs["plain_other_seed"]
    ks = ks_2samp(a["sline_len"], b["sline_len"])
    ks0 = ks_2samp(a["sline_len"], c["sline_len"])
    print(f"[learned] mean length {a['sline_len'].mean():.1f} vs {b['sline_len'].mean():.1f}; KS p={ks.pvalue:.3g} "
          f"(plain vs plain under another seed: p={ks0.pvalue:.3g})", flush=True)
    assert ks.pvalue > 0.05


def test_learned_million_mixed_seeds():
    """The learned proposal on the 1M-seed uneven-length workload (81x106x76,
    half the seeds in FA > 0.2, half in 0.05-0.15 border voxels), trained on
    pairs the plain getter records from 20,000 mask seeds: the same
    streamlines per seed as the plain getter (phase 1 does not use the
    proposal), lengths within the slot, finite points, a small share of
    envelope misses, and length and density statistics of the plain
    getter's (KS p > 0.01, r > 0.98)."""
    from scipy.stats import ks_2samp

    from gpustreamlines_amd import learned as LP
    from gpustreamlines_amd import workloads as W

    w = W.build("ptt")
    seeds = W.seeds_for(w, 1_000_000, placement="mixed")
    print("[learned] 1M mixed: recording + training", flush=True)
    dt = _tracker(w)
    try:
        pairs, n = LP.record_pairs(dt, W.seeds_for(w, 20000, placement="mask", seed=7), capacity=3_000_000,
                                   rate=0.1, rng_chunk=25000)
        tr = LP.Trainer(dt, batch=4096)
        tr.fit(pairs, n, epochs=2)
        params = tr.params()
        tr.free()
        pairs.free()
        print("[learned] 1M mixed: plain pass", flush=True)
        plain = dt.track(seeds, rng_chunk=25000)
        dt.set_proposal(LP.to_bf16(params))
        dt.proposal_counters(reset=True)
        print("[learned] 1M mixed: learned pass", flush=True)
        got = dt.track(seeds, rng_chunk=25000)
        cnt = dt.proposal_counters(reset=True)
    finally:
        dt.close()
    lens, lp = got["sline_len"].astype(np.int64), plain["sline_len"].astype(np.int64)
    assert np.array_equal(got["counts"], plain["counts"])
    assert len(lens) == len(lp) and lens.min() >= 1 and lens.max() <= 2 * w.params["max_sline_len"] - 1
    assert np.isfinite(got["points"]).all()
    shape = tuple(w.params[k] for k in ("dimx", "dimy", "dimz"))
    r = np.corrcoef(_density(plain, shape), _density(got, shape))[0, 1]
    ks = ks_2samp(lp, lens)
    miss = cnt[3] / max(1, cnt[0] + cnt[1])
    print(f"[learned] 1M mixed: mean length {lp.mean():.2f} vs {lens.mean():.2f}, KS p {ks.pvalue:.3g}, "
          f"density r {r:.5f}, envelope misses per learned sub-step {miss:.2e}", flush=True)
    assert miss < 0.05
    assert ks.pvalue > 0.01
    assert r > 0.98
