"""The batched bootstrap kernel (CUSLINES_BOOT_MFMA: 16 seeds in lockstep,
refit and ODF projection as fp32 MFMA products) against the exact kernel and
the CPU oracle on the same seeds and random streams.  Its products sum in a
different order, so the bar is the tolerance one: phase 1 (shared, exact)
equal; the streamline count equal; most streamlines bit-identical (a
streamline is identical up to the first step whose peak choice flips at a
near-tie); the length distributions indistinguishable (two-sample KS p >
0.05); voxel-visit densities correlated above 0.98."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _track(w, seeds, mode, rng_chunk=0, occ=0, full=False):
    from gpustreamlines_amd._device import DeviceTracker

    dt = DeviceTracker(w.params, device=0)
    try:
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot)
        dt.set_boot_mode(mode, occ, full_products=full)
        return dt.track(seeds, rng_chunk=rng_chunk)
    finally:
        dt.close()


@pytest.mark.parametrize("kind", ["csa", "opdt"])
def test_lowrank_form_detected(kind, small_workloads):
    """The getter matrices (H the hat matrix of the order-6 basis, R its
    leveraged centred residuals) have the low-rank form, of rank 28; a
    perturbed R does not, and the full-products flag turns it off."""
    from gpustreamlines_amd._device import DeviceTracker

    w = small_workloads[kind]
    dt = DeviceTracker(w.params, device=0)
    try:
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot)
        dt.set_boot_mode(1)
        assert dt.boot_lowrank() == 28
        dt.set_boot_mode(1, full_products=True)
        assert dt.boot_lowrank() == 0
        bad = dict(w.boot)
        bad["R"] = np.array(bad["R"], dtype=np.float32, copy=True)
        bad["R"][3, 5] += 1e-3
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=bad)
        dt.set_boot_mode(1)
        assert dt.boot_lowrank() == 0
    finally:
        dt.close()


def _density(pts, shape):
    shp = np.array(shape)
    idx = np.clip(np.rint(pts).astype(np.int64), 0, shp - 1)
    lin = (idx[:, 0] * shp[1] + idx[:, 1]) * shp[2] + idx[:, 2]
    return np.bincount(lin, minlength=int(np.prod(shp))).astype(np.float64)


def _compare(a, b, shape, min_identical):
    from scipy.stats import ks_2samp

    assert np.array_equal(a["counts"], b["counts"])
    assert np.array_equal(a["sline_seed"], b["sline_seed"])
    la, lb = a["sline_len"].astype(np.int64), b["sline_len"].astype(np.int64)
    assert len(la) == len(lb) > 0
    pa, pb = a["point_offsets"], b["point_offsets"]
    same = np.array([la[i] == lb[i] and np.array_equal(a["points"][pa[i]:pa[i + 1]], b["points"][pb[i]:pb[i + 1]])
                     for i in range(len(la))])
    ks = ks_2samp(la, lb)
    r = np.corrcoef(_density(a["points"], shape), _density(b["points"], shape))[0, 1]
    print(f"identical {same.mean():.4f} of {len(la)}, mean len {la.mean():.2f} vs {lb.mean():.2f}, "
          f"KS p {ks.pvalue:.3g}, density r {r:.5f}", flush=True)
    assert same.mean() >= min_identical
    assert ks.pvalue > 0.05
    assert r > 0.98
    assert np.isfinite(a["points"]).all()


@pytest.mark.parametrize("kind", ["csa", "opdt"])
@pytest.mark.parametrize("occ,full", [(2, False), (4, False), (4, True)])
def test_mfma_boot_matches_oracle_small(kind, occ, full, small_workloads):
    from gpustreamlines_amd import workloads as W
    from oracle.oracle import Oracle

    w = small_workloads[kind]
    seeds = W.seeds_for(w, 400, placement="mixed", seed=13)
    print(f"[boot-mfma] {kind} occ={occ} full={full} small seeds={len(seeds)}", flush=True)
    g = _track(w, seeds, 1, rng_chunk=0, occ=occ, full=full)
    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot).track(seeds, rng_chunk=0)
    from conftest import packed_points

    lens = o["sline_len"].astype(np.int64)
    ref = {"counts": o["counts"], "sline_seed": o["sline_seed"], "sline_len": o["sline_len"],
           "points": packed_points(o), "point_offsets": np.concatenate([[0], np.cumsum(lens)])}
    shape = tuple(w.params[k] for k in ("dimx", "dimy", "dimz"))
    _compare(g, ref, shape, 0.7)


@pytest.mark.parametrize("kind", ["csa", "opdt"])
def test_mfma_boot_matches_exact_full_volume(kind):
    from gpustreamlines_amd import workloads as W

    w = W.build(kind)
    seeds = W.seeds_for(w, 20000, placement="mask", seed=1)
    print(f"[boot-mfma] {kind} 81x106x76 seeds={len(seeds)}", flush=True)
    a = _track(w, seeds, 1, rng_chunk=25000)
    b = _track(w, seeds, 0, rng_chunk=25000)
    shape = tuple(w.params[k] for k in ("dimx", "dimy", "dimz"))
    _compare(a, b, shape, 0.7)


def test_mfma_boot_uneven_lengths_properties():
    """1M-seed uneven-length workload: whole-output properties."""
    from gpustreamlines_amd import workloads as W

    w = W.build("csa")
    seeds = W.seeds_for(w, 1_000_000, placement="mixed")
    print("[boot-mfma] csa 1M mixed", flush=True)
    g = _track(w, seeds, 1, rng_chunk=25000)
    counts, lens = g["counts"].astype(np.int64), g["sline_len"].astype(np.int64)
    L = w.params["max_sline_len"]
    assert len(lens) == counts.sum() == g["offsets"][-1]
    assert np.array_equal(np.repeat(np.arange(len(seeds)), counts), g["sline_seed"])
    assert lens.min() >= 1 and lens.max() <= 2 * L - 1
    assert (lens == 1).any() and lens.max() > 300
    assert np.isfinite(g["points"]).all()
