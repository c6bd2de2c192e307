"""GPU parity: the gfx950 path through the C ABI against the CPU oracle on
identical inputs and random streams.  Integer outputs (direction counts,
sphere-vertex indices, seed indices, lengths) and float32 points must be
bit-identical: the kernels and the oracle perform the same float
operations in the same order."""

import numpy as np
import pytest

from conftest import packed_points

pytestmark = pytest.mark.gpu

NSEEDS = {"prob": 256, "ptt": 48, "csa": 128, "opdt": 128}


def _run_both(w, seeds, rng_chunk=0, launch=(0, 0, 0), blocks=True, lanes=0, flat=True):
    from gpustreamlines_amd._device import DeviceTracker
    from oracle.oracle import Oracle

    print(f"[parity] {w.kind} seeds={len(seeds)} chunk={rng_chunk} launch={launch} flat={flat}", flush=True)
    dt = DeviceTracker(w.params, device=0)
    try:
        dt.set_launch(*launch)
        dt.set_ptt_layout(blocks)
        dt.set_lanes_per_seed(lanes)
        dt.set_stepping(flat)
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
        g = dt.track(seeds, rng_chunk=rng_chunk, keep_dirs=True)
    finally:
        dt.close()
    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot,
               ptt_lut=w.ptt_lut).track(seeds, rng_chunk=rng_chunk)
    return g, o


def _assert_same(g, o):
    n = len(g["counts"])
    assert np.array_equal(g["counts"], o["counts"]), "phase-1 direction counts differ"
    for i in range(n):
        c = o["counts"][i]
        assert np.array_equal(g["dirs"][i, :c], o["dirs"][i, :c]), f"seed {i}: peak vertex indices differ"
    assert np.array_equal(g["offsets"], o["offsets"])
    assert np.array_equal(g["sline_seed"], o["sline_seed"]), "streamline seed indices differ"
    bad = np.nonzero(g["sline_len"] != o["sline_len"])[0]
    assert len(bad) == 0, f"{len(bad)} streamline lengths differ, first {bad[:5]}"
    assert np.array_equal(g["points"], packed_points(o)), "streamline points differ"


@pytest.mark.parametrize("kind", ["prob", "csa", "opdt", "ptt"])
def test_track_bit_exact(small_workloads, kind):
    from gpustreamlines_amd import workloads as W

    w = small_workloads[kind]
    seeds = W.seeds_for(w, NSEEDS[kind], placement="mixed")
    g, o = _run_both(w, seeds)
    assert len(o["sline_len"]) > 0
    _assert_same(g, o)


@pytest.mark.parametrize("flat", [False, True])
@pytest.mark.parametrize("kind", ["ptt", "prob", "csa", "opdt"])
def test_stepping_modes_bit_exact(small_workloads, kind, flat):
    """Phase 2 on the queue with the reference's per-seed loop nesting and
    with one flat loop of steps per slot (a slot whose pass ends starts its
    next pass, direction or seed at once); mixed seeds, so slots of a wave
    end their streamlines at different times.  Chunk-relative streams of 40
    seeds; bootstrap on the exact kernel."""
    from gpustreamlines_amd import workloads as W

    w = small_workloads[kind]
    seeds = W.seeds_for(w, 96, placement="mixed", seed=31)
    g, o = _run_both(w, seeds, rng_chunk=40, flat=flat)
    assert len(o["sline_len"]) > len(seeds)
    _assert_same(g, o)


@pytest.mark.parametrize("kind,launch", [("prob", (1, 0, 0)), ("prob", (0, 64, 1)), ("ptt", (1, 128, 0)),
                                         ("csa", (0, 64, 2)), ("opdt", (1, 0, 0))])
def test_launch_configs_bit_exact(small_workloads, kind, launch):
    """Static and persistent-queue schedules, any block size: same output."""
    from gpustreamlines_amd import workloads as W

    w = small_workloads[kind]
    seeds = W.seeds_for(w, NSEEDS[kind] // 2, placement="mixed", seed=11)
    g, o = _run_both(w, seeds, launch=launch)
    _assert_same(g, o)


def test_ptt_texel_gather_path_bit_exact(small_workloads):
    """PTT taps gathered texel by texel (no corner blocks): same output."""
    from gpustreamlines_amd import workloads as W

    w = small_workloads["ptt"]
    seeds = W.seeds_for(w, 32, placement="mixed", seed=12)
    g, o = _run_both(w, seeds, blocks=False)
    _assert_same(g, o)


@pytest.mark.parametrize("flat", [True, False])
@pytest.mark.parametrize("lanes", [8, 16, 32])
def test_ptt_lanes_per_seed_bit_exact(small_workloads, lanes, flat):
    """PTT with 8, 16 or 32 lanes per seed (virtual random streams), flat
    stepping over the spatial order or the per-seed nesting, equals the
    oracle; chunk-relative streams of 24 seeds."""
    from gpustreamlines_amd import workloads as W

    w = small_workloads["ptt"]
    seeds = W.seeds_for(w, 72, placement="mixed", seed=13)
    g, o = _run_both(w, seeds, rng_chunk=24, lanes=lanes, flat=flat)
    _assert_same(g, o)


def test_chunk_relative_streams(small_workloads):
    """rng_chunk numbers seeds within chunks: tracking 2 chunks of 64 in one
    call equals tracking each chunk alone (cu_propagate_seeds.py semantics)."""
    from gpustreamlines_amd import workloads as W

    w = small_workloads["prob"]
    seeds = W.seeds_for(w, 128, placement="mask", seed=7)
    g, o = _run_both(w, seeds, rng_chunk=64)
    _assert_same(g, o)
    g2, _ = _run_both(w, seeds[64:], rng_chunk=0)
    n0 = g["offsets"][64]
    assert np.array_equal(g["sline_len"][n0:], g2["sline_len"])


@pytest.mark.parametrize("kind", ["prob", "ptt", "csa"])
def test_edge_seeds_bit_exact(small_workloads, kind):
    """Seeds outside the volume, on its low-edge half voxel, in zero-ODF /
    low-FA voxels and duplicated seeds: same counts, lengths and points as
    the oracle (outside seeds give no direction)."""
    from gpustreamlines_amd import workloads as W

    w = small_workloads[kind]
    X, Y, Z = (w.params[k] for k in ("dimx", "dimy", "dimz"))
    inside = W.seeds_for(w, 6, placement="mixed", seed=3)
    edge = np.array([[-0.6, 3, 3], [-0.25, 10, 10], [X - 0.51, Y / 2, Z / 2], [X + 4.0, 1, 1],
                     [1, 1, 1], [0, 0, 0]], np.float32)
    seeds = np.concatenate([inside, edge, inside[:2]]).astype(np.float32)
    g, o = _run_both(w, seeds)
    assert g["counts"][len(inside) + 3] == 0  # outside the volume
    _assert_same(g, o)


def test_empty_seed_batch(small_workloads):
    from gpustreamlines_amd._device import DeviceTracker

    w = small_workloads["prob"]
    dt = DeviceTracker(w.params, device=0)
    try:
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges)
        g = dt.track(np.zeros((0, 3), np.float32))
    finally:
        dt.close()
    assert len(g["counts"]) == 0 and len(g["sline_len"]) == 0 and len(g["points"]) == 0


@pytest.mark.parametrize("kind", ["prob", "ptt"])
def test_max_length_truncation_bit_exact(small_workloads, kind):
    """A short max_sline_len truncates streamlines at the same step on both
    sides (generate_streamlines_cuda.cu:412-418 step-count formula)."""
    import dataclasses

    from gpustreamlines_amd import workloads as W

    w0 = small_workloads[kind]
    w = dataclasses.replace(w0, params=dict(w0.params, max_sline_len=12))
    seeds = W.seeds_for(w, 24, placement="mask", seed=4)
    g, o = _run_both(w, seeds)
    assert g["sline_len"].max() <= 2 * 12 - 1
    assert (g["sline_len"] >= 12).any()
    _assert_same(g, o)


@pytest.mark.parametrize("kind", ["prob", "csa"])
def test_peaks_edge_path_bit_exact(small_workloads, kind):
    """Vertex 0 joined to its 9 nearest other vertices has more than 8
    neighbours, so the library takes the edge form of the peak test instead
    of its 8-neighbour table; the oracle sees the same edge list."""
    import dataclasses

    from gpustreamlines_amd import workloads as W

    w = small_workloads[kind]
    d = w.vertices.astype(np.float64) @ w.vertices[0].astype(np.float64)
    near = [int(j) for j in np.argsort(-d) if j != 0][:9]
    edges = np.concatenate([w.edges, np.array([[0, j] for j in near], np.int32)]).astype(np.int32)
    params = dict(w.params, nedges=len(edges))
    w2 = dataclasses.replace(w, edges=edges, params=params)
    seeds = W.seeds_for(w2, NSEEDS[kind], placement="mixed")
    g, o = _run_both(w2, seeds)
    assert len(o["sline_len"]) > 0
    _assert_same(g, o)


def test_ptt_with_negative_odf_values_bit_exact(small_workloads):
    """ODF data with negative values (not prepared by PttDirectionGetter,
    which clips at 0): the library then measures every candidate's full
    support instead of stopping at the acceptance threshold, and must still
    match the oracle."""
    import dataclasses

    from gpustreamlines_amd import workloads as W

    w = small_workloads["ptt"]
    data = w.dataf.copy()
    data[data == 0] = -0.02
    w2 = dataclasses.replace(w, dataf=data)
    seeds = W.seeds_for(w2, NSEEDS["ptt"], placement="mixed")
    g, o = _run_both(w2, seeds)
    assert len(o["sline_len"]) > 0
    _assert_same(g, o)


@pytest.mark.parametrize("kind", ["csa", "opdt"])
def test_boot_odd_matrix_sizes_bit_exact(kind):
    """90 gradient directions and SH order 8 (45 coefficients): no matrix row
    length is a multiple of 4, so every matrix-vector product takes the
    row-major path instead of the lane-interleaved one."""
    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd.boot_utils import prepare_csa, prepare_opdt
    from gpustreamlines_amd.phantom import make_phantom
    from gpustreamlines_amd.sphere import single_shell_gtab, symmetric_sphere

    shape = (24, 26, 22)
    ph = make_phantom(shape, gtab=single_shell_gtab(n_b0=6, n_dirs=90), seed=3)
    verts, edges = symmetric_sphere(724)
    prep = (prepare_csa if kind == "csa" else prepare_opdt)(ph.gtab, verts, sh_order_max=8)
    boot = {k: np.ascontiguousarray(prep[k], dtype=np.float32)
            for k in ("H", "R", "delta_b", "delta_q", "sampling_matrix")}
    boot["b0s_mask"] = np.asarray(prep["b0s_mask"], dtype=np.int32)
    assert boot["H"].shape[0] % 4 != 0 and boot["delta_b"].shape[0] % 4 != 0
    p = W.base_params(shape, ph.dwi.shape[3], len(verts), len(edges), 0.25, kind.upper())
    p["delta_nr"] = boot["delta_b"].shape[0]
    w = W.Workload(kind, p, ph.dwi, ph.fa, verts, edges, boot=boot, extra=dict(phantom=ph))
    seeds = W.seeds_for(w, 64, placement="mixed")
    g, o = _run_both(w, seeds)
    assert len(o["sline_len"]) > 0
    _assert_same(g, o)
