"""Bounded chunks (cuslines_chunk_async: phase 1, scan, phase 2 and
compaction enqueued against a fixed slot capacity, no host read) and their
HIP graphs (cuslines_chunk_graph_create / cuslines_graph_launch), against
the step-by-step protocol of the same library (phase1, scan, phase2,
compact with the streamline count read back in between), which the other
GPU tests hold bit-exact to the oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = ["prob", "ptt", "csa", "csa-mfma"]


def _tracker(small_workloads, kind):
    from gpustreamlines_amd._device import DeviceTracker

    w = small_workloads[kind.split("-")[0]]
    dt = DeviceTracker(w.params, device=0)
    dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
    if kind.endswith("mfma"):
        dt.set_boot_mode(DeviceTracker.BOOT_MFMA)
    return w, dt


def _prefix_equal(ref, got, tracked):
    """got (a chunk's tracked prefix of `tracked` slots) equals ref's first slots."""
    npts = int(ref["point_offsets"][tracked])
    assert np.array_equal(got["sline_len"], ref["sline_len"][:tracked]), "lengths differ"
    assert np.array_equal(got["point_offsets"], ref["point_offsets"][: tracked + 1]), "point offsets differ"
    assert np.array_equal(got["points"], ref["points"][:npts]), "points differ"
    if "sline_seed" in got:
        assert np.array_equal(got["sline_seed"], ref["sline_seed"][:tracked]), "seed indices differ"


@pytest.mark.parametrize("kind", KINDS)
def test_chunk_matches_step_by_step(kind, small_workloads):
    from gpustreamlines_amd import workloads as W

    w, dt = _tracker(small_workloads, kind)
    try:
        seeds = W.seeds_for(w, 96, placement="mixed", seed=13)
        print(f"[chunk] {kind} seeds={len(seeds)}", flush=True)
        ref = dt.track(seeds, rng_chunk=40)
        nsl = len(ref["sline_len"])
        got = dt.track_chunk(seeds, rng_chunk=40)
        assert got["totals"] == (nsl, nsl, int(ref["point_offsets"][-1]))
        assert np.array_equal(got["offsets"], ref["offsets"])
        _prefix_equal(ref, got, nsl)
    finally:
        dt.close()


@pytest.mark.parametrize("kind", KINDS)
def test_chunk_slot_capacity_bounds_the_seeds_tracked(kind, small_workloads):
    """With fewer slots than streamlines the seeds whose slots all lie below
    the capacity are tracked, with their full results; the totals name the
    count needed."""
    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd._device import ChunkBuffers

    w, dt = _tracker(small_workloads, kind)
    try:
        seeds = W.seeds_for(w, 96, placement="mixed", seed=17)
        ref = dt.track(seeds, rng_chunk=0)
        offsets = ref["offsets"]
        nsl = int(offsets[-1])
        for cap in (1, nsl // 2 + 1, nsl - 1, nsl, nsl + 5):
            print(f"[chunk] {kind} bounded cap={cap} of {nsl}", flush=True)
            expect = int(offsets[offsets <= cap].max())
            cb = ChunkBuffers(dt, len(seeds), cap)
            try:
                cb.sd.upload(seeds)
                cb.enqueue(len(seeds), 0)
                got = cb.collect(len(seeds))
                assert got["totals"] == (nsl, expect, int(ref["point_offsets"][expect]))
                _prefix_equal(ref, got, expect)
                # slots past the tracked prefix read length 0
                tail = cb.sld.download(np.zeros(cap, dtype=np.int32))[expect:]
                assert not tail.any()
            finally:
                cb.close()
    finally:
        dt.close()


@pytest.mark.parametrize("kind", KINDS)
def test_chunk_graph_replays_match_eager(kind, small_workloads):
    """One captured chunk replayed several times, on alternating seed sets
    uploaded between replays, with eager chunks on the same buffers in
    between: every replay equals the eager result for its seeds."""
    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd._device import ChunkBuffers

    w, dt = _tracker(small_workloads, kind)
    try:
        a = W.seeds_for(w, 64, placement="mixed", seed=3)
        b = W.seeds_for(w, 64, placement="mixed", seed=4)
        cb = ChunkBuffers(dt, len(a), 8 * len(a))
        try:
            ref = {}
            for name, s in (("a", a), ("b", b)):
                cb.sd.upload(s)
                cb.enqueue(len(s), 32)
                ref[name] = cb.collect(len(s))
                assert ref[name]["totals"][0] == ref[name]["totals"][1] > 0
            for rep in range(3):
                for name, s, graph in (("a", a, True), ("b", b, False), ("b", b, True), ("a", a, False)):
                    print(f"[chunk] {kind} replay {rep} seeds {name} graph={graph}", flush=True)
                    cb.sd.upload(s)
                    cb.enqueue(len(s), 32, graph=graph)
                    got = cb.collect(len(s))
                    assert got["totals"] == ref[name]["totals"]
                    for k in ("offsets", "sline_seed", "sline_len", "point_offsets", "points"):
                        assert np.array_equal(got[k], ref[name][k]), f"replay {rep} {name}: {k} differs"
            assert cb.graph is not None
        finally:
            cb.close()
    finally:
        dt.close()


def test_graph_holds_the_tracker(small_workloads):
    """While a chunk graph lives, its tracker refuses new data and
    destruction; after the graph is destroyed both work again."""
    from gpustreamlines_amd import _lib
    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd._device import ChunkBuffers

    w, dt = _tracker(small_workloads, "prob")
    try:
        seeds = W.seeds_for(w, 32, placement="mask", seed=9)
        cb = ChunkBuffers(dt, len(seeds), 256)
        cb.sd.upload(seeds)
        cb.enqueue(len(seeds), 0, graph=True)
        cb.totals()
        with pytest.raises(_lib.CuslinesError):
            dt.upload(w.dataf, w.stop_map, w.vertices, w.edges)
        assert dt.L.cuslines_tracker_destroy(dt.h) == 3  # CUSLINES_ESTATE
        cb.close()
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges)
    finally:
        dt.close()


def test_pipeline_reruns_chunks_that_overflow(small_workloads):
    """The chunk pipeline with too few slots reserved: every chunk runs
    again with more, and the output equals the step-by-step protocol's."""
    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd.propagate_seeds import ChunkPipeline

    w, dt = _tracker(small_workloads, "ptt")
    try:
        seeds = W.seeds_for(w, 90, placement="mixed", seed=23)
        chunk = 30
        ref_pts, ref_lens = [], []
        for k in range(0, len(seeds), chunk):
            r = dt.track(seeds[k:k + chunk], rng_chunk=0)
            ref_pts.append(r["points"])
            ref_lens.append(r["sline_len"].astype(np.int64))
        pipe = ChunkPipeline(dt, depth=2)
        try:
            out = list(pipe.run(seeds, chunk, slots_per_seed=0.25))
            assert pipe.reruns >= 1
        finally:
            pipe.close()
        assert np.array_equal(np.concatenate([o[1] for o in out]), np.concatenate(ref_lens))
        assert np.array_equal(np.concatenate([o[0] for o in out]), np.concatenate(ref_pts))
    finally:
        dt.close()
