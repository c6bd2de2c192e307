"""The drop-in Python API end to end on the GPU: GPUTracker with each getter,
several seed chunks (the reference's random streams restart per chunk), and
the packed result of generate_streamlines against the oracle run chunk by
chunk on the same prepared inputs."""

import numpy as np
import pytest

from conftest import packed_points

pytestmark = pytest.mark.gpu


def _oracle_by_chunks(w, seeds, chunk):
    from oracle.oracle import Oracle

    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
    pts, lens = [], []
    for k in range(0, len(seeds), chunk):
        r = o.track(seeds[k:k + chunk])
        if len(r["sline_len"]):
            pts.append(packed_points(r))
            lens.append(r["sline_len"])
    return np.concatenate(pts), np.concatenate(lens)


@pytest.mark.parametrize("launch_seeds,pipeline,graphs", [(0, 1, False), (0, 4, True), (0, 4, False),
                                                           (250000, 1, False), (64, 2, True), (64, 3, False)])
@pytest.mark.parametrize("kind", ["prob", "ptt", "csa"])
def test_gpu_tracker_generate_streamlines(small_workloads, kind, launch_seeds, pipeline, graphs):
    """launch_seeds 0: one chunk per launch (the reference's protocol), one
    chunk after the other (pipeline 1) or 4 chunks in flight on forked
    handles as bounded chunks, eager or as HIP graphs; 250000: all chunks in
    one launch with chunk-relative streams; 64 with a pipeline: launches of
    two chunks (chunk-relative streams) as bounded chunks, two in flight."""
    from gpustreamlines_amd import (BootDirectionGetter, GPUTracker, ProbDirectionGetter, PttDirectionGetter)
    from gpustreamlines_amd import workloads as W

    w = small_workloads[kind]
    seeds = W.seeds_for(w, 70, placement="mixed", seed=21)
    chunk = 32
    p = w.params
    if kind == "prob":
        dg, data = ProbDirectionGetter(), w.dataf
    elif kind == "ptt":
        dg, data = PttDirectionGetter(), small_workloads["prob"].dataf  # raw ODF; the getter prepares it
    else:
        dg, data = BootDirectionGetter("CSA", 1.0, **{k: v for k, v in w.boot.items()}, mfma=False), w.dataf
    print(f"[api] {kind} launch_seeds={launch_seeds} pipeline={pipeline} graphs={graphs}", flush=True)
    with GPUTracker(dg, data, w.stop_map, p["tc_threshold"], w.vertices, w.edges, step_size=p["step_size"],
                    chunk_size=chunk, launch_seeds=launch_seeds, chunk_pipeline=pipeline,
                    chunk_graphs=graphs) as tr:
        out = tr.generate_streamlines(seeds)
    pts, lens = _oracle_by_chunks(w, seeds, chunk)
    assert np.array_equal(out._lengths, lens)
    assert np.array_equal(out._data, pts)


def test_ptt_prepare_on_device_matches_host(small_workloads):
    """PttDirectionGetter(prepare_on_device=True): clip, stop-mask zeroing,
    per-voxel normalisation and the nearest-vertex LUT on the GPU
    (cuslines_ptt_prepare) give the host recipe's LUT and the same
    streamlines, on a raw ODF with negative values."""
    import time

    from gpustreamlines_amd import GPUTracker, PttDirectionGetter
    from gpustreamlines_amd import workloads as W

    w = small_workloads["prob"]
    raw = (w.dataf - np.float32(0.05)).astype(np.float32)  # negative lobes, for the clip
    seeds = W.seeds_for(w, 80, placement="mixed", seed=5)
    p = w.params
    out, luts, prep = {}, {}, {}
    for name, dev in (("host", False), ("device", True)):
        dg = PttDirectionGetter(prepare_on_device=dev)
        t0 = time.perf_counter()
        tr = GPUTracker(dg, raw, w.stop_map, p["tc_threshold"], w.vertices, w.edges, step_size=p["step_size"],
                        chunk_size=40)
        with tr:
            prep[name] = time.perf_counter() - t0
            out[name] = tr.generate_streamlines(seeds)
        luts[name] = dg.sphere_vertices_lut_h
    print(f"[api] PTT preparation + upload: host {prep['host']:.2f}s, device {prep['device']:.2f}s", flush=True)
    assert np.array_equal(luts["host"], luts["device"])
    assert len(out["host"]._lengths) > 0
    assert np.array_equal(out["host"]._lengths, out["device"]._lengths)
    assert np.array_equal(out["host"]._data, out["device"]._data)


def test_fork_shares_data_and_refuses_setters(small_workloads):
    """A fork tracks with its parent's device data on its own stream (same
    output); setters that replace device data are refused on the fork and on
    the parent while the fork lives, and the parent cannot be destroyed
    before it."""
    from gpustreamlines_amd import _lib
    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd._device import DeviceTracker

    w = small_workloads["prob"]
    seeds = W.seeds_for(w, 64, placement="mixed", seed=3)
    dt = DeviceTracker(w.params, device=0)
    try:
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges)
        ref = dt.track(seeds)
        f = dt.fork()
        try:
            got = f.track(seeds)
            for k in ("sline_len", "points"):
                assert np.array_equal(ref[k], got[k])
            stop = np.ascontiguousarray(w.stop_map, np.float32)
            assert f.L.cuslines_set_stop_map(f.h, stop.ctypes.data) == 3  # CUSLINES_ESTATE
            assert dt.L.cuslines_set_stop_map(dt.h, stop.ctypes.data) == 3
            assert dt.L.cuslines_tracker_destroy(dt.h) == 3
            assert "forks" in _lib.last_error()
        finally:
            f.close()
        assert np.array_equal(dt.track(seeds)["points"], ref["points"])
    finally:
        dt.close()
