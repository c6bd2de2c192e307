"""The launch autotuner on the GPU: it times every candidate launch shape
on a seed sample, applies the fastest, and every shape gives the same
streamlines (for the MFMA bootstrap kernel too: each seed's products are
computed in the same order whatever the workgroup shape)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["ptt", "prob", "csa", "csa-mfma"])
def test_autotune_applies_a_shape_with_identical_output(small_workloads, kind):
    from gpustreamlines_amd import autotune
    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd._device import DeviceTracker

    w = small_workloads[kind.split("-")[0]]
    seeds = W.seeds_for(w, 96, placement="mixed")
    print(f"[autotune] {kind}", flush=True)
    dt = DeviceTracker(w.params, device=0)
    try:
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
        mode = 1 if kind.endswith("mfma") else 0
        if w.boot is not None:
            dt.set_boot_mode(mode)
        ref = dt.track(seeds)
        best, times = autotune.tune(dt, seeds)
        assert len(times) == len(list(autotune.candidates(w.params["model_type"], mode)))
        assert all(t > 0 for t in times.values())
        assert tuple(sorted(best.items())) in times
        got = dt.track(seeds)
    finally:
        dt.close()
    assert np.array_equal(ref["sline_len"], got["sline_len"])
    assert np.array_equal(ref["points"], got["points"])
