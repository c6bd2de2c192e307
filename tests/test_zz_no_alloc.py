"""Phase 2 through cuslines_phase2_noalloc_async (the entry point that
allocates nothing: its seed-order workspace is the library's static
per-device array) against cuslines_phase2_async on the same inputs, on two
calls in a row, the second with fewer seeds.  Same kernels, so bit-exact."""


import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    for k in ("counts", "offsets", "sline_seed", "sline_len", "point_offsets", "points"):
        assert np.array_equal(a[k], b[k]), f"{k} differs"


@pytest.mark.parametrize("kind", ["csa", "prob", "ptt"])
def test_noalloc_phase2_matches_phase2(kind, small_workloads):
    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd._device import DeviceTracker

    w = small_workloads[kind]
    seeds = W.seeds_for(w, 96, placement="mixed", seed=11)
    print(f"[noalloc] {kind} seeds={len(seeds)} then {len(seeds) // 3}", flush=True)
    dt = DeviceTracker(w.params, device=0)
    try:
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
        for sub in (seeds, seeds[: len(seeds) // 3]):
            ref = dt.track(sub, rng_chunk=40)
            got = dt.track(sub, rng_chunk=40, noalloc=True)
            assert len(ref["sline_len"]) > len(sub) // 2
            _same(ref, got)
    finally:
        dt.close()


def test_noalloc_rejects_too_many_seeds(small_workloads):
    """More than CUSLINES_NOALLOC_MAX_SEEDS seeds is refused before any
    launch (the buffers are never touched)."""
    from gpustreamlines_amd import _lib
    from gpustreamlines_amd._device import DeviceTracker

    w = small_workloads["csa"]
    dt = DeviceTracker(w.params, device=0)
    try:
        rc = dt.L.cuslines_phase2_noalloc_async(dt.h, None, 1_048_577, 0, None, None, None, None, None)
        assert rc == 1  # CUSLINES_EINVAL
        assert "CUSLINES_NOALLOC_MAX_SEEDS" in _lib.last_error()
    finally:
        dt.close()
