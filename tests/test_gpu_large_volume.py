"""The large multi-shell workload (145x174x145 at 1.25 mm, 6 b0 + 3 x 90 DWI,
SH order 8 with 45 coefficients, hr = 270, step 0.4 voxel) on the GPU:
bootstrap (CSA) exact kernel bit-exact against the CPU oracle, the MFMA
kernel against the exact kernel under the tolerance bar, and PTT bit-exact
against the oracle, each on seeds of the FA > 0.2 mask.  The oracle runs a
few hundred seeds (seconds on the host)."""

import numpy as np
import pytest

from conftest import packed_points

pytestmark = pytest.mark.gpu

CHUNK = 25000


def _gpu(w, seeds, mode=None, keep_dirs=True):
    from gpustreamlines_amd._device import DeviceTracker

    dt = DeviceTracker(w.params, device=0)
    try:
        dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
        if mode is not None:
            dt.set_boot_mode(mode)
        return dt.track(seeds, rng_chunk=CHUNK, keep_dirs=keep_dirs)
    finally:
        dt.close()


def _oracle(w, seeds):
    from oracle.oracle import Oracle

    return Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot,
                  ptt_lut=w.ptt_lut).track(seeds, rng_chunk=CHUNK)


def _same(g, o):
    n = len(o["counts"])
    assert np.array_equal(g["counts"][:n], o["counts"])
    nsl = int(o["offsets"][n])
    assert np.array_equal(g["sline_len"][:nsl], o["sline_len"])
    npts = int(g["point_offsets"][nsl])
    assert np.array_equal(g["points"][:npts], packed_points(o))


def test_large_csa():
    from scipy.stats import ks_2samp

    from gpustreamlines_amd import workloads as W

    print("[large] building the 145x174x145 multi-shell phantom", flush=True)
    w = W.build_large("csa")
    assert w.params["dimt"] == 276 and w.params["delta_nr"] == 45 and w.boot["H"].shape == (270, 270)
    seeds = W.seeds_for(w, 5000, placement="mask")
    print("[large] csa exact vs oracle, 150 seeds", flush=True)
    exact = _gpu(w, seeds, mode=0)
    _same(exact, _oracle(w, seeds[:150]))
    print("[large] csa mfma vs exact, 5000 seeds", flush=True)
    m = _gpu(w, seeds, mode=1)
    assert np.array_equal(m["counts"], exact["counts"])
    la, lb = m["sline_len"].astype(np.int64), exact["sline_len"].astype(np.int64)
    assert len(la) == len(lb)
    same = (la == lb).mean()
    ks = ks_2samp(la, lb)
    print(f"[large] csa mfma: {len(la)} streamlines, equal lengths {same:.3f}, KS p {ks.pvalue:.3g}", flush=True)
    assert same > 0.7 and ks.pvalue > 0.05


def test_large_ptt():
    from gpustreamlines_amd import workloads as W

    print("[large] preparing the order-8 ODF", flush=True)
    w = W.build_large("ptt")
    seeds = W.seeds_for(w, 2000, placement="mask")
    print("[large] ptt vs oracle, 100 seeds", flush=True)
    g = _gpu(w, seeds[:100])
    _same(g, _oracle(w, seeds[:100]))
    print("[large] ptt 2000 seeds", flush=True)
    g = _gpu(w, seeds, keep_dirs=False)
    assert len(g["sline_len"]) >= 2000 and np.isfinite(g["points"]).all()
