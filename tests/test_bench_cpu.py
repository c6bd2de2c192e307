"""bench.py's host-side contract, checked without a GPU: the one-GPU guard,
the algorithmic bytes per step that its roofline figure divides by time, and
the committed traffic summaries its default lines report."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert r.stdout == ""
    assert "one GPU only" in r.stderr


def test_algorithmic_bytes_per_step():
    sys.path.insert(0, ROOT)
    import bench

    # 724-channel ODF: 8 corners x 724 x 4 B, 8-texel stop tap, 12-byte point
    assert bench.alg_bytes_per_step("prob", {"dimt": 724}) == 23_212
    # 170 DWI channels (the bootstrap getters gather the signal)
    assert bench.alg_bytes_per_step("csa", {"dimt": 170}) == 5_484
    # PTT at 10 sub-steps per output point: the taps every sub-step reads,
    # one tangent tap and 24 vertices x 8 probes, each 8 texels + a LUT word
    assert bench.alg_bytes_per_step("ptt", {"step_frac": 10, "dimt": 724}) == 69_812
    per_sub = (bench.alg_bytes_per_step("ptt", {"step_frac": 10}) - 12) / 10
    assert per_sub == 193 * 36 + 32
    # the reference's work adds one full round of 32 candidates x 8 probes
    assert bench.ptt_reference_bytes_per_step({"step_frac": 10}) == 161_972


def test_default_lines_find_their_measured_traffic():
    """The committed FETCH_SIZE/WRITE_SIZE summaries are matched by workload
    and kernel name; the default PTT line and the default prob and CSA
    lines must find theirs, or `roofline.traffic` would silently be null."""
    import types

    sys.path.insert(0, ROOT)
    import bench

    for getter, lowrank in (("ptt", False), ("prob", False), ("csa", True)):
        args = types.SimpleNamespace(getter=getter, boot_mode="mfma", boot_occ=0, ptt_lanes=0, stepping="flat")
        workload = f"{getter}_1000000_seeds_mixed_81x106x76_step0.25vox"
        t = bench.measured_traffic(workload, bench.kernel_name(args, lowrank))
        assert t is not None, (getter, bench.kernel_name(args, lowrank))
        assert t.get("traffic_bytes_unscaled", t["traffic_bytes"]) > 0
