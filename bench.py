"""Throughput of the tracking hot path on one MI355X.

A step is one pass of the hot path (phase 1: initial peak directions, the
on-device scan of direction counts, phase 2: tracking) over one batch of
synthetic seeds already resident in HBM.  Default workload: PTT on the
81x106x76 multi-tensor phantom (160 DWI + 10 b0, CSA ODF on a 724-vertex
sphere), 1M seeds, half in the FA > 0.2 mask and half in 0.05 <= FA <= 0.15
border voxels, step 0.25 voxel (0.5 mm at 2 mm), random streams numbered per
25,000-seed chunk.  Prints one JSON line.

Timing: HIP events recorded on the tracker's own stream around each phase;
steps = sum over streamlines of (points - 1).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
CHUNK = 25000


def alg_bytes_per_step(kind: str, p: dict) -> float:
    """Compulsory HBM bytes per output step with no cross-step reuse: the
    8-corner gather of the data channels the getter reads, the 8-texel stop
    map tap and the 12-byte point store.  PTT counts every probe tap of one
    rejection round (1 + 24*8 + 32*8 taps of 8 texels + LUT) per sub-step."""
    stop_tap = 8 * 4
    if kind == "ptt":
        taps = 1 + 24 * 8 + 32 * 8
        return p["step_frac"] * (taps * (8 * 4 + 4) + stop_tap) + 12
    return 8 * p["dimt"] * 4 + stop_tap + 12


MODEL_IDS = {"opdt": 0, "csa": 1, "prob": 2, "ptt": 3}
# Seeds of the CPU baseline sample: about 10-20 s of oracle work on 16 host threads.
CPU_SAMPLE = {"ptt": 2000, "prob": 250000, "csa": 8000, "opdt": 14000}


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summaries (profiles/*_traffic.json,
    copy-calibrated), or None when no summary covers this workload."""
    import glob

    for f in sorted(glob.glob(os.path.join(HERE, "profiles", "*_traffic.json"))):
        with open(f) as fh:
            t = json.load(fh)
        if t.get("workload") == workload and t.get("kernel") == kernel:
            t["source"] = os.path.relpath(f, HERE)
            return t
    return None


class HipEvents:
    """hipEvent timing on an explicit stream (the library's own)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, C.c_void_p(stream)) == 0

    def ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        t = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(t), a, b) == 0
        return float(t.value)


def cpu_baseline(w, seeds, n_sample):
    """The CPU oracle (C++/OpenMP restatement of the reference path) on the
    first n_sample seeds of the same workload, all host threads."""
    from oracle.oracle import Oracle, num_threads

    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
    sample = seeds[:n_sample]
    t0 = time.perf_counter()
    r = o.track(sample, rng_chunk=CHUNK)
    dt = time.perf_counter() - t0
    steps = int((r["sline_len"].astype(np.int64) - 1).clip(min=0).sum())
    return dict(value=steps / dt, unit="steps/s", cores=num_threads(), kind="port",
                sample=f"{n_sample} seeds of the same workload ({steps} steps, {len(r['sline_len'])} "
                       f"streamlines) in {dt:.1f} s")




def _density(dt, sld, sl, nsl, lens, shape):
    """Voxel visit counts of every point of the pass held in the padded slots
    (compacted on the device, then counted on the host)."""
    npts = int(lens.sum())
    pod = dt.alloc(8 * (nsl + 1))
    ptd = dt.alloc(12 * max(npts, 1))
    try:
        dt.compact_async(sld.ptr, nsl, sl.ptr, pod.ptr, ptd.ptr)
        dt.synchronize()
        pts = ptd.download(np.zeros((npts, 3), np.float32))
    finally:
        pod.free()
        ptd.free()
    shp = np.array(shape)
    idx = np.clip(np.rint(pts).astype(np.int64), 0, shp - 1)
    lin = (idx[:, 0] * shp[1] + idx[:, 1]) * shp[2] + idx[:, 2]
    return np.bincount(lin, minlength=int(np.prod(shp))).astype(np.float64)


def measure_learned(args, w, dt, step, plain_lens, sld, sl, nsl, L):
    """Train the BF16 proposal network on the GPU from pairs the plain getter
    records (100,000 seeds in the mask, 6% of its sub-steps: 20 M pairs,
    4 epochs of Adam),
    then time the same pass with the learned proposal (no warmup pass of its
    own).  Compares the streamline lengths (two-sample KS) and the voxel
    visit density of all points with the plain pass."""
    from scipy.stats import ks_2samp

    from gpustreamlines_amd import learned as LP
    from gpustreamlines_amd import workloads as W

    shape = tuple(w.params[k] for k in ("dimx", "dimy", "dimz"))
    d_plain = _density(dt, sld, sl, nsl, plain_lens, shape)
    t0 = time.perf_counter()
    pairs, npairs = LP.record_pairs(dt, W.seeds_for(w, 100000, placement="mask", seed=7), capacity=20_000_000,
                                    rate=0.06, rng_chunk=CHUNK)
    tr = LP.Trainer(dt, batch=4096)
    loss0 = tr.loss_grad(pairs.ptr, 4096)[0]
    hist = tr.fit(pairs, npairs, epochs=4)
    params = tr.params()
    tr.free()
    pairs.free()
    t_train = time.perf_counter() - t0
    print(f"[bench] learned proposal: {npairs} pairs, loss {loss0:.3f} -> {hist[-1]:.3f}, {t_train:.1f}s",
          file=sys.stderr, flush=True)
    dt.set_proposal(LP.to_bf16(params))
    dt.proposal_stats(reset=True)  # no warmup pass: the plain warmup already loaded the code object
    times = []
    for i in range(args.steps):
        r = step()
        times.append(r)
        print(f"[bench] learned step {i}: phase2 {r[2]:.1f} ms", file=sys.stderr, flush=True)
    stats = dt.proposal_stats(reset=True) // max(1, args.steps)
    dt.set_proposal(None)
    lens = sld.download(np.zeros(nsl, dtype=np.int32)).astype(np.int64)
    d_learned = _density(dt, sld, sl, nsl, lens, shape)
    t_step = float(np.mean([t[3] for t in times])) / 1e3
    steps = int((lens - 1).clip(min=0).sum())
    ks = ks_2samp(plain_lens, lens)
    return {
        "steps_per_s": steps / t_step,
        "ms_per_step": t_step * 1e3,
        "phase2_ms": float(np.mean([t[2] for t in times])),
        "points": int(lens.sum()),
        "first_round_accepted_substeps": int(stats[0]),
        "fallback_substeps": int(stats[1]),
        "plain_substeps_after_switch": int(stats[2]),
        "first_round_acceptance": float(stats[0] / max(1, stats[0] + stats[1])),
        "mean_length_plain": float(plain_lens.mean()),
        "mean_length_learned": float(lens.mean()),
        "ks_length_statistic": float(ks.statistic),
        "ks_length_p": float(ks.pvalue),
        "density_corr": float(np.corrcoef(d_plain, d_learned)[0, 1]),
        "train_pairs": int(npairs),
        "train_loss": [float(loss0)] + [float(h) for h in hist],
        "train_s": t_train,
    }


def measure_end_to_end(args, w, seeds, kernel_steps):
    """The user-facing path: GPUTracker.generate_streamlines over chunks of
    CHUNK seeds (seed upload, phase 1, scan, count read, phase 2, compaction,
    point download, length filter and packing), timed after one warm-up
    chunk.  Data preparation (PTT tiling and LUT) happens in the constructor
    and is reported apart; no file is written."""
    from gpustreamlines_amd import (BootDirectionGetter, GPUTracker, ProbDirectionGetter, PttDirectionGetter)

    p = w.params
    if args.getter == "prob":
        dg, data = ProbDirectionGetter(), w.dataf
    elif args.getter == "ptt":
        dg, data = PttDirectionGetter(), w.extra["odf"]
    else:
        dg, data = BootDirectionGetter(args.getter.upper(), p["min_signal"], **w.boot), w.dataf
    t0 = time.perf_counter()
    tr = GPUTracker(dg, data, w.stop_map, p["tc_threshold"], w.vertices, w.edges, max_angle=p["max_angle"],
                    step_size=p["step_size"], relative_peak_thresh=p["relative_peak_thresh"],
                    min_separation_angle=p["min_separation_angle"], chunk_size=CHUNK)
    with tr:
        prep = time.perf_counter() - t0
        tr.generate_streamlines(seeds[:CHUNK])
        t1 = time.perf_counter()
        out = tr.generate_streamlines(seeds)
        el = time.perf_counter() - t1
    lens = np.asarray(out._lengths, dtype=np.int64)
    steps = int((lens - 1).clip(min=0).sum())
    print(f"[bench] end to end: {el:.2f}s, {steps / el / 1e6:.2f} M steps/s (prep {prep:.1f}s)", file=sys.stderr,
          flush=True)
    return {"steps_per_s": steps / el, "s": el, "streamlines": int(len(lens)), "points": int(lens.sum()),
            "chunks": -(-len(seeds) // CHUNK), "prep_s": prep, "same_steps_as_kernel_pass": steps == kernel_steps}


def copy_calibration(nbytes):
    """One hipMemcpy device-to-device of nbytes: a plain copy kernel of known
    size, whose FETCH_SIZE/WRITE_SIZE scale the tracking kernels' counters."""
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--getter", default="ptt", choices=["ptt", "prob", "csa", "opdt"])
    ap.add_argument("--seeds", type=int, default=1_000_000)
    ap.add_argument("--placement", default="mixed", choices=["mixed", "mask"])
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--proposal", default="both", choices=["plain", "learned", "both"],
                    help="PTT: time the plain getter, the learned proposal, or both (value = plain)")
    ap.add_argument("--end-to-end", action="store_true",
                    help="also time GPUTracker.generate_streamlines over 25,000-seed chunks (host copies included)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: multi-GPU runs are not supported (seeds would shard "
              "contiguously per GPU with no exchange)", file=sys.stderr)
        sys.exit(2)

    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd._device import DeviceTracker

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    t_setup = time.perf_counter()
    w = W.build(args.getter)
    seeds = W.seeds_for(w, args.seeds, placement=args.placement)
    n = len(seeds)
    nv = w.params["nverts"]
    L = w.params["max_sline_len"]
    dt = DeviceTracker(w.params, device=0)
    dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
    sd = dt.alloc(seeds.nbytes)
    sd.upload(seeds)
    cd = dt.alloc(4 * (n + 1))
    dd = dt.alloc(4 * n * nv)
    od = dt.alloc(4 * (n + 1))
    stream = dt.stream()
    ev = HipEvents()
    e = [ev.create() for _ in range(4)]
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s getter={args.getter} seeds={n}", file=sys.stderr,
          flush=True)

    # Phase 1 + scan once to size the streamline buffers (same seeds every step).
    dt.phase1_async(sd.ptr, n, CHUNK, cd.ptr, dd.ptr)
    dt.scan_async(cd.ptr, n, od.ptr)
    offs = od.download(np.zeros(n + 1, dtype=np.int32))
    nsl = int(offs[n])
    ssd = dt.alloc(4 * nsl)
    sld = dt.alloc(4 * nsl)
    sl = dt.alloc(4 * 3 * 2 * L * nsl)
    print(f"[bench] {nsl} streamlines, slot buffer {4 * 3 * 2 * L * nsl / 2**30:.1f} GiB", file=sys.stderr, flush=True)

    def step():
        ev.record(e[0], stream)
        dt.phase1_async(sd.ptr, n, CHUNK, cd.ptr, dd.ptr)
        ev.record(e[1], stream)
        dt.scan_async(cd.ptr, n, od.ptr)
        ev.record(e[2], stream)
        dt.phase2_async(sd.ptr, n, CHUNK, od.ptr, dd.ptr, ssd.ptr, sld.ptr, sl.ptr)
        ev.record(e[3], stream)
        dt.synchronize()
        return ev.ms(e[0], e[1]), ev.ms(e[1], e[2]), ev.ms(e[2], e[3]), ev.ms(e[0], e[3])

    for i in range(args.warmup):
        r = step()
        print(f"[bench] warmup {i}: {r[3] / 1e3:.2f}s", file=sys.stderr, flush=True)
    dt.synchronize()
    times = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = step()
        times.append(r)
        print(f"[bench] step {i}: phase1 {r[0]:.1f} ms scan {r[1]:.2f} ms phase2 {r[2]:.1f} ms", file=sys.stderr,
              flush=True)
    wall = time.perf_counter() - t0
    lens = sld.download(np.zeros(nsl, dtype=np.int32)).astype(np.int64)
    learned = None
    if args.getter == "ptt" and args.proposal != "plain":
        try:
            learned = measure_learned(args, w, dt, step, lens, sld, sl, nsl, L)
        except Exception as exc:  # the plain line is still reported
            print(f"[bench] learned proposal failed: {exc!r}", file=sys.stderr, flush=True)
            learned = {"error": repr(exc)}
    steps_per_pass = int((lens - 1).clip(min=0).sum())
    end_to_end = None
    if args.end_to_end:
        for b in (ssd, sld, sl):
            b.free()
        ssd = sld = sl = None
        end_to_end = measure_end_to_end(args, w, seeds, steps_per_pass)
    t_step = np.array([t[3] for t in times]) / 1e3
    t_p2 = np.array([t[2] for t in times]) / 1e3
    ms_per_step = float(t_step.mean() * 1e3)
    value = steps_per_pass / float(t_step.mean())
    abytes = alg_bytes_per_step(args.getter, w.params)
    achieved = abytes * steps_per_pass / float(t_p2.mean()) / 1e9

    res = {
        "metric": "streamline steps/sec at 1M seeds, plain vs learned-proposal PTT, 1 MI355X",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.getter}_{n}_seeds_{args.placement}_81x106x76_step0.25vox",
            "getter": args.getter,
            "seeds": n,
            "streamlines": nsl,
            "points": int(lens.sum()),
            "streamlines_per_s": nsl / float(t_step.mean()),
            "phase1_ms": float(np.mean([t[0] for t in times])),
            "scan_ms": float(np.mean([t[1] for t in times])),
            "phase2_ms": float(t_p2.mean() * 1e3),
            "wall_s": wall,
            "proposal": "plain",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": f"phase2_queue_kernel<{MODEL_IDS[args.getter]}, 0>",
            "alg_bytes_per_step": abytes,
            "alg_bytes_per_launch": abytes * steps_per_pass,
        },
    }
    if end_to_end is not None:
        res["end_to_end"] = end_to_end
    if learned is not None:
        res["learned"] = learned
        if args.proposal == "learned" and "steps_per_s" in learned:
            res["value"] = learned["steps_per_s"]
            res["ms_per_step"] = learned["ms_per_step"]
            res["config"]["proposal"] = "learned"
    traffic = measured_traffic(res["config"]["workload"], res["roofline"]["kernel"])
    if traffic is not None:
        res["roofline"]["traffic"] = traffic["traffic_bytes"]
        res["roofline"]["traffic_source"] = traffic["source"]
    for b in (sd, cd, dd, od, ssd, sld, sl):
        if b is not None:
            b.free()
    dt.close()
    if not args.no_cpu:
        ns = args.cpu_sample or CPU_SAMPLE[args.getter]
        print(f"[bench] cpu baseline on {ns} seeds", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(w, seeds, ns)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
