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
    map tap and the 12-byte point store.  For PTT, the taps every sub-step
    reads whatever the data: the tangent tap and the vertex phase's 24
    curves x 8 probes (1 + 24*8 taps of one 32-byte corner block and a 4-byte
    LUT entry each), per sub-step; the rejection rounds' taps depend on the
    data and are left out (ptt_reference_bytes_per_step counts one full round)."""
    stop_tap = 8 * 4
    if kind == "ptt":
        taps = 1 + 24 * 8
        return p["step_frac"] * (taps * (8 * 4 + 4) + stop_tap) + 12
    return 8 * p["dimt"] * 4 + stop_tap + 12


def ptt_reference_bytes_per_step(p: dict) -> float:
    """The reference's PTT work per output step: the vertex phase and one
    full rejection round of 32 candidates x 8 probes, every tap read."""
    taps = 1 + 24 * 8 + 32 * 8
    return p["step_frac"] * (taps * (8 * 4 + 4) + 8 * 4) + 12


MODEL_IDS = {"opdt": 0, "csa": 1, "prob": 2, "ptt": 3}
# Seeds of the CPU baseline samples: about 10 s of oracle work each, on all
# host threads (16 on the GPU host) and on one thread.
CPU_SAMPLE = {"ptt": 1500, "prob": 200000, "csa": 6000, "opdt": 10000}
CPU_SAMPLE_1T = {"ptt": 80, "prob": 15000, "csa": 450, "opdt": 800}
# Wall-clock seconds held back for the legs after the plain passes.
CPU_RESERVE_S = 45.0
LEARNED_PASSES = 3


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summaries (profiles/*_traffic.json), or
    None when no summary covers this workload."""
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


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, seeds, n_sample, n_sample_1t):
    """The CPU oracle (C++/OpenMP restatement of the reference path) on the
    first n_sample seeds of the same workload on all host threads, and on the
    first n_sample_1t seeds on one thread."""
    from oracle.oracle import Oracle, num_threads, set_num_threads

    o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)

    def run(ns):
        t0 = time.perf_counter()
        r = o.track(seeds[:ns], rng_chunk=CHUNK)
        dt = time.perf_counter() - t0
        steps = int((r["sline_len"].astype(np.int64) - 1).clip(min=0).sum())
        return steps, len(r["sline_len"]), dt

    cores = num_threads()
    steps, nsl, dt = run(n_sample)
    set_num_threads(1)
    steps1, nsl1, dt1 = run(n_sample_1t)
    set_num_threads(cores)
    return dict(value=steps / dt, unit="steps/s", cores=cores, kind="port",
                sample=f"first {n_sample} seeds of the same workload ({steps} steps, {nsl} streamlines) in "
                       f"{dt:.1f} s on {cores} threads",
                streamlines_per_s=nsl / dt,
                one_thread={"value": steps1 / dt1, "unit": "steps/s", "cores": 1,
                            "sample": f"first {n_sample_1t} seeds ({steps1} steps, {nsl1} streamlines) in {dt1:.1f} s"},
                cpu_model=cpu_model())


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
    npass = max(1, min(args.steps, LEARNED_PASSES))
    for i in range(npass):
        r = step()
        times.append(r)
        print(f"[bench] learned step {i}: phase2 {r[2]:.1f} ms", file=sys.stderr, flush=True)
    stats = dt.proposal_stats(reset=True) // npass
    dt.set_proposal(None)
    lens = sld.download(np.zeros(nsl, dtype=np.int32)).astype(np.int64)
    d_learned = _density(dt, sld, sl, nsl, lens, shape)
    t_step = float(np.median([t[3] for t in times])) / 1e3
    steps = int((lens - 1).clip(min=0).sum())
    ks = ks_2samp(plain_lens, lens)
    return {
        "steps": npass,
        "steps_per_s": steps / t_step,
        "ms_per_step": t_step * 1e3,
        "phase2_ms": float(np.median([t[2] for t in times])),
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
    chunk; with chunks in flight the count read is gone (bounded chunks).
    Data preparation (PTT tiling and LUT) happens in the constructor
    and is reported apart; no file is written."""
    from gpustreamlines_amd import (BootDirectionGetter, GPUTracker, ProbDirectionGetter, PttDirectionGetter)

    p = w.params
    if args.getter == "prob":
        dg, data = ProbDirectionGetter(), w.dataf
    elif args.getter == "ptt":
        dg, data = PttDirectionGetter(), w.extra["odf"]
    else:
        dg, data = BootDirectionGetter(args.getter.upper(), p["min_signal"], **w.boot,
                                       mfma=args.boot_mode == "mfma"), w.dataf
    t0 = time.perf_counter()
    tr = GPUTracker(dg, data, w.stop_map, p["tc_threshold"], w.vertices, w.edges, max_angle=p["max_angle"],
                    step_size=p["step_size"], relative_peak_thresh=p["relative_peak_thresh"],
                    min_separation_angle=p["min_separation_angle"], chunk_size=CHUNK,
                    launch_seeds=args.launch_seeds, chunk_pipeline=args.chunk_pipeline,
                    chunk_graphs=bool(args.chunk_graphs))
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
            "chunks": -(-len(seeds) // CHUNK), "launch_seeds": args.launch_seeds,
            "chunk_pipeline": args.chunk_pipeline, "chunk_graphs": bool(args.chunk_graphs), "prep_s": prep,
            "same_steps_as_kernel_pass": steps == kernel_steps}


def kernel_name(args, lowrank=False):
    """The phase-2 kernel the plain pass launches (as rocprofv3 names it)."""
    if args.getter in ("csa", "opdt") and args.boot_mode == "mfma":
        lr = "true" if lowrank else "false"
        return f"phase2_boot_batch_kernel<{MODEL_IDS[args.getter]}, {args.boot_occ or 4}, {lr}>"
    if args.getter == "ptt" and args.ptt_lanes:
        kind = "phase2_ptt_flat_kernel" if args.stepping == "flat" else "phase2_ptt_queue_kernel"
        return f"{kind}<{args.ptt_lanes}>"
    v = 4 if args.stepping == "flat" else 0
    return f"phase2_queue_kernel<{MODEL_IDS[args.getter]}, {v}>"


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
    t_start = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed passes (capped by --budget, at least 3 kept)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes (capped by --budget, at least 1 kept)")
    ap.add_argument("--budget", type=float, default=360.0,
                    help="wall-clock seconds for the whole run; the pass counts are cut to fit")
    ap.add_argument("--getter", default="ptt", choices=["ptt", "prob", "csa", "opdt"])
    ap.add_argument("--seeds", type=int, default=1_000_000)
    ap.add_argument("--placement", default="mixed", choices=["mixed", "mask"])
    ap.add_argument("--cpu-sample", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--boot-mode", default="mfma", choices=["mfma", "exact"],
                    help="bootstrap getters: batched MFMA kernel or the bit-exact per-seed kernel")
    ap.add_argument("--boot-full-products", action="store_true",
                    help="MFMA bootstrap: the full [H; R] product instead of H.d = U (U^T d) and R.d from it")
    ap.add_argument("--boot-occ", type=int, default=0, choices=[0, 2, 4],
                    help="MFMA bootstrap kernel variant: waves per SIMD (0 = library default)")
    ap.add_argument("--proposal", default="both", choices=["plain", "learned", "both"],
                    help="PTT: time the plain getter, the learned proposal, or both (value = plain)")
    ap.add_argument("--end-to-end", dest="end_to_end", action="store_true", default=True,
                    help="also time GPUTracker.generate_streamlines over 25,000-seed chunks (host copies included; "
                         "the default)")
    ap.add_argument("--no-end-to-end", dest="end_to_end", action="store_false",
                    help="skip the end-to-end timing")
    ap.add_argument("--launch-seeds", type=int, default=250000,
                    help="end to end: seeds per launch (0: one launch per 25,000-seed chunk, the reference's protocol)")
    ap.add_argument("--chunk-pipeline", type=int, default=8,
                    help="end to end with one launch per chunk: chunks in flight at once (1: one after the other)")
    ap.add_argument("--ptt-lanes", type=int, default=0, choices=[0, 8, 16, 32],
                    help="PTT: lanes per seed (0: the shared 32-lane kernel)")
    ap.add_argument("--stepping", default="flat", choices=["flat", "per-seed"],
                    help="phase 2 on the queue: one loop of steps per seed slot (flat) or the reference's "
                         "per-seed nesting")
    ap.add_argument("--chunk-graphs", type=int, default=1, choices=[0, 1],
                    help="end to end with chunks in flight: each chunk as one HIP graph launch (0: eager launches)")
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

    def elapsed():
        return time.perf_counter() - t_start

    w = W.build(args.getter)
    seeds = W.seeds_for(w, args.seeds, placement=args.placement)
    n = len(seeds)
    nv = w.params["nverts"]
    L = w.params["max_sline_len"]
    dt = DeviceTracker(w.params, device=0)
    dt.upload(w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot, ptt_lut=w.ptt_lut)
    boot = args.getter in ("csa", "opdt")
    if boot:
        dt.set_boot_mode(dt.BOOT_MFMA if args.boot_mode == "mfma" else dt.BOOT_EXACT, args.boot_occ,
                         full_products=args.boot_full_products)
    dt.set_stepping(args.stepping == "flat")
    if args.getter == "ptt":
        dt.set_lanes_per_seed(args.ptt_lanes)
    sd = dt.alloc(seeds.nbytes)
    sd.upload(seeds)
    cd = dt.alloc(4 * (n + 1))
    dd = dt.alloc(4 * n * nv)
    od = dt.alloc(4 * (n + 1))
    stream = dt.stream()
    ev = HipEvents()
    e = [ev.create() for _ in range(4)]
    setup_s = elapsed()
    print(f"[bench] setup {setup_s:.1f}s getter={args.getter} seeds={n}", file=sys.stderr, flush=True)

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

    learned_leg = args.getter == "ptt" and args.proposal != "plain"
    cpu_reserve = 0.0 if args.no_cpu else CPU_RESERVE_S
    # Warm-up: the first pass loads the code objects; more only while they
    # take under a tenth of the budget.
    n_warm = 0
    t_pass = None
    while n_warm < max(1, args.warmup):
        t0 = time.perf_counter()
        r = step()
        t_pass = time.perf_counter() - t0
        n_warm += 1
        print(f"[bench] warmup {n_warm - 1}: {r[3] / 1e3:.2f}s", file=sys.stderr, flush=True)
        if n_warm >= args.warmup or elapsed() - setup_s + t_pass > 0.1 * args.budget:
            break
    dt.synchronize()
    # Passes that fit: the plain passes, then (PTT) the learned leg (training
    # ~30 s + LEARNED_PASSES passes), then the CPU legs.
    after = cpu_reserve + 10.0
    if learned_leg:
        after += 45.0 + LEARNED_PASSES * t_pass
    if args.end_to_end:  # data preparation, a warm-up chunk and the chunks at about the kernel's rate
        after += 25.0 + 1.4 * t_pass
    fit = int((args.budget - elapsed() - after) // max(t_pass, 1e-3))
    n_steps = max(min(args.steps, 3), min(args.steps, fit))
    times = []
    t0 = time.perf_counter()
    for i in range(n_steps):
        r = step()
        times.append(r)
        print(f"[bench] step {i}: phase1 {r[0]:.1f} ms scan {r[1]:.2f} ms phase2 {r[2]:.1f} ms", file=sys.stderr,
              flush=True)
    wall = time.perf_counter() - t0
    lens = sld.download(np.zeros(nsl, dtype=np.int32)).astype(np.int64)
    steps_per_pass = int((lens - 1).clip(min=0).sum())
    t_step = np.array([t[3] for t in times]) / 1e3
    t_p2 = np.array([t[2] for t in times]) / 1e3
    med_step = float(np.median(t_step))
    med_p2 = float(np.median(t_p2))
    alg_launch = alg_bytes_per_step(args.getter, w.params) * steps_per_pass
    achieved = alg_launch / med_p2 / 1e9
    print(f"[bench] plain: {steps_per_pass / med_step / 1e6:.2f} M steps/s, {achieved:.0f} GB/s algorithmic",
          file=sys.stderr, flush=True)
    learned = None
    if learned_leg:
        try:
            learned = measure_learned(args, w, dt, step, lens, sld, sl, nsl, L)
        except Exception as exc:  # the plain line is still reported
            print(f"[bench] learned proposal failed: {exc!r}", file=sys.stderr, flush=True)
            learned = {"error": repr(exc)}
    end_to_end = None
    if args.end_to_end:
        for b in (ssd, sld, sl):
            b.free()
        ssd = sld = sl = None
        end_to_end = measure_end_to_end(args, w, seeds, steps_per_pass)

    res = {
        "metric": "streamline steps/sec at 1M seeds, plain vs learned-proposal PTT, 1 MI355X",
        "value": steps_per_pass / med_step,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": n_steps,
        "warmup": n_warm,
        "ms_per_step": med_step * 1e3,
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
            "streamlines_per_s": nsl / med_step,
            "phase1_ms": float(np.median([t[0] for t in times])),
            "scan_ms": float(np.median([t[1] for t in times])),
            "phase2_ms": med_p2 * 1e3,
            "pass_ms_all": [float(x * 1e3) for x in t_step],
            "timing": "median of the timed passes (HIP events on the tracker's stream)",
            "requested_steps": args.steps,
            "requested_warmup": args.warmup,
            "budget_s": args.budget,
            "setup_s": setup_s,
            "wall_s": wall,
            "proposal": "plain",
            "boot_mode": args.boot_mode if boot else None,
            "boot_lowrank": dt.boot_lowrank() if boot else None,
            "stepping": args.stepping,
            "ptt_lanes": args.ptt_lanes if args.getter == "ptt" else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": kernel_name(args, boot and dt.boot_lowrank() > 0),
            "alg_bytes_per_step": alg_launch / max(steps_per_pass, 1),
            "alg_bytes_per_launch": alg_launch,
        },
    }
    if args.getter == "ptt":
        ref_step = ptt_reference_bytes_per_step(w.params)
        res["roofline"]["alg_bytes_source"] = ("taps every sub-step reads (tangent tap + 24 vertex curves x 8 "
                                               "probes); rejection-round taps left out")
        res["roofline"]["reference_work_bytes_per_step"] = ref_step
        res["roofline"]["reference_work_frac"] = ref_step * steps_per_pass / med_p2 / 1e9 / HBM_PEAK_GBS
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
        # the counters as read (FETCH_SIZE + WRITE_SIZE, no calibration factor)
        res["roofline"]["traffic"] = traffic.get("traffic_bytes_unscaled", traffic["traffic_bytes"])
        res["roofline"]["traffic_source"] = traffic["source"]
    for b in (sd, cd, dd, od, ssd, sld, sl):
        if b is not None:
            b.free()
    dt.close()
    print("[bench] preliminary: " + json.dumps(res), file=sys.stderr, flush=True)
    if not args.no_cpu:
        ns = args.cpu_sample or CPU_SAMPLE[args.getter]
        ns1 = max(1, ns * CPU_SAMPLE_1T[args.getter] // CPU_SAMPLE[args.getter]) if args.cpu_sample \
            else CPU_SAMPLE_1T[args.getter]
        print(f"[bench] cpu baseline on {ns} seeds (all threads) and {ns1} (one thread)", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(w, seeds, ns, ns1)
    else:
        res["cpu_baseline"] = None
    res["config"]["total_s"] = elapsed()
    # the timed passes and the whole run against the 600 s a driver allows one bench command
    timed_s = res["ms_per_step"] * res["steps"] / 1e3
    res["consistency"] = {"timed_s": timed_s, "ms_per_step_x_steps_s": timed_s, "run_s": res["config"]["total_s"],
                          "run_limit_s": 600.0, "fits_in_driver_run": res["config"]["total_s"] < 600.0}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
