"""Launch-shape autotuner for phase 2.

Every launch shape gives bit-identical streamlines (the random streams are
keyed by seed and virtual lane, never by hardware slot), so the tuner only
measures: it runs phase 2 of a seed sample under each candidate shape on the
tracker's stream, times it with HIP events, and applies the fastest.

Tuned knobs: schedule (persistent queue or static), threads per block (seed
slots per workgroup), resident blocks per CU (the occupancy maximum or
fewer workgroups, each keeping its LDS slots), the stepping of the queue
kernels (one flat loop of steps per seed slot, or the reference's per-seed
nesting) and, for PTT, lanes per seed (seeds per wavefront).  For the
batched MFMA bootstrap kernel, whose
workgroup is 16 seed slots of LDS and one product tile column each, the
knobs are its register/LDS tiling variant (2 waves per SIMD with 256
registers, or 4 with 128 and two workgroups per CU) and the workgroups per
CU; the bootstrap mode itself (exact or MFMA) is the caller's choice and is
kept.
"""

from __future__ import annotations

import ctypes as C
import itertools

import numpy as np

from ._globals import MAX_SLINE_LEN


class _Events:
    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        for f, a in (("hipEventCreate", [C.POINTER(C.c_void_p)]), ("hipEventRecord", [C.c_void_p, C.c_void_p]),
                     ("hipEventSynchronize", [C.c_void_p]),
                     ("hipEventElapsedTime", [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]),
                     ("hipEventDestroy", [C.c_void_p])):
            getattr(self.hip, f).argtypes = a

    def pair(self):
        a, b = C.c_void_p(), C.c_void_p()
        self.hip.hipEventCreate(C.byref(a))
        self.hip.hipEventCreate(C.byref(b))
        return a, b

    def destroy(self, *events):
        for e in events:
            self.hip.hipEventDestroy(e)

    def ms(self, a, b):
        self.hip.hipEventSynchronize(b)
        t = C.c_float()
        self.hip.hipEventElapsedTime(C.byref(t), a, b)
        return float(t.value)


def candidates(model_type: int, boot_mode: int = 0):
    if model_type == 3:  # PTT
        for lanes, threads in itertools.product((0, 16, 8), (256, 128)):
            yield dict(schedule=0, threads=threads, blocks_per_cu=0, lanes=lanes, flat=True)
        yield dict(schedule=0, threads=256, blocks_per_cu=2, lanes=0, flat=True)
        yield dict(schedule=0, threads=256, blocks_per_cu=0, lanes=0, flat=False)
        yield dict(schedule=1, threads=256, blocks_per_cu=0, lanes=0, flat=False)
    elif model_type in (0, 1) and boot_mode == 1:  # batched MFMA bootstrap
        for occ, bpc in ((4, 0), (4, 1), (2, 0)):
            yield dict(schedule=0, threads=0, blocks_per_cu=bpc, lanes=0, waves_per_simd=occ, flat=True)
    else:
        threads = (128, 64) if model_type in (0, 1) else (256, 128, 64)
        for t in threads:
            yield dict(schedule=0, threads=t, blocks_per_cu=0, lanes=0, flat=True)
        yield dict(schedule=0, threads=threads[0], blocks_per_cu=2, lanes=0, flat=True)
        yield dict(schedule=0, threads=threads[0], blocks_per_cu=0, lanes=0, flat=False)
        yield dict(schedule=1, threads=threads[0], blocks_per_cu=0, lanes=0, flat=False)


def apply(dt, cfg):
    dt.set_launch(cfg["schedule"], cfg["threads"], cfg["blocks_per_cu"])
    dt.set_stepping(cfg.get("flat", True))
    if dt.params["model_type"] == 3:
        dt.set_lanes_per_seed(cfg["lanes"])
    if "waves_per_simd" in cfg:
        dt.set_boot_mode(dt.BOOT_MFMA, cfg["waves_per_simd"])


def tune(dt, seeds: np.ndarray, rng_chunk: int = 0, verbose: bool = False):
    """Time phase 2 of ``seeds`` under each candidate shape on device tracker
    ``dt``; apply and return the fastest ``(config, {config: ms})``."""
    seeds = np.ascontiguousarray(seeds, dtype=np.float32)
    n = len(seeds)
    nv = dt.params["nverts"]
    bufs = []
    try:
        sd = dt.alloc(seeds.nbytes); bufs.append(sd)
        cd = dt.alloc(4 * (n + 1)); bufs.append(cd)
        dd = dt.alloc(4 * n * nv); bufs.append(dd)
        od = dt.alloc(4 * (n + 1)); bufs.append(od)
        sd.upload(seeds)
        dt.phase1_async(sd.ptr, n, rng_chunk, cd.ptr, dd.ptr)
        dt.scan_async(cd.ptr, n, od.ptr)
        nsl = int(od.download(np.zeros(n + 1, dtype=np.int32))[n])
        ssd = dt.alloc(4 * max(nsl, 1)); bufs.append(ssd)
        sld = dt.alloc(4 * max(nsl, 1)); bufs.append(sld)
        L = int(dt.params.get("max_sline_len", MAX_SLINE_LEN))  # slot = 2 L points
        sl = dt.alloc(4 * 3 * 2 * L * max(nsl, 1)); bufs.append(sl)
        ev = _Events()
        a, b = ev.pair()
        stream = C.c_void_p(dt.stream())
        times = {}
        try:
            for cfg in candidates(dt.params["model_type"], getattr(dt, "boot_mode", 0)):
                apply(dt, cfg)
                ev.hip.hipEventRecord(a, stream)
                dt.phase2_async(sd.ptr, n, rng_chunk, od.ptr, dd.ptr, ssd.ptr, sld.ptr, sl.ptr)
                ev.hip.hipEventRecord(b, stream)
                dt.synchronize()
                key = tuple(sorted(cfg.items()))
                times[key] = ev.ms(a, b)
                if verbose:
                    print(f"[autotune] {cfg} {times[key]:.1f} ms", flush=True)
        finally:
            ev.destroy(a, b)
        best = dict(min(times, key=times.get))
        apply(dt, best)
        return best, times
    finally:
        for buf in bufs:
            buf.free()
