"""One tracking problem resident on one HIP device, driven through the C ABI.

``DeviceTracker`` owns a ``cuslines_tracker`` handle (its device buffers and
HIP stream) and runs the two-phase seed protocol of the reference
(cuslines/cuda_python/cu_propagate_seeds.py:206-233): phase 1 (initial peak
directions per seed), an on-device exclusive scan of the per-seed counts,
phase 2 (tracking into padded slots) and an on-device compaction of the
slots into one packed point array, so only valid points cross PCIe.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceBuffer:
    """Device allocation owned by a ``DeviceTracker``."""

    def __init__(self, dt: "DeviceTracker", nbytes: int):
        self.dt = dt
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _lib.check(dt.L.cuslines_device_alloc(dt.h, max(self.nbytes, 4), C.byref(p)), "device_alloc")
        self.ptr = p.value

    def free(self):
        if self.ptr is not None and self.dt.h:
            _lib.check(self.dt.L.cuslines_device_free(self.dt.h, C.c_void_p(self.ptr)), "device_free", hard=False)
        self.ptr = None

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("host array larger than the device buffer")
        _lib.check(self.dt.L.cuslines_memcpy_htod(self.dt.h, C.c_void_p(self.ptr), a.ctypes.data, a.nbytes), "htod")

    def download(self, out: np.ndarray):
        if not out.flags.c_contiguous or not out.flags.writeable:
            raise ValueError("download needs a writeable C-contiguous host array")
        if out.nbytes > self.nbytes:
            raise ValueError("host array larger than the device buffer")
        _lib.check(self.dt.L.cuslines_memcpy_dtoh(self.dt.h, out.ctypes.data, C.c_void_p(self.ptr), out.nbytes), "dtoh")
        return out


class DeviceTracker:
    def __init__(self, params: dict, device: int = 0):
        self.L = _lib.load()
        self.params = dict(params)
        self._ps = _lib.params_struct(params)
        h = C.c_void_p()
        _lib.check(self.L.cuslines_tracker_create(C.byref(self._ps), int(device), C.byref(h)), "tracker_create")
        self.h = h.value
        self.device = device

    def fork(self) -> "DeviceTracker":
        """A handle sharing this tracker's device data with its own stream
        (cuslines_tracker_fork): consecutive chunks on different forks run
        concurrently.  Close forks before this tracker."""
        f = DeviceTracker.__new__(DeviceTracker)
        f.L, f.params, f._ps, f.device = self.L, dict(self.params), self._ps, self.device
        h = C.c_void_p()
        _lib.check(self.L.cuslines_tracker_fork(self.h, C.byref(h)), "tracker_fork")
        f.h = h.value
        return f

    # -- setup ---------------------------------------------------------------
    def upload(self, dataf, stop_map, vertices, edges, boot=None, ptt_lut=None):
        """``dataf`` None: the data come later (``ptt_prepare``)."""
        s = np.ascontiguousarray(stop_map, dtype=np.float32)
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        e = np.ascontiguousarray(edges, dtype=np.int32)
        if dataf is not None:
            d = np.ascontiguousarray(dataf, dtype=np.float32)
            _lib.check(self.L.cuslines_set_data(self.h, d.ctypes.data), "set_data")
        _lib.check(self.L.cuslines_set_stop_map(self.h, s.ctypes.data), "set_stop_map")
        _lib.check(self.L.cuslines_set_sphere(self.h, v.ctypes.data, e.ctypes.data), "set_sphere")
        if boot is not None:
            self.set_boot(**boot)
        if ptt_lut is not None:
            self.set_lut(ptt_lut)

    def set_boot(self, H, R, delta_b, delta_q, sampling_matrix, b0s_mask, **_):
        mats = [np.ascontiguousarray(m, dtype=np.float32) for m in (H, R, delta_b, delta_q, sampling_matrix)]
        b0 = np.ascontiguousarray(b0s_mask, dtype=np.int32)
        _lib.check(self.L.cuslines_set_boot_matrices(self.h, *[m.ctypes.data for m in mats], b0.ctypes.data),
                   "set_boot_matrices")

    def set_lut(self, lut):
        lut = np.ascontiguousarray(lut, dtype=np.int32)
        _lib.check(self.L.cuslines_set_ptt_lut(self.h, lut.ctypes.data), "set_ptt_lut")

    def ptt_prepare(self, odf, stop_map, stop_threshold, sphere_symm=False):
        """PTT data preparation on the device (cuslines_ptt_prepare): raw ODF
        in, prepared volume, LUT and corner blocks on the device; returns the
        LUT (int32 [lut_res]^3)."""
        o = np.ascontiguousarray(odf, dtype=np.float32)
        s = np.ascontiguousarray(stop_map, dtype=np.float32)
        r = int(self.params["lut_res"])
        lut = np.zeros((r, r, r), dtype=np.int32)
        _lib.check(self.L.cuslines_ptt_prepare(self.h, o.ctypes.data, s.ctypes.data, float(stop_threshold),
                                               int(bool(sphere_symm)), lut.ctypes.data), "ptt_prepare")
        return lut

    def close(self):
        if getattr(self, "h", None):
            _lib.check(self.L.cuslines_tracker_destroy(self.h), "tracker_destroy", hard=False)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers and stream -------------------------------------------------
    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def synchronize(self):
        _lib.check(self.L.cuslines_synchronize(self.h), "synchronize")

    def set_launch(self, schedule: int = 0, threads: int = 0, blocks_per_cu: int = 0):
        """Phase-2 launch: schedule 0 = persistent work queue, 1 = static."""
        _lib.check(self.L.cuslines_set_launch(self.h, int(schedule), int(threads), int(blocks_per_cu)), "set_launch")

    def set_lanes_per_seed(self, lanes: int = 0):
        """PTT phase 2: lanes per seed (8, 16, 32; 0 = shared 32-lane kernel)."""
        _lib.check(self.L.cuslines_set_lanes_per_seed(self.h, int(lanes)), "set_lanes_per_seed")

    def set_stepping(self, flat: bool = True):
        """Phase 2 on the queue: one loop of steps per seed slot (True,
        default) or the per-seed nesting of the reference."""
        _lib.check(self.L.cuslines_set_stepping(self.h, int(bool(flat))), "set_stepping")

    def set_ptt_layout(self, corner_blocks: bool = True):
        """PTT taps from 32-byte corner blocks (True) or 8 texel gathers."""
        _lib.check(self.L.cuslines_set_ptt_layout(self.h, int(bool(corner_blocks))), "set_ptt_layout")

    # -- learned PTT proposal -------------------------------------------------
    def set_proposal(self, mlp_bf16):
        """bf16 parameters (uint16 array) of the proposal network, or None."""
        if mlp_bf16 is None:
            _lib.check(self.L.cuslines_set_proposal(self.h, None), "set_proposal")
            return
        w = np.ascontiguousarray(mlp_bf16, dtype=np.uint16)
        if w.size != _lib.MLP_NPARAM:
            raise ValueError(f"expected {_lib.MLP_NPARAM} parameters, got {w.size}")
        _lib.check(self.L.cuslines_set_proposal(self.h, w.ctypes.data), "set_proposal")

    def set_recording(self, pairs_d=None, capacity=0, count_d=None, rate=1.0):
        _lib.check(self.L.cuslines_set_recording(self.h, C.c_void_p(pairs_d), int(capacity), C.c_void_p(count_d),
                                                 float(rate)), "set_recording")

    def proposal_stats(self, reset: bool = True) -> np.ndarray:
        out = (C.c_int64 * 3)()
        _lib.check(self.L.cuslines_proposal_stats(self.h, out, int(reset)), "proposal_stats")
        return np.array(list(out), dtype=np.int64)

    def proposal_counters(self, reset: bool = True) -> np.ndarray:
        """proposal_stats plus [3]: learned face draws that broke the
        envelope of the exact correction."""
        out = (C.c_int64 * 4)()
        _lib.check(self.L.cuslines_proposal_counters(self.h, out, int(reset)), "proposal_counters")
        return np.array(list(out), dtype=np.int64)

    BOOT_EXACT, BOOT_MFMA = 0, 1

    BOOT_FULL_PRODUCTS = 0x10000

    def set_boot_mode(self, mode: int, waves_per_simd: int = 0, full_products: bool = False):
        """Bootstrap phase-2 kernel: BOOT_EXACT (bit-identical to the CPU
        oracle) or BOOT_MFMA (16 seeds in lockstep, matrix products on the
        matrix cores; equal within float rounding).  waves_per_simd (2 or 4,
        0 = the default 4) picks the MFMA kernel's occupancy variant;
        full_products keeps the full [H; R] product where the low-rank form
        (H = U U^T, R its leveraged centred residual matrix) would apply."""
        m = int(mode) | (int(waves_per_simd) << 8) | (self.BOOT_FULL_PRODUCTS if full_products else 0)
        _lib.check(self.L.cuslines_set_boot_mode(self.h, m), "set_boot_mode")
        self.boot_mode = int(mode)

    def boot_lowrank(self) -> int:
        """Rank of H used by the MFMA kernel's low-rank form (0: full products)."""
        r = C.c_int32()
        _lib.check(self.L.cuslines_boot_lowrank(self.h, C.byref(r)), "boot_lowrank")
        return int(r.value)

    def boot_counters(self, reset: bool = True) -> np.ndarray:
        """MFMA bootstrap kernel: [slot attempts, workgroup rounds] since the
        last reset."""
        out = (C.c_int64 * 2)()
        _lib.check(self.L.cuslines_boot_counters(self.h, out, int(reset)), "boot_counters")
        return np.array(list(out), dtype=np.int64)

    def set_tap_count(self, enable: bool = True):
        """PTT phase 2 counts its ODF taps and sub-steps (same results)."""
        _lib.check(self.L.cuslines_set_tap_count(self.h, int(bool(enable))), "set_tap_count")

    def tap_count(self, reset: bool = True) -> np.ndarray:
        """[taps, sub-steps] counted since the last reset."""
        out = (C.c_int64 * 2)()
        _lib.check(self.L.cuslines_tap_count(self.h, out, int(reset)), "tap_count")
        return np.array(list(out), dtype=np.int64)

    def mlp_workspace_bytes(self, n: int) -> int:
        b = C.c_size_t(0)
        _lib.check(self.L.cuslines_mlp_workspace_bytes(int(n), C.byref(b)), "mlp_workspace_bytes")
        return b.value

    def mlp_loss_grad_async(self, params_d, pairs_d, n, loss_d, grad_d, work_d):
        _lib.check(self.L.cuslines_mlp_loss_grad_async(self.h, C.c_void_p(params_d), C.c_void_p(pairs_d), int(n),
                                                       C.c_void_p(loss_d), C.c_void_p(grad_d), C.c_void_p(work_d)),
                   "mlp_loss_grad")

    def mlp_adam_async(self, params_d, grad_d, m_d, v_d, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _lib.check(self.L.cuslines_mlp_adam_async(self.h, C.c_void_p(params_d), C.c_void_p(grad_d), C.c_void_p(m_d),
                                                  C.c_void_p(v_d), int(step), float(lr), float(beta1), float(beta2),
                                                  float(eps)), "mlp_adam")

    def stream(self) -> int:
        s = C.c_void_p()
        _lib.check(self.L.cuslines_get_stream(self.h, C.byref(s)), "get_stream")
        return s.value or 0

    # -- phases on device buffers -----------------------------------------
    def phase1_async(self, seeds_d, n, rng_chunk, counts_d, dirs_d):
        _lib.check(self.L.cuslines_phase1_async(self.h, C.c_void_p(seeds_d), int(n), int(rng_chunk),
                                                C.c_void_p(counts_d), C.c_void_p(dirs_d)), "phase1")

    def scan_async(self, counts_d, n, offsets_d):
        _lib.check(self.L.cuslines_scan_async(self.h, C.c_void_p(counts_d), int(n), C.c_void_p(offsets_d)), "scan")

    def phase2_async(self, seeds_d, n, rng_chunk, offsets_d, dirs_d, sseed_d, slen_d, sline_d):
        _lib.check(self.L.cuslines_phase2_async(self.h, C.c_void_p(seeds_d), int(n), int(rng_chunk),
                                                C.c_void_p(offsets_d), C.c_void_p(dirs_d), C.c_void_p(sseed_d),
                                                C.c_void_p(slen_d), C.c_void_p(sline_d)), "phase2")

    def phase2_noalloc_async(self, seeds_d, n, rng_chunk, offsets_d, dirs_d, sseed_d, slen_d, sline_d):
        """phase2_async without any allocation by the library: the seed-order
        workspace is the library's static per-device array (n <= 1,048,576)."""
        _lib.check(self.L.cuslines_phase2_noalloc_async(self.h, C.c_void_p(seeds_d), int(n), int(rng_chunk),
                                                        C.c_void_p(offsets_d), C.c_void_p(dirs_d),
                                                        C.c_void_p(sseed_d), C.c_void_p(slen_d),
                                                        C.c_void_p(sline_d)), "phase2_noalloc")

    def compact_async(self, slen_d, nsl, sline_d, poff_d, pts_d):
        _lib.check(self.L.cuslines_compact_async(self.h, C.c_void_p(slen_d), int(nsl), C.c_void_p(sline_d),
                                                 C.c_void_p(poff_d), C.c_void_p(pts_d)), "compact")

    # -- the whole protocol on host arrays ----------------------------------
    def track(self, seeds, rng_chunk: int = 0, keep_dirs: bool = False, keep_slots: bool = False,
              noalloc: bool = False) -> dict:
        """Phase 1, scan, phase 2 and compaction for ``seeds`` (float32 [n,3]).
        ``noalloc`` runs phase 2 through the entry point that allocates nothing.

        Returns host arrays: ``counts`` [n], ``offsets`` [n+1],
        ``sline_seed``/``sline_len`` [nsl], ``points`` [sum(len),3] and
        ``point_offsets`` int64 [nsl+1]; ``dirs`` [n, nverts] with
        ``keep_dirs``; padded ``slots`` with ``keep_slots``.
        """
        seeds = np.ascontiguousarray(seeds, dtype=np.float32).reshape(-1, 3)
        n = len(seeds)
        nv = self.params["nverts"]
        L = self.params.get("max_sline_len", 501)
        out = {}
        bufs = []
        try:
            sd = self.alloc(seeds.nbytes); bufs.append(sd)
            cd = self.alloc(4 * (n + 1)); bufs.append(cd)
            dd = self.alloc(4 * n * nv); bufs.append(dd)
            od = self.alloc(4 * (n + 1)); bufs.append(od)
            sd.upload(seeds)
            self.phase1_async(sd.ptr, n, rng_chunk, cd.ptr, dd.ptr)
            self.scan_async(cd.ptr, n, od.ptr)
            offsets = od.download(np.zeros(n + 1, dtype=np.int32))
            counts = cd.download(np.zeros(n, dtype=np.int32))
            nsl = int(offsets[n])
            out["counts"], out["offsets"] = counts, offsets
            if keep_dirs:
                out["dirs"] = dd.download(np.zeros((n, nv), dtype=np.int32))
            ssd = self.alloc(4 * nsl); bufs.append(ssd)
            sld = self.alloc(4 * nsl); bufs.append(sld)
            sl = self.alloc(4 * 3 * 2 * L * nsl); bufs.append(sl)
            p2 = self.phase2_noalloc_async if noalloc else self.phase2_async
            p2(sd.ptr, n, rng_chunk, od.ptr, dd.ptr, ssd.ptr, sld.ptr, sl.ptr)
            slen = sld.download(np.zeros(nsl, dtype=np.int32))
            out["sline_seed"] = ssd.download(np.zeros(nsl, dtype=np.int32))
            out["sline_len"] = slen
            if keep_slots:
                out["slots"] = sl.download(np.zeros((nsl, 2 * L, 3), dtype=np.float32))
            npts = int(slen.astype(np.int64).sum())
            pod = self.alloc(8 * (nsl + 1)); bufs.append(pod)
            ptd = self.alloc(4 * 3 * npts); bufs.append(ptd)
            self.compact_async(sld.ptr, nsl, sl.ptr, pod.ptr, ptd.ptr)
            out["point_offsets"] = pod.download(np.zeros(nsl + 1, dtype=np.int64))
            out["points"] = ptd.download(np.zeros((npts, 3), dtype=np.float32))
        finally:
            for b in bufs:
                b.free()
        return out

    def track_chunk(self, seeds, rng_chunk: int = 0, slot_cap: int | None = None, graph: bool = False,
                    launches: int = 1) -> dict:
        """``track`` as one bounded chunk (cuslines_chunk_async, or its HIP
        graph with ``graph``; ``launches`` > 1 replays the graph that many
        times, re-uploading the seeds before each).  ``slot_cap`` defaults to
        8 slots per seed.  Returns ``totals`` (streamlines of the chunk,
        slots tracked, points) and the tracked prefix: ``offsets`` [n+1],
        ``sline_seed``/``sline_len`` [tracked], ``points``, ``point_offsets``.
        """
        seeds = np.ascontiguousarray(seeds, dtype=np.float32).reshape(-1, 3)
        n = len(seeds)
        cb = ChunkBuffers(self, n, slot_cap if slot_cap is not None else max(8 * n, 1))
        try:
            for _ in range(max(1, int(launches))):
                cb.sd.upload(seeds)
                cb.enqueue(n, rng_chunk, graph=graph)
                res = cb.collect(n)
        finally:
            cb.close()
        return res


class ChunkBuffers:
    """Device buffers of bounded chunks (cuslines_chunk_async) of at most
    ``nmax`` seeds into ``cap`` slots on one handle, and the HIP graph of a
    chunk of exactly ``nmax`` seeds (captured at its first use)."""

    def __init__(self, dt: DeviceTracker, nmax: int, cap: int):
        self.dt, self.nmax, self.cap = dt, int(nmax), int(cap)
        nv = dt.params["nverts"]
        L = int(dt.params.get("max_sline_len", 501))
        ws = C.c_size_t()
        _lib.check(dt.L.cuslines_chunk_workspace_bytes(dt.h, self.nmax, self.cap, C.byref(ws)), "chunk_workspace")
        self.bufs = []
        try:
            self.sd = self._alloc(12 * self.nmax)
            self.cd = self._alloc(4 * (self.nmax + 1))
            self.dd = self._alloc(4 * self.nmax * nv)
            self.od = self._alloc(4 * (self.nmax + 1))
            self.ssd = self._alloc(4 * self.cap)
            self.sld = self._alloc(4 * self.cap)
            self.sl = self._alloc(4 * 3 * 2 * L * self.cap)
            self.pod = self._alloc(8 * (self.cap + 1))
            self.ptd = self._alloc(4 * 3 * (2 * L - 1) * self.cap)
            self.tot = self._alloc(8 * 3)
            self.ws = self._alloc(ws.value)
        except Exception:
            self.close()
            raise
        self.graph = None
        self.graph_chunk = None

    def _alloc(self, nbytes):
        b = self.dt.alloc(nbytes)
        self.bufs.append(b)
        return b

    def _args(self, n, rng_chunk):
        P = C.c_void_p
        return (self.dt.h, P(self.sd.ptr), int(n), int(rng_chunk), P(self.cd.ptr), P(self.dd.ptr), P(self.od.ptr),
                self.cap, P(self.ssd.ptr), P(self.sld.ptr), P(self.sl.ptr), P(self.pod.ptr), P(self.ptd.ptr),
                P(self.tot.ptr), P(self.ws.ptr), self.ws.nbytes)

    def enqueue(self, n: int, rng_chunk: int, graph: bool = False):
        """Enqueue the chunk of the first ``n`` seeds in ``sd`` (no host
        read); with ``graph`` and ``n == nmax``, as a graph launch."""
        if n > self.nmax:
            raise ValueError("chunk larger than its buffers")
        if graph and n == self.nmax:
            if self.graph is None or self.graph_chunk != rng_chunk:
                self.drop_graph()
                g = C.c_void_p()
                _lib.check(self.dt.L.cuslines_chunk_graph_create(*self._args(n, rng_chunk), C.byref(g)),
                           "chunk_graph_create")
                self.graph, self.graph_chunk = g.value, rng_chunk
            _lib.check(self.dt.L.cuslines_graph_launch(C.c_void_p(self.graph)), "graph_launch")
        else:
            _lib.check(self.dt.L.cuslines_chunk_async(*self._args(n, rng_chunk)), "chunk")

    def totals(self):
        """(streamlines of the chunk, slots tracked, points), after the chunk."""
        t = self.tot.download(np.zeros(3, dtype=np.int64))
        return int(t[0]), int(t[1]), int(t[2])

    def collect(self, n: int, with_seeds: bool = True, points_out=None) -> dict:
        """The chunk's results on the host; the points go into
        ``points_out`` (float32 [npts, 3], C-contiguous) when given."""
        nsl, tracked, npts = self.totals()
        out = {"totals": (nsl, tracked, npts)}
        out["offsets"] = self.od.download(np.zeros(n + 1, dtype=np.int32))
        out["sline_len"] = self.sld.download(np.zeros(tracked, dtype=np.int32))
        if with_seeds:
            out["sline_seed"] = self.ssd.download(np.zeros(tracked, dtype=np.int32))
        out["point_offsets"] = self.pod.download(np.zeros(tracked + 1, dtype=np.int64))
        if points_out is None:
            points_out = np.zeros((npts, 3), dtype=np.float32)
        elif points_out.shape != (npts, 3) or points_out.dtype != np.float32:
            raise ValueError(f"points_out must be float32 [{npts}, 3]")
        out["points"] = self.ptd.download(points_out)
        return out

    def drop_graph(self):
        if self.graph is not None:
            _lib.check(self.dt.L.cuslines_graph_destroy(C.c_void_p(self.graph)), "graph_destroy", hard=False)
            self.graph = None

    def close(self):
        self.drop_graph()
        for b in self.bufs:
            b.free()
        self.bufs = []
