"""Seed-chunk propagation (cuslines/cuda_python/cu_propagate_seeds.py).

Per chunk and per GPU share: upload seeds, phase 1, exclusive scan of the
per-seed direction counts, phase 2, then compaction of the padded point
slots.  Differences from the reference, both on the device side of the
boundary: the scan runs on the GPU (the reference copies counts to the host
and scans in a Python loop, :95-129) and only valid points are copied back
(the reference copies every 1,002-point slot, :171-191).  ``slines[n]`` is
therefore the packed float32 [n_points, 3] array of GPU ``n`` and
``sline_offs[n]`` its int64 [nSlines+1] point offsets.
"""

from __future__ import annotations

import logging
import math

import numpy as np

from ._globals import MAX_SLINE_LEN, REAL_SIZE, THR_X_BL, THR_X_SL
from .array_sequence import MEGABYTE, make_array_sequence

logger = logging.getLogger("GPUStreamlines")


def div_up(a, b):
    return (a + b - 1) // b


def seed_shares(nseeds: int, ngpus: int):
    """(first seed, seed count) of each GPU's contiguous share of a chunk:
    ceil(nseeds / ngpus) seeds per GPU, the last GPUs taking what is left
    (cu_propagate_seeds.py:45-54)."""
    per = div_up(nseeds, ngpus) if ngpus > 0 else 0
    return [(min(ii * per, nseeds), min(per, max(0, nseeds - ii * per))) for ii in range(ngpus)]


def streamline_offsets(counts):
    """Exclusive scan of the per-GPU streamline counts: where each GPU's
    streamlines start in the chunk's merged output (the one exchange the
    seed split needs; N integers)."""
    counts = np.asarray(counts, dtype=np.int64)
    return np.concatenate([[0], np.cumsum(counts)])


class SeedBatchPropagator:
    def __init__(self, gpu_tracker, minlen=0, maxlen=np.inf):
        self.gpu_tracker = gpu_tracker
        self.ngpus = gpu_tracker.ngpus
        self.minlen = minlen
        self.maxlen = maxlen
        self.nSlines = np.zeros(self.ngpus, dtype=np.int32)
        self.slines = [None] * self.ngpus
        self.sline_lens = [None] * self.ngpus
        self.sline_offs = [None] * self.ngpus
        self.sline_seeds = [None] * self.ngpus
        # Random streams restart every rng_chunk seeds of a launch (None: one
        # chunk per launch, as the reference); set when one launch carries
        # several chunks (GenericTracker, one GPU).
        self.rng_chunk = None
        # When set (an array_sequence.PackedSink), each GPU's points are
        # downloaded straight into it instead of into arrays of their own.
        self.sink = None
        self._reset_buffers()

    def _reset_buffers(self):
        n = self.ngpus
        self.seeds_d = [None] * n
        self.slinesOffs_d = [None] * n   # phase-1 counts
        self.slinesScan_d = [None] * n   # exclusive offsets
        self.shDirTemp0_d = [None] * n
        self.slineSeed_d = [None] * n
        self.slineLen_d = [None] * n
        self.sline_d = [None] * n

    def _switch_device(self, n):
        nseeds_gpu = seed_shares(self.nseeds, self.ngpus)[n][1]
        block = (THR_X_SL, THR_X_BL // THR_X_SL, 1)
        grid = (div_up(nseeds_gpu, THR_X_BL // THR_X_SL), 1, 1)
        return nseeds_gpu, block, grid

    def _free(self):
        for lst in (self.seeds_d, self.slinesOffs_d, self.slinesScan_d, self.shDirTemp0_d,
                    self.slineSeed_d, self.slineLen_d, self.sline_d):
            for b in lst:
                if b is not None:
                    b.free()
        self._reset_buffers()

    def propagate(self, seeds):
        seeds = np.ascontiguousarray(seeds, dtype=np.float32).reshape(-1, 3)
        self.nseeds = len(seeds)
        self.nseeds_per_gpu = (self.nseeds + self.gpu_tracker.ngpus - 1) // self.gpu_tracker.ngpus
        tr = self.gpu_tracker
        dg = tr.dg
        try:
            for ii in range(self.ngpus):
                nseeds_gpu, block, grid = self._switch_device(ii)
                dt = tr.device_trackers[ii]
                lo = seed_shares(self.nseeds, self.ngpus)[ii][0]
                part = seeds[lo: lo + nseeds_gpu]
                self.seeds_d[ii] = dt.alloc(part.nbytes)
                self.slinesOffs_d[ii] = dt.alloc(4 * (nseeds_gpu + 1))
                self.slinesScan_d[ii] = dt.alloc(4 * (nseeds_gpu + 1))
                self.shDirTemp0_d[ii] = dt.alloc(4 * max(1, nseeds_gpu) * tr.samplm_nr)
                if nseeds_gpu:
                    self.seeds_d[ii].upload(part)
            for ii in range(self.ngpus):
                nseeds_gpu, block, grid = self._switch_device(ii)
                if nseeds_gpu == 0:
                    continue
                dg.getNumStreamlines(ii, nseeds_gpu, block, grid, self)
            for ii in range(self.ngpus):
                nseeds_gpu, _, _ = self._switch_device(ii)
                dt = tr.device_trackers[ii]
                if nseeds_gpu == 0:
                    self.nSlines[ii] = 0
                    continue
                dt.scan_async(self.slinesOffs_d[ii].ptr, nseeds_gpu, self.slinesScan_d[ii].ptr)
                tot = np.zeros(nseeds_gpu + 1, dtype=np.int32)
                self.slinesScan_d[ii].download(tot)
                self.nSlines[ii] = int(tot[nseeds_gpu])
            for ii in range(self.ngpus):
                nseeds_gpu, block, grid = self._switch_device(ii)
                dt = tr.device_trackers[ii]
                nsl = int(self.nSlines[ii])
                self.slineSeed_d[ii] = dt.alloc(4 * nsl)
                self.slineLen_d[ii] = dt.alloc(4 * nsl)
                slot_len = int(dt.params.get("max_sline_len", MAX_SLINE_LEN))  # slot = 2 * slot_len points
                self.sline_d[ii] = dt.alloc(REAL_SIZE * 2 * 3 * slot_len * nsl)
                if nseeds_gpu and nsl:
                    dg.generateStreamlines(ii, nseeds_gpu, block, grid, self)
            self._cleanup()
        finally:
            self._free()

    def _cleanup(self):
        tr = self.gpu_tracker
        for ii in range(self.ngpus):
            dt = tr.device_trackers[ii]
            nsl = int(self.nSlines[ii])
            lens = np.zeros(nsl, dtype=np.int32)
            seeds_of = np.zeros(nsl, dtype=np.int32)
            if nsl:
                self.slineLen_d[ii].download(lens)
                self.slineSeed_d[ii].download(seeds_of)
            npts = int(lens.astype(np.int64).sum())
            offs = np.zeros(nsl + 1, dtype=np.int64)
            if self.sink is not None:
                pts = self.sink.take(npts, seed_shares(self.nseeds, self.ngpus)[ii][1])
            else:
                pts = np.zeros((npts, 3), dtype=np.float32)
            if nsl:
                pod = dt.alloc(8 * (nsl + 1))
                ptd = dt.alloc(12 * max(npts, 1))
                try:
                    dt.compact_async(self.slineLen_d[ii].ptr, nsl, self.sline_d[ii].ptr, pod.ptr, ptd.ptr)
                    pod.download(offs)
                    if npts:
                        ptd.download(pts)
                finally:
                    pod.free()
                    ptd.free()
            self.slines[ii] = pts
            self.sline_lens[ii] = lens
            self.sline_offs[ii] = offs
            self.sline_seeds[ii] = seeds_of
        self.gpu_tracker.rng_offset += self.nseeds

    def _keep(self, npts):
        return not (npts < self.minlen or npts > self.maxlen)

    def get_buffer_size(self):
        """Megabytes needed for the kept streamlines (:235-243)."""
        size = 0
        for ii in range(self.ngpus):
            lens = self.sline_lens[ii]
            if lens is None:
                continue
            keep = (lens >= self.minlen) & (lens <= self.maxlen)
            size += int(lens[keep].astype(np.int64).sum()) * 3 * REAL_SIZE
        return math.ceil(size / MEGABYTE)

    def as_generator(self):
        def _yield_slines():
            for ii in range(self.ngpus):
                pts, lens, offs = self.slines[ii], self.sline_lens[ii], self.sline_offs[ii]
                if lens is None:
                    continue
                for jj in range(int(self.nSlines[ii])):
                    npts = int(lens[jj])
                    if not self._keep(npts):
                        continue
                    yield pts[offs[jj]: offs[jj] + npts]

        return _yield_slines()

    def as_array_sequence(self):
        return make_array_sequence(self.as_generator(), self.get_buffer_size())


class ChunkPipeline:
    """The reference's chunk protocol (one phase-1/scan/phase-2 pass per
    chunk, random streams numbered per chunk) on ``depth`` handles that
    share one GPU's device data (``DeviceTracker.fork``), so consecutive
    chunks run concurrently: a chunk's phase 2 fills the GPU while the
    previous chunks' longest streamlines finish, instead of every chunk
    paying the persistent kernel's tail alone (cu_propagate_seeds.py:206-233
    runs chunk after chunk on one stream).

    A chunk is one bounded chunk (cuslines_chunk_async): phase 1, scan,
    phase 2 and compaction enqueued with a fixed slot capacity, with no host
    read between them (cu_propagate_seeds.py:95-201 reads the streamline
    count back to size the slots), and with ``graphs`` replayed as one HIP
    graph per handle.  Chunk k goes to handle k % depth; results are
    collected in chunk order.  A chunk whose streamlines exceed the slots
    runs again with more (the same results).  Buffers grow and are reused
    across chunks (freeing memory would synchronise the device)."""

    def __init__(self, dt, depth: int = 4, minlen=0, maxlen=np.inf, graphs: bool = True, rng_chunk=None):
        from ._device import ChunkBuffers

        self._CB = ChunkBuffers
        self.dt = dt
        # random streams restart every rng_chunk seeds of a chunk (None: once
        # per chunk, the reference's protocol; several reference chunks per
        # launch otherwise, as SeedBatchPropagator.rng_chunk)
        self.rng_chunk = rng_chunk
        self.handles = [dt] + [dt.fork() for _ in range(max(1, int(depth)) - 1)]
        self.cbs = [None] * len(self.handles)
        self.pending = [None] * len(self.handles)
        self.minlen, self.maxlen = minlen, maxlen
        self.graphs = bool(graphs)
        self.reruns = 0

    def _cb(self, k, n, cap):
        cb = self.cbs[k]
        if cb is None or cb.nmax < n or cb.cap < cap:
            if cb is not None:
                nmax, cap = max(n, cb.nmax), max(cap, 3 * cb.cap // 2)
                cb.close()
            else:
                nmax = n
            self.cbs[k] = None
            cb = self.cbs[k] = self._CB(self.handles[k], nmax, cap)
        return cb

    def _launch(self, k, seeds, slot_hint):
        n = len(seeds)
        cb = self._cb(k, n, max(slot_hint, 1))
        cb.sd.upload(seeds)
        cb.enqueue(n, self.rng_chunk or n, graph=self.graphs)
        self.pending[k] = seeds

    def _collect(self, k, sink=None):
        seeds = self.pending[k]
        self.pending[k] = None
        n = len(seeds)
        cb = self.cbs[k]
        nsl, tracked, npts = cb.totals()
        if tracked < nsl:  # the slots were too few: the same chunk with more
            self.reruns += 1
            cb = self._cb(k, n, nsl)
            cb.sd.upload(seeds)
            cb.enqueue(n, self.rng_chunk or n, graph=False)
            nsl, tracked, npts = cb.totals()
            assert tracked == nsl
        if sink is not None:
            r = cb.collect(n, with_seeds=False, points_out=sink.take(npts, n))
            return r["points"], r["sline_len"].astype(np.int64)
        if not nsl:
            return np.zeros((0, 3), np.float32), np.zeros(0, np.int64)
        r = cb.collect(n, with_seeds=False)
        pts, lens = r["points"], r["sline_len"].astype(np.int64)
        keep = (lens >= self.minlen) & (lens <= self.maxlen)
        if not keep.all():
            pts, lens = pts[np.repeat(keep, lens)], lens[keep]
        return pts, lens

    def run(self, seeds, chunk: int, slots_per_seed: float = 8, sink=None):
        """(points, lengths) of every chunk of ``chunk`` seeds, in order.
        ``slots_per_seed`` sizes the slots reserved up front (streamlines per
        seed are about 1 to 8; a chunk with more runs again).  With ``sink``
        (an array_sequence.PackedSink) each chunk's points are downloaded
        into it and yielded unfiltered, for the caller to ``sink.add`` before
        the next chunk; the sink's result applies the length filter."""
        seeds = np.ascontiguousarray(seeds, dtype=np.float32).reshape(-1, 3)
        nchunks = -(-len(seeds) // chunk) if len(seeds) else 0
        depth = len(self.handles)
        hint = max(1, int(slots_per_seed * chunk))
        order = []
        for c in range(nchunks):
            k = c % depth
            if self.pending[k] is not None:
                yield self._collect(k, sink)
                order.pop(0)
            self._launch(k, seeds[c * chunk:(c + 1) * chunk], hint)
            order.append(k)
        for k in order:
            yield self._collect(k, sink)

    def close(self):
        for k, cb in enumerate(self.cbs):
            if cb is not None:
                cb.close()
            self.cbs[k] = None
        for h in self.handles[1:]:
            h.close()
        self.handles = self.handles[:1]
