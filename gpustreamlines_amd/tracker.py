"""GPUTracker: device residency of one tracking problem
(cuslines/cuda_python/cu_tractography.py:36-300).

Same constructor signature, validation order, context-manager lifetime and
chunked output methods as the reference.  Each GPU gets one
``DeviceTracker`` (a ``cuslines_tracker`` handle of libcuslines_hip.so with
its own stream); the volume, stop map, sphere and getter data are uploaded
on ``__enter__`` and freed on ``__exit__``.
"""

from __future__ import annotations

import logging
from math import radians

import numpy as np

from . import _lib
from ._device import DeviceTracker
from .direction_getters import BootDirectionGetter, GPUDirectionGetter
from .generic_tracker import GenericTracker
from .propagate_seeds import SeedBatchPropagator

logger = logging.getLogger("GPUStreamlines")


class GPUTracker(GenericTracker):
    def __init__(self, dg: GPUDirectionGetter, dataf: np.ndarray, stop_map: np.ndarray,
                 stop_threshold: float, sphere_vertices: np.ndarray, sphere_edges: np.ndarray,
                 sphere_symm: bool = False, max_angle: float = radians(60), step_size: float = 0.5,
                 min_pts=0, max_pts=np.inf, relative_peak_thresh: float = 0.25,
                 min_separation_angle: float = radians(45), ngpus: int = 1, rng_seed: int = 0,
                 rng_offset: int = 0, chunk_size: int = 25000, launch_seeds: int = 250000,
                 chunk_pipeline: int = 8, chunk_graphs: bool = True):
        """Same parameters as the reference GPUTracker
        (cu_tractography.py:57-109); ``step_size`` is in voxels.
        ``launch_seeds`` (not in the reference): on one GPU, chunks are
        grouped so that one launch tracks about this many seeds; streams stay
        chunk-relative, so the output does not depend on it (0: one chunk per
        launch).  ``chunk_pipeline`` (not in the reference): with one chunk
        per launch, this many chunks run concurrently on handles sharing the
        device data; with grouped launches, two launches run at once, so one
        launch's points are downloaded while the next one tracks (1: one
        after the other, with slots sized from the read-back streamline
        count: the least memory; two 250,000-seed launches in flight hold
        about 98 GB of worst-case slots); the output does not depend on it.
        ``chunk_graphs`` (not in the reference): those launches run as one
        HIP graph per handle (phase 1, scan, phase 2 and compaction with no
        host read between them)."""
        _lib.load()
        self.dimx, self.dimy, self.dimz, self.dimt = dataf.shape
        if hasattr(dg, "prepare_data"):
            self.dataf = dg.prepare_data(dataf, stop_map, stop_threshold, sphere_vertices, sphere_symm)
        else:
            self.dataf = np.ascontiguousarray(dataf, dtype=np.float32)
        self.metric_map = np.ascontiguousarray(stop_map, dtype=np.float32)
        self.sphere_vertices = np.ascontiguousarray(sphere_vertices, dtype=np.float32)
        self.sphere_edges = np.ascontiguousarray(sphere_edges, dtype=np.int32)
        self.nedges = int(sphere_edges.shape[0])
        if isinstance(dg, BootDirectionGetter):
            self.samplm_nr = int(dg.sampling_matrix.shape[0])
            if len(self.sphere_vertices) != self.samplm_nr:
                raise ValueError("sampling_matrix rows must match the sphere vertex count")
        else:
            if len(self.sphere_vertices) != self.dimt:
                raise ValueError(
                    f"Number of sphere vertices ({len(self.sphere_vertices)}) must"
                    f" match last dimension of dataf ({self.dimt}), "
                    "because dataf should be an ODF when using prob or ptt tracking")
            self.samplm_nr = self.dimt
        self.n32dimt = ((self.dimt + 31) // 32) * 32
        self.dg = dg
        self.max_angle = np.float32(max_angle)
        self.tc_threshold = np.float32(stop_threshold)
        self.step_size = np.float32(step_size)
        self.relative_peak_thresh = np.float32(relative_peak_thresh)
        self.min_separation_angle = np.float32(min_separation_angle)
        self.ngpus = int(ngpus)
        self.rng_seed = int(rng_seed)
        self.rng_offset = int(rng_offset)
        self.chunk_size = int(chunk_size)
        self.launch_seeds = int(launch_seeds)
        self.chunk_pipeline = int(chunk_pipeline)
        self.chunk_graphs = bool(chunk_graphs)
        self.sphere_symm = bool(sphere_symm)
        avail = _lib.device_count()
        if self.ngpus > avail:
            raise RuntimeError(f"Requested {self.ngpus} GPUs but only {avail} available")
        logger.info("Creating GPUTracker with %d GPUs...", self.ngpus)
        self.device_trackers = []
        self.seed_propagator = SeedBatchPropagator(gpu_tracker=self, minlen=min_pts, maxlen=max_pts)
        self._allocated = False
        self.dg.compile_program(self)

    def __enter__(self):
        self._allocate()
        return self

    def _allocate(self):
        if self._allocated:
            return
        params = self.dg.macros
        for ii in range(self.ngpus):
            dt = DeviceTracker(params, device=ii)
            self.device_trackers.append(dt)
            on_device = getattr(self.dg, "prepare_on_device", False)
            dt.upload(None if on_device else self.dataf, self.metric_map, self.sphere_vertices, self.sphere_edges)
            self.dg.allocate_on_gpu(ii)
        self._allocated = True

    def __exit__(self, exc_type, exc, tb):
        logger.info("Destroying GPUTracker and freeing GPU memory...")
        for n, dt in enumerate(self.device_trackers):
            try:
                self.dg.deallocate_on_gpu(n)
            finally:
                dt.close()
        self.device_trackers = []
        self._allocated = False
        return False

    def _require(self):
        if not self._allocated:
            self._allocate()

    def generate_streamlines(self, seeds):
        self._require()
        return super().generate_streamlines(seeds)

    def generate_sft(self, seeds, ref_img):
        self._require()
        return super().generate_sft(seeds, ref_img)

    def generate_trx(self, seeds, ref_img):
        self._require()
        return super().generate_trx(seeds, ref_img)
