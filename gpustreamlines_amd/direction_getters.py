"""Direction-getter plugins (host side) for the gfx950 tracking kernels.

Mirrors the reference's plugin protocol (``GPUDirectionGetter``,
cuslines/cuda_python/cu_direction_getters.py:35-144) and its three getters
(``BootDirectionGetter`` :147-364, ``ProbDirectionGetter`` :367-414,
``PttDirectionGetter`` :417-648).  The kernels are precompiled for gfx950
and specialised by a runtime parameter block (``cuslines_params``) instead
of NVRTC macros; ``compile_program``/``set_macros`` build that block.
"""

from __future__ import annotations

import logging
import math
from abc import ABC, abstractmethod

import numpy as np

from . import _lib
from ._globals import MAX_SLINE_LEN
from .boot_utils import prepare_csa, prepare_opdt

logger = logging.getLogger("GPUStreamlines")

MODEL_TYPES = {"OPDT": 0, "CSA": 1, "PROB": 2, "PTT": 3}


class GPUDirectionGetter(ABC):
    """Plugin interface of one direction getter (cu_direction_getters.py:35-51)."""

    model_type = None

    @abstractmethod
    def getNumStreamlines(self, n, nseeds_gpu, block, grid, sp):
        """Phase 1 on device ``n`` for ``sp``'s current seed share."""

    @abstractmethod
    def generateStreamlines(self, n, nseeds_gpu, block, grid, sp):
        """Phase 2 on device ``n`` for ``sp``'s current seed share."""

    def set_macros(self, gpu_tracker):
        pass

    def allocate_on_gpu(self, n):
        pass

    def deallocate_on_gpu(self, n):
        pass

    def compile_program(self, gpu_tracker, debug: bool = False):
        """Build the runtime parameter block of the precompiled kernels
        (replaces the NVRTC macro set, cu_direction_getters.py:70-110)."""
        _lib.load()
        t = gpu_tracker
        self.gpu_tracker = t
        self.macros = dict(
            dimx=t.dimx, dimy=t.dimy, dimz=t.dimz, dimt=t.dimt,
            model_type=self.model_type,
            nverts=t.samplm_nr, nedges=t.nedges, sphere_symm=int(t.sphere_symm),
            max_sline_len=MAX_SLINE_LEN, delta_nr=getattr(self, "delta_nr", 0),
            rng_seed=t.rng_seed, step_size=t.step_size, max_angle=t.max_angle,
            tc_threshold=t.tc_threshold, relative_peak_thresh=t.relative_peak_thresh,
            min_separation_angle=t.min_separation_angle,
            min_signal=float(getattr(self, "min_signal", 1.0)),
        )
        self.set_macros(t)
        if debug:
            logger.info("GPUStreamlines parameters: %s", self.macros)
        return self.macros

    # Phase launches through the device tracker of GPU ``n`` -------------
    def _phase1(self, n, nseeds_gpu, sp):
        dt = sp.gpu_tracker.device_trackers[n]
        dt.phase1_async(sp.seeds_d[n].ptr, nseeds_gpu, sp.rng_chunk or nseeds_gpu, sp.slinesOffs_d[n].ptr,
                        sp.shDirTemp0_d[n].ptr)

    def _phase2(self, n, nseeds_gpu, sp):
        dt = sp.gpu_tracker.device_trackers[n]
        dt.phase2_async(sp.seeds_d[n].ptr, nseeds_gpu, sp.rng_chunk or nseeds_gpu, sp.slinesScan_d[n].ptr,
                        sp.shDirTemp0_d[n].ptr, sp.slineSeed_d[n].ptr, sp.slineLen_d[n].ptr,
                        sp.sline_d[n].ptr)


class BootDirectionGetter(GPUDirectionGetter):
    """Residual-bootstrap getter for the OPDT or CSA model
    (cu_direction_getters.py:147-215)."""

    def __init__(self, model_type: str, min_signal: float, H: np.ndarray, R: np.ndarray,
                 delta_b: np.ndarray, delta_q: np.ndarray, sampling_matrix: np.ndarray,
                 b0s_mask: np.ndarray, mfma: bool = True):
        """``mfma`` (not in the reference): True runs phase 2 on the batched
        matrix-core kernel (16 seeds in lockstep; results equal the
        reference's within float rounding, peak choices flip only at
        near-ties); False runs the per-seed kernel that is bit-identical to
        the CPU oracle."""
        self.mfma = bool(mfma)
        mt = str(model_type).upper()
        if mt not in ("OPDT", "CSA"):
            raise ValueError(f"Invalid model_type {model_type}, must be one of 'OPDT', 'CSA'")
        self.model_type = MODEL_TYPES[mt]
        self.H = np.ascontiguousarray(H, dtype=np.float32)
        self.R = np.ascontiguousarray(R, dtype=np.float32)
        self.delta_b = np.ascontiguousarray(delta_b, dtype=np.float32)
        self.delta_q = np.ascontiguousarray(delta_q, dtype=np.float32)
        self.delta_nr = int(delta_b.shape[0])
        self.min_signal = np.float32(min_signal)
        self.sampling_matrix = np.ascontiguousarray(sampling_matrix, dtype=np.float32)
        self.b0s_mask = np.ascontiguousarray(b0s_mask, dtype=np.int32)
        hr = int((self.b0s_mask == 0).sum())
        if self.H.shape != (hr, hr) or self.R.shape != (hr, hr):
            raise ValueError("H and R must be (n_dwi, n_dwi)")
        if self.delta_b.shape != (self.delta_nr, hr) or self.delta_q.shape != (self.delta_nr, hr):
            raise ValueError("delta_b and delta_q must be (n_coef, n_dwi)")
        if self.sampling_matrix.shape[1] != self.delta_nr:
            raise ValueError("sampling_matrix must be (n_vertices, n_coef)")

    @classmethod
    def from_dipy_opdt(cls, gtab, sphere, sh_order_max=6, full_basis=False, sh_lambda=0.006, min_signal=1):
        return cls(**prepare_opdt(gtab, sphere, sh_order_max, full_basis, sh_lambda, min_signal))

    @classmethod
    def from_dipy_csa(cls, gtab, sphere, sh_order_max=6, full_basis=False, sh_lambda=0.006, min_signal=1):
        return cls(**prepare_csa(gtab, sphere, sh_order_max, full_basis, sh_lambda, min_signal))

    def boot_arrays(self):
        return dict(H=self.H, R=self.R, delta_b=self.delta_b, delta_q=self.delta_q,
                    sampling_matrix=self.sampling_matrix, b0s_mask=self.b0s_mask)

    def allocate_on_gpu(self, n):
        dt = self.gpu_tracker.device_trackers[n]
        dt.set_boot(**self.boot_arrays())
        dt.set_boot_mode(dt.BOOT_MFMA if self.mfma else dt.BOOT_EXACT)

    def getNumStreamlines(self, n, nseeds_gpu, block, grid, sp):
        self._phase1(n, nseeds_gpu, sp)

    def generateStreamlines(self, n, nseeds_gpu, block, grid, sp):
        self._phase2(n, nseeds_gpu, sp)


class ProbDirectionGetter(GPUDirectionGetter):
    """Probabilistic getter on ODF amplitudes (cu_direction_getters.py:367-414)."""

    model_type = MODEL_TYPES["PROB"]

    def __init__(self):
        pass

    def getNumStreamlines(self, n, nseeds_gpu, block, grid, sp):
        self._phase1(n, nseeds_gpu, sp)

    def generateStreamlines(self, n, nseeds_gpu, block, grid, sp):
        self._phase2(n, nseeds_gpu, sp)


def calculate_tight_fit(dimt: int):
    """Power-of-two (nx, ny, nz) slot tiling with nx*ny*nz >= dimt, doubling
    the smallest axis first (cu_direction_getters.py:550-562)."""
    nx, ny, nz = 1, 1, 1
    while nx * ny * nz < dimt:
        if nx <= ny and nx <= nz:
            nx *= 2
        elif ny <= nx and ny <= nz:
            ny *= 2
        else:
            nz *= 2
    return nx, ny, nz


def nearest_vertex_lut(vertices, res=128, sphere_symm=False):
    """Index of the nearest sphere vertex for every point of the
    ``linspace(-1, 1, res)^3`` grid (cu_direction_getters.py:601-623)."""
    from scipy.spatial import KDTree

    coords = np.linspace(-1, 1, res)
    gx, gy, gz = np.meshgrid(coords, coords, coords, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    v = np.asarray(vertices, dtype=np.float64)
    if sphere_symm:
        tree = KDTree(np.concatenate([v, -v], axis=0))
        _, idx = tree.query(pts)
        idx = np.concatenate([np.arange(len(v)), np.arange(len(v))])[idx]
    else:
        _, idx = KDTree(v).query(pts)
    return np.ascontiguousarray(idx.reshape(res, res, res), dtype=np.int32)


class PttDirectionGetter(ProbDirectionGetter):
    """Parallel Transport Tractography getter (cu_direction_getters.py:417-648)."""

    model_type = MODEL_TYPES["PTT"]

    def __init__(self, odf_lut_res: int = 128, probe_length: float = 0.25,
                 target_short_step: float = 0.025, probe_quality: int = 8, proposal=None,
                 prepare_on_device: bool = False):
        """``proposal`` (not in the reference): parameters of a trained
        learned-proposal network (float32, or bf16 bit patterns as uint16;
        see gpustreamlines_amd.learned).  None keeps the reference's face-CDF
        proposal.  ``prepare_on_device`` (not in the reference): clip,
        mask, normalise and build the LUT on the GPU (cuslines_ptt_prepare)
        instead of on the host; same results."""
        self.prepare_on_device = bool(prepare_on_device)
        self.odf_lut_res = int(odf_lut_res)
        self.probe_length = probe_length
        self.probe_quality = int(probe_quality)
        self.target_short_step = target_short_step
        self.sphere_vertices_lut_h = None
        self.proposal = proposal

    def set_macros(self, gpu_tracker):
        step = np.float32(gpu_tracker.step_size)
        self.macros.update(
            tile_log2_x=self.log2_x, tile_log2_y=self.log2_y, tile_log2_z=self.log2_z,
            probe_quality=self.probe_quality, probe_step_size=float(self.probe_length),
            step_frac=int(np.round(step / self.target_short_step)),
            max_curvature=float(8.0 * np.sin(np.float32(gpu_tracker.max_angle) / 2.0)),
            lut_res=self.odf_lut_res,
        )

    def prepare_data(self, dataf, stop_map, stop_threshold, sphere_vertices, sphere_symm):
        """Clip ODFs at 0, zero them where the stop map is below threshold,
        normalise each voxel to max 1, choose the vertex tiling and build the
        nearest-vertex LUT (cu_direction_getters.py:531-625).  The tiling is
        kept as index arithmetic; the data stay [X][Y][Z][T] on the host."""
        dimx, dimy, dimz, dimt = dataf.shape
        nx, ny, nz = calculate_tight_fit(dimt)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.log2_x, self.log2_y, self.log2_z = (int(math.log2(nx)), int(math.log2(ny)), int(math.log2(nz)))
        if self.prepare_on_device:
            # the raw ODF goes to the device in allocate_on_gpu
            self._device_inputs = (np.asarray(stop_map, dtype=np.float32), float(stop_threshold), bool(sphere_symm))
            return np.ascontiguousarray(dataf, dtype=np.float32)
        dataf = np.asarray(dataf, dtype=np.float32).clip(min=0)
        dataf[np.asarray(stop_map) < stop_threshold, :] = 0
        mx = dataf.max(axis=3, keepdims=True)
        np.divide(dataf, mx, out=dataf, where=mx > 0)
        self.sphere_vertices_lut_h = nearest_vertex_lut(sphere_vertices, self.odf_lut_res, sphere_symm)
        return np.ascontiguousarray(dataf, dtype=np.float32)

    def allocate_on_gpu(self, n):
        dt = self.gpu_tracker.device_trackers[n]
        if self.prepare_on_device:
            stop, thr, symm = self._device_inputs
            self.sphere_vertices_lut_h = dt.ptt_prepare(self.gpu_tracker.dataf, stop, thr, symm)
        else:
            dt.set_lut(self.sphere_vertices_lut_h)
        if self.proposal is not None:
            from .learned import to_bf16

            w = np.asarray(self.proposal)
            dt.set_proposal(w if w.dtype == np.uint16 else to_bf16(w))
