"""Synthetic tracking workloads (inputs for both the GPU path and checkers).

A workload is the set of arrays and parameters one direction getter tracks
on: the phantom volume (DWI for bootstrap, CSA ODF amplitudes on the sphere
for prob/PTT, clipped at 0 and zero below the FA threshold, as the
reference CLI prepares its ODF, run_gpu_streamlines.py:357-380), the FA stop
map, the 724-vertex sphere and edges, the getter matrices or PTT LUT, the
parameter block of ``cuslines_params`` and the seeds.
"""

from __future__ import annotations

import functools
from dataclasses import dataclass, field

import numpy as np

from . import shm
from ._globals import MAX_SLINE_LEN
from .boot_utils import prepare_csa, prepare_opdt
from .direction_getters import MODEL_TYPES, PttDirectionGetter
from .phantom import make_phantom, mixed_seeds, seeds_in_mask
from .sphere import symmetric_sphere

STOP_THRESHOLD = 0.1


@dataclass
class Workload:
    kind: str
    params: dict
    dataf: np.ndarray
    stop_map: np.ndarray
    vertices: np.ndarray
    edges: np.ndarray
    boot: dict | None = None
    ptt_lut: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


@functools.lru_cache(maxsize=4)
def _volume(shape, seed):
    ph = make_phantom(shape, seed=seed)
    verts, edges = symmetric_sphere(724)
    return ph, verts, edges


@functools.lru_cache(maxsize=4)
def _odf(shape, seed):
    ph, verts, _ = _volume(shape, seed)
    b0 = ph.gtab.b0s_mask
    fit = shm.csa_fit_matrix(ph.gtab.bvecs[~b0], 6, 0.006)
    _, th, phi = shm.cart2sphere(*verts.astype(np.float64).T)
    smat, _, _ = shm.real_sh_descoteaux(6, th, phi)
    odf = shm.csa_odf(ph.dwi, b0, fit, smat).clip(min=0)
    odf[ph.fa < STOP_THRESHOLD, :] = 0
    return np.ascontiguousarray(odf, dtype=np.float32)


def base_params(shape, dimt, nverts, nedges, step_size, model):
    return dict(dimx=shape[0], dimy=shape[1], dimz=shape[2], dimt=dimt, model_type=MODEL_TYPES[model],
                nverts=nverts, nedges=nedges, sphere_symm=0, max_sline_len=MAX_SLINE_LEN,
                rng_seed=0, step_size=step_size, max_angle=np.radians(60), tc_threshold=STOP_THRESHOLD,
                relative_peak_thresh=0.25, min_separation_angle=np.radians(45), min_signal=1.0)


def build(kind: str, shape=(81, 106, 76), step_size=0.25, seed=0) -> Workload:
    """``kind`` in {"prob", "ptt", "csa", "opdt"}.  ``step_size`` in voxels
    (0.5 mm at 2 mm voxels = 0.25)."""
    shape = tuple(int(s) for s in shape)
    ph, verts, edges = _volume(shape, seed)
    if kind in ("csa", "opdt"):
        prep = (prepare_csa if kind == "csa" else prepare_opdt)(ph.gtab, verts)
        boot = {k: np.ascontiguousarray(prep[k], dtype=np.float32)
                for k in ("H", "R", "delta_b", "delta_q", "sampling_matrix")}
        boot["b0s_mask"] = np.asarray(prep["b0s_mask"], dtype=np.int32)
        p = base_params(shape, ph.dwi.shape[3], len(verts), len(edges), step_size, kind.upper())
        p["delta_nr"] = boot["delta_b"].shape[0]
        return Workload(kind, p, ph.dwi, ph.fa, verts, edges, boot=boot, extra=dict(phantom=ph))
    odf = _odf(shape, seed)
    if kind == "prob":
        p = base_params(shape, odf.shape[3], len(verts), len(edges), step_size, "PROB")
        return Workload(kind, p, odf, ph.fa, verts, edges, extra=dict(phantom=ph))
    if kind == "ptt":
        dg = PttDirectionGetter()
        data = dg.prepare_data(odf, ph.fa, STOP_THRESHOLD, verts, False)
        p = base_params(shape, odf.shape[3], len(verts), len(edges), step_size, "PTT")
        p.update(tile_log2_x=dg.log2_x, tile_log2_y=dg.log2_y, tile_log2_z=dg.log2_z,
                 probe_quality=dg.probe_quality, probe_step_size=dg.probe_length,
                 step_frac=int(np.round(np.float32(step_size) / dg.target_short_step)),
                 max_curvature=float(8.0 * np.sin(np.float32(p["max_angle"]) / 2.0)),
                 lut_res=dg.odf_lut_res)
        return Workload(kind, p, data, ph.fa, verts, edges, ptt_lut=dg.sphere_vertices_lut_h,
                        extra=dict(phantom=ph, odf=odf))
    raise ValueError(f"unknown workload kind {kind!r}")


def seeds_for(w: Workload, n: int, placement: str = "mask", seed: int = 1) -> np.ndarray:
    """``placement``: "mask" (uniform in FA > 0.2) or "mixed" (half in the mask,
    half in 0.05 <= FA <= 0.15 border voxels)."""
    ph = w.extra["phantom"]
    if placement == "mask":
        return seeds_in_mask(ph.mask, n, seed=seed)
    if placement == "mixed":
        return mixed_seeds(ph.fa, ph.mask, n, seed=seed)
    raise ValueError(placement)


LARGE_SHAPE = (145, 174, 145)


@functools.lru_cache(maxsize=1)
def _large_volume(seed):
    from .phantom import make_large_phantom

    ph = make_large_phantom(seed=seed)
    verts, edges = symmetric_sphere(724)
    return ph, verts, edges


def build_large(kind: str, step_size=0.4, seed=0) -> Workload:
    """The large multi-shell workload: 145x174x145 at 1.25 mm, 6 b0 + 3 x 90
    DWI (b = 1000/2000/3000), SNR 25, SH order 8 (45 coefficients), step
    0.4 voxel.  ``kind`` in {"csa", "opdt", "ptt"}; the PTT input is the CSA
    ODF of order 8 on the 724-vertex sphere, clipped at 0 and zero below the
    FA threshold, prepared as the reference's getter does."""
    ph, verts, edges = _large_volume(seed)
    shape = LARGE_SHAPE
    if kind in ("csa", "opdt"):
        prep = (prepare_csa if kind == "csa" else prepare_opdt)(ph.gtab, verts, sh_order_max=8)
        boot = {k: np.ascontiguousarray(prep[k], dtype=np.float32)
                for k in ("H", "R", "delta_b", "delta_q", "sampling_matrix")}
        boot["b0s_mask"] = np.asarray(prep["b0s_mask"], dtype=np.int32)
        p = base_params(shape, ph.dwi.shape[3], len(verts), len(edges), step_size, kind.upper())
        p["delta_nr"] = boot["delta_b"].shape[0]
        return Workload(kind, p, ph.dwi, ph.fa, verts, edges, boot=boot, extra=dict(phantom=ph))
    if kind == "ptt":
        b0 = ph.gtab.b0s_mask
        fit = shm.csa_fit_matrix(ph.gtab.bvecs[~b0], 8, 0.006)
        _, th, phi = shm.cart2sphere(*verts.astype(np.float64).T)
        smat, _, _ = shm.real_sh_descoteaux(8, th, phi)
        odf = shm.csa_odf(ph.dwi, b0, fit, smat).clip(min=0)
        odf[ph.fa < STOP_THRESHOLD, :] = 0
        odf = np.ascontiguousarray(odf, dtype=np.float32)
        dg = PttDirectionGetter()
        data = dg.prepare_data(odf, ph.fa, STOP_THRESHOLD, verts, False)
        p = base_params(shape, odf.shape[3], len(verts), len(edges), step_size, "PTT")
        p.update(tile_log2_x=dg.log2_x, tile_log2_y=dg.log2_y, tile_log2_z=dg.log2_z,
                 probe_quality=dg.probe_quality, probe_step_size=dg.probe_length,
                 step_frac=int(np.round(np.float32(step_size) / dg.target_short_step)),
                 max_curvature=float(8.0 * np.sin(np.float32(p["max_angle"]) / 2.0)),
                 lut_res=dg.odf_lut_res)
        return Workload(kind, p, data, ph.fa, verts, edges, ptt_lut=dg.sphere_vertices_lut_h, extra=dict(phantom=ph))
    raise ValueError(f"unknown large workload kind {kind!r}")
