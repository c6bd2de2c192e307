"""ctypes binding of libcuslines_hip.so (declared in include/cuslines_hip.h).

The library is built in-tree for gfx950 by ``gpustreamlines_amd.build``.  If
it is missing, loading raises: there is no CPU fallback behind this API.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CUSLINES_HIP_LIB") or os.path.join(HERE, "libcuslines_hip.so")

ABI_VERSION = 1
MLP_NPARAM = 24856  # CUSLINES_MLP_NPARAM
PAIR_FLOATS = 64    # CUSLINES_PAIR_FLOATS


class cuslines_params(C.Structure):
    _fields_ = [
        ("dimx", C.c_int32), ("dimy", C.c_int32), ("dimz", C.c_int32), ("dimt", C.c_int32),
        ("model_type", C.c_int32), ("nverts", C.c_int32), ("nedges", C.c_int32),
        ("sphere_symm", C.c_int32), ("max_sline_len", C.c_int32), ("delta_nr", C.c_int32),
        ("rng_seed", C.c_uint64),
        ("step_size", C.c_float), ("max_angle", C.c_float), ("tc_threshold", C.c_float),
        ("relative_peak_thresh", C.c_float), ("min_separation_angle", C.c_float),
        ("min_signal", C.c_float),
        ("step_frac", C.c_int32), ("probe_quality", C.c_int32), ("lut_res", C.c_int32),
        ("tile_log2_x", C.c_int32), ("tile_log2_y", C.c_int32), ("tile_log2_z", C.c_int32),
        ("probe_step_size", C.c_float), ("max_curvature", C.c_float),
    ]


PARAM_DEFAULTS = dict(
    sphere_symm=0, max_sline_len=501, delta_nr=0, rng_seed=0, step_size=0.5,
    max_angle=np.radians(60), tc_threshold=0.1, relative_peak_thresh=0.25,
    min_separation_angle=np.radians(45), min_signal=1.0, step_frac=1, probe_quality=8,
    lut_res=128, tile_log2_x=0, tile_log2_y=0, tile_log2_z=0, probe_step_size=0.25,
    max_curvature=0.0,
)


def params_struct(params: dict) -> cuslines_params:
    """Fill a ``cuslines_params`` from a dict (floats rounded to float32, as
    the reference rounds its constructor arguments)."""
    p = dict(PARAM_DEFAULTS)
    p.update(params)
    s = cuslines_params()
    for name, typ in cuslines_params._fields_:
        v = p[name]
        if typ is C.c_float:
            v = float(np.float32(v))
        else:
            v = int(v)
        setattr(s, name, v)
    return s


class CuslinesError(RuntimeError):
    """Nonzero status from the HIP library (mirrors checkCudaErrors' hard
    mode, cuslines/cuda_python/cutils.py:56-75)."""


_lib = None


def load():
    """Load the in-tree gfx950 library; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with gpustreamlines_amd.build.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp, sz = C.c_void_p, C.c_size_t
    i32 = C.c_int32
    sigs = {
        "cuslines_abi_version": [],
        "cuslines_last_error": [C.c_char_p, sz],
        "cuslines_device_count": [C.POINTER(C.c_int)],
        "cuslines_tracker_create": [C.POINTER(cuslines_params), C.c_int, C.POINTER(vp)],
        "cuslines_tracker_destroy": [vp],
        "cuslines_tracker_fork": [vp, C.POINTER(vp)],
        "cuslines_set_data": [vp, vp],
        "cuslines_set_stop_map": [vp, vp],
        "cuslines_set_sphere": [vp, vp, vp],
        "cuslines_set_boot_matrices": [vp, vp, vp, vp, vp, vp, vp],
        "cuslines_set_ptt_lut": [vp, vp],
        "cuslines_ptt_prepare": [vp, vp, vp, C.c_float, C.c_int, vp],
        "cuslines_device_alloc": [vp, sz, C.POINTER(vp)],
        "cuslines_device_free": [vp, vp],
        "cuslines_memcpy_htod": [vp, vp, vp, sz],
        "cuslines_memcpy_dtoh": [vp, vp, vp, sz],
        "cuslines_synchronize": [vp],
        "cuslines_get_stream": [vp, C.POINTER(vp)],
        "cuslines_set_launch": [vp, C.c_int, C.c_int, C.c_int],
        "cuslines_set_ptt_layout": [vp, C.c_int],
        "cuslines_set_lanes_per_seed": [vp, C.c_int],
        "cuslines_set_stepping": [vp, C.c_int],
        "cuslines_set_boot_mode": [vp, C.c_int],
        "cuslines_boot_counters": [vp, C.POINTER(C.c_int64), C.c_int],
        "cuslines_boot_lowrank": [vp, C.POINTER(i32)],
        "cuslines_hat_lowrank": [vp, vp, i32, C.POINTER(i32)],
        "cuslines_set_proposal": [vp, vp],
        "cuslines_set_recording": [vp, vp, i32, vp, C.c_float],
        "cuslines_proposal_stats": [vp, C.POINTER(C.c_int64), C.c_int],
        "cuslines_proposal_counters": [vp, C.POINTER(C.c_int64), C.c_int],
        "cuslines_set_tap_count": [vp, C.c_int],
        "cuslines_tap_count": [vp, C.POINTER(C.c_int64), C.c_int],
        "cuslines_mlp_workspace_bytes": [i32, C.POINTER(sz)],
        "cuslines_mlp_loss_grad_async": [vp, vp, vp, i32, vp, vp, vp],
        "cuslines_mlp_adam_async": [vp, vp, vp, vp, vp, i32, C.c_float, C.c_float, C.c_float, C.c_float],
        "cuslines_phase1_async": [vp, vp, i32, i32, vp, vp],
        "cuslines_scan_async": [vp, vp, i32, vp],
        "cuslines_phase2_async": [vp, vp, i32, i32, vp, vp, vp, vp, vp],
        "cuslines_phase2_noalloc_async": [vp, vp, i32, i32, vp, vp, vp, vp, vp],
        "cuslines_compact_async": [vp, vp, i32, vp, vp, vp],
        "cuslines_chunk_workspace_bytes": [vp, i32, i32, C.POINTER(sz)],
        "cuslines_chunk_async": [vp, vp, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, sz],
        "cuslines_chunk_graph_create": [vp, vp, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, sz,
                                        C.POINTER(vp)],
        "cuslines_graph_launch": [vp],
        "cuslines_graph_destroy": [vp],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = args
    if L.cuslines_abi_version() != ABI_VERSION:
        raise ImportError("libcuslines_hip.so ABI version mismatch")
    _lib = L
    return L


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    load().cuslines_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int, what: str = "", hard: bool = True):
    if status != 0:
        msg = f"cuslines HIP error {status} in {what}: {last_error()}"
        if hard:
            raise CuslinesError(msg)
        import logging

        logging.getLogger("GPUStreamlines").warning(msg)


def device_count() -> int:
    n = C.c_int(0)
    check(load().cuslines_device_count(C.byref(n)), "device_count")
    return n.value


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data
