"""ctypes loader for the CPU oracle (test infrastructure only).

Only the tests, the repository's smoke check and bench.py's CPU-baseline leg
import this module.  It exposes the same phase protocol as the gfx950 library on
host arrays, so a parity test runs one input through both and compares.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcuslines_oracle.so")
# CUSLINES_ORACLE_LIB selects another build of the same source, e.g. the
# sanitizer build of `make -C oracle sanitize`.
LOAD_PATH = os.environ.get("CUSLINES_ORACLE_LIB") or LIB_PATH

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if LOAD_PATH == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LOAD_PATH)
        vp, fp, ip, up = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_uint32)
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [vp]
        L.oracle_destroy.argtypes = [vp]
        for name, args in [
            ("oracle_set_data", [vp, fp]),
            ("oracle_set_stop_map", [vp, fp]),
            ("oracle_set_sphere", [vp, fp, ip]),
            ("oracle_set_boot_matrices", [vp, fp, fp, fp, fp, fp, ip]),
            ("oracle_set_ptt_lut", [vp, ip]),
            ("oracle_phase1", [vp, fp, C.c_int, C.c_int, ip, ip]),
            ("oracle_phase2", [vp, fp, C.c_int, C.c_int, ip, ip, ip, ip, fp]),
            ("oracle_peaks", [vp, fp, ip]),
            ("oracle_closest_peak", [vp, fp, C.c_int, ip, fp]),
            ("oracle_trilinear", [vp, fp, fp]),
            ("oracle_check_point", [vp, fp]),
            ("oracle_boot_odf", [vp, fp, C.c_int, C.c_int, fp, fp]),
            ("oracle_set_proposal", [vp, C.POINTER(C.c_uint16)]),
            ("oracle_proposal_stats", [vp, C.POINTER(C.c_longlong), C.c_int]),
            ("oracle_proposal_counters", [vp, C.POINTER(C.c_longlong), C.c_int]),
            ("oracle_ptt_learned_substep", [vp, fp, fp, fp, C.c_longlong, fp, fp, C.POINTER(C.c_longlong)]),
        ]:
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = args
        L.oracle_philox_draws.argtypes = [C.c_uint64, C.c_uint64, C.c_int, up]
        L.oracle_philox_mixed.argtypes = [C.c_uint64, C.c_uint64, C.c_int, ip, fp]
        L.oracle_log.restype = C.c_float
        L.oracle_log.argtypes = [C.c_float]
        L.oracle_sincos.argtypes = [C.c_float, fp, fp]
        L.oracle_rsqrt.restype = C.c_float
        L.oracle_rsqrt.argtypes = [C.c_float]
        L.oracle_sortnet_partner.restype = C.c_int
        L.oracle_sortnet_partner.argtypes = [C.c_int, C.c_int]
        L.oracle_sortnet_layers.restype = C.c_int
        L.oracle_warp_sort_dec.argtypes = [fp, ip]
        L.oracle_prefix_sum.argtypes = [fp, C.c_int]
        L.oracle_interp4.restype = C.c_float
        L.oracle_interp4.argtypes = [vp, fp, fp]
        L.oracle_prepare_propagator.argtypes = [C.c_float, C.c_float, C.c_float, fp]
        L.oracle_propagate_frame.argtypes = [fp, fp, fp]
        L.oracle_num_threads.restype = C.c_int
        L.oracle_set_num_threads.argtypes = [C.c_int]
        L.oracle_prob_select.restype = C.c_int
        L.oracle_prob_select.argtypes = [fp, C.c_int, C.c_float]
        L.oracle_prob_sample.restype = C.c_int
        L.oracle_prob_sample.argtypes = [vp, fp, fp, C.c_int, C.c_longlong, ip, fp]
        L.oracle_ptt_substep.restype = C.c_int
        L.oracle_ptt_substep.argtypes = [vp, fp, fp, C.c_int, C.c_longlong, fp, fp, ip, fp, fp]
        L.oracle_ptt_init_frame.restype = C.c_int
        L.oracle_ptt_init_frame.argtypes = [vp, fp, fp, C.c_longlong, fp]
        L.oracle_mlp_forward.argtypes = [C.POINTER(C.c_uint16), fp, fp]
        L.oracle_mlp_features.argtypes = [vp, fp, fp, fp, fp]
        L.oracle_vertex_supports.argtypes = [vp, fp, fp, fp]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Oracle:
    """One tracking problem on the CPU oracle (see gpustreamlines_amd.tracker
    for the parameter set, shared through ``cuslines_params``)."""

    def __init__(self, params, dataf, stop_map, vertices, edges, boot=None, ptt_lut=None):
        from gpustreamlines_amd._lib import params_struct

        self.L = lib()
        self.params = params
        self._ps = params_struct(params)
        self.h = self.L.oracle_create(C.byref(self._ps))
        if not self.h:
            raise ValueError("oracle rejected the parameters")
        self._keep = []
        d = np.ascontiguousarray(dataf, dtype=np.float32)
        s = np.ascontiguousarray(stop_map, dtype=np.float32)
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        e = np.ascontiguousarray(edges, dtype=np.int32)
        self.L.oracle_set_data(self.h, _f(d))
        self.L.oracle_set_stop_map(self.h, _f(s))
        self.L.oracle_set_sphere(self.h, _f(v), _i(e))
        if boot is not None:
            mats = [np.ascontiguousarray(boot[k], dtype=np.float32)
                    for k in ("H", "R", "delta_b", "delta_q", "sampling_matrix")]
            b0 = np.ascontiguousarray(boot["b0s_mask"], dtype=np.int32)
            self.L.oracle_set_boot_matrices(self.h, *[_f(m) for m in mats], _i(b0))
        if ptt_lut is not None:
            lut = np.ascontiguousarray(ptt_lut, dtype=np.int32)
            self.L.oracle_set_ptt_lut(self.h, _i(lut))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oracle_destroy(self.h)
            self.h = None

    def phase1(self, seeds, rng_chunk=0):
        seeds = np.ascontiguousarray(seeds, dtype=np.float32)
        n = len(seeds)
        counts = np.zeros(n + 1, dtype=np.int32)
        dirs = np.zeros((n, self.params["nverts"]), dtype=np.int32)
        self.L.oracle_phase1(self.h, _f(seeds), n, rng_chunk, _i(counts), _i(dirs))
        return counts, dirs

    def phase2(self, seeds, offsets, dirs, rng_chunk=0):
        seeds = np.ascontiguousarray(seeds, dtype=np.float32)
        n = len(seeds)
        nsl = int(offsets[n])
        L = self.params["max_sline_len"]
        sseed = np.full(nsl, -1, dtype=np.int32)
        slen = np.zeros(nsl, dtype=np.int32)
        sl = np.zeros((nsl, 2 * L, 3), dtype=np.float32)
        self.L.oracle_phase2(self.h, _f(seeds), n, rng_chunk, _i(np.ascontiguousarray(offsets, dtype=np.int32)),
                             _i(np.ascontiguousarray(dirs, dtype=np.int32)), _i(sseed), _i(slen), _f(sl))
        return sseed, slen, sl

    def track(self, seeds, rng_chunk=0):
        """Both phases with the host exclusive scan in between."""
        counts, dirs = self.phase1(seeds, rng_chunk)
        n = len(seeds)
        offsets = np.zeros(n + 1, dtype=np.int32)
        np.cumsum(counts[:n], out=offsets[1:])
        sseed, slen, sl = self.phase2(seeds, offsets, dirs, rng_chunk)
        return dict(counts=counts[:n], dirs=dirs, offsets=offsets, sline_seed=sseed,
                    sline_len=slen, slines=sl)

    # unit-level helpers ---------------------------------------------------
    def peaks(self, odf):
        odf = np.ascontiguousarray(odf, dtype=np.float32)
        out = np.zeros(self.params["nverts"], dtype=np.int32)
        n = self.L.oracle_peaks(self.h, _f(odf), _i(out))
        return out[:n]

    def closest_peak(self, direction, peaks):
        """closest_peak_d (boot.cu:115-193): (found, direction) for vertex
        indices ``peaks`` of the sphere."""
        d = np.ascontiguousarray(direction, dtype=np.float32)
        pk = np.ascontiguousarray(peaks, dtype=np.int32)
        out = np.zeros(3, dtype=np.float32)
        ok = self.L.oracle_closest_peak(self.h, _f(d), len(pk), _i(pk), _f(out))
        return ok, out

    def trilinear(self, pt):
        pt = np.ascontiguousarray(pt, dtype=np.float32)
        out = np.zeros(self.params["dimt"], dtype=np.float32)
        rv = self.L.oracle_trilinear(self.h, _f(pt), _f(out))
        return rv, out

    def check_point(self, pt):
        pt = np.ascontiguousarray(pt, dtype=np.float32)
        return self.L.oracle_check_point(self.h, _f(pt))

    def set_learned(self, mlp_bf16):
        """Learned PTT proposal: bf16 parameters (uint16), or None for the
        plain getter."""
        if mlp_bf16 is None:
            self.L.oracle_set_proposal(self.h, None)
        else:
            w = np.ascontiguousarray(mlp_bf16, dtype=np.uint16)
            self.L.oracle_set_proposal(self.h, w.ctypes.data_as(C.POINTER(C.c_uint16)))

    def mlp_features(self, pos, direction, frame):
        x = np.zeros(40, np.float32)
        self.L.oracle_mlp_features(self.h, _f(np.asarray(pos, np.float32)), _f(np.asarray(direction, np.float32)),
                                   _f(np.asarray(frame, np.float32)), _f(x))
        return x

    def vertex_supports(self, pos, frame):
        out = np.zeros(24, np.float32)
        self.L.oracle_vertex_supports(self.h, _f(np.asarray(pos, np.float32)), _f(np.asarray(frame, np.float32)),
                                      _f(out))
        return out

    def proposal_stats(self, reset=True):
        out = (C.c_longlong * 3)()
        self.L.oracle_proposal_stats(self.h, out, int(reset))
        return np.array(list(out), dtype=np.int64)

    def proposal_counters(self, reset=True):
        """proposal_stats plus [3]: face draws that broke the envelope."""
        out = (C.c_longlong * 4)()
        self.L.oracle_proposal_counters(self.h, out, int(reset))
        return np.array(list(out), dtype=np.int64)

    def boot_odf(self, pt, slot=0, phase=0):
        """One bootstrap refit at pt: (status, odf[nverts], signal[dimt])."""
        pt = np.ascontiguousarray(pt, dtype=np.float32)
        odf = np.zeros(self.params["nverts"], dtype=np.float32)
        sig = np.zeros(self.params["dimt"], dtype=np.float32)
        rv = self.L.oracle_boot_odf(self.h, _f(pt), int(slot), int(phase), _f(odf), _f(sig))
        return rv, odf, sig

    def prob_sample(self, pt, direction, nstreams, sub0=0):
        """Non-initial prob draws at ``pt`` with incoming ``direction`` on
        ``nstreams`` lane-0 streams: (vertex index or -1, direction) each."""
        idx = np.zeros(nstreams, np.int32)
        dirs = np.zeros((nstreams, 3), np.float32)
        rc = self.L.oracle_prob_sample(self.h, _f(np.asarray(pt, np.float32)), _f(np.asarray(direction, np.float32)),
                                       int(nstreams), int(sub0), _i(idx), _f(dirs))
        assert rc == 0
        return idx, dirs

    def ptt_substep(self, pos, frame, init=False, sub0=0):
        """One PTT sub-step; returns (ok, dict(k [32,2], sup [32], win, round,
        vert_pdf [24], frame [9]))."""
        k = np.zeros(64, np.float32)
        sup = np.zeros(32, np.float32)
        wr = np.zeros(2, np.int32)
        vp = np.zeros(24, np.float32)
        fo = np.zeros(9, np.float32)
        fr = np.ascontiguousarray(frame, np.float32).reshape(9)
        ok = self.L.oracle_ptt_substep(self.h, _f(np.asarray(pos, np.float32)), _f(fr), int(bool(init)), int(sub0),
                                       _f(k), _f(sup), _i(wr), _f(vp), _f(fo))
        return ok, dict(k=k.reshape(32, 2), sup=sup, win=int(wr[0]), round=int(wr[1]), vert_pdf=vp, frame=fo)

    def ptt_learned_substep(self, pos, frame, sub0=0):
        """One learned sub-step (the network set with set_learned):
        (status 1 learned win / 2 plain fallback / 0 none, winner k [2],
        frame after [9], counters [4])."""
        k = np.zeros(2, np.float32)
        fo = np.zeros(9, np.float32)
        cnt = (C.c_longlong * 4)()
        fr = np.ascontiguousarray(frame, np.float32).reshape(9)
        st = self.L.oracle_ptt_learned_substep(self.h, _f(np.asarray(pos, np.float32)), _f(fr), _f(fr[:3].copy()),
                                               int(sub0), _f(k), _f(fo), cnt)
        return st, k, fo, np.array(list(cnt), dtype=np.int64)

    def ptt_init_frame(self, seed, first_step, sub0=0):
        fo = np.zeros(18, np.float32)
        ok = self.L.oracle_ptt_init_frame(self.h, _f(np.asarray(seed, np.float32)),
                                          _f(np.asarray(first_step, np.float32)), int(sub0), _f(fo))
        return ok, fo

    def interp4(self, pos, frame):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        frame = np.ascontiguousarray(frame, dtype=np.float32)
        return self.L.oracle_interp4(self.h, _f(pos), _f(frame))


def philox_draws(seed, sub, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_philox_draws(seed, sub, n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def philox_mixed(seed, sub, kinds):
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    out = np.zeros(len(kinds), dtype=np.float32)
    lib().oracle_philox_mixed(seed, sub, len(kinds), _i(kinds), _f(out))
    return out


def ulog(x):
    return lib().oracle_log(float(x))


def usincos(x):
    s, c = C.c_float(), C.c_float()
    lib().oracle_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def mlp_forward(mlp_bf16, x):
    w = np.ascontiguousarray(mlp_bf16, dtype=np.uint16)
    x = np.ascontiguousarray(x, dtype=np.float32)
    o = np.zeros(24, dtype=np.float32)
    lib().oracle_mlp_forward(w.ctypes.data_as(C.POINTER(C.c_uint16)), _f(x), _f(o))
    return o


def rsqrt(x):
    return float(lib().oracle_rsqrt(float(x)))


def sortnet_table():
    L = lib()
    n = L.oracle_sortnet_layers()
    return [[L.oracle_sortnet_partner(i, j) for j in range(32)] for i in range(n)]


def warp_sort_dec(keys, vals):
    k = np.ascontiguousarray(keys, dtype=np.float32).copy()
    v = np.ascontiguousarray(vals, dtype=np.int32).copy()
    lib().oracle_warp_sort_dec(_f(k), _i(v))
    return k, v


def prob_select(cdf, sel):
    c = np.ascontiguousarray(cdf, np.float32)
    return lib().oracle_prob_select(_f(c), len(c), float(sel))


def prefix_sum(a):
    a = np.ascontiguousarray(a, dtype=np.float32).copy()
    lib().oracle_prefix_sum(_f(a), len(a))
    return a


def prepare_propagator(k1, k2, arclen):
    p = np.zeros(9, dtype=np.float32)
    lib().oracle_prepare_propagator(k1, k2, arclen, _f(p))
    return p


def propagate_frame(prop, frame):
    prop = np.ascontiguousarray(prop, dtype=np.float32)
    f = np.ascontiguousarray(frame, dtype=np.float32).copy()
    d = np.zeros(3, dtype=np.float32)
    lib().oracle_propagate_frame(_f(prop), _f(f), _f(d))
    return f, d


def num_threads():
    return lib().oracle_num_threads()


def set_num_threads(n: int):
    """OpenMP threads of later oracle calls (in this process)."""
    lib().oracle_set_num_threads(int(n))
