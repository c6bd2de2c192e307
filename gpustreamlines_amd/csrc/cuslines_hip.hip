// cuslines_hip.hip -- gfx950 kernels and C ABI of the per-seed tracking hot
// path: bootstrap (OPDT/CSA), probabilistic and PTT direction getters, the
// stepping loop and the two-phase seed protocol of dipy/GPUStreamlines.
//
// Layout in HBM (one tracker):
//   prob/boot data  float [X][Y][Z][T32]   channels padded to a multiple of
//                                          32 so each half-wave gather of 32
//                                          channels is one aligned 128-B line
//   PTT ODF         float [V][X][Y][Z]     vertex-major: the 8 texels of one
//                                          probe tap share z-adjacent lines
//   stop map        float [X][Y][Z]
//   sphere          float [V][3], edges packed (u | v << 16)
//   LUT             uint16 [R][R][R]       nearest sphere vertex
//
// Execution model: wave64, two seed slots per wave (one per 32-lane half).
// The 32 lanes of a slot hold the reference's 32 per-lane random streams and
// reproduce its reduction orders, so the results are bit-identical to the
// CPU oracle (oracle/cuslines_oracle.cpp) on the same inputs.
#include "device_math.h"

#include "../../include/cuslines_hip.h"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace cusl;

namespace {

constexpr double PMF_THRESHOLD_P = 0.05;               // _globals.py:4
constexpr float REAL_MAX = 3.40282346638528859812e+38f;
constexpr float REAL_MIN = -REAL_MAX;                  // globals.h:41-42
constexpr float K_SMALL = 0.0001f;                     // ptt.cuh:10
constexpr float PTT_NORM_EPS = 1e-8f;                  // ptt.cuh:12
constexpr float PTT_MIN_NORM2 = 1e-16f;                // (1/PTT_INV_NORM_EPS)^2, ptt.cuh:13
constexpr float CSA_C0 = 0.28209479177387814f;         // boot.cu:363
constexpr int DISC_VERT_CNT = 24;                      // disc.h:20
constexpr int DISC_FACE_CNT = 31;                      // disc.h:21
constexpr int SORT_LAYERS = 15;
constexpr int QUEUES = 8;  // persistent-queue counters, one per XCD
constexpr int VPROP_K = 24 * 9;                        // PTT table: curvatures after the 24 propagators
constexpr int VPROP_FLOATS = VPROP_K + 2 * 24;

// PTT disc tables (quality 2): numeric data of cuslines/cuda_c/disc.h:20-84
// (nibrary, BSD-3-Clause, Copyright (c) 2024 Dogu Baran Aydogan).
__constant__ float c_disc_vert[DISC_VERT_CNT * 2] = {
    -0.99680788f, -0.07983759f, -0.94276539f, 0.33345677f, -0.87928469f, -0.47629658f,
    -0.72856617f, 0.68497542f,  -0.60006556f, -0.79995082f, -0.54129995f, -0.02761342f,
    -0.39271207f, 0.37117272f,  -0.39217391f, 0.91989110f,  -0.36362884f, -0.40757367f,
    -0.22391316f, -0.97460910f, -0.00130022f, 0.53966106f,  0.00000000f,  0.00000000f,
    0.00973999f,  0.99995257f,  0.01606516f,  -0.54289908f, 0.21342395f,  -0.97695968f,
    0.38192071f,  -0.38666136f, 0.38897094f,  0.37442837f,  0.40696681f,  0.91344295f,
    0.54387161f,  -0.01477123f, 0.59119367f,  -0.80652963f, 0.73955688f,  0.67309406f,
    0.87601150f,  -0.48229022f, 0.94617928f,  0.32364298f,  0.99585368f,  -0.09096944f};
__constant__ int c_disc_face[DISC_FACE_CNT * 3] = {
    9, 8, 4,    11, 16, 10, 5, 8, 11,   5, 1, 0,    18, 16, 11, 11, 15, 18, 13, 8, 9,  11, 8, 13,
    13, 15, 11, 22, 18, 23, 22, 20, 16, 16, 18, 22, 16, 20, 17, 12, 10, 17, 17, 10, 16, 15, 19, 21,
    23, 18, 21, 21, 18, 15, 2, 4, 8,    2, 5, 0,    8, 5, 2,    7, 10, 12,  6, 7, 3,    10, 7, 6,
    3, 1, 6,    1, 5, 6,    11, 10, 6,  6, 5, 11,   14, 19, 15, 15, 13, 14, 14, 13, 9};

// Partner lane of every layer of the 32-key sort network, lane-major: lane g
// reads its 16 bytes (byte L = partner in layer L) with one load.  Filled at
// tracker creation from the generator below.
__constant__ uint4 c_sortnet4[32];

// Knuth's merge exchange (TAOCP 5.2.2 Algorithm M) for 32 keys with each
// comparator placed in the earliest layer its inputs allow: the network of
// the reference's warp sort tables (cuwsort.cuh:37-52).
void make_sortnet(unsigned char out[SORT_LAYERS][32], int *nlayers) {
    int ready[32] = {0};
    for (int L = 0; L < SORT_LAYERS; ++L)
        for (int l = 0; l < 32; ++l) out[L][l] = (unsigned char)l;
    int nl = 0;
    const int n = 32, t = 5;
    for (int p = 1 << (t - 1); p > 0; p >>= 1) {
        int q = 1 << (t - 1), r = 0, d = p;
        for (;;) {
            for (int i = 0; i < n - d; ++i) {
                if ((i & p) != r) continue;
                const int a = i, b = i + d;
                const int L = std::max(ready[a], ready[b]);
                out[L][a] = (unsigned char)b;
                out[L][b] = (unsigned char)a;
                ready[a] = ready[b] = L + 1;
                nl = std::max(nl, L + 1);
            }
            if (q != p) { d = q - p; q >>= 1; r = p; } else break;
        }
    }
    *nlayers = nl;
}

struct DevParams {
    int dimx, dimy, dimz, dimt, t32;
    int nverts, nv32, nedges, sphere_symm, max_sline_len, model;
    int hr, hr32, delta_nr;
    int hr16, dn16, nv16;  // batched bootstrap: hr, delta_nr, nverts rounded up to 16
    uint32_t seed_lo, seed_hi;
    float tc_threshold, rel_peak, cos_max_angle, cos_min_sep, min_signal;
    double step_d;
    // PTT
    int step_frac, probe_quality, lut_res, nx, ny, nz, log2x, log2y;
    float probe_step, max_curv, min_support_f, ptt_seed_thr_f;
    float ptt_sub_step;  // (float)(STEP_SIZE / STEP_FRAC)
    int lut_brick;               // PTT LUT stored as 4x4x4 bricks
    int ptt_nonneg;              // PTT ODF has no negative value: rejection rounds may stop early
    uint32_t rec_thresh;  // record a sub-step when its hash is below this
    int rec_cap;          // capacity of the pair buffer
    int slot_cap;         // bounded chunk: phase 2 skips seeds whose streamlines end past this slot (0: none)
    int rk16;             // batched bootstrap: rank of H rounded up to 16 when H.d = U (U^T d) (0: [H; R] product)
};

struct DevBufs {
    const float *data;
    const float *stop;
    const float *verts;
    const uint32_t *edges;
    const uint16_t *nbr;  // [nverts][8] neighbours, padded with the vertex itself (nullptr: edge path)
    const float *H, *R, *db, *dq, *smat;
    const float *hr_mat, *fit_mat, *odf_mat;  // batched bootstrap: [H;R], [dq|-db], smat, zero padded
    const float *ut_mat, *u_mat, *invlev;      // batched bootstrap, low-rank form: U^T, U, 1/sqrt(1 - H_ii)
    const int *b0;
    const int *dwi_idx;
    const uint32_t *lut;  // PTT LUT entry: vertex | (direction rank << 16)
    const float *blocks;  // PTT corner blocks (nullptr: gather 8 texels)
    const float *fprop;    // learned proposal: per disc vertex, propagator of 4 probe steps [24][9]
    const float *vprop;    // PTT: per disc vertex, propagator of one probe step [24][9], then (k1, k2) [24][2]
    const uint16_t *mlp;   // proposal network, bf16, lane-interleaved (IL_NPARAM)
    float *rec;            // training pairs [rec_cap][PAIR_FLOATS]
    int *rec_count;
    unsigned long long *stats;  // learned sub-steps, fallbacks, plain sub-steps
};

// -------------------------------------------------------------------------
// Sorting, scans and peak extraction (per half-wave)
// -------------------------------------------------------------------------

// warp_sort<32,32,WSORT_DIR_DEC> (cuwsort.cuh:104-168) on (key, value).
__device__ __forceinline__ void sort_dec(float &k, int &v) {
    const int g = lane32();
    const uint4 net = c_sortnet4[g];
#pragma unroll
    for (int L = 0; L < SORT_LAYERS; ++L) {
        const uint32_t w = (L < 4) ? net.x : (L < 8) ? net.y : (L < 12) ? net.z : net.w;
        const int src = (int)((w >> (8 * (L & 3))) & 0xffu);
        const float a = shfl32(k, src);
        const int b = shfl32(v, src);
        if (g >= src) {
            if (a < k) { k = a; v = b; }
        } else {
            if (a > k) { k = a; v = b; }
        }
    }
}

// prefix_sum_sh_d (utils.cu:81-120): blocks of 32, lane 0 adds the previous
// block's last value, Hillis-Steele inside the block.
__device__ void prefix_sum(float *a, int len) {
    const int l = lane32();
    float carry = 0.0f;
    for (int j = 0; j < len; j += 32) {
        float v = (j + l < len) ? a[j + l] : 0.0f;
        if (j != 0 && l == 0) v = v + carry;
#pragma unroll
        for (int i = 1; i < 32; i <<= 1) {
            const float o = shfl_up32(v, i);
            if (l >= i && j + l < len) v = v + o;
        }
        if (j + l < len) a[j + l] = v;
        carry = shfl32(v, 31);
    }
    WAVE_SYNC();
}

// peak_directions_d (tracking_helpers.cu:75-245).  odf, marks and cand are
// this slot's LDS arrays; the kept vertex indices end in cand[0..return).
// marks holds 2 bits per vertex, 16 vertices per word: bit 1 "has a strictly
// larger neighbour", bit 0 "has a strictly smaller one"; the reference's
// positive mark is exactly "bit 0 only".
__device__ __forceinline__ uint32_t peak_mark(const uint32_t *marks, int j) {
    return (marks[j >> 4] >> (2 * (j & 15))) & 3u;
}

__device__ int sort_dedup_peaks(const DevParams &P, const float *odf, uint16_t *cand, const float *sv, int n);

__device__ int peak_directions(const DevParams &P, const DevBufs &B, const float *odf, uint32_t *marks,
                               uint16_t *cand, const float *sv) {
    const int l = lane32();
    const int nv = P.nverts;
    float mn = REAL_MAX;
    for (int j = l; j < nv; j += 32) mn = (mn < odf[j]) ? mn : odf[j];
    mn = min32(mn);
    mn = (0.0f > mn) ? 0.0f : mn;
    if (B.nbr) {
        // Vertex-centred form of the edge marks below: a vertex is a peak when
        // some neighbour is strictly smaller and none strictly larger, i.e.
        // when the smallest neighbour is below it and the largest not above
        // it (fminf / fmaxf pass over a NaN neighbour as the comparisons do).
        // Bit r of pk: vertex l + 32 r (nverts <= 1024).  Groups of 4
        // vertices per lane: the 4 adjacency rows are loaded before the first
        // is used, so their latencies overlap.
        uint32_t pk = 0u;
        float m = REAL_MIN;
        for (int r0 = 0, j0 = l; j0 < nv; r0 += 4, j0 += 128) {
            uint4 qs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                qs[u] = reinterpret_cast<const uint4 *>(B.nbr)[min(j0 + 32 * u, nv - 1)];  // past the end: dropped
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + 32 * u;
                if (j >= nv) break;
                const uint32_t w[4] = {qs[u].x, qs[u].y, qs[u].z, qs[u].w};
                const float x = odf[j];
                float lo = odf[w[0] & 0xffffu], hi = lo;
#pragma unroll
                for (int k = 1; k < 8; ++k) {
                    const float y = odf[(w[k >> 1] >> (16 * (k & 1))) & 0xffffu];
                    lo = fminf(lo, y);
                    hi = fmaxf(hi, y);
                }
                if (lo < x && !(hi > x)) {
                    pk |= 1u << (r0 + u);
                    const float xm = x + (-mn);
                    m = (m > xm) ? m : xm;
                }
            }
        }
        m = max32(m);
        const float comp = (float)((double)P.rel_peak * (double)m);
        int n = 0;
        for (int r = 0, j0 = 0; j0 < nv; ++r, j0 += 32) {
            const int j = j0 + l;
            const bool keep = (j < nv) && ((pk >> r) & 1u) && ((odf[j] - mn) >= comp);
            const uint32_t msk = half_ballot(keep);
            if (keep) cand[n + __popc(msk & ((1u << l) - 1u))] = (uint16_t)j;
            n += __popc(msk);
        }
        WAVE_SYNC();
        return sort_dedup_peaks(P, odf, cand, sv, n);
    }
    for (int w = l; w < (nv + 15) >> 4; w += 32) marks[w] = 0u;
    WAVE_SYNC();
    for (int e = l; e < P.nedges; e += 32) {
        const uint32_t uv = B.edges[e];
        const int u = (int)(uv & 0xffffu), v = (int)(uv >> 16);
        const float a = odf[u], b = odf[v];
        if (a < b) {
            atomicOr(&marks[u >> 4], 2u << (2 * (u & 15)));
            atomicOr(&marks[v >> 4], 1u << (2 * (v & 15)));
        } else if (b < a) {
            atomicOr(&marks[v >> 4], 2u << (2 * (v & 15)));
            atomicOr(&marks[u >> 4], 1u << (2 * (u & 15)));
        }
    }
    WAVE_SYNC();
    float m = REAL_MIN;
    for (int j = l; j < nv; j += 32)
        if (peak_mark(marks, j) == 1u) { const float x = odf[j] + (-mn); m = (m > x) ? m : x; }
    m = max32(m);
    const float comp = (float)((double)P.rel_peak * (double)m);
    int n = 0;
    for (int j0 = 0; j0 < nv; j0 += 32) {
        const int j = j0 + l;
        const bool keep = (j < nv) && (peak_mark(marks, j) == 1u) && ((odf[j] - mn) >= comp);
        const uint32_t msk = half_ballot(keep);
        if (keep) cand[n + __popc(msk & ((1u << l) - 1u))] = (uint16_t)j;
        n += __popc(msk);
    }
    WAVE_SYNC();
    return sort_dedup_peaks(P, odf, cand, sv, n);
}

// The tail of peak_directions_d (tracking_helpers.cu:182-245) on the n
// candidates in cand: descending sort when n < 32, then the greedy
// minimum-separation filter on lane 0.
__device__ int sort_dedup_peaks(const DevParams &P, const float *odf, uint16_t *cand, const float *sv, int n) {
    const int l = lane32();
    if (n < 32) {
        float k = REAL_MIN;
        int v = 0;
        if (l < n) { v = cand[l]; k = odf[v]; }
        sort_dec(k, v);
        if (l < n) cand[l] = (uint16_t)v;
        WAVE_SYNC();
    }
    if (n == 0) return 0;
    if (n <= 32) {
        // remove_similar_vertices (tracking_helpers.cu:195-223) on 32 lanes:
        // lane i holds candidate i's direction and marks which earlier
        // candidates j lie within the minimum separation (the same dot
        // product and test as the greedy loop below); the greedy pass then
        // keeps candidate i iff none of its marked candidates was kept, and
        // the kept indices are compacted in order.  Same kept set as the
        // single-lane loop, without its chain of dependent vertex loads.
        int ci = 0;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        if (l < n) {
            ci = cand[l];
            ax = sv[3 * ci]; ay = sv[3 * ci + 1]; az = sv[3 * ci + 2];
        }
        uint32_t conf = 0u;
        for (int j = 0; j < n - 1; ++j) {
            const float bx = shfl32(ax, j), by = shfl32(ay, j), bz = shfl32(az, j);
            float c = dot3(ax, ay, az, bx, by, bz);
            if (P.sphere_symm) c = fabsf(c);
            if (j < l && c > P.cos_min_sep) conf |= 1u << j;
        }
        uint32_t keep = 1u;
        for (int i = 1; i < n; ++i) {
            const uint32_t ci_conf = (uint32_t)shfl32((int)conf, i);
            if ((ci_conf & keep) == 0u) keep |= 1u << i;
        }
        WAVE_SYNC();
        if (l < n && ((keep >> l) & 1u)) cand[__popc(keep & ((1u << l) - 1u))] = (uint16_t)ci;
        WAVE_SYNC();
        return __popc(keep);
    }
    int kept = 1;
    if (l == 0) {
        // remove_similar_vertices, single lane (tracking_helpers.cu:195-223),
        // compacting the kept indices in place.
        for (int i = 1; i < n; ++i) {
            const int ci = cand[i];
            const float ax = sv[3 * ci], ay = sv[3 * ci + 1], az = sv[3 * ci + 2];
            int j = 0;
            for (; j < kept; ++j) {
                const int cj = cand[j];
                float c = dot3(ax, ay, az, sv[3 * cj], sv[3 * cj + 1], sv[3 * cj + 2]);
                if (P.sphere_symm) c = fabsf(c);
                if (c > P.cos_min_sep) break;
            }
            if (j == kept) cand[kept++] = (uint16_t)ci;
        }
    }
    kept = shfl32(kept, 0);
    WAVE_SYNC();
    return kept;
}

// -------------------------------------------------------------------------
// Interpolation and the stop check
// -------------------------------------------------------------------------

// trilinear_interp_d (tracking_helpers.cu:3-61) over `dimt` channels with
// channel stride `ld`; writes out[0..dimt).  UNROLL channel iterations are
// in flight together (8 corner loads each): the probabilistic getter, whose
// step is this 23 KB gather, uses 3.
template <int UNROLL = 1>
__device__ int trilinear(const DevParams &P, const float *__restrict__ data, int ld, int dimt,
                         float px, float py, float pz, float *out) {
    if (px < -0.5f || px + 0.5f >= (float)P.dimx || py < -0.5f || py + 0.5f >= (float)P.dimy ||
        pz < -0.5f || pz + 0.5f >= (float)P.dimz)
        return -1;
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    const float wx1 = px - fx, wy1 = py - fy, wz1 = pz - fz;
    const float wx0 = 1.0f - wx1, wy0 = 1.0f - wy1, wz0 = 1.0f - wz1;
    const int x0 = (int)((0.0f > fx) ? 0.0f : fx), y0 = (int)((0.0f > fy) ? 0.0f : fy),
              z0 = (int)((0.0f > fz) ? 0.0f : fz);
    const int x1 = min(P.dimx - 1, x0 + 1), y1 = min(P.dimy - 1, y0 + 1), z1 = min(P.dimz - 1, z0 + 1);
    const float w[8] = {(wx0 * wy0) * wz0, (wx0 * wy0) * wz1, (wx0 * wy1) * wz0, (wx0 * wy1) * wz1,
                        (wx1 * wy0) * wz0, (wx1 * wy0) * wz1, (wx1 * wy1) * wz0, (wx1 * wy1) * wz1};
    const size_t DY = (size_t)P.dimy, DZ = (size_t)P.dimz;
    const float *c[8] = {
        data + (((size_t)x0 * DY + y0) * DZ + z0) * ld, data + (((size_t)x0 * DY + y0) * DZ + z1) * ld,
        data + (((size_t)x0 * DY + y1) * DZ + z0) * ld, data + (((size_t)x0 * DY + y1) * DZ + z1) * ld,
        data + (((size_t)x1 * DY + y0) * DZ + z0) * ld, data + (((size_t)x1 * DY + y0) * DZ + z1) * ld,
        data + (((size_t)x1 * DY + y1) * DZ + z0) * ld, data + (((size_t)x1 * DY + y1) * DZ + z1) * ld};
    const int l = lane32();
#pragma unroll UNROLL
    for (int t = l; t < dimt; t += 32) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = c[k][t];
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc = __builtin_fmaf(w[k], v[k], acc);
        out[t] = acc;
    }
    return 0;
}

// Texture coordinate -> (first texel, 8-bit weight of the second texel),
// after the CUDA guide's linear filtering: xB = x - 0.5, i = floor(xB),
// weight = frac(xB) with 8 fractional bits.
__device__ __forceinline__ float tex_weight(float c, int *i0) {
    const float b = c - 0.5f;
    const float fl = floorf(b);
    *i0 = (int)fl;
    return __builtin_rintf((b - fl) * 256.0f) * 0.00390625f;
}

// check_point_d (tracking_helpers.cu:63-73): linear filter of the stop map,
// border colour -1, texture x = array z (cu_tractography.py:237-241).
__device__ int check_point(const DevParams &P, const float *__restrict__ stop, float px, float py,
                           float pz) {
    int iz, iy, ix;
    const float a = tex_weight(pz, &iz), b = tex_weight(py, &iy), g = tex_weight(px, &ix);
    const float a0 = 1.0f - a, b0 = 1.0f - b, g0 = 1.0f - g;
    const int DX = P.dimx, DY = P.dimy, DZ = P.dimz;
    auto fetch = [&](int z, int y, int x) -> float {
        if (z < 0 || z >= DZ || y < 0 || y >= DY || x < 0 || x >= DX) return -1.0f;
        return stop[((size_t)x * DY + y) * DZ + z];
    };
    float acc = ((a0 * b0) * g0) * fetch(iz, iy, ix);
    acc = __builtin_fmaf((a * b0) * g0, fetch(iz + 1, iy, ix), acc);
    acc = __builtin_fmaf((a0 * b) * g0, fetch(iz, iy + 1, ix), acc);
    acc = __builtin_fmaf((a * b) * g0, fetch(iz + 1, iy + 1, ix), acc);
    acc = __builtin_fmaf((a0 * b0) * g, fetch(iz, iy, ix + 1), acc);
    acc = __builtin_fmaf((a * b0) * g, fetch(iz + 1, iy, ix + 1), acc);
    acc = __builtin_fmaf((a0 * b) * g, fetch(iz, iy + 1, ix + 1), acc);
    acc = __builtin_fmaf((a * b) * g, fetch(iz + 1, iy + 1, ix + 1), acc);
    if (acc == -1.0f) return CUSLINES_OUTSIDEIMAGE;
    return (acc > P.tc_threshold) ? CUSLINES_TRACKPOINT : CUSLINES_ENDPOINT;
}

// Texel index I of one axis of the vertex-tiled PTT texture -> (tile, local),
// clamp address mode; `hint` is the tile the query was built from.
__device__ __forceinline__ void axis_split(int I, int D, int ntiles, int hint, int *tile, int *loc) {
    I = min(max(I, 0), ntiles * D - 1);
    int t = hint, c = I - hint * D;
    while (c < 0) { --t; c += D; }
    while (c >= D) { ++t; c -= D; }
    *tile = t;
    *loc = c;
}

// Linear-filtered sample of the (nx,ny,nz)-tiled ODF texture at big-texture
// coordinates (xq,yq,zq) built from tile (tx,ty,tz); slots >= nverts read 0.
__device__ float tiled_sample(const DevParams &P, const float *__restrict__ odfv, float xq, float yq,
                              float zq, int tx, int ty, int tz) {
    int i, j, k;
    const float a = tex_weight(xq, &i), b = tex_weight(yq, &j), g = tex_weight(zq, &k);
    const float a0 = 1.0f - a, b0 = 1.0f - b, g0 = 1.0f - g;
    int txa, txb, lxa, lxb, tya, tyb, lya, lyb, tza, tzb, lza, lzb;
    axis_split(i, P.dimx, P.nx, tx, &txa, &lxa);
    axis_split(i + 1, P.dimx, P.nx, tx, &txb, &lxb);
    axis_split(j, P.dimy, P.ny, ty, &tya, &lya);
    axis_split(j + 1, P.dimy, P.ny, ty, &tyb, &lyb);
    axis_split(k, P.dimz, P.nz, tz, &tza, &lza);
    axis_split(k + 1, P.dimz, P.nz, tz, &tzb, &lzb);
    const int DX = P.dimx, DY = P.dimy, DZ = P.dimz, NV = P.nverts, nx = P.nx, ny = P.ny;
    auto fetch = [&](int ti, int li, int tj, int lj, int tk, int lk) -> float {
        const int v = ti + nx * (tj + ny * tk);
        if (v >= NV) return 0.0f;
        return odfv[(((size_t)v * DX + li) * DY + lj) * DZ + lk];
    };
    float acc = ((a0 * b0) * g0) * fetch(txa, lxa, tya, lya, tza, lza);
    acc = __builtin_fmaf((a * b0) * g0, fetch(txb, lxb, tya, lya, tza, lza), acc);
    acc = __builtin_fmaf((a0 * b) * g0, fetch(txa, lxa, tyb, lyb, tza, lza), acc);
    acc = __builtin_fmaf((a * b) * g0, fetch(txb, lxb, tyb, lyb, tza, lza), acc);
    acc = __builtin_fmaf((a0 * b0) * g, fetch(txa, lxa, tya, lya, tzb, lzb), acc);
    acc = __builtin_fmaf((a * b0) * g, fetch(txb, lxb, tya, lya, tzb, lzb), acc);
    acc = __builtin_fmaf((a0 * b) * g, fetch(txa, lxa, tyb, lyb, tzb, lzb), acc);
    acc = __builtin_fmaf((a * b) * g, fetch(txb, lxb, tyb, lyb, tzb, lzb), acc);
    return acc;
}

// interp4_d (ptt.cu:24-42) in two halves, so a caller can issue the LUT
// read of one tap before it consumes the previous tap's block.
// lut_index: linear index of the nearest-vertex LUT cell of the tangent
// (point sampling, normalised coordinates, clamp).  Index products use
// 24-bit multiplies (full rate; every operand is below 2^24: LUT coordinates,
// tile indices, cell coordinates).
__device__ __forceinline__ uint32_t lut_index(const DevParams &P, float f0, float f1, float f2) {
    const int res = P.lut_res;
    const float rf = (float)res, rmax = (float)(res - 1);
    const float u0 = __builtin_fmaf(f0, 0.5f, 0.5f), u1 = __builtin_fmaf(f1, 0.5f, 0.5f),
                u2 = __builtin_fmaf(f2, 0.5f, 0.5f);
    const uint32_t i0 = (uint32_t)fminf(fmaxf(floorf(u0 * rf), 0.0f), rmax);
    const uint32_t i1 = (uint32_t)fminf(fmaxf(floorf(u1 * rf), 0.0f), rmax);
    const uint32_t i2 = (uint32_t)fminf(fmaxf(floorf(u2 * rf), 0.0f), rmax);
    if (P.lut_brick) {
        const uint32_t nb = (uint32_t)res >> 2;
        return (__umul24(__umul24(i0 >> 2, nb) + (i1 >> 2), nb) + (i2 >> 2)) * 64u + ((i0 & 3u) << 4) +
               ((i1 & 3u) << 2) + (i2 & 3u);
    }
    return __umul24(__umul24(i0, (uint32_t)res) + i1, (uint32_t)res) + i2;
}

// The linear-filtered sample of a vertex's tile of the ODF texture at
// (px, py, pz); `lv` is the vertex's LUT entry (vertex | rank << 16).  One
// 32-byte corner block when the cell is inside the volume (its 8 texels, the
// tile's clamp already applied), else the 8-texel gather.  Blocks are stored
// cell-major, the vertices of a cell in the rank order, so one 128-byte line
// holds a cell's blocks for four neighbouring directions.
//
// In two halves, so that a probe loop can issue one tap's block read, do
// other work and blend afterwards: tap_begin computes the weights and issues
// the read (fast = the cell has a corner block), tap_end blends (or gathers
// the 8 texels when there is no block).
struct Tap {
    float a, b, g;
    float4 lo, hi;
    float xq, yq, zq;
    uint32_t lv;
    bool fast;
};

__device__ __forceinline__ Tap tap_begin(const DevParams &P, const DevBufs &B, uint32_t lv, float px, float py,
                                         float pz) {
    Tap t;
    t.lv = lv;
    const int idx = (int)(lv & 0xffffu);
    const int tx = idx & (P.nx - 1);
    const int ty = (idx >> P.log2x) & (P.ny - 1);
    const int tz = idx >> (P.log2x + P.log2y);
    const int ox = (int)__umul24((uint32_t)tx, (uint32_t)P.dimx);
    const int oy = (int)__umul24((uint32_t)ty, (uint32_t)P.dimy);
    const int oz = (int)__umul24((uint32_t)tz, (uint32_t)P.dimz);
    t.xq = (float)ox + px;
    t.yq = (float)oy + py;
    t.zq = (float)oz + pz;
    t.fast = false;
    if (B.blocks) {
        int i, j, k;
        t.a = tex_weight(t.xq, &i);
        t.b = tex_weight(t.yq, &j);
        t.g = tex_weight(t.zq, &k);
        const int cx = i - ox, cy = j - oy, cz = k - oz;
        if ((unsigned)(cx + 1) <= (unsigned)P.dimx && (unsigned)(cy + 1) <= (unsigned)P.dimy &&
            (unsigned)(cz + 1) <= (unsigned)P.dimz) {
            const uint32_t inner =
                __umul24(__umul24((uint32_t)(cx + 1), (uint32_t)(P.dimy + 1)) + (uint32_t)(cy + 1),
                         (uint32_t)(P.dimz + 1)) + (uint32_t)(cz + 1);
            const uint64_t cell = (uint64_t)inner * (uint32_t)P.nverts + (lv >> 16);
            const float4 *blk = reinterpret_cast<const float4 *>(B.blocks + cell * 8);
            t.lo = blk[0];
            t.hi = blk[1];
            t.fast = true;
        }
    }
    return t;
}

// Linear filter of a corner block (texel order x fastest, then y, z).
__device__ __forceinline__ float corner_blend(float a, float b, float g, const float4 &lo, const float4 &hi) {
    const float a0 = 1.0f - a, b0 = 1.0f - b, g0 = 1.0f - g;
    float acc = ((a0 * b0) * g0) * lo.x;
    acc = __builtin_fmaf((a * b0) * g0, lo.y, acc);
    acc = __builtin_fmaf((a0 * b) * g0, lo.z, acc);
    acc = __builtin_fmaf((a * b) * g0, lo.w, acc);
    acc = __builtin_fmaf((a0 * b0) * g, hi.x, acc);
    acc = __builtin_fmaf((a * b0) * g, hi.y, acc);
    acc = __builtin_fmaf((a0 * b) * g, hi.z, acc);
    acc = __builtin_fmaf((a * b) * g, hi.w, acc);
    return acc;
}

__device__ __forceinline__ float tap_end(const DevParams &P, const DevBufs &B, const Tap &t) {
    if (t.fast) return corner_blend(t.a, t.b, t.g, t.lo, t.hi);
    const int idx = (int)(t.lv & 0xffffu);
    return tiled_sample(P, B.data, t.xq, t.yq, t.zq, idx & (P.nx - 1), (idx >> P.log2x) & (P.ny - 1),
                        idx >> (P.log2x + P.log2y));
}

__device__ __forceinline__ float tap_sample(const DevParams &P, const DevBufs &B, uint32_t lv, float px, float py,
                                            float pz) {
    return tap_end(P, B, tap_begin(P, B, lv, px, py, pz));
}

// A tap for loops that hold several in flight: tap_begin's weights and
// block read, with a cell that has no block gathered at once (its value in
// lo.x), so a tap in flight holds no more than its weights and block.
struct TapB {
    float a, b, g;
    float4 lo, hi;
    bool fast;
};

__device__ __forceinline__ TapB tapb_begin(const DevParams &P, const DevBufs &B, uint32_t lv, float px, float py,
                                           float pz) {
    const Tap t = tap_begin(P, B, lv, px, py, pz);
    TapB r;
    r.a = t.a; r.b = t.b; r.g = t.g;
    r.lo = t.lo; r.hi = t.hi;
    r.fast = t.fast;
    if (!t.fast) r.lo.x = tap_end(P, B, t);
    return r;
}

__device__ __forceinline__ float tapb_end(const TapB &t) {
    return t.fast ? corner_blend(t.a, t.b, t.g, t.lo, t.hi) : t.lo.x;
}

__device__ __forceinline__ float interp4(const DevParams &P, const DevBufs &B, float px, float py,
                                         float pz, float f0, float f1, float f2) {
    return tap_sample(P, B, B.lut[lut_index(P, f0, f1, f2)], px, py, pz);
}


// -------------------------------------------------------------------------
// Half-wave matrix-vector products with the reference's per-row summation
// (ndotp_d, boot.cu:195-235): per row, 32 lane partials over j = l, l+32, ...
// (fma, in order), then the 16/8/4/2/1 pair tree.
// XF selects the elementwise transform of the vector: 0 none, 1 OPDT, 2 CSA.
// -------------------------------------------------------------------------
template <int XF>
__device__ __forceinline__ float gemv_xf(float x) {
    if (XF == 1) {
        const float lx = ulog(x);
        return ((-lx) * (1.5f + lx)) * x;
    } else if (XF == 2) {
        const float c = fminf(fmaxf(x, 0.001f), 0.999f);
        return ulog(-ulog(c));
    }
    return x;
}

// One row per lane: lane r of the slot sums row i0 + r itself, keeping the
// 32 partials in registers and adding them with the pair tree, so a row
// costs M fma + 31 adds (a shuffle reduction across lanes costs ~4 VALU per
// row-lane pair more).  When M % 4 == 0 the matrix is stored lane-interleaved
// (mat_offset: the 4-column pieces of 32 consecutive rows side by side), so
// the 16-byte row loads of one instruction are 512 contiguous bytes; else
// row-major.  x comes from LDS as broadcast reads and must be 16-byte
// aligned, already transformed (gemv_xf).
__host__ __device__ inline size_t mat_offset(int M, int i, int j) {
    return (((size_t)(i >> 5) * (size_t)(M >> 2) + (size_t)(j >> 2)) << 7) + (size_t)((i & 31) << 2) + (size_t)(j & 3);
}

__device__ void gemv_lanerow(int N, int M, const float *x, const float *__restrict__ mat, float *out) {
    const int l = lane32();
    const bool vec = (M & 3) == 0;
    for (int i0 = 0; i0 < N; i0 += 32) {
        const int i = i0 + l;
        if (i < N) {
            const float *row = vec ? mat + mat_offset(M, i, 0) : mat + (size_t)i * M;
            float p[32];
#pragma unroll
            for (int q = 0; q < 32; ++q) p[q] = 0.0f;
            for (int j0 = 0; j0 < M; j0 += 32) {
                // All loads of the block are issued before the first use: a
                // column past M re-reads the block's last valid one, and its
                // product is dropped by the select, so no load sits behind a
                // branch (and none leaves the matrix).
                const int nj = min(32, M - j0);
                if (vec) {  // M % 4 == 0: a vector is all in or all out
                    const int nq = nj >> 2;
                    float4 m[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        m[q] = *reinterpret_cast<const float4 *>(row + ((size_t)(j0 + 4 * min(q, nq - 1)) << 5));
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const float4 v = *reinterpret_cast<const float4 *>(x + j0 + 4 * q);
                        const bool in = q < nq;
                        const float a0 = __builtin_fmaf(v.x, m[q].x, p[4 * q]);
                        const float a1 = __builtin_fmaf(v.y, m[q].y, p[4 * q + 1]);
                        const float a2 = __builtin_fmaf(v.z, m[q].z, p[4 * q + 2]);
                        const float a3 = __builtin_fmaf(v.w, m[q].w, p[4 * q + 3]);
                        p[4 * q] = in ? a0 : p[4 * q];
                        p[4 * q + 1] = in ? a1 : p[4 * q + 1];
                        p[4 * q + 2] = in ? a2 : p[4 * q + 2];
                        p[4 * q + 3] = in ? a3 : p[4 * q + 3];
                    }
                } else {
                    float m[32];
#pragma unroll
                    for (int q = 0; q < 32; ++q) m[q] = row[j0 + min(q, nj - 1)];
#pragma unroll
                    for (int q = 0; q < 32; ++q) {
                        const float a = __builtin_fmaf(x[j0 + q], m[q], p[q]);
                        p[q] = (q < nj) ? a : p[q];
                    }
                }
            }
#pragma unroll
            for (int d = 16; d >= 1; d >>= 1)
#pragma unroll
                for (int q = 0; q < d; ++q) p[q] = p[q] + p[q + d];
            out[i] = p[0];
        }
    }
    WAVE_SYNC();
}

// -------------------------------------------------------------------------
// Probabilistic getter (generate_streamlines_cuda.cu:49-275)
// -------------------------------------------------------------------------
struct ProbLds {
    float *pmf;
    uint32_t *marks;
    uint16_t *cand;
};

template <bool IS_START>
__device__ int get_direction_prob(const DevParams &P, const DevBufs &B, const float *sv, ProbLds &S,
                                  Philox &rng, float dx, float dy, float dz, float px, float py,
                                  float pz, float *ndx, float *ndy, float *ndz) {
    const int l = lane32();
    const int nv = P.nverts;
    float *pmf = S.pmf;
    if (trilinear<3>(P, B.data, P.t32, P.dimt, px, py, pz, pmf) != 0) return 0;
    WAVE_SYNC();
    float mx = REAL_MIN;
    for (int i = l; i < nv; i += 32) mx = (mx > pmf[i]) ? mx : pmf[i];
    mx = max32(mx);
    const float thr = (float)(PMF_THRESHOLD_P * (double)mx);
    if (IS_START) {
        for (int i = l; i < nv; i += 32)
            if (pmf[i] < thr) pmf[i] = 0.0f;
        WAVE_SYNC();
        return peak_directions(P, B, pmf, S.marks, S.cand, sv);
    }
    for (int i = l; i < nv; i += 32) {
        float v = pmf[i];
        if (v < thr) v = 0.0f;
        float d = dot3(dx, dy, dz, sv[3 * i], sv[3 * i + 1], sv[3 * i + 2]);
        if (P.sphere_symm) d = fabsf(d);
        if (d < P.cos_max_angle) v = 0.0f;
        pmf[i] = v;
    }
    WAVE_SYNC();
    prefix_sum(pmf, nv);
    const float last = pmf[nv - 1];
    if (last == 0.0f) return 0;
    float sel = 0.0f;
    if (l == 0) sel = rng.uniform() * last;
    sel = shfl32(sel, 0);
    int low = 0, high = nv - 1;
    while ((high - low) >= 32) {
        const int mid = (low + high) / 2;
        if (pmf[mid] < sel) low = mid; else high = mid;
    }
    const bool pr = (low + l <= high) ? (sel < pmf[low + l]) : false;
    const uint32_t msk = half_ballot(pr);
    int ind = low + __ffs(msk) - 1;
    if (ind < 0) ind = 0;
    const float vx = sv[3 * ind], vy = sv[3 * ind + 1], vz = sv[3 * ind + 2];
    if (dot3(dx, dy, dz, vx, vy, vz) > 0.0f) { *ndx = vx; *ndy = vy; *ndz = vz; }
    else { *ndx = -vx; *ndy = -vy; *ndz = -vz; }
    WAVE_SYNC();
    return 1;
}

// -------------------------------------------------------------------------
// Bootstrap getter (boot.cu:19-113, 115-193, 330-608)
// -------------------------------------------------------------------------
struct BootLds {
    float *h;          // region C: H.d, then the ODF
    float *vox, *msk;  // region A during the refit: vox [t32], msk [hr32]
    float *r;          // region A: R.d, then the model coefficients
    uint32_t *marks;   // region A once the ODF is computed: peak marks, 2 bits per vertex
    uint16_t *cand;    // region A after the marks: peak candidates
};

// closest_peak_d (boot.cu:115-193): lane-strided best |dot| with strict
// improvement, then an xor tree with the same rule; lane 0's winner counts.
__device__ int closest_peak(const DevParams &P, const float *sv, float dx, float dy, float dz, int npeaks,
                            const uint16_t *peaks, float *ox, float *oy, float *oz) {
    const int l = lane32();
    float cd = 0.0f;
    int ci = -1;
    for (int j = l; j < npeaks; j += 32) {
        const int v = peaks[j];
        const float d = dot3(dx, dy, dz, sv[3 * v], sv[3 * v + 1], sv[3 * v + 2]);
        if (fabsf(d) > fabsf(cd)) { cd = d; ci = j; }
    }
    for (int m = 16; m; m >>= 1) {
        const float od = shfl_xor32(cd, m);
        const int oi = shfl_xor32(ci, m);
        if (fabsf(od) > fabsf(cd)) { cd = od; ci = oi; }
    }
    cd = shfl32(cd, 0);
    ci = shfl32(ci, 0);
    if (ci >= 0) {
        const int v = peaks[ci];
        if (cd >= P.cos_max_angle) { *ox = sv[3 * v]; *oy = sv[3 * v + 1]; *oz = sv[3 * v + 2]; return 1; }
        if (cd <= -P.cos_max_angle) { *ox = -sv[3 * v]; *oy = -sv[3 * v + 1]; *oz = -sv[3 * v + 2]; return 1; }
    }
    return 0;
}

template <int MODEL>
__device__ int get_direction_boot(const DevParams &P, const DevBufs &B, const float *sv, BootLds &S,
                                  Philox &rng, int nattempts, float dx, float dy, float dz, float px,
                                  float py, float pz, float *ndx, float *ndy, float *ndz) {
    const int l = lane32();
    const int hr = P.hr, dn = P.delta_nr, nv = P.nverts, T = P.dimt;
    for (int att = 0; att < nattempts; ++att) {
        const int rv = trilinear(P, B.data, P.t32, T, px, py, pz, S.vox);
        WAVE_SYNC();
        if (rv == 0) {
            for (int m = l; m < hr; m += 32) S.msk[m] = S.vox[B.dwi_idx[m]];
            WAVE_SYNC();
            gemv_lanerow(hr, hr, S.msk, B.R, S.r);
            gemv_lanerow(hr, hr, S.msk, B.H, S.h);
            for (int j = l; j < hr; j += 32) {
                const int src = (int)(rng.next() % (uint32_t)hr);
                S.h[j] = S.h[j] + S.r[src];
            }
            WAVE_SYNC();
            for (int m = l; m < hr; m += 32) S.vox[B.dwi_idx[m]] = S.h[m];
            WAVE_SYNC();
            float part = 0.0f;
            int cnt = 0;
            for (int j = l; j < T; j += 32) {
                float v = S.vox[j];
                v = (P.min_signal > v) ? P.min_signal : v;
                S.vox[j] = v;
                if (B.b0[j]) { ++cnt; part = part + v; }
            }
            const float sum = tree_sum32(part);
            for (int m = 16; m; m >>= 1) cnt += shfl_xor32(cnt, m);
            const float denom = sum / (float)cnt;
            for (int j = l; j < T; j += 32) S.vox[j] = S.vox[j] / denom;
            WAVE_SYNC();
            for (int m = l; m < hr; m += 32) S.msk[m] = S.vox[B.dwi_idx[m]];
            WAVE_SYNC();
            // the model's transform of the signal, staged in h (free here)
            for (int j = l; j < hr; j += 32) S.h[j] = gemv_xf<(MODEL == CUSLINES_OPDT) ? 1 : 2>(S.msk[j]);
            WAVE_SYNC();
            gemv_lanerow(dn, hr, S.h, B.dq, S.r);
            if (MODEL == CUSLINES_OPDT) {
                gemv_lanerow(dn, hr, S.msk, B.db, S.h);
                for (int j = l; j < dn; j += 32) S.r[j] = S.r[j] - S.h[j];
            } else {
                if (l == 0) S.r[0] = CSA_C0;
            }
            WAVE_SYNC();
            gemv_lanerow(nv, dn, S.r, B.smat, S.h);
        } else {
            for (int j = l; j < nv; j += 32) S.h[j] = 0.0f;
            WAVE_SYNC();
        }
        float mx = REAL_MIN;
        for (int j = l; j < nv; j += 32) mx = (mx > S.h[j]) ? mx : S.h[j];
        mx = max32(mx);
        const float thr = (float)(PMF_THRESHOLD_P * (double)mx);
        for (int j = l; j < nv; j += 32)
            if (S.h[j] < thr) S.h[j] = 0.0f;
        WAVE_SYNC();
        const int ndir = peak_directions(P, B, S.h, S.marks, S.cand, sv);
        if (nattempts == 1) return ndir;
        if (ndir > 0 && closest_peak(P, sv, dx, dy, dz, ndir, S.cand, ndx, ndy, ndz)) {
            WAVE_SYNC();
            return 1;
        }
        WAVE_SYNC();
    }
    return 0;
}

// -------------------------------------------------------------------------
// PTT getter (ptt.cu)
// -------------------------------------------------------------------------
// norm3_d (ptt.cu:1-13): scale to unit length when 1/|v| is finite, positive
// and below 1e8 (|v|^2 > 1e-16), else the FAIL axis.
template <int FAIL>
__device__ __forceinline__ void norm3(float &x, float &y, float &z) {
    const float n2 = __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x));
    if (n2 > PTT_MIN_NORM2 && n2 <= 3.40282347e+38f) {
        const float inv = rsqrt_nr(n2);
        x *= inv; y *= inv; z *= inv;
    } else {
        x = (FAIL == 0) ? 1.0f : 0.0f;
        y = (FAIL == 1) ? 1.0f : 0.0f;
        z = (FAIL == 2) ? 1.0f : 0.0f;
    }
}

template <int FAIL>
__device__ __forceinline__ void crossnorm3(float *d, const float *a, const float *b) {
    d[0] = __builtin_fmaf(a[1], b[2], -(a[2] * b[1]));
    d[1] = __builtin_fmaf(a[2], b[0], -(a[0] * b[2]));
    d[2] = __builtin_fmaf(a[0], b[1], -(a[1] * b[0]));
    norm3<FAIL>(d[0], d[1], d[2]);
}

__device__ __forceinline__ void prepare_propagator(float k1, float k2, float arclen, float *p) {
    if (fabsf(k1) < K_SMALL && fabsf(k2) < K_SMALL) {
        p[0] = arclen; p[1] = 0.0f; p[2] = 0.0f; p[3] = 1.0f; p[4] = 0.0f;
        p[5] = 0.0f; p[6] = 0.0f; p[7] = 0.0f; p[8] = 1.0f;
        return;
    }
    if (fabsf(k1) < K_SMALL) k1 = K_SMALL;
    if (fabsf(k2) < K_SMALL) k2 = K_SMALL;
    const float k = __builtin_sqrtf(__builtin_fmaf(k2, k2, k1 * k1));
    float sinkt, coskt;
    usincos(k * arclen, &sinkt, &coskt);
    const float kk = 1.0f / (k * k);
    p[0] = sinkt / k;
    p[1] = (k1 * (1.0f - coskt)) * kk;
    p[2] = (k2 * (1.0f - coskt)) * kk;
    p[3] = coskt;
    p[4] = (k1 * sinkt) / k;
    p[5] = (k2 * sinkt) / k;
    p[6] = -p[5];
    p[7] = ((k1 * k2) * (coskt - 1.0f)) * kk;
    p[8] = __builtin_fmaf(k2 * k2, coskt, k1 * k1) * kk;
}

__device__ __forceinline__ void propagate_frame(const float *p, float *f, float *direc) {
    float tmp[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        direc[i] = __builtin_fmaf(p[2], f[6 + i], __builtin_fmaf(p[1], f[3 + i], p[0] * f[i]));
        tmp[i] = __builtin_fmaf(p[5], f[6 + i], __builtin_fmaf(p[4], f[3 + i], p[3] * f[i]));
        f[6 + i] = __builtin_fmaf(p[8], f[6 + i], __builtin_fmaf(p[7], f[3 + i], p[6] * f[i]));
    }
    norm3<0>(tmp[0], tmp[1], tmp[2]);
    crossnorm3<1>(f + 3, f + 6, tmp);
    crossnorm3<2>(f + 6, tmp, f + 3);
    f[0] = tmp[0]; f[1] = tmp[1]; f[2] = tmp[2];
}

template <bool INIT>
__device__ __forceinline__ void probing_frame(const float *frame, Philox &rng, float *pf) {
    if (INIT) {
        pf[0] = frame[0]; pf[1] = frame[1]; pf[2] = frame[2];
        norm3<0>(pf[0], pf[1], pf[2]);
        pf[3] = rng.normal();
        pf[4] = rng.normal();
        pf[5] = rng.normal();
        const float dt = __builtin_fmaf(pf[5], pf[2], __builtin_fmaf(pf[4], pf[1], pf[3] * pf[0]));
        pf[3] = __builtin_fmaf(-dt, pf[0], pf[3]);
        pf[4] = __builtin_fmaf(-dt, pf[1], pf[4]);
        pf[5] = __builtin_fmaf(-dt, pf[2], pf[5]);
        const float n2 = __builtin_fmaf(pf[5], pf[5], __builtin_fmaf(pf[4], pf[4], pf[3] * pf[3]));
        if (n2 < PTT_NORM_EPS) {
            const float ax = fabsf(pf[0]), ay = fabsf(pf[1]), az = fabsf(pf[2]);
            if (ax <= ay && ax <= az) { pf[3] = 0.0f; pf[4] = pf[2]; pf[5] = -pf[1]; }
            else if (ay <= az) { pf[3] = -pf[2]; pf[4] = 0.0f; pf[5] = pf[0]; }
            else { pf[3] = pf[1]; pf[4] = -pf[0]; pf[5] = 0.0f; }
        }
        norm3<1>(pf[3], pf[4], pf[5]);
        crossnorm3<2>(pf + 6, pf, pf + 3);
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) pf[i] = frame[i];
    }
}

// calculate_data_support_d (ptt.cu:156-192)
// The frames of successive probes do not depend on the sampled values, so
// probe i's block read is issued, then probe i+1's frame is propagated and
// its LUT read issued, and only then is probe i's block blended: each load's
// latency overlaps the other probe's work.
// With ALLOW_WEAK_LINK (ptt.cuh:8) the support is first_val plus the taps,
// all >= 0 (the ODF is clipped at 0 and linear filtering keeps the sign),
// so the running sum never decreases: once it reaches `stop_at` the
// candidate passes whatever the remaining taps add, and the lane stops
// probing.  Callers that need the value itself pass +infinity.  `ntap`
// (instrumented kernels only; nullptr otherwise) counts the taps read.
__device__ __forceinline__ float data_support(const DevParams &P, const DevBufs &B, float support,
                                              float px, float py, float pz, float k1, float k2, float *pf,
                                              float stop_at, uint32_t *ntap = nullptr,
                                              const float *prop_pre = nullptr) {
    // already at the threshold: no tap can change the outcome, and the
    // probe frames are dropped by every caller that passes a finite stop_at
    if (support >= stop_at) return support;
    float prop[9], direc[3];
    if (prop_pre) {
#pragma unroll
        for (int i = 0; i < 9; ++i) prop[i] = prop_pre[i];
    } else {
        prepare_propagator(k1, k2, P.probe_step, prop);
    }
    propagate_frame(prop, pf, direc);
    px += direc[0]; py += direc[1]; pz += direc[2];
    uint32_t idx = B.lut[lut_index(P, pf[0], pf[1], pf[2])];
    for (int i = 0; i < P.probe_quality; ++i) {
        if (support >= stop_at) break;
        if (ntap) ++*ntap;
        const Tap t = tap_begin(P, B, idx, px, py, pz);
        if (i + 1 < P.probe_quality) {
            propagate_frame(prop, pf, direc);
            px += direc[0]; py += direc[1]; pz += direc[2];
            idx = B.lut[lut_index(P, pf[0], pf[1], pf[2])];
        }
        support += tap_end(P, B, t);
    }
    return support;
}

struct PttLds {
    float *frame;     // 18 floats: [0..8] backward frame, [9..17] forward frame
    float *vert_pdf;  // 32
    float *face_cdf;  // 32
    float *act;       // learned proposal: inputs [64], hidden 1 [128], hidden 2 [128], outputs [32]
};

// Face CDF over the 31 disc faces from S.vert_pdf (prefix_sum_sh_d, one
// block of 32); returns the total.
__device__ __forceinline__ float ptt_face_cdf(PttLds &S) {
    const int l = lane32();
    float face = 0.0f;
    if (l < DISC_FACE_CNT) {
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const float vv = S.vert_pdf[c_disc_face[l * 3 + jj]];
            if (vv == 0.0f) { face = 0.0f; break; }
            face = face + vv;
        }
    }
#pragma unroll
    for (int i = 1; i < 32; i <<= 1) {
        const float o = shfl_up32(face, i);
        if (l >= i && l < DISC_FACE_CNT) face = face + o;
    }
    if (l < DISC_FACE_CNT) S.face_cdf[l] = face;
    const float last = shfl32(face, DISC_FACE_CNT - 1);
    WAVE_SYNC();
    return last;
}

// One PTT sub-step (ptt.cu:194-412) as a single loop over phases: phase 0
// measures the 24 vertex supports (lanes < 24) into S.vert_pdf, phases
// 1..rounds are rejection rounds of 32 candidates drawn from the face CDF.
// Every phase ends in the same data_support call, so the two seeds of a
// wavefront probe together whichever phase each one is in.  first_phase 1
// starts from supports already in S.vert_pdf.  Returns 1 with the new
// direction, 0 when no vertex passes, the CDF is empty or no round has a
// winner.  `wk` (optional) receives the winner's curvature, `raw` (optional,
// LDS [24]) the vertex supports before the 0.4 cut.
template <bool INIT, bool CNT = false>
__device__ int ptt_substep(const DevParams &P, const DevBufs &B, PttLds &S, float *frame, const float *fr,
                           Philox &rng, float first_val, int first_phase, int rounds, float dx, float dy, float dz,
                           float px, float py, float pz, float *ndx, float *ndy, float *ndz, float *wk, float *raw,
                           uint32_t *ntap = nullptr) {
    const int l = lane32();
    const double mc = (double)P.max_curv;
    float last = 0.0f;
    for (int ph = first_phase; ph <= rounds; ++ph) {
        if (ph == 1) {
            last = ptt_face_cdf(S);
            if (last == 0.0f) return 0;
        }
        float k1 = 0.0f, k2 = 0.0f;
        bool active = true;
        if (ph == 0) {
            active = l < DISC_VERT_CNT;
            if (active) {
                if (B.vprop) {
                    k1 = B.vprop[VPROP_K + 2 * l];
                    k2 = B.vprop[VPROP_K + 2 * l + 1];
                } else {
                    k1 = (float)((double)c_disc_vert[2 * l] * mc);
                    k2 = (float)((double)c_disc_vert[2 * l + 1] * mc);
                }
            }
        } else {
            float r1 = rng.uniform(), r2 = rng.uniform();
            if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
            const float tmp = rng.uniform() * last;
            // first face jj < 30 with cdf >= tmp, else the last (the
            // reference's linear scan): the cdf is a prefix sum of values
            // >= 0, so it never decreases and a binary search finds the same
            // face in 5 reads (a NaN tmp fails every test in both)
            int jj = 0, hi = DISC_FACE_CNT - 1;
            while (jj < hi) {
                const int mid = (jj + hi) >> 1;
                if (S.face_cdf[mid] >= tmp) hi = mid; else jj = mid + 1;
            }
            const int f0 = c_disc_face[jj * 3], f1 = c_disc_face[jj * 3 + 1], f2 = c_disc_face[jj * 3 + 2];
            float vx0, vx1, vx2, vy0, vy1, vy2;
            if (B.vprop) {  // the same products, computed once per tracker
                const float *vk = B.vprop + VPROP_K;
                vx0 = vk[2 * f0]; vx1 = vk[2 * f1]; vx2 = vk[2 * f2];
                vy0 = vk[2 * f0 + 1]; vy1 = vk[2 * f1 + 1]; vy2 = vk[2 * f2 + 1];
            } else {
                vx0 = (float)(mc * (double)c_disc_vert[f0 * 2]);
                vx1 = (float)(mc * (double)c_disc_vert[f1 * 2]);
                vx2 = (float)(mc * (double)c_disc_vert[f2 * 2]);
                vy0 = (float)(mc * (double)c_disc_vert[f0 * 2 + 1]);
                vy1 = (float)(mc * (double)c_disc_vert[f1 * 2 + 1]);
                vy2 = (float)(mc * (double)c_disc_vert[f2 * 2 + 1]);
            }
            k1 = __builtin_fmaf(r2, vx2 - vx0, __builtin_fmaf(r1, vx1 - vx0, vx0));
            k2 = __builtin_fmaf(r2, vy2 - vy0, __builtin_fmaf(r1, vy1 - vy0, vy0));
        }
        float pf[9];
        float sup = 0.0f;
        if (active) {
            probing_frame<INIT>(fr, rng, pf);
            // rejection rounds of a non-initial sub-step only test sup >= min_support_f
            const float stop_at = (ph >= 1 && !INIT && P.ptt_nonneg) ? P.min_support_f : __builtin_inff();
            // the vertex phase's curves are the disc vertices: their one-step
            // propagators are the same at every sub-step (computed once by
            // vertex_propagators with the same prepare_propagator)
            sup = data_support(P, B, first_val, px, py, pz, k1, k2, pf, stop_at, CNT ? ntap : nullptr,
                               (ph == 0 && B.vprop) ? B.vprop + 9 * l : nullptr);
        }
        if (ph == 0) {
            const bool ok = active && !(sup < P.min_support_f);
            if (active) {
                S.vert_pdf[l] = ok ? sup : 0.0f;
                if (raw) raw[l] = sup;
            }
            if (half_ballot(ok) == 0u) return 0;
            WAVE_SYNC();
            continue;
        }
        uint32_t msk;
        if (INIT) {
            float mx = sup;
            for (int m = 16; m; m >>= 1) mx = fmaxf(mx, shfl_xor32(mx, m));
            msk = half_ballot(sup == mx && mx >= P.min_support_f);
        } else {
            // >= as ptt.cu:384: a NaN support never wins a round (the vertex
            // phase above keeps the reference's `<` test, which lets it in)
            msk = half_ballot(sup >= P.min_support_f);
        }
        if (msk != 0u) {
            const int win = __ffs(msk) - 1;
            float ox = dx, oy = dy, oz = dz;
            if (l == win) {
                if (!INIT) {
                    float prop[9], d[3];
                    prepare_propagator(k1, k2, P.ptt_sub_step, prop);
#pragma unroll
                    for (int i = 0; i < 9; ++i) pf[i] = fr[i];
                    propagate_frame(prop, pf, d);
                    norm3<0>(d[0], d[1], d[2]);
                    ox = d[0]; oy = d[1]; oz = d[2];
                }
#pragma unroll
                for (int i = 0; i < 9; ++i) frame[i] = pf[i];
            }
            *ndx = shfl32(ox, win);
            *ndy = shfl32(oy, win);
            *ndz = shfl32(oz, win);
            if (wk) { wk[0] = shfl32(k1, win); wk[1] = shfl32(k2, win); }
            WAVE_SYNC();
            return 1;
        }
    }
    return 0;
}

// get_direction_ptt_d (ptt.cu:194-412); `frame` is the 9-float frame in use.
// CNT (instrumented kernels): every tap read is counted in *ntap.
template <bool INIT, bool CNT = false>
__device__ int get_direction_ptt(const DevParams &P, const DevBufs &B, PttLds &S, float *frame,
                                 Philox &rng, float dx, float dy, float dz, float px, float py, float pz,
                                 float *ndx, float *ndy, float *ndz, float *wk = nullptr, float *raw = nullptr,
                                 uint32_t *ntap = nullptr) {
    const int l = lane32();
    if (INIT) {
        if (l == 0) { frame[0] = dx; frame[1] = dy; frame[2] = dz; }
        WAVE_SYNC();
    }
    float fr[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) fr[i] = frame[i];
    const float first_val = interp4(P, B, px, py, pz, fr[0], fr[1], fr[2]);
    if (CNT) ++*ntap;
    return ptt_substep<INIT, CNT>(P, B, S, frame, fr, rng, first_val, 0, 1024 / 32, dx, dy, dz, px, py, pz, ndx, ndy,
                                  ndz, wk, raw, ntap);
}

// init_frame_ptt_d (ptt.cu:415-474); frame[15..17] is set to the binormal.
template <bool CNT = false>
__device__ bool init_frame_ptt(const DevParams &P, const DevBufs &B, PttLds &S, Philox &rng,
                               float fx, float fy, float fz, float sx, float sy, float sz, uint32_t *ntap = nullptr) {
    const int l = lane32();
    float tx, ty, tz;
    int att = 0;
    bool ok = false;
    for (; att < 2; ++att) {  // +dir first, then -dir
        const float sg = att ? -1.0f : 1.0f;
        ok = get_direction_ptt<true, CNT>(P, B, S, S.frame, rng, sg * fx, sg * fy, sg * fz, sx, sy, sz, &tx, &ty, &tz,
                                          nullptr, nullptr, ntap);
        if (ok) break;
    }
    if (!ok) return false;
    if (l == 0) {
        if (att == 0)
            for (int i = 0; i < 6; ++i) S.frame[i] = -S.frame[i];
        for (int i = 0; i < 6; ++i) S.frame[9 + i] = -S.frame[i];
        for (int i = 0; i < 3; ++i) S.frame[15 + i] = S.frame[6 + i];
    }
    WAVE_SYNC();
    return true;
}

// -------------------------------------------------------------------------
// Learned PTT proposal (no counterpart in the reference).
//
// The network predicts the 24 vertex supports that the plain getter measures
// with 8 probe taps each, from two taps per vertex; the reference's face CDF
// and acceptance rule then run on the prediction.  The MLP runs on the 32
// lanes of the seed slot: lane l owns hidden neurons l, l+32, l+64, l+96 and
// output l; each neuron is an fma chain over its inputs in order, starting
// from the bias.  Parameters are bf16, arithmetic fp32.
// -------------------------------------------------------------------------
constexpr int MLP_IN = CUSLINES_MLP_IN, MLP_H = CUSLINES_MLP_HIDDEN, MLP_OUT = CUSLINES_MLP_OUT;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_IN, MLP_W2 = MLP_B1 + MLP_H,
              MLP_B2 = MLP_W2 + MLP_H * MLP_H, MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_OUT * MLP_H,
              MLP_NPARAM = MLP_B3 + MLP_OUT;
// The network as the tracking kernel reads it: each weight matrix stored
// lane-interleaved (mlp_il_offset: the 8-weight pieces of 32 consecutive
// rows side by side, rows padded to 32), so lane j's 16-byte row loads of
// one instruction are 512 contiguous bytes; biases after the matrices.
constexpr int MLP_OUT32 = 32;
constexpr int IL_W1 = 0, IL_W2 = IL_W1 + MLP_H * MLP_IN, IL_W3 = IL_W2 + MLP_H * MLP_H,
              IL_B1 = IL_W3 + MLP_OUT32 * MLP_H, IL_B2 = IL_B1 + MLP_H, IL_B3 = IL_B2 + MLP_H,
              IL_NPARAM = IL_B3 + MLP_OUT32;
constexpr int MLP_LDS_FLOATS = ((IL_NPARAM + 1) / 2 + 3) & ~3;  // the network staged in LDS (learned kernels)
__host__ __device__ inline int mlp_il_offset(int K, int j, int k) {
    return (((j >> 5) * (K >> 3) + (k >> 3)) << 8) + ((j & 31) << 3) + (k & 7);
}
static_assert(MLP_NPARAM == CUSLINES_MLP_NPARAM, "MLP layout");
static_assert(MLP_OUT == DISC_VERT_CNT, "one output per disc vertex");
constexpr int ACT_X = 0, ACT_H1 = 64, ACT_H2 = 192, ACT_O = 320, ACT_FLOATS = 352;
constexpr int PTT_SLOT_FLOATS = 96 + ACT_FLOATS;
constexpr int FEAT_STEPS = 4;  // probe steps between the two feature taps

__device__ __forceinline__ float bf16f(uint32_t h) { return __uint_as_float(h << 16); }

// Inputs at the start of a sub-step into x[0..39] (LDS).  Lane v < 24: ODF
// taps after 4 and 8 probe steps along vertex v's curve (one propagator of
// arc length 4 probe steps, applied twice), summed.
__device__ __forceinline__ void mlp_features(const DevParams &P, const DevBufs &B, float px, float py, float pz,
                                             float dx, float dy, float dz, const float *fr, float first_val,
                                             float *x) {
    const int l = lane32();
    if (l < DISC_VERT_CNT) {
        float pf[9], prop[9], d[3];
#pragma unroll
        for (int i = 0; i < 9; ++i) pf[i] = fr[i];
#pragma unroll
        for (int i = 0; i < 9; ++i) prop[i] = B.fprop[9 * l + i];
        // both frames first (they do not depend on the taps), then both LUT
        // reads, then both block reads, so the two taps' loads overlap
        propagate_frame(prop, pf, d);
        const float q1x = px + d[0], q1y = py + d[1], q1z = pz + d[2];
        const uint32_t lv1 = B.lut[lut_index(P, pf[0], pf[1], pf[2])];
        propagate_frame(prop, pf, d);
        const float q2x = q1x + d[0], q2y = q1y + d[1], q2z = q1z + d[2];
        const uint32_t lv2 = B.lut[lut_index(P, pf[0], pf[1], pf[2])];
        const Tap a = tap_begin(P, B, lv1, q1x, q1y, q1z);
        const Tap b = tap_begin(P, B, lv2, q2x, q2y, q2z);
        const float t1 = tap_end(P, B, a);
        const float t2 = tap_end(P, B, b);
        x[l] = t1 + t2;
    } else if (l == DISC_VERT_CNT) {
        x[l] = first_val;
    } else if (l < 28) {
        x[l] = 0.0f;
    } else if (l < 31) {
        x[l] = (l == 28) ? dx : (l == 29) ? dy : dz;
    }
    if (l < 9) x[31 + l] = fr[l];
    WAVE_SYNC();
}

// One neuron: bias + sum_k w[k] in[k] as an fma chain, K a multiple of 8.
// Address-space-qualified pointers for the out-of-line network: its
// arguments arrive as generic pointers, and generic (flat) loads are slower
// than ds loads.  The learned kernels stage the network in LDS
// (stage_network), so weights and activations are both LDS reads.
#define AS_LDS __attribute__((address_space(3)))
typedef const AS_LDS uint16_t GWeight;
typedef AS_LDS float LAct;
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef uint32_t vu4 __attribute__((ext_vector_type(4)));

// Four neurons of one lane (rows j, j + 32, j + 64, j + 96 of the layer),
// each summed in input order like a single neuron; w: row j in the
// lane-interleaved layout (mlp_il_offset(K, j, 0)), the other rows 32 rows
// further, i.e. (K / 8) * 256 weights on.
template <int K>
__device__ __forceinline__ void mlp_neuron4(GWeight *__restrict__ w, GWeight *bias, const LAct *in, float *acc) {
    const AS_LDS vu4 *w8 = reinterpret_cast<const AS_LDS vu4 *>(w);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = bf16f(bias[32 * r]);
#pragma unroll 2
    for (int k8 = 0; k8 < K / 8; ++k8) {
        const vf4 a = *reinterpret_cast<const AS_LDS vf4 *>(in + 8 * k8);
        const vf4 b = *reinterpret_cast<const AS_LDS vf4 *>(in + 8 * k8 + 4);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const vu4 q = w8[k8 * 32 + r * (K / 8) * 32];
            float c = acc[r];
            c = __builtin_fmaf(bf16f(q.x & 0xffffu), a.x, c);
            c = __builtin_fmaf(bf16f(q.x >> 16), a.y, c);
            c = __builtin_fmaf(bf16f(q.y & 0xffffu), a.z, c);
            c = __builtin_fmaf(bf16f(q.y >> 16), a.w, c);
            c = __builtin_fmaf(bf16f(q.z & 0xffffu), b.x, c);
            c = __builtin_fmaf(bf16f(q.z >> 16), b.y, c);
            c = __builtin_fmaf(bf16f(q.w & 0xffffu), b.z, c);
            c = __builtin_fmaf(bf16f(q.w >> 16), b.w, c);
            acc[r] = c;
        }
    }
}

// w: the neuron's row in the lane-interleaved layout (mlp_il_offset(K, j, 0)).
template <int K>
__device__ __forceinline__ float mlp_neuron(GWeight *__restrict__ w, uint16_t bias, const LAct *in) {
    float acc = bf16f(bias);
    const AS_LDS vu4 *w8 = reinterpret_cast<const AS_LDS vu4 *>(w);
#pragma unroll 2
    for (int k8 = 0; k8 < K / 8; ++k8) {
        const vu4 q = w8[k8 * 32];
        const vf4 a = *reinterpret_cast<const AS_LDS vf4 *>(in + 8 * k8);
        const vf4 b = *reinterpret_cast<const AS_LDS vf4 *>(in + 8 * k8 + 4);
        acc = __builtin_fmaf(bf16f(q.x & 0xffffu), a.x, acc);
        acc = __builtin_fmaf(bf16f(q.x >> 16), a.y, acc);
        acc = __builtin_fmaf(bf16f(q.y & 0xffffu), a.z, acc);
        acc = __builtin_fmaf(bf16f(q.y >> 16), a.w, acc);
        acc = __builtin_fmaf(bf16f(q.z & 0xffffu), b.x, acc);
        acc = __builtin_fmaf(bf16f(q.z >> 16), b.y, acc);
        acc = __builtin_fmaf(bf16f(q.w & 0xffffu), b.z, acc);
        acc = __builtin_fmaf(bf16f(q.w >> 16), b.w, acc);
    }
    return acc;
}

// Kept out of line so the tracking loop keeps its register budget.  Both
// W_generic (the staged network) and act_generic (the slot's activations)
// point into LDS.
__device__ __noinline__ void mlp_forward(const uint16_t *__restrict__ W_generic, float *act_generic) {
    GWeight *W = (GWeight *)W_generic;
    LAct *act = (LAct *)act_generic;
    const int l = lane32();
    float v[4];
    mlp_neuron4<MLP_IN>(W + IL_W1 + mlp_il_offset(MLP_IN, l, 0), W + IL_B1 + l, act + ACT_X, v);
#pragma unroll
    for (int r = 0; r < 4; ++r) act[ACT_H1 + l + 32 * r] = (v[r] > 0.0f) ? v[r] : 0.0f;
    WAVE_SYNC();
    mlp_neuron4<MLP_H>(W + IL_W2 + mlp_il_offset(MLP_H, l, 0), W + IL_B2 + l, act + ACT_H1, v);
#pragma unroll
    for (int r = 0; r < 4; ++r) act[ACT_H2 + l + 32 * r] = (v[r] > 0.0f) ? v[r] : 0.0f;
    WAVE_SYNC();
    if (l < MLP_OUT)
        act[ACT_O + l] = mlp_neuron<MLP_H>(W + IL_W3 + mlp_il_offset(MLP_H, l, 0), W[IL_B3 + l], act + ACT_H2);
    WAVE_SYNC();
}

// Learned proposal with an exact correction.  The reference draws a face
// with probability proportional to its value F (the sum of its 3 vertex
// supports, 0 when one is below 0.4) and a point uniform in it, then keeps
// the lowest lane whose candidate has support >= 0.4 (ptt.cu:270-410).  Here
// the face is drawn from the network's estimate instead,
//   g(f) = sum of the 3 predicted supports (clipped at 0) + a floor
//          (PROPOSAL_FLOOR x the mean predicted face value),
// and accepted with probability F(f) / (PROPOSAL_M g(f)) before the
// reference's support test; F(f) needs only the exact supports of the drawn
// faces' vertices (8 probes each, measured once per sub-step and shared by
// the lanes).  An accepted candidate is then distributed as F(f) x uniform
// x [support >= 0.4], the reference's sampler, whenever F <= PROPOSAL_M g
// on every face; draws where that bound fails are counted (stats[3]).
// Up to LEARNED_ROUNDS rounds of 32 candidates, then the plain sub-step.
constexpr float PROPOSAL_M = 2.0f;
constexpr float PROPOSAL_FLOOR = 0.5f;
constexpr int LEARNED_ROUNDS = 4;

// Inclusive prefix sum over the 31 faces (lanes >= 31 keep 0), the
// Hillis-Steele order of ptt_face_cdf; returns the total on every lane.
__device__ __forceinline__ float face_scan(float v, float *out) {
    const int l = lane32();
    if (l >= DISC_FACE_CNT) v = 0.0f;
#pragma unroll
    for (int i = 1; i < 32; i <<= 1) {
        const float o = shfl_up32(v, i);
        if (l >= i && l < DISC_FACE_CNT) v = v + o;
    }
    if (out && l < DISC_FACE_CNT) out[l] = v;
    return shfl32(v, DISC_FACE_CNT - 1);
}

__device__ int get_direction_ptt_learned(const DevParams &P, const DevBufs &B, PttLds &S, float *frame, Philox &rng,
                                         bool learned, float dx, float dy, float dz, float px, float py, float pz,
                                         float *ndx, float *ndy, float *ndz, int *fell, uint32_t *misses) {
    const int l = lane32();
    float fr[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) fr[i] = frame[i];
    const float first_val = interp4(P, B, px, py, pz, fr[0], fr[1], fr[2]);
    *fell = 0;
    if (learned) {
        mlp_features(P, B, px, py, pz, dx, dy, dz, fr, first_val, S.act + ACT_X);
        mlp_forward(B.mlp, S.act);
        float *E = S.act + ACT_X;   // exact vertex supports (the inputs are dead now)
        float *G = S.act + ACT_H1;  // proposal face values (so are the hidden activations)
        float gh = 0.0f;
        if (l < DISC_FACE_CNT) {
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                const float sv = S.act[ACT_O + c_disc_face[l * 3 + jj]];
                gh = gh + ((sv > 0.0f) ? sv : 0.0f);
            }
        }
        const float tot = face_scan(gh, nullptr);
        const float g = gh + (PROPOSAL_FLOOR * (tot / (float)DISC_FACE_CNT) + 1e-3f);
        if (l < DISC_FACE_CNT) G[l] = g;
        const float last = face_scan(g, S.face_cdf);
        WAVE_SYNC();
        const double mc = (double)P.max_curv;
        uint32_t known = 0u;
        for (int round = 0; round < LEARNED_ROUNDS; ++round) {
            float r1 = rng.uniform(), r2 = rng.uniform();
            if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
            const float tmp = rng.uniform() * last;
            const float ua = rng.uniform();
            int jj = 0, hi = DISC_FACE_CNT - 1;
            while (jj < hi) {
                const int mid = (jj + hi) >> 1;
                if (S.face_cdf[mid] >= tmp) hi = mid; else jj = mid + 1;
            }
            const int f0 = c_disc_face[jj * 3], f1 = c_disc_face[jj * 3 + 1], f2 = c_disc_face[jj * 3 + 2];
            uint32_t need = (1u << f0) | (1u << f1) | (1u << f2);
#pragma unroll
            for (int m = 16; m; m >>= 1) need |= (uint32_t)shfl_xor32((int)need, m);
            const uint32_t todo = need & ~known;
            if (l < DISC_VERT_CNT && ((todo >> l) & 1u)) {
                float pf[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) pf[i] = fr[i];
                const float k1v = (float)((double)c_disc_vert[2 * l] * mc);
                const float k2v = (float)((double)c_disc_vert[2 * l + 1] * mc);
                E[l] = data_support(P, B, first_val, px, py, pz, k1v, k2v, pf, __builtin_inff(), nullptr,
                                    B.vprop ? B.vprop + 9 * l : nullptr);
            }
            known |= need;
            WAVE_SYNC();
            // the reference's face value from the exact supports (vertex
            // phase test `< 0.4`, ptt_face_cdf's zero rule)
            float F = 0.0f;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const float e = E[c_disc_face[jj * 3 + q]];
                const float vv = (e < P.min_support_f) ? 0.0f : e;
                if (vv == 0.0f) { F = 0.0f; break; }
                F = F + vv;
            }
            const float mg = PROPOSAL_M * G[jj];
            const bool acc = ua * mg < F;
            const float vx0 = (float)(mc * (double)c_disc_vert[f0 * 2]);
            const float vx1 = (float)(mc * (double)c_disc_vert[f1 * 2]);
            const float vx2 = (float)(mc * (double)c_disc_vert[f2 * 2]);
            const float vy0 = (float)(mc * (double)c_disc_vert[f0 * 2 + 1]);
            const float vy1 = (float)(mc * (double)c_disc_vert[f1 * 2 + 1]);
            const float vy2 = (float)(mc * (double)c_disc_vert[f2 * 2 + 1]);
            const float k1 = __builtin_fmaf(r2, vx2 - vx0, __builtin_fmaf(r1, vx1 - vx0, vx0));
            const float k2 = __builtin_fmaf(r2, vy2 - vy0, __builtin_fmaf(r1, vy1 - vy0, vy0));
            float sup = 0.0f;
            if (acc) {
                float pf[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) pf[i] = fr[i];
                const float stop_at = P.ptt_nonneg ? P.min_support_f : __builtin_inff();
                sup = data_support(P, B, first_val, px, py, pz, k1, k2, pf, stop_at);
            }
            *misses += (uint32_t)__popc(half_ballot(F > mg));
            const uint32_t msk = half_ballot(acc && sup >= P.min_support_f);
            if (msk != 0u) {
                const int win = __ffs(msk) - 1;
                float ox = dx, oy = dy, oz = dz;
                if (l == win) {
                    float prop[9], d[3], pf[9];
                    prepare_propagator(k1, k2, P.ptt_sub_step, prop);
#pragma unroll
                    for (int i = 0; i < 9; ++i) pf[i] = fr[i];
                    propagate_frame(prop, pf, d);
                    norm3<0>(d[0], d[1], d[2]);
                    ox = d[0]; oy = d[1]; oz = d[2];
#pragma unroll
                    for (int i = 0; i < 9; ++i) frame[i] = pf[i];
                }
                *ndx = shfl32(ox, win);
                *ndy = shfl32(oy, win);
                *ndz = shfl32(oz, win);
                WAVE_SYNC();
                return 1;
            }
        }
        *fell = 1;
    }
    return ptt_substep<false>(P, B, S, frame, fr, rng, first_val, 0, 1024 / 32, dx, dy, dz, px, py, pz, ndx, ndy,
                              ndz, nullptr, nullptr);
}

// Training pair of a plain sub-step: inputs at the sub-step's start and the
// 24 measured vertex supports.
__device__ void record_pair(const DevParams &P, const DevBufs &B, PttLds &S, float px, float py, float pz, float dx,
                            float dy, float dz, const float *fr0, const float *raw) {
    const int l = lane32();
    const float first_val = interp4(P, B, px, py, pz, fr0[0], fr0[1], fr0[2]);
    mlp_features(P, B, px, py, pz, dx, dy, dz, fr0, first_val, S.act + ACT_X);
    int idx = 0;
    if (l == 0) idx = atomicAdd(B.rec_count, 1);
    idx = shfl32(idx, 0);
    if (idx < P.rec_cap) {
        float *o = B.rec + (size_t)idx * CUSLINES_PAIR_FLOATS;
        for (int j = l; j < MLP_IN; j += 32) o[j] = S.act[ACT_X + j];
        if (l < DISC_VERT_CNT) o[MLP_IN + l] = raw[l];
    }
    WAVE_SYNC();
}

__device__ __forceinline__ uint32_t mix32(uint32_t h) {  // murmur3 finaliser
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}

// -------------------------------------------------------------------------
// Tracker (generate_streamlines_cuda.cu:277-421, boot.cu:720-824)
// -------------------------------------------------------------------------
struct SlotLds {
    ProbLds prob;
    BootLds boot;
    PttLds ptt;
};

// V (PTT only): 0 plain getter, 1 learned proposal, 2 plain getter recording
// training pairs (tag identifies seed, direction and pass for the hash), 3
// plain getter counting its taps and sub-steps into *ntap and *nsub.
template <int MODEL, int V = 0>
__device__ int tracker(const DevParams &P, const DevBufs &B, const float *sv, SlotLds &S, Philox &rng,
                       float sx, float sy, float sz, float dx, float dy, float dz, float *frame,
                       float *__restrict__ sl, uint32_t tag = 0, uint32_t *ntap = nullptr,
                       uint32_t *nsub = nullptr) {
    const int l = lane32();
    const int step_frac = (MODEL == CUSLINES_PTT) ? P.step_frac : 1;
    const int MAXL = P.max_sline_len;
    const double sub = P.step_d / (double)step_frac;
    float px = sx, py = sy, pz = sz;
    if (l == 0) { sl[0] = px; sl[1] = py; sl[2] = pz; }
    int ltries = 0, lfalls = 0, nplain = 0;
    uint32_t lmiss = 0;
    bool learned = (V == 1);
    int i;
    for (i = 1; i < MAXL * step_frac; ++i) {
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        int ndir;
        if (MODEL == CUSLINES_PROB)
            ndir = get_direction_prob<false>(P, B, sv, S.prob, rng, dx, dy, dz, px, py, pz, &nx, &ny, &nz);
        else if (MODEL == CUSLINES_PTT && V == 1) {
            int fell = 0;
            ndir = get_direction_ptt_learned(P, B, S.ptt, frame, rng, learned, dx, dy, dz, px, py, pz, &nx, &ny, &nz,
                                             &fell, &lmiss);
            if (learned) {
                ++ltries;
                lfalls += fell;
                if (ltries >= 16 && 4 * lfalls > ltries) learned = false;
            } else {
                ++nplain;
            }
        } else if (MODEL == CUSLINES_PTT && V == 2) {
            float fr0[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) fr0[q] = frame[q];
            const bool rec = mix32(tag * 0x9E3779B1u + (uint32_t)i) < P.rec_thresh;
            ndir = get_direction_ptt<false>(P, B, S.ptt, frame, rng, dx, dy, dz, px, py, pz, &nx, &ny, &nz, nullptr,
                                            rec ? S.ptt.act + ACT_O : nullptr);
            if (rec) record_pair(P, B, S.ptt, px, py, pz, dx, dy, dz, fr0, S.ptt.act + ACT_O);
        } else if (MODEL == CUSLINES_PTT && V == 3) {
            ++*nsub;
            ndir = get_direction_ptt<false, true>(P, B, S.ptt, frame, rng, dx, dy, dz, px, py, pz, &nx, &ny, &nz,
                                                  nullptr, nullptr, ntap);
        } else if (MODEL == CUSLINES_PTT)
            ndir = get_direction_ptt<false>(P, B, S.ptt, frame, rng, dx, dy, dz, px, py, pz, &nx, &ny, &nz);
        else
            ndir = get_direction_boot<MODEL>(P, B, sv, S.boot, rng, 5, dx, dy, dz, px, py, pz, &nx, &ny, &nz);
        if (ndir == 0) break;
        dx = nx; dy = ny; dz = nz;
        px = (float)__builtin_fma((double)(dx / 1.0f), sub, (double)px);
        py = (float)__builtin_fma((double)(dy / 1.0f), sub, (double)py);
        pz = (float)__builtin_fma((double)(dz / 1.0f), sub, (double)pz);
        if (l == 0 && (i % step_frac) == 0) {
            float *o = sl + 3 * (size_t)(i / step_frac);
            o[0] = px; o[1] = py; o[2] = pz;
        }
        const int tissue = check_point(P, B.stop, px, py, pz);
        if (tissue == CUSLINES_ENDPOINT || tissue == CUSLINES_INVALIDPOINT || tissue == CUSLINES_OUTSIDEIMAGE)
            break;
    }
    if (V == 1 && B.stats && l == 0) {
        atomicAdd(B.stats, (unsigned long long)(ltries - lfalls));
        atomicAdd(B.stats + 1, (unsigned long long)lfalls);
        atomicAdd(B.stats + 2, (unsigned long long)nplain);
        atomicAdd(B.stats + 3, (unsigned long long)lmiss);
    }
    int nsteps = i / step_frac;
    if ((i % step_frac) != 0 && i < step_frac * (MAXL - 1)) {
        ++nsteps;
        if (l == 0) {
            float *o = sl + 3 * (size_t)nsteps;
            o[0] = px; o[1] = py; o[2] = pz;
        }
    }
    return nsteps;
}

// -------------------------------------------------------------------------
// PTT with L lanes per seed (L in {8, 16, 32}; 64/L seeds per wave).
//
// The reference gives every seed 32 lanes, each with its own Philox stream
// (subsequence 32*slot + lane + 1).  Here those 32 streams are virtual lanes
// whose state -- the next draw index k and, for a cached Box-Muller value,
// the index c of the draw pair it came from (-1: none) -- lives in LDS, and
// draws are computed from (subsequence, k) directly (counter-based Philox).
// A physical lane p evaluates virtual lanes p, p+L, p+2L, ... in batches.
//  - vertex phase: all 24 candidates (their supports build the face CDF);
//  - INIT rejection: all 32 candidates (the winner is the arg-max);
//  - non-INIT rejection: batches in lane order, stopping at the first batch
//    that holds a candidate with enough support.  The winner is the lowest
//    virtual lane that passes, as in the reference, and every virtual lane's
//    stream still advances by its 3 uniforms per round.
// The vertex propagators depend only on the disc and MAX_CURVATURE and are
// computed once per block.
// -------------------------------------------------------------------------
template <int L>
struct Grp {
    static __device__ __forceinline__ int lane() { return threadIdx.x & (L - 1); }
    static __device__ __forceinline__ uint32_t ballot(bool p) {
        const unsigned long long b = __ballot(p);
        const int base = (threadIdx.x & 63) & ~(L - 1);
        return (uint32_t)((b >> base) & (L == 32 ? 0xffffffffull : ((1ull << L) - 1)));
    }
    static __device__ __forceinline__ float shfl(float v, int src) { return __shfl(v, src, L); }
    static __device__ __forceinline__ int shfl(int v, int src) { return __shfl(v, src, L); }
    static __device__ __forceinline__ float shfl_xor(float v, int m) { return __shfl_xor(v, m, L); }
    static __device__ __forceinline__ int shfl_xor(int v, int m) { return __shfl_xor(v, m, L); }
};

// Philox draw k of subsequence `sub`, with a one-block cache per lane.
struct DrawCache {
    uint32_t sub, blk, r0, r1, r2, r3;
    int valid;
};

__device__ __forceinline__ uint32_t philox_draw(uint32_t key0, uint32_t key1, uint32_t sub, uint32_t k,
                                                DrawCache &dc) {
    const uint32_t blk = k >> 2;
    if (!(dc.valid && dc.sub == sub && dc.blk == blk)) {
        uint32_t x0 = blk, x1 = 0, x2 = sub, x3 = 0;
        uint32_t a = key0, b = key1;
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            const uint32_t lo0 = 0xD2511F53u * x0, hi0 = __umulhi(0xD2511F53u, x0);
            const uint32_t lo1 = 0xCD9E8D57u * x2, hi1 = __umulhi(0xCD9E8D57u, x2);
            const uint32_t n0 = hi1 ^ x1 ^ a, n2 = hi0 ^ x3 ^ b;
            x0 = n0; x1 = lo1; x2 = n2; x3 = lo0;
            a += 0x9E3779B9u;
            b += 0xBB67AE85u;
        }
        dc.r0 = x0; dc.r1 = x1; dc.r2 = x2; dc.r3 = x3;
        dc.sub = sub; dc.blk = blk; dc.valid = 1;
    }
    const int w = k & 3;
    return (w == 0) ? dc.r0 : (w == 1) ? dc.r1 : (w == 2) ? dc.r2 : dc.r3;
}

__device__ __forceinline__ float uniform_of(uint32_t x) { return 2.32830644e-10f + (float)x * 2.32830644e-10f; }

// Box-Muller pair of draws (x, y): the sine term is returned first and the
// cosine term cached, as Philox::normal does.
__device__ __forceinline__ void box_muller(uint32_t x, uint32_t y, float *sn_out, float *cs_out) {
    const float u = 2.32830644e-10f + (float)x * 2.32830644e-10f;
    const float v = 1.46291807e-09f + (float)y * 1.46291807e-09f;
    const float s = __builtin_sqrtf(-2.0f * ulog(u));
    float sn, cs;
    usincos(v, &sn, &cs);
    *sn_out = sn * s;
    *cs_out = cs * s;
}

struct VLane {
    uint32_t key0, key1, sub;  // stream of this virtual lane
    uint32_t k;                // next draw index
    int c;                     // draw pair of the cached normal, -1 if none
    __device__ __forceinline__ float uniform(DrawCache &dc) { return uniform_of(philox_draw(key0, key1, sub, k++, dc)); }
    __device__ __forceinline__ float normal(DrawCache &dc) {
        float sn, cs;
        if (c >= 0) {
            box_muller(philox_draw(key0, key1, sub, (uint32_t)c, dc), philox_draw(key0, key1, sub, (uint32_t)c + 1, dc),
                       &sn, &cs);
            c = -1;
            return cs;
        }
        box_muller(philox_draw(key0, key1, sub, k, dc), philox_draw(key0, key1, sub, k + 1, dc), &sn, &cs);
        c = (int)k;
        k += 2;
        return sn;
    }
};

// LDS of one PTT seed group.
struct PttG {
    float *frame;     // 18
    float *vert_pdf;  // 24
    float *face_cdf;  // 32
    uint32_t *rk;     // 32 virtual-lane draw indices
    int *rc;          // 32 cached-normal pair indices
};
constexpr int PTTG_FLOATS = 160;

// Block-shared constants: 24 vertex propagators (9 each) and the scaled
// disc vertices (MAX_CURVATURE * DISC_VERT, 48).
struct PttConst {
    float *vprop;  // 24 * 9
    float *vk;     // 48
};
constexpr int PTTC_FLOATS = 24 * 9 + 48;

__device__ void ptt_const_init(const DevParams &P, PttConst &C) {
    const double mc = (double)P.max_curv;
    for (int i = threadIdx.x; i < 48; i += blockDim.x) C.vk[i] = (float)(mc * (double)c_disc_vert[i]);
    for (int i = threadIdx.x; i < 24; i += blockDim.x) {
        const float k1 = (float)((double)c_disc_vert[2 * i] * mc);
        const float k2 = (float)((double)c_disc_vert[2 * i + 1] * mc);
        prepare_propagator(k1, k2, P.probe_step, C.vprop + 9 * i);
    }
}

__device__ __forceinline__ void probe_frame_init(const float *frame, VLane &vl, DrawCache &dc, float *pf) {
    pf[0] = frame[0]; pf[1] = frame[1]; pf[2] = frame[2];
    norm3<0>(pf[0], pf[1], pf[2]);
    pf[3] = vl.normal(dc);
    pf[4] = vl.normal(dc);
    pf[5] = vl.normal(dc);
    const float dt = __builtin_fmaf(pf[5], pf[2], __builtin_fmaf(pf[4], pf[1], pf[3] * pf[0]));
    pf[3] = __builtin_fmaf(-dt, pf[0], pf[3]);
    pf[4] = __builtin_fmaf(-dt, pf[1], pf[4]);
    pf[5] = __builtin_fmaf(-dt, pf[2], pf[5]);
    const float n2 = __builtin_fmaf(pf[5], pf[5], __builtin_fmaf(pf[4], pf[4], pf[3] * pf[3]));
    if (n2 < PTT_NORM_EPS) {
        const float ax = fabsf(pf[0]), ay = fabsf(pf[1]), az = fabsf(pf[2]);
        if (ax <= ay && ax <= az) { pf[3] = 0.0f; pf[4] = pf[2]; pf[5] = -pf[1]; }
        else if (ay <= az) { pf[3] = -pf[2]; pf[4] = 0.0f; pf[5] = pf[0]; }
        else { pf[3] = pf[1]; pf[4] = -pf[0]; pf[5] = 0.0f; }
    }
    norm3<1>(pf[3], pf[4], pf[5]);
    crossnorm3<2>(pf + 6, pf, pf + 3);
}

// Support of one candidate with a precomputed propagator; probing stops
// once the running support reaches `stop_at` (see data_support).
__device__ __forceinline__ float support_prop(const DevParams &P, const DevBufs &B, float support, float px,
                                              float py, float pz, const float *prop, float *pf, float stop_at) {
    for (int i = 0; i < P.probe_quality; ++i) {
        if (support >= stop_at) break;
        float direc[3];
        propagate_frame(prop, pf, direc);
        px += direc[0]; py += direc[1]; pz += direc[2];
        support += interp4(P, B, px, py, pz, pf[0], pf[1], pf[2]);
    }
    return support;
}

template <int L, bool INIT>
__device__ int get_direction_ptt_g(const DevParams &P, const DevBufs &B, const PttConst &K, PttG &S, float *frame,
                                   uint32_t key0, uint32_t key1, uint32_t sub0, DrawCache &dc, float dx, float dy,
                                   float dz, float px, float py, float pz, float *ndx, float *ndy, float *ndz) {
    using G = Grp<L>;
    const int p = G::lane();
    if (INIT) {
        if (p == 0) { frame[0] = dx; frame[1] = dy; frame[2] = dz; }
        WAVE_SYNC();
    }
    float fr[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) fr[i] = frame[i];
    const float first_val = interp4(P, B, px, py, pz, fr[0], fr[1], fr[2]);
    // vertex phase: virtual lanes 0..23
    bool found = false;
    constexpr int NC = (DISC_VERT_CNT + L - 1) / L;  // curves per lane
    if (INIT || NC > 2) {
        for (int v = p; v < DISC_VERT_CNT; v += L) {
            float pf[9];
            if (INIT) {
                VLane vl{key0, key1, sub0 + (uint32_t)v, S.rk[v], S.rc[v]};
                probe_frame_init(fr, vl, dc, pf);
                S.rk[v] = vl.k;
                S.rc[v] = vl.c;
            } else {
#pragma unroll
                for (int i = 0; i < 9; ++i) pf[i] = fr[i];
            }
            const float sup = support_prop(P, B, first_val, px, py, pz, K.vprop + 9 * v, pf, __builtin_inff());
            const bool ok = !(sup < P.min_support_f);
            S.vert_pdf[v] = ok ? sup : 0.0f;
            found |= ok;
        }
    } else {
        // this lane's NC <= 2 curves v = p, p + L probed in lockstep: each
        // probe step issues every curve's LUT read, then every curve's block
        // read, then blends, so the curves' load latencies overlap (with 3
        // curves per lane the registers of three taps in flight spill, which
        // costs more than it hides).  Each curve's arithmetic is
        // support_prop's, in its order.
        float pf[NC][9], qx[NC], qy[NC], qz[NC], sup[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
#pragma unroll
            for (int i = 0; i < 9; ++i) pf[c][i] = fr[i];
            qx[c] = px; qy[c] = py; qz[c] = pz;
            sup[c] = first_val;
        }
        for (int i = 0; i < P.probe_quality; ++i) {
            uint32_t lv[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                if (p + c * L < DISC_VERT_CNT) {
                    float direc[3];
                    propagate_frame(K.vprop + 9 * (p + c * L), pf[c], direc);
                    qx[c] += direc[0]; qy[c] += direc[1]; qz[c] += direc[2];
                    lv[c] = B.lut[lut_index(P, pf[c][0], pf[c][1], pf[c][2])];
                }
            }
            TapB t[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (p + c * L < DISC_VERT_CNT) t[c] = tapb_begin(P, B, lv[c], qx[c], qy[c], qz[c]);
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (p + c * L < DISC_VERT_CNT) sup[c] += tapb_end(t[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int v = p + c * L;
            if (v < DISC_VERT_CNT) {
                const bool ok = !(sup[c] < P.min_support_f);
                S.vert_pdf[v] = ok ? sup[c] : 0.0f;
                found |= ok;
            }
        }
    }
    if (G::ballot(found) == 0u) return 0;
    WAVE_SYNC();
    // faces and the 32-wide Hillis-Steele scan of prefix_sum_sh_d, per lane
    float cdf[DISC_FACE_CNT];
#pragma unroll
    for (int f = 0; f < DISC_FACE_CNT; ++f) {
        float face = 0.0f;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const float vv = S.vert_pdf[c_disc_face[f * 3 + jj]];
            if (vv == 0.0f) { face = 0.0f; break; }
            face = face + vv;
        }
        cdf[f] = face;
    }
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
#pragma unroll
        for (int f = DISC_FACE_CNT - 1; f >= d; --f) cdf[f] = cdf[f] + cdf[f - d];
    }
    const float last = cdf[DISC_FACE_CNT - 1];
    if (last == 0.0f) return 0;
    for (int round = 0; round < 1024 / 32; ++round) {
        float best = -__builtin_inff();
        int bestv = -1;
        int win = -1;
        float wk1 = 0.0f, wk2 = 0.0f, wpf[9];
        for (int b = 0; b < 32 / L; ++b) {
            const int v = b * L + p;
            VLane vl{key0, key1, sub0 + (uint32_t)v, S.rk[v], INIT ? S.rc[v] : -1};
            float r1 = vl.uniform(dc), r2 = vl.uniform(dc);
            if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
            const float tmp = vl.uniform(dc) * last;
            int jj = DISC_FACE_CNT - 1;
#pragma unroll
            for (int f = DISC_FACE_CNT - 2; f >= 0; --f)
                if (cdf[f] >= tmp) jj = f;
            const int f0 = c_disc_face[jj * 3], f1 = c_disc_face[jj * 3 + 1], f2 = c_disc_face[jj * 3 + 2];
            const float vx0 = K.vk[f0 * 2], vx1 = K.vk[f1 * 2], vx2 = K.vk[f2 * 2];
            const float vy0 = K.vk[f0 * 2 + 1], vy1 = K.vk[f1 * 2 + 1], vy2 = K.vk[f2 * 2 + 1];
            const float k1 = __builtin_fmaf(r2, vx2 - vx0, __builtin_fmaf(r1, vx1 - vx0, vx0));
            const float k2 = __builtin_fmaf(r2, vy2 - vy0, __builtin_fmaf(r1, vy1 - vy0, vy0));
            float pf[9];
            if (INIT) {
                probe_frame_init(fr, vl, dc, pf);
                S.rk[v] = vl.k;
                S.rc[v] = vl.c;
            } else {
#pragma unroll
                for (int i = 0; i < 9; ++i) pf[i] = fr[i];
            }
            float prop[9];
            prepare_propagator(k1, k2, P.probe_step, prop);
            const float sup = support_prop(P, B, first_val, px, py, pz, prop, pf,
                                           (!INIT && P.ptt_nonneg) ? P.min_support_f : __builtin_inff());
            if (INIT) {
                // batches visit this lane's virtual lanes in increasing order, so
                // keeping the first of equal supports keeps the lowest one
                if (bestv < 0 || sup > best) {
                    best = sup;
                    bestv = v;
#pragma unroll
                    for (int i = 0; i < 9; ++i) wpf[i] = pf[i];
                }
            } else {
                const uint32_t m = G::ballot(sup >= P.min_support_f);
                if (m != 0u) {
                    win = b * L + (__ffs(m) - 1);
                    if (v == win) {
                        wk1 = k1; wk2 = k2;
                    }
                    break;
                }
            }
        }
        if (INIT) {
            // arg-max over the 32 virtual lanes, ties to the lowest virtual lane
            float m = best;
            int mv = bestv;
            for (int o = L / 2; o; o >>= 1) {
                const float om = G::shfl_xor(m, o);
                const int ov = G::shfl_xor(mv, o);
                if (om > m || (om == m && ov < mv)) { m = om; mv = ov; }
            }
            WAVE_SYNC();
            if (m >= P.min_support_f) {
                if (bestv == mv) {
#pragma unroll
                    for (int i = 0; i < 9; ++i) frame[i] = wpf[i];
                }
                *ndx = dx; *ndy = dy; *ndz = dz;
                WAVE_SYNC();
                return 1;
            }
        } else {
            // every virtual lane drew its 3 uniforms this round
            for (int v = p; v < 32; v += L) S.rk[v] += 3u;
            WAVE_SYNC();
            if (win >= 0) {
                float ox = 0.0f, oy = 0.0f, oz = 0.0f;
                if ((win & (L - 1)) == p) {
                    float prop[9], d[3], pf[9];
                    prepare_propagator(wk1, wk2, P.ptt_sub_step, prop);
#pragma unroll
                    for (int i = 0; i < 9; ++i) pf[i] = fr[i];
                    propagate_frame(prop, pf, d);
                    norm3<0>(d[0], d[1], d[2]);
                    ox = d[0]; oy = d[1]; oz = d[2];
#pragma unroll
                    for (int i = 0; i < 9; ++i) frame[i] = pf[i];
                }
                const int wl = win & (L - 1);
                *ndx = G::shfl(ox, wl);
                *ndy = G::shfl(oy, wl);
                *ndz = G::shfl(oz, wl);
                WAVE_SYNC();
                return 1;
            }
        }
    }
    return 0;
}

template <int L>
__device__ bool init_frame_ptt_g(const DevParams &P, const DevBufs &B, const PttConst &K, PttG &S, uint32_t key0,
                                 uint32_t key1, uint32_t sub0, DrawCache &dc, float fx, float fy, float fz, float sx,
                                 float sy, float sz) {
    const int p = Grp<L>::lane();
    float tx, ty, tz;
    int att = 0;
    bool ok = false;
    for (; att < 2; ++att) {
        const float sg = att ? -1.0f : 1.0f;
        ok = get_direction_ptt_g<L, true>(P, B, K, S, S.frame, key0, key1, sub0, dc, sg * fx, sg * fy, sg * fz, sx, sy,
                                          sz, &tx, &ty, &tz);
        if (ok) break;
    }
    if (!ok) return false;
    if (p == 0) {
        if (att == 0)
            for (int i = 0; i < 6; ++i) S.frame[i] = -S.frame[i];
        for (int i = 0; i < 6; ++i) S.frame[9 + i] = -S.frame[i];
        for (int i = 0; i < 3; ++i) S.frame[15 + i] = S.frame[6 + i];
    }
    WAVE_SYNC();
    return true;
}

template <int L>
__device__ int tracker_ptt_g(const DevParams &P, const DevBufs &B, const PttConst &K, PttG &S, float *frame,
                             uint32_t key0, uint32_t key1, uint32_t sub0, DrawCache &dc, float sx, float sy, float sz,
                             float dx, float dy, float dz, float *__restrict__ sl) {
    const int p = Grp<L>::lane();
    const int step_frac = P.step_frac;
    const int MAXL = P.max_sline_len;
    const double sub = P.step_d / (double)step_frac;
    float px = sx, py = sy, pz = sz;
    if (p == 0) { sl[0] = px; sl[1] = py; sl[2] = pz; }
    int i;
    for (i = 1; i < MAXL * step_frac; ++i) {
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        const int ndir = get_direction_ptt_g<L, false>(P, B, K, S, frame, key0, key1, sub0, dc, dx, dy, dz, px, py, pz,
                                                       &nx, &ny, &nz);
        if (ndir == 0) break;
        dx = nx; dy = ny; dz = nz;
        px = (float)__builtin_fma((double)(dx / 1.0f), sub, (double)px);
        py = (float)__builtin_fma((double)(dy / 1.0f), sub, (double)py);
        pz = (float)__builtin_fma((double)(dz / 1.0f), sub, (double)pz);
        if (p == 0 && (i % step_frac) == 0) {
            float *o = sl + 3 * (size_t)(i / step_frac);
            o[0] = px; o[1] = py; o[2] = pz;
        }
        const int tissue = check_point(P, B.stop, px, py, pz);
        if (tissue == CUSLINES_ENDPOINT || tissue == CUSLINES_INVALIDPOINT || tissue == CUSLINES_OUTSIDEIMAGE) break;
    }
    int nsteps = i / step_frac;
    if ((i % step_frac) != 0 && i < step_frac * (MAXL - 1)) {
        ++nsteps;
        if (p == 0) {
            float *o = sl + 3 * (size_t)nsteps;
            o[0] = px; o[1] = py; o[2] = pz;
        }
    }
    return nsteps;
}

template <int L>
__device__ void track_seed_ptt_g(const DevParams &P, const DevBufs &B, const PttConst &K, PttG &S, int s,
                                 const float *__restrict__ seeds, int rng_chunk, const int *__restrict__ offsets,
                                 const int *__restrict__ dirs, int *__restrict__ sline_seed, int *__restrict__ sline_len,
                                 float *__restrict__ sline) {
    const int p = Grp<L>::lane();
    const float sx = seeds[3 * s], sy = seeds[3 * s + 1], sz = seeds[3 * s + 2];
    const uint32_t key0 = P.seed_lo, key1 = P.seed_hi;
    const uint32_t sub0 = 32u * (uint32_t)(s % rng_chunk) + 1u;
    for (int v = p; v < 32; v += L) { S.rk[v] = 0u; S.rc[v] = -1; }
    WAVE_SYNC();
    DrawCache dc;
    dc.valid = 0;
    const int ndir = (P.slot_cap > 0 && offsets[s + 1] > P.slot_cap) ? 0 : offsets[s + 1] - offsets[s];
    int off = offsets[s];
    const size_t slot_f = (size_t)2 * P.max_sline_len * 3;
    for (int i = 0; i < ndir; ++i, ++off) {
        const int vi = dirs[(size_t)s * P.nverts + i];
        const float fx = B.verts[3 * vi], fy = B.verts[3 * vi + 1], fz = B.verts[3 * vi + 2];
        float *cur = sline + (size_t)off * slot_f;
        if (p == 0) sline_seed[off] = s;
        if (!init_frame_ptt_g<L>(P, B, K, S, key0, key1, sub0, dc, fx, fy, fz, sx, sy, sz)) {
            if (p == 0) {
                sline_len[off] = 1;
                cur[0] = sx; cur[1] = sy; cur[2] = sz;
            }
            continue;
        }
        int nb = 0, nf = 0;
        for (int pass = 0; pass < 2; ++pass) {
            const float sg = pass ? 1.0f : -1.0f;
            float *out = pass ? cur + 3 * (size_t)(nb - 1) : cur;
            const int ns = tracker_ptt_g<L>(P, B, K, S, S.frame + 9 * pass, key0, key1, sub0, dc, sx, sy, sz, sg * fx,
                                            sg * fy, sg * fz, out);
            if (pass == 0) {
                nb = ns;
                MEM_SYNC();
                for (int j = p; j < nb / 2; j += L) {
                    float *a = cur + 3 * (size_t)j, *b = cur + 3 * (size_t)(nb - 1 - j);
                    const float a0 = a[0], a1 = a[1], a2 = a[2];
                    a[0] = b[0]; a[1] = b[1]; a[2] = b[2];
                    b[0] = a0; b[1] = a1; b[2] = a2;
                }
                MEM_SYNC();
            } else {
                nf = ns;
            }
        }
        if (p == 0) sline_len[off] = nb - 1 + nf;
    }
}

// track_seed_ptt_g's work for a stream of seeds with the stepping loop
// flattened, as flat_slot does for 32-lane slots: one loop iteration is one
// sub-step of whichever pass the L-lane group holds, and a group whose pass
// ends finishes it and starts its next pass, direction or seed at once, so
// the 64/L seeds of a wavefront keep sub-stepping together.  Seeds come from
// the XCD-ordered queues over the spatial order.  `st` (16 floats of the
// group's LDS) holds the seed's index, direction count, next direction,
// first slot, current slot, backward length and stream base (ints), its
// seed point and the direction (floats).
template <int L>
__device__ void flat_group_ptt(const DevParams &P, const DevBufs &B, const PttConst &K, PttG &S, float *st,
                               const float *__restrict__ seeds, int nseeds, int rng_chunk,
                               const int *__restrict__ offsets, const int *__restrict__ dirs,
                               int *__restrict__ sline_seed, int *__restrict__ sline_len, float *__restrict__ sline,
                               const int *__restrict__ order, int *__restrict__ work) {
    using G = Grp<L>;
    const int p = G::lane();
    const int xcd = blockIdx.x & (QUEUES - 1);
    const int step_frac = P.step_frac, MAXL = P.max_sline_len;
    const int nsub_max = MAXL * step_frac;
    const double sub = P.step_d / (double)step_frac;
    const size_t slot_f = (size_t)2 * MAXL * 3;
    const uint32_t key0 = P.seed_lo, key1 = P.seed_hi;
    int *si = reinterpret_cast<int *>(st);  // s, ndir, next direction, offsets[s], slot, nb, sub0
    float *sf = st + 8;                     // seed point [0..2], direction [3..5]
    if (p == 0) { si[1] = 0; si[2] = 0; }
    WAVE_SYNC();
    DrawCache dc;
    dc.valid = 0;
    int qk = 0;
    bool have = false;
    int pass = 0, i = 1, ri = 1, qi = 0;  // sub-step i = qi * step_frac + ri
    float px = 0.0f, py = 0.0f, pz = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
    float *sl = sline;
    for (;;) {
        if (!have) {
            while (qk < QUEUES) {
                const int di = si[2];
                if (di >= si[1]) {
                    const int qx = (xcd + qk) & (QUEUES - 1);
                    const int lo = (int)((long long)nseeds * qx / QUEUES);
                    const int hi = (int)((long long)nseeds * (qx + 1) / QUEUES);
                    int q = 0;
                    if (p == 0) q = atomicAdd(work + qx, 1);
                    q = G::shfl(q, 0);
                    if (lo + q >= hi) { ++qk; continue; }
                    const int s = order ? order[lo + q] : lo + q;
                    const int nd = (P.slot_cap > 0 && offsets[s + 1] > P.slot_cap) ? 0 : offsets[s + 1] - offsets[s];
                    WAVE_SYNC();
                    for (int v = p; v < 32; v += L) { S.rk[v] = 0u; S.rc[v] = -1; }
                    if (p == 0) {
                        si[0] = s; si[1] = nd; si[2] = 0; si[3] = offsets[s];
                        si[6] = (int)(32u * (uint32_t)(s % rng_chunk) + 1u);
                        sf[0] = seeds[3 * s]; sf[1] = seeds[3 * s + 1]; sf[2] = seeds[3 * s + 2];
                    }
                    dc.valid = 0;
                    WAVE_SYNC();
                    continue;
                }
                const int s = si[0], off = si[3] + di;
                const uint32_t sub0 = (uint32_t)si[6];
                const float sx = sf[0], sy = sf[1], sz = sf[2];
                const int vi = dirs[(size_t)s * P.nverts + di];
                const float fx = B.verts[3 * vi], fy = B.verts[3 * vi + 1], fz = B.verts[3 * vi + 2];
                float *cur = sline + (size_t)off * slot_f;
                WAVE_SYNC();
                if (p == 0) {
                    sline_seed[off] = s;
                    si[2] = di + 1; si[4] = off;
                    sf[3] = fx; sf[4] = fy; sf[5] = fz;
                }
                WAVE_SYNC();
                if (!init_frame_ptt_g<L>(P, B, K, S, key0, key1, sub0, dc, fx, fy, fz, sx, sy, sz)) {
                    if (p == 0) {
                        sline_len[off] = 1;
                        cur[0] = sx; cur[1] = sy; cur[2] = sz;
                    }
                    continue;
                }
                pass = 0;
                px = sx; py = sy; pz = sz;
                dx = -fx; dy = -fy; dz = -fz;
                sl = cur;
                if (p == 0) { sl[0] = px; sl[1] = py; sl[2] = pz; }
                i = 1; qi = (step_frac == 1) ? 1 : 0; ri = (step_frac == 1) ? 0 : 1;
                have = true;
                break;
            }
            if (!have) break;  // the queues are empty
        }
        // one sub-step of the pass (tracker_ptt_g's loop body)
        bool end = i >= nsub_max;
        if (!end) {
            float nx = 0.0f, ny = 0.0f, nz = 0.0f;
            const int nd = get_direction_ptt_g<L, false>(P, B, K, S, S.frame + (pass ? 9 : 0), key0, key1,
                                                         (uint32_t)si[6], dc, dx, dy, dz, px, py, pz, &nx, &ny, &nz);
            if (nd == 0) {
                end = true;
            } else {
                dx = nx; dy = ny; dz = nz;
                px = (float)__builtin_fma((double)(dx / 1.0f), sub, (double)px);
                py = (float)__builtin_fma((double)(dy / 1.0f), sub, (double)py);
                pz = (float)__builtin_fma((double)(dz / 1.0f), sub, (double)pz);
                if (p == 0 && ri == 0) {
                    float *o = sl + 3 * (size_t)qi;
                    o[0] = px; o[1] = py; o[2] = pz;
                }
                const int tissue = check_point(P, B.stop, px, py, pz);
                if (tissue == CUSLINES_ENDPOINT || tissue == CUSLINES_INVALIDPOINT || tissue == CUSLINES_OUTSIDEIMAGE) {
                    end = true;
                } else {
                    ++i;
                    if (++ri == step_frac) { ri = 0; ++qi; }
                    end = i >= nsub_max;
                }
            }
        }
        if (!end) continue;
        int nsteps = qi;
        if (ri != 0 && i < step_frac * (MAXL - 1)) {
            ++nsteps;
            if (p == 0) {
                float *o = sl + 3 * (size_t)nsteps;
                o[0] = px; o[1] = py; o[2] = pz;
            }
        }
        const int off = si[4];
        float *cur = sline + (size_t)off * slot_f;
        if (pass == 0) {
            const int nb = nsteps;
            MEM_SYNC();
            for (int j = p; j < nb / 2; j += L) {
                float *a = cur + 3 * (size_t)j, *b = cur + 3 * (size_t)(nb - 1 - j);
                const float a0 = a[0], a1 = a[1], a2 = a[2];
                a[0] = b[0]; a[1] = b[1]; a[2] = b[2];
                b[0] = a0; b[1] = a1; b[2] = a2;
            }
            MEM_SYNC();
            if (p == 0) si[5] = nb;
            pass = 1;
            px = sf[0]; py = sf[1]; pz = sf[2];
            dx = sf[3]; dy = sf[4]; dz = sf[5];
            sl = cur + 3 * (size_t)(nb - 1);
            if (p == 0) { sl[0] = px; sl[1] = py; sl[2] = pz; }
            i = 1; qi = (step_frac == 1) ? 1 : 0; ri = (step_frac == 1) ? 0 : 1;
            WAVE_SYNC();
        } else {
            if (p == 0) sline_len[off] = si[5] - 1 + nsteps;
            have = false;
        }
    }
}

// flat_group_ptt over workgroups of 256 threads (256/L seed groups).
template <int L>
__global__ void __launch_bounds__(256, 3) phase2_ptt_flat_kernel(DevParams P, DevBufs B,
                                                                 const float *__restrict__ seeds, int nseeds,
                                                                 int rng_chunk, const int *__restrict__ offsets,
                                                                 const int *__restrict__ dirs,
                                                                 int *__restrict__ sline_seed,
                                                                 int *__restrict__ sline_len,
                                                                 float *__restrict__ sline,
                                                                 const int *__restrict__ order,
                                                                 int *__restrict__ work) {
    extern __shared__ float smem[];
    PttConst K;
    K.vprop = smem;
    K.vk = smem + 24 * 9;
    ptt_const_init(P, K);
    __syncthreads();
    const int g = threadIdx.x / L;
    float *base = smem + PTTC_FLOATS + g * PTTG_FLOATS;
    PttG S;
    S.frame = base;
    S.vert_pdf = base + 18;
    S.face_cdf = base + 48;
    S.rk = reinterpret_cast<uint32_t *>(base + 80);
    S.rc = reinterpret_cast<int *>(base + 112);
    flat_group_ptt<L>(P, B, K, S, base + 144, seeds, nseeds, rng_chunk, offsets, dirs, sline_seed, sline_len, sline,
                      order, work);
}

// Persistent PTT phase 2 with L lanes per seed.
template <int L>
__global__ void __launch_bounds__(256) phase2_ptt_queue_kernel(DevParams P, DevBufs B, const float *__restrict__ seeds,
                                                               int nseeds, int rng_chunk,
                                                               const int *__restrict__ offsets,
                                                               const int *__restrict__ dirs,
                                                               int *__restrict__ sline_seed,
                                                               int *__restrict__ sline_len, float *__restrict__ sline,
                                                               int *__restrict__ work) {
    extern __shared__ float smem[];
    PttConst K;
    K.vprop = smem;
    K.vk = smem + 24 * 9;
    ptt_const_init(P, K);
    __syncthreads();
    const int g = threadIdx.x / L;
    float *base = smem + PTTC_FLOATS + g * PTTG_FLOATS;
    PttG S;
    S.frame = base;
    S.vert_pdf = base + 18;
    S.face_cdf = base + 48;
    S.rk = reinterpret_cast<uint32_t *>(base + 80);
    S.rc = reinterpret_cast<int *>(base + 112);
    const int p = Grp<L>::lane();
    for (;;) {
        int s = 0;
        if (p == 0) s = atomicAdd(work, 1);
        s = Grp<L>::shfl(s, 0);
        if (s >= nseeds) break;
        track_seed_ptt_g<L>(P, B, K, S, s, seeds, rng_chunk, offsets, dirs, sline_seed, sline_len, sline);
    }
}

// -------------------------------------------------------------------------
// Kernels
// -------------------------------------------------------------------------


// Floats of LDS per seed slot.  Peak marks take nv32/16 words, peak
// candidates nv32/2 floats (uint16 indices).
__host__ __device__ inline int marks_floats(int nv32) { return ((nv32 >> 4) + 3) & ~3; }
__host__ __device__ inline int cand_floats(int nv32) { return ((nv32 >> 1) + 3) & ~3; }

// Bootstrap slot: region C holds H.d and then the ODF; region A holds vox,
// msk and R.d / the model coefficients while the refit runs, and the peak
// marks and candidates once the ODF is computed (A is dead by then).
__host__ __device__ inline int boot_region_a(int nv32, int t32, int hr32) {
    const int refit = t32 + 2 * hr32;
    const int peaks = marks_floats(nv32) + cand_floats(nv32);
    return (refit > peaks) ? refit : peaks;
}

__host__ __device__ inline int slot_floats(int model, bool phase1, int nv32, int t32, int hr32) {
    if (model == CUSLINES_PTT && !phase1) return PTT_SLOT_FLOATS;
    if (model == CUSLINES_PROB && !phase1) return nv32;
    if (model == CUSLINES_PROB || model == CUSLINES_PTT) return nv32 + marks_floats(nv32) + cand_floats(nv32);
    const int c = (nv32 > hr32) ? nv32 : hr32;
    return c + boot_region_a(nv32, t32, hr32);
}

template <int MODEL>
__device__ void bind_slot(const DevParams &P, bool phase1, float *base, SlotLds &S) {
    if (MODEL == CUSLINES_PTT && !phase1) {
        S.ptt.frame = base;
        S.ptt.vert_pdf = base + 32;
        S.ptt.face_cdf = base + 64;
        S.ptt.act = base + 96;
    } else if (MODEL == CUSLINES_PROB || MODEL == CUSLINES_PTT) {
        S.prob.pmf = base;
        S.prob.marks = reinterpret_cast<uint32_t *>(base + P.nv32);
        S.prob.cand = reinterpret_cast<uint16_t *>(base + P.nv32 + marks_floats(P.nv32));
    } else {
        const int c = (P.nv32 > P.hr32) ? P.nv32 : P.hr32;
        float *a = base + c;
        S.boot.h = base;
        S.boot.vox = a;
        S.boot.msk = a + P.t32;
        S.boot.r = a + P.t32 + P.hr32;
        S.boot.marks = reinterpret_cast<uint32_t *>(a);
        S.boot.cand = reinterpret_cast<uint16_t *>(a + marks_floats(P.nv32));
    }
}

// Sphere vertices are staged in LDS for prob and PTT phase 1; bootstrap
// reads them from global memory (L1/L2), leaving its LDS to the seed slots.
template <int MODEL>
__host__ __device__ inline bool verts_in_lds() { return MODEL == CUSLINES_PROB || MODEL == CUSLINES_PTT; }

__device__ __forceinline__ int verts_floats(const DevParams &P) { return (3 * P.nverts + 3) & ~3; }

// getNumStreamlines{Prob,Boot,Ptt}_k (generate_streamlines_cuda.cu:423-462,
// boot.cu:610-718, ptt_init.cu:1-63): initial peak directions per seed.
template <int MODEL>
__global__ void __launch_bounds__(256) phase1_kernel(DevParams P, DevBufs B, const float *__restrict__ seeds,
                                                     int nseeds, int rng_chunk, int *__restrict__ counts,
                                                     int *__restrict__ dirs) {
    extern __shared__ float smem[];
    const float *sv = B.verts;
    int svf = 0;
    if (verts_in_lds<MODEL>()) {
        for (int i = threadIdx.x; i < 3 * P.nverts; i += blockDim.x) smem[i] = B.verts[i];
        sv = smem;
        svf = verts_floats(P);
    }
    __syncthreads();
    const int nslots = blockDim.x >> 5;
    const int slot = threadIdx.x >> 5;
    const int s = blockIdx.x * nslots + slot;
    if (s >= nseeds) return;
    const int l = lane32();
    float *base = smem + svf + slot * slot_floats(MODEL, true, P.nv32, P.t32, P.hr32);
    SlotLds S;
    bind_slot<MODEL>(P, true, base, S);
    const float px = seeds[3 * s], py = seeds[3 * s + 1], pz = seeds[3 * s + 2];
    Philox rng;
    rng.init(((uint64_t)P.seed_hi << 32) | P.seed_lo, (uint64_t)32 * (uint64_t)(s % rng_chunk) + l);
    int n;
    const uint16_t *kept;
    if (MODEL == CUSLINES_PROB) {
        float a, b, c;
        n = get_direction_prob<true>(P, B, sv, S.prob, rng, 0.0f, 0.0f, 0.0f, px, py, pz, &a, &b, &c);
        kept = S.prob.cand;
    } else if (MODEL == CUSLINES_PTT) {
        float *pmf = S.prob.pmf;
        for (int i = l; i < P.nverts; i += 32) {
            const int tx = i & (P.nx - 1), ty = (i >> P.log2x) & (P.ny - 1), tz = i >> (P.log2x + P.log2y);
            const float v = tiled_sample(P, B.data, (float)(tx * P.dimx) + px, (float)(ty * P.dimy) + py,
                                         (float)(tz * P.dimz) + pz, tx, ty, tz);
            pmf[i] = (v < P.ptt_seed_thr_f) ? 0.0f : v;
        }
        WAVE_SYNC();
        n = peak_directions(P, B, pmf, S.prob.marks, S.prob.cand, sv);
        kept = S.prob.cand;
    } else {
        float a, b, c;
        n = get_direction_boot<MODEL>(P, B, sv, S.boot, rng, 1, 0.0f, 0.0f, 0.0f, px, py, pz, &a, &b, &c);
        kept = S.boot.cand;
    }
    int *out = dirs + (size_t)s * P.nverts;
    for (int k = l; k < n; k += 32) out[k] = kept[k];
    if (l == 0) counts[s] = n;
}

// genStreamlinesMerge{Prob,Boot}_k body for one seed (generate_streamlines_cuda.cu:484-621,
// boot.cu:855-979): per direction, backward track, reverse, forward track.
template <int MODEL, int V = 0>
__device__ void track_seed(const DevParams &P, const DevBufs &B, const float *sv, SlotLds &S, int s,
                           const float *__restrict__ seeds, int rng_chunk, const int *__restrict__ offsets,
                           const int *__restrict__ dirs, int *__restrict__ sline_seed, int *__restrict__ sline_len,
                           float *__restrict__ sline) {
    const int l = lane32();
    const float sx = seeds[3 * s], sy = seeds[3 * s + 1], sz = seeds[3 * s + 2];
    Philox rng;
    rng.init(((uint64_t)P.seed_hi << 32) | P.seed_lo, (uint64_t)32 * (uint64_t)(s % rng_chunk) + l + 1);
    const int ndir = (P.slot_cap > 0 && offsets[s + 1] > P.slot_cap) ? 0 : offsets[s + 1] - offsets[s];
    int off = offsets[s];
    const size_t slot_f = (size_t)2 * P.max_sline_len * 3;
    uint32_t ntap = 0, nsub = 0;  // V == 3: this lane's taps, the seed's sub-steps
    for (int i = 0; i < ndir; ++i, ++off) {
        const int vi = dirs[(size_t)s * P.nverts + i];
        const float fx = B.verts[3 * vi], fy = B.verts[3 * vi + 1], fz = B.verts[3 * vi + 2];
        float *cur = sline + (size_t)off * slot_f;
        if (l == 0) sline_seed[off] = s;
        if (MODEL == CUSLINES_PTT) {
            if (!init_frame_ptt<V == 3>(P, B, S.ptt, rng, fx, fy, fz, sx, sy, sz, &ntap)) {
                if (l == 0) {
                    sline_len[off] = 1;
                    cur[0] = sx; cur[1] = sy; cur[2] = sz;
                }
                continue;
            }
        }
        int nb = 0, nf = 0;
        for (int pass = 0; pass < 2; ++pass) {  // backward (-dir), then forward (+dir)
            const float sg = pass ? 1.0f : -1.0f;
            float *frame = (MODEL == CUSLINES_PTT) ? S.ptt.frame + 9 * pass : nullptr;
            float *out = pass ? cur + 3 * (size_t)(nb - 1) : cur;
            const int ns = tracker<MODEL, V>(P, B, sv, S, rng, sx, sy, sz, sg * fx, sg * fy, sg * fz, frame, out,
                                             ((uint32_t)s * 64u + (uint32_t)i) * 2u + (uint32_t)pass, &ntap, &nsub);
            if (pass == 0) {
                nb = ns;
                MEM_SYNC();
                for (int j = l; j < nb / 2; j += 32) {
                    float *a = cur + 3 * (size_t)j, *b = cur + 3 * (size_t)(nb - 1 - j);
                    const float a0 = a[0], a1 = a[1], a2 = a[2];
                    a[0] = b[0]; a[1] = b[1]; a[2] = b[2];
                    b[0] = a0; b[1] = a1; b[2] = a2;
                }
                MEM_SYNC();
            } else {
                nf = ns;
            }
        }
        if (l == 0) sline_len[off] = nb - 1 + nf;
    }
    if (V == 3 && B.stats) {
        unsigned long long taps = ntap;
        for (int m = 16; m; m >>= 1) taps += __shfl_xor(taps, m, 32);
        if (l == 0) {
            atomicAdd(B.stats, taps);
            atomicAdd(B.stats + 1, (unsigned long long)nsub);
        }
    }
}

// The learned kernels (V == 1, 5) copy the proposal network (IL_NPARAM bf16,
// lane-interleaved) into the first MLP_LDS_FLOATS floats of `lds` and point
// Bl.mlp there; every slot of the workgroup reads it from LDS.  Returns the
// floats used.  The caller synchronises the workgroup before use.
template <int V>
__device__ __forceinline__ int stage_network(const DevBufs &B, float *lds, DevBufs &Bl) {
    Bl = B;
    if (V != 1 && V != 5) return 0;
    const uint4 *src = reinterpret_cast<const uint4 *>(B.mlp);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    for (int i = threadIdx.x; i < IL_NPARAM / 8; i += blockDim.x) dst[i] = src[i];
    Bl.mlp = reinterpret_cast<const uint16_t *>(lds);
    return MLP_LDS_FLOATS;
}

// Phase 2 of one 32-lane slot with the stepping loop flattened (the queue
// kernel's V == 4).  In track_seed a slot whose streamline pass, direction
// or seed ends leaves its loops and waits at their exit until the wave's
// other slot leaves them too; here one iteration of a single loop is one
// step (PTT: sub-step) of whichever pass the slot holds, and a slot whose
// pass ends finishes it (last point, reversal, length) and starts its next
// pass, direction or seed at once, so the wave's two slots keep stepping
// together.  Each seed's work, its order and its random draws are those of
// track_seed (genStreamlinesMerge{Prob,Boot,Ptt}, boot.cu:826-980,
// generate_streamlines_cuda.cu:464-621), so the results are identical.
// `st` (FLAT_STATE_FLOATS of the slot's LDS) holds the seed's index,
// direction count, next direction, first slot, current slot and backward
// length (ints), its seed point and the direction (floats).
constexpr int FLAT_STATE_FLOATS = 16;

template <int MODEL, bool LEARNED = false>
__device__ void flat_slot(const DevParams &P, const DevBufs &B, const float *sv, SlotLds &S, float *st,
                          const float *__restrict__ seeds, int nseeds, int rng_chunk, const int *__restrict__ offsets,
                          const int *__restrict__ dirs, int *__restrict__ sline_seed, int *__restrict__ sline_len,
                          float *__restrict__ sline, const int *__restrict__ order, int *__restrict__ work) {
    constexpr bool PTT = MODEL == CUSLINES_PTT;
    const int l = lane32();
    const int xcd = blockIdx.x & (QUEUES - 1);
    const int step_frac = PTT ? P.step_frac : 1, MAXL = P.max_sline_len;
    const int nsub_max = MAXL * step_frac;
    const double sub = P.step_d / (double)step_frac;
    const size_t slot_f = (size_t)2 * MAXL * 3;
    int *si = reinterpret_cast<int *>(st);  // s, ndir, next direction, offsets[s], slot, nb
    float *sf = st + 8;                     // seed point [0..2], direction [3..5]
    if (l == 0) { si[1] = 0; si[2] = 0; }
    WAVE_SYNC();
    Philox rng;
    int qk = 0;
    bool have = false;
    int pass = 0, i = 1, ri = 1, qi = 0;  // step i = qi * step_frac + ri
    float px = 0.0f, py = 0.0f, pz = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
    float *sl = sline;
    // learned proposal (tracker's per-pass state): proposal on, its tries,
    // fallbacks to the plain sub-step, plain sub-steps and envelope misses
    bool learned = LEARNED;
    int ltries = 0, lfalls = 0, nplain = 0;
    uint32_t lmiss = 0;
    for (;;) {
        if (!have) {
            // the next direction of this seed, or of the next seeds
            while (qk < QUEUES) {
                const int di = si[2];
                if (di >= si[1]) {
                    const int qx = (xcd + qk) & (QUEUES - 1);
                    const int lo = (int)((long long)nseeds * qx / QUEUES);
                    const int hi = (int)((long long)nseeds * (qx + 1) / QUEUES);
                    int q = 0;
                    if (l == 0) q = atomicAdd(work + qx, 1);
                    q = shfl32(q, 0);
                    if (lo + q >= hi) { ++qk; continue; }
                    const int s = order ? order[lo + q] : lo + q;
                    const int nd = (P.slot_cap > 0 && offsets[s + 1] > P.slot_cap) ? 0 : offsets[s + 1] - offsets[s];
                    rng.init(((uint64_t)P.seed_hi << 32) | P.seed_lo, (uint64_t)32 * (uint64_t)(s % rng_chunk) + l + 1);
                    WAVE_SYNC();
                    if (l == 0) {
                        si[0] = s; si[1] = nd; si[2] = 0; si[3] = offsets[s];
                        sf[0] = seeds[3 * s]; sf[1] = seeds[3 * s + 1]; sf[2] = seeds[3 * s + 2];
                    }
                    WAVE_SYNC();
                    continue;
                }
                const int s = si[0], off = si[3] + di;
                const float sx = sf[0], sy = sf[1], sz = sf[2];
                const int vi = dirs[(size_t)s * P.nverts + di];
                const float fx = B.verts[3 * vi], fy = B.verts[3 * vi + 1], fz = B.verts[3 * vi + 2];
                float *cur = sline + (size_t)off * slot_f;
                WAVE_SYNC();
                if (l == 0) {
                    sline_seed[off] = s;
                    si[2] = di + 1; si[4] = off;
                    sf[3] = fx; sf[4] = fy; sf[5] = fz;
                }
                WAVE_SYNC();
                if constexpr (PTT) {
                    if (!init_frame_ptt(P, B, S.ptt, rng, fx, fy, fz, sx, sy, sz)) {
                        if (l == 0) {
                            sline_len[off] = 1;
                            cur[0] = sx; cur[1] = sy; cur[2] = sz;
                        }
                        continue;
                    }
                }
                // backward pass (-dir; PTT: on the frame at [0..8])
                pass = 0;
                px = sx; py = sy; pz = sz;
                dx = -fx; dy = -fy; dz = -fz;
                sl = cur;
                if (l == 0) { sl[0] = px; sl[1] = py; sl[2] = pz; }
                i = 1; qi = (step_frac == 1) ? 1 : 0; ri = (step_frac == 1) ? 0 : 1;
                if (LEARNED) { learned = true; ltries = lfalls = nplain = 0; lmiss = 0; }
                have = true;
                break;
            }
            if (!have) break;  // the queues are empty
        }
        // one step of the pass (tracker's loop body)
        bool end = i >= nsub_max;
        if (!end) {
            float nx = 0.0f, ny = 0.0f, nz = 0.0f;
            int nd;
            if constexpr (MODEL == CUSLINES_PROB) {
                nd = get_direction_prob<false>(P, B, sv, S.prob, rng, dx, dy, dz, px, py, pz, &nx, &ny, &nz);
            } else if constexpr (PTT && LEARNED) {
                int fell = 0;
                nd = get_direction_ptt_learned(P, B, S.ptt, S.ptt.frame + (pass ? 9 : 0), rng, learned, dx, dy, dz, px,
                                               py, pz, &nx, &ny, &nz, &fell, &lmiss);
                if (learned) {
                    ++ltries;
                    lfalls += fell;
                    if (ltries >= 16 && 4 * lfalls > ltries) learned = false;
                } else {
                    ++nplain;
                }
            } else if constexpr (PTT) {
                nd = get_direction_ptt<false>(P, B, S.ptt, S.ptt.frame + (pass ? 9 : 0), rng, dx, dy, dz, px, py, pz,
                                              &nx, &ny, &nz);
            } else
                nd = get_direction_boot<MODEL>(P, B, sv, S.boot, rng, 5, dx, dy, dz, px, py, pz, &nx, &ny, &nz);
            if (nd == 0) {
                end = true;
            } else {
                dx = nx; dy = ny; dz = nz;
                px = (float)__builtin_fma((double)(dx / 1.0f), sub, (double)px);
                py = (float)__builtin_fma((double)(dy / 1.0f), sub, (double)py);
                pz = (float)__builtin_fma((double)(dz / 1.0f), sub, (double)pz);
                if (l == 0 && ri == 0) {
                    float *o = sl + 3 * (size_t)qi;
                    o[0] = px; o[1] = py; o[2] = pz;
                }
                const int tissue = check_point(P, B.stop, px, py, pz);
                if (tissue == CUSLINES_ENDPOINT || tissue == CUSLINES_INVALIDPOINT || tissue == CUSLINES_OUTSIDEIMAGE) {
                    end = true;
                } else {
                    ++i;
                    if (++ri == step_frac) { ri = 0; ++qi; }
                    end = i >= nsub_max;
                }
            }
        }
        if (!end) continue;
        // the pass ends at step i (tracker's tail)
        if (LEARNED && B.stats && l == 0) {
            atomicAdd(B.stats, (unsigned long long)(ltries - lfalls));
            atomicAdd(B.stats + 1, (unsigned long long)lfalls);
            atomicAdd(B.stats + 2, (unsigned long long)nplain);
            atomicAdd(B.stats + 3, (unsigned long long)lmiss);
        }
        int nsteps = qi;
        if (ri != 0 && i < step_frac * (MAXL - 1)) {
            ++nsteps;
            if (l == 0) {
                float *o = sl + 3 * (size_t)nsteps;
                o[0] = px; o[1] = py; o[2] = pz;
            }
        }
        const int off = si[4];
        float *cur = sline + (size_t)off * slot_f;
        if (pass == 0) {
            const int nb = nsteps;
            MEM_SYNC();
            for (int j = l; j < nb / 2; j += 32) {
                float *a = cur + 3 * (size_t)j, *b = cur + 3 * (size_t)(nb - 1 - j);
                const float a0 = a[0], a1 = a[1], a2 = a[2];
                a[0] = b[0]; a[1] = b[1]; a[2] = b[2];
                b[0] = a0; b[1] = a1; b[2] = a2;
            }
            MEM_SYNC();
            if (l == 0) si[5] = nb;
            // forward pass (+dir; PTT: on the frame at [9..17]), from the seed point at index nb - 1
            pass = 1;
            px = sf[0]; py = sf[1]; pz = sf[2];
            dx = sf[3]; dy = sf[4]; dz = sf[5];
            sl = cur + 3 * (size_t)(nb - 1);
            if (l == 0) { sl[0] = px; sl[1] = py; sl[2] = pz; }
            i = 1; qi = (step_frac == 1) ? 1 : 0; ri = (step_frac == 1) ? 0 : 1;
            if (LEARNED) { learned = true; ltries = lfalls = nplain = 0; lmiss = 0; }
            WAVE_SYNC();
        } else {
            if (l == 0) sline_len[off] = si[5] - 1 + nsteps;
            have = false;
        }
    }
}

// One seed per 32-lane slot, static assignment.
template <int MODEL, int V = 0>
__global__ void __launch_bounds__(256) phase2_kernel(DevParams P, DevBufs B, const float *__restrict__ seeds,
                                                     int nseeds, int rng_chunk, const int *__restrict__ offsets,
                                                     const int *__restrict__ dirs, int *__restrict__ sline_seed,
                                                     int *__restrict__ sline_len, float *__restrict__ sline) {
    extern __shared__ float smem[];
    const float *sv = B.verts;
    int svf = 0;
    if (MODEL == CUSLINES_PROB) {  // the cone test reads every vertex at every step
        for (int i = threadIdx.x; i < 3 * P.nverts; i += blockDim.x) smem[i] = B.verts[i];
        sv = smem;
        svf = verts_floats(P);
    }
    DevBufs BL;
    svf += stage_network<V>(B, smem + svf, BL);
    __syncthreads();
    const int nslots = blockDim.x >> 5;
    const int slot = threadIdx.x >> 5;
    const int s = blockIdx.x * nslots + slot;
    if (s >= nseeds) return;
    SlotLds S;
    bind_slot<MODEL>(P, false, smem + svf + slot * slot_floats(MODEL, false, P.nv32, P.t32, P.hr32), S);
    track_seed<MODEL, V>(P, BL, sv, S, s, seeds, rng_chunk, offsets, dirs, sline_seed, sline_len, sline);
}

// Persistent work queue: a fixed grid of slots; each 32-lane slot takes the
// next seed from a counter as soon as its previous seed is done, so seeds of
// very uneven length keep every slot busy.  `order` (or the identity) lists
// the seeds in processing order; `work` holds QUEUES zeroed counters.  The
// order changes no result: random streams and output slots follow the seed
// index.
// Plain PTT is held to 128 VGPRs (4 waves per SIMD): the few spills cost less
// than the latency hiding gained (+8% at 200k seeds).  Bootstrap likewise
// (4 waves per SIMD, which its 5 KB of LDS per slot allows: +13% CSA).  The
// flat PTT loop (V == 4) needs about 170 VGPRs and at 128 spills inside its
// loop, so it runs at 3 waves per SIMD (168 VGPRs, no spills); its learned
// form (V == 5) likewise, in 768-thread workgroups (3 waves per SIMD in one
// workgroup, so one staged copy of the network serves 24 slots).
template <int MODEL, int V = 0>
__global__ void __launch_bounds__((MODEL == CUSLINES_PTT && V == 1)   ? 512
                                  : (MODEL == CUSLINES_PTT && V == 5) ? 768
                                                                      : 256,
                                  (MODEL == CUSLINES_PTT && (V == 4 || V == 5)) ? 3
                                  : (MODEL == CUSLINES_PTT && V != 2)           ? 4
                                                                                : (MODEL <= CUSLINES_PROB ? 4 : 1))
    phase2_queue_kernel(DevParams P, DevBufs B, const float *__restrict__ seeds,
                                                           int nseeds, int rng_chunk, const int *__restrict__ offsets,
                                                           const int *__restrict__ dirs, int *__restrict__ sline_seed,
                                                           int *__restrict__ sline_len, float *__restrict__ sline,
                                                           const int *__restrict__ order, int *__restrict__ work) {
    extern __shared__ float smem[];
    const float *sv = B.verts;
    int svf = 0;
    if (MODEL == CUSLINES_PROB) {  // the cone test reads every vertex at every step
        for (int i = threadIdx.x; i < 3 * P.nverts; i += blockDim.x) smem[i] = B.verts[i];
        sv = smem;
        svf = verts_floats(P);
    }
    DevBufs BL;
    svf += stage_network<V>(B, smem + svf, BL);
    __syncthreads();
    const int slot = threadIdx.x >> 5;
    const int l = lane32();
    SlotLds S;
    const int stride =
        slot_floats(MODEL, false, P.nv32, P.t32, P.hr32) + ((V == 4 || V == 5) ? FLAT_STATE_FLOATS : 0);
    bind_slot<MODEL>(P, false, smem + svf + slot * stride, S);
    if constexpr (V == 4 || V == 5) {
        flat_slot<MODEL, V == 5>(P, BL, sv, S, smem + svf + slot * stride + stride - FLAT_STATE_FLOATS, seeds,
                                 nseeds, rng_chunk, offsets, dirs, sline_seed, sline_len, sline, order, work);
        return;
    }
    // Workgroups are dispatched to the 8 XCDs round-robin.  Each XCD first
    // drains its own eighth of the (spatially sorted) seed order, so an XCD's
    // L2 serves one region of the volume, then helps with the others.
    const int xcd = blockIdx.x & (QUEUES - 1);
    for (int k = 0; k < QUEUES; ++k) {
        const int qx = (xcd + k) & (QUEUES - 1);
        const int lo = (int)((long long)nseeds * qx / QUEUES);
        const int hi = (int)((long long)nseeds * (qx + 1) / QUEUES);
        for (;;) {
            int q = 0;
            if (l == 0) q = atomicAdd(work + qx, 1);
            q = shfl32(q, 0);
            if (lo + q >= hi) break;
            const int s = order ? order[lo + q] : lo + q;
            track_seed<MODEL, V>(P, BL, sv, S, s, seeds, rng_chunk, offsets, dirs, sline_seed, sline_len, sline);
        }
    }
}

// -------------------------------------------------------------------------
// Bootstrap phase 2, batched on the matrix cores.
//
// Same algorithm as the scalar path above (get_direction_boot_d,
// boot.cu:393-608; tracker_boot_d, boot.cu:720-824; genStreamlinesMergeBoot_k,
// boot.cu:826-980), organised for MFMA: a 512-thread workgroup holds
// BB_SLOTS = 16 seed slots, one per 32-lane half-wave, that advance in
// lockstep -- every round, each active slot makes one getter attempt.  The
// per-slot work (gather, resampling, normalisation, peaks, closest peak, the
// step and its stop check) runs on the slot's own 32 lanes with the scalar
// path's code; the matrix products of the attempt run batched over the 16
// slots on v_mfma_f32_16x16x4_f32 (fp32 operands, fp32 accumulation):
//   phase B  [H; R] (2 hr16 x hr16) . X (hr16 x 16)    -> H.d, R.d
//   phase D  [dq | -db] (dn16 x 2 hr16) . [xf; msk]    -> model coefficients
//   phase E  S (nv16 x dn16) . C (dn16 x 16)           -> ODF on the sphere
// Each matrix element read from L2 thus serves 16 seeds instead of one.  The
// products sum in another order than the reference's lane-strided dot
// products, so results agree with the scalar path within float rounding:
// peak indices are the same except at near-ties, and a streamline is
// identical up to the first step whose peak choice flips.
// Matrices (uploaded zero padded, row-major): hr_mat [2 hr16][hr16],
// fit_mat [dn16][2 hr16], odf_mat [nv16][dn16].  LDS per slot: the scalar
// slot's regions, the stride padded to 4 mod 32 floats so the 16 slots' rows
// sit in distinct LDS banks for the MFMA operand reads and result writes.
// -------------------------------------------------------------------------
constexpr int BB_SLOTS = 16, BB_THREADS = 512, BB_WAVES = BB_THREADS / 64;
constexpr int BB_MAX_ATTEMPTS = 5;  // NATTEMPTS of tracker_boot_d (boot.cu:757)

typedef float bb_v4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline int bb_slot_stride(int nv32, int t32, int hr32) {
    return ((slot_floats(CUSLINES_CSA, false, nv32, t32, hr32) + 27) & ~31) + 4;
}

// Tiles t = wave, wave + 8, ... (t < ntiles) of Y = A . X over the 16 slots.
// A: global row-major [ntiles * 16][lda], K (<= lda) a multiple of 16, zero
// padded.  Slot n's operand is X[n * ss + k] (k < K), or X2[n * ss + k - ksplit]
// for k >= ksplit; its result rows go to Y[n * ss + row], or Y2[n * ss + row -
// ysplit] for row >= ysplit, rows >= ymax dropped.  Lane l of a wave supplies
// A[l % 16][k] and X_{l % 16}[k] of each step and holds Y rows 4 (l / 16) + r;
// within a 16-wide chunk of K, lane group q = l / 16 reads the contiguous k =
// 16 c + 4 q + (0..3) as one 16-byte load and MFMA step s takes component s.
// wmax (optional): this wave's maximum of each slot's result rows < ymax,
// written to wmax[wave * 16 + slot].
template <int MAXT, bool CSA_ROW0 = false>
__device__ __forceinline__ void bb_gemm(const float *__restrict__ A, int lda, int K, int ntiles, const float *X,
                                        const float *X2, int ksplit, float *Y, float *Y2, int ysplit, int ymax,
                                        int ss, float *wmax = nullptr) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = lane & 15, q = lane >> 4;
    bb_v4 acc[MAXT];
#pragma unroll
    for (int i = 0; i < MAXT; ++i) acc[i] = bb_v4{0.0f, 0.0f, 0.0f, 0.0f};
    const float *ap = A + (size_t)n * lda + 4 * q;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const float *xs = (k0 < ksplit) ? X + n * ss + k0 : X2 + n * ss + (k0 - ksplit);
        const float4 b = *reinterpret_cast<const float4 *>(xs + 4 * q);
        float4 a[MAXT];
#pragma unroll
        for (int i = 0; i < MAXT; ++i) {
            const int t = w + BB_WAVES * i;
            if (t < ntiles) a[i] = *reinterpret_cast<const float4 *>(ap + (size_t)t * 16 * lda + k0);
        }
#pragma unroll
        for (int i = 0; i < MAXT; ++i) {
            const int t = w + BB_WAVES * i;
            if (t < ntiles) {
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].x, b.x, acc[i], 0, 0, 0);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].y, b.y, acc[i], 0, 0, 0);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].z, b.z, acc[i], 0, 0, 0);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].w, b.w, acc[i], 0, 0, 0);
            }
        }
    }
    if (wmax) {
        float pm = REAL_MIN;
#pragma unroll
        for (int i = 0; i < MAXT; ++i) {
            const int t = w + BB_WAVES * i;
            const int row = 16 * t + 4 * q;
            if (t < ntiles) {
                if (row < ymax) pm = fmaxf(pm, acc[i].x);
                if (row + 1 < ymax) pm = fmaxf(pm, acc[i].y);
                if (row + 2 < ymax) pm = fmaxf(pm, acc[i].z);
                if (row + 3 < ymax) pm = fmaxf(pm, acc[i].w);
            }
        }
        pm = fmaxf(pm, __shfl_xor(pm, 16));
        pm = fmaxf(pm, __shfl_xor(pm, 32));
        if (q == 0) wmax[w * 16 + n] = pm;
    }
#pragma unroll
    for (int i = 0; i < MAXT; ++i) {
        const int t = w + BB_WAVES * i;
        if (t < ntiles) {
            int row = 16 * t + 4 * q;
            if (CSA_ROW0 && row == 0) acc[i].x = CSA_C0;  // CSA's constant term (boot.cu:363-368)
            float *y = Y + n * ss;
            if (row >= ysplit) { y = Y2 + n * ss; row -= ysplit; }
            if (row + 3 < ymax) {
                *reinterpret_cast<float4 *>(y + row) = float4{acc[i].x, acc[i].y, acc[i].z, acc[i].w};
            } else {
                if (row < ymax) y[row] = acc[i].x;
                if (row + 1 < ymax) y[row + 1] = acc[i].y;
                if (row + 2 < ymax) y[row + 2] = acc[i].z;
            }
        }
    }
}

// The models' transforms of the signal (gemv_xf) with the hardware base-2
// logarithm: the batched kernel's results already differ from the
// reference's in the last bits, and two series logarithms per channel (CSA)
// are a large share of its per-slot work.
template <int XF>
__device__ __forceinline__ float bb_xf(float x) {
    constexpr float LN2 = 0.693147180559945309f;
    if (XF == 1) {
        const float lx = __builtin_amdgcn_logf(x) * LN2;
        return ((-lx) * (1.5f + lx)) * x;
    }
    const float c = fminf(fmaxf(x, 0.001f), 0.999f);
    return __builtin_amdgcn_logf(-(__builtin_amdgcn_logf(c) * LN2)) * LN2;
}

// peak_directions_d on the ODF thresholded at `thr` (values below it read as
// 0, tracking_helpers.cu:75-245 after generate_streamlines_cuda.cu's
// threshold) without writing the thresholded array: the adjacency form of
// peak_directions with the threshold applied to every value it reads.  A
// peak is never below thr (its thresholded value would be 0 with no
// strictly smaller neighbour), so the kept candidates' values, the sort
// keys, are the raw ODF values.  Needs the adjacency table (B.nbr).
__device__ int bb_peaks(const DevParams &P, const DevBufs &B, const float *odf, float thr, uint16_t *cand,
                        const float *sv) {
    const int l = lane32();
    const int nv = P.nverts;
    uint32_t pk = 0u;
    float mn = REAL_MAX, m = REAL_MIN;
    if (thr > 0.0f) {
        // With thr > 0 the thresholded test reduces to the raw values: a
        // vertex whose value x is below thr (or NaN) reads as 0 and has no
        // strictly smaller thresholded neighbour, so it is no peak; for x >=
        // thr, a neighbour y is smaller after thresholding iff y < x and
        // larger iff y > x.  So x is a peak iff x >= thr, the minimum of its
        // raw neighbours is below x and their maximum is not above it
        // (fminf / fmaxf pass over a NaN neighbour as the comparisons do).
        // The largest peak value is kept on the way: max(x + c) over the
        // peaks is max(x) + c, rounding being monotone.
        float pmax = REAL_MIN;
        for (int r0 = 0, j0 = l; j0 < nv; r0 += 4, j0 += 128) {
            uint4 qs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = min(j0 + 32 * u, nv - 1);  // past the end: re-read a valid row, result dropped
                qs[u] = reinterpret_cast<const uint4 *>(B.nbr)[j];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + 32 * u;
                if (j >= nv) break;
                const uint32_t w[4] = {qs[u].x, qs[u].y, qs[u].z, qs[u].w};
                const float x = odf[j];
                const float tx = (x < thr) ? 0.0f : x;
                mn = (mn < tx) ? mn : tx;
                float lo = odf[w[0] & 0xffffu], hi = lo;
#pragma unroll
                for (int k = 1; k < 8; ++k) {
                    const float y = odf[(w[k >> 1] >> (16 * (k & 1))) & 0xffffu];
                    lo = fminf(lo, y);
                    hi = fmaxf(hi, y);
                }
                if (x >= thr && lo < x && !(hi > x)) {
                    pk |= 1u << (r0 + u);
                    pmax = fmaxf(pmax, x);
                }
            }
        }
        mn = min32(mn);
        mn = (0.0f > mn) ? 0.0f : mn;
        pmax = max32(pmax);
        if (pmax != REAL_MIN) m = pmax + (-mn);
    } else {
        for (int r0 = 0, j0 = l; j0 < nv; r0 += 4, j0 += 128) {
            uint4 qs[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = min(j0 + 32 * u, nv - 1);
                qs[u] = reinterpret_cast<const uint4 *>(B.nbr)[j];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + 32 * u;
                if (j >= nv) break;
                const uint32_t w[4] = {qs[u].x, qs[u].y, qs[u].z, qs[u].w};
                float x = odf[j];
                x = (x < thr) ? 0.0f : x;
                mn = (mn < x) ? mn : x;
                bool smaller = false, larger = false;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    float y = odf[(w[k >> 1] >> (16 * (k & 1))) & 0xffffu];
                    y = (y < thr) ? 0.0f : y;
                    smaller = smaller || (y < x);
                    larger = larger || (x < y);
                }
                if (smaller && !larger) pk |= 1u << (r0 + u);
            }
        }
        mn = min32(mn);
        mn = (0.0f > mn) ? 0.0f : mn;
        for (int r = 0, j = l; j < nv; ++r, j += 32)
            if ((pk >> r) & 1u) {
                const float xm = odf[j] + (-mn);
                m = (m > xm) ? m : xm;
            }
        m = max32(m);
    }
    const float comp = (float)((double)P.rel_peak * (double)m);
    int n = 0;
    for (int r = 0, j0 = 0; j0 < nv; ++r, j0 += 32) {
        const int j = j0 + l;
        const bool keep = (j < nv) && ((pk >> r) & 1u) && ((odf[j] - mn) >= comp);
        const uint32_t msk = half_ballot(keep);
        if (keep) cand[n + __popc(msk & ((1u << l) - 1u))] = (uint16_t)j;
        n += __popc(msk);
    }
    WAVE_SYNC();
    return sort_dedup_peaks(P, odf, cand, sv, n);
}

// Per-slot state of the batched kernel (registers of the slot's 32 lanes).
struct BbSlot {
    int s;          // seed (-1: none yet)
    int qk;         // queue cursor: queues qk.. of this XCD's order still to drain
    int di, ndir;   // direction of the seed being tracked, directions of the seed
    int off;        // streamline index of direction di
    int pass;       // 0 backward, 1 forward
    int i, att;     // step index of the pass, attempt of the step
    int nb;         // backward points (pass 1)
    bool done;      // queues drained
    float sx, sy, sz, fx, fy, fz, px, py, pz, dx, dy, dz;
};

// Start pass `pass` of direction st.di: the tracker's loop head
// (boot.cu:738-745): the point at index 0 of the pass is the seed.
__device__ __forceinline__ void bb_start_pass(const DevParams &P, BbSlot &st, float *__restrict__ sline) {
    const float sg = st.pass ? 1.0f : -1.0f;
    st.px = st.sx; st.py = st.sy; st.pz = st.sz;
    st.dx = sg * st.fx; st.dy = sg * st.fy; st.dz = sg * st.fz;
    st.i = 1;
    st.att = 0;
    float *o = sline + (size_t)st.off * (2 * (size_t)P.max_sline_len * 3) + (st.pass ? 3 * (size_t)(st.nb - 1) : 0);
    if (lane32() == 0) { o[0] = st.px; o[1] = st.py; o[2] = st.pz; }
}

// Next seed with at least one direction from the XCD-ordered queues, or
// st.done.  Streamlines of seeds without directions need no work; in a
// bounded chunk (P.slot_cap) neither do seeds whose slots end past the cap.
__device__ void bb_next_seed(const DevParams &P, const DevBufs &B, BbSlot &st, Philox &rng, int nseeds,
                             const float *__restrict__ seeds, int rng_chunk, const int *__restrict__ offsets,
                             const int *__restrict__ dirs, const int *__restrict__ order, int *__restrict__ work,
                             float *__restrict__ sline) {
    const int l = lane32();
    const int xcd = blockIdx.x & (QUEUES - 1);
    while (st.qk < QUEUES) {
        const int qx = (xcd + st.qk) & (QUEUES - 1);
        const int lo = (int)((long long)nseeds * qx / QUEUES);
        const int hi = (int)((long long)nseeds * (qx + 1) / QUEUES);
        int q = 0;
        if (l == 0) q = atomicAdd(work + qx, 1);
        q = shfl32(q, 0);
        if (lo + q >= hi) { ++st.qk; continue; }
        const int s = order ? order[lo + q] : lo + q;
        const int nd = (P.slot_cap > 0 && offsets[s + 1] > P.slot_cap) ? 0 : offsets[s + 1] - offsets[s];
        if (nd == 0) continue;
        st.s = s;
        st.ndir = nd;
        st.di = 0;
        st.off = offsets[s];
        st.sx = seeds[3 * s]; st.sy = seeds[3 * s + 1]; st.sz = seeds[3 * s + 2];
        rng.init(((uint64_t)P.seed_hi << 32) | P.seed_lo, (uint64_t)32 * (uint64_t)(s % rng_chunk) + l + 1);
        const int vi = dirs[(size_t)s * P.nverts];
        st.fx = B.verts[3 * vi]; st.fy = B.verts[3 * vi + 1]; st.fz = B.verts[3 * vi + 2];
        st.pass = 0;
        st.nb = 0;
        return;
    }
    st.done = true;
}

// A pass ended with `ns` points: reverse the backward half in place, start
// the forward pass, or close the streamline and move to the next direction
// or seed (genStreamlinesMergeBoot_k, boot.cu:855-979).
__device__ void bb_end_pass(const DevParams &P, const DevBufs &B, BbSlot &st, Philox &rng, int ns, int nseeds,
                            const float *__restrict__ seeds, int rng_chunk, const int *__restrict__ offsets,
                            const int *__restrict__ dirs, const int *__restrict__ order, int *__restrict__ work,
                            int *__restrict__ sline_seed, int *__restrict__ sline_len, float *__restrict__ sline) {
    const int l = lane32();
    const size_t slot_f = 2 * (size_t)P.max_sline_len * 3;
    float *cur = sline + (size_t)st.off * slot_f;
    if (st.pass == 0) {
        st.nb = ns;
        MEM_SYNC();
        for (int j = l; j < ns / 2; j += 32) {
            float *a = cur + 3 * (size_t)j, *b = cur + 3 * (size_t)(ns - 1 - j);
            const float a0 = a[0], a1 = a[1], a2 = a[2];
            a[0] = b[0]; a[1] = b[1]; a[2] = b[2];
            b[0] = a0; b[1] = a1; b[2] = a2;
        }
        MEM_SYNC();
        st.pass = 1;
        bb_start_pass(P, st, sline);
        return;
    }
    if (l == 0) {
        sline_seed[st.off] = st.s;
        sline_len[st.off] = st.nb - 1 + ns;
    }
    ++st.di;
    ++st.off;
    if (st.di < st.ndir) {
        const int vi = dirs[(size_t)st.s * P.nverts + st.di];
        st.fx = B.verts[3 * vi]; st.fy = B.verts[3 * vi + 1]; st.fz = B.verts[3 * vi + 2];
        st.pass = 0;
        st.nb = 0;
        bb_start_pass(P, st, sline);
        return;
    }
    bb_next_seed(P, B, st, rng, nseeds, seeds, rng_chunk, offsets, dirs, order, work, sline);
    if (!st.done) bb_start_pass(P, st, sline);
}

// OCC: waves per SIMD the kernel is compiled for (2: 256 VGPRs and one
// workgroup per CU; 4: 128 VGPRs with spills and two workgroups per CU, the
// default: 99.8 against 71.9 M steps/s for CSA at 1M seeds when the two were
// compared, before the min/max peak test and the low-rank form, as the two
// workgroups' lockstep phases overlap).
// LR: the low-rank form of the [H; R] product (P.rk16 > 0), a separate
// instantiation so that each form's kernel carries only its own registers.
template <int MODEL, int OCC, bool LR>
__global__ void __launch_bounds__(BB_THREADS, OCC)
    phase2_boot_batch_kernel(DevParams P, DevBufs B, const float *__restrict__ seeds, int nseeds, int rng_chunk,
                             const int *__restrict__ offsets, const int *__restrict__ dirs,
                             int *__restrict__ sline_seed, int *__restrict__ sline_len, float *__restrict__ sline,
                             const int *__restrict__ order, int *__restrict__ work) {
    extern __shared__ float smem[];
    const int ss = bb_slot_stride(P.nv32, P.t32, P.hr32);
    const int slot = threadIdx.x >> 5;
    const int l = lane32();
    SlotLds SL;
    bind_slot<MODEL>(P, false, smem + slot * ss, SL);
    BootLds &S = SL.boot;
    // Slot n's arrays are the same offsets from smem + n * ss.
    const int off_h = 0;
    const int off_vox = (int)(S.vox - smem) - slot * ss, off_msk = (int)(S.msk - smem) - slot * ss,
              off_r = (int)(S.r - smem) - slot * ss;
    const int hr = P.hr, T = P.dimt, nv = P.nverts, hr16 = P.hr16;
    const int MAXL = P.max_sline_len;
    float *wmax = smem + BB_SLOTS * ss;  // [BB_WAVES][16] partial ODF maxima of phase E
    BbSlot st;
    st.s = -1; st.qk = 0; st.done = false; st.nb = 0; st.pass = 0;
    Philox rng;
    bb_next_seed(P, B, st, rng, nseeds, seeds, rng_chunk, offsets, dirs, order, work, sline);
    if (!st.done) bb_start_pass(P, st, sline);
    unsigned long long attempts = 0, rounds = 0;  // slot attempts and rounds of this workgroup
    for (;;) {
        // ---- phase A: gather the signal at the point -----------------------
        const bool active = !st.done;
        bool data = false;
        if (active) {
            data = trilinear<3>(P, B.data, P.t32, T, st.px, st.py, st.pz, S.vox) == 0;
            WAVE_SYNC();
            if (data) {
                for (int m = l; m < hr16; m += 32) S.msk[m] = (m < hr) ? S.vox[B.dwi_idx[m]] : 0.0f;
            }
        }
        const int nact = __syncthreads_count(active) / 32;  // active slots (32 lanes each)
        if (nact == 0) break;
        attempts += (unsigned long long)nact;
        ++rounds;
        // ---- phase B: H.d and R.d for the 16 slots --------------------------
        if (LR) {
            // low-rank form: z = U^T d into the R.d region, then H.d = U z;
            // R.d follows in phase C
            bb_gemm<1>(B.ut_mat, hr16, hr16, P.rk16 / 16, smem + off_msk, smem + off_msk, hr16, smem + off_r,
                       smem + off_r, 1 << 30, P.rk16, ss);
            __syncthreads();
            bb_gemm<(320 / 16 + BB_WAVES - 1) / BB_WAVES>(B.u_mat, P.rk16, P.rk16, hr16 / 16, smem + off_r,
                                                          smem + off_r, 1 << 30, smem + off_h, smem + off_h, 1 << 30,
                                                          hr16, ss);
        } else {
            bb_gemm<(2 * 320 / 16 + BB_WAVES - 1) / BB_WAVES>(B.hr_mat, hr16, hr16, 2 * hr16 / 16, smem + off_msk,
                                                              smem + off_msk, hr16, smem + off_h, smem + off_r, hr16,
                                                              2 * hr16, ss);
        }
        __syncthreads();
        // ---- phase C: residual bootstrap, normalisation, model transform -----
        if (LR && active && data) {
            // R.d = v - mean(v), v = (d - H.d) / sqrt(1 - H_ii)
            float part = 0.0f;
            for (int j = l; j < hr; j += 32) {
                const float v = (S.msk[j] - S.h[j]) * B.invlev[j];
                S.r[j] = v;
                part = part + v;
            }
            const float mean = tree_sum32(part) / (float)hr;
            for (int j = l; j < hr; j += 32) S.r[j] = S.r[j] - mean;
            WAVE_SYNC();
        }
        if (active && data) {
            for (int j = l; j < hr; j += 32) {
                const int src = (int)(rng.next() % (uint32_t)hr);
                S.h[j] = S.h[j] + S.r[src];
            }
            WAVE_SYNC();
            for (int m = l; m < hr; m += 32) S.vox[B.dwi_idx[m]] = S.h[m];
            WAVE_SYNC();
            float part = 0.0f;
            int cnt = 0;
            for (int j = l; j < T; j += 32) {
                float v = S.vox[j];
                v = (P.min_signal > v) ? P.min_signal : v;
                S.vox[j] = v;
                if (B.b0[j]) { ++cnt; part = part + v; }
            }
            const float sum = tree_sum32(part);
            for (int m = 16; m; m >>= 1) cnt += shfl_xor32(cnt, m);
            const float inv = (float)cnt / sum;
            for (int j = l; j < T; j += 32) S.vox[j] = S.vox[j] * inv;
            WAVE_SYNC();
            for (int m = l; m < hr16; m += 32) {
                const float v = (m < hr) ? S.vox[B.dwi_idx[m]] : 0.0f;
                S.msk[m] = v;
                S.h[m] = (m < hr) ? bb_xf<(MODEL == CUSLINES_OPDT) ? 1 : 2>(v) : 0.0f;
            }
        }
        __syncthreads();
        // ---- phase D: model coefficients (CSA: dq.xf; OPDT: dq.xf - db.msk) --
        bb_gemm<1, MODEL == CUSLINES_CSA>(B.fit_mat, 2 * hr16, (MODEL == CUSLINES_OPDT) ? 2 * hr16 : hr16,
                                           P.dn16 / 16, smem + off_h, smem + off_msk, hr16, smem + off_vox,
                                           smem + off_vox, 1 << 30, P.dn16, ss);
        __syncthreads();
        // ---- phase E: the ODF on the sphere ---------------------------------
        bb_gemm<(768 / 16 + BB_WAVES - 1) / BB_WAVES>(B.odf_mat, P.dn16, P.dn16, P.nv16 / 16, smem + off_vox,
                                                      smem + off_vox, 1 << 30, smem + off_h, smem + off_h, 1 << 30,
                                                      nv, ss, wmax);
        __syncthreads();
        // ---- phase F: peaks, closest peak, the step --------------------------
        if (active) {
            int ns = -1;  // >= 0: the pass ends with ns points
            if (!data) {
                // the signal is zero at this point: every remaining attempt
                // fails the same way without a random draw
                ns = st.i;
            } else {
                // the ODF's maximum from the 8 waves' partial maxima of phase E
                float mx = (l < BB_WAVES) ? wmax[l * 16 + slot] : REAL_MIN;
                mx = max32(mx);
                const float thr = (float)(PMF_THRESHOLD_P * (double)mx);
                int npk;
                if (B.nbr) {
                    npk = bb_peaks(P, B, S.h, thr, S.cand, B.verts);
                } else {
                    for (int j = l; j < nv; j += 32)
                        if (S.h[j] < thr) S.h[j] = 0.0f;
                    WAVE_SYNC();
                    npk = peak_directions(P, B, S.h, S.marks, S.cand, B.verts);
                }
                float nx = 0.0f, ny = 0.0f, nz = 0.0f;
                const bool ok = npk > 0 && closest_peak(P, B.verts, st.dx, st.dy, st.dz, npk, S.cand, &nx, &ny, &nz);
                WAVE_SYNC();
                if (!ok) {
                    if (++st.att == BB_MAX_ATTEMPTS) ns = st.i;
                } else {
                    st.att = 0;
                    st.dx = nx; st.dy = ny; st.dz = nz;
                    st.px = (float)__builtin_fma((double)(nx / 1.0f), P.step_d, (double)st.px);
                    st.py = (float)__builtin_fma((double)(ny / 1.0f), P.step_d, (double)st.py);
                    st.pz = (float)__builtin_fma((double)(nz / 1.0f), P.step_d, (double)st.pz);
                    if (l == 0) {
                        float *o = sline + (size_t)st.off * (2 * (size_t)MAXL * 3) +
                                   (st.pass ? 3 * (size_t)(st.nb - 1) : 0) + 3 * (size_t)st.i;
                        o[0] = st.px; o[1] = st.py; o[2] = st.pz;
                    }
                    const int tissue = check_point(P, B.stop, st.px, st.py, st.pz);
                    if (tissue == CUSLINES_ENDPOINT || tissue == CUSLINES_INVALIDPOINT ||
                        tissue == CUSLINES_OUTSIDEIMAGE)
                        ns = st.i;
                    else if (++st.i == MAXL)
                        ns = MAXL;
                }
            }
            if (ns >= 0)
                bb_end_pass(P, B, st, rng, ns, nseeds, seeds, rng_chunk, offsets, dirs, order, work, sline_seed,
                            sline_len, sline);
        }
        // no barrier here: phase A writes only this slot's region A, which
        // phase E of every wave has finished reading (barrier above)
    }
    if (threadIdx.x == 0 && B.stats) {
        atomicAdd(B.stats, attempts);
        atomicAdd(B.stats + 1, rounds);
    }
}

__global__ void zero_ints(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <typename TO>
__global__ void zero_scalar(TO *p) {
    *p = 0;
}

// Spatial processing order of the persistent queue: seeds bucketed by the
// Morton code of their 2^brick_log2-voxel brick (counting sort).
__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // bits 0..9 to every third bit
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void seed_keys(DevParams P, const float *__restrict__ seeds, int n, int brick_log2, int *__restrict__ key,
                          int *__restrict__ count) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t c[3];
    const int dims[3] = {P.dimx, P.dimy, P.dimz};
    for (int a = 0; a < 3; ++a) {
        const float f = floorf(seeds[3 * s + a]);
        const int i = (f != f || f < 0.0f) ? 0 : (f > (float)(dims[a] - 1) ? dims[a] - 1 : (int)f);
        c[a] = (uint32_t)i >> brick_log2;
    }
    const int k = (int)(spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2));
    key[s] = k;
    atomicAdd(count + k, 1);
}

__global__ void seed_scatter(const int *__restrict__ key, int n, const int *__restrict__ start, int *__restrict__ cursor,
                             int *__restrict__ order) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int k = key[s];
    order[start[k] + atomicAdd(cursor + k, 1)] = s;
}

// -------------------------------------------------------------------------
// Training of the proposal network (no counterpart in the reference).  Plain
// per-element kernels: the network is 26,920 parameters and a batch of a few
// thousand pairs, so a training step is microseconds of arithmetic.
// -------------------------------------------------------------------------
__device__ __forceinline__ float bf16_round(float x) {  // to the nearest bf16, ties to even
    uint32_t b = __float_as_uint(x);
    if ((b & 0x7f800000u) == 0x7f800000u) return x;
    b += 0x7fffu + ((b >> 16) & 1u);
    return __uint_as_float(b & 0xffff0000u);
}

// out[s][j] = act(b[j] + sum_k W[j][k] in[s][k]), parameters rounded to bf16.
template <int K, int N, bool RELU>
__global__ void mlp_fwd_layer(const float *__restrict__ prm, int w_off, int b_off, const float *__restrict__ in,
                              int in_ld, int n, float *__restrict__ out) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * N) return;
    const int smp = (int)(idx / N), j = (int)(idx % N);
    const float *w = prm + w_off + (size_t)j * K;
    const float *x = in + (size_t)smp * in_ld;
    float acc = bf16_round(prm[b_off + j]);
    for (int k = 0; k < K; ++k) acc = __builtin_fmaf(bf16_round(w[k]), x[k], acc);
    out[idx] = RELU ? ((acc > 0.0f) ? acc : 0.0f) : acc;
}

// Per pair: squared error of the 24 predicted supports, averaged over the
// 24, and its gradient with respect to the outputs divided by n.
__global__ void mlp_loss_out(const float *__restrict__ O, const float *__restrict__ pairs, int n,
                             float *__restrict__ dO, float *__restrict__ loss) {
    const int smp = blockIdx.x * blockDim.x + threadIdx.x;
    if (smp >= n) return;
    const float *o = O + (size_t)smp * MLP_OUT;
    const float *t = pairs + (size_t)smp * CUSLINES_PAIR_FLOATS + MLP_IN;
    float *g = dO + (size_t)smp * MLP_OUT;
    const float sc = 2.0f / ((float)MLP_OUT * (float)n);
    float acc = 0.0f;
    for (int j = 0; j < MLP_OUT; ++j) {
        const float e = o[j] - t[j];
        acc = __builtin_fmaf(e, e, acc);
        g[j] = e * sc;
    }
    loss[smp] = acc / (float)MLP_OUT;
}

// din[s][k] = (sum_j W[j][k] dout[s][j]) where hin[s][k] > 0, else 0.
template <int N, int K>
__global__ void mlp_bwd_input(const float *__restrict__ prm, int w_off, const float *__restrict__ dout,
                              const float *__restrict__ hin, int n, float *__restrict__ din) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * K) return;
    const int smp = (int)(idx / K), k = (int)(idx % K);
    const float *g = dout + (size_t)smp * N;
    float acc = 0.0f;
    for (int j = 0; j < N; ++j) acc = __builtin_fmaf(bf16_round(prm[w_off + (size_t)j * K + k]), g[j], acc);
    din[idx] = (hin[idx] > 0.0f) ? acc : 0.0f;
}

// grad W[j][k] = sum_s dout[s][j] in[s][k]; column k == K is the bias.
// Pass 1: blockIdx.y takes samples [y * WGRAD_CHUNK, (y + 1) * WGRAD_CHUNK)
// and writes its partial sums to part[y][N * (K + 1)]; pass 2 adds the
// partials in chunk order (deterministic).
constexpr int WGRAD_CHUNK = 128;
template <int N, int K>
__global__ void mlp_wgrad_part(const float *__restrict__ dout, const float *__restrict__ in, int in_ld, int n,
                               float *__restrict__ part) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * (K + 1)) return;
    const int j = idx / (K + 1), k = idx % (K + 1);
    const int s0 = blockIdx.y * WGRAD_CHUNK, s1 = min(n, s0 + WGRAD_CHUNK);
    float acc = 0.0f;
    if (k == K) {
        for (int smp = s0; smp < s1; ++smp) acc += dout[(size_t)smp * N + j];
    } else {
        for (int smp = s0; smp < s1; ++smp)
            acc = __builtin_fmaf(dout[(size_t)smp * N + j], in[(size_t)smp * in_ld + k], acc);
    }
    part[(size_t)blockIdx.y * (N * (K + 1)) + idx] = acc;
}

template <int N, int K>
__global__ void mlp_wgrad_sum(const float *__restrict__ part, int nchunk, float *__restrict__ grad, int w_off,
                              int b_off) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * (K + 1)) return;
    float acc = 0.0f;
    for (int c = 0; c < nchunk; ++c) acc += part[(size_t)c * (N * (K + 1)) + idx];
    const int j = idx / (K + 1), k = idx % (K + 1);
    if (k == K) grad[b_off + j] = acc;
    else grad[w_off + (size_t)j * K + k] = acc;
}

__global__ void __launch_bounds__(256) mean_kernel(const float *__restrict__ x, int n, float *__restrict__ out) {
    __shared__ float part[256];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) acc += x[i];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0] / (float)n;
}

// Adam (Kingma and Ba, 2015); c1 = 1 - beta1^step, c2 = 1 - beta2^step.
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (__builtin_sqrtf(vi / c2) + eps);
}

// Propagators of arc length FEAT_STEPS probe steps for the 24 disc vertices
// (the curves along which the learned proposal's inputs are tapped).
// Per disc vertex, the propagator of `steps` probe steps (FEAT_STEPS: the
// learned proposal's features; 1: the vertex phase's probes, followed at
// VPROP_K by the vertices' curvatures (k1, k2)).
__global__ void feature_propagators(DevParams P, int steps, float *__restrict__ out) {
    const int l = threadIdx.x;
    if (l < DISC_VERT_CNT) {
        const float k1 = (float)((double)c_disc_vert[2 * l] * (double)P.max_curv);
        const float k2 = (float)((double)c_disc_vert[2 * l + 1] * (double)P.max_curv);
        float p[9];
        prepare_propagator(k1, k2, (steps == 1) ? P.probe_step : (float)steps * P.probe_step, p);
        for (int i = 0; i < 9; ++i) out[9 * l + i] = p[i];
        if (steps == 1) {  // and the vertex's curvature, the disc scaled by MAX_CURVATURE
            out[VPROP_K + 2 * l] = k1;
            out[VPROP_K + 2 * l + 1] = k2;
        }
    }
}

// PTT corner blocks: for texel cell (cx,cy,cz) in [-1, D-1]^3 and direction
// rank r (vertex rank_vertex[r]), the 8 texels one linear-filtered tap of
// that vertex's tile reads (x fastest), with the tiling's bleed into
// neighbouring tiles and the clamp at the texture edge.  Block index =
// cell * nverts + r.
__global__ void build_corner_blocks(DevParams P, const float *__restrict__ odfv,
                                    const uint16_t *__restrict__ rank_vertex, float *__restrict__ blocks,
                                    long long ncells) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncells) return;
    const int DX = P.dimx, DY = P.dimy, DZ = P.dimz;
    long long t = idx;
    const int v = (int)rank_vertex[t % P.nverts]; t /= P.nverts;
    const int cz = (int)(t % (DZ + 1)) - 1; t /= (DZ + 1);
    const int cy = (int)(t % (DY + 1)) - 1; t /= (DY + 1);
    const int cx = (int)(t % (DX + 1)) - 1;
    const int tx = v & (P.nx - 1), ty = (v >> P.log2x) & (P.ny - 1), tz = v >> (P.log2x + P.log2y);
    int txa, txb, lxa, lxb, tya, tyb, lya, lyb, tza, tzb, lza, lzb;
    axis_split(tx * DX + cx, DX, P.nx, tx, &txa, &lxa);
    axis_split(tx * DX + cx + 1, DX, P.nx, tx, &txb, &lxb);
    axis_split(ty * DY + cy, DY, P.ny, ty, &tya, &lya);
    axis_split(ty * DY + cy + 1, DY, P.ny, ty, &tyb, &lyb);
    axis_split(tz * DZ + cz, DZ, P.nz, tz, &tza, &lza);
    axis_split(tz * DZ + cz + 1, DZ, P.nz, tz, &tzb, &lzb);
    const int NV = P.nverts, nx = P.nx, ny = P.ny;
    auto fetch = [&](int ti, int li, int tj, int lj, int tk, int lk) -> float {
        const int w = ti + nx * (tj + ny * tk);
        if (w >= NV) return 0.0f;
        return odfv[(((size_t)w * DX + li) * DY + lj) * DZ + lk];
    };
    float4 lo, hi;
    lo.x = fetch(txa, lxa, tya, lya, tza, lza);
    lo.y = fetch(txb, lxb, tya, lya, tza, lza);
    lo.z = fetch(txa, lxa, tyb, lyb, tza, lza);
    lo.w = fetch(txb, lxb, tyb, lyb, tza, lza);
    hi.x = fetch(txa, lxa, tya, lya, tzb, lzb);
    hi.y = fetch(txb, lxb, tya, lya, tzb, lzb);
    hi.z = fetch(txa, lxa, tyb, lyb, tzb, lzb);
    hi.w = fetch(txb, lxb, tyb, lyb, tzb, lzb);
    float4 *out = reinterpret_cast<float4 *>(blocks + (size_t)idx * 8);
    out[0] = lo;
    out[1] = hi;
}

// PttDirectionGetter.prepare_data (cu_direction_getters.py:531-625) on the
// device, one thread per voxel, in the host recipe's order: clip at 0 (NaN
// kept and -0.0 made +0.0, as numpy's clip does), zero the voxel where the stop value is
// below the threshold, then divide by the voxel's maximum when it is > 0 (a
// NaN maximum, as numpy's max propagates it, leaves the voxel unchanged).
__global__ void ptt_normalise(float *__restrict__ odf, const float *__restrict__ stop, float thr, long long nvox,
                              int T) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    float *o = odf + (size_t)v * T;
    const bool zero = stop[v] < thr;
    float mx = 0.0f;
    bool first = true;
    for (int t = 0; t < T; ++t) {
        float x = o[t];
        x = zero ? 0.0f : (x > 0.0f || x != x) ? x : 0.0f;  // -0.0 becomes +0.0, as in numpy
        o[t] = x;
        if (first || x != x || x > mx) { if (!(mx != mx)) mx = x; }
        first = false;
    }
    if (mx > 0.0f)
        for (int t = 0; t < T; ++t) o[t] = o[t] / mx;
}

// Sets *flag when some value of x[0..n) is not >= 0 (negative or NaN).
__global__ void flag_not_nonneg(const float *__restrict__ x, long long n, int *__restrict__ flag) {
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        bad |= !(x[i] >= 0.0f);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// Nearest sphere vertex of every point of the linspace(-1, 1, res)^3 grid
// ([i0][i1][i2], i2 fastest), as PttDirectionGetter.prepare_data's KDTree
// query (cu_direction_getters.py:601-623): grid values i * (2 / (res - 1)) - 1
// in double with the last one exactly 1 (numpy's linspace), squared
// Euclidean distances in double, the first of equal minima; with sphere_symm
// the antipodes are candidates too, after all the vertices.
__global__ void ptt_lut_kernel(const float *__restrict__ verts, int nv, int res, int symm, int32_t *__restrict__ lut) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)res * res * res;
    if (idx >= n) return;
    const int i2 = (int)(idx % res), i1 = (int)((idx / res) % res), i0 = (int)(idx / ((long long)res * res));
    const double step = 2.0 / (double)(res - 1);
    auto coord = [&](int i) -> double {
        if (i == res - 1) return 1.0;
        const double y = (double)i * step;
        return y + (-1.0);
    };
    const double px = coord(i0), py = coord(i1), pz = coord(i2);
    double best = 0.0;
    int bi = -1;
    for (int pass = 0; pass < (symm ? 2 : 1); ++pass) {
        const double sg = pass ? -1.0 : 1.0;
        for (int v = 0; v < nv; ++v) {
            const double dx = px - sg * (double)verts[3 * v], dy = py - sg * (double)verts[3 * v + 1],
                         dz = pz - sg * (double)verts[3 * v + 2];
            const double d = (dx * dx + dy * dy) + dz * dz;
            if (bi < 0 || d < best) { best = d; bi = v; }
        }
    }
    lut[idx] = bi;
}

// PTT data: [X][Y][Z][T] -> vertex-major [T][X][Y][Z] (32x32 LDS tiles).
__global__ void transpose_vox_t(const float *__restrict__ src, float *__restrict__ dst, long long nvox, int T) {
    __shared__ float tile[32][33];
    const long long v0 = (long long)blockIdx.x * 32;
    const int t0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const long long v = v0 + k;
        const int t = t0 + tx;
        if (v < nvox && t < T) tile[k][tx] = src[v * T + t];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int t = t0 + k;
        const long long v = v0 + tx;
        if (v < nvox && t < T) dst[(long long)t * nvox + v] = tile[tx][k];
    }
}

// Exclusive scans of int32 counts (1024 elements per block, three passes).
template <typename TO>
__device__ TO block_incl_scan(TO x, TO *wsum) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const TO o = __shfl_up(x, d, 64);
        if (lane >= d) x += o;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        TO y = (lane < (int)(blockDim.x >> 6)) ? wsum[lane] : (TO)0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const TO o = __shfl_up(y, d, 64);
            if (lane >= d) y += o;
        }
        if (lane < (int)(blockDim.x >> 6)) wsum[lane] = y;
    }
    __syncthreads();
    if (w > 0) x += wsum[w - 1];
    __syncthreads();
    return x;
}

template <typename TO>
__global__ void __launch_bounds__(1024) scan_pass1(const int *__restrict__ in, int n, TO *__restrict__ bsum) {
    __shared__ TO wsum[16];
    const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    const TO x = (i < n) ? (TO)in[i] : (TO)0;
    const TO inc = block_incl_scan<TO>(x, wsum);
    if (threadIdx.x == 1023) bsum[blockIdx.x] = inc;
}

template <typename TO>
__global__ void __launch_bounds__(1024) scan_pass2(TO *__restrict__ bsum, int nb) {
    __shared__ TO wsum[16];
    __shared__ TO carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int i = b0 + threadIdx.x;
        const TO x = (i < nb) ? bsum[i] : (TO)0;
        const TO inc = block_incl_scan<TO>(x, wsum);
        const TO c = carry;
        if (i < nb) bsum[i] = c + inc - x;  // exclusive
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + inc;
        __syncthreads();
    }
}

template <typename TO>
__global__ void __launch_bounds__(1024) scan_pass3(const int *__restrict__ in, int n, const TO *__restrict__ bsum,
                                                   TO *__restrict__ out) {
    __shared__ TO wsum[16];
    const long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    const TO x = (i < n) ? (TO)in[i] : (TO)0;
    const TO inc = block_incl_scan<TO>(x, wsum);
    const TO base = bsum[blockIdx.x];
    if (i < n) out[i] = base + inc - x;
    if (i == n - 1) out[n] = base + inc;
}

// Copy each streamline's points from its padded slot into the packed array.
__global__ void compact_kernel(const int *__restrict__ len, int nsl, const float *__restrict__ sline,
                               size_t slot_f, const long long *__restrict__ off, float *__restrict__ pts) {
    const int lane = threadIdx.x & 63;
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int k = w; k < nsl; k += nw) {
        const int nf = 3 * len[k];
        const float *src = sline + (size_t)k * slot_f;
        float *dst = pts + 3 * (size_t)off[k];
        for (int p = lane; p < nf; p += 64) dst[p] = src[p];
    }
}

// Totals of a bounded chunk (one thread): the chunk's streamlines
// (offsets[n]), the slots phase 2 filled -- the largest offset <= cap, as
// the offsets never decrease and a seed is tracked when its last slot is
// below the cap -- and the points compacted (point_off[cap]).
__global__ void chunk_totals(const int *__restrict__ offsets, int n, int cap, const long long *__restrict__ point_off,
                             long long *__restrict__ totals) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= cap) lo = mid; else hi = mid - 1;
    }
    totals[0] = offsets[n];
    totals[1] = offsets[lo];
    totals[2] = point_off[cap];
}

}  // namespace

// Workspace of cuslines_phase2_noalloc_async: the spatial seed order of at
// most CUSLINES_NOALLOC_MAX_SEEDS seeds (order and keys), its bucket counts
// and starts (at most 2^21 buckets) and the scan's block sums.  One copy per
// device for the life of the process, shared by every call on that device.
constexpr size_t SCRATCH_INTS = 2 * (size_t)CUSLINES_NOALLOC_MAX_SEEDS + 2 * ((size_t)1 << 21) + 1 + 2048;
constexpr size_t SCRATCH_BYTES = SCRATCH_INTS * sizeof(int);
static __device__ __align__(256) int scratch[SCRATCH_INTS];

// =========================================================================
// Host side: C ABI
// =========================================================================
thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

// roctx range over one C-ABI call (phases show up by name in rocprofv3
// --marker-trace timelines).
struct RoctxRange {
    explicit RoctxRange(const char *name) { roctxRangePushA(name); }
    ~RoctxRange() { roctxRangePop(); }
};

#define HIPC(expr)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return fail(CUSLINES_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct cuslines_tracker {
    cuslines_params p{};
    DevParams dp{};
    int device = 0;
    hipStream_t stream = nullptr;
    float *d_data = nullptr, *d_stop = nullptr, *d_verts = nullptr;
    uint32_t *d_edges = nullptr;
    uint16_t *d_nbr = nullptr;  // vertex adjacency, 8 per vertex (nullptr: a degree above 8 or > 1024 vertices)
    float *d_H = nullptr, *d_R = nullptr, *d_db = nullptr, *d_dq = nullptr, *d_smat = nullptr;
    float *d_hrm = nullptr, *d_fitm = nullptr, *d_odfm = nullptr;  // batched bootstrap matrices
    float *d_ut = nullptr, *d_u = nullptr, *d_invlev = nullptr;   // low-rank form of [H; R] (rk16 > 0)
    int rk16 = 0, hrank = 0;                                      // rank of H and rounded to 16 (0: not that form)
    int boot_full = 0;                                            // batched bootstrap: always the [H; R] product
    int boot_mode = CUSLINES_BOOT_EXACT;
    int boot_occ = 4;  // batched bootstrap: waves per SIMD of the kernel variant
    int *d_b0 = nullptr, *d_dwi = nullptr;
    uint32_t *d_lut = nullptr;
    std::vector<uint16_t> rank_vertex;  // PTT: vertex of each direction rank (corner-block order)
    int *d_work = nullptr;      // work-queue counters of the persistent phase-2 kernels (QUEUES)
    void *scratch = nullptr;    // device address of this device's `scratch` array
    float *d_blocks = nullptr;  // PTT corner blocks
    int use_blocks = 1;
    int schedule = CUSLINES_SCHED_QUEUE;
    int threads = 0;        // 0: per-getter default
    int blocks_per_cu = 0;  // 0: occupancy maximum
    int num_cu = 1;
    int ptt_lanes = 0;      // PTT: lanes per seed (8, 16, 32; 0 = shared 32-lane kernel)
    int flat = 1;           // phase 2 on the queue: one loop of steps per seed slot (flat_slot)
    float *d_fprop = nullptr;            // learned proposal: feature propagators [24][9]
    float *d_vprop = nullptr;            // PTT vertex-phase propagators [24][9]
    uint16_t *d_mlp = nullptr;           // learned proposal (nullptr: plain getter)
    unsigned long long *d_stats = nullptr;
    int count_taps = 0;  // PTT phase 2 counts its taps and sub-steps into d_stats
    float *rec = nullptr;                // caller's training-pair buffer (nullptr: no recording)
    int *rec_count = nullptr;
    bool has_data = false, has_stop = false, has_sphere = false, has_boot = false, has_lut = false;
    cuslines_tracker *parent = nullptr;  // fork: the tracker whose device data it shares
    int nforks = 0;                      // live forks of this tracker
    int ngraphs = 0;                     // live chunk graphs on this tracker (a parent: and on its forks)
};

// A captured chunk (cuslines_chunk_graph_create): its kernel nodes hold the
// tracker's device pointers and launch settings as they were at capture.
struct cuslines_graph {
    cuslines_tracker *t;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

static float smallest_float_at_least(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

static int ceil32(int x) { return (x + 31) & ~31; }

static DevBufs bufs_of(const cuslines_tracker *t) {
    DevBufs b;
    b.data = t->d_data;
    b.stop = t->d_stop;
    b.verts = t->d_verts;
    b.edges = t->d_edges;
    b.nbr = t->d_nbr;
    b.H = t->d_H; b.R = t->d_R; b.db = t->d_db; b.dq = t->d_dq; b.smat = t->d_smat;
    b.hr_mat = t->d_hrm; b.fit_mat = t->d_fitm; b.odf_mat = t->d_odfm;
    b.ut_mat = t->d_ut; b.u_mat = t->d_u; b.invlev = t->d_invlev;
    b.b0 = t->d_b0;
    b.dwi_idx = t->d_dwi;
    b.lut = t->d_lut;
    b.blocks = t->use_blocks ? t->d_blocks : nullptr;
    b.fprop = t->d_fprop;
    b.vprop = t->d_vprop;
    b.mlp = t->d_mlp;
    b.rec = t->rec;
    b.rec_count = t->rec_count;
    b.stats = t->d_stats;
    return b;
}

static void dfree(void *p) {
    if (p) (void)hipFree(p);
}

// Device address of `scratch` on each device, looked up once per device with
// that device current (each device holds its own copy of the array).
static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr;

static int scratch_of_device(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CUSLINES_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return CUSLINES_OK;
}

static int check_ready(const cuslines_tracker *t) {
    if (!t->has_data || !t->has_stop || !t->has_sphere)
        return fail(CUSLINES_ESTATE, "data, stop map and sphere must be uploaded first");
    if ((t->p.model_type == CUSLINES_OPDT || t->p.model_type == CUSLINES_CSA) && !t->has_boot)
        return fail(CUSLINES_ESTATE, "bootstrap matrices must be uploaded first");
    if (t->p.model_type == CUSLINES_PTT && !t->has_lut) return fail(CUSLINES_ESTATE, "PTT LUT must be uploaded first");
    return CUSLINES_OK;
}

// Setters that replace device data are refused on forks (they share their
// parent's buffers) and on trackers with live forks.
static int check_owner(const cuslines_tracker *t) {
    if (t->parent) return fail(CUSLINES_ESTATE, "a fork shares its parent's data: configure the parent");
    if (t->nforks) return fail(CUSLINES_ESTATE, "live forks share this tracker's data: destroy them first");
    if (t->ngraphs) return fail(CUSLINES_ESTATE, "live chunk graphs hold this tracker's data: destroy them first");
    return CUSLINES_OK;
}

template <int MODEL>
static int launch_phase1(cuslines_tracker *t, const float *seeds, int n, int chunk, int *counts, int *dirs) {
    const DevParams &d = t->dp;
    const int threads = 128;
    const int nslots = threads / 32;
    const size_t vf = verts_in_lds<MODEL>() ? (size_t)((3 * d.nverts + 3) & ~3) : 0;
    const size_t lds = sizeof(float) * (vf + (size_t)nslots * slot_floats(MODEL, true, d.nv32, d.t32, d.hr32));
    if (lds > 160 * 1024) return fail(CUSLINES_EINVAL, "sphere too large for LDS");
    HIPC(hipFuncSetAttribute((const void *)phase1_kernel<MODEL>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const unsigned grid = (unsigned)((n + nslots - 1) / nslots);
    hipLaunchKernelGGL(phase1_kernel<MODEL>, dim3(grid), dim3(threads), lds, t->stream, d, bufs_of(t), seeds, n,
                       chunk, counts, dirs);
    HIPC(hipGetLastError());
    return CUSLINES_OK;
}

static int spatial_order(cuslines_tracker *t, const float *seeds, int n, int **order, int *ws);

// PTT with L lanes per seed: with flat stepping (t->flat, the default) the
// flat grouped kernel over the XCD-ordered queues of the spatial seed order
// (`ws` as launch_phase2_v), else the per-seed kernel over one counter.
template <int L>
static int launch_ptt_l(cuslines_tracker *t, int threads, const float *seeds, int n, int chunk, const int *offsets,
                        const int *dirs, int *sseed, int *slen, float *sl, int *ws) {
    const DevParams &d = t->dp;
    const size_t lds = sizeof(float) * (size_t)(PTTC_FLOATS + (threads / L) * PTTG_FLOATS);
    const bool flat = t->flat != 0;
    const void *kfn = flat ? (const void *)phase2_ptt_flat_kernel<L> : (const void *)phase2_ptt_queue_kernel<L>;
    HIPC(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, threads, lds));
    if (per_cu < 1) per_cu = 1;
    if (t->blocks_per_cu > 0 && t->blocks_per_cu < per_cu) per_cu = t->blocks_per_cu;
    long long grid = (long long)per_cu * t->num_cu;
    const long long need = ((long long)n * L + threads - 1) / threads;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    int *order = nullptr;
    if (flat)
        if (int rc = spatial_order(t, seeds, n, &order, ws)) return rc;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(QUEUES), 0, t->stream, t->d_work, QUEUES);
    if (flat)
        hipLaunchKernelGGL(phase2_ptt_flat_kernel<L>, dim3((unsigned)grid), dim3(threads), lds, t->stream, d,
                           bufs_of(t), seeds, n, chunk, offsets, dirs, sseed, slen, sl, order, t->d_work);
    else
        hipLaunchKernelGGL(phase2_ptt_queue_kernel<L>, dim3((unsigned)grid), dim3(threads), lds, t->stream, d,
                           bufs_of(t), seeds, n, chunk, offsets, dirs, sseed, slen, sl, t->d_work);
    const hipError_t e = hipGetLastError();
    if (flat && !ws) (void)hipFreeAsync(order, t->stream);
    if (e != hipSuccess) return fail(CUSLINES_EHIP, std::string("phase 2 (PTT groups): ") + hipGetErrorString(e));
    return CUSLINES_OK;
}

static int launch_ptt_groups(cuslines_tracker *t, const float *seeds, int n, int chunk, const int *offsets,
                             const int *dirs, int *sseed, int *slen, float *sl, int *ws) {
    const int threads = t->threads > 0 ? t->threads : 256;
    switch (t->ptt_lanes) {
        case 8: return launch_ptt_l<8>(t, threads, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        case 16: return launch_ptt_l<16>(t, threads, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        default: return launch_ptt_l<32>(t, threads, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
    }
}

template <typename TO>
static int run_scan(cuslines_tracker *t, const int *in, int n, TO *out, TO *bsum_ws = nullptr);

// Buckets of the spatial seed order: 4-voxel bricks (coarser when a
// dimension exceeds 512 voxels), at most 2^21 of them.
static int order_buckets(const DevParams &d, int *brick_log2_out) {
    int mx = d.dimx > d.dimy ? d.dimx : d.dimy;
    mx = mx > d.dimz ? mx : d.dimz;
    int brick_log2 = 2;
    while (((mx - 1) >> brick_log2) >= 128) ++brick_log2;
    int bits = 0;
    while ((1 << bits) <= ((mx - 1) >> brick_log2)) ++bits;
    *brick_log2_out = brick_log2;
    return 1 << (3 * bits);
}

// ints of the seed-order workspace: order and keys (n each), bucket counts
// (nb), bucket starts (nb + 1) and the scan's block sums.
static size_t order_ws_ints(int n, int nb) {
    return 2 * (size_t)n + 2 * (size_t)nb + 1 + (size_t)((nb + 1023) / 1024);
}

// Seeds in Morton order of their brick, by a counting sort on the tracker's
// stream.  With `ws` (order_ws_ints(n, nb) ints of device memory) every
// buffer is carved from it and *order points into it; without, they are
// stream-ordered allocations and the caller frees *order with hipFreeAsync.
static int spatial_order(cuslines_tracker *t, const float *seeds, int n, int **order, int *ws) {
    const DevParams &d = t->dp;
    int brick_log2 = 0;
    const int nb = order_buckets(d, &brick_log2);
    int *key = nullptr, *count = nullptr, *start = nullptr, *bsum = nullptr;
    if (ws) {
        *order = ws;
        key = ws + n;
        count = key + n;
        start = count + nb;
        bsum = start + nb + 1;
    } else {
        *order = nullptr;
        hipError_t e = hipMallocAsync((void **)order, (size_t)n * sizeof(int), t->stream);
        if (e == hipSuccess) e = hipMallocAsync((void **)&key, (size_t)n * sizeof(int), t->stream);
        if (e == hipSuccess) e = hipMallocAsync((void **)&count, (size_t)nb * sizeof(int), t->stream);
        if (e == hipSuccess) e = hipMallocAsync((void **)&start, ((size_t)nb + 1) * sizeof(int), t->stream);
        if (e != hipSuccess) {
            for (int *p : {*order, key, count, start})
                if (p) (void)hipFreeAsync(p, t->stream);
            *order = nullptr;
            return fail(CUSLINES_ENOMEM, std::string("seed order: ") + hipGetErrorString(e));
        }
    }
    const dim3 gb((unsigned)((nb + 255) / 256)), gs((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(zero_ints, gb, dim3(256), 0, t->stream, count, nb);
    hipLaunchKernelGGL(seed_keys, gs, dim3(256), 0, t->stream, d, seeds, n, brick_log2, key, count);
    int rc = run_scan<int>(t, count, nb, start, bsum);
    if (rc == CUSLINES_OK) {
        hipLaunchKernelGGL(zero_ints, gb, dim3(256), 0, t->stream, count, nb);
        hipLaunchKernelGGL(seed_scatter, gs, dim3(256), 0, t->stream, key, n, start, count, *order);
    }
    hipError_t e = hipGetLastError();
    if (!ws) {
        (void)hipFreeAsync(key, t->stream);
        (void)hipFreeAsync(count, t->stream);
        (void)hipFreeAsync(start, t->stream);
    }
    if (rc) return rc;
    if (e != hipSuccess) return fail(CUSLINES_EHIP, std::string("seed order: ") + hipGetErrorString(e));
    return CUSLINES_OK;
}

template <int MODEL, int V>
static int launch_phase2_v(cuslines_tracker *t, const float *seeds, int n, int chunk, const int *offsets,
                           const int *dirs, int *sseed, int *slen, float *sl, int *ws) {
    const DevParams &d = t->dp;
    int threads = t->threads;
    if (threads <= 0) threads = (MODEL == CUSLINES_OPDT || MODEL == CUSLINES_CSA) ? 128 : 256;
    // learned: 512-thread workgroups of the queue kernel share one staged
    // copy of the network (50 KB of LDS) among 16 slots
    if (V == 1 && t->threads <= 0 && t->schedule != CUSLINES_SCHED_STATIC) threads = 512;
    if (V == 5 && t->threads <= 0) threads = 768;
    const int nslots = threads / 32;
    const size_t vf = ((MODEL == CUSLINES_PROB) ? (size_t)((3 * d.nverts + 3) & ~3) : 0) +
                      ((V == 1 || V == 5) ? (size_t)MLP_LDS_FLOATS : 0);
    const size_t lds = sizeof(float) * (vf + (size_t)nslots * (slot_floats(MODEL, false, d.nv32, d.t32, d.hr32) +
                                                             ((V == 4 || V == 5) ? FLAT_STATE_FLOATS : 0)));
    if (lds > 160 * 1024) return fail(CUSLINES_EINVAL, "sphere too large for LDS");
    if (t->schedule == CUSLINES_SCHED_STATIC) {
        HIPC(hipFuncSetAttribute((const void *)phase2_kernel<MODEL, V>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
        const unsigned grid = (unsigned)((n + nslots - 1) / nslots);
        hipLaunchKernelGGL((phase2_kernel<MODEL, V>), dim3(grid), dim3(threads), lds, t->stream, d, bufs_of(t), seeds,
                           n, chunk, offsets, dirs, sseed, slen, sl);
        HIPC(hipGetLastError());
        return CUSLINES_OK;
    }
    const void *kfn = (const void *)phase2_queue_kernel<MODEL, V>;
    HIPC(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, threads, lds));
    if (per_cu < 1) per_cu = 1;
    if (t->blocks_per_cu > 0 && t->blocks_per_cu < per_cu) per_cu = t->blocks_per_cu;
    long long grid = (long long)per_cu * t->num_cu;
    const long long need = (n + nslots - 1) / nslots;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    int *order = nullptr;
    if (int rc = spatial_order(t, seeds, n, &order, ws)) return rc;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(QUEUES), 0, t->stream, t->d_work, QUEUES);
    hipLaunchKernelGGL((phase2_queue_kernel<MODEL, V>), dim3((unsigned)grid), dim3(threads), lds, t->stream, d,
                       bufs_of(t), seeds, n, chunk, offsets, dirs, sseed, slen, sl, order, t->d_work);
    hipError_t e = hipGetLastError();
    if (!ws) (void)hipFreeAsync(order, t->stream);
    if (e != hipSuccess) return fail(CUSLINES_EHIP, std::string("phase 2: ") + hipGetErrorString(e));
    return CUSLINES_OK;
}

// Limits of the batched bootstrap kernel's tile counts per wave.
static bool boot_batch_fits(const DevParams &d) {
    return d.hr16 <= 320 && d.dn16 <= 128 && d.nv16 <= 768 && d.dn16 <= d.t32;
}

template <int MODEL>
static int launch_boot_batch(cuslines_tracker *t, const float *seeds, int n, int chunk, const int *offsets,
                             const int *dirs, int *sseed, int *slen, float *sl, int *ws) {
    DevParams d = t->dp;
    d.rk16 = t->boot_full ? 0 : t->rk16;  // H.d = U (U^T d) and R.d from it
    if (!boot_batch_fits(d))
        return fail(CUSLINES_EINVAL, "batched bootstrap needs hr <= 320, delta_nr <= 128 and nverts <= 768");
    const size_t lds = sizeof(float) * ((size_t)BB_SLOTS * (size_t)bb_slot_stride(d.nv32, d.t32, d.hr32) +
                                        (size_t)BB_WAVES * 16);
    if (lds > 160 * 1024) return fail(CUSLINES_EINVAL, "batched bootstrap slots too large for LDS");
    const bool lr = d.rk16 > 0;
    const void *kfn = (t->boot_occ == 4)
                          ? (lr ? (const void *)phase2_boot_batch_kernel<MODEL, 4, true>
                                : (const void *)phase2_boot_batch_kernel<MODEL, 4, false>)
                          : (lr ? (const void *)phase2_boot_batch_kernel<MODEL, 2, true>
                                : (const void *)phase2_boot_batch_kernel<MODEL, 2, false>);
    HIPC(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, BB_THREADS, lds));
    if (per_cu < 1) per_cu = 1;
    if (t->blocks_per_cu > 0 && t->blocks_per_cu < per_cu) per_cu = t->blocks_per_cu;
    long long grid = (long long)per_cu * t->num_cu;
    const long long need = (n + BB_SLOTS - 1) / BB_SLOTS;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    if (!t->d_stats && !ws) {  // the kernel's attempt and round counters (not allocated by the no-alloc entry)
        HIPC(hipMalloc(&t->d_stats, 4 * sizeof(unsigned long long)));
        HIPC(hipMemsetAsync(t->d_stats, 0, 4 * sizeof(unsigned long long), t->stream));
    }
    int *order = nullptr;
    if (int rc = spatial_order(t, seeds, n, &order, ws)) return rc;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(QUEUES), 0, t->stream, t->d_work, QUEUES);
    const DevBufs bb = bufs_of(t);
    const dim3 g((unsigned)grid), b(BB_THREADS);
    if (t->boot_occ == 4 && lr)
        hipLaunchKernelGGL((phase2_boot_batch_kernel<MODEL, 4, true>), g, b, lds, t->stream, d, bb, seeds, n, chunk,
                           offsets, dirs, sseed, slen, sl, order, t->d_work);
    else if (t->boot_occ == 4)
        hipLaunchKernelGGL((phase2_boot_batch_kernel<MODEL, 4, false>), g, b, lds, t->stream, d, bb, seeds, n, chunk,
                           offsets, dirs, sseed, slen, sl, order, t->d_work);
    else if (lr)
        hipLaunchKernelGGL((phase2_boot_batch_kernel<MODEL, 2, true>), g, b, lds, t->stream, d, bb, seeds, n, chunk,
                           offsets, dirs, sseed, slen, sl, order, t->d_work);
    else
        hipLaunchKernelGGL((phase2_boot_batch_kernel<MODEL, 2, false>), g, b, lds, t->stream, d, bb, seeds, n, chunk,
                           offsets, dirs, sseed, slen, sl, order, t->d_work);
    hipError_t e = hipGetLastError();
    if (!ws) (void)hipFreeAsync(order, t->stream);
    if (e != hipSuccess) return fail(CUSLINES_EHIP, std::string("phase 2 (batched): ") + hipGetErrorString(e));
    return CUSLINES_OK;
}

template <int MODEL>
static int launch_phase2(cuslines_tracker *t, const float *seeds, int n, int chunk, const int *offsets,
                         const int *dirs, int *sseed, int *slen, float *sl, int *ws) {
    if constexpr (MODEL == CUSLINES_OPDT || MODEL == CUSLINES_CSA) {
        if (t->boot_mode == CUSLINES_BOOT_MFMA && t->schedule != CUSLINES_SCHED_STATIC)
            return launch_boot_batch<MODEL>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
    }
    if constexpr (MODEL == CUSLINES_PTT) {
        if (t->d_mlp && t->flat && t->schedule != CUSLINES_SCHED_STATIC)
            return launch_phase2_v<MODEL, 5>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        if (t->d_mlp)
            return launch_phase2_v<MODEL, 1>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        if (t->rec)
            return launch_phase2_v<MODEL, 2>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        if (t->count_taps)
            return launch_phase2_v<MODEL, 3>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        if (t->ptt_lanes > 0 && t->schedule != CUSLINES_SCHED_STATIC)
            return launch_ptt_groups(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
    }
    if (t->flat && t->schedule != CUSLINES_SCHED_STATIC)
        return launch_phase2_v<MODEL, 4>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
    return launch_phase2_v<MODEL, 0>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
}

template <typename TO>
static int run_scan(cuslines_tracker *t, const int *in, int n, TO *out, TO *bsum_ws) {
    if (n <= 0) {
        hipLaunchKernelGGL(zero_scalar<TO>, dim3(1), dim3(1), 0, t->stream, out);
        HIPC(hipGetLastError());
        return CUSLINES_OK;
    }
    const int nb = (n + 1023) / 1024;
    TO *bsum = bsum_ws;
    if (!bsum) HIPC(hipMallocAsync((void **)&bsum, (size_t)nb * sizeof(TO), t->stream));
    hipLaunchKernelGGL(scan_pass1<TO>, dim3(nb), dim3(1024), 0, t->stream, in, n, bsum);
    hipLaunchKernelGGL(scan_pass2<TO>, dim3(1), dim3(1024), 0, t->stream, bsum, nb);
    hipLaunchKernelGGL(scan_pass3<TO>, dim3(nb), dim3(1024), 0, t->stream, in, n, bsum, out);
    hipError_t e = hipGetLastError();
    if (!bsum_ws) (void)hipFreeAsync(bsum, t->stream);
    if (e != hipSuccess) return fail(CUSLINES_EHIP, std::string("scan: ") + hipGetErrorString(e));
    return CUSLINES_OK;
}

static int phase1_models(cuslines_tracker *t, const float *seeds, int n, int chunk, int *counts, int *dirs) {
    switch (t->p.model_type) {
        case CUSLINES_OPDT: return launch_phase1<CUSLINES_OPDT>(t, seeds, n, chunk, counts, dirs);
        case CUSLINES_CSA: return launch_phase1<CUSLINES_CSA>(t, seeds, n, chunk, counts, dirs);
        case CUSLINES_PROB: return launch_phase1<CUSLINES_PROB>(t, seeds, n, chunk, counts, dirs);
        default: return launch_phase1<CUSLINES_PTT>(t, seeds, n, chunk, counts, dirs);
    }
}

// Phase 2 of the handle's model; `ws` as launch_phase2_v (nullptr:
// stream-ordered allocations).
static int phase2_models(cuslines_tracker *t, const float *seeds, int n, int chunk, const int *offsets,
                         const int *dirs, int *sseed, int *slen, float *sl, int *ws) {
    switch (t->p.model_type) {
        case CUSLINES_OPDT:
            return launch_phase2<CUSLINES_OPDT>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        case CUSLINES_CSA:
            return launch_phase2<CUSLINES_CSA>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        case CUSLINES_PROB:
            return launch_phase2<CUSLINES_PROB>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
        default:
            return launch_phase2<CUSLINES_PTT>(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, ws);
    }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace of a bounded chunk: the seed order's (order_ws_ints), then the
// block sums of the streamline-count scan (int) and of the point scan (int64).
static size_t chunk_ws_bytes(const DevParams &d, int n, int cap) {
    int bl = 0;
    return align256(order_ws_ints(n, order_buckets(d, &bl)) * sizeof(int)) +
           align256((size_t)((n + 1023) / 1024) * sizeof(int)) + (size_t)((cap + 1023) / 1024) * sizeof(long long);
}

// The stream work of cuslines_chunk_async (arguments already checked):
// phase 1, the scan, the lengths of the `cap` slots cleared, phase 2 of the
// seeds that fit (DevParams::slot_cap), the point scan and compaction over
// the `cap` slots and the totals.  No allocation and no host read.
static int chunk_enqueue(cuslines_tracker *t, const float *seeds, int n, int chunk, int *counts, int *dirs,
                         int *offsets, int cap, int *sseed, int *slen, float *sl, long long *poff, float *pts,
                         long long *totals, char *ws) {
    int bl = 0;
    int *order_ws = reinterpret_cast<int *>(ws);
    char *p = ws + align256(order_ws_ints(n, order_buckets(t->dp, &bl)) * sizeof(int));
    int *scan_ws = reinterpret_cast<int *>(p);
    long long *pscan_ws = reinterpret_cast<long long *>(p + align256((size_t)((n + 1023) / 1024) * sizeof(int)));
    if (n > 0)
        if (int rc = phase1_models(t, seeds, n, chunk, counts, dirs)) return rc;
    if (int rc = run_scan<int>(t, counts, n, offsets, scan_ws)) return rc;
    hipLaunchKernelGGL(zero_ints, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, t->stream, slen, cap);
    HIPC(hipGetLastError());
    if (n > 0) {
        t->dp.slot_cap = cap;  // copied into the launches' arguments
        const int rc = phase2_models(t, seeds, n, chunk, offsets, dirs, sseed, slen, sl, order_ws);
        t->dp.slot_cap = 0;
        if (rc) return rc;
    }
    if (int rc = run_scan<long long>(t, slen, cap, poff, pscan_ws)) return rc;
    const size_t slot_f = (size_t)2 * t->dp.max_sline_len * 3;
    const unsigned grid = (unsigned)std::min<long long>(((long long)cap + 3) / 4, 65536);
    hipLaunchKernelGGL(compact_kernel, dim3(grid), dim3(256), 0, t->stream, slen, cap, sl, slot_f, poff, pts);
    hipLaunchKernelGGL(chunk_totals, dim3(1), dim3(1), 0, t->stream, offsets, n, cap, poff, totals);
    HIPC(hipGetLastError());
    return CUSLINES_OK;
}

struct chunk_args {
    const float *seeds;
    int n, chunk;
    int *counts, *dirs, *offsets;
    int cap;
    int *sseed, *slen;
    float *sl;
    long long *poff;
    float *pts;
    long long *totals;
    char *ws;
};

static int chunk_check(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk, int32_t *counts_d,
                       int32_t *dirs_d, int32_t *offsets_d, int32_t slot_cap, int32_t *sline_seed_d,
                       int32_t *sline_len_d, float *sline_d, int64_t *point_off_d, float *points_d, int64_t *totals_d,
                       void *ws_d, size_t ws_bytes, chunk_args *a) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (nseeds < 0 || rng_chunk < 0) return fail(CUSLINES_EINVAL, "negative size");
    if (slot_cap < 1) return fail(CUSLINES_EINVAL, "slot_cap must be at least 1");
    if ((nseeds > 0 && (!seeds_d || !counts_d || !dirs_d)) || !offsets_d || !sline_seed_d || !sline_len_d ||
        !sline_d || !point_off_d || !points_d || !totals_d || !ws_d)
        return fail(CUSLINES_EINVAL, "null buffer");
    if (int rc = check_ready(t)) return rc;
    const size_t need = chunk_ws_bytes(t->dp, nseeds, slot_cap);
    if (ws_bytes < need)
        return fail(CUSLINES_EINVAL, "workspace of " + std::to_string(ws_bytes) + " bytes, " + std::to_string(need) +
                                         " needed (cuslines_chunk_workspace_bytes)");
    if ((reinterpret_cast<uintptr_t>(ws_d) & 255) != 0) return fail(CUSLINES_EINVAL, "workspace not 256-byte aligned");
    *a = chunk_args{seeds_d, nseeds, rng_chunk > 0 ? rng_chunk : nseeds, counts_d, dirs_d, offsets_d, slot_cap,
                    sline_seed_d, sline_len_d, sline_d, reinterpret_cast<long long *>(point_off_d), points_d,
                    reinterpret_cast<long long *>(totals_d), static_cast<char *>(ws_d)};
    return CUSLINES_OK;
}

static int chunk_run(cuslines_tracker *t, const chunk_args &a) {
    return chunk_enqueue(t, a.seeds, a.n, a.chunk, a.counts, a.dirs, a.offsets, a.cap, a.sseed, a.slen, a.sl, a.poff,
                         a.pts, a.totals, a.ws);
}

// Both phase-2 entry points.  noalloc: the seed-order workspace is this
// device's `scratch` array instead of stream-ordered allocations.
static int phase2_entry(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk,
                        const int32_t *offsets_d, const int32_t *dirs_d, int32_t *sline_seed_d, int32_t *sline_len_d,
                        float *sline_d, bool noalloc) {
    RoctxRange range_("cuslines_phase2");
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (nseeds < 0 || rng_chunk < 0) return fail(CUSLINES_EINVAL, "negative size");
    if (noalloc) {
        if (!t->scratch) return fail(CUSLINES_ESTATE, "no scratch array on this device");
        if (nseeds > CUSLINES_NOALLOC_MAX_SEEDS)
            return fail(CUSLINES_EINVAL, "nseeds exceeds CUSLINES_NOALLOC_MAX_SEEDS (" +
                                             std::to_string(CUSLINES_NOALLOC_MAX_SEEDS) + ")");
    }
    if (nseeds == 0) return CUSLINES_OK;
    if (!seeds_d || !offsets_d || !dirs_d || !sline_seed_d || !sline_len_d || !sline_d)
        return fail(CUSLINES_EINVAL, "null buffer");
    int rc = check_ready(t);
    if (rc) return rc;
    int *ws = nullptr;
    if (noalloc) {
        int brick_log2 = 0;
        if (order_ws_ints(nseeds, order_buckets(t->dp, &brick_log2)) * sizeof(int) > SCRATCH_BYTES)
            return fail(CUSLINES_EINVAL, "seed-order workspace exceeds the scratch array");
        ws = static_cast<int *>(t->scratch);
    }
    HIPC(hipSetDevice(t->device));
    const int chunk = rng_chunk > 0 ? rng_chunk : nseeds;
    return phase2_models(t, seeds_d, nseeds, chunk, offsets_d, dirs_d, sline_seed_d, sline_len_d, sline_d, ws);
}

extern "C" {

int cuslines_abi_version(void) { return CUSLINES_ABI_VERSION; }

int cuslines_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return CUSLINES_EINVAL;
    const size_t n = std::min(len - 1, g_last_error.size());
    std::memcpy(buf, g_last_error.data(), n);
    buf[n] = '\0';
    return CUSLINES_OK;
}

int cuslines_device_count(int *count) {
    if (!count) return fail(CUSLINES_EINVAL, "count is null");
    int n = 0;
    HIPC(hipGetDeviceCount(&n));
    *count = n;
    return CUSLINES_OK;
}

int cuslines_tracker_create(const cuslines_params *p, int device, cuslines_tracker **out) {
    if (!p || !out) return fail(CUSLINES_EINVAL, "null argument");
    *out = nullptr;
    if (p->dimx <= 0 || p->dimy <= 0 || p->dimz <= 0 || p->dimt <= 0)
        return fail(CUSLINES_EINVAL, "volume dimensions must be positive");
    if (p->model_type < CUSLINES_OPDT || p->model_type > CUSLINES_PTT)
        return fail(CUSLINES_EINVAL, "invalid model type");
    if (p->nverts <= 0 || p->nverts > 65535 || p->nedges <= 0)
        return fail(CUSLINES_EINVAL, "sphere must have 1..65535 vertices and at least one edge");
    if (p->max_sline_len < 2) return fail(CUSLINES_EINVAL, "max_sline_len must be >= 2");
    if ((p->model_type == CUSLINES_PROB || p->model_type == CUSLINES_PTT) && p->nverts != p->dimt)
        return fail(CUSLINES_EINVAL, "prob/PTT data must hold one ODF value per sphere vertex");
    if (p->model_type == CUSLINES_PTT) {
        if (p->step_frac < 1 || p->probe_quality < 1 || p->lut_res < 2)
            return fail(CUSLINES_EINVAL, "invalid PTT parameters");
        const long long slots = 1LL << (p->tile_log2_x + p->tile_log2_y + p->tile_log2_z);
        if (slots < p->dimt) return fail(CUSLINES_EINVAL, "PTT tiling has fewer slots than vertices");
        // limits of the 24-bit index products in interp4
        const long long X1 = p->dimx + 1LL, Y1 = p->dimy + 1LL, Z1 = p->dimz + 1LL;
        if (p->lut_res > 1024 || X1 * Y1 >= (1LL << 24) || X1 * Y1 * Z1 >= (1LL << 32) ||
            (1LL << p->tile_log2_x) * p->dimx >= (1LL << 24) || (1LL << p->tile_log2_y) * p->dimy >= (1LL << 24) ||
            (1LL << p->tile_log2_z) * p->dimz >= (1LL << 24))
            return fail(CUSLINES_EINVAL, "PTT volume or LUT too large for the tap index arithmetic");
    }
    if ((p->model_type == CUSLINES_OPDT || p->model_type == CUSLINES_CSA) && p->delta_nr <= 0)
        return fail(CUSLINES_EINVAL, "bootstrap needs delta_nr > 0");
    int ndev = 0;
    HIPC(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(CUSLINES_EINVAL, "device index out of range");
    HIPC(hipSetDevice(device));
    cuslines_tracker *t = new cuslines_tracker();
    t->p = *p;
    t->device = device;
    hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete t;
        return fail(CUSLINES_EHIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    }
    e = hipMalloc(&t->d_work, QUEUES * sizeof(int));
    if (e == hipSuccess) e = hipDeviceGetAttribute(&t->num_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        dfree(t->d_work);
        (void)hipStreamDestroy(t->stream);
        delete t;
        return fail(CUSLINES_EHIP, std::string("tracker setup: ") + hipGetErrorString(e));
    }
    if (int rc = scratch_of_device(device, &t->scratch)) {
        dfree(t->d_work);
        (void)hipStreamDestroy(t->stream);
        delete t;
        return rc;
    }
    unsigned char net[SORT_LAYERS][32];
    int nl = 0;
    make_sortnet(net, &nl);
    uint32_t packed_net[32][4] = {};
    for (int L = 0; L < SORT_LAYERS; ++L)
        for (int g = 0; g < 32; ++g) packed_net[g][L >> 2] |= (uint32_t)net[L][g] << (8 * (L & 3));
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_sortnet4), packed_net, sizeof(packed_net));
    if (e != hipSuccess) {
        dfree(t->d_work);
        (void)hipStreamDestroy(t->stream);
        delete t;
        return fail(CUSLINES_EHIP, std::string("sort network upload: ") + hipGetErrorString(e));
    }
    DevParams &d = t->dp;
    d.dimx = p->dimx; d.dimy = p->dimy; d.dimz = p->dimz; d.dimt = p->dimt; d.t32 = ceil32(p->dimt);
    d.nverts = p->nverts; d.nv32 = ceil32(p->nverts); d.nedges = p->nedges;
    d.sphere_symm = p->sphere_symm ? 1 : 0;
    d.max_sline_len = p->max_sline_len;
    d.model = p->model_type;
    d.hr = 0; d.hr32 = 0; d.delta_nr = p->delta_nr;
    d.seed_lo = (uint32_t)p->rng_seed; d.seed_hi = (uint32_t)(p->rng_seed >> 32);
    d.tc_threshold = p->tc_threshold;
    d.rel_peak = p->relative_peak_thresh;
    d.cos_max_angle = (float)std::cos((double)p->max_angle);
    d.cos_min_sep = (float)std::cos((double)p->min_separation_angle);
    d.min_signal = p->min_signal;
    d.step_d = (double)p->step_size;
    d.step_frac = (p->model_type == CUSLINES_PTT) ? p->step_frac : 1;
    d.probe_quality = p->probe_quality;
    d.lut_res = p->lut_res;
    d.log2x = p->tile_log2_x; d.log2y = p->tile_log2_y;
    d.nx = 1 << p->tile_log2_x; d.ny = 1 << p->tile_log2_y; d.nz = 1 << p->tile_log2_z;
    d.probe_step = p->probe_step_size;
    d.max_curv = p->max_curvature;
    d.min_support_f = smallest_float_at_least((double)p->probe_quality * PMF_THRESHOLD_P);
    d.ptt_seed_thr_f = smallest_float_at_least(PMF_THRESHOLD_P);
    d.ptt_sub_step = (float)(d.step_d / (double)d.step_frac);
    *out = t;
    return CUSLINES_OK;
}

int cuslines_tracker_destroy(cuslines_tracker *t) {
    if (!t) return CUSLINES_OK;
    if (t->nforks) return fail(CUSLINES_ESTATE, "destroy the tracker's forks first");
    if (t->ngraphs) return fail(CUSLINES_ESTATE, "destroy the tracker's chunk graphs first");
    (void)hipSetDevice(t->device);
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    if (t->parent) {  // a fork owns only its stream and counters
        --t->parent->nforks;
        dfree(t->d_work);
        dfree(t->d_stats);
        if (t->stream) (void)hipStreamDestroy(t->stream);
        delete t;
        return CUSLINES_OK;
    }
    dfree(t->d_data); dfree(t->d_stop); dfree(t->d_verts); dfree(t->d_edges); dfree(t->d_nbr);
    dfree(t->d_H); dfree(t->d_R); dfree(t->d_db); dfree(t->d_dq); dfree(t->d_smat);
    dfree(t->d_hrm); dfree(t->d_fitm); dfree(t->d_odfm);
    dfree(t->d_ut); dfree(t->d_u); dfree(t->d_invlev);
    dfree(t->d_b0); dfree(t->d_dwi); dfree(t->d_lut); dfree(t->d_work); dfree(t->d_blocks);
    dfree(t->d_fprop); dfree(t->d_vprop); dfree(t->d_mlp); dfree(t->d_stats);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
    return CUSLINES_OK;
}

int cuslines_tracker_fork(cuslines_tracker *t, cuslines_tracker **out) {
    if (!t || !out) return fail(CUSLINES_EINVAL, "null argument");
    *out = nullptr;
    if (t->parent) return fail(CUSLINES_EINVAL, "fork the parent tracker, not a fork");
    HIPC(hipSetDevice(t->device));
    cuslines_tracker *f = new cuslines_tracker(*t);  // shares every device buffer pointer
    f->parent = t;
    f->nforks = 0;
    f->ngraphs = 0;
    f->stream = nullptr;
    f->d_work = nullptr;
    f->d_stats = nullptr;  // learned statistics and tap counts stay with the parent
    f->rec = nullptr;
    f->rec_count = nullptr;
    f->count_taps = 0;
    hipError_t e = hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&f->d_work, QUEUES * sizeof(int));
    if (e != hipSuccess) {
        dfree(f->d_work);
        if (f->stream) (void)hipStreamDestroy(f->stream);
        delete f;
        return fail(CUSLINES_EHIP, std::string("tracker fork: ") + hipGetErrorString(e));
    }
    ++t->nforks;
    *out = f;
    return CUSLINES_OK;
}

// PTT corner blocks from the transposed ODF and the LUT's direction ranks
// (both must be present).  Without room for them taps gather their 8 texels.
// The low-rank form of the bootstrap's [H; R] products.  H is the hat
// matrix of the signal basis, a symmetric projection of rank r (the basis
// size): H = U U^T with U (hr x r) orthonormal, and R = lcr(H) =
// L (I - H) - 1 mu^T with L = diag(1 / sqrt(1 - H_ii)) (1 where H_ii >= 1)
// and mu the column means of L (I - H) (boot_utils.py:86-94, DIPY's
// lcr_matrix), so R.d = v - mean(v) with v = L (d - H.d).  U comes from
// Gram-Schmidt (twice, in double) over H's columns; the form is used only
// when U U^T and lcr(H) reproduce the given H and R to 1e-5.  Returns the
// rank (0 when the matrices do not have this form), U row-major [hr][r] and
// the leverage factors.
static int hat_lowrank(const float *H, const float *R, int hr, std::vector<double> &U, std::vector<double> &invlev) {
    std::vector<std::vector<double>> basis;
    double hmax = 0.0;
    for (size_t i = 0; i < (size_t)hr * hr; ++i) hmax = std::max(hmax, std::fabs((double)H[i]));
    if (hmax == 0.0) return 0;
    std::vector<double> v(hr);
    for (int j = 0; j < hr && (int)basis.size() < 64; ++j) {
        double n0 = 0.0;
        for (int i = 0; i < hr; ++i) { v[i] = H[(size_t)i * hr + j]; n0 += v[i] * v[i]; }
        if (n0 == 0.0) continue;
        for (int pass = 0; pass < 2; ++pass)
            for (const auto &u : basis) {
                double dot = 0.0;
                for (int i = 0; i < hr; ++i) dot += u[i] * v[i];
                for (int i = 0; i < hr; ++i) v[i] -= dot * u[i];
            }
        double n = 0.0;
        for (int i = 0; i < hr; ++i) n += v[i] * v[i];
        if (n > 1e-6 * n0 && n > 1e-12) {
            const double inv = 1.0 / std::sqrt(n);
            for (int i = 0; i < hr; ++i) v[i] *= inv;
            basis.push_back(v);
        }
    }
    const int r = (int)basis.size();
    if (r == 0 || r >= hr || r > 64) return 0;
    // A column added late can have a small residual, which the float32
    // rounding of H then dominates: refine the basis twice by U <- orth(H U)
    // (H is a projection, so H U lies in its range and is close to
    // orthonormal; Gram-Schmidt on it amplifies nothing)
    for (int it = 0; it < 2; ++it) {
        std::vector<std::vector<double>> nb;
        for (int k = 0; k < r; ++k) {
            for (int i = 0; i < hr; ++i) {
                double acc = 0.0;
                for (int j = 0; j < hr; ++j) acc += (double)H[(size_t)i * hr + j] * basis[k][j];
                v[i] = acc;
            }
            for (int pass = 0; pass < 2; ++pass)
                for (const auto &u : nb) {
                    double dot = 0.0;
                    for (int i = 0; i < hr; ++i) dot += u[i] * v[i];
                    for (int i = 0; i < hr; ++i) v[i] -= dot * u[i];
                }
            double n = 0.0;
            for (int i = 0; i < hr; ++i) n += v[i] * v[i];
            if (n < 0.25) return 0;  // H U lost a direction: not a projection
            const double inv = 1.0 / std::sqrt(n);
            for (int i = 0; i < hr; ++i) v[i] *= inv;
            nb.push_back(v);
        }
        basis.swap(nb);
    }
    // H = U U^T ?
    for (int i = 0; i < hr; ++i)
        for (int j = 0; j < hr; ++j) {
            double acc = 0.0;
            for (int k = 0; k < r; ++k) acc += basis[k][i] * basis[k][j];
            if (std::fabs(acc - (double)H[(size_t)i * hr + j]) > 1e-5 * hmax) return 0;
        }
    // R = lcr(H) ?
    invlev.assign(hr, 1.0);
    for (int i = 0; i < hr; ++i) {
        const double hii = H[(size_t)i * hr + i];
        if (hii < 1.0) invlev[i] = 1.0 / std::sqrt(1.0 - hii);
    }
    double rmax = 1.0;
    for (size_t i = 0; i < (size_t)hr * hr; ++i) rmax = std::max(rmax, std::fabs((double)R[i]));
    for (int j = 0; j < hr; ++j) {
        double mu = 0.0;
        for (int i = 0; i < hr; ++i) mu += ((i == j ? 1.0 : 0.0) - (double)H[(size_t)i * hr + j]) * invlev[i];
        mu /= hr;
        for (int i = 0; i < hr; ++i) {
            const double lij = ((i == j ? 1.0 : 0.0) - (double)H[(size_t)i * hr + j]) * invlev[i] - mu;
            if (std::fabs(lij - (double)R[(size_t)i * hr + j]) > 1e-5 * rmax) return 0;
        }
    }
    U.assign((size_t)hr * r, 0.0);
    for (int i = 0; i < hr; ++i)
        for (int k = 0; k < r; ++k) U[(size_t)i * r + k] = basis[k][i];
    return r;
}

static int build_blocks(cuslines_tracker *t) {
    const DevParams &d = t->dp;
    dfree(t->d_blocks);
    t->d_blocks = nullptr;
    const long long ncells = (long long)d.nverts * (d.dimx + 1) * (d.dimy + 1) * (d.dimz + 1);
    uint16_t *rv = nullptr;
    HIPC(hipMalloc(&rv, t->rank_vertex.size() * sizeof(uint16_t)));
    hipError_t e = hipMemcpy(rv, t->rank_vertex.data(), t->rank_vertex.size() * sizeof(uint16_t),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess && hipMalloc(&t->d_blocks, (size_t)ncells * 8 * sizeof(float)) == hipSuccess) {
        hipLaunchKernelGGL(build_corner_blocks, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, t->stream, d,
                           t->d_data, rv, t->d_blocks, ncells);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    } else if (e == hipSuccess) {
        (void)hipGetLastError();
        t->d_blocks = nullptr;
    }
    dfree(rv);
    if (e != hipSuccess) return fail(CUSLINES_EHIP, std::string("PTT blocks: ") + hipGetErrorString(e));
    return CUSLINES_OK;
}

// PTT: the [X][Y][Z][T] ODF in device buffer `src` (left unchanged) becomes
// the tracker's vertex-major copy; corner blocks follow when the LUT is set.
static int ptt_take_data(cuslines_tracker *t, const float *src, int nonneg) {
    const DevParams &d = t->dp;
    const size_t nvox = (size_t)d.dimx * d.dimy * d.dimz;
    const size_t bytes = nvox * d.dimt * sizeof(float);
    dfree(t->d_data);
    t->d_data = nullptr;
    t->has_data = false;
    t->dp.ptt_nonneg = nonneg;
    if (hipMalloc(&t->d_data, bytes) != hipSuccess) {
        (void)hipGetLastError();
        t->d_data = nullptr;
        return fail(CUSLINES_ENOMEM, "PTT data allocation failed");
    }
    dim3 grid((unsigned)((nvox + 31) / 32), (unsigned)((d.dimt + 31) / 32));
    hipLaunchKernelGGL(transpose_vox_t, grid, dim3(256), 0, t->stream, src, t->d_data, (long long)nvox, d.dimt);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) return fail(CUSLINES_EHIP, std::string("PTT data transpose: ") + hipGetErrorString(e));
    t->has_data = true;
    if (t->has_lut) return build_blocks(t);
    return CUSLINES_OK;
}

int cuslines_set_data(cuslines_tracker *t, const float *dataf) {
    RoctxRange range_("cuslines_set_data");
    if (!t || !dataf) return fail(CUSLINES_EINVAL, "null argument");
    if (int rc = check_owner(t)) return rc;
    HIPC(hipSetDevice(t->device));
    const DevParams &d = t->dp;
    const size_t nvox = (size_t)d.dimx * d.dimy * d.dimz;
    dfree(t->d_data);
    t->d_data = nullptr;
    t->has_data = false;
    if (t->p.model_type == CUSLINES_PTT) {
        // The early stop of rejection-round probes (data_support) relies on
        // taps >= 0; the reference's prepare_data clips the ODF at 0.
        const size_t nval = nvox * (size_t)d.dimt;
        // a NaN counts as negative: a NaN tap makes the full support NaN,
        // which fails the acceptance test, so no early stop may skip it
        const int nonneg = std::none_of(dataf, dataf + nval, [](float v) { return !(v >= 0.0f); }) ? 1 : 0;
        float *tmp = nullptr;
        HIPC(hipMalloc(&tmp, nval * sizeof(float)));
        hipError_t e = hipMemcpy(tmp, dataf, nval * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) { dfree(tmp); return fail(CUSLINES_EHIP, std::string("PTT data upload: ") + hipGetErrorString(e)); }
        const int rc = ptt_take_data(t, tmp, nonneg);
        dfree(tmp);
        return rc;
    } else {
        const size_t pitch = (size_t)d.t32 * sizeof(float);
        HIPC(hipMalloc(&t->d_data, nvox * pitch));
        HIPC(hipMemset(t->d_data, 0, nvox * pitch));
        HIPC(hipMemcpy2D(t->d_data, pitch, dataf, (size_t)d.dimt * sizeof(float), (size_t)d.dimt * sizeof(float),
                         nvox, hipMemcpyHostToDevice));
    }
    t->has_data = true;
    return CUSLINES_OK;
}

int cuslines_set_stop_map(cuslines_tracker *t, const float *stop_map) {
    if (!t || !stop_map) return fail(CUSLINES_EINVAL, "null argument");
    if (int rc = check_owner(t)) return rc;
    HIPC(hipSetDevice(t->device));
    const size_t n = (size_t)t->dp.dimx * t->dp.dimy * t->dp.dimz;
    // A failed re-upload must not leave the tracker ready with freed buffers.
    t->has_stop = false;
    dfree(t->d_stop);
    t->d_stop = nullptr;
    HIPC(hipMalloc(&t->d_stop, n * sizeof(float)));
    HIPC(hipMemcpy(t->d_stop, stop_map, n * sizeof(float), hipMemcpyHostToDevice));
    t->has_stop = true;
    return CUSLINES_OK;
}

int cuslines_set_sphere(cuslines_tracker *t, const float *vertices, const int32_t *edges) {
    if (!t || !vertices || !edges) return fail(CUSLINES_EINVAL, "null argument");
    if (int rc = check_owner(t)) return rc;
    HIPC(hipSetDevice(t->device));
    const int nv = t->dp.nverts, ne = t->dp.nedges;
    std::vector<uint32_t> packed(ne);
    for (int e = 0; e < ne; ++e) {
        const int u = edges[2 * e], v = edges[2 * e + 1];
        if (u < 0 || u >= nv || v < 0 || v >= nv) return fail(CUSLINES_EINVAL, "edge vertex index out of range");
        packed[e] = (uint32_t)u | ((uint32_t)v << 16);
    }
    t->has_sphere = false;
    dfree(t->d_verts); dfree(t->d_edges); dfree(t->d_nbr);
    t->d_verts = nullptr; t->d_edges = nullptr; t->d_nbr = nullptr;
    HIPC(hipMalloc(&t->d_verts, (size_t)nv * 3 * sizeof(float)));
    HIPC(hipMemcpy(t->d_verts, vertices, (size_t)nv * 3 * sizeof(float), hipMemcpyHostToDevice));
    HIPC(hipMalloc(&t->d_edges, (size_t)ne * sizeof(uint32_t)));
    HIPC(hipMemcpy(t->d_edges, packed.data(), (size_t)ne * sizeof(uint32_t), hipMemcpyHostToDevice));
    // Adjacency for the peak test (peak_directions): up to 8 distinct
    // neighbours per vertex, the rest of the row the vertex itself.
    std::vector<uint16_t> nbr((size_t)nv * 8);
    std::vector<int> deg(nv, 0);
    bool fits = nv <= 1024;
    for (int v = 0; v < nv; ++v)
        for (int k = 0; k < 8; ++k) nbr[(size_t)v * 8 + k] = (uint16_t)v;
    for (int e = 0; e < ne && fits; ++e) {
        const int a = edges[2 * e], b = edges[2 * e + 1];
        for (int side = 0; side < 2 && fits; ++side) {
            const int x = side ? b : a, y = side ? a : b;
            bool seen = false;
            for (int k = 0; k < deg[x]; ++k) seen = seen || nbr[(size_t)x * 8 + k] == y;
            if (seen) continue;
            if (deg[x] == 8) { fits = false; break; }
            nbr[(size_t)x * 8 + deg[x]++] = (uint16_t)y;
        }
    }
    if (fits) {
        HIPC(hipMalloc(&t->d_nbr, nbr.size() * sizeof(uint16_t)));
        HIPC(hipMemcpy(t->d_nbr, nbr.data(), nbr.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
    t->has_sphere = true;
    return CUSLINES_OK;
}

int cuslines_set_boot_matrices(cuslines_tracker *t, const float *H, const float *R, const float *delta_b,
                               const float *delta_q, const float *sampling_matrix, const int32_t *b0s_mask) {
    if (!t || !H || !R || !delta_b || !delta_q || !sampling_matrix || !b0s_mask)
        return fail(CUSLINES_EINVAL, "null argument");
    if (int rc = check_owner(t)) return rc;
    if (t->p.model_type != CUSLINES_OPDT && t->p.model_type != CUSLINES_CSA)
        return fail(CUSLINES_ESTATE, "bootstrap matrices need an OPDT or CSA tracker");
    HIPC(hipSetDevice(t->device));
    const int T = t->dp.dimt, dn = t->dp.delta_nr, nv = t->dp.nverts;
    std::vector<int> b0(T), dwi;
    int nb0 = 0;
    for (int j = 0; j < T; ++j) {
        b0[j] = b0s_mask[j] ? 1 : 0;
        if (b0[j]) ++nb0; else dwi.push_back(j);
    }
    const int hr = (int)dwi.size();
    if (hr == 0 || nb0 == 0) return fail(CUSLINES_EINVAL, "bootstrap needs at least one b0 and one DWI volume");
    t->has_boot = false;
    t->dp.hr = hr;
    t->dp.hr32 = ceil32(hr);
    float **dst[5] = {&t->d_H, &t->d_R, &t->d_db, &t->d_dq, &t->d_smat};
    const float *src[5] = {H, R, delta_b, delta_q, sampling_matrix};
    const int rows[5] = {hr, hr, dn, dn, nv}, cols[5] = {hr, hr, hr, hr, dn};
    for (int k = 0; k < 5; ++k) {
        // gemv_lanerow's layout: lane-interleaved when the row length is a
        // multiple of 4 (rows padded to a multiple of 32 with zeros), else row-major
        const int N = rows[k], M = cols[k];
        std::vector<float> h;
        if ((M & 3) == 0) {
            h.assign((size_t)ceil32(N) * M, 0.0f);
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < M; ++j) h[mat_offset(M, i, j)] = src[k][(size_t)i * M + j];
        } else {
            h.assign(src[k], src[k] + (size_t)N * M);
        }
        dfree(*dst[k]);
        *dst[k] = nullptr;
        HIPC(hipMalloc(dst[k], h.size() * sizeof(float)));
        HIPC(hipMemcpy(*dst[k], h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    // the batched kernel's zero-padded row-major copies: [H; R] [2 hr16][hr16],
    // [dq | -db] [dn16][2 hr16], sampling matrix [nv16][dn16]
    {
        const int hr16 = (hr + 15) & ~15, dn16 = (dn + 15) & ~15, nv16 = (nv + 15) & ~15;
        t->dp.hr16 = hr16; t->dp.dn16 = dn16; t->dp.nv16 = nv16;
        std::vector<float> hrm((size_t)2 * hr16 * hr16, 0.0f), fit((size_t)dn16 * 2 * hr16, 0.0f),
            odf((size_t)nv16 * dn16, 0.0f);
        for (int i = 0; i < hr; ++i)
            for (int j = 0; j < hr; ++j) {
                hrm[(size_t)i * hr16 + j] = H[(size_t)i * hr + j];
                hrm[(size_t)(hr16 + i) * hr16 + j] = R[(size_t)i * hr + j];
            }
        for (int i = 0; i < dn; ++i)
            for (int j = 0; j < hr; ++j) {
                fit[(size_t)i * 2 * hr16 + j] = delta_q[(size_t)i * hr + j];
                fit[(size_t)i * 2 * hr16 + hr16 + j] = -delta_b[(size_t)i * hr + j];
            }
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < dn; ++j) odf[(size_t)i * dn16 + j] = sampling_matrix[(size_t)i * dn + j];
        // low-rank form: U^T [rk16][hr16], U [hr16][rk16], 1/lev [hr16], zero padded
        std::vector<double> U, invlev;
        const int r = hat_lowrank(H, R, hr, U, invlev);
        const int rk16 = (r + 15) & ~15;
        std::vector<float> ut((size_t)std::max(rk16, 16) * hr16, 0.0f), um((size_t)hr16 * std::max(rk16, 16), 0.0f),
            il((size_t)hr16, 0.0f);
        for (int i = 0; i < hr && r > 0; ++i) {
            for (int k = 0; k < r; ++k) {
                ut[(size_t)k * hr16 + i] = (float)U[(size_t)i * r + k];
                um[(size_t)i * rk16 + k] = (float)U[(size_t)i * r + k];
            }
            il[i] = (float)invlev[i];
        }
        t->rk16 = r > 0 ? rk16 : 0;
        t->hrank = r;
        float **bdst[6] = {&t->d_hrm, &t->d_fitm, &t->d_odfm, &t->d_ut, &t->d_u, &t->d_invlev};
        const std::vector<float> *bsrc[6] = {&hrm, &fit, &odf, &ut, &um, &il};
        for (int k = 0; k < 6; ++k) {
            dfree(*bdst[k]);
            *bdst[k] = nullptr;
            HIPC(hipMalloc(bdst[k], bsrc[k]->size() * sizeof(float)));
            HIPC(hipMemcpy(*bdst[k], bsrc[k]->data(), bsrc[k]->size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    dfree(t->d_b0); dfree(t->d_dwi);
    t->d_b0 = nullptr; t->d_dwi = nullptr;
    HIPC(hipMalloc(&t->d_b0, (size_t)T * sizeof(int)));
    HIPC(hipMemcpy(t->d_b0, b0.data(), (size_t)T * sizeof(int), hipMemcpyHostToDevice));
    HIPC(hipMalloc(&t->d_dwi, (size_t)hr * sizeof(int)));
    HIPC(hipMemcpy(t->d_dwi, dwi.data(), (size_t)hr * sizeof(int), hipMemcpyHostToDevice));
    t->has_boot = true;
    return CUSLINES_OK;
}

int cuslines_set_ptt_lut(cuslines_tracker *t, const int32_t *lut) {
    if (!t || !lut) return fail(CUSLINES_EINVAL, "null argument");
    if (int rc = check_owner(t)) return rc;
    if (t->p.model_type != CUSLINES_PTT) return fail(CUSLINES_ESTATE, "LUT needs a PTT tracker");
    HIPC(hipSetDevice(t->device));
    const size_t r = (size_t)t->dp.lut_res;
    const int nv = t->dp.nverts;
    std::vector<uint32_t> h(r * r * r);
    // Direction ranks: vertices numbered in the order they first appear along
    // a Morton traversal of the LUT grid, so consecutive ranks are
    // neighbouring directions; the corner blocks of a cell are stored in this
    // order (tap_sample).
    std::vector<int> rank(nv, -1);
    std::vector<uint16_t> rank_vertex;
    rank_vertex.reserve(nv);
    {
        int bits = 0;
        while ((size_t(1) << bits) < r) ++bits;
        const uint64_t ncode = uint64_t(1) << (3 * bits);
        for (uint64_t c = 0; c < ncode; ++c) {
            uint64_t a[3] = {0, 0, 0};
            for (int b = 0; b < bits; ++b)
                for (int k = 0; k < 3; ++k) a[k] |= ((c >> (3 * b + 2 - k)) & 1u) << b;
            if (a[0] >= r || a[1] >= r || a[2] >= r) continue;
            const int32_t v = lut[(a[0] * r + a[1]) * r + a[2]];
            if (v < 0 || v >= nv) return fail(CUSLINES_EINVAL, "LUT vertex index out of range");
            if (rank[v] < 0) { rank[v] = (int)rank_vertex.size(); rank_vertex.push_back((uint16_t)v); }
        }
        for (int v = 0; v < nv; ++v)
            if (rank[v] < 0) { rank[v] = (int)rank_vertex.size(); rank_vertex.push_back((uint16_t)v); }
    }
    // Unit directions touch a thin shell of the cube: with res % 4 == 0 the
    // table is stored as 4x4x4 bricks of 128 bytes (one cache line), so the
    // shell spans ~10x fewer lines than with rows along the last axis.
    const bool brick = (r % 4) == 0;
    const size_t nb = r / 4;
    for (size_t i0 = 0; i0 < r; ++i0)
        for (size_t i1 = 0; i1 < r; ++i1)
            for (size_t i2 = 0; i2 < r; ++i2) {
                const int32_t v = lut[(i0 * r + i1) * r + i2];
                const size_t dst = brick ? (((i0 >> 2) * nb + (i1 >> 2)) * nb + (i2 >> 2)) * 64 +
                                               ((i0 & 3) << 4) + ((i1 & 3) << 2) + (i2 & 3)
                                         : (i0 * r + i1) * r + i2;
                h[dst] = (uint32_t)v | ((uint32_t)rank[v] << 16);
            }
    t->has_lut = false;
    t->dp.lut_brick = brick ? 1 : 0;
    dfree(t->d_lut);
    t->d_lut = nullptr;
    HIPC(hipMalloc(&t->d_lut, h.size() * sizeof(uint32_t)));
    HIPC(hipMemcpy(t->d_lut, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    t->rank_vertex.swap(rank_vertex);
    if (!t->d_vprop) {
        HIPC(hipMalloc(&t->d_vprop, VPROP_FLOATS * sizeof(float)));
        hipLaunchKernelGGL(feature_propagators, dim3(1), dim3(32), 0, t->stream, t->dp, 1, t->d_vprop);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(t->stream));
    }
    t->has_lut = true;
    if (t->has_data) return build_blocks(t);
    return CUSLINES_OK;
}

int cuslines_device_alloc(cuslines_tracker *t, size_t bytes, void **dptr) {
    if (!t || !dptr) return fail(CUSLINES_EINVAL, "null argument");
    HIPC(hipSetDevice(t->device));
    *dptr = nullptr;
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) return fail(CUSLINES_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return CUSLINES_OK;
}

int cuslines_device_free(cuslines_tracker *t, void *dptr) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    HIPC(hipSetDevice(t->device));
    if (dptr) HIPC(hipFree(dptr));
    return CUSLINES_OK;
}

int cuslines_memcpy_htod(cuslines_tracker *t, void *dst, const void *src, size_t bytes) {
    if (!t || (bytes && (!dst || !src))) return fail(CUSLINES_EINVAL, "null argument");
    HIPC(hipSetDevice(t->device));
    if (bytes) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, t->stream));
    HIPC(hipStreamSynchronize(t->stream));
    return CUSLINES_OK;
}

int cuslines_memcpy_dtoh(cuslines_tracker *t, void *dst, const void *src, size_t bytes) {
    if (!t || (bytes && (!dst || !src))) return fail(CUSLINES_EINVAL, "null argument");
    HIPC(hipSetDevice(t->device));
    if (bytes) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, t->stream));
    HIPC(hipStreamSynchronize(t->stream));
    return CUSLINES_OK;
}

int cuslines_synchronize(cuslines_tracker *t) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    HIPC(hipSetDevice(t->device));
    HIPC(hipStreamSynchronize(t->stream));
    return CUSLINES_OK;
}

int cuslines_set_launch(cuslines_tracker *t, int schedule, int threads, int blocks_per_cu) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (schedule != CUSLINES_SCHED_QUEUE && schedule != CUSLINES_SCHED_STATIC)
        return fail(CUSLINES_EINVAL, "unknown schedule");
    if (threads != 0 && (threads < 64 || threads > 256 || threads % 64 != 0))
        return fail(CUSLINES_EINVAL, "threads per block must be 0 or a multiple of 64 in [64, 256]");
    if (blocks_per_cu < 0) return fail(CUSLINES_EINVAL, "blocks_per_cu must be >= 0");
    t->schedule = schedule;
    t->threads = threads;
    t->blocks_per_cu = blocks_per_cu;
    return CUSLINES_OK;
}

int cuslines_set_boot_mode(cuslines_tracker *t, int mode) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (t->p.model_type != CUSLINES_OPDT && t->p.model_type != CUSLINES_CSA)
        return fail(CUSLINES_EINVAL, "the bootstrap mode is for OPDT and CSA trackers");
    if ((mode & 0xff) != CUSLINES_BOOT_EXACT && (mode & 0xff) != CUSLINES_BOOT_MFMA)
        return fail(CUSLINES_EINVAL, "unknown mode");
    const int occ = (mode >> 8) & 0xff;
    if (occ != 0 && occ != 2 && occ != 4) return fail(CUSLINES_EINVAL, "waves per SIMD must be 2 or 4");
    if ((mode & ~(0xffff | CUSLINES_BOOT_FULL_PRODUCTS)) != 0) return fail(CUSLINES_EINVAL, "unknown mode flag");
    t->boot_mode = mode & 0xff;
    t->boot_occ = occ ? occ : 4;
    t->boot_full = (mode & CUSLINES_BOOT_FULL_PRODUCTS) ? 1 : 0;
    return CUSLINES_OK;
}

int cuslines_hat_lowrank(const float *H, const float *R, int32_t hr, int32_t *rank) {
    if (!H || !R || !rank) return fail(CUSLINES_EINVAL, "null argument");
    if (hr < 1) return fail(CUSLINES_EINVAL, "hr must be positive");
    std::vector<double> U, invlev;
    *rank = hat_lowrank(H, R, hr, U, invlev);
    return CUSLINES_OK;
}

int cuslines_boot_lowrank(cuslines_tracker *t, int32_t *rank) {
    if (!t || !rank) return fail(CUSLINES_EINVAL, "null argument");
    *rank = (t->has_boot && !t->boot_full) ? t->hrank : 0;
    return CUSLINES_OK;
}

int cuslines_set_stepping(cuslines_tracker *t, int flat) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (flat != 0 && flat != 1) return fail(CUSLINES_EINVAL, "stepping must be 0 (per seed) or 1 (flat)");
    t->flat = flat;
    return CUSLINES_OK;
}

int cuslines_set_lanes_per_seed(cuslines_tracker *t, int lanes) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (lanes != 0 && lanes != 8 && lanes != 16 && lanes != 32)
        return fail(CUSLINES_EINVAL, "lanes per seed must be 0, 8, 16 or 32");
    t->ptt_lanes = lanes;
    return CUSLINES_OK;
}

int cuslines_ptt_prepare(cuslines_tracker *t, const float *odf, const float *stop_map, float stop_threshold,
                         int sphere_symm, int32_t *lut_out) {
    RoctxRange range_("cuslines_ptt_prepare");
    if (!t || !odf || !stop_map) return fail(CUSLINES_EINVAL, "null argument");
    if (int rc = check_owner(t)) return rc;
    if (t->p.model_type != CUSLINES_PTT) return fail(CUSLINES_ESTATE, "PTT preparation needs a PTT tracker");
    if (!t->has_sphere) return fail(CUSLINES_ESTATE, "upload the sphere before preparing PTT data");
    if (t->dp.lut_res < 2) return fail(CUSLINES_EINVAL, "lut_res must be at least 2");
    HIPC(hipSetDevice(t->device));
    const DevParams &d = t->dp;
    const size_t nvox = (size_t)d.dimx * d.dimy * d.dimz;
    const size_t nval = nvox * (size_t)d.dimt;
    const size_t nlut = (size_t)d.lut_res * d.lut_res * d.lut_res;
    float *raw = nullptr, *stop = nullptr;
    int32_t *lut = nullptr;
    hipError_t e = hipMalloc(&raw, nval * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&stop, nvox * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&lut, nlut * sizeof(int32_t));
    if (e == hipSuccess) e = hipMemcpy(raw, odf, nval * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(stop, stop_map, nvox * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(ptt_normalise, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, t->stream, raw, stop,
                           stop_threshold, (long long)nvox, d.dimt);
        hipLaunchKernelGGL(ptt_lut_kernel, dim3((unsigned)((nlut + 255) / 256)), dim3(256), 0, t->stream, t->d_verts,
                           d.nverts, d.lut_res, sphere_symm ? 1 : 0, lut);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    std::vector<int32_t> hl(nlut);
    if (e == hipSuccess) e = hipMemcpy(hl.data(), lut, nlut * sizeof(int32_t), hipMemcpyDeviceToHost);
    dfree(stop);
    dfree(lut);
    if (e != hipSuccess) {
        dfree(raw);
        return fail(CUSLINES_EHIP, std::string("PTT preparation: ") + hipGetErrorString(e));
    }
    // the clipped ODF has no negative value, but NaN inputs stay NaN
    int nonneg = 0;
    {
        int *flag = nullptr;
        e = hipMalloc(&flag, sizeof(int));
        if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(int), t->stream);
        if (e == hipSuccess) {
            const size_t nb = std::min<size_t>((nval + 255) / 256, 65536);
            hipLaunchKernelGGL(flag_not_nonneg, dim3((unsigned)nb), dim3(256), 0, t->stream, raw, (long long)nval, flag);
            e = hipGetLastError();
        }
        int hf = 1;
        if (e == hipSuccess) e = hipMemcpyAsync(&hf, flag, sizeof(int), hipMemcpyDeviceToHost, t->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
        dfree(flag);
        if (e != hipSuccess) {
            dfree(raw);
            return fail(CUSLINES_EHIP, std::string("PTT preparation: ") + hipGetErrorString(e));
        }
        nonneg = hf ? 0 : 1;
    }
    int rc = ptt_take_data(t, raw, nonneg);
    dfree(raw);
    if (rc) return rc;
    rc = cuslines_set_ptt_lut(t, hl.data());
    if (rc) return rc;
    if (lut_out) std::memcpy(lut_out, hl.data(), nlut * sizeof(int32_t));
    return CUSLINES_OK;
}

int cuslines_set_ptt_layout(cuslines_tracker *t, int corner_blocks) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (int rc = check_owner(t)) return rc;
    t->use_blocks = corner_blocks ? 1 : 0;
    return CUSLINES_OK;
}

static int ensure_feature_propagators(cuslines_tracker *t) {
    if (t->d_fprop) return CUSLINES_OK;
    HIPC(hipMalloc(&t->d_fprop, DISC_VERT_CNT * 9 * sizeof(float)));
    hipLaunchKernelGGL(feature_propagators, dim3(1), dim3(32), 0, t->stream, t->dp, FEAT_STEPS, t->d_fprop);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(t->stream));
    return CUSLINES_OK;
}

int cuslines_set_proposal(cuslines_tracker *t, const uint16_t *mlp_bf16) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (int rc = check_owner(t)) return rc;
    HIPC(hipSetDevice(t->device));
    if (!mlp_bf16) {
        dfree(t->d_mlp);
        t->d_mlp = nullptr;
        return CUSLINES_OK;
    }
    if (t->p.model_type != CUSLINES_PTT) return fail(CUSLINES_EINVAL, "the learned proposal is for PTT");
    if (int rc = ensure_feature_propagators(t)) return rc;
    std::vector<uint16_t> il(IL_NPARAM, 0);
    auto put = [&](int dst, int src, int N, int K) {
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < K; ++k) il[dst + mlp_il_offset(K, j, k)] = mlp_bf16[src + j * K + k];
    };
    put(IL_W1, MLP_W1, MLP_H, MLP_IN);
    put(IL_W2, MLP_W2, MLP_H, MLP_H);
    put(IL_W3, MLP_W3, MLP_OUT, MLP_H);
    for (int j = 0; j < MLP_H; ++j) { il[IL_B1 + j] = mlp_bf16[MLP_B1 + j]; il[IL_B2 + j] = mlp_bf16[MLP_B2 + j]; }
    for (int j = 0; j < MLP_OUT; ++j) il[IL_B3 + j] = mlp_bf16[MLP_B3 + j];
    if (!t->d_mlp) HIPC(hipMalloc(&t->d_mlp, IL_NPARAM * sizeof(uint16_t)));
    HIPC(hipMemcpy(t->d_mlp, il.data(), IL_NPARAM * sizeof(uint16_t), hipMemcpyHostToDevice));
    if (!t->d_stats) {
        HIPC(hipMalloc(&t->d_stats, 4 * sizeof(unsigned long long)));
        HIPC(hipMemset(t->d_stats, 0, 4 * sizeof(unsigned long long)));
    }
    return CUSLINES_OK;
}

int cuslines_set_recording(cuslines_tracker *t, float *pairs_d, int32_t capacity, int32_t *count_d, float rate) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (int rc = check_owner(t)) return rc;
    if (!pairs_d) {
        t->rec = nullptr;
        t->rec_count = nullptr;
        return CUSLINES_OK;
    }
    if (!count_d || capacity <= 0 || !(rate > 0.0f && rate <= 1.0f))
        return fail(CUSLINES_EINVAL, "recording needs count_d, capacity > 0 and 0 < rate <= 1");
    if (t->p.model_type != CUSLINES_PTT) return fail(CUSLINES_EINVAL, "recording is for PTT");
    if (int rc = ensure_feature_propagators(t)) return rc;
    t->rec = pairs_d;
    t->rec_count = count_d;
    t->dp.rec_cap = capacity;
    const double th = (double)rate * 4294967296.0;
    t->dp.rec_thresh = (th >= 4294967295.0) ? 0xffffffffu : (uint32_t)th;
    return CUSLINES_OK;
}

int cuslines_proposal_stats(cuslines_tracker *t, int64_t *out3, int reset) {
    if (!t || !out3) return fail(CUSLINES_EINVAL, "null argument");
    HIPC(hipSetDevice(t->device));
    unsigned long long h[4] = {0, 0, 0, 0};
    if (t->d_stats) {
        HIPC(hipStreamSynchronize(t->stream));
        HIPC(hipMemcpy(h, t->d_stats, sizeof(h), hipMemcpyDeviceToHost));
        if (reset) HIPC(hipMemset(t->d_stats, 0, sizeof(h)));
    }
    for (int i = 0; i < 3; ++i) out3[i] = (int64_t)h[i];
    return CUSLINES_OK;
}

int cuslines_proposal_counters(cuslines_tracker *t, int64_t *out4, int reset) {
    if (!t || !out4) return fail(CUSLINES_EINVAL, "null argument");
    HIPC(hipSetDevice(t->device));
    unsigned long long h[4] = {0, 0, 0, 0};
    if (t->d_stats) {
        HIPC(hipStreamSynchronize(t->stream));
        HIPC(hipMemcpy(h, t->d_stats, sizeof(h), hipMemcpyDeviceToHost));
        if (reset) HIPC(hipMemset(t->d_stats, 0, sizeof(h)));
    }
    for (int i = 0; i < 4; ++i) out4[i] = (int64_t)h[i];
    return CUSLINES_OK;
}

int cuslines_boot_counters(cuslines_tracker *t, int64_t *out2, int reset) {
    if (!t || !out2) return fail(CUSLINES_EINVAL, "null argument");
    if (t->p.model_type != CUSLINES_OPDT && t->p.model_type != CUSLINES_CSA)
        return fail(CUSLINES_EINVAL, "bootstrap counters are for OPDT and CSA trackers");
    HIPC(hipSetDevice(t->device));
    unsigned long long h[4] = {0, 0, 0, 0};
    if (t->d_stats) {
        HIPC(hipStreamSynchronize(t->stream));
        HIPC(hipMemcpy(h, t->d_stats, sizeof(h), hipMemcpyDeviceToHost));
        if (reset) HIPC(hipMemset(t->d_stats, 0, sizeof(h)));
    }
    out2[0] = (int64_t)h[0];
    out2[1] = (int64_t)h[1];
    return CUSLINES_OK;
}

int cuslines_set_tap_count(cuslines_tracker *t, int enable) {
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (t->p.model_type != CUSLINES_PTT) return fail(CUSLINES_EINVAL, "tap counting is for PTT");
    // Switched off: the instrumented phase-2 kernel stopped responding on
    // its first GPU run (48 seeds, 32x36x30 phantom) and the cause has not
    // been found, so it is never launched.
    if (enable) return fail(CUSLINES_ESTATE, "the instrumented PTT kernel (tap counting) is switched off");
    HIPC(hipSetDevice(t->device));
    if (enable && !t->d_stats) {
        HIPC(hipMalloc(&t->d_stats, 4 * sizeof(unsigned long long)));
        HIPC(hipMemset(t->d_stats, 0, 4 * sizeof(unsigned long long)));
    }
    t->count_taps = enable ? 1 : 0;
    return CUSLINES_OK;
}

int cuslines_tap_count(cuslines_tracker *t, int64_t *out2, int reset) {
    if (!t || !out2) return fail(CUSLINES_EINVAL, "null argument");
    if (t->d_mlp) return fail(CUSLINES_ESTATE, "the counters hold the learned proposal's statistics");
    HIPC(hipSetDevice(t->device));
    unsigned long long h[4] = {0, 0, 0, 0};
    if (t->d_stats) {
        HIPC(hipStreamSynchronize(t->stream));
        HIPC(hipMemcpy(h, t->d_stats, sizeof(h), hipMemcpyDeviceToHost));
        if (reset) HIPC(hipMemset(t->d_stats, 0, sizeof(h)));
    }
    out2[0] = (int64_t)h[0];
    out2[1] = (int64_t)h[1];
    return CUSLINES_OK;
}

// H1 H2 dH2 dH1, O dO, loss; then the weight-gradient partials of the largest layer
static size_t mlp_ws_floats(int n) {
    const size_t nchunk = (size_t)((n + WGRAD_CHUNK - 1) / WGRAD_CHUNK);
    return (size_t)n * (4 * MLP_H + 2 * MLP_OUT + 1) + nchunk * MLP_H * (MLP_H + 1);
}

// The weight-gradient partials take one grid row per WGRAD_CHUNK pairs, and a
// grid has at most 65535 rows.
constexpr int MLP_MAX_PAIRS = 65535 * WGRAD_CHUNK;

int cuslines_mlp_workspace_bytes(int32_t n, size_t *bytes) {
    if (!bytes || n <= 0) return fail(CUSLINES_EINVAL, "n > 0 and bytes required");
    if (n > MLP_MAX_PAIRS) return fail(CUSLINES_EINVAL, "at most 8,388,480 pairs per training call");
    *bytes = mlp_ws_floats(n) * sizeof(float);
    return CUSLINES_OK;
}

int cuslines_mlp_loss_grad_async(cuslines_tracker *t, const float *params_d, const float *pairs_d, int32_t n,
                                 float *loss_d, float *grad_d, void *work_d) {
    if (!t || !params_d || !pairs_d || !loss_d || !grad_d || !work_d || n <= 0)
        return fail(CUSLINES_EINVAL, "null argument or n <= 0");
    if (n > MLP_MAX_PAIRS) return fail(CUSLINES_EINVAL, "at most 8,388,480 pairs per training call");
    HIPC(hipSetDevice(t->device));
    float *H1 = static_cast<float *>(work_d);
    float *H2 = H1 + (size_t)n * MLP_H;
    float *O = H2 + (size_t)n * MLP_H;
    float *dO = O + (size_t)n * MLP_OUT;
    float *dH2 = dO + (size_t)n * MLP_OUT;
    float *dH1 = dH2 + (size_t)n * MLP_H;
    float *ls = dH1 + (size_t)n * MLP_H;
    float *part = ls + n;
    const unsigned nchunk = (unsigned)((n + WGRAD_CHUNK - 1) / WGRAD_CHUNK);
    const float *X = pairs_d;
    const int LD = CUSLINES_PAIR_FLOATS;
    auto g1 = [](long long cnt) { return dim3((unsigned)((cnt + 255) / 256)); };
    hipStream_t st = t->stream;
    hipLaunchKernelGGL((mlp_fwd_layer<MLP_IN, MLP_H, true>), g1((long long)n * MLP_H), dim3(256), 0, st, params_d,
                       MLP_W1, MLP_B1, X, LD, n, H1);
    hipLaunchKernelGGL((mlp_fwd_layer<MLP_H, MLP_H, true>), g1((long long)n * MLP_H), dim3(256), 0, st, params_d,
                       MLP_W2, MLP_B2, H1, MLP_H, n, H2);
    hipLaunchKernelGGL((mlp_fwd_layer<MLP_H, MLP_OUT, false>), g1((long long)n * MLP_OUT), dim3(256), 0, st,
                       params_d, MLP_W3, MLP_B3, H2, MLP_H, n, O);
    hipLaunchKernelGGL(mlp_loss_out, g1(n), dim3(256), 0, st, O, X, n, dO, ls);
    auto wgrad = [&](auto part_k, auto sum_k, int tiles, const float *dout, const float *in, int ld, int w, int b) {
        hipLaunchKernelGGL(part_k, dim3((unsigned)((tiles + 255) / 256), nchunk), dim3(256), 0, st, dout, in, ld, n,
                           part);
        hipLaunchKernelGGL(sum_k, g1(tiles), dim3(256), 0, st, part, (int)nchunk, grad_d, w, b);
    };
    wgrad(mlp_wgrad_part<MLP_OUT, MLP_H>, mlp_wgrad_sum<MLP_OUT, MLP_H>, MLP_OUT * (MLP_H + 1), dO, H2, MLP_H, MLP_W3,
          MLP_B3);
    hipLaunchKernelGGL((mlp_bwd_input<MLP_OUT, MLP_H>), g1((long long)n * MLP_H), dim3(256), 0, st, params_d, MLP_W3,
                       dO, H2, n, dH2);
    wgrad(mlp_wgrad_part<MLP_H, MLP_H>, mlp_wgrad_sum<MLP_H, MLP_H>, MLP_H * (MLP_H + 1), dH2, H1, MLP_H, MLP_W2,
          MLP_B2);
    hipLaunchKernelGGL((mlp_bwd_input<MLP_H, MLP_H>), g1((long long)n * MLP_H), dim3(256), 0, st, params_d, MLP_W2,
                       dH2, H1, n, dH1);
    wgrad(mlp_wgrad_part<MLP_H, MLP_IN>, mlp_wgrad_sum<MLP_H, MLP_IN>, MLP_H * (MLP_IN + 1), dH1, X, LD, MLP_W1,
          MLP_B1);
    hipLaunchKernelGGL(mean_kernel, dim3(1), dim3(256), 0, st, ls, n, loss_d);
    HIPC(hipGetLastError());
    return CUSLINES_OK;
}

int cuslines_mlp_adam_async(cuslines_tracker *t, float *params_d, const float *grad_d, float *m_d, float *v_d,
                            int32_t step, float lr, float beta1, float beta2, float eps) {
    if (!t || !params_d || !grad_d || !m_d || !v_d || step < 1) return fail(CUSLINES_EINVAL, "bad Adam arguments");
    HIPC(hipSetDevice(t->device));
    const float c1 = (float)(1.0 - std::pow((double)beta1, (double)step));
    const float c2 = (float)(1.0 - std::pow((double)beta2, (double)step));
    hipLaunchKernelGGL(adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, t->stream, params_d, grad_d, m_d,
                       v_d, MLP_NPARAM, lr, beta1, beta2, eps, c1, c2);
    HIPC(hipGetLastError());
    return CUSLINES_OK;
}

int cuslines_get_stream(cuslines_tracker *t, void **stream) {
    if (!t || !stream) return fail(CUSLINES_EINVAL, "null argument");
    *stream = (void *)t->stream;
    return CUSLINES_OK;
}

int cuslines_phase1_async(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk,
                          int32_t *counts_d, int32_t *dirs_d) {
    RoctxRange range_("cuslines_phase1");
    if (!t) return fail(CUSLINES_EINVAL, "null tracker");
    if (nseeds < 0 || rng_chunk < 0) return fail(CUSLINES_EINVAL, "negative size");
    if (nseeds == 0) return CUSLINES_OK;
    if (!seeds_d || !counts_d || !dirs_d) return fail(CUSLINES_EINVAL, "null buffer");
    int rc = check_ready(t);
    if (rc) return rc;
    HIPC(hipSetDevice(t->device));
    return phase1_models(t, seeds_d, nseeds, rng_chunk > 0 ? rng_chunk : nseeds, counts_d, dirs_d);
}

int cuslines_scan_async(cuslines_tracker *t, const int32_t *counts_d, int32_t nseeds, int32_t *offsets_d) {
    RoctxRange range_("cuslines_scan");
    if (!t || !offsets_d || (nseeds > 0 && !counts_d)) return fail(CUSLINES_EINVAL, "null argument");
    if (nseeds < 0) return fail(CUSLINES_EINVAL, "negative size");
    HIPC(hipSetDevice(t->device));
    return run_scan<int>(t, counts_d, nseeds, offsets_d);
}

int cuslines_phase2_async(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk,
                          const int32_t *offsets_d, const int32_t *dirs_d, int32_t *sline_seed_d,
                          int32_t *sline_len_d, float *sline_d) {
    return phase2_entry(t, seeds_d, nseeds, rng_chunk, offsets_d, dirs_d, sline_seed_d, sline_len_d, sline_d, false);
}

int cuslines_phase2_noalloc_async(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk,
                                  const int32_t *offsets_d, const int32_t *dirs_d, int32_t *sline_seed_d,
                                  int32_t *sline_len_d, float *sline_d) {
    return phase2_entry(t, seeds_d, nseeds, rng_chunk, offsets_d, dirs_d, sline_seed_d, sline_len_d, sline_d, true);
}

int cuslines_compact_async(cuslines_tracker *t, const int32_t *sline_len_d, int32_t nslines, const float *sline_d,
                           int64_t *point_off_d, float *points_d) {
    RoctxRange range_("cuslines_compact");
    if (!t || !point_off_d || (nslines > 0 && (!sline_len_d || !sline_d || !points_d)))
        return fail(CUSLINES_EINVAL, "null argument");
    if (nslines < 0) return fail(CUSLINES_EINVAL, "negative size");
    HIPC(hipSetDevice(t->device));
    int rc = run_scan<long long>(t, sline_len_d, nslines, reinterpret_cast<long long *>(point_off_d));
    if (rc) return rc;
    if (nslines == 0) return CUSLINES_OK;
    const size_t slot_f = (size_t)2 * t->dp.max_sline_len * 3;
    const unsigned grid = (unsigned)std::min<long long>(((long long)nslines + 3) / 4, 65536);
    hipLaunchKernelGGL(compact_kernel, dim3(grid), dim3(256), 0, t->stream, sline_len_d, nslines, sline_d, slot_f,
                       reinterpret_cast<const long long *>(point_off_d), points_d);
    HIPC(hipGetLastError());
    return CUSLINES_OK;
}

int cuslines_chunk_workspace_bytes(cuslines_tracker *t, int32_t nseeds, int32_t slot_cap, size_t *bytes) {
    if (!t || !bytes) return fail(CUSLINES_EINVAL, "null argument");
    if (nseeds < 0 || slot_cap < 1) return fail(CUSLINES_EINVAL, "negative size or slot_cap < 1");
    *bytes = chunk_ws_bytes(t->dp, nseeds, slot_cap);
    return CUSLINES_OK;
}

int cuslines_chunk_async(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk,
                         int32_t *counts_d, int32_t *dirs_d, int32_t *offsets_d, int32_t slot_cap,
                         int32_t *sline_seed_d, int32_t *sline_len_d, float *sline_d, int64_t *point_off_d,
                         float *points_d, int64_t *totals_d, void *ws_d, size_t ws_bytes) {
    RoctxRange range_("cuslines_chunk");
    chunk_args a;
    if (int rc = chunk_check(t, seeds_d, nseeds, rng_chunk, counts_d, dirs_d, offsets_d, slot_cap, sline_seed_d,
                             sline_len_d, sline_d, point_off_d, points_d, totals_d, ws_d, ws_bytes, &a))
        return rc;
    HIPC(hipSetDevice(t->device));
    return chunk_run(t, a);
}

int cuslines_chunk_graph_create(cuslines_tracker *t, const float *seeds_d, int32_t nseeds, int32_t rng_chunk,
                                int32_t *counts_d, int32_t *dirs_d, int32_t *offsets_d, int32_t slot_cap,
                                int32_t *sline_seed_d, int32_t *sline_len_d, float *sline_d, int64_t *point_off_d,
                                float *points_d, int64_t *totals_d, void *ws_d, size_t ws_bytes,
                                cuslines_graph **out) {
    RoctxRange range_("cuslines_chunk_graph_create");
    if (!out) return fail(CUSLINES_EINVAL, "null argument");
    *out = nullptr;
    chunk_args a;
    if (int rc = chunk_check(t, seeds_d, nseeds, rng_chunk, counts_d, dirs_d, offsets_d, slot_cap, sline_seed_d,
                             sline_len_d, sline_d, point_off_d, points_d, totals_d, ws_d, ws_bytes, &a))
        return rc;
    HIPC(hipSetDevice(t->device));
    HIPC(hipStreamBeginCapture(t->stream, hipStreamCaptureModeRelaxed));
    const int rc = chunk_run(t, a);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(t->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : fail(CUSLINES_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(CUSLINES_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    *out = new cuslines_graph{t, g, x};
    ++t->ngraphs;
    if (t->parent) ++t->parent->ngraphs;
    return CUSLINES_OK;
}

int cuslines_graph_launch(cuslines_graph *g) {
    if (!g) return fail(CUSLINES_EINVAL, "null graph");
    HIPC(hipSetDevice(g->t->device));
    HIPC(hipGraphLaunch(g->exec, g->t->stream));
    return CUSLINES_OK;
}

int cuslines_graph_destroy(cuslines_graph *g) {
    if (!g) return CUSLINES_OK;
    (void)hipSetDevice(g->t->device);
    (void)hipStreamSynchronize(g->t->stream);  // a launch may still run
    const hipError_t e1 = hipGraphExecDestroy(g->exec);
    const hipError_t e2 = hipGraphDestroy(g->graph);
    --g->t->ngraphs;
    if (g->t->parent) --g->t->parent->ngraphs;
    delete g;
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(CUSLINES_EHIP, "hipGraph destroy failed");
    return CUSLINES_OK;
}

}  // extern "C"
