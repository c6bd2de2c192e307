/*
 * cuslines_oracle.cpp -- CPU restatement of the reference tracking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the gfx950
 * library (gpustreamlines_amd/csrc) and the timed CPU baseline of bench.py.
 * Nothing in the product loads it.  Only the tests, the repository's smoke
 * check and bench.py's CPU-baseline leg use it.
 *
 * It restates the semantics of dipy/GPUStreamlines' CUDA path
 * (the .cu files of cuslines/cuda_c) per seed, with the 32-lane slice of the reference
 * spelled out wherever lane structure decides a result: per-lane random
 * streams, the lane-strided partial sums and their shuffle trees, the
 * blocked Hillis-Steele scan, the ballot/ffs selections and the sort network.
 *
 * Floating point: IEEE float32 with -ffp-contract=off; every fused
 * multiply-add is an explicit fmaf placed where the reference's a*b+c
 * patterns are contracted.  Transcendentals (log, sin, cos) are this
 * project's own float routines, stated below, because the reference's fast
 * intrinsics (__logf/__sinf/__cosf) are hardware approximations that cannot
 * be restated; the GPU library uses the same routines, and tests pin them to
 * float64 libm.  Where the reference compares a float against a macro double
 * literal the comparison is done in double, as C++ promotion does there.
 *
 * Parity unpinned (nothing available here fixes them):
 *  - normals: curand_normal's own algorithm is not published; the normals
 *    here are Box-Muller over this file's log/sin/cos (pinned to their own
 *    float64 formula and to the uniforms, which rocRAND's known-answer
 *    vectors of the same Philox4x32-10 API pin), so they are not cuRAND's
 *    bits.
 *  - texture filtering: the CUDA guide fixes the 9-bit fixed-point weights
 *    (8 fractional bits), the half-texel shift and the border colour; the
 *    rounding of the fraction to 8 bits (nearest, ties to even, here) is
 *    not documented.
 *
 * Deliberate deviations (undefined behaviour in the reference):
 *  - boot tracking: the reference writes the peak list of a non-initial step
 *    into a one-element shared slot (boot.cu:773-789 passes
 *    __sh_new_dir + tidy as the peaks buffer), racing with the other seed of
 *    the block; here every seed has its own peak buffer.
 *  - PTT forward frame: init_frame_ptt_d copies only 6 of 9 frame floats
 *    (ptt.cu:466-471), so the forward pass starts from a stale binormal in
 *    shared memory; here frame[15..17] = frame[6..8] (the unflipped binormal
 *    the comment at ptt.cu:469 describes).
 *  - prob sampling: when no CDF entry exceeds the draw the reference indexes
 *    sphere_vertices[low-1] (generate_streamlines_cuda.cu:199-201), which is
 *    out of bounds when low == 0; here the index is clamped at 0.
 */
#include "../include/cuslines_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int LANES = 32;                 // THR_X_SL, threads per seed
constexpr double PMF_THRESHOLD_P = 0.05;  // cuslines/cuda_python/_globals.py:4
constexpr float REAL_MAX = 3.40282346638528859812e+38f;
constexpr float REAL_MIN = -REAL_MAX;     // globals.h:41-42
constexpr float K_SMALL = 0.0001f;        // ptt.cuh:10
constexpr float PTT_NORM_EPS = 1e-8f;     // ptt.cuh:12
constexpr float PTT_MIN_NORM2 = 1e-16f;   // (1/PTT_INV_NORM_EPS)^2, ptt.cuh:13
constexpr float CSA_C0 = 0.28209479177387814f;  // boot.cu:363

// ---------------------------------------------------------------------------
// Float routines shared (as an algorithm) with the kernels.
// ---------------------------------------------------------------------------
inline float dot3(const float *a, const float *b) {
    return std::fmaf(a[2], b[2], std::fmaf(a[1], b[1], a[0] * b[0]));
}

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// Natural logarithm for positive finite x: x = m * 2^e with m in
// [sqrt(1/2), sqrt(2)), f = m - 1, s = f/(2+f),
// log(1+f) = 2 atanh(s) = 2s + 2s^3/3 + 2s^5/5 + 2s^7/7 + 2s^9/9, and
// log x = e*ln2_hi + (log(1+f) + e*ln2_lo).
float ulog(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (std::isinf(x)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x *= 8388608.0f; e = -23; }
    uint32_t b = f2u(x);
    e += (int)((b >> 23) & 0xff) - 127;
    b = (b & 0x007fffffu) | 0x3f800000u;
    if (b > 0x3fb504f3u) { b -= 0x00800000u; e += 1; }
    const float f = u2f(b) - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float r = std::fmaf(z, 0.22222222f, 0.28571429f);   // 2/9, 2/7
    r = std::fmaf(z, r, 0.4f);                            // 2/5
    r = std::fmaf(z, r, 0.66666669f);                     // 2/3
    const float l1p = std::fmaf(s * z, r, 2.0f * s);
    const float fe = (float)e;
    return std::fmaf(fe, 0.693145751953125f, std::fmaf(fe, 1.4286068203094172e-06f, l1p));
}

inline float bits_float(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}


// sin and cos of |x| < ~1e4: q = rint(x*2/pi), r = x - q*pi/2 with a
// three-part Cody-Waite split of pi/2, Taylor polynomials on |r| <= pi/4
// (sin to r^11, cos to r^12), quadrant rotation.
void usincos(float x, float *sn, float *cs) {
    const float q = std::rint(x * 0.636619772f);
    float r = std::fmaf(-q, 1.5703125f, x);
    r = std::fmaf(-q, 4.83751297e-04f, r);
    r = std::fmaf(-q, 7.54978942e-08f, r);
    const float z = r * r;
    float ps = std::fmaf(z, -2.50521084e-08f, 2.75573192e-06f);
    ps = std::fmaf(z, ps, -1.98412698e-04f);
    ps = std::fmaf(z, ps, 8.33333333e-03f);
    ps = std::fmaf(z, ps, -1.66666667e-01f);
    const float s = std::fmaf(r * z, ps, r);
    float pc = std::fmaf(z, 2.08767570e-09f, -2.75573192e-07f);
    pc = std::fmaf(z, pc, 2.48015873e-05f);
    pc = std::fmaf(z, pc, -1.38888889e-03f);
    pc = std::fmaf(z, pc, 4.16666667e-02f);
    pc = std::fmaf(z, pc, -0.5f);
    const float c = std::fmaf(z, pc, 1.0f);
    const int iq = ((int)q) & 3;
    switch (iq) {
        case 0: *sn = s; *cs = c; break;
        case 1: *sn = c; *cs = -s; break;
        case 2: *sn = -s; *cs = -c; break;
        default: *sn = -c; *cs = s; break;
    }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) with the cuRAND-API
// stream layout as stated by rocRAND's open implementation: key = seed,
// subsequence in counter words z:w, draw index k in x:y (k/4) and word k%4;
// uniform = 2^-32 + u*2^-32; normal = Box-Muller returning the sine term and
// caching the cosine term (cuslines calls: boot.cu:482, ptt.cu:81-83,
// ptt.cu:332-339, generate_streamlines_cuda.cu:184).
// ---------------------------------------------------------------------------
struct Philox {
    uint32_t k0 = 0, k1 = 0;
    uint64_t sub = 0, ctr = 0;
    uint32_t res[4] = {0, 0, 0, 0};
    int idx = 0;
    bool has_normal = false;
    float normal_cache = 0.0f;

    static void block(const uint32_t in[4], uint32_t key0, uint32_t key1, uint32_t out[4]) {
        uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
        for (int r = 0; r < 10; ++r) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
            const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ key0;
            const uint32_t n1 = (uint32_t)p1;
            const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ key1;
            const uint32_t n3 = (uint32_t)p0;
            c0 = n0; c1 = n1; c2 = n2; c3 = n3;
            key0 += 0x9E3779B9u;
            key1 += 0xBB67AE85u;
        }
        out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
    }
    void refill() {
        const uint32_t c[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), (uint32_t)sub,
                               (uint32_t)(sub >> 32)};
        block(c, k0, k1, res);
    }
    void init(uint64_t seed, uint64_t subsequence) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        sub = subsequence;
        ctr = 0;
        idx = 0;
        has_normal = false;
        refill();
    }
    uint32_t next() {
        const uint32_t r = res[idx++];
        if (idx == 4) { idx = 0; ++ctr; refill(); }
        return r;
    }
    float uniform() { return 2.32830644e-10f + (float)next() * 2.32830644e-10f; }
    float normal() {
        if (has_normal) { has_normal = false; return normal_cache; }
        const uint32_t x = next();
        const uint32_t y = next();
        const float u = 2.32830644e-10f + (float)x * 2.32830644e-10f;
        const float v = 1.46291807e-09f + (float)y * 1.46291807e-09f;
        const float s = std::sqrt(-2.0f * ulog(u));
        float sn, cs;
        usincos(v, &sn, &cs);
        normal_cache = cs * s;
        has_normal = true;
        return sn * s;
    }
};

// ---------------------------------------------------------------------------
// Sort network: Knuth's merge exchange (TAOCP 5.2.2 Algorithm M) for 32
// keys, comparators grouped into layers as early as their inputs allow.
// This generates the reference's swap32 table (cuwsort.cuh:37-52).
// ---------------------------------------------------------------------------
struct SortNet {
    int nlayers = 0;
    int partner[32][32];
    SortNet() {
        int ready[32] = {0};
        for (auto &row : partner) for (int l = 0; l < 32; ++l) row[l] = l;
        const int n = 32, t = 5;
        for (int p = 1 << (t - 1); p > 0; p >>= 1) {
            int q = 1 << (t - 1), r = 0, d = p;
            for (;;) {
                for (int i = 0; i < n - d; ++i) {
                    if ((i & p) != r) continue;
                    const int a = i, b = i + d;
                    const int L = std::max(ready[a], ready[b]);
                    partner[L][a] = b;
                    partner[L][b] = a;
                    ready[a] = ready[b] = L + 1;
                    nlayers = std::max(nlayers, L + 1);
                }
                if (q != p) { d = q - p; q >>= 1; r = p; } else break;
            }
        }
    }
};
const SortNet &sortnet() { static const SortNet s; return s; }

// warp_sort<32,32,WSORT_DIR_DEC> on (key, value) pairs (cuwsort.cuh:104-168).
void warp_sort_dec(float *k, int *v) {
    const SortNet &sn = sortnet();
    float nk[32];
    int nv[32];
    for (int L = 0; L < sn.nlayers; ++L) {
        for (int g = 0; g < 32; ++g) {
            const int src = sn.partner[L][g];
            const float a = k[src];
            const int b = v[src];
            nk[g] = k[g];
            nv[g] = v[g];
            if ((g < src) == false) {       // (gid < srclane) == WSORT_DIR_DEC
                if (a < k[g]) { nk[g] = a; nv[g] = b; }
            } else {
                if (a > k[g]) { nk[g] = a; nv[g] = b; }
            }
        }
        std::memcpy(k, nk, sizeof(nk));
        std::memcpy(v, nv, sizeof(nv));
    }
}

// ---------------------------------------------------------------------------
// Problem state
// ---------------------------------------------------------------------------
struct Problem {
    cuslines_params p{};
    std::vector<float> dataf, stop, verts, H, R, delta_b, delta_q, smat;
    std::vector<int32_t> edges, b0s, lut;
    int hr_side = 0;
    float cos_max_angle = 0, cos_min_sep = 0;
    double step_d = 0;  // STEP_SIZE as the double literal the macro holds
    int nx = 1, ny = 1, nz = 1;
    // learned PTT proposal (no counterpart in the reference)
    std::vector<uint16_t> mlp;     // bf16 parameters; empty: plain getter
    mutable long long stats[4] = {0, 0, 0, 0};
};

// Lane-strided FMA partial sums reduced by the shuffle tree that lane 0 sees
// (ndotp_d, boot.cu:195-235): partial[l] accumulates j = l, l+32, ...; then
// pairs at distance 16, 8, 4, 2, 1.
inline float tree_sum(float *part) {
    for (int off = 16; off >= 1; off >>= 1)
        for (int l = 0; l < off; ++l) part[l] = part[l] + part[l + off];
    return part[0];
}

void ndotp(int N, int M, const float *v, const float *mat, float *out) {
    for (int i = 0; i < N; ++i) {
        float part[LANES] = {0};
        for (int j = 0; j < M; ++j) part[j % LANES] = std::fmaf(v[j], mat[(size_t)i * M + j], part[j % LANES]);
        out[i] = tree_sum(part);
    }
}

void ndotp_log_opdt(int N, int M, const float *v, const float *mat, float *out) {
    for (int i = 0; i < N; ++i) {
        float part[LANES] = {0};
        for (int j = 0; j < M; ++j) {
            const float x = v[j];
            const float lx = ulog(x);
            const float t = ((-lx) * (1.5f + lx)) * x;
            part[j % LANES] = std::fmaf(t, mat[(size_t)i * M + j], part[j % LANES]);
        }
        out[i] = tree_sum(part);
    }
}

void ndotp_log_csa(int N, int M, const float *v, const float *mat, float *out) {
    for (int i = 0; i < N; ++i) {
        float part[LANES] = {0};
        for (int j = 0; j < M; ++j) {
            const float x = std::min(std::max(v[j], 0.001f), 0.999f);
            const float t = ulog(-ulog(x));
            part[j % LANES] = std::fmaf(t, mat[(size_t)i * M + j], part[j % LANES]);
        }
        out[i] = tree_sum(part);
    }
}

// prefix_sum_sh_d (utils.cu:81-120): 32-element blocks, lane 0 carries the
// previous block's last value, Hillis-Steele inside the block.
void prefix_sum(float *a, int len) {
    for (int j = 0; j < len; j += LANES) {
        if (j != 0) a[j] = a[j] + a[j - 1];
        const int cnt = std::min(LANES, len - j);
        float v[LANES];
        for (int t = 0; t < cnt; ++t) v[t] = a[j + t];
        for (int i = 1; i < LANES; i *= 2) {
            float old[LANES];
            std::memcpy(old, v, sizeof(float) * cnt);
            for (int t = i; t < cnt; ++t) v[t] = v[t] + old[t - i];
        }
        for (int t = 0; t < cnt; ++t) a[j + t] = v[t];
    }
}

// trilinear_interp_d (tracking_helpers.cu:3-61).
int trilinear(const Problem &P, const float *data, int dimt, const float *pt, float *out) {
    const int DX = P.p.dimx, DY = P.p.dimy, DZ = P.p.dimz;
    if (pt[0] < -0.5f || pt[0] + 0.5f >= (float)DX || pt[1] < -0.5f || pt[1] + 0.5f >= (float)DY ||
        pt[2] < -0.5f || pt[2] + 0.5f >= (float)DZ)
        return -1;
    long long c[3][2];
    float w[3][2];
    const int D[3] = {DX, DY, DZ};
    for (int a = 0; a < 3; ++a) {
        const float fl = std::floor(pt[a]);
        w[a][1] = pt[a] - fl;
        w[a][0] = 1.0f - w[a][1];
        c[a][0] = (long long)std::max(0.0f, fl);
        c[a][1] = std::min((long long)D[a] - 1, c[a][0] + 1);
    }
    for (int t = 0; t < dimt; ++t) {
        float acc = 0.0f;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) {
                    const float wt = (w[0][i] * w[1][j]) * w[2][k];
                    const size_t off = (((size_t)c[0][i] * DY + c[1][j]) * DZ + c[2][k]) * dimt + t;
                    acc = std::fmaf(wt, data[off], acc);
                }
        out[t] = acc;
    }
    return 0;
}

// Linear texture filtering on unnormalised coordinates, emulated in
// software after the CUDA programming guide's definition: xB = x - 0.5,
// i = floor(xB), alpha = frac(xB) held with 8 fractional bits; the eight
// texel weights multiply out exactly and are accumulated fastest axis first.
inline float tex_weight(float c, int *i0) {
    const float b = c - 0.5f;
    const float fl = std::floor(b);
    *i0 = (int)fl;
    const float a = b - fl;
    return std::rint(a * 256.0f) * 0.00390625f;
}

template <class Fetch>
float tex3d_linear(const Fetch &fetch, float x, float y, float z) {
    int i, j, k;
    const float a = tex_weight(x, &i), b = tex_weight(y, &j), g = tex_weight(z, &k);
    const float a0 = 1.0f - a, b0 = 1.0f - b, g0 = 1.0f - g;
    float acc = ((a0 * b0) * g0) * fetch(i, j, k);
    acc = std::fmaf((a * b0) * g0, fetch(i + 1, j, k), acc);
    acc = std::fmaf((a0 * b) * g0, fetch(i, j + 1, k), acc);
    acc = std::fmaf((a * b) * g0, fetch(i + 1, j + 1, k), acc);
    acc = std::fmaf((a0 * b0) * g, fetch(i, j, k + 1), acc);
    acc = std::fmaf((a * b0) * g, fetch(i + 1, j, k + 1), acc);
    acc = std::fmaf((a0 * b) * g, fetch(i, j + 1, k + 1), acc);
    acc = std::fmaf((a * b) * g, fetch(i + 1, j + 1, k + 1), acc);
    return acc;
}

// check_point_d (tracking_helpers.cu:63-73): tex3D(stop, z, y, x) with
// border colour -1 (cutils.py:111-125); texture x is the array's z axis.
int check_point(const Problem &P, const float *pt) {
    const int DX = P.p.dimx, DY = P.p.dimy, DZ = P.p.dimz;
    const float *s = P.stop.data();
    auto fetch = [&](int iz, int iy, int ix) -> float {
        if (iz < 0 || iz >= DZ || iy < 0 || iy >= DY || ix < 0 || ix >= DX) return -1.0f;
        return s[((size_t)ix * DY + iy) * DZ + iz];
    };
    const float val = tex3d_linear(fetch, pt[2], pt[1], pt[0]);
    if (val == -1.0f) return CUSLINES_OUTSIDEIMAGE;
    return (val > P.p.tc_threshold) ? CUSLINES_TRACKPOINT : CUSLINES_ENDPOINT;
}

// peak_directions_d (tracking_helpers.cu:75-245).  Writes kept vertex
// indices to out_idx and returns their count.
int peak_directions(const Problem &P, const float *odf, int *out_idx, int *scratch) {
    const int n_v = P.p.nverts;
    float odf_min = REAL_MAX;
    for (int j = 0; j < n_v; ++j) odf_min = (odf_min < odf[j]) ? odf_min : odf[j];
    odf_min = (0.0f > odf_min) ? 0.0f : odf_min;  // MAX(0, odf_min)
    int *mark = scratch;
    std::fill(mark, mark + n_v, 0);
    for (int e = 0; e < P.p.nedges; ++e) {
        const int u = P.edges[2 * e], v = P.edges[2 * e + 1];
        const float uv = odf[u], vv = odf[v];
        if (uv < vv) { mark[u] = -1; mark[v] |= 1; }
        else if (vv < uv) { mark[v] = -1; mark[u] |= 1; }
    }
    float m = REAL_MIN;
    for (int j = 0; j < n_v; ++j)
        if (mark[j] > 0) { const float x = odf[j] + (-odf_min); m = (m > x) ? m : x; }
    const float comp = (float)((double)P.p.relative_peak_thresh * (double)m);
    int n = 0;
    int *ind = scratch + n_v;
    for (int j = 0; j < n_v; ++j)
        if (mark[j] > 0 && (odf[j] - odf_min) >= comp) ind[n++] = j;
    if (n < LANES) {
        float k[LANES];
        int v[LANES];
        for (int l = 0; l < LANES; ++l) {
            k[l] = (l < n) ? odf[ind[l]] : REAL_MIN;
            v[l] = (l < n) ? ind[l] : 0;
        }
        warp_sort_dec(k, v);
        for (int l = 0; l < n; ++l) ind[l] = v[l];
    }
    if (n == 0) return 0;
    const float *V = P.verts.data();
    out_idx[0] = ind[0];
    int kept = 1;
    for (int i = 1; i < n; ++i) {
        const float *abc = V + 3 * ind[i];
        int j = 0;
        for (; j < kept; ++j) {
            float c = dot3(abc, V + 3 * out_idx[j]);
            if (P.p.sphere_symm) c = std::fabs(c);
            if (c > P.cos_min_sep) break;
        }
        if (j == kept) out_idx[kept++] = ind[i];
    }
    return kept;
}

struct SeedScratch {
    std::vector<float> a, b, c, d;
    std::vector<int> iscr, peaks;
    Philox rng[LANES];
    float frame[18];
    float mix[8 * 12];
    long long stats[4] = {0, 0, 0, 0};
    void size(const Problem &P) {
        const size_t n = (size_t)std::max(P.p.dimt, P.p.nverts) + 64;
        a.assign(n, 0.f); b.assign(n, 0.f); c.assign(n, 0.f); d.assign(n, 0.f);
        iscr.assign(2 * n, 0);
        peaks.assign(n, 0);
    }
};

// ---------------------------------------------------------------------------
// Probabilistic getter (generate_streamlines_cuda.cu:49-275)
// ---------------------------------------------------------------------------
// Inverse-CDF selection (generate_streamlines_cuda.cu:184-201): bisection of
// the inclusive prefix sum down to a window of at most 32 entries, then the
// first entry of the window whose CDF exceeds sel (lane ballot + ffs).  An
// empty ballot gives low - 1 (index -1 clamped to 0, the reference reads
// pmf[-1] there).
int prob_select(const float *cdf, int n, float sel) {
    int low = 0, high = n - 1;
    while ((high - low) >= LANES) {
        const int mid = (low + high) / 2;
        if (cdf[mid] < sel) low = mid; else high = mid;
    }
    int ind = low - 1;
    for (int t = 0; t < LANES; ++t)
        if (low + t <= high && sel < cdf[low + t]) { ind = low + t; break; }
    if (ind < 0) ind = 0;
    return ind;
}

int get_direction_prob(const Problem &P, SeedScratch &S, bool is_start, const float *dir,
                       const float *pt, float *new_dir, int *peaks_out, int *ind_out = nullptr) {
    const int n = P.p.nverts;
    float *pmf = S.a.data();
    if (trilinear(P, P.dataf.data(), P.p.dimt, pt, pmf) != 0) return 0;
    float mx = REAL_MIN;
    for (int i = 0; i < n; ++i) mx = (mx > pmf[i]) ? mx : pmf[i];
    const float thr = (float)(PMF_THRESHOLD_P * (double)mx);
    for (int i = 0; i < n; ++i) if (pmf[i] < thr) pmf[i] = 0.0f;
    if (is_start) return peak_directions(P, pmf, peaks_out, S.iscr.data());
    const float *V = P.verts.data();
    for (int i = 0; i < n; ++i) {
        float d = dot3(dir, V + 3 * i);
        if (P.p.sphere_symm) d = std::fabs(d);
        if (d < P.cos_max_angle) pmf[i] = 0.0f;
    }
    prefix_sum(pmf, n);
    const float last = pmf[n - 1];
    if (last == 0.0f) return 0;
    const float sel = S.rng[0].uniform() * last;
    const int ind = prob_select(pmf, n, sel);
    if (ind_out) *ind_out = ind;
    const float *v = V + 3 * ind;
    if (dot3(dir, v) > 0.0f) { new_dir[0] = v[0]; new_dir[1] = v[1]; new_dir[2] = v[2]; }
    else { new_dir[0] = -v[0]; new_dir[1] = -v[1]; new_dir[2] = -v[2]; }
    return 1;
}

// ---------------------------------------------------------------------------
// Bootstrap getter (boot.cu:19-113, 115-193, 393-608)
// ---------------------------------------------------------------------------
int mask_get(const Problem &P, const float *plain, float *masked) {
    int w = 0;
    for (int j = 0; j < P.p.dimt; ++j) if (!P.b0s[j]) masked[w++] = plain[j];
    return w;
}
void mask_put(const Problem &P, const float *masked, float *plain) {
    int w = 0;
    for (int j = 0; j < P.p.dimt; ++j) if (!P.b0s[j]) plain[j] = masked[w++];
}
float avg_mask(const Problem &P, const float *data) {
    float part[LANES] = {0};
    int cnt[LANES] = {0};
    for (int j = 0; j < P.p.dimt; ++j)
        if (P.b0s[j]) { cnt[j % LANES]++; part[j % LANES] = part[j % LANES] + data[j]; }
    int c = 0;
    for (int l = 0; l < LANES; ++l) c += cnt[l];
    return tree_sum(part) / (float)c;
}

int closest_peak(const Problem &P, const float *dir, int npeaks, const int *peaks, float *out) {
    float cd[LANES];
    int ci[LANES];
    const float *V = P.verts.data();
    for (int l = 0; l < LANES; ++l) {
        cd[l] = 0.0f; ci[l] = -1;
        for (int j = l; j < npeaks; j += LANES) {
            const float d = dot3(dir, V + 3 * peaks[j]);
            if (std::fabs(d) > std::fabs(cd[l])) { cd[l] = d; ci[l] = j; }
        }
    }
    for (int off = 16; off >= 1; off >>= 1) {
        float nd[LANES];
        int ni[LANES];
        for (int l = 0; l < LANES; ++l) {
            nd[l] = cd[l]; ni[l] = ci[l];
            const int o = l ^ off;
            if (std::fabs(cd[o]) > std::fabs(cd[l])) { nd[l] = cd[o]; ni[l] = ci[o]; }
        }
        std::memcpy(cd, nd, sizeof(cd));
        std::memcpy(ci, ni, sizeof(ci));
    }
    if (ci[0] >= 0) {
        const float *v = V + 3 * peaks[ci[0]];
        if (cd[0] >= P.cos_max_angle) { out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; return 1; }
        if (cd[0] <= -P.cos_max_angle) { out[0] = -v[0]; out[1] = -v[1]; out[2] = -v[2]; return 1; }
    }
    return 0;
}

// One bootstrap refit (boot.cu:447-545): trilinear DWI gather, residual
// bootstrap h = H d + (R d)[perm], min_signal clip, b0 normalisation, model
// fit (OPDT or CSA) and ODF on the sphere into odf[nverts] (all zero outside
// the volume).  The normalised bootstrap signal is left in S.a[0..dimt).
int boot_odf(const Problem &P, SeedScratch &S, const float *pt, float *odf) {
    const int hr = P.hr_side;
    const int dn = P.p.delta_nr;
    const int nv = P.p.nverts;
    float *vox = S.a.data(), *msk = S.b.data(), *r = S.c.data();
    const int rv = trilinear(P, P.dataf.data(), P.p.dimt, pt, vox);
    mask_get(P, vox, msk);
    if (rv != 0) {
        for (int j = 0; j < nv; ++j) odf[j] = 0.0f;
        return rv;
    }
    ndotp(hr, hr, msk, P.R.data(), r);
    ndotp(hr, hr, msk, P.H.data(), odf);
    for (int j = 0; j < hr; ++j) {
        const int src = (int)(S.rng[j % LANES].next() % (uint32_t)hr);
        odf[j] = odf[j] + r[src];
    }
    mask_put(P, odf, vox);
    for (int j = 0; j < P.p.dimt; ++j) vox[j] = (P.p.min_signal > vox[j]) ? P.p.min_signal : vox[j];
    const float denom = avg_mask(P, vox);
    for (int j = 0; j < P.p.dimt; ++j) vox[j] = vox[j] / denom;
    mask_get(P, vox, msk);
    if (P.p.model_type == CUSLINES_OPDT) {
        ndotp_log_opdt(dn, hr, msk, P.delta_q.data(), r);
        ndotp(dn, hr, msk, P.delta_b.data(), odf);
        for (int j = 0; j < dn; ++j) r[j] = r[j] - odf[j];
    } else {
        ndotp_log_csa(dn, hr, msk, P.delta_q.data(), r);
        r[0] = CSA_C0;
    }
    ndotp(nv, dn, r, P.smat.data(), odf);
    return 0;
}

int get_direction_boot(const Problem &P, SeedScratch &S, int nattempts, const float *dir,
                       const float *pt, float *new_dir, int *peaks_out) {
    const int nv = P.p.nverts;
    float *h = S.d.data();
    int *peaks = peaks_out ? peaks_out : S.peaks.data();
    for (int att = 0; att < nattempts; ++att) {
        boot_odf(P, S, pt, h);
        float mx = REAL_MIN;
        for (int j = 0; j < nv; ++j) mx = (mx > h[j]) ? mx : h[j];
        const float thr = (float)(PMF_THRESHOLD_P * (double)mx);
        for (int j = 0; j < nv; ++j) if (h[j] < thr) h[j] = 0.0f;
        const int ndir = peak_directions(P, h, peaks, S.iscr.data());
        if (nattempts == 1) return ndir;
        if (ndir > 0 && closest_peak(P, dir, ndir, peaks, new_dir)) return 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// PTT getter (ptt.cu)
// ---------------------------------------------------------------------------
// 1/sqrt(n2) for finite n2 > 1e-16: Lomont's initial approximation
// (0x5f375a86 - bits/2, "Fast Inverse Square Root", 2003) refined by three
// Newton-Raphson steps y <- y (1.5 - (n2/2) y^2).  It stands in for
// rnorm3df, whose implementation is not restated.
float rsqrt_nr(float n2) {
    uint32_t b;
    std::memcpy(&b, &n2, 4);
    b = 0x5f375a86u - (b >> 1);
    float y;
    std::memcpy(&y, &b, 4);
    const float h = 0.5f * n2;
    y = y * std::fmaf(-(h * y), y, 1.5f);
    y = y * std::fmaf(-(h * y), y, 1.5f);
    y = y * std::fmaf(-(h * y), y, 1.5f);
    return y;
}

// norm3_d (ptt.cu:1-13): unit length when 1/|v| is finite, positive and
// below PTT_INV_NORM_EPS = 1e8, i.e. |v|^2 > 1e-16; else the FAIL axis.
template <int FAIL>
void norm3(float *v) {
    const float n2 = std::fmaf(v[2], v[2], std::fmaf(v[1], v[1], v[0] * v[0]));
    if (n2 > PTT_MIN_NORM2 && n2 <= REAL_MAX) {
        const float inv = rsqrt_nr(n2);
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    } else {
        v[0] = v[1] = v[2] = 0.0f;
        v[FAIL] = 1.0f;
    }
}
template <int FAIL>
void crossnorm3(float *d, const float *a, const float *b) {
    d[0] = std::fmaf(a[1], b[2], -(a[2] * b[1]));
    d[1] = std::fmaf(a[2], b[0], -(a[0] * b[2]));
    d[2] = std::fmaf(a[0], b[1], -(a[1] * b[0]));
    norm3<FAIL>(d);
}

int lut_lookup(const Problem &P, const float *frame) {
    const int res = P.p.lut_res;
    int ix[3];
    for (int a = 0; a < 3; ++a) {
        const float u = std::fmaf(frame[a], 0.5f, 0.5f);
        const float x = std::floor(u * (float)res);
        const int i = (int)std::fmin(std::fmax(x, 0.0f), (float)(res - 1));
        ix[a] = i;
    }
    return P.lut[((size_t)ix[0] * res + ix[1]) * res + ix[2]];
}

// interp4_d (ptt.cu:24-42): vertex from the nearest-vertex LUT, then the
// linear-filtered tiled ODF texture (clamp address mode; tiles of
// nx x ny x nz vertex slots; slots >= nverts are zero padding).
float interp4(const Problem &P, const float *pos, const float *frame) {
    const int idx = lut_lookup(P, frame);
    const int lx = P.p.tile_log2_x, ly = P.p.tile_log2_y;
    const int tx = idx & (P.nx - 1);
    const int ty = (idx >> lx) & (P.ny - 1);
    const int tz = idx >> (lx + ly);
    const int DX = P.p.dimx, DY = P.p.dimy, DZ = P.p.dimz, T = P.p.dimt;
    const float xq = (float)(tx * DX) + pos[0];
    const float yq = (float)(ty * DY) + pos[1];
    const float zq = (float)(tz * DZ) + pos[2];
    const int W = P.nx * DX, Hh = P.ny * DY, Dd = P.nz * DZ;
    const float *d = P.dataf.data();
    auto fetch = [&](int I, int J, int K) -> float {
        I = std::min(std::max(I, 0), W - 1);
        J = std::min(std::max(J, 0), Hh - 1);
        K = std::min(std::max(K, 0), Dd - 1);
        const int v = (I / DX) + P.nx * ((J / DY) + P.ny * (K / DZ));
        if (v >= T) return 0.0f;
        return d[(((size_t)(I % DX) * DY + (J % DY)) * DZ + (K % DZ)) * T + v];
    };
    return tex3d_linear(fetch, xq, yq, zq);
}

void prepare_propagator(float k1, float k2, float arclen, float *p) {
    if (std::fabs(k1) < K_SMALL && std::fabs(k2) < K_SMALL) {
        p[0] = arclen; p[1] = 0; p[2] = 0; p[3] = 1; p[4] = 0; p[5] = 0; p[6] = 0; p[7] = 0; p[8] = 1;
        return;
    }
    if (std::fabs(k1) < K_SMALL) k1 = K_SMALL;
    if (std::fabs(k2) < K_SMALL) k2 = K_SMALL;
    const float k = std::sqrt(std::fmaf(k2, k2, k1 * k1));
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
    p[8] = std::fmaf(k2 * k2, coskt, k1 * k1) * kk;
}

void random_normal(Philox &rng, float *pf) {
    pf[3] = rng.normal();
    pf[4] = rng.normal();
    pf[5] = rng.normal();
    const float dt = std::fmaf(pf[5], pf[2], std::fmaf(pf[4], pf[1], pf[3] * pf[0]));
    pf[3] = std::fmaf(-dt, pf[0], pf[3]);
    pf[4] = std::fmaf(-dt, pf[1], pf[4]);
    pf[5] = std::fmaf(-dt, pf[2], pf[5]);
    const float n2 = std::fmaf(pf[5], pf[5], std::fmaf(pf[4], pf[4], pf[3] * pf[3]));
    if (n2 < PTT_NORM_EPS) {
        const float ax = std::fabs(pf[0]), ay = std::fabs(pf[1]), az = std::fabs(pf[2]);
        if (ax <= ay && ax <= az) { pf[3] = 0.0f; pf[4] = pf[2]; pf[5] = -pf[1]; }
        else if (ay <= az) { pf[3] = -pf[2]; pf[4] = 0.0f; pf[5] = pf[0]; }
        else { pf[3] = pf[1]; pf[4] = -pf[0]; pf[5] = 0.0f; }
    }
}

void get_probing_frame(bool init, const float *frame, Philox &rng, float *pf) {
    if (init) {
        pf[0] = frame[0]; pf[1] = frame[1]; pf[2] = frame[2];
        norm3<0>(pf);
        random_normal(rng, pf);
        norm3<1>(pf + 3);
        crossnorm3<2>(pf + 6, pf, pf + 3);
    } else {
        for (int i = 0; i < 9; ++i) pf[i] = frame[i];
    }
}

void propagate_frame(const float *p, float *f, float *direc) {
    float tmp[3];
    for (int i = 0; i < 3; ++i) {
        direc[i] = std::fmaf(p[2], f[6 + i], std::fmaf(p[1], f[3 + i], p[0] * f[i]));
        tmp[i] = std::fmaf(p[5], f[6 + i], std::fmaf(p[4], f[3 + i], p[3] * f[i]));
        f[6 + i] = std::fmaf(p[8], f[6 + i], std::fmaf(p[7], f[3 + i], p[6] * f[i]));
    }
    norm3<0>(tmp);
    crossnorm3<1>(f + 3, f + 6, tmp);
    crossnorm3<2>(f + 6, tmp, f + 3);
    f[0] = tmp[0]; f[1] = tmp[1]; f[2] = tmp[2];
}

float data_support(const Problem &P, float support, const float *pos, float k1, float k2, float *pf) {
    float prop[9], direc[3];
    prepare_propagator(k1, k2, P.p.probe_step_size, prop);
    float pp[3] = {pos[0], pos[1], pos[2]};
    for (int i = 0; i < P.p.probe_quality; ++i) {
        propagate_frame(prop, pf, direc);
        pp[0] += direc[0]; pp[1] += direc[1]; pp[2] += direc[2];
        support += interp4(P, pp, pf);
    }
    return support;
}

// ---------------------------------------------------------------------------
// PTT disc tables (quality 2) -- numeric data of the reference's
// cuslines/cuda_c/disc.h:20-84 (from nibrary, BSD-3-Clause,
// Copyright (c) 2024 Dogu Baran Aydogan): vertex (k1,k2) coordinates on the
// unit disc and the triangle list of their triangulation.
// ---------------------------------------------------------------------------
const float DISC_VERT[24 * 2] = {
    -0.99680788f, -0.07983759f, -0.94276539f, 0.33345677f, -0.87928469f, -0.47629658f,
    -0.72856617f, 0.68497542f,  -0.60006556f, -0.79995082f, -0.54129995f, -0.02761342f,
    -0.39271207f, 0.37117272f,  -0.39217391f, 0.91989110f,  -0.36362884f, -0.40757367f,
    -0.22391316f, -0.97460910f, -0.00130022f, 0.53966106f,  0.00000000f,  0.00000000f,
    0.00973999f,  0.99995257f,  0.01606516f,  -0.54289908f, 0.21342395f,  -0.97695968f,
    0.38192071f,  -0.38666136f, 0.38897094f,  0.37442837f,  0.40696681f,  0.91344295f,
    0.54387161f,  -0.01477123f, 0.59119367f,  -0.80652963f, 0.73955688f,  0.67309406f,
    0.87601150f,  -0.48229022f, 0.94617928f,  0.32364298f,  0.99585368f,  -0.09096944f};
const int DISC_FACE[31 * 3] = {
    9, 8, 4,    11, 16, 10, 5, 8, 11,   5, 1, 0,    18, 16, 11, 11, 15, 18, 13, 8, 9,  11, 8, 13,
    13, 15, 11, 22, 18, 23, 22, 20, 16, 16, 18, 22, 16, 20, 17, 12, 10, 17, 17, 10, 16, 15, 19, 21,
    23, 18, 21, 21, 18, 15, 2, 4, 8,    2, 5, 0,    8, 5, 2,    7, 10, 12,  6, 7, 3,    10, 7, 6,
    3, 1, 6,    1, 5, 6,    11, 10, 6,  6, 5, 11,   14, 19, 15, 15, 13, 14, 14, 13, 9};
constexpr int DISC_VERT_CNT = 24, DISC_FACE_CNT = 31;

// Face CDF over the 31 disc faces from vert_pdf and up to `rounds` rounds
// of 32 candidates (ptt.cu:270-410).  1 with new_dir (and the frame
// advanced), 0 when the CDF is empty or no round has a winner.
// What one sub-step drew and measured (for the oracle's own distribution
// tests): the last round's 32 candidate curvatures and supports, the winning
// lane (-1: none) and the round it won in.
struct PttTrace {
    float k[2 * LANES];
    float sup[LANES];
    int win = -1, round = -1;
};

int ptt_sample(const Problem &P, SeedScratch &S, bool init, const float *dir, float *frame, const float *pos,
               float first_val, const float *vert_pdf, int rounds, float *new_dir, PttTrace *tr = nullptr) {
    const double min_support = (double)P.p.probe_quality * PMF_THRESHOLD_P;
    float cdf[DISC_FACE_CNT];
    for (int ii = 0; ii < DISC_FACE_CNT; ++ii) {
        cdf[ii] = 0.0f;
        for (int jj = 0; jj < 3; ++jj) {
            const float vv = vert_pdf[DISC_FACE[ii * 3 + jj]];
            if (vv == 0.0f) { cdf[ii] = 0.0f; break; }
            cdf[ii] = cdf[ii] + vv;
        }
    }
    prefix_sum(cdf, DISC_FACE_CNT);
    const float last = cdf[DISC_FACE_CNT - 1];
    if (last == 0.0f) return 0;
    const double mc = (double)P.p.max_curvature;
    for (int round = 0; round < rounds; ++round) {
        float sup[LANES], k1s[LANES], k2s[LANES], pfs[LANES][9];
        for (int t = 0; t < LANES; ++t) {
            Philox &rng = S.rng[t];
            float r1 = rng.uniform(), r2 = rng.uniform();
            if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
            const float tmp = rng.uniform() * last;
            int jj = 0;
            for (; jj < DISC_FACE_CNT - 1; ++jj) if (cdf[jj] >= tmp) break;
            const int *fc = DISC_FACE + jj * 3;
            const float vx0 = (float)(mc * (double)DISC_VERT[fc[0] * 2]);
            const float vx1 = (float)(mc * (double)DISC_VERT[fc[1] * 2]);
            const float vx2 = (float)(mc * (double)DISC_VERT[fc[2] * 2]);
            const float vy0 = (float)(mc * (double)DISC_VERT[fc[0] * 2 + 1]);
            const float vy1 = (float)(mc * (double)DISC_VERT[fc[1] * 2 + 1]);
            const float vy2 = (float)(mc * (double)DISC_VERT[fc[2] * 2 + 1]);
            k1s[t] = std::fmaf(r2, vx2 - vx0, std::fmaf(r1, vx1 - vx0, vx0));
            k2s[t] = std::fmaf(r2, vy2 - vy0, std::fmaf(r1, vy1 - vy0, vy0));
            get_probing_frame(init, frame, rng, pfs[t]);
            sup[t] = data_support(P, first_val, pos, k1s[t], k2s[t], pfs[t]);
        }
        int win = -1;
        if (tr) {
            for (int t = 0; t < LANES; ++t) { tr->k[2 * t] = k1s[t]; tr->k[2 * t + 1] = k2s[t]; tr->sup[t] = sup[t]; }
            tr->round = round;
        }
        if (init) {
            float mx = sup[0];
            for (int t = 1; t < LANES; ++t) mx = std::fmax(mx, sup[t]);
            for (int t = 0; t < LANES; ++t)
                if (sup[t] == mx && (double)mx >= min_support) { win = t; break; }
        } else {
            for (int t = 0; t < LANES; ++t) if ((double)sup[t] >= min_support) { win = t; break; }
        }
        if (win != -1) {
            if (tr) tr->win = win;
            if (init) {
                new_dir[0] = dir[0]; new_dir[1] = dir[1]; new_dir[2] = dir[2];
                for (int i = 0; i < 9; ++i) frame[i] = pfs[win][i];
            } else {
                float prop[9], d[3], pf[9];
                prepare_propagator(k1s[win], k2s[win], (float)(P.step_d / P.p.step_frac), prop);
                for (int i = 0; i < 9; ++i) pf[i] = frame[i];
                propagate_frame(prop, pf, d);
                norm3<0>(d);
                new_dir[0] = d[0]; new_dir[1] = d[1]; new_dir[2] = d[2];
                for (int i = 0; i < 9; ++i) frame[i] = pf[i];
            }
            return 1;
        }
    }
    return 0;
}

// The 24 vertex supports of a sub-step (ptt.cu:232-260) into vert_pdf (0
// below 0.4; raw values into `raw` when given).  Returns whether any passed.
bool ptt_vertex_supports(const Problem &P, SeedScratch &S, bool init, const float *frame, const float *pos,
                         float first_val, float *vert_pdf, float *raw) {
    const double min_support = (double)P.p.probe_quality * PMF_THRESHOLD_P;
    bool found = false;
    for (int ii = 0; ii < DISC_VERT_CNT; ++ii) {  // lane ii
        const float k1 = (float)((double)DISC_VERT[ii * 2] * (double)P.p.max_curvature);
        const float k2 = (float)((double)DISC_VERT[ii * 2 + 1] * (double)P.p.max_curvature);
        float pf[9];
        get_probing_frame(init, frame, S.rng[ii], pf);
        const float sup = data_support(P, first_val, pos, k1, k2, pf);
        if (raw) raw[ii] = sup;
        if ((double)sup < min_support) vert_pdf[ii] = 0.0f;
        else { vert_pdf[ii] = sup; found = true; }
    }
    return found;
}

int get_direction_ptt(const Problem &P, SeedScratch &S, bool init, const float *dir, float *frame,
                      const float *pos, float *new_dir, float *raw = nullptr, PttTrace *tr = nullptr) {
    if (init) { frame[0] = dir[0]; frame[1] = dir[1]; frame[2] = dir[2]; }
    const float first_val = interp4(P, pos, frame);
    float vert_pdf[DISC_VERT_CNT];
    if (!ptt_vertex_supports(P, S, init, frame, pos, first_val, vert_pdf, raw)) return 0;
    return ptt_sample(P, S, init, dir, frame, pos, first_val, vert_pdf, 1024 / LANES, new_dir, tr);
}

bool init_frame_ptt(const Problem &P, SeedScratch &S, const float *first_step, const float *seed) {
    float tmp[3];
    float *frame = S.frame;
    bool ok = get_direction_ptt(P, S, true, first_step, frame, seed, tmp);
    if (!ok) {
        const float neg[3] = {-first_step[0], -first_step[1], -first_step[2]};
        ok = get_direction_ptt(P, S, true, neg, frame, seed, tmp);
        if (!ok) return false;
    } else {
        for (int i = 0; i < 6; ++i) frame[i] = -frame[i];
    }
    for (int i = 0; i < 6; ++i) frame[9 + i] = -frame[i];
    for (int i = 0; i < 3; ++i) frame[15 + i] = frame[6 + i];
    return true;
}

// ---------------------------------------------------------------------------
// Learned PTT proposal (no counterpart in the reference; the layout and
// semantics are those of include/cuslines_hip.h).  Same operation order as
// the kernels: neurons are fma chains over their inputs in order starting
// from the bias.
// ---------------------------------------------------------------------------
constexpr int MLP_IN = CUSLINES_MLP_IN, MLP_H = CUSLINES_MLP_HIDDEN, MLP_OUT = CUSLINES_MLP_OUT;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_IN, MLP_W2 = MLP_B1 + MLP_H,
              MLP_B2 = MLP_W2 + MLP_H * MLP_H, MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_OUT * MLP_H;
constexpr int FEAT_STEPS = 4;

inline float bf16f(uint16_t h) { return bits_float((uint32_t)h << 16); }

// Inputs: per disc vertex the sum of the ODF taps after 4 and 8 probe steps
// along its curve (one propagator of 4 probe steps, applied twice), the ODF
// at the point, three zeros, the incoming direction and the frame.
void mlp_features(const Problem &P, const float *pos, const float *dir, const float *frame, float first_val,
                  float *x) {
    for (int v = 0; v < DISC_VERT_CNT; ++v) {
        const float k1 = (float)((double)DISC_VERT[v * 2] * (double)P.p.max_curvature);
        const float k2 = (float)((double)DISC_VERT[v * 2 + 1] * (double)P.p.max_curvature);
        float prop[9], pf[9], d[3];
        prepare_propagator(k1, k2, (float)FEAT_STEPS * P.p.probe_step_size, prop);
        for (int i = 0; i < 9; ++i) pf[i] = frame[i];
        float q[3] = {pos[0], pos[1], pos[2]};
        propagate_frame(prop, pf, d);
        q[0] += d[0]; q[1] += d[1]; q[2] += d[2];
        const float t1 = interp4(P, q, pf);
        propagate_frame(prop, pf, d);
        q[0] += d[0]; q[1] += d[1]; q[2] += d[2];
        const float t2 = interp4(P, q, pf);
        x[v] = t1 + t2;
    }
    x[24] = first_val;
    x[25] = x[26] = x[27] = 0.0f;
    for (int j = 0; j < 3; ++j) x[28 + j] = dir[j];
    for (int j = 0; j < 9; ++j) x[31 + j] = frame[j];
}

float mlp_neuron(const uint16_t *w, uint16_t bias, const float *in, int K) {
    float acc = bf16f(bias);
    for (int k = 0; k < K; ++k) acc = std::fmaf(bf16f(w[k]), in[k], acc);
    return acc;
}

void mlp_forward(const uint16_t *W, const float *x, float *o) {
    float h1[MLP_H], h2[MLP_H];
    for (int j = 0; j < MLP_H; ++j) {
        const float v = mlp_neuron(W + MLP_W1 + j * MLP_IN, W[MLP_B1 + j], x, MLP_IN);
        h1[j] = (v > 0.0f) ? v : 0.0f;
    }
    for (int j = 0; j < MLP_H; ++j) {
        const float v = mlp_neuron(W + MLP_W2 + j * MLP_H, W[MLP_B2 + j], h1, MLP_H);
        h2[j] = (v > 0.0f) ? v : 0.0f;
    }
    for (int j = 0; j < MLP_OUT; ++j) o[j] = mlp_neuron(W + MLP_W3 + j * MLP_H, W[MLP_B3 + j], h2, MLP_H);
}

// Learned sub-step with the exact correction (same operations as the
// kernel's get_direction_ptt_learned): faces drawn from the network's
// estimate g(f) = clipped predicted supports summed + PROPOSAL_FLOOR x their
// mean, accepted with probability F(f) / (PROPOSAL_M g(f)) where F is the
// reference's face value from the exact supports of the drawn faces'
// vertices, then the reference's support test; lowest passing lane wins.  Up
// to LEARNED_ROUNDS rounds, then (*fell = 1) the plain sub-step.  Draws whose
// F exceeds PROPOSAL_M g are counted in S.stats[3].
constexpr float PROPOSAL_M = 2.0f;
constexpr float PROPOSAL_FLOOR = 0.5f;
constexpr int LEARNED_ROUNDS = 4;

int get_direction_ptt_learned(const Problem &P, SeedScratch &S, bool learned, const float *dir, float *frame,
                              const float *pos, float *new_dir, int *fell, float *k_win = nullptr) {
    const double min_support = (double)P.p.probe_quality * PMF_THRESHOLD_P;
    const float min_support_f = (float)min_support < min_support ? std::nextafter((float)min_support, INFINITY)
                                                                  : (float)min_support;
    const float first_val = interp4(P, pos, frame);
    float vert_pdf[DISC_VERT_CNT];
    *fell = 0;
    if (learned) {
        float x[MLP_IN], o[MLP_OUT];
        mlp_features(P, pos, dir, frame, first_val, x);
        mlp_forward(P.mlp.data(), x, o);
        float gh[DISC_FACE_CNT], g[DISC_FACE_CNT], cdf[DISC_FACE_CNT], tmpa[DISC_FACE_CNT];
        for (int f = 0; f < DISC_FACE_CNT; ++f) {
            gh[f] = 0.0f;
            for (int jj = 0; jj < 3; ++jj) {
                const float sv = o[DISC_FACE[f * 3 + jj]];
                gh[f] = gh[f] + ((sv > 0.0f) ? sv : 0.0f);
            }
            tmpa[f] = gh[f];
        }
        prefix_sum(tmpa, DISC_FACE_CNT);
        const float tot = tmpa[DISC_FACE_CNT - 1];
        const float fl = PROPOSAL_FLOOR * (tot / (float)DISC_FACE_CNT) + 1e-3f;
        for (int f = 0; f < DISC_FACE_CNT; ++f) { g[f] = gh[f] + fl; cdf[f] = g[f]; }
        prefix_sum(cdf, DISC_FACE_CNT);
        const float last = cdf[DISC_FACE_CNT - 1];
        const double mc = (double)P.p.max_curvature;
        float E[DISC_VERT_CNT];
        bool known[DISC_VERT_CNT] = {};
        for (int round = 0; round < LEARNED_ROUNDS; ++round) {
            int jjs[LANES];
            float k1s[LANES], k2s[LANES], uas[LANES];
            bool need[DISC_VERT_CNT] = {};
            for (int t = 0; t < LANES; ++t) {
                Philox &rng = S.rng[t];
                float r1 = rng.uniform(), r2 = rng.uniform();
                if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
                const float tmp = rng.uniform() * last;
                uas[t] = rng.uniform();
                int jj = 0;
                for (; jj < DISC_FACE_CNT - 1; ++jj) if (cdf[jj] >= tmp) break;
                jjs[t] = jj;
                const int *fc = DISC_FACE + jj * 3;
                for (int q = 0; q < 3; ++q) need[fc[q]] = true;
                const float vx0 = (float)(mc * (double)DISC_VERT[fc[0] * 2]);
                const float vx1 = (float)(mc * (double)DISC_VERT[fc[1] * 2]);
                const float vx2 = (float)(mc * (double)DISC_VERT[fc[2] * 2]);
                const float vy0 = (float)(mc * (double)DISC_VERT[fc[0] * 2 + 1]);
                const float vy1 = (float)(mc * (double)DISC_VERT[fc[1] * 2 + 1]);
                const float vy2 = (float)(mc * (double)DISC_VERT[fc[2] * 2 + 1]);
                k1s[t] = std::fmaf(r2, vx2 - vx0, std::fmaf(r1, vx1 - vx0, vx0));
                k2s[t] = std::fmaf(r2, vy2 - vy0, std::fmaf(r1, vy1 - vy0, vy0));
            }
            for (int v = 0; v < DISC_VERT_CNT; ++v) {
                if (!need[v] || known[v]) continue;
                const float k1 = (float)((double)DISC_VERT[v * 2] * mc);
                const float k2 = (float)((double)DISC_VERT[v * 2 + 1] * mc);
                float pf[9];
                for (int i = 0; i < 9; ++i) pf[i] = frame[i];
                E[v] = data_support(P, first_val, pos, k1, k2, pf);
                known[v] = true;
            }
            int win = -1;
            for (int t = 0; t < LANES; ++t) {
                const int *fc = DISC_FACE + jjs[t] * 3;
                float F = 0.0f;
                for (int q = 0; q < 3; ++q) {
                    const float e = E[fc[q]];
                    const float vv = (e < min_support_f) ? 0.0f : e;
                    if (vv == 0.0f) { F = 0.0f; break; }
                    F = F + vv;
                }
                const float mg = PROPOSAL_M * g[jjs[t]];
                if (F > mg) ++S.stats[3];
                const bool acc = uas[t] * mg < F;
                if (acc && win < 0) {
                    float pf[9];
                    for (int i = 0; i < 9; ++i) pf[i] = frame[i];
                    if ((double)data_support(P, first_val, pos, k1s[t], k2s[t], pf) >= min_support) win = t;
                }
            }
            if (win >= 0) {
                if (k_win) { k_win[0] = k1s[win]; k_win[1] = k2s[win]; }
                float prop[9], d[3], pf[9];
                prepare_propagator(k1s[win], k2s[win], (float)(P.step_d / P.p.step_frac), prop);
                for (int i = 0; i < 9; ++i) pf[i] = frame[i];
                propagate_frame(prop, pf, d);
                norm3<0>(d);
                new_dir[0] = d[0]; new_dir[1] = d[1]; new_dir[2] = d[2];
                for (int i = 0; i < 9; ++i) frame[i] = pf[i];
                return 1;
            }
        }
        *fell = 1;
    }
    if (!ptt_vertex_supports(P, S, false, frame, pos, first_val, vert_pdf, nullptr)) return 0;
    return ptt_sample(P, S, false, dir, frame, pos, first_val, vert_pdf, 1024 / LANES, new_dir);
}

// ---------------------------------------------------------------------------
// Tracker (generate_streamlines_cuda.cu:277-421, boot.cu:720-824)
// ---------------------------------------------------------------------------
int tracker(const Problem &P, SeedScratch &S, const float *seed, const float *first_step,
            float *ptt_frame, float *sl) {
    const int model = P.p.model_type;
    const int step_frac = (model == CUSLINES_PTT) ? P.p.step_frac : 1;
    const int MAXL = P.p.max_sline_len;
    const double sub = P.step_d / (double)step_frac;
    float pt[3] = {seed[0], seed[1], seed[2]};
    float dir[3] = {first_step[0], first_step[1], first_step[2]};
    sl[0] = pt[0]; sl[1] = pt[1]; sl[2] = pt[2];
    int tissue = CUSLINES_TRACKPOINT;
    const bool use_mlp = (model == CUSLINES_PTT) && !P.mlp.empty();
    bool learned = use_mlp;
    int ltries = 0, lfalls = 0, nplain = 0;
    int i;
    for (i = 1; i < MAXL * step_frac; ++i) {
        float nd[3];
        int ndir;
        if (model == CUSLINES_PROB) ndir = get_direction_prob(P, S, false, dir, pt, nd, nullptr);
        else if (model == CUSLINES_PTT && use_mlp) {
            // learned proposal; the plain getter when a round fails or the
            // streamline left the proposal
            int fell = 0;
            ndir = get_direction_ptt_learned(P, S, learned, dir, ptt_frame, pt, nd, &fell);
            if (learned) {
                ++ltries;
                lfalls += fell;
                if (ltries >= 16 && 4 * lfalls > ltries) learned = false;
            } else {
                ++nplain;
            }
        }
        else if (model == CUSLINES_PTT) ndir = get_direction_ptt(P, S, false, dir, ptt_frame, pt, nd);
        else ndir = get_direction_boot(P, S, 5, dir, pt, nd, nullptr);
        if (ndir == 0) break;
        dir[0] = nd[0]; dir[1] = nd[1]; dir[2] = nd[2];
        for (int a = 0; a < 3; ++a) pt[a] = (float)std::fma((double)(dir[a] / 1.0f), sub, (double)pt[a]);
        if (i % step_frac == 0) {
            float *o = sl + 3 * (size_t)(i / step_frac);
            o[0] = pt[0]; o[1] = pt[1]; o[2] = pt[2];
        }
        tissue = check_point(P, pt);
        if (tissue == CUSLINES_ENDPOINT || tissue == CUSLINES_INVALIDPOINT ||
            tissue == CUSLINES_OUTSIDEIMAGE)
            break;
    }
    S.stats[0] += ltries - lfalls;
    S.stats[1] += lfalls;
    S.stats[2] += nplain;
    int nsteps = i / step_frac;
    if ((i % step_frac) != 0 && i < step_frac * (MAXL - 1)) {
        ++nsteps;
        float *o = sl + 3 * (size_t)nsteps;
        o[0] = pt[0]; o[1] = pt[1]; o[2] = pt[2];
    }
    return nsteps;
}

void init_lane_rngs(const Problem &P, SeedScratch &S, long long slid, int plus) {
    for (int l = 0; l < LANES; ++l)
        S.rng[l].init(P.p.rng_seed, (uint64_t)(LANES * slid + l + plus));
}

bool check_params(const cuslines_params *p) {
    return p && p->dimx > 0 && p->dimy > 0 && p->dimz > 0 && p->dimt > 0 && p->nverts > 0 &&
           p->nedges > 0 && p->max_sline_len > 1 && p->model_type >= 0 && p->model_type <= 3;
}

}  // namespace


// ---------------------------------------------------------------------------
// C ABI (mirrors the product's phase entry points on host arrays)
// ---------------------------------------------------------------------------
extern "C" {

struct oracle_problem;

oracle_problem *oracle_create(const cuslines_params *p) {
    if (!check_params(p)) return nullptr;
    Problem *P = new Problem();
    P->p = *p;
    P->cos_max_angle = (float)std::cos((double)p->max_angle);
    P->cos_min_sep = (float)std::cos((double)p->min_separation_angle);
    P->step_d = (double)p->step_size;
    if (p->model_type == CUSLINES_PTT) {
        P->nx = 1 << p->tile_log2_x;
        P->ny = 1 << p->tile_log2_y;
        P->nz = 1 << p->tile_log2_z;
    }
    return reinterpret_cast<oracle_problem *>(P);
}

void oracle_destroy(oracle_problem *h) { delete reinterpret_cast<Problem *>(h); }

int oracle_set_data(oracle_problem *h, const float *dataf) {
    Problem *P = reinterpret_cast<Problem *>(h);
    const size_t n = (size_t)P->p.dimx * P->p.dimy * P->p.dimz * P->p.dimt;
    P->dataf.assign(dataf, dataf + n);
    return 0;
}
int oracle_set_stop_map(oracle_problem *h, const float *s) {
    Problem *P = reinterpret_cast<Problem *>(h);
    const size_t n = (size_t)P->p.dimx * P->p.dimy * P->p.dimz;
    P->stop.assign(s, s + n);
    return 0;
}
int oracle_set_sphere(oracle_problem *h, const float *v, const int32_t *e) {
    Problem *P = reinterpret_cast<Problem *>(h);
    P->verts.assign(v, v + 3 * (size_t)P->p.nverts);
    P->edges.assign(e, e + 2 * (size_t)P->p.nedges);
    return 0;
}
int oracle_set_boot_matrices(oracle_problem *h, const float *H, const float *R, const float *db,
                             const float *dq, const float *sm, const int32_t *b0) {
    Problem *P = reinterpret_cast<Problem *>(h);
    P->b0s.assign(b0, b0 + P->p.dimt);
    int hr = 0;
    for (int j = 0; j < P->p.dimt; ++j) hr += P->b0s[j] ? 0 : 1;
    P->hr_side = hr;
    P->H.assign(H, H + (size_t)hr * hr);
    P->R.assign(R, R + (size_t)hr * hr);
    P->delta_b.assign(db, db + (size_t)P->p.delta_nr * hr);
    P->delta_q.assign(dq, dq + (size_t)P->p.delta_nr * hr);
    P->smat.assign(sm, sm + (size_t)P->p.nverts * P->p.delta_nr);
    return 0;
}
int oracle_set_proposal(oracle_problem *h, const uint16_t *w) {
    Problem *P = reinterpret_cast<Problem *>(h);
    if (!w) P->mlp.clear();
    else P->mlp.assign(w, w + CUSLINES_MLP_NPARAM);
    return 0;
}
int oracle_proposal_stats(oracle_problem *h, long long *out, int reset) {
    Problem *P = reinterpret_cast<Problem *>(h);
    for (int q = 0; q < 4; ++q) {
        if (q < 3) out[q] = P->stats[q];
        if (reset) P->stats[q] = 0;
    }
    return 0;
}
int oracle_proposal_counters(oracle_problem *h, long long *out, int reset) {
    Problem *P = reinterpret_cast<Problem *>(h);
    for (int q = 0; q < 4; ++q) {
        out[q] = P->stats[q];
        if (reset) P->stats[q] = 0;
    }
    return 0;
}
// Unit entry points of the learned proposal for the oracle's own tests.
void oracle_mlp_forward(const uint16_t *w, const float *x, float *o) { mlp_forward(w, x, o); }
int oracle_mlp_features(oracle_problem *h, const float *pos, const float *dir, const float *frame, float *x) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    mlp_features(P, pos, dir, frame, interp4(P, pos, frame), x);
    return 0;
}
// The 24 measured vertex supports of a non-initial sub-step (no random draws).
int oracle_vertex_supports(oracle_problem *h, const float *pos, const float *frame, float *raw) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    SeedScratch S;
    float vp[DISC_VERT_CNT];
    ptt_vertex_supports(P, S, false, frame, pos, interp4(P, pos, frame), vp, raw);
    return 0;
}
int oracle_set_ptt_lut(oracle_problem *h, const int32_t *lut) {
    Problem *P = reinterpret_cast<Problem *>(h);
    const size_t r = (size_t)P->p.lut_res;
    P->lut.assign(lut, lut + r * r * r);
    return 0;
}

// Phase 1: counts[i] and dirs[i*nverts + k] (vertex indices).
int oracle_phase1(oracle_problem *h, const float *seeds, int nseeds, int rng_chunk, int32_t *counts,
                  int32_t *dirs) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    if (rng_chunk <= 0) rng_chunk = nseeds;
#pragma omp parallel
    {
        SeedScratch S;
        S.size(P);
#pragma omp for schedule(dynamic, 16)
        for (int s = 0; s < nseeds; ++s) {
            const float *pt = seeds + 3 * (size_t)s;
            int *out = dirs + (size_t)s * P.p.nverts;
            int n = 0;
            init_lane_rngs(P, S, s % rng_chunk, 0);
            if (P.p.model_type == CUSLINES_PROB) {
                const float z[3] = {0, 0, 0};
                float nd[3];
                n = get_direction_prob(P, S, true, z, pt, nd, out);
            } else if (P.p.model_type == CUSLINES_PTT) {
                // getNumStreamlinesPtt_k (ptt_init.cu:35-56)
                float *pmf = S.a.data();
                const int lx = P.p.tile_log2_x, ly = P.p.tile_log2_y;
                const int DX = P.p.dimx, DY = P.p.dimy, DZ = P.p.dimz, T = P.p.dimt;
                const int W = P.nx * DX, Hh = P.ny * DY, Dd = P.nz * DZ;
                const float *d = P.dataf.data();
                auto fetch = [&](int I, int J, int K) -> float {
                    I = std::min(std::max(I, 0), W - 1);
                    J = std::min(std::max(J, 0), Hh - 1);
                    K = std::min(std::max(K, 0), Dd - 1);
                    const int v = (I / DX) + P.nx * ((J / DY) + P.ny * (K / DZ));
                    if (v >= T) return 0.0f;
                    return d[(((size_t)(I % DX) * DY + (J % DY)) * DZ + (K % DZ)) * T + v];
                };
                for (int i = 0; i < T; ++i) {
                    const int tx = i & (P.nx - 1), ty = (i >> lx) & (P.ny - 1), tz = i >> (lx + ly);
                    const float v = tex3d_linear(fetch, (float)(tx * DX) + pt[0], (float)(ty * DY) + pt[1],
                                                 (float)(tz * DZ) + pt[2]);
                    pmf[i] = ((double)v < PMF_THRESHOLD_P) ? 0.0f : v;
                }
                n = peak_directions(P, pmf, out, S.iscr.data());
            } else {
                const float z[3] = {0, 0, 0};
                float nd[3];
                n = get_direction_boot(P, S, 1, z, pt, nd, out);
            }
            counts[s] = n;
        }
    }
    return 0;
}

// Phase 2: tracks every (seed, direction) into padded slots.
int oracle_phase2(oracle_problem *h, const float *seeds, int nseeds, int rng_chunk,
                  const int32_t *offsets, const int32_t *dirs, int32_t *sline_seed,
                  int32_t *sline_len, float *sline) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    if (rng_chunk <= 0) rng_chunk = nseeds;
    const size_t slot = (size_t)2 * P.p.max_sline_len * 3;
#pragma omp parallel
    {
        SeedScratch S;
        S.size(P);
#pragma omp for schedule(dynamic, 4)
        for (int s = 0; s < nseeds; ++s) {
            const float *seed = seeds + 3 * (size_t)s;
            init_lane_rngs(P, S, s % rng_chunk, 1);
            const int ndir = offsets[s + 1] - offsets[s];
            int off = offsets[s];
            for (int i = 0; i < ndir; ++i) {
                const float *fs = P.verts.data() + 3 * (size_t)dirs[(size_t)s * P.p.nverts + i];
                float *cur = sline + (size_t)off * slot;
                sline_seed[off] = s;
                if (P.p.model_type == CUSLINES_PTT) {
                    if (!init_frame_ptt(P, S, fs, seed)) {
                        sline_len[off] = 1;
                        cur[0] = seed[0]; cur[1] = seed[1]; cur[2] = seed[2];
                        ++off;
                        continue;
                    }
                }
                const float neg[3] = {-fs[0], -fs[1], -fs[2]};
                const int nb = tracker(P, S, seed, neg, S.frame, cur);
                for (int j = 0; j < nb / 2; ++j)
                    for (int a = 0; a < 3; ++a) std::swap(cur[3 * j + a], cur[3 * (nb - 1 - j) + a]);
                const int nf = tracker(P, S, seed, fs, S.frame + 9, cur + 3 * (size_t)(nb - 1));
                sline_len[off] = nb - 1 + nf;
                ++off;
            }
        }
#pragma omp critical
        for (int q = 0; q < 4; ++q) P.stats[q] += S.stats[q];
    }
    return 0;
}

// --- unit-level entry points for the oracle's own known-answer tests -------
void oracle_philox_draws(uint64_t seed, uint64_t subseq, int n, uint32_t *out) {
    Philox p;
    p.init(seed, subseq);
    for (int i = 0; i < n; ++i) out[i] = p.next();
}
void oracle_philox_mixed(uint64_t seed, uint64_t subseq, int n, const int *kinds, float *out) {
    // kinds: 0 uniform, 1 normal
    Philox p;
    p.init(seed, subseq);
    for (int i = 0; i < n; ++i) out[i] = kinds[i] ? p.normal() : p.uniform();
}
float oracle_log(float x) { return ulog(x); }
void oracle_sincos(float x, float *s, float *c) { usincos(x, s, c); }
float oracle_rsqrt(float x) { return rsqrt_nr(x); }
int oracle_sortnet_partner(int layer, int lane) {
    const SortNet &s = sortnet();
    if (layer < 0 || layer >= s.nlayers || lane < 0 || lane >= 32) return -1;
    return s.partner[layer][lane];
}
int oracle_sortnet_layers(void) { return sortnet().nlayers; }
void oracle_warp_sort_dec(float *k, int *v) { warp_sort_dec(k, v); }
void oracle_prefix_sum(float *a, int len) { prefix_sum(a, len); }
int oracle_peaks(oracle_problem *h, const float *odf, int32_t *out) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    std::vector<int> scr(2 * (size_t)P.p.nverts + 64);
    return peak_directions(P, odf, out, scr.data());
}
int oracle_closest_peak(oracle_problem *h, const float *dir, int npeaks, const int32_t *peaks, float *out) {
    return closest_peak(*reinterpret_cast<Problem *>(h), dir, npeaks, peaks, out);
}
int oracle_trilinear(oracle_problem *h, const float *pt, float *out) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    return trilinear(P, P.dataf.data(), P.p.dimt, pt, out);
}
int oracle_check_point(oracle_problem *h, const float *pt) {
    return check_point(*reinterpret_cast<Problem *>(h), pt);
}
float oracle_interp4(oracle_problem *h, const float *pos, const float *frame) {
    return interp4(*reinterpret_cast<Problem *>(h), pos, frame);
}
// One bootstrap refit at pt with the random streams of seed slot `slot` in
// phase `phase` (0: getNumStreamlines, 1: genStreamlines): the ODF before
// thresholding into odf[nverts] and the normalised bootstrap signal into
// signal[dimt].  Returns the trilinear status (0 inside the volume).
int oracle_boot_odf(oracle_problem *h, const float *pt, int slot, int phase, float *odf, float *signal) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    SeedScratch S;
    S.size(P);
    init_lane_rngs(P, S, slot, phase);
    const int rv = boot_odf(P, S, pt, S.d.data());
    for (int j = 0; j < P.p.nverts; ++j) odf[j] = S.d[j];
    for (int j = 0; j < P.p.dimt; ++j) signal[j] = S.a[j];
    return rv;
}
void oracle_prepare_propagator(float k1, float k2, float arclen, float *p) {
    prepare_propagator(k1, k2, arclen, p);
}
void oracle_propagate_frame(const float *p, float *f, float *d) { propagate_frame(p, f, d); }
// --- sampler entry points for the oracle's distribution tests ------------
int oracle_prob_select(const float *cdf, int n, float sel) { return prob_select(cdf, n, sel); }

// The prob getter's step draw (non-initial) at point pt with incoming
// direction dir, on nstreams independent lane-0 streams (subsequences
// sub0, sub0 + 1, ...): idx_out[k] = chosen vertex or -1 (no direction),
// dirs_out[3k..] the returned direction.
int oracle_prob_sample(oracle_problem *h, const float *pt, const float *dir, int nstreams, long long sub0,
                       int32_t *idx_out, float *dirs_out) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    if (P.p.model_type != CUSLINES_PROB) return 1;
#pragma omp parallel
    {
        SeedScratch S;
        S.size(P);
#pragma omp for schedule(static)
        for (int k = 0; k < nstreams; ++k) {
            S.rng[0].init(P.p.rng_seed, (uint64_t)(sub0 + k));
            int ind = -1;
            float nd[3] = {0, 0, 0};
            if (!get_direction_prob(P, S, false, dir, pt, nd, nullptr, &ind)) ind = -1;
            idx_out[k] = ind;
            for (int a = 0; a < 3; ++a) dirs_out[3 * k + a] = nd[a];
        }
    }
    return 0;
}

// One PTT sub-step at pos with the 9-float frame (init: frame[0..2] is the
// direction, the rest is drawn), lanes on subsequences sub0 + lane.  Returns
// the getter's result; trace: k[64] candidate curvatures, sup[32] supports
// of the deciding round, then win, round; vert_pdf[24] the vertex values
// after the 0.4 cut; frame_out[9] the frame afterwards.
int oracle_ptt_substep(oracle_problem *h, const float *pos, const float *frame, int init, long long sub0,
                       float *k_out, float *sup_out, int32_t *win_round, float *vert_pdf, float *frame_out) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    if (P.p.model_type != CUSLINES_PTT) return -1;
    SeedScratch S;
    S.size(P);
    for (int l = 0; l < LANES; ++l) S.rng[l].init(P.p.rng_seed, (uint64_t)(sub0 + l));
    float fr[18];
    for (int i = 0; i < 9; ++i) fr[i] = frame[i];
    const float first_val = interp4(P, pos, fr);
    float vp[DISC_VERT_CNT];
    PttTrace tr;
    int ok = 0;
    if (init) fr[0] = frame[0], fr[1] = frame[1], fr[2] = frame[2];
    if (ptt_vertex_supports(P, S, init != 0, fr, pos, first_val, vp, nullptr)) {
        float nd[3];
        ok = ptt_sample(P, S, init != 0, frame, fr, pos, first_val, vp, 1024 / LANES, nd, &tr);
    }
    for (int v = 0; v < DISC_VERT_CNT; ++v) vert_pdf[v] = vp[v];
    for (int t = 0; t < 2 * LANES; ++t) k_out[t] = tr.k[t];
    for (int t = 0; t < LANES; ++t) sup_out[t] = tr.sup[t];
    win_round[0] = tr.win;
    win_round[1] = tr.round;
    for (int i = 0; i < 9; ++i) frame_out[i] = fr[i];
    return ok;
}

// One non-initial learned sub-step at pos with frame on lanes sub0 + lane:
// returns 1 (a learned round won; k_out its curvature), 2 (the plain
// fallback ran and found a direction), 0 (no direction); frame_out the
// frame after it, counters_out the sub-step's proposal counters.
int oracle_ptt_learned_substep(oracle_problem *h, const float *pos, const float *frame, const float *dir,
                               long long sub0, float *k_out, float *frame_out, long long *counters_out) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    if (P.p.model_type != CUSLINES_PTT || P.mlp.empty()) return -1;
    SeedScratch S;
    S.size(P);
    for (int l = 0; l < LANES; ++l) S.rng[l].init(P.p.rng_seed, (uint64_t)(sub0 + l));
    float fr[9];
    for (int i = 0; i < 9; ++i) fr[i] = frame[i];
    float nd[3];
    int fell = 0;
    k_out[0] = k_out[1] = 0.0f;
    const int ok = get_direction_ptt_learned(P, S, true, dir, fr, pos, nd, &fell, k_out);
    for (int i = 0; i < 9; ++i) frame_out[i] = fr[i];
    for (int q = 0; q < 4; ++q) counters_out[q] = S.stats[q];
    return ok ? (fell ? 2 : 1) : 0;
}

// init_frame_ptt_d for one seed and first step on lanes sub0 + lane:
// returns 1 and the 18-float frame pair, or 0.
int oracle_ptt_init_frame(oracle_problem *h, const float *seed, const float *first_step, long long sub0,
                          float *frame_out) {
    const Problem &P = *reinterpret_cast<Problem *>(h);
    if (P.p.model_type != CUSLINES_PTT) return -1;
    SeedScratch S;
    S.size(P);
    for (int l = 0; l < LANES; ++l) S.rng[l].init(P.p.rng_seed, (uint64_t)(sub0 + l));
    const bool ok = init_frame_ptt(P, S, first_step, seed);
    for (int i = 0; i < 18; ++i) frame_out[i] = S.frame[i];
    return ok ? 1 : 0;
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
}
