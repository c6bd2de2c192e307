// device_math.h -- float routines, random streams and half-wave primitives
// for the gfx950 tracking kernels.
//
// Execution model: a wave64 carries two seed slots, one per 32-lane half.
// Every per-seed decision is uniform inside its half, so the two halves may
// diverge freely; cross-lane operations stay inside a half (width 32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SLOT_LANES 32

#define WAVE_SYNC()                                            \
    do {                                                       \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
        __builtin_amdgcn_wave_barrier();                       \
    } while (0)

// Orders memory traffic of one wave around cross-lane reuse of global data
// (points written by one lane and read back by the other lanes of a slot).
#define MEM_SYNC()                                             \
    do {                                                       \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); \
        __builtin_amdgcn_wave_barrier();                       \
    } while (0)

namespace cusl {

__device__ __forceinline__ int lane32() { return threadIdx.x & 31; }

// Ballot of the calling half-wave, as a 32-bit mask of its lanes.
__device__ __forceinline__ uint32_t half_ballot(bool p) {
    const unsigned long long b = __ballot(p);
    return (uint32_t)(b >> (threadIdx.x & 32));
}

__device__ __forceinline__ float shfl32(float v, int src) { return __shfl(v, src, 32); }
__device__ __forceinline__ int shfl32(int v, int src) { return __shfl(v, src, 32); }
__device__ __forceinline__ float shfl_xor32(float v, int m) { return __shfl_xor(v, m, 32); }
__device__ __forceinline__ int shfl_xor32(int v, int m) { return __shfl_xor(v, m, 32); }
__device__ __forceinline__ float shfl_up32(float v, int d) { return __shfl_up(v, d, 32); }

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return __builtin_fmaf(az, bz, __builtin_fmaf(ay, by, ax * bx));
}

// Sum of 32 lane partials in the pairing order of a shuffle-down tree
// (16, 8, 4, 2, 1).  Every lane of the half receives the same bits.
__device__ __forceinline__ float tree_sum32(float v) {
    v = v + shfl_xor32(v, 16);
    v = v + shfl_xor32(v, 8);
    v = v + shfl_xor32(v, 4);
    v = v + shfl_xor32(v, 2);
    v = v + shfl_xor32(v, 1);
    return v;
}

__device__ __forceinline__ float max32(float v) {
    for (int m = 16; m; m >>= 1) {
        const float o = shfl_xor32(v, m);
        v = (v > o) ? v : o;
    }
    return v;
}

__device__ __forceinline__ float min32(float v) {
    for (int m = 16; m; m >>= 1) {
        const float o = shfl_xor32(v, m);
        v = (v < o) ? v : o;
    }
    return v;
}

// 1/sqrt(n2) for finite n2 > 1e-16: Lomont's initial approximation
// (0x5f375a86 - bits/2, "Fast Inverse Square Root", 2003) and three
// Newton-Raphson steps y <- y (1.5 - (n2/2) y^2), each with one fused
// multiply-add.  Relative error ~1e-7; same sequence as the CPU oracle.
__device__ __forceinline__ float rsqrt_nr(float n2) {
    float y = __uint_as_float(0x5f375a86u - (__float_as_uint(n2) >> 1));
    const float h = 0.5f * n2;
    y = y * __builtin_fmaf(-(h * y), y, 1.5f);
    y = y * __builtin_fmaf(-(h * y), y, 1.5f);
    y = y * __builtin_fmaf(-(h * y), y, 1.5f);
    return y;
}

// Natural logarithm of a positive float: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// s = (m-1)/(m+1), log m = 2 atanh s (series to s^9), log x = e ln2 + log m
// with ln2 split in two.  Same operation sequence as the CPU oracle.
__device__ __forceinline__ float ulog(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (__builtin_isinf(x)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x *= 8388608.0f; e = -23; }
    uint32_t b = __float_as_uint(x);
    e += (int)((b >> 23) & 0xff) - 127;
    b = (b & 0x007fffffu) | 0x3f800000u;
    if (b > 0x3fb504f3u) { b -= 0x00800000u; e += 1; }
    const float f = __uint_as_float(b) - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float r = __builtin_fmaf(z, 0.22222222f, 0.28571429f);
    r = __builtin_fmaf(z, r, 0.4f);
    r = __builtin_fmaf(z, r, 0.66666669f);
    const float l1p = __builtin_fmaf(s * z, r, 2.0f * s);
    const float fe = (float)e;
    return __builtin_fmaf(fe, 0.693145751953125f, __builtin_fmaf(fe, 1.4286068203094172e-06f, l1p));
}

// sin and cos: quadrant reduction by pi/2 (three-part Cody-Waite split) and
// Taylor polynomials on |r| <= pi/4.  Same sequence as the CPU oracle.
__device__ __forceinline__ void usincos(float x, float *sn, float *cs) {
    const float q = __builtin_rintf(x * 0.636619772f);
    float r = __builtin_fmaf(-q, 1.5703125f, x);
    r = __builtin_fmaf(-q, 4.83751297e-04f, r);
    r = __builtin_fmaf(-q, 7.54978942e-08f, r);
    const float z = r * r;
    float ps = __builtin_fmaf(z, -2.50521084e-08f, 2.75573192e-06f);
    ps = __builtin_fmaf(z, ps, -1.98412698e-04f);
    ps = __builtin_fmaf(z, ps, 8.33333333e-03f);
    ps = __builtin_fmaf(z, ps, -1.66666667e-01f);
    const float s = __builtin_fmaf(r * z, ps, r);
    float pc = __builtin_fmaf(z, 2.08767570e-09f, -2.75573192e-07f);
    pc = __builtin_fmaf(z, pc, 2.48015873e-05f);
    pc = __builtin_fmaf(z, pc, -1.38888889e-03f);
    pc = __builtin_fmaf(z, pc, 4.16666667e-02f);
    pc = __builtin_fmaf(z, pc, -0.5f);
    const float c = __builtin_fmaf(z, pc, 1.0f);
    const int iq = ((int)q) & 3;
    if (iq == 0) { *sn = s; *cs = c; }
    else if (iq == 1) { *sn = c; *cs = -s; }
    else if (iq == 2) { *sn = -s; *cs = -c; }
    else { *sn = -c; *cs = s; }
}

// Philox4x32-10 (Salmon et al., SC'11) with the cuRAND-API stream layout
// (key = seed; subsequence in counter words z:w; draw k in x:y = k/4, word
// k%4).  One stream per lane of a seed slot, like the reference's
// curand_init(RNG_SEED, 32*slot + lane [+1], 0).
struct Philox {
    uint32_t k0, k1, s0, s1, c0, c1;
    uint32_t r0, r1, r2, r3;
    int idx;
    int has_normal;
    float normal_cache;

    __device__ __forceinline__ void refill() {
        uint32_t x0 = c0, x1 = c1, x2 = s0, x3 = s1;
        uint32_t key0 = k0, key1 = k1;
#pragma unroll
        for (int r = 0; r < 10; ++r) {
            const uint32_t lo0 = 0xD2511F53u * x0, hi0 = __umulhi(0xD2511F53u, x0);
            const uint32_t lo1 = 0xCD9E8D57u * x2, hi1 = __umulhi(0xCD9E8D57u, x2);
            const uint32_t n0 = hi1 ^ x1 ^ key0, n2 = hi0 ^ x3 ^ key1;
            x0 = n0; x1 = lo1; x2 = n2; x3 = lo0;
            key0 += 0x9E3779B9u;
            key1 += 0xBB67AE85u;
        }
        r0 = x0; r1 = x1; r2 = x2; r3 = x3;
    }
    __device__ __forceinline__ void init(uint64_t seed, uint64_t sub) {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        s0 = (uint32_t)sub; s1 = (uint32_t)(sub >> 32);
        c0 = 0; c1 = 0; idx = 0; has_normal = 0; normal_cache = 0.0f;
        refill();
    }
    __device__ __forceinline__ uint32_t next() {
        const uint32_t r = (idx == 0) ? r0 : (idx == 1) ? r1 : (idx == 2) ? r2 : r3;
        if (++idx == 4) {
            idx = 0;
            if (++c0 == 0) ++c1;
            refill();
        }
        return r;
    }
    __device__ __forceinline__ float uniform() {
        return 2.32830644e-10f + (float)next() * 2.32830644e-10f;
    }
    __device__ __forceinline__ float normal() {
        if (has_normal) { has_normal = 0; return normal_cache; }
        const uint32_t x = next();
        const uint32_t y = next();
        const float u = 2.32830644e-10f + (float)x * 2.32830644e-10f;
        const float v = 1.46291807e-09f + (float)y * 1.46291807e-09f;
        const float s = __builtin_sqrtf(-2.0f * ulog(u));
        float sn, cs;
        usincos(v, &sn, &cs);
        normal_cache = cs * s;
        has_normal = 1;
        return sn * s;
    }
};

}  // namespace cusl
