// ud_device.h -- device helpers shared by the splatting kernels (gfx950).
//
// All translation units are compiled with -ffp-contract=off: the tree build and
// the sampler must round exactly like their CPU restatement, so every fused
// multiply-add below is written out explicitly with fmaf().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define UD_PI 3.14159265358979323846

// pointer into device (global) memory: lets the compiler emit global_load
// instead of flat_load for data reached through structs and non-inlined calls
#define UD_GLOBAL __attribute__((address_space(1)))

namespace ud {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// FP32 exp / log with a fixed operation sequence (identical on CPU and GPU).
// exp: x = k ln2 + r, |r| <= ln2/2 (two-part ln2), degree-7 Taylor polynomial
// of e^r in Horner form, scaled by 2^k built from the exponent bits.
// Arguments below -87 flush to 0 (results stay normal numbers).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pexpf(float x)
{
    if (x != x) return x;
    if (x > 88.0f) return __builtin_huge_valf();
    if (x < -87.0f) return 0.0f;
    const float log2e = 1.44269504088896341f;
    const float ln2hi = 0.693145751953125f;
    const float ln2lo = 1.428606765330187045e-06f;
    float kf = rintf(x * log2e);
    float r = fmaf(-kf, ln2hi, x);
    r = fmaf(-kf, ln2lo, r);
    float p = 1.0f / 5040.0f;
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int k = (int)kf;
    float sc = __uint_as_float((uint32_t)(k + 127) << 23);
    return p * sc;
}

// log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)).
__device__ __forceinline__ float plogf(float x)
{
    if (x != x) return x;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_huge_valf();
    if (x == __builtin_huge_valf()) return x;
    uint32_t u = __float_as_uint(x);
    int e = 0;
    if (u < 0x00800000u) { u = __float_as_uint(x * 16777216.0f); e = -24; }
    e += (int)((u >> 23) & 0xffu) - 127;
    float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 1.0f / 9.0f;
    p = fmaf(p, s2, 1.0f / 7.0f);
    p = fmaf(p, s2, 1.0f / 5.0f);
    p = fmaf(p, s2, 1.0f / 3.0f);
    p = fmaf(p, s2, 1.0f);
    float lm = 2.0f * s * p;
    const float ln2hi = 0.693145751953125f;
    const float ln2lo = 1.428606765330187045e-06f;
    float ef = (float)e;
    return fmaf(ef, ln2hi, fmaf(ef, ln2lo, lm));
}

// log(1 - u) for u in [0, 1]: the log1p series -(u + u^2/2 + ... + u^8/8) for
// u < 1/16 (truncation error < u^9 / 9 relative), plogf(1 - u) otherwise.
__device__ __forceinline__ float plog1mf(float u)
{
    if (u >= 1.0f) return -__builtin_huge_valf();
    if (u < 0.0625f) {
        float p = 1.0f / 8.0f;
        p = p * u + 1.0f / 7.0f;
        p = p * u + 1.0f / 6.0f;
        p = p * u + 1.0f / 5.0f;
        p = p * u + 1.0f / 4.0f;
        p = p * u + 1.0f / 3.0f;
        p = p * u + 0.5f;
        p = p * u + 1.0f;
        return -(p * u);
    }
    return plogf(1.0f - u);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 counter-based generator (Salmon et al., SC'11).
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32 x 32 -> 64-bit product per multiplier (a single v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        u32x4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
    }
    return c;
}

// uniform in (0, 1]: the 24 high bits plus one, times 2^-24 (exact)
__device__ __forceinline__ float u01(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f;
}

// ---------------------------------------------------------------------------
// Small vector helpers (explicit operation order: ((a0 b0 + a1 b1) + a2 b2)).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float dot3(const float* a, const float* b)
{
    float x = a[0] * b[0];
    x = x + a[1] * b[1];
    x = x + a[2] * b[2];
    return x;
}

// toScreenspace (FP32, the order of the sampler/tree restatement)
__device__ __forceinline__ void screen_f32(float Wf, float Hf, float ff, const float* p, float& sx, float& sy)
{
    float t = p[0] * ff;
    t = -t;
    sx = t / p[2] + 0.5f * Wf;
    float u = p[1] * ff;
    sy = u / p[2] + 0.5f * Hf;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// wavefront-wide inclusive prefix sum of an int (64 lanes)
__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (l >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// XCD-aware block remap: the hardware dispatches consecutive workgroups
// round-robin over the 8 XCDs; remap so each XCD works on a contiguous chunk
// of tiles (neighbouring pixel tiles share tree nodes in that XCD's L2).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks)
{
    const uint32_t nx = 8;
    if (nblocks < nx) return bid;
    uint32_t per = nblocks / nx, rem = nblocks % nx;
    uint32_t x = bid % nx, i = bid / nx;
    // xcd x owns chunk [start_x, start_x + len_x)
    uint32_t start = x * per + (x < rem ? x : rem);
    uint32_t len = per + (x < rem ? 1u : 0u);
    return (i < len) ? start + i : bid;
}

// Chunked round-robin remap for per-tile passes whose cost varies across the
// image: runs of LC consecutive tiles are dealt to the 8 XCDs in turn (XCD x
// takes runs x, x + 8, ...), so every XCD gets a sample of the whole image (its
// dense and its empty parts) while a run's neighbouring tiles still share that
// XCD's L2.  The last nblocks mod 8·LC blocks keep their own tile.  A permutation
// of [0, nblocks).
template <uint32_t LC>
__device__ __forceinline__ uint32_t xcd_chunk_remap(uint32_t bid, uint32_t nblocks)
{
    const uint32_t full = nblocks / (8u * LC) * (8u * LC);
    if (bid >= full) return bid;
    const uint32_t x = bid % 8u, i = bid / 8u;
    return (x + 8u * (i / LC)) * LC + i % LC;
}

} // namespace ud
