// Device-side building blocks shared by every kernel of the step: counter-based Philox,
// the engine's own expf, Poisson inverse-CDF sampling, trace flush and weight clamp.
//
// All arithmetic here is plain IEEE-754 binary32/binary64 (the library is compiled with
// -ffp-contract=off), so the CPU restatement of the step reproduces it bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "buryn_amd.h"

namespace bn {

// Traces below this are flushed to zero (kernel_neuron_group_evolve.cu:23,
// kernel_connection_propagate.cu:14, kernel_triplet_plasticity_update.cu:24).
constexpr float kTraceFlush = 1e-30f;

__device__ __forceinline__ float flush_trace(float v) { return v < kTraceFlush ? 0.0f : v; }

// Clamp as triplet_clamp_weight / istdp_clamp_weight (kernel_triplet_plasticity_update.cu:6-14).
__device__ __forceinline__ float clamp_weight(float w, float lo, float hi) {
    if (w < lo) return lo;
    if (w > hi) return hi;
    return w;
}

// ---------------------------------------------------------------------------------------
// Philox4x32-10: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11.  Multipliers and Weyl key increments are the
// published constants of the method.
// ---------------------------------------------------------------------------------------
struct U32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c.x;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = U32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

// Background draws of one neuron at one step: counter (step, stream, global id, instance).
__device__ __forceinline__ U32x4 background_draws(const bn_device_clock& clk, uint32_t step,
                                                  uint32_t global_id, uint32_t instance) {
    return philox4x32_10(U32x4{step, (uint32_t)BN_RNG_STREAM_BACKGROUND, global_id, instance},
                         clk.rng_key[0], clk.rng_key[1]);
}

// Poisson sample by inversion of a host-built CDF table: min{k : u < table[k]}, else len.
__device__ __forceinline__ uint32_t poisson_from_table(uint32_t u, const uint32_t* table,
                                                       uint32_t len) {
    uint32_t k = 0;
    while (k < len && u >= table[k]) ++k;
    return k;
}

// poisson_from_table with the first four entries already in registers (loaded with the other
// inputs of the step): min{k : u < table[k]}, the same count.
__device__ __forceinline__ uint32_t poisson_lookup(uint32_t u, const uint32_t* table, uint32_t len,
                                                   const uint32_t (&head)[4]) {
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
        if (k >= len || u < head[k]) return k;
    return poisson_from_table(u, table + 4, len - 4) + 4;
}

// ---------------------------------------------------------------------------------------
// expf of the engine: Cody–Waite argument reduction x = k*ln2 + r with ln2 split into
// 0.693359375 (exact, 9 significant bits) and -2.12194440e-4, exp(r) by the degree-7
// Taylor polynomial in Horner form (|r| <= 0.35, truncation < 6e-9 relative), and an
// exact power-of-two scaling (two steps where 2^k is not a normal float).
// Worst-case error against the exact exponential is about 2 ulp.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float pow2_float(int e) {  // e in [-126, 127]
    return __int_as_float((e + 127) << 23);
}

__device__ __forceinline__ float bn_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __int_as_float(0x7f800000);
    if (x < -103.97208404541015625f) return 0.0f;
    const float kf = __builtin_rintf(x * 1.44269502162933349609375f);
    const int k = (int)kf;
    float r = x - kf * 0.693359375f;
    r = r - kf * -2.12194440e-4f;
    float p = 1.98412698e-4f;                // 1/7!
    p = p * r + 1.38888889e-3f;              // 1/6!
    p = p * r + 8.33333377e-3f;              // 1/5!
    p = p * r + 4.16666679e-2f;              // 1/4!
    p = p * r + 1.66666672e-1f;              // 1/3!
    p = p * r + 0.5f;                        // 1/2!
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    if (k > 127) return (p * pow2_float(k - 1)) * 2.0f;
    if (k < -126) return (p * pow2_float(k + 126)) * pow2_float(-126);
    return p * pow2_float(k);
}

}  // namespace bn
