// Per-neuron device functions shared by the step kernels and the differentiable window:
// the EIF update, the background update, and the combination of the FP64 propagate partials.
#pragma once

#include <hip/hip_runtime.h>

#include "bn_device.h"
#include "bn_step.h"

namespace bn {

// kernel_neuron_group_evolve.cu:85-123 — refractory pin, or forward-Euler EIF step, on the neuron
// parameters as values (NeuronGroupView field names).
__device__ __forceinline__ void eif_core(float timestep, float g_leak, float membrane_capacitance,
                                         float depolarization_slope_factor, float v_effective_spike_threshold,
                                         float v_excitatory_reversal, float v_inhibitory_reversal, float v_threshold,
                                         float v_reset, float v_rest, uint32_t refractory_steps, float ge, float gi,
                                         float istim, float& V, uint32_t& refr, uint32_t& spike) {
    if (refr > 0) {
        refr = refr - 1;
        spike = 0;
        V = v_reset;
        return;
    }
    const float vv = V;
    const float linear_leak_term = -1.0f * g_leak * (vv - v_rest);
    const float exponential_depolarization_term =
        g_leak * depolarization_slope_factor *
        bn_expf((vv - v_effective_spike_threshold) / depolarization_slope_factor);
    const float excitatory_synaptic_term = -1.0f * ge * (vv - v_excitatory_reversal);
    const float inhibitory_synaptic_term = -1.0f * gi * (vv - v_inhibitory_reversal);
    const float dv = (timestep / membrane_capacitance) *
                     (linear_leak_term + exponential_depolarization_term +
                      excitatory_synaptic_term + inhibitory_synaptic_term + istim);
    const float v_next = vv + dv;
    if (v_next >= v_threshold) {
        refr = refractory_steps;
        spike = 1;
        V = v_threshold;
    } else {
        spike = 0;
        V = v_next;
    }
}

__device__ __forceinline__ void eif_step(const NeuronGroupView& v, float ge, float gi, float istim,
                                         float& V, uint32_t& refr, uint32_t& spike) {
    eif_core(v.timestep, v.g_leak, v.membrane_capacitance, v.depolarization_slope_factor,
             v.v_effective_spike_threshold, v.v_excitatory_reversal, v.v_inhibitory_reversal, v.v_threshold,
             v.v_reset, v.v_rest, v.refractory_steps, ge, gi, istim, V, refr, spike);
}

// kernel_neuron_group_evolve.cu:11-83 on values already in registers: decay (+flush) of the
// four background traces, the Poisson shot counts, the increments, the background conductances.
// head_e / head_i: the tables' first four entries when the caller loaded them with its other
// inputs (the same counts, without a dependent load in the common case), or null.
__device__ __forceinline__ void background_advance(const NeuronGroupView& v, uint32_t gid, uint32_t step,
                                                   const bn_device_clock& clk, const uint32_t* tab_e,
                                                   const uint32_t* tab_i, float& re, float& ri, float& de,
                                                   float& di, float& bg_ge, float& bg_gi,
                                                   const uint32_t (*head_e)[4] = nullptr,
                                                   const uint32_t (*head_i)[4] = nullptr) {
    re = flush_trace(re * v.background_excitatory_postsynaptic_r_trace_decay);
    ri = flush_trace(ri * v.background_inhibitory_postsynaptic_r_trace_decay);
    de = flush_trace(de * v.background_excitatory_postsynaptic_d_trace_decay);
    di = flush_trace(di * v.background_inhibitory_postsynaptic_d_trace_decay);
    uint32_t ke = 0, ki = 0;
    if (v.poisson_table_excitatory_len > 0 || v.poisson_table_inhibitory_len > 0) {
        const U32x4 r = background_draws(clk, step, gid, v.rng_instance);
        if (v.mean_background_excitatory_spikes > 0.0f)
            ke = head_e ? poisson_lookup(r.x, tab_e, v.poisson_table_excitatory_len, *head_e)
                        : poisson_from_table(r.x, tab_e, v.poisson_table_excitatory_len);
        if (v.mean_background_inhibitory_spikes > 0.0f)
            ki = head_i ? poisson_lookup(r.y, tab_i, v.poisson_table_inhibitory_len, *head_i)
                        : poisson_from_table(r.y, tab_i, v.poisson_table_inhibitory_len);
    }
    if (ke > 0u) {
        const float inc = (float)ke;
        re += inc;
        de += inc;
    }
    bg_ge = (de - re) * v.background_excitatory_scaling_factor;
    if (ki > 0u) {
        const float inc = (float)ki;
        ri += inc;
        di += inc;
    }
    bg_gi = (di - ri) * v.background_inhibitory_scaling_factor;
}

// Sum of FP64 GEMV partials into an FP32 conductance.  Order of the combination: the chunks
// of a source are cut into kFinParts contiguous parts of ceil(chunks / kFinParts); each part is
// summed in increasing chunk order from 0.0, then g' = fl32(g + part_0 + part_1 + ...) over the
// non-empty parts, in FP64.
__device__ __forceinline__ uint32_t fin_part_size(uint32_t chunks) { return (chunks + kFinParts - 1) / kFinParts; }
constexpr uint32_t kFinBatch = 16;  // partial loads issued together

__device__ __forceinline__ double fin_part_sum(const FinalizeSource& s, uint32_t j, uint32_t c0, uint32_t c1) {
    const double* p = s.partials + j;
    const size_t ld = s.ld;
    double acc = 0.0;
    for (uint32_t c = c0; c < c1; c += kFinBatch) {
        double v[kFinBatch];
#pragma unroll
        for (uint32_t k = 0; k < kFinBatch; ++k) v[k] = c + k < c1 ? p[(size_t)(c + k) * ld] : 0.0;
#pragma unroll
        for (uint32_t k = 0; k < kFinBatch; ++k)
            if (c + k < c1) acc += v[k];
    }
    return acc;
}

// Part [a0, a1) of source A and part [b0, b1) of source B with their loads issued together;
// each is summed exactly as fin_part_sum sums it.
__device__ __forceinline__ void fin_part_sum_pair(const FinalizeSource& A, uint32_t ja, uint32_t a0, uint32_t a1,
                                                  const FinalizeSource& Bs, uint32_t jb, uint32_t b0, uint32_t b1,
                                                  double& sa, double& sb) {
    const double* pa = A.partials + ja;
    const double* pb = Bs.partials + jb;
    double acc_a = 0.0, acc_b = 0.0;
    for (uint32_t ca = a0, cb = b0; ca < a1 || cb < b1; ca += kFinBatch, cb += kFinBatch) {
        double va[kFinBatch], vb[kFinBatch];
#pragma unroll
        for (uint32_t k = 0; k < kFinBatch; ++k) {
            va[k] = ca + k < a1 ? pa[(size_t)(ca + k) * A.ld] : 0.0;
            vb[k] = cb + k < b1 ? pb[(size_t)(cb + k) * Bs.ld] : 0.0;
        }
#pragma unroll
        for (uint32_t k = 0; k < kFinBatch; ++k) {
            if (ca + k < a1) acc_a += va[k];
            if (cb + k < b1) acc_b += vb[k];
        }
    }
    sa = acc_a;
    sb = acc_b;
}

constexpr uint32_t kRangeBatch = 8;

// Sums of R chunk ranges: range r adds p[r][c * ld[r]] for c in [c0[r], c1[r]) in increasing c,
// from 0.0 (the order fin_part_sum uses); the loads of a batch of every range are in flight
// together.
template <int R>
__device__ __forceinline__ void range_sums(const double* const (&p)[R], const uint32_t (&ld)[R],
                                           const uint32_t (&c0)[R], const uint32_t (&c1)[R], double (&out)[R]) {
    uint32_t n = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        out[r] = 0.0;
        n = max(n, c1[r] > c0[r] ? c1[r] - c0[r] : 0u);
    }
    for (uint32_t b = 0; b < n; b += kRangeBatch) {
        double v[R][kRangeBatch];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (uint32_t k = 0; k < kRangeBatch; ++k) {
                const uint32_t c = c0[r] + b + k;
                // Through a global-segment pointer: the partials are device memory, and a
                // pointer that went through uni() would otherwise take flat loads.
                const __attribute__((address_space(1))) double* g =
                    (const __attribute__((address_space(1))) double*)p[r];
                v[r][k] = c < c1[r] ? g[(size_t)c * ld[r]] : 0.0;
            }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (uint32_t k = 0; k < kRangeBatch; ++k)
                if (c0[r] + b + k < c1[r]) out[r] += v[r][k];
    }
}

// A plan field as a wave-uniform value.  Used where a lane selects among the fields of several
// groups: the selected loads then stay scalar loads of the kernel arguments (a select between
// two loads would otherwise become one load through a per-lane address into them).
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) { return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(x))); }
template <typename T>
__device__ __forceinline__ T* uni(T* p) {
    // readfirstlane returns int: each half goes through uint32_t so the low half is not sign-extended.
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return reinterpret_cast<T*>((hi << 32) | lo);
}

// One thread, every part in order (component path and g materialisation).
__device__ __forceinline__ float finalize_g(float g, const FinalizeSource& s, uint32_t j) {
    const uint32_t per = s.per ? s.per : fin_part_size(s.chunks);
    double acc = (double)g;
    for (uint32_t c0 = 0; per && c0 < s.chunks; c0 += per) acc += fin_part_sum(s, j, c0, min(c0 + per, s.chunks));
    return (float)acc;
}

}  // namespace bn
