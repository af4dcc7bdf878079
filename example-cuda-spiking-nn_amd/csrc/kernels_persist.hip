// The persistent step: one cooperative launch advances the whole network by n steps
// (bn_persist.h describes the work and the buffers).  Each block owns a fixed list of items
// (connection, 256-column tile, segment of rows); per step it streams its items' weights once,
// using them for the propagate (kernel_connection_propagate.cu:79-94) and, in the same pass,
// applying that step's plasticity to them (kernel_triplet_plasticity_update.cu:54-93,
// kernel_istdp_plasticity_update.cu:42-80, kernel_stdp_plasticity_update.cu); the block that
// completes a column tile of a post group evolves those neurons for the next step
// (kernel_neuron_group_evolve.cu:11-130, kernel_connection_propagate.cu:8-46, the plasticity
// trace steps), and the block that completes a segment of the lazily normalised matrix updates
// its rows' sums and scales (kernel_triplet_plasticity_update.cu:117-145 as W = diag(s) M).
//
// Every wait is bounded: past the limit a block raises the state's error word (2) and leaves,
// and every other block leaves at its next wait, so the launch always ends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "bn_device.h"
#include "bn_errors.h"
#include "bn_neuron.h"
#include "bn_persist.h"

namespace bn {

constexpr uint32_t kPartBatch = 8;        // segment partials loaded together by the evolve

__device__ __forceinline__ uint32_t pbuf2(uint32_t n, uint32_t k) { return (n - 1u - k) & 1u; }
__device__ __forceinline__ uint32_t pbuf3(uint32_t n, uint32_t k) { return (n - 1u - k) % 3u; }

__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__device__ __forceinline__ uint32_t load_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0 of a block: wait until *cnt >= target.  False when the launch is being abandoned
// (this block or another ran past the limit).
__device__ __forceinline__ bool p_wait(const uint32_t* cnt, uint32_t target, uint32_t* status, uint32_t limit) {
    for (uint32_t it = 0;; ++it) {
        if (load_relaxed(cnt) >= target) return true;
        if ((it & 15u) == 15u && load_relaxed(status) >= 2u) return false;
        if (it >= limit) {
            atomicMax(status, 2u);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Whole block: thread 0 waits, the outcome is broadcast, then the acquire.
__device__ __forceinline__ bool p_block_wait(const PPlan& P, uint32_t target, uint32_t* flag) {
    if (threadIdx.x == 0) *flag = p_wait(P.tails_done, target, P.status, P.wait_limit) ? 1u : 0u;
    __syncthreads();
    const bool ok = *flag != 0u;
    fence_acquire();
    __syncthreads();
    return ok;
}

// Every thread's stores released, then thread 0 adds one to *cnt and returns its old value to
// every thread through `slot`.
__device__ __forceinline__ uint32_t p_block_arrive(uint32_t* cnt, uint32_t* slot) {
    fence_release();
    __syncthreads();
    if (threadIdx.x == 0) *slot = atomicAdd(cnt, 1u);
    __syncthreads();
    const uint32_t old = *slot;
    __syncthreads();
    return old;
}

// ---------------------------------------------------------------------------------------
// Evolve of one column tile of a group (threads 0..255, one neuron each)
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void p_evolve(const PPlan& P, uint32_t gi, uint32_t q, uint32_t n, uint32_t k, uint32_t step,
                         bool stim_prev, const bn_device_clock& clk) {
    const PGroup& G = P.g[gi];
    const NeuronGroupView& v = G.v;
    const uint32_t t = threadIdx.x;
    const uint32_t lid = q * kPCols + t;
    if (t >= (uint32_t)kPCols || lid >= v.num_neurons) return;
    const uint32_t gid = v.offset + lid;
    const uint32_t bo = pbuf2(n, k);                    // buffers this evolve writes
    const uint32_t bi = k == 0 ? 0u : pbuf2(n, k - 1);  // and the previous step's
    const uint32_t co = pbuf3(n, k);
    // Every input of the neuron, loads issued together.
    float V = v.membrane_potentials[lid];
    uint32_t refr = v.refractory_counter[lid];
    const float ge_base = v.g_excitatory[lid];
    const float gi_base = v.g_inhibitory[lid];
    float istim = v.i_stimulus[lid];
    float re = v.background_excitatory_postsynaptic_r_traces[lid];
    float ri = v.background_inhibitory_postsynaptic_r_traces[lid];
    float de = v.background_excitatory_postsynaptic_d_traces[lid];
    float di = v.background_inhibitory_postsynaptic_d_traces[lid];
    const uint32_t spike_prev = P.spikes[gid];
    uint32_t he[4], hi[4];
#pragma unroll
    for (uint32_t kk = 0; kk < 4; ++kk) {
        he[kk] = kk < v.poisson_table_excitatory_len ? v.poisson_table_excitatory[kk] : 0u;
        hi[kk] = kk < v.poisson_table_inhibitory_len ? v.poisson_table_inhibitory[kk] : 0u;
    }
    float tr_r[kMaxTraceSets], tr_d[kMaxTraceSets];
#pragma unroll
    for (int s = 0; s < kMaxTraceSets; ++s) {
        tr_r[s] = tr_d[s] = 0.0f;
        const PTraceSet& T = P.ts[s];
        if ((uint32_t)s < P.n_ts && gid - T.pre_off < T.n_pre && gid >= T.pre_off) {
            tr_r[s] = T.r[gid - T.pre_off];
            tr_d[s] = T.d[gid - T.pre_off];
        }
    }
    float pa[kMaxPlast][2], pb[kMaxPlast][2];  // [plast][pre, post]: fast and slow traces
#pragma unroll
    for (int s = 0; s < kMaxPlast; ++s) {
        pa[s][0] = pa[s][1] = pb[s][0] = pb[s][1] = 0.0f;
        const PPlast& D = P.pl[s];
        if ((uint32_t)s >= P.n_pl) continue;
        const bool trip = D.kind == kPlastTriplet;
        if (gid >= D.pre_off && gid - D.pre_off < D.n_pre) {
            pa[s][0] = D.pre[bi][gid - D.pre_off];
            if (trip) pb[s][0] = D.tpre[bi][gid - D.pre_off];
        }
        if (gid >= D.post_off && gid - D.post_off < D.n_post) {
            pa[s][1] = D.post[bi][gid - D.post_off];
            if (trip) pb[s][1] = D.tpost[bi][gid - D.post_off];
        }
    }
    // Conductances: the base plus the segment partials of the step's propagate, in order.
    float ge = ge_base, gi_ = gi_base;
    for (uint32_t side = 0; side < 2; ++side) {
        const uint32_t src = side == 0 ? G.src_e : G.src_i;
        if (src == kPNone) continue;
        const PConn& C = P.c[src];
        const double* p = C.partials + lid;
        double acc = (double)(side == 0 ? ge_base : gi_base);
        for (uint32_t s0 = 0; s0 < C.nseg; s0 += kPartBatch) {
            double pv[kPartBatch];
#pragma unroll
            for (uint32_t kk = 0; kk < kPartBatch; ++kk) pv[kk] = s0 + kk < C.nseg ? p[(size_t)(s0 + kk) * C.ld] : 0.0;
#pragma unroll
            for (uint32_t kk = 0; kk < kPartBatch; ++kk)
                if (s0 + kk < C.nseg) acc += pv[kk];
        }
        if (side == 0) ge = (float)acc;
        else gi_ = (float)acc;
    }
    // stimulus.cpp:46-60 for this neuron: the previous step's events in list order.
    if (stim_prev && P.ev_steps && step - 1 < P.ev_steps) {
        const uint32_t e1 = min(P.ev_start[step], P.ev_count);
        bool hit = false;
        for (uint32_t e = P.ev_start[step - 1]; e < e1; ++e)
            if (P.ev_ids[e] + P.ev_offsets[e] == gid) {
                istim += P.ev_amps[e];
                hit = true;
            }
        if (hit) v.i_stimulus[lid] = istim;
    }
    // Background (kernel_neuron_group_evolve.cu:11-83).
    re = flush_trace(re * v.background_excitatory_postsynaptic_r_trace_decay);
    ri = flush_trace(ri * v.background_inhibitory_postsynaptic_r_trace_decay);
    de = flush_trace(de * v.background_excitatory_postsynaptic_d_trace_decay);
    di = flush_trace(di * v.background_inhibitory_postsynaptic_d_trace_decay);
    uint32_t ke = 0, ki = 0;
    if (v.poisson_table_excitatory_len > 0 || v.poisson_table_inhibitory_len > 0) {
        const U32x4 r = background_draws(clk, step, gid, v.rng_instance);
        if (v.mean_background_excitatory_spikes > 0.0f)
            ke = poisson_lookup(r.x, v.poisson_table_excitatory, v.poisson_table_excitatory_len, he);
        if (v.mean_background_inhibitory_spikes > 0.0f)
            ki = poisson_lookup(r.y, v.poisson_table_inhibitory, v.poisson_table_inhibitory_len, hi);
    }
    if (ke > 0u) {
        const float inc = (float)ke;
        re += inc;
        de += inc;
    }
    const float bg_ge = (de - re) * v.background_excitatory_scaling_factor;
    if (ki > 0u) {
        const float inc = (float)ki;
        ri += inc;
        di += inc;
    }
    const float bg_gi = (di - ri) * v.background_inhibitory_scaling_factor;
    // EIF (kernel_neuron_group_evolve.cu:85-130).
    uint32_t spike = 0;
    eif_step(v, ge, gi_, istim, V, refr, spike);
    v.membrane_potentials[lid] = V;
    v.refractory_counter[lid] = refr;
    P.spikes[gid] = spike;
    v.g_excitatory[lid] = bg_ge;
    v.g_inhibitory[lid] = bg_gi;
    v.background_excitatory_postsynaptic_r_traces[lid] = re;
    v.background_inhibitory_postsynaptic_r_traces[lid] = ri;
    v.background_excitatory_postsynaptic_d_traces[lid] = de;
    v.background_inhibitory_postsynaptic_d_traces[lid] = di;
    v.background_g_excitatory_synaptic_conductances[lid] = bg_ge;
    v.background_g_inhibitory_synaptic_conductances[lid] = bg_gi;
    // Plasticity traces: the previous step's +1 then this step's decay.
#pragma unroll
    for (int s = 0; s < kMaxPlast; ++s) {
        const PPlast& D = P.pl[s];
        if ((uint32_t)s >= P.n_pl) continue;
        const bool trip = D.kind == kPlastTriplet;
        if (gid >= D.pre_off && gid - D.pre_off < D.n_pre) {
            const uint32_t i = gid - D.pre_off;
            D.pre[bo][i] = flush_trace((spike_prev == 1 ? pa[s][0] + 1.0f : pa[s][0]) * D.pre_decay);
            if (trip) D.tpre[bo][i] = flush_trace((spike_prev == 1 ? pb[s][0] + 1.0f : pb[s][0]) * D.tpre_decay);
        }
        if (gid >= D.post_off && gid - D.post_off < D.n_post) {
            const uint32_t j = gid - D.post_off;
            D.post[bo][j] = flush_trace((spike_prev == 1 ? pa[s][1] + 1.0f : pa[s][1]) * D.post_decay);
            if (trip) D.tpost[bo][j] = flush_trace((spike_prev == 1 ? pb[s][1] + 1.0f : pb[s][1]) * D.tpost_decay);
        }
    }
    // Connection traces (kernel_connection_propagate.cu:8-46).
#pragma unroll
    for (int s = 0; s < kMaxTraceSets; ++s) {
        const PTraceSet& T = P.ts[s];
        if ((uint32_t)s < P.n_ts && gid >= T.pre_off && gid - T.pre_off < T.n_pre) {
            const uint32_t i = gid - T.pre_off;
            float dd = flush_trace(tr_d[s] * T.decay_d);
            float rr = flush_trace(tr_r[s] * T.decay_r);
            if (spike == 1) {
                rr += T.K;
                dd += T.K;
            }
            T.d[i] = dd;
            T.r[i] = rr;
            T.x[bo][i] = dd - rr;
        }
    }
    // Spike bits (cleared two steps ahead) and the monitors' rings.
    if (spike == 1) {
        const uint32_t bit = 1u << (gid & 31u);
        atomicOr(&P.bits[co][gid >> 5], bit);
        if (P.spike_ring) atomicOr(&P.spike_ring[(size_t)(step % P.interval) * P.words + (gid >> 5)], bit);
    }
    if (P.v_ring) {
        const int32_t slot = P.vrec_index[gid];
        if (slot >= 0) P.v_ring[(size_t)(step % P.interval) * P.n_vrec + (uint32_t)slot] = V;
    }
}

// ---------------------------------------------------------------------------------------
// Row sums and scales of one segment of the lazily normalised matrix after the step's
// plasticity: R_i = 0 + T_0 + T_1 + ... (pre neuron spiked: the tiles' sums of the new row) or
// R_i + dT_0 + dT_1 + ... (the tiles' sums of the row's LTP changes); s_i = target_i / R_i.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void p_row_tail(const PPlan& P, uint32_t seg, uint32_t cb) {
    const PPlast& D = P.pl[P.lazy_pl];
    const PConn& C = P.c[D.conn];
    const uint32_t r0 = seg * kPWaves * C.rb;
    const uint32_t r1 = min(r0 + kPWaves * C.rb, C.n_pre);
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += kPThreads) {
        const uint32_t g = C.pre_off + i;
        const bool spiked = (P.bits[cb][g >> 5] >> (g & 31u)) & 1u;
        double R = spiked ? 0.0 : D.R[i];
        for (uint32_t q = 0; q < C.ncol; ++q) R += D.dR[(size_t)q * C.n_pre + i];
        D.R[i] = R;
        if (R > 0.0) D.s[i] = (float)((double)D.target[i] / R);
    }
}

// ---------------------------------------------------------------------------------------
// One item: chunk sums of the 8 waves, the plasticity of the step on the entries, the segment
// partial.
// ---------------------------------------------------------------------------------------
template <bool kNT>
__device__ __forceinline__ float4 p_load4(const float* p) {
    if constexpr (kNT) {
        typedef float f4v __attribute__((ext_vector_type(4)));
        const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
        return make_float4(t.x, t.y, t.z, t.w);
    } else {
        return *reinterpret_cast<const float4*>(p);
    }
}

__device__ __forceinline__ float f4_get(const float4& v, uint32_t c) {
    return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}
__device__ __forceinline__ void f4_set(float4& v, uint32_t c, float x) {
    if (c == 0) v.x = x;
    else if (c == 1) v.y = x;
    else if (c == 2) v.z = x;
    else v.w = x;
}

// Entry update in true units (lazy: w = s*m, stored back as w'/s), as apply_delta.
__device__ __forceinline__ float p_apply(bool lazy, float m, float si, float delta, float lo, float hi) {
    if (lazy) return clamp_weight(si * m + delta, lo, hi) / si;
    return clamp_weight(m + delta, lo, hi);
}

// The plasticity of the step on one row i of an item (lane: columns j0 .. j0+3): per entry the
// LTP of a spiking post column then the LTD of a spiking pre row (index self-skip), stores of the
// touched entries, and for the lazy matrix the row's tile sum (pre spiked: of the new entries;
// else of the LTP changes) into dR_tile[i].
__device__ __forceinline__ void p_plast_row(const PPlast& D, bool lazy, bool triplet, bool sp, uint32_t i, uint32_t b,
                                            uint32_t j0, bool active, uint32_t n_post, uint32_t colspike,
                                            const float (&post_j)[4], const float (&fac_ltp)[4], const float4 m4,
                                            float* dst, double* dR_tile) {
    const float si = lazy ? D.s[i] : 1.0f;
    const float pre_i = D.pre[b][i];
    const float fac_i = triplet ? D.p.ltd_A + (D.p.ltd_B * D.tpre[b][i]) : 0.0f;
    const bool update = !lazy || si != 0.0f;
    float4 nv = m4;
    uint32_t touched = 0;
    double lsum = 0.0;
#pragma unroll
    for (uint32_t c4 = 0; c4 < 4; ++c4) {
        const uint32_t j = j0 + c4;
        if (!active || j >= n_post) continue;
        const float m = f4_get(m4, c4);
        float val = m;
        bool t = false;
        if (update && j != i) {
            if ((colspike >> c4) & 1u) {
                const float delta = triplet ? pre_i * fac_ltp[c4] : D.p.ltp_A * pre_i;
                val = p_apply(lazy, val, si, delta, D.p.weight_min, D.p.weight_max);
                t = true;
            }
            if (sp) {
                float d;
                if (triplet) d = -1.0f * post_j[c4] * fac_i;
                else if (D.kind == kPlastIstdp) d = -1.0f * D.p.ltd_A * (post_j[c4] - D.p.ltd_B);
                else d = -1.0f * D.p.ltd_A * post_j[c4];
                val = p_apply(lazy, val, si, d, D.p.weight_min, D.p.weight_max);
                t = true;
            }
        }
        if (lazy) lsum += sp ? (double)val : (t ? (double)val - (double)m : 0.0);
        if (t) {
            f4_set(nv, c4, val);
            touched |= 1u << c4;
        }
    }
    if (touched == 0xFu) {
        *reinterpret_cast<float4*>(dst) = nv;
    } else if (touched) {
#pragma unroll
        for (uint32_t c4 = 0; c4 < 4; ++c4)
            if ((touched >> c4) & 1u) dst[c4] = f4_get(nv, c4);
    }
    if (lazy) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
        if ((threadIdx.x & 63u) == 0) dR_tile[i] = lsum;
    }
}

template <bool kNT>
__device__ __forceinline__ void p_item(const PPlan& P, const PItem& it, uint32_t n, uint32_t k, double (*red)[kPCols]) {
    const PConn& C = P.c[it.conn];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t b = pbuf2(n, k), cb = pbuf3(n, k);
    const uint32_t rb = C.rb;
    const uint32_t i0 = it.seg * kPWaves * rb + wave * rb;
    const uint32_t i1 = min(i0 + rb, C.n_pre);
    const uint32_t j0 = (uint32_t)it.col * kPCols + 4u * lane;
    const bool active = j0 < C.n_post;
    const float* __restrict__ x = C.x[b];
    const float* __restrict__ s = C.s;
    const bool plastic = C.plast != kPNone;
    const PPlast& D = P.pl[plastic ? C.plast : 0];
    const bool lazy = plastic && D.lazy;
    const bool triplet = plastic && D.kind == kPlastTriplet;
    // Columns of the lane: spike flags of the step, post traces, LTP factors.
    uint32_t colspike = 0;
    float post_j[4] = {0.0f, 0.0f, 0.0f, 0.0f}, fac_ltp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (plastic && active) {
#pragma unroll
        for (uint32_t c4 = 0; c4 < 4; ++c4) {
            const uint32_t j = j0 + c4;
            if (j >= C.n_post) continue;
            const uint32_t g = C.post_off + j;
            if ((P.bits[cb][g >> 5] >> (g & 31u)) & 1u) colspike |= 1u << c4;
            post_j[c4] = D.post[b][j];
            if (triplet) fac_ltp[c4] = D.p.ltp_A + (D.p.ltp_B * D.tpost[b][j]);
        }
    }
    const bool any_col = plastic && __ballot(colspike != 0u) != 0ull;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    float* wrow = C.W + (size_t)i0 * C.ld + j0;
    for (uint32_t r0 = i0; r0 < i1; r0 += kPBatch) {
        const uint32_t nb = min((uint32_t)kPBatch, i1 - r0);
        // Loads of the batch issued together (rows past the chunk re-read its first row and are
        // not used; inactive lanes read nothing).
        float4 w[kPBatch];
        float xs[kPBatch];
#pragma unroll
        for (uint32_t kk = 0; kk < (uint32_t)kPBatch; ++kk) {
            const uint32_t r = kk < nb ? kk : 0u;
            w[kk] = active ? p_load4<kNT>(wrow + (size_t)r * C.ld) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            xs[kk] = s ? x[r0 + r] * s[r0 + r] : x[r0 + r];
        }
#pragma unroll
        for (uint32_t kk = 0; kk < (uint32_t)kPBatch; ++kk) {
            if (kk >= nb) continue;
            const double xd = (double)xs[kk];
            a0 += xd * (double)w[kk].x;
            a1 += xd * (double)w[kk].y;
            a2 += xd * (double)w[kk].z;
            a3 += xd * (double)w[kk].w;
        }
        if (plastic) {
            // Spike flags of the batch's pre neurons (<= 16 rows: at most two words).
            const uint32_t g0 = C.pre_off + r0, gl = g0 + nb - 1;
            const uint64_t two = (uint64_t)P.bits[cb][g0 >> 5] | ((uint64_t)P.bits[cb][gl >> 5] << 32);
            const uint32_t rowbits = (uint32_t)(((g0 >> 5) == (gl >> 5) ? (two & 0xFFFFFFFFull) : two) >> (g0 & 31u)) &
                                     ((1u << nb) - 1u);
#pragma unroll
            for (uint32_t kk = 0; kk < (uint32_t)kPBatch; ++kk) {
                const bool sp = (rowbits >> kk) & 1u;
                if (kk < nb && (sp || any_col))
                    p_plast_row(D, lazy, triplet, sp, r0 + kk, b, j0, active, C.n_post, colspike, post_j, fac_ltp,
                                w[kk], wrow + (size_t)kk * C.ld, D.dR + (size_t)it.col * C.n_pre);
                else if (kk < nb && lazy && lane == 0)
                    D.dR[(size_t)it.col * C.n_pre + r0 + kk] = 0.0;
            }
        }
        wrow += (size_t)kPBatch * C.ld;
    }
    // Segment partial: the 8 chunk sums in chunk order from 0.0.
    red[wave][4 * lane + 0] = a0;
    red[wave][4 * lane + 1] = a1;
    red[wave][4 * lane + 2] = a2;
    red[wave][4 * lane + 3] = a3;
    __syncthreads();
    if (threadIdx.x < (uint32_t)kPCols) {
        const uint32_t col = (uint32_t)it.col * kPCols + threadIdx.x;
        const uint32_t seg_rows = min(kPWaves * rb, C.n_pre - it.seg * kPWaves * rb);
        const uint32_t chunks = (seg_rows + rb - 1) / rb;
        if (col < C.n_post) {
            double acc = 0.0;
            for (uint32_t w8 = 0; w8 < chunks; ++w8) acc += red[w8][threadIdx.x];
            C.partials[(size_t)it.seg * C.ld + col] = acc;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// The launch
// ---------------------------------------------------------------------------------------
template <bool kNT>
__global__ __launch_bounds__(kPThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_persist(const PPlan* __restrict__ plan, uint32_t n) {
    __shared__ double red[kPWaves][kPCols];
    __shared__ uint32_t flag, slot;
    const PPlan& P = *plan;
    const bn_device_clock clk = *P.clock;
    const uint32_t t0 = clk.clock_step;
    const uint32_t per_step = P.n_coltiles + P.n_rowtails;
    const uint32_t blk = blockIdx.x;
    const uint32_t ib = P.block_begin[blk], ie = P.block_begin[blk + 1];
    // Phase 0: the evolve of the launch's first step, one column tile per block.
    for (uint32_t ct = blk; ct < P.n_coltiles; ct += gridDim.x) {
        uint32_t gi = 0;
        while (gi + 1 < P.n_groups && ct >= P.g[gi + 1].col_base) ++gi;
        p_evolve(P, gi, ct - P.g[gi].col_base, n, 0, t0, false, clk);
        __syncthreads();
        (void)p_block_arrive(P.tails_done, &slot);
    }
    for (uint32_t k = 0; k < n; ++k) {
        // Every tail of the previous step (the evolve of step k included) has finished.
        if (!p_block_wait(P, P.n_coltiles + k * per_step, &flag)) return;
        if (blk == 0 && k + 2 < n) {
            // Spike bits of evolve k + 2 and its ring slot, cleared ahead (their last readers
            // are the items of step k - 1).
            uint32_t* bits = P.bits[pbuf3(n, k + 2)];
            for (uint32_t w = threadIdx.x; w < P.words; w += kPThreads) bits[w] = 0u;
            if (P.spike_ring) {
                uint32_t* row = P.spike_ring + (size_t)((t0 + k + 2) % P.interval) * P.words;
                for (uint32_t w = threadIdx.x; w < P.words; w += kPThreads) row[w] = 0u;
            }
        }
        for (uint32_t e = ib; e < ie; ++e) {
            const PItem it = P.items[e];
            p_item<kNT>(P, it, n, k, red);
            const PConn& C = P.c[it.conn];
            const PGroup& G = P.g[C.post_group];
            const uint32_t ct = G.col_base + it.col;
            const uint32_t old_c = p_block_arrive(&P.col_count[ct], &slot);
            const bool last_col = old_c == (k + 1) * G.expect - 1u;
            bool last_row = false;
            if (P.lazy_pl != kPNone && C.plast == P.lazy_pl) {
                const uint32_t old_r = p_block_arrive(&P.row_count[it.seg], &slot);
                last_row = old_r == (k + 1) * C.ncol - 1u;
            }
            if (last_col) {
                fence_acquire();
                if (k + 1 < n) p_evolve(P, C.post_group, it.col, n, k + 1, t0 + k + 1, true, clk);
                __syncthreads();
                (void)p_block_arrive(P.tails_done, &slot);
            }
            if (last_row) {
                fence_acquire();
                p_row_tail(P, it.seg, pbuf3(n, k));
                __syncthreads();
                (void)p_block_arrive(P.tails_done, &slot);
            }
        }
    }
    // The last step's stimulus events and the clock advance of the whole launch, once every tail
    // of the launch has finished (so every block has read the clock and no evolve is pending).
    if (blk == 0) {
        if (!p_block_wait(P, P.n_coltiles + n * per_step, &flag)) return;
        if (threadIdx.x == 0) {
            const uint32_t last = t0 + n - 1;
            if (P.ev_steps && last < P.ev_steps) {
                const uint32_t e1 = min(P.ev_start[last + 1], P.ev_count);
                for (uint32_t e = P.ev_start[last]; e < e1; ++e) P.i_stimulus[P.ev_ids[e] + P.ev_offsets[e]] += P.ev_amps[e];
            }
            bn_device_clock c = clk;
            c.clock_step = t0 + n;
            c.next_event_idx = P.ev_steps ? P.ev_start[min(t0 + n, P.ev_steps)] : clk.next_event_idx;
            *P.clock_rw = c;
        }
    }
}

int persistent_max_blocks(uint32_t* blocks) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_persist<false>),
                                                         kPThreads, 0);
    if (e == hipSuccess) {
        int per_cu_nt = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_nt, reinterpret_cast<const void*>(k_persist<true>),
                                                         kPThreads, 0);
        per_cu = std::min(per_cu, per_cu_nt);
    }
    *blocks = e == hipSuccess ? (uint32_t)std::max(0, cus * per_cu) : 0u;
    return (int)e;
}

int launch_persistent(const PPlan* d_plan, const PPlan& hp, uint32_t n, uint32_t nontemporal, bn_stream_t s) {
    if (n == 0 || hp.blocks == 0) return 0;
    void* args[] = {(void*)&d_plan, (void*)&n};
    const void* f = nontemporal ? reinterpret_cast<const void*>(k_persist<true>)
                                : reinterpret_cast<const void*>(k_persist<false>);
    return (int)hipLaunchCooperativeKernel(f, dim3(hp.blocks), dim3(kPThreads), args, 0, s);
}

}  // namespace bn
