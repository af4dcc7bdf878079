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
#include "bn_persist_dev.h"

namespace bn {

constexpr uint32_t kPartBatch = 8;        // segment partials loaded together by the evolve

__device__ __forceinline__ uint32_t pbuf2(uint32_t n, uint32_t k) { return (n - 1u - k) & 1u; }
__device__ __forceinline__ uint32_t pbuf3(uint32_t n, uint32_t k) { return (n - 1u - k) % 3u; }

// Thread 0 of a block: wait until *cnt >= target.  False when the launch is being abandoned
// (this block or another ran past the limit).
__device__ __forceinline__ bool p_wait(const uint32_t* cnt, uint32_t target, uint32_t* status, uint32_t limit) {
    for (uint32_t it = 0;; ++it) {
        if (ld_c(cnt) >= target) return true;
        if ((it & 15u) == 15u && ld_c(status) >= 2u) return false;
        if (it >= limit) {
            max_c(status, 2u);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Block-shared memory: the chunk sums of an item and the staged rows of its segment.
constexpr uint32_t kPSegMax = kPWaves * kPMaxRows;   // rows of a segment
struct PShared {
    double red[kPWaves][kPCols];   // chunk sums of the item's columns
    double dr[kPSegMax];           // lazy matrix: the row's tile sum
    float xs[kPSegMax];            // x_i * s_i (x_i without row scales)
    float pre[kPSegMax];           // plastic matrix: pre trace of the row
    float fac[kPSegMax];           // triplet: ltd_A + ltd_B * slow pre trace
    float si[kPSegMax];            // lazy matrix: row scale (else 1)
    uint32_t rbits[kPSegMax / 32]; // spike flags of the rows (bit r of word r/32: row r)
    float gsum[kPCols];            // evolve: g_i of the tile's neurons
    uint32_t flag, slot;
};

// Whole block: thread 0 waits, the outcome is broadcast.
__device__ __forceinline__ bool p_block_wait(const PPlan& P, uint32_t target, PShared& S) {
    if (threadIdx.x == 0) S.flag = p_wait(P.tails_done, target, P.status, P.wait_limit) ? 1u : 0u;
    __syncthreads();
    const bool ok = S.flag != 0u;
    __syncthreads();
    return ok;
}

// Every store of every thread of the block complete, then thread 0 adds one to *cnt; its old
// value is returned to every thread.
__device__ __forceinline__ uint32_t p_block_arrive(uint32_t* cnt, PShared& S) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) S.slot = add_c(cnt, 1u);
    __syncthreads();
    const uint32_t old = S.slot;
    __syncthreads();
    return old;
}

// g = fl32(base + p_0 + p_1 + ...) over the segment partials of column j of a connection, in
// FP64 in segment order; the first 16 partials are requested together.
__device__ __forceinline__ float p_gsum(const PConn& C, uint32_t j, float base) {
    const double* p = C.partials + j;
    double acc = (double)base;
    double pv[16];
#pragma unroll
    for (uint32_t kk = 0; kk < 16; ++kk) pv[kk] = kk < C.nseg ? ld_c(p + (size_t)kk * C.ld) : 0.0;
#pragma unroll
    for (uint32_t kk = 0; kk < 16; ++kk)
        if (kk < C.nseg) acc += pv[kk];
    for (uint32_t s0 = 16; s0 < C.nseg; s0 += kPartBatch) {
        double qv[kPartBatch];
#pragma unroll
        for (uint32_t kk = 0; kk < kPartBatch; ++kk) qv[kk] = s0 + kk < C.nseg ? ld_c(p + (size_t)(s0 + kk) * C.ld) : 0.0;
#pragma unroll
        for (uint32_t kk = 0; kk < kPartBatch; ++kk)
            if (s0 + kk < C.nseg) acc += qv[kk];
    }
    return (float)acc;
}

// ---------------------------------------------------------------------------------------
// Evolve of one column tile of a group.  Every thread of the block calls it: the upper half
// (waves 4-7) sums g_i of the tile's 256 neurons while the lower half loads everything else and
// sums g_e, so that all the loads of the tail are in flight together.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void p_evolve(const PPlan& P, uint32_t gi, uint32_t q, uint32_t n, uint32_t k, uint32_t step,
                                         bool stim_prev, const bn_device_clock& clk, PShared& S) {
    const PGroup& G = P.g[gi];
    const NeuronGroupView& v = G.v;
    const uint32_t t = threadIdx.x & (uint32_t)(kPCols - 1);
    const uint32_t half = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8);
    const uint32_t lid = q * kPCols + t;
    const bool valid = lid < v.num_neurons;
    if (half == 1) {
        if (valid) {
            const float base = ld_c(v.g_inhibitory + lid);
            S.gsum[t] = G.src_i != kPNone ? p_gsum(P.c[G.src_i], lid, base) : base;
        }
        __syncthreads();
        return;
    }
    const uint32_t gid = v.offset + lid;
    const uint32_t bo = pbuf2(n, k);                    // buffers this evolve writes
    const uint32_t bi = k == 0 ? 0u : pbuf2(n, k - 1);  // and the previous step's
    const uint32_t co = pbuf3(n, k);
    float V = 0.0f, istim = 0.0f, re = 0.0f, ri = 0.0f, de = 0.0f, di = 0.0f, ge_base = 0.0f;
    uint32_t refr = 0, spike_prev = 0;
    uint32_t he[4] = {0u, 0u, 0u, 0u}, hi[4] = {0u, 0u, 0u, 0u};
    float tr_r[kMaxTraceSets], tr_d[kMaxTraceSets];
    float pa[kMaxPlast][2], pb[kMaxPlast][2];  // [plast][pre, post]: fast and slow traces
    float ge = 0.0f;
    if (valid) {
        // Every input of the neuron, loads issued together.
        V = ld_c(v.membrane_potentials + lid);
        refr = ld_c(v.refractory_counter + lid);
        ge_base = ld_c(v.g_excitatory + lid);
        istim = ld_c(v.i_stimulus + lid);
        re = ld_c(v.background_excitatory_postsynaptic_r_traces + lid);
        ri = ld_c(v.background_inhibitory_postsynaptic_r_traces + lid);
        de = ld_c(v.background_excitatory_postsynaptic_d_traces + lid);
        di = ld_c(v.background_inhibitory_postsynaptic_d_traces + lid);
        spike_prev = ld_c(P.spikes + gid);
#pragma unroll
        for (uint32_t kk = 0; kk < 4; ++kk) {
            he[kk] = kk < v.poisson_table_excitatory_len ? gp(v.poisson_table_excitatory)[kk] : 0u;
            hi[kk] = kk < v.poisson_table_inhibitory_len ? gp(v.poisson_table_inhibitory)[kk] : 0u;
        }
#pragma unroll
        for (int s = 0; s < kMaxTraceSets; ++s) {
            tr_r[s] = tr_d[s] = 0.0f;
            const PTraceSet& T = P.ts[s];
            if ((uint32_t)s < P.n_ts && gid - T.pre_off < T.n_pre && gid >= T.pre_off) {
                tr_r[s] = ld_c(T.r + (gid - T.pre_off));
                tr_d[s] = ld_c(T.d + (gid - T.pre_off));
            }
        }
#pragma unroll
        for (int s = 0; s < kMaxPlast; ++s) {
            pa[s][0] = pa[s][1] = pb[s][0] = pb[s][1] = 0.0f;
            const PPlast& D = P.pl[s];
            if ((uint32_t)s >= P.n_pl) continue;
            const bool trip = D.kind == kPlastTriplet;
            if (gid >= D.pre_off && gid - D.pre_off < D.n_pre) {
                pa[s][0] = ld_c(D.pre[bi] + (gid - D.pre_off));
                if (trip) pb[s][0] = ld_c(D.tpre[bi] + (gid - D.pre_off));
            }
            if (gid >= D.post_off && gid - D.post_off < D.n_post) {
                pa[s][1] = ld_c(D.post[bi] + (gid - D.post_off));
                if (trip) pb[s][1] = ld_c(D.tpost[bi] + (gid - D.post_off));
            }
        }
        // g_e: the base plus the segment partials of the step's propagate, in order.
        ge = G.src_e != kPNone ? p_gsum(P.c[G.src_e], lid, ge_base) : ge_base;
    }
    __syncthreads();
    if (!valid) return;
    const float gi_ = S.gsum[t];
    // stimulus.cpp:46-60 for this neuron: the previous step's events in list order.
    if (stim_prev && P.ev_steps && step - 1 < P.ev_steps) {
        const uint32_t e1 = min(P.ev_start[step], P.ev_count);
        bool hit = false;
        for (uint32_t e = P.ev_start[step - 1]; e < e1; ++e)
            if (P.ev_ids[e] + P.ev_offsets[e] == gid) {
                istim += P.ev_amps[e];
                hit = true;
            }
        if (hit) st_c(v.i_stimulus + lid, istim);
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
    st_c(v.membrane_potentials + lid, V);
    st_c(v.refractory_counter + lid, refr);
    st_c(P.spikes + gid, spike);
    st_c(v.g_excitatory + lid, bg_ge);
    st_c(v.g_inhibitory + lid, bg_gi);
    st_c(v.background_excitatory_postsynaptic_r_traces + lid, re);
    st_c(v.background_inhibitory_postsynaptic_r_traces + lid, ri);
    st_c(v.background_excitatory_postsynaptic_d_traces + lid, de);
    st_c(v.background_inhibitory_postsynaptic_d_traces + lid, di);
    st_c(v.background_g_excitatory_synaptic_conductances + lid, bg_ge);
    st_c(v.background_g_inhibitory_synaptic_conductances + lid, bg_gi);
    // Plasticity traces: the previous step's +1 then this step's decay.
#pragma unroll
    for (int s = 0; s < kMaxPlast; ++s) {
        const PPlast& D = P.pl[s];
        if ((uint32_t)s >= P.n_pl) continue;
        const bool trip = D.kind == kPlastTriplet;
        if (gid >= D.pre_off && gid - D.pre_off < D.n_pre) {
            const uint32_t i = gid - D.pre_off;
            st_c(D.pre[bo] + i, flush_trace((spike_prev == 1 ? pa[s][0] + 1.0f : pa[s][0]) * D.pre_decay));
            if (trip) st_c(D.tpre[bo] + i, flush_trace((spike_prev == 1 ? pb[s][0] + 1.0f : pb[s][0]) * D.tpre_decay));
        }
        if (gid >= D.post_off && gid - D.post_off < D.n_post) {
            const uint32_t j = gid - D.post_off;
            st_c(D.post[bo] + j, flush_trace((spike_prev == 1 ? pa[s][1] + 1.0f : pa[s][1]) * D.post_decay));
            if (trip) st_c(D.tpost[bo] + j, flush_trace((spike_prev == 1 ? pb[s][1] + 1.0f : pb[s][1]) * D.tpost_decay));
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
            st_c(T.d + i, dd);
            st_c(T.r + i, rr);
            st_c(T.x[bo] + i, dd - rr);
        }
    }
    // Spike bits (cleared two steps ahead) and the monitors' rings (read after the launch).
    if (spike == 1) {
        const uint32_t bit = 1u << (gid & 31u);
        or_c(P.bits[co] + (gid >> 5), bit);
        if (P.spike_ring) or_c(P.spike_ring + (size_t)(step % P.interval) * P.words + (gid >> 5), bit);
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
        const bool spiked = (ld_c(P.bits[cb] + (g >> 5)) >> (g & 31u)) & 1u;
        double R = spiked ? 0.0 : ld_c(D.R + i);
        for (uint32_t q0 = 0; q0 < C.ncol; q0 += kPartBatch) {
            double dv[kPartBatch];
#pragma unroll
            for (uint32_t kk = 0; kk < kPartBatch; ++kk)
                dv[kk] = q0 + kk < C.ncol ? ld_c(D.dR + (size_t)(q0 + kk) * C.n_pre + i) : 0.0;
#pragma unroll
            for (uint32_t kk = 0; kk < kPartBatch; ++kk)
                if (q0 + kk < C.ncol) R += dv[kk];
        }
        st_c(D.R + i, R);
        if (R > 0.0) st_c(D.s + i, (float)((double)gp(D.target)[i] / R));
    }
}

// ---------------------------------------------------------------------------------------
// One item of the streamed kernel: chunk sums of the 8 waves, the plasticity of the step on
// the entries, the segment partial.
// ---------------------------------------------------------------------------------------
template <bool kNT>
__device__ __forceinline__ v4f p_load4(const BN_GAS float* p) {
    if constexpr (kNT) return __builtin_nontemporal_load(reinterpret_cast<const BN_GAS v4f*>(p));
    else return *reinterpret_cast<const BN_GAS v4f*>(p);
}

template <bool kNT>
__device__ __forceinline__ void p_item(const PPlan& P, uint32_t conn, uint32_t col, uint32_t seg, uint32_t n,
                                       uint32_t k, PShared& S) {
    const PConn& C = P.c[conn];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t b = pbuf2(n, k), cb = pbuf3(n, k);
    const PRows R = p_rows(C, col, seg);
    const uint32_t s0 = R.s0, seg_rows = R.seg_rows, c0 = R.c0, c1 = R.c1, j0 = R.j0;
    const bool active = R.active;
    const bool plastic = C.plast != kPNone;
    const PPlast& D = P.pl[plastic ? C.plast : 0];
    const bool lazy = plastic && D.lazy;
    const bool triplet = plastic && D.kind == kPlastTriplet;
    const PRule rule = p_rule(D);
    BN_GAS float* wrow = gp(C.W) + (size_t)(s0 + c0) * C.ld + j0;
    // The first batch of weights is requested before the staging loads.
    v4f w[kPBatch];
    {
        const uint32_t nb = c1 > c0 ? min((uint32_t)kPBatch, c1 - c0) : 0u;
#pragma unroll
        for (uint32_t kk = 0; kk < (uint32_t)kPBatch; ++kk) {
            const uint32_t r = kk < nb ? kk : 0u;
            w[kk] = active && nb ? p_load4<kNT>(wrow + (size_t)r * C.ld) : v4f{0.0f, 0.0f, 0.0f, 0.0f};
        }
    }
    // The segment's rows into LDS: x_i * s_i and, for a plastic matrix, the row's pre trace,
    // LTD factor, scale and spike flag.
    for (uint32_t r = threadIdx.x; r < seg_rows; r += kPThreads) {
        const uint32_t i = s0 + r;
        const float xv = ld_c(C.x[b] + i);
        S.xs[r] = C.s ? xv * ld_c(C.s + i) : xv;
        if (plastic) {
            S.pre[r] = ld_c(D.pre[b] + i);
            S.fac[r] = triplet ? D.p.ltd_A + (D.p.ltd_B * ld_c(D.tpre[b] + i)) : 0.0f;
            S.si[r] = lazy ? ld_c(D.s + i) : 1.0f;
        }
    }
    if (plastic) {
        for (uint32_t wd = threadIdx.x; wd < (seg_rows + 31u) / 32u; wd += kPThreads) {
            const uint32_t g = C.pre_off + s0 + 32u * wd;   // global id of row 32 wd
            const uint32_t sh = g & 31u;
            const uint32_t lo = ld_c(P.bits[cb] + (g >> 5));
            const uint32_t hi = sh && (g >> 5) + 1u < P.words ? ld_c(P.bits[cb] + (g >> 5) + 1u) : 0u;
            S.rbits[wd] = sh ? (lo >> sh) | (hi << (32u - sh)) : lo;
        }
    }
    // Columns of the lane: spike flags of the step, post traces, LTP factors.
    uint32_t colspike = 0;
    float post_j[4] = {0.0f, 0.0f, 0.0f, 0.0f}, fac_ltp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (plastic && active) {
#pragma unroll
        for (uint32_t c4 = 0; c4 < 4; ++c4) {
            const uint32_t j = j0 + c4;
            if (j >= C.n_post) continue;
            const uint32_t g = C.post_off + j;
            if ((ld_c(P.bits[cb] + (g >> 5)) >> (g & 31u)) & 1u) colspike |= 1u << c4;
            post_j[c4] = ld_c(D.post[b] + j);
            if (triplet) fac_ltp[c4] = D.p.ltp_A + (D.p.ltp_B * ld_c(D.tpost[b] + j));
        }
    }
    const bool any_col = plastic && __ballot(colspike != 0u) != 0ull;
    __syncthreads();
    // Spike flags of rows r .. r+7 of the segment (bit kk: row r + kk).
    auto row_bits8 = [&](uint32_t r) -> uint32_t {
        const uint32_t wl = r >> 5, sh = r & 31u;
        const uint32_t lo = S.rbits[wl];
        const uint32_t hi = sh > 24u && wl + 1u < kPSegMax / 32u ? S.rbits[wl + 1u] : 0u;
        return ((sh ? (lo >> sh) | (hi << (32u - sh)) : lo)) & 0xFFu;
    };
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    {
        for (uint32_t r0 = c0; r0 < c1; r0 += kPBatch) {
            const uint32_t nb = min((uint32_t)kPBatch, c1 - r0);
            if (r0 != c0) {
                // Loads of the batch issued together (rows past the chunk re-read its first row
                // and are not used; inactive lanes read nothing).
#pragma unroll
                for (uint32_t kk = 0; kk < (uint32_t)kPBatch; ++kk) {
                    const uint32_t r = kk < nb ? kk : 0u;
                    w[kk] = active ? p_load4<kNT>(wrow + (size_t)r * C.ld) : v4f{0.0f, 0.0f, 0.0f, 0.0f};
                }
            }
#pragma unroll
            for (uint32_t kk = 0; kk < (uint32_t)kPBatch; ++kk) {
                if (kk >= nb) continue;
                const double xd = (double)S.xs[r0 + kk];
                a0 += xd * (double)w[kk].x;
                a1 += xd * (double)w[kk].y;
                a2 += xd * (double)w[kk].z;
                a3 += xd * (double)w[kk].w;
            }
            if (plastic) {
                const uint32_t rowbits = row_bits8(r0);
#pragma unroll
                for (uint32_t kk = 0; kk < (uint32_t)kPBatch; ++kk) {
                    if (kk >= nb) continue;
                    const uint32_t r = r0 + kk;
                    const bool sp = (rowbits >> kk) & 1u;
                    double lsum = 0.0;
                    if (sp || any_col) {
                        v4f nv = w[kk];
                        const uint32_t touched = p_plast_row(rule, lazy, triplet, sp, s0 + r, S.si[r], S.pre[r], S.fac[r], j0,
                                                             active, C.n_post, colspike, post_j, fac_ltp, nv, lsum);
                        p_store_row(wrow + (size_t)kk * C.ld, nv, touched);
                    }
                    if (lazy && lane == 0) S.dr[r] = lsum;
                }
            }
            wrow += (size_t)kPBatch * C.ld;
        }
    }
    // Segment partial: the chunk sums in chunk order from 0.0; the lazy matrix's row tile sums.
    S.red[wave][4 * lane + 0] = a0;
    S.red[wave][4 * lane + 1] = a1;
    S.red[wave][4 * lane + 2] = a2;
    S.red[wave][4 * lane + 3] = a3;
    __syncthreads();
    if (threadIdx.x < (uint32_t)kPCols) {
        const uint32_t j = col * kPCols + threadIdx.x;
        const uint32_t chunks = (seg_rows + C.rb - 1) / C.rb;
        if (j < C.n_post) {
            double acc = 0.0;
            for (uint32_t w8 = 0; w8 < chunks; ++w8) acc += S.red[w8][threadIdx.x];
            st_c(C.partials + (size_t)seg * C.ld + j, acc);
        }
    }
    if (lazy)
        for (uint32_t r = threadIdx.x; r < seg_rows; r += kPThreads) st_c(D.dR + (size_t)col * C.n_pre + s0 + r, S.dr[r]);
}

// ---------------------------------------------------------------------------------------
// The launch
// ---------------------------------------------------------------------------------------
template <bool kNT>
__global__ __launch_bounds__(kPThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_persist(const PPlan* __restrict__ plan, uint32_t n) {
    const PPlan& P = *plan;
    __shared__ PShared S;
    const bn_device_clock clk = *P.clock;
    const uint32_t t0 = clk.clock_step;
    const uint32_t per_step = P.n_coltiles + P.n_rowtails;
    const uint32_t blk = blockIdx.x;
    const uint32_t ib = P.block_begin[blk], ie = P.block_begin[blk + 1];
    // Phase 0: the evolve of the launch's first step, one column tile per block.
    for (uint32_t ct = blk; ct < P.n_coltiles; ct += gridDim.x) {
        uint32_t gi = 0;
        while (gi + 1 < P.n_groups && ct >= P.g[gi + 1].col_base) ++gi;
        p_evolve(P, gi, ct - P.g[gi].col_base, n, 0, t0, false, clk, S);
        (void)p_block_arrive(P.tails_done, S);
    }
    for (uint32_t k = 0; k < n; ++k) {
        // Every tail of the previous step (the evolve of step k included) has finished.
        if (!p_block_wait(P, P.n_coltiles + k * per_step, S)) break;
        if (blk == 0 && k + 2 < n) {
            // Spike bits of evolve k + 2 and its ring slot, cleared ahead (their last readers
            // are the items of step k - 1).
            uint32_t* bits = P.bits[pbuf3(n, k + 2)];
            for (uint32_t w = threadIdx.x; w < P.words; w += kPThreads) st_c(bits + w, 0u);
            if (P.spike_ring) {
                uint32_t* row = P.spike_ring + (size_t)((t0 + k + 2) % P.interval) * P.words;
                for (uint32_t w = threadIdx.x; w < P.words; w += kPThreads) st_c(row + w, 0u);
            }
        }
        for (uint32_t e = ib; e < ie; ++e) {
            const PItem it = P.items[e];
            const uint32_t conn = __builtin_amdgcn_readfirstlane(it.conn);
            const uint32_t col = __builtin_amdgcn_readfirstlane(it.col);
            const uint32_t seg = __builtin_amdgcn_readfirstlane(it.seg);
            p_item<kNT>(P, conn, col, seg, n, k, S);
            const PConn& C = P.c[conn];
            const PGroup& G = P.g[C.post_group];
            const uint32_t ct = G.col_base + col;
            const uint32_t old_c = p_block_arrive(P.col_count + ct, S);
            const bool last_col = old_c == (k + 1) * G.expect - 1u;
            bool last_row = false;
            if (P.lazy_pl != kPNone && C.plast == P.lazy_pl) {
                const uint32_t old_r = p_block_arrive(P.row_count + seg, S);
                last_row = old_r == (k + 1) * C.ncol - 1u;
            }
            if (last_col) {
                if (k + 1 < n) p_evolve(P, C.post_group, col, n, k + 1, t0 + k + 1, true, clk, S);
                (void)p_block_arrive(P.tails_done, S);
            }
            if (last_row) {
                p_row_tail(P, seg, pbuf3(n, k));
                (void)p_block_arrive(P.tails_done, S);
            }
        }
    }
    // The last step's stimulus events and the clock advance of the whole launch, once every tail
    // of the launch has finished (so every block has read the clock and no evolve is pending).
    if (blk == 0) {
        if (!p_block_wait(P, P.n_coltiles + n * per_step, S)) return;
        if (threadIdx.x == 0) {
            const uint32_t last = t0 + n - 1;
            if (P.ev_steps && last < P.ev_steps) {
                const uint32_t e1 = min(P.ev_start[last + 1], P.ev_count);
                for (uint32_t e = P.ev_start[last]; e < e1; ++e) {
                    float* dst = P.i_stimulus + P.ev_ids[e] + P.ev_offsets[e];
                    st_c(dst, ld_c(dst) + P.ev_amps[e]);
                }
            }
            bn_device_clock c = clk;
            c.clock_step = t0 + n;
            c.next_event_idx = P.ev_steps ? P.ev_start[min(t0 + n, P.ev_steps)] : clk.next_event_idx;
            *P.clock_rw = c;
        }
    }
}

static const void* persist_kernel(uint32_t nontemporal) {
    return nontemporal ? reinterpret_cast<const void*>(k_persist<true>) : reinterpret_cast<const void*>(k_persist<false>);
}

int persistent_max_blocks(uint32_t resident_rows, uint32_t* blocks) {
    if (resident_rows) return persistent_ll_max_blocks(resident_rows, blocks);
    int dev = 0, cus = 0, per_cu = 0, per_cu_nt = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel(0), kPThreads, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_nt, persist_kernel(1), kPThreads, 0);
    *blocks = e == hipSuccess ? (uint32_t)std::max(0, cus * std::min(per_cu, per_cu_nt)) : 0u;
    return (int)e;
}

int launch_persistent(const PPlan* d_plan, const PPlan& hp, uint32_t n, uint32_t nontemporal, bn_stream_t s) {
    if (n == 0 || hp.blocks == 0) return 0;
    if (hp.resident_rows) return launch_persistent_ll(d_plan, hp, n, s);
    void* args[] = {(void*)&d_plan, (void*)&n};
    return (int)hipLaunchCooperativeKernel(persist_kernel(nontemporal), dim3(hp.blocks), dim3(kPThreads), args, 0, s);
}

}  // namespace bn
