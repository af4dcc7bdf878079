// The persistent step with resident weights (bn_persist.h): one cooperative launch runs n
// steps; each block owns one item (connection, 256-column tile, segment of 8 chunks of at most
// 24 rows) whose weight entries stay in the registers of their threads for the whole launch.
// Data passes between blocks as tagged words: a 32-bit value and the step that produced it in
// one 64-bit agent-coherent store, so a reader polls the data it needs (no counter, and a
// writer never waits for its stores to complete).  Per step there are two hand-offs:
//   records   the block that evolves a column tile (its evolver: the block of the item
//             (source connection, that tile, segment 0)) writes, per neuron, what the items
//             read: x of each trace set, the plasticity traces, and the spike flags packed 32 to
//             a word per tile;
//   partials  each item writes its segment partials, and for the lazily normalised matrix its
//             rows' tile sums; the evolver of the column tile polls the partials, and the blocks
//             of the same segment poll the tile sums at the next step to update their own copy
//             of the rows' sums and scales.
// An evolver keeps the state of its neurons (membrane, background, traces) in memory that only
// it reads and writes during the launch.  Tagged buffers have three slots by step; a block
// rewrites a slot only after every block has finished the step that last read it (a count of
// finished steps, raised without waiting).  Every poll is bounded: past the limit a block raises
// the state's error word (2) and leaves, and every other block leaves at its next poll.
//
// The arithmetic is the persistent step's: chunk sums in FP64 from 0.0 in row order, segment
// partials as the chunk sums in chunk order, g = fl32(g + partials in segment order), the
// plasticity per entry as LTP then LTD (kernel_triplet_plasticity_update.cu:54-93,
// kernel_istdp_plasticity_update.cu:42-80), lazy row sums in column-tile order, the evolve of
// kernel_neuron_group_evolve.cu:11-130 and the trace steps.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bn_device.h"
#include "bn_errors.h"
#include "bn_neuron.h"
#include "bn_persist.h"
#include "bn_persist_dev.h"

namespace bn {

__device__ __forceinline__ uint32_t l_tag(uint32_t a) { return a + 1u; }  // tag of step a (0: never written)
__device__ __forceinline__ uint32_t l_slot(uint32_t a) { return a % 3u; }
__device__ __forceinline__ void l_put(uint64_t* p, uint32_t a, uint32_t bits) {
    st_c(p, ((uint64_t)l_tag(a) << 32) | bits);
}
// A double as two words: the low half at p, the high half `stride` words further (so the
// stores of a wave's lanes to consecutive p fill whole lines).
__device__ __forceinline__ void l_put_d(uint64_t* p, uint32_t stride, uint32_t a, double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    l_put(p, a, (uint32_t)b);
    l_put(p + stride, a, (uint32_t)(b >> 32));
}
__device__ __forceinline__ double l_double(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Evolver state in LDS, per neuron of the tile: membrane, refractory count (bits), stimulus
// current, background traces, background conductances (the next evolve's g bases), spike flag
// (bits), then r, d of each trace set, then pre, slow pre, post, slow post of each plasticity,
// then the neuron's slot in the membrane-potential monitor (bits; -1: none).
enum : uint32_t {
    kEvV, kEvRefr, kEvIstim, kEvRe, kEvRi, kEvDe, kEvDi, kEvGe, kEvGi, kEvSpike,
    kEvTr, kEvPl = kEvTr + 2 * kMaxTraceSets, kEvVrec = kEvPl + 4 * kMaxPlast, kEvVars
};
constexpr uint32_t kLPollPart = 9;    // segment partials an evolve polls per thread and round

struct LShared {
    // Small arrays first: their addresses stay within the 64 KB a DS instruction's offset reaches
    // from one base register.
    float ev[kEvVars][kPCols];              // evolver: its neurons' state (see l_state)
    float gsum[kPCols];
    uint32_t heads[8];                      // evolver: the first 4 entries of its group's Poisson tables (e, i)
    float xs[kLSegMax], pre[kLSegMax], fac[kLSegMax];
    float post[kPCols], tpost[kPCols];      // post traces of the column tile
    float s[kLSegMax];                      // lazy matrix: the segment's row scales (kept across steps)
    float target[kLSegMax];                 // and normalisation targets
    double xd[kLSegMax];                    // x_i * s_i (x_i without row scales) in FP64
    double dr[kLSegMax];                    // this item's row tile sums
    double R[kLSegMax];                     // lazy matrix: the segment's row sums (kept across steps)
    uint32_t rbits[kLSegMax / 32], rbits_prev[kLSegMax / 32];
    uint32_t rwords[kLSegMax / 32 + 2];     // spike words (32 neurons of the pre group) covering the rows
    uint32_t cbits[8];                      // spike words of the column tile
    uint32_t abort;
    double red[kPWaves][kPCols];            // chunk sums
    union {
        double drin[kLDrJobs * kPThreads];  // staging: the previous step's row tile sums [tile][row]
        double pv[kLPartMax][kPCols];       // evolve: the tile's segment partials [segment][neuron]
        v4f pbuf[kPWaves][8 * 64];          // plasticity: 8 resident rows of each wave
    };
};

// Polls this thread's words (bit i of `want`: word addr(i), tag tag_lo for i < 3 else tag_hi)
// until each carries its tag, handing the low half of each to sink(i, bits).  Every thread of
// the block calls it; false for every thread when a thread gave up (past the wait limit) or the
// launch is being abandoned.
template <int kM, typename Addr, typename Sink>
__device__ __forceinline__ bool l_poll(const PPlan& P, LShared& S, uint32_t want, uint32_t tag_lo, uint32_t tag_hi,
                                       Addr addr, Sink sink) {
    uint32_t pending = want;
    for (uint32_t it = 0; pending; ++it) {
        uint64_t v[kM];
#pragma unroll
        for (int i = 0; i < kM; ++i) v[i] = (pending >> i) & 1u ? ld_c(addr(i)) : 0ull;
#pragma unroll
        for (int i = 0; i < kM; ++i)
            if (((pending >> i) & 1u) && (uint32_t)(v[i] >> 32) == (i < 3 ? tag_lo : tag_hi)) {
                sink(i, (uint32_t)v[i]);
                pending &= ~(1u << i);
            }
        if (!pending) break;
        if (it >= P.wait_limit) {
            max_c(P.status, 2u);
            S.abort = 1u;
            break;
        }
        if ((it & 15u) == 15u && ld_c(P.status) >= 2u) {
            S.abort = 1u;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __syncthreads();
    return S.abort == 0u;
}

// Thread 0: *cnt >= target (`seen`: a value of *cnt loaded ahead).  Every thread calls it.
__device__ __forceinline__ bool l_wait(const PPlan& P, LShared& S, const uint32_t* cnt, uint32_t target, uint32_t seen) {
    if (threadIdx.x == 0 && seen < target) {
        for (uint32_t it = 0;; ++it) {
            if (ld_c(cnt) >= target) break;
            if (it >= P.wait_limit) {
                max_c(P.status, 2u);
                S.abort = 1u;
                break;
            }
            if ((it & 15u) == 15u && ld_c(P.status) >= 2u) {
                S.abort = 1u;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    return S.abort == 0u;
}

// ---------------------------------------------------------------------------------------
// The evolver's neurons: their state lives in LDS for the whole launch (read from the state's
// buffers before the first evolve, written back after the last).
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void l_state(const PPlan& P, LShared& S, uint32_t gi, uint32_t q, bool store) {
    const NeuronGroupView& v = P.g[gi].v;
    const uint32_t t = threadIdx.x;
    const uint32_t lid = q * kPCols + t;
    if (t >= (uint32_t)kPCols || lid >= v.num_neurons) return;
    const uint32_t gid = v.offset + lid;
    float* const g[] = {v.membrane_potentials, reinterpret_cast<float*>(v.refractory_counter), v.i_stimulus,
                        v.background_excitatory_postsynaptic_r_traces, v.background_inhibitory_postsynaptic_r_traces,
                        v.background_excitatory_postsynaptic_d_traces, v.background_inhibitory_postsynaptic_d_traces,
                        v.g_excitatory, v.g_inhibitory, reinterpret_cast<float*>(P.spikes + v.offset)};
#pragma unroll
    for (uint32_t k = 0; k < kEvTr; ++k) {
        if (store) gp(g[k])[lid] = S.ev[k][t];
        else S.ev[k][t] = gp(g[k])[lid];
    }
    if (store) {
        gp(v.background_g_excitatory_synaptic_conductances)[lid] = S.ev[kEvGe][t];
        gp(v.background_g_inhibitory_synaptic_conductances)[lid] = S.ev[kEvGi][t];
    } else {
        S.ev[kEvVrec][t] = __uint_as_float(P.v_ring ? (uint32_t)P.vrec_index[gid] : 0xFFFFFFFFu);
        if (t < 4) {
            S.heads[t] = t < v.poisson_table_excitatory_len ? gp(v.poisson_table_excitatory)[t] : 0u;
            S.heads[4 + t] = t < v.poisson_table_inhibitory_len ? gp(v.poisson_table_inhibitory)[t] : 0u;
        }
    }
#pragma unroll
    for (int s = 0; s < kMaxTraceSets; ++s) {
        const PTraceSet& T = P.ts[s];
        if ((uint32_t)s >= P.n_ts || gid < T.pre_off || gid - T.pre_off >= T.n_pre) continue;
        const uint32_t i = gid - T.pre_off;
        float& r = S.ev[kEvTr + 2 * s][t];
        float& d = S.ev[kEvTr + 2 * s + 1][t];
        if (store) {
            gp(T.r)[i] = r;
            gp(T.d)[i] = d;
            gp(T.x[0])[i] = d - r;
        } else {
            r = gp(T.r)[i];
            d = gp(T.d)[i];
        }
    }
#pragma unroll
    for (int s = 0; s < kMaxPlast; ++s) {
        const PPlast& D = P.pl[s];
        if ((uint32_t)s >= P.n_pl) continue;
        const bool trip = D.kind == kPlastTriplet;
        float* const w[] = {D.pre[0], D.tpre[0], D.post[0], D.tpost[0]};
        const bool in[] = {gid >= D.pre_off && gid - D.pre_off < D.n_pre, gid >= D.post_off && gid - D.post_off < D.n_post};
        const uint32_t idx[] = {gid - D.pre_off, gid - D.post_off};
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            if (!in[u >> 1] || ((u & 1u) && !trip)) continue;
            float& x = S.ev[kEvPl + 4 * s + u][t];
            if (store) gp(w[u])[idx[u >> 1]] = x;
            else x = gp(w[u])[idx[u >> 1]];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Evolve of column tile q of group gi for step a (launch step ke = a - t0).  Every thread calls
// it: all 512 threads poll the tile's segment partials of step a - 1 (interleaved over the
// segments), the upper half sums g_i and the lower half sums g_e and runs the neurons.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ bool l_evolve(const PPlan& P, LShared& S, uint32_t gi, uint32_t q, uint32_t a, uint32_t ke,
                                         uint32_t n, const bn_device_clock& clk, uint32_t done_seen) {
    const PGroup& G = P.g[gi];
    const NeuronGroupView& v = G.v;
    const uint32_t t = threadIdx.x & (uint32_t)(kPCols - 1);
    const uint32_t half = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8);
    const uint32_t lid = q * kPCols + t;
    const bool valid = lid < v.num_neurons;
    const uint32_t gid = v.offset + lid;
    // The tile's segment partials of step a - 1: sources e then i, segment order.
    const uint32_t ne = G.src_e != kPNone ? P.c[G.src_e].nseg : 0u;
    const uint32_t ni = G.src_i != kPNone ? P.c[G.src_i].nseg : 0u;
    const uint32_t nall = ne + ni;
    if (ke == 0) {
        // The launch's first evolve: the pending partials are in the state's buffers.
        for (uint32_t s = half; s < nall; s += 2) {
            const PConn& C = P.c[s < ne ? G.src_e : G.src_i];
            const uint32_t sg = s < ne ? s : s - ne;
            S.pv[s][t] = valid ? gp(C.partials)[(size_t)sg * C.ld + lid] : 0.0;
        }
        __syncthreads();
    } else {
        const uint32_t ap = a - 1u, sl = l_slot(ap);
        uint32_t* pv32 = reinterpret_cast<uint32_t*>(&S.pv[0][0]);
        for (uint32_t r0 = 0; r0 < nall; r0 += 2u * kLPollPart) {
            // Round: segments r0 + half + 2u, u < kLPollPart (words 2u, 2u + 1: low, high).
            uint32_t want = 0;
            for (uint32_t u = 0; u < kLPollPart; ++u)
                if (valid && r0 + half + 2u * u < nall) want |= 3u << (2u * u);
            auto addr = [&](int i) -> const uint64_t* {
                const uint32_t sg_all = r0 + half + 2u * (uint32_t)(i >> 1);
                const uint32_t src = sg_all < ne ? G.src_e : G.src_i;
                const uint32_t sg = sg_all < ne ? sg_all : sg_all - ne;
                return P.ll_part[src] + (((size_t)sl * P.c[src].nseg + sg) * 2u + (uint32_t)(i & 1)) * P.c[src].ld + lid;
            };
            auto sink = [&](int i, uint32_t bits) {
                pv32[((r0 + half + 2u * (uint32_t)(i >> 1)) * kPCols + t) * 2u + (uint32_t)(i & 1)] = bits;
            };
            if (!l_poll<(int)(2 * kLPollPart)>(P, S, want, l_tag(ap), l_tag(ap), addr, sink)) return false;
        }
    }
    // g_e (lower half) and g_i (upper half): the base (the previous evolve's background
    // conductance) plus the partials in segment order.
    float g = 0.0f;
    if (valid) {
        const float base = S.ev[half == 0 ? kEvGe : kEvGi][t];
        g = base;
        const uint32_t src = half == 0 ? G.src_e : G.src_i;
        if (src != kPNone) {
            double acc = (double)base;
            const uint32_t s0 = half == 0 ? 0u : ne, s1 = half == 0 ? ne : nall;
            for (uint32_t s = s0; s < s1; ++s) acc += S.pv[s][t];
            g = (float)acc;
        }
        if (half == 1) S.gsum[t] = g;
    }
    // Slot reuse: every block has finished launch step ke - 3 (its items read the slot's records).
    if (!l_wait(P, S, P.done, ke >= 3 ? (ke - 2u) * P.blocks : 0u, done_seen)) return false;
    const uint32_t sl = l_slot(a);
    if (half == 0) {
    uint32_t spike = 0;
    if (valid) {
        const float ge = g, gi_ = S.gsum[t];
        float V = S.ev[kEvV][t], istim = S.ev[kEvIstim][t];
        float re = S.ev[kEvRe][t], ri = S.ev[kEvRi][t], de = S.ev[kEvDe][t], di = S.ev[kEvDi][t];
        uint32_t refr = __float_as_uint(S.ev[kEvRefr][t]);
        const uint32_t spike_prev = __float_as_uint(S.ev[kEvSpike][t]);
        // stimulus.cpp:46-60 for this neuron: the previous step's events in list order.
        if (ke > 0 && P.ev_steps && a - 1 < P.ev_steps) {
            const uint32_t e1 = min(P.ev_start[a], P.ev_count);
            for (uint32_t e = P.ev_start[a - 1]; e < e1; ++e)
                if (P.ev_ids[e] + P.ev_offsets[e] == gid) istim += P.ev_amps[e];
            S.ev[kEvIstim][t] = istim;
        }
        // Background (kernel_neuron_group_evolve.cu:11-83).
        uint32_t he[4], hi[4];
#pragma unroll
        for (uint32_t kk = 0; kk < 4; ++kk) {
            he[kk] = S.heads[kk];
            hi[kk] = S.heads[4 + kk];
        }
        re = flush_trace(re * v.background_excitatory_postsynaptic_r_trace_decay);
        ri = flush_trace(ri * v.background_inhibitory_postsynaptic_r_trace_decay);
        de = flush_trace(de * v.background_excitatory_postsynaptic_d_trace_decay);
        di = flush_trace(di * v.background_inhibitory_postsynaptic_d_trace_decay);
        uint32_t ke_ = 0, ki_ = 0;
        if (v.poisson_table_excitatory_len > 0 || v.poisson_table_inhibitory_len > 0) {
            const U32x4 r = background_draws(clk, a, gid, v.rng_instance);
            if (v.mean_background_excitatory_spikes > 0.0f)
                ke_ = poisson_lookup(r.x, v.poisson_table_excitatory, v.poisson_table_excitatory_len, he);
            if (v.mean_background_inhibitory_spikes > 0.0f)
                ki_ = poisson_lookup(r.y, v.poisson_table_inhibitory, v.poisson_table_inhibitory_len, hi);
        }
        if (ke_ > 0u) {
            const float inc = (float)ke_;
            re += inc;
            de += inc;
        }
        const float bg_ge = (de - re) * v.background_excitatory_scaling_factor;
        if (ki_ > 0u) {
            const float inc = (float)ki_;
            ri += inc;
            di += inc;
        }
        const float bg_gi = (di - ri) * v.background_inhibitory_scaling_factor;
        // EIF (kernel_neuron_group_evolve.cu:85-130).
        eif_step(v, ge, gi_, istim, V, refr, spike);
        S.ev[kEvV][t] = V;
        S.ev[kEvRefr][t] = __uint_as_float(refr);
        S.ev[kEvSpike][t] = __uint_as_float(spike);
        S.ev[kEvGe][t] = bg_ge;
        S.ev[kEvGi][t] = bg_gi;
        S.ev[kEvRe][t] = re;
        S.ev[kEvRi][t] = ri;
        S.ev[kEvDe][t] = de;
        S.ev[kEvDi][t] = di;
        // Plasticity traces: the previous step's +1 then this step's decay; the records.
#pragma unroll
        for (int s = 0; s < kMaxPlast; ++s) {
            const PPlast& D = P.pl[s];
            if ((uint32_t)s >= P.n_pl) continue;
            const bool trip = D.kind == kPlastTriplet;
            const bool in[] = {gid >= D.pre_off && gid - D.pre_off < D.n_pre, gid >= D.post_off && gid - D.post_off < D.n_post};
            const uint32_t idx[] = {gid - D.pre_off, gid - D.post_off};
            const uint32_t len[] = {D.n_pre, D.n_post};
            const float dec[] = {D.pre_decay, D.tpre_decay, D.post_decay, D.tpost_decay};
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                if (!in[u >> 1] || ((u & 1u) && !trip)) continue;
                float& x = S.ev[kEvPl + 4 * s + u][t];
                x = flush_trace((spike_prev == 1 ? x + 1.0f : x) * dec[u]);
                l_put(P.ll_pl[s][u] + (size_t)sl * len[u >> 1] + idx[u >> 1], a, __float_as_uint(x));
            }
        }
        // Connection traces (kernel_connection_propagate.cu:8-46) and their x records.
#pragma unroll
        for (int s = 0; s < kMaxTraceSets; ++s) {
            const PTraceSet& T = P.ts[s];
            if ((uint32_t)s < P.n_ts && gid >= T.pre_off && gid - T.pre_off < T.n_pre) {
                const uint32_t i = gid - T.pre_off;
                float& r = S.ev[kEvTr + 2 * s][t];
                float& d = S.ev[kEvTr + 2 * s + 1][t];
                float dd = flush_trace(d * T.decay_d);
                float rr = flush_trace(r * T.decay_r);
                if (spike == 1) {
                    rr += T.K;
                    dd += T.K;
                }
                d = dd;
                r = rr;
                l_put(P.ll_x[s] + (size_t)sl * T.n_pre + i, a, __float_as_uint(dd - rr));
            }
        }
        // Monitors' rings (slots cleared before the launch) and, at the launch's last evolve,
        // the state's spike bits (cleared before the launch).
        if (spike == 1) {
            const uint32_t bit = 1u << (gid & 31u);
            if (P.spike_ring) or_c(P.spike_ring + (size_t)(a % P.interval) * P.words + (gid >> 5), bit);
            if (ke + 1 == n) or_c(P.bits[0] + (gid >> 5), bit);
        }
        if (P.v_ring) {
            const int32_t slot = (int32_t)__float_as_uint(S.ev[kEvVrec][t]);
            if (slot >= 0) P.v_ring[(size_t)(a % P.interval) * P.n_vrec + (uint32_t)slot] = V;
        }
    }
    // The tile's spike words: 32 neurons each, two per wave.
    const uint64_t m = __ballot(spike == 1);
    const uint32_t lane = threadIdx.x & 63u;
    if (lane == 0 || lane == 32) {
        const uint32_t word = (threadIdx.x >> 5);  // 0..7 within the tile
        l_put(P.ll_spk + ((size_t)sl * P.n_coltiles + G.col_base + q) * 8u + word, a,
              lane == 0 ? (uint32_t)m : (uint32_t)(m >> 32));
    }
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// The block's item at step a (launch step k), first part: poll what it reads, update the lazy
// rows' sums and scales, the chunk sums from the resident rows, the segment partials.
// ---------------------------------------------------------------------------------------
template <int kRes>
__device__ __forceinline__ bool l_item_front(const PPlan& P, LShared& S, uint32_t conn, uint32_t col, uint32_t seg,
                                             uint32_t a, uint32_t k, uint32_t n, const v4f (&w)[kRes]) {
    const PConn& C = P.c[conn];
    const PRows Rw = p_rows(C, col, seg);
    const uint32_t s0 = Rw.s0, seg_rows = Rw.seg_rows, c0 = Rw.c0, c1 = Rw.c1;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool plastic = C.plast != kPNone;
    const uint32_t pl = plastic ? C.plast : 0u;
    const PPlast& D = P.pl[pl];
    const bool lazy = plastic && D.lazy;
    const bool triplet = plastic && D.kind == kPlastTriplet;
    const PGroup& Gpre = P.g[C.pre_group];
    const PGroup& Gpost = P.g[C.post_group];
    const PTraceSet& T = P.ts[C.ts];
    const uint32_t sl = l_slot(a);
    const size_t spk_base = (size_t)sl * P.n_coltiles * 8u;
    const uint32_t gw0 = s0 >> 5;                                  // first spike word of the rows
    const uint32_t nw = ((s0 + seg_rows - 1u) >> 5) - gw0 + 1u;    // words covering the rows
    const bool upper = tid >= (uint32_t)kPCols;
    const uint32_t c = tid - (uint32_t)kPCols;                     // upper half: column of the tile
    const uint32_t jc = col * kPCols + c;
    // Tile sums of the previous step: jobs d = tid + 512 u over (tile, row).
    const bool with_dr = lazy && k > 0;
    const uint32_t ndr = with_dr ? C.ncol * seg_rows : 0u;
    uint32_t want = 0;
    if (!upper) {
        if (tid < seg_rows) want |= plastic ? (triplet ? 7u : 3u) : 1u;
    } else {
        if (plastic && jc < C.n_post) want |= triplet ? 3u : 1u;
        if (plastic && (c < 8u || c - 8u < nw)) want |= 4u;
    }
    for (uint32_t u = 0; u < kLDrJobs; ++u)
        if (tid + kPThreads * u < ndr) want |= 3u << (3u + 2u * u);
    const uint32_t ap = a - 1u, slp = l_slot(ap);
    auto addr = [&](int i) -> const uint64_t* {
        if (i >= 3) {
            const uint32_t d = tid + kPThreads * (uint32_t)((i - 3) >> 1);
            const uint32_t qd = d / seg_rows, r = d - qd * seg_rows;
            return P.ll_dr + (((size_t)slp * C.ncol + qd) * 2u + (uint32_t)((i - 3) & 1)) * C.n_pre + s0 + r;
        }
        if (!upper) {
            if (i == 0) return P.ll_x[C.ts] + (size_t)sl * T.n_pre + s0 + tid;
            return P.ll_pl[pl][i == 1 ? 0 : 1] + (size_t)sl * D.n_pre + s0 + tid;
        }
        if (i < 2) return P.ll_pl[pl][i == 0 ? 2 : 3] + (size_t)sl * D.n_post + jc;
        return c < 8u ? P.ll_spk + spk_base + (size_t)(Gpost.col_base + col) * 8u + c
                      : P.ll_spk + spk_base + (size_t)Gpre.col_base * 8u + gw0 + (c - 8u);
    };
    // Thread 0 also reads the count of finished steps now, for the slot check before the writes.
    const uint32_t seen = tid == 0 ? ld_c(P.done) : 0u;
    uint32_t* drin32 = reinterpret_cast<uint32_t*>(S.drin);
    auto sink = [&](int i, uint32_t bits) {
        if (i >= 3) {
            drin32[(tid + kPThreads * (uint32_t)((i - 3) >> 1)) * 2u + (uint32_t)((i - 3) & 1)] = bits;
        } else if (!upper) {
            if (i == 0) S.xs[tid] = __uint_as_float(bits);
            else if (i == 1) S.pre[tid] = __uint_as_float(bits);
            else S.fac[tid] = D.p.ltd_A + (D.p.ltd_B * __uint_as_float(bits));
        } else {
            if (i == 0) S.post[c] = __uint_as_float(bits);
            else if (i == 1) S.tpost[c] = __uint_as_float(bits);
            else if (c < 8u) S.cbits[c] = bits;
            else S.rwords[c - 8u] = bits;
        }
    };
    if (!l_poll<(int)(3 + 2 * kLDrJobs)>(P, S, want, l_tag(a), l_tag(ap), addr, sink)) return false;
    if (!upper && tid < seg_rows && plastic && !triplet) S.fac[tid] = 0.0f;
    __syncthreads();
    // Row spike flags of the segment; the lazy rows' sums and scales after the previous step's
    // plasticity (R_i = 0 or R_i, plus the tiles' sums in tile order; s_i = target_i / R_i);
    // x_i * s_i.
    if (plastic && tid < (seg_rows + 31u) / 32u) {
        const uint32_t sh = s0 & 31u;
        const uint32_t lo = S.rwords[tid];
        const uint32_t hi = sh && tid + 1u < nw ? S.rwords[tid + 1u] : 0u;
        S.rbits[tid] = sh ? (lo >> sh) | (hi << (32u - sh)) : lo;
    }
    if (tid < seg_rows) {
        if (with_dr) {
            const bool spiked = (S.rbits_prev[tid >> 5] >> (tid & 31u)) & 1u;
            double R = spiked ? 0.0 : S.R[tid];
            for (uint32_t q0 = 0; q0 < C.ncol; q0 += 8) {
                double t8[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u) t8[u] = q0 + u < C.ncol ? S.drin[(q0 + u) * seg_rows + tid] : 0.0;
#pragma unroll
                for (uint32_t u = 0; u < 8; ++u)
                    if (q0 + u < C.ncol) R += t8[u];
            }
            S.R[tid] = R;
            if (R > 0.0) S.s[tid] = (float)((double)S.target[tid] / R);
        }
        S.xd[tid] = (double)(lazy ? S.xs[tid] * S.s[tid] : S.xs[tid]);
    }
    __syncthreads();
    // Chunk sum over the resident rows.
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    // (The FP64 product of two FP32 values is exact, so fma(x, w, acc) rounds as acc + x * w.)
#pragma unroll
    for (uint32_t k8 = 0; k8 < (uint32_t)kRes; k8 += 8) {
        double xb[8];
#pragma unroll
        for (uint32_t kk = 0; kk < 8; ++kk) xb[kk] = c0 + k8 + kk < c1 ? S.xd[c0 + k8 + kk] : 0.0;
#pragma unroll
        for (uint32_t kk = 0; kk < 8; ++kk) {
            if (c0 + k8 + kk >= c1) continue;
            a0 = fma(xb[kk], (double)w[k8 + kk].x, a0);
            a1 = fma(xb[kk], (double)w[k8 + kk].y, a1);
            a2 = fma(xb[kk], (double)w[k8 + kk].z, a2);
            a3 = fma(xb[kk], (double)w[k8 + kk].w, a3);
        }
    }
    S.red[wave][4 * lane + 0] = a0;
    S.red[wave][4 * lane + 1] = a1;
    S.red[wave][4 * lane + 2] = a2;
    S.red[wave][4 * lane + 3] = a3;
    // Slot reuse: every block has finished launch step k - 2 (the slot's last readers).  The
    // last step's partials go to the state's buffers, which the launch's first evolves read.
    if (!l_wait(P, S, P.done, k >= 3 ? (k - 1u) * P.blocks : 0u, seen)) return false;
    if (k + 1 == n && !l_wait(P, S, P.done + 1, P.n_coltiles, 0u)) return false;
    // Segment partial: the chunk sums in chunk order from 0.0 (the last step's go to the state's
    // buffers).
    if (tid < (uint32_t)kPCols) {
        const uint32_t j = col * kPCols + tid;
        const uint32_t chunks = (seg_rows + C.rb - 1) / C.rb;
        if (j < C.n_post) {
            double acc = 0.0;
            for (uint32_t w8 = 0; w8 < chunks; ++w8) acc += S.red[w8][tid];
            if (k + 1 < n) l_put_d(P.ll_part[conn] + ((size_t)sl * C.nseg + seg) * 2u * C.ld + j, C.ld, a, acc);
            else gp(C.partials)[(size_t)seg * C.ld + j] = acc;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// Second part, off the step's critical path: the step's plasticity on the resident rows (after
// the partials, which used the old entries), the lazy matrix's row tile sums.  The rows that
// need it go through LDS 8 at a time, so one copy of the row update serves every row.
// ---------------------------------------------------------------------------------------
template <int kRes>
__device__ __forceinline__ void l_item_back(const PPlan& P, LShared& S, uint32_t conn, uint32_t col, uint32_t seg,
                                            uint32_t a, v4f (&w)[kRes], uint32_t& dirty) {
    const PConn& C = P.c[conn];
    const PRows Rw = p_rows(C, col, seg);
    const uint32_t s0 = Rw.s0, seg_rows = Rw.seg_rows, c0 = Rw.c0, c1 = Rw.c1, j0 = Rw.j0;
    const bool active = Rw.active;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool plastic = C.plast != kPNone;
    if (!plastic) return;
    const PPlast& D = P.pl[C.plast];
    const bool lazy = D.lazy;
    const bool triplet = D.kind == kPlastTriplet;
    const PRule rule = p_rule(D);
    // Columns of the lane: spike flags of the step, post traces, LTP factors.
    uint32_t colspike = 0;
    float post_j[4] = {0.0f, 0.0f, 0.0f, 0.0f}, fac_ltp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (active) {
#pragma unroll
        for (uint32_t c4 = 0; c4 < 4; ++c4) {
            const uint32_t l = 4u * lane + c4;
            if (j0 + c4 >= C.n_post) continue;
            if ((S.cbits[l >> 5] >> (l & 31u)) & 1u) colspike |= 1u << c4;
            post_j[c4] = S.post[l];
            if (triplet) fac_ltp[c4] = D.p.ltp_A + (D.p.ltp_B * S.tpost[l]);
        }
    }
    const bool any_col = __ballot(colspike != 0u) != 0ull;
    // Spike flags of the wave's rows (bit kk: row c0 + kk).
    uint32_t chunk_bits = 0;
#pragma unroll
    for (uint32_t r8 = 0; r8 < (uint32_t)kRes; r8 += 8) {
        if (c0 + r8 >= c1) continue;
        const uint32_t r0 = c0 + r8, wl = r0 >> 5, sh = r0 & 31u;
        const uint32_t lo = S.rbits[wl];
        const uint32_t hi = sh > 24u && wl + 1u < kLSegMax / 32u ? S.rbits[wl + 1u] : 0u;
        chunk_bits |= ((sh ? (lo >> sh) | (hi << (32u - sh)) : lo) & 0xFFu) << r8;
    }
    const uint32_t nrows = c1 > c0 ? c1 - c0 : 0u;
    chunk_bits &= nrows >= 32u ? 0xFFFFFFFFu : ((1u << nrows) - 1u);
    const uint32_t nr = c1 > c0 ? c1 - c0 : 0u;
    if (lazy && lane < nr) S.dr[c0 + lane] = 0.0;
    // The usual case: no row of the chunk spiked and at most two lanes hold the spiking columns.
    // Those lanes' rows go through LDS and lane r of the wave updates row r (LTP of the spiking
    // columns), so the update runs across the wave rather than down one lane; a row's tile sum is
    // the sum of the two lanes' changes (the only non-zero terms of the lane reduction).
    const uint64_t col_lanes = __ballot(colspike != 0u);
    const uint32_t ncl = (uint32_t)__popcll(col_lanes);
    if (any_col && !chunk_bits && ncl <= 2u) {
        v4f* T = &S.pbuf[wave][0];
        const uint32_t la = (uint32_t)__ffsll((long long)col_lanes) - 1u;
        const uint64_t rest = col_lanes & (col_lanes - 1ull);
        const uint32_t lb = rest ? (uint32_t)__ffsll((long long)rest) - 1u : 64u;
        const uint32_t cs_a = (uint32_t)__shfl((int)colspike, (int)la, 64);
        const uint32_t cs_b = lb < 64u ? (uint32_t)__shfl((int)colspike, (int)lb, 64) : 0u;
        if (lane == la || lane == lb) {
            const uint32_t base = lane == la ? 0u : 32u;
#pragma unroll
            for (uint32_t kk = 0; kk < (uint32_t)kRes; ++kk)
                if (c0 + kk < c1) T[base + kk] = w[kk];
        }
        uint32_t t = 0;
        if (lane < nr) {
            const uint32_t r = c0 + lane, i = s0 + r;
            const float si = lazy ? S.s[r] : 1.0f;
            double rs = 0.0;
            if (!lazy || si != 0.0f) {
                const float pre_i = S.pre[r];
#pragma unroll
                for (uint32_t h = 0; h < 2; ++h) {
                    const uint32_t cs = h == 0 ? cs_a : cs_b;
                    if (!cs) continue;
                    const uint32_t lh = h == 0 ? la : lb;
                    const uint32_t jh = col * kPCols + 4u * lh;
                    v4f m4 = T[32u * h + lane];
                    double rh = 0.0;
#pragma unroll
                    for (uint32_t c4 = 0; c4 < 4; ++c4) {
                        if (!((cs >> c4) & 1u) || jh + c4 == i) continue;
                        const float m = m4[c4];
                        const float fl = triplet ? D.p.ltp_A + (D.p.ltp_B * S.tpost[4u * lh + c4]) : 0.0f;
                        const float delta = triplet ? pre_i * fl : rule.ltp_A * pre_i;
                        const float val = p_apply(lazy, m, si, delta, rule.weight_min, rule.weight_max);
                        rh += (double)val - (double)m;
                        m4[c4] = val;
                        t |= 1u << h;
                    }
                    T[32u * h + lane] = m4;
                    rs = h == 0 ? rh : rs + rh;
                }
            }
            if (lazy) S.dr[r] = rs;
        }
        const uint32_t rows_a = (uint32_t)__ballot(t & 1u), rows_b = (uint32_t)__ballot((t >> 1) & 1u);
        if (lane == la || lane == lb) {
            const uint32_t base = lane == la ? 0u : 32u;
#pragma unroll
            for (uint32_t kk = 0; kk < (uint32_t)kRes; ++kk)
                if (c0 + kk < c1) w[kk] = T[base + kk];
            dirty |= lane == la ? rows_a : rows_b;
        }
    } else if (any_col || chunk_bits) {
        // Rows through LDS 8 at a time: a row whose pre neuron spiked takes the whole row update
        // (LTP then LTD, every column); any other row only the LTP of the spiking columns, on the
        // lanes that hold them.
        v4f* pb = &S.pbuf[wave][0];
#pragma unroll
        for (uint32_t b8 = 0; b8 < (uint32_t)kRes; b8 += 8) {
            if (c0 + b8 >= c1) continue;
            if (!any_col && ((chunk_bits >> b8) & 0xFFu) == 0u) continue;
#pragma unroll
            for (uint32_t kk = 0; kk < 8; ++kk) pb[kk * 64u + lane] = w[b8 + kk];
#pragma unroll 1
            for (uint32_t kk = 0; kk < 8; ++kk) {
                const uint32_t r = c0 + b8 + kk;
                if (r >= c1) break;
                const bool sp = (chunk_bits >> (b8 + kk)) & 1u;
                if (!sp && !any_col) continue;
                const float si = lazy ? S.s[r] : 1.0f;
                const uint32_t i = s0 + r;
                v4f m4 = pb[kk * 64u + lane];
                double lsum = 0.0;
                uint32_t touched = 0;
                if (sp) {
                    touched = p_plast_row(rule, lazy, triplet, true, i, si, S.pre[r], S.fac[r], j0, active, C.n_post,
                                          colspike, post_j, fac_ltp, m4, lsum);
                } else {
                    double rs = 0.0;
                    if (colspike && (!lazy || si != 0.0f)) {
                        const float pre_i = S.pre[r];
#pragma unroll
                        for (uint32_t c4 = 0; c4 < 4; ++c4) {
                            if (!((colspike >> c4) & 1u) || j0 + c4 == i) continue;
                            const float m = m4[c4];
                            const float delta = triplet ? pre_i * fac_ltp[c4] : rule.ltp_A * pre_i;
                            const float val = p_apply(lazy, m, si, delta, rule.weight_min, rule.weight_max);
                            rs += (double)val - (double)m;
                            m4[c4] = val;
                            touched |= 1u << c4;
                        }
                    }
                    if (lazy) {
                        // The lanes' sums combined as the butterfly of the row update combines
                        // them; with at most two non-zero lanes that is their sum.
                        const uint64_t nz = __ballot(rs != 0.0);
                        if (__popcll(nz) <= 2) {
                            double t = 0.0;
                            if (nz) {
                                const int l0 = __ffsll((long long)nz) - 1;
                                t = __shfl(rs, l0, 64);
                                const uint64_t rest = nz & (nz - 1);
                                if (rest) t = t + __shfl(rs, __ffsll((long long)rest) - 1, 64);
                            }
                            lsum = t;
                        } else {
#pragma unroll
                            for (int off = 32; off >= 1; off >>= 1) rs += __shfl_xor(rs, off, 64);
                            lsum = rs;
                        }
                    }
                }
                if (touched) {
                    pb[kk * 64u + lane] = m4;
                    dirty |= 1u << (b8 + kk);
                }
                if (lazy && lane == 0) S.dr[r] = lsum;
            }
#pragma unroll
            for (uint32_t kk = 0; kk < 8; ++kk) w[b8 + kk] = pb[kk * 64u + lane];
        }
    }
    __syncthreads();
    if (lazy && tid < seg_rows)
        l_put_d(P.ll_dr + ((size_t)l_slot(a) * C.ncol + col) * 2u * C.n_pre + s0 + tid, C.n_pre, a, S.dr[tid]);
    if (tid < (seg_rows + 31u) / 32u) S.rbits_prev[tid] = S.rbits[tid];
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// The launch.  A block owns one item, or (when the device has room) only the evolves of one
// column tile; an item block without room for a separate evolver also evolves a tile.
// ---------------------------------------------------------------------------------------
template <int kRes>
__global__ __launch_bounds__(kPThreads) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_persist_ll(const PPlan* __restrict__ plan, uint32_t n) {
    __shared__ LShared S;
    const PPlan& P = *plan;
    const bn_device_clock clk = *P.clock;
    const uint32_t t0 = clk.clock_step;
    const uint32_t blk = blockIdx.x;
    const uint32_t ib = P.block_begin[blk];
    const bool has_item = ib < P.block_begin[blk + 1];
    const PItem it = P.items[has_item ? ib : 0u];
    const uint32_t conn = __builtin_amdgcn_readfirstlane(it.conn);
    const uint32_t col = __builtin_amdgcn_readfirstlane(it.col);
    const uint32_t seg = __builtin_amdgcn_readfirstlane(it.seg);
    const uint32_t tile = __builtin_amdgcn_readfirstlane(P.block_tile[blk]);
    const PConn& C = P.c[conn];
    const PRows Rw = p_rows(C, col, seg);
    const bool plastic = has_item && C.plast != kPNone;
    const PPlast& D = P.pl[plastic ? C.plast : 0u];
    const bool lazy = plastic && D.lazy;
    if (threadIdx.x == 0) S.abort = 0u;
    // The item's rows into registers; the lazy rows' sums, scales and targets into LDS.
    v4f w[kRes];
    {
        const BN_GAS float* wrow = gp(C.W) + (size_t)(Rw.s0 + Rw.c0) * C.ld + Rw.j0;
#pragma unroll
        for (uint32_t kk = 0; kk < (uint32_t)kRes; ++kk)
            w[kk] = has_item && Rw.active && Rw.c0 + kk < Rw.c1
                        ? *reinterpret_cast<const BN_GAS v4f*>(wrow + (size_t)kk * C.ld)
                        : v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
    uint32_t dirty = 0;
    if (lazy)
        for (uint32_t r = threadIdx.x; r < Rw.seg_rows; r += kPThreads) {
            S.R[r] = gp(D.R)[Rw.s0 + r];
            S.s[r] = gp(D.s)[Rw.s0 + r];
            S.target[r] = gp(D.target)[Rw.s0 + r];
        }
    uint32_t gi = 0, q = 0;
    if (tile != kPNone) {
        while (gi + 1 < P.n_groups && tile >= P.g[gi + 1].col_base) ++gi;
        q = tile - P.g[gi].col_base;
        l_state(P, S, gi, q, false);
    }
    __syncthreads();
    bool ok = true;
    // The launch's first evolve (counted in the second word of the step count).
    if (tile != kPNone) {
        ok = l_evolve(P, S, gi, q, t0, 0, n, clk, 0u);
        __syncthreads();
        if (ok && threadIdx.x == 0) (void)add_c(P.done + 1, 1u);
    }
    uint32_t k = 0;
    for (; ok && k < n; ++k) {
        if (has_item) ok = l_item_front<kRes>(P, S, conn, col, seg, t0 + k, k, n, w);
        if (ok && tile != kPNone && k + 1 < n) {
            const uint32_t seen = threadIdx.x == 0 ? ld_c(P.done) : 0u;
            ok = l_evolve(P, S, gi, q, t0 + k + 1, k + 1, n, clk, seen);
        }
        if (ok && has_item) l_item_back<kRes>(P, S, conn, col, seg, t0 + k, w, dirty);
        __syncthreads();
        if (ok && threadIdx.x == 0) (void)add_c(P.done, 1u);
    }
    if (!ok) return;
    // After the last step: the last step's stimulus events for the evolver's neurons and its
    // neurons' state back to the state's buffers; the lazy segment's final row sums and scales
    // (the block of column tile 0); the resident rows that changed.
    const uint32_t last = t0 + n - 1u;
    if (tile != kPNone) {
        const NeuronGroupView& v = P.g[gi].v;
        const uint32_t lid = q * kPCols + threadIdx.x;
        if (threadIdx.x < (uint32_t)kPCols && lid < v.num_neurons && P.ev_steps && last < P.ev_steps) {
            const uint32_t gid = v.offset + lid;
            const uint32_t e1 = min(P.ev_start[last + 1], P.ev_count);
            for (uint32_t e = P.ev_start[last]; e < e1; ++e)
                if (P.ev_ids[e] + P.ev_offsets[e] == gid) S.ev[kEvIstim][threadIdx.x] += P.ev_amps[e];
        }
        l_state(P, S, gi, q, true);
    }
    if (lazy && col == 0) {
        const uint32_t seg_rows = Rw.seg_rows, tid = threadIdx.x;
        const uint32_t ndr = C.ncol * seg_rows;
        uint32_t want = 0;
        for (uint32_t u = 0; u < kLDrJobs; ++u)
            if (tid + kPThreads * u < ndr) want |= 3u << (3u + 2u * u);
        const uint32_t sl = l_slot(last);
        auto addr = [&](int i) -> const uint64_t* {
            const uint32_t d = tid + kPThreads * (uint32_t)((i - 3) >> 1);
            const uint32_t qd = d / seg_rows, r = d - qd * seg_rows;
            return P.ll_dr + (((size_t)sl * C.ncol + qd) * 2u + (uint32_t)((i - 3) & 1)) * C.n_pre + Rw.s0 + r;
        };
        uint32_t* drin32 = reinterpret_cast<uint32_t*>(S.drin);
        auto sink = [&](int i, uint32_t bits) {
            drin32[(tid + kPThreads * (uint32_t)((i - 3) >> 1)) * 2u + (uint32_t)((i - 3) & 1)] = bits;
        };
        if (!l_poll<(int)(3 + 2 * kLDrJobs)>(P, S, want, l_tag(last), l_tag(last), addr, sink)) return;
        if (tid < seg_rows) {
            const bool spiked = (S.rbits_prev[tid >> 5] >> (tid & 31u)) & 1u;
            double R = spiked ? 0.0 : S.R[tid];
            for (uint32_t qq = 0; qq < C.ncol; ++qq) R += S.drin[qq * seg_rows + tid];
            gp(D.R)[Rw.s0 + tid] = R;
            gp(D.s)[Rw.s0 + tid] = R > 0.0 ? (float)((double)S.target[tid] / R) : S.s[tid];
        }
    }
    if (dirty) {
        BN_GAS float* wrow = gp(C.W) + (size_t)(Rw.s0 + Rw.c0) * C.ld + Rw.j0;
#pragma unroll
        for (uint32_t kk = 0; kk < (uint32_t)kRes; ++kk)
            if ((dirty >> kk) & 1u) *reinterpret_cast<BN_GAS v4f*>(wrow + (size_t)kk * C.ld) = w[kk];
    }
    // The clock advance of the whole launch, once every block has finished every step (so every
    // block has read the clock).
    if (blk == 0) {
        if (!l_wait(P, S, P.done, n * P.blocks, 0u)) return;
        if (threadIdx.x == 0) {
            bn_device_clock c = clk;
            c.clock_step = t0 + n;
            c.next_event_idx = P.ev_steps ? P.ev_start[min(t0 + n, P.ev_steps)] : clk.next_event_idx;
            *P.clock_rw = c;
        }
    }
}

static const void* persist_ll_kernel(uint32_t resident_rows) {
    switch (resident_rows) {
        case 16: return reinterpret_cast<const void*>(k_persist_ll<16>);
        case 24: return reinterpret_cast<const void*>(k_persist_ll<24>);
        default: return nullptr;
    }
}

int persistent_ll_max_blocks(uint32_t resident_rows, uint32_t* blocks) {
    const void* f = persist_ll_kernel(resident_rows);
    if (!f) return (int)hipErrorInvalidValue;
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kPThreads, 0);
    *blocks = e == hipSuccess ? (uint32_t)std::max(0, cus * per_cu) : 0u;
    return (int)e;
}

int launch_persistent_ll(const PPlan* d_plan, const PPlan& hp, uint32_t n, bn_stream_t s) {
    const void* f = persist_ll_kernel(hp.resident_rows);
    if (!f) return (int)hipErrorInvalidValue;
    void* args[] = {(void*)&d_plan, (void*)&n};
    return (int)hipLaunchCooperativeKernel(f, dim3(hp.blocks), dim3(kPThreads), args, 0, s);
}

}  // namespace bn
