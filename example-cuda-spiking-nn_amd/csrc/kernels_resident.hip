// The resident step (bn_resident.h): one cooperative launch runs n steps of a network whose
// weights stay in the registers of the co-resident blocks for the whole launch.
//
// Per step a, every block (owning C post columns of one group, all pre rows of them):
//   poll      the spike words of step a of every block (one wave) and, for the rows of the
//             lazily normalised matrix, their scales of step a (the rows' threads);
//   rows      each thread advances its rows' connection traces and plasticity pre traces from
//             the spike bits (kernel_connection_propagate.cu:8-46, the trace step of
//             kernel_triplet_plasticity_update.cu:16-50 with the previous step's +1 first);
//   propagate FP64 column sums of x_i (s_i) W_ij over the thread's rows, the lane butterfly,
//             the waves in order: g of the block's neurons (kernel_connection_propagate.cu:79-94);
//   plastic   LTP of the block's spiking columns and LTD of the spiking rows on the entries in
//             registers, LTP then LTD per entry (kernel_triplet_plasticity_update.cu:54-93,
//             kernel_istdp_plasticity_update.cu:42-80, kernel_stdp_plasticity_update.cu);
//             the lazy matrix's row-sum parts go to the rows' owners;
//   evolve    the block's neurons for step a + 1 (kernel_neuron_group_evolve.cu:11-130, the
//             stimulus events of step a first, stimulus.cpp:46-60) and their spike word;
//   owner     for its own rows of the lazy matrix: the parts of the row sums, R_i, s_i
//             (kernel_triplet_plasticity_update.cu:117-145 kept as W = diag(s) M), the scales
//             of step a + 1.
// Before the first step the block loads its entries, its rows' traces and its neurons' state
// and runs the evolve of the launch's first step; after the last it stores them back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "bn_device.h"
#include "bn_errors.h"
#include "bn_neuron.h"
#include "bn_persist_dev.h"
#include "bn_resident.h"

namespace bn {

namespace {

__device__ __forceinline__ uint32_t r_tag(uint32_t a) { return a + 1u; }  // 0: never written
__device__ __forceinline__ uint32_t r_slot(uint32_t a) { return a % 3u; }
__device__ __forceinline__ uint64_t r_word(uint32_t a, uint32_t bits) { return ((uint64_t)r_tag(a) << 32) | bits; }

template <int K>
struct RShared {
    // The block's neurons (its post columns).
    float V[kRCols], istim[kRCols], re[kRCols], ri[kRCols], de[kRCols], di[kRCols];
    float bge[kRCols], bgi[kRCols], ge[kRCols], gi[kRCols];
    uint32_t refr[kRCols], spike[kRCols];
    float post[2][3][kRCols];          // post traces of step a in [a & 1]: E->E fast, E->E slow, I->E
    int32_t vrec[kRCols];              // slot in the membrane-potential ring, or -1
    double R[kRCols];                  // lazy E->E: the block's rows' sums, scales, targets
    float s[kRCols], target[kRCols];
    uint32_t heads[8];                 // first four entries of the group's Poisson tables (e, i)
    float rk[2][5];                    // per pre group: r decay, d decay, K, pre decay, slow pre decay
    uint32_t abort;
    // The threads' rows, [row][thread].
    float r[K][kRThreads], d[K][kRThreads], pre[K][kRThreads], tpre[K][kRThreads], sc[K][kRThreads];
    float xs[K][kRThreads];            // the rows' x_i (s_i) of the step, FP32
    uint32_t spk[kRMaxBlocks];         // spike words of the step, block order
    double red[kRWaves][kRCols];       // the waves' column sums
    float zy[kRWaves][2][kRCols];      // per wave and step: LTP factors Z_c, LTD factors Y_c
    double part[kRCols][kRMaxBlocks];  // owner: the polled row-sum parts, [row][rank]
};

// Polls up to M words per thread (bit m of `want`: the word at addr(m)) until each carries the
// tag, handing its low half to sink(m, bits).  Every thread of the block calls it; false for
// every thread when a wait ran past its limit here or in another block.
template <int M, int K, typename Addr, typename Sink>
__device__ __forceinline__ bool r_poll(const RPlan& P, RShared<K>& S, uint32_t want, uint32_t tag, Addr addr,
                                       Sink sink) {
    uint32_t pending = want;
    for (uint32_t it = 0; pending; ++it) {
        uint64_t v[M];
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = (pending >> m) & 1u ? ld_c(addr(m)) : 0ull;
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (((pending >> m) & 1u) && (uint32_t)(v[m] >> 32) == tag) {
                sink(m, (uint32_t)v[m]);
                pending &= ~(1u << m);
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

// As r_poll for the threads that call it, without the closing barrier: the caller's next
// block barrier is followed by the check of S.abort.
template <int M, int K, typename Addr, typename Sink>
__device__ __forceinline__ void r_poll_nb(const RPlan& P, RShared<K>& S, uint32_t want, uint32_t tag, Addr addr,
                                          Sink sink) {
    uint32_t pending = want;
    for (uint32_t it = 0; pending; ++it) {
        uint64_t v[M];
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = (pending >> m) & 1u ? ld_c(addr(m)) : 0ull;
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (((pending >> m) & 1u) && (uint32_t)(v[m] >> 32) == tag) {
                sink(m, (uint32_t)v[m]);
                pending &= ~(1u << m);
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
}

// Index of the k-th set bit (k < popcount) of the 256-bit mask m[0..3].
__device__ __forceinline__ uint32_t r_nth_set256(const uint64_t (&m)[4], uint32_t k) {
    uint32_t base = 0;
    uint64_t x = m[0];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        const uint32_t c = (uint32_t)__popcll(x);
        if (k < c) break;
        k -= c;
        base += 64u;
        x = m[w + 1];
    }
#pragma unroll
    for (uint32_t width = 32; width >= 1; width >>= 1) {
        const uint64_t lo = x & ((1ull << width) - 1ull);
        const uint32_t c = (uint32_t)__popcll(lo);
        if (k >= c) {
            k -= c;
            base += width;
            x >>= width;
        } else {
            x = lo;
        }
    }
    return base;
}

// Global id of pre-local row 0 of the E rows (0) or the I rows (ne).
__device__ __forceinline__ uint32_t pre_base_of(const RPlan& P, bool eside) { return eside ? 0u : P.ne; }

// Spike bit of global neuron gid in the step's block words.
template <int K>
__device__ __forceinline__ uint32_t r_spiked(const RPlan& P, const RShared<K>& S, uint32_t gid) {
    const uint32_t lid = gid < P.ne ? gid : gid - P.ne;
    const uint32_t blk = (gid < P.ne ? 0u : P.be) + lid / kRCols;
    return (S.spk[blk] >> (lid % kRCols)) & 1u;
}

// Word wd of the step's spike bits in global ids (bit g & 31 of word g >> 5), from the block
// words: an E block word holds 16 aligned neurons, an I block word the 16 from ne + 16 * (its
// I index), which need not be aligned to the words.
template <int K>
__device__ __forceinline__ uint32_t r_global_word(const RPlan& P, const RShared<K>& S, uint32_t wd) {
    const uint32_t lo = 32u * wd, hi = min(lo + 32u, P.n);
    uint32_t word = 0;
    if (lo < P.ne) {
        word = S.spk[lo / (uint32_t)kRCols];
        if (lo + (uint32_t)kRCols < P.ne) word |= S.spk[lo / (uint32_t)kRCols + 1u] << kRCols;
    }
    const uint32_t ilo = max(lo, P.ne);
    if (ilo < hi)
        for (uint32_t l = (ilo - P.ne) & ~(uint32_t)(kRCols - 1); l < hi - P.ne; l += (uint32_t)kRCols) {
            const uint32_t bw = S.spk[P.be + l / (uint32_t)kRCols];
            const int sh = (int)(P.ne + l) - (int)lo;   // bit of I neuron l in the word (-15 .. 31)
            word |= sh >= 0 ? bw << sh : bw >> -sh;
        }
    return word;
}

// Position of the k-th set bit of m (k < popcount(m)).
__device__ __forceinline__ uint32_t r_nth_bit(uint32_t m, uint32_t k) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t b = 0; b < (uint32_t)kRCols; ++b) {
        const bool set = (m >> b) & 1u;
        if (set && k == 0u) pos = b;
        if (set) --k;
    }
    return pos;
}

// ---------------------------------------------------------------------------------------
// Evolve of the block's neurons for step a: threads [0, ncols).  Events [e0, e1) are the
// stimulus events of step a - 1 (none for the launch's first step: those are in the state).
// Returns the thread's spike; the spike word of the block is published by the caller.
// ---------------------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ uint32_t r_evolve(const RPlan& P, RShared<K>& S, const NeuronGroupView& v, bool epost,
                                             uint32_t j0, uint32_t c, uint32_t a, uint32_t e0, uint32_t e1,
                                             const bn_device_clock& clk) {
    const uint32_t gid = v.offset + j0 + c;
    float istim = S.istim[c];
    // stimulus.cpp:46-60 for this neuron: the events in list order.
    for (uint32_t e = e0; e < e1; ++e)
        if (P.ev_ids[e] + P.ev_offsets[e] == gid) istim += P.ev_amps[e];
    // Background (kernel_neuron_group_evolve.cu:11-83).
    uint32_t he[4], hi[4];
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
        he[q] = S.heads[q];
        hi[q] = S.heads[4 + q];
    }
    float re = flush_trace(S.re[c] * v.background_excitatory_postsynaptic_r_trace_decay);
    float ri = flush_trace(S.ri[c] * v.background_inhibitory_postsynaptic_r_trace_decay);
    float de = flush_trace(S.de[c] * v.background_excitatory_postsynaptic_d_trace_decay);
    float di = flush_trace(S.di[c] * v.background_inhibitory_postsynaptic_d_trace_decay);
    uint32_t ke = 0, ki = 0;
    if (v.poisson_table_excitatory_len > 0 || v.poisson_table_inhibitory_len > 0) {
        const U32x4 r = background_draws(clk, a, gid, v.rng_instance);
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
    // EIF (kernel_neuron_group_evolve.cu:85-130) on the g of the last propagate.
    float V = S.V[c];
    uint32_t refr = S.refr[c], spike = 0;
    eif_step(v, S.ge[c], S.gi[c], istim, V, refr, spike);
    // Post traces of the plasticities (E neurons): the previous step's +1, then the decay.
    if (epost) {
        const bool was = S.spike[c] == 1u;
        const float(&pp)[3][kRCols] = S.post[(a - 1u) & 1u];
        float(&pn)[3][kRCols] = S.post[a & 1u];
        if (P.ee_kind != kRNone) {
            pn[0][c] = flush_trace((was ? pp[0][c] + 1.0f : pp[0][c]) * P.ee_dec[2]);
            if (P.ee_kind == kPlastTriplet) pn[1][c] = flush_trace((was ? pp[1][c] + 1.0f : pp[1][c]) * P.ee_dec[3]);
        }
        if (P.ie_kind != kRNone) pn[2][c] = flush_trace((was ? pp[2][c] + 1.0f : pp[2][c]) * P.ie_dec[1]);
    }
    S.V[c] = V;
    S.refr[c] = refr;
    S.spike[c] = spike;
    S.istim[c] = istim;
    S.re[c] = re;
    S.ri[c] = ri;
    S.de[c] = de;
    S.di[c] = di;
    S.bge[c] = bg_ge;
    S.bgi[c] = bg_gi;
    if (P.v_ring && S.vrec[c] >= 0) gp(P.v_ring)[(size_t)(a % P.interval) * P.n_vrec + (uint32_t)S.vrec[c]] = V;
    return spike;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// The launch.  K: rows per thread (the plan's `rows`).  Roles within a step: the last wave
// finalises g and evolves the block's neurons (the spike word every block waits for) while the
// other waves apply the plasticity and, in E-post blocks, poll their rows' row-sum parts.
// ---------------------------------------------------------------------------------------
typedef float v16f __attribute__((ext_vector_type(16)));
static_assert(kRCols == 16, "a row of the block's entries is one 16-wide vector");
constexpr uint32_t kREvolveWave = kRWaves - 1;

template <int K>
__global__ __launch_bounds__(kRThreads) void k_resident(const RPlan* __restrict__ plan, uint32_t n) {
    __shared__ RShared<K> S;
    const RPlan& P = *plan;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t b = blockIdx.x;
    const bool epost = b < P.be;
    const uint32_t j0 = (epost ? b : b - P.be) * (uint32_t)kRCols;   // first post-local column
    const uint32_t npost = epost ? P.ne : P.ni;
    const uint32_t ncols = min((uint32_t)kRCols, npost - j0);
    const NeuronGroupView& G = P.g[epost ? 0 : 1];
    // The thread's rows: pre group `side`, rows tl, tl + T, ...
    const bool eside = wave < P.e_waves;
    const uint32_t side = eside ? 0u : 1u;
    const uint32_t T = 64u * (eside ? P.e_waves : (uint32_t)kRWaves - P.e_waves);
    const uint32_t tl = eside ? tid : tid - 64u * P.e_waves;
    const uint32_t npre = eside ? P.ne : P.ni;
    const uint32_t conn = epost ? (eside ? 0u : 2u) : (eside ? 1u : 3u);
    const uint32_t pkind = !epost ? kRNone : (eside ? P.ee_kind : P.ie_kind);
    const bool lazy = epost && eside && P.ee_lazy;
    const bool triplet = pkind == kPlastTriplet;
    const PlasticityParams& PP = eside ? P.ee_p : P.ie_p;
    const bool evolver = wave == kREvolveWave;
    const bn_device_clock clk = *P.clock;
    const uint32_t t0 = clk.clock_step;
    if (tid == 0) S.abort = 0u;
    // The launch's loads go out together, the small ones first (the rows' traces and scales,
    // the block's neurons), then the weights; the LDS copies of the small ones wait only for
    // those.  Indices past a group are clamped to row 0 (the value is then discarded), so no
    // lane branches around a load.
    float rv[K], dv[K], pv[K], tv[K], sv[K];
    uint32_t spv[K];
    {
        const BN_GAS float* pr = gp(P.r[side]);
        const BN_GAS float* pd = gp(P.d[side]);
        const BN_GAS uint32_t* ps = gp(P.spikes) + pre_base_of(P, eside);
        const BN_GAS float* ppre = gp(eside ? P.ee_tr[0] : P.ie_tr[0]);
        const BN_GAS float* ptp = gp(P.ee_tr[1]);
        const BN_GAS float* psc = gp(P.s);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t i = tl + T * (uint32_t)k;
            const uint32_t ii = i < npre ? i : 0u;
            rv[k] = pr[ii];
            dv[k] = pd[ii];
            spv[k] = ps[ii];
            pv[k] = pkind != kRNone ? ppre[ii] : 0.0f;
            tv[k] = triplet ? ptp[ii] : 0.0f;
            sv[k] = lazy ? psc[ii] : 1.0f;
        }
    }
    // The block's neurons (threads [0, ncols)).
    const bool holder = tid < ncols;
    const uint32_t lid0 = j0 + (holder ? tid : 0u);
    float nV = 0.0f, nistim = 0.0f, nre = 0.0f, nri = 0.0f, nde = 0.0f, ndi = 0.0f, nge = 0.0f, ngi = 0.0f;
    float np0 = 0.0f, np1 = 0.0f, np2 = 0.0f, ns = 0.0f, ntg = 0.0f;
    uint32_t nrefr = 0, nspike = 0;
    int32_t nvrec = -1;
    double nR = 0.0;
    if (wave == 0) {
        const NeuronGroupView& v = G;
        nV = gp(v.membrane_potentials)[lid0];
        nrefr = gp(v.refractory_counter)[lid0];
        nistim = gp(v.i_stimulus)[lid0];
        nre = gp(v.background_excitatory_postsynaptic_r_traces)[lid0];
        nri = gp(v.background_inhibitory_postsynaptic_r_traces)[lid0];
        nde = gp(v.background_excitatory_postsynaptic_d_traces)[lid0];
        ndi = gp(v.background_inhibitory_postsynaptic_d_traces)[lid0];
        nge = gp(v.g_excitatory)[lid0];
        ngi = gp(v.g_inhibitory)[lid0];
        nspike = gp(P.spikes)[v.offset + lid0];
        if (epost && P.ee_kind != kRNone) np0 = gp(P.ee_tr[2])[lid0];
        if (epost && P.ee_kind == kPlastTriplet) np1 = gp(P.ee_tr[3])[lid0];
        if (epost && P.ie_kind != kRNone) np2 = gp(P.ie_tr[1])[lid0];
        if (P.v_ring) nvrec = gp(P.vrec_index)[v.offset + lid0];
        if (epost && P.ee_lazy) {
            nR = gp(P.R)[lid0];
            ns = gp(P.s)[lid0];
            ntg = gp(P.target)[lid0];
        }
    }
    // The entries of the block's columns in the thread's rows (padding columns and rows: 0).
    v16f w[K];
    {
        const BN_GAS float* W = gp(P.W[conn]);
        const uint32_t ld = P.ld[conn];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t i = tl + T * (uint32_t)k;
            const bool ok = i < npre;
            const BN_GAS v4f* src = reinterpret_cast<const BN_GAS v4f*>(W + (size_t)(ok ? i : 0u) * ld + j0);
#pragma unroll
            for (int q = 0; q < kRCols / 4; ++q) {
                const v4f x = src[q];
#pragma unroll
                for (int e = 0; e < 4; ++e) w[k][4 * q + e] = ok && (uint32_t)(4 * q + e) < ncols ? x[e] : 0.0f;
            }
        }
    }
    // The rows' traces and scales; bit k of sp_prev: row k's neuron spiked at the step before.
    uint32_t sp_prev = 0;
    // Entries the launch changes: rows with an LTD (per thread) and columns with an LTP (per
    // block); only those are stored back after the last step.
    uint32_t dirty_rows = 0, dirty_cols = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const bool ok = tl + T * (uint32_t)k < npre;
        S.r[k][tid] = ok ? rv[k] : 0.0f;
        S.d[k][tid] = ok ? dv[k] : 0.0f;
        S.pre[k][tid] = ok ? pv[k] : 0.0f;
        S.tpre[k][tid] = ok ? tv[k] : 0.0f;
        S.sc[k][tid] = ok ? sv[k] : 1.0f;
        if (ok && spv[k] == 1u) sp_prev |= 1u << k;
    }
    // The block's neurons (post traces in the buffer of the step before the launch).
    if (holder) {
        float(&pp)[3][kRCols] = S.post[(t0 - 1u) & 1u];
        S.V[tid] = nV;
        S.refr[tid] = nrefr;
        S.istim[tid] = nistim;
        S.re[tid] = nre;
        S.ri[tid] = nri;
        S.de[tid] = nde;
        S.di[tid] = ndi;
        S.ge[tid] = nge;
        S.gi[tid] = ngi;
        S.spike[tid] = nspike;
        pp[0][tid] = np0;
        pp[1][tid] = np1;
        pp[2][tid] = np2;
        S.vrec[tid] = nvrec;
        S.R[tid] = nR;
        S.s[tid] = ns;
        S.target[tid] = ntg;
    }
    if (tid < 4) {
        S.heads[tid] = tid < G.poisson_table_excitatory_len ? gp(G.poisson_table_excitatory)[tid] : 0u;
        S.heads[4 + tid] = tid < G.poisson_table_inhibitory_len ? gp(G.poisson_table_inhibitory)[tid] : 0u;
    }
    if (tid < 2) {
        S.rk[tid][0] = P.decay_r[tid];
        S.rk[tid][1] = P.decay_d[tid];
        S.rk[tid][2] = P.K[tid];
        S.rk[tid][3] = tid == 0 ? P.ee_dec[0] : P.ie_dec[0];
        S.rk[tid][4] = P.ee_dec[1];
    }
    // The spike bit of the thread's row q: bit (tl % 16) of block word spk_row0 + q * spk_step
    // (a thread's rows are 16-aligned apart, so they share the bit); vmask: rows within the group.
    const uint32_t spk_row0 = (eside ? 0u : P.be) + tl / (uint32_t)kRCols;
    const uint32_t spk_step = T / (uint32_t)kRCols, spk_bit = tl % (uint32_t)kRCols;
    uint32_t vmask = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) vmask |= (tl + T * (uint32_t)q < npre ? 1u : 0u) << q;
    // Every load of the state has returned before the block publishes anything: another block
    // stores its part of the state back after the launch's last step, which can follow this
    // block's first spike word closely.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // The evolve of the launch's first step.
    if (evolver) {
        const uint32_t sp = lane < ncols ? r_evolve(P, S, G, epost, j0, lane, t0, 0u, 0u, clk) : 0u;
        const uint32_t bits = (uint32_t)__ballot(sp == 1u);
        if (lane == 0) st_c(P.spk + (size_t)r_slot(t0) * P.blocks + b, r_word(t0, bits));
    }
    uint32_t ev0 = 0, ev1 = 0;   // stimulus events of step a (applied by the evolve of step a + 1)
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t a = t0 + k;
        // The row and column bases count as rewritten each step (an empty register constraint,
        // no instruction): otherwise the compiler hoists every entry's self-skip and bounds test
        // out of the step loop and keeps those lane masks in spilled scalar registers.
        uint32_t tl_k = tl, j0_k = j0, ncols_k = ncols;
        asm volatile("" : "+v"(tl_k), "+s"(j0_k), "+s"(ncols_k));
        if (P.ev_steps && a < P.ev_steps) {  // loaded now, used after the poll
            ev0 = P.ev_start[a];
            ev1 = min(P.ev_start[a + 1], P.ev_count);
        } else {
            ev0 = ev1 = 0;
        }
        // ---- poll: every block's spike word of step a (the lazy rows' scales, which come
        // later, are polled after the rows' traces).
        {
            uint32_t want = 0;
            if (wave == 0)
                for (uint32_t m = 0; m < kRMaxBlocks / 64; ++m)
                    if (lane + 64u * m < P.blocks) want |= 1u << m;
            const uint32_t sl = r_slot(a);
            auto addr = [&](int m) -> const uint64_t* {
                return P.spk + (size_t)sl * P.blocks + lane + 64u * (uint32_t)m;
            };
            auto sink = [&](int m, uint32_t bits) { S.spk[lane + 64u * (uint32_t)m] = bits; };
            if (!r_poll<4>(P, S, want, r_tag(a), addr, sink)) return;
        }
        // ---- the monitors' spike ring and (last step) the state's spike bits, written by the
        // first wave of the I-post blocks (word w by I-post block w % bi): those blocks poll
        // nothing else, and a wave's later loads would wait for these stores.
        if (!epost && wave == 0 && (P.spike_ring || k + 1 == n)) {
            for (uint32_t w = (b - P.be) + P.bi * lane; w < P.words; w += P.bi * 64u) {
                const uint32_t word = r_global_word(P, S, w);
                if (P.spike_ring) gp(P.spike_ring)[(size_t)(a % P.interval) * P.words + w] = word;
                if (k + 1 == n) gp(P.bits)[w] = word;
            }
        }
        // ---- rows: connection traces with the spikes of step a, plasticity pre traces (the
        // previous step's +1 first), x_i for the propagate.
        uint32_t sp_now = 0;
        {
            const float dr = S.rk[side][0], dd = S.rk[side][1], Kn = S.rk[side][2];
            const float dpre = S.rk[side][3], dtpre = S.rk[side][4];
            float rv[K], dv[K];
            uint32_t wv[K];
#pragma unroll
            for (int q = 0; q < K; ++q) {
                wv[q] = S.spk[min(spk_row0 + spk_step * (uint32_t)q, kRMaxBlocks - 1u)];
                rv[q] = S.r[q][tid];
                dv[q] = S.d[q][tid];
            }
#pragma unroll
            for (int q = 0; q < K; ++q) sp_now |= ((wv[q] >> spk_bit) & 1u) << q;
            sp_now &= vmask;
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const bool sp = (sp_now >> q) & 1u;
                float d = flush_trace(dv[q] * dd);
                float r = flush_trace(rv[q] * dr);
                if (sp) {
                    r += Kn;
                    d += Kn;
                }
                S.d[q][tid] = d;
                S.r[q][tid] = r;
                S.xs[q][tid] = d - r;
            }
            if (pkind != kRNone) {
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const bool was = (sp_prev >> q) & 1u;
                    const float pr = S.pre[q][tid];
                    const float tp = S.tpre[q][tid];
                    S.pre[q][tid] = flush_trace((was ? pr + 1.0f : pr) * dpre);
                    if (triplet) S.tpre[q][tid] = flush_trace((was ? tp + 1.0f : tp) * dtpre);
                }
            }
        }
        // ---- the lazy rows' scales of step a, each thread its own rows (published by the rows'
        // owners after the plasticity of step a - 1; the step's first launch step has them from
        // the state).  No barrier: the propagate reads only the thread's rows, and a wait that
        // gives up is seen at the step's last barrier.
        if (lazy && k > 0) {
            const uint32_t sl = r_slot(a);
            auto addr = [&](int m) -> const uint64_t* { return P.sx + (size_t)sl * P.ne + tl_k + T * (uint32_t)m; };
            auto sink = [&](int m, uint32_t bits) { S.sc[m][tid] = __uint_as_float(bits); };
            r_poll_nb<K>(P, S, vmask, r_tag(a), addr, sink);
        }
        // ---- propagate: column sums of x_i (s_i) W_ij over the thread's rows (FP64 from 0.0,
        // row order), the lane butterfly (32,16,8,4,2,1) run as a reduce-scatter, the waves in
        // order.
        {
            double acc[kRCols];
#pragma unroll
            for (int c = 0; c < kRCols; ++c) acc[c] = 0.0;
            // (The FP64 product of two FP32 values is exact, so fma(x, w, acc) rounds as acc + x * w.
            // Rows past the group hold x = 0 and entries 0: they add +0.)
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const float xf = S.xs[q][tid];
                const double xq = (double)(lazy ? xf * S.sc[q][tid] : xf);
#pragma unroll
                for (int c = 0; c < kRCols; ++c) acc[c] = fma(xq, (double)w[q][c], acc[c]);
            }
            // The entries count as rewritten here (an empty register constraint, no instruction):
            // otherwise the compiler keeps their FP64 conversions from the sums above alive until
            // the plasticity below converts the same values again, twice the entries' registers.
#pragma unroll
            for (int q = 0; q < K; ++q) asm volatile("" : "+v"(w[q]));
            double u8[8], u4[4], u2[2];
            const bool h32 = lane & 32u, h16 = lane & 16u, h8 = lane & 8u, h4 = lane & 4u;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const double send = h32 ? acc[m] : acc[m + 8];
                const double keep = h32 ? acc[m + 8] : acc[m];
                u8[m] = keep + __shfl_xor(send, 32, 64);
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double send = h16 ? u8[m] : u8[m + 4];
                const double keep = h16 ? u8[m + 4] : u8[m];
                u4[m] = keep + __shfl_xor(send, 16, 64);
            }
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const double send = h8 ? u4[m] : u4[m + 2];
                const double keep = h8 ? u4[m + 2] : u4[m];
                u2[m] = keep + __shfl_xor(send, 8, 64);
            }
            double u1;
            {
                const double send = h4 ? u2[0] : u2[1];
                const double keep = h4 ? u2[1] : u2[0];
                u1 = keep + __shfl_xor(send, 4, 64);
            }
            u1 = u1 + __shfl_xor(u1, 2, 64);
            u1 = u1 + __shfl_xor(u1, 1, 64);
            if ((lane & 3u) == 0u) {
                const uint32_t col = (h32 ? 8u : 0u) + (h16 ? 4u : 0u) + (h8 ? 2u : 0u) + (h4 ? 1u : 0u);
                S.red[wave][col] = u1;
            }
        }
        __syncthreads();
        const uint32_t pa = a & 1u;
        const uint32_t own = __builtin_amdgcn_readfirstlane(S.spk[b]);   // the block's spiking columns
        // ---- evolver wave: g of the block's neurons, the evolve of step a + 1 (with the events
        // of step a) and its spike word.
        if (evolver) {
            if (lane < ncols) {
                double ge = (double)S.bge[lane];
                for (uint32_t q = 0; q < P.e_waves; ++q) ge += S.red[q][lane];
                double gi = (double)S.bgi[lane];
                for (uint32_t q = P.e_waves; q < (uint32_t)kRWaves; ++q) gi += S.red[q][lane];
                S.ge[lane] = (float)ge;
                S.gi[lane] = (float)gi;
            }
            if (k + 1 < n) {
                const uint32_t sp = lane < ncols ? r_evolve(P, S, G, epost, j0, lane, a + 1, ev0, ev1, clk) : 0u;
                const uint32_t bits = (uint32_t)__ballot(sp == 1u);
                if (lane == 0) st_c(P.spk + (size_t)r_slot(a + 1) * P.blocks + b, r_word(a + 1, bits));
            }
        }
        // ---- plasticity of step a on the entries in registers: LTP of the block's spiking
        // columns (a uniform index into each row's vector), then LTD of the spiking rows (one
        // pass per row slot that spikes on some lane of the wave).  Every rule is written as
        //   LTP: delta = pre_i * Z_c   (triplet: Z_c = ltp_A + ltp_B * slow post_c; else ltp_A)
        //   LTD: dl = X_i * Y_c        (triplet: X_i = ltd_A + ltd_B * slow pre_i, Y_c = -post_c;
        //                               iSTDP: X = -ltd_A, Y_c = post_c - ltd_B; pair: X = -ltd_A,
        //                               Y_c = post_c)
        // (the reference's products with their factors swapped, the same roundings), and the
        // wave's Z_c, Y_c go through its own LDS slots once per step.
        // (Entered by whole waves: lanes 0..15 fill the wave's factors for every lane.)
        if (pkind != kRNone && (own != 0u || __ballot(sp_now != 0u) != 0ull)) {
            const float wmin = PP.weight_min, wmax = PP.weight_max;
            float(&zc)[kRCols] = S.zy[wave][0];
            float(&yc)[kRCols] = S.zy[wave][1];
            if (lane < (uint32_t)kRCols) {
                const float(&post)[3][kRCols] = S.post[pa];
                const float pj = post[eside ? 0 : 2][lane];
                zc[lane] = triplet ? PP.ltp_A + (PP.ltp_B * post[1][lane]) : PP.ltp_A;
                yc[lane] = triplet ? -pj : (pkind == kPlastIstdp ? pj - PP.ltd_B : pj);
            }
            const float xw = -PP.ltd_A;   // X of iSTDP and pair
            double change[K];
            // The rows' pre traces and scales, loaded together before the column loop (up to 7
            // rows per thread; at 8 the registers are short and they are read per use).
            constexpr bool kPreload = K <= 7;
            float pr[K], sr[K];
#pragma unroll
            for (int q = 0; q < K; ++q) {
                change[q] = 0.0;
                pr[q] = kPreload ? S.pre[q][tid] : 0.0f;
                sr[q] = kPreload && lazy ? S.sc[q][tid] : 1.0f;
            }
            // One straight-line form over the rows, so that the rows' chains interleave: a row
            // of a matrix that is not lazily normalised has the scale 1 (1 * m and a quotient by
            // 1 are exact), and an entry that is not updated computes a value that the select
            // drops.
            for (uint32_t cm = own; cm; cm &= cm - 1u) {
                const uint32_t c = (uint32_t)__builtin_ctz(cm);   // uniform
                const uint32_t j = j0_k + c;
                const float z = zc[c];
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const uint32_t i = tl_k + T * (uint32_t)q;
                    const float m = w[q][c];
                    const float delta = (kPreload ? pr[q] : S.pre[q][tid]) * z;
                    const float si = kPreload ? sr[q] : (lazy ? S.sc[q][tid] : 1.0f);
                    float val = clamp_weight(si * m + delta, wmin, wmax) / si;
                    asm volatile("" : "+v"(val));   // computed on every lane, not under a branch
                    const bool ok = ((vmask >> q) & 1u) && j != i && si != 0.0f;
                    const float nw = ok ? val : m;
                    w[q][c] = nw;
                    change[q] += (double)nw - (double)m;
                }
            }
            // The lazy matrix's row-sum parts of the rows whose neuron did not spike (their LTP
            // changes, when a column spiked); a spiking row's part follows its LTD below.
            if (lazy && own != 0u) {
                uint64_t* const drow = P.dr + ((size_t)r_slot(a) * P.be + b) * P.ne * 2u + (size_t)tl_k * 2u;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    if (!((vmask >> q) & 1u) || ((sp_now >> q) & 1u)) continue;
                    const uint64_t bits = (uint64_t)__double_as_longlong(change[q]);
                    uint64_t* dst = drow + (size_t)T * 2u * (uint32_t)q;
                    st_c(dst, r_word(a, (uint32_t)bits));
                    st_c(dst + 1, r_word(a, (uint32_t)(bits >> 32)));
                }
            }
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const bool sp = (sp_now >> q) & 1u;
                if (__ballot(sp) == 0ull) continue;   // no lane of the wave has row q spiking
                const uint32_t i = tl_k + T * (uint32_t)q;
                float y[kRCols];
#pragma unroll
                for (int c = 0; c < kRCols; ++c) y[c] = yc[c];
                const float x = triplet ? PP.ltd_A + (PP.ltd_B * S.tpre[q][tid]) : xw;
                double part = 0.0;
                if (lazy) {
                    const float si = S.sc[q][tid];
                    const bool upd = sp && si != 0.0f;
#pragma unroll
                    for (int c = 0; c < kRCols; ++c) {
                        const float m = w[q][c];
                        const float v = clamp_weight(si * m + x * y[c], wmin, wmax) / si;
                        const bool ok = upd && (uint32_t)c < ncols_k && j0_k + (uint32_t)c != i;
                        w[q][c] = ok ? v : m;
                        if ((uint32_t)c < ncols_k) part += (double)w[q][c];
                    }
                    if (sp) {
                        const uint64_t bits = (uint64_t)__double_as_longlong(part);
                        uint64_t* dst = P.drs + (((size_t)r_slot(a) * P.ne + i) * P.be + b) * 2u;
                        st_c(dst, r_word(a, (uint32_t)bits));
                        st_c(dst + 1, r_word(a, (uint32_t)(bits >> 32)));
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < kRCols; ++c) {
                        const float m = w[q][c];
                        const float v = clamp_weight(m + x * y[c], wmin, wmax);
                        const bool ok = sp && (uint32_t)c < ncols_k && j0_k + (uint32_t)c != i;
                        w[q][c] = ok ? v : m;
                    }
                }
            }
        }
        // ---- owner (E-post blocks): the parts of the block's rows of the lazy matrix after
        // step a's plasticity, R_i and s_i of step a + 1.  One wave does it: the first wave of I
        // rows, which stores nothing before its polls (a wave's loads wait for its own earlier
        // stores), and runs beside the evolver.
        const bool owner = epost && P.ee_lazy;
        const uint32_t owner_wave = P.e_waves < kREvolveWave ? P.e_waves : 0u;
        if (owner && wave == owner_wave) {
            const uint32_t rows_mask = (1u << ncols_k) - 1u;
            const uint32_t spm = own & rows_mask, qm = ~own & rows_mask;
            // Contributors: the E-post blocks with a spiking column (block order).
            uint64_t cmask[4];
            uint32_t nc = 0;
#pragma unroll
            for (uint32_t m = 0; m < 4; ++m) {
                const uint32_t q = lane + 64u * m;
                cmask[m] = __ballot(q < P.be && S.spk[q] != 0u);
                nc += (uint32_t)__popcll(cmask[m]);
            }
            // Items: (contributor rank, quiet row) for the rows whose neuron did not spike, then
            // (spiking row, E-post block); lane l polls items l, l + 64, ... (neighbouring lanes
            // read neighbouring rows of one contributor, or neighbouring blocks of one row).
            const uint32_t nq = (uint32_t)__popc(qm), ns = (uint32_t)__popc(spm);
            const uint32_t nquiet = nq * nc, total = nquiet + ns * P.be;
            const uint32_t sl = r_slot(a);
            for (uint32_t base = 0; base < total; base += 4u * 64u) {
                uint32_t row_of[4], rank_of[4];
                const uint64_t* src_of[4];
                uint32_t want = 0;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t it = base + lane + 64u * (uint32_t)m;
                    uint32_t row = 0, rank = 0;
                    const uint64_t* src = P.dr;
                    if (it < nquiet) {
                        row = r_nth_bit(qm, it % nq);
                        rank = it / nq;
                        const uint32_t blk = r_nth_set256(cmask, rank);
                        src = P.dr + (((size_t)sl * P.be + blk) * P.ne + j0_k + row) * 2u;
                    } else if (it < total) {
                        const uint32_t i2 = it - nquiet;
                        row = r_nth_bit(spm, i2 / P.be);
                        rank = i2 % P.be;
                        src = P.drs + (((size_t)sl * P.ne + j0_k + row) * P.be + rank) * 2u;
                    }
                    row_of[m] = row;
                    rank_of[m] = rank;
                    src_of[m] = src;
                    if (it < total) want |= 3u << (2 * m);
                }
                auto addr = [&](int m) -> const uint64_t* { return src_of[m >> 1] + (m & 1); };
                auto sink = [&](int m, uint32_t bits) {
                    reinterpret_cast<uint32_t*>(&S.part[row_of[m >> 1]][rank_of[m >> 1]])[m & 1] = bits;
                };
                r_poll_nb<8>(P, S, want, r_tag(a), addr, sink);
            }
            // A spiking row: R = the xor butterfly over 64 lane sums, lane l summing the parts of
            // blocks l, l + 64, ... in order.  Another row: R plus its contributors' parts in
            // block order.  Then s_i = target_i / R_i (R_i > 0), published for step a + 1.
            for (uint32_t m = 0; m < ns; ++m) {
                const uint32_t r = r_nth_bit(spm, m);
                double acc = 0.0;
                for (uint32_t u = lane; u < P.be; u += 64u) acc += S.part[r][u];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
                if (lane == r) S.R[r] = acc;
            }
            if (lane < ncols_k) {
                double R = S.R[lane];
                if (!((spm >> lane) & 1u))
                    for (uint32_t u = 0; u < nc; ++u) R += S.part[lane][u];
                S.R[lane] = R;
                float sv = S.s[lane];
                if (R > 0.0) sv = (float)((double)S.target[lane] / R);
                S.s[lane] = sv;
                if (k + 1 < n)
                    st_c(P.sx + (size_t)r_slot(a + 1) * P.ne + j0_k + lane, r_word(a + 1, __float_as_uint(sv)));
            }
        }
        __syncthreads();
        if (S.abort) return;
        if (pkind != kRNone) {
            dirty_rows |= sp_now;
            dirty_cols |= own;
        }
        sp_prev = sp_now;
    }
    // ---- after the last step: the block's neurons (with the last step's stimulus events), its
    // rows' sums and scales, the rows' traces (block 0 holds every row), the plastic entries.
    const uint32_t pl = (t0 + n - 1u) & 1u;
    if (tid < ncols) {
        const uint32_t lid = j0 + tid;
        const NeuronGroupView& v = G;
        const uint32_t gid = v.offset + lid;
        float istim = S.istim[tid];
        for (uint32_t e = ev0; e < ev1; ++e)   // the events of the last step
            if (P.ev_ids[e] + P.ev_offsets[e] == gid) istim += P.ev_amps[e];
        gp(v.membrane_potentials)[lid] = S.V[tid];
        gp(v.refractory_counter)[lid] = S.refr[tid];
        gp(v.i_stimulus)[lid] = istim;
        gp(v.background_excitatory_postsynaptic_r_traces)[lid] = S.re[tid];
        gp(v.background_inhibitory_postsynaptic_r_traces)[lid] = S.ri[tid];
        gp(v.background_excitatory_postsynaptic_d_traces)[lid] = S.de[tid];
        gp(v.background_inhibitory_postsynaptic_d_traces)[lid] = S.di[tid];
        gp(v.background_g_excitatory_synaptic_conductances)[lid] = S.bge[tid];
        gp(v.background_g_inhibitory_synaptic_conductances)[lid] = S.bgi[tid];
        gp(v.g_excitatory)[lid] = S.ge[tid];
        gp(v.g_inhibitory)[lid] = S.gi[tid];
        gp(P.spikes)[gid] = S.spike[tid];
        if (epost && P.ee_kind != kRNone) {
            gp(P.ee_tr[2])[lid] = S.post[pl][0][tid];
            if (P.ee_kind == kPlastTriplet) gp(P.ee_tr[3])[lid] = S.post[pl][1][tid];
        }
        if (epost && P.ie_kind != kRNone) gp(P.ie_tr[1])[lid] = S.post[pl][2][tid];
        if (epost && P.ee_lazy) {
            gp(P.R)[lid] = S.R[tid];
            gp(P.s)[lid] = S.s[tid];
        }
    }
    if (b == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const uint32_t i = tl + T * (uint32_t)q;
            if (i >= npre) continue;
            gp(P.r[side])[i] = S.r[q][tid];
            gp(P.d[side])[i] = S.d[q][tid];
            gp(P.x[side])[i] = S.d[q][tid] - S.r[q][tid];
            if (pkind != kRNone) gp(eside ? P.ee_tr[0] : P.ie_tr[0])[i] = S.pre[q][tid];
            if (triplet) gp(P.ee_tr[1])[i] = S.tpre[q][tid];
        }
    }
    if (pkind != kRNone && (dirty_cols != 0u || dirty_rows != 0u)) {
        // A row whose neuron spiked, or every row when a column spiked: the row's segment (four
        // 16-byte stores when the block has all 16 columns).
        BN_GAS float* W = gp(P.W[conn]);
        const uint32_t ld = P.ld[conn];
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const uint32_t i = tl + T * (uint32_t)q;
            if (i >= npre || (dirty_cols == 0u && !((dirty_rows >> q) & 1u))) continue;
            BN_GAS float* dst = W + (size_t)i * ld + j0;
            if (ncols == (uint32_t)kRCols) {
#pragma unroll
                for (int c4 = 0; c4 < kRCols / 4; ++c4)
                    reinterpret_cast<BN_GAS v4f*>(dst)[c4] =
                        v4f{w[q][4 * c4], w[q][4 * c4 + 1], w[q][4 * c4 + 2], w[q][4 * c4 + 3]};
            } else {
#pragma unroll
                for (int c = 0; c < kRCols; ++c)
                    if ((uint32_t)c < ncols) dst[c] = w[q][c];
            }
        }
    }
    // The clock after the n steps (step and stimulus cursor), by block 0 at its end: every block
    // read the clock before it published its first spike word, which block 0 has seen.
    if (b == 0 && tid == 0) {
        bn_device_clock c = clk;
        c.clock_step = t0 + n;
        if (P.ev_steps) c.next_event_idx = P.ev_start[min(c.clock_step, P.ev_steps)];
        *P.clock = c;
    }
}

namespace {
const void* resident_kernel(uint32_t rows) {
    switch (rows) {
        case 1: return reinterpret_cast<const void*>(k_resident<1>);
        case 2: return reinterpret_cast<const void*>(k_resident<2>);
        case 3: return reinterpret_cast<const void*>(k_resident<3>);
        case 4: return reinterpret_cast<const void*>(k_resident<4>);
        case 5: return reinterpret_cast<const void*>(k_resident<5>);
        case 6: return reinterpret_cast<const void*>(k_resident<6>);
        case 7: return reinterpret_cast<const void*>(k_resident<7>);
        case 8: return reinterpret_cast<const void*>(k_resident<8>);
        default: return nullptr;
    }
}
}  // namespace

void resident_shape(uint32_t ne, uint32_t ni, uint32_t* rows, uint32_t* e_waves) {
    uint32_t best = 0, best_w = 0;
    for (uint32_t we = 1; we < (uint32_t)kRWaves; ++we) {
        const uint32_t ke = (ne + 64u * we - 1u) / (64u * we);
        const uint32_t ki = (ni + 64u * (kRWaves - we) - 1u) / (64u * (kRWaves - we));
        const uint32_t k = std::max(ke, ki);
        if (best == 0 || k < best) {
            best = k;
            best_w = we;
        }
    }
    *rows = best <= (uint32_t)kRMaxRows ? best : 0u;
    *e_waves = best_w;
}

int resident_max_blocks(uint32_t rows, uint32_t* blocks) {
    const void* f = resident_kernel(rows);
    if (!f) return (int)hipErrorInvalidValue;
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kRThreads, 0);
    *blocks = e == hipSuccess ? (uint32_t)std::max(0, cus * per_cu) : 0u;
    return (int)e;
}

int launch_resident(const RPlan* d_plan, const RPlan& hp, uint32_t n, bool cooperative, bn_stream_t s) {
    const void* f = resident_kernel(hp.rows);
    if (!f || n == 0) return (int)hipErrorInvalidValue;
    void* args[] = {(void*)&d_plan, (void*)&n};
    if (!cooperative) return (int)hipLaunchKernel(f, dim3(hp.blocks), dim3(kRThreads), args, 0, s);
    return (int)hipLaunchCooperativeKernel(f, dim3(hp.blocks), dim3(kRThreads), args, 0, s);
}

}  // namespace bn
