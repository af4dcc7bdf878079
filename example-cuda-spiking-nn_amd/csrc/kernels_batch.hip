// Batched shared-weight step for CDNA4 (gfx950): B instances of one network with frozen
// weights.  Per step two launches:
//   k_batch_evolve  every (neuron, instance): partial sums -> g, EIF step, background Poisson
//                   noise (Philox keyed per instance), connection traces, stimulus pulse;
//   k_batch_gemm    for every connection G[post][B] = W^T[post][pre] * X[pre][B] on the FP32
//                   matrix cores (v_mfma_f32_32x32x2f32), FP32 accumulation.
// Per-neuron semantics follow the same reference lines as the plain step
// (kernel_neuron_group_evolve.cu:11-130, kernel_connection_propagate.cu:8-46,79-94,
// stimulus.cpp:46-60); only the propagate's summation differs (FP32 matrix-core order).
#include <hip/hip_runtime.h>

#include "bn_batch.h"
#include "bn_device.h"
#include "bn_errors.h"

namespace bn {

// kernel_neuron_group_evolve.cu:85-123 (same arithmetic as the plain step's eif_step).
__device__ __forceinline__ void batch_eif(const NeuronGroupView& v, float ge, float gi, float istim, float& V,
                                          uint32_t& refr, uint32_t& spike) {
    if (refr > 0) {
        refr = refr - 1;
        spike = 0;
        V = v.v_reset;
        return;
    }
    const float vv = V;
    const float linear_leak_term = -1.0f * v.g_leak * (vv - v.v_rest);
    const float exponential_depolarization_term =
        v.g_leak * v.depolarization_slope_factor *
        bn_expf((vv - v.v_effective_spike_threshold) / v.depolarization_slope_factor);
    const float excitatory_synaptic_term = -1.0f * ge * (vv - v.v_excitatory_reversal);
    const float inhibitory_synaptic_term = -1.0f * gi * (vv - v.v_inhibitory_reversal);
    const float dv = (v.timestep / v.membrane_capacitance) *
                     (linear_leak_term + exponential_depolarization_term + excitatory_synaptic_term +
                      inhibitory_synaptic_term + istim);
    const float v_next = vv + dv;
    if (v_next >= v.v_threshold) {
        refr = v.refractory_steps;
        spike = 1;
        V = v.v_threshold;
    } else {
        spike = 0;
        V = v_next;
    }
}

// One thread per (neuron n, instance b); b is the fast index, so a wave is one neuron.
__global__ __launch_bounds__(256) void k_batch_evolve(BatchEvolvePlan P) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = __builtin_amdgcn_readfirstlane(t / P.B);
    const uint32_t b = t - n * P.B;
    if (n >= P.N) return;
    const uint32_t step = *P.step;
    const size_t k = (size_t)n * P.B + b;
    uint32_t spike = 0;
    // The connection traces of this neuron in every set, loaded with the state below (the
    // stores of the state would otherwise hold these loads back).
    float tr_d[kBatchMaxGroups], tr_r[kBatchMaxGroups];
#pragma unroll
    for (int s = 0; s < kBatchMaxGroups; ++s) {
        const BatchTraceSet& T = P.ts[s];
        tr_d[s] = tr_r[s] = 0.0f;
        if ((uint32_t)s < P.n_ts && n >= T.pre_offset && n < T.pre_offset + T.n_pre) {
            const size_t kk = (size_t)(n - T.pre_offset) * P.B + b;
            tr_d[s] = T.d[kk];
            tr_r[s] = T.r[kk];
        }
    }
    for (uint32_t g = 0; g < P.n_groups; ++g) {
        const BatchGroup& G = P.g[g];
        const NeuronGroupView& v = G.v;
        if (n < v.offset || n >= v.offset + v.num_neurons) continue;
        const uint32_t lid = n - v.offset;
        // Issue every load first.
        float ge = P.ge[k], gi = P.gi[k], istim = P.istim[k], V = P.V[k];
        uint32_t refr = P.refr[k];
        float re = P.bre[k], de = P.bde[k], ri = P.bri[k], di = P.bdi[k];
        const float amp = P.stim_amp ? P.stim_amp[k] : 0.0f;
        uint32_t he[4], hi[4];  // the first Poisson table entries (wave-uniform)
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            he[q] = q < v.poisson_table_excitatory_len ? v.poisson_table_excitatory[q] : 0u;
            hi[q] = q < v.poisson_table_inhibitory_len ? v.poisson_table_inhibitory[q] : 0u;
        }
        float pe[8], pi[8];
        const uint32_t ne = G.has_fin_e ? min(G.fin_e.groups, 8u) : 0u;
        const uint32_t ni = G.has_fin_i ? min(G.fin_i.groups, 8u) : 0u;
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) {
            pe[q] = q < ne ? G.fin_e.part[((size_t)q * v.num_neurons + lid) * P.B + b] : 0.0f;
            pi[q] = q < ni ? G.fin_i.part[((size_t)q * v.num_neurons + lid) * P.B + b] : 0.0f;
        }
        // g = background of the last step + the propagate partials in group order.
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) {
            if (q < ne) ge += pe[q];
            if (q < ni) gi += pi[q];
        }
        for (uint32_t q = 8; G.has_fin_e && q < G.fin_e.groups; ++q)
            ge += G.fin_e.part[((size_t)q * v.num_neurons + lid) * P.B + b];
        for (uint32_t q = 8; G.has_fin_i && q < G.fin_i.groups; ++q)
            gi += G.fin_i.part[((size_t)q * v.num_neurons + lid) * P.B + b];
        batch_eif(v, ge, gi, istim, V, refr, spike);
        // This step's stimulus phase (stimulus.cpp:46-60), seen by the next evolve: +amp at
        // stim_on, -amp at stim_off.
        if (amp != 0.0f) {
            if (step == P.stim_on[b]) istim += amp;
            if (step == P.stim_off[b]) istim += -amp;
        }
        // Background (kernel_neuron_group_evolve.cu:11-83), Philox keyed by the instance.
        re = flush_trace(re * v.background_excitatory_postsynaptic_r_trace_decay);
        ri = flush_trace(ri * v.background_inhibitory_postsynaptic_r_trace_decay);
        de = flush_trace(de * v.background_excitatory_postsynaptic_d_trace_decay);
        di = flush_trace(di * v.background_inhibitory_postsynaptic_d_trace_decay);
        uint32_t ke = 0, ki = 0;
        if (v.poisson_table_excitatory_len > 0 || v.poisson_table_inhibitory_len > 0) {
            const U32x4 r = philox4x32_10(U32x4{step, (uint32_t)BN_RNG_STREAM_BACKGROUND, n, v.rng_instance},
                                          P.keys[2 * b], P.keys[2 * b + 1]);
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
        P.V[k] = V;
        P.refr[k] = refr;
        P.spikes[k] = spike;
        P.istim[k] = istim;
        P.bre[k] = re;
        P.bde[k] = de;
        P.bri[k] = ri;
        P.bdi[k] = di;
        P.ge[k] = bg_ge;
        P.gi[k] = bg_gi;
    }
    // Connection traces (kernel_connection_propagate.cu:8-46) of this neuron in every set.
#pragma unroll
    for (int s = 0; s < kBatchMaxGroups; ++s) {
        const BatchTraceSet& T = P.ts[s];
        if ((uint32_t)s >= P.n_ts || n < T.pre_offset || n >= T.pre_offset + T.n_pre) continue;
        const size_t kk = (size_t)(n - T.pre_offset) * P.B + b;
        float dd = flush_trace(tr_d[s] * T.decay_d);
        float rr = flush_trace(tr_r[s] * T.decay_r);
        if (spike == 1) {
            rr += T.K;
            dd += T.K;
        }
        T.d[kk] = dd;
        T.r[kk] = rr;
        T.x[kk] = dd - rr;
    }
}

// GEMM tile of one wave: 64 post neurons x 64 instances (four 32x32 accumulators: two post
// halves x two instance halves) over a chunk of kc pre rows, two rows per
// v_mfma_f32_32x32x2f32; each W load feeds two operations and so does each X load.  Operand
// lanes: A[m][k] = W[k][m] with m = lane % 32 and k = lane / 32; B[k][n] = X[k][n] with
// n = lane % 32.  Result lane layout: acc[i] is row m = 8*(i/4) + 4*(lane/32) + i%4, column
// lane % 32.
typedef float v16f __attribute__((ext_vector_type(16)));
// Word 3 of a raw buffer resource on CDNA3/CDNA4 (32-bit data format, no swizzle, no stride).
constexpr int kBufferRsrcWord3 = 0x00020000;

__global__ __launch_bounds__(kBatchGemmWaves * 64) void k_batch_gemm(BatchGemmPlan P) {
    __shared__ float red[kBatchGemmWaves][4][16][64];
    const uint32_t blk = blockIdx.x;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t ci = 0;
    while (ci + 1 < P.n_conn && blk >= P.c[ci].block_end) ++ci;
    const BatchGemmConn& C = P.c[ci];
    // Block -> (group, m tile, n pair); consecutive blocks walk the m tiles of one row range.
    const uint32_t t = blk - C.block_begin;
    const uint32_t np = t % C.n_pairs;
    const uint32_t mt = (t / C.n_pairs) % C.m_tiles;
    const uint32_t grp = t / (C.n_pairs * C.m_tiles);
    const uint32_t chunk = grp * kBatchGemmWaves + wave;
    const uint32_t m0 = mt * kBatchGemmTileM, n0 = np * 64;
    const uint32_t B = P.B;
    const uint32_t ml = m0 + (lane & 31);  // ml + 32 < ld: the padded tail of a row holds zeros
    const uint32_t kl = lane >> 5;
    if (chunk < C.n_chunks) {
        v16f acc00 = {}, acc01 = {}, acc10 = {}, acc11 = {};  // acc[post half][instance half]
        const uint32_t i0 = chunk * C.kc;
        const uint32_t i1 = min(i0 + C.kc, C.n_pre);
        const uint32_t last = i1 - 1;
        // Buffer loads: a scalar resource per matrix (base and byte size, both wave-uniform) and
        // one 32-bit byte offset per load, the +32 column in the instruction's immediate offset
        // (weights 25 MB, traces 0.8 MB at 3,125 neurons: far below 4 GiB).
        const __amdgpu_buffer_rsrc_t wr =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(C.W), 0, (int)(C.n_pre * C.ld * 4u), kBufferRsrcWord3);
        const __amdgpu_buffer_rsrc_t xr =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(C.X), 0, (int)(C.n_pre * B * 4u), kBufferRsrcWord3);
        const uint32_t ldw = C.ld, xcol = n0 + (lane & 31);
        // Batches of kGemmPairs row pairs, double-buffered: the loads of the next batch are in
        // flight while the matrix cores work on this one.  Every load is issued (rows past the
        // chunk are clamped to its last row and their W operands are zeroed), so the in-order
        // load counter lets each batch wait only for its own loads.
        constexpr int kGemmPairs = 8;
        constexpr uint32_t kRows = 2 * kGemmPairs;
        typedef float Buf[4][kGemmPairs];  // W low half, W high half, X low half, X high half
        Buf ba, bb;
        auto load = [&](uint32_t base, Buf& b) {
#pragma unroll
            for (int u = 0; u < kGemmPairs; ++u) {
                const uint32_t row = min(base + 2 * u + kl, last);
                const uint32_t wo = (row * ldw + ml) * 4u, xo = (row * B + xcol) * 4u;
                b[0][u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(wr, wo, 0, 0));
                b[1][u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(wr, wo + 128u, 0, 0));
                b[2][u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(xr, xo, 0, 0));
                b[3][u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(xr, xo + 128u, 0, 0));
            }
        };
        auto mma = [&](uint32_t base, const Buf& b) {
#pragma unroll
            for (int u = 0; u < kGemmPairs; ++u) {
                const bool ok = base + 2 * u + kl <= last;
                const float w0 = ok ? b[0][u] : 0.0f, w1 = ok ? b[1][u] : 0.0f;
                acc00 = __builtin_amdgcn_mfma_f32_32x32x2f32(w0, b[2][u], acc00, 0, 0, 0);
                acc01 = __builtin_amdgcn_mfma_f32_32x32x2f32(w0, b[3][u], acc01, 0, 0, 0);
                acc10 = __builtin_amdgcn_mfma_f32_32x32x2f32(w1, b[2][u], acc10, 0, 0, 0);
                acc11 = __builtin_amdgcn_mfma_f32_32x32x2f32(w1, b[3][u], acc11, 0, 0, 0);
            }
        };
        // Pairs of batches and one loop exit (an odd count gets a trailing batch whose W
        // operands are all zero: it adds exact zeros).  The scheduling barriers keep the
        // compiler from sinking a prefetch below the MFMAs.
        const uint32_t npairs = ((i1 - i0 + kRows - 1) / kRows + 1) / 2;
        uint32_t i = i0;
        load(i, ba);
        for (uint32_t p = 0; p < npairs; ++p) {
            load(min(i + kRows, last), bb);
            __builtin_amdgcn_sched_barrier(0);
            mma(i, ba);
            i += kRows;
            load(min(i + kRows, last), ba);
            __builtin_amdgcn_sched_barrier(0);
            mma(i, bb);
            i += kRows;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            red[wave][0][r][lane] = acc00[r];
            red[wave][1][r][lane] = acc01[r];
            red[wave][2][r][lane] = acc10[r];
            red[wave][3][r][lane] = acc11[r];
        }
    }
    // (A wave past the last chunk stores nothing: its slots are past nw below.)
    __syncthreads();
    // Wave q adds the chunks of the block for accumulator q, in chunk order, and writes one
    // partial per group.
    static_assert(kBatchGemmWaves == 4, "one wave per accumulator in the block reduction");
    {
        const uint32_t q = wave;
        const uint32_t nw = min((uint32_t)kBatchGemmWaves, C.n_chunks - grp * kBatchGemmWaves);
        float* out = C.part + (size_t)grp * C.n_post * B;
        const uint32_t mh = q >> 1, h = q & 1;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float s = red[0][q][r][lane];
            for (uint32_t w = 1; w < nw; ++w) s += red[w][q][r][lane];
            const uint32_t m = m0 + 32 * mh + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
            if (m < C.n_post) out[(size_t)m * B + n0 + 32 * h + (lane & 31)] = s;
        }
    }
    // The last block advances the step counter (no block of this launch reads it).
    if (blk == gridDim.x - 1 && threadIdx.x == 0) *P.step = *P.step + 1;
}

int launch_batch_evolve(const BatchEvolvePlan& plan, bn_stream_t s) {
    const uint64_t threads = (uint64_t)plan.N * plan.B;
    hipLaunchKernelGGL(k_batch_evolve, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, plan);
    return (int)hipGetLastError();
}

int launch_batch_gemm(const BatchGemmPlan& plan, bn_stream_t s) {
    if (plan.total_blocks == 0) return 0;
    hipLaunchKernelGGL(k_batch_gemm, dim3(plan.total_blocks), dim3(kBatchGemmWaves * 64), 0, s, plan);
    return (int)hipGetLastError();
}

}  // namespace bn
