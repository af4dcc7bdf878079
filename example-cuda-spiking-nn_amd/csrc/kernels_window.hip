// Differentiable window for CDNA4 (gfx950): forward step with per-step records, and the
// reverse pass of the window loss (adjoint).  Forward per step: the propagate (the plain step's
// k_gemv, in its part-sum form where the matrices allow it) then k_win_forward (the plain
// evolve's layout: 64 neurons per block, the partial sums split over 8 waves).  Reverse per step: k_win_reverse (lambda x = W lambda g
// of the next step, one block per 4 pre rows, then the traces, rates and EIF adjoint of those
// neurons); per checkpoint segment, k_win_outer adds sum_t x_{t-1} (outer) lambda g_t to dW.
//
// Forward semantics: the plain step without plasticity or stimulus
// (kernel_neuron_group_evolve.cu:11-130, kernel_connection_propagate.cu:8-94).  Reverse: the
// chain rule of that step with the spike derivative ds/dU = 1/(beta (1 + |U - theta|/beta)^2)
// (beta = 0: spikes held fixed), the reset differentiated with the spike held fixed
// (dV/dU = 1 - s), refractory steps passing no gradient, flushed traces passing none.
#include <hip/hip_runtime.h>

#include "bn_errors.h"
#include "bn_neuron.h"
#include "bn_window.h"

namespace bn {

// A block owns 64 neurons and has kFinParts waves, each with one role whose loads all go out at
// the start: wave 0 loads the neuron's state, traces and background g and the EIF parameters of
// its group, wave 1 runs the background update, waves 2.. sum the parts of the propagate partials
// of g_e and g_i (a block straddling E|I selects each lane's sources per lane); after the
// barrier wave 0 combines the parts in order and runs the EIF step, the traces, the records.
constexpr uint32_t kWinPartWave0 = 2;
constexpr uint32_t kWinPartWaves = kFinParts - kWinPartWave0;

__global__ __launch_bounds__(kEvolveNeurons * kFinParts) void k_win_forward(WinForwardPlan P) {
    __shared__ double fin[2][kFinParts][kEvolveNeurons];
    __shared__ float bgl[2][kEvolveNeurons];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t j = blockIdx.x * kEvolveNeurons + lane;
    const bool live = j < P.N;
    const bool exc = j < P.NE;
    const WinGroup& G0 = P.g[0];
    const WinGroup& G1 = P.g[1];
    const uint32_t lid = exc ? j : j - P.NE;
    const WinState& S = P.st;
    float V = 0.0f, d0 = 0.0f, r0 = 0.0f, bg0 = 0.0f, bg1 = 0.0f;
    uint32_t refr = 0;
    float e_dt = 0.0f, e_c = 0.0f, e_slope = 0.0f, e_vt = 0.0f, e_ee = 0.0f, e_ei = 0.0f, e_vthr = 0.0f,
          e_vreset = 0.0f, e_vrest = 0.0f;
    uint32_t e_refr = 0;
    float g_leak = 0.0f, decay_e_d = 0.0f;
    int32_t vs = -1;
    if (wave == 0) {
        g_leak = P.prm->g_leak;
        decay_e_d = P.prm->decay_e_d;
        if (live) {
            vs = P.vrec_slot[j];
            V = S.V[j];
            refr = S.refr[j];
            d0 = (exc ? S.dE : S.dI)[lid];
            r0 = (exc ? S.rE : S.rI)[lid];
            bg0 = S.bg_ge[j];
            bg1 = S.bg_gi[j];
        }
        const NeuronGroupView& A = G0.v;
        const NeuronGroupView& B = G1.v;
        e_dt = exc ? uni(A.timestep) : uni(B.timestep);
        e_c = exc ? uni(A.membrane_capacitance) : uni(B.membrane_capacitance);
        e_slope = exc ? uni(A.depolarization_slope_factor) : uni(B.depolarization_slope_factor);
        e_vt = exc ? uni(A.v_effective_spike_threshold) : uni(B.v_effective_spike_threshold);
        e_ee = exc ? uni(A.v_excitatory_reversal) : uni(B.v_excitatory_reversal);
        e_ei = exc ? uni(A.v_inhibitory_reversal) : uni(B.v_inhibitory_reversal);
        e_vthr = exc ? uni(A.v_threshold) : uni(B.v_threshold);
        e_vreset = exc ? uni(A.v_reset) : uni(B.v_reset);
        e_vrest = exc ? uni(A.v_rest) : uni(B.v_rest);
        e_refr = exc ? uni(A.refractory_steps) : uni(B.refractory_steps);
    } else if (wave == 1) {
        if (live) {
            const NeuronGroupView& v = exc ? G0.v : G1.v;
            uint32_t he[4], hi[4];   // the Poisson tables' first entries, loaded with the traces
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                he[q] = q < v.poisson_table_excitatory_len ? v.poisson_table_excitatory[q] : 0u;
                hi[q] = q < v.poisson_table_inhibitory_len ? v.poisson_table_inhibitory[q] : 0u;
            }
            float re = S.bre[j], ri = S.bri[j], de = S.bde[j], di = S.bdi[j], bg_ge, bg_gi;
            background_advance(v, j, P.step, P.clock, v.poisson_table_excitatory, v.poisson_table_inhibitory, re, ri,
                               de, di, bg_ge, bg_gi, &he, &hi);
            S.bre[j] = re, S.bri[j] = ri, S.bde[j] = de, S.bdi[j] = di;
            bgl[0][lane] = bg_ge;
            bgl[1][lane] = bg_gi;
        }
    } else if (live) {
        // Parts pw and pw + kWinPartWaves of both sources, their loads in flight together.
        const uint32_t pw = wave - kWinPartWave0, pb = pw + kWinPartWaves;
        const bool has_b = pb < (uint32_t)kFinParts;
        const double* pe = (exc ? uni(G0.fin_e.partials) : uni(G1.fin_e.partials)) + lid;
        const double* pi = (exc ? uni(G0.fin_i.partials) : uni(G1.fin_i.partials)) + lid;
        const uint32_t lde = exc ? uni(G0.fin_e.ld) : uni(G1.fin_e.ld), ldi = exc ? uni(G0.fin_i.ld) : uni(G1.fin_i.ld);
        const uint32_t ce = exc ? uni(G0.fin_e.chunks) : uni(G1.fin_e.chunks);
        const uint32_t ci = exc ? uni(G0.fin_i.chunks) : uni(G1.fin_i.chunks);
        const uint32_t se = fin_part_size(ce), si = fin_part_size(ci);
        const double* const ptr[4] = {pe, pi, pe, pi};
        const uint32_t ld[4] = {lde, ldi, lde, ldi};
        const uint32_t c0[4] = {min(pw * se, ce), min(pw * si, ci), has_b ? min(pb * se, ce) : 0u,
                                has_b ? min(pb * si, ci) : 0u};
        const uint32_t c1[4] = {min(c0[0] + se, ce), min(c0[1] + si, ci), has_b ? min(c0[2] + se, ce) : 0u,
                                has_b ? min(c0[3] + si, ci) : 0u};
        double out[4];
        range_sums<4>(ptr, ld, c0, c1, out);
        if (pw * se < ce) fin[0][pw][lane] = out[0];
        if (pw * si < ci) fin[1][pw][lane] = out[1];
        if (has_b && pb * se < ce) fin[0][pb][lane] = out[2];
        if (has_b && pb * si < ci) fin[1][pb][lane] = out[3];
    }
    __syncthreads();
    if (wave != 0) return;
    uint32_t spike = 0;
    if (live) {
        // g of this step: background of the last step + the propagate partials (finalize_g's
        // order: fl32(g + part_0 + part_1 + ...) in FP64).
        float g2[2];
        for (int q = 0; q < 2; ++q) {
            const uint32_t chunks = q == 0 ? (exc ? uni(G0.fin_e.chunks) : uni(G1.fin_e.chunks))
                                           : (exc ? uni(G0.fin_i.chunks) : uni(G1.fin_i.chunks));
            const uint32_t per = fin_part_size(chunks);
            double acc = (double)(q == 0 ? bg0 : bg1);
            for (uint32_t w = 0; per && w * per < chunks; ++w) acc += fin[q][w][lane];
            g2[q] = (float)acc;
        }
        const float ge = g2[0], gi = g2[1];
        float* dT = exc ? S.dE : S.dI;
        float* rT = exc ? S.rE : S.rI;
        if (P.rec.Vprev) {
            const size_t k = (size_t)P.slot * P.N + j;
            P.rec.Vprev[k] = V;
            P.rec.ge[k] = ge;
            P.rec.gi[k] = gi;
            P.rec.refr[k] = refr > 0 ? 1 : 0;
            const uint32_t n = exc ? P.NE : P.NI;
            (exc ? P.rec.dE : P.rec.dI)[(size_t)P.slot * n + lid] = d0;
            (exc ? P.rec.rE : P.rec.rI)[(size_t)P.slot * n + lid] = r0;
        }
        eif_core(e_dt, g_leak, e_c, e_slope, e_vt, e_ee, e_ei, e_vthr, e_vreset, e_vrest, e_refr, ge, gi, 0.0f,
                 V, refr, spike);
        S.V[j] = V;
        S.refr[j] = refr;
        S.bg_ge[j] = bgl[0][lane];
        S.bg_gi[j] = bgl[1][lane];
        // Connection traces (kernel_connection_propagate.cu:8-46).
        const float ad = exc ? decay_e_d : P.decay_i_d, ar = exc ? P.decay_e_r : P.decay_i_r;
        float dd = flush_trace(d0 * ad), rr = flush_trace(r0 * ar);
        if (spike == 1) {
            dd += P.K;
            rr += P.K;
        }
        dT[lid] = dd;
        rT[lid] = rr;
        (exc ? S.xE : S.xI)[lid] = dd - rr;
        if (P.rec.Vprev) P.rec.s[(size_t)P.slot * P.N + j] = (float)spike;
        if (P.raster) P.raster[j] = (uint8_t)spike;
        if (vs >= 0 && P.v_rec) P.v_rec[vs] = V;
    }
    if (P.counts) {
        const uint64_t be = __ballot(live && spike && exc), bi = __ballot(live && spike && !exc);
        if (lane == 0) {
            if (be) atomicAdd(&P.counts[0], (uint32_t)__popcll(be));
            if (bi) atomicAdd(&P.counts[1], (uint32_t)__popcll(bi));
        }
    }
}

__global__ __launch_bounds__(256) void k_win_copy(WinCopyPlan P) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < P.N) {
        P.dst.V[j] = P.src.V[j];
        P.dst.refr[j] = P.src.refr[j];
        P.dst.bg_ge[j] = P.src.bg_ge[j];
        P.dst.bg_gi[j] = P.src.bg_gi[j];
        P.dst.bre[j] = P.src.bre[j];
        P.dst.bde[j] = P.src.bde[j];
        P.dst.bri[j] = P.src.bri[j];
        P.dst.bdi[j] = P.src.bdi[j];
    }
    if (j < P.NE) {
        P.dst.dE[j] = P.src.dE[j];
        P.dst.rE[j] = P.src.rE[j];
        P.dst.xE[j] = P.src.xE[j];
    }
    if (j < P.NI) {
        P.dst.dI[j] = P.src.dI[j];
        P.dst.rI[j] = P.src.rI[j];
        P.dst.xI[j] = P.src.xI[j];
    }
}

// What the reverse step of neuron j reads besides lambda x_j, its accumulators included (loaded
// before the row product so that their latency overlaps it; only this thread writes them).
struct ReverseInputs {
    float d0, r0, Vp, ge, gi, s;
    float ad, gl;                     // the group's d decay, g_leak (device parameters)
    uint32_t refr;
    double carry_d, carry_r, carry_V, dv, lrho, acc_alpha, acc_gleak;
};

__device__ __forceinline__ ReverseInputs load_reverse_inputs(const WinReversePlan& P, uint32_t j) {
    ReverseInputs in;
    const bool exc = j < P.NE;
    const uint32_t lid = exc ? j : j - P.NE;
    const size_t k = (size_t)P.slot * P.N + j;
    const uint32_t n = exc ? P.NE : P.NI;
    in.d0 = (exc ? P.rec.dE : P.rec.dI)[(size_t)P.slot * n + lid];
    in.r0 = (exc ? P.rec.rE : P.rec.rI)[(size_t)P.slot * n + lid];
    in.ad = exc ? P.prm->decay_e_d : P.decay_i_d;
    in.gl = P.prm->g_leak;
    in.lrho = P.lrho[exc ? 0 : 1];
    in.acc_alpha = exc ? P.acc_alpha[j] : 0.0;
    in.acc_gleak = P.acc_gleak[j];
    in.carry_d = P.carry_d[j];
    in.carry_r = P.carry_r[j];
    in.carry_V = P.carry_V[j];
    const int32_t vs = P.vrec_slot[j];
    in.dv = vs >= 0 ? P.dv_direct[vs] : 0.0;
    in.refr = P.rec.refr[k];
    in.Vp = P.rec.Vprev[k];
    in.ge = P.rec.ge[k];
    in.gi = P.rec.gi[k];
    in.s = P.rec.s[k];
    return in;
}

// Per neuron j at one step, given lambda x_j: the adjoint of its connection traces (as a pre
// neuron), then of its EIF update (as a post neuron).
__device__ __forceinline__ void neuron_reverse(const WinReversePlan& P, uint32_t j, double lx, const ReverseInputs& in) {
    const bool exc = j < P.NE;
    const size_t k = (size_t)P.slot * P.N + j;
    // Traces: d_t = flush(d_{t-1} a_d) + K s_t, r likewise, x_t = d_t - r_t.
    const float d0 = in.d0, r0 = in.r0;
    const float ad = in.ad, ar = exc ? P.decay_e_r : P.decay_i_r;
    const double ld = lx + in.carry_d;
    const double lr = -lx + in.carry_r;
    const bool nd = !(d0 * ad < kTraceFlush), nr = !(r0 * ar < kTraceFlush);
    if (exc && nd) P.acc_alpha[j] = in.acc_alpha + ld * (double)d0;
    P.carry_d[j] = nd ? (double)ad * ld : 0.0;
    P.carry_r[j] = nr ? (double)ar * lr : 0.0;
    // Spike: traces and the smoothed population rate.
    const double ls = (double)P.K * (ld + lr) + in.lrho * (exc ? P.rate_coef_e : P.rate_coef_i);
    const double lv = in.carry_V + in.dv;
    if (in.refr) {
        P.carry_V[j] = 0.0;
        P.lg_out_e[j] = 0.0;
        P.lg_out_i[j] = 0.0;
        P.rec.le[k] = 0.0f;
        P.rec.li[k] = 0.0f;
        return;
    }
    const NeuronGroupView& v = P.g[exc ? 0 : 1].v;
    const double Vp = in.Vp, ge = in.ge, gi = in.gi, s = in.s;
    // U of the step in FP64 from its recorded FP32 inputs (the spike itself is the record's).
    const double DT = v.depolarization_slope_factor, gl = in.gl;
    const double ex = exp((Vp - (double)v.v_effective_spike_threshold) / DT);
    const double F = -gl * (Vp - (double)v.v_rest) + gl * DT * ex - ge * (Vp - (double)v.v_excitatory_reversal) -
                     gi * (Vp - (double)v.v_inhibitory_reversal);
    const double kd = (double)v.timestep / (double)v.membrane_capacitance;
    const double U = Vp + kd * F;
    double sg = 0.0;
    if (P.beta > 0.0f) {
        const double z = 1.0 + fabs(U - (double)P.theta) / (double)P.beta;
        sg = 1.0 / ((double)P.beta * z * z);
    }
    const double lU = lv * (1.0 - s) + ls * sg;
    P.carry_V[j] = lU * (1.0 + kd * (-gl + gl * ex - ge - gi));
    const double le = -lU * kd * (Vp - (double)v.v_excitatory_reversal);
    const double li = -lU * kd * (Vp - (double)v.v_inhibitory_reversal);
    P.lg_out_e[j] = le;
    P.lg_out_i[j] = li;
    P.rec.le[k] = (float)le;
    P.rec.li[k] = (float)li;
    P.acc_gleak[j] = in.acc_gleak + lU * kd * (-(Vp - (double)v.v_rest) + DT * ex);
}

// One block per kRevRows neurons of one group: lambda x_j = sum_k W[j][k] lambda g_k over the
// two matrices leaving j (E rows: EE and EI with lambda g_e; I rows: IE and II with lambda g_i).
// The block's kRevSplit waves take interleaved float4 columns of the rows (each lambda g value
// loaded once for the block's rows), so that about 3,000 waves keep the weight reads in flight;
// FP64 sums per lane, the lane butterfly, then the waves' sums in wave order; lane r of wave 0
// runs the reverse step of the block's row r.
constexpr uint32_t kRevRows = 4, kRevSplit = 4;
__global__ __launch_bounds__(64 * kRevSplit) void k_win_reverse(WinReversePlan P) {
    __shared__ double part[kRevSplit][kRevRows];
    const uint32_t blk = blockIdx.x;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t blk_e = (P.NE + kRevRows - 1) / kRevRows, blk_i = (P.NI + kRevRows - 1) / kRevRows;
    if (blk >= blk_e + blk_i) return;   // whole blocks
    const bool exc = blk < blk_e;
    const uint32_t l0 = (exc ? blk : blk - blk_e) * kRevRows;       // first group-local row
    const uint32_t nrows = min(kRevRows, (exc ? P.NE : P.NI) - l0);
    const uint32_t g0 = exc ? l0 : P.NE + l0;                       // first global id
    const double* lg = exc ? P.lg_in_e : P.lg_in_i;
    ReverseInputs in{};
    if (wave == 0 && lane < nrows) in = load_reverse_inputs(P, g0 + lane);
    double acc[kRevRows] = {};
    for (int h = 0; h < 2; ++h) {
        const int c = (exc ? 0 : 2) + h;   // to E posts, then to I posts
        const uint32_t n_post = h == 0 ? P.NE : P.NI;
        const double* l = lg + (h == 0 ? 0 : P.NE);
        const uint32_t n4 = n_post >> 2;
        const float* wr[kRevRows];
#pragma unroll
        for (uint32_t r = 0; r < kRevRows; ++r) wr[r] = P.W[c] + (size_t)(l0 + min(r, nrows - 1)) * P.ld[c];
#pragma unroll 2
        for (uint32_t q = wave * 64 + lane; q < n4; q += 64 * kRevSplit) {
            const double l_0 = l[4 * q], l_1 = l[4 * q + 1], l_2 = l[4 * q + 2], l_3 = l[4 * q + 3];
#pragma unroll
            for (uint32_t r = 0; r < kRevRows; ++r) {
                const float4 wv = reinterpret_cast<const float4*>(wr[r])[q];
                acc[r] += (double)wv.x * l_0;
                acc[r] += (double)wv.y * l_1;
                acc[r] += (double)wv.z * l_2;
                acc[r] += (double)wv.w * l_3;
            }
        }
        for (uint32_t jj = 4 * n4 + wave * 64 + lane; jj < n_post; jj += 64 * kRevSplit)
#pragma unroll
            for (uint32_t r = 0; r < kRevRows; ++r) acc[r] += (double)wr[r][jj] * l[jj];
    }
#pragma unroll
    for (uint32_t r = 0; r < kRevRows; ++r)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc[r] += __shfl_xor(acc[r], off, 64);
    if (lane == 0)
#pragma unroll
        for (uint32_t r = 0; r < kRevRows; ++r) part[wave][r] = acc[r];
    __syncthreads();
    if (wave != 0 || lane >= nrows) return;
    double mine = part[0][lane];
#pragma unroll
    for (uint32_t v = 1; v < kRevSplit; ++v) mine += part[v][lane];
    neuron_reverse(P, g0 + lane, mine, in);
}

// dW[i][j] += sum_slots x_slot[i] * lambda_slot[j], x = d - r at the start of the step: the
// segment's outer products as one GEMM, dW += X^T L (X: slots x pre, L: slots x post), on the
// FP32 matrix cores.  A block owns a 64 x 64 tile of dW and stages 16 slots of its X and L
// columns in LDS at a time; wave q accumulates the 32 x 32 quarter (rows 32 * (q >> 1), columns
// 32 * (q & 1)) with v_mfma_f32_32x32x2f32, two slots per operation.  Operand lanes:
// A[m][k] = X[slot k][row m] with m = lane % 32, k = lane / 32; B[k][n] = L[slot k][column n]
// with n = lane % 32.  Result: acc[r] is row 8 * (r / 4) + 4 * (lane / 32) + r % 4, column
// lane % 32.  FP32 sums (the slots past the segment are staged as zeros).
typedef float v16f_w __attribute__((ext_vector_type(16)));
__global__ __launch_bounds__(256) void k_win_outer(WinOuterPlan P) {
    __shared__ float xs[16][64];
    __shared__ float ls[16][64];
    const uint32_t ti = blockIdx.y * 64, tj = blockIdx.x * 64;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t mi = 32u * (wave >> 1) + (lane & 31), nj = 32u * (wave & 1) + (lane & 31), kk = lane >> 5;
    v16f_w acc = {};
    for (uint32_t s0 = 0; s0 < P.slots; s0 += 16) {
        for (uint32_t e = threadIdx.x; e < 16 * 64; e += 256) {
            const uint32_t sl = e / 64, c = e % 64;
            const uint32_t slot = s0 + sl;
            float xv = 0.0f, lv = 0.0f;
            if (slot < P.slots) {
                if (ti + c < P.n_pre) {
                    const size_t q = (size_t)slot * P.pre_stride + P.pre_offset + ti + c;
                    xv = P.d[q] - P.r[q];
                }
                if (tj + c < P.n_post) lv = P.lam[(size_t)slot * P.lam_stride + P.post_offset + tj + c];
            }
            xs[sl][c] = xv;
            ls[sl][c] = lv;
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < 8; ++p)
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(xs[2 * p + kk][mi], ls[2 * p + kk][nj], acc, 0, 0, 0);
        __syncthreads();
    }
    const uint32_t j = tj + nj;
    if (j >= P.n_post) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const uint32_t i = ti + 32u * (wave >> 1) + 8u * (r / 4) + 4u * kk + (r % 4);
        if (i < P.n_pre) P.dW[(size_t)i * P.ld + j] += acc[r];
    }
}

int launch_win_forward(const WinForwardPlan& p, bn_stream_t s) {
    hipLaunchKernelGGL(k_win_forward, dim3((p.N + kEvolveNeurons - 1) / kEvolveNeurons), dim3(kEvolveNeurons * kFinParts),
                       0, s, p);
    return (int)hipGetLastError();
}

int launch_win_copy(const WinCopyPlan& p, bn_stream_t s) {
    hipLaunchKernelGGL(k_win_copy, dim3((p.N + 255) / 256), dim3(256), 0, s, p);
    return (int)hipGetLastError();
}

int launch_win_reverse(const WinReversePlan& p, bn_stream_t s) {
    const uint32_t blocks = (p.NE + kRevRows - 1) / kRevRows + (p.NI + kRevRows - 1) / kRevRows;
    hipLaunchKernelGGL(k_win_reverse, dim3(blocks), dim3(64 * kRevSplit), 0, s, p);
    return (int)hipGetLastError();
}

int launch_win_outer(const WinOuterPlan& p, bn_stream_t s) {
    if (p.slots == 0) return 0;
    hipLaunchKernelGGL(k_win_outer, dim3((p.n_post + 63) / 64, (p.n_pre + 63) / 64), dim3(256), 0, s, p);
    return (int)hipGetLastError();
}

}  // namespace bn
