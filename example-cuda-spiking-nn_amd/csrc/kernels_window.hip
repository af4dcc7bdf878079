// Differentiable window for CDNA4 (gfx950): forward step with per-step records, and the
// reverse pass of the window loss (adjoint).  Forward per step: the propagate (k_gemv, shared
// with the plain step) then k_win_forward.  Reverse per step: k_win_rows (lambda x = W lambda g
// of the next step, one wave per pre row) then k_win_backward (traces, rates, EIF adjoint per
// neuron); per checkpoint segment, k_win_outer adds sum_t x_{t-1} (outer) lambda g_t to dW.
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

__global__ __launch_bounds__(256) void k_win_forward(WinForwardPlan P) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    const bool exc = j < P.NE;
    const WinGroup& G = P.g[exc ? 0 : 1];
    const NeuronGroupView& v = G.v;
    const uint32_t lid = exc ? j : j - P.NE;
    const WinState& S = P.st;
    // g of this step: background of the last step + the propagate partials of the last step.
    const float ge = finalize_g(S.bg_ge[j], G.fin_e, lid);
    const float gi = finalize_g(S.bg_gi[j], G.fin_i, lid);
    float V = S.V[j];
    uint32_t refr = S.refr[j];
    float* dT = exc ? S.dE : S.dI;
    float* rT = exc ? S.rE : S.rI;
    const float d0 = dT[lid], r0 = rT[lid];
    if (P.rec.Vprev) {
        const size_t k = (size_t)P.slot * P.N + j;
        P.rec.Vprev[k] = V;
        P.rec.ge[k] = ge;
        P.rec.gi[k] = gi;
        P.rec.refr[k] = refr > 0 ? 1 : 0;
        float* rd = exc ? P.rec.dE : P.rec.dI;
        float* rr = exc ? P.rec.rE : P.rec.rI;
        const uint32_t n = exc ? P.NE : P.NI;
        rd[(size_t)P.slot * n + lid] = d0;
        rr[(size_t)P.slot * n + lid] = r0;
    }
    uint32_t spike = 0;
    eif_step(v, ge, gi, 0.0f, V, refr, spike);
    float re = S.bre[j], ri = S.bri[j], de = S.bde[j], di = S.bdi[j], bg_ge, bg_gi;
    background_advance(v, j, P.step, P.clock, v.poisson_table_excitatory, v.poisson_table_inhibitory, re, ri, de, di,
                       bg_ge, bg_gi);
    S.V[j] = V;
    S.refr[j] = refr;
    S.bre[j] = re, S.bri[j] = ri, S.bde[j] = de, S.bdi[j] = di;
    S.bg_ge[j] = bg_ge;
    S.bg_gi[j] = bg_gi;
    // Connection traces (kernel_connection_propagate.cu:8-46).
    const float ad = exc ? P.decay_e_d : P.decay_i_d, ar = exc ? P.decay_e_r : P.decay_i_r;
    float dd = flush_trace(d0 * ad), rr = flush_trace(r0 * ar);
    if (spike == 1) {
        dd += P.K;
        rr += P.K;
    }
    dT[lid] = dd;
    rT[lid] = rr;
    (exc ? S.xE : S.xI)[lid] = dd - rr;
    if (P.rec.Vprev) P.rec.s[(size_t)P.slot * P.N + j] = (float)spike;
    if (spike) atomicAdd(&P.counts[exc ? 0 : 1], 1u);
    if (P.raster) P.raster[j] = (uint8_t)spike;
    const int32_t vs = P.vrec_slot[j];
    if (vs >= 0) P.v_rec[vs] = V;
}

// lambda x_i = sum_j W[i][j] lambda g_j over the two matrices leaving pre neuron i (E rows: EE
// and EI with lambda g_e; I rows: IE and II with lambda g_i); one wave per row, FP64 sums.
__global__ __launch_bounds__(256) void k_win_rows(WinRowPlan P) {
    const uint32_t row = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t N = P.NE + P.NI;
    if (row >= N) return;
    const bool exc = row < P.NE;
    const uint32_t il = exc ? row : row - P.NE;
    const float* lg = exc ? P.lg_e : P.lg_i;
    double acc = 0.0;
    for (int h = 0; h < 2; ++h) {
        const int c = (exc ? 0 : 2) + h;   // to E posts, then to I posts
        const uint32_t n_post = h == 0 ? P.NE : P.NI;
        const float* w = P.W[c] + (size_t)il * P.ld[c];
        const float* l = lg + (h == 0 ? 0 : P.NE);
        for (uint32_t j = lane; j < n_post; j += 64) acc += (double)w[j] * (double)l[j];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) P.lx[row] = acc;
}

// Per neuron j at one step: the adjoint of its connection traces (as a pre neuron), then of its
// EIF update (as a post neuron).
__global__ __launch_bounds__(256) void k_win_backward(WinBackwardPlan P) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.N) return;
    const bool exc = j < P.NE;
    const uint32_t lid = exc ? j : j - P.NE;
    const size_t k = (size_t)P.slot * P.N + j;
    // Traces: d_t = flush(d_{t-1} a_d) + K s_t, r likewise, x_t = d_t - r_t.
    const uint32_t n = exc ? P.NE : P.NI;
    const float d0 = (exc ? P.rec.dE : P.rec.dI)[(size_t)P.slot * n + lid];
    const float r0 = (exc ? P.rec.rE : P.rec.rI)[(size_t)P.slot * n + lid];
    const float ad = exc ? P.decay_e_d : P.decay_i_d, ar = exc ? P.decay_e_r : P.decay_i_r;
    const double lx = P.lx[j];
    const double ld = lx + (double)P.carry_d[j];
    const double lr = -lx + (double)P.carry_r[j];
    const bool nd = !(d0 * ad < kTraceFlush), nr = !(r0 * ar < kTraceFlush);
    if (exc && nd) P.acc_alpha[j] += ld * (double)d0;
    P.carry_d[j] = nd ? (float)((double)ad * ld) : 0.0f;
    P.carry_r[j] = nr ? (float)((double)ar * lr) : 0.0f;
    // Spike: traces and the smoothed population rate.
    const double ls = (double)P.K * (ld + lr) + (exc ? (double)P.lrho_e * P.rate_coef_e : (double)P.lrho_i * P.rate_coef_i);
    double lv = (double)P.carry_V[j];
    const int32_t vs = P.vrec_slot[j];
    if (vs >= 0) lv += (double)P.dv_direct[vs];
    if (P.rec.refr[k]) {
        P.carry_V[j] = 0.0f;
        P.lg_e[j] = 0.0f;
        P.lg_i[j] = 0.0f;
        P.rec.le[k] = 0.0f;
        P.rec.li[k] = 0.0f;
        return;
    }
    const NeuronGroupView& v = P.g[exc ? 0 : 1].v;
    const float Vp = P.rec.Vprev[k], ge = P.rec.ge[k], gi = P.rec.gi[k], s = P.rec.s[k];
    // U of the forward step, recomputed with the forward's arithmetic.
    const float ex = bn_expf((Vp - v.v_effective_spike_threshold) / v.depolarization_slope_factor);
    const float F = -1.0f * v.g_leak * (Vp - v.v_rest) + v.g_leak * v.depolarization_slope_factor * ex -
                    1.0f * ge * (Vp - v.v_excitatory_reversal) - 1.0f * gi * (Vp - v.v_inhibitory_reversal);
    const float kappa = v.timestep / v.membrane_capacitance;
    const float U = Vp + kappa * F;
    double sg = 0.0;
    if (P.beta > 0.0f) {
        const double z = 1.0 + fabs((double)U - (double)P.theta) / (double)P.beta;
        sg = 1.0 / ((double)P.beta * z * z);
    }
    const double lU = lv * (1.0 - (double)s) + ls * sg;
    const double kd = (double)kappa, exd = (double)ex, gl = (double)v.g_leak;
    P.carry_V[j] = (float)(lU * (1.0 + kd * (-gl + gl * exd - (double)ge - (double)gi)));
    const float le = (float)(-lU * kd * ((double)Vp - (double)v.v_excitatory_reversal));
    const float li = (float)(-lU * kd * ((double)Vp - (double)v.v_inhibitory_reversal));
    P.lg_e[j] = le;
    P.lg_i[j] = li;
    P.rec.le[k] = le;
    P.rec.li[k] = li;
    P.acc_gleak[j] += lU * kd * (-((double)Vp - (double)v.v_rest) + (double)v.depolarization_slope_factor * exd);
}

// dW[i][j] += sum_slots x_slot[i] * lambda_slot[j], x = d - r at the start of the step.  A block
// owns a 64 x 64 tile; each thread 4 x 4 entries; FP32 sums over the segment's slots.
__global__ __launch_bounds__(256) void k_win_outer(WinOuterPlan P) {
    __shared__ float xs[16][64];
    __shared__ float ls[16][64];
    const uint32_t ti = blockIdx.y * 64, tj = blockIdx.x * 64;
    const uint32_t tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    float acc[4][4] = {};
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
        for (int sl = 0; sl < 16; ++sl) {
            float a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = xs[sl][ty * 4 + u];
                b[u] = ls[sl][tx * 4 + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int w = 0; w < 4; ++w) acc[u][w] += a[u] * b[w];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const uint32_t i = ti + ty * 4 + u;
        if (i >= P.n_pre) continue;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t jj = tj + tx * 4 + w;
            if (jj < P.n_post) P.dW[(size_t)i * P.ld + jj] += acc[u][w];
        }
    }
}

int launch_win_forward(const WinForwardPlan& p, bn_stream_t s) {
    hipLaunchKernelGGL(k_win_forward, dim3((p.N + 255) / 256), dim3(256), 0, s, p);
    return (int)hipGetLastError();
}

int launch_win_rows(const WinRowPlan& p, bn_stream_t s) {
    hipLaunchKernelGGL(k_win_rows, dim3((p.NE + p.NI + 3) / 4), dim3(256), 0, s, p);
    return (int)hipGetLastError();
}

int launch_win_backward(const WinBackwardPlan& p, bn_stream_t s) {
    hipLaunchKernelGGL(k_win_backward, dim3((p.N + 255) / 256), dim3(256), 0, s, p);
    return (int)hipGetLastError();
}

int launch_win_outer(const WinOuterPlan& p, bn_stream_t s) {
    if (p.slots == 0) return 0;
    hipLaunchKernelGGL(k_win_outer, dim3((p.n_post + 63) / 64, (p.n_pre + 63) / 64), dim3(256), 0, s, p);
    return (int)hipGetLastError();
}

}  // namespace bn
