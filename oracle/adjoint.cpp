// ORACLE — test infrastructure only.  FP64 restatement of the differentiable window: the step
// of the plasticity-free, stimulus-free network (the same reference semantics as oracle.cpp:
// kernel_neuron_group_evolve.cu:11-130, kernel_connection_propagate.cu:8-94, system.cpp:43-63)
// run forward from a given state, the loss of the window, and its reverse-mode gradient with
// respect to the four weight matrices, g_leak and tau_decay_e.
//
// Loss (T steps): L = (1/T) sum_t [ sum_p (rho_p,t - rho*_p,t)^2
//                                  + (1/n_v) sum_k ((V_k,t - V*_k,t) / 1 mV)^2 ]
//   rho_p,t = a rho_p,t-1 + (1 - a) (spikes of population p at t / N_p) / dt, rho_p,0 = 0,
//   a = exp(-dt / rate_tau), rates in Hz; p = E, I.
// Spike derivative: ds/dU = 1 / (beta (1 + |U - v_threshold| / beta)^2) (fast sigmoid, slope
// 1/beta at threshold) or 0 (beta = 0: spikes held fixed, the exact derivative where no
// perturbation flips a spike).  The spike derivative enters through the paths of the spike
// itself (connection traces, population rates); in the reset V = s v_threshold + (1 - s) U
// the spike is held fixed (dV/dU = 1 - s).  Refractory steps (V pinned to v_reset) pass no
// gradient.
#include <omp.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

struct WinRun {
    // dimensions
    uint32_t NE, NI, N, T;
    // per-step records [t][neuron] (t = 0 .. T-1 for steps 1 .. T)
    std::vector<double> Vprev, ge, gi, U, S;
    std::vector<uint8_t> refr_flag;
    std::vector<double> dE_prev, rE_prev, dI_prev, rI_prev;  // traces at the start of each step
    std::vector<double> V_rec;                               // [t][n_vrec]
    std::vector<double> rates;                               // [t][2]
};

inline uint32_t poisson_lookup(uint32_t u, const std::vector<uint32_t>& t) {
    uint32_t k = 0;
    while (k < t.size() && u >= t[k]) ++k;
    return k;
}

void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        if (round) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
    }
}

inline double flush(double v) { return v < 1e-30 ? 0.0 : v; }


struct WinConst {
    uint32_t NE, NI, N, T, nv;
    double dt, kappa, Vr, VT, DT, Ee, Ei, theta, K, aE_d, aE_r, aI_d, aI_r, a_rate, beta, g_leak, tau_decay_e;
};

WinConst window_constants(const oracle_params* p, double g_leak, double tau_decay_e, const oracle_window_spec* spec) {
    WinConst c;
    c.NE = p->num_excitatory_neurons, c.NI = p->num_inhibitory_neurons, c.N = c.NE + c.NI, c.T = spec->steps;
    c.nv = spec->n_vrec;
    c.dt = p->timestep;
    c.kappa = c.dt / (double)p->membrane_capacitance;
    c.Vr = p->v_rest, c.VT = p->v_effective_spike_threshold, c.DT = p->depolarization_slope_factor;
    c.Ee = p->v_excitatory_reversal, c.Ei = p->v_inhibitory_reversal, c.theta = p->v_threshold;
    c.K = p->postsynaptic_kernel_normalization_value;
    c.aE_d = std::exp(-c.dt / tau_decay_e), c.aE_r = std::exp(-c.dt / (double)p->tau_rise_e);
    c.aI_d = std::exp(-c.dt / (double)p->tau_decay_i), c.aI_r = std::exp(-c.dt / (double)p->tau_rise_i);
    c.a_rate = std::exp(-c.dt / spec->rate_tau);
    c.beta = spec->surrogate_beta;
    c.g_leak = g_leak;
    c.tau_decay_e = tau_decay_e;
    return c;
}

// Reverse pass over the records of a forward run: dW (dense [pre][post]), dL/dg_leak and
// dL/dtau_decay_e (through a_d = exp(-dt / tau_decay_e): dL/dtau = dL/da_d a_d dt / tau^2).
void window_reverse(const WinConst& wc, const double* const* W, const WinRun& R, const oracle_window_spec* spec,
                    double* const* dW, double* dg_leak, double* dtau_decay_e) {
    const uint32_t NE = wc.NE, NI = wc.NI, N = wc.N, T = wc.T, nv = wc.nv;
    const double dt = wc.dt, kappa = wc.kappa, Vr = wc.Vr, VT = wc.VT, DT = wc.DT, Ee = wc.Ee, Ei = wc.Ei;
    const double theta = wc.theta, K = wc.K, aE_d = wc.aE_d, aE_r = wc.aE_r, aI_d = wc.aI_d, aI_r = wc.aI_r;
    const double a_rate = wc.a_rate, beta = wc.beta, g_leak = wc.g_leak, tau = wc.tau_decay_e;
    const uint32_t pre_of[4] = {NE, NE, NI, NI}, post_of[4] = {NE, NI, NE, NI};
    // Reverse pass.
    for (int c = 0; c < 4; ++c) std::fill(dW[c], dW[c] + (size_t)pre_of[c] * post_of[c], 0.0);
    std::vector<double> lamV(N, 0.0), lge(N, 0.0), lgi(N, 0.0), lge_next(N, 0.0), lgi_next(N, 0.0);
    std::vector<double> cdE(NE, 0.0), crE(NE, 0.0), cdI(NI, 0.0), crI(NI, 0.0), lsTrace(N, 0.0);
    std::vector<double> lxE(NE), lxI(NI);
    std::vector<double> vdirect(N, 0.0);
    std::vector<int> vslot(N, -1);
    for (uint32_t k = 0; k < nv; ++k) vslot[spec->vrec_ids[k]] = (int)k;
    double lrho[2] = {0.0, 0.0};
    double dgl = 0.0, dalpha = 0.0;
    for (int64_t t = (int64_t)T - 1; t >= 0; --t) {
        // lambda x_t = W lambda g_{t+1}
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)N; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const bool e = i < NE;
            const uint32_t il = e ? i : i - NE;
            const double* W1 = W[e ? 0 : 2];  // to E posts
            const double* W2 = W[e ? 1 : 3];  // to I posts
            const std::vector<double>& l = e ? lge_next : lgi_next;
            double acc = 0.0;
            for (uint32_t j = 0; j < NE; ++j) acc += W1[(size_t)il * NE + j] * l[j];
            for (uint32_t j = 0; j < NI; ++j) acc += W2[(size_t)il * NI + j] * l[NE + j];
            if (e) lxE[il] = acc;
            else lxI[il] = acc;
        }
        // traces
        for (uint32_t i = 0; i < NE; ++i) {
            const double d0 = R.dE_prev[(size_t)t * NE + i], r0 = R.rE_prev[(size_t)t * NE + i];
            const double ld = lxE[i] + cdE[i], lr = -lxE[i] + crE[i];
            const double nd = d0 * aE_d < 1e-30 ? 0.0 : 1.0, nr = r0 * aE_r < 1e-30 ? 0.0 : 1.0;
            dalpha += ld * d0 * nd;
            cdE[i] = aE_d * nd * ld;
            crE[i] = aE_r * nr * lr;
            lsTrace[i] = K * (ld + lr);
        }
        for (uint32_t i = 0; i < NI; ++i) {
            const double d0 = R.dI_prev[(size_t)t * NI + i], r0 = R.rI_prev[(size_t)t * NI + i];
            const double ld = lxI[i] + cdI[i], lr = -lxI[i] + crI[i];
            const double nd = d0 * aI_d < 1e-30 ? 0.0 : 1.0, nr = r0 * aI_r < 1e-30 ? 0.0 : 1.0;
            cdI[i] = aI_d * nd * ld;
            crI[i] = aI_r * nr * lr;
            lsTrace[NE + i] = K * (ld + lr);
        }
        // rates
        for (int q = 0; q < 2; ++q) {
            const double tgt = spec->target_rates ? spec->target_rates[(size_t)t * 2 + q] : 0.0;
            lrho[q] = 2.0 / T * (R.rates[(size_t)t * 2 + q] - tgt) + a_rate * lrho[q];
        }
        // neurons
        for (uint32_t j = 0; j < N; ++j) {
            const size_t k = (size_t)t * N + j;
            const int q = j < NE ? 0 : 1;
            const double ls = lsTrace[j] + lrho[q] * (1.0 - a_rate) / ((q == 0 ? NE : NI) * dt);
            double lv = lamV[j];
            if (vslot[j] >= 0) {
                const double tgt = spec->target_v ? spec->target_v[(size_t)t * nv + vslot[j]] : 0.0;
                const double v = R.V_rec[(size_t)t * nv + vslot[j]];
                lv += 2.0 / ((double)T * nv) * (v - tgt) * 1e6;
            }
            if (R.refr_flag[k]) {
                lamV[j] = 0.0;
                lge[j] = lgi[j] = 0.0;
                continue;
            }
            const double U = R.U[k], s = R.S[k], Vp = R.Vprev[k];
            double sg = 0.0;
            if (beta > 0.0) {
                const double z = 1.0 + std::fabs(U - theta) / beta;
                sg = 1.0 / (beta * z * z);
            }
            const double lU = lv * (1.0 - s) + ls * sg;
            const double ex = std::exp((Vp - VT) / DT);
            lamV[j] = lU * (1.0 + kappa * (-g_leak + g_leak * ex - R.ge[k] - R.gi[k]));
            lge[j] = -lU * kappa * (Vp - Ee);
            lgi[j] = -lU * kappa * (Vp - Ei);
            dgl += lU * kappa * (-(Vp - Vr) + DT * ex);
        }
        // dW += x_{t-1} (outer) lambda g_t
        const double* dEp = &R.dE_prev[(size_t)t * NE];
        const double* rEp = &R.rE_prev[(size_t)t * NE];
        const double* dIp = &R.dI_prev[(size_t)t * NI];
        const double* rIp = &R.rI_prev[(size_t)t * NI];
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)N; ++ii) {
            const uint32_t i = (uint32_t)ii;
            const bool e = i < NE;
            const uint32_t il = e ? i : i - NE;
            const double x = e ? dEp[il] - rEp[il] : dIp[il] - rIp[il];
            double* D1 = dW[e ? 0 : 2] + (size_t)il * NE;
            double* D2 = dW[e ? 1 : 3] + (size_t)il * NI;
            const std::vector<double>& l = e ? lge : lgi;
            for (uint32_t j = 0; j < NE; ++j) D1[j] += x * l[j];
            for (uint32_t j = 0; j < NI; ++j) D2[j] += x * l[NE + j];
        }
        lge_next.swap(lge);
        lgi_next.swap(lgi);
    }
    if (dg_leak) *dg_leak = dgl;
    if (dtau_decay_e) *dtau_decay_e = dalpha * aE_d * dt / (tau * tau);
}

// FP64 forward from s0 with the loss; with dW, the reverse pass; with capture, the records.
int window_impl(const oracle_params* p, const double* const* W, double g_leak, double tau_decay_e,
                const oracle_window_state* s0, const oracle_window_spec* spec, double* loss, double* const* dW,
                double* dg_leak, double* dtau_decay_e, double* rates_out, double* v_out, uint8_t* spikes_out,
                WinRun* capture) {
    const uint32_t NE = p->num_excitatory_neurons, NI = p->num_inhibitory_neurons, N = NE + NI, T = spec->steps;
    const double dt = p->timestep;
    const double kappa = dt / (double)p->membrane_capacitance;
    const double Vr = p->v_rest, VT = p->v_effective_spike_threshold, DT = p->depolarization_slope_factor;
    const double Ee = p->v_excitatory_reversal, Ei = p->v_inhibitory_reversal;
    const double theta = p->v_threshold, vreset = p->v_reset;
    const double K = p->postsynaptic_kernel_normalization_value;
    const double aE_d = std::exp(-dt / tau_decay_e), aE_r = std::exp(-dt / (double)p->tau_rise_e);
    const double aI_d = std::exp(-dt / (double)p->tau_decay_i), aI_r = std::exp(-dt / (double)p->tau_rise_i);
    const double a_rate = std::exp(-dt / spec->rate_tau);
    const uint32_t nv = spec->n_vrec;

    // Poisson tables and background parameters per group (same tables as the engine).
    std::vector<uint32_t> tab[2][2];
    for (int g = 0; g < 2; ++g)
        for (int k = 0; k < 2; ++k) {
            const float mean = k == 0 ? p->group[g].mean_background_excitatory_spikes : p->group[g].mean_background_inhibitory_spikes;
            tab[g][k].resize(256);
            tab[g][k].resize(oracle_poisson_table(mean, tab[g][k].data(), 256));
        }
    const uint32_t k0 = (uint32_t)(int64_t)p->rng_seed, k1 = (uint32_t)((int64_t)p->rng_seed >> 32);

    // State.
    std::vector<double> V(N), bge(N), bgi(N), bre(N), bde(N), bri(N), bdi(N);
    std::vector<uint32_t> refr(N);
    std::vector<double> dE(NE), rE(NE), dI(NI), rI(NI);
    for (uint32_t j = 0; j < N; ++j) {
        V[j] = s0->V[j], refr[j] = s0->refr[j], bge[j] = s0->bg_ge[j], bgi[j] = s0->bg_gi[j];
        bre[j] = s0->bre[j], bde[j] = s0->bde[j], bri[j] = s0->bri[j], bdi[j] = s0->bdi[j];
    }
    for (uint32_t i = 0; i < NE; ++i) dE[i] = s0->d_e[i], rE[i] = s0->r_e[i];
    for (uint32_t i = 0; i < NI; ++i) dI[i] = s0->d_i[i], rI[i] = s0->r_i[i];

    WinRun R;
    R.NE = NE, R.NI = NI, R.N = N, R.T = T;
    const bool grad = dW != nullptr || capture != nullptr;
    if (grad) {
        for (auto* v : {&R.Vprev, &R.ge, &R.gi, &R.U, &R.S}) v->assign((size_t)T * N, 0.0);
        R.refr_flag.assign((size_t)T * N, 0);
        R.dE_prev.assign((size_t)T * NE, 0.0), R.rE_prev.assign((size_t)T * NE, 0.0);
        R.dI_prev.assign((size_t)T * NI, 0.0), R.rI_prev.assign((size_t)T * NI, 0.0);
    }
    R.V_rec.assign((size_t)T * nv, 0.0);
    R.rates.assign((size_t)T * 2, 0.0);
    std::vector<double> ge(N), gi(N), xE(NE), xI(NI);
    double rho[2] = {0.0, 0.0};
    double L = 0.0;

    for (uint32_t t = 0; t < T; ++t) {
        const uint32_t step = s0->clock_step + t;
        for (uint32_t i = 0; i < NE; ++i) xE[i] = dE[i] - rE[i];
        for (uint32_t i = 0; i < NI; ++i) xI[i] = dI[i] - rI[i];
        // g at this step: background of the last step + W^T x of the last step.
#pragma omp parallel for schedule(static)
        for (int64_t jj = 0; jj < (int64_t)N; ++jj) {
            const uint32_t j = (uint32_t)jj;
            const bool e = j < NE;
            const uint32_t jl = e ? j : j - NE;
            const double* We = W[e ? 0 : 1];
            const double* Wi = W[e ? 2 : 3];
            const uint32_t np = e ? NE : NI;
            double se = 0.0, si = 0.0;
            for (uint32_t i = 0; i < NE; ++i) se += We[(size_t)i * np + jl] * xE[i];
            for (uint32_t i = 0; i < NI; ++i) si += Wi[(size_t)i * np + jl] * xI[i];
            ge[j] = bge[j] + se;
            gi[j] = bgi[j] + si;
        }
        if (grad) {
            std::memcpy(&R.dE_prev[(size_t)t * NE], dE.data(), NE * sizeof(double));
            std::memcpy(&R.rE_prev[(size_t)t * NE], rE.data(), NE * sizeof(double));
            std::memcpy(&R.dI_prev[(size_t)t * NI], dI.data(), NI * sizeof(double));
            std::memcpy(&R.rI_prev[(size_t)t * NI], rI.data(), NI * sizeof(double));
        }
        uint32_t count[2] = {0, 0};
        std::vector<uint8_t> s(N, 0);
        for (uint32_t j = 0; j < N; ++j) {
            const size_t k = (size_t)t * N + j;
            const double Vp = V[j];
            if (grad) {
                R.Vprev[k] = Vp, R.ge[k] = ge[j], R.gi[k] = gi[j];
                R.refr_flag[k] = refr[j] > 0;
            }
            if (refr[j] > 0) {
                refr[j] -= 1;
                V[j] = vreset;
                s[j] = 0;
            } else {
                const double F = -g_leak * (Vp - Vr) + g_leak * DT * std::exp((Vp - VT) / DT) - ge[j] * (Vp - Ee) -
                                 gi[j] * (Vp - Ei);
                const double U = Vp + kappa * F;
                const bool sp = spec->forced_spikes ? spec->forced_spikes[(size_t)t * N + j] != 0 : U >= theta;
                s[j] = sp;
                V[j] = sp ? theta : U;
                refr[j] = sp ? p->refractory_steps : 0;
                if (grad) R.U[k] = U;
            }
            if (grad) R.S[k] = s[j];
            count[j < NE ? 0 : 1] += s[j];
            // Background (param-independent): decay, Poisson counts, conductances.
            const oracle_group_params& G = p->group[j < NE ? 0 : 1];
            double re = flush(bre[j] * (double)G.background_excitatory_postsynaptic_r_trace_decay);
            double ri = flush(bri[j] * (double)G.background_inhibitory_postsynaptic_r_trace_decay);
            double de = flush(bde[j] * (double)G.background_excitatory_postsynaptic_d_trace_decay);
            double di = flush(bdi[j] * (double)G.background_inhibitory_postsynaptic_d_trace_decay);
            const int g = j < NE ? 0 : 1;
            if (!tab[g][0].empty() || !tab[g][1].empty()) {
                uint32_t c[4] = {step, 0u, j, p->rng_instance};
                philox(c, k0, k1);
                const uint32_t ke = G.mean_background_excitatory_spikes > 0.0f ? poisson_lookup(c[0], tab[g][0]) : 0;
                const uint32_t ki = G.mean_background_inhibitory_spikes > 0.0f ? poisson_lookup(c[1], tab[g][1]) : 0;
                re += ke, de += ke, ri += ki, di += ki;
            }
            bre[j] = re, bde[j] = de, bri[j] = ri, bdi[j] = di;
            bge[j] = (de - re) * (double)G.background_excitatory_scaling_factor;
            bgi[j] = (di - ri) * (double)G.background_inhibitory_scaling_factor;
        }
        for (uint32_t i = 0; i < NE; ++i) {
            dE[i] = flush(dE[i] * aE_d) + K * s[i];
            rE[i] = flush(rE[i] * aE_r) + K * s[i];
        }
        for (uint32_t i = 0; i < NI; ++i) {
            dI[i] = flush(dI[i] * aI_d) + K * s[NE + i];
            rI[i] = flush(rI[i] * aI_r) + K * s[NE + i];
        }
        if (spikes_out) std::memcpy(spikes_out + (size_t)t * N, s.data(), N);
        // Loss terms.
        double lt = 0.0;
        for (int q = 0; q < 2; ++q) {
            rho[q] = a_rate * rho[q] + (1.0 - a_rate) * ((double)count[q] / (q == 0 ? NE : NI)) / dt;
            R.rates[(size_t)t * 2 + q] = rho[q];
            const double tgt = spec->target_rates ? spec->target_rates[(size_t)t * 2 + q] : 0.0;
            lt += (rho[q] - tgt) * (rho[q] - tgt);
        }
        for (uint32_t k = 0; k < nv; ++k) {
            const double v = V[spec->vrec_ids[k]];
            R.V_rec[(size_t)t * nv + k] = v;
            const double tgt = spec->target_v ? spec->target_v[(size_t)t * nv + k] : 0.0;
            lt += ((v - tgt) * 1e3) * ((v - tgt) * 1e3) / nv;
        }
        L += lt / T;
    }
    if (loss) *loss = L;
    if (rates_out) std::memcpy(rates_out, R.rates.data(), R.rates.size() * sizeof(double));
    if (v_out) std::memcpy(v_out, R.V_rec.data(), R.V_rec.size() * sizeof(double));
    if (capture) *capture = R;
    if (!dW) return 0;

    window_reverse(window_constants(p, g_leak, tau_decay_e, spec), W, R, spec, dW, dg_leak, dtau_decay_e);
    return 0;
}

}  // namespace

extern "C" int oracle_window(const oracle_params* p, const double* const* W, double g_leak, double tau_decay_e,
                             const oracle_window_state* s0, const oracle_window_spec* spec, double* loss,
                             double* const* dW, double* dg_leak, double* dtau_decay_e, double* rates_out,
                             double* v_out, uint8_t* spikes_out) {
    return window_impl(p, W, g_leak, tau_decay_e, s0, spec, loss, dW, dg_leak, dtau_decay_e, rates_out, v_out,
                       spikes_out, nullptr);
}

extern "C" int oracle_window_records(const oracle_params* p, const double* const* W, double g_leak,
                                     double tau_decay_e, const oracle_window_state* s0,
                                     const oracle_window_spec* spec, float* Vprev, float* ge, float* gi,
                                     uint8_t* refr, uint8_t* spikes, float* d_e, float* r_e, float* d_i, float* r_i,
                                     float* v_rec) {
    WinRun R;
    const int rc = window_impl(p, W, g_leak, tau_decay_e, s0, spec, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, &R);
    if (rc) return rc;
    auto put = [](float* dst, const std::vector<double>& src) {
        for (size_t k = 0; k < src.size(); ++k) dst[k] = (float)src[k];
    };
    put(Vprev, R.Vprev), put(ge, R.ge), put(gi, R.gi);
    for (size_t k = 0; k < R.S.size(); ++k) spikes[k] = R.S[k] != 0.0, refr[k] = R.refr_flag[k];
    put(d_e, R.dE_prev), put(r_e, R.rE_prev), put(d_i, R.dI_prev), put(r_i, R.rI_prev), put(v_rec, R.V_rec);
    return 0;
}

extern "C" int oracle_window_along(const oracle_params* p, const double* const* W, double g_leak, double tau_decay_e,
                                   const oracle_window_trajectory* tr, const oracle_window_spec* spec, double* loss,
                                   double* const* dW, double* dg_leak, double* dtau_decay_e) {
    const WinConst wc = window_constants(p, g_leak, tau_decay_e, spec);
    const uint32_t NE = wc.NE, NI = wc.NI, N = wc.N, T = wc.T, nv = wc.nv;
    WinRun R;
    R.NE = NE, R.NI = NI, R.N = N, R.T = T;
    const size_t TN = (size_t)T * N;
    R.Vprev.resize(TN), R.ge.resize(TN), R.gi.resize(TN), R.U.assign(TN, 0.0), R.S.resize(TN);
    R.refr_flag.resize(TN);
    for (size_t k = 0; k < TN; ++k) {
        R.Vprev[k] = tr->Vprev[k], R.ge[k] = tr->ge[k], R.gi[k] = tr->gi[k];
        R.S[k] = tr->spikes[k], R.refr_flag[k] = tr->refr[k];
        if (!R.refr_flag[k]) {
            const double Vp = R.Vprev[k];
            const double F = -g_leak * (Vp - wc.Vr) + g_leak * wc.DT * std::exp((Vp - wc.VT) / wc.DT) -
                             R.ge[k] * (Vp - wc.Ee) - R.gi[k] * (Vp - wc.Ei);
            R.U[k] = Vp + wc.kappa * F;
        }
    }
    R.dE_prev.assign(tr->d_e, tr->d_e + (size_t)T * NE), R.rE_prev.assign(tr->r_e, tr->r_e + (size_t)T * NE);
    R.dI_prev.assign(tr->d_i, tr->d_i + (size_t)T * NI), R.rI_prev.assign(tr->r_i, tr->r_i + (size_t)T * NI);
    R.V_rec.assign(tr->v_rec, tr->v_rec + (size_t)T * nv);
    R.rates.assign((size_t)T * 2, 0.0);
    double rho[2] = {0.0, 0.0}, L = 0.0;
    for (uint32_t t = 0; t < T; ++t) {
        uint32_t count[2] = {0, 0};
        for (uint32_t j = 0; j < N; ++j) count[j < NE ? 0 : 1] += tr->spikes[(size_t)t * N + j] != 0;
        double lt = 0.0;
        for (int q = 0; q < 2; ++q) {
            rho[q] = wc.a_rate * rho[q] + (1.0 - wc.a_rate) * ((double)count[q] / (q == 0 ? NE : NI)) / wc.dt;
            R.rates[(size_t)t * 2 + q] = rho[q];
            const double tgt = spec->target_rates ? spec->target_rates[(size_t)t * 2 + q] : 0.0;
            lt += (rho[q] - tgt) * (rho[q] - tgt);
        }
        for (uint32_t k = 0; k < nv; ++k) {
            const double v = R.V_rec[(size_t)t * nv + k];
            const double tgt = spec->target_v ? spec->target_v[(size_t)t * nv + k] : 0.0;
            lt += ((v - tgt) * 1e3) * ((v - tgt) * 1e3) / nv;
        }
        L += lt / T;
    }
    if (loss) *loss = L;
    if (dW) window_reverse(wc, W, R, spec, dW, dg_leak, dtau_decay_e);
    return 0;
}
