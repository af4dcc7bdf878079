// ORACLE — test infrastructure only (see oracle.h).  CPU restatement of the reference's
// per-step semantics.  Each function cites the reference lines it restates.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// Parts the propagate chunk partials are combined in (the engine sums one part per wave).
constexpr uint32_t kFinParts = 8;
// Parts the spiking-post list of a non-spiking row's LTP is cut into (one wave each).
constexpr uint32_t kQuietParts = 4;

// ---------------------------------------------------------------------------------------
// Primitives
// ---------------------------------------------------------------------------------------
// Philox4x32-10, Salmon et al. SC'11 (published round function and constants).
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

float bits_to_float(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
float pow2f(int e) { return bits_to_float((uint32_t)(e + 127) << 23); }

// Cody–Waite reduction + degree-7 Taylor, binary32 throughout (same method as the kernels).
float expf_engine(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return bits_to_float(0x7f800000u);
    if (x < -103.97208404541015625f) return 0.0f;
    const float kf = std::rint(x * 1.44269502162933349609375f);
    const int k = (int)kf;
    float r = x - kf * 0.693359375f;
    r = r - kf * -2.12194440e-4f;
    float p = 1.98412698e-4f;
    p = p * r + 1.38888889e-3f;
    p = p * r + 8.33333377e-3f;
    p = p * r + 4.16666679e-2f;
    p = p * r + 1.66666672e-1f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    if (k > 127) return (p * pow2f(k - 1)) * 2.0f;
    if (k < -126) return (p * pow2f(k + 126)) * pow2f(-126);
    return p * pow2f(k);
}

uint32_t poisson_table(float mean, uint32_t* t, uint32_t cap) {
    if (!(mean > 0.0f)) return 0;
    const double lam = (double)mean;
    double p = std::exp(-lam);
    double cdf = p;
    for (uint32_t k = 0; k < cap; ++k) {
        const double scaled = std::ldexp(cdf, 32);
        const uint32_t v = scaled >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)std::floor(scaled);
        t[k] = v;
        if (v == 0xFFFFFFFFu) return k + 1;
        p = p * lam / (double)(k + 1);
        cdf = cdf + p;
    }
    return cap;
}

uint32_t poisson_sample(uint32_t u, const std::vector<uint32_t>& t) {
    uint32_t k = 0;
    while (k < t.size() && u >= t[k]) ++k;
    return k;
}

template <typename R>
R flush(R v) {
    return v < (R)1e-30f ? (R)0 : v;
}
template <typename R>
R clampw(R w, R lo, R hi) {
    if (w < lo) return lo;
    if (w > hi) return hi;
    return w;
}
inline float ex(float x) { return expf_engine(x); }
inline double ex(double x) { return std::exp(x); }

// Normalisation row sum (engine order): 256 interleaved FP64 partials, xor butterfly in each
// group of 64, the four group sums added in order.
double canonical_row_sum(const float* row, uint32_t n) {
    double part[256];
    for (int t = 0; t < 256; ++t) {
        double a = 0.0;
        for (uint32_t j = (uint32_t)t; j < n; j += 256) a += (double)row[j];
        part[t] = a;
    }
    double wave[4];
    for (int w = 0; w < 4; ++w) {
        double v[64];
        for (int l = 0; l < 64; ++l) v[l] = part[64 * w + l];
        for (int off = 32; off >= 1; off >>= 1) {
            double nv[64];
            for (int l = 0; l < 64; ++l) nv[l] = v[l] + v[l ^ off];
            std::memcpy(v, nv, sizeof v);
        }
        wave[w] = v[0];
    }
    return ((wave[0] + wave[1]) + wave[2]) + wave[3];
}

// The xor butterfly (32,16,8,4,2,1) over 64 lane values: the value every lane holds at the end.
double butterfly64(double (&v)[64]) {
    for (int off = 32; off >= 1; off >>= 1) {
        double nv[64];
        for (int l = 0; l < 64; ++l) nv[l] = v[l] + v[l ^ off];
        std::memcpy(v, nv, sizeof v);
    }
    return v[0];
}

// Sum over one 256-column tile of a row in the persistent step's order: lane l sums its four
// columns from 0.0, the 64 lane sums combined by the xor butterfly (32,16,8,4,2,1).  val(j)
// gives the term of column j (columns >= n contribute nothing).
template <typename Val>
double tile_sum(uint32_t q, uint32_t n, Val val) {
    double v[64];
    for (uint32_t l = 0; l < 64; ++l) {
        double a = 0.0;
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t j = q * 256 + 4 * l + k;
            if (j < n) a += val(j);
        }
        v[l] = a;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        double nv[64];
        for (int l = 0; l < 64; ++l) nv[l] = v[l] + v[l ^ off];
        std::memcpy(v, nv, sizeof v);
    }
    return v[0];
}

// ---------------------------------------------------------------------------------------
// Simulation
// ---------------------------------------------------------------------------------------
enum { kTriplet = 0, kIstdp = 1, kPair = 2 };

template <typename R>
struct Sim {
    oracle_params p{};
    uint32_t NE = 0, NI = 0, N = 0;
    uint32_t step = 0;
    // neuron state (global ids)
    std::vector<R> V, ge, gi, istim, bg_ge, bg_gi, bre, bde, bri, bdi;
    std::vector<uint32_t> refr, spikes;
    struct Group {
        uint32_t off, n;
        oracle_group_params gp;
        std::vector<uint32_t> tab_e, tab_i;
    } grp[2];
    struct Conn {
        uint32_t pre_g, post_g, pre_off, n_pre, post_off, n_post, row_block;
        bool inh;
        R decay_r, decay_d, K;
        std::vector<R> W;  // dense [pre][post]
        std::vector<R> r, d, x;
    } conn[4];
    struct Plast {
        int kind;
        uint32_t c;  // connection index
        R ltp_A, ltd_A, ltp_B, ltd_B, pre_decay, post_decay, tpre_decay, tpost_decay, wmin, wmax;
        std::vector<R> pre, post, tpre, tpost;  // local ids
        std::vector<R> target;                  // pre_weight_sums
        // lazy normalisation (FP32 only)
        bool lazy = false;
        std::vector<float> s;
        std::vector<double> Rs;
    };
    std::vector<Plast> plast;
    std::vector<uint32_t> stim_counts, stim_ids, stim_offs;
    std::vector<float> stim_amps;
    uint32_t stim_t = 0, stim_next = 0;
    static constexpr bool kF32 = sizeof(R) == 4;

    R dec(float dt, float tau) const {
        if (kF32) return (R)std::exp(-1.0f * dt / tau);  // connection.cpp:46-47, main_buryn.cpp:295-298
        return (R)std::exp(-(double)dt / (double)tau);
    }

    void init(const oracle_params& prm, const float* w[4]) {
        p = prm;
        NE = p.num_excitatory_neurons;
        NI = p.num_inhibitory_neurons;
        N = NE + NI;
        auto z = [&](std::vector<R>& v) { v.assign(N, (R)0); };
        V.assign(N, (R)p.v_rest);  // neuron_group.cpp:21
        z(ge), z(gi), z(istim), z(bg_ge), z(bg_gi), z(bre), z(bde), z(bri), z(bdi);
        refr.assign(N, 0);
        spikes.assign(N, 0);
        for (int g = 0; g < 2; ++g) {
            grp[g].off = g == 0 ? 0 : NE;
            grp[g].n = g == 0 ? NE : NI;
            grp[g].gp = p.group[g];
            std::vector<uint32_t> t(256);
            t.resize(poisson_table(grp[g].gp.mean_background_excitatory_spikes, t.data(), 256));
            grp[g].tab_e = t;
            t.assign(256, 0);
            t.resize(poisson_table(grp[g].gp.mean_background_inhibitory_spikes, t.data(), 256));
            grp[g].tab_i = t;
        }
        // main_buryn.cpp:210-272 — EE, EI, IE, II
        const uint32_t pre_g[4] = {0, 0, 1, 1}, post_g[4] = {0, 1, 0, 1};
        for (int c = 0; c < 4; ++c) {
            Conn& C = conn[c];
            C.pre_g = pre_g[c];
            C.post_g = post_g[c];
            C.pre_off = grp[C.pre_g].off;
            C.n_pre = grp[C.pre_g].n;
            C.post_off = grp[C.post_g].off;
            C.n_post = grp[C.post_g].n;
            C.row_block = p.gemv_row_block[c] ? p.gemv_row_block[c] : C.n_pre;
            C.inh = C.pre_g == 1;
            const bool e = C.pre_g == 0;
            C.decay_r = dec(p.timestep, e ? p.tau_rise_e : p.tau_rise_i);
            C.decay_d = dec(p.timestep, e ? p.tau_decay_e : p.tau_decay_i);
            C.K = (R)p.postsynaptic_kernel_normalization_value;
            C.W.resize((size_t)C.n_pre * C.n_post);
            for (size_t k = 0; k < C.W.size(); ++k) C.W[k] = (R)w[c][k];
            C.r.assign(C.n_pre, (R)0);
            C.d.assign(C.n_pre, (R)0);
            C.x.assign(C.n_pre, (R)0);
        }
        if (p.e_e_plasticity_enabled) {
            Plast P{};
            // main_buryn.cpp:286-305 (triplet) or the reference's StdpPlasticity component.
            P.kind = p.e_e_pair_stdp ? kPair : kTriplet;
            P.c = 0;
            P.ltp_A = (R)p.e_e_ltp_A, P.ltd_A = (R)p.e_e_ltd_A, P.ltp_B = (R)p.e_e_ltp_B, P.ltd_B = (R)p.e_e_ltd_B;
            P.pre_decay = dec(p.timestep, p.e_e_tau_ltp);
            P.post_decay = dec(p.timestep, p.e_e_tau_ltd);
            P.tpre_decay = dec(p.timestep, p.e_e_triplet_tau_ltp);
            P.tpost_decay = dec(p.timestep, p.e_e_triplet_tau_ltd);
            if (P.kind == kPair) P.ltp_B = P.ltd_B = P.tpre_decay = P.tpost_decay = (R)0;
            P.wmin = (R)p.e_e_weight_min, P.wmax = (R)p.e_e_weight_max;
            plast.push_back(P);
        }
        if (p.i_e_plasticity_enabled) {
            Plast P{};
            P.kind = kIstdp;
            P.c = 2;
            P.ltp_A = (R)p.i_e_ltp_A, P.ltd_A = (R)p.i_e_ltd_A, P.ltp_B = (R)0, P.ltd_B = (R)p.i_e_ltd_B;
            P.pre_decay = dec(p.timestep, p.i_e_tau_ltp);
            P.post_decay = dec(p.timestep, p.i_e_tau_ltd);
            P.tpre_decay = P.tpost_decay = (R)0;
            P.wmin = (R)p.i_e_weight_min, P.wmax = (R)p.i_e_weight_max;
            plast.push_back(P);
        }
        for (Plast& P : plast) {
            const Conn& C = conn[P.c];
            P.pre.assign(C.n_pre, (R)0);
            P.post.assign(C.n_post, (R)0);
            P.tpre.assign(C.n_pre, (R)0);
            P.tpost.assign(C.n_post, (R)0);
            if (P.kind == kTriplet) {
                // triplet_plasticity.cpp:22-30: column sums of the initial matrix, FP32 sequential
                P.target.assign(C.n_post, (R)0);
                for (uint32_t j = 0; j < C.n_post; ++j) {
                    R sum = (R)0;
                    for (uint32_t i = 0; i < C.n_pre; ++i) sum += (R)w[P.c][(size_t)i * C.n_post + j];
                    P.target[j] = sum;
                }
                if (kF32 && p.lazy_normalize) {
                    P.lazy = true;
                    P.s.assign(C.n_pre, 1.0f);
                    P.Rs.resize(C.n_pre);
                    for (uint32_t i = 0; i < C.n_pre; ++i)
                        P.Rs[i] = canonical_row_sum(reinterpret_cast<const float*>(&C.W[(size_t)i * C.n_post]), C.n_post);
                }
            }
        }
    }

    // kernel_neuron_group_evolve.cu:11-83 then :85-130, for one group.
    void evolve_group(const Group& G) {
        const oracle_group_params& gp = G.gp;
        const R dt = (R)p.timestep, C = (R)p.membrane_capacitance, gl = (R)p.g_leak;
        const R dT = (R)p.depolarization_slope_factor, VT = (R)p.v_effective_spike_threshold;
        const R Ee = (R)p.v_excitatory_reversal, Ei = (R)p.v_inhibitory_reversal;
        const R vth = (R)p.v_threshold, vreset = (R)p.v_reset, vrest = (R)p.v_rest;
#pragma omp parallel for schedule(static)
        for (int64_t li = 0; li < (int64_t)G.n; ++li) {
            const uint32_t n = G.off + (uint32_t)li;
            // background
            R re = flush(bre[n] * (R)gp.background_excitatory_postsynaptic_r_trace_decay);
            R ri = flush(bri[n] * (R)gp.background_inhibitory_postsynaptic_r_trace_decay);
            R de = flush(bde[n] * (R)gp.background_excitatory_postsynaptic_d_trace_decay);
            R di = flush(bdi[n] * (R)gp.background_inhibitory_postsynaptic_d_trace_decay);
            uint32_t c4[4] = {step, 0u, n, p.rng_instance};
            philox(c4, (uint32_t)(uint64_t)(int64_t)p.rng_seed, (uint32_t)((uint64_t)(int64_t)p.rng_seed >> 32));
            const uint32_t ke = gp.mean_background_excitatory_spikes > 0.0f ? poisson_sample(c4[0], G.tab_e) : 0u;
            const uint32_t ki = gp.mean_background_inhibitory_spikes > 0.0f ? poisson_sample(c4[1], G.tab_i) : 0u;
            if (ke > 0u) {
                re += (R)ke;
                de += (R)ke;
            }
            const R bge = (de - re) * (R)gp.background_excitatory_scaling_factor;
            if (ki > 0u) {
                ri += (R)ki;
                di += (R)ki;
            }
            const R bgi = (di - ri) * (R)gp.background_inhibitory_scaling_factor;
            bre[n] = re, bri[n] = ri, bde[n] = de, bdi[n] = di;
            bg_ge[n] = bge, bg_gi[n] = bgi;
            // EIF
            if (refr[n] > 0) {
                refr[n] -= 1;
                spikes[n] = 0;
                V[n] = vreset;
            } else {
                const R v = V[n];
                const R linear = (R)-1.0f * gl * (v - vrest);
                const R expo = gl * dT * ex((v - VT) / dT);
                const R exc = (R)-1.0f * ge[n] * (v - Ee);
                const R inh = (R)-1.0f * gi[n] * (v - Ei);
                const R dv = (dt / C) * (linear + expo + exc + inh + istim[n]);
                const R vn = v + dv;
                if (vn >= vth) {
                    refr[n] = p.refractory_steps;
                    spikes[n] = 1;
                    V[n] = vth;
                } else {
                    spikes[n] = 0;
                    V[n] = vn;
                }
            }
            ge[n] = bge;
            gi[n] = bgi;
        }
    }

    // kernel_connection_propagate.cu:8-46 then :79-94 (GEMV in the engine's order).
    void propagate(int c) {
        Conn& C = conn[c];
        for (uint32_t i = 0; i < C.n_pre; ++i) {
            R d = flush(C.d[i] * C.decay_d);
            R r = flush(C.r[i] * C.decay_r);
            if (spikes[C.pre_off + i] == 1) {
                r += C.K;
                d += C.K;
            }
            C.d[i] = d;
            C.r[i] = r;
            C.x[i] = d - r;
        }
        std::vector<R>& g = C.inh ? gi : ge;
        std::vector<R> xs(C.x);
        for (const Plast& P : plast)
            if (P.lazy && P.c == (uint32_t)c)
                for (uint32_t i = 0; i < C.n_pre; ++i) xs[i] = (R)((float)C.x[i] * P.s[i]);
        const uint32_t NB = 256;
        const int64_t nblk = (C.n_post + NB - 1) / NB;
#pragma omp parallel for schedule(static)
        for (int64_t b = 0; b < nblk; ++b) {
            const uint32_t j0 = (uint32_t)b * NB, j1 = std::min(j0 + NB, C.n_post);
            double acc[NB], part[NB];
            for (uint32_t j = j0; j < j1; ++j) acc[j - j0] = (double)g[C.post_off + j];
            if (kF32 && p.resident_columns) {
                // Resident step order: thread t of the T = 64 * waves threads of this pre group
                // sums rows t, t + T, ... from 0.0; the lane sums of a wave by the xor
                // butterfly; g + wave_0 + wave_1 + ... in wave order.
                const uint32_t waves = C.pre_g == 0 ? p.resident_e_waves : 8 - p.resident_e_waves;
                const uint32_t T = 64 * waves;
                static thread_local std::vector<double> lane;
                lane.assign((size_t)64 * NB, 0.0);
                for (uint32_t w = 0; w < waves; ++w) {
                    std::fill(lane.begin(), lane.end(), 0.0);
                    for (uint32_t l = 0; l < 64; ++l) {
                        double* a = &lane[(size_t)l * NB];
                        for (uint32_t i = 64 * w + l; i < C.n_pre; i += T) {
                            const double xi = (double)xs[i];
                            if (xi == 0.0 && !p.dense_gemv) continue;  // adds +0 in the engine: no effect
                            const R* row = &C.W[(size_t)i * C.n_post];
                            for (uint32_t j = j0; j < j1; ++j) a[j - j0] += xi * (double)row[j];
                        }
                    }
                    for (uint32_t j = j0; j < j1; ++j) {
                        double v[64];
                        for (uint32_t l = 0; l < 64; ++l) v[l] = lane[(size_t)l * NB + (j - j0)];
                        acc[j - j0] += butterfly64(v);
                    }
                }
            } else if (kF32) {
                // Engine order: chunk partials (row_block rows each, increasing i, from 0.0);
                // the chunks cut into kFinParts contiguous parts of ceil(chunks/kFinParts),
                // each part summed in chunk order from 0.0; then g + part_0 + part_1 + ...
                const uint32_t chunks = (C.n_pre + C.row_block - 1) / C.row_block;
                const uint32_t per =
                    p.gemv_part_chunks[c] ? p.gemv_part_chunks[c] : (chunks + kFinParts - 1) / kFinParts;
                double qsum[NB];
                for (uint32_t c0 = 0; c0 < chunks; c0 += per) {
                    const uint32_t c1 = std::min(c0 + per, chunks);
                    for (uint32_t j = j0; j < j1; ++j) qsum[j - j0] = 0.0;
                    for (uint32_t c = c0; c < c1; ++c) {
                        const uint32_t i0 = c * C.row_block, i1 = std::min(i0 + C.row_block, C.n_pre);
                        for (uint32_t j = j0; j < j1; ++j) part[j - j0] = 0.0;
                        for (uint32_t i = i0; i < i1; ++i) {
                            const double xi = (double)xs[i];
                            if (xi == 0.0 && !p.dense_gemv) continue;  // adds +0 in the engine: no effect
                            const R* row = &C.W[(size_t)i * C.n_post];
                            for (uint32_t j = j0; j < j1; ++j) part[j - j0] += xi * (double)row[j];
                        }
                        for (uint32_t j = j0; j < j1; ++j) qsum[j - j0] += part[j - j0];
                    }
                    for (uint32_t j = j0; j < j1; ++j) acc[j - j0] += qsum[j - j0];
                }
            } else {
                for (uint32_t i = 0; i < C.n_pre; ++i) {
                    const double xi = (double)xs[i];
                    if (xi == 0.0 && !p.dense_gemv) continue;
                    const R* row = &C.W[(size_t)i * C.n_post];
                    for (uint32_t j = j0; j < j1; ++j) acc[j - j0] += xi * (double)row[j];
                }
            }
            for (uint32_t j = j0; j < j1; ++j) g[C.post_off + j] = (R)acc[j - j0];
        }
    }

    // stimulus.cpp:46-60 + kernel_stimulus_apply.cu:7-21 (events applied in list order)
    int stimulus() {
        if (stim_t >= stim_counts.size()) return 0;
        const uint32_t n = stim_counts[stim_t];
        if ((uint64_t)stim_next + n > stim_ids.size()) return 1;
        for (uint32_t e = stim_next; e < stim_next + n; ++e)
            istim[stim_ids[e] + stim_offs[e]] += (R)stim_amps[e];
        stim_next += n;
        ++stim_t;
        return 0;
    }

    R ltp_delta(const Plast& P, uint32_t i, uint32_t j) const {
        if (P.kind == kTriplet) return P.pre[i] * (P.ltp_A + (P.ltp_B * P.tpost[j]));
        return P.ltp_A * P.pre[i];  // kernel_istdp_plasticity_update.cu:55, kernel_stdp_plasticity_update.cu:55
    }
    R ltd_delta(const Plast& P, uint32_t i, uint32_t j) const {
        if (P.kind == kTriplet) return (R)-1.0f * P.post[j] * (P.ltd_A + (P.ltd_B * P.tpre[i]));
        if (P.kind == kPair) return (R)-1.0f * P.ltd_A * P.post[j];  // kernel_stdp_plasticity_update.cu:77
        return (R)-1.0f * P.ltd_A * (P.post[j] - P.ltd_B);
    }

    // kernel_triplet_plasticity_update.cu:181-216 / kernel_istdp_plasticity_update.cu:103-135
    void plasticity(Plast& P) {
        Conn& C = conn[P.c];
        const uint32_t np = C.n_pre, nq = C.n_post;
        for (uint32_t i = 0; i < np; ++i) {
            P.pre[i] = flush(P.pre[i] * P.pre_decay);
            if (P.kind == kTriplet) P.tpre[i] = flush(P.tpre[i] * P.tpre_decay);
        }
        for (uint32_t j = 0; j < nq; ++j) {
            P.post[j] = flush(P.post[j] * P.post_decay);
            if (P.kind == kTriplet) P.tpost[j] = flush(P.tpost[j] * P.tpost_decay);
        }
        const uint32_t* pre_sp = &spikes[C.pre_off];
        const uint32_t* post_sp = &spikes[C.post_off];
        if (!P.lazy) {
            // LTP: every column of a spiking post (index-based self skip)
            for (uint32_t j = 0; j < nq; ++j) {
                if (post_sp[j] != 1) continue;
#pragma omp parallel for schedule(static)
                for (int64_t ii = 0; ii < (int64_t)np; ++ii) {
                    const uint32_t i = (uint32_t)ii;
                    if (i == j) continue;
                    R& w = C.W[(size_t)i * nq + j];
                    w = clampw(w + ltp_delta(P, i, j), P.wmin, P.wmax);
                }
            }
            // LTD: every row of a spiking pre
            for (uint32_t i = 0; i < np; ++i) {
                if (pre_sp[i] != 1) continue;
                R* row = &C.W[(size_t)i * nq];
                for (uint32_t j = 0; j < nq; ++j) {
                    if (i == j) continue;
                    row[j] = clampw(row[j] + ltd_delta(P, i, j), P.wmin, P.wmax);
                }
            }
            if (P.kind == kTriplet) {
                // kernel_triplet_plasticity_update.cu:117-145
#pragma omp parallel for schedule(static)
                for (int64_t ii = 0; ii < (int64_t)np; ++ii) {
                    const uint32_t i = (uint32_t)ii;
                    R* row = &C.W[(size_t)i * nq];
                    R sum;
                    if (kF32) {
                        sum = (R)canonical_row_sum(reinterpret_cast<const float*>(row), nq);
                    } else {
                        double s = 0.0;
                        for (uint32_t j = 0; j < nq; ++j) s += (double)row[j];
                        sum = (R)s;
                    }
                    if (sum <= (R)0) continue;
                    const R scale = P.target[i] / sum;
                    for (uint32_t j = 0; j < nq; ++j) row[j] *= scale;
                }
            }
        } else {
            lazy_plasticity(P, C, pre_sp, post_sp);
        }
        for (uint32_t i = 0; i < np; ++i)
            if (pre_sp[i] == 1) {
                P.pre[i] += (R)1;
                if (P.kind == kTriplet) P.tpre[i] += (R)1;
            }
        for (uint32_t j = 0; j < nq; ++j)
            if (post_sp[j] == 1) {
                P.post[j] += (R)1;
                if (P.kind == kTriplet) P.tpost[j] += (R)1;
            }
    }

    void lazy_plasticity(Plast& P, Conn& C, const uint32_t* pre_sp, const uint32_t* post_sp) {
        const uint32_t np = C.n_pre, nq = C.n_post;
        float* M = reinterpret_cast<float*>(C.W.data());
        std::vector<uint32_t> posts;
        for (uint32_t j = 0; j < nq; ++j)
            if (post_sp[j] == 1) posts.push_back(j);
        // LTP over every spiking post column (rows of spiking pres included: their sums are
        // recomputed below).  Engine order of each row's change of M: the spiking-post list
        // (increasing post order) cut into kQuietParts contiguous parts of ceil(count/4); each
        // part's changes summed in list order from 0.0; R += part_0 + part_1 + ... (non-empty).
        const uint32_t count = (uint32_t)posts.size();
        const uint32_t per = (count + kQuietParts - 1) / kQuietParts;
        std::vector<double> dq((size_t)kQuietParts * np, 0.0);
        // change[(e, i)] = (double)M' - (double)M of the LTP entry (i, posts[e]) (tile order).
        const bool keep_changes = p.lazy_tile_order || p.resident_columns;
        std::vector<double> change(keep_changes ? (size_t)count * np : 0, 0.0);
        for (uint32_t e = 0; e < count; ++e) {
            const uint32_t j = posts[e], q = e / per;
            for (uint32_t i = 0; i < np; ++i) {
                if (i == j || P.s[i] == 0.0f) continue;
                float& m = M[(size_t)i * nq + j];
                const float w = clampw<float>(P.s[i] * m + (float)ltp_delta(P, i, j), (float)P.wmin, (float)P.wmax);
                const float mn = w / P.s[i];
                dq[(size_t)q * np + i] += (double)mn - (double)m;
                if (keep_changes) change[(size_t)e * np + i] = (double)mn - (double)m;
                m = mn;
            }
        }
        change_ = std::move(change);
        if (p.resident_columns) {
            lazy_resident_sums(P, M, np, nq, pre_sp, post_sp, posts);
            return;
        }
        if (p.lazy_tile_order) {
            lazy_tile_order_sums(P, M, np, nq, pre_sp, post_sp, posts);
            return;
        }
        for (uint32_t i = 0; i < np; ++i) {
            float* row = &M[(size_t)i * nq];
            if (pre_sp[i] == 1) {
                if (P.s[i] != 0.0f)
                    for (uint32_t j = 0; j < nq; ++j) {
                        if (i == j) continue;
                        const float w = clampw<float>(P.s[i] * row[j] + (float)ltd_delta(P, i, j), (float)P.wmin, (float)P.wmax);
                        row[j] = w / P.s[i];
                    }
                P.Rs[i] = canonical_row_sum(row, nq);
            } else {
                double Rn = P.Rs[i];
                for (uint32_t q = 0; q < kQuietParts; ++q)
                    if (q * per < count) Rn += dq[(size_t)q * np + i];
                P.Rs[i] = Rn;
            }
            if (P.Rs[i] > 0.0) P.s[i] = (float)((double)P.target[i] / P.Rs[i]);
        }
    }

    std::vector<double> change_;  // LTP changes of the last lazy update, [spiking post][row]

    // The persistent step's order of the lazy row sums (see oracle.h).  The LTP pass above has
    // already updated the spiking-post columns (change_ holds each entry's change).
    void lazy_tile_order_sums(Plast& P, float* M, uint32_t np, uint32_t nq, const uint32_t* pre_sp,
                              const uint32_t* post_sp, const std::vector<uint32_t>& posts) {
        const uint32_t ntiles = (nq + 255) / 256;
        std::vector<int32_t> post_rank(nq, -1);
        for (uint32_t e = 0; e < posts.size(); ++e) post_rank[posts[e]] = (int32_t)e;
        (void)post_sp;
        for (uint32_t i = 0; i < np; ++i) {
            float* row = &M[(size_t)i * nq];
            if (pre_sp[i] == 1) {
                if (P.s[i] != 0.0f)
                    for (uint32_t j = 0; j < nq; ++j) {
                        if (i == j) continue;
                        const float w = clampw<float>(P.s[i] * row[j] + (float)ltd_delta(P, i, j), (float)P.wmin, (float)P.wmax);
                        row[j] = w / P.s[i];
                    }
                double rs = 0.0;
                for (uint32_t q = 0; q < ntiles; ++q) rs += tile_sum(q, nq, [&](uint32_t j) { return (double)row[j]; });
                P.Rs[i] = rs;
            } else {
                double rs = P.Rs[i];
                for (uint32_t q = 0; q < ntiles; ++q)
                    rs += tile_sum(q, nq, [&](uint32_t j) {
                        const int32_t e = post_rank[j];
                        return e < 0 ? 0.0 : change_[(size_t)e * np + i];
                    });
                P.Rs[i] = rs;
            }
            if (P.Rs[i] > 0.0) P.s[i] = (float)((double)P.target[i] / P.Rs[i]);
        }
    }

    // The resident step's order of the lazy row sums (see oracle.h): blocks of C post columns.
    void lazy_resident_sums(Plast& P, float* M, uint32_t np, uint32_t nq, const uint32_t* pre_sp,
                            const uint32_t* post_sp, const std::vector<uint32_t>& posts) {
        const uint32_t Cb = p.resident_columns;
        const uint32_t nblk = (nq + Cb - 1) / Cb;
        std::vector<int32_t> post_rank(nq, -1);
        for (uint32_t e = 0; e < posts.size(); ++e) post_rank[posts[e]] = (int32_t)e;
        for (uint32_t i = 0; i < np; ++i) {
            float* row = &M[(size_t)i * nq];
            if (pre_sp[i] == 1) {
                if (P.s[i] != 0.0f)
                    for (uint32_t j = 0; j < nq; ++j) {
                        if (i == j) continue;
                        const float w = clampw<float>(P.s[i] * row[j] + (float)ltd_delta(P, i, j), (float)P.wmin, (float)P.wmax);
                        row[j] = w / P.s[i];
                    }
                // Block parts P_b; lane l of 64 sums P_l, P_{l+64}, ... from 0.0; xor butterfly.
                double lane[64];
                for (uint32_t l = 0; l < 64; ++l) {
                    double acc = 0.0;
                    for (uint32_t b = l; b < nblk; b += 64) {
                        double pb = 0.0;
                        for (uint32_t j = b * Cb; j < std::min(b * Cb + Cb, nq); ++j) pb += (double)row[j];
                        acc += pb;
                    }
                    lane[l] = acc;
                }
                P.Rs[i] = butterfly64(lane);
            } else {
                double rs = P.Rs[i];
                for (uint32_t b = 0; b < nblk; ++b) {
                    bool any = false;
                    double db = 0.0;
                    for (uint32_t j = b * Cb; j < std::min(b * Cb + Cb, nq); ++j)
                        if (post_sp[j] == 1) {
                            any = true;
                            db += change_[(size_t)post_rank[j] * np + i];
                        }
                    if (any) rs += db;
                }
                P.Rs[i] = rs;
            }
            if (P.Rs[i] > 0.0) P.s[i] = (float)((double)P.target[i] / P.Rs[i]);
        }
    }

    int run_step() {
        evolve_group(grp[0]);
        evolve_group(grp[1]);
        for (int c = 0; c < 4; ++c) propagate(c);
        const int rc = stimulus();
        for (Plast& P : plast) plasticity(P);
        ++step;
        return rc;
    }
};

}  // namespace

template <typename R, typename T>
void copy_out(const std::vector<R>& v, T* dst) {
    if (!dst) return;
    for (size_t k = 0; k < v.size(); ++k) dst[k] = (T)v[k];
}

struct oracle_sim {
    int precision;
    Sim<float> f;
    Sim<double> d;
};

extern "C" {

oracle_sim* oracle_create(const oracle_params* p, int precision, const float* w_ee, const float* w_ei,
                          const float* w_ie, const float* w_ii, const uint32_t* stim_counts, uint32_t num_timesteps,
                          const uint32_t* stim_ids, const uint32_t* stim_offsets, const float* stim_amps,
                          uint32_t num_events) {
    oracle_sim* s = new oracle_sim();
    s->precision = precision;
    const float* w[4] = {w_ee, w_ei, w_ie, w_ii};
    auto fill = [&](auto& S) {
        S.init(*p, w);
        if (stim_counts) S.stim_counts.assign(stim_counts, stim_counts + num_timesteps);
        if (num_events) {
            S.stim_ids.assign(stim_ids, stim_ids + num_events);
            S.stim_offs.assign(stim_offsets, stim_offsets + num_events);
            S.stim_amps.assign(stim_amps, stim_amps + num_events);
        }
    };
    if (precision == 64) fill(s->d);
    else fill(s->f);
    return s;
}

void oracle_destroy(oracle_sim* s) { delete s; }

int oracle_step(oracle_sim* s, uint32_t n) {
    int rc = 0;
    for (uint32_t k = 0; k < n; ++k) rc |= s->precision == 64 ? s->d.run_step() : s->f.run_step();
    return rc;
}

uint32_t oracle_clock_step(const oracle_sim* s) { return s->precision == 64 ? s->d.step : s->f.step; }

void oracle_get_state(const oracle_sim* s, float* v, float* ge, float* gi, float* istim, uint32_t* refr,
                      uint32_t* spikes, float* bg_ge, float* bg_gi) {
    auto get = [&](const auto& S) {
        copy_out(S.V, v), copy_out(S.ge, ge), copy_out(S.gi, gi), copy_out(S.istim, istim);
        copy_out(S.refr, refr), copy_out(S.spikes, spikes), copy_out(S.bg_ge, bg_ge), copy_out(S.bg_gi, bg_gi);
    };
    if (s->precision == 64) get(s->d);
    else get(s->f);
}

void oracle_get_state64(const oracle_sim* s, double* v, double* ge, double* gi) {
    auto get = [&](const auto& S) { copy_out(S.V, v), copy_out(S.ge, ge), copy_out(S.gi, gi); };
    if (s->precision == 64) get(s->d);
    else get(s->f);
}

void oracle_get_weights(const oracle_sim* s, uint32_t idx, float* dst) {
    auto get = [&](const auto& S) {
        const auto& C = S.conn[idx];
        for (const auto& P : S.plast)
            if (P.lazy && P.c == idx) {
                for (uint32_t i = 0; i < C.n_pre; ++i)
                    for (uint32_t j = 0; j < C.n_post; ++j)
                        dst[(size_t)i * C.n_post + j] = P.s[i] * (float)C.W[(size_t)i * C.n_post + j];
                return;
            }
        copy_out(C.W, dst);
    };
    if (s->precision == 64) get(s->d);
    else get(s->f);
}

void oracle_get_connection_traces(const oracle_sim* s, uint32_t idx, float* r, float* d, float* x) {
    auto get = [&](const auto& S) { copy_out(S.conn[idx].r, r), copy_out(S.conn[idx].d, d), copy_out(S.conn[idx].x, x); };
    if (s->precision == 64) get(s->d);
    else get(s->f);
}

void oracle_get_plasticity_traces(const oracle_sim* s, uint32_t off, float* pre, float* post, float* tpre,
                                  float* tpost) {
    auto get = [&](const auto& S) {
        const uint32_t N = S.N;
        for (float* a : {pre, post, tpre, tpost})
            if (a) std::fill(a, a + N, 0.0f);
        if (off >= S.plast.size()) return;
        const auto& P = S.plast[off];
        const auto& C = S.conn[P.c];
        if (pre) copy_out(P.pre, pre + C.pre_off);
        if (post) copy_out(P.post, post + C.post_off);
        if (tpre) copy_out(P.tpre, tpre + C.pre_off);
        if (tpost) copy_out(P.tpost, tpost + C.post_off);
    };
    if (s->precision == 64) get(s->d);
    else get(s->f);
}

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox(c, key[0], key[1]);
    std::memcpy(out, c, sizeof c);
}

float oracle_expf(float x) { return expf_engine(x); }

uint32_t oracle_poisson_table(float mean, uint32_t* table, uint32_t capacity) {
    return poisson_table(mean, table, capacity);
}

}  // extern "C"
