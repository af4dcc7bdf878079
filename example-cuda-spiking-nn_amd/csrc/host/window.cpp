// Differentiable window (bn_window_* of include/buryn_amd.h): the plasticity-free step run for
// T steps from a snapshot of a bn_sim, the window loss, and its gradient by the adjoint with
// checkpoints every S steps (forward stores the state at each segment start; the reverse pass
// recomputes a segment with per-step records, walks it backwards, then adds the segment's
// x (outer) lambda_g products to dW).  The forward records the last segment itself, and the
// records of the other segments are doubled, so that the recompute of segment s-1 runs on a
// second stream while segment s is walked backwards: two independent chains of launches in one
// graph, which share only read-only data (weights, parameters).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../bn_window.h"
#include "api_util.h"
#include "network.h"
#include "sim_internal.h"

using namespace bn;

namespace {
uint32_t round_up_w(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }
}  // namespace

struct bn_window {
    Config cfg;
    uint32_t NE = 0, NI = 0, N = 0, T = 0, S = 0, n_seg = 0, nv = 0;
    uint32_t start_step = 0;
    float rate_tau = 5e-3f, beta = 1e-3f;
    float g_leak = 0.0f, tau_decay_e = 0.0f;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;        // the backward's recompute chain (captured into the graph)
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<void*> allocs;
    float* W[4] = {};
    uint32_t ld[4] = {}, pre[4] = {}, post[4] = {};
    WinGroup grp[2]{};
    WinState st{}, snap{};
    std::vector<WinState> ckpt;
    WinRecords rec[3]{};               // records of segment s: rec_of(s)
    GemvPlan gemv{};
    bn_device_clock clock{};
    int32_t* vrec_slot = nullptr;
    std::vector<uint32_t> vrec_ids;
    uint32_t* d_counts = nullptr;      // [T][2]
    float* d_vrec = nullptr;           // [T][nv]
    uint8_t* d_raster = nullptr;       // [T][N]
    double* d_dv = nullptr;            // [T][nv]
    double* d_lrho = nullptr;          // [T][2]
    WinParams* d_params = nullptr;
    // Forward (T steps with checkpoints) and backward (every segment's recompute and reverse
    // walk) captured once per window; parameters, targets and loss terms live in device
    // memory, so the graphs stay valid.
    hipGraph_t fwd_graph = nullptr, bwd_graph = nullptr;  // kept for the executables' lifetime
    hipGraphExec_t fwd_exec = nullptr, bwd_exec = nullptr;
    float* staging = nullptr;           // pinned host buffer for gradient downloads
    hipEvent_t ev[4] = {};              // around the last forward and backward graph launches
    float fwd_ms = 0.0f, bwd_ms = 0.0f;
    double *carry_V = nullptr, *carry_d = nullptr, *carry_r = nullptr;
    double *lg_e = nullptr, *lg_i = nullptr;  // [2][N]: lambda g of step t in buffer t % 2
    double *acc_gleak = nullptr, *acc_alpha = nullptr;
    float* dW[4] = {};
    std::vector<double> target_rates, target_v, lrho;  // [T][2], [T][nv], [T][2]
    std::vector<double> rates;
    double loss = 0.0;
    bool have_forward = false;

    template <typename T_>
    T_* alloc(size_t n, const char* what) {
        void* p = nullptr;
        BN_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T_)), std::string("hipMalloc ") + what);
        allocs.push_back(p);
        BN_HIP_CHECK(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T_), stream), "hipMemsetAsync");
        return static_cast<T_*>(p);
    }
    ~bn_window() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (staging) (void)hipHostFree(staging);
        if (fwd_exec) (void)hipGraphExecDestroy(fwd_exec);
        if (bwd_exec) (void)hipGraphExecDestroy(bwd_exec);
        if (fwd_graph) (void)hipGraphDestroy(fwd_graph);
        if (bwd_graph) (void)hipGraphDestroy(bwd_graph);
        for (void* p : allocs) (void)hipFree(p);
        for (hipEvent_t e : {fork, join})
            if (e) (void)hipEventDestroy(e);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    WinState alloc_state(const char* what) {
        WinState s{};
        s.V = alloc<float>(N, what);
        s.refr = alloc<uint32_t>(N, what);
        s.bg_ge = alloc<float>(N, what);
        s.bg_gi = alloc<float>(N, what);
        s.bre = alloc<float>(N, what);
        s.bde = alloc<float>(N, what);
        s.bri = alloc<float>(N, what);
        s.bdi = alloc<float>(N, what);
        s.dE = alloc<float>(NE, what);
        s.rE = alloc<float>(NE, what);
        s.xE = alloc<float>(NE, what);
        s.dI = alloc<float>(NI, what);
        s.rI = alloc<float>(NI, what);
        s.xI = alloc<float>(NI, what);
        return s;
    }
    void copy_state(const WinState& dst, const WinState& src, hipStream_t s) {
        WinCopyPlan p{dst, src, NE, NI, N};
        BN_HIP_CHECK((hipError_t)launch_win_copy(p, s), "window state copy launch");
    }
    void copy_state(const WinState& dst, const WinState& src) { copy_state(dst, src, stream); }

    void apply_params() {
        WinParams h{};
        h.g_leak = g_leak;
        h.decay_e_d = std::exp(-1.0f * cfg.timestep / tau_decay_e);  // connection.cpp:46-47
        BN_HIP_CHECK(hipMemcpyAsync(d_params, &h, sizeof h, hipMemcpyHostToDevice, stream), "window params");
        BN_HIP_CHECK(hipStreamSynchronize(stream), "window params sync");
    }
    // The last segment's records are written by the forward and read by every backward pass;
    // the other segments are recomputed into two alternating sets.
    const WinRecords& rec_of(uint32_t seg) const { return seg + 1 == n_seg ? rec[2] : rec[seg & 1]; }
    // record: per-step records for the backward; outputs: spike counts, recorded V, raster.
    WinForwardPlan fwd_plan(uint32_t t, bool record, bool outputs) const {
        WinForwardPlan p{};
        p.g[0] = grp[0];
        p.g[1] = grp[1];
        p.st = st;
        if (record) p.rec = rec_of(t / S);
        p.prm = d_params;
        p.NE = NE, p.NI = NI, p.N = N;
        p.slot = t % S;
        p.step = start_step + t;
        p.decay_e_r = std::exp(-1.0f * cfg.timestep / cfg.tau_rise_e);  // connection.cpp:46-47
        p.decay_i_r = std::exp(-1.0f * cfg.timestep / cfg.tau_rise_i);
        p.decay_i_d = std::exp(-1.0f * cfg.timestep / cfg.tau_decay_i);
        p.K = cfg.postsynaptic_kernel_normalization_value;
        p.clock = clock;
        p.vrec_slot = vrec_slot;
        if (outputs) {
            p.counts = d_counts + 2 * (size_t)t;
            p.v_rec = d_vrec + (size_t)nv * t;
            p.raster = d_raster + (size_t)N * t;
        }
        return p;
    }
    void step_forward(uint32_t t, bool record, bool outputs, hipStream_t s) {
        BN_HIP_CHECK((hipError_t)launch_gemv(gemv, s), "window propagate launch");
        BN_HIP_CHECK((hipError_t)launch_win_forward(fwd_plan(t, record, outputs), s), "window forward launch");
    }
    // Segment seg from its checkpoint, with per-step records into rec_of(seg).
    void recompute(uint32_t seg, hipStream_t s) {
        const uint32_t t0 = seg * S, t1 = std::min(T, t0 + S);
        copy_state(st, ckpt[seg], s);
        for (uint32_t t = t0; t < t1; ++t) step_forward(t, true, false, s);
    }

    template <typename F>
    hipGraphExec_t capture(hipGraph_t& graph, F&& body) {
        BN_HIP_CHECK(hipStreamSynchronize(stream), "pre-capture sync");
        BN_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            body();
        } catch (...) {
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            throw;
        }
        BN_HIP_CHECK(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        BN_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        return exec;
    }

    void forward() {
        apply_params();
        BN_HIP_CHECK(hipMemsetAsync(d_counts, 0, 2 * (size_t)T * sizeof(uint32_t), stream), "clear counts");
        if (!fwd_exec)
            fwd_exec = capture(fwd_graph, [&] {
                copy_state(st, snap);
                for (uint32_t t = 0; t < T; ++t) {
                    if (t % S == 0) copy_state(ckpt[t / S], st);
                    step_forward(t, t / S + 1 == n_seg, true, stream);
                }
            });
        BN_HIP_CHECK(hipEventRecord(ev[0], stream), "hipEventRecord");
        BN_HIP_CHECK(hipGraphLaunch(fwd_exec, stream), "window forward graph");
        BN_HIP_CHECK(hipEventRecord(ev[1], stream), "hipEventRecord");
        std::vector<uint32_t> counts(2 * (size_t)T);
        std::vector<float> vrec(std::max<size_t>(1, (size_t)T * nv));
        BN_HIP_CHECK(hipMemcpyAsync(counts.data(), d_counts, counts.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream), "counts");
        if (nv) BN_HIP_CHECK(hipMemcpyAsync(vrec.data(), d_vrec, (size_t)T * nv * sizeof(float), hipMemcpyDeviceToHost, stream), "v");
        BN_HIP_CHECK(hipStreamSynchronize(stream), "forward sync");
        BN_HIP_CHECK(hipEventElapsedTime(&fwd_ms, ev[0], ev[1]), "hipEventElapsedTime");
        // Loss (FP64): smoothed population rates and recorded voltages against the targets.
        const double dt = cfg.timestep, a = std::exp(-dt / (double)rate_tau);
        rates.assign(2 * (size_t)T, 0.0);
        double rho[2] = {0.0, 0.0}, L = 0.0;
        std::vector<double> dv((size_t)T * nv);
        for (uint32_t t = 0; t < T; ++t) {
            for (int q = 0; q < 2; ++q) {
                rho[q] = a * rho[q] + (1.0 - a) * ((double)counts[2 * t + q] / (q == 0 ? NE : NI)) / dt;
                rates[2 * t + q] = rho[q];
                const double e = rho[q] - target_rates[2 * t + q];
                L += e * e / T;
            }
            for (uint32_t k = 0; k < nv; ++k) {
                const double e = ((double)vrec[(size_t)t * nv + k] - target_v[(size_t)t * nv + k]) * 1e3;
                L += e * e / ((double)T * nv);
                dv[(size_t)t * nv + k] = 2.0 / ((double)T * nv) * e * 1e3;
            }
        }
        loss = L;
        lrho.assign(2 * (size_t)T, 0.0);
        double carry[2] = {0.0, 0.0};
        for (int64_t t = (int64_t)T - 1; t >= 0; --t)
            for (int q = 0; q < 2; ++q) {
                carry[q] = 2.0 / T * (rates[2 * t + q] - target_rates[2 * t + q]) + a * carry[q];
                lrho[2 * t + q] = carry[q];
            }
        BN_HIP_CHECK(hipMemcpyAsync(d_lrho, lrho.data(), lrho.size() * sizeof(double), hipMemcpyHostToDevice, stream), "lrho");
        if (nv) BN_HIP_CHECK(hipMemcpyAsync(d_dv, dv.data(), dv.size() * sizeof(double), hipMemcpyHostToDevice, stream), "dv");
        BN_HIP_CHECK(hipStreamSynchronize(stream), "loss terms sync");
        have_forward = true;
    }

    void enqueue_backward() {
        const double dt = cfg.timestep, a = std::exp(-dt / (double)rate_tau);
        for (int64_t seg = (int64_t)n_seg - 1; seg >= 0; --seg) {
            const uint32_t t0 = (uint32_t)seg * S, t1 = std::min(T, t0 + S), slots = t1 - t0;
            const WinRecords& rec = rec_of((uint32_t)seg);
            // Segment seg-1 is recomputed on the side stream meanwhile (the last segment was
            // recorded by the forward).  It writes the state, the propagate partials and
            // rec_of(seg - 1); the fork comes after segment seg+1's reverse walk and dW update
            // (the last readers of those records) and after segment seg's recompute (the last
            // user of the state and partials).
            if (seg > 0) {
                BN_HIP_CHECK(hipEventRecord(fork, stream), "backward fork");
                BN_HIP_CHECK(hipStreamWaitEvent(side, fork, 0), "backward fork wait");
                recompute((uint32_t)seg - 1, side);
                BN_HIP_CHECK(hipEventRecord(join, side), "backward join");
            }
            for (int64_t sl = (int64_t)slots - 1; sl >= 0; --sl) {
                const uint32_t t = t0 + (uint32_t)sl;
                WinReversePlan rp{};
                rp.g[0] = grp[0];
                rp.g[1] = grp[1];
                rp.rec = rec;
                rp.prm = d_params;
                for (int c = 0; c < 4; ++c) rp.W[c] = W[c], rp.ld[c] = ld[c];
                rp.NE = NE, rp.NI = NI, rp.N = N, rp.slot = (uint32_t)sl;
                const WinForwardPlan fp = fwd_plan(t, true, false);
                rp.decay_e_r = fp.decay_e_r, rp.decay_i_r = fp.decay_i_r, rp.decay_i_d = fp.decay_i_d;
                rp.K = fp.K;
                rp.theta = cfg.v_threshold;
                rp.beta = beta;
                rp.rate_coef_e = (1.0 - a) / (NE * dt);
                rp.rate_coef_i = (1.0 - a) / (NI * dt);
                rp.lrho = d_lrho + 2 * (size_t)t;
                rp.dv_direct = d_dv + (size_t)nv * t;
                rp.vrec_slot = vrec_slot;
                const size_t in = (size_t)((t + 1) & 1) * N, out = (size_t)(t & 1) * N;
                rp.lg_in_e = lg_e + in, rp.lg_in_i = lg_i + in;
                rp.lg_out_e = lg_e + out, rp.lg_out_i = lg_i + out;
                rp.carry_V = carry_V, rp.carry_d = carry_d, rp.carry_r = carry_r;
                rp.acc_gleak = acc_gleak, rp.acc_alpha = acc_alpha;
                BN_HIP_CHECK((hipError_t)launch_win_reverse(rp, stream), "window reverse launch");
            }
            // dW_c += sum over the segment of x_{t-1} (outer) lambda g_t.
            for (int c = 0; c < 4; ++c) {
                WinOuterPlan op{};
                const bool e_pre = c < 2, e_post = c % 2 == 0;
                op.dW = dW[c];
                op.ld = ld[c];
                op.n_pre = pre[c];
                op.n_post = post[c];
                op.slots = slots;
                op.d = e_pre ? rec.dE : rec.dI;
                op.r = e_pre ? rec.rE : rec.rI;
                op.pre_stride = e_pre ? NE : NI;
                op.pre_offset = 0;
                op.lam = e_pre ? rec.le : rec.li;
                op.lam_stride = N;
                op.post_offset = e_post ? 0 : NE;
                BN_HIP_CHECK((hipError_t)launch_win_outer(op, stream), "window outer launch");
            }
            if (seg > 0) BN_HIP_CHECK(hipStreamWaitEvent(stream, join, 0), "backward join wait");
        }
    }

    void backward(double* dgl, double* dtau) {
        if (!have_forward) throw Error(BN_ERR_RUNTIME, "bn_window_backward needs a forward first");
        for (int c = 0; c < 4; ++c)
            BN_HIP_CHECK(hipMemsetAsync(dW[c], 0, (size_t)pre[c] * ld[c] * sizeof(float), stream), "clear dW");
        for (double* p : {carry_V, carry_d, carry_r})
            BN_HIP_CHECK(hipMemsetAsync(p, 0, N * sizeof(double), stream), "clear adjoint");
        for (double* p : {lg_e, lg_i}) BN_HIP_CHECK(hipMemsetAsync(p, 0, 2 * N * sizeof(double), stream), "clear adjoint");
        BN_HIP_CHECK(hipMemsetAsync(acc_gleak, 0, N * sizeof(double), stream), "clear acc");
        BN_HIP_CHECK(hipMemsetAsync(acc_alpha, 0, NE * sizeof(double), stream), "clear acc");
        if (!bwd_exec) bwd_exec = capture(bwd_graph, [&] { enqueue_backward(); });
        BN_HIP_CHECK(hipEventRecord(ev[2], stream), "hipEventRecord");
        BN_HIP_CHECK(hipGraphLaunch(bwd_exec, stream), "window backward graph");
        BN_HIP_CHECK(hipEventRecord(ev[3], stream), "hipEventRecord");
        const double dt = cfg.timestep;
        std::vector<double> ag(N), aa(NE);
        BN_HIP_CHECK(hipMemcpyAsync(ag.data(), acc_gleak, N * sizeof(double), hipMemcpyDeviceToHost, stream), "acc");
        BN_HIP_CHECK(hipMemcpyAsync(aa.data(), acc_alpha, NE * sizeof(double), hipMemcpyDeviceToHost, stream), "acc");
        BN_HIP_CHECK(hipStreamSynchronize(stream), "backward sync");
        BN_HIP_CHECK(hipEventElapsedTime(&bwd_ms, ev[2], ev[3]), "hipEventElapsedTime");
        double sg = 0.0, sa = 0.0;
        for (double v : ag) sg += v;
        for (double v : aa) sa += v;
        const double ad = std::exp(-dt / (double)tau_decay_e);
        if (dgl) *dgl = sg;
        if (dtau) *dtau = sa * ad * dt / ((double)tau_decay_e * tau_decay_e);
    }
};

namespace {

void build_window(bn_window* w, bn_sim* sim, const bn_window_spec* spec) {
    const Config& c = sim->cfg;
    if (c.e_e_plasticity_enabled || c.i_e_plasticity_enabled)
        throw Error(BN_ERR_INVALID_ARGUMENT, "the differentiable window needs a plasticity-free network");
    if (spec->steps == 0 || spec->checkpoint_interval == 0)
        throw Error(BN_ERR_INVALID_ARGUMENT, "window steps and checkpoint interval must be positive");
    w->cfg = c;
    w->NE = c.num_excitatory_neurons, w->NI = c.num_inhibitory_neurons, w->N = w->NE + w->NI;
    w->T = spec->steps, w->S = spec->checkpoint_interval, w->n_seg = (w->T + w->S - 1) / w->S;
    w->rate_tau = spec->rate_tau, w->beta = spec->surrogate_beta;
    w->g_leak = c.g_leak, w->tau_decay_e = c.tau_decay_e;
    const uint32_t N = w->N, NE = w->NE, NI = w->NI, T = w->T, S = w->S;
    BN_HIP_CHECK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking), "hipStreamCreate");
    BN_HIP_CHECK(hipStreamCreateWithFlags(&w->side, hipStreamNonBlocking), "hipStreamCreate");
    for (hipEvent_t& e : w->ev) BN_HIP_CHECK(hipEventCreate(&e), "hipEventCreate");
    for (hipEvent_t* e : {&w->fork, &w->join})
        BN_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
    // Weights: the simulation's current ones, rows padded to 64 floats.
    const uint32_t pres[4] = {NE, NE, NI, NI}, posts[4] = {NE, NI, NE, NI};
    for (int k = 0; k < 4; ++k) {
        w->pre[k] = pres[k], w->post[k] = posts[k], w->ld[k] = round_up_w(posts[k], 64);
        w->W[k] = w->alloc<float>((size_t)pres[k] * w->ld[k], "window weights");
        w->dW[k] = w->alloc<float>((size_t)pres[k] * w->ld[k], "window dW");
        std::vector<float> h((size_t)pres[k] * posts[k]);
        BN_CALL(bn_sim_get_weights(sim, k, h.data()));
        BN_HIP_CHECK(hipMemcpy2DAsync(w->W[k], w->ld[k] * sizeof(float), h.data(), posts[k] * sizeof(float),
                                      posts[k] * sizeof(float), pres[k], hipMemcpyHostToDevice, w->stream), "upload W");
        BN_HIP_CHECK(hipStreamSynchronize(w->stream), "upload sync");
    }
    // Groups: parameters and Poisson tables.
    for (int g = 0; g < 2; ++g) {
        const NeuronGroupParams p = group_params(c, g == 1);
        NeuronGroupView& v = w->grp[g].v;
        std::memset(&v, 0, sizeof v);
        v.timestep = p.timestep, v.g_leak = p.g_leak, v.membrane_capacitance = p.membrane_capacitance;
        v.depolarization_slope_factor = p.depolarization_slope_factor;
        v.v_effective_spike_threshold = p.v_effective_spike_threshold;
        v.v_excitatory_reversal = p.v_excitatory_reversal, v.v_inhibitory_reversal = p.v_inhibitory_reversal;
        v.v_threshold = p.v_threshold, v.v_reset = p.v_reset, v.v_rest = p.v_rest;
        v.refractory_steps = p.refractory_steps, v.is_inhibitory = p.is_inhibitory;
        v.mean_background_excitatory_spikes = p.mean_background_excitatory_spikes;
        v.mean_background_inhibitory_spikes = p.mean_background_inhibitory_spikes;
        v.background_excitatory_postsynaptic_r_trace_decay = p.background_excitatory_postsynaptic_r_trace_decay;
        v.background_excitatory_postsynaptic_d_trace_decay = p.background_excitatory_postsynaptic_d_trace_decay;
        v.background_inhibitory_postsynaptic_r_trace_decay = p.background_inhibitory_postsynaptic_r_trace_decay;
        v.background_inhibitory_postsynaptic_d_trace_decay = p.background_inhibitory_postsynaptic_d_trace_decay;
        v.background_excitatory_scaling_factor = p.background_excitatory_scaling_factor;
        v.background_inhibitory_scaling_factor = p.background_inhibitory_scaling_factor;
        v.num_neurons = p.num_neurons, v.offset = p.offset;
        for (int kind = 0; kind < 2; ++kind) {
            std::vector<uint32_t> t(256);
            uint32_t len = 0;
            BN_CALL(bn_poisson_table(kind == 0 ? p.mean_background_excitatory_spikes : p.mean_background_inhibitory_spikes,
                                     t.data(), 256, &len));
            uint32_t* d = nullptr;
            if (len) {
                d = w->alloc<uint32_t>(len, "poisson table");
                BN_HIP_CHECK(hipMemcpyAsync(d, t.data(), len * sizeof(uint32_t), hipMemcpyHostToDevice, w->stream), "table");
                BN_HIP_CHECK(hipStreamSynchronize(w->stream), "table sync");
            }
            (kind == 0 ? v.poisson_table_excitatory : v.poisson_table_inhibitory) = d;
            (kind == 0 ? v.poisson_table_excitatory_len : v.poisson_table_inhibitory_len) = len;
        }
    }
    // State: snapshot of the simulation, the running state and one checkpoint per segment.
    w->snap = w->alloc_state("window snapshot");
    w->st = w->alloc_state("window state");
    for (uint32_t k = 0; k < w->n_seg; ++k) w->ckpt.push_back(w->alloc_state("window checkpoint"));
    {
        std::vector<float> V(N), bge(N), bgi(N), bre(N), bde(N), bri(N), bdi(N);
        std::vector<uint32_t> refr(N);
        bn_neuron_state_host h{};
        h.membrane_potentials = V.data();
        h.refractory_counter = refr.data();
        h.background_g_excitatory_synaptic_conductances = bge.data();
        h.background_g_inhibitory_synaptic_conductances = bgi.data();
        h.background_excitatory_postsynaptic_r_traces = bre.data();
        h.background_excitatory_postsynaptic_d_traces = bde.data();
        h.background_inhibitory_postsynaptic_r_traces = bri.data();
        h.background_inhibitory_postsynaptic_d_traces = bdi.data();
        BN_CALL(bn_sim_get_neuron_state(sim, &h));
        std::vector<float> rE(NE), dE(NE), xE(NE), rI(NI), dI(NI), xI(NI);
        BN_CALL(bn_sim_get_connection_traces(sim, 0, rE.data(), dE.data(), xE.data()));
        BN_CALL(bn_sim_get_connection_traces(sim, 2, rI.data(), dI.data(), xI.data()));
        auto up = [&](void* d, const void* s, size_t bytes) {
            BN_HIP_CHECK(hipMemcpyAsync(d, s, bytes, hipMemcpyHostToDevice, w->stream), "snapshot upload");
        };
        const WinState& s = w->snap;
        up(s.V, V.data(), N * 4), up(s.refr, refr.data(), N * 4), up(s.bg_ge, bge.data(), N * 4);
        up(s.bg_gi, bgi.data(), N * 4), up(s.bre, bre.data(), N * 4), up(s.bde, bde.data(), N * 4);
        up(s.bri, bri.data(), N * 4), up(s.bdi, bdi.data(), N * 4);
        up(s.dE, dE.data(), NE * 4), up(s.rE, rE.data(), NE * 4), up(s.xE, xE.data(), NE * 4);
        up(s.dI, dI.data(), NI * 4), up(s.rI, rI.data(), NI * 4), up(s.xI, xI.data(), NI * 4);
        BN_HIP_CHECK(hipStreamSynchronize(w->stream), "snapshot sync");
        bn_sim_info info{};
        BN_CALL(bn_sim_get_info(sim, &info));
        // Before the first step the stored conductances are the initial ones, not a
        // background + propagate split.
        if (info.clock_step == 0) throw Error(BN_ERR_INVALID_ARGUMENT, "bn_window_create: run the simulation at least one step first");
        w->start_step = info.clock_step;
    }
    const int64_t key = c.background_noise_seed >= 0 ? (int64_t)c.background_noise_seed : (int64_t)c.rng_seed;
    w->clock.rng_key[0] = (uint32_t)key;
    w->clock.rng_key[1] = (uint32_t)((uint64_t)key >> 32);
    // Propagate plan over the window's weights and traces, with the simulation's row blocks
    // (the same FP64 partial sums, so the window's first steps equal the plain step's).
    const float* xs[4] = {w->st.xE, w->st.xE, w->st.xI, w->st.xI};
    // Part sums where every matrix allows them (rows padded to 16 bytes, at most
    // kGemvPartChunksMax chunks per part): the propagate combines each part of a column's chunks
    // itself, in the same order, so the forward reads at most kFinParts sums per source.
    bool parts = true;
    for (int k = 0; k < 4; ++k) {
        const uint32_t rb = sim->state->gemv_row_block(k);
        const uint32_t chunks = (pres[k] + rb - 1) / rb;
        const uint32_t per = (chunks + kFinParts - 1) / kFinParts;
        if (w->ld[k] % 4 != 0 || (uintptr_t)w->W[k] % 16 != 0 || per > kGemvPartChunksMax) parts = false;
    }
    auto part_count = [](uint32_t chunks) {
        const uint32_t per = (chunks + kFinParts - 1) / kFinParts;
        return per ? (chunks + per - 1) / per : 0u;
    };
    uint32_t tiles = 0;
    for (int k = 0; k < 4; ++k) {
        GemvConn& C = w->gemv.c[k];
        C.W = w->W[k];
        C.x = xs[k];
        C.row_scale = nullptr;
        C.ld = w->ld[k];
        C.n_pre = pres[k];
        C.n_post = posts[k];
        C.row_block = sim->state->gemv_row_block(k);
        C.n_chunks = (pres[k] + C.row_block - 1) / C.row_block;
        C.partials = w->alloc<double>((size_t)C.n_chunks * C.ld, "window partials");
        C.n_coltiles = (posts[k] + kGemvTileCols - 1) / kGemvTileCols;
        C.tile_begin = tiles;
        tiles += (parts ? part_count(C.n_chunks) : C.n_chunks) * C.n_coltiles;
        C.tile_end = tiles;
    }
    w->gemv.n_conn = 4;
    w->gemv.parts = parts ? 1u : 0u;
    w->gemv.total_tiles = tiles;
    w->gemv.nontemporal = 0;
    // (part sums: one partials row per part, each combined as a part of one chunk)
    auto src = [&](int k) {
        const GemvConn& C = w->gemv.c[k];
        return FinalizeSource{C.partials, parts ? part_count(C.n_chunks) : C.n_chunks, w->ld[k]};
    };
    w->grp[0].fin_e = src(0);
    w->grp[0].fin_i = src(2);
    w->grp[1].fin_e = src(1);
    w->grp[1].fin_i = src(3);
    // Records of three segments (rec_of), outputs, adjoint buffers.
    for (WinRecords& r : w->rec) {
        r.Vprev = w->alloc<float>((size_t)S * N, "records");
        r.ge = w->alloc<float>((size_t)S * N, "records");
        r.gi = w->alloc<float>((size_t)S * N, "records");
        r.s = w->alloc<float>((size_t)S * N, "records");
        r.refr = w->alloc<uint8_t>((size_t)S * N, "records");
        r.dE = w->alloc<float>((size_t)S * NE, "records");
        r.rE = w->alloc<float>((size_t)S * NE, "records");
        r.dI = w->alloc<float>((size_t)S * NI, "records");
        r.rI = w->alloc<float>((size_t)S * NI, "records");
        r.le = w->alloc<float>((size_t)S * N, "records");
        r.li = w->alloc<float>((size_t)S * N, "records");
    }
    w->nv = spec->n_vrec;
    std::vector<int32_t> slot(N, -1);
    for (uint32_t k = 0; k < w->nv; ++k) {
        const uint32_t id = spec->vrec_ids[k];
        if (id >= NE) throw Error(BN_ERR_OUT_OF_RANGE, "recorded ids must be excitatory");
        slot[id] = (int32_t)k;
        w->vrec_ids.push_back(id);
    }
    w->vrec_slot = w->alloc<int32_t>(N, "vrec slots");
    BN_HIP_CHECK(hipMemcpyAsync(w->vrec_slot, slot.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, w->stream), "slots");
    w->d_counts = w->alloc<uint32_t>(2 * (size_t)T, "counts");
    w->d_vrec = w->alloc<float>((size_t)T * std::max(1u, w->nv), "v");
    w->d_lrho = w->alloc<double>(2 * (size_t)T, "rate adjoint");
    w->d_params = w->alloc<WinParams>(1, "window params");
    w->d_raster = w->alloc<uint8_t>((size_t)T * N, "raster");
    w->d_dv = w->alloc<double>((size_t)T * std::max(1u, w->nv), "dv");
    w->carry_V = w->alloc<double>(N, "adjoint");
    w->carry_d = w->alloc<double>(N, "adjoint");
    w->carry_r = w->alloc<double>(N, "adjoint");
    w->lg_e = w->alloc<double>(2 * (size_t)N, "adjoint");
    w->lg_i = w->alloc<double>(2 * (size_t)N, "adjoint");
    w->acc_gleak = w->alloc<double>(N, "adjoint");
    w->acc_alpha = w->alloc<double>(NE, "adjoint");
    w->target_rates.assign(2 * (size_t)T, 0.0);
    w->target_v.assign((size_t)T * w->nv, 0.0);
    BN_HIP_CHECK(hipStreamSynchronize(w->stream), "window setup sync");
}

}  // namespace

extern "C" {

int bn_window_create(bn_sim* sim, const bn_window_spec* spec, bn_window** out) {
    if (!sim || !spec || !out) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_create: null argument");
    *out = nullptr;
    if (sim->system && (sim->system->persistent() || sim->system->resident()))
        return set_error(BN_ERR_INVALID_ARGUMENT,
                         "bn_window_create: the window starts from the graph step's state (not BN_SIM_PERSISTENT or BN_SIM_RESIDENT)");
    auto w = std::make_unique<bn_window>();
    const int rc = guarded([&] { build_window(w.get(), sim, spec); });
    if (rc) return rc;
    *out = w.release();
    return BN_OK;
}

void bn_window_destroy(bn_window* w) { delete w; }

int bn_window_set_params(bn_window* w, const float* const* weights, float g_leak, float tau_decay_e) {
    if (!w) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_set_params: null window");
    if (!(tau_decay_e > 0.0f)) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_set_params: tau_decay_e must be > 0");
    return guarded([&] {
        if (weights)
            for (int k = 0; k < 4; ++k)
                if (weights[k])
                    BN_HIP_CHECK(hipMemcpy2DAsync(w->W[k], w->ld[k] * sizeof(float), weights[k], w->post[k] * sizeof(float),
                                                  w->post[k] * sizeof(float), w->pre[k], hipMemcpyDefault, w->stream),
                                 "set weights");
        w->g_leak = g_leak;
        w->tau_decay_e = tau_decay_e;
        w->have_forward = false;
        BN_HIP_CHECK(hipStreamSynchronize(w->stream), "set params sync");
    });
}

int bn_window_get_weights(bn_window* w, uint32_t idx, float* dst) {
    if (!w || !dst || idx >= 4) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_get_weights: bad argument");
    return guarded([&] {
        BN_HIP_CHECK(hipMemcpy2DAsync(dst, w->post[idx] * sizeof(float), w->W[idx], w->ld[idx] * sizeof(float),
                                      w->post[idx] * sizeof(float), w->pre[idx], hipMemcpyDefault, w->stream), "get weights");
        BN_HIP_CHECK(hipStreamSynchronize(w->stream), "get weights sync");
    });
}

int bn_window_set_targets(bn_window* w, const double* target_rates, const double* target_v) {
    if (!w) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_set_targets: null window");
    if (target_rates) std::memcpy(w->target_rates.data(), target_rates, w->target_rates.size() * sizeof(double));
    else std::fill(w->target_rates.begin(), w->target_rates.end(), 0.0);
    if (target_v) std::memcpy(w->target_v.data(), target_v, w->target_v.size() * sizeof(double));
    else std::fill(w->target_v.begin(), w->target_v.end(), 0.0);
    w->have_forward = false;
    return BN_OK;
}

int bn_window_forward(bn_window* w, double* loss, double* rates, float* v, uint8_t* raster) {
    if (!w) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_forward: null window");
    return guarded([&] {
        w->forward();
        if (loss) *loss = w->loss;
        if (rates) std::memcpy(rates, w->rates.data(), w->rates.size() * sizeof(double));
        if (v && w->nv)
            BN_HIP_CHECK(hipMemcpyAsync(v, w->d_vrec, (size_t)w->T * w->nv * sizeof(float), hipMemcpyDeviceToHost, w->stream), "v");
        if (raster)
            BN_HIP_CHECK(hipMemcpyAsync(raster, w->d_raster, (size_t)w->T * w->N, hipMemcpyDeviceToHost, w->stream), "raster");
        BN_HIP_CHECK(hipStreamSynchronize(w->stream), "forward outputs sync");
    });
}

int bn_window_backward(bn_window* w, float* const* dW, double* dg_leak, double* dtau_decay_e) {
    if (!w) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_backward: null window");
    return guarded([&] {
        w->backward(dg_leak, dtau_decay_e);
        if (dW)
            for (int k = 0; k < 4; ++k)
                if (dW[k])
                    BN_HIP_CHECK(hipMemcpy2DAsync(dW[k], w->post[k] * sizeof(float), w->dW[k], w->ld[k] * sizeof(float),
                                                  w->post[k] * sizeof(float), w->pre[k], hipMemcpyDefault, w->stream),
                                 "dW out");
        BN_HIP_CHECK(hipStreamSynchronize(w->stream), "backward outputs sync");
    });
}

int bn_window_get_gradient(bn_window* w, uint32_t idx, float* dst) {
    if (!w || !dst || idx >= 4) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_get_gradient: bad argument");
    return guarded([&] {
        const size_t n = (size_t)w->pre[idx] * w->post[idx];
        if (!w->staging) {
            size_t most = 0;
            for (int k = 0; k < 4; ++k) most = std::max(most, (size_t)w->pre[k] * w->post[k]);
            BN_HIP_CHECK(hipHostMalloc((void**)&w->staging, std::max<size_t>(most, 1) * sizeof(float), hipHostMallocDefault),
                         "hipHostMalloc");
        }
        BN_HIP_CHECK(hipMemcpy2DAsync(w->staging, w->post[idx] * sizeof(float), w->dW[idx], w->ld[idx] * sizeof(float),
                                      w->post[idx] * sizeof(float), w->pre[idx], hipMemcpyDeviceToHost, w->stream),
                     "gradient download");
        BN_HIP_CHECK(hipStreamSynchronize(w->stream), "gradient download sync");
        std::memcpy(dst, w->staging, n * sizeof(float));
    });
}

int bn_window_get_info(bn_window* w, bn_window_info* info) {
    if (!w || !info) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_window_get_info: null argument");
    std::memset(info, 0, sizeof *info);
    info->steps = w->T;
    info->checkpoint_interval = w->S;
    info->start_step = w->start_step;
    info->total_neurons = w->N;
    info->num_excitatory_neurons = w->NE;
    info->num_inhibitory_neurons = w->NI;
    info->n_vrec = w->nv;
    info->g_leak = w->g_leak;
    info->tau_decay_e = w->tau_decay_e;
    info->forward_ms = w->fwd_ms;
    info->backward_ms = w->bwd_ms;
    return BN_OK;
}

}  // extern "C"
