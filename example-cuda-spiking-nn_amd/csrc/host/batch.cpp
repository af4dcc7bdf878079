// Batched shared-weight engine: B instances of one plasticity-free network, advanced together
// (evolve of every (neuron, instance), then one GEMM per connection on the matrix cores), ten
// steps per hipGraph.  extern "C" bn_batch_* API of include/buryn_amd.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "../bn_batch.h"
#include "api_util.h"
#include "config.h"
#include "network.h"

using namespace bn;

namespace {

uint32_t round_up_u(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }
uint32_t cdiv_u(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

struct bn_batch {
    Config cfg;
    uint32_t n_inst = 0, B = 0, N = 0, NE = 0, NI = 0;
    uint32_t step = 0;
    uint32_t steps_per_graph = 100;  // graph boundaries cost about as much as a step; long graphs amortise them
    uint32_t gemm_rows = 0;  // rows per wave chunk (GEMM tile height along pre)
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    BatchEvolvePlan ev{};
    BatchGemmPlan gm{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float* part[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t part_cap[4] = {0, 0, 0, 0};  // floats allocated per connection partial buffer
    uint32_t pre[4] = {0, 0, 0, 0}, post[4] = {0, 0, 0, 0}, ld[4] = {0, 0, 0, 0};
    float* W[4] = {nullptr, nullptr, nullptr, nullptr};

    template <typename T>
    T* alloc(size_t n, const char* what) {
        void* p = nullptr;
        BN_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), std::string("hipMalloc ") + what);
        allocs.push_back(p);
        BN_HIP_CHECK(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), stream), "hipMemsetAsync");
        return static_cast<T*>(p);
    }
    template <typename T>
    void upload(T* dst, const std::vector<T>& src) {
        if (src.empty()) return;
        BN_HIP_CHECK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, stream), "H2D");
        BN_HIP_CHECK(hipStreamSynchronize(stream), "H2D sync");  // src is pageable
    }
    void destroy_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~bn_batch() {
        destroy_graph();
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // GEMM plan for `rows` pre rows per wave chunk (0 = heuristic: about 2,048 waves).
    void plan_gemm(uint32_t rows) {
        if (rows == 0) {
            uint64_t work = 0;
            for (int c = 0; c < 4; ++c) work += (uint64_t)cdiv_u(post[c], kBatchGemmTileM) * (B / 64) * pre[c];
            rows = (uint32_t)std::min<uint64_t>(512, std::max<uint64_t>(8, round_up_u((uint32_t)(work / 2048), 8)));
        }
        if (rows % 8 != 0 || rows == 0) throw Error(BN_ERR_INVALID_ARGUMENT, "GEMM rows per wave must be a positive multiple of 8");
        gemm_rows = rows;
        gm = BatchGemmPlan{};
        gm.B = B;
        gm.n_conn = 4;
        gm.step = ev.step;
        uint32_t blocks = 0;
        const BatchTraceSet* ts[4] = {&ev.ts[0], &ev.ts[0], &ev.ts[1], &ev.ts[1]};
        for (int c = 0; c < 4; ++c) {
            BatchGemmConn& C = gm.c[c];
            C.W = W[c];
            C.X = ts[c]->x;
            C.ld = ld[c];
            C.n_pre = pre[c];
            C.n_post = post[c];
            C.kc = rows;
            C.n_chunks = cdiv_u(pre[c], rows);
            C.groups = cdiv_u(C.n_chunks, kBatchGemmWaves);
            C.m_tiles = cdiv_u(post[c], kBatchGemmTileM);
            C.n_pairs = B / 64;
            const size_t need = (size_t)C.groups * post[c] * B;
            if (need > part_cap[c]) throw Error(BN_ERR_INVALID_ARGUMENT, "GEMM rows per wave below the allocated minimum");
            C.part = part[c];
            C.block_begin = blocks;
            blocks += C.groups * C.m_tiles * C.n_pairs;
            C.block_end = blocks;
        }
        gm.total_blocks = blocks;
        // Evolve reads the partials of the post group's connections: EE, IE -> E; EI, II -> I.
        ev.g[0].fin_e = BatchPartials{part[0], gm.c[0].groups};
        ev.g[0].fin_i = BatchPartials{part[2], gm.c[2].groups};
        ev.g[1].fin_e = BatchPartials{part[1], gm.c[1].groups};
        ev.g[1].fin_i = BatchPartials{part[3], gm.c[3].groups};
        for (int g = 0; g < 2; ++g) ev.g[g].has_fin_e = ev.g[g].has_fin_i = 1;
    }

    void enqueue_step() {
        BN_HIP_CHECK((hipError_t)launch_batch_evolve(ev, stream), "batched evolve launch");
        BN_HIP_CHECK((hipError_t)launch_batch_gemm(gm, stream), "batched GEMM launch");
    }

    void capture() {
        destroy_graph();
        BN_HIP_CHECK(hipStreamSynchronize(stream), "pre-capture sync");
        BN_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        for (uint32_t k = 0; k < steps_per_graph; ++k) enqueue_step();
        BN_HIP_CHECK(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
        BN_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }

    void run(uint32_t n) {
        uint32_t k = 0;
        for (; exec && k + steps_per_graph <= n; k += steps_per_graph)
            BN_HIP_CHECK(hipGraphLaunch(exec, stream), "hipGraphLaunch");
        for (; k < n; ++k) enqueue_step();
        step += n;
    }
};

namespace {

void build_batch(bn_batch* bt, const bn_batch_instance* inst) {
    const Config& c = bt->cfg;
    if (c.e_e_plasticity_enabled || c.i_e_plasticity_enabled)
        throw Error(BN_ERR_INVALID_ARGUMENT, "the batched shared-weight form needs frozen weights (plasticity off)");
    bt->NE = c.num_excitatory_neurons;
    bt->NI = c.num_inhibitory_neurons;
    bt->N = bt->NE + bt->NI;
    bt->B = round_up_u(bt->n_inst, kBatchAlign);
    const uint32_t N = bt->N, B = bt->B;
    BN_HIP_CHECK(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking), "hipStreamCreate");

    // Shared weights from the config's generator (rng_seed), padded rows of 64 floats.
    Network net = generate_inputs(c, false);
    const uint32_t pres[4] = {bt->NE, bt->NE, bt->NI, bt->NI}, posts[4] = {bt->NE, bt->NI, bt->NE, bt->NI};
    for (int k = 0; k < 4; ++k) {
        bt->pre[k] = pres[k];
        bt->post[k] = posts[k];
        bt->ld[k] = round_up_u(posts[k], 64);
        std::vector<float> w((size_t)pres[k] * bt->ld[k], 0.0f);
        for (uint32_t i = 0; i < pres[k]; ++i)
            std::memcpy(&w[(size_t)i * bt->ld[k]], &net.w[k][(size_t)i * posts[k]], posts[k] * sizeof(float));
        bt->W[k] = bt->alloc<float>(w.size(), "weights");
        bt->upload(bt->W[k], w);
    }

    // Neuron state [N][B].
    BatchEvolvePlan& ev = bt->ev;
    const size_t NB = (size_t)N * B;
    ev.N = N;
    ev.B = B;
    ev.V = bt->alloc<float>(NB, "V");
    ev.ge = bt->alloc<float>(NB, "g_e");
    ev.gi = bt->alloc<float>(NB, "g_i");
    ev.istim = bt->alloc<float>(NB, "i_stimulus");
    ev.refr = bt->alloc<uint32_t>(NB, "refractory");
    ev.spikes = bt->alloc<uint32_t>(NB, "spikes");
    ev.bre = bt->alloc<float>(NB, "bg r e");
    ev.bde = bt->alloc<float>(NB, "bg d e");
    ev.bri = bt->alloc<float>(NB, "bg r i");
    ev.bdi = bt->alloc<float>(NB, "bg d i");
    ev.step = bt->alloc<uint32_t>(1, "step");
    std::vector<float> v0(NB);
    for (int g = 0; g < 2; ++g) {
        const NeuronGroupParams p = group_params(c, g == 1);
        for (uint32_t n = p.offset; n < p.offset + p.num_neurons; ++n)
            for (uint32_t b = 0; b < B; ++b) v0[(size_t)n * B + b] = p.v_rest;  // neuron_group.cpp:21
        NeuronGroupView& v = ev.g[g].v;
        std::memset(&v, 0, sizeof v);
        v.timestep = p.timestep;
        v.g_leak = p.g_leak;
        v.membrane_capacitance = p.membrane_capacitance;
        v.depolarization_slope_factor = p.depolarization_slope_factor;
        v.v_effective_spike_threshold = p.v_effective_spike_threshold;
        v.v_excitatory_reversal = p.v_excitatory_reversal;
        v.v_inhibitory_reversal = p.v_inhibitory_reversal;
        v.v_threshold = p.v_threshold;
        v.v_reset = p.v_reset;
        v.v_rest = p.v_rest;
        v.refractory_steps = p.refractory_steps;
        v.is_inhibitory = p.is_inhibitory;
        v.mean_background_excitatory_spikes = p.mean_background_excitatory_spikes;
        v.mean_background_inhibitory_spikes = p.mean_background_inhibitory_spikes;
        v.background_excitatory_postsynaptic_r_trace_decay = p.background_excitatory_postsynaptic_r_trace_decay;
        v.background_excitatory_postsynaptic_d_trace_decay = p.background_excitatory_postsynaptic_d_trace_decay;
        v.background_inhibitory_postsynaptic_r_trace_decay = p.background_inhibitory_postsynaptic_r_trace_decay;
        v.background_inhibitory_postsynaptic_d_trace_decay = p.background_inhibitory_postsynaptic_d_trace_decay;
        v.background_excitatory_scaling_factor = p.background_excitatory_scaling_factor;
        v.background_inhibitory_scaling_factor = p.background_inhibitory_scaling_factor;
        v.num_neurons = p.num_neurons;
        v.offset = p.offset;
        for (int kind = 0; kind < 2; ++kind) {
            std::vector<uint32_t> t(256);
            uint32_t len = 0;
            BN_CALL(bn_poisson_table(kind == 0 ? p.mean_background_excitatory_spikes : p.mean_background_inhibitory_spikes,
                                     t.data(), 256, &len));
            t.resize(len);
            uint32_t* d = len ? bt->alloc<uint32_t>(len, "poisson table") : nullptr;
            if (len) bt->upload(d, t);
            (kind == 0 ? v.poisson_table_excitatory : v.poisson_table_inhibitory) = d;
            (kind == 0 ? v.poisson_table_excitatory_len : v.poisson_table_inhibitory_len) = len;
        }
    }
    ev.n_groups = 2;
    bt->upload(ev.V, v0);

    // Connection traces: one set per pre group (EE/EI share tau_e, IE/II tau_i; main_buryn.cpp:210-272).
    for (int s = 0; s < 2; ++s) {
        BatchTraceSet& T = ev.ts[s];
        const bool exc = s == 0;
        T.pre_offset = exc ? 0 : bt->NE;
        T.n_pre = exc ? bt->NE : bt->NI;
        T.decay_r = std::exp(-1.0f * c.timestep / (exc ? c.tau_rise_e : c.tau_rise_i));   // connection.cpp:46-47
        T.decay_d = std::exp(-1.0f * c.timestep / (exc ? c.tau_decay_e : c.tau_decay_i));
        T.K = c.postsynaptic_kernel_normalization_value;
        T.r = bt->alloc<float>((size_t)T.n_pre * B, "r traces");
        T.d = bt->alloc<float>((size_t)T.n_pre * B, "d traces");
        T.x = bt->alloc<float>((size_t)T.n_pre * B, "x traces");
    }
    ev.n_ts = 2;

    // Instances: noise keys and stimulus pulses (padding instances: key 0, no stimulus).
    std::vector<uint32_t> keys(2 * (size_t)B, 0u), on(B, 0xFFFFFFFFu), off(B, 0xFFFFFFFFu);
    std::vector<float> amp(NB, 0.0f);
    bool any_stim = false;
    const uint32_t T = c.simulation_steps;
    for (uint32_t b = 0; b < bt->n_inst; ++b) {
        const bn_batch_instance& in = inst[b];
        keys[2 * b] = (uint32_t)in.noise_seed;
        keys[2 * b + 1] = (uint32_t)(in.noise_seed >> 32);
        if (in.num_stimulus_ids == 0) continue;
        if (!in.stimulus_ids) throw Error(BN_ERR_INVALID_ARGUMENT, "instance " + std::to_string(b) + ": stimulus ids are null");
        if (in.stimulus_start > in.stimulus_end)
            throw Error(BN_ERR_INVALID_ARGUMENT, "instance " + std::to_string(b) + ": stimulus start after end");
        // The plain schedule drops events at or after simulation_steps (generate_stimulus_schedules.cpp).
        on[b] = in.stimulus_start < T ? in.stimulus_start : 0xFFFFFFFFu;
        off[b] = in.stimulus_end + 1 < T ? in.stimulus_end + 1 : 0xFFFFFFFFu;
        for (uint32_t q = 0; q < in.num_stimulus_ids; ++q) {
            const uint32_t id = in.stimulus_ids[q];
            if (id >= bt->NE) throw Error(BN_ERR_OUT_OF_RANGE, "instance " + std::to_string(b) + ": stimulus id out of range");
            if (amp[(size_t)id * B + b] != 0.0f)
                throw Error(BN_ERR_INVALID_ARGUMENT, "instance " + std::to_string(b) + ": duplicate stimulus id");
            amp[(size_t)id * B + b] = in.stimulus_amplitude;
            any_stim = true;
        }
    }
    uint32_t* d_keys = bt->alloc<uint32_t>(keys.size(), "keys");
    bt->upload(d_keys, keys);
    ev.keys = d_keys;
    if (any_stim) {
        float* d_amp = bt->alloc<float>(NB, "stimulus amplitudes");
        uint32_t* d_on = bt->alloc<uint32_t>(B, "stimulus on");
        uint32_t* d_off = bt->alloc<uint32_t>(B, "stimulus off");
        bt->upload(d_amp, amp);
        bt->upload(d_on, on);
        bt->upload(d_off, off);
        ev.stim_amp = d_amp;
        ev.stim_on = d_on;
        ev.stim_off = d_off;
    }

    // GEMM partial buffers sized for the smallest chunk height (8 rows).
    for (int k = 0; k < 4; ++k) {
        const uint32_t groups = cdiv_u(cdiv_u(bt->pre[k], 8), kBatchGemmWaves);
        bt->part_cap[k] = (size_t)groups * bt->post[k] * B;
        bt->part[k] = bt->alloc<float>(bt->part_cap[k], "GEMM partials");
    }
    bt->plan_gemm(0);
    BN_HIP_CHECK(hipStreamSynchronize(bt->stream), "batch setup sync");
    bt->capture();
}

}  // namespace

extern "C" {

int bn_batch_create(const char* config_path, uint32_t num_instances, const bn_batch_instance* instances,
                    bn_batch** out) {
    if (!out) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_create: out is null");
    *out = nullptr;
    if (num_instances == 0 || !instances) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_create: no instances");
    auto bt = std::make_unique<bn_batch>();
    const int rc = guarded([&] {
        bt->cfg = load_config_json(config_path ? config_path : "");
        bt->n_inst = num_instances;
        build_batch(bt.get(), instances);
    });
    if (rc) return rc;
    *out = bt.release();
    return BN_OK;
}

void bn_batch_destroy(bn_batch* bt) { delete bt; }

int bn_batch_workload_instance(uint32_t instance, uint32_t num_excitatory, uint32_t num_ids, float max_amplitude,
                               uint32_t* ids, float* amplitude) {
    if (!ids || !amplitude) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_workload_instance: null output");
    if (num_ids > num_excitatory) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_workload_instance: more ids than neurons");
    // Recorded per-instance draw: mt19937 seeded with the instance index; num_ids distinct ids
    // by a partial Fisher-Yates shuffle of [0, num_excitatory), then amplitude = max * U(0,1).
    std::mt19937 gen(instance);
    std::vector<uint32_t> pool(num_excitatory);
    for (uint32_t k = 0; k < num_excitatory; ++k) pool[k] = k;
    for (uint32_t k = 0; k < num_ids; ++k) {
        std::uniform_int_distribution<uint32_t> pick(k, num_excitatory - 1);
        std::swap(pool[k], pool[pick(gen)]);
        ids[k] = pool[k];
    }
    std::sort(ids, ids + num_ids);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    *amplitude = max_amplitude * u(gen);
    return BN_OK;
}

int bn_batch_run(bn_batch* bt, uint32_t num_steps) {
    if (!bt) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_run: null batch");
    return guarded([&] { bt->run(num_steps); });
}

int bn_batch_synchronize(bn_batch* bt) {
    if (!bt) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_synchronize: null batch");
    return guarded([&] { BN_HIP_CHECK(hipStreamSynchronize(bt->stream), "bn_batch_synchronize"); });
}

int bn_batch_run_timed(bn_batch* bt, uint32_t num_steps, float* ms) {
    if (!bt || !ms) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_run_timed: null argument");
    return guarded([&] {
        hipEvent_t a, b;
        BN_HIP_CHECK(hipEventCreate(&a), "hipEventCreate");
        BN_HIP_CHECK(hipEventCreate(&b), "hipEventCreate");
        BN_HIP_CHECK(hipStreamSynchronize(bt->stream), "pre-timing sync");
        BN_HIP_CHECK(hipEventRecord(a, bt->stream), "hipEventRecord");
        bt->run(num_steps);
        BN_HIP_CHECK(hipEventRecord(b, bt->stream), "hipEventRecord");
        BN_HIP_CHECK(hipEventSynchronize(b), "hipEventSynchronize");
        BN_HIP_CHECK(hipEventElapsedTime(ms, a, b), "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    });
}

int bn_batch_time_gemm(bn_batch* bt, uint32_t num_launches, float* ms) {
    if (!bt || !ms) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_time_gemm: null argument");
    return guarded([&] {
        // The GEMM advances the step counter and overwrites the pending partials: both are
        // saved and restored, so the batch's state is unchanged.
        size_t total = 0;
        for (int c = 0; c < 4; ++c) total += bt->part_cap[c];
        float* save = nullptr;
        uint32_t step_save = 0;
        BN_HIP_CHECK(hipMalloc(&save, total * sizeof(float)), "hipMalloc(partials copy)");
        size_t off = 0;
        for (int c = 0; c < 4; ++c) {
            BN_HIP_CHECK(hipMemcpyAsync(save + off, bt->part[c], bt->part_cap[c] * sizeof(float), hipMemcpyDeviceToDevice,
                                        bt->stream), "save partials");
            off += bt->part_cap[c];
        }
        BN_HIP_CHECK(hipMemcpyAsync(&step_save, bt->ev.step, sizeof(uint32_t), hipMemcpyDeviceToHost, bt->stream), "step");
        hipEvent_t a, b;
        BN_HIP_CHECK(hipEventCreate(&a), "hipEventCreate");
        BN_HIP_CHECK(hipEventCreate(&b), "hipEventCreate");
        BN_HIP_CHECK(hipStreamSynchronize(bt->stream), "pre-timing sync");
        BN_HIP_CHECK(hipEventRecord(a, bt->stream), "hipEventRecord");
        for (uint32_t k = 0; k < num_launches; ++k)
            BN_HIP_CHECK((hipError_t)launch_batch_gemm(bt->gm, bt->stream), "batched GEMM launch");
        BN_HIP_CHECK(hipEventRecord(b, bt->stream), "hipEventRecord");
        BN_HIP_CHECK(hipEventSynchronize(b), "hipEventSynchronize");
        BN_HIP_CHECK(hipEventElapsedTime(ms, a, b), "hipEventElapsedTime");
        off = 0;
        for (int c = 0; c < 4; ++c) {
            BN_HIP_CHECK(hipMemcpyAsync(bt->part[c], save + off, bt->part_cap[c] * sizeof(float), hipMemcpyDeviceToDevice,
                                        bt->stream), "restore partials");
            off += bt->part_cap[c];
        }
        BN_HIP_CHECK(hipMemcpyAsync(bt->ev.step, &step_save, sizeof(uint32_t), hipMemcpyHostToDevice, bt->stream), "step");
        BN_HIP_CHECK(hipStreamSynchronize(bt->stream), "timing sync");
        (void)hipFree(save);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    });
}

int bn_batch_set_gemm_rows(bn_batch* bt, uint32_t rows_per_wave) {
    if (!bt) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_set_gemm_rows: null batch");
    return guarded([&] {
        // The partial sums pending for the next evolve follow the current tiling (their
        // number per post neuron), so the tiling may only change before the first step.
        if (bt->step != 0) throw Error(BN_ERR_RUNTIME, "the GEMM tiling can only change before the first step");
        bt->plan_gemm(rows_per_wave);
        bt->capture();
    });
}

int bn_batch_get_info(bn_batch* bt, bn_batch_info* info) {
    if (!bt || !info) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_get_info: null argument");
    std::memset(info, 0, sizeof *info);
    info->num_instances = bt->n_inst;
    info->padded_instances = bt->B;
    info->num_excitatory_neurons = bt->NE;
    info->num_inhibitory_neurons = bt->NI;
    info->total_neurons = bt->N;
    info->clock_step = bt->step;
    info->gemm_rows_per_wave = bt->gemm_rows;
    info->gemm_blocks = bt->gm.total_blocks;
    for (int c = 0; c < 4; ++c) {
        info->weight_pre[c] = bt->pre[c];
        info->weight_post[c] = bt->post[c];
        info->gemm_groups[c] = bt->gm.c[c].groups;
    }
    info->kernels_per_step = 2;
    return BN_OK;
}

int bn_batch_get_state(bn_batch* bt, uint32_t instance, const bn_neuron_state_host* d) {
    if (!bt || !d) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_batch_get_state: null argument");
    if (instance >= bt->n_inst) return set_error(BN_ERR_OUT_OF_RANGE, "bn_batch_get_state: instance out of range");
    return guarded([&] {
        const uint32_t N = bt->N, B = bt->B;
        const size_t NB = (size_t)N * B;
        auto column = [&](const void* src, size_t elem, void* dst) {
            if (!dst) return;
            std::vector<unsigned char> all(NB * elem);
            BN_HIP_CHECK(hipMemcpyAsync(all.data(), src, NB * elem, hipMemcpyDeviceToHost, bt->stream), "D2H state");
            BN_HIP_CHECK(hipStreamSynchronize(bt->stream), "D2H sync");
            for (uint32_t n = 0; n < N; ++n)
                std::memcpy((unsigned char*)dst + (size_t)n * elem, &all[((size_t)n * B + instance) * elem], elem);
        };
        const BatchEvolvePlan& ev = bt->ev;
        column(ev.V, 4, d->membrane_potentials);
        column(ev.istim, 4, d->i_stimulus);
        column(ev.refr, 4, d->refractory_counter);
        column(ev.spikes, 4, d->spikes);
        column(ev.ge, 4, d->background_g_excitatory_synaptic_conductances);
        column(ev.gi, 4, d->background_g_inhibitory_synaptic_conductances);
        column(ev.bre, 4, d->background_excitatory_postsynaptic_r_traces);
        column(ev.bde, 4, d->background_excitatory_postsynaptic_d_traces);
        column(ev.bri, 4, d->background_inhibitory_postsynaptic_r_traces);
        column(ev.bdi, 4, d->background_inhibitory_postsynaptic_d_traces);
        // g after the last propagate: the background conductance plus the pending partials,
        // added in group order in FP32 exactly as the next evolve does.
        for (int which = 0; which < 2; ++which) {
            float* dst = which == 0 ? d->g_excitatory : d->g_inhibitory;
            if (!dst) continue;
            column(which == 0 ? ev.ge : ev.gi, 4, dst);
            for (int g = 0; g < 2; ++g) {
                const BatchGroup& G = ev.g[g];
                const BatchPartials& F = which == 0 ? G.fin_e : G.fin_i;
                const uint32_t n_post = G.v.num_neurons;
                std::vector<float> part((size_t)F.groups * n_post * B);
                BN_HIP_CHECK(hipMemcpyAsync(part.data(), F.part, part.size() * sizeof(float), hipMemcpyDeviceToHost,
                                            bt->stream), "D2H partials");
                BN_HIP_CHECK(hipStreamSynchronize(bt->stream), "D2H sync");
                for (uint32_t j = 0; j < n_post; ++j) {
                    float gv = dst[G.v.offset + j];
                    for (uint32_t q = 0; q < F.groups; ++q) gv += part[((size_t)q * n_post + j) * B + instance];
                    dst[G.v.offset + j] = gv;
                }
            }
        }
    });
}

}  // extern "C"
