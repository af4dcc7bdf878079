// extern "C" whole-simulation API (include/buryn_amd.h) and the network wiring of the
// reference driver (cpp_src/main_buryn.cpp:31-343).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>

#include "../bn_errors.h"
#include "config.h"
#include "engine.h"
#include "generation.h"
#include "api_util.h"
#include "network.h"
#include "sim_internal.h"

namespace bn {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

namespace {


// main_buryn.cpp:31-74
void clean_up(const std::string& output_directory, const std::string& description, const std::string& config_path) {
    std::filesystem::path dir = output_directory;
    if (std::filesystem::exists(dir) && std::filesystem::is_directory(dir)) {
        for (const auto& entry : std::filesystem::directory_iterator(dir)) std::filesystem::remove_all(entry.path());
    } else {
        std::filesystem::create_directories(dir);
    }
    std::ofstream d(dir / "description.txt");
    if (!d) throw std::runtime_error("Failed to write description file in output directory");
    d << description;
    for (const char* sub : {"stimulus_schedules", "recordings", "locations", "videos", "weights", "plots"})
        std::filesystem::create_directory(dir / sub);
    std::filesystem::path cfg(config_path);
    if (!std::filesystem::exists(cfg)) throw std::runtime_error("Config file does not exist: " + config_path);
    std::filesystem::copy_file(cfg, dir / cfg.filename(), std::filesystem::copy_options::overwrite_existing);
}

// main_buryn.cpp:76-96
void validate_plasticity_params(const PlasticityParams& p, bool requires_triplet, const std::string& name) {
    if (p.pre_trace_decay < 0.0f || p.pre_trace_decay > 1.0f)
        throw std::runtime_error(name + ": pre_trace_decay must be in [0, 1]");
    if (p.post_trace_decay < 0.0f || p.post_trace_decay > 1.0f)
        throw std::runtime_error(name + ": post_trace_decay must be in [0, 1]");
    if (requires_triplet) {
        if (p.triplet_pre_trace_decay < 0.0f || p.triplet_pre_trace_decay > 1.0f)
            throw std::runtime_error(name + ": triplet_pre_trace_decay must be in [0, 1]");
        if (p.triplet_post_trace_decay < 0.0f || p.triplet_post_trace_decay > 1.0f)
            throw std::runtime_error(name + ": triplet_post_trace_decay must be in [0, 1]");
    }
    if (p.weight_min > p.weight_max) throw std::runtime_error(name + ": weight_min must be <= weight_max");
}

}  // namespace
}  // namespace bn

using namespace bn;


namespace {

// Steps per captured graph of the fused step: a graph launch costs about a microsecond at its
// boundary, so long graphs amortise it (3,125 neurons: 55.5k steps/s at 10, 60.0k at 50, 60.6k at
// 100); 100 divides the 10,000-step recording interval.
constexpr uint32_t kDefaultStepsPerGraph = 100;

void build(bn_sim* sim, const std::string& config_path) {
    const Config& c = sim->cfg;
    const uint32_t flags = sim->flags;
    const bool quiet = flags & BN_SIM_QUIET;
    const bool files = flags & BN_SIM_WRITE_FILES;
    const bool monitors = !(flags & BN_SIM_NO_MONITORS);
    const bool component = flags & BN_SIM_COMPONENT_PATH;
    if (component && (flags & BN_SIM_LAZY_NORMALIZE))
        throw Error(BN_ERR_INVALID_ARGUMENT, "lazy normalisation needs the fused step (not the component path)");
    if (files) clean_up(c.output_directory, c.description, config_path);
    const uint32_t n_plast = (uint32_t)c.e_e_plasticity_enabled + (uint32_t)c.i_e_plasticity_enabled;
    StateOptions opt;
    opt.reference_histories = component && monitors;
    opt.fused_rings = !component && monitors;
    if (opt.fused_rings) {
        const auto& ids = c.voltage_recording_neuron_ids;
        if (ids.empty())
            for (uint32_t k = 0; k < c.num_excitatory_neurons; ++k) opt.vrec_global_ids.push_back(k);
        for (uint32_t id : ids)
            if (id < c.num_excitatory_neurons) opt.vrec_global_ids.push_back(id);
    }
    opt.lazy_normalize = (flags & BN_SIM_LAZY_NORMALIZE) && c.e_e_plasticity_enabled;
    opt.overlap_steps = !(flags & BN_SIM_SEQUENTIAL_GRAPH);
    opt.reference_layout = flags & BN_SIM_REFERENCE_LAYOUT;
    opt.rng_seed = c.background_noise_seed >= 0 ? (int64_t)c.background_noise_seed : (int64_t)c.rng_seed;
    for (int k = 0; k < 4; ++k) opt.gemv_row_block[k] = c.gemv_row_blocks[k];
    opt.gemv_nontemporal = c.gemv_nontemporal_loads;
    const uint32_t num_stimulus_events =
        (uint32_t)(c.stimulus_selected_input_ids.size() * c.stimulus_start_clock_steps.size() * 2);  // main_buryn.cpp:112
    sim->state.reset(new GpuSimulationState(c.simulation_steps, c.num_excitatory_neurons, c.num_inhibitory_neurons,
                                            num_stimulus_events, n_plast, opt));
    GpuSimulationState& S = *sim->state;
    if (!quiet) std::cout << S.memory_report() << "\n" << "Generating neuron locations and weights...\n";
    Network net = generate_inputs(c, files);
    StimulusGenerationParams sp;
    sp.max_timestep = c.simulation_steps;
    sp.selected_input_ids = c.stimulus_selected_input_ids;
    sp.selected_input_post_group_offsets = c.stimulus_selected_input_post_group_offsets;
    sp.start_clock_steps = c.stimulus_start_clock_steps;
    sp.end_clock_steps = c.stimulus_end_clock_steps;
    sp.stimulus_input_values = c.stimulus_input_values;
    StimulusGenerationResult st = partial_group_multiple_square_waves(files ? c.stimulus_schedules_folder : "", sp);

    sim->system.reset(new System(S));
    System& sys = *sim->system;
    sys.quiet = quiet;
    auto ge = std::make_unique<NeuronGroup>(net.loc_e, group_params(c, false), S);
    auto gi = std::make_unique<NeuronGroup>(net.loc_i, group_params(c, true), S);
    NeuronGroup& E = *ge;
    NeuronGroup& I = *gi;
    sys.add_neuron_group(std::move(ge));
    sys.add_neuron_group(std::move(gi));
    // main_buryn.cpp:196-272 — EE, EI, IE, II
    auto conn = [&](NeuronGroup& pre, NeuronGroup& post, uint32_t idx, bool exc) {
        ConnectionParams p;
        p.timestep = c.timestep;
        p.decay_time = exc ? c.tau_decay_e : c.tau_decay_i;
        p.rise_time = exc ? c.tau_rise_e : c.tau_rise_i;
        p.postsynaptic_kernel_normalization_value = c.postsynaptic_kernel_normalization_value;
        p.weight_matrix_idx = idx;
        auto cn = std::make_unique<Connection>(pre, post, net.w[idx], p, S);
        Connection& ref = *cn;
        sys.add_connection(std::move(cn));
        return &ref;
    };
    Connection* ee = conn(E, E, 0, true);
    conn(E, I, 1, true);
    Connection* ie = conn(I, E, 2, false);
    conn(I, I, 3, false);
    sys.add_stimulus(std::make_unique<Stimulus>(st.timestep_event_counts, st.neuron_ids, st.post_group_offsets,
                                                st.amplitude_changes, S));
    // main_buryn.cpp:286-321
    uint32_t off = 0;
    const float dt = c.timestep;
    if (c.e_e_plasticity_enabled) {
        PlasticityParams p{};
        p.ltp_A = c.e_e_plasticity_ltp_A;
        p.ltd_A = c.e_e_plasticity_ltd_A;
        p.ltp_B = c.e_e_plasticity_ltp_B;
        p.ltd_B = c.e_e_plasticity_ltd_B;
        p.weight_min = c.e_e_plasticity_weight_min;
        p.weight_max = c.e_e_plasticity_weight_max;
        p.pre_trace_decay = std::exp(-1.0f * dt / c.e_e_plasticity_tau_ltp);
        p.post_trace_decay = std::exp(-1.0f * dt / c.e_e_plasticity_tau_ltd);
        p.triplet_pre_trace_decay = std::exp(-1.0f * dt / c.e_e_plasticity_triplet_tau_ltp);
        p.triplet_post_trace_decay = std::exp(-1.0f * dt / c.e_e_plasticity_triplet_tau_ltd);
        p.plasticity_offset = off;
        if (c.e_e_plasticity_rule == "pair") {
            // The reference's pair-STDP component on E->E (its PlasticityParams use ltp_A, ltd_A,
            // the two trace decays and the weight bounds only).
            p.ltp_B = p.ltd_B = 0.0f;
            p.triplet_pre_trace_decay = p.triplet_post_trace_decay = 0.0f;
            sys.add_plasticity(std::make_unique<StdpPlasticity>(*ee, p, S));
            validate_plasticity_params(p, false, "e_e excitatory pair Stdp plasticity");
        } else {
            sys.add_plasticity(std::make_unique<TripletPlasticity>(*ee, p, S));
            validate_plasticity_params(p, true, "e_e excitatory Triplet plasticity");
        }
        ++off;
    }
    if (c.i_e_plasticity_enabled) {
        PlasticityParams p{};
        p.ltp_A = c.i_e_plasticity_ltp_A;
        p.ltd_A = c.i_e_plasticity_ltd_A;
        p.ltd_B = c.i_e_plasticity_ltd_B;
        p.weight_min = c.i_e_plasticity_weight_min;
        p.weight_max = c.i_e_plasticity_weight_max;
        p.pre_trace_decay = std::exp(-1.0f * dt / c.i_e_plasticity_tau_ltp);
        p.post_trace_decay = std::exp(-1.0f * dt / c.i_e_plasticity_tau_ltd);
        p.plasticity_offset = off;
        sys.add_plasticity(std::make_unique<IstdpPlasticity>(*ie, p, S));
        validate_plasticity_params(p, false, "i_e inhibitory Istdp plasticity");
        ++off;
    }
    if (monitors) {
        const uint32_t T = S.options().recording_interval_steps;
        sys.add_monitor(std::make_unique<SpikeRecorder>(E, files ? c.e_spike_recording_file : "", T, S));
        sys.add_monitor(std::make_unique<SpikeRecorder>(I, files ? c.i_spike_recording_file : "", T, S));
        sys.add_monitor(std::make_unique<VoltageRecorder>(E, files ? c.voltage_recording_file : "",
                                                          c.voltage_recording_neuron_ids, T, S));
        sys.add_monitor(std::make_unique<WeightRecorder>(*ee, files ? c.e_e_weight_recording_file : "", S));
        for (auto& m : sys.monitors()) m->quiet = quiet;
    }
    if (flags & BN_SIM_GEMV_PART_SUMS) {
        if (component) throw Error(BN_ERR_INVALID_ARGUMENT, "part sums are a fused-step option (not the component path)");
        sys.set_gemv_part_sums(true);
    }
    if (!component) sys.enable_fused(!(flags & BN_SIM_NO_GRAPH), kDefaultStepsPerGraph);
    if (flags & BN_SIM_PERSISTENT) {
        if (component) throw Error(BN_ERR_INVALID_ARGUMENT, "the persistent step is a fused step (not the component path)");
        sys.enable_persistent({0, 0, 0, 0});
    }
    if (flags & BN_SIM_RESIDENT) {
        if (component) throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step is a fused step (not the component path)");
        sys.enable_resident();
    }
    BN_HIP_CHECK(hipMalloc(&sim->d_tmp_ge, S.total_neurons() * sizeof(float)), "hipMalloc g temp");
    BN_HIP_CHECK(hipMalloc(&sim->d_tmp_gi, S.total_neurons() * sizeof(float)), "hipMalloc g temp");
    BN_HIP_CHECK(hipStreamSynchronize(S.stream()), "build sync");
}

int create(Config cfg, uint32_t flags, const std::string& config_path, bn_sim** out) {
    if (!out) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_create: out is null");
    *out = nullptr;
    std::unique_ptr<bn_sim> sim(new bn_sim());
    sim->cfg = std::move(cfg);
    sim->flags = flags;
    const int rc = guarded([&] { build(sim.get(), config_path); });
    if (rc) return rc;
    *out = sim.release();
    return BN_OK;
}

// Two HIP events destroyed on every exit path (timing helpers throw through guarded()).
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        BN_HIP_CHECK(hipEventCreate(&a), "hipEventCreate");
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            throw Error(BN_ERR_HIP, "hipEventCreate: failed");
        }
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
};

template <typename T>
void d2h(T* dst, const T* src, size_t n, hipStream_t s) {
    if (dst) BN_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s), "D2H state");
}

}  // namespace

extern "C" {

const char* bn_last_error(void) { return g_last_error.c_str(); }

// Inverse-CDF table of Poisson(mean): entry k = floor(2^32 * P(X <= k)), capped at 2^32-1.
int bn_poisson_table(float mean, uint32_t* table, uint32_t capacity, uint32_t* len) {
    if (!len) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_poisson_table: len is null");
    *len = 0;
    if (!(mean > 0.0f)) return BN_OK;
    if (mean > 64.0f) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_poisson_table: mean above 64 is not supported");
    if (!table) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_poisson_table: table is null");
    const double lam = (double)mean;
    double p = std::exp(-lam);
    double cdf = p;
    for (uint32_t k = 0; k < capacity; ++k) {
        const double scaled = std::ldexp(cdf, 32);
        const uint32_t v = scaled >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)std::floor(scaled);
        table[k] = v;
        *len = k + 1;
        if (v == 0xFFFFFFFFu) return BN_OK;
        p = p * lam / (double)(k + 1);
        cdf = cdf + p;
    }
    return set_error(BN_ERR_INVALID_ARGUMENT, "bn_poisson_table: capacity too small for this mean");
}

int bn_sim_create(const char* config_path, uint32_t flags, bn_sim** out) {
    Config cfg;
    const int rc = guarded([&] { cfg = load_config_json(config_path ? config_path : "config.json"); });
    if (rc) return rc;
    return create(std::move(cfg), flags, config_path ? config_path : "config.json", out);
}

int bn_sim_create_from_json_text(const char* json_text, uint32_t flags, bn_sim** out) {
    if (!json_text) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_create_from_json_text: text is null");
    Config cfg;
    int rc = guarded([&] { cfg = load_config_json_text(json_text); });
    if (rc) return rc;
    std::string path;
    if (flags & BN_SIM_WRITE_FILES) {
        rc = guarded([&] {
            std::filesystem::create_directories(cfg.output_directory);
            path = (std::filesystem::path(cfg.output_directory) / "config.json").string();
            std::ofstream(path + ".in") << json_text;
            path += ".in";
        });
        if (rc) return rc;
    }
    return create(std::move(cfg), flags, path, out);
}

void bn_sim_destroy(bn_sim* sim) { delete sim; }

int bn_sim_run(bn_sim* sim, uint32_t num_steps) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_run: sim is null");
    return guarded([&] { sim->system->run(num_steps); });
}

int bn_sim_step_components(bn_sim* sim, uint32_t num_steps) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_step_components: sim is null");
    return guarded([&] {
        for (uint32_t k = 0; k < num_steps; ++k) sim->system->step_the_clock(sim->system->clock_step());
    });
}

int bn_sim_synchronize(bn_sim* sim) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_synchronize: sim is null");
    return guarded([&] {
        BN_HIP_CHECK(hipStreamSynchronize(sim->state->stream()), "hipStreamSynchronize");
        sim->state->check_status();
    });
}

int bn_sim_get_info(bn_sim* sim, bn_sim_info* info) {
    if (!sim || !info) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_get_info: null argument");
    const GpuSimulationState& S = *sim->state;
    std::memset(info, 0, sizeof *info);
    info->num_excitatory_neurons = S.num_excitatory_neurons();
    info->num_inhibitory_neurons = S.num_inhibitory_neurons();
    info->total_neurons = S.total_neurons();
    info->clock_step = sim->system->clock_step();
    info->simulation_steps = sim->cfg.simulation_steps;
    info->num_plasticities = S.num_plasticities();
    for (uint32_t k = 0; k < 4; ++k) {
        info->weight_matrix_ld[k] = S.weight_ld(k);
        info->gemv_row_block[k] = S.gemv_row_block(k);
        info->weight_pre[k] = S.weight_pre(k);
        info->weight_post[k] = S.weight_post(k);
    }
    info->flags = sim->flags;
    info->kernels_per_step = sim->system->kernels_per_step();
    info->device_bytes = S.total_allocated_bytes();
    info->steps_per_graph = sim->system->graph_steps();
    info->graph_launches = sim->system->graph_launches();
    info->persistent_blocks = sim->system->persistent_blocks();
    info->persistent_items = sim->system->persistent_items();
    info->persistent_resident_rows = sim->system->persistent_resident_rows();
    if (sim->system->resident()) {
        const RPlan& R = sim->system->resident_plan();
        info->resident_blocks = R.blocks;
        info->resident_columns = kRCols;
        info->resident_e_waves = R.e_waves;
        info->resident_rows = R.rows;
    }
    info->gemv_part_sums = sim->system->gemv_part_sums() ? 1u : 0u;
    if (sim->system->persistent()) {
        info->gemv_part_chunks = kPWaves;
        info->kernels_per_step = 1;
        for (uint32_t k = 0; k < 4; ++k) info->gemv_row_block[k] = sim->system->persistent_rows(k);
    }
    return BN_OK;
}

int bn_sim_get_neuron_state(bn_sim* sim, const bn_neuron_state_host* d) {
    if (!sim || !d) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_get_neuron_state: null argument");
    return guarded([&] {
        GpuSimulationState& S = *sim->state;
        hipStream_t s = S.stream();
        const size_t N = S.total_neurons();
        sim->system->materialize_g(sim->d_tmp_ge, sim->d_tmp_gi);
        d2h(d->membrane_potentials, S.d_membrane_potentials(), N, s);
        d2h(d->g_excitatory, sim->d_tmp_ge, N, s);
        d2h(d->g_inhibitory, sim->d_tmp_gi, N, s);
        d2h(d->i_stimulus, S.d_i_stimulus(), N, s);
        d2h(d->refractory_counter, S.d_refractory_counter(), N, s);
        d2h(d->spikes, S.d_spikes(), N, s);
        d2h(d->background_g_excitatory_synaptic_conductances, S.d_background_g_excitatory_synaptic_conductances(), N, s);
        d2h(d->background_g_inhibitory_synaptic_conductances, S.d_background_g_inhibitory_synaptic_conductances(), N, s);
        d2h(d->background_excitatory_postsynaptic_r_traces, S.d_background_excitatory_postsynaptic_r_traces(), N, s);
        d2h(d->background_excitatory_postsynaptic_d_traces, S.d_background_excitatory_postsynaptic_d_traces(), N, s);
        d2h(d->background_inhibitory_postsynaptic_r_traces, S.d_background_inhibitory_postsynaptic_r_traces(), N, s);
        d2h(d->background_inhibitory_postsynaptic_d_traces, S.d_background_inhibitory_postsynaptic_d_traces(), N, s);
        BN_HIP_CHECK(hipStreamSynchronize(s), "state download sync");
        S.check_status();
    });
}

int bn_sim_get_weights(bn_sim* sim, uint32_t idx, float* dst) {
    if (!sim || !dst) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_get_weights: null argument");
    return guarded([&] {
        const std::vector<float> w = sim->state->download_weight_matrix(idx);
        std::memcpy(dst, w.data(), w.size() * sizeof(float));
    });
}

int bn_sim_get_connection_traces(bn_sim* sim, uint32_t idx, float* r, float* d, float* x) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_get_connection_traces: sim is null");
    return guarded([&] {
        GpuSimulationState& S = *sim->state;
        if (idx >= 4) throw Error(BN_ERR_OUT_OF_RANGE, "connection index out of range");
        const uint32_t n = S.weight_pre(idx);
        const TraceSet* ts = sim->system->trace_set_for_connection(idx);
        hipStream_t s = S.stream();
        d2h(r, ts ? ts->r : S.d_postsynaptic_r_traces(idx), n, s);
        d2h(d, ts ? ts->d : S.d_postsynaptic_d_traces(idx), n, s);
        d2h(x, ts ? ts->x : S.d_summed_postsynaptic_conductances(idx), n, s);
        BN_HIP_CHECK(hipStreamSynchronize(s), "trace download sync");
    });
}

int bn_sim_get_plasticity_traces(bn_sim* sim, uint32_t off, float* pre, float* post, float* tpre, float* tpost) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_get_plasticity_traces: sim is null");
    return guarded([&] {
        GpuSimulationState& S = *sim->state;
        if (off >= S.num_plasticities()) throw Error(BN_ERR_OUT_OF_RANGE, "plasticity offset out of range");
        const size_t N = S.total_neurons(), base = (size_t)off * N;
        hipStream_t s = S.stream();
        d2h(pre, S.d_pre_traces() + base, N, s);
        d2h(post, S.d_post_traces() + base, N, s);
        d2h(tpre, S.d_triplet_pre_traces() + base, N, s);
        d2h(tpost, S.d_triplet_post_traces() + base, N, s);
        std::vector<uint32_t> spikes(N);
        d2h(spikes.data(), S.d_spikes(), N, s);
        BN_HIP_CHECK(hipStreamSynchronize(s), "trace download sync");
        if (sim->system->fused()) {
            // The fused step keeps the +1 of a spike for the next step's decay; present the
            // traces as the reference holds them after the step (trace + 1 where spiked).
            for (const auto& p : sim->system->plasticities()) {
                const PlasticityView& v = p->gpu_view();
                if (v.plasticity_offset != off) continue;
                auto add = [&](float* a, uint32_t lo, uint32_t n) {
                    if (!a) return;
                    for (uint32_t g = lo; g < lo + n; ++g)
                        if (spikes[g] == 1) a[g] = a[g] + 1.0f;
                };
                add(pre, v.pre_group_offset, v.pre_group_num_neurons);
                add(post, v.post_group_offset, v.post_group_num_neurons);
                if (p->kind() == kPlastTriplet) {
                    add(tpre, v.pre_group_offset, v.pre_group_num_neurons);
                    add(tpost, v.post_group_offset, v.post_group_num_neurons);
                }
            }
        }
    });
}

int bn_sim_get_spike_raster(bn_sim* sim, uint32_t first_step, uint32_t num_steps, uint32_t* dst) {
    if (!sim || (!dst && num_steps)) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_get_spike_raster: null argument");
    return guarded([&] {
        if (!sim->system->fused()) throw Error(BN_ERR_RUNTIME, "spike raster: the component path keeps [neuron][slot] histories");
        BN_HIP_CHECK(hipStreamSynchronize(sim->state->stream()), "raster sync");
        const std::vector<uint32_t> rows =
            sim->state->download_spike_ring_rows(first_step, num_steps, sim->system->clock_step());
        std::memcpy(dst, rows.data(), rows.size() * sizeof(uint32_t));
    });
}

int bn_sim_run_timed(bn_sim* sim, uint32_t num_steps, float* ms) {
    if (!sim || !ms) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_run_timed: null argument");
    return guarded([&] {
        System& sys = *sim->system;
        hipStream_t s = sim->state->stream();
        EventPair ev;
        BN_HIP_CHECK(hipStreamSynchronize(s), "pre-timing sync");
        BN_HIP_CHECK(hipEventRecord(ev.a, s), "hipEventRecord");
        if (sys.fused()) sys.launch_fused_steps(num_steps);
        else
            for (uint32_t k = 0; k < num_steps; ++k) sys.step_the_clock(sys.clock_step());
        BN_HIP_CHECK(hipEventRecord(ev.b, s), "hipEventRecord");
        BN_HIP_CHECK(hipEventSynchronize(ev.b), "hipEventSynchronize");
        BN_HIP_CHECK(hipEventElapsedTime(ms, ev.a, ev.b), "hipEventElapsedTime");
        sim->state->check_status();
    });
}

int bn_sim_save_state(bn_sim* sim, const char* path) {
    if (!sim || !path) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_save_state: null argument");
    return guarded([&] {
        const std::vector<uint8_t> bytes = sim->state->snapshot();
        std::FILE* f = std::fopen(path, "wb");
        if (!f) throw Error(BN_ERR_RUNTIME, std::string("cannot open ") + path);
        const size_t w = std::fwrite(bytes.data(), 1, bytes.size(), f);
        const int rc = std::fclose(f);
        if (w != bytes.size() || rc != 0) throw Error(BN_ERR_RUNTIME, std::string("cannot write ") + path);
    });
}

int bn_sim_load_state(bn_sim* sim, const char* path) {
    if (!sim || !path) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_load_state: null argument");
    return guarded([&] {
        std::FILE* f = std::fopen(path, "rb");
        if (!f) throw Error(BN_ERR_RUNTIME, std::string("cannot open ") + path);
        std::vector<uint8_t> bytes;
        uint8_t buf[1 << 16];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + n);
        std::fclose(f);
        sim->state->restore(bytes);
        bn_device_clock clk{};
        BN_HIP_CHECK(hipMemcpy(&clk, sim->state->d_clock(), sizeof clk, hipMemcpyDeviceToHost), "clock download");
        sim->system->set_clock_step(clk.clock_step);
    });
}

namespace {
int measure_read_bandwidth(uint64_t bytes, uint32_t repeats, bool in_one_launch, float* gb_per_s) {
    if (!gb_per_s || bytes < 16 || repeats == 0)
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_measure_read_bandwidth: bad argument");
    return guarded([&] {
        bytes &= ~(uint64_t)15;
        // One launch reads at most 4 GiB (well under a second at HBM rate): a dispatch running for
        // seconds looks like a hung queue.
        if (in_one_launch) repeats = (uint32_t)std::min<uint64_t>(repeats, std::max<uint64_t>(1, (4ull << 30) / bytes));
        void* buf = nullptr;
        float* sink = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        const uint32_t blocks = 4096;
        auto cleanup = [&] {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            if (s) (void)hipStreamDestroy(s);
            if (sink) (void)hipFree(sink);
            if (buf) (void)hipFree(buf);
        };
        try {
            BN_HIP_CHECK(hipMalloc(&buf, bytes), "hipMalloc(probe buffer)");
            BN_HIP_CHECK(hipMalloc(&sink, blocks * sizeof(float)), "hipMalloc(probe sink)");
            BN_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
            BN_HIP_CHECK(hipMemsetAsync(buf, 0, bytes, s), "hipMemsetAsync");
            BN_HIP_CHECK(hipEventCreate(&a), "hipEventCreate");
            BN_HIP_CHECK(hipEventCreate(&b), "hipEventCreate");
            BN_HIP_CHECK((hipError_t)launch_read_probe(buf, bytes, sink, blocks, s), "probe launch");
            BN_HIP_CHECK(hipEventRecord(a, s), "hipEventRecord");
            if (in_one_launch) {
                BN_HIP_CHECK((hipError_t)launch_read_probe_passes(buf, bytes, repeats, sink, blocks, s), "probe launch");
            } else {
                for (uint32_t k = 0; k < repeats; ++k)
                    BN_HIP_CHECK((hipError_t)launch_read_probe(buf, bytes, sink, blocks, s), "probe launch");
            }
            BN_HIP_CHECK(hipEventRecord(b, s), "hipEventRecord");
            BN_HIP_CHECK(hipEventSynchronize(b), "hipEventSynchronize");
            float ms = 0.0f;
            BN_HIP_CHECK(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
            *gb_per_s = (float)((double)bytes * repeats / (ms * 1e-3) / 1e9);
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}
}  // namespace

int bn_measure_read_bandwidth(uint64_t bytes, uint32_t repeats, float* gb_per_s) {
    return measure_read_bandwidth(bytes, repeats, false, gb_per_s);
}

int bn_measure_read_bandwidth_one_launch(uint64_t bytes, uint32_t passes, float* gb_per_s) {
    return measure_read_bandwidth(bytes, passes, true, gb_per_s);
}

int bn_sim_set_persistent_rows(bn_sim* sim, const uint32_t rows[4]) {
    if (!sim || !rows) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_set_persistent_rows: null argument");
    return guarded([&] {
        if (!sim->system->persistent()) throw Error(BN_ERR_RUNTIME, "bn_sim_set_persistent_rows: not a persistent simulation");
        sim->system->enable_persistent({rows[0], rows[1], rows[2], rows[3]});
    });
}

int bn_sim_set_persistent_resident(bn_sim* sim, int allow) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_set_persistent_resident: null simulation");
    return guarded([&] { sim->system->set_persistent_resident(allow != 0); });
}

int bn_sim_set_async_recording(bn_sim* sim, int async) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_set_async_recording: null simulation");
    return guarded([&] { sim->system->set_async_recording(async != 0); });
}

int bn_sim_record_times(bn_sim* sim, double* call_seconds, double* writer_seconds, uint32_t* record_steps) {
    if (!sim || !call_seconds || !writer_seconds || !record_steps)
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_record_times: null argument");
    return guarded([&] {
        *call_seconds = sim->system->record_call_seconds();
        *writer_seconds = sim->system->record_writer_seconds();
        *record_steps = sim->system->record_steps();
    });
}

int bn_sim_set_resident_cooperative(bn_sim* sim, int cooperative) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_set_resident_cooperative: null simulation");
    return guarded([&] { sim->system->set_resident_cooperative(cooperative != 0); });
}

int bn_sim_set_persistent_wait_limit(bn_sim* sim, uint32_t polls) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_set_persistent_wait_limit: null simulation");
    return guarded([&] { sim->system->set_persistent_wait_limit(polls); });
}

int bn_sim_set_steps_per_graph(bn_sim* sim, uint32_t steps_per_graph) {
    if (!sim) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_set_steps_per_graph: null simulation");
    return guarded([&] { sim->system->set_steps_per_graph(steps_per_graph); });
}

int bn_sim_time_propagate_kernel(bn_sim* sim, uint32_t num_steps, float* kernel_ms, float* step_ms) {
    if (!sim || !kernel_ms || !step_ms) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_sim_time_propagate_kernel: null argument");
    return guarded([&] {
        System& sys = *sim->system;
        if (!sys.fused()) throw Error(BN_ERR_RUNTIME, "kernel timing needs the fused step");
        if (sys.persistent() || sys.resident())
            throw Error(BN_ERR_RUNTIME, "kernel timing: this step is one launch per run (time bn_sim_run_timed)");
        hipStream_t s = sim->state->stream();
        EventPair steps, kernels;   // destroyed on every exit, errors included
        // (1) num_steps whole steps launched one by one (no graph).
        BN_HIP_CHECK(hipStreamSynchronize(s), "pre-timing sync");
        BN_HIP_CHECK(hipEventRecord(steps.a, s), "hipEventRecord");
        for (uint32_t k = 0; k < num_steps; ++k) sys.launch_fused_step_eager(nullptr, nullptr);
        BN_HIP_CHECK(hipEventRecord(steps.b, s), "hipEventRecord");
        // (2) num_steps back-to-back propagate launches over the current weights.  The pending
        // partial sums belong to the next step: saved before and restored after.
        GemvPlan G = sys.gemv_plan();
        G.pc.list = nullptr;  // the propagate alone
        size_t total = 0;
        for (uint32_t c = 0; c < G.n_conn; ++c) total += (size_t)G.c[c].n_chunks * G.c[c].ld;
        double* save = nullptr;
        if (total) BN_HIP_CHECK(hipMalloc(&save, total * sizeof(double)), "hipMalloc(partials copy)");
        // Freed on every exit after the stream has drained (the copies may still read it).
        const auto release = [s](double* p) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(p);
        };
        std::unique_ptr<double, decltype(release)> save_owner(save, release);
        size_t off = 0;
        for (uint32_t c = 0; c < G.n_conn; ++c) {
            const size_t n = (size_t)G.c[c].n_chunks * G.c[c].ld;
            BN_HIP_CHECK(hipMemcpyAsync(save + off, G.c[c].partials, n * sizeof(double), hipMemcpyDeviceToDevice, s),
                         "save partials");
            off += n;
        }
        BN_HIP_CHECK(hipEventRecord(kernels.a, s), "hipEventRecord");
        for (uint32_t k = 0; k < num_steps; ++k)
            BN_HIP_CHECK((hipError_t)launch_gemv(G, s), "propagate launch");
        BN_HIP_CHECK(hipEventRecord(kernels.b, s), "hipEventRecord");
        off = 0;
        for (uint32_t c = 0; c < G.n_conn; ++c) {
            const size_t n = (size_t)G.c[c].n_chunks * G.c[c].ld;
            BN_HIP_CHECK(hipMemcpyAsync(G.c[c].partials, save + off, n * sizeof(double), hipMemcpyDeviceToDevice, s),
                         "restore partials");
            off += n;
        }
        BN_HIP_CHECK(hipStreamSynchronize(s), "timing sync");
        save_owner.reset();
        float total_ms = 0.0f, kern = 0.0f;
        BN_HIP_CHECK(hipEventElapsedTime(&total_ms, steps.a, steps.b), "hipEventElapsedTime");
        BN_HIP_CHECK(hipEventElapsedTime(&kern, kernels.a, kernels.b), "hipEventElapsedTime");
        *kernel_ms = kern;
        *step_ms = total_ms;
        sim->state->check_status();
    });
}

int bn_config_dump_json(const char* config_path, char* buf, size_t buf_size, size_t* needed) {
    std::string text;
    const int rc = guarded([&] { text = config_to_json(load_config_json(config_path ? config_path : "")); });
    if (rc) return rc;
    if (needed) *needed = text.size() + 1;
    if (buf && buf_size > 0) {
        const size_t n = std::min(buf_size - 1, text.size());
        std::memcpy(buf, text.data(), n);
        buf[n] = '\0';
    }
    return BN_OK;
}

int bn_set_connectivity_cache(const char* directory) {
    return guarded([&] { set_connectivity_cache_dir(directory ? directory : ""); });
}

int bn_generate_weight_matrix(const char* config_path, uint32_t idx, float* dst, uint32_t* pre, uint32_t* post) {
    if (idx >= 4) return set_error(BN_ERR_OUT_OF_RANGE, "Invalid weight matrix index: " + std::to_string(idx));
    return guarded([&] {
        const Config c = load_config_json(config_path ? config_path : "");
        const uint32_t np = (idx < 2) ? c.num_excitatory_neurons : c.num_inhibitory_neurons;
        const uint32_t nq = (idx % 2 == 0) ? c.num_excitatory_neurons : c.num_inhibitory_neurons;
        if (pre) *pre = np;
        if (post) *post = nq;
        if (!dst) return;
        const auto le = generate_2d_grid_locations("", c.num_excitatory_neurons, c.excitatory_location_spacing,
                                                   c.excitatory_side_points, c.excitatory_side_points);
        const auto li = generate_2d_grid_locations("", c.num_inhibitory_neurons, c.inhibitory_location_spacing,
                                                   c.inhibitory_side_points, c.inhibitory_side_points);
        const WeightGenerationParams* gp[4] = {&c.weight_generation_params_e_to_e, &c.weight_generation_params_e_to_i,
                                               &c.weight_generation_params_i_to_e, &c.weight_generation_params_i_to_i};
        const auto& a = idx < 2 ? le : li;
        const auto& b = idx % 2 == 0 ? le : li;
        const std::vector<float> w = distance_2d_weights("", a, b, *gp[idx], c.rng_seed, idx);
        std::memcpy(dst, w.data(), w.size() * sizeof(float));
    });
}

}  // extern "C"
