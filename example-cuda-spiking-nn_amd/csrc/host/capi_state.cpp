// extern "C" device-state object (include/buryn_amd.h, bn_state_*): the engine's
// GpuSimulationState behind the C ABI, for a reference-side caller that keeps its own components
// and builds views as cpp_src/gpu_core/gpu_simulation_state.h:35-81 and the components'
// make_gpu_view functions do (neuron_group.cpp:64-76, connection.cpp:44-87, stimulus.cpp,
// triplet_plasticity.cpp:35-85).
#include <cstring>
#include <memory>
#include <vector>

#include "../bn_errors.h"
#include "api_util.h"
#include "engine.h"

struct bn_state {
    std::unique_ptr<bn::GpuSimulationState> s;
    uint32_t flags = 0;
};

using namespace bn;

namespace {

template <typename T>
std::vector<T> host_or_zeros(const T* src, size_t n) {
    return src ? std::vector<T>(src, src + n) : std::vector<T>(n, T(0));
}

GpuSimulationState& state_of(bn_state* st) {
    if (!st || !st->s) throw Error(BN_ERR_INVALID_ARGUMENT, "bn_state: null state");
    return *st->s;
}

}  // namespace

extern "C" {

int bn_state_create(uint32_t total_simulation_length, uint32_t num_excitatory_neurons, uint32_t num_inhibitory_neurons,
                    uint32_t num_stimulus_events, uint32_t num_plasticities, uint64_t rng_seed, uint32_t flags,
                    bn_state** out) {
    if (!out) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_create: out is null");
    *out = nullptr;
    if (num_excitatory_neurons == 0 || num_inhibitory_neurons == 0)
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_create: both neuron groups must be non-empty");
    return guarded([&] {
        StateOptions opt;
        opt.reference_histories = (flags & BN_STATE_HISTORIES) != 0;
        opt.fused_rings = false;
        opt.reference_layout = (flags & BN_STATE_REFERENCE_LAYOUT) != 0;
        opt.rng_seed = (int64_t)rng_seed;
        std::unique_ptr<bn_state> st(new bn_state());
        st->flags = flags;
        st->s.reset(new GpuSimulationState(total_simulation_length, num_excitatory_neurons, num_inhibitory_neurons,
                                           num_stimulus_events, num_plasticities, opt));
        *out = st.release();
    });
}

void bn_state_destroy(bn_state* state) { delete state; }

int bn_state_get_buffers(bn_state* state, bn_state_buffers* b) {
    if (!b) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_get_buffers: out is null");
    return guarded([&] {
        GpuSimulationState& S = state_of(state);
        std::memset(b, 0, sizeof *b);
        b->total_neurons = S.total_neurons();
        b->num_excitatory_neurons = S.num_excitatory_neurons();
        b->num_inhibitory_neurons = S.num_inhibitory_neurons();
        b->recording_interval_steps = S.options().recording_interval_steps;
        b->membrane_potentials = S.d_membrane_potentials();
        b->g_excitatory = S.d_g_excitatory();
        b->g_inhibitory = S.d_g_inhibitory();
        b->i_stimulus = S.d_i_stimulus();
        b->refractory_counter = S.d_refractory_counter();
        b->spikes = S.d_spikes();
        b->background_g_excitatory_synaptic_conductances = S.d_background_g_excitatory_synaptic_conductances();
        b->background_g_inhibitory_synaptic_conductances = S.d_background_g_inhibitory_synaptic_conductances();
        b->background_excitatory_postsynaptic_r_traces = S.d_background_excitatory_postsynaptic_r_traces();
        b->background_excitatory_postsynaptic_d_traces = S.d_background_excitatory_postsynaptic_d_traces();
        b->background_inhibitory_postsynaptic_r_traces = S.d_background_inhibitory_postsynaptic_r_traces();
        b->background_inhibitory_postsynaptic_d_traces = S.d_background_inhibitory_postsynaptic_d_traces();
        b->clock = S.d_clock();
        for (uint32_t k = 0; k < 4; ++k) {
            b->weight_matrix[k] = S.d_w(k);
            b->weight_matrix_ld[k] = S.weight_ld(k);
            b->weight_pre[k] = S.weight_pre(k);
            b->weight_post[k] = S.weight_post(k);
            b->postsynaptic_r_traces[k] = S.d_postsynaptic_r_traces(k);
            b->postsynaptic_d_traces[k] = S.d_postsynaptic_d_traces(k);
            b->summed_postsynaptic_conductances[k] = S.d_summed_postsynaptic_conductances(k);
            b->gemv_partials[k] = S.d_gemv_partials(k);
            b->gemv_row_block[k] = S.gemv_row_block(k);
        }
        b->timestep_event_count = S.d_timestep_event_count();
        b->stimulus_event_neuron_ids = S.d_stimulus_event_neuron_ids();
        b->stimulus_event_post_group_offsets = S.d_stimulus_event_post_group_offsets();
        b->stimulus_event_amplitude_changes = S.d_stimulus_event_amplitude_changes();
        b->pre_traces = S.d_pre_traces();
        b->post_traces = S.d_post_traces();
        b->triplet_pre_traces = S.d_triplet_pre_traces();
        b->triplet_post_traces = S.d_triplet_post_traces();
        b->pre_weight_sums = S.d_pre_weight_sums();
        b->spike_history = S.d_spike_history();
        b->membrane_potential_history = S.d_membrane_potential_history();
    });
}

int bn_state_get_stream(bn_state* state, bn_stream_t* out) {
    if (!out) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_get_stream: out is null");
    return guarded([&] { *out = state_of(state).stream(); });
}

int bn_state_total_allocated_bytes(bn_state* state, uint64_t* bytes) {
    if (!bytes) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_total_allocated_bytes: out is null");
    return guarded([&] { *bytes = state_of(state).total_allocated_bytes(); });
}

int bn_state_poisson_table(bn_state* state, uint32_t group, uint32_t kind, float mean, const uint32_t** table,
                           uint32_t* len) {
    if (!table || !len) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_poisson_table: null output");
    if (group > 1 || kind > 1) return set_error(BN_ERR_OUT_OF_RANGE, "bn_state_poisson_table: group and kind are 0 or 1");
    return guarded([&] {
        GpuSimulationState& S = state_of(state);
        S.set_poisson_table(group, kind, mean);
        *table = S.d_poisson_table(group, kind);
        *len = S.poisson_table_len(group, kind);
    });
}

int bn_state_upload_neuron_group_view(bn_state* state, const NeuronGroupView* v, const bn_neuron_state_host* h) {
    if (!v) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_upload_neuron_group_view: view is null");
    return guarded([&] {
        GpuSimulationState& S = state_of(state);
        const size_t n = v->num_neurons;
        const bn_neuron_state_host z{};
        const bn_neuron_state_host& a = h ? *h : z;
        NeuronGroupInitHostData d;
        d.membrane_potentials = host_or_zeros<float>(a.membrane_potentials, n);
        d.g_excitatory = host_or_zeros<float>(a.g_excitatory, n);
        d.g_inhibitory = host_or_zeros<float>(a.g_inhibitory, n);
        d.i_stimulus = host_or_zeros<float>(a.i_stimulus, n);
        d.refractory_counter = host_or_zeros<uint32_t>(a.refractory_counter, n);
        d.spikes = host_or_zeros<uint32_t>(a.spikes, n);
        d.background_g_excitatory_synaptic_conductances =
            host_or_zeros<float>(a.background_g_excitatory_synaptic_conductances, n);
        d.background_g_inhibitory_synaptic_conductances =
            host_or_zeros<float>(a.background_g_inhibitory_synaptic_conductances, n);
        d.background_excitatory_postsynaptic_r_traces = host_or_zeros<float>(a.background_excitatory_postsynaptic_r_traces, n);
        d.background_excitatory_postsynaptic_d_traces = host_or_zeros<float>(a.background_excitatory_postsynaptic_d_traces, n);
        d.background_inhibitory_postsynaptic_r_traces = host_or_zeros<float>(a.background_inhibitory_postsynaptic_r_traces, n);
        d.background_inhibitory_postsynaptic_d_traces = host_or_zeros<float>(a.background_inhibitory_postsynaptic_d_traces, n);
        // The view's slices must be the state's own (the reference uploads through the view).
        if (v->membrane_potentials != S.d_membrane_potentials() + v->offset)
            throw Error(BN_ERR_INVALID_ARGUMENT, "upload_neuron_group_view: the view does not point into this state");
        S.upload_neuron_group_view(*v, d);
    });
}

int bn_state_upload_connection_view(bn_state* state, const ConnectionView* v, const float* weights, const float* r,
                                    const float* d, const float* x) {
    if (!v || !weights) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_upload_connection_view: null argument");
    return guarded([&] {
        GpuSimulationState& S = state_of(state);
        const uint32_t k = v->weight_matrix_idx;
        (void)S.lookup_weight_matrix_name(k);
        if (v->pre_group_num_neurons != S.weight_pre(k) || v->post_group_num_neurons != S.weight_post(k))
            throw Error(BN_ERR_INVALID_ARGUMENT, "upload_connection_view: view shape differs from weight matrix " +
                                                     S.lookup_weight_matrix_name(k));
        const size_t np = S.weight_pre(k);
        ConnectionInitHostData h;
        h.weight_matrix.assign(weights, weights + np * S.weight_post(k));
        h.postsynaptic_r_traces = host_or_zeros<float>(r, np);
        h.postsynaptic_d_traces = host_or_zeros<float>(d, np);
        h.summed_postsynaptic_conductances = host_or_zeros<float>(x, np);
        S.upload_connection_view(*v, h);
    });
}

int bn_state_upload_stimulus_view(bn_state* state, const StimulusView* v, const uint32_t* counts, uint32_t num_timesteps,
                                  const uint32_t* ids, const uint32_t* offsets, const float* amps, uint32_t num_events) {
    if (!v) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_upload_stimulus_view: view is null");
    if ((num_timesteps && !counts) || (num_events && (!ids || !offsets || !amps)))
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_upload_stimulus_view: null host array");
    return guarded([&] {
        GpuSimulationState& S = state_of(state);
        StimulusInitHostData h;
        h.timestep_event_count.assign(counts, counts + num_timesteps);
        h.stimulus_event_neuron_ids.assign(ids, ids + num_events);
        h.stimulus_event_post_group_offsets.assign(offsets, offsets + num_events);
        h.stimulus_event_amplitude_changes.assign(amps, amps + num_events);
        for (uint32_t e = 0; e < num_events; ++e)
            if ((uint64_t)ids[e] + offsets[e] >= S.total_neurons())
                throw Error(BN_ERR_OUT_OF_RANGE, "stimulus event neuron id out of range");
        S.upload_stimulus_view(*v, h);
    });
}

int bn_state_upload_plasticity_view(bn_state* state, const PlasticityView* v, const float* pre, const float* post,
                                    const float* tpre, const float* tpost, const float* wsums) {
    if (!v) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_upload_plasticity_view: view is null");
    return guarded([&] {
        (void)state_of(state);
        PlasticityInitHostData h;
        const size_t np = v->pre_group_num_neurons, nq = v->post_group_num_neurons;
        h.pre_traces = host_or_zeros<float>(pre, np);
        h.post_traces = host_or_zeros<float>(post, nq);
        h.triplet_pre_traces = host_or_zeros<float>(tpre, np);
        h.triplet_post_traces = host_or_zeros<float>(tpost, nq);
        // One target per pre row (triplet_plasticity.cpp:22-30 fills it with the initial column sums).
        h.pre_weight_sums = host_or_zeros<float>(wsums, np);
        state_of(state).upload_plasticity_view(*v, h);
    });
}

int bn_state_download_and_flush_spike_history(bn_state* state, uint32_t offset, uint32_t num_neurons,
                                              uint32_t num_timesteps, uint32_t* dst) {
    if (!dst) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_download_and_flush_spike_history: dst is null");
    return guarded([&] {
        GpuSimulationState& S = state_of(state);
        if (!S.d_spike_history()) throw Error(BN_ERR_RUNTIME, "the state has no histories (BN_STATE_HISTORIES)");
        const std::vector<uint32_t> h = S.download_and_flush_spike_history(offset, num_neurons, num_timesteps);
        std::memcpy(dst, h.data(), h.size() * sizeof(uint32_t));
    });
}

int bn_state_download_and_flush_membrane_potential_history(bn_state* state, uint32_t offset, uint32_t num_neurons,
                                                           uint32_t num_timesteps, float* dst) {
    if (!dst) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_download_and_flush_membrane_potential_history: dst is null");
    return guarded([&] {
        GpuSimulationState& S = state_of(state);
        if (!S.d_membrane_potential_history()) throw Error(BN_ERR_RUNTIME, "the state has no histories (BN_STATE_HISTORIES)");
        const std::vector<float> h = S.download_and_flush_membrane_potential_history(offset, num_neurons, num_timesteps);
        std::memcpy(dst, h.data(), h.size() * sizeof(float));
    });
}

int bn_state_download_weight_matrix(bn_state* state, uint32_t idx, float* dst) {
    if (!dst) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_download_weight_matrix: dst is null");
    return guarded([&] {
        const std::vector<float> w = state_of(state).download_weight_matrix(idx);
        std::memcpy(dst, w.data(), w.size() * sizeof(float));
    });
}

int bn_state_download(bn_state* state, const void* src, void* dst, size_t bytes) {
    if ((!src || !dst) && bytes) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_state_download: null pointer");
    return guarded([&] {
        GpuSimulationState& S = state_of(state);
        if (!bytes) return;
        BN_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S.stream()), "bn_state_download");
        BN_HIP_CHECK(hipStreamSynchronize(S.stream()), "bn_state_download sync");
    });
}

}  // extern "C"
