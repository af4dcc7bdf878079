// Host engine: the reference's component API (same class names, methods and step order) on
// top of the extern "C" launchers, plus the fused, graph-captured step.
//
// reference: cpp_src/gpu_core/gpu_simulation_state.h, cpp_src/neuron_groups/neuron_group.h,
// cpp_src/connections/connection.h, cpp_src/stimuli/stimulus.h, cpp_src/plasticities/*.h,
// cpp_src/monitors/*.h, cpp_src/core/system.h
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <exception>
#include <functional>
#include <fstream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../bn_persist.h"
#include "../bn_resident.h"
#include "../bn_step.h"
#include "buryn_amd.h"
#include "generation.h"

namespace bn {

constexpr uint32_t kRecordingIntervalSteps = 10000;  // buryn_definitions.h:13

// --------------------------------------------------------------------------------------
// GpuSimulationState — owns every device buffer (gpu_simulation_state.cu:35-104)
// --------------------------------------------------------------------------------------
struct StateOptions {
    bool reference_histories = true;   // [neuron][slot] spike/voltage histories (component path)
    bool fused_rings = false;          // spike-bit ring + voltage ring of recorded neurons
    std::vector<uint32_t> vrec_global_ids;
    bool lazy_normalize = false;
    uint32_t recording_interval_steps = kRecordingIntervalSteps;
    int64_t rng_seed = 314;  // Philox key (buryn::rng_seed, buryn_definitions.h:11)
    std::array<uint32_t, 4> gemv_row_block = {0, 0, 0, 0};  // 0 = heuristic
    int gemv_nontemporal = -1;  // weight loads past the caches: 1 yes, 0 no, -1 when W >> MALL
    bool overlap_steps = true;  // captured graphs overlap plasticity(t-1) with evolve(t)
    bool reference_layout = false;  // weight rows unpadded (row stride = post), as the reference allocates
};

class GpuSimulationState {
public:
    GpuSimulationState(uint32_t total_simulation_length, uint32_t num_excitatory_neurons,
                       uint32_t num_inhibitory_neurons, uint32_t num_stimulus_events,
                       uint32_t num_plasticities, StateOptions opt = StateOptions());
    ~GpuSimulationState();
    GpuSimulationState(const GpuSimulationState&) = delete;
    GpuSimulationState& operator=(const GpuSimulationState&) = delete;

    uint64_t total_allocated_bytes() const { return allocated_; }
    std::string memory_report() const;
    hipStream_t stream() const { return stream_; }
    const StateOptions& options() const { return opt_; }

    float* d_membrane_potentials() const { return v_; }
    float* d_g_excitatory() const { return ge_; }
    float* d_g_inhibitory() const { return gi_; }
    float* d_i_stimulus() const { return istim_; }
    uint32_t* d_refractory_counter() const { return refr_; }
    uint32_t* d_spikes() const { return spikes_; }
    float* d_background_g_excitatory_synaptic_conductances() const { return bg_ge_; }
    float* d_background_g_inhibitory_synaptic_conductances() const { return bg_gi_; }
    float* d_background_excitatory_postsynaptic_r_traces() const { return bre_; }
    float* d_background_excitatory_postsynaptic_d_traces() const { return bde_; }
    float* d_background_inhibitory_postsynaptic_r_traces() const { return bri_; }
    float* d_background_inhibitory_postsynaptic_d_traces() const { return bdi_; }
    bn_device_clock* d_clock() const { return clock_; }
    uint32_t* d_status() const { return status_; }

    float* d_w(uint32_t idx) const { return w_[idx]; }
    float* d_w_ee() const { return w_[0]; }
    float* d_w_ei() const { return w_[1]; }
    float* d_w_ie() const { return w_[2]; }
    float* d_w_ii() const { return w_[3]; }
    uint32_t weight_ld(uint32_t idx) const { return ld_[idx]; }
    uint32_t weight_pre(uint32_t idx) const { return pre_[idx]; }
    uint32_t weight_post(uint32_t idx) const { return post_[idx]; }
    float* d_postsynaptic_r_traces(uint32_t idx) const { return tr_r_[idx]; }
    float* d_postsynaptic_d_traces(uint32_t idx) const { return tr_d_[idx]; }
    float* d_summed_postsynaptic_conductances(uint32_t idx) const { return tr_x_[idx]; }
    double* d_gemv_partials(uint32_t idx) const { return partials_[idx]; }
    uint32_t gemv_row_block(uint32_t idx) const { return row_block_[idx]; }

    uint32_t* d_timestep_event_count() const { return stim_count_; }
    uint32_t* d_stimulus_event_neuron_ids() const { return stim_ids_; }
    uint32_t* d_stimulus_event_post_group_offsets() const { return stim_offs_; }
    float* d_stimulus_event_amplitude_changes() const { return stim_amps_; }

    float* d_pre_traces() const { return pre_tr_; }
    float* d_post_traces() const { return post_tr_; }
    float* d_triplet_pre_traces() const { return tpre_tr_; }
    float* d_triplet_post_traces() const { return tpost_tr_; }
    float* d_pre_weight_sums() const { return pre_wsum_; }

    uint32_t* d_spike_history() const { return spike_hist_; }
    float* d_membrane_potential_history() const { return v_hist_; }
    uint32_t* d_spike_bits() const { return spike_bits_; }
    uint32_t* d_spike_bits_ring() const { return spike_ring_; }
    float* d_v_ring() const { return v_ring_; }
    const int32_t* d_vrec_index() const { return vrec_index_; }
    uint32_t spike_words() const { return words_; }
    uint32_t n_vrec() const { return (uint32_t)opt_.vrec_global_ids.size(); }

    // lazy normalisation of the triplet matrix (plasticity_offset 0)
    float* d_row_scale() const { return row_scale_; }
    double* d_row_sum() const { return row_sum_; }

    // Poisson tables per group (0 E, 1 I) and kind (0 exc, 1 inh)
    const uint32_t* d_poisson_table(uint32_t group, uint32_t kind) const { return poisson_[group][kind]; }
    uint32_t poisson_table_len(uint32_t group, uint32_t kind) const { return poisson_len_[group][kind]; }
    void set_poisson_table(uint32_t group, uint32_t kind, float mean);

    void upload_neuron_group_view(const NeuronGroupView& v, const struct NeuronGroupInitHostData& h);
    void upload_connection_view(const ConnectionView& v, const struct ConnectionInitHostData& h);
    void upload_stimulus_view(const StimulusView& v, const struct StimulusInitHostData& h);
    void upload_plasticity_view(const PlasticityView& v, const struct PlasticityInitHostData& h);

    std::vector<uint32_t> download_and_flush_spike_history(uint32_t offset, uint32_t num_neurons, uint32_t num_timesteps);
    std::vector<float> download_and_flush_membrane_potential_history(uint32_t offset, uint32_t num_neurons,
                                                                     uint32_t num_timesteps);
    // Fused-path rings: [interval][words] spike bits (downloaded once per record step and
    // zeroed), and [interval][n_vrec] voltages.
    const std::vector<uint32_t>& spike_ring_snapshot(uint32_t clock_step);
    std::vector<float> download_and_flush_v_ring();
    std::vector<uint32_t> download_spike_ring_rows(uint32_t first_step, uint32_t num_steps, uint32_t clock_step) const;
    // The rings' rows up to `step` went to the recorders (their copies were queued on the stream).
    void mark_rings_flushed(uint32_t step) { ring_snapshot_step_ = step; }
    // Dense [pre][post] true weights (lazy normalisation: fl32(s_i * M_ij)).
    std::vector<float> download_weight_matrix(uint32_t weight_matrix_idx) const;
    std::string lookup_weight_matrix_name(uint32_t weight_matrix_idx) const;
    void setup_lazy_normalization(uint32_t weight_matrix_idx);  // after W_EE upload
    bool lazy_active() const { return lazy_active_; }
    // Every device buffer of the state, in allocation order (snapshot / resume).
    std::vector<uint8_t> snapshot() const;
    void restore(const std::vector<uint8_t>& bytes);

    uint32_t total_simulation_length() const { return total_simulation_length_; }
    uint32_t num_excitatory_neurons() const { return ne_; }
    uint32_t num_inhibitory_neurons() const { return ni_; }
    uint32_t total_neurons() const { return n_; }
    uint32_t num_stimulus_events() const { return num_stimulus_events_; }
    uint32_t num_plasticities() const { return num_plasticities_; }
    void check_status();  // throws if a kernel flagged an error

private:
    template <typename T>
    T* alloc(size_t n, const char* what);
    void require_upload_bounds(const std::string& field, uint32_t offset, uint32_t count) const;

    StateOptions opt_;
    uint32_t total_simulation_length_, ne_, ni_, n_, num_stimulus_events_, num_plasticities_;
    hipStream_t stream_ = nullptr;
    std::vector<void*> allocations_;
    std::vector<size_t> allocation_bytes_;
    uint64_t allocated_ = 0;

    float *v_, *ge_, *gi_, *istim_, *bg_ge_, *bg_gi_, *bre_, *bde_, *bri_, *bdi_;
    uint32_t *refr_, *spikes_;
    bn_device_clock* clock_;
    uint32_t* status_;
    float* w_[4];
    uint32_t ld_[4], pre_[4], post_[4], row_block_[4];
    float *tr_r_[4], *tr_d_[4], *tr_x_[4];
    double* partials_[4];
    uint32_t *stim_count_ = nullptr, *stim_ids_ = nullptr, *stim_offs_ = nullptr;
    float* stim_amps_ = nullptr;
    float *pre_tr_, *post_tr_, *tpre_tr_, *tpost_tr_, *pre_wsum_;
    uint32_t* spike_hist_ = nullptr;
    float* v_hist_ = nullptr;
    uint32_t* spike_bits_ = nullptr;
    uint32_t* spike_ring_ = nullptr;
    float* v_ring_ = nullptr;
    int32_t* vrec_index_ = nullptr;
    uint32_t words_ = 0;
    float* row_scale_ = nullptr;
    double* row_sum_ = nullptr;
    bool lazy_active_ = false;
    uint32_t lazy_matrix_ = 0;
    uint32_t* poisson_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    uint32_t poisson_len_[2][2] = {{0, 0}, {0, 0}};
    std::vector<uint32_t> ring_snapshot_;
    int64_t ring_snapshot_step_ = -1;
};

// --------------------------------------------------------------------------------------
// Components
// --------------------------------------------------------------------------------------
struct NeuronGroupParams {  // neuron_group_params.h:7-33
    float timestep = 1e-4f, g_leak = 10.0f, membrane_capacitance = 0.1f, depolarization_slope_factor = 2.0f;
    float v_effective_spike_threshold = -50.0f, v_excitatory_reversal = 0.0f, v_inhibitory_reversal = -80.0f;
    float v_threshold = -45.0f, v_reset = -55.0f, v_rest = -65.0f;
    uint32_t refractory_steps = 0;
    bool is_inhibitory = false;
    float mean_background_excitatory_spikes = 0.0f, mean_background_inhibitory_spikes = 0.0f;
    float background_excitatory_postsynaptic_r_trace_decay = 0.0f, background_excitatory_postsynaptic_d_trace_decay = 0.0f;
    float background_inhibitory_postsynaptic_r_trace_decay = 0.0f, background_inhibitory_postsynaptic_d_trace_decay = 0.0f;
    float background_excitatory_scaling_factor = 0.0f, background_inhibitory_scaling_factor = 0.0f;
    uint32_t offset = 0, num_neurons = 0;
};

struct NeuronGroupInitHostData {
    std::vector<float> membrane_potentials, g_excitatory, g_inhibitory, i_stimulus;
    std::vector<uint32_t> refractory_counter, spikes;
    std::vector<float> background_g_excitatory_synaptic_conductances, background_g_inhibitory_synaptic_conductances;
    std::vector<float> background_excitatory_postsynaptic_r_traces, background_excitatory_postsynaptic_d_traces;
    std::vector<float> background_inhibitory_postsynaptic_r_traces, background_inhibitory_postsynaptic_d_traces;
};

class NeuronGroup {
public:
    NeuronGroup(std::vector<Location> neuron_locations, NeuronGroupParams params, GpuSimulationState& gpu_state);
    NeuronGroupInitHostData initialize_data();
    void make_gpu_view();
    void upload_data(const NeuronGroupInitHostData& h);
    void evolve();
    uint32_t offset() const { return params_.offset; }
    uint32_t num_neurons() const { return params_.num_neurons; }
    bool is_inhibitory() const { return params_.is_inhibitory; }
    const NeuronGroupView& gpu_view() const { return gpu_view_; }
    const std::vector<Location>& neuron_locations() const { return neuron_locations_; }

private:
    GpuSimulationState& gpu_state_;
    NeuronGroupParams params_;
    std::vector<Location> neuron_locations_;
    NeuronGroupView gpu_view_{};
};

struct ConnectionParams {  // connection_params.h:7-15
    float timestep = 0.0f, decay_time = 0.0f, rise_time = 0.0f, postsynaptic_kernel_normalization_value = 0.0f;
    uint32_t weight_matrix_idx = 0;
};

struct ConnectionInitHostData {
    std::vector<float> weight_matrix;  // dense [pre][post]
    std::vector<float> postsynaptic_r_traces, postsynaptic_d_traces, summed_postsynaptic_conductances;
};

class Connection {
public:
    // weight_matrix: dense row-major [pre][post]
    Connection(NeuronGroup& pre_group, NeuronGroup& post_group, const std::vector<float>& weight_matrix,
               ConnectionParams params, GpuSimulationState& gpu_state);
    void make_gpu_view();
    void propagate();
    uint32_t weight_matrix_idx() const { return params_.weight_matrix_idx; }
    uint32_t pre_group_offset() const { return pre_group_.offset(); }
    uint32_t post_group_offset() const { return post_group_.offset(); }
    uint32_t pre_group_num_neurons() const { return pre_group_.num_neurons(); }
    uint32_t post_group_num_neurons() const { return post_group_.num_neurons(); }
    const ConnectionView& gpu_view() const { return gpu_view_; }
    const NeuronGroup& pre_group() const { return pre_group_; }
    const NeuronGroup& post_group() const { return post_group_; }

private:
    NeuronGroup& pre_group_;
    NeuronGroup& post_group_;
    ConnectionParams params_;
    ConnectionView gpu_view_{};
    GpuSimulationState& gpu_state_;
};

struct StimulusInitHostData {
    std::vector<uint32_t> timestep_event_count, stimulus_event_neuron_ids, stimulus_event_post_group_offsets;
    std::vector<float> stimulus_event_amplitude_changes;
};

class Stimulus {
public:
    Stimulus(std::vector<uint32_t> timestep_event_count, std::vector<uint32_t> stimulus_event_neuron_ids,
             std::vector<uint32_t> stimulus_event_post_group_offsets, std::vector<float> stimulus_event_amplitude_changes,
             GpuSimulationState& gpu_state);
    void apply();  // host cursor, as stimulus.cpp:46-60
    const StimulusView& gpu_view() const { return gpu_view_; }

private:
    std::vector<uint32_t> timestep_event_count_;
    uint32_t next_event_idx_ = 0, current_timestep_ = 0, total_stimulus_events_ = 0;
    StimulusView gpu_view_{};
    GpuSimulationState& gpu_state_;
};

struct PlasticityInitHostData {
    std::vector<float> pre_traces, post_traces, triplet_pre_traces, triplet_post_traces, pre_weight_sums;
};

class Plasticity {
public:
    Plasticity(Connection& connection, PlasticityParams params, GpuSimulationState& gpu_state, uint32_t kind);
    virtual ~Plasticity() = default;
    virtual void update() = 0;
    uint32_t kind() const { return kind_; }
    virtual bool normalizes() const { return false; }
    const PlasticityView& gpu_view() const { return gpu_view_; }
    const PlasticityParams& params() const { return params_; }
    const Connection& connection() const { return connection_; }

protected:
    void make_gpu_view();
    Connection& connection_;
    PlasticityParams params_;
    PlasticityView gpu_view_{};
    GpuSimulationState& gpu_state_;
    uint32_t kind_;
};

class TripletPlasticity : public Plasticity {  // triplet_plasticity.cpp
public:
    TripletPlasticity(Connection& connection, PlasticityParams params, GpuSimulationState& gpu_state);
    void update() override;
    bool normalizes() const override { return true; }
};

class IstdpPlasticity : public Plasticity {  // istdp_plasticity.cpp
public:
    IstdpPlasticity(Connection& connection, PlasticityParams params, GpuSimulationState& gpu_state);
    void update() override;
};

class StdpPlasticity : public Plasticity {  // stdp_plasticity.cpp
public:
    StdpPlasticity(Connection& connection, PlasticityParams params, GpuSimulationState& gpu_state);
    void update() override;
};

// --------------------------------------------------------------------------------------
// Monitors (monitor.h, spike_recorder.h, voltage_recorder.h, weight_recorder.h)
// --------------------------------------------------------------------------------------
class Monitor {
public:
    explicit Monitor(GpuSimulationState& s) : gpu_state_(s) {}
    virtual ~Monitor() = default;
    virtual void historicize(uint32_t clock_step) = 0;
    virtual void record(uint32_t clock_step) = 0;
    // Recorders of the fused step's rings: write the interval ending at clock_step from host
    // copies of the spike-bit ring [interval][words] and the voltage ring [interval][n_vrec]
    // (called by the recording pipeline's writer thread).
    virtual bool from_rings() const { return false; }
    virtual void record_from_rings(uint32_t, const uint32_t*, const float*) {}
    bool fused = false;  // rings are written by the fused step
    bool quiet = false;

protected:
    GpuSimulationState& gpu_state_;
};

class SpikeRecorder : public Monitor {
public:
    SpikeRecorder(const NeuronGroup& group, const std::string& output_filepath, uint32_t recording_interval_steps,
                  GpuSimulationState& gpu_state);
    void historicize(uint32_t clock_step) override;
    void record(uint32_t clock_step) override;
    bool from_rings() const override { return fused; }
    void record_from_rings(uint32_t clock_step, const uint32_t* spike_ring, const float* v_ring) override;
    bool good() const { return output_.good(); }

private:
    const NeuronGroup& group_;
    std::ofstream output_;
    uint32_t recording_interval_steps_;
    SpikeRecorderView gpu_view_{};
};

class VoltageRecorder : public Monitor {
public:
    VoltageRecorder(const NeuronGroup& group, const std::string& output_filepath, std::vector<uint32_t> selected_neurons,
                    uint32_t recording_interval_steps, GpuSimulationState& gpu_state);
    void historicize(uint32_t clock_step) override;
    void record(uint32_t clock_step) override;
    bool from_rings() const override { return fused; }
    void record_from_rings(uint32_t clock_step, const uint32_t* spike_ring, const float* v_ring) override;
    const std::vector<uint32_t>& selected_neurons() const { return selected_neurons_; }

private:
    void write_files(uint32_t clock_step, const float* v, uint32_t stride_n, const int64_t* col);
    const NeuronGroup& group_;
    std::string output_filepath_;
    std::vector<uint32_t> selected_neurons_;
    uint32_t recording_interval_steps_;
    VoltageRecorderView gpu_view_{};
};

class WeightRecorder : public Monitor {
public:
    WeightRecorder(const Connection& connection, const std::string& output_filepath, GpuSimulationState& gpu_state);
    void historicize(uint32_t) override {}
    void record(uint32_t clock_step) override;

private:
    const Connection& connection_;
    std::string output_filepath_;
};

// --------------------------------------------------------------------------------------
// Recording pipeline of the fused step: at a record step the rings' copies to pinned host
// memory (double-buffered) and their clearing are queued on the state's stream, and a writer
// thread writes the recorders' files while the next interval's steps run.
// --------------------------------------------------------------------------------------
class RecordPipeline {
public:
    RecordPipeline(GpuSimulationState& s, std::vector<Monitor*> monitors);
    ~RecordPipeline();
    RecordPipeline(const RecordPipeline&) = delete;
    RecordPipeline& operator=(const RecordPipeline&) = delete;
    // Queue the interval ending at `step` (waits only while both host buffers are being written).
    void submit(uint32_t step);
    // Wait until every queued interval is written; rethrows the writer's first error.
    void drain();
    double writer_seconds() const { return writer_seconds_; }   // formatting and writing, all intervals
    uint32_t intervals() const { return intervals_; }

private:
    struct Slot {
        uint32_t* spk = nullptr;
        float* v = nullptr;
        hipEvent_t ready = nullptr;
        bool busy = false;
    };
    void writer();
    GpuSimulationState& s_;
    std::vector<Monitor*> monitors_;
    size_t spk_elems_ = 0, v_elems_ = 0;
    Slot slot_[2];
    int next_ = 0;
    int device_ = 0;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::pair<int, uint32_t>> jobs_;
    bool stop_ = false;
    std::exception_ptr error_;
    double writer_seconds_ = 0.0;
    uint32_t intervals_ = 0;
    std::thread thread_;
};

// --------------------------------------------------------------------------------------
// System (system.h) + the fused step
// --------------------------------------------------------------------------------------
class System {
public:
    explicit System(GpuSimulationState& s) : state_(s) {}
    ~System();
    void add_neuron_group(std::unique_ptr<NeuronGroup> g);
    void add_connection(std::unique_ptr<Connection> c);
    void add_stimulus(std::unique_ptr<Stimulus> s);
    void add_plasticity(std::unique_ptr<Plasticity> p);
    void add_monitor(std::unique_ptr<Monitor> m);
    void add_step_callback(std::function<void(uint32_t)> cb);

    // Component path: the reference's order, one launcher per component (system.cpp:43-63).
    void step_the_clock(uint32_t clock_step);
    // Reference run loop (system.cpp:65-83); uses the fused step when enabled.
    void run(uint32_t total_clock_steps);

    // Fused step: build the plans, optionally capture `steps_per_graph` steps in a hipGraph.
    void enable_fused(bool use_graph, uint32_t steps_per_graph);
    // Steps per captured graph of the fused step (even: the overlapped form); re-captures the
    // graph only (plans and state are untouched).
    void set_steps_per_graph(uint32_t steps_per_graph);
    uint32_t steps_per_graph() const { return steps_per_graph_; }
    uint32_t graph_steps() const { return fused_ && use_graph_ ? steps_per_graph_ : 0; }
    uint32_t graph_launches() const { return graph_launches_; }
    bool fused() const { return fused_; }
    void launch_fused_steps(uint32_t n);           // enqueue n steps (graph or eager)
    void launch_fused_step_eager(hipEvent_t gemv_begin, hipEvent_t gemv_end);
    const GemvPlan& gemv_plan() const { return gemv_; }
    // The fused propagate writes part sums (GemvPlan::parts) when the shapes allow; set before
    // enable_fused.
    void set_gemv_part_sums(bool on) { gemv_part_sums_ = on; }
    bool gemv_part_sums() const { return gemv_.parts != 0; }
    const EvolvePlan& evolve_plan() const { return evolve_; }
    const PlastStepPlan& plast_step_plan() const { return plast_step_; }
    bool has_norm_tail() const { return tail_.norm_rows != 0; }
    // Captured graphs overlap the plasticity of step t-1 with the evolve of step t.
    bool overlapped() const { return overlap_; }
    // The plans of the overlapped form for graphs of `steps_per_graph` (even) steps, built on
    // first use (also without a graph of this system's own, for the multi-network step).
    void prepare_overlap(uint32_t steps_per_graph);
    const EvolvePlan& evolve_even() const { return evolve_even_; }
    const EvolvePlan& evolve_odd() const { return evolve_odd_; }
    const PlastStepPlan& plast_staging() const { return plast_staging_; }
    const PlastStepPlan& plast_current() const { return plast_current_; }
    const PlastStepPlan& plast_final() const { return plast_final_; }
    // Steps run by an external launcher on this system's plans (several networks in one launch).
    void note_external_steps(uint32_t n) { clock_step_ += n; }
    // After a state restore: the host's step counter follows the restored device clock.
    void set_clock_step(uint32_t step) { clock_step_ = step; }
    uint32_t kernels_per_step() const { return kernels_per_step_; }
    // Conductances after the last propagate, for state downloads (fused path keeps them as
    // base + pending GEMV partials).
    void materialize_g(float* d_ge_out, float* d_gi_out);
    // Trace set used by the fused step for connection idx (EE/EI share one, IE/II another).
    const TraceSet* trace_set_for_connection(uint32_t idx) const;

    // Persistent step: one cooperative launch runs a whole run of steps (bn_persist.h).  rows:
    // rows per wave chunk for each connection (0: chosen from the network's shape and the
    // device's co-resident blocks).  Re-plans only; the state is untouched.
    void enable_persistent(const std::array<uint32_t, 4>& rows);
    bool persistent() const { return persist_; }
    uint32_t persistent_blocks() const { return persist_ ? persist_plan_.blocks : 0u; }
    uint32_t persistent_rows(uint32_t c) const { return persist_ ? persist_plan_.c[c].rb : 0u; }
    uint32_t persistent_items() const { return persist_items_; }
    // Rows per chunk of the resident-weight kernel (0: weights streamed each step).
    uint32_t persistent_resident_rows() const { return persist_ ? persist_plan_.resident_rows : 0u; }
    // Allow (default) or forbid resident weights; the plan is rebuilt.
    void set_persistent_resident(bool allow);
    // Polls a block of the persistent step makes before it abandons a wait (and the launch).
    void set_persistent_wait_limit(uint32_t polls);

    // Resident step: one cooperative launch per run, the weights in registers for the whole run
    // (bn_resident.h).  Builds the plan; the state is untouched.
    void enable_resident();
    bool resident() const { return resident_; }
    // Cooperative launch (default) or an ordinary one (same kernel; co-residency then rests on
    // an idle device, the bounded waits turning a shortfall into an error).
    void set_resident_cooperative(bool cooperative);
    bool resident_cooperative() const { return resident_cooperative_; }
    const RPlan& resident_plan() const { return resident_plan_; }

    uint32_t clock_step() const { return clock_step_; }
    // Fused step: recordings written by the pipeline's writer thread while the steps run (default)
    // or on the calling thread at each record step, as the component path does.
    void set_async_recording(bool on);
    bool async_recording() const { return async_recording_; }
    // Host seconds spent on recordings: on the calling thread at record steps, and (pipeline)
    // in the writer thread; and the number of record steps.
    double record_call_seconds() const { return record_call_seconds_; }
    double record_writer_seconds() const { return pipeline_ ? pipeline_->writer_seconds() : 0.0; }
    uint32_t record_steps() const { return record_steps_; }
    bool quiet = false;
    const std::vector<std::unique_ptr<NeuronGroup>>& neuron_groups() const { return neuron_groups_; }
    const std::vector<std::unique_ptr<Connection>>& connections() const { return connections_; }
    const std::vector<std::unique_ptr<Plasticity>>& plasticities() const { return plasticities_; }
    const std::vector<std::unique_ptr<Monitor>>& monitors() const { return monitors_; }

private:
    void build_plans();
    void build_overlap_plans();
    void build_event_starts();
    void launch_persistent_steps(uint32_t n);
    void launch_resident_steps(uint32_t n);
    void* persist_alloc(size_t bytes);
    void enqueue_fused_step(hipEvent_t gemv_begin, hipEvent_t gemv_end);
    void enqueue_overlapped_steps(uint32_t n);

    GpuSimulationState& state_;
    std::vector<std::unique_ptr<NeuronGroup>> neuron_groups_;
    std::vector<std::unique_ptr<Connection>> connections_;
    std::vector<std::unique_ptr<Stimulus>> stimuli_;
    std::vector<std::unique_ptr<Plasticity>> plasticities_;
    std::vector<std::unique_ptr<Monitor>> monitors_;
    std::vector<std::function<void(uint32_t)>> step_callbacks_;
    uint32_t clock_step_ = 0;
    uint32_t graph_launches_ = 0;

    bool fused_ = false, use_graph_ = false;
    uint32_t steps_per_graph_ = 1;
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    EvolvePlan evolve_{};
    GemvPlan gemv_{};
    bool gemv_part_sums_ = false;
    PlastPlan plast_{};
    PlastStepPlan plast_step_{};
    TailPlan tail_{};
    std::vector<int> conn_trace_set_;
    uint32_t kernels_per_step_ = 0;
    void capture_graph();  // the graph of steps_per_graph_ steps in the current form
    // Overlapped graphs: even steps of a graph evolve from the current spike bits and
    // plasticity traces into the staging ones, odd steps back; the plasticity of a step reads
    // the buffers its evolve wrote; the last plasticity of a graph carries the stimulus/clock
    // tail.
    bool overlap_ = false;
    EvolvePlan evolve_even_{}, evolve_odd_{};
    PlastStepPlan plast_staging_{}, plast_current_{}, plast_final_{};
    GemvPlan gemv_even_{}, gemv_odd_{};   // overlapped graphs: the compaction reads the step's bits
    uint32_t* pre_list_buf_ = nullptr;    // [kPreListCap + kMaxPlast]: PreCompact list and counts
    uint32_t* d_ev_start_ = nullptr;
    bool overlap_built_ = false;
    std::vector<void*> staging_;
    // Persistent step
    bool persist_ = false;
    PPlan persist_plan_{};
    PPlan* d_persist_plan_ = nullptr;
    uint32_t* persist_counters_ = nullptr;
    size_t persist_counter_words_ = 0;
    uint32_t persist_nontemporal_ = 0;
    uint32_t persist_items_ = 0;
    bool persist_resident_ = true;
    std::array<uint32_t, 4> persist_rows_{};
    uint32_t persist_wait_limit_ = 1u << 18;  // about a quarter second of polling
    std::vector<void*> persist_bufs_;
    std::vector<std::pair<void*, size_t>> persist_ll_bufs_;  // tagged-word buffers (64-bit words)
    uint32_t persist_ll_next_ = 0;                          // the clock after the last resident launch
    // Resident step
    bool resident_ = false;
    RPlan resident_plan_{};
    RPlan* d_resident_plan_ = nullptr;
    std::vector<std::pair<void*, size_t>> resident_bufs_;  // tagged-word buffers (64-bit words)
    uint32_t resident_next_ = 0;                           // the clock after the last launch
    bool resident_cooperative_ = true;
    bool async_recording_ = true;
    std::unique_ptr<RecordPipeline> pipeline_;
    double record_call_seconds_ = 0.0;
    uint32_t record_steps_ = 0;
};

}  // namespace bn
