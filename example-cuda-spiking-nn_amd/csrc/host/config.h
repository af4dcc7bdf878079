// Config of one simulation run: the reference's flat JSON schema with the same keys and
// defaults (reference cpp_src/core/config.h:10-154), loaded with the same rules
// (cpp_src/core/config.cpp:68-318): missing keys keep defaults, global weight-generation
// keys are broadcast to the four matrices and then overridden per prefix, relative output
// paths are prefixed with output_directory, and the same validation errors are raised.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bn {

// reference cpp_src/generation/generate_weights.h:17-27
struct WeightGenerationParams {
    float x_location_length;
    float y_location_length;
    float max_distance;
    float spread_scale;
    float max_weight_value;
    bool weight_value_normal_dropoff;
    bool probabilistic_normal_not_uniform;
    float probability_multiplier;
    bool no_diagonal_connections;
    // extension: nonzero weights drawn as max_weight_value * U(0,1) from a second mt19937
    // stream (the connection mask stream is unchanged)
    bool weight_value_uniform;
};

struct Config {
    uint32_t simulation_steps = 10000;
    std::string output_directory = "../data/buryn_data/buryn_outputs_";
    std::string description;
    int rng_seed = 314;
    float timestep = 1e-4f;

    uint32_t num_excitatory_neurons = 100;
    uint32_t num_inhibitory_neurons = 100;
    float g_leak = 0.05e-6f;
    float membrane_capacitance = 1.0e-9f;
    float depolarization_slope_factor = 0.0065625f;
    float v_effective_spike_threshold = -0.0606250f;
    float v_excitatory_reversal = 0.0f;
    float v_inhibitory_reversal = -0.08f;
    float v_threshold = -0.04f;
    float v_reset = -0.06f;
    float v_rest = -0.07f;
    uint32_t refractory_steps = 50;

    float group_e_mean_background_excitatory_spikes = 0.0f;
    float group_e_mean_background_inhibitory_spikes = 0.0f;
    float group_e_background_excitatory_postsynaptic_r_trace_decay = 0.0f;
    float group_e_background_excitatory_postsynaptic_d_trace_decay = 0.0f;
    float group_e_background_inhibitory_postsynaptic_r_trace_decay = 0.0f;
    float group_e_background_inhibitory_postsynaptic_d_trace_decay = 0.0f;
    float group_e_background_excitatory_scaling_factor = 0.0f;
    float group_e_background_inhibitory_scaling_factor = 0.0f;
    float group_i_mean_background_excitatory_spikes = 0.0f;
    float group_i_mean_background_inhibitory_spikes = 0.0f;
    float group_i_background_excitatory_postsynaptic_r_trace_decay = 0.0f;
    float group_i_background_excitatory_postsynaptic_d_trace_decay = 0.0f;
    float group_i_background_inhibitory_postsynaptic_r_trace_decay = 0.0f;
    float group_i_background_inhibitory_postsynaptic_d_trace_decay = 0.0f;
    float group_i_background_excitatory_scaling_factor = 0.0f;
    float group_i_background_inhibitory_scaling_factor = 0.0f;

    std::string excitatory_locations_file = "../data/buryn_data/buryn_outputs_/locations/group_e_locations.txt";
    std::string inhibitory_locations_file = "../data/buryn_data/buryn_outputs_/locations/group_i_locations.txt";
    float excitatory_location_spacing = 1.0f;
    float inhibitory_location_spacing = 2.0f;
    uint32_t excitatory_side_points = 10;
    uint32_t inhibitory_side_points = 10;
    float side_length = 10.0f;

    float e_max_weight_value = 0.2235e-9f;
    float i_max_weight_value = 0.0578e-9f;
    float max_connection_distance = 3.0f;
    float connection_spread_scale = 1.0f;
    float connection_probability_multiplier = 1.0f;
    WeightGenerationParams weight_generation_params_e_to_e = {10.0f, 10.0f, 3.0f, 1.0f, 0.2235e-9f, false, true, 1.0f, true, false};
    WeightGenerationParams weight_generation_params_e_to_i = {10.0f, 10.0f, 3.0f, 1.0f, 0.2235e-9f, false, true, 1.0f, false, false};
    WeightGenerationParams weight_generation_params_i_to_e = {10.0f, 10.0f, 3.0f, 1.0f, 0.0578e-9f, false, false, 1.0f, false, false};
    WeightGenerationParams weight_generation_params_i_to_i = {10.0f, 10.0f, 3.0f, 1.0f, 0.0578e-9f, false, false, 1.0f, false, false};

    std::string weights_e_to_i_filename = "../data/buryn_data/buryn_outputs_/weights/weights_e_to_i.txt";
    std::string weights_i_to_e_filename = "../data/buryn_data/buryn_outputs_/weights/weights_i_to_e.txt";
    std::string weights_e_to_e_filename = "../data/buryn_data/buryn_outputs_/weights/weights_e_to_e.txt";
    std::string weights_i_to_i_filename = "../data/buryn_data/buryn_outputs_/weights/weights_i_to_i.txt";
    float tau_decay_e = 2e-3f;
    float tau_rise_e = 0.3e-3f;
    float tau_decay_i = 3e-3f;
    float tau_rise_i = 0.3e-3f;
    float postsynaptic_kernel_normalization_value = 1.0f;

    bool e_e_plasticity_enabled = true;
    float e_e_plasticity_ltp_A = 1.0e-12f;
    float e_e_plasticity_ltd_A = 1.0e-12f;
    float e_e_plasticity_ltp_B = 1.0e-12f;
    float e_e_plasticity_ltd_B = 1.0e-12f;
    float e_e_plasticity_tau_ltp = 0.02f;
    float e_e_plasticity_tau_ltd = 0.02f;
    float e_e_plasticity_triplet_tau_ltp = 0.02f;
    float e_e_plasticity_triplet_tau_ltd = 0.02f;
    float e_e_plasticity_weight_min = 0.0f;
    float e_e_plasticity_weight_max = 0.2235e-9f;
    bool i_e_plasticity_enabled = true;
    float i_e_plasticity_ltp_A = 1.0e-12f;
    float i_e_plasticity_ltd_A = 1.0e-12f;
    float i_e_plasticity_ltd_B = 1.0e-12f;
    float i_e_plasticity_tau_ltp = 0.02f;
    float i_e_plasticity_tau_ltd = 0.02f;
    float i_e_plasticity_weight_min = 0.0f;
    float i_e_plasticity_weight_max = 0.2235e-9f;

    std::vector<float> stimulus_input_values = {0.25e-9f};
    std::string stimulus_schedules_folder = "../data/buryn_data/buryn_outputs_/stimulus_schedules";
    std::vector<uint32_t> stimulus_selected_input_ids = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
    std::vector<uint32_t> stimulus_selected_input_post_group_offsets = std::vector<uint32_t>(10, 0);
    std::vector<uint32_t> stimulus_start_clock_steps = {5000};
    std::vector<uint32_t> stimulus_end_clock_steps = {6000};

    std::string e_spike_recording_file = "../data/buryn_data/buryn_outputs_/recordings/spikes_e.txt";
    std::string i_spike_recording_file = "../data/buryn_data/buryn_outputs_/recordings/spikes_i.txt";
    std::string voltage_recording_file = "../data/buryn_data/buryn_outputs_/recordings/voltages_step0.txt";
    std::vector<uint32_t> voltage_recording_neuron_ids = {10, 426};
    std::string e_e_weight_recording_file = "../data/buryn_data/buryn_outputs_/weights/weights_e_to_e_recordings.txt";

    // extension (engine tuning): rows of W summed per propagate partial for E->E, E->I, I->E,
    // I->I; 0 picks the engine's heuristic.  Fixes the summation order, so it is part of the
    // numerical definition of a run.
    std::vector<uint32_t> gemv_row_blocks = {0, 0, 0, 0};
    // extension (engine tuning): weight loads of the propagate kernel bypass the caches (1),
    // use them (0), or the engine decides from the size of the weight set (-1).
    int gemv_nontemporal_loads = -1;
    // extension: key of the background-noise Philox streams; -1 uses rng_seed (which also seeds
    // the connectivity), so one weight set can be run with several noise streams.
    long long background_noise_seed = -1;
    // extension: the E->E learning rule.  "triplet" (the reference's wiring, main_buryn.cpp:286-305)
    // or "pair": the reference's StdpPlasticity component (stdp_plasticity.cpp,
    // kernel_stdp_plasticity_update.cu) on E->E with e_e_plasticity_ltp_A, _ltd_A, _tau_ltp,
    // _tau_ltd, _weight_min and _weight_max (no triplet terms, no row normalisation).
    std::string e_e_plasticity_rule = "triplet";
};

Config load_config_json(const std::string& path);
Config load_config_json_text(const std::string& text);
// Every field as one JSON object (used to compare with the reference loader).
std::string config_to_json(const Config& c);

}  // namespace bn
