// Test infrastructure only — never linked into the product.
//
// Driver around the REFERENCE's own host-side generator and config loader, compiled from the
// reference sources where they lie (see oracle/Makefile, target `ref`): it loads a config
// with the reference's load_config_json (cpp_src/core/config.cpp:68-318), generates locations
// (generate_location.cpp:5-47), the four dense weight matrices (generate_weights.cpp:10-107)
// and the stimulus schedule (generate_stimulus_schedules.cpp:20-228), and writes
//   <out>/config_dump.json    every Config field (same layout as the engine's dump)
//   <out>/w<k>.bin             raw little-endian float32 [pre][post], k = 0 EE,1 EI,2 IE,3 II
//   <out>/stim_*.bin           timestep_event_counts, neuron_ids, post_group_offsets, amplitudes
//   <out>/ref_files/...        the reference's own text files (locations, weights, schedules)
// The golden fixtures under tests/golden/ are digests of these outputs.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "buryn_definitions.h"
#include "config.h"
#include "generate_location.h"
#include "generate_stimulus_schedules.h"
#include "generate_weights.h"

namespace {

std::string num(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}
std::string q(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if (c == '\n') {
            o += "\\n";
            continue;
        }
        o += c;
    }
    return o + "\"";
}

struct Dump {
    std::ostringstream o;
    bool first = true;
    void key(const std::string& k) {
        o << (first ? "{" : ",") << "\n  " << q(k) << ": ";
        first = false;
    }
    void f(const std::string& k, float v) { key(k); o << num((double)v); }
    void u(const std::string& k, unsigned long long v) { key(k); o << v; }
    void i(const std::string& k, long long v) { key(k); o << v; }
    void b(const std::string& k, bool v) { key(k); o << (v ? "true" : "false"); }
    void s(const std::string& k, const std::string& v) { key(k); o << q(v); }
    void fa(const std::string& k, const std::vector<float>& v) {
        key(k);
        o << "[";
        for (size_t n = 0; n < v.size(); ++n) o << (n ? ", " : "") << num((double)v[n]);
        o << "]";
    }
    void ua(const std::string& k, const std::vector<uint32_t>& v) {
        key(k);
        o << "[";
        for (size_t n = 0; n < v.size(); ++n) o << (n ? ", " : "") << v[n];
        o << "]";
    }
    void w(const std::string& k, const generate_weights::WeightGenerationParams& p) {
        f(k + ".x_location_length", p.x_location_length);
        f(k + ".y_location_length", p.y_location_length);
        f(k + ".max_distance", p.max_distance);
        f(k + ".spread_scale", p.spread_scale);
        f(k + ".max_weight_value", p.max_weight_value);
        b(k + ".weight_value_normal_dropoff", p.weight_value_normal_dropoff);
        b(k + ".probabilistic_normal_not_uniform", p.probabilistic_normal_not_uniform);
        f(k + ".probability_multiplier", p.probability_multiplier);
        b(k + ".no_diagonal_connections", p.no_diagonal_connections);
    }
};

std::string dump(const Config& c) {
    Dump d;
#define F(x) d.f(#x, c.x);
#define U(x) d.u(#x, c.x);
#define S(x) d.s(#x, c.x);
#define B(x) d.b(#x, c.x);
    U(simulation_steps) S(output_directory) S(description) d.i("rng_seed", c.rng_seed); F(timestep)
    U(num_excitatory_neurons) U(num_inhibitory_neurons) F(g_leak) F(membrane_capacitance)
    F(depolarization_slope_factor) F(v_effective_spike_threshold) F(v_excitatory_reversal)
    F(v_inhibitory_reversal) F(v_threshold) F(v_reset) F(v_rest) U(refractory_steps)
    F(group_e_mean_background_excitatory_spikes) F(group_e_mean_background_inhibitory_spikes)
    F(group_e_background_excitatory_postsynaptic_r_trace_decay) F(group_e_background_excitatory_postsynaptic_d_trace_decay)
    F(group_e_background_inhibitory_postsynaptic_r_trace_decay) F(group_e_background_inhibitory_postsynaptic_d_trace_decay)
    F(group_e_background_excitatory_scaling_factor) F(group_e_background_inhibitory_scaling_factor)
    F(group_i_mean_background_excitatory_spikes) F(group_i_mean_background_inhibitory_spikes)
    F(group_i_background_excitatory_postsynaptic_r_trace_decay) F(group_i_background_excitatory_postsynaptic_d_trace_decay)
    F(group_i_background_inhibitory_postsynaptic_r_trace_decay) F(group_i_background_inhibitory_postsynaptic_d_trace_decay)
    F(group_i_background_excitatory_scaling_factor) F(group_i_background_inhibitory_scaling_factor)
    S(excitatory_locations_file) S(inhibitory_locations_file) F(excitatory_location_spacing)
    F(inhibitory_location_spacing) U(excitatory_side_points) U(inhibitory_side_points) F(side_length)
    F(e_max_weight_value) F(i_max_weight_value) F(max_connection_distance) F(connection_spread_scale)
    F(connection_probability_multiplier)
    d.w("weight_generation_params_e_to_e", c.weight_generation_params_e_to_e);
    d.w("weight_generation_params_e_to_i", c.weight_generation_params_e_to_i);
    d.w("weight_generation_params_i_to_e", c.weight_generation_params_i_to_e);
    d.w("weight_generation_params_i_to_i", c.weight_generation_params_i_to_i);
    S(weights_e_to_i_filename) S(weights_i_to_e_filename) S(weights_e_to_e_filename) S(weights_i_to_i_filename)
    F(tau_decay_e) F(tau_rise_e) F(tau_decay_i) F(tau_rise_i) F(postsynaptic_kernel_normalization_value)
    B(e_e_plasticity_enabled) F(e_e_plasticity_ltp_A) F(e_e_plasticity_ltd_A) F(e_e_plasticity_ltp_B)
    F(e_e_plasticity_ltd_B) F(e_e_plasticity_tau_ltp) F(e_e_plasticity_tau_ltd) F(e_e_plasticity_triplet_tau_ltp)
    F(e_e_plasticity_triplet_tau_ltd) F(e_e_plasticity_weight_min) F(e_e_plasticity_weight_max)
    B(i_e_plasticity_enabled) F(i_e_plasticity_ltp_A) F(i_e_plasticity_ltd_A) F(i_e_plasticity_ltd_B)
    F(i_e_plasticity_tau_ltp) F(i_e_plasticity_tau_ltd) F(i_e_plasticity_weight_min) F(i_e_plasticity_weight_max)
    d.fa("stimulus_input_values", c.stimulus_input_values);
    S(stimulus_schedules_folder)
    d.ua("stimulus_selected_input_ids", c.stimulus_selected_input_ids);
    d.ua("stimulus_selected_input_post_group_offsets", c.stimulus_selected_input_post_group_offsets);
    d.ua("stimulus_start_clock_steps", c.stimulus_start_clock_steps);
    d.ua("stimulus_end_clock_steps", c.stimulus_end_clock_steps);
    S(e_spike_recording_file) S(i_spike_recording_file) S(voltage_recording_file)
    d.ua("voltage_recording_neuron_ids", c.voltage_recording_neuron_ids);
    S(e_e_weight_recording_file)
#undef F
#undef U
#undef S
#undef B
    d.o << "\n}\n";
    return d.o.str();
}

template <typename T>
void write_bin(const std::string& path, const std::vector<T>& v) {
    std::ofstream o(path, std::ios::binary);
    o.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

void write_matrix(const std::string& path, const std::vector<std::vector<float>>& m) {
    std::ofstream o(path, std::ios::binary);
    for (const auto& row : m) o.write(reinterpret_cast<const char*>(row.data()), (std::streamsize)(row.size() * sizeof(float)));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: ref_hostgen <config.json> <outdir> [text]\n";
        return 2;
    }
    const std::string out = argv[2];
    const bool text = argc > 3 && std::string(argv[3]) == "text";
    Config c = load_config_json(argv[1]);
    {
        std::ofstream o(out + "/config_dump.json");
        o << dump(c);
    }
    buryn::rng_seed = c.rng_seed;
    buryn::timestep = c.timestep;
    const std::string rf = out + "/ref_files/";
    auto le = generate_location::generate_2d_grid_locations(rf + "group_e_locations.txt", c.num_excitatory_neurons,
                                                           c.excitatory_location_spacing, c.excitatory_side_points,
                                                           c.excitatory_side_points);
    auto li = generate_location::generate_2d_grid_locations(rf + "group_i_locations.txt", c.num_inhibitory_neurons,
                                                           c.inhibitory_location_spacing, c.inhibitory_side_points,
                                                           c.inhibitory_side_points);
    const std::string sink = "/dev/null";
    auto w0 = generate_weights::distance_2d_weights_file(text ? rf + "weights_e_to_e.txt" : sink, le, le, c.weight_generation_params_e_to_e);
    auto w1 = generate_weights::distance_2d_weights_file(text ? rf + "weights_e_to_i.txt" : sink, le, li, c.weight_generation_params_e_to_i);
    auto w2 = generate_weights::distance_2d_weights_file(text ? rf + "weights_i_to_e.txt" : sink, li, le, c.weight_generation_params_i_to_e);
    auto w3 = generate_weights::distance_2d_weights_file(text ? rf + "weights_i_to_i.txt" : sink, li, li, c.weight_generation_params_i_to_i);
    write_matrix(out + "/w0.bin", w0);
    write_matrix(out + "/w1.bin", w1);
    write_matrix(out + "/w2.bin", w2);
    write_matrix(out + "/w3.bin", w3);
    generate_stimulus_schedules::StimulusGenerationParams sp{};
    sp.max_timestep = c.simulation_steps;
    sp.selected_input_ids = c.stimulus_selected_input_ids;
    sp.selected_input_post_group_offsets = c.stimulus_selected_input_post_group_offsets;
    sp.start_clock_steps = c.stimulus_start_clock_steps;
    sp.end_clock_steps = c.stimulus_end_clock_steps;
    sp.stimulus_input_values = c.stimulus_input_values;
    auto st = generate_stimulus_schedules::partial_group_multiple_square_waves(rf + "stimulus_schedules", sp);
    write_bin(out + "/stim_counts.bin", st.timestep_event_counts);
    write_bin(out + "/stim_ids.bin", st.neuron_ids);
    write_bin(out + "/stim_offsets.bin", st.post_group_offsets);
    write_bin(out + "/stim_amps.bin", st.amplitude_changes);
    return 0;
}
