#include "config.h"

#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "json.h"

namespace bn {
namespace {

using json::Value;

const char* kind_name(const Value& v) {
    switch (v.kind) {
        case Value::Null: return "null";
        case Value::Bool: return "boolean";
        case Value::Integer:
        case Value::Unsigned:
        case Value::Float: return "number";
        case Value::String: return "string";
        case Value::Array: return "array";
        case Value::Object: return "object";
    }
    return "?";
}

// Arithmetic conversion of a JSON value as a JSON library's get<T>() does: numbers are cast,
// booleans become 0/1, anything else is a type error.
template <typename T>
T to_arith(const Value& v, const std::string& key) {
    switch (v.kind) {
        case Value::Integer: return static_cast<T>(v.i);
        case Value::Unsigned: return static_cast<T>(v.u);
        case Value::Float: return static_cast<T>(v.f);
        case Value::Bool: return static_cast<T>(v.b);
        default:
            throw std::runtime_error("type must be number, but is " + std::string(kind_name(v)) +
                                     " (key '" + key + "')");
    }
}

template <typename T>
T value_or(const Value& j, const std::string& key, T def) {
    const Value* v = j.find(key);
    return v ? to_arith<T>(*v, key) : def;
}

bool bool_or(const Value& j, const std::string& key, bool def) {
    const Value* v = j.find(key);
    if (!v) return def;
    if (v->kind != Value::Bool)
        throw std::runtime_error("type must be boolean, but is " + std::string(kind_name(*v)) +
                                 " (key '" + key + "')");
    return v->b;
}

std::string string_or(const Value& j, const std::string& key, const std::string& def) {
    const Value* v = j.find(key);
    if (!v) return def;
    if (v->kind != Value::String)
        throw std::runtime_error("type must be string, but is " + std::string(kind_name(*v)) +
                                 " (key '" + key + "')");
    return v->s;
}

// config.cpp:210-288 — integer arrays: only integers, none negative.
void read_u32_array(const Value& j, const std::string& key, std::vector<uint32_t>& out) {
    const Value* a = j.find(key);
    if (!a) return;
    if (a->kind != Value::Array) throw std::runtime_error(key + " must be an array");
    for (const Value& e : a->arr) {
        if (!e.is_integer()) throw std::runtime_error(key + " must contain only integers");
        if (e.kind == Value::Integer && e.i < 0)
            throw std::runtime_error(key + " must not contain negative values");
    }
    out.clear();
    for (const Value& e : a->arr) out.push_back(static_cast<uint32_t>(e.kind == Value::Integer ? (uint64_t)e.i : e.u));
}

// config.cpp:13-44
void apply_weight_generation_overrides(WeightGenerationParams& w, const Value& j, const std::string& prefix) {
    w.max_weight_value = value_or<float>(j, prefix + "_max_weight_value", w.max_weight_value);
    w.max_distance = value_or<float>(j, prefix + "_max_distance", w.max_distance);
    w.spread_scale = value_or<float>(j, prefix + "_spread_scale", w.spread_scale);
    w.weight_value_normal_dropoff = bool_or(j, prefix + "_weight_value_normal_dropoff", w.weight_value_normal_dropoff);
    w.probabilistic_normal_not_uniform =
        bool_or(j, prefix + "_probabilistic_normal_not_uniform", w.probabilistic_normal_not_uniform);
    w.probability_multiplier = value_or<float>(j, prefix + "_probability_multiplier", w.probability_multiplier);
    w.no_diagonal_connections = bool_or(j, prefix + "_no_diagonal_connections", w.no_diagonal_connections);
    w.x_location_length = value_or<float>(j, prefix + "_x_location_length", w.x_location_length);
    w.y_location_length = value_or<float>(j, prefix + "_y_location_length", w.y_location_length);
    w.weight_value_uniform = bool_or(j, prefix + "_weight_value_uniform", w.weight_value_uniform);
}

// config.cpp:46-66
std::string prefix_output_path(const std::string& output_directory, const std::string& path_value) {
    if (path_value.empty()) return path_value;
    std::filesystem::path path(path_value);
    if (path.is_absolute()) return path_value;
    std::filesystem::path output_dir(output_directory);
    if (!output_directory.empty()) {
        auto pi = path.begin();
        auto oi = output_dir.begin();
        if (pi != path.end() && oi != output_dir.end() && *pi == *oi) return path_value;
    }
    return (output_dir / path).lexically_normal().string();
}

Config from_value(const Value& j) {
    if (j.kind != Value::Object) throw std::runtime_error("Config JSON must be an object at the top level");
    Config c;
    c.simulation_steps = value_or<uint32_t>(j, "simulation_steps", c.simulation_steps);
    c.output_directory = string_or(j, "output_directory", c.output_directory);
    c.description = string_or(j, "description", c.description);
    c.rng_seed = value_or<int>(j, "rng_seed", c.rng_seed);
    c.timestep = value_or<float>(j, "timestep", c.timestep);

    c.num_excitatory_neurons = value_or<uint32_t>(j, "num_excitatory_neurons", c.num_excitatory_neurons);
    c.num_inhibitory_neurons = value_or<uint32_t>(j, "num_inhibitory_neurons", c.num_inhibitory_neurons);
    c.g_leak = value_or<float>(j, "g_leak", c.g_leak);
    c.membrane_capacitance = value_or<float>(j, "membrane_capacitance", c.membrane_capacitance);
    c.depolarization_slope_factor = value_or<float>(j, "depolarization_slope_factor", c.depolarization_slope_factor);
    c.v_effective_spike_threshold = value_or<float>(j, "v_effective_spike_threshold", c.v_effective_spike_threshold);
    c.v_excitatory_reversal = value_or<float>(j, "v_excitatory_reversal", c.v_excitatory_reversal);
    c.v_inhibitory_reversal = value_or<float>(j, "v_inhibitory_reversal", c.v_inhibitory_reversal);
    c.v_threshold = value_or<float>(j, "v_threshold", c.v_threshold);
    c.v_reset = value_or<float>(j, "v_reset", c.v_reset);
    c.v_rest = value_or<float>(j, "v_rest", c.v_rest);
    c.refractory_steps = value_or<uint32_t>(j, "refractory_steps", c.refractory_steps);

#define BN_GROUP_KEYS(g)                                                                                   \
    c.group_##g##_mean_background_excitatory_spikes =                                                       \
        value_or<float>(j, "group_" #g "_mean_background_excitatory_spikes", c.group_##g##_mean_background_excitatory_spikes); \
    c.group_##g##_mean_background_inhibitory_spikes =                                                       \
        value_or<float>(j, "group_" #g "_mean_background_inhibitory_spikes", c.group_##g##_mean_background_inhibitory_spikes); \
    c.group_##g##_background_excitatory_postsynaptic_r_trace_decay = value_or<float>(                     \
        j, "group_" #g "_background_excitatory_postsynaptic_r_trace_decay", c.group_##g##_background_excitatory_postsynaptic_r_trace_decay); \
    c.group_##g##_background_excitatory_postsynaptic_d_trace_decay = value_or<float>(                     \
        j, "group_" #g "_background_excitatory_postsynaptic_d_trace_decay", c.group_##g##_background_excitatory_postsynaptic_d_trace_decay); \
    c.group_##g##_background_inhibitory_postsynaptic_r_trace_decay = value_or<float>(                     \
        j, "group_" #g "_background_inhibitory_postsynaptic_r_trace_decay", c.group_##g##_background_inhibitory_postsynaptic_r_trace_decay); \
    c.group_##g##_background_inhibitory_postsynaptic_d_trace_decay = value_or<float>(                     \
        j, "group_" #g "_background_inhibitory_postsynaptic_d_trace_decay", c.group_##g##_background_inhibitory_postsynaptic_d_trace_decay); \
    c.group_##g##_background_excitatory_scaling_factor =                                                    \
        value_or<float>(j, "group_" #g "_background_excitatory_scaling_factor", c.group_##g##_background_excitatory_scaling_factor); \
    c.group_##g##_background_inhibitory_scaling_factor =                                                    \
        value_or<float>(j, "group_" #g "_background_inhibitory_scaling_factor", c.group_##g##_background_inhibitory_scaling_factor);
    BN_GROUP_KEYS(e)
    BN_GROUP_KEYS(i)
#undef BN_GROUP_KEYS

    c.excitatory_locations_file = string_or(j, "excitatory_locations_file", c.excitatory_locations_file);
    c.inhibitory_locations_file = string_or(j, "inhibitory_locations_file", c.inhibitory_locations_file);
    c.excitatory_location_spacing = value_or<float>(j, "excitatory_location_spacing", c.excitatory_location_spacing);
    c.inhibitory_location_spacing = value_or<float>(j, "inhibitory_location_spacing", c.inhibitory_location_spacing);
    c.excitatory_side_points = value_or<uint32_t>(j, "excitatory_side_points", c.excitatory_side_points);
    c.inhibitory_side_points = value_or<uint32_t>(j, "inhibitory_side_points", c.inhibitory_side_points);
    c.side_length = value_or<float>(j, "side_length", c.side_length);

    c.e_max_weight_value = value_or<float>(j, "e_max_weight_value", c.e_max_weight_value);
    c.i_max_weight_value = value_or<float>(j, "i_max_weight_value", c.i_max_weight_value);
    c.max_connection_distance = value_or<float>(j, "max_connection_distance", c.max_connection_distance);
    c.connection_spread_scale = value_or<float>(j, "connection_spread_scale", c.connection_spread_scale);
    c.connection_probability_multiplier =
        value_or<float>(j, "connection_probability_multiplier", c.connection_probability_multiplier);
    WeightGenerationParams* all[4] = {&c.weight_generation_params_e_to_e, &c.weight_generation_params_e_to_i,
                                      &c.weight_generation_params_i_to_e, &c.weight_generation_params_i_to_i};
    for (int m = 0; m < 4; ++m) {
        all[m]->max_weight_value = m < 2 ? c.e_max_weight_value : c.i_max_weight_value;
        all[m]->max_distance = c.max_connection_distance;
        all[m]->spread_scale = c.connection_spread_scale;
        all[m]->probability_multiplier = c.connection_probability_multiplier;
        all[m]->x_location_length = c.side_length;
        all[m]->y_location_length = c.side_length;
    }
    apply_weight_generation_overrides(c.weight_generation_params_e_to_e, j, "e_to_e");
    apply_weight_generation_overrides(c.weight_generation_params_e_to_i, j, "e_to_i");
    apply_weight_generation_overrides(c.weight_generation_params_i_to_e, j, "i_to_e");
    apply_weight_generation_overrides(c.weight_generation_params_i_to_i, j, "i_to_i");

    c.weights_e_to_i_filename = string_or(j, "weights_e_to_i_filename", string_or(j, "weights_e_to_i", c.weights_e_to_i_filename));
    c.weights_i_to_e_filename = string_or(j, "weights_i_to_e_filename", string_or(j, "weights_i_to_e", c.weights_i_to_e_filename));
    c.weights_e_to_e_filename = string_or(j, "weights_e_to_e_filename", string_or(j, "weights_e_to_e", c.weights_e_to_e_filename));
    c.weights_i_to_i_filename = string_or(j, "weights_i_to_i_filename", string_or(j, "weights_i_to_i", c.weights_i_to_i_filename));
    c.tau_decay_e = value_or<float>(j, "tau_decay_e", c.tau_decay_e);
    c.tau_rise_e = value_or<float>(j, "tau_rise_e", c.tau_rise_e);
    c.tau_decay_i = value_or<float>(j, "tau_decay_i", c.tau_decay_i);
    c.tau_rise_i = value_or<float>(j, "tau_rise_i", c.tau_rise_i);
    c.postsynaptic_kernel_normalization_value =
        value_or<float>(j, "postsynaptic_kernel_normalization_value", c.postsynaptic_kernel_normalization_value);

    c.e_e_plasticity_enabled = bool_or(j, "e_e_plasticity_enabled", c.e_e_plasticity_enabled);
    c.e_e_plasticity_ltp_A = value_or<float>(j, "e_e_plasticity_ltp_A", c.e_e_plasticity_ltp_A);
    c.e_e_plasticity_ltd_A = value_or<float>(j, "e_e_plasticity_ltd_A", c.e_e_plasticity_ltd_A);
    c.e_e_plasticity_ltp_B = value_or<float>(j, "e_e_plasticity_ltp_B", c.e_e_plasticity_ltp_B);
    c.e_e_plasticity_ltd_B = value_or<float>(j, "e_e_plasticity_ltd_B", c.e_e_plasticity_ltd_B);
    c.e_e_plasticity_tau_ltp = value_or<float>(j, "e_e_plasticity_tau_ltp", c.e_e_plasticity_tau_ltp);
    c.e_e_plasticity_tau_ltd = value_or<float>(j, "e_e_plasticity_tau_ltd", c.e_e_plasticity_tau_ltd);
    c.e_e_plasticity_triplet_tau_ltp = value_or<float>(j, "e_e_plasticity_triplet_tau_ltp", c.e_e_plasticity_triplet_tau_ltp);
    c.e_e_plasticity_triplet_tau_ltd = value_or<float>(j, "e_e_plasticity_triplet_tau_ltd", c.e_e_plasticity_triplet_tau_ltd);
    c.e_e_plasticity_weight_min = value_or<float>(j, "e_e_plasticity_weight_min", c.e_e_plasticity_weight_min);
    c.e_e_plasticity_weight_max = value_or<float>(j, "e_e_plasticity_weight_max", c.e_e_plasticity_weight_max);
    c.i_e_plasticity_enabled = bool_or(j, "i_e_plasticity_enabled", c.i_e_plasticity_enabled);
    c.i_e_plasticity_ltp_A = value_or<float>(j, "i_e_plasticity_ltp_A", c.i_e_plasticity_ltp_A);
    c.i_e_plasticity_ltd_A = value_or<float>(j, "i_e_plasticity_ltd_A", c.i_e_plasticity_ltd_A);
    c.i_e_plasticity_ltd_B = value_or<float>(j, "i_e_plasticity_ltd_B", c.i_e_plasticity_ltd_B);
    c.i_e_plasticity_tau_ltp = value_or<float>(j, "i_e_plasticity_tau_ltp", c.i_e_plasticity_tau_ltp);
    c.i_e_plasticity_tau_ltd = value_or<float>(j, "i_e_plasticity_tau_ltd", c.i_e_plasticity_tau_ltd);
    c.i_e_plasticity_weight_min = value_or<float>(j, "i_e_plasticity_weight_min", c.i_e_plasticity_weight_min);
    c.i_e_plasticity_weight_max = value_or<float>(j, "i_e_plasticity_weight_max", c.i_e_plasticity_weight_max);

    if (const Value* a = j.find("stimulus_input_values")) {
        if (a->kind != Value::Array) throw std::runtime_error("stimulus_input_values must be an array");
        c.stimulus_input_values.clear();
        for (const Value& e : a->arr) c.stimulus_input_values.push_back(to_arith<float>(e, "stimulus_input_values"));
    }
    c.stimulus_schedules_folder = string_or(j, "stimulus_schedules_folder", c.stimulus_schedules_folder);
    read_u32_array(j, "stimulus_selected_input_ids", c.stimulus_selected_input_ids);
    read_u32_array(j, "stimulus_selected_input_post_group_offsets", c.stimulus_selected_input_post_group_offsets);
    read_u32_array(j, "stimulus_start_clock_steps", c.stimulus_start_clock_steps);
    read_u32_array(j, "stimulus_end_clock_steps", c.stimulus_end_clock_steps);

    c.e_spike_recording_file = string_or(j, "e_spike_recording_file", c.e_spike_recording_file);
    c.i_spike_recording_file = string_or(j, "i_spike_recording_file", c.i_spike_recording_file);
    c.voltage_recording_file = string_or(j, "voltage_recording_file", c.voltage_recording_file);
    read_u32_array(j, "voltage_recording_neuron_ids", c.voltage_recording_neuron_ids);
    c.e_e_weight_recording_file = string_or(j, "e_e_weight_recording_file", c.e_e_weight_recording_file);
    read_u32_array(j, "gemv_row_blocks", c.gemv_row_blocks);
    if (c.gemv_row_blocks.size() != 4) throw std::runtime_error("gemv_row_blocks must hold 4 values (EE, EI, IE, II)");
    c.gemv_nontemporal_loads = value_or<int>(j, "gemv_nontemporal_loads", c.gemv_nontemporal_loads);
    c.background_noise_seed = value_or<long long>(j, "background_noise_seed", c.background_noise_seed);
    c.e_e_plasticity_rule = string_or(j, "e_e_plasticity_rule", c.e_e_plasticity_rule);
    if (c.e_e_plasticity_rule != "triplet" && c.e_e_plasticity_rule != "pair")
        throw std::runtime_error("e_e_plasticity_rule must be \"triplet\" or \"pair\"");

    const std::string& od = c.output_directory;
    c.excitatory_locations_file = prefix_output_path(od, c.excitatory_locations_file);
    c.inhibitory_locations_file = prefix_output_path(od, c.inhibitory_locations_file);
    c.weights_e_to_i_filename = prefix_output_path(od, c.weights_e_to_i_filename);
    c.weights_i_to_e_filename = prefix_output_path(od, c.weights_i_to_e_filename);
    c.weights_e_to_e_filename = prefix_output_path(od, c.weights_e_to_e_filename);
    c.weights_i_to_i_filename = prefix_output_path(od, c.weights_i_to_i_filename);
    c.stimulus_schedules_folder = prefix_output_path(od, c.stimulus_schedules_folder);
    c.e_spike_recording_file = prefix_output_path(od, c.e_spike_recording_file);
    c.i_spike_recording_file = prefix_output_path(od, c.i_spike_recording_file);
    c.voltage_recording_file = prefix_output_path(od, c.voltage_recording_file);
    c.e_e_weight_recording_file = prefix_output_path(od, c.e_e_weight_recording_file);

    if (c.stimulus_selected_input_post_group_offsets.empty())
        c.stimulus_selected_input_post_group_offsets = std::vector<uint32_t>(c.stimulus_selected_input_ids.size(), 0);
    if (c.stimulus_selected_input_post_group_offsets.size() != c.stimulus_selected_input_ids.size())
        throw std::runtime_error("stimulus_selected_input_post_group_offsets size must match stimulus_selected_input_ids size");
    if (c.num_excitatory_neurons == 0 || c.num_inhibitory_neurons == 0)
        throw std::runtime_error("Neuron counts must be > 0");
    if (c.membrane_capacitance <= 0.0f) throw std::runtime_error("membrane_capacitance must be > 0");
    return c;
}

}  // namespace

Config load_config_json_text(const std::string& text) { return from_value(json::parse(text)); }

Config load_config_json(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open config file: " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return load_config_json_text(ss.str());
}

namespace {
struct Dump {
    std::ostringstream o;
    bool first = true;
    void key(const std::string& k) {
        o << (first ? "{" : ",") << "\n  " << json::quote(k) << ": ";
        first = false;
    }
    void f(const std::string& k, float v) { key(k); o << json::number((double)v); }
    void u(const std::string& k, uint64_t v) { key(k); o << v; }
    void i(const std::string& k, int64_t v) { key(k); o << v; }
    void b(const std::string& k, bool v) { key(k); o << (v ? "true" : "false"); }
    void s(const std::string& k, const std::string& v) { key(k); o << json::quote(v); }
    void fa(const std::string& k, const std::vector<float>& v) {
        key(k);
        o << "[";
        for (size_t n = 0; n < v.size(); ++n) o << (n ? ", " : "") << json::number((double)v[n]);
        o << "]";
    }
    void ua(const std::string& k, const std::vector<uint32_t>& v) {
        key(k);
        o << "[";
        for (size_t n = 0; n < v.size(); ++n) o << (n ? ", " : "") << v[n];
        o << "]";
    }
    void w(const std::string& k, const WeightGenerationParams& p) {
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
}  // namespace

std::string config_to_json(const Config& c) {
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
    d.ua("gemv_row_blocks", c.gemv_row_blocks);
    d.i("gemv_nontemporal_loads", c.gemv_nontemporal_loads);
    d.i("background_noise_seed", c.background_noise_seed);
    S(e_e_plasticity_rule)
#undef F
#undef U
#undef S
#undef B
    d.o << "\n}\n";
    return d.o.str();
}

}  // namespace bn
