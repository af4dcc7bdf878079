// Network inputs shared by the plain and the batched engines (main_buryn.cpp:151-272).
#include "network.h"

namespace bn {

NeuronGroupParams group_params(const Config& c, bool inhibitory) {
    NeuronGroupParams p;
    p.timestep = c.timestep;
    p.g_leak = c.g_leak;
    p.membrane_capacitance = c.membrane_capacitance;
    p.depolarization_slope_factor = c.depolarization_slope_factor;
    p.v_effective_spike_threshold = c.v_effective_spike_threshold;
    p.v_excitatory_reversal = c.v_excitatory_reversal;
    p.v_inhibitory_reversal = c.v_inhibitory_reversal;
    p.v_threshold = c.v_threshold;
    p.v_reset = c.v_reset;
    p.v_rest = c.v_rest;
    p.refractory_steps = c.refractory_steps;
    p.is_inhibitory = inhibitory;
    p.offset = inhibitory ? c.num_excitatory_neurons : 0;
    p.num_neurons = inhibitory ? c.num_inhibitory_neurons : c.num_excitatory_neurons;
#define BN_G(field) p.field = inhibitory ? c.group_i_##field : c.group_e_##field
    BN_G(mean_background_excitatory_spikes);
    BN_G(mean_background_inhibitory_spikes);
    BN_G(background_excitatory_postsynaptic_r_trace_decay);
    BN_G(background_excitatory_postsynaptic_d_trace_decay);
    BN_G(background_inhibitory_postsynaptic_r_trace_decay);
    BN_G(background_inhibitory_postsynaptic_d_trace_decay);
    BN_G(background_excitatory_scaling_factor);
    BN_G(background_inhibitory_scaling_factor);
#undef BN_G
    return p;
}

Network generate_inputs(const Config& c, bool write_files) {
    Network net;
    net.loc_e = generate_2d_grid_locations(write_files ? c.excitatory_locations_file : "", c.num_excitatory_neurons,
                                           c.excitatory_location_spacing, c.excitatory_side_points,
                                           c.excitatory_side_points);
    net.loc_i = generate_2d_grid_locations(write_files ? c.inhibitory_locations_file : "", c.num_inhibitory_neurons,
                                           c.inhibitory_location_spacing, c.inhibitory_side_points,
                                           c.inhibitory_side_points);
    const WeightGenerationParams* gp[4] = {&c.weight_generation_params_e_to_e, &c.weight_generation_params_e_to_i,
                                           &c.weight_generation_params_i_to_e, &c.weight_generation_params_i_to_i};
    const std::string* files[4] = {&c.weights_e_to_e_filename, &c.weights_e_to_i_filename, &c.weights_i_to_e_filename,
                                   &c.weights_i_to_i_filename};
    const std::vector<Location>* pre[4] = {&net.loc_e, &net.loc_e, &net.loc_i, &net.loc_i};
    const std::vector<Location>* post[4] = {&net.loc_e, &net.loc_i, &net.loc_e, &net.loc_i};
    for (uint32_t k = 0; k < 4; ++k)
        net.w[k] = distance_2d_weights(write_files ? *files[k] : "", *pre[k], *post[k], *gp[k], c.rng_seed, k);
    return net;
}

}  // namespace bn
