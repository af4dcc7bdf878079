// Host-side inputs of the step: neuron grid locations, dense [pre][post] connectivity and the
// stimulus event schedule.  Same algorithms, same random-number call sequence and the same
// text formats as the reference generator (cpp_src/generation/*.cpp,
// cpp_src/utils/math_functions.cpp), so connectivity is bit-identical for the same config.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "config.h"

namespace bn {

using Location = std::array<float, 2>;

// generate_location.cpp:5-47.  output_filepath empty = do not write the file.
std::vector<Location> generate_2d_grid_locations(const std::string& output_filepath, uint32_t num_neurons,
                                                 float spacing, uint32_t rows, uint32_t cols);

// math_functions.cpp:8-16
float euclidean_distance_pbc(float i, float j, float ip, float jp, float Lx, float Ly);

// generate_weights.cpp:10-107: dense row-major [pre][post].  `filename` empty = no text file.
// `matrix_idx` (0 EE, 1 EI, 2 IE, 3 II) only seeds the extension's weight-value stream.
// Directory of the binary connectivity cache ("" = off; default: $BN_CONNECTIVITY_CACHE).
void set_connectivity_cache_dir(const std::string& dir);
const std::string& connectivity_cache_dir();

std::vector<float> distance_2d_weights(const std::string& filename, const std::vector<Location>& pre,
                                       const std::vector<Location>& post, const WeightGenerationParams& params,
                                       int rng_seed, uint32_t matrix_idx);

// generate_stimulus_schedules.cpp:20-228
struct StimulusGenerationParams {
    uint32_t max_timestep = 0;
    std::vector<uint32_t> selected_input_ids;
    std::vector<uint32_t> selected_input_post_group_offsets;
    std::vector<uint32_t> start_clock_steps;
    std::vector<uint32_t> end_clock_steps;
    std::vector<float> stimulus_input_values;
};
struct StimulusGenerationResult {
    std::vector<uint32_t> timestep_event_counts;
    std::vector<uint32_t> neuron_ids;
    std::vector<uint32_t> post_group_offsets;
    std::vector<float> amplitude_changes;
};
// folder empty = do not write the schedule files.
StimulusGenerationResult partial_group_multiple_square_waves(const std::string& folder,
                                                             StimulusGenerationParams& params);

}  // namespace bn
