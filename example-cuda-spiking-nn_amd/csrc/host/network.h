// Network inputs shared by the plain and the batched engines: per-group parameters from the
// config, and the host-generated locations and dense weights (main_buryn.cpp:151-272).
#pragma once

#include <vector>

#include "config.h"
#include "engine.h"
#include "generation.h"

namespace bn {

NeuronGroupParams group_params(const Config& c, bool inhibitory);

struct Network {
    std::vector<Location> loc_e, loc_i;
    std::vector<float> w[4];  // EE, EI, IE, II, dense row-major [pre][post]
};

Network generate_inputs(const Config& c, bool write_files);

}  // namespace bn
