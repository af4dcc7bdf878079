// The object behind the C ABI's bn_sim handle (one network: state + component system).
#pragma once

#include <memory>

#include "config.h"
#include "engine.h"

struct bn_sim {
    bn::Config cfg;
    uint32_t flags = 0;
    std::unique_ptr<bn::GpuSimulationState> state;
    std::unique_ptr<bn::System> system;
    float* d_tmp_ge = nullptr;
    float* d_tmp_gi = nullptr;
    ~bn_sim() {
        system.reset();
        if (d_tmp_ge) (void)hipFree(d_tmp_ge);
        if (d_tmp_gi) (void)hipFree(d_tmp_gi);
        state.reset();
    }
};
