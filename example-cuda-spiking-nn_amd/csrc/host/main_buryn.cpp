// Driver with the reference's command line (cpp_src/main_buryn.cpp:98-101): argv[1] is the
// config path (default "config.json").  Optional flags after it:
//   --component   step through the component launchers (System::step_the_clock)
//   --lazy        lazy triplet normalisation (W_EE = diag(s) M)
//   --no-graph    launch the fused step eagerly
//   --no-files    do not write the output directory
#include <cstdio>
#include <cstring>
#include <string>

#include "buryn_amd.h"

int main(int argc, char** argv) {
    const char* config = argc > 1 ? argv[1] : "config.json";
    uint32_t flags = BN_SIM_WRITE_FILES;
    for (int k = 2; k < argc; ++k) {
        const std::string a = argv[k];
        if (a == "--component") flags |= BN_SIM_COMPONENT_PATH;
        else if (a == "--lazy") flags |= BN_SIM_LAZY_NORMALIZE;
        else if (a == "--no-graph") flags |= BN_SIM_NO_GRAPH;
        else if (a == "--no-files") flags &= ~(uint32_t)BN_SIM_WRITE_FILES;
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    bn_sim* sim = nullptr;
    if (bn_sim_create(config, flags, &sim) != 0) {
        std::fprintf(stderr, "terminate called after throwing an instance of 'std::runtime_error'\n  what():  %s\n",
                     bn_last_error());
        return 1;
    }
    bn_sim_info info;
    bn_sim_get_info(sim, &info);
    std::printf("Running simulation...\n");
    if (bn_sim_run(sim, info.simulation_steps) != 0 || bn_sim_synchronize(sim) != 0) {
        std::fprintf(stderr, "terminate called after throwing an instance of 'std::runtime_error'\n  what():  %s\n",
                     bn_last_error());
        bn_sim_destroy(sim);
        return 1;
    }
    std::printf("Simulation complete.\n");
    bn_sim_destroy(sim);
    return 0;
}
