// Writes Philox4x32-10 outputs of rocRAND's engine (an independent open-source implementation
// of the published method) for counters laid out as the engine draws its background noise:
// counter = (step, 0, global neuron id, instance), key = (seed low, seed high).
// In rocRAND terms: subsequence = neuron | instance << 32, offset = 4 * step.
// Build and run: hipcc -O1 tests/golden/rocrand_philox_vectors.cpp -o /tmp/rp && /tmp/rp > tests/golden/philox_rocrand.json
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

int main() {
    const unsigned long long seeds[] = {0ull, 1ull, 751ull, 0xFFFFFFFFFFFFFFF9ull /* (uint64)(int)-7 */, 0x123456789abcdefull};
    const unsigned steps[] = {0u, 1u, 2u, 1000u, 10000u, 4294967295u};
    const unsigned neurons[] = {0u, 1u, 2499u, 3124u, 10124u};
    const unsigned instances[] = {0u, 63u};
    std::printf("[\n");
    bool first = true;
    for (unsigned long long seed : seeds)
        for (unsigned step : steps)
            for (unsigned n : neurons)
                for (unsigned inst : instances) {
                    const unsigned long long subseq = (unsigned long long)n | ((unsigned long long)inst << 32);
                    rocrand_device::philox4x32_10_engine e(seed, subseq, 4ull * step);
                    const uint4 r = e.next4();
                    std::printf("%s {\"seed\": %llu, \"step\": %u, \"neuron\": %u, \"instance\": %u, \"out\": [%u, %u, %u, %u]}",
                                first ? "" : ",\n", seed, step, n, inst, r.x, r.y, r.z, r.w);
                    first = false;
                }
    std::printf("\n]\n");
    return 0;
}
