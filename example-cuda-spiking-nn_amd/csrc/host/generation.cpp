#include "generation.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <random>
#include <stdexcept>
#include <unordered_set>

namespace bn {

std::vector<Location> generate_2d_grid_locations(const std::string& output_filepath, uint32_t num_neurons,
                                                 float spacing, uint32_t rows, uint32_t cols) {
    std::vector<Location> locations;
    if (num_neurons == 0) return locations;
    if (rows == 0 || cols == 0) throw std::invalid_argument("rows and cols must be > 0");
    locations.resize(num_neurons);
    if (rows < (num_neurons + cols - 1) / cols)
        throw std::invalid_argument("grid too small: need at least num_neurons cells");
    std::ofstream output;
    if (!output_filepath.empty()) {
        output.open(output_filepath, std::ios::trunc);
        if (!output.is_open() || !output.good())
            throw std::runtime_error("failed to open output file: " + output_filepath);
        output << "neuron_id, locationdim_0, locationdim_1\n";
    }
    for (uint32_t id = 0; id < num_neurons; ++id) {
        const float x = static_cast<float>(id % cols) * spacing;
        const float y = static_cast<float>(id / cols) * spacing;
        locations[id] = {x, y};
        if (output.is_open()) output << id << ", " << x << ", " << y << "\n";
    }
    if (output.is_open() && !output.good())
        throw std::runtime_error("failed while writing output file: " + output_filepath);
    return locations;
}

float euclidean_distance_pbc(float i, float j, float ip, float jp, float Lx, float Ly) {
    float dx = std::fabs(i - ip);
    float dy = std::fabs(j - jp);
    dx = std::min(dx, Lx - dx);
    dy = std::min(dy, Ly - dy);
    // Sum of squares with one rounding for the first product: the reference is built with
    // -O3 -march=native (Makefile:15), where GCC contracts dx*dx + dy*dy into fma(dx, dx, dy*dy).
    return std::sqrt(std::fma(dx, dx, dy * dy));
}

namespace {
// Fast text writer for "i, j, w" lines with the iostream default float format (%g, 6 digits).
struct LineWriter {
    FILE* f = nullptr;
    std::vector<char> buf;
    size_t n = 0;
    explicit LineWriter(const std::string& path) {
        f = std::fopen(path.c_str(), "w");
        buf.resize(1 << 22);
    }
    ~LineWriter() { close(); }
    bool ok() const { return f != nullptr; }
    void put(const char* s, size_t len) {
        if (n + len > buf.size()) flush();
        std::memcpy(buf.data() + n, s, len);
        n += len;
    }
    void line(uint32_t a, uint32_t b, float w) {
        char tmp[96];
        const int len = std::snprintf(tmp, sizeof tmp, "%u, %u, %g\n", a, b, (double)w);
        put(tmp, (size_t)len);
    }
    void flush() {
        if (f && n) std::fwrite(buf.data(), 1, n, f);
        n = 0;
    }
    void close() {
        if (f) {
            flush();
            std::fclose(f);
            f = nullptr;
        }
    }
};
}  // namespace

// ---------------------------------------------------------------------------------------
// Binary connectivity cache (no counterpart in the reference, which regenerates every matrix on
// every run, generate_weights.cpp:39-100).  A matrix is fixed by its locations, its generation
// parameters, the seed and its index; with a cache directory set (set_connectivity_cache_dir,
// or the BN_CONNECTIVITY_CACHE environment variable), each generated matrix is stored as
// "w_<key>.bnw" (key: FNV-1a 64, over 8-byte words, of those inputs) and read back by later runs
// instead of drawn again.  The file holds a header (magic, format, key, shape, the same hash of
// the entries) and the FP32 entries; a file whose header or checksum does not match is regenerated and rewritten.
// ---------------------------------------------------------------------------------------
namespace {
constexpr char kCacheMagic[8] = {'B', 'N', 'W', 'E', 'I', 'G', 'H', 'T'};
constexpr uint32_t kCacheFormat = 2;

std::string& cache_dir_ref() {
    static std::string dir = [] {
        const char* e = std::getenv("BN_CONNECTIVITY_CACHE");
        return std::string(e ? e : "");
    }();
    return dir;
}

struct Fnv {   // FNV-1a 64 over 8-byte words (then the tail bytes)
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        size_t k = 0;
        for (; k + 8 <= n; k += 8) {
            uint64_t word;
            std::memcpy(&word, b + k, 8);
            h = (h ^ word) * 1099511628211ull;
        }
        for (; k < n; ++k) h = (h ^ b[k]) * 1099511628211ull;
    }
    template <typename T>
    void add(const T& v) { add(&v, sizeof v); }
};

struct CacheHeader {
    char magic[8];
    uint32_t format, n_pre, n_post, pad;
    uint64_t key, sum;
};

uint64_t matrix_key(const std::vector<Location>& pre, const std::vector<Location>& post,
                    const WeightGenerationParams& p, int rng_seed, uint32_t matrix_idx) {
    Fnv f;
    f.add(kCacheFormat);
    const uint32_t np = (uint32_t)pre.size(), nq = (uint32_t)post.size();
    f.add(np);
    f.add(nq);
    f.add(pre.data(), pre.size() * sizeof(Location));
    f.add(post.data(), post.size() * sizeof(Location));
    for (float x : {p.x_location_length, p.y_location_length, p.max_distance, p.spread_scale, p.max_weight_value,
                    p.probability_multiplier})
        f.add(x);
    for (bool x : {p.weight_value_normal_dropoff, p.probabilistic_normal_not_uniform, p.no_diagonal_connections,
                   p.weight_value_uniform})
        f.add((uint8_t)x);
    f.add((int32_t)rng_seed);
    f.add(matrix_idx);
    return f.h;
}

bool cache_load(const std::string& path, uint64_t key, uint32_t n_pre, uint32_t n_post, std::vector<float>& w) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    CacheHeader h{};
    bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, kCacheMagic, 8) == 0 &&
              h.format == kCacheFormat && h.key == key && h.n_pre == n_pre && h.n_post == n_post;
    if (ok) {
        w.assign((size_t)n_pre * n_post, 0.0f);
        ok = std::fread(w.data(), sizeof(float), w.size(), f) == w.size();
        if (ok) {
            Fnv s;
            s.add(w.data(), w.size() * sizeof(float));
            ok = s.h == h.sum;
        }
    }
    std::fclose(f);
    return ok;
}

void cache_store(const std::string& path, uint64_t key, uint32_t n_pre, uint32_t n_post, const std::vector<float>& w) {
    CacheHeader h{};
    std::memcpy(h.magic, kCacheMagic, 8);
    h.format = kCacheFormat;
    h.n_pre = n_pre;
    h.n_post = n_post;
    h.key = key;
    Fnv s;
    s.add(w.data(), w.size() * sizeof(float));
    h.sum = s.h;
    // Written under a temporary name and renamed, so a reader never sees a partial file.
    const std::string tmp = path + ".tmp" + std::to_string((uintptr_t)&h);
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;   // an unwritable cache directory only costs the next run a regeneration
    const bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(w.data(), sizeof(float), w.size(), f) == w.size();
    const bool closed = std::fclose(f) == 0;
    std::error_code ec;
    if (ok && closed) std::filesystem::rename(tmp, path, ec);
    if (!(ok && closed) || ec) std::filesystem::remove(tmp, ec);
}
}  // namespace

void set_connectivity_cache_dir(const std::string& dir) { cache_dir_ref() = dir; }
const std::string& connectivity_cache_dir() { return cache_dir_ref(); }

std::vector<float> distance_2d_weights(const std::string& filename, const std::vector<Location>& pre,
                                       const std::vector<Location>& post, const WeightGenerationParams& params,
                                       int rng_seed, uint32_t matrix_idx) {
    const uint32_t n_pre = (uint32_t)pre.size();
    const uint32_t n_post = (uint32_t)post.size();
    if (n_pre == 0 || n_post == 0) throw std::invalid_argument("Neuron groups must not be empty.");
    std::unique_ptr<LineWriter> out;
    if (!filename.empty()) {
        out.reset(new LineWriter(filename));
        if (!out->ok()) throw std::runtime_error("Failed to create weights file: " + filename);
        const char* hdr = "pre_neuron_id, post_neuron_id, weight\n";
        out->put(hdr, std::strlen(hdr));
    }
    std::vector<float> w;
    std::string cache_path;
    uint64_t key = 0;
    if (!cache_dir_ref().empty()) {
        key = matrix_key(pre, post, params, rng_seed, matrix_idx);
        char name[40];
        std::snprintf(name, sizeof name, "w_%016llx.bnw", (unsigned long long)key);
        cache_path = (std::filesystem::path(cache_dir_ref()) / name).string();
        if (cache_load(cache_path, key, n_pre, n_post, w)) {
            if (out)
                for (uint32_t i = 0; i < n_pre; ++i)
                    for (uint32_t j = 0; j < n_post; ++j) out->line(i, j, w[(size_t)i * n_post + j]);
            return w;
        }
    }
    w.assign((size_t)n_pre * n_post, 0.0f);
    // The reference re-seeds one mt19937 per matrix (generate_weights.cpp:36-37).
    std::mt19937 rng(rng_seed);
    std::uniform_real_distribution<float> unif(0.0f, 1.0f);
    // Extension stream for uniform weight values, separate so the mask stream is unchanged.
    std::mt19937 rng_value((uint32_t)rng_seed ^ (0x5bd1e995u * (matrix_idx + 1u)));
    std::uniform_real_distribution<float> unif_value(0.0f, 1.0f);
    for (uint32_t i = 0; i < n_pre; ++i) {
        for (uint32_t j = 0; j < n_post; ++j) {
            const float distance = euclidean_distance_pbc(pre[i][0], pre[i][1], post[j][0], post[j][1],
                                                          params.x_location_length, params.y_location_length);
            float weight = 0.0f;
            bool draw = true;
            if (params.no_diagonal_connections && i == j) draw = false;
            else if (std::abs(distance) > params.max_distance) draw = false;
            if (draw) {
                const float u = unif(rng);
                float p = params.probability_multiplier;
                if (params.probabilistic_normal_not_uniform)
                    p *= std::exp(-1.0f * (distance * distance) / (params.spread_scale));
                if (u < p) {
                    weight = params.max_weight_value;
                    if (params.weight_value_normal_dropoff)
                        weight *= std::exp(-1.0f * (distance * distance) / (params.spread_scale));
                    if (params.weight_value_uniform) weight = params.max_weight_value * unif_value(rng_value);
                    if (weight < 0)
                        throw std::runtime_error(
                            "generate_weights::distance_2d_weights_file: negative weight value calculated (pre=" +
                            std::to_string(i) + ", post=" + std::to_string(j) + ")");
                }
            }
            w[(size_t)i * n_post + j] = weight;
            if (out) out->line(i, j, weight);
        }
    }
    if (!cache_path.empty()) {
        std::error_code ec;
        std::filesystem::create_directories(cache_dir_ref(), ec);
        cache_store(cache_path, key, n_pre, n_post, w);
    }
    return w;
}

namespace {
bool write_u32_lines(const std::filesystem::path& p, const std::vector<uint32_t>& v) {
    std::ofstream out(p);
    if (!out.is_open()) return false;
    for (uint32_t x : v) out << x << '\n';
    return true;
}
bool write_f32_lines(const std::filesystem::path& p, const std::vector<float>& v) {
    std::ofstream out(p);
    if (!out.is_open()) return false;
    for (float x : v) out << x << '\n';
    return true;
}
}  // namespace

StimulusGenerationResult partial_group_multiple_square_waves(const std::string& folder,
                                                             StimulusGenerationParams& params) {
    StimulusGenerationResult result;
    if (params.selected_input_ids.size() != params.selected_input_post_group_offsets.size()) {
        std::cerr << "Selected neuron vectors must have the same size: selected_input_ids="
                  << params.selected_input_ids.size()
                  << ", selected_input_post_group_offsets=" << params.selected_input_post_group_offsets.size() << "\n";
        return result;
    }
    if (params.start_clock_steps.size() != params.end_clock_steps.size()) {
        std::cerr << "Multi square wave parameter vectors must have the same size: start_clock_steps="
                  << params.start_clock_steps.size() << ", end_clock_steps=" << params.end_clock_steps.size() << "\n";
        return result;
    }
    if (params.stimulus_input_values.size() != params.start_clock_steps.size()) {
        if (params.stimulus_input_values.size() == 1) {
            params.stimulus_input_values.resize(params.start_clock_steps.size(), params.stimulus_input_values[0]);
        } else {
            std::cerr << "Multi square wave parameter vectors must have the same size: stimulus_input_values="
                      << params.stimulus_input_values.size() << ", start_clock_steps=" << params.start_clock_steps.size()
                      << "\n";
            return result;
        }
    }
    for (size_t i = 0; i < params.start_clock_steps.size(); ++i) {
        if (params.start_clock_steps[i] > params.end_clock_steps[i]) {
            std::cerr << "Square wave start has to be before end: start_clock_step (" << params.start_clock_steps[i]
                      << ") > end_clock_step (" << params.end_clock_steps[i] << ")\n";
            return result;
        }
    }
    if (params.max_timestep == 0) {
        std::cerr << "params.max_timestep must be > 0\n";
        return result;
    }
    const uint32_t T = params.max_timestep;
    std::filesystem::path dir(folder);
    if (!folder.empty()) {
        std::error_code ec;
        std::filesystem::create_directories(dir, ec);
        if (ec) {
            std::cerr << "Failed to create stimulus schedules folder: " << dir << "\n";
            return result;
        }
    }
    std::vector<std::pair<uint32_t, uint32_t>> selected;  // (post_group_offset, neuron id), first occurrence
    std::unordered_set<uint64_t> seen;
    for (size_t i = 0; i < params.selected_input_ids.size(); ++i) {
        const uint32_t nid = params.selected_input_ids[i];
        const uint32_t off = params.selected_input_post_group_offsets[i];
        if (seen.insert(((uint64_t)off << 32) | nid).second) selected.emplace_back(off, nid);
    }
    if (selected.empty()) {
        result.timestep_event_counts.assign(T, 0u);
        if (!folder.empty()) {
            write_u32_lines(dir / "timestep_event_counts.txt", result.timestep_event_counts);
            write_u32_lines(dir / "stimulus_event_neuron_ids.txt", result.neuron_ids);
            write_u32_lines(dir / "stimulus_event_post_group_offsets.txt", result.post_group_offsets);
            write_f32_lines(dir / "stimulus_event_delta_amplitudes.txt", result.amplitude_changes);
        }
        return result;
    }
    struct Ev {
        uint32_t t, off, neuron;
        float delta;
    };
    std::vector<Ev> events;
    for (const auto& sel : selected) {
        for (size_t w = 0; w < params.start_clock_steps.size(); ++w) {
            const uint32_t start = params.start_clock_steps[w];
            const uint32_t off_t = params.end_clock_steps[w] + 1;
            const float amp = params.stimulus_input_values[w];
            if (start < T) events.push_back(Ev{start, sel.first, sel.second, +amp});
            if (off_t < T) events.push_back(Ev{off_t, sel.first, sel.second, -amp});
        }
    }
    std::sort(events.begin(), events.end(), [](const Ev& a, const Ev& b) {
        if (a.t != b.t) return a.t < b.t;
        if (a.off != b.off) return a.off < b.off;
        return a.neuron < b.neuron;
    });
    for (size_t k = 1; k < events.size(); ++k) {
        if (events[k].t == events[k - 1].t && events[k].off == events[k - 1].off &&
            events[k].neuron == events[k - 1].neuron)
            throw std::runtime_error("Error: duplicate events detected for time step " + std::to_string(events[k].t) +
                                     ", post_group_offset " + std::to_string(events[k].off) + ", neuron_id " +
                                     std::to_string(events[k].neuron));
    }
    result.timestep_event_counts.assign(T, 0u);
    for (const Ev& e : events) result.timestep_event_counts[e.t] += 1u;
    for (const Ev& e : events) {
        result.neuron_ids.push_back(e.neuron);
        result.post_group_offsets.push_back(e.off);
        result.amplitude_changes.push_back(e.delta);
    }
    if (!folder.empty()) {
        const bool ok1 = write_u32_lines(dir / "timestep_event_counts.txt", result.timestep_event_counts);
        const bool ok2 = write_u32_lines(dir / "stimulus_event_neuron_ids.txt", result.neuron_ids);
        const bool ok3 = write_u32_lines(dir / "stimulus_event_post_group_offsets.txt", result.post_group_offsets);
        const bool ok4 = write_f32_lines(dir / "stimulus_event_delta_amplitudes.txt", result.amplitude_changes);
        if (!ok1) std::cerr << "Failed to write timestep_event_counts.txt\n";
        if (!ok2) std::cerr << "Failed to write stimulus_event_neuron_ids.txt\n";
        if (!ok3) std::cerr << "Failed to write stimulus_event_post_group_offsets.txt\n";
        if (!ok4) std::cerr << "Failed to write stimulus_event_delta_amplitudes.txt\n";
        std::ofstream out(dir / "stimulus_events_compact.txt");
        if (out.is_open()) {
            out << "# idx time_step post_group_offset neuron_id delta\n";
            uint32_t idx = 0;
            for (const Ev& e : events) out << idx++ << " " << e.t << " " << e.off << " " << e.neuron << " " << e.delta << "\n";
        }
    }
    return result;
}

}  // namespace bn
