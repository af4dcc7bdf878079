#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <numeric>
#include <sstream>

#include "../bn_errors.h"

namespace bn {

namespace {
uint32_t round_up(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }
uint32_t cdivu(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Rows per GEMV chunk, one value for all four matrices: the largest power of two <= 512 that
// still gives at least 768 tiles (one wave each, ~3 per CU) over the whole weight set.  Long
// row runs per wave stream HBM best (measured on the 410 MB set: 512 rows, 800 tiles ahead of
// 128 rows, 6,300 tiles); small sets keep enough tiles to cover the latency.
uint32_t default_row_block(const uint32_t* pres, const uint32_t* posts) {
    uint32_t rb = 512;
    for (; rb > 16; rb >>= 1) {
        uint64_t tiles = 0;
        for (int k = 0; k < 4; ++k) tiles += (uint64_t)cdivu(posts[k], kGemvTileCols) * cdivu(pres[k], rb);
        if (tiles >= 768) break;
    }
    return rb;
}
}  // namespace

// ======================================================================================
// GpuSimulationState
// ======================================================================================
template <typename T>
T* GpuSimulationState::alloc(size_t n, const char* what) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    BN_HIP_CHECK(hipMalloc(&p, bytes), std::string("hipMalloc ") + what);
    BN_HIP_CHECK(hipMemsetAsync(p, 0, bytes, stream_), std::string("hipMemset ") + what);
    allocations_.push_back(p);
    allocation_bytes_.push_back(bytes);
    allocated_ += bytes;
    return static_cast<T*>(p);
}

// Snapshot layout: "BNSTATE1", uint64 buffer count, the byte size of each buffer, then the
// buffers in allocation order (the same config and flags allocate the same buffers).
static const char kSnapshotMagic[8] = {'B', 'N', 'S', 'T', 'A', 'T', 'E', '1'};

std::vector<uint8_t> GpuSimulationState::snapshot() const {
    size_t total = 16 + 8 * allocations_.size();
    for (size_t b : allocation_bytes_) total += b;
    std::vector<uint8_t> out(total);
    std::memcpy(out.data(), kSnapshotMagic, 8);
    const uint64_t count = allocations_.size();
    std::memcpy(out.data() + 8, &count, 8);
    size_t off = 16;
    for (size_t b : allocation_bytes_) {
        const uint64_t v = b;
        std::memcpy(out.data() + off, &v, 8);
        off += 8;
    }
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "snapshot sync");
    for (size_t k = 0; k < allocations_.size(); ++k) {
        BN_HIP_CHECK(hipMemcpy(out.data() + off, allocations_[k], allocation_bytes_[k], hipMemcpyDeviceToHost),
                     "snapshot download");
        off += allocation_bytes_[k];
    }
    return out;
}

void GpuSimulationState::restore(const std::vector<uint8_t>& in) {
    if (in.size() < 16 || std::memcmp(in.data(), kSnapshotMagic, 8) != 0)
        throw Error(BN_ERR_INVALID_ARGUMENT, "not a state snapshot");
    uint64_t count = 0;
    std::memcpy(&count, in.data() + 8, 8);
    if (count != allocations_.size() || in.size() < 16 + 8 * count)
        throw Error(BN_ERR_INVALID_ARGUMENT, "the snapshot was taken from a different network or flags");
    size_t off = 16 + 8 * count, total = off;
    for (size_t k = 0; k < count; ++k) {
        uint64_t b = 0;
        std::memcpy(&b, in.data() + 16 + 8 * k, 8);
        if (b != allocation_bytes_[k])
            throw Error(BN_ERR_INVALID_ARGUMENT, "the snapshot was taken from a different network or flags");
        total += b;
    }
    if (in.size() != total) throw Error(BN_ERR_INVALID_ARGUMENT, "truncated state snapshot");
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "restore sync");
    for (size_t k = 0; k < count; ++k) {
        BN_HIP_CHECK(hipMemcpy(allocations_[k], in.data() + off, allocation_bytes_[k], hipMemcpyHostToDevice),
                     "snapshot upload");
        off += allocation_bytes_[k];
    }
    ring_snapshot_step_ = -1;
}

GpuSimulationState::GpuSimulationState(uint32_t total_simulation_length, uint32_t num_excitatory_neurons,
                                       uint32_t num_inhibitory_neurons, uint32_t num_stimulus_events,
                                       uint32_t num_plasticities, StateOptions opt)
    : opt_(std::move(opt)),
      total_simulation_length_(total_simulation_length),
      ne_(num_excitatory_neurons),
      ni_(num_inhibitory_neurons),
      n_(num_excitatory_neurons + num_inhibitory_neurons),
      num_stimulus_events_(num_stimulus_events),
      num_plasticities_(num_plasticities) {
    BN_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    const uint32_t N = n_;
    v_ = alloc<float>(N, "membrane_potentials");
    ge_ = alloc<float>(N, "g_excitatory");
    gi_ = alloc<float>(N, "g_inhibitory");
    istim_ = alloc<float>(N, "i_stimulus");
    refr_ = alloc<uint32_t>(N, "refractory_counter");
    spikes_ = alloc<uint32_t>(N, "spikes");
    bg_ge_ = alloc<float>(N, "background_g_excitatory");
    bg_gi_ = alloc<float>(N, "background_g_inhibitory");
    bre_ = alloc<float>(N, "background_excitatory_r_traces");
    bde_ = alloc<float>(N, "background_excitatory_d_traces");
    bri_ = alloc<float>(N, "background_inhibitory_r_traces");
    bdi_ = alloc<float>(N, "background_inhibitory_d_traces");
    clock_ = alloc<bn_device_clock>(1, "clock");
    status_ = alloc<uint32_t>(1, "status");
    // Weights: EE, EI, IE, II, row-major [pre][post] with a 64-float (256 B) aligned row stride
    // (or the reference's unpadded rows).
    const uint32_t pres[4] = {ne_, ne_, ni_, ni_}, posts[4] = {ne_, ni_, ne_, ni_};
    for (int k = 0; k < 4; ++k) {
        pre_[k] = pres[k];
        post_[k] = posts[k];
        ld_[k] = opt_.reference_layout ? posts[k] : round_up(posts[k], 64);
        row_block_[k] = opt_.gemv_row_block[k] ? opt_.gemv_row_block[k] : default_row_block(pres, posts);
        w_[k] = alloc<float>((size_t)pres[k] * ld_[k], "weight_matrix");
        tr_r_[k] = alloc<float>(pres[k], "postsynaptic_r_traces");
        tr_d_[k] = alloc<float>(pres[k], "postsynaptic_d_traces");
        tr_x_[k] = alloc<float>(pres[k], "summed_postsynaptic_conductances");
        partials_[k] = alloc<double>((size_t)cdivu(pres[k], row_block_[k]) * ld_[k], "gemv_partials");
    }
    stim_count_ = alloc<uint32_t>(total_simulation_length_, "timestep_event_count");
    stim_ids_ = alloc<uint32_t>(num_stimulus_events_, "stimulus_event_neuron_ids");
    stim_offs_ = alloc<uint32_t>(num_stimulus_events_, "stimulus_event_post_group_offsets");
    stim_amps_ = alloc<float>(num_stimulus_events_, "stimulus_event_amplitude_changes");
    const size_t np = (size_t)num_plasticities_ * N;
    pre_tr_ = alloc<float>(np, "pre_traces");
    post_tr_ = alloc<float>(np, "post_traces");
    tpre_tr_ = alloc<float>(np, "triplet_pre_traces");
    tpost_tr_ = alloc<float>(np, "triplet_post_traces");
    pre_wsum_ = alloc<float>(np, "pre_weight_sums");
    const uint32_t T = opt_.recording_interval_steps;
    if (opt_.reference_histories) {
        spike_hist_ = alloc<uint32_t>((size_t)N * T, "spike_history");
        v_hist_ = alloc<float>((size_t)N * T, "membrane_potential_history");
    }
    words_ = cdivu(N, 32);
    spike_bits_ = alloc<uint32_t>(words_, "spike_bits");
    if (opt_.fused_rings) {
        spike_ring_ = alloc<uint32_t>((size_t)words_ * T, "spike_bits_ring");
        const size_t nv = opt_.vrec_global_ids.size();
        v_ring_ = alloc<float>(std::max<size_t>(nv, 1) * T, "voltage_ring");
        vrec_index_ = alloc<int32_t>(N, "voltage_ring_index");
        std::vector<int32_t> idx(N, -1);
        for (size_t k = 0; k < nv; ++k) {
            if (opt_.vrec_global_ids[k] >= N) throw Error(BN_ERR_OUT_OF_RANGE, "recorded neuron id out of range");
            idx[opt_.vrec_global_ids[k]] = (int32_t)k;
        }
        BN_HIP_CHECK(hipMemcpyAsync(vrec_index_, idx.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, stream_),
                     "H2D voltage ring index");
    }
    if (opt_.lazy_normalize) {
        row_scale_ = alloc<float>(ne_, "row_scale");
        row_sum_ = alloc<double>(ne_, "row_sum");
    }
    // The reference seeds cuRAND with (unsigned long long)rng_seed (kernel_init_neuron_rng_states.cu:20-25).
    bn_device_clock c{};
    const uint64_t key = (uint64_t)opt_.rng_seed;
    c.rng_key[0] = (uint32_t)key;
    c.rng_key[1] = (uint32_t)(key >> 32);
    BN_HIP_CHECK(hipMemcpyAsync(clock_, &c, sizeof c, hipMemcpyHostToDevice, stream_), "H2D clock");
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "state init sync");
}

GpuSimulationState::~GpuSimulationState() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : allocations_) (void)hipFree(p);
    for (auto& g : poisson_)
        for (auto* t : g)
            if (t) (void)hipFree(t);
    if (stream_) (void)hipStreamDestroy(stream_);
}

std::string GpuSimulationState::memory_report() const {
    std::ostringstream o;
    o << "Total GPU Memory Allocated: " << allocated_ << " bytes (" << (double)allocated_ / (1024.0 * 1024.0) << " MiB)";
    return o.str();
}

void GpuSimulationState::set_poisson_table(uint32_t group, uint32_t kind, float mean) {
    std::vector<uint32_t> t(256);
    uint32_t len = 0;
    BN_CALL(bn_poisson_table(mean, t.data(), 256, &len));
    if (poisson_[group][kind]) {
        (void)hipFree(poisson_[group][kind]);
        poisson_[group][kind] = nullptr;
    }
    poisson_len_[group][kind] = len;
    if (len) {
        void* p = nullptr;
        BN_HIP_CHECK(hipMalloc(&p, len * sizeof(uint32_t)), "hipMalloc poisson table");
        BN_HIP_CHECK(hipMemcpyAsync(p, t.data(), len * sizeof(uint32_t), hipMemcpyHostToDevice, stream_),
                     "H2D poisson table");
        BN_HIP_CHECK(hipStreamSynchronize(stream_), "poisson table sync");  // `t` is pageable and local
        poisson_[group][kind] = static_cast<uint32_t*>(p);
    }
}

void GpuSimulationState::require_upload_bounds(const std::string& field, uint32_t offset, uint32_t count) const {
    if (offset > n_ || count > n_ - offset)
        throw Error(BN_ERR_INVALID_ARGUMENT, field + " slice is out of bounds. offset=" + std::to_string(offset) +
                                                 ", count=" + std::to_string(count) +
                                                 ", total_neurons=" + std::to_string(n_));
}

void GpuSimulationState::upload_neuron_group_view(const NeuronGroupView& v, const NeuronGroupInitHostData& h) {
    const uint32_t n = v.num_neurons;
    const size_t sizes[12] = {h.membrane_potentials.size(), h.g_excitatory.size(), h.g_inhibitory.size(),
                              h.i_stimulus.size(), h.refractory_counter.size(), h.spikes.size(),
                              h.background_g_excitatory_synaptic_conductances.size(),
                              h.background_g_inhibitory_synaptic_conductances.size(),
                              h.background_excitatory_postsynaptic_r_traces.size(),
                              h.background_excitatory_postsynaptic_d_traces.size(),
                              h.background_inhibitory_postsynaptic_r_traces.size(),
                              h.background_inhibitory_postsynaptic_d_traces.size()};
    for (size_t s : sizes)
        if (s != n)
            throw Error(BN_ERR_INVALID_ARGUMENT,
                        "upload_neuron_group_view host data vectors must all match gpu_view.num_neurons");
    require_upload_bounds("neuron_group", v.offset, n);
    auto up = [&](void* dst, const void* src, size_t bytes, const char* what) {
        BN_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), what);
    };
    up(v.membrane_potentials, h.membrane_potentials.data(), n * 4, "H2D membrane_potentials slice");
    up(v.g_excitatory, h.g_excitatory.data(), n * 4, "H2D g_excitatory slice");
    up(v.g_inhibitory, h.g_inhibitory.data(), n * 4, "H2D g_inhibitory slice");
    up(v.i_stimulus, h.i_stimulus.data(), n * 4, "H2D i_stimulus slice");
    up(v.refractory_counter, h.refractory_counter.data(), n * 4, "H2D refractory_counter slice");
    up(v.spikes, h.spikes.data(), n * 4, "H2D spikes slice");
    up(v.background_g_excitatory_synaptic_conductances, h.background_g_excitatory_synaptic_conductances.data(), n * 4,
       "H2D background_g_excitatory slice");
    up(v.background_g_inhibitory_synaptic_conductances, h.background_g_inhibitory_synaptic_conductances.data(), n * 4,
       "H2D background_g_inhibitory slice");
    up(v.background_excitatory_postsynaptic_r_traces, h.background_excitatory_postsynaptic_r_traces.data(), n * 4,
       "H2D background r traces");
    up(v.background_excitatory_postsynaptic_d_traces, h.background_excitatory_postsynaptic_d_traces.data(), n * 4,
       "H2D background d traces");
    up(v.background_inhibitory_postsynaptic_r_traces, h.background_inhibitory_postsynaptic_r_traces.data(), n * 4,
       "H2D background r traces");
    up(v.background_inhibitory_postsynaptic_d_traces, h.background_inhibitory_postsynaptic_d_traces.data(), n * 4,
       "H2D background d traces");
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "upload sync");
}

void GpuSimulationState::upload_connection_view(const ConnectionView& v, const ConnectionInitHostData& h) {
    // gpu_simulation_state.cu:221-229
    for (float w : h.weight_matrix)
        if (w < 0.0f)
            throw Error(BN_ERR_INVALID_ARGUMENT,
                        "connection weight_matrix must be non-negative to preserve inhibitory routing contract");
    const uint32_t k = v.weight_matrix_idx;
    (void)lookup_weight_matrix_name(k);
    const uint32_t pre = pre_[k], post = post_[k], ld = ld_[k];
    if (h.weight_matrix.size() != (size_t)pre * post)
        throw Error(BN_ERR_INVALID_ARGUMENT, "upload_connection_view weight matrix size mismatch");
    BN_HIP_CHECK(hipMemcpy2DAsync(w_[k], ld * sizeof(float), h.weight_matrix.data(), post * sizeof(float),
                                  post * sizeof(float), pre, hipMemcpyHostToDevice, stream_),
                 "H2D weight matrix");
    BN_HIP_CHECK(hipMemcpyAsync(tr_r_[k], h.postsynaptic_r_traces.data(), pre * 4, hipMemcpyHostToDevice, stream_),
                 "H2D postsynaptic_r_traces");
    BN_HIP_CHECK(hipMemcpyAsync(tr_d_[k], h.postsynaptic_d_traces.data(), pre * 4, hipMemcpyHostToDevice, stream_),
                 "H2D postsynaptic_d_traces");
    BN_HIP_CHECK(hipMemcpyAsync(tr_x_[k], h.summed_postsynaptic_conductances.data(), pre * 4, hipMemcpyHostToDevice,
                                stream_),
                 "H2D summed_postsynaptic_conductances");
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "upload sync");
}

void GpuSimulationState::upload_stimulus_view(const StimulusView&, const StimulusInitHostData& h) {
    // gpu_simulation_state.cu:163-178
    const uint32_t count = num_stimulus_events_;
    if (h.timestep_event_count.size() != total_simulation_length_ || h.stimulus_event_neuron_ids.size() != count ||
        h.stimulus_event_post_group_offsets.size() != count || h.stimulus_event_amplitude_changes.size() != count)
        throw Error(BN_ERR_INVALID_ARGUMENT,
                    "upload_stimulus_view host data vectors must match gpu_view.num_stimulus_events and "
                    "total_simulation_length");
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) BN_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "H2D stimulus");
    };
    up(stim_count_, h.timestep_event_count.data(), total_simulation_length_ * 4);
    up(stim_offs_, h.stimulus_event_post_group_offsets.data(), count * 4);
    up(stim_ids_, h.stimulus_event_neuron_ids.data(), count * 4);
    up(stim_amps_, h.stimulus_event_amplitude_changes.data(), count * 4);
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "upload sync");
}

void GpuSimulationState::upload_plasticity_view(const PlasticityView& v, const PlasticityInitHostData& h) {
    if (h.pre_traces.size() != v.pre_group_num_neurons || h.post_traces.size() != v.post_group_num_neurons)
        throw Error(BN_ERR_INVALID_ARGUMENT,
                    "upload_plasticity_view host data vectors must match gpu_view.pre_group_num_neurons and "
                    "gpu_view.post_group_num_neurons");
    auto up = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) BN_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "H2D plasticity");
    };
    up(v.pre_traces, h.pre_traces.data(), v.pre_group_num_neurons * 4);
    up(v.post_traces, h.post_traces.data(), v.post_group_num_neurons * 4);
    if (!(h.triplet_pre_traces.empty() && h.triplet_post_traces.empty())) {
        if (h.triplet_pre_traces.size() != v.pre_group_num_neurons ||
            h.triplet_post_traces.size() != v.post_group_num_neurons)
            throw Error(BN_ERR_INVALID_ARGUMENT,
                        "upload_plasticity_view host data triplet trace vectors must match gpu_view.pre_group_num_neurons "
                        "and gpu_view.post_group_num_neurons");
        up(v.triplet_pre_traces, h.triplet_pre_traces.data(), v.pre_group_num_neurons * 4);
        up(v.triplet_post_traces, h.triplet_post_traces.data(), v.post_group_num_neurons * 4);
        up(v.pre_weight_sums, h.pre_weight_sums.data(), v.pre_group_num_neurons * 4);
    }
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "upload sync");
}

std::string GpuSimulationState::lookup_weight_matrix_name(uint32_t idx) const {
    static const char* names[4] = {"w_ee", "w_ei", "w_ie", "w_ii"};
    if (idx >= 4) throw Error(BN_ERR_OUT_OF_RANGE, "Invalid weight matrix index: " + std::to_string(idx));
    return names[idx];
}

std::vector<uint32_t> GpuSimulationState::download_and_flush_spike_history(uint32_t offset, uint32_t num_neurons,
                                                                           uint32_t num_timesteps) {
    if (num_neurons == 0) return {};
    if (offset + num_neurons > n_)
        throw Error(BN_ERR_OUT_OF_RANGE, "download_and_flush_spike_history: neuron range out of bounds");
    const size_t elems = (size_t)num_neurons * num_timesteps;
    std::vector<uint32_t> host(elems);
    uint32_t* src = spike_hist_ + (size_t)offset * num_timesteps;
    BN_HIP_CHECK(hipMemcpyAsync(host.data(), src, elems * 4, hipMemcpyDeviceToHost, stream_),
                 "hipMemcpy spike history device->host failed");
    BN_CALL(launch_fill_u32(src, 0u, elems, stream_));
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "spike history sync");
    return host;
}

std::vector<float> GpuSimulationState::download_and_flush_membrane_potential_history(uint32_t offset,
                                                                                     uint32_t num_neurons,
                                                                                     uint32_t num_timesteps) {
    if (num_neurons == 0) return {};
    if (offset + num_neurons > n_)
        throw Error(BN_ERR_OUT_OF_RANGE, "download_and_flush_membrane_potential_history: neuron range out of bounds");
    const size_t elems = (size_t)num_neurons * num_timesteps;
    std::vector<float> host(elems);
    float* src = v_hist_ + (size_t)offset * num_timesteps;
    BN_HIP_CHECK(hipMemcpyAsync(host.data(), src, elems * 4, hipMemcpyDeviceToHost, stream_),
                 "hipMemcpy membrane potential history device->host failed");
    BN_CALL(launch_fill_u32(reinterpret_cast<uint32_t*>(src), 0u, elems, stream_));
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "membrane potential history sync");
    return host;
}

const std::vector<uint32_t>& GpuSimulationState::spike_ring_snapshot(uint32_t clock_step) {
    if (ring_snapshot_step_ != (int64_t)clock_step) {
        const size_t elems = (size_t)words_ * opt_.recording_interval_steps;
        ring_snapshot_.resize(elems);
        BN_HIP_CHECK(hipMemcpyAsync(ring_snapshot_.data(), spike_ring_, elems * 4, hipMemcpyDeviceToHost, stream_),
                     "D2H spike ring");
        BN_CALL(launch_fill_u32(spike_ring_, 0u, elems, stream_));
        BN_HIP_CHECK(hipStreamSynchronize(stream_), "spike ring sync");
        ring_snapshot_step_ = clock_step;
    }
    return ring_snapshot_;
}

// Rows [first_step, first_step + num_steps) of the spike-bit ring: step s lives in slot
// s % interval until the slot is written again or the ring is flushed by a record step.
std::vector<uint32_t> GpuSimulationState::download_spike_ring_rows(uint32_t first_step, uint32_t num_steps,
                                                                   uint32_t clock_step) const {
    if (!spike_ring_) throw Error(BN_ERR_RUNTIME, "spike raster: no spike ring (the simulation runs without monitors)");
    const uint32_t T = opt_.recording_interval_steps;
    const uint64_t end = (uint64_t)first_step + num_steps;
    if (end > clock_step) throw Error(BN_ERR_OUT_OF_RANGE, "spike raster: steps past the last executed step");
    if ((uint64_t)clock_step - first_step > T)
        throw Error(BN_ERR_OUT_OF_RANGE, "spike raster: steps older than one recording interval were overwritten");
    if (ring_snapshot_step_ >= 0 && (int64_t)first_step <= ring_snapshot_step_)
        throw Error(BN_ERR_OUT_OF_RANGE, "spike raster: steps up to the last record step were flushed to the recorders");
    std::vector<uint32_t> host((size_t)num_steps * words_);
    for (uint32_t k = 0; k < num_steps;) {
        const uint32_t slot = (first_step + k) % T;
        const uint32_t run = std::min(num_steps - k, T - slot);
        BN_HIP_CHECK(hipMemcpyAsync(host.data() + (size_t)k * words_, spike_ring_ + (size_t)slot * words_,
                                    (size_t)run * words_ * 4, hipMemcpyDeviceToHost, stream_),
                     "D2H spike ring rows");
        k += run;
    }
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "spike ring rows sync");
    return host;
}

std::vector<float> GpuSimulationState::download_and_flush_v_ring() {
    const size_t elems = (size_t)std::max<uint32_t>(n_vrec(), 1) * opt_.recording_interval_steps;
    std::vector<float> host(elems);
    BN_HIP_CHECK(hipMemcpyAsync(host.data(), v_ring_, elems * 4, hipMemcpyDeviceToHost, stream_), "D2H voltage ring");
    BN_CALL(launch_fill_u32(reinterpret_cast<uint32_t*>(v_ring_), 0u, elems, stream_));
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "voltage ring sync");
    return host;
}

std::vector<float> GpuSimulationState::download_weight_matrix(uint32_t idx) const {
    (void)lookup_weight_matrix_name(idx);
    const uint32_t pre = pre_[idx], post = post_[idx];
    std::vector<float> host((size_t)pre * post);
    BN_HIP_CHECK(hipMemcpy2DAsync(host.data(), post * sizeof(float), w_[idx], ld_[idx] * sizeof(float),
                                  post * sizeof(float), pre, hipMemcpyDeviceToHost, stream_),
                 "D2H weight matrix");
    std::vector<float> scale;
    if (lazy_active_ && idx == lazy_matrix_) {
        scale.resize(pre);
        BN_HIP_CHECK(hipMemcpyAsync(scale.data(), row_scale_, pre * 4, hipMemcpyDeviceToHost, stream_), "D2H row scale");
    }
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "D2H weight sync");
    if (!scale.empty())
        for (uint32_t i = 0; i < pre; ++i)
            for (uint32_t j = 0; j < post; ++j) host[(size_t)i * post + j] = scale[i] * host[(size_t)i * post + j];
    return host;
}

void GpuSimulationState::setup_lazy_normalization(uint32_t idx) {
    if (!opt_.lazy_normalize) return;
    lazy_active_ = true;
    lazy_matrix_ = idx;
}

void GpuSimulationState::check_status() {
    uint32_t st = 0;
    BN_HIP_CHECK(hipMemcpyAsync(&st, status_, 4, hipMemcpyDeviceToHost, stream_), "D2H status");
    BN_HIP_CHECK(hipStreamSynchronize(stream_), "status sync");
    if (st == 1u) throw Error(BN_ERR_RUNTIME, "Stimulus event indexing exceeded uploaded event buffers");
    if (st == 2u)
        throw Error(BN_ERR_RUNTIME, "persistent step: a block waited past its limit and the launch was abandoned "
                                    "(the simulation state is no longer valid)");
}

// ======================================================================================
// NeuronGroup (neuron_group.cpp:4-85)
// ======================================================================================
NeuronGroup::NeuronGroup(std::vector<Location> neuron_locations, NeuronGroupParams params,
                         GpuSimulationState& gpu_state)
    : gpu_state_(gpu_state), params_(params) {
    if (neuron_locations.size() != params_.num_neurons)
        throw Error(BN_ERR_INVALID_ARGUMENT, "Size of neuron_locations does not match num_neurons");
    neuron_locations_ = std::move(neuron_locations);
    NeuronGroupInitHostData h = initialize_data();
    make_gpu_view();
    upload_data(h);
}

NeuronGroupInitHostData NeuronGroup::initialize_data() {
    const uint32_t n = params_.num_neurons;
    NeuronGroupInitHostData h;
    h.membrane_potentials.assign(n, params_.v_rest);
    h.g_excitatory.assign(n, 0.0f);
    h.g_inhibitory.assign(n, 0.0f);
    h.i_stimulus.assign(n, 0.0f);
    h.refractory_counter.assign(n, 0);
    h.spikes.assign(n, 0);
    h.background_g_excitatory_synaptic_conductances.assign(n, 0.0f);
    h.background_g_inhibitory_synaptic_conductances.assign(n, 0.0f);
    h.background_excitatory_postsynaptic_r_traces.assign(n, 0.0f);
    h.background_excitatory_postsynaptic_d_traces.assign(n, 0.0f);
    h.background_inhibitory_postsynaptic_r_traces.assign(n, 0.0f);
    h.background_inhibitory_postsynaptic_d_traces.assign(n, 0.0f);
    return h;
}

void NeuronGroup::make_gpu_view() {
    NeuronGroupView& v = gpu_view_;
    const NeuronGroupParams& p = params_;
    v.timestep = p.timestep;
    v.g_leak = p.g_leak;
    v.membrane_capacitance = p.membrane_capacitance;
    v.depolarization_slope_factor = p.depolarization_slope_factor;
    v.v_effective_spike_threshold = p.v_effective_spike_threshold;
    v.v_excitatory_reversal = p.v_excitatory_reversal;
    v.v_inhibitory_reversal = p.v_inhibitory_reversal;
    v.v_threshold = p.v_threshold;
    v.v_reset = p.v_reset;
    v.v_rest = p.v_rest;
    v.refractory_steps = p.refractory_steps;
    v.is_inhibitory = p.is_inhibitory;
    v.mean_background_excitatory_spikes = p.mean_background_excitatory_spikes;
    v.mean_background_inhibitory_spikes = p.mean_background_inhibitory_spikes;
    v.background_excitatory_postsynaptic_r_trace_decay = p.background_excitatory_postsynaptic_r_trace_decay;
    v.background_excitatory_postsynaptic_d_trace_decay = p.background_excitatory_postsynaptic_d_trace_decay;
    v.background_inhibitory_postsynaptic_r_trace_decay = p.background_inhibitory_postsynaptic_r_trace_decay;
    v.background_inhibitory_postsynaptic_d_trace_decay = p.background_inhibitory_postsynaptic_d_trace_decay;
    v.background_excitatory_scaling_factor = p.background_excitatory_scaling_factor;
    v.background_inhibitory_scaling_factor = p.background_inhibitory_scaling_factor;
    v.num_neurons = p.num_neurons;
    v.offset = p.offset;
    GpuSimulationState& s = gpu_state_;
    v.membrane_potentials = s.d_membrane_potentials() + p.offset;
    v.g_excitatory = s.d_g_excitatory() + p.offset;
    v.g_inhibitory = s.d_g_inhibitory() + p.offset;
    v.i_stimulus = s.d_i_stimulus() + p.offset;
    v.refractory_counter = s.d_refractory_counter() + p.offset;
    v.spikes = s.d_spikes() + p.offset;
    v.background_g_excitatory_synaptic_conductances = s.d_background_g_excitatory_synaptic_conductances() + p.offset;
    v.background_g_inhibitory_synaptic_conductances = s.d_background_g_inhibitory_synaptic_conductances() + p.offset;
    v.background_excitatory_postsynaptic_r_traces = s.d_background_excitatory_postsynaptic_r_traces() + p.offset;
    v.background_excitatory_postsynaptic_d_traces = s.d_background_excitatory_postsynaptic_d_traces() + p.offset;
    v.background_inhibitory_postsynaptic_r_traces = s.d_background_inhibitory_postsynaptic_r_traces() + p.offset;
    v.background_inhibitory_postsynaptic_d_traces = s.d_background_inhibitory_postsynaptic_d_traces() + p.offset;
    v.rng_clock = s.d_clock();
    const uint32_t g = p.is_inhibitory ? 1 : 0;
    s.set_poisson_table(g, 0, p.mean_background_excitatory_spikes);
    s.set_poisson_table(g, 1, p.mean_background_inhibitory_spikes);
    v.poisson_table_excitatory = s.d_poisson_table(g, 0);
    v.poisson_table_inhibitory = s.d_poisson_table(g, 1);
    v.poisson_table_excitatory_len = s.poisson_table_len(g, 0);
    v.poisson_table_inhibitory_len = s.poisson_table_len(g, 1);
    v.rng_instance = 0;
}

void NeuronGroup::upload_data(const NeuronGroupInitHostData& h) { gpu_state_.upload_neuron_group_view(gpu_view_, h); }

void NeuronGroup::evolve() { BN_CALL(bn_neuron_group_evolve(gpu_view_, gpu_state_.stream())); }

// ======================================================================================
// Connection (connection.cpp:3-95)
// ======================================================================================
Connection::Connection(NeuronGroup& pre_group, NeuronGroup& post_group, const std::vector<float>& weight_matrix,
                       ConnectionParams params, GpuSimulationState& gpu_state)
    : pre_group_(pre_group), post_group_(post_group), params_(params), gpu_state_(gpu_state) {
    if (weight_matrix.size() != (size_t)pre_group_.num_neurons() * post_group_.num_neurons() || weight_matrix.empty())
        throw Error(BN_ERR_INVALID_ARGUMENT,
                    "Weight matrix size does not match the number of connections between pre and post groups.");
    ConnectionInitHostData h;
    h.weight_matrix = weight_matrix;
    h.postsynaptic_r_traces.assign(pre_group_.num_neurons(), 0.0f);
    h.postsynaptic_d_traces.assign(pre_group_.num_neurons(), 0.0f);
    h.summed_postsynaptic_conductances.assign(pre_group_.num_neurons(), 0.0f);
    make_gpu_view();
    gpu_state_.upload_connection_view(gpu_view_, h);
}

void Connection::make_gpu_view() {
    ConnectionView& v = gpu_view_;
    v.postsynaptic_r_trace_decay = std::exp(-1.0f * params_.timestep / params_.rise_time);
    v.postsynaptic_d_trace_decay = std::exp(-1.0f * params_.timestep / params_.decay_time);
    v.postsynaptic_kernel_normalization_value = params_.postsynaptic_kernel_normalization_value;
    v.weight_matrix_idx = params_.weight_matrix_idx;
    v.pre_group_is_inhibtory = pre_group_.is_inhibitory();
    v.pre_group_offset = pre_group_.offset();
    v.post_group_offset = post_group_.offset();
    v.pre_group_num_neurons = pre_group_.num_neurons();
    v.post_group_num_neurons = post_group_.num_neurons();
    const uint32_t k = params_.weight_matrix_idx;
    (void)gpu_state_.lookup_weight_matrix_name(k);
    if (gpu_state_.weight_pre(k) != v.pre_group_num_neurons || gpu_state_.weight_post(k) != v.post_group_num_neurons)
        throw Error(BN_ERR_INVALID_ARGUMENT, "Connection groups do not match the shape of weight matrix " +
                                                 gpu_state_.lookup_weight_matrix_name(k));
    v.weight_matrix = gpu_state_.d_w(k);
    v.postsynaptic_r_traces = gpu_state_.d_postsynaptic_r_traces(k);
    v.postsynaptic_d_traces = gpu_state_.d_postsynaptic_d_traces(k);
    v.summed_postsynaptic_conductances = gpu_state_.d_summed_postsynaptic_conductances(k);
    v.post_group_g_excitatory = gpu_state_.d_g_excitatory() + post_group_.offset();
    v.post_group_g_inhibitory = gpu_state_.d_g_inhibitory() + post_group_.offset();
    v.spikes = gpu_state_.d_spikes() + pre_group_.offset();
    v.weight_matrix_ld = gpu_state_.weight_ld(k);
    v.gemv_partials = gpu_state_.d_gemv_partials(k);
    v.gemv_row_block = gpu_state_.gemv_row_block(k);
}

void Connection::propagate() { BN_CALL(bn_connection_propagate(gpu_view_, gpu_state_.stream())); }

// ======================================================================================
// Stimulus (stimulus.cpp:3-60)
// ======================================================================================
Stimulus::Stimulus(std::vector<uint32_t> timestep_event_count, std::vector<uint32_t> ids, std::vector<uint32_t> offs,
                   std::vector<float> amps, GpuSimulationState& gpu_state)
    : gpu_state_(gpu_state) {
    StimulusInitHostData h;
    h.timestep_event_count = std::move(timestep_event_count);
    h.stimulus_event_neuron_ids = std::move(ids);
    h.stimulus_event_post_group_offsets = std::move(offs);
    h.stimulus_event_amplitude_changes = std::move(amps);
    timestep_event_count_ = h.timestep_event_count;
    total_stimulus_events_ = (uint32_t)h.stimulus_event_neuron_ids.size();
    gpu_view_.g_excitatory = gpu_state_.d_g_excitatory();
    gpu_view_.g_inhibitory = gpu_state_.d_g_inhibitory();
    gpu_view_.i_stimulus = gpu_state_.d_i_stimulus();
    gpu_view_.timestep_event_count = gpu_state_.d_timestep_event_count();
    gpu_view_.stimulus_event_neuron_ids = gpu_state_.d_stimulus_event_neuron_ids();
    gpu_view_.stimulus_event_post_group_offsets = gpu_state_.d_stimulus_event_post_group_offsets();
    gpu_view_.stimulus_event_amplitude_changes = gpu_state_.d_stimulus_event_amplitude_changes();
    gpu_view_.num_timesteps = (uint32_t)timestep_event_count_.size();
    gpu_view_.num_events = total_stimulus_events_;
    gpu_state_.upload_stimulus_view(gpu_view_, h);
    // Event global ids must stay inside the neuron arrays.
    for (size_t e = 0; e < h.stimulus_event_neuron_ids.size(); ++e)
        if ((uint64_t)h.stimulus_event_neuron_ids[e] + h.stimulus_event_post_group_offsets[e] >= gpu_state_.total_neurons())
            throw Error(BN_ERR_OUT_OF_RANGE, "stimulus event neuron id out of range");
}

void Stimulus::apply() {
    if (current_timestep_ >= timestep_event_count_.size()) return;
    const uint32_t n = timestep_event_count_[current_timestep_];
    if (next_event_idx_ + n > total_stimulus_events_)
        throw Error(BN_ERR_RUNTIME, "Stimulus event indexing exceeded uploaded event buffers");
    BN_CALL(bn_stimulus_apply(gpu_view_, next_event_idx_, n, gpu_state_.stream()));
    next_event_idx_ += n;
    ++current_timestep_;
}

// ======================================================================================
// Plasticities (plasticity.cpp, triplet_plasticity.cpp, istdp_plasticity.cpp, stdp_plasticity.cpp)
// ======================================================================================
Plasticity::Plasticity(Connection& connection, PlasticityParams params, GpuSimulationState& gpu_state, uint32_t kind)
    : connection_(connection), params_(params), gpu_state_(gpu_state), kind_(kind) {}

void Plasticity::make_gpu_view() {
    PlasticityView& v = gpu_view_;
    v.ltp_A = params_.ltp_A;
    v.ltd_A = params_.ltd_A;
    v.ltp_B = params_.ltp_B;
    v.ltd_B = params_.ltd_B;
    v.weight_min = params_.weight_min;
    v.weight_max = params_.weight_max;
    v.pre_trace_decay = params_.pre_trace_decay;
    v.post_trace_decay = params_.post_trace_decay;
    v.triplet_pre_trace_decay = params_.triplet_pre_trace_decay;
    v.triplet_post_trace_decay = params_.triplet_post_trace_decay;
    v.plasticity_offset = params_.plasticity_offset;
    v.weight_matrix_idx = connection_.weight_matrix_idx();
    v.pre_group_offset = connection_.pre_group_offset();
    v.post_group_offset = connection_.post_group_offset();
    v.pre_group_num_neurons = connection_.pre_group_num_neurons();
    v.post_group_num_neurons = connection_.post_group_num_neurons();
    GpuSimulationState& s = gpu_state_;
    if (params_.plasticity_offset >= s.num_plasticities())
        throw Error(BN_ERR_OUT_OF_RANGE, "plasticity_offset exceeds the number of allocated plasticities");
    v.pre_spikes = s.d_spikes() + v.pre_group_offset;
    v.post_spikes = s.d_spikes() + v.post_group_offset;
    v.weight_matrix = s.d_w(v.weight_matrix_idx);
    v.weight_matrix_ld = s.weight_ld(v.weight_matrix_idx);
    const size_t base = (size_t)v.plasticity_offset * s.total_neurons();
    v.pre_traces = s.d_pre_traces() + base + v.pre_group_offset;
    v.post_traces = s.d_post_traces() + base + v.post_group_offset;
    v.triplet_pre_traces = s.d_triplet_pre_traces() + base + v.pre_group_offset;
    v.triplet_post_traces = s.d_triplet_post_traces() + base + v.post_group_offset;
    v.pre_weight_sums = s.d_pre_weight_sums() + base + v.post_group_offset;
}

TripletPlasticity::TripletPlasticity(Connection& connection, PlasticityParams params, GpuSimulationState& gpu_state)
    : Plasticity(connection, params, gpu_state, kPlastTriplet) {
    PlasticityInitHostData h;
    const uint32_t np = connection_.pre_group_num_neurons(), nq = connection_.post_group_num_neurons();
    h.pre_traces.assign(np, 0.0f);
    h.post_traces.assign(nq, 0.0f);
    h.triplet_pre_traces.assign(np, 0.0f);
    h.triplet_post_traces.assign(nq, 0.0f);
    // triplet_plasticity.cpp:22-30: initial column sums (FP32, sequential over pre)
    h.pre_weight_sums.assign(nq, 0.0f);
    const std::vector<float> w = gpu_state_.download_weight_matrix(connection_.weight_matrix_idx());
    for (uint32_t j = 0; j < nq; ++j) {
        float sum = 0.0f;
        for (uint32_t i = 0; i < np; ++i) sum += w[(size_t)i * nq + j];
        h.pre_weight_sums[j] = sum;
    }
    make_gpu_view();
    gpu_state_.upload_plasticity_view(gpu_view_, h);
}

void TripletPlasticity::update() {
    BN_CALL(bn_triplet_plasticity_update_normalize(gpu_view_, params_, gpu_state_.stream()));
}

IstdpPlasticity::IstdpPlasticity(Connection& connection, PlasticityParams params, GpuSimulationState& gpu_state)
    : Plasticity(connection, params, gpu_state, kPlastIstdp) {
    PlasticityInitHostData h;
    h.pre_traces.assign(connection_.pre_group_num_neurons(), 0.0f);
    h.post_traces.assign(connection_.post_group_num_neurons(), 0.0f);
    make_gpu_view();
    gpu_state_.upload_plasticity_view(gpu_view_, h);
}

void IstdpPlasticity::update() { BN_CALL(bn_istdp_plasticity_update(gpu_view_, params_, gpu_state_.stream())); }

StdpPlasticity::StdpPlasticity(Connection& connection, PlasticityParams params, GpuSimulationState& gpu_state)
    : Plasticity(connection, params, gpu_state, kPlastStdp) {
    PlasticityInitHostData h;
    h.pre_traces.assign(connection_.pre_group_num_neurons(), 0.0f);
    h.post_traces.assign(connection_.post_group_num_neurons(), 0.0f);
    make_gpu_view();
    gpu_state_.upload_plasticity_view(gpu_view_, h);
}

void StdpPlasticity::update() { BN_CALL(bn_stdp_plasticity_update(gpu_view_, params_, gpu_state_.stream())); }

// ======================================================================================
// Monitors
// ======================================================================================
SpikeRecorder::SpikeRecorder(const NeuronGroup& group, const std::string& output_filepath,
                             uint32_t recording_interval_steps, GpuSimulationState& gpu_state)
    : Monitor(gpu_state), group_(group), recording_interval_steps_(recording_interval_steps) {
    if (!output_filepath.empty()) output_.open(output_filepath, std::ios::trunc);
    gpu_view_.num_neurons = group_.num_neurons();
    gpu_view_.offset = group_.offset();
    gpu_view_.spikes = gpu_state_.d_spikes() + group_.offset();
    gpu_view_.spike_history =
        gpu_state_.d_spike_history() ? gpu_state_.d_spike_history() + (size_t)group_.offset() * recording_interval_steps
                                     : nullptr;
    if (output_.good()) output_ << "clock_step, neuron_ids\n";  // spike_recorder.cpp:22-28
}

void SpikeRecorder::historicize(uint32_t clock_step) {
    if (fused) return;
    BN_CALL(bn_historicize_spikes(gpu_view_.spikes, gpu_view_.spike_history, gpu_view_.num_neurons,
                                  recording_interval_steps_, clock_step, gpu_state_.stream()));
}

// spike_recorder.cpp:40-60 — slot 0 holds the record step itself but is labelled C - interval.
void SpikeRecorder::record(uint32_t clock_step) {
    const uint32_t T = recording_interval_steps_;
    if (fused) {
        const std::vector<uint32_t>& ring = gpu_state_.spike_ring_snapshot(clock_step);
        record_from_rings(clock_step, ring.data(), nullptr);
        return;
    }
    if (clock_step < T) return;
    const uint32_t first = clock_step - T;
    std::string buf;
    buf.reserve(1 << 20);
    char tmp[32];
    const std::vector<uint32_t> h = gpu_state_.download_and_flush_spike_history(gpu_view_.offset, gpu_view_.num_neurons, T);
    for (uint32_t inner = 0; inner < T; ++inner) {
        buf += std::to_string(first + inner) + ", ";
        for (uint32_t n = 0; n < gpu_view_.num_neurons; ++n)
            if (h[(size_t)n * T + inner] == 1) {
                const int len = std::snprintf(tmp, sizeof tmp, "%u ", n);
                buf.append(tmp, (size_t)len);
            }
        buf += "\n";
    }
    if (output_.good()) {
        output_ << buf;
        output_.flush();
    }
}

// The same lines from the fused step's spike-bit ring ([interval][words], bit g & 31 of word
// g >> 5 for global neuron g), the set bits of the group's words scanned in order.
void SpikeRecorder::record_from_rings(uint32_t clock_step, const uint32_t* ring, const float*) {
    const uint32_t T = recording_interval_steps_;
    if (clock_step < T || !output_.good()) return;
    const uint32_t first = clock_step - T;
    const uint32_t W = gpu_state_.spike_words();
    const uint32_t lo = gpu_view_.offset, hi = lo + gpu_view_.num_neurons;
    std::string buf;
    buf.reserve(1 << 20);
    char tmp[32];
    for (uint32_t inner = 0; inner < T; ++inner) {
        buf += std::to_string(first + inner) + ", ";
        const uint32_t* row = ring + (size_t)inner * W;
        for (uint32_t wd = lo >> 5; wd <= (hi - 1) >> 5 && lo < hi; ++wd) {
            uint32_t bits = row[wd];
            while (bits) {
                const uint32_t g = 32u * wd + (uint32_t)__builtin_ctz(bits);
                bits &= bits - 1u;
                if (g < lo || g >= hi) continue;
                const int len = std::snprintf(tmp, sizeof tmp, "%u ", g - lo);
                buf.append(tmp, (size_t)len);
            }
        }
        buf += "\n";
    }
    output_ << buf;
    output_.flush();
}

VoltageRecorder::VoltageRecorder(const NeuronGroup& group, const std::string& output_filepath,
                                 std::vector<uint32_t> selected_neurons, uint32_t recording_interval_steps,
                                 GpuSimulationState& gpu_state)
    : Monitor(gpu_state),
      group_(group),
      output_filepath_(output_filepath),
      selected_neurons_(std::move(selected_neurons)),
      recording_interval_steps_(recording_interval_steps) {
    if (selected_neurons_.empty()) {
        selected_neurons_.resize(group_.num_neurons());
        std::iota(selected_neurons_.begin(), selected_neurons_.end(), 0u);
    } else {
        for (uint32_t id : selected_neurons_)
            if (id >= group_.num_neurons())
                throw Error(BN_ERR_OUT_OF_RANGE,
                            "Selected neuron ID " + std::to_string(id) + " is out of range for the neuron group.");
    }
    gpu_view_.num_neurons = group_.num_neurons();
    gpu_view_.offset = group_.offset();
    gpu_view_.membrane_potentials = gpu_state_.d_membrane_potentials() + group_.offset();
    gpu_view_.membrane_potential_history =
        gpu_state_.d_membrane_potential_history()
            ? gpu_state_.d_membrane_potential_history() + (size_t)group_.offset() * recording_interval_steps
            : nullptr;
}

void VoltageRecorder::historicize(uint32_t clock_step) {
    if (fused) return;
    BN_CALL(bn_historicize_membrane_potentials(gpu_view_.membrane_potentials, gpu_view_.membrane_potential_history,
                                               gpu_view_.num_neurons, recording_interval_steps_, clock_step,
                                               gpu_state_.stream()));
}

// voltage_recorder.cpp:54-77 — one file per history slot, name index computed in uint32.
void VoltageRecorder::record(uint32_t clock_step) {
    const uint32_t T = recording_interval_steps_;
    if (fused) {
        const std::vector<float> h = gpu_state_.download_and_flush_v_ring();
        record_from_rings(clock_step, nullptr, h.data());
        return;
    }
    if (!gpu_state_.options().reference_histories) return;
    const std::vector<float> h =
        gpu_state_.download_and_flush_membrane_potential_history(gpu_view_.offset, gpu_view_.num_neurons, T);
    write_files(clock_step, h.data(), 0, nullptr);
}

// From the fused step's voltage ring ([interval][n_vrec], columns in the order of the
// simulation's recorded ids).
void VoltageRecorder::record_from_rings(uint32_t clock_step, const uint32_t*, const float* v_ring) {
    std::vector<int64_t> col(selected_neurons_.size());
    const auto& ids = gpu_state_.options().vrec_global_ids;
    for (size_t k = 0; k < selected_neurons_.size(); ++k) {
        const uint32_t g = gpu_view_.offset + selected_neurons_[k];
        col[k] = std::find(ids.begin(), ids.end(), g) - ids.begin();
    }
    write_files(clock_step, v_ring, gpu_state_.n_vrec(), col.data());
}

// The interval's files: one per history slot hs, named with k = C - T + hs + 1 in uint32.
// Values: v[hs * stride_n + col[q]] from a ring (col given), else v[id * T + hs] (histories).
// The files are independent, so they are written by up to 8 threads over slot ranges.
void VoltageRecorder::write_files(uint32_t clock_step, const float* v, uint32_t stride_n, const int64_t* col) {
    const uint32_t T = recording_interval_steps_;
    if (output_filepath_.empty()) return;
    const size_t pos = output_filepath_.rfind(".txt");
    if (!quiet) std::cout << "Recording membrane potentials at clock step " << clock_step << "...\n";
    auto write_range = [&](uint32_t h0, uint32_t h1) {
        std::string text;
        char tmp[64];
        for (uint32_t hs = h0; hs < h1; ++hs) {
            std::string name = output_filepath_;
            const uint32_t k = (clock_step - T) + hs + 1;
            name.insert(pos == std::string::npos ? name.size() : pos, "_step" + std::to_string(k));
            text.assign("neuron_id, voltage\n");
            for (size_t q = 0; q < selected_neurons_.size(); ++q) {
                const uint32_t id = selected_neurons_[q];
                const float val = col ? v[(size_t)hs * stride_n + (size_t)col[q]] : v[(size_t)id * T + hs];
                std::ostringstream os;   // the stream formatting of the reference's `out << v`
                os << val;
                const int len = std::snprintf(tmp, sizeof tmp, "%u, ", id);
                text.append(tmp, (size_t)len).append(os.str()).append("\n");
            }
            FILE* f = std::fopen(name.c_str(), "w");
            if (!f) return;
            std::fwrite(text.data(), 1, text.size(), f);
            std::fclose(f);
        }
    };
    const uint32_t nt = std::min<uint32_t>(8, std::max<uint32_t>(1, T / 256));
    if (nt <= 1) {
        write_range(0, T);
        return;
    }
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < nt; ++t) pool.emplace_back(write_range, T * t / nt, T * (t + 1) / nt);
    for (auto& th : pool) th.join();
}

WeightRecorder::WeightRecorder(const Connection& connection, const std::string& output_filepath,
                               GpuSimulationState& gpu_state)
    : Monitor(gpu_state), connection_(connection), output_filepath_(output_filepath) {}

// weight_recorder.cpp:25-51
void WeightRecorder::record(uint32_t clock_step) {
    if (clock_step % 100000 != 0) return;
    if (output_filepath_.empty()) return;
    std::string name = output_filepath_;
    const size_t pos = name.rfind(".txt");
    name.insert(pos == std::string::npos ? name.size() : pos, "_step" + std::to_string(clock_step));
    FILE* f = std::fopen(name.c_str(), "w");
    if (!quiet) std::cout << "clock step " << clock_step << " saved weights: " << name << "\n";
    if (!f) return;
    const std::vector<float> w = gpu_state_.download_weight_matrix(connection_.weight_matrix_idx());
    std::fputs("pre_neuron_id, post_neuron_id, weight\n", f);
    const uint32_t np = connection_.pre_group_num_neurons(), nq = connection_.post_group_num_neurons();
    std::vector<char> buf;
    buf.reserve(1 << 22);
    char tmp[96];
    for (uint32_t i = 0; i < np; ++i) {
        for (uint32_t j = 0; j < nq; ++j) {
            const int len = std::snprintf(tmp, sizeof tmp, "%u, %u, %g\n", i, j, (double)w[(size_t)i * nq + j]);
            buf.insert(buf.end(), tmp, tmp + len);
        }
        if (buf.size() > (1u << 22)) {
            std::fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
}

}  // namespace bn
