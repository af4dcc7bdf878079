// System: the reference's step orchestrator (cpp_src/core/system.cpp:43-83) plus the fused,
// graph-captured step of the whole network.
#include <chrono>
#include <cstring>
#include <iostream>

#include "../bn_errors.h"
#include "engine.h"

namespace bn {

System::~System() {
    pipeline_.reset();   // the writer finishes its queued intervals first
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (void* p : staging_) (void)hipFree(p);
    for (void* p : persist_bufs_) (void)hipFree(p);
    if (pre_list_buf_) (void)hipFree(pre_list_buf_);
}

void System::add_neuron_group(std::unique_ptr<NeuronGroup> g) {
    if (g) neuron_groups_.push_back(std::move(g));
}
void System::add_connection(std::unique_ptr<Connection> c) {
    if (c) connections_.push_back(std::move(c));
}
void System::add_stimulus(std::unique_ptr<Stimulus> s) {
    if (s) stimuli_.push_back(std::move(s));
}
void System::add_plasticity(std::unique_ptr<Plasticity> p) {
    if (p) plasticities_.push_back(std::move(p));
}
void System::add_monitor(std::unique_ptr<Monitor> m) {
    if (m) monitors_.push_back(std::move(m));
}
void System::add_step_callback(std::function<void(uint32_t)> cb) {
    if (cb) step_callbacks_.push_back(std::move(cb));
}

// system.cpp:43-63, then the device clock advances (counter-based RNG).
void System::step_the_clock(uint32_t clock_step) {
    if (fused_) throw Error(BN_ERR_RUNTIME, "step_the_clock: the fused step is enabled for this system");
    for (auto& g : neuron_groups_) g->evolve();
    for (auto& c : connections_) c->propagate();
    for (auto& s : stimuli_) s->apply();
    for (auto& p : plasticities_) p->update();
    for (auto& m : monitors_) m->historicize(clock_step);
    BN_CALL(bn_clock_advance(state_.d_clock(), state_.stream()));
    clock_step_ = clock_step + 1;
}

void System::set_async_recording(bool on) {
    if (pipeline_) pipeline_->drain();
    async_recording_ = on;
}

// system.cpp:65-83.  Fused step with the recording pipeline: a record step queues its rings'
// copies and the writer thread writes the files while the next interval's steps run; the run
// returns once the files are written, as the reference's does.
void System::run(uint32_t total_clock_steps) {
    using clock = std::chrono::steady_clock;
    const auto start = clock::now();
    const uint32_t interval = state_.options().recording_interval_steps;
    uint32_t step = clock_step_;
    const uint32_t end = clock_step_ + total_clock_steps;
    std::vector<Monitor*> ring_monitors;
    for (auto& m : monitors_)
        if (m->from_rings()) ring_monitors.push_back(m.get());
    const bool piped = fused_ && async_recording_ && !ring_monitors.empty();
    if (piped && !pipeline_) pipeline_ = std::make_unique<RecordPipeline>(state_, ring_monitors);
    while (step < end) {
        uint32_t n = 1;
        if (fused_ && step_callbacks_.empty()) {
            // Run up to and including the next record step in one batch.
            const uint32_t next_record = (step % interval == 0) ? step : (step / interval + 1) * interval;
            n = std::min(end, next_record + 1) - step;
        }
        for (uint32_t k = 0; k < n; ++k)
            for (auto& cb : step_callbacks_) cb(step + k);
        if (fused_) launch_fused_steps(n);
        else step_the_clock(step);
        step += n;
        clock_step_ = step;
        const uint32_t last = step - 1;
        if (last % interval == 0) {
            const auto r0 = clock::now();
            if (piped) {
                pipeline_->submit(last);
                for (auto& m : monitors_)   // others (the weight dump) read the state now
                    if (!m->from_rings()) m->record(last);
            } else {
                BN_HIP_CHECK(hipStreamSynchronize(state_.stream()), "step sync");
                state_.check_status();
                for (auto& m : monitors_) m->record(last);
            }
            const std::chrono::duration<double> rt = clock::now() - r0;
            record_call_seconds_ += rt.count();
            ++record_steps_;
            const std::chrono::duration<float> elapsed = clock::now() - start;
            if (!quiet) std::cout << "Completed clock step: " << last << " | elapsed: " << elapsed.count() << "s\n";
        }
    }
    if (piped) {
        pipeline_->drain();
        BN_HIP_CHECK(hipStreamSynchronize(state_.stream()), "step sync");
        state_.check_status();
    }
}

const TraceSet* System::trace_set_for_connection(uint32_t idx) const {
    if (!fused_ || idx >= conn_trace_set_.size()) return nullptr;
    return &evolve_.ts[conn_trace_set_[idx]];
}

void System::build_plans() {
    GpuSimulationState& S = state_;
    if (neuron_groups_.size() > (size_t)kMaxGroups) throw Error(BN_ERR_INVALID_ARGUMENT, "too many neuron groups");
    if (connections_.size() > (size_t)kMaxConns) throw Error(BN_ERR_INVALID_ARGUMENT, "too many connections");
    if (plasticities_.size() > (size_t)kMaxPlast) throw Error(BN_ERR_INVALID_ARGUMENT, "too many plasticities");
    if (stimuli_.size() > 1) throw Error(BN_ERR_INVALID_ARGUMENT, "the fused step supports one stimulus");
    evolve_ = EvolvePlan{};
    gemv_ = GemvPlan{};
    plast_ = PlastPlan{};
    tail_ = TailPlan{};
    conn_trace_set_.assign(connections_.size(), -1);

    // Part sums: every connection's rows padded to 16 bytes and at most kGemvPartChunksMax
    // chunks per part.
    bool parts = gemv_part_sums_;
    for (size_t c = 0; c < connections_.size() && parts; ++c) {
        const ConnectionView& v = connections_[c]->gpu_view();
        const uint32_t chunks = (v.pre_group_num_neurons + v.gemv_row_block - 1) / v.gemv_row_block;
        const uint32_t per = (chunks + kFinParts - 1) / kFinParts;
        if (v.weight_matrix_ld % 4 != 0 || (uintptr_t)v.weight_matrix % 16 != 0 || per > kGemvPartChunksMax) parts = false;
    }
    auto part_count = [](uint32_t chunks) {
        const uint32_t per = (chunks + kFinParts - 1) / kFinParts;
        return per ? (chunks + per - 1) / per : 0u;
    };

    // Groups and the GEMV partial sources that finalise their conductances.
    for (size_t g = 0; g < neuron_groups_.size(); ++g) evolve_.g[g].v = neuron_groups_[g]->gpu_view();
    evolve_.n_groups = (uint32_t)neuron_groups_.size();
    evolve_.total_neurons = S.total_neurons();
    for (size_t c = 0; c < connections_.size(); ++c) {
        const ConnectionView& v = connections_[c]->gpu_view();
        uint32_t pg = 0;
        while (pg < evolve_.n_groups && evolve_.g[pg].v.offset != v.post_group_offset) ++pg;
        if (pg == evolve_.n_groups) throw Error(BN_ERR_INVALID_ARGUMENT, "connection post group not found");
        EvolveGroup& G = evolve_.g[pg];
        const uint32_t chunks = (v.pre_group_num_neurons + v.gemv_row_block - 1) / v.gemv_row_block;
        // (part sums: one row per part, each combined as a part of one chunk)
        const FinalizeSource src{v.gemv_partials, parts ? part_count(chunks) : chunks, v.weight_matrix_ld};
        if (v.pre_group_is_inhibtory) {
            if (G.n_fin_i >= (uint32_t)kMaxSources) throw Error(BN_ERR_INVALID_ARGUMENT, "too many inhibitory sources");
            G.fin_i[G.n_fin_i++] = src;
        } else {
            if (G.n_fin_e >= (uint32_t)kMaxSources) throw Error(BN_ERR_INVALID_ARGUMENT, "too many excitatory sources");
            G.fin_e[G.n_fin_e++] = src;
        }
        // Connections with the same pre group and trace constants share one trace set.
        int ts = -1;
        for (uint32_t t = 0; t < evolve_.n_ts; ++t) {
            const TraceSet& T = evolve_.ts[t];
            if (T.pre_offset == v.pre_group_offset && T.n_pre == v.pre_group_num_neurons &&
                T.decay_r == v.postsynaptic_r_trace_decay && T.decay_d == v.postsynaptic_d_trace_decay &&
                T.K == v.postsynaptic_kernel_normalization_value)
                ts = (int)t;
        }
        if (ts < 0) {
            if (evolve_.n_ts >= (uint32_t)kMaxTraceSets) throw Error(BN_ERR_INVALID_ARGUMENT, "too many trace sets");
            ts = (int)evolve_.n_ts++;
            evolve_.ts[ts] = TraceSet{v.pre_group_offset,         v.pre_group_num_neurons,
                                      v.postsynaptic_r_trace_decay, v.postsynaptic_d_trace_decay,
                                      v.postsynaptic_kernel_normalization_value,
                                      v.postsynaptic_r_traces,      v.postsynaptic_d_traces,
                                      v.summed_postsynaptic_conductances};
        }
        conn_trace_set_[c] = ts;
    }
    // Plasticity trace decays and the plasticity plan.
    int lazy_idx = -1;
    for (size_t k = 0; k < plasticities_.size(); ++k) {
        const Plasticity& P = *plasticities_[k];
        const PlasticityView& v = P.gpu_view();
        const PlasticityParams& pp = P.params();
        PlastTraces& Q = evolve_.pt[k];
        Q.kind = P.kind();
        Q.pre_offset = v.pre_group_offset;
        Q.n_pre = v.pre_group_num_neurons;
        Q.post_offset = v.post_group_offset;
        Q.n_post = v.post_group_num_neurons;
        Q.pre_decay = pp.pre_trace_decay;
        Q.post_decay = pp.post_trace_decay;
        Q.trip_pre_decay = pp.triplet_pre_trace_decay;
        Q.trip_post_decay = pp.triplet_post_trace_decay;
        Q.pre = Q.pre_out = v.pre_traces;
        Q.post = Q.post_out = v.post_traces;
        Q.trip_pre = Q.trip_pre_out = v.triplet_pre_traces;
        Q.trip_post = Q.trip_post_out = v.triplet_post_traces;
        PlastDev& D = plast_.p[k];
        D.kind = P.kind();
        D.normalize = P.normalizes() ? 1u : 0u;
        D.p = pp;
        D.W = v.weight_matrix;
        D.ld = v.weight_matrix_ld;
        D.n_pre = v.pre_group_num_neurons;
        D.n_post = v.post_group_num_neurons;
        D.pre_offset = v.pre_group_offset;
        D.post_offset = v.post_group_offset;
        D.pre_spikes = v.pre_spikes;
        D.post_spikes = v.post_spikes;
        D.pre = v.pre_traces;
        D.post = v.post_traces;
        D.trip_pre = v.triplet_pre_traces;
        D.trip_post = v.triplet_post_traces;
        D.target = v.pre_weight_sums;
        D.bits = S.d_spike_bits();
        if (D.normalize && D.n_post > 64u * kNormThreads)
            throw Error(BN_ERR_INVALID_ARGUMENT, "post group larger than the normalisation kernel supports");
        if (D.normalize && S.options().lazy_normalize && lazy_idx < 0) {
            if (v.weight_matrix_idx != 0 || D.n_pre != S.num_excitatory_neurons())
                throw Error(BN_ERR_INVALID_ARGUMENT, "lazy normalisation supports the E->E matrix");
            D.lazy = 1;
            D.s = S.d_row_scale();
            D.R = S.d_row_sum();
            lazy_idx = (int)k;
            S.setup_lazy_normalization(v.weight_matrix_idx);
        }
        if (D.n_post > 64u * kNormThreads)
            throw Error(BN_ERR_INVALID_ARGUMENT, "post group larger than the plasticity kernel supports");
        plast_.items_post += D.n_post;
        plast_.items_pre += D.n_pre;
    }
    evolve_.n_pt = plast_.n = (uint32_t)plasticities_.size();
    evolve_.clock = S.d_clock();
    evolve_.spike_bits = S.d_spike_bits();
    evolve_.spike_bits_ring = S.d_spike_bits_ring();
    evolve_.v_ring = S.d_v_ring();
    evolve_.vrec_index = S.d_vrec_index();
    evolve_.n_vrec = S.n_vrec();
    evolve_.interval = S.options().recording_interval_steps;
    evolve_.words = S.spike_words();

    // GEMV over every connection, one wave per (row chunk, 256 columns) tile.
    uint32_t tiles = 0;
    for (size_t c = 0; c < connections_.size(); ++c) {
        const ConnectionView& v = connections_[c]->gpu_view();
        GemvConn& C = gemv_.c[c];
        C.W = v.weight_matrix;
        C.x = evolve_.ts[conn_trace_set_[c]].x;
        C.row_scale = nullptr;
        if (lazy_idx >= 0 && plast_.p[lazy_idx].W == v.weight_matrix) C.row_scale = S.d_row_scale();
        C.partials = v.gemv_partials;
        C.ld = v.weight_matrix_ld;
        C.n_pre = v.pre_group_num_neurons;
        C.n_post = v.post_group_num_neurons;
        C.row_block = v.gemv_row_block;
        C.n_chunks = (C.n_pre + C.row_block - 1) / C.row_block;
        C.n_coltiles = (C.n_post + kGemvTileCols - 1) / kGemvTileCols;
        C.tile_begin = tiles;
        tiles += (parts ? part_count(C.n_chunks) : C.n_chunks) * C.n_coltiles;
        C.tile_end = tiles;
    }
    gemv_.parts = parts ? 1u : 0u;
    gemv_.n_conn = (uint32_t)connections_.size();
    gemv_.total_tiles = tiles;
    // Weight sets well above the 256 MB MALL are streamed with non-temporal loads; smaller
    // ones are left to stay cache-resident from step to step.
    uint64_t w_bytes = 0;
    for (size_t c = 0; c < connections_.size(); ++c) w_bytes += 4ull * gemv_.c[c].n_pre * gemv_.c[c].ld;
    const int nt = S.options().gemv_nontemporal;
    gemv_.nontemporal = nt < 0 ? (w_bytes > (192ull << 20) ? 1u : 0u) : (nt ? 1u : 0u);

    // Stimulus events + clock advance: in the plasticity kernel, or after the faithful
    // normalisation when there is one.
    StimTail st{};
    if (!stimuli_.empty()) {
        const StimulusView& sv = stimuli_[0]->gpu_view();
        st.i_stimulus = sv.i_stimulus;
        st.timestep_event_count = sv.timestep_event_count;
        st.event_ids = sv.stimulus_event_neuron_ids;
        st.event_offsets = sv.stimulus_event_post_group_offsets;
        st.event_amplitudes = sv.stimulus_event_amplitude_changes;
        st.num_timesteps = sv.num_timesteps;
        st.num_events = sv.num_events;
    }
    st.clock = S.d_clock();
    st.status = S.d_status();
    st.active = 1;
    tail_.pl = plast_;
    tail_.norm_rows = 0;
    for (uint32_t k = 0; k < tail_.pl.n; ++k) {
        PlastDev& D = tail_.pl.p[k];
        if (D.lazy) D.normalize = 0;
        if (D.normalize) tail_.norm_rows += D.n_pre;
    }
    tail_.total_neurons = S.total_neurons();
    tail_.increments = 0;  // folded into the next step's trace decay (k_evolve)
    tail_.tail = st;
    plast_step_.pl = plast_;
    plast_step_.row_blocks = 0;  // each plasticity's rows padded to whole blocks
    for (uint32_t k = 0; k < plast_.n; ++k)
        plast_step_.row_blocks += (plast_.p[k].n_pre + kPlastQuietRows - 1) / kPlastQuietRows;
    plast_step_.spiking_blocks = std::max<uint32_t>(1, std::min<uint32_t>(256, plast_.items_pre));
    plast_step_.tail = st;
    plast_step_.tail.active = tail_.norm_rows == 0 ? 1 : 0;
    // The propagate launch compacts the step's spiking pre neurons for the plasticity launch.
    gemv_.pc = PreCompact{};
    plast_step_.pre_list = plast_step_.pre_counts = nullptr;
    if (plast_.n > 0) {
        if (!pre_list_buf_)
            BN_HIP_CHECK(hipMalloc(&pre_list_buf_, (kPreListCap + kMaxPlast) * sizeof(uint32_t)), "hipMalloc(pre list)");
        gemv_.pc.bits = S.d_spike_bits();
        gemv_.pc.n = plast_.n;
        for (uint32_t k = 0; k < plast_.n; ++k) {
            gemv_.pc.pre_offset[k] = plast_.p[k].pre_offset;
            gemv_.pc.n_pre[k] = plast_.p[k].n_pre;
        }
        gemv_.pc.list = pre_list_buf_;
        gemv_.pc.counts = pre_list_buf_ + kPreListCap;
        plast_step_.pre_list = gemv_.pc.list;
        plast_step_.pre_counts = gemv_.pc.counts;
    }

    kernels_per_step_ = 1 + (gemv_.total_tiles ? 1 : 0) + 1 + (tail_.norm_rows ? 1 : 0);
    if (lazy_idx >= 0) BN_HIP_CHECK((hipError_t)launch_lazy_init(plast_.p[lazy_idx], S.stream()), "lazy normalisation init");
}

void System::prepare_overlap(uint32_t steps_per_graph) {
    if (!fused_ || tail_.norm_rows || steps_per_graph == 0 || steps_per_graph % 2)
        throw Error(BN_ERR_INVALID_ARGUMENT, "the overlapped step needs the fused step, lazy or no normalisation "
                                             "and an even number of steps per graph");
    if (!overlap_built_) build_overlap_plans();
    plast_final_.tail.step_offset = steps_per_graph - 1;
    plast_final_.tail.advance = steps_per_graph;
}

// First event of every step (prefix sums of timestep_event_count): the stimulus events of any
// step found without the host cursor.
void System::build_event_starts() {
    if (d_ev_start_) return;
    const StimTail& st = plast_step_.tail;
    if (!st.timestep_event_count || !st.num_timesteps) return;
    std::vector<uint32_t> counts(st.num_timesteps), start(st.num_timesteps + 1, 0);
    BN_HIP_CHECK(hipMemcpy(counts.data(), st.timestep_event_count, counts.size() * sizeof(uint32_t),
                           hipMemcpyDeviceToHost), "event counts");
    for (uint32_t t = 0; t < st.num_timesteps; ++t) start[t + 1] = start[t] + counts[t];
    void* p = nullptr;
    BN_HIP_CHECK(hipMalloc(&p, start.size() * sizeof(uint32_t)), "hipMalloc event starts");
    staging_.push_back(p);
    d_ev_start_ = static_cast<uint32_t*>(p);
    BN_HIP_CHECK(hipMemcpy(d_ev_start_, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
                 "event starts");
}

void System::build_overlap_plans() {
    hipStream_t s = state_.stream();
    overlap_built_ = true;
    evolve_even_ = evolve_odd_ = evolve_;
    plast_staging_ = plast_current_ = plast_step_;
    auto alloc = [&](size_t bytes) {
        void* p = nullptr;
        BN_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 4)), "hipMalloc staging");
        staging_.push_back(p);
        return p;
    };
    // Staging twins of the plasticity traces: even steps write them, odd steps read them.
    for (uint32_t k = 0; k < evolve_.n_pt; ++k) {
        const PlastTraces& Q = evolve_.pt[k];
        PlastTraces& E = evolve_even_.pt[k];
        PlastTraces& O = evolve_odd_.pt[k];
        PlastDev& D = plast_staging_.pl.p[k];
        auto twin = [&](float* cur, uint32_t n, float*& even_out, float*& odd_in, float*& plast_in) {
            if (!cur || n == 0) return;
            float* st = static_cast<float*>(alloc((size_t)n * sizeof(float)));
            even_out = st;
            odd_in = st;
            plast_in = st;
        };
        twin(Q.pre, Q.n_pre, E.pre_out, O.pre, D.pre);
        twin(Q.post, Q.n_post, E.post_out, O.post, D.post);
        if (Q.kind == kPlastTriplet) {
            twin(Q.trip_pre, Q.n_pre, E.trip_pre_out, O.trip_pre, D.trip_pre);
            twin(Q.trip_post, Q.n_post, E.trip_post_out, O.trip_post, D.trip_post);
        }
    }
    uint32_t* bits = static_cast<uint32_t*>(alloc((size_t)std::max<uint32_t>(1, evolve_.words) * sizeof(uint32_t)));
    evolve_even_.spike_bits = bits;
    for (uint32_t k = 0; k < plast_staging_.pl.n; ++k) plast_staging_.pl.p[k].bits = bits;
    // Even steps' spikes are in the staging bits: their propagate compacts those.
    gemv_even_ = gemv_odd_ = gemv_;
    if (gemv_even_.pc.list) gemv_even_.pc.bits = bits;
    plast_staging_.tail.active = plast_current_.tail.active = 0;
    // The stimulus: event starts per step for the per-neuron application and the final tail.
    plast_final_ = plast_current_;
    plast_final_.tail = plast_step_.tail;
    plast_final_.tail.active = 1;
    const StimTail& st = plast_step_.tail;
    build_event_starts();
    if (d_ev_start_) {
        for (EvolvePlan* e : {&evolve_even_, &evolve_odd_}) {
            e->ev_start = d_ev_start_;
            e->ev_ids = st.event_ids;
            e->ev_offsets = st.event_offsets;
            e->ev_amps = st.event_amplitudes;
            e->ev_steps = st.num_timesteps;
            e->ev_count = st.num_events;
        }
        plast_final_.tail.ev_start = d_ev_start_;
    }
    BN_HIP_CHECK(hipStreamSynchronize(s), "staging sync");
}

// A graph of n (even) steps: k_evolve, k_gemv for the first, then (evolve of t + plasticity of
// t-1, propagate of t) for the others, and the plasticity of the last step with the stimulus
// events of that step and the clock advance of the whole graph.
void System::enqueue_overlapped_steps(uint32_t n) {
    hipStream_t s = state_.stream();
    for (uint32_t k = 0; k < n; ++k) {
        EvolvePlan e = k % 2 == 0 ? evolve_even_ : evolve_odd_;
        e.step_offset = k;
        e.stim_prev = k > 0 && e.ev_start ? 1u : 0u;
        const PlastStepPlan* prev = k == 0 ? nullptr : (k - 1) % 2 == 0 ? &plast_staging_ : &plast_current_;
        BN_HIP_CHECK((hipError_t)launch_evolve_plast(e, prev, s), "overlapped evolve launch");
        BN_HIP_CHECK((hipError_t)launch_gemv(k % 2 == 0 ? gemv_even_ : gemv_odd_, s), "overlapped propagate launch");
    }
    BN_HIP_CHECK((hipError_t)launch_plasticity(plast_final_, s), "overlapped plasticity launch");
}

void System::enqueue_fused_step(hipEvent_t gemv_begin, hipEvent_t gemv_end) {
    hipStream_t s = state_.stream();
    BN_HIP_CHECK((hipError_t)launch_evolve(evolve_, s), "fused evolve launch");
    if (gemv_begin) BN_HIP_CHECK(hipEventRecord(gemv_begin, s), "event record");
    BN_HIP_CHECK((hipError_t)launch_gemv(gemv_, s), "fused gemv launch");
    if (gemv_end) BN_HIP_CHECK(hipEventRecord(gemv_end, s), "event record");
    BN_HIP_CHECK((hipError_t)launch_plasticity(plast_step_, s), "fused plasticity launch");
    if (tail_.norm_rows) BN_HIP_CHECK((hipError_t)launch_norm_tail(tail_, s), "fused normalisation launch");
}

void System::enable_fused(bool use_graph, uint32_t steps_per_graph) {
    for (auto& cb : step_callbacks_) (void)cb;
    build_plans();
    fused_ = true;
    for (auto& m : monitors_) m->fused = true;
    use_graph_ = use_graph;
    steps_per_graph_ = std::max<uint32_t>(1, steps_per_graph);
    capture_graph();
}

void System::capture_graph() {
    if (graph_exec_) BN_HIP_CHECK(hipGraphExecDestroy(graph_exec_), "hipGraphExecDestroy");
    if (graph_) BN_HIP_CHECK(hipGraphDestroy(graph_), "hipGraphDestroy");
    graph_exec_ = nullptr;
    graph_ = nullptr;
    kernels_per_step_ = 1 + (gemv_.total_tiles ? 1 : 0) + 1 + (tail_.norm_rows ? 1 : 0);
    overlap_ = use_graph_ && state_.options().overlap_steps && tail_.norm_rows == 0 && steps_per_graph_ % 2 == 0;
    if (overlap_) {
        prepare_overlap(steps_per_graph_);
        kernels_per_step_ = 2;
    }
    if (use_graph_) {
        hipStream_t s = state_.stream();
        BN_HIP_CHECK(hipStreamSynchronize(s), "pre-capture sync");
        BN_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        if (overlap_) enqueue_overlapped_steps(steps_per_graph_);
        else
            for (uint32_t k = 0; k < steps_per_graph_; ++k) enqueue_fused_step(nullptr, nullptr);
        BN_HIP_CHECK(hipStreamEndCapture(s, &graph_), "hipStreamEndCapture");
        BN_HIP_CHECK(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
}

void System::set_steps_per_graph(uint32_t steps_per_graph) {
    if (!fused_) throw Error(BN_ERR_RUNTIME, "steps per graph: the component path has no graph");
    if (steps_per_graph == 0) throw Error(BN_ERR_INVALID_ARGUMENT, "steps per graph must be at least 1");
    if (persist_ || resident_) {  // these steps launch whole runs, no graphs
        steps_per_graph_ = steps_per_graph;
        return;
    }
    BN_HIP_CHECK(hipStreamSynchronize(state_.stream()), "pre-capture sync");
    steps_per_graph_ = steps_per_graph;
    capture_graph();
}

void System::launch_fused_steps(uint32_t n) {
    if (persist_) {
        launch_persistent_steps(n);
        return;
    }
    if (resident_) {
        launch_resident_steps(n);
        return;
    }
    hipStream_t s = state_.stream();
    uint32_t k = 0;
    if (use_graph_)
        for (; k + steps_per_graph_ <= n; k += steps_per_graph_) {
            BN_HIP_CHECK(hipGraphLaunch(graph_exec_, s), "hipGraphLaunch");
            ++graph_launches_;
        }
    for (; k < n; ++k) enqueue_fused_step(nullptr, nullptr);
    clock_step_ += n;
}

void System::launch_fused_step_eager(hipEvent_t gemv_begin, hipEvent_t gemv_end) {
    enqueue_fused_step(gemv_begin, gemv_end);
    clock_step_ += 1;
}

void System::materialize_g(float* d_ge_out, float* d_gi_out) {
    hipStream_t s = state_.stream();
    if (resident_) {
        // The resident step stores g after the propagate itself: no pending partials.
        for (uint32_t g = 0; g < 2; ++g) {
            const NeuronGroupView& v = resident_plan_.g[g];
            BN_HIP_CHECK((hipError_t)launch_materialize_g(d_ge_out + v.offset, v.g_excitatory, nullptr, 0,
                                                          v.num_neurons, s),
                         "materialize g_e (resident)");
            BN_HIP_CHECK((hipError_t)launch_materialize_g(d_gi_out + v.offset, v.g_inhibitory, nullptr, 0,
                                                          v.num_neurons, s),
                         "materialize g_i (resident)");
        }
        return;
    }
    if (persist_) {
        // Pending segment partials of the last persistent step: g + P_0 + P_1 + ... in order.
        const PPlan& P = persist_plan_;
        for (uint32_t g = 0; g < P.n_groups; ++g) {
            const NeuronGroupView& v = P.g[g].v;
            for (uint32_t side = 0; side < 2; ++side) {
                const uint32_t c = side == 0 ? P.g[g].src_e : P.g[g].src_i;
                FinalizeSource src{};
                uint32_t n_src = 0;
                if (c != kPNone) {
                    src = FinalizeSource{P.c[c].partials, P.c[c].nseg, P.c[c].ld, 1u};
                    n_src = 1;
                }
                BN_HIP_CHECK((hipError_t)launch_materialize_g((side == 0 ? d_ge_out : d_gi_out) + v.offset,
                                                              side == 0 ? v.g_excitatory : v.g_inhibitory, &src, n_src,
                                                              v.num_neurons, s),
                             "materialize g (persistent)");
            }
        }
        return;
    }
    for (uint32_t g = 0; g < evolve_.n_groups || (!fused_ && g < neuron_groups_.size()); ++g) {
        const NeuronGroupView& v = fused_ ? evolve_.g[g].v : neuron_groups_[g]->gpu_view();
        const uint32_t ne = fused_ ? evolve_.g[g].n_fin_e : 0, ni = fused_ ? evolve_.g[g].n_fin_i : 0;
        BN_HIP_CHECK((hipError_t)launch_materialize_g(d_ge_out + v.offset, v.g_excitatory,
                                                      fused_ ? evolve_.g[g].fin_e : nullptr, ne, v.num_neurons, s),
                     "materialize g_e");
        BN_HIP_CHECK((hipError_t)launch_materialize_g(d_gi_out + v.offset, v.g_inhibitory,
                                                      fused_ ? evolve_.g[g].fin_i : nullptr, ni, v.num_neurons, s),
                     "materialize g_i");
    }
}

}  // namespace bn
