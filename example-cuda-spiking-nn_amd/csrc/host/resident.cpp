// The resident step's plan (bn_resident.h) from the fused step's plans, and its launches.
#include <algorithm>
#include <string>

#include "../bn_errors.h"
#include "engine.h"

namespace bn {

void System::enable_resident() {
    if (!fused_) throw Error(BN_ERR_RUNTIME, "the resident step needs the fused step (not the component path)");
    if (persist_) throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step and the persistent step exclude each other");
    if (tail_.norm_rows)
        throw Error(BN_ERR_INVALID_ARGUMENT,
                    "the resident step keeps E->E normalisation lazily (BN_SIM_LAZY_NORMALIZE); the rescale of every "
                    "row every step is not supported there");
    GpuSimulationState& S = state_;
    BN_HIP_CHECK(hipStreamSynchronize(S.stream()), "resident plan sync");
    build_event_starts();

    RPlan P{};
    // Groups: E (global ids from 0) and I.
    if (evolve_.n_groups != 2 || evolve_.g[0].v.offset != 0 || evolve_.g[1].v.offset != evolve_.g[0].v.num_neurons)
        throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step takes one excitatory and one inhibitory group");
    P.g[0] = evolve_.g[0].v;
    P.g[1] = evolve_.g[1].v;
    P.ne = P.g[0].num_neurons;
    P.ni = P.g[1].num_neurons;
    P.n = P.ne + P.ni;
    if (P.ne == 0 || P.ni == 0) throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step needs both groups");
    // Connections EE, EI, IE, II (by their pre and post groups).
    if (gemv_.n_conn != 4) throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step takes the four connections");
    int slot_of[4] = {-1, -1, -1, -1};
    for (uint32_t c = 0; c < 4; ++c) {
        const ConnectionView& v = connections_[c]->gpu_view();
        const uint32_t pre = v.pre_group_offset == 0 ? 0u : 1u, post = v.post_group_offset == 0 ? 0u : 1u;
        if ((pre == 1u) != (bool)v.pre_group_is_inhibtory)
            throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step needs inhibitory connections from the I group");
        const int k = (int)(pre * 2 + post);  // EE 0, EI 1, IE 2, II 3
        if (slot_of[k] >= 0) throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step takes one connection per group pair");
        slot_of[k] = (int)c;
        const GemvConn& G = gemv_.c[c];
        const uint32_t padded = (G.n_post + kRCols - 1) / kRCols * kRCols;
        if (G.ld % 4 != 0 || G.ld < padded || (uintptr_t)G.W % 16 != 0)
            throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step needs padded weight rows (not BN_SIM_REFERENCE_LAYOUT)");
        P.W[k] = const_cast<float*>(G.W);
        P.ld[k] = G.ld;
    }
    // Connection traces: one trace set per pre group.
    for (uint32_t side = 0; side < 2; ++side) {
        const int ta = conn_trace_set_[slot_of[side * 2]], tb = conn_trace_set_[slot_of[side * 2 + 1]];
        if (ta != tb)
            throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step needs equal trace constants for the connections of a pre group");
        const TraceSet& T = evolve_.ts[ta];
        P.decay_r[side] = T.decay_r;
        P.decay_d[side] = T.decay_d;
        P.K[side] = T.K;
        P.r[side] = T.r;
        P.d[side] = T.d;
        P.x[side] = T.x;
    }
    // Plasticities: E->E (triplet, lazily normalised, or pair) and I->E (iSTDP).
    P.ee_kind = P.ie_kind = kRNone;
    for (uint32_t k = 0; k < plast_.n; ++k) {
        const PlastDev& D = plast_.p[k];
        const PlastTraces& Q = evolve_.pt[k];
        if (D.W == P.W[0] && (D.kind == kPlastTriplet || D.kind == kPlastStdp)) {
            if (D.kind == kPlastTriplet && !D.lazy)
                throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step needs BN_SIM_LAZY_NORMALIZE for the triplet rule");
            P.ee_kind = D.kind;
            P.ee_lazy = D.lazy;
            P.ee_p = D.p;
            P.ee_dec[0] = Q.pre_decay;
            P.ee_dec[1] = Q.trip_pre_decay;
            P.ee_dec[2] = Q.post_decay;
            P.ee_dec[3] = Q.trip_post_decay;
            P.ee_tr[0] = D.pre;
            P.ee_tr[1] = D.trip_pre;
            P.ee_tr[2] = D.post;
            P.ee_tr[3] = D.trip_post;
            if (D.lazy) {
                P.s = D.s;
                P.R = D.R;
                P.target = D.target;
            }
        } else if (D.W == P.W[2] && D.kind == kPlastIstdp) {
            P.ie_kind = D.kind;
            P.ie_p = D.p;
            P.ie_dec[0] = Q.pre_decay;
            P.ie_dec[1] = Q.post_decay;
            P.ie_tr[0] = D.pre;
            P.ie_tr[1] = D.post;
        } else {
            throw Error(BN_ERR_INVALID_ARGUMENT, "the resident step takes plasticity on E->E and I->E only");
        }
    }
    // Shape: rows per thread, waves of E rows, blocks.
    resident_shape(P.ne, P.ni, &P.rows, &P.e_waves);
    P.be = (P.ne + kRCols - 1) / kRCols;
    P.bi = (P.ni + kRCols - 1) / kRCols;
    P.blocks = P.be + P.bi;
    uint32_t max_blocks = 0;
    if (P.rows) BN_HIP_CHECK((hipError_t)resident_max_blocks(P.rows, &max_blocks), "resident step occupancy");
    if (!P.rows || P.blocks > std::min(max_blocks, kRMaxBlocks))
        throw Error(BN_ERR_INVALID_ARGUMENT,
                    "the network does not fit the resident step (" + std::to_string(P.blocks) + " blocks of " +
                        std::to_string(kRCols) + " columns, at most " + std::to_string(std::min(max_blocks, kRMaxBlocks)) +
                        " co-resident; at most " + std::to_string(kRMaxRows) + " rows per thread)");
    // State.
    P.clock = S.d_clock();
    P.spikes = S.d_spikes();
    P.bits = S.d_spike_bits();
    P.words = S.spike_words();
    P.spike_ring = S.d_spike_bits_ring();
    P.v_ring = S.d_v_ring();
    P.vrec_index = S.d_vrec_index();
    P.n_vrec = S.n_vrec();
    P.interval = S.options().recording_interval_steps;
    const StimTail& T = plast_step_.tail;
    P.ev_start = d_ev_start_;
    P.ev_ids = T.event_ids;
    P.ev_offsets = T.event_offsets;
    P.ev_amps = T.event_amplitudes;
    P.ev_steps = d_ev_start_ ? T.num_timesteps : 0;
    P.ev_count = T.num_events;
    P.status = S.d_status();
    P.wait_limit = persist_wait_limit_;
    // Exchange buffers (zero: no tag).
    resident_bufs_.clear();
    auto words = [&](size_t count) {
        uint64_t* p = static_cast<uint64_t*>(persist_alloc(count * sizeof(uint64_t)));
        resident_bufs_.push_back({p, count});
        return p;
    };
    P.spk = words((size_t)3 * P.blocks);
    P.sx = words((size_t)3 * P.ne);
    P.dr = P.ee_lazy ? words((size_t)3 * P.be * P.ne * 2) : nullptr;
    P.drs = P.ee_lazy ? words((size_t)3 * P.ne * P.be * 2) : nullptr;
    d_resident_plan_ = static_cast<RPlan*>(persist_alloc(sizeof(RPlan)));
    BN_HIP_CHECK(hipMemcpy(d_resident_plan_, &P, sizeof(RPlan), hipMemcpyHostToDevice), "resident plan");
    resident_plan_ = P;
    resident_ = true;
    resident_next_ = 0;
    kernels_per_step_ = 1;
    BN_HIP_CHECK(hipStreamSynchronize(S.stream()), "resident plan sync");
}

void System::set_resident_cooperative(bool cooperative) {
    if (!resident_) throw Error(BN_ERR_RUNTIME, "not a resident simulation");
    resident_cooperative_ = cooperative;
}

// One launch per run (its tagged words carry their steps, so consecutive launches
// need no clearing); after a state restore moved the clock back, the words are cleared first.
void System::launch_resident_steps(uint32_t n) {
    if (n == 0) return;
    hipStream_t s = state_.stream();
    if (clock_step_ < resident_next_)
        for (const auto& b : resident_bufs_)
            BN_HIP_CHECK((hipError_t)launch_fill_u32(static_cast<uint32_t*>(b.first), 0u, b.second * 2, s),
                         "resident tagged words");
    BN_HIP_CHECK((hipError_t)launch_resident(d_resident_plan_, resident_plan_, n, resident_cooperative_, s),
                 "resident step launch");
    clock_step_ += n;
    resident_next_ = clock_step_;
}

}  // namespace bn
