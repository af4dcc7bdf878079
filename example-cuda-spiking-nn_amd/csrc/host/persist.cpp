// Host side of the persistent step (bn_persist.h): the device plan, the schedule of items over
// co-resident blocks, and the launches.
#include <algorithm>
#include <functional>
#include <queue>
#include <string>
#include <vector>

#include "../bn_errors.h"
#include "engine.h"

namespace bn {

namespace {

uint32_t cdiv32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

struct Sched {
    std::vector<PItem> items;          // ordered by block
    std::vector<uint32_t> begin;       // [blocks + 1]
    uint64_t max_bytes = 0;
    uint32_t max_items = 0;
};

// Items of every connection for row chunks of rb[c] rows, spread over at most max_blocks blocks:
// largest first onto the least loaded block (bytes of weights streamed).
Sched schedule(const PPlan& P, uint32_t max_blocks) {
    struct It {
        PItem it;
        uint64_t bytes;
    };
    std::vector<It> all;
    for (uint32_t c = 0; c < P.n_conn; ++c) {
        const PConn& C = P.c[c];
        const uint32_t seg_rows = kPWaves * C.rb;
        for (uint32_t seg = 0; seg < C.nseg; ++seg)
            for (uint32_t q = 0; q < C.ncol; ++q) {
                const uint64_t rows = std::min<uint64_t>(seg_rows, C.n_pre - (uint64_t)seg * seg_rows);
                const uint64_t cols = std::min<uint64_t>(kPCols, C.n_post - (uint64_t)q * kPCols);
                all.push_back({PItem{(uint16_t)c, (uint16_t)q, seg}, rows * cols * 4});
            }
    }
    std::stable_sort(all.begin(), all.end(), [](const It& a, const It& b) { return a.bytes > b.bytes; });
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(max_blocks, (uint32_t)all.size()));
    using Load = std::pair<uint64_t, uint32_t>;  // (bytes, block), least loaded first, lowest index on ties
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (uint32_t b = 0; b < blocks; ++b) heap.push({0, b});
    std::vector<std::vector<PItem>> per(blocks);
    std::vector<uint64_t> load(blocks, 0);
    for (const It& x : all) {
        const Load l = heap.top();
        heap.pop();
        per[l.second].push_back(x.it);
        load[l.second] = l.first + x.bytes;
        heap.push({load[l.second], l.second});
    }
    Sched s;
    s.begin.push_back(0);
    for (uint32_t b = 0; b < blocks; ++b) {
        s.items.insert(s.items.end(), per[b].begin(), per[b].end());
        s.begin.push_back((uint32_t)s.items.size());
        s.max_bytes = std::max(s.max_bytes, load[b]);
        s.max_items = std::max(s.max_items, (uint32_t)per[b].size());
    }
    return s;
}

}  // namespace

void* System::persist_alloc(size_t bytes) {
    void* p = nullptr;
    BN_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)), "hipMalloc persistent step");
    persist_bufs_.push_back(p);
    BN_HIP_CHECK(hipMemsetAsync(p, 0, std::max<size_t>(bytes, 16), state_.stream()), "hipMemset persistent step");
    // The zeroing has finished before the caller fills the buffer: its plans, item tables and
    // block maps go up by hipMemcpy on the null stream, which does not wait for the state's
    // stream (created non-blocking), so a zeroing still queued there could land after them.
    BN_HIP_CHECK(hipStreamSynchronize(state_.stream()), "hipMemset persistent step sync");
    return p;
}

void System::set_persistent_resident(bool allow) {
    if (!persist_) throw Error(BN_ERR_RUNTIME, "not a persistent simulation");
    persist_resident_ = allow;
    enable_persistent(persist_rows_);
}

void System::enable_persistent(const std::array<uint32_t, 4>& rows) {
    if (!fused_) throw Error(BN_ERR_RUNTIME, "the persistent step needs the fused step (not the component path)");
    if (tail_.norm_rows)
        throw Error(BN_ERR_INVALID_ARGUMENT,
                    "the persistent step keeps E->E normalisation lazily (BN_SIM_LAZY_NORMALIZE); the rescale of "
                    "every row every step is not supported there");
    if (stimuli_.size() > 1) throw Error(BN_ERR_INVALID_ARGUMENT, "the persistent step supports one stimulus");
    GpuSimulationState& S = state_;
    hipStream_t st = S.stream();
    BN_HIP_CHECK(hipStreamSynchronize(st), "persistent plan sync");
    for (void* p : persist_bufs_) (void)hipFree(p);
    persist_bufs_.clear();
    d_persist_plan_ = nullptr;
    persist_counters_ = nullptr;
    build_event_starts();

    PPlan P{};
    // Groups (their views as the fused evolve has them).
    P.n_groups = evolve_.n_groups;
    for (uint32_t g = 0; g < P.n_groups; ++g) {
        P.g[g].v = evolve_.g[g].v;
        P.g[g].src_e = P.g[g].src_i = kPNone;
    }
    // Trace sets with their staging x.
    P.n_ts = evolve_.n_ts;
    for (uint32_t t = 0; t < P.n_ts; ++t) {
        const TraceSet& T = evolve_.ts[t];
        PTraceSet& Q = P.ts[t];
        Q.pre_off = T.pre_offset;
        Q.n_pre = T.n_pre;
        Q.decay_r = T.decay_r;
        Q.decay_d = T.decay_d;
        Q.K = T.K;
        Q.r = T.r;
        Q.d = T.d;
        Q.x[0] = T.x;
        Q.x[1] = static_cast<float*>(persist_alloc((size_t)T.n_pre * sizeof(float)));
    }
    // Plasticities with staging traces.
    P.n_pl = plast_.n;
    P.lazy_pl = kPNone;
    for (uint32_t k = 0; k < P.n_pl; ++k) {
        const PlastDev& D = plast_.p[k];
        const PlastTraces& Q = evolve_.pt[k];
        PPlast& L = P.pl[k];
        L.kind = D.kind;
        L.lazy = D.lazy;
        L.p = D.p;
        L.conn = kPNone;
        for (uint32_t c = 0; c < gemv_.n_conn; ++c)
            if (gemv_.c[c].W == D.W) L.conn = c;
        if (L.conn == kPNone) throw Error(BN_ERR_INVALID_ARGUMENT, "plasticity without its connection");
        L.pre_off = D.pre_offset;
        L.n_pre = D.n_pre;
        L.post_off = D.post_offset;
        L.n_post = D.n_post;
        L.pre[0] = D.pre;
        L.post[0] = D.post;
        L.pre[1] = static_cast<float*>(persist_alloc((size_t)D.n_pre * sizeof(float)));
        L.post[1] = static_cast<float*>(persist_alloc((size_t)D.n_post * sizeof(float)));
        if (D.kind == kPlastTriplet) {
            L.tpre[0] = D.trip_pre;
            L.tpost[0] = D.trip_post;
            L.tpre[1] = static_cast<float*>(persist_alloc((size_t)D.n_pre * sizeof(float)));
            L.tpost[1] = static_cast<float*>(persist_alloc((size_t)D.n_post * sizeof(float)));
        }
        L.pre_decay = Q.pre_decay;
        L.post_decay = Q.post_decay;
        L.tpre_decay = Q.trip_pre_decay;
        L.tpost_decay = Q.trip_post_decay;
        if (D.lazy) {
            L.s = D.s;
            L.R = D.R;
            L.target = D.target;
            P.lazy_pl = k;
        }
    }
    // Connections.
    P.n_conn = gemv_.n_conn;
    for (uint32_t c = 0; c < P.n_conn; ++c) {
        const GemvConn& G = gemv_.c[c];
        const ConnectionView& v = connections_[c]->gpu_view();
        PConn& C = P.c[c];
        if (G.ld % 4 != 0 || (uintptr_t)G.W % 16 != 0)
            throw Error(BN_ERR_INVALID_ARGUMENT, "the persistent step needs weight rows padded to 16 bytes");
        C.W = const_cast<float*>(G.W);
        const int ts = conn_trace_set_[c];
        C.x[0] = P.ts[ts].x[0];
        C.x[1] = P.ts[ts].x[1];
        C.s = G.row_scale;
        C.partials = G.partials;
        C.ld = G.ld;
        C.n_pre = G.n_pre;
        C.n_post = G.n_post;
        C.pre_off = v.pre_group_offset;
        C.post_off = v.post_group_offset;
        C.ncol = cdiv32(C.n_post, kPCols);
        uint32_t pg = 0;
        while (pg < P.n_groups && P.g[pg].v.offset != v.post_group_offset) ++pg;
        if (pg == P.n_groups) throw Error(BN_ERR_INVALID_ARGUMENT, "connection post group not found");
        C.post_group = pg;
        uint32_t qg = 0;
        while (qg < P.n_groups && P.g[qg].v.offset != v.pre_group_offset) ++qg;
        if (qg == P.n_groups) throw Error(BN_ERR_INVALID_ARGUMENT, "connection pre group not found");
        C.pre_group = qg;
        C.ts = (uint32_t)ts;
        uint32_t& src = v.pre_group_is_inhibtory ? P.g[pg].src_i : P.g[pg].src_e;
        if (src != kPNone) throw Error(BN_ERR_INVALID_ARGUMENT, "the persistent step takes one connection per conductance");
        src = c;
        C.plast = kPNone;
        for (uint32_t k = 0; k < P.n_pl; ++k)
            if (P.pl[k].conn == c) C.plast = k;
    }
    // Rows per chunk.  Resident weights (unless switched off) when the items fit one per
    // co-resident block with chunks of at most 32 rows: the smallest such height, so the most
    // blocks.  Otherwise streamed weights: the requested height, or the candidate whose schedule
    // streams the fewest bytes on the busiest block (plus a per-item and a per-segment
    // allowance).  A segment's partials must fit the state's partial buffers
    // (ceil(pre / gemv_row_block) rows).
    auto min_rows = [&](uint32_t c) { return cdiv32(S.gemv_row_block(c), kPWaves); };
    const bool fixed = rows[0] || rows[1] || rows[2] || rows[3];
    auto set_rows = [&](uint32_t cand) {
        uint32_t top = 0;
        for (uint32_t c = 0; c < P.n_conn; ++c) {
            const uint32_t want = fixed ? (rows[c] ? rows[c] : rows[0]) : cand;
            P.c[c].rb = std::max(std::max(want, min_rows(c)), 1u);
            P.c[c].nseg = cdiv32(P.c[c].n_pre, (uint64_t)kPWaves * P.c[c].rb);
            top = std::max(top, P.c[c].rb);
        }
        return top;
    };
    auto item_count = [&] {
        uint64_t items = 0;
        for (uint32_t c = 0; c < P.n_conn; ++c) items += (uint64_t)P.c[c].ncol * P.c[c].nseg;
        return items;
    };
    // The resident kernel's limits: a group's two sources have at most kLPartMax segments, and
    // the lazy matrix's tile sums of a segment fit the polls of one block.
    auto resident_fits = [&](uint32_t res, uint32_t res_blocks, uint32_t top) {
        if (top > res || item_count() > res_blocks) return false;
        for (uint32_t g = 0; g < P.n_groups; ++g) {
            uint32_t nall = 0;
            for (uint32_t c = 0; c < P.n_conn; ++c)
                if (P.c[c].post_group == g) nall += P.c[c].nseg;
            if (nall > kLPartMax) return false;
        }
        if (P.lazy_pl != kPNone) {
            const PConn& C = P.c[P.pl[P.lazy_pl].conn];
            if ((uint64_t)C.ncol * kPWaves * C.rb > (uint64_t)kLDrJobs * kPThreads) return false;
        }
        return true;
    };
    const uint32_t cands[] = {4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128};
    P.resident_rows = 0;
    uint32_t max_blocks = 0;
    if (persist_resident_) {
        for (uint32_t res : kPResidentRows) {
            uint32_t res_blocks = 0;
            BN_HIP_CHECK((hipError_t)persistent_max_blocks(res, &res_blocks), "persistent step occupancy");
            for (uint32_t cand : cands) {
                if (!fixed && cand > res) break;
                const uint32_t top = set_rows(cand);
                if (resident_fits(res, res_blocks, top)) {
                    P.resident_rows = res;
                    max_blocks = res_blocks;
                    break;
                }
                if (fixed) break;
            }
            if (P.resident_rows) break;
        }
    }
    if (!P.resident_rows) {
        BN_HIP_CHECK((hipError_t)persistent_max_blocks(0, &max_blocks), "persistent step occupancy");
        if (max_blocks == 0) throw Error(BN_ERR_RUNTIME, "the persistent step cannot be resident on this device");
        uint64_t best_cost = ~0ull;
        uint32_t best_rb = cands[0];
        for (uint32_t cand : cands) {
            set_rows(cand);
            const Sched sc = schedule(P, max_blocks);
            uint32_t seg_e = 0;
            for (uint32_t g = 0; g < P.n_groups; ++g) {
                uint32_t n = 0;
                for (uint32_t side = 0; side < 2; ++side) {
                    const uint32_t c = side == 0 ? P.g[g].src_e : P.g[g].src_i;
                    if (c != kPNone) n += P.c[c].nseg;
                }
                seg_e = std::max(seg_e, n);
            }
            const uint64_t cost = sc.max_bytes + 24576ull * sc.max_items + 8192ull * seg_e;
            if (cost < best_cost) {
                best_cost = cost;
                best_rb = cand;
            }
            if (fixed) break;
        }
        set_rows(best_rb);
    }
    for (uint32_t c = 0; c < P.n_conn; ++c) {
        if (P.c[c].rb > kPMaxRows)
            throw Error(BN_ERR_INVALID_ARGUMENT, "persistent step: at most " + std::to_string(kPMaxRows) +
                                                     " rows per chunk");
        if ((uint64_t)P.c[c].nseg > cdiv32(P.c[c].n_pre, S.gemv_row_block(c)))
            throw Error(BN_ERR_INVALID_ARGUMENT, "persistent step: segment partials exceed the partial buffers");
    }
    // Column tiles per group, items per column tile.
    P.n_coltiles = 0;
    for (uint32_t g = 0; g < P.n_groups; ++g) {
        PGroup& G = P.g[g];
        G.ncol = cdiv32(G.v.num_neurons, kPCols);
        G.col_base = P.n_coltiles;
        P.n_coltiles += G.ncol;
        G.expect = 0;
        for (uint32_t side = 0; side < 2; ++side) {
            const uint32_t c = side == 0 ? G.src_e : G.src_i;
            if (c != kPNone) G.expect += P.c[c].nseg;
        }
        // Its evolve runs when its column tiles complete: a group nothing projects to has none.
        if (G.expect == 0) throw Error(BN_ERR_INVALID_ARGUMENT, "the persistent step needs a connection into every group");
    }
    P.n_rowtails = 0;
    if (P.lazy_pl != kPNone) {
        PPlast& L = P.pl[P.lazy_pl];
        const PConn& C = P.c[L.conn];
        L.dR = static_cast<double*>(persist_alloc((size_t)C.ncol * C.n_pre * sizeof(double)));
        P.n_rowtails = C.nseg;
    }
    Sched sc = schedule(P, max_blocks);
    if (P.resident_rows && sc.max_items != 1) throw Error(BN_ERR_RUNTIME, "persistent step: resident plan without one item per block");
    // Resident weights with room on the device: one more block per column tile that only evolves.
    const bool own_evolvers = P.resident_rows && sc.items.size() + P.n_coltiles <= max_blocks;
    if (own_evolvers)
        for (uint32_t t = 0; t < P.n_coltiles; ++t) sc.begin.push_back((uint32_t)sc.items.size());
    P.blocks = (uint32_t)sc.begin.size() - 1;
    persist_items_ = (uint32_t)sc.items.size();
    PItem* d_items = static_cast<PItem*>(persist_alloc(sc.items.size() * sizeof(PItem)));
    uint32_t* d_begin = static_cast<uint32_t*>(persist_alloc(sc.begin.size() * sizeof(uint32_t)));
    BN_HIP_CHECK(hipMemcpy(d_items, sc.items.data(), sc.items.size() * sizeof(PItem), hipMemcpyHostToDevice), "items");
    BN_HIP_CHECK(hipMemcpy(d_begin, sc.begin.data(), sc.begin.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
                 "block ranges");
    P.items = d_items;
    P.block_begin = d_begin;
    // Counters: column tiles, row segments, finished tails; finished steps and first evolves.
    persist_counter_words_ = (size_t)P.n_coltiles + P.n_rowtails + 3;
    persist_counters_ = static_cast<uint32_t*>(persist_alloc(persist_counter_words_ * sizeof(uint32_t)));
    P.col_count = persist_counters_;
    P.row_count = persist_counters_ + P.n_coltiles;
    P.tails_done = persist_counters_ + P.n_coltiles + P.n_rowtails;
    P.done = P.tails_done + 1;
    // Resident weights: the evolver of each column tile (the block of the item (source
    // connection, tile, segment 0)) and the tagged-word buffers, three slots each.
    persist_ll_bufs_.clear();
    persist_ll_next_ = 0;
    if (P.resident_rows) {
        std::vector<uint32_t> tile_of(P.blocks, kPNone);
        const uint32_t item_blocks = (uint32_t)sc.items.size();
        if (own_evolvers)
            for (uint32_t t = 0; t < P.n_coltiles; ++t) tile_of[item_blocks + t] = t;
        for (uint32_t g = 0; g < P.n_groups && !own_evolvers; ++g) {
            const uint32_t src = P.g[g].src_e != kPNone ? P.g[g].src_e : P.g[g].src_i;
            for (uint32_t q = 0; q < P.g[g].ncol; ++q) {
                uint32_t b = 0;
                while (b < item_blocks && !(sc.items[sc.begin[b]].conn == src && sc.items[sc.begin[b]].col == q &&
                                            sc.items[sc.begin[b]].seg == 0))
                    ++b;
                if (b == item_blocks) throw Error(BN_ERR_RUNTIME, "persistent step: no block for a column tile's evolve");
                tile_of[b] = P.g[g].col_base + q;
            }
        }
        uint32_t* d_tile = static_cast<uint32_t*>(persist_alloc(tile_of.size() * sizeof(uint32_t)));
        BN_HIP_CHECK(hipMemcpy(d_tile, tile_of.data(), tile_of.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
                     "evolver tiles");
        P.block_tile = d_tile;
        auto words = [&](size_t count) {
            uint64_t* p = static_cast<uint64_t*>(persist_alloc(3 * count * sizeof(uint64_t)));
            persist_ll_bufs_.push_back({p, 3 * count});
            return p;
        };
        for (uint32_t t = 0; t < P.n_ts; ++t) P.ll_x[t] = words(P.ts[t].n_pre);
        for (uint32_t k = 0; k < P.n_pl; ++k) {
            const bool trip = P.pl[k].kind == kPlastTriplet;
            P.ll_pl[k][0] = words(P.pl[k].n_pre);
            P.ll_pl[k][1] = trip ? words(P.pl[k].n_pre) : nullptr;
            P.ll_pl[k][2] = words(P.pl[k].n_post);
            P.ll_pl[k][3] = trip ? words(P.pl[k].n_post) : nullptr;
        }
        P.ll_spk = words((size_t)P.n_coltiles * 8);
        for (uint32_t c = 0; c < P.n_conn; ++c) P.ll_part[c] = words((size_t)P.c[c].nseg * P.c[c].ld * 2);
        if (P.lazy_pl != kPNone) {
            const PConn& C = P.c[P.pl[P.lazy_pl].conn];
            P.ll_dr = words((size_t)C.ncol * C.n_pre * 2);
        }
    }
    P.status = S.d_status();
    P.wait_limit = persist_wait_limit_;
    // State.
    P.total_neurons = S.total_neurons();
    P.words = S.spike_words();
    P.clock = S.d_clock();
    P.clock_rw = S.d_clock();
    P.spikes = S.d_spikes();
    P.bits[0] = S.d_spike_bits();
    P.bits[1] = static_cast<uint32_t*>(persist_alloc((size_t)std::max<uint32_t>(1, P.words) * sizeof(uint32_t)));
    P.bits[2] = static_cast<uint32_t*>(persist_alloc((size_t)std::max<uint32_t>(1, P.words) * sizeof(uint32_t)));
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
    P.i_stimulus = S.d_i_stimulus();
    d_persist_plan_ = static_cast<PPlan*>(persist_alloc(sizeof(PPlan)));
    BN_HIP_CHECK(hipMemcpy(d_persist_plan_, &P, sizeof(PPlan), hipMemcpyHostToDevice), "persistent plan");
    persist_plan_ = P;
    persist_rows_ = rows;
    persist_nontemporal_ = gemv_.nontemporal;
    persist_ = true;
    kernels_per_step_ = 1;
    BN_HIP_CHECK(hipStreamSynchronize(st), "persistent plan sync");
}

void System::set_persistent_wait_limit(uint32_t polls) {
    if (!persist_ && !resident_) throw Error(BN_ERR_RUNTIME, "not a persistent or resident simulation");
    BN_HIP_CHECK(hipStreamSynchronize(state_.stream()), "persistent plan sync");
    persist_wait_limit_ = polls;
    if (persist_) {
        persist_plan_.wait_limit = polls;
        BN_HIP_CHECK(hipMemcpy(d_persist_plan_, &persist_plan_, sizeof(PPlan), hipMemcpyHostToDevice), "persistent plan");
    } else {
        resident_plan_.wait_limit = polls;
        BN_HIP_CHECK(hipMemcpy(d_resident_plan_, &resident_plan_, sizeof(RPlan), hipMemcpyHostToDevice), "resident plan");
    }
}

// Runs of at most kPersistMaxSteps steps per launch; before each launch the counters, the spike
// bits of its first two evolves and their ring slots are cleared.
void System::launch_persistent_steps(uint32_t n) {
    constexpr uint32_t kPersistMaxSteps = 1u << 20;
    hipStream_t s = state_.stream();
    const PPlan& P = persist_plan_;
    while (n > 0) {
        uint32_t m = std::min(n, kPersistMaxSteps);
        BN_HIP_CHECK((hipError_t)launch_fill_u32(persist_counters_, 0u, persist_counter_words_, s), "persistent counters");
        if (P.resident_rows) {
            // Resident weights: the launch's ring slots and the state's spike bits are cleared
            // here (the evolves only set bits), so a launch covers at most one ring of slots; a
            // clock that went back (a restored state) clears the tagged words, whose tags are steps.
            if (P.spike_ring) m = std::min(m, P.interval);
            if (clock_step_ < persist_ll_next_)
                for (const auto& b : persist_ll_bufs_)
                    BN_HIP_CHECK((hipError_t)launch_fill_u32(static_cast<uint32_t*>(b.first), 0u, b.second * 2, s),
                                 "persistent tagged words");
            BN_HIP_CHECK((hipError_t)launch_fill_u32(P.bits[0], 0u, P.words, s), "persistent spike bits");
            if (P.spike_ring)
                for (uint32_t k = 0; k < m; ++k)
                    BN_HIP_CHECK((hipError_t)launch_fill_u32(P.spike_ring + (size_t)((clock_step_ + k) % P.interval) * P.words,
                                                             0u, P.words, s),
                                 "persistent ring slot");
            BN_HIP_CHECK((hipError_t)launch_persistent(d_persist_plan_, P, m, persist_nontemporal_, s), "persistent step launch");
            clock_step_ += m;
            persist_ll_next_ = clock_step_;
            n -= m;
            continue;
        }
        for (uint32_t k = 0; k < std::min<uint32_t>(m, 2); ++k) {
            const uint32_t buf = (m - 1 - k) % 3;
            BN_HIP_CHECK((hipError_t)launch_fill_u32(P.bits[buf], 0u, P.words, s), "persistent spike bits");
            if (P.spike_ring)
                BN_HIP_CHECK((hipError_t)launch_fill_u32(P.spike_ring + (size_t)((clock_step_ + k) % P.interval) * P.words,
                                                         0u, P.words, s),
                             "persistent ring slot");
        }
        BN_HIP_CHECK((hipError_t)launch_persistent(d_persist_plan_, P, m, persist_nontemporal_, s), "persistent step launch");
        clock_step_ += m;
        n -= m;
    }
}

}  // namespace bn
