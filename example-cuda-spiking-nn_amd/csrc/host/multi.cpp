// Several independent plastic networks of any sizes advanced together: every network keeps its
// own state, weights, plasticity and fused-step plans (exactly as a single bn_sim), and one
// launch per phase (evolve, propagate, plasticity) runs that phase for all of them, ten steps
// per hipGraph.  Each network therefore computes bit for bit what it computes alone.  With
// several streams, the networks are split into sets of about equal weight bytes stepped on
// their own streams, so one set's latency-bound evolve and plasticity launches overlap another
// set's bandwidth-bound propagate.
// extern "C" bn_multi_* API of include/buryn_amd.h.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "api_util.h"
#include "sim_internal.h"

using namespace bn;

namespace {
constexpr uint32_t kMultiSpikingBlocks = 16;  // blocks per network for its spiking rows
}  // namespace

struct bn_multi {
    struct Group {
        MultiStep step{};
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        uint64_t bytes = 0;
        uint32_t networks = 0;
        hipStream_t stream = nullptr;    // the main stream, or a lane of its own
    };
    std::vector<bn_sim*> sims;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;       // plans and the all-network tables (kept until destroy)
    std::vector<void*> group_allocs; // tables of the current schedule
    MultiStep step{};                // every network in one launch per phase
    std::vector<Group> groups;       // the schedule: network groups, each stepped by its own graph
    uint32_t steps_per_graph = 10;
    uint64_t group_bytes = 0;        // weight bytes per group (0 = one group)
    uint32_t num_streams = 1;        // > 1: concurrent network sets, one stream each
    bool overlap_enabled = false;    // graphs overlap plasticity(t-1) with evolve(t)
    bool overlapped = false;         // the current graphs do
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> lane_events;
    hipEvent_t main_event = nullptr;
    uint32_t clock = 0;
    const EvolvePlan* d_ev = nullptr;
    const GemvPlan* d_gm = nullptr;
    const PlastStepPlan* d_pl = nullptr;
    MultiOverlap ov{};                // overlapped form (graphs of an even number of steps)
    std::vector<PlastStepPlan> pl_final;
    PlastStepPlan* d_pl_final = nullptr;
    std::vector<uint32_t> nb_ev, nb_gm, nb_pl;
    std::vector<uint64_t> w_bytes;

    ~bn_multi() {
        clear_groups();
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipStream_t l : lanes) (void)hipStreamDestroy(l);
        for (hipEvent_t e : lane_events) (void)hipEventDestroy(e);
        if (main_event) (void)hipEventDestroy(main_event);
        for (void* p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
        for (bn_sim* s : sims) bn_sim_destroy(s);
    }

    template <typename T>
    T* upload(const std::vector<T>& v, const char* what, std::vector<void*>& owner) {
        void* p = nullptr;
        BN_HIP_CHECK(hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)), std::string("hipMalloc ") + what);
        owner.push_back(p);
        if (!v.empty())
            BN_HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), std::string("upload ") + what);
        return static_cast<T*>(p);
    }

    // Block tables of one phase over the networks `nets` (global network ids).
    MultiPhase phase(const std::vector<uint32_t>& blocks_per_net, const std::vector<uint32_t>& nets, const char* what,
                     std::vector<void*>& owner) {
        std::vector<uint32_t> net_of, first(sims.size(), 0);
        for (uint32_t n : nets) {
            first[n] = (uint32_t)net_of.size();
            net_of.insert(net_of.end(), blocks_per_net[n], n);
        }
        MultiPhase ph;
        ph.blocks = (uint32_t)net_of.size();
        ph.block_net = upload(net_of, what, owner);
        ph.net_first_block = upload(first, what, owner);
        return ph;
    }

    MultiStep make_step(const std::vector<uint32_t>& nets, std::vector<void*>& owner) {
        MultiStep st{};
        st.evolve = d_ev;
        st.gemv = d_gm;
        st.plast = d_pl;
        st.ph_evolve = phase(nb_ev, nets, "evolve block table", owner);
        st.ph_gemv = phase(nb_gm, nets, "propagate block table", owner);
        st.ph_plast = phase(nb_pl, nets, "plasticity block table", owner);
        uint64_t bytes = 0;
        for (uint32_t n : nets) bytes += w_bytes[n];
        // Weight sets well above the MALL stream past the caches; smaller ones stay resident.
        st.nontemporal = bytes > (192ull << 20) ? 1u : 0u;
        st.ov = ov;
        return st;
    }

    void build() {
        BN_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        std::vector<EvolvePlan> ev, ev_even, ev_odd;
        std::vector<GemvPlan> gm;
        std::vector<PlastStepPlan> pl, pl_stage, pl_cur;
        for (bn_sim* s : sims) {
            System& sys = *s->system;
            if (!sys.fused()) throw Error(BN_ERR_RUNTIME, "multi-network step needs the fused step of every network");
            if (sys.gemv_part_sums())
                throw Error(BN_ERR_INVALID_ARGUMENT, "multi-network step: a network was created with BN_SIM_GEMV_PART_SUMS");
            if (sys.has_norm_tail())
                throw Error(BN_ERR_INVALID_ARGUMENT,
                            "multi-network step supports lazy or no normalisation (BN_SIM_LAZY_NORMALIZE)");
            BN_HIP_CHECK(hipStreamSynchronize(s->state->stream()), "network setup sync");
            ev.push_back(sys.evolve_plan());
            gm.push_back(sys.gemv_plan());
            // Fewer blocks share a network's spiking rows than when it runs alone (a network has
            // a few spiking rows per step and the launch holds every network's blocks); which
            // block takes a row does not change the row's result.
            auto fewer = [](PlastStepPlan p) {
                p.spiking_blocks = std::min<uint32_t>(p.spiking_blocks, kMultiSpikingBlocks);
                p.pre_list = p.pre_counts = nullptr;  // the multi-network propagate does not compact
                return p;
            };
            pl.push_back(fewer(sys.plast_step_plan()));
            sys.prepare_overlap(2);
            ev_even.push_back(sys.evolve_even());
            ev_odd.push_back(sys.evolve_odd());
            pl_stage.push_back(fewer(sys.plast_staging()));
            pl_cur.push_back(fewer(sys.plast_current()));
            pl_final.push_back(fewer(sys.plast_final()));
            nb_ev.push_back((sys.evolve_plan().total_neurons + kEvolveNeurons - 1) / kEvolveNeurons);
            nb_gm.push_back((sys.gemv_plan().total_tiles + 3) / 4);
            nb_pl.push_back(pl.back().row_blocks + pl.back().spiking_blocks);
            uint64_t wb = 0;
            for (uint32_t c = 0; c < sys.gemv_plan().n_conn; ++c)
                wb += 4ull * sys.gemv_plan().c[c].n_pre * sys.gemv_plan().c[c].ld;
            w_bytes.push_back(wb);
        }
        d_ev = upload(ev, "evolve plans", allocs);
        d_gm = upload(gm, "propagate plans", allocs);
        d_pl = upload(pl, "plasticity plans", allocs);
        ov.evolve_even = upload(ev_even, "evolve plans", allocs);
        ov.evolve_odd = upload(ev_odd, "evolve plans", allocs);
        ov.plast_staging = upload(pl_stage, "plasticity plans", allocs);
        ov.plast_current = upload(pl_cur, "plasticity plans", allocs);
        d_pl_final = upload(pl_final, "plasticity plans", allocs);
        ov.plast_final = d_pl_final;
        std::vector<uint32_t> all(sims.size());
        for (uint32_t n = 0; n < all.size(); ++n) all[n] = n;
        step = make_step(all, allocs);
        schedule(0, 10, 1);
    }

    void clear_groups() {
        join();
        for (Group& g : groups) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            if (g.graph) (void)hipGraphDestroy(g.graph);
        }
        groups.clear();
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : group_allocs) (void)hipFree(p);
        group_allocs.clear();
    }

    // Networks are independent, so the steps of different networks need not be interleaved:
    // groups of networks whose weights fit `bytes` (0 = all in one group) each advance
    // `spg` steps per graph launch, one group after the other, so a group's weights can stay
    // in the MALL across its steps.
    void schedule(uint64_t bytes, uint32_t spg, uint32_t nstreams) {
        if (spg == 0) throw Error(BN_ERR_INVALID_ARGUMENT, "steps per graph must be positive");
        if (nstreams == 0 || nstreams > 4) throw Error(BN_ERR_INVALID_ARGUMENT, "streams must be 1 to 4");
        clear_groups();
        group_bytes = nstreams > 1 ? 0 : bytes;
        steps_per_graph = spg;
        num_streams = std::min<uint32_t>(nstreams, (uint32_t)sims.size());
        std::vector<std::vector<uint32_t>> sets(1);
        if (num_streams > 1) {
            // Heaviest networks first, each to the lightest set; each set keeps network order.
            sets.assign(num_streams, {});
            std::vector<uint64_t> load(num_streams, 0);
            std::vector<uint32_t> order(sims.size());
            for (uint32_t n = 0; n < order.size(); ++n) order[n] = n;
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return w_bytes[a] > w_bytes[b]; });
            for (uint32_t n : order) {
                const size_t k = std::min_element(load.begin(), load.end()) - load.begin();
                sets[k].push_back(n);
                load[k] += w_bytes[n];
            }
            for (auto& set : sets) std::sort(set.begin(), set.end());
            while (lanes.size() < num_streams) {
                hipStream_t l = nullptr;
                BN_HIP_CHECK(hipStreamCreateWithFlags(&l, hipStreamNonBlocking), "hipStreamCreate");
                lanes.push_back(l);
                hipEvent_t e = nullptr;
                BN_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
                lane_events.push_back(e);
            }
        } else {
            uint64_t acc = 0;
            for (uint32_t n = 0; n < sims.size(); ++n) {
                if (bytes && !sets.back().empty() && acc + w_bytes[n] > bytes) {
                    sets.emplace_back();
                    acc = 0;
                }
                sets.back().push_back(n);
                acc += w_bytes[n];
            }
        }
        // The last plasticity of an overlapped graph advances the clock by the graph's steps.
        overlapped = overlap_enabled && spg % 2 == 0;
        if (overlapped) {
            for (PlastStepPlan& p : pl_final) {
                p.tail.step_offset = spg - 1;
                p.tail.advance = spg;
            }
            BN_HIP_CHECK(hipMemcpy(d_pl_final, pl_final.data(), pl_final.size() * sizeof(PlastStepPlan),
                                   hipMemcpyHostToDevice), "final plasticity plans");
        }
        for (size_t i = 0; i < sets.size(); ++i) {
            const auto& set = sets[i];
            Group g;
            g.step = make_step(set, group_allocs);
            g.networks = (uint32_t)set.size();
            g.stream = num_streams > 1 ? lanes[i] : stream;
            for (uint32_t n : set) g.bytes += w_bytes[n];
            BN_HIP_CHECK(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            if (overlapped)
                BN_HIP_CHECK((hipError_t)launch_multi_overlapped(g.step, steps_per_graph, g.stream),
                             "multi-network overlapped steps launch");
            else
                for (uint32_t k = 0; k < steps_per_graph; ++k)
                    BN_HIP_CHECK((hipError_t)launch_multi_step(g.step, g.stream), "multi-network step launch");
            BN_HIP_CHECK(hipStreamEndCapture(g.stream, &g.graph), "hipStreamEndCapture");
            BN_HIP_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            groups.push_back(g);
        }
    }

    // Lanes start after the work already on the main stream, and the main stream continues
    // after the lanes' work, so everything else (timing events, downloads) orders as before.
    void fork() {
        if (num_streams <= 1) return;
        if (!main_event) BN_HIP_CHECK(hipEventCreateWithFlags(&main_event, hipEventDisableTiming), "hipEventCreate");
        BN_HIP_CHECK(hipEventRecord(main_event, stream), "hipEventRecord");
        for (size_t i = 0; i < num_streams; ++i) BN_HIP_CHECK(hipStreamWaitEvent(lanes[i], main_event, 0), "hipStreamWaitEvent");
    }
    void join() {
        if (num_streams <= 1 || lanes.empty()) return;
        for (size_t i = 0; i < num_streams; ++i) {
            BN_HIP_CHECK(hipEventRecord(lane_events[i], lanes[i]), "hipEventRecord");
            BN_HIP_CHECK(hipStreamWaitEvent(stream, lane_events[i], 0), "hipStreamWaitEvent");
        }
    }

    void run(uint32_t n) {
        fork();
        uint32_t k = 0;
        for (; k + steps_per_graph <= n; k += steps_per_graph)
            for (Group& g : groups) BN_HIP_CHECK(hipGraphLaunch(g.exec, g.stream), "hipGraphLaunch");
        for (Group& g : groups)
            for (uint32_t r = k; r < n; ++r)
                BN_HIP_CHECK((hipError_t)launch_multi_step(g.step, g.stream), "multi-network step launch");
        join();
        for (bn_sim* s : sims) s->system->note_external_steps(n);
        clock += n;
    }
};

extern "C" {

int bn_multi_create(const char* const* config_paths, uint32_t num_networks, uint32_t flags, bn_multi** out) {
    if (!out) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_create: out is null");
    *out = nullptr;
    if (!config_paths || num_networks == 0) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_create: no networks");
    if (flags & (BN_SIM_COMPONENT_PATH | BN_SIM_WRITE_FILES))
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_create: the fused step only, without files");
    auto m = std::make_unique<bn_multi>();
    for (uint32_t k = 0; k < num_networks; ++k) {
        bn_sim* s = nullptr;
        const int rc = bn_sim_create(config_paths[k], flags | BN_SIM_NO_MONITORS | BN_SIM_NO_GRAPH | BN_SIM_QUIET, &s);
        if (rc) return set_error(rc, "network " + std::to_string(k) + ": " + bn_last_error());
        m->sims.push_back(s);
    }
    const int rc = guarded([&] { m->build(); });
    if (rc) return rc;
    *out = m.release();
    return BN_OK;
}

void bn_multi_destroy(bn_multi* m) { delete m; }

int bn_multi_run(bn_multi* m, uint32_t num_steps) {
    if (!m) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_run: null handle");
    return guarded([&] { m->run(num_steps); });
}

int bn_multi_run_timed(bn_multi* m, uint32_t num_steps, float* ms) {
    if (!m || !ms) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_run_timed: null argument");
    return guarded([&] {
        hipEvent_t a, b;
        BN_HIP_CHECK(hipEventCreate(&a), "hipEventCreate");
        BN_HIP_CHECK(hipEventCreate(&b), "hipEventCreate");
        BN_HIP_CHECK(hipStreamSynchronize(m->stream), "pre-timing sync");
        BN_HIP_CHECK(hipEventRecord(a, m->stream), "hipEventRecord");
        m->run(num_steps);
        BN_HIP_CHECK(hipEventRecord(b, m->stream), "hipEventRecord");
        BN_HIP_CHECK(hipEventSynchronize(b), "hipEventSynchronize");
        BN_HIP_CHECK(hipEventElapsedTime(ms, a, b), "hipEventElapsedTime");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    });
}

int bn_multi_time_propagate(bn_multi* m, uint32_t num_launches, float* ms) {
    if (!m || !ms) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_time_propagate: null argument");
    return guarded([&] {
        // Back-to-back propagate launches over every network; the pending partial sums belong
        // to the next step and are saved before and restored after.
        std::vector<std::pair<double*, size_t>> bufs;
        size_t total = 0;
        for (bn_sim* s : m->sims) {
            const GemvPlan& G = s->system->gemv_plan();
            for (uint32_t c = 0; c < G.n_conn; ++c) {
                bufs.emplace_back(G.c[c].partials, (size_t)G.c[c].n_chunks * G.c[c].ld);
                total += bufs.back().second;
            }
        }
        double* save = nullptr;
        BN_HIP_CHECK(hipMalloc(&save, std::max<size_t>(1, total) * sizeof(double)), "hipMalloc(partials copy)");
        size_t off = 0;
        for (auto& b : bufs) {
            BN_HIP_CHECK(hipMemcpyAsync(save + off, b.first, b.second * sizeof(double), hipMemcpyDeviceToDevice, m->stream),
                         "save partials");
            off += b.second;
        }
        hipEvent_t a, e;
        BN_HIP_CHECK(hipEventCreate(&a), "hipEventCreate");
        BN_HIP_CHECK(hipEventCreate(&e), "hipEventCreate");
        BN_HIP_CHECK(hipStreamSynchronize(m->stream), "pre-timing sync");
        BN_HIP_CHECK(hipEventRecord(a, m->stream), "hipEventRecord");
        for (uint32_t k = 0; k < num_launches; ++k)
            BN_HIP_CHECK((hipError_t)launch_multi_propagate(m->step, m->stream), "propagate launch");
        BN_HIP_CHECK(hipEventRecord(e, m->stream), "hipEventRecord");
        BN_HIP_CHECK(hipEventSynchronize(e), "hipEventSynchronize");
        BN_HIP_CHECK(hipEventElapsedTime(ms, a, e), "hipEventElapsedTime");
        off = 0;
        for (auto& b : bufs) {
            BN_HIP_CHECK(hipMemcpyAsync(b.first, save + off, b.second * sizeof(double), hipMemcpyDeviceToDevice, m->stream),
                         "restore partials");
            off += b.second;
        }
        BN_HIP_CHECK(hipStreamSynchronize(m->stream), "timing sync");
        (void)hipFree(save);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(e);
    });
}

int bn_multi_set_schedule(bn_multi* m, uint64_t group_weight_bytes, uint32_t steps_per_graph) {
    if (!m) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_set_schedule: null handle");
    return guarded([&] {
        m->overlap_enabled = false;
        m->schedule(group_weight_bytes, steps_per_graph, 1);
    });
}

int bn_multi_set_streams(bn_multi* m, uint32_t streams, uint32_t steps_per_graph, uint32_t overlap) {
    if (!m) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_set_streams: null handle");
    if (overlap && steps_per_graph % 2)
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_set_streams: overlapped graphs need an even step count");
    return guarded([&] {
        m->overlap_enabled = overlap != 0;
        m->schedule(0, steps_per_graph, streams);
    });
}

int bn_multi_synchronize(bn_multi* m) {
    if (!m) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_synchronize: null handle");
    return guarded([&] { BN_HIP_CHECK(hipStreamSynchronize(m->stream), "bn_multi_synchronize"); });
}

// The network's own bn_sim, for its state, weights, traces and info (after the steps so far;
// valid until bn_multi_destroy).  The multi stream is synchronised first.
int bn_multi_network(bn_multi* m, uint32_t index, bn_sim** out) {
    if (!m || !out) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_network: null argument");
    if (index >= m->sims.size()) return set_error(BN_ERR_OUT_OF_RANGE, "bn_multi_network: index out of range");
    return guarded([&] {
        BN_HIP_CHECK(hipStreamSynchronize(m->stream), "multi sync");
        *out = m->sims[index];
    });
}

int bn_multi_get_info(bn_multi* m, bn_multi_info* info) {
    if (!m || !info) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_multi_get_info: null argument");
    std::memset(info, 0, sizeof *info);
    info->num_networks = (uint32_t)m->sims.size();
    info->clock_step = m->clock;
    info->evolve_blocks = m->step.ph_evolve.blocks;
    info->propagate_blocks = m->step.ph_gemv.blocks;
    info->plasticity_blocks = m->step.ph_plast.blocks;
    info->nontemporal = m->step.nontemporal;
    info->kernels_per_step = (m->overlapped ? 2u : 3u) * (uint32_t)m->groups.size();
    info->num_groups = (uint32_t)m->groups.size();
    info->num_streams = m->num_streams;
    info->overlapped = m->overlapped ? 1u : 0u;
    info->steps_per_graph = m->steps_per_graph;
    info->group_weight_bytes = m->group_bytes;
    return BN_OK;
}

}  // extern "C"
