// Internal (C++) interface between the host engine and the fused-step kernels.
//
// A step of the whole network is a fixed sequence of launches, each described by a plan
// struct passed by value as the kernel argument (so the sequence can be captured once in a
// hipGraph and replayed: nothing in a plan changes between steps; the step index lives in
// the device clock).
//
//   faithful step (reference-exact triplet normalisation):
//     k_evolve      every neuron: finalise g from last step's GEMV partials, EIF step,
//                   background shot noise, connection pre-traces, plasticity trace decays,
//                   spike bits and monitor rings
//     k_gemv        every connection: FP64 partial sums of W^T x per row chunk
//     k_ltp         LTP column walks for spiking post neurons
//     k_ltd         LTD row walks for spiking pre neurons
//     k_norm_tail   triplet row normalisation, plasticity trace increments, stimulus
//                   events of this step, clock advance
//
//   lazy step (W_EE = diag(s) * M; normalisation by per-row scale factors):
//     k_evolve -> k_fused_ee (GEMV + LTP/LTD of the lazily normalised matrix, fused in one
//     pass) together with k_gemv for the other connections -> k_rowscale_tail
#pragma once

#include <stdint.h>

#include "buryn_amd.h"

namespace bn {

constexpr int kMaxGroups = 4;
constexpr int kMaxConns = 4;
constexpr int kMaxSources = 2;
constexpr int kMaxPlast = 4;
constexpr int kMaxTraceSets = 4;
constexpr int kGemvTileCols = 256;   // 64 lanes x float4
constexpr int kNormThreads = 256;    // threads per row in the normalisation sum
constexpr int kNormRegs = 64;        // columns per thread kept in registers (n_post <= 16384)

enum PlastKind : uint32_t { kPlastTriplet = 0, kPlastIstdp = 1, kPlastStdp = 2 };

struct FinalizeSource {
    const double* partials;  // [chunks][ld], indexed by post-local id
    uint32_t chunks;
    uint32_t ld;
};

struct EvolveGroup {
    NeuronGroupView v;
    FinalizeSource fin_e[kMaxSources];
    FinalizeSource fin_i[kMaxSources];
    uint32_t n_fin_e;
    uint32_t n_fin_i;
};

// Pre-synaptic difference-of-exponentials traces of one pre group (shared by every
// connection with the same pre group and the same rise/decay/normalisation constants).
struct TraceSet {
    uint32_t pre_offset, n_pre;
    float decay_r, decay_d, K;
    float* r;
    float* d;
    float* x;
};

struct PlastTraces {
    uint32_t kind;
    uint32_t pre_offset, n_pre, post_offset, n_post;
    float pre_decay, post_decay, trip_pre_decay, trip_post_decay;
    float* pre;        // indexed by pre-local id
    float* post;       // indexed by post-local id
    float* trip_pre;
    float* trip_post;
};

struct EvolvePlan {
    EvolveGroup g[kMaxGroups];
    TraceSet ts[kMaxTraceSets];
    PlastTraces pt[kMaxPlast];
    uint32_t n_groups, n_ts, n_pt, total_neurons;
    const bn_device_clock* clock;
    uint32_t* spike_bits;        // [words] bits of this step (global ids)
    uint32_t* spike_bits_ring;   // [interval][words] or nullptr
    float* v_ring;               // [interval][n_vrec] or nullptr
    const int32_t* vrec_index;   // [total_neurons] slot in v_ring or -1
    uint32_t n_vrec, interval, words;
    uint32_t skip_plast_decay;   // 1: plasticity trace decays are done elsewhere
};

struct GemvConn {
    const float* W;
    const float* x;          // per pre-local id
    const float* row_scale;  // nullptr or s_i (lazy normalisation): x' = x_i * s_i
    double* partials;        // [n_chunks][ld]
    uint32_t ld, n_pre, n_post, row_block, n_chunks, n_coltiles, tile_begin, tile_end;
};

struct GemvPlan {
    GemvConn c[kMaxConns];
    uint32_t n_conn;
    uint32_t total_tiles;    // one wave per tile
};

struct PlastDev {
    uint32_t kind;
    uint32_t normalize;
    PlasticityParams p;
    float* W;
    uint32_t ld, n_pre, n_post, pre_offset, post_offset;
    const uint32_t* pre_spikes;   // pre-local
    const uint32_t* post_spikes;  // post-local
    float* pre;
    float* post;
    float* trip_pre;
    float* trip_post;
    const float* target;          // pre_weight_sums (triplet normalisation targets)
    // lazy normalisation (W = diag(s) M): only for a normalising triplet plasticity
    uint32_t lazy;
    uint32_t kmax;                // capacity of `delta` in spiking posts
    float* s;                     // [n_pre] row scale factors
    double* R;                    // [n_pre] FP64 row sums of M
    double* delta;                // [kmax][n_pre] LTP deltas of M, by rank of the spiking post
    const uint32_t* bits;         // spike bits of this step (global ids)
};

struct PlastPlan {
    PlastDev p[kMaxPlast];
    uint32_t n;
    uint32_t items_post;  // sum of n_post (LTP work items)
    uint32_t items_pre;   // sum of n_pre (LTD work items)
};

struct TailPlan {
    PlastPlan pl;
    uint32_t norm_rows;          // rows normalised (sum over normalising plasticities)
    uint32_t total_neurons;
    uint32_t increments;         // 1: apply plasticity trace increments here
    // stimulus (device cursor)
    float* i_stimulus;           // global
    const uint32_t* timestep_event_count;
    const uint32_t* event_ids;
    const uint32_t* event_offsets;
    const float* event_amplitudes;
    uint32_t num_timesteps, num_events;
    bn_device_clock* clock;
    uint32_t* status;            // device error word (1 = stimulus cursor overflow)
    uint32_t advance_clock;
};

// Launchers used by the host engine (return hipError_t as int).
int launch_evolve(const EvolvePlan& plan, bn_stream_t s);
int launch_gemv(const GemvPlan& plan, bn_stream_t s);
int launch_ltp(const PlastPlan& plan, bn_stream_t s);
int launch_ltd(const PlastPlan& plan, bn_stream_t s);
int launch_norm_tail(const TailPlan& plan, bn_stream_t s);
// Lazy normalisation: LTD rows of the lazy matrix, row-sum bookkeeping and scale factors.
int launch_lazy_rows(const PlastPlan& plan, bn_stream_t s);
// R_i = canonical row sum of W, s_i = 1 (set-up of the lazy normalisation).
int launch_lazy_init(const PlastDev& d, bn_stream_t s);
// g[j] = fl32((double)g[j] + sum_p partials[p][j]) for every source, in order.
int launch_materialize_g(float* g_dst, const float* g_base, const FinalizeSource* src,
                         uint32_t n_src, uint32_t n, bn_stream_t s);
int launch_fill_u32(uint32_t* p, uint32_t value, size_t n, bn_stream_t s);

}  // namespace bn
