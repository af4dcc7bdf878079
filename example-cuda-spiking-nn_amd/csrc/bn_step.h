// Internal (C++) interface between the host engine and the fused-step kernels.
//
// A step of the whole network is a fixed sequence of launches, each described by a plan
// struct passed by value as the kernel argument, so the sequence is captured once in a
// hipGraph and replayed: nothing in a plan changes between steps (the step index lives in the
// device clock).
//
//   k_evolve       every neuron: finalise g from the last step's GEMV partials, EIF step,
//                  background shot noise, connection pre-traces, plasticity trace decay (with
//                  the previous step's spike increments folded in), spike bits, monitor rings
//   k_gemv         every connection: FP64 partial sums of W^T x per row chunk
//   k_plasticity   LTP/LTD of every plasticity, organised by pre row (a thread per row for
//                  rows whose pre neuron did not spike, a block per spiking row); the lazily
//                  normalised matrix also updates its row sums and scale factors here; the
//                  last block applies this step's stimulus events and advances the clock
//   k_norm_tail    (reference-exact normalisation only) triplet row normalisation, then the
//                  stimulus and clock advance instead of k_plasticity
//
// Overlapped form of the same steps (captured graphs with lazy or no normalisation): the
// plasticity of step t-1 touches only weights, row sums and scales and reads the spike bits and
// plasticity traces of step t-1, so it runs in the launch of the evolve of step t
// (k_evolve_plast) while that evolve writes its spike bits and traces to the other buffer of a
// pair (even steps of the graph: current -> staging, odd steps: staging -> current; graphs have
// an even number of steps, so they end on the current buffers).  The stimulus events of step
// t-1 are applied per neuron by the evolve of step t, and the clock advances once per graph,
// each launch of the graph knowing its step offset.  Per graph of n steps:
//   k_evolve(t0) k_gemv(t0) [k_evolve_plast(t, t-1) k_gemv(t)] x (n-1) k_plasticity(t0+n-1, tail)
// The arithmetic of every neuron and weight is the sequential step's.
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
constexpr int kFinParts = 8;         // parts the chunk partials of a source are combined in
constexpr int kEvolveNeurons = 64;   // neurons per evolve block (kFinParts waves each)
constexpr int kPlastQuietRows = 64;  // rows per quiet-row block of the plasticity kernel
constexpr int kNormThreads = 256;    // threads per row in the row sums
constexpr int kNormRegs = 64;        // columns per thread kept in registers (n_post <= 16384)

enum PlastKind : uint32_t { kPlastTriplet = 0, kPlastIstdp = 1, kPlastStdp = 2 };

struct FinalizeSource {
    const double* partials;  // [chunks][ld], indexed by post-local id
    uint32_t chunks;
    uint32_t ld;
    uint32_t per;            // chunks per combination part (0: ceil(chunks / kFinParts))
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
    float* pre_out;    // where the step writes them (the same arrays, or staging copies)
    float* post_out;
    float* trip_pre_out;
    float* trip_post_out;
};

struct EvolvePlan {
    EvolveGroup g[kMaxGroups];
    TraceSet ts[kMaxTraceSets];
    PlastTraces pt[kMaxPlast];
    uint32_t n_groups, n_ts, n_pt, total_neurons;
    const bn_device_clock* clock;
    uint32_t* spike_bits;        // [words] bits of this step (global ids), or their staging copy
    uint32_t* spike_bits_ring;   // [interval][words] or nullptr
    float* v_ring;               // [interval][n_vrec] or nullptr
    const int32_t* vrec_index;   // [total_neurons] slot in v_ring or -1
    uint32_t n_vrec, interval, words;
    uint32_t step_offset;        // the step is clock_step + step_offset
    // Overlapped graphs: the stimulus events of the previous step, applied per neuron (in list
    // order) before its EIF step.
    uint32_t stim_prev;
    const uint32_t* ev_start;    // [ev_steps + 1] first event of each step
    const uint32_t* ev_ids;
    const uint32_t* ev_offsets;
    const float* ev_amps;
    uint32_t ev_steps, ev_count;
};

struct GemvConn {
    const float* W;
    const float* x;          // per pre-local id
    const float* row_scale;  // nullptr or s_i (lazy normalisation): x' = x_i * s_i
    double* partials;        // [n_chunks][ld]
    uint32_t ld, n_pre, n_post, row_block, n_chunks, n_coltiles, tile_begin, tile_end;
};

// The spiking pre neurons of every plasticity, compacted by one extra block of the propagate
// launch of the same step (list == nullptr: none), for the plasticity of that step to read
// instead of compacting the spike bits again: list = pre-local ids in plasticity order, each
// plasticity's in increasing order (entries past kPreListCap are not stored), counts[k] = the
// count of plasticity k.
constexpr uint32_t kPreListCap = 4096;
struct PreCompact {
    const uint32_t* bits;    // spike bits of the step (global ids)
    uint32_t n;              // plasticities
    uint32_t pre_offset[kMaxPlast], n_pre[kMaxPlast];
    uint32_t* list;          // [kPreListCap]
    uint32_t* counts;        // [kMaxPlast]
};

constexpr uint32_t kGemvPartChunksMax = 16;   // chunks per part a part-sum block holds in LDS

struct GemvPlan {
    GemvConn c[kMaxConns];
    uint32_t n_conn;
    uint32_t parts;          // 1: blocks of part sums (tile = (part, 256 columns)), partials row p = part p
    uint32_t total_tiles;    // one wave per tile (parts: one block per tile)
    uint32_t nontemporal;    // stream the weights past the caches (weight set >> MALL)
    PreCompact pc;           // pc.list == nullptr: no compaction block
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
    const uint32_t* bits;         // spike bits of this step (global ids)
    // lazy normalisation (W = diag(s) M) of a normalising triplet plasticity
    uint32_t lazy;
    float* s;                     // [n_pre] row scale factors
    double* R;                    // [n_pre] FP64 row sums of M
};

struct PlastPlan {
    PlastDev p[kMaxPlast];
    uint32_t n;
    uint32_t items_post;  // sum of n_post
    uint32_t items_pre;   // sum of n_pre
};

// Stimulus events of the step being executed (device cursor) and the clock advance.
struct StimTail {
    float* i_stimulus;           // global
    const uint32_t* timestep_event_count;
    const uint32_t* event_ids;
    const uint32_t* event_offsets;
    const float* event_amplitudes;
    uint32_t num_timesteps, num_events;
    bn_device_clock* clock;
    uint32_t* status;            // device error word (1 = stimulus cursor overflow)
    uint32_t active;             // 0: nothing to do in this kernel
    // End of an overlapped graph: the events of step clock_step + step_offset, found through
    // ev_start (the cursor then points past them), and a clock advance of `advance` steps.
    // Sequential steps: ev_start nullptr (the cursor's events), offset 0, advance 0 (= one step).
    const uint32_t* ev_start;
    uint32_t step_offset, advance;
};

struct PlastStepPlan {
    PlastPlan pl;
    uint32_t row_blocks;         // blocks of kPlastQuietRows rows (rows whose pre did not spike)
    uint32_t spiking_blocks;     // blocks sharing the rows whose pre neuron spiked
    StimTail tail;
    // The spiking pre neurons as compacted by this step's propagate launch (PreCompact), or
    // nullptr: the spiking-row blocks compact the spike bits themselves.
    const uint32_t* pre_list;
    const uint32_t* pre_counts;
};

struct TailPlan {
    PlastPlan pl;
    uint32_t norm_rows;          // rows normalised (sum over normalising plasticities)
    uint32_t total_neurons;
    uint32_t increments;         // 1: apply plasticity trace increments here
    StimTail tail;
};

// Launchers used by the host engine (return hipError_t as int).
int launch_evolve(const EvolvePlan& plan, bn_stream_t s);

// Several networks' fused steps in one launch per phase: per-network plans in device memory
// and, per phase, a block -> network table and each network's first block.
struct MultiPhase {
    const uint32_t* block_net;
    const uint32_t* net_first_block;
    uint32_t blocks;
};
// Per-network plans of the overlapped graph form (device memory; null: not prepared).
struct MultiOverlap {
    const EvolvePlan* evolve_even;     // even steps of a graph: current -> staging buffers
    const EvolvePlan* evolve_odd;      // odd steps: staging -> current
    const PlastStepPlan* plast_staging;  // plasticity of an even step
    const PlastStepPlan* plast_current;  // of an odd step
    const PlastStepPlan* plast_final;    // of the last step, with the stimulus/clock tail
};
struct MultiStep {
    const EvolvePlan* evolve;
    const GemvPlan* gemv;
    const PlastStepPlan* plast;
    MultiPhase ph_evolve, ph_gemv, ph_plast;
    uint32_t nontemporal;
    MultiOverlap ov;
};
int launch_multi_step(const MultiStep& m, bn_stream_t s);
// n (even) overlapped steps of every network of m (the graph form of a single network's
// overlapped steps, with the networks' launches merged per phase).
int launch_multi_overlapped(const MultiStep& m, uint32_t n, bn_stream_t s);
int launch_multi_propagate(const MultiStep& m, bn_stream_t s);
int launch_gemv(const GemvPlan& plan, bn_stream_t s);
// Overlapped step: evolve of step t with the plasticity of step t-1 (plast = nullptr: evolve
// only).
int launch_evolve_plast(const EvolvePlan& ev, const PlastStepPlan* plast, bn_stream_t s);
int launch_plasticity(const PlastStepPlan& plan, bn_stream_t s);
int launch_ltp(const PlastPlan& plan, bn_stream_t s);
int launch_ltd(const PlastPlan& plan, bn_stream_t s);
int launch_norm_tail(const TailPlan& plan, bn_stream_t s);
// R_i = canonical row sum of W, s_i = 1 (set-up of the lazy normalisation).
int launch_lazy_init(const PlastDev& d, bn_stream_t s);
// g[j] = fl32((double)g[j] + sum_p partials[p][j]) for every source, in order.
int launch_materialize_g(float* g_dst, const float* g_base, const FinalizeSource* src,
                         uint32_t n_src, uint32_t n, bn_stream_t s);
// dst[i] = src[i] + (spikes[i] == 1 ? 1 : 0): plasticity traces with this step's increments.
int launch_add_increments(float* dst, const float* src, const uint32_t* spikes, uint32_t n, bn_stream_t s);
int launch_fill_u32(uint32_t* p, uint32_t value, size_t n, bn_stream_t s);
int launch_read_probe(const void* buf, uint64_t bytes, float* sink, uint32_t blocks, bn_stream_t s);
// `passes` passes over the buffer in one launch of `blocks` blocks (each pass from a rotated start).
int launch_read_probe_passes(const void* buf, uint64_t bytes, uint32_t passes, float* sink, uint32_t blocks,
                             bn_stream_t s);

}  // namespace bn
