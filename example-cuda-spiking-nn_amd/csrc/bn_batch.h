// Launch plans of the batched shared-weight step: B instances of one network (same frozen
// weights, own noise key and stimulus each) advanced together.  Neuron state is laid out
// [neuron][instance] (instance fastest), so a wave of 64 lanes is one neuron in 64 instances
// and the pre-synaptic activity of all instances is the [pre][B] operand of one GEMM per
// connection on MFMA.
#pragma once

#include <stdint.h>

#include "buryn_amd.h"

namespace bn {

constexpr int kBatchMaxGroups = 2;
constexpr int kBatchMaxConns = 4;
constexpr int kBatchGemmWaves = 4;   // row chunks (waves) per GEMM block, summed in LDS
constexpr int kBatchGemmTileM = 64;  // post neurons per GEMM wave tile (ld is a multiple of 64)
constexpr int kBatchAlign = 64;      // instance count padding (one wave = 64 instances)

struct BatchPartials {
    const float* part;   // [groups][n_post][B] FP32 partial sums of one connection
    uint32_t groups;
};

struct BatchGroup {
    NeuronGroupView v;   // scalar parameters and the Poisson tables (state pointers unused)
    BatchPartials fin_e, fin_i;
    uint32_t has_fin_e, has_fin_i;
};

struct BatchTraceSet {
    uint32_t pre_offset, n_pre;
    float decay_r, decay_d, K;
    float* r;            // [n_pre][B]
    float* d;
    float* x;
};

struct BatchEvolvePlan {
    BatchGroup g[kBatchMaxGroups];
    BatchTraceSet ts[kBatchMaxGroups];
    uint32_t n_groups, n_ts, N, B;
    float* V;
    float* ge;           // background conductance of the last step; the GEMM partials add to it
    float* gi;
    float* istim;
    uint32_t* refr;
    uint32_t* spikes;
    float* bre;
    float* bde;
    float* bri;
    float* bdi;
    const uint32_t* keys;        // [B][2] Philox key per instance
    const float* stim_amp;       // [N][B] pulse amplitude of (neuron, instance), 0 if not stimulated
    const uint32_t* stim_on;     // [B] step whose stimulus phase adds the pulse
    const uint32_t* stim_off;    // [B] step whose stimulus phase removes it
    uint32_t* step;              // device step counter (advanced by the GEMM's last block)
};

struct BatchGemmConn {
    const float* W;      // [n_pre][ld] row-major
    const float* X;      // [n_pre][B]
    float* part;         // [groups][n_post][B]
    uint32_t ld, n_pre, n_post, kc, n_chunks, groups, m_tiles, n_pairs, block_begin, block_end;
};

struct BatchGemmPlan {
    BatchGemmConn c[kBatchMaxConns];
    uint32_t n_conn, total_blocks, B;
    uint32_t* step;
};

int launch_batch_evolve(const BatchEvolvePlan& plan, bn_stream_t s);
int launch_batch_gemm(const BatchGemmPlan& plan, bn_stream_t s);

}  // namespace bn
