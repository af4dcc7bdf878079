// Internal (C++) interface of the persistent step: one cooperative launch runs n whole steps of
// the network (kernels_persist.hip).
//
// Work of a step t:
//   tiles      the four connections cut into items (connection, 256-column tile, segment of
//              kPWaves * rb pre rows); wave w of the block owning an item sums rows
//              [seg*8*rb + w*rb, +rb) of its 256 columns in FP64 from 0.0 (a chunk), the block
//              adds its 8 chunk sums in chunk order from 0.0 (a segment partial) and stores it.
//              The same pass applies the plasticity of step t to the entries it holds (LTP of
//              spiking post columns, LTD of spiking pre rows, in the reference's LTP-then-LTD
//              order per entry) after using their old values, and, for a lazily normalised
//              matrix, the tile's part of each row's change of R (column-tile order).
//              With resident weights (one item per block, chunks of at most 24 rows) the
//              entries stay in the registers of their thread for the whole launch and only
//              the rows that changed are stored, at the end of the launch.
//   tails      the block that finishes the last item of a column tile of a post group runs the
//              evolve of step t+1 for those neurons (partials combined, background, EIF,
//              connection and plasticity traces, spike bits, rings); the block that finishes the
//              last column tile of a segment of the lazy matrix updates R_i and s_i of its rows.
//   ordering   a step's tiles start once every tail of the previous step has finished
//              (one counter of finished tails); counters only grow within a launch.
//   coherence  what one block writes for another (partials, row-sum parts, neuron state,
//              traces, row scales, spike bits) moves through agent-coherent loads and stores;
//              a block completes its stores before it raises a counter.  No fence writes back
//              or invalidates an L2.  Weight entries stay with the thread that owns them.
// Buffers read by one step while the next step's evolve writes them are doubled (x, plasticity
// traces: evolve k of a launch of n writes buffer (n-1-k) & 1, 0 being the state's own) or
// tripled (spike bits, set with atomic OR after being cleared two steps ahead: buffer
// (n-1-k) % 3), so that the last evolve of a launch writes the state's own buffers.
#pragma once

#include <stdint.h>

#include "bn_step.h"

namespace bn {

constexpr int kPWaves = 8;                  // waves per block = chunks per segment
constexpr int kPThreads = kPWaves * 64;     // 512
constexpr int kPCols = 256;                 // columns per tile (64 lanes x float4)
constexpr int kPBatch = 8;                  // rows of a chunk loaded together
constexpr uint32_t kPMaxRows = 128;         // rows per chunk at most (a segment is staged in LDS)
constexpr uint32_t kPNone = 0xFFFFFFFFu;

struct PConn {
    float* W;                   // [n_pre][ld]
    const float* x[2];          // pre activity, pre-local, [buffer]
    const float* s;             // row scale factors of a lazily normalised matrix, or nullptr
    double* partials;           // [nseg][ld] segment partials of the step
    uint32_t ld, n_pre, n_post, pre_off, post_off;
    uint32_t rb, nseg, ncol;    // rows per chunk, segments of kPWaves*rb rows, column tiles
    uint32_t post_group;        // index into PPlan::g
    uint32_t pre_group;
    uint32_t ts;                // index into PPlan::ts: the trace set whose x it reads
    uint32_t plast;             // index into PPlan::pl, or kPNone
};

struct PPlast {
    uint32_t kind;              // PlastKind
    uint32_t lazy;
    PlasticityParams p;
    uint32_t conn;
    uint32_t pre_off, n_pre, post_off, n_post;
    float* pre[2];              // [buffer], pre-local (triplet x_fast / iSTDP pre)
    float* post[2];             // post-local (y_fast / post)
    float* tpre[2];             // triplet slow traces (nullptr otherwise)
    float* tpost[2];
    float pre_decay, post_decay, tpre_decay, tpost_decay;
    float* s;                   // lazy: [n_pre] row scales, R [n_pre] FP64 row sums
    double* R;
    const float* target;        // lazy: [n_pre] normalisation targets
    double* dR;                 // lazy: [ncol][n_pre] the tiles' parts of the row updates
};

struct PTraceSet {
    uint32_t pre_off, n_pre;
    float decay_r, decay_d, K;
    float* r;
    float* d;
    float* x[2];
};

struct PGroup {
    NeuronGroupView v;
    uint32_t src_e, src_i;      // connections whose partials make g_e / g_i (kPNone: none)
    uint32_t ncol;              // column tiles of 256 over the group
    uint32_t col_base;          // index of its first column tile in the column counters
    uint32_t expect;            // items per column tile (sum of nseg of the sources)
};

struct PItem {
    uint16_t conn, col;
    uint32_t seg;
};

struct PPlan {
    PGroup g[kMaxGroups];
    PConn c[kMaxConns];
    PPlast pl[kMaxPlast];
    PTraceSet ts[kMaxTraceSets];
    uint32_t n_groups, n_conn, n_pl, n_ts, total_neurons, words;
    uint32_t n_coltiles;        // column tiles of all groups (evolve tails per step)
    uint32_t lazy_pl;           // the lazily normalised plasticity, or kPNone
    uint32_t n_rowtails;        // segments of its matrix (row tails per step)
    // schedule
    uint32_t resident_rows;     // 0: weights streamed each step; else one item per block, its
                                // rows held in registers (a chunk of at most this many rows)
    const PItem* items;
    const uint32_t* block_begin;    // [blocks + 1]
    uint32_t blocks;
    // counters (zeroed before each launch)
    uint32_t* col_count;        // [n_coltiles]
    uint32_t* row_count;        // [n_rowtails]
    uint32_t* tails_done;
    uint32_t* status;           // the state's error word (2: a wait ran past its limit)
    uint32_t wait_limit;        // polls before a wait gives up
    // state
    const bn_device_clock* clock;
    bn_device_clock* clock_rw;
    uint32_t* spikes;           // per-neuron spike flags (global ids)
    uint32_t* bits[3];          // spike bits [buffer], bits[0] the state's own
    uint32_t* spike_ring;       // [interval][words] or nullptr
    float* v_ring;              // [interval][n_vrec] or nullptr
    const int32_t* vrec_index;
    uint32_t n_vrec, interval;
    // stimulus: events of step s are [ev_start[s], ev_start[s+1]) of the lists
    const uint32_t* ev_start;
    const uint32_t* ev_ids;
    const uint32_t* ev_offsets;
    const float* ev_amps;
    uint32_t ev_steps, ev_count;
    float* i_stimulus;          // global
    // Resident weights: tagged 64-bit words (a 32-bit value and the step that produced it), in
    // three slots by step (kernels_persist_ll.hip).
    const uint32_t* block_tile;         // [blocks] column tile the block evolves, or kPNone
    uint64_t* ll_x[kMaxTraceSets];      // [3][n_pre] x of the trace set's neurons
    uint64_t* ll_pl[kMaxPlast][4];      // [3][n] pre, slow pre ([3][n_pre]), post, slow post ([3][n_post])
    uint64_t* ll_spk;                   // [3][n_coltiles][8] spike flags of a tile, 32 a word
    uint64_t* ll_part[kMaxConns];       // [3][nseg][2][ld] segment partials (low, high words)
    uint64_t* ll_dr;                    // [3][ncol][2][n_pre] lazy matrix: row tile sums (low, high words)
    uint32_t* done;                     // [0] steps finished, summed over the blocks; [1] first
                                        // evolves finished
};

// Limits of the resident-weight kernel.
constexpr uint32_t kLSegMax = kPWaves * 32;   // rows of a segment (chunks of at most 32 rows)
constexpr uint32_t kLPartMax = 32;            // segment partials of a group's two sources
constexpr uint32_t kLDrJobs = 6;              // row tile sums polled per thread

// n steps from the state (cooperative launch of host_plan.blocks x kPThreads threads; the
// counters, the spike bits of the first two evolves and their ring slots must be zero).
int launch_persistent(const PPlan* d_plan, const PPlan& host_plan, uint32_t n, uint32_t nontemporal, bn_stream_t s);
// Largest number of co-resident blocks of the persistent kernel (resident_rows 0: streamed
// weights; 16, 24: resident weights) on the current device.
int persistent_max_blocks(uint32_t resident_rows, uint32_t* blocks);
int persistent_ll_max_blocks(uint32_t resident_rows, uint32_t* blocks);
int launch_persistent_ll(const PPlan* d_plan, const PPlan& host_plan, uint32_t n, bn_stream_t s);
constexpr uint32_t kPResidentRows[] = {16, 24};

}  // namespace bn
