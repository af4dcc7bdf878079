// Internal (C++) interface of the resident step: one cooperative launch runs n whole steps of a
// network whose four weight matrices fit in the registers of the co-resident blocks
// (kernels_resident.hip).
//
// Ownership.  Block b owns C post columns of one post group (E-post blocks first, then I-post
// blocks) and holds, in the registers of its threads, every pre row of those columns: the
// excitatory rows on the first WE waves (thread t of those 64*WE threads: rows t, t + 64*WE, ...)
// and the inhibitory rows on the other 8 - WE waves.  The block evolves its own C neurons, so
// the propagate -> evolve hand-off of a step stays inside the block.
//
// Replicas.  Every quantity a row needs that depends only on the pre neuron's spike history
// (connection traces r, d and x = d - r, the plasticity pre traces) is kept by each thread for
// its own rows and advanced from the spike bits alone, bit for bit as the neuron's evolve
// would.  So a step exchanges only:
//   spk   per block and step, one tagged word: the spike bits of its C neurons;
//   sx    per lazily normalised row and step, one tagged word: its scale s_i, published by the
//         block owning neuron i (the row's R_i, s_i and target live there);
//   dr    per E-post block, row and step, two tagged words (an FP64 value, low half first): the
//         block's part of the row-sum update of the lazy matrix, polled by the row's owner.  For
//         a row whose pre neuron did not spike, when one of the block's columns spiked: the sum
//         of those columns' LTP changes ([slot][block][row]); for a row whose pre neuron spiked:
//         the sum of its new entries over the block's columns ([slot][row][block], so that the
//         owner's loads of one row's parts are contiguous).
// A tagged word is a 32-bit value and the step that produced it in one 64-bit agent-coherent
// store; readers poll the words they need, in three slots by step.  A block rewrites a slot only
// after every block has published its spike word of a later step, which each block does only
// after reading the slot (so no counters and no barriers across blocks).  Every poll is
// bounded: past the limit a block raises the state's error word (2) and every block leaves.
//
// Arithmetic (restated by the oracle, resident_columns / resident_e_waves): the GEMV of a column
// sums the rows of each thread in order from 0.0 in FP64, the 64 lane sums of a wave by the xor
// butterfly (32,16,8,4,2,1), and g' = fl32(((double)g + wave_0) + wave_1 + ...); the lazy row
// sums by blocks of C columns in block order; everything else as the fused step.
#pragma once

#include <stdint.h>

#include "bn_step.h"

namespace bn {

constexpr int kRThreads = 512;          // threads per block (8 waves)
constexpr int kRWaves = kRThreads / 64;
constexpr int kRCols = 16;              // post columns per block
constexpr int kRMaxRows = 8;            // rows per thread at most
constexpr uint32_t kRMaxBlocks = 256;   // blocks at most (the spike words fit one wave's polls)
constexpr uint32_t kRNone = 0xFFFFFFFFu;

struct RPlan {
    // Network: E neurons are global ids [0, ne), I neurons [ne, ne + ni).
    uint32_t ne, ni, n;
    uint32_t be, bi, blocks;      // E-post blocks, I-post blocks, all
    uint32_t e_waves;             // waves holding E rows (the other waves hold I rows)
    uint32_t rows;                // rows per thread (the kernel's template height)
    NeuronGroupView g[2];         // E, I
    // Weights [pre][ld]: EE, EI, IE, II.
    float* W[4];
    uint32_t ld[4];
    // Connection traces of the two pre groups (E: EE and EI, I: IE and II), pre-local.
    float decay_r[2], decay_d[2], K[2];
    float* r[2];
    float* d[2];
    float* x[2];
    // E->E plasticity (kPlastTriplet, kPlastStdp or kRNone) and I->E iSTDP (kPlastIstdp or kRNone).
    uint32_t ee_kind, ee_lazy, ie_kind;
    PlasticityParams ee_p, ie_p;
    float ee_dec[4];              // decays of pre, slow pre, post, slow post (E)
    float ie_dec[2];              // decays of pre (I), post (E)
    float* ee_tr[4];              // their state buffers (pre-local / post-local, no last +1)
    float* ie_tr[2];
    float* s;                     // lazy E->E: row scales, FP64 row sums, targets
    double* R;
    const float* target;
    // State.
    bn_device_clock* clock;
    uint32_t* spikes;             // per neuron
    uint32_t* bits;               // spike bits of the last step [words]
    uint32_t words;
    uint32_t* spike_ring;         // [interval][words] or nullptr
    float* v_ring;                // [interval][n_vrec] or nullptr
    const int32_t* vrec_index;
    uint32_t n_vrec, interval;
    // Stimulus events of step s: [ev_start[s], ev_start[s + 1]) of the lists.
    const uint32_t* ev_start;
    const uint32_t* ev_ids;
    const uint32_t* ev_offsets;
    const float* ev_amps;
    uint32_t ev_steps, ev_count;
    // Exchange (tagged 64-bit words, three slots by step).
    uint64_t* spk;                // [3][blocks]
    uint64_t* sx;                 // [3][ne]
    uint64_t* dr;                 // [3][be][ne][2]: quiet rows' parts
    uint64_t* drs;                // [3][ne][be][2]: spiking rows' parts
    uint32_t* status;             // the state's error word
    uint32_t wait_limit;          // polls before a wait gives up
};

// Rows per thread and waves of E rows for a network (0 rows: it does not fit a block's threads).
void resident_shape(uint32_t ne, uint32_t ni, uint32_t* rows, uint32_t* e_waves);
// Largest number of co-resident blocks of the kernel for `rows` rows per thread.
int resident_max_blocks(uint32_t rows, uint32_t* blocks);
// n steps from the state (one launch of host_plan.blocks blocks, cooperative or ordinary), the
// clock advanced by n steps with its stimulus cursor.
int launch_resident(const RPlan* d_plan, const RPlan& host_plan, uint32_t n, bool cooperative, bn_stream_t s);

}  // namespace bn
