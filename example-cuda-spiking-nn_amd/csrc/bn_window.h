// Plans of the differentiable window: the plasticity-free, stimulus-free step run forward from
// a fixed state with per-step records, and its reverse pass (adjoint) for the gradient of the
// window loss with respect to the four weight matrices, g_leak and tau_decay_e.
#pragma once

#include <stdint.h>

#include "bn_step.h"

namespace bn {

struct WinGroup {
    NeuronGroupView v;          // parameters (g_leak = the differentiated value) and Poisson tables
    FinalizeSource fin_e, fin_i;
};

// State of the window (global ids; traces per pre group).
struct WinState {
    float* V;
    uint32_t* refr;
    float* bg_ge;               // background conductance of the last step (the base of g)
    float* bg_gi;
    float* bre;
    float* bde;
    float* bri;
    float* bdi;
    float* dE;                  // [NE] connection traces of the excitatory pre group
    float* rE;
    float* xE;
    float* dI;                  // [NI]
    float* rI;
    float* xI;
};

// Records of one segment (slot = step within the segment), [slot][neuron].
struct WinRecords {
    float* Vprev;
    float* ge;                  // g at the step (background + propagate)
    float* gi;
    float* s;                   // spike 0/1
    uint8_t* refr;              // refractory at the step (no gradient)
    float* dE;                  // traces at the start of the step, [slot][NE] / [slot][NI]
    float* rE;
    float* dI;
    float* rI;
    float* le;                  // lambda g_e of the step, [slot][N] (backward)
    float* li;
};

struct WinForwardPlan {
    WinGroup g[2];
    WinState st;
    WinRecords rec;             // rec.Vprev == nullptr: no records (plain forward)
    uint32_t NE, NI, N, slot, step;   // slot in the segment, absolute step (noise counter)
    float decay_e_r, decay_e_d, decay_i_r, decay_i_d, K;
    bn_device_clock clock;      // Philox key (clock_step unused: `step` is the counter)
    uint32_t* counts;           // [2] spike counts of E, I this step (window-global array offset)
    float* v_rec;               // [n_vrec] V of the recorded neurons this step
    const int32_t* vrec_slot;   // [N] index in v_rec or -1
    uint8_t* raster;            // [N] spikes this step (window-global array offset) or nullptr
};

struct WinBackwardPlan {
    WinGroup g[2];
    WinRecords rec;
    uint32_t NE, NI, N, slot;
    float decay_e_r, decay_e_d, decay_i_r, decay_i_d, K;
    float theta, beta;          // beta = 0: spikes held fixed
    float rate_coef_e, rate_coef_i;   // (1 - a) / (N_p dt)
    float lrho_e, lrho_i;       // lambda of the smoothed rates at this step
    const float* dv_direct;     // [n_vrec] dL/dV of the recorded neurons at this step
    const int32_t* vrec_slot;
    const double* lx;           // [N] lambda x of this step (row products with lambda g of the next)
    float* carry_V;             // [N]
    float* carry_d;             // [N] (pre neuron = global id)
    float* carry_r;
    float* lg_e;                // [N] lambda g of this step (output; read by the next row products)
    float* lg_i;
    double* acc_gleak;          // [N]
    double* acc_alpha;          // [NE]
};

struct WinRowPlan {
    const float* W[4];          // EE, EI, IE, II, row stride ld[c]
    uint32_t ld[4];
    uint32_t NE, NI;
    const float* lg_e;          // [N] lambda g_e of the next step
    const float* lg_i;
    double* lx;                 // [N]
};

struct WinOuterPlan {
    float* dW;                  // [pre][ld] accumulated
    uint32_t ld, n_pre, n_post, slots;
    const float* d;             // [slot][pre_stride] traces at the start of each step
    const float* r;
    uint32_t pre_stride, pre_offset;
    const float* lam;           // [slot][lam_stride] lambda g of each step (global ids)
    uint32_t lam_stride, post_offset;
};

int launch_win_forward(const WinForwardPlan& p, bn_stream_t s);
int launch_win_rows(const WinRowPlan& p, bn_stream_t s);
int launch_win_backward(const WinBackwardPlan& p, bn_stream_t s);
int launch_win_outer(const WinOuterPlan& p, bn_stream_t s);

}  // namespace bn
