// Plans of the differentiable window: the plasticity-free, stimulus-free step run forward from
// a fixed state with per-step records, and its reverse pass (adjoint) for the gradient of the
// window loss with respect to the four weight matrices, g_leak and tau_decay_e.
#pragma once

#include <stdint.h>

#include "bn_step.h"

namespace bn {

// The differentiated scalars, read from device memory so that captured graphs stay valid when
// they change.
struct WinParams {
    float g_leak;
    float decay_e_d;            // exp(-dt / tau_decay_e)
};

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
    const WinParams* prm;
    uint32_t NE, NI, N, slot, step;   // slot in the segment, absolute step (noise counter)
    float decay_e_r, decay_i_r, decay_i_d, K;
    bn_device_clock clock;      // Philox key (clock_step unused: `step` is the counter)
    uint32_t* counts;           // [2] spike counts of E, I this step, or nullptr
    float* v_rec;               // [n_vrec] V of the recorded neurons this step, or nullptr
    const int32_t* vrec_slot;   // [N] index in v_rec or -1
    uint8_t* raster;            // [N] spikes this step, or nullptr
};

// Copy of the whole window state (checkpoints).
struct WinCopyPlan {
    WinState dst, src;
    uint32_t NE, NI, N;
};

// One reverse step: lambda x = W lambda g of the next step (row products), then the adjoint of
// every neuron's connection traces (as a pre neuron) and EIF update (as a post neuron).
struct WinReversePlan {
    WinGroup g[2];
    WinRecords rec;
    const WinParams* prm;
    const float* W[4];          // EE, EI, IE, II, row stride ld[c]
    uint32_t ld[4];
    uint32_t NE, NI, N, slot;
    float decay_e_r, decay_i_r, decay_i_d, K;
    float theta, beta;          // beta = 0: spikes held fixed
    double rate_coef_e, rate_coef_i;  // (1 - a) / (N_p dt)
    const double* lrho;         // [2] lambda of the smoothed rates (E, I) at this step
    const double* dv_direct;    // [n_vrec] dL/dV of the recorded neurons at this step
    const int32_t* vrec_slot;
    const double* lg_in_e;      // [N] lambda g of the next step (zero after the last step)
    const double* lg_in_i;
    double* lg_out_e;           // [N] lambda g of this step (a different buffer)
    double* lg_out_i;
    double* carry_V;            // [N]
    double* carry_d;            // [N] (pre neuron = global id)
    double* carry_r;
    double* acc_gleak;          // [N]
    double* acc_alpha;          // [NE]
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
int launch_win_copy(const WinCopyPlan& p, bn_stream_t s);
int launch_win_reverse(const WinReversePlan& p, bn_stream_t s);
int launch_win_outer(const WinOuterPlan& p, bn_stream_t s);

}  // namespace bn
