/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the reference for checks / CPU baseline; the product
 * never links or calls it.
 *
 * A CPU (C++/OpenMP) restatement of the reference's per-step semantics
 * (cpp_src/core/system.cpp:43-63 order: evolve -> propagate -> stimulus -> plasticity -> monitors)
 * with the engine's documented choices where the reference is unpinned (cuRAND, cuBLAS):
 *   - background noise: Philox4x32-10 counter (step, 0, global id, instance), key = rng_seed,
 *     Poisson by inversion of a floor(2^32 * CDF) table;
 *   - expf: Cody–Waite reduction + degree-7 Taylor (same algorithm as the kernels);
 *   - GEMV: FP64 accumulation of exact FP32 products, rows summed sequentially inside chunks of
 *     gemv_row_block rows; the chunk sums cut into 8 contiguous parts of ceil(chunks/8), each
 *     part summed in chunk order from 0.0, the parts added to (double)g in order, one rounding
 *     to FP32;
 *   - triplet row normalisation sum: 256 interleaved FP64 partials, xor butterfly inside each
 *     64-lane group (32,16,8,4,2,1), the four group sums added in order.
 *   - lazy normalisation (lazy_normalize = 1): the triplet matrix is kept as W = diag(s) M.  A
 *     touched entry is updated in true units, w' = clamp(s_i*M_ij + delta), M_ij = w'/s_i (rows
 *     with s_i = 0 stay untouched); R_i = sum_j M_ij is kept in FP64: rows whose pre neuron
 *     spiked (LTD) are re-summed in the normalisation order above, other rows add their LTP
 *     changes (double)M'_ij - (double)M_ij: the list of spiking posts (increasing post order)
 *     is cut into 4 contiguous parts of ceil(count/4), each part's changes summed in list order
 *     from 0.0, the parts added to R_i in order; then s_i = fl32(target_i / R_i) when R_i > 0.
 *   - persistent step (gemv_part_chunks = 8, lazy_tile_order = 1): the chunk sums are combined in
 *     parts of exactly 8 chunks (the 8 waves of one block, summed in chunk order); the lazy row
 *     sums follow the column tiles of 256 columns: in tile q, lane l (0..63) sums columns
 *     q*256 + 4l .. q*256 + 4l + 3 in order from 0.0, the 64 lane sums are combined by the xor
 *     butterfly (32,16,8,4,2,1), giving T_q; a row whose pre neuron spiked gets R_i = 0.0 + T_0 +
 *     T_1 + ... over the new row; another row sums its LTP changes (double)M'_ij - (double)M_ij
 *     the same way per tile and gets R_i = R_i + dT_0 + dT_1 + ....
 *   - resident step (resident_columns = C > 0, resident_e_waves = WE): the GEMV of a post column
 *     sums the excitatory pre rows over T = 64*WE threads and the inhibitory ones over
 *     T = 64*(8 - WE): thread t (wave t / 64, lane t % 64) sums rows t, t + T, t + 2T, ... in
 *     order from 0.0, the 64 lane sums of a wave are combined by the xor butterfly
 *     (32,16,8,4,2,1), and g' = fl32(((double)g + wave_0) + wave_1 + ...) in wave order.  The lazy
 *     row sums follow blocks of C post columns (block b: columns C*b .. C*b + C - 1): a row whose
 *     pre neuron spiked gets R_i = the xor butterfly over 64 lane sums, lane l summing P_l,
 *     P_{l+64}, ... in order from 0.0, P_b the block's new entries summed in column order from
 *     0.0; another row gets R_i = R_i + D_b + ... over the blocks b holding a spiking post, D_b
 *     its LTP changes (double)M'_ij - (double)M_ij summed in column order.
 * precision 32 reproduces the engine bit for bit; precision 64 runs every quantity in FP64
 * (libm exp, plain sequential sums) as an independent accuracy reference.
 *
 *   - pair STDP (e_e_pair_stdp = 1): the reference's StdpPlasticity on E->E
 *     (kernel_stdp_plasticity_update.cu:16-137): decay pre then post traces with flush, LTP of
 *     every column of a spiking post (w += ltp_A * pre_i, clamp), LTD of every row of a spiking
 *     pre (w += (-1 * ltd_A) * post_j, clamp), self entries i == j skipped, then the +1 trace
 *     increments; no triplet terms, no normalisation.
 * Parity pins (tests/test_oracle_*.py): Philox known-answer vectors of the published method and
 * rocRAND's engine; expf against libm/FP64 within 2 ulp; Poisson table against scipy's CDF;
 * traces against the closed form K(d^k - r^k); EIF single step against an FP64 hand
 * computation; GEMV against FP64 numpy; plasticity deltas for isolated pre/post pairs; row
 * normalisation against numpy column sums; connectivity against the reference generator.
 */
#ifndef BURYN_ORACLE_H
#define BURYN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_group_params {
    float mean_background_excitatory_spikes;
    float mean_background_inhibitory_spikes;
    float background_excitatory_postsynaptic_r_trace_decay;
    float background_excitatory_postsynaptic_d_trace_decay;
    float background_inhibitory_postsynaptic_r_trace_decay;
    float background_inhibitory_postsynaptic_d_trace_decay;
    float background_excitatory_scaling_factor;
    float background_inhibitory_scaling_factor;
} oracle_group_params;

typedef struct oracle_params {
    uint32_t num_excitatory_neurons;
    uint32_t num_inhibitory_neurons;
    int32_t rng_seed;
    uint32_t rng_instance;
    float timestep;
    float g_leak;
    float membrane_capacitance;
    float depolarization_slope_factor;
    float v_effective_spike_threshold;
    float v_excitatory_reversal;
    float v_inhibitory_reversal;
    float v_threshold;
    float v_reset;
    float v_rest;
    uint32_t refractory_steps;
    oracle_group_params group[2]; /* 0 = E, 1 = I */
    float tau_decay_e, tau_rise_e, tau_decay_i, tau_rise_i;
    float postsynaptic_kernel_normalization_value;
    uint32_t e_e_plasticity_enabled;
    float e_e_ltp_A, e_e_ltd_A, e_e_ltp_B, e_e_ltd_B;
    float e_e_tau_ltp, e_e_tau_ltd, e_e_triplet_tau_ltp, e_e_triplet_tau_ltd;
    float e_e_weight_min, e_e_weight_max;
    uint32_t i_e_plasticity_enabled;
    float i_e_ltp_A, i_e_ltd_A, i_e_ltd_B, i_e_tau_ltp, i_e_tau_ltd, i_e_weight_min, i_e_weight_max;
    uint32_t gemv_row_block[4];   /* EE, EI, IE, II */
    uint32_t lazy_normalize;      /* 1: W_EE = diag(s) M with per-row scale factors */
    uint32_t dense_gemv;          /* 1: visit every pre row (no skip of x = 0 rows; same result) */
    uint32_t gemv_part_chunks[4]; /* chunks per combination part (0: ceil(chunks / 8)) */
    uint32_t lazy_tile_order;     /* 1: lazy row sums in the column-tile order (see above) */
    uint32_t e_e_pair_stdp;       /* 1: pair STDP on E->E (kernel_stdp_plasticity_update.cu) */
    uint32_t resident_columns;    /* > 0: the resident step's orders (see above), post columns per block */
    uint32_t resident_e_waves;    /* its waves (of 8) holding excitatory pre rows */
} oracle_params;

typedef struct oracle_sim oracle_sim;

/* Weights: dense row-major [pre][post] (EE N_E x N_E, EI N_E x N_I, IE N_I x N_E, II N_I x N_I).
 * Stimulus arrays as produced by the schedule generator (may be NULL when num_events == 0 and
 * stim_counts == NULL). precision: 32 or 64. */
oracle_sim* oracle_create(const oracle_params* p, int precision, const float* w_ee, const float* w_ei,
                          const float* w_ie, const float* w_ii, const uint32_t* stim_counts,
                          uint32_t num_timesteps, const uint32_t* stim_ids, const uint32_t* stim_offsets,
                          const float* stim_amps, uint32_t num_events);
void oracle_destroy(oracle_sim* s);
/* Runs n steps; returns 0, or 1 when the stimulus cursor runs past the event tables. */
int oracle_step(oracle_sim* s, uint32_t n);
uint32_t oracle_clock_step(const oracle_sim* s);
/* Global-id arrays of length N (any pointer may be NULL). g_* are after the last propagate. */
void oracle_get_state(const oracle_sim* s, float* v, float* ge, float* gi, float* istim, uint32_t* refr,
                      uint32_t* spikes, float* bg_ge, float* bg_gi);
void oracle_get_state64(const oracle_sim* s, double* v, double* ge, double* gi);
/* True weights (lazy mode: s_i * M), dense [pre][post]. */
void oracle_get_weights(const oracle_sim* s, uint32_t idx, float* dst);
void oracle_get_connection_traces(const oracle_sim* s, uint32_t idx, float* r, float* d, float* x);
/* plasticity_offset 0/1 in enable order; arrays of length N (global ids). */
void oracle_get_plasticity_traces(const oracle_sim* s, uint32_t plasticity_offset, float* pre, float* post,
                                  float* trip_pre, float* trip_post);

/* Differentiable window (adjoint.cpp): FP64 forward of the plasticity-free, stimulus-free step
 * from a given state, the window loss, and its gradient.  State after the step before the
 * window, global ids (length N) or per pre group (r_e, d_e: N_E; r_i, d_i: N_I). */
typedef struct oracle_window_state {
    const float* V;
    const uint32_t* refr;
    const float* bg_ge;
    const float* bg_gi;
    const float* bre;
    const float* bde;
    const float* bri;
    const float* bdi;
    const float* r_e;
    const float* d_e;
    const float* r_i;
    const float* d_i;
    uint32_t clock_step;          /* step index of the window's first step (noise counter) */
} oracle_window_state;

typedef struct oracle_window_spec {
    uint32_t steps;
    uint32_t n_vrec;
    const uint32_t* vrec_ids;      /* excitatory ids whose V enters the loss */
    double rate_tau;               /* s, smoothing of the population rates */
    double surrogate_beta;         /* V; 0 = spikes held fixed (exact derivative) */
    const double* target_rates;    /* [steps][2] Hz (E, I), or NULL (zero) */
    const double* target_v;        /* [steps][n_vrec] V, or NULL (zero) */
    const uint8_t* forced_spikes;  /* [steps][N] raster to follow instead of thresholding, or NULL */
} oracle_window_spec;

/* W[c]: dense [pre][post] FP64 (EE, EI, IE, II).  dW = NULL: forward and loss only.  Outputs
 * may be NULL.  rates_out [steps][2], v_out [steps][n_vrec], spikes_out [steps][N]. */
int oracle_window(const oracle_params* p, const double* const* W, double g_leak, double tau_decay_e,
                  const oracle_window_state* s0, const oracle_window_spec* spec, double* loss, double* const* dW,
                  double* dg_leak, double* dtau_decay_e, double* rates_out, double* v_out, uint8_t* spikes_out);

/* A recorded forward trajectory of the window (for instance the FP32 step's): per step t the
 * state before the step and what the step produced. */
typedef struct oracle_window_trajectory {
    const float* Vprev;            /* [steps][N] V before the step */
    const float* ge;               /* [steps][N] conductances the step uses */
    const float* gi;
    const uint8_t* refr;           /* [steps][N] 1 = refractory at the step */
    const uint8_t* spikes;         /* [steps][N] */
    const float* d_e;              /* [steps][N_E] connection traces before the step */
    const float* r_e;
    const float* d_i;              /* [steps][N_I] */
    const float* r_i;
    const float* v_rec;            /* [steps][n_vrec] V after the step */
} oracle_window_trajectory;

/* The records of oracle_window's FP64 forward, rounded to FP32 (trajectory layout above;
 * outputs sized for spec->steps). */
int oracle_window_records(const oracle_params* p, const double* const* W, double g_leak, double tau_decay_e,
                          const oracle_window_state* s0, const oracle_window_spec* spec, float* Vprev, float* ge,
                          float* gi, uint8_t* refr, uint8_t* spikes, float* d_e, float* r_e, float* d_i, float* r_i,
                          float* v_rec);

/* The loss and FP64 reverse pass of oracle_window along a given trajectory instead of the
 * oracle's own FP64 forward (spec->forced_spikes unused). */
int oracle_window_along(const oracle_params* p, const double* const* W, double g_leak, double tau_decay_e,
                        const oracle_window_trajectory* traj, const oracle_window_spec* spec, double* loss,
                        double* const* dW, double* dg_leak, double* dtau_decay_e);

/* Primitives (for the pinning tests). */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float oracle_expf(float x);
/* Returns the table length (0 for mean <= 0); fills table[0..len). */
uint32_t oracle_poisson_table(float mean, uint32_t* table, uint32_t capacity);

#ifdef __cplusplus
}
#endif

#endif
