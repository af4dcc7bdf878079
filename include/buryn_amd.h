/*
 * buryn_amd.h — C ABI of the MI355X-native per-step engine for the conductance-based EIF
 * spiking network of paulazoo/example-cuda-spiking-nn ("buryn").
 *
 * Everything here is plain C: POD structs, plain pointers, sizes, and an opaque HIP stream
 * handle.  No C++ or torch type crosses this boundary, and no exception either: every entry
 * point returns 0 on success or a nonzero code, and bn_last_error() returns the message.
 * The C++ host layer (csrc/host) rethrows these as std::runtime_error("<what>: <msg>"),
 * the same convention as the reference's cuda_check (cpp_src/gpu_core/cuda_utils.h:10-20).
 *
 * Layer map (reference → here):
 *   kernel_*_launch(XView)            → bn_*(XView, bn_stream_t)          (drop-in launchers)
 *   GpuSimulationState + components   → csrc/host C++ classes (same names) + bn_sim_* below
 *   System::step_the_clock / run      → bn_sim_run (fused, graph-captured step) and
 *                                       bn_sim_step_components (component-by-component step)
 *
 * Field names of the *View structs follow the reference exactly, including the misspelled
 * ConnectionView::pre_group_is_inhibtory.  Fields marked "extension" are additions this
 * engine needs (row stride, scratch for the deterministic GEMV, counter-based RNG clock).
 */
#ifndef BURYN_AMD_H
#define BURYN_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t is `struct ihipStream_t*`; this alias keeps the header free of HIP includes. */
typedef struct ihipStream_t* bn_stream_t;

/* ----------------------------------------------------------------------------------------
 * Error handling
 * ------------------------------------------------------------------------------------- */
enum {
    BN_OK = 0,
    BN_ERR_INVALID_ARGUMENT = 1001, /* std::invalid_argument in the reference */
    BN_ERR_OUT_OF_RANGE = 1002,     /* std::out_of_range in the reference */
    BN_ERR_RUNTIME = 1003,          /* std::runtime_error in the reference */
    BN_ERR_HIP = 1004               /* a HIP runtime call failed; message carries hipGetErrorString */
};
/* Message of the last failing call on this thread ("<what>: <detail>"). */
const char* bn_last_error(void);

/* ----------------------------------------------------------------------------------------
 * Device clock (replaces the per-neuron curandStatePhilox4_32_10_t array,
 * reference: cpp_src/gpu_core/gpu_simulation_state.cu:50, kernel_init_neuron_rng_states.cu:6-26)
 *
 * The background noise is drawn from a stateless Philox4x32-10 keyed by rng_key (the run's
 * rng_seed) with counter (clock_step, stream, global neuron id, instance).  The clock lives
 * in device memory so a whole step can be captured in a hipGraph and replayed: kernels read
 * clock_step, the step-tail kernel advances it.  next_event_idx is the stimulus cursor that
 * the reference keeps on the host (stimulus.h:42-44).
 * ------------------------------------------------------------------------------------- */
typedef struct bn_device_clock {
    uint32_t clock_step;      /* index of the step being executed */
    uint32_t next_event_idx;  /* first stimulus event not yet applied */
    uint32_t rng_key[2];      /* Philox key: low and high word of (uint64)rng_seed */
} bn_device_clock;

/* Philox counter word 1 ("stream") values. */
enum { BN_RNG_STREAM_BACKGROUND = 0 };

/* ----------------------------------------------------------------------------------------
 * Views — reference cpp_src/<component>/<name>_view.h and <name>_params.h
 * ------------------------------------------------------------------------------------- */

/* reference: cpp_src/neuron_groups/neuron_group_view.h:8-49 */
typedef struct NeuronGroupView {
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
    bool is_inhibitory;
    float mean_background_excitatory_spikes;
    float mean_background_inhibitory_spikes;
    float background_excitatory_postsynaptic_r_trace_decay;
    float background_excitatory_postsynaptic_d_trace_decay;
    float background_inhibitory_postsynaptic_r_trace_decay;
    float background_inhibitory_postsynaptic_d_trace_decay;
    float background_excitatory_scaling_factor;
    float background_inhibitory_scaling_factor;

    uint32_t num_neurons;
    uint32_t offset;

    float* membrane_potentials;
    float* g_excitatory;
    float* g_inhibitory;
    float* i_stimulus;
    uint32_t* refractory_counter;
    uint32_t* spikes;
    float* background_g_excitatory_synaptic_conductances;
    float* background_g_inhibitory_synaptic_conductances;
    float* background_excitatory_postsynaptic_r_traces;
    float* background_excitatory_postsynaptic_d_traces;
    float* background_inhibitory_postsynaptic_r_traces;
    float* background_inhibitory_postsynaptic_d_traces;
    /* extension: replaces `void* neuron_rng_states` (counter-based RNG, see bn_device_clock) */
    const bn_device_clock* rng_clock;
    /* extension: Poisson inverse-CDF tables built by bn_poisson_table() (length 0 = mean 0) */
    const uint32_t* poisson_table_excitatory;
    const uint32_t* poisson_table_inhibitory;
    uint32_t poisson_table_excitatory_len;
    uint32_t poisson_table_inhibitory_len;
    /* extension: Philox counter word 3; 0 for a single network, the instance id when batched */
    uint32_t rng_instance;
} NeuronGroupView;

/* reference: cpp_src/connections/connection_view.h:8-31 (cublas_handle dropped) */
typedef struct ConnectionView {
    float postsynaptic_r_trace_decay;
    float postsynaptic_d_trace_decay;
    float postsynaptic_kernel_normalization_value;
    bool pre_group_is_inhibtory; /* sic — reference field name */

    uint32_t weight_matrix_idx;
    uint32_t pre_group_offset;
    uint32_t post_group_offset;
    uint32_t pre_group_num_neurons;
    uint32_t post_group_num_neurons;

    float* weight_matrix;  /* row-major [pre][post], row stride weight_matrix_ld */
    float* postsynaptic_r_traces;
    float* postsynaptic_d_traces;
    float* summed_postsynaptic_conductances;
    float* post_group_g_excitatory;
    float* post_group_g_inhibitory;
    uint32_t* spikes;
    /* extension: row stride in floats (0 = post_group_num_neurons, the reference layout).
     * Must be a multiple of 4 with 16-byte aligned rows for the vectorised GEMV, else the
     * launcher falls back to a scalar-load GEMV. */
    uint32_t weight_matrix_ld;
    /* extension: deterministic GEMV scratch, ceil(pre/gemv_row_block) * ld doubles */
    double* gemv_partials;
    uint32_t gemv_row_block;
} ConnectionView;

/* reference: cpp_src/stimuli/stimulus_view.h:8-16 */
typedef struct StimulusView {
    float* g_excitatory;
    float* g_inhibitory;
    float* i_stimulus;
    uint32_t* timestep_event_count;
    uint32_t* stimulus_event_neuron_ids;
    uint32_t* stimulus_event_post_group_offsets;
    float* stimulus_event_amplitude_changes;
    /* extension: lengths of the two tables above, used by the device-side cursor */
    uint32_t num_timesteps;
    uint32_t num_events;
} StimulusView;

/* reference: cpp_src/plasticities/plasticity_view.h:8-38 */
typedef struct PlasticityView {
    float ltp_A;
    float ltd_A;
    float ltp_B;
    float ltd_B;
    float pre_trace_decay;
    float post_trace_decay;
    float triplet_pre_trace_decay;
    float triplet_post_trace_decay;
    float weight_min;
    float weight_max;

    uint32_t weight_matrix_idx;
    uint32_t plasticity_offset;
    uint32_t pre_group_offset;
    uint32_t post_group_offset;
    uint32_t pre_group_num_neurons;
    uint32_t post_group_num_neurons;

    float* pre_traces;
    float* post_traces;
    float* triplet_pre_traces;
    float* triplet_post_traces;
    uint32_t* pre_spikes;
    uint32_t* post_spikes;
    float* weight_matrix;
    float* pre_weight_sums;
    /* extension: row stride of weight_matrix in floats (0 = post_group_num_neurons) */
    uint32_t weight_matrix_ld;
} PlasticityView;

/* reference: cpp_src/plasticities/plasticity_params.h:7-20 */
typedef struct PlasticityParams {
    float ltp_A;
    float ltd_A;
    float ltp_B;
    float ltd_B;
    float pre_trace_decay;
    float post_trace_decay;
    float triplet_pre_trace_decay;
    float triplet_post_trace_decay;
    float weight_min;
    float weight_max;
    uint32_t plasticity_offset;
} PlasticityParams;

/* reference: cpp_src/monitors/spike_recorder_view.h:8-16 */
typedef struct SpikeRecorderView {
    uint32_t num_neurons;
    uint32_t offset;
    uint32_t* spikes;
    uint32_t* spike_history;
} SpikeRecorderView;

/* reference: cpp_src/monitors/voltage_recorder_view.h:8-16 */
typedef struct VoltageRecorderView {
    uint32_t num_neurons;
    uint32_t offset;
    float* membrane_potentials;
    float* membrane_potential_history;
} VoltageRecorderView;

/* ----------------------------------------------------------------------------------------
 * Drop-in launchers.  Same work and order as the reference launchers; they are asynchronous
 * on `stream` (the reference ends each one in cudaDeviceSynchronize).
 * ------------------------------------------------------------------------------------- */

/* Builds the inverse-CDF table of Poisson(mean) used by the background noise: entry k is
 * floor(2^32 * P(X <= k)); a 32-bit uniform u maps to min{k : u < table[k]} (or len).
 * *len = 0 when mean <= 0 (no draw, as the reference skips curand_poisson).
 * Supported means: 0 <= mean <= 64. */
int bn_poisson_table(float mean, uint32_t* table, uint32_t capacity, uint32_t* len);

/* Replaces kernel_neuron_group_evolve_launch (neuron_groups/kernel_neuron_group_evolve.h:10):
 * background shot-noise conductances then the EIF step, fused in one kernel. */
int bn_neuron_group_evolve(NeuronGroupView view, bn_stream_t stream);

/* Replaces kernel_connection_propagate_launch (connections/kernel_connection_propagate.h:10):
 * per-pre r/d traces, x = d - r, then g_post += W^T x with a hand-written GEMV
 * (FP64 accumulation of exact FP32 products: per column, rows are summed sequentially inside
 * chunks of gemv_row_block rows, then the chunk sums are added to g in chunk order). */
int bn_connection_propagate(ConnectionView view, bn_stream_t stream);

/* No counterpart in the reference: bn_connection_propagate for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  view.gemv_partials is ignored: the FP64 partial sums go to a scratch array that the
 * library holds per GPU (8 MiB, 1,048,576 doubles, until the process ends).  Same kernels, same
 * results as bn_connection_propagate with the same gemv_row_block.  Rejected with
 * BN_ERR_INVALID_ARGUMENT when ceil(pre / gemv_row_block) * row stride exceeds 1,048,576.  Calls
 * no allocator.  Calls that overlap on one GPU (other streams or threads) race on the scratch. */
int bn_connection_propagate_no_alloc(ConnectionView view, bn_stream_t stream);

/* Replaces kernel_stimulus_apply_launch (stimuli/kernel_stimulus_apply.h:8-10). */
int bn_stimulus_apply(StimulusView view, uint32_t next_event_idx, uint32_t num_events_to_apply,
                      bn_stream_t stream);
/* Same with the cursor on the device: applies the events of clock->clock_step and advances
 * clock->next_event_idx (graph-capturable). */
int bn_stimulus_apply_device_cursor(StimulusView view, bn_device_clock* clock, bn_stream_t stream);

/* Replace kernel_{triplet,istdp,stdp}_plasticity_update*_launch
 * (plasticities/kernel_triplet_plasticity_update.h:9-10, kernel_istdp_plasticity_update.h:9,
 *  kernel_stdp_plasticity_update.h:9). */
int bn_triplet_plasticity_update(PlasticityView view, PlasticityParams params, bn_stream_t stream);
int bn_triplet_plasticity_update_normalize(PlasticityView view, PlasticityParams params,
                                           bn_stream_t stream);
int bn_istdp_plasticity_update(PlasticityView view, PlasticityParams params, bn_stream_t stream);
int bn_stdp_plasticity_update(PlasticityView view, PlasticityParams params, bn_stream_t stream);

/* Replace kernel_historicize_{spikes,membrane_potentials}_launch
 * (monitors/kernel_historicize_spikes.h:7-12, kernel_historicize_membrane_potentials.h:7-12).
 * Layout as the reference: history[neuron * recording_interval_steps + step % interval]. */
int bn_historicize_spikes(const uint32_t* d_spikes, uint32_t* spikes_history, uint32_t total_neurons,
                          uint32_t recording_interval_steps, uint32_t clock_step, bn_stream_t stream);
int bn_historicize_membrane_potentials(const float* d_membrane_potentials,
                                       float* membrane_potential_history, uint32_t total_neurons,
                                       uint32_t recording_interval_steps, uint32_t clock_step,
                                       bn_stream_t stream);

/* Replaces kernel_init_neuron_rng_states_launch (gpu_core/kernel_init_neuron_rng_states.h:7-11):
 * there is no per-neuron RNG state to initialise; this sets clock_step = 0, next_event_idx = 0
 * and the Philox key to the two words of rng_seed, asynchronously on `stream`. */
int bn_clock_init(bn_device_clock* clock, uint64_t rng_seed, bn_stream_t stream);
/* Advances clock->clock_step by one (the last node of a captured step). */
int bn_clock_advance(bn_device_clock* clock, bn_stream_t stream);

/* ----------------------------------------------------------------------------------------
 * Whole-simulation API (what the reference's main_buryn.cpp + System::run provide).
 * ------------------------------------------------------------------------------------- */
typedef struct bn_sim bn_sim;

enum {
    BN_SIM_WRITE_FILES = 1u << 0,       /* write locations/weights/schedules/recordings like the
                                           reference (output_directory is cleaned first) */
    BN_SIM_NO_MONITORS = 1u << 1,       /* no spike/voltage/weight recorders */
    BN_SIM_LAZY_NORMALIZE = 1u << 2,    /* triplet normalisation kept as per-row scale factors
                                           (W = s_i * M) instead of rewriting W_EE every step */
    BN_SIM_COMPONENT_PATH = 1u << 3,    /* step through the component launchers (drop-in path) */
    BN_SIM_NO_GRAPH = 1u << 4,          /* launch the fused step eagerly instead of a hipGraph */
    BN_SIM_QUIET = 1u << 5,             /* no progress prints */
    BN_SIM_SEQUENTIAL_GRAPH = 1u << 6,  /* graphs run each step's launches one after the other
                                         * instead of overlapping the plasticity of step t-1
                                         * with the evolve of step t (same results) */
    BN_SIM_REFERENCE_LAYOUT = 1u << 7,  /* weight rows unpadded (row stride = post), as the
                                         * reference allocates them (same results, slower GEMV
                                         * where post % 4 != 0) */
    BN_SIM_PERSISTENT = 1u << 8,        /* the persistent step: one cooperative launch per run of
                                         * steps; plasticity applied inside the weight pass, the
                                         * evolve run by the block that completes a column tile.
                                         * Needs BN_SIM_LAZY_NORMALIZE when E->E normalises.  Its
                                         * summation orders (bn_sim_info.gemv_part_chunks = 8,
                                         * lazy row sums by 256-column tile) differ from the
                                         * graph step's, so results differ in rounding only. */
    BN_SIM_GEMV_PART_SUMS = 1u << 9,    /* the fused step's propagate reduces each part of a
                                         * connection's row chunks (the 8 parts the evolve
                                         * combines) inside a block, so the evolve reads one FP64
                                         * row per part instead of one per chunk.  Same results;
                                         * not for bn_multi networks. */
    BN_SIM_RESIDENT = 1u << 10          /* the resident step: one cooperative launch per run of
                                         * steps with every weight in the registers of its blocks
                                         * (each block owns 16 post columns and evolves those
                                         * neurons); only spike words, lazy row scales and
                                         * row-sum parts cross blocks.  Needs
                                         * BN_SIM_LAZY_NORMALIZE when E->E normalises, padded rows,
                                         * and a network whose blocks fit the device (about 4,000
                                         * neurons).  Its summation orders (bn_sim_info.resident_*)
                                         * differ from the graph step's: results differ in
                                         * rounding only. */
};

typedef struct bn_sim_info {
    uint32_t num_excitatory_neurons;
    uint32_t num_inhibitory_neurons;
    uint32_t total_neurons;
    uint32_t clock_step;             /* steps executed so far */
    uint32_t simulation_steps;
    uint32_t num_plasticities;
    uint32_t weight_matrix_ld[4];    /* EE, EI, IE, II row strides (floats) */
    uint32_t gemv_row_block[4];      /* rows per GEMV partial chunk (FP64 summation order) */
    uint32_t weight_pre[4];
    uint32_t weight_post[4];
    uint32_t flags;
    uint32_t kernels_per_step;       /* launches in one fused step */
    uint64_t device_bytes;           /* device memory held by the simulation */
    uint32_t steps_per_graph;        /* steps per captured graph of the fused step (0: no graph) */
    uint32_t graph_launches;         /* captured graphs launched so far */
    uint32_t persistent_blocks;      /* blocks of the persistent step's launch (0: not persistent) */
    uint32_t persistent_items;       /* its (connection, column tile, row segment) work items */
    uint32_t gemv_part_chunks;       /* chunk sums per combination part (0: ceil(chunks / 8));
                                        the persistent step's gemv_row_block is its chunk height */
    uint32_t persistent_resident_rows; /* persistent step with the weights held in registers for
                                        the whole launch: rows per chunk of its kernel (0: the
                                        weights are streamed from memory every step) */
    uint32_t gemv_part_sums;         /* 1: the fused propagate writes one FP64 row per part of
                                        each connection's row chunks (BN_SIM_GEMV_PART_SUMS) */
    uint32_t resident_blocks;        /* blocks of the resident step's launch (0: not resident) */
    uint32_t resident_columns;       /* its post columns per block */
    uint32_t resident_e_waves;       /* its waves (of 8) holding excitatory pre rows */
    uint32_t resident_rows;          /* its pre rows per thread */
} bn_sim_info;

/* Host pointers to fill; any may be NULL.  Arrays are indexed by global neuron id (length
 * total_neurons).  g_* are the conductances after the last step's propagate. */
typedef struct bn_neuron_state_host {
    float* membrane_potentials;
    float* g_excitatory;
    float* g_inhibitory;
    float* i_stimulus;
    uint32_t* refractory_counter;
    uint32_t* spikes;
    float* background_g_excitatory_synaptic_conductances;
    float* background_g_inhibitory_synaptic_conductances;
    float* background_excitatory_postsynaptic_r_traces;
    float* background_excitatory_postsynaptic_d_traces;
    float* background_inhibitory_postsynaptic_r_traces;
    float* background_inhibitory_postsynaptic_d_traces;
} bn_neuron_state_host;

int bn_sim_create(const char* config_path, uint32_t flags, bn_sim** out);
int bn_sim_create_from_json_text(const char* json_text, uint32_t flags, bn_sim** out);
void bn_sim_destroy(bn_sim* sim);
/* Runs `num_steps` steps with the reference's System::run semantics (record() at
 * step % recording_interval_steps == 0 when monitors are on). Asynchronous unless monitors
 * record. */
int bn_sim_run(bn_sim* sim, uint32_t num_steps);
/* Same, through System::step_the_clock and the component launchers. */
int bn_sim_step_components(bn_sim* sim, uint32_t num_steps);
int bn_sim_synchronize(bn_sim* sim);
int bn_sim_get_info(bn_sim* sim, bn_sim_info* info);
int bn_sim_get_neuron_state(bn_sim* sim, const bn_neuron_state_host* dst);
/* True weights of matrix idx (0 EE, 1 EI, 2 IE, 3 II) as dense row-major [pre][post]. */
int bn_sim_get_weights(bn_sim* sim, uint32_t weight_matrix_idx, float* dst);
/* Per-pre connection traces of connection idx: r, d, x (length pre); any may be NULL. */
int bn_sim_get_connection_traces(bn_sim* sim, uint32_t idx, float* r, float* d, float* x);
/* Plasticity traces (length total_neurons each, global ids); any may be NULL. */
int bn_sim_get_plasticity_traces(bn_sim* sim, uint32_t plasticity_offset, float* pre, float* post,
                                 float* triplet_pre, float* triplet_post);
/* Spike bits of steps [first_step, first_step + num_steps) from the device ring the fused step
 * writes every step (the monitors' spike history; needs monitors on): dst [num_steps][words],
 * words = ceil(total_neurons / 32), bit (g & 31) of word g >> 5 = global neuron g spiked.  The
 * steps must have run, lie within the last recording interval, and follow the last record step
 * (whose record flushes the ring). */
int bn_sim_get_spike_raster(bn_sim* sim, uint32_t first_step, uint32_t num_steps, uint32_t* dst);
/* Times `num_steps` steps between two HIP events on the simulation stream; *ms = elapsed. */
int bn_sim_run_timed(bn_sim* sim, uint32_t num_steps, float* ms);
/* Steps per captured graph of the fused step (no counterpart in the reference, whose host loop
 * launches every kernel): default 100; an even count runs the overlapped form (the plasticity of
 * step t-1 in the launch of the evolve of step t), an odd one the sequential form.  Re-captures
 * the graph only; the state is untouched.  Runs shorter than a graph are stepped eagerly. */
int bn_sim_set_steps_per_graph(bn_sim* sim, uint32_t steps_per_graph);
/* Kernel timing on the simulation stream.  *step_ms = wall time of `num_steps` whole steps
 * launched one by one (no graph; the simulation advances by num_steps).  *kernel_ms = wall time
 * of `num_steps` back-to-back launches of the propagate (GEMV) kernel alone over the current
 * weights; the pending propagate partials are saved before and restored after, so the state of
 * the simulation is unchanged by this part. */
int bn_sim_time_propagate_kernel(bn_sim* sim, uint32_t num_steps, float* kernel_ms, float* step_ms);
/* Rows per wave chunk of the persistent step for each connection (0 = chosen from the shape and
 * the device); re-plans the work only.  Changes the GEMV summation order (results move by
 * rounding).  Persistent simulations only. */
int bn_sim_set_persistent_rows(bn_sim* sim, const uint32_t rows[4]);
/* Polls a block of the persistent or resident step makes while it waits for another block's
 * words before it gives up: the launch then ends early on every block, and the next
 * synchronisation reports BN_ERR_RUNTIME (the state is no longer valid).  Default 262,144 (about
 * a quarter second; a step's waits last microseconds). */
int bn_sim_set_persistent_wait_limit(bn_sim* sim, uint32_t polls);
/* Allow (1, the default) or forbid (0) resident weights in the persistent step: when every work
 * item fits a block of its own with chunks of at most 32 rows, each thread keeps its weight
 * entries in registers for the whole launch (the plasticity updates them there) and stores the
 * rows that changed when the launch ends.  Same results either way; re-plans the work.
 * Persistent simulations only. */
int bn_sim_set_persistent_resident(bn_sim* sim, int allow);
/* How the resident step (BN_SIM_RESIDENT) launches its run (no counterpart in the reference):
 * cooperatively (1, the default: the runtime admits the launch only with every block resident at
 * once, and orders it against other cooperative launches) or as an ordinary launch (0: about 14 us
 * less per launch on MI355X; the blocks are co-resident only if no other kernel holds the CUs,
 * and when one does, the bounded waits end the launch with BN_ERR_RUNTIME instead of a hang).
 * The kernel and its results are the same.  Resident simulations only. */
int bn_sim_set_resident_cooperative(bn_sim* sim, int cooperative);
/* Recordings of the fused step (no counterpart in the reference, which writes them on the stepping
 * thread, spike_recorder.cpp:40-60, voltage_recorder.cpp:54-77): written by a writer thread while
 * the next interval's steps run (1, the default; a record step queues the rings' copies to pinned
 * host buffers on the simulation stream, and bn_sim_run returns once the files are written) or on
 * the calling thread at each record step (0).  The files are the same either way. */
int bn_sim_set_async_recording(bn_sim* sim, int async);
/* Host time spent on recordings since creation: seconds on the calling thread at record steps
 * (waits for the device included), seconds the writer thread spent formatting and writing, and
 * the number of record steps. */
int bn_sim_record_times(bn_sim* sim, double* call_seconds, double* writer_seconds, uint32_t* record_steps);
/* State snapshot / resume (no counterpart in the reference): every device buffer of the
 * simulation (neuron and background state, weights, connection and plasticity traces, the
 * pending propagate partials, lazy row sums and scales, monitor rings, the device clock and
 * stimulus cursor) to a file, and back into a simulation created from the same config and
 * flags (checked: a different network or flag set is refused).  A restored simulation continues
 * exactly as the saved one would have. */
int bn_sim_save_state(bn_sim* sim, const char* path);
int bn_sim_load_state(bn_sim* sim, const char* path);
/* Measurement aid (no counterpart in the reference): the read bandwidth of a plain streaming
 * kernel over a device buffer of `bytes` bytes read `repeats` times back to back after one
 * warm-up pass (a working set under the 256 MB MALL stays cache-resident between passes).  GB/s
 * from HIP events; allocates and frees its buffer. */
int bn_measure_read_bandwidth(uint64_t bytes, uint32_t repeats, float* gb_per_s);
/* The same read rate with all `passes` passes inside one launch (4,096 blocks of 256 threads, each
 * pass starting from a rotated offset so that no XCD re-reads its own L2 lines): the steady-state
 * streaming rate without launch ramp and drain.  The passes are capped so that the launch reads
 * at most 4 GiB (well under a second; a multi-second dispatch reads as a hung queue), and the
 * rate uses the passes run.  Measurement aid, no counterpart in the reference. */
int bn_measure_read_bandwidth_one_launch(uint64_t bytes, uint32_t passes, float* gb_per_s);
/* Writes the JSON of the parsed config (all fields, after defaults/overrides/prefixing). */
int bn_config_dump_json(const char* config_path, char* buf, size_t buf_size, size_t* needed);

/* Host-side inputs (reference cpp_src/generation/generate_weights.cpp): dense row-major [pre][post] matrix of
 * matrix idx for the given config, without writing files. */
int bn_generate_weight_matrix(const char* config_path, uint32_t weight_matrix_idx, float* dst,
                              uint32_t* pre, uint32_t* post);
/* Binary connectivity cache (no counterpart in the reference, which draws every matrix again on
 * every run, generate_weights.cpp:39-100): with a directory set, each generated weight matrix is
 * stored there once (keyed by its locations, generation parameters, seed and index, checked by
 * a checksum) and read back by later simulations and bn_generate_weight_matrix calls.  NULL or
 * "" turns it off.  Process-wide; the default is the BN_CONNECTIVITY_CACHE environment variable. */
int bn_set_connectivity_cache(const char* directory);

/* -------------------------------------------------------------------------------------
 * Device state object: replaces GpuSimulationState (gpu_core/gpu_simulation_state.h:19-103,
 * gpu_simulation_state.cu:35-104) for a reference-side caller that keeps its own components and
 * calls the drop-in launchers above.  Every device buffer of one network is allocated once by
 * bn_state_create and owned by the state; the caller builds its *View structs from
 * bn_state_buffers as the reference's make_gpu_view functions do from the d_*() getters.  The
 * Poisson tables, the GEMV scratch (gemv_partials) and the device clock (in place of the
 * per-neuron cuRAND states) are parts of the state.
 * ------------------------------------------------------------------------------------- */
typedef struct bn_state bn_state;

enum {
    BN_STATE_REFERENCE_LAYOUT = 1u << 0, /* weight rows unpadded: row stride = post (the reference's) */
    BN_STATE_HISTORIES = 1u << 1         /* [neuron][slot] spike and voltage histories of
                                            recording_interval_steps slots (the monitors') */
};

typedef struct bn_state_buffers {
    uint32_t total_neurons;
    uint32_t num_excitatory_neurons;
    uint32_t num_inhibitory_neurons;
    uint32_t recording_interval_steps;
    /* neuron state, global ids (gpu_simulation_state.h:38-50) */
    float* membrane_potentials;
    float* g_excitatory;
    float* g_inhibitory;
    float* i_stimulus;
    uint32_t* refractory_counter;
    uint32_t* spikes;
    float* background_g_excitatory_synaptic_conductances;
    float* background_g_inhibitory_synaptic_conductances;
    float* background_excitatory_postsynaptic_r_traces;
    float* background_excitatory_postsynaptic_d_traces;
    float* background_inhibitory_postsynaptic_r_traces;
    float* background_inhibitory_postsynaptic_d_traces;
    bn_device_clock* clock;             /* in place of d_neuron_rng_states */
    /* per weight matrix 0 EE, 1 EI, 2 IE, 3 II (gpu_simulation_state.h:52-67) */
    float* weight_matrix[4];            /* [pre][post], row stride weight_matrix_ld */
    uint32_t weight_matrix_ld[4];
    uint32_t weight_pre[4];
    uint32_t weight_post[4];
    float* postsynaptic_r_traces[4];
    float* postsynaptic_d_traces[4];
    float* summed_postsynaptic_conductances[4];
    double* gemv_partials[4];           /* ConnectionView::gemv_partials */
    uint32_t gemv_row_block[4];
    /* stimulus (gpu_simulation_state.h:69-72) */
    uint32_t* timestep_event_count;
    uint32_t* stimulus_event_neuron_ids;
    uint32_t* stimulus_event_post_group_offsets;
    float* stimulus_event_amplitude_changes;
    /* plasticity, [num_plasticities][total_neurons] each (gpu_simulation_state.h:74-78) */
    float* pre_traces;
    float* post_traces;
    float* triplet_pre_traces;
    float* triplet_post_traces;
    float* pre_weight_sums;
    /* monitors (gpu_simulation_state.h:80-81); NULL without BN_STATE_HISTORIES */
    uint32_t* spike_history;
    float* membrane_potential_history;
} bn_state_buffers;

int bn_state_create(uint32_t total_simulation_length, uint32_t num_excitatory_neurons,
                    uint32_t num_inhibitory_neurons, uint32_t num_stimulus_events, uint32_t num_plasticities,
                    uint64_t rng_seed, uint32_t flags, bn_state** out);
void bn_state_destroy(bn_state* state);
int bn_state_get_buffers(bn_state* state, bn_state_buffers* out);
/* The stream every upload/download of the state runs on (launchers may use it or any other). */
int bn_state_get_stream(bn_state* state, bn_stream_t* out);
int bn_state_total_allocated_bytes(bn_state* state, uint64_t* bytes);
/* Poisson table of group (0 E, 1 I) and kind (0 excitatory, 1 inhibitory background) for
 * `mean`; *table and *len receive the device table for the NeuronGroupView (len 0: no draws). */
int bn_state_poisson_table(bn_state* state, uint32_t group, uint32_t kind, float mean, const uint32_t** table,
                           uint32_t* len);
/* upload_*_view equivalents (gpu_simulation_state.h:84-88): host arrays into the view's slices
 * of the state, with the reference's bounds and size checks.  NULL arrays upload zeros. */
int bn_state_upload_neuron_group_view(bn_state* state, const NeuronGroupView* view, const bn_neuron_state_host* host);
/* weights: dense [pre][post] host array (non-negative, gpu_simulation_state.cu:221-229). */
int bn_state_upload_connection_view(bn_state* state, const ConnectionView* view, const float* weights,
                                    const float* r_traces, const float* d_traces, const float* x);
int bn_state_upload_stimulus_view(bn_state* state, const StimulusView* view, const uint32_t* timestep_event_count,
                                  uint32_t num_timesteps, const uint32_t* neuron_ids, const uint32_t* post_group_offsets,
                                  const float* amplitude_changes, uint32_t num_events);
/* pre_traces, triplet_pre_traces, pre_weight_sums: pre_group_num_neurons each; the post traces post_group_num_neurons. */
int bn_state_upload_plasticity_view(bn_state* state, const PlasticityView* view, const float* pre_traces,
                                    const float* post_traces, const float* triplet_pre_traces,
                                    const float* triplet_post_traces, const float* pre_weight_sums);
/* download_* equivalents (gpu_simulation_state.h:90-93): [num_neurons][num_timesteps] slices of
 * the histories (zeroed after the copy) and the dense [pre][post] weights. */
int bn_state_download_and_flush_spike_history(bn_state* state, uint32_t offset, uint32_t num_neurons,
                                              uint32_t num_timesteps, uint32_t* dst);
int bn_state_download_and_flush_membrane_potential_history(bn_state* state, uint32_t offset, uint32_t num_neurons,
                                                           uint32_t num_timesteps, float* dst);
int bn_state_download_weight_matrix(bn_state* state, uint32_t weight_matrix_idx, float* dst);
/* Synchronous copy of `bytes` from a device address of the state to host memory. */
int bn_state_download(bn_state* state, const void* device_src, void* host_dst, size_t bytes);

/* -------------------------------------------------------------------------------------
 * Batched shared-weight form: B instances of one network whose weights are frozen
 * (plasticity off), each with its own background-noise key and stimulus pulse, advanced
 * together.  The propagate of all instances is one GEMM per connection on the FP32 matrix
 * cores (FP32 accumulation), so each instance matches a plain run of the same config with
 * background_noise_seed = its noise seed and the same pulse, to FP32 summation-order
 * rounding.  No counterpart in the reference (it runs one network per process).
 * ------------------------------------------------------------------------------------- */
typedef struct bn_batch bn_batch;

typedef struct bn_batch_instance {
    uint64_t noise_seed;             /* Philox key of this instance's background noise */
    float stimulus_amplitude;        /* square pulse: +amplitude in the stimulus phase of step */
    uint32_t stimulus_start;         /* stimulus_start, -amplitude in that of stimulus_end + 1 */
    uint32_t stimulus_end;           /* (events at or after simulation_steps are dropped, as the
                                        plain schedule does) */
    uint32_t num_stimulus_ids;       /* excitatory-local ids receiving the pulse (distinct) */
    const uint32_t* stimulus_ids;
} bn_batch_instance;

typedef struct bn_batch_info {
    uint32_t num_instances;
    uint32_t padded_instances;       /* instances padded to a multiple of 64 (one wave) */
    uint32_t num_excitatory_neurons;
    uint32_t num_inhibitory_neurons;
    uint32_t total_neurons;
    uint32_t clock_step;
    uint32_t gemm_rows_per_wave;     /* pre rows per MFMA wave chunk (the tuned tile height) */
    uint32_t gemm_blocks;
    uint32_t gemm_groups[4];         /* FP32 partial sums per post neuron, EE, EI, IE, II */
    uint32_t weight_pre[4];
    uint32_t weight_post[4];
    uint32_t kernels_per_step;
} bn_batch_info;

/* config_path: the plain network's JSON config (plasticity must be off; rng_seed seeds the
 * shared connectivity). */
int bn_batch_create(const char* config_path, uint32_t num_instances, const bn_batch_instance* instances,
                    bn_batch** out);
void bn_batch_destroy(bn_batch* batch);
int bn_batch_run(bn_batch* batch, uint32_t num_steps);
int bn_batch_synchronize(bn_batch* batch);
int bn_batch_run_timed(bn_batch* batch, uint32_t num_steps, float* ms);
/* Wall time of `num_launches` back-to-back GEMM launches; the batch's state is unchanged. */
int bn_batch_time_gemm(bn_batch* batch, uint32_t num_launches, float* ms);
/* Rows of pre per MFMA wave chunk (multiple of 8; 0 = heuristic); before the first step only. */
int bn_batch_set_gemm_rows(bn_batch* batch, uint32_t rows_per_wave);
int bn_batch_get_info(bn_batch* batch, bn_batch_info* info);
/* State of one instance, indexed by global neuron id (g_* after the last propagate). */
int bn_batch_get_state(bn_batch* batch, uint32_t instance, const bn_neuron_state_host* dst);
/* The recorded per-instance stimulus draw of the shared-weight workload: mt19937 seeded with
 * `instance`, num_ids distinct excitatory ids (partial Fisher-Yates, returned sorted) and
 * amplitude = max_amplitude * U(0,1). */
int bn_batch_workload_instance(uint32_t instance, uint32_t num_excitatory, uint32_t num_ids, float max_amplitude,
                               uint32_t* ids, float* amplitude);

/* -------------------------------------------------------------------------------------
 * Several independent plastic networks (any sizes, own weights, seeds and plasticity) advanced
 * together: one launch per phase (evolve, propagate, plasticity) covers every network, ten
 * steps per hipGraph.  Each network keeps its own bn_sim (state, weights, plans) and computes
 * bit for bit what it computes alone.  No counterpart in the reference (one network per
 * process there).
 * ------------------------------------------------------------------------------------- */
typedef struct bn_multi bn_multi;

typedef struct bn_multi_info {
    uint32_t num_networks;
    uint32_t clock_step;
    uint32_t evolve_blocks;          /* blocks of one launch of each phase, all networks */
    uint32_t propagate_blocks;
    uint32_t plasticity_blocks;
    uint32_t nontemporal;            /* weight loads past the caches (total weights > 192 MB) */
    uint32_t kernels_per_step;       /* launches per step of every network (3 per group) */
    uint32_t num_groups;             /* schedule: network groups stepped one after the other */
    uint32_t steps_per_graph;
    uint64_t group_weight_bytes;     /* weight bytes per group (0 = one group) */
    uint32_t num_streams;            /* concurrent network sets (bn_multi_set_streams) */
    uint32_t overlapped;             /* graphs run the plasticity of step t-1 beside the evolve of t */
} bn_multi_info;

/* flags: BN_SIM_* of every network (BN_SIM_LAZY_NORMALIZE is required when a network
 * normalises; monitors, files and the component path are not available here). */
int bn_multi_create(const char* const* config_paths, uint32_t num_networks, uint32_t flags, bn_multi** out);
void bn_multi_destroy(bn_multi* multi);
int bn_multi_run(bn_multi* multi, uint32_t num_steps);
int bn_multi_run_timed(bn_multi* multi, uint32_t num_steps, float* ms);
int bn_multi_synchronize(bn_multi* multi);
/* Schedule: networks are grouped in order so that a group's weights stay under
 * group_weight_bytes (0 = one group); each group advances steps_per_graph steps per graph
 * launch, the groups one after the other (networks are independent, so results do not depend
 * on the schedule).  Sized to the 256 MB MALL, a group's weights stay cache-resident across
 * its steps. */
int bn_multi_set_schedule(bn_multi* multi, uint64_t group_weight_bytes, uint32_t steps_per_graph);
/* Concurrent schedule: the networks split into `streams` (1-4) sets of about equal weight bytes,
 * each set stepped by its own graphs on its own stream, so that one set's latency-bound evolve
 * and plasticity launches run beside another set's propagate.  overlap = 1 (even
 * steps_per_graph): each graph runs the plasticity of step t-1 in the launch of the evolve of
 * step t, as a single network's graphs do.  Results do not change. */
int bn_multi_set_streams(bn_multi* multi, uint32_t streams, uint32_t steps_per_graph, uint32_t overlap);
/* Wall time of `num_launches` back-to-back propagate launches over every network; the state of
 * every network is unchanged (pending partial sums saved and restored). */
int bn_multi_time_propagate(bn_multi* multi, uint32_t num_launches, float* ms);
/* The bn_sim of network `index` (owned by the multi object; synchronises first) for state,
 * weight and trace downloads. */
int bn_multi_network(bn_multi* multi, uint32_t index, bn_sim** out);
int bn_multi_get_info(bn_multi* multi, bn_multi_info* info);

/* -------------------------------------------------------------------------------------
 * Differentiable window: `steps` plasticity-free, stimulus-free steps of a network from the
 * state of a bn_sim, taken when the window is created; the loss
 *   L = (1/T) sum_t [ sum_{p=E,I} (rho_p,t - rho*_p,t)^2 + (1/n_vrec) sum_k ((V_k,t - V*_k,t) / 1 mV)^2 ],
 *   rho_p,t = a rho_p,t-1 + (1 - a) (spikes of p at t / N_p) / dt,  a = exp(-dt / rate_tau),
 * and its gradient with respect to the four weight matrices, g_leak and tau_decay_e by the
 * adjoint (reverse) pass: the forward stores the state every checkpoint_interval steps, the
 * backward recomputes each segment with per-step records and walks it backwards.  The spike
 * derivative is the fast sigmoid 1 / (beta (1 + |U - v_threshold| / beta)^2), beta =
 * surrogate_beta (0: spikes held fixed — the exact derivative wherever no spike flips); the
 * reset passes gradient with the spike held fixed, refractory steps pass none.  No
 * counterpart in the reference (no gradient code there); the step it differentiates is
 * kernel_neuron_group_evolve.cu:11-130 with kernel_connection_propagate.cu:8-94 and the
 * trace decay exp(-dt / tau_decay_e) of connection.cpp:46-47.
 * ------------------------------------------------------------------------------------- */
typedef struct bn_window bn_window;

typedef struct bn_window_spec {
    uint32_t steps;                  /* T */
    uint32_t checkpoint_interval;    /* steps per recompute segment (100) */
    uint32_t n_vrec;                 /* recorded excitatory neurons entering the loss */
    const uint32_t* vrec_ids;        /* [n_vrec] global ids < N_E */
    float rate_tau;                  /* s (5e-3) */
    float surrogate_beta;            /* V (1e-3 = slope 1/(1 mV)); 0 = spikes held fixed */
} bn_window_spec;

typedef struct bn_window_info {
    uint32_t steps;
    uint32_t checkpoint_interval;
    uint32_t start_step;             /* noise counter of the first step */
    uint32_t total_neurons;
    uint32_t num_excitatory_neurons;
    uint32_t num_inhibitory_neurons;
    uint32_t n_vrec;
    float g_leak;                    /* current parameters */
    float tau_decay_e;
    float forward_ms;                /* device time of the last forward's steps (HIP events) */
    float backward_ms;               /* device time of the last backward's recompute + reverse */
} bn_window_info;

/* The network must have both plasticities off.  Weights, g_leak and tau_decay_e start at the
 * simulation's values. */
int bn_window_create(bn_sim* sim, const bn_window_spec* spec, bn_window** out);
void bn_window_destroy(bn_window* window);
/* weights: NULL or 4 pointers (EE, EI, IE, II; any may be NULL = keep), dense [pre][post]
 * FP32, host or device memory. */
int bn_window_set_params(bn_window* window, const float* const* weights, float g_leak, float tau_decay_e);
int bn_window_get_weights(bn_window* window, uint32_t weight_matrix_idx, float* dst);
/* target_rates [steps][2] Hz (E, I), target_v [steps][n_vrec] V; NULL = zero. */
int bn_window_set_targets(bn_window* window, const double* target_rates, const double* target_v);
/* Runs the window from the start state and evaluates L.  Outputs may be NULL: rates
 * [steps][2] (FP64), v [steps][n_vrec], raster [steps][N] (0/1). */
int bn_window_forward(bn_window* window, double* loss, double* rates, float* v, uint8_t* raster);
/* Gradient of the last forward's L.  dW: NULL or 4 pointers (NULL = skip), dense [pre][post]
 * FP32, host or device memory. */
int bn_window_backward(bn_window* window, float* const* dW, double* dg_leak, double* dtau_decay_e);
/* dL/dW of weight matrix idx from the last backward into host memory [pre][post] (through a
 * pinned staging buffer: the fast path to the host). */
int bn_window_get_gradient(bn_window* window, uint32_t weight_matrix_idx, float* dst);
int bn_window_get_info(bn_window* window, bn_window_info* info);

#ifdef __cplusplus
}
#endif

#endif /* BURYN_AMD_H */
