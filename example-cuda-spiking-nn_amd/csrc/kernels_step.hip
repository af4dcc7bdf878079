// Hand-written HIP kernels of the per-step hot path for CDNA4 (gfx950), and the extern "C"
// launcher layer of include/buryn_amd.h.
//
// Reference semantics followed (paths relative to the reference repository):
//   background + EIF      cpp_src/neuron_groups/kernel_neuron_group_evolve.cu:11-130
//   connection traces     cpp_src/connections/kernel_connection_propagate.cu:8-46
//   GEMV (cublasSgemv)    cpp_src/connections/kernel_connection_propagate.cu:79-94
//   stimulus              cpp_src/stimuli/kernel_stimulus_apply.cu:7-21, stimulus.cpp:46-60
//   triplet STDP          cpp_src/plasticities/kernel_triplet_plasticity_update.cu:16-216
//   iSTDP                 cpp_src/plasticities/kernel_istdp_plasticity_update.cu:16-135
//   pair STDP             cpp_src/plasticities/kernel_stdp_plasticity_update.cu:16-137
//   monitors              cpp_src/monitors/kernel_historicize_{spikes,membrane_potentials}.cu
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "bn_device.h"
#include "bn_errors.h"
#include "bn_step.h"
#include "bn_neuron.h"

namespace bn {

// =======================================================================================
// Per-neuron device functions (one neuron, reference order of operations)
// =======================================================================================

__device__ __forceinline__ void background_store(const NeuronGroupView& v, uint32_t lid, float re, float ri,
                                                 float de, float di, float bg_ge, float bg_gi) {
    v.background_excitatory_postsynaptic_r_traces[lid] = re;
    v.background_inhibitory_postsynaptic_r_traces[lid] = ri;
    v.background_excitatory_postsynaptic_d_traces[lid] = de;
    v.background_inhibitory_postsynaptic_d_traces[lid] = di;
    v.background_g_excitatory_synaptic_conductances[lid] = bg_ge;
    v.background_g_inhibitory_synaptic_conductances[lid] = bg_gi;
}

// The whole background update of one neuron, memory to memory (component path).
__device__ __forceinline__ void background_step(const NeuronGroupView& v, uint32_t lid,
                                                uint32_t step, const bn_device_clock& clk,
                                                float& bg_ge, float& bg_gi) {
    float re = v.background_excitatory_postsynaptic_r_traces[lid];
    float ri = v.background_inhibitory_postsynaptic_r_traces[lid];
    float de = v.background_excitatory_postsynaptic_d_traces[lid];
    float di = v.background_inhibitory_postsynaptic_d_traces[lid];
    background_advance(v, v.offset + lid, step, clk, v.poisson_table_excitatory, v.poisson_table_inhibitory,
                       re, ri, de, di, bg_ge, bg_gi);
    background_store(v, lid, re, ri, de, di, bg_ge, bg_gi);
}

// kernel_connection_propagate.cu:8-46 for one pre neuron.
__device__ __forceinline__ void connection_trace_step(float* r, float* d, float* x, float decay_r,
                                                      float decay_d, float K, uint32_t i,
                                                      uint32_t spike) {
    float dd = flush_trace(d[i] * decay_d);
    float rr = flush_trace(r[i] * decay_r);
    if (spike == 1) {
        rr += K;
        dd += K;
    }
    d[i] = dd;
    r[i] = rr;
    x[i] = dd - rr;
}

// Part sums of one block's 64 neurons: wave w of the block sums part w of every source.
// fin[slot][w][lane], slots 0..kMaxSources-1 for g_e sources, kMaxSources.. for g_i.
typedef double FinParts[2 * kMaxSources][kFinParts][kEvolveNeurons];

__device__ __forceinline__ float combine_parts(float g, uint32_t chunks, const double (*part)[kEvolveNeurons],
                                               uint32_t lane) {
    const uint32_t per = fin_part_size(chunks);
    double acc = (double)g;
#pragma unroll
    for (uint32_t w = 0; w < (uint32_t)kFinParts; ++w)
        if (w * per < chunks) acc += part[w][lane];
    return (float)acc;
}

// Plasticity trace: the previous step's +1 increment (kernel_triplet_plasticity_update.cu:95-115)
// then this step's decay with flush (:16-50), the same two roundings as the reference.
__device__ __forceinline__ float trace_step(float tr, uint32_t spiked_prev, float decay) {
    if (spiked_prev == 1) tr += 1.0f;
    return flush_trace(tr * decay);
}

// =======================================================================================
// k_evolve: every neuron of every group in one launch
// =======================================================================================
// A block owns 64 neurons (lane = neuron) and has kFinParts waves, each with one role whose
// loads are all issued at the start (one memory round trip per wave): wave 0 loads the neuron
// state and the connection traces, wave 1 runs the background update (Philox draws, Poisson
// counts, the four background traces), wave 2 the plasticity trace step (it needs only the
// previous step's spikes), and waves 3.. sum the parts of the propagate partials.  After the
// barrier wave 0 combines the parts and runs the EIF step, the connection traces and the
// monitors.

constexpr uint32_t kPartWave0 = 3;                          // first wave that sums partials
constexpr uint32_t kPartWaves = kFinParts - kPartWave0;     // parts p, p + kPartWaves, ... per wave
// The parts of this wave (pw, pw + kPartWaves, ... < kFinParts) of every source of the group,
// stored in fin[slot][part][lane] (slots 0.. for g_e sources, kMaxSources.. for g_i).
// One excitatory and one inhibitory source, given per lane (a block may straddle two groups):
// parts pw and pw + kPartWaves of both, four ranges with their loads in flight together.
__device__ __forceinline__ void evolve_parts_pair(const double* pe, uint32_t lde, uint32_t ce, const double* pi,
                                                  uint32_t ldi, uint32_t ci, uint32_t pw, uint32_t lane,
                                                  FinParts& fin) {
    const uint32_t se = fin_part_size(ce), si = fin_part_size(ci);
    const uint32_t pb = pw + kPartWaves;
    const bool has_b = pb < (uint32_t)kFinParts;
    const double* const ptr[4] = {pe, pi, pe, pi};
    const uint32_t ld[4] = {lde, ldi, lde, ldi};
    const uint32_t c0[4] = {min(pw * se, ce), min(pw * si, ci), has_b ? min(pb * se, ce) : 0u,
                            has_b ? min(pb * si, ci) : 0u};
    const uint32_t c1[4] = {min(c0[0] + se, ce), min(c0[1] + si, ci), has_b ? min(c0[2] + se, ce) : 0u,
                            has_b ? min(c0[3] + si, ci) : 0u};
    double out[4];
    range_sums<4>(ptr, ld, c0, c1, out);
    if (pw * se < ce) fin[0][pw][lane] = out[0];
    if (pw * si < ci) fin[kMaxSources][pw][lane] = out[1];
    if (has_b && pb * se < ce) fin[0][pb][lane] = out[2];
    if (has_b && pb * si < ci) fin[kMaxSources][pb][lane] = out[3];
}

// Any number of sources, one group.
__device__ __forceinline__ void evolve_parts_source(const FinalizeSource& F, uint32_t slot, uint32_t lid, uint32_t pw,
                                                    uint32_t lane, FinParts& fin) {
    const uint32_t per = fin_part_size(F.chunks);
    for (uint32_t q = pw; q < (uint32_t)kFinParts; q += kPartWaves) {
        const uint32_t c0 = q * per;
        if (c0 < F.chunks) fin[slot][q][lane] = fin_part_sum(F, lid, c0, min(c0 + per, F.chunks));
    }
}
__device__ __forceinline__ void evolve_parts(const EvolveGroup& G, uint32_t lid, uint32_t pw, uint32_t lane,
                                             FinParts& fin) {
    for (uint32_t s = 0; s < G.n_fin_e; ++s) evolve_parts_source(G.fin_e[s], s, lid, pw, lane, fin);
    for (uint32_t s = 0; s < G.n_fin_i; ++s) evolve_parts_source(G.fin_i[s], kMaxSources + s, lid, pw, lane, fin);
}

// The groups of a block's first and last neuron (wave-uniform).  Groups are contiguous global id
// ranges, so a block whose every live lane lies in one of the two straddles at most these two.
__device__ __forceinline__ void block_groups(const EvolvePlan& P, uint32_t blk, uint32_t& g_lo, uint32_t& g_hi) {
    const uint32_t first = blk * kEvolveNeurons;
    const uint32_t last = min(first + (uint32_t)kEvolveNeurons, P.total_neurons) - 1;
    g_lo = g_hi = 0;
#pragma unroll
    for (int g = 0; g < kMaxGroups; ++g) {  // unrolled: the offset and size loads go out together
        const uint32_t o = P.g[g].v.offset, n = P.g[g].v.num_neurons;
        const bool on = (uint32_t)g < P.n_groups;
        if (on && first >= o && first - o < n) g_lo = g;
        if (on && last >= o && last - o < n) g_hi = g;
    }
}

__device__ __forceinline__ void evolve_block(const EvolvePlan& P, uint32_t blk, uint32_t step_offset,
                                             uint32_t stim_prev) {
    __shared__ FinParts fin;
    __shared__ float bgl[2][kEvolveNeurons];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t gid = blk * kEvolveNeurons + lane;
    const bool live = gid < P.total_neurons;

    // Wave 0: per-neuron inputs of the EIF step and the connection traces.
    float V = 0.0f, ge = 0.0f, gi_ = 0.0f, istim = 0.0f;
    // The lane's group fields used after the barrier (wave 0; see below).
    bool two_groups = false, in_group = false;
    uint32_t lid = 0, n_fe = 0, n_fi = 0, chunks_e[kMaxSources] = {}, chunks_i[kMaxSources] = {};
    float e_dt = 0.0f, e_gl = 0.0f, e_c = 0.0f, e_slope = 0.0f, e_vt = 0.0f, e_ee = 0.0f, e_ei = 0.0f,
          e_vthr = 0.0f, e_vreset = 0.0f, e_vrest = 0.0f;
    uint32_t e_refr = 0;
    float *pV = nullptr, *pGe = nullptr, *pGi = nullptr, *pIs = nullptr;
    uint32_t *pR = nullptr, *pS = nullptr;
    uint32_t refr = 0;
    int32_t vslot = -1;
    float tr_r[kMaxTraceSets], tr_d[kMaxTraceSets];
#pragma unroll
    for (int t = 0; t < kMaxTraceSets; ++t) tr_r[t] = tr_d[t] = 0.0f;
    if (wave == 0) {
        if (live) {
#pragma unroll
            for (int g = 0; g < kMaxGroups; ++g) {
                const NeuronGroupView& v = P.g[g].v;
                if ((uint32_t)g < P.n_groups && gid >= v.offset && gid < v.offset + v.num_neurons) {
                    const uint32_t lid = gid - v.offset;
                    V = v.membrane_potentials[lid];
                    refr = v.refractory_counter[lid];
                    ge = v.g_excitatory[lid];
                    gi_ = v.g_inhibitory[lid];
                    istim = v.i_stimulus[lid];
                }
            }
#pragma unroll
            for (int t = 0; t < kMaxTraceSets; ++t) {
                const TraceSet& T = P.ts[t];
                if ((uint32_t)t < P.n_ts && gid >= T.pre_offset && gid < T.pre_offset + T.n_pre) {
                    tr_r[t] = T.r[gid - T.pre_offset];
                    tr_d[t] = T.d[gid - T.pre_offset];
                }
            }
            if (P.v_ring) vslot = P.vrec_index[gid];
        }
        // The lane's EIF parameters, state pointers and source chunk counts from the block's (at
        // most two) groups, loaded before the barrier.
        uint32_t g_lo, g_hi;
        block_groups(P, blk, g_lo, g_hi);
        const EvolveGroup& GA = P.g[g_lo];
        const EvolveGroup& GB = P.g[g_hi];
        const uint32_t oa = uni(GA.v.offset), ob = uni(GB.v.offset);
        const bool inA = live && gid >= oa && gid - oa < uni(GA.v.num_neurons);
        const bool inB = !inA && live && gid >= ob && gid - ob < uni(GB.v.num_neurons);
        two_groups = __ballot(live && !inA && !inB) == 0;
        in_group = inA || inB;
        const bool b = inB;
        lid = b ? gid - ob : gid - oa;
        e_dt = b ? uni(GB.v.timestep) : uni(GA.v.timestep);
        e_gl = b ? uni(GB.v.g_leak) : uni(GA.v.g_leak);
        e_c = b ? uni(GB.v.membrane_capacitance) : uni(GA.v.membrane_capacitance);
        e_slope = b ? uni(GB.v.depolarization_slope_factor) : uni(GA.v.depolarization_slope_factor);
        e_vt = b ? uni(GB.v.v_effective_spike_threshold) : uni(GA.v.v_effective_spike_threshold);
        e_ee = b ? uni(GB.v.v_excitatory_reversal) : uni(GA.v.v_excitatory_reversal);
        e_ei = b ? uni(GB.v.v_inhibitory_reversal) : uni(GA.v.v_inhibitory_reversal);
        e_vthr = b ? uni(GB.v.v_threshold) : uni(GA.v.v_threshold);
        e_vreset = b ? uni(GB.v.v_reset) : uni(GA.v.v_reset);
        e_vrest = b ? uni(GB.v.v_rest) : uni(GA.v.v_rest);
        e_refr = b ? uni(GB.v.refractory_steps) : uni(GA.v.refractory_steps);
        pV = b ? uni(GB.v.membrane_potentials) : uni(GA.v.membrane_potentials);
        pGe = b ? uni(GB.v.g_excitatory) : uni(GA.v.g_excitatory);
        pGi = b ? uni(GB.v.g_inhibitory) : uni(GA.v.g_inhibitory);
        pIs = b ? uni(GB.v.i_stimulus) : uni(GA.v.i_stimulus);
        pR = b ? uni(GB.v.refractory_counter) : uni(GA.v.refractory_counter);
        pS = b ? uni(GB.v.spikes) : uni(GA.v.spikes);
        n_fe = b ? uni(GB.n_fin_e) : uni(GA.n_fin_e);
        n_fi = b ? uni(GB.n_fin_i) : uni(GA.n_fin_i);
#pragma unroll
        for (int q = 0; q < kMaxSources; ++q) {
            chunks_e[q] = b ? uni(GB.fin_e[q].chunks) : uni(GA.fin_e[q].chunks);
            chunks_i[q] = b ? uni(GB.fin_i[q].chunks) : uni(GA.fin_i[q].chunks);
        }
    } else if (wave >= kPartWave0) {
        // Waves 3..: part sums of the propagate partials.  The common layout (one excitatory and
        // one inhibitory source per group) with the lane's sources selected per lane from the
        // block's (at most two) groups.
        uint32_t g_lo, g_hi;
        block_groups(P, blk, g_lo, g_hi);
        const EvolveGroup& GA = P.g[g_lo];
        const EvolveGroup& GB = P.g[g_hi];
        const uint32_t oa = uni(GA.v.offset), ob = uni(GB.v.offset);
        const bool inA = live && gid >= oa && gid - oa < uni(GA.v.num_neurons);
        const bool inB = !inA && live && gid >= ob && gid - ob < uni(GB.v.num_neurons);
        const bool pair = uni(GA.n_fin_e) == 1 && uni(GA.n_fin_i) == 1 && uni(GB.n_fin_e) == 1 &&
                          uni(GB.n_fin_i) == 1 && __ballot(live && !inA && !inB) == 0;
        if (pair) {
            if (inA || inB) {
                const double* pe = inB ? uni(GB.fin_e[0].partials) + (gid - ob) : uni(GA.fin_e[0].partials) + (gid - oa);
                const double* pi = inB ? uni(GB.fin_i[0].partials) + (gid - ob) : uni(GA.fin_i[0].partials) + (gid - oa);
                const uint32_t lde = inB ? uni(GB.fin_e[0].ld) : uni(GA.fin_e[0].ld);
                const uint32_t ldi = inB ? uni(GB.fin_i[0].ld) : uni(GA.fin_i[0].ld);
                const uint32_t ce = inB ? uni(GB.fin_e[0].chunks) : uni(GA.fin_e[0].chunks);
                const uint32_t ci = inB ? uni(GB.fin_i[0].chunks) : uni(GA.fin_i[0].chunks);
                evolve_parts_pair(pe, lde, ce, pi, ldi, ci, wave - kPartWave0, lane, fin);
            }
        } else if (live) {
            for (uint32_t g = 0; g < P.n_groups; ++g) {
                const EvolveGroup& G = P.g[g];
                if (gid < G.v.offset || gid >= G.v.offset + G.v.num_neurons) continue;
                evolve_parts(G, gid - G.v.offset, wave - kPartWave0, lane, fin);
            }
        }
    } else if (wave == 1) {
        // Wave 1: background (kernel_neuron_group_evolve.cu:11-83).
        if (live) {
            for (uint32_t g = 0; g < P.n_groups; ++g) {
                const NeuronGroupView& v = P.g[g].v;
                if (gid < v.offset || gid >= v.offset + v.num_neurons) continue;
                const uint32_t lid = gid - v.offset;
                float re = v.background_excitatory_postsynaptic_r_traces[lid];
                float ri = v.background_inhibitory_postsynaptic_r_traces[lid];
                float de = v.background_excitatory_postsynaptic_d_traces[lid];
                float di = v.background_inhibitory_postsynaptic_d_traces[lid];
                uint32_t he[4], hi[4];
#pragma unroll
                for (uint32_t k = 0; k < 4; ++k) {
                    he[k] = k < v.poisson_table_excitatory_len ? v.poisson_table_excitatory[k] : 0u;
                    hi[k] = k < v.poisson_table_inhibitory_len ? v.poisson_table_inhibitory[k] : 0u;
                }
                const bn_device_clock clk = *P.clock;
                const uint32_t step = clk.clock_step + step_offset;
                re = flush_trace(re * v.background_excitatory_postsynaptic_r_trace_decay);
                ri = flush_trace(ri * v.background_inhibitory_postsynaptic_r_trace_decay);
                de = flush_trace(de * v.background_excitatory_postsynaptic_d_trace_decay);
                di = flush_trace(di * v.background_inhibitory_postsynaptic_d_trace_decay);
                uint32_t ke = 0, ki = 0;
                if (v.poisson_table_excitatory_len > 0 || v.poisson_table_inhibitory_len > 0) {
                    const U32x4 r = background_draws(clk, step, gid, v.rng_instance);
                    if (v.mean_background_excitatory_spikes > 0.0f)
                        ke = poisson_lookup(r.x, v.poisson_table_excitatory, v.poisson_table_excitatory_len, he);
                    if (v.mean_background_inhibitory_spikes > 0.0f)
                        ki = poisson_lookup(r.y, v.poisson_table_inhibitory, v.poisson_table_inhibitory_len, hi);
                }
                if (ke > 0u) {
                    const float inc = (float)ke;
                    re += inc;
                    de += inc;
                }
                const float bg_ge = (de - re) * v.background_excitatory_scaling_factor;
                if (ki > 0u) {
                    const float inc = (float)ki;
                    ri += inc;
                    di += inc;
                }
                const float bg_gi = (di - ri) * v.background_inhibitory_scaling_factor;
                background_store(v, lid, re, ri, de, di, bg_ge, bg_gi);
                bgl[0][lane] = bg_ge;
                bgl[1][lane] = bg_gi;
            }
        }
    } else {
        // Wave 2: plasticity traces, +1 of the previous step then this step's decay.
        if (live) {
            uint32_t spike_prev = 0;
#pragma unroll
            for (int g = 0; g < kMaxGroups; ++g) {
                const NeuronGroupView& v = P.g[g].v;
                if ((uint32_t)g < P.n_groups && gid >= v.offset && gid < v.offset + v.num_neurons)
                    spike_prev = v.spikes[gid - v.offset];
            }
            float a[kMaxPlast][2], b[kMaxPlast][2];
#pragma unroll
            for (int q = 0; q < kMaxPlast; ++q) {
                const PlastTraces& Q = P.pt[q];
                a[q][0] = b[q][0] = a[q][1] = b[q][1] = 0.0f;
                if ((uint32_t)q >= P.n_pt) continue;
                const bool trip = Q.kind == kPlastTriplet;
                if (gid >= Q.pre_offset && gid < Q.pre_offset + Q.n_pre) {
                    a[q][0] = Q.pre[gid - Q.pre_offset];
                    if (trip) b[q][0] = Q.trip_pre[gid - Q.pre_offset];
                }
                if (gid >= Q.post_offset && gid < Q.post_offset + Q.n_post) {
                    a[q][1] = Q.post[gid - Q.post_offset];
                    if (trip) b[q][1] = Q.trip_post[gid - Q.post_offset];
                }
            }
#pragma unroll
            for (int q = 0; q < kMaxPlast; ++q) {
                const PlastTraces& Q = P.pt[q];
                if ((uint32_t)q >= P.n_pt) continue;
                const bool trip = Q.kind == kPlastTriplet;
                if (gid >= Q.pre_offset && gid < Q.pre_offset + Q.n_pre) {
                    const uint32_t i = gid - Q.pre_offset;
                    Q.pre_out[i] = trace_step(a[q][0], spike_prev, Q.pre_decay);
                    if (trip) Q.trip_pre_out[i] = trace_step(b[q][0], spike_prev, Q.trip_pre_decay);
                }
                if (gid >= Q.post_offset && gid < Q.post_offset + Q.n_post) {
                    const uint32_t j = gid - Q.post_offset;
                    Q.post_out[j] = trace_step(a[q][1], spike_prev, Q.post_decay);
                    if (trip) Q.trip_post_out[j] = trace_step(b[q][1], spike_prev, Q.trip_post_decay);
                }
            }
        }
    }
    __syncthreads();
    if (wave != 0) return;

    const bn_device_clock clk = *P.clock;
    const uint32_t step = clk.clock_step + step_offset;
    uint32_t spike = 0;
    if (two_groups) {
        if (in_group) {
            if (stim_prev && P.ev_steps && step - 1 < P.ev_steps) {
                // stimulus.cpp:46-60 for this neuron: the previous step's events in list order.
                const uint32_t e1 = min(P.ev_start[step], P.ev_count);
                bool hit = false;
                for (uint32_t e = P.ev_start[step - 1]; e < e1; ++e)
                    if (P.ev_ids[e] + P.ev_offsets[e] == gid) {
                        istim += P.ev_amps[e];
                        hit = true;
                    }
                if (hit) pIs[lid] = istim;
            }
#pragma unroll
            for (int q = 0; q < kMaxSources; ++q)
                if ((uint32_t)q < n_fe) ge = combine_parts(ge, chunks_e[q], fin[q], lane);
#pragma unroll
            for (int q = 0; q < kMaxSources; ++q)
                if ((uint32_t)q < n_fi) gi_ = combine_parts(gi_, chunks_i[q], fin[kMaxSources + q], lane);
            eif_core(e_dt, e_gl, e_c, e_slope, e_vt, e_ee, e_ei, e_vthr, e_vreset, e_vrest, e_refr, ge, gi_, istim,
                     V, refr, spike);
            pV[lid] = V;
            pR[lid] = refr;
            pS[lid] = spike;
            pGe[lid] = bgl[0][lane];
            pGi[lid] = bgl[1][lane];
        }
    } else if (live) {
        for (uint32_t g = 0; g < P.n_groups; ++g) {
            const EvolveGroup& G = P.g[g];
            const NeuronGroupView& v = G.v;
            if (gid < v.offset || gid >= v.offset + v.num_neurons) continue;
            const uint32_t lid = gid - v.offset;
            if (stim_prev && P.ev_steps && step - 1 < P.ev_steps) {
                // stimulus.cpp:46-60 for this neuron: the previous step's events in list order.
                const uint32_t e1 = min(P.ev_start[step], P.ev_count);
                bool hit = false;
                for (uint32_t e = P.ev_start[step - 1]; e < e1; ++e)
                    if (P.ev_ids[e] + P.ev_offsets[e] == gid) {
                        istim += P.ev_amps[e];
                        hit = true;
                    }
                if (hit) v.i_stimulus[lid] = istim;
            }
            for (uint32_t q = 0; q < G.n_fin_e; ++q) ge = combine_parts(ge, G.fin_e[q].chunks, fin[q], lane);
            for (uint32_t q = 0; q < G.n_fin_i; ++q) gi_ = combine_parts(gi_, G.fin_i[q].chunks, fin[kMaxSources + q], lane);
            eif_step(v, ge, gi_, istim, V, refr, spike);
            v.membrane_potentials[lid] = V;
            v.refractory_counter[lid] = refr;
            v.spikes[lid] = spike;
            v.g_excitatory[lid] = bgl[0][lane];
            v.g_inhibitory[lid] = bgl[1][lane];
        }
    }
    if (live) {
#pragma unroll
        for (int t = 0; t < kMaxTraceSets; ++t) {
            const TraceSet& T = P.ts[t];
            if ((uint32_t)t < P.n_ts && gid >= T.pre_offset && gid < T.pre_offset + T.n_pre) {
                const uint32_t i = gid - T.pre_offset;
                float dd = flush_trace(tr_d[t] * T.decay_d);
                float rr = flush_trace(tr_r[t] * T.decay_r);
                if (spike == 1) {
                    rr += T.K;
                    dd += T.K;
                }
                T.d[i] = dd;
                T.r[i] = rr;
                T.x[i] = dd - rr;
            }
        }
    }
    // Spike bits of this step (64 global ids per wave -> two 32-bit words).
    const uint64_t ballot = __ballot(spike == 1);
    if (lane == 0 || lane == 32) {
        const uint32_t word = gid >> 5;
        if (word < P.words) {
            const uint32_t bits = lane == 0 ? (uint32_t)ballot : (uint32_t)(ballot >> 32);
            P.spike_bits[word] = bits;
            if (P.spike_bits_ring) P.spike_bits_ring[(size_t)(step % P.interval) * P.words + word] = bits;
        }
    }
    if (vslot >= 0) P.v_ring[(size_t)(step % P.interval) * P.n_vrec + vslot] = V;
}

// The plan is read in place from the kernel-argument segment, where the first explicit argument
// starts at offset 0: reading the by-value parameter itself lets the compiler copy the whole
// plan into private memory once the kernel reads enough of its fields.
__device__ __forceinline__ const EvolvePlan& evolve_plan_arg() {
    return *(const EvolvePlan*)__builtin_amdgcn_kernarg_segment_ptr();
}

__global__ __launch_bounds__(kFinParts * 64) void k_evolve(EvolvePlan) {
    const EvolvePlan& P = evolve_plan_arg();
    evolve_block(P, blockIdx.x, P.step_offset, P.stim_prev);
}

// Several networks in one launch: block -> (network, block of that network) from the tables.
__global__ __launch_bounds__(kFinParts * 64) void k_evolve_multi(const EvolvePlan* __restrict__ plans,
                                                                 const uint32_t* __restrict__ block_net,
                                                                 const uint32_t* __restrict__ net_first_block) {
    const uint32_t net = __builtin_amdgcn_readfirstlane(block_net[blockIdx.x]);
    evolve_block(plans[net], blockIdx.x - net_first_block[net], 0, 0);
}

// =======================================================================================
// k_gemv: FP64 partial sums, one wave per tile of (row chunk, 256 columns).
//   partial[chunk][j] = sum_{i in chunk, increasing} (double)x'_i * (double)W[i][j]
// Products of two binary32 numbers are exact in binary64, so the only rounding is the FP64
// accumulation in this fixed order, the part sums of the combination, then one rounding to
// FP32 when g is finalised.
// =======================================================================================
// x and the row scales are read-only during a propagate launch (the previous launch writes
// them): read through the constant address space, so that their wave-uniform loads stay scalar
// loads however the rest of the launch writes memory.
typedef const __attribute__((address_space(4))) float ConstFloat;
template <bool kScaled>
__device__ __forceinline__ float row_x(const float* __restrict__ x, const float* __restrict__ s, uint32_t i) {
    const ConstFloat* xc = (const ConstFloat*)x;
    const ConstFloat* sc = (const ConstFloat*)s;
    return kScaled ? xc[i] * sc[i] : xc[i];
}

// One tile: every index is wave-uniform, so x (and s) come through the scalar cache and the
// eight float4 row loads of an iteration are all in flight together.
template <bool kNT>
__device__ __forceinline__ float4 load_w4(const float* p) {
    if constexpr (kNT) {
        typedef float f4v __attribute__((ext_vector_type(4)));
        const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
        return make_float4(t.x, t.y, t.z, t.w);
    } else {
        return *reinterpret_cast<const float4*>(p);
    }
}

// Sums of rows [i0, i1) of the lane's 4 columns j0 .. j0+3, each in FP64 from 0.0 in row order.
template <bool kScaled, bool kNT>
__device__ __forceinline__ void gemv_tile_sums(const float* __restrict__ W, const float* __restrict__ x,
                                               const float* __restrict__ s, uint32_t ld, uint32_t i0, uint32_t i1,
                                               uint32_t j0, double& o0, double& o1, double& o2, double& o3) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const float* wp = W + (size_t)i0 * ld + j0;
    uint32_t i = i0;
    for (; i + 8 <= i1; i += 8) {
        float4 w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) w[k] = load_w4<kNT>(wp + (size_t)k * ld);
        float xs[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) xs[k] = row_x<kScaled>(x, s, i + k);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double xd = (double)xs[k];
            a0 += xd * (double)w[k].x;
            a1 += xd * (double)w[k].y;
            a2 += xd * (double)w[k].z;
            a3 += xd * (double)w[k].w;
        }
        wp += (size_t)8 * ld;
    }
    for (; i < i1; ++i) {
        const float4 w = load_w4<kNT>(wp);
        const double xd = (double)row_x<kScaled>(x, s, i);
        a0 += xd * (double)w.x;
        a1 += xd * (double)w.y;
        a2 += xd * (double)w.z;
        a3 += xd * (double)w.w;
        wp += ld;
    }
    o0 = a0;
    o1 = a1;
    o2 = a2;
    o3 = a3;
}

template <bool kScaled, bool kNT>
__device__ __forceinline__ void gemv_tile(const float* __restrict__ W, const float* __restrict__ x,
                                          const float* __restrict__ s, double* __restrict__ partials,
                                          uint32_t ld, uint32_t i0, uint32_t i1, uint32_t j0, uint32_t chunk) {
    double a0, a1, a2, a3;
    gemv_tile_sums<kScaled, kNT>(W, x, s, ld, i0, i1, j0, a0, a1, a2, a3);
    double* out = partials + (size_t)chunk * ld + j0;
    out[0] = a0;
    out[1] = a1;
    out[2] = a2;
    out[3] = a3;
}

template <bool kNT>
__device__ __forceinline__ void gemv_block(const GemvPlan& P, uint32_t blk) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blk * 4 + (threadIdx.x >> 6));
    const uint32_t lane = threadIdx.x & 63;
    if (wave >= P.total_tiles) return;
    uint32_t ci = 0;
    while (ci + 1 < P.n_conn && wave >= P.c[ci].tile_end) ++ci;
    const GemvConn& C = P.c[ci];
    const uint32_t t = wave - C.tile_begin;
    const uint32_t chunk = t / C.n_coltiles;
    const uint32_t q = t - chunk * C.n_coltiles;
    const uint32_t j0 = q * kGemvTileCols + 4 * lane;
    const uint32_t i0 = chunk * C.row_block;
    const uint32_t i1 = min(i0 + C.row_block, C.n_pre);
    if (j0 >= C.n_post) return;
    if (C.row_scale) gemv_tile<true, kNT>(C.W, C.x, C.row_scale, C.partials, C.ld, i0, i1, j0, chunk);
    else gemv_tile<false, kNT>(C.W, C.x, nullptr, C.partials, C.ld, i0, i1, j0, chunk);
}

// The propagate launch's extra last block (GemvPlan::pc), defined with the compaction helpers.
__device__ __forceinline__ void compact_pre_to_global(const PreCompact& pc);

template <bool kNT>
__global__ __launch_bounds__(256) void k_gemv(GemvPlan P) {
    if (P.pc.list && blockIdx.x == gridDim.x - 1) {
        compact_pre_to_global(P.pc);
        return;
    }
    gemv_block<kNT>(P, blockIdx.x);
}

// Part sums (the fused step's propagate): a block of kFinParts waves owns one part of a
// connection's row chunks -- part p is chunks [p * per, min((p + 1) * per, chunks)), per =
// ceil(chunks / kFinParts), the parts the evolve combines -- and one 256-column tile.  Each wave
// sums its chunks (chunk w, w + kFinParts, ... of the part) into LDS; the block adds the part's
// chunk sums in chunk order from 0.0, which is the part sum the evolve forms from the chunk
// partials, and stores it as row p.  The evolve then reads one row per part instead of one per
// chunk (same arithmetic).
template <bool kNT>
__device__ __forceinline__ void gemv_part_block(const GemvPlan& P, uint32_t blk) {
    __shared__ double cs[kGemvPartChunksMax][kGemvTileCols];
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (blk >= P.total_tiles) return;
    uint32_t ci = 0;
    while (ci + 1 < P.n_conn && blk >= P.c[ci].tile_end) ++ci;
    const GemvConn& C = P.c[ci];
    const uint32_t t = blk - C.tile_begin;
    const uint32_t part = t / C.n_coltiles;
    const uint32_t q = t - part * C.n_coltiles;
    const uint32_t per = fin_part_size(C.n_chunks);
    const uint32_t c0 = part * per, c1 = min(c0 + per, C.n_chunks);
    const uint32_t j0 = q * kGemvTileCols + 4 * lane;
    for (uint32_t c = c0 + wave; c < c1; c += (uint32_t)kFinParts) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        const uint32_t i0 = c * C.row_block, i1 = min(i0 + C.row_block, C.n_pre);
        if (j0 < C.n_post) {
            if (C.row_scale) gemv_tile_sums<true, kNT>(C.W, C.x, C.row_scale, C.ld, i0, i1, j0, a0, a1, a2, a3);
            else gemv_tile_sums<false, kNT>(C.W, C.x, nullptr, C.ld, i0, i1, j0, a0, a1, a2, a3);
        }
        double* row = cs[c - c0];
        row[4 * lane + 0] = a0;
        row[4 * lane + 1] = a1;
        row[4 * lane + 2] = a2;
        row[4 * lane + 3] = a3;
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)kGemvTileCols) {
        const uint32_t j = q * kGemvTileCols + threadIdx.x;
        if (j < C.n_post) {
            double acc = 0.0;
            for (uint32_t c = c0; c < c1; ++c) acc += cs[c - c0][threadIdx.x];
            C.partials[(size_t)part * C.ld + j] = acc;
        }
    }
}

template <bool kNT>
__global__ __launch_bounds__(kFinParts * 64) void k_gemv_parts(GemvPlan P) {
    if (P.pc.list && blockIdx.x == gridDim.x - 1) {
        compact_pre_to_global(P.pc);
        return;
    }
    gemv_part_block<kNT>(P, blockIdx.x);
}

template <bool kNT>
__global__ __launch_bounds__(256) void k_gemv_multi(const GemvPlan* __restrict__ plans,
                                                    const uint32_t* __restrict__ block_net,
                                                    const uint32_t* __restrict__ net_first_block) {
    const uint32_t net = __builtin_amdgcn_readfirstlane(block_net[blockIdx.x]);
    gemv_block<kNT>(plans[net], blockIdx.x - net_first_block[net]);
}

// Scalar-load variant for row strides that are not 16-byte aligned (reference layout with
// post % 4 != 0).  Same tiles, same summation order.
__global__ __launch_bounds__(256) void k_gemv_scalar(GemvPlan P) {
    if (P.pc.list && blockIdx.x == gridDim.x - 1) {
        compact_pre_to_global(P.pc);
        return;
    }
    const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (wave >= P.total_tiles) return;
    uint32_t ci = 0;
    while (ci + 1 < P.n_conn && wave >= P.c[ci].tile_end) ++ci;
    const GemvConn& C = P.c[ci];
    const uint32_t t = wave - C.tile_begin;
    const uint32_t chunk = t / C.n_coltiles;
    const uint32_t q = t - chunk * C.n_coltiles;
    const uint32_t i0 = chunk * C.row_block;
    const uint32_t i1 = min(i0 + C.row_block, C.n_pre);
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t j = q * kGemvTileCols + 4 * lane + k;
        if (j >= C.n_post) return;
        double a = 0.0;
        for (uint32_t i = i0; i < i1; ++i) {
            const float xi = C.row_scale ? C.x[i] * C.row_scale[i] : C.x[i];
            a += (double)xi * (double)C.W[(size_t)i * C.ld + j];
        }
        C.partials[(size_t)chunk * C.ld + j] = a;
    }
}

// g_dst[j] = fl32((double)g_base[j] + partials...) for each source in order.
__global__ __launch_bounds__(256) void k_materialize_g(float* g_dst, const float* g_base,
                                                      FinalizeSource s0, FinalizeSource s1,
                                                      uint32_t n_src, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float g = g_base[j];
    if (n_src > 0) g = finalize_g(g, s0, j);
    if (n_src > 1) g = finalize_g(g, s1, j);
    g_dst[j] = g;
}

__global__ __launch_bounds__(256) void k_add_increments(float* dst, const float* src, const uint32_t* spikes,
                                                        uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t = src[i];
    if (spikes[i] == 1) t += 1.0f;
    dst[i] = t;
}

// =======================================================================================
// Plasticity
// =======================================================================================

// LTP increment of entry (i, j): kernel_triplet_plasticity_update.cu:64-66,
// kernel_istdp_plasticity_update.cu:55, kernel_stdp_plasticity_update.cu:55.
__device__ __forceinline__ float ltp_factor(const PlastDev& D, uint32_t j) {
    return D.kind == kPlastTriplet ? D.p.ltp_A + (D.p.ltp_B * D.trip_post[j]) : D.p.ltp_A;
}
__device__ __forceinline__ float ltp_delta(const PlastDev& D, uint32_t i, float fac_j) {
    return D.kind == kPlastTriplet ? D.pre[i] * fac_j : D.p.ltp_A * D.pre[i];
}
// LTD increment of entry (i, j): kernel_triplet_plasticity_update.cu:85-87,
// kernel_istdp_plasticity_update.cu:77, kernel_stdp_plasticity_update.cu:77.
__device__ __forceinline__ float ltd_factor(const PlastDev& D, uint32_t i) {
    return D.kind == kPlastTriplet ? D.p.ltd_A + (D.p.ltd_B * D.trip_pre[i]) : 0.0f;
}
__device__ __forceinline__ float ltd_delta(const PlastDev& D, uint32_t j, float fac_i) {
    if (D.kind == kPlastTriplet) return -1.0f * D.post[j] * fac_i;
    if (D.kind == kPlastIstdp) return -1.0f * D.p.ltd_A * (D.post[j] - D.p.ltd_B);
    return -1.0f * D.p.ltd_A * D.post[j];
}
// One update of a stored value m in true units (lazy: w = s*m, stored back as w'/s).
__device__ __forceinline__ float apply_delta(const PlastDev& D, float m, float si, float delta) {
    if (D.lazy) return clamp_weight(si * m + delta, D.p.weight_min, D.p.weight_max) / si;
    return clamp_weight(m + delta, D.p.weight_min, D.p.weight_max);
}

// Canonical row sum: thread t of 256 sums columns t, t+256, ... in FP64, xor butterfly inside
// each wave (32,16,8,4,2,1), the four wave sums added in wave order.  `vals` holds the
// thread's columns.
// Combination of the per-thread FP64 partials of a row (thread t of 256 holds columns t, t+256,
// ... summed in increasing order): xor butterfly inside each wave (32,16,8,4,2,1), then the four
// wave sums added in wave order.  Every thread returns the same total.
__device__ __forceinline__ double block_reduce_sum(double acc, double* lds4) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((t & 63) == 0 && t < kNormThreads) lds4[t >> 6] = acc;
    __syncthreads();
    const double total = ((lds4[0] + lds4[1]) + lds4[2]) + lds4[3];
    __syncthreads();
    return total;
}

// Canonical row sum of a row held in registers (NR columns per thread).
template <int NR>
__device__ __forceinline__ double block_row_sum(const float* vals, uint32_t n, double* lds4) {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < NR; ++k)
        if (threadIdx.x + (uint32_t)k * kNormThreads < n) acc += (double)vals[k];
    return block_reduce_sum(acc, lds4);
}

// Spiking neurons of this step as a list in LDS: the set bits of global range [g0, g1), as
// offsets from g0 in increasing order, compacted by the whole block (popcount + block scan).
// Returns the count (the same in every thread); entries past `cap` are not stored.
constexpr uint32_t kListCap = 4096;
constexpr uint32_t kMaxBlockWaves = 8;  // blocks of 256 or 512 threads (the latter: overlapped step)
__device__ __forceinline__ bool spike_bit(const uint32_t* bits, uint32_t gid) { return (bits[gid >> 5] >> (gid & 31)) & 1u; }

__device__ __forceinline__ uint32_t block_compact_bits(const uint32_t* bits, uint32_t g0, uint32_t g1,
                                                       uint32_t* list, uint32_t cap, uint32_t* wsum) {
    if (g1 <= g0) return 0;
    const uint32_t w0 = g0 >> 5, nw = ((g1 - 1) >> 5) - w0 + 1;
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint32_t total = 0;
    for (uint32_t base = 0; base < nw; base += blockDim.x) {
        const uint32_t wi = base + t;
        const uint32_t w = w0 + wi;
        uint32_t word = 0;
        if (wi < nw) {
            word = bits[w];
            if (w == (g0 >> 5)) word &= 0xFFFFFFFFu << (g0 & 31);
            if (w == ((g1 - 1) >> 5)) word &= 0xFFFFFFFFu >> (31 - ((g1 - 1) & 31));
        }
        const uint32_t c = (uint32_t)__popc(word);
        uint32_t v = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t n = __shfl_up(v, off, 64);
            if (lane >= (uint32_t)off) v += n;
        }
        if (lane == 63) wsum[wv] = v;
        __syncthreads();
        uint32_t wbase = 0, btot = 0;
        const uint32_t nwaves = blockDim.x >> 6;
#pragma unroll
        for (uint32_t k = 0; k < kMaxBlockWaves; ++k) {
            if (k >= nwaves) break;
            if (k < wv) wbase += wsum[k];
            btot += wsum[k];
        }
        uint32_t pos = total + wbase + v - c;
        while (word) {
            const uint32_t b = (uint32_t)__ffs(word) - 1;
            word &= word - 1;
            if (pos < cap) list[pos] = w * 32 + b - g0;
            ++pos;
        }
        total += btot;
        __syncthreads();
    }
    return total;
}

// PreCompact: the spiking pre neurons of every plasticity into the global list, in order.
__device__ __forceinline__ void compact_pre_to_global(const PreCompact& pc) {
    __shared__ uint32_t wsum[kMaxBlockWaves];
    uint32_t base = 0;
    for (uint32_t k = 0; k < pc.n; ++k) {
        const uint32_t g0 = pc.pre_offset[k], g1 = g0 + pc.n_pre[k];
        const uint32_t at = min(base, kPreListCap);
        const uint32_t count = block_compact_bits(pc.bits, g0, g1, pc.list + at, kPreListCap - at, wsum);
        if (threadIdx.x == 0) pc.counts[k] = count;
        base += count;
    }
}

// Spiking pre neurons of every plasticity in one pass: the set bits of each plasticity's pre
// range [pre_offset, pre_offset + n_pre) of its spike bits, as pre-local ids, in one list in
// plasticity order (each range in increasing order); counts[k] (LDS) receives the count of
// plasticity k.  Returns the total (the same in every thread); entries past `cap` are not stored.
__device__ __forceinline__ uint32_t block_compact_pre(const PlastPlan& pl, uint32_t* list, uint32_t cap,
                                                      uint32_t* wsum, uint32_t* counts) {
    uint32_t vend[kMaxPlast];  // end of plasticity k's words in the concatenated word sequence
    uint32_t nv = 0;
#pragma unroll
    for (int k = 0; k < kMaxPlast; ++k) {
        if ((uint32_t)k < pl.n && pl.p[k].n_pre > 0)
            nv += ((pl.p[k].pre_offset + pl.p[k].n_pre - 1) >> 5) - (pl.p[k].pre_offset >> 5) + 1;
        vend[k] = nv;
    }
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (t < (uint32_t)kMaxPlast) counts[t] = 0;
    __syncthreads();
    uint32_t total = 0;
    for (uint32_t base = 0; base < nv; base += blockDim.x) {
        const uint32_t vi = base + t;
        uint32_t word = 0, w = 0, k = 0, g0 = 0;
        if (vi < nv) {
            uint32_t vstart = 0;
#pragma unroll
            for (int kk = 0; kk + 1 < kMaxPlast; ++kk)
                if (vi >= vend[kk]) {
                    k = kk + 1;
                    vstart = vend[kk];
                }
            const uint32_t* bits = nullptr;
            uint32_t g1 = 0;
#pragma unroll
            for (int kk = 0; kk < kMaxPlast; ++kk)
                if ((uint32_t)kk == k) {
                    bits = uni(pl.p[kk].bits);
                    g0 = uni(pl.p[kk].pre_offset);
                    g1 = g0 + uni(pl.p[kk].n_pre);
                }
            w = (g0 >> 5) + (vi - vstart);
            word = bits[w];
            if (w == (g0 >> 5)) word &= 0xFFFFFFFFu << (g0 & 31);
            if (w == ((g1 - 1) >> 5)) word &= 0xFFFFFFFFu >> (31 - ((g1 - 1) & 31));
        }
        const uint32_t c = (uint32_t)__popc(word);
        uint32_t v = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t m = __shfl_up(v, off, 64);
            if (lane >= (uint32_t)off) v += m;
        }
        if (lane == 63) wsum[wv] = v;
        if (c) atomicAdd(&counts[k], c);
        __syncthreads();
        uint32_t wbase = 0, btot = 0;
        const uint32_t nwaves = blockDim.x >> 6;
#pragma unroll
        for (uint32_t q = 0; q < kMaxBlockWaves; ++q) {
            if (q >= nwaves) break;
            if (q < wv) wbase += wsum[q];
            btot += wsum[q];
        }
        uint32_t pos = total + wbase + v - c;
        while (word) {
            const uint32_t b = (uint32_t)__ffs(word) - 1;
            word &= word - 1;
            if (pos < cap) list[pos] = w * 32 + b - g0;
            ++pos;
        }
        total += btot;
        __syncthreads();
    }
    return total;
}

// LTP entries (i, js[0..n)) of a quiet row, in list order: quiet_row_load loads the entries and
// the post traces of their factors (ltp_factor of each j) together, quiet_row_apply updates and
// stores them.
constexpr int kLtpBatch = 16;
__device__ __forceinline__ void quiet_row_load(const PlastDev& D, const float* row, const uint32_t* js, int n,
                                               float (&m)[kLtpBatch], float (&fs)[kLtpBatch]) {
    const bool triplet = D.kind == kPlastTriplet;
#pragma unroll
    for (int k = 0; k < kLtpBatch; ++k)
        if (k < n) {
            m[k] = row[js[k]];
            fs[k] = triplet ? D.trip_post[js[k]] : 0.0f;
        }
}
__device__ __forceinline__ void quiet_row_apply(const PlastDev& D, uint32_t i, float* row, float si, float pre_i,
                                                const uint32_t* js, int n, const float (&m)[kLtpBatch],
                                                const float (&fs)[kLtpBatch], double& dsum) {
    const bool triplet = D.kind == kPlastTriplet;
#pragma unroll
    for (int k = 0; k < kLtpBatch; ++k) {
        if (k >= n || js[k] == i) continue;
        const float delta = triplet ? pre_i * (D.p.ltp_A + (D.p.ltp_B * fs[k])) : D.p.ltp_A * pre_i;
        const float mn = apply_delta(D, m[k], si, delta);
        row[js[k]] = mn;
        if (D.lazy) dsum += (double)mn - (double)m[k];
    }
}

// Row i whose pre neuron did NOT spike: the LTP entries of this step's spiking posts.  The list
// of spiking posts (increasing post order, LDS) is cut into kQuietParts contiguous parts of
// ceil(count / kQuietParts); this call does part q in list order (straight from the spike bits
// when the list overflowed) and returns the part's sum of the changes of M (lazy), summed from
// 0.0.  The row's scale, its pre trace and the first batch of entries are loaded together; a
// row with s_i = 0 is left as it is.
constexpr uint32_t kQuietParts = 4;
constexpr uint32_t kQuietRows = kPlastQuietRows;
__device__ __forceinline__ double plasticity_quiet_part(const PlastDev& D, uint32_t i, const uint32_t* list,
                                                        uint32_t count, uint32_t q) {
    const bool lazy = D.lazy != 0;
    double dsum = 0.0;
    const uint32_t per = (count + kQuietParts - 1) / kQuietParts;
    const uint32_t e0 = q * per, e1 = min(count, e0 + per);
    if (e0 >= e1) return 0.0;
    const float si = lazy ? D.s[i] : 1.0f;
    float* row = D.W + (size_t)i * D.ld;
    const float pre_i = D.pre[i];
    uint32_t js[kLtpBatch];
    if (count <= kListCap) {
        for (uint32_t k0 = e0; k0 < e1; k0 += kLtpBatch) {
            const int n = (int)min((uint32_t)kLtpBatch, e1 - k0);
#pragma unroll
            for (int k = 0; k < kLtpBatch; ++k) js[k] = k < n ? list[k0 + k] : 0u;
            float m[kLtpBatch], fs[kLtpBatch];
            quiet_row_load(D, row, js, n, m, fs);
            if (lazy && si == 0.0f) return 0.0;
            quiet_row_apply(D, i, row, si, pre_i, js, n, m, fs, dsum);
        }
    } else {
        if (lazy && si == 0.0f) return 0.0;
        const uint32_t g0 = D.post_offset, g1 = D.post_offset + D.n_post;
        uint32_t rank = 0;
        for (uint32_t w = g0 >> 5; w <= (g1 - 1) >> 5 && rank < e1; ++w) {
            uint32_t bits = D.bits[w];
            if (w == (g0 >> 5)) bits &= 0xFFFFFFFFu << (g0 & 31);
            if (w == ((g1 - 1) >> 5)) bits &= 0xFFFFFFFFu >> (31 - ((g1 - 1) & 31));
            while (bits) {
                const uint32_t b = (uint32_t)__ffs(bits) - 1;
                bits &= bits - 1;
                if (rank >= e0 && rank < e1) {
                    js[0] = w * 32 + b - g0;
                    float m[kLtpBatch], fs[kLtpBatch];
                    quiet_row_load(D, row, js, 1, m, fs);
                    quiet_row_apply(D, i, row, si, pre_i, js, 1, m, fs, dsum);
                }
                ++rank;
            }
        }
    }
    return dsum;
}

// Row i whose pre neuron spiked, by the whole block: per entry the LTP increment (if the post
// neuron spiked) then the LTD increment, as the reference's LTP-then-LTD kernel order.
// Lazy: the row's sum of M is recomputed in the canonical order.
__device__ __forceinline__ void plasticity_spiking_row(const PlastDev& D, uint32_t i, double* lds4) {
    const bool lazy = D.lazy != 0;
    const float si = lazy ? D.s[i] : 1.0f;
    const float target = lazy ? D.target[i] : 0.0f;
    const bool update = !lazy || si != 0.0f;
    const float fac_i = ltd_factor(D, i);
    float* row = D.W + (size_t)i * D.ld;
    const bool triplet = D.kind == kPlastTriplet;
    const float pre_i = D.pre[i];
    double acc = 0.0;
    // Batches of 16 columns per thread: all loads first, then updates/stores in column order.
    constexpr int kB = 16;
    const uint32_t j_first = threadIdx.x < kNormThreads ? threadIdx.x : D.n_post;  // threads past 256 idle
    for (uint32_t j0 = j_first; j0 < D.n_post; j0 += kB * kNormThreads) {
        float m[kB], post_j[kB], tpost_j[kB];
        uint32_t spw[kB];
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const uint32_t j = j0 + (uint32_t)k * kNormThreads;
            if (j < D.n_post) {
                m[k] = row[j];
                spw[k] = D.bits[(D.post_offset + j) >> 5];
                post_j[k] = D.post[j];
                tpost_j[k] = triplet ? D.trip_post[j] : 0.0f;
            }
        }
#pragma unroll
        for (int k = 0; k < kB; ++k) {
            const uint32_t j = j0 + (uint32_t)k * kNormThreads;
            if (j >= D.n_post) continue;
            float v = m[k];
            if (update && j != i) {
                if ((spw[k] >> ((D.post_offset + j) & 31)) & 1u) {
                    const float fac = triplet ? D.p.ltp_A + (D.p.ltp_B * tpost_j[k]) : D.p.ltp_A;
                    v = apply_delta(D, v, si, triplet ? pre_i * fac : D.p.ltp_A * pre_i);
                }
                float d;
                if (triplet) d = -1.0f * post_j[k] * fac_i;
                else if (D.kind == kPlastIstdp) d = -1.0f * D.p.ltd_A * (post_j[k] - D.p.ltd_B);
                else d = -1.0f * D.p.ltd_A * post_j[k];
                v = apply_delta(D, v, si, d);
                row[j] = v;
            }
            acc += (double)v;
        }
    }
    if (lazy) {
        const double R = block_reduce_sum(acc, lds4);
        if (threadIdx.x == 0) {
            D.R[i] = R;
            if (R > 0.0) D.s[i] = (float)((double)target / R);
        }
    }
}

// Stimulus events of clock_step in list order (stimulus.cpp:46-60), then the clock advance.
// One thread: the event count per step is small, and sequential order makes the sum
// independent of scheduling (the reference uses atomicAdd).
__device__ __forceinline__ void stimulus_and_clock(const StimTail& T) {
    bn_device_clock* clk = T.clock;
    const uint32_t base = clk->clock_step;
    const uint32_t s = base + T.step_offset;
    if (T.timestep_event_count && s < T.num_timesteps) {
        const uint32_t n = T.timestep_event_count[s];
        const uint32_t cur = T.ev_start ? T.ev_start[s] : clk->next_event_idx;
        if ((uint64_t)cur + n > T.num_events) {
            *T.status = 1u;
        } else {
            for (uint32_t e = cur; e < cur + n; ++e)
                T.i_stimulus[T.event_ids[e] + T.event_offsets[e]] += T.event_amplitudes[e];
            clk->next_event_idx = cur + n;
        }
    }
    clk->clock_step = base + (T.advance ? T.advance : 1u);
}

// Blocks [0, row_blocks): kQuietRows rows each of the pre rows of all plasticities, each
// plasticity's rows padded to a multiple of kQuietRows so that a block serves one plasticity
// (rows whose pre did not spike).  Blocks after: the rows whose pre spiked, one block each.
// The last block also applies the stimulus and advances the clock (no block reads the clock).
__device__ __forceinline__ void plasticity_block(const PlastStepPlan& P, uint32_t b, uint32_t nblocks) {
    __shared__ double lds4[4];
    __shared__ uint32_t wsum[kMaxBlockWaves];
    __shared__ uint32_t list[kListCap];
    __shared__ double qsum[kQuietParts][64];
    __shared__ uint32_t rcount[kMaxPlast];
    if (b < P.row_blocks) {
        // kQuietRows rows per block, lane = row; wave q applies part q of the spiking-post list.
        const uint32_t r0 = b * kQuietRows;
        const uint32_t q = threadIdx.x >> 6, lane = threadIdx.x & 63;
        uint32_t base = 0;
        for (uint32_t k = 0; k < P.pl.n; ++k) {
            const PlastDev& D = P.pl.p[k];
            if (r0 < base + D.n_pre && r0 + kQuietRows > base) {
                // The row's own spike bit is loaded with the words the list is compacted from.
                const uint32_t r = r0 + lane;
                const bool in_rows = r >= base && r < base + D.n_pre;
                const uint32_t own = in_rows ? D.pre_offset + (r - base) : 0u;
                const uint32_t own_word = in_rows ? D.bits[own >> 5] : 0u;
                const uint32_t count =
                    block_compact_bits(D.bits, D.post_offset, D.post_offset + D.n_post, list, kListCap, wsum);
                const bool mine = in_rows && !((own_word >> (own & 31)) & 1u);
                double R0 = 0.0;
                float target = 0.0f;
                if (D.lazy && q == 0 && mine) {
                    R0 = D.R[r - base];
                    target = D.target[r - base];
                }
                const double dq = mine && count ? plasticity_quiet_part(D, r - base, list, count, q) : 0.0;
                if (D.lazy) {
                    if (q < kQuietParts) qsum[q][lane] = dq;
                    __syncthreads();
                    if (q == 0 && mine) {
                        // R_i += the part sums in part order (non-empty parts only).
                        const uint32_t per = (count + kQuietParts - 1) / kQuietParts;
                        double R = R0;
                        for (uint32_t p = 0; p < kQuietParts; ++p)
                            if (p * per < count) R += qsum[p][lane];
                        D.R[r - base] = R;
                        if (R > 0.0) D.s[r - base] = (float)((double)target / R);
                    }
                }
                __syncthreads();
            }
            base += (D.n_pre + kQuietRows - 1) / kQuietRows * kQuietRows;
        }
    } else if (b < P.row_blocks + P.spiking_blocks) {
        // Spiking rows of all plasticities in one sequence (plasticity order, increasing row);
        // block sb takes the entries e = sb, sb + spiking_blocks, ...  The spiking pre neurons of
        // every plasticity are compacted in one pass.
        const uint32_t sb = b - P.row_blocks;
        const uint32_t G = P.spiking_blocks;
        // The list compacted by this step's propagate launch (the counts and this block's first
        // entry in one round trip; spiking_blocks <= kPreListCap), or a compaction here.
        const bool global_list = P.pre_list != nullptr;
        uint32_t cnt[kMaxPlast];
        uint32_t first_entry = 0;
        if (global_list) {
#pragma unroll
            for (int k = 0; k < kMaxPlast; ++k) cnt[k] = P.pre_counts[k];
            first_entry = P.pre_list[sb];
        } else {
            block_compact_pre(P.pl, list, kListCap, wsum, rcount);
#pragma unroll
            for (int k = 0; k < kMaxPlast; ++k) cnt[k] = rcount[k];
        }
        const uint32_t cap = global_list ? kPreListCap : kListCap;
        uint32_t base = 0;
#pragma unroll
        for (int k = 0; k < kMaxPlast; ++k) {
            if ((uint32_t)k >= P.pl.n) break;
            const PlastDev& D = P.pl.p[k];
            const uint32_t count = cnt[k];
            const uint32_t first = base + (sb + G - base % G) % G;
            if (base + count <= cap) {
                for (uint32_t e = first; e < base + count; e += G) {
                    const uint32_t i = !global_list ? list[e] : e == sb ? first_entry : P.pre_list[e];
                    plasticity_spiking_row(D, i, lds4);
                }
            } else {
                uint32_t rank = 0;
                for (uint32_t i = 0; i < D.n_pre; ++i)
                    if (spike_bit(D.bits, D.pre_offset + i)) {
                        if ((base + rank) % G == sb) plasticity_spiking_row(D, i, lds4);
                        ++rank;
                    }
            }
            base += count;
        }
    }
    if (P.tail.active && b == nblocks - 1 && threadIdx.x == 0) stimulus_and_clock(P.tail);
}

__global__ __launch_bounds__(256) void k_plasticity(PlastStepPlan P) { plasticity_block(P, blockIdx.x, gridDim.x); }

// Overlapped step: blocks [0, n_evolve) evolve step t, the rest run the plasticity of step t-1
// (its 256-thread body; the upper four waves take part in its barriers only).
__global__ __launch_bounds__(kFinParts * 64) void k_evolve_plast(EvolvePlan, PlastStepPlan P, uint32_t n_evolve) {
    const EvolvePlan& E = evolve_plan_arg();  // the first argument
    if (blockIdx.x < n_evolve) evolve_block(E, blockIdx.x, E.step_offset, E.stim_prev);
    else plasticity_block(P, blockIdx.x - n_evolve, gridDim.x - n_evolve);
}

// The same for several networks: evolve blocks of every network (step offset and previous-step
// stimulus of the launch), then the plasticity blocks of every network (plast = nullptr: none).
__global__ __launch_bounds__(kFinParts * 64) void k_evolve_plast_multi(const EvolvePlan* __restrict__ ev, MultiPhase pe,
                                                                       const PlastStepPlan* __restrict__ pl, MultiPhase pp,
                                                                       uint32_t step_offset, uint32_t stim_prev) {
    if (blockIdx.x < pe.blocks) {
        const uint32_t net = __builtin_amdgcn_readfirstlane(pe.block_net[blockIdx.x]);
        evolve_block(ev[net], blockIdx.x - pe.net_first_block[net], step_offset, stim_prev);
    } else {
        const uint32_t b = blockIdx.x - pe.blocks;
        const uint32_t net = __builtin_amdgcn_readfirstlane(pp.block_net[b]);
        const PlastStepPlan& P = pl[net];
        plasticity_block(P, b - pp.net_first_block[net], P.row_blocks + P.spiking_blocks);
    }
}

__global__ __launch_bounds__(256) void k_plasticity_multi(const PlastStepPlan* __restrict__ plans,
                                                          const uint32_t* __restrict__ block_net,
                                                          const uint32_t* __restrict__ net_first_block) {
    const uint32_t net = __builtin_amdgcn_readfirstlane(block_net[blockIdx.x]);
    const PlastStepPlan& P = plans[net];
    plasticity_block(P, blockIdx.x - net_first_block[net], P.row_blocks + P.spiking_blocks);
}


__device__ __forceinline__ bool locate_item(const PlastPlan& P, uint32_t item, bool post,
                                            uint32_t& pi, uint32_t& idx) {
    for (uint32_t k = 0; k < P.n; ++k) {
        const uint32_t n = post ? P.p[k].n_post : P.p[k].n_pre;
        if (item < n) {
            pi = k;
            idx = item;
            return true;
        }
        item -= n;
    }
    return false;
}

// Component-path LTP: one block per spiking post column, lanes over pre rows
// (kernel_triplet_plasticity_update.cu:54-72 with the index-based self skip).
__global__ __launch_bounds__(256) void k_ltp(PlastPlan P) {
    for (uint32_t item = blockIdx.x; item < P.items_post; item += gridDim.x) {
        uint32_t pi, j;
        if (!locate_item(P, item, true, pi, j)) return;
        const PlastDev& D = P.p[pi];
        if (D.post_spikes[j] != 1) continue;
        const float fac = ltp_factor(D, j);
        for (uint32_t i = threadIdx.x; i < D.n_pre; i += blockDim.x) {
            if (i == j) continue;
            float* w = D.W + (size_t)i * D.ld + j;
            *w = clamp_weight(*w + ltp_delta(D, i, fac), D.p.weight_min, D.p.weight_max);
        }
    }
}

// Component-path LTD: one block per spiking pre row, lanes over post columns
// (kernel_triplet_plasticity_update.cu:74-93).
__global__ __launch_bounds__(256) void k_ltd(PlastPlan P) {
    for (uint32_t item = blockIdx.x; item < P.items_pre; item += gridDim.x) {
        uint32_t pi, i;
        if (!locate_item(P, item, false, pi, i)) return;
        const PlastDev& D = P.p[pi];
        if (D.pre_spikes[i] != 1) continue;
        const float fac = ltd_factor(D, i);
        float* row = D.W + (size_t)i * D.ld;
        for (uint32_t j = threadIdx.x; j < D.n_post; j += blockDim.x) {
            if (i == j) continue;
            row[j] = clamp_weight(row[j] + ltd_delta(D, j, fac), D.p.weight_min, D.p.weight_max);
        }
    }
}

// Triplet row normalisation (kernel_triplet_plasticity_update.cu:117-145) with the canonical
// FP64 row sum: skip if sum <= 0, else scale the row by target/sum (target = pre_weight_sums[i],
// the initial column sum, triplet_plasticity.cpp:22-30).
template <int NR>
__device__ __forceinline__ void normalize_row(const PlastDev& D, uint32_t i, double* lds4) {
    float* row = D.W + (size_t)i * D.ld;
    float vals[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const uint32_t j = threadIdx.x + (uint32_t)k * kNormThreads;
        vals[k] = j < D.n_post ? row[j] : 0.0f;
    }
    const double total = block_row_sum<NR>(vals, D.n_post, lds4);
    const float current_sum = (float)total;
    if (current_sum <= 0.0f) return;
    const float scale = D.target[i] / current_sum;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const uint32_t j = threadIdx.x + (uint32_t)k * kNormThreads;
        if (j < D.n_post) row[j] = vals[k] * scale;
    }
}

__device__ __forceinline__ void increment_traces(const PlastPlan& P, uint32_t gid) {
    // kernel_triplet_plasticity_update.cu:95-115 (and istdp/stdp equivalents)
    for (uint32_t k = 0; k < P.n; ++k) {
        const PlastDev& D = P.p[k];
        if (gid >= D.pre_offset && gid < D.pre_offset + D.n_pre) {
            const uint32_t i = gid - D.pre_offset;
            if (D.pre_spikes[i] == 1) {
                D.pre[i] += 1.0f;
                if (D.kind == kPlastTriplet) D.trip_pre[i] += 1.0f;
            }
        }
        if (gid >= D.post_offset && gid < D.post_offset + D.n_post) {
            const uint32_t j = gid - D.post_offset;
            if (D.post_spikes[j] == 1) {
                D.post[j] += 1.0f;
                if (D.kind == kPlastTriplet) D.trip_post[j] += 1.0f;
            }
        }
    }
}

// Blocks [0, norm_rows): one row each.  Blocks after: trace increments for 256 neurons each
// (component path).  The last block also applies the stimulus and advances the clock when the
// tail is active (no other block reads the clock in this kernel).
template <int NR>
__global__ __launch_bounds__(256) void k_norm_tail(TailPlan T) {
    __shared__ double lds4[4];
    const uint32_t b = blockIdx.x;
    if (b < T.norm_rows) {
        uint32_t rem = b;
        for (uint32_t k = 0; k < T.pl.n; ++k) {
            const PlastDev& D = T.pl.p[k];
            if (!D.normalize) continue;
            if (rem < D.n_pre) {
                normalize_row<NR>(D, rem, lds4);
                break;
            }
            rem -= D.n_pre;
        }
    } else {
        const uint32_t gid = (b - T.norm_rows) * blockDim.x + threadIdx.x;
        if (T.increments && gid < T.total_neurons) increment_traces(T.pl, gid);
    }
    if (T.tail.active && b == gridDim.x - 1 && threadIdx.x == 0) stimulus_and_clock(T.tail);
}

template <int NR>
__global__ __launch_bounds__(256) void k_lazy_init(PlastDev D) {
    __shared__ double lds4[4];
    const uint32_t i = blockIdx.x;
    if (i >= D.n_pre) return;
    float vals[NR];
    const float* row = D.W + (size_t)i * D.ld;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const uint32_t j = threadIdx.x + (uint32_t)k * kNormThreads;
        vals[k] = j < D.n_post ? row[j] : 0.0f;
    }
    const double R = block_row_sum<NR>(vals, D.n_post, lds4);
    if (threadIdx.x == 0) {
        D.R[i] = R;
        D.s[i] = 1.0f;
    }
}

// =======================================================================================
// Component (drop-in) kernels
// =======================================================================================
__global__ __launch_bounds__(256) void k_group_evolve(NeuronGroupView v) {
    const uint32_t lid = blockIdx.x * blockDim.x + threadIdx.x;
    if (lid >= v.num_neurons) return;
    const bn_device_clock clk = *v.rng_clock;
    float V = v.membrane_potentials[lid];
    uint32_t refr = v.refractory_counter[lid];
    uint32_t spike = 0;
    eif_step(v, v.g_excitatory[lid], v.g_inhibitory[lid], v.i_stimulus[lid], V, refr, spike);
    v.membrane_potentials[lid] = V;
    v.refractory_counter[lid] = refr;
    v.spikes[lid] = spike;
    float bg_ge, bg_gi;
    background_step(v, lid, clk.clock_step, clk, bg_ge, bg_gi);
    v.g_excitatory[lid] = bg_ge;
    v.g_inhibitory[lid] = bg_gi;
}

__global__ __launch_bounds__(256) void k_connection_traces(ConnectionView v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.pre_group_num_neurons) return;
    connection_trace_step(v.postsynaptic_r_traces, v.postsynaptic_d_traces,
                          v.summed_postsynaptic_conductances, v.postsynaptic_r_trace_decay,
                          v.postsynaptic_d_trace_decay, v.postsynaptic_kernel_normalization_value,
                          i, v.spikes[i]);
}

__global__ __launch_bounds__(256) void k_plast_decay(PlastPlan P) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = 0; k < P.n; ++k) {
        const PlastDev& D = P.p[k];
        if (idx < D.n_pre) {
            D.pre[idx] = flush_trace(D.pre[idx] * D.p.pre_trace_decay);
            if (D.kind == kPlastTriplet)
                D.trip_pre[idx] = flush_trace(D.trip_pre[idx] * D.p.triplet_pre_trace_decay);
        }
        if (idx < D.n_post) {
            D.post[idx] = flush_trace(D.post[idx] * D.p.post_trace_decay);
            if (D.kind == kPlastTriplet)
                D.trip_post[idx] = flush_trace(D.trip_post[idx] * D.p.triplet_post_trace_decay);
        }
    }
}

__global__ __launch_bounds__(256) void k_plast_increment(PlastPlan P, uint32_t total_neurons) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < total_neurons) increment_traces(P, gid);
}

__global__ void k_stimulus_apply(StimulusView v, uint32_t first, uint32_t n) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (uint32_t e = first; e < first + n; ++e)
        v.i_stimulus[v.stimulus_event_neuron_ids[e] + v.stimulus_event_post_group_offsets[e]] +=
            v.stimulus_event_amplitude_changes[e];
}

__global__ void k_stimulus_device_cursor(StimulusView v, bn_device_clock* clk) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint32_t s = clk->clock_step;
    if (s >= v.num_timesteps) return;
    const uint32_t n = v.timestep_event_count[s];
    const uint32_t cur = clk->next_event_idx;
    if ((uint64_t)cur + n > v.num_events) return;
    for (uint32_t e = cur; e < cur + n; ++e)
        v.i_stimulus[v.stimulus_event_neuron_ids[e] + v.stimulus_event_post_group_offsets[e]] +=
            v.stimulus_event_amplitude_changes[e];
    clk->next_event_idx = cur + n;
}

__global__ void k_clock_advance(bn_device_clock* clk) {
    if (blockIdx.x == 0 && threadIdx.x == 0) clk->clock_step = clk->clock_step + 1;
}

__global__ void k_clock_init(bn_device_clock* clk, uint32_t key_lo, uint32_t key_hi) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk->clock_step = 0;
        clk->next_event_idx = 0;
        clk->rng_key[0] = key_lo;
        clk->rng_key[1] = key_hi;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_historicize(const T* src, T* hist, uint32_t n,
                                                     uint32_t interval, uint32_t step) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    hist[(size_t)id * interval + step % interval] = src[id];
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t value, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// =======================================================================================
// Host-side launch helpers
// =======================================================================================
// Read-bandwidth probe: every float4 of the buffer read once per pass (grid-stride), folded into
// one value per thread so the loads are kept.
// Streaming read of a buffer: kProbeLoads independent float4 loads per thread and pass, all in
// flight together (the grid is sized so that one pass covers the buffer when it can); with the
// propagate's cache policy (non-temporal loads above 192 MB).
constexpr uint32_t kProbeLoads = 8;
template <bool kNT>
__global__ __launch_bounds__(256) void k_read_probe(const float4* __restrict__ p, size_t n4, float* sink) {
    float acc[kProbeLoads] = {};
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (kProbeLoads - 1) * stride < n4; i += kProbeLoads * stride) {
        float4 v[kProbeLoads];
#pragma unroll
        for (uint32_t u = 0; u < kProbeLoads; ++u) v[u] = load_w4<kNT>(reinterpret_cast<const float*>(p + i + u * stride));
#pragma unroll
        for (uint32_t u = 0; u < kProbeLoads; ++u) acc[u] += (v[u].x + v[u].y) + (v[u].z + v[u].w);
    }
    for (; i < n4; i += stride) acc[0] += (p[i].x + p[i].y) + (p[i].z + p[i].w);
    float a = 0.0f;
#pragma unroll
    for (uint32_t u = 0; u < kProbeLoads; ++u) a += acc[u];
    if (a == 12345.678f) sink[blockIdx.x] = a;  // practically never; the sum must not be dead
}

// Many passes over the buffer inside one launch (no launch ramp or drain between passes).  Pass
// p starts at element p * (n4 / 8 + 4099) (mod n4): blocks read a different part of the buffer
// every pass, so a block's XCD does not find last pass's lines in its own L2 and the reads are
// served past the L2 (from the MALL when the buffer fits it).  Accesses stay coalesced except
// where a wave straddles the wrap.
template <bool kNT>
__global__ __launch_bounds__(256) void k_read_probe_passes(const float4* __restrict__ p, size_t n4, uint32_t passes,
                                                           float* sink) {
    float acc[kProbeLoads] = {};
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t rot = n4 / 8 + 4099;
    size_t base = 0;
    for (uint32_t pass = 0; pass < passes; ++pass) {
        size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        for (; i + (kProbeLoads - 1) * stride < n4; i += kProbeLoads * stride) {
            float4 v[kProbeLoads];
#pragma unroll
            for (uint32_t u = 0; u < kProbeLoads; ++u) {
                size_t k = i + u * stride + base;
                if (k >= n4) k -= n4;
                v[u] = load_w4<kNT>(reinterpret_cast<const float*>(p + k));
            }
#pragma unroll
            for (uint32_t u = 0; u < kProbeLoads; ++u) acc[u] += (v[u].x + v[u].y) + (v[u].z + v[u].w);
        }
        for (; i < n4; i += stride) {
            size_t k = i + base;
            if (k >= n4) k -= n4;
            acc[0] += (p[k].x + p[k].y) + (p[k].z + p[k].w);
        }
        base += rot;
        while (base >= n4) base -= n4;
    }
    float a = 0.0f;
#pragma unroll
    for (uint32_t u = 0; u < kProbeLoads; ++u) a += acc[u];
    if (a == 12345.678f) sink[blockIdx.x] = a;  // practically never; the sum must not be dead
}

int launch_read_probe_passes(const void* buf, uint64_t bytes, uint32_t passes, float* sink, uint32_t blocks,
                             bn_stream_t s) {
    const uint64_t n4 = bytes / 16;
    if (bytes > (192ull << 20))
        hipLaunchKernelGGL(k_read_probe_passes<true>, dim3(blocks), dim3(256), 0, s, static_cast<const float4*>(buf),
                           (size_t)n4, passes, sink);
    else
        hipLaunchKernelGGL(k_read_probe_passes<false>, dim3(blocks), dim3(256), 0, s, static_cast<const float4*>(buf),
                           (size_t)n4, passes, sink);
    return (int)hipGetLastError();
}

// `blocks` bounds the grid (and the sink's length); fewer are launched when kProbeLoads loads
// per thread cover the buffer in one pass.
int launch_read_probe(const void* buf, uint64_t bytes, float* sink, uint32_t blocks, bn_stream_t s) {
    const uint64_t n4 = bytes / 16;
    const uint64_t one_pass = (n4 + 256ull * kProbeLoads - 1) / (256ull * kProbeLoads);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(blocks, one_pass));
    if (bytes > (192ull << 20))
        hipLaunchKernelGGL(k_read_probe<true>, dim3(grid), dim3(256), 0, s, static_cast<const float4*>(buf), (size_t)n4, sink);
    else
        hipLaunchKernelGGL(k_read_probe<false>, dim3(grid), dim3(256), 0, s, static_cast<const float4*>(buf), (size_t)n4, sink);
    return (int)hipGetLastError();
}

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

int launch_evolve(const EvolvePlan& plan, bn_stream_t s) {
    static_assert(sizeof(EvolvePlan) < 3800, "kernel argument too large");
    if (plan.total_neurons == 0) return 0;
    hipLaunchKernelGGL(k_evolve, dim3(cdiv(plan.total_neurons, kEvolveNeurons)), dim3(kFinParts * 64), 0, s, plan);
    return (int)hipGetLastError();
}

static bool gemv_vector_ok(const GemvPlan& plan) {
    for (uint32_t c = 0; c < plan.n_conn; ++c) {
        if (plan.c[c].ld % 4 != 0) return false;
        if (((uintptr_t)plan.c[c].W) % 16 != 0) return false;
    }
    return true;
}

int launch_evolve_plast(const EvolvePlan& ev, const PlastStepPlan* plast, bn_stream_t s) {
    static_assert(sizeof(EvolvePlan) + sizeof(PlastStepPlan) + 4 < 3800, "kernel argument too large");
    const uint32_t ne = cdiv(ev.total_neurons, kEvolveNeurons);
    const uint32_t np = plast ? plast->row_blocks + plast->spiking_blocks : 0;
    if (np == 0)
        hipLaunchKernelGGL(k_evolve, dim3(ne), dim3(kFinParts * 64), 0, s, ev);
    else
        hipLaunchKernelGGL(k_evolve_plast, dim3(ne + np), dim3(kFinParts * 64), 0, s, ev, *plast, ne);
    return (int)hipGetLastError();
}

int launch_gemv(const GemvPlan& plan, bn_stream_t s) {
    if (plan.total_tiles == 0 && !plan.pc.list) return 0;
    if (plan.parts) {
        // One part tile per block, and the compaction block last when the plan has one.
        const uint32_t blocks = plan.total_tiles + (plan.pc.list ? 1u : 0u);
        if (plan.nontemporal)
            hipLaunchKernelGGL(k_gemv_parts<true>, dim3(blocks), dim3(kFinParts * 64), 0, s, plan);
        else
            hipLaunchKernelGGL(k_gemv_parts<false>, dim3(blocks), dim3(kFinParts * 64), 0, s, plan);
        return (int)hipGetLastError();
    }
    // Four tiles per block, and the compaction block last when the plan has one.
    const uint32_t blocks = cdiv(plan.total_tiles, 4) + (plan.pc.list ? 1u : 0u);
    if (gemv_vector_ok(plan) && plan.nontemporal)
        hipLaunchKernelGGL(k_gemv<true>, dim3(blocks), dim3(256), 0, s, plan);
    else if (gemv_vector_ok(plan))
        hipLaunchKernelGGL(k_gemv<false>, dim3(blocks), dim3(256), 0, s, plan);
    else
        hipLaunchKernelGGL(k_gemv_scalar, dim3(blocks), dim3(256), 0, s, plan);
    return (int)hipGetLastError();
}

// Registers per thread for a row of n columns held by 256 threads (n <= 64 * 256).
static int row_regs(uint32_t n) { return n <= 16u * kNormThreads ? 16 : n <= 32u * kNormThreads ? 32 : 64; }
static uint32_t max_post(const PlastPlan& p) {
    uint32_t m = 0;
    for (uint32_t k = 0; k < p.n; ++k) m = std::max(m, p.p[k].n_post);
    return m;
}

int launch_multi_propagate(const MultiStep& m, bn_stream_t s) {
    if (!m.ph_gemv.blocks) return 0;
    if (m.nontemporal)
        hipLaunchKernelGGL(k_gemv_multi<true>, dim3(m.ph_gemv.blocks), dim3(256), 0, s, m.gemv, m.ph_gemv.block_net,
                           m.ph_gemv.net_first_block);
    else
        hipLaunchKernelGGL(k_gemv_multi<false>, dim3(m.ph_gemv.blocks), dim3(256), 0, s, m.gemv, m.ph_gemv.block_net,
                           m.ph_gemv.net_first_block);
    return (int)hipGetLastError();
}

int launch_multi_step(const MultiStep& m, bn_stream_t s) {
    if (m.ph_evolve.blocks)
        hipLaunchKernelGGL(k_evolve_multi, dim3(m.ph_evolve.blocks), dim3(kFinParts * 64), 0, s, m.evolve,
                           m.ph_evolve.block_net, m.ph_evolve.net_first_block);
    if (m.ph_gemv.blocks) {
        if (m.nontemporal)
            hipLaunchKernelGGL(k_gemv_multi<true>, dim3(m.ph_gemv.blocks), dim3(256), 0, s, m.gemv,
                               m.ph_gemv.block_net, m.ph_gemv.net_first_block);
        else
            hipLaunchKernelGGL(k_gemv_multi<false>, dim3(m.ph_gemv.blocks), dim3(256), 0, s, m.gemv,
                               m.ph_gemv.block_net, m.ph_gemv.net_first_block);
    }
    if (m.ph_plast.blocks)
        hipLaunchKernelGGL(k_plasticity_multi, dim3(m.ph_plast.blocks), dim3(256), 0, s, m.plast,
                           m.ph_plast.block_net, m.ph_plast.net_first_block);
    return (int)hipGetLastError();
}

int launch_multi_overlapped(const MultiStep& m, uint32_t n, bn_stream_t s) {
    if (!m.ov.evolve_even || n % 2) return (int)hipErrorInvalidValue;
    for (uint32_t k = 0; k < n; ++k) {
        const EvolvePlan* ev = k % 2 == 0 ? m.ov.evolve_even : m.ov.evolve_odd;
        const PlastStepPlan* pl = k == 0 ? nullptr : (k - 1) % 2 == 0 ? m.ov.plast_staging : m.ov.plast_current;
        const uint32_t blocks = m.ph_evolve.blocks + (pl ? m.ph_plast.blocks : 0);
        if (blocks)
            hipLaunchKernelGGL(k_evolve_plast_multi, dim3(blocks), dim3(kFinParts * 64), 0, s, ev, m.ph_evolve, pl,
                               m.ph_plast, k, k > 0 ? 1u : 0u);
        if (m.ph_gemv.blocks) {
            if (m.nontemporal)
                hipLaunchKernelGGL(k_gemv_multi<true>, dim3(m.ph_gemv.blocks), dim3(256), 0, s, m.gemv,
                                   m.ph_gemv.block_net, m.ph_gemv.net_first_block);
            else
                hipLaunchKernelGGL(k_gemv_multi<false>, dim3(m.ph_gemv.blocks), dim3(256), 0, s, m.gemv,
                                   m.ph_gemv.block_net, m.ph_gemv.net_first_block);
        }
    }
    if (m.ph_plast.blocks)
        hipLaunchKernelGGL(k_plasticity_multi, dim3(m.ph_plast.blocks), dim3(256), 0, s, m.ov.plast_final,
                           m.ph_plast.block_net, m.ph_plast.net_first_block);
    return (int)hipGetLastError();
}

int launch_plasticity(const PlastStepPlan& plan, bn_stream_t s) {
    static_assert(sizeof(PlastStepPlan) < 3800, "kernel argument too large");
    const uint32_t blocks = plan.row_blocks + plan.spiking_blocks;
    if (blocks == 0) return 0;
    hipLaunchKernelGGL(k_plasticity, dim3(blocks), dim3(256), 0, s, plan);
    return (int)hipGetLastError();
}

static uint32_t plast_grid(uint32_t items) { return items < 2048 ? (items ? items : 1) : 2048; }

int launch_ltp(const PlastPlan& plan, bn_stream_t s) {
    if (plan.n == 0 || plan.items_post == 0) return 0;
    hipLaunchKernelGGL(k_ltp, dim3(plast_grid(plan.items_post)), dim3(256), 0, s, plan);
    return (int)hipGetLastError();
}

int launch_ltd(const PlastPlan& plan, bn_stream_t s) {
    if (plan.n == 0 || plan.items_pre == 0) return 0;
    hipLaunchKernelGGL(k_ltd, dim3(plast_grid(plan.items_pre)), dim3(256), 0, s, plan);
    return (int)hipGetLastError();
}

int launch_norm_tail(const TailPlan& plan, bn_stream_t s) {
    static_assert(sizeof(TailPlan) < 3800, "kernel argument too large");
    const uint32_t blocks =
        plan.norm_rows + (plan.increments ? cdiv(plan.total_neurons, 256) : 0) + (plan.tail.active ? 1 : 0);
    if (blocks == 0) return 0;
    switch (row_regs(max_post(plan.pl))) {
        case 16: hipLaunchKernelGGL(k_norm_tail<16>, dim3(blocks), dim3(256), 0, s, plan); break;
        case 32: hipLaunchKernelGGL(k_norm_tail<32>, dim3(blocks), dim3(256), 0, s, plan); break;
        default: hipLaunchKernelGGL(k_norm_tail<64>, dim3(blocks), dim3(256), 0, s, plan); break;
    }
    return (int)hipGetLastError();
}

int launch_lazy_init(const PlastDev& d, bn_stream_t s) {
    if (d.n_pre == 0) return 0;
    switch (row_regs(d.n_post)) {
        case 16: hipLaunchKernelGGL(k_lazy_init<16>, dim3(d.n_pre), dim3(256), 0, s, d); break;
        case 32: hipLaunchKernelGGL(k_lazy_init<32>, dim3(d.n_pre), dim3(256), 0, s, d); break;
        default: hipLaunchKernelGGL(k_lazy_init<64>, dim3(d.n_pre), dim3(256), 0, s, d); break;
    }
    return (int)hipGetLastError();
}

int launch_materialize_g(float* g_dst, const float* g_base, const FinalizeSource* src,
                         uint32_t n_src, uint32_t n, bn_stream_t s) {
    if (n == 0) return 0;
    FinalizeSource s0{}, s1{};
    if (n_src > 0) s0 = src[0];
    if (n_src > 1) s1 = src[1];
    hipLaunchKernelGGL(k_materialize_g, dim3(cdiv(n, 256)), dim3(256), 0, s, g_dst, g_base, s0, s1,
                       n_src, n);
    return (int)hipGetLastError();
}

int launch_add_increments(float* dst, const float* src, const uint32_t* spikes, uint32_t n, bn_stream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_add_increments, dim3(cdiv(n, 256)), dim3(256), 0, s, dst, src, spikes, n);
    return (int)hipGetLastError();
}

int launch_fill_u32(uint32_t* p, uint32_t value, size_t n, bn_stream_t s) {
    if (n == 0) return 0;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, value, n);
    return (int)hipGetLastError();
}

// Single-plasticity plan from the reference's view/params pair.
static PlastPlan plan_from_view(const PlasticityView& v, const PlasticityParams& p, uint32_t kind,
                                bool normalize) {
    PlastPlan P{};
    PlastDev& D = P.p[0];
    D.kind = kind;
    D.normalize = normalize ? 1u : 0u;
    D.p = p;
    D.W = v.weight_matrix;
    D.ld = v.weight_matrix_ld ? v.weight_matrix_ld : v.post_group_num_neurons;
    D.n_pre = v.pre_group_num_neurons;
    D.n_post = v.post_group_num_neurons;
    D.pre_offset = v.pre_group_offset;
    D.post_offset = v.post_group_offset;
    D.pre_spikes = v.pre_spikes;
    D.post_spikes = v.post_spikes;
    D.pre = v.pre_traces;
    D.post = v.post_traces;
    D.trip_pre = v.triplet_pre_traces;
    D.trip_post = v.triplet_post_traces;
    D.target = v.pre_weight_sums;
    P.n = 1;
    P.items_post = D.n_post;
    P.items_pre = D.n_pre;
    return P;
}

// decay -> LTP -> LTD -> (normalise) -> increments, as the reference launchers
// (kernel_triplet_plasticity_update.cu:181-216, kernel_istdp_plasticity_update.cu:103-135).
static int plasticity_update(const PlasticityView& v, const PlasticityParams& p, uint32_t kind,
                             bool normalize, bn_stream_t s, const char* what) {
    if (normalize && v.post_group_num_neurons > (uint32_t)(kNormRegs * kNormThreads))
        return set_error(BN_ERR_INVALID_ARGUMENT,
                         std::string(what) + ": post group larger than the normalisation kernel supports");
    PlastPlan P = plan_from_view(v, p, kind, normalize);
    const uint32_t nmax = std::max(P.p[0].n_pre, P.p[0].n_post);
    if (nmax == 0) return BN_OK;
    hipLaunchKernelGGL(k_plast_decay, dim3(cdiv(nmax, 256)), dim3(256), 0, s, P);
    BN_HIP_TRY(hipGetLastError(), what);
    BN_HIP_TRY((hipError_t)launch_ltp(P, s), what);
    BN_HIP_TRY((hipError_t)launch_ltd(P, s), what);
    if (normalize) {
        TailPlan T{};
        T.pl = P;
        T.norm_rows = P.p[0].n_pre;
        BN_HIP_TRY((hipError_t)launch_norm_tail(T, s), what);
    }
    const uint32_t hi = std::max(P.p[0].pre_offset + P.p[0].n_pre, P.p[0].post_offset + P.p[0].n_post);
    hipLaunchKernelGGL(k_plast_increment, dim3(cdiv(hi, 256)), dim3(256), 0, s, P, hi);
    BN_HIP_TRY(hipGetLastError(), what);
    return BN_OK;
}

}  // namespace bn

// =======================================================================================
// extern "C" launcher layer
// =======================================================================================
using namespace bn;

extern "C" int bn_neuron_group_evolve(NeuronGroupView v, bn_stream_t s) {
    if (v.num_neurons == 0) return BN_OK;
    if (!v.rng_clock) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_neuron_group_evolve: rng_clock is null");
    hipLaunchKernelGGL(k_group_evolve, dim3(cdiv(v.num_neurons, 256)), dim3(256), 0, s, v);
    BN_HIP_TRY(hipGetLastError(), "kernel_neuron_group_evolve launch");
    return BN_OK;
}

// kernel_connection_propagate.cu:8-94 for one connection: traces, the FP64 chunk partials of the
// GEMV into `partials` ([ceil(pre / row_block)][ld] doubles), then g += the partials in order.
static int connection_propagate(ConnectionView v, double* partials, uint32_t row_block, bn_stream_t s) {
    hipLaunchKernelGGL(k_connection_traces, dim3(cdiv(v.pre_group_num_neurons, 256)), dim3(256), 0, s, v);
    BN_HIP_TRY(hipGetLastError(), "kernel_connection_traces launch");
    GemvPlan P{};
    GemvConn& C = P.c[0];
    C.W = v.weight_matrix;
    C.x = v.summed_postsynaptic_conductances;
    C.row_scale = nullptr;
    C.partials = partials;
    C.ld = v.weight_matrix_ld ? v.weight_matrix_ld : v.post_group_num_neurons;
    C.n_pre = v.pre_group_num_neurons;
    C.n_post = v.post_group_num_neurons;
    C.row_block = row_block;
    C.n_chunks = cdiv(C.n_pre, C.row_block);
    C.n_coltiles = cdiv(C.n_post, kGemvTileCols);
    C.tile_begin = 0;
    C.tile_end = C.n_chunks * C.n_coltiles;
    P.n_conn = 1;
    P.total_tiles = C.tile_end;
    BN_HIP_TRY((hipError_t)launch_gemv(P, s), "gemv connection accumulate");
    float* target = v.pre_group_is_inhibtory ? v.post_group_g_inhibitory : v.post_group_g_excitatory;
    FinalizeSource src{partials, C.n_chunks, C.ld};
    BN_HIP_TRY((hipError_t)launch_materialize_g(target, target, &src, 1, C.n_post, s), "gemv finalize");
    return BN_OK;
}

extern "C" int bn_connection_propagate(ConnectionView v, bn_stream_t s) {
    if (v.pre_group_num_neurons == 0 || v.post_group_num_neurons == 0) return BN_OK;
    if (!v.gemv_partials || v.gemv_row_block == 0)
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_connection_propagate: gemv_partials/gemv_row_block not set");
    return connection_propagate(v, v.gemv_partials, v.gemv_row_block, s);
}

// ---------------------------------------------------------------------------------------
// Allocation-free propagate: the FP64 GEMV partials live in `scratch`, a file-scope device array
// of kNoAllocScratchDoubles doubles (8 MiB per GPU, held until the process ends), instead of the
// view's gemv_partials.  Host code reaches the array only through the address that
// hipGetSymbolAddress returns on each device (the array's name is a host-side stand-in there);
// it is looked up once per device, with that device current, on the entry point's first call
// there.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two
// streams, two threads) race on it.  The kernels write every partial they later read, so
// nothing needs clearing between calls.
// ---------------------------------------------------------------------------------------
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[kNoAllocMaxDevices] = {};

static int scratch_on_current_device(double** out) {
    int dev = -1;
    BN_HIP_TRY(hipGetDevice(&dev), "bn_connection_propagate_no_alloc: hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return set_error(BN_ERR_RUNTIME, "bn_connection_propagate_no_alloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        double* p = nullptr;
        BN_HIP_TRY(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)),
                   "bn_connection_propagate_no_alloc: hipGetSymbolAddress(scratch)");
        if (!p) return set_error(BN_ERR_RUNTIME, "bn_connection_propagate_no_alloc: null scratch address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return BN_OK;
}

extern "C" int bn_connection_propagate_no_alloc(ConnectionView v, bn_stream_t s) {
    if (v.pre_group_num_neurons == 0 || v.post_group_num_neurons == 0) return BN_OK;
    if (v.gemv_row_block == 0)
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_connection_propagate_no_alloc: gemv_row_block not set");
    const uint64_t ld = v.weight_matrix_ld ? v.weight_matrix_ld : v.post_group_num_neurons;
    const uint64_t need = (uint64_t)cdiv(v.pre_group_num_neurons, v.gemv_row_block) * ld;
    if (need > kNoAllocScratchDoubles)
        return set_error(BN_ERR_INVALID_ARGUMENT,
                         "bn_connection_propagate_no_alloc: ceil(pre / gemv_row_block) * row stride = " +
                             std::to_string(need) + " partial sums exceed the scratch of " +
                             std::to_string(kNoAllocScratchDoubles) + " (use a larger gemv_row_block)");
    double* partials = nullptr;
    const int rc = scratch_on_current_device(&partials);
    if (rc != BN_OK) return rc;
    return connection_propagate(v, partials, v.gemv_row_block, s);
}

extern "C" int bn_stimulus_apply(StimulusView v, uint32_t next_event_idx, uint32_t n, bn_stream_t s) {
    if (n == 0) return BN_OK;
    hipLaunchKernelGGL(k_stimulus_apply, dim3(1), dim3(64), 0, s, v, next_event_idx, n);
    BN_HIP_TRY(hipGetLastError(), "kernel_stimulus_apply launch");
    return BN_OK;
}

extern "C" int bn_stimulus_apply_device_cursor(StimulusView v, bn_device_clock* clock, bn_stream_t s) {
    if (!clock) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_stimulus_apply_device_cursor: clock is null");
    hipLaunchKernelGGL(k_stimulus_device_cursor, dim3(1), dim3(64), 0, s, v, clock);
    BN_HIP_TRY(hipGetLastError(), "kernel_stimulus_apply launch");
    return BN_OK;
}

extern "C" int bn_triplet_plasticity_update(PlasticityView v, PlasticityParams p, bn_stream_t s) {
    return plasticity_update(v, p, kPlastTriplet, false, s, "kernel_triplet_plasticity_update");
}
extern "C" int bn_triplet_plasticity_update_normalize(PlasticityView v, PlasticityParams p, bn_stream_t s) {
    return plasticity_update(v, p, kPlastTriplet, true, s, "kernel_triplet_plasticity_update");
}
extern "C" int bn_istdp_plasticity_update(PlasticityView v, PlasticityParams p, bn_stream_t s) {
    return plasticity_update(v, p, kPlastIstdp, false, s, "kernel_istdp_plasticity_update");
}
extern "C" int bn_stdp_plasticity_update(PlasticityView v, PlasticityParams p, bn_stream_t s) {
    return plasticity_update(v, p, kPlastStdp, false, s, "kernel_stdp_plasticity_update");
}

extern "C" int bn_historicize_spikes(const uint32_t* d_spikes, uint32_t* hist, uint32_t n,
                                     uint32_t interval, uint32_t step, bn_stream_t s) {
    if (n == 0) return BN_OK;
    if (interval == 0) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_historicize_spikes: interval is 0");
    hipLaunchKernelGGL(k_historicize<uint32_t>, dim3(cdiv(n, 256)), dim3(256), 0, s, d_spikes, hist, n,
                       interval, step);
    BN_HIP_TRY(hipGetLastError(), "kernel_historicize_spikes launch");
    return BN_OK;
}

extern "C" int bn_historicize_membrane_potentials(const float* d_v, float* hist, uint32_t n,
                                                  uint32_t interval, uint32_t step, bn_stream_t s) {
    if (n == 0) return BN_OK;
    if (interval == 0)
        return set_error(BN_ERR_INVALID_ARGUMENT, "bn_historicize_membrane_potentials: interval is 0");
    hipLaunchKernelGGL(k_historicize<float>, dim3(cdiv(n, 256)), dim3(256), 0, s, d_v, hist, n, interval,
                       step);
    BN_HIP_TRY(hipGetLastError(), "kernel_historicize_membrane_potentials launch");
    return BN_OK;
}

extern "C" int bn_clock_init(bn_device_clock* clock, uint64_t rng_seed, bn_stream_t s) {
    if (!clock) return set_error(BN_ERR_INVALID_ARGUMENT, "bn_clock_init: clock is null");
    hipLaunchKernelGGL(k_clock_init, dim3(1), dim3(64), 0, s, clock, (uint32_t)rng_seed, (uint32_t)(rng_seed >> 32));
    BN_HIP_TRY(hipGetLastError(), "kernel_init_neuron_rng_states launch");
    return BN_OK;
}

extern "C" int bn_clock_advance(bn_device_clock* clock, bn_stream_t s) {
    hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, s, clock);
    BN_HIP_TRY(hipGetLastError(), "clock advance launch");
    return BN_OK;
}
