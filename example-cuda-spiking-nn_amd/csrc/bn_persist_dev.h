// Device helpers shared by the two persistent-step kernels (kernels_persist.hip: weights
// streamed each step; kernels_persist_ll.hip: weights resident in registers).
#pragma once

#include <hip/hip_runtime.h>

#include "bn_device.h"
#include "bn_persist.h"

namespace bn {

// Global address space: accesses through these pointers are global_ rather than flat_
// instructions (a flat access also counts against the LDS/scalar counter and makes the
// compiler wait for LDS traffic as well).
#define BN_GAS __attribute__((address_space(1)))
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ const BN_GAS T* gp(const T* p) { return (const BN_GAS T*)p; }
template <typename T>
__device__ __forceinline__ BN_GAS T* gp(T* p) { return (BN_GAS T*)p; }

// Values that one block writes and another block reads within the launch are moved with
// agent-coherent loads and stores (relaxed atomics at agent scope: they do not stop in the L2
// of one XCD), so no fence has to write back or invalidate a whole L2.  A block publishes them
// by completing every store before it raises a counter (p_block_arrive), and a block that has
// seen the counter reads them with the same kind of load.  The weight entries of an item are
// read and written only by the thread that owns them (the block owning an item keeps it for
// the whole launch), so they take ordinary loads and stores.
template <typename T>
__device__ __forceinline__ T ld_c(const T* p) {
    return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_c(T* p, T v) {
    __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add_c(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void or_c(uint32_t* p, uint32_t v) {
    (void)__hip_atomic_fetch_or(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void max_c(uint32_t* p, uint32_t v) {
    (void)__hip_atomic_fetch_max(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Entry update in true units (lazy: w = s*m, stored back as w'/s), as apply_delta.
__device__ __forceinline__ float p_apply(bool lazy, float m, float si, float delta, float lo, float hi) {
    if (lazy) return clamp_weight(si * m + delta, lo, hi) / si;
    return clamp_weight(m + delta, lo, hi);
}

// The plasticity of the step on one row i of an item (lane: columns j0 .. j0+3): per entry the
// LTP of a spiking post column then the LTD of a spiking pre row (index self-skip), in true
// units.  m4 becomes the row's new entries and the mask of touched entries is returned; for the
// lazy matrix lsum is the row's tile sum (pre spiked: of the new entries; else of the LTP
// changes), summed over the wave's lanes.
// The rule's constants used per entry, read once per item rather than per row.
struct PRule {
    float ltp_A, ltd_A, ltd_B, weight_min, weight_max;
    uint32_t kind;
};
__device__ __forceinline__ PRule p_rule(const PPlast& D) {
    return PRule{D.p.ltp_A, D.p.ltd_A, D.p.ltd_B, D.p.weight_min, D.p.weight_max, D.kind};
}

__device__ __forceinline__ uint32_t p_plast_row(const PRule& D, bool lazy, bool triplet, bool sp, uint32_t i, float si,
                                                float pre_i, float fac_i, uint32_t j0, bool active, uint32_t n_post,
                                                uint32_t colspike, const float (&post_j)[4], const float (&fac_ltp)[4],
                                                v4f& m4, double& lsum) {
    const bool update = !lazy || si != 0.0f;
    uint32_t touched = 0;
    double rs = 0.0;
#pragma unroll
    for (uint32_t c4 = 0; c4 < 4; ++c4) {
        if (!active || j0 + c4 >= n_post) continue;
        const uint32_t j = j0 + c4;
        const float m = m4[c4];
        float val = m;
        bool t = false;
        if (update && j != i) {
            if ((colspike >> c4) & 1u) {
                const float delta = triplet ? pre_i * fac_ltp[c4] : D.ltp_A * pre_i;
                val = p_apply(lazy, val, si, delta, D.weight_min, D.weight_max);
                t = true;
            }
            if (sp) {
                float d;
                if (triplet) d = -1.0f * post_j[c4] * fac_i;
                else if (D.kind == kPlastIstdp) d = -1.0f * D.ltd_A * (post_j[c4] - D.ltd_B);
                else d = -1.0f * D.ltd_A * post_j[c4];
                val = p_apply(lazy, val, si, d, D.weight_min, D.weight_max);
                t = true;
            }
        }
        if (lazy) rs += sp ? (double)val : (t ? (double)val - (double)m : 0.0);
        if (t) {
            m4[c4] = val;
            touched |= 1u << c4;
        }
    }
    if (lazy) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) rs += __shfl_xor(rs, off, 64);
    }
    lsum = rs;
    return touched;
}

// Stores of the touched entries of a row.
__device__ __forceinline__ void p_store_row(BN_GAS float* dst, const v4f& nv, uint32_t touched) {
    if (touched == 0xFu) {
        *reinterpret_cast<BN_GAS v4f*>(dst) = nv;
    } else if (touched) {
#pragma unroll
        for (uint32_t c4 = 0; c4 < 4; ++c4)
            if ((touched >> c4) & 1u) dst[c4] = nv[c4];
    }
}

// Where the rows of an item sit for the thread (wave: chunk of rows, lane: 4 columns).
struct PRows {
    uint32_t s0, seg_rows, c0, c1, j0;
    bool active;
};

__device__ __forceinline__ PRows p_rows(const PConn& C, uint32_t col, uint32_t seg) {
    PRows r;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    r.s0 = seg * kPWaves * C.rb;
    r.seg_rows = min(kPWaves * C.rb, C.n_pre - r.s0);
    r.c0 = wave * C.rb;
    r.c1 = min(r.c0 + C.rb, r.seg_rows);
    r.j0 = col * kPCols + 4u * (threadIdx.x & 63u);
    r.active = r.j0 < C.n_post;
    return r;
}

}  // namespace bn
