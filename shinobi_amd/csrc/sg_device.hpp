// sg_device.hpp -- wave64 / block primitives for gfx950: prefix sums,
// ballot ranking and the decoupled look-back used by the single-pass
// count+scan+emit kernels and by the radix sort.
#pragma once

#include "sg_common.hpp"

namespace sg {

SG_D int lane_id() { return (int)__lane_id(); }
SG_D int wave_id() { return (int)(threadIdx.x >> 6); }
SG_D uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// inclusive prefix sum across the 64 lanes of a wave
SG_D uint32_t wave_incl_scan(uint32_t v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (l >= d) v += t;
    }
    return v;
}

SG_D uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// exclusive prefix sum over a block of NW waves; lds must hold NW words.
// Returns the exclusive prefix of this thread, *total = block sum.
template <int NW>
SG_D uint32_t block_excl_scan(uint32_t v, uint32_t* total, uint32_t* lds) {
    uint32_t inc = wave_incl_scan(v);
    const int w = wave_id();
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        uint32_t x = lds[i];
        pre += (i < w) ? x : 0u;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// ------------------------------------------------------------ look-back
// Status word per tile: high 32 bits = flag, low 32 bits = value.
constexpr uint64_t LB_AGG = 1ull << 32;
constexpr uint64_t LB_INC = 2ull << 32;
constexpr uint32_t kSpinLimit = 1u << 24;

SG_D void lb_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
SG_D uint64_t lb_load(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by every lane of ONE wave of tile `tile` (> 0) after it published
// its aggregate.  Returns the exclusive prefix of the tile (all lanes).
// Waits only on tiles with smaller ids, which were handed out earlier by the
// tile counter and therefore are resident or done: every wait ends.  A wait
// that exceeds kSpinLimit polls sets ST_SPIN and counts the tile as zero.
SG_D uint32_t lb_prefix(const uint64_t* st, uint32_t tile, uint32_t* status) {
    uint32_t prefix = 0;
    int64_t base = (int64_t)tile - 1;
    const int l = lane_id();
    while (base >= 0) {
        int64_t j = base - l;
        uint64_t s = LB_INC;  // lanes past the start count as inclusive 0
        if (j >= 0) {
            uint32_t spins = 0;
            s = lb_load(st + j);
            while ((s >> 32) == 0) {
                if (++spins > kSpinLimit) {
                    atomicOr(status, ST_SPIN);
                    s = LB_INC;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                s = lb_load(st + j);
            }
        }
        uint64_t incm = __ballot((s >> 32) == 2);
        uint32_t v = (uint32_t)s;
        if (incm) {
            int k = __ffsll((long long)incm) - 1;   // nearest inclusive predecessor
            prefix += wave_sum(l <= k ? v : 0u);
            break;
        }
        prefix += wave_sum(v);
        base -= 64;
    }
    return prefix;
}

// Block-level wrapper: publish aggregate, look back with wave 0, publish the
// inclusive value, broadcast the exclusive prefix through LDS.
SG_D uint32_t lb_tile_prefix(uint64_t* st, uint32_t tile, uint32_t agg, uint32_t* status,
                             uint32_t* lds_word) {
    if (threadIdx.x < 64) {
        uint32_t pre = 0;
        if (tile == 0) {
            if (threadIdx.x == 0) lb_store(st, LB_INC | (uint64_t)agg);
        } else {
            if (threadIdx.x == 0) lb_store(st + tile, LB_AGG | (uint64_t)agg);
            pre = lb_prefix(st, tile, status);
            if (threadIdx.x == 0) lb_store(st + tile, LB_INC | (uint64_t)(uint32_t)(pre + agg));
        }
        if (threadIdx.x == 0) *lds_word = pre;
    }
    __syncthreads();
    return *lds_word;
}

// dynamic tile id: tiles are handed out in launch order of their blocks'
// first instruction, which makes look-back waits safe.
SG_D uint32_t next_tile(uint32_t* ctr, uint32_t* lds_word) {
    if (threadIdx.x == 0) *lds_word = atomicAdd(ctr, 1u);
    __syncthreads();
    uint32_t t = *lds_word;
    __syncthreads();
    return t;
}

}  // namespace sg
