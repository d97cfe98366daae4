// sg_device.hpp -- wave64 / block primitives for gfx950.
//
// Cross-block prefix sums use reduce-then-scan: a count kernel writes one
// sum per block and the consuming kernel folds the sums of its predecessor
// blocks itself.  No block ever waits on another block: on MI355X the eight
// XCDs have private L2s, so a polled flag is a round trip to the memory-side
// cache and same-address atomics serialize (~11 ns each), which makes
// decoupled look-back chains and atomic tile tickets the slowest part of a
// pass.
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

template <int NW>
SG_D uint32_t block_sum(uint32_t v, uint32_t* lds) {
    const uint32_t s = wave_sum(v);
    if (lane_id() == 0) lds[wave_id()] = s;
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) tot += lds[i];
    __syncthreads();
    return tot;
}

// sum of sums[0..b) computed by the whole block (the "scan" half of
// reduce-then-scan, done redundantly by every consumer block; the sums of a
// pass are a few KB and stay in L2)
template <int NW>
SG_D uint32_t block_prefix_of(const uint32_t* __restrict__ sums, uint32_t b, uint32_t* lds) {
    uint32_t acc = 0;
    for (uint32_t j = threadIdx.x; j < b; j += NW * 64) acc += sums[j];
    return block_sum<NW>(acc, lds);
}

// lanes of the wave with ok set and the same low db bits of d (a wave64
// "match"): one ballot per bit
SG_D uint64_t digit_peers(bool ok, uint32_t d, int db) {
    uint64_t peers = __ballot(ok);
    for (int b = 0; b < db; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// ------------------------------------------- correctly rounded 1/x and a/b
// Both return exactly what IEEE division returns, in fewer instructions
// than the compiler's general division sequence (div_scale, rcp, five
// fused multiply-adds, div_fmas, div_fixup).
//
// rcp_rn: v_rcp_f32 (within 1 ulp) refined by one Newton step whose
// residual 1 - x*r is exact in a fused multiply-add; the step then rounds
// to the correctly rounded reciprocal (P. Markstein, "Computation of
// elementary functions on the IBM RISC System/6000 processor", IBM J. Res.
// Dev. 34(1), 1990).  Checked against 1.f/x on gfx950 for all 2^32 inputs:
// equal for every |x| in [2^-126, 2^126); the rest (zero, subnormal, huge,
// inf, NaN) takes the IEEE division.
SG_D bool rcp_fast_range(float x) { return fabsf(x) >= 0x1p-126f && fabsf(x) < 0x1p126f; }

SG_D float rcp_rn(float x) {
    if (rcp_fast_range(x)) {
        const float r = __builtin_amdgcn_rcpf(x);
        return fmaf(fmaf(-x, r, 1.f), r, r);
    }
    return 1.f / x;
}

// div_rn: a/b given y = 1/b correctly rounded: q = a*y is within one ulp
// of a/b, the remainder a - b*q is exact in a fused multiply-add, and
// q + remainder*y rounded once is a/b correctly rounded (Markstein 1990).
// Valid while nothing under- or overflows: b and a nonzero have binary
// exponents in [-60, 60] (checked on gfx950 against a/b for every b over
// 1.3e10 pairs); other operands take the IEEE division.  For a = +-0 the
// quotient is q = a*y itself (the correction step would turn -0 into +0).
SG_D bool div_fast_range(float x) { return fabsf(x) >= 0x1p-60f && fabsf(x) < 0x1p61f; }

SG_D float div_rn(float a, float b, float y, bool b_ok) {
    if (b_ok && (a == 0.f || div_fast_range(a))) {
        const float q = a * y;
        const float f = fmaf(fmaf(-b, q, a), y, q);
        return a == 0.f ? q : f;
    }
    return a / b;
}

}  // namespace sg
