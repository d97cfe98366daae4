// sg_sort.hip -- stable LSD radix sort of (key, value) pairs and a
// single-pass prefix scan, written for gfx950 wave64.
//
// Sort (semantics of chag::pp::sort as driven by src/Utils/Sort.cu:28-56):
// one histogram pass over the keys builds the digit histograms of every
// pass; each pass is then ONE kernel (Onesweep-style): a 4096-key tile is
// ranked in registers with wave64 ballot matching per digit, per-digit tile
// counts are published to a decoupled look-back array, and keys and values
// scatter straight to their final slots.  Ranking walks items in tile order
// and tiles are handed out in order, so the sort is stable.
//
// Scan (src/Utils/Scan.h:35-68): 2048 values per tile, blocked per thread,
// wave64 shuffle scan + decoupled look-back.
#include "sg_device.hpp"
#include "sg_kernels.hpp"

namespace sg {

constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;   // 4096 keys
constexpr uint32_t SS_AGG = 1u << 30, SS_INC = 2u << 30, SS_VAL = (1u << 30) - 1u;

int sort_passes(int key_bits, int radix_bits) {
    if (key_bits < 1) key_bits = 1;
    return (key_bits + radix_bits - 1) / radix_bits;
}

size_t sort_scratch_words(uint32_t cap, int key_bits, int radix_bits) {
    const size_t R = (size_t)1 << radix_bits;
    const size_t passes = (size_t)sort_passes(key_bits, radix_bits);
    const size_t tiles = ((size_t)cap + kSortTile - 1) / kSortTile;
    size_t w = passes * R + passes * tiles * R;
    return (w + 3) & ~(size_t)3;   // 16-byte multiple
}

template <int RB>
__global__ __launch_bounds__(kBlock) void k_sort_hist(const uint32_t* __restrict__ keys, const uint32_t* d_n,
                                                      uint32_t cap, int passes, uint32_t* hist) {
    constexpr int R = 1 << RB;
    __shared__ uint32_t s[8 * R];
    for (int i = threadIdx.x; i < passes * R; i += kBlock) s[i] = 0;
    __syncthreads();
    const uint32_t n = min(*d_n, cap);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&s[p * R + ((k >> (p * RB)) & (R - 1))], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * R; i += kBlock) {
        const uint32_t v = s[i];
        if (v) atomicAdd(&hist[i], v);
    }
}

SG_D uint32_t ss_load(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
SG_D void ss_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int RB>
__global__ __launch_bounds__(kBlock) void k_sort_pass(const uint32_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      const uint32_t* d_n, uint32_t cap,
                                                      const uint32_t* __restrict__ hist, uint32_t* sstat,
                                                      int shift, uint32_t* tile_ctr, uint32_t* status) {
    constexpr int R = 1 << RB;
    __shared__ uint32_t s_cnt[4][R];
    __shared__ uint32_t s_goff[R];
    __shared__ uint32_t s_scan[4];
    __shared__ uint32_t s_word;
    const uint32_t n = min(*d_n, cap);
    const uint32_t tile = next_tile(tile_ctr, &s_word);
    const uint32_t base = tile * kSortTile;
    if (base >= n) return;

    // digit start offsets of this pass
    const uint32_t hv = threadIdx.x < R ? hist[threadIdx.x] : 0u;
    uint32_t htot;
    const uint32_t dstart = block_excl_scan<kBlock / 64>(hv, &htot, s_scan);
    for (int i = threadIdx.x; i < 4 * R; i += kBlock) (&s_cnt[0][0])[i] = 0u;
    __syncthreads();

    const int w = wave_id(), l = lane_id();
    const uint32_t wbase = base + (uint32_t)w * (kSortItems * 64);
    uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t pos = wbase + k * 64 + l;
        const bool ok = pos < n;
        key[k] = ok ? kin[pos] : 0u;
        val[k] = ok ? vin[pos] : 0u;
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t pos = wbase + k * 64 + l;
        const bool ok = pos < n;
        const uint32_t d = (key[k] >> shift) & (R - 1);
        uint64_t m = __ballot(ok);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)__popcll(m & lt);
        const uint32_t cnt = (uint32_t)__popcll(m);
        const uint32_t cur = ok ? s_cnt[w][d] : 0u;
        rank[k] = cur + below;
        if (ok && l == __ffsll((long long)m) - 1) s_cnt[w][d] = cur + cnt;
    }
    __syncthreads();

    if (threadIdx.x < R) {
        const int d = threadIdx.x;
        const uint32_t c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
        s_cnt[0][d] = 0u;
        s_cnt[1][d] = c0;
        s_cnt[2][d] = c0 + c1;
        s_cnt[3][d] = c0 + c1 + c2;
        const uint32_t total = c0 + c1 + c2 + c3;
        uint32_t* st = sstat + (size_t)tile * R + d;
        uint32_t prefix = 0;
        if (tile == 0) {
            ss_store(st, SS_INC | total);
        } else {
            ss_store(st, SS_AGG | total);
            int64_t j = (int64_t)tile - 1;
            while (j >= 0) {
                const uint32_t* p = sstat + (size_t)j * R + d;
                uint32_t s = ss_load(p);
                uint32_t spins = 0;
                while ((s >> 30) == 0) {
                    if (++spins > kSpinLimit) {
                        atomicOr(status, ST_SPIN);
                        s = SS_INC;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    s = ss_load(p);
                }
                prefix += s & SS_VAL;
                if ((s >> 30) == 2) break;
                --j;
            }
            ss_store(st, SS_INC | ((prefix + total) & SS_VAL));
        }
        s_goff[d] = dstart + prefix;
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t pos = wbase + k * 64 + l;
        if (pos < n) {
            const uint32_t d = (key[k] >> shift) & (R - 1);
            const uint32_t dst = s_goff[d] + s_cnt[w][d] + rank[k];
            kout[dst] = key[k];
            vout[dst] = val[k];
        }
    }
}

template <int RB>
static void sort_impl(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const uint32_t* d_n, uint32_t cap,
                      int key_bits, uint32_t* scratch, uint32_t* tile_ctr, uint32_t* status, hipStream_t st) {
    constexpr int R = 1 << RB;
    const int passes = sort_passes(key_bits, RB);
    const uint32_t tiles = (cap + kSortTile - 1) / kSortTile;
    if (tiles == 0) return;
    uint32_t* hist = scratch;
    uint32_t* sstat = scratch + (size_t)passes * R;
    uint32_t hblocks = (cap + kBlock * 16 - 1) / (kBlock * 16);
    hblocks = hblocks < 1 ? 1 : (hblocks > 1024 ? 1024 : hblocks);
    hipLaunchKernelGGL(k_sort_hist<RB>, dim3(hblocks), dim3(kBlock), 0, st, k0, d_n, cap, passes, hist);
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_sort_pass<RB>, dim3(tiles), dim3(kBlock), 0, st, ki, vi, ko, vo, d_n, cap,
                           hist + (size_t)p * R, sstat + (size_t)p * tiles * R, p * RB, tile_ctr + p, status);
        uint32_t* t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
}

void launch_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const uint32_t* d_n, uint32_t cap,
                       int key_bits, int radix_bits, uint32_t* scratch, uint32_t* tile_ctr, uint32_t* status,
                       hipStream_t st) {
    switch (radix_bits) {
        case 4: sort_impl<4>(k0, v0, k1, v1, d_n, cap, key_bits, scratch, tile_ctr, status, st); break;
        case 6: sort_impl<6>(k0, v0, k1, v1, d_n, cap, key_bits, scratch, tile_ctr, status, st); break;
        default: sort_impl<8>(k0, v0, k1, v1, d_n, cap, key_bits, scratch, tile_ctr, status, st); break;
    }
}

// ------------------------------------------------------------------ scan
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

size_t scan_scratch_bytes(uint32_t n) {
    const size_t tiles = ((size_t)n + kScanTile - 1) / kScanTile;
    return ((tiles + 2) * sizeof(uint64_t) + 15) & ~(size_t)15;
}

__global__ void k_zero_u64(uint64_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_scan(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                 uint32_t n, int inclusive, uint64_t* scratch) {
    __shared__ uint32_t s_scan[kBlock / 64];
    __shared__ uint32_t s_word;
    uint32_t* ctr = reinterpret_cast<uint32_t*>(scratch);
    uint32_t* status = ctr + 1;
    uint64_t* lb = scratch + 2;
    const uint32_t tile = next_tile(ctr, &s_word);
    const uint32_t base = tile * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t i = base + k;
        v[k] = i < n ? in[i] : 0u;
        sum += v[k];
    }
    uint32_t agg;
    const uint32_t local = block_excl_scan<kBlock / 64>(sum, &agg, s_scan);
    const uint32_t prefix = lb_tile_prefix(lb, tile, agg, status, &s_word);
    uint32_t run = prefix + local;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t i = base + k;
        const uint32_t x = v[k];
        if (i < n) out[i] = inclusive ? run + x : run;
        run += x;
    }
}

void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, bool inclusive, void* scratch, hipStream_t st) {
    if (n == 0) return;
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    uint64_t* s = reinterpret_cast<uint64_t*>(scratch);
    hipLaunchKernelGGL(k_zero_u64, dim3((tiles + 2 + 255) / 256), dim3(256), 0, st, s, (size_t)tiles + 2);
    hipLaunchKernelGGL(k_scan, dim3(tiles), dim3(kBlock), 0, st, in, out, n, inclusive ? 1 : 0, s);
}

}  // namespace sg
