// sg_sort.hip -- stable LSD radix sort of (key, value) pairs and a prefix
// scan for gfx950, both reduce-then-scan (no inter-block waiting).
//
// Sort (semantics of chag::pp::sort as driven by src/Utils/Sort.cu:28-56),
// per pass over 4096-key tiles:
//   up     per-tile digit counts -> counts[digit][tile]
//   scan   one block per digit: exclusive scan of its row, and its total
//   down   scans the digit totals into bases, ranks the tile's keys per digit
//          with wave64 ballot matching (items in tile order, so the pass is
//          stable), stages the tile in LDS in digit order and writes every
//          digit run contiguously
// Only ceil(key_bits / radix_bits) passes run (the reference sorts 32-bit
// keys two bits at a time, 48 launches).
//
// Scan (src/Utils/Scan.h:35-68): block sums, one-block scan of the sums,
// then each block rescans its tile from its offset.
#include "sg_device.hpp"
#include "sg_kernels.hpp"

namespace sg {

constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;   // 4096 keys

int sort_passes(int key_bits, int radix_bits) {
    if (key_bits < 1) key_bits = 1;
    return (key_bits + radix_bits - 1) / radix_bits;
}

static size_t sort_tiles(uint32_t cap) { return ((size_t)cap + kSortTile - 1) / kSortTile; }

size_t sort_scratch_words(uint32_t cap, int key_bits, int radix_bits) {
    const size_t R = (size_t)1 << radix_bits;
    const size_t passes = (size_t)sort_passes(key_bits, radix_bits);
    size_t w = passes * R + R * (sort_tiles(cap) + 1);
    return (w + 3) & ~(size_t)3;   // 16-byte multiple
}

// per-tile digit counts of one pass -> counts[d * T + tile].  Each thread
// counts 16 consecutive keys and adds runs of equal digits with one LDS
// atomic, so runs of equal keys (cell-ordered emission) do not serialise.
// (the kernels are bodies taking the tile / digit index `bid`, called by
// the single-sort kernels with blockIdx.x and by the multi-grid kernels
// with the block's index within its grid)
template <int RB>
SG_D void sort_up_body(const uint32_t* __restrict__ keys, const uint32_t* d_n, uint32_t cap, int shift,
                       uint32_t* __restrict__ counts, uint32_t T, uint32_t bid) {
    constexpr int R = 1 << RB;
    __shared__ uint32_t s[R];
    const uint32_t n = min(*d_n, cap);
    const uint32_t base = bid * kSortTile;
    if (base >= n) return;
    for (int i = threadIdx.x; i < R; i += kBlock) s[i] = 0;
    __syncthreads();
    const uint32_t i0 = base + threadIdx.x * kSortItems;
    uint32_t k[kSortItems];
    if (i0 + kSortItems <= n) {
#pragma unroll
        for (int q = 0; q < kSortItems / 4; ++q) {
            const uint4 v = reinterpret_cast<const uint4*>(keys + i0)[q];
            k[4 * q] = v.x; k[4 * q + 1] = v.y; k[4 * q + 2] = v.z; k[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kSortItems; ++j) k[j] = i0 + j < n ? keys[i0 + j] : 0u;
    }
    const uint32_t m = i0 < n ? min((uint32_t)kSortItems, n - i0) : 0u;
    uint32_t rd = 0, rc = 0;
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        if ((uint32_t)j < m) {
            const uint32_t d = (k[j] >> shift) & (R - 1);
            if (rc && d != rd) {
                atomicAdd(&s[rd], rc);
                rc = 0;
            }
            rd = d;
            ++rc;
        }
    }
    if (rc) atomicAdd(&s[rd], rc);
    __syncthreads();
    for (int d = threadIdx.x; d < R; d += kBlock) counts[(size_t)d * T + bid] = s[d];
}

template <int RB>
__global__ __launch_bounds__(kBlock) void k_sort_up(const uint32_t* __restrict__ keys, const uint32_t* d_n,
                                                    uint32_t cap, int shift, uint32_t* __restrict__ counts,
                                                    uint32_t T) {
    sort_up_body<RB>(keys, d_n, cap, shift, counts, T, blockIdx.x);
}

// one block per digit d: counts[d][t] -> sum over t' < t (in place), and
// totals[d] = the digit's count over all tiles
template <int RB>
SG_D void sort_scan_body(const uint32_t* d_n, uint32_t cap, uint32_t* __restrict__ totals,
                         uint32_t* __restrict__ counts, uint32_t T, uint32_t bid) {
    __shared__ uint32_t s_scan[kBlock / 64];
    const uint32_t n = min(*d_n, cap);
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    const uint32_t d = bid;
    uint32_t run = 0;
    uint32_t* row = counts + (size_t)d * T;
    // chunks of 256*8 tiles: each thread scans 8 consecutive values
    for (uint32_t c0 = 0; c0 < tiles; c0 += kBlock * 8) {
        const uint32_t t0 = c0 + threadIdx.x * 8;
        uint32_t v[8], sum = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = t0 + k < tiles ? row[t0 + k] : 0u;
            sum += v[k];
        }
        uint32_t tot;
        uint32_t pre = run + block_excl_scan<kBlock / 64>(sum, &tot, s_scan);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (t0 + k < tiles) row[t0 + k] = pre;
            pre += v[k];
        }
        run += tot;
    }
    if (threadIdx.x == 0) totals[d] = run;
}

template <int RB>
__global__ __launch_bounds__(kBlock) void k_sort_scan(const uint32_t* d_n, uint32_t cap,
                                                      uint32_t* __restrict__ totals,
                                                      uint32_t* __restrict__ counts, uint32_t T) {
    sort_scan_body<RB>(d_n, cap, totals, counts, T, blockIdx.x);
}

template <int RB>
SG_D void sort_down_body(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                         uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, const uint32_t* d_n,
                         uint32_t cap, int shift, const uint32_t* __restrict__ totals,
                         const uint32_t* __restrict__ offsets, uint32_t T, uint32_t bid) {
    constexpr int R = 1 << RB;
    __shared__ uint32_t s_cnt[4][R];
    __shared__ uint32_t s_goff[R];
    __shared__ uint32_t s_tdo[R];
    __shared__ uint32_t s_k[kSortTile];
    __shared__ uint32_t s_v[kSortTile];
    __shared__ uint32_t s_scan[4];
    const uint32_t n = min(*d_n, cap);
    const uint32_t tile = bid;
    const uint32_t base = tile * kSortTile;
    if (base >= n) return;
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
    // stable rank inside the wave: lanes with the same digit found by
    // RB ballots; items processed in tile order
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
    uint32_t total = 0;
    if (threadIdx.x < R) {
        const int d = threadIdx.x;
        const uint32_t c0 = s_cnt[0][d], c1 = s_cnt[1][d], c2 = s_cnt[2][d], c3 = s_cnt[3][d];
        s_cnt[0][d] = 0u;
        s_cnt[1][d] = c0;
        s_cnt[2][d] = c0 + c1;
        s_cnt[3][d] = c0 + c1 + c2;
        total = c0 + c1 + c2 + c3;
    }
    uint32_t tsum;
    const uint32_t tdo = block_excl_scan<kBlock / 64>(total, &tsum, s_scan);
    // digit bases: exclusive scan of the digit totals
    const uint32_t dbase = block_excl_scan<kBlock / 64>(threadIdx.x < R ? totals[threadIdx.x] : 0u, &tsum, s_scan);
    if (threadIdx.x < R) {
        s_tdo[threadIdx.x] = tdo;
        s_goff[threadIdx.x] = dbase + offsets[(size_t)threadIdx.x * T + tile] - tdo;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t pos = wbase + k * 64 + l;
        if (pos < n) {
            const uint32_t d = (key[k] >> shift) & (R - 1);
            const uint32_t lp = s_tdo[d] + s_cnt[w][d] + rank[k];
            s_k[lp] = key[k];
            s_v[lp] = val[k];
        }
    }
    __syncthreads();
    const uint32_t cnt_tile = min((uint32_t)kSortTile, n - base);
    for (uint32_t j = threadIdx.x; j < cnt_tile; j += kBlock) {
        const uint32_t k = s_k[j];
        const uint32_t g = s_goff[(k >> shift) & (R - 1)] + j;
        kout[g] = k;
        vout[g] = s_v[j];
    }
}

template <int RB>
__global__ __launch_bounds__(kBlock) void k_sort_down(const uint32_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      const uint32_t* d_n, uint32_t cap, int shift,
                                                      const uint32_t* __restrict__ totals,
                                                      const uint32_t* __restrict__ offsets, uint32_t T) {
    sort_down_body<RB>(kin, vin, kout, vout, d_n, cap, shift, totals, offsets, T, blockIdx.x);
}

// pass `pass` of the sorts of many grids (UniJob): scratch layout as
// sort_impl (per-pass digit totals, then the counts matrix)
template <int RB>
__global__ __launch_bounds__(kBlock) void k_multi_sort_up(const UniJob* __restrict__ jobs,
                                                          const MultiEntry* __restrict__ e, uint32_t ne, int pass) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    sort_up_body<RB>(j.k[pass & 1], &j.h->P1, j.cap_P, pass * RB, j.sort + (size_t)j.passes * (1 << RB), j.T, bid);
}

template <int RB>
__global__ __launch_bounds__(kBlock) void k_multi_sort_scan(const UniJob* __restrict__ jobs,
                                                            const MultiEntry* __restrict__ e, uint32_t ne, int pass) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    sort_scan_body<RB>(&j.h->P1, j.cap_P, j.sort + (size_t)pass * (1 << RB), j.sort + (size_t)j.passes * (1 << RB),
                       j.T, bid);
}

template <int RB>
__global__ __launch_bounds__(kBlock) void k_multi_sort_down(const UniJob* __restrict__ jobs,
                                                            const MultiEntry* __restrict__ e, uint32_t ne, int pass) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    const int i = pass & 1, o = i ^ 1;
    sort_down_body<RB>(j.k[i], j.v[i], j.k[o], j.v[o], &j.h->P1, j.cap_P, pass * RB,
                       j.sort + (size_t)pass * (1 << RB), j.sort + (size_t)j.passes * (1 << RB), j.T, bid);
}

template <int RB>
static void sort_impl(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const uint32_t* d_n, uint32_t cap,
                      int key_bits, uint32_t* scratch, hipStream_t st) {
    constexpr int R = 1 << RB;
    const int passes = sort_passes(key_bits, RB);
    const uint32_t T = (uint32_t)sort_tiles(cap);
    if (T == 0) return;
    uint32_t* totals = scratch;
    uint32_t* counts = scratch + (size_t)passes * R;
    uint32_t *ki = k0, *vi = v0, *ko = k1, *vo = v1;
    for (int p = 0; p < passes; ++p) {
        uint32_t* tot = totals + (size_t)p * R;
        hipLaunchKernelGGL(k_sort_up<RB>, dim3(T), dim3(kBlock), 0, st, ki, d_n, cap, p * RB, counts, T);
        hipLaunchKernelGGL(k_sort_scan<RB>, dim3(R), dim3(kBlock), 0, st, d_n, cap, tot, counts, T);
        hipLaunchKernelGGL(k_sort_down<RB>, dim3(T), dim3(kBlock), 0, st, ki, vi, ko, vo, d_n, cap, p * RB, tot,
                           counts, T);
        uint32_t* t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
}

uint32_t multi_sort_blocks(const UniJob& j, MultiSortStep step) {
    return step == MS_SCAN ? (1u << j.rb) : j.T;
}

template <int RB>
static void multi_sort_impl(MultiSortStep step, int pass, const UniJob* jobs, const MultiEntry* e, uint32_t ne,
                            uint32_t blocks, hipStream_t st) {
    const dim3 g(blocks), b(kBlock);
    if (step == MS_UP) hipLaunchKernelGGL(k_multi_sort_up<RB>, g, b, 0, st, jobs, e, ne, pass);
    else if (step == MS_SCAN) hipLaunchKernelGGL(k_multi_sort_scan<RB>, g, b, 0, st, jobs, e, ne, pass);
    else hipLaunchKernelGGL(k_multi_sort_down<RB>, g, b, 0, st, jobs, e, ne, pass);
}

void launch_multi_sort(MultiSortStep step, int rb, int pass, const UniJob* jobs, const MultiEntry* e, uint32_t ne,
                       uint32_t blocks, hipStream_t st) {
    if (ne == 0 || blocks == 0) return;
    switch (rb) {
        case 4: multi_sort_impl<4>(step, pass, jobs, e, ne, blocks, st); break;
        case 6: multi_sort_impl<6>(step, pass, jobs, e, ne, blocks, st); break;
        default: multi_sort_impl<8>(step, pass, jobs, e, ne, blocks, st); break;
    }
}

void launch_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const uint32_t* d_n, uint32_t cap,
                       int key_bits, int radix_bits, uint32_t* scratch, hipStream_t st) {
    switch (radix_bits) {
        case 4: sort_impl<4>(k0, v0, k1, v1, d_n, cap, key_bits, scratch, st); break;
        case 6: sort_impl<6>(k0, v0, k1, v1, d_n, cap, key_bits, scratch, st); break;
        default: sort_impl<8>(k0, v0, k1, v1, d_n, cap, key_bits, scratch, st); break;
    }
}

// ------------------------------------------------------------------ scan
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;   // 4096 values

size_t scan_scratch_bytes(uint32_t n) {
    const size_t tiles = ((size_t)n + kScanTile - 1) / kScanTile;
    return ((tiles + 4) * sizeof(uint32_t) + 15) & ~(size_t)15;
}

__global__ __launch_bounds__(kBlock) void k_scan_up(const uint32_t* __restrict__ in, uint32_t n,
                                                    uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_red[kBlock / 64];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t acc = 0;
    for (uint32_t i = base + threadIdx.x; i < min(n, base + (uint32_t)kScanTile); i += kBlock) acc += in[i];
    const uint32_t s = block_sum<kBlock / 64>(acc, s_red);
    if (threadIdx.x == 0) sums[blockIdx.x] = s;
}

// exclusive scan of the tile sums in place (one block of 1024 threads)
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* sums, uint32_t tiles) {
    __shared__ uint32_t s_scan[16];
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < tiles; c0 += 1024 * 4) {
        const uint32_t t0 = c0 + threadIdx.x * 4;
        uint32_t v[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = t0 + k < tiles ? sums[t0 + k] : 0u;
            sum += v[k];
        }
        uint32_t tot;
        uint32_t pre = run + block_excl_scan<16>(sum, &tot, s_scan);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (t0 + k < tiles) sums[t0 + k] = pre;
            pre += v[k];
        }
        run += tot;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                      uint32_t n, int inclusive, const uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_scan[kBlock / 64];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
    if (base + kScanItems <= n) {
#pragma unroll
        for (int k = 0; k < kScanItems; k += 4) {
            const uint4 q = *reinterpret_cast<const uint4*>(in + base + k);
            v[k] = q.x; v[k + 1] = q.y; v[k + 2] = q.z; v[k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) v[k] = base + k < n ? in[base + k] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    uint32_t agg;
    uint32_t run = sums[blockIdx.x] + block_excl_scan<kBlock / 64>(sum, &agg, s_scan);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t x = v[k];
        v[k] = inclusive ? run + x : run;
        run += x;
    }
    if (base + kScanItems <= n) {
#pragma unroll
        for (int k = 0; k < kScanItems; k += 4)
            *reinterpret_cast<uint4*>(out + base + k) = make_uint4(v[k], v[k + 1], v[k + 2], v[k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) out[base + k] = v[k];
    }
}

// in and out must be 16-byte aligned (device allocations are)
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, bool inclusive, void* scratch, hipStream_t st) {
    if (n == 0) return;
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    uint32_t* sums = reinterpret_cast<uint32_t*>(scratch);
    hipLaunchKernelGGL(k_scan_up, dim3(tiles), dim3(kBlock), 0, st, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, sums, tiles);
    hipLaunchKernelGGL(k_scan_down, dim3(tiles), dim3(kBlock), 0, st, in, out, n, inclusive ? 1 : 0, sums);
}

}  // namespace sg
