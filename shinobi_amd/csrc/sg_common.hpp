// sg_common.hpp -- types and device helpers shared by the gfx950 kernels of
// the grid-acceleration hot path.
//
// Floating point: every kernel is compiled with -ffp-contract=off and without
// fast-math, so each binary32 operation rounds once, in the order written.
// The helpers below keep the reference's evaluation order
// (src/Core/Algebra.hpp:485-510, 570-577, 606-609).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SG_HD __host__ __device__ __forceinline__
#define SG_D __device__ __forceinline__

namespace sg {

constexpr float kEps = 0.00001f;            // src/CUDAStdAfx.h:96
constexpr uint32_t kNotEmpty = 0x40000000u;  // src/Control/Structure/TwoLevelGrid.hpp:66
constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr int kBlock = 256;                  // threads per block for build kernels

// ---------------------------------------------------------------- status bits
enum : uint32_t {
    ST_INVALID = 1u << 0,        // empty scene / zero extent / resolution < 1
    ST_LEAF_RES = 1u << 1,       // leaf resolution outside [1,255]
    ST_TOO_MANY_CELLS = 1u << 2, // C >= 2^24 (float key products inexact)
    ST_NEG_EXTENT = 1u << 3,     // negative leaf extent
    ST_CAP_P1 = 1u << 4,         // capacity overflow: top pairs
    ST_CAP_C1 = 1u << 5,         // top cells
    ST_CAP_C2 = 1u << 6,         // leaf cells
    ST_CAP_P2 = 1u << 7,         // leaf pairs
};

// -------------------------------------------------------------- device header
// One per grid, in device memory.  Written by the build kernels, read by the
// traversal kernels; the host reads it back only in sg_grid_get_info.
struct alignas(256) GridHeader {
    uint32_t bmin_key[3], bmax_key[3];   // order-preserving float keys (atomics)
    float bmin[3], bmax[3];
    int32_t res[3];
    float cs[3];         // (max-min)/res           build and traversal cell size
    float rcp_build[3];  // 1/cs                    build (FastSimpleGridBuilder.h:117)
    float cell_rcp[3];   // res/(max-min)           traversal (SimpleGrid.h:104-107)
    uint32_t n_tris;
    uint32_t C1, P1, P1_valid, C2, P2, P2_valid;
    uint32_t need_P1, need_C2, need_P2;  // true totals (exceed cap_* on overflow)
    uint32_t status;
    uint32_t leaf_sort;  // accepted leaf pairs the global sort must order (0: no global sort needed)
    uint32_t oversize;   // leaf tiles too large for the LDS sort met by this build
    // leaf cell size 1/(cell_rcp[d]*c) along d for a top cell of leaf
    // resolution c: the subCellSize of the two-level traversal
    // (src/Control/Algorithm/TwoLevelGridTraverser.h:499-510), computed once
    // per build instead of three divisions per entered top cell
    float leaf_size[3][256];
};

// Build parameters passed by value to the kernels.
struct BuildArgs {
    float lambda;
    int32_t top_divisor;       // 1 = uniform grid
    float lambda_leaf;
    uint32_t leaf_min_refs;
    int32_t fixed_res[3];      // all > 0: fixed (rebuild)
    uint32_t cap_C1, cap_P1, cap_C2, cap_P2;
    uint32_t keep_debug;       // write per-triangle / per-pair counts
};

// ------------------------------------------------------------------ vec3
struct v3 {
    float x, y, z;
};
SG_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
SG_HD v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
SG_HD v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
SG_HD v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
SG_HD v3 operator/(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
SG_HD v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
SG_HD v3 vmin(v3 a, v3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
SG_HD v3 vmax(v3 a, v3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
SG_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SG_HD v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
SG_HD v3 normalize(v3 a) {
    float r = 1.0f / sqrtf(dot(a, a));
    return scl(a, r);
}
SG_HD float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
SG_HD void setc(v3& a, int i, float f) {
    if (i == 0) a.x = f;
    else if (i == 1) a.y = f;
    else a.z = f;
}

// float -> integer with the saturating semantics of the GPU conversion
// instructions (NaN -> 0); equal to C casts within range.
SG_HD uint32_t f2u(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}
SG_HD int32_t f2i(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

// order-preserving map float -> uint32 for atomic min/max
SG_HD uint32_t fkey(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
SG_HD float fkey_inv(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __builtin_bit_cast(float, u);
}

// -------------------------------------------------------- deterministic pow
// x^y for x > 0 from IEEE binary64 +,-,*,/ only (atanh series for log, Taylor
// series for exp, exact exponent split), so host and device produce the same
// bits.  cbrtpow() stands in for powf(x, 0.3333333f) of
// src/Control/Algorithm/GridBuilder.h:49-50 and
// src/Control/Algorithm/TwoLevelBuildKernels.h:57-58.
SG_HD double det_pow(double x, double y) {
    const double atanh_c[14] = {
        0x1.0000000000000p+0, 0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3,
        0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111111p-4,
        0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5, 0x1.8618618618618p-5, 0x1.642c8590b2164p-5,
        0x1.47ae147ae147bp-5, 0x1.2f684bda12f68p-5};
    const double exp_c[18] = {
        0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3,
        0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
        0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
        0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41,
        0x1.ae7f3e733b81fp-45, 0x1.952c77030ad4ap-49};
    const double ln2_hi = 0x1.62e42fef00000p-1, ln2_lo = 0x1.473de00000000p-34;
    const double inv_ln2 = 0x1.71547652b82fep+0;
    if (x != x || y != y) return __builtin_nan("");
    if (x <= 0.0) return x == 0.0 ? 0.0 : __builtin_nan("");
    if (x == __builtin_inf()) return __builtin_inf();
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {
        x = x * 0x1p54;
        b = __builtin_bit_cast(uint64_t, x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = __builtin_bit_cast(double, (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    double z = (m - 1.0) / (m + 1.0);
    double z2 = z * z;
    double s = atanh_c[13];
    for (int k = 12; k >= 0; --k) s = s * z2 + atanh_c[k];
    double lnm = (2.0 * z) * s;
    double de = (double)e;
    double lnx = (de * ln2_hi + lnm) + de * ln2_lo;
    double w = y * lnx;
    if (w > 709.0) return __builtin_inf();
    if (w < -708.0) return 0.0;
    double n = __builtin_floor(w * inv_ln2 + 0.5);
    double r = (w - n * ln2_hi) - n * ln2_lo;
    double p = exp_c[17];
    for (int k = 16; k >= 0; --k) p = p * r + exp_c[k];
    int ni = (int)n;
    return p * __builtin_bit_cast(double, (uint64_t)(ni + 1023) << 52);
}
SG_HD float cbrtpow(float x) { return (float)det_pow((double)x, (double)0.3333333f); }

// sin/cos of 2*pi*u (double Taylor series after an exact quarter-turn
// reduction); stands in for sinCos(2*M_PI*u), src/Utils/HemisphereSamplers.hpp:64
SG_HD void sincos_2pi(float u, float* so, float* co) {
    const double sc[10] = {
        0x1.0000000000000p+0, -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
        0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33, -0x1.ae7f3e733b81fp-41,
        0x1.952c77030ad4ap-49, -0x1.2f49b46814157p-57};
    const double cc[10] = {
        0x1.0000000000000p+0, -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
        0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29, -0x1.93974a8c07c9dp-37,
        0x1.ae7f3e733b81fp-45, -0x1.6827863b97d97p-53};
    const double two_pi = 0x1.921fb54442d18p+2;
    double a = (double)u;
    double q = __builtin_floor(a * 4.0 + 0.5);
    double f = a - q * 0.25;
    double th = f * two_pi;
    double t2 = th * th;
    double ps = sc[9], pc = cc[9];
    for (int k = 8; k >= 0; --k) {
        ps = ps * t2 + sc[k];
        pc = pc * t2 + cc[k];
    }
    double sn = th * ps, cs = pc;
    int qi = ((int)q) & 3;
    double s, c;
    if (qi == 0) { s = sn; c = cs; }
    else if (qi == 1) { s = cs; c = -sn; }
    else if (qi == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
    *so = (float)s;
    *co = (float)c;
}

// SplitMix64 counter hash (ray / scene generators, bounce sampling)
SG_HD uint64_t sm64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
SG_HD uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) { return sm64(sm64(sm64(seed) + a) + b); }
SG_HD float u01(uint64_t h) { return (float)(uint32_t)(h >> 40) * 0x1p-24f; }

// ---------------------------------------------------------- grid parameters
// Resolution (GridBuilder.h:41-62; two-level FastTwoLevelGridBuilder.h:87-102)
// and the derived cell sizes.  Returns false for an invalid scene.
SG_HD bool derive_params(const float bmin[3], const float bmax[3], uint32_t n, float lambda,
                         int divisor, const int32_t fixed[3], int32_t res[3], float cs[3],
                         float rcpb[3], float crcp[3]) {
    v3 d = mk(bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]);
    if (n == 0 || !(d.x > 0.0f) || !(d.y > 0.0f) || !(d.z > 0.0f)) return false;
    if (fixed[0] > 0 && fixed[1] > 0 && fixed[2] > 0) {
        res[0] = fixed[0];
        res[1] = fixed[1];
        res[2] = fixed[2];
    } else {
        float volume = d.x * d.y * d.z;
        float m = cbrtpow(lambda * (float)n / volume);
        v3 dm = scl(d, m);
        res[0] = f2i(dm.x);
        res[1] = f2i(dm.y);
        res[2] = f2i(dm.z);
        if (divisor > 1) {
            for (int i = 0; i < 3; ++i) {
                int r = res[i] / divisor;
                res[i] = r > 2 ? r : 2;
            }
        }
    }
    if (res[0] < 1 || res[1] < 1 || res[2] < 1) return false;
    v3 rf = mk((float)res[0], (float)res[1], (float)res[2]);
    v3 mn = mk(bmin[0], bmin[1], bmin[2]), mx = mk(bmax[0], bmax[1], bmax[2]);
    v3 c = (mx - mn) / rf;
    v3 r = mk(1.f, 1.f, 1.f) / c;
    v3 cr = rf / (mx - mn);
    cs[0] = c.x; cs[1] = c.y; cs[2] = c.z;
    rcpb[0] = r.x; rcpb[1] = r.y; rcpb[2] = r.z;
    crcp[0] = cr.x; crcp[1] = cr.y; crcp[2] = cr.z;
    return true;
}

}  // namespace sg
