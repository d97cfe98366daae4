// sg_gen.cpp -- seeded synthetic inputs for the grid-acceleration workloads
// (host code).  Every value comes from a SplitMix64 counter hash of
// (seed, stream, index), so any element can be regenerated on its own and the
// output does not depend on the thread count.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <thread>
#include <vector>

#include "../../include/shinobi_grid.h"

namespace {

inline uint64_t sm64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) { return sm64(sm64(sm64(seed) + a) + b); }
inline float u01(uint64_t h) { return (float)(uint32_t)(h >> 40) * 0x1p-24f; }

template <class F>
void parallel(uint32_t n, F f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (n < 65536) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t)
        th.emplace_back([=] { f((uint32_t)((uint64_t)n * t / nt), (uint32_t)((uint64_t)n * (t + 1) / nt)); });
    f(0, (uint32_t)((uint64_t)n / nt));
    for (auto& x : th) x.join();
}

struct V {
    float x, y, z;
};
inline V mk(float x, float y, float z) { return V{x, y, z}; }
inline V sub(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V add(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V scl(V a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V divs(V a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline V norm(V a) { return scl(a, 1.f / sqrtf(dot(a, a))); }

// lattice value noise, one octave
double value_noise(uint64_t seed, int oct, double x, double y) {
    const double f = (double)(4 << oct);
    const double fx = x * f, fy = y * f;
    const double ix = floor(fx), iy = floor(fy);
    const double tx = fx - ix, ty = fy - iy;
    const double sx = tx * tx * (3.0 - 2.0 * tx), sy = ty * ty * (3.0 - 2.0 * ty);
    auto lat = [&](int64_t a, int64_t b) {
        return (double)u01(hash3(seed, 1000 + (uint64_t)oct, ((uint64_t)(uint32_t)a << 32) | (uint32_t)b));
    };
    const int64_t a = (int64_t)ix, b = (int64_t)iy;
    const double v00 = lat(a, b), v10 = lat(a + 1, b), v01 = lat(a, b + 1), v11 = lat(a + 1, b + 1);
    const double v0 = v00 + (v10 - v00) * sx, v1 = v01 + (v11 - v01) * sx;
    return v0 + (v1 - v0) * sy;
}

}  // namespace

extern "C" {

// uniform-in-cube soup: centroid ~ U[0,1]^3, each vertex = centroid +
// U[-edge,edge]^3 (the 10,000-triangle workload uses edge = 0.01)
sg_status sg_gen_soup(uint64_t seed, uint32_t n, float edge, float* const soa9[9]) {
    for (int i = 0; i < 9; ++i)
        if (!soa9[i]) return SG_ERR_INVALID_ARG;
    parallel(n, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            float c[3];
            for (int a = 0; a < 3; ++a) c[a] = u01(hash3(seed, i, (uint64_t)a));
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) {
                    const float u = u01(hash3(seed, i, 3 + (uint64_t)(3 * v + a)));
                    soa9[3 * v + a][i] = c[a] + (u * 2.0f - 1.0f) * edge;
                }
        }
    });
    return SG_OK;
}

float sg_heightfield_height(uint64_t seed, float x, float y, float amplitude) {
    double s = 0.0, w = 0.0, a = 1.0;
    for (int o = 0; o < 8; ++o) {
        s += a * value_noise(seed, o, (double)x, (double)y);
        w += a;
        a *= 0.5;
    }
    return (float)((double)amplitude * (s / w));
}

// height field over [0,1]^2: grid x grid vertices, (grid-1)^2 quads, two
// triangles per quad ((i,j),(i+1,j),(i+1,j+1)) and ((i,j),(i+1,j+1),(i,j+1)),
// heights = amplitude * normalised sum of 8 octaves of value noise
sg_status sg_gen_heightfield(uint64_t seed, uint32_t grid, float amplitude, float* const soa9[9]) {
    if (grid < 2) return SG_ERR_INVALID_ARG;
    for (int i = 0; i < 9; ++i)
        if (!soa9[i]) return SG_ERR_INVALID_ARG;
    const uint32_t g = grid;
    std::vector<float> h((size_t)g * g);
    parallel(g, [&](uint32_t b, uint32_t e) {
        for (uint32_t j = b; j < e; ++j)
            for (uint32_t i = 0; i < g; ++i)
                h[(size_t)j * g + i] =
                    sg_heightfield_height(seed, (float)i / (float)(g - 1), (float)j / (float)(g - 1), amplitude);
    });
    const uint32_t q = g - 1;
    parallel(q, [&](uint32_t b, uint32_t e) {
        for (uint32_t j = b; j < e; ++j)
            for (uint32_t i = 0; i < q; ++i) {
                const float x0 = (float)i / (float)(g - 1), x1 = (float)(i + 1) / (float)(g - 1);
                const float y0 = (float)j / (float)(g - 1), y1 = (float)(j + 1) / (float)(g - 1);
                const float h00 = h[(size_t)j * g + i], h10 = h[(size_t)j * g + i + 1];
                const float h01 = h[(size_t)(j + 1) * g + i], h11 = h[(size_t)(j + 1) * g + i + 1];
                const size_t t0 = 2 * ((size_t)j * q + i), t1 = t0 + 1;
                const float A[2][9] = {{x0, y0, h00, x1, y0, h10, x1, y1, h11}, {x0, y0, h00, x1, y1, h11, x0, y1, h01}};
                for (int k = 0; k < 9; ++k) {
                    soa9[k][t0] = A[0][k];
                    soa9[k][t1] = A[1][k];
                }
            }
    });
    return SG_OK;
}

// dense cluster in a large scene: n_small triangles with centroids uniform in
// the sphere of radius 0.02 around (0.5,0.5,0.5) and vertex offsets
// U[-1e-4,1e-4]^3, then n_large triangles with centroids uniform in [0,10]^3
// and offsets U[-0.3,0.3]^3
sg_status sg_gen_cluster(uint64_t seed, uint32_t n_small, uint32_t n_large, float* const soa9[9]) {
    for (int i = 0; i < 9; ++i)
        if (!soa9[i]) return SG_ERR_INVALID_ARG;
    const uint32_t n = n_small + n_large;
    parallel(n, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            float c[3];
            float h;
            if (i < n_small) {
                // uniform in the ball: radius r = R u^(1/3), direction uniform
                const double u = u01(hash3(seed, i, 0)), z = 2.0 * u01(hash3(seed, i, 1)) - 1.0;
                const double ph = 6.283185307179586 * u01(hash3(seed, i, 2));
                const double r = 0.02 * cbrt(u), s = sqrt(1.0 - z * z);
                c[0] = (float)(0.5 + r * s * cos(ph));
                c[1] = (float)(0.5 + r * s * sin(ph));
                c[2] = (float)(0.5 + r * z);
                h = 1e-4f;
            } else {
                for (int a = 0; a < 3; ++a) c[a] = 10.0f * u01(hash3(seed, i, (uint64_t)a));
                h = 0.3f;
            }
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) {
                    const float u = u01(hash3(seed, i, 3 + (uint64_t)(3 * v + a)));
                    soa9[3 * v + a][i] = c[a] + (u * 2.0f - 1.0f) * h;
                }
        }
    });
    return SG_OK;
}

// Camera::init (src/Primitive/Camera.h:41-60): normalised forward, right =
// ~(forward x up), up = ~(forward x right); row / col vectors from the
// vertical opening angle and the aspect ratio
sg_status sg_pinhole_init(const float eye[3], const float fwd[3], const float up[3], float fov_deg, uint32_t res_x,
                          uint32_t res_y, sg_pinhole* out) {
    if (!eye || !fwd || !up || !out || res_x == 0 || res_y == 0) return SG_ERR_INVALID_ARG;
    const float aspect = (float)res_x / (float)res_y;
    const V f = norm(mk(fwd[0], fwd[1], fwd[2]));
    const V r = norm(cross(f, mk(up[0], up[1], up[2])));
    const V u = norm(cross(f, r));
    const float ang = fov_deg * 3.14159265358979323846f / 180.f;
    const float tn = tanf(ang / 2.f);
    const V row = scl(scl(scl(r, 2.f), tn), aspect);
    const V col = scl(scl(u, 2.f), tn);
    const V stepX = divs(row, (float)res_x), stepY = divs(col, (float)res_y);
    const V topLeft = sub(sub(f, divs(row, 2.f)), divs(col, 2.f));
    const V e = mk(eye[0], eye[1], eye[2]);
    const V* src[4] = {&e, &topLeft, &stepX, &stepY};
    float* dst[4] = {out->position, out->top_left, out->step_x, out->step_y};
    for (int k = 0; k < 4; ++k) {
        dst[k][0] = src[k]->x;
        dst[k][1] = src[k]->y;
        dst[k][2] = src[k]->z;
    }
    out->res_x = res_x;
    out->res_y = res_y;
    return SG_OK;
}

// Camera::getDirection (src/Primitive/Camera.h:67-70) with the
// RegularPixelSampler<1,1> of src/Primitive/PixelSampler.h:34-54 and
// RegularPrimaryRayGenerator (src/Engine/RayGenerators.h:36-61); t = FLT_MAX
sg_status sg_gen_camera_rays(const float eye[3], const float fwd[3], const float up[3], float fov_deg,
                             uint32_t res_x, uint32_t res_y, float* const rays7[7]) {
    for (int i = 0; i < 7; ++i)
        if (!rays7[i]) return SG_ERR_INVALID_ARG;
    sg_pinhole cam;
    const sg_status st = sg_pinhole_init(eye, fwd, up, fov_deg, res_x, res_y, &cam);
    if (st != SG_OK) return st;
    const V topLeft = mk(cam.top_left[0], cam.top_left[1], cam.top_left[2]);
    const V stepX = mk(cam.step_x[0], cam.step_x[1], cam.step_x[2]);
    const V stepY = mk(cam.step_y[0], cam.step_y[1], cam.step_y[2]);
    const uint32_t n = res_x * res_y;
    const float rx = (float)res_x;
    parallel(n, [&](uint32_t b, uint32_t e) {
        for (uint32_t id = b; id < e; ++id) {
            float py = (truncf(0.f) + 0.5f) / 1.f;
            float px = 0.f - truncf(0.f) + 0.5f / 1.f;
            const float tmp = (float)id / rx;
            py += truncf(tmp);
            px += (tmp - truncf(tmp)) * rx;
            const V d = add(add(topLeft, scl(stepX, px)), scl(stepY, py));
            rays7[0][id] = eye[0];
            rays7[1][id] = eye[1];
            rays7[2][id] = eye[2];
            rays7[3][id] = d.x;
            rays7[4][id] = d.y;
            rays7[5][id] = d.z;
            rays7[6][id] = 3.402823466e+38f;
        }
    });
    return SG_OK;
}

sg_status sg_gen_random_rays(uint64_t seed, uint32_t n, float lo, float hi, float* const rays7[7]) {
    for (int i = 0; i < 7; ++i)
        if (!rays7[i]) return SG_ERR_INVALID_ARG;
    parallel(n, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            for (int a = 0; a < 3; ++a) rays7[a][i] = lo + (hi - lo) * u01(hash3(seed, i, (uint64_t)a));
            const double z = 2.0 * u01(hash3(seed, i, 3)) - 1.0;
            const double ph = 6.283185307179586 * u01(hash3(seed, i, 4));
            const double s = sqrt(1.0 - z * z);
            rays7[3][i] = (float)(s * cos(ph));
            rays7[4][i] = (float)(s * sin(ph));
            rays7[5][i] = (float)z;
            rays7[6][i] = 3.402823466e+38f;
        }
    });
    return SG_OK;
}

}  // extern "C"
