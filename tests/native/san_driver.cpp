// Host-code workout for the AddressSanitizer / UndefinedBehaviorSanitizer
// build (tests/test_sanitizers.py): the oracle (builders, traversals,
// referees, bounce, primitives) on ordinary and degenerate inputs, the
// synthetic generators, the OBJ and camera.cfg readers on the given files,
// and the serving loop's first-fit arena under random alloc / release.
// Exits non-zero on a failed check; the sanitizers abort on their findings.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <random>
#include <vector>

#include "../../include/shinobi_grid.h"
#include "../../oracle/sg_oracle.h"
#include "../../shinobi_amd/csrc/sg_arena.hpp"

#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            exit(1);                                                     \
        }                                                                \
    } while (0)

struct Soa {
    std::vector<float> a;
    uint32_t n;
    int rows;
    Soa(int r, uint32_t n_) : a((size_t)r * (n_ ? n_ : 1)), n(n_), rows(r) {}
    float* row(int k) { return a.data() + (size_t)k * (n ? n : 1); }
    std::vector<float*> ptrs() {
        std::vector<float*> p(rows);
        for (int k = 0; k < rows; ++k) p[k] = row(k);
        return p;
    }
};

static or_params params(int nthreads, float lambda_leaf = 2.0f) {
    or_params p;
    p.lambda = 5.0f;
    p.top_divisor = 6;
    p.lambda_leaf = lambda_leaf;
    p.leaf_min_refs = 16;
    p.res_override[0] = p.res_override[1] = p.res_override[2] = 0;
    p.nthreads = nthreads;
    return p;
}

// build both grid kinds over `tri`, trace `rays` every way, bounce
static void workout(Soa& tri, Soa& rays) {
    std::vector<float*> tp = tri.ptrs(), rp = rays.ptrs();
    const float* const* v = (const float* const*)tp.data();
    const float* const* r = (const float* const*)rp.data();
    const uint32_t nr = rays.n;
    std::vector<uint32_t> prim(nr ? nr : 1), stats(4 * (size_t)(nr ? nr : 1)), prim2(nr ? nr : 1);
    std::vector<float> t(nr ? nr : 1), buf(8 * (size_t)(nr ? nr : 1));
    std::vector<double> t64(nr ? nr : 1);
    for (int kind = 0; kind < 2; ++kind) {
        or_params p = params(3);
        if (kind == 0) p.top_divisor = 1;
        or_grid* g = kind ? or_build_twolevel(v, tri.n, &p) : or_build_uniform(v, tri.n, &p);
        CHECK(g != nullptr);
        if (g->status == OR_OK || g->status == OR_ERR_LEAF_RES || g->status == OR_ERR_NEG_EXTENT) {
            or_trace(g, v, r, nr, prim.data(), t.data(), stats.data(), 3);
            or_trace_store(g, v, r, nr, prim2.data(), t.data(), buf.data(), 2);
            for (uint32_t i = 0; i < nr; ++i) CHECK(prim[i] == prim2[i]);
            or_trace_shadow(g, v, r, nr, prim.data(), t.data(), nullptr, 3);
            // rebuild semantics: the same scene at a fixed resolution
            or_params q = p;
            for (int k = 0; k < 3; ++k) q.res_override[k] = g->res[k];
            or_grid* h = kind ? or_build_twolevel(v, tri.n, &q) : or_build_uniform(v, tri.n, &q);
            CHECK(h != nullptr && h->P1 == g->P1 && h->C1 == g->C1);
            or_grid_free(h);
        }
        or_grid_free(g);
    }
    if (tri.n <= 4000) {
        or_brute_f32(v, tri.n, r, nr, prim.data(), t.data(), 3);
        or_brute_f64(v, tri.n, r, nr, prim.data(), t64.data(), 3);
    }
    // bounce from the closest hits of a two-level grid
    or_params p = params(2);
    or_grid* g = or_build_twolevel(v, tri.n, &p);
    if (g && g->status == OR_OK) {
        or_trace(g, v, r, nr, prim.data(), t.data(), nullptr, 2);
        Soa out(7, nr);
        std::vector<float*> op = out.ptrs();
        std::vector<uint32_t> src(nr ? nr : 1);
        const uint32_t m = or_bounce(v, r, prim.data(), t.data(), nr, 9, op.data(), src.data());
        CHECK(m <= nr);
        for (uint32_t i = 0; i < m; ++i) CHECK(src[i] < nr && prim[src[i]] != OR_MISS);
    }
    or_grid_free(g);
}

static void oracle_cases() {
    // ordinary scenes from the generators
    {
        Soa tri(9, 3000), rays(7, 5000);
        std::vector<float*> tp = tri.ptrs(), rp = rays.ptrs();
        CHECK(sg_gen_soup(3, tri.n, 0.03f, tp.data()) == SG_OK);
        CHECK(sg_gen_random_rays(4, rays.n, 0.f, 1.f, rp.data()) == SG_OK);
        workout(tri, rays);
    }
    {
        Soa tri(9, 2 * 63 * 63), rays(7, 64 * 48);
        std::vector<float*> tp = tri.ptrs(), rp = rays.ptrs();
        CHECK(sg_gen_heightfield(5, 64, 0.2f, tp.data()) == SG_OK);
        const float eye[3] = {0.5f, -0.3f, 0.6f}, fwd[3] = {0.f, 0.8f, -0.5f}, up[3] = {0.f, 0.f, 1.f};
        CHECK(sg_gen_camera_rays(eye, fwd, up, 60.f, 64, 48, rp.data()) == SG_OK);
        sg_pinhole cam;
        CHECK(sg_pinhole_init(eye, fwd, up, 60.f, 64, 48, &cam) == SG_OK);
        workout(tri, rays);
    }
    {
        // dense cluster: oversize leaf tiles
        Soa tri(9, 20000 + 100), rays(7, 4000);
        std::vector<float*> tp = tri.ptrs(), rp = rays.ptrs();
        CHECK(sg_gen_cluster(6, 20000, 100, tp.data()) == SG_OK);
        CHECK(sg_gen_random_rays(7, rays.n, 0.45f, 0.55f, rp.data()) == SG_OK);
        workout(tri, rays);
    }
    // degenerate scenes: one triangle, a flat scene, coincident vertices
    {
        Soa tri(9, 1), rays(7, 64);
        const float v[9] = {0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f};
        for (int k = 0; k < 9; ++k) tri.row(k)[0] = v[k];
        std::vector<float*> rp = rays.ptrs();
        CHECK(sg_gen_random_rays(8, rays.n, -0.5f, 1.5f, rp.data()) == SG_OK);
        workout(tri, rays);
        for (int k = 0; k < 9; ++k) tri.row(k)[0] = 0.25f;   // zero-extent bounds
        workout(tri, rays);
    }
    // rays that the traversal must survive: NaN, zero and infinite
    // components, origins inside and on the boundary, zero tmax
    {
        Soa tri(9, 500), rays(7, 12);
        std::vector<float*> tp = tri.ptrs();
        CHECK(sg_gen_soup(9, tri.n, 0.05f, tp.data()) == SG_OK);
        const float nan = nanf(""), inf = INFINITY;
        const float R[12][7] = {
            {0.5f, 0.5f, 0.5f, 1.f, 0.f, 0.f, 1e30f}, {0.5f, 0.5f, 0.5f, 0.f, 0.f, 0.f, 1e30f},
            {nan, 0.5f, 0.5f, 1.f, 1.f, 1.f, 1e30f},  {0.5f, 0.5f, 0.5f, nan, 1.f, 0.f, 1e30f},
            {0.5f, 0.5f, 0.5f, inf, 1.f, 0.f, 1e30f}, {-1.f, -1.f, -1.f, 1.f, 1.f, 1.f, 1e30f},
            {0.f, 0.f, 0.f, 1.f, 1.f, 1.f, 0.f},      {2.f, 2.f, 2.f, 1.f, 1.f, 1.f, 1e30f},
            {0.5f, 0.5f, -5.f, 0.f, 0.f, 1.f, 1e30f}, {0.5f, 0.5f, 0.5f, -0.f, 0.f, 1.f, 1e30f},
            {0.5f, 0.5f, 0.5f, 1e-30f, 1.f, 1e-30f, inf}, {0.1f, 0.9f, 0.3f, 0.3f, -0.2f, 0.9f, nan}};
        for (int i = 0; i < 12; ++i)
            for (int k = 0; k < 7; ++k) rays.row(k)[i] = R[i][k];
        workout(tri, rays);
    }
    // primitives: scans with wrap-around, sorts with duplicates, ranges
    {
        std::mt19937 rng(11);
        for (size_t n : {0, 1, 2, 1000, 100003}) {
            std::vector<uint32_t> a(n + 1), b(n + 1), k(n + 1), val(n + 1);
            for (size_t i = 0; i < n; ++i) {
                a[i] = rng();
                k[i] = rng() % 97;
                val[i] = (uint32_t)i;
            }
            or_exclusive_scan_u32(a.data(), b.data(), n);
            or_inclusive_scan_u32(a.data(), b.data(), n);
            CHECK(or_sort_pairs(k.data(), val.data(), n, 3) == 0);
            for (size_t i = 1; i < n; ++i) CHECK(k[i - 1] < k[i] || (k[i - 1] == k[i] && val[i - 1] < val[i]));
            std::vector<uint32_t> cells(2 * 97);
            or_ranges(k.data(), n, cells.data(), 97);
        }
    }
}

static void reader_cases(const char* obj_path, const char* bad_obj, const char* cam_path) {
    for (const char* path : {obj_path, bad_obj}) {
        sg_obj* o = nullptr;
        sg_obj_info info;
        const sg_status st = sg_obj_load(path, &o, &info);
        if (st != SG_OK) {
            CHECK(o == nullptr);
            continue;
        }
        Soa tri(9, info.n_triangles);
        std::vector<float*> tp = tri.ptrs();
        CHECK(sg_obj_triangles(o, tp.data()) == SG_OK);
        sg_obj_destroy(o);
    }
    sg_obj* o = nullptr;
    sg_obj_info info;
    CHECK(sg_obj_load("/nonexistent/scene.obj", &o, &info) == SG_ERR_IO);
    sg_camera cam;
    CHECK(sg_camera_load(cam_path, &cam) == SG_OK);
    CHECK(sg_camera_load("/nonexistent/camera.cfg", &cam) == SG_ERR_IO);
}

static void arena_cases() {
    std::mt19937 rng(12);
    const size_t size = 1 << 20;
    sg::Arena a(size);
    std::map<size_t, size_t> live;   // offset -> bytes
    for (int step = 0; step < 200000; ++step) {
        if (live.empty() || rng() % 2) {
            const size_t bytes = 1 + rng() % 40000;
            size_t off = 0;
            if (a.alloc(bytes, &off)) {
                CHECK(off + bytes <= size);
                auto nx = live.lower_bound(off);
                CHECK(nx == live.end() || off + bytes <= nx->first);
                if (nx != live.begin()) {
                    auto pv = std::prev(nx);
                    CHECK(pv->first + pv->second <= off);
                }
                live[off] = bytes;
            }
        } else {
            auto it = live.begin();
            std::advance(it, rng() % live.size());
            a.release(it->first, it->second);
            live.erase(it);
        }
        // free blocks never touch each other (coalesced) nor a live block
        size_t end = 0;
        bool first = true;
        for (auto& fb : a.free_) {
            CHECK(first || fb.first > end);
            first = false;
            end = fb.first + fb.second;
            CHECK(end <= size);
        }
    }
    for (auto& kv : live) a.release(kv.first, kv.second);
    CHECK(a.free_.size() == 1 && a.free_.begin()->first == 0 && a.free_.begin()->second == size);
}

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s scene.obj bad.obj camera.cfg\n", argv[0]);
        return 2;
    }
    oracle_cases();
    reader_cases(argv[1], argv[2], argv[3]);
    arena_cases();
    printf("sanitized host code: all checks passed\n");
    return 0;
}
