// Drop-in use of include/shinobi/grid.hpp the way the reference's
// StaticRTEngine::build (src/Engine/StaticRTEngine.h:225-237) drives its
// builder, followed by one trace.  Prints one JSON line with the grid sizes
// and checksums of the grid arrays and hits for the GPU test to compare.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <numeric>
#include <vector>

#include "shinobi/grid.hpp"

using namespace shinobi;

static unsigned long long checksum(const std::vector<uint32_t>& v) {
    unsigned long long h = 1469598103934665603ull;
    for (uint32_t x : v) h = (h ^ x) * 1099511628211ull;
    return h;
}

int main(int argc, char** argv) {
    const uint32_t n = argc > 1 ? (uint32_t)atoi(argv[1]) : 20000;
    const uint32_t nr = argc > 2 ? (uint32_t)atoi(argv[2]) : 50000;
    try {
        std::vector<float> soa(9 * (size_t)n);
        float* p[9];
        for (int i = 0; i < 9; ++i) p[i] = soa.data() + (size_t)i * n;
        check(sg_gen_soup(17, n, 0.02f, p), "sg_gen_soup");
        std::vector<Triangle> tris(n);
        for (uint32_t i = 0; i < n; ++i)
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) tris[i].vertices[v][a] = p[3 * v + a][i];

        Context& ctx = Context::global();
        FaceSoup storage(ctx);
        TwoLevelGrid grid;
        FastTwoLevelGridBuilder<FaceSoup> builder;
        builder.params.lambda_leaf = 2.0f;
        builder.init(grid, storage, tris);
        builder.build(grid, storage);
        GridParameters gp = grid.getParameters();
        std::vector<TwoLevelGridCell> cells = grid.cells();
        std::vector<uint32_t> leaves = grid.leaves();
        std::vector<uint32_t> idx = grid.indices();

        std::vector<float> r(7 * (size_t)nr);
        float* rp[7];
        for (int i = 0; i < 7; ++i) rp[i] = r.data() + (size_t)i * nr;
        check(sg_gen_random_rays(18, nr, 0.0f, 1.0f, rp), "sg_gen_random_rays");
        void *d_r = nullptr, *d_prim = nullptr, *d_t = nullptr;
        check(sg_malloc(ctx.get(), r.size() * 4, &d_r), "sg_malloc");
        check(sg_malloc(ctx.get(), (size_t)nr * 4, &d_prim), "sg_malloc");
        check(sg_malloc(ctx.get(), (size_t)nr * 4, &d_t), "sg_malloc");
        check(sg_memcpy_htod(ctx.get(), d_r, r.data(), r.size() * 4, nullptr), "htod");
        sg_rays rays;
        for (int i = 0; i < 3; ++i) {
            rays.o[i] = (const float*)d_r + (size_t)i * nr;
            rays.d[i] = (const float*)d_r + (size_t)(3 + i) * nr;
        }
        rays.tmax = (const float*)d_r + (size_t)6 * nr;
        trace(grid, ctx, rays, nr, (uint32_t*)d_prim, (float*)d_t);
        std::vector<uint32_t> prim(nr);
        check(sg_memcpy_dtoh(ctx.get(), prim.data(), d_prim, (size_t)nr * 4, nullptr), "dtoh");
        size_t hits = 0;
        for (uint32_t x : prim) hits += x != 0xFFFFFFFFu;
        std::vector<uint32_t> cw(cells.size() * 2);
        memcpy(cw.data(), cells.data(), cw.size() * 4);
        printf("{\"res\": [%d, %d, %d], \"leavesCount\": %u, \"instancesCount\": %u, \"cells\": %llu, "
               "\"leaves\": %llu, \"indices\": %llu, \"prim\": %llu, \"hits\": %zu, \"cell_rcp_x\": %.9g}\n",
               gp.res[0], gp.res[1], gp.res[2], grid.leavesCount, grid.instancesCount, checksum(cw),
               checksum(leaves), checksum(idx), checksum(prim), hits, gp.cellSizeRCP.x);
        sg_free(ctx.get(), d_r);
        sg_free(ctx.get(), d_prim);
        sg_free(ctx.get(), d_t);
    } catch (const Error& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
