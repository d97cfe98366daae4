// shinobi/grid.hpp -- header-only C++ layer that restores the reference's
// host-side concepts on top of the C ABI (shinobi_grid.h), so code written
// against the Shinobi builder / storage / grid types compiles against the
// gfx950 implementation with the same call pattern:
//
//   FaceSoup storage;                               // src/Control/Structure/FaceSoup.h:41-247
//   TwoLevelGrid grid;                              // src/Control/Structure/TwoLevelGrid.h:39-358
//   FastTwoLevelGridBuilder<FaceSoup> builder;      // src/Control/Algorithm/FastTwoLevelGridBuilder.h:45-783
//   builder.init(grid, storage, triangles);         // bounds, upload, resolution
//   builder.build(grid, storage);                   // device build
//   builder.rebuild(grid, storage, stream);         // same resolution, new vertices
//   trace(grid, rays, prim, t);                     // trace kernel contract, RayTracingKernels.h:37-136
//
// Differences a caller sees: errors throw shinobi::Error (the reference
// printed and exited); the grid owns its device buffers (no texture
// binding); FaceSoup::indices is read through grid.indices() after a build.
#pragma once

#include <math.h>
#include <stdint.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../shinobi_grid.h"

// The value types below (vec3f, BBox, Triangle, GridParameters,
// TwoLevelGridCell) are also the device-side types of shinobi/device.hpp:
// compiled by hipcc their members are host and device functions.
#if defined(__HIPCC__)
#define SHINOBI_HD __host__ __device__ inline
#else
#define SHINOBI_HD inline
#endif
// First statement of a function whose binary32 arithmetic must round once
// per operation (no fused multiply-add), whatever -ffp-contract says.
#if defined(__clang__)
#define SHINOBI_EXACT _Pragma("clang fp contract(off)")
#else
#define SHINOBI_EXACT
#endif

namespace shinobi {

class Error : public std::runtime_error {
public:
    Error(sg_status s, const std::string& what)
        : std::runtime_error(what + ": " + sg_status_str(s)), status(s) {}
    sg_status status;
};

inline void check(sg_status s, const char* what) {
    if (s != SG_OK) throw Error(s, what);
}

// src/Core/Algebra.hpp:407-610: per-component arithmetic, dot product
// (x*x' + y*y') + z*z', cross product (operator%), component-wise min / max
// (fminf / fmaxf, as the reference's device code), all in the reference's
// evaluation order; division is IEEE division (the reference's fastDivide)
struct vec3f {
    float x, y, z;
    SHINOBI_HD float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    SHINOBI_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    SHINOBI_HD static vec3f rep(float f) { return vec3f{f, f, f}; }
    SHINOBI_HD vec3f operator+(const vec3f& b) const { SHINOBI_EXACT return vec3f{x + b.x, y + b.y, z + b.z}; }
    SHINOBI_HD vec3f operator-(const vec3f& b) const { SHINOBI_EXACT return vec3f{x - b.x, y - b.y, z - b.z}; }
    SHINOBI_HD vec3f operator*(const vec3f& b) const { SHINOBI_EXACT return vec3f{x * b.x, y * b.y, z * b.z}; }
    SHINOBI_HD vec3f operator/(const vec3f& b) const { SHINOBI_EXACT return vec3f{x / b.x, y / b.y, z / b.z}; }
    SHINOBI_HD vec3f operator*(float f) const { SHINOBI_EXACT return vec3f{x * f, y * f, z * f}; }
    SHINOBI_HD vec3f operator/(float f) const { SHINOBI_EXACT return vec3f{x / f, y / f, z / f}; }
    SHINOBI_HD vec3f operator-() const { return vec3f{-x, -y, -z}; }
    SHINOBI_HD float dot(const vec3f& b) const { SHINOBI_EXACT return x * b.x + y * b.y + z * b.z; }
    SHINOBI_HD vec3f cross(const vec3f& b) const {
        SHINOBI_EXACT
        return vec3f{y * b.z - z * b.y, z * b.x - x * b.z, x * b.y - y * b.x};
    }
    SHINOBI_HD vec3f operator%(const vec3f& b) const { return cross(b); }
    SHINOBI_HD static vec3f min(const vec3f& a, const vec3f& b) {
        return vec3f{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)};
    }
    SHINOBI_HD static vec3f max(const vec3f& a, const vec3f& b) {
        return vec3f{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)};
    }
    SHINOBI_HD static vec3f fastDivide(const vec3f& a, const vec3f& b) { return a / b; }
};
SHINOBI_HD vec3f operator*(float f, const vec3f& v) { SHINOBI_EXACT return vec3f{f * v.x, f * v.y, f * v.z}; }

// src/Primitive/BBox.hpp:34-125
struct BBox {
    vec3f min, max;
    // BBox::fastClip (BBox.hpp:60-75): slab entry / exit distances for rcp(dir)
    SHINOBI_HD void fastClip(const vec3f& org, const vec3f& dirRCP, float& entry, float& exit) const {
        SHINOBI_EXACT
        const vec3f t1 = (min - org) * dirRCP;
        vec3f tMax = (max - org) * dirRCP;
        const vec3f tMin = vec3f::min(t1, tMax);
        tMax = vec3f::max(t1, tMax);
        entry = fmaxf(fmaxf(tMin.x, tMin.y), tMin.z);
        exit = fminf(fminf(tMax.x, tMax.y), tMax.z);
    }
};

// src/Primitive/Triangle.hpp:33-36
struct Triangle {
    vec3f vertices[3];
};

// src/Control/Structure/Grid.hpp:34-62
struct GridParameters {
    int res[3];
    vec3f cellSize;
    vec3f cellSizeRCP;
    BBox bounds;
    SHINOBI_HD vec3f getCellSize() const { return cellSize; }
    SHINOBI_HD vec3f getCellSizeRCP() const { return cellSizeRCP; }
};

// src/Control/Structure/TwoLevelGrid.hpp:34-152: identical 8-byte layout,
// data.x = first leaf, data.y = resX | resY << 8 | resZ << 16 | flags
struct TwoLevelGridCell {
    uint32_t x, y;
    enum : uint32_t { SHIFTX = 0, SHIFTY = 8, SHIFTZ = 16, MASK = 0xFF, LARGENOTEMPTYMASK = 0x40000000u };
    SHINOBI_HD uint32_t notEmpty() const { return y & LARGENOTEMPTYMASK; }
    SHINOBI_HD uint32_t operator[](uint32_t id) const { return (y >> (id * 8)) & MASK; }
    SHINOBI_HD uint32_t getLeafRangeBegin() const { return x; }
};
static_assert(sizeof(TwoLevelGridCell) == 8, "TwoLevelGridCell must stay 8 bytes");

// a stream-ordered context shared by storage, grids and traces
class Context {
public:
    explicit Context(int device = 0) { check(sg_ctx_create(device, &ctx_), "sg_ctx_create"); }
    ~Context() {
        if (ctx_) sg_ctx_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    sg_ctx* get() const { return ctx_; }
    static Context& global() {
        static Context c(0);
        return c;
    }

private:
    sg_ctx* ctx_ = nullptr;
};

// Primitive storage (FaceSoup): triangles as nine SoA float arrays on the device.
class FaceSoup {
public:
    explicit FaceSoup(Context& ctx = Context::global()) : ctx_(&ctx) {}
    ~FaceSoup() {
        if (scene_) sg_scene_destroy(scene_);
    }
    FaceSoup(const FaceSoup&) = delete;
    FaceSoup& operator=(const FaceSoup&) = delete;

    // FaceSoup::upload (FaceSoup.h:171-247)
    void upload(const std::vector<Triangle>& tris, void* stream = nullptr) {
        std::vector<float> soa = to_soa(tris);
        const float* p[9];
        for (int i = 0; i < 9; ++i) p[i] = soa.data() + (size_t)i * tris.size();
        if (scene_) sg_scene_destroy(scene_);
        scene_ = nullptr;
        check(sg_scene_create(ctx_->get(), p, (uint32_t)tris.size(), 0, stream, &scene_), "sg_scene_create");
        n_ = (uint32_t)tris.size();
    }
    // FaceSoup::uploadVertices (FaceSoup.h:249-339): same topology, new positions
    void uploadVertices(const std::vector<Triangle>& tris, void* stream = nullptr) {
        if (!scene_ || tris.size() != n_) throw Error(SG_ERR_INVALID_ARG, "uploadVertices");
        std::vector<float> soa = to_soa(tris);
        const float* p[9];
        for (int i = 0; i < 9; ++i) p[i] = soa.data() + (size_t)i * tris.size();
        check(sg_scene_update(scene_, p, 0, stream), "sg_scene_update");
    }
    uint32_t size() const { return n_; }
    sg_scene* handle() const { return scene_; }
    Context& context() const { return *ctx_; }

private:
    static std::vector<float> to_soa(const std::vector<Triangle>& t) {
        std::vector<float> s(9 * t.size());
        for (size_t i = 0; i < t.size(); ++i)
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) s[(size_t)(3 * v + a) * t.size() + i] = t[i].vertices[v][a];
        return s;
    }
    Context* ctx_;
    sg_scene* scene_ = nullptr;
    uint32_t n_ = 0;
};

// Common part of SimpleGrid / TwoLevelGrid: owns the sg_grid.
class GridBase {
public:
    ~GridBase() {
        if (grid_) sg_grid_destroy(grid_);
    }
    GridBase() = default;
    GridBase(const GridBase&) = delete;
    GridBase& operator=(const GridBase&) = delete;

    int resX = 0, resY = 0, resZ = 0;
    BBox bounds{};

    sg_grid* handle() const { return grid_; }
    sg_grid_info info() const {
        sg_grid_info i;
        check(sg_grid_get_info(grid_, &i), "sg_grid_get_info");
        return i;
    }
    // SimpleGrid::getParameters / TwoLevelGrid::getParameters
    GridParameters getParameters() const {
        sg_grid_info i = info();
        GridParameters p;
        for (int k = 0; k < 3; ++k) {
            p.res[k] = i.res[k];
            p.cellSize[k] = i.cell_size[k];
            p.cellSizeRCP[k] = i.cell_size_rcp[k];
            p.bounds.min[k] = i.bmin[k];
            p.bounds.max[k] = i.bmax[k];
        }
        return p;
    }
    std::vector<uint32_t> download(int which) const {
        sg_grid_info i = info();
        size_t n = 0;
        switch (which) {
            case SG_ARRAY_TOP_KEYS: case SG_ARRAY_TOP_VALS: case SG_ARRAY_PAIR_COUNTS: n = i.P1; break;
            case SG_ARRAY_TOP_RANGES: case SG_ARRAY_TOP_CELLS: n = 2 * (size_t)i.C1; break;
            case SG_ARRAY_LEAF_KEYS: case SG_ARRAY_LEAF_VALS: n = i.P2; break;
            case SG_ARRAY_LEAVES: n = 2 * (size_t)i.C2; break;
            case SG_ARRAY_INDICES: n = i.kind == SG_GRID_TWO_LEVEL ? i.P2 : i.P1; break;
            case SG_ARRAY_TRI_COUNTS: n = i.n_tris; break;
            case SG_ARRAY_LEAF_COUNTS: n = i.C1; break;
            default: throw Error(SG_ERR_INVALID_ARG, "download");
        }
        std::vector<uint32_t> out(n ? n : 1);
        check(sg_grid_download(grid_, which, out.data(), out.size() * 4), "sg_grid_download");
        out.resize(n);
        return out;
    }
    // FaceSoup::indices of the reference: reference index -> triangle id
    std::vector<uint32_t> indices() const { return download(SG_ARRAY_INDICES); }
    // device pointers of the built grid and its scene, for kernels written
    // against shinobi/device.hpp (waits for the grid's stream)
    sg_grid_device_view deviceView() const {
        sg_grid_device_view v;
        check(sg_grid_get_device_view(grid_, &v), "sg_grid_get_device_view");
        return v;
    }

protected:
    void adopt(sg_grid* g) {
        if (grid_) sg_grid_destroy(grid_);
        grid_ = g;
        sg_grid_info i = info();
        resX = i.res[0];
        resY = i.res[1];
        resZ = i.res[2];
        for (int k = 0; k < 3; ++k) {
            bounds.min[k] = i.bmin[k];
            bounds.max[k] = i.bmax[k];
        }
    }
    template <class> friend class FastSimpleGridBuilder;
    template <class> friend class FastTwoLevelGridBuilder;
    sg_grid* grid_ = nullptr;
};

// src/Control/Structure/SimpleGrid.h:41-282
class SimpleGrid : public GridBase {
public:
    typedef uint32_t Cell[2];
    std::vector<uint32_t> cells() const { return download(SG_ARRAY_TOP_RANGES); }
};

// src/Control/Structure/TwoLevelGrid.h:39-358
class TwoLevelGrid : public GridBase {
public:
    typedef TwoLevelGridCell t_Cell;
    uint32_t leavesCount = 0, instancesCount = 0;
    std::vector<TwoLevelGridCell> cells() const {
        std::vector<uint32_t> w = download(SG_ARRAY_TOP_CELLS);
        std::vector<TwoLevelGridCell> c(w.size() / 2);
        std::memcpy(c.data(), w.data(), w.size() * 4);
        return c;
    }
    std::vector<uint32_t> leaves() const { return download(SG_ARRAY_LEAVES); }
};

// Builder concept of src/Control/Algorithm/FastTwoLevelGridBuilder.h:56-740
template <class tPrimitiveStorage>
class FastTwoLevelGridBuilder {
public:
    FastTwoLevelGridBuilder() { sg_build_params_default(&params, SG_GRID_TWO_LEVEL); }
    sg_build_params params;
    void setNumPrimitives(uint32_t n) { num_ = n; }
    void init(TwoLevelGrid& grid, tPrimitiveStorage& storage, const std::vector<Triangle>& tris) {
        storage.upload(tris);
        num_ = storage.size();
        (void)grid;
    }
    void build(TwoLevelGrid& grid, tPrimitiveStorage& storage, void* stream = nullptr) {
        sg_grid* g = nullptr;
        check(sg_build(storage.context().get(), storage.handle(), &params, stream, &g), "sg_build");
        grid.adopt(g);
        sg_grid_info i = grid.info();
        grid.leavesCount = i.C2;
        grid.instancesCount = i.P2;
    }
    void rebuild(TwoLevelGrid& grid, tPrimitiveStorage& storage, void* stream = nullptr) {
        check(sg_rebuild(storage.context().get(), storage.handle(), grid.handle(), stream), "sg_rebuild");
    }

private:
    uint32_t num_ = 0;
};

// Builder concept of src/Control/Algorithm/FastSimpleGridBuilder.h:54-452
template <class tPrimitiveStorage>
class FastSimpleGridBuilder {
public:
    FastSimpleGridBuilder() { sg_build_params_default(&params, SG_GRID_UNIFORM); }
    sg_build_params params;
    void setNumPrimitives(uint32_t n) { num_ = n; }
    void init(SimpleGrid& grid, tPrimitiveStorage& storage, const std::vector<Triangle>& tris) {
        storage.upload(tris);
        num_ = storage.size();
        (void)grid;
    }
    void build(SimpleGrid& grid, tPrimitiveStorage& storage, void* stream = nullptr) {
        sg_grid* g = nullptr;
        check(sg_build(storage.context().get(), storage.handle(), &params, stream, &g), "sg_build");
        grid.adopt(g);
    }
    void rebuild(SimpleGrid& grid, tPrimitiveStorage& storage, void* stream = nullptr) {
        check(sg_rebuild(storage.context().get(), storage.handle(), grid.handle(), stream), "sg_rebuild");
    }

private:
    uint32_t num_ = 0;
};

// The trace kernel contract (src/Engine/RayTracingKernels.h:37-136) with
// SimpleIntersector and the grid's traverser: SoA device rays in, device
// triangle ids (0xFFFFFFFF = miss) and t out.
inline void trace(const GridBase& grid, Context& ctx, const sg_rays& rays, uint32_t n, uint32_t* d_prim,
                  float* d_t, void* stream = nullptr) {
    check(sg_trace_closest(ctx.get(), grid.handle(), &rays, n, nullptr, d_prim, d_t, nullptr, stream),
          "sg_trace_closest");
}

// The shadow-ray contract (TwoLevelGridShadowTraverser /
// UniformGridShadowTraverser): segments in, occluded where t < 0.9999.
inline void traceShadow(const GridBase& grid, Context& ctx, const sg_rays& rays, uint32_t n, uint32_t* d_prim,
                        float* d_t, void* stream = nullptr) {
    check(sg_trace_shadow(ctx.get(), grid.handle(), &rays, n, nullptr, d_prim, d_t, nullptr, stream),
          "sg_trace_shadow");
}

}  // namespace shinobi
