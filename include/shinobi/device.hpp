// shinobi/device.hpp -- the reference's device-side concepts for kernels
// compiled with hipcc for gfx950, so a renderer kernel written against
// Shinobi's Storage / Intersector / Traverser / RayGenerator / OutputBuffer
// concepts (src/Engine/RayTracingTypes.h:84-128) compiles against this
// library's grids:
//
//   FaceSoupView                     storage concept of FaceSoup
//                                    (src/Control/Structure/FaceSoup.h:41-165):
//                                    indices, operator[](refIdx), operator()(primId),
//                                    getMaterialId
//   SimpleIntersector<Storage>       (src/Intersector/SimpleIntersector.hpp:290-337)
//   TwoLevelGridTraverser<Control, Storage, Intersector>
//                                    (src/Control/Algorithm/TwoLevelGridTraverser.h:423-673)
//   UniformGridTraverser<Control, Storage, Intersector>
//                                    (src/Control/Algorithm/SimpleGridTraverser.h:242-403)
//   Camera, RegularPixelSampler<1,1>, RegularPrimaryRayGenerator
//                                    (src/Primitive/Camera.h:36-70,
//                                    src/Primitive/PixelSampler.h:34-54,
//                                    src/Engine/RayGenerators.h:36-61)
//   SimpleRayBuffer                  (src/Engine/RayBuffers.h:38-140)
//
// Where the reference reads the grid through bound textures and a
// __constant__ GridParameters (src/Control/Structure/TwoLevelGrid.h:36-37),
// the control structure here carries the grid's device arrays as well:
// TwoLevelGridControl / UniformGridControl are GridParameters plus the cell
// (and leaf) arrays, filled from sg_grid_get_device_view.  Textures' clamp
// addressing (top cells) and zero out-of-range reads (leaves) are explicit.
// The traversers keep the reference's signature and its contract that
// rcp(dir) sits at ((vec3f*)sharedMem)[threadId1D32()].
//
// Arithmetic: every binary32 operation rounds once, in the reference's
// order.  The header turns floating-point contraction off in its own
// functions (SHINOBI_EXACT, a clang fp pragma), which hipcc's default
// -ffp-contract=fast-honor-pragmas (and =on, =off) respects; an explicit
// -ffp-contract=fast ignores such pragmas and must not be used.  Division
// must be IEEE division: hipcc's default -fhip-fp32-correctly-rounded-
// divide-sqrt, without -ffast-math.  Results
// are then bit-identical to sg_trace_closest: the library's kernel runs the
// same per-ray sequence of operations, reorganised for whole waves.  Like the
// library's traversal, each level's walk stops after 2 (resX + resY + resZ)
// + 16 steps, which ends only walks the reference would never end (NaN or
// infinite rays), and a miss leaves bestHit at 0xFFFFFFFF.
#pragma once

#include <float.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "grid.hpp"

namespace shinobi {

constexpr float EPS = 0.00001f;                 // src/CUDAStdAfx.h:96
constexpr uint32_t NO_HIT = 0xFFFFFFFFu;

// the thread's slot in a block's shared arrays (src/CUDAStdAfx.h:204-206,
// threadIdx.x + threadIdx.y * 32 for the reference's 32-wide blocks)
__device__ inline uint32_t threadId1D32() { return threadIdx.x + threadIdx.y * blockDim.x; }

// --------------------------------------------------------------- storage
// FaceSoup as the kernels see it: indices maps a reference (a position in
// the grid's cell ranges) to a triangle; the triangles are the scene's SoA
// vertex arrays; materials (optional) holds one material id per triangle.
struct FaceSoupView {
    const uint32_t* indices;
    const float* v[9];
    const int* materials;

    // FaceSoup::operator()(primId), FaceSoup.h:110-157
    __device__ Triangle operator()(uint32_t primId) const {
        Triangle t;
        for (int k = 0; k < 3; ++k) t.vertices[k] = vec3f{v[3 * k][primId], v[3 * k + 1][primId], v[3 * k + 2][primId]};
        return t;
    }
    // FaceSoup::operator[](refIdx), FaceSoup.h:54-108: the triangle of reference refIdx
    __device__ Triangle operator[](uint32_t refIdx) const { return (*this)(indices[refIdx]); }
    // FaceSoup::getMaterialId, FaceSoup.h:164-168
    __device__ int getMaterialId(uint32_t primId) const { return materials ? materials[primId] : 0; }
};

inline FaceSoupView faceSoupView(const sg_grid_device_view& g, const int* materials = nullptr) {
    FaceSoupView s;
    s.indices = g.indices;
    for (int i = 0; i < 9; ++i) s.v[i] = g.tri[i];
    s.materials = materials;
    return s;
}

// ------------------------------------------------------ control structures
inline GridParameters gridParameters(const sg_grid_device_view& g) {
    GridParameters p;
    for (int k = 0; k < 3; ++k) {
        p.res[k] = g.res[k];
        p.cellSize[k] = g.cell_size[k];
        p.cellSizeRCP[k] = g.cell_size_rcp[k];
        p.bounds.min[k] = g.bmin[k];
        p.bounds.max[k] = g.bmax[k];
    }
    return p;
}

struct TwoLevelGridControl : GridParameters {
    const TwoLevelGridCell* cells;   // res[0] * res[1] * res[2] top cells, x fastest
    const uint2* leaves;             // numLeaves leaf ranges
    uint32_t numLeaves;
};
inline TwoLevelGridControl twoLevelGridControl(const sg_grid_device_view& g) {
    TwoLevelGridControl c;
    static_cast<GridParameters&>(c) = gridParameters(g);
    c.cells = reinterpret_cast<const TwoLevelGridCell*>(g.cells);
    c.leaves = reinterpret_cast<const uint2*>(g.leaves);
    c.numLeaves = g.n_leaves;
    return c;
}

struct UniformGridControl : GridParameters {
    const uint2* cells;              // res[0] * res[1] * res[2] reference ranges, x fastest
};
inline UniformGridControl uniformGridControl(const sg_grid_device_view& g) {
    UniformGridControl c;
    static_cast<GridParameters&>(c) = gridParameters(g);
    c.cells = reinterpret_cast<const uint2*>(g.cells);
    return c;
}

namespace detail {
// tex3D's clamp addressing at integral coordinates
__device__ inline uint32_t clampCell(float c, int res) { return (uint32_t)fminf(fmaxf(c, 0.f), (float)(res - 1)); }
__device__ inline uint32_t cellIndex(const GridParameters& p, const vec3f& c) {
    return clampCell(c.x, p.res[0]) + clampCell(c.y, p.res[1]) * (uint32_t)p.res[0] +
           clampCell(c.z, p.res[2]) * (uint32_t)p.res[0] * (uint32_t)p.res[1];
}
// float -> uint32 as the GPU conversion saturates (NaN and negatives -> 0)
__device__ inline uint32_t toIndex(float f) {
    if (!(f > 0.f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}
// MIN_DIMENSION (TwoLevelGridTraverser.h:633-634)
__device__ inline int minDimension(const vec3f& t) { return (t.x < t.y) ? ((t.x < t.z) ? 0 : 2) : ((t.y < t.z) ? 1 : 2); }
}  // namespace detail

// ------------------------------------------------------------ intersector
// SimpleIntersector with the Moller-Trumbore test of ray directions given
// as reciprocals (SimpleIntersector.hpp:50-89, entry point :326-337): the
// direction is recomputed as 1 / rcp, a hit needs alpha >= 0, beta >= 0,
// alpha + beta <= 1 and EPS < t < the current t (strict: the first of equal
// distances stays), and bestHit becomes the reference index.
template <class tStorage>
class SimpleIntersector {
public:
    static const uint32_t SHAREDMEMSIZE = 0u;
    static const uint32_t GLOBALMEMSIZE = 0u;

    __device__ void operator()(const vec3f& aRayOrg, const vec3f& aRayDirRCP, float& oRayT, const uint2& aIdRange,
                               const tStorage& aScene, uint32_t* /*aSharedMem*/, uint32_t& oBestHit) const {
        SHINOBI_EXACT
        const vec3f rayDir = vec3f::fastDivide(vec3f::rep(1.f), aRayDirRCP);
        for (uint32_t it = aIdRange.x; it != aIdRange.y; ++it) {
            const Triangle tri = aScene[it];
            const vec3f org = tri.vertices[0];
            const vec3f edge1 = tri.vertices[1] - org;
            const vec3f edge2 = tri.vertices[2] - org;
            const vec3f pvec = rayDir % edge2;
            const float detRCP = 1.f / edge1.dot(pvec);
            vec3f tvec = aRayOrg - org;
            const float alpha = detRCP * tvec.dot(pvec);
            tvec = tvec % edge1;
            const float beta = detRCP * tvec.dot(rayDir);
            const float dist = detRCP * edge2.dot(tvec);
            if (alpha >= 0.f && beta >= 0.f && alpha + beta <= 1.f && dist > EPS && dist < oRayT) {
                oRayT = dist;
                oBestHit = it;
            }
        }
    }
};

// -------------------------------------------------------------- traversers
// TwoLevelGridTraverser::traverse (TwoLevelGridTraverser.h:445-584) with
// traverseInit (:586-626), traverseMacroCell (:628-649) and traverseCell
// (:651-671), one ray per thread.
template <class tControlStructure, class tStorageStructure, class tIntersector>
class TwoLevelGridTraverser {
public:
    static const uint32_t SHAREDMEMSIZE = 3u;   // floats per thread: rcp(dir)

    __device__ void traverse(const vec3f& aRayOrg, float& oRayT, uint32_t& oBestHit,
                             const tControlStructure& aGrid, const tStorageStructure& aScene,
                             uint32_t* aSharedMem) const {
        SHINOBI_EXACT
        const vec3f& rcp = ((vec3f*)aSharedMem)[threadId1D32()];
        oBestHit = NO_HIT;
        bool flag1 = oRayT > 0.f;
        vec3f tMax{0.f, 0.f, 0.f}, cellId{0.f, 0.f, 0.f};
        float tEntry, tExit;
        aGrid.bounds.fastClip(aRayOrg, rcp, tEntry, tExit);
        if (tExit <= tEntry || tExit < 0.f) {
            flag1 = false;
        } else {
            tEntry = fmaxf(0.f, tEntry);
            const vec3f entryPt = aRayOrg + vec3f::fastDivide(vec3f::rep(tEntry + EPS), rcp);
            cellId = (entryPt - aGrid.bounds.min) * aGrid.getCellSizeRCP();
            cellId = vec3f{floorf(cellId.x), floorf(cellId.y), floorf(cellId.z)};
            const vec3f up{rcp.x > 0.f ? 1.f : 0.f, rcp.y > 0.f ? 1.f : 0.f, rcp.z > 0.f ? 1.f : 0.f};
            tMax = ((cellId + up) * aGrid.getCellSize() + aGrid.bounds.min - aRayOrg) * rcp;
            if (!((cellId.x != (float)(rcp.x > 0.f ? aGrid.res[0] : -1)) &&
                  (cellId.y != (float)(rcp.y > 0.f ? aGrid.res[1] : -1)) &&
                  (cellId.z != (float)(rcp.z > 0.f ? aGrid.res[2] : -1))))
                flag1 = false;
        }
        const uint32_t cap = 2u * (uint32_t)(aGrid.res[0] + aGrid.res[1] + aGrid.res[2]) + 16u;
        uint32_t guard = 0;
        tIntersector intersect;
        while (flag1 && guard++ < cap) {
            const TwoLevelGridCell cell = aGrid.cells[detail::cellIndex(aGrid, cellId)];
            if (cell.notEmpty()) {
                // sub-grid set-up (:490-528)
                const vec3f entryPoint = aRayOrg + vec3f::fastDivide(vec3f::rep(tEntry), rcp);
                const vec3f cellRes{(float)cell[0], (float)cell[1], (float)cell[2]};
                const vec3f subCellSizeRCP = aGrid.getCellSizeRCP() * cellRes;
                const vec3f minBound = cellId * aGrid.getCellSize() + aGrid.bounds.min;
                vec3f subId{floorf((entryPoint - minBound).x * subCellSizeRCP.x),
                            floorf((entryPoint - minBound).y * subCellSizeRCP.y),
                            floorf((entryPoint - minBound).z * subCellSizeRCP.z)};
                const vec3f subCellSize{1.f / subCellSizeRCP.x, 1.f / subCellSizeRCP.y, 1.f / subCellSizeRCP.z};
                subId = vec3f::min(cellRes - vec3f::rep(1.f), vec3f::max(subId, vec3f::rep(0.f)));
                vec3f subMax;
                subMax.x = ((subId.x + (float)(rcp.x > 0.f ? 1 : 0)) * subCellSize.x + minBound.x - aRayOrg.x) * rcp.x;
                subMax.y = ((subId.y + (float)(rcp.y > 0.f ? 1 : 0)) * subCellSize.y + minBound.y - aRayOrg.y) * rcp.y;
                subMax.z = ((subId.z + (float)(rcp.z > 0.f ? 1 : 0)) * subCellSize.z + minBound.z - aRayOrg.z) * rcp.z;
                const uint32_t icap = 2u * (cell[0] + cell[1] + cell[2]) + 16u;
                uint32_t iguard = 0;
                bool flag2 = true;
                while (flag2 && iguard++ < icap) {
                    // leaf fetch (:537-539); reads past the leaves return an empty range
                    const uint32_t li = cell.getLeafRangeBegin() +
                                        detail::toIndex(subId.x + (float)cell[0] * subId.y +
                                                        (float)cell[0] * (float)cell[1] * subId.z);
                    const uint2 range = li < aGrid.numLeaves ? aGrid.leaves[li] : make_uint2(0u, 0u);
                    intersect(aRayOrg, rcp, oRayT, range, aScene, aSharedMem, oBestHit);
                    if (!(oRayT > subMax.x || oRayT > subMax.y || oRayT > subMax.z)) break;
                    // traverseCell (:651-671)
                    const int d = detail::minDimension(subMax);
                    subId[d] += rcp[d] > 0.f ? 1.f : -1.f;
                    subMax[d] += subCellSize[d] * fabsf(rcp[d]);
                    flag2 = !(fabsf(subId[d] - cellRes[d]) < 0.1f || fabsf(subId[d] + 1.f) < 0.1f);
                }
            }
            if (!(oRayT > tMax.x || oRayT > tMax.y || oRayT > tMax.z)) break;
            // traverseMacroCell (:628-649)
            const int d = detail::minDimension(tMax);
            tEntry = tMax[d];
            cellId[d] += rcp[d] > 0.f ? 1.f : -1.f;
            tMax[d] += aGrid.cellSize[d] * fabsf(rcp[d]);
            flag1 = !(fabsf(cellId[d] - (float)aGrid.res[d]) < 0.1f || fabsf(cellId[d] + 1.f) < 0.1f);
        }
    }
};

// UniformGridTraverser::traverse (SimpleGridTraverser.h:274-401): the entry
// point steps EPS inside only for tEntry >= 0, cells are IEEE-divided
// positions times cellSizeRCP.
template <class tControlStructure, class tStorageStructure, class tIntersector>
class UniformGridTraverser {
public:
    static const uint32_t SHAREDMEMSIZE = 3u;

    __device__ void traverse(const vec3f& aRayOrg, float& oRayT, uint32_t& oBestHit,
                             const tControlStructure& aGrid, const tStorageStructure& aScene,
                             uint32_t* aSharedMem) const {
        SHINOBI_EXACT
        const vec3f& rcp = ((vec3f*)aSharedMem)[threadId1D32()];
        oBestHit = NO_HIT;
        bool flag = false;
        vec3f tMax{0.f, 0.f, 0.f}, cellId{0.f, 0.f, 0.f};
        if (oRayT > 0.f) {
            float tEntry, tExit;
            aGrid.bounds.fastClip(aRayOrg, rcp, tEntry, tExit);
            if (!(tExit <= tEntry || tExit < 0.f)) {
                const vec3f pt = (tEntry >= 0.f) ? aRayOrg + vec3f::fastDivide(vec3f::rep(tEntry + EPS), rcp) : aRayOrg;
                cellId = (pt - aGrid.bounds.min) * aGrid.getCellSizeRCP();
                cellId = vec3f{floorf(cellId.x), floorf(cellId.y), floorf(cellId.z)};
                const vec3f up{rcp.x > 0.f ? 1.f : 0.f, rcp.y > 0.f ? 1.f : 0.f, rcp.z > 0.f ? 1.f : 0.f};
                tMax = ((cellId + up) * aGrid.getCellSize() + aGrid.bounds.min - aRayOrg) * rcp;
                flag = (cellId.x != (float)(rcp.x > 0.f ? aGrid.res[0] : -1)) &&
                       (cellId.y != (float)(rcp.y > 0.f ? aGrid.res[1] : -1)) &&
                       (cellId.z != (float)(rcp.z > 0.f ? aGrid.res[2] : -1));
            }
        }
        const uint32_t cap = 2u * (uint32_t)(aGrid.res[0] + aGrid.res[1] + aGrid.res[2]) + 16u;
        uint32_t guard = 0;
        tIntersector intersect;
        while (flag && guard++ < cap) {
            const uint2 range = aGrid.cells[detail::cellIndex(aGrid, cellId)];
            intersect(aRayOrg, rcp, oRayT, range, aScene, aSharedMem, oBestHit);
            if (!(oRayT > tMax.x || oRayT > tMax.y || oRayT > tMax.z)) break;
            const int d = detail::minDimension(tMax);
            cellId[d] += rcp[d] > 0.f ? 1.f : -1.f;
            tMax[d] += aGrid.cellSize[d] * fabsf(rcp[d]);
            flag = !(fabsf(cellId[d] - (float)aGrid.res[d]) < 0.1f || fabsf(cellId[d] + 1.f) < 0.1f);
        }
    }
};

// ------------------------------------------------------- ray generation
// Camera after Camera::init (Camera.h:41-60); init runs on the host through
// the library (sg_pinhole_init), getDirection on the device.
struct Camera {
    vec3f mCenter, mTopLeft, mStepX, mStepY;

    static Camera fromPinhole(const sg_pinhole& p) {
        Camera c;
        for (int k = 0; k < 3; ++k) {
            c.mCenter[k] = p.position[k];
            c.mTopLeft[k] = p.top_left[k];
            c.mStepX[k] = p.step_x[k];
            c.mStepY[k] = p.step_y[k];
        }
        return c;
    }
    SHINOBI_HD vec3f getPosition() const { return mCenter; }
    // Camera::getDirection (Camera.h:67-70)
    SHINOBI_HD vec3f getDirection(float aX, float aY) const { SHINOBI_EXACT return mTopLeft + aX * mStepX + aY * mStepY; }
};

// RegularPixelSampler (PixelSampler.h:34-54)
template <int taNumSamplesX, int taNumSamplesY>
struct RegularPixelSampler {
    float resX, resY;
    SHINOBI_HD void getPixelCoords(float aPixelId, float aSampleId, float& oPixelX, float& oPixelY) const {
        SHINOBI_EXACT
        float tmp = aSampleId / (float)taNumSamplesX;
        oPixelY = (truncf(tmp) + 0.5f) / (float)taNumSamplesY;
        oPixelX = tmp - truncf(tmp) + 0.5f / (float)taNumSamplesX;
        tmp = aPixelId / resX;
        oPixelY += truncf(tmp);
        oPixelX += (tmp - truncf(tmp)) * resX;
    }
};

// RegularPrimaryRayGenerator<taSafe> (RayGenerators.h:36-61); the camera,
// sampler and image id are members instead of __constant__ symbols
template <bool taSafe>
struct RegularPrimaryRayGenerator {
    Camera camera;
    RegularPixelSampler<1, 1> sampler;
    uint32_t imageId;

    __device__ float operator()(vec3f& oRayOrg, vec3f& oRayDir, uint32_t aRayId, uint32_t aNumRays) const {
        oRayOrg = camera.getPosition();
        float screenX, screenY;
        const uint32_t id = taSafe ? (aRayId < aNumRays - 1u ? aRayId : aNumRays - 1u) : aRayId;
        sampler.getPixelCoords((float)id, (float)imageId, screenX, screenY);
        oRayDir = camera.getDirection(screenX, screenY);
        return FLT_MAX;
    }
};

// ---------------------------------------------------------- output buffer
// SimpleRayBuffer (RayBuffers.h:38-140): 8 words per ray, SoA at stride numRays
class SimpleRayBuffer {
    void* mMemoryPtr;

public:
    SHINOBI_HD explicit SimpleRayBuffer(void* aMemPtr) : mMemoryPtr(aMemPtr) {}
    SHINOBI_HD void* getData() { return mMemoryPtr; }

    __device__ void store(const vec3f& aRayOrg, const vec3f& aRayDir, const float aRayT, const uint32_t aBestHit,
                          const uint32_t aRayId, const uint32_t aNumRays) {
        SHINOBI_EXACT
        const vec3f hitPoint = aRayOrg + (aRayT - EPS) * aRayDir;
        float* out = (float*)mMemoryPtr + aRayId;
        const float w[7] = {hitPoint.x, hitPoint.y, hitPoint.z, aRayDir.x, aRayDir.y, aRayDir.z, aRayT};
        for (int k = 0; k < 7; ++k) out[(size_t)k * aNumRays] = w[k];
        *((uint32_t*)out + (size_t)7 * aNumRays) = aBestHit;
    }
    __device__ void load(vec3f& oRayOrg, vec3f& oRayDir, float& oRayT, uint32_t& oBestHit, const uint32_t aRayId,
                         const uint32_t aNumRays) const {
        const float* in = (const float*)mMemoryPtr + aRayId;
        oRayOrg = vec3f{in[0], in[(size_t)aNumRays], in[(size_t)2 * aNumRays]};
        oRayDir = vec3f{in[(size_t)3 * aNumRays], in[(size_t)4 * aNumRays], in[(size_t)5 * aNumRays]};
        oRayT = in[(size_t)6 * aNumRays];
        oBestHit = *((const uint32_t*)in + (size_t)7 * aNumRays);
    }
};

}  // namespace shinobi
