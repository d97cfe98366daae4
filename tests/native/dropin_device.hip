// A renderer's trace kernel written against the reference's device concepts
// (the shape of trace<> in src/Engine/RayTracingKernels.h:37-136: a ray
// generator, a traverser templated on control structure, storage and
// intersector, rcp(dir) kept in shared memory, an output buffer), compiled
// against include/shinobi/device.hpp only.  Test code: it checks that the
// drop-in concepts reproduce the library's traversal bit for bit.
#include <shinobi/device.hpp>

using namespace shinobi;

namespace {

// rays from SoA device arrays (the library's ray layout) as a RayGenerator
struct SoARayGenerator {
    const float* a[7];
    __device__ float operator()(vec3f& oRayOrg, vec3f& oRayDir, uint32_t aRayId, uint32_t /*aNumRays*/) const {
        oRayOrg = vec3f{a[0][aRayId], a[1][aRayId], a[2][aRayId]};
        oRayDir = vec3f{a[3][aRayId], a[4][aRayId], a[5][aRayId]};
        return a[6][aRayId];
    }
};

constexpr uint32_t kThreads = 64;

template <class tControlStructure, template <class, class, class> class tTraverser, class tRayGenerator>
__global__ __launch_bounds__(kThreads) void trace(FaceSoupView aStorage, tRayGenerator aRayGenerator,
                                                  SimpleRayBuffer oOutputBuffer, tControlStructure aGrid,
                                                  uint32_t aNumRays) {
    typedef SimpleIntersector<FaceSoupView> t_Intersector;
    typedef tTraverser<tControlStructure, FaceSoupView, t_Intersector> t_Traverser;
    __shared__ uint32_t sharedMem[kThreads * t_Traverser::SHAREDMEMSIZE];
    for (uint32_t myRayIndex = blockIdx.x * blockDim.x + threadIdx.x; myRayIndex < aNumRays;
         myRayIndex += gridDim.x * blockDim.x) {
        t_Traverser traverser;
        vec3f rayOrg;
        vec3f& rayDir = ((vec3f*)sharedMem)[threadId1D32()];
        float rayT = aRayGenerator(rayOrg, rayDir, myRayIndex, aNumRays);
        rayDir.x = 1.f / rayDir.x;
        rayDir.y = 1.f / rayDir.y;
        rayDir.z = 1.f / rayDir.z;
        uint32_t bestHit;
        traverser.traverse(rayOrg, rayT, bestHit, aGrid, aStorage, sharedMem);
        rayDir.x = 1.f / rayDir.x;
        rayDir.y = 1.f / rayDir.y;
        rayDir.z = 1.f / rayDir.z;
        oOutputBuffer.store(rayOrg, rayDir, rayT, bestHit, myRayIndex, aNumRays);
    }
}

template <class tRayGenerator>
int launch(const sg_grid_device_view* g, const tRayGenerator& gen, uint32_t n, float* d_raybuf) {
    if (!g || !d_raybuf || n == 0) return 1;
    const dim3 grid((n + kThreads - 1) / kThreads), block(kThreads);
    const FaceSoupView storage = faceSoupView(*g);
    SimpleRayBuffer out(d_raybuf);
    if (g->kind == SG_GRID_TWO_LEVEL) {
        hipLaunchKernelGGL((trace<TwoLevelGridControl, TwoLevelGridTraverser, tRayGenerator>), grid, block, 0, 0,
                           storage, gen, out, twoLevelGridControl(*g), n);
    } else {
        hipLaunchKernelGGL((trace<UniformGridControl, UniformGridTraverser, tRayGenerator>), grid, block, 0, 0,
                           storage, gen, out, uniformGridControl(*g), n);
    }
    if (hipGetLastError() != hipSuccess) return 4;
    return hipDeviceSynchronize() == hipSuccess ? 0 : 4;
}

}  // namespace

extern "C" {

// SoA device rays (ox oy oz dx dy dz tmax), SimpleRayBuffer out (8 n floats)
int dropin_trace_rays(const sg_grid_device_view* g, const float* const rays7[7], uint32_t n, float* d_raybuf) {
    SoARayGenerator gen;
    for (int k = 0; k < 7; ++k) gen.a[k] = rays7[k];
    return launch(g, gen, n, d_raybuf);
}

// camera rays from RegularPrimaryRayGenerator, SimpleRayBuffer out
int dropin_trace_camera(const sg_grid_device_view* g, const sg_pinhole* cam, float* d_raybuf) {
    if (!cam) return 1;
    RegularPrimaryRayGenerator<true> gen;
    gen.camera = Camera::fromPinhole(*cam);
    gen.sampler.resX = (float)cam->res_x;
    gen.sampler.resY = (float)cam->res_y;
    gen.imageId = 0;
    return launch(g, gen, cam->res_x * cam->res_y, d_raybuf);
}

}  // extern "C"
