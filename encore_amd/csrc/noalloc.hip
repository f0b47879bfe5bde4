// noalloc.hip -- the allocation-free closest-hit entry point.
//
// rq_intersect_closest allocates (and keeps, per stream) the work counter and
// the per-lane spill stacks of its persistent launch.  Callers that cannot
// allocate device memory -- code under hipGraph capture, tight serving loops,
// harnesses that forbid it -- use rq_intersect_closest_static instead: the
// same kernels, with that scratch in the file-scope array below, which every
// GPU the library runs on holds until the process ends.
//
// Host code never passes the array's name to a kernel or a copy: in host code
// it names a host-side stand-in.  bind_static_scratch, called by
// rq_device_open with the device current, asks the runtime for this GPU's
// address of the array once, and every use goes through that address.
#include <hip/hip_runtime.h>

#include "rq_internal.h"

namespace rq {

// Work counter (first 256 bytes) and spill stacks of the static entry point.
static __device__ __align__(256) uint32_t scratch[kStaticScratchBytes / sizeof(uint32_t)];

void bind_static_scratch(rq_device* dev) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        throw Error(RQ_ERR_DEVICE, std::string("hipGetSymbolAddress(static scratch): ") + hipGetErrorString(e));
    }
    dev->static_scratch = static_cast<uint32_t*>(p);
}

}  // namespace rq

extern "C" rq_status rq_intersect_closest_static(rq_scene* scene, const rq_ray* d_rays, uint32_t n, rq_hit* d_hits,
                                                 void* stream) {
    struct Ctx {
        rq_scene* scene;
        const rq_ray* rays;
        uint32_t n;
        rq_hit* hits;
        hipStream_t s;
    } c{scene, d_rays, n, d_hits, (hipStream_t)stream};
    return rq::guard_call(
        [](void* p) {
            Ctx& c = *static_cast<Ctx*>(p);
            if (c.scene == nullptr) throw rq::Error(RQ_ERR_INVALID_ARG, "scene is null");
            if (c.n != 0 && (c.rays == nullptr || c.hits == nullptr))
                throw rq::Error(RQ_ERR_INVALID_ARG, "null ray or hit buffer");
            if (c.scene->dev->static_scratch == nullptr)
                throw rq::Error(RQ_ERR_DEVICE, "static scratch not bound (device not opened with rq_device_open)");
            rq::trace_closest_in(c.scene, c.rays, c.n, c.hits, c.s, c.scene->dev->static_scratch,
                                 rq::kStaticScratchBytes);
        },
        &c);
}
