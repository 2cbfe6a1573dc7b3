// pt_trace_device_noalloc: the ray-batch traversal of pt_trace_device for
// callers that cannot allocate device memory (code under stream capture, tight
// serving loops). The traversal kernels need two buffers beyond the caller's
// rays and hits: the work counter the persistent waves take rays from, and the
// per-lane spill area of the traversal stack. Here both live in one
// file-scope device array, so a call allocates nothing.
//
// The array's name is a host-side stand-in in host code: its device address
// is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to kernels.
#include <mutex>

#include "api_internal.hpp"

namespace {

// Scratch per device: the work counter in the first 256 bytes, the traversal
// stack's spill area after it. 64 MiB holds the full traversal grid of a
// tree up to 8 levels deep (3 * 8 + 1 + 4 = 29 entries a lane) and a
// reduced grid for deeper trees; see noalloc_grid.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kSpillOffset = 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (the current device).
void* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw pt::status_error(PT_ERR_LIMIT, "pt_trace_device_noalloc: device number beyond 63");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw pt::status_error(PT_ERR_DEVICE, "pt_trace_device_noalloc: no address for the scratch array");
        g_scratch[device] = p;
    }
    return g_scratch[device];
}

// Clears the work counter with a kernel (no memset node if the stream is
// being captured).
__global__ void k_clear_counter(uint32_t* counter) {
    if (threadIdx.x == 0) *counter = 0;
}

}  // namespace

namespace ptapi {

// Traversal blocks that fit the scratch spill area for this scene: the
// pt_trace_device grid, or fewer for scenes whose trees need deeper stacks.
// 0 if not even one block per compute unit fits.
int noalloc_grid(pt_scene* s, int S) {
    const size_t lane_bytes = size_t(s->max_stack) * sizeof(uint2);
    const size_t lanes = (kScratchBytes - kSpillOffset) / lane_bytes;
    const int fit = static_cast<int>(lanes / ptd::TRACE_BLOCK);
    const int grid = std::min(trace_grid_for(s, S), fit);
    return grid >= s->cus ? grid : 0;
}

}  // namespace ptapi

extern "C" {

pt_status pt_trace_device_noalloc(pt_scene* s, const void* d_rays, size_t n, int any_hit, void* d_hits,
                                  void* stream) {
    if (!s) return pt::fail(PT_ERR_INVALID, "pt_trace_device_noalloc: null scene");
    if (n == 0) return PT_OK;
    if (!d_rays || !d_hits) return pt::fail(PT_ERR_INVALID, "pt_trace_device_noalloc: null buffer");
    if (n >= (size_t(1) << 32) - 64) return pt::fail(PT_ERR_LIMIT, "pt_trace_device_noalloc: more than 2^32 rays");
    PT_GUARD({
        int dev = -1;
        HIPC(hipGetDevice(&dev));
        if (dev != s->device)
            throw pt::status_error(PT_ERR_INVALID, "pt_trace_device_noalloc: the scene's device is not current");
        const int S = ptapi::stack_entries_for(s, 0);
        const int grid = ptapi::noalloc_grid(s, S);
        if (grid == 0)
            throw pt::status_error(PT_ERR_LIMIT, "pt_trace_device_noalloc: the scene's trees need a traversal stack "
                                                 "deeper than the 64 MiB scratch holds for one block per compute unit");
        unsigned char* base = static_cast<unsigned char*>(scratch_address(dev));
        uint32_t* work = reinterpret_cast<uint32_t*>(base);
        uint2* spill = reinterpret_cast<uint2*>(base + kSpillOffset);
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
        k_clear_counter<<<1, 64, 0, st>>>(work);
        HIPC(hipGetLastError());
        HIPC(ptd::launch_trace(s->view, static_cast<const float4*>(d_rays), static_cast<uint32_t>(n),
                               static_cast<ptd::Hit*>(d_hits), any_hit != 0, spill, work, nullptr, S, grid, st));
        if (!stream) HIPC(hipStreamSynchronize(st));
    });
}

}  // extern "C"
