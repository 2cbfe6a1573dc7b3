// Persistent traversal kernels over a batch of rays (pt_trace): each wave
// takes 64 rays at a time from a shared counter until the batch is done.
#include "kernels.hpp"
#include "traverse.hpp"

namespace ptd {

template <int S, bool ANY, bool COUNT, bool INST>
__global__ __launch_bounds__(TRACE_BLOCK)
__attribute__((amdgpu_waves_per_eu(trace_occupancy<S, INST>::waves))) void k_trace(SceneView sc, const float4* __restrict__ rays, uint32_t n,
                                                        Hit* __restrict__ hits, uint2* spill, uint32_t* work,
                                                        unsigned long long* counts) {
    __shared__ uint2 lds[S * TRACE_BLOCK];
    auto st = lane_stack<S, TRACE_BLOCK>(lds, spill);
    const uint32_t lane = threadIdx.x & 63;
    Visits vis;
    const uint32_t groups = fetch_groups(n);
    while (true) {
        const uint32_t base = wave_fetch(work, 64 * groups);
        if (base >= n) break;
        for (uint32_t g = 0; g < groups; ++g) {
            const uint32_t q = base + 64 * g + lane;
            if (q >= n) break;
            const float4 a = rays[2 * q], b = rays[2 * q + 1];
            Hit h;
            if (ANY) {
                const bool blocked = trace_ray<true, COUNT, INST>(sc, xyz(a), xyz(b), a.w, b.w, h, st, &vis);
                h.t = 0;
                h.u = h.v = 0;
                h.kind = blocked ? 1u : 0u;
                h.prim = 0;
                h.inst = NONE;
            } else {
                trace_ray<false, COUNT, INST>(sc, xyz(a), xyz(b), a.w, b.w, h, st, &vis);
            }
            hits[q] = h;
        }
    }
    if (COUNT) add_visits(counts, vis);
}

template <int S, bool INST>
static void launch_si(const SceneView& sc, const float4* rays, uint32_t n, Hit* hits, bool any_hit, uint2* spill,
                      uint32_t* work, unsigned long long* counts, int grid, hipStream_t stream) {
    if (counts) {
        if (any_hit) k_trace<S, true, true, INST><<<grid, TRACE_BLOCK, 0, stream>>>(sc, rays, n, hits, spill, work, counts);
        else k_trace<S, false, true, INST><<<grid, TRACE_BLOCK, 0, stream>>>(sc, rays, n, hits, spill, work, counts);
    } else {
        if (any_hit) k_trace<S, true, false, INST><<<grid, TRACE_BLOCK, 0, stream>>>(sc, rays, n, hits, spill, work, counts);
        else k_trace<S, false, false, INST><<<grid, TRACE_BLOCK, 0, stream>>>(sc, rays, n, hits, spill, work, counts);
    }
}

template <int S>
static hipError_t launch_s(const SceneView& sc, const float4* rays, uint32_t n, Hit* hits, bool any_hit, uint2* spill,
                           uint32_t* work, unsigned long long* counts, int grid, hipStream_t stream) {
    if (sc.inst_root != NONE) launch_si<S, true>(sc, rays, n, hits, any_hit, spill, work, counts, grid, stream);
    else launch_si<S, false>(sc, rays, n, hits, any_hit, spill, work, counts, grid, stream);
    return hipGetLastError();
}

hipError_t launch_trace(const SceneView& sc, const float4* rays, uint32_t n, Hit* hits, bool any_hit, uint2* spill,
                        uint32_t* work, unsigned long long* counts, int stack_entries, int grid, hipStream_t stream) {
    switch (stack_entries) {
        case 8: return launch_s<8>(sc, rays, n, hits, any_hit, spill, work, counts, grid, stream);
        case 16: return launch_s<16>(sc, rays, n, hits, any_hit, spill, work, counts, grid, stream);
        case 32: return launch_s<32>(sc, rays, n, hits, any_hit, spill, work, counts, grid, stream);
        default: return hipErrorInvalidValue;
    }
}

int trace_blocks_per_cu(int stack_entries) {
    int blocks = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (stack_entries) {
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_trace<8, false, false, false>, TRACE_BLOCK, 0); break;
        case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_trace<16, false, false, false>, TRACE_BLOCK, 0); break;
        case 32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_trace<32, false, false, false>, TRACE_BLOCK, 0); break;
        default: break;
    }
    return e == hipSuccess && blocks > 0 ? blocks : 1;
}

}  // namespace ptd
