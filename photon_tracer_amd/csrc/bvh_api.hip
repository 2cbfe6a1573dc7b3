// The C ABI's BVH builds (pt_bvh_build, pt_pack_ex): bvh::build
// (bvh.hpp:39-52) over caller boxes on host threads or on the GPU, and the
// scene packer with the GPU builder. Host code only; the kernels are in
// bvh_gpu.hip.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "api_internal.hpp"
#include "bvh_gpu.hpp"
#include "host.hpp"

namespace ptapi {
namespace {

constexpr size_t kMaxBoxes = size_t(1) << 30;

// Device buffers kept between builds (grown, never shrunk, until the
// process ends): allocating a few hundred MB per call cost more than the
// build of 1M boxes itself. One build at a time per process.
struct Pool {
    std::mutex m;
    dbuf boxes, order, nodes, scratch;
};
Pool& pool() {
    static Pool* p = new Pool;   // never destroyed: no device frees after the runtime is gone
    return *p;
}

}  // namespace

// build_bvh's tree, built by the device builder.
pt::bvh_tree device_bvh(const std::vector<pt::box3>& boxes) {
    static_assert(sizeof(pt::box3) == 24, "boxes are 6 floats");
    pt::bvh_tree out;
    const size_t n = boxes.size();
    if (n == 0) return out;
    if (n > ptd::bvhg::kMaxDeviceBoxes)
        throw pt::status_error(PT_ERR_LIMIT, "device BVH build: more than 2^28 primitives");
    // With a NaN, std::min's result depends on the order it is taken in, so
    // the parallel build could not promise the host's tree: refused.
    bool negative_zero = false;
    for (const pt::box3& b : boxes)
        for (int a = 0; a < 3; ++a) {
            if (std::isnan(b.lo[a]) || std::isnan(b.hi[a]))
                throw pt::status_error(PT_ERR_INVALID, "device BVH build: a box has a NaN coordinate");
            negative_zero |= (b.lo[a] == 0 && std::signbit(b.lo[a])) || (b.hi[a] == 0 && std::signbit(b.hi[a]));
        }
    // -0 and +0 compare equal, so std::min and std::max keep whichever of the
    // two comes first in the reference's order, while the device reductions
    // (atomics on order-preserving keys, shuffles) always pick one sign. The
    // node words could then differ in a sign bit: such boxes are built by the
    // host builder, which takes the reference's order.
    if (negative_zero) return pt::build_bvh(boxes);
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.m);
    P.boxes.ensure(24 * n);
    P.order.ensure(4 * n);
    P.nodes.ensure(sizeof(pt::bvh_node) * n);
    const size_t scratch = ptd::bvhg::scratch_bytes(static_cast<uint32_t>(n));
    P.scratch.ensure(scratch);
    HIPC(hipMemcpy(P.boxes.p, boxes.data(), 24 * n, hipMemcpyHostToDevice));
    uint32_t k = 0;
    HIPC(ptd::bvhg::build(P.boxes.as<float>(), static_cast<uint32_t>(n), P.order.as<uint32_t>(), P.nodes.p, &k,
                          P.scratch.p, scratch, nullptr));
    if (k == 0 || k > n) throw pt::status_error(PT_ERR_DEVICE, "BVH build: bad node count from the device");
    out.nodes.resize(k);
    out.order.resize(n);
    HIPC(hipMemcpy(out.nodes.data(), P.nodes.p, sizeof(pt::bvh_node) * k, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(out.order.data(), P.order.p, 4 * n, hipMemcpyDeviceToHost));
    return out;
}

}  // namespace ptapi

extern "C" {

pt_status pt_bvh_build(const float* boxes, size_t n, int flags, uint32_t** nodes, size_t* n_nodes, uint32_t** order,
                       double* ms) {
    if (!nodes || !n_nodes || !order) return pt::fail(PT_ERR_INVALID, "pt_bvh_build: null output pointer");
    *nodes = nullptr;
    *order = nullptr;
    *n_nodes = 0;
    if (n > 0 && !boxes) return pt::fail(PT_ERR_INVALID, "pt_bvh_build: null boxes");
    if (flags & ~PT_BVH_DEVICE) return pt::fail(PT_ERR_INVALID, "pt_bvh_build: unknown flags");
    if (n > ptapi::kMaxBoxes) return pt::fail(PT_ERR_LIMIT, "pt_bvh_build: more than 2^30 boxes");
    PT_GUARD({
        std::vector<pt::box3> b(n);
        if (n) std::memcpy(b.data(), boxes, 24 * n);
        const auto t0 = std::chrono::steady_clock::now();
        const pt::bvh_tree t = (flags & PT_BVH_DEVICE) ? ptapi::device_bvh(b) : pt::build_bvh(b);
        const auto t1 = std::chrono::steady_clock::now();
        size_t n_order = 0;
        uint32_t* words = static_cast<uint32_t*>(std::malloc(sizeof(pt::bvh_node) * t.nodes.size() + 4));
        if (!words) throw std::bad_alloc();
        if (!t.nodes.empty()) std::memcpy(words, t.nodes.data(), sizeof(pt::bvh_node) * t.nodes.size());
        try {
            *order = pt::copy_out(t.order, &n_order);
        } catch (...) {
            std::free(words);
            throw;
        }
        *nodes = words;
        *n_nodes = t.nodes.size();
        if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    });
}

pt_status pt_pack_ex(const uint32_t* ptsd, size_t n_words, int flags, uint32_t** ptgp, size_t* ptgp_words) {
    if (!ptsd || !ptgp || !ptgp_words) return pt::fail(PT_ERR_INVALID, "pt_pack_ex: null pointer");
    if (flags & ~PT_PACK_DEVICE_BVH) return pt::fail(PT_ERR_INVALID, "pt_pack_ex: unknown flags");
    PT_GUARD({
        const pt::scene_desc s = pt::ptsd_read(ptsd, n_words);
        const pt::bvh_builder device = ptapi::device_bvh;
        *ptgp = pt::copy_out(pt::pack_ptgp(s, (flags & PT_PACK_DEVICE_BVH) ? &device : nullptr), ptgp_words);
    });
}

}  // extern "C"
