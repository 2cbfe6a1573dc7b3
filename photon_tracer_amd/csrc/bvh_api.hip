// The C ABI's BVH builds (pt_bvh_build, pt_pack_ex): bvh::build
// (bvh.hpp:39-52) over caller boxes on host threads or on the GPU, and the
// scene packer with the GPU builder. Host code only; the kernels are in
// bvh_gpu.hip.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdlib>
#include <cstring>

#include "api_internal.hpp"
#include "bvh_gpu.hpp"
#include "host.hpp"

namespace ptapi {
namespace {

constexpr size_t kMaxBoxes = size_t(1) << 30;

struct owned {
    dbuf b;
    ~owned() { b.release(); }
};

}  // namespace

// build_bvh's tree, built by the device builder.
pt::bvh_tree device_bvh(const std::vector<pt::box3>& boxes) {
    static_assert(sizeof(pt::box3) == 24, "boxes are 6 floats");
    pt::bvh_tree out;
    const size_t n = boxes.size();
    if (n == 0) return out;
    if (n > kMaxBoxes) throw pt::status_error(PT_ERR_LIMIT, "BVH build: more than 2^30 primitives");
    owned b, order, nodes;
    b.b.ensure(24 * n);
    order.b.ensure(4 * n);
    nodes.b.ensure(sizeof(pt::bvh_node) * n);
    HIPC(hipMemcpy(b.b.p, boxes.data(), 24 * n, hipMemcpyHostToDevice));
    uint32_t k = 0;
    HIPC(ptd::bvhg::build(b.b.as<float>(), static_cast<uint32_t>(n), order.b.as<uint32_t>(), nodes.b.p, &k, nullptr));
    if (k == 0 || k > n) throw pt::status_error(PT_ERR_DEVICE, "BVH build: bad node count from the device");
    out.nodes.resize(k);
    out.order.resize(n);
    HIPC(hipMemcpy(out.nodes.data(), nodes.b.p, sizeof(pt::bvh_node) * k, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(out.order.data(), order.b.p, 4 * n, hipMemcpyDeviceToHost));
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
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<pt::box3> b(n);
        if (n) std::memcpy(b.data(), boxes, 24 * n);
        const pt::bvh_tree t = (flags & PT_BVH_DEVICE) ? ptapi::device_bvh(b) : pt::build_bvh(b);
        const std::vector<uint32_t> words(reinterpret_cast<const uint32_t*>(t.nodes.data()),
                                          reinterpret_cast<const uint32_t*>(t.nodes.data() + t.nodes.size()));
        size_t n_words = 0, n_order = 0;
        *nodes = pt::copy_out(words, &n_words);
        try {
            *order = pt::copy_out(t.order, &n_order);
        } catch (...) {
            std::free(*nodes);
            *nodes = nullptr;
            throw;
        }
        *n_nodes = t.nodes.size();
        if (ms) *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
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
