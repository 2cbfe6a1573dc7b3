// The C ABI's BVH builds (pt_bvh_build, pt_pack_ex): bvh::build
// (bvh.hpp:39-52) over caller boxes on host threads or on the GPU, and the
// scene packer with the GPU builder. Host code only; the kernels are in
// bvh_gpu.hip.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "api_internal.hpp"
#include "bvh_gpu.hpp"
#include "host.hpp"

namespace ptapi {
namespace {

constexpr size_t kMaxBoxes = size_t(1) << 30;

// Device buffers kept between builds (grown, never shrunk, until the
// process ends): allocating a few hundred MB per call cost more than the
// build of 1M boxes itself. The pinned host buffer stages the tree on its
// way back (a copy from pageable memory runs at half the link's rate). One
// build at a time per process.
struct Pool {
    std::mutex m;
    dbuf boxes, order, nodes, scratch;
    void* host = nullptr;   // pinned, host_bytes long
    size_t host_bytes = 0;
    void ensure_host(size_t n) {
        if (n <= host_bytes && host) return;
        if (host) (void)hipHostFree(host);
        host = nullptr;
        host_bytes = 0;
        HIPC(hipHostMalloc(&host, n, hipHostMallocDefault));
        host_bytes = n;
    }
};
Pool& pool() {
    static Pool* p = new Pool;   // never destroyed: no device frees after the runtime is gone
    return *p;
}

// fn(lo, hi) over [0, bytes) in pieces of at least 4 MB on up to 16 host
// threads (one piece: the calling thread alone).
template <class Fn>
void parallel_bytes(size_t bytes, Fn fn) {
    constexpr size_t kPiece = size_t(4) << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t pieces = std::min<size_t>(std::min(hw, 16u), bytes / kPiece);
    if (pieces < 2) {
        if (bytes) fn(size_t(0), bytes);
        return;
    }
    auto part = [&](size_t k) { fn(bytes * k / pieces, bytes * (k + 1) / pieces); };
    std::vector<std::thread> th;
    th.reserve(pieces - 1);
    for (size_t k = 1; k < pieces; ++k) th.emplace_back(part, k);
    part(0);
    for (std::thread& t : th) t.join();
}

// memcpy split over host threads: the destinations here are fresh
// allocations, whose first touch (a page fault per 4 KB) costs more than
// the copy itself; spread over threads, the faults are taken in parallel.
void parallel_copy(void* dst, const void* src, size_t bytes) {
    parallel_bytes(bytes, [&](size_t lo, size_t hi) {
        std::memcpy(static_cast<char*>(dst) + lo, static_cast<const char*>(src) + lo, hi - lo);
    });
}

// The BOX_* flags of n boxes in host memory (a NaN coordinate, a -0.0
// coordinate), scanned branch-free on host threads, so bad boxes are refused
// before any device call.
uint32_t host_box_flags(const float* boxes, size_t n) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(boxes);
    std::atomic<uint32_t> flags{0};
    parallel_bytes(24 * n, [&](size_t lo, size_t hi) {
        uint32_t nan = 0, nz = 0;
        for (size_t i = lo / 4; i < hi / 4; ++i) {
            const uint32_t u = w[i];
            nan |= (u & 0x7fffffffu) > 0x7f800000u ? 1u : 0u;
            nz |= u == 0x80000000u ? 1u : 0u;
        }
        flags.fetch_or((nan ? ptd::bvhg::BOX_NAN : 0u) | (nz ? ptd::bvhg::BOX_NEG_ZERO : 0u));
    });
    return flags.load();
}

// The device build over n caller boxes (host memory, 6 floats each): on
// success the tree is in the pool, *k nodes in P.nodes and the order in
// P.order, and its words are staged in P.host (nodes, then order). Returns
// false for boxes with a -0.0 coordinate, which the host builder takes: -0
// and +0 compare equal, so std::min and std::max keep whichever of the two
// comes first in the reference's order, while the device reductions (atomics
// on order-preserving keys, shuffles) always pick one sign. With a NaN,
// std::min's result depends on the order it is taken in, so the parallel
// build could not promise the host's tree: refused, before any device call.
// The build checks the boxes again on the device as it starts.
bool device_build(Pool& P, const float* boxes, size_t n, uint32_t* k) {
    if (n > ptd::bvhg::kMaxDeviceBoxes)
        throw pt::status_error(PT_ERR_LIMIT, "device BVH build: more than 2^28 primitives");
    const uint32_t host_flags = host_box_flags(boxes, n);
    if (host_flags & ptd::bvhg::BOX_NAN)
        throw pt::status_error(PT_ERR_INVALID, "device BVH build: a box has a NaN coordinate");
    if (host_flags & ptd::bvhg::BOX_NEG_ZERO) return false;
    P.boxes.ensure(24 * n);
    P.order.ensure(4 * n);
    P.nodes.ensure(sizeof(pt::bvh_node) * n);
    const size_t scratch = ptd::bvhg::scratch_bytes(static_cast<uint32_t>(n));
    P.scratch.ensure(scratch);
    HIPC(hipMemcpy(P.boxes.p, boxes, 24 * n, hipMemcpyHostToDevice));
    uint32_t flags = 0;
    *k = 0;
    HIPC(ptd::bvhg::build(P.boxes.as<float>(), static_cast<uint32_t>(n), P.order.as<uint32_t>(), P.nodes.p, k,
                          &flags, P.scratch.p, scratch, nullptr));
    if (flags & ptd::bvhg::BOX_NAN)
        throw pt::status_error(PT_ERR_INVALID, "device BVH build: a box has a NaN coordinate");
    if (flags & ptd::bvhg::BOX_NEG_ZERO) return false;
    if (*k == 0 || *k > n) throw pt::status_error(PT_ERR_DEVICE, "BVH build: bad node count from the device");
    const size_t node_bytes = sizeof(pt::bvh_node) * *k;
    P.ensure_host(node_bytes + 4 * n);
    HIPC(hipMemcpyAsync(P.host, P.nodes.p, node_bytes, hipMemcpyDeviceToHost, nullptr));
    HIPC(hipMemcpyAsync(static_cast<char*>(P.host) + node_bytes, P.order.p, 4 * n, hipMemcpyDeviceToHost, nullptr));
    HIPC(hipStreamSynchronize(nullptr));
    return true;
}

}  // namespace

// build_bvh's tree, built by the device builder (the host builder for boxes
// with a -0.0 coordinate).
pt::bvh_tree device_bvh(const std::vector<pt::box3>& boxes) {
    static_assert(sizeof(pt::box3) == 24, "boxes are 6 floats");
    pt::bvh_tree out;
    const size_t n = boxes.size();
    if (n == 0) return out;
    Pool& P = pool();
    std::lock_guard<std::mutex> lock(P.m);
    uint32_t k = 0;
    if (!device_build(P, reinterpret_cast<const float*>(boxes.data()), n, &k)) return pt::build_bvh(boxes);
    const pt::bvh_node* nd = static_cast<const pt::bvh_node*>(P.host);
    const uint32_t* ord = reinterpret_cast<const uint32_t*>(nd + k);
    out.nodes.assign(nd, nd + k);
    out.order.assign(ord, ord + n);
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
        uint32_t* words = nullptr;
        size_t n_nodes_out = 0;
        bool done = false;
        if ((flags & PT_BVH_DEVICE) && n > 0) {
            // Straight from the caller's boxes to the device, and from the
            // pinned staging copy into the returned arrays.
            ptapi::Pool& P = ptapi::pool();
            std::lock_guard<std::mutex> lock(P.m);
            uint32_t k = 0;
            if (ptapi::device_build(P, boxes, n, &k)) {
                const size_t node_bytes = sizeof(pt::bvh_node) * k;
                words = static_cast<uint32_t*>(std::malloc(node_bytes + 4));
                uint32_t* ord = static_cast<uint32_t*>(std::malloc(4 * n + 4));
                if (!words || !ord) {
                    std::free(words);
                    std::free(ord);
                    throw std::bad_alloc();
                }
                ptapi::parallel_copy(words, P.host, node_bytes);
                ptapi::parallel_copy(ord, static_cast<const char*>(P.host) + node_bytes, 4 * n);
                *order = ord;
                n_nodes_out = k;
                done = true;
            }
        }
        if (!done) {
            std::vector<pt::box3> b(n);
            if (n) std::memcpy(b.data(), boxes, 24 * n);
            const pt::bvh_tree t = pt::build_bvh(b);
            size_t n_order = 0;
            words = static_cast<uint32_t*>(std::malloc(sizeof(pt::bvh_node) * t.nodes.size() + 4));
            if (!words) throw std::bad_alloc();
            if (!t.nodes.empty()) std::memcpy(words, t.nodes.data(), sizeof(pt::bvh_node) * t.nodes.size());
            try {
                *order = pt::copy_out(t.order, &n_order);
            } catch (...) {
                std::free(words);
                throw;
            }
            n_nodes_out = t.nodes.size();
        }
        *nodes = words;
        *n_nodes = n_nodes_out;
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
