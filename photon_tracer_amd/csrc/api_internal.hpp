// Host-side state shared by the C ABI's translation units (api.hip,
// guide_api.hip): device buffers, the scene handle, the guide handle.
#pragma once

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "../../include/photon_tracer.h"
#include "guide_kernels.hpp"
#include "kernels.hpp"
#include "status.hpp"

using ptd::NONE;

namespace ptapi {

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw pt::status_error(e == hipErrorOutOfMemory ? PT_ERR_NOMEM : PT_ERR_DEVICE,
                               std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

struct dbuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        release();
        HIPC(hipMalloc(&p, n ? n : 16));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// The tail launch takes over a frame's paths once a bounce's queues hold
// fewer than this many per lane of the traversal grid. Measured at 4, 8, 16,
// 32 and 64: 32 made the 640x360 sphere field 20% faster (its tail starts
// after the camera rays) and the 64-job batch 3.7% faster, and left the 1080p
// 1M-triangle frame unchanged. Scenes with instances keep 4: their tail walks
// one query per lane without refill, and 32 made the instancing scene 8%
// slower.
inline constexpr uint32_t kTailPathsPerLane = 32;
inline constexpr uint32_t kTailPathsPerLaneInstanced = 4;
inline uint32_t tail_threshold(int trace_grid, bool instanced) {
    return static_cast<uint32_t>(trace_grid) * ptd::TRACE_BLOCK *
           (instanced ? kTailPathsPerLaneInstanced : kTailPathsPerLane);
}

// Paths in flight per batch: up to 2^27 (a 1080p frame at 64 spp in one
// batch), or what a quarter of the free HBM holds at kPathBytes each.
inline constexpr size_t kMaxPaths = size_t(1) << 27;
inline constexpr size_t kPathBytes = 2 * 36 + 24 + 48 + 48;

struct render_bufs {
    dbuf ro[2], rd[2], rpath[2], hits, so, sd, sl, pa, pb, pkey, pq, paux;
    dbuf accum, aux_accum, ctrl, frame, totals, out_rgb, out_aux;
    dbuf gout, gmis, grec, gwt, gblk, gblk_sum, gtot, gsamples;   // guided and training renders
    std::vector<dbuf*> all() {
        return {&ro[0], &ro[1], &rd[0], &rd[1], &rpath[0], &rpath[1], &hits, &so, &sd, &sl, &pa, &pb, &pkey, &pq,
                &paux, &accum, &aux_accum, &ctrl, &frame, &totals, &out_rgb, &out_aux, &gout, &gmis, &grec, &gwt, &gblk,
                &gblk_sum, &gtot, &gsamples};
    }
    void release() {
        for (dbuf* b : all()) b->release();
    }
    size_t bytes() {
        size_t n = 0;
        for (dbuf* b : all()) n += b->bytes;
        return n;
    }
};

}  // namespace ptapi

using ptapi::dbuf;
using ptapi::hip_check;

struct pt_scene {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    std::vector<uint32_t> host;  // the PTGP words
    dbuf words, leaf, spill, work, io_rays, io_hits, counts;
    ptd::SceneView view{};
    int max_stack = 0;   // entries a lane's stack can reach in this scene
    // New number of each PTGP node in the device copy (the trees' top
    // levels first); node_map[k] for node k of the caller's buffer.
    std::vector<uint32_t> node_map;
    uint32_t device_node(uint32_t k) const { return k == ptd::NONE ? k : node_map[k]; }
    int tuned_stack = 0, tuned_blocks = 0;
    uint64_t tuned_paths = 0;   // paths one render batch may hold (pt_scene_autotune; 0: memory-sized)
    ptapi::render_bufs rb;
    pt_guide* guide = nullptr;   // attached learned guide, if any
    std::vector<pt_guide*> guides;   // every guide made on this scene (they use its stream)
    // The last captured bounce-loop graph and the launch it was captured for.
    struct graph_cache {
        hipGraphExec_t exec = nullptr;
        ptd::FrameParams key{};
        int S = 0, trace_grid = 0, shade_grid = 0;
    } graph;
    // The bounce at which the last graph render of shape tail_hint_key had
    // the tail launch take over (0: it did not in every batch).
    ptd::FrameParams tail_hint_key{};
    uint32_t tail_hint = 0;
    void drop_graph() {
        if (graph.exec) (void)hipGraphExecDestroy(graph.exec);
        graph.exec = nullptr;
    }
    ~pt_scene();
    void release_all() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        rb.release();
        for (dbuf* b : {&words, &leaf, &spill, &work, &io_rays, &io_hits, &counts}) b->release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};


// A learned guide: parameters, Adam state, training buffers, the position
// box its inputs are normalised to.
struct pt_guide {
    pt_scene* scene = nullptr;
    dbuf p32, m, v, w16, wt16;
    dbuf partials, losses, grad, samples, io_a, io_b, io_c, io_d;
    dbuf records, train_count;   // PT_RENDER_TRAIN: per-path vertex records, compacted training samples
    float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    float lr = 1e-3f;
    int steps_per_batch = 0, minibatch = 16384;
    float weight_clamp = 16.0f;  // training weights are clamped at this many times their batch mean
    int bounces = ptd::MAX_BOUNCES;   // diffuse vertices of the first `bounces` bounces sample the guide
    uint32_t step = 0;           // Adam's step count
    uint64_t stat_steps = 0, stat_samples = 0;
    float stat_loss = 0;
    ptd::guide::GuideView view() const {
        ptd::guide::GuideView g;
        g.w16 = w16.as<_Float16>();
        g.wt16 = wt16.as<_Float16>();
        g.p32 = p32.as<float>();
        g.lo = make_float4(lo[0], lo[1], lo[2], 0);
        g.inv_ext = make_float4(1 / std::max(hi[0] - lo[0], 1e-6f), 1 / std::max(hi[1] - lo[1], 1e-6f),
                                1 / std::max(hi[2] - lo[2], 1e-6f), 0);
        return g;
    }
    ptd::guide::GuideParamsDev params() const {
        return {p32.as<float>(), m.as<float>(), v.as<float>(), w16.as<_Float16>(), wt16.as<_Float16>()};
    }
    ~pt_guide();
};

namespace ptapi {
// Adam steps of a guide over n device samples (3 float4 each), in
// minibatches of g->minibatch, steps_per_batch of them (0: one pass).
void guide_train_on(pt_guide* g, const float4* samples, uint32_t n);

// Checks a PTGP v1 buffer as pt_scene_create does (throws status_error).
void check_ptgp(const uint32_t* w, size_t n);
// LDS stack entries for a render (8, 16 or 32): requested, else tuned, else
// the default for the scene.
int stack_entries_for(pt_scene* s, int requested);
// Traversal grid: blocks per CU (tuned, else the occupancy limit) x CUs.
int trace_grid_for(pt_scene* s, int S);
// The traversal stacks' spill area for a grid of that many blocks.
void ensure_spill(pt_scene* s, int grid);
}  // namespace ptapi
