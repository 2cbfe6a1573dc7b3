// Launch interface between the C ABI (api.hip) and the kernel translation
// units (trace.hip, render.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"
#include "guide.hpp"

namespace ptd {

// Per-bounce control block in HBM: queue sizes and the work counters the
// persistent kernels take items from. Zeroed by k_finish at the end of every
// batch, so a captured graph can be replayed.
struct Ctrl {
    uint32_t n_rays;      // closest-hit queue size of this bounce
    uint32_t n_shadow;    // shadow queue size of this bounce
    uint32_t n_backface;  // light samples that count as rays but need no trace
    uint32_t w_extend;    // next item of the traversal stage (shadow + closest-hit queues) to hand out
    uint32_t w_unused;
    uint32_t pad[3];
    uint32_t w_seg[8];    // the traversal stage's work counters, one per segment of its queues (XCD_SEGMENTS)
};
// The traversal stage splits its queues into this many contiguous segments,
// one per XCD: each XCD's waves take work from their own segment (neighbouring
// pixels' rays, whose nodes then stay in that XCD's L2) and from the others
// once theirs is done.
#ifndef PT_XCD_SEGMENTS
#define PT_XCD_SEGMENTS 8
#endif
constexpr uint32_t XCD_SEGMENTS = PT_XCD_SEGMENTS;
constexpr int MAX_BOUNCES = 64;

// Per-render values the kernels read from HBM rather than from their
// arguments, so that one captured graph serves every render of the same shape.
struct FrameConst {
    uint64_t seed;
    int32_t first_sample;
    uint32_t batch;       // batch number, advanced by k_finish
    uint32_t tail_from;   // bounce at which the tail launch took over, or ~0u
    uint32_t pad;
};

// Samples (float4) one accumulation tile of a shared-queue pass holds in LDS.
constexpr uint32_t POOL_TILE = 1024;

// Shared-queue batches: one pass of the bounce loop carries the paths of
// many jobs at once, in slices. A slice is one job's pixels for a run of its
// samples (ns a pixel from sample index index0 on); its paths are numbered
// pixel-major from path_base, as one render numbers a batch's. Everything a
// kernel reads of a job comes from its path's slice, found from the path
// number by slice_of's search, whose result is a valid slice for any number.
struct PoolSlice {
    float4 cam[4];          // origin, corner, across, up
    uint64_t seed;
    uint32_t W, H;          // the job's image size
    uint32_t ns, index0;    // samples a pixel in this slice, the first one's number
    uint32_t root;          // the job's tree root (device node), or NONE
    uint32_t large_off, n_large;    // its run of the merged large-sphere list
    uint32_t lights_off, n_lights;  // its run of the merged light list
    uint32_t flags;         // 1 sky, 2 sample lights (the job's NO_NEE applied)
    uint32_t pix_off;       // the job's first pixel in the sums
    uint32_t tile_pix;      // pixels per accumulation tile: tile_pix * ns <= POOL_TILE, or 1
    uint32_t pad[2];
};
static_assert(sizeof(PoolSlice) == 128, "slice records are 128 bytes");

struct FrameParams {
    SceneView sc;
    int W, H, spp, depth;
    int j_lo, j_hi;           // rows rendered: j in [j_lo, j_hi), counted from the bottom
    int spb;                  // samples per pixel per batch
    uint32_t P, npix;         // paths per batch, pixels in the band
    int nee, want_aux;
    // ray queues (double buffered by bounce parity): (o, tmin), (d, tmax), path
    float4* ro[2];
    float4* rd[2];
    uint32_t* rpath[2];
    Hit* hits;
    // shadow queue: (o, tmin), (d, tmax), (L, path)
    float4* so;
    float4* sd;
    float4* sl;
    // path state: (throughput, rng counter), (radiance, flags), rng key, (pixel seed, sample index)
    float4* pa;
    float4* pb;
    uint2* pkey;
    uint2* pq;
    float4* paux;             // 2 per path: (first albedo, 0), (first normal, 0)
    float4* accum;            // per pixel: rgb sum
    float4* aux_accum;        // 2 per pixel: (albedo sum, lum^2 sum), (normal sum, 0)
    Ctrl* ctrl;               // MAX_BOUNCES + 1 entries; ctrl[MAX_BOUNCES] holds the tail launch's counts
    FrameConst* frame;        // seed, first sample, batch number
    unsigned long long* totals;  // [0] closest-hit rays, [1] shadow rays, [2] primary rays
    uint2* spill;             // traversal stack spill area
    unsigned long long* visits;  // null, or seven totals the traversal stage counts into (Visits), then
                                 // the diffuse hits the guide's inference evaluated
    uint32_t tail_threshold;  // queues below this many rays go to the tail launch (0: never)
    uint32_t camera_packets;  // 1: camera rays of a scene without instances are walked as wave packets
    uint32_t tail_last;       // bounce at which the tail launch takes over whatever the queue sizes, and
                              // the last bounce launched (0: every bounce launched, the tail by queue size)
    // The learned guide: 0 off, 1 diffuse bounces sample it (one-sample MIS
    // with the cosine lobe), 2 they also record training vertices.
    int guide_mode;
    guide::GuideView gv;
    uint32_t* gout;           // per slot of the current bounce's queue: the encoded mixture, a word per lobe
    float4* grec;             // GUIDE_RECORDS per path, 4 float4 each (layout at record_vertex)
    float* gwt;               // per record: its training weight
    uint32_t* gblk;           // per block of 256 records: count of weighted records, then its offset
    float* gblk_sum;          // per block: sum of its weights
    uint32_t* gtot;           // [0] training samples written, [1] float bits of their weight sum
    float4* gsamples;         // compacted training samples, 3 float4 each
    float gclamp;             // training weights are clamped at gclamp x their batch mean
    int guide_bounces;        // diffuse vertices of bounces [0, guide_bounces) sample the guide
    float* gmis;              // guided renders: per path, the product of its guided MIS factors
    // Shared-queue passes (0 slices: a single render): n_slices slices, their
    // first paths (slice_path[0] = 0, increasing, slice_path[n_slices] = P),
    // first pass pixels (slice_pix, the same shape, ending at npix) and first
    // accumulation tiles (slice_tile, ending at the pass's tile count).
    const PoolSlice* slices;
    const uint32_t* slice_path;
    const uint32_t* slice_pix;
    const uint32_t* slice_tile;
    uint32_t n_slices;
};
constexpr uint32_t GUIDE_RECORDS = 2;   // diffuse vertices per path recorded for training
constexpr uint32_t GUIDE_TARGET_BLOCK = 256;


// Traversal kernels for ray batches (pt_trace).
// counts: null, or six 64-bit visit totals (Visits) to add the batch's into.
hipError_t launch_trace(const SceneView& sc, const float4* rays, uint32_t n, Hit* hits, bool any_hit, uint2* spill,
                        uint32_t* work, unsigned long long* counts, int stack_entries, int grid, hipStream_t stream);
// Occupancy of the traversal kernels: blocks of 256 per CU (pt_trace's
// kernel, and the path loop's traversal stage).
int trace_blocks_per_cu(int stack_entries);
int path_blocks_per_cu(int stack_entries, bool inst);
constexpr int TRACE_BLOCK = 256;

// The bounce loop of one batch of paths. ev: null, or 2 * depth events
// recorded around each bounce's traversal launch.
hipError_t launch_batch(const FrameParams& fp, int stack_entries, int trace_grid, int shade_grid, hipStream_t stream,
                        hipEvent_t* ev);
hipError_t launch_resolve(const FrameParams& fp, float* rgb, float* aux, hipStream_t stream);
// One shared-queue pass: the bounce loop over every slice's paths (no
// instances, guide or aux), then each pass pixel's samples added onto its
// job's sums. The caller clears the sums first (launch_clear_frame over all
// of them) and resolves each job after the last pass.
hipError_t launch_pool_pass(const FrameParams& fp, int stack_entries, int trace_grid, int shade_grid,
                            hipStream_t stream);
hipError_t launch_progressive_add(const float* img, float* sum, uint32_t npix, int n, bool first, hipStream_t stream);
hipError_t launch_progressive_resolve(const float* sum, float* rgb, uint32_t npix, int done, hipStream_t stream);
// After a training batch: its records' weights, compacted in path order into
// fp.gsamples (count and weight sum in fp.gtot).
hipError_t launch_guide_targets(const FrameParams& fp, hipStream_t stream);
hipError_t launch_clear_frame(const FrameParams& fp, hipStream_t stream);

}  // namespace ptd
