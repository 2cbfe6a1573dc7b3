// The wavefront path loop: render_row_linear / ray_color (renderer.hpp:100-149,
// 366-416) split into stages over SoA queues in HBM.
//
//   extend   closest hit of every queued ray            (geometry::hit);
//            at bounce 0 the batch's camera rays, made in the kernel
//   shade    emission, scatter, light sample, roulette  (ray_color's body);
//            at bounce 0 the path state starts in registers
//   shadow   any hit of every queued light sample       (geometry::occluded)
//   ...      extend/shade/shadow again for each bounce, up to max_bounces
//   accumulate  per pixel, the batch's samples in sample order
//   finish   ray counts into the totals, control blocks zeroed
//
// A path's state is indexed by its path number, not its queue slot, and every
// random number it draws comes from its own counter-based stream, so the
// image does not depend on the order in which queues fill. Queues are filled
// by wave-aggregated appends (ballot, popcount, one atomic per wave).
#include "kernels.hpp"
#include "traverse.hpp"

namespace ptd {

namespace {

constexpr int SHADE_BLOCK = 1024;
constexpr int SHADE_WAVES = SHADE_BLOCK / 64;
// The guide's inference: blocks of eight independent waves sharing one LDS
// copy of the network (42 KB), 2 blocks per CU: 4 waves per SIMD, what the
// kernel's 128 registers allow.
constexpr int INFER_BLOCK = 512;
constexpr int INFER_BLOCKS_PER_CU = 2;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63; }

// Slots in K queues for the lanes of a block where pred[k] holds: each wave
// counts its lanes with a ballot, the block adds up its waves in LDS, and one
// thread takes the block's run of each queue with a single atomic (the
// counters are shared by the whole grid across all XCDs, where atomics on one
// address are slow). Called by every thread of the block.
template <int K>
struct BlockAppend {
    uint32_t count[K][SHADE_WAVES];
    uint32_t base[K][SHADE_WAVES];
};

template <int K>
__device__ __forceinline__ void block_append(BlockAppend<K>& sh, uint32_t* const* counters, const bool* pred,
                                             uint32_t* slot) {
    const uint32_t w = threadIdx.x >> 6;
    unsigned long long mask[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        mask[k] = __ballot(pred[k]);
        if (lane_id() == 0) sh.count[k][w] = static_cast<uint32_t>(__popcll(mask[k]));
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        uint32_t total = 0;
        for (int i = 0; i < SHADE_WAVES; ++i) total += sh.count[k][i];
        uint32_t at = total ? atomicAdd(counters[k], total) : 0;
        for (int i = 0; i < SHADE_WAVES; ++i) {
            sh.base[k][i] = at;
            at += sh.count[k][i];
        }
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane_id()) - 1ull;
#pragma unroll
    for (int k = 0; k < K; ++k) slot[k] = sh.base[k][w] + static_cast<uint32_t>(__popcll(mask[k] & below));
    __syncthreads();  // sh is reused by the next iteration
}

// sample_2d (sampler.hpp:107-124): QMC for pairs up to 2 * qmc_bounces,
// else two numbers from the path's stream.
__device__ __forceinline__ void sample_2d(uint32_t pseed, uint32_t index, int pair, uint64_t key, uint32_t& cnt,
                                          float& a, float& b) {
    if (pair > 2 * QMC_BOUNCES) {
        a = rand01(key, cnt);
        b = rand01(key, cnt);
        return;
    }
    qmc_pair(pseed, index, pair, a, b);
}

__device__ __forceinline__ f3 reflect(f3 v, f3 n) { return v - n * (2 * dot(v, n)); }

__device__ __forceinline__ f3 refract(f3 uv, f3 n, float ratio) {
    const float cos_theta = fminf(dot(-uv, n), 1.0f);
    const f3 r_perp = (uv + n * cos_theta) * ratio;
    const f3 r_parallel = n * -sqrtf(fabsf(1 - dot(r_perp, r_perp)));
    return r_perp + r_parallel;
}

__device__ __forceinline__ float schlick(float cosine, float ratio) {
    float r0 = (1 - ratio) / (1 + ratio);
    r0 = r0 * r0;
    const float x = 1 - cosine, x2 = x * x;
    return r0 + (1 - r0) * x2 * x2 * x;
}

}  // namespace

// Path p of the current batch: its pixel, sample number and camera ray
// (renderer.hpp:379-385, camera.hpp:45-47). Paths are numbered pixel-major
// within the band, spb samples of a pixel next to each other, so a wave's
// camera rays are neighbours. Recomputed where needed rather than stored:
// bounce 0 reads no ray or path state from HBM.
struct CameraSample {
    f3 o, d;
    uint32_t pseed, index;
    int i, j;
};

__device__ __forceinline__ CameraSample camera_sample(const FrameParams& fp, const FrameConst& fc, uint32_t p) {
    CameraSample cs;
    const uint32_t pixel = p / fp.spb, s = p % fp.spb;
    cs.i = static_cast<int>(pixel % fp.W);
    const int k = static_cast<int>(pixel / fp.W);   // row within the band, top first
    cs.j = fp.j_hi - 1 - k;
    cs.index = static_cast<uint32_t>(fc.first_sample) + fc.batch * fp.spb + s;
    cs.pseed = pixel_seed(fc.seed, cs.i, cs.j);
    float du, dv;
    qmc_pair(cs.pseed, cs.index, 0, du, dv);
    const float u = (static_cast<float>(cs.i) + du) / static_cast<float>(fp.W - 1);
    const float v = (static_cast<float>(cs.j) + dv) / static_cast<float>(fp.H - 1);
    cs.o = xyz(fp.sc.cam[0]);
    cs.d = xyz(fp.sc.cam[1]) + xyz(fp.sc.cam[2]) * u + xyz(fp.sc.cam[3]) * v - cs.o;
    return cs;
}

// Shared-queue passes. The slice of pass path p (with slice_pix, of pass
// pixel p): the last slice whose first entry is at most p. base[0] = 0, so
// the answer lies in [0, n - 1] for any p, and only base[0, n) is read; the
// host launches these kernels with n >= 1.
__device__ __forceinline__ uint32_t slice_of(const uint32_t* base, uint32_t n, uint32_t p) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (base[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A lane's last slice and its path range: queues keep paths roughly in
// path order, so the next path is mostly in the same slice.
// A read-only word at a wave-uniform index, through the scalar cache.
__device__ __forceinline__ uint32_t scalar_u32(const uint32_t* base, uint32_t i) {
    const __attribute__((address_space(4))) uint32_t* q = (const __attribute__((address_space(4))) uint32_t*)(base);
    return q[i];
}

// slice_of for a wave-uniform p, by scalar loads.
__device__ __forceinline__ uint32_t slice_of_uniform(const uint32_t* base, uint32_t n, uint32_t p) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (scalar_u32(base, mid) <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A lane's last slice and its path range. A lookup outside the range first
// tries the slice of the first lane that looks one up (the wave's paths are
// mostly one slice's), found once for the wave by scalar loads, and searches
// on its own only when its path lies outside that one too.
struct SliceCache {
    uint32_t k = 0, lo = 1, hi = 0;   // an empty range: the first lookup searches
    __device__ __forceinline__ uint32_t find(const FrameParams& fp, uint32_t p) {
        if (!(p >= lo && p < hi)) {
            const uint32_t k0 = slice_of_uniform(fp.slice_path, fp.n_slices, __builtin_amdgcn_readfirstlane(p));
            const uint32_t lo0 = scalar_u32(fp.slice_path, k0), hi0 = scalar_u32(fp.slice_path, k0 + 1);
            if (p >= lo0 && p < hi0) {
                k = k0;
                lo = lo0;
                hi = hi0;
            } else {
                k = slice_of(fp.slice_path, fp.n_slices, p);
                lo = fp.slice_path[k];
                hi = fp.slice_path[k + 1];   // the array holds n_slices + 1 entries
            }
        }
        return k;
    }
};

// Slice k for the lanes that ask; one read through the scalar unit when they
// all ask for the same slice (a wave's paths are mostly one job's).
__device__ __forceinline__ PoolSlice read_slice(const FrameParams& fp, uint32_t k) {
    const uint32_t k0 = __builtin_amdgcn_readfirstlane(k);
    if (__ballot(k != k0) == 0) return fp.slices[k0];
    return fp.slices[k];
}

// camera_sample for path pl of a slice (counted from the slice's first
// path): the job's whole frame, its samples numbered from index0.
__device__ __forceinline__ CameraSample camera_sample_pool(const PoolSlice& S, uint32_t pl) {
    CameraSample cs;
    const uint32_t pixel = pl / S.ns, s = pl % S.ns;
    cs.i = static_cast<int>(pixel % S.W);
    cs.j = static_cast<int>(S.H) - 1 - static_cast<int>(pixel / S.W);
    cs.index = S.index0 + s;
    cs.pseed = pixel_seed(S.seed, cs.i, cs.j);
    float du, dv;
    qmc_pair(cs.pseed, cs.index, 0, du, dv);
    const float u = (static_cast<float>(cs.i) + du) / static_cast<float>(static_cast<int>(S.W) - 1);
    const float v = (static_cast<float>(cs.j) + dv) / static_cast<float>(static_cast<int>(S.H) - 1);
    cs.o = xyz(S.cam[0]);
    cs.d = xyz(S.cam[1]) + xyz(S.cam[2]) * u + xyz(S.cam[3]) * v - cs.o;
    return cs;
}

// The scene a slice's rays are traced against: the merged scene with the
// job's tree root and large spheres.
__device__ __forceinline__ SceneView slice_scene(const SceneView& sc, const PoolSlice& S) {
    SceneView v = sc;
    v.flat_root = S.root;
    v.large = sc.large + S.large_off;
    v.n_large = S.n_large;
    return v;
}

// What shading reads of a job besides the shared arrays: the sky flag, NEE
// and the lights NEE picks from.
struct LightView {
    uint32_t flags;
    bool nee;
    const uint32_t* lights;
    uint32_t n_lights;
};
__device__ __forceinline__ LightView light_view(const FrameParams& fp) {
    return LightView{fp.sc.flags, fp.nee != 0, fp.sc.lights, fp.sc.n_lights};
}
__device__ __forceinline__ LightView light_view(const FrameParams& fp, const PoolSlice& S) {
    return LightView{S.flags, (S.flags & 2u) != 0, fp.sc.lights + S.lights_off, S.n_lights};
}

// An unoccluded light sample of bounce bs: its radiance joins its path and,
// while the guide trains, the record of the vertex that took it, whose
// radiance then counts as arrived before the vertex's scattered ray.
__device__ __forceinline__ void light_sample_arrives(const FrameParams& fp, uint32_t item, int bs) {
    const float4 l = fp.sl[item];
    const uint32_t p = __float_as_uint(l.w);
    float4 r = fp.pb[p];
    r.x = r.x + l.x;
    r.y = r.y + l.y;
    r.z = r.z + l.z;
    fp.pb[p] = r;
    if (fp.guide_mode == 2) {
        const uint32_t k = (__float_as_uint(r.w) >> 1) & 3u;
        if (k > 0) {
            float4* R = fp.grec + (static_cast<size_t>(p) * GUIDE_RECORDS + k - 1) * 4;
            if (__float_as_uint(R[0].w) == static_cast<uint32_t>(bs)) {
                float4 t = R[3];
                t.y = t.y + l.x;
                t.z = t.z + l.y;
                t.w = t.w + l.z;
                R[3] = t;
            }
        }
    }
}

// The traversal stage of bounce b: the light samples queued by bounce b - 1
// (any hit, geometry::occluded: an unoccluded sample adds its radiance to its
// path) and the rays of bounce b (closest hit, geometry::hit), drained by one
// persistent grid from one counter, shadow rays first. PRIMARY: bounce 0,
// whose rays are the batch's camera rays, made here; it has no shadow rays.
//
// The traversal stage walks with lane refill: each lane advances its own
// query one stack entry at a time, and once at least `lanes` lanes of a wave
// have finished theirs, those lanes take the next items from the wave's run
// of the queue (one atomic per run of 64 * fetch_groups items). A wave then
// stays busy with most of its lanes, instead of waiting for its longest walk
// before taking 64 new ones. Starting a query costs the whole wave the setup
// code (a camera ray: the QMC jitter and the camera; a queued ray: its loads
// and three divisions), so lanes wait until most of the wave is idle: 48 of
// 64 for queued rays (the speed is flat from 40 to 60), and all 64 for camera
// rays, whose wave is one pixel's samples with walks alike (the whole wave
// takes its next 64 together: 3% faster than refilling at 56 once the QMC
// jitter got cheap). Measured on the 1M-triangle icosphere at 1080p. Each
// query still goes through the same steps, so the hits are the same.
template <bool PRIMARY>
struct refill_threshold {
#ifndef PT_REFILL_LANES
#define PT_REFILL_LANES 48
#endif
    static constexpr uint32_t lanes = PRIMARY ? 64 : PT_REFILL_LANES;
};

// Nodes of the trees' top levels each traversal block keeps in LDS: 40 x
// 144 bytes next to the 8-entry stack ring's 16 KB lets 7 blocks of 256
// share a CU's 160 KB.
#ifndef PT_NODE_CACHE
#define PT_NODE_CACHE 40
#endif
constexpr uint32_t NODE_CACHE = PT_NODE_CACHE;

#ifndef PT_LEAF_BATCH_P
#define PT_LEAF_BATCH_P 0
#endif
#ifndef PT_LEAF_BATCH_S
#define PT_LEAF_BATCH_S 12
#endif
template <bool PRIMARY>
struct leaf_batch {
    static constexpr uint32_t lanes = PRIMARY ? PT_LEAF_BATCH_P : PT_LEAF_BATCH_S;
};

// Traversal queues of at least this many items per lane of the grid are
// split into XCD_SEGMENTS segments.
#ifndef PT_SEGMENT_MIN
#define PT_SEGMENT_MIN 64
#endif
constexpr uint32_t SEGMENT_MIN = PT_SEGMENT_MIN;

// Scenes with instances walk with lane refill too (the instance tree's steps
// in the same walk); false keeps the per-wave fetch of 64 queries for them.
constexpr bool INST_REFILL = true;

template <int S, bool COUNT, bool PRIMARY, bool INST, bool POOL, class Stack>
__device__ __forceinline__ void extend_refill(const FrameParams& fp, const SceneView& sc, int b, Ctrl* c,
                                              uint32_t n_shadow, uint32_t n, Stack& st, Visits& vis) {
    const float4* ro = fp.ro[b & 1];
    const float4* rd = fp.rd[b & 1];
    FrameConst fc;
    if (PRIMARY) fc = *fp.frame;
    const uint32_t chunk = 64 * fetch_groups(n);
    const unsigned long long below = (1ull << lane_id()) - 1ull;
    // Segment x of the work: shadow items [ns * x / S, ns * (x + 1) / S), then
    // closest-hit items n_shadow + [nr * x / S, nr * (x + 1) / S); pool and
    // pool_end are offsets within the wave's current segment. Blocks are
    // dispatched to the XCDs in turn, so block b's XCD is b % 8. Queues of
    // fewer than SEGMENT_MIN items per lane of the grid stay one segment
    // (measured: a 640x360 frame's camera rays took twice as long split).
    const uint32_t nr = n - n_shadow;
    const uint32_t nseg = uint64_t(n) >= uint64_t(SEGMENT_MIN) * gridDim.x * blockDim.x ? XCD_SEGMENTS : 1u;
    uint32_t seg = blockIdx.x % nseg;
    uint32_t seg_done = 0;
    uint32_t s_lo = 0, s_len = 0, r_lo = 0, r_len = 0;
    auto seg_bounds = [&](uint32_t x) {
        s_lo = static_cast<uint32_t>(uint64_t(n_shadow) * x / nseg);
        s_len = static_cast<uint32_t>(uint64_t(n_shadow) * (x + 1) / nseg) - s_lo;
        r_lo = static_cast<uint32_t>(uint64_t(nr) * x / nseg);
        r_len = static_cast<uint32_t>(uint64_t(nr) * (x + 1) / nseg) - r_lo;
    };
    seg_bounds(seg);
    uint32_t pool = 0, pool_end = 0;   // the wave's run of items not yet started
    bool drained = false;              // the queue has no more items
    bool live = false, any = false;
    uint32_t item = 0;
    Walk<INST> w;
    SliceCache slc;
    while (true) {
        const unsigned long long idle = __ballot(!live);
        const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (drained) {
            if (n_idle == 64) break;
        } else if (n_idle >= refill_threshold<PRIMARY>::lanes) {
            while (pool == pool_end && !drained) {
                const uint32_t len = s_len + r_len;
                const uint32_t got = wave_fetch(&c->w_seg[seg], chunk);
                if (got < len) {
                    pool = got;
                    pool_end = min(got + chunk, len);
                } else if (++seg_done == nseg) {
                    drained = true;
                } else {
                    seg = (seg + 1) % nseg;
                    seg_bounds(seg);
                }
            }
            if (!live) {
                const uint32_t r = static_cast<uint32_t>(__popcll(idle & below));
                if (pool + r < pool_end) {
                    const uint32_t off = pool + r;
                    item = off < s_len ? s_lo + off : n_shadow + r_lo + (off - s_len);
                    any = !PRIMARY && item < n_shadow;
                    f3 o, d;
                    float tmin = T_MIN, tmax = T_MAX;
                    if constexpr (POOL) {
                        // The ray's path and slice: the job's camera, tree and large spheres.
                        uint32_t p = item;
                        if (!PRIMARY) {
                            p = any ? reinterpret_cast<const uint32_t*>(fp.sl)[4 * size_t(item) + 3]
                                    : fp.rpath[b & 1][item - n_shadow];
                        }
                        const PoolSlice sl = read_slice(fp, slc.find(fp, p));
                        if (PRIMARY) {
                            const CameraSample cs = camera_sample_pool(sl, item - slc.lo);
                            o = cs.o;
                            d = cs.d;
                        } else {
                            const float4 a = any ? fp.so[item] : ro[item - n_shadow];
                            const float4 bb = any ? fp.sd[item] : rd[item - n_shadow];
                            o = xyz(a);
                            d = xyz(bb);
                            tmin = a.w;
                            tmax = bb.w;
                        }
                        live = !walk_begin<COUNT, INST>(slice_scene(sc, sl), w, o, d, tmin, tmax, any, st, &vis);
                    } else {
                        if (PRIMARY) {
                            const CameraSample cs = camera_sample(fp, fc, item);
                            o = cs.o;
                            d = cs.d;
                        } else {
                            const float4 a = any ? fp.so[item] : ro[item - n_shadow];
                            const float4 bb = any ? fp.sd[item] : rd[item - n_shadow];
                            o = xyz(a);
                            d = xyz(bb);
                            tmin = a.w;
                            tmax = bb.w;
                        }
                        live = !walk_begin<COUNT, INST>(sc, w, o, d, tmin, tmax, any, st, &vis);
                    }
                    if (!live) {   // ended at once: only an any-hit query on a large sphere
                        item |= 0x80000000u;
                    }
                }
            }
            pool = min(pool + n_idle, pool_end);
        }
        bool done = (item & 0x80000000u) != 0;
        // A lane whose next entry is a leaf waits while fewer than
        // LEAF_BATCH lanes are at leaves and others still have nodes to
        // visit, so the leaf tests' loads run for more lanes at once; each
        // lane's own sequence of steps is unchanged.
        bool step = live;
        if (leaf_batch<PRIMARY>::lanes > 0) {
            const bool at_leaf = live && (w.e.x >> COUNT_SHIFT) != 0;
            const uint32_t n_leaf = static_cast<uint32_t>(__popcll(__ballot(at_leaf)));
            const bool nodes_left = __ballot(live && !at_leaf) != 0;
            if (at_leaf && nodes_left && n_leaf < leaf_batch<PRIMARY>::lanes) step = false;
        }
        // Waves stepping only light samples (any hit) push hit children
        // without ranking them. Scenes with instances keep the ranking: the
        // extra ballot made the instancing scene 1.4% slower, against 0.7%
        // faster for the 1M-triangle mesh.
        const bool ordered = PRIMARY || INST || __ballot(step && !any) != 0;
        if (step && walk_step<COUNT, INST>(sc, w, any, st, &vis, ordered)) {
            live = false;
            done = true;
        }
        if (done) {
            item &= 0x7fffffffu;
            if (any) {
                if (w.kind == 0) {   // unoccluded: the light sample reaches its path
                    light_sample_arrives(fp, item, b - 1);
                }
            } else {
                Hit h;
                h.t = w.tmax;
                h.u = w.hu;
                h.v = w.hv;
                h.kind = w.kind;
                h.prim = w.prim;
                if constexpr (INST) h.inst = w.inst;
                else h.inst = NONE;
                fp.hits[item - n_shadow] = h;
            }
            item = 0;
        }
    }
}

// Camera rays of scenes without instances, walked a wave at a time as a
// packet. A wave's 64 camera rays are 64 consecutive paths of the batch:
// one pixel's samples, or the samples of a few neighbouring pixels, so they
// visit nearly the same nodes. The wave keeps one stack in LDS of (node,
// mask of the lanes whose rays entered it), reads each node's and leaf's
// words once through the scalar cache, and every lane in the mask tests its
// own ray against the four children or the leaf's primitives, with its own
// shrinking t_max, in the arithmetic of the per-lane walk. Children are
// taken nearest first by the entry distances of the first lane that entered
// any of them. A lane leaves out a child only when its slab test fails, so
// it tests every primitive its own walk could reach before its nearest hit
// and ends with the same nearest hit, except where two candidates are tied:
// two primitives at the same t (a ray through an edge shared by two leaves),
// or one whose t is rounded below its box's entry, where the visiting order
// decides which one the walk keeps. The packets start at multiples of 64
// paths, so a packet's rays depend on the batch's paths only, not on the
// launch shape. PT_RENDER_PER_RAY walks camera rays one per lane instead.
constexpr int PK_DEPTH = 256;   // entries: the deepest stack pt_scene_create accepts
struct PacketStack {
    uint32_t code[PK_DEPTH];
    uint32_t lo[PK_DEPTH], hi[PK_DEPTH];   // the lanes' mask
};
// A read-only global row read at a wave-uniform address: the constant
// address space lets the compiler load it through the scalar cache.
__device__ __forceinline__ float4 scalar_row(const void* base, size_t i) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const __attribute__((address_space(4))) v4* q = (const __attribute__((address_space(4))) v4*)(base);
    const v4 r = q[i];
    return make_float4(r.x, r.y, r.z, r.w);
}

template <bool COUNT>
__device__ __forceinline__ void packet_walk(const SceneView& sc, bool on, f3 o, f3 d, float tmin, float& tmax,
                                            uint32_t& kind, uint32_t& prim, float& hu, float& hv, PacketStack& ps,
                                            Visits* vis) {
    const uint32_t lane = lane_id();
    if (COUNT && on) ++vis->queries;
    // Large spheres, in double, per lane (geometry.hpp:117-118, 157-160).
    for (uint32_t k = 0; k < sc.n_large; ++k) {
        if (COUNT && on) ++vis->large;
        const uint32_t s = sc.large[k];
        if (on && sphere_hit_d(sc.spheres[2 * s], o, d, tmin, tmax)) {
            kind = 1;
            prim = s;
        }
    }
    const f3 inv = safe_rcp3(d);
    unsigned long long mask = __ballot(on);
    uint32_t code = sc.flat_root;
    if (mask == 0 || code == NONE) return;
    int top = 0;
    while (true) {
        code = __builtin_amdgcn_readfirstlane(code);   // uniform: its words load through the scalar cache
        const bool lon = ((mask >> lane) & 1ull) != 0;
        const uint32_t first = code & FIRST_MASK, count = code >> COUNT_SHIFT;
        bool next = false;
        if (count == 0) {
            if (COUNT && lon) ++vis->nodes;
            const size_t nd = static_cast<size_t>(first) * 8;
            const float4 lx = scalar_row(sc.nodes, nd), ly = scalar_row(sc.nodes, nd + 1),
                         lz = scalar_row(sc.nodes, nd + 2), hx = scalar_row(sc.nodes, nd + 3),
                         hy = scalar_row(sc.nodes, nd + 4), hz = scalar_row(sc.nodes, nd + 5),
                         cw = scalar_row(sc.nodes, nd + 6);
            const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w},
                        lza[4] = {lz.x, lz.y, lz.z, lz.w};
            const float hxa[4] = {hx.x, hx.y, hx.z, hx.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w},
                        hza[4] = {hz.x, hz.y, hz.z, hz.w};
            const uint32_t cd[4] = {__float_as_uint(cw.x), __float_as_uint(cw.y), __float_as_uint(cw.z),
                                    __float_as_uint(cw.w)};
            unsigned long long m[4];
            float tn[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                // visit_node's slab test, in the same operation order
                const float x0 = (lxa[c] - o.x) * inv.x, x1 = (hxa[c] - o.x) * inv.x;
                const float y0 = (lya[c] - o.y) * inv.y, y1 = (hya[c] - o.y) * inv.y;
                const float z0 = (lza[c] - o.z) * inv.z, z1 = (hza[c] - o.z) * inv.z;
                const float near = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
                const float far = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
                tn[c] = near;
                m[c] = mask & __ballot(near <= far);   // lanes in the packet whose slab test passes
            }
            const unsigned long long any = m[0] | m[1] | m[2] | m[3];
            if (any != 0) {
                // The order: the entry distances of the first lane that entered a child.
                const int leader = __builtin_ctzll(any);
                float lt[4];
                int nv = 0;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float t = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(tn[c]), leader));
                    lt[c] = ((m[c] >> leader) & 1ull) ? t : __builtin_inff();
                    nv += m[c] != 0 ? 1 : 0;
                }
                // rank 0: taken next; the others pushed so the nearer pop first
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (m[c] == 0) continue;
                    int rank = 0;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (e != c && m[e] != 0 && (lt[e] < lt[c] || (lt[e] == lt[c] && e < c))) ++rank;
                    if (rank == 0) {
                        code = cd[c];
                        mask = m[c];
                        next = true;
                    } else {
                        const int at = top + (nv - 1 - rank);
                        if (lane == 0) {
                            ps.code[at] = cd[c];
                            ps.lo[at] = static_cast<uint32_t>(m[c]);
                            ps.hi[at] = static_cast<uint32_t>(m[c] >> 32);
                        }
                    }
                }
                top += nv - 1;
            }
        } else {
            // A leaf's primitives in order (geometry.hpp:119-127).
            for (uint32_t i = first; i < first + count; ++i) {
                const float4 r0 = scalar_row(sc.leaf, 3 * size_t(i)), r1 = scalar_row(sc.leaf, 3 * size_t(i) + 1);
                if (COUNT && lon) {
                    ++vis->leafrefs;
                    if (__float_as_uint(r1.w) != 0u) ++vis->spheres;
                    else ++vis->tris;
                }
                if (__float_as_uint(r1.w) != 0u) {
                    if (lon && sphere_hit_f(make_float4(r0.x, r0.y, r0.z, r1.x), o, d, tmin, tmax)) {
                        kind = 2;
                        prim = __float_as_uint(r0.w);
                    }
                } else {
                    const float4 r2 = scalar_row(sc.leaf, 3 * size_t(i) + 2);
                    if (lon && tri_hit(xyz(r0), xyz(r1), xyz(r2), o, d, tmin, tmax, hu, hv)) {
                        kind = 3;
                        prim = __float_as_uint(r0.w);
                    }
                }
            }
        }
        if (!next) {
            if (top == 0) break;
            --top;
            __builtin_amdgcn_wave_barrier();
            code = __builtin_amdgcn_readfirstlane(ps.code[top]);
            const uint32_t lo = __builtin_amdgcn_readfirstlane(ps.lo[top]), hi = __builtin_amdgcn_readfirstlane(ps.hi[top]);
            mask = (static_cast<unsigned long long>(hi) << 32) | lo;
        }
    }
}

// The camera-ray launch with packet walks: the waves take runs of the frame's
// paths (XCD segments as extend_refill's) and trace 64 at a time.
template <bool COUNT>
__device__ __forceinline__ void extend_packet(const FrameParams& fp, const SceneView& sc, Ctrl* c, uint32_t n,
                                              PacketStack& ps, Visits& vis) {
    const FrameConst fc = *fp.frame;
    const uint32_t chunk = 64 * fetch_groups(n);
    const uint32_t nseg = uint64_t(n) >= uint64_t(SEGMENT_MIN) * gridDim.x * blockDim.x ? XCD_SEGMENTS : 1u;
    uint32_t seg = blockIdx.x % nseg, seg_done = 0;
    while (true) {
        // Segment bounds rounded to multiples of 64 paths (the last ends at n).
        const uint32_t s_lo = static_cast<uint32_t>(uint64_t(n) * seg / nseg) & ~63u;
        const uint32_t s_hi = seg + 1 == nseg ? n : static_cast<uint32_t>(uint64_t(n) * (seg + 1) / nseg) & ~63u;
        const uint32_t s_len = s_hi - s_lo;
        const uint32_t got = wave_fetch(&c->w_seg[seg], chunk);
        if (got >= s_len) {
            if (++seg_done == nseg) break;
            seg = (seg + 1) % nseg;
            continue;
        }
        const uint32_t run_end = s_lo + min(got + chunk, s_len);
        for (uint32_t base = s_lo + got; base < run_end; base += 64) {
            const uint32_t item = base + lane_id();
            const bool on = item < run_end;
            f3 o = mk(0, 0, 0), d = mk(0, 0, 1);
            if (on) {
                const CameraSample cs = camera_sample(fp, fc, item);
                o = cs.o;
                d = cs.d;
            }
            float tmax = T_MAX, hu = 0, hv = 0;
            uint32_t kind = 0, prim = 0;
            packet_walk<COUNT>(sc, on, o, d, T_MIN, tmax, kind, prim, hu, hv, ps, &vis);
            if (on) {
                Hit h;
                h.t = tmax;
                h.u = hu;
                h.v = hv;
                h.kind = kind;
                h.prim = prim;
                h.inst = NONE;
                fp.hits[item] = h;
            }
        }
    }
}

template <bool COUNT>
__global__ __launch_bounds__(TRACE_BLOCK) void k_camera_packet(FrameParams fp, int b) {
    __shared__ PacketStack stk[TRACE_BLOCK / 64];
    Visits vis;
    Ctrl* c = fp.ctrl + b;
    if (blockIdx.x == 0 && threadIdx.x == 0) c->n_rays = fp.P;
    extend_packet<COUNT>(fp, fp.sc, c, fp.P, stk[threadIdx.x >> 6], vis);
    if (COUNT) add_visits(fp.visits, vis);
}

// The camera rays of a shared-queue pass as packets: a run of 64 paths can
// hold the paths of more than one job's slice, so the run is walked once per
// slice among its lanes, each time as a packet of that slice's lanes over the
// job's tree (slice_scene), its rays from its camera (camera_sample_pool).
__global__ __launch_bounds__(TRACE_BLOCK) void k_camera_packet_pool(FrameParams fp, int b) {
    __shared__ PacketStack stk[TRACE_BLOCK / 64];
    PacketStack& ps = stk[threadIdx.x >> 6];
    Visits vis;
    Ctrl* c = fp.ctrl + b;
    const uint32_t n = fp.P;
    if (blockIdx.x == 0 && threadIdx.x == 0) c->n_rays = n;
    const uint32_t chunk = 64 * fetch_groups(n);
    const uint32_t nseg = uint64_t(n) >= uint64_t(SEGMENT_MIN) * gridDim.x * blockDim.x ? XCD_SEGMENTS : 1u;
    uint32_t seg = blockIdx.x % nseg, seg_done = 0;
    SliceCache slc;
    while (true) {
        const uint32_t s_lo = static_cast<uint32_t>(uint64_t(n) * seg / nseg) & ~63u;
        const uint32_t s_hi = seg + 1 == nseg ? n : static_cast<uint32_t>(uint64_t(n) * (seg + 1) / nseg) & ~63u;
        const uint32_t s_len = s_hi - s_lo;
        const uint32_t got = wave_fetch(&c->w_seg[seg], chunk);
        if (got >= s_len) {
            if (++seg_done == nseg) break;
            seg = (seg + 1) % nseg;
            continue;
        }
        const uint32_t run_end = s_lo + min(got + chunk, s_len);
        for (uint32_t base = s_lo + got; base < run_end; base += 64) {
            const uint32_t item = base + lane_id();
            const bool on = item < run_end;
            const uint32_t k = on ? slc.find(fp, item) : 0u;
            unsigned long long todo = __ballot(on);
            while (todo != 0) {
                const uint32_t k0 = __builtin_amdgcn_readlane(k, __builtin_ctzll(todo));
                const bool mine = on && k == k0;
                todo &= ~__ballot(mine);
                const PoolSlice sl = fp.slices[k0];
                f3 o = mk(0, 0, 0), d = mk(0, 0, 1);
                if (mine) {
                    const CameraSample cs = camera_sample_pool(sl, item - slc.lo);
                    o = cs.o;
                    d = cs.d;
                }
                float tmax = T_MAX, hu = 0, hv = 0;
                uint32_t kind = 0, prim = 0;
                packet_walk<false>(slice_scene(fp.sc, sl), mine, o, d, T_MIN, tmax, kind, prim, hu, hv, ps, &vis);
                if (mine) {
                    Hit h;
                    h.t = tmax;
                    h.u = hu;
                    h.v = hv;
                    h.kind = kind;
                    h.prim = prim;
                    h.inst = NONE;
                    fp.hits[item] = h;
                }
            }
        }
    }
}

template <int S, bool COUNT, bool INST, bool PRIMARY, bool POOL = false>
__global__ __launch_bounds__(TRACE_BLOCK)
__attribute__((amdgpu_waves_per_eu(trace_occupancy<S, INST, true>::waves))) void k_extend(FrameParams fp, int b) {
    if (!PRIMARY && fp.frame->tail_from <= static_cast<uint32_t>(b)) return;  // the tail launch has these rays
    __shared__ uint2 lds[S * TRACE_BLOCK];
    __shared__ float4 node_cache[NODE_CACHE * (NODE_CACHE_STRIDE / 16)];
    const SceneView sc = cache_top_nodes(fp.sc, node_cache, NODE_CACHE);
    auto st = lane_stack<S, TRACE_BLOCK>(lds, fp.spill);
    Visits vis;
    Ctrl* c = fp.ctrl + b;
    const uint32_t n_shadow = PRIMARY ? 0u : fp.ctrl[b - 1].n_shadow;
    const uint32_t n_rays = PRIMARY ? fp.P : c->n_rays;
    const uint32_t n = n_shadow + n_rays;
    if (PRIMARY && blockIdx.x == 0 && threadIdx.x == 0) c->n_rays = fp.P;
    if constexpr (!INST || INST_REFILL || POOL) {
        extend_refill<S, COUNT, PRIMARY, INST, POOL>(fp, sc, b, c, n_shadow, n, st, vis);
        if (COUNT) add_visits(fp.visits, vis);
        return;
    }
    const float4* ro = fp.ro[b & 1];
    const float4* rd = fp.rd[b & 1];
    FrameConst fc;
    if (PRIMARY) fc = *fp.frame;
    const uint32_t groups = fetch_groups(n);
    while (true) {
        const uint32_t base = wave_fetch(&c->w_extend, 64 * groups);
        if (base >= n) break;
        for (uint32_t g = 0; g < groups; ++g) {
            const uint32_t item = base + 64 * g + lane_id();
            if (item >= n) break;
            Hit h;
            if (!PRIMARY && item < n_shadow) {
                const float4 a = fp.so[item], bb = fp.sd[item];
                if (!trace_ray<true, COUNT, INST>(fp.sc, xyz(a), xyz(bb), a.w, bb.w, h, st, &vis)) {
                    light_sample_arrives(fp, item, b - 1);
                }
                continue;
            }
            const uint32_t q = item - n_shadow;
            f3 o, d;
            float tmin = T_MIN, tmax = T_MAX;
            if (PRIMARY) {
                const CameraSample cs = camera_sample(fp, fc, q);
                o = cs.o;
                d = cs.d;
            } else {
                const float4 a = ro[q], bb = rd[q];
                o = xyz(a);
                d = xyz(bb);
                tmin = a.w;
                tmax = bb.w;
            }
            trace_ray<false, COUNT, INST>(fp.sc, o, d, tmin, tmax, h, st, &vis);
            fp.hits[q] = h;
        }
    }
    if (COUNT) add_visits(fp.visits, vis);
}

// The state a path carries from bounce to bounce.
struct PathState {
    f3 thr, rad;        // throughput, radiance so far
    uint32_t cnt;       // draws taken from the path's random stream
    bool sampled;       // this bounce took a light sample (renderer.hpp:106-108,130)
    uint64_t key;       // the path's random stream
    uint2 pq;           // QMC pixel seed and sample index
    uint32_t nrec;      // training records written (guide training)
    float mis;          // guided renders: the product of the guided MIS factors in thr
};

// The path's flag word (pb[p].w): bit 0 `sampled`, bits 1-2 the records written.
__device__ __forceinline__ float path_flags(const PathState& ps) {
    return __uint_as_float((ps.sampled ? 1u : 0u) | (ps.nrec << 1));
}
__device__ __forceinline__ void read_flags(float w, PathState& ps) {
    const uint32_t f = __float_as_uint(w);
    ps.sampled = (f & 1u) != 0;
    ps.nrec = (f >> 1) & 3u;
}

// What a bounce produced: the next ray (cont: from pnt along dir) and a light
// sample to test (shadow: from pnt along sh_d up to sh_tmax, adding sh_l if
// unoccluded); backface: a light sample that counts as a ray but is zero
// without a test (renderer.hpp:72,76-77).
struct ShadeResult {
    bool cont, shadow, backface;
    f3 pnt, dir, sh_d, sh_l;
    float sh_tmax;
    uint32_t light;   // the light sample's light (its index in the light list)
};

// The hit record of a ray's closest hit (geometry.hpp:135-151,
// sphere.hpp:50-58, triangle.hpp:63-74, geometry.hpp:231-244): point,
// front-facing shading normal, material, and whether it is a light that
// NEE samples.
struct Surface {
    f3 pnt, nrm;
    bool front, sampled_light;
    uint32_t mat;
    float tu, tv;   // texture coordinates (triangle.hpp:72-73; 0 on spheres, hittable.hpp:14)
};

// A unit vector from the hardware's reciprocal square root (within a few
// ulp of the divisions normalize() takes): for the guide network's inputs,
// which are rounded to FP16 after the encoding.
__device__ __forceinline__ f3 normalize_fast(f3 a) { return a * __builtin_amdgcn_rsqf(dot(a, a)); }

// FAST: the shading normal from normalize_fast (the guide's inference, which
// needs only the point, the normal and the material).
template <bool FAST = false>
__device__ __forceinline__ Surface surface_at(const FrameParams& fp, const Hit& h, f3 o, f3 d, bool nee) {
    const SceneView& sc = fp.sc;
    Surface sf;
    const f3 pnt = o + d * h.t;
    sf.pnt = pnt;
    f3 nrm;
    bool front;
    uint32_t mat;
    bool sampled_light = false;
    sf.tu = sf.tv = 0;
    if (h.kind <= 2) {
        const float4 cr = sc.spheres[2 * h.prim];
        mat = __float_as_uint(sc.spheres[2 * h.prim + 1].x);
        const f3 outward = (pnt - xyz(cr)) * (1 / cr.w);
        front = dot(d, outward) < 0;
        nrm = front ? outward : -outward;
        if (h.kind == 2 && nee) {
            const float4 em = sc.mats[3 * mat + 1];
            sampled_light = em.x > 0 || em.y > 0 || em.z > 0;
        }
    } else {
        const float4* tr = sc.tris + 7 * h.prim;
        mat = __float_as_uint(tr[0].w);
        const f3 n0 = xyz(tr[3]), n1 = xyz(tr[4]), n2 = xyz(tr[5]), fn = xyz(tr[6]);
        const float b0 = 1 - h.u - h.v;
        const f3 sn = FAST ? normalize_fast(n0 * b0 + n1 * h.u + n2 * h.v) : normalize(n0 * b0 + n1 * h.u + n2 * h.v);
        // uv in the rows' w words: (e1, u0) (e2, v0) (n0, u1) (n1, v1) (n2, u2) (fn, v2)
        sf.tu = tr[1].w * b0 + tr[3].w * h.u + tr[5].w * h.v;
        sf.tv = tr[2].w * b0 + tr[4].w * h.u + tr[6].w * h.v;
        if (h.inst == NONE) {
            front = dot(d, fn) < 0;
            nrm = front ? sn : -sn;
        } else {
            const float4* in = sc.inst + 7 * h.inst;
            const float4 r0 = in[3], r1 = in[4], r2 = in[5];
            const f3 ld = mk(r0.x * d.x + r0.y * d.y + r0.z * d.z, r1.x * d.x + r1.y * d.y + r1.z * d.z,
                             r2.x * d.x + r2.y * d.y + r2.z * d.z);
            front = dot(ld, fn) < 0;
            const f3 ln = front ? sn : -sn;
            // the transpose of to_object (transform.hpp:23-27)
            const f3 wn = mk(r0.x * ln.x + r1.x * ln.y + r2.x * ln.z, r0.y * ln.x + r1.y * ln.y + r2.y * ln.z,
                             r0.z * ln.x + r1.z * ln.y + r2.z * ln.z);
            nrm = FAST ? normalize_fast(wn) : normalize(wn);
            const uint32_t im = __float_as_uint(in[6].y);
            if (im != NONE) mat = im;
        }
    }
    sf.nrm = nrm;
    sf.front = front;
    sf.mat = mat;
    sf.sampled_light = sampled_light;
    return sf;
}

// Training record of a diffuse vertex (4 float4): (point, bounce),
// (normal, dir.x), (dir.y, dir.z, thr.x, thr.y), (thr.z, base): thr is the
// throughput before the vertex scatters, base the path's radiance once the
// vertex's own emission and light sample are in, so that (final radiance -
// base) / thr is what arrived along dir.
__device__ __forceinline__ void record_vertex(const FrameParams& fp, uint32_t p, uint32_t k, int b, f3 pnt, f3 nrm,
                                              f3 dir, f3 thr, f3 rad) {
    float4* R = fp.grec + (static_cast<size_t>(p) * GUIDE_RECORDS + k) * 4;
    R[0] = make_float4(pnt.x, pnt.y, pnt.z, __uint_as_float(static_cast<uint32_t>(b)));
    R[1] = make_float4(nrm.x, nrm.y, nrm.z, dir.x);
    R[2] = make_float4(dir.y, dir.z, thr.x, thr.y);
    R[3] = make_float4(thr.z, rad.x, rad.y, rad.z);
}

// ray_color's loop body for bounce b (renderer.hpp:109-146): the ray from o
// along d, whose closest hit is h, of path p. go: the guide's outputs for
// this hit when diffuse bounces sample the guide, else null.
template <bool GUIDED = false>
__device__ __forceinline__ ShadeResult shade_hit(const FrameParams& fp, int b, uint32_t p, const Hit& h, f3 o, f3 d,
                                                 PathState& ps, const uint32_t* go, const LightView& lv) {
    const SceneView& sc = fp.sc;
    ShadeResult r;
    r.cont = r.shadow = r.backface = false;
    r.pnt = r.dir = r.sh_d = r.sh_l = mk(0, 0, 0);
    r.sh_tmax = 0;
    r.light = 0;
    const bool was_sampled = ps.sampled;
    ps.sampled = false;
    if (h.kind == 0) {
        // renderer.hpp:24-29
        f3 bg = mk(0, 0, 0);
        if (lv.flags & 1u) {
            const f3 unit = normalize(d);
            const float t = 0.5f * (unit.y + 1);
            bg = mk(1, 1, 1) * (1 - t) + mk(0.5f, 0.7f, 1.0f) * t;
        }
        ps.rad = ps.rad + mul(ps.thr, bg);
        if (fp.want_aux && b == 0) {
            fp.paux[2 * p] = make_float4(1, 1, 1, 0);
            fp.paux[2 * p + 1] = make_float4(0, 0, 0, 0);
        }
        return r;
    }
    const Surface sf = surface_at(fp, h, o, d, lv.nee);
    const f3 pnt = sf.pnt, nrm = sf.nrm;
    const bool front = sf.front, sampled_light = sf.sampled_light;
    const uint32_t mat = sf.mat;
    r.pnt = pnt;
    const float4 m0 = sc.mats[3 * mat], m1 = sc.mats[3 * mat + 1], m2 = sc.mats[3 * mat + 2];
    if (front && !(was_sampled && sampled_light)) ps.rad = ps.rad + mul(ps.thr, xyz(m1));

    // material::scatter (material.hpp:59-137)
    const int type = static_cast<int>(m0.w);
    f3 att = xyz(m0), dir = mk(0, 0, 0);
    // A textured matte's albedo times its texture (material.hpp:72).
    const uint32_t tex = __float_as_uint(m2.y);
    if (type == 0 && tex != 0) att = mul(att, texture_sample(sc, tex - 1, sf.tu, sf.tv));
    bool ok = true;
    f3 wgt = att;          // the scattered ray's throughput factor (att unless guided)
    bool go_on = true;
    if (type == 0) {
        float a, bq;
        sample_2d(ps.pq.x, ps.pq.y, 1 + 2 * b, ps.key, ps.cnt, a, bq);
        const float rr = sqrtf(a), phi = 2 * PI_F * bq;
        const f3 nn = nrm;
        const float sign = copysignf(1.0f, nn.z);
        const float pp = -1 / (sign + nn.z), qq = nn.x * nn.y * pp;
        const f3 t1 = mk(1 + sign * nn.x * nn.x * pp, sign * qq, -sign * nn.x);
        const f3 t2 = mk(qq, sign + nn.y * nn.y * pp, -nn.y);
        dir = t1 * (rr * cosf(phi)) + t2 * (rr * sinf(phi)) + nn * sqrtf(fmaxf(0.0f, 1 - a));
        if (GUIDED && go) {
            // One-sample MIS: the cosine lobe or the guide's mixture, 50/50,
            // weighted by the mixed pdf, so the estimate stays unbiased for
            // any guide. A direction below the surface carries nothing on.
            if (rand01(ps.key, ps.cnt) >= 0.5f) {
                const float u0 = rand01(ps.key, ps.cnt), u1 = rand01(ps.key, ps.cnt), u2 = rand01(ps.key, ps.cnt);
                dir = guide::decoded_sample(go, u0, u1, u2);
            }
            const float cosv = dot(dir, nrm);
            if (cosv > 0) {
                const float pc = cosv * (1 / PI_F);
                const float f = pc / (0.5f * (pc + guide::decoded_pdf(go, dir)));
                wgt = att * f;
                ps.mis *= f;
                if (fp.guide_mode == 2 && ps.nrec < GUIDE_RECORDS) {
                    record_vertex(fp, p, ps.nrec, b, pnt, nrm, dir, ps.thr, ps.rad);
                    ++ps.nrec;
                }
            } else {
                go_on = false;
            }
        }
    } else if (type == 1) {
        const f3 reflected = reflect(normalize(d), nrm);
        f3 ball = mk(0, 0, 0);
        for (int tries = 0; tries < 64; ++tries) {  // vec3.hpp:121-127
            const float x = -1 + 2 * rand01(ps.key, ps.cnt);
            const float y = -1 + 2 * rand01(ps.key, ps.cnt);
            const float z = -1 + 2 * rand01(ps.key, ps.cnt);
            ball = mk(x, y, z);
            if (dot(ball, ball) < 1) break;
        }
        dir = reflected + ball * m1.w;
        ok = dot(dir, nrm) > 0;
    } else if (type == 2) {
        att = mk(1, 1, 1);
        const float ior = m2.x;
        const float ratio = front ? (1 / ior) : ior;
        const f3 ud = normalize(d);
        const float cos_theta = fminf(dot(-ud, nrm), 1.0f);
        const float sin_theta = sqrtf(1 - cos_theta * cos_theta);
        const bool cannot_refract = ratio * sin_theta > 1;
        const bool refl = cannot_refract || schlick(cos_theta, ratio) > rand01(ps.key, ps.cnt);
        dir = refl ? reflect(ud, nrm) : refract(ud, nrm, ratio);
    } else {
        ok = false;
    }
    if (fp.want_aux && b == 0) {
        const f3 fa = ok && type != 2 ? att : mk(1, 1, 1);
        fp.paux[2 * p] = make_float4(fa.x, fa.y, fa.z, 0);
        fp.paux[2 * p + 1] = make_float4(nrm.x, nrm.y, nrm.z, 0);
    }
    if (!ok) return r;
    const bool nee = type == 0 && lv.nee && lv.n_lights > 0 && b + 1 < fp.depth;
    if (nee) {
        // direct_light (renderer.hpp:41-80)
        const uint32_t nl = lv.n_lights;
        const uint32_t pick = min(static_cast<uint32_t>(rand01(ps.key, ps.cnt) * nl), nl - 1);
        r.light = pick;
        const uint32_t ls = lv.lights[pick];
        const float4 lcr = sc.spheres[2 * ls];
        const uint32_t lmat = __float_as_uint(sc.spheres[2 * ls + 1].x);
        const f3 to_center = xyz(lcr) - pnt;
        const float dist2 = dot(to_center, to_center), r2 = lcr.w * lcr.w;
        if (dist2 > r2) {
            const float sin2_max = r2 / dist2;
            const float cos_max = sqrtf(1 - sin2_max);
            const float omc = sin2_max / (1 + cos_max);
            const f3 w = to_center * (1 / sqrtf(dist2));
            const f3 ax = fabsf(w.x) > 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
            const f3 uu = normalize(cross(w, ax)), vv = cross(w, uu);
            float s1, s2;
            sample_2d(ps.pq.x, ps.pq.y, 2 + 2 * b, ps.key, ps.cnt, s1, s2);
            const float cos_t = 1 - s1 * omc;
            const float sin_t = sqrtf(fmaxf(0.0f, 1 - cos_t * cos_t));
            const float phi = 2 * PI_F * s2;
            const f3 ldir = uu * (cosf(phi) * sin_t) + vv * (sinf(phi) * sin_t) + w * cos_t;
            const float cos_surface = dot(ldir, nrm);
            float t_light = __builtin_huge_valf();
            if (cos_surface > 0 && sphere_hit_f(lcr, pnt, ldir, 0.001f, t_light)) {
                const f3 lp = pnt + ldir * t_light;
                const f3 outward = (lp - xyz(lcr)) * (1 / lcr.w);
                if (dot(ldir, outward) < 0) {
                    const float pdf = 1 / (2 * PI_F * omc);
                    const f3 contrib =
                        xyz(sc.mats[3 * lmat + 1]) * (cos_surface / PI_F / pdf * static_cast<float>(nl));
                    r.shadow = true;
                    r.sh_d = ldir;
                    r.sh_tmax = t_light * 0.9999f;
                    r.sh_l = mul(mul(ps.thr, att), contrib);
                } else {
                    r.backface = true;
                }
            }
        }
    }
    ps.sampled = nee;
    if (!go_on) return r;
    if (type != 0) wgt = att;
    ps.thr = mul(ps.thr, wgt);
    if (b + 1 >= 3) {  // roulette_after (renderer.hpp:22,140-146)
        float pm = fmaxf(ps.thr.x, fmaxf(ps.thr.y, ps.thr.z));
        // A guided render's paths survive at least as often as the cosine
        // lobe's would: the guided MIS factors that lowered the throughput are
        // divided back out of the survival probability (factors above 1 are
        // not, so no path survives less often than without the guide). A path
        // the guide sent toward bright light has a small weight, and a roulette
        // on that weight would kill most such paths and scale the few
        // survivors up, putting back the variance guiding removed. Any
        // survival probability divided out keeps the estimator unbiased.
        if constexpr (GUIDED) pm = pm / fminf(ps.mis, 1.0f);
        if (pm < 1) {
            if (rand01(ps.key, ps.cnt) >= pm) return r;
            ps.thr = ps.thr * (1 / pm);
        }
    }
    if (b + 1 < fp.depth) {
        r.cont = true;
        r.dir = dir;
    }
    return r;
}

// The guide's mixture for the shade stage of bounce b: for every queued hit
// on a diffuse (Lambertian) surface, the network at its point and normal,
// encoded as one 32-bit word per lobe at gout[slot] (octahedral axis,
// log kappa, weight relative to the largest lobe's). The block first copies the network's weights
// into LDS. Each wave then walks its runs of 64 slots and stages the
// diffuse hits among them (their five encoding scalars and slot) in its own
// LDS list; whenever 64 are staged the network runs on them, in registers
// from the encoding to the outputs, so the matrix unit sees full batches
// whatever the share of metal, glass and lights among the hits; the last
// partial batch runs at the end. A query's outputs do not depend on which
// others share its batch. Weights read from the caches instead of LDS made
// the kernel 2.5x slower: their per-layer addresses, hoisted out of the loop
// by the compiler, cost half the registers.
struct alignas(16) InferEntry {
    float e[5];
    uint32_t slot;
    uint32_t pad[2];
};
constexpr int INFER_WAVES = INFER_BLOCK / 64;

// The network and decoding for staged entries [0, count) of a wave's list
// (count <= 64): lane (c, kh) encodes entries c and 32 + c and writes their
// lobes 4 kh .. 4 kh + 3.
__device__ __forceinline__ void infer_staged(const FrameParams& fp, const guide::LdsNet& net, const InferEntry* st,
                                             uint32_t count) {
    const int c = guide::lane_c(), kh = guide::lane_kh();
    float q0[5], q1[5];
    const InferEntry& a = st[c];
    const InferEntry& bq = st[32 + c];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        q0[v] = a.e[v];
        q1[v] = bq.e[v];
    }
    const uint32_t slot[2] = {a.slot, bq.slot};
    guide::RegActs X;
    guide::encode_blobs(q0, q1, X);
    guide::f16v O[2];
    guide::forward_regs(guide::LdsWeights{net}, X, O);
    // The lane holds whole lobes (2 g4 + kh, g4 < 4) of its two entries: each
    // is encoded once here for the shade stage, its weight relative to the
    // largest of the eight, the other half's four lobes taken from the
    // partner lane.
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        float mx = O[nb][3];
#pragma unroll
        for (int g4 = 1; g4 < 4; ++g4) mx = fmaxf(mx, O[nb][4 * g4 + 3]);
        mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
        if (static_cast<uint32_t>(32 * nb + c) < count) {
            uint32_t* out = fp.gout + static_cast<size_t>(slot[nb]) * guide::LOBES + kh;
#pragma unroll
            for (int g4 = 0; g4 < 4; ++g4)
                out[2 * g4] = guide::encode_lobe(O[nb][4 * g4], O[nb][4 * g4 + 1], O[nb][4 * g4 + 2],
                                                 __expf(O[nb][4 * g4 + 3] - mx));
        }
    }
}

template <bool PRIMARY>
__global__ __launch_bounds__(INFER_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_guide_infer(FrameParams fp, int b) {
    if (fp.frame->tail_from <= static_cast<uint32_t>(b)) return;
    const uint32_t n = PRIMARY ? fp.P : fp.ctrl[b].n_rays;
    FrameConst fc;
    if (PRIMARY) fc = *fp.frame;
    const int cur = b & 1;
    __shared__ guide::LdsNet net;
    __shared__ InferEntry stage[INFER_WAVES][128];
    guide::load_net(fp.gv, net);
    InferEntry* st = stage[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t stride = gridDim.x * INFER_BLOCK;
    uint32_t staged = 0;   // entries in the wave's list (fewer than 64 between runs)
    // One loop with one call of the network, so it is compiled once: a run
    // of slots is staged, then a full batch runs, or, once the slots are
    // done, the last partial one.
    for (uint32_t base = blockIdx.x * INFER_BLOCK + (threadIdx.x & ~63u);; base += stride) {
        const bool more = base < n;
        if (more) {
            const uint32_t q = base + lane;
            bool diffuse = false;
            f3 pnt = mk(0, 0, 0), nrm = mk(0, 0, 1);
            if (q < n) {
                const Hit h = fp.hits[q];
                if (h.kind != 0) {
                    f3 o, d;
                    if (PRIMARY) {
                        const CameraSample cs = camera_sample(fp, fc, q);
                        o = cs.o;
                        d = cs.d;
                    } else {
                        o = xyz(fp.ro[cur][q]);
                        d = xyz(fp.rd[cur][q]);
                    }
                    const Surface sf = surface_at<true>(fp, h, o, d, fp.nee != 0);
                    diffuse = static_cast<int>(fp.sc.mats[3 * sf.mat].w) == 0;
                    pnt = sf.pnt;
                    nrm = sf.nrm;
                }
            }
            const unsigned long long rows = __ballot(diffuse);
            if (rows == 0) continue;   // the whole wave agrees
            if (fp.visits && lane == 0) atomicAdd(fp.visits + 7, static_cast<unsigned long long>(__popcll(rows)));
            if (diffuse) {
                InferEntry& en = st[staged + static_cast<uint32_t>(__popcll(rows & below))];
                guide::encode_scalars(fp.gv, pnt, nrm, en.e);
                en.slot = q;
            }
            staged += static_cast<uint32_t>(__popcll(rows));
            __builtin_amdgcn_wave_barrier();
            if (staged < 64) continue;
        } else if (staged == 0) {
            break;
        }
        const uint32_t count = staged < 64 ? staged : 64;
        infer_staged(fp, net, st, count);
        // The entries past the batch move to the front of the list.
        const uint32_t rest = staged - count;
        uint4 m0 = make_uint4(0, 0, 0, 0), m1 = m0;
        if (lane < rest) {
            const uint4* from = reinterpret_cast<const uint4*>(st + 64 + lane);
            m0 = from[0];
            m1 = from[1];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < rest) {
            uint4* to = reinterpret_cast<uint4*>(st + lane);
            to[0] = m0;
            to[1] = m1;
        }
        __builtin_amdgcn_wave_barrier();
        staged = rest;
        if (!more) break;
    }
}

// The shade stage of bounce b over its queue. PRIMARY: bounce 0; the camera
// ray and the fresh path state are rebuilt from the path number instead of
// read back. GUIDED: diffuse vertices sample the guide (its own kernel, so
// the plain one carries none of that code's registers).
template <bool PRIMARY, bool GUIDED = false, bool POOL = false>
__global__ __launch_bounds__(SHADE_BLOCK) void k_shade(FrameParams fp, int b) {
    if (fp.frame->tail_from <= static_cast<uint32_t>(b)) return;  // the tail launch has taken these paths
    Ctrl* c = fp.ctrl + b;
    const uint32_t n = PRIMARY ? fp.P : c->n_rays;
    FrameConst fc;
    if (PRIMARY) fc = *fp.frame;
    const int cur = b & 1, nxt = cur ^ 1;
    __shared__ BlockAppend<4> sh;
    // GUIDED: each thread's encoded mixture (eight 32-bit words), copied from
    // HBM with two 16-byte loads into an LDS row of 9 words (an odd stride:
    // the lobe loops' reads, one word at a time, fall in distinct banks),
    // instead of gathered from HBM inside the lobe loops.
    constexpr int GROW = guide::LOBES + 1;
    __shared__ uint32_t gtile[GUIDED ? SHADE_BLOCK * GROW : 1];
    // Light samples are appended in two runs per block, those to even- and
    // to odd-numbered lights, so a traversal wave mostly holds rays toward
    // one light (directions alike) rather than a random mix.
    uint32_t* const counters[4] = {&fp.ctrl[b + 1].n_rays, &c->n_shadow, &c->n_shadow, &c->n_backface};
    const uint32_t stride = gridDim.x * SHADE_BLOCK;
    SliceCache slc;
    // The block steps through the queue together (block_append synchronises).
    for (uint32_t bbase = blockIdx.x * SHADE_BLOCK; bbase < n; bbase += stride) {
        const uint32_t q = bbase + threadIdx.x;
        ShadeResult r;
        r.cont = r.shadow = r.backface = false;
        uint32_t p = 0;
        if (q < n) {
            const uint32_t* go = nullptr;
            if (GUIDED && b < fp.guide_bounces) {
                const uint4* src = reinterpret_cast<const uint4*>(fp.gout + static_cast<size_t>(q) * guide::LOBES);
                uint32_t* row = gtile + threadIdx.x * GROW;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const uint4 v = src[i];
                    row[4 * i] = v.x;
                    row[4 * i + 1] = v.y;
                    row[4 * i + 2] = v.z;
                    row[4 * i + 3] = v.w;
                }
                go = row;
            }
            const Hit h = fp.hits[q];
            f3 o, d;
            PathState ps;
            LightView lv;
            PoolSlice sl;
            if (POOL) {
                p = PRIMARY ? q : fp.rpath[cur][q];
                sl = read_slice(fp, slc.find(fp, p));
                lv = light_view(fp, sl);
            } else {
                lv = light_view(fp);
            }
            if (PRIMARY) {
                p = q;
                const CameraSample cs = POOL ? camera_sample_pool(sl, p - slc.lo) : camera_sample(fp, fc, p);
                o = cs.o;
                d = cs.d;
                ps.thr = mk(1, 1, 1);
                ps.rad = mk(0, 0, 0);
                ps.cnt = 0;
                ps.sampled = false;
                ps.nrec = 0;
                ps.mis = 1;
                ps.key = path_key(POOL ? sl.seed : fc.seed, cs.i, cs.j, cs.index);
                ps.pq = make_uint2(cs.pseed, cs.index);
            } else {
                p = fp.rpath[cur][q];
                const float4 o4 = fp.ro[cur][q], d4 = fp.rd[cur][q];
                const float4 A = fp.pa[p], B = fp.pb[p];
                const uint2 kk = fp.pkey[p];
                ps.pq = fp.pq[p];
                ps.key = (static_cast<uint64_t>(kk.y) << 32) | kk.x;
                ps.thr = xyz(A);
                ps.rad = xyz(B);
                ps.cnt = __float_as_uint(A.w);
                read_flags(B.w, ps);
                ps.mis = GUIDED ? fp.gmis[p] : 1.0f;
                o = xyz(o4);
                d = xyz(d4);
            }
            r = shade_hit<GUIDED>(fp, b, p, h, o, d, ps, go, lv);
            // The radiance always (accumulate reads it); the rest only for a
            // path that goes on.
            fp.pb[p] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, path_flags(ps));
            if (r.cont) {
                fp.pa[p] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, __uint_as_float(ps.cnt));
                if (PRIMARY) {
                    fp.pkey[p] = make_uint2(static_cast<uint32_t>(ps.key), static_cast<uint32_t>(ps.key >> 32));
                    fp.pq[p] = ps.pq;
                }
                if (GUIDED) fp.gmis[p] = ps.mis;
            }
        }
        // Every thread of the block reaches the appends.
        const bool odd = (r.light & 1u) != 0;
        const bool pred[4] = {r.cont, r.shadow && !odd, r.shadow && odd, r.backface};
        uint32_t slot[4];
        block_append<4>(sh, counters, pred, slot);
        if (r.cont) {
            fp.ro[nxt][slot[0]] = make_float4(r.pnt.x, r.pnt.y, r.pnt.z, T_MIN);
            fp.rd[nxt][slot[0]] = make_float4(r.dir.x, r.dir.y, r.dir.z, T_MAX);
            fp.rpath[nxt][slot[0]] = p;
        }
        if (r.shadow) {
            const uint32_t ss = odd ? slot[2] : slot[1];
            fp.so[ss] = make_float4(r.pnt.x, r.pnt.y, r.pnt.z, 0.001f);
            fp.sd[ss] = make_float4(r.sh_d.x, r.sh_d.y, r.sh_d.z, r.sh_tmax);
            fp.sl[ss] = make_float4(r.sh_l.x, r.sh_l.y, r.sh_l.z, __uint_as_float(p));
        }
    }
}

// Whether bounce b starts the tail: the queues left (rays of bounce b and
// the light samples of bounce b - 1) are too small to fill the GPU stage by
// stage, and no earlier bounce started it.
__device__ __forceinline__ bool tail_starts(const FrameParams& fp, int b) {
    if (fp.frame->tail_from < static_cast<uint32_t>(b)) return false;
    if (fp.guide_mode && b < fp.guide_bounces) return false;   // the tail shades without the guide
    if (fp.tail_last != 0 && static_cast<uint32_t>(b) == fp.tail_last) return true;
    return fp.ctrl[b - 1].n_shadow + fp.ctrl[b].n_rays < fp.tail_threshold;
}

// Tail, part 1: bounce b - 1's light samples, so that their radiance is in
// place before part 2 carries the paths on.
template <int S, bool INST, bool POOL = false>
__global__ __launch_bounds__(TRACE_BLOCK) void k_tail_shadow(FrameParams fp, int b) {
    if (!tail_starts(fp, b)) return;
    __shared__ uint2 lds[S * TRACE_BLOCK];
    auto st = lane_stack<S, TRACE_BLOCK>(lds, fp.spill);
    const uint32_t n = fp.ctrl[b - 1].n_shadow;
    uint32_t* work = &fp.ctrl[MAX_BOUNCES].w_unused;
    const uint32_t groups = fetch_groups(n);
    SliceCache slc;
    while (true) {
        const uint32_t base = wave_fetch(work, 64 * groups);
        if (base >= n) break;
        for (uint32_t g = 0; g < groups; ++g) {
            const uint32_t item = base + 64 * g + lane_id();
            if (item >= n) break;
            const float4 a = fp.so[item], bb = fp.sd[item];
            Hit h;
            bool hit;
            if constexpr (POOL) {
                const uint32_t p = reinterpret_cast<const uint32_t*>(fp.sl)[4 * size_t(item) + 3];
                const SceneView v = slice_scene(fp.sc, read_slice(fp, slc.find(fp, p)));
                hit = trace_ray<true, false, INST>(v, xyz(a), xyz(bb), a.w, bb.w, h, st);
            } else {
                hit = trace_ray<true, false, INST>(fp.sc, xyz(a), xyz(bb), a.w, bb.w, h, st);
            }
            if (!hit) {
                light_sample_arrives(fp, item, b - 1);
            }
        }
    }
}

// Tail, part 2: every path still going at bounce b taken to its end, lane by
// lane: closest hit, shade, light sample, next bounce. The later bounces'
// stage launches then return at once. Rays are counted as the stages count
// them, into ctrl[MAX_BOUNCES].
//
// A lane advances its path one bounce per iteration, and once TAIL_REFILL of
// the wave's lanes have ended theirs, those lanes take the next paths from
// the wave's run of the queue: path lengths vary with Russian roulette, and a
// wave that waited for its longest path kept few lanes busy. Each path still
// goes through the same steps, so the image is the same. (Measured: the tail
// launch of the 1M-triangle frame went from 1.28 to 1.25 ms; what is left is
// the serial chain of the longest paths' remaining bounces.)
constexpr uint32_t TAIL_REFILL = 8;

template <int S, bool INST, bool GUIDED = false, bool POOL = false>
__global__ __launch_bounds__(TRACE_BLOCK) void k_tail(FrameParams fp, int b) {
    if (!tail_starts(fp, b)) return;
    Ctrl* c = fp.ctrl + b;
    const uint32_t n = c->n_rays;
    if (blockIdx.x == 0 && threadIdx.x == 0) fp.frame->tail_from = b;  // read by later launches only
    __shared__ uint2 lds[S * TRACE_BLOCK];
    auto st = lane_stack<S, TRACE_BLOCK>(lds, fp.spill);
    uint32_t rays = 0, shadows = 0, backfaces = 0;
    const int cur = b & 1;
    const uint32_t chunk = 64 * fetch_groups(n);
    const unsigned long long below = (1ull << lane_id()) - 1ull;
    uint32_t pool = 0, pool_end = 0;   // the wave's run of paths not yet started
    bool drained = false, live = false;
    // The lane's path: its number, state, the ray of bounce bb.
    uint32_t p = 0;
    int bb = 0;
    PathState ps;
    f3 o = mk(0, 0, 0), d = mk(0, 0, 1);
    float tmin = T_MIN, tmax = T_MAX;
    // The lane's path's job: its scene and lights (the single render's otherwise).
    SliceCache slc;
    SceneView view = fp.sc;
    LightView lv = light_view(fp);
    while (true) {
        const unsigned long long idle = __ballot(!live);
        const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (drained) {
            if (n_idle == 64) break;
        } else if (n_idle >= TAIL_REFILL) {
            if (pool == pool_end) {
                pool = wave_fetch(&c->w_extend, chunk);
                pool_end = pool < n ? min(pool + chunk, n) : pool;
                drained = pool >= n;
            }
            if (!live) {
                const uint32_t r = static_cast<uint32_t>(__popcll(idle & below));
                if (pool + r < pool_end) {
                    const uint32_t q = pool + r;
                    p = fp.rpath[cur][q];
                    const float4 o4 = fp.ro[cur][q], d4 = fp.rd[cur][q];
                    const float4 A = fp.pa[p], B = fp.pb[p];
                    const uint2 kk = fp.pkey[p];
                    ps.pq = fp.pq[p];
                    ps.key = (static_cast<uint64_t>(kk.y) << 32) | kk.x;
                    ps.thr = xyz(A);
                    ps.rad = xyz(B);
                    ps.cnt = __float_as_uint(A.w);
                    read_flags(B.w, ps);
                    ps.mis = GUIDED ? fp.gmis[p] : 1.0f;
                    o = xyz(o4);
                    d = xyz(d4);
                    tmin = o4.w;
                    tmax = d4.w;
                    bb = b;
                    live = true;
                    if constexpr (POOL) {
                        const PoolSlice sl = read_slice(fp, slc.find(fp, p));
                        view = slice_scene(fp.sc, sl);
                        lv = light_view(fp, sl);
                    }
                }
            }
            pool = min(pool + n_idle, pool_end);
        }
        if (!live) continue;
        Hit h;
        trace_ray<false, false, INST>(view, o, d, tmin, tmax, h, st);
        if (bb > b) ++rays;  // bounce b's rays were counted when they were queued
        const ShadeResult r = shade_hit<GUIDED>(fp, bb, p, h, o, d, ps, nullptr, lv);   // never samples the guide
        if (r.backface) ++backfaces;
        if (r.shadow) {
            ++shadows;
            Hit hs;
            if (!trace_ray<true, false, INST>(view, r.pnt, r.sh_d, 0.001f, r.sh_tmax, hs, st)) {
                ps.rad.x = ps.rad.x + r.sh_l.x;
                ps.rad.y = ps.rad.y + r.sh_l.y;
                ps.rad.z = ps.rad.z + r.sh_l.z;
            }
        }
        if (r.cont && bb + 1 < fp.depth) {
            o = r.pnt;
            d = r.dir;
            tmin = T_MIN;
            tmax = T_MAX;
            ++bb;
        } else {
            fp.pb[p] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, path_flags(ps));
            live = false;
        }
    }
    // Per wave, one atomic per count.
    uint32_t v[3] = {rays, shadows, backfaces};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    if (lane_id() == 0) {
        Ctrl* t = fp.ctrl + MAX_BOUNCES;
        if (v[0]) atomicAdd(&t->n_rays, v[0]);
        if (v[1]) atomicAdd(&t->n_shadow, v[1]);
        if (v[2]) atomicAdd(&t->n_backface, v[2]);
    }
}

// Per pixel, the batch's samples added in sample order onto the running sum
// (renderer.hpp:379-396). A pixel's samples are adjacent (spb float4), so a
// block first copies a tile of pixels' samples into LDS with coalesced loads,
// then one thread per pixel adds them up in order. The aux sums, and batches
// with more samples than a tile holds, take the direct loop.
constexpr uint32_t ACC_TILE = 1024;  // float4 per tile

__global__ __launch_bounds__(256) void k_accumulate(FrameParams fp) {
    // A pixel's row in the tile is padded by one float4 when spb >= 8 (then
    // at most ACC_TILE / 8 pixels per tile), so the adders, reading at a
    // stride of spb float4, spread over the LDS banks.
    __shared__ float4 tile[ACC_TILE + ACC_TILE / 8];
    const uint32_t spb = static_cast<uint32_t>(fp.spb);
    if (!fp.want_aux && spb <= ACC_TILE) {
        const uint32_t tp = ACC_TILE / spb;                 // pixels per tile
        const uint32_t pitch = spb + (spb >= 8 ? 1u : 0u);
        for (uint32_t px0 = blockIdx.x * tp; px0 < fp.npix; px0 += gridDim.x * tp) {
            const uint32_t np = min(tp, fp.npix - px0);
            const uint32_t count = np * spb;
            const float4* src = fp.pb + static_cast<size_t>(px0) * spb;
            for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) tile[(k / spb) * pitch + k % spb] = src[k];
            __syncthreads();
            for (uint32_t q = threadIdx.x; q < np; q += blockDim.x) {
                const uint32_t px = px0 + q;
                const float4 acc = fp.accum[px];
                f3 sum = xyz(acc);
                const float4* t = tile + q * pitch;
                for (uint32_t s = 0; s < spb; ++s) sum = sum + xyz(t[s]);
                fp.accum[px] = make_float4(sum.x, sum.y, sum.z, 0);
            }
            __syncthreads();
        }
        return;
    }
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < fp.npix; px += gridDim.x * blockDim.x) {
        float4 acc = fp.accum[px];
        f3 sum = xyz(acc);
        float4 a0 = make_float4(0, 0, 0, 0), a1 = make_float4(0, 0, 0, 0);
        if (fp.want_aux) {
            a0 = fp.aux_accum[2 * px];
            a1 = fp.aux_accum[2 * px + 1];
        }
        for (uint32_t s = 0; s < spb; ++s) {
            const uint32_t p = px * spb + s;
            const f3 c = xyz(fp.pb[p]);
            sum = sum + c;
            if (fp.want_aux) {
                const float4 fa = fp.paux[2 * p], fn = fp.paux[2 * p + 1];
                a0.x = a0.x + fa.x;
                a0.y = a0.y + fa.y;
                a0.z = a0.z + fa.z;
                a1.x = a1.x + fn.x;
                a1.y = a1.y + fn.y;
                a1.z = a1.z + fn.z;
                const float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;  // renderer.hpp:364
                a0.w += lum * lum;
            }
        }
        fp.accum[px] = make_float4(sum.x, sum.y, sum.z, 0);
        if (fp.want_aux) {
            fp.aux_accum[2 * px] = a0;
            fp.aux_accum[2 * px + 1] = a1;
        }
    }
}

// Shared-queue passes: each pass pixel's slice samples added in sample order
// onto its job's pixel sum; a job's slices come in sample order, pass after
// pass, so every pixel sums its samples in the order a single render does.
// A block takes a tile at a time: tile_pix pixels of one slice, whose
// samples lie next to each other (a pixel's ns samples adjacent), copied into
// LDS with coalesced loads and summed by one thread per pixel. A slice with
// more samples a pixel than a tile holds is summed straight from HBM.
__global__ __launch_bounds__(256) void k_accumulate_pool(FrameParams fp) {
    __shared__ float4 tile[POOL_TILE + POOL_TILE / 8];
    const uint32_t n_tiles = fp.slice_tile[fp.n_slices];
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint32_t k = slice_of(fp.slice_tile, fp.n_slices, t);
        const PoolSlice& sl = fp.slices[k];
        const uint32_t ns = sl.ns, tp = sl.tile_pix;
        const uint32_t npix = fp.slice_pix[k + 1] - fp.slice_pix[k];
        const uint32_t px0 = (t - fp.slice_tile[k]) * tp;
        const uint32_t np = min(tp, npix - px0);
        const float4* src = fp.pb + fp.slice_path[k] + static_cast<size_t>(px0) * ns;
        float4* dst = fp.accum + sl.pix_off + px0;
        if (ns <= POOL_TILE) {
            // A pixel's row is padded by one float4 when ns >= 8 (then at most
            // POOL_TILE / 8 pixels a tile), so the adders spread over the banks.
            const uint32_t pitch = ns + (ns >= 8 ? 1u : 0u);
            const uint32_t count = np * ns;
            for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) tile[(i / ns) * pitch + i % ns] = src[i];
            __syncthreads();
            for (uint32_t q = threadIdx.x; q < np; q += blockDim.x) {
                const float4 acc = dst[q];
                f3 sum = xyz(acc);
                const float4* r = tile + q * pitch;
                for (uint32_t s = 0; s < ns; ++s) sum = sum + xyz(r[s]);
                dst[q] = make_float4(sum.x, sum.y, sum.z, 0);
            }
            __syncthreads();
        } else if (threadIdx.x == 0) {   // tp is 1
            const float4 acc = dst[0];
            f3 sum = xyz(acc);
            for (uint32_t s = 0; s < ns; ++s) sum = sum + xyz(src[s]);
            dst[0] = make_float4(sum.x, sum.y, sum.z, 0);
        }
    }
}

__global__ void k_finish(FrameParams fp) {
    if (threadIdx.x != 0) return;
    unsigned long long rays = 0, shadow = 0;
    for (int b = 0; b <= fp.depth && b < MAX_BOUNCES; ++b) {
        rays += fp.ctrl[b].n_rays;
        shadow += fp.ctrl[b].n_shadow + fp.ctrl[b].n_backface;
    }
    const Ctrl& t = fp.ctrl[MAX_BOUNCES];  // the tail launch's counts
    rays += t.n_rays;
    shadow += t.n_shadow + t.n_backface;
    fp.totals[0] += rays;
    fp.totals[1] += shadow;
    fp.totals[2] += fp.ctrl[0].n_rays;
    // The latest bounce at which the tail launch took over in any batch (0xffff: a batch without it).
    const unsigned long long tf = fp.frame->tail_from == ~0u ? 0xffffull : fp.frame->tail_from;
    fp.totals[3] = fp.totals[3] > tf ? fp.totals[3] : tf;
    for (int b = 0; b <= fp.depth && b < MAX_BOUNCES; ++b) fp.ctrl[b] = Ctrl{};
    fp.ctrl[MAX_BOUNCES] = Ctrl{};
    fp.frame->batch += 1;
    fp.frame->tail_from = ~0u;
}

__global__ __launch_bounds__(256) void k_clear_frame(FrameParams fp) {
    const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px < fp.npix) {
        fp.accum[px] = make_float4(0, 0, 0, 0);
        if (fp.want_aux) {
            fp.aux_accum[2 * px] = make_float4(0, 0, 0, 0);
            fp.aux_accum[2 * px + 1] = make_float4(0, 0, 0, 0);
        }
    }
    if (px == 0) {
        fp.frame->batch = 0;
        fp.frame->tail_from = ~0u;
        fp.totals[0] = fp.totals[1] = fp.totals[2] = fp.totals[3] = 0;
        for (int b = 0; b <= fp.depth && b < MAX_BOUNCES; ++b) fp.ctrl[b] = Ctrl{};
        fp.ctrl[MAX_BOUNCES] = Ctrl{};
    }
}

// Averages into the caller's layout: 3 floats a pixel, the band's top row
// first (renderer.hpp:397-407).
__global__ __launch_bounds__(256) void k_resolve(FrameParams fp, float* rgb, float* aux) {
    const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= fp.npix) return;
    const float scale = 1.0f / static_cast<float>(fp.spp);
    const float4 acc = fp.accum[px];
    rgb[3 * px] = acc.x * scale;
    rgb[3 * px + 1] = acc.y * scale;
    rgb[3 * px + 2] = acc.z * scale;
    if (aux) {
        const float4 a0 = fp.aux_accum[2 * px], a1 = fp.aux_accum[2 * px + 1];
        float* o = aux + 7 * static_cast<size_t>(px);
        o[0] = a0.x * scale;
        o[1] = a0.y * scale;
        o[2] = a0.z * scale;
        o[3] = a1.x * scale;
        o[4] = a1.y * scale;
        o[5] = a1.z * scale;
        o[6] = a0.w * scale;
    }
}

template <int S, bool INST>
static void launch_loop(const FrameParams& fp, int trace_grid, int shade_grid, hipStream_t stream, hipEvent_t* ev) {
    const int acc_grid = shade_grid * 8;
    // Every bounce of a guided render shades with the guided kernels (which
    // sample the guide only below guide_bounces), since its roulette divides
    // the guided MIS factors out of the throughput.
    const bool guided = fp.guide_mode != 0;
    for (int b = 0; b < fp.depth; ++b) {
        if (b > 0) {
            k_tail_shadow<S, INST><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            if (guided) k_tail<S, INST, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            else k_tail<S, INST><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            // The tail has every path from here on: no launch after it has work.
            if (fp.tail_last != 0 && static_cast<uint32_t>(b) == fp.tail_last) break;
        }
        if (ev) (void)hipEventRecord(ev[2 * b], stream);
        if (b == 0) {
            if (!INST && fp.camera_packets) {
                if (fp.visits) k_camera_packet<true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
                else k_camera_packet<false><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            } else if (fp.visits) {
                k_extend<S, true, INST, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            } else {
                k_extend<S, false, INST, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            }
        } else {
            if (fp.visits) k_extend<S, true, INST, false><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            else k_extend<S, false, INST, false><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
        }
        if (ev) (void)hipEventRecord(ev[2 * b + 1], stream);
        if (fp.guide_mode && b < fp.guide_bounces) {
            const int infer_grid = (shade_grid / 2) * INFER_BLOCKS_PER_CU;
            if (ev) (void)hipEventRecord(ev[2 * (fp.depth + b)], stream);
            if (b == 0) k_guide_infer<true><<<infer_grid, INFER_BLOCK, 0, stream>>>(fp, b);
            else k_guide_infer<false><<<infer_grid, INFER_BLOCK, 0, stream>>>(fp, b);
            if (ev) (void)hipEventRecord(ev[2 * (fp.depth + b) + 1], stream);
        }
        if (b == 0) {
            if (guided) k_shade<true, true><<<shade_grid, SHADE_BLOCK, 0, stream>>>(fp, b);
            else k_shade<true><<<shade_grid, SHADE_BLOCK, 0, stream>>>(fp, b);
        } else {
            if (guided) k_shade<false, true><<<shade_grid, SHADE_BLOCK, 0, stream>>>(fp, b);
            else k_shade<false><<<shade_grid, SHADE_BLOCK, 0, stream>>>(fp, b);
        }
    }
    // No light samples are taken at the last bounce (renderer.hpp:127-130),
    // so no shadow rays are left over.
    k_accumulate<<<acc_grid, 256, 0, stream>>>(fp);
    k_finish<<<1, 64, 0, stream>>>(fp);
}

// Training weight of record k of path p, F/q = lum((final radiance - base)
// / thr): the radiance that arrived along the sampled direction, times the
// BSDF and cosine, over the pdf it was sampled with. Per block of 256
// records: how many have weight, and their sum (fixed-order reductions).
__global__ __launch_bounds__(GUIDE_TARGET_BLOCK) void k_guide_weights(FrameParams fp) {
    __shared__ uint32_t wc[GUIDE_TARGET_BLOCK / 64];
    __shared__ float wsum[GUIDE_TARGET_BLOCK / 64];
    const uint32_t nrec = fp.P * GUIDE_RECORDS;
    const uint32_t r = blockIdx.x * GUIDE_TARGET_BLOCK + threadIdx.x;
    float w = 0;
    if (r < nrec) {
        const uint32_t p = r / GUIDE_RECORDS, k = r % GUIDE_RECORDS;
        const float4 B = fp.pb[p];
        if (k < ((__float_as_uint(B.w) >> 1) & 3u)) {
            const float4 r2 = fp.grec[static_cast<size_t>(r) * 4 + 2], r3 = fp.grec[static_cast<size_t>(r) * 4 + 3];
            const float tx = r2.z, ty = r2.w, tz = r3.x;
            const float ax = fmaxf(B.x - r3.y, 0.0f), ay = fmaxf(B.y - r3.z, 0.0f), az = fmaxf(B.z - r3.w, 0.0f);
            const float ex = tx > 0 ? ax / tx : 0.0f, ey = ty > 0 ? ay / ty : 0.0f, ez = tz > 0 ? az / tz : 0.0f;
            w = 0.2126f * ex + 0.7152f * ey + 0.0722f * ez;
            if (!(w > 0.0f) || !(w < 3.0e38f)) w = 0.0f;   // zero, NaN and infinity carry nothing
        }
        fp.gwt[r] = w;
    }
    const uint32_t c = static_cast<uint32_t>(__popcll(__ballot(w > 0.0f)));
    float sm = w;
    for (int off = 32; off > 0; off >>= 1) sm += __shfl_xor(sm, off, 64);
    if (lane_id() == 0) {
        wc[threadIdx.x >> 6] = c;
        wsum[threadIdx.x >> 6] = sm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tc = 0;
        float ts = 0;
        for (int i = 0; i < GUIDE_TARGET_BLOCK / 64; ++i) {
            tc += wc[i];
            ts += wsum[i];
        }
        fp.gblk[blockIdx.x] = tc;
        fp.gblk_sum[blockIdx.x] = ts;
    }
}

// The blocks' counts become offsets (exclusive scan, one block, each thread
// a contiguous run of blocks); gtot gets the count and the weight sum.
__global__ __launch_bounds__(1024) void k_guide_scan(FrameParams fp, uint32_t nblk) {
    __shared__ uint32_t part[1024];
    __shared__ float psum[1024];
    const uint32_t chunk = (nblk + 1023) / 1024;
    const uint32_t lo = min(threadIdx.x * chunk, nblk), hi = min(lo + chunk, nblk);
    uint32_t c = 0;
    float sm = 0;
    for (uint32_t i = lo; i < hi; ++i) {
        c += fp.gblk[i];
        sm += fp.gblk_sum[i];
    }
    part[threadIdx.x] = c;
    psum[threadIdx.x] = sm;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        float total = 0;
        for (int t = 0; t < 1024; ++t) {
            const uint32_t v = part[t];
            part[t] = run;
            run += v;
            total += psum[t];
        }
        fp.gtot[0] = run;
        fp.gtot[1] = __float_as_uint(total);
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t v = fp.gblk[i];
        fp.gblk[i] = run;
        run += v;
    }
}

// The weighted records, in record order, as training samples: (point,
// weight / batch mean, clamped), (normal, 0), (direction, 0).
__global__ __launch_bounds__(GUIDE_TARGET_BLOCK) void k_guide_compact(FrameParams fp) {
    __shared__ uint32_t wc[GUIDE_TARGET_BLOCK / 64];
    const uint32_t nrec = fp.P * GUIDE_RECORDS;
    const uint32_t r = blockIdx.x * GUIDE_TARGET_BLOCK + threadIdx.x;
    const float w = r < nrec ? fp.gwt[r] : 0.0f;
    const unsigned long long m = __ballot(w > 0.0f);
    if (lane_id() == 0) wc[threadIdx.x >> 6] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (!(w > 0.0f)) return;
    uint32_t slot = fp.gblk[blockIdx.x];
    for (uint32_t i = 0; i < (threadIdx.x >> 6); ++i) slot += wc[i];
    slot += static_cast<uint32_t>(__popcll(m & ((1ull << lane_id()) - 1ull)));
    const float total = __uint_as_float(fp.gtot[1]);
    const float mean = total / static_cast<float>(fp.gtot[0]);
    const float ws = fminf(w / mean, fp.gclamp);
    const float4* R = fp.grec + static_cast<size_t>(r) * 4;
    const float4 r0 = R[0], r1 = R[1], r2 = R[2];
    float4* S = fp.gsamples + static_cast<size_t>(slot) * 3;
    S[0] = make_float4(r0.x, r0.y, r0.z, ws);
    S[1] = make_float4(r1.x, r1.y, r1.z, 0.0f);
    S[2] = make_float4(r1.w, r2.x, r2.y, 0.0f);
}

hipError_t launch_guide_targets(const FrameParams& fp, hipStream_t stream) {
    const uint32_t nrec = fp.P * GUIDE_RECORDS;
    const uint32_t nblk = (nrec + GUIDE_TARGET_BLOCK - 1) / GUIDE_TARGET_BLOCK;
    k_guide_weights<<<nblk, GUIDE_TARGET_BLOCK, 0, stream>>>(fp);
    k_guide_scan<<<1, 1024, 0, stream>>>(fp, nblk);
    k_guide_compact<<<nblk, GUIDE_TARGET_BLOCK, 0, stream>>>(fp);
    return hipGetLastError();
}

hipError_t launch_batch(const FrameParams& fp, int stack_entries, int trace_grid, int shade_grid, hipStream_t stream,
                        hipEvent_t* ev) {
    if (fp.depth < 1 || fp.depth >= MAX_BOUNCES) return hipErrorInvalidValue;
    const bool inst = fp.sc.inst_root != NONE;
    switch (stack_entries) {
        case 8:
            if (inst) launch_loop<8, true>(fp, trace_grid, shade_grid, stream, ev);
            else launch_loop<8, false>(fp, trace_grid, shade_grid, stream, ev);
            break;
        case 16:
            if (inst) launch_loop<16, true>(fp, trace_grid, shade_grid, stream, ev);
            else launch_loop<16, false>(fp, trace_grid, shade_grid, stream, ev);
            break;
        case 32:
            if (inst) launch_loop<32, true>(fp, trace_grid, shade_grid, stream, ev);
            else launch_loop<32, false>(fp, trace_grid, shade_grid, stream, ev);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int S>
static void launch_pool_loop(const FrameParams& fp, int trace_grid, int shade_grid, hipStream_t stream) {
    for (int b = 0; b < fp.depth; ++b) {
        if (b > 0) {
            k_tail_shadow<S, false, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            k_tail<S, false, false, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            k_extend<S, false, false, false, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            k_shade<false, false, true><<<shade_grid, SHADE_BLOCK, 0, stream>>>(fp, b);
        } else {
            if (fp.camera_packets) k_camera_packet_pool<<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            else k_extend<S, false, false, true, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            k_shade<true, false, true><<<shade_grid, SHADE_BLOCK, 0, stream>>>(fp, b);
        }
    }
    k_accumulate_pool<<<shade_grid * 8, 256, 0, stream>>>(fp);
    k_finish<<<1, 64, 0, stream>>>(fp);
}

hipError_t launch_pool_pass(const FrameParams& fp, int stack_entries, int trace_grid, int shade_grid,
                            hipStream_t stream) {
    if (fp.depth < 1 || fp.depth >= MAX_BOUNCES || fp.n_slices < 1 || !fp.slices || !fp.slice_path ||
        !fp.slice_pix || !fp.slice_tile || fp.guide_mode != 0 || fp.want_aux || fp.sc.inst_root != NONE)
        return hipErrorInvalidValue;
    switch (stack_entries) {
        case 8: launch_pool_loop<8>(fp, trace_grid, shade_grid, stream); break;
        case 16: launch_pool_loop<16>(fp, trace_grid, shade_grid, stream); break;
        case 32: launch_pool_loop<32>(fp, trace_grid, shade_grid, stream); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Progressive renders: a step's image (the average of its n samples, 3
// floats a pixel) added into the running sum, and the running average out.
__global__ __launch_bounds__(256) void k_prog_add(const float* img, float* sum, uint32_t n3, float n, int first) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n3) return;
    const float v = img[k] * n;
    sum[k] = first ? v : sum[k] + v;
}

__global__ __launch_bounds__(256) void k_prog_resolve(const float* sum, float* rgb, uint32_t n3, float done) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n3) return;
    rgb[k] = sum[k] / done;
}

hipError_t launch_progressive_add(const float* img, float* sum, uint32_t npix, int n, bool first, hipStream_t stream) {
    const uint32_t n3 = 3 * npix;
    k_prog_add<<<(n3 + 255) / 256, 256, 0, stream>>>(img, sum, n3, static_cast<float>(n), first ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_progressive_resolve(const float* sum, float* rgb, uint32_t npix, int done, hipStream_t stream) {
    const uint32_t n3 = 3 * npix;
    k_prog_resolve<<<(n3 + 255) / 256, 256, 0, stream>>>(sum, rgb, n3, static_cast<float>(done));
    return hipGetLastError();
}

hipError_t launch_resolve(const FrameParams& fp, float* rgb, float* aux, hipStream_t stream) {
    k_resolve<<<(fp.npix + 255) / 256, 256, 0, stream>>>(fp, rgb, aux);
    return hipGetLastError();
}

template <int S, bool INST>
static int extend_occupancy() {
    int blocks = 0;
    const hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_extend<S, false, INST, false>, TRACE_BLOCK, 0);
    return e == hipSuccess && blocks > 0 ? blocks : 1;
}

int path_blocks_per_cu(int stack_entries, bool inst) {
    switch (stack_entries) {
        case 8: return inst ? extend_occupancy<8, true>() : extend_occupancy<8, false>();
        case 16: return inst ? extend_occupancy<16, true>() : extend_occupancy<16, false>();
        case 32: return inst ? extend_occupancy<32, true>() : extend_occupancy<32, false>();
        default: return 1;
    }
}

hipError_t launch_clear_frame(const FrameParams& fp, hipStream_t stream) {
    k_clear_frame<<<(fp.npix + 255) / 256, 256, 0, stream>>>(fp);
    return hipGetLastError();
}

}  // namespace ptd
