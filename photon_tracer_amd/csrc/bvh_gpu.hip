// Device BVH builder. Same tree as bvh::build (bvh.hpp:39-52,126-247) and
// the host builder, node for node and word for word:
//
//  * Segments longer than kSmall are split level by level across the whole
//    GPU: each level bounds, bins and partitions every open segment, chunk by
//    chunk (kChunk primitives per workgroup), with float min/max merged as
//    order-preserving integer keys (min and max do not depend on the order
//    they are taken in, so the boxes equal the reference's sequential grows).
//  * Segments of at most kSmall primitives are finished by one wave each,
//    depth first, with the segment's boxes and centroids held in LDS.
//  * The 4-wide collapse (bvh.hpp:205-242) then runs over all binary nodes
//    at once: each node's widest-first child list, the collapsed subtree
//    sizes bottom-up and the reference's preorder numbering top-down, one
//    launch per binary depth.
//
// The split arithmetic is the reference's, operation for operation, with no
// FMA contraction (the build flags turn it off): bin_of, half_area and the
// SAH cost round as they do on the host. std::partition's element order
// (libstdc++'s two-ended swap for bidirectional iterators, which the
// reference and the host builder run) is reproduced exactly: the k-th
// element from the left that belongs right is swapped with the k-th element
// from the right that belongs left, for every k up to the number of such
// pairs.
#include "bvh_gpu.hpp"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace ptd {
namespace bvhg {
namespace {

constexpr int kBins = 16;           // bvh.hpp:116-120
constexpr uint32_t kMaxLeaf = 8;
constexpr uint32_t kMaxDepth = 64;
constexpr float kTraversalCost = 1, kIntersectionCost = 1;

// Longest segment one wave finishes in LDS. The wave's work is a chain of
// dependent LDS steps per node, so what pays is waves per CU: 128 keeps the
// wave's LDS near 8 KB. For 1M boxes the subtree kernel took 19.6 ms at 1024
// (46 KB, 3 waves per CU), 3.5 ms at 256 (14 KB) and 3.0 ms at 128, where one
// more level pass costs 0.13 ms.
constexpr uint32_t kSmall = 128;
constexpr uint32_t kChunk = 4096;   // primitives per workgroup in the level passes
constexpr int kBlock = 256;
constexpr int kBinWords = 3 * kBins * 7;   // per axis and bin: min xyz, max xyz (keys), count

constexpr uint32_t KEY_POS_INF = 0xff800000u;   // fkey(+inf): the empty box's min
constexpr uint32_t KEY_NEG_INF = 0x007fffffu;   // fkey(-inf): the empty box's max

struct BNode {   // binary node (bvh.hpp:109-113)
    float lo[3], hi[3];
    uint32_t first, count;
};

struct alignas(16) WNode {   // 4-wide node (bvh.hpp:27-32)
    float lo[3][4];
    float hi[3][4];
    uint32_t first[4];
    uint32_t count[4];
};
static_assert(sizeof(WNode) == 128, "node layout");

struct Seg {   // one open segment of a level pass
    uint32_t first, count, depth, node;
    uint32_t chunk_begin, chunk_count;
    uint32_t kind;   // k_decide: 0 leaf, 1 partition on (axis, bin), 2 halve
    int32_t axis, bin;
    uint32_t mid, left;
    float lo, extent;   // centroid bounds along the split axis
    uint32_t pad[3];
};

struct SmallSeg {   // a segment one wave finishes
    uint32_t first, count, depth, node;
    uint32_t base;   // its subtree's nodes: [base, base + 2 * count - 2)
    uint32_t pad[3];
};

constexpr uint32_t NO_DEPTH = 0xffffffffu;   // depth_of of an unused node slot

// The level passes' control words on the device, so the host need not read
// each level's segments back: the open segments and chunks of the level
// with parity p, the small segments so far, an error flag (a list would
// overflow: cannot happen, the lists are sized for disjoint segments), and
// the node slots the whole tree takes.
enum : uint32_t { LC_NSEG = 0, LC_NCH = 2, LC_NSMALL = 4, LC_ERR = 5, LC_TOTAL = 6, LC_WORDS = 8 };

struct SegAcc {
    uint32_t key[12];   // box min, box max, centroid min, centroid max
    uint32_t bins[kBinWords];
};

// Order-preserving map of floats to unsigned integers (no NaN arrives here).
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fval(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

struct FBox {
    float lo[3], hi[3];
};

__device__ __forceinline__ void box_empty(FBox& b) {
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = __builtin_inff();
        b.hi[a] = -__builtin_inff();
    }
}

// aabb::grow (aabb.hpp:15-20): std::min(a, b) keeps a unless b < a.
__device__ __forceinline__ void box_grow(FBox& b, const float* lo, const float* hi) {
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = lo[a] < b.lo[a] ? lo[a] : b.lo[a];
        b.hi[a] = b.hi[a] < hi[a] ? hi[a] : b.hi[a];
    }
}

// aabb::half_area (aabb.hpp:26-30).
__device__ __forceinline__ float half_area(const FBox& b) {
    if (b.lo[0] > b.hi[0]) return 0;
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// bvh::bin_of (bvh.hpp:244-247).
__device__ __forceinline__ int bin_of(float c, float lo, float extent) {
    const int b = static_cast<int>(kBins * (c - lo) / extent);
    return b < 0 ? 0 : b >= kBins ? kBins - 1 : b;
}

__device__ __forceinline__ void load_box(const float* __restrict__ boxes, uint32_t p, float* lo, float* hi) {
    const float2* b = reinterpret_cast<const float2*>(boxes + 6 * size_t(p));
    const float2 a = b[0], c = b[1], d = b[2];
    lo[0] = a.x;
    lo[1] = a.y;
    lo[2] = c.x;
    hi[0] = c.y;
    hi[1] = d.x;
    hi[2] = d.y;
}

// One axis of the SAH sweep (bvh.hpp:149-170) over its 16 bins (7 words
// each), taking a boundary only when it is strictly cheaper than best_cost.
__device__ __forceinline__ void sweep_axis(const uint32_t* bins, float parent_area, int axis, float& best_cost,
                                           int& best_axis, int& best_bin) {
    float right_area[kBins];
    uint32_t right_count[kBins];
    FBox acc;
    box_empty(acc);
    uint32_t n = 0;
#pragma unroll
    for (int b = kBins - 1; b > 0; --b) {
        const uint32_t* w = bins + 7 * b;
        const float lo[3] = {fval(w[0]), fval(w[1]), fval(w[2])};
        const float hi[3] = {fval(w[3]), fval(w[4]), fval(w[5])};
        box_grow(acc, lo, hi);
        n += w[6];
        right_area[b] = half_area(acc);
        right_count[b] = n;
    }
    box_empty(acc);
    n = 0;
#pragma unroll
    for (int b = 0; b < kBins - 1; ++b) {
        const uint32_t* w = bins + 7 * b;
        const float lo[3] = {fval(w[0]), fval(w[1]), fval(w[2])};
        const float hi[3] = {fval(w[3]), fval(w[4]), fval(w[5])};
        box_grow(acc, lo, hi);
        n += w[6];
        if (n == 0 || right_count[b + 1] == 0) continue;
        const float cost = kTraversalCost + kIntersectionCost *
                                                (half_area(acc) * n + right_area[b + 1] * right_count[b + 1]) /
                                                parent_area;
        if (cost < best_cost) {
            best_cost = cost;
            best_axis = axis;
            best_bin = b;
        }
    }
}

// Lane exchanges within rows of 16 lanes by DPP (a VALU operand modifier,
// no trip through LDS): quad_perm swaps of neighbours one and two apart,
// mirrors within half rows and rows, and shifts along a row. A source lane
// outside the row leaves the lane its own value. Across rows, shuffles.
constexpr int DPP_QUAD_XOR1 = 0xB1;         // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_XOR2 = 0x4E;         // quad_perm [2, 3, 0, 1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141;  // lane i <- 7 - i within each half row
constexpr int DPP_ROW_MIRROR = 0x140;       // lane i <- 15 - i within each row
constexpr int DPP_ROW_SHL = 0x100;          // + n: lane i <- lane i + n of its row
constexpr int DPP_ROW_SHR = 0x110;          // + n: lane i <- lane i - n of its row

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u(uint32_t x) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(x), static_cast<int>(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float x) {
    return __uint_as_float(dpp_u<CTRL>(__float_as_uint(x)));
}

// Reductions over the wave: every lane gets the result. The lanes are
// combined in pairs whose values are the same whatever order min, max and
// integer sums take them in.
template <class Op, class T, class X>
__device__ __forceinline__ T wave_reduce(T x, Op op, X exch) {
    x = op(x, exch(std::integral_constant<int, DPP_QUAD_XOR1>(), x));
    x = op(x, exch(std::integral_constant<int, DPP_QUAD_XOR2>(), x));
    x = op(x, exch(std::integral_constant<int, DPP_ROW_HALF_MIRROR>(), x));
    x = op(x, exch(std::integral_constant<int, DPP_ROW_MIRROR>(), x));
    x = op(x, __shfl_xor(x, 16, 64));
    x = op(x, __shfl_xor(x, 32, 64));
    return x;
}
struct FExch {
    template <int C>
    __device__ __forceinline__ float operator()(std::integral_constant<int, C>, float v) const { return dpp_f<C>(v); }
};
struct UExch {
    template <int C>
    __device__ __forceinline__ uint32_t operator()(std::integral_constant<int, C>, uint32_t v) const {
        return dpp_u<C>(v);
    }
};
__device__ __forceinline__ float wave_min(float x) {
    return wave_reduce(x, [](float a, float b) { return b < a ? b : a; }, FExch());
}
__device__ __forceinline__ float wave_max(float x) {
    return wave_reduce(x, [](float a, float b) { return a < b ? b : a; }, FExch());
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
    return wave_reduce(x, [](uint32_t a, uint32_t b) { return a + b; }, UExch());
}

// Box and centroid bounds of 12 running values: [0,3) box min, [3,6) box
// max, [6,9) centroid min, [9,12) centroid max.
__device__ __forceinline__ void bounds_init(float* v) {
    for (int a = 0; a < 3; ++a) {
        v[a] = __builtin_inff();
        v[3 + a] = -__builtin_inff();
        v[6 + a] = __builtin_inff();
        v[9 + a] = -__builtin_inff();
    }
}
__device__ __forceinline__ void bounds_add(float* v, const float* lo, const float* hi, const float* c) {
    for (int a = 0; a < 3; ++a) {
        v[a] = lo[a] < v[a] ? lo[a] : v[a];
        v[3 + a] = v[3 + a] < hi[a] ? hi[a] : v[3 + a];
        v[6 + a] = c[a] < v[6 + a] ? c[a] : v[6 + a];
        v[9 + a] = v[9 + a] < c[a] ? c[a] : v[9 + a];
    }
}
__device__ __forceinline__ void bounds_wave(float* v) {
    for (int a = 0; a < 3; ++a) {
        v[a] = wave_min(v[a]);
        v[3 + a] = wave_max(v[3 + a]);
        v[6 + a] = wave_min(v[6 + a]);
        v[9 + a] = wave_max(v[9 + a]);
    }
}

__device__ __forceinline__ uint32_t bin_init_word(int i) {
    const int k = i % 7;
    return k < 3 ? KEY_POS_INF : k < 6 ? KEY_NEG_INF : 0u;
}

// ---------------------------------------------------------------------------
// Set-up: centroids (aabb::center, aabb.hpp:22), identity order, counters,
// and the boxes' flags in counters[2] (zeroed before the launch): BOX_NAN
// if a coordinate is NaN, BOX_NEG_ZERO if one is -0.0, one atomic per wave
// that found either.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_setup(const float* __restrict__ boxes, uint32_t n, float4* centers,
                                                   uint32_t* order, uint32_t* counters, uint32_t* depth_of,
                                                   uint32_t cap) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    for (uint32_t k = i; k < cap; k += gridDim.x * kBlock) depth_of[k] = NO_DEPTH;
    if (i == 0) {
        counters[0] = 1;   // node 0 is the root
        counters[1] = 0;   // deepest binary depth
    }
    uint32_t flags = 0;
    if (i < n) {
        float lo[3], hi[3];
        load_box(boxes, i, lo, hi);
        centers[i] = make_float4((lo[0] + hi[0]) * 0.5f, (lo[1] + hi[1]) * 0.5f, (lo[2] + hi[2]) * 0.5f, 0.f);
        order[i] = i;
        for (int a = 0; a < 3; ++a) {
            flags |= (__builtin_isnan(lo[a]) || __builtin_isnan(hi[a])) ? BOX_NAN : 0u;
            flags |= (__float_as_uint(lo[a]) == 0x80000000u || __float_as_uint(hi[a]) == 0x80000000u) ? BOX_NEG_ZERO
                                                                                                     : 0u;
        }
    }
    const unsigned long long nan_lanes = __ballot((flags & BOX_NAN) != 0);
    const unsigned long long zero_lanes = __ballot((flags & BOX_NEG_ZERO) != 0);
    if ((threadIdx.x & 63) == 0 && (nan_lanes | zero_lanes))
        atomicOr(&counters[2], (nan_lanes ? BOX_NAN : 0u) | (zero_lanes ? BOX_NEG_ZERO : 0u));
}

// ---------------------------------------------------------------------------
// Level passes over segments longer than kSmall.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_seg_init(SegAcc* acc, const uint32_t* lc, int p) {
    if (blockIdx.x >= lc[LC_NSEG + p]) return;
    SegAcc& a = acc[blockIdx.x];
    for (int i = threadIdx.x; i < 12; i += kBlock) a.key[i] = (i % 6) < 3 ? KEY_POS_INF : KEY_NEG_INF;
    for (int i = threadIdx.x; i < kBinWords; i += kBlock) a.bins[i] = bin_init_word(i);
}

__device__ __forceinline__ uint32_t chunk_end(const uint2& ch, const Seg& s) {
    return min(ch.y + kChunk, s.first + s.count);
}

// bvh.hpp:128-132: the segment's box and centroid bounds.
__global__ __launch_bounds__(kBlock) void k_bounds(const float* __restrict__ boxes, const float4* __restrict__ centers,
                                                    const uint32_t* __restrict__ order, const uint2* __restrict__ chunks,
                                                    const Seg* __restrict__ segs, SegAcc* acc,
        const uint32_t* lc, int p) {
    if (blockIdx.x >= lc[LC_NCH + p]) return;
    __shared__ float part[kBlock / 64][12];
    const uint2 ch = chunks[blockIdx.x];
    const Seg s = segs[ch.x];
    const uint32_t end = chunk_end(ch, s);
    float v[12];
    bounds_init(v);
    for (uint32_t i = ch.y + threadIdx.x; i < end; i += kBlock) {
        const uint32_t p = order[i];
        float lo[3], hi[3];
        load_box(boxes, p, lo, hi);
        const float4 c4 = centers[p];
        const float c[3] = {c4.x, c4.y, c4.z};
        bounds_add(v, lo, hi, c);
    }
    bounds_wave(v);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int k = 0; k < 12; ++k) part[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        const bool is_min = (k % 6) < 3;
        float x = part[0][k];
        for (int w = 1; w < kBlock / 64; ++w) {
            const float y = part[w][k];
            x = is_min ? (y < x ? y : x) : (x < y ? y : x);
        }
        if (is_min)
            atomicMin(&acc[ch.x].key[k], fkey(x));
        else
            atomicMax(&acc[ch.x].key[k], fkey(x));
    }
}

// bvh.hpp:141-148: per axis with a positive centroid extent, the 16 bins'
// boxes and counts.
__global__ __launch_bounds__(kBlock) void k_bin(const float* __restrict__ boxes, const float4* __restrict__ centers,
                                                 const uint32_t* __restrict__ order, const uint2* __restrict__ chunks,
                                                 const Seg* __restrict__ segs, SegAcc* acc,
        const uint32_t* lc, int p) {
    if (blockIdx.x >= lc[LC_NCH + p]) return;
    __shared__ uint32_t bins[kBinWords];
    const uint2 ch = chunks[blockIdx.x];
    const Seg s = segs[ch.x];
    const uint32_t end = chunk_end(ch, s);
    float clo[3], cext[3];
    for (int a = 0; a < 3; ++a) {
        clo[a] = fval(acc[ch.x].key[6 + a]);
        cext[a] = fval(acc[ch.x].key[9 + a]) - clo[a];
    }
    for (int i = threadIdx.x; i < kBinWords; i += kBlock) bins[i] = bin_init_word(i);
    __syncthreads();
    for (uint32_t i = ch.y + threadIdx.x; i < end; i += kBlock) {
        const uint32_t p = order[i];
        float lo[3], hi[3];
        load_box(boxes, p, lo, hi);
        const float4 c4 = centers[p];
        const float c[3] = {c4.x, c4.y, c4.z};
        for (int a = 0; a < 3; ++a) {
            if (cext[a] <= 0) continue;
            uint32_t* w = bins + 7 * (a * kBins + bin_of(c[a], clo[a], cext[a]));
            for (int k = 0; k < 3; ++k) {
                atomicMin(w + k, fkey(lo[k]));
                atomicMax(w + 3 + k, fkey(hi[k]));
            }
            atomicAdd(w + 6, 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kBinWords; i += kBlock) {
        const uint32_t v = bins[i];
        const int k = i % 7;
        uint32_t* g = &acc[ch.x].bins[i];
        if (k < 3) {
            if (v != KEY_POS_INF) atomicMin(g, v);
        } else if (k < 6) {
            if (v != KEY_NEG_INF) atomicMax(g, v);
        } else if (v) {
            atomicAdd(g, v);
        }
    }
}

// bvh.hpp:126-191 per segment: the node's box, the leaf test, the SAH split
// and the child pair.
__global__ __launch_bounds__(64) void k_decide(Seg* segs, const SegAcc* __restrict__ acc, const uint32_t* lc, int p,
                                               BNode* nodes, uint32_t* depth_of, uint32_t* counters) {
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= lc[LC_NSEG + p]) return;
    Seg s = segs[i];
    const SegAcc& A = acc[i];
    FBox box, cen;
    for (int a = 0; a < 3; ++a) {
        box.lo[a] = fval(A.key[a]);
        box.hi[a] = fval(A.key[3 + a]);
        cen.lo[a] = fval(A.key[6 + a]);
        cen.hi[a] = fval(A.key[9 + a]);
    }
    BNode& nd = nodes[s.node];
    for (int a = 0; a < 3; ++a) {
        nd.lo[a] = box.lo[a];
        nd.hi[a] = box.hi[a];
    }
    depth_of[s.node] = s.depth;
    atomicMax(&counters[1], s.depth);
    s.kind = 0;
    bool leaf = s.count == 1 || s.depth + 1 >= kMaxDepth;
    float best_cost = __builtin_inff();
    int best_axis = -1, best_bin = 0;
    if (!leaf) {
        const float parent = half_area(box);
        for (int axis = 0; axis < 3; ++axis) {
            const float lo = cen.lo[axis], extent = cen.hi[axis] - lo;
            if (extent <= 0) continue;
            sweep_axis(A.bins + 7 * kBins * axis, parent, axis, best_cost, best_axis, best_bin);
        }
        leaf = s.count <= kMaxLeaf && !(best_cost < s.count * kIntersectionCost);
    }
    if (leaf) {
        nd.first = s.first;
        nd.count = s.count;
        segs[i] = s;
        return;
    }
    if (best_axis >= 0) {
        s.kind = 1;
        s.axis = best_axis;
        s.bin = best_bin;
        s.lo = cen.lo[best_axis];
        s.extent = cen.hi[best_axis] - s.lo;
        uint32_t l = 0;
        for (int b = 0; b <= best_bin; ++b) l += A.bins[7 * (best_axis * kBins + b) + 6];
        s.mid = s.first + l;
    } else {
        s.kind = 2;
        s.mid = s.first + s.count / 2;
    }
    s.left = atomicAdd(&counters[0], 2u);
    nd.first = s.left;
    nd.count = 0;
    segs[i] = s;
}

__device__ __forceinline__ bool goes_left(const float4* __restrict__ centers, uint32_t p, const Seg& s) {
    const float* c = reinterpret_cast<const float*>(centers + p);
    return bin_of(c[s.axis], s.lo, s.extent) <= s.bin;
}

// Block-wide sums of two counters, returned to every thread.
__device__ __forceinline__ void block_sum2(uint32_t& a, uint32_t& b, uint32_t (*part)[2]) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        part[wave][0] = a;
        part[wave][1] = b;
    }
    __syncthreads();
    a = b = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        a += part[w][0];
        b += part[w][1];
    }
}

// Per chunk: elements left of the partition point that belong right, and
// elements right of it that belong left.
__global__ __launch_bounds__(kBlock) void k_part_count(const float4* __restrict__ centers,
                                                        const uint32_t* __restrict__ order,
                                                        const uint2* __restrict__ chunks, const Seg* __restrict__ segs,
                                                        uint32_t* cnt,
        const uint32_t* lc, int p) {
    if (blockIdx.x >= lc[LC_NCH + p]) return;
    __shared__ uint32_t part[kBlock / 64][2];
    const uint2 ch = chunks[blockIdx.x];
    const Seg s = segs[ch.x];
    if (s.kind != 1) return;
    const uint32_t end = chunk_end(ch, s);
    uint32_t nl = 0, nr = 0;
    for (uint32_t i = ch.y + threadIdx.x; i < end; i += kBlock) {
        const bool left = goes_left(centers, order[i], s);
        nl += (i < s.mid && !left) ? 1u : 0u;
        nr += (i >= s.mid && left) ? 1u : 0u;
    }
    block_sum2(nl, nr, part);
    if (threadIdx.x == 0) {
        cnt[2 * blockIdx.x] = nl;
        cnt[2 * blockIdx.x + 1] = nr;
    }
}

// Positions of the misplaced elements in scan order: those left of the
// partition point from the segment's start, those right of it from its end.
__global__ __launch_bounds__(kBlock) void k_part_pos(const float4* __restrict__ centers,
                                                      const uint32_t* __restrict__ order,
                                                      const uint2* __restrict__ chunks, const Seg* __restrict__ segs,
                                                      const uint32_t* __restrict__ cnt, uint32_t* pos,
        const uint32_t* lc, int p) {
    if (blockIdx.x >= lc[LC_NCH + p]) return;
    __shared__ uint32_t part[kBlock / 64][2];
    const uint2 ch = chunks[blockIdx.x];
    const Seg s = segs[ch.x];
    if (s.kind != 1) return;
    const uint32_t end = chunk_end(ch, s);
    uint32_t ol = 0, orr = 0;
    for (uint32_t c = s.chunk_begin + threadIdx.x; c < blockIdx.x; c += kBlock) {
        ol += cnt[2 * c];
        orr += cnt[2 * c + 1];
    }
    block_sum2(ol, orr, part);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1;
    for (uint32_t base = ch.y; base < end; base += kBlock) {
        const uint32_t i = base + threadIdx.x;
        bool lf = false, rt = false;
        if (i < end) {
            const bool left = goes_left(centers, order[i], s);
            lf = i < s.mid && !left;
            rt = i >= s.mid && left;
        }
        const unsigned long long ml = __ballot(lf), mr = __ballot(rt);
        __syncthreads();
        if (lane == 0) {
            part[wave][0] = __popcll(ml);
            part[wave][1] = __popcll(mr);
        }
        __syncthreads();
        uint32_t bl = 0, br = 0, tl = 0, tr = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            if (w < wave) {
                bl += part[w][0];
                br += part[w][1];
            }
            tl += part[w][0];
            tr += part[w][1];
        }
        if (lf) pos[s.first + ol + bl + __popcll(ml & below)] = i;
        if (rt) pos[s.first + s.count - 1 - (orr + br + __popcll(mr & below))] = i;
        ol += tl;
        orr += tr;
    }
}

// The swaps: the r-th misplaced element from the left with the r-th from the
// right.
__global__ __launch_bounds__(kBlock) void k_part_swap(const uint2* __restrict__ chunks, const Seg* __restrict__ segs,
                                                       const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ pos,
                                                       uint32_t* order,
        const uint32_t* lc, int p) {
    if (blockIdx.x >= lc[LC_NCH + p]) return;
    __shared__ uint32_t part[kBlock / 64][2];
    const uint2 ch = chunks[blockIdx.x];
    const Seg s = segs[ch.x];
    if (s.kind != 1) return;
    uint32_t total = 0, unused = 0;
    for (uint32_t c = s.chunk_begin + threadIdx.x; c < s.chunk_begin + s.chunk_count; c += kBlock) total += cnt[2 * c];
    block_sum2(total, unused, part);
    const uint32_t r0 = ch.y - s.first;
    const uint32_t r1 = min(r0 + kChunk, total);
    for (uint32_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
        const uint32_t a = pos[s.first + r], b = pos[s.first + s.count - total + r];
        const uint32_t x = order[a];
        order[a] = order[b];
        order[b] = x;
    }
}

// ---------------------------------------------------------------------------
// The next level's lists, on the device: the children of the level's split
// segments in segment order, those longer than kSmall as the next level's
// segments (with their chunks) and the others appended to the small list,
// the order the host loop that came before this kernel used. One block of
// kPlanBlock threads walks the level's segments in tiles.
// ---------------------------------------------------------------------------
constexpr int kPlanBlock = 1024;

// Exclusive block-wide prefix sums of three counters, and their totals.
__device__ __forceinline__ void block_scan3(uint32_t (&v)[3], uint32_t (&tot)[3], uint32_t (*part)[3]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t x = v[k];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        incl[k] = x;
    }
    __syncthreads();
    if (lane == 63)
        for (int k = 0; k < 3; ++k) part[wave][k] = incl[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t before = 0, all = 0;
        for (int w = 0; w < kPlanBlock / 64; ++w) {
            const uint32_t t = part[w][k];
            before += w < wave ? t : 0u;
            all += t;
        }
        v[k] = before + incl[k] - v[k];
        tot[k] = all;
    }
}

__global__ __launch_bounds__(kPlanBlock) void k_plan(const Seg* __restrict__ cur, uint32_t* lc, int p, Seg* next,
                                                      uint2* next_chunks, SmallSeg* small, uint32_t max_segs,
                                                      uint32_t max_chunks, uint32_t max_small) {
    __shared__ uint32_t part[kPlanBlock / 64][3];
    const uint32_t nseg = lc[LC_NSEG + p];
    uint32_t n_big = 0, n_ch = 0, n_small = lc[LC_NSMALL];
    bool overflow = false;
    for (uint32_t tile = 0; tile < nseg; tile += kPlanBlock) {
        const uint32_t i = tile + threadIdx.x;
        uint32_t first[2] = {0, 0}, count[2] = {0, 0}, node[2] = {0, 0}, depth = 0;
        int kids = 0;
        if (i < nseg) {
            const Seg sg = cur[i];
            if (sg.kind != 0) {
                kids = 2;
                depth = sg.depth + 1;
                first[0] = sg.first;
                count[0] = sg.mid - sg.first;
                node[0] = sg.left;
                first[1] = sg.mid;
                count[1] = sg.first + sg.count - sg.mid;
                node[1] = sg.left + 1;
            }
        }
        uint32_t v[3] = {0, 0, 0}, tot[3];   // big children, their chunks, small children
        for (int c = 0; c < kids; ++c) {
            if (count[c] > kSmall) {
                ++v[0];
                v[1] += (count[c] + kChunk - 1) / kChunk;
            } else {
                ++v[2];
            }
        }
        block_scan3(v, tot, part);
        for (int c = 0; c < kids; ++c) {
            if (count[c] > kSmall) {
                const uint32_t at = n_big + v[0]++, ch0 = n_ch + v[1];
                const uint32_t nc = (count[c] + kChunk - 1) / kChunk;
                v[1] += nc;
                if (at >= max_segs || ch0 + nc > max_chunks) {
                    overflow = true;
                    continue;
                }
                Seg ns{};
                ns.first = first[c];
                ns.count = count[c];
                ns.depth = depth;
                ns.node = node[c];
                ns.chunk_begin = ch0;
                ns.chunk_count = nc;
                next[at] = ns;
                for (uint32_t k = 0; k < nc; ++k) next_chunks[ch0 + k] = make_uint2(at, first[c] + k * kChunk);
            } else {
                const uint32_t at = n_small + v[2]++;
                if (at >= max_small) {
                    overflow = true;
                    continue;
                }
                SmallSeg ss{};
                ss.first = first[c];
                ss.count = count[c];
                ss.depth = depth;
                ss.node = node[c];
                small[at] = ss;
            }
        }
        n_big += tot[0];
        n_ch += tot[1];
        n_small += tot[2];
    }
    if (overflow) lc[LC_ERR] = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        lc[LC_NSEG + (p ^ 1)] = n_big;
        lc[LC_NCH + (p ^ 1)] = n_ch;
        lc[LC_NSMALL] = n_small;
    }
}

// Each small segment's node range, [base, base + 2 count - 2), after the
// level passes' nodes (counters[0] of them), in list order; lc[LC_TOTAL]:
// every node slot the tree takes.
__global__ __launch_bounds__(kPlanBlock) void k_small_bases(SmallSeg* small, uint32_t* lc, const uint32_t* counters) {
    __shared__ uint32_t part[kPlanBlock / 64][3];
    const uint32_t ns = lc[LC_NSMALL];
    uint32_t carry = counters[0];
    for (uint32_t tile = 0; tile < ns; tile += kPlanBlock) {
        const uint32_t i = tile + threadIdx.x;
        const uint32_t c = i < ns ? small[i].count : 1u;
        uint32_t v[3] = {2 * c - 2, 0, 0}, tot[3];
        block_scan3(v, tot, part);
        if (i < ns) small[i].base = carry + v[0];
        carry += tot[0];
    }
    if (threadIdx.x == 0) lc[LC_TOTAL] = carry;
}

// ---------------------------------------------------------------------------
// Segments of at most kSmall primitives: one wave builds the whole subtree,
// depth first (bvh.hpp:126-197), on LDS copies of its boxes and centroids.
// Its new nodes come from its own range (a subtree over c primitives has at
// most 2c - 2 nodes below its root), not from a shared counter: one global
// atomic per split serialised the waves.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_small(const float* __restrict__ boxes, const float4* __restrict__ centers,
                                              uint32_t* order, const SmallSeg* __restrict__ small, BNode* nodes,
                                              uint32_t* depth_of, uint32_t* counters) {
    __shared__ float s_lo[3][kSmall], s_hi[3][kSmall], s_c[3][kSmall];
    __shared__ uint32_t s_order[kSmall];
    __shared__ uint16_t s_perm[kSmall], s_lst[kSmall];
    __shared__ uint32_t s_bins[kBinWords];
    __shared__ uint32_t s_stack[4][kMaxDepth + 8];   // first, count, depth, node

    const SmallSeg sg = small[blockIdx.x];
    const uint32_t g0 = sg.first, total = sg.count;
    uint32_t next = sg.base;
    const int lane = threadIdx.x;
    for (uint32_t i = lane; i < total; i += 64) {
        const uint32_t p = order[g0 + i];
        float lo[3], hi[3];
        load_box(boxes, p, lo, hi);
        const float4 c = centers[p];
        s_order[i] = p;
        for (int a = 0; a < 3; ++a) {
            s_lo[a][i] = lo[a];
            s_hi[a][i] = hi[a];
        }
        s_c[0][i] = c.x;
        s_c[1][i] = c.y;
        s_c[2][i] = c.z;
        s_perm[i] = static_cast<uint16_t>(i);
    }
    if (lane == 0) {
        s_stack[0][0] = 0;
        s_stack[1][0] = total;
        s_stack[2][0] = sg.depth;
        s_stack[3][0] = sg.node;
    }
    uint32_t sp = 1, deepest = 0;
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1;
    while (sp > 0) {
        --sp;
        const uint32_t f = s_stack[0][sp], c = s_stack[1][sp], d = s_stack[2][sp], node = s_stack[3][sp];
        __syncthreads();
        deepest = max(deepest, d);
        float v[12];
        bounds_init(v);
        for (uint32_t k = f + lane; k < f + c; k += 64) {
            const uint32_t li = s_perm[k];
            const float lo[3] = {s_lo[0][li], s_lo[1][li], s_lo[2][li]};
            const float hi[3] = {s_hi[0][li], s_hi[1][li], s_hi[2][li]};
            const float cc[3] = {s_c[0][li], s_c[1][li], s_c[2][li]};
            bounds_add(v, lo, hi, cc);
        }
        bounds_wave(v);
        if (lane == 0) {
            for (int a = 0; a < 3; ++a) {
                nodes[node].lo[a] = v[a];
                nodes[node].hi[a] = v[3 + a];
            }
            depth_of[node] = d;
        }
        if (c == 1 || d + 1 >= kMaxDepth) {
            if (lane == 0) {
                nodes[node].first = g0 + f;
                nodes[node].count = c;
            }
            continue;
        }
        float clo[3], cext[3];
        for (int a = 0; a < 3; ++a) {
            clo[a] = v[6 + a];
            cext[a] = v[9 + a] - clo[a];
        }
        for (int i = lane; i < kBinWords; i += 64) s_bins[i] = bin_init_word(i);
        __syncthreads();
        for (uint32_t k = f + lane; k < f + c; k += 64) {
            const uint32_t li = s_perm[k];
            for (int a = 0; a < 3; ++a) {
                if (cext[a] <= 0) continue;
                uint32_t* w = s_bins + 7 * (a * kBins + bin_of(s_c[a][li], clo[a], cext[a]));
                for (int q = 0; q < 3; ++q) {
                    atomicMin(w + q, fkey(s_lo[q][li]));
                    atomicMax(w + 3 + q, fkey(s_hi[q][li]));
                }
                atomicAdd(w + 6, 1u);
            }
        }
        __syncthreads();
        // The SAH sweep (bvh.hpp:149-170) on 48 lanes, one bin boundary each
        // (lane = 16 * axis + bin): prefix and suffix boxes and counts by
        // shuffles within each axis's 16 lanes (min, max and integer sums do
        // not depend on the order they are taken in), each boundary's cost
        // by the reference's expression, then the cheapest boundary, the
        // first of equals in (axis, bin) order: the one the reference's
        // sequential strict-less scan keeps.
        FBox box;
        for (int a = 0; a < 3; ++a) {
            box.lo[a] = v[a];
            box.hi[a] = v[3 + a];
        }
        const int sa = lane >> 4, sb = lane & 15;
        const float ext_l = sa == 0 ? cext[0] : sa == 1 ? cext[1] : cext[2];
        const bool axis_ok = lane < 48 && !(ext_l <= 0);
        FBox pre, suf;
        uint32_t pc = 0;
        box_empty(pre);
        if (lane < 48) {
            const uint32_t* w = s_bins + 7 * lane;
            for (int q = 0; q < 3; ++q) {
                pre.lo[q] = fval(w[q]);
                pre.hi[q] = fval(w[3 + q]);
            }
            pc = w[6];
        }
        suf = pre;
        uint32_t sc = pc;
        // One step of the scans: lanes dl or more into the row take the
        // prefix dl lanes before them, lanes dl or more from its end the
        // suffix dl lanes after them (row shifts; a row is an axis).
        auto scan_step = [&](auto dlc) {
            constexpr int dl = decltype(dlc)::value;
            FBox up, dn;
            for (int q = 0; q < 3; ++q) {
                up.lo[q] = dpp_f<DPP_ROW_SHR + dl>(pre.lo[q]);
                up.hi[q] = dpp_f<DPP_ROW_SHR + dl>(pre.hi[q]);
                dn.lo[q] = dpp_f<DPP_ROW_SHL + dl>(suf.lo[q]);
                dn.hi[q] = dpp_f<DPP_ROW_SHL + dl>(suf.hi[q]);
            }
            const uint32_t upc = dpp_u<DPP_ROW_SHR + dl>(pc), dnc = dpp_u<DPP_ROW_SHL + dl>(sc);
            if (sb >= dl) {
                box_grow(pre, up.lo, up.hi);
                pc += upc;
            }
            if (sb + dl < kBins) {
                box_grow(suf, dn.lo, dn.hi);
                sc += dnc;
            }
        };
        static_assert(kBins == 16, "one bin a lane of a 16-lane row");
        scan_step(std::integral_constant<int, 1>());
        scan_step(std::integral_constant<int, 2>());
        scan_step(std::integral_constant<int, 4>());
        scan_step(std::integral_constant<int, 8>());
        FBox right;
        for (int q = 0; q < 3; ++q) {
            right.lo[q] = dpp_f<DPP_ROW_SHL + 1>(suf.lo[q]);
            right.hi[q] = dpp_f<DPP_ROW_SHL + 1>(suf.hi[q]);
        }
        const uint32_t rc = dpp_u<DPP_ROW_SHL + 1>(sc);
        float cand_cost = __builtin_inff();
        int cand = 64;
        if (axis_ok && sb < kBins - 1 && pc != 0 && rc != 0) {
            const float cost = kTraversalCost + kIntersectionCost *
                                                    (half_area(pre) * pc + half_area(right) * rc) / half_area(box);
            if (cost < __builtin_inff()) {
                cand_cost = cost;
                cand = lane;
            }
        }
        // The least (cost, lane) over the wave: the pairs are combined in an
        // order that does not change the least.
        auto take = [&](float oc, int oi) {
            if (oi < 64 && (cand >= 64 || oc < cand_cost || (oc == cand_cost && oi < cand))) {
                cand_cost = oc;
                cand = oi;
            }
        };
        auto take_dpp = [&](auto ctrl) {
            constexpr int C = decltype(ctrl)::value;
            take(dpp_f<C>(cand_cost), static_cast<int>(dpp_u<C>(static_cast<uint32_t>(cand))));
        };
        take_dpp(std::integral_constant<int, DPP_QUAD_XOR1>());
        take_dpp(std::integral_constant<int, DPP_QUAD_XOR2>());
        take_dpp(std::integral_constant<int, DPP_ROW_HALF_MIRROR>());
        take_dpp(std::integral_constant<int, DPP_ROW_MIRROR>());
        take(__shfl_xor(cand_cost, 16, 64), __shfl_xor(cand, 16, 64));
        take(__shfl_xor(cand_cost, 32, 64), __shfl_xor(cand, 32, 64));
        cand = __builtin_amdgcn_readfirstlane(cand);   // the same on every lane
        const float best_cost = cand < 64 ? cand_cost : __builtin_inff();
        const int best_axis = cand < 64 ? cand >> 4 : -1, best_bin = cand & 15;
        const uint32_t left_count = __builtin_amdgcn_readlane(pc, cand < 64 ? cand : 0);
        if (c <= kMaxLeaf && !(best_cost < c * kIntersectionCost)) {
            if (lane == 0) {
                nodes[node].first = g0 + f;
                nodes[node].count = c;
            }
            continue;
        }
        uint32_t mid;
        if (best_axis >= 0) {
            mid = f + left_count;
            const float lo = clo[best_axis], ext = cext[best_axis];
            uint32_t nl = 0, nr = 0;
            for (uint32_t base = f; base < f + c; base += 64) {
                const uint32_t k = base + lane;
                bool lf = false, rt = false;
                if (k < f + c) {
                    const bool left = bin_of(s_c[best_axis][s_perm[k]], lo, ext) <= best_bin;
                    lf = k < mid && !left;
                    rt = k >= mid && left;
                }
                const unsigned long long ml = __ballot(lf), mr = __ballot(rt);
                if (lf) s_lst[f + nl + __popcll(ml & below)] = static_cast<uint16_t>(k);
                if (rt) s_lst[f + c - 1 - (nr + __popcll(mr & below))] = static_cast<uint16_t>(k);
                nl += __popcll(ml);
                nr += __popcll(mr);
            }
            __syncthreads();
            for (uint32_t r = lane; r < nl; r += 64) {
                const uint32_t a = s_lst[f + r], b = s_lst[f + c - nl + r];
                const uint16_t x = s_perm[a];
                s_perm[a] = s_perm[b];
                s_perm[b] = x;
            }
            __syncthreads();
        } else {
            mid = f + c / 2;
        }
        const uint32_t left = next;
        next += 2;
        if (lane == 0) {
            nodes[node].first = left;
            nodes[node].count = 0;
            s_stack[0][sp] = mid;
            s_stack[1][sp] = f + c - mid;
            s_stack[2][sp] = d + 1;
            s_stack[3][sp] = left + 1;
            s_stack[0][sp + 1] = f;
            s_stack[1][sp + 1] = mid - f;
            s_stack[2][sp + 1] = d + 1;
            s_stack[3][sp + 1] = left;
        }
        sp += 2;
        __syncthreads();
    }
    __syncthreads();
    for (uint32_t i = lane; i < total; i += 64) order[g0 + i] = s_order[s_perm[i]];
    if (lane == 0) atomicMax(&counters[1], deepest);
}

// ---------------------------------------------------------------------------
// 4-wide collapse (bvh.hpp:205-242).
// ---------------------------------------------------------------------------
__device__ __forceinline__ FBox node_box(const BNode& b) {
    FBox x;
    for (int a = 0; a < 3; ++a) {
        x.lo[a] = b.lo[a];
        x.hi[a] = b.hi[a];
    }
    return x;
}

// Every node's child list as collapse(b) would open it: the interior child
// of largest area (the first of equals) until four or only leaves remain.
__global__ __launch_bounds__(kBlock) void k_clist(const BNode* __restrict__ nodes, const uint32_t* __restrict__ depth_of,
                                                   uint32_t total, int4* clist) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= total || depth_of[b] == NO_DEPTH) return;
    int child[4] = {static_cast<int>(b), -1, -1, -1};
    int n = 1;
    while (n < 4) {
        int widest = -1;
        float widest_area = -1;
        for (int i = 0; i < n; ++i) {
            const BNode c = nodes[child[i]];
            if (c.count != 0) continue;
            const float area = half_area(node_box(c));
            if (area > widest_area) {
                widest = i;
                widest_area = area;
            }
        }
        if (widest < 0) break;
        const uint32_t left = nodes[child[widest]].first;
        child[widest] = static_cast<int>(left);
        child[n++] = static_cast<int>(left + 1);
    }
    clist[b] = make_int4(child[0], child[1], child[2], child[3]);
}

__device__ __forceinline__ int child_at(const int4& c, int i) {
    return i == 0 ? c.x : i == 1 ? c.y : i == 2 ? c.z : c.w;
}

// Collapsed subtree sizes of the interior nodes at binary depth d (their
// collapsed children are deeper, so already sized).
__global__ __launch_bounds__(kBlock) void k_wsize(const BNode* __restrict__ nodes, const uint32_t* __restrict__ depth_of,
                                                   const int4* __restrict__ clist, uint32_t total, uint32_t d,
                                                   uint32_t* wsize) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= total || depth_of[b] != d || nodes[b].count != 0) return;
    const int4 cl = clist[b];
    uint32_t w = 1;
    for (int i = 0; i < 4; ++i) {
        const int c = child_at(cl, i);
        if (c >= 0 && nodes[c].count == 0) w += wsize[c];
    }
    wsize[b] = w;
}

// Preorder numbers (collapse's emplace order) of the collapsed children of
// the numbered interior nodes at binary depth d.
__global__ __launch_bounds__(kBlock) void k_index(const BNode* __restrict__ nodes, const uint32_t* __restrict__ depth_of,
                                                   const int4* __restrict__ clist, uint32_t total, uint32_t d,
                                                   const uint32_t* __restrict__ wsize, int32_t* index) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= total || depth_of[b] != d || nodes[b].count != 0) return;
    const int32_t at = index[b];
    if (at < 0) return;
    const int4 cl = clist[b];
    uint32_t next = static_cast<uint32_t>(at) + 1;
    for (int i = 0; i < 4; ++i) {
        const int c = child_at(cl, i);
        if (c < 0 || nodes[c].count != 0) continue;
        index[c] = static_cast<int32_t>(next);
        next += wsize[c];
    }
}

__global__ __launch_bounds__(kBlock) void k_index_init(uint32_t total, int32_t* index) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b < total) index[b] = b == 0 ? 0 : -1;
}

// The 4-wide nodes (bvh.hpp:222-241): empty slots are +inf boxes with
// first = count = 0.
__global__ __launch_bounds__(kBlock) void k_emit(const BNode* __restrict__ nodes, const int4* __restrict__ clist,
                                                  const int32_t* __restrict__ index, uint32_t total, WNode* out) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= total) return;
    const int32_t at = index[b];
    if (at < 0) return;
    const int4 cl = clist[b];
    WNode w;
    const float inf = __builtin_inff();
    for (int i = 0; i < 4; ++i) {
        const int c = child_at(cl, i);
        if (c < 0) {
            for (int a = 0; a < 3; ++a) {
                w.lo[a][i] = inf;
                w.hi[a][i] = inf;
            }
            w.first[i] = 0;
            w.count[i] = 0;
            continue;
        }
        const BNode n = nodes[c];
        for (int a = 0; a < 3; ++a) {
            w.lo[a][i] = n.lo[a];
            w.hi[a][i] = n.hi[a];
        }
        if (n.count > 0) {
            w.first[i] = n.first;
            w.count[i] = n.count;
        } else {
            w.first[i] = static_cast<uint32_t>(index[c]);
            w.count[i] = 0;
        }
    }
    out[at] = w;
}

// The build's scratch, carved from the caller's block in 256-byte aligned
// pieces.
struct DevMem {
    char* base = nullptr;
    size_t size = 0, used = 0;
    static size_t round(size_t b) { return (b + 255) & ~size_t(255); }
    template <class T>
    hipError_t get(T** p, size_t count) {
        const size_t b = round(count ? count * sizeof(T) : 16);
        if (used + b > size) return hipErrorOutOfMemory;
        *p = reinterpret_cast<T*>(base + used);
        used += b;
        return hipSuccess;
    }
};

#define BVH_TRY(x)                              \
    do {                                        \
        const hipError_t e_ = (x);              \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

inline uint32_t blocks(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

size_t scratch_bytes(uint32_t n) {
    // Binary node slots: the level passes' nodes (at most 2n - 1) and each
    // small segment's range (2 per primitive).
    const size_t cap = 4 * size_t(n);
    const size_t max_segs = n / kSmall + 2;
    const size_t max_chunks = n / kChunk + max_segs + 2;
    return DevMem::round(sizeof(float4) * n) + DevMem::round(4 * size_t(n)) + DevMem::round(sizeof(BNode) * cap) +
           3 * DevMem::round(4 * cap) + DevMem::round(sizeof(int4) * cap) + DevMem::round(16) +
           2 * DevMem::round(sizeof(Seg) * max_segs) + DevMem::round(sizeof(SegAcc) * max_segs) +
           2 * DevMem::round(sizeof(uint2) * max_chunks) + DevMem::round(8 * max_chunks) +
           DevMem::round(sizeof(SmallSeg) * size_t(n)) + DevMem::round(4 * LC_WORDS);
}

hipError_t build(const float* d_boxes, uint32_t n, uint32_t* d_order, void* d_nodes, uint32_t* n_nodes,
                 uint32_t* box_flags, void* d_scratch, size_t scratch_size, hipStream_t st) {
    if (n == 0 || n > kMaxDeviceBoxes || !d_scratch || scratch_size < scratch_bytes(n) || !box_flags)
        return hipErrorInvalidValue;
    const uint32_t cap = 4 * n;   // binary node slots, as scratch_bytes counts them
    const uint32_t max_segs = n / kSmall + 2;
    const uint32_t max_chunks = n / kChunk + max_segs + 2;
    DevMem mem;
    mem.base = static_cast<char*>(d_scratch);
    mem.size = scratch_size;
    float4* centers;
    uint32_t *pos, *depth_of, *counters, *wsize, *cnt, *lc;
    BNode* nodes;
    int4* clist;
    int32_t* index;
    BVH_TRY(mem.get(&centers, n));
    BVH_TRY(mem.get(&pos, n));
    BVH_TRY(mem.get(&nodes, cap));
    BVH_TRY(mem.get(&depth_of, cap));
    BVH_TRY(mem.get(&clist, cap));
    BVH_TRY(mem.get(&wsize, cap));
    BVH_TRY(mem.get(&index, cap));
    BVH_TRY(mem.get(&counters, 4));
    Seg* segs[2];
    SegAcc* acc;
    uint2* chunks[2];
    BVH_TRY(mem.get(&segs[0], max_segs));
    BVH_TRY(mem.get(&segs[1], max_segs));
    BVH_TRY(mem.get(&acc, max_segs));
    BVH_TRY(mem.get(&chunks[0], max_chunks));
    BVH_TRY(mem.get(&chunks[1], max_chunks));
    BVH_TRY(mem.get(&cnt, 2 * max_chunks));
    SmallSeg* d_small;   // small segments are disjoint and non-empty: at most n
    BVH_TRY(mem.get(&d_small, n));
    BVH_TRY(mem.get(&lc, LC_WORDS));

    BVH_TRY(hipMemsetAsync(counters, 0, 4 * sizeof(uint32_t), st));
    k_setup<<<blocks(n), kBlock, 0, st>>>(d_boxes, n, centers, d_order, counters, depth_of, cap);
    BVH_TRY(hipGetLastError());

    // The root segment: a level pass's if longer than kSmall, else the one
    // small segment. The level passes then run without the host: each level
    // ends with k_plan writing the next level's lists and counts, and the
    // kernels of a level read their counts from lc (grids sized for the
    // most a level can hold: at most 2^L segments at level L, and a chunk
    // per kChunk primitives plus one per segment).
    uint32_t h_lc[LC_WORDS] = {};
    Seg root{};
    root.count = n;
    std::vector<uint2> root_chunks;
    if (n > kSmall) {
        root.chunk_count = (n + kChunk - 1) / kChunk;
        for (uint32_t k = 0; k < root.chunk_count; ++k) root_chunks.push_back(make_uint2(0, k * kChunk));
        h_lc[LC_NSEG] = 1;
        h_lc[LC_NCH] = root.chunk_count;
        BVH_TRY(hipMemcpyAsync(segs[0], &root, sizeof(Seg), hipMemcpyHostToDevice, st));
        BVH_TRY(hipMemcpyAsync(chunks[0], root_chunks.data(), root_chunks.size() * sizeof(uint2),
                               hipMemcpyHostToDevice, st));
    } else {
        SmallSeg one{};
        one.count = n;
        h_lc[LC_NSMALL] = 1;
        BVH_TRY(hipMemcpyAsync(d_small, &one, sizeof(SmallSeg), hipMemcpyHostToDevice, st));
    }
    BVH_TRY(hipMemcpyAsync(lc, h_lc, sizeof(h_lc), hipMemcpyHostToDevice, st));
    // Boxes with a NaN or a -0.0 coordinate stop here (the caller decides).
    uint32_t flags = 0;
    BVH_TRY(hipMemcpyAsync(&flags, counters + 2, 4, hipMemcpyDeviceToHost, st));
    BVH_TRY(hipStreamSynchronize(st));
    *box_flags = flags;
    *n_nodes = 0;
    if (flags) return hipSuccess;

    if (n > kSmall) {
        // The host looks at the counts only now and then: first after the
        // levels a balanced split of n would take, then every two levels
        // (a level with nothing open costs a few empty launches).
        uint32_t check_at = 1;
        while ((uint64_t(kSmall) << check_at) < n) ++check_at;
        int p = 0;
        for (uint32_t level = 0;; ++level) {
            if (level > kMaxDepth) return hipErrorInvalidValue;   // cannot happen: depth is capped at kMaxDepth
            const uint32_t seg_grid = level < 31 ? std::min<uint32_t>(max_segs, 1u << level) : max_segs;
            const uint32_t ch_grid = static_cast<uint32_t>(
                std::min<uint64_t>(max_chunks, uint64_t(n) / kChunk + 1 + seg_grid));
            k_seg_init<<<seg_grid, kBlock, 0, st>>>(acc, lc, p);
            k_bounds<<<ch_grid, kBlock, 0, st>>>(d_boxes, centers, d_order, chunks[p], segs[p], acc, lc, p);
            k_bin<<<ch_grid, kBlock, 0, st>>>(d_boxes, centers, d_order, chunks[p], segs[p], acc, lc, p);
            k_decide<<<(seg_grid + 63) / 64, 64, 0, st>>>(segs[p], acc, lc, p, nodes, depth_of, counters);
            k_part_count<<<ch_grid, kBlock, 0, st>>>(centers, d_order, chunks[p], segs[p], cnt, lc, p);
            k_part_pos<<<ch_grid, kBlock, 0, st>>>(centers, d_order, chunks[p], segs[p], cnt, pos, lc, p);
            k_part_swap<<<ch_grid, kBlock, 0, st>>>(chunks[p], segs[p], cnt, pos, d_order, lc, p);
            k_plan<<<1, kPlanBlock, 0, st>>>(segs[p], lc, p, segs[p ^ 1], chunks[p ^ 1], d_small, max_segs, max_chunks,
                                             n);
            BVH_TRY(hipGetLastError());
            p ^= 1;
            if (level + 1 >= check_at) {
                BVH_TRY(hipMemcpyAsync(h_lc, lc, sizeof(h_lc), hipMemcpyDeviceToHost, st));
                BVH_TRY(hipStreamSynchronize(st));
                if (h_lc[LC_ERR]) return hipErrorInvalidValue;
                if (h_lc[LC_NSEG + p] == 0) break;
                check_at = level + 3;
            }
        }
    }

    k_small_bases<<<1, kPlanBlock, 0, st>>>(d_small, lc, counters);
    BVH_TRY(hipMemcpyAsync(h_lc, lc, sizeof(h_lc), hipMemcpyDeviceToHost, st));
    BVH_TRY(hipStreamSynchronize(st));
    const uint32_t n_small = h_lc[LC_NSMALL], total = h_lc[LC_TOTAL];
    if (h_lc[LC_ERR] || n_small == 0 || n_small > n || total > cap) return hipErrorInvalidValue;
    k_small<<<n_small, 64, 0, st>>>(d_boxes, centers, d_order, d_small, nodes, depth_of, counters);
    BVH_TRY(hipGetLastError());
    uint32_t hc[2];
    BVH_TRY(hipMemcpyAsync(hc, counters, 8, hipMemcpyDeviceToHost, st));
    BVH_TRY(hipStreamSynchronize(st));
    const uint32_t deepest = hc[1];
    if (deepest >= kMaxDepth) return hipErrorInvalidValue;

    k_clist<<<blocks(total), kBlock, 0, st>>>(nodes, depth_of, total, clist);
    k_index_init<<<blocks(total), kBlock, 0, st>>>(total, index);
    for (int d = static_cast<int>(deepest); d >= 0; --d)
        k_wsize<<<blocks(total), kBlock, 0, st>>>(nodes, depth_of, clist, total, static_cast<uint32_t>(d), wsize);
    for (uint32_t d = 0; d <= deepest; ++d)
        k_index<<<blocks(total), kBlock, 0, st>>>(nodes, depth_of, clist, total, d, wsize, index);
    k_emit<<<blocks(total), kBlock, 0, st>>>(nodes, clist, index, total, static_cast<WNode*>(d_nodes));
    BVH_TRY(hipGetLastError());
    uint32_t root_count = 0, root_size = 0;
    BVH_TRY(hipMemcpyAsync(&root_count, &nodes[0].count, 4, hipMemcpyDeviceToHost, st));
    BVH_TRY(hipMemcpyAsync(&root_size, wsize, 4, hipMemcpyDeviceToHost, st));
    BVH_TRY(hipStreamSynchronize(st));
    *n_nodes = root_count > 0 ? 1 : root_size;
    return hipSuccess;
}

}  // namespace bvhg
}  // namespace ptd
