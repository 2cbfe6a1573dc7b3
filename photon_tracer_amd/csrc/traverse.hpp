// Closest-hit and any-hit traversal of the packed scene for one ray per lane.
//
// The walk is the reference's (bvh.hpp:57-106, geometry.hpp:111-190,
// 212-229): large spheres in double first, then the flat tree, then the
// instance tree, whose leaves move the ray into each instance's space and walk
// that mesh's tree. Entries are popped nearest first, an entry whose entry
// distance is not below the current t_max is skipped, hit children are pushed
// farthest first with ties kept in child order, and leaves test primitives in
// order, so the primitive found on an exact tie is the reference's.
//
// The stack is a short ring in LDS (S entries per lane, lane-interleaved so a
// wave's accesses spread over all banks) backed by a per-lane spill area in
// HBM for the deep end: pushing onto a full ring moves its oldest entry out,
// popping past the ring's bottom reads it back. Entries are 8 bytes: the node
// or leaf start in the low 28 bits and a leaf's count above, and the entry t.
#pragma once

#include "device.hpp"

namespace ptd {

// Explicit address spaces, so the ring is read with ds_read and the spill
// area with global loads (a generic pointer selected between the two becomes
// a flat load, which waits on every outstanding memory operation).
// Entries are kept as 64-bit words: the node code low, the entry t high.
using lds_u64 = __attribute__((address_space(3))) unsigned long long;
using gbl_u64 = __attribute__((address_space(1))) unsigned long long;

__device__ __forceinline__ unsigned long long pack_entry(uint32_t code, float t) {
    return (static_cast<unsigned long long>(__float_as_uint(t)) << 32) | code;
}
__device__ __forceinline__ uint2 unpack_entry(unsigned long long e) {
    return make_uint2(static_cast<uint32_t>(e), static_cast<uint32_t>(e >> 32));
}

template <int S, int BLOCK>
struct LaneStack {
    lds_u64* lds;      // this lane's slot 0; slot k is lds[k * BLOCK]
    gbl_u64* spill;    // this lane's entry 0 in HBM; entry k is spill[k * stride]
    uint32_t stride;
    int top, base;     // entries [0, base) are in HBM, [base, top) in the ring

    __device__ __forceinline__ void reset() { top = base = 0; }
    // A node's hit children are pushed together: reserve() makes room for m
    // entries (the oldest move out to HBM), put() writes entry top + k,
    // advance() counts the m entries in.
    __device__ __forceinline__ void reserve(int m) {
        while (top - base + m > S) {
            __builtin_nontemporal_store(lds[(base & (S - 1)) * BLOCK], spill + static_cast<size_t>(base) * stride);
            ++base;
        }
    }
    __device__ __forceinline__ void put(int k, uint32_t code, float t) {
        lds[((top + k) & (S - 1)) * BLOCK] = pack_entry(code, t);
    }
    __device__ __forceinline__ void advance(int m) { top += m; }
    __device__ __forceinline__ uint2 pop() {
        --top;
        unsigned long long v = lds[(top & (S - 1)) * BLOCK];   // the ring slot, read unconditionally
        if (top < base) {                                      // below the ring: the entry was spilled
            base = top;
            // A load of its own kind, so the compiler cannot merge it with
            // the ring read into one flat load.
            v = __builtin_nontemporal_load(spill + static_cast<size_t>(top) * stride);
        }
        return unpack_entry(v);
    }
};

constexpr uint32_t FIRST_MASK = 0x0fffffffu;

// Waves per SIMD the traversal kernels are compiled for. The 8-entry ring
// without the instance walk fits 64 VGPRs (8 waves, 16 KB of LDS per block
// of 4 waves); the others keep the compiler's own budget. The path loop's
// traversal stage (REFILL: lane refill, camera rays or queued rays and light
// samples) gets 7 waves: at 64 VGPRs it spills to scratch, at 72 it does not
// and runs faster.
template <int S, bool INST, bool REFILL = false>
struct trace_occupancy {
    static constexpr int waves = (S == 8 && !INST) ? (REFILL ? 7 : 8) : 1;
};

// Visit counts of one query, kept only by the counting kernel variants: node
// slab tests, leaf references, triangle tests, sphere tests, instance
// entries, large-sphere tests. They price a query's algorithmic bytes.
struct Visits {
    uint32_t nodes = 0, leafrefs = 0, tris = 0, spheres = 0, instances = 0, large = 0, queries = 0;
};
constexpr int COUNT_SHIFT = 28;

// The slab test of a node's four children, then the hit children pushed
// farthest first (bvh.hpp:80-104). IEEE min/max stand in for SSE's
// minps/maxps; they differ only where 0 * inf makes a NaN (a ray with a zero
// direction component starting exactly on a box face). The nearest hit child,
// which the reference pushes last and pops at once, is returned in `next`
// instead of going through the stack; returns whether there is one.
//
// Nodes below sc.n_cached come from the launch's LDS copy (the top levels
// of the trees, which nearly every ray visits), the others from HBM through
// the caches: the traversal's limit is the vector memory pipeline, seven
// 16-byte loads per lane and node step. The LDS rows are read by inline
// assembly that waits for its own loads, so the compiler cannot merge the two
// kinds of read into one flat load of a selected address.
typedef float lds_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 as_f4(lds_f4 v) { return make_float4(v.x, v.y, v.z, v.w); }

//
// ordered = false is for any-hit queries: whether anything is hit in
// [tmin, tmax] does not depend on the order the children are taken in (no
// hit moves tmax before the query ends), so the hit children are pushed in
// child order and the last is taken next, without ranking their entry
// distances. The caller passes the same value on every lane of a wave.
template <class Stack>
__device__ __forceinline__ bool visit_node(const SceneView& sc, uint32_t idx, f3 o, f3 inv, float tmin, float tmax,
                                           Stack& st, uint2& next, bool ordered = true) {
    float4 lx, ly, lz, hx, hy, hz;
    uint4 cd;   // first | count << 28 (the device copy of the nodes)
    if (idx < sc.n_cached) {
        const uint32_t a = sc.cache + idx * NODE_CACHE_STRIDE;
        lds_f4 r0, r1, r2, r3, r4, r5, r6;
        asm volatile(
            "ds_read_b128 %0, %7\n\t"
            "ds_read_b128 %1, %7 offset:16\n\t"
            "ds_read_b128 %2, %7 offset:32\n\t"
            "ds_read_b128 %3, %7 offset:48\n\t"
            "ds_read_b128 %4, %7 offset:64\n\t"
            "ds_read_b128 %5, %7 offset:80\n\t"
            "ds_read_b128 %6, %7 offset:96\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(r0), "=&v"(r1), "=&v"(r2), "=&v"(r3), "=&v"(r4), "=&v"(r5), "=&v"(r6)
            : "v"(a));
        lx = as_f4(r0);
        ly = as_f4(r1);
        lz = as_f4(r2);
        hx = as_f4(r3);
        hy = as_f4(r4);
        hz = as_f4(r5);
        cd = make_uint4(__float_as_uint(r6.x), __float_as_uint(r6.y), __float_as_uint(r6.z), __float_as_uint(r6.w));
    } else {
        const Node* nd = sc.nodes + idx;
        lx = nd->lo[0];
        ly = nd->lo[1];
        lz = nd->lo[2];
        hx = nd->hi[0];
        hy = nd->hi[1];
        hz = nd->hi[2];
        cd = nd->first;
    }
    const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w}, lza[4] = {lz.x, lz.y, lz.z, lz.w};
    const float hxa[4] = {hx.x, hx.y, hx.z, hx.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
    const uint32_t code[4] = {cd.x, cd.y, cd.z, cd.w};
    float tn[4];
    bool h[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float x0 = (lxa[c] - o.x) * inv.x, x1 = (hxa[c] - o.x) * inv.x;
        const float y0 = (lya[c] - o.y) * inv.y, y1 = (hya[c] - o.y) * inv.y;
        const float z0 = (lza[c] - o.z) * inv.z, z1 = (hza[c] - o.z) * inv.z;
        const float near = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
        const float far = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
        tn[c] = near;
        h[c] = near <= far;
    }
    // Position of each hit child in the farthest-first order: those entered
    // later come first, and on equal entry the lower child number first.
    int rank[4] = {0, 0, 0, 0};
    if (ordered) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int b2 = a + 1; b2 < 4; ++b2) {
                const bool later = tn[b2] > tn[a];     // b2 is pushed before a
                rank[a] += (h[b2] && later) ? 1 : 0;
                rank[b2] += (h[a] && !later) ? 1 : 0;
            }
        }
    } else {
        rank[1] = h[0] ? 1 : 0;
        rank[2] = rank[1] + (h[1] ? 1 : 0);
        rank[3] = rank[2] + (h[2] ? 1 : 0);
    }
    const int n = (h[0] ? 1 : 0) + (h[1] ? 1 : 0) + (h[2] ? 1 : 0) + (h[3] ? 1 : 0);
    if (n == 0) return false;
    st.reserve(n - 1);
    uint32_t ncd = 0;
    float nt = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (h[c] && rank[c] < n - 1) st.put(rank[c], code[c], tn[c]);
        const bool sel = h[c] && rank[c] == n - 1;
        ncd = sel ? code[c] : ncd;
        nt = sel ? tn[c] : nt;
    }
    st.advance(n - 1);
    next = make_uint2(ncd, __float_as_uint(nt));
    return true;
}

__device__ __forceinline__ f3 safe_rcp3(f3 d) { return {rcp_rn(d.x), rcp_rn(d.y), rcp_rn(d.z)}; }

// Returns whether something was hit. With ANY, stops at the first hit
// (geometry::occluded); otherwise fills `hit` with the nearest. Any-hit
// walks of scenes without instances push children unranked; with instances
// the nearest-first order finds the occluder sooner and stays.
// INST = false is the variant for scenes without instances: the instance
// walk, and the registers it keeps live, are compiled out.
template <bool ANY, bool COUNT = false, bool INST = true, class Stack>
__device__ __forceinline__ bool trace_ray(const SceneView& sc, f3 o, f3 d, float tmin, float tmax, Hit& hit,
                                          Stack& st, Visits* vis = nullptr) {
    uint32_t kind = 0, prim = 0, inst = NONE;
    float hu = 0, hv = 0;
    if (COUNT) ++vis->queries;

    // Large spheres, in double (geometry.hpp:117-118, 157-160).
    for (uint32_t k = 0; k < sc.n_large; ++k) {
        if (COUNT) ++vis->large;
        const uint32_t s = sc.large[k];
        if (sphere_hit_d(sc.spheres[2 * s], o, d, tmin, tmax)) {
            if (ANY) return true;
            kind = 1;
            prim = s;
        }
    }
    const f3 inv = safe_rcp3(d);

    // The flat tree: spheres then triangles (geometry.hpp:119-127).
    if (sc.flat_root != NONE) {
        st.reset();
        uint2 e = make_uint2(sc.flat_root, __float_as_uint(tmin));
        bool have = true;
        while (true) {
            if (!have) {
                if (st.top == 0) break;
                e = st.pop();
            }
            have = false;
            if (__uint_as_float(e.y) >= tmax) continue;
            const uint32_t first = e.x & FIRST_MASK, count = e.x >> COUNT_SHIFT;
            if (count == 0) {
                if (COUNT) ++vis->nodes;
                have = visit_node(sc, first, o, inv, tmin, tmax, st, e, !ANY || INST);
                continue;
            }
            for (uint32_t i = first; i < first + count; ++i) {
                const float4 r0 = sc.leaf[3 * i], r1 = sc.leaf[3 * i + 1];
                if (COUNT) {
                    ++vis->leafrefs;
                    if (__float_as_uint(r1.w) != 0u) ++vis->spheres;
                    else ++vis->tris;
                }
                if (__float_as_uint(r1.w) != 0u) {
                    const float4 cr = make_float4(r0.x, r0.y, r0.z, r1.x);
                    if (sphere_hit_f(cr, o, d, tmin, tmax)) {
                        if (ANY) return true;
                        kind = 2;
                        prim = __float_as_uint(r0.w);
                        inst = NONE;
                    }
                } else {
                    const float4 r2 = sc.leaf[3 * i + 2];
                    if (tri_hit(xyz(r0), xyz(r1), xyz(r2), o, d, tmin, tmax, hu, hv)) {
                        if (ANY) return true;
                        kind = 3;
                        prim = __float_as_uint(r0.w);
                        inst = NONE;
                    }
                }
            }
        }
    }

    // The instance tree; a leaf walks each instance's mesh tree in object
    // space, where t is the same (geometry.hpp:212-229).
    if (INST && sc.inst_root != NONE) {
        st.reset();
        uint2 e = make_uint2(sc.inst_root, __float_as_uint(tmin));
        bool have = true;
        bool in_blas = false;
        int blas_base = 0;
        uint32_t li = 0, lend = 0, cur = 0;
        f3 lo = o, ld = d, linv = inv;
        auto enter = [&](uint32_t k) {
            if (COUNT) ++vis->instances;
            cur = k;
            const float4 r0 = sc.inst[7 * k + 3], r1 = sc.inst[7 * k + 4], r2 = sc.inst[7 * k + 5];
            const uint32_t root = __float_as_uint(sc.inst[7 * k + 6].x);
            lo = mk(r0.x * o.x + r0.y * o.y + r0.z * o.z + r0.w, r1.x * o.x + r1.y * o.y + r1.z * o.z + r1.w,
                    r2.x * o.x + r2.y * o.y + r2.z * o.z + r2.w);
            ld = mk(r0.x * d.x + r0.y * d.y + r0.z * d.z, r1.x * d.x + r1.y * d.y + r1.z * d.z,
                    r2.x * d.x + r2.y * d.y + r2.z * d.z);
            linv = safe_rcp3(ld);
            blas_base = st.top;
            in_blas = true;
            e = make_uint2(root, __float_as_uint(tmin));
            have = root != NONE;
        };
        while (true) {
            if (!have) {
                if (in_blas && st.top == blas_base) {
                    if (++li < lend) {
                        enter(sc.order[li]);
                        continue;
                    }
                    in_blas = false;
                }
                if (st.top == 0) break;
                e = st.pop();
            }
            have = false;
            if (__uint_as_float(e.y) >= tmax) continue;
            const uint32_t first = e.x & FIRST_MASK, count = e.x >> COUNT_SHIFT;
            if (count == 0) {
                if (COUNT) ++vis->nodes;
                if (in_blas) have = visit_node(sc, first, lo, linv, tmin, tmax, st, e, true);
                else have = visit_node(sc, first, o, inv, tmin, tmax, st, e, true);
                continue;
            }
            if (COUNT) vis->leafrefs += count;
            if (!in_blas) {
                li = first;
                lend = first + count;
                enter(sc.order[li]);
                continue;
            }
            for (uint32_t i = first; i < first + count; ++i) {
                if (COUNT) ++vis->tris;
                const float4 r0 = sc.leaf[3 * i], r1 = sc.leaf[3 * i + 1], r2 = sc.leaf[3 * i + 2];
                if (tri_hit(xyz(r0), xyz(r1), xyz(r2), lo, ld, tmin, tmax, hu, hv)) {
                    if (ANY) return true;
                    kind = 3;
                    prim = __float_as_uint(r0.w);
                    inst = cur;
                }
            }
        }
    }
    if (ANY) return false;
    hit.t = tmax;
    hit.u = hu;
    hit.v = hv;
    hit.kind = kind;
    hit.prim = prim;
    hit.inst = inst;
    return kind != 0;
}

// One ray's walk, advanced one stack entry per step, so a lane whose ray has
// ended can start another while the rest of its wave is still walking. A ray
// goes through exactly the steps trace_ray takes for it, with the same
// results: the large spheres, the flat tree, then (INST) the instance tree,
// whose leaves enter each instance's mesh tree in object space.
template <bool INST>
struct WalkInst {};
template <>
struct WalkInst<true> {
    f3 lo, ld, linv;        // the ray in the entered instance's object space
    uint32_t li, lend;      // the instance-tree leaf's slots still to enter: [li, lend)
    uint32_t cur, inst;     // the instance entered; the instance of the nearest hit (NONE: none)
    int blas_base;          // stack depth when the instance was entered
    bool tlas, in_blas;     // walking the instance tree; inside an instance's mesh tree
};

template <bool INST = false>
struct Walk : WalkInst<INST> {
    f3 o, d, inv;
    float tmin, tmax, hu, hv;
    uint32_t kind, prim;
    uint2 e;       // the entry to take next when `have`
    bool have;
};

// Moves a walk on to the next stack entry it will process: the pending
// entry, or the ones popped after it, skipping every entry whose entry
// distance is not below the nearest hit so far (bvh.hpp:74); INST: also the
// instance tree's bookkeeping between entries (the next instance of a leaf,
// the switch from the flat tree to the instance tree). These are the steps
// trace_ray takes between two node or leaf visits, so a walk step always
// visits a node or a leaf and the step that empties the stack ends the walk.
// Returns true when the walk has ended.
template <bool COUNT, bool INST, class Stack>
__device__ __forceinline__ bool walk_advance(const SceneView& sc, Walk<INST>& w, Stack& st, Visits* vis);

// Starts a walk: the large spheres in double, then the flat tree's root.
// Returns true if the query has already ended (an any-hit query that hit a
// large sphere, or nothing to walk).
template <bool COUNT, bool INST = false, class Stack>
__device__ __forceinline__ bool walk_begin(const SceneView& sc, Walk<INST>& w, f3 o, f3 d, float tmin, float tmax,
                                           bool any, Stack& st, Visits* vis) {
    if (COUNT) ++vis->queries;
    w.o = o;
    w.d = d;
    w.tmin = tmin;
    w.tmax = tmax;
    w.hu = w.hv = 0;
    w.kind = w.prim = 0;
    if constexpr (INST) {
        w.inst = NONE;
        w.tlas = w.in_blas = false;
    }
    for (uint32_t k = 0; k < sc.n_large; ++k) {
        if (COUNT) ++vis->large;
        const uint32_t s = sc.large[k];
        if (sphere_hit_d(sc.spheres[2 * s], o, d, tmin, w.tmax)) {
            w.kind = 1;
            w.prim = s;
            if (any) return true;
        }
    }
    w.inv = safe_rcp3(d);
    st.reset();
    w.e = make_uint2(sc.flat_root, __float_as_uint(tmin));
    w.have = sc.flat_root != NONE;
    return walk_advance<COUNT, INST>(sc, w, st, vis);
}

// Enters instance k of an instance-tree leaf: the ray to its object space
// (the same t there), its mesh tree's root next (trace_ray's `enter`).
template <bool COUNT, class Stack>
__device__ __forceinline__ void walk_enter(const SceneView& sc, Walk<true>& w, uint32_t k, Stack& st, Visits* vis) {
    if (COUNT) ++vis->instances;
    w.cur = k;
    const float4 r0 = sc.inst[7 * k + 3], r1 = sc.inst[7 * k + 4], r2 = sc.inst[7 * k + 5];
    const uint32_t root = __float_as_uint(sc.inst[7 * k + 6].x);
    const f3 o = w.o, d = w.d;
    w.lo = mk(r0.x * o.x + r0.y * o.y + r0.z * o.z + r0.w, r1.x * o.x + r1.y * o.y + r1.z * o.z + r1.w,
              r2.x * o.x + r2.y * o.y + r2.z * o.z + r2.w);
    w.ld = mk(r0.x * d.x + r0.y * d.y + r0.z * d.z, r1.x * d.x + r1.y * d.y + r1.z * d.z,
              r2.x * d.x + r2.y * d.y + r2.z * d.z);
    w.linv = safe_rcp3(w.ld);
    w.blas_base = st.top;
    w.in_blas = true;
    w.e = make_uint2(root, __float_as_uint(w.tmin));
    w.have = root != NONE;
}

template <bool COUNT, bool INST, class Stack>
__device__ __forceinline__ bool walk_advance(const SceneView& sc, Walk<INST>& w, Stack& st, Visits* vis) {
    while (true) {
        if (w.have) {
            if (!(__uint_as_float(w.e.y) >= w.tmax)) return false;
            w.have = false;   // behind the nearest hit: skipped
        }
        if constexpr (INST) {
            if (w.in_blas && st.top == w.blas_base) {   // the entered instance's mesh tree is done
                if (++w.li < w.lend) {
                    walk_enter<COUNT>(sc, w, sc.order[w.li], st, vis);
                    continue;
                }
                w.in_blas = false;
            }
            if (st.top == 0) {
                if (w.tlas || sc.inst_root == NONE) return true;
                w.tlas = true;   // the flat tree is done: the instance tree next
                st.reset();
                w.e = make_uint2(sc.inst_root, __float_as_uint(w.tmin));
                w.have = true;
                continue;
            }
        } else {
            if (st.top == 0) return true;
        }
        w.e = st.pop();
        w.have = true;
    }
}

// One node or leaf of the walk (the entry walk_advance left pending): a
// node's slab test, a leaf's primitives, or (INST) an instance-tree leaf's
// first instance entered; then walk_advance to the next entry. Returns true
// when the walk has ended (every tree walked, or an any-hit query found a
// hit); kind != 0 then says something was hit.
template <bool COUNT, bool INST = false, class Stack>
__device__ __forceinline__ bool walk_step(const SceneView& sc, Walk<INST>& w, bool any, Stack& st, Visits* vis,
                                          bool ordered = true) {
    w.have = false;
    const uint32_t first = w.e.x & FIRST_MASK, count = w.e.x >> COUNT_SHIFT;
    if (count == 0) {
        if (COUNT) ++vis->nodes;
        bool in_blas = false;
        if constexpr (INST) in_blas = w.in_blas;
        if (in_blas) {
            if constexpr (INST) w.have = visit_node(sc, first, w.lo, w.linv, w.tmin, w.tmax, st, w.e, ordered);
        } else {
            w.have = visit_node(sc, first, w.o, w.inv, w.tmin, w.tmax, st, w.e, ordered);
        }
        return walk_advance<COUNT, INST>(sc, w, st, vis);
    }
    if constexpr (INST) {
        if (w.tlas) {
            if (COUNT) vis->leafrefs += count;
            if (!w.in_blas) {   // an instance-tree leaf: enter its instances in turn
                w.li = first;
                w.lend = first + count;
                walk_enter<COUNT>(sc, w, sc.order[first], st, vis);
                return walk_advance<COUNT, INST>(sc, w, st, vis);
            }
            for (uint32_t i = first; i < first + count; ++i) {   // a mesh leaf, in object space
                if (COUNT) ++vis->tris;
                const float4 r0 = sc.leaf[3 * i], r1 = sc.leaf[3 * i + 1], r2 = sc.leaf[3 * i + 2];
                if (tri_hit(xyz(r0), xyz(r1), xyz(r2), w.lo, w.ld, w.tmin, w.tmax, w.hu, w.hv)) {
                    w.kind = 3;
                    w.prim = __float_as_uint(r0.w);
                    w.inst = w.cur;
                    if (any) return true;
                }
            }
            return walk_advance<COUNT, INST>(sc, w, st, vis);
        }
    }
    for (uint32_t i = first; i < first + count; ++i) {
        const float4 r0 = sc.leaf[3 * i], r1 = sc.leaf[3 * i + 1];
        if (COUNT) {
            ++vis->leafrefs;
            if (__float_as_uint(r1.w) != 0u) ++vis->spheres;
            else ++vis->tris;
        }
        bool hit;
        if (__float_as_uint(r1.w) != 0u) {
            hit = sphere_hit_f(make_float4(r0.x, r0.y, r0.z, r1.x), w.o, w.d, w.tmin, w.tmax);
            if (hit) w.kind = 2;
        } else {
            const float4 r2 = sc.leaf[3 * i + 2];
            hit = tri_hit(xyz(r0), xyz(r1), xyz(r2), w.o, w.d, w.tmin, w.tmax, w.hu, w.hv);
            if (hit) w.kind = 3;
        }
        if (hit) {
            w.prim = __float_as_uint(r0.w);
            if constexpr (INST) w.inst = NONE;
            if (any) return true;
        }
    }
    return walk_advance<COUNT, INST>(sc, w, st, vis);
}

// Copies nodes [0, min(n_cache, n_top)) of the device node array into a block's LDS node
// cache (7 rows of 16 bytes each, NODE_CACHE_STRIDE apart) and returns the
// scene view that reads them from there. Called by every thread of the
// block; ends with a barrier.
__device__ __forceinline__ SceneView cache_top_nodes(const SceneView& sc, float4* cache, uint32_t n_cache) {
    SceneView v = sc;
    const uint32_t n = min(n_cache, sc.n_top);   // the device copy numbers its top nodes first
    const float4* src = reinterpret_cast<const float4*>(sc.nodes);
    for (uint32_t i = threadIdx.x; i < 7 * n; i += blockDim.x) {
        const uint32_t k = i / 7, row = i % 7;
        cache[k * (NODE_CACHE_STRIDE / 16) + row] = src[8 * k + row];
    }
    __syncthreads();
    v.n_cached = n;
    v.cache = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) float4*)(cache)));
    return v;
}

// One lane's stack over the block's LDS ring array and the grid's spill area.
template <int S, int BLOCK>
__device__ __forceinline__ LaneStack<S, BLOCK> lane_stack(uint2* lds, uint2* spill) {
    LaneStack<S, BLOCK> st;
    st.lds = (lds_u64*)(lds) + threadIdx.x;
    const uint32_t stride = gridDim.x * BLOCK;
    st.spill = (gbl_u64*)(spill) + (blockIdx.x * BLOCK + threadIdx.x);
    st.stride = stride;
    st.top = st.base = 0;
    return st;
}

// Adds a lane's visit counts into six 64-bit totals, one atomic per wave and counter.
__device__ __forceinline__ void add_visits(unsigned long long* totals, const Visits& v) {
    const uint32_t x[7] = {v.nodes, v.leafrefs, v.tris, v.spheres, v.instances, v.large, v.queries};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        unsigned long long s = x[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if ((threadIdx.x & 63) == 0 && s) atomicAdd(totals + k, s);
    }
}

// A wave takes `n` items at a time from a shared counter.
__device__ __forceinline__ uint32_t wave_fetch(uint32_t* counter, uint32_t n) {
    uint32_t base = 0;
    if ((threadIdx.x & 63) == 0) base = atomicAdd(counter, n);
    return __shfl(base, 0, 64);
}

// Groups of 64 items a wave takes per fetch from a queue of n items. One
// counter is shared by every wave of the grid across all XCDs, where each
// atomic on it is slow; large queues are handed out up to 512 items at a
// time, sized so every wave still makes about 16 fetches (a finer split of
// the last pieces, which decide when the kernel ends).
__device__ __forceinline__ uint32_t fetch_groups(uint32_t n) {
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    const uint32_t g = n / (waves * 64u * 16u);
    return g < 1 ? 1 : g > 8 ? 8 : g;
}

}  // namespace ptd
