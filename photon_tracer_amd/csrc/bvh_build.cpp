// Binned-SAH binary build, then collapse to the 4-wide node layout.
//
// The result must equal the reference's tree node for node (bvh.hpp:39-52,
// 126-247), since the packed scene is a drop-in for pack_for_gpu's: the same
// 16 bins, unit SAH costs, max leaf 8, max depth 64, the same split test and
// tie rules, and std::partition for the split (its element order is
// implementation-defined, so it is the same library routine the reference
// compiles against). The recursion is replaced by an explicit work list, and
// the two halves of large nodes are built on separate threads: each subtree
// only touches its own range of `order`, and the node numbering comes from
// the collapse pass, which runs afterwards in the reference's preorder.
#include "host.hpp"

#include <algorithm>
#include <thread>

namespace pt {
namespace {

constexpr int kBins = 16;
constexpr int kMaxLeaf = 8;
constexpr int kMaxDepth = 64;
constexpr float kTraversalCost = 1, kIntersectionCost = 1;

struct bnode {
    box3 box;
    uint32_t first = 0;   // interior: left child (right = left + 1) in `kids`; leaf: start in order
    uint32_t count = 0;   // primitives in a leaf, 0 for interior
    int32_t left = -1;    // index of the left child node (this builder's numbering)
};

inline int bin_of(float c, float lo, float extent) {
    const int b = static_cast<int>(kBins * (c - lo) / extent);
    return b < 0 ? 0 : b >= kBins ? kBins - 1 : b;
}

struct builder {
    const std::vector<box3>& boxes;
    std::vector<v3> centers;
    std::vector<uint32_t> order;

    explicit builder(const std::vector<box3>& b) : boxes(b) {}

    // One binary subtree, built depth first into `out` (a per-thread arena).
    // Mirrors bvh::split (bvh.hpp:126-197).
    struct arena {
        std::vector<bnode> nodes;
    };

    // Decides the split of [first, first+count); returns mid, or UINT32_MAX for a leaf.
    uint32_t split_range(uint32_t first, uint32_t count, int depth, box3& box_out) {
        box3 box, centroids;
        for (uint32_t i = first; i < first + count; ++i) {
            box.grow(boxes[order[i]]);
            box3 c;
            c.lo = c.hi = centers[order[i]];
            centroids.grow(c);
        }
        box_out = box;
        if (count == 1 || depth + 1 >= kMaxDepth) return UINT32_MAX;

        float best_cost = box3::inf();
        int best_axis = -1, best_bin = 0;
        for (int axis = 0; axis < 3; ++axis) {
            const float lo = centroids.lo[axis], extent = centroids.hi[axis] - lo;
            if (extent <= 0) continue;
            box3 bin_box[kBins];
            uint32_t bin_count[kBins] = {};
            for (uint32_t i = first; i < first + count; ++i) {
                const int b = bin_of(centers[order[i]][axis], lo, extent);
                bin_box[b].grow(boxes[order[i]]);
                ++bin_count[b];
            }
            float right_area[kBins];
            uint32_t right_count[kBins];
            box3 acc;
            uint32_t n = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bin_box[b]);
                n += bin_count[b];
                right_area[b] = acc.half_area();
                right_count[b] = n;
            }
            acc = box3();
            n = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                acc.grow(bin_box[b]);
                n += bin_count[b];
                if (n == 0 || right_count[b + 1] == 0) continue;
                const float cost = kTraversalCost + kIntersectionCost *
                                   (acc.half_area() * n + right_area[b + 1] * right_count[b + 1]) / box.half_area();
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_bin = b;
                }
            }
        }
        if (count <= kMaxLeaf && !(best_cost < count * kIntersectionCost)) return UINT32_MAX;

        if (best_axis >= 0) {
            const float lo = centroids.lo[best_axis], extent = centroids.hi[best_axis] - lo;
            const int axis = best_axis, bin = best_bin;
            uint32_t* m = std::partition(order.data() + first, order.data() + first + count, [&](uint32_t p) {
                return bin_of(centers[p][axis], lo, extent) <= bin;
            });
            return static_cast<uint32_t>(m - order.data());
        }
        return first + count / 2;
    }

    // Builds the subtree for [first, first+count) at `depth` into arena `a`,
    // returning its root's index in a.nodes. Iterative, children in order.
    int32_t build_subtree(arena& a, uint32_t first, uint32_t count, int depth, int parallel_levels);
};

int32_t builder::build_subtree(arena& a, uint32_t first, uint32_t count, int depth, int parallel_levels) {
    const int32_t root = static_cast<int32_t>(a.nodes.size());
    a.nodes.push_back({});
    box3 box;
    const uint32_t mid = split_range(first, count, depth, box);
    a.nodes[root].box = box;
    if (mid == UINT32_MAX) {
        a.nodes[root].first = first;
        a.nodes[root].count = count;
        return root;
    }
    const uint32_t lcount = mid - first, rcount = first + count - mid;
    int32_t l, r;
    if (parallel_levels > 0 && count > 200000) {
        // Disjoint ranges of `order`: the right half builds on its own thread
        // into its own arena and is spliced in after.
        arena right_arena;
        int32_t r_local = -1;
        std::thread t([&] { r_local = build_subtree(right_arena, mid, rcount, depth + 1, parallel_levels - 1); });
        l = build_subtree(a, first, lcount, depth + 1, parallel_levels - 1);
        t.join();
        const int32_t base = static_cast<int32_t>(a.nodes.size());
        for (bnode n : right_arena.nodes) {
            if (n.left >= 0) {  // interior: both child links are arena indices
                n.left += base;
                n.first += static_cast<uint32_t>(base);
            }
            a.nodes.push_back(n);
        }
        r = r_local + base;
    } else {
        l = build_subtree(a, first, lcount, depth + 1, 0);
        r = build_subtree(a, mid, rcount, depth + 1, 0);
    }
    a.nodes[root].left = l;
    a.nodes[root].first = static_cast<uint32_t>(r);  // right child index kept here for interior nodes
    a.nodes[root].count = 0;
    return root;
}

struct collapser {
    const std::vector<bnode>& bin;
    std::vector<bvh_node>& nodes;

    // bvh::collapse (bvh.hpp:205-242): open the widest interior child until
    // there are four children or only leaves.
    uint32_t collapse(int32_t b) {
        int32_t child[4] = {b, 0, 0, 0};
        int n = 1;
        while (n < 4) {
            int widest = -1;
            float widest_area = -1;
            for (int i = 0; i < n; ++i) {
                const bnode& c = bin[child[i]];
                if (c.count == 0 && c.box.half_area() > widest_area) {
                    widest = i;
                    widest_area = c.box.half_area();
                }
            }
            if (widest < 0) break;
            const bnode& w = bin[child[widest]];
            const int32_t left = w.left, right = static_cast<int32_t>(w.first);
            child[widest] = left;
            child[n++] = right;
        }
        const uint32_t index = static_cast<uint32_t>(nodes.size());
        nodes.emplace_back();
        bvh_node nd;
        const float inf = std::numeric_limits<float>::infinity();
        for (int i = 0; i < 4; ++i) {
            for (int a = 0; a < 3; ++a) {
                nd.lo[a][i] = i < n ? bin[child[i]].box.lo[a] : inf;
                nd.hi[a][i] = i < n ? bin[child[i]].box.hi[a] : inf;
            }
            nd.first[i] = 0;
            nd.count[i] = 0;
            if (i >= n) continue;
            const bnode& c = bin[child[i]];
            if (c.count > 0) {
                nd.first[i] = c.first;
                nd.count[i] = c.count;
            } else {
                nd.first[i] = collapse(child[i]);
            }
        }
        nodes[index] = nd;
        return index;
    }
};

}  // namespace

bvh_tree build_bvh(const std::vector<box3>& boxes) {
    bvh_tree out;
    out.order.resize(boxes.size());
    for (uint32_t i = 0; i < out.order.size(); ++i) out.order[i] = i;
    if (boxes.empty()) return out;
    builder b(boxes);
    b.centers.resize(boxes.size());
    for (size_t i = 0; i < boxes.size(); ++i) b.centers[i] = boxes[i].center();
    b.order = std::move(out.order);
    builder::arena a;
    a.nodes.reserve(2 * boxes.size());
    const int levels = boxes.size() > 400000 ? 3 : 0;
    const int32_t root = b.build_subtree(a, 0, static_cast<uint32_t>(boxes.size()), 0, levels);
    out.order = std::move(b.order);
    out.nodes.reserve(a.nodes.size() / 3 + 1);
    collapser c{a.nodes, out.nodes};
    c.collapse(root);
    return out;
}

int tree_depth(const std::vector<bvh_node>& nodes, uint32_t n) {
    int d = 0;
    for (int c = 0; c < 4; ++c)
        if (nodes[n].count[c] == 0 && nodes[n].lo[0][c] != std::numeric_limits<float>::infinity())
            d = std::max(d, tree_depth(nodes, nodes[n].first[c]));
    return d + 1;
}

box3 sphere_box(const sphere_rec& s) {
    const float r = std::fabs(s.radius);
    box3 b;
    b.lo = s.center - v3(r, r, r);
    b.hi = s.center + v3(r, r, r);
    return b;
}

box3 triangle_box(const triangle_rec& t) {
    const float pad = 1e-4f;
    box3 b;
    for (int a = 0; a < 3; ++a) {
        // std::min / std::max over an initializer list: the first smallest / largest.
        float lo = t.p[0][a], hi = t.p[0][a];
        for (int k = 1; k < 3; ++k) {
            if (t.p[k][a] < lo) lo = t.p[k][a];
            if (hi < t.p[k][a]) hi = t.p[k][a];
        }
        b.lo[a] = lo - pad;
        b.hi[a] = hi + pad;
    }
    return b;
}

}  // namespace pt
