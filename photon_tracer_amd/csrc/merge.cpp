// Several flat PTGP v1 scenes as one PTGP v1 buffer, for batched renders:
// one device scene view, whose kernels start each query at its job's root.
#include <algorithm>
#include <limits>

#include "host.hpp"

namespace pt {

namespace {

// The sections of one input scene (the PTGP v1 header, gpu_pack.hpp:177-203).
struct ptgp_sections {
    const uint32_t* w;
    const uint32_t* at[8];
    uint32_t count[8];
    uint32_t flat_root, inst_root, n_small;
};

ptgp_sections sections_of(const ptgp_in& s) {
    static const uint32_t elem[8] = {128, 4, 32, 112, 112, 48, 4, 4};
    ptgp_sections r{};
    r.w = s.w;
    if (!s.w || s.n < 48 || s.w[0] != kPtgpMagic || s.w[1] != 1) throw std::runtime_error("not a PTGP v1 buffer");
    for (int k = 0; k < 8; ++k) {
        const uint32_t off = s.w[2 + 2 * k], bytes = s.w[3 + 2 * k];
        if (off % 4 || uint64_t(off) + bytes > 4 * uint64_t(s.n) || bytes % elem[k])
            throw std::runtime_error("PTGP section outside the buffer");
        r.at[k] = s.w + off / 4;
        r.count[k] = bytes / elem[k];
    }
    r.flat_root = s.w[18];
    r.inst_root = s.w[19];
    r.n_small = s.w[20];
    return r;
}

// The box of everything under node k: the union of its children's boxes.
void node_box(const bvh_node& n, float lo[3], float hi[3]) {
    const float inf = std::numeric_limits<float>::infinity();
    for (int a = 0; a < 3; ++a) {
        lo[a] = inf;
        hi[a] = -inf;
    }
    for (int c = 0; c < 4; ++c) {
        if (n.lo[0][c] == inf) continue;   // empty child (bvh.hpp's padding)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], n.lo[a][c]);
            hi[a] = std::max(hi[a], n.hi[a][c]);
        }
    }
}

}  // namespace

std::vector<uint32_t> merge_ptgp(const std::vector<ptgp_in>& in, std::vector<merge_part>& parts) {
    const float inf = std::numeric_limits<float>::infinity();
    const size_t K = in.size();
    std::vector<ptgp_sections> S(K);
    uint64_t n_nodes = 0, n_order = 0, n_small = 0, n_big = 0, n_tri = 0, n_mat = 0, n_lights = 0, n_large = 0;
    for (size_t k = 0; k < K; ++k) {
        S[k] = sections_of(in[k]);
        const ptgp_sections& s = S[k];
        if (s.inst_root != ~0u || s.count[4] != 0) throw std::runtime_error("batched scenes must have no instances");
        if (s.w[38] || s.w[39] || s.w[40]) throw std::runtime_error("batched scenes must have no textures");
        if (s.n_small > s.count[2]) throw std::runtime_error("more small spheres than spheres");
        n_nodes += s.count[0];
        n_order += s.count[1];
        n_small += s.n_small;
        n_big += s.count[2] - s.n_small;
        n_tri += s.count[3];
        n_mat += s.count[5];
        n_lights += s.count[6];
        n_large += s.count[7];
    }
    if (n_nodes + 2 * K + 2 >= (1u << 28) || n_order >= (1u << 28) || n_small + n_tri >= 0xffffffffull)
        throw std::runtime_error("batched scenes too large together");

    std::vector<bvh_node> nodes;
    std::vector<uint32_t> order, spheres(8 * (n_small + n_big)), tris, mats, lights, large;
    nodes.reserve(n_nodes + K);
    order.reserve(n_order);
    tris.reserve(28 * n_tri);
    mats.reserve(12 * n_mat);
    parts.assign(K, merge_part{});

    uint32_t node_off = 0, order_off = 0, small_off = 0, big_off = 0, tri_off = 0, mat_off = 0;
    std::vector<uint32_t> roots;
    for (size_t k = 0; k < K; ++k) {
        const ptgp_sections& s = S[k];
        const uint32_t ns = s.n_small;
        // A scene's sphere s: its small spheres first, then its others, each
        // group after the earlier scenes' (the merged buffer keeps every small
        // sphere before every other, as pack_for_gpu does).
        auto sphere_at = [&](uint32_t x) {
            return x < ns ? small_off + x : static_cast<uint32_t>(n_small) + big_off + (x - ns);
        };
        for (uint32_t x = 0; x < s.count[2]; ++x) {
            uint32_t* o = &spheres[8 * size_t(sphere_at(x))];
            std::copy(s.at[2] + 8 * size_t(x), s.at[2] + 8 * size_t(x) + 8, o);
            o[4] += mat_off;
        }
        for (uint32_t t = 0; t < s.count[3]; ++t) {
            const uint32_t* r = s.at[3] + 28 * size_t(t);
            tris.insert(tris.end(), r, r + 28);
            tris[tris.size() - 28 + 3] += mat_off;
        }
        mats.insert(mats.end(), s.at[5], s.at[5] + 12 * size_t(s.count[5]));
        merge_part& p = parts[k];
        p.lights_off = static_cast<uint32_t>(lights.size());
        p.n_lights = s.count[6];
        for (uint32_t q = 0; q < s.count[6]; ++q) lights.push_back(sphere_at(s.at[6][q]));
        p.large_off = static_cast<uint32_t>(large.size());
        p.n_large = s.count[7];
        for (uint32_t q = 0; q < s.count[7]; ++q) large.push_back(sphere_at(s.at[7][q]));
        // Flat tree leaves hold sphere numbers below n_small and n_small +
        // triangle numbers above (gpu_pack.hpp:96-110).
        for (uint32_t q = 0; q < s.count[1]; ++q) {
            const uint32_t x = s.at[1][q];
            order.push_back(x < ns ? small_off + x : static_cast<uint32_t>(n_small) + tri_off + (x - ns));
        }
        // Links rebased as pack_for_gpu's append does (gpu_pack.hpp:58-73).
        for (uint32_t q = 0; q < s.count[0]; ++q) {
            bvh_node n;
            std::memcpy(&n, s.at[0] + 32 * size_t(q), sizeof n);
            for (int c = 0; c < 4; ++c) {
                if (n.count[c] > 0) n.first[c] += order_off;
                else if (n.lo[0][c] != inf) n.first[c] += node_off;
            }
            nodes.push_back(n);
        }
        p.root = s.flat_root == ~0u ? ~0u : s.flat_root + node_off;
        if (p.root != ~0u) roots.push_back(p.root);
        p.flags = s.w[21];
        std::memcpy(p.camera, s.w + 22, sizeof p.camera);
        node_off += s.count[0];
        order_off += s.count[1];
        small_off += ns;
        big_off += s.count[2] - ns;
        tri_off += s.count[3];
        mat_off += s.count[5];
    }

    // Nodes above the scenes' roots, four children each, so that the merged
    // flat root reaches every scene's tree (a render of the merged scene sees
    // them all; batched renders start at their job's root).
    std::vector<uint32_t> level = roots;
    while (level.size() > 1) {
        std::vector<uint32_t> up;
        for (size_t i = 0; i < level.size(); i += 4) {
            bvh_node n;
            for (int c = 0; c < 4; ++c) {
                const bool has = i + c < level.size();
                float lo[3] = {inf, inf, inf}, hi[3] = {inf, inf, inf};
                if (has) node_box(nodes[level[i + c]], lo, hi);
                // A child whose tree has no boxes at all still gets a finite
                // (point) box, so it is not read as an empty slot.
                for (int a = 0; a < 3; ++a) {
                    if (has && lo[a] > hi[a]) lo[a] = hi[a] = 0;
                    n.lo[a][c] = lo[a];
                    n.hi[a][c] = hi[a];
                }
                n.first[c] = has ? level[i + c] : 0;
                n.count[c] = 0;
            }
            up.push_back(static_cast<uint32_t>(nodes.size()));
            nodes.push_back(n);
        }
        level.swap(up);
    }
    const uint32_t flat_root = level.empty() ? ~0u : level[0];

    std::vector<uint32_t> out(48, 0);
    out[0] = kPtgpMagic;
    out[1] = 1;
    auto add = [&](int slot, const void* data, size_t bytes) {
        if (4 * out.size() + bytes + 16 > 0xffffffffull)
            throw std::runtime_error("batched scenes too large for the PTGP v1 32-bit offsets");
        out[2 + 2 * slot] = static_cast<uint32_t>(4 * out.size());
        out[3 + 2 * slot] = static_cast<uint32_t>(bytes);
        const size_t at = out.size();
        out.resize(at + (bytes + 15) / 16 * 4, 0);
        if (bytes) std::memcpy(&out[at], data, bytes);
    };
    add(0, nodes.data(), nodes.size() * sizeof(bvh_node));
    add(1, order.data(), order.size() * 4);
    add(2, spheres.data(), spheres.size() * 4);
    add(3, tris.data(), tris.size() * 4);
    add(4, nullptr, 0);
    add(5, mats.data(), mats.size() * 4);
    add(6, lights.data(), lights.size() * 4);
    add(7, large.data(), large.size() * 4);
    out[18] = flat_root;
    out[19] = ~0u;
    out[20] = static_cast<uint32_t>(n_small);
    // The merged scene's own flags and camera: the first scene's.
    if (K) std::memcpy(&out[21], S[0].w + 21, 17 * 4);
    return out;
}

}  // namespace pt
