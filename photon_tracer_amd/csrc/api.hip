// C ABI: device scenes, ray queries and rendering.
//
// A pt_scene holds the PTGP buffer in HBM as uploaded (the sections are used
// in place), plus one derived array: each leaf slot's primitive data in a
// dense 48-byte record, so a leaf's primitives are read contiguously rather
// than through order[] into 112-byte triangle records.
#include <algorithm>
#include <cstring>
#include <limits>
#include <vector>

#include "api_internal.hpp"

using ptapi::kMaxPaths;
using ptapi::kPathBytes;
using ptapi::render_bufs;

namespace {

struct section {
    uint32_t offset, bytes;
};

float fbits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Checks a PTGP v1 buffer (gpu_pack.hpp:177-203) well enough that no kernel
// can read outside it: section bounds, every reachable tree link and leaf
// range, primitive and material numbers. Returns the stack bound and builds
// the leaf records.
struct ptgp_check {
    const uint32_t* w;
    size_t n;
    section sec[8];
    uint32_t n_nodes, n_order, n_sph, n_tri, n_inst, n_mat, n_lights, n_large;
    uint32_t flat_root, inst_root, n_small;
    uint32_t tex_off = 0, tex_bytes = 0, n_tex = 0;
    std::vector<float> leaf;   // 12 floats per order slot
    int depth_flat = 0, depth_inst = 0, depth_blas = 0;

    const uint32_t* at(int s) const { return w + sec[s].offset / 4; }
    const float* node_f(uint32_t k) const { return reinterpret_cast<const float*>(at(0)) + 32 * k; }
    const uint32_t* node_u(uint32_t k) const { return at(0) + 32 * k; }

    static void bad(const std::string& m) { throw pt::status_error(PT_ERR_INVALID, "PTGP: " + m); }

    void header() {
        if (!w || n < 48) bad("buffer shorter than the 48-word header");
        if (w[0] != 0x50475450u) bad("bad magic (expected \"PTGP\")");
        if (w[1] != 1) bad("unsupported version");
        const uint32_t elem[8] = {128, 4, 32, 112, 112, 48, 4, 4};
        for (int s = 0; s < 8; ++s) {
            sec[s] = {w[2 + 2 * s], w[3 + 2 * s]};
            if (sec[s].offset % 16 || sec[s].offset < 192) bad("section offset not 16-byte aligned past the header");
            if (uint64_t(sec[s].offset) + sec[s].bytes > 4 * uint64_t(n)) bad("section runs past the buffer");
            if (sec[s].bytes % elem[s]) bad("section length not a whole number of records");
        }
        n_nodes = sec[0].bytes / 128;
        n_order = sec[1].bytes / 4;
        n_sph = sec[2].bytes / 32;
        n_tri = sec[3].bytes / 112;
        n_inst = sec[4].bytes / 112;
        n_mat = sec[5].bytes / 48;
        n_lights = sec[6].bytes / 4;
        n_large = sec[7].bytes / 4;
        flat_root = w[18];
        inst_root = w[19];
        n_small = w[20];
        if (n_small > n_sph) bad("more small spheres than spheres");
        if (n_nodes >= (1u << 28) || n_order >= (1u << 28)) bad("more than 2^28 nodes or leaf slots");
        // The texture section, this library's extension in words the
        // reference writes as 0: header words 38-40 (offset, bytes, count).
        tex_off = w[38];
        tex_bytes = w[39];
        n_tex = w[40];
        if (tex_off || tex_bytes || n_tex) {
            if (tex_off % 16 || tex_off < 192 || uint64_t(tex_off) + tex_bytes > 4 * uint64_t(n))
                bad("texture section out of the buffer or not 16-byte aligned");
            if (n_tex == 0 || tex_bytes % 16 || tex_bytes < 16ull * n_tex) bad("texture section size");
            const uint64_t texels = (tex_bytes - 16ull * n_tex) / 16;
            const uint32_t* info = w + tex_off / 4;
            for (uint32_t t = 0; t < n_tex; ++t) {
                const uint32_t tw = info[4 * t], th = info[4 * t + 1], first = info[4 * t + 2];
                if (tw < 1 || th < 1 || tw > 65536 || th > 65536) bad("texture size out of range");
                if (uint64_t(first) + uint64_t(tw) * th > texels) bad("texture texels out of the section");
            }
        }
        for (uint32_t m = 0; m < n_mat; ++m) {
            const float type = fbits(at(5)[12 * m + 3]);
            if (!(type == 0 || type == 1 || type == 2 || type == 3)) bad("material type not 0-3");
            const uint32_t tex = at(5)[12 * m + 9];
            if (tex && (type != 0 || tex > n_tex)) bad("material texture out of range or not on a matte material");
        }
        for (uint32_t s = 0; s < n_sph; ++s)
            if (at(2)[8 * s + 4] >= n_mat) bad("sphere material out of range");
        for (uint32_t t = 0; t < n_tri; ++t)
            if (at(3)[28 * t + 3] >= n_mat) bad("triangle material out of range");
        for (uint32_t k = 0; k < n_lights; ++k)
            if (at(6)[k] >= n_small) bad("light is not a small sphere");
        for (uint32_t k = 0; k < n_large; ++k)
            if (at(7)[k] >= n_sph) bad("large sphere out of range");
        for (uint32_t k = 0; k < n_inst; ++k) {
            const uint32_t root = at(4)[28 * k + 24], mat = at(4)[28 * k + 25];
            if (root != NONE && root >= n_nodes) bad("instance root out of range");
            if (mat != NONE && mat >= n_mat) bad("instance material out of range");
        }
        leaf.assign(size_t(n_order) * 12, 0.0f);
    }

    void put_triangle(uint32_t slot, uint32_t t) {
        const float* tr = reinterpret_cast<const float*>(at(3)) + 28 * t;
        float* r = &leaf[12 * size_t(slot)];
        r[0] = tr[0];
        r[1] = tr[1];
        r[2] = tr[2];
        std::memcpy(&r[3], &t, 4);
        r[4] = tr[4];
        r[5] = tr[5];
        r[6] = tr[6];
        r[7] = 0;
        r[8] = tr[8];
        r[9] = tr[9];
        r[10] = tr[10];
        r[11] = 0;
    }
    void put_sphere(uint32_t slot, uint32_t s) {
        const float* sp = reinterpret_cast<const float*>(at(2)) + 8 * s;
        float* r = &leaf[12 * size_t(slot)];
        r[0] = sp[0];
        r[1] = sp[1];
        r[2] = sp[2];
        std::memcpy(&r[3], &s, 4);
        r[4] = sp[3];
        const uint32_t one = 1;
        std::memcpy(&r[7], &one, 4);
    }

    // Walks one tree from root; kind 0 flat, 1 mesh, 2 instances. Returns its
    // depth in 4-wide levels; collects mesh roots met in instance leaves.
    int walk(uint32_t root, int kind, std::vector<uint8_t>& seen, std::vector<uint32_t>* blas_roots) {
        if (root == NONE) return 0;
        if (root >= n_nodes) bad("tree root out of range");
        struct item {
            uint32_t node;
            int level;
        };
        std::vector<item> todo{{root, 1}};
        std::vector<uint32_t> visited;
        int depth = 0;
        const float inf = std::numeric_limits<float>::infinity();
        while (!todo.empty()) {
            const item it = todo.back();
            todo.pop_back();
            if (seen[it.node]) bad("tree has a cycle or shared node");
            seen[it.node] = 1;
            visited.push_back(it.node);
            depth = std::max(depth, it.level);
            if (it.level > 64) bad("tree deeper than 64 levels");
            const float* f = node_f(it.node);
            const uint32_t* u = node_u(it.node);
            for (int c = 0; c < 4; ++c) {
                bool empty = true;
                for (int a = 0; a < 6; ++a) empty = empty && f[4 * a + c] == inf;
                const uint32_t first = u[24 + c], count = u[28 + c];
                if (count > 0) {
                    if (count > 15) bad("leaf holds more than 15 primitives");
                    if (uint64_t(first) + count > n_order) bad("leaf range out of range");
                    for (uint32_t i = first; i < first + count; ++i) {
                        const uint32_t p = at(1)[i];
                        if (kind == 0) {
                            if (p < n_small) put_sphere(i, p);
                            else if (p - n_small < n_tri) put_triangle(i, p - n_small);
                            else bad("flat leaf primitive out of range");
                        } else if (kind == 1) {
                            if (p >= n_tri) bad("mesh leaf triangle out of range");
                            put_triangle(i, p);
                        } else {
                            if (p >= n_inst) bad("instance leaf out of range");
                            if (blas_roots) blas_roots->push_back(at(4)[28 * p + 24]);
                        }
                    }
                } else if (!empty) {
                    if (first >= n_nodes) bad("child node out of range");
                    todo.push_back({first, it.level + 1});
                }
            }
        }
        for (uint32_t v : visited) seen[v] = 0;
        return depth;
    }

    int run() {
        header();
        std::vector<uint8_t> seen(n_nodes, 0);
        depth_flat = walk(flat_root, 0, seen, nullptr);
        std::vector<uint32_t> roots;
        depth_inst = walk(inst_root, 2, seen, &roots);
        std::sort(roots.begin(), roots.end());
        roots.erase(std::unique(roots.begin(), roots.end()), roots.end());
        for (uint32_t r : roots) depth_blas = std::max(depth_blas, walk(r, 1, seen, nullptr));
        const int flat_need = 3 * depth_flat + 1;
        const int inst_need = depth_inst ? 3 * depth_inst + 1 + 3 * depth_blas + 1 : 0;
        return std::max(flat_need, inst_need) + 4;
    }
};

// The device copy of a checked PTGP buffer's nodes, in place in `dev` (a
// copy of the buffer): the first TOP_NODES nodes met breadth first from the
// flat root and the instance root get numbers 0, 1, 2 ..., the rest keep
// their order after them, and every node link (child codes, the instances'
// mesh roots) follows; each child slot's first and count become one word
// (first | count << 28, the traversal stack's entry code) in place of
// `first`, so a node step loads 112 of the node's 128 bytes. Returns the new
// number of each node.
std::vector<uint32_t> device_nodes(const ptgp_check& chk, std::vector<uint32_t>& dev) {
    const uint32_t n = chk.n_nodes;
    const uint32_t* old = chk.at(0);
    const float inf = std::numeric_limits<float>::infinity();
    auto internal_child = [&](uint32_t k, int c) {
        const float* f = reinterpret_cast<const float*>(old) + 32 * size_t(k);
        bool empty = true;
        for (int a = 0; a < 6; ++a) empty = empty && f[4 * a + c] == inf;
        return !empty && old[32 * size_t(k) + 28 + c] == 0;
    };
    std::vector<uint32_t> map(n, NONE), top;
    std::vector<uint32_t> queue;
    for (uint32_t r : {chk.flat_root, chk.inst_root})
        if (r != NONE) queue.push_back(r);
    for (size_t q = 0; q < queue.size() && top.size() < ptd::TOP_NODES; ++q) {
        const uint32_t k = queue[q];
        if (map[k] != NONE) continue;
        map[k] = static_cast<uint32_t>(top.size());
        top.push_back(k);
        for (int c = 0; c < 4; ++c)
            if (internal_child(k, c)) queue.push_back(old[32 * size_t(k) + 24 + c]);
    }
    uint32_t next = static_cast<uint32_t>(top.size());
    for (uint32_t k = 0; k < n; ++k)
        if (map[k] == NONE) map[k] = next++;
    uint32_t* nd = dev.data() + chk.sec[0].offset / 4;
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t* to = nd + 32 * size_t(map[k]);
        std::memcpy(to, old + 32 * size_t(k), 128);
        for (int c = 0; c < 4; ++c) {
            uint32_t first = to[24 + c];
            if (internal_child(k, c)) first = map[first];
            to[24 + c] = first | (to[28 + c] << 28);
        }
    }
    uint32_t* in = dev.data() + chk.sec[4].offset / 4;
    for (uint32_t k = 0; k < chk.n_inst; ++k) {
        uint32_t& root = in[28 * size_t(k) + 24];
        if (root != NONE) root = map[root];
    }
    return map;
}

int choose_spb(int spp, size_t npix, int requested, size_t free_bytes, size_t extra_path_bytes, uint64_t tuned) {
    if (requested > 0) {
        if (spp % requested) throw pt::status_error(PT_ERR_INVALID, "samples_per_batch must divide spp");
        return requested;
    }
    size_t target = std::min(kMaxPaths, free_bytes / 4 / (kPathBytes + extra_path_bytes));
    if (tuned) target = std::min<size_t>(target, tuned);
    const size_t cap = std::max<size_t>(1, target / std::max<size_t>(1, npix));
    int best = 1;
    for (int d = 1; d <= spp; ++d)
        if (spp % d == 0 && size_t(d) <= cap) best = d;
    return best;
}

}  // namespace

namespace ptapi {

void check_ptgp(const uint32_t* w, size_t n) {
    ptgp_check chk{w, n};
    chk.run();
}

int stack_entries_for(pt_scene* s, int requested) {
    // Without instances the 8-entry ring runs at 8 waves per SIMD; with them
    // two trees are live at once and 16 entries spill less.
    const int fallback = s->view.inst_root == NONE ? 8 : 16;
    int S = requested > 0 ? requested : (s->tuned_stack > 0 ? s->tuned_stack : fallback);
    if (S != 8 && S != 16 && S != 32) throw pt::status_error(PT_ERR_INVALID, "stack_entries must be 8, 16 or 32");
    return S;
}

int trace_grid_for(pt_scene* s, int S) {
    const int per_cu = s->tuned_blocks > 0 ? s->tuned_blocks : ptd::trace_blocks_per_cu(S);
    return std::max(1, s->cus * per_cu);
}

void ensure_spill(pt_scene* s, int grid) {
    const size_t lanes = size_t(grid) * ptd::TRACE_BLOCK;
    s->spill.ensure(lanes * size_t(s->max_stack) * sizeof(uint2));
}

}  // namespace ptapi

using ptapi::ensure_spill;
using ptapi::stack_entries_for;
using ptapi::trace_grid_for;

extern "C" {

pt_status pt_device_info(char* name, size_t name_len, int* compute_units) {
    PT_GUARD({
        int dev = 0;
        HIPC(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIPC(hipGetDeviceProperties(&prop, dev));
        if (name && name_len) {
            std::strncpy(name, prop.gcnArchName, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (compute_units) *compute_units = prop.multiProcessorCount;
    });
}

pt_status pt_ptgp_depth(const uint32_t* ptgp, size_t n_words, int* depth) {
    if (!depth) return pt::fail(PT_ERR_INVALID, "pt_ptgp_depth: null output");
    PT_GUARD({
        ptgp_check chk{ptgp, n_words};
        chk.run();
        *depth = std::max({chk.depth_flat, chk.depth_inst, chk.depth_blas});
    });
}

pt_status pt_scene_create(const uint32_t* ptgp, size_t n_words, pt_scene** out) {
    if (!out) return pt::fail(PT_ERR_INVALID, "pt_scene_create: null output");
    *out = nullptr;
    pt_scene* s = nullptr;
    try {
        ptgp_check chk{ptgp, n_words};
        const int need = chk.run();
        if (need > 256) throw pt::status_error(PT_ERR_LIMIT, "PTGP: trees too deep for the traversal stack");
        s = new pt_scene();
        s->max_stack = need;
        HIPC(hipGetDevice(&s->device));
        hipDeviceProp_t prop;
        HIPC(hipGetDeviceProperties(&prop, s->device));
        s->cus = prop.multiProcessorCount;
        HIPC(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->host.assign(ptgp, ptgp + n_words);
        std::vector<uint32_t> dev(ptgp, ptgp + n_words);
        s->node_map = device_nodes(chk, dev);
        s->words.ensure(n_words * 4);
        HIPC(hipMemcpy(s->words.p, dev.data(), n_words * 4, hipMemcpyHostToDevice));
        s->leaf.ensure(chk.leaf.size() * 4);
        if (!chk.leaf.empty()) HIPC(hipMemcpy(s->leaf.p, chk.leaf.data(), chk.leaf.size() * 4, hipMemcpyHostToDevice));
        s->work.ensure(64);
        const uint32_t* dw = s->words.as<uint32_t>();
        auto sec = [&](int k) { return dw + chk.sec[k].offset / 4; };
        ptd::SceneView& v = s->view;
        v.nodes = reinterpret_cast<const ptd::Node*>(sec(0));
        v.order = sec(1);
        v.leaf = s->leaf.as<float4>();
        v.spheres = reinterpret_cast<const float4*>(sec(2));
        v.tris = reinterpret_cast<const float4*>(sec(3));
        v.inst = reinterpret_cast<const float4*>(sec(4));
        v.mats = reinterpret_cast<const float4*>(sec(5));
        v.lights = sec(6);
        v.large = sec(7);
        v.flat_root = s->device_node(chk.flat_root);
        v.inst_root = s->device_node(chk.inst_root);
        v.n_cached = 0;
        v.cache = 0;
        v.n_top = std::min<uint32_t>(ptd::TOP_NODES, chk.n_nodes);
        v.n_tex = chk.n_tex;
        v.tex_info = chk.n_tex ? reinterpret_cast<const uint4*>(dw + chk.tex_off / 4) : nullptr;
        v.texels = chk.n_tex ? reinterpret_cast<const float4*>(dw + chk.tex_off / 4) + chk.n_tex : nullptr;
        v.n_small = chk.n_small;
        v.flags = ptgp[21];
        v.n_lights = chk.n_lights;
        v.n_large = chk.n_large;
        for (int r = 0; r < 4; ++r)
            v.cam[r] = make_float4(fbits(ptgp[22 + 4 * r]), fbits(ptgp[23 + 4 * r]), fbits(ptgp[24 + 4 * r]),
                                   fbits(ptgp[25 + 4 * r]));
        *out = s;
        return PT_OK;
    } catch (const pt::status_error& e) {
        delete s;
        return pt::fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        delete s;
        return pt::fail(PT_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        delete s;
        return pt::fail(PT_ERR_INVALID, e.what());
    }
}

pt_status pt_scene_destroy(pt_scene* scene) {
    if (!scene) return PT_OK;
    PT_GUARD({
        (void)hipStreamSynchronize(scene->stream);
        delete scene;
    });
}

pt_status pt_trace_device(pt_scene* s, const void* d_rays, size_t n, int any_hit, void* d_hits, float* elapsed_ms,
                          uint64_t* visits) {
    if (!s) return pt::fail(PT_ERR_INVALID, "pt_trace_device: null scene");
    if (n == 0) {
        if (elapsed_ms) *elapsed_ms = 0;
        return PT_OK;
    }
    if (!d_rays || !d_hits) return pt::fail(PT_ERR_INVALID, "pt_trace_device: null buffer");
    if (n >= (size_t(1) << 32) - 64) return pt::fail(PT_ERR_LIMIT, "pt_trace_device: more than 2^32 rays");
    PT_GUARD({
        const int S = stack_entries_for(s, 0);
        const int grid = trace_grid_for(s, S);
        ensure_spill(s, grid);
        HIPC(hipMemsetAsync(s->work.p, 0, 4, s->stream));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (elapsed_ms) {
            HIPC(hipEventCreate(&e0));
            HIPC(hipEventCreate(&e1));
            HIPC(hipEventRecord(e0, s->stream));
        }
        unsigned long long* counts = nullptr;
        if (visits) {
            s->counts.ensure(64);
            HIPC(hipMemsetAsync(s->counts.p, 0, 64, s->stream));
            counts = s->counts.as<unsigned long long>();
        }
        HIPC(ptd::launch_trace(s->view, static_cast<const float4*>(d_rays), static_cast<uint32_t>(n),
                               static_cast<ptd::Hit*>(d_hits), any_hit != 0, s->spill.as<uint2>(), s->work.as<uint32_t>(),
                               counts, S, grid, s->stream));
        if (elapsed_ms) {
            HIPC(hipEventRecord(e1, s->stream));
            HIPC(hipEventSynchronize(e1));
            HIPC(hipEventElapsedTime(elapsed_ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        if (visits) {
            unsigned long long v[6];
            HIPC(hipMemcpyAsync(v, s->counts.p, 48, hipMemcpyDeviceToHost, s->stream));
            HIPC(hipStreamSynchronize(s->stream));
            for (int k = 0; k < 6; ++k) visits[k] = v[k];
        }
        HIPC(hipStreamSynchronize(s->stream));
    });
}

pt_status pt_trace(pt_scene* s, const pt_ray* rays, size_t n, int any_hit, pt_hit* hits) {
    if (!s) return pt::fail(PT_ERR_INVALID, "pt_trace: null scene");
    if (n == 0) return PT_OK;
    if (!rays || !hits) return pt::fail(PT_ERR_INVALID, "pt_trace: null buffer");
    static_assert(sizeof(pt_ray) == 32 && sizeof(pt_hit) == 24 && sizeof(ptd::Hit) == 24, "ABI record sizes");
    PT_GUARD({
        s->io_rays.ensure(n * sizeof(pt_ray));
        s->io_hits.ensure(n * sizeof(pt_hit));
        HIPC(hipMemcpyAsync(s->io_rays.p, rays, n * sizeof(pt_ray), hipMemcpyHostToDevice, s->stream));
        const pt_status st = pt_trace_device(s, s->io_rays.p, n, any_hit, s->io_hits.p, nullptr, nullptr);
        if (st != PT_OK) throw pt::status_error(st, pt_last_error());
        HIPC(hipMemcpyAsync(hits, s->io_hits.p, n * sizeof(pt_hit), hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
    });
}

pt_status pt_render(pt_scene* s, const pt_render_params* prm, float* rgb, float* aux, pt_render_stats* stats) {
    if (!s || !prm || !rgb) return pt::fail(PT_ERR_INVALID, "pt_render: null argument");
    PT_GUARD({
        const pt_render_params& P = *prm;
        if (P.width < 2 || P.height < 2 || P.width > 32768 || P.height > 32768)
            throw pt::status_error(PT_ERR_INVALID, "pt_render: width and height must be in [2, 32768]");
        if (P.spp < 1) throw pt::status_error(PT_ERR_INVALID, "pt_render: spp must be at least 1");
        if (P.max_bounces < 1 || P.max_bounces >= ptd::MAX_BOUNCES)
            throw pt::status_error(PT_ERR_INVALID, "pt_render: max_bounces must be in [1, 63]");
        int j_lo = P.row_begin, j_hi = P.row_end;
        if (j_lo == 0 && j_hi == 0) j_hi = P.height;
        if (j_lo < 0 || j_hi > P.height || j_lo >= j_hi)
            throw pt::status_error(PT_ERR_INVALID, "pt_render: row range outside the image");
        if (P.first_sample < 0) throw pt::status_error(PT_ERR_INVALID, "pt_render: first_sample must be >= 0");
        const size_t npix = size_t(P.width) * size_t(j_hi - j_lo);
        size_t free_bytes = 0, total_bytes = 0;
        HIPC(hipMemGetInfo(&free_bytes, &total_bytes));
        free_bytes += s->rb.bytes();  // buffers this scene already holds are reused
        const bool training = (P.flags & PT_RENDER_TRAIN) != 0;
        const bool guided = training || (P.flags & PT_RENDER_GUIDE) != 0;
        pt_guide* G = s->guide;
        if (guided && !G) throw pt::status_error(PT_ERR_INVALID, "pt_render: guided render without a guide (pt_scene_set_guide)");
        // Per path: the guide's outputs; while training, its records, their
        // weights and the compacted samples.
        const size_t guide_bytes = (guided ? 64 + 4 : 0) + (training ? ptd::GUIDE_RECORDS * (64 + 4 + 48) : 0);
        const int spb = choose_spb(P.spp, npix, P.samples_per_batch, free_bytes, guide_bytes, s->tuned_paths);
        const size_t paths = npix * size_t(spb);
        if (paths >= (size_t(1) << 31)) throw pt::status_error(PT_ERR_LIMIT, "pt_render: batch too large");
        const int S = stack_entries_for(s, P.stack_entries);
        const int trace_grid = trace_grid_for(s, S);
        const int shade_grid = std::max(1, s->cus * 2);  // blocks of 1024
        ensure_spill(s, trace_grid);
        const bool want_aux = aux != nullptr;

        render_bufs& B = s->rb;
        for (int k = 0; k < 2; ++k) {
            B.ro[k].ensure(paths * 16);
            B.rd[k].ensure(paths * 16);
            B.rpath[k].ensure(paths * 4);
        }
        B.hits.ensure(paths * sizeof(ptd::Hit));
        B.so.ensure(paths * 16);
        B.sd.ensure(paths * 16);
        B.sl.ensure(paths * 16);
        B.pa.ensure(paths * 16);
        B.pb.ensure(paths * 16);
        B.pkey.ensure(paths * 8);
        B.pq.ensure(paths * 8);
        if (want_aux) {
            B.paux.ensure(paths * 32);
            B.aux_accum.ensure(npix * 32);
            B.out_aux.ensure(npix * 7 * 4);
        }
        B.accum.ensure(npix * 16);
        B.ctrl.ensure(sizeof(ptd::Ctrl) * (ptd::MAX_BOUNCES + 1));
        B.frame.ensure(sizeof(ptd::FrameConst));
        B.totals.ensure(32);
        B.out_rgb.ensure(npix * 12);
        HIPC(hipMemsetAsync(B.ctrl.p, 0, sizeof(ptd::Ctrl) * (ptd::MAX_BOUNCES + 1), s->stream));

        ptd::FrameParams fp{};
        fp.sc = s->view;
        if (P.flags & PT_RENDER_NO_NEE) fp.sc.flags &= ~2u;
        fp.W = P.width;
        fp.H = P.height;
        fp.spp = P.spp;
        fp.depth = P.max_bounces;
        fp.j_lo = j_lo;
        fp.j_hi = j_hi;
        fp.spb = spb;
        fp.P = static_cast<uint32_t>(paths);
        fp.npix = static_cast<uint32_t>(npix);
        fp.nee = (fp.sc.flags & 2u) ? 1 : 0;
        fp.want_aux = want_aux ? 1 : 0;
        for (int k = 0; k < 2; ++k) {
            fp.ro[k] = B.ro[k].as<float4>();
            fp.rd[k] = B.rd[k].as<float4>();
            fp.rpath[k] = B.rpath[k].as<uint32_t>();
        }
        fp.hits = B.hits.as<ptd::Hit>();
        fp.so = B.so.as<float4>();
        fp.sd = B.sd.as<float4>();
        fp.sl = B.sl.as<float4>();
        fp.pa = B.pa.as<float4>();
        fp.pb = B.pb.as<float4>();
        fp.pkey = B.pkey.as<uint2>();
        fp.pq = B.pq.as<uint2>();
        fp.paux = want_aux ? B.paux.as<float4>() : nullptr;
        fp.accum = B.accum.as<float4>();
        fp.aux_accum = want_aux ? B.aux_accum.as<float4>() : nullptr;
        fp.ctrl = B.ctrl.as<ptd::Ctrl>();
        fp.frame = B.frame.as<ptd::FrameConst>();
        fp.totals = B.totals.as<unsigned long long>();
        fp.spill = s->spill.as<uint2>();
        // Below four rays per resident traversal lane, a queue no longer fills
        // the GPU stage by stage and the tail launch takes the paths.
        fp.tail_threshold = (P.flags & PT_RENDER_NO_TAIL) ? 0u : ptapi::tail_threshold(trace_grid, s->view.inst_root != NONE);
        fp.camera_packets = (P.flags & PT_RENDER_PER_RAY) ? 0u : 1u;
        if (guided) {
            fp.guide_mode = training ? 2 : 1;
            fp.gv = G->view();
            fp.gclamp = G->weight_clamp;
            fp.guide_bounces = G->bounces;
            B.gout.ensure(paths * ptd::guide::LOBES * 4);
            fp.gout = B.gout.as<uint32_t>();
            B.gmis.ensure(paths * 4);
            fp.gmis = B.gmis.as<float>();
            if (training) {
                const size_t nrec = paths * ptd::GUIDE_RECORDS;
                const size_t nblk = (nrec + ptd::GUIDE_TARGET_BLOCK - 1) / ptd::GUIDE_TARGET_BLOCK;
                B.grec.ensure(nrec * 64);
                B.gwt.ensure(nrec * 4);
                B.gblk.ensure(nblk * 4);
                B.gblk_sum.ensure(nblk * 4);
                B.gtot.ensure(16);
                B.gsamples.ensure(nrec * 48);
                fp.grec = B.grec.as<float4>();
                fp.gwt = B.gwt.as<float>();
                fp.gblk = B.gblk.as<uint32_t>();
                fp.gblk_sum = B.gblk_sum.as<float>();
                fp.gtot = B.gtot.as<uint32_t>();
                fp.gsamples = B.gsamples.as<float4>();
            }
        }
        const bool counting = stats && (P.flags & PT_RENDER_COUNT);
        if (counting) {
            s->counts.ensure(64);
            HIPC(hipMemsetAsync(s->counts.p, 0, 64, s->stream));
            fp.visits = s->counts.as<unsigned long long>();
        }

        const int batches = P.spp / spb;
        const bool timing = stats && (P.flags & PT_RENDER_TIMING);
        const bool use_graph = !timing && !training && !(P.flags & PT_RENDER_NO_GRAPH) && !counting;
        // A graph of a shape rendered before whose tail launch took over at
        // bounce t ends there: the tail takes over at t whatever the queue
        // sizes (its paths come out the same), and the launches after it,
        // which would return at once, are left out.
        const ptd::FrameParams shape_key = fp;
        if (use_graph && s->tail_hint != 0 && std::memcmp(&s->tail_hint_key, &shape_key, sizeof fp) == 0)
            fp.tail_last = s->tail_hint;
        if (use_graph) {
            // The batch's whole bounce loop as one graph, replayed per batch
            // and kept for the next render of the same shape; kernels read
            // queue sizes, the seed and the batch number from HBM. Captured
            // before the render's first event, so render_ms is device time
            // and not the host's capture.
            auto& g = s->graph;
            const bool reuse = g.exec && g.S == S && g.trace_grid == trace_grid && g.shade_grid == shade_grid &&
                               std::memcmp(&g.key, &fp, sizeof fp) == 0;
            if (!reuse) {
                s->drop_graph();
                hipGraph_t graph;
                HIPC(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                const hipError_t le = ptd::launch_batch(fp, S, trace_grid, shade_grid, s->stream, nullptr);
                const hipError_t ce = hipStreamEndCapture(s->stream, &graph);
                HIPC(le);
                HIPC(ce);
                const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) g.exec = nullptr;
                HIPC(ie);
                g.key = fp;
                g.S = S;
                g.trace_grid = trace_grid;
                g.shade_grid = shade_grid;
            }
        }
        hipEvent_t e0, e1;
        HIPC(hipEventCreate(&e0));
        HIPC(hipEventCreate(&e1));
        const ptd::FrameConst fc{P.seed, P.first_sample, 0u, ~0u, 0u};
        HIPC(hipMemcpyAsync(B.frame.p, &fc, sizeof fc, hipMemcpyHostToDevice, s->stream));
        HIPC(ptd::launch_clear_frame(fp, s->stream));
        HIPC(hipEventRecord(e0, s->stream));
        double extend_ms = 0, shadow_ms = 0, guide_ms = 0;
        if (timing) {
            // Eager, with an event pair around every traversal launch and,
            // in guided renders, every inference launch (pairs 2 depth ..).
            std::vector<hipEvent_t> ev(4 * P.max_bounces);
            for (auto& e : ev) HIPC(hipEventCreate(&e));
            const int guided_bounces = fp.guide_mode ? std::min(P.max_bounces, fp.guide_bounces) : 0;
            for (int b = 0; b < batches; ++b) {
                HIPC(ptd::launch_batch(fp, S, trace_grid, shade_grid, s->stream, ev.data()));
                HIPC(hipStreamSynchronize(s->stream));
                for (int k = 0; k < P.max_bounces; ++k) {
                    float ms = 0;
                    HIPC(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
                    extend_ms += ms;
                }
                for (int k = 0; k < guided_bounces; ++k) {
                    float ms = 0;
                    HIPC(hipEventElapsedTime(&ms, ev[2 * (P.max_bounces + k)], ev[2 * (P.max_bounces + k) + 1]));
                    guide_ms += ms;
                }
            }
            for (auto& e : ev) (void)hipEventDestroy(e);
        } else if (training) {
            // Eager: after each batch its records become samples, and the
            // guide takes Adam steps over them before the next batch.
            for (int b = 0; b < batches; ++b) {
                HIPC(ptd::launch_batch(fp, S, trace_grid, shade_grid, s->stream, nullptr));
                HIPC(ptd::launch_guide_targets(fp, s->stream));
                uint32_t tot[2] = {0, 0};
                HIPC(hipMemcpyAsync(tot, B.gtot.p, sizeof tot, hipMemcpyDeviceToHost, s->stream));
                HIPC(hipStreamSynchronize(s->stream));
                ptapi::guide_train_on(G, fp.gsamples, tot[0]);
            }
        } else if ((P.flags & PT_RENDER_NO_GRAPH) || counting) {
            for (int b = 0; b < batches; ++b)
                HIPC(ptd::launch_batch(fp, S, trace_grid, shade_grid, s->stream, nullptr));
        } else {
            for (int b = 0; b < batches; ++b) HIPC(hipGraphLaunch(s->graph.exec, s->stream));
        }
        HIPC(hipEventRecord(e1, s->stream));
        HIPC(hipEventSynchronize(e1));
        float render_ms = 0;
        HIPC(hipEventElapsedTime(&render_ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);

        if (P.flags & PT_RENDER_DEVICE_OUTPUT) {
            // rgb and aux are device addresses: the image stays in HBM.
            HIPC(ptd::launch_resolve(fp, rgb, aux, s->stream));
        } else {
            HIPC(ptd::launch_resolve(fp, B.out_rgb.as<float>(), want_aux ? B.out_aux.as<float>() : nullptr, s->stream));
            HIPC(hipMemcpyAsync(rgb, B.out_rgb.p, npix * 12, hipMemcpyDeviceToHost, s->stream));
            if (want_aux) HIPC(hipMemcpyAsync(aux, B.out_aux.p, npix * 28, hipMemcpyDeviceToHost, s->stream));
        }
        unsigned long long totals[4] = {0, 0, 0, 0};
        HIPC(hipMemcpyAsync(totals, B.totals.p, sizeof totals, hipMemcpyDeviceToHost, s->stream));
        HIPC(hipStreamSynchronize(s->stream));
        if (use_graph && fp.tail_last == 0) {
            // totals[3]: the latest bounce at which the tail took over in any batch (0xffff: not in every one).
            const bool hint = totals[3] >= 1 && totals[3] < static_cast<unsigned long long>(P.max_bounces);
            s->tail_hint = hint ? static_cast<uint32_t>(totals[3]) : 0u;
            s->tail_hint_key = shape_key;
        }
        if (stats) {
            stats->rays = totals[0] + totals[1];
            stats->primary_rays = totals[2];
            stats->shadow_rays = totals[1];
            stats->render_ms = render_ms;
            stats->extend_ms = extend_ms;
            stats->shadow_ms = shadow_ms;
            unsigned long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (counting) HIPC(hipMemcpy(v, s->counts.p, sizeof v, hipMemcpyDeviceToHost));
            for (int k = 0; k < 7; ++k) stats->visits[k] = v[k];
            stats->batches = batches;
            stats->samples_per_batch = spb;
            stats->guide_ms = guide_ms;
            stats->guide_queries = v[7];
        }
    });
}

pt_status pt_scene_autotune(pt_scene* s, const pt_render_params* params, pt_tune_result* result) {
    if (!s || !params) return pt::fail(PT_ERR_INVALID, "pt_scene_autotune: null argument");
    if (params->flags & (PT_RENDER_TRAIN | PT_RENDER_TIMING | PT_RENDER_COUNT | PT_RENDER_DEVICE_OUTPUT))
        return pt::fail(PT_ERR_INVALID, "pt_scene_autotune: TRAIN, TIMING, COUNT and DEVICE_OUTPUT renders are not tuned");
    const int old_stack = s->tuned_stack, old_blocks = s->tuned_blocks;
    const uint64_t old_paths = s->tuned_paths;
    try {
        pt_render_params P = *params;
        P.stack_entries = 0;
        P.samples_per_batch = 0;
        const int rows = (P.row_begin == 0 && P.row_end == 0) ? P.height : P.row_end - P.row_begin;
        std::vector<float> img(3 * size_t(std::max(0, rows)) * size_t(std::max(0, P.width)));
        struct cand {
            int S, per_cu;
            uint64_t paths;
            double mrays;
        };
        std::vector<cand> tried;
        uint64_t auto_paths = 0;
        // Each candidate renders the frame as given twice (the first captures
        // its graph) and is timed by the second's device time.
        auto rate = [&](int S, int per_cu, uint64_t paths) {
            s->tuned_stack = S;
            s->tuned_blocks = per_cu;
            s->tuned_paths = paths;
            pt_render_stats st{};
            for (int k = 0; k < 2; ++k) {
                const pt_status e = pt_render(s, &P, img.data(), nullptr, &st);
                if (e != PT_OK) throw pt::status_error(e, pt_last_error());
            }
            if (!paths && !auto_paths) auto_paths = uint64_t(st.samples_per_batch) * (img.size() / 3);
            tried.push_back({S, per_cu, paths, st.render_ms > 0 ? double(st.rays) / (st.render_ms * 1e3) : 0.0});
        };
        auto best = [&] {
            return *std::max_element(tried.begin(), tried.end(), [](const cand& x, const cand& y) { return x.mrays < y.mrays; });
        };
        // Stage 1: LDS stack entries x traversal blocks per CU.
        const bool inst = s->view.inst_root != NONE;
        for (int S : {8, 16, 32}) {
            const int occ = ptd::path_blocks_per_cu(S, inst);
            std::vector<int> blocks{occ};
            if (occ > 2) blocks.push_back(occ - 1);
            if (occ > 4) blocks.push_back(occ / 2);
            for (int pc : blocks) rate(S, pc, 0);
        }
        cand c = best();
        // Stage 2: smaller batches (queue sizes) than the memory-sized one.
        for (uint64_t div : {2, 4})
            if (auto_paths / div >= img.size() / 3) rate(c.S, c.per_cu, auto_paths / div);
        c = best();
        s->tuned_stack = c.S;
        s->tuned_blocks = c.per_cu;
        s->tuned_paths = c.paths;
        if (result) {
            *result = pt_tune_result{};
            result->stack_entries = c.S;
            result->blocks_per_cu = c.per_cu;
            result->streams = 1;
            result->max_paths = c.paths;
            result->n_tried = static_cast<int32_t>(std::min<size_t>(tried.size(), 32));
            for (int k = 0; k < result->n_tried; ++k) {
                result->tried[k].stack_entries = tried[k].S;
                result->tried[k].blocks_per_cu = tried[k].per_cu;
                result->tried[k].streams = 1;
                result->tried[k].max_paths = tried[k].paths;
                result->tried[k].mrays_per_s = tried[k].mrays;
            }
        }
        return PT_OK;
    } catch (const pt::status_error& e) {
        s->tuned_stack = old_stack;
        s->tuned_blocks = old_blocks;
        s->tuned_paths = old_paths;
        return pt::fail(e.code, e.what());
    } catch (const std::exception& e) {
        s->tuned_stack = old_stack;
        s->tuned_blocks = old_blocks;
        s->tuned_paths = old_paths;
        return pt::fail(PT_ERR_INVALID, e.what());
    }
}

pt_status pt_device_alloc(size_t bytes, void** ptr) {
    if (!ptr) return pt::fail(PT_ERR_INVALID, "pt_device_alloc: null output");
    *ptr = nullptr;
    PT_GUARD(HIPC(hipMalloc(ptr, bytes ? bytes : 16)));
}

pt_status pt_device_free(void* ptr) { PT_GUARD(if (ptr) HIPC(hipFree(ptr))); }

pt_status pt_device_copy(void* dst, const void* src, size_t bytes, int kind) {
    if ((!dst || !src) && bytes) return pt::fail(PT_ERR_INVALID, "pt_device_copy: null pointer");
    if (kind < 0 || kind > 2) return pt::fail(PT_ERR_INVALID, "pt_device_copy: kind must be 0, 1 or 2");
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    PT_GUARD(if (bytes) HIPC(hipMemcpy(dst, src, bytes, k)));
}

}  // extern "C"
