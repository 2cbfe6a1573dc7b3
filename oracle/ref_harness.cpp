// Test infrastructure, not product code: a C-callable harness over the
// reference's own headers, compiled from where they lie (see oracle/Makefile;
// the output goes to oracle/_ref/ only). It lets the tests
//   - build the reference's geometry from a PTSD scene and pack it with the
//     reference's pack_for_gpu, to compare with this library's packer;
//   - build the reference's own scenes (build_world, renderer.hpp:335-343) and
//     export them as PTSD and PTGP;
//   - answer ray queries with the reference's traversal and intersection
//     code, reporting which primitive was hit;
//   - render with the reference's render_row_linear on host threads, the way
//     its CLI schedules rows (main.cpp:127-170), for the CPU baseline;
//   - draw the reference's QMC samples.
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "denoise.hpp"
#include "gpu_pack.hpp"
#include "renderer.hpp"

namespace {

thread_local std::string last_error;

float fb(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
uint32_t ub(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

uint32_t* copy_out(const std::vector<uint32_t>& v, size_t* n) {
    uint32_t* p = static_cast<uint32_t*>(std::malloc(v.size() * 4 + 4));
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * 4);
    *n = v.size();
    return p;
}

struct ref_world {
    geometry world;
    camera cam;
    uint32_t n_flat_tris = 0;
    std::vector<uint32_t> mesh_base;  // PTGP triangle number of each mesh's first triangle
};

// PTSD (this library's scene description; layout in photon_tracer.h) into the
// reference's geometry, then geometry::build(true).
ref_world* from_ptsd(const uint32_t* w, size_t n) {
    if (n < 32 || w[0] != 0x44535450u || w[1] != 1) throw std::runtime_error("not a PTSD v1 buffer");
    auto* rw = new ref_world();
    geometry& g = rw->world;
    g.sky = (w[2] & 1u) != 0;
    g.sample_lights = (w[2] & 2u) != 0;
    const point3 from(fb(w[3]), fb(w[4]), fb(w[5])), at(fb(w[6]), fb(w[7]), fb(w[8]));
    rw->cam = camera(from, at, static_cast<double>(fb(w[9])));
    const uint32_t n_mat = w[10], n_sph = w[11], n_large = w[12], n_tri = w[13], n_mesh = w[14], n_inst = w[15];
    // The textures (PTSD word 17 of them) follow the instances: the reference's
    // image_texture made from each image8, as its OBJ loader makes them.
    std::vector<std::shared_ptr<const image_texture>> textures;
    {
        size_t t = 32 + 12 * size_t(n_mat) + 5 * size_t(n_sph + n_large) + 28 * size_t(n_tri) + n_mesh +
                   28 * size_t(w[16]) + 16 * size_t(n_inst);
        for (uint32_t i = 0; i < w[17]; ++i) {
            if (t + 2 > n) throw std::runtime_error("PTSD textures truncated");
            image8 img;
            img.width = static_cast<int>(w[t]);
            img.height = static_cast<int>(w[t + 1]);
            img.rgb.resize(size_t(img.width) * img.height * 3);
            if (t + 2 + (img.rgb.size() + 3) / 4 > n) throw std::runtime_error("PTSD textures truncated");
            std::memcpy(img.rgb.data(), w + t + 2, img.rgb.size());
            t += 2 + (img.rgb.size() + 3) / 4;
            textures.push_back(std::make_shared<image_texture>(img));
        }
    }
    size_t k = 32;
    std::vector<const material*> mats;
    for (uint32_t m = 0; m < n_mat; ++m, k += 12) {
        const color albedo(fb(w[k]), fb(w[k + 1]), fb(w[k + 2]));
        const uint32_t type = w[k + 3];
        const color emission(fb(w[k + 4]), fb(w[k + 5]), fb(w[k + 6]));
        const float fuzz = fb(w[k + 7]), ior = fb(w[k + 8]);
        const uint32_t tex = w[k + 9];
        if (type == 0) mats.push_back(g.own(std::make_shared<lambertian>(albedo, tex ? textures.at(tex - 1) : nullptr)));
        else if (type == 1) mats.push_back(g.own(std::make_shared<metal>(albedo, fuzz)));
        else if (type == 2) mats.push_back(g.own(std::make_shared<dielectric>(ior)));
        else mats.push_back(g.own(std::make_shared<diffuse_light>(emission)));
    }
    for (uint32_t s = 0; s < n_sph; ++s, k += 5)
        g.spheres.emplace_back(point3(fb(w[k]), fb(w[k + 1]), fb(w[k + 2])), fb(w[k + 3]), mats.at(w[k + 4]));
    for (uint32_t s = 0; s < n_large; ++s, k += 5)
        g.large_spheres.emplace_back(point3(fb(w[k]), fb(w[k + 1]), fb(w[k + 2])), fb(w[k + 3]), mats.at(w[k + 4]));
    auto read_tri = [&](size_t at) {
        auto v = [&](size_t o) { return vec3(fb(w[at + o]), fb(w[at + o + 1]), fb(w[at + o + 2])); };
        triangle t(v(0), v(3), v(6), v(9), v(12), v(15), mats.at(w[at + 24]));
        for (int c = 0; c < 6; ++c) t.uv[c] = fb(w[at + 18 + c]);
        return t;
    };
    for (uint32_t t = 0; t < n_tri; ++t, k += 28) g.triangles.push_back(read_tri(k));
    rw->n_flat_tris = n_tri;
    std::vector<uint32_t> sizes(n_mesh);
    for (auto& s : sizes) s = w[k++];
    uint32_t base = n_tri;
    for (uint32_t m = 0; m < n_mesh; ++m) {
        std::vector<triangle> tris;
        for (uint32_t t = 0; t < sizes[m]; ++t, k += 28) tris.push_back(read_tri(k));
        rw->mesh_base.push_back(base);
        base += sizes[m];
        g.add_mesh(std::move(tris));
    }
    for (uint32_t i = 0; i < n_inst; ++i, k += 16) {
        affine a;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) a.m[r][c] = fb(w[k + 2 + 3 * r + c]);
        a.t = vec3(fb(w[k + 11]), fb(w[k + 12]), fb(w[k + 13]));
        g.add_instance(w[k], a, w[k + 1] == ~0u ? nullptr : mats.at(w[k + 1]));
    }
    for (const auto& tex : textures) k += 2 + (tex->source.rgb.size() + 3) / 4;
    if (k != n) throw std::runtime_error("PTSD size mismatch");
    g.build(true);
    return rw;
}

// The reference's cameras for its scenes (renderer.hpp:212-218).
void camera_params(scene_id id, float from[3], float at[3], float* vfov) {
    struct c { float f[3], a[3], v; };
    c x = {{0, 0, 0}, {0, 0, -1}, 90};
    if (id == scene_id::field) x = {{0, 1.0f, 1.2f}, {0, -0.3f, -1.2f}, 55};
    if (id == scene_id::mesh) x = {{0, 0.5f, 1.2f}, {0, -0.15f, -1.2f}, 45};
    if (id == scene_id::room) x = {{0, 0, 0.3f}, {0, -0.15f, -3.0f}, 62};
    if (id == scene_id::crowd) x = {{0, 2.4f, 3.2f}, {0, -0.6f, -7.0f}, 52};
    for (int i = 0; i < 3; ++i) {
        from[i] = x.f[i];
        at[i] = x.a[i];
    }
    *vfov = x.v;
}

geometry reference_scene(int id, int crowd_side, int flatten, const char* obj_path, double turn_deg) {
    std::string text;
    std::string dir;
    if (obj_path && *obj_path) {
        std::ifstream f(obj_path);
        if (!f) throw std::runtime_error(std::string("cannot open ") + obj_path);
        std::ostringstream b;
        b << f.rdbuf();
        text = b.str();
        const std::string p(obj_path);
        dir = p.find('/') == std::string::npos ? "" : p.substr(0, p.find_last_of('/') + 1);
    }
    const file_reader read = [dir](const std::string& name) {
        std::ifstream f(dir + name, std::ios::binary);
        if (!f) return std::string();
        std::ostringstream b;
        b << f.rdbuf();
        return b.str();
    };
    return build_world(static_cast<scene_id>(id), true, text, read, turn_deg, crowd_side, flatten != 0);
}

// The reference geometry's flat arrays as PTSD.
std::vector<uint32_t> to_ptsd(const geometry& g, const float from[3], const float at[3], float vfov) {
    std::map<const material*, uint32_t> index;
    std::vector<const material*> order;
    auto mat = [&](const material* m) {
        auto it = index.find(m);
        if (it != index.end()) return it->second;
        const uint32_t k = static_cast<uint32_t>(order.size());
        index.emplace(m, k);
        order.push_back(m);
        return k;
    };
    std::vector<uint32_t> body;
    auto put = [&](float f) { body.push_back(ub(f)); };
    std::vector<uint32_t> sph, lrg, tri, mesh_sizes, mesh_tris, inst;
    auto sphere_words = [&](std::vector<uint32_t>& o, const sphere& s) {
        for (int c = 0; c < 3; ++c) o.push_back(ub(s.center.e[c]));
        o.push_back(ub(s.radius));
        o.push_back(mat(s.mat));
    };
    auto tri_words = [&](std::vector<uint32_t>& o, const triangle& t) {
        const vec3 v[6] = {t.p0, t.p1, t.p2, t.n0, t.n1, t.n2};
        for (const auto& x : v)
            for (int c = 0; c < 3; ++c) o.push_back(ub(x.e[c]));
        for (int c = 0; c < 6; ++c) o.push_back(ub(t.uv[c]));
        o.push_back(mat(t.mat));
        o.push_back(0);
        o.push_back(0);
        o.push_back(0);
    };
    for (const auto& s : g.spheres) sphere_words(sph, s);
    for (const auto& s : g.large_spheres) sphere_words(lrg, s);
    for (const auto& t : g.triangles) tri_words(tri, t);
    for (const auto& m : g.meshes) {
        mesh_sizes.push_back(static_cast<uint32_t>(m.triangles.size()));
        for (const auto& t : m.triangles) tri_words(mesh_tris, t);
    }
    for (const auto& in : g.instances) {
        inst.push_back(in.mesh);
        inst.push_back(in.mat ? mat(in.mat) : ~0u);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) inst.push_back(ub(in.to_world.m[r][c]));
        for (int c = 0; c < 3; ++c) inst.push_back(ub(in.to_world.t.e[c]));
        inst.push_back(0);
        inst.push_back(0);
    }
    std::vector<uint32_t> out(32, 0);
    out[0] = 0x44535450u;
    out[1] = 1;
    out[2] = (g.sky ? 1u : 0u) | (g.sample_lights ? 2u : 0u);
    for (int c = 0; c < 3; ++c) {
        out[3 + c] = ub(from[c]);
        out[6 + c] = ub(at[c]);
    }
    out[9] = ub(vfov);
    out[10] = static_cast<uint32_t>(order.size());
    out[11] = static_cast<uint32_t>(g.spheres.size());
    out[12] = static_cast<uint32_t>(g.large_spheres.size());
    out[13] = static_cast<uint32_t>(g.triangles.size());
    out[14] = static_cast<uint32_t>(g.meshes.size());
    out[15] = static_cast<uint32_t>(g.instances.size());
    out[16] = static_cast<uint32_t>(mesh_tris.size() / 28);
    std::map<const image_texture*, uint32_t> tex_index;   // texture number + 1
    std::vector<const image_texture*> tex_order;
    for (const material* m : order) {
        color albedo(0, 0, 0);
        uint32_t type = 3, tex = 0;
        float fuzz = 0, ior = 1.5f;
        if (auto* l = dynamic_cast<const lambertian*>(m)) {
            type = 0;
            albedo = l->albedo;
            if (l->texture) {
                auto it = tex_index.find(l->texture.get());
                if (it == tex_index.end()) {
                    tex_order.push_back(l->texture.get());
                    it = tex_index.emplace(l->texture.get(), static_cast<uint32_t>(tex_order.size())).first;
                }
                tex = it->second;
            }
        } else if (auto* me = dynamic_cast<const metal*>(m)) {
            type = 1;
            albedo = me->albedo;
            fuzz = me->fuzz;
        } else if (auto* d = dynamic_cast<const dielectric*>(m)) {
            type = 2;
            ior = d->ior;
        }
        for (int c = 0; c < 3; ++c) put(albedo.e[c]);
        body.push_back(type);
        for (int c = 0; c < 3; ++c) put(m->emission.e[c]);
        put(fuzz);
        put(ior);
        body.push_back(tex);
        body.push_back(0);
        body.push_back(0);
    }
    for (auto* part : {&sph, &lrg, &tri, &mesh_sizes, &mesh_tris, &inst}) body.insert(body.end(), part->begin(), part->end());
    for (const image_texture* t : tex_order) {   // the decoded image, as image_texture keeps it (texture.hpp:36-37)
        const image8& img = t->source;
        body.push_back(static_cast<uint32_t>(img.width));
        body.push_back(static_cast<uint32_t>(img.height));
        const size_t at = body.size();
        body.resize(at + (img.rgb.size() + 3) / 4, 0);
        std::memcpy(&body[at], img.rgb.data(), img.rgb.size());
    }
    out[17] = static_cast<uint32_t>(tex_order.size());
    out.insert(out.end(), body.begin(), body.end());
    return out;
}

struct rhit {
    uint32_t kind = 0, index = 0, sub = 0;
    float t = 0, u = 0, v = 0;
};

// geometry::hit (geometry.hpp:111-152) restated over its public members, so
// that the primitive is reported; the reference's own traversal and
// intersection routines do the work.
rhit ref_closest(const geometry& g, const ray& r, real t_min, real t_max) {
    rhit h;
    for (uint32_t i = 0; i < g.large_spheres.size(); ++i)
        if (g.large_spheres[i].intersect<double>(r, t_min, t_max)) {
            h.kind = 1;
            h.index = i;
        }
    const uint32_t n_spheres = static_cast<uint32_t>(g.spheres.size());
    real u = 0, v = 0;
    auto test = [&](uint32_t p, real& t_far) {
        if (p < n_spheres) {
            if (g.spheres[p].intersect(r, t_min, t_far)) {
                h.kind = 2;
                h.index = p;
            }
        } else if (g.triangles[p - n_spheres].intersect(r, t_min, t_far, u, v)) {
            h.kind = 3;
            h.index = p - n_spheres;
        }
    };
    auto test_instance = [&](uint32_t k, real& t_far) {
        const auto& in = g.instances[k];
        const auto& m = g.meshes[in.mesh];
        const ray local(in.to_object.point(r.origin()), in.to_object.vector(r.direction()));
        bool found = false;
        real ht = t_far, hu = 0, hv = 0;
        uint32_t tri = 0;
        m.tree.traverse(local, t_min, ht, [&](uint32_t q, real& t_mesh) {
            if (m.triangles[q].intersect(local, t_min, t_mesh, hu, hv)) {
                found = true;
                tri = q;
            }
        });
        if (found) {
            h.kind = 4;
            h.index = k;
            h.sub = tri;
            t_far = ht;
            u = hu;
            v = hv;
        }
    };
    g.flat_tree().traverse(r, t_min, t_max, test);
    if (!g.instances.empty()) g.instances_tree().traverse(r, t_min, t_max, test_instance);
    h.t = t_max;
    h.u = u;
    h.v = v;
    return h;
}

}  // namespace

extern "C" {

const char* ref_last_error() { return last_error.c_str(); }
void ref_free(void* p) { std::free(p); }

void* ref_world_create(const uint32_t* ptsd, size_t n) {
    try {
        return from_ptsd(ptsd, n);
    } catch (const std::exception& e) {
        last_error = e.what();
        return nullptr;
    }
}

void ref_world_destroy(void* w) { delete static_cast<ref_world*>(w); }

// serialize(pack_for_gpu(world, cam)); fails (returns 1) where the reference
// refuses a deep tree.
int ref_world_pack(void* w, uint32_t** out, size_t* n) {
    try {
        auto* rw = static_cast<ref_world*>(w);
        *out = copy_out(serialize(pack_for_gpu(rw->world, rw->cam)), n);
        return 0;
    } catch (const std::exception& e) {
        last_error = e.what();
        return 1;
    }
}

// A tree's nodes (32 words each) and order: which = -1 flat, -2 instances,
// k >= 0 mesh k.
int ref_world_tree(void* w, int which, uint32_t** nodes, size_t* n_nodes_words, uint32_t** order, size_t* n_order) {
    auto* rw = static_cast<ref_world*>(w);
    const bvh* t = which == -1 ? &rw->world.flat_tree() : which == -2 ? &rw->world.instances_tree()
                   : which < static_cast<int>(rw->world.meshes.size()) ? &rw->world.meshes[which].tree : nullptr;
    if (!t) {
        last_error = "no such tree";
        return 1;
    }
    std::vector<uint32_t> nw(t->nodes.size() * 32);
    if (!nw.empty()) std::memcpy(nw.data(), t->nodes.data(), nw.size() * 4);
    *nodes = copy_out(nw, n_nodes_words);
    *order = copy_out(t->order, n_order);
    return 0;
}

// Camera rows of the world's camera: origin, corner, across, up (12 floats).
void ref_world_camera(void* w, float* out) {
    auto* rw = static_cast<ref_world*>(w);
    const vec3 rows[4] = {rw->cam.position(), rw->cam.corner(), rw->cam.across(), rw->cam.up()};
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 3; ++c) out[3 * r + c] = rows[r].e[c];
}

// One of the reference's scenes (0 materials, 1 field, 2 mesh, 3 room,
// 4 crowd) with its camera, as PTSD, and packed by the reference as PTGP.
int ref_scene(int id, int crowd_side, int flatten, const char* obj_path, double turn_deg, uint32_t** ptsd,
              size_t* n_ptsd, uint32_t** ptgp, size_t* n_ptgp) {
    try {
        const geometry g = reference_scene(id, crowd_side, flatten, obj_path, turn_deg);
        float from[3], at[3], vfov;
        camera_params(static_cast<scene_id>(id), from, at, &vfov);
        *ptsd = copy_out(to_ptsd(g, from, at, vfov), n_ptsd);
        *ptgp = copy_out(serialize(pack_for_gpu(g, make_camera(static_cast<scene_id>(id)))), n_ptgp);
        return 0;
    } catch (const std::exception& e) {
        last_error = e.what();
        return 1;
    }
}

// Ray queries. rays: n x 8 floats (o, tmin, d, tmax). Outputs per ray: kind
// (0 miss, 1 large sphere, 2 sphere, 3 flat triangle, 4 instance), index
// (large-sphere, sphere, triangle or instance number), sub (the instance's
// mesh triangle), t, u, v; for any-hit queries only kind (0 or 1).
// Also checks the restated closest hit against geometry::hit's own t and the
// any hit against geometry::occluded; returns the number of disagreements.
long ref_world_trace(void* w, const float* rays, size_t n, int any_hit, uint32_t* kind, uint32_t* index, uint32_t* sub,
                     float* t, float* u, float* v, int threads) {
    auto* rw = static_cast<ref_world*>(w);
    const geometry& g = rw->world;
    std::atomic<long> bad{0};
    std::atomic<size_t> next{0};
    auto worker = [&] {
        for (size_t b = next.fetch_add(1024); b < n; b = next.fetch_add(1024)) {
            for (size_t i = b; i < std::min(n, b + 1024); ++i) {
                const float* q = rays + 8 * i;
                const ray r(point3(q[0], q[1], q[2]), vec3(q[4], q[5], q[6]));
                if (any_hit) {
                    const bool occ = g.occluded(r, q[3], q[7]);
                    kind[i] = occ ? 1 : 0;
                    index[i] = sub[i] = 0;
                    t[i] = u[i] = v[i] = 0;
                    continue;
                }
                const rhit h = ref_closest(g, r, q[3], q[7]);
                hit_record rec;
                const bool found = g.hit(r, q[3], q[7], rec);
                if (found != (h.kind != 0) || (found && rec.t != h.t)) ++bad;
                kind[i] = h.kind;
                index[i] = h.index;
                sub[i] = h.sub;
                t[i] = h.t;
                u[i] = h.u;
                v[i] = h.v;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int k = 0; k < std::max(1, threads); ++k) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    return bad.load();
}

// render_row_linear over rows [j_lo, j_hi), scheduled as main.cpp:127-170:
// one std::thread per worker claiming rows from an atomic counter, the
// per-thread ray counters summed at the end, and the time taken from thread
// start to join. rgb: 3 floats a pixel, the band's top row first.
int ref_world_render(void* w, int width, int height, int spp, int max_bounces, uint64_t seed, int first_sample,
                     int nee, int threads, int j_lo, int j_hi, float* rgb, float* aux, uint64_t* rays, double* seconds) {
    auto* rw = static_cast<ref_world*>(w);
    geometry& g = rw->world;
    const bool saved = g.sample_lights;
    if (!nee) g.sample_lights = false;
    std::atomic<int> next_row{j_lo};
    std::atomic<long long> total{0};
    auto worker = [&] {
        for (int j = next_row.fetch_add(1); j < j_hi; j = next_row.fetch_add(1)) {
            const size_t row = static_cast<size_t>(j_hi - 1 - j) * width;
            render_row_linear(g, rw->cam, j, width, height, spp, max_bounces, rgb + 3 * row, seed, first_sample,
                              aux ? aux + aux_floats * row : nullptr);
        }
        total += rays_traced;
    };
    const auto start = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int k = 0; k < std::max(1, threads); ++k) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    *rays = static_cast<uint64_t>(total.load());
    g.sample_lights = saved;
    return 0;
}

// The reference sampler: begin_sample then sample_2d of pair `pair` (pair 0,
// or 1 + 2k + offset through set_bounce(k)); returns 1 where it fell back to
// the random stream.
int ref_sample_2d(uint64_t seed, int i, int j, uint32_t index, int pair, float* a, float* b) {
    begin_sample(seed, i, j, index);
    if (pair == 0) {
        sample_2d(0, *a, *b);
    } else {
        const int k = (pair - 1) / 2, off = (pair - 1) % 2;
        set_bounce(k);
        sample_2d(off, *a, *b);
    }
    return pair > 2 * qmc_bounces ? 1 : 0;
}

// The reference's image_texture (texture.hpp:13-33) over w x h sRGB bytes,
// sampled at n (u, v) pairs: 3 floats each.
void ref_texture_sample(const uint8_t* rgb, int w, int h, const float* uv, size_t n, float* out) {
    image8 img;
    img.width = w;
    img.height = h;
    img.rgb.assign(rgb, rgb + 3 * static_cast<size_t>(w) * h);
    const image_texture tex(img);
    for (size_t i = 0; i < n; ++i) {
        const color c = tex.sample(uv[2 * i], uv[2 * i + 1]);
        out[3 * i] = c.x();
        out[3 * i + 1] = c.y();
        out[3 * i + 2] = c.z();
    }
}

// decode_image (image_io.hpp:295) of a file's bytes: 0 and the image (rgb
// allocated with malloc, free with ref_free), or 1 and the error message.
int ref_decode_image(const uint8_t* bytes, size_t n, int* w, int* h, uint8_t** rgb) {
    try {
        const image8 img = decode_image(std::string(reinterpret_cast<const char*>(bytes), n));
        *w = img.width;
        *h = img.height;
        *rgb = static_cast<uint8_t*>(std::malloc(std::max<size_t>(1, img.rgb.size())));
        if (!img.rgb.empty()) std::memcpy(*rgb, img.rgb.data(), img.rgb.size());
        return 0;
    } catch (const std::exception& e) {
        last_error = e.what();
        return 1;
    }
}

// std::thread::hardware_concurrency(), as main.cpp:116-118 uses.
int ref_hardware_threads() { return static_cast<int>(std::thread::hardware_concurrency()); }

// The CLI's --denoise step (main.cpp:172-186): albedo and normal from the
// render's aux floats, each pixel's variance of the mean from the mean
// squared luminance, then the reference's filter (denoise.hpp) on `threads`
// threads. rgb: 3 floats a pixel, aux: aux_floats a pixel, out: 3 a pixel.
int ref_denoise(int width, int height, int spp, const float* rgb, const float* aux, float* out, int threads,
                double* seconds) {
    const auto start = std::chrono::steady_clock::now();
    const size_t n = static_cast<size_t>(width) * height;
    std::vector<float> albedo(3 * n), normal(3 * n), variance(n);
    for (size_t p = 0; p < n; ++p) {
        const float* a = &aux[aux_floats * p];
        for (int c = 0; c < 3; ++c) {
            albedo[3 * p + c] = a[c];
            normal[3 * p + c] = a[3 + c];
        }
        const float mean = luminance(color(rgb[3 * p], rgb[3 * p + 1], rgb[3 * p + 2]));
        variance[p] = (a[6] - mean * mean) / spp;
    }
    denoise({width, height, rgb, albedo.data(), normal.data(), variance.data()}, out, threads);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    return 0;
}

}  // extern "C"
