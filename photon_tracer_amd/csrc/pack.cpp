// PTSD read/write, the reference camera, instance transforms, and packing a
// scene into the reference's PTGP v1 buffer.
//
// Packing follows geometry::build (geometry.hpp:74-104) for what is built,
// and pack_for_gpu / serialize (gpu_pack.hpp:84-203) for the layout: the
// result is byte-identical to the reference's for any scene the reference's
// packer accepts, and it also accepts trees deeper than its 10-level cap.
#include "host.hpp"

#include <algorithm>
#include <thread>
#include <unordered_map>

namespace pt {

namespace {

inline float fbits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint32_t ubits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

struct reader {
    const uint32_t* w;
    size_t n, at;
    uint32_t u() {
        if (at >= n) throw std::runtime_error("PTSD buffer is truncated");
        return w[at++];
    }
    float f() { return fbits(u()); }
    v3 vec() {
        const float a = f(), b = f(), c = f();
        return {a, b, c};
    }
};

void put(std::vector<uint32_t>& o, float f) { o.push_back(ubits(f)); }
void put(std::vector<uint32_t>& o, v3 v) {
    put(o, v.x);
    put(o, v.y);
    put(o, v.z);
}

void write_triangle(std::vector<uint32_t>& o, const triangle_rec& t) {
    for (int k = 0; k < 3; ++k) put(o, t.p[k]);
    for (int k = 0; k < 3; ++k) put(o, t.n[k]);
    for (int k = 0; k < 6; ++k) put(o, t.uv[k]);
    o.push_back(t.material);
    o.push_back(0);
    o.push_back(0);
    o.push_back(0);
}

triangle_rec read_triangle(reader& r) {
    triangle_rec t;
    for (int k = 0; k < 3; ++k) t.p[k] = r.vec();
    for (int k = 0; k < 3; ++k) t.n[k] = r.vec();
    for (int k = 0; k < 6; ++k) t.uv[k] = r.f();
    t.material = r.u();
    r.u();
    r.u();
    r.u();
    return t;
}

}  // namespace

std::vector<uint32_t> ptsd_write(const scene_desc& s) {
    std::vector<uint32_t> o(kPtsdHeader, 0);
    o[0] = kPtsdMagic;
    o[1] = 1;
    o[2] = (s.sky ? 1u : 0u) | (s.sample_lights ? 2u : 0u);
    o[3] = ubits(s.lookfrom.x);
    o[4] = ubits(s.lookfrom.y);
    o[5] = ubits(s.lookfrom.z);
    o[6] = ubits(s.lookat.x);
    o[7] = ubits(s.lookat.y);
    o[8] = ubits(s.lookat.z);
    o[9] = ubits(s.vfov);
    o[10] = static_cast<uint32_t>(s.materials.size());
    o[11] = static_cast<uint32_t>(s.spheres.size());
    o[12] = static_cast<uint32_t>(s.large.size());
    o[13] = static_cast<uint32_t>(s.triangles.size());
    o[14] = static_cast<uint32_t>(s.meshes.size());
    o[15] = static_cast<uint32_t>(s.instances.size());
    size_t mesh_tris = 0;
    for (const auto& m : s.meshes) mesh_tris += m.size();
    o[16] = static_cast<uint32_t>(mesh_tris);
    o[17] = static_cast<uint32_t>(s.textures.size());
    for (const auto& m : s.materials) {
        put(o, m.albedo);
        o.push_back(m.type);
        put(o, m.emission);
        put(o, m.fuzz);
        put(o, m.ior);
        o.push_back(m.texture);
        o.push_back(0);
        o.push_back(0);
    }
    for (const auto* list : {&s.spheres, &s.large})
        for (const auto& sp : *list) {
            put(o, sp.center);
            put(o, sp.radius);
            o.push_back(sp.material);
        }
    for (const auto& t : s.triangles) write_triangle(o, t);
    for (const auto& m : s.meshes) o.push_back(static_cast<uint32_t>(m.size()));
    for (const auto& m : s.meshes)
        for (const auto& t : m) write_triangle(o, t);
    for (const auto& in : s.instances) {
        o.push_back(in.mesh);
        o.push_back(in.material);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) put(o, in.to_world.m[r][c]);
        put(o, in.to_world.t);
        o.push_back(0);
        o.push_back(0);
    }
    for (const auto& im : s.textures) {
        o.push_back(im.width);
        o.push_back(im.height);
        const size_t at = o.size();
        o.resize(at + (im.rgb.size() + 3) / 4, 0);
        if (!im.rgb.empty()) std::memcpy(&o[at], im.rgb.data(), im.rgb.size());
    }
    return o;
}

scene_desc ptsd_read(const uint32_t* w, size_t n) {
    if (!w || n < static_cast<size_t>(kPtsdHeader)) throw std::runtime_error("PTSD buffer is too short");
    if (w[0] != kPtsdMagic) throw std::runtime_error("not a PTSD buffer (bad magic)");
    if (w[1] != 1) throw std::runtime_error("unsupported PTSD version");
    scene_desc s;
    s.sky = (w[2] & 1u) != 0;
    s.sample_lights = (w[2] & 2u) != 0;
    s.lookfrom = {fbits(w[3]), fbits(w[4]), fbits(w[5])};
    s.lookat = {fbits(w[6]), fbits(w[7]), fbits(w[8])};
    s.vfov = fbits(w[9]);
    const uint32_t n_mat = w[10], n_sph = w[11], n_large = w[12], n_tri = w[13], n_mesh = w[14], n_inst = w[15];
    const uint32_t n_tex = w[17];
    const size_t fixed = kPtsdHeader + size_t(n_mat) * kMatWords + size_t(n_sph + n_large) * kSphWords +
                         size_t(n_tri) * kTriWords + n_mesh + size_t(w[16]) * kTriWords + size_t(n_inst) * kInstWords;
    // The textures' sizes come from their own headers.
    size_t need = fixed;
    for (uint32_t k = 0; k < n_tex; ++k) {
        if (need + 2 > n) throw std::runtime_error("PTSD buffer size does not match its header");
        if (w[need] == 0 || w[need + 1] == 0 || w[need] > 65536 || w[need + 1] > 65536)
            throw std::runtime_error("PTSD texture size out of range (1 to 65536 a side)");
        const uint64_t texels = uint64_t(w[need]) * w[need + 1];
        need += 2 + size_t((texels * 3 + 3) / 4);
    }
    if (need != n) throw std::runtime_error("PTSD buffer size does not match its header");
    reader r{w, n, kPtsdHeader};
    s.materials.resize(n_mat);
    for (auto& m : s.materials) {
        m.albedo = r.vec();
        m.type = r.u();
        m.emission = r.vec();
        m.fuzz = r.f();
        m.ior = r.f();
        m.texture = r.u();
        r.u();
        r.u();
        if (m.type > 3) throw std::runtime_error("PTSD material type out of range");
        if (m.texture > n_tex || (m.texture && m.type != 0))
            throw std::runtime_error("PTSD material texture out of range or not on a matte material");
    }
    auto read_sphere = [&](sphere_rec& sp) {
        sp.center = r.vec();
        sp.radius = r.f();
        sp.material = r.u();
        if (sp.material >= n_mat) throw std::runtime_error("PTSD sphere material out of range");
    };
    s.spheres.resize(n_sph);
    for (auto& sp : s.spheres) read_sphere(sp);
    s.large.resize(n_large);
    for (auto& sp : s.large) read_sphere(sp);
    s.triangles.resize(n_tri);
    for (auto& t : s.triangles) {
        t = read_triangle(r);
        if (t.material >= n_mat) throw std::runtime_error("PTSD triangle material out of range");
    }
    std::vector<uint32_t> sizes(n_mesh);
    size_t total = 0;
    for (auto& c : sizes) {
        c = r.u();
        total += c;
    }
    if (total != w[16]) throw std::runtime_error("PTSD mesh sizes do not add up");
    s.meshes.resize(n_mesh);
    for (uint32_t k = 0; k < n_mesh; ++k) {
        s.meshes[k].resize(sizes[k]);
        for (auto& t : s.meshes[k]) {
            t = read_triangle(r);
            if (t.material >= n_mat) throw std::runtime_error("PTSD mesh triangle material out of range");
        }
    }
    s.instances.resize(n_inst);
    for (auto& in : s.instances) {
        in.mesh = r.u();
        in.material = r.u();
        for (int rr = 0; rr < 3; ++rr)
            for (int c = 0; c < 3; ++c) in.to_world.m[rr][c] = r.f();
        in.to_world.t = r.vec();
        r.u();
        r.u();
        if (in.mesh >= n_mesh) throw std::runtime_error("PTSD instance mesh out of range");
        if (in.material != ~0u && in.material >= n_mat)
            throw std::runtime_error("PTSD instance material out of range");
    }
    s.textures.resize(n_tex);
    for (auto& im : s.textures) {
        im.width = r.u();
        im.height = r.u();
        im.rgb.resize(size_t(im.width) * im.height * 3);
        std::memcpy(im.rgb.data(), w + r.at, im.rgb.size());
        r.at += (im.rgb.size() + 3) / 4;
    }
    return s;
}

void camera_rows(v3 lookfrom, v3 lookat, double vfov_deg, float out[16]) {
    // camera.hpp:18-34: viewport in double, the basis in float, and the
    // double sizes converted to float where they scale a vec3.
    const double aspect_ratio = 16.0 / 9.0;
    const double viewport_height = 2.0 * std::tan(vfov_deg * M_PI / 360.0);
    const double viewport_width = aspect_ratio * viewport_height;
    const v3 w = normalize(lookfrom - lookat);
    const v3 u = normalize(cross(v3(0, 1, 0), w));
    const v3 v = cross(w, u);
    const v3 origin = lookfrom;
    const v3 horizontal = u * static_cast<float>(viewport_width);
    const v3 vertical = v * static_cast<float>(viewport_height);
    const v3 corner = origin - (horizontal * 0.5f) - (vertical * 0.5f) - w;
    const v3 rows[4] = {origin, corner, horizontal, vertical};
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 3; ++c) out[4 * r + c] = rows[r][c];
        out[4 * r + 3] = 0;
    }
}

affine_rec affine_rec::inverse() const {
    double a[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = m[i][j];
    const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    const double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    const double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    const double inv[3][3] = {
        {c00 / det, (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det, (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det},
        {c01 / det, (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det, (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det},
        {c02 / det, (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det, (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det}};
    affine_rec r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = static_cast<float>(inv[i][j]);
    r.t = -r.vector(t);
    return r;
}

box3 affine_rec::box(const box3& b) const {
    box3 out;
    for (int k = 0; k < 8; ++k) {
        const v3 c(k & 1 ? b.hi.x : b.lo.x, k & 2 ? b.hi.y : b.lo.y, k & 4 ? b.hi.z : b.lo.z);
        const v3 p = point(c);
        box3 pb;
        pb.lo = pb.hi = p;
        out.grow(pb);
    }
    return out;
}

namespace {

// Appends a tree with rebased links (gpu_pack.hpp:58-73), without the
// reference's depth cap.
template <class Map>
uint32_t append_tree(std::vector<bvh_node>& nodes, std::vector<uint32_t>& order, const bvh_tree& tree, Map&& map_leaf) {
    if (tree.nodes.empty()) return ~0u;
    const uint32_t node_base = static_cast<uint32_t>(nodes.size()), order_base = static_cast<uint32_t>(order.size());
    const float inf = std::numeric_limits<float>::infinity();
    for (bvh_node n : tree.nodes) {
        for (int c = 0; c < 4; ++c) {
            if (n.count[c] > 0) n.first[c] += order_base;
            else if (n.lo[0][c] != inf) n.first[c] += node_base;
        }
        nodes.push_back(n);
    }
    for (uint32_t p : tree.order) order.push_back(map_leaf(p));
    return node_base;
}

}  // namespace

std::vector<uint32_t> pack_ptgp(const scene_desc& s, const bvh_builder* builder) {
    const auto build = [&](const std::vector<box3>& boxes) { return builder ? (*builder)(boxes) : build_bvh(boxes); };
    // geometry::build (geometry.hpp:74-104): lights, mesh trees, flat tree, instance tree.
    std::vector<uint32_t> lights;
    for (uint32_t i = 0; i < s.spheres.size(); ++i)
        if (s.materials[s.spheres[i].material].is_light()) lights.push_back(i);

    std::vector<bvh_tree> mesh_trees(s.meshes.size());
    std::vector<box3> mesh_bounds(s.meshes.size());
    {
        // Meshes are independent: build them side by side on host threads,
        // or one after another with a given builder.
        std::vector<std::thread> pool;
        for (size_t k = 0; k < s.meshes.size(); ++k) {
            const auto one = [&, k] {
                std::vector<box3> boxes;
                boxes.reserve(s.meshes[k].size());
                box3 bounds;
                for (const auto& t : s.meshes[k]) {
                    boxes.push_back(triangle_box(t));
                    bounds.grow(boxes.back());
                }
                mesh_bounds[k] = bounds;
                mesh_trees[k] = build(boxes);
            };
            if (builder) {
                one();
                continue;
            }
            pool.emplace_back(one);
            if (pool.size() >= 8) {
                for (auto& t : pool) t.join();
                pool.clear();
            }
        }
        for (auto& t : pool) t.join();
    }
    bvh_tree flat, inst_tree;
    {
        std::vector<box3> boxes;
        boxes.reserve(s.spheres.size() + s.triangles.size());
        for (const auto& sp : s.spheres) boxes.push_back(sphere_box(sp));
        for (const auto& t : s.triangles) boxes.push_back(triangle_box(t));
        flat = build(boxes);
        boxes.clear();
        for (const auto& in : s.instances) boxes.push_back(in.to_world.box(mesh_bounds[in.mesh]));
        inst_tree = build(boxes);
    }

    // pack_for_gpu (gpu_pack.hpp:84-175)
    std::vector<bvh_node> nodes;
    std::vector<uint32_t> order;
    std::vector<float> spheres, triangles, instances, materials;
    std::vector<uint32_t> large;
    std::unordered_map<uint32_t, uint32_t> index;
    auto material_of = [&](uint32_t id) -> uint32_t {
        auto found = index.find(id);
        if (found != index.end()) return found->second;
        const material_rec& m = s.materials[id];
        // Only what each kind carries: gpu_pack.hpp:93-106.
        float type = 0, fuzz = 0, ior = 1.5f;
        v3 albedo(0, 0, 0);
        if (m.type == 0) {
            albedo = m.albedo;
        } else if (m.type == 1) {
            type = 1;
            albedo = m.albedo;
            fuzz = m.fuzz < 1 ? m.fuzz : 1;  // metal's constructor clamps (material.hpp:81)
        } else if (m.type == 2) {
            type = 2;
            ior = m.ior;
        } else {
            type = 3;
        }
        const v3 e = m.type == 3 ? m.emission : v3(0, 0, 0);
        // Word 9: a matte material's texture (number + 1, 0 none): this
        // library's extension of PTGP v1, where the reference writes 0.
        const float tex = fbits(m.type == 0 ? m.texture : 0u);
        const float row[12] = {albedo.x, albedo.y, albedo.z, type, e.x, e.y, e.z, fuzz, ior, tex, 0, 0};
        materials.insert(materials.end(), row, row + 12);
        const uint32_t k = static_cast<uint32_t>(index.size());
        index.emplace(id, k);
        return k;
    };
    auto pack_sphere = [&](const sphere_rec& sp) {
        const float row[8] = {sp.center.x, sp.center.y, sp.center.z, sp.radius, fbits(material_of(sp.material)), 0, 0, 0};
        spheres.insert(spheres.end(), row, row + 8);
    };
    auto pack_triangle = [&](const triangle_rec& t) {
        // triangle.hpp:97-100: edges from p0, face normal from their cross product.
        const v3 e1 = t.p[1] - t.p[0], e2 = t.p[2] - t.p[0];
        const v3 fn = normalize(cross(e1, e2));
        const float row[28] = {t.p[0].x, t.p[0].y, t.p[0].z, fbits(material_of(t.material)),
                               e1.x,     e1.y,     e1.z,     t.uv[0],
                               e2.x,     e2.y,     e2.z,     t.uv[1],
                               t.n[0].x, t.n[0].y, t.n[0].z, t.uv[2],
                               t.n[1].x, t.n[1].y, t.n[1].z, t.uv[3],
                               t.n[2].x, t.n[2].y, t.n[2].z, t.uv[4],
                               fn.x,     fn.y,     fn.z,     t.uv[5]};
        triangles.insert(triangles.end(), row, row + 28);
    };

    for (const auto& sp : s.spheres) pack_sphere(sp);
    const uint32_t small_spheres = static_cast<uint32_t>(s.spheres.size());
    for (const auto& sp : s.large) {
        large.push_back(static_cast<uint32_t>(spheres.size() / 8));
        pack_sphere(sp);
    }
    for (const auto& t : s.triangles) pack_triangle(t);
    const uint32_t flat_root = append_tree(nodes, order, flat, [](uint32_t p) { return p; });

    std::vector<uint32_t> mesh_root;
    for (size_t k = 0; k < s.meshes.size(); ++k) {
        const uint32_t tri_base = static_cast<uint32_t>(triangles.size() / 28);
        for (const auto& t : s.meshes[k]) pack_triangle(t);
        mesh_root.push_back(append_tree(nodes, order, mesh_trees[k], [tri_base](uint32_t q) { return tri_base + q; }));
    }
    for (const auto& in : s.instances) {
        const affine_rec to_object = in.to_world.inverse();
        float row[28] = {};
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                row[4 * r + c] = in.to_world.m[r][c];
                row[12 + 4 * r + c] = to_object.m[r][c];
            }
            row[4 * r + 3] = in.to_world.t[r];
            row[12 + 4 * r + 3] = to_object.t[r];
        }
        row[24] = fbits(mesh_root[in.mesh]);
        row[25] = fbits(in.material != ~0u ? material_of(in.material) : ~0u);
        instances.insert(instances.end(), row, row + 28);
    }
    const uint32_t instance_root = append_tree(nodes, order, inst_tree, [](uint32_t q) { return q; });

    float camera[16];
    camera_rows(s.lookfrom, s.lookat, s.vfov, camera);

    // serialize (gpu_pack.hpp:177-203)
    std::vector<uint32_t> out(48, 0);
    out[0] = kPtgpMagic;
    out[1] = 1;
    auto add = [&](int slot, const void* data, size_t bytes) {
        if (4 * out.size() > 0xffffffffull || bytes > 0xffffffffull)
            throw std::runtime_error("scene too large for the PTGP v1 32-bit offsets");
        out[2 + 2 * slot] = static_cast<uint32_t>(4 * out.size());
        out[3 + 2 * slot] = static_cast<uint32_t>(bytes);
        const size_t words = (bytes + 15) / 16 * 4;
        const size_t at = out.size();
        out.resize(at + words, 0);
        if (bytes) std::memcpy(&out[at], data, bytes);
    };
    add(0, nodes.data(), nodes.size() * sizeof(bvh_node));
    add(1, order.data(), order.size() * 4);
    add(2, spheres.data(), spheres.size() * 4);
    add(3, triangles.data(), triangles.size() * 4);
    add(4, instances.data(), instances.size() * 4);
    add(5, materials.data(), materials.size() * 4);
    add(6, lights.data(), lights.size() * 4);
    add(7, large.data(), large.size() * 4);
    out[18] = flat_root;
    out[19] = instance_root;
    out[20] = small_spheres;
    out[21] = (s.sky ? 1u : 0u) | (s.sample_lights ? 2u : 0u);
    std::memcpy(&out[22], camera, sizeof camera);
    if (!s.textures.empty()) {
        // The texture section (header words 38-40: byte offset, length and
        // number of textures; 0 in the reference's PTGP, which drops
        // textures, gpu_pack.hpp:96):
        // per texture (width, height, first texel, 0), then every texture's
        // texels as linear RGB float4 (a = 0), rows top first. The sRGB bytes
        // are linearised as image_texture's table does (texture.hpp:13-20).
        float linear[256];
        for (int v = 0; v < 256; ++v) {
            const float c = v / 255.0f;
            linear[v] = c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f);
        }
        std::vector<uint32_t> sec;
        uint64_t first = 0;
        for (const auto& im : s.textures) {
            sec.insert(sec.end(), {im.width, im.height, static_cast<uint32_t>(first), 0u});
            first += uint64_t(im.width) * im.height;
            if (first >= (uint64_t(1) << 32)) throw std::runtime_error("textures too large for PTGP");
        }
        sec.reserve(sec.size() + 4 * first);
        for (const auto& im : s.textures)
            for (size_t k = 0; k < size_t(im.width) * im.height; ++k) {
                for (int c = 0; c < 3; ++c) sec.push_back(ubits(linear[im.rgb[3 * k + c]]));
                sec.push_back(0);
            }
        while (out.size() % 4) out.push_back(0);   // float4 texels: the section starts 16-byte aligned
        const size_t at = out.size();
        if (4 * (at + sec.size()) > 0xffffffffull) throw std::runtime_error("scene too large for the PTGP 32-bit offsets");
        out[38] = static_cast<uint32_t>(4 * at);
        out[39] = static_cast<uint32_t>(4 * sec.size());
        out[40] = static_cast<uint32_t>(s.textures.size());
        out.insert(out.end(), sec.begin(), sec.end());
    }
    return out;
}

}  // namespace pt
