// OBJ text to triangles (load_obj, obj.hpp:114-218), for scenes that the
// packer and the device path take as PTSD. Instead of a file reader, the
// caller passes the text of one MTL library, which stands in for every
// library the OBJ names (none: every face takes the given material, as the
// browser demo loads). MTL materials map onto the four kinds as
// make_material does (obj.hpp:89-112); texture maps are not carried (PTGP
// has none, gpu_pack.hpp:96), so a textured matte keeps its Kd.
//
// The arithmetic follows the reference's types step by step, so the
// triangles come out bit for bit: coordinates read as double and stored as
// float; the turn about the vertical axis in double; the height scale
// computed in double and applied in float (vec3 * real); face normals and
// their per-vertex sums in float, in face order; float normalisation.
#include <array>
#include <cmath>
#include <map>
#include <sstream>

#include "host.hpp"

namespace pt {
namespace {

struct corner {
    int v, vt, vn;   // -1 where the file gave none
};

v3 cross3(const v3& a, const v3& b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
v3 sub3(const v3& a, const v3& b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
v3 add3(const v3& a, const v3& b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
float dot3(const v3& a, const v3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
v3 unit3(const v3& a) {   // vec3::normalize (vec3.hpp:95-98)
    const float length = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3(a.x / length, a.y / length, a.z / length);
}

// A face corner's index: 1-based, or negative counting back from the end of
// its list so far; -1 when the field is empty. Out of range is an error
// here (the reference leaves it undefined).
int corner_index(const std::string& s, size_t count) {
    if (s.empty()) return -1;
    const int i = std::stoi(s);
    const long k = i > 0 ? long(i) - 1 : long(count) + i;
    if (k < 0 || k >= long(count)) throw std::runtime_error("OBJ face index out of range: " + s);
    return static_cast<int>(k);
}

struct mtl_rec {
    double kd[3] = {0.8, 0.8, 0.8}, ks[3] = {0, 0, 0}, ke[3] = {0, 0, 0};
    double ns = 0, ni = 1.5, d = 1;
    int illum = 2;
};

// read_mtl (obj.hpp:49-87), less map_Kd.
void read_mtl(const std::string& text, std::map<std::string, mtl_rec>& out) {
    std::istringstream in(text);
    std::string line;
    mtl_rec* cur = nullptr;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "newmtl") {
            std::string name;
            ls >> name;
            cur = &out[name];
            *cur = mtl_rec();
        } else if (!cur) {
            continue;
        } else if (tag == "Kd" || tag == "Ks" || tag == "Ke") {
            double r, g, b;
            if (!(ls >> r)) continue;
            if (!(ls >> g >> b)) g = b = r;   // one value means gray
            double* c = tag == "Kd" ? cur->kd : tag == "Ks" ? cur->ks : cur->ke;
            c[0] = r;
            c[1] = g;
            c[2] = b;
        } else if (tag == "Ns") {
            ls >> cur->ns;
        } else if (tag == "Ni") {
            ls >> cur->ni;
        } else if (tag == "d") {
            ls >> cur->d;
        } else if (tag == "Tr") {
            double tr;
            if (ls >> tr) cur->d = 1 - tr;
        } else if (tag == "illum") {
            ls >> cur->illum;
        }
    }
}

// make_material (obj.hpp:89-112): emission makes a light; dissolve below 1
// or a glass model (4, 6, 7, 9) glass; a reflection model (3, 5) metal with
// a fuzz from the shininess; anything else matte. Appended to s.materials.
uint32_t make_material(scene_desc& s, const mtl_rec& m) {
    auto nonzero = [](const double* c) { return c[0] > 0 || c[1] > 0 || c[2] > 0; };
    auto color = [](const double* c) {
        return v3(static_cast<float>(c[0]), static_cast<float>(c[1]), static_cast<float>(c[2]));
    };
    material_rec r;
    if (nonzero(m.ke)) {
        r.type = 3;
        r.emission = color(m.ke);
    } else if (m.d < 1 || m.illum == 4 || m.illum == 6 || m.illum == 7 || m.illum == 9) {
        r.type = 2;
        r.ior = static_cast<float>(m.ni > 1 ? m.ni : 1.5);
    } else if (m.illum == 3 || m.illum == 5) {
        r.type = 1;
        r.albedo = color(nonzero(m.ks) ? m.ks : m.kd);
        const float f = static_cast<float>(std::min(1.0, std::sqrt(2 / (m.ns + 2))));
        r.fuzz = f < 1 ? f : 1;   // metal's constructor clamp (material.hpp:83)
    } else {
        r.type = 0;
        r.albedo = color(m.kd);
    }
    s.materials.push_back(r);
    return static_cast<uint32_t>(s.materials.size() - 1);
}

}  // namespace

void add_obj(scene_desc& s, const std::string& text, v3 base, double height, uint32_t material, double turn_deg,
             const std::string* mtl_text) {
    if (material >= s.materials.size()) throw std::runtime_error("OBJ material index out of range");
    std::map<std::string, mtl_rec> library;
    std::map<std::string, uint32_t> made;
    uint32_t current = material;
    std::vector<v3> verts, file_normals;
    std::vector<std::array<float, 2>> uvs;
    struct face {
        std::array<corner, 3> c;
        uint32_t material;
    };
    std::vector<face> faces;

    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "v" || tag == "vn") {
            double x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            (tag == "v" ? verts : file_normals)
                .emplace_back(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z));
        } else if (tag == "vt") {
            double u = 0, v = 0;
            ls >> u >> v;
            uvs.push_back({static_cast<float>(u), static_cast<float>(v)});
        } else if (tag == "mtllib" && mtl_text) {
            std::string name;
            while (ls >> name) read_mtl(*mtl_text, library);
        } else if (tag == "usemtl") {
            std::string name;
            ls >> name;
            const auto found = library.find(name);
            if (found == library.end()) {
                current = material;
            } else {
                auto m = made.find(name);
                if (m == made.end()) m = made.emplace(name, make_material(s, found->second)).first;
                current = m->second;
            }
        } else if (tag == "f") {
            std::vector<corner> poly;
            std::string tok;
            while (ls >> tok) {
                const size_t s1 = tok.find('/'), s2 = s1 == std::string::npos ? s1 : tok.find('/', s1 + 1);
                corner c;
                c.v = corner_index(tok.substr(0, s1), verts.size());
                if (c.v < 0) throw std::runtime_error("OBJ face corner without a vertex");
                c.vt = s1 == std::string::npos ? -1 : corner_index(tok.substr(s1 + 1, s2 - s1 - 1), uvs.size());
                c.vn = s2 == std::string::npos ? -1 : corner_index(tok.substr(s2 + 1), file_normals.size());
                poly.push_back(c);
            }
            for (size_t k = 2; k < poly.size(); ++k) faces.push_back({{poly[0], poly[k - 1], poly[k]}, current});
        }
    }
    if (verts.empty() || faces.empty()) return;

    const double turn = turn_deg * M_PI / 180, c_turn = std::cos(turn), s_turn = std::sin(turn);
    auto rotate = [&](v3& p) {
        p = v3(static_cast<float>(c_turn * p.x + s_turn * p.z), p.y, static_cast<float>(-s_turn * p.x + c_turn * p.z));
    };
    for (auto& v : verts) rotate(v);
    for (auto& n : file_normals) rotate(n);

    v3 lo = verts[0], hi = verts[0];
    for (const auto& v : verts) {
        lo = v3(std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z));
        hi = v3(std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z));
    }
    const double scale = height / (hi.y - lo.y);
    const float fscale = static_cast<float>(scale);
    const v3 bottom(static_cast<float>((lo.x + hi.x) * 0.5), lo.y, static_cast<float>((lo.z + hi.z) * 0.5));
    for (auto& v : verts) {
        const v3 d = sub3(v, bottom);
        v = add3(base, v3(d.x * fscale, d.y * fscale, d.z * fscale));
    }

    // Area-weighted vertex normals: the cross product's length is twice the
    // face's area.
    std::vector<v3> smooth(verts.size());
    for (const auto& face : faces) {
        const auto& f = face.c;
        const v3 n = cross3(sub3(verts[f[1].v], verts[f[0].v]), sub3(verts[f[2].v], verts[f[0].v]));
        for (const corner& c : f) smooth[c.v] = add3(smooth[c.v], n);
    }

    s.triangles.reserve(s.triangles.size() + faces.size());
    for (const auto& face : faces) {
        const auto& f = face.c;
        const v3 &a = verts[f[0].v], &b = verts[f[1].v], &c = verts[f[2].v];
        const v3 fn = cross3(sub3(b, a), sub3(c, a));
        if (dot3(fn, fn) < 1e-24) continue;   // degenerate
        const bool from_file = f[0].vn >= 0 && f[1].vn >= 0 && f[2].vn >= 0;
        triangle_rec t;
        t.p[0] = a;
        t.p[1] = b;
        t.p[2] = c;
        for (int k = 0; k < 3; ++k) {
            t.n[k] = unit3(from_file ? file_normals[f[k].vn] : smooth[f[k].v]);
            if (f[k].vt >= 0) {
                t.uv[2 * k] = uvs[f[k].vt][0];
                t.uv[2 * k + 1] = uvs[f[k].vt][1];
            }
        }
        t.material = face.material;
        s.triangles.push_back(t);
    }
}

}  // namespace pt
