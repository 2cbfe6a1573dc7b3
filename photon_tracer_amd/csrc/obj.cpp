// Wavefront OBJ/MTL import into a PTSD scene (the reference's load_obj,
// obj.hpp:114-218, and its material mapping, obj.hpp:38-112).
//
// The import runs in four passes over data of its own:
//   scan     one pass over the text with a cursor that splits words and
//            reads numbers in place: positions, texture coordinates, file
//            normals, and a face table (three corners a triangle, fans for
//            polygons) with each face's material slot; MTL text is scanned
//            the same way into surface specifications;
//   place    the turn about the vertical axis, the fit to the height and the
//            set-down on the base point;
//   normals  area-weighted vertex normals for faces without file normals;
//   emit     triangles (degenerate ones dropped) appended to the scene.
//
// What must match the reference bit for bit is the arithmetic of `place`,
// `normals` and `emit`, because the triangles are compared word for word
// with the reference's (tests/test_host.py, teapot and Spot): coordinates
// parsed as double and stored as float; the turn computed in double from
// float coordinates; the bounding box and the height difference in float,
// the scale in double and applied as a float; face normals as float cross
// products summed per vertex in face order; float normalisation
// (vec3.hpp:95-98); the 1e-24 degenerate test on the float cross product.
// The reading of the text is this file's own: numbers are read with
// std::from_chars (locale independent), a field that is not a number reads
// as 0 for positions and the MTL's scalar fields (as a failed stream
// extraction stores 0), and out-of-range face indices are errors where the
// reference's are undefined.
//
// Texture maps: an MTL material's map_Kd names an image; images the caller
// passes are matched by that name (the map_Kd line's last word) and become
// PTSD textures of matte materials, as the reference's lambertian takes
// `Kd x map_Kd` (obj.hpp:103-111). Without a matching image the material
// keeps its Kd alone.
#include <array>
#include <charconv>
#include <cmath>
#include <string_view>
#include <unordered_map>

#include "host.hpp"

namespace pt {
namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Words and numbers of one line, read in place.
class line_cursor {
  public:
    explicit line_cursor(std::string_view line) : p_(line.data()), end_(line.data() + line.size()) {}

    // The next whitespace-delimited word ("" at the end of the line).
    std::string_view word() {
        while (p_ < end_ && is_space(*p_)) ++p_;
        const char* start = p_;
        while (p_ < end_ && !is_space(*p_)) ++p_;
        return {start, static_cast<size_t>(p_ - start)};
    }
    // Every remaining word; the last one ("" if none).
    std::string_view last_word() {
        std::string_view last;
        for (std::string_view w = word(); !w.empty(); w = word()) last = w;
        return last;
    }
    // The next word read as a number; false (and `out` untouched) if there is
    // no word or it does not start with one.
    bool number(double& out) {
        const std::string_view w = word();
        return parse(w, out);
    }
    static bool parse(std::string_view w, double& out) {
        const char* b = w.data();
        const char* e = b + w.size();
        if (b < e && *b == '+') ++b;
        return b < e && std::from_chars(b, e, out).ec == std::errc();
    }

  private:
    const char* p_;
    const char* end_;
};

// Calls f(line) for every line of the text.
template <class F>
void for_each_line(std::string_view text, F&& f) {
    size_t at = 0;
    while (at <= text.size()) {
        size_t nl = text.find('\n', at);
        if (nl == std::string_view::npos) nl = text.size();
        f(text.substr(at, nl - at));
        at = nl + 1;
    }
}

// ---------------------------------------------------------------------------
// MTL: the fields the four material kinds read.
// ---------------------------------------------------------------------------
struct surface_spec {
    std::array<double, 3> diffuse{0.8, 0.8, 0.8}, specular{0, 0, 0}, emissive{0, 0, 0};
    double shininess = 0, ior = 1.5, dissolve = 1;
    int illum = 2;
    std::string diffuse_map;   // map_Kd's image name, "" if none
};

using surface_library = std::unordered_map<std::string, surface_spec>;

void scan_library(std::string_view text, surface_library& lib) {
    surface_spec* cur = nullptr;
    for_each_line(text, [&](std::string_view line) {
        line_cursor c(line);
        const std::string_view key = c.word();
        if (key == "newmtl") {
            cur = &lib[std::string(c.word())];
            *cur = surface_spec();
            return;
        }
        if (!cur) return;
        auto scalar = [&](double& field) {   // a failed read stores 0
            double x;
            field = c.number(x) ? x : 0.0;
        };
        if (key == "Kd" || key == "Ks" || key == "Ke") {
            std::array<double, 3>& rgb = key == "Kd" ? cur->diffuse : key == "Ks" ? cur->specular : cur->emissive;
            double r, g, b;
            if (!c.number(r)) return;
            const bool full = c.number(g) && c.number(b);
            rgb = full ? std::array<double, 3>{r, g, b} : std::array<double, 3>{r, r, r};   // one value: gray
        } else if (key == "Ns") {
            scalar(cur->shininess);
        } else if (key == "Ni") {
            scalar(cur->ior);
        } else if (key == "d") {
            scalar(cur->dissolve);
        } else if (key == "Tr") {
            double tr;
            if (c.number(tr)) cur->dissolve = 1 - tr;
        } else if (key == "illum") {
            double m;
            cur->illum = c.number(m) ? static_cast<int>(m) : 0;
        } else if (key == "map_Kd") {
            const std::string_view name = c.last_word();   // options come before the name
            if (!name.empty()) cur->diffuse_map.assign(name);
        }
    });
}

// The material kind a specification makes, tested in the reference's order
// (obj.hpp:89-112): emission -> light; dissolve below 1 or a transparent
// illumination model (4, 6, 7, 9) -> glass; a reflective model (3, 5) ->
// metal, tinted by Ks (else Kd), with a fuzz of sqrt(2 / (Ns + 2)) capped at
// 1; anything else matte.
material_rec material_of(const surface_spec& m) {
    auto any = [](const std::array<double, 3>& c) { return c[0] > 0 || c[1] > 0 || c[2] > 0; };
    auto rgb = [](const std::array<double, 3>& c) {
        return v3(static_cast<float>(c[0]), static_cast<float>(c[1]), static_cast<float>(c[2]));
    };
    material_rec r;
    const bool transparent_model = m.illum == 4 || m.illum == 6 || m.illum == 7 || m.illum == 9;
    if (any(m.emissive)) {
        r.type = 3;
        r.emission = rgb(m.emissive);
    } else if (m.dissolve < 1 || transparent_model) {
        r.type = 2;
        r.ior = static_cast<float>(m.ior > 1 ? m.ior : 1.5);
    } else if (m.illum == 3 || m.illum == 5) {
        r.type = 1;
        r.albedo = rgb(any(m.specular) ? m.specular : m.diffuse);
        const float fuzz = static_cast<float>(std::min(1.0, std::sqrt(2 / (m.shininess + 2))));
        r.fuzz = fuzz < 1 ? fuzz : 1;   // metal's constructor clamp (material.hpp:83)
    } else {
        r.type = 0;
        r.albedo = rgb(m.diffuse);
    }
    return r;
}

// ---------------------------------------------------------------------------
// OBJ scan: vertex data and a face table.
// ---------------------------------------------------------------------------
struct face_table {
    std::vector<std::array<int32_t, 3>> pos, uv, nrm;   // -1: the corner has none
    std::vector<uint32_t> material;
    size_t size() const { return pos.size(); }
};

struct obj_data {
    std::vector<v3> positions, normals;
    std::vector<std::array<float, 2>> uvs;
    face_table faces;
};

// A corner field's index into a list of `count` entries so far: 1-based, or
// negative counting back from the end; -1 for an empty field.
int32_t corner_index(std::string_view field, size_t count) {
    if (field.empty()) return -1;
    const char* b = field.data();
    const char* e = b + field.size();
    if (*b == '+') ++b;
    long i = 0;
    if (std::from_chars(b, e, i).ec != std::errc())
        throw std::runtime_error("OBJ face corner is not a number: " + std::string(field));
    const long k = i > 0 ? i - 1 : static_cast<long>(count) + i;
    if (k < 0 || k >= static_cast<long>(count))
        throw std::runtime_error("OBJ face index out of range: " + std::string(field));
    return static_cast<int32_t>(k);
}

// The scan. `resolve(name)` gives the material slot a usemtl name selects;
// `library_line` is called for each mtllib line that names a library.
template <class Resolve, class Library>
obj_data scan_obj(std::string_view text, uint32_t fallback, Resolve&& resolve, Library&& library_line) {
    obj_data d;
    uint32_t current = fallback;
    std::vector<std::array<int32_t, 3>> poly;
    for_each_line(text, [&](std::string_view line) {
        line_cursor c(line);
        const std::string_view key = c.word();
        if (key == "v" || key == "vn") {
            double xyz[3] = {0, 0, 0};
            for (double& x : xyz)
                if (!c.number(x)) x = 0;
            (key == "v" ? d.positions : d.normals)
                .emplace_back(static_cast<float>(xyz[0]), static_cast<float>(xyz[1]), static_cast<float>(xyz[2]));
        } else if (key == "vt") {
            double uv[2] = {0, 0};
            for (double& x : uv)
                if (!c.number(x)) x = 0;
            d.uvs.push_back({static_cast<float>(uv[0]), static_cast<float>(uv[1])});
        } else if (key == "mtllib") {
            if (!c.word().empty()) library_line();
        } else if (key == "usemtl") {
            current = resolve(c.word());
        } else if (key == "f") {
            poly.clear();
            for (std::string_view corner = c.word(); !corner.empty(); corner = c.word()) {
                // v, v/vt, v//vn or v/vt/vn
                const size_t a = corner.find('/');
                const size_t b = a == std::string_view::npos ? a : corner.find('/', a + 1);
                std::array<int32_t, 3> k{-1, -1, -1};
                k[0] = corner_index(corner.substr(0, a), d.positions.size());
                if (k[0] < 0) throw std::runtime_error("OBJ face corner without a vertex");
                if (a != std::string_view::npos) k[1] = corner_index(corner.substr(a + 1, b - a - 1), d.uvs.size());
                if (b != std::string_view::npos) k[2] = corner_index(corner.substr(b + 1), d.normals.size());
                poly.push_back(k);
            }
            for (size_t i = 2; i < poly.size(); ++i) {   // a fan around the first corner
                const std::array<int32_t, 3>* tri[3] = {&poly[0], &poly[i - 1], &poly[i]};
                std::array<int32_t, 3> pos, uv, nrm;
                for (int j = 0; j < 3; ++j) {
                    pos[j] = (*tri[j])[0];
                    uv[j] = (*tri[j])[1];
                    nrm[j] = (*tri[j])[2];
                }
                d.faces.pos.push_back(pos);
                d.faces.uv.push_back(uv);
                d.faces.nrm.push_back(nrm);
                d.faces.material.push_back(current);
            }
        }
    });
    return d;
}

// ---------------------------------------------------------------------------
// place, normals, emit: the reference's arithmetic (see the file comment).
// ---------------------------------------------------------------------------
v3 cross_f(v3 a, v3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

void place(obj_data& d, v3 base, double height, double turn_deg) {
    const double angle = turn_deg * M_PI / 180, cs = std::cos(angle), sn = std::sin(angle);
    auto turn = [&](v3& p) {
        const double x = p.x, z = p.z;
        p = v3(static_cast<float>(cs * x + sn * z), p.y, static_cast<float>(-sn * x + cs * z));
    };
    for (v3& p : d.positions) turn(p);
    for (v3& n : d.normals) turn(n);
    v3 lo = d.positions[0], hi = d.positions[0];
    for (const v3& p : d.positions)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    const float fit = static_cast<float>(height / (hi.y - lo.y));
    const v3 foot(static_cast<float>((lo.x + hi.x) * 0.5), lo.y, static_cast<float>((lo.z + hi.z) * 0.5));
    for (v3& p : d.positions) p = base + (p - foot) * fit;
}

// The sum of the face normals (cross products: twice the faces' areas) at
// each vertex, in face order.
std::vector<v3> vertex_normals(const obj_data& d) {
    std::vector<v3> sum(d.positions.size());
    for (size_t f = 0; f < d.faces.size(); ++f) {
        const auto& k = d.faces.pos[f];
        const v3 p0 = d.positions[k[0]];
        const v3 n = cross_f(d.positions[k[1]] - p0, d.positions[k[2]] - p0);
        for (int j = 0; j < 3; ++j) sum[k[j]] = sum[k[j]] + n;
    }
    return sum;
}

void emit(scene_desc& s, const obj_data& d, const std::vector<v3>& smooth) {
    s.triangles.reserve(s.triangles.size() + d.faces.size());
    for (size_t f = 0; f < d.faces.size(); ++f) {
        const auto &pk = d.faces.pos[f], &tk = d.faces.uv[f], &nk = d.faces.nrm[f];
        triangle_rec t;
        for (int j = 0; j < 3; ++j) t.p[j] = d.positions[pk[j]];
        const v3 fn = cross_f(t.p[1] - t.p[0], t.p[2] - t.p[0]);
        if (dot(fn, fn) < 1e-24) continue;   // no area
        const bool file_normals = nk[0] >= 0 && nk[1] >= 0 && nk[2] >= 0;
        for (int j = 0; j < 3; ++j) {
            t.n[j] = normalize(file_normals ? d.normals[nk[j]] : smooth[pk[j]]);
            if (tk[j] >= 0) {
                t.uv[2 * j] = d.uvs[tk[j]][0];
                t.uv[2 * j + 1] = d.uvs[tk[j]][1];
            }
        }
        t.material = d.faces.material[f];
        s.triangles.push_back(t);
    }
}

}  // namespace

void add_obj(scene_desc& s, const std::string& text, v3 base, double height, uint32_t material, double turn_deg,
             const std::string* mtl_text, const std::vector<image_rec>* images) {
    if (material >= s.materials.size()) throw std::runtime_error("OBJ material index out of range");
    surface_library lib;
    std::unordered_map<std::string, uint32_t> made;      // material name -> scene material, made on first use
    std::unordered_map<std::string, uint32_t> textures;  // image name -> scene texture, added on first use
    auto texture_of = [&](const std::string& name) -> uint32_t {
        if (!images || name.empty()) return 0;
        const auto seen = textures.find(name);
        if (seen != textures.end()) return seen->second;
        for (const image_rec& im : *images)
            if (im.name == name) {
                s.textures.push_back(im);
                const uint32_t t = static_cast<uint32_t>(s.textures.size());   // texture number + 1
                textures.emplace(name, t);
                return t;
            }
        return 0;
    };
    auto resolve = [&](std::string_view name) -> uint32_t {
        const std::string key(name);
        const auto spec = lib.find(key);
        if (spec == lib.end()) return material;
        const auto done = made.find(key);
        if (done != made.end()) return done->second;
        material_rec m = material_of(spec->second);
        if (m.type == 0) m.texture = texture_of(spec->second.diffuse_map);
        const uint32_t k = s.add_material(m);
        made.emplace(key, k);
        return k;
    };
    // The MTL text stands in for every library an mtllib line names.
    auto library_line = [&] {
        if (mtl_text) scan_library(*mtl_text, lib);
    };
    obj_data d = scan_obj(text, material, resolve, library_line);
    if (d.positions.empty() || d.faces.size() == 0) return;
    place(d, base, height, turn_deg);
    emit(s, d, vertex_normals(d));
}

}  // namespace pt
