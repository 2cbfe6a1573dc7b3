// Seeded procedural scenes for the benchmark configurations.
//
// None of these scenes exist in the reference (its own are built at
// renderer.hpp:158-343); they are described by the benchmark and generated
// here into the reference's flat arrays (PTSD), in a fixed draw order from a
// documented generator (splitmix64), so every build places the same scene.
// Smooth normals are area weighted as the reference's OBJ loader makes them
// (obj.hpp:198-204): the sum of the unnormalised face cross products.
#include "host.hpp"

#include <algorithm>
#include <array>
#include <map>

namespace pt {
namespace {

struct rng64 {
    uint64_t s;
    explicit rng64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double u() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double u(double a, double b) { return a + (b - a) * u(); }
};

inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// 3-octave value noise in [-1, 1]: hashed lattice values, smoothstep
// trilinear interpolation, base frequency f, each octave doubling the
// frequency and halving the amplitude.
struct value_noise {
    uint32_t seed;
    double lattice(int x, int y, int z) const {
        uint32_t h = hash32(static_cast<uint32_t>(x) * 0x8da6b343u ^ hash32(static_cast<uint32_t>(y) * 0xd8163841u ^
                                                                              hash32(static_cast<uint32_t>(z) * 0xcb1ab31fu ^ seed)));
        return h * (2.0 / 4294967295.0) - 1.0;
    }
    double octave(double x, double y, double z) const {
        const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        const int ix = static_cast<int>(fx), iy = static_cast<int>(fy), iz = static_cast<int>(fz);
        auto s = [](double t) { return t * t * (3 - 2 * t); };
        const double tx = s(x - fx), ty = s(y - fy), tz = s(z - fz);
        double c[2][2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d);
        auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
        const double x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
        const double x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
        return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
    }
    double operator()(double x, double y, double z, double f) const {
        double sum = 0, amp = 1, norm = 0;
        for (int o = 0; o < 3; ++o) {
            sum += amp * octave(x * f + 17.0 * o, y * f + 31.0 * o, z * f + 47.0 * o);
            norm += amp;
            amp *= 0.5;
            f *= 2;
        }
        return sum / norm;
    }
};

struct indexed_mesh {
    std::vector<std::array<double, 3>> verts;
    std::vector<std::array<double, 3>> dirs;   // displacement direction per vertex
    std::vector<std::array<uint32_t, 3>> faces;
};

// Geodesic icosphere of frequency n: each icosahedron face split into n^2
// triangles, points pushed out to the unit sphere; 20 n^2 triangles. Points
// on an edge are computed once, from the edge's lower-numbered corner, so
// neighbouring faces share them exactly.
indexed_mesh geodesic_sphere(int n) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double iv[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                              {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    int fv[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                     {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                     {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    indexed_mesh m;
    auto unit = [](double x, double y, double z) {
        const double l = std::sqrt(x * x + y * y + z * z);
        return std::array<double, 3>{x / l, y / l, z / l};
    };
    for (auto& v : iv) m.verts.push_back(unit(v[0], v[1], v[2]));
    // Outward winding: flip any face whose normal points inward.
    for (auto& f : fv) {
        const auto &a = m.verts[f[0]], &b = m.verts[f[1]], &c = m.verts[f[2]];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
        if (nx * (a[0] + b[0] + c[0]) + ny * (a[1] + b[1] + c[1]) + nz * (a[2] + b[2] + c[2]) < 0) std::swap(f[1], f[2]);
    }
    std::map<std::pair<int, int>, uint32_t> edge_first;  // first vertex id of an edge's interior points
    auto edge_point = [&](int a, int b, int k) -> uint32_t {
        // k steps of n from a towards b (0 < k < n)
        if (a > b) {
            std::swap(a, b);
            k = n - k;
        }
        auto key = std::make_pair(a, b);
        auto it = edge_first.find(key);
        if (it == edge_first.end()) {
            const uint32_t first = static_cast<uint32_t>(m.verts.size());
            const auto &A = m.verts[a], &B = m.verts[b];
            const std::array<double, 3> Ac = {iv[a][0], iv[a][1], iv[a][2]}, Bc = {iv[b][0], iv[b][1], iv[b][2]};
            (void)A;
            (void)B;
            for (int s = 1; s < n; ++s) {
                const double w = static_cast<double>(s) / n;
                m.verts.push_back(unit(Ac[0] + (Bc[0] - Ac[0]) * w, Ac[1] + (Bc[1] - Ac[1]) * w, Ac[2] + (Bc[2] - Ac[2]) * w));
            }
            it = edge_first.emplace(key, first).first;
        }
        return it->second + static_cast<uint32_t>(k - 1);
    };
    for (auto& f : fv) {
        const int A = f[0], B = f[1], C = f[2];
        // Grid (i, j): i steps toward B, j toward C.
        std::vector<uint32_t> id((n + 1) * (n + 1), ~0u);
        auto at = [&](int i, int j) -> uint32_t& { return id[i * (n + 1) + j]; };
        for (int i = 0; i <= n; ++i)
            for (int j = 0; i + j <= n; ++j) {
                uint32_t v;
                if (i == 0 && j == 0) v = A;
                else if (i == n) v = B;
                else if (j == n) v = C;
                else if (j == 0) v = edge_point(A, B, i);
                else if (i == 0) v = edge_point(A, C, j);
                else if (i + j == n) v = edge_point(B, C, j);
                else {
                    const double wa = static_cast<double>(n - i - j) / n, wb = static_cast<double>(i) / n,
                                 wc = static_cast<double>(j) / n;
                    v = static_cast<uint32_t>(m.verts.size());
                    m.verts.push_back(unit(iv[A][0] * wa + iv[B][0] * wb + iv[C][0] * wc,
                                           iv[A][1] * wa + iv[B][1] * wb + iv[C][1] * wc,
                                           iv[A][2] * wa + iv[B][2] * wb + iv[C][2] * wc));
                }
                at(i, j) = v;
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; i + j < n; ++j) {
                m.faces.push_back({at(i, j), at(i + 1, j), at(i, j + 1)});
                if (i + j < n - 1) m.faces.push_back({at(i + 1, j), at(i + 1, j + 1), at(i, j + 1)});
            }
    }
    m.dirs = m.verts;
    return m;
}

// Torus in the xz plane: major radius R, minor r, nu x nv quads, 2 nu nv triangles.
indexed_mesh torus(double R, double r, int nu, int nv) {
    indexed_mesh m;
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            const double u = 2 * M_PI * i / nu, v = 2 * M_PI * j / nv;
            const double cx = std::cos(u), cz = std::sin(u);
            m.verts.push_back({(R + r * std::cos(v)) * cx, r * std::sin(v), (R + r * std::cos(v)) * cz});
            m.dirs.push_back({std::cos(v) * cx, std::sin(v), std::cos(v) * cz});
        }
    auto id = [&](int i, int j) { return static_cast<uint32_t>((i % nu) * nv + (j % nv)); };
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            m.faces.push_back({id(i, j), id(i, j + 1), id(i + 1, j + 1)});
            m.faces.push_back({id(i, j), id(i + 1, j + 1), id(i + 1, j)});
        }
    // Outward winding, checked against the analytic normal.
    for (auto& f : m.faces) {
        const auto &a = m.verts[f[0]], &b = m.verts[f[1]], &c = m.verts[f[2]];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
        const auto& d = m.dirs[f[0]];
        if (nx * d[0] + ny * d[1] + nz * d[2] < 0) std::swap(f[1], f[2]);
    }
    return m;
}

// Displaces every vertex along its direction by amplitude * noise, scales
// and moves it, then emits triangles with area-weighted smooth normals.
// `material_of(face)` picks each face's material.
template <class MatFn>
std::vector<triangle_rec> finish_mesh(indexed_mesh& m, const value_noise& noise, double freq, double amplitude,
                                      double scale, v3 offset, MatFn&& material_of) {
    std::vector<v3> p(m.verts.size());
    for (size_t i = 0; i < m.verts.size(); ++i) {
        const auto &v = m.verts[i], &d = m.dirs[i];
        const double h = amplitude * noise(v[0], v[1], v[2], freq);
        p[i] = v3(static_cast<float>((v[0] + d[0] * h) * scale + offset.x),
                  static_cast<float>((v[1] + d[1] * h) * scale + offset.y),
                  static_cast<float>((v[2] + d[2] * h) * scale + offset.z));
    }
    std::vector<v3> smooth(p.size());
    for (const auto& f : m.faces) {
        const v3 n = cross(p[f[1]] - p[f[0]], p[f[2]] - p[f[0]]);
        for (uint32_t c : f) smooth[c] = smooth[c] + n;
    }
    std::vector<triangle_rec> out;
    out.reserve(m.faces.size());
    for (size_t k = 0; k < m.faces.size(); ++k) {
        const auto& f = m.faces[k];
        triangle_rec t;
        for (int c = 0; c < 3; ++c) {
            t.p[c] = p[f[c]];
            t.n[c] = normalize(smooth[f[c]]);
        }
        t.material = material_of(static_cast<uint32_t>(k));
        out.push_back(t);
    }
    return out;
}

material_rec lambertian(double r, double g, double b) {
    material_rec m;
    m.type = 0;
    m.albedo = v3(static_cast<float>(r), static_cast<float>(g), static_cast<float>(b));
    return m;
}
material_rec metal(double r, double g, double b, double fuzz) {
    material_rec m = lambertian(r, g, b);
    m.type = 1;
    m.fuzz = static_cast<float>(fuzz);
    return m;
}
material_rec glass(double ior) {
    material_rec m;
    m.type = 2;
    m.ior = static_cast<float>(ior);
    return m;
}
material_rec light(double e) {
    material_rec m;
    m.type = 3;
    m.emission = v3(static_cast<float>(e), static_cast<float>(e), static_cast<float>(e));
    return m;
}

// Config 0: the sphere field.
scene_desc sphere_field(uint64_t seed, int grid_lo, int grid_hi) {
    scene_desc s;
    rng64 rng(seed ^ 0x5eed0000ull);
    s.large.push_back({v3(0, -1000, 0), 1000, s.add_material(lambertian(0.5, 0.5, 0.5))});
    for (int a = grid_lo; a < grid_hi; ++a)
        for (int b = grid_lo; b < grid_hi; ++b) {
            const double jx = rng.u(-0.45, 0.45), jz = rng.u(-0.45, 0.45);
            const double pick = rng.u();
            material_rec m;
            if (pick < 0.8) {
                double c[3];
                for (double& x : c) {
                    x = rng.u();
                    x *= rng.u();
                }
                m = lambertian(c[0], c[1], c[2]);
            } else if (pick < 0.95) {
                double c[3];
                for (double& x : c) x = rng.u(0.5, 1);
                m = metal(c[0], c[1], c[2], rng.u(0, 0.5));
            } else {
                m = glass(1.5);
            }
            s.spheres.push_back({v3(static_cast<float>(a + 0.5 + jx), 0.2f, static_cast<float>(b + 0.5 + jz)), 0.2f,
                                 s.add_material(m)});
        }
    s.spheres.push_back({v3(0, 1, 0), 1, s.add_material(glass(1.5))});
    s.spheres.push_back({v3(-4, 1, 0), 1, s.add_material(lambertian(0.4, 0.2, 0.1))});
    s.spheres.push_back({v3(4, 1, 0), 1, s.add_material(metal(0.7, 0.6, 0.5, 0.0))});
    s.spheres.push_back({v3(-2, 3, 2.5f), 0.5f, s.add_material(light(8))});
    s.lookfrom = v3(13, 2, 3);
    s.lookat = v3(0, 0, 0);
    s.vfov = 20;
    return s;
}

// A displaced geodesic sphere of frequency n, 70% lambertian / 30% metal by
// face hash, radius 1 at the origin, with two r = 0.1 lights of radiance 50.
scene_desc displaced_sphere(uint64_t seed, int n) {
    scene_desc s;
    const uint32_t matte = s.add_material(lambertian(0.65, 0.6, 0.55));
    const uint32_t shiny = s.add_material(metal(0.9, 0.8, 0.6, 0.2));
    indexed_mesh m = geodesic_sphere(n);
    const value_noise noise{hash32(static_cast<uint32_t>(seed) ^ 0x1234567u)};
    const uint32_t fs = hash32(static_cast<uint32_t>(seed >> 32) ^ 0x9abcdefu);
    s.triangles = finish_mesh(m, noise, 4.0, 0.05, 1.0, v3(0, 0, 0), [&](uint32_t f) {
        return (hash32(f ^ fs) % 10u) < 7u ? matte : shiny;
    });
    const uint32_t lamp = s.add_material(light(50));
    s.spheres.push_back({v3(1.6f, 1.4f, 1.2f), 0.1f, lamp});
    s.spheres.push_back({v3(-1.5f, 0.8f, 1.5f), 0.1f, lamp});
    s.lookfrom = v3(0, 0.3f, 3.6f);
    s.lookat = v3(0, 0, 0);
    s.vfov = 40;
    return s;
}

// Shoemake's uniform random rotation from three uniforms.
void random_rotation(rng64& rng, float m[3][3]) {
    const double u1 = rng.u(), u2 = rng.u(), u3 = rng.u();
    const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    const double qx = a * std::sin(2 * M_PI * u2), qy = a * std::cos(2 * M_PI * u2);
    const double qz = b * std::sin(2 * M_PI * u3), qw = b * std::cos(2 * M_PI * u3);
    const double r[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                            {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                            {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = static_cast<float>(r[i][j]);
}

// Config 2: three meshes placed by 4096 instances.
scene_desc instanced(uint64_t seed, double scale) {
    scene_desc s;
    rng64 rng(seed ^ 0x1257a2ull);
    const uint32_t base = s.add_material(lambertian(0.7, 0.7, 0.7));
    const value_noise noise{hash32(static_cast<uint32_t>(seed) ^ 0x51f15eedu)};
    auto dim = [&](double v) { return std::max(3, static_cast<int>(std::lround(v * std::sqrt(scale)))); };
    {
        indexed_mesh m = torus(1.0, 0.35, dim(250), dim(100));  // 50,000 triangles at scale 1
        s.meshes.push_back(finish_mesh(m, noise, 3.0, 0.05, 1.0, v3(0, 0, 0), [&](uint32_t) { return base; }));
    }
    {
        indexed_mesh m = geodesic_sphere(dim(100));  // 200,000 triangles
        s.meshes.push_back(finish_mesh(m, noise, 4.0, 0.05, 1.0, v3(0, 0, 0), [&](uint32_t) { return base; }));
    }
    {
        indexed_mesh m = torus(1.0, 0.2, dim(1000), dim(250));  // 500,000 triangles
        s.meshes.push_back(finish_mesh(m, noise, 5.0, 0.05, 1.0, v3(0, 0, 0), [&](uint32_t) { return base; }));
    }
    const int n_inst = std::max(8, static_cast<int>(std::lround(4096 * std::min(1.0, scale * 4))));
    for (int k = 0; k < n_inst; ++k) {
        instance_rec in;
        in.mesh = static_cast<uint32_t>(rng.next() % 3);
        float rot[3][3];
        random_rotation(rng, rot);
        const double sc = rng.u(0.2, 2.0);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) in.to_world.m[i][j] = static_cast<float>(rot[i][j] * sc);
        in.to_world.t = v3(static_cast<float>(rng.u(-50, 50)), static_cast<float>(rng.u(-50, 50)),
                           static_cast<float>(rng.u(-50, 50)));
        if (rng.u() < 0.5) {
            const double pick = rng.u();
            double c[3];
            for (double& x : c) x = rng.u();
            if (pick < 0.6) in.material = s.add_material(lambertian(c[0] * 0.8, c[1] * 0.8, c[2] * 0.8));
            else if (pick < 0.9) in.material = s.add_material(metal(0.5 + 0.5 * c[0], 0.5 + 0.5 * c[1], 0.5 + 0.5 * c[2], 0.2 * rng.u()));
            else in.material = s.add_material(glass(1.5));
        }
        s.instances.push_back(in);
    }
    s.lookfrom = v3(0, 20, 130);
    s.lookat = v3(0, 0, 0);
    s.vfov = 45;
    return s;
}

void add_quad(scene_desc& s, v3 p, v3 e1, v3 e2, uint32_t mat) {
    // renderer.hpp:235-238, flat shaded: vertex normals = the face normal.
    auto tri = [&](v3 a, v3 b, v3 c) {
        triangle_rec t;
        t.p[0] = a;
        t.p[1] = b;
        t.p[2] = c;
        const v3 fn = normalize(cross(b - a, c - a));
        t.n[0] = t.n[1] = t.n[2] = fn;
        t.material = mat;
        s.triangles.push_back(t);
    };
    tri(p, p + e1, p + e1 + e2);
    tri(p, p + e1 + e2, p + e2);
}

// Config 4: a closed room (10 x 4 x 8) holding a displaced mesh, lit only by
// two small lamps in an alcove behind a partition with a 0.5 x 0.5 opening.
scene_desc guided_room(uint64_t seed, double scale) {
    scene_desc s;
    s.sky = false;
    const uint32_t white = s.add_material(lambertian(0.7, 0.7, 0.7));
    const uint32_t warm = s.add_material(lambertian(0.7, 0.45, 0.3));
    const float x0 = -5, x1 = 5, y0 = 0, y1 = 4, z0 = -4, z1 = 4;
    add_quad(s, v3(x0, y0, z0), v3(x1 - x0, 0, 0), v3(0, 0, z1 - z0), white);  // floor
    add_quad(s, v3(x0, y1, z0), v3(x1 - x0, 0, 0), v3(0, 0, z1 - z0), white);  // ceiling
    add_quad(s, v3(x0, y0, z0), v3(x1 - x0, 0, 0), v3(0, y1 - y0, 0), white);  // back
    add_quad(s, v3(x0, y0, z1), v3(x1 - x0, 0, 0), v3(0, y1 - y0, 0), white);  // front
    add_quad(s, v3(x0, y0, z0), v3(0, 0, z1 - z0), v3(0, y1 - y0, 0), warm);   // left
    add_quad(s, v3(x1, y0, z0), v3(0, 0, z1 - z0), v3(0, y1 - y0, 0), white);  // right
    // Partition at x = 3.5 with a 0.5 x 0.5 opening at y in [1.75, 2.25], z in [-0.25, 0.25].
    const float px = 3.5f, hy0 = 1.75f, hy1 = 2.25f, hz0 = -0.25f, hz1 = 0.25f;
    add_quad(s, v3(px, y0, z0), v3(0, 0, z1 - z0), v3(0, hy0 - y0, 0), white);     // below the opening
    add_quad(s, v3(px, hy1, z0), v3(0, 0, z1 - z0), v3(0, y1 - hy1, 0), white);    // above
    add_quad(s, v3(px, hy0, z0), v3(0, 0, hz0 - z0), v3(0, hy1 - hy0, 0), white);  // beside, -z
    add_quad(s, v3(px, hy0, hz1), v3(0, 0, z1 - hz1), v3(0, hy1 - hy0, 0), white); // beside, +z
    // The mesh is diffuse (two albedos by face hash): the scene tests how
    // diffuse interreflection finds light through the opening, the transport
    // a guide for diffuse bounces addresses.
    const uint32_t matte = s.add_material(lambertian(0.65, 0.6, 0.55));
    const uint32_t matte2 = s.add_material(lambertian(0.45, 0.5, 0.6));
    const int n = std::max(2, static_cast<int>(std::lround(224 * std::sqrt(scale))));
    indexed_mesh m = geodesic_sphere(n);
    const value_noise noise{hash32(static_cast<uint32_t>(seed) ^ 0x7001u)};
    const uint32_t fs = hash32(static_cast<uint32_t>(seed >> 32) ^ 0x7002u);
    auto mesh = finish_mesh(m, noise, 4.0, 0.08, 1.2, v3(-1.0f, 1.25f, 0.0f), [&](uint32_t f) {
        return (hash32(f ^ fs) % 10u) < 7u ? matte : matte2;
    });
    s.triangles.insert(s.triangles.end(), mesh.begin(), mesh.end());
    const uint32_t lamp = s.add_material(light(2000));
    s.spheres.push_back({v3(4.5f, 2.0f, 0.9f), 0.05f, lamp});
    s.spheres.push_back({v3(4.5f, 2.0f, -0.9f), 0.05f, lamp});
    s.lookfrom = v3(-4.2f, 2.6f, 3.6f);
    s.lookat = v3(0.5f, 1.4f, -0.5f);
    s.vfov = 65;
    return s;
}

}  // namespace

scene_desc generate_config(int config, uint64_t seed, double scale) {
    if (!(scale > 0) || scale > 1) throw std::runtime_error("scale must be in (0, 1]");
    switch (config) {
        case 0: {
            // 22 x 22 at scale 1; smaller grids for tests.
            const int half = std::max(1, static_cast<int>(std::lround(11 * std::sqrt(scale))));
            return sphere_field(seed, -half, half);
        }
        case 1: {
            const int n = std::max(1, static_cast<int>(std::lround(224 * std::sqrt(scale))));  // 1,003,520 triangles
            return displaced_sphere(seed, n);
        }
        case 2: return instanced(seed, scale);
        case 4: return guided_room(seed, scale);
        default: throw std::runtime_error("unknown scene configuration (0, 1, 2 or 4)");
    }
}

scene_desc generate_job(uint64_t seed, double scale, job_desc& d) {
    if (!(scale > 0) || scale > 1) throw std::runtime_error("scale must be in (0, 1]");
    rng64 rng(seed * 0x2545f4914f6cdd1dull + 0x0b5e5ull);
    d.seed = seed;
    d.kind = rng.u() < 0.5 ? 0 : 1;
    static const int res[4][2] = {{256, 256}, {512, 512}, {1024, 576}, {1024, 1024}};
    static const int spps[3] = {4, 16, 64};
    const int r = static_cast<int>(rng.next() % 4), k = static_cast<int>(rng.next() % 3);
    d.width = res[r][0];
    d.height = res[r][1];
    d.spp = spps[k];
    if (d.kind == 0) {
        const uint32_t count = 50 + static_cast<uint32_t>(rng.next() % 1951);  // U{50..2000}
        d.primitives = count;
        scene_desc s;
        rng64 g(seed ^ 0xba7c0000ull);
        s.large.push_back({v3(0, -1000, 0), 1000, s.add_material(lambertian(0.5, 0.5, 0.5))});
        const double side = std::sqrt(static_cast<double>(count)) * 1.1;
        for (uint32_t i = 0; i < count; ++i) {
            const double pick = g.u();
            double c[3];
            for (double& x : c) x = g.u();
            material_rec m = pick < 0.8 ? lambertian(c[0] * c[0], c[1] * c[1], c[2] * c[2])
                           : pick < 0.95 ? metal(0.5 + 0.5 * c[0], 0.5 + 0.5 * c[1], 0.5 + 0.5 * c[2], 0.5 * g.u())
                                         : glass(1.5);
            const float rad = static_cast<float>(g.u(0.15, 0.35));
            s.spheres.push_back({v3(static_cast<float>(g.u(-side / 2, side / 2)), rad,
                                    static_cast<float>(g.u(-side / 2, side / 2))),
                                 rad, s.add_material(m)});
        }
        s.spheres.push_back({v3(0, static_cast<float>(side * 0.4 + 2), 0), 0.5f, s.add_material(light(8))});
        s.lookfrom = v3(static_cast<float>(side * 0.8), static_cast<float>(side * 0.35 + 1), static_cast<float>(side * 0.8));
        s.lookat = v3(0, 0, 0);
        s.vfov = 40;
        return s;
    }
    const uint32_t tris = 10000 + static_cast<uint32_t>(rng.next() % 990001);  // U[10k, 1M]
    const uint32_t scaled = std::max<uint32_t>(20, static_cast<uint32_t>(tris * scale));
    const int n = std::max(1, static_cast<int>(std::lround(std::sqrt(scaled / 20.0))));
    d.primitives = static_cast<uint32_t>(20 * n * n);
    return displaced_sphere(seed ^ 0x10b5ull, n);
}

}  // namespace pt
