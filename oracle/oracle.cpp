// ORACLE -- test infrastructure only. A plain C++ restatement, on the CPU, of
// the reference's render path over the packed PTGP scene. Only the tests, the
// smoke check and the benchmark's CPU-baseline leg may load it, and only as the
// checker: the product never calls it.
//
// What it restates (reference file:line):
//   or_trace      geometry::hit / occluded (geometry.hpp:111-190, 212-229),
//                 bvh::traverse (bvh.hpp:57-106) with SSE min/max semantics
//                 (simd4.hpp:15-18), sphere::intersect (sphere.hpp:25-48),
//                 triangle::intersect (triangle.hpp:40-61); counts node,
//                 leaf, primitive and instance visits per query.
//   or_trace_f64  the same closest-hit query in double precision, each
//                 primitive tested in FP64 and the tree used only to cull
//                 (boxes padded), as an independent check of t and primitive;
//                 the ray enters an instance's object space in double, or
//                 (float_ray) in float as the reference moves it.
//   or_sample_2d  the shuffled Owen-scrambled Sobol sampler
//                 (sampler.hpp:35-127), with the reference's byte tables.
//   or_render     ray_color, direct_light, the materials and
//                 render_row_impl (renderer.hpp:24-149,366-409;
//                 material.hpp:59-137; camera.hpp:45-47), drawing the
//                 non-QMC numbers from the per-path counter-based stream the
//                 GPU path defines (key = f(seed, pixel, sample), splitmix64),
//                 since the reference's per-thread sequential xoshiro stream
//                 (vec3.hpp:11-52) cannot be reproduced by a parallel tracer;
//                 textured mattes multiply the albedo by
//   or_texture_sample  image_texture::sample (texture.hpp:22-33) over the
//                 PTGP texture section's texels.
//
// Pinning (tests/test_oracle.py, tests/test_textures.py): or_trace against
// the reference's own traversal (oracle/_ref) bit for bit, and against
// or_trace_f64; or_sample_2d against the reference sampler;
// or_texture_sample against the reference's image_texture bit for bit;
// or_render against the reference renderer statistically, against closed
// forms (white furnace, one sphere light) and against the reference's own
// render fixtures (the textured "spot crowd" among them).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t NONE = 0xffffffffu;

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

struct V {
    float x, y, z;
};
V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V operator-(V a) { return {-a.x, -a.y, -a.z}; }
V operator*(V a, float t) { return {a.x * t, a.y * t, a.z * t}; }
V mul(V a, V b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V unit(V a) {
    const float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / l, a.y / l, a.z / l};
}

// SSE minps/maxps: the second operand unless the first compares less/greater.
float smin(float a, float b) { return a < b ? a : b; }
float smax(float a, float b) { return a > b ? a : b; }

struct scene {
    const uint32_t* w;
    const float* nodes;   // 32 words per node
    const uint32_t* nodes_u;
    const uint32_t* order;
    const float* sph;     // 8 per sphere
    const float* tri;     // 28 per triangle
    const float* inst;    // 28 per instance
    const float* mat;     // 12 per material
    const uint32_t* lights;
    const uint32_t* large;
    uint32_t n_nodes, n_order, n_sph, n_tri, n_inst, n_mat, n_lights, n_large;
    uint32_t flat_root, inst_root, n_small, flags;
    const uint32_t* tex_info = nullptr;   // per texture: width, height, first texel, 0
    const float* texels = nullptr;        // linear RGB + 0 per texel
    uint32_t n_tex = 0;
    V cam[4];

    explicit scene(const uint32_t* words) : w(words) {
        auto sec = [&](int s) { return words + words[2 + 2 * s] / 4; };
        auto len = [&](int s) { return words[3 + 2 * s]; };
        nodes = reinterpret_cast<const float*>(sec(0));
        nodes_u = sec(0);
        order = sec(1);
        sph = reinterpret_cast<const float*>(sec(2));
        tri = reinterpret_cast<const float*>(sec(3));
        inst = reinterpret_cast<const float*>(sec(4));
        mat = reinterpret_cast<const float*>(sec(5));
        lights = sec(6);
        large = sec(7);
        n_nodes = len(0) / 128;
        n_order = len(1) / 4;
        n_sph = len(2) / 32;
        n_tri = len(3) / 112;
        n_inst = len(4) / 112;
        n_mat = len(5) / 48;
        n_lights = len(6) / 4;
        n_large = len(7) / 4;
        flat_root = words[18];
        inst_root = words[19];
        n_small = words[20];
        flags = words[21];
        for (int r = 0; r < 4; ++r) cam[r] = {fb(words[22 + 4 * r]), fb(words[23 + 4 * r]), fb(words[24 + 4 * r])};
        n_tex = words[40];
        if (n_tex) {
            tex_info = words + words[38] / 4;
            texels = reinterpret_cast<const float*>(tex_info + 4 * n_tex);
        }
    }
};

// sphere::intersect (sphere.hpp:25-48), in float or double.
template <class P>
bool sphere_hit(const float* s, V o, V d, float t_min, float& t_max) {
    const P ox = P(o.x) - s[0], oy = P(o.y) - s[1], oz = P(o.z) - s[2];
    const P dx = d.x, dy = d.y, dz = d.z;
    const P a = dx * dx + dy * dy + dz * dz;
    const P half_b = ox * dx + oy * dy + oz * dz;
    const P c = ox * ox + oy * oy + oz * oz - P(s[3]) * s[3];
    const P disc = half_b * half_b - a * c;
    if (disc < 0) return false;
    const P sq = std::sqrt(disc);
    P root = (-half_b - sq) / a;
    if (root < t_min || root > t_max) {
        root = (-half_b + sq) / a;
        if (root < t_min || root > t_max) return false;
    }
    t_max = static_cast<float>(root);
    return true;
}

// triangle::intersect (triangle.hpp:40-61) on a packed triangle.
bool tri_hit(const float* t, V o, V d, float t_min, float& t_max, float& uo, float& vo) {
    const V p0{t[0], t[1], t[2]}, e1{t[4], t[5], t[6]}, e2{t[8], t[9], t[10]};
    const V pvec = cross(d, e2);
    const float det = dot(e1, pvec);
    if (std::fabs(det) < 1e-12f) return false;
    const float inv_det = 1 / det;
    const V tvec = o - p0;
    const float u = dot(tvec, pvec) * inv_det;
    if (u < 0 || u > 1) return false;
    const V qvec = cross(tvec, e1);
    const float v = dot(d, qvec) * inv_det;
    if (v < 0 || u + v > 1) return false;
    const float tt = dot(e2, qvec) * inv_det;
    if (tt < t_min || tt > t_max) return false;
    t_max = tt;
    uo = u;
    vo = v;
    return true;
}

struct counters {
    uint64_t nodes = 0, leafrefs = 0, tris = 0, spheres = 0, instances = 0, large = 0;
};

struct entry {
    uint32_t first, count;
    float t;
};

// bvh::traverse (bvh.hpp:57-106), calling test(primitive, t_max).
template <class Test>
void traverse(const scene& s, uint32_t root, V o, V d, float t_min, float& t_max, counters& c, Test&& test) {
    if (root == NONE) return;
    const float ix = 1 / d.x, iy = 1 / d.y, iz = 1 / d.z;
    std::vector<entry> stack;
    stack.reserve(64);
    stack.push_back({root, 0, t_min});
    while (!stack.empty()) {
        const entry e = stack.back();
        stack.pop_back();
        if (e.t >= t_max) continue;
        if (e.count > 0) {
            for (uint32_t i = e.first; i < e.first + e.count; ++i) {
                ++c.leafrefs;
                test(s.order[i], t_max);
            }
            continue;
        }
        ++c.nodes;
        const float* nd = s.nodes + 32 * size_t(e.first);
        const uint32_t* nu = s.nodes_u + 32 * size_t(e.first);
        entry hit[4];
        int n = 0;
        for (int ch = 0; ch < 4; ++ch) {
            const float x0 = (nd[0 + ch] - o.x) * ix, x1 = (nd[12 + ch] - o.x) * ix;
            const float y0 = (nd[4 + ch] - o.y) * iy, y1 = (nd[16 + ch] - o.y) * iy;
            const float z0 = (nd[8 + ch] - o.z) * iz, z1 = (nd[20 + ch] - o.z) * iz;
            const float tn = smax(smax(smin(x0, x1), smin(y0, y1)), smax(smin(z0, z1), t_min));
            const float tf = smin(smin(smax(x0, x1), smax(y0, y1)), smin(smax(z0, z1), t_max));
            if (!(tn <= tf)) continue;
            const entry h{nu[24 + ch], nu[28 + ch], tn};
            int k = n++;
            for (; k > 0 && hit[k - 1].t < h.t; --k) hit[k] = hit[k - 1];
            hit[k] = h;
        }
        for (int k = 0; k < n; ++k) stack.push_back(hit[k]);
    }
}

struct hit_out {
    float t = 0, u = 0, v = 0;
    uint32_t kind = 0, prim = 0, inst = NONE;
};

// geometry::hit and geometry::occluded over the packed scene.
bool trace(const scene& s, V o, V d, float t_min, float t_max, bool any, hit_out& h, counters& c) {
    h = hit_out();
    for (uint32_t k = 0; k < s.n_large; ++k) {
        ++c.large;
        const uint32_t sp = s.large[k];
        float t = t_max;
        if (sphere_hit<double>(s.sph + 8 * sp, o, d, t_min, t)) {
            if (any) {
                h.kind = 1;
                return true;
            }
            t_max = t;
            h.kind = 1;
            h.prim = sp;
        }
    }
    bool blocked = false;
    float u = 0, v = 0;
    auto test = [&](uint32_t p, float& t_far) {
        if (blocked) return;
        if (p < s.n_small) {
            ++c.spheres;
            float t = t_far;
            if (sphere_hit<float>(s.sph + 8 * p, o, d, t_min, t)) {
                if (any) {
                    blocked = true;
                    t_far = t_min;
                    return;
                }
                t_far = t;
                h.kind = 2;
                h.prim = p;
                h.inst = NONE;
            }
        } else {
            ++c.tris;
            float t = t_far;
            if (tri_hit(s.tri + 28 * size_t(p - s.n_small), o, d, t_min, t, u, v)) {
                if (any) {
                    blocked = true;
                    t_far = t_min;
                    return;
                }
                t_far = t;
                h.kind = 3;
                h.prim = p - s.n_small;
                h.inst = NONE;
                h.u = u;
                h.v = v;
            }
        }
    };
    float tf = t_max;
    traverse(s, s.flat_root, o, d, t_min, tf, c, test);
    if (blocked) {
        h.kind = 1;
        return true;
    }
    if (!any) t_max = tf;
    if (s.inst_root != NONE) {
        float ti = any ? t_max : t_max;
        auto test_instance = [&](uint32_t k, float& t_far) {
            if (blocked) return;
            ++c.instances;
            const float* in = s.inst + 28 * size_t(k);
            const float* m = in + 12;  // to_object rows
            const V lo{m[0] * o.x + m[1] * o.y + m[2] * o.z + m[3], m[4] * o.x + m[5] * o.y + m[6] * o.z + m[7],
                       m[8] * o.x + m[9] * o.y + m[10] * o.z + m[11]};
            const V ld{m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
                       m[8] * d.x + m[9] * d.y + m[10] * d.z};
            const uint32_t root = ub(in[24]);
            float ht = t_far, hu = 0, hv = 0;
            bool found = false;
            uint32_t which = 0;
            traverse(s, root, lo, ld, t_min, ht, c, [&](uint32_t q, float& t_mesh) {
                if (found && any) return;
                ++c.tris;
                if (tri_hit(s.tri + 28 * size_t(q), lo, ld, t_min, t_mesh, hu, hv)) {
                    found = true;
                    which = q;
                    if (any) t_mesh = t_min;
                }
            });
            if (!found) return;
            if (any) {
                blocked = true;
                t_far = t_min;
                return;
            }
            t_far = ht;
            h.kind = 3;
            h.prim = which;
            h.inst = k;
            h.u = hu;
            h.v = hv;
        };
        traverse(s, s.inst_root, o, d, t_min, ti, c, test_instance);
        if (blocked) {
            h.kind = 1;
            return true;
        }
        if (!any) t_max = ti;
    }
    if (any) return false;
    h.t = t_max;
    return h.kind != 0;
}

// image_texture::sample (texture.hpp:22-33): bilinear lookup at (u, v),
// repeating, v up the image whose rows are stored top first, over the linear
// texels the packer made with image_texture's table (texture.hpp:13-20).
V texture_sample(const scene& s, uint32_t t, float u, float v) {
    const uint32_t* info = s.tex_info + 4 * t;
    const int w = static_cast<int>(info[0]), h = static_cast<int>(info[1]);
    const float* tx = s.texels + 4 * size_t(info[2]);
    const float x = (u - std::floor(u)) * w - 0.5f;
    const float y = (1 - (v - std::floor(v))) * h - 0.5f;
    // A NaN or infinite coordinate reads texel 0 with NaN weights (finite
    // ones put x in [-0.5, w - 0.5]); the reference converts NaN to int with
    // its platform's result and multiplies by the same NaN weights.
    auto cell = [](float c) {
        const float f = std::floor(c);
        return f >= -1.0f && f <= 65536.0f ? static_cast<int>(f) : 0;
    };
    const int x0 = cell(x), y0 = cell(y);
    const float fx = x - x0, fy = y - y0;
    auto texel = [&](int xi, int yi) {
        xi = ((xi % w) + w) % w;
        yi = ((yi % h) + h) % h;
        const float* c = tx + 4 * (size_t(yi) * w + xi);
        return V{c[0], c[1], c[2]};
    };
    return texel(x0, y0) * ((1 - fx) * (1 - fy)) + texel(x0 + 1, y0) * (fx * (1 - fy)) +
           texel(x0, y0 + 1) * ((1 - fx) * fy) + texel(x0 + 1, y0 + 1) * (fx * fy);
}

// ---- FP64 closest hit: every candidate primitive in double -----------------

struct dv {
    double x, y, z;
};
dv dsub(dv a, dv b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double ddot(dv a, dv b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
dv dcross(dv a, dv b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

bool sphere64(const float* s, dv o, dv d, double t_min, double& t_max) {
    const dv oc = dsub(o, {s[0], s[1], s[2]});
    const double a = ddot(d, d), hb = ddot(oc, d), c = ddot(oc, oc) - double(s[3]) * s[3];
    const double disc = hb * hb - a * c;
    if (disc < 0) return false;
    const double sq = std::sqrt(disc);
    double r = (-hb - sq) / a;
    if (r < t_min || r > t_max) {
        r = (-hb + sq) / a;
        if (r < t_min || r > t_max) return false;
    }
    t_max = r;
    return true;
}

// Moller-Trumbore in double, from the triangle's own corners p0, p0+e1, p0+e2.
bool tri64(const float* t, dv o, dv d, double t_min, double& t_max) {
    const dv p0{t[0], t[1], t[2]}, e1{t[4], t[5], t[6]}, e2{t[8], t[9], t[10]};
    const dv pv = dcross(d, e2);
    const double det = ddot(e1, pv);
    if (std::fabs(det) < 1e-300) return false;
    const dv tv = dsub(o, p0);
    const double u = ddot(tv, pv) / det;
    if (u < 0 || u > 1) return false;
    const dv qv = dcross(tv, e1);
    const double v = ddot(d, qv) / det;
    if (v < 0 || u + v > 1) return false;
    const double tt = ddot(e2, qv) / det;
    if (tt < t_min || tt > t_max) return false;
    t_max = tt;
    return true;
}

// Every node box padded, then a double slab test: a conservative cull.
template <class Test>
void traverse64(const scene& s, uint32_t root, dv o, dv d, double t_min, double& t_max, Test&& test) {
    if (root == NONE) return;
    std::vector<uint32_t> stack{root};
    const double inf = std::numeric_limits<double>::infinity();
    while (!stack.empty()) {
        const uint32_t n = stack.back();
        stack.pop_back();
        const float* nd = s.nodes + 32 * size_t(n);
        const uint32_t* nu = s.nodes_u + 32 * size_t(n);
        for (int ch = 0; ch < 4; ++ch) {
            if (nd[ch] == std::numeric_limits<float>::infinity()) continue;
            double tn = t_min * 0.999 - 1e-6, tf = t_max * 1.001 + 1e-6;
            const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
            bool miss = false;
            for (int a = 0; a < 3 && !miss; ++a) {
                const double lo = nd[4 * a + ch], hi = nd[12 + 4 * a + ch];
                const double pad = 1e-5 * (std::fabs(lo) + std::fabs(hi) + 1.0);
                const double l = lo - pad, h = hi + pad;
                if (dd[a] == 0) {
                    if (oo[a] < l || oo[a] > h) miss = true;
                    continue;
                }
                double t0 = (l - oo[a]) / dd[a], t1 = (h - oo[a]) / dd[a];
                if (t0 > t1) std::swap(t0, t1);
                tn = std::max(tn, t0);
                tf = std::min(tf, t1);
                if (tn > tf) miss = true;
            }
            if (miss || tf == -inf) continue;
            if (nu[28 + ch] > 0) {
                for (uint32_t i = nu[24 + ch]; i < nu[24 + ch] + nu[28 + ch]; ++i) test(s.order[i], t_max);
            } else {
                stack.push_back(nu[24 + ch]);
            }
        }
    }
}

// float_ray: move the ray into an instance's object space in float, as the
// reference does (geometry.hpp:213-217), and only then intersect in double;
// otherwise the object-space ray is computed in double too.
void trace64(const scene& s, V of, V df, double t_min, double t_max, double& t_out, uint32_t& kind, uint32_t& prim,
             uint32_t& inst, bool float_ray) {
    const dv o{of.x, of.y, of.z}, d{df.x, df.y, df.z};
    kind = 0;
    prim = 0;
    inst = NONE;
    for (uint32_t k = 0; k < s.n_large; ++k)
        if (sphere64(s.sph + 8 * s.large[k], o, d, t_min, t_max)) {
            kind = 1;
            prim = s.large[k];
        }
    traverse64(s, s.flat_root, o, d, t_min, t_max, [&](uint32_t p, double& tm) {
        if (p < s.n_small) {
            if (sphere64(s.sph + 8 * p, o, d, t_min, tm)) {
                kind = 2;
                prim = p;
                inst = NONE;
            }
        } else if (tri64(s.tri + 28 * size_t(p - s.n_small), o, d, t_min, tm)) {
            kind = 3;
            prim = p - s.n_small;
            inst = NONE;
        }
    });
    traverse64(s, s.inst_root, o, d, t_min, t_max, [&](uint32_t k, double& tm) {
        const float* in = s.inst + 28 * size_t(k);
        const float* m = in + 12;
        dv lo{m[0] * o.x + m[1] * o.y + m[2] * o.z + m[3], m[4] * o.x + m[5] * o.y + m[6] * o.z + m[7],
              m[8] * o.x + m[9] * o.y + m[10] * o.z + m[11]};
        dv ld{m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
              m[8] * d.x + m[9] * d.y + m[10] * d.z};
        if (float_ray) {
            const V fo{m[0] * of.x + m[1] * of.y + m[2] * of.z + m[3], m[4] * of.x + m[5] * of.y + m[6] * of.z + m[7],
                       m[8] * of.x + m[9] * of.y + m[10] * of.z + m[11]};
            const V fd{m[0] * df.x + m[1] * df.y + m[2] * df.z, m[4] * df.x + m[5] * df.y + m[6] * df.z,
                       m[8] * df.x + m[9] * df.y + m[10] * df.z};
            lo = {fo.x, fo.y, fo.z};
            ld = {fd.x, fd.y, fd.z};
        }
        traverse64(s, ub(in[24]), lo, ld, t_min, tm, [&](uint32_t q, double& tq) {
            if (tri64(s.tri + 28 * size_t(q), lo, ld, t_min, tq)) {
                kind = 3;
                prim = q;
                inst = k;
            }
        });
    });
    t_out = t_max;
}

// ---- sampler (sampler.hpp:35-127), with the reference's byte tables ---------

uint32_t reverse_bits(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}
uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
uint32_t lk_permute(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
struct sobol_tables {
    uint32_t t[4][256];
    sobol_tables() {
        uint32_t v[32];
        v[0] = 1u << 31;
        for (int k = 1; k < 32; ++k) v[k] = v[k - 1] ^ (v[k - 1] >> 1);
        for (int byte = 0; byte < 4; ++byte)
            for (uint32_t b = 0; b < 256; ++b) {
                uint32_t r = 0;
                for (int k = 0; k < 8; ++k)
                    if (b & (1u << k)) r ^= v[8 * byte + k];
                t[byte][b] = reverse_bits(r);
            }
    }
};
const sobol_tables& tables() {
    static const sobol_tables t;
    return t;
}
uint32_t sobol_1_reversed(uint32_t i) {
    const auto& t = tables().t;
    return t[0][i & 255] ^ t[1][(i >> 8) & 255] ^ t[2][(i >> 16) & 255] ^ t[3][i >> 24];
}
uint32_t pixel_seed(uint64_t seed, int i, int j) {
    return hash_u32(static_cast<uint32_t>(i) * 0x9e3779b1u ^
                    hash_u32(static_cast<uint32_t>(j) ^ hash_u32(static_cast<uint32_t>(seed) ^ static_cast<uint32_t>(seed >> 32))));
}
void qmc(uint32_t pseed, uint32_t index, int pair, float& a, float& b) {
    const uint32_t seed = hash_u32(pseed + static_cast<uint32_t>(pair) * 0x68bc21ebu);
    const uint32_t i = reverse_bits(lk_permute(reverse_bits(index), seed));
    const uint32_t x = reverse_bits(lk_permute(i, seed ^ 0xa511e9b3u));
    const uint32_t y = reverse_bits(lk_permute(sobol_1_reversed(i), seed ^ 0x63d83595u));
    a = static_cast<float>(x >> 8) * 0x1.0p-24f;
    b = static_cast<float>(y >> 8) * 0x1.0p-24f;
}

// ---- the per-path stream (the GPU path's definition) ------------------------

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
uint64_t path_key(uint64_t seed, int i, int j, uint32_t index) {
    const uint64_t pix = (static_cast<uint64_t>(static_cast<uint32_t>(j)) << 32) | static_cast<uint32_t>(i);
    return mix64(seed + 0x9e3779b97f4a7c15ull) ^ mix64(pix ^ 0x2545f4914f6cdd1dull) ^
           mix64(static_cast<uint64_t>(index) * 0xd1b54a32d192ed03ull + 0x5851f42d4c957f2dull);
}
struct path_rng {
    uint64_t key;
    uint32_t counter = 0;
    float next() {
        const uint64_t z = mix64(key + static_cast<uint64_t>(++counter) * 0x9e3779b97f4a7c15ull);
        return static_cast<float>(static_cast<uint32_t>(z >> 40)) * 0x1.0p-24f;
    }
};

// ---- the integrator ---------------------------------------------------------

constexpr float PI_F = 3.14159265358979323846f;

V reflect(V v, V n) { return v - n * (2 * dot(v, n)); }

struct sample_result {
    V color;
    V albedo, normal;
    uint64_t rays = 0;
};

// ray_color (renderer.hpp:100-149) with direct_light (41-80).
sample_result path(const scene& s, V o, V d, int depth, bool nee, uint32_t pseed, uint32_t index, path_rng& rng) {
    sample_result out;
    V thr{1, 1, 1}, rad{0, 0, 0};
    bool sampled = false;
    counters c;
    out.albedo = {1, 1, 1};
    out.normal = {0, 0, 0};
    auto sample_2d = [&](int pair, float& a, float& b) {
        if (pair > 4) {
            a = rng.next();
            b = rng.next();
        } else {
            qmc(pseed, index, pair, a, b);
        }
    };
    for (int bounce = 0; bounce < depth; ++bounce) {
        ++out.rays;
        hit_out h;
        if (!trace(s, o, d, 0.001f, 1000.0f, false, h, c)) {
            V bg{0, 0, 0};
            if (s.flags & 1u) {
                const V u = unit(d);
                const float t = 0.5f * (u.y + 1);
                bg = V{1, 1, 1} * (1 - t) + V{0.5f, 0.7f, 1.0f} * t;
            }
            out.color = rad + mul(thr, bg);
            return out;
        }
        const V p = o + d * h.t;
        V n;
        bool front;
        float tex_u = 0, tex_v = 0;   // hittable.hpp:14: 0 unless a triangle sets them
        uint32_t m;
        bool sampled_light = false;
        if (h.kind <= 2) {
            const float* sp = s.sph + 8 * h.prim;
            m = ub(sp[4]);
            const V outward = (p - V{sp[0], sp[1], sp[2]}) * (1 / sp[3]);
            front = dot(d, outward) < 0;
            n = front ? outward : -outward;
            if (h.kind == 2 && nee) {
                const float* e = s.mat + 12 * m + 4;
                sampled_light = e[0] > 0 || e[1] > 0 || e[2] > 0;
            }
        } else {
            const float* t = s.tri + 28 * size_t(h.prim);
            m = ub(t[3]);
            const V n0{t[12], t[13], t[14]}, n1{t[16], t[17], t[18]}, n2{t[20], t[21], t[22]}, fn{t[24], t[25], t[26]};
            const V sn = unit(n0 * (1 - h.u - h.v) + n1 * h.u + n2 * h.v);
            // triangle::fill's texture coordinates (triangle.hpp:72-73); the
            // packed uv sit in the w words of rows 1-6 (gpu_pack.hpp:116-129)
            tex_u = t[7] * (1 - h.u - h.v) + t[15] * h.u + t[23] * h.v;
            tex_v = t[11] * (1 - h.u - h.v) + t[19] * h.u + t[27] * h.v;
            if (h.inst == NONE) {
                front = dot(d, fn) < 0;
                n = front ? sn : -sn;
            } else {
                const float* in = s.inst + 28 * size_t(h.inst);
                const float* r = in + 12;
                const V ld{r[0] * d.x + r[1] * d.y + r[2] * d.z, r[4] * d.x + r[5] * d.y + r[6] * d.z,
                           r[8] * d.x + r[9] * d.y + r[10] * d.z};
                front = dot(ld, fn) < 0;
                const V ln = front ? sn : -sn;
                n = unit(V{r[0] * ln.x + r[4] * ln.y + r[8] * ln.z, r[1] * ln.x + r[5] * ln.y + r[9] * ln.z,
                           r[2] * ln.x + r[6] * ln.y + r[10] * ln.z});
                if (ub(in[25]) != NONE) m = ub(in[25]);
            }
        }
        const float* M = s.mat + 12 * m;
        const V albedo{M[0], M[1], M[2]}, emission{M[4], M[5], M[6]};
        const int type = static_cast<int>(M[3]);
        if (front && !(sampled && sampled_light)) rad = rad + mul(thr, emission);
        V att = albedo, dir{0, 0, 0};
        const uint32_t tex = ub(M[9]);
        if (type == 0 && tex != 0) att = mul(att, texture_sample(s, tex - 1, tex_u, tex_v));   // material.hpp:72
        bool ok = true;
        if (type == 0) {
            float a, b;
            sample_2d(1 + 2 * bounce, a, b);
            const float r = std::sqrt(a), phi = 2 * PI_F * b;
            const float sign = std::copysign(1.0f, n.z);
            const float pp = -1 / (sign + n.z), qq = n.x * n.y * pp;
            const V t1{1 + sign * n.x * n.x * pp, sign * qq, -sign * n.x};
            const V t2{qq, sign + n.y * n.y * pp, -n.y};
            dir = t1 * (r * std::cos(phi)) + t2 * (r * std::sin(phi)) + n * std::sqrt(std::fmax(0.0f, 1 - a));
        } else if (type == 1) {
            const V refl = reflect(unit(d), n);
            V ball{0, 0, 0};
            for (int tries = 0; tries < 64; ++tries) {
                const float x = -1 + 2 * rng.next();
                const float y = -1 + 2 * rng.next();
                const float z = -1 + 2 * rng.next();
                ball = {x, y, z};
                if (dot(ball, ball) < 1) break;
            }
            dir = refl + ball * M[7];
            ok = dot(dir, n) > 0;
        } else if (type == 2) {
            att = {1, 1, 1};
            const float ior = M[8];
            const float ratio = front ? (1 / ior) : ior;
            const V ud = unit(d);
            const float cos_t = std::fmin(dot(-ud, n), 1.0f);
            const float sin_t = std::sqrt(1 - cos_t * cos_t);
            const bool cannot = ratio * sin_t > 1;
            float r0 = (1 - ratio) / (1 + ratio);
            r0 = r0 * r0;
            const float x = 1 - cos_t, x2 = x * x;
            const bool refl = cannot || (r0 + (1 - r0) * x2 * x2 * x) > rng.next();
            if (refl) {
                dir = reflect(ud, n);
            } else {
                const float ct = std::fmin(dot(-ud, n), 1.0f);
                const V perp = (ud + n * ct) * ratio;
                const V par = n * -std::sqrt(std::fabs(1 - dot(perp, perp)));
                dir = perp + par;
            }
        } else {
            ok = false;
        }
        if (bounce == 0) {
            out.albedo = ok && type != 2 ? att : V{1, 1, 1};
            out.normal = n;
        }
        if (!ok) {
            out.color = rad;
            return out;
        }
        sampled = type == 0 && nee && s.n_lights > 0 && bounce + 1 < depth;
        if (sampled) {
            const uint32_t nl = s.n_lights;
            const uint32_t pick = std::min(static_cast<uint32_t>(rng.next() * nl), nl - 1);
            const float* L = s.sph + 8 * s.lights[pick];
            const V lc{L[0], L[1], L[2]};
            const V tc = lc - p;
            const float dist2 = dot(tc, tc), r2 = L[3] * L[3];
            if (dist2 > r2) {
                const float sin2 = r2 / dist2, cos_max = std::sqrt(1 - sin2), omc = sin2 / (1 + cos_max);
                const V w = tc * (1 / std::sqrt(dist2));
                const V ax = std::fabs(w.x) > 0.9f ? V{0, 1, 0} : V{1, 0, 0};
                const V uu = unit(cross(w, ax)), vv = cross(w, uu);
                float s1, s2;
                sample_2d(2 + 2 * bounce, s1, s2);
                const float cos_t = 1 - s1 * omc;
                const float sin_t = std::sqrt(std::fmax(0.0f, 1 - cos_t * cos_t));
                const float phi = 2 * PI_F * s2;
                const V ldir = uu * (std::cos(phi) * sin_t) + vv * (std::sin(phi) * sin_t) + w * cos_t;
                const float cs = dot(ldir, n);
                float tl = std::numeric_limits<float>::infinity();
                if (cs > 0 && sphere_hit<float>(L, p, ldir, 0.001f, tl)) {
                    ++out.rays;
                    hit_out sh;
                    const bool occ = trace(s, p, ldir, 0.001f, tl * 0.9999f, true, sh, c);
                    const V lp = p + ldir * tl;
                    const V outward = (lp - lc) * (1 / L[3]);
                    if (!occ && dot(ldir, outward) < 0) {
                        const float pdf = 1 / (2 * PI_F * omc);
                        const float* E = s.mat + 12 * ub(L[4]) + 4;
                        const V contrib = V{E[0], E[1], E[2]} * (cs / PI_F / pdf * static_cast<float>(nl));
                        rad = rad + mul(mul(thr, att), contrib);
                    }
                }
            }
        }
        thr = mul(thr, att);
        o = p;
        d = dir;
        if (bounce + 1 >= 3) {
            const float pm = std::fmax(thr.x, std::fmax(thr.y, thr.z));
            if (pm < 1) {
                if (rng.next() >= pm) {
                    out.color = rad;
                    return out;
                }
                thr = thr * (1 / pm);
            }
        }
    }
    out.color = rad;
    return out;
}

}  // namespace

extern "C" {

// Closest (any = 0) or any (any = 1) hit for n rays of 8 floats (o, tmin, d,
// tmax). counts (optional, 6 entries): nodes visited, leaf references,
// triangle tests, sphere tests, instance visits, large-sphere tests, summed.
void or_trace(const uint32_t* ptgp, const float* rays, size_t n, int any, float* t, float* u, float* v, uint32_t* kind,
              uint32_t* prim, uint32_t* inst, uint64_t* counts, int threads) {
    const scene s(ptgp);
    std::atomic<size_t> next{0};
    std::vector<counters> per(std::max(1, threads));
    auto worker = [&](int id) {
        for (size_t b = next.fetch_add(256); b < n; b = next.fetch_add(256))
            for (size_t i = b; i < std::min(n, b + 256); ++i) {
                const float* q = rays + 8 * i;
                hit_out h;
                const bool f = trace(s, {q[0], q[1], q[2]}, {q[4], q[5], q[6]}, q[3], q[7], any != 0, h, per[id]);
                t[i] = any ? 0 : h.t;
                u[i] = any ? 0 : h.u;
                v[i] = any ? 0 : h.v;
                kind[i] = any ? (f ? 1 : 0) : h.kind;
                prim[i] = any ? 0 : h.prim;
                inst[i] = any ? NONE : h.inst;
            }
    };
    std::vector<std::thread> pool;
    for (int k = 0; k < std::max(1, threads); ++k) pool.emplace_back(worker, k);
    for (auto& th : pool) th.join();
    if (counts) {
        for (int k = 0; k < 6; ++k) counts[k] = 0;
        for (const auto& c : per) {
            counts[0] += c.nodes;
            counts[1] += c.leafrefs;
            counts[2] += c.tris;
            counts[3] += c.spheres;
            counts[4] += c.instances;
            counts[5] += c.large;
        }
    }
}

void or_trace_f64(const uint32_t* ptgp, const float* rays, size_t n, double* t, uint32_t* kind, uint32_t* prim,
                  uint32_t* inst, int threads, int float_ray) {
    const scene s(ptgp);
    std::atomic<size_t> next{0};
    auto worker = [&] {
        for (size_t b = next.fetch_add(256); b < n; b = next.fetch_add(256))
            for (size_t i = b; i < std::min(n, b + 256); ++i) {
                const float* q = rays + 8 * i;
                trace64(s, {q[0], q[1], q[2]}, {q[4], q[5], q[6]}, q[3], q[7], t[i], kind[i], prim[i], inst[i],
                        float_ray != 0);
            }
    };
    std::vector<std::thread> pool;
    for (int k = 0; k < std::max(1, threads); ++k) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
}

// image_texture::sample of texture t of a PTGP buffer's texture section at
// n (u, v) pairs: 3 floats each.
void or_texture_sample(const uint32_t* ptgp, uint32_t t, const float* uv, size_t n, float* out) {
    const scene s(ptgp);
    for (size_t i = 0; i < n; ++i) {
        const V c = texture_sample(s, t, uv[2 * i], uv[2 * i + 1]);
        out[3 * i] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
    }
}

void or_sample_2d(uint64_t seed, int i, int j, uint32_t index, int pair, float* a, float* b) {
    qmc(pixel_seed(seed, i, j), index, pair, *a, *b);
}

uint64_t or_path_key(uint64_t seed, int i, int j, uint32_t index) { return path_key(seed, i, j, index); }

// render_row_impl (renderer.hpp:366-409) over rows [j_lo, j_hi); rgb and aux
// (optional, 7 floats) per pixel, the band's top row first. Returns rays.
uint64_t or_render(const uint32_t* ptgp, int width, int height, int spp, int depth, uint64_t seed, int first_sample,
                   int nee_flag, int j_lo, int j_hi, float* rgb, float* aux, int threads) {
    const scene s(ptgp);
    const bool nee = nee_flag && (s.flags & 2u);
    std::atomic<int> next{j_lo};
    std::atomic<uint64_t> rays{0};
    auto worker = [&] {
        uint64_t local = 0;
        for (int j = next.fetch_add(1); j < j_hi; j = next.fetch_add(1)) {
            const size_t row = size_t(j_hi - 1 - j) * width;
            const float scale = 1.0f / spp;
            for (int i = 0; i < width; ++i) {
                V sum{0, 0, 0}, asum{0, 0, 0}, nsum{0, 0, 0};
                float lum2 = 0;
                const uint32_t ps = pixel_seed(seed, i, j);
                for (int sidx = 0; sidx < spp; ++sidx) {
                    const uint32_t index = static_cast<uint32_t>(first_sample + sidx);
                    float du, dv;
                    qmc(ps, index, 0, du, dv);
                    const float uu = (static_cast<float>(i) + du) / static_cast<float>(width - 1);
                    const float vv = (static_cast<float>(j) + dv) / static_cast<float>(height - 1);
                    const V d = s.cam[1] + s.cam[2] * uu + s.cam[3] * vv - s.cam[0];
                    path_rng rng{path_key(seed, i, j, index)};
                    const sample_result r = path(s, s.cam[0], d, depth, nee, ps, index, rng);
                    local += r.rays;
                    sum = sum + r.color;
                    asum = asum + r.albedo;
                    nsum = nsum + r.normal;
                    const float lum = 0.2126f * r.color.x + 0.7152f * r.color.y + 0.0722f * r.color.z;
                    lum2 += lum * lum;
                }
                float* o = rgb + 3 * (row + i);
                o[0] = sum.x * scale;
                o[1] = sum.y * scale;
                o[2] = sum.z * scale;
                if (aux) {
                    float* a = aux + 7 * (row + i);
                    a[0] = asum.x * scale;
                    a[1] = asum.y * scale;
                    a[2] = asum.z * scale;
                    a[3] = nsum.x * scale;
                    a[4] = nsum.y * scale;
                    a[5] = nsum.z * scale;
                    a[6] = lum2 * scale;
                }
            }
        }
        rays += local;
    };
    std::vector<std::thread> pool;
    for (int k = 0; k < std::max(1, threads); ++k) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
    return rays.load();
}

}  // extern "C"
