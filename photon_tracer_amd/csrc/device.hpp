// Device-side scene view, intersection routines, sampler and RNG for gfx950.
//
// Everything on the intersection path is written to round exactly as the
// reference's CPU code does: the translation units are compiled with
// -ffp-contract=off (no FMA contraction), float division and sqrt are the
// correctly rounded ones, large spheres are intersected in double, and every
// expression keeps the reference's operand order. That makes a hit's t, u, v
// and primitive the reference's own, bit for bit, on the same ray.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptd {

constexpr uint32_t NONE = 0xffffffffu;
constexpr float T_MIN = 0.001f;   // renderer.hpp:114
constexpr float T_MAX = 1000.0f;
constexpr float PI_F = 3.14159265358979323846f;  // real(M_PI)

// PTGP v1 as read on the device (gpu_pack.hpp:21-33).
struct Node {
    float4 lo[3];   // lo[axis] = the four children's lower bounds on that axis
    float4 hi[3];
    uint4 first;
    uint4 count;
};

struct SceneView {
    const Node* nodes;
    const uint32_t* order;
    const float4* leaf;      // 3 float4 per order slot: the primitive's intersection data
    const float4* spheres;   // 2 float4 per sphere: (center, radius), (material bits, 0, 0, 0)
    const float4* tris;      // 7 float4 per triangle (gpu_pack.hpp:118-129)
    const float4* inst;      // 7 float4 per instance: to_world rows, to_object rows, (root, material, 0, 0)
    const float4* mats;      // 3 float4 per material: (albedo, type), (emission, fuzz), (ior, 0, 0, 0)
    const uint32_t* lights;  // sphere numbers
    const uint32_t* large;   // sphere numbers
    uint32_t flat_root, inst_root, n_small, flags;  // flags: 1 sky, 2 sample lights
    uint32_t n_lights, n_large;
    // Nodes [0, n_cached) are also in the launch's LDS node cache, at LDS
    // byte address `cache` + 144 * node (0: no cache). The device copy of a
    // scene numbers the top levels of its trees first, breadth first, so the
    // nodes every ray visits are the cached ones.
    uint32_t n_cached, cache;
    uint32_t n_top;          // nodes the host numbered first (at most TOP_NODES), available for the cache
    uint32_t n_tex;          // textures (PTGP texture section; 0: none)
    uint32_t pad[2];
    const uint4* tex_info;   // per texture: width, height, first texel, 0
    const float4* texels;    // linear RGB (a = 0), rows top first
    float4 cam[4];           // origin, corner, across, up
};
constexpr uint32_t NODE_CACHE_STRIDE = 144;   // bytes per cached node: 7 rows of 16, padded against bank conflicts
constexpr uint32_t TOP_NODES = 64;            // nodes a scene's device copy numbers first (the cache's upper bound)

struct Hit {
    float t, u, v;
    uint32_t kind;   // 0 miss, 1 large sphere, 2 sphere, 3 triangle
    uint32_t prim;   // PTGP sphere or triangle number
    uint32_t inst;   // instance number, or NONE
};

// ---------------------------------------------------------------------------
// Small float3 helpers with the reference's operation order (vec3.hpp).
// ---------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float t) { return {a.x * t, a.y * t, a.z * t}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 normalize(f3 a) {
    const float len = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / len, a.y / len, a.z / len};
}

// 1/x rounded to nearest, as the IEEE division 1.0f / x gives it. The
// hardware reciprocal refined by one Newton step with FMA is the correctly
// rounded reciprocal for 2^-126 <= |x| < 2^126: checked against the division
// for every mantissa and sign at every exponent of that range on gfx950 (the
// only differences are denormal inputs and results). Zeros, infinities, NaN,
// denormals and |x| >= 2^126 take the division.
__device__ __forceinline__ float rcp_rn(float x) {
    const float ax = fabsf(x);
    if (__builtin_expect(ax >= 0x1p-126f && ax < 0x1p126f, 1)) {
        const float r = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
    }
    return 1.0f / x;
}

// ---------------------------------------------------------------------------
// Primitive tests.
// ---------------------------------------------------------------------------

// sphere::intersect<float> (sphere.hpp:25-48): on a hit in [t_min, t_max],
// moves t_max to it.
__device__ __forceinline__ bool sphere_hit_f(float4 cr, f3 o, f3 d, float t_min, float& t_max) {
    const float ox = o.x - cr.x, oy = o.y - cr.y, oz = o.z - cr.z;
    const float a = d.x * d.x + d.y * d.y + d.z * d.z;
    const float half_b = ox * d.x + oy * d.y + oz * d.z;
    const float c = ox * ox + oy * oy + oz * oz - cr.w * cr.w;
    const float disc = half_b * half_b - a * c;
    if (disc < 0) return false;
    const float sq = sqrtf(disc);
    float root = (-half_b - sq) / a;
    if (root < t_min || root > t_max) {
        root = (-half_b + sq) / a;
        if (root < t_min || root > t_max) return false;
    }
    t_max = root;
    return true;
}

// sphere::intersect<double> (sphere.hpp:25-48), used for the large spheres
// (geometry.hpp:117-118,158-160).
__device__ __forceinline__ bool sphere_hit_d(float4 cr, f3 o, f3 d, float t_min, float& t_max) {
    const double ox = double(o.x) - cr.x, oy = double(o.y) - cr.y, oz = double(o.z) - cr.z;
    const double dx = d.x, dy = d.y, dz = d.z;
    const double a = dx * dx + dy * dy + dz * dz;
    const double half_b = ox * dx + oy * dy + oz * dz;
    const double c = ox * ox + oy * oy + oz * oz - double(cr.w) * cr.w;
    // Outside the sphere and moving away: both roots are at most 0, below
    // any t_min > 0. In IEEE arithmetic disc <= fl(half_b^2) here and
    // sqrt(fl(half_b^2)) = half_b, so the full test would reject both roots
    // too; this skips its square root and divisions (the sky half of a
    // ground sphere's queries).
    if (half_b >= 0 && c > 0 && t_min > 0) return false;
    const double disc = half_b * half_b - a * c;
    if (disc < 0) return false;
    const double sq = sqrt(disc);
    double root = (-half_b - sq) / a;
    if (root < t_min || root > t_max) {
        root = (-half_b + sq) / a;
        if (root < t_min || root > t_max) return false;
    }
    t_max = static_cast<float>(root);
    return true;
}

// triangle::intersect (triangle.hpp:40-61): Moller-Trumbore.
__device__ __forceinline__ bool tri_hit(f3 p0, f3 e1, f3 e2, f3 o, f3 d, float t_min, float& t_max, float& u_out,
                                        float& v_out) {
    const f3 pvec = cross(d, e2);
    const float det = dot(e1, pvec);
    if (fabsf(det) < 1e-12f) return false;
    const float inv_det = rcp_rn(det);
    const f3 tvec = o - p0;
    const float u = dot(tvec, pvec) * inv_det;
    if (u < 0 || u > 1) return false;
    const f3 qvec = cross(tvec, e1);
    const float v = dot(d, qvec) * inv_det;
    if (v < 0 || u + v > 1) return false;
    const float t = dot(e2, qvec) * inv_det;
    if (t < t_min || t > t_max) return false;
    t_max = t;
    u_out = u;
    v_out = v;
    return true;
}

// image_texture::sample (texture.hpp:22-33): bilinear, repeating outside
// [0, 1), v running up the image whose rows are stored top first; the
// reference's float operations in its order (no FMA contraction).
__device__ __forceinline__ f3 texture_sample(const SceneView& sc, uint32_t t, float u, float v) {
    const uint4 info = sc.tex_info[t];
    const int w = static_cast<int>(info.x), h = static_cast<int>(info.y);
    const float4* tx = sc.texels + info.z;
    const float x = (u - floorf(u)) * static_cast<float>(w) - 0.5f;
    const float y = (1 - (v - floorf(v))) * static_cast<float>(h) - 0.5f;
    // Finite coordinates give x in [-0.5, w - 0.5], y in [-0.5, h - 0.5]; a
    // NaN or infinite one reads texel 0 (its weights are NaN, as on the CPU).
    auto cell = [](float c) { const float f = floorf(c); return f >= -1.0f && f <= 65536.0f ? static_cast<int>(f) : 0; };
    const int x0 = cell(x), y0 = cell(y);
    const float fx = x - static_cast<float>(x0), fy = y - static_cast<float>(y0);
    auto texel = [&](int xi, int yi) {
        xi = ((xi % w) + w) % w;
        yi = ((yi % h) + h) % h;
        const float4 c = tx[static_cast<size_t>(yi) * w + xi];
        return f3{c.x, c.y, c.z};
    };
    return texel(x0, y0) * ((1 - fx) * (1 - fy)) + texel(x0 + 1, y0) * (fx * (1 - fy)) +
           texel(x0, y0 + 1) * ((1 - fx) * fy) + texel(x0 + 1, y0 + 1) * (fx * fy);
}

// ---------------------------------------------------------------------------
// Sampler: the reference's shuffled Owen-scrambled Sobol pairs
// (sampler.hpp:35-127), a pure function of (pixel seed, index, pair).
// ---------------------------------------------------------------------------
constexpr int QMC_BOUNCES = 2;  // sampler.hpp:24

__device__ __forceinline__ uint32_t reverse_bits(uint32_t x) { return __builtin_bitreverse32(x); }

__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint32_t lk_permute(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

// The second Sobol dimension of index i, bit-reversed: XOR of the direction
// numbers v_k (v_0 = 1 << 31, v_{k+1} = v_k ^ (v_k >> 1)) of i's set bits,
// then reversed (sampler.hpp:63-94). v_k is row k of Pascal's triangle mod 2,
// high bit first, so bit j of the reversed result is the XOR of i's bits k
// with C(k, j) odd, which by Lucas' theorem are the k whose binary digits
// include j's: a superset-XOR transform over the 5-bit bit positions, one
// shift, mask and XOR per position bit (both sides are linear over GF(2) and
// agree on the 32 unit vectors).
__device__ __forceinline__ uint32_t sobol_1_reversed(uint32_t x) {
    x ^= (x >> 1) & 0x55555555u;
    x ^= (x >> 2) & 0x33333333u;
    x ^= (x >> 4) & 0x0f0f0f0fu;
    x ^= (x >> 8) & 0x00ff00ffu;
    x ^= (x >> 16) & 0x0000ffffu;
    return x;
}

// begin_sample (sampler.hpp:98-104)
__device__ __forceinline__ uint32_t pixel_seed(uint64_t seed, int i, int j) {
    return hash_u32(static_cast<uint32_t>(i) * 0x9e3779b1u ^
                    hash_u32(static_cast<uint32_t>(j) ^ hash_u32(static_cast<uint32_t>(seed) ^ static_cast<uint32_t>(seed >> 32))));
}

// sample_2d for a QMC pair (pair <= 2 * QMC_BOUNCES), sampler.hpp:114-123.
__device__ __forceinline__ void qmc_pair(uint32_t pseed, uint32_t index, int pair, float& a, float& b) {
    const uint32_t seed = hash_u32(pseed + static_cast<uint32_t>(pair) * 0x68bc21ebu);
    const uint32_t i = reverse_bits(lk_permute(reverse_bits(index), seed));
    const uint32_t x = reverse_bits(lk_permute(i, seed ^ 0xa511e9b3u));
    const uint32_t y = reverse_bits(lk_permute(sobol_1_reversed(i), seed ^ 0x63d83595u));
    a = static_cast<float>(x >> 8) * 0x1.0p-24f;
    b = static_cast<float>(y >> 8) * 0x1.0p-24f;
}

// ---------------------------------------------------------------------------
// Per-path random stream. The reference draws from one xoshiro256+ stream
// per thread and row (vec3.hpp:11-52, renderer.hpp:349-351), which a
// wavefront cannot reproduce; here every path has its own counter-based
// stream keyed by (seed, pixel, sample): splitmix64 of key + counter. Floats
// take the top 24 bits, as random_real does (vec3.hpp:50).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t path_key(uint64_t seed, int i, int j, uint32_t index) {
    const uint64_t pix = (static_cast<uint64_t>(static_cast<uint32_t>(j)) << 32) | static_cast<uint32_t>(i);
    return mix64(seed + 0x9e3779b97f4a7c15ull) ^ mix64(pix ^ 0x2545f4914f6cdd1dull) ^
           mix64(static_cast<uint64_t>(index) * 0xd1b54a32d192ed03ull + 0x5851f42d4c957f2dull);
}

__device__ __forceinline__ float rand01(uint64_t key, uint32_t& counter) {
    const uint64_t z = mix64(key + static_cast<uint64_t>(++counter) * 0x9e3779b97f4a7c15ull);
    return static_cast<float>(static_cast<uint32_t>(z >> 40)) * 0x1.0p-24f;
}

}  // namespace ptd
