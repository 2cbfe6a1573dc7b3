// Host-side scene layer: the PTSD scene description, the binned-SAH 4-wide
// BVH builder and the PTGP v1 packer.
//
// Floating point here must round exactly as the reference's host code does
// (g++ -O3 on x86-64, SSE scalar, no FMA contraction), because the packed
// boxes, edges and normals are compared byte for byte with the reference's
// pack_for_gpu. This file is compiled with -ffp-contract=off.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace pt {

// ---------------------------------------------------------------------------
// Small float vector with the reference's operation order (vec3.hpp:58-118).
// ---------------------------------------------------------------------------
struct v3 {
    float x = 0, y = 0, z = 0;
    v3() = default;
    v3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
    float& operator[](int i) { return i == 0 ? x : i == 1 ? y : z; }
};
inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
inline v3 operator*(v3 a, float t) { return {a.x * t, a.y * t, a.z * t}; }
inline v3 mul(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline v3 cross(v3 a, v3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline v3 normalize(v3 a) {
    const float len = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / len, a.y / len, a.z / len};
}

struct box3 {
    v3 lo{inf(), inf(), inf()}, hi{-inf(), -inf(), -inf()};
    static float inf() { return std::numeric_limits<float>::infinity(); }
    // aabb::grow (aabb.hpp:15-20): std::min(a, b) returns a unless b < a.
    void grow(const box3& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = b.lo[a] < lo[a] ? b.lo[a] : lo[a];
            hi[a] = hi[a] < b.hi[a] ? b.hi[a] : hi[a];
        }
    }
    v3 center() const { return (lo + hi) * 0.5f; }
    // aabb::half_area (aabb.hpp:24-28)
    float half_area() const {
        if (lo.x > hi.x) return 0;
        const v3 d = hi - lo;
        return d.x * d.y + d.y * d.z + d.z * d.x;
    }
};

// ---------------------------------------------------------------------------
// PTSD: this library's scene description, the reference geometry's flat
// arrays (geometry.hpp:15-58) in one word buffer.
//
//   word 0 magic 'PTSD' (0x44535450), 1 version (1), 2 flags (1 sky, 2 NEE),
//   3-5 lookfrom, 6-8 lookat, 9 vfov (degrees, float), 10 n_materials,
//   11 n_spheres, 12 n_large, 13 n_triangles, 14 n_meshes, 15 n_instances,
//   16 mesh triangles in total, 17 n_textures, 18-31 reserved (0); then the arrays:
//   materials  12 words: albedo(3) type emission(3) fuzz ior texture 0 0
//              (texture: 0 none, else texture number + 1; matte only)
//   spheres     5 words: center(3) radius material
//   large       5 words: as spheres
//   triangles  28 words: p0 p1 p2 n0 n1 n2 (18) uv(6) material 0 0 0
//   mesh sizes  n_meshes words, then each mesh's triangles (28 words each)
//   instances  16 words: mesh material(~0u: none) m[3][3] (row major) t(3) 0 0
//   textures   width height, then width x height x 3 bytes of sRGB (rows top
//              first, as image8 holds them, image_io.hpp), padded to a word
// Material types follow gpu_pack.hpp:93-106: 0 lambertian, 1 metal,
// 2 dielectric, 3 light.
// ---------------------------------------------------------------------------
constexpr uint32_t kPtsdMagic = 0x44535450u;
constexpr uint32_t kPtgpMagic = 0x50475450u;
constexpr int kPtsdHeader = 32;
constexpr int kMatWords = 12, kSphWords = 5, kTriWords = 28, kInstWords = 16;

struct material_rec {
    v3 albedo;
    uint32_t type = 0;
    v3 emission;
    float fuzz = 0, ior = 1.5f;
    uint32_t texture = 0;   // matte: 0 none, else texture number + 1 (albedo x texture, material.hpp:72)
    // Only lights carry emission (material.hpp:97-102); material::emits (material.hpp:29).
    bool is_light() const { return type == 3 && (emission.x > 0 || emission.y > 0 || emission.z > 0); }
};
struct sphere_rec {
    v3 center;
    float radius = 0;
    uint32_t material = 0;
};
struct triangle_rec {
    v3 p[3];
    v3 n[3];
    float uv[6] = {0, 0, 0, 0, 0, 0};
    uint32_t material = 0;
};
struct affine_rec {
    float m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    v3 t;
    // transform.hpp:17-21
    v3 vector(v3 v) const {
        return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
    }
    v3 point(v3 p) const { return vector(p) + t; }
    affine_rec inverse() const;   // transform.hpp:42-59, cofactors in double
    box3 box(const box3& b) const;  // transform.hpp:62-70
};
struct instance_rec {
    uint32_t mesh = 0;
    uint32_t material = ~0u;
    affine_rec to_world;
};

// An 8-bit sRGB image (image8, image_io.hpp): rows top first, 3 bytes a
// texel. `name` is only used to match OBJ/MTL map_Kd names.
struct image_rec {
    std::string name;
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgb;
};

// decode_image (image_io.hpp:295): a PNG or PPM file's bytes, told apart by
// their first bytes, as 8-bit sRGB (image.cpp). Throws on what it cannot read.
image_rec decode_image(const uint8_t* bytes, size_t size);

struct scene_desc {
    bool sky = true, sample_lights = true;
    v3 lookfrom{0, 0, 0}, lookat{0, 0, -1};
    float vfov = 90;
    std::vector<material_rec> materials;
    std::vector<sphere_rec> spheres, large;
    std::vector<triangle_rec> triangles;
    std::vector<std::vector<triangle_rec>> meshes;
    std::vector<instance_rec> instances;
    std::vector<image_rec> textures;

    uint32_t add_material(const material_rec& m) {
        materials.push_back(m);
        return static_cast<uint32_t>(materials.size() - 1);
    }
};

std::vector<uint32_t> ptsd_write(const scene_desc& s);
scene_desc ptsd_read(const uint32_t* w, size_t n);

// The camera rows of the reference's look-at camera (camera.hpp:18-34):
// origin, lower-left corner, across, up; 4 floats each (w = 0).
void camera_rows(v3 lookfrom, v3 lookat, double vfov_deg, float out[16]);

// ---------------------------------------------------------------------------
// 4-wide BVH in the reference's node layout (bvh.hpp:27-32).
// ---------------------------------------------------------------------------
struct alignas(16) bvh_node {
    float lo[3][4];
    float hi[3][4];
    uint32_t first[4];
    uint32_t count[4];
};
static_assert(sizeof(bvh_node) == 128, "node layout");

struct bvh_tree {
    std::vector<bvh_node> nodes;
    std::vector<uint32_t> order;
};

// Binned SAH build then 4-wide collapse (bvh.hpp:39-52,126-247).
bvh_tree build_bvh(const std::vector<box3>& boxes);

// Levels of 4-wide nodes below and including node n (gpu_pack.hpp:41-47).
int tree_depth(const std::vector<bvh_node>& nodes, uint32_t n);

// Primitive boxes (sphere.hpp:60-64, triangle.hpp:76-83).
box3 sphere_box(const sphere_rec& s);
box3 triangle_box(const triangle_rec& t);

// PTSD -> PTGP v1 words (geometry::build + pack_for_gpu + serialize). With
// a builder, every tree is built by it (one call at a time) instead of
// build_bvh on host threads; it must return build_bvh's tree.
using bvh_builder = std::function<bvh_tree(const std::vector<box3>&)>;
std::vector<uint32_t> pack_ptgp(const scene_desc& s, const bvh_builder* builder = nullptr);

// load_obj (obj.hpp:114-218): the OBJ text's faces appended to s.triangles,
// with materials from the MTL text (standing in for every library the OBJ
// names; appended to s.materials) or `material` (obj.cpp). A matte
// material's map_Kd that names one of `images` gets that image as a texture
// (appended to s.textures).
void add_obj(scene_desc& s, const std::string& text, v3 base, double height, uint32_t material, double turn_deg,
             const std::string* mtl_text = nullptr, const std::vector<image_rec>* images = nullptr);

// A malloc'd copy of v for the C ABI (free with pt_free); *n = v.size().
uint32_t* copy_out(const std::vector<uint32_t>& v, size_t* n);

// Batched renders (merge.cpp): several flat PTGP v1 scenes (no instances) as
// one, every array concatenated with its indices rebased and a few 4-wide
// nodes over the scenes' roots as the merged flat root, so the result is an
// ordinary PTGP v1 buffer. parts: where each scene's tree root, large spheres
// and lights went, with its header flags and camera rows.
struct ptgp_in {
    const uint32_t* w;
    size_t n;
};
struct merge_part {
    uint32_t root, n_large, large_off, n_lights, lights_off, flags;
    float camera[16];
};
std::vector<uint32_t> merge_ptgp(const std::vector<ptgp_in>& in, std::vector<merge_part>& parts);

// Scene generators for the benchmark configurations (scenegen.cpp).
scene_desc generate_config(int config, uint64_t seed, double scale);
struct job_desc {
    int kind, width, height, spp;
    uint32_t primitives;
    uint64_t seed;
};
scene_desc generate_job(uint64_t seed, double scale, job_desc& desc);

}  // namespace pt
