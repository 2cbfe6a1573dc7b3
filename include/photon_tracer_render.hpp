/* photon_tracer_render.hpp -- the reference renderer's row call over this
 * library's GPU path: a drop-in for render_row_linear (renderer.hpp:411-416)
 * where the reference's geometry and camera are at hand.
 *
 * Include it after the reference's renderer.hpp and gpu_pack.hpp (it uses
 * geometry, camera, pack_for_gpu, serialize, aux_floats and the thread-local
 * rays_traced counter); link -lphoton_tracer_amd.
 *
 *   const photon_tracer_gpu::world g(world, cam);          // packs once (gpu_pack.hpp:84-203)
 *   render_row_linear(g, j, width, height, spp, max_bounces, row, seed);   // found by argument lookup
 *   photon_tracer_gpu::render_frame(g, width, height, spp, max_bounces, image, seed);
 *
 * render_row_linear keeps the reference's signature and meaning with the
 * world/camera pair replaced by the packed scene: row j counted from the
 * bottom, width linear RGB floats (and aux_floats a pixel of aux when aux is
 * not null), the average of spp samples numbered from first_sample, the rays
 * it traced added to this thread's rays_traced. Every (seed, pixel, sample)
 * draws its own numbers, so a frame comes out the same however its rows are
 * called and render_frame gives those rows in one launch sequence, top row
 * first as the CLI lays the image out (main.cpp:148-151). Errors throw
 * std::runtime_error with the library's message, as the reference throws. */
#ifndef PHOTON_TRACER_RENDER_HPP
#define PHOTON_TRACER_RENDER_HPP

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "photon_tracer.h"

namespace photon_tracer_gpu {

inline void check(pt_status s) {
    if (s != PT_OK) throw std::runtime_error(pt_last_error());
}

/* The device copy of one geometry and camera: the reference's own PTGP
 * (serialize(pack_for_gpu(world, cam))), so its trees are limited to the
 * reference packer's depth (gpu_pack.hpp:36,61-62); deeper scenes go through
 * PTSD and pt_pack. NEE follows world.sample_lights. */
class world {
public:
    world(const geometry& g, const camera& cam) : nee_(g.sample_lights) {
        const std::vector<uint32_t> words = serialize(pack_for_gpu(g, cam));
        check(pt_scene_create(words.data(), words.size(), &scene_));
    }
    ~world() { pt_scene_destroy(scene_); }
    world(const world&) = delete;
    world& operator=(const world&) = delete;

    /* Rows [row_begin, row_end) counted from the bottom, written top row first. */
    void render_rows(int row_begin, int row_end, int width, int height, int spp, int max_bounces, float* out,
                     uint64_t seed, int first_sample, float* aux) const {
        pt_render_params p{};
        p.width = width;
        p.height = height;
        p.spp = spp;
        p.max_bounces = max_bounces;
        p.seed = seed;
        p.first_sample = first_sample;
        p.row_begin = row_begin;
        p.row_end = row_end;
        p.flags = nee_ ? 0 : PT_RENDER_NO_NEE;
        pt_render_stats st{};
        check(pt_render(scene_, &p, out, aux, &st));
        rays_traced += static_cast<long long>(st.rays);
    }
    pt_scene* scene() const { return scene_; }

private:
    pt_scene* scene_ = nullptr;
    bool nee_;
};

/* render_row_linear (renderer.hpp:411-416) for one row j on the GPU. */
inline void render_row_linear(const world& g, int j, int width, int height, int samples_per_pixel, int max_bounces,
                              float* out, uint64_t seed, int first_sample = 0, float* aux = nullptr) {
    g.render_rows(j, j + 1, width, height, samples_per_pixel, max_bounces, out, seed, first_sample, aux);
}

/* Every row of the frame in one call: the rows render_row_linear gives, top
 * row first (out: width * height * 3 floats; aux: width * height * aux_floats). */
inline void render_frame(const world& g, int width, int height, int samples_per_pixel, int max_bounces, float* out,
                         uint64_t seed, int first_sample = 0, float* aux = nullptr) {
    g.render_rows(0, height, width, height, samples_per_pixel, max_bounces, out, seed, first_sample, aux);
}

}  // namespace photon_tracer_gpu

#endif
