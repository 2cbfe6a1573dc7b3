// Test infrastructure, not product code: the reference-side binding
// (include/photon_tracer_render.hpp) compiled against the reference's own
// headers from where they lie (oracle/Makefile; output into oracle/_ref/
// only) and linked with libphoton_tracer_amd.so, as a maintainer would build
// it next to renderer.hpp.
//
//   render_shim_check <scene> <width> <height> <spp> <seed> <out.bin>
//
// builds the reference scene (build_world, renderer.hpp:335-343: 0 materials,
// 1 field, 3 room; make_camera, renderer.hpp:212-218), renders it through the
// header twice -- every row j by its own render_row_linear call, placed top
// row first as the CLI does (main.cpp:148-151), then the whole frame by
// render_frame -- and writes both images (float32, rows top first) to
// out.bin. Prints "rays <row by row> <frame>" from rays_traced. Exit 2 with
// the library's message when a call throws (no GPU, say).
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "gpu_pack.hpp"
#include "renderer.hpp"
#include "photon_tracer_render.hpp"

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s scene width height spp seed out.bin\n", argv[0]);
        return 1;
    }
    const auto id = static_cast<scene_id>(std::atoi(argv[1]));
    const int width = std::atoi(argv[2]), height = std::atoi(argv[3]), spp = std::atoi(argv[4]);
    const uint64_t seed = std::strtoull(argv[5], nullptr, 10);
    try {
        const geometry world = build_world(id, true);
        const camera cam = make_camera(id);
        const photon_tracer_gpu::world g(world, cam);
        const size_t row = 3 * static_cast<size_t>(width);
        std::vector<float> rows(row * height), frame(row * height);
        rays_traced = 0;
        for (int j = 0; j < height; ++j)
            render_row_linear(g, j, width, height, spp, 10, &rows[row * (height - 1 - j)], seed);
        const long long by_rows = rays_traced;
        rays_traced = 0;
        photon_tracer_gpu::render_frame(g, width, height, spp, 10, frame.data(), seed);
        std::FILE* f = std::fopen(argv[6], "wb");
        if (!f) throw std::runtime_error("cannot write the output");
        std::fwrite(rows.data(), 4, rows.size(), f);
        std::fwrite(frame.data(), 4, frame.size(), f);
        std::fclose(f);
        std::printf("rays %lld %lld\n", by_rows, static_cast<long long>(rays_traced));
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
}
