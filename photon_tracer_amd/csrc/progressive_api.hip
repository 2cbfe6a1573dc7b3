// The C ABI's progressive driver (pt_progressive_*): the browser demo's
// render loop (gpu.js:199-231) over pt_render's first_sample continuation
// (renderer.hpp:345-356). Host code only.
#include <hip/hip_runtime.h>

#include "api_internal.hpp"

struct pt_progressive {
    pt_scene* scene = nullptr;
    pt_render_params params{};
    int rows = 0, done = 0, batch = 1;
    dbuf step_img, sum, out_img;
};

namespace {

// gpu.js:225-226: a step under 20 ms doubles the next (up to 64 samples), one
// over 60 ms halves it.
constexpr double kGrowBelowMs = 20, kShrinkAboveMs = 60;
constexpr int kMaxStep = 64;

}  // namespace

extern "C" {

pt_status pt_progressive_create(pt_scene* scene, const pt_render_params* params, pt_progressive** out) {
    if (!out) return pt::fail(PT_ERR_INVALID, "pt_progressive_create: null output");
    *out = nullptr;
    if (!scene || !params) return pt::fail(PT_ERR_INVALID, "pt_progressive_create: null scene or params");
    const pt_render_params& P = *params;
    if (P.width < 2 || P.height < 2 || P.spp < 1 || P.first_sample < 0)
        return pt::fail(PT_ERR_INVALID, "pt_progressive_create: width and height must be >= 2, spp >= 1");
    if (P.flags & (PT_RENDER_TIMING | PT_RENDER_COUNT | PT_RENDER_TRAIN))
        return pt::fail(PT_ERR_INVALID, "pt_progressive_create: timing, counting and training renders are single calls");
    const int j0 = (P.row_begin == 0 && P.row_end == 0) ? 0 : P.row_begin;
    const int j1 = (P.row_begin == 0 && P.row_end == 0) ? P.height : P.row_end;
    if (j0 < 0 || j1 > P.height || j0 >= j1) return pt::fail(PT_ERR_INVALID, "pt_progressive_create: bad row band");
    PT_GUARD({
        auto p = std::make_unique<pt_progressive>();
        p->scene = scene;
        p->params = P;
        p->rows = j1 - j0;
        const size_t n = static_cast<size_t>(P.width) * p->rows;
        p->step_img.ensure(12 * n);
        p->sum.ensure(12 * n);
        if (!(P.flags & PT_RENDER_DEVICE_OUTPUT)) p->out_img.ensure(12 * n);
        *out = p.release();
    });
}

pt_status pt_progressive_step(pt_progressive* p, float* rgb, int* done, int* step_spp, pt_render_stats* stats) {
    if (!p) return pt::fail(PT_ERR_INVALID, "pt_progressive_step: null handle");
    if (p->done >= p->params.spp) return pt::fail(PT_ERR_INVALID, "pt_progressive_step: the render is complete");
    pt_render_params step = p->params;
    const int n = std::min(p->batch, p->params.spp - p->done);
    step.spp = n;
    step.first_sample = p->params.first_sample + p->done;
    step.flags = p->params.flags | PT_RENDER_DEVICE_OUTPUT;
    step.samples_per_batch = 0;
    pt_render_stats st{};
    const pt_status r = pt_render(p->scene, &step, p->step_img.as<float>(), nullptr, &st);
    if (r != PT_OK) return r;   // pt_render has set the message
    PT_GUARD({
        const uint32_t npix = static_cast<uint32_t>(p->params.width) * static_cast<uint32_t>(p->rows);
        hipStream_t s = p->scene->stream;
        HIPC(ptd::launch_progressive_add(p->step_img.as<float>(), p->sum.as<float>(), npix, n, p->done == 0, s));
        p->done += n;
        if (st.render_ms < kGrowBelowMs && p->batch < kMaxStep) p->batch *= 2;
        else if (st.render_ms > kShrinkAboveMs && p->batch > 1) p->batch = std::max(1, p->batch >> 1);
        if (rgb) {
            const bool device = (p->params.flags & PT_RENDER_DEVICE_OUTPUT) != 0;
            float* to = device ? rgb : p->out_img.as<float>();
            HIPC(ptd::launch_progressive_resolve(p->sum.as<float>(), to, npix, p->done, s));
            if (!device) HIPC(hipMemcpyAsync(rgb, to, 12 * size_t(npix), hipMemcpyDeviceToHost, s));
        }
        HIPC(hipStreamSynchronize(s));
        if (done) *done = p->done;
        if (step_spp) *step_spp = n;
        if (stats) *stats = st;
    });
}

pt_status pt_progressive_destroy(pt_progressive* p) {
    if (!p) return PT_OK;
    PT_GUARD({
        p->step_img.release();
        p->sum.release();
        p->out_img.release();
        delete p;
    });
}

}  // extern "C"
