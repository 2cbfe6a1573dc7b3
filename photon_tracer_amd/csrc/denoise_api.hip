// The C ABI's denoiser (pt_denoise): the CLI's --denoise step
// (main.cpp:172-186) over the render's own color and aux outputs, on the
// device. Host code only; the kernels are in denoise.hip.
#include <hip/hip_runtime.h>

#include <climits>

#include "api_internal.hpp"
#include "denoise_kernels.hpp"

extern "C" {

pt_status pt_denoise(int width, int height, int spp, const float* rgb, const float* aux, float* out, int flags,
                     double* device_ms) {
    if (width < 1 || height < 1 || spp < 1) return pt::fail(PT_ERR_INVALID, "pt_denoise: width, height and spp must be >= 1");
    if (!rgb || !aux || !out) return pt::fail(PT_ERR_INVALID, "pt_denoise: null image pointer");
    if (static_cast<long long>(width) * height > INT_MAX / 8)
        return pt::fail(PT_ERR_LIMIT, "pt_denoise: image too large");
    if (flags & ~PT_DENOISE_DEVICE) return pt::fail(PT_ERR_INVALID, "pt_denoise: unknown flags");
    if (out == rgb) return pt::fail(PT_ERR_INVALID, "pt_denoise: out must not alias rgb");
    PT_GUARD({
        const size_t n = static_cast<size_t>(width) * height;
        ptapi::dbuf light[2], var[2], normal, albedo, drgb, daux, dout;
        struct releaser {
            std::vector<ptapi::dbuf*> bufs;
            ~releaser() {
                for (auto* b : bufs) b->release();
            }
        } rel{{&light[0], &light[1], &var[0], &var[1], &normal, &albedo, &drgb, &daux, &dout}};
        for (auto& l : light) l.ensure(16 * n);
        for (auto& v : var) v.ensure(4 * n);
        normal.ensure(16 * n);
        albedo.ensure(16 * n);
        const bool device = (flags & PT_DENOISE_DEVICE) != 0;
        const float* in_rgb = rgb;
        const float* in_aux = aux;
        float* to = out;
        if (!device) {
            drgb.ensure(12 * n);
            daux.ensure(28 * n);
            dout.ensure(12 * n);
            HIPC(hipMemcpy(drgb.p, rgb, 12 * n, hipMemcpyHostToDevice));
            HIPC(hipMemcpy(daux.p, aux, 28 * n, hipMemcpyHostToDevice));
            in_rgb = drgb.as<float>();
            in_aux = daux.as<float>();
            to = dout.as<float>();
        }
        const ptd::dn::DenoiseBufs b{{light[0].as<float4>(), light[1].as<float4>()},
                                     {var[0].as<float>(), var[1].as<float>()},
                                     normal.as<float4>(),
                                     albedo.as<float4>()};
        struct events {
            hipEvent_t e[2] = {nullptr, nullptr};
            ~events() {
                for (hipEvent_t x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } ev;
        HIPC(hipEventCreate(&ev.e[0]));
        HIPC(hipEventCreate(&ev.e[1]));
        HIPC(hipEventRecord(ev.e[0], nullptr));
        HIPC(ptd::dn::launch_denoise(width, height, spp, in_rgb, in_aux, to, b, nullptr));
        HIPC(hipEventRecord(ev.e[1], nullptr));
        HIPC(hipEventSynchronize(ev.e[1]));
        float ms = 0;
        HIPC(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
        if (device_ms) *device_ms = ms;
        if (!device) HIPC(hipMemcpy(out, dout.p, 12 * n, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
