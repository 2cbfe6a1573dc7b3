// Launch interface of the learned guide between the C ABI (api.hip) and its
// kernels (guide.hip), and what the two share: the parameter buffers, the
// Adam settings and the LDS layout of the training pass.
#pragma once

#include <hip/hip_runtime.h>

#include "guide.hpp"

namespace ptd {
namespace guide {

// Device buffers of a guide's parameters.
struct GuideParamsDev {
    float* p32;        // FP32 masters
    float* m;          // Adam first moments
    float* v;          // Adam second moments
    _Float16* w16;     // FP16 mirror, [out][in]
    _Float16* wt16;    // FP16 mirror, [in][out]
};

struct AdamConfig {
    float lr, beta1, beta2, eps;
    float bias1, bias2;   // 1 - beta1^t, 1 - beta2^t for this step t
};

// The backward pass carries d loss / d output * LOSS_SCALE in FP16,
// saturated at +-60000; the FP32 weight gradients are scaled back.
constexpr float LOSS_SCALE = 1.0f;

__device__ __forceinline__ _Float16 to_half_sat(float v) {
    return static_cast<_Float16>(fminf(fmaxf(v, -60000.0f), 60000.0f));
}

// LDS of one training wave: the five layer inputs row-major and transposed,
// two ping-pong gradient tiles (row-major and transposed), the outputs.
struct TrainLds {
    _Float16 x[5][64 * XS];
    _Float16 xt[5][64 * XS];
    _Float16 dy[2][64 * XS];
    _Float16 dyt[2][64 * XS];
    float o[64 * OS];
};

// Parameter j's FP16 copies, when it is a weight (biases stay FP32).
__device__ __forceinline__ void mirror_one(const GuideParamsDev& pd, uint32_t j, float p) {
    const int l = j < static_cast<uint32_t>(OUT_W) ? static_cast<int>(j / LAYER_STRIDE) : 4;
    const int off = static_cast<int>(j) - w_off(l);
    const int nout = n_out(l);
    if (off >= nout * IN) return;
    const int r = off / IN, c = off % IN;
    const _Float16 h = static_cast<_Float16>(p);
    pd.w16[j] = h;
    pd.wt16[w_off(l) + c * nout + r] = h;
}

// Host-side launchers (guide.hip). Queries and samples are device pointers.
hipError_t launch_guide_eval(const GuideView& g, const float4* xs, const float4* ns, uint32_t count, float* out,
                             hipStream_t stream);
hipError_t launch_guide_pdf(const GuideView& g, const float4* xs, const float4* ns, const float4* dirs,
                            uint32_t count, float* pdf, hipStream_t stream);
hipError_t launch_guide_sample(const GuideView& g, const float4* xs, const float4* ns, const float4* u,
                               uint32_t count, float4* out, hipStream_t stream);
// One training pass over count samples: nblocks = ceil(count / 64) gradient
// rows in partials, one summed loss per block in losses.
hipError_t launch_guide_train(const GuideView& g, const float4* samples, uint32_t count, float* partials,
                              float* losses, hipStream_t stream);
hipError_t launch_guide_update(const GuideParamsDev& pd, const float* partials, uint32_t nblocks, float inv_count,
                               const AdamConfig& cfg, float* grad_out, hipStream_t stream);
hipError_t launch_guide_mirror(const GuideParamsDev& pd, hipStream_t stream);

}  // namespace guide
}  // namespace ptd
