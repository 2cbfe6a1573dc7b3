// The denoiser's kernels: demodulation, five a-trous passes, remodulation.
//
// The filter is the reference's (denoise.hpp:55-139), with its operations in
// its order (no FMA contraction), so the result differs from the reference's
// only where the device's exponential rounds differently from the host's:
// the lighting (color / albedo floored at 0.01) is blurred five times with a
// 5x5 B3-spline kernel whose taps are 1, 2, 4, 8 and 16 pixels apart; each
// tap is weighted by the normals' cosine to the 64th power, and by one
// exponential of the albedo difference and of the luminance difference in
// units of the center's noise (its variance blurred over 3x3); the variance
// is filtered along with the color. A pixel is one thread; the image arrays
// (a 1080p frame's are 100 MB) stay in the Infinity Cache across the taps.
#include "denoise_kernels.hpp"

namespace ptd {
namespace dn {

namespace {

constexpr int PASSES = 5;
constexpr float SIGMA_LUMINANCE = 4;
constexpr float ALBEDO_SIGMA2 = 0.01f;
constexpr float INV_ALBEDO_SIGMA2 = 1 / ALBEDO_SIGMA2;
constexpr int BX = 16, BY = 16;

__device__ __forceinline__ float lum(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

__device__ __forceinline__ float kernel_weight(int d) {
    return d == 0 ? 3.0f / 8 : d == 1 ? 1.0f / 4 : 1.0f / 16;
}

// Per pixel: the variance of the mean from the aux luminance square
// (main.cpp:181-182), then the demodulated lighting, its variance and the
// unit normal (denoise.hpp:62-75).
__global__ __launch_bounds__(256) void k_dn_prep(int n, int spp, const float* rgb, const float* aux, DenoiseBufs b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float* a = aux + 7 * static_cast<size_t>(p);
    const float* c = rgb + 3 * static_cast<size_t>(p);
    const float mean = lum(c[0], c[1], c[2]);
    const float variance = (a[6] - mean * mean) / static_cast<float>(spp);
    float l[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) l[k] = c[k] / fmaxf(a[k], 0.01f);
    b.light[0][p] = make_float4(l[0], l[1], l[2], lum(l[0], l[1], l[2]));
    const float al = fmaxf(lum(a[0], a[1], a[2]), 0.01f);
    b.var[0][p] = fmaxf(variance, 0.0f) / (al * al);
    const float len = sqrtf(a[3] * a[3] + a[4] * a[4] + a[5] * a[5]);
    b.normal[p] = len > 1e-6f ? make_float4(a[3] / len, a[4] / len, a[5] / len, 0) : make_float4(0, 0, 0, 0);
    b.albedo[p] = make_float4(a[0], a[1], a[2], 0);
}

// One pass at tap spacing `step` (denoise.hpp:77-132), from buffers `src`
// into the other pair.
__global__ __launch_bounds__(BX * BY) void k_dn_pass(int w, int h, int step, int src, DenoiseBufs b) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    if (x >= w || y >= h) return;
    const float4* light = b.light[src];
    const float* var = b.var[src];
    const size_t p = static_cast<size_t>(y) * w + x;
    const float4 np = b.normal[p], ap = b.albedo[p];
    const bool p_sky = np.x == 0 && np.y == 0 && np.z == 0;
    const float lp = light[p].w;
    float v = 0, vw = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int qx = x + dx, qy = y + dy;
            if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
            const float k = (dx ? 0.5f : 1) * (dy ? 0.5f : 1);
            v += k * var[static_cast<size_t>(qy) * w + qx];
            vw += k;
        }
    const float scale = 1 / (SIGMA_LUMINANCE * sqrtf(v / vw) + 1e-4f);
    float s0 = 0, s1 = 0, s2 = 0, sum_w = 0, sum_var = 0;
    for (int dy = -2; dy <= 2; ++dy) {
        const int qy = y + dy * step;
        if (qy < 0 || qy >= h) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int qx = x + dx * step;
            if (qx < 0 || qx >= w) continue;
            const size_t q = static_cast<size_t>(qy) * w + qx;
            const float4 nq = b.normal[q], aq = b.albedo[q], lq = light[q];
            const bool q_sky = nq.x == 0 && nq.y == 0 && nq.z == 0;
            if (p_sky != q_sky) continue;
            float weight = kernel_weight(dx < 0 ? -dx : dx) * kernel_weight(dy < 0 ? -dy : dy);
            if (!p_sky) {
                float nd = fmaxf(np.x * nq.x + np.y * nq.y + np.z * nq.z, 0.0f);
#pragma unroll
                for (int k = 0; k < 6; ++k) nd *= nd;
                weight *= nd;
            }
            const float da0 = ap.x - aq.x, da1 = ap.y - aq.y, da2 = ap.z - aq.z;
            weight *= expf(-(da0 * da0 + da1 * da1 + da2 * da2) * INV_ALBEDO_SIGMA2 - fabsf(lp - lq.w) * scale);
            s0 += weight * lq.x;
            s1 += weight * lq.y;
            s2 += weight * lq.z;
            sum_w += weight;
            sum_var += weight * weight * var[q];
        }
    }
    const float o0 = s0 / sum_w, o1 = s1 / sum_w, o2 = s2 / sum_w;   // the center tap always has weight
    b.light[src ^ 1][p] = make_float4(o0, o1, o2, lum(o0, o1, o2));
    b.var[src ^ 1][p] = sum_var / (sum_w * sum_w);
}

// Remodulation (denoise.hpp:135-137).
__global__ __launch_bounds__(256) void k_dn_finish(int n, int src, DenoiseBufs b, float* out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 l = b.light[src][p], a = b.albedo[p];
    float* o = out + 3 * static_cast<size_t>(p);
    o[0] = l.x * fmaxf(a.x, 0.01f);
    o[1] = l.y * fmaxf(a.y, 0.01f);
    o[2] = l.z * fmaxf(a.z, 0.01f);
}

}  // namespace

hipError_t launch_denoise(int width, int height, int spp, const float* rgb, const float* aux, float* out,
                          const DenoiseBufs& b, hipStream_t stream) {
    const int n = width * height;
    if (n <= 0) return hipSuccess;
    k_dn_prep<<<(n + 255) / 256, 256, 0, stream>>>(n, spp, rgb, aux, b);
    const dim3 grid((width + BX - 1) / BX, (height + BY - 1) / BY), block(BX, BY);
    int src = 0;
    for (int pass = 0; pass < PASSES; ++pass) {
        k_dn_pass<<<grid, block, 0, stream>>>(width, height, 1 << pass, src, b);
        src ^= 1;
    }
    k_dn_finish<<<(n + 255) / 256, 256, 0, stream>>>(n, src, b, out);
    return hipGetLastError();
}

}  // namespace dn
}  // namespace ptd
