// Kernels of the learned path guide: batched forward passes (network
// outputs, mixture pdf, mixture samples), the training pass (forward and
// backward on FP16 MFMA tiles, one wave per 64 samples, the FP32 weight
// gradient of each block written to its own row), and the update (the rows
// summed in a fixed order, then Adam on the FP32 masters and the FP16
// mirrors rewritten). No step depends on thread timing, so training is
// deterministic.
#include "guide_kernels.hpp"

namespace ptd {
namespace guide {

namespace {

// Runs the network on the block's queries (one wave) as the path loop's
// inference does -- weights copied to LDS, activations in registers -- and
// leaves query threadIdx.x's 32 outputs at o + threadIdx.x * OS.
struct EvalLds {
    LdsNet net;
    float o[64 * OS];
};

__device__ __forceinline__ void eval_block(const GuideView& g, EvalLds& s, f3 x, f3 n) {
    load_net(g, s.net);
    RegActs X;
    encode_regs(g, x, n, X);
    f16v O[2];
    forward_regs(LdsWeights{s.net}, X, O);
    const int c = lane_c(), kh = lane_kh();
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
#pragma unroll
        for (int g4 = 0; g4 < 4; ++g4) {
#pragma unroll
            for (int t = 0; t < 4; ++t) s.o[(32 * nb + c) * OS + 8 * g4 + 4 * kh + t] = O[nb][4 * g4 + t];
        }
    }
    __syncthreads();
}

}  // namespace

// Raw network outputs of `count` queries (position, normal): 32 floats each.
__global__ __launch_bounds__(64) void k_guide_eval(GuideView g, const float4* xs, const float4* ns, uint32_t count,
                                                    float* out) {
    __shared__ EvalLds s;
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const bool ok = i < count;
    const f3 x = ok ? xyz(xs[i]) : mk(0, 0, 0), n = ok ? xyz(ns[i]) : mk(0, 0, 1);
    eval_block(g, s, x, n);
    if (ok) {
        for (int j = 0; j < OUT; ++j) out[static_cast<size_t>(i) * OUT + j] = s.o[threadIdx.x * OS + j];
    }
}

// Mixture pdf of direction dirs[i] at query i.
__global__ __launch_bounds__(64) void k_guide_pdf(GuideView g, const float4* xs, const float4* ns,
                                                   const float4* dirs, uint32_t count, float* pdf) {
    __shared__ EvalLds s;
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const bool ok = i < count;
    const f3 x = ok ? xyz(xs[i]) : mk(0, 0, 0), n = ok ? xyz(ns[i]) : mk(0, 0, 1);
    eval_block(g, s, x, n);
    if (!ok) return;
    pdf[i] = mixture_pdf(s.o + threadIdx.x * OS, xyz(dirs[i]));
}

// One direction per query from the mixture, with its pdf: u[i] = (u0, u1, u2).
__global__ __launch_bounds__(64) void k_guide_sample(GuideView g, const float4* xs, const float4* ns,
                                                      const float4* u, uint32_t count, float4* out) {
    __shared__ EvalLds s;
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const bool ok = i < count;
    const f3 x = ok ? xyz(xs[i]) : mk(0, 0, 0), n = ok ? xyz(ns[i]) : mk(0, 0, 1);
    eval_block(g, s, x, n);
    if (!ok) return;
    const float* o = s.o + threadIdx.x * OS;
    const float4 uu = u[i];
    const f3 w = mixture_sample(o, uu.x, uu.y, uu.z);
    out[i] = make_float4(w.x, w.y, w.z, mixture_pdf(o, w));
}

// The training pass over `count` samples, 64 per block. A sample is three
// float4: (position, weight), (normal, 0), (direction, 0); its loss is
// weight * -log p(direction). Each block writes the gradient of its
// samples' summed loss to partials[blockIdx.x * N_PARAMS ...] and the summed
// loss to losses[blockIdx.x].
__global__ __launch_bounds__(64) void k_guide_train(GuideView g, const float4* samples, uint32_t count,
                                                     float* partials, float* losses) {
    __shared__ TrainLds s;
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const bool ok = i < count;
    float4 s0 = make_float4(0, 0, 0, 0), s1 = make_float4(0, 0, 1, 0), s2 = make_float4(0, 0, 1, 0);
    if (ok) {
        s0 = samples[3 * static_cast<size_t>(i)];
        s1 = samples[3 * static_cast<size_t>(i) + 1];
        s2 = samples[3 * static_cast<size_t>(i) + 2];
    }
    const float ws = ok ? s0.w : 0.0f;
    _Float16* row = s.x[0] + threadIdx.x * XS;
    encode(g, xyz(s0), xyz(s1), row);
    for (int f = 0; f < IN; ++f) s.xt[0][f * XS + threadIdx.x] = row[f];
    __syncthreads();
    _Float16* X[5] = {s.x[0], s.x[1], s.x[2], s.x[3], s.x[4]};
    _Float16* XT[5] = {s.xt[0], s.xt[1], s.xt[2], s.xt[3], s.xt[4]};
    forward(g, X, XT, s.o);

    float o[OUT], d[OUT];
    for (int j = 0; j < OUT; ++j) o[j] = s.o[threadIdx.x * OS + j];
    float loss = loss_grad(o, xyz(s2), ws, d);
    for (int j = 0; j < OUT; ++j) {
        const _Float16 h = to_half_sat(d[j] * LOSS_SCALE);
        s.dy[0][threadIdx.x * XS + j] = h;
        s.dyt[0][j * XS + threadIdx.x] = h;
    }
    __syncthreads();

    float* P = partials + static_cast<size_t>(blockIdx.x) * N_PARAMS;
    const float un = 1.0f / LOSS_SCALE;
    layer_backward<OUT>(g, 4, s.dy[0], s.dyt[0], s.x[4], s.xt[4], P + w_off(4), P + b_off(4), s.dy[1], s.dyt[1], un);
    __syncthreads();
    layer_backward<HID>(g, 3, s.dy[1], s.dyt[1], s.x[3], s.xt[3], P + w_off(3), P + b_off(3), s.dy[0], s.dyt[0], un);
    __syncthreads();
    layer_backward<HID>(g, 2, s.dy[0], s.dyt[0], s.x[2], s.xt[2], P + w_off(2), P + b_off(2), s.dy[1], s.dyt[1], un);
    __syncthreads();
    layer_backward<HID>(g, 1, s.dy[1], s.dyt[1], s.x[1], s.xt[1], P + w_off(1), P + b_off(1), s.dy[0], s.dyt[0], un);
    __syncthreads();
    layer_backward<HID>(g, 0, s.dy[0], s.dyt[0], s.x[0], s.xt[0], P + w_off(0), P + b_off(0), nullptr, nullptr, un);

    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (threadIdx.x == 0) losses[blockIdx.x] = loss;
}

// Sums the blocks' gradient rows and scales by inv_count. With grad_out:
// writes the gradient and stops. Otherwise one Adam step (Kingma and Ba) on
// the FP32 masters, then the FP16 mirrors. A block takes UPDATE_PARAMS
// parameters; its UPDATE_GROUPS groups of threads each add a run of rows in
// row order, and the groups' sums are added in group order: a fixed order,
// so the step is deterministic, with many blocks and short runs (one thread
// per parameter over all rows left most of the GPU idle, the update being a
// chain of dependent loads).
constexpr int UPDATE_PARAMS = 16, UPDATE_GROUPS = 16;
__global__ __launch_bounds__(UPDATE_PARAMS * UPDATE_GROUPS) void k_guide_update(GuideParamsDev pd, const float* partials,
                                                                                uint32_t nblocks, float inv_count,
                                                                                AdamConfig cfg, float* grad_out) {
    __shared__ float part[UPDATE_GROUPS][UPDATE_PARAMS];
    const uint32_t jj = threadIdx.x % UPDATE_PARAMS, grp = threadIdx.x / UPDATE_PARAMS;
    const uint32_t j = blockIdx.x * UPDATE_PARAMS + jj;
    const uint32_t run = (nblocks + UPDATE_GROUPS - 1) / UPDATE_GROUPS;
    const uint32_t b0 = grp * run, b1 = min(nblocks, b0 + run);
    float gsum = 0;
    if (j < static_cast<uint32_t>(N_PARAMS))
        for (uint32_t b = b0; b < b1; ++b) gsum += partials[static_cast<size_t>(b) * N_PARAMS + j];
    part[grp][jj] = gsum;
    __syncthreads();
    if (grp != 0 || j >= static_cast<uint32_t>(N_PARAMS)) return;
    gsum = part[0][jj];
    for (int k = 1; k < UPDATE_GROUPS; ++k) gsum += part[k][jj];
    const float gr = gsum * inv_count;
    if (grad_out) {
        grad_out[j] = gr;
        return;
    }
    const float m = cfg.beta1 * pd.m[j] + (1 - cfg.beta1) * gr;
    const float v = cfg.beta2 * pd.v[j] + (1 - cfg.beta2) * gr * gr;
    pd.m[j] = m;
    pd.v[j] = v;
    const float mh = m / cfg.bias1, vh = v / cfg.bias2;
    const float p = pd.p32[j] - cfg.lr * mh / (sqrtf(vh) + cfg.eps);
    pd.p32[j] = p;
    mirror_one(pd, j, p);
}

// Rewrites the FP16 mirrors from the FP32 masters (after weights are set).
__global__ void k_guide_mirror(GuideParamsDev pd) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < static_cast<uint32_t>(N_PARAMS)) mirror_one(pd, j, pd.p32[j]);
}

}  // namespace guide
}  // namespace ptd

namespace ptd {
namespace guide {

static uint32_t blocks_of(uint32_t count, uint32_t per) { return (count + per - 1) / per; }

hipError_t launch_guide_eval(const GuideView& g, const float4* xs, const float4* ns, uint32_t count, float* out,
                             hipStream_t stream) {
    if (count == 0) return hipSuccess;
    k_guide_eval<<<blocks_of(count, 64), 64, 0, stream>>>(g, xs, ns, count, out);
    return hipGetLastError();
}

hipError_t launch_guide_pdf(const GuideView& g, const float4* xs, const float4* ns, const float4* dirs,
                            uint32_t count, float* pdf, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    k_guide_pdf<<<blocks_of(count, 64), 64, 0, stream>>>(g, xs, ns, dirs, count, pdf);
    return hipGetLastError();
}

hipError_t launch_guide_sample(const GuideView& g, const float4* xs, const float4* ns, const float4* u,
                               uint32_t count, float4* out, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    k_guide_sample<<<blocks_of(count, 64), 64, 0, stream>>>(g, xs, ns, u, count, out);
    return hipGetLastError();
}

hipError_t launch_guide_train(const GuideView& g, const float4* samples, uint32_t count, float* partials,
                              float* losses, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    k_guide_train<<<blocks_of(count, 64), 64, 0, stream>>>(g, samples, count, partials, losses);
    return hipGetLastError();
}

hipError_t launch_guide_update(const GuideParamsDev& pd, const float* partials, uint32_t nblocks, float inv_count,
                               const AdamConfig& cfg, float* grad_out, hipStream_t stream) {
    k_guide_update<<<blocks_of(N_PARAMS, UPDATE_PARAMS), UPDATE_PARAMS * UPDATE_GROUPS, 0, stream>>>(
        pd, partials, nblocks, inv_count, cfg, grad_out);
    return hipGetLastError();
}

hipError_t launch_guide_mirror(const GuideParamsDev& pd, hipStream_t stream) {
    k_guide_mirror<<<blocks_of(N_PARAMS, 256), 256, 0, stream>>>(pd);
    return hipGetLastError();
}

}  // namespace guide
}  // namespace ptd
