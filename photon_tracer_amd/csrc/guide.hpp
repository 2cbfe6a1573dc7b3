// The learned path guide on the device: input encoding, the MLP on FP16
// MFMA tiles with FP32 accumulation, the 8-lobe von Mises-Fisher mixture it
// outputs, and the gradient of the training loss.
//
// Network: one-blob position (3 x 16) + one-blob normal angles (2 x 8) = 64
// inputs; four hidden layers of 64 with ReLU; 32 outputs = 8 lobes x
// (theta logit, phi logit, log kappa, weight logit). Parameters are kept as
// FP32 masters (Adam updates them) and mirrored in FP16, row-major [out][in]
// and transposed [in][out], for the MFMA operands.
//
// A wave handles 64 samples. Every product is computed transposed, Y^T =
// W X^T, on v_mfma_f32_32x32x16_f16: the A operand is a row of W (8 inputs a
// lane), the B operand is a sample's row of X (8 inputs a lane), and the
// 32x32 result leaves each lane one sample's column of 16 output features.
// Training writes them back to LDS as four 8-byte runs of the next layer's
// row (the backward pass needs every layer's activations); inference keeps
// them in registers and reads the weights from an LDS copy (forward_regs).
#pragma once

#include "device.hpp"

namespace ptd {
namespace guide {

constexpr int LOBES = 8;
constexpr int IN = 64, HID = 64, OUT = 4 * LOBES;
constexpr int N_LAYERS = 5;
constexpr int LAYER_STRIDE = HID * IN + HID;      // hidden layer: weights [64][64] then bias [64]
constexpr int OUT_W = 4 * LAYER_STRIDE;           // output layer weights [32][64]
constexpr int OUT_B = OUT_W + OUT * HID;
constexpr int N_PARAMS = OUT_B + OUT;             // 18,720

__host__ __device__ constexpr int w_off(int l) { return l < 4 ? l * LAYER_STRIDE : OUT_W; }
__host__ __device__ constexpr int b_off(int l) { return l < 4 ? l * LAYER_STRIDE + HID * IN : OUT_B; }
__host__ __device__ constexpr int n_out(int l) { return l < 4 ? HID : OUT; }

// log kappa is clamped to [-3, 10]: kappa in [0.05, 22026], a lobe down to
// about 0.4 degrees across (a 0.5 m opening seen from 10 m is 3 degrees).
constexpr float LOG_KAPPA_MIN = -3.0f, LOG_KAPPA_MAX = 10.0f;

// Row stride (halves) of a 64-wide activation tile in LDS: 144 bytes keeps
// 16-byte alignment and spreads a wave's row reads over the banks.
constexpr int XS = 72;
constexpr int OS = 33;   // row stride (floats) of the FP32 output tile

struct GuideView {
    const _Float16* w16;    // FP16 weights, [out][in] at w_off(l)
    const _Float16* wt16;   // FP16 weights transposed, [in][out] at w_off(l)
    const float* p32;       // FP32 master parameters (biases are read from here)
    float4 lo;              // position box: lower corner
    float4 inv_ext;         // 1 / box extent per axis
};

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// One-blob encoding: a value in [0,1] spread over `bins` Gaussian bumps of
// width one bin centred at (i + 0.5) / bins, exp(-d^2 / 2) for d = v bins -
// (i + 0.5), computed as 2^-(d k)^2 with k = sqrt(log2(e) / 2) folded into
// both terms of d, on the hardware's 2^x (v_exp_f32). Its error, a few units
// in the last place of FP32, is far below the FP16 rounding every input goes
// through.
constexpr float BLOB_K = 0.8493218002880191f;   // sqrt(log2(e) / 2)

// Eight neighbouring blobs, exp2(-(vk - c_j)^2) for c_j = c0 + j BLOB_K, from
// two exponentials: with d_j = vk - c_j, d_j^2 - d_{j+1}^2 = K (2 d_j - K), so
// b_{j+1} = b_j r_j with r_j = exp2(K (2 d_j - K)) and r_{j+1} = r_j
// exp2(-2 K^2) = r_j / e (2 K^2 = log2 e). Each product adds half an ulp: the
// run is within 14 ulps of the direct values, far below the FP16 rounding
// they go through. On the encoding's ranges (vk in [0, 16 K], c0 a bin
// centre) r_0 lies in [2^-13, 2^22]; where b_0 underflows, every value of the
// run is below 2^-39, zero in FP16 as the direct values are.
constexpr float BLOB_Q = 0.36787944117144233f;   // exp2(-2 BLOB_K^2) = 1/e
__device__ __forceinline__ void blob_run8(float vk, float c0, float (&b)[8]) {
    const float d0 = vk - c0;
    b[0] = __builtin_amdgcn_exp2f(-(d0 * d0));
    float r = __builtin_amdgcn_exp2f(BLOB_K * (2.0f * d0 - BLOB_K));
#pragma unroll
    for (int j = 1; j < 8; ++j) {
        b[j] = b[j - 1] * r;
        r = r * BLOB_Q;
    }
}

// The normal's angles for the encoding, from polynomial approximations of
// Abramowitz and Stegun's Handbook of Mathematical Functions (Hastings'
// fits): arccos by 4.4.46 (|error| <= 2e-8 on [0, 1], reflected for x < 0)
// and arctan by 4.4.49 (|error| <= 2e-8 on [0, 1]) after folding the
// quadrant into [0, 1]. Their error, with FP32 evaluation and the hardware
// reciprocal a few 1e-7, is far below the FP16 rounding of the one-blob
// inputs they feed; training and inference share them. atan2 keeps IEEE's
// signed-zero cases: atan2(+-0, -0) = +-pi, atan2(+-0, +0) = +-0.
__device__ __forceinline__ float enc_acos(float x) {
    const float a = fabsf(x);
    float p = -0.0012624911f;
    p = p * a + 0.0066700901f;
    p = p * a - 0.0170881256f;
    p = p * a + 0.0308918810f;
    p = p * a - 0.0501743046f;
    p = p * a + 0.0889789874f;
    p = p * a - 0.2145988016f;
    p = p * a + 1.5707963050f;
    const float r = __builtin_amdgcn_sqrtf(1.0f - a) * p;
    return x < 0 ? PI_F - r : r;
}
__device__ __forceinline__ float enc_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float hi = fmaxf(ax, ay), lo = fminf(ax, ay);
    const float t = hi > 0 ? lo * __builtin_amdgcn_rcpf(hi) : 0.0f;
    const float t2 = t * t;
    float p = 0.0028662257f;
    p = p * t2 - 0.0161657367f;
    p = p * t2 + 0.0429096138f;
    p = p * t2 - 0.0752896400f;
    p = p * t2 + 0.1065626393f;
    p = p * t2 - 0.1420889944f;
    p = p * t2 + 0.1999355085f;
    p = p * t2 - 0.3333314528f;
    float r = t + t * (p * t2);      // arctan(t), t in [0, 1]
    if (ay > ax) r = 0.5f * PI_F - r;
    if (__builtin_signbit(x)) r = PI_F - r;
    return copysignf(r, y);
}

// The five values a shading point is encoded from: its position normalised
// to the box per axis, and the normal's polar angle / pi and azimuth / (2 pi).
__device__ __forceinline__ void encode_scalars(const GuideView& g, f3 x, f3 n, float e[5]) {
    e[0] = clamp01((x.x - g.lo.x) * g.inv_ext.x);
    e[1] = clamp01((x.y - g.lo.y) * g.inv_ext.y);
    e[2] = clamp01((x.z - g.lo.z) * g.inv_ext.z);
    e[3] = clamp01(enc_acos(fminf(fmaxf(n.z, -1.0f), 1.0f)) * (1 / PI_F));
    e[4] = clamp01((enc_atan2(n.y, n.x) + PI_F) * (1 / (2 * PI_F)));
}

// The 64 inputs of a shading point, as one row: 16 bins per position axis,
// then 8 bins of the polar and 8 of the azimuth fraction.
__device__ __forceinline__ void encode(const GuideView& g, f3 x, f3 n, _Float16* row) {
    float e[5];
    encode_scalars(g, x, n, e);
    float b[8];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            blob_run8(e[a] * (16.0f * BLOB_K), (static_cast<float>(8 * h) + 0.5f) * BLOB_K, b);
#pragma unroll
            for (int i = 0; i < 8; ++i) row[16 * a + 8 * h + i] = static_cast<_Float16>(b[i]);
        }
    }
#pragma unroll
    for (int a = 3; a < 5; ++a) {
        blob_run8(e[a] * (8.0f * BLOB_K), 0.5f * BLOB_K, b);
#pragma unroll
        for (int i = 0; i < 8; ++i) row[48 + 8 * (a - 3) + i] = static_cast<_Float16>(b[i]);
    }
}

// ---------------------------------------------------------------------------
// The mixture.
// ---------------------------------------------------------------------------
struct Lobe {
    f3 mu;
    float kappa, w;
    float st, ct, sp, cp;   // sin/cos of the lobe's polar and azimuth angles
    float sa, sb;           // sigmoid of the two angle logits
    bool kappa_free;        // log kappa inside its clamp (its gradient passes)
};

__device__ __forceinline__ float sigmoid(float v) { return 1.0f / (1.0f + expf(-v)); }

__device__ __forceinline__ void mixture(const float* o, Lobe* L) {
    float mx = o[3];
#pragma unroll
    for (int k = 1; k < LOBES; ++k) mx = fmaxf(mx, o[4 * k + 3]);
    float e[LOBES], sum = 0;
#pragma unroll
    for (int k = 0; k < LOBES; ++k) {
        e[k] = expf(o[4 * k + 3] - mx);
        sum += e[k];
    }
#pragma unroll
    for (int k = 0; k < LOBES; ++k) {
        Lobe& l = L[k];
        l.w = e[k] / sum;
        l.sa = sigmoid(o[4 * k]);
        l.sb = sigmoid(o[4 * k + 1]);
        const float th = PI_F * l.sa, ph = 2 * PI_F * l.sb;
        l.st = sinf(th);
        l.ct = cosf(th);
        l.sp = sinf(ph);
        l.cp = cosf(ph);
        l.mu = mk(l.st * l.cp, l.st * l.sp, l.ct);
        const float lk = o[4 * k + 2];
        l.kappa_free = lk > LOG_KAPPA_MIN && lk < LOG_KAPPA_MAX;
        l.kappa = expf(fminf(fmaxf(lk, LOG_KAPPA_MIN), LOG_KAPPA_MAX));
    }
}

// vMF density on the sphere: kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (mu.w - 1)}.
__device__ __forceinline__ float vmf_pdf(f3 mu, float kappa, f3 w) {
    return kappa / (2 * PI_F * -expm1f(-2 * kappa)) * expf(kappa * (dot(mu, w) - 1));
}
__device__ __forceinline__ float vmf_pdf(const Lobe& l, f3 w) { return vmf_pdf(l.mu, l.kappa, w); }

// The mixture from the network's outputs as pt_guide_pdf / pt_guide_sample
// evaluate it (the path loop decodes the same lobes once per hit, below):
// the hardware's fast exponential, logarithm, sine and cosine (arguments here
// are bounded: angles in (0, 2 pi), exponents at most 2 e^10), where the
// trainer's gradient keeps the correctly rounded library functions. The
// sampler and the pdf share these functions and the lobe axis is
// renormalised, so a sample is weighted by the density it was drawn from.
// Quotients take the hardware reciprocal (within 1 ulp): __fdividef is the
// full correctly rounded division sequence on this target, ten instructions.
__device__ __forceinline__ float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fast_sigmoid(float v) { return __builtin_amdgcn_rcpf(1.0f + __expf(-v)); }

__device__ __forceinline__ float fast_vmf_pdf(f3 mu, float kappa, f3 w) {
    return fast_div(kappa, 2 * PI_F * (1.0f - __expf(-2 * kappa))) * __expf(kappa * (dot(mu, w) - 1));
}

// Lobe k's axis and concentration from the outputs o (FP32 or FP16).
template <class T>
__device__ __forceinline__ void lobe_at(const T* o, int k, f3& mu, float& kappa) {
    const float th = PI_F * fast_sigmoid(static_cast<float>(o[4 * k]));
    const float ph = 2 * PI_F * fast_sigmoid(static_cast<float>(o[4 * k + 1]));
    float st, ct, sp, cp;
    __sincosf(th, &st, &ct);
    __sincosf(ph, &sp, &cp);
    const f3 m = mk(st * cp, st * sp, ct);
    mu = m * rsqrtf(dot(m, m));
    kappa = __expf(fminf(fmaxf(static_cast<float>(o[4 * k + 2]), LOG_KAPPA_MIN), LOG_KAPPA_MAX));
}

// The softmax of the weight logits as (max, sum of exp(logit - max)).
template <class T>
__device__ __forceinline__ void weight_norm(const T* o, float& mx, float& sum) {
    mx = static_cast<float>(o[3]);
#pragma unroll 1
    for (int k = 1; k < LOBES; ++k) mx = fmaxf(mx, static_cast<float>(o[4 * k + 3]));
    sum = 0;
#pragma unroll 1
    for (int k = 0; k < LOBES; ++k) sum += __expf(static_cast<float>(o[4 * k + 3]) - mx);
}

// Mixture pdf at w, one lobe at a time (the loops stay rolled, so few
// registers are live in the shade kernels, which run 16 waves a block).
template <class T>
__device__ __forceinline__ float mixture_pdf(const T* o, f3 w) {
    float mx, sum;
    weight_norm(o, mx, sum);
    float p = 0;
#pragma unroll 1
    for (int k = 0; k < LOBES; ++k) {
        f3 mu;
        float kappa;
        lobe_at(o, k, mu, kappa);
        p += __expf(static_cast<float>(o[4 * k + 3]) - mx) * fast_vmf_pdf(mu, kappa, w);
    }
    return fast_div(p, sum);
}

// A direction from lobe (nn, kappa): the inversion c = 1 + log(u1 + (1 - u1)
// e^{-2 kappa}) / kappa for the cosine to the axis, a uniform azimuth u2.
__device__ __forceinline__ f3 vmf_sample(f3 nn, float kappa, float u1, float u2) {
    const float cw = fminf(1.0f, fmaxf(-1.0f, 1 + fast_div(__logf(u1 + (1 - u1) * __expf(-2 * kappa)), kappa)));
    const float s = sqrtf(fmaxf(0.0f, 1 - cw * cw));
    const float phi = 2 * PI_F * u2;
    float sphi, cphi;
    __sincosf(phi, &sphi, &cphi);
    const float sign = copysignf(1.0f, nn.z);
    const float pp = -__builtin_amdgcn_rcpf(sign + nn.z), qq = nn.x * nn.y * pp;
    const f3 t1 = mk(1 + sign * nn.x * nn.x * pp, sign * qq, -sign * nn.x);
    const f3 t2 = mk(qq, sign + nn.y * nn.y * pp, -nn.y);
    return t1 * (s * cphi) + t2 * (s * sphi) + nn * cw;
}

// A direction from the mixture: u0 picks the lobe by weight (the first whose
// running weight passes u0, the last if rounding leaves u0 above the sum),
// (u1, u2) sample it.
template <class T>
__device__ __forceinline__ f3 mixture_sample(const T* o, float u0, float u1, float u2) {
    float mx, sum;
    weight_norm(o, mx, sum);
    const float target = u0 * sum;
    float acc = 0;
    int sel = LOBES - 1;
    bool chosen = false;
#pragma unroll 1
    for (int k = 0; k < LOBES; ++k) {
        acc += __expf(static_cast<float>(o[4 * k + 3]) - mx);
        if (!chosen && target < acc) {
            sel = k;
            chosen = true;
        }
    }
    f3 nn;
    float kappa;
    lobe_at(o, sel, nn, kappa);
    return vmf_sample(nn, kappa, u1, u2);
}

// ---------------------------------------------------------------------------
// The mixture as the path loop carries it from the inference stage to the
// shade stage: one 32-bit word per lobe -- the axis in octahedral form with
// 10 bits a coordinate (bits 0-19, within about 0.2 degrees), log kappa over
// its clamp range in 6 bits (bits 20-25, kappa within 11%), and the lobe's
// softmax weight relative to the mixture's largest in 6 bits (bits 26-31;
// lobes under 1/126 of it are dropped). The inference kernel encodes each
// lobe once (its lanes hold whole lobes); the shade stage's sampler and pdf
// decode the same words the same way and weight by the density they sample,
// so the estimator stays unbiased whatever the rounding.
// ---------------------------------------------------------------------------
constexpr float MIX_KAPPA_STEP = (LOG_KAPPA_MAX - LOG_KAPPA_MIN) / 63.0f;

// Octahedral fold of the lower hemisphere (z < 0) onto the square's corners.
__device__ __forceinline__ void oct_fold(float& px, float& py) {
    const float qx = copysignf(1.0f - fabsf(py), px), qy = copysignf(1.0f - fabsf(px), py);
    px = qx;
    py = qy;
}

// The word of a lobe from its outputs (theta logit, phi logit, log kappa)
// and its weight relative to the largest lobe's, rel in (0, 1].
__device__ __forceinline__ uint32_t encode_lobe(float o0, float o1, float o2, float rel) {
    // The angles in turns (theta = pi sigmoid, phi = 2 pi sigmoid), the unit
    // the hardware's sine and cosine take.
    const float th = 0.5f * fast_sigmoid(o0), ph = fast_sigmoid(o1);
    const float st = __builtin_amdgcn_sinf(th), ct = __builtin_amdgcn_cosf(th);
    const float sp = __builtin_amdgcn_sinf(ph), cp = __builtin_amdgcn_cosf(ph);
    const float x = st * cp, y = st * sp;
    const float inv = __builtin_amdgcn_rcpf(fabsf(x) + fabsf(y) + fabsf(ct));
    float px = x * inv, py = y * inv;
    if (ct < 0) oct_fold(px, py);
    const uint32_t ux = static_cast<uint32_t>(fminf(fmaxf(px * 511.5f + 512.0f, 0.0f), 1023.0f));
    const uint32_t uy = static_cast<uint32_t>(fminf(fmaxf(py * 511.5f + 512.0f, 0.0f), 1023.0f));
    const float lk = fminf(fmaxf(o2, LOG_KAPPA_MIN), LOG_KAPPA_MAX) - LOG_KAPPA_MIN;
    const uint32_t uk = static_cast<uint32_t>(fminf(lk * (1.0f / MIX_KAPPA_STEP) + 0.5f, 63.0f));
    const uint32_t uw = static_cast<uint32_t>(fminf(fmaxf(rel * 63.0f + 0.5f, 0.0f), 63.0f));
    return ux | (uy << 10) | (uk << 20) | (uw << 26);
}

__device__ __forceinline__ float decoded_weight(const uint32_t* q, int k) { return static_cast<float>(q[k] >> 26); }

// Lobe k of an encoded mixture: its unit axis, kappa and (unnormalised) weight.
// The octahedral point is never the origin (its coordinates are odd
// multiples of 1/1023), so the axis always normalises.
__device__ __forceinline__ void decoded_lobe(const uint32_t* q, int k, f3& mu, float& kappa, float& w) {
    const uint32_t v = q[k];
    float px = static_cast<float>(v & 1023u) * (2.0f / 1023.0f) - 1.0f;
    float py = static_cast<float>((v >> 10) & 1023u) * (2.0f / 1023.0f) - 1.0f;
    const float pz = 1.0f - fabsf(px) - fabsf(py);
    if (pz < 0) oct_fold(px, py);
    const f3 m = mk(px, py, pz);
    mu = m * rsqrtf(dot(m, m));
    kappa = __expf(LOG_KAPPA_MIN + static_cast<float>((v >> 20) & 63u) * MIX_KAPPA_STEP);
    w = static_cast<float>(v >> 26);
}

__device__ __forceinline__ float decoded_weight_sum(const uint32_t* q) {
    float sum = 0;
#pragma unroll
    for (int k = 0; k < LOBES; ++k) sum += decoded_weight(q, k);
    return sum;
}

__device__ __forceinline__ float decoded_pdf(const uint32_t* q, f3 w) {
    float p = 0;
#pragma unroll 1
    for (int k = 0; k < LOBES; ++k) {
        f3 mu;
        float kappa, wk;
        decoded_lobe(q, k, mu, kappa, wk);
        p += wk * fast_vmf_pdf(mu, kappa, w);
    }
    return fast_div(p, decoded_weight_sum(q));
}

__device__ __forceinline__ f3 decoded_sample(const uint32_t* q, float u0, float u1, float u2) {
    const float target = u0 * decoded_weight_sum(q);
    float acc = 0;
    int sel = LOBES - 1;
    bool chosen = false;
#pragma unroll
    for (int k = 0; k < LOBES; ++k) {
        acc += decoded_weight(q, k);
        if (!chosen && target < acc) {
            sel = k;
            chosen = true;
        }
    }
    f3 nn;
    float kappa, wk;
    decoded_lobe(q, sel, nn, kappa, wk);
    return vmf_sample(nn, kappa, u1, u2);
}

// Gradient of the sample's loss ws * -log p(w) with respect to the 32
// outputs. Returns the loss.
__device__ __forceinline__ float loss_grad(const float* o, f3 w, float ws, float* d) {
    Lobe L[LOBES];
    mixture(o, L);
    float v[LOBES], p = 0;
#pragma unroll
    for (int k = 0; k < LOBES; ++k) {
        v[k] = vmf_pdf(L[k], w);
        p += L[k].w * v[k];
    }
    p = fmaxf(p, 1e-30f);
#pragma unroll
    for (int k = 0; k < LOBES; ++k) {
        const Lobe& l = L[k];
        const float r = l.w * v[k] / p;   // responsibility
        // log v = log kappa - log 2pi - log(1 - e^{-2 kappa}) + kappa (mu.w - 1)
        const float dlv_dk = 1 / l.kappa - 2 / expm1f(2 * l.kappa) + (dot(l.mu, w) - 1);
        const f3 dmu_dth = mk(l.ct * l.cp, l.ct * l.sp, -l.st);
        const f3 dmu_dph = mk(-l.st * l.sp, l.st * l.cp, 0.0f);
        const float g = -ws * r;   // d loss / d log v
        d[4 * k] = g * l.kappa * dot(w, dmu_dth) * (PI_F * l.sa * (1 - l.sa));
        d[4 * k + 1] = g * l.kappa * dot(w, dmu_dph) * (2 * PI_F * l.sb * (1 - l.sb));
        d[4 * k + 2] = l.kappa_free ? g * dlv_dk * l.kappa : 0.0f;
        d[4 * k + 3] = ws * (l.w - r);
    }
    return -ws * logf(p);
}

// ---------------------------------------------------------------------------
// MLP tiles. X, Y, dY: row-major [64 samples][XS] FP16 in LDS; XT, YT, dYT:
// feature-major [features][XS] (sample i at column i). Lane = (c, kh):
// c = lane & 31 picks the sample (or feature) within a 32-block, kh = lane >> 5
// the half of each 16-wide k chunk. The two operands index k the same way,
// so the sum runs over all k whatever order the unit takes them in.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_c() { return threadIdx.x & 31; }
__device__ __forceinline__ int lane_kh() { return (threadIdx.x >> 5) & 1; }

// Layer l forward: Y = act(X W^T + b). NOUT = 64 (hidden, ReLU) or 32 (the
// output layer, no activation, FP32 into O). YT may be null.
template <int NOUT>
__device__ __forceinline__ void layer_forward(const GuideView& g, int l, const _Float16* X, _Float16* Y,
                                              _Float16* YT, float* O) {
    const int c = lane_c(), kh = lane_kh();
    const _Float16* W = g.w16 + w_off(l);
    const float* B = g.p32 + b_off(l);
#pragma unroll
    for (int mb = 0; mb < NOUT / 32; ++mb) {
        h8 a[4];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) a[kb] = *reinterpret_cast<const h8*>(W + (mb * 32 + c) * IN + kb * 16 + kh * 8);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            f16v acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0;
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                const h8 bx = *reinterpret_cast<const h8*>(X + (nb * 32 + c) * XS + kb * 16 + kh * 8);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], bx, acc, 0, 0, 0);
            }
            const int i = nb * 32 + c;
#pragma unroll
            for (int g4 = 0; g4 < 4; ++g4) {
                const int m0 = mb * 32 + 8 * g4 + 4 * kh;
                float v[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) v[t] = acc[4 * g4 + t] + B[m0 + t];
                if (O) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) O[i * OS + m0 + t] = v[t];
                } else {
                    h4 hv;
#pragma unroll
                    for (int t = 0; t < 4; ++t) hv[t] = static_cast<_Float16>(fmaxf(v[t], 0.0f));
                    *reinterpret_cast<h4*>(Y + i * XS + m0) = hv;
                    if (YT) {
#pragma unroll
                        for (int t = 0; t < 4; ++t) YT[(m0 + t) * XS + i] = hv[t];
                    }
                }
            }
        }
    }
}

// The whole network on a wave's 64 encoded rows X[0]; hidden activations in
// X[1..4] (and their transposes in XT[0..4] when XT is not null); outputs in O.
__device__ __forceinline__ void forward(const GuideView& g, _Float16* const* X, _Float16* const* XT, float* O) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        layer_forward<HID>(g, l, X[l], X[l + 1], XT ? XT[l + 1] : nullptr, nullptr);
        __syncthreads();
    }
    layer_forward<OUT>(g, 4, X[4], nullptr, nullptr, O);
    __syncthreads();
}

// Layer l backward, given dY (= d loss / d layer output, row-major, NOUT
// wide) and its transpose dYT: the FP32 weight gradient dW = dY^T X into
// gw[j * 64 + k], the bias gradient into gb, and, when l > 0, the gradient
// with respect to the layer's input masked by the ReLU that made it
// (X > 0), into dX (row-major) and dXT.
template <int NOUT>
__device__ __forceinline__ void layer_backward(const GuideView& g, int l, const _Float16* dY, const _Float16* dYT,
                                               const _Float16* X, const _Float16* XT, float* gw, float* gb,
                                               _Float16* dX, _Float16* dXT, float unscale) {
    const int c = lane_c(), kh = lane_kh();
    // dW[j][k] = sum_i dY[i][j] X[i][k]: M = j, N = k, K = samples.
#pragma unroll
    for (int mb = 0; mb < NOUT / 32; ++mb) {
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            f16v acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0;
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                const h8 a = *reinterpret_cast<const h8*>(dYT + (mb * 32 + c) * XS + kb * 16 + kh * 8);
                const h8 bx = *reinterpret_cast<const h8*>(XT + (nb * 32 + c) * XS + kb * 16 + kh * 8);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, bx, acc, 0, 0, 0);
            }
            const int k = nb * 32 + c;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int j = mb * 32 + 8 * (r / 4) + 4 * kh + (r % 4);
                gw[j * IN + k] = acc[r] * unscale;
            }
        }
    }
    // Bias: sum over the samples, one feature per lane.
    if (threadIdx.x < NOUT) {
        float s = 0;
        for (int i = 0; i < 64; ++i) s += static_cast<float>(dYT[threadIdx.x * XS + i]);
        gb[threadIdx.x] = s * unscale;
    }
    if (l == 0) return;
    // dX^T = W^T dY^T: M = input feature k, N = sample, K = output j.
    const _Float16* WT = g.wt16 + w_off(l);
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        h8 a[NOUT / 16];
#pragma unroll
        for (int kb = 0; kb < NOUT / 16; ++kb)
            a[kb] = *reinterpret_cast<const h8*>(WT + (mb * 32 + c) * NOUT + kb * 16 + kh * 8);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            f16v acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0;
#pragma unroll
            for (int kb = 0; kb < NOUT / 16; ++kb) {
                const h8 bx = *reinterpret_cast<const h8*>(dY + (nb * 32 + c) * XS + kb * 16 + kh * 8);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], bx, acc, 0, 0, 0);
            }
            const int i = nb * 32 + c;
#pragma unroll
            for (int g4 = 0; g4 < 4; ++g4) {
                const int m0 = mb * 32 + 8 * g4 + 4 * kh;
                const h4 xm = *reinterpret_cast<const h4*>(X + i * XS + m0);
                h4 hv;
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    hv[t] = static_cast<float>(xm[t]) > 0.0f ? static_cast<_Float16>(acc[4 * g4 + t])
                                                             : static_cast<_Float16>(0.0f);
                *reinterpret_cast<h4*>(dX + i * XS + m0) = hv;
#pragma unroll
                for (int t = 0; t < 4; ++t) dXT[(m0 + t) * XS + i] = hv[t];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Register-resident forward pass (inference). A wave's 64 queries stay in
// VGPRs from the encoding to the outputs: no LDS, no barrier. Lane (c, kh)
// holds, for the queries c and 32 + c (nb = 0, 1), the 8 values of each
// 16-wide k chunk kb that its half kh feeds the matrix unit. A layer's 32x32
// result leaves lane (c, kh) the outputs 8 g4 + 4 kh + t (g4, t < 4) of row
// block mb for its two queries; they stay where they are and become the next
// layer's chunk kb = 2 mb + (g4 >> 1), element 4 (g4 & 1) + t. The product
// sums over every k of a chunk whatever their order, so the next layer's
// weight operand just takes its columns in the same order: element 4 s + t of
// chunk kb is column 16 kb + 8 s + 4 kh + t (two 8-byte loads).
// ---------------------------------------------------------------------------
typedef h8 RegActs[2][4];

// The encoded inputs in that layout (the first layer's chunks hold columns in
// their natural order: chunk kb, half kh = inputs 16 kb + 8 kh + j). Same
// arithmetic as encode(), so the same FP16 inputs.
// From the five encoding scalars of the lane's two queries: q0 of query c,
// q1 of query 32 + c.
__device__ __forceinline__ void encode_blobs(const float (&q0)[5], const float (&q1)[5], RegActs& X) {
    const int kh = lane_kh();
    const float c16 = (static_cast<float>(8 * kh) + 0.5f) * BLOB_K;   // first bin centre of this lane's half
    float b[8];
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        const float(&q)[5] = nb == 0 ? q0 : q1;
#pragma unroll
        for (int kb = 0; kb < 3; ++kb) {
            blob_run8(q[kb] * (16.0f * BLOB_K), c16, b);
#pragma unroll
            for (int j = 0; j < 8; ++j) X[nb][kb][j] = static_cast<_Float16>(b[j]);
        }
        blob_run8((kh ? q[4] : q[3]) * (8.0f * BLOB_K), 0.5f * BLOB_K, b);
#pragma unroll
        for (int j = 0; j < 8; ++j) X[nb][3][j] = static_cast<_Float16>(b[j]);
    }
}

// The lane's own query (32 kh + c) at point x with normal n: its scalars,
// and the partner lane's (the other half's) for the lane's other query.
__device__ __forceinline__ void encode_regs(const GuideView& g, f3 x, f3 n, RegActs& X) {
    float e[5], q0[5], q1[5];
    encode_scalars(g, x, n, e);
    const bool upper = lane_kh() != 0;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const float o = __shfl_xor(e[v], 32, 64);
        q0[v] = upper ? o : e[v];
        q1[v] = upper ? e[v] : o;
    }
    encode_blobs(q0, q1, X);
}

// ReLU of 8 values rounded to FP16: the conversion (round to nearest even),
// then a 16-bit integer max with 0 -- a negative value or -0 has its sign bit
// set and reads as a negative integer, 0 and positive values keep their bits.
typedef short s8 __attribute__((ext_vector_type(8)));
typedef float f8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ h8 relu_h8(f8 v) {
    const s8 bits = __builtin_bit_cast(s8, __builtin_convertvector(v, h8));
    return __builtin_bit_cast(h8, __builtin_elementwise_max(bits, s8(0)));
}

// The network's FP16 weights and FP32 biases copied to LDS once per block:
// rows padded to WROW halves (a wave's 16-byte row reads spread over the
// banks), rows of all layers one after another (layer l's row j at
// 64 l + j), and the columns of layers 1..4 stored in the register order,
// so every weight operand is one 16-byte read at a lane base plus an
// immediate offset.
constexpr int WROW = IN + 8;
constexpr int NET_ROWS = 4 * HID + OUT;
struct LdsNet {
    alignas(16) _Float16 w[NET_ROWS * WROW];
    alignas(16) float b[NET_ROWS];
};

// Fills s from the mirrors; all threads of the block take part (then a
// barrier). Column 16 kb + 8 s + 4 kh + t of a row of layers 1..4 goes to
// 16 kb + 8 kh + 4 s + t.
__device__ __forceinline__ void load_net(const GuideView& g, LdsNet& s) {
    for (uint32_t idx = threadIdx.x; idx < static_cast<uint32_t>(NET_ROWS) * (IN / 8); idx += blockDim.x) {
        const uint32_t r = idx >> 3, k8 = idx & 7u;
        const int l = r < 4u * HID ? static_cast<int>(r / HID) : 4;
        const uint32_t row = r - static_cast<uint32_t>(l) * HID;
        const h8 v = *reinterpret_cast<const h8*>(g.w16 + w_off(l) + (row * IN + 8u * k8));
        _Float16* d = s.w + r * WROW;
        if (l == 0) {
            *reinterpret_cast<h8*>(d + 8u * k8) = v;
        } else {
            const uint32_t kb = k8 >> 1, sb = k8 & 1u;
            *reinterpret_cast<h4*>(d + 16u * kb + 4u * sb) = h4{v[0], v[1], v[2], v[3]};
            *reinterpret_cast<h4*>(d + 16u * kb + 8u + 4u * sb) = h4{v[4], v[5], v[6], v[7]};
        }
    }
    for (uint32_t r = threadIdx.x; r < static_cast<uint32_t>(NET_ROWS); r += blockDim.x) {
        const int l = r < 4u * HID ? static_cast<int>(r / HID) : 4;
        s.b[r] = g.p32[b_off(l) + (r - static_cast<uint32_t>(l) * HID)];
    }
    __syncthreads();
}

struct LdsWeights {
    const LdsNet& s;
    // Every row starts on a 16-byte boundary (WROW halves = 144 bytes), and
    // so does every read: one ds_read_b128 at the lane's base plus an
    // immediate offset.
    __device__ __forceinline__ h8 a(int l, int mb, int kb) const {
        const uint32_t r = static_cast<uint32_t>(l * HID + mb * 32 + lane_c());
        const h8* p = reinterpret_cast<const h8*>(s.w + (r * WROW + 16u * kb + 8u * lane_kh()));
        return *static_cast<const h8*>(__builtin_assume_aligned(p, 16));
    }
    __device__ __forceinline__ float4 bias(int l, int mb, int g4) const {
        const float4* p = reinterpret_cast<const float4*>(
            s.b + (static_cast<uint32_t>(l * HID + mb * 32) + 4u * lane_kh() + 8u * g4));
        return *static_cast<const float4*>(__builtin_assume_aligned(p, 16));
    }
};

// Row block mb of layer l for both query blocks: acc[nb] = W X^T + b.
template <class Src>
__device__ __forceinline__ void block_product(const Src& src, int l, const RegActs& X, int mb, f16v (&acc)[2]) {
    h8 a[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) a[kb] = src.a(l, mb, kb);
#pragma unroll
    for (int g4 = 0; g4 < 4; ++g4) {
        const float4 bb = src.bias(l, mb, g4);
        acc[0][4 * g4] = bb.x;
        acc[0][4 * g4 + 1] = bb.y;
        acc[0][4 * g4 + 2] = bb.z;
        acc[0][4 * g4 + 3] = bb.w;
    }
    acc[1] = acc[0];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[kb], X[nb][kb], acc[nb], 0, 0, 0);
    }
}

// The whole network on X; O[nb][4 g4 + t] is output 8 g4 + 4 kh + t of
// query 32 nb + c (FP32, before any rounding). Src supplies the weight
// operands and biases (LdsWeights).
template <class Src>
__device__ __forceinline__ void forward_regs(const Src& src, RegActs& X, f16v (&O)[2]) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        RegActs Y;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            f16v acc[2];
            block_product(src, l, X, mb, acc);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    f8 v;
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = acc[nb][8 * h + j];
                    Y[nb][2 * mb + h] = relu_h8(v);
                }
            }
        }
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) X[nb][kb] = Y[nb][kb];
        }
    }
    block_product(src, 4, X, 0, O);
}

}  // namespace guide
}  // namespace ptd
