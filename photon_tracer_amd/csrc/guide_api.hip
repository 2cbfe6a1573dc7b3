// C ABI of the learned path guide: creation and weight initialisation,
// batched queries (outputs, pdf, samples) and training steps on host-given
// samples. Renders that use or train the guide are in api.hip (pt_render).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "api_internal.hpp"

namespace G = ptd::guide;

namespace {

constexpr float kBeta1 = 0.9f, kBeta2 = 0.99f, kEps = 1e-8f;

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// He-uniform weights for the ReLU layers, a tenth of that for the output
// layer (the first mixtures are then near-uniform), zero biases.
std::vector<float> initial_weights(uint64_t seed) {
    std::vector<float> p(G::N_PARAMS, 0.0f);
    uint64_t s = seed ^ 0x6a09e667f3bcc909ull;
    for (int l = 0; l < G::N_LAYERS; ++l) {
        const float bound = std::sqrt(6.0f / G::IN) * (l == 4 ? 0.1f : 1.0f);
        const int nw = G::n_out(l) * G::IN;
        for (int i = 0; i < nw; ++i) {
            const double u = double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            p[G::w_off(l) + i] = float((2 * u - 1) * bound);
        }
    }
    return p;
}

// The union of the boxes of a root node's used child slots.
void root_bounds(const std::vector<uint32_t>& w, uint32_t root, float* lo, float* hi, bool& any) {
    if (root == NONE) return;
    const uint32_t off = w[2] / 4;   // nodes section (gpu_pack.hpp:177-196)
    const float* f = reinterpret_cast<const float*>(w.data() + off + 32 * size_t(root));
    for (int c = 0; c < 4; ++c) {
        const float l[3] = {f[c], f[4 + c], f[8 + c]}, h[3] = {f[12 + c], f[16 + c], f[20 + c]};
        if (!(l[0] <= h[0] && l[1] <= h[1] && l[2] <= h[2])) continue;   // empty slot: +inf box
        for (int a = 0; a < 3; ++a) {
            lo[a] = any ? std::min(lo[a], l[a]) : l[a];
            hi[a] = any ? std::max(hi[a], h[a]) : h[a];
        }
        any = true;
    }
}

void check_guide(pt_guide* g) {
    if (!g || !g->scene) throw pt::status_error(PT_ERR_INVALID, "guide is NULL");
}

void upload(dbuf& b, const void* host, size_t bytes, hipStream_t st) {
    b.ensure(bytes);
    HIPC(hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, st));
}

// Points and normals (3 floats each) as float4 arrays on the device.
void upload_points(dbuf& b, const float* v, size_t n, hipStream_t st) {
    if (!v) throw pt::status_error(PT_ERR_INVALID, "NULL array");
    std::vector<float> f4(4 * n);
    for (size_t i = 0; i < n; ++i) {
        f4[4 * i] = v[3 * i];
        f4[4 * i + 1] = v[3 * i + 1];
        f4[4 * i + 2] = v[3 * i + 2];
        f4[4 * i + 3] = 0;
    }
    upload(b, f4.data(), f4.size() * 4, st);
}

void check_count(size_t n) {
    if (n > (size_t(1) << 26)) throw pt::status_error(PT_ERR_INVALID, "more than 2^26 guide queries in one call");
}

// One training pass over n device samples: gradient rows, then either the
// mean gradient (grad_out) or an Adam step. Returns the number of rows.
uint32_t train_pass(pt_guide* g, const float4* samples, uint32_t n, float* grad_out) {
    hipStream_t st = g->scene->stream;
    const uint32_t nb = (n + 63) / 64;
    g->partials.ensure(size_t(nb) * G::N_PARAMS * sizeof(float));
    g->losses.ensure(size_t(nb) * sizeof(float));
    HIPC(G::launch_guide_train(g->view(), samples, n, g->partials.as<float>(), g->losses.as<float>(), st));
    G::AdamConfig cfg{};
    cfg.lr = g->lr;
    cfg.beta1 = kBeta1;
    cfg.beta2 = kBeta2;
    cfg.eps = kEps;
    if (!grad_out) {
        ++g->step;
        cfg.bias1 = 1 - std::pow(kBeta1, float(g->step));
        cfg.bias2 = 1 - std::pow(kBeta2, float(g->step));
    } else {
        cfg.bias1 = cfg.bias2 = 1;
    }
    HIPC(G::launch_guide_update(g->params(), g->partials.as<float>(), nb, 1.0f / float(n), cfg, grad_out, st));
    return nb;
}

float mean_loss(pt_guide* g, uint32_t nb, uint32_t n) {
    std::vector<float> l(nb);
    HIPC(hipMemcpyAsync(l.data(), g->losses.p, nb * sizeof(float), hipMemcpyDeviceToHost, g->scene->stream));
    HIPC(hipStreamSynchronize(g->scene->stream));
    double s = 0;
    for (float v : l) s += v;
    return float(s / n);
}

}  // namespace

namespace ptapi {

void guide_train_on(pt_guide* g, const float4* samples, uint32_t n) {
    if (n == 0) return;
    const uint32_t mb = uint32_t(std::max(64, g->minibatch));
    const uint32_t passes = (n + mb - 1) / mb;
    const uint32_t steps = g->steps_per_batch > 0 ? uint32_t(g->steps_per_batch) : passes;
    uint32_t nb = 0, m = 0;
    for (uint32_t k = 0; k < steps; ++k) {
        const uint32_t off = (k % passes) * mb;
        m = std::min(mb, n - off);
        nb = train_pass(g, samples + size_t(off) * 3, m, nullptr);
        g->stat_samples += m;
    }
    g->stat_steps += steps;
    g->stat_loss = mean_loss(g, nb, m);
}

}  // namespace ptapi

// A scene going away first leaves its guides usable only for destruction:
// their calls then fail with "guide is NULL".
pt_scene::~pt_scene() {
    for (pt_guide* g : guides) g->scene = nullptr;
    release_all();
}

pt_guide::~pt_guide() {
    if (scene) {
        if (scene->guide == this) scene->guide = nullptr;
        auto& v = scene->guides;
        v.erase(std::remove(v.begin(), v.end(), this), v.end());
    }
    if (scene && scene->stream) (void)hipStreamSynchronize(scene->stream);
    for (dbuf* b : {&p32, &m, &v, &w16, &wt16, &partials, &losses, &grad, &samples, &io_a, &io_b, &io_c, &io_d,
                    &records, &train_count})
        b->release();
}

extern "C" {

pt_status pt_guide_create(pt_scene* scene, const pt_guide_params* params, pt_guide** out) {
    PT_GUARD(
        if (!scene || !out) throw pt::status_error(PT_ERR_INVALID, "scene or output is NULL");
        *out = nullptr;
        pt_guide_params p{};
        if (params) p = *params;
        auto g = std::make_unique<pt_guide>();
        g->scene = scene;
        g->lr = p.learning_rate > 0 ? p.learning_rate : 1e-3f;
        g->steps_per_batch = std::max(0, p.steps_per_batch);
        g->minibatch = p.minibatch > 0 ? p.minibatch : 16384;
        g->weight_clamp = p.weight_clamp > 0 ? p.weight_clamp : 16.0f;
        if (p.bounces < 0 || p.bounces > 63) throw pt::status_error(PT_ERR_INVALID, "guide bounces must be in [0, 63]");
        g->bounces = p.bounces > 0 ? p.bounces : ptd::MAX_BOUNCES;
        const bool box_given = p.box_lo[0] != 0 || p.box_lo[1] != 0 || p.box_lo[2] != 0 || p.box_hi[0] != 0 ||
                               p.box_hi[1] != 0 || p.box_hi[2] != 0;
        if (box_given) {
            for (int a = 0; a < 3; ++a) {
                if (!(p.box_hi[a] > p.box_lo[a])) throw pt::status_error(PT_ERR_INVALID, "empty guide box");
                g->lo[a] = p.box_lo[a];
                g->hi[a] = p.box_hi[a];
            }
        } else {
            bool any = false;
            root_bounds(scene->host, scene->host[18], g->lo, g->hi, any);
            root_bounds(scene->host, scene->host[19], g->lo, g->hi, any);
            if (!any) {
                for (int a = 0; a < 3; ++a) {
                    g->lo[a] = 0;
                    g->hi[a] = 1;
                }
            }
        }
        hipStream_t st = scene->stream;
        const std::vector<float> w = initial_weights(p.seed);
        upload(g->p32, w.data(), w.size() * 4, st);
        g->m.ensure(G::N_PARAMS * 4);
        g->v.ensure(G::N_PARAMS * 4);
        HIPC(hipMemsetAsync(g->m.p, 0, G::N_PARAMS * 4, st));
        HIPC(hipMemsetAsync(g->v.p, 0, G::N_PARAMS * 4, st));
        g->w16.ensure(G::N_PARAMS * 2);
        g->wt16.ensure(G::N_PARAMS * 2);
        HIPC(G::launch_guide_mirror(g->params(), st));
        HIPC(hipStreamSynchronize(st));
        scene->guides.push_back(g.get());
        *out = g.release();)
}

pt_status pt_guide_destroy(pt_guide* guide) {
    PT_GUARD(delete guide;)
}

pt_status pt_scene_set_guide(pt_scene* scene, pt_guide* guide) {
    PT_GUARD(
        if (!scene) throw pt::status_error(PT_ERR_INVALID, "scene is NULL");
        if (guide && guide->scene != scene) throw pt::status_error(PT_ERR_INVALID, "the guide belongs to another scene");
        scene->guide = guide;)
}

pt_status pt_guide_get_weights(pt_guide* g, float* params, size_t n) {
    PT_GUARD(
        check_guide(g);
        if (!params || n != size_t(G::N_PARAMS))
            throw pt::status_error(PT_ERR_INVALID, "weights: need PT_GUIDE_PARAMS floats");
        HIPC(hipMemcpyAsync(params, g->p32.p, n * 4, hipMemcpyDeviceToHost, g->scene->stream));
        HIPC(hipStreamSynchronize(g->scene->stream));)
}

pt_status pt_guide_set_weights(pt_guide* g, const float* params, size_t n) {
    PT_GUARD(
        check_guide(g);
        if (!params || n != size_t(G::N_PARAMS))
            throw pt::status_error(PT_ERR_INVALID, "weights: need PT_GUIDE_PARAMS floats");
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(params[i])) throw pt::status_error(PT_ERR_INVALID, "weights: not finite");
        hipStream_t st = g->scene->stream;
        upload(g->p32, params, n * 4, st);
        HIPC(hipMemsetAsync(g->m.p, 0, G::N_PARAMS * 4, st));
        HIPC(hipMemsetAsync(g->v.p, 0, G::N_PARAMS * 4, st));
        g->step = 0;
        HIPC(G::launch_guide_mirror(g->params(), st));
        HIPC(hipStreamSynchronize(st));)
}

pt_status pt_guide_eval(pt_guide* g, const float* pos, const float* nrm, size_t n, float* out) {
    PT_GUARD(
        check_guide(g);
        check_count(n);
        if (!out) throw pt::status_error(PT_ERR_INVALID, "output is NULL");
        if (n == 0) return PT_OK;
        hipStream_t st = g->scene->stream;
        upload_points(g->io_a, pos, n, st);
        upload_points(g->io_b, nrm, n, st);
        g->io_c.ensure(n * G::OUT * 4);
        HIPC(G::launch_guide_eval(g->view(), g->io_a.as<float4>(), g->io_b.as<float4>(), uint32_t(n),
                                  g->io_c.as<float>(), st));
        HIPC(hipMemcpyAsync(out, g->io_c.p, n * G::OUT * 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));)
}

pt_status pt_guide_pdf(pt_guide* g, const float* pos, const float* nrm, const float* dir, size_t n, float* pdf) {
    PT_GUARD(
        check_guide(g);
        check_count(n);
        if (!pdf) throw pt::status_error(PT_ERR_INVALID, "output is NULL");
        if (n == 0) return PT_OK;
        hipStream_t st = g->scene->stream;
        upload_points(g->io_a, pos, n, st);
        upload_points(g->io_b, nrm, n, st);
        upload_points(g->io_c, dir, n, st);
        g->io_d.ensure(n * 4);
        HIPC(G::launch_guide_pdf(g->view(), g->io_a.as<float4>(), g->io_b.as<float4>(), g->io_c.as<float4>(),
                                 uint32_t(n), g->io_d.as<float>(), st));
        HIPC(hipMemcpyAsync(pdf, g->io_d.p, n * 4, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));)
}

pt_status pt_guide_sample(pt_guide* g, const float* pos, const float* nrm, const float* u, size_t n,
                          float* dir_pdf) {
    PT_GUARD(
        check_guide(g);
        check_count(n);
        if (!dir_pdf) throw pt::status_error(PT_ERR_INVALID, "output is NULL");
        if (n == 0) return PT_OK;
        hipStream_t st = g->scene->stream;
        upload_points(g->io_a, pos, n, st);
        upload_points(g->io_b, nrm, n, st);
        upload_points(g->io_c, u, n, st);
        g->io_d.ensure(n * 16);
        HIPC(G::launch_guide_sample(g->view(), g->io_a.as<float4>(), g->io_b.as<float4>(), g->io_c.as<float4>(),
                                    uint32_t(n), g->io_d.as<float4>(), st));
        HIPC(hipMemcpyAsync(dir_pdf, g->io_d.p, n * 16, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));)
}

pt_status pt_guide_gradient(pt_guide* g, const float* samples, size_t n, float* grad, float* loss) {
    PT_GUARD(
        check_guide(g);
        check_count(n);
        if (!samples || !grad || n == 0) throw pt::status_error(PT_ERR_INVALID, "need samples and a gradient array");
        hipStream_t st = g->scene->stream;
        upload(g->samples, samples, n * 48, st);
        g->grad.ensure(G::N_PARAMS * 4);
        const uint32_t nb = train_pass(g, g->samples.as<float4>(), uint32_t(n), g->grad.as<float>());
        HIPC(hipMemcpyAsync(grad, g->grad.p, G::N_PARAMS * 4, hipMemcpyDeviceToHost, st));
        const float l = mean_loss(g, nb, uint32_t(n));
        if (loss) *loss = l;)
}

pt_status pt_guide_step(pt_guide* g, const float* samples, size_t n, float* loss) {
    PT_GUARD(
        check_guide(g);
        check_count(n);
        if (!samples || n == 0) throw pt::status_error(PT_ERR_INVALID, "need samples");
        upload(g->samples, samples, n * 48, g->scene->stream);
        const uint32_t nb = train_pass(g, g->samples.as<float4>(), uint32_t(n), nullptr);
        const float l = mean_loss(g, nb, uint32_t(n));
        ++g->stat_steps;
        g->stat_samples += n;
        if (loss) *loss = l;)
}

pt_status pt_guide_stats(pt_guide* g, uint64_t* steps, uint64_t* samples, float* last_loss) {
    PT_GUARD(
        check_guide(g);
        if (steps) *steps = g->stat_steps;
        if (samples) *samples = g->stat_samples;
        if (last_loss) *last_loss = g->stat_loss;)
}

}  // extern "C"
