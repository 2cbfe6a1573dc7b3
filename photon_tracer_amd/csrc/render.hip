// The wavefront path loop: render_row_linear / ray_color (renderer.hpp:100-149,
// 366-416) split into stages over SoA queues in HBM.
//
//   extend   closest hit of every queued ray            (geometry::hit);
//            at bounce 0 the batch's camera rays, made in the kernel
//   shade    emission, scatter, light sample, roulette  (ray_color's body);
//            at bounce 0 the path state starts in registers
//   shadow   any hit of every queued light sample       (geometry::occluded)
//   ...      extend/shade/shadow again for each bounce, up to max_bounces
//   accumulate  per pixel, the batch's samples in sample order
//   finish   ray counts into the totals, control blocks zeroed
//
// A path's state is indexed by its path number, not its queue slot, and every
// random number it draws comes from its own counter-based stream, so the
// image does not depend on the order in which queues fill. Queues are filled
// by wave-aggregated appends (ballot, popcount, one atomic per wave).
#include "kernels.hpp"
#include "traverse.hpp"

namespace ptd {

namespace {

constexpr int SHADE_BLOCK = 1024;
constexpr int SHADE_WAVES = SHADE_BLOCK / 64;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63; }

// Slots in K queues for the lanes of a block where pred[k] holds: each wave
// counts its lanes with a ballot, the block adds up its waves in LDS, and one
// thread takes the block's run of each queue with a single atomic (the
// counters are shared by the whole grid across all XCDs, where atomics on one
// address are slow). Called by every thread of the block.
template <int K>
struct BlockAppend {
    uint32_t count[K][SHADE_WAVES];
    uint32_t base[K][SHADE_WAVES];
};

template <int K>
__device__ __forceinline__ void block_append(BlockAppend<K>& sh, uint32_t* const* counters, const bool* pred,
                                             uint32_t* slot) {
    const uint32_t w = threadIdx.x >> 6;
    unsigned long long mask[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        mask[k] = __ballot(pred[k]);
        if (lane_id() == 0) sh.count[k][w] = static_cast<uint32_t>(__popcll(mask[k]));
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        uint32_t total = 0;
        for (int i = 0; i < SHADE_WAVES; ++i) total += sh.count[k][i];
        uint32_t at = total ? atomicAdd(counters[k], total) : 0;
        for (int i = 0; i < SHADE_WAVES; ++i) {
            sh.base[k][i] = at;
            at += sh.count[k][i];
        }
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane_id()) - 1ull;
#pragma unroll
    for (int k = 0; k < K; ++k) slot[k] = sh.base[k][w] + static_cast<uint32_t>(__popcll(mask[k] & below));
    __syncthreads();  // sh is reused by the next iteration
}

// sample_2d (sampler.hpp:107-124): QMC for pairs up to 2 * qmc_bounces,
// else two numbers from the path's stream.
__device__ __forceinline__ void sample_2d(uint32_t pseed, uint32_t index, int pair, uint64_t key, uint32_t& cnt,
                                          float& a, float& b) {
    if (pair > 2 * QMC_BOUNCES) {
        a = rand01(key, cnt);
        b = rand01(key, cnt);
        return;
    }
    qmc_pair(pseed, index, pair, a, b);
}

__device__ __forceinline__ f3 reflect(f3 v, f3 n) { return v - n * (2 * dot(v, n)); }

__device__ __forceinline__ f3 refract(f3 uv, f3 n, float ratio) {
    const float cos_theta = fminf(dot(-uv, n), 1.0f);
    const f3 r_perp = (uv + n * cos_theta) * ratio;
    const f3 r_parallel = n * -sqrtf(fabsf(1 - dot(r_perp, r_perp)));
    return r_perp + r_parallel;
}

__device__ __forceinline__ float schlick(float cosine, float ratio) {
    float r0 = (1 - ratio) / (1 + ratio);
    r0 = r0 * r0;
    const float x = 1 - cosine, x2 = x * x;
    return r0 + (1 - r0) * x2 * x2 * x;
}

}  // namespace

// Path p of the current batch: its pixel, sample number and camera ray
// (renderer.hpp:379-385, camera.hpp:45-47). Paths are numbered pixel-major
// within the band, spb samples of a pixel next to each other, so a wave's
// camera rays are neighbours. Recomputed where needed rather than stored:
// bounce 0 reads no ray or path state from HBM.
struct CameraSample {
    f3 o, d;
    uint32_t pseed, index;
    int i, j;
};

__device__ __forceinline__ CameraSample camera_sample(const FrameParams& fp, const FrameConst& fc, uint32_t p) {
    CameraSample cs;
    const uint32_t pixel = p / fp.spb, s = p % fp.spb;
    cs.i = static_cast<int>(pixel % fp.W);
    const int k = static_cast<int>(pixel / fp.W);   // row within the band, top first
    cs.j = fp.j_hi - 1 - k;
    cs.index = static_cast<uint32_t>(fc.first_sample) + fc.batch * fp.spb + s;
    cs.pseed = pixel_seed(fc.seed, cs.i, cs.j);
    float du, dv;
    qmc_pair(cs.pseed, cs.index, 0, du, dv);
    const float u = (static_cast<float>(cs.i) + du) / static_cast<float>(fp.W - 1);
    const float v = (static_cast<float>(cs.j) + dv) / static_cast<float>(fp.H - 1);
    cs.o = xyz(fp.sc.cam[0]);
    cs.d = xyz(fp.sc.cam[1]) + xyz(fp.sc.cam[2]) * u + xyz(fp.sc.cam[3]) * v - cs.o;
    return cs;
}

// The traversal stage of bounce b: the light samples queued by bounce b - 1
// (any hit, geometry::occluded: an unoccluded sample adds its radiance to its
// path) and the rays of bounce b (closest hit, geometry::hit), drained by one
// persistent grid from one counter, shadow rays first. PRIMARY: bounce 0,
// whose rays are the batch's camera rays, made here; it has no shadow rays.
//
// Scenes without instances walk with lane refill: each lane advances its own
// query one stack entry at a time, and once at least REFILL lanes of a wave
// have finished theirs, those lanes take the next items from the wave's run
// of the queue (one atomic per run of 64 * fetch_groups items). A wave then
// stays busy with most of its lanes, instead of waiting for its longest walk
// before taking 64 new ones. Starting a query costs the whole wave the setup
// code (a camera ray: the QMC jitter and the camera; a queued ray: its loads
// and three divisions), so lanes wait until most of the wave is idle: 56 of
// 64 for camera rays, 48 for queued ones (measured on the 1M-triangle
// icosphere at 1080p; the speed is flat from 40 to 60). Each query still
// goes through the same steps, so the hits are the same.
template <bool PRIMARY>
struct refill_threshold {
    static constexpr uint32_t lanes = PRIMARY ? 56 : 48;
};

template <int S, bool COUNT, bool PRIMARY, class Stack>
__device__ __forceinline__ void extend_refill(const FrameParams& fp, int b, Ctrl* c, uint32_t n_shadow, uint32_t n,
                                              Stack& st, Visits& vis) {
    const float4* ro = fp.ro[b & 1];
    const float4* rd = fp.rd[b & 1];
    FrameConst fc;
    if (PRIMARY) fc = *fp.frame;
    const uint32_t chunk = 64 * fetch_groups(n);
    const unsigned long long below = (1ull << lane_id()) - 1ull;
    uint32_t pool = 0, pool_end = 0;   // the wave's run of items not yet started
    bool drained = false;              // the queue has no more items
    bool live = false, any = false;
    uint32_t item = 0;
    Walk w;
    while (true) {
        const unsigned long long idle = __ballot(!live);
        const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (drained) {
            if (n_idle == 64) break;
        } else if (n_idle >= refill_threshold<PRIMARY>::lanes) {
            if (pool == pool_end) {
                pool = wave_fetch(&c->w_extend, chunk);
                pool_end = pool < n ? min(pool + chunk, n) : pool;
                drained = pool >= n;
            }
            if (!live) {
                const uint32_t r = static_cast<uint32_t>(__popcll(idle & below));
                if (pool + r < pool_end) {
                    item = pool + r;
                    any = !PRIMARY && item < n_shadow;
                    f3 o, d;
                    float tmin = T_MIN, tmax = T_MAX;
                    if (PRIMARY) {
                        const CameraSample cs = camera_sample(fp, fc, item);
                        o = cs.o;
                        d = cs.d;
                    } else {
                        const float4 a = any ? fp.so[item] : ro[item - n_shadow];
                        const float4 bb = any ? fp.sd[item] : rd[item - n_shadow];
                        o = xyz(a);
                        d = xyz(bb);
                        tmin = a.w;
                        tmax = bb.w;
                    }
                    live = !walk_begin<COUNT>(fp.sc, w, o, d, tmin, tmax, any, st, &vis);
                    if (!live) {   // ended at once: only an any-hit query on a large sphere
                        item |= 0x80000000u;
                    }
                }
            }
            pool = min(pool + n_idle, pool_end);
        }
        bool done = (item & 0x80000000u) != 0;
        if (live && walk_step<COUNT>(fp.sc, w, any, st, &vis)) {
            live = false;
            done = true;
        }
        if (done) {
            item &= 0x7fffffffu;
            if (any) {
                if (w.kind == 0) {   // unoccluded: the light sample reaches its path
                    const float4 l = fp.sl[item];
                    const uint32_t p = __float_as_uint(l.w);
                    float4 r = fp.pb[p];
                    r.x = r.x + l.x;
                    r.y = r.y + l.y;
                    r.z = r.z + l.z;
                    fp.pb[p] = r;
                }
            } else {
                Hit h;
                h.t = w.tmax;
                h.u = w.hu;
                h.v = w.hv;
                h.kind = w.kind;
                h.prim = w.prim;
                h.inst = NONE;
                fp.hits[item - n_shadow] = h;
            }
            item = 0;
        }
    }
}

template <int S, bool COUNT, bool INST, bool PRIMARY>
__global__ __launch_bounds__(TRACE_BLOCK)
__attribute__((amdgpu_waves_per_eu(trace_occupancy<S, INST, true>::waves))) void k_extend(FrameParams fp, int b) {
    if (!PRIMARY && fp.frame->tail_from <= static_cast<uint32_t>(b)) return;  // the tail launch has these rays
    __shared__ uint2 lds[S * TRACE_BLOCK];
    auto st = lane_stack<S, TRACE_BLOCK>(lds, fp.spill);
    Visits vis;
    Ctrl* c = fp.ctrl + b;
    const uint32_t n_shadow = PRIMARY ? 0u : fp.ctrl[b - 1].n_shadow;
    const uint32_t n_rays = PRIMARY ? fp.P : c->n_rays;
    const uint32_t n = n_shadow + n_rays;
    if (PRIMARY && blockIdx.x == 0 && threadIdx.x == 0) c->n_rays = fp.P;
    if constexpr (!INST) {
        extend_refill<S, COUNT, PRIMARY>(fp, b, c, n_shadow, n, st, vis);
        if (COUNT) add_visits(fp.visits, vis);
        return;
    }
    const float4* ro = fp.ro[b & 1];
    const float4* rd = fp.rd[b & 1];
    FrameConst fc;
    if (PRIMARY) fc = *fp.frame;
    const uint32_t groups = fetch_groups(n);
    while (true) {
        const uint32_t base = wave_fetch(&c->w_extend, 64 * groups);
        if (base >= n) break;
        for (uint32_t g = 0; g < groups; ++g) {
            const uint32_t item = base + 64 * g + lane_id();
            if (item >= n) break;
            Hit h;
            if (!PRIMARY && item < n_shadow) {
                const float4 a = fp.so[item], bb = fp.sd[item];
                if (!trace_ray<true, COUNT, INST>(fp.sc, xyz(a), xyz(bb), a.w, bb.w, h, st, &vis)) {
                    const float4 l = fp.sl[item];
                    const uint32_t p = __float_as_uint(l.w);
                    float4 r = fp.pb[p];
                    r.x = r.x + l.x;
                    r.y = r.y + l.y;
                    r.z = r.z + l.z;
                    fp.pb[p] = r;
                }
                continue;
            }
            const uint32_t q = item - n_shadow;
            f3 o, d;
            float tmin = T_MIN, tmax = T_MAX;
            if (PRIMARY) {
                const CameraSample cs = camera_sample(fp, fc, q);
                o = cs.o;
                d = cs.d;
            } else {
                const float4 a = ro[q], bb = rd[q];
                o = xyz(a);
                d = xyz(bb);
                tmin = a.w;
                tmax = bb.w;
            }
            trace_ray<false, COUNT, INST>(fp.sc, o, d, tmin, tmax, h, st, &vis);
            fp.hits[q] = h;
        }
    }
    if (COUNT) add_visits(fp.visits, vis);
}

// The state a path carries from bounce to bounce.
struct PathState {
    f3 thr, rad;        // throughput, radiance so far
    uint32_t cnt;       // draws taken from the path's random stream
    bool sampled;       // this bounce took a light sample (renderer.hpp:106-108,130)
    uint64_t key;       // the path's random stream
    uint2 pq;           // QMC pixel seed and sample index
};

// What a bounce produced: the next ray (cont: from pnt along dir) and a light
// sample to test (shadow: from pnt along sh_d up to sh_tmax, adding sh_l if
// unoccluded); backface: a light sample that counts as a ray but is zero
// without a test (renderer.hpp:72,76-77).
struct ShadeResult {
    bool cont, shadow, backface;
    f3 pnt, dir, sh_d, sh_l;
    float sh_tmax;
};

// ray_color's loop body for bounce b (renderer.hpp:109-146): the ray from o
// along d, whose closest hit is h, of path p.
__device__ __forceinline__ ShadeResult shade_hit(const FrameParams& fp, int b, uint32_t p, const Hit& h, f3 o, f3 d,
                                                 PathState& ps) {
    const SceneView& sc = fp.sc;
    ShadeResult r;
    r.cont = r.shadow = r.backface = false;
    r.pnt = r.dir = r.sh_d = r.sh_l = mk(0, 0, 0);
    r.sh_tmax = 0;
    const bool was_sampled = ps.sampled;
    ps.sampled = false;
    if (h.kind == 0) {
        // renderer.hpp:24-29
        f3 bg = mk(0, 0, 0);
        if (sc.flags & 1u) {
            const f3 unit = normalize(d);
            const float t = 0.5f * (unit.y + 1);
            bg = mk(1, 1, 1) * (1 - t) + mk(0.5f, 0.7f, 1.0f) * t;
        }
        ps.rad = ps.rad + mul(ps.thr, bg);
        if (fp.want_aux && b == 0) {
            fp.paux[2 * p] = make_float4(1, 1, 1, 0);
            fp.paux[2 * p + 1] = make_float4(0, 0, 0, 0);
        }
        return r;
    }
    // The hit record (geometry.hpp:135-151, sphere.hpp:50-58,
    // triangle.hpp:63-74, geometry.hpp:231-244).
    const f3 pnt = o + d * h.t;
    r.pnt = pnt;
    f3 nrm;
    bool front;
    uint32_t mat;
    bool sampled_light = false;
    if (h.kind <= 2) {
        const float4 cr = sc.spheres[2 * h.prim];
        mat = __float_as_uint(sc.spheres[2 * h.prim + 1].x);
        const f3 outward = (pnt - xyz(cr)) * (1 / cr.w);
        front = dot(d, outward) < 0;
        nrm = front ? outward : -outward;
        if (h.kind == 2 && fp.nee) {
            const float4 em = sc.mats[3 * mat + 1];
            sampled_light = em.x > 0 || em.y > 0 || em.z > 0;
        }
    } else {
        const float4* tr = sc.tris + 7 * h.prim;
        mat = __float_as_uint(tr[0].w);
        const f3 n0 = xyz(tr[3]), n1 = xyz(tr[4]), n2 = xyz(tr[5]), fn = xyz(tr[6]);
        const float b0 = 1 - h.u - h.v;
        const f3 sn = normalize(n0 * b0 + n1 * h.u + n2 * h.v);
        if (h.inst == NONE) {
            front = dot(d, fn) < 0;
            nrm = front ? sn : -sn;
        } else {
            const float4* in = sc.inst + 7 * h.inst;
            const float4 r0 = in[3], r1 = in[4], r2 = in[5];
            const f3 ld = mk(r0.x * d.x + r0.y * d.y + r0.z * d.z, r1.x * d.x + r1.y * d.y + r1.z * d.z,
                             r2.x * d.x + r2.y * d.y + r2.z * d.z);
            front = dot(ld, fn) < 0;
            const f3 ln = front ? sn : -sn;
            // the transpose of to_object (transform.hpp:23-27)
            nrm = normalize(mk(r0.x * ln.x + r1.x * ln.y + r2.x * ln.z, r0.y * ln.x + r1.y * ln.y + r2.y * ln.z,
                               r0.z * ln.x + r1.z * ln.y + r2.z * ln.z));
            const uint32_t im = __float_as_uint(in[6].y);
            if (im != NONE) mat = im;
        }
    }
    const float4 m0 = sc.mats[3 * mat], m1 = sc.mats[3 * mat + 1], m2 = sc.mats[3 * mat + 2];
    if (front && !(was_sampled && sampled_light)) ps.rad = ps.rad + mul(ps.thr, xyz(m1));

    // material::scatter (material.hpp:59-137)
    const int type = static_cast<int>(m0.w);
    f3 att = xyz(m0), dir = mk(0, 0, 0);
    bool ok = true;
    if (type == 0) {
        float a, bq;
        sample_2d(ps.pq.x, ps.pq.y, 1 + 2 * b, ps.key, ps.cnt, a, bq);
        const float rr = sqrtf(a), phi = 2 * PI_F * bq;
        const f3 nn = nrm;
        const float sign = copysignf(1.0f, nn.z);
        const float pp = -1 / (sign + nn.z), qq = nn.x * nn.y * pp;
        const f3 t1 = mk(1 + sign * nn.x * nn.x * pp, sign * qq, -sign * nn.x);
        const f3 t2 = mk(qq, sign + nn.y * nn.y * pp, -nn.y);
        dir = t1 * (rr * cosf(phi)) + t2 * (rr * sinf(phi)) + nn * sqrtf(fmaxf(0.0f, 1 - a));
    } else if (type == 1) {
        const f3 reflected = reflect(normalize(d), nrm);
        f3 ball = mk(0, 0, 0);
        for (int tries = 0; tries < 64; ++tries) {  // vec3.hpp:121-127
            const float x = -1 + 2 * rand01(ps.key, ps.cnt);
            const float y = -1 + 2 * rand01(ps.key, ps.cnt);
            const float z = -1 + 2 * rand01(ps.key, ps.cnt);
            ball = mk(x, y, z);
            if (dot(ball, ball) < 1) break;
        }
        dir = reflected + ball * m1.w;
        ok = dot(dir, nrm) > 0;
    } else if (type == 2) {
        att = mk(1, 1, 1);
        const float ior = m2.x;
        const float ratio = front ? (1 / ior) : ior;
        const f3 ud = normalize(d);
        const float cos_theta = fminf(dot(-ud, nrm), 1.0f);
        const float sin_theta = sqrtf(1 - cos_theta * cos_theta);
        const bool cannot_refract = ratio * sin_theta > 1;
        const bool refl = cannot_refract || schlick(cos_theta, ratio) > rand01(ps.key, ps.cnt);
        dir = refl ? reflect(ud, nrm) : refract(ud, nrm, ratio);
    } else {
        ok = false;
    }
    if (fp.want_aux && b == 0) {
        const f3 fa = ok && type != 2 ? att : mk(1, 1, 1);
        fp.paux[2 * p] = make_float4(fa.x, fa.y, fa.z, 0);
        fp.paux[2 * p + 1] = make_float4(nrm.x, nrm.y, nrm.z, 0);
    }
    if (!ok) return r;
    const bool nee = type == 0 && fp.nee && sc.n_lights > 0 && b + 1 < fp.depth;
    if (nee) {
        // direct_light (renderer.hpp:41-80)
        const uint32_t nl = sc.n_lights;
        const uint32_t pick = min(static_cast<uint32_t>(rand01(ps.key, ps.cnt) * nl), nl - 1);
        const uint32_t ls = sc.lights[pick];
        const float4 lcr = sc.spheres[2 * ls];
        const uint32_t lmat = __float_as_uint(sc.spheres[2 * ls + 1].x);
        const f3 to_center = xyz(lcr) - pnt;
        const float dist2 = dot(to_center, to_center), r2 = lcr.w * lcr.w;
        if (dist2 > r2) {
            const float sin2_max = r2 / dist2;
            const float cos_max = sqrtf(1 - sin2_max);
            const float omc = sin2_max / (1 + cos_max);
            const f3 w = to_center * (1 / sqrtf(dist2));
            const f3 ax = fabsf(w.x) > 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
            const f3 uu = normalize(cross(w, ax)), vv = cross(w, uu);
            float s1, s2;
            sample_2d(ps.pq.x, ps.pq.y, 2 + 2 * b, ps.key, ps.cnt, s1, s2);
            const float cos_t = 1 - s1 * omc;
            const float sin_t = sqrtf(fmaxf(0.0f, 1 - cos_t * cos_t));
            const float phi = 2 * PI_F * s2;
            const f3 ldir = uu * (cosf(phi) * sin_t) + vv * (sinf(phi) * sin_t) + w * cos_t;
            const float cos_surface = dot(ldir, nrm);
            float t_light = __builtin_huge_valf();
            if (cos_surface > 0 && sphere_hit_f(lcr, pnt, ldir, 0.001f, t_light)) {
                const f3 lp = pnt + ldir * t_light;
                const f3 outward = (lp - xyz(lcr)) * (1 / lcr.w);
                if (dot(ldir, outward) < 0) {
                    const float pdf = 1 / (2 * PI_F * omc);
                    const f3 contrib =
                        xyz(sc.mats[3 * lmat + 1]) * (cos_surface / PI_F / pdf * static_cast<float>(nl));
                    r.shadow = true;
                    r.sh_d = ldir;
                    r.sh_tmax = t_light * 0.9999f;
                    r.sh_l = mul(mul(ps.thr, att), contrib);
                } else {
                    r.backface = true;
                }
            }
        }
    }
    ps.sampled = nee;
    ps.thr = mul(ps.thr, att);
    if (b + 1 >= 3) {  // roulette_after (renderer.hpp:22,140-146)
        const float pm = fmaxf(ps.thr.x, fmaxf(ps.thr.y, ps.thr.z));
        if (pm < 1) {
            if (rand01(ps.key, ps.cnt) >= pm) return r;
            ps.thr = ps.thr * (1 / pm);
        }
    }
    if (b + 1 < fp.depth) {
        r.cont = true;
        r.dir = dir;
    }
    return r;
}

// The shade stage of bounce b over its queue. PRIMARY: bounce 0; the camera
// ray and the fresh path state are rebuilt from the path number instead of
// read back.
template <bool PRIMARY>
__global__ __launch_bounds__(SHADE_BLOCK) void k_shade(FrameParams fp, int b) {
    if (fp.frame->tail_from <= static_cast<uint32_t>(b)) return;  // the tail launch has taken these paths
    Ctrl* c = fp.ctrl + b;
    const uint32_t n = PRIMARY ? fp.P : c->n_rays;
    FrameConst fc;
    if (PRIMARY) fc = *fp.frame;
    const int cur = b & 1, nxt = cur ^ 1;
    __shared__ BlockAppend<3> sh;
    uint32_t* const counters[3] = {&fp.ctrl[b + 1].n_rays, &c->n_shadow, &c->n_backface};
    const uint32_t stride = gridDim.x * SHADE_BLOCK;
    // The block steps through the queue together (block_append synchronises).
    for (uint32_t bbase = blockIdx.x * SHADE_BLOCK; bbase < n; bbase += stride) {
        const uint32_t q = bbase + threadIdx.x;
        ShadeResult r;
        r.cont = r.shadow = r.backface = false;
        uint32_t p = 0;
        if (q < n) {
            const Hit h = fp.hits[q];
            f3 o, d;
            PathState ps;
            if (PRIMARY) {
                p = q;
                const CameraSample cs = camera_sample(fp, fc, p);
                o = cs.o;
                d = cs.d;
                ps.thr = mk(1, 1, 1);
                ps.rad = mk(0, 0, 0);
                ps.cnt = 0;
                ps.sampled = false;
                ps.key = path_key(fc.seed, cs.i, cs.j, cs.index);
                ps.pq = make_uint2(cs.pseed, cs.index);
            } else {
                p = fp.rpath[cur][q];
                const float4 o4 = fp.ro[cur][q], d4 = fp.rd[cur][q];
                const float4 A = fp.pa[p], B = fp.pb[p];
                const uint2 kk = fp.pkey[p];
                ps.pq = fp.pq[p];
                ps.key = (static_cast<uint64_t>(kk.y) << 32) | kk.x;
                ps.thr = xyz(A);
                ps.rad = xyz(B);
                ps.cnt = __float_as_uint(A.w);
                ps.sampled = (__float_as_uint(B.w) & 1u) != 0;
                o = xyz(o4);
                d = xyz(d4);
            }
            r = shade_hit(fp, b, p, h, o, d, ps);
            // The radiance always (accumulate reads it); the rest only for a
            // path that goes on.
            fp.pb[p] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, __uint_as_float(ps.sampled ? 1u : 0u));
            if (r.cont) {
                fp.pa[p] = make_float4(ps.thr.x, ps.thr.y, ps.thr.z, __uint_as_float(ps.cnt));
                if (PRIMARY) {
                    fp.pkey[p] = make_uint2(static_cast<uint32_t>(ps.key), static_cast<uint32_t>(ps.key >> 32));
                    fp.pq[p] = ps.pq;
                }
            }
        }
        // Every thread of the block reaches the appends.
        const bool pred[3] = {r.cont, r.shadow, r.backface};
        uint32_t slot[3];
        block_append<3>(sh, counters, pred, slot);
        if (r.cont) {
            fp.ro[nxt][slot[0]] = make_float4(r.pnt.x, r.pnt.y, r.pnt.z, T_MIN);
            fp.rd[nxt][slot[0]] = make_float4(r.dir.x, r.dir.y, r.dir.z, T_MAX);
            fp.rpath[nxt][slot[0]] = p;
        }
        if (r.shadow) {
            fp.so[slot[1]] = make_float4(r.pnt.x, r.pnt.y, r.pnt.z, 0.001f);
            fp.sd[slot[1]] = make_float4(r.sh_d.x, r.sh_d.y, r.sh_d.z, r.sh_tmax);
            fp.sl[slot[1]] = make_float4(r.sh_l.x, r.sh_l.y, r.sh_l.z, __uint_as_float(p));
        }
    }
}

// Whether bounce b starts the tail: the queues left (rays of bounce b and
// the light samples of bounce b - 1) are too small to fill the GPU stage by
// stage, and no earlier bounce started it.
__device__ __forceinline__ bool tail_starts(const FrameParams& fp, int b) {
    if (fp.frame->tail_from < static_cast<uint32_t>(b)) return false;
    return fp.ctrl[b - 1].n_shadow + fp.ctrl[b].n_rays < fp.tail_threshold;
}

// Tail, part 1: bounce b - 1's light samples, so that their radiance is in
// place before part 2 carries the paths on.
template <int S, bool INST>
__global__ __launch_bounds__(TRACE_BLOCK) void k_tail_shadow(FrameParams fp, int b) {
    if (!tail_starts(fp, b)) return;
    __shared__ uint2 lds[S * TRACE_BLOCK];
    auto st = lane_stack<S, TRACE_BLOCK>(lds, fp.spill);
    const uint32_t n = fp.ctrl[b - 1].n_shadow;
    uint32_t* work = &fp.ctrl[MAX_BOUNCES].w_unused;
    const uint32_t groups = fetch_groups(n);
    while (true) {
        const uint32_t base = wave_fetch(work, 64 * groups);
        if (base >= n) break;
        for (uint32_t g = 0; g < groups; ++g) {
            const uint32_t item = base + 64 * g + lane_id();
            if (item >= n) break;
            const float4 a = fp.so[item], bb = fp.sd[item];
            Hit h;
            if (!trace_ray<true, false, INST>(fp.sc, xyz(a), xyz(bb), a.w, bb.w, h, st)) {
                const float4 l = fp.sl[item];
                const uint32_t p = __float_as_uint(l.w);
                float4 r = fp.pb[p];
                r.x = r.x + l.x;
                r.y = r.y + l.y;
                r.z = r.z + l.z;
                fp.pb[p] = r;
            }
        }
    }
}

// Tail, part 2: every path still going at bounce b taken to its end, lane by
// lane: closest hit, shade, light sample, next bounce. The later bounces'
// stage launches then return at once. Rays are counted as the stages count
// them, into ctrl[MAX_BOUNCES].
template <int S, bool INST>
__global__ __launch_bounds__(TRACE_BLOCK) void k_tail(FrameParams fp, int b) {
    if (!tail_starts(fp, b)) return;
    Ctrl* c = fp.ctrl + b;
    const uint32_t n = c->n_rays;
    if (blockIdx.x == 0 && threadIdx.x == 0) fp.frame->tail_from = b;  // read by later launches only
    __shared__ uint2 lds[S * TRACE_BLOCK];
    auto st = lane_stack<S, TRACE_BLOCK>(lds, fp.spill);
    uint32_t rays = 0, shadows = 0, backfaces = 0;
    const int cur = b & 1;
    const uint32_t groups = fetch_groups(n);
    while (true) {
        const uint32_t base = wave_fetch(&c->w_extend, 64 * groups);
        if (base >= n) break;
        for (uint32_t g = 0; g < groups; ++g) {
            const uint32_t q = base + 64 * g + lane_id();
            if (q >= n) break;
            const uint32_t p = fp.rpath[cur][q];
            const float4 o4 = fp.ro[cur][q], d4 = fp.rd[cur][q];
            const float4 A = fp.pa[p], B = fp.pb[p];
            const uint2 kk = fp.pkey[p];
            PathState ps;
            ps.pq = fp.pq[p];
            ps.key = (static_cast<uint64_t>(kk.y) << 32) | kk.x;
            ps.thr = xyz(A);
            ps.rad = xyz(B);
            ps.cnt = __float_as_uint(A.w);
            ps.sampled = (__float_as_uint(B.w) & 1u) != 0;
            f3 o = xyz(o4), d = xyz(d4);
            float tmin = o4.w, tmax = d4.w;
            for (int bb = b; bb < fp.depth; ++bb) {
                Hit h;
                trace_ray<false, false, INST>(fp.sc, o, d, tmin, tmax, h, st);
                if (bb > b) ++rays;  // bounce b's rays were counted when they were queued
                const ShadeResult r = shade_hit(fp, bb, p, h, o, d, ps);
                if (r.backface) ++backfaces;
                if (r.shadow) {
                    ++shadows;
                    Hit hs;
                    if (!trace_ray<true, false, INST>(fp.sc, r.pnt, r.sh_d, 0.001f, r.sh_tmax, hs, st)) {
                        ps.rad.x = ps.rad.x + r.sh_l.x;
                        ps.rad.y = ps.rad.y + r.sh_l.y;
                        ps.rad.z = ps.rad.z + r.sh_l.z;
                    }
                }
                if (!r.cont) break;
                o = r.pnt;
                d = r.dir;
                tmin = T_MIN;
                tmax = T_MAX;
            }
            fp.pb[p] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, 0.0f);
        }
    }
    // Per wave, one atomic per count.
    uint32_t v[3] = {rays, shadows, backfaces};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    if (lane_id() == 0) {
        Ctrl* t = fp.ctrl + MAX_BOUNCES;
        if (v[0]) atomicAdd(&t->n_rays, v[0]);
        if (v[1]) atomicAdd(&t->n_shadow, v[1]);
        if (v[2]) atomicAdd(&t->n_backface, v[2]);
    }
}

// Per pixel, the batch's samples added in sample order onto the running sum
// (renderer.hpp:379-396). A pixel's samples are adjacent (spb float4), so a
// block first copies a tile of pixels' samples into LDS with coalesced loads,
// then one thread per pixel adds them up in order. The aux sums, and batches
// with more samples than a tile holds, take the direct loop.
constexpr uint32_t ACC_TILE = 1024;  // float4 per tile

__global__ __launch_bounds__(256) void k_accumulate(FrameParams fp) {
    // A pixel's row in the tile is padded by one float4 when spb >= 8 (then
    // at most ACC_TILE / 8 pixels per tile), so the adders, reading at a
    // stride of spb float4, spread over the LDS banks.
    __shared__ float4 tile[ACC_TILE + ACC_TILE / 8];
    const uint32_t spb = static_cast<uint32_t>(fp.spb);
    if (!fp.want_aux && spb <= ACC_TILE) {
        const uint32_t tp = ACC_TILE / spb;                 // pixels per tile
        const uint32_t pitch = spb + (spb >= 8 ? 1u : 0u);
        for (uint32_t px0 = blockIdx.x * tp; px0 < fp.npix; px0 += gridDim.x * tp) {
            const uint32_t np = min(tp, fp.npix - px0);
            const uint32_t count = np * spb;
            const float4* src = fp.pb + static_cast<size_t>(px0) * spb;
            for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) tile[(k / spb) * pitch + k % spb] = src[k];
            __syncthreads();
            for (uint32_t q = threadIdx.x; q < np; q += blockDim.x) {
                const uint32_t px = px0 + q;
                const float4 acc = fp.accum[px];
                f3 sum = xyz(acc);
                const float4* t = tile + q * pitch;
                for (uint32_t s = 0; s < spb; ++s) sum = sum + xyz(t[s]);
                fp.accum[px] = make_float4(sum.x, sum.y, sum.z, 0);
            }
            __syncthreads();
        }
        return;
    }
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < fp.npix; px += gridDim.x * blockDim.x) {
        float4 acc = fp.accum[px];
        f3 sum = xyz(acc);
        float4 a0 = make_float4(0, 0, 0, 0), a1 = make_float4(0, 0, 0, 0);
        if (fp.want_aux) {
            a0 = fp.aux_accum[2 * px];
            a1 = fp.aux_accum[2 * px + 1];
        }
        for (uint32_t s = 0; s < spb; ++s) {
            const uint32_t p = px * spb + s;
            const f3 c = xyz(fp.pb[p]);
            sum = sum + c;
            if (fp.want_aux) {
                const float4 fa = fp.paux[2 * p], fn = fp.paux[2 * p + 1];
                a0.x = a0.x + fa.x;
                a0.y = a0.y + fa.y;
                a0.z = a0.z + fa.z;
                a1.x = a1.x + fn.x;
                a1.y = a1.y + fn.y;
                a1.z = a1.z + fn.z;
                const float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;  // renderer.hpp:364
                a0.w += lum * lum;
            }
        }
        fp.accum[px] = make_float4(sum.x, sum.y, sum.z, 0);
        if (fp.want_aux) {
            fp.aux_accum[2 * px] = a0;
            fp.aux_accum[2 * px + 1] = a1;
        }
    }
}

__global__ void k_finish(FrameParams fp) {
    if (threadIdx.x != 0) return;
    unsigned long long rays = 0, shadow = 0;
    for (int b = 0; b <= fp.depth && b < MAX_BOUNCES; ++b) {
        rays += fp.ctrl[b].n_rays;
        shadow += fp.ctrl[b].n_shadow + fp.ctrl[b].n_backface;
    }
    const Ctrl& t = fp.ctrl[MAX_BOUNCES];  // the tail launch's counts
    rays += t.n_rays;
    shadow += t.n_shadow + t.n_backface;
    fp.totals[0] += rays;
    fp.totals[1] += shadow;
    fp.totals[2] += fp.ctrl[0].n_rays;
    for (int b = 0; b <= fp.depth && b < MAX_BOUNCES; ++b) fp.ctrl[b] = Ctrl{};
    fp.ctrl[MAX_BOUNCES] = Ctrl{};
    fp.frame->batch += 1;
    fp.frame->tail_from = ~0u;
}

__global__ __launch_bounds__(256) void k_clear_frame(FrameParams fp) {
    const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px < fp.npix) {
        fp.accum[px] = make_float4(0, 0, 0, 0);
        if (fp.want_aux) {
            fp.aux_accum[2 * px] = make_float4(0, 0, 0, 0);
            fp.aux_accum[2 * px + 1] = make_float4(0, 0, 0, 0);
        }
    }
    if (px == 0) {
        fp.frame->batch = 0;
        fp.frame->tail_from = ~0u;
        fp.totals[0] = fp.totals[1] = fp.totals[2] = 0;
        for (int b = 0; b <= fp.depth && b < MAX_BOUNCES; ++b) fp.ctrl[b] = Ctrl{};
        fp.ctrl[MAX_BOUNCES] = Ctrl{};
    }
}

// Averages into the caller's layout: 3 floats a pixel, the band's top row
// first (renderer.hpp:397-407).
__global__ __launch_bounds__(256) void k_resolve(FrameParams fp, float* rgb, float* aux) {
    const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= fp.npix) return;
    const float scale = 1.0f / static_cast<float>(fp.spp);
    const float4 acc = fp.accum[px];
    rgb[3 * px] = acc.x * scale;
    rgb[3 * px + 1] = acc.y * scale;
    rgb[3 * px + 2] = acc.z * scale;
    if (aux) {
        const float4 a0 = fp.aux_accum[2 * px], a1 = fp.aux_accum[2 * px + 1];
        float* o = aux + 7 * static_cast<size_t>(px);
        o[0] = a0.x * scale;
        o[1] = a0.y * scale;
        o[2] = a0.z * scale;
        o[3] = a1.x * scale;
        o[4] = a1.y * scale;
        o[5] = a1.z * scale;
        o[6] = a0.w * scale;
    }
}

template <int S, bool INST>
static void launch_loop(const FrameParams& fp, int trace_grid, int shade_grid, hipStream_t stream, hipEvent_t* ev) {
    const int acc_grid = shade_grid * 8;
    for (int b = 0; b < fp.depth; ++b) {
        if (b > 0) {
            k_tail_shadow<S, INST><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            k_tail<S, INST><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
        }
        if (ev) (void)hipEventRecord(ev[2 * b], stream);
        if (b == 0) {
            if (fp.visits) k_extend<S, true, INST, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            else k_extend<S, false, INST, true><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
        } else {
            if (fp.visits) k_extend<S, true, INST, false><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
            else k_extend<S, false, INST, false><<<trace_grid, TRACE_BLOCK, 0, stream>>>(fp, b);
        }
        if (ev) (void)hipEventRecord(ev[2 * b + 1], stream);
        if (b == 0) k_shade<true><<<shade_grid, SHADE_BLOCK, 0, stream>>>(fp, b);
        else k_shade<false><<<shade_grid, SHADE_BLOCK, 0, stream>>>(fp, b);
    }
    // No light samples are taken at the last bounce (renderer.hpp:127-130),
    // so no shadow rays are left over.
    k_accumulate<<<acc_grid, 256, 0, stream>>>(fp);
    k_finish<<<1, 64, 0, stream>>>(fp);
}

hipError_t launch_batch(const FrameParams& fp, int stack_entries, int trace_grid, int shade_grid, hipStream_t stream,
                        hipEvent_t* ev) {
    if (fp.depth < 1 || fp.depth >= MAX_BOUNCES) return hipErrorInvalidValue;
    const bool inst = fp.sc.inst_root != NONE;
    switch (stack_entries) {
        case 8:
            if (inst) launch_loop<8, true>(fp, trace_grid, shade_grid, stream, ev);
            else launch_loop<8, false>(fp, trace_grid, shade_grid, stream, ev);
            break;
        case 16:
            if (inst) launch_loop<16, true>(fp, trace_grid, shade_grid, stream, ev);
            else launch_loop<16, false>(fp, trace_grid, shade_grid, stream, ev);
            break;
        case 32:
            if (inst) launch_loop<32, true>(fp, trace_grid, shade_grid, stream, ev);
            else launch_loop<32, false>(fp, trace_grid, shade_grid, stream, ev);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_resolve(const FrameParams& fp, float* rgb, float* aux, hipStream_t stream) {
    k_resolve<<<(fp.npix + 255) / 256, 256, 0, stream>>>(fp, rgb, aux);
    return hipGetLastError();
}

hipError_t launch_clear_frame(const FrameParams& fp, hipStream_t stream) {
    k_clear_frame<<<(fp.npix + 255) / 256, 256, 0, stream>>>(fp);
    return hipGetLastError();
}

}  // namespace ptd
