// C ABI: batched renders (many jobs in one launch sequence) and the
// autotuner of their launch shape. Host code only.
//
// A batch holds one merged device scene (merge.cpp). A job is rendered by the
// same bounce-loop kernels as pt_render, over a view of the merged scene:
// its tree root, its runs of the large-sphere and light lists, its flags and
// camera. Jobs are spread over a few HIP streams (longest first, each to the
// least loaded stream); every stream's whole job sequence -- frame clears,
// bounce loops, resolves -- is one captured graph, launched once per render,
// and the streams run side by side, so one job's tail overlaps the next
// jobs' full queues. Each stream owns its ray queues, path state, control
// block and traversal spill area; each job its pixels' sums, its counters
// and its frame constants.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "api_internal.hpp"
#include "host.hpp"

using ptapi::kMaxPaths;
using ptapi::kPathBytes;

namespace {

constexpr int kMaxStreams = 8;

// One stream of a batch and what its kernels write.
struct batch_stream {
    hipStream_t stream = nullptr;
    ptapi::render_bufs rb;   // ray queues and path state, sized for the stream's largest job batch
    dbuf ctrl, spill;
    hipGraphExec_t exec = nullptr;
    std::vector<ptd::FrameParams> key;   // the launches the graph was captured for
    int S = 0, trace_grid = 0, shade_grid = 0, batches_key = 0;
    bool pool_key = false;               // the graph is a shared-queue sequence
    const float* out_key = nullptr;   // the output the graph's resolves write
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        rb.release();
        ctrl.release();
        spill.release();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

}  // namespace

struct pt_batch {
    pt_scene* scene = nullptr;
    std::vector<pt::merge_part> parts;
    std::vector<batch_stream> streams;
    dbuf frames, totals, accum, out, pool_tab;
    int tuned_stack = 0, tuned_blocks = 0, tuned_streams = 0;
    uint64_t tuned_paths = 0;
    ~pt_batch() {
        for (auto& s : streams) s.release();
        for (dbuf* b : {&frames, &totals, &accum, &out, &pool_tab}) b->release();
        if (scene) (void)pt_scene_destroy(scene);
    }
};

namespace {

std::vector<uint32_t> merge_checked(const uint32_t* const* ptgps, const size_t* n_words, int n,
                                    std::vector<pt::merge_part>& parts) {
    if (n < 1 || !ptgps || !n_words) throw pt::status_error(PT_ERR_INVALID, "batch: no scenes");
    std::vector<pt::ptgp_in> in(n);
    for (int k = 0; k < n; ++k) {
        if (!ptgps[k]) throw pt::status_error(PT_ERR_INVALID, "batch: null scene buffer");
        ptapi::check_ptgp(ptgps[k], n_words[k]);
        if (ptgps[k][19] != NONE) throw pt::status_error(PT_ERR_INVALID, "batch: scenes must have no instance tree");
        in[k] = {ptgps[k], n_words[k]};
    }
    try {
        return pt::merge_ptgp(in, parts);
    } catch (const std::runtime_error& e) {
        throw pt::status_error(PT_ERR_LIMIT, std::string("batch: ") + e.what());
    }
}

// The launch plan of one job: where its pixels go, its samples per batch,
// its stream.
struct job_plan {
    size_t npix = 0, pix_off = 0;
    int spb = 1, batches = 1, stream = 0;
    uint64_t paths = 0;   // npix * spp: the job's work
};

// Samples per pixel per batch: the largest divisor of spp whose batch fits
// cap paths (the kernels number a batch's samples first_sample + batch *
// spb + s, so spb divides spp).
int job_spb(int spp, size_t npix, uint64_t cap) {
    const uint64_t fit = std::max<uint64_t>(1, cap / std::max<size_t>(1, npix));
    int best = 1;
    for (int d = 1; d <= spp; ++d)
        if (spp % d == 0 && uint64_t(d) <= fit) best = d;
    return best;
}

void check_jobs(const pt_batch* b, const pt_batch_job* jobs, int n_jobs) {
    for (int j = 0; j < n_jobs; ++j) {
        const pt_batch_job& J = jobs[j];
        if (J.scene < 0 || J.scene >= static_cast<int>(b->parts.size()))
            throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: job scene out of range");
        if (J.width < 2 || J.height < 2 || J.width > 32768 || J.height > 32768)
            throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: width and height must be in [2, 32768]");
        if (J.spp < 1) throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: spp must be at least 1");
        if (J.first_sample < 0 || int64_t(J.first_sample) + J.spp >= (int64_t(1) << 31))
            throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: first_sample out of range");
        if (J.flags & ~(PT_RENDER_NO_NEE | PT_JOB_CAMERA))
            throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: job flags other than NO_NEE and JOB_CAMERA");
    }
}

// The job's view of the merged scene: its tree, large spheres, lights,
// flags and camera; the arrays themselves are shared.
ptd::SceneView job_view(const pt_batch* b, const pt_batch_job& J) {
    const pt::merge_part& part = b->parts[J.scene];
    ptd::SceneView v = b->scene->view;
    v.flat_root = b->scene->device_node(part.root);
    v.inst_root = NONE;
    v.large = b->scene->view.large + part.large_off;
    v.n_large = part.n_large;
    v.lights = b->scene->view.lights + part.lights_off;
    v.n_lights = part.n_lights;
    v.flags = part.flags & ((J.flags & PT_RENDER_NO_NEE) ? ~2u : ~0u);
    const float* cam = (J.flags & PT_JOB_CAMERA) ? J.camera : part.camera;
    for (int r = 0; r < 4; ++r) v.cam[r] = make_float4(cam[4 * r], cam[4 * r + 1], cam[4 * r + 2], 0.0f);
    return v;
}

// Paths a stream's batch (or a shared-queue pass) may hold when the caller
// and the tuner leave it open: a quarter of the free HBM shared by the
// streams, at most max_paths.
uint64_t default_paths(pt_batch* b, int n_streams, uint64_t max_paths) {
    size_t free_bytes = 0, total_bytes = 0;
    HIPC(hipMemGetInfo(&free_bytes, &total_bytes));
    for (auto& s : b->streams) free_bytes += s.rb.bytes();
    return std::max<uint64_t>(1, std::min<uint64_t>(max_paths, free_bytes / 4 / kPathBytes / n_streams));
}

struct render_shape {
    int S, per_cu, n_streams;
    uint64_t cap;
};

// Paths one shared-queue pass may hold: every job shares one set of queues,
// so they may take twice a single render's (2^28 paths, 52 GB of queues and
// path state).
constexpr uint64_t kMaxPoolPaths = uint64_t(kMaxPaths) * 2;

render_shape shape_for(pt_batch* b, const pt_batch_params& P, const pt_batch_job* jobs, int n_jobs) {
    render_shape r{};
    pt_scene* s = b->scene;
    const bool pool = !(P.flags & PT_BATCH_STREAMS);
    r.S = ptapi::stack_entries_for(s, P.stack_entries > 0 ? P.stack_entries : b->tuned_stack);
    r.per_cu = P.blocks_per_cu > 0 ? P.blocks_per_cu
               : b->tuned_blocks > 0 ? b->tuned_blocks
                                      : ptd::path_blocks_per_cu(r.S, false);
    if (r.per_cu > 64) throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: blocks_per_cu above 64");
    r.n_streams = P.streams > 0 ? P.streams : b->tuned_streams > 0 ? b->tuned_streams : std::min(2, n_jobs);
    r.n_streams = std::max(1, std::min({r.n_streams, n_jobs, kMaxStreams}));
    if (P.streams > kMaxStreams) throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: more than 8 streams");
    const uint64_t max_cap = pool ? kMaxPoolPaths : kMaxPaths;
    r.cap = P.max_paths ? P.max_paths
            : b->tuned_paths ? b->tuned_paths
            : std::min<uint64_t>(max_cap, default_paths(b, pool ? 1 : r.n_streams, max_cap));
    if (r.cap > max_cap)
        throw pt::status_error(PT_ERR_LIMIT, pool ? "pt_render_batch: max_paths above 2^28"
                                                  : "pt_render_batch: max_paths above 2^27 (per-job streams)");
    for (int j = 0; j < n_jobs; ++j) {
        const uint64_t np = uint64_t(jobs[j].width) * uint64_t(jobs[j].height);
        if (np > r.cap) {
            if (P.max_paths) throw pt::status_error(PT_ERR_LIMIT, "pt_render_batch: a job has more pixels than max_paths");
            r.cap = np;
        }
    }
    return r;
}

// A slice of a shared-queue pass: samples [s0, s0 + ns) of every pixel of job j.
struct slice_plan {
    int job, s0, ns;
};

// Shared-queue passes of at most cap paths: the jobs in order, each job's
// samples in order, split where a pass fills up, so a job's slices come in
// sample order across passes (and every pixel of a slice in one pass).
std::vector<std::vector<slice_plan>> plan_passes(const pt_batch_job* jobs, int n_jobs, const std::vector<job_plan>& jp,
                                                 uint64_t cap) {
    std::vector<std::vector<slice_plan>> passes(1);
    uint64_t room = cap;
    for (int j = 0; j < n_jobs; ++j) {
        const uint64_t npix = jp[j].npix;
        int s0 = 0;
        while (s0 < jobs[j].spp) {
            const uint64_t fit = room / npix;
            if (fit == 0) {   // npix <= cap (shape_for), so a fresh pass holds at least one sample
                passes.emplace_back();
                room = cap;
                continue;
            }
            const int ns = static_cast<int>(std::min<uint64_t>(uint64_t(jobs[j].spp - s0), fit));
            passes.back().push_back({j, s0, ns});
            room -= uint64_t(ns) * npix;
            s0 += ns;
        }
    }
    return passes;
}

// Every job in shared-queue passes on one stream: the pass's slices of all
// jobs go through one bounce loop, so the GPU stays full until the last pass
// and the queue tails of small jobs overlap with everything else. A job's
// image is pt_render's bit for bit: each path draws from its own (seed,
// pixel, sample) streams, and each pixel's samples are summed in sample
// order. The whole sequence -- clear, passes, the jobs' resolves -- is one
// captured graph, kept for later renders of the same plan; the slice tables
// it reads are uploaded before each launch.
void render_pool(pt_batch* b, const pt_batch_job* jobs, int n_jobs, const pt_batch_params& P, const render_shape& sh,
                 float* rgb, pt_render_stats* stats) {
    pt_scene* s = b->scene;
    const int trace_grid = s->cus * sh.per_cu;
    const int shade_grid = std::max(1, s->cus * 2);
    std::vector<job_plan> jp(n_jobs);
    size_t total_pix = 0;
    for (int j = 0; j < n_jobs; ++j) {
        jp[j].npix = size_t(jobs[j].width) * size_t(jobs[j].height);
        jp[j].pix_off = total_pix;
        total_pix += jp[j].npix;
    }
    if (total_pix >= (size_t(1) << 31)) throw pt::status_error(PT_ERR_LIMIT, "pt_render_batch: too many pixels");
    const std::vector<std::vector<slice_plan>> passes = plan_passes(jobs, n_jobs, jp, sh.cap);

    // The slice tables of every pass in one upload: per pass its slices,
    // then the first paths and first pixels of its slices (n + 1 each).
    std::vector<size_t> tab_off(passes.size());
    size_t tab_bytes = 0, max_paths = 1;
    for (size_t t = 0; t < passes.size(); ++t) {
        tab_off[t] = tab_bytes;
        const size_t n = passes[t].size();
        tab_bytes += n * sizeof(ptd::PoolSlice) + ((3 * (n + 1) * 4 + 127) & ~size_t(127));
    }
    std::vector<unsigned char> tab(tab_bytes, 0);
    const uint32_t n_nodes = static_cast<uint32_t>(s->node_map.size());
    int max_ns = 0;
    std::vector<uint32_t> pass_paths(passes.size()), pass_pix(passes.size());
    for (size_t t = 0; t < passes.size(); ++t) {
        const size_t n = passes[t].size();
        auto* sl = reinterpret_cast<ptd::PoolSlice*>(tab.data() + tab_off[t]);
        auto* spath = reinterpret_cast<uint32_t*>(tab.data() + tab_off[t] + n * sizeof(ptd::PoolSlice));
        uint32_t* spix = spath + n + 1;
        uint32_t* stile = spix + n + 1;
        uint64_t path = 0, pix = 0, tiles = 0;
        for (size_t k = 0; k < n; ++k) {
            const slice_plan& q = passes[t][k];
            const pt_batch_job& J = jobs[q.job];
            const ptd::SceneView v = job_view(b, J);
            ptd::PoolSlice& o = sl[k];
            o = ptd::PoolSlice{};
            for (int r = 0; r < 4; ++r) o.cam[r] = v.cam[r];
            o.seed = J.seed;
            o.W = static_cast<uint32_t>(J.width);
            o.H = static_cast<uint32_t>(J.height);
            o.ns = static_cast<uint32_t>(q.ns);
            o.index0 = static_cast<uint32_t>(J.first_sample + q.s0);
            o.root = v.flat_root;
            o.large_off = static_cast<uint32_t>(v.large - s->view.large);
            o.n_large = v.n_large;
            o.lights_off = static_cast<uint32_t>(v.lights - s->view.lights);
            o.n_lights = v.n_lights;
            o.flags = v.flags;
            o.pix_off = static_cast<uint32_t>(jp[q.job].pix_off);
            const uint64_t npix = jp[q.job].npix;
            o.tile_pix = static_cast<uint32_t>(
                uint64_t(q.ns) <= ptd::POOL_TILE ? std::min<uint64_t>(ptd::POOL_TILE / q.ns, npix) : 1);
            spath[k] = static_cast<uint32_t>(path);
            spix[k] = static_cast<uint32_t>(pix);
            stile[k] = static_cast<uint32_t>(tiles);
            // What the kernels rely on, checked before anything is launched:
            // every index a slice leads to lies in its array.
            const bool ok = q.ns >= 1 && J.width >= 2 && J.height >= 2 &&
                            (o.root == NONE || o.root < n_nodes) &&
                            uint64_t(o.large_off) + o.n_large <= s->view.n_large &&
                            uint64_t(o.lights_off) + o.n_lights <= s->view.n_lights &&
                            uint64_t(o.pix_off) + jp[q.job].npix <= total_pix && o.tile_pix >= 1 &&
                            (uint64_t(o.tile_pix) * o.ns <= ptd::POOL_TILE || o.tile_pix == 1);
            if (!ok) throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: internal error in a slice table");
            path += npix * uint64_t(q.ns);
            pix += npix;
            tiles += (npix + o.tile_pix - 1) / o.tile_pix;
            max_ns = std::max(max_ns, q.ns);
        }
        if (n == 0 || path > sh.cap || path >= (uint64_t(1) << 32))
            throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: internal error in a pass plan");
        spath[n] = static_cast<uint32_t>(path);
        spix[n] = static_cast<uint32_t>(pix);
        stile[n] = static_cast<uint32_t>(tiles);
        pass_paths[t] = static_cast<uint32_t>(path);
        pass_pix[t] = static_cast<uint32_t>(pix);
        max_paths = std::max<size_t>(max_paths, path);
    }

    // Buffers: the first stream's queues and path state, sized for the
    // largest pass; the jobs' sums, one set of counters and frame constants.
    if (b->streams.empty()) b->streams.resize(1);
    batch_stream& st = b->streams[0];
    if (!st.stream) HIPC(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
    ptapi::render_bufs& B = st.rb;
    for (int q = 0; q < 2; ++q) {
        B.ro[q].ensure(max_paths * 16);
        B.rd[q].ensure(max_paths * 16);
        B.rpath[q].ensure(max_paths * 4);
    }
    B.hits.ensure(max_paths * sizeof(ptd::Hit));
    B.so.ensure(max_paths * 16);
    B.sd.ensure(max_paths * 16);
    B.sl.ensure(max_paths * 16);
    B.pa.ensure(max_paths * 16);
    B.pb.ensure(max_paths * 16);
    B.pkey.ensure(max_paths * 8);
    B.pq.ensure(max_paths * 8);
    st.ctrl.ensure(sizeof(ptd::Ctrl) * (ptd::MAX_BOUNCES + 1));
    st.spill.ensure(size_t(trace_grid) * ptd::TRACE_BLOCK * size_t(s->max_stack) * sizeof(uint2));
    b->frames.ensure(sizeof(ptd::FrameConst));
    b->totals.ensure(4 * sizeof(unsigned long long));
    b->accum.ensure(total_pix * 16);
    b->out.ensure(total_pix * 12);
    b->pool_tab.ensure(std::max<size_t>(tab_bytes, 128));
    const ptd::FrameConst fc0{0, 0, 0u, ~0u, 0u};
    HIPC(hipMemcpy(b->frames.p, &fc0, sizeof fc0, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(b->pool_tab.p, tab.data(), tab_bytes, hipMemcpyHostToDevice));

    const bool device_out = (P.flags & PT_RENDER_DEVICE_OUTPUT) != 0;
    float* out = device_out ? rgb : b->out.as<float>();
    ptd::FrameParams base{};
    base.sc = s->view;
    base.depth = P.max_bounces;
    base.spb = 1;
    for (int q = 0; q < 2; ++q) {
        base.ro[q] = B.ro[q].as<float4>();
        base.rd[q] = B.rd[q].as<float4>();
        base.rpath[q] = B.rpath[q].as<uint32_t>();
    }
    base.hits = B.hits.as<ptd::Hit>();
    base.so = B.so.as<float4>();
    base.sd = B.sd.as<float4>();
    base.sl = B.sl.as<float4>();
    base.pa = B.pa.as<float4>();
    base.pb = B.pb.as<float4>();
    base.pkey = B.pkey.as<uint2>();
    base.pq = B.pq.as<uint2>();
    base.accum = b->accum.as<float4>();
    base.ctrl = st.ctrl.as<ptd::Ctrl>();
    base.frame = b->frames.as<ptd::FrameConst>();
    base.totals = b->totals.as<unsigned long long>();
    base.spill = st.spill.as<uint2>();
    base.tail_threshold = (P.flags & PT_RENDER_NO_TAIL) ? 0u : ptapi::tail_threshold(trace_grid, false);
    base.camera_packets = (P.flags & PT_RENDER_PER_RAY) ? 0u : 1u;
    // The launches: clear every sum, the passes, each job's resolve.
    std::vector<ptd::FrameParams> launches;
    ptd::FrameParams clear = base;
    clear.npix = static_cast<uint32_t>(total_pix);
    launches.push_back(clear);
    for (size_t t = 0; t < passes.size(); ++t) {
        ptd::FrameParams fp = base;
        const size_t n = passes[t].size();
        const unsigned char* at = b->pool_tab.as<unsigned char>() + tab_off[t];
        fp.slices = reinterpret_cast<const ptd::PoolSlice*>(at);
        fp.slice_path = reinterpret_cast<const uint32_t*>(at + n * sizeof(ptd::PoolSlice));
        fp.slice_pix = fp.slice_path + n + 1;
        fp.slice_tile = fp.slice_pix + n + 1;
        fp.n_slices = static_cast<uint32_t>(n);
        fp.P = pass_paths[t];
        fp.npix = pass_pix[t];
        launches.push_back(fp);
    }
    for (int j = 0; j < n_jobs; ++j) {
        ptd::FrameParams r = base;
        r.accum = b->accum.as<float4>() + jp[j].pix_off;
        r.npix = static_cast<uint32_t>(jp[j].npix);
        r.spp = jobs[j].spp;
        launches.push_back(r);
    }
    auto enqueue = [&]() -> hipError_t {
        hipError_t e = ptd::launch_clear_frame(launches[0], st.stream);
        for (size_t t = 0; e == hipSuccess && t < passes.size(); ++t)
            e = ptd::launch_pool_pass(launches[1 + t], sh.S, trace_grid, shade_grid, st.stream);
        for (int j = 0; e == hipSuccess && j < n_jobs; ++j)
            e = ptd::launch_resolve(launches[1 + passes.size() + j], out + 3 * jp[j].pix_off, nullptr, st.stream);
        return e;
    };
    const bool graph = !(P.flags & PT_RENDER_NO_GRAPH);
    if (graph) {
        const bool reuse = st.exec && st.pool_key && st.S == sh.S && st.trace_grid == trace_grid &&
                           st.shade_grid == shade_grid && st.out_key == out && st.key.size() == launches.size() &&
                           std::memcmp(st.key.data(), launches.data(), launches.size() * sizeof(ptd::FrameParams)) == 0;
        if (!reuse) {
            st.drop_graph();
            hipGraph_t g;
            HIPC(hipStreamBeginCapture(st.stream, hipStreamCaptureModeThreadLocal));
            const hipError_t le = enqueue();
            const hipError_t ce = hipStreamEndCapture(st.stream, &g);
            HIPC(le);
            HIPC(ce);
            const hipError_t ie = hipGraphInstantiate(&st.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) st.exec = nullptr;
            HIPC(ie);
            st.key = launches;
            st.pool_key = true;
            st.S = sh.S;
            st.trace_grid = trace_grid;
            st.shade_grid = shade_grid;
            st.out_key = out;
        }
    }
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipEventRecord(e0, st.stream));
    if (graph) HIPC(hipGraphLaunch(st.exec, st.stream));
    else HIPC(enqueue());
    HIPC(hipEventRecord(e1, st.stream));
    HIPC(hipEventSynchronize(e1));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!device_out) HIPC(hipMemcpy(rgb, b->out.p, total_pix * 12, hipMemcpyDeviceToHost));
    unsigned long long tot[4] = {};
    HIPC(hipMemcpy(tot, b->totals.p, sizeof tot, hipMemcpyDeviceToHost));
    if (stats) {
        *stats = pt_render_stats{};
        stats->rays = tot[0] + tot[1];
        stats->shadow_rays = tot[1];
        stats->primary_rays = tot[2];
        stats->batches = static_cast<int32_t>(passes.size());
        stats->samples_per_batch = max_ns;
        stats->render_ms = ms;
    }
}

}  // namespace

extern "C" {

pt_status pt_batch_merge(const uint32_t* const* ptgps, const size_t* n_words, int n, uint32_t** merged,
                         size_t* merged_words, pt_batch_part* parts) {
    if (!merged || !merged_words) return pt::fail(PT_ERR_INVALID, "pt_batch_merge: null output");
    *merged = nullptr;
    *merged_words = 0;
    PT_GUARD({
        std::vector<pt::merge_part> P;
        const std::vector<uint32_t> w = merge_checked(ptgps, n_words, n, P);
        if (parts) {
            for (int k = 0; k < n; ++k) {
                pt_batch_part& o = parts[k];
                o.root = P[k].root;
                o.n_large = P[k].n_large;
                o.large_off = P[k].large_off;
                o.n_lights = P[k].n_lights;
                o.lights_off = P[k].lights_off;
                o.flags = P[k].flags;
                std::memcpy(o.camera, P[k].camera, sizeof o.camera);
            }
        }
        uint32_t* p = static_cast<uint32_t*>(std::malloc(w.size() * 4 + 4));
        if (!p) throw std::bad_alloc();
        std::memcpy(p, w.data(), w.size() * 4);
        *merged = p;
        *merged_words = w.size();
    });
}

pt_status pt_batch_create(const uint32_t* const* ptgps, const size_t* n_words, int n, pt_batch** out) {
    if (!out) return pt::fail(PT_ERR_INVALID, "pt_batch_create: null output");
    *out = nullptr;
    PT_GUARD({
        std::vector<pt::merge_part> parts;
        const std::vector<uint32_t> w = merge_checked(ptgps, n_words, n, parts);
        pt_scene* s = nullptr;
        const pt_status st = pt_scene_create(w.data(), w.size(), &s);
        if (st != PT_OK) throw pt::status_error(st, pt_last_error());
        pt_batch* b = new pt_batch();
        b->scene = s;
        b->parts = std::move(parts);
        *out = b;
    });
}

pt_status pt_batch_destroy(pt_batch* b) {
    if (!b) return PT_OK;
    PT_GUARD(delete b);
}

pt_status pt_render_batch(pt_batch* b, const pt_batch_job* jobs, int n_jobs, const pt_batch_params* prm, float* rgb,
                          pt_render_stats* stats) {
    if (!b || !jobs || !prm || !rgb || n_jobs < 1) return pt::fail(PT_ERR_INVALID, "pt_render_batch: null argument");
    PT_GUARD({
        const pt_batch_params& P = *prm;
        pt_scene* s = b->scene;
        if (P.max_bounces < 1 || P.max_bounces >= ptd::MAX_BOUNCES)
            throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: max_bounces must be in [1, 63]");
        if (P.flags & ~(PT_RENDER_NO_GRAPH | PT_RENDER_DEVICE_OUTPUT | PT_RENDER_NO_TAIL | PT_BATCH_STREAMS |
                        PT_RENDER_PER_RAY))
            throw pt::status_error(PT_ERR_INVALID, "pt_render_batch: flags other than NO_GRAPH, DEVICE_OUTPUT, "
                                                   "NO_TAIL, BATCH_STREAMS, PER_RAY");
        check_jobs(b, jobs, n_jobs);
        const render_shape sh = shape_for(b, P, jobs, n_jobs);
        if (!(P.flags & PT_BATCH_STREAMS)) {
            render_pool(b, jobs, n_jobs, P, sh, rgb, stats);
            return PT_OK;
        }
        const int trace_grid = s->cus * sh.per_cu;
        const int shade_grid = std::max(1, s->cus * 2);

        // Per job: pixels, samples per batch; longest jobs first, each to the
        // least loaded stream.
        std::vector<job_plan> jp(n_jobs);
        size_t total_pix = 0;
        for (int j = 0; j < n_jobs; ++j) {
            jp[j].npix = size_t(jobs[j].width) * size_t(jobs[j].height);
            jp[j].pix_off = total_pix;
            total_pix += jp[j].npix;
            jp[j].spb = job_spb(jobs[j].spp, jp[j].npix, sh.cap);
            jp[j].batches = jobs[j].spp / jp[j].spb;
            jp[j].paths = uint64_t(jp[j].npix) * uint64_t(jobs[j].spp);
        }
        if (total_pix >= (size_t(1) << 31)) throw pt::status_error(PT_ERR_LIMIT, "pt_render_batch: too many pixels");
        std::vector<int> order(n_jobs);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return jp[x].paths > jp[y].paths; });
        std::vector<uint64_t> load(sh.n_streams, 0);
        std::vector<std::vector<int>> per_stream(sh.n_streams);
        std::vector<size_t> stream_paths(sh.n_streams, 0);
        for (int j : order) {
            const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
            jp[j].stream = k;
            load[k] += jp[j].paths;
            per_stream[k].push_back(j);
            stream_paths[k] = std::max(stream_paths[k], jp[j].npix * size_t(jp[j].spb));
        }

        // Buffers: per stream the queues and path state, per job its sums,
        // counters and frame constants.
        if (static_cast<int>(b->streams.size()) < sh.n_streams) b->streams.resize(sh.n_streams);
        for (int k = 0; k < sh.n_streams; ++k) {
            batch_stream& st = b->streams[k];
            if (!st.stream) HIPC(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
            const size_t paths = stream_paths[k];
            ptapi::render_bufs& B = st.rb;
            for (int q = 0; q < 2; ++q) {
                B.ro[q].ensure(paths * 16);
                B.rd[q].ensure(paths * 16);
                B.rpath[q].ensure(paths * 4);
            }
            B.hits.ensure(paths * sizeof(ptd::Hit));
            B.so.ensure(paths * 16);
            B.sd.ensure(paths * 16);
            B.sl.ensure(paths * 16);
            B.pa.ensure(paths * 16);
            B.pb.ensure(paths * 16);
            B.pkey.ensure(paths * 8);
            B.pq.ensure(paths * 8);
            st.ctrl.ensure(sizeof(ptd::Ctrl) * (ptd::MAX_BOUNCES + 1));
            st.spill.ensure(size_t(trace_grid) * ptd::TRACE_BLOCK * size_t(s->max_stack) * sizeof(uint2));
        }
        b->frames.ensure(size_t(n_jobs) * sizeof(ptd::FrameConst));
        b->totals.ensure(size_t(n_jobs) * 4 * sizeof(unsigned long long));
        b->accum.ensure(total_pix * 16);
        b->out.ensure(total_pix * 12);
        std::vector<ptd::FrameConst> fc(n_jobs);
        for (int j = 0; j < n_jobs; ++j) fc[j] = ptd::FrameConst{jobs[j].seed, jobs[j].first_sample, 0u, ~0u, 0u};
        HIPC(hipMemcpy(b->frames.p, fc.data(), fc.size() * sizeof(ptd::FrameConst), hipMemcpyHostToDevice));

        const bool device_out = (P.flags & PT_RENDER_DEVICE_OUTPUT) != 0;
        float* out = device_out ? rgb : b->out.as<float>();
        // The launches of one job, as pt_render makes them for it.
        auto job_params = [&](int j) {
            const batch_stream& st = b->streams[jp[j].stream];
            const ptapi::render_bufs& B = st.rb;
            ptd::FrameParams fp{};
            fp.sc = job_view(b, jobs[j]);
            fp.W = jobs[j].width;
            fp.H = jobs[j].height;
            fp.spp = jobs[j].spp;
            fp.depth = P.max_bounces;
            fp.j_lo = 0;
            fp.j_hi = jobs[j].height;
            fp.spb = jp[j].spb;
            fp.P = static_cast<uint32_t>(jp[j].npix * size_t(jp[j].spb));
            fp.npix = static_cast<uint32_t>(jp[j].npix);
            fp.nee = (fp.sc.flags & 2u) ? 1 : 0;
            for (int q = 0; q < 2; ++q) {
                fp.ro[q] = B.ro[q].as<float4>();
                fp.rd[q] = B.rd[q].as<float4>();
                fp.rpath[q] = B.rpath[q].as<uint32_t>();
            }
            fp.hits = B.hits.as<ptd::Hit>();
            fp.so = B.so.as<float4>();
            fp.sd = B.sd.as<float4>();
            fp.sl = B.sl.as<float4>();
            fp.pa = B.pa.as<float4>();
            fp.pb = B.pb.as<float4>();
            fp.pkey = B.pkey.as<uint2>();
            fp.pq = B.pq.as<uint2>();
            fp.accum = b->accum.as<float4>() + jp[j].pix_off;
            fp.ctrl = st.ctrl.as<ptd::Ctrl>();
            fp.frame = b->frames.as<ptd::FrameConst>() + j;
            fp.totals = b->totals.as<unsigned long long>() + 4 * size_t(j);
            fp.spill = st.spill.as<uint2>();
            fp.tail_threshold =
                (P.flags & PT_RENDER_NO_TAIL) ? 0u : ptapi::tail_threshold(trace_grid, false);   // batched scenes are flat
            fp.camera_packets = (P.flags & PT_RENDER_PER_RAY) ? 0u : 1u;
            return fp;
        };
        // A stream's whole sequence: for each of its jobs, clear, the
        // batches' bounce loops, resolve into the output.
        auto enqueue = [&](int k, const std::vector<ptd::FrameParams>& fps) -> hipError_t {
            const hipStream_t strm = b->streams[k].stream;
            for (size_t q = 0; q < fps.size(); ++q) {
                const int j = per_stream[k][q];
                const ptd::FrameParams& fp = fps[q];
                hipError_t e = ptd::launch_clear_frame(fp, strm);
                for (int t = 0; e == hipSuccess && t < jp[j].batches; ++t)
                    e = ptd::launch_batch(fp, sh.S, trace_grid, shade_grid, strm, nullptr);
                if (e == hipSuccess) e = ptd::launch_resolve(fp, out + 3 * jp[j].pix_off, nullptr, strm);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        };

        std::vector<std::vector<ptd::FrameParams>> fps(sh.n_streams);
        for (int k = 0; k < sh.n_streams; ++k)
            for (int j : per_stream[k]) fps[k].push_back(job_params(j));
        const bool graph = !(P.flags & PT_RENDER_NO_GRAPH);
        if (graph) {
            for (int k = 0; k < sh.n_streams; ++k) {
                batch_stream& st = b->streams[k];
                int batches = 0;
                for (int j : per_stream[k]) batches = batches * 131 + jp[j].batches;
                const bool reuse = st.exec && !st.pool_key && st.S == sh.S && st.trace_grid == trace_grid &&
                                   st.shade_grid == shade_grid && st.batches_key == batches && st.out_key == out &&
                                   st.key.size() == fps[k].size() &&
                                   std::memcmp(st.key.data(), fps[k].data(), fps[k].size() * sizeof(ptd::FrameParams)) == 0;
                if (reuse) continue;
                st.drop_graph();
                hipGraph_t g;
                HIPC(hipStreamBeginCapture(st.stream, hipStreamCaptureModeThreadLocal));
                const hipError_t le = enqueue(k, fps[k]);
                const hipError_t ce = hipStreamEndCapture(st.stream, &g);
                HIPC(le);
                HIPC(ce);
                const hipError_t ie = hipGraphInstantiate(&st.exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess) st.exec = nullptr;
                HIPC(ie);
                st.key = fps[k];
                st.pool_key = false;
                st.S = sh.S;
                st.trace_grid = trace_grid;
                st.shade_grid = shade_grid;
                st.batches_key = batches;
                st.out_key = out;
            }
        }
        // All streams start after one event and each records its end.
        hipEvent_t e0;
        HIPC(hipEventCreate(&e0));
        std::vector<hipEvent_t> e1(sh.n_streams);
        for (auto& e : e1) HIPC(hipEventCreate(&e));
        HIPC(hipEventRecord(e0, b->streams[0].stream));
        for (int k = 0; k < sh.n_streams; ++k) {
            batch_stream& st = b->streams[k];
            if (k) HIPC(hipStreamWaitEvent(st.stream, e0, 0));
            if (graph) HIPC(hipGraphLaunch(st.exec, st.stream));
            else HIPC(enqueue(k, fps[k]));
            HIPC(hipEventRecord(e1[k], st.stream));
        }
        float render_ms = 0;
        for (int k = 0; k < sh.n_streams; ++k) {
            HIPC(hipEventSynchronize(e1[k]));
            float ms = 0;
            HIPC(hipEventElapsedTime(&ms, e0, e1[k]));
            render_ms = std::max(render_ms, ms);
        }
        (void)hipEventDestroy(e0);
        for (auto& e : e1) (void)hipEventDestroy(e);

        if (!device_out) HIPC(hipMemcpy(rgb, b->out.p, total_pix * 12, hipMemcpyDeviceToHost));
        std::vector<unsigned long long> tot(size_t(n_jobs) * 4);
        HIPC(hipMemcpy(tot.data(), b->totals.p, tot.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        if (stats) {
            *stats = pt_render_stats{};
            for (int j = 0; j < n_jobs; ++j) {
                stats->rays += tot[4 * j] + tot[4 * j + 1];
                stats->shadow_rays += tot[4 * j + 1];
                stats->primary_rays += tot[4 * j + 2];
                stats->batches += jp[j].batches;
                stats->samples_per_batch = std::max(stats->samples_per_batch, jp[j].spb);
            }
            stats->render_ms = render_ms;
        }
    });
}

pt_status pt_batch_autotune(pt_batch* b, const pt_batch_job* jobs, int n_jobs, int max_bounces, pt_tune_result* result) {
    if (!b || !jobs || n_jobs < 1) return pt::fail(PT_ERR_INVALID, "pt_batch_autotune: null argument");
    PT_GUARD({
        check_jobs(b, jobs, n_jobs);
        // Each candidate renders the jobs as given: small renders are bound
        // by launch latency and queue tails and would pick a shape that
        // starves the full render (a caller wanting a quicker tune passes
        // jobs with fewer samples).
        std::vector<pt_batch_job> short_jobs(jobs, jobs + n_jobs);
        size_t total = 0;
        for (auto& j : short_jobs) total += size_t(j.width) * size_t(j.height) * 3;
        std::vector<float> img(total);
        struct cand {
            int S, per_cu, n_streams;
            uint64_t paths;
            double mrays;
        };
        std::vector<cand> tried;
        auto rate = [&](int S, int per_cu, int n_streams, uint64_t paths) {
            pt_batch_params p{};
            p.max_bounces = max_bounces;
            p.stack_entries = S;
            p.blocks_per_cu = per_cu;
            p.streams = n_streams;
            p.max_paths = paths;
            pt_render_stats st{};
            // the first render at a new shape captures its graphs; time the second
            for (int k = 0; k < 2; ++k) {
                const pt_status e = pt_render_batch(b, short_jobs.data(), n_jobs, &p, img.data(), &st);
                if (e != PT_OK) throw pt::status_error(e, pt_last_error());
            }
            const double r = st.render_ms > 0 ? double(st.rays) / (st.render_ms * 1e3) : 0.0;
            tried.push_back({S, per_cu, n_streams, paths, r});
        };
        auto best = [&] {
            return *std::max_element(tried.begin(), tried.end(), [](const cand& x, const cand& y) { return x.mrays < y.mrays; });
        };
        // Stage 1: LDS stack entries x traversal blocks per CU (shared-queue
        // passes, which use one stream).
        for (int S : {8, 16, 32}) {
            const int occ = ptd::path_blocks_per_cu(S, false);
            std::vector<int> blocks{occ};
            if (occ > 2) blocks.push_back(occ - 1);
            if (occ > 4) blocks.push_back(occ / 2);
            for (int pc : blocks) rate(S, pc, 1, 0);
        }
        cand c = best();
        // Stage 2: smaller passes (queue sizes) at the best shape.
        for (int e : {25, 26, 27}) rate(c.S, c.per_cu, 1, uint64_t(1) << e);
        c = best();
        b->tuned_stack = c.S;
        b->tuned_blocks = c.per_cu;
        b->tuned_paths = c.paths;
        if (result) {
            *result = pt_tune_result{};
            result->stack_entries = c.S;
            result->blocks_per_cu = c.per_cu;
            result->streams = c.n_streams;
            result->max_paths = c.paths;
            result->n_tried = static_cast<int32_t>(std::min<size_t>(tried.size(), 32));
            for (int k = 0; k < result->n_tried; ++k) {
                result->tried[k].stack_entries = tried[k].S;
                result->tried[k].blocks_per_cu = tried[k].per_cu;
                result->tried[k].streams = tried[k].n_streams;
                result->tried[k].max_paths = tried[k].paths;
                result->tried[k].mrays_per_s = tried[k].mrays;
            }
        }
    });
}

}  // extern "C"
