/*
 * photon_tracer.h -- C ABI of the MI355X path-tracing core.
 *
 * The core is a drop-in for the render path of the reference CPU path tracer
 * (apollo-2006/photon_tracer). Every entry point below replaces one call of
 * that path; the comment on each names the reference interface it stands in
 * for (file:line in the reference tree).
 *
 * Conventions
 *   - Plain C: pointers and sizes, no C++ types, no exceptions across the ABI.
 *   - Every function returns a pt_status; on failure pt_last_error() gives the
 *     message (thread-local), where the reference throws std::runtime_error
 *     (gpu_pack.hpp:62,86) and its CLI returns 1 (main.cpp:103-108).
 *   - The caller owns host arrays. A pt_scene owns its device buffers, freed by
 *     pt_scene_destroy. Output goes to caller memory.
 *   - One HIP stream per scene handle: calls on one handle from several host
 *     threads must be serialised (the reference is reentrant per thread over a
 *     const geometry, renderer.hpp:17-19).
 *
 * Data formats
 *   PTSD ("scene description") is this library's own word buffer holding the
 *   reference `geometry` flat arrays (geometry.hpp:15-58) and the camera
 *   parameters (camera.hpp:18-34). Layout in pt_layout comments below.
 *   PTGP v1 is the reference's packed GPU scene, byte for byte
 *   (gpu_pack.hpp:21-33,77-82,177-203): 48-word header, 8 sections, 16-byte
 *   aligned. The reference drops textures (gpu_pack.hpp:96); this library
 *   carries them in words the reference writes as zero, so an untextured
 *   scene's buffer is unchanged: header words 38-40 give the byte offset,
 *   length and count of a texture section (per texture width, height, first
 *   texel, 0; then linear RGB texels as 4 floats, rows top first), and word
 *   9 of a matte material's 12 holds its texture's number + 1 (0: none).
 *   pt_pack produces PTGP from PTSD with this library's own binned-SAH
 *   4-wide builder, which emits the reference's node layout (bvh.hpp:27-32),
 *   without the reference's 10-level depth cap (gpu_pack.hpp:36,61-62).
 */
#ifndef PHOTON_TRACER_H
#define PHOTON_TRACER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PT_OK = 0,
    PT_ERR_INVALID = 1,   /* bad argument or malformed buffer */
    PT_ERR_DEVICE = 2,    /* a HIP call failed */
    PT_ERR_NOMEM = 3,     /* host or device allocation failed */
    PT_ERR_LIMIT = 4      /* a size beyond what the library supports */
} pt_status;

typedef struct pt_scene pt_scene;

/* Message of the last failure on this thread ("" if none). */
const char* pt_last_error(void);

/* Library version string and the device it runs on (name, CU count). */
const char* pt_version(void);
pt_status pt_device_info(char* name, size_t name_len, int* compute_units);

/* Frees a buffer returned by pt_generate / pt_pack. */
void pt_free(void* p);

/* ---------------------------------------------------------------------------
 * Scene description and packing (host side).
 * ------------------------------------------------------------------------- */

/* Seeded procedural scenes of the benchmark configurations.
 *   config 0: sphere field, 1: displaced icosphere mesh, 2: instanced meshes,
 *   4: closed room with a displaced mesh and lights behind a partition.
 *   (3, the batch, is a list of config-0/1 scenes: see pt_generate_job.)
 * `scale` shrinks the triangle counts for tests (1.0 = the full sizes).
 * Replaces the reference's scene builders (renderer.hpp:158-343) for scenes
 * the reference does not have. Returns a PTSD buffer (free with pt_free). */
pt_status pt_generate(int config, uint64_t seed, double scale, uint32_t** ptsd, size_t* n_words);

/* One job of the batched workload: a config-0 or config-1 type scene with a
 * seeded primitive count, and its resolution and spp. */
typedef struct {
    int32_t kind;          /* 0 spheres, 1 triangle mesh */
    int32_t width, height;
    int32_t spp;
    uint32_t primitives;   /* sphere or triangle count */
    uint64_t seed;
} pt_job_desc;
pt_status pt_generate_job(uint64_t seed, double scale, pt_job_desc* desc, uint32_t** ptsd, size_t* n_words);

/* load_obj (obj.hpp:114-218): the faces of OBJ text (polygons split into
 * fans; v, vt, vn and negative indices), turned turn_deg about the vertical
 * axis, scaled to `height` and set down with the bottom of their bounding box
 * centred on `base`, appended to a PTSD scene's triangles. Normals from the
 * file where every corner of a face has one, otherwise area-weighted vertex
 * normals; degenerate faces dropped. In place of the reference's file reader,
 * mtl_text (may be NULL) is the text of the MTL library that every mtllib
 * line names: usemtl picks its materials, mapped as make_material maps them
 * (obj.hpp:89-112) and appended to the scene's materials; faces before any
 * usemtl, or naming a material the library lacks, take `material` (an index
 * into the scene's materials). Texture maps are left out here (the
 * materials are untextured, as the reference makes them when its reader
 * finds no image; pt_ptsd_add_obj_textured takes the images). Returns a new
 * PTSD buffer (free with pt_free). */
pt_status pt_ptsd_add_obj(const uint32_t* ptsd, size_t n_words, const char* obj_text, size_t text_len,
                          const char* mtl_text, size_t mtl_len, const float base[3], double height, uint32_t material,
                          double turn_deg, uint32_t** out, size_t* out_words);

/* An 8-bit sRGB image as the reference's image8 holds one (image_io.hpp:
 * rows top first, 3 bytes a texel), named as an MTL map_Kd line names it. */
typedef struct {
    const char* name;
    uint32_t width, height;     /* 1 .. 65536 */
    const uint8_t* rgb;         /* width * height * 3 bytes */
} pt_image;

/* decode_image (image_io.hpp:295): a PNG or PPM file's bytes (told apart by
 * their first bytes) as 8-bit sRGB, 3 bytes a pixel, rows top first, for
 * pt_image. PNG: not interlaced, 8 or 16 bits a sample (16-bit samples keep
 * their high byte), gray, gray + alpha, RGB, RGBA or an 8-bit palette, alpha
 * dropped (image_io.hpp:185-259); PPM: P3 or P6, maximum value 1 to 255,
 * scaled to 255 (image_io.hpp:262-291). 1 to 65536 pixels a side. rgb is
 * allocated by the library (free with pt_free). PT_ERR_INVALID with the
 * reason for what it cannot read. */
pt_status pt_decode_image(const void* bytes, size_t n_bytes, uint32_t* width, uint32_t* height, uint8_t** rgb);

/* pt_ptsd_add_obj with texture maps: a matte MTL material whose map_Kd names
 * one of `images` (by the map_Kd line's last word, the file name the
 * reference's file reader opens) gets that image as its texture, so its
 * albedo is Kd times the bilinear, repeating, sRGB-decoded texture at the
 * hit's texture coordinates (load_obj + make_material, obj.hpp:103-111;
 * image_texture, texture.hpp:13-32). In place of the reference's file reader
 * the caller passes the images, decoded (pt_decode_image does what the
 * reference's decode_image does). Each image used is appended to the scene's
 * textures once; images no map_Kd names are not added. */
pt_status pt_ptsd_add_obj_textured(const uint32_t* ptsd, size_t n_words, const char* obj_text, size_t text_len,
                                   const char* mtl_text, size_t mtl_len, const pt_image* images, size_t n_images,
                                   const float base[3], double height, uint32_t material, double turn_deg,
                                   uint32_t** out, size_t* out_words);

/* Builds every BVH of a PTSD scene (binned SAH, 16 bins, max leaf 8, 4-wide
 * collapse; bvh.hpp:39-52,116-247; geometry.hpp:74-104) and packs the scene
 * as PTGP v1 (pack_for_gpu + serialize, gpu_pack.hpp:84-203). */
pt_status pt_pack(const uint32_t* ptsd, size_t n_words, uint32_t** ptgp, size_t* ptgp_words);

/* pt_pack with options. PT_PACK_DEVICE_BVH builds every tree (mesh BLASes,
 * the flat tree, the instance TLAS) with the GPU builder of pt_bvh_build; the
 * PTGP words are the same as pt_pack's. */
enum { PT_PACK_DEVICE_BVH = 1 };
pt_status pt_pack_ex(const uint32_t* ptsd, size_t n_words, int flags, uint32_t** ptgp, size_t* ptgp_words);

/* bvh::build over caller boxes (bvh.hpp:39-52): binned SAH with 16 bins,
 * unit costs, max leaf 8, max depth 64, then the widest-first 4-wide collapse
 * (bvh.hpp:126-247). boxes: n boxes of 6 floats (min xyz, max xyz: the aabb
 * layout, aabb.hpp:9). Returns the 4-wide nodes (32 words each: lo[3][4],
 * hi[3][4], first[4], count[4]; bvh.hpp:27-32) and the primitive order, both
 * allocated by the library (free with pt_free); n = 0 gives no nodes.
 * flags: PT_BVH_DEVICE builds on the GPU (level-parallel binning and
 * partition, one wave per small subtree, parallel collapse) with the same
 * result word for word; 0 builds on host threads. ms: wall-clock of the
 * whole call, the device's transfers and allocations and the copies into the
 * returned arrays included (may be NULL). At most 2^30 boxes (2^28 for the
 * device build, PT_ERR_LIMIT); the device build refuses boxes with a NaN
 * coordinate (PT_ERR_INVALID) and hands boxes with a -0.0 coordinate to the
 * host builder (the reference keeps whichever of -0.0 and +0.0 comes first;
 * the device's reductions cannot), so the words are the same either way. */
enum { PT_BVH_DEVICE = 1 };
pt_status pt_bvh_build(const float* boxes, size_t n, int flags, uint32_t** nodes, size_t* n_nodes, uint32_t** order,
                       double* ms);

/* Depth (levels of 4-wide nodes) of the deepest tree in a PTGP buffer. */
pt_status pt_ptgp_depth(const uint32_t* ptgp, size_t n_words, int* depth);

/* ---------------------------------------------------------------------------
 * Device scene.
 * ------------------------------------------------------------------------- */

/* Copies a PTGP v1 buffer to the device. Replaces the page's GpuRenderer
 * loadScene (web/gpu.js:59-86), which uploads serialize(pack_for_gpu(...)).
 * Accepts trees deeper than the reference's WebGPU stack allows. */
pt_status pt_scene_create(const uint32_t* ptgp, size_t n_words, pt_scene** scene);
pt_status pt_scene_destroy(pt_scene* scene);

/* ---------------------------------------------------------------------------
 * Ray queries: the closest hit of geometry::hit (geometry.hpp:111-152) and
 * the any hit of geometry::occluded (geometry.hpp:156-190), over a batch.
 * ------------------------------------------------------------------------- */

/* A ray: origin, t_min, direction, t_max (32 bytes). */
typedef struct {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} pt_ray;

/* kind values of pt_hit */
enum { PT_MISS = 0, PT_LARGE_SPHERE = 1, PT_SPHERE = 2, PT_TRIANGLE = 3 };

/* A hit: t, barycentrics, what was hit. `prim` is the PTGP sphere number
 * (kinds 1-2) or PTGP triangle number (kind 3); `inst` is the instance number
 * for a triangle reached through the instance tree, else 0xffffffff. For any
 * hit queries only `kind` (miss or not) is meaningful. */
typedef struct {
    float t, u, v;
    uint32_t kind;
    uint32_t prim;
    uint32_t inst;
} pt_hit;

/* Host arrays in and out. any_hit = 0: closest hit; 1: any hit. */
pt_status pt_trace(pt_scene* scene, const pt_ray* rays, size_t n, int any_hit, pt_hit* hits);

/* Device arrays in and out (rays and hits already in HBM), on the scene's
 * stream; `elapsed_ms` (optional) gets the traversal kernel's time from HIP
 * events on that stream. `visits` (optional, 6 entries) runs the counting
 * variant of the kernel and gets the batch's totals of node slab tests, leaf
 * references, triangle tests, sphere tests, instance entries and large-sphere
 * tests, which price its algorithmic bytes. */
pt_status pt_trace_device(pt_scene* scene, const void* d_rays, size_t n, int any_hit, void* d_hits,
                          float* elapsed_ms, uint64_t* visits);

/* pt_trace_device without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * The traversal's work counter and stack spill area live in a 64 MiB array
 * the library holds on each device it runs on until the process ends; the
 * caller passes the rays and hits (device addresses, as pt_trace_device).
 * stream: a hipStream_t to enqueue on, returning without waiting (the call may
 * be captured into a graph), or NULL for the scene's stream, waited for before
 * returning. The scene's device must be current. Every call on one device
 * shares that device's array: calls that overlap on one device race (enqueue
 * them on one stream). Scenes whose trees need more stack than the array
 * holds for one traversal block per compute unit are refused (PT_ERR_LIMIT);
 * deep trees run on a smaller grid than pt_trace_device's. Hits are bit for
 * bit those of pt_trace_device. */
pt_status pt_trace_device_noalloc(pt_scene* scene, const void* d_rays, size_t n, int any_hit, void* d_hits,
                                  void* stream);

/* ---------------------------------------------------------------------------
 * Rendering: render_row_linear (renderer.hpp:411-416) for a whole frame, or a
 * band of rows, as a wavefront path loop on the GPU.
 * ------------------------------------------------------------------------- */

typedef struct {
    int32_t width, height;    /* image size (the camera is the PTGP one) */
    int32_t spp;              /* samples per pixel */
    int32_t max_bounces;      /* 10 in the reference CLI (main.cpp:84) */
    uint64_t seed;
    int32_t first_sample;     /* progressive continuation (renderer.hpp:345-356) */
    int32_t row_begin;        /* rows [row_begin, row_end) counted from the   */
    int32_t row_end;          /* bottom (j), as render_row_linear; 0,0 = all  */
    int32_t flags;            /* PT_RENDER_* */
    int32_t samples_per_batch;/* paths in flight per pixel per batch; 0 = auto */
    int32_t stack_entries;    /* LDS short-stack entries per lane; 0 = auto   */
} pt_render_params;

enum {
    PT_RENDER_NO_NEE = 1,        /* world.sample_lights = false (main.cpp --no-nee) */
    PT_RENDER_NO_GRAPH = 2,      /* launch the bounce loop eagerly, not as a hipGraph */
    PT_RENDER_TIMING = 4,        /* eager, with HIP events around every traversal launch */
    PT_RENDER_COUNT = 8,         /* eager, the traversal stages count their visits (stats->visits) */
    PT_RENDER_DEVICE_OUTPUT = 16,/* rgb and aux are device addresses (pt_device_alloc): no copy to the host */
    PT_RENDER_NO_TAIL = 32,      /* every bounce stage by stage: no tail launch for the last, small queues */
    PT_RENDER_GUIDE = 64,        /* diffuse bounces sample the attached guide (pt_scene_set_guide) 50/50
                                    with the cosine lobe, one-sample MIS */
    PT_RENDER_TRAIN = 128,       /* also record each path's diffuse vertices and, after every batch, take
                                    Adam steps on the guide over them (implies PT_RENDER_GUIDE) */
    PT_RENDER_PER_RAY = 256      /* camera rays walk the tree one per lane, in each ray's own nearest-first
                                    order (renderer.hpp's per-pixel loop). Without it, camera rays of a
                                    scene without instances walk it 64 at a time as packets: the same
                                    nearest hit except where two candidates tie at one t (an edge shared
                                    by two leaves), where the packet's order picks one; a packet is 64
                                    consecutive paths of a batch, so ties can resolve differently when
                                    samples_per_batch or the row band changes */
};

/* Render statistics. */
typedef struct {
    uint64_t rays;             /* counted as the reference: renderer.hpp:19,72,110 */
    uint64_t primary_rays;
    uint64_t shadow_rays;
    double render_ms;          /* device time of the render, HIP events */
    double extend_ms;          /* with PT_RENDER_TIMING: time in the traversal stages
                                  (closest-hit rays and the previous bounce's shadow rays) */
    double shadow_ms;          /* reserved, 0 (shadow rays are traced in the traversal stage) */
    uint64_t visits[7];        /* with PT_RENDER_COUNT, over the traversal-stage launches: node,
                                  leaf-ref, triangle, sphere, instance and large-sphere visits,
                                  and the number of queries those launches traced */
    int32_t batches;           /* batches the samples were split into */
    int32_t samples_per_batch; /* samples per pixel in each batch */
    double guide_ms;           /* guided renders with PT_RENDER_TIMING: time in the guide's inference
                                  launches (the network at every guided bounce) */
    uint64_t guide_queries;    /* guided renders with PT_RENDER_COUNT: diffuse hits the guide's
                                  inference evaluated the network for */
} pt_render_stats;

/* rgb: 3 floats a pixel, top row first, for the rows rendered (the band's top
 * row first). aux: NULL, or 7 floats a pixel (albedo, normal, luminance^2;
 * renderer.hpp:358-362). stats: optional. */
pt_status pt_render(pt_scene* scene, const pt_render_params* params, float* rgb, float* aux,
                    pt_render_stats* stats);

/* ---------------------------------------------------------------------------
 * Progressive rendering: the browser demo's driver (gpu.js:199-231) over
 * pt_render's first_sample continuation (renderer.hpp:345-356). A render of
 * params->spp samples is taken in steps that start at one sample per pixel
 * and double while a step's device time stays under 20 ms (up to 64), halving
 * above 60 ms; the running sum stays in HBM, so each step costs its own
 * samples only. Each sample is the one a single pt_render of params->spp
 * samples would draw (same seed, pixel and sample number).
 * ------------------------------------------------------------------------- */
typedef struct pt_progressive pt_progressive;
/* params as for pt_render (spp: the total); the scene must outlive the handle. */
pt_status pt_progressive_create(pt_scene* scene, const pt_render_params* params, pt_progressive** out);
/* Renders the next step. done: samples per pixel rendered so far (== spp when
 * the render is complete: further steps are refused); step_spp: this step's.
 * rgb: NULL, or the running average, 3 floats a pixel (a device address with
 * PT_RENDER_DEVICE_OUTPUT in params->flags). stats: this step's, optional. */
pt_status pt_progressive_step(pt_progressive* p, float* rgb, int* done, int* step_spp, pt_render_stats* stats);
pt_status pt_progressive_destroy(pt_progressive* p);

/* ---------------------------------------------------------------------------
 * Denoising: the CLI's --denoise step (main.cpp:172-186) and the reference's
 * filter (denoise.hpp:55-139): an edge-avoiding a-trous wavelet filter over
 * the lighting (color divided by the first-hit albedo), steered by each
 * pixel's noise, five passes of a 5x5 kernel with taps 1 to 16 pixels apart.
 * rgb: 3 floats a pixel as pt_render writes them; aux: pt_render's 7 aux
 * floats a pixel (albedo, normal, mean squared luminance); spp: the samples
 * behind them (each pixel's variance of the mean is the sample variance over
 * spp). out: 3 floats a pixel, not aliasing rgb. With PT_DENOISE_DEVICE all
 * three are device addresses; otherwise host memory, copied over and back.
 * device_ms (optional): the kernels' time. Working buffers (72 bytes a pixel,
 * and 52 more for the copies of host images) are allocated and freed by each
 * call.
 * ------------------------------------------------------------------------- */
enum { PT_DENOISE_DEVICE = 1 };
pt_status pt_denoise(int width, int height, int spp, const float* rgb, const float* aux, float* out, int flags,
                     double* device_ms);

/* Device memory for callers without the HIP headers (ray and hit buffers for
 * pt_trace_device, images for PT_RENDER_DEVICE_OUTPUT). No counterpart in the
 * reference, which has no device. */
pt_status pt_device_alloc(size_t bytes, void** ptr);
pt_status pt_device_free(void* ptr);
/* kind: 0 device to host, 1 host to device, 2 device to device. */
pt_status pt_device_copy(void* dst, const void* src, size_t bytes, int kind);

/* ---------------------------------------------------------------------------
 * Batched rendering (the config-3 workload): many jobs -- scenes, views, image
 * sizes and spp budgets of uneven size -- in one call. The jobs' scenes are
 * merged into one device scene. By default the jobs' paths share the ray
 * queues: passes of the bounce loop each carry slices of many jobs (a job's
 * pixels for a run of its samples), every kernel reading a path's camera,
 * tree root, large spheres and lights from its slice, so the GPU stays full
 * until the last pass; the whole sequence is one captured graph. With
 * PT_BATCH_STREAMS each job runs pt_render's bounce loop over its view of
 * the merged scene instead, the jobs spread over a few HIP streams that run
 * side by side, each stream's job sequence one captured graph. The reference
 * renders one image per call (render_row_linear, renderer.hpp:411-416); a
 * caller with many jobs calls it once per job. Every job's image is bit for
 * bit what pt_render gives for the same scene, size, spp, seed and
 * first_sample.
 * ------------------------------------------------------------------------- */
typedef struct pt_batch pt_batch;

/* Where one scene of a batch sits in the merged scene (pt_batch_merge). */
typedef struct {
    uint32_t root;              /* flat tree root node, or 0xffffffff (no tree) */
    uint32_t n_large, large_off;/* its large spheres: merged large-sphere list [large_off, +n_large) */
    uint32_t n_lights, lights_off;
    uint32_t flags;             /* its PTGP header flags (1 sky, 2 sample lights) */
    float camera[16];           /* its PTGP camera rows: origin, corner, across, up (w = 0) */
} pt_batch_part;

/* Host only: several flat PTGP v1 scenes (no instances) as one PTGP v1
 * buffer (free with pt_free): arrays concatenated with rebased indices, and
 * 4-wide nodes over the scenes' roots as its flat root. parts: n entries. */
pt_status pt_batch_merge(const uint32_t* const* ptgps, const size_t* n_words, int n, uint32_t** merged,
                         size_t* merged_words, pt_batch_part* parts);

/* Uploads the merged scene of n flat PTGP v1 scenes (each checked as
 * pt_scene_create checks one). */
pt_status pt_batch_create(const uint32_t* const* ptgps, const size_t* n_words, int n, pt_batch** batch);
pt_status pt_batch_destroy(pt_batch* batch);

enum { PT_JOB_CAMERA = 0x100 };   /* pt_batch_job.flags: use the job's camera rows, not its scene's */

typedef struct {
    int32_t scene;              /* which of the batch's scenes */
    int32_t width, height;      /* image size, 2 .. 32768 */
    int32_t spp;                /* samples per pixel */
    uint64_t seed;
    int32_t first_sample;       /* progressive continuation, as pt_render_params */
    int32_t flags;              /* PT_RENDER_NO_NEE, PT_JOB_CAMERA */
    float camera[16];           /* with PT_JOB_CAMERA: origin, corner, across, up (4 floats each, w unused) */
} pt_batch_job;

/* pt_batch_params.flags: each job's own bounce loops, jobs spread over HIP
 * streams, instead of shared-queue passes (the default). */
enum { PT_BATCH_STREAMS = 0x200 };

typedef struct {
    int32_t max_bounces;        /* every job's bounce cap, 1 .. 63 */
    int32_t flags;              /* PT_RENDER_NO_GRAPH, PT_RENDER_DEVICE_OUTPUT, PT_RENDER_NO_TAIL, PT_BATCH_STREAMS,
                                   PT_RENDER_PER_RAY (camera rays walked per lane, as pt_render with the flag;
                                   without it they are walked as 64-ray packets per job) */
    int32_t stack_entries;      /* 8, 16 or 32; 0 = tuned (pt_batch_autotune) or 8 */
    int32_t blocks_per_cu;      /* traversal blocks per CU; 0 = tuned or the occupancy limit */
    int32_t streams;            /* HIP streams the jobs are spread over, 1 .. 8; 0 = tuned or 2 */
    int32_t reserved;
    uint64_t max_paths;         /* paths one pass (or, with PT_BATCH_STREAMS, one batch of a job) may hold: the
                                   queue size, at most 2^28 (2^27); 0 = tuned or memory-sized */
} pt_batch_params;

/* Renders every job. rgb: the jobs' images one after another in job order,
 * each 3 floats a pixel, top row first (a host array, or with
 * PT_RENDER_DEVICE_OUTPUT a device one). Shared queues: passes of at most
 * max_paths paths, filled with the jobs in order, each job's samples in
 * order and split where a pass fills. PT_BATCH_STREAMS: a job's samples go
 * in batches of spb samples a pixel, spb the largest divisor of its spp whose
 * batch fits max_paths. Either way a pixel's samples are summed in sample
 * order, so the split changes no bit of the result. stats: rays as pt_render
 * counts them, summed over the jobs; batches: passes (shared queues) or
 * batches summed over the jobs; render_ms: device time of the whole call. */
pt_status pt_render_batch(pt_batch* batch, const pt_batch_job* jobs, int n_jobs, const pt_batch_params* params,
                          float* rgb, pt_render_stats* stats);

/* The autotuner: renders the jobs in shared-queue passes (as given: pass
 * fewer samples for a quicker tune) under candidate launch shapes -- LDS
 * stack entries (8, 16, 32) x traversal blocks per CU, then pass sizes
 * (2^25, 2^26, 2^27 paths, and the memory-sized default) -- and keeps the
 * fastest in the batch for later renders whose params leave them 0 (streams
 * stay untuned: 2 with PT_BATCH_STREAMS).
 * result (optional): the chosen shape and the measured rate of each
 * candidate. */
typedef struct {
    int32_t stack_entries, blocks_per_cu, streams, n_tried;
    uint64_t max_paths;
    struct {
        int32_t stack_entries, blocks_per_cu, streams, reserved;
        uint64_t max_paths;
        double mrays_per_s;
    } tried[32];
} pt_tune_result;
pt_status pt_batch_autotune(pt_batch* batch, const pt_batch_job* jobs, int n_jobs, int max_bounces,
                            pt_tune_result* result);

/* The autotuner for one scene's renders: renders the frame `params` describe
 * (as given; pass fewer samples for a quicker tune) under candidate launch
 * shapes -- LDS stack entries (8, 16, 32) x traversal blocks per CU, then
 * batches of a half and a quarter of the memory-sized one -- timing each by
 * its device time, and keeps the fastest in the scene for its later renders
 * and traces whose params leave stack_entries and samples_per_batch 0. The
 * shape changes no bit of an image. result (optional): the chosen shape
 * (max_paths: paths a batch may hold, 0 = memory-sized; streams: 1) and the
 * rate of each candidate. The reference has no tuner: its CLI splits rows
 * over hardware threads (main.cpp:116-158). */
pt_status pt_scene_autotune(pt_scene* scene, const pt_render_params* params, pt_tune_result* result);

/* ---------------------------------------------------------------------------
 * Learned path guide (the config-4 guided tracer). The reference has no path
 * guiding: these entry points have no counterpart there. The guide is an MLP
 * (one-blob position 3x16 + normal angles 2x8 = 64 inputs, four hidden layers
 * of 64 with ReLU, 32 outputs = 8 von Mises-Fisher lobes x (polar logit,
 * azimuth logit, log kappa, weight logit)), FP16 weights with FP32
 * accumulation on the GPU's matrix cores, trained by Adam on the tracer's own
 * radiance samples. A diffuse bounce then samples it 50/50 with the cosine
 * lobe and weighs by the mixed pdf (one-sample MIS), so images stay unbiased.
 * ------------------------------------------------------------------------- */
typedef struct pt_guide pt_guide;

#define PT_GUIDE_PARAMS 18720   /* FP32 parameters: 4 x (64x64 + 64) + 32x64 + 32 */
#define PT_GUIDE_OUTPUTS 32

typedef struct {
    uint64_t seed;              /* weight initialisation */
    float learning_rate;        /* Adam step size; 0 means 1e-3 */
    float box_lo[3], box_hi[3]; /* position box of the input encoding; all zero: the scene's tree bounds */
    int32_t steps_per_batch;    /* PT_RENDER_TRAIN: Adam steps after each batch; 0 means one pass over its records */
    int32_t minibatch;          /* samples per Adam step; 0 means 16384 */
    float weight_clamp;         /* training weights are clamped at this many times their batch mean; 0 means 16 */
    int32_t bounces;            /* diffuse vertices of bounces [0, bounces) sample the guide (later ones the
                                   cosine lobe alone, and the tail launch may take them); 0 means every bounce */
} pt_guide_params;

pt_status pt_guide_create(pt_scene* scene, const pt_guide_params* params, pt_guide** guide);
pt_status pt_guide_destroy(pt_guide* guide);
/* Attaches a guide to a scene for PT_RENDER_GUIDE / PT_RENDER_TRAIN renders (NULL detaches). */
pt_status pt_scene_set_guide(pt_scene* scene, pt_guide* guide);
/* The FP32 parameters, layer by layer: weights [out][in] row-major, then biases. */
pt_status pt_guide_get_weights(pt_guide* guide, float* params, size_t n);
/* Sets the parameters and clears Adam's moments and step count. */
pt_status pt_guide_set_weights(pt_guide* guide, const float* params, size_t n);
/* Network outputs at n shading points (pos, nrm: 3 floats each): 32 floats each. */
pt_status pt_guide_eval(pt_guide* guide, const float* pos, const float* nrm, size_t n, float* out);
/* Mixture pdf (per steradian) of direction dir[i] (3 floats) at point i. */
pt_status pt_guide_pdf(pt_guide* guide, const float* pos, const float* nrm, const float* dir, size_t n, float* pdf);
/* One mixture sample per point from u[i] = 3 uniforms: direction and pdf (4 floats each). */
pt_status pt_guide_sample(pt_guide* guide, const float* pos, const float* nrm, const float* u, size_t n,
                          float* dir_pdf);
/* Training samples are 12 floats: position, weight, normal, 0, direction, 0; a sample's loss is
 * weight * -log pdf(direction). pt_guide_gradient gives the mean gradient over n samples (and the
 * mean loss); pt_guide_step takes one Adam step with it. */
pt_status pt_guide_gradient(pt_guide* guide, const float* samples, size_t n, float* grad, float* loss);
pt_status pt_guide_step(pt_guide* guide, const float* samples, size_t n, float* loss);
/* Training statistics of PT_RENDER_TRAIN renders since the last call: Adam steps taken, samples
 * used, mean loss of the last step. */
pt_status pt_guide_stats(pt_guide* guide, uint64_t* steps, uint64_t* samples, float* last_loss);

#ifdef __cplusplus
}
#endif

#endif /* PHOTON_TRACER_H */
