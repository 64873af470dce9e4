/*
 * lts.h — C ABI of the MI355X path tracer (liblts.so).
 *
 * Plain C types and host pointers only. Every entry point returns 0 on success and a negative
 * LTS_E_* code on failure (the message is available from lts_last_error); none of them exits
 * the process. Calls are synchronous on return. One context per host thread; a context owns all
 * of its device memory and one HIP stream on the device it was created for.
 *
 * Each declaration names the reference routine it replaces (paths inside
 * tierChampion/LightTransportSimulation/src).
 */
#ifndef LTS_H
#define LTS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LTS_OK 0
#define LTS_E_ARG -1      /* invalid argument */
#define LTS_E_IO -2       /* file missing or malformed */
#define LTS_E_HIP -3      /* HIP runtime error */
#define LTS_E_STATE -4    /* call out of order (e.g. render before load) */
#define LTS_E_NOMEM -5    /* device allocation failed */

typedef struct lts_ctx lts_ctx;

/* Canonical LBVH arrays (caller-allocated, n primitives). Children use the unified index space
 * of the reference's flattened tree: interior i in [0, n-2], leaf k at n-1+k
 * (acceleration/BVH.cuh:311-329). Bounds are {min x, y, z, max x, y, z}. */
typedef struct lts_bvh_export {
    uint32_t* codes_sorted; /* n     Morton codes after the stable sort           */
    int32_t* perm;          /* n     sorted position -> primitive number          */
    int32_t* left;          /* n-1   left child of interior i                     */
    int32_t* right;         /* n-1   right child of interior i                    */
    float* int_bounds;      /* 6(n-1)                                             */
    int32_t* leaf_prim;     /* n     primitive referenced by leaf k               */
    float* leaf_bounds;     /* 6n                                                 */
    int32_t* parent;        /* 2n-1  parent in the unified index space, -1 at root */
    float scene_bounds[6];  /* global bounds used for the Morton quantisation     */
} lts_bvh_export;

typedef struct lts_bvh_timing {
    double total_ms;   /* bounds -> pack, device time of the whole build           */
    double sort_ms;    /* the four radix-sort passes                               */
    double bounds_ms;  /* primitive bounds + scene-bound reduction + Morton codes  */
    double tree_ms;    /* Karras hierarchy + bottom-up refit + node packing        */
} lts_bvh_timing;

/* Integrators. */
#define LTS_INTEGRATOR_MEGAKERNEL 0
#define LTS_INTEGRATOR_WAVEFRONT 1
#define LTS_INTEGRATOR_CACHED 2
/* Bidirectional path tracing (replaces BidirectionalPathIntegrator,
 * intergrators/BidirectionalPathIntegrator.cuh:1-665, which the reference disables at
 * LightTransportSimulation.cu:116-121): every strategy with a camera subpath of at least two
 * vertices, balance-heuristic MIS, max_bounce at most 8, no Russian roulette (rr_start unused).
 * Same film and sample layout as the megakernel. */
#define LTS_INTEGRATOR_BDPT 3

typedef struct lts_render_params {
    int32_t spp;          /* samples per pixel, 1 .. 2^29                          */
    int32_t max_bounce;   /* PathTracingIntegrator MAX_BOUNCE                      */
    int32_t rr_start;     /* Russian roulette from this bounce                     */
    uint32_t seed;        /* sampler seed (replaces time(0))                       */
    int32_t stratified;   /* 1: film sample per Sampler::stratifiedSample          */
    int32_t integrator;   /* LTS_INTEGRATOR_*                                      */
    int32_t spp_per_pass; /* samples per pass; 0 = automatic (wavefront and cached: */
                          /* 2^27 paths per pass; megakernel: 2^28 / 12 B)          */
    /* wavefront only (0 / -1 = default or the autotuned value) */
    int32_t trace_block;        /* block size of the persistent traversal kernels: 64, 128, 256 */
    int32_t persistent_blocks;  /* grid of the persistent traversal kernels                     */
    int32_t sort_materials;     /* -1 default (on), 0 off, 1 counting-sort shading by material  */
    int32_t use_graph;          /* -1 default (on), 0 off: capture each pass in a hipGraph       */
    int32_t autotune;           /* 1: time candidate tunings on this scene before rendering:    */
                                /*    trace block, persistent grid, material sort, tile rows and */
                                /*    (with spp_per_pass 0) the pass size, whole or half; the    */
                                /*    choice is kept per scene, film, pass size and spp in the   */
                                /*    context, and in the file named by $LTS_TUNE_CACHE if set  */
    int32_t profile_phases;     /* 1: time the wavefront phases with HIP events (stats.phase_ms); */
                                /*    passes are then launched directly, not replayed as graphs  */
    int32_t batch_group_pixels; /* lts_render_batch: scenes are grouped, largest first, until a   */
                                /*    group holds this many pixels; a group shares every launch   */
                                /*    of its passes. 0: 2^20; -1: all scenes in one group         */
    int32_t count_traversal;    /* 1: count render-node fetches and triangle tests (a separate,    */
                                /*    untimed instantiation of the traversal): stats.node_visits / */
                                /*    tri_tests for the path rays; the wavefront integrators also   */
                                /*    count shadow and MIS rays (stats.ray_nodes / ray_tris)       */
    int32_t sample_begin;       /* progressive rendering: samples [sample_begin, sample_end) of  */
    int32_t sample_end;         /*    the spp-sample image (sample_end 0: spp). sample_begin > 0  */
                                /*    keeps the film: the previous render of this context (or    */
                                /*    lts_film_load) must have ended at sample_begin with the same */
                                /*    scene, spp, seed, stratification, bounces, roulette and      */
                                /*    integrator. The image then equals one render of             */
                                /*    [0, sample_end) bit for bit. lts_render_batch: full range.  */
    int32_t tile_rows;          /* wavefront: path order within a sample, bands of this many rows  */
                                /*    walked column by column (a wave of 64 paths covers a        */
                                /*    tile_rows x 64/tile_rows tile); 0: 8 or the autotuned value; */
                                /*    1 .. 64. The image does not depend on it.                   */
} lts_render_params;

/* Film checkpoint (no reference counterpart: the reference renders all samples in one launch).
 * lts_film_save writes the context's film (per-pixel rgb sums and filter weights), the next sample
 * index and the sampler parameters of the renders that filled it; lts_film_load restores them into
 * a context whose loaded scene has the same film size, so a render with sample_begin = next sample
 * continues it. lts_film_resolve applies getRender (rendering/Camera.cu:256-289) to the current
 * film: rgb_out receives width * height * 3 floats, top row first. */
int lts_film_save(lts_ctx* ctx, const char* path);
int lts_film_load(lts_ctx* ctx, const char* path);
int lts_film_resolve(lts_ctx* ctx, float* rgb_out);

/* Wavefront phases timed when profile_phases is set (indices into lts_stats.phase_ms). */
#define LTS_PHASE_EXTEND 0      /* all traversal: path rays with the shadow and MIS rays traced */
                                /* in the same launches (one per bounce)                      */
#define LTS_PHASE_SORT 1        /* material sort of the shading queue                        */
#define LTS_PHASE_SHADE 2
#define LTS_PHASE_SHADOW 3      /* 0: shadow rays are traced with the path rays (EXTEND)      */
#define LTS_PHASE_MIS 4         /* 0: MIS-BSDF rays are traced with the path rays (EXTEND)    */
#define LTS_PHASE_FINALIZE 5
#define LTS_PHASE_CACHE_INFER 6 /* radiance-cache inference (cached integrator)              */
#define LTS_PHASE_CACHE_TRAIN 7 /* training-batch compaction and Adam steps                   */

typedef struct lts_stats {
    double render_ms;       /* device time of the integrator kernels (HIP events)  */
    double resolve_ms;      /* device time of the film gather                       */
    double kernel_avg_ms;   /* mean duration of one integrator launch               */
    int64_t launches;       /* integrator launches                                  */
    int64_t samples;        /* camera samples traced                                */
    double msamples_per_s;  /* samples / (render_ms + resolve_ms)                   */
    int64_t node_visits;    /* count_traversal: render-node fetches of the path rays */
    int64_t tri_tests;
    int32_t trace_block;    /* wavefront tuning that was used                      */
    int32_t persistent_blocks;
    int32_t sort_materials;
    int32_t use_graph;
    double autotune_ms;     /* time spent in the autotuner by this call             */
    double nrc_loss;        /* cached: loss of the last training step               */
    int64_t nrc_queries;    /* cached: cache queries of the last pass               */
    int32_t nrc_train_samples; /* cached: training samples collected in the last pass */
    int32_t nrc_steps;      /* cached: Adam steps taken by the cache so far          */
    double phase_ms[8];     /* device time per LTS_PHASE_* summed over the passes     */
    int64_t phase_launches[8]; /* kernel launches per phase                          */
    int32_t batch_groups;   /* lts_render_batch: groups the scenes were rendered in    */
    int32_t pad;
    int64_t ray_nodes[3];   /* count_traversal: 4-wide render-node fetches of the path, shadow and  */
    int64_t ray_tris[3];    /*    MIS rays (megakernel: path rays only), and their triangle tests  */
    int32_t tile_rows;      /* wavefront: band height of the path order that was used            */
    int32_t spp_per_pass;   /* wavefront: samples per pass that were used                         */
} lts_stats;

/* Context. */
int lts_create(lts_ctx** out, int hip_device);
void lts_destroy(lts_ctx* ctx);
/* With ctx NULL: the message of the calling thread's last failed call that takes no context. */
const char* lts_last_error(const lts_ctx* ctx);
const char* lts_version(void);

/* Replaces readApplicationParameters (core/FileIO.cuh:234-257): the positional app.params. */
int lts_read_app_params(const char* path, int32_t* width, int32_t* height, int32_t* spp, int32_t* max_bounce,
                        int32_t* rr_start, int32_t* format, char* output, char* scene, char* subject, int32_t cap);

/* Replaces parseScene (core/SceneLoader.cuh:195-360): reads <dir><scene>.scene and
 * <dir><subject>.subject (subject may be "" or NULL), meshes, materials, images and lights,
 * builds the LBVH on the device (sorted-leaf layout) and keeps everything resident. */
int lts_load_scene(lts_ctx* ctx, const char* project_dir, const char* scene, const char* subject,
                   int32_t width, int32_t height);
int lts_scene_info(lts_ctx* ctx, int32_t* n_tris, int32_t* n_lights, int32_t* n_meshes);

/* Render tree of the scenes this context loads from now on. Rendering does not require the
 * reference's tree: any binary tree over the Morton-sorted leaves finds the same closest hits.
 * LTS_TREE_PLOC (default) clusters the sorted leaves by PLOC (surface-area nearest neighbours,
 * radius 8) and falls back to the Karras tree when the PLOC tree is too deep for the traversal
 * stack; LTS_TREE_KARRAS renders with the LBVH itself (BVH::intersect's tree, BVH.cuh:365-398).
 * The binary tree is collapsed into 4-wide nodes greedily (each node opens its largest-area
 * interior child until it has four children) or, with LTS_TREE_COLLAPSE_EVEN or when the greedy
 * tree has too many levels for the traversal stack, by taking every interior node of even depth.
 * The exported LBVH (lts_bvh_build*) is always the bit-exact Karras tree. No reference
 * counterpart; the reference renders with its LBVH. */
#define LTS_TREE_KARRAS 0
#define LTS_TREE_PLOC 1
#define LTS_TREE_COLLAPSE_EVEN 0x100 /* or-ed into the kind */
int lts_set_render_tree(lts_ctx* ctx, int32_t kind);
/* Texture pipeline of the scenes this context loads from now on. LTS_TEXTURE_REFERENCE (default)
 * reproduces the reference's image textures: its blocked texel layout with the aliasing row index
 * (core/MemoryHelper.cuh:88-96), pyramid levels filled only near the origin
 * (Mipmap_kernel.cu:29-30) and the fixed lookup width 0.01 (ImageTexture.cuh:8). LTS_TEXTURE_FILTERED
 * is the pipeline done right (no reference counterpart): a full box-filtered pyramid down to 1x1,
 * and trilinear lookups whose width comes from the ray differentials of the primary hit
 * (Interaction.cuh:118-160, PBRT's isotropic width); secondary hits, which have no differentials,
 * read the finest level. */
#define LTS_TEXTURE_REFERENCE 0
#define LTS_TEXTURE_FILTERED 1
int lts_set_texture_filter(lts_ctx* ctx, int32_t mode);

/* Tree the loaded scene renders with (LTS_TREE_COLLAPSE_EVEN set when its 4-wide nodes come from
 * the even-depth collapse), its depth in edges and its 4-wide node count. */
int lts_render_tree_info(lts_ctx* ctx, int32_t* kind, int32_t* depth, int32_t* nodes4);

/* Replaces Scene::intersect / BVH::intersect (acceleration/BVH.cuh:365-398) as a query: the closest
 * hit of each of n rays (rays[6 i .. 6 i + 5] = origin, direction, world space, tMax infinite) in the
 * loaded scene, by the renderer's own traversal of its render tree. prim[i] receives the hit
 * triangle's primitive number (the order of the scene's meshes, as lts_scene_info counts them) or
 * -1, t[i] the distance in units of the direction's length (infinity for a miss). */
int lts_trace_rays(lts_ctx* ctx, const float* rays, int32_t n, int32_t* prim, float* t);

/* Replaces CreateBVHTree (acceleration/BVH.cuh:411-452) for a triangle soup of n triangles
 * (9 floats each, host memory). mode 0 reproduces the reference's leaf assignment (leaf k holds
 * primitive k, BVH.cuh:242); mode 1 assigns leaf k to the k-th primitive in Morton order.
 * Results stay on the device; `out` (may be NULL) receives the canonical arrays. */
int lts_bvh_build(lts_ctx* ctx, const float* tri_xyz, int32_t n, int32_t mode, lts_bvh_export* out,
                  lts_bvh_timing* timing);
/* Same build on a soup already resident in device memory (d_tri_xyz is a device pointer, or NULL
 * for the soup uploaded by the last lts_bvh_build with the same n). One build runs with events
 * between its stages (timing's stage figures); with repeat > 1 the build is then captured as a
 * HIP graph and replayed `repeat` times back to back, and timing.total_ms is the mean replay. */
int lts_bvh_build_device(lts_ctx* ctx, const float* d_tri_xyz, int32_t n, int32_t mode, int32_t repeat,
                         lts_bvh_timing* timing);
/* Copies the arrays of the most recent build to host memory. */
int lts_bvh_export_last(lts_ctx* ctx, lts_bvh_export* out);

/* Replaces the world-transform loops of the TriangleMesh constructor (shapes/Triangle.cu:26-41):
 * transforms n_points points and n_normals normals (3 floats each, host memory) in place on the
 * context's GPU. `subject` holds the scene file's subject transform (core/SceneLoader.cuh:33-193):
 * translation (3), rotation axis (3), angle in radians (taken modulo 2 pi), uniform scale. Points
 * map to s q p q* + t with q = (cos a/2, axis sin a/2); normals to normalize(q* n q), normalised
 * with the reference's fast inverse square root. lts_load_scene applies the same kernel to every
 * mesh (identity for the scene's own meshes). */
int lts_mesh_transform(lts_ctx* ctx, const float* subject, float* points, int32_t n_points, float* normals,
                       int32_t n_normals);

/* No counterpart in the reference: the LBVH build for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Builds the
 * tree lts_bvh_build_device builds (mode 0 or 1) for a device soup of n triangles, 1 <= n <=
 * LTS_BVH_STATIC_MAX_PRIMS, without allocating anything: the temporaries live in one static device
 * buffer of the library (LTS_BVH_STATIC_SCRATCH_BYTES per device, held until the process ends) and
 * the results go to the caller's device arrays: d_codes_sorted (n), d_perm (n), d_left (n-1),
 * d_right (n-1), d_parent (2n-1) and d_bounds (6(2n-1) floats {min xyz, max xyz} in the unified
 * index space: interior i at 6i, leaf k at 6(n-1+k); leaf k holds primitive k in mode 0 and
 * d_perm[k] in mode 1). scene_bounds (host, may be NULL) receives the global bounds of the Morton
 * quantisation. Runs on the current device and `hip_stream` (a hipStream_t; NULL for the default
 * stream) and returns when the build is done. Every call on one device shares its buffer: calls
 * that overlap on one device race. LTS_E_ARG beyond LTS_BVH_STATIC_MAX_PRIMS. */
#define LTS_BVH_STATIC_MAX_PRIMS (1 << 18)
#define LTS_BVH_STATIC_SCRATCH_BYTES (24u << 20)
int lts_bvh_build_static(const float* d_tri_xyz, int32_t n, int32_t mode, uint32_t* d_codes_sorted, int32_t* d_perm,
                         int32_t* d_left, int32_t* d_right, int32_t* d_parent, float* d_bounds, float* scene_bounds,
                         void* hip_stream);

/* Replaces PathTracingKernel + Camera::getRender (intergrators/IntegratorKernels.cu:17-23,
 * rendering/Camera.cu:256-289). rgb_out (may be NULL) receives W*H*3 floats in getRender row
 * order; sample_rgb (may be NULL) receives every sample's radiance, [spp][H][W][3]. */
int lts_render(lts_ctx* ctx, const lts_render_params* params, float* rgb_out, float* sample_rgb, lts_stats* stats);

/* Batched form of lts_render with the wavefront integrator (the reference renders one scene per
 * process, LightTransportSimulation.cu:23-146): the scenes loaded in ctxs[0..n-1] (contexts of one
 * device) are rendered on ctxs[0]'s stream and buffers in groups (params.batch_group_pixels); the
 * scenes of a group are rendered together, every queue kernel of a pass working on the paths of all
 * of them. Each scene's image equals what lts_render gives for it alone; rgb_out[i] (entries may be
 * NULL) receives scene i in getRender row order; stats cover the whole batch. */
int lts_render_batch(lts_ctx* const* ctxs, int32_t n, const lts_render_params* params, float* const* rgb_out,
                     lts_stats* stats);

/* ---- Learned radiance cache. No reference counterpart: the reference traces every path to the
 * end. A multilayer perceptron (64-dim encoding of a 16-float record -> 5 hidden layers of 64,
 * ReLU, no biases -> RGB) is trained online on the radiance of path suffixes and queried where a
 * cached render cuts paths short (LTS_INTEGRATOR_CACHED). Weights are fp32 masters; the MFMA
 * kernels use bf16 or fp16 copies with fp32 accumulation. Record layout (LTS_NRC_RAW floats):
 * position normalised to the scene bounds (3), outgoing direction theta/pi and phi/2pi + 1/2 (2),
 * normal theta/pi and phi/2pi + 1/2 (2), roughness (1), diffuse reflectance (3), specular
 * reflectance (3), 2 unused. Parameters: W_l[in][out] row-major for l = 0..4 (64 x 64), then the
 * 64 x 16 output matrix whose columns 3..15 are padding and stay zero. */
#define LTS_NRC_PARAMS 21504
#define LTS_NRC_RAW 16

typedef struct lts_nrc_config {
    int32_t precision;      /* 0 bf16 (default), 1 fp16                                             */
    int32_t batch;          /* training samples per Adam step (default 65536)                       */
    int32_t query_bounce;   /* cached renders query the cache at the first non-specular hit of      */
                            /* this bounce or later (default 2)                                     */
    int32_t train_paths;    /* paths per pass traced to the end as training paths (default 262144,  */
                            /* at most 1/8 of the pass's paths)                                     */
    int32_t warmup_passes;  /* the first passes of cached rendering after this call are traced       */
                            /* without queries while the cache trains (default 4, -1: none)        */
    uint32_t seed;          /* weight initialisation (Xavier-uniform from a SplitMix64 stream)      */
    float lr, beta1, beta2, eps;  /* Adam (defaults 1e-2, 0.9, 0.99, 1e-15)                          */
    int32_t train_steps;    /* Adam steps per cached pass over the batches collected in the pass,   */
                            /* cycling through them (default 16)                                    */
    int32_t reserved[5];
} lts_nrc_config;

/* (Re)initialises the cache of this context: weights from the seed, zero Adam state. NULL or zero
 * fields take the defaults. */
int lts_nrc_configure(lts_ctx* ctx, const lts_nrc_config* config);
/* Replaces the weights (LTS_NRC_PARAMS floats; padding entries are ignored) and resets Adam. */
int lts_nrc_set_weights(lts_ctx* ctx, const float* params);
int lts_nrc_get_weights(lts_ctx* ctx, float* params);
/* Network outputs for n records (host arrays: raw n x LTS_NRC_RAW, out n x 3). */
int lts_nrc_infer(lts_ctx* ctx, const float* raw, int32_t n, float* out);
/* One Adam step on n <= batch samples with the relative L2 loss
 * (1/n) sum_i sum_c (y_ic - t_ic)^2 / (y_ic^2 + 0.01), the denominator held constant.
 * loss (may be NULL) receives the loss before the step, grad (may be NULL) its gradient. */
int lts_nrc_train_step(lts_ctx* ctx, const float* raw, const float* target, int32_t n, float* loss, float* grad);

/* No reference counterpart: a plain device-to-device copy kernel over `bytes` bytes, timed with
 * HIP events (mean of `repeat` launches). Used to calibrate HBM traffic counters and to measure
 * the achievable copy bandwidth next to the kernels' roofline figures. */
int lts_device_copy_benchmark(lts_ctx* ctx, uint64_t bytes, int32_t repeat, double* ms_per_copy);

/* Replaces writeToPPM / writeToPFM (core/ImageO.cuh:16-78). format 0: P3 PPM, 1: PFM (the
 * reference's two formats); 2: Radiance RGBE .hdr, the format of the reference's unused outputFile
 * (core/ImageO.cuh:80-92); 3: 8-bit PNG with the PPM tone map int(pow(clamp(x), 1/2.2) * 255 + 0.5).
 * rgb rows are top first, as lts_render returns them. */
int lts_write_image(const char* path, const float* rgb, int32_t width, int32_t height, int32_t format);

/* Replaces loadImageFile_s (core/ImageI.cuh:11-37): decodes an 8-bit image file (PNG, binary PPM
 * or baseline JPEG, the last to the bytes of the reference's stb_image 2.27) to linear RGB with
 * pow(c / 255, 2.2), bottom row first, as the texture loader does. *width and *height receive the size; rgb_out (may be NULL to query the size)
 * receives width * height * 3 floats when cap_floats is large enough, else LTS_E_ARG. LTS_E_IO for
 * a missing, malformed or unsupported file. No GPU use. */
int lts_load_image(const char* path, int32_t* width, int32_t* height, float* rgb_out, int64_t cap_floats);

/* Replaces parseMeshFile + binMeshWrite (core/FileIO.cuh:68-158, 20-45): converts
 * <mesh_stem>.obj (v, vt, vn and triangular "f p/uv/n" statements) to <mesh_stem>.bin, the mesh
 * format lts_load_scene reads. lts_load_scene does the same by itself for a mesh that has only an
 * .obj, as the reference's parseMesh (core/FileIO.cuh:212-231) does. *n_tris (may be NULL)
 * receives the triangle count. LTS_E_IO for a missing, malformed or unwritable file. No GPU use. */
int lts_obj_to_bin(const char* mesh_stem, int32_t* n_tris);

#ifdef __cplusplus
}
#endif

#endif /* LTS_H */
