/*
 * ff_abi.h -- C ABI of firefly_amd, the MI355X (gfx950) wavefront path tracer.
 *
 * Plain C: opaque handles, explicit sizes, caller-visible device buffers,
 * HIP streams passed as void*.  Every call returns an ff_status and never
 * exits the process; ff_last_error() gives the message of the calling
 * thread's last failure.  One context per host thread; handles are not
 * thread-safe, except that ff_scene_create only reads its context and may
 * run concurrently on one context from several host threads.
 *
 * Each entry point names the reference interface it replaces
 * (JakeKurtz/Firefly, CUDA-RayTracer/ = CRT/).
 */
#ifndef FF_ABI_H
#define FF_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FF_ABI_VERSION 2

typedef enum {
    FF_OK = 0,
    FF_ERR_INVALID = 1,     /* bad handle, size, pointer or parameter      */
    FF_ERR_HIP = 2,         /* HIP runtime error (message has the detail)  */
    FF_ERR_ALLOC = 3,       /* device or host allocation failed            */
    FF_ERR_UNSUPPORTED = 4, /* request outside what this build supports    */
    FF_ERR_CAPACITY = 5     /* request exceeds a documented capacity limit */
} ff_status;

typedef struct ff_context ff_context;
typedef struct ff_scene ff_scene;
typedef struct ff_envmap ff_envmap;

/* 32-byte flattened BVH node, same fields as LinearBVHNode (CRT/BVH.h:63-72):
 * bounds, first primitive (leaf) or second child (interior), primitive count
 * (0 = interior), split axis. */
typedef struct {
    float bmin[3];
    float bmax[3];
    int32_t offset;
    uint16_t n_prims;
    uint8_t axis;
    uint8_t pad;
} ff_bvh_node;

/* Material record (16 floats), fields of dMaterial (CRT/dMaterial.cuh:9-31):
 * [0] type, [1..3] base colour, [4] roughness, [5] metallic, [6..8] fresnel f0,
 * [9] ior, [10] albedo texture (-1 none, 3 channels), [11] roughness texture
 * (-1 none, 1 channel), [12] metallic texture (-1 none, 1 channel; the
 * metallicRoughnessTexture blue / metallicTexture lookups of get_metallic,
 * CRT/dMaterial.cu:44-59), [13] normal texture (-1 none, 3 channels, tangent
 * space mapped by 2t-1 as get_normal, CRT/dMaterial.cu:61-87; needs the
 * scene's tangents), [14..15] reserved. */
#define FF_MAT_MIX 0        /* reference wf_mat_mix lobe pair (diffuse + GGX)  */
#define FF_MAT_DIFFUSE 1    /* reference diffuse lobe (wf_mat_diffuse)         */
#define FF_MAT_CONDUCTOR 2  /* reference GGX lobe (wf_mat_cook), metallic      */
#define FF_MAT_DIELECTRIC 3 /* smooth dielectric (no reference counterpart)    */
#define FF_MAT_FLOATS 16

/* Scene description in host memory; copied at ff_scene_create. */
typedef struct {
    int32_t n_tris;
    const float *positions;    /* [n][3][3] */
    const float *normals;      /* [n][3][3] */
    const float *uvs;          /* [n][3][2] (may be NULL) */
    const int32_t *material_ids; /* [n] */
    int32_t n_materials;
    const float *materials;    /* [n_materials][16] */
    int32_t n_textures;
    const float *const *textures; /* per texture [h][w][c] float */
    const int32_t *tex_w;
    const int32_t *tex_h;
    const int32_t *tex_c;      /* 1 or 3 */
    int32_t max_prims_in_node; /* SAH leaf size, 8 in the reference (CRT/dScene.cpp:372) */
    int32_t builder;           /* FF_BVH_SAH (0) or FF_BVH_LBVH */
    const float *tangents;     /* [n][3][6]: per vertex tangent xyz, bitangent xyz
                                  (Vertex::tangent/bitangent, CRT/dVertex.cuh:8-9);
                                  may be NULL when no material has a normal texture */
} ff_scene_desc;

/* BVH builders.  FF_BVH_SAH: the reference's host SAH tree (BVHAccel,
 * CRT/BVH.cpp:81-307), the default.  FF_BVH_LBVH: a linear BVH built on the
 * GPU (Morton order, Karras hierarchy; the reference declares HLBVH,
 * CRT/BVH.h:23, but never builds it), leaves of at most max_prims_in_node
 * triangles.  No reference counterpart for the tree itself; closest hits are
 * the same triangles (the closest hit does not depend on the tree). */
#define FF_BVH_SAH 0
#define FF_BVH_LBVH 1
/* FF_BVH_PLOC: built on the GPU by locally-ordered clustering of the
 * Morton-sorted triangles (Meister and Bittner, IEEE TVCG 2018: each round,
 * mutual nearest neighbours by union surface area within 16 positions
 * merge), subtrees collapsed into leaves of at most max_prims_in_node
 * triangles where the SAH cost (traversal 0.125 per node, 1 per triangle, as
 * CRT/BVH.cpp:237-262 prices them) says so.  No reference counterpart; the
 * SAH-quality alternative to FF_BVH_LBVH. */
#define FF_BVH_PLOC 2

/* dCamera (CRT/dCamera.cuh:9-22): inverse(proj * view), column-major. */
typedef struct {
    float inv_view_proj[16];
    float lens_radius;
    float focal_distance;
    float exposure;
    float pad;
} ff_camera;

/* One render request (replaces PathTracer's compile-time MAX_PATH_LENGTH and
 * the per-frame bounce loop, CRT/PathTracer.h:20-21, CRT/PathTracer.cpp:59-79). */
typedef struct {
    int32_t width, height;
    int32_t spp;          /* samples per pixel rendered by this call          */
    int32_t max_depth;    /* a path ends when its length exceeds max_depth    */
    uint32_t seed;
    int32_t spp_batch;    /* samples of each pixel in flight at once (film layers) */
    int32_t sample_offset;/* index of the first sample (progressive rendering)*/
    int32_t flags;        /* FF_RENDER_* */
} ff_render_req;

#define FF_RENDER_USE_GRAPH 1 /* capture the bounce loop in a hipGraph              */
#define FF_RENDER_TIMING 2    /* time every stage launch with hipEvents (stats); the
                                 stages then run one after another (otherwise each
                                 iteration's shadow traversal overlaps its extension) */

typedef struct {
    double rays_primary;   /* camera rays (closest hit)                     */
    double rays_extension; /* bounce >= 1 extension rays                     */
    double rays_shadow;    /* any-hit rays (light + BRDF-sampled visibility) */
    double paths;          /* completed samples                              */
    double ext_nodes, ext_tris;       /* boxes / triangles tested by extend  */
    double shadow_nodes, shadow_tris; /* boxes / triangles tested by shadow  */
    /* FF_RENDER_TIMING only: summed launch durations (ms) and launch counts */
    double ms_extend, ms_shadow, ms_logic, ms_shade, ms_other;
    double launches_extend, launches_shadow;
    /* BRDF-sample visibility rays not traced because the light ray of their
     * vertex was occluded (its visibility gates the whole NEE sum; the
     * reference's BRDF_L traces them anyway, CRT/dMaterial.cu:404-448) */
    double rays_shadow_gated;
} ff_ray_stats;

/* ---- errors / version ---- */
const char *ff_last_error(void);
int32_t ff_abi_version(void);

/* ---- context (device selection + reusable wavefront buffers) ----
 * Replaces PathTracer's allocation of Paths/Queues (CRT/PathTracer.cpp:98-140). */
ff_status ff_context_create(int32_t device, ff_context **out);
ff_status ff_context_destroy(ff_context *ctx);

/* ---- linear BVH built on the GPU (the FF_BVH_LBVH scene builder) ----
 * No reference counterpart: the reference declares SplitMethod::HLBVH
 * (CRT/BVH.h:23) and never builds it.  Same node format and outputs as
 * ff_bvh_build_host; build_ms_out (may be NULL) receives the device time of
 * the build kernels (key sort included, host transfers excluded). */
ff_status ff_bvh_build_gpu(ff_context *ctx, const float *positions, int32_t n_tris, int32_t max_prims,
                           ff_bvh_node *nodes_out, int32_t *ordered_out, int32_t *n_nodes_out,
                           int32_t *max_depth_out, float *build_ms_out);

/* ---- PLOC BVH built on the GPU (the FF_BVH_PLOC scene builder) ----
 * No reference counterpart (the reference's HLBVH, CRT/BVH.h:23, is never
 * built).  Same arguments, node format and outputs as ff_bvh_build_gpu. */
ff_status ff_bvh_build_ploc(ff_context *ctx, const float *positions, int32_t n_tris, int32_t max_prims,
                            ff_bvh_node *nodes_out, int32_t *ordered_out, int32_t *n_nodes_out,
                            int32_t *max_depth_out, float *build_ms_out);

/* ---- host SAH BVH build (no GPU needed) ----
 * Replaces BVHAccel(info, SplitMethod::SAH, maxPrims) + init_BVH_triangle_info
 * (CRT/BVH.cpp:52-64, CRT/dScene.cpp:324-379).  nodes_out holds 2n-1 nodes. */
ff_status ff_bvh_build_host(const float *positions, int32_t n_tris, int32_t max_prims,
                            ff_bvh_node *nodes_out, int32_t *ordered_out,
                            int32_t *n_nodes_out, int32_t *max_depth_out);

/* ---- scene: BVH build on the host, triangles reordered and uploaded ----
 * Replaces dScene::load_scene (CRT/dScene.cpp:60-80) with load_models,
 * process_mesh and reorder_primitives (CRT/kernel.cu:162-191,701-723). */
ff_status ff_scene_create(ff_context *ctx, const ff_scene_desc *desc, ff_scene **out);
ff_status ff_scene_destroy(ff_scene *scene);
ff_status ff_scene_info(const ff_scene *scene, int32_t *n_tris, int32_t *n_nodes,
                        int32_t *max_depth);
ff_status ff_scene_get_bvh(const ff_scene *scene, ff_bvh_node *nodes_out,
                           int32_t *ordered_out);

/* ---- environment light: FP64 marginal/conditional CDF build on the GPU ----
 * Replaces add_environment_light and its four CDF kernels
 * (CRT/kernel.cu:65-149,742-771).  rgb is host [h][w][3] float. */
ff_status ff_envmap_create(ff_context *ctx, const float *rgb, int32_t w, int32_t h,
                           ff_envmap **out);
ff_status ff_envmap_read(const ff_envmap *env, double *marginal_y, double *marginal_p,
                         double *conds, float *pdf, double *denom);
ff_status ff_envmap_destroy(ff_envmap *env);
/* Parameter updates between renders (inverse rendering loops); data may be
 * host or device memory, copied on `stream`.  ff_envmap_update rebuilds the
 * CDFs on the device (texels must be finite, non-negative and not all zero).
 * No reference counterpart: the reference loads its scene once. */
ff_status ff_envmap_update(ff_envmap *env, const float *rgb, void *stream);
ff_status ff_scene_set_texture(ff_scene *scene, int32_t texture, const float *data, void *stream);

/* ---- standalone wavefront stages on caller device buffers ----
 * rays: device [n][8] floats (o.x o.y o.z - d.x d.y d.z -).
 * ff_trace_closest replaces wf_extend/intersect (CRT/kernel.cu:487-523,
 *   CRT/dTriangle.cu:140-213): hits device [n][4] = (t, u, v, slot as int32);
 *   slot -1 and t = 1e32 for a miss.  two_sided != 0 uses the non-culling
 *   test of dTriangle::intersect (CRT/dTriangle.cu:63-80).
 * ff_trace_any replaces wf_shadow/intersect_shadows (CRT/kernel.cu:525-543,
 *   CRT/dTriangle.cu:215-259): occluded device [n] bytes. */
/* counts (device, 2 x uint64, may be NULL) accumulates boxes and triangles tested. */
ff_status ff_trace_closest(ff_scene *scene, const float *rays, uint32_t n, float *hits,
                           int32_t two_sided, uint64_t *counts, void *stream);
ff_status ff_trace_any(ff_scene *scene, const float *rays, uint32_t n, uint8_t *occluded,
                       uint64_t *counts, void *stream);
/* ff_trace_closest without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as ff_trace_closest; its work counters and traversal stack
 * overflow live in a fixed per-device scratch of 64 MiB + 24 KiB that the
 * library holds from context creation until the process ends, so it accepts
 * BVHs of depth <= 39 (two-child traversal variants) or <= 26 (4-wide
 * variants) and returns FF_ERR_CAPACITY beyond that.  Calls on one device
 * share that scratch: calls that overlap on one device (different streams,
 * threads or contexts) race and must be serialised by the caller. */
ff_status ff_trace_closest_noalloc(ff_scene *scene, const float *rays, uint32_t n, float *hits,
                                   int32_t two_sided, uint64_t *counts, void *stream);
/* camera-ray generation only (wf_generate + dCamera::gen_ray,
 * CRT/kernel.cu:457-485, CRT/dCamera.cu:10-36): rays device [h*w][8] for
 * sample index `sample`. */
ff_status ff_generate_rays(ff_context *ctx, const ff_camera *cam, int32_t width,
                           int32_t height, uint32_t seed, int32_t sample, float *rays,
                           void *stream);

/* ---- full render: plain per-request wavefront ----
 * Replaces PathTracer::draw -> wavefront_pathtrace (CRT/PathTracer.cpp:59-79,
 * CRT/kernel.cu:658-676).  film: device [h][w][4]; the call ADDS the radiance
 * sums of req->spp samples to xyz and req->spp to w (clear it first, as
 * PathTracer::clear_buffer does, CRT/PathTracer.cpp:150-170). */
ff_status ff_render(ff_context *ctx, ff_scene *scene, ff_envmap *env, const ff_camera *cam,
                    const ff_render_req *req, float *film, void *stream);
/* ---- batched requests: many scene/camera requests of uneven resolution,
 * spp and depth in one wavefront (spp_batch path slots per pixel of every
 * request, 0 meaning 1; a request's film equals what ff_render produces for
 * it with the same spp_batch).  All scenes must belong to
 * ctx; films must be distinct device [h][w][4] buffers (ADDED to, like
 * ff_render).  flags: FF_RENDER_USE_GRAPH captures one batch of launches in a
 * hipGraph and replays it once per sample index.  No reference counterpart:
 * the reference renders one frame per draw (CRT/PathTracer.cpp:59-79). */
typedef struct ff_grads ff_grads;
typedef struct {
    ff_scene *scene;
    ff_envmap *env;
    ff_camera camera;
    int32_t width, height, spp, max_depth;
    uint32_t seed;
    int32_t sample_offset;
    float *film;
    float *path_radiance;     /* optional: [spp][h][w][4] radiance of every path  */
    const float *d_image;     /* ff_adjoint_batch: d loss / d image, [h][w][4]    */
    const ff_grads *grads;    /* ff_adjoint_batch: gradient buffers               */
    int32_t spp_batch;        /* samples of each pixel in flight (0 = 1)          */
} ff_batch_item;
ff_status ff_render_batch(ff_context *ctx, const ff_batch_item *items, int32_t n, int32_t flags,
                          void *stream);

/* ---- differentiable rendering: path-replay adjoint ----
 * No reference counterpart (the reference is not differentiable).  The image
 * of a request is the mean of its spp samples (film.xyz / spp).
 * ff_render_paths is ff_render that also stores every path's radiance,
 * path_radiance device [spp][h][w][4] (xyz used).  ff_adjoint replays those
 * paths with the same random numbers (same request: scene, environment,
 * camera, size, spp, depth, seed, sample offset; spp_batch may differ) and
 * ADDS d loss / d parameter to the gradient buffers, given d_image = d loss /
 * d image, device [h][w][4] (xyz used).  The gradient is that of the
 * detached-sampling estimator: directions, pdfs, MIS weights and Russian
 * roulette stay those of the forward; BSDF values (albedo and roughness
 * textures) and environment radiance (texels) are differentiated.  Terms are
 * accumulated with FP32 atomics; with FF_ADJOINT_DETERMINISTIC in the flags
 * the replay runs twice (per-element bounds, then int64 fixed-point sums at
 * per-element scales) and the result is bit-reproducible.
 * ff_adjoint_batch does the same for batch items (their path_radiance,
 * d_image and grads), sharing gradient buffers across items when they point
 * to the same memory. */
struct ff_grads {
    float *env;               /* device [env_h][env_w][3], or NULL             */
    int32_t n_textures;       /* entries in textures (<= the scene's textures)  */
    float *const *textures;   /* host array of device [h][w][c] buffers (NULL entries skipped) */
};
#define FF_ADJOINT_DETERMINISTIC 4
ff_status ff_render_paths(ff_context *ctx, ff_scene *scene, ff_envmap *env, const ff_camera *cam,
                          const ff_render_req *req, float *film, float *path_radiance, void *stream);
ff_status ff_adjoint(ff_context *ctx, ff_scene *scene, ff_envmap *env, const ff_camera *cam,
                     const ff_render_req *req, const float *path_radiance, const float *d_image,
                     const ff_grads *grads, void *stream);
ff_status ff_adjoint_batch(ff_context *ctx, const ff_batch_item *items, int32_t n, int32_t flags,
                           void *stream);

/* ---- traversal kernel variants (workgroup size, LDS stack entries, the
 * number of idle lanes at which a wave of the persistent traversal fetches
 * new rays, the loop shape: 0 = one step per iteration, 1 = while-while,
 * 2 = speculative while-while (leaves postponed while other lanes walk),
 * 3, 5, 6, 7 = leaf batching (a lane queues the leaf it reaches and walks on;
 * queued leaves are tested together once 8, 8, 4 or 16 lanes hold one; shape
 * 3 queues two leaves per lane, the others one), and the
 * BVH record width: 2 = the reference's binary nodes two children at a time,
 * 4 = the same tree collapsed to 4-wide records, visited in the same leaf
 * order so hits are identical), chosen per context by the autotuner
 * (firefly_amd/autotune.py); variant 0 is the default.  The context's variant
 * applies to its renders and to ff_trace_closest / ff_trace_any.  No
 * reference counterpart (the reference has one fixed shape).  Renders and
 * traces on one context share its work counters and buffers: they must not
 * overlap on different streams. */
int32_t ff_traversal_variant_count(void);
ff_status ff_traversal_variant_info(int32_t variant, int32_t *block, int32_t *lds_entries,
                                    int32_t *refill, int32_t *while_while, int32_t *bvh_width);
ff_status ff_set_traversal_variant(ff_context *ctx, int32_t variant);
/* A second variant for the bounce iterations from `from_iteration` on (1 =
 * camera rays), whose launches trace the few surviving paths: the shapes that
 * win on the large first launches lose on small ones.  variant -1 turns it
 * off (every launch uses the context's variant).  No reference counterpart. */
ff_status ff_set_traversal_tail(ff_context *ctx, int32_t variant, int32_t from_iteration);

/* ray counts of the last ff_render on this context (synchronises its stream). */
ff_status ff_last_render_stats(ff_context *ctx, ff_ray_stats *out);

/* ---- film resolve (dFilm): average and Reinhard tonemap ----
 * Replaces the resolve at the end of wf_logic (CRT/kernel.cu:447-454):
 * avg = sum/n; c = avg*exposure; c/(c+1); rgba = (int)(255*c).  Either output
 * may be NULL. */
ff_status ff_film_resolve(const float *film, int32_t width, int32_t height, float exposure,
                          float *avg_rgb, uint8_t *rgba8, void *stream);

#ifdef __cplusplus
}
#endif
#endif
