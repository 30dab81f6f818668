/*
 * ff_oracle.h -- CPU oracle for the Firefly wavefront path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithm (JakeKurtz/Firefly, CUDA-RayTracer/) used as the checker for the
 * HIP implementation in firefly_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function cites the reference file:line it restates.  How each function
 * is pinned (closed forms, brute force, invariants) is listed in the header of
 * ff_oracle.c; functions without an independent pin are marked
 * "parity unpinned" there.
 */
#ifndef FF_ORACLE_H
#define FF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 32-byte flattened BVH node; same fields as LinearBVHNode (CRT/BVH.h:63-72). */
typedef struct {
    float bmin[3];
    float bmax[3];
    int32_t offset;   /* leaf: first primitive slot; interior: second child index */
    uint16_t n_prims; /* 0 => interior */
    uint8_t axis;
    uint8_t pad;
} or_node;

/* Material record: 16 floats.  [0] type (0 mix, 1 diffuse, 2 conductor,
 * 3 dielectric), [1..3] base colour, [4] roughness, [5] metallic,
 * [6..8] fresnel f0, [9] ior, [10] albedo texture (-1 none),
 * [11] roughness texture (-1 none), [12] metallic texture (-1 none),
 * [13] normal texture (-1 none; needs tan), [14..15] reserved.
 * Fields follow dMaterial (CRT/dMaterial.cuh:9-31). */
#define OR_MAT_MIX 0
#define OR_MAT_DIFFUSE 1
#define OR_MAT_CONDUCTOR 2
#define OR_MAT_DIELECTRIC 3

typedef struct {
    int32_t n_tris;
    const float *pos;      /* [n][3 verts][3]  original triangle order   */
    const float *nrm;      /* [n][3][3]                                  */
    const float *uv;       /* [n][3][2]                                  */
    const int32_t *mat;    /* [n]                                        */
    int32_t n_mats;
    const float *mats;     /* [n_mats][16]                               */
    int32_t n_tex;
    const float *const *tex; /* per texture: [h][w][c] float texels     */
    const int32_t *tex_w, *tex_h, *tex_c;
    int32_t env_w, env_h;  /* 0 => constant environment colour          */
    const float *env_rgb;  /* [h][w][3], row j at v = (j+0.5)/h          */
    float env_color[3];
    float cam_inv_vp[16];  /* column-major: column c, row r at [c*4+r]   */
    float lens_radius, focal, exposure;
    int32_t n_nodes;       /* BVH from or_bvh_build over pos             */
    const or_node *nodes;
    const int32_t *ordered;/* BVH slot -> original triangle              */
    const float *tan;      /* [n][3][6] tangent, bitangent per vertex, or NULL */
} or_scene;

typedef struct {
    int32_t width, height, spp, max_depth;
    uint32_t seed;
    int32_t spp_batch;     /* K: sample s accumulates into film layer s % K */
    int32_t n_threads;
    int32_t row_begin, row_end; /* rows rendered; [0,height) for a full frame */
} or_params;

typedef struct {
    double rays_primary;   /* camera rays (closest hit)                  */
    double rays_extension; /* bounce >= 1 extension rays (closest hit)   */
    double rays_shadow;    /* any-hit rays                               */
    double paths;          /* completed samples                          */
    double rays_shadow_gated; /* BRDF-sample visibility rays not traced: their
                                 light ray was occluded                  */
} or_stats;

/* ---- RNG (counter based, keyed by seed/pixel/sample/bounce/dim) ---- */
uint32_t or_rng_base(uint32_t seed, uint32_t pixel, uint32_t sample);
float or_rng(uint32_t base, uint32_t bounce, uint32_t dim);

/* ---- BVH ---- */
int32_t or_bvh_build(const float *pos, int32_t n, int32_t max_prims,
                     or_node *nodes_out, int32_t *ordered_out);
int32_t or_bvh_max_depth(const or_node *nodes, int32_t n_nodes);
/* linear BVH, as the GPU builder makes it (Morton order, Karras hierarchy) */
int32_t or_lbvh_build(const float *pos, int32_t n, int32_t max_prims,
                      or_node *nodes_out, int32_t *ordered_out);
/* the GPU PLOC builder restated (no reference counterpart); same outputs */
int32_t or_ploc_build(const float *pos, int32_t n, int32_t max_prims,
                      or_node *nodes_out, int32_t *ordered_out);

/* ---- traversal: rays are [n][6] (o.xyz, d.xyz); tris are [slot][9] (v0,v1,v2) ---- */
void or_trace_closest(const or_node *nodes, const float *tris, const float *rays,
                      int32_t n, int32_t two_sided, float *t, float *u, float *v,
                      int32_t *id);
void or_trace_closest_brute(int32_t n_tris, const float *tris, const float *rays,
                            int32_t n, int32_t two_sided, float *t, int32_t *id);
void or_trace_any(const or_node *nodes, const float *tris, const float *rays,
                  int32_t n, uint8_t *occluded);
void or_trace_any_brute(int32_t n_tris, const float *tris, const float *rays,
                        int32_t n, uint8_t *occluded);
/* the double transcendentals the oracle and the GPU share, element-wise:
 * fn 0 sin, 1 cos, 2 atan, 3 atan2(y, x), 4 asin, 5 acos (y used by atan2 only) */
void or_sincos_d(double x, double *s, double *c);
double or_atan_d(double x);
double or_atan2_d(double y, double x);
double or_asin_d(double x);
double or_acos_d(double x);
void or_math_array(int32_t fn, const double *x, const double *y, int32_t n, double *out);

/* width 2: two child boxes per interior node; width 4: the four grandchild
 * boxes per interior node (the 4-wide records) */
void or_count_traversal(const or_node *nodes, const float *tris, const float *rays,
                        int32_t n, int32_t any_hit, int32_t prune, int32_t width,
                        int64_t *nodes_tested, int64_t *tris_tested);
void or_trace_walk(const or_node *nodes, const float *tris, const float *rays, int32_t n,
                   int32_t any_hit, int32_t prune, int32_t width, float *t, int32_t *id);

/* ---- environment CDF (FP64) and sampling ---- */
void or_env_build(const float *rgb, int32_t w, int32_t h, double *marginal_y,
                  double *marginal_p, double *conds, float *pdf, double *denom);
int32_t or_upper_bound(const double *list, int32_t size, float val);
/* test hook: renders use these tables (of a w x h map, as or_env_build lays
 * them out) instead of building their own; NULL marginal_y = build again.
 * The pointers must stay valid while renders use them. */
void or_set_env_tables(const double *marginal_y, const double *marginal_p, const double *conds,
                       const float *pdf, int32_t w, int32_t h);
void or_env_sample_dir(const double *marginal_y, const double *conds, int32_t w,
                       int32_t h, float ex, float ey, int32_t *x, int32_t *y,
                       float *dir);
void or_bilinear(const float *tex, int32_t w, int32_t h, int32_t c, float u,
                 float v, float *out);
float or_env_pdf(const float *pdf, int32_t w, int32_t h, const float *wi);
void or_env_L(const float *rgb, int32_t w, int32_t h, const float *dir, float *out);

/* ---- camera ---- */
/* test hook: plant mistake m (1-7, 0 = none) in the path estimator; m = 8
 * traces the BRDF-sample visibility rays also where the light ray is occluded
 * (the reference's ray count, same image) */
void or_set_mutation(int32_t m);
/* ConcentricSampleDisk (CRT/dRandom.cu:104-126) of n points u[n][2]: out[n][2] */
void or_concentric(const float *u, int32_t n, float *out);
void or_gen_ray(const or_scene *s, int32_t width, int32_t height, int32_t px,
                int32_t py, uint32_t rng_base, float *ray6);

/* ---- BSDF pieces (reference lobes) ---- */
double or_ggx_ndf(const float *n, const float *h, float r);
double or_geo_atten(const float *wi, const float *wo, const float *n, float r);
float or_spec_pdf(const float *n, const float *wi, const float *wo, float r);
void or_spec_f(const float *n, const float *wi, const float *wo, float r, float m,
               const float *albedo, const float *fres, float *out);
void or_diff_f(const float *wi, const float *wo, float m, const float *albedo,
               const float *fres, float *out);
void or_diff_sample(const float *n, float e0, float e1, float *out);
void or_spec_sample(const float *n, const float *wo, float r, float e0, float e1,
                    float *out);
float or_fresnel_dielectric(float cos_i, float eta_i, float eta_t);

/* ---- full render: film is [h][w][4] = (sum r, sum g, sum b, samples) ---- */
void or_render(const or_scene *s, const or_params *p, float *film, or_stats *st);

/* ---- differentiable render (FP64): detached-sampling estimator of each path
 * as a function of the albedo / roughness textures and environment texels ----
 * or_diff_adjoint: d_image [H][W][3] = d loss / d (mean image); gradients are
 * ADDED to g_env [env_h][env_w][3] and g_tex[t] [h][w][c] (NULL: skipped);
 * image (may be NULL) receives the FP64 mean image. */
void or_diff_adjoint(const or_scene *s, const or_params *p, const float *d_image, double *image,
                     double *g_env, double *const *g_tex);
typedef struct or_tapes or_tapes;
or_tapes *or_tapes_record(const or_scene *s, const or_params *p);
void or_tapes_eval(const or_tapes *t, const or_scene *s, const float *const *tex, const float *env,
                   int32_t row_begin, int32_t row_end, double *image);
int32_t or_tapes_len(const or_tapes *t, int32_t pixel, int32_t sample);
void or_tapes_free(or_tapes *t);

#ifdef __cplusplus
}
#endif
#endif
