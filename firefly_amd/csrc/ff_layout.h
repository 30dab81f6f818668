// ff_layout.h -- device data layout shared by the kernels and the host driver.
//
// HBM layout (all arrays are structure-of-arrays indexed by path id or BVH
// slot; float4 fields keep every per-path record a 16-byte aligned load):
//   nodes   : the reference's binary BVH as 64-byte two-child records
//             (both child boxes + child refs; 32 B per box tested)
//   tris    : 9 floats per slot, v0 | e1 | e2 (36 B, positions only)
//   nrm/uv  : shading attributes per slot, read once per closest hit
//   paths   : ray (2 x float4, carrying the rng base and the state word in
//             their w), hit, beta+state, Lp, fext, nee[3], sray[3][2], vis[3];
//             queues of u32 path ids with device counters.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FF_BLOCK 256          // threads per workgroup of the stage kernels (4 waves of 64)
#define FF_TRAV_THREADS 524288 // threads of the persistent traversal grids (any block size)
#define FF_LDS_MIN 8          // smallest LDS stack of a traversal variant (sizes the overflow area)
// Stack overflow entries per traversal lane in the fixed device scratch of
// ff_trace_closest_noalloc (with the work counters: 64 MiB + 24 KiB per device)
#define FF_STATIC_OVF_ENTRIES 32
#define FF_GRID_CAP 4096      // grid-stride cap of the other stage kernels
#define QSTAGE_CAP 2048       // queue entries a workgroup stages in LDS before appending them

// path state word (beta.w bits; after a shading pass, fext.w)
#define ST_LEN_MASK 0xffu
#define ST_INSIDE (1u << 8)
#define ST_PREV_DELTA (1u << 9)
#define ST_NEED_L (1u << 10)
#define ST_NEED_S (1u << 11)
#define ST_NEED_D (1u << 12)
#define ST_DEAD (1u << 13)   // the continuation had zero f or pdf
// Per-triangle material word: the material's type (0-3) in the top two bits,
// so wf_logic routes a hit with one load.
#define MAT_TYPE_SHIFT 30
#define MAT_INDEX_MASK 0x3fffffffu

// device counters
#define C_EXT0 0
#define C_EXT1 1
#define C_SHADOW 2
#define C_MAT0 3
#define C_COUNT 16

// Work counters of the persistent traversal kernels: the queue is split into
// FETCH_SEGS segments with one counter each, every counter on its own 128-B
// line (device-scope atomics on one address serialise at the memory side).
#define FETCH_SEGS 64
#define FETCH_STRIDE 32
#define F_EXT 0
#define F_SHD 1
#define F_TRACE 2
#define FETCH_WORDS (3 * FETCH_SEGS * FETCH_STRIDE)
#define FETCH_CHUNK 256  // largest claim of one wave from a work counter

// ray statistics
#define S_PRIMARY 0
#define S_EXTENSION 1
#define S_SHADOW 2
#define S_PATHS 3
#define S_EXT_NODES 4
#define S_EXT_TRIS 5
#define S_SHD_NODES 6
#define S_SHD_TRIS 7
#define S_SHD_GATED 8  // BRDF-sample visibility rays not traced (light ray occluded)
#define S_COUNT 9
// ray statistics are replicated per workgroup group on separate lines
#define S_REPLICAS 64
#define S_STRIDE 16
#define S_WORDS (S_REPLICAS * S_STRIDE)

#define FF_MAT_FLOATS_DEV 16

#define MAT_MIX 0
#define MAT_DIFFUSE 1
#define MAT_CONDUCTOR 2
#define MAT_DIELECTRIC 3

struct ff_dev_camera {
    float inv_vp[16];
    float lens_radius, focal, exposure, pad;
};

struct ff_tex_desc {
    const float* data;
    int32_t w, h, c, pad;
};

struct ff_dev_scene {
    const float4* nodes;       // two-child records, 4 x float4 each
    const float4* nodes4;      // 4-wide records, 8 x float4 each
    const int32_t* leaf_count; // per slot, for leaves of 15 or more primitives
    const float* root;         // root box lo.xyz hi.xyz, root ref
    const float* tris;
    const float4* attr;        // per slot 4 x float4: normals [0..8], uvs [9..14], material word [15]
    const float* tan;          // per slot [3][6] tangent, bitangent (null: no normal textures)
    const uint32_t* mat;       // per slot: material index | material type << MAT_TYPE_SHIFT
    const float* mats;
    const ff_tex_desc* tex;
    int32_t n_nodes, n_tris, n_mats, n_tex, max_depth;
};

// Guide tables of the inverse-CDF searches: for bucket g of [0, 1) (width
// 1 / ENV_GUIDE) the first and last candidate index of the upper bound (two
// buckets per texel of a 2048-texel row, so a search is about one step).
// Entry with bit 31 set: the list is not monotone (NaN rows), search it whole.
// Bits 29 and 30 of every entry of a list: list[0] != 0 and list[1] != 0, the
// two values the search's trailing adjustment tests (no second fetch there).
#define ENV_GUIDE 4096
#define ENV_GUIDE_FULL 0x80000000u
#define ENV_GUIDE_NZ0 0x20000000u
#define ENV_GUIDE_NZ1 0x40000000u
#define ENV_GUIDE_INDEX 0x1fffffffu

struct ff_dev_env {
    const float* rgb;
    const float4* rgbp;       // [h][w] (r, g, b, pdf): the texel and pdf lookups of rendering
    const double* marg_y;
    const double* conds;
    const float* pdf;
    const uint32_t* guide_m;  // [ENV_GUIDE + 1] over marg_y
    const uint32_t* guide_c;  // [h][ENV_GUIDE + 1] over the rows of conds
    int32_t w, h;
};

// Differentiable rendering.  A gradient target is a caller's float buffer
// (environment texels [h][w][3] or a texture [h][w][c]) plus its offset in
// the call's arena of per-element bound / fixed-point accumulators.
struct ff_dev_gtarget {
    float* buf;
    uint64_t off;
};
#define GT_ENV 0      // target 0: environment; target 1 + t: scene texture t
// gradient accumulation modes of an adjoint pass
#define GM_FLOAT 0    // FP32 atomics into the caller's buffers
#define GM_BOUND 1    // deterministic pass 1: per-element max |term| and term count
#define GM_FIXED 2    // deterministic pass 2: int64 fixed-point sums at per-element scales

// request modes
#define RQ_STORE_PATHS 1u  // forward: keep each path's radiance (path_L)
#define RQ_ADJOINT 2u      // path-replay adjoint of a stored forward

// One render request: what a path slot needs to know about its frame.
// Slots [slot_base, slot_base + W*H*K) belong to the request; slot - slot_base
// = layer * W*H + pixel.  K == 1 accumulates straight into the film (one
// writer per pixel); K > 1 accumulates into layers resolved at the end.
struct ff_dev_request {
    ff_dev_scene S;
    ff_dev_env E;
    ff_dev_camera C;
    float4* film;      // [H][W] (sums, count)
    float4* layers;    // [K][H*W] when K > 1
    uint32_t slot_base, W, H, spp, depth, K, seed, sample_offset;
    uint32_t mode;     // RQ_*
    float inv_spp;     // adjoint: 1 / spp (the image is the mean of this call's samples)
    float4* path_L;    // [spp][H][W] radiance of every path (RQ_STORE_PATHS / RQ_ADJOINT)
    const float4* d_img;  // adjoint: d loss / d image [H][W] (xyz)
    const ff_dev_gtarget* gt;  // adjoint: [1 + n_tex] targets (buf null = not wanted)
};

// device-resident batch state (graph-replay safe): [0] batch index
#define B_BATCH 0
#define B_COUNT 4

struct ff_paths {
    float4* ray;     // [P][2]  (o, rng base bits) (d, state word bits: 0 on camera rays; its
                     //         inside flag makes the closest-hit test two-sided)
    float4* hit;     // [P]     (t, u, v, slot)
    float4* beta;    // [P]     (beta.xyz, state)
    float4* Lp;      // [P]     (path radiance, -)
    float4* fext;    // [P]     f * cos / pdf of the continuation (0 when f or pdf is zero), state word
    float4* nee;     // [P][3]  light term, spec term, diffuse term
    float4* sray;    // [P][3][2] shadow rays
    uint8_t* vis;    // [P][3]
    uint32_t* preq;  // [P]     request of the slot (stored only when n_req > 1)
    uint32_t* q_ext[2];
    uint32_t* q_mat[4];
    uint32_t* q_shadow;  // [P] events with a light ray: path << 2 | needS | needD << 1
    uint32_t* counters;  // [C_COUNT]
    uint32_t* fetch;     // [FETCH_WORDS] traversal work counters
    uint32_t* qbase;     // [n_req + 1] first extension-queue slot of each request's new samples
    float4* layers;      // [P] film layer accumulators (path id = layer*WH + pixel)
    unsigned long long* stats;  // [S_REPLICAS][S_STRIDE], S_COUNT used per replica
    int32_t* ovf;        // traversal stack overflow [depth][FF_TRAV_THREADS] of the extension stage
    int32_t* ovf_shd;    // the same for the shadow stage (the two run at the same time)
    uint32_t* bstate;    // [B_COUNT]
    const ff_dev_request* reqs;
    uint32_t n_req;
    uint32_t n_slots;    // total slots over all requests
    int32_t uni_scene;   // every request of the call renders the same scene
    int32_t uni_env;     // every request of the call uses the same environment
    // adjoint replay
    float4* dlt;         // [P]     d loss / d path radiance (xyz)
    uint32_t* g_max;     // arena: max |term| bits per gradient element (GM_BOUND)
    uint32_t* g_cnt;     // arena: number of terms per element (GM_BOUND)
    unsigned long long* g_fx;  // arena: fixed-point sums (GM_FIXED)
    int32_t gmode;       // GM_*
};
