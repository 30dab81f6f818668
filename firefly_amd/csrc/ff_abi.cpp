// ff_abi.cpp -- implementation of the C ABI in include/ff/ff_abi.h.
//
// Host side of the wavefront tracer: scene upload (host SAH BVH, triangles
// reordered into BVH-slot order), the FP64 environment CDF build, and the
// per-request bounce loop that replaces PathTracer::draw ->
// wavefront_pathtrace (CRT/PathTracer.cpp:59-79, CRT/kernel.cu:658-676).
// Errors are reported as ff_status plus a thread-local message; nothing here
// exits the process (the reference's checkCudaErrors exits, CRT/CudaHelpers.cpp:3-11).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <limits>
#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ff/ff_abi.h"
#include "ff_internal.h"

namespace {

thread_local std::string g_err;

ff_status fail(ff_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? FF_ERR_ALLOC : FF_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, sizeof(T) * count);
}
template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

struct PathStore {
    size_t cap = 0;
    ff_paths p{};
    void release() {
        dfree(p.ray); dfree(p.hit); dfree(p.beta); dfree(p.Lp); dfree(p.fext); dfree(p.nee);
        dfree(p.sray); dfree(p.vis); dfree(p.preq); dfree(p.q_ext[0]); dfree(p.q_ext[1]);
        for (int m = 0; m < 4; m++) dfree(p.q_mat[m]);
        dfree(p.q_shadow); dfree(p.layers); dfree(p.dlt);
        cap = 0;
    }
};

}  // namespace

struct StageRecord {
    int stage;  // 0 extend, 1 shadow, 2 logic, 3 shade, 4 other
    hipEvent_t begin, end;
};

struct ff_context {
    int device = 0;
    std::vector<hipEvent_t> event_pool;
    size_t events_used = 0;
    std::vector<StageRecord> records;
    PathStore paths;
    uint32_t* counters = nullptr;
    unsigned long long* stats = nullptr;
    int32_t* ovf = nullptr;
    size_t ovf_entries = 0;  // stack entries per grid thread
    hipStream_t last_stream = nullptr;
    hipStream_t own_stream = nullptr;  // graph capture when the caller passes the null stream
    // a second stream for the shadow traversal, which runs beside the
    // extension traversal of the same iteration (fork / join events)
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    bool have_stats = false;
    uint32_t* bstate = nullptr;
    ff_dev_request* d_reqs = nullptr;
    uint32_t* d_qbase = nullptr;  // [req_cap + 1]
    size_t req_cap = 0;
    uint32_t* fetch = nullptr;    // [FETCH_WORDS] segmented traversal work counters
    int32_t* scratch = nullptr;   // this device's fixed trace scratch (ff_trace_closest_noalloc)
    int trav_variant = 0;
    int trav_tail = -1;      // variant of the late bounce iterations (-1: none)
    int trav_tail_from = 0;  // first iteration that uses it
    // adjoint: gradient target tables and the deterministic-reduction arena
    ff_dev_gtarget* d_gt = nullptr;
    size_t gt_cap = 0;
    uint32_t* g_max = nullptr;
    uint32_t* g_cnt = nullptr;
    unsigned long long* g_fx = nullptr;
    size_t g_cap = 0;
};

struct ff_scene {
    ff_context* ctx = nullptr;
    int32_t n_tris = 0, n_mats = 0, max_depth = 0;
    std::vector<ffh::BvhNode> nodes;
    std::vector<int32_t> ordered;
    float4* d_nodes = nullptr;       // two-child records
    float4* d_nodes4 = nullptr;      // 4-wide records
    int32_t* d_leaf_count = nullptr;
    float* d_root = nullptr;
    float* d_tris = nullptr;
    float4* d_attr = nullptr;        // per slot: normals, uvs, material word (64 B)
    float* d_tan = nullptr;          // per slot [3][6] tangent frames, or none
    uint32_t* d_mat = nullptr;
    float* d_mats = nullptr;
    std::vector<float*> d_tex;
    std::vector<ff_tex_desc> tex_desc;  // host copy (sizes)
    ff_tex_desc* d_tex_desc = nullptr;
    int32_t n_tex = 0;
    bool has_type[4] = {false, false, false, false};
    ff_dev_scene view() const {
        ff_dev_scene s;
        s.nodes = d_nodes; s.nodes4 = d_nodes4; s.leaf_count = d_leaf_count; s.root = d_root; s.tris = d_tris; s.attr = d_attr; s.mat = d_mat;
        s.tan = d_tan;
        s.mats = d_mats; s.tex = d_tex_desc; s.n_nodes = (int32_t)nodes.size(); s.n_tris = n_tris;
        s.n_mats = n_mats; s.n_tex = n_tex; s.max_depth = max_depth;
        return s;
    }
};

struct ff_envmap {
    int device = 0;
    int32_t w = 0, h = 0;
    float* rgb = nullptr;
    float* lum = nullptr;
    double* row_sum = nullptr;
    double* denom = nullptr;
    double* marg_p = nullptr;
    double* marg_y = nullptr;
    double* conds = nullptr;
    float* pdf = nullptr;
    uint32_t* guide_m = nullptr;
    uint32_t* guide_c = nullptr;
    float4* rgbp = nullptr;
    ff_dev_env view() const {
        ff_dev_env e;
        e.rgb = rgb; e.rgbp = rgbp; e.marg_y = marg_y; e.conds = conds; e.pdf = pdf; e.w = w; e.h = h;
        e.guide_m = guide_m; e.guide_c = guide_c;
        return e;
    }
};

namespace {

ff_status ensure_paths(ff_context* c, size_t P) {
    if (c->paths.cap >= P) return FF_OK;
    c->paths.release();
    ff_paths& p = c->paths.p;
    HIP_TRY(dalloc(&p.ray, 2 * P));
    HIP_TRY(dalloc(&p.hit, P));
    HIP_TRY(dalloc(&p.beta, P));
    HIP_TRY(dalloc(&p.Lp, P));
    HIP_TRY(dalloc(&p.fext, P));
    HIP_TRY(dalloc(&p.nee, 3 * P));
    HIP_TRY(dalloc(&p.sray, 6 * P));
    HIP_TRY(dalloc(&p.vis, 3 * P));
    HIP_TRY(dalloc(&p.preq, P));
    HIP_TRY(dalloc(&p.q_ext[0], P));
    HIP_TRY(dalloc(&p.q_ext[1], P));
    for (int m = 0; m < 4; m++) HIP_TRY(dalloc(&p.q_mat[m], P));
    HIP_TRY(dalloc(&p.q_shadow, P));
    HIP_TRY(dalloc(&p.layers, P));
    HIP_TRY(dalloc(&p.dlt, P));
    c->paths.cap = P;
    return FF_OK;
}

// traversal stack overflow area: entries beyond the smallest LDS stack per grid thread
// Stack entries a traversal can hold: one per binary level for the two-child
// records; up to three per 4-wide record (two binary levels) for the others.
// Two areas: the extension and shadow traversals of an iteration run at the
// same time, each with its own.
ff_status ensure_ovf(ff_context* c, int32_t max_depth) {
    int32_t entries = std::max(max_depth, 3 * ((max_depth + 1) / 2));
    size_t need = entries > FF_LDS_MIN ? (size_t)(entries - FF_LDS_MIN + 1) : 1;
    if (c->ovf && c->ovf_entries >= need) return FF_OK;
    dfree(c->ovf);
    HIP_TRY(dalloc(&c->ovf, 2 * need * (size_t)FF_TRAV_THREADS));
    c->ovf_entries = need;
    return FF_OK;
}

// The reference's pre-order binary BVH re-packed as 64-byte two-child
// records: record k holds both child boxes of the k-th interior node, the
// child refs and the split axis.  Leaf refs are ~(first slot << 4 | count),
// count 15 meaning "read leaf_count[first slot]".
void pack_two_child(const std::vector<ffh::BvhNode>& nodes, std::vector<float4>& wide,
                    std::vector<int32_t>& leaf_count, float root[8], size_t n_tris) {
    std::vector<int32_t> index(nodes.size(), -1);
    int32_t nw = 0;
    for (size_t i = 0; i < nodes.size(); i++)
        if (nodes[i].n_prims == 0) index[i] = nw++;
    leaf_count.assign(std::max<size_t>(n_tris, 1), 0);
    auto ref_of = [&](size_t i) -> int32_t {
        const ffh::BvhNode& nd = nodes[i];
        if (nd.n_prims == 0) return index[i];
        int32_t c = nd.n_prims < 15 ? nd.n_prims : 15;
        if (c == 15) leaf_count[nd.offset] = nd.n_prims;
        return ~((nd.offset << 4) | c);
    };
    wide.assign((size_t)nw * 4, make_float4(0.f, 0.f, 0.f, 0.f));
    for (size_t i = 0; i < nodes.size(); i++) {
        if (nodes[i].n_prims != 0) continue;
        const ffh::BvhNode& A = nodes[i + 1];
        const ffh::BvhNode& B = nodes[nodes[i].offset];
        float4* w = &wide[(size_t)index[i] * 4];
        w[0] = make_float4(A.bmin[0], A.bmin[1], A.bmin[2], A.bmax[0]);
        w[1] = make_float4(A.bmax[1], A.bmax[2], B.bmin[0], B.bmin[1]);
        w[2] = make_float4(B.bmin[2], B.bmax[0], B.bmax[1], B.bmax[2]);
        int32_t m[4] = {ref_of(i + 1), ref_of((size_t)nodes[i].offset), (int32_t)nodes[i].axis, 0};
        std::memcpy(&w[3], m, 16);
    }
    std::memcpy(root, nodes[0].bmin, 12);
    std::memcpy(root + 3, nodes[0].bmax, 12);
    int32_t r = ref_of(0);
    std::memcpy(root + 6, &r, 4);
    root[7] = 0.f;
}

// The binary tree collapsed to 4-wide records of 128 bytes.  The record of
// interior node N holds the boxes of N's grandchildren: slots 0-1 under N's
// first child, 2-3 under its second (a child that is a leaf keeps its own
// box in the first slot of its pair and leaves the second empty), plus the
// three split axes (N, first child, second child) that order them:
//   f[0..5] = lo.x, lo.y, lo.z, hi.x, hi.y, hi.z of slots 0..3
//   f[6]    = refs: >= 0 record index, < 0 leaf ~(first slot << 4 | count)
//   f[7]    = (axis N, axis first child, axis second child, 0)
// An empty slot has the box (+inf, -inf), which every slab test misses, and
// the ref of a leaf with no triangles.  Records are numbered in pre-order.
void pack_four(const std::vector<ffh::BvhNode>& nodes, std::vector<float4>& recs,
               std::vector<int32_t>& leaf_count, float root[8], size_t n_tris) {
    leaf_count.assign(std::max<size_t>(n_tris, 1), 0);
    std::vector<int32_t> index(nodes.size(), -1);
    std::vector<size_t> order;  // binary node of each record, pre-order
    if (nodes[0].n_prims == 0) {
        std::vector<size_t> stack{0};
        while (!stack.empty()) {
            size_t i = stack.back();
            stack.pop_back();
            index[i] = (int32_t)order.size();
            order.push_back(i);
            size_t kids[2] = {i + 1, (size_t)nodes[i].offset};
            std::vector<size_t> slots;
            for (size_t c : kids) {
                if (nodes[c].n_prims == 0) {
                    slots.push_back(c + 1);
                    slots.push_back((size_t)nodes[c].offset);
                } else {
                    slots.push_back(c);
                    slots.push_back(SIZE_MAX);
                }
            }
            for (size_t k = slots.size(); k-- > 0;)
                if (slots[k] != SIZE_MAX && nodes[slots[k]].n_prims == 0) stack.push_back(slots[k]);
        }
    }
    auto leaf_ref = [&](const ffh::BvhNode& nd) -> int32_t {
        int32_t c = nd.n_prims < 15 ? nd.n_prims : 15;
        if (c == 15) leaf_count[nd.offset] = nd.n_prims;
        return ~((nd.offset << 4) | c);
    };
    auto ref_of = [&](size_t i) -> int32_t {
        return nodes[i].n_prims == 0 ? index[i] : leaf_ref(nodes[i]);
    };
    const float inf = std::numeric_limits<float>::infinity();
    recs.assign(order.size() * 8, make_float4(0.f, 0.f, 0.f, 0.f));
    for (size_t r = 0; r < order.size(); r++) {
        size_t i = order[r];
        size_t kids[2] = {i + 1, (size_t)nodes[i].offset};
        float lo[3][4], hi[3][4];
        int32_t ref[4], ax[4] = {(int32_t)nodes[i].axis, 0, 0, 0};
        for (int side = 0; side < 2; side++) {
            size_t c = kids[side];
            size_t sl[2];
            bool two = nodes[c].n_prims == 0;
            if (two) {
                sl[0] = c + 1;
                sl[1] = (size_t)nodes[c].offset;
                ax[1 + side] = (int32_t)nodes[c].axis;
            } else {
                sl[0] = c;
                sl[1] = SIZE_MAX;
            }
            for (int k = 0; k < 2; k++) {
                int s = 2 * side + k;
                if (sl[k] == SIZE_MAX) {
                    for (int a = 0; a < 3; a++) { lo[a][s] = inf; hi[a][s] = -inf; }
                    ref[s] = -1;  // a leaf of zero triangles
                } else {
                    const ffh::BvhNode& nd = nodes[sl[k]];
                    for (int a = 0; a < 3; a++) { lo[a][s] = nd.bmin[a]; hi[a][s] = nd.bmax[a]; }
                    ref[s] = ref_of(sl[k]);
                }
            }
        }
        float4* w = &recs[r * 8];
        for (int a = 0; a < 3; a++) {
            w[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
            w[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
        }
        std::memcpy(&w[6], ref, 16);
        std::memcpy(&w[7], ax, 16);
    }
    std::memcpy(root, nodes[0].bmin, 12);
    std::memcpy(root + 3, nodes[0].bmax, 12);
    int32_t r0 = ref_of(0);
    std::memcpy(root + 6, &r0, 4);
    root[7] = 0.f;
}

ff_dev_camera dev_camera(const ff_camera* cam) {
    ff_dev_camera d;
    std::memcpy(d.inv_vp, cam->inv_view_proj, sizeof(d.inv_vp));
    d.lens_radius = cam->lens_radius;
    d.focal = cam->focal_distance;
    d.exposure = cam->exposure;
    d.pad = 0.f;
    return d;
}

}  // namespace

extern "C" {

const char* ff_last_error(void) { return g_err.c_str(); }
int32_t ff_abi_version(void) { return FF_ABI_VERSION; }

ff_status ff_context_create(int32_t device, ff_context** out) {
    if (!out) return fail(FF_ERR_INVALID, "ff_context_create: out is NULL");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(FF_ERR_INVALID, "ff_context_create: device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device));
    ff_context* c = new ff_context();
    c->device = device;
    if (dalloc(&c->counters, C_COUNT) != hipSuccess || dalloc(&c->stats, S_WORDS) != hipSuccess ||
        dalloc(&c->fetch, FETCH_WORDS) != hipSuccess || hipMemset(c->fetch, 0, 4 * FETCH_WORDS) != hipSuccess ||
        dalloc(&c->bstate, B_COUNT) != hipSuccess) {
        dfree(c->counters);
        dfree(c->stats);
        dfree(c->fetch);
        dfree(c->bstate);
        delete c;
        return fail(FF_ERR_ALLOC, "ff_context_create: counter allocation failed");
    }
    // the fixed trace scratch's address on this device (current here)
    hipError_t se = ffk::static_scratch(device, &c->scratch);
    if (se != hipSuccess) {
        ff_context_destroy(c);
        return fail(FF_ERR_HIP, "ff_context_create: address of the fixed trace scratch: %s", hipGetErrorString(se));
    }
    *out = c;
    return FF_OK;
}

ff_status ff_context_destroy(ff_context* c) {
    if (!c) return fail(FF_ERR_INVALID, "ff_context_destroy: NULL context");
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    c->paths.release();
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    dfree(c->bstate);
    dfree(c->d_reqs);
    dfree(c->d_qbase);
    dfree(c->fetch);
    dfree(c->counters);
    dfree(c->stats);
    dfree(c->ovf);
    dfree(c->d_gt);
    dfree(c->g_max);
    dfree(c->g_cnt);
    dfree(c->g_fx);
    delete c;
    return FF_OK;
}

ff_status ff_bvh_build_host(const float* positions, int32_t n_tris, int32_t max_prims, ff_bvh_node* nodes_out,
                            int32_t* ordered_out, int32_t* n_nodes_out, int32_t* max_depth_out) {
    if (!positions || n_tris <= 0 || max_prims <= 0)
        return fail(FF_ERR_INVALID, "ff_bvh_build_host: need positions, n_tris > 0, max_prims > 0");
    std::vector<ffh::BvhNode> nodes;
    std::vector<int32_t> ordered;
    ffh::build_bvh_sah(positions, n_tris, max_prims, nodes, ordered);
    if (nodes_out) std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(ffh::BvhNode));
    if (ordered_out) std::memcpy(ordered_out, ordered.data(), ordered.size() * sizeof(int32_t));
    if (n_nodes_out) *n_nodes_out = (int32_t)nodes.size();
    if (max_depth_out) *max_depth_out = ffh::bvh_max_depth(nodes);
    return FF_OK;
}

ff_status ff_bvh_build_gpu(ff_context* ctx, const float* positions, int32_t n_tris, int32_t max_prims,
                           ff_bvh_node* nodes_out, int32_t* ordered_out, int32_t* n_nodes_out, int32_t* max_depth_out,
                           float* build_ms_out) {
    if (!ctx || !positions || n_tris <= 0 || max_prims <= 0)
        return fail(FF_ERR_INVALID, "ff_bvh_build_gpu: need a context, positions, n_tris > 0, max_prims > 0");
    if (n_tris >= (1 << 27))
        return fail(FF_ERR_CAPACITY, "ff_bvh_build_gpu: %d triangles (limit 2^27 - 1)", n_tris);
    for (size_t k = 0; k < (size_t)n_tris * 9; k++)
        if (!std::isfinite(positions[k]))
            return fail(FF_ERR_INVALID, "ff_bvh_build_gpu: triangle %d has a non-finite vertex coordinate", (int)(k / 9));
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<ffh::BvhNode> nodes;
    std::vector<int32_t> ordered;
    float ms = 0.f;
    HIP_TRY(ffk::build_lbvh(positions, n_tris, max_prims, nodes, ordered, &ms));
    if (nodes_out) std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(ffh::BvhNode));
    if (ordered_out) std::memcpy(ordered_out, ordered.data(), ordered.size() * sizeof(int32_t));
    if (n_nodes_out) *n_nodes_out = (int32_t)nodes.size();
    if (max_depth_out) *max_depth_out = ffh::bvh_max_depth(nodes);
    if (build_ms_out) *build_ms_out = ms;
    return FF_OK;
}

ff_status ff_bvh_build_ploc(ff_context* ctx, const float* positions, int32_t n_tris, int32_t max_prims,
                            ff_bvh_node* nodes_out, int32_t* ordered_out, int32_t* n_nodes_out, int32_t* max_depth_out,
                            float* build_ms_out) {
    if (!ctx || !positions || n_tris <= 0 || max_prims <= 0)
        return fail(FF_ERR_INVALID, "ff_bvh_build_ploc: need a context, positions, n_tris > 0, max_prims > 0");
    if (n_tris >= (1 << 27))
        return fail(FF_ERR_CAPACITY, "ff_bvh_build_ploc: %d triangles (limit 2^27 - 1)", n_tris);
    for (size_t k = 0; k < (size_t)n_tris * 9; k++)
        if (!std::isfinite(positions[k]))
            return fail(FF_ERR_INVALID, "ff_bvh_build_ploc: triangle %d has a non-finite vertex coordinate", (int)(k / 9));
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<ffh::BvhNode> nodes;
    std::vector<int32_t> ordered;
    float ms = 0.f;
    HIP_TRY(ffk::build_ploc(positions, n_tris, max_prims, nodes, ordered, &ms));
    if (nodes_out) std::memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(ffh::BvhNode));
    if (ordered_out) std::memcpy(ordered_out, ordered.data(), ordered.size() * sizeof(int32_t));
    if (n_nodes_out) *n_nodes_out = (int32_t)nodes.size();
    if (max_depth_out) *max_depth_out = ffh::bvh_max_depth(nodes);
    if (build_ms_out) *build_ms_out = ms;
    return FF_OK;
}

ff_status ff_scene_create(ff_context* ctx, const ff_scene_desc* d, ff_scene** out) {
    if (!ctx || !d || !out) return fail(FF_ERR_INVALID, "ff_scene_create: NULL argument");
    *out = nullptr;
    if (d->n_tris <= 0 || !d->positions || !d->normals || !d->material_ids)
        return fail(FF_ERR_INVALID, "ff_scene_create: empty scene or missing positions/normals/material ids");
    if (d->n_materials <= 0 || !d->materials) return fail(FF_ERR_INVALID, "ff_scene_create: no materials");
    if (d->builder != FF_BVH_SAH && d->builder != FF_BVH_LBVH && d->builder != FF_BVH_PLOC)
        return fail(FF_ERR_INVALID, "ff_scene_create: unknown BVH builder %d", d->builder);
    if ((uint32_t)d->n_materials > MAT_INDEX_MASK)
        return fail(FF_ERR_CAPACITY, "ff_scene_create: %d materials (limit 2^30 - 1)", d->n_materials);
    if (d->n_tris >= (1 << 27))
        return fail(FF_ERR_CAPACITY, "ff_scene_create: %d triangles (limit 2^27 - 1)", d->n_tris);
    // finite vertices give boxes with lo <= hi, which the traversal's box test relies on
    for (size_t k = 0; k < (size_t)d->n_tris * 9; k++)
        if (!std::isfinite(d->positions[k]))
            return fail(FF_ERR_INVALID, "ff_scene_create: triangle %d has a non-finite vertex coordinate",
                        (int)(k / 9));
    for (int32_t i = 0; i < d->n_tris; i++)
        if (d->material_ids[i] < 0 || d->material_ids[i] >= d->n_materials)
            return fail(FF_ERR_INVALID, "ff_scene_create: triangle %d material %d out of range", i, d->material_ids[i]);
    for (int32_t m = 0; m < d->n_materials; m++) {
        const float* M = d->materials + (size_t)m * FF_MAT_FLOATS;
        int type = (int)M[0];
        if (type < 0 || type > 3) return fail(FF_ERR_INVALID, "ff_scene_create: material %d type %d", m, type);
        for (int k = 10; k <= 13; k++) {
            int t = (int)M[k];
            if (t >= d->n_textures || t < -1)
                return fail(FF_ERR_INVALID, "ff_scene_create: material %d texture %d out of range", m, t);
            if (t >= 0 && d->tex_c[t] != (k == 10 || k == 13 ? 3 : 1))
                return fail(FF_ERR_INVALID, "ff_scene_create: material %d texture %d has %d channels", m, t, d->tex_c[t]);
        }
        if ((int)M[13] >= 0 && !d->tangents)
            return fail(FF_ERR_INVALID, "ff_scene_create: material %d has a normal texture but the scene no tangents", m);
    }
    for (int32_t t = 0; t < d->n_textures; t++)
        if (!d->textures || !d->textures[t] || d->tex_w[t] <= 0 || d->tex_h[t] <= 0 ||
            (d->tex_c[t] != 1 && d->tex_c[t] != 3))
            return fail(FF_ERR_INVALID, "ff_scene_create: texture %d invalid", t);
    HIP_TRY(hipSetDevice(ctx->device));

    ff_scene* s = new ff_scene();
    s->ctx = ctx;
    s->n_tris = d->n_tris;
    s->n_mats = d->n_materials;
    int32_t max_prims = d->max_prims_in_node > 0 ? d->max_prims_in_node : 8;
    if (d->builder == FF_BVH_LBVH || d->builder == FF_BVH_PLOC) {
        hipError_t be = d->builder == FF_BVH_LBVH
                            ? ffk::build_lbvh(d->positions, d->n_tris, max_prims, s->nodes, s->ordered)
                            : ffk::build_ploc(d->positions, d->n_tris, max_prims, s->nodes, s->ordered);
        if (be != hipSuccess) {
            ff_scene_destroy(s);
            return fail(FF_ERR_HIP, "ff_scene_create: GPU BVH build failed: %s", hipGetErrorString(be));
        }
    } else {
        ffh::build_bvh_sah(d->positions, d->n_tris, max_prims, s->nodes, s->ordered);
    }
    s->max_depth = ffh::bvh_max_depth(s->nodes);
    size_t n = (size_t)d->n_tris;
    // reorder_primitives (CRT/kernel.cu:184-191,701-712) done on the host while
    // packing the positions-only records v0 | e1 | e2.
    std::vector<float> tris(9 * n), attr(16 * n, 0.f), tan(d->tangents ? 18 * n : 0);
    std::vector<uint32_t> mat(n);
    for (size_t i = 0; i < n; i++) {
        size_t src = (size_t)s->ordered[i];
        const float* p = d->positions + src * 9;
        float* t = &tris[i * 9];
        for (int k = 0; k < 3; k++) {
            t[k] = p[k];
            t[3 + k] = p[3 + k] - p[k];
            t[6 + k] = p[6 + k] - p[k];
        }
        std::memcpy(&attr[i * 16], d->normals + src * 9, 36);
        if (d->uvs) std::memcpy(&attr[i * 16 + 9], d->uvs + src * 6, 24);
        if (d->tangents) std::memcpy(&tan[i * 18], d->tangents + src * 18, 72);
        uint32_t m = (uint32_t)d->material_ids[src];
        mat[i] = m | ((uint32_t)(int)d->materials[(size_t)m * FF_MAT_FLOATS] << MAT_TYPE_SHIFT);
        std::memcpy(&attr[i * 16 + 15], &mat[i], 4);
    }
    for (int32_t m = 0; m < d->n_materials; m++) s->has_type[(int)d->materials[(size_t)m * FF_MAT_FLOATS]] = true;
    auto cleanup = [&](ff_status st) { ff_scene_destroy(s); return st; };
    std::vector<float4> wide;
    std::vector<int32_t> leaf_count;
    float root[8];
    pack_two_child(s->nodes, wide, leaf_count, root, n);
    std::vector<float4> four;
    std::vector<int32_t> leaf_count4;
    float root4[8];
    pack_four(s->nodes, four, leaf_count4, root4, n);
    // the traversal addresses records by 32-bit byte offsets (64-B and 128-B records)
    if (wide.size() / 4 >= (size_t(1) << 26) || four.size() / 8 >= (size_t(1) << 25))
        return cleanup(fail(FF_ERR_CAPACITY, "ff_scene_create: %zu BVH records exceed the 2^25 4-wide / 2^26 "
                                            "two-child limit of the traversal kernels", wide.size() / 4));
    if (dalloc(&s->d_nodes, wide.size()) != hipSuccess || dalloc(&s->d_nodes4, four.size()) != hipSuccess ||
        dalloc(&s->d_leaf_count, leaf_count.size()) != hipSuccess ||
        dalloc(&s->d_root, 8) != hipSuccess || dalloc(&s->d_tris, 9 * n) != hipSuccess ||
        dalloc(&s->d_attr, 4 * n) != hipSuccess ||
        dalloc(&s->d_mat, n) != hipSuccess || dalloc(&s->d_mats, (size_t)FF_MAT_FLOATS * d->n_materials) != hipSuccess ||
        (d->tangents && dalloc(&s->d_tan, 18 * n) != hipSuccess))
        return cleanup(fail(FF_ERR_ALLOC, "ff_scene_create: device allocation failed"));
    hipError_t e = hipSuccess;
    if (!wide.empty()) e = hipMemcpy(s->d_nodes, wide.data(), wide.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess && !four.empty())
        e = hipMemcpy(s->d_nodes4, four.data(), four.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(s->d_leaf_count, leaf_count.data(), leaf_count.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_root, root, sizeof(root), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_tris, tris.data(), 36 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_attr, attr.data(), 64 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess && d->tangents) e = hipMemcpy(s->d_tan, tan.data(), 72 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d_mat, mat.data(), 4 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(s->d_mats, d->materials, sizeof(float) * FF_MAT_FLOATS * d->n_materials, hipMemcpyHostToDevice);
    if (e != hipSuccess) return cleanup(fail(FF_ERR_HIP, "ff_scene_create: upload: %s", hipGetErrorString(e)));
    s->n_tex = d->n_textures;
    std::vector<ff_tex_desc> descs((size_t)std::max(1, d->n_textures));
    for (int32_t t = 0; t < d->n_textures; t++) {
        size_t cnt = (size_t)d->tex_w[t] * d->tex_h[t] * d->tex_c[t];
        float* p = nullptr;
        if (dalloc(&p, cnt) != hipSuccess) return cleanup(fail(FF_ERR_ALLOC, "ff_scene_create: texture alloc"));
        s->d_tex.push_back(p);
        if (hipMemcpy(p, d->textures[t], cnt * 4, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(FF_ERR_HIP, "ff_scene_create: texture upload"));
        descs[t].data = p; descs[t].w = d->tex_w[t]; descs[t].h = d->tex_h[t]; descs[t].c = d->tex_c[t]; descs[t].pad = 0;
    }
    s->tex_desc = descs;
    if (dalloc(&s->d_tex_desc, descs.size()) != hipSuccess ||
        hipMemcpy(s->d_tex_desc, descs.data(), descs.size() * sizeof(ff_tex_desc), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(FF_ERR_ALLOC, "ff_scene_create: texture table"));
    // the context is only read here (its device), so scenes may be created
    // concurrently from several host threads; stack space is sized per trace
    *out = s;
    return FF_OK;
}

ff_status ff_scene_destroy(ff_scene* s) {
    if (!s) return fail(FF_ERR_INVALID, "ff_scene_destroy: NULL scene");
    (void)hipDeviceSynchronize();
    dfree(s->d_nodes); dfree(s->d_nodes4); dfree(s->d_leaf_count); dfree(s->d_root); dfree(s->d_tris); dfree(s->d_attr); dfree(s->d_tan); dfree(s->d_mat); dfree(s->d_mats);
    for (auto& p : s->d_tex) dfree(p);
    dfree(s->d_tex_desc);
    delete s;
    return FF_OK;
}

ff_status ff_scene_info(const ff_scene* s, int32_t* n_tris, int32_t* n_nodes, int32_t* max_depth) {
    if (!s) return fail(FF_ERR_INVALID, "ff_scene_info: NULL scene");
    if (n_tris) *n_tris = s->n_tris;
    if (n_nodes) *n_nodes = (int32_t)s->nodes.size();
    if (max_depth) *max_depth = s->max_depth;
    return FF_OK;
}

ff_status ff_scene_get_bvh(const ff_scene* s, ff_bvh_node* nodes_out, int32_t* ordered_out) {
    if (!s) return fail(FF_ERR_INVALID, "ff_scene_get_bvh: NULL scene");
    if (nodes_out) std::memcpy(nodes_out, s->nodes.data(), s->nodes.size() * 32);
    if (ordered_out) std::memcpy(ordered_out, s->ordered.data(), s->ordered.size() * 4);
    return FF_OK;
}

ff_status ff_envmap_create(ff_context* ctx, const float* rgb, int32_t w, int32_t h, ff_envmap** out) {
    if (!ctx || !rgb || !out) return fail(FF_ERR_INVALID, "ff_envmap_create: NULL argument");
    *out = nullptr;
    if (w < 2 || h < 2) return fail(FF_ERR_INVALID, "ff_envmap_create: map must be at least 2x2 (got %dx%d)", w, h);
    if ((uint32_t)w > ENV_GUIDE_INDEX || (uint32_t)h > ENV_GUIDE_INDEX)
        return fail(FF_ERR_INVALID, "ff_envmap_create: map side above %u texels", ENV_GUIDE_INDEX);
    double total = 0.0;
    for (size_t i = 0; i < (size_t)w * h * 3; i++) {
        if (!(rgb[i] >= 0.f) || rgb[i] > 3.0e38f)
            return fail(FF_ERR_INVALID, "ff_envmap_create: texel value %zu is negative or not finite", i / 3);
        total += rgb[i];
    }
    if (total == 0.0) return fail(FF_ERR_INVALID, "ff_envmap_create: all-zero map has no sampling distribution");
    HIP_TRY(hipSetDevice(ctx->device));
    ff_envmap* e = new ff_envmap();
    e->device = ctx->device;
    e->w = w;
    e->h = h;
    size_t wh = (size_t)w * h;
    if (dalloc(&e->rgb, 3 * wh) != hipSuccess || dalloc(&e->lum, wh) != hipSuccess ||
        dalloc(&e->row_sum, h) != hipSuccess || dalloc(&e->denom, 1) != hipSuccess ||
        dalloc(&e->marg_p, h) != hipSuccess || dalloc(&e->marg_y, h) != hipSuccess ||
        dalloc(&e->conds, wh) != hipSuccess || dalloc(&e->pdf, wh) != hipSuccess ||
        dalloc(&e->guide_m, ENV_GUIDE + 1) != hipSuccess ||
        dalloc(&e->guide_c, (size_t)h * (ENV_GUIDE + 1)) != hipSuccess || dalloc(&e->rgbp, wh) != hipSuccess) {
        ff_envmap_destroy(e);
        return fail(FF_ERR_ALLOC, "ff_envmap_create: device allocation failed");
    }
    hipError_t err = hipMemcpy(e->rgb, rgb, 12 * wh, hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = ffk::launch_env_build(e->rgb, w, h, e->lum, e->row_sum, e->denom, e->marg_p, e->marg_y, e->conds,
                                    e->pdf, e->guide_m, e->guide_c, e->rgbp, nullptr);
    if (err == hipSuccess) err = hipDeviceSynchronize();
    if (err != hipSuccess) {
        ff_envmap_destroy(e);
        return fail(FF_ERR_HIP, "ff_envmap_create: CDF build: %s", hipGetErrorString(err));
    }
    *out = e;
    return FF_OK;
}

ff_status ff_envmap_update(ff_envmap* e, const float* rgb, void* stream) {
    if (!e || !rgb) return fail(FF_ERR_INVALID, "ff_envmap_update: NULL argument");
    HIP_TRY(hipSetDevice(e->device));
    size_t wh = (size_t)e->w * e->h;
    hipStream_t hs = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(e->rgb, rgb, 12 * wh, hipMemcpyDefault, hs));
    HIP_TRY(ffk::launch_env_build(e->rgb, e->w, e->h, e->lum, e->row_sum, e->denom, e->marg_p, e->marg_y, e->conds,
                                  e->pdf, e->guide_m, e->guide_c, e->rgbp, hs));
    return FF_OK;
}

ff_status ff_scene_set_texture(ff_scene* s, int32_t t, const float* data, void* stream) {
    if (!s || !data) return fail(FF_ERR_INVALID, "ff_scene_set_texture: NULL argument");
    if (t < 0 || t >= s->n_tex) return fail(FF_ERR_INVALID, "ff_scene_set_texture: texture %d of %d", t, s->n_tex);
    HIP_TRY(hipSetDevice(s->ctx->device));
    const ff_tex_desc& d = s->tex_desc[(size_t)t];
    HIP_TRY(hipMemcpyAsync(s->d_tex[(size_t)t], data, sizeof(float) * (size_t)d.w * d.h * d.c, hipMemcpyDefault,
                           (hipStream_t)stream));
    return FF_OK;
}

ff_status ff_envmap_read(const ff_envmap* e, double* marginal_y, double* marginal_p, double* conds, float* pdf,
                         double* denom) {
    if (!e) return fail(FF_ERR_INVALID, "ff_envmap_read: NULL env");
    size_t wh = (size_t)e->w * e->h;
    if (marginal_y) HIP_TRY(hipMemcpy(marginal_y, e->marg_y, 8 * (size_t)e->h, hipMemcpyDeviceToHost));
    if (marginal_p) HIP_TRY(hipMemcpy(marginal_p, e->marg_p, 8 * (size_t)e->h, hipMemcpyDeviceToHost));
    if (conds) HIP_TRY(hipMemcpy(conds, e->conds, 8 * wh, hipMemcpyDeviceToHost));
    if (pdf) HIP_TRY(hipMemcpy(pdf, e->pdf, 4 * wh, hipMemcpyDeviceToHost));
    if (denom) HIP_TRY(hipMemcpy(denom, e->denom, 8, hipMemcpyDeviceToHost));
    return FF_OK;
}

ff_status ff_envmap_destroy(ff_envmap* e) {
    if (!e) return fail(FF_ERR_INVALID, "ff_envmap_destroy: NULL env");
    (void)hipDeviceSynchronize();
    dfree(e->rgb); dfree(e->lum); dfree(e->row_sum); dfree(e->denom); dfree(e->marg_p); dfree(e->marg_y);
    dfree(e->conds); dfree(e->pdf); dfree(e->guide_m); dfree(e->guide_c); dfree(e->rgbp);
    delete e;
    return FF_OK;
}

ff_status ff_trace_closest(ff_scene* s, const float* rays, uint32_t n, float* hits, int32_t two_sided,
                           uint64_t* counts, void* stream) {
    if (!s || (!rays && n) || (!hits && n)) return fail(FF_ERR_INVALID, "ff_trace_closest: NULL argument");
    if (n == 0) return FF_OK;
    HIP_TRY(hipSetDevice(s->ctx->device));
    ff_status st = ensure_ovf(s->ctx, s->max_depth);
    if (st != FF_OK) return st;
    HIP_TRY(ffk::launch_trace_closest(s->view(), (const float4*)rays, n, (float4*)hits, two_sided, s->ctx->fetch,
                                      s->ctx->ovf,
                                      (unsigned long long*)counts, s->ctx->trav_variant, (hipStream_t)stream));
    return FF_OK;
}

ff_status ff_trace_closest_noalloc(ff_scene* s, const float* rays, uint32_t n, float* hits, int32_t two_sided,
                                   uint64_t* counts, void* stream) {
    if (!s || (!rays && n) || (!hits && n)) return fail(FF_ERR_INVALID, "ff_trace_closest_noalloc: NULL argument");
    if (n == 0) return FF_OK;
    ff_context* c = s->ctx;
    // stack entries a traversal can hold with the variant's record width
    // (one per binary level; up to three per 4-wide record), beyond the
    // smallest LDS stack of any variant
    const int32_t d = s->max_depth;
    const int32_t width = ffk::kTravWidth[c->trav_variant];
    const int32_t entries = width == 4 ? std::max(d, 3 * ((d + 1) / 2)) : d;
    const int32_t need = entries > FF_LDS_MIN ? entries - FF_LDS_MIN + 1 : 1;
    if (need > FF_STATIC_OVF_ENTRIES)
        return fail(FF_ERR_CAPACITY,
                    "ff_trace_closest_noalloc: a BVH of depth %d needs %d stack overflow entries per lane with "
                    "%d-wide records; the fixed scratch holds %d (BVH depth <= %d with two-child records, <= %d "
                    "with 4-wide ones); use ff_trace_closest",
                    d, need, width, FF_STATIC_OVF_ENTRIES, FF_STATIC_OVF_ENTRIES + FF_LDS_MIN - 1,
                    2 * ((FF_STATIC_OVF_ENTRIES + FF_LDS_MIN - 1) / 3));
    if (!c->scratch) return fail(FF_ERR_HIP, "ff_trace_closest_noalloc: no fixed scratch on device %d", c->device);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(ffk::launch_trace_closest(s->view(), (const float4*)rays, n, (float4*)hits, two_sided,
                                      (uint32_t*)c->scratch, c->scratch + FETCH_WORDS, (unsigned long long*)counts,
                                      c->trav_variant, (hipStream_t)stream));
    return FF_OK;
}

ff_status ff_trace_any(ff_scene* s, const float* rays, uint32_t n, uint8_t* occluded, uint64_t* counts,
                       void* stream) {
    if (!s || (!rays && n) || (!occluded && n)) return fail(FF_ERR_INVALID, "ff_trace_any: NULL argument");
    if (n == 0) return FF_OK;
    HIP_TRY(hipSetDevice(s->ctx->device));
    ff_status st = ensure_ovf(s->ctx, s->max_depth);
    if (st != FF_OK) return st;
    HIP_TRY(ffk::launch_trace_any(s->view(), (const float4*)rays, n, occluded, s->ctx->fetch, s->ctx->ovf,
                                  (unsigned long long*)counts, s->ctx->trav_variant, (hipStream_t)stream));
    return FF_OK;
}

ff_status ff_generate_rays(ff_context* ctx, const ff_camera* cam, int32_t width, int32_t height, uint32_t seed,
                           int32_t sample, float* rays, void* stream) {
    if (!ctx || !cam || !rays) return fail(FF_ERR_INVALID, "ff_generate_rays: NULL argument");
    if (width <= 0 || height <= 0 || sample < 0) return fail(FF_ERR_INVALID, "ff_generate_rays: bad size");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(ffk::launch_gen_rays(dev_camera(cam), width, height, seed, (uint32_t)sample, (float4*)rays,
                                 (hipStream_t)stream));
    return FF_OK;
}

}  // extern "C"

namespace {

struct HostReq {
    ff_scene* s;
    ff_envmap* e;
    ff_camera cam;
    int32_t W, H, spp, depth, K, sample_offset;
    uint32_t seed;
    float* film;
    uint32_t mode = 0;               // RQ_*
    float* path_L = nullptr;         // [spp][H][W][4]
    const float* d_img = nullptr;    // adjoint: [H][W][4]
    const ff_grads* grads = nullptr; // adjoint
};

ff_status check_req(const char* who, ff_context* ctx, const HostReq& q) {
    if (!q.s || !q.e || (!q.film && !(q.mode & RQ_ADJOINT)))
        return fail(FF_ERR_INVALID, "%s: NULL scene, environment or film", who);
    if ((q.mode & (RQ_STORE_PATHS | RQ_ADJOINT)) && !q.path_L)
        return fail(FF_ERR_INVALID, "%s: NULL path radiance buffer", who);
    if ((q.mode & RQ_ADJOINT) && (!q.d_img || !q.grads))
        return fail(FF_ERR_INVALID, "%s: NULL d_image or gradient targets", who);
    if ((q.mode & RQ_ADJOINT) && (q.grads->n_textures < 0 || q.grads->n_textures > q.s->n_tex ||
                                  (q.grads->n_textures > 0 && !q.grads->textures)))
        return fail(FF_ERR_INVALID, "%s: %d gradient textures for a scene with %d", who, q.grads->n_textures,
                    q.s->n_tex);
    if (q.s->ctx != ctx) return fail(FF_ERR_INVALID, "%s: scene belongs to another context", who);
    if (q.W <= 0 || q.H <= 0 || q.spp <= 0 || q.depth < 1 || q.depth > 250 || q.sample_offset < 0 || q.K < 1)
        return fail(FF_ERR_INVALID, "%s: invalid request (%dx%d, spp %d, depth %d)", who, q.W, q.H, q.spp, q.depth);
    return FF_OK;
}

// The wavefront over a table of requests (one frame, or many uneven ones).
// Per batch b every request renders samples [b*K, b*K + K) of each pixel;
// the batch's bounce loop runs max depth + 1 iterations of
// extend -> shadow -> logic -> shade, all stage kernels reading their work
// counts from device counters (no host synchronisation inside a render).
ff_status render_requests(ff_context* ctx, const std::vector<HostReq>& reqs, int32_t flags, hipStream_t hs,
                          const char* who) {
    uint64_t n_slots = 0;
    int32_t max_tree = 0, max_depth = 0;
    int64_t n_batches = 0;
    bool has_type[4] = {false, false, false, false};
    for (const HostReq& q : reqs) {
        n_slots += (uint64_t)q.W * (uint64_t)q.H * (uint64_t)q.K;
        max_tree = std::max(max_tree, q.s->max_depth);
        max_depth = std::max(max_depth, q.depth);
        n_batches = std::max<int64_t>(n_batches, (q.spp + q.K - 1) / q.K);
        for (int m = 0; m < 4; m++) has_type[m] = has_type[m] || q.s->has_type[m];
    }
    if (n_slots > (1ull << 30))
        return fail(FF_ERR_CAPACITY, "%s: %llu paths in flight exceed 2^30", who, (unsigned long long)n_slots);
    HIP_TRY(hipSetDevice(ctx->device));
    ff_status st = ensure_paths(ctx, (size_t)n_slots);
    if (st != FF_OK) return st;
    st = ensure_ovf(ctx, max_tree);
    if (st != FF_OK) return st;
    const bool timing = (flags & FF_RENDER_TIMING) != 0;
    const bool graph = (flags & FF_RENDER_USE_GRAPH) != 0 && !timing;
    if (graph && !hs) {
        if (!ctx->own_stream) HIP_TRY(hipStreamCreate(&ctx->own_stream));
        hs = ctx->own_stream;
    }
    if (!ctx->side) {
        HIP_TRY(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming));
    }
    HIP_TRY(ffk::prepare_traversal(ctx->trav_variant));
    if (ctx->trav_tail >= 0) HIP_TRY(ffk::prepare_traversal(ctx->trav_tail));
    ff_paths p = ctx->paths.p;
    const bool adjoint = (reqs[0].mode & RQ_ADJOINT) != 0;
    for (const HostReq& q : reqs)
        if (((q.mode & RQ_ADJOINT) != 0) != adjoint)
            return fail(FF_ERR_INVALID, "%s: forward and adjoint requests cannot share a call", who);
    const bool deterministic = adjoint && (flags & FF_ADJOINT_DETERMINISTIC) != 0;

    // gradient targets: [1 + n_tex] per request; one arena slot range per
    // distinct caller buffer (requests may share a texture's gradient)
    struct Target { float* buf; uint64_t off, n; };
    std::vector<Target> targets;
    std::vector<ff_dev_gtarget> gt;
    std::vector<size_t> gt_base(reqs.size(), 0);
    uint64_t arena = 0;
    if (adjoint) {
        auto target_of = [&](float* buf, uint64_t n, ff_dev_gtarget& out) -> ff_status {
            out.buf = buf;
            out.off = 0;
            if (!buf) return FF_OK;
            for (const Target& t : targets)
                if (t.buf == buf) {
                    if (t.n != n) return fail(FF_ERR_INVALID, "%s: one gradient buffer bound to targets of %llu and %llu elements",
                                              who, (unsigned long long)t.n, (unsigned long long)n);
                    out.off = t.off;
                    return FF_OK;
                }
            targets.push_back({buf, arena, n});
            out.off = arena;
            arena += n;
            return FF_OK;
        };
        for (size_t r = 0; r < reqs.size(); r++) {
            const HostReq& q = reqs[r];
            gt_base[r] = gt.size();
            ff_dev_gtarget e{};
            ff_status st = target_of(q.grads->env, (uint64_t)q.e->w * q.e->h * 3, e);
            if (st != FF_OK) return st;
            gt.push_back(e);
            for (int32_t t = 0; t < q.s->n_tex; t++) {
                ff_dev_gtarget g{};
                float* buf = t < q.grads->n_textures ? q.grads->textures[t] : nullptr;
                const ff_tex_desc& d = q.s->tex_desc[(size_t)t];
                st = target_of(buf, (uint64_t)d.w * d.h * d.c, g);
                if (st != FF_OK) return st;
                gt.push_back(g);
            }
        }
        if (ctx->gt_cap < gt.size()) {
            dfree(ctx->d_gt);
            HIP_TRY(dalloc(&ctx->d_gt, gt.size()));
            ctx->gt_cap = gt.size();
        }
        if (deterministic && ctx->g_cap < arena) {
            dfree(ctx->g_max); dfree(ctx->g_cnt); dfree(ctx->g_fx);
            ctx->g_cap = 0;
            HIP_TRY(dalloc(&ctx->g_max, arena));
            HIP_TRY(dalloc(&ctx->g_cnt, arena));
            HIP_TRY(dalloc(&ctx->g_fx, arena));
            ctx->g_cap = arena;
        }
    }

    // request table (uploaded once the stream's earlier work is done with it)
    std::vector<ff_dev_request> table(reqs.size());
    uint64_t base = 0;
    for (size_t r = 0; r < reqs.size(); r++) {
        const HostReq& q = reqs[r];
        ff_dev_request& d = table[r];
        d.S = q.s->view();
        d.E = q.e->view();
        d.C = dev_camera(&q.cam);
        d.film = (float4*)q.film;
        d.layers = p.layers + base;
        d.slot_base = (uint32_t)base;
        d.W = (uint32_t)q.W; d.H = (uint32_t)q.H; d.spp = (uint32_t)q.spp; d.depth = (uint32_t)q.depth;
        d.K = (uint32_t)q.K; d.seed = q.seed; d.sample_offset = (uint32_t)q.sample_offset;
        d.mode = q.mode;
        d.inv_spp = 1.f / (float)q.spp;
        d.path_L = (float4*)q.path_L;
        d.d_img = (const float4*)q.d_img;
        d.gt = adjoint ? ctx->d_gt + gt_base[r] : nullptr;
        base += (uint64_t)q.W * q.H * q.K;
    }
    if (ctx->req_cap < table.size()) {
        dfree(ctx->d_reqs);
        dfree(ctx->d_qbase);
        ctx->req_cap = 0;
        HIP_TRY(dalloc(&ctx->d_reqs, table.size()));
        HIP_TRY(dalloc(&ctx->d_qbase, table.size() + 1));
        ctx->req_cap = table.size();
    }
    HIP_TRY(hipStreamSynchronize(hs));
    HIP_TRY(hipMemcpy(ctx->d_reqs, table.data(), table.size() * sizeof(ff_dev_request), hipMemcpyHostToDevice));
    if (adjoint && !gt.empty())
        HIP_TRY(hipMemcpy(ctx->d_gt, gt.data(), gt.size() * sizeof(ff_dev_gtarget), hipMemcpyHostToDevice));
    p.counters = ctx->counters;
    p.fetch = ctx->fetch;
    p.qbase = ctx->d_qbase;
    p.stats = ctx->stats;
    p.ovf = ctx->ovf;
    p.ovf_shd = ctx->ovf + ctx->ovf_entries * (size_t)FF_TRAV_THREADS;
    p.bstate = ctx->bstate;
    p.reqs = ctx->d_reqs;
    p.n_req = (uint32_t)table.size();
    p.n_slots = (uint32_t)n_slots;
    p.uni_scene = 1;
    p.uni_env = 1;
    for (const HostReq& q : reqs) {
        if (q.s != reqs[0].s) p.uni_scene = 0;
        if (q.e != reqs[0].e) p.uni_env = 0;
    }
    p.g_max = ctx->g_max;
    p.g_cnt = ctx->g_cnt;
    p.g_fx = ctx->g_fx;
    p.gmode = GM_FLOAT;
    // the context's variant, and from trav_tail_from on the tail variant
    auto var_at = [&](int32_t it) { return ctx->trav_tail >= 0 && it >= ctx->trav_tail_from ? ctx->trav_tail : ctx->trav_variant; };

    ctx->records.clear();
    ctx->events_used = 0;
    // Records a begin/end event pair around one stage launch when timing.
    auto timed_on = [&](hipStream_t ts, int stage, auto&& launch) -> hipError_t {
        if (!timing) return launch();
        while (ctx->event_pool.size() < ctx->events_used + 2) {
            hipEvent_t e;
            hipError_t ce = hipEventCreate(&e);
            if (ce != hipSuccess) return ce;
            ctx->event_pool.push_back(e);
        }
        StageRecord rec{stage, ctx->event_pool[ctx->events_used], ctx->event_pool[ctx->events_used + 1]};
        ctx->events_used += 2;
        hipError_t e = hipEventRecord(rec.begin, ts);
        if (e == hipSuccess) e = launch();
        if (e == hipSuccess) e = hipEventRecord(rec.end, ts);
        ctx->records.push_back(rec);
        return e;
    };
    auto timed = [&](int stage, auto&& launch) -> hipError_t { return timed_on(hs, stage, launch); };
    // the first batch stores every layer of a request whose spp covers its
    // samples in flight (wf_logic writes it without reading); only a request
    // with fewer samples than layers keeps layers that no path writes
    if (!adjoint)
        for (size_t r = 0; r < reqs.size(); r++)
            if (reqs[r].K > 1 && reqs[r].spp < reqs[r].K)
                HIP_TRY(ffk::launch_fill4(table[r].layers, (uint32_t)((uint64_t)reqs[r].W * reqs[r].H * reqs[r].K), hs));
    HIP_TRY(ffk::launch_fill_u64(p.stats, S_WORDS, hs));

    // one batch of the forward: identical launches for every batch (the index is on the device)
    auto batch = [&]() -> ff_status {
        HIP_TRY(ffk::launch_batch_begin(p, 0, 0, hs));
        HIP_TRY(timed(4, [&] { return ffk::launch_generate(p, hs); }));
        int cur = 0;
        for (int32_t it = 1; it <= max_depth + 1; it++) {
            // the iteration's extension rays and its vertices' shadow rays are
            // independent: the shadow traversal runs on the side stream, so
            // either kernel's waves fill the other's tail (with per-launch
            // timing the two run one after another, so each time is its own)
            hipStream_t ss = timing ? hs : ctx->side;
            if (!timing) {
                HIP_TRY(hipEventRecord(ctx->fork_ev, hs));
                HIP_TRY(hipStreamWaitEvent(ss, ctx->fork_ev, 0));
            }
            HIP_TRY(timed(0, [&] { return ffk::launch_extend(p, cur, var_at(it), hs); }));
            HIP_TRY(timed_on(ss, 1, [&] { return ffk::launch_shadow(p, var_at(it), ss); }));
            if (!timing) {
                HIP_TRY(hipEventRecord(ctx->join_ev, ss));
                HIP_TRY(hipStreamWaitEvent(hs, ctx->join_ev, 0));
            }
            uint32_t mask = (1u << (cur == 0 ? C_EXT1 : C_EXT0)) | (1u << C_SHADOW) | (0xfu << C_MAT0);
            HIP_TRY(ffk::launch_reset(p.counters, mask, p.fetch, (1u << F_EXT) | (1u << F_SHD), hs));
            HIP_TRY(timed(2, [&] { return ffk::launch_logic(p, cur, it, p.n_slots, hs); }));
            if (it <= max_depth) {
                for (int m = 0; m < 4; m++)
                    if (has_type[m]) HIP_TRY(timed(3, [&] { return ffk::launch_shade(m, p, 1 - cur, it == 1, p.n_slots, hs); }));
            }
            cur = 1 - cur;
        }
        return FF_OK;
    };
    // one batch of the adjoint replay: a vertex's shadow rays are traced
    // between its NEE replay and its gradient kernel
    auto batch_adj = [&]() -> ff_status {
        HIP_TRY(ffk::launch_batch_begin(p, 0, 0, hs));
        HIP_TRY(timed(4, [&] { return ffk::launch_generate(p, hs); }));
        int cur = 0;
        for (int32_t it = 1; it <= max_depth + 1; it++) {
            HIP_TRY(timed(0, [&] { return ffk::launch_extend(p, cur, var_at(it), hs); }));
            uint32_t mask = (1u << (cur == 0 ? C_EXT1 : C_EXT0)) | (1u << C_SHADOW) | (0xfu << C_MAT0);
            HIP_TRY(ffk::launch_reset(p.counters, mask, p.fetch, (1u << F_EXT) | (1u << F_SHD), hs));
            HIP_TRY(timed(2, [&] { return ffk::launch_logic_adj(p, cur, p.n_slots, hs); }));
            if (it <= max_depth) {
                for (int m = 0; m < 3; m++)
                    if (has_type[m]) HIP_TRY(timed(3, [&] { return ffk::launch_shade_nee(m, p, p.n_slots, hs); }));
                HIP_TRY(timed(1, [&] { return ffk::launch_shadow(p, var_at(it), hs); }));
                for (int m = 0; m < 4; m++)
                    if (has_type[m])
                        HIP_TRY(timed(3, [&] { return ffk::launch_shade_adj(m, p, 1 - cur, p.n_slots, hs); }));
            }
            cur = 1 - cur;
        }
        return FF_OK;
    };
    auto run = [&]() -> ff_status {
        HIP_TRY(ffk::launch_batch_begin(p, 1, 0xffffffffu, hs));  // the first batch_begin makes it 0
        if (graph) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            HIP_TRY(hipStreamBeginCapture(hs, hipStreamCaptureModeThreadLocal));
            ff_status bst = adjoint ? batch_adj() : batch();
            hipError_t ce = hipStreamEndCapture(hs, &g);
            if (bst != FF_OK) {
                if (g) (void)hipGraphDestroy(g);
                return bst;
            }
            HIP_TRY(ce);
            hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(g);
                HIP_TRY(ie);
            }
            hipError_t le = hipSuccess;
            for (int64_t b = 0; b < n_batches && le == hipSuccess; b++) le = hipGraphLaunch(ge, hs);
            (void)hipGraphExecDestroy(ge);
            (void)hipGraphDestroy(g);
            HIP_TRY(le);
        } else {
            for (int64_t b = 0; b < n_batches; b++) {
                ff_status bst = adjoint ? batch_adj() : batch();
                if (bst != FF_OK) return bst;
            }
        }
        return FF_OK;
    };
    if (!deterministic) {
        ff_status st = run();
        if (st != FF_OK) return st;
    } else {
        // pass 1 bounds every gradient element's terms; pass 2 sums them as
        // int64 fixed point at per-element scales (order-independent sums)
        HIP_TRY(ffk::launch_fill_u32(p.g_max, arena, hs));
        HIP_TRY(ffk::launch_fill_u32(p.g_cnt, arena, hs));
        HIP_TRY(ffk::launch_fill_u32((uint32_t*)p.g_fx, 2 * arena, hs));
        p.gmode = GM_BOUND;
        ff_status st = run();
        if (st != FF_OK) return st;
        p.gmode = GM_FIXED;
        st = run();
        if (st != FF_OK) return st;
        for (const Target& t : targets)
            HIP_TRY(ffk::launch_fixed_to_float(p.g_fx + t.off, p.g_max + t.off, p.g_cnt + t.off, t.n, t.buf, hs));
    }
    if (!adjoint)
        for (size_t r = 0; r < reqs.size(); r++)
            if (reqs[r].K > 1)
                HIP_TRY(timed(4, [&] {
                    return ffk::launch_resolve_layers(table[r].layers, (uint32_t)reqs[r].K,
                                                      (uint32_t)((uint64_t)reqs[r].W * reqs[r].H), (float4*)reqs[r].film, hs);
                }));
    ctx->last_stream = hs;
    ctx->have_stats = true;
    return FF_OK;
}

}  // namespace

extern "C" {

ff_status ff_render(ff_context* ctx, ff_scene* s, ff_envmap* env, const ff_camera* cam, const ff_render_req* r,
                    float* film, void* stream) {
    if (!ctx || !s || !env || !cam || !r || !film) return fail(FF_ERR_INVALID, "ff_render: NULL argument");
    HostReq q{s, env, *cam, r->width, r->height, r->spp, r->max_depth,
              r->spp_batch > 0 ? std::min(r->spp_batch, std::max(r->spp, 1)) : 1, r->sample_offset, r->seed, film};
    ff_status st = check_req("ff_render", ctx, q);
    if (st != FF_OK) return st;
    return render_requests(ctx, {q}, r->flags, (hipStream_t)stream, "ff_render");
}

ff_status ff_render_batch(ff_context* ctx, const ff_batch_item* items, int32_t n, int32_t flags, void* stream) {
    if (!ctx || !items || n <= 0) return fail(FF_ERR_INVALID, "ff_render_batch: NULL argument or empty batch");
    std::vector<HostReq> reqs;
    reqs.reserve((size_t)n);
    for (int32_t i = 0; i < n; i++) {
        const ff_batch_item& it = items[i];
        HostReq q{it.scene, it.env, it.camera, it.width, it.height, it.spp, it.max_depth,
                  it.spp_batch > 0 ? std::min(it.spp_batch, std::max(it.spp, 1)) : 1, it.sample_offset, it.seed,
                  it.film};
        if (it.path_radiance) {
            q.mode = RQ_STORE_PATHS;
            q.path_L = it.path_radiance;
        }
        ff_status st = check_req("ff_render_batch", ctx, q);
        if (st != FF_OK) return st;
        for (int32_t j = 0; j < i; j++)
            if (items[j].film == it.film) return fail(FF_ERR_INVALID, "ff_render_batch: items %d and %d share a film", j, i);
        reqs.push_back(q);
    }
    return render_requests(ctx, reqs, flags, (hipStream_t)stream, "ff_render_batch");
}

ff_status ff_render_paths(ff_context* ctx, ff_scene* s, ff_envmap* env, const ff_camera* cam,
                          const ff_render_req* r, float* film, float* path_radiance, void* stream) {
    if (!ctx || !s || !env || !cam || !r || !film || !path_radiance)
        return fail(FF_ERR_INVALID, "ff_render_paths: NULL argument");
    HostReq q{s, env, *cam, r->width, r->height, r->spp, r->max_depth,
              r->spp_batch > 0 ? std::min(r->spp_batch, std::max(r->spp, 1)) : 1, r->sample_offset, r->seed, film};
    q.mode = RQ_STORE_PATHS;
    q.path_L = path_radiance;
    ff_status st = check_req("ff_render_paths", ctx, q);
    if (st != FF_OK) return st;
    return render_requests(ctx, {q}, r->flags, (hipStream_t)stream, "ff_render_paths");
}

ff_status ff_adjoint(ff_context* ctx, ff_scene* s, ff_envmap* env, const ff_camera* cam, const ff_render_req* r,
                     const float* path_radiance, const float* d_image, const ff_grads* grads, void* stream) {
    if (!ctx || !s || !env || !cam || !r || !path_radiance || !d_image || !grads)
        return fail(FF_ERR_INVALID, "ff_adjoint: NULL argument");
    HostReq q{s, env, *cam, r->width, r->height, r->spp, r->max_depth,
              r->spp_batch > 0 ? std::min(r->spp_batch, std::max(r->spp, 1)) : 1, r->sample_offset, r->seed,
              nullptr};
    q.mode = RQ_ADJOINT;
    q.path_L = (float*)path_radiance;
    q.d_img = d_image;
    q.grads = grads;
    ff_status st = check_req("ff_adjoint", ctx, q);
    if (st != FF_OK) return st;
    return render_requests(ctx, {q}, r->flags, (hipStream_t)stream, "ff_adjoint");
}

ff_status ff_adjoint_batch(ff_context* ctx, const ff_batch_item* items, int32_t n, int32_t flags, void* stream) {
    if (!ctx || !items || n <= 0) return fail(FF_ERR_INVALID, "ff_adjoint_batch: NULL argument or empty batch");
    std::vector<HostReq> reqs;
    reqs.reserve((size_t)n);
    for (int32_t i = 0; i < n; i++) {
        const ff_batch_item& it = items[i];
        HostReq q{it.scene, it.env, it.camera, it.width, it.height, it.spp, it.max_depth,
                  it.spp_batch > 0 ? std::min(it.spp_batch, std::max(it.spp, 1)) : 1, it.sample_offset, it.seed,
                  nullptr};
        q.mode = RQ_ADJOINT;
        q.path_L = it.path_radiance;
        q.d_img = it.d_image;
        q.grads = it.grads;
        ff_status st = check_req("ff_adjoint_batch", ctx, q);
        if (st != FF_OK) return st;
        reqs.push_back(q);
    }
    return render_requests(ctx, reqs, flags, (hipStream_t)stream, "ff_adjoint_batch");
}

ff_status ff_set_traversal_variant(ff_context* ctx, int32_t variant) {
    if (!ctx || variant < 0 || variant >= ffk::kTravVariants)
        return fail(FF_ERR_INVALID, "ff_set_traversal_variant: variant %d of %d", variant, ffk::kTravVariants);
    ctx->trav_variant = variant;
    return FF_OK;
}

ff_status ff_set_traversal_tail(ff_context* ctx, int32_t variant, int32_t from_iteration) {
    if (!ctx || variant < -1 || variant >= ffk::kTravVariants || from_iteration < 1)
        return fail(FF_ERR_INVALID, "ff_set_traversal_tail: variant %d (of %d), from iteration %d", variant,
                    ffk::kTravVariants, from_iteration);
    ctx->trav_tail = variant;
    ctx->trav_tail_from = from_iteration;
    return FF_OK;
}

ff_status ff_traversal_variant_info(int32_t variant, int32_t* block, int32_t* lds_entries, int32_t* refill,
                                    int32_t* while_while, int32_t* bvh_width) {
    if (variant < 0 || variant >= ffk::kTravVariants)
        return fail(FF_ERR_INVALID, "ff_traversal_variant_info: variant %d of %d", variant, ffk::kTravVariants);
    if (block) *block = ffk::kTravBlock[variant];
    if (lds_entries) *lds_entries = ffk::kTravLds[variant];
    if (refill) *refill = ffk::kTravRefill[variant];
    if (while_while) *while_while = ffk::kTravWhileWhile[variant];
    if (bvh_width) *bvh_width = ffk::kTravWidth[variant];
    return FF_OK;
}

int32_t ff_traversal_variant_count(void) { return ffk::kTravVariants; }

ff_status ff_last_render_stats(ff_context* ctx, ff_ray_stats* out) {
    if (!ctx || !out) return fail(FF_ERR_INVALID, "ff_last_render_stats: NULL argument");
    if (!ctx->have_stats) return fail(FF_ERR_INVALID, "ff_last_render_stats: no render on this context");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->last_stream));
    std::vector<unsigned long long> rep(S_WORDS);
    HIP_TRY(hipMemcpy(rep.data(), ctx->stats, sizeof(unsigned long long) * S_WORDS, hipMemcpyDeviceToHost));
    unsigned long long h[S_COUNT] = {0};
    for (int r = 0; r < S_REPLICAS; r++)
        for (int k = 0; k < S_COUNT; k++) h[k] += rep[(size_t)r * S_STRIDE + k];
    std::memset(out, 0, sizeof(*out));
    out->rays_primary = (double)h[S_PRIMARY];
    out->rays_extension = (double)h[S_EXTENSION];
    out->rays_shadow = (double)h[S_SHADOW];
    out->paths = (double)h[S_PATHS];
    out->ext_nodes = (double)h[S_EXT_NODES];
    out->ext_tris = (double)h[S_EXT_TRIS];
    out->shadow_nodes = (double)h[S_SHD_NODES];
    out->shadow_tris = (double)h[S_SHD_TRIS];
    out->rays_shadow_gated = (double)h[S_SHD_GATED];
    double* ms[5] = {&out->ms_extend, &out->ms_shadow, &out->ms_logic, &out->ms_shade, &out->ms_other};
    for (const StageRecord& rec : ctx->records) {
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, rec.begin, rec.end));
        *ms[rec.stage] += (double)t;
        if (rec.stage == 0) out->launches_extend += 1;
        if (rec.stage == 1) out->launches_shadow += 1;
    }
    return FF_OK;
}

ff_status ff_film_resolve(const float* film, int32_t width, int32_t height, float exposure, float* avg_rgb,
                          uint8_t* rgba8, void* stream) {
    if (!film || width <= 0 || height <= 0) return fail(FF_ERR_INVALID, "ff_film_resolve: bad arguments");
    HIP_TRY(ffk::launch_film_resolve((const float4*)film, (uint32_t)((uint64_t)width * height), exposure, avg_rgb,
                                     rgba8, (hipStream_t)stream));
    return FF_OK;
}

}  // extern "C"
