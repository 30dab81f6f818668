// ff_lbvh.hip -- linear BVH built on the GPU (opt-in scene builder).
//
// The reference declares SplitMethod::HLBVH but builds only the host SAH tree
// (CRT/BVH.h:23, CRT/BVH.cpp:81-307).  This builder makes a tree of the same
// node format (flattened depth first, first child at +1, second child offset,
// leaves over a contiguous range of the slot -> triangle permutation), so
// everything after the build -- record packing, primitive reordering,
// traversal -- is shared with the SAH path:
//   1. centroid bounds (float min/max, exact, order independent);
//   2. 30-bit Morton code of each triangle's box centre in those bounds,
//      key = code << 32 | triangle index (unique keys);
//   3. radix sort of the keys (rocPRIM);
//   4. Karras' hierarchy (Maximizing parallelism in the construction of
//      BVHs, octrees and k-d trees, HPG 2012): internal node i covers a
//      contiguous key range found from common-prefix lengths, split where the
//      prefix grows;
//   5. subtrees of at most max_prims triangles become leaves (their slots are
//      contiguous in key order);
//   6. bottom-up boxes and subtree sizes, the second child to arrive at a
//      node finishing it;
//   7. each output node finds its depth-first index by walking to the root.
// Every step is deterministic; the oracle restates it sequentially
// (or_lbvh_build) and the two trees are compared node for node.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <stdint.h>

#include <vector>

#include "ff_internal.h"

namespace ffk {
namespace {

constexpr uint32_t LB_BLOCK = 256;

// order-preserving map of a float to an unsigned key (for atomic min/max)
__device__ __forceinline__ uint32_t f2key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// min / max by one comparison (the first operand on ties, so signed zeros
// come out as the oracle's)
__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return b > a ? b : a; }

// Box of triangle i; its centre is formed as the reference forms
// BVHPrimitiveInfo's centroid (0.5 * bmin + 0.5 * bmax, CRT/BVH.h:36-40).
__device__ __forceinline__ void tri_box(const float* __restrict__ pos, uint32_t i, float lo[3], float hi[3]) {
    const float* p = pos + (size_t)i * 9;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float x = p[a], y = p[3 + a], z = p[6 + a];
        lo[a] = mn(mn(x, y), z);
        hi[a] = mx(mx(x, y), z);
    }
}

__global__ __launch_bounds__(LB_BLOCK) void k_centroid_bounds(const float* __restrict__ pos, uint32_t n,
                                                             uint32_t* cb /* [6]: min xyz, max xyz keys */) {
    __shared__ uint32_t s[6][LB_BLOCK / 64];
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * LB_BLOCK + threadIdx.x; i < n; i += gridDim.x * LB_BLOCK) {
        float lo[3], hi[3];
        tri_box(pos, i, lo, hi);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            uint32_t k = f2key(0.5f * lo[a] + 0.5f * hi[a]);
            mn[a] = min(mn[a], k);
            mx[a] = max(mx[a], k);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
        for (int off = 32; off > 0; off >>= 1) {
            mn[a] = min(mn[a], (uint32_t)__shfl_xor((int)mn[a], off));
            mx[a] = max(mx[a], (uint32_t)__shfl_xor((int)mx[a], off));
        }
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; a++) {
            s[a][threadIdx.x >> 6] = mn[a];
            s[3 + a][threadIdx.x >> 6] = mx[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        uint32_t v = s[threadIdx.x][0];
        for (uint32_t w = 1; w < LB_BLOCK / 64; w++)
            v = threadIdx.x < 3 ? min(v, s[threadIdx.x][w]) : max(v, s[threadIdx.x][w]);
        if (threadIdx.x < 3) atomicMin(&cb[threadIdx.x], v);
        else atomicMax(&cb[threadIdx.x], v);
    }
}

__global__ void k_bounds_init(uint32_t* cb) {
    if (threadIdx.x < 6) cb[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
}

// 10 bits -> every third bit of 30
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
// quantised coordinate of c in [lo, lo + ext]: floor(1024 * (c - lo) / ext), clamped to [0, 1023]
__device__ __forceinline__ uint32_t quant10(float c, float lo, float ext) {
    float u = ext > 0.f ? (c - lo) / ext : 0.f;
    float q = u * 1024.f;
    if (!(q > 0.f)) q = 0.f;
    if (q > 1023.f) q = 1023.f;
    return (uint32_t)q;
}

__global__ __launch_bounds__(LB_BLOCK) void k_morton(const float* __restrict__ pos, uint32_t n,
                                                    const uint32_t* __restrict__ cb, uint64_t* __restrict__ keys) {
    float lo[3], ext[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lo[a] = key2f(cb[a]);
        ext[a] = key2f(cb[3 + a]) - lo[a];
    }
    for (uint32_t i = blockIdx.x * LB_BLOCK + threadIdx.x; i < n; i += gridDim.x * LB_BLOCK) {
        float bl[3], bh[3];
        tri_box(pos, i, bl, bh);
        uint32_t q[3];
#pragma unroll
        for (int a = 0; a < 3; a++) q[a] = quant10(0.5f * bl[a] + 0.5f * bh[a], lo[a], ext[a]);
        uint32_t code = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
        keys[i] = ((uint64_t)code << 32) | i;
    }
}

// common prefix length of keys i and j; -1 outside [0, n)
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll((long long)(k[i] ^ k[j]));
}

// Internal node i of Karras' hierarchy: its key range, split and children
// (c >= 0 internal node, c < 0 primitive ~c), parent links and split axis.
__global__ __launch_bounds__(LB_BLOCK) void k_karras(const uint64_t* __restrict__ k, int n, int2* __restrict__ child,
                                                    int2* __restrict__ range, int* __restrict__ par_int,
                                                    int* __restrict__ par_leaf, uint8_t* __restrict__ axis) {
    int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (i >= n - 1) return;
    int d = delta(k, n, i, i + 1) > delta(k, n, i, i - 1) ? 1 : -1;
    int dmin = delta(k, n, i, i - d);
    int lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int lo = 0, hi = lmax;  // largest l in [0, lmax) with delta(i, i + l d) > dmin
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (delta(k, n, i, i + mid * d) > dmin) lo = mid;
        else hi = mid;
    }
    int l = lo, j = i + l * d;
    int dnode = delta(k, n, i, j);
    lo = 0;
    hi = l;  // largest s in [0, l) with delta(i, i + s d) > dnode
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (delta(k, n, i, i + mid * d) > dnode) lo = mid;
        else hi = mid;
    }
    int g = i + lo * d + (d < 0 ? -1 : 0);
    int first = min(i, j), last = max(i, j);
    int left = first == g ? ~g : g, right = last == g + 1 ? ~(g + 1) : g + 1;
    child[i] = make_int2(left, right);
    range[i] = make_int2(first, last);
    if (left >= 0) par_int[left] = i; else par_leaf[~left] = i;
    if (right >= 0) par_int[right] = i; else par_leaf[~right] = i;
    // the split bit is the highest bit where the range's first and last keys differ
    int p = 63 - dnode;
    axis[i] = p >= 32 ? (uint8_t)(2 - (p - 32) % 3) : (uint8_t)0;
}

struct Tree {
    const uint64_t* keys;
    const float* pos;
    const int2* child;
    const int2* range;
    const int* par_int;
    const int* par_leaf;
    const uint8_t* axis;
    float* ibox;    // [n-1][6] box of an internal node that is an output node
    float* lbox;    // [n][6] box of a primitive that is an output leaf
    int* isize;     // [n-1] output nodes in the subtree (1 for a collapsed leaf)
    uint32_t* flag; // [n-1] arrivals
    int n, max_prims;
};

__device__ __forceinline__ int rsize(const Tree& T, int v) { return T.range[v].y - T.range[v].x + 1; }
__device__ __forceinline__ void box_union(float* o, const float* a, const float* b) {
    for (int c = 0; c < 3; c++) {
        o[c] = mn(a[c], b[c]);
        o[3 + c] = mx(a[3 + c], b[3 + c]);
    }
}
__device__ __forceinline__ void range_box(const Tree& T, int a, int b, float* o) {
    float lo[3], hi[3];
    tri_box(T.pos, (uint32_t)(T.keys[a] & 0xffffffffu), lo, hi);
    for (int s = a + 1; s <= b; s++) {
        float l2[3], h2[3];
        tri_box(T.pos, (uint32_t)(T.keys[s] & 0xffffffffu), l2, h2);
        for (int c = 0; c < 3; c++) {
            lo[c] = mn(lo[c], l2[c]);
            hi[c] = mx(hi[c], h2[c]);
        }
    }
    for (int c = 0; c < 3; c++) {
        o[c] = lo[c];
        o[3 + c] = hi[c];
    }
}

// One thread per primitive: the primitive's output leaf (itself, or the
// highest ancestor holding at most max_prims primitives) is finished by the
// leaf's first primitive, which then climbs the kept internal nodes; at each
// the second arrival forms the box and size.
__global__ __launch_bounds__(LB_BLOCK) void k_refit(Tree T) {
    int p = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (p >= T.n) return;
    int u = ~p;  // output leaf candidate (primitive)
    int par = T.n > 1 ? T.par_leaf[p] : -1;
    while (par >= 0 && rsize(T, par) <= T.max_prims) {
        u = par;
        par = T.par_int[par];
    }
    if (u >= 0) {
        if (T.range[u].x != p) return;  // the range's first primitive builds the leaf
        range_box(T, T.range[u].x, T.range[u].y, T.ibox + (size_t)u * 6);
        T.isize[u] = 1;
    } else {
        range_box(T, p, p, T.lbox + (size_t)p * 6);
    }
    while (par >= 0) {
        __threadfence();
        if (atomicAdd(&T.flag[par], 1u) == 0u) return;  // the sibling finishes the node
        __threadfence();
        int2 c = T.child[par];
        const float* a = c.x >= 0 ? T.ibox + (size_t)c.x * 6 : T.lbox + (size_t)(~c.x) * 6;
        const float* b = c.y >= 0 ? T.ibox + (size_t)c.y * 6 : T.lbox + (size_t)(~c.y) * 6;
        float o[6];
        box_union(o, a, b);  // the sibling's writes are visible after the fences
        for (int q = 0; q < 6; q++) T.ibox[(size_t)par * 6 + q] = o[q];
        int sa = c.x >= 0 ? T.isize[c.x] : 1;
        int sb = c.y >= 0 ? T.isize[c.y] : 1;
        T.isize[par] = 1 + sa + sb;
        par = T.par_int[par];
    }
}

__device__ __forceinline__ bool kept(const Tree& T, int v) { return rsize(T, v) > T.max_prims; }

// One thread per internal node and per primitive: an output node finds its
// depth-first index (first child at +1, second after the first's subtree)
// and writes its record.
__global__ __launch_bounds__(LB_BLOCK) void k_emit(Tree T, ffh::BvhNode* __restrict__ out,
                                                  int32_t* __restrict__ ordered) {
    int t = blockIdx.x * LB_BLOCK + threadIdx.x;
    int n = T.n;
    if (t < n) ordered[t] = (int32_t)(T.keys[t] & 0xffffffffu);
    if (t >= 2 * n - 1) return;
    bool prim = t >= n - 1;
    int v = prim ? ~(t - (n - 1)) : t;  // node id: internal >= 0, primitive < 0
    int par = prim ? (n > 1 ? T.par_leaf[~v] : -1) : T.par_int[v];
    bool is_out = prim ? (par < 0 || kept(T, par)) : (kept(T, v) || par < 0 || kept(T, par));
    if (!is_out) return;
    int idx = 0;
    for (int cur = v, pp = par; pp >= 0;) {
        int2 c = T.child[pp];
        if (c.y == cur) idx += 1 + (c.x >= 0 ? T.isize[c.x] : 1);
        else idx += 1;
        cur = pp;
        pp = T.par_int[pp];
    }
    ffh::BvhNode r;
    const float* b = prim ? T.lbox + (size_t)(~v) * 6 : T.ibox + (size_t)v * 6;
    for (int c = 0; c < 3; c++) {
        r.bmin[c] = b[c];
        r.bmax[c] = b[3 + c];
    }
    r.pad = 0;
    if (!prim && kept(T, v)) {
        int2 c = T.child[v];
        r.offset = idx + 1 + (c.x >= 0 ? T.isize[c.x] : 1);
        r.n_prims = 0;
        r.axis = T.axis[v];
    } else {
        int a = prim ? ~v : T.range[v].x, z = prim ? ~v : T.range[v].y;
        r.offset = a;
        r.n_prims = (uint16_t)(z - a + 1);
        r.axis = 0;
    }
    out[idx] = r;
}

template <class X>
hipError_t alloc(X** p, size_t n) {
    return hipMalloc((void**)p, sizeof(X) * (n ? n : 1));
}

}  // namespace

// positions on the host [n][3][3]; nodes and ordered come back on the host;
// gpu_ms (optional): device time from the bounds pass to the emitted tree
hipError_t build_lbvh(const float* positions, int32_t n, int32_t max_prims, std::vector<ffh::BvhNode>& nodes,
                      std::vector<int32_t>& ordered, float* gpu_ms) {
    if (n <= 0) return hipErrorInvalidValue;
    if (max_prims < 1) max_prims = 1;
    if (max_prims > 65535) max_prims = 65535;  // a leaf's count is 16 bits
    hipStream_t s = nullptr;
    float *d_pos = nullptr, *d_ibox = nullptr, *d_lbox = nullptr;
    uint64_t *d_keys = nullptr, *d_sorted = nullptr;
    uint32_t *d_cb = nullptr, *d_flag = nullptr;
    int2 *d_child = nullptr, *d_range = nullptr;
    int *d_par_int = nullptr, *d_par_leaf = nullptr, *d_isize = nullptr, *d_ord = nullptr;
    uint8_t* d_axis = nullptr;
    ffh::BvhNode* d_nodes = nullptr;
    void* d_tmp = nullptr;
    size_t m = (size_t)n, tmp_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess && gpu_ms) e = hipEventCreate(&ev0);
    if (e == hipSuccess && gpu_ms) e = hipEventCreate(&ev1);
    auto ok = [&](hipError_t x) { if (e == hipSuccess) e = x; return e == hipSuccess; };
    if (ok(alloc(&d_pos, 9 * m)) && ok(alloc(&d_keys, m)) && ok(alloc(&d_sorted, m)) && ok(alloc(&d_cb, 6)) &&
        ok(alloc(&d_child, m)) && ok(alloc(&d_range, m)) && ok(alloc(&d_par_int, m)) && ok(alloc(&d_par_leaf, m)) &&
        ok(alloc(&d_axis, m)) && ok(alloc(&d_ibox, 6 * m)) && ok(alloc(&d_lbox, 6 * m)) && ok(alloc(&d_isize, m)) &&
        ok(alloc(&d_flag, m)) && ok(alloc(&d_nodes, 2 * m)) && ok(alloc(&d_ord, m)) &&
        ok(hipMemcpyAsync(d_pos, positions, sizeof(float) * 9 * m, hipMemcpyHostToDevice, s)) &&
        ok(hipMemsetAsync(d_flag, 0, sizeof(uint32_t) * m, s)) &&
        ok(hipMemsetAsync(d_par_int, 0xff, sizeof(int) * m, s)) &&  // the root's parent: -1
        ok(rocprim::radix_sort_keys(nullptr, tmp_bytes, d_keys, d_sorted, m, 0, 62, s)) &&
        ok(hipMalloc(&d_tmp, tmp_bytes ? tmp_bytes : 1))) {
        uint32_t g = (uint32_t)((m + LB_BLOCK - 1) / LB_BLOCK);
        uint32_t gb = g < 4096 ? g : 4096;
        if (ev0) ok(hipEventRecord(ev0, s));
        hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, d_cb);
        hipLaunchKernelGGL(k_centroid_bounds, dim3(gb), dim3(LB_BLOCK), 0, s, d_pos, (uint32_t)n, d_cb);
        hipLaunchKernelGGL(k_morton, dim3(gb), dim3(LB_BLOCK), 0, s, d_pos, (uint32_t)n, d_cb, d_keys);
        if (ok(hipGetLastError()) && ok(rocprim::radix_sort_keys(d_tmp, tmp_bytes, d_keys, d_sorted, m, 0, 62, s))) {
            if (n > 1)
                hipLaunchKernelGGL(k_karras, dim3((uint32_t)((m - 1 + LB_BLOCK - 1) / LB_BLOCK)), dim3(LB_BLOCK), 0, s,
                                   d_sorted, n, d_child, d_range, d_par_int, d_par_leaf, d_axis);
            Tree T{d_sorted, d_pos, d_child, d_range, d_par_int, d_par_leaf, d_axis,
                   d_ibox, d_lbox, d_isize, d_flag, n, max_prims};
            hipLaunchKernelGGL(k_refit, dim3(g), dim3(LB_BLOCK), 0, s, T);
            hipLaunchKernelGGL(k_emit, dim3((uint32_t)((2 * m - 1 + LB_BLOCK - 1) / LB_BLOCK)), dim3(LB_BLOCK), 0, s,
                               T, d_nodes, d_ord);
            if (ev1) ok(hipEventRecord(ev1, s));
            int32_t n_out = 1;
            if (ok(hipGetLastError()) && n > 1 && (int32_t)(m) > max_prims)
                ok(hipMemcpyAsync(&n_out, d_isize, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            if (ok(hipStreamSynchronize(s))) {
                if (ev0) ok(hipEventElapsedTime(gpu_ms, ev0, ev1));
                nodes.resize((size_t)n_out);
                ordered.resize(m);
                if (ok(hipMemcpy(nodes.data(), d_nodes, sizeof(ffh::BvhNode) * (size_t)n_out, hipMemcpyDeviceToHost)))
                    ok(hipMemcpy(ordered.data(), d_ord, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
            }
        }
    }
    for (void* p : {(void*)d_pos, (void*)d_keys, (void*)d_sorted, (void*)d_cb, (void*)d_child, (void*)d_range,
                    (void*)d_par_int, (void*)d_par_leaf, (void*)d_axis, (void*)d_ibox, (void*)d_lbox, (void*)d_isize,
                    (void*)d_flag, (void*)d_nodes, (void*)d_ord, d_tmp})
        if (p) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (s) (void)hipStreamDestroy(s);
    return e;
}

}  // namespace ffk
