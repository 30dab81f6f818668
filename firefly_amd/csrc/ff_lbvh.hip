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
#include <rocprim/device/device_scan.hpp>

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

// ------------------------------------------------------------------ PLOC
// Parallel locally-ordered clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 2018) over the Morton-sorted triangles: every round, each
// cluster finds its nearest neighbour among the PLOC_R clusters on either
// side in the current order (the smallest surface area of the two boxes'
// union; ties to the lower position), mutual neighbours merge into a new node
// at the lower position, and the array is compacted in order.  Node ids:
// triangles 0..n-1 in sorted order, then internal nodes in creation order
// (a round's merges by position), so every child precedes its parent and a
// node's cost, leaf decision and child order are fixed when it is made.
constexpr int PLOC_R = 16;

struct PNodes {
    float* box;       // [2n-1][6] lo xyz, hi xyz
    int* first;       // [2n-1] children in visit order (-1: triangle)
    int* second;
    int* parent;      // [2n-1] (-1: root)
    int* count;       // [2n-1] triangles in the subtree
    float* cost;      // [2n-1] SAH cost of the subtree (half areas)
    int* osize;       // [2n-1] output nodes of the subtree (1 for a leaf)
    uint8_t* axis;    // [2n-1] split axis of an interior node
    uint8_t* leaf;    // [2n-1] 1: an output leaf (a triangle or a collapsed subtree)
    int max_prims;
};

// half the surface area of a box (the SAH's area, up to the factor 2)
__device__ __forceinline__ float half_area(const float* b) {
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return (dx * dy + dy * dz) + dz * dx;
}

__global__ __launch_bounds__(LB_BLOCK) void k_ploc_leaves(const uint64_t* __restrict__ keys,
                                                         const float* __restrict__ pos, int n, PNodes N,
                                                         int* __restrict__ C) {
    int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (i >= n) return;
    float lo[3], hi[3];
    tri_box(pos, (uint32_t)(keys[i] & 0xffffffffu), lo, hi);
    float* b = N.box + (size_t)i * 6;
    for (int c = 0; c < 3; c++) {
        b[c] = lo[c];
        b[3 + c] = hi[c];
    }
    N.first[i] = -1;
    N.second[i] = -1;
    N.parent[i] = -1;
    N.count[i] = 1;
    N.cost[i] = half_area(b);
    N.osize[i] = 1;
    N.axis[i] = 0;
    N.leaf[i] = 1;
    C[i] = i;
}

// nearest neighbour of every cluster within PLOC_R positions; the block's
// clusters and the windows' margins are staged in LDS
__global__ __launch_bounds__(LB_BLOCK) void k_ploc_nn(const int* __restrict__ C, int m, const float* __restrict__ box,
                                                     int* __restrict__ nn) {
    __shared__ float sb[(LB_BLOCK + 2 * PLOC_R) * 6];
    int b0 = (int)blockIdx.x * (int)LB_BLOCK - PLOC_R;
    for (int t = threadIdx.x; t < (int)LB_BLOCK + 2 * PLOC_R; t += LB_BLOCK) {
        int q = b0 + t;
        if (q >= 0 && q < m) {
            const float* src = box + (size_t)C[q] * 6;
            for (int c = 0; c < 6; c++) sb[t * 6 + c] = src[c];
        }
    }
    __syncthreads();
    int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (i >= m) return;
    const float* a = &sb[(i - b0) * 6];
    float best = 0.f;
    int bj = -1;
    int j0 = i - PLOC_R < 0 ? 0 : i - PLOC_R, j1 = i + PLOC_R > m - 1 ? m - 1 : i + PLOC_R;
    for (int j = j0; j <= j1; j++) {
        if (j == i) continue;
        const float* b = &sb[(j - b0) * 6];
        float u[6];
        for (int c = 0; c < 3; c++) {
            u[c] = mn(a[c], b[c]);
            u[3 + c] = mx(a[3 + c], b[3 + c]);
        }
        float ar = half_area(u);
        if (bj < 0 || ar < best) {
            best = ar;
            bj = j;
        }
    }
    nn[i] = bj;
}

// per position: bit 0 keep (the cluster or its merge survives here), bit 32
// new (a merge is created here)
__global__ __launch_bounds__(LB_BLOCK) void k_ploc_flags(const int* __restrict__ nn, int m,
                                                        unsigned long long* __restrict__ flags) {
    int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (i >= m) return;
    int j = nn[i];
    bool mutual = nn[j] == i;
    unsigned long long keep = (!mutual || i < j) ? 1ull : 0ull;
    unsigned long long made = (mutual && i < j) ? 1ull : 0ull;
    flags[i] = keep | (made << 32);
}

__global__ __launch_bounds__(LB_BLOCK) void k_ploc_merge(const int* __restrict__ C, const int* __restrict__ nn, int m,
                                                        const unsigned long long* __restrict__ flags,
                                                        const unsigned long long* __restrict__ scan, int base,
                                                        PNodes N, int* __restrict__ Cout) {
    int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (i >= m) return;
    unsigned long long f = flags[i];
    if (!(f & 1ull)) return;
    int dst = (int)(scan[i] & 0xffffffffull);
    if (!(f >> 32)) {
        Cout[dst] = C[i];
        return;
    }
    int id = base + (int)(scan[i] >> 32);
    int a = C[i], b = C[nn[i]];  // a at the lower position
    const float* ba = N.box + (size_t)a * 6;
    const float* bb = N.box + (size_t)b * 6;
    float* o = N.box + (size_t)id * 6;
    float u[6];
    for (int c = 0; c < 3; c++) {
        u[c] = mn(ba[c], bb[c]);
        u[3 + c] = mx(ba[3 + c], bb[3 + c]);
    }
    for (int c = 0; c < 6; c++) o[c] = u[c];
    // split axis: the largest separation of the children's centres (lowest
    // axis on ties); the child with the lower centre on it is visited first
    int ax = 0;
    float sep = 0.f, sgn = 0.f;
    for (int c = 0; c < 3; c++) {
        float d = (bb[c] + bb[3 + c]) - (ba[c] + ba[3 + c]);
        float ad = d < 0.f ? -d : d;
        if (c == 0 || ad > sep) {
            sep = ad;
            sgn = d;
            ax = c;
        }
    }
    int f1 = sgn < 0.f ? b : a, f2 = sgn < 0.f ? a : b;
    int cnt = N.count[a] + N.count[b];
    float A = half_area(u);
    float c_split = (0.125f * A + N.cost[a]) + N.cost[b];
    float c_leaf = (float)cnt * A;
    bool collapse = cnt <= N.max_prims && c_leaf <= c_split;
    N.first[id] = f1;
    N.second[id] = f2;
    N.parent[id] = -1;
    N.parent[a] = id;
    N.parent[b] = id;
    N.count[id] = cnt;
    N.cost[id] = collapse ? c_leaf : c_split;
    N.leaf[id] = collapse ? 1 : 0;
    N.osize[id] = collapse ? 1 : 1 + N.osize[f1] + N.osize[f2];
    N.axis[id] = (uint8_t)ax;
    Cout[dst] = id;
}

// Every node finds its depth-first output index and the depth-first index
// of its first triangle by walking to the root; triangles place themselves
// in the slot order, output nodes write their records.
__global__ __launch_bounds__(LB_BLOCK) void k_ploc_emit(PNodes N, const uint64_t* __restrict__ keys, int n,
                                                       ffh::BvhNode* __restrict__ out, int32_t* __restrict__ ordered) {
    int v = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (v >= 2 * n - 1) return;
    int idx = 0, poff = 0;
    bool is_out = true;
    for (int cur = v, p = N.parent[v]; p >= 0; cur = p, p = N.parent[p]) {
        if (N.leaf[p]) is_out = false;
        int f = N.first[p];
        if (N.second[p] == cur) {
            idx += 1 + N.osize[f];
            poff += N.count[f];
        } else {
            idx += 1;
        }
    }
    if (v < n) ordered[poff] = (int32_t)(keys[v] & 0xffffffffu);
    if (!is_out) return;
    ffh::BvhNode r;
    const float* b = N.box + (size_t)v * 6;
    for (int c = 0; c < 3; c++) {
        r.bmin[c] = b[c];
        r.bmax[c] = b[3 + c];
    }
    r.pad = 0;
    if (N.leaf[v]) {
        r.offset = poff;
        r.n_prims = (uint16_t)N.count[v];
        r.axis = 0;
    } else {
        r.offset = idx + 1 + N.osize[N.first[v]];
        r.n_prims = 0;
        r.axis = N.axis[v];
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

// PLOC tree (see the kernels above) in the SAH builder's node format: the
// Morton keys of build_lbvh, then clustering rounds until one cluster is left
// (the cluster count read back after each round), then the depth-first
// emission.  positions on the host; gpu_ms as for build_lbvh.
hipError_t build_ploc(const float* positions, int32_t n, int32_t max_prims, std::vector<ffh::BvhNode>& nodes,
                      std::vector<int32_t>& ordered, float* gpu_ms) {
    if (n <= 0) return hipErrorInvalidValue;
    if (max_prims < 1) max_prims = 1;
    if (max_prims > 65535) max_prims = 65535;
    hipStream_t s = nullptr;
    float *d_pos = nullptr, *d_box = nullptr, *d_cost = nullptr;
    uint64_t *d_keys = nullptr, *d_sorted = nullptr;
    unsigned long long *d_flags = nullptr, *d_scan = nullptr;
    uint32_t* d_cb = nullptr;
    int *d_first = nullptr, *d_second = nullptr, *d_parent = nullptr, *d_count = nullptr, *d_osize = nullptr;
    int *d_C0 = nullptr, *d_C1 = nullptr, *d_nn = nullptr, *d_ord = nullptr;
    uint8_t *d_axis = nullptr, *d_leaf = nullptr;
    ffh::BvhNode* d_nodes = nullptr;
    void *d_tmp = nullptr, *d_tmp2 = nullptr;
    size_t m0 = (size_t)n, nn2 = 2 * m0 - 1, tmp_bytes = 0, tmp2_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess && gpu_ms) e = hipEventCreate(&ev0);
    if (e == hipSuccess && gpu_ms) e = hipEventCreate(&ev1);
    auto ok = [&](hipError_t x) { if (e == hipSuccess) e = x; return e == hipSuccess; };
    if (ok(alloc(&d_pos, 9 * m0)) && ok(alloc(&d_keys, m0)) && ok(alloc(&d_sorted, m0)) && ok(alloc(&d_cb, 6)) &&
        ok(alloc(&d_box, 6 * nn2)) && ok(alloc(&d_cost, nn2)) && ok(alloc(&d_first, nn2)) &&
        ok(alloc(&d_second, nn2)) && ok(alloc(&d_parent, nn2)) && ok(alloc(&d_count, nn2)) &&
        ok(alloc(&d_osize, nn2)) && ok(alloc(&d_axis, nn2)) && ok(alloc(&d_leaf, nn2)) && ok(alloc(&d_C0, m0)) &&
        ok(alloc(&d_C1, m0)) && ok(alloc(&d_nn, m0)) && ok(alloc(&d_flags, m0)) && ok(alloc(&d_scan, m0)) &&
        ok(alloc(&d_nodes, nn2)) && ok(alloc(&d_ord, m0)) &&
        ok(hipMemcpyAsync(d_pos, positions, sizeof(float) * 9 * m0, hipMemcpyHostToDevice, s)) &&
        ok(rocprim::radix_sort_keys(nullptr, tmp_bytes, d_keys, d_sorted, m0, 0, 62, s)) &&
        ok(hipMalloc(&d_tmp, tmp_bytes ? tmp_bytes : 1)) &&
        ok(rocprim::exclusive_scan(nullptr, tmp2_bytes, d_flags, d_scan, 0ull, m0, rocprim::plus<unsigned long long>(),
                                   s)) &&
        ok(hipMalloc(&d_tmp2, tmp2_bytes ? tmp2_bytes : 1))) {
        uint32_t g = (uint32_t)((m0 + LB_BLOCK - 1) / LB_BLOCK);
        uint32_t gb = g < 4096 ? g : 4096;
        PNodes N{d_box, d_first, d_second, d_parent, d_count, d_cost, d_osize, d_axis, d_leaf, max_prims};
        if (ev0) ok(hipEventRecord(ev0, s));
        hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, d_cb);
        hipLaunchKernelGGL(k_centroid_bounds, dim3(gb), dim3(LB_BLOCK), 0, s, d_pos, (uint32_t)n, d_cb);
        hipLaunchKernelGGL(k_morton, dim3(gb), dim3(LB_BLOCK), 0, s, d_pos, (uint32_t)n, d_cb, d_keys);
        if (ok(hipGetLastError()) && ok(rocprim::radix_sort_keys(d_tmp, tmp_bytes, d_keys, d_sorted, m0, 0, 62, s))) {
            hipLaunchKernelGGL(k_ploc_leaves, dim3(g), dim3(LB_BLOCK), 0, s, d_sorted, d_pos, n, N, d_C0);
            int m = n, base = n;
            int *C = d_C0, *Cn = d_C1;
            while (m > 1 && ok(hipGetLastError())) {
                uint32_t gm = (uint32_t)((m + LB_BLOCK - 1) / LB_BLOCK);
                hipLaunchKernelGGL(k_ploc_nn, dim3(gm), dim3(LB_BLOCK), 0, s, C, m, d_box, d_nn);
                hipLaunchKernelGGL(k_ploc_flags, dim3(gm), dim3(LB_BLOCK), 0, s, d_nn, m, d_flags);
                if (!ok(hipGetLastError()) ||
                    !ok(rocprim::exclusive_scan(d_tmp2, tmp2_bytes, d_flags, d_scan, 0ull, (size_t)m,
                                                rocprim::plus<unsigned long long>(), s)))
                    break;
                hipLaunchKernelGGL(k_ploc_merge, dim3(gm), dim3(LB_BLOCK), 0, s, C, d_nn, m, d_flags, d_scan, base, N,
                                   Cn);
                unsigned long long tail[2] = {0, 0};
                if (!ok(hipGetLastError()) ||
                    !ok(hipMemcpyAsync(&tail[0], d_scan + (m - 1), 8, hipMemcpyDeviceToHost, s)) ||
                    !ok(hipMemcpyAsync(&tail[1], d_flags + (m - 1), 8, hipMemcpyDeviceToHost, s)) ||
                    !ok(hipStreamSynchronize(s)))
                    break;
                unsigned long long tot = tail[0] + tail[1];
                int made = (int)(tot >> 32), kept = (int)(tot & 0xffffffffull);
                if (made < 1 || kept >= m) {  // a round always merges at least one pair
                    e = hipErrorUnknown;
                    break;
                }
                base += made;
                m = kept;
                int* t = C;
                C = Cn;
                Cn = t;
            }
            if (e == hipSuccess)
                hipLaunchKernelGGL(k_ploc_emit, dim3((uint32_t)((nn2 + LB_BLOCK - 1) / LB_BLOCK)), dim3(LB_BLOCK), 0, s,
                                   N, d_sorted, n, d_nodes, d_ord);
            if (ev1) ok(hipEventRecord(ev1, s));
            int32_t n_out = 1;
            if (ok(hipGetLastError()))
                ok(hipMemcpyAsync(&n_out, d_osize + (nn2 - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
            if (ok(hipStreamSynchronize(s))) {
                if (ev0) ok(hipEventElapsedTime(gpu_ms, ev0, ev1));
                nodes.resize((size_t)n_out);
                ordered.resize(m0);
                if (ok(hipMemcpy(nodes.data(), d_nodes, sizeof(ffh::BvhNode) * (size_t)n_out, hipMemcpyDeviceToHost)))
                    ok(hipMemcpy(ordered.data(), d_ord, sizeof(int32_t) * m0, hipMemcpyDeviceToHost));
            }
        }
    }
    for (void* p : {(void*)d_pos, (void*)d_keys, (void*)d_sorted, (void*)d_cb, (void*)d_box, (void*)d_cost,
                    (void*)d_first, (void*)d_second, (void*)d_parent, (void*)d_count, (void*)d_osize, (void*)d_axis,
                    (void*)d_leaf, (void*)d_C0, (void*)d_C1, (void*)d_nn, (void*)d_flags, (void*)d_scan,
                    (void*)d_nodes, (void*)d_ord, d_tmp, d_tmp2})
        if (p) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (s) (void)hipStreamDestroy(s);
    return e;
}

}  // namespace ffk
