// ff_bvh.cpp -- host SAH BVH builder for the wavefront tracer.
//
// Restates BVHAccel's SAH build (CRT/BVH.cpp:52-331) over the triangle bounds
// that dScene::init_BVH_triangle_info forms (CRT/dScene.cpp:324-362), so the
// flattened node array and slot order are the ones the reference's traversal
// order is defined on:
//   * centroid-bounds maximum-extent axis; leaf for 1 primitive or a
//     degenerate centroid extent;
//   * 12 buckets, cost 0.125 + (n0*SA0 + n1*SA1)/SA with SA in double of the
//     float diagonal products, first minimum wins;
//   * leaf when n <= maxPrims and minCost >= n;
//   * two-pointer partition (the classic algorithm of std::partition on
//     bidirectional iterators) by bucket <= split bucket; two primitives are
//     ordered by centroid (nth_element on two elements);
//   * depth-first flattening, first child at +1.
// Min/max use the reference host build's ternary fminf/fmaxf
// (CRT/dMath.cuh:40-48) and Union's min/max-then-construct sequence
// (CRT/BVH.cpp:26-50).
//
// Parallel build, same output: the two children of a node own disjoint
// primitive ranges, so subtrees of large nodes are built by separate threads
// into node lists of their own and spliced in depth-first order afterwards
// (child offsets and leaf slots relocated).  Every node is computed by the
// same code on the same range as in a one-thread build, so the node array and
// slot order are bit-identical.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <future>
#include <thread>
#include <vector>

#include "ff_internal.h"

namespace ffh {
namespace {

inline float fmin_r(float a, float b) { return a < b ? a : b; }
inline float fmax_r(float a, float b) { return a > b ? a : b; }

struct Box {
    float lo[3], hi[3];
    static Box empty() {
        Box b;
        for (int k = 0; k < 3; k++) { b.lo[k] = FLT_MAX; b.hi[k] = -FLT_MAX; }
        return b;
    }
    static Box make(const float* p1, const float* p2) {
        Box b;
        for (int k = 0; k < 3; k++) { b.lo[k] = fmin_r(p1[k], p2[k]); b.hi[k] = fmax_r(p1[k], p2[k]); }
        return b;
    }
    Box merged(const float* p) const {
        float l[3], h[3];
        for (int k = 0; k < 3; k++) { l[k] = fmin_r(lo[k], p[k]); h[k] = fmax_r(hi[k], p[k]); }
        return make(l, h);
    }
    Box merged(const Box& o) const {
        float l[3], h[3];
        for (int k = 0; k < 3; k++) { l[k] = fmin_r(lo[k], o.lo[k]); h[k] = fmax_r(hi[k], o.hi[k]); }
        return make(l, h);
    }
    double area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0 * (double)(dx * dy + dx * dz + dy * dz);
    }
    int widest() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx > dy && dx > dz) return 0;
        if (dy > dz) return 1;
        return 2;
    }
};

struct Prim {
    int32_t id;
    Box box;
    float c[3];
};

constexpr int kBuckets = 12;

class Builder {
public:
    Builder(std::vector<Prim>& prims, int32_t max_prims, std::vector<BvhNode>& nodes,
            std::vector<int32_t>& ordered)
        : prims_(prims), max_prims_(max_prims), nodes_(nodes), ordered_(ordered) {}

    void build(int start, int end) {
        int32_t idx = (int32_t)nodes_.size();
        nodes_.emplace_back();
        Box bounds;
        int dim, mid;
        if (!split(start, end, bounds, dim, mid)) return leaf(idx, start, end, bounds);
        build(start, mid);
        int32_t second = (int32_t)nodes_.size();
        build(mid, end);
        interior(nodes_[idx], node_box(idx + 1), node_box(second), second, dim);
    }

    // Bounds of [start, end) and, unless the range becomes a leaf (false),
    // the split axis and the partition point (the range is partitioned).
    bool split(int start, int end, Box& bounds, int& dim, int& mid) {
        bounds = Box::empty();
        for (int i = start; i < end; i++) bounds = bounds.merged(prims_[i].box);
        int n = end - start;
        if (n == 1) return false;
        Box cb = Box::empty();
        for (int i = start; i < end; i++) cb = cb.merged(prims_[i].c);
        dim = cb.widest();
        if (cb.hi[dim] == cb.lo[dim]) return false;
        if (n <= 2) {
            mid = (start + end) / 2;
            if (prims_[start + 1].c[dim] < prims_[start].c[dim]) std::swap(prims_[start], prims_[start + 1]);
        } else {
            int count[kBuckets] = {0};
            Box bb[kBuckets];
            for (auto& b : bb) b = Box::empty();
            for (int i = start; i < end; i++) {
                int b = bucket(cb, dim, prims_[i].c);
                count[b]++;
                bb[b] = bb[b].merged(prims_[i].box);
            }
            double sa = bounds.area();
            float best = 0.f;
            int best_b = 0;
            for (int s = 0; s < kBuckets - 1; s++) {
                Box b0 = Box::empty(), b1 = Box::empty();
                int c0 = 0, c1 = 0;
                for (int j = 0; j <= s; j++) { b0 = b0.merged(bb[j]); c0 += count[j]; }
                for (int j = s + 1; j < kBuckets; j++) { b1 = b1.merged(bb[j]); c1 += count[j]; }
                float cost = (float)((double)0.125f + (c0 * b0.area() + c1 * b1.area()) / sa);
                if (s == 0 || cost < best) { best = cost; best_b = s; }
            }
            if (!(n > max_prims_ || best < (float)n)) return false;
            mid = partition(start, end, cb, dim, best_b);
        }
        return true;
    }

    static void interior(BvhNode& nd, const Box& c0, const Box& c1, int32_t second, int dim) {
        Box ib = c0.merged(c1);
        std::memcpy(nd.bmin, ib.lo, 12);
        std::memcpy(nd.bmax, ib.hi, 12);
        nd.offset = second;
        nd.n_prims = 0;
        nd.axis = (uint8_t)dim;
        nd.pad = 0;
    }
    static Box box_of(const BvhNode& nd) {
        Box b;
        std::memcpy(b.lo, nd.bmin, 12);
        std::memcpy(b.hi, nd.bmax, 12);
        return b;
    }
    void leaf(int32_t idx, int start, int end, const Box& b) {
        BvhNode& nd = nodes_[idx];
        nd.offset = (int32_t)ordered_.size();
        for (int i = start; i < end; i++) ordered_.push_back(prims_[i].id);
        std::memcpy(nd.bmin, b.lo, 12);
        std::memcpy(nd.bmax, b.hi, 12);
        nd.n_prims = (uint16_t)(end - start);
        nd.axis = 0;
        nd.pad = 0;
    }

private:
    static int bucket(const Box& cb, int dim, const float* c) {
        float o = c[dim] - cb.lo[dim];
        if (cb.hi[dim] > cb.lo[dim]) o /= cb.hi[dim] - cb.lo[dim];
        int b = (int)(12.0f * o);
        return b == kBuckets ? kBuckets - 1 : b;
    }
    int partition(int first, int last, const Box& cb, int dim, int split) {
        auto left = [&](int i) { return bucket(cb, dim, prims_[i].c) <= split; };
        for (;;) {
            for (;;) {
                if (first == last) return first;
                if (left(first)) ++first;
                else break;
            }
            --last;
            for (;;) {
                if (first == last) return first;
                if (!left(last)) --last;
                else break;
            }
            std::swap(prims_[first], prims_[last]);
            ++first;
        }
    }
    Box node_box(int32_t i) const { return box_of(nodes_[i]); }

    std::vector<Prim>& prims_;
    int32_t max_prims_;
    std::vector<BvhNode>& nodes_;
    std::vector<int32_t>& ordered_;
};

// A subtree built on its own: node indices and leaf slots start at 0.
struct SubTree {
    std::vector<BvhNode> nodes;
    std::vector<int32_t> ordered;
};

// Appends `t` (built on its own) to `out`, moving its interior offsets and
// leaf slots to where it lands.
void splice(SubTree& out, const SubTree& t) {
    size_t nb = out.nodes.size(), ob = out.ordered.size();
    out.nodes.resize(nb + t.nodes.size());
    BvhNode* dst = out.nodes.data() + nb;
    for (size_t i = 0; i < t.nodes.size(); i++) {
        BvhNode nd = t.nodes[i];
        nd.offset += nd.n_prims == 0 ? (int32_t)nb : (int32_t)ob;
        dst[i] = nd;
    }
    out.ordered.insert(out.ordered.end(), t.ordered.begin(), t.ordered.end());
}

// Builds [start, end) into `out`; large ranges build their first child on
// another thread (up to `spawn` more levels deep) while this one builds the
// second.
void build_parallel(std::vector<Prim>& prims, int32_t max_prims, int start, int end, int spawn, int min_par,
                    SubTree& out) {
    if (spawn <= 0 || end - start < min_par) {
        Builder(prims, max_prims, out.nodes, out.ordered).build(start, end);
        return;
    }
    Builder b(prims, max_prims, out.nodes, out.ordered);
    out.nodes.emplace_back();
    Box bounds;
    int dim, mid;
    if (!b.split(start, end, bounds, dim, mid)) {
        b.leaf(0, start, end, bounds);
        return;
    }
    SubTree left, right;
    left.nodes.reserve((size_t)2 * (mid - start));
    right.nodes.reserve((size_t)2 * (end - mid));
    auto task = std::async(std::launch::async, build_parallel, std::ref(prims), max_prims, start, mid, spawn - 1,
                           min_par, std::ref(left));
    build_parallel(prims, max_prims, mid, end, spawn - 1, min_par, right);
    task.get();
    int32_t second = 1 + (int32_t)left.nodes.size();
    Builder::interior(out.nodes[0], Builder::box_of(left.nodes[0]), Builder::box_of(right.nodes[0]), second, dim);
    splice(out, left);
    splice(out, right);
}

}  // namespace

void build_bvh_sah(const float* positions, int32_t n, int32_t max_prims, std::vector<BvhNode>& nodes,
                   std::vector<int32_t>& ordered) {
    nodes.clear();
    ordered.clear();
    if (n <= 0) return;
    std::vector<Prim> prims((size_t)n);
    auto init = [&](int32_t lo, int32_t hi) {
        for (int32_t i = lo; i < hi; i++) {
            const float* t = positions + (size_t)i * 9;
            Box b = Box::make(t, t + 3).merged(t + 6);
            prims[i].id = i;
            prims[i].box = b;
            for (int k = 0; k < 3; k++) prims[i].c[k] = 0.5f * b.lo[k] + 0.5f * b.hi[k];
        }
    };
    unsigned hw = std::thread::hardware_concurrency();
    int n_init = n < 65536 ? 1 : (int)std::min(hw ? hw : 1u, 16u);
    std::vector<std::thread> th;
    for (int k = 1; k < n_init; k++)
        th.emplace_back(init, (int32_t)((int64_t)n * k / n_init), (int32_t)((int64_t)n * (k + 1) / n_init));
    init(0, (int32_t)((int64_t)n / n_init));
    for (auto& x : th) x.join();
    // subtree threads: up to 4x the cores (32 at most) for balance, none for
    // ranges under 16384 primitives
    int spawn = 0;
    while ((1u << spawn) < 4u * (hw ? hw : 1u) && spawn < 5) spawn++;
    SubTree t;
    t.nodes.reserve((size_t)2 * n);
    t.ordered.reserve((size_t)n);
    build_parallel(prims, max_prims, 0, n, spawn, 16384, t);
    nodes.swap(t.nodes);
    ordered.swap(t.ordered);
}

int32_t bvh_max_depth(const std::vector<BvhNode>& nodes) {
    if (nodes.empty()) return 0;
    // iterative depth-first walk: (node, depth)
    std::vector<std::pair<int32_t, int32_t>> st;
    st.emplace_back(0, 1);
    int32_t best = 0;
    while (!st.empty()) {
        auto [i, d] = st.back();
        st.pop_back();
        if (d > best) best = d;
        if (nodes[i].n_prims == 0) {
            st.emplace_back(i + 1, d + 1);
            st.emplace_back(nodes[i].offset, d + 1);
        }
    }
    return best;
}

}  // namespace ffh
