// ff_kernels.hip -- gfx950 kernels of the wavefront path tracer and their
// host launchers.  Stage structure follows the reference's wavefront
// (CRT/kernel.cu:283-676): generate, extend, shadow, logic, material queues,
// plus the FP64 environment CDF build and the film resolve.
//
// Every stage kernel serves a table of render requests: a path slot knows its
// request (scene, environment, camera, depth, film) through P.preq, so one
// wavefront renders a single frame or many uneven requests at once.  The
// batch index lives in device memory (P.bstate) so the per-sample sequence of
// launches is identical from batch to batch and can be replayed as a graph.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "ff_device.h"
#include "ff_internal.h"

using namespace ffd;

namespace {

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t gstride() { return gridDim.x * blockDim.x; }

__device__ __forceinline__ Bvh bvh_of(const ff_dev_scene& S) {
    Bvh b;
    b.nodes = S.nodes;
    b.nodes4 = S.nodes4;
    b.leaf_count = S.leaf_count;
    b.root = S.root;
    return b;
}

// With UNI every request of the call shades the same scene and environment:
// they are read from request 0, so their pointers stay in scalar registers.
template <bool UNI>
__device__ __forceinline__ const ff_dev_scene& scene_of(const ff_paths& P, const ff_dev_request& R) {
    return UNI ? P.reqs[0].S : R.S;
}
template <bool UNI>
__device__ __forceinline__ const ff_dev_env& env_of(const ff_paths& P, const ff_dev_request& R) {
    return UNI ? P.reqs[0].E : R.E;
}

__device__ __forceinline__ EnvView env_view(const ff_dev_env& e) {
    EnvView v;
    v.rgbp = e.rgbp; v.marg_y = e.marg_y; v.conds = e.conds; v.w = e.w; v.h = e.h;
    v.guide_m = e.guide_m; v.guide_c = e.guide_c;
    return v;
}

// this workgroup's replica of the ray statistics
__device__ __forceinline__ unsigned long long* stats_rep(const ff_paths& P) {
    return P.stats + (size_t)(blockIdx.x % S_REPLICAS) * S_STRIDE;
}

// request of a path: a one-request call does not store it
__device__ __forceinline__ uint32_t req_of(const ff_paths& P, uint32_t id) {
    return P.n_req == 1 ? 0u : P.preq[id];
}

// request owning a slot: last r with slot_base <= id
__device__ __forceinline__ uint32_t find_req(const ff_dev_request* __restrict__ reqs, uint32_t n, uint32_t id) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (reqs[mid].slot_base <= id) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ============================================================ batch state
// Starts a batch: batch index += 1 (or = value), queue and work counters
// cleared, and the extension queue laid out for the batch's new samples:
// request r's active slots (its first min(K, spp - b*K) layers, a contiguous
// slot range) occupy queue entries [qbase[r], qbase[r] + count), so
// generate places every path without an atomic.
__global__ void k_batch_begin(ff_paths P, int set, uint32_t value) {
    uint32_t i = threadIdx.x;
    if (i < C_COUNT) P.counters[i] = 0;
    for (uint32_t k = i; k < 3 * FETCH_SEGS; k += blockDim.x) P.fetch[(size_t)k * FETCH_STRIDE] = 0;
    __syncthreads();
    if (i == 0) {
        uint32_t b = set ? value : P.bstate[B_BATCH] + 1u;
        P.bstate[B_BATCH] = b;
        if (set) return;  // (re)initialisation only
        uint32_t q = 0;
        for (uint32_t r = 0; r < P.n_req; r++) {
            const ff_dev_request& R = P.reqs[r];
            int64_t rem = (int64_t)R.spp - (int64_t)b * R.K;
            uint32_t layers = rem <= 0 ? 0u : (rem < (int64_t)R.K ? (uint32_t)rem : R.K);
            P.qbase[r] = q;
            q += layers * R.W * R.H;
        }
        P.qbase[P.n_req] = q;
        P.counters[C_EXT0] = q;
    }
}

// ================================================================ generate
// wf_generate + dCamera::gen_ray (CRT/kernel.cu:457-485): every active slot
// of the batch starts its next sample; its extension-queue entry is
// qbase[request] + (slot - slot_base), so the queue is in slot order.
__global__ __launch_bounds__(FF_BLOCK) void k_generate(ff_paths P) {
    const uint32_t b = P.bstate[B_BATCH];
    for (uint32_t id = gtid(); id < P.n_slots; id += gstride()) {
        uint32_t r = find_req(P.reqs, P.n_req, id);
        const ff_dev_request& R = P.reqs[r];
        uint32_t WH = R.W * R.H, local = id - R.slot_base;
        uint32_t layer = local / WH, pix = local % WH;
        int64_t rem = (int64_t)R.spp - (int64_t)b * R.K;
        bool act = rem > 0 && (int64_t)layer < rem && layer < R.K;
        if (!act) continue;
        uint32_t sample = R.sample_offset + b * R.K + layer;
        uint32_t base = rng_base(R.seed, pix, sample);
        f3 o, d;
        gen_ray(R.C, (int)R.W, (int)R.H, (int)(pix % R.W), (int)(pix / R.W), base, o, d);
        P.ray[2 * id] = f4(o, __uint_as_float(base));
        P.ray[2 * id + 1] = f4(d, 0.f);  // state word 0: outside every surface
        if (P.n_req > 1) P.preq[id] = r;
        if (R.mode & RQ_ADJOINT) {
            // replay: L_rem starts as the stored radiance of this path
            P.beta[id] = make_float4(1.f, 1.f, 1.f, __uint_as_float(ST_PREV_DELTA));
            uint32_t sl = b * R.K + layer;
            P.Lp[id] = R.path_L[(size_t)sl * WH + pix];
            float4 g = R.d_img[pix];
            P.dlt[id] = make_float4(g.x * R.inv_spp, g.y * R.inv_spp, g.z * R.inv_spp, 0.f);
        }
        // a forward camera path starts at beta = 1, L = 0, state ST_PREV_DELTA:
        // wf_logic's first pass takes those as constants, so they are not stored
        P.q_ext[0][P.qbase[r] + local] = id;
    }
}

// =============================================================== traversal
// Ray sources of the persistent traversal kernels: load() binds queue entry k
// to a ray and (unless kUniform) its scene, hit()/occluded() write the result.
// kUniform: every ray of the launch uses one scene, bound once per wave by
// scene0() (the BVH and triangle pointers then stay in scalar registers and
// a refill needs no request-table lookup).
template <bool UNI>
struct ExtSrc {  // wf_extend (CRT/kernel.cu:487-523): closest hit of every queued path
    static constexpr bool kUniform = UNI;
    static constexpr bool kChain = false;
    ff_paths P;
    const uint32_t* queue;
    __device__ __forceinline__ void scene0(Bvh& B, const float*& tris) const {
        const ff_dev_scene& S = P.reqs[0].S;
        B = bvh_of(S);
        tris = S.tris;
    }
    __device__ __forceinline__ void load(uint32_t k, uint32_t& id, uint32_t&, Bvh& B, const float*& tris, f3& o,
                                         f3& d, bool& two) const {
        id = queue[k];
        if (!UNI) {
            const ff_dev_scene& S = P.reqs[req_of(P, id)].S;
            B = bvh_of(S);
            tris = S.tris;
        }
        o = xyz(P.ray[2 * id]);
        float4 r1 = P.ray[2 * id + 1];
        d = xyz(r1);
        two = (__float_as_uint(r1.w) & ST_INSIDE) != 0;
    }
    __device__ __forceinline__ void hit(uint32_t id, const Hit& h) const {
        P.hit[id] = make_float4(h.t, h.u, h.v, __int_as_float(h.id));
    }
    __device__ __forceinline__ void occluded(uint32_t, uint32_t, bool) const {}
    __device__ __forceinline__ bool advance(uint32_t&, bool, uint32_t&) const { return false; }
    __device__ __forceinline__ void load_next(uint32_t, uint32_t, f3&, f3&) const {}
    __device__ __forceinline__ void count(uint32_t, uint32_t) const {}
};
// A shadow-queue entry is one shading event, path << 2 | needS | needD << 1:
// its light ray first, then -- only when the light ray is unoccluded, since
// its visibility gates the whole NEE sum -- the BRDF-sample visibility rays
// (spec, then diffuse) in the same lane.  aux = ray (0 light, 1 spec,
// 2 diffuse) | rays still to trace << 2.  The reference traces the BRDF-sample
// rays inside BRDF_L either way (CRT/dMaterial.cu:404-448); where the light
// ray is occluded their results are never used, so no pixel changes.
template <bool UNI>
struct ShadowSrc {  // wf_shadow (CRT/kernel.cu:525-543): any hit of every queued shadow ray
    static constexpr bool kUniform = UNI;
    static constexpr bool kChain = true;
    ff_paths P;
    __device__ __forceinline__ void scene0(Bvh& B, const float*& tris) const {
        const ff_dev_scene& S = P.reqs[0].S;
        B = bvh_of(S);
        tris = S.tris;
    }
    __device__ __forceinline__ void load(uint32_t k, uint32_t& id, uint32_t& aux, Bvh& B, const float*& tris, f3& o,
                                         f3& d, bool& two) const {
        uint32_t e = P.q_shadow[k];
        id = e >> 2;
        aux = (e & 3u) << 2;
        if (!UNI) {
            const ff_dev_scene& S = P.reqs[req_of(P, id)].S;
            B = bvh_of(S);
            tris = S.tris;
        }
        o = xyz(P.sray[6 * id]);
        d = xyz(P.sray[6 * id + 1]);
        two = false;
    }
    __device__ __forceinline__ void hit(uint32_t, const Hit&) const {}
    __device__ __forceinline__ void occluded(uint32_t id, uint32_t aux, bool occ) const {
        P.vis[3 * id + (aux & 3u)] = occ ? 0 : 1;
    }
    // after a ray with result occ: whether the path has another ray (aux then names it)
    __device__ __forceinline__ bool advance(uint32_t& aux, bool occ, uint32_t& gated) const {
        uint32_t cur = aux & 3u, rest = aux >> 2;
        if (cur == 0u && occ) gated += (uint32_t)__popc(rest);
        if ((cur == 0u && occ) || rest == 0u) return false;
        uint32_t nxt = (rest & 1u) ? 1u : 2u;
        aux = nxt | ((rest & ~nxt) << 2);
        return true;
    }
    __device__ __forceinline__ void load_next(uint32_t id, uint32_t aux, f3& o, f3& d) const {
        o = xyz(P.sray[6 * id + 2 * (aux & 3u)]);
        d = xyz(P.sray[6 * id + 2 * (aux & 3u) + 1]);
    }
    __device__ __forceinline__ void count(uint32_t rays, uint32_t gated) const {
        wave_sum_add(stats_rep(P) + S_SHADOW, rays);
        wave_sum_add(stats_rep(P) + S_SHD_GATED, gated);
    }
};
struct RaySrc {  // a plain ray array (ff_trace_closest / ff_trace_any)
    static constexpr bool kUniform = true;
    static constexpr bool kChain = false;
    ff_dev_scene S;
    const float4* rays;
    float4* hits;
    uint8_t* occ;
    int two_sided;
    __device__ __forceinline__ void scene0(Bvh& B, const float*& tris) const {
        B = bvh_of(S);
        tris = S.tris;
    }
    __device__ __forceinline__ void load(uint32_t k, uint32_t& id, uint32_t&, Bvh&, const float*&, f3& o, f3& d,
                                         bool& two) const {
        id = k;
        o = xyz(rays[2 * k]);
        d = xyz(rays[2 * k + 1]);
        two = two_sided != 0;
    }
    __device__ __forceinline__ void hit(uint32_t id, const Hit& h) const {
        hits[id] = make_float4(h.t, h.u, h.v, __int_as_float(h.id));
    }
    __device__ __forceinline__ void occluded(uint32_t id, uint32_t, bool o) const { occ[id] = o ? 1 : 0; }
    __device__ __forceinline__ bool advance(uint32_t&, bool, uint32_t&) const { return false; }
    __device__ __forceinline__ void load_next(uint32_t, uint32_t, f3&, f3&) const {}
    __device__ __forceinline__ void count(uint32_t, uint32_t) const {}
};

// Persistent traversal.  The grid holds only the lanes the chip keeps
// resident; a lane whose ray is finished takes the next queue entry from the
// segmented work counters (per wave: one ballot and one atomic for all idle
// lanes), so a wave no longer lasts as long as its slowest ray and no second
// round of waves is needed.  A wave refills once REFILL of its lanes are idle
// (or all are), trading idle lanes against refill stalls.  Every wave leaves
// the loop once all segments are exhausted and its own rays are finished.
// WW selects the loop shape: 0 = one step (interior record or leaf) per
// lane per iteration; 1 = "while-while": lanes at interior records step
// until every busy lane has reached a leaf, then all leaves are tested
// together (fewer divergent switches between the two code paths);
// 2 = speculative while-while (Aila and Laine, "Understanding the Efficiency
// of Ray Traversal on GPUs", HPG 2009): a lane that reaches a leaf while it
// holds none postpones it and keeps walking interior records until every
// lane holds a leaf or has no interior record left, then the postponed (and
// current) leaves are tested together.  Each lane still tests its leaves in
// its visit order, so hits are those of the other shapes; pruning then sees
// the closest hit of the leaves already tested, so the boxes tested depend
// on the lanes that share the wave.  3 = leaf batching: a lane standing on a
// leaf queues it (two at most) and keeps walking; the queued leaves are
// tested together once 8 lanes hold one, or when no lane can walk on -- the
// leaf code then runs for many lanes at once instead of for the few that
// reach a leaf in any one step.  Leaves are still tested in each lane's visit
// order.  5, 6, 7 = one queued leaf per lane (fewer registers), tested
// together once 8, 4 or 16 lanes hold one.
// WIDE walks the 4-wide records (half the dependent interior steps, same
// leaf order) instead of the two-child records.
template <int BS, int LDS, int REFILL, bool ANY, int WW, bool WIDE, class Src>
__device__ __forceinline__ void trav_persistent(const Src& src, uint32_t n, uint32_t* fetch, int32_t* ovf,
                                                unsigned long long* cnt, int* smem) {
    SegFetch F;
    F.init(fetch, n, blockIdx.x * (BS / 64) + (threadIdx.x >> 6));
    Stack st;
    st.lds = (lds_int*)&smem[threadIdx.x];
    st.lds_stride = BS;
    st.lds_entries = LDS;
    st.ovf = (glb_int*)(ovf + gtid());
    st.ovf_stride = FF_TRAV_THREADS;
    uint32_t cn = 0, ct = 0, id = 0, aux = 0, c_rays = 0, c_gated = 0;
    bool busy = false, more = true, two = false;
    bool walk = false;  // WW >= 2: the lane has records or leaves left to visit
    int pend = 0;       // WW >= 2: postponed leaf (0 = none; leaf refs are negative)
    int pend2 = 0;      // WW >= 3: a second postponed leaf, tested after pend
    Bvh B;
    B.nodes = nullptr; B.nodes4 = nullptr; B.leaf_count = nullptr; B.root = nullptr;
    const float* tris = nullptr;
    if (Src::kUniform) src.scene0(B, tris);
    TravRay R;
    Hit h = no_hit();  // closest hit: h.t here, u, v and slot in the lane's LDS words (hs)
    float lim = prune_lim(h.t);  // pruning limit of the closest hit so far (kept beside it)
    HitSlots hs;  // closest-hit kernels allocate LDS * BS + 3 * BS words of smem
    hs.u = ANY ? nullptr : (lds_float*)&smem[LDS * BS + threadIdx.x];
    hs.v = ANY ? nullptr : (lds_float*)&smem[LDS * BS + BS + threadIdx.x];
    hs.id = ANY ? nullptr : (lds_int*)&smem[LDS * BS + 2 * BS + threadIdx.x];
    // closest-hit sources: the lane's leaf test, the hit reset and the result
    auto leaf_c = [&](const TravRay& L) {
        leaf_closest_slots(B, tris, L, h.t, hs, two, ct);
        lim = prune_lim(h.t);
    };
    auto reset_hit = [&]() {
        h = no_hit();
        lim = prune_lim(h.t);
        if (!ANY) {
            *hs.u = h.u;
            *hs.v = h.v;
            *hs.id = h.id;
        }
    };
    auto hit_out = [&]() {
        Hit r;
        r.t = h.t;
        r.u = *hs.u;
        r.v = *hs.v;
        r.id = *hs.id;
        src.hit(id, r);
    };
    // an any-hit ray is finished with result occ: the result is written and,
    // for a source that chains rays, `nxt` marks that the lane's path has
    // another ray, which the lane starts at the top of the next iteration
    // instead of taking a queue entry
    bool nxt = false;
    auto end_any = [&](bool occ) {
        src.occluded(id, aux, occ);
        nxt = Src::kChain && src.advance(aux, occ, c_gated);
    };
    for (;;) {
        if (Src::kChain && __ballot(nxt) != 0) {
            if (nxt) {
                f3 o, d;
                src.load_next(id, aux, o, d);
                c_rays++;
                busy = trav_begin(B, o, d, R, st, cn);
                walk = busy;
                pend = 0;
                pend2 = 0;
                nxt = false;
                if (!busy) end_any(false);
            }
        }
        uint64_t idle = __ballot(!busy && !nxt);
        uint32_t n_idle = (uint32_t)__popcll(idle);
        if (more && n_idle >= (uint32_t)REFILL) {
            uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            uint32_t got;
            bool done;
            uint32_t first = F.take_chunked(n_idle, got, done);
            if (done) more = false;
            if (!busy && !nxt && rank < got) {
                f3 o, d;
                src.load(first + rank, id, aux, B, tris, o, d, two);
                c_rays++;
                if (!ANY) reset_hit();
                busy = trav_begin(B, o, d, R, st, cn);
                walk = busy;
                pend = 0;
                pend2 = 0;
                if (!busy) {
                    if (ANY) end_any(false);
                    else hit_out();
                }
            }
        }
        if (__ballot(busy || nxt) == 0) {
            if (!more) break;
            continue;
        }
        if (WW == 2) {
            for (;;) {
                if (busy && walk && R.ref < 0 && pend == 0) {  // postpone the first leaf
                    pend = R.ref;
                    if (st.sp == 0) walk = false;
                    else R.ref = st.pop();
                }
                bool at_int = busy && walk && R.ref >= 0;
                bool seeking = busy && walk && pend == 0;
                if (__ballot(at_int) == 0 || __ballot(seeking) == 0) break;
                if (at_int && !visit_node<WIDE, !ANY>(B, R, lim, st, cn)) {
                    if (st.sp == 0) walk = false;
                    else R.ref = st.pop();
                }
            }
            if (busy) {
                bool occ = false;
                if (pend != 0) {  // the postponed leaf, then the current one
                    TravRay L = R;
                    L.ref = pend;
                    if (ANY) occ = leaf_any(B, tris, L, ct);
                    else leaf_c(L);
                    pend = 0;
                }
                if (!occ && walk && R.ref < 0) {
                    if (ANY) occ = leaf_any(B, tris, R, ct);
                    else leaf_c(R);
                    if (st.sp == 0) walk = false;
                    else R.ref = st.pop();
                }
                if (occ || !walk) {
                    if (ANY) end_any(occ);
                    else hit_out();
                    busy = false;
                }
            }
        } else if (WW >= 3) {
            // leaf batching: a lane standing on a leaf queues it (two at most)
            // and walks on; the queued leaves are tested together once
            // LEAF_T lanes hold one, or when no lane can walk on.  Each lane
            // tests its leaves in its visit order.
            constexpr uint32_t LEAF_T = WW == 7 ? 16u : (WW == 6 ? 4u : 8u);
            constexpr bool TWO = WW == 3;  // shapes 5-7 queue one leaf per lane
            if (busy && walk && R.ref < 0 && (TWO ? pend2 == 0 : pend == 0)) {
                if (pend == 0) pend = R.ref;
                else pend2 = R.ref;
                if (st.sp == 0) walk = false;
                else R.ref = st.pop();
            }
            if (busy && walk && R.ref >= 0 && !visit_node<WIDE, !ANY>(B, R, lim, st, cn)) {
                if (st.sp == 0) walk = false;
                else R.ref = st.pop();
            }
            bool has = busy && pend != 0;
            uint32_t n_pend = (uint32_t)__popcll(__ballot(has));
            bool occ = false;
            if (n_pend >= LEAF_T || (n_pend > 0 && __ballot(busy && walk && R.ref >= 0) == 0)) {
                if (has) {
                    TravRay L = R;
                    L.ref = pend;
                    if (ANY) occ = leaf_any(B, tris, L, ct);
                    else leaf_c(L);
                    pend = TWO ? pend2 : 0;
                    pend2 = 0;
                }
            }
            if (busy && (occ || (!walk && pend == 0))) {
                if (ANY) end_any(occ);
                else hit_out();
                busy = false;
            }
        } else if (WW == 1) {
            for (;;) {  // interior records until every busy lane is at a leaf
                bool at_int = busy && R.ref >= 0;
                if (__ballot(at_int) == 0) break;
                if (at_int && !visit_node<WIDE, !ANY>(B, R, lim, st, cn)) {
                    if (st.sp == 0) {
                        if (ANY) end_any(false);
                        else hit_out();
                        busy = false;
                    } else {
                        R.ref = st.pop();
                    }
                }
            }
            if (busy) {  // at a leaf
                if (ANY) {
                    bool occ = leaf_any(B, tris, R, ct);
                    if (occ || st.sp == 0) {
                        end_any(occ);
                        busy = false;
                    } else {
                        R.ref = st.pop();
                    }
                } else {
                    leaf_c(R);
                    if (st.sp == 0) {
                        hit_out();
                        busy = false;
                    } else {
                        R.ref = st.pop();
                    }
                }
            }
        } else if (busy) {
            if (ANY) {
                int r = any_step<WIDE>(B, tris, R, st, cn, ct);
                if (r != 0) {
                    end_any(r == 1);
                    busy = false;
                }
            } else {
                // closest_step with the hit's u, v and slot in LDS
                bool cont = true;
                if (R.ref >= 0 && visit_node<WIDE, true>(B, R, lim, st, cn)) {
                } else {
                    if (R.ref < 0) leaf_c(R);
                    if (st.sp == 0) cont = false;
                    else R.ref = st.pop();
                }
                if (!cont) {
                    hit_out();
                    busy = false;
                }
            }
        }
    }
    if (cnt) {
        wave_sum_add(&cnt[0], cn);
        wave_sum_add(&cnt[1], ct);
    }
    src.count(c_rays, c_gated);
}

// The one-scene extend kernels of the one-step and leaf-batching shapes run
// 8 waves per SIMD (64 VGPRs, a few bytes of spills; the hit's u, v and slot
// live in LDS) where their LDS stack allows that many; the others keep the
// compiler's choice.
template <int BS, int LDS, int RF, int WW, bool WIDE, bool UNI>
__global__ __launch_bounds__(BS, (UNI && (WW == 0 || WW >= 5) && !WIDE && LDS <= 16 && BS >= 128) ? 8 : 1) void k_extend(
    ff_paths P, int q) {
    __shared__ int smem[LDS * BS + 3 * BS];  // stack, then the hit words
    ExtSrc<UNI> src{P, P.q_ext[q]};
    trav_persistent<BS, LDS, RF, false, WW, WIDE>(src, P.counters[q == 0 ? C_EXT0 : C_EXT1],
                                            P.fetch + F_EXT * FETCH_SEGS * FETCH_STRIDE, P.ovf,
                                            stats_rep(P) + S_EXT_NODES, smem);
}
template <int BS, int LDS, int RF, int WW, bool WIDE, bool UNI>
__global__ __launch_bounds__(BS) void k_shadow(ff_paths P) {
    __shared__ int smem[LDS * BS];
    ShadowSrc<UNI> src{P};
    trav_persistent<BS, LDS, RF, true, WW, WIDE>(src, P.counters[C_SHADOW], P.fetch + F_SHD * FETCH_SEGS * FETCH_STRIDE,
                                           P.ovf_shd, stats_rep(P) + S_SHD_NODES, smem);
}

// counters consumed by the previous stages are cleared by a kernel (graph-safe):
// queue counters by `mask`, traversal work counters by kind (`fetch_mask`)
__global__ void k_reset_counters(uint32_t* c, uint32_t mask, uint32_t* fetch, uint32_t fetch_mask) {
    uint32_t i = threadIdx.x;
    if (c && i < C_COUNT && ((mask >> i) & 1u)) c[i] = 0;
    if (fetch && i < 3 * FETCH_SEGS && ((fetch_mask >> (i / FETCH_SEGS)) & 1u)) fetch[(size_t)i * FETCH_STRIDE] = 0;
}

// =================================================================== logic
// wf_logic (CRT/kernel.cu:297-455): NEE accumulation gated by the light
// ray's visibility, throughput update, Russian roulette for len > 3,
// environment radiance for escaped rays after the camera (or a delta event),
// termination into the film, routing of live paths to material queues.
// Each thread takes two paths of a 512-entry tile (entries t and t + 256, so
// every load of a wave stays coalesced), and the loads are issued level by
// level for both before any result is used: queue -> path record -> what the
// record asks for (material word, direction, visibilities with the NEE terms
// they gate, throughput factor, RNG base) -> the rare third level (escaped
// delta rays' environment lookup, inside-absorption triangle).  The NEE terms
// are read with their visibility bytes rather than after them.  The
// arithmetic of every path is the one-path wf_logic's, operation for
// operation.  In the first batch of a samples-in-flight request every slot is
// written exactly once (each slot holds one path per batch and every path
// terminates within its batch), so the layer is stored without being read.
// A path that survives its first vertex still holds the start record
// (beta = 1, L = 0): iteration 1 does not store it, and iteration 2 (and the
// first shading pass) use the constants instead of reading it.
template <bool UNI, int NP>
__global__ __launch_bounds__(FF_BLOCK, NP == 1 ? 8 : 5) void k_logic(ff_paths P, int q, int iter) {
    const bool primary = iter == 1;
    __shared__ QStage<FF_BLOCK, 4, QSTAGE_CAP> st;
    const int qidx[4] = {C_MAT0, C_MAT0 + 1, C_MAT0 + 2, C_MAT0 + 3};
    uint32_t* const qs[4] = {P.q_mat[0], P.q_mat[1], P.q_mat[2], P.q_mat[3]};
    st.init();
    const uint32_t n = P.counters[q == 0 ? C_EXT0 : C_EXT1];
    const uint32_t* queue = P.q_ext[q];
    const bool first_batch = P.bstate[B_BATCH] == 0;
    constexpr uint32_t TILE = NP * FF_BLOCK;
    const uint32_t n_round = (n + TILE - 1) / TILE * TILE;  // uniform trip count per workgroup
    uint32_t c_prim = 0, c_ext = 0, c_paths = 0;
    for (uint32_t t0 = blockIdx.x * TILE; t0 < n_round; t0 += gridDim.x * TILE) {
        bool live[NP];
        uint32_t id[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            uint32_t i = t0 + (uint32_t)j * FF_BLOCK + threadIdx.x;
            live[j] = i < n;
            id[j] = live[j] ? queue[i] : 0u;
        }
        // level 1: the path record
        float4 bw[NP], lp4[NP], hh[NP], fe[NP];
        uint32_t rq[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            // a camera path's record is the constant start state (not stored)
            bw[j] = make_float4(1.f, 1.f, 1.f, __uint_as_float(ST_PREV_DELTA));
            lp4[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (iter > 2) {
                bw[j] = P.beta[id[j]];
                lp4[j] = P.Lp[id[j]];
            }
            hh[j] = P.hit[id[j]];
            rq[j] = req_of(P, id[j]);
            // after a shading pass the state word travels in fext.w:
            // (f * cos / pdf of the continuation, state | ST_DEAD where f or pdf is zero)
            fe[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!primary) fe[j] = P.fext[id[j]];
        }
        // level 2: what the record asks for
        uint32_t stw[NP], len[NP], mw[NP], rb[NP];
        int slot[NP];
        bool found[NP], need_d[NP];
        f3 d[NP], nL[NP], nS[NP], nD[NP];
        uint8_t v0[NP], v1[NP], v2[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const ff_dev_request& R = P.reqs[rq[j]];
            const ff_dev_scene& S = scene_of<UNI>(P, R);
            stw[j] = __float_as_uint(primary ? bw[j].w : fe[j].w);
            len[j] = (stw[j] & ST_LEN_MASK) + 1;
            slot[j] = __float_as_int(hh[j].w);
            found[j] = live[j] && slot[j] >= 0;
            need_d[j] = live[j] && ((found[j] && (stw[j] & ST_INSIDE)) || (!found[j] && (stw[j] & ST_PREV_DELTA)));
            mw[j] = 0u;
            if (found[j]) mw[j] = ldg(S.mat + slot[j]);
            d[j] = V3(0.f, 0.f, 0.f);
            if (need_d[j]) d[j] = xyz(P.ray[2 * id[j] + 1]);
            bool ext = live[j] && len[j] > 1;
            v0[j] = v1[j] = v2[j] = 0;
            nL[j] = nS[j] = nD[j] = V3(0.f, 0.f, 0.f);
            if (ext && (stw[j] & ST_NEED_L)) {
                v0[j] = P.vis[3 * id[j] + 0];
                nL[j] = xyz(P.nee[3 * id[j]]);
            }
            if (ext && (stw[j] & ST_NEED_S)) {
                v1[j] = P.vis[3 * id[j] + 1];
                nS[j] = xyz(P.nee[3 * id[j] + 1]);
            }
            if (ext && (stw[j] & ST_NEED_D)) {
                v2[j] = P.vis[3 * id[j] + 2];
                nD[j] = xyz(P.nee[3 * id[j] + 2]);
            }
            rb[j] = 0u;
            if (ext && found[j] && len[j] > 3) rb[j] = __float_as_uint(ldg(&P.ray[2 * id[j]].w));
        }
        // the NEE sum of the previous vertex, gated by its visibilities
        f3 lb[NP];
        bool visL[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            uint32_t sw = stw[j];
            visL[j] = live[j] && len[j] > 1 && (sw & ST_NEED_L) && v0[j];
            lb[j] = nL[j];
            if ((sw & ST_NEED_S) && v1[j]) lb[j] = vadd(lb[j], nS[j]);
            if ((sw & ST_NEED_D) && v2[j]) lb[j] = vadd(lb[j], nD[j]);
        }
        // level 3 and the per-path arithmetic, in wf_logic's order: an escaped
        // ray's environment term, then the NEE sum, then (a hit) throughput
        // and roulette
        int route[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            route[j] = -1;
            if (!live[j]) continue;
            const ff_dev_request& R = P.reqs[rq[j]];
            const ff_dev_scene& S = scene_of<UNI>(P, R);
            uint32_t sw = stw[j], ln = len[j];
            f3 beta = xyz(bw[j]);
            f3 Lp = xyz(lp4[j]);
            bool tm = (ln > R.depth) || !found[j];
            int mtype = -1;
            if (found[j]) {
                mtype = (int)(mw[j] >> MAT_TYPE_SHIFT);
                if ((sw & ST_INSIDE) && mtype != MAT_DIELECTRIC) {
                    Tri T = load_tri(S.tris, slot[j]);
                    if (!(vdot(d[j], vcross(T.e1, T.e2)) < 0.f)) tm = true;  // absorbed
                }
                if (visL[j]) Lp = vadd(Lp, vmul(lb[j], beta));
                if (ln > 1) {
                    if (sw & ST_DEAD) {
                        tm = true;
                    } else {
                        beta = vmul(beta, xyz(fe[j]));
                        if (ln > 3) {
                            float qrr = fmaxf(0.05f, 1.f - beta.y);
                            if (rng(rb[j], ln, DIM_RR) < qrr) tm = true;
                        }
                    }
                }
            } else {
                if (sw & ST_PREV_DELTA) Lp = vadd(Lp, vmul(beta, env_L(env_view(env_of<UNI>(P, R)), d[j])));
                if (visL[j]) Lp = vadd(Lp, vmul(lb[j], beta));
            }
            if (tm) {
                uint32_t local = id[j] - R.slot_base, WH = R.W * R.H;
                if (R.K > 1 && first_batch) {
                    stg4(R.layers + local, make_float4(0.f + Lp.x, 0.f + Lp.y, 0.f + Lp.z, 0.f + 1.f));
                } else {
                    float4* acc = R.K > 1 ? &R.layers[local] : &R.film[local % WH];
                    float4 L = ldg4(acc);
                    stg4(acc, make_float4(L.x + Lp.x, L.y + Lp.y, L.z + Lp.z, L.w + 1.f));
                }
                if (R.mode & RQ_STORE_PATHS) {
                    uint32_t sl = P.bstate[B_BATCH] * R.K + local / WH;
                    stg4(R.path_L + (size_t)sl * WH + local % WH, f4(Lp, 0.f));
                }
            } else {
                uint32_t nst = (sw & (ST_INSIDE | ST_PREV_DELTA)) | ln;
                if (!primary) {
                    P.beta[id[j]] = f4(beta, __uint_as_float(nst));
                    P.Lp[id[j]] = f4(Lp, 0.f);
                }
                route[j] = mtype;
            }
            c_prim += ln == 1 ? 1u : 0u;
            c_ext += ln == 1 ? 0u : 1u;
            c_paths += tm ? 1u : 0u;
        }
#pragma unroll
        for (int j = 0; j < NP; j++) st.add(route[j], id[j]);
        st.flush_if_full(P.counters, qidx, qs, NP);  // room for the next tile's NP adds
    }
    st.flush(P.counters, qidx, qs);
    unsigned long long* sr = stats_rep(P);
    wave_sum_add(sr + S_PRIMARY, c_prim);
    wave_sum_add(sr + S_EXTENSION, c_ext);
    wave_sum_add(sr + S_PATHS, c_paths);
}

// ======================================================= material queues
// Every shaded event continues to the next extension stage, so the next
// extension queue is the material queues one after another: material queue
// TYPE starts after the entries of the types below it.  The first workgroup
// adds the queue's length to the extension counter (one atomic per launch).
template <int TYPE>
__device__ __forceinline__ uint32_t ext_base(const ff_paths& P, int next_q, uint32_t n) {
    uint32_t b = 0u;
#pragma unroll
    for (int t = 0; t < TYPE; t++) b += P.counters[C_MAT0 + t];
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(&P.counters[next_q == 0 ? C_EXT0 : C_EXT1], n);
    return b;
}

// Everything the shading of one vertex produces.  With GRAD the partial
// derivatives of its terms are produced too (path-replay adjoint).
struct ShadeOut {
    f3 A, Bs, Cd;       // light-sample term and BRDF-sampled spec / diffuse terms (before visibility)
    f3 wi, wi_s, wi_d;  // their directions
    f3 lo, so, dd_o;    // shadow ray origins
    bool needL, needS, needD;
    f3 fext;
    float pdf_ext, cos_ext;
    f3 next_o, next_d;
    uint32_t nst;       // state of the continuation
    uint32_t base;      // the path's rng base (rides with the continuation ray)
    // GRAD only
    bool a_on, rough_live;
    int at, rt;
    float tu, tv;
    f3 dA_alb, dB_alb, dC_alb, dA_r, dB_r;  // d term_c / d albedo_c, d term_c / d roughness
    f3 kA, kB, kC;                         // d term_c / d Le_c
    f3 dext_alb, dext_r;
};

// Shading of one material class (wf_mat_mix / wf_mat_diffuse / wf_mat_cook,
// CRT/kernel.cu:545-651, lobes of CRT/dMaterial.cu).  The environment light
// sample of wf_logic (CRT/kernel.cu:435-439) is drawn here; the visibility
// rays of BRDF_L's BRDF-sampled MIS terms (CRT/dMaterial.cu:404-448) are
// deferred to the shadow stage with their contributions kept per path.
// Where a vertex's state word comes from (length, inside, previous delta):
#define SW_RECORD 0   // the path record's beta.w (adjoint replay)
#define SW_START 1    // the first vertex of a forward path: the start record's
#define SW_RAY 2      // a later forward vertex: the ray's .w holds the previous
                      // shading word, which wf_logic turns into this vertex's
                      // word by adding one to the length and keeping the flags
template <int TYPE, bool GRAD, bool UNI>
__device__ __forceinline__ void shade_core(const ff_paths& P, uint32_t id, ShadeOut& so, int sw_src = SW_RECORD) {
    const ff_dev_request& R = P.reqs[req_of(P, id)];
    const ff_dev_scene& S = scene_of<UNI>(P, R);
    EnvView E = env_view(env_of<UNI>(P, R));
    float4 hh = P.hit[id];
    int slot = __float_as_int(hh.w);
    float4 r0 = P.ray[2 * id], r1 = P.ray[2 * id + 1];
    f3 o = xyz(r0), d = xyz(r1);
    uint32_t stw;
    if (sw_src == SW_START) {
        stw = ST_PREV_DELTA | 1u;
    } else if (sw_src == SW_RAY) {
        uint32_t w = __float_as_uint(r1.w);
        stw = (w & (ST_INSIDE | ST_PREV_DELTA)) | ((w & ST_LEN_MASK) + 1u);
    } else {
        stw = __float_as_uint(P.beta[id].w);
    }
    uint32_t len = stw & ST_LEN_MASK;
    bool inside = (stw & ST_INSIDE) != 0;
    uint32_t base = __float_as_uint(r0.w);
    so.base = base;
    f3 Pp = vadd(o, vscl(d, hh.x));
    // the slot's 64-B shading record: normals n0 n1 n2, uvs t0 t1 t2, material word
    const float4* A = S.attr + (size_t)slot * 4;
    float4 a0 = ldg4(A), a1 = ldg4(A + 1), a2 = ldg4(A + 2), a3 = ldg4(A + 3);
    float w0 = (1.f - hh.y) - hh.z;
    f3 nrm = vnorm(vadd(vadd(vscl(V3(a0.w, a1.x, a1.y), hh.y), vscl(V3(a1.z, a1.w, a2.x), hh.z)),
                        vscl(V3(a0.x, a0.y, a0.z), w0)));
    float tu = (a2.w * hh.y + a3.y * hh.z) + a2.y * w0;
    float tv = (a3.x * hh.y + a3.z * hh.z) + a2.z * w0;
    // the 64-B material record (64-B aligned) as four 16-B loads
    const float4* M = (const float4*)(S.mats + (size_t)(__float_as_uint(a3.w) & MAT_INDEX_MASK) * FF_MAT_FLOATS_DEV);
    float4 m0 = ldg4(M), m1 = ldg4(M + 1), m2 = ldg4(M + 2), m3 = ldg4(M + 3);
    f3 alb = V3(m0.y, m0.z, m0.w);
    float rough = m1.x, metal = m1.y;
    f3 fres = V3(m1.z, m1.w, m2.x);
    int at = (int)m2.z, rt = (int)m2.w, mt = (int)m3.x;
    bool rough_live = false;
    if (at >= 0) {
        ff_tex_desc td = S.tex[at];
        float t3[3];
        bilinear(td.data, td.w, td.h, 3, tu, tv, t3);
        alb = ld3(t3);
    }
    if (rt >= 0) {
        ff_tex_desc td = S.tex[rt];
        float t1[1];
        bilinear(td.data, td.w, td.h, 1, tu, tv, t1);
        rough = fmaxf(t1[0], 0.01f);
        rough_live = t1[0] > 0.01f;
    }
    if (mt >= 0) {
        ff_tex_desc td = S.tex[mt];
        float t1[1];
        bilinear(td.data, td.w, td.h, 1, tu, tv, t1);
        metal = t1[0];
    }
    // the light shadow ray leaves along the interpolated normal (CRT/kernel.cu:437);
    // everything else uses the normal-mapped one (get_normal, CRT/dMaterial.cu:61-87)
    f3 n_geo = nrm;
    int nt = (int)m3.y;
    if (nt >= 0) {
        const float* T = S.tan + (size_t)slot * 18;
        f3 tg = vnorm(vadd(vadd(vscl(ld3(T + 6), hh.y), vscl(ld3(T + 12), hh.z)), vscl(ld3(T), w0)));
        f3 bt = vnorm(vadd(vadd(vscl(ld3(T + 9), hh.y), vscl(ld3(T + 15), hh.z)), vscl(ld3(T + 3), w0)));
        ff_tex_desc td = S.tex[nt];
        float t3[3];
        bilinear(td.data, td.w, td.h, 3, tu, tv, t3);
        float mx = t3[0] * 2.f - 1.f, my = t3[1] * 2.f - 1.f, mz = t3[2] * 2.f - 1.f;
        nrm = vnorm(vadd(vadd(vscl(tg, mx), vscl(bt, my)), vscl(nrm, mz)));
    }
    f3 wo = vneg(d);
    f3 zero = V3(0, 0, 0);
    so.A = zero; so.Bs = zero; so.Cd = zero;
    so.wi = zero; so.wi_s = zero; so.wi_d = zero;
    so.lo = Pp; so.so = Pp; so.dd_o = Pp;
    so.needL = so.needS = so.needD = false;
    if (GRAD) {
        so.a_on = false;
        so.rough_live = rough_live;
        so.at = at; so.rt = rt; so.tu = tu; so.tv = tv;
        so.dA_alb = zero; so.dB_alb = zero; so.dC_alb = zero; so.dA_r = zero; so.dB_r = zero;
        so.kA = zero; so.kB = zero; so.kC = zero; so.dext_alb = zero; so.dext_r = zero;
    }
    uint32_t nst = (stw & (ST_INSIDE | ST_PREV_DELTA | ST_LEN_MASK));
    if (TYPE == MAT_DIELECTRIC) {
        Tri T = load_tri(S.tris, slot);
        f3 gn = vnorm(vcross(T.e1, T.e2));
        float cs = vdot(d, gn);
        bool entering = cs < 0.f;
        f3 nn = entering ? gn : vneg(gn);
        float cos_i = fabsf(cs);
        if (cos_i > 1.f) cos_i = 1.f;
        float ior = m2.y;
        float eta_i = entering ? 1.f : ior, eta_t = entering ? ior : 1.f;
        float F = fresnel_dielectric(cos_i, eta_i, eta_t);
        if (rng(base, len, DIM_DIEL) < F) {
            so.next_d = vnorm(vadd(d, vscl(nn, 2.f * cos_i)));
            so.next_o = vadd(Pp, vscl(nn, 1e-4f));
        } else {
            float eta = eta_i / eta_t;
            float k = 1.f - eta * eta * (1.f - cos_i * cos_i);
            float cos_t = sqrtf(fmaxf(k, 0.f));
            so.next_d = vnorm(vadd(vscl(d, eta), vscl(nn, eta * cos_i - cos_t)));
            so.next_o = vsub(Pp, vscl(nn, 1e-4f));
            inside = !inside;
        }
        so.fext = V3(1.f, 1.f, 1.f);
        so.pdf_ext = 1.f;
        so.cos_ext = 1.f;
        so.nst = (nst & ST_LEN_MASK) | ST_PREV_DELTA | (inside ? ST_INSIDE : 0u);
        return;
    }
    constexpr bool do_spec = (TYPE == MAT_MIX || TYPE == MAT_CONDUCTOR);
    constexpr bool do_diff = (TYPE == MAT_MIX || TYPE == MAT_DIFFUSE);
    f3 wi = env_sample(E, Pp, rng(base, len, DIM_EX), rng(base, len, DIM_EY));
    so.wi = wi;
    double diff_pdf = (double)(1.f / (2.f * F_PI));
    // light-sample term of BRDF_L (CRT/dMaterial.cu:376-402)
    f3 Li = env_L(E, wi);
    float ndwi = vdot(nrm, wi);
    if (ndwi > 0.f) {
        double sp = do_spec ? (double)spec_pdf(nrm, wi, wo, rough) : 0.0;
        double lp = (double)env_pdf(E, wi);
        // lp is infinite where sin(acos(wi.z)) == 0; the term's limit there is 0
        if (lp > 0.f && lp < (double)INFINITY && !vzero(Li)) {
            double dwt = power_heur(lp, diff_pdf), swt = power_heur(lp, sp);
            if (do_spec) {
                f3 f = vscl(spec_f(nrm, wi, wo, rough, metal, alb, fres), ndwi);
                if (!vzero(f)) {
                    so.A = vadd(so.A, vdiv(vmul(vscl(f, (float)swt), Li), (float)lp));
                    if (GRAD) {
                        LobeGrad g = spec_grad(nrm, wi, wo, rough, metal, alb, fres);
                        float s = ndwi * (float)swt / (float)lp;
                        so.a_on = true;
                        so.kA = vadd(so.kA, vscl(f, (float)swt / (float)lp));
                        so.dA_alb = vadd(so.dA_alb, vmul(vscl(g.da, s), Li));
                        so.dA_r = vadd(so.dA_r, vmul(vscl(g.dr, s), Li));
                    }
                }
            }
            if (do_diff) {
                f3 f = vscl(diff_f(wi, wo, metal, alb, fres), ndwi);
                if (!vzero(f)) {
                    so.A = vadd(so.A, vdiv(vmul(vscl(f, (float)dwt), Li), (float)lp));
                    if (GRAD) {
                        f3 da = diff_grad(wi, wo, metal, alb, fres);
                        float s = ndwi * (float)dwt / (float)lp;
                        so.a_on = true;
                        so.kA = vadd(so.kA, vscl(f, (float)dwt / (float)lp));
                        so.dA_alb = vadd(so.dA_alb, vmul(vscl(da, s), Li));
                    }
                }
            }
        }
    }
    // BRDF-sampled MIS terms (CRT/dMaterial.cu:404-448)
    f3 wi_s = zero, wi_d = zero;
    if (do_spec) wi_s = spec_sample(nrm, wo, rough, rng(base, len, DIM_S0), rng(base, len, DIM_S1));
    if (do_diff) wi_d = diff_sample(nrm, rng(base, len, DIM_D0), rng(base, len, DIM_D1));
    so.wi_s = wi_s;
    so.wi_d = wi_d;
    if (do_spec) {
        float nd = vdot(nrm, wi_s);
        if (nd > 0.f) {
            double sp = (double)spec_pdf(nrm, wi_s, wo, rough);
            double lp = (double)env_pdf(E, wi_s);
            double wt = power_heur(sp, lp);
            f3 f = vscl(spec_f(nrm, wi_s, wo, rough, metal, alb, fres), nd);
            f3 L2 = env_L(E, wi_s);
            if (!vzero(f) && sp > 0.f && !vzero(L2)) {
                so.Bs = vdiv(vscl(vmul(f, L2), (float)wt), (float)sp);
                so.needS = true;
                so.so = vadd(Pp, vscl(wi_s, 0.001f));
                if (GRAD) {
                    LobeGrad g = spec_grad(nrm, wi_s, wo, rough, metal, alb, fres);
                    float s = nd * (float)wt / (float)sp;
                    so.kB = vscl(f, (float)wt / (float)sp);
                    so.dB_alb = vmul(vscl(g.da, s), L2);
                    so.dB_r = vmul(vscl(g.dr, s), L2);
                }
            }
        }
    }
    if (do_diff) {
        float nd = vdot(nrm, wi_d);
        if (nd > 0.f) {
            double lp = (double)env_pdf(E, wi_d);
            double wt = power_heur(diff_pdf, lp);
            f3 f = vscl(diff_f(wi_d, wo, metal, alb, fres), nd);
            f3 L3 = env_L(E, wi_d);
            if (!vzero(f) && diff_pdf > 0.f && !vzero(L3)) {
                so.Cd = vdiv(vscl(vmul(f, L3), (float)wt), (float)diff_pdf);
                so.needD = true;
                // the mixed material offsets this ray along wi_spec (CRT/dMaterial.cu:431)
                so.dd_o = vadd(Pp, vscl(TYPE == MAT_MIX ? wi_s : wi_d, 0.001f));
                if (GRAD) {
                    f3 da = diff_grad(wi_d, wo, metal, alb, fres);
                    float s = nd * (float)wt / (float)diff_pdf;
                    so.kC = vscl(f, (float)wt / (float)diff_pdf);
                    so.dC_alb = vmul(vscl(da, s), L3);
                }
            }
        }
    }
    so.needL = !vzero(so.A) || so.needS || so.needD;
    so.lo = vadd(Pp, vscl(n_geo, 0.001f));
    f3 dir;
    if (TYPE == MAT_MIX) {
        dir = (rng(base, len, DIM_LOBE) < 0.5f) ? wi_d : wi_s;
        so.fext = vadd(spec_f(nrm, dir, wo, rough, metal, alb, fres), diff_f(dir, wo, metal, alb, fres));
        float spf = spec_pdf(nrm, dir, wo, rough);
        float dpf = (float)diff_pdf;
        so.pdf_ext = (float)(0.5 * (double)(spf + dpf));
    } else if (TYPE == MAT_DIFFUSE) {
        dir = wi_d;
        so.fext = diff_f(dir, wo, metal, alb, fres);
        so.pdf_ext = (float)diff_pdf;
    } else {
        dir = wi_s;
        so.fext = spec_f(nrm, dir, wo, rough, metal, alb, fres);
        so.pdf_ext = spec_pdf(nrm, dir, wo, rough);
    }
    if (GRAD) {
        if (do_spec) {
            LobeGrad g = spec_grad(nrm, dir, wo, rough, metal, alb, fres);
            so.dext_alb = g.da;
            so.dext_r = g.dr;
        }
        if (do_diff) so.dext_alb = vadd(so.dext_alb, diff_grad(dir, wo, metal, alb, fres));
    }
    so.cos_ext = fmaxf(0.0f, vdot(nrm, dir));
    so.next_o = Pp;
    so.next_d = dir;
    so.nst = (nst & (ST_LEN_MASK | ST_INSIDE)) | (so.needL ? ST_NEED_L : 0u) | (so.needS ? ST_NEED_S : 0u) |
             (so.needD ? ST_NEED_D : 0u);
}

template <int TYPE, bool UNI>
__global__ __launch_bounds__(FF_BLOCK) void k_shade(ff_paths P, int next_q, int first) {
    __shared__ QStage<FF_BLOCK, 1, QSTAGE_CAP> st;
    const uint32_t n = P.counters[C_MAT0 + TYPE];
    const uint32_t* queue = P.q_mat[TYPE];
    const uint32_t n_round = (n + FF_BLOCK - 1) / FF_BLOCK * FF_BLOCK;
    const int qidx[1] = {C_SHADOW};
    uint32_t* const qs[1] = {P.q_shadow};
    const uint32_t ebase = ext_base<TYPE>(P, next_q, n);
    st.init();
    for (uint32_t i = gtid(); i < n_round; i += gstride()) {
        bool live = i < n;
        uint32_t id = live ? queue[i] : 0u;
        bool needL = false, needS = false, needD = false;
        if (live) {
            ShadeOut so;
            shade_core<TYPE, false, UNI>(P, id, so, first ? SW_START : SW_RAY);
            needL = so.needL; needS = so.needS; needD = so.needD;
            if (needL) {
                P.sray[6 * id + 0] = f4(so.lo, 0.f);
                P.sray[6 * id + 1] = f4(so.wi, 0.f);
            }
            if (needS) {
                P.sray[6 * id + 2] = f4(so.so, 0.f);
                P.sray[6 * id + 3] = f4(so.wi_s, 0.f);
            }
            if (needD) {
                P.sray[6 * id + 4] = f4(so.dd_o, 0.f);
                P.sray[6 * id + 5] = f4(so.wi_d, 0.f);
            }
            // wf_logic reads a term only where the state word asks for it
            // (needS and needD imply needL)
            if (needL) P.nee[3 * id] = f4(so.A, 0.f);
            if (needS) P.nee[3 * id + 1] = f4(so.Bs, 0.f);
            if (needD) P.nee[3 * id + 2] = f4(so.Cd, 0.f);
            // the throughput factor wf_logic applies, f * cos / pdf, formed here
            // with the same operations (logic then reads 16 B instead of 32)
            // and the continuation's state word rides in its w (the path
            // record's beta is left to wf_logic)
            bool dead = vzero(so.fext) || so.pdf_ext == 0.f;
            uint32_t sw = dead ? so.nst | ST_DEAD : so.nst;
            P.fext[id] = dead ? make_float4(0.f, 0.f, 0.f, __uint_as_float(sw))
                              : f4(vdiv(vscl(so.fext, so.cos_ext), so.pdf_ext), __uint_as_float(sw));
            P.ray[2 * id] = f4(so.next_o, __uint_as_float(so.base));
            P.ray[2 * id + 1] = f4(so.next_d, __uint_as_float(sw));
        }
        // every event continues (wf_logic decides its fate): the material
        // queues follow one another in the next extension queue.  One
        // shadow-queue entry per event with a light ray (needS / needD imply
        // needL); the shadow stage traces and counts its rays.
        if (live) P.q_ext[next_q][ebase + i] = id;
        st.add(needL ? 0 : -1, (id << 2) | (needS ? 1u : 0u) | (needD ? 2u : 0u));
        st.flush_if_full(P.counters, qidx, qs);
    }
    st.flush(P.counters, qidx, qs);
}

// ================================================== path-replay adjoint
// The adjoint replays every stored forward path with the same random
// numbers (same camera ray, same hits, same samples) and, at each event,
// moves d loss / d radiance onto the parameters the event's radiance depends
// on: Le through environment texels, f through albedo / roughness texels.
// L_rem (in Lp) starts as the path's stored radiance and loses each term as
// it is replayed, so at a vertex it is the radiance that flows through the
// vertex's continuation; its derivative w.r.t. f is L_rem * (df/dtheta) / f
// (path replay backpropagation).  Sampling is detached: directions, pdfs,
// MIS weights and Russian roulette are those of the forward.

// One term of d loss / d parameter into gradient target element i.
__device__ __forceinline__ void grad_add(const ff_paths& P, const ff_dev_gtarget& T, uint32_t i, float v) {
    if (v == 0.f) return;
    if (P.gmode == GM_FLOAT) {
        unsafeAtomicAdd(T.buf + i, v);
    } else if (P.gmode == GM_BOUND) {
        uint64_t a = T.off + i;
        atomicMax(&P.g_max[a], __float_as_uint(fabsf(v)));
        atomicAdd(&P.g_cnt[a], 1u);
    } else {
        uint64_t a = T.off + i;
        int e = fixed_exponent(P.g_max[a], P.g_cnt[a]);
        long long qv = __double2ll_rn(ldexp((double)v, e));
        atomicAdd(&P.g_fx[a], (unsigned long long)qv);
    }
}
// d loss / d Le(dir) = val (per channel) onto the environment texels
__device__ __forceinline__ void grad_env(const ff_paths& P, const ff_dev_request& R, f3 dir, f3 val) {
    const ff_dev_gtarget& T = R.gt[GT_ENV];
    if (!T.buf) return;
    float u, v;
    sph_map(dir, u, v);
    Taps tp = taps_of(R.E.w, R.E.h, u, v);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        grad_add(P, T, 3 * tp.i[k] + 0, tp.w[k] * val.x);
        grad_add(P, T, 3 * tp.i[k] + 1, tp.w[k] * val.y);
        grad_add(P, T, 3 * tp.i[k] + 2, tp.w[k] * val.z);
    }
}
// d loss / d (texture t value at (u, v)) = val[0..c) onto its texels
template <int C>
__device__ __forceinline__ void grad_tex(const ff_paths& P, const ff_dev_request& R, int t, float u, float v,
                                         const float* val) {
    const ff_dev_gtarget& T = R.gt[1 + t];
    if (!T.buf) return;
    const ff_tex_desc& td = R.S.tex[t];
    Taps tp = taps_of(td.w, td.h, u, v);
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int c = 0; c < C; c++) grad_add(P, T, C * tp.i[k] + c, tp.w[k] * val[c]);
}

// wf_logic of the replay: same termination, Russian roulette and routing as
// k_logic; an escaping delta/camera ray's Le term is replayed.  The vertex's
// NEE terms and throughput update were replayed by k_shade_adj.
template <bool UNI>
__global__ __launch_bounds__(FF_BLOCK) void k_logic_adj(ff_paths P, int q) {
    __shared__ QStage<FF_BLOCK, 4, QSTAGE_CAP> st;
    const int qidx[4] = {C_MAT0, C_MAT0 + 1, C_MAT0 + 2, C_MAT0 + 3};
    uint32_t* const qs[4] = {P.q_mat[0], P.q_mat[1], P.q_mat[2], P.q_mat[3]};
    st.init();
    const uint32_t n = P.counters[q == 0 ? C_EXT0 : C_EXT1];
    const uint32_t* queue = P.q_ext[q];
    const uint32_t n_round = (n + FF_BLOCK - 1) / FF_BLOCK * FF_BLOCK;  // uniform trip count per workgroup
    uint32_t c_prim = 0, c_ext = 0, c_paths = 0;
    for (uint32_t i = gtid(); i < n_round; i += gstride()) {
        bool live = i < n;
        int route = -1;
        bool term = false, primary = false;
        uint32_t id = live ? queue[i] : 0u;
        if (live) {
            const ff_dev_request& R = P.reqs[req_of(P, id)];
            const ff_dev_scene& S = scene_of<UNI>(P, R);
            float4 bw = P.beta[id];
            uint32_t stw = __float_as_uint(bw.w);
            uint32_t len = (stw & ST_LEN_MASK) + 1;
            primary = len == 1;
            f3 beta = xyz(bw);
            f3 Lrem = xyz(P.Lp[id]);
            float4 hh = P.hit[id];
            int slot = __float_as_int(hh.w);
            bool found = slot >= 0;
            term = (len > R.depth) || !found || (stw & ST_DEAD);
            int mtype = -1;
            // the direction is read only by the two cases that use it
            f3 d = V3(0.f, 0.f, 0.f);
            if ((found && (stw & ST_INSIDE)) || (!found && (stw & ST_PREV_DELTA))) d = xyz(P.ray[2 * id + 1]);
            if (found) {
                mtype = (int)(S.mat[slot] >> MAT_TYPE_SHIFT);
                if ((stw & ST_INSIDE) && mtype != MAT_DIELECTRIC) {
                    Tri T = load_tri(S.tris, slot);
                    if (!(vdot(d, vcross(T.e1, T.e2)) < 0.f)) term = true;  // absorbed
                }
            }
            if (!found && (stw & ST_PREV_DELTA)) {
                Lrem = vsub(Lrem, vmul(beta, env_L(env_view(env_of<UNI>(P, R)), d)));
                grad_env(P, R, d, vmul(xyz(P.dlt[id]), beta));
            }
            if (!(stw & ST_DEAD) && len > 3) {
                float qrr = fmaxf(0.05f, 1.f - beta.y);
                if (rng(__float_as_uint(ldg(&P.ray[2 * id].w)), len, DIM_RR) < qrr) term = true;
            }
            if (!term) {
                uint32_t nst = (stw & (ST_INSIDE | ST_PREV_DELTA)) | len;
                P.beta[id] = f4(beta, __uint_as_float(nst));
                P.Lp[id] = f4(Lrem, 0.f);
                route = mtype;
            }
        }
        c_prim += (live && primary) ? 1u : 0u;
        c_ext += (live && !primary) ? 1u : 0u;
        c_paths += (live && term) ? 1u : 0u;
        st.add(route, id);
        st.flush_if_full(P.counters, qidx, qs);
    }
    st.flush(P.counters, qidx, qs);
    unsigned long long* sr = stats_rep(P);
    wave_sum_add(sr + S_PRIMARY, c_prim);
    wave_sum_add(sr + S_EXTENSION, c_ext);
    wave_sum_add(sr + S_PATHS, c_paths);
}

// Replay of a vertex's shadow rays: the same rays k_shade queues.
template <int TYPE, bool UNI>
__global__ __launch_bounds__(FF_BLOCK) void k_shade_nee(ff_paths P) {
    __shared__ QStage<FF_BLOCK, 1, QSTAGE_CAP> st;
    const uint32_t n = P.counters[C_MAT0 + TYPE];
    const uint32_t* queue = P.q_mat[TYPE];
    const uint32_t n_round = (n + FF_BLOCK - 1) / FF_BLOCK * FF_BLOCK;
    const int qidx[1] = {C_SHADOW};
    uint32_t* const qs[1] = {P.q_shadow};
    st.init();
    for (uint32_t i = gtid(); i < n_round; i += gstride()) {
        bool live = i < n;
        uint32_t id = live ? queue[i] : 0u;
        bool needL = false, needS = false, needD = false;
        if (live) {
            ShadeOut so;
            shade_core<TYPE, false, UNI>(P, id, so);
            needL = so.needL; needS = so.needS; needD = so.needD;
            if (needL) {
                P.sray[6 * id + 0] = f4(so.lo, 0.f);
                P.sray[6 * id + 1] = f4(so.wi, 0.f);
            }
            if (needS) {
                P.sray[6 * id + 2] = f4(so.so, 0.f);
                P.sray[6 * id + 3] = f4(so.wi_s, 0.f);
            }
            if (needD) {
                P.sray[6 * id + 4] = f4(so.dd_o, 0.f);
                P.sray[6 * id + 5] = f4(so.wi_d, 0.f);
            }
        }
        st.add(needL ? 0 : -1, (id << 2) | (needS ? 1u : 0u) | (needD ? 2u : 0u));
        st.flush_if_full(P.counters, qidx, qs);
    }
    st.flush(P.counters, qidx, qs);
}

// Replay of a vertex once its shadow rays are traced: its visibility-gated
// NEE terms leave L_rem (with their gradients), then the continuation's
// gradient L_rem * (df/dtheta) / f, the throughput update of wf_logic and the
// next extension ray.
template <int TYPE, bool UNI>
__global__ __launch_bounds__(FF_BLOCK) void k_shade_adj(ff_paths P, int next_q) {
    const uint32_t n = P.counters[C_MAT0 + TYPE];
    const uint32_t* queue = P.q_mat[TYPE];
    const uint32_t n_round = (n + FF_BLOCK - 1) / FF_BLOCK * FF_BLOCK;
    const uint32_t ebase = ext_base<TYPE>(P, next_q, n);
    for (uint32_t i = gtid(); i < n_round; i += gstride()) {
        bool live = i < n;
        uint32_t id = live ? queue[i] : 0u;
        if (live) {
            const ff_dev_request& R = P.reqs[req_of(P, id)];
            ShadeOut so;
            shade_core<TYPE, TYPE != MAT_DIELECTRIC, UNI>(P, id, so);
            f3 beta = xyz(P.beta[id]);
            uint32_t nst = so.nst;
            if (TYPE != MAT_DIELECTRIC) {
                f3 Lrem = xyz(P.Lp[id]);
                f3 dl = xyz(P.dlt[id]);
                bool visL = so.needL && P.vis[3 * id + 0];
                bool visS = visL && so.needS && P.vis[3 * id + 1];
                bool visD = visL && so.needD && P.vis[3 * id + 2];
                if (visL) {
                    f3 lb = so.A;
                    if (visS) lb = vadd(lb, so.Bs);
                    if (visD) lb = vadd(lb, so.Cd);
                    Lrem = vsub(Lrem, vmul(lb, beta));
                    f3 w = vmul(dl, beta);
                    if (so.at >= 0) {
                        f3 ga = so.dA_alb;
                        if (visS) ga = vadd(ga, so.dB_alb);
                        if (visD) ga = vadd(ga, so.dC_alb);
                        float g3[3] = {w.x * ga.x, w.y * ga.y, w.z * ga.z};
                        grad_tex<3>(P, R, so.at, so.tu, so.tv, g3);
                    }
                    if (so.rt >= 0 && so.rough_live) {
                        f3 gr = so.dA_r;
                        if (visS) gr = vadd(gr, so.dB_r);
                        float g1 = vdot(w, gr);
                        grad_tex<1>(P, R, so.rt, so.tu, so.tv, &g1);
                    }
                    if (so.a_on) grad_env(P, R, so.wi, vmul(w, so.kA));
                    if (visS) grad_env(P, R, so.wi_s, vmul(w, so.kB));
                    if (visD) grad_env(P, R, so.wi_d, vmul(w, so.kC));
                }
                bool dead = vzero(so.fext) || so.pdf_ext == 0.f;
                if (!dead) {
                    f3 qv = V3(so.fext.x != 0.f ? dl.x * Lrem.x / so.fext.x : 0.f,
                               so.fext.y != 0.f ? dl.y * Lrem.y / so.fext.y : 0.f,
                               so.fext.z != 0.f ? dl.z * Lrem.z / so.fext.z : 0.f);
                    if (so.at >= 0) {
                        float g3[3] = {qv.x * so.dext_alb.x, qv.y * so.dext_alb.y, qv.z * so.dext_alb.z};
                        grad_tex<3>(P, R, so.at, so.tu, so.tv, g3);
                    }
                    if (so.rt >= 0 && so.rough_live) {
                        float g1 = vdot(qv, so.dext_r);
                        grad_tex<1>(P, R, so.rt, so.tu, so.tv, &g1);
                    }
                    beta = vmul(beta, vdiv(vscl(so.fext, so.cos_ext), so.pdf_ext));
                }
                nst = (nst & (ST_LEN_MASK | ST_INSIDE)) | (dead ? ST_DEAD : 0u);
                P.Lp[id] = f4(Lrem, 0.f);
            }
            P.beta[id] = f4(beta, __uint_as_float(nst));
            P.ray[2 * id] = f4(so.next_o, __uint_as_float(so.base));
            P.ray[2 * id + 1] = f4(so.next_d, __uint_as_float(nst));
        }
        if (live) P.q_ext[next_q][ebase + i] = id;
    }
}

// Adjoint-mode converter: fixed-point sums -> caller's float gradients.
__global__ void k_fixed_to_float(const unsigned long long* fx, const uint32_t* mx, const uint32_t* cnt,
                                 uint64_t n, float* out) {
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        long long v = (long long)fx[i];
        if (v == 0) continue;
        int e = fixed_exponent(mx[i], cnt[i]);
        out[i] += (float)ldexp((double)v, -e);
    }
}

// ============================================================ film layers
__global__ void k_fill4(float4* p, uint32_t n) {
    for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void k_fill_u32(uint32_t* p, uint64_t n) {
    for (uint64_t i = gtid(); i < n; i += gstride()) p[i] = 0u;
}
__global__ void k_fill_u64(unsigned long long* p, uint32_t n) {
    for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = 0ull;
}
// film[pixel] += sum over layers k (in k order) of layers[k*WH + pixel]
__global__ void k_resolve_layers(const float4* layers, uint32_t K, uint32_t WH, float4* film) {
    for (uint32_t p = gtid(); p < WH; p += gstride()) {
        float4 s = layers[p];
        for (uint32_t k = 1; k < K; k++) {
            float4 l = layers[(size_t)k * WH + p];
            s.x = s.x + l.x; s.y = s.y + l.y; s.z = s.z + l.z; s.w = s.w + l.w;
        }
        float4 f = film[p];
        film[p] = make_float4(f.x + s.x, f.y + s.y, f.z + s.z, f.w + s.w);
    }
}

// ======================================================== standalone stages
struct TraceArgs {  // one ff_trace_closest / ff_trace_any call
    ff_dev_scene S;
    const float4* rays;
    uint32_t n;
    float4* hits;
    uint8_t* occ;
    int two_sided;
    uint32_t* fetch;
    int32_t* ovf;
    unsigned long long* counts;
};
template <int BS, int LDS, int RF, int WW, bool WIDE>
__global__ __launch_bounds__(BS, (WW == 0 && !WIDE && LDS <= 16 && BS >= 128) ? 7 : 1) void k_trace_closest(
    TraceArgs a) {
    __shared__ int smem[LDS * BS + 3 * BS];  // stack, then the hit words
    RaySrc src{a.S, a.rays, a.hits, nullptr, a.two_sided};
    trav_persistent<BS, LDS, RF, false, WW, WIDE>(src, a.n, a.fetch, a.ovf, a.counts, smem);
}
template <int BS, int LDS, int RF, int WW, bool WIDE>
__global__ __launch_bounds__(BS) void k_trace_any(TraceArgs a) {
    __shared__ int smem[LDS * BS];
    RaySrc src{a.S, a.rays, nullptr, a.occ, 0};
    trav_persistent<BS, LDS, RF, true, WW, WIDE>(src, a.n, a.fetch, a.ovf, a.counts, smem);
}
__global__ void k_gen_rays(ff_dev_camera cam, int32_t W, int32_t H, uint32_t seed, uint32_t sample,
                           float4* rays) {
    uint32_t WH = (uint32_t)W * (uint32_t)H;
    for (uint32_t pix = gtid(); pix < WH; pix += gstride()) {
        f3 o, d;
        gen_ray(cam, W, H, (int)(pix % (uint32_t)W), (int)(pix / (uint32_t)W), rng_base(seed, pix, sample), o, d);
        rays[2 * pix] = f4(o, 0.f);
        rays[2 * pix + 1] = f4(d, 0.f);
    }
}

// ========================================================= environment CDF
// FP64 CDF build replacing the serial <<<1,1>>> passes of
// g_compute_pdf_denom / g_compute_marginal_dist / g_compute_conditional_dist /
// g_write_pdf_texture (CRT/kernel.cu:65-149).  Same per-texel terms; sums are
// fixed-order tree reductions and blocked scans (deterministic, FP64).
__global__ void k_env_lum(const float* rgb, int32_t w, int32_t h, float* lum) {
    uint32_t n = (uint32_t)w * (uint32_t)h;
    for (uint32_t k = gtid(); k < n; k += gstride()) {
        int i = (int)(k % (uint32_t)w), j = (int)(k / (uint32_t)w);
        float s[3];
        bilinear(rgb, w, h, 3, (float)i / (float)w, (float)j / (float)h, s);
        lum[k] = lum_of(s);
    }
}

__device__ __forceinline__ double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = FF_BLOCK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// one workgroup per row: row_sum[j] = sum_i lum(i,j) * sin_theta_j
__global__ __launch_bounds__(FF_BLOCK) void k_env_rows(const float* lum, int32_t w, int32_t h, double* row_sum) {
    __shared__ double red[FF_BLOCK];
    int j = blockIdx.x;
    float v = (float)j / (float)h;
    double st = (double)f_sin(F_PI * v);
    double acc = 0.0;
    for (int i = threadIdx.x; i < w; i += FF_BLOCK) acc += (double)lum[(size_t)j * w + i] * st;
    double s = block_sum(acc, red);
    if (threadIdx.x == 0) row_sum[j] = s;
}

// single workgroup: denom = sum of the row sums (h values)
__global__ __launch_bounds__(FF_BLOCK) void k_env_denom(const double* row_sum, int32_t h, double* denom_out) {
    __shared__ double red[FF_BLOCK];
    double acc = 0.0;
    for (int j = threadIdx.x; j < h; j += FF_BLOCK) acc += row_sum[j];
    double denom = block_sum(acc, red);
    if (threadIdx.x == 0) *denom_out = denom;
}

// one workgroup per row: marginal_p[y] = sum_x lum * (sin_theta_y / denom)
// (g_compute_marginal_dist's term), then the row's conditional CDF (blocked
// scan of lum * sin_theta_y / (denom * marginal_p[y])) and its pdf texels
__global__ __launch_bounds__(FF_BLOCK) void k_env_conditional(const float* lum, int32_t w, int32_t h,
                                                              const double* denom_p, double* marg_p,
                                                              double* conds, float* pdf) {
    __shared__ double red[FF_BLOCK];
    int y = blockIdx.x;
    double denom = *denom_p;
    float v = (float)y / (float)h;
    double st = (double)f_sin(F_PI * v);
    const float* row = lum + (size_t)y * w;
    double stn = st / denom;
    double mp = 0.0;
    for (int x = threadIdx.x; x < w; x += FF_BLOCK) mp += (double)row[x] * stn;
    mp = block_sum(mp, red);
    if (threadIdx.x == 0) marg_p[y] = mp;
    double val = st / (denom * mp);
    double* crow = conds + (size_t)y * w;
    int c = (w + FF_BLOCK - 1) / FF_BLOCK;
    int lo = threadIdx.x * c, hi = lo + c < w ? lo + c : w;
    double local = 0.0;
    for (int x = lo; x < hi; x++) local += (double)row[x] * val;
    red[threadIdx.x] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int t = 0; t < FF_BLOCK; t++) { double q = red[t]; red[t] = run; run += q; }
    }
    __syncthreads();
    double run = red[threadIdx.x];
    for (int x = lo; x < hi; x++) { run += (double)row[x] * val; crow[x] = run; }
    float scale = (float)(w * h) / (2.f * F_PI * F_PI);
    for (int x = threadIdx.x; x < w; x += FF_BLOCK) {
        float p = (float)(((double)row[x] * st) / denom);
        p *= scale;
        pdf[(size_t)y * w + x] = p;
    }
}

// the render-time texel array: (r, g, b, pdf) per texel
__global__ void k_env_pack(const float* rgb, const float* pdf, uint32_t n, float4* rgbp) {
    for (uint32_t k = gtid(); k < n; k += gstride())
        rgbp[k] = make_float4(rgb[3 * (size_t)k], rgb[3 * (size_t)k + 1], rgb[3 * (size_t)k + 2], pdf[k]);
}

// single workgroup: marginal_y = inclusive scan of marginal_p (thread t
// owns rows [t*c, (t+1)*c), block offsets scanned in order)
__global__ __launch_bounds__(FF_BLOCK) void k_env_marginal_scan(const double* marg_p, int32_t h, double* marg_y) {
    __shared__ double red[FF_BLOCK];
    int c = (h + FF_BLOCK - 1) / FF_BLOCK;
    int lo = threadIdx.x * c, hi = lo + c < h ? lo + c : h;
    double local = 0.0;
    for (int j = lo; j < hi; j++) local += marg_p[j];
    red[threadIdx.x] = local;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int t = 0; t < FF_BLOCK; t++) { double x = red[t]; red[t] = run; run += x; }
    }
    __syncthreads();
    double run = red[threadIdx.x];
    for (int j = lo; j < hi; j++) { run += marg_p[j]; marg_y[j] = run; }
}

// ============================================================ film resolve
// Resolve of wf_logic (CRT/kernel.cu:447-454): average, exposure, Reinhard,
// (int)(255*c) per channel.
__global__ void k_film_resolve(const float4* film, uint32_t n, float exposure, float* avg, uint8_t* rgba) {
    for (uint32_t p = gtid(); p < n; p += gstride()) {
        float4 f = film[p];
        float cnt = f.w;
        float r = 0.f, g = 0.f, b = 0.f;
        if (cnt >= 1.f) { r = f.x / cnt; g = f.y / cnt; b = f.z / cnt; }
        if (avg) { avg[3 * p] = r; avg[3 * p + 1] = g; avg[3 * p + 2] = b; }
        if (rgba) {
            float cr = r * exposure, cg = g * exposure, cb = b * exposure;
            cr = cr / (cr + 1.f); cg = cg / (cg + 1.f); cb = cb / (cb + 1.f);
            rgba[4 * p] = (uint8_t)(int)(255 * cr);
            rgba[4 * p + 1] = (uint8_t)(int)(255 * cg);
            rgba[4 * p + 2] = (uint8_t)(int)(255 * cb);
            rgba[4 * p + 3] = 255;
        }
    }
}

inline uint32_t grid_for(uint64_t n, uint32_t cap) {
    uint64_t g = (n + FF_BLOCK - 1) / FF_BLOCK;
    if (g < 1) g = 1;
    return (uint32_t)(g < cap ? g : cap);
}

// Workgroups of a traversal kernel that the current device keeps resident:
// CUs x blocks per CU at the kernel's occupancy, capped so every lane has its
// overflow stack column.  Cached per (kernel, device).
template <class K>
uint32_t resident_grid(K kernel, int bs) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, uint32_t> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_pair((const void*)kernel, dev);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int nb = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, bs, 0) != hipSuccess || nb < 1) nb = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 1;
    uint64_t g = (uint64_t)nb * (uint64_t)cus, cap = FF_TRAV_THREADS / bs;
    if (g > cap) g = cap;
    cache[key] = (uint32_t)g;
    return (uint32_t)g;
}

// traversal variants (block size, LDS stack entries, refill threshold, loop
// shape 0-4, BVH record width)
#define FF_TRAV_VARIANT_LIST(X)  \
    X(0, 256, 16, 32, 0, false)  \
    X(1, 256, 16, 16, 0, false)  \
    X(2, 256, 16, 64, 0, false)  \
    X(3, 128, 16, 32, 0, false)  \
    X(4, 256, 8, 32, 0, false)   \
    X(5, 256, 24, 32, 0, false)  \
    X(6, 64, 32, 32, 0, false)   \
    X(7, 128, 24, 32, 0, false)  \
    X(8, 256, 16, 32, 1, false)  \
    X(9, 64, 32, 32, 1, false)   \
    X(10, 256, 16, 16, 1, false) \
    X(11, 128, 24, 32, 1, false) \
    X(12, 256, 16, 64, 0, true)  \
    X(13, 256, 16, 32, 0, true)  \
    X(14, 64, 32, 32, 0, true)   \
    X(15, 256, 16, 32, 1, true)  \
    X(16, 256, 16, 64, 2, false) \
    X(17, 256, 16, 32, 2, false) \
    X(18, 256, 16, 8, 0, false)  \
    X(19, 256, 24, 16, 0, false) \
    X(20, 256, 16, 16, 3, false) \
    X(21, 256, 16, 32, 3, false) \
    X(22, 256, 16, 16, 5, false) \
    X(23, 256, 16, 32, 5, false) \
    X(24, 256, 16, 16, 6, false) \
    X(25, 256, 16, 16, 7, false) \
    X(26, 128, 16, 16, 5, false)

// which: 0 extend, 1 shadow, 2 grid sizes only (before a graph capture),
// 3 ff_trace_closest, 4 ff_trace_any
template <int BS, int LDS, int RF, int WW, bool WIDE>
hipError_t trav_launch(int which, const ff_paths* P, int q, const TraceArgs* T, hipStream_t s) {
    if (which == 0 && P->uni_scene) {
        uint32_t g = resident_grid(k_extend<BS, LDS, RF, WW, WIDE, true>, BS);
        hipLaunchKernelGGL((k_extend<BS, LDS, RF, WW, WIDE, true>), dim3(g), dim3(BS), 0, s, *P, q);
    } else if (which == 0) {
        uint32_t g = resident_grid(k_extend<BS, LDS, RF, WW, WIDE, false>, BS);
        hipLaunchKernelGGL((k_extend<BS, LDS, RF, WW, WIDE, false>), dim3(g), dim3(BS), 0, s, *P, q);
    } else if (which == 1 && P->uni_scene) {
        uint32_t g = resident_grid(k_shadow<BS, LDS, RF, WW, WIDE, true>, BS);
        hipLaunchKernelGGL((k_shadow<BS, LDS, RF, WW, WIDE, true>), dim3(g), dim3(BS), 0, s, *P);
    } else if (which == 1) {
        uint32_t g = resident_grid(k_shadow<BS, LDS, RF, WW, WIDE, false>, BS);
        hipLaunchKernelGGL((k_shadow<BS, LDS, RF, WW, WIDE, false>), dim3(g), dim3(BS), 0, s, *P);
    } else if (which == 3) {
        uint32_t g = resident_grid(k_trace_closest<BS, LDS, RF, WW, WIDE>, BS);
        hipLaunchKernelGGL((k_trace_closest<BS, LDS, RF, WW, WIDE>), dim3(g), dim3(BS), 0, s, *T);
    } else if (which == 4) {
        uint32_t g = resident_grid(k_trace_any<BS, LDS, RF, WW, WIDE>, BS);
        hipLaunchKernelGGL((k_trace_any<BS, LDS, RF, WW, WIDE>), dim3(g), dim3(BS), 0, s, *T);
    } else {
        (void)resident_grid(k_extend<BS, LDS, RF, WW, WIDE, true>, BS);
        (void)resident_grid(k_shadow<BS, LDS, RF, WW, WIDE, true>, BS);
        (void)resident_grid(k_extend<BS, LDS, RF, WW, WIDE, false>, BS);
        (void)resident_grid(k_shadow<BS, LDS, RF, WW, WIDE, false>, BS);
    }
    return hipGetLastError();
}

hipError_t trav_dispatch(int which, int variant, const ff_paths* P, int q, const TraceArgs* T, hipStream_t s) {
    switch (variant) {
#define FF_TRAV_CASE(v, bs, lds, rf, ww, wide) \
    case v: return trav_launch<bs, lds, rf, ww, wide>(which, P, q, T, s);
        FF_TRAV_VARIANT_LIST(FF_TRAV_CASE)
#undef FF_TRAV_CASE
        default: return hipErrorInvalidValue;
    }
}

// Guide tables of the environment's inverse-CDF searches (block r < h: row
// r of the conditional CDF, block h: the marginal).  Entry g is the upper
// bound of g / ENV_GUIDE (first element greater, compared in double as the
// sampler compares), so the upper bound of any val in [g, g+1) / ENV_GUIDE
// lies in [entry g, entry g+1].  A list with a NaN or a decreasing step gets
// flagged entries (searched whole).  Every entry also carries list[0] != 0
// and list[1] != 0 (ENV_GUIDE_NZ0 / ENV_GUIDE_NZ1).
__global__ void k_env_guide(const double* marg_y, const double* conds, int32_t w, int32_t h, uint32_t* guide_m,
                            uint32_t* guide_c) {
    const bool marg = (int)blockIdx.x == h;
    const double* list = marg ? marg_y : conds + (size_t)blockIdx.x * w;
    const int size = marg ? h : w;
    uint32_t* out = marg ? guide_m : guide_c + (size_t)blockIdx.x * (ENV_GUIDE + 1);
    int mine = 0;
    for (int i = threadIdx.x; i < size; i += blockDim.x) {
        double v = list[i];
        if (v != v || (i > 0 && v < list[i - 1])) mine = 1;
    }
    const int bad = __syncthreads_or(mine);
    const uint32_t nz = (list[0] != 0.0 ? ENV_GUIDE_NZ0 : 0u) | (list[1] != 0.0 ? ENV_GUIDE_NZ1 : 0u);
    for (int g = threadIdx.x; g <= ENV_GUIDE; g += blockDim.x) {
        float val = (float)g / (float)ENV_GUIDE;
        int left = 0, right = size;
        while (left < right) {
            int middle = (right - left) / 2 + left;
            if ((double)val >= list[middle]) left = middle + 1;
            else right = middle;
        }
        out[g] = (bad ? ENV_GUIDE_FULL : (uint32_t)left) | nz;
    }
}

}  // namespace

// ================================================== fixed trace scratch
// Device memory of ff_trace_closest_noalloc, the entry point for callers
// that cannot allocate (graph capture, serving loops): the traversal work
// counters (FETCH_WORDS words) followed by the stack overflow area
// (FF_STATIC_OVF_ENTRIES entries for each of the FF_TRAV_THREADS lanes a
// persistent traversal grid holds).  Every device the library runs on holds
// its own copy until the process ends; calls on one device share it, so such
// calls must not overlap there.  Host code uses the address from
// hipGetSymbolAddress, never the array's name.
static __device__ __align__(256) int32_t scratch[FETCH_WORDS + (size_t)FF_STATIC_OVF_ENTRIES * FF_TRAV_THREADS];

// ===================================================================== host
namespace ffk {

#define FF_TRAV_ONE(v, bs, lds, rf, ww, wide) +1
const int kTravVariants = 0 FF_TRAV_VARIANT_LIST(FF_TRAV_ONE);
#undef FF_TRAV_ONE
#define FF_TRAV_BS(v, bs, lds, rf, ww, wide) bs,
#define FF_TRAV_LDS(v, bs, lds, rf, ww, wide) lds,
#define FF_TRAV_RF(v, bs, lds, rf, ww, wide) rf,
#define FF_TRAV_WW(v, bs, lds, rf, ww, wide) ww,
#define FF_TRAV_WIDTH(v, bs, lds, rf, ww, wide) (wide ? 4 : 2),
const int kTravBlock[] = {FF_TRAV_VARIANT_LIST(FF_TRAV_BS)};
const int kTravLds[] = {FF_TRAV_VARIANT_LIST(FF_TRAV_LDS)};
const int kTravRefill[] = {FF_TRAV_VARIANT_LIST(FF_TRAV_RF)};
const int kTravWhileWhile[] = {FF_TRAV_VARIANT_LIST(FF_TRAV_WW)};
const int kTravWidth[] = {FF_TRAV_VARIANT_LIST(FF_TRAV_WIDTH)};
#undef FF_TRAV_BS
#undef FF_TRAV_LDS
#undef FF_TRAV_RF
#undef FF_TRAV_WW
#undef FF_TRAV_WIDTH

// every request of the call shades the same scene and environment
static bool uni_shading(const ff_paths& P) { return P.uni_scene && P.uni_env; }

hipError_t launch_batch_begin(const ff_paths& P, int set, uint32_t value, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(256), 0, s, P, set, value);
    return hipGetLastError();
}
hipError_t launch_generate(const ff_paths& P, hipStream_t s) {
    hipLaunchKernelGGL(k_generate, dim3(grid_for(P.n_slots, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, P);
    return hipGetLastError();
}
hipError_t launch_extend(const ff_paths& P, int q, int variant, hipStream_t s) {
    return trav_dispatch(0, variant, &P, q, nullptr, s);
}
hipError_t launch_shadow(const ff_paths& P, int variant, hipStream_t s) {
    return trav_dispatch(1, variant, &P, 0, nullptr, s);
}
hipError_t prepare_traversal(int variant) { return trav_dispatch(2, variant, nullptr, 0, nullptr, nullptr); }
hipError_t launch_reset(uint32_t* counters, uint32_t mask, uint32_t* fetch, uint32_t fetch_mask, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(256), 0, s, counters, mask, fetch, fetch_mask);
    return hipGetLastError();
}
template <bool UNI>
hipError_t logic_launch(const ff_paths& P, int q, int iter, uint32_t n_max, hipStream_t s) {
    hipLaunchKernelGGL((k_logic<UNI, 2>), dim3(grid_for(n_max, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, P, q, iter);
    return hipGetLastError();
}
hipError_t launch_logic(const ff_paths& P, int q, int iter, uint32_t n_max, hipStream_t s) {
    return uni_shading(P) ? logic_launch<true>(P, q, iter, n_max, s) : logic_launch<false>(P, q, iter, n_max, s);
}
template <bool UNI>
hipError_t shade_launch(int type, const ff_paths& P, int next_q, int first, uint32_t n_max, hipStream_t s) {
    dim3 g(grid_for(n_max, FF_GRID_CAP)), b(FF_BLOCK);
    switch (type) {
        case MAT_MIX: hipLaunchKernelGGL((k_shade<MAT_MIX, UNI>), g, b, 0, s, P, next_q, first); break;
        case MAT_DIFFUSE: hipLaunchKernelGGL((k_shade<MAT_DIFFUSE, UNI>), g, b, 0, s, P, next_q, first); break;
        case MAT_CONDUCTOR: hipLaunchKernelGGL((k_shade<MAT_CONDUCTOR, UNI>), g, b, 0, s, P, next_q, first); break;
        default: hipLaunchKernelGGL((k_shade<MAT_DIELECTRIC, UNI>), g, b, 0, s, P, next_q, first); break;
    }
    return hipGetLastError();
}
hipError_t launch_shade(int type, const ff_paths& P, int next_q, int first, uint32_t n_max, hipStream_t s) {
    return uni_shading(P) ? shade_launch<true>(type, P, next_q, first, n_max, s)
                          : shade_launch<false>(type, P, next_q, first, n_max, s);
}
template <bool UNI>
hipError_t logic_adj_launch(const ff_paths& P, int q, uint32_t n_max, hipStream_t s) {
    hipLaunchKernelGGL(k_logic_adj<UNI>, dim3(grid_for(n_max, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, P, q);
    return hipGetLastError();
}
hipError_t launch_logic_adj(const ff_paths& P, int q, uint32_t n_max, hipStream_t s) {
    return uni_shading(P) ? logic_adj_launch<true>(P, q, n_max, s) : logic_adj_launch<false>(P, q, n_max, s);
}
template <bool UNI>
hipError_t shade_nee_launch(int type, const ff_paths& P, uint32_t n_max, hipStream_t s) {
    dim3 g(grid_for(n_max, FF_GRID_CAP)), b(FF_BLOCK);
    switch (type) {
        case MAT_MIX: hipLaunchKernelGGL((k_shade_nee<MAT_MIX, UNI>), g, b, 0, s, P); break;
        case MAT_DIFFUSE: hipLaunchKernelGGL((k_shade_nee<MAT_DIFFUSE, UNI>), g, b, 0, s, P); break;
        case MAT_CONDUCTOR: hipLaunchKernelGGL((k_shade_nee<MAT_CONDUCTOR, UNI>), g, b, 0, s, P); break;
        default: return hipSuccess;  // a dielectric vertex has no NEE rays
    }
    return hipGetLastError();
}
hipError_t launch_shade_nee(int type, const ff_paths& P, uint32_t n_max, hipStream_t s) {
    return uni_shading(P) ? shade_nee_launch<true>(type, P, n_max, s) : shade_nee_launch<false>(type, P, n_max, s);
}
template <bool UNI>
hipError_t shade_adj_launch(int type, const ff_paths& P, int next_q, uint32_t n_max, hipStream_t s) {
    dim3 g(grid_for(n_max, FF_GRID_CAP)), b(FF_BLOCK);
    switch (type) {
        case MAT_MIX: hipLaunchKernelGGL((k_shade_adj<MAT_MIX, UNI>), g, b, 0, s, P, next_q); break;
        case MAT_DIFFUSE: hipLaunchKernelGGL((k_shade_adj<MAT_DIFFUSE, UNI>), g, b, 0, s, P, next_q); break;
        case MAT_CONDUCTOR: hipLaunchKernelGGL((k_shade_adj<MAT_CONDUCTOR, UNI>), g, b, 0, s, P, next_q); break;
        default: hipLaunchKernelGGL((k_shade_adj<MAT_DIELECTRIC, UNI>), g, b, 0, s, P, next_q); break;
    }
    return hipGetLastError();
}
hipError_t launch_shade_adj(int type, const ff_paths& P, int next_q, uint32_t n_max, hipStream_t s) {
    return uni_shading(P) ? shade_adj_launch<true>(type, P, next_q, n_max, s)
                          : shade_adj_launch<false>(type, P, next_q, n_max, s);
}
hipError_t launch_fixed_to_float(const unsigned long long* fx, const uint32_t* mx, const uint32_t* cnt, uint64_t n,
                                 float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fixed_to_float, dim3(grid_for(n, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, fx, mx, cnt, n, out);
    return hipGetLastError();
}
hipError_t launch_fill4(float4* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill4, dim3(grid_for(n, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, p, n);
    return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, p, n);
    return hipGetLastError();
}
hipError_t launch_fill_u64(unsigned long long* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, p, n);
    return hipGetLastError();
}
hipError_t launch_resolve_layers(const float4* layers, uint32_t K, uint32_t WH, float4* film, hipStream_t s) {
    hipLaunchKernelGGL(k_resolve_layers, dim3(grid_for(WH, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, layers, K, WH,
                       film);
    return hipGetLastError();
}
hipError_t launch_trace_closest(const ff_dev_scene& S, const float4* rays, uint32_t n, float4* hits,
                                int two_sided, uint32_t* fetch, int32_t* ovf, unsigned long long* counts,
                                int variant, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(256), 0, s, nullptr, 0u, fetch, 1u << F_TRACE);
    TraceArgs T{S, rays, n, hits, nullptr, two_sided, fetch + F_TRACE * FETCH_SEGS * FETCH_STRIDE, ovf, counts};
    return trav_dispatch(3, variant, nullptr, 0, &T, s);
}
// Device address of `scratch` on `device` (looked up once per device, with
// that device current, by ff_context_create).
hipError_t static_scratch(int device, int32_t** out) {
    static std::mutex mu;
    static std::map<int, int32_t*> addr;
    std::lock_guard<std::mutex> lock(mu);
    *out = nullptr;
    auto it = addr.find(device);
    if (it != addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    addr[device] = (int32_t*)p;
    *out = (int32_t*)p;
    return hipSuccess;
}
hipError_t launch_trace_any(const ff_dev_scene& S, const float4* rays, uint32_t n, uint8_t* occ, uint32_t* fetch,
                            int32_t* ovf, unsigned long long* counts, int variant, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(256), 0, s, nullptr, 0u, fetch, 1u << F_TRACE);
    TraceArgs T{S, rays, n, nullptr, occ, 0, fetch + F_TRACE * FETCH_SEGS * FETCH_STRIDE, ovf, counts};
    return trav_dispatch(4, variant, nullptr, 0, &T, s);
}
hipError_t launch_gen_rays(const ff_dev_camera& cam, int W, int H, uint32_t seed, uint32_t sample, float4* rays,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_gen_rays, dim3(grid_for((uint64_t)W * H, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, cam, W, H,
                       seed, sample, rays);
    return hipGetLastError();
}
hipError_t launch_env_build(const float* rgb, int w, int h, float* lum, double* row_sum, double* denom,
                            double* marg_p, double* marg_y, double* conds, float* pdf, uint32_t* guide_m,
                            uint32_t* guide_c, float4* rgbp, hipStream_t s) {
    hipLaunchKernelGGL(k_env_lum, dim3(grid_for((uint64_t)w * h, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, rgb, w, h,
                       lum);
    hipLaunchKernelGGL(k_env_rows, dim3(h), dim3(FF_BLOCK), 0, s, lum, w, h, row_sum);
    hipLaunchKernelGGL(k_env_denom, dim3(1), dim3(FF_BLOCK), 0, s, row_sum, h, denom);
    hipLaunchKernelGGL(k_env_conditional, dim3(h), dim3(FF_BLOCK), 0, s, lum, w, h, denom, marg_p, conds, pdf);
    hipLaunchKernelGGL(k_env_marginal_scan, dim3(1), dim3(FF_BLOCK), 0, s, marg_p, h, marg_y);
    hipLaunchKernelGGL(k_env_guide, dim3(h + 1), dim3(FF_BLOCK), 0, s, marg_y, conds, w, h, guide_m, guide_c);
    hipLaunchKernelGGL(k_env_pack, dim3(grid_for((uint64_t)w * h, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, rgb, pdf,
                       (uint32_t)((size_t)w * h), rgbp);
    return hipGetLastError();
}
hipError_t launch_film_resolve(const float4* film, uint32_t n, float exposure, float* avg, uint8_t* rgba,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_film_resolve, dim3(grid_for(n, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, film, n, exposure,
                       avg, rgba);
    return hipGetLastError();
}

}  // namespace ffk
