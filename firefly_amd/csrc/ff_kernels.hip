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
    b.leaf_count = S.leaf_count;
    b.root = S.root;
    return b;
}

__device__ __forceinline__ EnvView env_view(const ff_dev_env& e) {
    EnvView v;
    v.rgb = e.rgb; v.marg_y = e.marg_y; v.conds = e.conds; v.pdf = e.pdf; v.w = e.w; v.h = e.h;
    return v;
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
// Starts a batch: batch index += 1 (or = value), all queue counters cleared.
__global__ void k_batch_begin(uint32_t* bstate, uint32_t* counters, int set, uint32_t value) {
    uint32_t i = threadIdx.x;
    if (i < C_COUNT) counters[i] = 0;
    if (i == 0) bstate[B_BATCH] = set ? value : bstate[B_BATCH] + 1u;
}

// ================================================================ generate
// wf_generate + dCamera::gen_ray (CRT/kernel.cu:457-485): every active slot
// of the batch starts its next sample and joins the extension queue.
__global__ __launch_bounds__(FF_BLOCK) void k_generate(ff_paths P) {
    const uint32_t b = P.bstate[B_BATCH];
    const uint32_t n_round = (P.n_slots + 63u) & ~63u;
    for (uint32_t id = gtid(); id < n_round; id += gstride()) {
        bool act = false;
        if (id < P.n_slots) {
            uint32_t r = find_req(P.reqs, P.n_req, id);
            const ff_dev_request& R = P.reqs[r];
            uint32_t WH = R.W * R.H, local = id - R.slot_base;
            uint32_t layer = local / WH, pix = local % WH;
            int64_t rem = (int64_t)R.spp - (int64_t)b * R.K;
            act = rem > 0 && (int64_t)layer < rem && layer < R.K;
            if (act) {
                uint32_t sample = R.sample_offset + b * R.K + layer;
                uint32_t base = rng_base(R.seed, pix, sample);
                f3 o, d;
                gen_ray(R.C, (int)R.W, (int)R.H, (int)(pix % R.W), (int)(pix / R.W), base, o, d);
                P.ray[2 * id] = f4(o, 0.f);
                P.ray[2 * id + 1] = f4(d, 0.f);
                P.rngb[id] = base;
                P.preq[id] = r;
                P.beta[id] = make_float4(1.f, 1.f, 1.f, __uint_as_float(ST_PREV_DELTA));
                P.Lp[id] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        uint32_t slot = wave_append(&P.counters[C_EXT0], act);
        if (act) P.q_ext[0][slot] = id;
    }
}

// =============================================================== traversal
// Ray sources of the persistent traversal kernels: load() binds queue entry k
// to a ray and its scene, hit()/occluded() write the result.
struct ExtSrc {  // wf_extend (CRT/kernel.cu:487-523): closest hit of every queued path
    ff_paths P;
    const uint32_t* queue;
    __device__ __forceinline__ void load(uint32_t k, uint32_t& id, Bvh& B, const float*& tris, f3& o, f3& d,
                                         bool& two) const {
        id = queue[k];
        const ff_dev_scene& S = P.reqs[P.preq[id]].S;
        B = bvh_of(S);
        tris = S.tris;
        o = xyz(P.ray[2 * id]);
        d = xyz(P.ray[2 * id + 1]);
        two = (__float_as_uint(P.beta[id].w) & ST_INSIDE) != 0;
    }
    __device__ __forceinline__ void hit(uint32_t id, const Hit& h) const {
        P.hit[id] = make_float4(h.t, h.u, h.v, __int_as_float(h.id));
    }
    __device__ __forceinline__ void occluded(uint32_t, bool) const {}
};
struct ShadowSrc {  // wf_shadow (CRT/kernel.cu:525-543): any hit of every queued shadow ray
    ff_paths P;
    __device__ __forceinline__ void load(uint32_t k, uint32_t& s, Bvh& B, const float*& tris, f3& o, f3& d,
                                         bool& two) const {
        s = P.q_shadow[k];
        const ff_dev_scene& S = P.reqs[P.preq[s / 3]].S;
        B = bvh_of(S);
        tris = S.tris;
        o = xyz(P.sray[2 * s]);
        d = xyz(P.sray[2 * s + 1]);
        two = false;
    }
    __device__ __forceinline__ void hit(uint32_t, const Hit&) const {}
    __device__ __forceinline__ void occluded(uint32_t s, bool occ) const { P.vis[s] = occ ? 0 : 1; }
};
struct RaySrc {  // a plain ray array (ff_trace_closest / ff_trace_any)
    ff_dev_scene S;
    const float4* rays;
    float4* hits;
    uint8_t* occ;
    int two_sided;
    __device__ __forceinline__ void load(uint32_t k, uint32_t& id, Bvh& B, const float*& tris, f3& o, f3& d,
                                         bool& two) const {
        id = k;
        B = bvh_of(S);
        tris = S.tris;
        o = xyz(rays[2 * k]);
        d = xyz(rays[2 * k + 1]);
        two = two_sided != 0;
    }
    __device__ __forceinline__ void hit(uint32_t id, const Hit& h) const {
        hits[id] = make_float4(h.t, h.u, h.v, __int_as_float(h.id));
    }
    __device__ __forceinline__ void occluded(uint32_t id, bool o) const { occ[id] = o ? 1 : 0; }
};

// Persistent traversal.  The grid holds only the lanes the chip keeps
// resident; a lane whose ray is finished takes the next queue entry from a
// shared counter (per wave: one ballot and one atomic for all idle lanes), so
// a wave no longer lasts as long as its slowest ray and no second round of
// waves is needed.  A wave refills once REFILL of its lanes are idle (or all
// are), trading idle lanes against refill stalls.  Every wave leaves the loop
// once the counter has passed n and its own rays are finished.
template <int BS, int LDS, int REFILL, bool ANY, class Src>
__device__ __forceinline__ void trav_persistent(const Src& src, uint32_t n, uint32_t* fetch, int32_t* ovf,
                                                unsigned long long* cnt, int* smem) {
    Stack st;
    st.lds = &smem[threadIdx.x];
    st.lds_stride = BS;
    st.lds_entries = LDS;
    st.ovf = ovf + gtid();
    st.ovf_stride = FF_TRAV_THREADS;
    uint32_t cn = 0, ct = 0, id = 0;
    bool busy = false, more = true, two = false;
    Bvh B;
    B.nodes = nullptr; B.leaf_count = nullptr; B.root = nullptr;
    const float* tris = nullptr;
    TravRay R;
    Hit h = no_hit();
    for (;;) {
        uint64_t idle = __ballot(!busy);
        uint32_t n_idle = (uint32_t)__popcll(idle);
        if (more && n_idle >= (uint32_t)REFILL) {
            uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(fetch, n_idle);
            base = __shfl(base, leader);
            if (base + n_idle >= n) more = false;
            if (!busy && base + rank < n) {
                f3 o, d;
                src.load(base + rank, id, B, tris, o, d, two);
                h = no_hit();
                busy = trav_begin(B, o, d, R, st, cn);
                if (!busy) {
                    if (ANY) src.occluded(id, false);
                    else src.hit(id, h);
                }
            }
        }
        if (__ballot(busy) == 0) {
            if (!more) break;
            continue;
        }
        if (busy) {
            if (ANY) {
                int r = any_step(B, tris, R, st, cn, ct);
                if (r != 0) {
                    src.occluded(id, r == 1);
                    busy = false;
                }
            } else if (!closest_step(B, tris, R, h, two, st, cn, ct)) {
                src.hit(id, h);
                busy = false;
            }
        }
    }
    if (cnt) {
        wave_sum_add(&cnt[0], cn);
        wave_sum_add(&cnt[1], ct);
    }
}

template <int BS, int LDS, int RF>
__global__ __launch_bounds__(BS) void k_extend(ff_paths P, int q) {
    __shared__ int smem[LDS * BS];
    ExtSrc src{P, P.q_ext[q]};
    trav_persistent<BS, LDS, RF, false>(src, P.counters[q == 0 ? C_EXT0 : C_EXT1], &P.counters[C_FETCH_EXT], P.ovf,
                                        &P.stats[S_EXT_NODES], smem);
}
template <int BS, int LDS, int RF>
__global__ __launch_bounds__(BS) void k_shadow(ff_paths P) {
    __shared__ int smem[LDS * BS];
    ShadowSrc src{P};
    trav_persistent<BS, LDS, RF, true>(src, P.counters[C_SHADOW], &P.counters[C_FETCH_SHD], P.ovf,
                                       &P.stats[S_SHD_NODES], smem);
}

// counters consumed by the previous stages are cleared by a kernel (graph-safe)
__global__ void k_reset_counters(uint32_t* c, uint32_t mask) {
    uint32_t i = threadIdx.x;
    if (i < C_COUNT && ((mask >> i) & 1u)) c[i] = 0;
}

// =================================================================== logic
// wf_logic (CRT/kernel.cu:297-455): NEE accumulation gated by the light
// ray's visibility, throughput update, Russian roulette for len > 3,
// environment radiance for escaped rays after the camera (or a delta event),
// termination into the film, routing of live paths to material queues.
__global__ __launch_bounds__(FF_BLOCK) void k_logic(ff_paths P, int q) {
    const uint32_t n = P.counters[q == 0 ? C_EXT0 : C_EXT1];
    const uint32_t* queue = P.q_ext[q];
    const uint32_t n_round = (n + 63u) & ~63u;
    for (uint32_t i = gtid(); i < n_round; i += gstride()) {
        bool live = i < n;
        int route = -1;
        bool term = false, primary = false;
        uint32_t id = live ? queue[i] : 0u;
        if (live) {
            const ff_dev_request& R = P.reqs[P.preq[id]];
            const ff_dev_scene& S = R.S;
            float4 bw = P.beta[id];
            uint32_t stw = __float_as_uint(bw.w);
            uint32_t len = (stw & ST_LEN_MASK) + 1;
            primary = len == 1;
            f3 beta = xyz(bw);
            f3 Lp = xyz(P.Lp[id]);
            float4 hh = P.hit[id];
            int slot = __float_as_int(hh.w);
            bool found = slot >= 0;
            term = (len > R.depth) || !found;
            int mtype = -1;
            f3 d = xyz(P.ray[2 * id + 1]);
            if (found) {
                mtype = (int)S.mats[(size_t)S.mat[slot] * FF_MAT_FLOATS_DEV];
                if ((stw & ST_INSIDE) && mtype != MAT_DIELECTRIC) {
                    Tri T = load_tri(S.tris, slot);
                    if (!(vdot(d, vcross(T.e1, T.e2)) < 0.f)) term = true;  // absorbed
                }
            }
            if (!found && (stw & ST_PREV_DELTA)) Lp = vadd(Lp, vmul(beta, env_L(env_view(R.E), d)));
            if (len > 1) {
                bool visL = (stw & ST_NEED_L) && P.vis[3 * id + 0];
                if (visL) {
                    f3 lb = xyz(P.nee[3 * id]);
                    if ((stw & ST_NEED_S) && P.vis[3 * id + 1]) lb = vadd(lb, xyz(P.nee[3 * id + 1]));
                    if ((stw & ST_NEED_D) && P.vis[3 * id + 2]) lb = vadd(lb, xyz(P.nee[3 * id + 2]));
                    Lp = vadd(Lp, vmul(lb, beta));
                }
                float4 fe = P.fext[id];
                f3 f = xyz(fe);
                float cos_ext = P.nee[3 * id].w;
                if (vzero(f) || fe.w == 0.f) {
                    term = true;
                } else {
                    beta = vmul(beta, vdiv(vscl(f, cos_ext), fe.w));
                    if (len > 3) {
                        float qrr = fmaxf(0.05f, 1.f - beta.y);
                        if (rng(P.rngb[id], len, DIM_RR) < qrr) term = true;
                    }
                }
            }
            if (term) {
                uint32_t local = id - R.slot_base;
                float4* acc = R.K > 1 ? &R.layers[local] : &R.film[local % (R.W * R.H)];
                float4 L = *acc;
                *acc = make_float4(L.x + Lp.x, L.y + Lp.y, L.z + Lp.z, L.w + 1.f);
            } else {
                uint32_t nst = (stw & (ST_INSIDE | ST_PREV_DELTA)) | len;
                P.beta[id] = f4(beta, __uint_as_float(nst));
                P.Lp[id] = f4(Lp, 0.f);
                route = mtype;
            }
        }
        wave_count(&P.stats[S_PRIMARY], live && primary);
        wave_count(&P.stats[S_EXTENSION], live && !primary);
        wave_count(&P.stats[S_PATHS], live && term);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            uint32_t slot = wave_append(&P.counters[C_MAT0 + m], route == m);
            if (route == m) P.q_mat[m][slot] = id;
        }
    }
}

// ======================================================= material queues
// Shading of one material class (wf_mat_mix / wf_mat_diffuse / wf_mat_cook,
// CRT/kernel.cu:545-651, lobes of CRT/dMaterial.cu).  The environment light
// sample of wf_logic (CRT/kernel.cu:435-439) is drawn here; the visibility
// rays of BRDF_L's BRDF-sampled MIS terms (CRT/dMaterial.cu:404-448) are
// deferred to the shadow stage with their contributions kept per path.
template <int TYPE>
__global__ __launch_bounds__(FF_BLOCK) void k_shade(ff_paths P, int next_q) {
    const uint32_t n = P.counters[C_MAT0 + TYPE];
    const uint32_t* queue = P.q_mat[TYPE];
    const uint32_t n_round = (n + 63u) & ~63u;
    uint32_t* ext_counter = &P.counters[next_q == 0 ? C_EXT0 : C_EXT1];
    for (uint32_t i = gtid(); i < n_round; i += gstride()) {
        bool live = i < n;
        uint32_t id = live ? queue[i] : 0u;
        bool needL = false, needS = false, needD = false;
        if (live) {
            const ff_dev_request& R = P.reqs[P.preq[id]];
            const ff_dev_scene& S = R.S;
            EnvView E = env_view(R.E);
            float4 bw = P.beta[id];
            uint32_t stw = __float_as_uint(bw.w);
            uint32_t len = stw & ST_LEN_MASK;
            bool inside = (stw & ST_INSIDE) != 0;
            float4 hh = P.hit[id];
            int slot = __float_as_int(hh.w);
            float4 r0 = P.ray[2 * id], r1 = P.ray[2 * id + 1];
            f3 o = xyz(r0), d = xyz(r1);
            uint32_t base = P.rngb[id];
            f3 Pp = vadd(o, vscl(d, hh.x));
            const float* N = S.nrm + (size_t)slot * 9;
            float w0 = (1.f - hh.y) - hh.z;
            f3 nrm = vnorm(vadd(vadd(vscl(ld3(N + 3), hh.y), vscl(ld3(N + 6), hh.z)), vscl(ld3(N), w0)));
            const float* U = S.uv + (size_t)slot * 6;
            float tu = (U[2] * hh.y + U[4] * hh.z) + U[0] * w0;
            float tv = (U[3] * hh.y + U[5] * hh.z) + U[1] * w0;
            const float* M = S.mats + (size_t)S.mat[slot] * FF_MAT_FLOATS_DEV;
            f3 alb = V3(M[1], M[2], M[3]);
            float rough = M[4], metal = M[5];
            f3 fres = V3(M[6], M[7], M[8]);
            int at = (int)M[10], rt = (int)M[11];
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
            }
            f3 wo = vneg(d);
            f3 next_o, next_d, fext;
            float pdf_ext, cos_ext;
            f3 A = V3(0, 0, 0), Bs = A, Cd = A;
            uint32_t nst = (stw & (ST_INSIDE | ST_PREV_DELTA | ST_LEN_MASK));
            if (TYPE == MAT_DIELECTRIC) {
                Tri T = load_tri(S.tris, slot);
                f3 gn = vnorm(vcross(T.e1, T.e2));
                float cs = vdot(d, gn);
                bool entering = cs < 0.f;
                f3 nn = entering ? gn : vneg(gn);
                float cos_i = fabsf(cs);
                if (cos_i > 1.f) cos_i = 1.f;
                float ior = M[9];
                float eta_i = entering ? 1.f : ior, eta_t = entering ? ior : 1.f;
                float F = fresnel_dielectric(cos_i, eta_i, eta_t);
                if (rng(base, len, DIM_DIEL) < F) {
                    next_d = vnorm(vadd(d, vscl(nn, 2.f * cos_i)));
                    next_o = vadd(Pp, vscl(nn, 1e-4f));
                } else {
                    float eta = eta_i / eta_t;
                    float k = 1.f - eta * eta * (1.f - cos_i * cos_i);
                    float cos_t = sqrtf(fmaxf(k, 0.f));
                    next_d = vnorm(vadd(vscl(d, eta), vscl(nn, eta * cos_i - cos_t)));
                    next_o = vsub(Pp, vscl(nn, 1e-4f));
                    inside = !inside;
                }
                fext = V3(1.f, 1.f, 1.f);
                pdf_ext = 1.f;
                cos_ext = 1.f;
                nst = (nst & ST_LEN_MASK) | ST_PREV_DELTA | (inside ? ST_INSIDE : 0u);
            } else {
                constexpr bool do_spec = (TYPE == MAT_MIX || TYPE == MAT_CONDUCTOR);
                constexpr bool do_diff = (TYPE == MAT_MIX || TYPE == MAT_DIFFUSE);
                f3 wi = env_sample(E, Pp, rng(base, len, DIM_EX), rng(base, len, DIM_EY));
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
                            if (!vzero(f)) A = vadd(A, vdiv(vmul(vscl(f, (float)swt), Li), (float)lp));
                        }
                        if (do_diff) {
                            f3 f = vscl(diff_f(wi, wo, metal, alb, fres), ndwi);
                            if (!vzero(f)) A = vadd(A, vdiv(vmul(vscl(f, (float)dwt), Li), (float)lp));
                        }
                    }
                }
                // BRDF-sampled MIS terms (CRT/dMaterial.cu:404-448)
                f3 wi_s = V3(0, 0, 0), wi_d = wi_s;
                if (do_spec) wi_s = spec_sample(nrm, wo, rough, rng(base, len, DIM_S0), rng(base, len, DIM_S1));
                if (do_diff) wi_d = diff_sample(nrm, rng(base, len, DIM_D0), rng(base, len, DIM_D1));
                f3 so = Pp, dd_o = Pp;
                if (do_spec) {
                    float nd = vdot(nrm, wi_s);
                    if (nd > 0.f) {
                        double sp = (double)spec_pdf(nrm, wi_s, wo, rough);
                        double lp = (double)env_pdf(E, wi_s);
                        double wt = power_heur(sp, lp);
                        f3 f = vscl(spec_f(nrm, wi_s, wo, rough, metal, alb, fres), nd);
                        f3 L2 = env_L(E, wi_s);
                        if (!vzero(f) && sp > 0.f && !vzero(L2)) {
                            Bs = vdiv(vscl(vmul(f, L2), (float)wt), (float)sp);
                            needS = true;
                            so = vadd(Pp, vscl(wi_s, 0.001f));
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
                            Cd = vdiv(vscl(vmul(f, L3), (float)wt), (float)diff_pdf);
                            needD = true;
                            // the mixed material offsets this ray along wi_spec (CRT/dMaterial.cu:431)
                            dd_o = vadd(Pp, vscl(TYPE == MAT_MIX ? wi_s : wi_d, 0.001f));
                        }
                    }
                }
                needL = !vzero(A) || needS || needD;
                f3 dir;
                if (TYPE == MAT_MIX) {
                    dir = (rng(base, len, DIM_LOBE) < 0.5f) ? wi_d : wi_s;
                    fext = vadd(spec_f(nrm, dir, wo, rough, metal, alb, fres), diff_f(dir, wo, metal, alb, fres));
                    float spf = spec_pdf(nrm, dir, wo, rough);
                    float dpf = (float)diff_pdf;
                    pdf_ext = (float)(0.5 * (double)(spf + dpf));
                } else if (TYPE == MAT_DIFFUSE) {
                    dir = wi_d;
                    fext = diff_f(dir, wo, metal, alb, fres);
                    pdf_ext = (float)diff_pdf;
                } else {
                    dir = wi_s;
                    fext = spec_f(nrm, dir, wo, rough, metal, alb, fres);
                    pdf_ext = spec_pdf(nrm, dir, wo, rough);
                }
                cos_ext = fmaxf(0.0f, vdot(nrm, dir));
                next_o = Pp;
                next_d = dir;
                nst = (nst & (ST_LEN_MASK | ST_INSIDE)) | (needL ? ST_NEED_L : 0u) |
                      (needS ? ST_NEED_S : 0u) | (needD ? ST_NEED_D : 0u);
                if (needL) {
                    P.sray[6 * id + 0] = f4(vadd(Pp, vscl(nrm, 0.001f)), 0.f);
                    P.sray[6 * id + 1] = f4(wi, 0.f);
                }
                if (needS) {
                    P.sray[6 * id + 2] = f4(so, 0.f);
                    P.sray[6 * id + 3] = f4(wi_s, 0.f);
                }
                if (needD) {
                    P.sray[6 * id + 4] = f4(dd_o, 0.f);
                    P.sray[6 * id + 5] = f4(wi_d, 0.f);
                }
            }
            P.nee[3 * id] = f4(A, cos_ext);
            P.nee[3 * id + 1] = f4(Bs, 0.f);
            P.nee[3 * id + 2] = f4(Cd, 0.f);
            P.fext[id] = f4(fext, pdf_ext);
            P.ray[2 * id] = f4(next_o, 0.f);
            P.ray[2 * id + 1] = f4(next_d, 0.f);
            P.beta[id].w = __uint_as_float(nst);
        }
        uint32_t es = wave_append(ext_counter, live);
        if (live) P.q_ext[next_q][es] = id;
        wave_count(&P.stats[S_SHADOW], needL);
        wave_count(&P.stats[S_SHADOW], needS);
        wave_count(&P.stats[S_SHADOW], needD);
        uint32_t s0 = wave_append(&P.counters[C_SHADOW], needL);
        if (needL) P.q_shadow[s0] = 3 * id + 0;
        uint32_t s1 = wave_append(&P.counters[C_SHADOW], needS);
        if (needS) P.q_shadow[s1] = 3 * id + 1;
        uint32_t s2 = wave_append(&P.counters[C_SHADOW], needD);
        if (needD) P.q_shadow[s2] = 3 * id + 2;
    }
}

// ============================================================ film layers
__global__ void k_fill4(float4* p, uint32_t n) {
    for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void k_fill_u64(unsigned long long* p, uint32_t n) {
    uint32_t i = threadIdx.x;
    if (i < n) p[i] = 0ull;
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
template <int BS, int LDS, int RF>
__global__ __launch_bounds__(BS) void k_trace_closest(ff_dev_scene S, const float4* rays, uint32_t n,
                                                      float4* hits, int two_sided, uint32_t* fetch, int32_t* ovf,
                                                      unsigned long long* counts) {
    __shared__ int smem[LDS * BS];
    RaySrc src{S, rays, hits, nullptr, two_sided};
    trav_persistent<BS, LDS, RF, false>(src, n, fetch, ovf, counts, smem);
}
template <int BS, int LDS, int RF>
__global__ __launch_bounds__(BS) void k_trace_any(ff_dev_scene S, const float4* rays, uint32_t n, uint8_t* occ,
                                                  uint32_t* fetch, int32_t* ovf, unsigned long long* counts) {
    __shared__ int smem[LDS * BS];
    RaySrc src{S, rays, nullptr, occ, 0};
    trav_persistent<BS, LDS, RF, true>(src, n, fetch, ovf, counts, smem);
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

// single workgroup: denom, marginal_p, marginal_y (blocked inclusive scan)
__global__ __launch_bounds__(FF_BLOCK) void k_env_marginal(const float* lum, const double* row_sum, int32_t w,
                                                           int32_t h, double* denom_out, double* marg_p,
                                                           double* marg_y) {
    __shared__ double red[FF_BLOCK];
    double acc = 0.0;
    for (int j = threadIdx.x; j < h; j += FF_BLOCK) acc += row_sum[j];
    double denom = block_sum(acc, red);
    if (threadIdx.x == 0) *denom_out = denom;
    // marginal_p[j] = sum_i lum * (sin_theta_j / denom): thread-per-row, serial in i
    for (int j = threadIdx.x; j < h; j += FF_BLOCK) {
        float v = (float)j / (float)h;
        double st = (double)f_sin(F_PI * v) / denom;
        double s = 0.0;
        for (int i = 0; i < w; i++) s += (double)lum[(size_t)j * w + i] * st;
        marg_p[j] = s;
    }
    __syncthreads();
    // blocked scan: thread t owns rows [t*c, (t+1)*c)
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

// one workgroup per row: conditional CDF (blocked scan) and pdf texels
__global__ __launch_bounds__(FF_BLOCK) void k_env_conditional(const float* lum, int32_t w, int32_t h,
                                                              const double* denom_p, const double* marg_p,
                                                              double* conds, float* pdf) {
    __shared__ double red[FF_BLOCK];
    int y = blockIdx.x;
    double denom = *denom_p;
    float v = (float)y / (float)h;
    double st = (double)f_sin(F_PI * v);
    double val = st / (denom * marg_p[y]);
    const float* row = lum + (size_t)y * w;
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

// traversal variants (block size, LDS stack entries, refill threshold)
#define FF_TRAV_VARIANT_LIST(X) \
    X(0, 256, 16, 32)           \
    X(1, 256, 16, 16)           \
    X(2, 256, 16, 64)           \
    X(3, 128, 16, 32)           \
    X(4, 256, 8, 32)            \
    X(5, 256, 24, 32)           \
    X(6, 64, 32, 32)            \
    X(7, 128, 24, 32)

template <int BS, int LDS, int RF>
hipError_t trav_launch(int which, const ff_paths& P, int q, hipStream_t s) {
    if (which == 0) {
        uint32_t g = resident_grid(k_extend<BS, LDS, RF>, BS);
        hipLaunchKernelGGL((k_extend<BS, LDS, RF>), dim3(g), dim3(BS), 0, s, P, q);
    } else if (which == 1) {
        uint32_t g = resident_grid(k_shadow<BS, LDS, RF>, BS);
        hipLaunchKernelGGL((k_shadow<BS, LDS, RF>), dim3(g), dim3(BS), 0, s, P);
    } else {  // grid sizes only (before a graph capture)
        (void)resident_grid(k_extend<BS, LDS, RF>, BS);
        (void)resident_grid(k_shadow<BS, LDS, RF>, BS);
    }
    return hipGetLastError();
}

hipError_t trav_dispatch(int which, int variant, const ff_paths& P, int q, hipStream_t s) {
    switch (variant) {
#define FF_TRAV_CASE(v, bs, lds, rf) \
    case v: return trav_launch<bs, lds, rf>(which, P, q, s);
        FF_TRAV_VARIANT_LIST(FF_TRAV_CASE)
#undef FF_TRAV_CASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

// ===================================================================== host
namespace ffk {

const int kTravVariants = 8;
#define FF_TRAV_BS(v, bs, lds, rf) bs,
#define FF_TRAV_LDS(v, bs, lds, rf) lds,
#define FF_TRAV_RF(v, bs, lds, rf) rf,
const int kTravBlock[kTravVariants] = {FF_TRAV_VARIANT_LIST(FF_TRAV_BS)};
const int kTravLds[kTravVariants] = {FF_TRAV_VARIANT_LIST(FF_TRAV_LDS)};
const int kTravRefill[kTravVariants] = {FF_TRAV_VARIANT_LIST(FF_TRAV_RF)};
#undef FF_TRAV_BS
#undef FF_TRAV_LDS
#undef FF_TRAV_RF

hipError_t launch_batch_begin(const ff_paths& P, int set, uint32_t value, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, s, P.bstate, P.counters, set, value);
    return hipGetLastError();
}
hipError_t launch_generate(const ff_paths& P, hipStream_t s) {
    hipLaunchKernelGGL(k_generate, dim3(grid_for(P.n_slots, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, P);
    return hipGetLastError();
}
hipError_t launch_extend(const ff_paths& P, int q, int variant, hipStream_t s) {
    return trav_dispatch(0, variant, P, q, s);
}
hipError_t launch_shadow(const ff_paths& P, int variant, hipStream_t s) { return trav_dispatch(1, variant, P, 0, s); }
hipError_t prepare_traversal(int variant) {
    ff_paths none{};
    return trav_dispatch(2, variant, none, 0, nullptr);
}
hipError_t launch_reset(uint32_t* counters, uint32_t mask, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counters, mask);
    return hipGetLastError();
}
hipError_t launch_logic(const ff_paths& P, int q, uint32_t n_max, hipStream_t s) {
    hipLaunchKernelGGL(k_logic, dim3(grid_for(n_max, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, P, q);
    return hipGetLastError();
}
hipError_t launch_shade(int type, const ff_paths& P, int next_q, uint32_t n_max, hipStream_t s) {
    dim3 g(grid_for(n_max, FF_GRID_CAP)), b(FF_BLOCK);
    switch (type) {
        case MAT_MIX: hipLaunchKernelGGL(k_shade<MAT_MIX>, g, b, 0, s, P, next_q); break;
        case MAT_DIFFUSE: hipLaunchKernelGGL(k_shade<MAT_DIFFUSE>, g, b, 0, s, P, next_q); break;
        case MAT_CONDUCTOR: hipLaunchKernelGGL(k_shade<MAT_CONDUCTOR>, g, b, 0, s, P, next_q); break;
        default: hipLaunchKernelGGL(k_shade<MAT_DIELECTRIC>, g, b, 0, s, P, next_q); break;
    }
    return hipGetLastError();
}
hipError_t launch_fill4(float4* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill4, dim3(grid_for(n, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, p, n);
    return hipGetLastError();
}
hipError_t launch_fill_u64(unsigned long long* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, p, n);
    return hipGetLastError();
}
hipError_t launch_resolve_layers(const float4* layers, uint32_t K, uint32_t WH, float4* film, hipStream_t s) {
    hipLaunchKernelGGL(k_resolve_layers, dim3(grid_for(WH, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, layers, K, WH,
                       film);
    return hipGetLastError();
}
hipError_t launch_trace_closest(const ff_dev_scene& S, const float4* rays, uint32_t n, float4* hits,
                                int two_sided, uint32_t* counters, int32_t* ovf, unsigned long long* counts,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counters, 1u << C_FETCH_TRACE);
    uint32_t g = resident_grid(k_trace_closest<256, 16, 32>, 256);
    hipLaunchKernelGGL((k_trace_closest<256, 16, 32>), dim3(g), dim3(256), 0, s, S, rays, n, hits, two_sided,
                       counters + C_FETCH_TRACE, ovf, counts);
    return hipGetLastError();
}
hipError_t launch_trace_any(const ff_dev_scene& S, const float4* rays, uint32_t n, uint8_t* occ, uint32_t* counters,
                            int32_t* ovf, unsigned long long* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counters, 1u << C_FETCH_TRACE);
    uint32_t g = resident_grid(k_trace_any<256, 16, 32>, 256);
    hipLaunchKernelGGL((k_trace_any<256, 16, 32>), dim3(g), dim3(256), 0, s, S, rays, n, occ,
                       counters + C_FETCH_TRACE, ovf, counts);
    return hipGetLastError();
}
hipError_t launch_gen_rays(const ff_dev_camera& cam, int W, int H, uint32_t seed, uint32_t sample, float4* rays,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_gen_rays, dim3(grid_for((uint64_t)W * H, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, cam, W, H,
                       seed, sample, rays);
    return hipGetLastError();
}
hipError_t launch_env_build(const float* rgb, int w, int h, float* lum, double* row_sum, double* denom,
                            double* marg_p, double* marg_y, double* conds, float* pdf, hipStream_t s) {
    hipLaunchKernelGGL(k_env_lum, dim3(grid_for((uint64_t)w * h, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, rgb, w, h,
                       lum);
    hipLaunchKernelGGL(k_env_rows, dim3(h), dim3(FF_BLOCK), 0, s, lum, w, h, row_sum);
    hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(FF_BLOCK), 0, s, lum, row_sum, w, h, denom, marg_p, marg_y);
    hipLaunchKernelGGL(k_env_conditional, dim3(h), dim3(FF_BLOCK), 0, s, lum, w, h, denom, marg_p, conds, pdf);
    return hipGetLastError();
}
hipError_t launch_film_resolve(const float4* film, uint32_t n, float exposure, float* avg, uint8_t* rgba,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_film_resolve, dim3(grid_for(n, FF_GRID_CAP)), dim3(FF_BLOCK), 0, s, film, n, exposure,
                       avg, rgba);
    return hipGetLastError();
}

}  // namespace ffk
