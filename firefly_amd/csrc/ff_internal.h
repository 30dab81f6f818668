// ff_internal.h -- host-side declarations shared by the C ABI implementation,
// the host BVH builder and the kernel launchers.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "ff_layout.h"

namespace ffk {
extern const int kTravVariants;
extern const int kTravLds[];
extern const int kTravBlock[];
extern const int kTravRefill[];
extern const int kTravWhileWhile[];
extern const int kTravWidth[];  // BVH record width of the variant (2 or 4)
// computes the grid sizes of a traversal variant (call before a graph capture)
hipError_t prepare_traversal(int variant);
hipError_t launch_batch_begin(const ff_paths& P, int set, uint32_t value, hipStream_t s);
hipError_t launch_generate(const ff_paths& P, hipStream_t s);
hipError_t launch_extend(const ff_paths& P, int q, int variant, hipStream_t s);
hipError_t launch_shadow(const ff_paths& P, int variant, hipStream_t s);
hipError_t launch_reset(uint32_t* counters, uint32_t mask, uint32_t* fetch, uint32_t fetch_mask, hipStream_t s);
hipError_t launch_logic(const ff_paths& P, int q, int iter, uint32_t n_max, hipStream_t s);
hipError_t launch_shade(int type, const ff_paths& P, int next_q, int first, uint32_t n_max, hipStream_t s);
hipError_t launch_logic_adj(const ff_paths& P, int q, uint32_t n_max, hipStream_t s);
hipError_t launch_shade_nee(int type, const ff_paths& P, uint32_t n_max, hipStream_t s);
hipError_t launch_shade_adj(int type, const ff_paths& P, int next_q, uint32_t n_max, hipStream_t s);
hipError_t launch_fixed_to_float(const unsigned long long* fx, const uint32_t* mx, const uint32_t* cnt, uint64_t n,
                                 float* out, hipStream_t s);
hipError_t launch_fill4(float4* p, uint32_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, uint32_t n, hipStream_t s);
hipError_t launch_resolve_layers(const float4* layers, uint32_t K, uint32_t WH, float4* film, hipStream_t s);
hipError_t launch_trace_closest(const ff_dev_scene& S, const float4* rays, uint32_t n,
                                float4* hits, int two_sided, uint32_t* fetch, int32_t* ovf,
                                unsigned long long* counts, int variant, hipStream_t s);
hipError_t launch_trace_any(const ff_dev_scene& S, const float4* rays, uint32_t n, uint8_t* occ,
                            uint32_t* fetch, int32_t* ovf, unsigned long long* counts, int variant,
                            hipStream_t s);
// device address of the fixed trace scratch on `device` (FETCH_WORDS work
// counter words, then FF_STATIC_OVF_ENTRIES * FF_TRAV_THREADS overflow entries)
hipError_t static_scratch(int device, int32_t** out);
hipError_t launch_gen_rays(const ff_dev_camera& cam, int W, int H, uint32_t seed,
                           uint32_t sample, float4* rays, hipStream_t s);
hipError_t launch_env_build(const float* rgb, int w, int h, float* lum, double* row_sum,
                            double* denom, double* marg_p, double* marg_y, double* conds,
                            float* pdf, uint32_t* guide_m, uint32_t* guide_c, float4* rgbp,
                            hipStream_t s);
hipError_t launch_film_resolve(const float4* film, uint32_t n, float exposure, float* avg,
                               uint8_t* rgba, hipStream_t s);
}  // namespace ffk

namespace ffh {
// 32-byte node as stored in HBM (CRT/BVH.h:63-72 field order).
struct BvhNode {
    float bmin[3];
    float bmax[3];
    int32_t offset;
    uint16_t n_prims;
    uint8_t axis;
    uint8_t pad;
};
static_assert(sizeof(BvhNode) == 32, "BVH node must stay 32 bytes");

// SAH BVH over triangles (positions [n][3][3]); returns nodes in flattened
// depth-first order and the slot -> triangle permutation.
void build_bvh_sah(const float* positions, int32_t n, int32_t max_prims,
                   std::vector<BvhNode>& nodes, std::vector<int32_t>& ordered);
int32_t bvh_max_depth(const std::vector<BvhNode>& nodes);
}  // namespace ffh

namespace ffk {
// Linear BVH on the GPU (Morton order, Karras hierarchy), in the SAH builder's
// node format; positions on the host, nodes and permutation returned there
// (gpu_ms, if given: device time of the build kernels).
hipError_t build_lbvh(const float* positions, int32_t n, int32_t max_prims, std::vector<ffh::BvhNode>& nodes,
                      std::vector<int32_t>& ordered, float* gpu_ms = nullptr);
// PLOC tree on the GPU (locally-ordered clustering of the Morton-sorted
// triangles, SAH leaf collapse), same node format and arguments.
hipError_t build_ploc(const float* positions, int32_t n, int32_t max_prims, std::vector<ffh::BvhNode>& nodes,
                      std::vector<int32_t>& ordered, float* gpu_ms = nullptr);
}  // namespace ffk
