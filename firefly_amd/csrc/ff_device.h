// ff_device.h -- device math, RNG, BSDF lobes, environment light and BVH
// traversal for the gfx950 wavefront path tracer.
//
// Every routine restates a reference routine (JakeKurtz/Firefly,
// CUDA-RayTracer/ = CRT/) cited beside it.  The whole library is compiled
// with -ffp-contract=off so each float/double operation is rounded on its own
// as the reference's arithmetic is written; the traversal results are
// therefore bit-identical to a host evaluation of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ff_layout.h"

namespace ffd {

// ------------------------------------------------------------ constants --
// Float constants as the reference declares them (CRT/dMath.cuh:78-128).
constexpr float F_PI = 3.14159265358979323846f;
constexpr float F_2PI = 6.28318530717958647692f;
constexpr float F_PI_2 = 1.57079632679489661923f;
constexpr float F_PI_4 = 0.785398163397448309615f;
constexpr float F_1_PI = 0.318309886183790671538f;
constexpr float F_1_2PI = 0.159154943091895335769f;
constexpr float K_EPS = 1e-6f;
constexpr float K_HUGE = 1e32f;

// RNG dimensions: bounce 0 is the camera, bounce L the L-th path vertex.
enum : uint32_t {
    DIM_JX = 0, DIM_JY = 1, DIM_LX = 2, DIM_LY = 3,
    DIM_RR = 0, DIM_EX = 1, DIM_EY = 2, DIM_S0 = 3, DIM_S1 = 4,
    DIM_D0 = 5, DIM_D1 = 6, DIM_LOBE = 7, DIM_DIEL = 8
};

// Conservative distance pruning added to the reference traversal (which
// tests every box the ray's line crosses).  Order of visits is unchanged, so
// the first closest candidate is the same triangle the reference returns.
constexpr float PRUNE_REL = 0x1p-12f;
constexpr float PRUNE_ABS = 1e-6f;
constexpr float PRUNE_BEHIND = 1e-4f;
// TravRay::negm flag: 1/d.x or o.x is not finite, so a slab's x distance can be NaN
constexpr uint32_t XNAN_RAY = 8u;

// Transcendentals are evaluated in double and rounded to float: the CUDA
// math library of the reference is closed.  The double routines below are
// this project's own, and the host oracle evaluates the same operations in
// the same order, so host and device agree bit for bit; they are within a
// few ulp of the C library in double, so after rounding to float they nearly
// always equal the correctly rounded value.  They need far fewer registers
// and instructions than general library routines (one division per atan2);
// sin/cos arguments of magnitude 2^19 or more go to the library, out of line.
//
// sin and cos: Cody-Waite reduction modulo pi/2 with a three-part split of
// pi/2 (33 + 33 + 53 bits, so k * part is exact for |k| < 2^20), then the
// Taylor series of sin to r^17 and of cos to r^18 on |r| <= pi/4 (first
// omitted terms below 1e-19; coefficients 1/n! rounded to double).
__device__ __noinline__ void sincos_lib(double x, double* s, double* c) {
    *s = sin(x);
    *c = cos(x);
}
__device__ __forceinline__ void sincos_d(double x, double& s, double& c) {
    if (!(fabs(x) < 0x1p19)) {
        sincos_lib(x, &s, &c);
        return;
    }
    if (x == 0.0) {  // keeps the sign of a zero
        s = x;
        c = 1.0;
        return;
    }
    double kf = rint(x * 0x1.45f306dc9c883p-1);  // x * 2/pi
    int q = (int)kf & 3;
    double r = ((x - kf * 0x1.921fb544p+0) - kf * 0x1.0b4611a6p-34) - kf * 0x1.3198a2e037073p-69;
    double z = r * r;
    double ps = 0x1.952c77030ad4ap-49;
    ps = -0x1.ae7f3e733b81fp-41 + z * ps;
    ps = 0x1.6124613a86d09p-33 + z * ps;
    ps = -0x1.ae64567f544e4p-26 + z * ps;
    ps = 0x1.71de3a556c734p-19 + z * ps;
    ps = -0x1.a01a01a01a01ap-13 + z * ps;
    ps = 0x1.1111111111111p-7 + z * ps;
    ps = -0x1.5555555555555p-3 + z * ps;
    double pc = -0x1.6827863b97d97p-53;
    pc = 0x1.ae7f3e733b81fp-45 + z * pc;
    pc = -0x1.93974a8c07c9dp-37 + z * pc;
    pc = 0x1.1eed8eff8d898p-29 + z * pc;
    pc = -0x1.27e4fb7789f5cp-22 + z * pc;
    pc = 0x1.a01a01a01a01ap-16 + z * pc;
    pc = -0x1.6c16c16c16c17p-10 + z * pc;
    pc = 0x1.5555555555555p-5 + z * pc;
    pc = -0.5 + z * pc;
    double sr = r + (r * z) * ps, cr = 1.0 + z * pc;
    s = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
    c = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}
// atan on [0, 1]: above tan(pi/12) the addition formula
// atan t = pi/6 + atan((t*sqrt3 - 1)/(t + sqrt3)); then the Taylor series to
// u^27 on |u| <= tan(pi/12) (first omitted term below 1e-18; coefficients
// (-1)^n/(2n+1) rounded to double).
__device__ __forceinline__ double atan_unit(double t) {
    bool shift = t > 0x1.126145e9ecd56p-2;
    double u = shift ? (t * 0x1.bb67ae8584caap+0 - 1.0) / (t + 0x1.bb67ae8584caap+0) : t;
    double z = u * u;
    double p = -0x1.2f684bda12f68p-5;
    p = 0x1.47ae147ae147bp-5 + z * p;
    p = -0x1.642c8590b2164p-5 + z * p;
    p = 0x1.8618618618618p-5 + z * p;
    p = -0x1.af286bca1af28p-5 + z * p;
    p = 0x1.e1e1e1e1e1e1ep-5 + z * p;
    p = -0x1.1111111111111p-4 + z * p;
    p = 0x1.3b13b13b13b14p-4 + z * p;
    p = -0x1.745d1745d1746p-4 + z * p;
    p = 0x1.c71c71c71c71cp-4 + z * p;
    p = -0x1.2492492492492p-3 + z * p;
    p = 0x1.999999999999ap-3 + z * p;
    p = -0x1.5555555555555p-2 + z * p;
    double r = u + (u * z) * p;
    return shift ? 0x1.0c152382d7366p-1 + r : r;
}
// atan: odd symmetry; |x| > 1 by atan x = pi/2 - atan(1/x)
__device__ __forceinline__ double atan_d(double x) {
    if (x != x) return x;
    double a = fabs(x);
    bool inv = a > 1.0;
    double r = atan_unit(inv ? 1.0 / a : a);
    if (inv) r = 0x1.921fb54442d18p+0 - r;
    return copysign(r, x);
}
// atan2 with the C library's conventions for zeros, signs and infinities:
// theta = atan(min/max) of the magnitudes (one division), reflected to
// pi/2 - theta above the diagonal and to pi - theta for x < 0, signed by y
__device__ __forceinline__ double atan2_d(double y, double x) {
    if (y != y || x != x) return x + y;
    double a = fabs(y), b = fabs(x);
    if (isinf(a) || isinf(b)) {
        double t = isinf(a) ? (isinf(b) ? (x > 0.0 ? 0x1.921fb54442d18p-1 : 0x1.2d97c7f3321d2p+1)
                                        : 0x1.921fb54442d18p+0)
                            : (x > 0.0 ? 0.0 : 0x1.921fb54442d18p+1);
        return copysign(t, y);
    }
    if (b == 0.0) {
        if (a == 0.0) return signbit(x) ? copysign(0x1.921fb54442d18p+1, y) : y;
        return copysign(0x1.921fb54442d18p+0, y);
    }
    bool inv = a > b;
    double r = atan_unit(inv ? b / a : a / b);
    if (inv) r = 0x1.921fb54442d18p+0 - r;
    if (signbit(x)) r = 0x1.921fb54442d18p+1 - r;
    return copysign(r, y);
}
// asin and acos through atan2 of (x, sqrt((1 - x)(1 + x)))
__device__ __forceinline__ double asin_d(double x) {
    if (!(fabs(x) <= 1.0)) return (x - x) / (x - x);
    return atan2_d(x, sqrt((1.0 - x) * (1.0 + x)));
}
__device__ __forceinline__ double acos_d(double x) {
    if (!(fabs(x) <= 1.0)) return (x - x) / (x - x);
    return atan2_d(sqrt((1.0 - x) * (1.0 + x)), x);
}
__device__ __forceinline__ double sin_d(double x) { double s, c; sincos_d(x, s, c); return s; }
__device__ __forceinline__ double cos_d(double x) { double s, c; sincos_d(x, s, c); return c; }
__device__ __forceinline__ float f_sin(float x) { return (float)sin_d((double)x); }
__device__ __forceinline__ float f_cos(float x) { return (float)cos_d((double)x); }
__device__ __forceinline__ float f_atan(float x) { return (float)atan_d((double)x); }
__device__ __forceinline__ float f_atan2(float y, float x) { return (float)atan2_d((double)y, (double)x); }
__device__ __forceinline__ float f_asin(float x) { return (float)asin_d((double)x); }
__device__ __forceinline__ float f_acos(float x) { return (float)acos_d((double)x); }

// --------------------------------------------------------------- vectors --
struct f3 { float x, y, z; };
__device__ __forceinline__ f3 V3(float a, float b, float c) { f3 r; r.x = a; r.y = b; r.z = c; return r; }
__device__ __forceinline__ f3 vadd(f3 a, f3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 vsub(f3 a, f3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 vmul(f3 a, f3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 vscl(f3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 vdiv(f3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 vneg(f3 a) { return V3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float vdot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 vcross(f3 a, f3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// normalize with 1/sqrtf (the reference's host rsqrtf, CRT/dMath.cuh:70-73,1300-1304)
__device__ __forceinline__ f3 vnorm(f3 v) { float inv = 1.0f / sqrtf(vdot(v, v)); return vscl(v, inv); }
__device__ __forceinline__ bool vzero(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
__device__ __forceinline__ f3 vreflect(f3 i, f3 n) { return vsub(i, vscl(vscl(n, 2.0f), vdot(n, i))); }
__device__ __forceinline__ f3 vlerp(f3 a, f3 b, float t) { return vadd(a, vscl(vsub(b, a), t)); }
__device__ __forceinline__ f3 xyz(float4 v) { return V3(v.x, v.y, v.z); }
__device__ __forceinline__ float4 f4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
// Loads through the global address space (global_load instead of flat_load:
// the scene pointers come from memory, so the compiler cannot infer it).
template <class T>
__device__ __forceinline__ T ldg(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldg4(const float4* p) {
    v4f v = *(const __attribute__((address_space(1))) v4f*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 ldg4i(const int4* p) {
    v4i v = *(const __attribute__((address_space(1))) v4i*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}
// a float4 store through the global address space
__device__ __forceinline__ void stg4(float4* p, float4 v) {
    v4f w = {v.x, v.y, v.z, v.w};
    *(__attribute__((address_space(1))) v4f*)p = w;
}
__device__ __forceinline__ f3 ld3(const float* p) { return V3(ldg(p), ldg(p + 1), ldg(p + 2)); }

// ------------------------------------------------------------------- RNG --
// Counter RNG keyed by (seed, pixel, sample, bounce, dim); replaces the
// clock()-seeded wang_hash of CRT/dRandom.cu:20-27.  Mixer: "lowbias32"
// (C. Wellons, hash-prospector, public domain).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t rng_base(uint32_t seed, uint32_t pixel, uint32_t sample) {
    uint32_t h = mix32(seed ^ 0x5bd1e995U);
    h = mix32(h ^ pixel);
    h = mix32(h + sample * 0x9E3779B9U);
    return h;
}
__device__ __forceinline__ float rng(uint32_t base, uint32_t bounce, uint32_t dim) {
    uint32_t h = mix32(mix32(base ^ ((bounce << 8) | dim)));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// ------------------------------------------------------------- textures --
// Texel columns and rows of a wrap-addressed bilinear tap: with the
// coordinate's fraction in [0, 1], floor(x) lies in [-1, w - 1], so the
// wrap is one conditional add (an index out of range, from a NaN
// coordinate, is taken as 0).
__device__ __forceinline__ void wrap_texels(float fx, float fy, int w, int h, int& i0, int& j0, int& i1,
                                            int& j1) {
    i0 = (int)fx;
    j0 = (int)fy;
    if (i0 < 0) i0 += w;
    if (j0 < 0) j0 += h;
    if ((unsigned)i0 >= (unsigned)w) i0 = 0;
    if ((unsigned)j0 >= (unsigned)h) j0 = 0;
    i1 = i0 + 1 == w ? 0 : i0 + 1;
    j1 = j0 + 1 == h ? 0 : j0 + 1;
}
// Software bilinear filter, wrap addressing on normalized coordinates
// (replaces tex2DLod on linear/wrap texture objects, CRT/dTexture.cu:251-255).
__device__ __forceinline__ void bilinear(const float* __restrict__ tex, int w, int h, int c,
                                         float u, float v, float* out) {
    float uu = u - floorf(u), vv = v - floorf(v);
    float x = uu * (float)w - 0.5f, y = vv * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int i0, j0, i1, j1;
    wrap_texels(fx, fy, w, h, i0, j0, i1, j1);
    float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
    for (int k = 0; k < c; k++) {
        float t00 = ldg(tex + ((size_t)j0 * w + i0) * c + k), t10 = ldg(tex + ((size_t)j0 * w + i1) * c + k);
        float t01 = ldg(tex + ((size_t)j1 * w + i0) * c + k), t11 = ldg(tex + ((size_t)j1 * w + i1) * c + k);
        out[k] = ((w00 * t00 + w10 * t10) + w01 * t01) + w11 * t11;
    }
}
// luminance (CRT/dMath.cuh:1503-1507): double expression returned as float
__device__ __forceinline__ float lum_of(const float* s) {
    return (float)(0.299 * (double)s[0] + 0.587 * (double)s[1] + 0.114 * (double)s[2]);
}

// ------------------------------------------------------------ env light --
// sample_spherical_direction / sample_spherical_map (CRT/dMath.cuh:1483-1501)
__device__ __forceinline__ f3 sph_dir(float u, float v) {
    float theta = F_2PI * u, phi = F_PI * v;
    float x = f_cos(theta) * f_sin(phi), y = f_cos(phi), z = f_sin(theta) * f_sin(phi);
    return V3(-x, -y, -z);
}
// The asin argument is clamped to [-1, 1]: a normalised float direction can
// have |y| = 1 + 1 ulp, where the reference's asin gives NaN radiance.
__device__ __forceinline__ void sph_map(f3 d, float& u, float& v) {
    u = f_atan2(d.z, d.x) * F_1_2PI + 0.5f;
    v = f_asin(fminf(fmaxf(d.y, -1.f), 1.f)) * F_1_PI + 0.5f;
}
// upper_bound with the trailing list[left <= val] adjustment (CRT/Helpers.cu:16-31)
__device__ __forceinline__ int upper_bound_ref(const double* __restrict__ list, int size, float val) {
    int left = 0, right = size;
    while (left < right) {
        int middle = (right - left) / 2 + left;
        if ((double)val >= ldg(list + middle)) left = middle + 1;
        else right = middle;
    }
    if (left < size && ldg(list + (((float)left <= val) ? 1 : 0)) != 0.0) left++;
    return left;
}

struct EnvView {
    const float4* rgbp;    // [h][w] (r, g, b, pdf)
    const double* marg_y;  // [h]
    const double* conds;   // [h][w]
    const uint32_t* guide_m, *guide_c;
    int w, h;
};
// upper_bound_ref over a non-decreasing list, starting from its guide table:
// the guide brackets the first element greater than val (the bracket of
// val's bucket), so the loop below finds the same index as a search of the
// whole list.  The trailing adjustment tests list[(float)left <= val], which
// for val in [0, 1) is list[1] when left == 0 and list[0] otherwise; the
// guide entries carry both tests.  Lists that are not monotone (guide entries
// flagged) are searched whole, and values outside [0, 1) are searched whole
// with the adjustment read from the list, exactly as upper_bound_ref does.
__device__ __forceinline__ int upper_bound_guided(const double* __restrict__ list, int size,
                                                  const uint32_t* __restrict__ guide, float val) {
    int left = 0, right = size;
    if (!(val >= 0.f && val < 1.f)) return upper_bound_ref(list, size, val);
    uint32_t g = (uint32_t)(val * (float)ENV_GUIDE);  // exact: ENV_GUIDE is a power of two
    if (g > ENV_GUIDE - 1) g = ENV_GUIDE - 1;
    uint32_t a = ldg(guide + g), b = ldg(guide + g + 1);
    if (((a | b) & ENV_GUIDE_FULL) == 0) {
        left = (int)(a & ENV_GUIDE_INDEX);
        right = (int)(b & ENV_GUIDE_INDEX);
    }
    while (left < right) {
        int middle = (right - left) / 2 + left;
        if ((double)val >= ldg(list + middle)) left = middle + 1;
        else right = middle;
    }
    if (left < size && (a & (left == 0 ? ENV_GUIDE_NZ1 : ENV_GUIDE_NZ0)) != 0u) left++;
    return left;
}
// dEnvironmentLight::L (CRT/dEnvironmentLight.cu:66-76): bilinear() over the
// (r, g, b) of the interleaved texels, one 16-B load per texel
__device__ __forceinline__ f3 env_L(const EnvView& e, f3 wi) {
    float u, v;
    sph_map(wi, u, v);
    float uu = u - floorf(u), vv = v - floorf(v);
    float x = uu * (float)e.w - 0.5f, y = vv * (float)e.h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int i0, j0, i1, j1;
    wrap_texels(fx, fy, e.w, e.h, i0, j0, i1, j1);
    float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
    float4 t00 = ldg4(e.rgbp + (size_t)j0 * e.w + i0), t10 = ldg4(e.rgbp + (size_t)j0 * e.w + i1);
    float4 t01 = ldg4(e.rgbp + (size_t)j1 * e.w + i0), t11 = ldg4(e.rgbp + (size_t)j1 * e.w + i1);
    return V3(((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x,
              ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y,
              ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z);
}
// dEnvironmentLight::get_pdf(isect, wi) (CRT/dEnvironmentLight.cu:83-103)
__device__ __forceinline__ float env_pdf(const EnvView& e, f3 wi) {
    float u, v;
    sph_map(wi, u, v);
    int px = (int)(u * (float)(e.w - 1)), py = (int)(v * (float)(e.h - 1));
    px = px < 0 ? 0 : (px > e.w - 1 ? e.w - 1 : px);
    py = py < 0 ? 0 : (py > e.h - 1 ? e.h - 1 : py);
    float z = wi.z < -1.f ? -1.f : (wi.z > 1.f ? 1.f : wi.z);
    return ldg4(e.rgbp + (size_t)py * e.w + px).w / f_sin(f_acos(z));
}
// dEnvironmentLight::get_direction (CRT/dEnvironmentLight.cu:41-58);
// indices clamped at 0 where the reference would index out of range.
__device__ __forceinline__ f3 env_sample(const EnvView& e, f3 P, float ex, float ey) {
    int y = (int)((float)upper_bound_guided(e.marg_y, e.h, e.guide_m, ey) - 1.f);
    if (y < 0) y = 0;
    int x = (int)((float)upper_bound_guided(e.conds + (size_t)y * e.w, e.w, e.guide_c + (size_t)y * (ENV_GUIDE + 1),
                                            ex) - 1.f);
    if (x < 0) x = 0;
    f3 dir = sph_dir((float)x / (float)e.w, (float)y / (float)e.h);
    f3 spt = vadd(V3(0.f, 0.f, 0.f), vscl(dir, 10000.f));
    return vnorm(vsub(spt, P));
}

// -------------------------------------------------------------- camera --
// dCamera::gen_ray (CRT/dCamera.cu:10-36) with Matrix4x4*float4
// (CRT/dMatrix.cuh:61-69) and ConcentricSampleDisk (CRT/dRandom.cu:104-126).
__device__ __forceinline__ void gen_ray(const ff_dev_camera& cam, int width, int height, int px,
                                        int py, uint32_t base, f3& o, f3& d) {
    float nx = (float)(2.0f * (((double)px + 0.5) / (double)width) - 1.0f);
    float ny = (float)(1.0f - 2.0f * (((double)py + 0.5) / (double)height));
    const float* m = cam.inv_vp;
    float a[4] = {nx, ny, -1.f, 1.f}, b[4] = {nx, ny, 1.f, 1.f}, pa[4], pb[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        pa[r] = m[0 * 4 + r] * a[0] + m[1 * 4 + r] * a[1] + m[2 * 4 + r] * a[2] + m[3 * 4 + r] * a[3];
        pb[r] = m[0 * 4 + r] * b[0] + m[1 * 4 + r] * b[1] + m[2 * 4 + r] * b[2] + m[3 * 4 + r] * b[3];
    }
    f3 pn = V3(pa[0] / pa[3], pa[1] / pa[3], pa[2] / pa[3]);
    f3 pf = V3(pb[0] / pb[3], pb[1] / pb[3], pb[2] / pb[3]);
    float sx = rng(base, 0, DIM_JX) * 0.005f, sy = rng(base, 0, DIM_JY) * 0.005f;
    o = vadd(pn, V3(sx, sy, 0.f));
    d = vnorm(vsub(pf, pn));
    if (cam.lens_radius > 0.f) {
        f3 pfoc = vadd(o, vscl(d, cam.focal));
        float u0 = rng(base, 0, DIM_LX), u1 = rng(base, 0, DIM_LY);
        float ox = 2.f * u0 - 1.f, oy = 2.f * u1 - 1.f, lx = 0.f, ly = 0.f;
        if (!(ox == 0 && oy == 0)) {
            float theta, r;
            if (fabsf(ox) > fabsf(oy)) { r = ox; theta = F_PI_4 * (oy / ox); }
            else { r = oy; theta = F_PI_2 - F_PI_4 * (ox / oy); }
            lx = r * f_cos(theta);
            ly = r * f_sin(theta);
        }
        o = vadd(o, V3(lx * cam.lens_radius, ly * cam.lens_radius, 0.f));
        d = vnorm(vsub(pfoc, o));
    }
}

// ---------------------------------------------------------------- BSDFs --
__device__ __forceinline__ float pow5(float x) { float x2 = x * x; return (x2 * x2) * x; }
// fresnel (CRT/dMaterial.cu:90-92)
__device__ __forceinline__ f3 fresnel_s(f3 f0, f3 h, f3 wo) {
    float p = pow5(1.f - fmaxf(0.f, vdot(h, wo)));
    return vadd(f0, vscl(V3(1.f - f0.x, 1.f - f0.y, 1.f - f0.z), p));
}
// power_heuristic with nf = ng = 1 (CRT/dMaterial.cu:114-119)
__device__ __forceinline__ double power_heur(double f, double g) { return (f * f) / (f * f + g * g); }
// ggxtr_ndf (CRT/dMaterial.cu:120-125)
__device__ __forceinline__ double ggx_ndf(f3 n, f3 h, float r) {
    double a2 = (double)(r * r * r * r);
    float nh = fmaxf(0.0f, vdot(n, h));
    double nh2 = (double)(nh * nh);
    double q = nh2 * (a2 - 1.0) + 1.0;
    return a2 / ((double)F_PI * (q * q));
}
// geo_atten (CRT/dMaterial.cu:126-137)
__device__ __forceinline__ double geo_att(f3 wi, f3 wo, f3 n, float r) {
    float rp = r + 1.f;
    double k = (double)((rp * rp) / 8.f);
    double nl = (double)fmaxf(vdot(n, wi), 0.0f), nv = (double)fmaxf(vdot(n, wo), 0.0f);
    double g1 = nl / (nl * (1.0 - k) + k), g2 = nv / (nv * (1.0 - k) + k);
    return g1 * g2;
}
// spec_get_pdf (CRT/dMaterial.cu:255-265)
__device__ __forceinline__ float spec_pdf(f3 n, f3 wi, f3 wo, float r) {
    f3 wh = vnorm(vadd(wo, wi));
    double whn = (double)fmaxf(vdot(wh, n), 0.f), wowh = (double)fmaxf(vdot(wo, wh), 0.f);
    double D = ggx_ndf(n, wh, r);
    float den = fmaxf((float)(4.0 * wowh), 0.001f);
    return (float)((D * whn) / (double)den);
}
// spec_f (CRT/dMaterial.cu:266-288)
__device__ __forceinline__ f3 spec_f(f3 n, f3 wi, f3 wo, float r, float m, f3 alb, f3 fres) {
    f3 f0 = vlerp(fres, alb, m);
    f3 wh = vnorm(vadd(wo, wi));
    double ndwi = (double)fmaxf(vdot(n, wi), 0.f), ndwo = (double)fmaxf(vdot(n, wo), 0.f);
    double D = ggx_ndf(n, wh, r), G = geo_att(wi, wo, n, r);
    f3 F = fresnel_s(f0, wh, wi);
    float den = fmaxf((float)(4.0 * ndwo * ndwi), 0.001f);
    return vdiv(vscl(F, (float)(D * G)), den);
}
// diff_f (CRT/dMaterial.cu:164-178)
__device__ __forceinline__ f3 diff_f(f3 wi, f3 wo, float m, f3 alb, f3 fres) {
    f3 f0 = vlerp(fres, alb, m);
    f3 wh = vnorm(vadd(wo, wi));
    f3 F = fresnel_s(f0, wh, wi);
    f3 kd = V3(1.f - F.x, 1.f - F.y, 1.f - F.z);
    kd = vscl(kd, (float)(1.0 - (double)m));
    return vscl(vmul(kd, alb), F_1_PI);
}
// get_orthogonal_vec (CRT/dMath.cuh:1450-1463)
__device__ __forceinline__ f3 ortho_axis(f3 in) {
    if (fabsf(in.x) < 0.57735026919f) return V3(1, 0, 0);
    if (fabsf(in.y) < 0.57735026919f) return V3(0, 1, 0);
    return V3(0, 0, 1);
}
// diff_sample (CRT/dMaterial.cu:139-159)
__device__ __forceinline__ f3 diff_sample(f3 N, float e0, float e1) {
    float sin_t = sqrtf(1.f - e0 * e0);
    float phi = 2.f * F_PI * e1;
    float x = sin_t * f_cos(phi), z = sin_t * f_sin(phi);
    f3 T = vnorm(vcross(N, ortho_axis(N)));
    f3 B = vnorm(vcross(N, T));
    return vnorm(vadd(vadd(vscl(T, x), vscl(N, e0)), vscl(B, z)));
}
// spec_sample (CRT/dMaterial.cu:229-254)
__device__ __forceinline__ f3 spec_sample(f3 N, f3 wo, float r, float e0, float e1) {
    double theta = (double)f_atan(r * r * sqrtf(e0 / (1.f - e0)));
    double phi = (double)((2 * F_PI) * e1);
    double st, ct, sp, cp;
    sincos_d(theta, st, ct);
    sincos_d(phi, sp, cp);
    f3 h = V3((float)(st * cp), (float)ct, (float)(st * sp));
    f3 T = vnorm(vcross(N, ortho_axis(N)));
    f3 B = vnorm(vcross(N, T));
    f3 smp = vadd(vadd(vscl(T, h.x), vscl(N, h.y)), vscl(B, h.z));
    return vneg(vreflect(wo, vnorm(smp)));
}
// Exact Fresnel reflectance of a smooth dielectric (textbook Fresnel
// equations).  No reference counterpart (the reference has no dielectric).
__device__ __forceinline__ float fresnel_dielectric(float cos_i, float eta_i, float eta_t) {
    float sin_i = sqrtf(fmaxf(0.f, 1.f - cos_i * cos_i));
    float sin_t = eta_i / eta_t * sin_i;
    if (sin_t >= 1.f) return 1.f;
    float cos_t = sqrtf(fmaxf(0.f, 1.f - sin_t * sin_t));
    float rpar = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t);
    float rper = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t);
    return 0.5f * (rpar * rpar + rper * rper);
}

// ------------------------------------------------------------- adjoint --
// Partial derivatives of the BSDF values for the path-replay adjoint.  The
// sampling (directions, pdfs, MIS weights) is detached, so only the values
// of spec_f / diff_f depend on albedo and roughness.  No reference
// counterpart (the reference is not differentiable).
struct LobeGrad {
    f3 da;  // d f_c / d albedo_c
    f3 dr;  // d f_c / d roughness
};
// spec_f = F * D * G / den with F = f0 + (1 - f0) p, f0 = fres + (albedo - fres) m,
// D = a2 / (pi q^2), a2 = r^4, q = nh^2 (a2 - 1) + 1, G = g(n.wi) g(n.wo),
// g(x) = x / (x (1 - k) + k), k = (r + 1)^2 / 8.
__device__ __forceinline__ LobeGrad spec_grad(f3 n, f3 wi, f3 wo, float r, float m, f3 alb, f3 fres) {
    // D and G in FP64 as ggx_ndf / geo_att evaluate them (q loses all float
    // precision near the lobe peak at low roughness)
    f3 f0 = vlerp(fres, alb, m);
    f3 wh = vnorm(vadd(wo, wi));
    double ndwi = (double)fmaxf(vdot(n, wi), 0.f), ndwo = (double)fmaxf(vdot(n, wo), 0.f);
    float nh = fmaxf(0.f, vdot(n, wh));
    double rd = (double)r;
    double a2 = (double)(r * r * r * r), nh2 = (double)(nh * nh);
    double q = nh2 * (a2 - 1.0) + 1.0;
    double D = a2 / ((double)F_PI * (q * q));
    double dD = 4.0 * rd * rd * rd * (q - 2.0 * a2 * nh2) / ((double)F_PI * q * q * q);
    float rp = r + 1.f;
    double k = (double)((rp * rp) / 8.f), dk = 0.25 * (rd + 1.0);
    double d1 = ndwi * (1.0 - k) + k, d2 = ndwo * (1.0 - k) + k;
    double g1 = ndwi / d1, g2 = ndwo / d2;
    double dg1 = -ndwi * (1.0 - ndwi) / (d1 * d1), dg2 = -ndwo * (1.0 - ndwo) / (d2 * d2);
    double G = g1 * g2, dG = (dg1 * g2 + g1 * dg2) * dk;
    float p = pow5(1.f - fmaxf(0.f, vdot(wh, wi)));
    float den = fmaxf((float)(4.0 * ndwo * ndwi), 0.001f);
    f3 F = vadd(f0, vscl(V3(1.f - f0.x, 1.f - f0.y, 1.f - f0.z), p));
    LobeGrad g;
    float sa = (float)((double)(m * (1.f - p)) * D * G / (double)den);
    g.da = V3(sa, sa, sa);
    g.dr = vscl(F, (float)((dD * G + D * dG) / (double)den));
    return g;
}
// diff_f = (1 - F) (1 - m) albedo / pi
__device__ __forceinline__ f3 diff_grad(f3 wi, f3 wo, float m, f3 alb, f3 fres) {
    f3 f0 = vlerp(fres, alb, m);
    f3 wh = vnorm(vadd(wo, wi));
    float p = pow5(1.f - fmaxf(0.f, vdot(wh, wi)));
    f3 F = vadd(f0, vscl(V3(1.f - f0.x, 1.f - f0.y, 1.f - f0.z), p));
    float c = (1.f - m) * F_1_PI, s = m * (1.f - p);
    return V3(c * ((1.f - F.x) - alb.x * s), c * ((1.f - F.y) - alb.y * s), c * ((1.f - F.z) - alb.z * s));
}

// The four texels and weights bilinear() reads at (u, v) (same arithmetic).
struct Taps {
    uint32_t i[4];
    float w[4];
};
__device__ __forceinline__ Taps taps_of(int w, int h, float u, float v) {
    float uu = u - floorf(u), vv = v - floorf(v);
    float x = uu * (float)w - 0.5f, y = vv * (float)h - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int i0, j0, i1, j1;
    wrap_texels(fx, fy, w, h, i0, j0, i1, j1);
    Taps t;
    t.i[0] = (uint32_t)(j0 * w + i0); t.i[1] = (uint32_t)(j0 * w + i1);
    t.i[2] = (uint32_t)(j1 * w + i0); t.i[3] = (uint32_t)(j1 * w + i1);
    t.w[0] = (1.f - a) * (1.f - b); t.w[1] = a * (1.f - b); t.w[2] = (1.f - a) * b; t.w[3] = a * b;
    return t;
}

// Scale exponent of one fixed-point gradient element: every partial sum of
// its terms is bounded by count * max|term| < 2^k, so terms are kept in
// units of 2^-(61 - k) and the int64 sum cannot overflow.
__device__ __forceinline__ int fixed_exponent(uint32_t max_bits, uint32_t count) {
    double b = (double)__uint_as_float(max_bits) * (double)count;
    if (!(b > 0.0)) return 0;
    int k;
    (void)frexp(b, &k);
    return 61 - k;
}

// ------------------------------------------------------------ traversal --

// Bounds3f::hit(ray, invDir, dirIsNeg) (CRT/Bounds3f.h:128-160) on one box:
// the line test without [0, tmax] clipping; returns the final tmin/tmax.
// The reference's comparisons are evaluated branch-free (all six plane
// distances first): same values, same outcome, and no load or branch is
// left waiting on the x/y test.
__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz, f3 o,
                                     f3 inv, bool nx, bool ny, bool nz, float& tn, float& tf) {
    float tmin = ((nx ? hx : lx) - o.x) * inv.x;
    float tmax = ((nx ? lx : hx) - o.x) * inv.x;
    float tymin = ((ny ? hy : ly) - o.y) * inv.y;
    float tymax = ((ny ? ly : hy) - o.y) * inv.y;
    float tzmin = ((nz ? hz : lz) - o.z) * inv.z;
    float tzmax = ((nz ? lz : hz) - o.z) * inv.z;
    bool ok_xy = !((tmin > tymax) || (tymin > tmax));
    float t1 = (tymin > tmin) ? tymin : tmin;
    float t2 = (tymax < tmax) ? tymax : tmax;
    bool ok_z = !((t1 > tzmax) || (tzmin > t2));
    tn = (tzmin > t1) ? tzmin : t1;
    tf = (tzmax < t2) ? tzmax : t2;
    return ok_xy & ok_z;
}

// The same test in fewer instructions, for boxes with lo <= hi (every box of
// a scene with finite vertices) and rays whose x distances cannot be NaN
// (finite 1/d.x and o.x): entry = max, exit = min of the three slab
// distances, hit = !(entry > exit).  A NaN y or z distance (0 * inf: origin
// on a slab plane, zero direction component) drops out of the reference's
// comparisons and out of fmaxf/fminf alike, so the outcome, entry and exit
// equal slab()'s.  (A NaN x distance would also hide the y distance from
// the reference's z comparison; such rays take slab().)
__device__ __forceinline__ bool slab_box(float lx, float ly, float lz, float hx, float hy, float hz, f3 o,
                                         f3 inv, bool nx, bool ny, bool nz, float& tn, float& tf) {
    float x0 = ((nx ? hx : lx) - o.x) * inv.x;
    float x1 = ((nx ? lx : hx) - o.x) * inv.x;
    float y0 = ((ny ? hy : ly) - o.y) * inv.y;
    float y1 = ((ny ? ly : hy) - o.y) * inv.y;
    float z0 = ((nz ? hz : lz) - o.z) * inv.z;
    float z1 = ((nz ? lz : hz) - o.z) * inv.z;
    tn = fmaxf(fmaxf(x0, y0), z0);
    tf = fminf(fminf(x1, y1), z1);
    return !(tn > tf);
}

// Triangle record in BVH-slot order: v0, e1 = v1 - v0, e2 = v2 - v0 (the
// edges the reference forms per test, CRT/dTriangle.cu:35-36).
struct Tri { f3 v0, e1, e2; };
__device__ __forceinline__ Tri load_tri(const float* __restrict__ t, int slot) {
    const float* p = t + (size_t)slot * 9;
    Tri r;
    r.v0 = V3(ldg(p + 0), ldg(p + 1), ldg(p + 2));
    r.e1 = V3(ldg(p + 3), ldg(p + 4), ldg(p + 5));
    r.e2 = V3(ldg(p + 6), ldg(p + 7), ldg(p + 8));
    return r;
}
// dTriangle::intersect, TEST_CULL branch (CRT/dTriangle.cu:34-62)
__device__ __forceinline__ bool mt_cull(const Tri& T, f3 o, f3 d, float& t, float& u, float& v) {
    f3 pvec = vcross(d, T.e2);
    float detf = vdot(T.e1, pvec);
    if (detf < K_EPS) return false;
    f3 tvec = vsub(o, T.v0);
    float uf = vdot(tvec, pvec);
    if (uf < 0.f || uf > detf) return false;
    f3 qvec = vcross(tvec, T.e1);
    float vf = vdot(d, qvec);
    if (vf < 0.f || (uf + vf) > detf) return false;
    float tf = vdot(T.e2, qvec);
    double inv_det = 1.0 / (double)detf;
    t = (float)((double)tf * inv_det);
    u = (float)((double)uf * inv_det);
    v = (float)((double)vf * inv_det);
    return true;
}
// dTriangle::intersect, non-culling branch (CRT/dTriangle.cu:63-80)
__device__ __forceinline__ bool mt_two(const Tri& T, f3 o, f3 d, float& t, float& u, float& v) {
    f3 pvec = vcross(d, T.e2);
    double det = (double)vdot(T.e1, pvec);
    if (det > (double)-K_EPS && det < (double)K_EPS) return false;
    double inv_det = 1.0 / det;
    f3 tvec = vsub(o, T.v0);
    float uf = (float)((double)vdot(tvec, pvec) * inv_det);
    if (uf < 0.f || uf > 1.f) return false;
    f3 qvec = vcross(tvec, T.e1);
    float vf = (float)((double)vdot(d, qvec) * inv_det);
    if (vf < 0.f || (double)(uf + vf) > 1.0) return false;
    t = (float)((double)vdot(T.e2, qvec) * inv_det);
    u = uf;
    v = vf;
    return true;
}
// dTriangle::shadow_hit (CRT/dTriangle.cu:110-138)
__device__ __forceinline__ bool mt_shadow(const Tri& T, f3 o, f3 d, float& t) {
    f3 pvec = vcross(d, T.e2);
    double det = (double)vdot(T.e1, pvec);
    if (det > (double)-K_EPS && det < (double)K_EPS) return false;
    double inv_det = 1.0 / det;
    f3 tvec = vsub(o, T.v0);
    double u = (double)vdot(tvec, pvec) * inv_det;
    if (u < 0.0 || u > 1.0) return false;
    f3 qvec = vcross(tvec, T.e1);
    double v = (double)vdot(d, qvec) * inv_det;
    if (v < 0.0 || u + v > 1.0) return false;
    float tf = (float)((double)vdot(T.e2, qvec) * inv_det);
    if (tf < 0.f) return false;
    t = tf;
    return true;
}

// Traversal stack: the first `lds_entries` entries of each lane live in LDS
// ([depth][lane], conflict-free), deeper entries in a per-lane global
// overflow area ([depth - lds_entries][grid thread]).
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(1))) int glb_int;
// Traversal stack: LDS entries first (ds_read / ds_write), then the lane's
// column of the global overflow area.  The pointers carry their address
// spaces so no access goes through a generic (flat) pointer.
struct Stack {
    lds_int* lds;      // &smem[threadIdx.x]
    glb_int* ovf;      // &overflow[grid thread]
    uint32_t lds_stride, ovf_stride;
    int lds_entries;
    int sp;
    __device__ __forceinline__ void push(int v) {
        if (sp < lds_entries) lds[sp * lds_stride] = v;
        else ovf[(size_t)(sp - lds_entries) * ovf_stride] = v;
        sp++;
    }
    __device__ __forceinline__ int pop() {
        sp--;
        int v;
        if (sp < lds_entries) v = lds[sp * lds_stride];
        else v = ovf[(size_t)(sp - lds_entries) * ovf_stride];
        return v;
    }
};

struct Hit { float t, u, v; int id; };
// pruning limit of the current closest hit
__device__ __forceinline__ float prune_lim(float t) { return t + t * PRUNE_REL + PRUNE_ABS; }

// Device BVH: the reference's binary tree (same nodes, same order) stored as
// 64-byte two-child records so one fetch tests both children's boxes:
//   a = (c0.lo.xyz, c0.hi.x)  b = (c0.hi.yz, c1.lo.xy)  c = (c1.lo.z, c1.hi.xyz)
//   m = (ref0, ref1, axis, -)
// ref >= 0: interior record index; ref < 0: leaf, ~ref = first slot << 4 |
// count (count 15 = read leaf_count[first slot]).  The root box is tested
// first, as the reference tests node 0.
struct Bvh {
    const float4* nodes;   // two-child records
    const float4* nodes4;  // 4-wide records (same tree, collapsed)
    const int32_t* leaf_count;
    const float* root;  // lo.xyz, hi.xyz, ref (int bits)
};

__device__ __forceinline__ void leaf_range(const Bvh& B, int ref, int& off, int& cnt) {
    int e = ~ref;
    off = e >> 4;
    cnt = e & 15;
    if (cnt == 15) cnt = ldg(B.leaf_count + off);
}

// Per-lane traversal state: the ray, its reciprocal direction and octant,
// and the next node reference to visit.
struct TravRay {
    f3 o, d, inv;
    bool nx, ny, nz;
    uint32_t negm;  // nx | ny << 1 | nz << 2 | XNAN_RAY
    int ref;
};

// Tests the root box (the reference tests node 0 first); false when the ray
// misses the scene or the scene lies wholly behind the origin.
__device__ __forceinline__ bool trav_begin(const Bvh& B, f3 o, f3 d, TravRay& R, Stack& st,
                                           uint32_t& cnt_nodes) {
    R.o = o;
    R.d = d;
    R.inv = V3(1.f / d.x, 1.f / d.y, 1.f / d.z);
    R.nx = R.inv.x < 0;
    R.ny = R.inv.y < 0;
    R.nz = R.inv.z < 0;
    R.negm = (R.nx ? 1u : 0u) | (R.ny ? 2u : 0u) | (R.nz ? 4u : 0u) |
             ((isfinite(R.inv.x) && isfinite(o.x)) ? 0u : XNAN_RAY);
    float tn, tf;
    cnt_nodes++;
    float4 r0 = ldg4((const float4*)B.root), r1 = ldg4((const float4*)B.root + 1);
    if (!slab(r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, o, R.inv, R.nx, R.ny, R.nz, tn, tf) || tf < -PRUNE_BEHIND)
        return false;
    R.ref = __float_as_int(r1.z);
    st.sp = 0;
    return true;
}

// One interior record: both child boxes are tested; the near child (by
// dirIsNeg[axis]) is visited next and the far one pushed when both are hit.
// Returns false when neither child is entered (the caller pops).
template <bool PRUNE>
__device__ __forceinline__ bool visit_interior(const Bvh& B, TravRay& R, float lim, Stack& st,
                                               uint32_t& cnt_nodes) {
    // 32-bit byte offset (records < 2^26, checked at scene creation): the
    // load adds it to the base without 64-bit address arithmetic
    const float4* n = (const float4*)((const char*)B.nodes + ((uint32_t)R.ref << 6));
    float4 a = ldg4(n), b = ldg4(n + 1), c = ldg4(n + 2);
    int4 m = ldg4i((const int4*)(n + 3));
    cnt_nodes += 2;
    float tn0, tf0, tn1, tf1;
    bool h0, h1;
    if (__ballot((R.negm & XNAN_RAY) != 0) == 0) {  // wave-uniform: no lane can see a NaN x distance
        h0 = slab_box(a.x, a.y, a.z, a.w, b.x, b.y, R.o, R.inv, R.nx, R.ny, R.nz, tn0, tf0);
        h1 = slab_box(b.z, b.w, c.x, c.y, c.z, c.w, R.o, R.inv, R.nx, R.ny, R.nz, tn1, tf1);
    } else {
        h0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, R.o, R.inv, R.nx, R.ny, R.nz, tn0, tf0);
        h1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, R.o, R.inv, R.nx, R.ny, R.nz, tn1, tf1);
    }
    if (PRUNE) {
        h0 = h0 && !(tn0 > lim || tf0 < -PRUNE_BEHIND);
        h1 = h1 && !(tn1 > lim || tf1 < -PRUNE_BEHIND);
    } else {
        h0 = h0 && !(tf0 < -PRUNE_BEHIND);
        h1 = h1 && !(tf1 < -PRUNE_BEHIND);
    }
    // both hit: the near child (by dirIsNeg[axis]) next, the far one pushed;
    // one hit: that one next (no direction test needed)
    if (h0 && h1) {
        bool neg = ((R.negm >> (uint32_t)m.z) & 1u) != 0;  // dirIsNeg[axis]
        st.push(neg ? m.x : m.y);
        R.ref = neg ? m.y : m.x;
        return true;
    }
    if (h0 || h1) {
        R.ref = h0 ? m.x : m.y;
        return true;
    }
    return false;
}

// intersect() closest hit (CRT/dTriangle.cu:140-213), one step: an interior
// record or a leaf's triangles.  A candidate replaces the record when
// 0 <= t < tmin.  Children whose boxes miss (or lie beyond the current hit by
// the conservative slack, or wholly behind the origin) are never visited.
// Returns false once the traversal is complete.
// One 4-wide record: the four grandchild boxes are tested at once and the
// hit ones are visited in the binary tree's near-first order (side by the
// parent's split axis, then each side by its own axis): the first is entered,
// the others pushed so that they pop in order.  Leaves are visited in the
// order the binary walk visits them, so the closest hit (ties included) is
// unchanged.  Empty slots (box +inf..-inf) never hit.
__device__ __forceinline__ float sel4(const float4& v, int k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}
__device__ __forceinline__ int sel4i(const int4& v, int k) {
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}
template <bool PRUNE>
__device__ __forceinline__ bool visit_interior4(const Bvh& B, TravRay& R, float lim, Stack& st,
                                                uint32_t& cnt_nodes) {
    const float4* n = (const float4*)((const char*)B.nodes4 + ((uint32_t)R.ref << 7));
    float4 lx = ldg4(n), ly = ldg4(n + 1), lz = ldg4(n + 2);
    float4 hx = ldg4(n + 3), hy = ldg4(n + 4), hz = ldg4(n + 5);
    int4 rf = ldg4i((const int4*)(n + 6)), ax = ldg4i((const int4*)(n + 7));
    bool hit[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float tn, tf;
        bool h = slab(sel4(lx, k), sel4(ly, k), sel4(lz, k), sel4(hx, k), sel4(hy, k), sel4(hz, k), R.o, R.inv,
                      R.nx, R.ny, R.nz, tn, tf);
        hit[k] = PRUNE ? (h && !(tn > lim || tf < -PRUNE_BEHIND)) : (h && !(tf < -PRUNE_BEHIND));
    }
    cnt_nodes += (rf.x != -1 ? 1u : 0u) + (rf.y != -1 ? 1u : 0u) + (rf.z != -1 ? 1u : 0u) + (rf.w != -1 ? 1u : 0u);
    bool negN = ((R.negm >> (uint32_t)ax.x) & 1u) != 0;
    bool negA = ((R.negm >> (uint32_t)ax.y) & 1u) != 0;
    bool negB = ((R.negm >> (uint32_t)ax.z) & 1u) != 0;
    int s0 = negN ? 2 : 0, s1 = negN ? 0 : 2;
    bool f0 = negN ? negB : negA, f1 = negN ? negA : negB;
    int o0 = s0 + (f0 ? 1 : 0), o1 = s0 + (f0 ? 0 : 1), o2 = s1 + (f1 ? 1 : 0), o3 = s1 + (f1 ? 0 : 1);
    bool h0 = o0 == 0 ? hit[0] : (o0 == 1 ? hit[1] : (o0 == 2 ? hit[2] : hit[3]));
    bool h1 = o1 == 0 ? hit[0] : (o1 == 1 ? hit[1] : (o1 == 2 ? hit[2] : hit[3]));
    bool h2 = o2 == 0 ? hit[0] : (o2 == 1 ? hit[1] : (o2 == 2 ? hit[2] : hit[3]));
    bool h3 = o3 == 0 ? hit[0] : (o3 == 1 ? hit[1] : (o3 == 2 ? hit[2] : hit[3]));
    int r0 = sel4i(rf, o0), r1 = sel4i(rf, o1), r2 = sel4i(rf, o2), r3 = sel4i(rf, o3);
    if (h3 && (h0 || h1 || h2)) st.push(r3);
    if (h2 && (h0 || h1)) st.push(r2);
    if (h1 && h0) st.push(r1);
    if (h0) { R.ref = r0; return true; }
    if (h1) { R.ref = r1; return true; }
    if (h2) { R.ref = r2; return true; }
    if (h3) { R.ref = r3; return true; }
    return false;
}
template <bool WIDE, bool PRUNE>
__device__ __forceinline__ bool visit_node(const Bvh& B, TravRay& R, float lim, Stack& st, uint32_t& cnt_nodes) {
    return WIDE ? visit_interior4<PRUNE>(B, R, lim, st, cnt_nodes)
                : visit_interior<PRUNE>(B, R, lim, st, cnt_nodes);
}

// the triangles of the leaf R.ref (closest-hit test of every one)
__device__ __forceinline__ void leaf_closest(const Bvh& B, const float* __restrict__ tris, const TravRay& R,
                                             Hit& h, bool two_sided, uint32_t& cnt_tris) {
    int off, np;
    leaf_range(B, R.ref, off, np);
    for (int i = 0; i < np; i++) {
        Tri T = load_tri(tris, off + i);
        cnt_tris++;
        float t = 0.f, u = 0.f, v = 0.f;
        bool ok = two_sided ? mt_two(T, R.o, R.d, t, u, v) : mt_cull(T, R.o, R.d, t, u, v);
        if (ok && !(t < 0.f) && t < h.t) { h.t = t; h.u = u; h.v = v; h.id = off + i; }
    }
}
// The closest hit of a persistent traversal lane: t in a register (every
// test reads it), u, v and the triangle slot in three LDS words of the lane
// (written only when a closer hit is found, read once at the end), which
// keeps the closest-hit kernels at 8 waves per SIMD.
typedef __attribute__((address_space(3))) float lds_float;
struct HitSlots {
    lds_float* u;
    lds_float* v;
    lds_int* id;
};
__device__ __forceinline__ void leaf_closest_slots(const Bvh& B, const float* __restrict__ tris, const TravRay& R,
                                                   float& t_hit, const HitSlots& hs, bool two_sided,
                                                   uint32_t& cnt_tris) {
    int off, np;
    leaf_range(B, R.ref, off, np);
    for (int i = 0; i < np; i++) {
        Tri T = load_tri(tris, off + i);
        cnt_tris++;
        float t = 0.f, u = 0.f, v = 0.f;
        bool ok = two_sided ? mt_two(T, R.o, R.d, t, u, v) : mt_cull(T, R.o, R.d, t, u, v);
        if (ok && !(t < 0.f) && t < t_hit) {
            t_hit = t;
            *hs.u = u;
            *hs.v = v;
            *hs.id = off + i;
        }
    }
}
// any-hit test of the leaf R.ref's triangles
__device__ __forceinline__ bool leaf_any(const Bvh& B, const float* __restrict__ tris, const TravRay& R,
                                         uint32_t& cnt_tris) {
    int off, np;
    leaf_range(B, R.ref, off, np);
    for (int i = 0; i < np; i++) {
        Tri T = load_tri(tris, off + i);
        cnt_tris++;
        float t = 0.f;
        if (mt_shadow(T, R.o, R.d, t) && t < K_HUGE) return true;
    }
    return false;
}

template <bool WIDE = false>
__device__ __forceinline__ bool closest_step(const Bvh& B, const float* __restrict__ tris, TravRay& R, Hit& h,
                                             bool two_sided, Stack& st, uint32_t& cnt_nodes,
                                             uint32_t& cnt_tris) {
    if (R.ref >= 0) {
        if (visit_node<WIDE, true>(B, R, prune_lim(h.t), st, cnt_nodes)) return true;
    } else {
        leaf_closest(B, tris, R, h, two_sided, cnt_tris);
    }
    if (st.sp == 0) return false;
    R.ref = st.pop();
    return true;
}

__device__ __forceinline__ Hit no_hit() {
    Hit h;
    h.t = K_HUGE; h.u = 0.f; h.v = 0.f; h.id = -1;
    return h;
}

__device__ __forceinline__ Hit closest_hit(const Bvh& B, const float* __restrict__ tris, f3 o, f3 d,
                                           bool two_sided, Stack& st, uint32_t& cnt_nodes,
                                           uint32_t& cnt_tris) {
    Hit h = no_hit();
    TravRay R;
    if (!trav_begin(B, o, d, R, st, cnt_nodes)) return h;
    while (closest_step(B, tris, R, h, two_sided, st, cnt_nodes, cnt_tris)) {}
    return h;
}

// intersect_shadows with tmin = K_HUGE (CRT/dTriangle.cu:215-259,
// CRT/dEnvironmentLight.cu:60-64), one step.  Returns 0 while work remains,
// 1 when a triangle with 0 <= t < 1e32 is found, 2 when the ray is unoccluded.
template <bool WIDE = false>
__device__ __forceinline__ int any_step(const Bvh& B, const float* __restrict__ tris, TravRay& R, Stack& st,
                                        uint32_t& cnt_nodes, uint32_t& cnt_tris) {
    if (R.ref >= 0) {
        if (visit_node<WIDE, false>(B, R, 0.f, st, cnt_nodes)) return 0;
    } else if (leaf_any(B, tris, R, cnt_tris)) {
        return 1;
    }
    if (st.sp == 0) return 2;
    R.ref = st.pop();
    return 0;
}

__device__ __forceinline__ bool any_hit(const Bvh& B, const float* __restrict__ tris, f3 o, f3 d, Stack& st,
                                        uint32_t& cnt_nodes, uint32_t& cnt_tris) {
    TravRay R;
    if (!trav_begin(B, o, d, R, st, cnt_nodes)) return false;
    int r;
    while ((r = any_step(B, tris, R, st, cnt_nodes, cnt_tris)) == 0) {}
    return r == 1;
}

// ----------------------------------------------------- wave64 utilities --
// Wave-aggregated queue append: one atomic per wave, lane offsets from the
// 64-bit ballot (replaces the 32-lane atomicAggInc, CRT/kernel.cu:24-36).
// Every lane of the wave must call it; returns the slot for lanes with pred.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    uint32_t base = 0;
    if (mask) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
    }
    return base + rank;
}
// Wave-aggregated 64-bit counter of lanes with pred.
__device__ __forceinline__ void wave_count(unsigned long long* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    uint64_t all = __ballot(1);
    if (mask) {
        int leader = __ffsll((unsigned long long)all) - 1;
        if ((int)__lane_id() == leader) atomicAdd(counter, (unsigned long long)__popcll(mask));
    }
}

// Sum of a per-lane value over a full wave, added once per wave.  Must be
// called with every lane of the wave active (after the grid-stride loop).
__device__ __forceinline__ void wave_sum_add(unsigned long long* counter, uint32_t v) {
    unsigned long long s = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (__lane_id() == 0) atomicAdd(counter, s);
}

// Workgroup-aggregated queue allocation: thread t asks for cnt[q] (0..3)
// slots of queue q; one atomic per queue per workgroup (device-scope atomics
// on one address serialise at the memory side, so per-wave atomics on a few
// queue counters throttle whole kernels).  Every thread of the workgroup must
// call it the same number of times.  lds: NQ * (BS / 64 + 1) words.
template <int BS, int NQ>
__device__ __forceinline__ void block_alloc(uint32_t* counters, const int* qidx, const uint32_t* cnt,
                                            uint32_t* first, uint32_t* lds) {
    constexpr int NW = BS / 64;
    const uint32_t w = threadIdx.x >> 6;
    uint32_t pre[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        uint64_t b0 = __ballot((cnt[q] & 1u) != 0), b1 = __ballot((cnt[q] & 2u) != 0);
        uint32_t r0 = __builtin_amdgcn_mbcnt_hi((uint32_t)(b0 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b0, 0u));
        uint32_t r1 = __builtin_amdgcn_mbcnt_hi((uint32_t)(b1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b1, 0u));
        pre[q] = r0 + 2u * r1;
        if (__lane_id() == 0) lds[q * (NW + 1) + w] = (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1);
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int q = threadIdx.x;
        uint32_t tot = 0;
        for (int k = 0; k < NW; k++) {
            uint32_t c = lds[q * (NW + 1) + k];
            lds[q * (NW + 1) + k] = tot;
            tot += c;
        }
        lds[q * (NW + 1) + NW] = tot ? atomicAdd(&counters[qidx[q]], tot) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) first[q] = lds[q * (NW + 1) + NW] + lds[q * (NW + 1) + w] + pre[q];
    __syncthreads();
}

// Queue appends staged in LDS.  A device-scope atomic on one address costs
// about 11 ns at the memory side however many workgroups issue it, so one
// atomic per queue per 256-path tile (a pass over 133 M paths: 520 k of them
// on each counter) bounds a kernel at several milliseconds.  A workgroup
// instead collects the entries of several tiles in LDS and appends each
// queue's run with one atomic when fewer than BS free places remain (and at
// the end): CAP / BS times fewer atomics, and the runs are written coalesced.
// Each lane adds at most one entry per add(), so the staged count grows by at
// most BS per add(): flush_if_full(..., adds) makes room for the next `adds`
// calls of add() (a workgroup that stages two entries per thread per tile
// passes 2).  Every thread of the workgroup calls add(), flush_if_full() and
// flush() the same number of times.
template <int BS, int NQ, int CAP>
struct QStage {
    static constexpr int NW = BS / 64;
    static_assert(CAP >= 4 * BS && NQ < BS && NQ <= 16, "staging room for several adds");
    uint32_t val[CAP];       // entry values
    uint32_t pos[CAP];       // queue << 28 | rank in the queue's run
    uint32_t wtot[NQ][NW];   // this add's entries per queue and wave
    uint32_t fill[NQ];       // entries staged per queue
    uint32_t base[NQ];       // first queue place of each run (flush)
    uint32_t total;          // entries staged

    __device__ __forceinline__ void init() {
        if (threadIdx.x < NQ) fill[threadIdx.x] = 0u;
        if (threadIdx.x == 0) total = 0u;
        __syncthreads();
    }
    // q: the lane's queue (-1: no entry), v: the entry
    __device__ __forceinline__ void add(int q, uint32_t v) {
        const uint32_t w = threadIdx.x >> 6;
        uint32_t rank_q = 0u;
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            uint64_t b = __ballot(q == k);
            uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
            if (q == k) rank_q = r;
            if (__lane_id() == 0) wtot[k][w] = (uint32_t)__popcll(b);
        }
        uint64_t ball = __ballot(q >= 0);
        uint32_t rank_all = __builtin_amdgcn_mbcnt_hi((uint32_t)(ball >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)ball, 0u));
        __syncthreads();
        if (q >= 0) {
            uint32_t off_all = 0u, off_q = 0u;
            for (uint32_t ww = 0; ww < w; ww++) {
#pragma unroll
                for (int k = 0; k < NQ; k++) off_all += wtot[k][ww];
                off_q += wtot[q][ww];
            }
            uint32_t slot = total + off_all + rank_all;
            val[slot] = v;
            pos[slot] = ((uint32_t)q << 28) | (fill[q] + off_q + rank_q);
        }
        __syncthreads();
        if (threadIdx.x <= (unsigned)NQ) {
            uint32_t t = 0u;
            if (threadIdx.x < (unsigned)NQ) {
                for (int ww = 0; ww < NW; ww++) t += wtot[threadIdx.x][ww];
                fill[threadIdx.x] += t;
            } else {
                for (int k = 0; k < NQ; k++)
                    for (int ww = 0; ww < NW; ww++) t += wtot[k][ww];
                total += t;
            }
        }
        __syncthreads();
    }
    // appends every staged run: queue q's entries go to queues[q] from the
    // place an atomic on counters[qidx[q]] returns
    __device__ __forceinline__ void flush(uint32_t* counters, const int* qidx, uint32_t* const* queues) {
        if (threadIdx.x < (unsigned)NQ) {
            uint32_t f = fill[threadIdx.x];
            base[threadIdx.x] = f ? atomicAdd(&counters[qidx[threadIdx.x]], f) : 0u;
        }
        __syncthreads();
        const uint32_t T = total;
        for (uint32_t k = threadIdx.x; k < T; k += BS) {
            uint32_t p = pos[k];
            uint32_t q = p >> 28;
            uint32_t* dst = queues[0];
#pragma unroll
            for (int j = 1; j < NQ; j++)
                if (q == (uint32_t)j) dst = queues[j];
            dst[base[q] + (p & 0x0fffffffu)] = val[k];
        }
        __syncthreads();
        if (threadIdx.x < (unsigned)NQ) fill[threadIdx.x] = 0u;
        if (threadIdx.x == 0) total = 0u;
        __syncthreads();
    }
    __device__ __forceinline__ void flush_if_full(uint32_t* counters, const int* qidx, uint32_t* const* queues,
                                                  uint32_t adds = 1u) {
        // total is the same for every thread; afterwards total + adds * BS <= CAP
        if (total + adds * BS > CAP) flush(counters, qidx, queues);
    }
};

// Segmented work fetch of the persistent traversal: queue [0, n) is split
// into FETCH_SEGS segments with a counter each; a wave draws from its current
// segment and, once that is exhausted, finds the next segment with work left
// by reading all FETCH_SEGS counters at once (lane k reads counter k with a
// device-coherent load, one round trip).  It reports `done` only when such a
// scan finds every segment exhausted.  Counters only grow, so a segment seen
// exhausted stays exhausted and a wave fails at most FETCH_SEGS times.
// Wave-uniform: every lane calls it with the same arguments.
static_assert(FETCH_SEGS == 64, "one segment per lane of the scan");
struct SegFetch {
    uint32_t* ctr;   // FETCH_SEGS counters, FETCH_STRIDE words apart
    uint32_t n, seg_len, seg;
    uint32_t res_first, res_left;  // items claimed by this wave and not yet handed out
    uint32_t chunk;                // items claimed per atomic
    __device__ __forceinline__ void init(uint32_t* c, uint32_t n_, uint32_t home) {
        ctr = c;
        n = n_;
        seg_len = (n + FETCH_SEGS - 1) / FETCH_SEGS;
        seg = home % FETCH_SEGS;
        res_first = 0;
        res_left = 0;
        // a wave claims up to 1/256 of a segment (at most FETCH_CHUNK) per
        // atomic, so refills of a few lanes rarely wait on the memory-side
        // atomic, and the last claims of a segment stay small
        uint32_t c8 = seg_len >> 8;
        chunk = c8 < FETCH_CHUNK ? c8 : FETCH_CHUNK;
    }
    // Hands out up to `want` items from the wave's reservoir, refilling it
    // from the counters when it is empty.  Same contract as take().
    __device__ __forceinline__ uint32_t take_chunked(uint32_t want, uint32_t& got, bool& done) {
        done = false;
        if (res_left == 0) {
            uint32_t g;
            uint32_t f = take(want > chunk ? want : chunk, g, done);
            res_first = __builtin_amdgcn_readfirstlane(f);
            res_left = __builtin_amdgcn_readfirstlane(g);
        }
        got = res_left < want ? res_left : want;
        uint32_t first = res_first;
        res_first += got;
        res_left -= got;
        if (res_left > 0) done = false;
        return first;
    }
    __device__ __forceinline__ uint32_t seg_size(uint32_t s) const {
        uint32_t lo = s * seg_len;
        return lo < n ? (n - lo < seg_len ? n - lo : seg_len) : 0u;
    }
    // Grants up to `want` items: returns the first item and the number
    // granted; got == 0 with done == false means "try again later".
    __device__ __forceinline__ uint32_t take(uint32_t want, uint32_t& got, bool& done) {
        got = 0;
        done = false;
        for (int attempt = 0; attempt < 2; attempt++) {
            uint32_t len = seg_size(seg);
            if (len > 0) {
                uint32_t base = 0;
                if (__lane_id() == 0) base = atomicAdd(ctr + (size_t)seg * FETCH_STRIDE, want);
                base = __shfl(base, 0);
                if (base < len) {
                    got = len - base < want ? len - base : want;
                    return seg * seg_len + base;
                }
            }
            // current segment exhausted: scan every counter (lane k -> segment k)
            uint32_t k = __lane_id();
            bool avail = false;
            if (k < FETCH_SEGS) {
                uint32_t lk = seg_size(k);
                uint32_t v = lk ? __hip_atomic_load(ctr + (size_t)k * FETCH_STRIDE, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT)
                                : 0u;
                avail = lk > 0 && v < lk;
            }
            uint64_t m = __ballot(avail);
            if (m == 0) {
                done = true;
                return 0;
            }
            // next segment with work, searching upwards from the current one
            uint32_t sh = (seg + 1) % 64u;
            uint64_t rot = (m >> sh) | (sh ? (m << (64u - sh)) : 0ull);
            seg = (sh + (uint32_t)(__ffsll((unsigned long long)rot) - 1)) % 64u;
        }
        return 0;
    }
};

}  // namespace ffd
