// Device-side kernel functions and wave helpers for the MI355X H2 engine.
//
// Kernel formulas follow src/H2Pack_3D_kernels.h: Coulomb :52-96 (1/r, the
// r = 0 entry is param[0] or 0), Gaussian :169-209 (exp(-l r^2)), Matern32
// :388-432 ((1 - t) e^t with t = -sqrt(3) l r), RPY :893-959 (3x3 tensor
// t1 I + t2 r^ r^T with three regimes of r against the radii).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define H2P_HD __device__ __attribute__((always_inline))
#include "h2p_exp2.h"

#define H2P_WAVE 64

namespace h2p {

typedef double d4 __attribute__((ext_vector_type(4)));   // one lane's share of a 16x16 FP64 MFMA tile
typedef float f4 __attribute__((ext_vector_type(4)));    // one lane's share of a 16x16 FP32 MFMA tile

// Kernel ids.  KGW / KMW are Gaussian / Matern-3/2 for FP64 JIT handles
// whose point extent can push the exponential's argument past the range of
// the unclamped exp2_neg: their kernel values clamp the argument first
// (bitwise the same values wherever the unclamped form is in range).
enum { KC = 0, KG = 1, KM = 2, KR = 3, KGW = 4, KMW = 5 };

__host__ __device__ constexpr int kbase(int kid) { return kid == KGW ? KG : kid == KMW ? KM : kid; }

// the table of exp2_neg, copied into LDS by the kernels that use it
__constant__ double c_e2tab[H2P_E2_N] = {H2P_E2_TABLE};

struct KParam {
    double p0;      // Coulomb diagonal value / Gaussian l / Matern l / RPY eta
    double p1;      // RPY: C = 1/(6 pi eta); others unused
};

// Scalar kernels as functions of r^2 (kd = 1).
template <int KID>
__device__ __forceinline__ double kval(double r2, const KParam &kp) {
    if constexpr (KID == KC) {
        return (r2 == 0.0) ? kp.p0 : rsqrt(r2);
    } else if constexpr (KID == KG) {
        return exp(-kp.p0 * r2);
    } else {
        double t = -1.7320508075688772 * kp.p0 * sqrt(r2);
        return (1.0 - t) * exp(t);
    }
}

template <int KID>
__device__ __forceinline__ float kvalf(float r2, float p0) {
    if constexpr (KID == KC) {
        return (r2 == 0.0f) ? p0 : __builtin_amdgcn_rsqf(r2);
    } else if constexpr (KID == KG) {
        // exp(-l r2) = 2^(-l r2 log2 e)
        return __builtin_amdgcn_exp2f(-p0 * 1.4426950408889634f * r2);
    } else {
        float t = -1.7320508075688772f * p0 * __builtin_sqrtf(r2);
        return (1.0f - t) * __builtin_amdgcn_exp2f(t * 1.4426950408889634f);
    }
}

// RPY tensor coefficients (t1, t2) and unit vector for one pair.
__device__ __forceinline__ void rpy_coef(double dx, double dy, double dz,
                                         double ta, double sa, double C,
                                         double &t1, double &t2,
                                         double &ux, double &uy, double &uz) {
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = sqrt(r2);
    double inv_r = (r == 0.0) ? 0.0 : 1.0 / r;
    double inv_r2 = inv_r * inv_r;
    ux = dx * inv_r; uy = dy * inv_r; uz = dz * inv_r;
    double tps = ta + sa, tms = ta - sa;
    if (r > tps) {
        double tmp0 = C * 0.75 * inv_r;
        double tmp1 = (ta * ta + sa * sa) * inv_r2;
        t1 = tmp0 * (1.0 + tmp1 * (1.0 / 3.0));
        t2 = tmp0 * (1.0 - tmp1);
    } else if (r > fabs(tms)) {
        double tmp0 = tms * tms;
        double tmp1 = (inv_r2 * inv_r * C) / (ta * sa * 32.0);
        double a = tmp0 + 3.0 * r2;
        t1 = tmp1 * (16.0 * r2 * r * tps - a * a);
        double b = tmp0 - r2;
        t2 = tmp1 * 3.0 * b * b;
    } else {
        t1 = C / fmax(ta, sa);
        t2 = 0.0;
    }
}

// 1/sqrt(q) for q > 0 in FP64 (defined below).
__device__ __forceinline__ double rsqrt_refined(double q);

// RPY coefficients as rpy_coef, for the matvec's kernel-evaluating launch:
// 1/r from the refined reciprocal square root and the radius divisions as
// products of the points' reciprocal radii (ita = 1/ta, isa = 1/sa), which
// leaves no FP64 division or square root in the pair loop; the result
// differs from the divided form by a few ulp.
__device__ __forceinline__ void rpy_coef_rcp(double dx, double dy, double dz,
                                             double ta, double sa, double ita, double isa, double C,
                                             double &t1, double &t2,
                                             double &ux, double &uy, double &uz) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const bool z = r2 == 0.0;
    const double inv_r = z ? 0.0 : rsqrt_refined(z ? 1.0 : r2);
    const double r = r2 * inv_r;
    const double inv_r2 = inv_r * inv_r;
    ux = dx * inv_r; uy = dy * inv_r; uz = dz * inv_r;
    const double tps = ta + sa, tms = ta - sa;
    if (r > tps) {
        const double tmp0 = C * 0.75 * inv_r;
        const double tmp1 = (ta * ta + sa * sa) * inv_r2;
        t1 = tmp0 * (1.0 + tmp1 * (1.0 / 3.0));
        t2 = tmp0 * (1.0 - tmp1);
    } else if (r > fabs(tms)) {
        const double tmp0 = tms * tms;
        const double tmp1 = (inv_r2 * inv_r * C) * (ita * isa * (1.0 / 32.0));
        const double a = tmp0 + 3.0 * r2;
        t1 = tmp1 * (16.0 * r2 * r * tps - a * a);
        const double b = tmp0 - r2;
        t2 = tmp1 * 3.0 * b * b;
    } else {
        t1 = C * fmin(ita, isa);
        t2 = 0.0;
    }
}

// The same in FP32 (the FP32-mixed RPY launch): 1/r from the hardware
// reciprocal square root (within 1 ulp), the radius divisions as products
// of the reciprocal radii.
__device__ __forceinline__ void rpy_coef_rcp(float dx, float dy, float dz,
                                             float ta, float sa, float ita, float isa, float C,
                                             float &t1, float &t2,
                                             float &ux, float &uy, float &uz) {
    const float r2 = dx * dx + dy * dy + dz * dz;
    const bool z = r2 == 0.0f;
    const float inv_r = z ? 0.0f : __builtin_amdgcn_rsqf(z ? 1.0f : r2);
    const float r = r2 * inv_r;
    const float inv_r2 = inv_r * inv_r;
    ux = dx * inv_r; uy = dy * inv_r; uz = dz * inv_r;
    const float tps = ta + sa, tms = ta - sa;
    if (r > tps) {
        const float tmp0 = C * 0.75f * inv_r;
        const float tmp1 = (ta * ta + sa * sa) * inv_r2;
        t1 = tmp0 * (1.0f + tmp1 * (1.0f / 3.0f));
        t2 = tmp0 * (1.0f - tmp1);
    } else if (r > fabsf(tms)) {
        const float tmp0 = tms * tms;
        const float tmp1 = (inv_r2 * inv_r * C) * (ita * isa * (1.0f / 32.0f));
        const float a = tmp0 + 3.0f * r2;
        t1 = tmp1 * (16.0f * r2 * r * tps - a * a);
        const float b = tmp0 - r2;
        t2 = tmp1 * 3.0f * b * b;
    } else {
        t1 = C * fminf(ita, isa);
        t2 = 0.0f;
    }
}

// 1/sqrt(q) for q > 0 in FP64: the hardware estimate refined by one
// Newton-Raphson step with its second-order term, e = 1 - q y^2,
// y <- y + y e (1/2 + 3/8 e); the error of the estimate is cubed.
__device__ __forceinline__ double rsqrt_refined(double q) {
    const double y = __builtin_amdgcn_rsq(q);
    const double e = __builtin_fma(-q * y, y, 1.0);
    return __builtin_fma(y * e, __builtin_fma(0.375, e, 0.5), y);
}

// Scalar kernels on prepared squared distances.  Coordinates are scaled
// before the distance is formed so that the kernels need no parameter:
//   Gaussian exp(-l r^2)          : scale sqrt(128 l log2 e) -> K = 2^(-q / 128)
//   Matern32 (1+t) e^{-t}, t=sqrt3 l r : scale sqrt3 l       -> K = (1+s) e^(-s), s = sqrt(q)
//   Coulomb 1/r                   : scale 1                  -> K = q == 0 ? diag : rsqrt(q)
// (kscale(kid, p0, true)).  e2t is the exp2_neg128 table (LDS), unused by
// Coulomb.
template <int KID>
__device__ __forceinline__ double kprep(double q, double diag, const double *e2t = nullptr) {
    if constexpr (KID == KC) {
        const bool z = q == 0.0;
        const double y = rsqrt_refined(z ? 1.0 : q);   // branch-free
        return z ? diag : y;
    } else if constexpr (KID == KG) {
        return exp2_neg128(q, e2t);
    } else if constexpr (KID == KGW) {
        return exp2_neg128_clamped(q, e2t);
    } else {
        static_assert(KID == KM || KID == KMW, "scalar kernels only");
        // sqrt(q) as q / sqrt(q); the clamp keeps q = 0 (and a rounded
        // negative q from the norm-expanded distance) finite, s ~ 1e-150
        const double qq = __builtin_fmax(q, 1e-300);
        const double s = qq * rsqrt_refined(qq);
        const double x = s * (128.0 * 1.4426950408889634);   // e^(-s) = 2^(-x / 128)
        const double e = KID == KM ? exp2_neg128(x, e2t) : exp2_neg128_clamped(x, e2t);
        return __builtin_fma(s, e, e);
    }
}

template <int KID>
__device__ __forceinline__ float kprep(float q, float diag) {
    if constexpr (KID == KC) {
        return (q == 0.0f) ? diag : __builtin_amdgcn_rsqf(q);
    } else if constexpr (KID == KG) {
        return __builtin_amdgcn_exp2f(-q);
    } else {
        const float qq = __builtin_fmaxf(q, 1e-30f);
        const float s = qq * __builtin_amdgcn_rsqf(qq);
        return (1.0f + s) * __builtin_amdgcn_exp2f(-1.4426950408889634f * s);
    }
}

// Coordinate scale of the prepared forms: the FP64 kernels above (f64) and
// the FP32 ones (Gaussian K = 2^(-q) there).
__host__ __device__ inline double kscale(int kid, double p0, bool f64) {
    if (kbase(kid) == KG) return __builtin_sqrt(p0 * 1.4426950408889634 * (f64 ? 128.0 : 1.0));
    if (kbase(kid) == KM) return 1.7320508075688772 * p0;
    return 1.0;
}

// Sum over the 64 lanes of four per-lane values at once (butterfly
// reduce-scatter): 7 exchanges instead of 4 x 6.  On return lane l holds
// the full sum of value (l >> 4) & 3.
template <typename T>
__device__ __forceinline__ T wave_sum4(T v0, T v1, T v2, T v3) {
    const int lane = threadIdx.x & 63;
    const bool b5 = lane & 32, b4 = lane & 16;
    T s0 = b5 ? v0 : v2, s1 = b5 ? v1 : v3;
    T k0 = b5 ? v2 : v0, k1 = b5 ? v3 : v1;
    T a0 = k0 + __shfl_xor(s0, 32, H2P_WAVE);
    T a1 = k1 + __shfl_xor(s1, 32, H2P_WAVE);
    T s = b4 ? a0 : a1, k = b4 ? a1 : a0;
    T c = k + __shfl_xor(s, 16, H2P_WAVE);
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) c += __shfl_xor(c, m, H2P_WAVE);
    return c;
}

// Sums over the 16 lanes of each lane group (lanes 16 g .. 16 g + 15) of
// four per-lane values at once.  On return the lanes with (l & 3) == 0 hold
// the full sum of value (l >> 2) & 3 of their group.
template <typename T>
__device__ __forceinline__ T group16_sum4(T v0, T v1, T v2, T v3) {
    const int lane = threadIdx.x & 63;
    const bool b3 = lane & 8, b2 = lane & 4;
    const T s0 = b3 ? v0 : v2, s1 = b3 ? v1 : v3;
    const T k0 = b3 ? v2 : v0, k1 = b3 ? v3 : v1;
    const T a0 = k0 + __shfl_xor(s0, 8, H2P_WAVE);
    const T a1 = k1 + __shfl_xor(s1, 8, H2P_WAVE);
    const T s = b2 ? a0 : a1, k = b2 ? a1 : a0;
    T c = k + __shfl_xor(s, 4, H2P_WAVE);
    c += __shfl_xor(c, 2, H2P_WAVE);
    c += __shfl_xor(c, 1, H2P_WAVE);
    return c;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, H2P_WAVE);
    return v;
}

__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, H2P_WAVE);
    return v;
}

}  // namespace h2p
