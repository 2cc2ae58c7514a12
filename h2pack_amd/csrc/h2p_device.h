// Device-side kernel functions and wave helpers for the MI355X H2 engine.
//
// Kernel formulas follow src/H2Pack_3D_kernels.h: Coulomb :52-96 (1/r, the
// r = 0 entry is param[0] or 0), Gaussian :169-209 (exp(-l r^2)), Matern32
// :388-432 ((1 - t) e^t with t = -sqrt(3) l r), RPY :893-959 (3x3 tensor
// t1 I + t2 r^ r^T with three regimes of r against the radii).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define H2P_WAVE 64

namespace h2p {

enum { KC = 0, KG = 1, KM = 2, KR = 3 };

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
