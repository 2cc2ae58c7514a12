// Base-2 exponential 2^(-x / 128) for x >= 0 in FP64, table-driven (Tang's
// method: P. T. P. Tang, "Table-driven implementation of the exponential
// function in IEEE floating-point arithmetic", ACM TOMS 15(2), 1989, here in
// base 2 with 128 table entries).  Used by the FP64 Gaussian and Matern-3/2
// kernel evaluations of the JIT block launches, whose arguments arrive
// already multiplied by 128 (the Gaussian through its coordinate scale);
// the reference evaluates the same kernels through its SIMD library's exp
// (src/H2Pack_3D_kernels.h:212-268, 433-497).
//
// With k = round(-x) and u = -x - k in [-1/2, 1/2] (both exact):
//     2^(-x/128) = 2^floor(k / 128) * 2^((k mod 128) / 128) * 2^(u / 128).
// k comes from the magic-number rounding t = 1.5 2^52 - x: t's last place is
// 1, so its 52-bit fraction field holds 2^51 + k, whose low 7 bits index
// the table and whose bits 7..38 are floor(k / 128) as a 32-bit integer.
// 2^(u/128) = 1 + u P(u) with P of degree 3, minimax for the relative error
// over u in [-1/2, 1/2] (truncation error 7.6e-17); the table holds
// 2^(j/128) rounded to nearest.  Max relative error ~2.3e-16 (checked
// against long double exp2 on the host).
//
// Range: exact reduction and a correct exponent need x < 2^38.  Callers
// that cannot bound x clamp it first (exp2_neg128_clamped): the result
// underflows to 0 well before.
//
// Per value: 3 adds for the reduction (each with at most one scalar
// operand), 2 integer operations for the table address, 1 bit-field
// extraction for the exponent, 5 FMA/multiplies for the polynomial and the
// table factor, 1 ldexp: 12 VALU instructions (the former degree-10
// Chebyshev form on [-1/2, 1/2] took 14).
#pragma once
#include <stdint.h>

#ifndef H2P_HD
#define H2P_HD
#endif

// 2^(j/128), j = 0..127, correctly rounded
#define H2P_E2_TABLE \
    0x1.0000000000000p+0, 0x1.0163da9fb3335p+0, 0x1.02c9a3e778061p+0, 0x1.04315e86e7f85p+0,  \
    0x1.059b0d3158574p+0, 0x1.0706b29ddf6dep+0, 0x1.0874518759bc8p+0, 0x1.09e3ecac6f383p+0,  \
    0x1.0b5586cf9890fp+0, 0x1.0cc922b7247f7p+0, 0x1.0e3ec32d3d1a2p+0, 0x1.0fb66affed31bp+0,  \
    0x1.11301d0125b51p+0, 0x1.12abdc06c31ccp+0, 0x1.1429aaea92de0p+0, 0x1.15a98c8a58e51p+0,  \
    0x1.172b83c7d517bp+0, 0x1.18af9388c8deap+0, 0x1.1a35beb6fcb75p+0, 0x1.1bbe084045cd4p+0,  \
    0x1.1d4873168b9aap+0, 0x1.1ed5022fcd91dp+0, 0x1.2063b88628cd6p+0, 0x1.21f49917ddc96p+0,  \
    0x1.2387a6e756238p+0, 0x1.251ce4fb2a63fp+0, 0x1.26b4565e27cddp+0, 0x1.284dfe1f56381p+0,  \
    0x1.29e9df51fdee1p+0, 0x1.2b87fd0dad990p+0, 0x1.2d285a6e4030bp+0, 0x1.2ecafa93e2f56p+0,  \
    0x1.306fe0a31b715p+0, 0x1.32170fc4cd831p+0, 0x1.33c08b26416ffp+0, 0x1.356c55f929ff1p+0,  \
    0x1.371a7373aa9cbp+0, 0x1.38cae6d05d866p+0, 0x1.3a7db34e59ff7p+0, 0x1.3c32dc313a8e5p+0,  \
    0x1.3dea64c123422p+0, 0x1.3fa4504ac801cp+0, 0x1.4160a21f72e2ap+0, 0x1.431f5d950a897p+0,  \
    0x1.44e086061892dp+0, 0x1.46a41ed1d0057p+0, 0x1.486a2b5c13cd0p+0, 0x1.4a32af0d7d3dep+0,  \
    0x1.4bfdad5362a27p+0, 0x1.4dcb299fddd0dp+0, 0x1.4f9b2769d2ca7p+0, 0x1.516daa2cf6642p+0,  \
    0x1.5342b569d4f82p+0, 0x1.551a4ca5d920fp+0, 0x1.56f4736b527dap+0, 0x1.58d12d497c7fdp+0,  \
    0x1.5ab07dd485429p+0, 0x1.5c9268a5946b7p+0, 0x1.5e76f15ad2148p+0, 0x1.605e1b976dc09p+0,  \
    0x1.6247eb03a5585p+0, 0x1.6434634ccc320p+0, 0x1.6623882552225p+0, 0x1.68155d44ca973p+0,  \
    0x1.6a09e667f3bcdp+0, 0x1.6c012750bdabfp+0, 0x1.6dfb23c651a2fp+0, 0x1.6ff7df9519484p+0,  \
    0x1.71f75e8ec5f74p+0, 0x1.73f9a48a58174p+0, 0x1.75feb564267c9p+0, 0x1.780694fde5d3fp+0,  \
    0x1.7a11473eb0187p+0, 0x1.7c1ed0130c132p+0, 0x1.7e2f336cf4e62p+0, 0x1.80427543e1a12p+0,  \
    0x1.82589994cce13p+0, 0x1.8471a4623c7adp+0, 0x1.868d99b4492edp+0, 0x1.88ac7d98a6699p+0,  \
    0x1.8ace5422aa0dbp+0, 0x1.8cf3216b5448cp+0, 0x1.8f1ae99157736p+0, 0x1.9145b0b91ffc6p+0,  \
    0x1.93737b0cdc5e5p+0, 0x1.95a44cbc8520fp+0, 0x1.97d829fde4e50p+0, 0x1.9a0f170ca07bap+0,  \
    0x1.9c49182a3f090p+0, 0x1.9e86319e32323p+0, 0x1.a0c667b5de565p+0, 0x1.a309bec4a2d33p+0,  \
    0x1.a5503b23e255dp+0, 0x1.a799e1330b358p+0, 0x1.a9e6b5579fdbfp+0, 0x1.ac36bbfd3f37ap+0,  \
    0x1.ae89f995ad3adp+0, 0x1.b0e07298db666p+0, 0x1.b33a2b84f15fbp+0, 0x1.b59728de5593ap+0,  \
    0x1.b7f76f2fb5e47p+0, 0x1.ba5b030a1064ap+0, 0x1.bcc1e904bc1d2p+0, 0x1.bf2c25bd71e09p+0,  \
    0x1.c199bdd85529cp+0, 0x1.c40ab5fffd07ap+0, 0x1.c67f12e57d14bp+0, 0x1.c8f6d9406e7b5p+0,  \
    0x1.cb720dcef9069p+0, 0x1.cdf0b555dc3fap+0, 0x1.d072d4a07897cp+0, 0x1.d2f87080d89f2p+0,  \
    0x1.d5818dcfba487p+0, 0x1.d80e316c98398p+0, 0x1.da9e603db3285p+0, 0x1.dd321f301b460p+0,  \
    0x1.dfc97337b9b5fp+0, 0x1.e264614f5a129p+0, 0x1.e502ee78b3ff6p+0, 0x1.e7a51fbc74c83p+0,  \
    0x1.ea4afa2a490dap+0, 0x1.ecf482d8e67f1p+0, 0x1.efa1bee615a27p+0, 0x1.f252b376bba97p+0,  \
    0x1.f50765b6e4540p+0, 0x1.f7bfdad9cbe14p+0, 0x1.fa7c1819e90d8p+0, 0x1.fd3c22b8f71f1p+0

namespace h2p {

#define H2P_E2_N 128

// 2^(-x / 128) for 0 <= x < 2^38 (see above); tab = H2P_E2_TABLE.
H2P_HD inline double exp2_neg128(double x, const double *tab) {
    const double magic = 0x1.8p52;
    const double t = magic - x;
    const double kd = t - magic;                   // k = round(-x), exact
    const double u = -x - kd;                      // exact
    uint64_t bits;
    __builtin_memcpy(&bits, &t, sizeof bits);
    const double T = tab[(uint32_t)bits & (H2P_E2_N - 1)];
    const int m = (int)(uint32_t)(bits >> 7);      // floor(k / 128)
    double p = __builtin_fma(u, 0x1.3b2ab31427f87p-35, 0x1.c6b09b16cfcd0p-26);
    p = __builtin_fma(p, u, 0x1.ebfbdff82cc9dp-17);
    p = __builtin_fma(p, u, 0x1.62e42fefa3686p-8);
    return __builtin_ldexp(__builtin_fma(T * u, p, T), m);
}

// The same for any x >= 0 (or +inf): x is clamped to 128 * 1100 first,
// where the result is already 0.
H2P_HD inline double exp2_neg128_clamped(double x, const double *tab) {
    return exp2_neg128(__builtin_fmin(x, 128.0 * 1100.0), tab);
}

}  // namespace h2p
