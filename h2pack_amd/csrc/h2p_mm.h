// Multi-RHS H2 product (H2P_matmul, src/H2Pack_matmul.c:455-583) on MFMA.
//
// Every phase of the matmul is a list of tasks of one form,
//     dst (M x nv)  (+)=  sum_s  op(A_s) (M x K_s) * X_s (K_s x nv),
// evaluated in the pull formulation of H2P_matmul_intmd_mult /
// H2P_matmul_dense_mult (one task per destination, each source block applied
// to the destination side only), so every output row has one writer and no
// partial buffers are needed:
//   upward   : Y0[node]      = U^T[node] * (PX[leaf rows] | Y0[children])
//   coupling : Y1[node]      = sum op(B) * (Y0[other] | PX[other leaf rows])
//   downward : Y1[children] += U[node] * Y1[node]           (level by level)
//   leaves   : PY[leaf rows] = U[leaf] * Y1[leaf] + sum op(D) * PX + sum op(B) * Y0
// A_s is a dense block (U, U^T, or a stored B/D block, transposed for the
// negative entries of the pair-to-block CSR) or, in JIT mode, kernel values
// evaluated into the tile: the kernels are symmetric, so op(B) = K(dst set,
// src set).  Panels are row-major with a row length of nvp (vectors padded to
// a multiple of 16).  The products run on v_mfma_f64_16x16x4f64: lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16] and receives
// D[4 r + l / 16][l % 16] in its register r.
#pragma once
#include "h2p_device.h"

namespace h2p {

enum { MM_PX = 0, MM_Y0 = 1, MM_Y1 = 2, MM_PY = 3 };
enum { MM_A_U = 0, MM_A_UT = 1, MM_A_BD = 2, MM_A_JIT = 3 };

struct MMSrc {
    long long a_off;   // dense: element offset of A in its blob
    int kind;          // MM_A_*
    int trans;         // dense: op(A) = A^T
    int lda;           // dense: row length of the stored A
    int K;             // inner dimension (vector entries)
    int c0, c1;        // JIT: first destination / source point (unified coordinates)
    int x_space, x_off;
    int shift;         // JIT: shift-table entry added to the source points (periodic images)
    int pad;
};

struct MMTask {
    int dst_space, dst_off, M, accumulate;
    int sbeg, send, pad0, pad1;
};

struct MMPanels {
    double *p[4];      // PX, Y0, Y1, PY
    int nvp;           // padded row length
};

// Tile geometry of one variant: RC rows of the destination per chunk (a
// multiple of 16, and of 48 for the 3x3 kernel so tiles hold whole points),
// KC inner entries per LDS stage, NV the widest panel the variant takes.
// The (RC/16) x (nvp/16) output tiles of a chunk are dealt to the four waves
// round-robin, so every wave issues MFMAs whatever the panel width.
template <int RC_, int KC_, int NV_>
struct MMCfg {
    static constexpr int RC = RC_, KC = KC_, NV = NV_;
    static_assert(RC % 16 == 0 && KC % 4 == 0 && NV % 16 == 0 && NV <= 64, "bad tile geometry");
};

#define H2P_MM_MAXNVP 64
#define H2P_MM_NVARIANT 4

// The variants the autotuner chooses from, per kernel dimension and panel
// width class (NV = 16, 32, 64); variant 0 is the default.
template <int KD, int NV, int V> struct MMVar;
template <int NV> struct MMVar<1, NV, 0> : MMCfg<64, 32, NV> {};
template <int NV> struct MMVar<1, NV, 1> : MMCfg<64, 64, NV> {};
template <int NV> struct MMVar<1, NV, 2> : MMCfg<128, 32, NV> {};
template <int NV> struct MMVar<1, NV, 3> : MMCfg<32, 64, NV> {};
template <int NV> struct MMVar<3, NV, 0> : MMCfg<48, 48, NV> {};
template <int NV> struct MMVar<3, NV, 1> : MMCfg<96, 24, NV> {};
template <int NV> struct MMVar<3, NV, 2> : MMCfg<96, 48, NV> {};
template <int NV> struct MMVar<3, NV, 3> : MMCfg<48, 24, NV> {};

// One stage of a task: the (RC x KC) operator tile of source si at inner
// offset kc and the (KC x nvp) operand tile, fetched into registers (loads
// or kernel evaluations), then stored to LDS.  The stages of a chunk run as
// a software pipeline: stage s + 1 is fetched before the MFMAs of stage s
// are issued, so its global-memory latency (or its kernel evaluations)
// overlaps the matrix work.
template <int KID, int KD, class G>
struct MMStage {
    static constexpr int RC = G::RC, KC = G::KC, NV = G::NV;
    // per thread: NA kernel-evaluated pairs (JIT; 9 entries each for the 3x3
    // kernel) or ND stored entries (dense), NX operand entries
    static constexpr int NA = ((RC / KD) * (KC / KD) + 255) / 256;
    static constexpr int ND = (RC * KC + 255) / 256;
    static constexpr int NX = (KC * NV + 255) / 256;
    static constexpr int NREG = (KD * KD * NA > ND) ? KD * KD * NA : ND;
    double a[NREG];
    double x[NX];

    __device__ __forceinline__ void fetch(const MMTask &t, const MMSrc &s, int mc, int kc, int nvp,
                                          const double *__restrict__ Ublob, const double *__restrict__ UTblob,
                                          const double *__restrict__ BDblob, const double *__restrict__ cx,
                                          long long cstride, const KParam &kp,
                                          const double *__restrict__ shifts, const MMPanels &P) {
        // rows of cx: x, y, z and, for the 3x3 kernel, radius and reciprocal radius
        const double *cy = cx + cstride, *cz = cx + 2 * cstride, *ca = cx + 3 * cstride, *cia = cx + 4 * cstride;
        if (s.kind == MM_A_JIT) {
            constexpr int RP = RC / KD, CP = KC / KD;
            const int p0 = mc / KD, q0 = kc / KD;
            const int np0 = t.M / KD - p0, nq0 = s.K / KD - q0;
            const double hx = shifts[3 * s.shift], hy = shifts[3 * s.shift + 1], hz = shifts[3 * s.shift + 2];
#pragma unroll
            for (int u = 0; u < NA; u++) {
                const int e = threadIdx.x + 256 * u;
                const int i = e / CP, j = e - i * CP;
                const bool ok = e < RP * CP && i < np0 && j < nq0;
                const int gi = s.c0 + p0 + (ok ? i : 0), gj = s.c1 + q0 + (ok ? j : 0);
                const double dx = cx[gi] - (cx[gj] + hx), dy = cy[gi] - (cy[gj] + hy), dz = cz[gi] - (cz[gj] + hz);
                if constexpr (KD == 1) {
                    a[u] = ok ? kval<KID>(dx * dx + dy * dy + dz * dz, kp) : 0.0;
                } else {
                    double t1, t2, ux, uy, uz;
                    rpy_coef_rcp(dx, dy, dz, ca[gi], ca[gj], cia[gi], cia[gj], kp.p1, t1, t2, ux, uy, uz);
                    const double uu[3] = {ux, uy, uz};
#pragma unroll
                    for (int p = 0; p < 3; p++)
#pragma unroll
                        for (int q = 0; q < 3; q++)
                            a[9 * u + 3 * p + q] = ok ? t2 * uu[p] * uu[q] + (p == q ? t1 : 0.0) : 0.0;
                }
            }
        } else {
            const double *A = (s.kind == MM_A_U ? Ublob : s.kind == MM_A_UT ? UTblob : BDblob) + s.a_off;
#pragma unroll
            for (int u = 0; u < ND; u++) {
                const int e = threadIdx.x + 256 * u;
                int i, j;
                if (!s.trans) { i = e / KC; j = e - i * KC; }      // consecutive lanes along a row of A
                else { j = e / RC; i = e - j * RC; }              // consecutive lanes along a row of A^T
                const int m = mc + i, k = kc + j;
                const bool ok = e < RC * KC && m < t.M && k < s.K;
                a[u] = ok ? (s.trans ? A[(long long)k * s.lda + m] : A[(long long)m * s.lda + k]) : 0.0;
            }
        }
        const double *X = P.p[s.x_space] + (long long)s.x_off * nvp;
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int e = threadIdx.x + 256 * u;
            const int k = e / nvp, c = e - k * nvp;
            x[u] = (e < KC * nvp && kc + k < s.K) ? X[(long long)(kc + k) * nvp + c] : 0.0;
        }
    }

    __device__ __forceinline__ void store(const MMSrc &s, int nvp, double (*As)[KC + 1], double (*Xs)[NV + 1]) const {
        if (s.kind == MM_A_JIT) {
            constexpr int RP = RC / KD, CP = KC / KD;
#pragma unroll
            for (int u = 0; u < NA; u++) {
                const int e = threadIdx.x + 256 * u;
                if (e >= RP * CP) break;
                const int i = e / CP, j = e - i * CP;
                if constexpr (KD == 1) {
                    As[i][j] = a[u];
                } else {
#pragma unroll
                    for (int p = 0; p < 3; p++)
#pragma unroll
                        for (int q = 0; q < 3; q++) As[3 * i + p][3 * j + q] = a[9 * u + 3 * p + q];
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < ND; u++) {
                const int e = threadIdx.x + 256 * u;
                if (e >= RC * KC) break;
                int i, j;
                if (!s.trans) { i = e / KC; j = e - i * KC; }
                else { j = e / RC; i = e - j * RC; }
                As[i][j] = a[u];
            }
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int e = threadIdx.x + 256 * u;
            if (e >= KC * nvp) break;
            const int k = e / nvp, c = e - k * nvp;
            Xs[k][c] = x[u];
        }
    }
};

// the stage after (si, kc): next inner chunk, or the first chunk of the next
// source with K > 0
__device__ __forceinline__ void mm_next(const MMSrc *__restrict__ srcs, int send, int KCs, int &si, int &kc) {
    kc += KCs;
    if (kc < srcs[si].K) return;
    kc = 0;
    for (si++; si < send && srcs[si].K <= 0; si++) {}
}

template <int KID, int KD, class G>
__device__ __forceinline__ void mm_body(const MMTask &t, const MMSrc *__restrict__ srcs,
                                        const double *__restrict__ Ublob,
                                        const double *__restrict__ UTblob,
                                        const double *__restrict__ BDblob,
                                        const double *__restrict__ cx, long long cstride,
                                        const KParam &kp, const double *__restrict__ shifts,
                                        const MMPanels &P) {
    constexpr int RC = G::RC, KC = G::KC, NV = G::NV;
    constexpr int MT = RC / 16;                         // M tiles per chunk
    constexpr int SLOTS = (MT * (NV / 16) + 3) / 4;     // output tiles per wave, at most
    __shared__ double As[RC][KC + 1];
    __shared__ double Xs[KC][NV + 1];
    const int nvp = P.nvp, NT = nvp >> 4;               // nvp <= NV (host-checked)
    const int T = MT * NT;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lq = lane >> 4;
    double *dst = P.p[t.dst_space] + (long long)t.dst_off * nvp;
    int s_first = t.sbeg;
    while (s_first < t.send && srcs[s_first].K <= 0) s_first++;
    MMStage<KID, KD, G> stg;
    for (int mc = 0; mc < t.M; mc += RC) {
        d4 acc[SLOTS];
#pragma unroll
        for (int j = 0; j < SLOTS; j++) acc[j] = (d4){0.0, 0.0, 0.0, 0.0};
        int si = s_first, kc = 0;
        if (si < t.send)
            stg.fetch(t, srcs[si], mc, kc, nvp, Ublob, UTblob, BDblob, cx, cstride, kp, shifts, P);
        while (si < t.send) {
            const MMSrc cur = srcs[si];
            __syncthreads();                 // the previous stage's MFMAs are done with LDS
            stg.store(cur, nvp, As, Xs);
            __syncthreads();
            int nsi = si, nkc = kc;
            mm_next(srcs, t.send, KC, nsi, nkc);
            if (nsi < t.send)                // next stage in flight during the MFMAs
                stg.fetch(t, srcs[nsi], mc, nkc, nvp, Ublob, UTblob, BDblob, cx, cstride, kp, shifts, P);
#pragma unroll
            for (int j = 0; j < SLOTS; j++) {
                const int tt = w + 4 * j;
                if (tt < T) {
                    const int mt = tt / NT, nt = tt - mt * NT;
#pragma unroll
                    for (int kk = 0; kk < KC; kk += 4)
                        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[mt * 16 + li][kk + lq],
                                                                      Xs[kk + lq][nt * 16 + li], acc[j], 0, 0, 0);
                }
            }
            si = nsi;
            kc = nkc;
        }
        // ---- write the chunk: register r of lane (li, lq) holds row 4 r + lq
        // of its M tile, column li (the f64 16x16x4 output layout)
#pragma unroll
        for (int j = 0; j < SLOTS; j++) {
            const int tt = w + 4 * j;
            if (tt < T) {
                const int mt = tt / NT, nt = tt - mt * NT;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m = mc + mt * 16 + 4 * r + lq;
                    if (m < t.M) {
                        double *o = dst + (long long)m * nvp + nt * 16 + li;
                        *o = t.accumulate ? *o + acc[j][r] : acc[j][r];
                    }
                }
            }
        }
        __syncthreads();                     // LDS reuse by the next chunk
    }
}

template <int KID, int KD, class G>
__global__ void __launch_bounds__(256) k_mm(const MMTask *__restrict__ tasks,
                                            const MMSrc *__restrict__ srcs,
                                            const double *__restrict__ Ublob,
                                            const double *__restrict__ UTblob,
                                            const double *__restrict__ BDblob,
                                            const double *__restrict__ cx, long long cstride,
                                            KParam kp, const double *__restrict__ shifts, MMPanels P) {
    mm_body<KID, KD, G>(tasks[blockIdx.x], srcs, Ublob, UTblob, BDblob, cx, cstride, kp, shifts, P);
}

// ---------------------------------------------------------------------------
// FP32-mixed products (handles of precision FP32MIX): the B and D sources
// -- kernel values evaluated in FP32 (JIT: the FP64 coordinate difference and
// squared distance rounded once to FP32, then the hardware exp2 / rsqrt as
// in the FP32-mixed matvec) or blocks stored in FP32 (AOT) -- and their
// operand rows rounded to FP32 run on v_mfma_f32_16x16x4f32; the FP32 tile
// sums of every 16 inner entries (four products, counted from the start of
// each source) are folded into FP64 accumulators, and the output panels
// stay FP64.  The fold points do not depend on the tile variant (KC is a
// multiple of 16), so every variant gives bitwise the same result.  The U phases keep the FP64 kernel.  Only
// B/D sources reach this kernel (the host splits the leaf tasks into a U
// part and a B/D part).
//
// Lane layouts: v_mfma_f32_16x16x4f32 takes A[l % 16][l / 16] and
// B[l / 16][l % 16] like the FP64 product, but returns D[4 (l / 16) + r]
// [l % 16] in register r, where the FP64 product returns row 4 r + l / 16.
// Feeding the A rows in the order sigma(4 a + b) = 4 b + a makes the FP32
// tile land in the FP64 accumulator layout, so the fold is lane-local.
__device__ __forceinline__ int mm32_sigma(int i) { return 4 * (i & 3) + (i >> 2); }

template <int KID, class G>
struct MMStage32 {
    static constexpr int RC = G::RC, KC = G::KC, NV = G::NV;
    static_assert(KC % 16 == 0, "FP32 stages fold every 16 inner entries");
    static constexpr int NA = (RC * KC + 255) / 256;
    static constexpr int NX = (KC * NV + 255) / 256;
    float a[NA];
    float x[NX];

    __device__ __forceinline__ void fetch(const MMTask &t, const MMSrc &s, int mc, int kc, int nvp,
                                          const float *__restrict__ BD32, const double *__restrict__ cx,
                                          long long cstride, float p0, const double *__restrict__ shifts,
                                          const MMPanels &P) {
        const double *cy = cx + cstride, *cz = cx + 2 * cstride;
        if (s.kind == MM_A_JIT) {
            const int np0 = t.M - mc, nq0 = s.K - kc;
            const double hx = shifts[3 * s.shift], hy = shifts[3 * s.shift + 1], hz = shifts[3 * s.shift + 2];
#pragma unroll
            for (int u = 0; u < NA; u++) {
                const int e = threadIdx.x + 256 * u;
                const int i = e / KC, j = e - i * KC;
                const bool ok = e < RC * KC && i < np0 && j < nq0;
                const int gi = s.c0 + mc + (ok ? i : 0), gj = s.c1 + kc + (ok ? j : 0);
                const double dx = cx[gi] - (cx[gj] + hx), dy = cy[gi] - (cy[gj] + hy), dz = cz[gi] - (cz[gj] + hz);
                const double r2 = dx * dx + dy * dy + dz * dz;
                a[u] = ok ? kvalf<KID>((float)r2, p0) : 0.0f;
            }
        } else {
            const float *A = BD32 + s.a_off;
#pragma unroll
            for (int u = 0; u < NA; u++) {
                const int e = threadIdx.x + 256 * u;
                int i, j;
                if (!s.trans) { i = e / KC; j = e - i * KC; }
                else { j = e / RC; i = e - j * RC; }
                const int m = mc + i, k = kc + j;
                const bool ok = e < RC * KC && m < t.M && k < s.K;
                a[u] = ok ? (s.trans ? A[(long long)k * s.lda + m] : A[(long long)m * s.lda + k]) : 0.0f;
            }
        }
        const double *X = P.p[s.x_space] + (long long)s.x_off * nvp;
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int e = threadIdx.x + 256 * u;
            const int k = e / nvp, c = e - k * nvp;
            x[u] = (e < KC * nvp && kc + k < s.K) ? (float)X[(long long)(kc + k) * nvp + c] : 0.0f;
        }
    }

    __device__ __forceinline__ void store(const MMSrc &s, int nvp, float (*As)[KC + 1], float (*Xs)[NV + 1]) const {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int e = threadIdx.x + 256 * u;
            if (e >= RC * KC) break;
            int i, j;
            if (s.kind == MM_A_JIT || !s.trans) { i = e / KC; j = e - i * KC; }
            else { j = e / RC; i = e - j * RC; }
            As[i][j] = a[u];
        }
#pragma unroll
        for (int u = 0; u < NX; u++) {
            const int e = threadIdx.x + 256 * u;
            if (e >= KC * nvp) break;
            const int k = e / nvp, c = e - k * nvp;
            Xs[k][c] = x[u];
        }
    }
};

template <int KID, class G>
__device__ __forceinline__ void mm32_body(const MMTask &t, const MMSrc *__restrict__ srcs,
                                          const float *__restrict__ BD32, const double *__restrict__ cx,
                                          long long cstride, float p0, const double *__restrict__ shifts,
                                          const MMPanels &P) {
    constexpr int RC = G::RC, KC = G::KC, NV = G::NV;
    constexpr int MT = RC / 16;
    constexpr int SLOTS = (MT * (NV / 16) + 3) / 4;
    __shared__ float As[RC][KC + 1];
    __shared__ float Xs[KC][NV + 1];
    const int nvp = P.nvp, NT = nvp >> 4;
    const int T = MT * NT;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lq = lane >> 4, ls = mm32_sigma(li);
    double *dst = P.p[t.dst_space] + (long long)t.dst_off * nvp;
    int s_first = t.sbeg;
    while (s_first < t.send && srcs[s_first].K <= 0) s_first++;
    MMStage32<KID, G> stg;
    for (int mc = 0; mc < t.M; mc += RC) {
        d4 acc[SLOTS];
#pragma unroll
        for (int j = 0; j < SLOTS; j++) acc[j] = (d4){0.0, 0.0, 0.0, 0.0};
        int si = s_first, kc = 0;
        if (si < t.send) stg.fetch(t, srcs[si], mc, kc, nvp, BD32, cx, cstride, p0, shifts, P);
        while (si < t.send) {
            const MMSrc cur = srcs[si];
            __syncthreads();
            stg.store(cur, nvp, As, Xs);
            __syncthreads();
            int nsi = si, nkc = kc;
            mm_next(srcs, t.send, KC, nsi, nkc);
            if (nsi < t.send) stg.fetch(t, srcs[nsi], mc, nkc, nvp, BD32, cx, cstride, p0, shifts, P);
#pragma unroll
            for (int j = 0; j < SLOTS; j++) {
                const int tt = w + 4 * j;
                if (tt < T) {
                    const int mt = tt / NT, nt = tt - mt * NT;
#pragma unroll
                    for (int k0 = 0; k0 < KC; k0 += 16) {
                        f4 a32 = (f4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                        for (int kk = k0; kk < k0 + 16; kk += 4)
                            a32 = __builtin_amdgcn_mfma_f32_16x16x4f32(As[mt * 16 + ls][kk + lq],
                                                                       Xs[kk + lq][nt * 16 + li], a32, 0, 0, 0);
#pragma unroll
                        for (int r = 0; r < 4; r++) acc[j][r] += (double)a32[r];
                    }
                }
            }
            si = nsi;
            kc = nkc;
        }
#pragma unroll
        for (int j = 0; j < SLOTS; j++) {
            const int tt = w + 4 * j;
            if (tt < T) {
                const int mt = tt / NT, nt = tt - mt * NT;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int m = mc + mt * 16 + 4 * r + lq;
                    if (m < t.M) {
                        double *o = dst + (long long)m * nvp + nt * 16 + li;
                        *o = t.accumulate ? *o + acc[j][r] : acc[j][r];
                    }
                }
            }
        }
        __syncthreads();
    }
}

template <int KID, class G>
__global__ void __launch_bounds__(256) k_mm32(const MMTask *__restrict__ tasks, const MMSrc *__restrict__ srcs,
                                              const float *__restrict__ BD32, const double *__restrict__ cx,
                                              long long cstride, float p0, const double *__restrict__ shifts,
                                              MMPanels P) {
    mm32_body<KID, G>(tasks[blockIdx.x], srcs, BD32, cx, cstride, p0, shifts, P);
}

// Caller panel (original order, row- or column-major, ld) -> PX (permuted,
// row-major nvp, zero padding), for columns [v0, v0 + nv).
__global__ void k_mm_gather(const double *__restrict__ X, long long ldx, int row_major, int v0, int nv,
                            const int *__restrict__ fwd, int N, double *__restrict__ PX, int nvp) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)N * nvp) return;
    const int i = (int)(e / nvp), c = (int)(e - (long long)i * nvp);
    double v = 0.0;
    if (c < nv) {
        const long long src = fwd[i];
        v = row_major ? X[src * ldx + v0 + c] : X[(long long)(v0 + c) * ldx + src];
    }
    PX[e] = v;
}

// PY -> caller panel: Y[fwd[i], v0 + c] = PY[i, c] (the backward permutation).
__global__ void k_mm_scatter(const double *__restrict__ PY, int nvp, const int *__restrict__ fwd, int N,
                             double *__restrict__ Y, long long ldy, int row_major, int v0, int nv) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)N * nv) return;
    const int i = (int)(e / nv), c = (int)(e - (long long)i * nv);
    const long long dstr = fwd[i];
    const double v = PY[(long long)i * nvp + c];
    if (row_major) Y[dstr * ldy + v0 + c] = v;
    else Y[(long long)(v0 + c) * ldy + dstr] = v;
}

}  // namespace h2p
