// Global kernels of the MI355X H2 engine (matvec path).
//
// Vector spaces on the device:
//   xin   = [ px (N, permuted input) | y0 (S, skeleton coefficients) ]
//   y1    = S skeleton accumulators (downward sweep)
//   arena = partial outputs of the grouped B/D block launch
// Skeleton offsets are assigned so that the children of every node occupy a
// contiguous range in child order; the stacked operand of a non-leaf U is
// therefore one contiguous slice of y0 / y1.
#pragma once
#include <type_traits>

#include "h2p_device.h"

namespace h2p {

// One B or D block of the grouped launch.
struct BlkTask {
    long long data_off;   // AOT: element offset of the row-major block
    long long out0, out1; // arena offsets of the two partial outputs (out1 < 0: none)
    int nrow, ncol;       // block dimensions in vector entries (points * kd)
    int c0, c1;           // JIT: first point of each set in the unified coordinates
    int n0, n1;           // JIT: point counts
    int in0, in1;         // xin offsets of the operands (in0: ncol long, in1: nrow long)
    int flags;            // bit 0: has out1
    int shift;            // index into the shift table, added to set 1's coordinates
                          // (periodic images; entry 0 is the zero shift)
};

// Upward-sweep task: y0[out] = U^T in, U row-major nrow x r.
#define H2P_FWD_MIN_WG 256    // a level's upward tasks narrow their column chunks to reach this many (one per CU)
struct FwdTask {
    long long u_off;
    int nrow, r, in_off, out_off;
    int col0;         // first of the output columns this task computes
    int ncol;         // at most this many columns from col0: 64, 128 or 256
};

// Segmented reduction into y1: y1[dst + e] = sum_k arena[contrib[k] + e].
struct RedTask {
    int dst, len, cbeg, cend;
};

// ---------------------------------------------------------------------------
__global__ void k_gather(const double *__restrict__ x, const int *__restrict__ idx,
                         double *__restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = x[idx[i]];
}

// Fixed-order segmented reduction of the block partials into y1.
// v + arena[contrib[k] + e] for k = kb, kb + step, ... < ke, added in list
// order.  The offsets and values of H2P_GATHER_AHEAD entries are loaded
// before their adds, so that many loads are in flight instead of one
// offset-then-value chain per entry; the sum is the plain loop's, bit for bit.
#define H2P_GATHER_AHEAD 8
__device__ __forceinline__ double ordered_gather_sum(double v, const long long *__restrict__ contrib,
                                                     const double *__restrict__ arena, int kb, int ke, int step,
                                                     int e) {
    int k = kb;
    for (; k + (H2P_GATHER_AHEAD - 1) * step < ke; k += H2P_GATHER_AHEAD * step) {
        double a[H2P_GATHER_AHEAD];
#pragma unroll
        for (int u = 0; u < H2P_GATHER_AHEAD; u++) a[u] = arena[contrib[k + u * step] + e];
#pragma unroll
        for (int u = 0; u < H2P_GATHER_AHEAD; u++) v += a[u];
    }
    for (; k < ke; k += step) v += arena[contrib[k] + e];
    return v;
}

__device__ __forceinline__ void reduce_y1_body(const RedTask &t, const long long *__restrict__ contrib,
                                               const double *__restrict__ arena, double *__restrict__ y1) {
    for (int e = threadIdx.x; e < t.len; e += blockDim.x)
        y1[t.dst + e] = ordered_gather_sum(0.0, contrib, arena, t.cbeg, t.cend, 1, e);
}

__global__ void __launch_bounds__(256) k_reduce_y1(const RedTask *__restrict__ tasks,
                                                   const long long *__restrict__ contrib,
                                                   const double *__restrict__ arena,
                                                   double *__restrict__ y1) {
    reduce_y1_body(tasks[blockIdx.x], contrib, arena, y1);
}

// ---------------------------------------------------------------------------
// Sweep kernels.
//
// Upward sweep (src/H2Pack_matvec.c:95-177): one workgroup per node and
// column chunk of its U (256 columns, or 128 / 64 on a level with fewer
// workgroups than CUs, so a node with a large basis spreads over several
// workgroups); lanes own the chunk's columns (NS strips of 64, only the
// strips that hold columns), the four waves take contiguous quarters of the
// rows and load UR = 32 / NS rows x NS strips before their FMAs (32 loads
// in flight per wave whatever the chunk width, 24 at NS = 3; then 8-row
// batches, then single rows; rows in order, so each column's sum runs over
// the rows in order whatever the unrolling and the chunk width); the four
// wave partials are added in LDS in a fixed order.
// R rows x NS strips loaded, then their FMAs in row order.
template <int R, int NS>
__device__ __forceinline__ void fwd2_rows(const double *__restrict__ Ut, int ld, const double *__restrict__ in, int i,
                                          const int *cc, double *acc) {
    double u[R][NS];
#pragma unroll
    for (int a = 0; a < R; a++)
#pragma unroll
        for (int k = 0; k < NS; k++) u[a][k] = Ut[(long long)(i + a) * ld + cc[k]];
#pragma unroll
    for (int a = 0; a < R; a++) {
        const double v = in[i + a];
#pragma unroll
        for (int k = 0; k < NS; k++) acc[k] += u[a][k] * v;
    }
}

template <int NS>
__device__ __forceinline__ void fwd2_strips(const FwdTask &t, const double *__restrict__ U,
                                            double *__restrict__ xin, double (*red)[4][H2P_WAVE]) {
    constexpr int UR = NS == 3 ? 8 : 32 / NS;
    // the wave index is the same on every lane: said so, the row addresses
    // stay scalar and each load needs only its column offset per lane
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double *Ut = U + t.u_off;
    const double *in = xin + t.in_off;
    const int q = (t.nrow + 3) >> 2;
    const int i0 = w * q, i1 = min(t.nrow, i0 + q);
    const int cs = t.col0;
    double acc[NS];
    int cc[NS];
    bool ok[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        acc[k] = 0.0;
        cc[k] = cs + k * H2P_WAVE + lane;
        ok[k] = cc[k] < t.r;
        cc[k] = ok[k] ? cc[k] : 0;
    }
    int i = i0;
    for (; i + UR - 1 < i1; i += UR) fwd2_rows<UR, NS>(Ut, t.r, in, i, cc, acc);
    if (UR > 8)
        for (; i + 7 < i1; i += 8) fwd2_rows<8, NS>(Ut, t.r, in, i, cc, acc);
    for (; i < i1; i++) {
        const double v = in[i];
#pragma unroll
        for (int k = 0; k < NS; k++) acc[k] += Ut[(long long)i * t.r + cc[k]] * v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NS; k++) red[k][w][lane] = acc[k];
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int k = 0; k < NS; k++)
            if (ok[k])
                xin[t.out_off + cc[k]] = (red[k][0][lane] + red[k][1][lane]) + (red[k][2][lane] + red[k][3][lane]);
    }
}

__device__ __forceinline__ void fwd2_body(const FwdTask &t, const double *__restrict__ U,
                                          double *__restrict__ xin) {
    __shared__ double red[4][4][H2P_WAVE];
    const int ns = (min(t.r - t.col0, t.ncol) + H2P_WAVE - 1) / H2P_WAVE;
    if (ns <= 1) fwd2_strips<1>(t, U, xin, red);
    else if (ns == 2) fwd2_strips<2>(t, U, xin, red);
    else if (ns == 3) fwd2_strips<3>(t, U, xin, red);
    else fwd2_strips<4>(t, U, xin, red);
}

__global__ void __launch_bounds__(256) k_fwd2(const FwdTask *__restrict__ tasks,
                                              const double *__restrict__ U,
                                              double *__restrict__ xin) {
    fwd2_body(tasks[blockIdx.x], U, xin);
}

// Downward sweep (src/H2Pack_matvec.c:767-861), one workgroup per 256-row chunk of a node:
// yp[child rows] = U (y1 + yp)[node], U^T stored row-major.  The push-down
// buffer yp is written exactly once per row (each child has one parent).
struct BwdTask2 {
    long long ut_off;
    int nrow, r, y1_off, push, out_off, row0;
};
#define H2P_MAXR 2048
// sum_c UT[c * ld] * yv[c] for one row of U (U^T stored row-major, so the
// lanes' rows are contiguous for each c): four accumulators taking every
// fourth c, 16 values loaded before their FMAs; each accumulator's FMAs run
// over its c in order, so the sum is the same for any unrolling.
__device__ __forceinline__ double ut_row_dot(const double *__restrict__ ut, int ld, int r, const double *yv) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int c = 0;
    for (; c + 15 < r; c += 16) {
        double u[16];
#pragma unroll
        for (int j = 0; j < 16; j++) u[j] = ut[(long long)(c + j) * ld];
#pragma unroll
        for (int j = 0; j < 16; j += 4) {
            a0 += u[j] * yv[c + j];
            a1 += u[j + 1] * yv[c + j + 1];
            a2 += u[j + 2] * yv[c + j + 2];
            a3 += u[j + 3] * yv[c + j + 3];
        }
    }
    for (; c + 3 < r; c += 4) {
        a0 += ut[(long long)c * ld] * yv[c];
        a1 += ut[(long long)(c + 1) * ld] * yv[c + 1];
        a2 += ut[(long long)(c + 2) * ld] * yv[c + 2];
        a3 += ut[(long long)(c + 3) * ld] * yv[c + 3];
    }
    for (; c < r; c++) a0 += ut[(long long)c * ld] * yv[c];
    return (a0 + a1) + (a2 + a3);
}

__device__ __forceinline__ void bwd2_body(const BwdTask2 &t, const double *__restrict__ UT,
                                          const double *__restrict__ y1, double *__restrict__ yp) {
    __shared__ double yv[H2P_MAXR];
    for (int c = threadIdx.x; c < t.r; c += blockDim.x)
        yv[c] = y1[t.y1_off + c] + (t.push ? yp[t.y1_off + c] : 0.0);
    __syncthreads();
    const int row = t.row0 + threadIdx.x;
    if (row >= t.nrow) return;
    yp[t.out_off + row] = ut_row_dot(UT + t.ut_off + row, t.nrow, t.r, yv);
}

__global__ void __launch_bounds__(256) k_bwd2(const BwdTask2 *__restrict__ tasks,
                                              const double *__restrict__ UT,
                                              const double *__restrict__ y1,
                                              double *__restrict__ yp) {
    bwd2_body(tasks[blockIdx.x], UT, y1, yp);
}

// Leaf outputs (src/H2Pack_matvec.c:767-861 for the leaves, and the backward
// permutation y[fwd[i]] = py[i] of src/H2Pack_matvec.c:1156-1159): the leaf's skeleton value is its own block
// partials plus its parent's push-down; then U y1 plus the point partials
// go straight to the caller's order.
struct LeafTask2 {
    long long ut_off;
    int nrow, r, s_off, push, vec_s;
    int cb1, ce1;   // contributions in skeleton space (length r)
    int cb0, ce0;   // contributions in point space (length nrow)
    int pad;
};
// shift != 0 adds shift * x (a diagonal nugget applied outside the H2
// product, as the PCG wrapper of examples/SPDHSS-H2/pcg_tests.c:15-21 does);
// px holds x in the permuted order.  A leaf's output row i is
//   ((U y1)_i) + its point-space contributions, in list order, + shift x_i,
// where y1 = push-down value + skeleton-space contributions, in list order.
// With a small rank (r <= blockDim / 2) the contributions are dealt to
// G = blockDim / r groups of r threads (group g sums contributions g, g + G,
// ...), so that G loads are in flight per skeleton entry instead of one
// dependent chain, and the G group sums are added in group order; part
// holds blockDim doubles.  The caller synchronises before reading yv.
__device__ __forceinline__ void leaf_stage_y1(const LeafTask2 &t, const long long *__restrict__ contrib,
                                              const double *__restrict__ arena, const double *__restrict__ yp,
                                              double *yv, double *part) {
    const int G = t.r > 0 ? min(8, (int)blockDim.x / t.r) : 1;
    if (G <= 1) {
        for (int c = threadIdx.x; c < t.r; c += blockDim.x)
            yv[c] = ordered_gather_sum(t.push ? yp[t.s_off + c] : 0.0, contrib, arena, t.cb1, t.ce1, 1, c);
        return;
    }
    const int g = threadIdx.x / t.r, c = threadIdx.x - g * t.r;
    if (g < G) part[g * t.r + c] = ordered_gather_sum(0.0, contrib, arena, t.cb1 + g, t.ce1, G, c);
    __syncthreads();
    for (int c2 = threadIdx.x; c2 < t.r; c2 += blockDim.x) {
        double v = t.push ? yp[t.s_off + c2] : 0.0;
        for (int g2 = 0; g2 < G; g2++) v += part[g2 * t.r + c2];
        yv[c2] = v;
    }
}

__device__ __forceinline__ double leaf_u_row(const LeafTask2 &t, const double *__restrict__ UT, const double *yv,
                                             int i) {
    return ut_row_dot(UT + t.ut_off + i, t.nrow, t.r, yv);
}

__device__ __forceinline__ double leaf_d_row(const LeafTask2 &t, const long long *__restrict__ contrib,
                                             const double *__restrict__ arena, const double *__restrict__ px,
                                             double shift, int i, double v) {
    v = ordered_gather_sum(v, contrib, arena, t.cb0, t.ce0, 1, i);
    if (shift != 0.0) v += shift * px[t.vec_s + i];
    return v;
}

__device__ __forceinline__ void leaf_final2_body(const LeafTask2 &t, const long long *__restrict__ contrib,
                                                 const double *__restrict__ arena,
                                                 const double *__restrict__ UT,
                                                 const double *__restrict__ yp,
                                                 const int *__restrict__ fwd, double *__restrict__ y,
                                                 const double *__restrict__ px, double shift) {
    __shared__ double yv[H2P_MAXR], part[256];
    leaf_stage_y1(t, contrib, arena, yp, yv, part);
    __syncthreads();
    for (int i = threadIdx.x; i < t.nrow; i += blockDim.x)
        y[fwd[t.vec_s + i]] = leaf_d_row(t, contrib, arena, px, shift, i, leaf_u_row(t, UT, yv, i));
}

__global__ void __launch_bounds__(256) k_leaf_final2(const LeafTask2 *__restrict__ tasks,
                                                     const long long *__restrict__ contrib,
                                                     const double *__restrict__ arena,
                                                     const double *__restrict__ UT,
                                                     const double *__restrict__ yp,
                                                     const int *__restrict__ fwd,
                                                     double *__restrict__ y,
                                                     const double *__restrict__ px, double shift) {
    leaf_final2_body(tasks[blockIdx.x], contrib, arena, UT, yp, fwd, y, px, shift);
}

// The leaf outputs in two launches, for handles whose D blocks are still
// running during the downward sweep: k_leaf_u writes (U y1)_i to yt (in
// the permuted order, no D contribution needed), k_leaf_d adds the
// point-space contributions once the D blocks are done.  Same operations in
// the same order as k_leaf_final2, so the same bits.
__global__ void __launch_bounds__(256) k_leaf_u(const LeafTask2 *__restrict__ tasks,
                                                const long long *__restrict__ contrib,
                                                const double *__restrict__ arena,
                                                const double *__restrict__ UT,
                                                const double *__restrict__ yp, double *__restrict__ yt) {
    __shared__ double yv[H2P_MAXR], part[256];
    const LeafTask2 t = tasks[blockIdx.x];
    leaf_stage_y1(t, contrib, arena, yp, yv, part);
    __syncthreads();
    for (int i = threadIdx.x; i < t.nrow; i += blockDim.x) yt[t.vec_s + i] = leaf_u_row(t, UT, yv, i);
}

__global__ void __launch_bounds__(256) k_leaf_d(const LeafTask2 *__restrict__ tasks,
                                                const long long *__restrict__ contrib,
                                                const double *__restrict__ arena,
                                                const double *__restrict__ yt,
                                                const int *__restrict__ fwd, double *__restrict__ y,
                                                const double *__restrict__ px, double shift) {
    const LeafTask2 t = tasks[blockIdx.x];
    for (int i = threadIdx.x; i < t.nrow; i += blockDim.x)
        y[fwd[t.vec_s + i]] = leaf_d_row(t, contrib, arena, px, shift, i, yt[t.vec_s + i]);
}

// ---------------------------------------------------------------------------
// Grouped bi-GEMV over stored blocks (AOT), the products of
// CBLAS_BI_GEMV (src/H2Pack_matvec.c:26-65) as used by the intermediate
// (:303-381) and dense (:864-934) phases.  One workgroup per block: lanes
// own columns (NS strips of 64), the four waves take every fourth row; the
// row sums are wave reductions, the column sums are reduced across waves in
// LDS in a fixed order.
template <typename TS, int NS>
__device__ __forceinline__ void blk_aot_pass(const BlkTask &t, int cbase, int ncols,
                                             const TS *__restrict__ A,
                                             const double *__restrict__ xin,
                                             double *__restrict__ arena,
                                             double (*red)[4][H2P_WAVE], bool first) {
    // TS = double: FP64 products.  TS = float (FP32-mixed): blocks stored and
    // multiplied in FP32; row sums leave the pass as FP32 into FP64 partials,
    // column sums are FP32 over 16 rows and folded into FP64 registers.
    // Each wave takes four consecutive rows per iteration (4 x NS loads in
    // flight per lane) and reduces the four row sums with one butterfly.
    // The wave index is the same on every lane: said so, the row addresses
    // stay scalar and each load needs only its column offset per lane.
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    TS x0[NS], acc1[NS];
    double acc1d[NS];
    int cc[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const int c = k * H2P_WAVE + lane;
        const bool ok = c < ncols;
        cc[k] = ok ? c : 0;
        x0[k] = ok ? (TS)xin[t.in0 + cbase + c] : (TS)0;
        acc1[k] = (TS)0;
        acc1d[k] = 0.0;
    }
    const double *in1 = xin + t.in1;
    const bool has1 = t.flags & 1;
    int tile = 0;
    for (int r0 = 4 * w; r0 < t.nrow; r0 += 16) {
        TS p[4], v1[4], a[4][NS];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = min(r0 + q, t.nrow - 1);     // rows past the end repeat the last row
            const bool ok = r0 + q < t.nrow;
            const TS *row = A + (long long)i * t.ncol + cbase;
#pragma unroll
            for (int k = 0; k < NS; k++) a[q][k] = row[cc[k]];
            v1[q] = (has1 && ok) ? (TS)in1[i] : (TS)0;
            p[q] = (TS)0;
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int k = 0; k < NS; k++) {
                p[q] += a[q][k] * x0[k];
                acc1[k] += a[q][k] * v1[q];
            }
        const TS s = wave_sum4<TS>(p[0], p[1], p[2], p[3]);
        const int myrow = r0 + ((lane >> 4) & 3);
        if ((lane & 15) == 0 && myrow < t.nrow) {
            double *o = arena + t.out0 + myrow;
            if (first) *o = (double)s;
            else *o += (double)s;
        }
        if constexpr (sizeof(TS) == 4) {
            if (++tile == 4) {
                tile = 0;
#pragma unroll
                for (int k = 0; k < NS; k++) {
                    acc1d[k] += (double)acc1[k];
                    acc1[k] = (TS)0;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NS; k++) acc1d[k] += (double)acc1[k];
    if (has1) {
#pragma unroll
        for (int k = 0; k < NS; k++) red[k][w][lane] = acc1d[k];
        __syncthreads();
        for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
            const int k = c >> 6, l = c & 63;
            arena[t.out1 + cbase + c] = (red[k][0][l] + red[k][1][l]) + (red[k][2][l] + red[k][3][l]);
        }
        __syncthreads();
    }
}

#define H2P_AOT_MAXNS 4
template <typename TS>
__device__ __forceinline__ void blocks_aot_body(const BlkTask &t, const TS *__restrict__ blob,
                                                const double *__restrict__ xin,
                                                double *__restrict__ arena) {
    __shared__ double red[H2P_AOT_MAXNS][4][H2P_WAVE];
    const TS *A = blob + t.data_off;
    for (int cb = 0; cb < t.ncol; cb += H2P_AOT_MAXNS * H2P_WAVE) {
        const int nc = min(t.ncol - cb, H2P_AOT_MAXNS * H2P_WAVE);
        const int ns = (nc + 63) >> 6;
        const bool first = (cb == 0);
        if (ns <= 1) blk_aot_pass<TS, 1>(t, cb, nc, A, xin, arena, red, first);
        else if (ns <= 2) blk_aot_pass<TS, 2>(t, cb, nc, A, xin, arena, red, first);
        else if (ns <= 3) blk_aot_pass<TS, 3>(t, cb, nc, A, xin, arena, red, first);
        else blk_aot_pass<TS, 4>(t, cb, nc, A, xin, arena, red, first);
        // the next column pass adds into out0 rows written by the same lanes
        __syncthreads();
    }
}

template <typename TS>
__global__ void __launch_bounds__(256) k_blocks_aot(const BlkTask *__restrict__ tasks,
                                                    const TS *__restrict__ blob,
                                                    const double *__restrict__ xin,
                                                    double *__restrict__ arena) {
    blocks_aot_body<TS>(tasks[blockIdx.x], blob, xin, arena);
}

__global__ void k_f64_to_f32(const double *__restrict__ in, float *__restrict__ out, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (float)in[i];
}

// ---------------------------------------------------------------------------
// Grouped kernel-evaluating bi-matvec for RPY (JIT), the krnl_bimv products
// of src/H2Pack_matvec.c:586-764 (B) and :1002-1145 (D) with the 3x3 tensor
// K = t1 I + t2 u u^T per point pair.
//  * Source points live in lanes, NS strips of 64 per pass, with their
//    coordinates (plus the task's periodic shift), radius and its
//    reciprocal in registers.
//  * Target points are staged in LDS per chunk (coordinates, radius, its
//    reciprocal, the transposed product's input); each wave takes four
//    targets per iteration and reduces their twelve row sums with three
//    butterflies (one per component).
//  * Lanes without a source point use a valid point with zero coefficients.
#define H2P_JIT_ROWS 256                 // target rows staged per chunk (all JIT launches)
#define H2P_JIT_MAXNS 4
#define H2P_RPY_MAXNS 2
// T = double: FP64.  T = float (FP32-mixed): coordinates relative to the
// task's first target point (the differences taken in FP64, once per point,
// then rounded), the tensor and the products in FP32; row sums leave the
// pass as FP32 into the FP64 partials, column sums are FP32 over 16 rows
// per lane and folded into FP64 registers.
template <typename T>
struct RpyRowT { T x, y, z, a, ia, v0, v1, v2; };
using RpyRow = RpyRowT<double>;

template <int NS, typename T>
__device__ __forceinline__ void rpy_pass(const BlkTask &t, int pbase, int npts,
                                         const double *__restrict__ cx, long long cstride, const KParam &kp,
                                         const double *__restrict__ shifts, const double *__restrict__ xin,
                                         double *__restrict__ arena, RpyRowT<T> *rows,
                                         double (*red)[3][4][H2P_WAVE], bool first) {
    constexpr bool F32 = sizeof(T) == 4;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // rows of cx: x, y, z, radius, reciprocal radius
    const double *cy = cx + cstride, *cz = cx + 2 * cstride, *ca = cx + 3 * cstride, *cia = cx + 4 * cstride;
    // set 1 may be a periodic image: its coordinates plus the task's shift
    double hx = shifts[3 * t.shift], hy = shifts[3 * t.shift + 1], hz = shifts[3 * t.shift + 2];
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if constexpr (F32) {
        ox = cx[t.c0];
        oy = cy[t.c0];
        oz = cz[t.c0];
    }
    T sx[NS], sy[NS], sz[NS], sa[NS], isa[NS], x0[NS][3], acc1[NS][3];
    double acc1d[NS][3];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const int p = k * H2P_WAVE + lane;
        const bool ok = p < npts;
        const int g = t.c1 + pbase + (ok ? p : 0);
        if constexpr (F32) {
            sx[k] = (T)(cx[g] + hx - ox);
            sy[k] = (T)(cy[g] + hy - oy);
            sz[k] = (T)(cz[g] + hz - oz);
        } else {
            sx[k] = cx[g] + hx;
            sy[k] = cy[g] + hy;
            sz[k] = cz[g] + hz;
        }
        sa[k] = (T)ca[g];
        isa[k] = (T)cia[g];
#pragma unroll
        for (int b = 0; b < 3; b++) {
            x0[k][b] = ok ? (T)xin[t.in0 + (pbase + p) * 3 + b] : (T)0;
            acc1[k][b] = (T)0;
            acc1d[k][b] = 0.0;
        }
    }
    const bool has1 = t.flags & 1;
    const T C = (T)kp.p1;
    int tile = 0;
    for (int rc = 0; rc < t.n0; rc += H2P_JIT_ROWS) {
        const int nr = min(t.n0 - rc, H2P_JIT_ROWS);
        const int nr_pad = (nr + 15) & ~15;
        __syncthreads();
        for (int q = threadIdx.x; q < nr_pad; q += blockDim.x) {
            const bool ok = q < nr;
            const int g = t.c0 + rc + (ok ? q : 0);
            RpyRowT<T> r;
            if constexpr (F32) {
                r.x = (T)(cx[g] - ox);
                r.y = (T)(cy[g] - oy);
                r.z = (T)(cz[g] - oz);
            } else {
                r.x = cx[g];
                r.y = cy[g];
                r.z = cz[g];
            }
            r.a = (T)ca[g];
            r.ia = (T)cia[g];
            const bool v = ok && has1;
            r.v0 = v ? (T)xin[t.in1 + (rc + q) * 3] : (T)0;
            r.v1 = v ? (T)xin[t.in1 + (rc + q) * 3 + 1] : (T)0;
            r.v2 = v ? (T)xin[t.in1 + (rc + q) * 3 + 2] : (T)0;
            rows[q] = r;
        }
        __syncthreads();
        for (int r0 = 4 * w; r0 < nr_pad; r0 += 16) {
            T p[4][3];
            // one target at a time (the barriers bound the registers the
            // scheduler may spend on interleaving targets)
#pragma unroll
            for (int a = 0; a < 4; a++) {
                __builtin_amdgcn_sched_barrier(0);
                const RpyRowT<T> R = rows[r0 + a];
                p[a][0] = p[a][1] = p[a][2] = (T)0;
#pragma unroll
                for (int k = 0; k < NS; k++) {
                    const bool ok = (k * H2P_WAVE + lane) < npts;
                    T t1, t2, ux, uy, uz;
                    rpy_coef_rcp(R.x - sx[k], R.y - sy[k], R.z - sz[k], R.a, sa[k], R.ia, isa[k], C,
                                 t1, t2, ux, uy, uz);
                    t1 = ok ? t1 : (T)0;
                    t2 = ok ? t2 : (T)0;
                    // K = t1 I + t2 u u^T (symmetric)
                    const T ud0 = ux * x0[k][0] + uy * x0[k][1] + uz * x0[k][2];
                    const T ud1 = ux * R.v0 + uy * R.v1 + uz * R.v2;
                    p[a][0] += t1 * x0[k][0] + t2 * ux * ud0;
                    p[a][1] += t1 * x0[k][1] + t2 * uy * ud0;
                    p[a][2] += t1 * x0[k][2] + t2 * uz * ud0;
                    acc1[k][0] += t1 * R.v0 + t2 * ux * ud1;
                    acc1[k][1] += t1 * R.v1 + t2 * uy * ud1;
                    acc1[k][2] += t1 * R.v2 + t2 * uz * ud1;
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            const int myrow = r0 + ((lane >> 4) & 3);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const T s = wave_sum4<T>(p[0][c], p[1][c], p[2][c], p[3][c]);
                if ((lane & 15) == 0 && myrow < nr) {
                    double *o = arena + t.out0 + (rc + myrow) * 3 + c;
                    if (first) *o = (double)s;
                    else *o += (double)s;
                }
            }
            if constexpr (F32) {
                if (++tile == 4) {
                    tile = 0;
#pragma unroll
                    for (int k = 0; k < NS; k++)
#pragma unroll
                        for (int b = 0; b < 3; b++) {
                            acc1d[k][b] += (double)acc1[k][b];
                            acc1[k][b] = (T)0;
                        }
                }
            }
        }
    }
    if (has1) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NS; k++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if constexpr (F32) red[k][b][w][lane] = acc1d[k][b] + (double)acc1[k][b];
                else red[k][b][w][lane] = acc1[k][b];
            }
        __syncthreads();
        for (int q = threadIdx.x; q < npts * 3; q += blockDim.x) {
            const int pnt = q / 3, b = q - pnt * 3;
            const int k = pnt >> 6, l = pnt & 63;
            arena[t.out1 + (pbase + pnt) * 3 + b] =
                (red[k][b][0][l] + red[k][b][1][l]) + (red[k][b][2][l] + red[k][b][3][l]);
        }
    }
}

template <int KID, int KD, typename T = double>
__device__ __forceinline__ void blocks_jit_body(const BlkTask &t, const double *__restrict__ cx,
                                                long long cstride, const KParam &kp,
                                                const double *__restrict__ shifts,
                                                const double *__restrict__ xin,
                                                double *__restrict__ arena) {
    static_assert(KID == KR && KD == 3, "the per-pass launch is the RPY kernel's");
    __shared__ RpyRowT<T> rows[H2P_JIT_ROWS];
    __shared__ double red[H2P_RPY_MAXNS][3][4][H2P_WAVE];
    for (int pb = 0; pb < t.n1; pb += H2P_RPY_MAXNS * H2P_WAVE) {
        const int np = min(t.n1 - pb, H2P_RPY_MAXNS * H2P_WAVE);
        const bool first = (pb == 0);
        if (np <= H2P_WAVE) rpy_pass<1, T>(t, pb, np, cx, cstride, kp, shifts, xin, arena, rows, red, first);
        else rpy_pass<2, T>(t, pb, np, cx, cstride, kp, shifts, xin, arena, rows, red, first);
        __syncthreads();
    }
}

template <int KID, int KD, typename T = double>
__global__ void __launch_bounds__(256) k_blocks_jit(const BlkTask *__restrict__ tasks,
                                                    const double *__restrict__ cx,
                                                    long long cstride, KParam kp,
                                                    const double *__restrict__ shifts,
                                                    const double *__restrict__ xin,
                                                    double *__restrict__ arena) {
    blocks_jit_body<KID, KD, T>(tasks[blockIdx.x], cx, cstride, kp, shifts, xin, arena);
}

// The root's periodic block (H2P_matvec_periodic, src/H2Pack_matvec_periodic.c:
// 455-470): y1[root] = per_blk y0[root], one wave per row, lanes over the
// columns, fixed-order wave reduction.
__global__ void __launch_bounds__(256) k_root_block(const double *__restrict__ per, int r,
                                                    const double *__restrict__ y0,
                                                    double *__restrict__ y1) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= r) return;
    const double *a = per + (long long)row * r;
    double s = 0.0;
    for (int c = lane; c < r; c += H2P_WAVE) s += a[c] * y0[c];
    s = wave_sum(s);
    if (lane == 0) y1[row] = s;
}

// ---------------------------------------------------------------------------
// Grouped kernel-evaluating bi-matvec for FP64 Coulomb.  1/r has the
// kernel's diagonal value at r = 0, so r = 0 must be exact: the distances
// are formed from coordinate differences.
//  * Source points (columns) live in lanes, NS strips of 64 per pass.
//  * Target points (rows) are staged in LDS; each wave takes four rows per
//    iteration and reduces the four row sums with one butterfly.
//  * Lanes without a source point use a valid point with x = 0, so every
//    kernel value stays finite and no per-pair select is needed.
struct JitRow { double x, y, z, v; };    // coordinates and the transposed product's input

template <int NS>
__device__ __forceinline__ void jitc_pass(const BlkTask &t, int pbase, int npts,
                                          const double *__restrict__ cx, long long cstride, double diag,
                                          const double *__restrict__ xin, double *__restrict__ arena,
                                          JitRow *rows, double (*red)[4][H2P_WAVE], bool first) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double *cy = cx + cstride, *cz = cx + 2 * cstride;
    double sx[NS], sy[NS], sz[NS], x0[NS], acc1[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        const int p = k * H2P_WAVE + lane;
        const bool ok = p < npts;
        const int g = t.c1 + pbase + (ok ? p : 0);
        sx[k] = cx[g];
        sy[k] = cy[g];
        sz[k] = cz[g];
        x0[k] = ok ? xin[t.in0 + pbase + p] : 0.0;
        acc1[k] = 0.0;
    }
    const bool has1 = t.flags & 1;
    for (int rc = 0; rc < t.n0; rc += H2P_JIT_ROWS) {
        const int nr = min(t.n0 - rc, H2P_JIT_ROWS);
        const int nr_pad = (nr + 15) & ~15;
        __syncthreads();
        for (int q = threadIdx.x; q < nr_pad; q += blockDim.x) {
            const bool ok = q < nr;
            const int g = t.c0 + rc + (ok ? q : 0);
            JitRow r;
            r.x = cx[g];
            r.y = cy[g];
            r.z = cz[g];
            r.v = (ok && has1) ? xin[t.in1 + rc + q] : 0.0;
            rows[q] = r;
        }
        __syncthreads();
        for (int r0 = 4 * w; r0 < nr_pad; r0 += 16) {
            double p[4];
            JitRow rr[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                rr[a] = rows[r0 + a];
                p[a] = 0.0;
            }
#pragma unroll
            for (int k = 0; k < NS; k++) {
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    const double dx = rr[a].x - sx[k], dy = rr[a].y - sy[k], dz = rr[a].z - sz[k];
                    const double kv = kprep<KC>(dx * dx + dy * dy + dz * dz, diag);
                    p[a] += kv * x0[k];
                    acc1[k] += kv * rr[a].v;
                }
            }
            const double s = wave_sum4<double>(p[0], p[1], p[2], p[3]);
            const int myrow = r0 + ((lane >> 4) & 3);
            if ((lane & 15) == 0 && myrow < nr) {
                double *o = arena + t.out0 + rc + myrow;
                if (first) *o = s;
                else *o += s;
            }
        }
    }
    if (has1) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NS; k++) red[k][w][lane] = acc1[k];
        __syncthreads();
        for (int c = threadIdx.x; c < npts; c += blockDim.x) {
            const int k = c >> 6, l = c & 63;
            arena[t.out1 + pbase + c] = (red[k][0][l] + red[k][1][l]) + (red[k][2][l] + red[k][3][l]);
        }
    }
}

__device__ __forceinline__ void blocks_jitc(const BlkTask &t, const double *__restrict__ cx, long long cstride,
                                            double diag, const double *__restrict__ xin,
                                            double *__restrict__ arena) {
    __shared__ JitRow rows[H2P_JIT_ROWS];
    __shared__ double red[H2P_JIT_MAXNS][4][H2P_WAVE];
    for (int pb = 0; pb < t.n1; pb += H2P_JIT_MAXNS * H2P_WAVE) {
        const int np = min(t.n1 - pb, H2P_JIT_MAXNS * H2P_WAVE);
        const int ns = (np + 63) >> 6;
        const bool first = (pb == 0);
        if (ns <= 1) jitc_pass<1>(t, pb, np, cx, cstride, diag, xin, arena, rows, red, first);
        else if (ns <= 2) jitc_pass<2>(t, pb, np, cx, cstride, diag, xin, arena, rows, red, first);
        else if (ns <= 3) jitc_pass<3>(t, pb, np, cx, cstride, diag, xin, arena, rows, red, first);
        else jitc_pass<4>(t, pb, np, cx, cstride, diag, xin, arena, rows, red, first);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FP64 norm-form kernels (Gaussian, Matern) with the distances on the matrix
// cores.  The squared distances of a tile of 16 target rows x 16 source
// columns are one v_mfma_f64_16x16x4f64,
//     q[i][j] = |r_i|^2 + sum_k A[i][k] B[k][j],
//     A[i] = (1, r_i),  B[.][j] = (|s_j|^2, -2 s_j),
// with |r_i|^2 as the accumulator input (the same terms, in the same order,
// as the VALU form |r|^2 + |s|^2 - 2 r.s), so the VALU only evaluates the
// kernel and the two products of each pair.  MFMA layout, lane l: it
// supplies A[l % 16][l / 16] and B[l / 16][l % 16] and receives
// q[4 r + l / 16][l % 16] in register r.
//  * A pass covers up to H2P_JITM_NT tiles of 16 source columns, staged in
//    LDS (B operand and input value); each lane keeps one column partial
//    sum per tile in registers.
//  * Target rows are staged in LDS in chunks; the NW cooperating waves take
//    the 16-row tiles w, w + NW, ...; per tile a lane accumulates four row
//    sums (rows 4 r + l / 16) over its columns, reduced over the 16 lanes
//    of each group once per tile.
//  * Column partial sums are reduced at the end of the pass over the four
//    lane groups (and, with NW = 4, the four waves) in a fixed order.
//  * Padded rows repeat a valid row with input 0 and padded columns a valid
//    source with input 0, so every kernel value is finite and the padding
//    adds exact zeros.
// Two launch units share one kernel: a workgroup unit (NW = 4) runs one
// large block; a wave unit runs four small blocks (n0 <= H2P_JITW_ROWS,
// n1 <= 16 H2P_JITM_NT), one per wave, with wave-private LDS and no
// workgroup barriers, so the short blocks do not wait on each other.
#define H2P_JITM_NT 8

// Runs a tile body for the rows that hold points: lane group g's register r
// holds tile row 4 r + g, so with vr valid rows the registers r >= ceil(vr /
// 4) hold only padding (a repeated point with input 0, whose products add
// exact zeros) and are skipped.  Same results, fewer kernel evaluations in a
// block's last row tile.
template <typename F>
__device__ __forceinline__ void tile_rows(int vr, F &&body) {
    if (vr <= 4) body(std::integral_constant<int, 1>{});
    else if (vr <= 8) body(std::integral_constant<int, 2>{});
    else body(std::integral_constant<int, 4>{});
}
#define H2P_JITW_ROWS 80
struct JitmRow { double a[4]; };   // (1, x, y, z) of a target; (|s|^2, -2 s) of a source
struct JitmNV { double n, v; };    // |r|^2 and the transposed product's input

// LDS of the norm-form kernels, in doubles: the workgroup unit uses rows
// (4 + 2) x H2P_JIT_ROWS, sources 5 x 16 NT and the wave partials
// 4 x NT x 64; a wave unit gives each wave (4 + 2) x H2P_JITW_ROWS + 5 x 16 NT.
#define H2P_JITM_LDS_WG (6 * H2P_JIT_ROWS + 5 * 16 * H2P_JITM_NT + 4 * H2P_JITM_NT * H2P_WAVE)
#define H2P_JITM_LDS_WAVE (6 * H2P_JITW_ROWS + 5 * 16 * H2P_JITM_NT)
#define H2P_JITM_LDS (H2P_JITM_LDS_WG > 4 * H2P_JITM_LDS_WAVE ? H2P_JITM_LDS_WG : 4 * H2P_JITM_LDS_WAVE)

// Orders the LDS accesses of one wave (LDS operations of a wave complete in
// issue order; the fences keep the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NW>
__device__ __forceinline__ void jitm_sync() {
    if constexpr (NW == 4) __syncthreads();
    else wave_lds_sync();
}

template <int KID, int NT, int NW>
__device__ __forceinline__ void jitm_pass(const BlkTask &t, int pbase, int npts,
                                          const double *__restrict__ cx, long long cstride,
                                          double scale, double diag, double ox, double oy, double oz,
                                          const double *__restrict__ xin, double *__restrict__ arena,
                                          JitmRow *rowa, JitmNV *rownv, JitmRow *srcb, double *srcx,
                                          double *red, const double *e2t, bool first) {
    constexpr int ROWS = NW == 4 ? H2P_JIT_ROWS : H2P_JITW_ROWS;
    const int lane = threadIdx.x & 63, w = NW == 4 ? (threadIdx.x >> 6) : 0;
    const int tid = NW == 4 ? threadIdx.x : lane;
    const int li = lane & 15, lg = lane >> 4;
    const double *cy = cx + cstride, *cz = cx + 2 * cstride;
    // the pass's sources: B operand (|s|^2, -2 s) and input value
    for (int q = tid; q < NT * 16; q += NW * H2P_WAVE) {
        const bool ok = q < npts;
        const int g = t.c1 + pbase + (ok ? q : 0);
        const double ux = (cx[g] - ox) * scale, uy = (cy[g] - oy) * scale, uz = (cz[g] - oz) * scale;
        JitmRow b;
        b.a[0] = ux * ux + uy * uy + uz * uz;
        b.a[1] = -2.0 * ux;
        b.a[2] = -2.0 * uy;
        b.a[3] = -2.0 * uz;
        srcb[q] = b;
        srcx[q] = ok ? xin[t.in0 + pbase + q] : 0.0;
    }
    double acc[NT];
#pragma unroll
    for (int c = 0; c < NT; c++) acc[c] = 0.0;
    const bool has1 = t.flags & 1;
    for (int rc = 0; rc < t.n0; rc += ROWS) {
        const int nr = min(t.n0 - rc, ROWS);
        const int ntile = (nr + 15) >> 4;
        jitm_sync<NW>();
        for (int q = tid; q < ntile * 16; q += NW * H2P_WAVE) {
            const bool ok = q < nr;
            const int g = t.c0 + rc + (ok ? q : 0);
            const double ux = (cx[g] - ox) * scale, uy = (cy[g] - oy) * scale, uz = (cz[g] - oz) * scale;
            JitmRow ra;
            ra.a[0] = 1.0;
            ra.a[1] = ux;
            ra.a[2] = uy;
            ra.a[3] = uz;
            rowa[q] = ra;
            JitmNV e;
            e.n = ux * ux + uy * uy + uz * uz;
            e.v = (ok && has1) ? xin[t.in1 + rc + q] : 0.0;
            rownv[q] = e;
        }
        jitm_sync<NW>();
        for (int tt = w; tt < ntile; tt += NW) {
            const double a = rowa[tt * 16 + li].a[lg];
            d4 cn;
            double v[4], p[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const JitmNV e = rownv[tt * 16 + 4 * r + lg];
                cn[r] = e.n;
                v[r] = e.v;
                p[r] = 0.0;
            }
            // the scheduling barriers keep the compiler from hoisting all NT
            // products ahead of the VALU work (each live result costs 8
            // registers); other waves cover the product's latency
            auto tile = [&](auto nrq) {
                constexpr int NR = decltype(nrq)::value;
#pragma unroll
                for (int c = 0; c < NT; c++) {
                    const double xc = srcx[c * 16 + li];
                    const d4 q = __builtin_amdgcn_mfma_f64_16x16x4f64(a, srcb[c * 16 + li].a[lg], cn, 0, 0, 0);
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        const double kv = kprep<KID>(q[r], diag, e2t);
                        p[r] += kv * xc;
                        acc[c] += kv * v[r];
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            };
            tile_rows(nr - tt * 16, tile);
            const double s = group16_sum4(p[0], p[1], p[2], p[3]);
            const int myrow = tt * 16 + 4 * ((lane >> 2) & 3) + lg;
            if ((lane & 3) == 0 && myrow < nr) {
                double *o = arena + t.out0 + rc + myrow;
                if (first) *o = s;
                else *o += s;
            }
        }
    }
    if (!has1) return;
    if constexpr (NW == 4) {
#pragma unroll
        for (int c = 0; c < NT; c++) red[(w * NT + c) * H2P_WAVE + lane] = acc[c];
        __syncthreads();
        for (int j = threadIdx.x; j < npts; j += blockDim.x) {
            const int c = j >> 4, m = j & 15;
            double s = 0.0;
#pragma unroll
            for (int w2 = 0; w2 < 4; w2++)
#pragma unroll
                for (int g = 0; g < 4; g++) s += red[(w2 * NT + c) * H2P_WAVE + g * 16 + m];
            arena[t.out1 + pbase + j] = s;
        }
    } else {
        // (g0 + g1) + (g2 + g3) over the four lane groups
#pragma unroll
        for (int c = 0; c < NT; c++) {
            double s = acc[c] + __shfl_xor(acc[c], 16, H2P_WAVE);
            s += __shfl_xor(s, 32, H2P_WAVE);
            const int j = c * 16 + li;
            if (lg == 0 && j < npts) arena[t.out1 + pbase + j] = s;
        }
    }
}

template <int KID, int NW>
__device__ __forceinline__ void jitm_task(const BlkTask &t, const double *__restrict__ cx,
                                          long long cstride, double scale, double diag,
                                          const double *__restrict__ xin, double *__restrict__ arena,
                                          double *lds, const double *e2t) {
    constexpr int ROWS = NW == 4 ? H2P_JIT_ROWS : H2P_JITW_ROWS;
    JitmRow *rowa = (JitmRow *)lds;
    JitmNV *rownv = (JitmNV *)(lds + 4 * ROWS);
    JitmRow *srcb = (JitmRow *)(lds + 6 * ROWS);
    double *srcx = lds + 6 * ROWS + 4 * 16 * H2P_JITM_NT;
    double *red = srcx + 16 * H2P_JITM_NT;   // workgroup unit only
    const double ox = cx[t.c0], oy = cx[cstride + t.c0], oz = cx[2 * cstride + t.c0];
    for (int pb = 0; pb < t.n1; pb += H2P_JITM_NT * 16) {
        const int np = min(t.n1 - pb, H2P_JITM_NT * 16);
        const bool first = (pb == 0);
        switch ((np + 15) >> 4) {
#define H2P_JITM_CASE(n) \
            case n: jitm_pass<KID, n, NW>(t, pb, np, cx, cstride, scale, diag, ox, oy, oz, xin, arena, rowa, rownv, srcb, srcx, red, e2t, first); break;
            H2P_JITM_CASE(1) H2P_JITM_CASE(2) H2P_JITM_CASE(3) H2P_JITM_CASE(4)
            H2P_JITM_CASE(5) H2P_JITM_CASE(6) H2P_JITM_CASE(7) H2P_JITM_CASE(8)
#undef H2P_JITM_CASE
            default: break;
        }
        jitm_sync<NW>();
    }
}

// Launch unit u of a tiled launch: u < nbig runs task u on the whole
// workgroup, later units run tasks nbig + 4 (u - nbig) + w, one per wave.
// The exp2_neg table is copied into LDS first (one copy per workgroup).
template <int KID>
__device__ __forceinline__ void blocks_jitm_unit(const BlkTask *__restrict__ tasks, int u, int nbig, int ntask,
                                                 const double *__restrict__ cx, long long cstride,
                                                 double scale, double diag, const double *__restrict__ xin,
                                                 double *__restrict__ arena) {
    __shared__ double lds[H2P_JITM_LDS];
    __shared__ double e2t[H2P_E2_N];
    if (threadIdx.x < H2P_E2_N) e2t[threadIdx.x] = c_e2tab[threadIdx.x];
    __syncthreads();
    if (u < nbig) {
        jitm_task<KID, 4>(tasks[u], cx, cstride, scale, diag, xin, arena, lds, e2t);
    } else {
        const int w = threadIdx.x >> 6;
        const int i = nbig + 4 * (u - nbig) + w;
        if (i < ntask)
            jitm_task<KID, 1>(tasks[i], cx, cstride, scale, diag, xin, arena, lds + w * H2P_JITM_LDS_WAVE, e2t);
    }
}

// ---------------------------------------------------------------------------
// FP32-mixed scalar kernels (Coulomb, Gaussian, Matern) in the same tiling
// and launch units, with the distances on the VALU from coordinate
// differences (FP32 keeps differences for accuracy, and Coulomb's r = 0 stays
// exact).  Lane l takes column l % 16 of a tile and rows 4 r + l / 16.
// Kernel values and products in FP32; a pass's row sums (up to
// 16 H2P_JITM_NT columns) leave it as FP32 into the FP64 partial outputs;
// each lane's column sums are FP32 over one staged row chunk and are then
// folded into FP64 registers.
struct JitfPt { float x, y, z, v; };   // prepared coordinates and input value

// A workgroup unit's pass covers up to H2P_JITF_NT tiles of 16 columns
// (one source per thread), a wave unit's up to H2P_JITM_NT: the wider pass
// halves the per-column-tile cost of the row reductions and row staging.
// LDS in doubles: workgroup unit rows 2 H2P_JIT_ROWS, sources 2 x 16 NT,
// wave partials 4 x NT x 16; a wave unit 2 H2P_JITW_ROWS + 2 x 16 NT per wave.
#define H2P_JITF_NT 16
#define H2P_JITF_LDS_WG (2 * H2P_JIT_ROWS + 2 * 16 * H2P_JITF_NT + 4 * H2P_JITF_NT * 16)
#define H2P_JITF_LDS_WAVE (2 * H2P_JITW_ROWS + 2 * 16 * H2P_JITM_NT)
#define H2P_JITF_LDS (H2P_JITF_LDS_WG > 4 * H2P_JITF_LDS_WAVE ? H2P_JITF_LDS_WG : 4 * H2P_JITF_LDS_WAVE)

template <int KID, int NT, int NW>
__device__ __forceinline__ void jitf_pass(const BlkTask &t, int pbase, int npts,
                                          const double *__restrict__ cx, long long cstride,
                                          double scale, double diag, double ox, double oy, double oz,
                                          const double *__restrict__ xin, double *__restrict__ arena,
                                          JitfPt *rows, JitfPt *srcs, double *red, bool first) {
    constexpr int ROWS = NW == 4 ? H2P_JIT_ROWS : H2P_JITW_ROWS;
    const int lane = threadIdx.x & 63, w = NW == 4 ? (threadIdx.x >> 6) : 0;
    const int tid = NW == 4 ? threadIdx.x : lane;
    const int li = lane & 15, lg = lane >> 4;
    const double *cy = cx + cstride, *cz = cx + 2 * cstride;
    const float dg = (float)diag;
    for (int q = tid; q < NT * 16; q += NW * H2P_WAVE) {
        const bool ok = q < npts;
        const int g = t.c1 + pbase + (ok ? q : 0);
        JitfPt e;
        e.x = (float)((cx[g] - ox) * scale);
        e.y = (float)((cy[g] - oy) * scale);
        e.z = (float)((cz[g] - oz) * scale);
        e.v = ok ? (float)xin[t.in0 + pbase + q] : 0.0f;
        srcs[q] = e;
    }
    float acc[NT];
#pragma unroll
    for (int c = 0; c < NT; c++) acc[c] = 0.0f;
    const bool has1 = t.flags & 1;
    for (int rc = 0; rc < t.n0; rc += ROWS) {
        const int nr = min(t.n0 - rc, ROWS);
        const int ntile = (nr + 15) >> 4;
        jitm_sync<NW>();
        for (int q = tid; q < ntile * 16; q += NW * H2P_WAVE) {
            const bool ok = q < nr;
            const int g = t.c0 + rc + (ok ? q : 0);
            JitfPt e;
            e.x = (float)((cx[g] - ox) * scale);
            e.y = (float)((cy[g] - oy) * scale);
            e.z = (float)((cz[g] - oz) * scale);
            e.v = (ok && has1) ? (float)xin[t.in1 + rc + q] : 0.0f;
            rows[q] = e;
        }
        jitm_sync<NW>();
        for (int tt = w; tt < ntile; tt += NW) {
            float rx[4], ry[4], rz[4], rv[4], p[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const JitfPt e = rows[tt * 16 + 4 * r + lg];
                rx[r] = e.x;
                ry[r] = e.y;
                rz[r] = e.z;
                rv[r] = e.v;
                p[r] = 0.0f;
            }
            // one column tile at a time (the barriers bound the registers
            // the scheduler may spend on interleaving tiles)
            auto tile = [&](auto nrq) {
                constexpr int NR = decltype(nrq)::value;
#pragma unroll
                for (int c = 0; c < NT; c++) {
                    const JitfPt e = srcs[c * 16 + li];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        const float dx = rx[r] - e.x, dy = ry[r] - e.y, dz = rz[r] - e.z;
                        const float kv = kprep<KID>(dx * dx + dy * dy + dz * dz, dg);
                        p[r] += kv * e.v;
                        acc[c] += kv * rv[r];
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            };
            tile_rows(nr - tt * 16, tile);
            const float s = group16_sum4(p[0], p[1], p[2], p[3]);
            const int myrow = tt * 16 + 4 * ((lane >> 2) & 3) + lg;
            if ((lane & 3) == 0 && myrow < nr) {
                double *o = arena + t.out0 + rc + myrow;
                if (first) *o = (double)s;
                else *o += (double)s;
            }
        }
        if (has1) {
            // the chunk's FP32 column sums in FP64, (g0 + g1) + (g2 + g3)
            // over the four lane groups, added in chunk order to the wave's
            // sums (in LDS for a workgroup unit, in the output for a wave unit)
#pragma unroll
            for (int c = 0; c < NT; c++) {
                double v = (double)acc[c];
                v += __shfl_xor(v, 16, H2P_WAVE);
                v += __shfl_xor(v, 32, H2P_WAVE);
                acc[c] = 0.0f;
                const int j = c * 16 + li;
                double *o = NW == 4 ? red + (w * NT + c) * 16 + li : arena + t.out1 + pbase + j;
                if (lg == 0 && (NW == 4 || j < npts)) *o = rc == 0 ? v : *o + v;
            }
        }
    }
    if (!has1) return;
    if constexpr (NW == 4) {   // the four waves in order
        __syncthreads();
        for (int j = threadIdx.x; j < npts; j += blockDim.x) {
            const int c = j >> 4, m = j & 15;
            double s = red[c * 16 + m];
#pragma unroll
            for (int w2 = 1; w2 < 4; w2++) s += red[(w2 * NT + c) * 16 + m];
            arena[t.out1 + pbase + j] = s;
        }
    }
}

template <int KID, int NW>
__device__ __forceinline__ void jitf_task(const BlkTask &t, const double *__restrict__ cx,
                                          long long cstride, double scale, double diag,
                                          const double *__restrict__ xin, double *__restrict__ arena,
                                          double *lds) {
    constexpr int ROWS = NW == 4 ? H2P_JIT_ROWS : H2P_JITW_ROWS;
    constexpr int NTX = NW == 4 ? H2P_JITF_NT : H2P_JITM_NT;
    JitfPt *rows = (JitfPt *)lds;
    JitfPt *srcs = (JitfPt *)(lds + 2 * ROWS);
    double *red = lds + 2 * ROWS + 2 * 16 * NTX;   // workgroup unit only
    const double ox = cx[t.c0], oy = cx[cstride + t.c0], oz = cx[2 * cstride + t.c0];
    for (int pb = 0; pb < t.n1; pb += NTX * 16) {
        const int np = min(t.n1 - pb, NTX * 16);
        const bool first = (pb == 0);
        const int nt = (np + 15) >> 4;
#define H2P_JITF_CASE(n) \
            case n: jitf_pass<KID, n, NW>(t, pb, np, cx, cstride, scale, diag, ox, oy, oz, xin, arena, rows, srcs, red, first); break;
        if constexpr (NW == 4) {
            switch (nt) {
                H2P_JITF_CASE(1) H2P_JITF_CASE(2) H2P_JITF_CASE(3) H2P_JITF_CASE(4)
                H2P_JITF_CASE(5) H2P_JITF_CASE(6) H2P_JITF_CASE(7) H2P_JITF_CASE(8)
                H2P_JITF_CASE(9) H2P_JITF_CASE(10) H2P_JITF_CASE(11) H2P_JITF_CASE(12)
                H2P_JITF_CASE(13) H2P_JITF_CASE(14) H2P_JITF_CASE(15) H2P_JITF_CASE(16)
                default: break;
            }
        } else {
            switch (nt) {
                H2P_JITF_CASE(1) H2P_JITF_CASE(2) H2P_JITF_CASE(3) H2P_JITF_CASE(4)
                H2P_JITF_CASE(5) H2P_JITF_CASE(6) H2P_JITF_CASE(7) H2P_JITF_CASE(8)
                default: break;
            }
        }
#undef H2P_JITF_CASE
        jitm_sync<NW>();
    }
}

template <int KID>
__device__ __forceinline__ void blocks_jitf_unit(const BlkTask *__restrict__ tasks, int u, int nbig, int ntask,
                                                 const double *__restrict__ cx, long long cstride,
                                                 double scale, double diag, const double *__restrict__ xin,
                                                 double *__restrict__ arena) {
    __shared__ double lds[H2P_JITF_LDS];
    if (u < nbig) {
        jitf_task<KID, 4>(tasks[u], cx, cstride, scale, diag, xin, arena, lds);
    } else {
        const int w = threadIdx.x >> 6;
        const int i = nbig + 4 * (u - nbig) + w;
        if (i < ntask)
            jitf_task<KID, 1>(tasks[i], cx, cstride, scale, diag, xin, arena, lds + w * H2P_JITF_LDS_WAVE);
    }
}

// Whether a task of the tiled launches runs as a wave unit.
__host__ __device__ inline bool jitm_wave_task(int n0, int n1) {
    return n0 <= H2P_JITW_ROWS && n1 <= 16 * H2P_JITM_NT;
}

// Launch unit u of the grouped kernel-evaluating launch: the FP64 norm-form
// kernels and the FP32-mixed kernels run workgroup and wave units, FP64
// Coulomb one task per workgroup (nbig = ntask).
template <int KID, typename T>
__device__ __forceinline__ void blocks_jit2_unit(const BlkTask *__restrict__ tasks, int u, int nbig, int ntask,
                                                 const double *__restrict__ cx, long long cstride,
                                                 double scale, double diag, const double *__restrict__ xin,
                                                 double *__restrict__ arena) {
    if constexpr (sizeof(T) == 4) {
        blocks_jitf_unit<KID>(tasks, u, nbig, ntask, cx, cstride, scale, diag, xin, arena);
    } else if constexpr (KID == KG || KID == KM || KID == KGW || KID == KMW) {
        blocks_jitm_unit<KID>(tasks, u, nbig, ntask, cx, cstride, scale, diag, xin, arena);
    } else {
        static_assert(KID == KC, "scalar kernels only");
        blocks_jitc(tasks[u], cx, cstride, diag, xin, arena);
    }
}

// Occupancy build of the grouped launch: the same body compiled for at
// least WPE waves per SIMD, so with fewer registers per lane.  Same
// arithmetic, same results.  PART (0: B blocks, 1: D blocks) only names the
// launch, so that the two launches of a matvec are told apart in profiles.
template <int KID, typename T, int WPE, int PART>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_blocks_jit2_occ(const BlkTask *__restrict__ tasks, int nbig, int ntask, const double *__restrict__ cx,
                  long long cstride, double scale, double diag, const double *__restrict__ xin,
                  double *__restrict__ arena) {
    blocks_jit2_unit<KID, T>(tasks, blockIdx.x, nbig, ntask, cx, cstride, scale, diag, xin, arena);
}

template <typename TS, int WPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_blocks_aot_occ(const BlkTask *__restrict__ tasks, const TS *__restrict__ blob,
                 const double *__restrict__ xin, double *__restrict__ arena) {
    blocks_aot_body<TS>(tasks[blockIdx.x], blob, xin, arena);
}

template <int KID, typename T, int PART>
__global__ void __launch_bounds__(256) k_blocks_jit2(const BlkTask *__restrict__ tasks, int nbig, int ntask,
                                                     const double *__restrict__ cx,
                                                     long long cstride, double scale, double diag,
                                                     const double *__restrict__ xin,
                                                     double *__restrict__ arena) {
    blocks_jit2_unit<KID, T>(tasks, blockIdx.x, nbig, ntask, cx, cstride, scale, diag, xin, arena);
}

// ---------------------------------------------------------------------------
// Device-side evaluation of AOT blocks (H2P_build_B_AOT / H2P_build_D_AOT,
// src/H2Pack_build.c:789-889, 1034-1140): one workgroup per block, threads
// over the row-major entries.
template <int KID, int KD, typename TS>
__global__ void __launch_bounds__(256) k_build_blocks(const BlkTask *__restrict__ tasks,
                                                      const double *__restrict__ cx,
                                                      long long cstride, KParam kp,
                                                      TS *__restrict__ blob) {
    // values are evaluated in FP64 and stored as TS
    const BlkTask t = tasks[blockIdx.x];
    const double *cy = cx + cstride, *cz = cx + 2 * cstride, *ca = cx + 3 * cstride;
    TS *A = blob + t.data_off;
    const long long total = (long long)t.n0 * t.n1;
    for (long long e = threadIdx.x; e < total; e += blockDim.x) {
        int i = (int)(e / t.n1), j = (int)(e - (long long)i * t.n1);
        int gi = t.c0 + i, gj = t.c1 + j;
        double dx = cx[gi] - cx[gj], dy = cy[gi] - cy[gj], dz = cz[gi] - cz[gj];
        if constexpr (KD == 1) {
            A[e] = (TS)kval<KID>(dx * dx + dy * dy + dz * dz, kp);
        } else {
            double t1, t2, ux, uy, uz;
            rpy_coef(dx, dy, dz, ca[gi], ca[gj], kp.p1, t1, t2, ux, uy, uz);
            double u[3] = {ux, uy, uz};
            const long long ld = (long long)t.n1 * 3;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
                    A[(long long)(3 * i + a) * ld + 3 * j + b] = (TS)(t2 * u[a] * u[b] + (a == b ? t1 : 0.0));
        }
    }
}

// ---------------------------------------------------------------------------
// Direct FP64 n-body rows (examples/direct_nbody.h): grid = (target blocks of
// 256, source chunks); partial sums per chunk, reduced in a fixed order.
template <int KID, int KD>
__global__ void __launch_bounds__(256) k_nbody_partial(const double *__restrict__ src, long long sld, int n_src,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ dst, long long dld, int n_dst,
                                                       KParam kp, int chunk, double *__restrict__ part) {
    __shared__ double s[4][256];
    __shared__ double sxv[3][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int js = blockIdx.y * chunk, je = min(n_src, js + chunk);
    double tx = 0, ty = 0, tz = 0, ta = 0;
    if (i < n_dst) {
        tx = dst[i]; ty = dst[dld + i]; tz = dst[2 * dld + i];
        if constexpr (KD == 3) ta = dst[3 * dld + i];
    }
    double acc[KD];
#pragma unroll
    for (int a = 0; a < KD; a++) acc[a] = 0.0;
    for (int j0 = js; j0 < je; j0 += 256) {
        int j = j0 + threadIdx.x;
        __syncthreads();
        if (j < je) {
            s[0][threadIdx.x] = src[j]; s[1][threadIdx.x] = src[sld + j]; s[2][threadIdx.x] = src[2 * sld + j];
            if constexpr (KD == 3) s[3][threadIdx.x] = src[3 * sld + j];
#pragma unroll
            for (int b = 0; b < KD; b++) sxv[b][threadIdx.x] = x[(long long)j * KD + b];
        }
        __syncthreads();
        int n = min(256, je - j0);
        for (int q = 0; q < n; q++) {
            double dx = tx - s[0][q], dy = ty - s[1][q], dz = tz - s[2][q];
            if constexpr (KD == 1) {
                acc[0] += kval<KID>(dx * dx + dy * dy + dz * dz, kp) * sxv[0][q];
            } else {
                double t1, t2, ux, uy, uz;
                rpy_coef(dx, dy, dz, ta, s[3][q], kp.p1, t1, t2, ux, uy, uz);
                double ud = ux * sxv[0][q] + uy * sxv[1][q] + uz * sxv[2][q];
                acc[0] += t1 * sxv[0][q] + t2 * ux * ud;
                acc[1] += t1 * sxv[1][q] + t2 * uy * ud;
                acc[2] += t1 * sxv[2][q] + t2 * uz * ud;
            }
        }
    }
    if (i < n_dst)
#pragma unroll
        for (int a = 0; a < KD; a++) part[((long long)blockIdx.y * n_dst + i) * KD + a] = acc[a];
}

__global__ void k_nbody_reduce(const double *__restrict__ part, int nchunk, int len,
                               double *__restrict__ y) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += part[(long long)c * len + e];
    y[e] = s;
}

}  // namespace h2p
