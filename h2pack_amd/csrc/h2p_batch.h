// Co-scheduled execution of many resident H2 matrices.
//
// One launch serves the same phase of every active request: a workgroup
// finds its (request, task) pair by binary search over a small table of
// items (first workgroup index, request, task array of that request), then
// runs the same device body as the single-matrix launch.  Each request keeps
// its own scratch (xin, arena, y1, yp), so a request's result does not
// depend on which other requests share the launches.
//
// The batched conjugate gradients restate the reference's PCG with no
// preconditioner (examples/PCG/pcg.c:9-152; the shift of
// examples/SPDHSS-H2/pcg_tests.c:15-21 applied inside the product): all
// scalars stay on the device, dot products are per-request fixed-order
// reductions (partials per workgroup, summed in order).
#pragma once
#include "h2p_kernels.h"
#include "h2p_mm.h"

namespace h2p {

struct ReqCtx {
    // matrix (from the handle)
    const double *coord;
    long long CT;
    double scale, diag;
    KParam kp;
    double *xin, *y1, *yp, *arena;
    const double *U, *UT;
    const void *blob;
    const int *fwd;
    const long long *contrib;
    const double *shifts;    // shift table (entry 0: zero)
    int N, pad;
    // this call
    const double *x;
    double *y;
    double shift;
};

struct BatchItem {
    int begin;       // first workgroup of the item
    int req;         // index into the ReqCtx table
    const void *tasks;
    int nbig, ntask; // grouped block launches: workgroup-unit tasks, all tasks of the item
    int pad0, pad1;
};

// items[0..n) sorted by begin, items[n].begin = total workgroups
__device__ __forceinline__ int find_item(const BatchItem *__restrict__ items, int n, int bid) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].begin <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_gather_b(const BatchItem *__restrict__ items, int n,
                                                  const ReqCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const ReqCtx &c = reqs[items[it].req];
    const int i = (blockIdx.x - items[it].begin) * 256 + threadIdx.x;
    if (i < c.N) c.xin[i] = c.x[c.fwd[i]];
}

__global__ void __launch_bounds__(256) k_fwd2_b(const BatchItem *__restrict__ items, int n,
                                                const ReqCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const ReqCtx &c = reqs[items[it].req];
    const FwdTask t = ((const FwdTask *)items[it].tasks)[blockIdx.x - items[it].begin];
    fwd2_body(t, c.U, c.xin);
}

// WPE: waves per SIMD the kernel is compiled for (1: no constraint); the
// tiled kernels run at 4, as the single-matrix launch's occupancy variant.
template <int KID, typename T, int WPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_blocks_jit2_b(const BatchItem *__restrict__ items, int n, const ReqCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const ReqCtx &c = reqs[items[it].req];
    const BatchItem &I = items[it];
    blocks_jit2_unit<KID, T>((const BlkTask *)I.tasks, blockIdx.x - I.begin, I.nbig, I.ntask,
                             c.coord, c.CT, c.scale, c.diag, c.xin, c.arena);
}

template <int KID, int KD, typename T = double>
__global__ void __launch_bounds__(256) k_blocks_jit_b(const BatchItem *__restrict__ items, int n,
                                                      const ReqCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const ReqCtx &c = reqs[items[it].req];
    const BlkTask t = ((const BlkTask *)items[it].tasks)[blockIdx.x - items[it].begin];
    blocks_jit_body<KID, KD, T>(t, c.coord, c.CT, c.kp, c.shifts, c.xin, c.arena);
}

template <typename TS>
__global__ void __launch_bounds__(256) k_blocks_aot_b(const BatchItem *__restrict__ items, int n,
                                                      const ReqCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const ReqCtx &c = reqs[items[it].req];
    const BlkTask t = ((const BlkTask *)items[it].tasks)[blockIdx.x - items[it].begin];
    blocks_aot_body<TS>(t, (const TS *)c.blob, c.xin, c.arena);
}

__global__ void __launch_bounds__(256) k_reduce_y1_b(const BatchItem *__restrict__ items, int n,
                                                     const ReqCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const ReqCtx &c = reqs[items[it].req];
    const RedTask t = ((const RedTask *)items[it].tasks)[blockIdx.x - items[it].begin];
    reduce_y1_body(t, c.contrib, c.arena, c.y1);
}

__global__ void __launch_bounds__(256) k_bwd2_b(const BatchItem *__restrict__ items, int n,
                                                const ReqCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const ReqCtx &c = reqs[items[it].req];
    const BwdTask2 t = ((const BwdTask2 *)items[it].tasks)[blockIdx.x - items[it].begin];
    bwd2_body(t, c.UT, c.y1, c.yp);
}

__global__ void __launch_bounds__(256) k_leaf_final2_b(const BatchItem *__restrict__ items, int n,
                                                       const ReqCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const ReqCtx &c = reqs[items[it].req];
    const LeafTask2 t = ((const LeafTask2 *)items[it].tasks)[blockIdx.x - items[it].begin];
    leaf_final2_body(t, c.contrib, c.arena, c.UT, c.yp, c.fwd, c.y, c.xin, c.shift);
}

// ---------------------------------------------------------------------------
// Co-scheduled multi-RHS products: one launch per matmul phase serves every
// request of the batch, each with its own panels.
struct MMCtx {
    const MMSrc *srcs;
    const double *shifts;
    const double *U, *UT, *BD, *coord;
    long long CT;
    KParam kp;
    MMPanels P;
    const int *fwd;
    int N, nv;           // points x kd, vectors of this chunk
    const double *X;     // caller panels
    double *Y;
    long long ldx, ldy;
    int row_major, v0;
};

template <int KID, int KD, class G>
__global__ void __launch_bounds__(256) k_mm_b(const BatchItem *__restrict__ items, int n,
                                              const MMCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const MMCtx &c = reqs[items[it].req];
    const MMTask t = ((const MMTask *)items[it].tasks)[blockIdx.x - items[it].begin];
    mm_body<KID, KD, G>(t, c.srcs, c.U, c.UT, c.BD, c.coord, c.CT, c.kp, c.shifts, c.P);
}

// The FP32-mixed product phases (coupling, and the B/D part of the leaves)
// of FP32-mixed handles; BD is then the FP32 block store (AOT).
template <int KID, class G>
__global__ void __launch_bounds__(256) k_mm32_b(const BatchItem *__restrict__ items, int n,
                                                const MMCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const MMCtx &c = reqs[items[it].req];
    const MMTask t = ((const MMTask *)items[it].tasks)[blockIdx.x - items[it].begin];
    mm32_body<KID, G>(t, c.srcs, (const float *)c.BD, c.coord, c.CT, (float)c.kp.p0, c.shifts, c.P);
}

__global__ void __launch_bounds__(256) k_mm_gather_b(const BatchItem *__restrict__ items, int n,
                                                     const MMCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const MMCtx &c = reqs[items[it].req];
    const int nvp = c.P.nvp;
    const long long e = (long long)(blockIdx.x - items[it].begin) * 256 + threadIdx.x;
    if (e >= (long long)c.N * nvp) return;
    const int i = (int)(e / nvp), col = (int)(e - (long long)i * nvp);
    double v = 0.0;
    if (col < c.nv) {
        const long long src = c.fwd[i];
        v = c.row_major ? c.X[src * c.ldx + c.v0 + col] : c.X[(long long)(c.v0 + col) * c.ldx + src];
    }
    c.P.p[MM_PX][e] = v;
}

__global__ void __launch_bounds__(256) k_mm_scatter_b(const BatchItem *__restrict__ items, int n,
                                                      const MMCtx *__restrict__ reqs) {
    const int it = find_item(items, n, blockIdx.x);
    const MMCtx &c = reqs[items[it].req];
    const long long e = (long long)(blockIdx.x - items[it].begin) * 256 + threadIdx.x;
    if (e >= (long long)c.N * c.nv) return;
    const int i = (int)(e / c.nv), col = (int)(e - (long long)i * c.nv);
    const long long dstr = c.fwd[i];
    const double v = c.P.p[MM_PY][(long long)i * c.P.nvp + col];
    if (c.row_major) c.Y[dstr * c.ldy + c.v0 + col] = v;
    else c.Y[(long long)(c.v0 + col) * c.ldy + dstr] = v;
}

// ---------------------------------------------------------------------------
// Batched CG.  Vectors are in the caller's (original) point order.
// done: 0 iterating, 1 finished (||r|| <= tol ||b|| or max_iter), 2 paused
// for a residual replacement (see k_cg_finish).
struct CGScal {
    double rr;        // ||r||^2 of the current residual
    double rr_prev;
    double alpha;
    double ps;        // p . s
    double stop;      // tol ||b||
    double bb;        // ||b||^2
    double rr_true;   // ||r||^2 when r was last formed as b - A x
    double replace;   // pause when rr < replace * rr_true (0: never)
    int iter, done, max_iter, pad;
};

struct CGCtx {
    double *x, *r, *p, *s;
    const double *b;
    double *part;     // two partial sums per workgroup of this request
    CGScal *sc;
    double tol;
    double replace;   // squared residual-drop ratio that asks for a replacement (0: never)
    int N, nblk, max_iter, pad;
};

#define H2P_CG_PER_BLOCK 1024   // entries per workgroup (256 threads x 4)

// per-workgroup partial of a . b over the request's slice, written in order
__device__ __forceinline__ double block_dot(const double *a, const double *b, int i0, int N) {
    __shared__ double red[4];
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = i0 + q * 256 + threadIdx.x;
        if (i < N) v += a[i] * b[i];
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

__device__ __forceinline__ int find_cg(const int *__restrict__ first_blk, int n, int bid) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first_blk[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// r = b - s (s = A x + shift x already formed), partials of r . r and b . b
// (pcg.c:29-50)
__global__ void __launch_bounds__(256) k_cg_init(const int *__restrict__ first_blk, int n,
                                                 const CGCtx *__restrict__ cgs) {
    const int q = find_cg(first_blk, n, blockIdx.x);
    const CGCtx &c = cgs[q];
    const int b = blockIdx.x - first_blk[q];
    const int i0 = b * H2P_CG_PER_BLOCK;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int i = i0 + k * 256 + threadIdx.x;
        if (i < c.N) c.r[i] = c.b[i] - c.s[i];
    }
    __syncthreads();
    const double rr = block_dot(c.r, c.r, i0, c.N);
    __syncthreads();
    const double bb = block_dot(c.b, c.b, i0, c.N);
    if (threadIdx.x == 0) {
        c.part[2 * b] = rr;
        c.part[2 * b + 1] = bb;
    }
}

__global__ void k_cg_init_finish(int n, const CGCtx *__restrict__ cgs) {
    if ((int)blockIdx.x >= n) return;   // one workgroup per request
    const CGCtx &c = cgs[blockIdx.x];
    if (threadIdx.x != 0) return;
    double rr = 0.0, bb = 0.0;
    for (int b = 0; b < c.nblk; b++) {
        rr += c.part[2 * b];
        bb += c.part[2 * b + 1];
    }
    CGScal *sc = c.sc;
    sc->rr = rr;
    sc->rr_prev = 1.0;
    sc->alpha = 0.0;
    sc->ps = 0.0;
    sc->bb = bb;
    sc->rr_true = rr;
    sc->replace = c.replace;
    // stop when ||r|| <= tol ||b||, compared on the norms as pcg.c:47-62 does
    sc->stop = c.tol * sqrt(bb);
    sc->iter = 0;
    sc->max_iter = c.max_iter;
    sc->done = (c.max_iter <= 0 || !(sqrt(rr) > sc->stop)) ? 1 : 0;
}

// p = r (first iteration, a copy as pcg.c:86 makes: p may hold anything) or
// p = r + (rr / rr_prev) p (pcg.c:88-90 with z = r)
__global__ void __launch_bounds__(256) k_cg_p(const int *__restrict__ first_blk, int n,
                                              const CGCtx *__restrict__ cgs) {
    const int q = find_cg(first_blk, n, blockIdx.x);
    const CGCtx &c = cgs[q];
    const CGScal sc = *c.sc;
    if (sc.done) return;
    const int i0 = (blockIdx.x - first_blk[q]) * H2P_CG_PER_BLOCK;
    const bool first = sc.iter == 0;
    const double beta = first ? 0.0 : sc.rr / sc.rr_prev;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = i0 + q * 256 + threadIdx.x;
        if (i < c.N) c.p[i] = first ? c.r[i] : c.r[i] + beta * c.p[i];
    }
}

// partials of p . s
__global__ void __launch_bounds__(256) k_cg_dot_ps(const int *__restrict__ first_blk, int n,
                                                   const CGCtx *__restrict__ cgs) {
    const int q = find_cg(first_blk, n, blockIdx.x);
    const CGCtx &c = cgs[q];
    const int b = blockIdx.x - first_blk[q];
    const double v = block_dot(c.p, c.s, b * H2P_CG_PER_BLOCK, c.N);
    if (threadIdx.x == 0) c.part[b] = v;
}

// alpha = rr / (p . s), one workgroup per request, partials summed in order
__global__ void k_cg_alpha(int n, const CGCtx *__restrict__ cgs) {
    if ((int)blockIdx.x >= n) return;   // one workgroup per request
    const CGCtx &c = cgs[blockIdx.x];
    if (threadIdx.x != 0) return;
    CGScal *sc = c.sc;
    if (sc->done) return;
    double ps = 0.0;
    for (int b = 0; b < c.nblk; b++) ps += c.part[b];
    sc->ps = ps;
    sc->alpha = sc->rr / ps;
}

// x += alpha p, r -= alpha s, partials of r . r  (pcg.c:110-121)
__global__ void __launch_bounds__(256) k_cg_update(const int *__restrict__ first_blk, int n,
                                                   const CGCtx *__restrict__ cgs) {
    const int q = find_cg(first_blk, n, blockIdx.x);
    const CGCtx &c = cgs[q];
    const CGScal sc = *c.sc;
    const int b = blockIdx.x - first_blk[q];
    const int i0 = b * H2P_CG_PER_BLOCK;
    if (!sc.done) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = i0 + q * 256 + threadIdx.x;
            if (i < c.N) {
                c.x[i] += sc.alpha * c.p[i];
                c.r[i] -= sc.alpha * c.s[i];
            }
        }
    }
    const double v = block_dot(c.r, c.r, i0, c.N);
    if (threadIdx.x == 0) c.part[b] = v;
}

// rr update, iteration count, stop test (pcg.c:78, 121-124).  A solve with
// a replacement ratio pauses (done = 2) instead of finishing, and also once
// its recurrence residual has dropped by that ratio since r was last formed
// as b - A x: the caller then replaces r (k_cg_replace_finish) and the solve
// goes on with the same search direction.  Where it pauses depends only on
// its own iterates, not on how often the caller looks.
__global__ void k_cg_finish(int n, const CGCtx *__restrict__ cgs) {
    if ((int)blockIdx.x >= n) return;   // one workgroup per request
    const CGCtx &c = cgs[blockIdx.x];
    if (threadIdx.x != 0) return;
    CGScal *sc = c.sc;
    if (sc->done) return;
    double rr = 0.0;
    for (int b = 0; b < c.nblk; b++) rr += c.part[b];
    sc->rr_prev = sc->rr;
    sc->rr = rr;
    sc->iter += 1;
    const bool converged = !(sqrt(rr) > sc->stop);
    if (sc->replace > 0.0 && sc->iter < sc->max_iter && (converged || rr < sc->replace * sc->rr_true))
        sc->done = 2;
    else if (converged || sc->iter >= sc->max_iter)
        sc->done = 1;
}

// Residual replacement (the "reliable updates" of G. L. G. Sleijpen and
// H. A. van der Vorst, Computing 56, 1996): after r = b - s with s = A x +
// shift x formed by a more accurate product (k_cg_init's first half), rr and
// the stop test are redone on it; rr_prev stays, so the next search
// direction is r + (rr / rr_prev) p as usual.  Run on a running solve it
// replaces r mid-course just the same.
__global__ void k_cg_replace_finish(int n, const CGCtx *__restrict__ cgs) {
    if ((int)blockIdx.x >= n) return;   // one workgroup per request
    const CGCtx &c = cgs[blockIdx.x];
    if (threadIdx.x != 0) return;
    CGScal *sc = c.sc;
    double rr = 0.0;
    for (int b = 0; b < c.nblk; b++) rr += c.part[2 * b];
    sc->rr = rr;
    sc->rr_true = rr;
    // a finished solve (max_iter) stays finished, now reporting this residual
    if (sc->done != 1) sc->done = (!(sqrt(rr) > sc->stop) || sc->iter >= sc->max_iter) ? 1 : 0;
}

}  // namespace h2p
