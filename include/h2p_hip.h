/*
 * h2p_hip.h -- C ABI of the MI355X (gfx950) H2 matrix-vector / matrix-matrix
 * engine.  Drop-in for the apply half of H2Pack:
 *
 *   reference                                         replacement
 *   void H2P_matvec(H2Pack_p, const DTYPE *x, DTYPE *y)
 *       src/H2Pack_matvec.h:17, src/H2Pack_matvec.c:1162      h2p_hip_matvec
 *   void H2P_matmul(H2Pack_p, CBLAS_LAYOUT, int n_vec,
 *                   const DTYPE *X, int ldx, DTYPE *Y, int ldy)
 *       src/H2Pack_matmul.h:24-27, src/H2Pack_matmul.c:455     h2p_hip_matmul
 *   direct_nbody(...)  examples/direct_nbody.h:2-54            h2p_hip_direct_nbody
 *
 * The H2 representation is built on the host exactly as before
 * (H2P_partition_points + proxy points + H2P_build); h2p_hip_upload mirrors
 * it into HBM once.  The representation is passed as plain arrays
 * (h2p_hip_h2pack) whose fields carry the meaning of the struct H2Pack
 * fields of the same names (src/H2Pack_typedef.h:88-207).
 *
 * Conventions: every int-returning call returns 0 on success and a negative
 * H2P_HIP_E* code on failure (h2p_hip_last_error() describes the last
 * failure of the calling thread); nothing aborts.  A handle is immutable
 * after upload except for its scratch vectors: calls on one handle must not
 * overlap, calls on different handles may.  Results are deterministic: no
 * atomics touch output or intermediate sums.
 */
#ifndef H2P_HIP_H
#define H2P_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kernel ids; parameters as in src/H2Pack_3D_kernels.h */
#define H2P_HIP_KRNL_COULOMB   0  /* 1/r, K(r=0) = param[0] or 0 (n_param = 0) */
#define H2P_HIP_KRNL_GAUSSIAN  1  /* exp(-param[0] r^2)                        */
#define H2P_HIP_KRNL_MATERN32  2  /* (1+sqrt3 l r) exp(-sqrt3 l r), l=param[0] */
#define H2P_HIP_KRNL_RPY       3  /* 3x3 RPY tensor, param[0] = eta, radius is
                                     coordinate row 3 (xpt_dim = 4)            */

/* evaluation mode of the B and D blocks (H2Pack's BD_JIT) */
#define H2P_HIP_AOT 0   /* stored blocks streamed from HBM */
#define H2P_HIP_JIT 1   /* blocks evaluated in registers from coordinates */

/* arithmetic of the B/D phase */
#define H2P_HIP_FP64    0  /* everything in FP64 */
#define H2P_HIP_FP32MIX 1  /* FP32 kernel values (JIT) or FP32-stored blocks
                              (AOT) and FP32 B/D products, their FP32 partial
                              sums folded into FP64; U sweeps and the output
                              accumulation stay FP64.  Matvec: all kernels,
                              JIT and AOT (periodic handles: FP64 only).
                              Matmul: the scalar kernels, JIT and AOT, on
                              FP32 matrix cores (RPY: H2P_HIP_ENOSYS, there
                              are no FP32 RPY matmul tiles). */

/* CBLAS_LAYOUT values */
#define H2P_HIP_ROW_MAJOR 101
#define H2P_HIP_COL_MAJOR 102

/* error codes */
#define H2P_HIP_EINVAL  -1   /* bad argument or inconsistent representation */
#define H2P_HIP_ENOMEM  -2   /* device or host allocation failed            */
#define H2P_HIP_EDEVICE -3   /* HIP runtime error                            */
#define H2P_HIP_ENOSYS  -4   /* combination not supported                    */

typedef struct {
    int id;            /* H2P_HIP_KRNL_* */
    int n_param;       /* number of meaningful entries of param */
    double param[4];
} h2p_kernel_desc;

/* Fields of struct H2Pack read by the apply path (src/H2Pack_typedef.h).
 * All arrays are host memory and are only read during h2p_hip_upload. */
typedef struct {
    int pt_dim, xpt_dim, krnl_dim, n_point, krnl_mat_size;
    int n_node, n_leaf_node, max_child, max_level, min_adm_level;
    const int *children;      /* n_node * max_child, -1 padded            */
    const int *n_child;       /* n_node                                   */
    const int *node_level;    /* n_node                                   */
    const int *pt_cluster;    /* 2 * n_node, inclusive point ranges        */
    const int *mat_cluster;   /* 2 * n_node, inclusive vector ranges       */
    const int *level_n_node;  /* max_level + 1                            */
    const int *level_nodes;   /* (max_level + 1) * n_leaf_node            */
    const int *leaf_nodes;    /* n_leaf_node: height_nodes row 0 (D order) */
    int n_r_adm_pair, n_r_inadm_pair;
    const int *r_adm_pairs;   /* 2 * n_r_adm_pair                         */
    const int *r_inadm_pairs; /* 2 * n_r_inadm_pair                       */
    const int *node_n_r_adm;  /* n_node                                   */
    const int *U_nrow, *U_ncol;        /* n_node                          */
    const double *const *U;            /* n_node row-major U[i]->data     */
    const int *J_npt;                  /* n_node, skeleton point counts   */
    const double *const *J_coord;      /* n_node, xpt_dim x J_npt          */
    const double *coord;               /* xpt_dim x n_point, sorted        */
    const int *fwd_pmt_idx, *bwd_pmt_idx;  /* krnl_mat_size              */
    const int *B_nrow, *B_ncol;        /* n_r_adm_pair                     */
    const size_t *B_ptr;               /* n_r_adm_pair + 1                 */
    const double *B_data;              /* NULL: evaluated on the device    */
    const int *D_nrow, *D_ncol;        /* n_leaf_node + n_r_inadm_pair     */
    const size_t *D_ptr;               /* n_leaf_node + n_r_inadm_pair + 1 */
    const double *D_data;              /* NULL: evaluated on the device    */
    /* Periodic mode (H2P_run_RPY_Ewald + H2P_build_periodic, JIT RPY only;
     * H2P_matvec_periodic, src/H2Pack_matvec_periodic.c:394).  Pairs are
     * one-sided (node 0 receives), node 1's coordinates are shifted by the
     * pair's lattice offset, and the root carries a dense periodic block
     * applied to its skeleton coefficients.  periodic = 0: the fields below
     * are ignored. */
    int periodic;
    const double *per_adm_shifts;      /* n_r_adm_pair x pt_dim            */
    const double *per_inadm_shifts;    /* n_r_inadm_pair x pt_dim          */
    const double *per_blk;             /* U_ncol[root]^2, row-major        */
} h2p_hip_h2pack;

typedef struct h2p_hip_handle h2p_hip_handle;

/* Statistics mirroring H2Pack's mat_size / JIT_flops / timers model
 * (src/H2Pack_typedef.h:210-246, src/H2Pack_typedef.c:447-514). */
typedef struct {
    size_t U_elems, B_elems, D_elems;   /* stored FP64 elements            */
    size_t hbm_bytes_model;             /* 8*(2 U + B + D) + vector terms  */
    double B_pairs, D_pairs;            /* kernel evaluations per matvec   */
    double jit_flops_model;             /* krnl_bimv_flops * pairs         */
    size_t device_bytes;                /* HBM held by the handle          */
    int n_block_tasks, n_levels;
    int mode, prec;
    /* mean per-phase device time in milliseconds over the calls profiled
     * since the previous h2p_hip_get_info (HIP events on the call's stream) */
    double t_fwd, t_blocks, t_reduce, t_bwd, t_final, t_total;
    int n_profiled;
    int mv_variant;      /* variant of the matvec block launch (h2p_hip_autotune) */
    /* t_blocks split into its two launches on JIT handles of the scalar
     * kernels (B blocks, D blocks; profiled calls run them in order);
     * elsewhere t_blocks_b = t_blocks and t_blocks_d = 0 */
    double t_blocks_b, t_blocks_d;
    /* 1 when the D blocks of a JIT handle of the scalar kernels run on a
     * second stream beside the sweeps, 0 when they run in order
     * (h2p_hip_autotune) */
    int mv_d_concurrent;
    /* 1 when the levels above the leaf level (their upward sweep, B blocks,
     * skeleton sums and downward sweep) run on a third stream beside the
     * deep levels' B blocks, 0 when everything runs in order
     * (h2p_hip_autotune) */
    int mv_level_split;
} h2p_hip_info;

/* Mirror an H2 representation into HBM.  mode = H2P_HIP_AOT / _JIT,
 * prec = H2P_HIP_FP64 / _FP32MIX.  In AOT mode, NULL B_data / D_data ask
 * the device to evaluate the blocks itself (same values as H2P_build_B_AOT /
 * H2P_build_D_AOT, src/H2Pack_build.c:789-889, 1034-1140). */
int h2p_hip_upload(const h2p_hip_h2pack *h2, const h2p_kernel_desc *krnl,
                   int mode, int prec, h2p_hip_handle **out);

/* y := A_H2 x, host pointers, length krnl_mat_size, original point order;
 * synchronous; same semantics as H2P_matvec. */
int h2p_hip_matvec(h2p_hip_handle *h, const double *x, double *y);

/* Same on device pointers, enqueued on `stream` (a hipStream_t, NULL =
 * default stream); returns once enqueued. */
int h2p_hip_matvec_dev(h2p_hip_handle *h, const double *d_x, double *d_y,
                       void *stream);

/* Y := A_H2 X for n_vec columns; layout H2P_HIP_ROW_MAJOR (ldx >= n_vec) or
 * H2P_HIP_COL_MAJOR (ldx >= krnl_mat_size); host pointers; synchronous; same
 * semantics as H2P_matmul. */
int h2p_hip_matmul(h2p_hip_handle *h, int layout, int n_vec,
                   const double *X, int ldx, double *Y, int ldy);

/* Same on device pointers, enqueued on `stream`. */
int h2p_hip_matmul_dev(h2p_hip_handle *h, int layout, int n_vec,
                       const double *d_X, int ldx, double *d_Y, int ldy,
                       void *stream);

/* h2p_hip_matmul_dev without any allocation.  No counterpart in the
 * reference: this is for callers that cannot allocate device memory during
 * the call (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same kernels and tile variants as
 * h2p_hip_matmul_dev, bitwise the same result, launched directly (no
 * graph); the panels (permuted X, skeleton coefficients, permuted Y) live
 * in one static device array of H2P_HIP_NOALLOC_BYTES per GPU instead of
 * the handle's lazily allocated panels.  The array is shared by every call
 * on that GPU, so calls that overlap on one GPU (any handle, any stream)
 * race; it needs (2 krnl_mat_size + 2 sum(U_ncol)) * min(64, n_vec rounded
 * up to 16) doubles, and larger requests return H2P_HIP_EINVAL. */
#define H2P_HIP_NOALLOC_BYTES ((size_t)256 << 20)
int h2p_hip_matmul_dev_noalloc(h2p_hip_handle *h, int layout, int n_vec,
                               const double *d_X, int ldx, double *d_Y,
                               int ldy, void *stream);

/* Autotuner of the matmul tiles.  Each matmul phase (upward sweep, coupling,
 * downward sweep, leaf outputs) runs as grouped MFMA launches whose tile
 * geometry (rows per chunk, inner entries per LDS stage) is one of
 * H2P_HIP_MM_VARIANTS compiled variants.  This call times every variant on
 * the handle's own blocks for panels of n_vec columns and keeps the fastest
 * per phase; choices are shared, within the process, by handles whose phase
 * has the same kernel, mode and block-size histogram.  Every variant
 * computes bitwise the same result.  Optional outputs: summed phase times
 * of the default variants and of the chosen ones, in milliseconds.
 * n_vec = 0 tunes the matvec instead: its grouped B/D launch has two
 * builds of the same body compiled for different occupancies (JIT: 4 waves
 * per SIMD by default, unconstrained registers as the alternative; AOT:
 * unconstrained by default, 5 (FP64) or 8 (FP32) waves as the
 * alternative), and on JIT handles of the scalar kernels the D blocks run
 * either beside the sweeps on a second stream (default) or in order;
 * again with identical results.
 * (H2Pack has no counterpart: its products are fixed BLAS / SIMD loops.) */
#define H2P_HIP_MM_VARIANTS 4
int h2p_hip_autotune(h2p_hip_handle *h, int n_vec, double *t_default_ms,
                     double *t_tuned_ms);

/* Pin the tile variant (0 .. H2P_HIP_MM_VARIANTS-1) of one matmul phase
 * (0..3, or -1 for all) for panels of n_vec columns; variant -1 returns the
 * current choice of the phase (phase -1: of phase 0) instead.  Phase 4 is
 * the matvec's block launch (variants 0 and 1; n_vec is ignored), phase 5
 * the matvec's schedule on JIT handles of the scalar kernels (bit 0: D
 * blocks in order instead of beside the sweeps, bit 1: no level split;
 * every schedule gives bitwise the same y). */
int h2p_hip_mm_variant(h2p_hip_handle *h, int phase, int n_vec, int variant);

/* Record per-phase HIP events on each call (up to 512 calls between reads;
 * recording does not synchronise). */
int h2p_hip_set_profiling(h2p_hip_handle *h, int on);

/* Sizes, the byte/flop model and, after profiled calls, the mean per-phase
 * times (waits for the profiled calls to finish). */
int h2p_hip_get_info(h2p_hip_handle *h, h2p_hip_info *info);

void h2p_hip_free(h2p_hip_handle *h);

const char *h2p_hip_last_error(void);

/* FP64 dense kernel product on the device: y[i] = sum_j K(dst_i, src_j) x_j
 * for n_dst target points (host pointers, coordinates xpt_dim x ld).  The
 * accuracy reference of examples/direct_nbody.h:2-54. */
int h2p_hip_direct_nbody(const h2p_kernel_desc *krnl, int xpt_dim,
                         const double *src, int src_ld, int n_src,
                         const double *x,
                         const double *dst, int dst_ld, int n_dst, double *y);

/* ---- Co-scheduled execution of many resident matrices (serving) --------
 *
 * These calls batch what a reference program does one matrix at a time:
 * H2P_matvec under the PCG shift wrapper (examples/SPDHSS-H2/pcg_tests.c:15-21,
 * y = A x + shift x), the unpreconditioned pcg() of examples/PCG/pcg.c:11-152,
 * and H2P_matmul.  One launch per phase serves every handle of the batch (a
 * workgroup finds its handle and task through a small table), so blocks of
 * different matrices run side by side.  Each handle keeps its own scratch and
 * device-side solver state: a handle may appear once per call, and must not
 * be used by another call (batched or not) that overlaps with it.  Results
 * do not depend on which other handles share the batch.
 *
 * A batch context holds pinned staging for the launch tables; one context
 * serves one thread, and its calls may go to different streams. */
typedef struct h2p_hip_batch h2p_hip_batch;

/* Device bytes h2p_hip_upload would hold for this representation, plus the
 * matmul panels for n_vec columns (0: none) and the CG state (with_cg != 0).
 * Allocates nothing on the device. */
int h2p_hip_footprint(const h2p_hip_h2pack *h2, const h2p_kernel_desc *krnl,
                      int mode, int prec, int n_vec, int with_cg, size_t *bytes);

int h2p_hip_batch_create(h2p_hip_batch **out);
void h2p_hip_batch_free(h2p_hip_batch *b);

/* y_i := A_i x_i + shift_i x_i for i < n (device pointers, caller order;
 * shift may be NULL for zeros). */
int h2p_hip_matvec_batch(h2p_hip_batch *b, int n, h2p_hip_handle *const *hs,
                         const double *const *d_x, double *const *d_y,
                         const double *shift, void *stream);

/* Y_i := A_i X_i for i < n; one layout for all, per-handle n_vec and leading
 * dimensions (device pointers).  Same arithmetic as h2p_hip_matmul_dev. */
int h2p_hip_matmul_batch(h2p_hip_batch *b, int n, h2p_hip_handle *const *hs,
                         int layout, const int *n_vec,
                         const double *const *d_X, const int *ldx,
                         double *const *d_Y, const int *ldy, void *stream);

/* Start CG solves of (A_i + shift_i I) x_i = b_i: r = b - (A + shift I) x0
 * with x0 = the current contents of d_x[i], stopping once ||r|| <= tol_i ||b||
 * or after max_iter_i iterations (pcg.c:29-62).  d_b and d_x must stay valid
 * until the solve is polled done; x is updated in place. */
int h2p_hip_cg_start(h2p_hip_batch *b, int n, h2p_hip_handle *const *hs,
                     const double *const *d_b, double *const *d_x,
                     const double *shift, const double *tol,
                     const int *max_iter, void *stream);

/* Advance every listed solve by n_iter iterations of pcg.c:63-123 (no
 * preconditioner); solves done or paused are left unchanged. */
int h2p_hip_cg_step(h2p_hip_batch *b, int n, h2p_hip_handle *const *hs,
                    int n_iter, void *stream);

/* Residual replacement for the next h2p_hip_cg_start on h (no counterpart in
 * the reference's pcg.c, which has one precision): with ratio > 0 the solve
 * pauses (poll reports done = 2) when its recurrence residual claims
 * convergence or has dropped by `ratio` since the residual was last formed
 * as b - A x, so that the caller can form it again with a more accurate
 * product (h2p_hip_cg_replace).  0 (the default) never pauses; ratio in
 * [0, 1). */
int h2p_hip_cg_set_replace(h2p_hip_handle *h, double ratio);

/* Replace the residual of each listed solve: r = b - (A' + shift I) x with
 * the product taken by hs_acc[i] (for instance an FP64 handle of the same
 * representation as the FP32-mixed hs[i]; the same order N), then redo the
 * stop test on it; the search direction is kept.  A paused solve resumes
 * (done = 0) or finishes (done = 1: ||r|| <= tol ||b|| or max_iter). */
int h2p_hip_cg_replace(h2p_hip_batch *b, int n, h2p_hip_handle *const *hs,
                       h2p_hip_handle *const *hs_acc, void *stream);

/* Wait for `stream` and report each solve's state: done flag (0 running,
 * 1 done, 2 paused for a residual replacement), iterations, relative residual
 * ||r|| / ||b|| of the recurrence (pcg.c:125-127; after a replacement, of the
 * replaced residual). */
int h2p_hip_cg_poll(h2p_hip_batch *b, int n, h2p_hip_handle *const *hs,
                    int *done, int *iter, double *relres, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* H2P_HIP_H */
