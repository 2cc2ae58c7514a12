/*
 * ORACLE -- test infrastructure only (see h2oracle.h).  CPU restatement of
 * the reference's H2P_matvec / H2P_matmul, its kernels and direct_nbody.
 * Plain C + OpenMP, no BLAS: the dense products the reference hands to CBLAS
 * are written out as loops.
 */
#include "h2oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define KRNL_COULOMB  0
#define KRNL_GAUSSIAN 1
#define KRNL_MATERN32 2
#define KRNL_RPY      3

/* ------------------------------------------------------------------------ */
/* Kernels: src/H2Pack_3D_kernels.h (Coulomb :52-96, Gaussian :169-209,      */
/* Matern32 :388-432, RPY :893-959).  Same block layout as krnl_eval.        */
/* ------------------------------------------------------------------------ */
void h2o_kernel_eval(int kid, const double *param, int n_param,
                     const double *c0, int ld0, int n0,
                     const double *c1, int ld1, int n1,
                     double *mat, int ldm)
{
    const double *x0 = c0, *y0 = c0 + ld0, *z0 = c0 + 2 * (size_t)ld0;
    const double *x1 = c1, *y1 = c1 + ld1, *z1 = c1 + 2 * (size_t)ld1;
    if (kid == KRNL_COULOMB) {
        const double dv = (n_param > 0) ? param[0] : 0.0;
        for (int i = 0; i < n0; i++) {
            double *row = mat + (size_t)i * ldm;
            for (int j = 0; j < n1; j++) {
                double dx = x0[i] - x1[j], dy = y0[i] - y1[j], dz = z0[i] - z1[j];
                double r2 = dx * dx + dy * dy + dz * dz;
                row[j] = (r2 == 0.0) ? dv : 1.0 / sqrt(r2);
            }
        }
    } else if (kid == KRNL_GAUSSIAN) {
        const double nl = -param[0];
        for (int i = 0; i < n0; i++) {
            double *row = mat + (size_t)i * ldm;
            for (int j = 0; j < n1; j++) {
                double dx = x0[i] - x1[j], dy = y0[i] - y1[j], dz = z0[i] - z1[j];
                row[j] = exp(nl * (dx * dx + dy * dy + dz * dz));
            }
        }
    } else if (kid == KRNL_MATERN32) {
        const double s = -1.7320508075688772 * param[0];
        for (int i = 0; i < n0; i++) {
            double *row = mat + (size_t)i * ldm;
            for (int j = 0; j < n1; j++) {
                double dx = x0[i] - x1[j], dy = y0[i] - y1[j], dz = z0[i] - z1[j];
                double r = sqrt(dx * dx + dy * dy + dz * dz) * s;
                row[j] = (1.0 - r) * exp(r);
            }
        }
    } else {
        const double *a0 = c0 + 3 * (size_t)ld0, *a1 = c1 + 3 * (size_t)ld1;
        const double C = 1.0 / (6.0 * M_PI * param[0]);
        for (int i = 0; i < n0; i++) {
            for (int j = 0; j < n1; j++) {
                double dx = x0[i] - x1[j], dy = y0[i] - y1[j], dz = z0[i] - z1[j];
                double ta = a0[i], sa = a1[j];
                double r2 = dx * dx + dy * dy + dz * dz;
                double r = sqrt(r2);
                double inv_r = (r == 0.0) ? 0.0 : 1.0 / r;
                double inv_r2 = inv_r * inv_r;
                dx *= inv_r; dy *= inv_r; dz *= inv_r;
                double t1, t2, tmp0, tmp1;
                double tps = ta + sa, tms = ta - sa;
                if (r > tps) {
                    tmp0 = C * 0.75 * inv_r;
                    tmp1 = (ta * ta + sa * sa) * inv_r2;
                    t1 = tmp0 * (1.0 + tmp1 / 3.0);
                    t2 = tmp0 * (1.0 - tmp1);
                } else if (r > fabs(tms)) {
                    tmp0 = tms * tms;
                    tmp1 = (inv_r2 * inv_r * C) / (ta * sa * 32.0);
                    t1 = tmp0 + 3.0 * r2;
                    t1 = tmp1 * (16.0 * r2 * r * tps - t1 * t1);
                    t2 = tmp0 - r2;
                    t2 = tmp1 * 3.0 * t2 * t2;
                } else {
                    t1 = C / (ta > sa ? ta : sa);
                    t2 = 0.0;
                }
                double *k = mat + 3 * (size_t)i * ldm + 3 * j;
                double u0 = t2 * dx, u1 = t2 * dy, u2 = t2 * dz;
                k[0] = u0 * dx + t1;     k[1] = u0 * dy;           k[2] = u0 * dz;
                k[ldm + 0] = u1 * dx;    k[ldm + 1] = u1 * dy + t1; k[ldm + 2] = u1 * dz;
                k[2 * ldm + 0] = u2 * dx; k[2 * ldm + 1] = u2 * dy; k[2 * ldm + 2] = u2 * dz + t1;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Dense helpers standing in for the CBLAS calls of the reference.           */
/* ------------------------------------------------------------------------ */

/* CBLAS_BI_GEMV, src/H2Pack_matvec.c:26-65:
 * x_out_0 += mat * x_in_0, x_out_1 += mat^T * x_in_1 */
static void bi_gemv(int nrow, int ncol, const double *mat, int ldm,
                    const double *xi0, const double *xi1, double *xo0, double *xo1)
{
    for (int i = 0; i < nrow; i++) {
        const double *row = mat + (size_t)i * ldm;
        const double v1 = xi1[i];
        double s = 0.0;
        #pragma omp simd reduction(+:s)
        for (int j = 0; j < ncol; j++) {
            s += row[j] * xi0[j];
            xo1[j] += v1 * row[j];
        }
        xo0[i] += s;
    }
}

/* y (= or +=) A^T x, A row-major nrow x ncol */
static void gemv_t(int nrow, int ncol, const double *A, int lda,
                   const double *x, double beta, double *y)
{
    if (beta == 0.0) memset(y, 0, sizeof(double) * ncol);
    for (int i = 0; i < nrow; i++) {
        const double *row = A + (size_t)i * lda;
        const double xi = x[i];
        #pragma omp simd
        for (int j = 0; j < ncol; j++) y[j] += row[j] * xi;
    }
}

/* y += A x, A row-major */
static void gemv_n_acc(int nrow, int ncol, const double *A, int lda,
                       const double *x, double *y)
{
    for (int i = 0; i < nrow; i++) {
        const double *row = A + (size_t)i * lda;
        double s = 0.0;
        #pragma omp simd reduction(+:s)
        for (int j = 0; j < ncol; j++) s += row[j] * x[j];
        y[i] += s;
    }
}

/* C (m x n, ldc) += op(A) (m x k) * B (k x n, ldb); opA = A (lda) or A^T */
static void gemm_acc(int transA, int m, int n, int k, const double *A, int lda,
                     const double *B, int ldb, double *C, int ldc)
{
    for (int i = 0; i < m; i++) {
        double *ci = C + (size_t)i * ldc;
        for (int l = 0; l < k; l++) {
            double a = transA ? A[(size_t)l * lda + i] : A[(size_t)i * lda + l];
            const double *bl = B + (size_t)l * ldb;
            #pragma omp simd
            for (int j = 0; j < n; j++) ci[j] += a * bl[j];
        }
    }
}

/* H2P_krnl_eval_bimv, src/H2Pack_matvec.c:558-582: evaluate the kernel block
 * in row blocks of about 128 KB and apply a bi-GEMV to each row block. */
static void krnl_eval_bimv(const h2o_h2pack *h,
                           const double *c0, int ld0, int n0,
                           const double *c1, int ld1, int n1,
                           const double *xi0, const double *xi1,
                           double *xo0, double *xo1, double *buf)
{
    const int kd = h->krnl_dim;
    const int ldm = n1 * kd;
    int blk_npt = (128 * 1024) / (int)(sizeof(double) * (ldm > 0 ? ldm : 1)) / kd;
    if (blk_npt < 1) blk_npt = 1;
    for (int s = 0; s < n0; s += blk_npt) {
        int npt = (s + blk_npt > n0) ? (n0 - s) : blk_npt;
        h2o_kernel_eval(h->krnl_id, h->krnl_param, h->n_krnl_param,
                        c0 + s, ld0, npt, c1, ld1, n1, buf, ldm);
        bi_gemv(npt * kd, ldm, buf, ldm, xi0, xi1 + s * kd, xo0 + s * kd, xo1);
    }
}

static size_t eval_buf_elems(const h2o_h2pack *h)
{
    /* 128 KB row block, at least one point row of the widest block */
    size_t maxcol = 0;
    for (int i = 0; i < h->n_leaf_node + h->n_r_inadm_pair; i++)
        if ((size_t)h->D_ncol[i] > maxcol) maxcol = h->D_ncol[i];
    for (int i = 0; i < h->n_r_adm_pair; i++)
        if ((size_t)h->B_ncol[i] > maxcol) maxcol = h->B_ncol[i];
    size_t e = 128 * 1024 / sizeof(double) + (size_t)h->krnl_dim * maxcol * h->krnl_dim;
    return e;
}

static int node_npt(const h2o_h2pack *h, int node)
{
    return h->pt_cluster[2 * node + 1] - h->pt_cluster[2 * node] + 1;
}

/* ------------------------------------------------------------------------ */
/* H2P_matvec                                                                */
/* ------------------------------------------------------------------------ */
int h2o_matvec(const h2o_h2pack *h, const double *x, double *y, int n_thread)
{
    const int N = h->krnl_mat_size, nn = h->n_node, kd = h->krnl_dim;
    const int nlf = h->n_leaf_node;
    if (n_thread <= 0) n_thread = omp_get_max_threads();
    size_t *soff = (size_t *)malloc(sizeof(size_t) * (nn + 1));
    soff[0] = 0;
    for (int i = 0; i < nn; i++) soff[i + 1] = soff[i] + h->U_ncol[i];
    const size_t S = soff[nn];
    double *pmt_x = (double *)malloc(sizeof(double) * N);
    double *pmt_y = (double *)calloc(N, sizeof(double));
    double *tby = (double *)calloc((size_t)n_thread * N, sizeof(double));
    double *y0 = (double *)calloc(S + 1, sizeof(double));
    double *y1 = (double *)calloc((size_t)n_thread * (S + 1), sizeof(double));
    int *y1set = (int *)calloc(nn, sizeof(int));
    const size_t ebuf = eval_buf_elems(h);
    double *ebufs = h->BD_JIT ? (double *)malloc(sizeof(double) * ebuf * n_thread) : NULL;
    if (!soff || !pmt_x || !pmt_y || !tby || !y0 || !y1 || !y1set || (h->BD_JIT && !ebufs)) {
        free(soff); free(pmt_x); free(pmt_y); free(tby); free(y0); free(y1); free(y1set); free(ebufs);
        return -1;
    }

    /* 1. forward permutation, src/H2Pack_matvec.c:1149-1152 */
    #pragma omp parallel for num_threads(n_thread)
    for (int i = 0; i < N; i++) pmt_x[i] = x[h->fwd_pmt_idx[i]];

    /* 3. upward sweep, src/H2Pack_matvec.c:95-177 */
    for (int lv = h->max_level; lv >= h->min_adm_level; lv--) {
        const int *lnodes = h->level_nodes + (size_t)lv * nlf;
        #pragma omp parallel for schedule(dynamic) num_threads(n_thread)
        for (int j = 0; j < h->level_n_node[lv]; j++) {
            int node = lnodes[j];
            int r = h->U_ncol[node];
            double *out = y0 + soff[node];
            if (r == 0) continue;
            if (h->n_child[node] == 0) {
                gemv_t(h->U_nrow[node], r, h->U[node], r, pmt_x + h->mat_cluster[2 * node], 0.0, out);
            } else {
                int srow = 0;
                for (int k = 0; k < h->n_child[node]; k++) {
                    int c = h->children[(size_t)node * h->max_child + k];
                    int rc = h->U_ncol[c];
                    gemv_t(rc, r, h->U[node] + (size_t)srow * r, r, y0 + soff[c], k == 0 ? 0.0 : 1.0, out);
                    srow += rc;
                }
            }
        }
    }

    /* 4. intermediate multiplication (thread-local y1 rows and y buffers),
     *    src/H2Pack_matvec.c:234-300 (y1 init/sum), :303-437 (AOT),
     *    :586-764 (JIT) */
    for (int i = 0; i < nn; i++) y1set[i] = h->node_n_r_adm[i] ? 1 : 0;
    #pragma omp parallel num_threads(n_thread)
    {
        int tid = omp_get_thread_num();
        double *ty = tby + (size_t)tid * N;
        double *ty1 = y1 + (size_t)tid * (S + 1);
        double *eb = ebufs ? ebufs + (size_t)tid * ebuf : NULL;
        #pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < h->n_r_adm_pair; i++) {
            int n0 = h->r_adm_pairs[2 * i], n1 = h->r_adm_pairs[2 * i + 1];
            int l0 = h->node_level[n0], l1 = h->node_level[n1];
            int br = h->B_nrow[i], bc = h->B_ncol[i];
            /* operands by level case, :333-379 */
            const double *in0 = (l0 > l1) ? pmt_x + h->mat_cluster[2 * n1] : y0 + soff[n1];
            const double *in1 = (l0 < l1) ? pmt_x + h->mat_cluster[2 * n0] : y0 + soff[n0];
            double *out0 = (l0 < l1) ? ty + h->mat_cluster[2 * n0] : ty1 + soff[n0];
            double *out1 = (l0 > l1) ? ty + h->mat_cluster[2 * n1] : ty1 + soff[n1];
            if (!h->BD_JIT) {
                bi_gemv(br, bc, h->B_data + h->B_ptr[i], bc, in0, in1, out0, out1);
            } else {
                const double *c0, *c1;
                int ld0, ld1, m0, m1;
                if (l0 < l1) { c0 = h->coord + h->pt_cluster[2 * n0]; ld0 = h->n_point; m0 = node_npt(h, n0); }
                else { c0 = h->J_coord[n0]; ld0 = h->J_npt[n0]; m0 = h->J_npt[n0]; }
                if (l0 > l1) { c1 = h->coord + h->pt_cluster[2 * n1]; ld1 = h->n_point; m1 = node_npt(h, n1); }
                else { c1 = h->J_coord[n1]; ld1 = h->J_npt[n1]; m1 = h->J_npt[n1]; }
                krnl_eval_bimv(h, c0, ld0, m0, c1, ld1, m1, in0, in1, out0, out1, eb);
            }
        }
        /* 6. dense multiplication, src/H2Pack_matvec.c:864-998 (AOT) and
         *    :1002-1145 (JIT) */
        #pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < nlf; i++) {
            int node = h->leaf_nodes[i];
            int vs = h->mat_cluster[2 * node];
            if (!h->BD_JIT) {
                gemv_n_acc(h->D_nrow[i], h->D_ncol[i], h->D_data + h->D_ptr[i], h->D_ncol[i], pmt_x + vs, ty + vs);
            } else {
                int m = node_npt(h, node);
                const double *c = h->coord + h->pt_cluster[2 * node];
                /* the diagonal block's second output is discarded (:1048-1059):
                 * evaluate row blocks and apply a plain GEMV */
                int ldm = m * kd;
                int blk = (128 * 1024) / (int)(sizeof(double) * ldm) / kd;
                if (blk < 1) blk = 1;
                for (int s = 0; s < m; s += blk) {
                    int np = (s + blk > m) ? (m - s) : blk;
                    h2o_kernel_eval(h->krnl_id, h->krnl_param, h->n_krnl_param, c + s, h->n_point, np, c, h->n_point, m, eb, ldm);
                    gemv_n_acc(np * kd, ldm, eb, ldm, pmt_x + vs, ty + vs + s * kd);
                }
            }
        }
        #pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < h->n_r_inadm_pair; i++) {
            int n0 = h->r_inadm_pairs[2 * i], n1 = h->r_inadm_pairs[2 * i + 1];
            int v0 = h->mat_cluster[2 * n0], v1 = h->mat_cluster[2 * n1];
            int di = nlf + i;
            if (!h->BD_JIT) {
                bi_gemv(h->D_nrow[di], h->D_ncol[di], h->D_data + h->D_ptr[di], h->D_ncol[di],
                        pmt_x + v1, pmt_x + v0, ty + v0, ty + v1);
            } else {
                krnl_eval_bimv(h, h->coord + h->pt_cluster[2 * n0], h->n_point, node_npt(h, n0),
                               h->coord + h->pt_cluster[2 * n1], h->n_point, node_npt(h, n1),
                               pmt_x + v1, pmt_x + v0, ty + v0, ty + v1, eb);
            }
        }
    }
    /* sum thread-local y1, src/H2Pack_matvec.c:272-300 */
    #pragma omp parallel for num_threads(n_thread)
    for (size_t k = 0; k < S; k++) {
        double s = 0.0;
        for (int t = 0; t < n_thread; t++) s += y1[(size_t)t * (S + 1) + k];
        y1[k] = s;
    }

    /* 5. downward sweep, src/H2Pack_matvec.c:767-861 (pmt_y written
     *    directly for leaves, children's y1 copied when unset, else added) */
    int maxrow = 1;
    for (int i = 0; i < nn; i++) if (h->U_nrow[i] > maxrow) maxrow = h->U_nrow[i];
    double *tmpall = (double *)malloc(sizeof(double) * (size_t)maxrow * n_thread);
    for (int lv = h->min_adm_level; lv <= h->max_level; lv++) {
        const int *lnodes = h->level_nodes + (size_t)lv * nlf;
        #pragma omp parallel for schedule(dynamic) num_threads(n_thread)
        for (int j = 0; j < h->level_n_node[lv]; j++) {
            int node = lnodes[j];
            if (!y1set[node]) continue;
            double *tmp = tmpall + (size_t)omp_get_thread_num() * maxrow;
            int nr = h->U_nrow[node], r = h->U_ncol[node];
            memset(tmp, 0, sizeof(double) * nr);
            gemv_n_acc(nr, r, h->U[node], r, y1 + soff[node], tmp);
            if (h->n_child[node] == 0) {
                double *dst = pmt_y + h->mat_cluster[2 * node];
                for (int k = 0; k < nr; k++) dst[k] += tmp[k];
            } else {
                int off = 0;
                for (int k = 0; k < h->n_child[node]; k++) {
                    int c = h->children[(size_t)node * h->max_child + k];
                    int rc = h->U_ncol[c];
                    double *dst = y1 + soff[c];
                    if (!y1set[c]) { memcpy(dst, tmp + off, sizeof(double) * rc); y1set[c] = 1; }
                    else for (int l = 0; l < rc; l++) dst[l] += tmp[off + l];
                    off += rc;
                }
            }
        }
    }
    free(tmpall);

    /* 7. reduce thread-local y, src/H2Pack_matvec.c:1249-1264 */
    #pragma omp parallel for num_threads(n_thread)
    for (int i = 0; i < N; i++) {
        double s = pmt_y[i];
        for (int t = 0; t < n_thread; t++) s += tby[(size_t)t * N + i];
        pmt_y[i] = s;
    }
    /* 8. backward permutation, src/H2Pack_matvec.c:1156-1159 */
    #pragma omp parallel for num_threads(n_thread)
    for (int i = 0; i < N; i++) y[i] = pmt_y[h->bwd_pmt_idx[i]];

    free(soff); free(pmt_x); free(pmt_y); free(tby); free(y0); free(y1); free(y1set); free(ebufs);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* H2P_matmul (pull formulation over the pair-to-block CSR)                   */
/* ------------------------------------------------------------------------ */

/* H2P_get_int_CSR_elem, src/H2Pack_utils.c:293-308 */
static int csr_get(const int *rp, const int *ci, const int *v, int row, int col)
{
    for (int i = rp[row]; i < rp[row + 1]; i++) if (ci[i] == col) return v[i];
    return 0;
}

/* H2P_get_Bij_block / H2P_get_Dij_block, src/H2Pack_utils.c:330-451: the
 * stored block for (node0_, node1_) into buf; *trans set when the caller
 * asked for the transposed orientation. */
static int get_block(const h2o_h2pack *h, int isB, int node0, int node1,
                     double *buf, int *nrow, int *ncol, int *trans)
{
    int idx = isB ? csr_get(h->B_p2i_rowptr, h->B_p2i_colidx, h->B_p2i_val, node0, node1)
                  : csr_get(h->D_p2i_rowptr, h->D_p2i_colidx, h->D_p2i_val, node0, node1);
    if (idx == 0) return -1;
    int a = node0, b = node1;
    *trans = 0;
    if (idx < 0) { *trans = 1; idx = -idx; a = node1; b = node0; }
    idx--;
    int kd = h->krnl_dim;
    *nrow = isB ? h->B_nrow[idx] : h->D_nrow[idx];
    *ncol = isB ? h->B_ncol[idx] : h->D_ncol[idx];
    if (!h->BD_JIT) {
        const double *src = isB ? h->B_data + h->B_ptr[idx] : h->D_data + h->D_ptr[idx];
        memcpy(buf, src, sizeof(double) * (size_t)(*nrow) * (*ncol));
        return 0;
    }
    const double *c0, *c1;
    int ld0, ld1, m0, m1;
    int la = h->node_level[a], lb = h->node_level[b];
    if (!isB || la < lb) { c0 = h->coord + h->pt_cluster[2 * a]; ld0 = h->n_point; m0 = node_npt(h, a); }
    else { c0 = h->J_coord[a]; ld0 = h->J_npt[a]; m0 = h->J_npt[a]; }
    if (!isB || la > lb) { c1 = h->coord + h->pt_cluster[2 * b]; ld1 = h->n_point; m1 = node_npt(h, b); }
    else { c1 = h->J_coord[b]; ld1 = h->J_npt[b]; m1 = h->J_npt[b]; }
    h2o_kernel_eval(h->krnl_id, h->krnl_param, h->n_krnl_param, c0, ld0, m0, c1, ld1, m1, buf, m1 * kd);
    return 0;
}

int h2o_matmul(const h2o_h2pack *h, int layout, int n_vec,
               const double *X, int ldx, double *Y, int ldy, int n_thread)
{
    const int N = h->krnl_mat_size, nn = h->n_node;
    const int MM = 128;   /* mm_max_n_vec default, src/H2Pack_typedef.c:108 */
    const int row_major = (layout == 101);
    if (n_thread <= 0) n_thread = omp_get_max_threads();
    size_t *soff = (size_t *)malloc(sizeof(size_t) * (nn + 1));
    soff[0] = 0;
    for (int i = 0; i < nn; i++) soff[i + 1] = soff[i] + h->U_ncol[i];
    const size_t S = soff[nn];
    size_t maxblk = 1;
    for (int i = 0; i < h->n_r_adm_pair; i++)
        if ((size_t)h->B_nrow[i] * h->B_ncol[i] > maxblk) maxblk = (size_t)h->B_nrow[i] * h->B_ncol[i];
    for (int i = 0; i < h->n_leaf_node + h->n_r_inadm_pair; i++)
        if ((size_t)h->D_nrow[i] * h->D_ncol[i] > maxblk) maxblk = (size_t)h->D_nrow[i] * h->D_ncol[i];
    int maxrow = 1;
    for (int i = 0; i < nn; i++) if (h->U_nrow[i] > maxrow) maxrow = h->U_nrow[i];
    double *px = (double *)malloc(sizeof(double) * (size_t)N * MM);
    double *py = (double *)malloc(sizeof(double) * (size_t)N * MM);
    double *y0 = (double *)malloc(sizeof(double) * (S + 1) * MM);
    double *y1 = (double *)malloc(sizeof(double) * (S + 1) * MM);
    int *y1set = (int *)calloc(nn, sizeof(int));
    double *bufs = (double *)malloc(sizeof(double) * (maxblk + (size_t)maxrow * MM) * n_thread);
    if (!soff || !px || !py || !y0 || !y1 || !y1set || !bufs) {
        free(soff); free(px); free(py); free(y0); free(y1); free(y1set); free(bufs);
        return -1;
    }
    for (int iv = 0; iv < n_vec; iv += MM) {
        const int nv = (iv + MM <= n_vec) ? MM : n_vec - iv;
        /* row permutation into a row-major panel (ld = nv),
         * src/H2Pack_matmul.c:422-436 */
        #pragma omp parallel for num_threads(n_thread)
        for (int i = 0; i < N; i++) {
            int src = h->fwd_pmt_idx[i];
            for (int v = 0; v < nv; v++)
                px[(size_t)i * nv + v] = row_major ? X[(size_t)src * ldx + iv + v]
                                                   : X[(size_t)(iv + v) * ldx + src];
        }
        memset(py, 0, sizeof(double) * (size_t)N * nv);
        /* fwd, src/H2Pack_matmul.c:43-112 */
        for (int lv = h->max_level; lv >= h->min_adm_level; lv--) {
            const int *ln = h->level_nodes + (size_t)lv * h->n_leaf_node;
            #pragma omp parallel for schedule(dynamic) num_threads(n_thread)
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                int node = ln[j], r = h->U_ncol[node];
                if (r == 0) continue;
                double *out = y0 + soff[node] * nv;
                memset(out, 0, sizeof(double) * (size_t)r * nv);
                if (h->n_child[node] == 0) {
                    gemm_acc(1, r, nv, h->U_nrow[node], h->U[node], r,
                             px + (size_t)h->mat_cluster[2 * node] * nv, nv, out, nv);
                } else {
                    int srow = 0;
                    for (int k = 0; k < h->n_child[node]; k++) {
                        int c = h->children[(size_t)node * h->max_child + k];
                        int rc = h->U_ncol[c];
                        gemm_acc(1, r, nv, rc, h->U[node] + (size_t)srow * r, r, y0 + soff[c] * nv, nv, out, nv);
                        srow += rc;
                    }
                }
            }
        }
        /* intmd (pull over B_p2i rows), src/H2Pack_matmul.c:140-245 */
        for (int i = 0; i < nn; i++) y1set[i] = h->node_n_r_adm[i] ? 1 : 0;
        #pragma omp parallel num_threads(n_thread)
        {
            double *buf = bufs + (size_t)omp_get_thread_num() * (maxblk + (size_t)maxrow * MM);
            #pragma omp for schedule(dynamic)
            for (int n0 = 0; n0 < nn; n0++) {
                int l0 = h->node_level[n0];
                double *y1n = y1 + soff[n0] * nv;
                if (h->U_ncol[n0] > 0) memset(y1n, 0, sizeof(double) * (size_t)h->U_ncol[n0] * nv);
                for (int e = h->B_p2i_rowptr[n0]; e < h->B_p2i_rowptr[n0 + 1]; e++) {
                    int n1 = h->B_p2i_colidx[e], l1 = h->node_level[n1];
                    int br, bc, tr;
                    get_block(h, 1, n0, n1, buf, &br, &bc, &tr);
                    int m = tr ? bc : br, k = tr ? br : bc;   /* op(B) is m x k */
                    if (l0 == l1)
                        gemm_acc(tr, m, nv, k, buf, bc, y0 + soff[n1] * nv, nv, y1n, nv);
                    if (l0 > l1)
                        gemm_acc(tr, m, nv, k, buf, bc, px + (size_t)h->mat_cluster[2 * n1] * nv, nv, y1n, nv);
                    if (l0 < l1)
                        gemm_acc(tr, m, nv, k, buf, bc, y0 + soff[n1] * nv, nv,
                                 py + (size_t)h->mat_cluster[2 * n0] * nv, nv);
                }
            }
        }
        /* bwd, src/H2Pack_matmul.c:248-350 */
        for (int lv = h->min_adm_level; lv <= h->max_level; lv++) {
            const int *ln = h->level_nodes + (size_t)lv * h->n_leaf_node;
            #pragma omp parallel for schedule(dynamic) num_threads(n_thread)
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                int node = ln[j];
                if (!y1set[node]) continue;
                double *tmp = bufs + (size_t)omp_get_thread_num() * (maxblk + (size_t)maxrow * MM) + maxblk;
                int nr = h->U_nrow[node], r = h->U_ncol[node];
                memset(tmp, 0, sizeof(double) * (size_t)nr * nv);
                gemm_acc(0, nr, nv, r, h->U[node], r, y1 + soff[node] * nv, nv, tmp, nv);
                if (h->n_child[node] == 0) {
                    double *dst = py + (size_t)h->mat_cluster[2 * node] * nv;
                    for (size_t k = 0; k < (size_t)nr * nv; k++) dst[k] += tmp[k];
                } else {
                    int off = 0;
                    for (int k = 0; k < h->n_child[node]; k++) {
                        int c = h->children[(size_t)node * h->max_child + k];
                        int rc = h->U_ncol[c];
                        double *dst = y1 + soff[c] * nv;
                        if (!y1set[c]) { memcpy(dst, tmp + (size_t)off * nv, sizeof(double) * (size_t)rc * nv); y1set[c] = 1; }
                        else for (size_t l = 0; l < (size_t)rc * nv; l++) dst[l] += tmp[(size_t)off * nv + l];
                        off += rc;
                    }
                }
            }
        }
        /* dense (pull over D_p2i rows), src/H2Pack_matmul.c:353-418 */
        #pragma omp parallel num_threads(n_thread)
        {
            double *buf = bufs + (size_t)omp_get_thread_num() * (maxblk + (size_t)maxrow * MM);
            #pragma omp for schedule(dynamic)
            for (int n0 = 0; n0 < nn; n0++) {
                double *dst = py + (size_t)h->mat_cluster[2 * n0] * nv;
                for (int e = h->D_p2i_rowptr[n0]; e < h->D_p2i_rowptr[n0 + 1]; e++) {
                    int n1 = h->D_p2i_colidx[e];
                    int dr, dc, tr;
                    get_block(h, 0, n0, n1, buf, &dr, &dc, &tr);
                    int m = tr ? dc : dr, k = tr ? dr : dc;
                    gemm_acc(tr, m, nv, k, buf, dc, px + (size_t)h->mat_cluster[2 * n1] * nv, nv, dst, nv);
                }
            }
        }
        /* backward row permutation, src/H2Pack_matmul.c:440-452 */
        #pragma omp parallel for num_threads(n_thread)
        for (int i = 0; i < N; i++) {
            int src = h->bwd_pmt_idx[i];
            for (int v = 0; v < nv; v++) {
                double val = py[(size_t)src * nv + v];
                if (row_major) Y[(size_t)i * ldy + iv + v] = val;
                else Y[(size_t)(iv + v) * ldy + i] = val;
            }
        }
    }
    free(soff); free(px); free(py); free(y0); free(y1); free(y1set); free(bufs);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* examples/direct_nbody.h:2-54: 256 x 256 kernel tiles, GEMV each.          */
/* ------------------------------------------------------------------------ */
void h2o_direct_nbody(int kid, const double *param, int n_param, int kd,
                      const double *src, int src_ld, int n_src, const double *x,
                      const double *dst, int dst_ld, int n_dst, double *y,
                      int n_thread)
{
    const int blk = 256;
    if (n_thread <= 0) n_thread = omp_get_max_threads();
    memset(y, 0, sizeof(double) * (size_t)n_dst * kd);
    #pragma omp parallel num_threads(n_thread)
    {
        double *buf = (double *)malloc(sizeof(double) * (size_t)blk * kd * blk * kd);
        #pragma omp for schedule(dynamic)
        for (int d = 0; d < n_dst; d += blk) {
            int nd = (d + blk > n_dst) ? n_dst - d : blk;
            for (int s = 0; s < n_src; s += blk) {
                int ns = (s + blk > n_src) ? n_src - s : blk;
                h2o_kernel_eval(kid, param, n_param, dst + d, dst_ld, nd, src + s, src_ld, ns, buf, ns * kd);
                gemv_n_acc(nd * kd, ns * kd, buf, ns * kd, x + (size_t)s * kd, y + (size_t)d * kd);
            }
        }
        free(buf);
    }
}
