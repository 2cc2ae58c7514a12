/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the reference's H2 apply path, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker
 * and as the timed CPU baseline.  The product never links or calls it.
 *
 * It restates, function by function:
 *   h2o_matvec       src/H2Pack_matvec.c:1162-1284 (H2P_matvec) with
 *                    fwd :95-177, intmd AOT :303-437, intmd JIT :586-764,
 *                    bwd :767-861, dense AOT :864-998, dense JIT :1002-1145,
 *                    thread-buffer reduce :1249-1274, permutations :1149-1159
 *   h2o_matmul       src/H2Pack_matmul.c:455-583 (H2P_matmul) with fwd :43-112,
 *                    intmd :140-245, bwd :248-350, dense :353-418,
 *                    row permutes :422-452, block lookup
 *                    src/H2Pack_utils.c:293-308, 330-451
 *   h2o_kernel_eval  src/H2Pack_3D_kernels.h Coulomb :52-96, Gaussian :169-209,
 *                    Matern32 :388-432, RPY :893-959
 *   h2o_direct_nbody examples/direct_nbody.h:2-54
 *
 * The reference itself cannot be compiled in this image (it needs cblas.h /
 * lapacke.h and the unvendored ASTER submodule), so this restatement is
 * pinned by the checks in tests/test_oracle.py instead: explicit dense
 * assembly of the H2 matrix on small trees, closed-form kernel values, the
 * FP64 direct n-body product against the build tolerance, matmul == matvec
 * per column, and AOT == JIT.
 */
#ifndef H2ORACLE_H
#define H2ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The H2Pack fields the apply path reads (src/H2Pack_typedef.h:88-207). */
typedef struct {
    int pt_dim, xpt_dim, krnl_dim, n_point, krnl_mat_size;
    int n_node, n_leaf_node, max_child, max_level, min_adm_level;
    const int *children, *n_child, *node_level, *pt_cluster, *mat_cluster;
    const int *level_n_node, *level_nodes, *leaf_nodes;
    int n_r_adm_pair, n_r_inadm_pair;
    const int *r_adm_pairs, *r_inadm_pairs, *node_n_r_adm;
    const int *U_nrow, *U_ncol;
    const double *const *U;
    const int *J_npt;
    const double *const *J_coord;     /* xpt_dim x J_npt, row-major */
    const double *coord;              /* xpt_dim x n_point */
    const int *fwd_pmt_idx, *bwd_pmt_idx;
    const int *B_nrow, *B_ncol;
    const size_t *B_ptr;
    const double *B_data;             /* NULL in JIT mode */
    const int *D_nrow, *D_ncol;
    const size_t *D_ptr;
    const double *D_data;             /* NULL in JIT mode */
    const int *B_p2i_rowptr, *B_p2i_colidx, *B_p2i_val;
    const int *D_p2i_rowptr, *D_p2i_colidx, *D_p2i_val;
    int krnl_id;
    double krnl_param[4];
    int n_krnl_param;                 /* 0 => Coulomb diagonal value 0 */
    int BD_JIT;
} h2o_h2pack;

void h2o_kernel_eval(int kid, const double *param, int n_param,
                     const double *c0, int ld0, int n0,
                     const double *c1, int ld1, int n1,
                     double *mat, int ldm);

int h2o_matvec(const h2o_h2pack *h, const double *x, double *y, int n_thread);

int h2o_matmul(const h2o_h2pack *h, int layout, int n_vec,
               const double *X, int ldx, double *Y, int ldy, int n_thread);

void h2o_direct_nbody(int kid, const double *param, int n_param, int kd,
                      const double *src, int src_ld, int n_src, const double *x,
                      const double *dst, int dst_ld, int n_dst, double *y,
                      int n_thread);

#ifdef __cplusplus
}
#endif
#endif
