/*
 * H2Pack_p -> MI355X engine binding (declared in include/h2p_hip_h2pack.h).
 *
 * Built together with the H2Pack sources: it includes their struct H2Pack
 * (src/H2Pack_typedef.h:88-207) and reads it field by field.  Nothing here
 * calls into H2Pack; the engine copies everything during h2p_hip_upload.
 */
#include <stdlib.h>
#include <string.h>

#include "H2Pack_typedef.h"
#include "h2p_hip_h2pack.h"

struct h2p_hip_h2pack_view {
    const double **U, **J_coord;
    int *U_nrow, *U_ncol, *J_npt;
    double *packed;   /* rows of any U / J_coord stored with ld > ncol */
};

void h2p_hip_h2pack_release(h2p_hip_h2pack_view *v) {
    if (!v) return;
    free((void *)v->U);
    free((void *)v->J_coord);
    free(v->U_nrow);
    free(v->U_ncol);
    free(v->J_npt);
    free(v->packed);
    free(v);
}

/* elements a packed copy of m needs (0 when m is already contiguous) */
static size_t pack_need(const H2P_dense_mat_p m) {
    if (!m || m->ncol <= 0 || m->nrow <= 0 || m->ld == m->ncol) return 0;
    return (size_t)m->nrow * (size_t)m->ncol;
}

/* m's data with row stride ncol: m->data itself, or a copy into *dst */
static const double *contiguous(const H2P_dense_mat_p m, double **dst) {
    if (!m || m->ncol <= 0 || m->nrow <= 0) return NULL;
    if (m->ld == m->ncol) return m->data;
    double *o = *dst;
    for (int i = 0; i < m->nrow; i++)
        memcpy(o + (size_t)i * m->ncol, m->data + (size_t)i * m->ld, sizeof(double) * (size_t)m->ncol);
    *dst += (size_t)m->nrow * m->ncol;
    return o;
}

int h2p_hip_h2pack_describe(H2Pack_p h, h2p_hip_h2pack *d, h2p_hip_h2pack_view **view) {
    if (!h || !d || !view) return H2P_HIP_EINVAL;
    *view = NULL;
    const int n = h->n_node;
    if (n <= 0 || !h->U || !h->J_coord) return H2P_HIP_EINVAL;
    h2p_hip_h2pack_view *v = (h2p_hip_h2pack_view *)calloc(1, sizeof *v);
    if (!v) return H2P_HIP_ENOMEM;
    v->U = (const double **)calloc((size_t)n, sizeof(double *));
    v->J_coord = (const double **)calloc((size_t)n, sizeof(double *));
    v->U_nrow = (int *)calloc((size_t)n, sizeof(int));
    v->U_ncol = (int *)calloc((size_t)n, sizeof(int));
    v->J_npt = (int *)calloc((size_t)n, sizeof(int));
    size_t need = 0;
    for (int i = 0; i < n; i++) need += pack_need(h->U[i]) + pack_need(h->J_coord[i]);
    v->packed = need ? (double *)malloc(sizeof(double) * need) : NULL;
    if (!v->U || !v->J_coord || !v->U_nrow || !v->U_ncol || !v->J_npt || (need && !v->packed)) {
        h2p_hip_h2pack_release(v);
        return H2P_HIP_ENOMEM;
    }
    double *pk = v->packed;
    for (int i = 0; i < n; i++) {
        /* nodes without a basis carry an empty U and J_coord
         * (src/H2Pack_build.c:241-263; rank 0: src/H2Pack_ID_compress.c:524-531) */
        const H2P_dense_mat_p U = h->U[i], J = h->J_coord[i];
        v->U_nrow[i] = U ? U->nrow : 0;
        v->U_ncol[i] = U ? U->ncol : 0;
        v->U[i] = contiguous(U, &pk);
        v->J_npt[i] = J ? J->ncol : 0;
        v->J_coord[i] = contiguous(J, &pk);
    }

    const int hss = h->is_HSS != 0;
    memset(d, 0, sizeof *d);
    d->pt_dim = h->pt_dim;
    d->xpt_dim = h->xpt_dim;
    d->krnl_dim = h->krnl_dim;
    d->n_point = h->n_point;
    d->krnl_mat_size = h->krnl_mat_size;
    d->n_node = n;
    d->n_leaf_node = h->n_leaf_node;
    d->max_child = h->max_child;
    d->max_level = h->max_level;
    d->min_adm_level = hss ? h->HSS_min_adm_level : h->min_adm_level;
    d->children = h->children;
    d->n_child = h->n_child;
    d->node_level = h->node_level;
    d->pt_cluster = h->pt_cluster;
    d->mat_cluster = h->mat_cluster;
    d->level_n_node = h->level_n_node;
    d->level_nodes = h->level_nodes;
    d->leaf_nodes = h->height_nodes;   /* height 0: the leaves, in D-block order */
    d->n_r_adm_pair = hss ? h->HSS_n_r_adm_pair : h->n_r_adm_pair;
    d->r_adm_pairs = hss ? h->HSS_r_adm_pairs : h->r_adm_pairs;
    d->n_r_inadm_pair = hss ? h->HSS_n_r_inadm_pair : h->n_r_inadm_pair;
    d->r_inadm_pairs = hss ? h->HSS_r_inadm_pairs : h->r_inadm_pairs;
    /* counted from the pair list H2P_build used (src/H2Pack_build.c:669-716) */
    d->node_n_r_adm = h->node_n_r_adm;
    d->U_nrow = v->U_nrow;
    d->U_ncol = v->U_ncol;
    d->U = v->U;
    d->J_npt = v->J_npt;
    d->J_coord = v->J_coord;
    d->coord = h->coord;
    d->fwd_pmt_idx = h->fwd_pmt_idx;
    d->bwd_pmt_idx = h->bwd_pmt_idx;
    d->B_nrow = h->B_nrow;
    d->B_ncol = h->B_ncol;
    d->B_ptr = h->B_ptr;
    d->B_data = h->BD_JIT ? NULL : h->B_data;
    d->D_nrow = h->D_nrow;
    d->D_ncol = h->D_ncol;
    d->D_ptr = h->D_ptr;
    d->D_data = h->BD_JIT ? NULL : h->D_data;
    if (h->is_RPY_Ewald) {
        d->periodic = 1;
        d->per_adm_shifts = h->per_adm_shifts;
        d->per_inadm_shifts = h->per_inadm_shifts;
        d->per_blk = h->per_blk;
    }
    *view = v;
    return 0;
}

int h2p_hip_h2pack_kernel(H2Pack_p h, int krnl_id, h2p_kernel_desc *k) {
    if (!h || !k) return H2P_HIP_EINVAL;
    memset(k, 0, sizeof *k);
    k->id = krnl_id;
    const double *p = (const double *)h->krnl_param;
    switch (krnl_id) {
        case H2P_HIP_KRNL_COULOMB:   /* NULL: the diagonal entry is 0 */
            if (p) {
                k->n_param = 1;
                k->param[0] = p[0];
            }
            return 0;
        case H2P_HIP_KRNL_GAUSSIAN:
        case H2P_HIP_KRNL_MATERN32:
        case H2P_HIP_KRNL_RPY:       /* eta; the radius is coordinate row 3 */
            if (!p) return H2P_HIP_EINVAL;
            k->n_param = 1;
            k->param[0] = p[0];
            return 0;
        default:
            return H2P_HIP_EINVAL;
    }
}

int h2p_hip_upload_h2pack(H2Pack_p h, int krnl_id, int prec, h2p_hip_handle **out) {
    if (!out) return H2P_HIP_EINVAL;
    *out = NULL;
    h2p_kernel_desc k;
    int rc = h2p_hip_h2pack_kernel(h, krnl_id, &k);
    if (rc) return rc;
    h2p_hip_h2pack d;
    h2p_hip_h2pack_view *v = NULL;
    if ((rc = h2p_hip_h2pack_describe(h, &d, &v))) return rc;
    rc = h2p_hip_upload(&d, &k, h->BD_JIT ? H2P_HIP_JIT : H2P_HIP_AOT, prec, out);
    h2p_hip_h2pack_release(v);
    return rc;
}
