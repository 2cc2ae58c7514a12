/*
 * Test helper: a struct H2Pack (the reference's own typedef,
 * src/H2Pack_typedef.h:88-207) filled from plain arrays the way H2P_build /
 * H2P_build_periodic / H2P_run_HSS leave it, so that the binding in
 * h2pack_amd/glue/h2p_hip_h2pack.c can be checked field by field and used
 * for an upload.  Compiled by tests/test_h2pack_binding.py against the
 * reference headers (skipped where they are absent).
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "H2Pack_typedef.h"
#include "h2p_hip_h2pack.h"

typedef struct {
    int pt_dim, xpt_dim, krnl_dim, n_point, krnl_mat_size;
    int n_node, n_leaf_node, max_child, max_level;
    int min_adm_level, HSS_min_adm_level;
    int n_r_adm_pair, n_r_inadm_pair, HSS_n_r_adm_pair, HSS_n_r_inadm_pair;
    int is_HSS, is_RPY_Ewald, BD_JIT;
    int ld_pad;   /* extra leading-dimension entries of every U and J_coord */
    int *children, *n_child, *node_level, *pt_cluster, *mat_cluster;
    int *level_n_node, *level_nodes, *height_nodes;
    int *r_adm_pairs, *r_inadm_pairs, *HSS_r_adm_pairs, *HSS_r_inadm_pairs, *node_n_r_adm;
    int *fwd_pmt_idx, *bwd_pmt_idx, *B_nrow, *B_ncol, *D_nrow, *D_ncol;
    size_t *B_ptr, *D_ptr;
    double *coord, *B_data, *D_data, *per_adm_shifts, *per_inadm_shifts, *per_blk;
    int *U_nrow, *U_ncol, *J_n;
    double **U, **J_coord;   /* contiguous rows (ld = ncol) */
    double *krnl_param;      /* NULL allowed (Coulomb) */
} h2s_src;

static H2P_dense_mat_p mat(int nrow, int ncol, int pad, const double *src) {
    H2P_dense_mat_p m = (H2P_dense_mat_p)calloc(1, sizeof(*m));
    m->nrow = nrow;
    m->ncol = ncol;
    m->ld = ncol > 0 ? ncol + pad : 0;
    m->size = nrow * (m->ld > 0 ? m->ld : 1);
    m->data = (double *)calloc((size_t)(m->size > 0 ? m->size : 1), sizeof(double));
    for (int i = 0; i < nrow && ncol > 0; i++) {
        memcpy(m->data + (size_t)i * m->ld, src + (size_t)i * ncol, sizeof(double) * ncol);
        for (int j = ncol; j < m->ld; j++) m->data[(size_t)i * m->ld + j] = -7.0;   /* never read */
    }
    return m;
}

H2Pack_p h2s_make(const h2s_src *s) {
    H2Pack_p h = (H2Pack_p)calloc(1, sizeof(struct H2Pack));
    h->n_thread = 1;
    h->pt_dim = s->pt_dim;
    h->xpt_dim = s->xpt_dim;
    h->krnl_dim = s->krnl_dim;
    h->n_point = s->n_point;
    h->krnl_mat_size = s->krnl_mat_size;
    h->n_node = s->n_node;
    h->root_idx = s->n_node - 1;
    h->n_leaf_node = s->n_leaf_node;
    h->max_child = s->max_child;
    h->max_level = s->max_level;
    h->min_adm_level = s->min_adm_level;
    h->HSS_min_adm_level = s->HSS_min_adm_level;
    h->n_r_adm_pair = s->n_r_adm_pair;
    h->n_r_inadm_pair = s->n_r_inadm_pair;
    h->HSS_n_r_adm_pair = s->HSS_n_r_adm_pair;
    h->HSS_n_r_inadm_pair = s->HSS_n_r_inadm_pair;
    h->n_UJ = s->n_node;
    h->n_B = s->is_HSS ? s->HSS_n_r_adm_pair : s->n_r_adm_pair;
    h->n_D = s->n_leaf_node + (s->is_HSS ? s->HSS_n_r_inadm_pair : s->n_r_inadm_pair);
    h->BD_JIT = s->BD_JIT;
    h->is_HSS = s->is_HSS;
    h->is_RPY = s->krnl_dim == 3;
    h->is_RPY_Ewald = s->is_RPY_Ewald;
    h->children = s->children;
    h->n_child = s->n_child;
    h->node_level = s->node_level;
    h->pt_cluster = s->pt_cluster;
    h->mat_cluster = s->mat_cluster;
    h->level_n_node = s->level_n_node;
    h->level_nodes = s->level_nodes;
    h->height_nodes = s->height_nodes;
    h->r_adm_pairs = s->r_adm_pairs;
    h->r_inadm_pairs = s->r_inadm_pairs;
    h->HSS_r_adm_pairs = s->HSS_r_adm_pairs;
    h->HSS_r_inadm_pairs = s->HSS_r_inadm_pairs;
    h->node_n_r_adm = s->node_n_r_adm;
    h->fwd_pmt_idx = s->fwd_pmt_idx;
    h->bwd_pmt_idx = s->bwd_pmt_idx;
    h->B_nrow = s->B_nrow;
    h->B_ncol = s->B_ncol;
    h->B_ptr = s->B_ptr;
    h->D_nrow = s->D_nrow;
    h->D_ncol = s->D_ncol;
    h->D_ptr = s->D_ptr;
    h->coord = s->coord;
    h->B_data = s->B_data;
    h->D_data = s->D_data;
    h->per_adm_shifts = s->per_adm_shifts;
    h->per_inadm_shifts = s->per_inadm_shifts;
    h->per_blk = s->per_blk;
    h->krnl_param = s->krnl_param;
    h->U = (H2P_dense_mat_p *)calloc((size_t)s->n_node, sizeof(H2P_dense_mat_p));
    h->J_coord = (H2P_dense_mat_p *)calloc((size_t)s->n_node, sizeof(H2P_dense_mat_p));
    for (int i = 0; i < s->n_node; i++) {
        h->U[i] = mat(s->U_nrow[i], s->U_ncol[i], s->ld_pad, s->U[i]);
        h->J_coord[i] = mat(s->J_n[i] > 0 ? s->xpt_dim : 0, s->J_n[i], s->ld_pad, s->J_coord[i]);
    }
    return h;
}

void h2s_free(H2Pack_p h) {
    if (!h) return;
    for (int i = 0; i < h->n_node; i++) {
        free(h->U[i]->data);
        free(h->U[i]);
        free(h->J_coord[i]->data);
        free(h->J_coord[i]);
    }
    free(h->U);
    free(h->J_coord);
    free(h);
}

int h2s_describe(H2Pack_p h, h2p_hip_h2pack *d, void **view) {
    return h2p_hip_h2pack_describe(h, d, (h2p_hip_h2pack_view **)view);
}

void h2s_release(void *view) { h2p_hip_h2pack_release((h2p_hip_h2pack_view *)view); }

int h2s_kernel(H2Pack_p h, int krnl_id, h2p_kernel_desc *k) { return h2p_hip_h2pack_kernel(h, krnl_id, k); }

int h2s_upload(H2Pack_p h, int krnl_id, int prec, h2p_hip_handle **out) {
    return h2p_hip_upload_h2pack(h, krnl_id, prec, out);
}

/* offsets of the fields the binding reads, for a layout sanity check */
size_t h2s_offsetof_HSS_r_adm_pairs(void) { return offsetof(struct H2Pack, HSS_r_adm_pairs); }
size_t h2s_sizeof_H2Pack(void) { return sizeof(struct H2Pack); }
