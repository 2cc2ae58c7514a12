/*
 * h2p_hip_h2pack.h -- binding of H2Pack's H2Pack_p to the MI355X engine.
 *
 * Include after the H2Pack headers (H2Pack.h or H2Pack_typedef.h, which
 * define struct H2Pack, src/H2Pack_typedef.h:88-207); compile
 * h2pack_amd/glue/h2p_hip_h2pack.c with the H2Pack sources and link
 * libh2p_hip.so.  An H2Pack program then replaces
 *
 *   H2P_matvec(h2pack, x, y)                 src/H2Pack_matvec.h:17
 *   H2P_matmul(h2pack, layout, n, X, ldx, Y, ldy)   src/H2Pack_matmul.h:24-27
 *   H2P_matvec_periodic(h2pack, x, y)        src/H2Pack_matvec_periodic.h
 *
 * by h2p_hip_upload_h2pack once after H2P_build / H2P_build_periodic, then
 * h2p_hip_matvec / h2p_hip_matmul on the returned handle.
 */
#ifndef H2P_HIP_H2PACK_H
#define H2P_HIP_H2PACK_H

#include "h2p_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-node pointer arrays (and packed copies of any U or J_coord whose
 * leading dimension exceeds its column count) that a filled h2p_hip_h2pack
 * points into; valid until h2p_hip_h2pack_release. */
typedef struct h2p_hip_h2pack_view h2p_hip_h2pack_view;

/* Fill *d with the fields of h that the apply path reads.  The pair lists
 * and the minimum admissible level are the ones H2P_matvec uses: in HSS mode
 * (h->is_HSS) HSS_r_adm_pairs / HSS_n_r_adm_pair, HSS_r_inadm_pairs /
 * HSS_n_r_inadm_pair and HSS_min_adm_level (src/H2Pack_matvec.c:101, 308,
 * 592, 773, 904, 1009), otherwise the H2 lists.  A periodic representation
 * (h->is_RPY_Ewald) adds its lattice shifts and the root's periodic block
 * (src/H2Pack_matvec_periodic.c:394).  B_data / D_data are passed in AOT
 * mode (h->BD_JIT == 0) and NULL in JIT mode.  d's arrays alias h's (and the
 * view's) memory.  Returns 0 or H2P_HIP_EINVAL / H2P_HIP_ENOMEM. */
int h2p_hip_h2pack_describe(H2Pack_p h, h2p_hip_h2pack *d, h2p_hip_h2pack_view **view);

void h2p_hip_h2pack_release(h2p_hip_h2pack_view *view);

/* Kernel descriptor for kernel id krnl_id (H2P_HIP_KRNL_*) from the
 * parameter array the plugin reads (h->krnl_param, src/H2Pack_3D_kernels.h:
 * Coulomb :58 diagonal value or NULL, Gaussian :174 and Matern32 :393 l,
 * RPY :900 eta).  The plugin functions are static in a header, so the kernel
 * is named by the caller, not derived from h->krnl_eval. */
int h2p_hip_h2pack_kernel(H2Pack_p h, int krnl_id, h2p_kernel_desc *k);

/* describe + h2p_hip_upload + release: mode from h->BD_JIT, prec =
 * H2P_HIP_FP64 or H2P_HIP_FP32MIX.  h may be freed after this returns. */
int h2p_hip_upload_h2pack(H2Pack_p h, int krnl_id, int prec, h2p_hip_handle **out);

#ifdef __cplusplus
}
#endif
#endif /* H2P_HIP_H2PACK_H */
