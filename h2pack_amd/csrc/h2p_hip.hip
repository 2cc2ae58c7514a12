// MI355X H2 matvec / matmul engine: upload, schedules and the C ABI of
// include/h2p_hip.h.
//
// Phase order of one matvec (src/H2Pack_matvec.c:1162-1284):
//   gather x -> px                      (H2P_permute_vector_forward)
//   upward sweep, one launch per level  (H2P_matvec_fwd_transform)
//   ONE grouped launch of every B and D block, each block read or
//   evaluated once, two partial outputs per block
//                                       (intmd_mult_AOT/JIT + dense_mult_AOT/JIT)
//   fixed-order reduction of partials into y1
//   downward sweep, one launch per level (H2P_matvec_bwd_transform)
//   leaf outputs: U y1 + partials, scattered to the caller's order
//                                       (thread reduce + H2P_permute_vector_backward)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "h2p_hip.h"
#include "h2p_kernels.h"
#include "h2p_mm.h"
#include "h2p_batch.h"

using namespace h2p;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return fail(H2P_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dev_upload(T **dptr, const std::vector<T> &v, size_t *acct) {
    *dptr = nullptr;
    size_t bytes = sizeof(T) * std::max<size_t>(v.size(), 1);
    if (hipMalloc(dptr, bytes) != hipSuccess) return fail(H2P_HIP_ENOMEM, "hipMalloc failed");
    *acct += bytes;
    if (!v.empty()) HIP_TRY(hipMemcpy(*dptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
int dev_alloc(T **dptr, size_t n, size_t *acct) {
    *dptr = nullptr;
    size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
    if (hipMalloc(dptr, bytes) != hipSuccess) return fail(H2P_HIP_ENOMEM, "hipMalloc failed");
    *acct += bytes;
    return 0;
}

struct Level {
    int off, n;   // slice of a task array
};

inline bool fwdidx_ok(int v, int n) { return v >= 0 && v < n; }

}  // namespace

// profiling events recorded per matvec call (see launch_interior)
#define H2P_EV_PER_CALL 7
// share of the D blocks' kernel pairs launched after the B blocks (the
// smallest blocks), overlapping the skeleton sums and the downward sweep;
// the rest overlaps the upward sweep.  H2P_HIP_D_LATE overrides it.
#define H2P_D_LATE 0.5

struct h2p_hip_handle {
    int mode = 0, prec = 0, kid = 0, kd = 1, xpt_dim = 3;
    int N = 0, n_point = 0, S = 0;
    long long CT = 0;        // unified coordinate stride (points + skeleton points)
    KParam kp{};
    // device arrays
    double *d_coord = nullptr;   // xpt_dim x CT (RPY: plus a row of reciprocal radii)
    double *d_U = nullptr, *d_UT = nullptr;
    void *d_blob = nullptr;      // AOT B then D (double, or float in FP32-mixed mode)
    double *d_xin = nullptr;     // N + S
    double *d_y1 = nullptr;      // S
    double *d_arena = nullptr;
    int *d_fwd = nullptr;
    BlkTask *d_blk = nullptr;
    FwdTask *d_fwdt = nullptr;
    BwdTask2 *d_bwdt = nullptr;
    RedTask *d_red = nullptr;
    LeafTask2 *d_leaf = nullptr;
    double *d_yp = nullptr;      // S: push-down values of the downward sweep
    double *d_yt = nullptr;      // N kd: leaf (U y1) rows of a split handle (k_leaf_u)
    long long *d_contrib = nullptr;
    double *d_mm_x = nullptr, *d_mm_y = nullptr;  // staging of the host-pointer calls
    size_t mm_cap = 0;
    // matmul (pull-form MFMA tasks, built at upload)
    MMTask *d_mmt = nullptr;
    MMSrc *d_mms = nullptr;
    std::vector<Level> mm_fwd_levels, mm_bwd_levels;
    std::vector<BlkTask> h_blk;   // host copies, for the autotuner's histograms
    std::vector<MMTask> h_mmt;
    std::vector<MMSrc> h_mms;
    Level mm_intmd{0, 0}, mm_leaf{0, 0};
    // FP32-mixed matmul (scalar kernels, JIT or FP32-stored AOT): the
    // coupling phase and the B/D part of the leaf outputs (mm_leaf32, after
    // the U part mm_leaf) run in the FP32 product kernel
    bool mm_f32 = false;
    Level mm_leaf32{0, 0};
    double *d_panels = nullptr;   // PX | Y0 | Y1 | PY, panel_nvp wide
    int panel_nvp = 0;
    int mm_var[4][3] = {};        // tile variant per matmul phase and panel width class
    int mv_var = 0;               // variant of the matvec block launch (0 default, 1 other occupancy)
    // FP64 Gaussian / Matern JIT: the point extent can push the exponential's
    // argument past the unclamped range, so the launches use KGW / KMW
    bool wide = false;
    // periodic mode: shift table (x, y, z per entry; entry 0 is zero) and the
    // root's periodic block (also the matmul's dense source for the root)
    int periodic = 0;
    double *d_shift = nullptr;
    double *d_per = nullptr;
    int per_r = 0, per_soff = 0;
    bool mm_ok = false;
    std::vector<Level> fwd_levels, bwd_levels;
    int n_blk = 0, n_red = 0, n_leaf = 0;
    // grouped block launches: part 0 = B blocks (all blocks when not split;
    // on a split handle the B blocks of the deep levels), parts 1 and 2 = D
    // blocks launched before and after the B blocks, part 3 = the B blocks
    // of the shallow levels; each part a contiguous range of d_blk with its
    // workgroup-unit tasks first
    struct BlkPart { int off = 0, n = 0, big = 0, units = 0; };
    BlkPart part[4];
    bool split = false;             // D blocks on aux_stream, concurrent with the sweeps
    bool d_serial = false;          // split handle whose D blocks run in order (h2p_hip_autotune)
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_mid = nullptr, ev_join = nullptr;
    // level split of a split handle: the levels above lv_cut (upward sweep,
    // their B blocks, skeleton sums and downward sweep) run on lv_stream
    // while the deep levels' B blocks run; lv_on = 0 runs them in order
    // (h2p_hip_autotune).  Shallow entries come first in red, deep ones
    // first in fwd_levels.
    int lv_cut = 0, lv_on = 0, lv_possible = 0;
    int n_fwd_deep = 0, n_bwd_shallow = 0, n_red_shallow = 0;
    hipStream_t lv_stream = nullptr;
    hipEvent_t ev_lv = nullptr, ev_side = nullptr;
    long long arena_len = 0;
    size_t device_bytes = 0;
    h2p_hip_info info{};
    // per-phase timing: one set of 6 events per profiled call, read back and
    // averaged by h2p_hip_get_info, so the profiled calls do not synchronise
    int profiling = 0;
    std::vector<hipEvent_t> ev_pool;
    int ev_used = 0;      // calls recorded since the last read
    int ev_cap = 0;       // calls the pool holds
    hipStream_t own_stream = nullptr;
    // the interior of a matvec (everything between the gather of x and the
    // leaf outputs) touches only handle-owned buffers: captured once, replayed
    int use_graph = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // the whole matvec is one graph: the gather of the caller's x, the
    // interior and the leaf outputs into the caller's y; the two boundary
    // kernel nodes get the next call's x / y when they change
    hipGraphNode_t gnode_gather = nullptr, gnode_leaf = nullptr;
    const double *graph_x = nullptr;
    double *graph_y = nullptr;
    // recorded after every graph replay of this handle (matvec and matmul),
    // so that dropping a graph waits for this handle's work only
    hipEvent_t ev_replay = nullptr;
    bool replay_pending = false;
    // the interior of a matmul chunk (all product phases on the panels), one
    // graph per panel width; dropped when the panels move or a variant changes
    struct MMGraph { int nvp; hipGraph_t g; hipGraphExec_t e; };
    std::vector<MMGraph> mm_graphs;
    // device-side CG state of the batched solver (allocated by the first start)
    double *d_cg = nullptr;       // r | p | s | partials | scalars
    double *cg_r = nullptr, *cg_p = nullptr, *cg_s = nullptr, *cg_part = nullptr;
    CGScal *cg_sc = nullptr;
    double *cg_x = nullptr;
    const double *cg_b = nullptr;
    double cg_shift = 0.0, cg_tol = 0.0;
    double cg_replace = 0.0;        // squared residual-drop ratio of the next solve (h2p_hip_cg_set_replace)
    int cg_max_iter = 0, cg_nblk = 0, cg_started = 0;
};

static void free_handle(h2p_hip_handle *h) {
    if (!h) return;
    void *ptrs[] = {h->d_coord, h->d_U, h->d_UT, h->d_blob, h->d_xin, h->d_y1, h->d_arena,
                    h->d_fwd, h->d_blk, h->d_fwdt, h->d_bwdt, h->d_red, h->d_leaf,
                    h->d_contrib, h->d_mm_x, h->d_mm_y, h->d_yp, h->d_yt, h->d_mmt, h->d_mms, h->d_panels,
                    h->d_cg, h->d_shift, h->d_per};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &e : h->ev_pool) (void)hipEventDestroy(e);
    for (auto &m : h->mm_graphs) {
        (void)hipGraphExecDestroy(m.e);
        (void)hipGraphDestroy(m.g);
    }
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    if (h->aux_stream) (void)hipStreamDestroy(h->aux_stream);
    if (h->lv_stream) (void)hipStreamDestroy(h->lv_stream);
    if (h->ev_lv) (void)hipEventDestroy(h->ev_lv);
    if (h->ev_side) (void)hipEventDestroy(h->ev_side);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_mid) (void)hipEventDestroy(h->ev_mid);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    if (h->ev_replay) (void)hipEventDestroy(h->ev_replay);
    delete h;
}

template <int KID, int KD>
static void launch_jit(h2p_hip_handle *h, hipStream_t s) {
    if (h->prec == H2P_HIP_FP32MIX)
        hipLaunchKernelGGL((k_blocks_jit<KID, KD, float>), dim3(h->n_blk), dim3(256), 0, s, h->d_blk,
                           h->d_coord, h->CT, h->kp, h->d_shift, h->d_xin, h->d_arena);
    else
        hipLaunchKernelGGL((k_blocks_jit<KID, KD, double>), dim3(h->n_blk), dim3(256), 0, s, h->d_blk,
                           h->d_coord, h->CT, h->kp, h->d_shift, h->d_xin, h->d_arena);
}

template <int KID, int KD>
static void launch_build(h2p_hip_handle *h, hipStream_t s) {
    if (h->prec == H2P_HIP_FP32MIX)
        hipLaunchKernelGGL((k_build_blocks<KID, KD, float>), dim3(h->n_blk), dim3(256), 0, s, h->d_blk,
                           h->d_coord, h->CT, h->kp, (float *)h->d_blob);
    else
        hipLaunchKernelGGL((k_build_blocks<KID, KD, double>), dim3(h->n_blk), dim3(256), 0, s, h->d_blk,
                           h->d_coord, h->CT, h->kp, (double *)h->d_blob);
}

// variant 0 (default): the kernel compiled for 4 waves per SIMD; variant 1:
// compiled without an occupancy target (more registers, 2-3 waves per SIMD)
template <int KID, typename T, int PART>
static void launch_jit2_part(h2p_hip_handle *h, hipStream_t s, const h2p_hip_handle::BlkPart &P) {
    if (P.n == 0) return;
    const BlkTask *tasks = h->d_blk + P.off;
    if constexpr (KID == KGW || KID == KMW) {   // one build for the clamped kernels
        hipLaunchKernelGGL((k_blocks_jit2_occ<KID, T, 4, PART>), dim3(P.units), dim3(256), 0, s, tasks, P.big, P.n,
                           h->d_coord, h->CT, kscale(KID, h->kp.p0, sizeof(T) == 8), h->kp.p0, h->d_xin, h->d_arena);
        return;
    }
    if (h->mv_var == 1)
        hipLaunchKernelGGL((k_blocks_jit2<KID, T, PART>), dim3(P.units), dim3(256), 0, s, tasks, P.big, P.n,
                           h->d_coord, h->CT, kscale(KID, h->kp.p0, sizeof(T) == 8), h->kp.p0, h->d_xin, h->d_arena);
    else
        hipLaunchKernelGGL((k_blocks_jit2_occ<KID, T, 4, PART>), dim3(P.units), dim3(256), 0, s, tasks, P.big, P.n,
                           h->d_coord, h->CT, kscale(KID, h->kp.p0, sizeof(T) == 8), h->kp.p0, h->d_xin, h->d_arena);
}

template <int KID, typename T>
static void launch_jit2(h2p_hip_handle *h, hipStream_t s, int part) {
    // the B parts (0, 3) share one kernel instantiation (template argument
    // 0), the D parts (1, 2) another
    if (part == 0 || part == 3) launch_jit2_part<KID, T, 0>(h, s, h->part[part]);
    else launch_jit2_part<KID, T, 1>(h, s, h->part[part]);
}

static void dispatch_jit(h2p_hip_handle *h, hipStream_t s, int part) {
    const bool f32 = h->prec == H2P_HIP_FP32MIX;
    switch (h->kid) {
        case KC: f32 ? launch_jit2<KC, float>(h, s, part) : launch_jit2<KC, double>(h, s, part); break;
        case KG:
            if (f32) launch_jit2<KG, float>(h, s, part);
            else if (h->wide) launch_jit2<KGW, double>(h, s, part);
            else launch_jit2<KG, double>(h, s, part);
            break;
        case KM:
            if (f32) launch_jit2<KM, float>(h, s, part);
            else if (h->wide) launch_jit2<KMW, double>(h, s, part);
            else launch_jit2<KM, double>(h, s, part);
            break;
        default:
            if (part == 0) launch_jit<KR, 3>(h, s);
            break;
    }
}

static void dispatch_build(h2p_hip_handle *h, hipStream_t s) {
    switch (h->kid) {
        case KC: launch_build<KC, 1>(h, s); break;
        case KG: launch_build<KG, 1>(h, s); break;
        case KM: launch_build<KM, 1>(h, s); break;
        default: launch_build<KR, 3>(h, s); break;
    }
}

// Part `part` of the grouped B/D launch of a matvec (AOT variant 1: the
// occupancy kernels, FP64 at 6 waves per SIMD, FP32 at 8).  AOT and RPY
// handles are not split: part 0 holds every block.
static void launch_blocks(h2p_hip_handle *h, hipStream_t s, int part) {
    if (h->part[part].n == 0) return;
    if (h->mode != H2P_HIP_AOT) {
        dispatch_jit(h, s, part);
        return;
    }
    const dim3 g(h->n_blk), b(256);
    if (h->mode == H2P_HIP_AOT && h->prec == H2P_HIP_FP32MIX) {
        if (h->mv_var == 1)
            hipLaunchKernelGGL((k_blocks_aot_occ<float, 8>), g, b, 0, s, h->d_blk, (const float *)h->d_blob,
                               h->d_xin, h->d_arena);
        else
            hipLaunchKernelGGL(k_blocks_aot<float>, g, b, 0, s, h->d_blk, (const float *)h->d_blob, h->d_xin,
                               h->d_arena);
    } else if (h->mode == H2P_HIP_AOT) {
        if (h->mv_var == 1)
            hipLaunchKernelGGL((k_blocks_aot_occ<double, 6>), g, b, 0, s, h->d_blk, (const double *)h->d_blob,
                               h->d_xin, h->d_arena);
        else
            hipLaunchKernelGGL(k_blocks_aot<double>, g, b, 0, s, h->d_blk, (const double *)h->d_blob, h->d_xin,
                               h->d_arena);
    }
}

// Interior of a matvec: upward sweep, grouped B/D launches, skeleton sums,
// downward sweep (reads xin[0, N) = px, writes yp / y1 / arena).  On a split
// handle the D blocks, which read only px, run on aux_stream and are joined
// at the end of the interior, so that their kernel evaluations overlap the
// HBM-bound sweeps: part 1 from the start of the interior (beside the
// upward sweep), part 2 once the deep B blocks are done (beside the
// skeleton sums, the downward sweep and the leaves' U y1 rows, k_leaf_u,
// which ends the interior of a split handle; k_leaf_d then adds the D
// contributions).  With the level split on (lv_on), the levels above lv_cut
// form a chain of their own on lv_stream once the deep levels' upward sweep
// is done -- their upward sweep, their B blocks (part 3), their skeleton
// sums and their downward sweep -- beside the deep levels' B blocks (part 0)
// on s; s then adds the deep skeleton sums, waits for the chain and runs the
// deep downward sweep.  Every kernel and every sum is the same in any
// schedule, so the results are bitwise the same.
// With profiling events (ev: 7 per call) every phase runs in order on s
// instead: ev[1] after the upward sweep, ev[2] after the B blocks, ev[3]
// after the D blocks, ev[4] after the skeleton sums, ev[5] after the
// downward sweep.
static void launch_fwd(h2p_hip_handle *h, hipStream_t s, int l0, int l1) {
    for (int k = l0; k < l1; k++) {
        const Level &L = h->fwd_levels[k];
        hipLaunchKernelGGL(k_fwd2, dim3(L.n), dim3(256), 0, s, h->d_fwdt + L.off, h->d_U, h->d_xin);
    }
}

static void launch_bwd(h2p_hip_handle *h, hipStream_t s, int l0, int l1) {
    for (int k = l0; k < l1; k++) {
        const Level &L = h->bwd_levels[k];
        hipLaunchKernelGGL(k_bwd2, dim3(L.n), dim3(256), 0, s, h->d_bwdt + L.off, h->d_UT, h->d_y1, h->d_yp);
    }
}

static void launch_reduce(h2p_hip_handle *h, hipStream_t s, int t0, int t1) {
    if (t1 > t0)
        hipLaunchKernelGGL(k_reduce_y1, dim3(t1 - t0), dim3(256), 0, s, h->d_red + t0, h->d_contrib, h->d_arena,
                           h->d_y1);
}

static void launch_interior(h2p_hip_handle *h, hipStream_t s, hipEvent_t *ev) {
    const bool fork = h->split && !h->d_serial && !ev && (h->part[1].n > 0 || h->part[2].n > 0);
    const bool lv = h->split && h->lv_on && h->lv_possible && !ev;
    const int nf = (int)h->fwd_levels.size(), nbw = (int)h->bwd_levels.size();
    const int fd = lv ? h->n_fwd_deep : nf, bs = lv ? h->n_bwd_shallow : 0, rs = lv ? h->n_red_shallow : 0;
    if (fork) {
        (void)hipEventRecord(h->ev_fork, s);
        (void)hipStreamWaitEvent(h->aux_stream, h->ev_fork, 0);
        launch_blocks(h, h->aux_stream, 1);
    }
    launch_fwd(h, s, 0, fd);
    if (lv) {
        hipStream_t q = h->lv_stream;
        (void)hipEventRecord(h->ev_lv, s);
        (void)hipStreamWaitEvent(q, h->ev_lv, 0);
        launch_fwd(h, q, fd, nf);
        launch_blocks(h, q, 3);
        launch_reduce(h, q, 0, rs);
        launch_bwd(h, q, 0, bs);
        (void)hipEventRecord(h->ev_side, q);
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    launch_blocks(h, s, 0);
    if (!lv) launch_blocks(h, s, 3);
    if (ev) (void)hipEventRecord(ev[2], s);
    if (fork) {
        if (h->part[2].n > 0) {
            (void)hipEventRecord(h->ev_mid, s);
            (void)hipStreamWaitEvent(h->aux_stream, h->ev_mid, 0);
            launch_blocks(h, h->aux_stream, 2);
        }
        (void)hipEventRecord(h->ev_join, h->aux_stream);
    } else {
        launch_blocks(h, s, 1);
        launch_blocks(h, s, 2);
    }
    if (ev) (void)hipEventRecord(ev[3], s);
    launch_reduce(h, s, rs, h->n_red);
    if (h->periodic)   // (periodic handles are not split)
        hipLaunchKernelGGL(k_root_block, dim3((h->per_r + 3) / 4), dim3(256), 0, s, h->d_per, h->per_r,
                           h->d_xin + h->N + h->per_soff, h->d_y1 + h->per_soff);
    if (ev) (void)hipEventRecord(ev[4], s);
    if (lv) (void)hipStreamWaitEvent(s, h->ev_side, 0);
    launch_bwd(h, s, bs, nbw);
    if (ev) (void)hipEventRecord(ev[5], s);
    if (h->split)
        hipLaunchKernelGGL(k_leaf_u, dim3(h->n_leaf), dim3(256), 0, s, h->d_leaf, h->d_contrib, h->d_arena, h->d_UT,
                           h->d_yp, h->d_yt);
    if (fork) (void)hipStreamWaitEvent(s, h->ev_join, 0);
}

// The leaf outputs of a matvec into the caller's y (the last launch).
static void launch_leaf_out(h2p_hip_handle *h, double *d_y, hipStream_t s) {
    if (h->split)
        hipLaunchKernelGGL(k_leaf_d, dim3(h->n_leaf), dim3(256), 0, s, h->d_leaf, h->d_contrib, h->d_arena,
                           h->d_yt, h->d_fwd, d_y, (const double *)h->d_xin, 0.0);
    else
        hipLaunchKernelGGL(k_leaf_final2, dim3(h->n_leaf), dim3(256), 0, s, h->d_leaf, h->d_contrib, h->d_arena,
                           h->d_UT, h->d_yp, h->d_fwd, d_y, (const double *)h->d_xin, 0.0);
}

// Point the graph's gather node at x and its leaf-output node at y.  The
// argument lists repeat the kernels' parameters one for one.
static int graph_set_xy(h2p_hip_handle *h, const double *d_x, double *d_y) {
    if (d_x != h->graph_x) {
        hipKernelNodeParams p{};
        HIP_TRY(hipGraphKernelNodeGetParams(h->gnode_gather, &p));
        const double *a0 = d_x;
        const int *a1 = h->d_fwd;
        double *a2 = h->d_xin;
        int a3 = h->N;
        void *args[] = {(void *)&a0, (void *)&a1, (void *)&a2, (void *)&a3};
        p.kernelParams = args;
        p.extra = nullptr;
        HIP_TRY(hipGraphExecKernelNodeSetParams(h->graph_exec, h->gnode_gather, &p));
        h->graph_x = d_x;
    }
    if (d_y != h->graph_y) {
        hipKernelNodeParams p{};
        HIP_TRY(hipGraphKernelNodeGetParams(h->gnode_leaf, &p));
        const LeafTask2 *a0 = h->d_leaf;
        const long long *a1 = h->d_contrib;
        const double *a2 = h->d_arena;
        const double *a3 = h->split ? (const double *)h->d_yt : (const double *)h->d_UT;
        const double *a4 = h->d_yp;
        const int *a5 = h->d_fwd;
        double *a6 = d_y;
        const double *a7 = h->d_xin;
        double a8 = 0.0;
        void *split_args[] = {(void *)&a0, (void *)&a1, (void *)&a2, (void *)&a3, (void *)&a5, (void *)&a6,
                              (void *)&a7, (void *)&a8};
        void *final_args[] = {(void *)&a0, (void *)&a1, (void *)&a2, (void *)&a3, (void *)&a4, (void *)&a5,
                              (void *)&a6, (void *)&a7, (void *)&a8};
        p.kernelParams = h->split ? split_args : final_args;
        p.extra = nullptr;
        HIP_TRY(hipGraphExecKernelNodeSetParams(h->graph_exec, h->gnode_leaf, &p));
        h->graph_y = d_y;
    }
    return 0;
}

// Capture gather + interior + leaf outputs once (for x, y) and find the two
// boundary kernel nodes.
static int ensure_graph(h2p_hip_handle *h, const double *d_x, double *d_y) {
    if (h->graph_exec) return graph_set_xy(h, d_x, d_y);
    hipStream_t cs = h->own_stream;
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_gather, dim3((h->N + 255) / 256), dim3(256), 0, cs, d_x, h->d_fwd, h->d_xin, h->N);
    launch_interior(h, cs, nullptr);
    launch_leaf_out(h, d_y, cs);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &g);
    if (e != hipSuccess) return fail(H2P_HIP_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    size_t n = 0;
    HIP_TRY(hipGraphGetNodes(g, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    HIP_TRY(hipGraphGetNodes(g, nodes.data(), &n));
    const void *f_gather = reinterpret_cast<const void *>(&k_gather);
    const void *f_leaf = h->split ? reinterpret_cast<const void *>(&k_leaf_d)
                                  : reinterpret_cast<const void *>(&k_leaf_final2);
    hipGraphNode_t ng = nullptr, nl = nullptr;
    int n_gather = 0, n_leaf = 0;
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType ty;
        HIP_TRY(hipGraphNodeGetType(nd, &ty));
        if (ty != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams p{};
        HIP_TRY(hipGraphKernelNodeGetParams(nd, &p));
        if (p.func == f_gather) { ng = nd; n_gather++; }
        if (p.func == f_leaf) { nl = nd; n_leaf++; }
    }
    if (n_gather != 1 || n_leaf != 1) {
        (void)hipGraphDestroy(g);
        return fail(H2P_HIP_EDEVICE, "captured matvec graph: boundary kernel nodes not found");
    }
    h->graph = g;
    h->gnode_gather = ng;
    h->gnode_leaf = nl;
    HIP_TRY(hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
    h->graph_x = d_x;
    h->graph_y = d_y;
    return 0;
}

// Record the end of a call's work on stream s, and wait for the handle's
// last recorded call (matvec, matmul or batched), if any.  Calls on one
// handle do not overlap, so the last one covers the earlier ones.
static void mark_use(h2p_hip_handle *h, hipStream_t s) {
    if (hipEventRecord(h->ev_replay, s) == hipSuccess) h->replay_pending = true;
}

static void wait_replays(h2p_hip_handle *h) {
    if (h->replay_pending) (void)hipEventSynchronize(h->ev_replay);
    h->replay_pending = false;
}

static void drop_matvec_graph(h2p_hip_handle *h) {
    if (!h->graph_exec) return;
    wait_replays(h);
    (void)hipGraphExecDestroy(h->graph_exec);
    (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->gnode_gather = h->gnode_leaf = nullptr;
    h->graph_x = nullptr;
    h->graph_y = nullptr;
}

// One matvec on device pointers (x, y in the caller's order).
static int run_matvec(h2p_hip_handle *h, const double *d_x, double *d_y, hipStream_t s) {
    const bool prof = h->profiling && h->ev_used < h->ev_cap;
    if (!prof && h->use_graph) {
        int rc = ensure_graph(h, d_x, d_y);
        if (rc) return rc;
        HIP_TRY(hipGraphLaunch(h->graph_exec, s));
        mark_use(h, s);
        return 0;
    }
    hipEvent_t *ev = prof ? &h->ev_pool[(size_t)h->ev_used * H2P_EV_PER_CALL] : nullptr;
    if (prof) HIP_TRY(hipEventRecord(ev[0], s));
    hipLaunchKernelGGL(k_gather, dim3((h->N + 255) / 256), dim3(256), 0, s, d_x, h->d_fwd, h->d_xin, h->N);
    launch_interior(h, s, ev);
    launch_leaf_out(h, d_y, s);
    HIP_TRY(hipGetLastError());
    if (prof) {
        HIP_TRY(hipEventRecord(ev[6], s));
        h->ev_used++;
    }
    mark_use(h, s);
    return 0;
}

// matmul launches: kernel dimension and id, panel width class (NV = 16, 32
// or 64 columns) and tile variant (chosen by h2p_hip_autotune)
static int nv_class(int nvp) { return nvp <= 16 ? 0 : nvp <= 32 ? 1 : 2; }

template <int KID, int KD, int NV, int V>
static void launch_mm_v(const h2p_hip_handle *h, const Level &L, const MMPanels &P, hipStream_t s) {
    // dense B/D sources: the stored blocks, or in periodic mode the root's block
    const double *bd = h->periodic ? h->d_per : (const double *)h->d_blob;
    hipLaunchKernelGGL((k_mm<KID, KD, MMVar<KD, NV, V>>), dim3(L.n), dim3(256), 0, s, h->d_mmt + L.off, h->d_mms,
                       h->d_U, h->d_UT, bd, h->d_coord, h->CT, h->kp, h->d_shift, P);
}

template <int KID, int KD, int NV>
static void launch_mm_nv(const h2p_hip_handle *h, const Level &L, const MMPanels &P, hipStream_t s, int v) {
    switch (v) {
        case 1: launch_mm_v<KID, KD, NV, 1>(h, L, P, s); break;
        case 2: launch_mm_v<KID, KD, NV, 2>(h, L, P, s); break;
        case 3: launch_mm_v<KID, KD, NV, 3>(h, L, P, s); break;
        default: launch_mm_v<KID, KD, NV, 0>(h, L, P, s); break;
    }
}

template <int KID, int KD>
static void launch_mm(const h2p_hip_handle *h, const Level &L, const MMPanels &P, hipStream_t s, int v) {
    switch (nv_class(P.nvp)) {
        case 0: launch_mm_nv<KID, KD, 16>(h, L, P, s, v); break;
        case 1: launch_mm_nv<KID, KD, 32>(h, L, P, s, v); break;
        default: launch_mm_nv<KID, KD, 64>(h, L, P, s, v); break;
    }
}

static void dispatch_mm(const h2p_hip_handle *h, const Level &L, const MMPanels &P, hipStream_t s, int v) {
    if (L.n <= 0) return;
    switch (h->kid) {
        case KC: launch_mm<KC, 1>(h, L, P, s, v); break;
        case KG: launch_mm<KG, 1>(h, L, P, s, v); break;
        case KM: launch_mm<KM, 1>(h, L, P, s, v); break;
        default: launch_mm<KR, 3>(h, L, P, s, v); break;
    }
}

template <int KID, int NV, int V>
static void launch_mm32_v(const h2p_hip_handle *h, const Level &L, const MMPanels &P, hipStream_t s) {
    hipLaunchKernelGGL((k_mm32<KID, MMVar<1, NV, V>>), dim3(L.n), dim3(256), 0, s, h->d_mmt + L.off, h->d_mms,
                       (const float *)h->d_blob, h->d_coord, h->CT, (float)h->kp.p0, h->d_shift, P);
}

template <int KID, int NV>
static void launch_mm32_nv(const h2p_hip_handle *h, const Level &L, const MMPanels &P, hipStream_t s, int v) {
    switch (v) {
        case 1: launch_mm32_v<KID, NV, 1>(h, L, P, s); break;
        case 2: launch_mm32_v<KID, NV, 2>(h, L, P, s); break;
        case 3: launch_mm32_v<KID, NV, 3>(h, L, P, s); break;
        default: launch_mm32_v<KID, NV, 0>(h, L, P, s); break;
    }
}

template <int KID>
static void launch_mm32(const h2p_hip_handle *h, const Level &L, const MMPanels &P, hipStream_t s, int v) {
    switch (nv_class(P.nvp)) {
        case 0: launch_mm32_nv<KID, 16>(h, L, P, s, v); break;
        case 1: launch_mm32_nv<KID, 32>(h, L, P, s, v); break;
        default: launch_mm32_nv<KID, 64>(h, L, P, s, v); break;
    }
}

static void dispatch_mm32(const h2p_hip_handle *h, const Level &L, const MMPanels &P, hipStream_t s, int v) {
    if (L.n <= 0) return;
    switch (h->kid) {
        case KC: launch_mm32<KC>(h, L, P, s, v); break;
        case KG: launch_mm32<KG>(h, L, P, s, v); break;
        default: launch_mm32<KM>(h, L, P, s, v); break;
    }
}

// the four phase classes of a matmul, as level lists for the FP64 product
// kernel and for the FP32 one (FP32-mixed handles)
enum { MMP_FWD = 0, MMP_INTMD = 1, MMP_BWD = 2, MMP_LEAF = 3 };
static std::vector<Level> mm_phase_levels(const h2p_hip_handle *h, int ph) {
    switch (ph) {
        case MMP_FWD: return h->mm_fwd_levels;
        case MMP_INTMD: return h->mm_f32 ? std::vector<Level>{} : std::vector<Level>{h->mm_intmd};
        case MMP_BWD: return h->mm_bwd_levels;
        default: return {h->mm_leaf};
    }
}

static std::vector<Level> mm_phase_levels32(const h2p_hip_handle *h, int ph) {
    if (!h->mm_f32) return {};
    if (ph == MMP_INTMD) return {h->mm_intmd};
    if (ph == MMP_LEAF) return {h->mm_leaf32};
    return {};
}

static void launch_mm_phase(const h2p_hip_handle *h, int ph, const MMPanels &P, hipStream_t s, int v) {
    for (const Level &L : mm_phase_levels(h, ph)) dispatch_mm(h, L, P, s, v);
    for (const Level &L : mm_phase_levels32(h, ph)) dispatch_mm32(h, L, P, s, v);
}

static void drop_mm_graphs(h2p_hip_handle *h) {
    if (h->mm_graphs.empty()) return;
    wait_replays(h);   // a replay may still be running
    for (auto &m : h->mm_graphs) {
        (void)hipGraphExecDestroy(m.e);
        (void)hipGraphDestroy(m.g);
    }
    h->mm_graphs.clear();
}

// all product phases of one matmul chunk with panels P, captured once per
// panel width and replayed
static int mm_graph(h2p_hip_handle *h, const MMPanels &P, hipGraphExec_t *out) {
    for (auto &m : h->mm_graphs)
        if (m.nvp == P.nvp) {
            *out = m.e;
            return 0;
        }
    hipStream_t cs = h->own_stream;
    const int c = nv_class(P.nvp);
    HIP_TRY(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int ph = 0; ph < 4; ph++) launch_mm_phase(h, ph, P, cs, h->mm_var[ph][c]);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &g);
    if (e != hipSuccess) return fail(H2P_HIP_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(H2P_HIP_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    h->mm_graphs.push_back({P.nvp, g, ge});
    *out = ge;
    return 0;
}

extern "C" {

const char *h2p_hip_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// Upload, or with dry != NULL only the validation and schedule building,
// returning in *dry the device bytes the upload would hold.
static int upload_impl(const h2p_hip_h2pack *g, const h2p_kernel_desc *krnl, int mode, int prec,
                       h2p_hip_handle **out, size_t *dry) {
    if (!out) return fail(H2P_HIP_EINVAL, "out is NULL");
    *out = nullptr;
    if (!g || !krnl) return fail(H2P_HIP_EINVAL, "NULL representation or kernel");
    if (mode != H2P_HIP_AOT && mode != H2P_HIP_JIT) return fail(H2P_HIP_EINVAL, "bad mode");
    if (prec != H2P_HIP_FP64 && prec != H2P_HIP_FP32MIX) return fail(H2P_HIP_EINVAL, "bad precision");
    if (krnl->id < 0 || krnl->id > 3) return fail(H2P_HIP_EINVAL, "bad kernel id");
    const int kd = g->krnl_dim;
    if (g->pt_dim != 3) return fail(H2P_HIP_EINVAL, "pt_dim must be 3");
    if ((krnl->id == KR) != (kd == 3) || (kd != 1 && kd != 3))
        return fail(H2P_HIP_EINVAL, "krnl_dim does not match the kernel");
    if (krnl->id == KR && g->xpt_dim != 4) return fail(H2P_HIP_EINVAL, "RPY needs xpt_dim 4 (x, y, z, radius)");
    if (g->n_point <= 0 || g->krnl_mat_size != kd * g->n_point || g->n_node <= 0)
        return fail(H2P_HIP_EINVAL, "inconsistent sizes");
    if ((long long)g->n_point * kd > 0x7fffffffLL) return fail(H2P_HIP_EINVAL, "too many points");
    if (g->periodic) {
        if (mode != H2P_HIP_JIT || krnl->id != KR || prec != H2P_HIP_FP64)
            return fail(H2P_HIP_ENOSYS, "the periodic mode is FP64 JIT with the RPY kernel (as H2P_build_periodic)");
        if (!g->per_blk || (g->n_r_adm_pair > 0 && !g->per_adm_shifts) || (g->n_r_inadm_pair > 0 && !g->per_inadm_shifts))
            return fail(H2P_HIP_EINVAL, "periodic representation without shifts or periodic block");
    }

    h2p_hip_handle *h = new h2p_hip_handle();
    h->mode = mode;
    h->prec = prec;
    h->kid = krnl->id;
    h->kd = kd;
    h->xpt_dim = g->xpt_dim;
    h->N = g->krnl_mat_size;
    h->n_point = g->n_point;
    if (krnl->id == KC) h->kp.p0 = krnl->n_param > 0 ? krnl->param[0] : 0.0;
    else h->kp.p0 = krnl->param[0];
    if (krnl->id == KR) h->kp.p1 = 1.0 / (6.0 * M_PI * krnl->param[0]);

    auto bail = [&](int rc) { free_handle(h); return rc; };

    const int nn = g->n_node, nlf = g->n_leaf_node, mc = g->max_child;
    const int N = h->N;
    // ---- tree helpers
    std::vector<int> parent(nn, -1);
    for (int i = 0; i < nn; i++)
        for (int k = 0; k < g->n_child[i]; k++) {
            int c = g->children[(size_t)i * mc + k];
            if (c < 0 || c >= nn) return bail(fail(H2P_HIP_EINVAL, "bad child index"));
            parent[c] = i;
        }
    int root = -1;
    for (int i = 0; i < nn; i++)
        if (parent[i] < 0) {
            if (root >= 0) return bail(fail(H2P_HIP_EINVAL, "more than one root"));
            root = i;
        }
    if (root < 0) return bail(fail(H2P_HIP_EINVAL, "no root"));
    auto npt = [&](int n) { return g->pt_cluster[2 * n + 1] - g->pt_cluster[2 * n] + 1; };
    for (int i = 0; i < nn; i++) {
        if (g->U_ncol[i] < 0 || g->U_nrow[i] < 0) return bail(fail(H2P_HIP_EINVAL, "negative U size"));
        if (g->U_ncol[i] > 0 && !g->U[i]) return bail(fail(H2P_HIP_EINVAL, "NULL U"));
        if (g->U_ncol[i] > 0 && g->U_ncol[i] != g->J_npt[i] * kd)
            return bail(fail(H2P_HIP_EINVAL, "U_ncol != kd * |J|"));
        if (g->mat_cluster[2 * i] != kd * g->pt_cluster[2 * i]) return bail(fail(H2P_HIP_EINVAL, "mat_cluster mismatch"));
    }
    // ---- skeleton offsets: BFS so that children of a node are contiguous
    std::vector<int> soff(nn, 0);
    {
        std::vector<int> q{root};
        int cur = 0;
        soff[root] = cur;
        cur += g->U_ncol[root];
        for (size_t qi = 0; qi < q.size(); qi++) {
            int n = q[qi];
            for (int k = 0; k < g->n_child[n]; k++) {
                int c = g->children[(size_t)n * mc + k];
                soff[c] = cur;
                cur += g->U_ncol[c];
                q.push_back(c);
            }
        }
        if ((int)q.size() != nn) return bail(fail(H2P_HIP_EINVAL, "tree is not connected"));
        h->S = cur;
    }
    const int S = h->S;
    // ---- U and U^T blobs
    std::vector<long long> uoff(nn, 0);
    long long utot = 0;
    for (int i = 0; i < nn; i++) {
        uoff[i] = utot;
        utot += (long long)g->U_nrow[i] * g->U_ncol[i];
    }
    std::vector<double> Ub(utot), UTb(utot);
    for (int i = 0; i < nn; i++) {
        const int nr = g->U_nrow[i], r = g->U_ncol[i];
        if (r == 0) continue;
        const double *src = g->U[i];
        std::memcpy(Ub.data() + uoff[i], src, sizeof(double) * nr * (size_t)r);
        double *dt = UTb.data() + uoff[i];
        for (int a = 0; a < nr; a++)
            for (int c = 0; c < r; c++) dt[(size_t)c * nr + a] = src[(size_t)a * r + c];
    }
    // ---- unified coordinates [points | skeleton points]
    std::vector<int> joff(nn, 0);
    long long jtot = 0;
    for (int i = 0; i < nn; i++) {
        joff[i] = (int)jtot;
        jtot += (g->U_ncol[i] > 0) ? g->J_npt[i] : 0;
    }
    const long long CT = g->n_point + jtot;
    h->CT = CT;
    const int xd = g->xpt_dim;
    // RPY: one more row, the reciprocal radii, so that the kernel-evaluating
    // launches need no FP64 division per pair
    const bool rcp_row = krnl->id == KR;
    std::vector<double> cu((size_t)(xd + (rcp_row ? 1 : 0)) * CT);
    for (int d = 0; d < xd; d++) {
        std::memcpy(cu.data() + (size_t)d * CT, g->coord + (size_t)d * g->n_point, sizeof(double) * g->n_point);
        for (int i = 0; i < nn; i++) {
            if (g->U_ncol[i] == 0) continue;
            const int m = g->J_npt[i];
            std::memcpy(cu.data() + (size_t)d * CT + g->n_point + joff[i], g->J_coord[i] + (size_t)d * m, sizeof(double) * m);
        }
    }
    if (rcp_row)
        for (long long i = 0; i < CT; i++) cu[(size_t)4 * CT + i] = 1.0 / cu[(size_t)3 * CT + i];
    // FP64 Gaussian / Matern JIT: the largest scaled distance bounds the
    // argument of the exponential; past 2^30 the launches clamp it (wide)
    if (mode == H2P_HIP_JIT && prec == H2P_HIP_FP64 && (krnl->id == KG || krnl->id == KM)) {
        double ext = 0.0;
        for (int d = 0; d < 3; d++) {
            const double *c = cu.data() + (size_t)d * CT;
            double lo = c[0], hi = c[0];
            for (long long i = 1; i < CT; i++) {
                lo = std::min(lo, c[i]);
                hi = std::max(hi, c[i]);
            }
            ext = std::max(ext, hi - lo);
        }
        const double rmax = 1.7320508075688772 * ext * kscale(krnl->id, h->kp.p0, false);
        const double arg = krnl->id == KG ? rmax * rmax : rmax * 1.4426950408889634;
        h->wide = !(arg < 0x1p30);   // also for non-finite coordinates or parameters
    }
    // ---- upward sweep tasks per level
    // level split (split handles): the tree level holding the most leaves
    // (the deepest on a tie) and everything below it are "deep"
    {
        std::vector<int> lcount(g->max_level + 1, 0);
        for (int i = 0; i < nlf; i++) {
            const int lv = g->node_level[g->leaf_nodes[i]];
            if (lv >= 0 && lv <= g->max_level) lcount[lv]++;
        }
        h->lv_cut = g->max_level;
        for (int lv = g->max_level; lv >= 0; lv--)
            if (lcount[lv] > lcount[h->lv_cut]) h->lv_cut = lv;
    }
    std::vector<FwdTask> fwdt;
    for (int lv = g->max_level; lv >= g->min_adm_level; lv--) {
        Level L{(int)fwdt.size(), 0};
        // column chunk of the level's tasks: 256 columns, narrower while the
        // level would have fewer than H2P_FWD_MIN_WG workgroups (each
        // workgroup keeps the same loads in flight at any width, so more of
        // them on a small level means more bytes in flight)
        int ncol = 4 * H2P_WAVE;
        while (ncol > H2P_WAVE) {
            long long nwg = 0;
            for (int j = 0; j < g->level_n_node[lv]; j++)
                nwg += (g->U_ncol[g->level_nodes[(size_t)lv * nlf + j]] + ncol - 1) / ncol;
            if (nwg >= H2P_FWD_MIN_WG) break;
            ncol >>= 1;
        }
        for (int j = 0; j < g->level_n_node[lv]; j++) {
            int n = g->level_nodes[(size_t)lv * nlf + j];
            if (g->U_ncol[n] == 0) continue;
            FwdTask t{};
            t.u_off = uoff[n];
            t.nrow = g->U_nrow[n];
            t.r = g->U_ncol[n];
            t.out_off = N + soff[n];
            if (g->n_child[n] == 0) {
                if (t.nrow != npt(n) * kd) return bail(fail(H2P_HIP_EINVAL, "leaf U_nrow != kd * points"));
                t.in_off = g->mat_cluster[2 * n];
            } else {
                int sum = 0;
                for (int k = 0; k < g->n_child[n]; k++) sum += g->U_ncol[g->children[(size_t)n * mc + k]];
                if (sum != t.nrow) return bail(fail(H2P_HIP_EINVAL, "non-leaf U_nrow != sum of child ranks"));
                t.in_off = N + soff[g->children[(size_t)n * mc]];
            }
            t.ncol = ncol;
            for (int c0 = 0; c0 < t.r; c0 += ncol) {
                t.col0 = c0;
                fwdt.push_back(t);
            }
        }
        L.n = (int)fwdt.size() - L.off;
        if (L.n > 0) {
            h->fwd_levels.push_back(L);
            if (lv >= h->lv_cut) h->n_fwd_deep++;
        }
    }
    // ---- grouped block tasks (B then D) and partial destinations
    std::vector<BlkTask> blk;
    std::vector<char> blk_deep;   // B tasks: skeleton side at level >= lv_cut
    std::vector<std::vector<long long>> dest_y1(nn), dest_leaf(nn);
    long long arena = 0;
    size_t btot = (g->n_r_adm_pair > 0) ? g->B_ptr[g->n_r_adm_pair] : 0;
    const int nD = nlf + g->n_r_inadm_pair;
    size_t dtot = g->D_ptr[nD];
    double B_pairs = 0, D_pairs = 0;
    const bool per = g->periodic != 0;
    h->periodic = per ? 1 : 0;
    // pair and leaf indices in range, and node_n_r_adm as H2P_build counts it
    // from the pair list (src/H2Pack_build.c:707-717: both nodes of a pair,
    // only node 0 in periodic mode), since it marks the nodes whose skeleton
    // sums the downward sweep reads
    {
        auto node_ok = [&](int v) { return v >= 0 && v < nn; };
        std::vector<int> cnt(nn, 0);
        for (int i = 0; i < g->n_r_adm_pair; i++) {
            const int n0 = g->r_adm_pairs[2 * i], n1 = g->r_adm_pairs[2 * i + 1];
            if (!node_ok(n0) || !node_ok(n1)) return bail(fail(H2P_HIP_EINVAL, "admissible pair index out of range"));
            cnt[n0]++;
            if (!per) cnt[n1]++;
        }
        for (int i = 0; i < 2 * g->n_r_inadm_pair; i++)
            if (!node_ok(g->r_inadm_pairs[i])) return bail(fail(H2P_HIP_EINVAL, "inadmissible pair index out of range"));
        for (int i = 0; i < nlf; i++)
            if (!node_ok(g->leaf_nodes[i])) return bail(fail(H2P_HIP_EINVAL, "leaf index out of range"));
        for (int n = 0; n < nn; n++)
            if (g->node_n_r_adm[n] != cnt[n])
                return bail(fail(H2P_HIP_EINVAL, "node_n_r_adm does not match r_adm_pairs"));
    }
    // shift table: entry 0 is the zero shift, then one entry per distinct offset
    std::vector<double> shtab{0.0, 0.0, 0.0};
    auto shift_index = [&](const double *v) {
        if (v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0) return 0;
        for (size_t k = 1; k < shtab.size() / 3; k++)
            if (shtab[3 * k] == v[0] && shtab[3 * k + 1] == v[1] && shtab[3 * k + 2] == v[2]) return (int)k;
        shtab.insert(shtab.end(), v, v + 3);
        return (int)(shtab.size() / 3 - 1);
    };
    std::vector<int> adm_shift(g->n_r_adm_pair, 0), inadm_shift(g->n_r_inadm_pair, 0);
    if (per) {
        for (int i = 0; i < g->n_r_adm_pair; i++) adm_shift[i] = shift_index(g->per_adm_shifts + 3 * (size_t)i);
        for (int i = 0; i < g->n_r_inadm_pair; i++) inadm_shift[i] = shift_index(g->per_inadm_shifts + 3 * (size_t)i);
    }
    for (int i = 0; i < g->n_r_adm_pair; i++) {
        const int n0 = g->r_adm_pairs[2 * i], n1 = g->r_adm_pairs[2 * i + 1];
        const int l0 = g->node_level[n0], l1 = g->node_level[n1];
        BlkTask t{};
        t.shift = adm_shift[i];
        t.nrow = g->B_nrow[i];
        t.ncol = g->B_ncol[i];
        if (l0 < l1) { t.c0 = g->pt_cluster[2 * n0]; t.n0 = npt(n0); t.in1 = g->mat_cluster[2 * n0]; }
        else { t.c0 = g->n_point + joff[n0]; t.n0 = g->J_npt[n0]; t.in1 = N + soff[n0]; }
        if (l0 > l1) { t.c1 = g->pt_cluster[2 * n1]; t.n1 = npt(n1); t.in0 = g->mat_cluster[2 * n1]; }
        else { t.c1 = g->n_point + joff[n1]; t.n1 = g->J_npt[n1]; t.in0 = N + soff[n1]; }
        if (t.nrow != t.n0 * kd || t.ncol != t.n1 * kd)
            return bail(fail(H2P_HIP_EINVAL, "B block size does not match its point sets"));
        if ((l0 >= l1 && g->U_ncol[n0] == 0) || (l0 <= l1 && g->U_ncol[n1] == 0)) {
            if (t.nrow * (long long)t.ncol != 0) return bail(fail(H2P_HIP_EINVAL, "B block on a node without U"));
        }
        if (t.nrow == 0 || t.ncol == 0) continue;
        t.data_off = (long long)g->B_ptr[i];
        t.out0 = arena; arena += t.nrow;
        (l0 < l1 ? dest_leaf[n0] : dest_y1[n0]).push_back(t.out0);
        if (per) {   // one-sided: node 0 receives only
            t.flags = 0;
            t.out1 = -1;
        } else {
            t.flags = 1;
            t.out1 = arena; arena += t.ncol;
            (l0 > l1 ? dest_leaf[n1] : dest_y1[n1]).push_back(t.out1);
        }
        B_pairs += (double)t.n0 * t.n1;
        blk.push_back(t);
        blk_deep.push_back(std::max(l0, l1) >= h->lv_cut ? 1 : 0);
    }
    const int nB = (int)blk.size();
    // the deep B blocks first, the shallow ones after them (stable)
    int nB_deep = nB;
    {
        std::vector<BlkTask> deep, shallow;
        for (int i = 0; i < nB; i++) (blk_deep[i] ? deep : shallow).push_back(blk[i]);
        nB_deep = (int)deep.size();
        std::copy(deep.begin(), deep.end(), blk.begin());
        std::copy(shallow.begin(), shallow.end(), blk.begin() + nB_deep);
    }
    for (int i = 0; i < nD; i++) {
        BlkTask t{};
        int n0, n1;
        if (i < nlf) n0 = n1 = g->leaf_nodes[i];
        else { n0 = g->r_inadm_pairs[2 * (i - nlf)]; n1 = g->r_inadm_pairs[2 * (i - nlf) + 1]; }
        if (g->n_child[n0] != 0 || g->n_child[n1] != 0) return bail(fail(H2P_HIP_EINVAL, "D block on a non-leaf"));
        t.nrow = g->D_nrow[i];
        t.ncol = g->D_ncol[i];
        t.c0 = g->pt_cluster[2 * n0]; t.n0 = npt(n0);
        t.c1 = g->pt_cluster[2 * n1]; t.n1 = npt(n1);
        if (t.nrow != t.n0 * kd || t.ncol != t.n1 * kd) return bail(fail(H2P_HIP_EINVAL, "D block size mismatch"));
        t.in0 = g->mat_cluster[2 * n1];
        t.in1 = g->mat_cluster[2 * n0];
        t.data_off = (long long)(btot + g->D_ptr[i]);
        t.out0 = arena; arena += t.nrow;
        dest_leaf[n0].push_back(t.out0);
        if (i >= nlf) t.shift = inadm_shift[i - nlf];
        if (i >= nlf && !per) {
            t.flags = 1;
            t.out1 = arena; arena += t.ncol;
            dest_leaf[n1].push_back(t.out1);
        } else {
            t.flags = 0;
            t.out1 = -1;
        }
        D_pairs += (double)t.n0 * t.n1;
        if (t.nrow == 0 || t.ncol == 0) continue;
        blk.push_back(t);
    }
    h->arena_len = arena;
    h->n_blk = (int)blk.size();
    // JIT handles of the scalar kernels launch the B blocks and the D blocks
    // separately (the D blocks read only the permuted input, so they overlap
    // the sweeps); AOT and RPY handles keep one grouped launch.  Each part in
    // longest-processing-time order; the tiled kernels (FP64 Gaussian and
    // Matern, FP32-mixed Coulomb, Gaussian and Matern) then run the small
    // blocks of a part as wave units of four, after its workgroup-unit blocks.
    h->split = mode == H2P_HIP_JIT && h->kid != KR && !per;
    const bool tiled = mode == H2P_HIP_JIT && h->kid != KR &&
                       (prec == H2P_HIP_FP32MIX || h->kid == KG || h->kid == KM);
    for (auto &P : h->part) P = h2p_hip_handle::BlkPart{};
    h->part[0].n = h->split ? nB_deep : h->n_blk;
    h->part[3].off = h->part[0].n;
    h->part[3].n = h->split ? nB - nB_deep : 0;
    h->part[1].off = h->split ? nB : h->n_blk;
    h->part[1].n = h->n_blk - h->part[1].off;
    for (int k : {0, 3, 1}) {
        auto &P = h->part[k];
        auto b0 = blk.begin() + P.off, b1 = b0 + P.n;
        std::stable_sort(b0, b1, [](const BlkTask &x, const BlkTask &y) {
            return (long long)x.nrow * x.ncol > (long long)y.nrow * y.ncol;
        });
        P.big = P.n;
        if (tiled)
            P.big = (int)(std::stable_partition(b0, b1, [](const BlkTask &t) { return !jitm_wave_task(t.n0, t.n1); }) -
                          b0);
        P.units = P.big + (P.n - P.big + 3) / 4;
    }
    // the D blocks' last tasks (the smallest, about d_late of their pairs)
    // form part 2, launched after the B blocks
    if (h->split && h->part[1].n > 0) {
        double d_late = H2P_D_LATE;
        if (const char *e = getenv("H2P_HIP_D_LATE")) d_late = atof(e);
        const h2p_hip_handle::BlkPart D = h->part[1];
        double tot = 0.0, acc = 0.0;
        for (int i = 0; i < D.n; i++) tot += (double)blk[D.off + i].nrow * blk[D.off + i].ncol;
        int cut = D.n;
        while (cut > 0 && acc + (double)blk[D.off + cut - 1].nrow * blk[D.off + cut - 1].ncol <= d_late * tot) {
            cut--;
            acc += (double)blk[D.off + cut].nrow * blk[D.off + cut].ncol;
        }
        auto &P1 = h->part[1], &P2 = h->part[2];
        P1.n = cut;
        P1.big = std::min(D.big, cut);
        P1.units = P1.big + (P1.n - P1.big + 3) / 4;
        P2.off = D.off + cut;
        P2.n = D.n - cut;
        P2.big = std::max(0, D.big - cut);
        P2.units = P2.big + (P2.n - P2.big + 3) / 4;
    }
    // ---- active nodes of the downward sweep (y1 "set", src/H2Pack_matvec.c:251-256, :802-838)
    std::vector<char> active(nn, 0), has_push(nn, 0);
    for (int lv = g->min_adm_level; lv <= g->max_level; lv++)
        for (int j = 0; j < g->level_n_node[lv]; j++) {
            int n = g->level_nodes[(size_t)lv * nlf + j];
            const int p = parent[n];
            has_push[n] = (p >= 0 && g->node_level[p] >= g->min_adm_level && active[p] && g->n_child[p] > 0) ? 1 : 0;
            // periodic: the root's y1 is its periodic block product
            const bool per_root = per && n == root;
            active[n] = ((g->node_n_r_adm[n] > 0 || has_push[n] || per_root) && g->U_ncol[n] > 0) ? 1 : 0;
            if (!active[n]) has_push[n] = 0;
        }
    // ---- non-leaf skeleton sums (leaves sum their own in the output kernel)
    // (the nodes above lv_cut first: on a split handle their sums belong to
    // the shallow levels' chain)
    std::vector<long long> contrib;
    std::vector<RedTask> red;
    for (int pass = 0; pass < 2; pass++)
        for (int n = 0; n < nn; n++) {
            if (g->U_ncol[n] == 0 || g->n_child[n] == 0) continue;
            if ((g->node_level[n] < h->lv_cut) != (pass == 0)) continue;
            RedTask t{};
            t.dst = soff[n];
            t.len = g->U_ncol[n];
            t.cbeg = (int)contrib.size();
            for (long long o : dest_y1[n]) contrib.push_back(o);
            t.cend = (int)contrib.size();
            red.push_back(t);
            if (pass == 0) h->n_red_shallow++;
        }
    h->n_red = (int)red.size();
    // ---- downward sweep: one task per 256-row chunk, level by level
    std::vector<BwdTask2> bwdt;
    for (int lv = g->min_adm_level; lv <= g->max_level; lv++) {
        Level L{(int)bwdt.size(), 0};
        for (int j = 0; j < g->level_n_node[lv]; j++) {
            int n = g->level_nodes[(size_t)lv * nlf + j];
            if (!active[n] || g->n_child[n] == 0) continue;
            if (g->U_ncol[n] > H2P_MAXR) return bail(fail(H2P_HIP_ENOSYS, "rank above 2048"));
            for (int r0 = 0; r0 < g->U_nrow[n]; r0 += 256) {
                BwdTask2 t{};
                t.ut_off = uoff[n];
                t.nrow = g->U_nrow[n];
                t.r = g->U_ncol[n];
                t.y1_off = soff[n];
                t.push = has_push[n];
                t.out_off = soff[g->children[(size_t)n * mc]];
                t.row0 = r0;
                bwdt.push_back(t);
            }
        }
        L.n = (int)bwdt.size() - L.off;
        if (L.n > 0) {
            h->bwd_levels.push_back(L);
            if (lv < h->lv_cut) h->n_bwd_shallow++;
        }
    }
    std::vector<LeafTask2> leaf;
    for (int i = 0; i < nlf; i++) {
        int n = g->leaf_nodes[i];
        LeafTask2 t{};
        t.nrow = npt(n) * kd;
        t.r = active[n] ? g->U_ncol[n] : 0;
        if (t.r > H2P_MAXR) return bail(fail(H2P_HIP_ENOSYS, "rank above 2048"));
        t.push = has_push[n];
        t.ut_off = uoff[n];
        t.s_off = soff[n];
        t.vec_s = g->mat_cluster[2 * n];
        t.cb1 = (int)contrib.size();
        if (t.r > 0)
            for (long long o : dest_y1[n]) contrib.push_back(o);
        t.ce1 = (int)contrib.size();
        t.cb0 = (int)contrib.size();
        for (long long o : dest_leaf[n]) contrib.push_back(o);
        t.ce0 = (int)contrib.size();
        leaf.push_back(t);
    }
    h->n_leaf = (int)leaf.size();
    // ---- matmul tasks (pull form, one task per destination)
    std::vector<MMTask> mmt;
    std::vector<MMSrc> mms;
    const bool jit = mode == H2P_HIP_JIT;
    const bool mm_f32 = prec == H2P_HIP_FP32MIX && krnl->id != KR && !per;
    h->mm_f32 = mm_f32;
    {
        // per-destination source lists, in pair order
        std::vector<std::vector<MMSrc>> y1_src(nn), leaf_src(nn);
        auto set_start = [&](int n, bool pts) { return pts ? g->pt_cluster[2 * n] : g->n_point + joff[n]; };
        for (int i = 0; i < g->n_r_adm_pair; i++) {
            const int n0 = g->r_adm_pairs[2 * i], n1 = g->r_adm_pairs[2 * i + 1];
            const int l0 = g->node_level[n0], l1 = g->node_level[n1];
            const int br = g->B_nrow[i], bc = g->B_ncol[i];
            if (br == 0 || bc == 0) continue;
            const bool p0 = l0 < l1, p1 = l0 > l1;      // side given by points
            MMSrc a{}, b{};
            a.kind = b.kind = jit ? MM_A_JIT : MM_A_BD;
            a.a_off = b.a_off = (long long)g->B_ptr[i];
            a.lda = b.lda = bc;
            // entry of n0: op = B
            a.trans = 0; a.K = bc;
            a.c0 = set_start(n0, p0); a.c1 = set_start(n1, p1);
            a.x_space = p1 ? MM_PX : MM_Y0; a.x_off = p1 ? g->mat_cluster[2 * n1] : soff[n1];
            a.shift = adm_shift[i];
            (p0 ? leaf_src[n0] : y1_src[n0]).push_back(a);
            if (per) continue;   // one-sided pairs
            // entry of n1: op = B^T
            b.trans = 1; b.K = br;
            b.c0 = set_start(n1, p1); b.c1 = set_start(n0, p0);
            b.x_space = p0 ? MM_PX : MM_Y0; b.x_off = p0 ? g->mat_cluster[2 * n0] : soff[n0];
            (p1 ? leaf_src[n1] : y1_src[n1]).push_back(b);
        }
        for (int i = 0; i < nD; i++) {
            int n0, n1;
            if (i < nlf) n0 = n1 = g->leaf_nodes[i];
            else { n0 = g->r_inadm_pairs[2 * (i - nlf)]; n1 = g->r_inadm_pairs[2 * (i - nlf) + 1]; }
            const int dr = g->D_nrow[i], dc = g->D_ncol[i];
            if (dr == 0 || dc == 0) continue;
            MMSrc a{};
            a.kind = jit ? MM_A_JIT : MM_A_BD;
            a.a_off = (long long)(btot + g->D_ptr[i]);
            a.lda = dc;
            a.trans = 0; a.K = dc;
            a.c0 = g->pt_cluster[2 * n0]; a.c1 = g->pt_cluster[2 * n1];
            a.x_space = MM_PX; a.x_off = g->mat_cluster[2 * n1];
            if (i >= nlf) a.shift = inadm_shift[i - nlf];
            leaf_src[n0].push_back(a);
            if (i >= nlf && !per) {
                MMSrc b = a;
                b.trans = 1; b.K = dr;
                b.c0 = g->pt_cluster[2 * n1]; b.c1 = g->pt_cluster[2 * n0];
                b.x_off = g->mat_cluster[2 * n0];
                leaf_src[n1].push_back(b);
            }
        }
        // a destination of M rows becomes tasks of at most MM_PIECE rows (a
        // multiple of every variant's row chunk, and of 3 for whole points),
        // each with its sources restricted to its rows: more workgroups for
        // the few large destinations of the upper levels, the same MFMA
        // sequence per output element, one writer per row
        const int MM_PIECE = kd == 3 ? 96 : 128;
        auto add_task = [&](int space, int off, int M, int accum, const std::vector<MMSrc> &src) {
            for (int m0 = 0; m0 < M || (M == 0 && m0 == 0); m0 += MM_PIECE) {
                MMTask t{};
                t.dst_space = space; t.dst_off = off + m0; t.M = std::min(MM_PIECE, M - m0); t.accumulate = accum;
                t.sbeg = (int)mms.size();
                for (MMSrc x : src) {
                    if (x.kind == MM_A_JIT) x.c0 += m0 / kd;
                    else x.a_off += x.trans ? (long long)m0 : (long long)m0 * x.lda;
                    mms.push_back(x);
                }
                t.send = (int)mms.size();
                mmt.push_back(t);
                if (M == 0) break;
            }
        };
        auto work = [&](const MMTask &t) {
            long long wsum = 0;
            for (int k = t.sbeg; k < t.send; k++) wsum += (long long)t.M * mms[k].K;
            return wsum;
        };
        // upward sweep, level by level
        for (int lv = g->max_level; lv >= g->min_adm_level; lv--) {
            Level L{(int)mmt.size(), 0};
            for (int j = 0; j < g->level_n_node[lv]; j++) {
                int n = g->level_nodes[(size_t)lv * nlf + j];
                if (g->U_ncol[n] == 0) continue;
                MMSrc a{};
                a.kind = MM_A_UT; a.a_off = uoff[n]; a.lda = g->U_nrow[n]; a.trans = 0; a.K = g->U_nrow[n];
                if (g->n_child[n] == 0) { a.x_space = MM_PX; a.x_off = g->mat_cluster[2 * n]; }
                else { a.x_space = MM_Y0; a.x_off = soff[g->children[(size_t)n * mc]]; }
                add_task(MM_Y0, soff[n], g->U_ncol[n], 0, std::vector<MMSrc>{a});
            }
            L.n = (int)mmt.size() - L.off;
            if (L.n > 0) h->mm_fwd_levels.push_back(L);
        }
        // periodic: the root's block, a dense source on its own skeleton
        // coefficients (its only coupling: the root has no admissible pair)
        if (per) {
            MMSrc r{};
            r.kind = MM_A_BD; r.a_off = 0; r.lda = g->U_ncol[root]; r.trans = 0; r.K = g->U_ncol[root];
            r.x_space = MM_Y0; r.x_off = soff[root];
            y1_src[root].push_back(r);
        }
        // coupling into Y1 (every node with a basis, so Y1 is always written)
        {
            Level L{(int)mmt.size(), 0};
            for (int n = 0; n < nn; n++)
                if (g->U_ncol[n] > 0) add_task(MM_Y1, soff[n], g->U_ncol[n], 0, y1_src[n]);
            L.n = (int)mmt.size() - L.off;
            std::stable_sort(mmt.begin() + L.off, mmt.end(), [&](const MMTask &x, const MMTask &y) { return work(x) > work(y); });
            h->mm_intmd = L;
        }
        // downward sweep, level by level
        for (int lv = g->min_adm_level; lv <= g->max_level; lv++) {
            Level L{(int)mmt.size(), 0};
            for (int j = 0; j < g->level_n_node[lv]; j++) {
                int n = g->level_nodes[(size_t)lv * nlf + j];
                if (!active[n] || g->n_child[n] == 0) continue;
                MMSrc a{};
                a.kind = MM_A_U; a.a_off = uoff[n]; a.lda = g->U_ncol[n]; a.trans = 0; a.K = g->U_ncol[n];
                a.x_space = MM_Y1; a.x_off = soff[n];
                add_task(MM_Y1, soff[g->children[(size_t)n * mc]], g->U_nrow[n], 1, std::vector<MMSrc>{a});
            }
            L.n = (int)mmt.size() - L.off;
            if (L.n > 0) h->mm_bwd_levels.push_back(L);
        }
        // leaf outputs; FP32-mixed handles split them into the U part (FP64
        // kernel, first writer) and the B/D part (FP32 kernel, accumulating)
        {
            Level L{(int)mmt.size(), 0};
            for (int i = 0; i < nlf; i++) {
                int n = g->leaf_nodes[i];
                std::vector<MMSrc> src;
                if (active[n]) {
                    MMSrc a{};
                    a.kind = MM_A_U; a.a_off = uoff[n]; a.lda = g->U_ncol[n]; a.trans = 0; a.K = g->U_ncol[n];
                    a.x_space = MM_Y1; a.x_off = soff[n];
                    src.push_back(a);
                }
                if (!mm_f32)
                    for (const MMSrc &x : leaf_src[n]) src.push_back(x);
                add_task(MM_PY, g->mat_cluster[2 * n], npt(n) * kd, 0, src);
            }
            L.n = (int)mmt.size() - L.off;
            std::stable_sort(mmt.begin() + L.off, mmt.end(), [&](const MMTask &x, const MMTask &y) { return work(x) > work(y); });
            h->mm_leaf = L;
            if (mm_f32) {
                Level L32{(int)mmt.size(), 0};
                for (int i = 0; i < nlf; i++) {
                    int n = g->leaf_nodes[i];
                    if (!leaf_src[n].empty()) add_task(MM_PY, g->mat_cluster[2 * n], npt(n) * kd, 1, leaf_src[n]);
                }
                L32.n = (int)mmt.size() - L32.off;
                std::stable_sort(mmt.begin() + L32.off, mmt.end(), [&](const MMTask &x, const MMTask &y) { return work(x) > work(y); });
                h->mm_leaf32 = L32;
            }
        }
    }
    // ---- every access of every task stays inside its buffer
    {
        const long long XIN = (long long)N + S, Y1 = S, AR = arena, NC = (long long)contrib.size();
        const long long BLOB = (long long)(btot + dtot);
        if (XIN > 0x7fffffffLL) return bail(fail(H2P_HIP_EINVAL, "vector space exceeds 2^31 entries"));
        auto in = [](long long off, long long len, long long cap) { return off >= 0 && len >= 0 && off + len <= cap; };
        for (const FwdTask &t : fwdt)
            if (!in(t.in_off, t.nrow, XIN) || !in(t.out_off, t.r, XIN) || !in(t.u_off, (long long)t.nrow * t.r, utot) ||
                t.col0 < 0 || t.col0 >= t.r || (t.ncol != 64 && t.ncol != 128 && t.ncol != 256))
                return bail(fail(H2P_HIP_EINVAL, "upward task out of range"));
        for (const BwdTask2 &t : bwdt)
            if (!in(t.y1_off, t.r, Y1) || !in(t.out_off, t.nrow, Y1) || !in(t.ut_off, (long long)t.nrow * t.r, utot) ||
                t.row0 < 0 || t.row0 >= t.nrow || t.r > H2P_MAXR)
                return bail(fail(H2P_HIP_EINVAL, "downward task out of range"));
        for (const RedTask &t : red) {
            if (!in(t.dst, t.len, Y1) || t.cbeg < 0 || t.cend < t.cbeg || t.cend > NC)
                return bail(fail(H2P_HIP_EINVAL, "reduction task out of range"));
            for (int k = t.cbeg; k < t.cend; k++)
                if (!in(contrib[k], t.len, AR)) return bail(fail(H2P_HIP_EINVAL, "reduction source out of range"));
        }
        for (const LeafTask2 &t : leaf) {
            if (!in(t.vec_s, t.nrow, N) || t.cb1 < 0 || t.ce1 < t.cb1 || t.cb0 != t.ce1 || t.ce0 < t.cb0 ||
                t.ce0 > NC || t.r > H2P_MAXR)
                return bail(fail(H2P_HIP_EINVAL, "leaf task out of range"));
            if (t.r > 0 && (!in(t.s_off, t.r, Y1) || !in(t.ut_off, (long long)t.nrow * t.r, utot)))
                return bail(fail(H2P_HIP_EINVAL, "leaf task out of range"));
            for (int k = t.cb1; k < t.ce1; k++)
                if (!in(contrib[k], t.r, AR)) return bail(fail(H2P_HIP_EINVAL, "leaf source out of range"));
            for (int k = t.cb0; k < t.ce0; k++)
                if (!in(contrib[k], t.nrow, AR)) return bail(fail(H2P_HIP_EINVAL, "leaf source out of range"));
        }
        for (const BlkTask &t : blk) {
            bool ok = in(t.in0, t.ncol, XIN) && in(t.out0, t.nrow, AR) && in(t.c0, t.n0, CT) && in(t.c1, t.n1, CT) &&
                      t.nrow == t.n0 * kd && t.ncol == t.n1 * kd;
            if (t.flags & 1) ok = ok && in(t.in1, t.nrow, XIN) && in(t.out1, t.ncol, AR);
            if (mode == H2P_HIP_AOT) ok = ok && in(t.data_off, (long long)t.nrow * t.ncol, BLOB);
            ok = ok && t.shift >= 0 && t.shift < (int)(shtab.size() / 3);
            if (!ok) return bail(fail(H2P_HIP_EINVAL, "block task out of range"));
        }
        for (int i = 0; i < N; i++)
            if (fwdidx_ok(g->fwd_pmt_idx[i], N) == false) return bail(fail(H2P_HIP_EINVAL, "fwd_pmt_idx out of range"));
        const long long rows_of[4] = {N, S, S, N};
        for (const MMTask &t : mmt) {
            bool ok = t.dst_space >= 0 && t.dst_space < 4 && in(t.dst_off, t.M, rows_of[t.dst_space]) &&
                      t.sbeg >= 0 && t.send >= t.sbeg && t.send <= (int)mms.size();
            for (int k = t.sbeg; ok && k < t.send; k++) {
                const MMSrc &x = mms[k];
                ok = x.x_space >= 0 && x.x_space < 4 && in(x.x_off, x.K, rows_of[x.x_space]) && x.K >= 0;
                if (!ok) break;
                if (x.kind == MM_A_JIT) {
                    ok = (t.M % kd == 0) && (x.K % kd == 0) && in(x.c0, t.M / kd, CT) && in(x.c1, x.K / kd, CT) &&
                         x.shift >= 0 && x.shift < (int)(shtab.size() / 3);
                } else {
                    const long long per_cap = (long long)g->U_ncol[root] * g->U_ncol[root];
                    const long long cap = (x.kind == MM_A_BD) ? (per ? per_cap : BLOB) : utot;
                    const long long rows = x.trans ? x.K : t.M, cols = x.trans ? t.M : x.K;
                    ok = (x.kind == MM_A_U || x.kind == MM_A_UT || x.kind == MM_A_BD) && x.lda >= cols &&
                         (rows == 0 || in(x.a_off, (rows - 1) * (long long)x.lda + cols, cap));
                }
            }
            if (!ok) return bail(fail(H2P_HIP_EINVAL, "matmul task out of range"));
        }
    }
    // ---- level split (after the bounds checks above): the two chains touch
    // what the schedule assumes.
    // The shallow chain (upward sweep above lv_cut, part 3, the first
    // n_red_shallow skeleton sums, the first n_bwd_shallow downward levels)
    // runs beside part 0 after the deep upward sweep; then s adds the deep
    // skeleton sums and runs the deep downward levels.  Checked here: every
    // skeleton coefficient a B block reads comes from its own chain's upward
    // sweep (or from the deep one, which both chains wait for), every
    // partial a skeleton sum adds comes from its own chain's B blocks, and
    // every skeleton sum a downward task reads comes from its own chain.
    h->lv_possible = (h->split && h->n_fwd_deep < (int)h->fwd_levels.size()) ? 1 : 0;
    if (h->lv_possible) {
        const int fd = h->n_fwd_deep, bs = h->n_bwd_shallow, rs = h->n_red_shallow;
        // y0 entries: 1 = deep upward sweep, 2 = shallow upward sweep
        std::vector<char> y0c((size_t)S, 0), y1c((size_t)S, 0);
        for (int k = 0; k < (int)h->fwd_levels.size(); k++) {
            const Level &L = h->fwd_levels[k];
            for (int i = L.off; i < L.off + L.n; i++) {
                const FwdTask &t = fwdt[i];
                for (int c = t.col0; c < std::min(t.r, t.col0 + t.ncol); c++)
                    y0c[(size_t)(t.out_off - N + c)] = k < fd ? 1 : 2;
            }
        }
        // partial ranges of the B blocks: (start, length, chain)
        std::vector<std::array<long long, 3>> pr;
        auto y0_ok = [&](int off, int len, int chain) {
            if (off < N) return true;   // px
            for (int e = 0; e < len; e++) {
                const char c = y0c[(size_t)(off - N + e)];
                if (c != 1 && c != chain) return false;
            }
            return true;
        };
        bool ok = true;
        for (int k : {0, 3}) {
            const int chain = k == 0 ? 1 : 2;
            const auto &P = h->part[k];
            for (int i = P.off; i < P.off + P.n && ok; i++) {
                const BlkTask &t = blk[i];
                ok = y0_ok(t.in0, t.ncol, chain) && (!(t.flags & 1) || y0_ok(t.in1, t.nrow, chain));
                pr.push_back({t.out0, (long long)t.nrow, chain});
                if (t.flags & 1) pr.push_back({t.out1, (long long)t.ncol, chain});
            }
        }
        std::sort(pr.begin(), pr.end());
        auto chain_of = [&](long long o) {
            auto it = std::upper_bound(pr.begin(), pr.end(), std::array<long long, 3>{o, LLONG_MAX, LLONG_MAX});
            if (it == pr.begin()) return 0LL;
            --it;
            return o < (*it)[0] + (*it)[1] ? (*it)[2] : 0LL;
        };
        for (int i = 0; i < h->n_red && ok; i++) {
            const RedTask &t = red[i];
            const int chain = i < rs ? 2 : 1;
            for (int c = t.cbeg; c < t.cend && ok; c++) ok = chain_of(contrib[c]) == chain;
            for (int e = 0; e < t.len; e++) y1c[(size_t)t.dst + e] = (char)chain;
        }
        for (int k = 0; k < (int)h->bwd_levels.size() && ok; k++) {
            const Level &L = h->bwd_levels[k];
            const int chain = k < bs ? 2 : 1;
            for (int i = L.off; i < L.off + L.n && ok; i++) ok = y1c[(size_t)bwdt[i].y1_off] == chain;
        }
        if (!ok) return bail(fail(H2P_HIP_EINVAL, "level split: a task reads data of the other chain"));
    }
    if (per && (g->n_child[root] == 0 || g->U_ncol[root] == 0))
        return bail(fail(H2P_HIP_EINVAL, "periodic mode needs a root with children and a basis"));
    const size_t per_elems = per ? (size_t)g->U_ncol[root] * g->U_ncol[root] : 0;
    if (dry) {
        auto vb = [](const auto &v) {
            return sizeof(typename std::decay_t<decltype(v)>::value_type) * std::max<size_t>(v.size(), 1);
        };
        size_t b = vb(cu) + vb(Ub) + vb(UTb) + sizeof(int) * std::max<size_t>(N, 1) + vb(blk) + vb(fwdt) +
                   vb(bwdt) + vb(red) + vb(leaf) + vb(contrib) + vb(mmt) + vb(mms) + vb(shtab) +
                   (per ? sizeof(double) * per_elems : 0);
        b += sizeof(double) * (std::max<size_t>((size_t)N + S, 1) + 2 * std::max<size_t>(S, 1) +
                               std::max<size_t>((size_t)arena, 1));
        if (h->split) b += sizeof(double) * std::max<size_t>(N, 1);
        if (mode == H2P_HIP_AOT)
            b += (prec == H2P_HIP_FP32MIX ? sizeof(float) : sizeof(double)) * std::max<size_t>(btot + dtot, 1);
        *dry = b;
        free_handle(h);
        return 0;
    }
    // ---- device memory
    size_t *acct = &h->device_bytes;
    std::vector<int> fwdidx(g->fwd_pmt_idx, g->fwd_pmt_idx + N);
    int rc;
    if ((rc = dev_upload(&h->d_coord, cu, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_U, Ub, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_UT, UTb, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_fwd, fwdidx, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_blk, blk, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_fwdt, fwdt, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_bwdt, bwdt, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_red, red, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_leaf, leaf, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_contrib, contrib, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_mmt, mmt, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_mms, mms, acct))) return bail(rc);
    if ((rc = dev_upload(&h->d_shift, shtab, acct))) return bail(rc);
    if (per) {
        std::vector<double> pv(g->per_blk, g->per_blk + per_elems);
        if ((rc = dev_upload(&h->d_per, pv, acct))) return bail(rc);
        h->per_r = g->U_ncol[root];
        h->per_soff = soff[root];
    }
    h->h_mmt = std::move(mmt);
    h->h_mms = std::move(mms);
    h->h_blk = blk;
    // FP32-mixed handles without FP32 matmul tiles (RPY, whose matmul tiles
    // are FP64 only) refuse matmul rather than run it in FP64
    h->mm_ok = !(prec == H2P_HIP_FP32MIX && !mm_f32);
    if ((rc = dev_alloc(&h->d_xin, (size_t)N + S, acct))) return bail(rc);
    if ((rc = dev_alloc(&h->d_y1, (size_t)S, acct))) return bail(rc);
    if ((rc = dev_alloc(&h->d_yp, (size_t)S, acct))) return bail(rc);
    if (h->split && (rc = dev_alloc(&h->d_yt, (size_t)N, acct))) return bail(rc);
    if ((rc = dev_alloc(&h->d_arena, (size_t)arena, acct))) return bail(rc);
    if (hipMemset(h->d_xin, 0, sizeof(double) * ((size_t)N + S)) != hipSuccess)
        return bail(fail(H2P_HIP_EDEVICE, "hipMemset failed"));
    if (mode == H2P_HIP_AOT) {
        const bool f32 = prec == H2P_HIP_FP32MIX;
        const size_t esz = f32 ? sizeof(float) : sizeof(double);
        {
            void *p = nullptr;
            size_t bytes = esz * std::max<size_t>(btot + dtot, 1);
            if (hipMalloc(&p, bytes) != hipSuccess) return bail(fail(H2P_HIP_ENOMEM, "hipMalloc of the AOT blocks failed"));
            h->d_blob = p;
            *acct += bytes;
        }
        if (g->B_data && g->D_data) {
            if (!f32) {
                double *blob = (double *)h->d_blob;
                if (btot && hipMemcpy(blob, g->B_data, sizeof(double) * btot, hipMemcpyHostToDevice) != hipSuccess)
                    return bail(fail(H2P_HIP_EDEVICE, "B upload failed"));
                if (dtot && hipMemcpy(blob + btot, g->D_data, sizeof(double) * dtot, hipMemcpyHostToDevice) != hipSuccess)
                    return bail(fail(H2P_HIP_EDEVICE, "D upload failed"));
            } else {
                // stage FP64 host blocks through a device buffer and round them to FP32
                const size_t chunk = (size_t)32 << 20;
                double *stage = nullptr;
                if (hipMalloc(&stage, sizeof(double) * chunk) != hipSuccess) return bail(fail(H2P_HIP_ENOMEM, "hipMalloc failed"));
                float *blob = (float *)h->d_blob;
                const double *srcs[2] = {g->B_data, g->D_data};
                const size_t lens[2] = {btot, dtot}, offs[2] = {0, btot};
                int err = 0;
                for (int part = 0; part < 2 && !err; part++)
                    for (size_t o = 0; o < lens[part] && !err; o += chunk) {
                        const size_t n = std::min(chunk, lens[part] - o);
                        if (hipMemcpy(stage, srcs[part] + o, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) err = 1;
                        else hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, 0, stage, blob + offs[part] + o, (long long)n);
                    }
                if (!err && hipDeviceSynchronize() != hipSuccess) err = 1;
                (void)hipFree(stage);
                if (err) return bail(fail(H2P_HIP_EDEVICE, "block upload failed"));
            }
        } else if (!g->B_data && !g->D_data) {
            if (h->n_blk > 0) dispatch_build(h, 0);
            if (hipDeviceSynchronize() != hipSuccess) return bail(fail(H2P_HIP_EDEVICE, "device block build failed"));
        } else {
            return bail(fail(H2P_HIP_EINVAL, "give both B_data and D_data or neither"));
        }
    }
    if (hipStreamCreate(&h->own_stream) != hipSuccess) return bail(fail(H2P_HIP_EDEVICE, "stream create failed"));
    if (hipEventCreateWithFlags(&h->ev_replay, hipEventDisableTiming) != hipSuccess)
        return bail(fail(H2P_HIP_EDEVICE, "event create failed"));
    if (h->split) {
        if (hipStreamCreate(&h->aux_stream) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_mid, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
            return bail(fail(H2P_HIP_EDEVICE, "auxiliary stream / event create failed"));
        if (h->lv_possible &&
            (hipStreamCreate(&h->lv_stream) != hipSuccess ||
             hipEventCreateWithFlags(&h->ev_lv, hipEventDisableTiming) != hipSuccess ||
             hipEventCreateWithFlags(&h->ev_side, hipEventDisableTiming) != hipSuccess))
            return bail(fail(H2P_HIP_EDEVICE, "level-split stream / event create failed"));
        h->lv_on = h->lv_possible;
    }
    // the upload is synchronous: callers may use the handle on any stream,
    // including streams that do not order with the null stream
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(H2P_HIP_EDEVICE, "upload did not complete"));
    {
        const char *ng = getenv("H2P_HIP_NO_GRAPH");
        h->use_graph = (ng && ng[0] == '1') ? 0 : 1;
    }
    // ---- statistics (H2Pack's mat_size / JIT_flops model)
    h2p_hip_info &I = h->info;
    I.U_elems = (size_t)utot;
    I.B_elems = btot;
    I.D_elems = dtot;
    // mat_size[MV_FWD/MID/BWD/DEN] of src/H2Pack_build.c:251-255, 755-758, 965-1002
    size_t vec = 0;
    for (int i = 0; i < nn; i++) vec += 2 * (size_t)(g->U_nrow[i] + g->U_ncol[i]);
    for (int i = 0; i < g->n_r_adm_pair; i++) vec += 2 * (size_t)(g->B_nrow[i] + g->B_ncol[i]);
    for (int i = 0; i < nD; i++) vec += (size_t)(i < nlf ? 1 : 2) * (g->D_nrow[i] + g->D_ncol[i]);
    const size_t blk_esz = (prec == H2P_HIP_FP32MIX) ? sizeof(float) : sizeof(double);
    I.hbm_bytes_model = 8 * (2 * (size_t)utot + vec) + (mode == H2P_HIP_AOT ? blk_esz * (btot + dtot) : 0);
    I.B_pairs = B_pairs;
    I.D_pairs = D_pairs;
    static const int flops[4] = {14, 14, 17, 82};
    I.jit_flops_model = flops[h->kid] * (B_pairs + D_pairs);
    I.device_bytes = h->device_bytes;
    I.n_block_tasks = h->n_blk;
    I.n_levels = g->max_level + 1;
    I.mode = mode;
    I.prec = prec;
    *out = h;
    return 0;
}

// matmul panel rows of a handle: PX (N) | Y0 (S) | Y1 (S) | PY (N)
static size_t panel_rows(int N, int S) { return 2 * (size_t)N + 2 * (size_t)S; }
static int panel_width(int n_vec) { return std::min(H2P_MM_MAXNVP, (n_vec + 15) & ~15); }

static int ensure_panels(h2p_hip_handle *h, int nvp) {
    if (h->panel_nvp >= nvp) return 0;
    const size_t rows = panel_rows(h->N, h->S);
    drop_mm_graphs(h);
    if (h->d_panels) {
        // the handle's last product (on whichever stream) may still read the
        // old panels
        wait_replays(h);
        (void)hipFree(h->d_panels);
        h->d_panels = nullptr;
        h->device_bytes -= sizeof(double) * rows * h->panel_nvp;
        h->panel_nvp = 0;
    }
    if (hipMalloc(&h->d_panels, sizeof(double) * rows * nvp) != hipSuccess)
        return fail(H2P_HIP_ENOMEM, "hipMalloc of the matmul panels failed");
    h->panel_nvp = nvp;
    h->device_bytes += sizeof(double) * rows * nvp;
    h->info.device_bytes = h->device_bytes;
    return 0;
}

static MMPanels panels_of(const h2p_hip_handle *h, int nvp, double *base) {
    MMPanels P;
    P.nvp = nvp;
    P.p[MM_PX] = base;
    P.p[MM_Y0] = P.p[MM_PX] + (size_t)h->N * nvp;
    P.p[MM_Y1] = P.p[MM_Y0] + (size_t)h->S * nvp;
    P.p[MM_PY] = P.p[MM_Y1] + (size_t)h->S * nvp;
    return P;
}

static int cg_blocks(int N) { return (N + H2P_CG_PER_BLOCK - 1) / H2P_CG_PER_BLOCK; }
static size_t cg_doubles(int N) {
    return 3 * (size_t)N + 2 * (size_t)cg_blocks(N) + (sizeof(CGScal) + 7) / 8;
}

extern "C" {

int h2p_hip_upload(const h2p_hip_h2pack *g, const h2p_kernel_desc *krnl, int mode, int prec,
                   h2p_hip_handle **out) {
    return upload_impl(g, krnl, mode, prec, out, nullptr);
}

int h2p_hip_footprint(const h2p_hip_h2pack *g, const h2p_kernel_desc *krnl, int mode, int prec, int n_vec,
                      int with_cg, size_t *bytes) {
    if (!bytes) return fail(H2P_HIP_EINVAL, "bytes is NULL");
    *bytes = 0;
    if (n_vec < 0) return fail(H2P_HIP_EINVAL, "n_vec < 0");
    h2p_hip_handle *dummy = nullptr;
    size_t b = 0;
    int rc = upload_impl(g, krnl, mode, prec, &dummy, &b);
    if (rc) return rc;
    // S is needed for the panels: the sum of the U column counts
    size_t S = 0;
    for (int i = 0; i < g->n_node; i++) S += (size_t)g->U_ncol[i];
    const int N = g->krnl_mat_size;
    if (n_vec > 0) b += sizeof(double) * panel_rows(N, (int)S) * panel_width(n_vec);
    if (with_cg) b += sizeof(double) * cg_doubles(N);
    *bytes = b;
    return 0;
}

int h2p_hip_matvec_dev(h2p_hip_handle *h, const double *d_x, double *d_y, void *stream) {
    if (!h || !d_x || !d_y) return fail(H2P_HIP_EINVAL, "NULL argument");
    return run_matvec(h, d_x, d_y, (hipStream_t)stream);
}

int h2p_hip_matvec(h2p_hip_handle *h, const double *x, double *y) {
    if (!h || !x || !y) return fail(H2P_HIP_EINVAL, "NULL argument");
    int rc;
    if (h->mm_cap < 1) {
        if (h->d_mm_x) { (void)hipFree(h->d_mm_x); h->d_mm_x = nullptr; }
        if (h->d_mm_y) { (void)hipFree(h->d_mm_y); h->d_mm_y = nullptr; }
        size_t dummy = 0;
        if ((rc = dev_alloc(&h->d_mm_x, (size_t)h->N, &dummy))) return rc;
        if ((rc = dev_alloc(&h->d_mm_y, (size_t)h->N, &dummy))) return rc;
        h->mm_cap = 1;
    }
    hipStream_t s = h->own_stream;
    HIP_TRY(hipMemcpyAsync(h->d_mm_x, x, sizeof(double) * h->N, hipMemcpyHostToDevice, s));
    if ((rc = run_matvec(h, h->d_mm_x, h->d_mm_y, s))) return rc;
    HIP_TRY(hipMemcpyAsync(y, h->d_mm_y, sizeof(double) * h->N, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"

// Argument checks shared by the matmul entry points.
static int check_matmul_args(const h2p_hip_handle *h, int layout, int n_vec, const void *X, int ldx, const void *Y,
                             int ldy) {
    if (!h || (n_vec > 0 && (!X || !Y))) return fail(H2P_HIP_EINVAL, "NULL argument");
    if (n_vec < 0) return fail(H2P_HIP_EINVAL, "n_vec < 0");
    if (n_vec == 0) return 0;
    const bool rm = layout == H2P_HIP_ROW_MAJOR;
    if (!rm && layout != H2P_HIP_COL_MAJOR) return fail(H2P_HIP_EINVAL, "bad layout");
    const int N = h->N;
    if (rm ? (ldx < n_vec || ldy < n_vec) : (ldx < N || ldy < N)) return fail(H2P_HIP_EINVAL, "bad leading dimension");
    if (!h->mm_ok) return fail(H2P_HIP_ENOSYS, "FP32-mixed matmul is available for the scalar kernels in JIT mode and FP32-stored AOT blocks of them");
    return 0;
}

// The chunks of one matmul on panels starting at `panels` (panel_rows x
// panel_width(n_vec) doubles).  graphs: replay the handle's captured chunk
// graphs; otherwise every phase is launched directly.
static int run_matmul(h2p_hip_handle *h, int layout, int n_vec, const double *d_X, int ldx, double *d_Y, int ldy,
                      hipStream_t s, double *panels, bool graphs) {
    const bool rm = layout == H2P_HIP_ROW_MAJOR;
    const int N = h->N;
    int rc = 0;
    // columns in chunks of at most 64 (the reference chunks by mm_max_n_vec,
    // src/H2Pack_matmul.c:497-580)
    for (int v0 = 0; v0 < n_vec; v0 += H2P_MM_MAXNVP) {
        const int nv = std::min(H2P_MM_MAXNVP, n_vec - v0);
        const int nvp = (nv + 15) & ~15;
        const MMPanels P = panels_of(h, nvp, panels);
        const long long ne = (long long)N * nvp;
        hipLaunchKernelGGL(k_mm_gather, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, d_X, (long long)ldx,
                           rm ? 1 : 0, v0, nv, h->d_fwd, N, P.p[MM_PX], nvp);
        const int c = nv_class(nvp);
        if (graphs) {
            hipGraphExec_t ge = nullptr;
            if ((rc = mm_graph(h, P, &ge))) return rc;
            HIP_TRY(hipGraphLaunch(ge, s));
        } else {
            for (int ph = 0; ph < 4; ph++) launch_mm_phase(h, ph, P, s, h->mm_var[ph][c]);
        }
        const long long no = (long long)N * nv;
        hipLaunchKernelGGL(k_mm_scatter, dim3((unsigned)((no + 255) / 256)), dim3(256), 0, s, P.p[MM_PY], nvp,
                           h->d_fwd, N, d_Y, (long long)ldy, rm ? 1 : 0, v0, nv);
        HIP_TRY(hipGetLastError());
    }
    mark_use(h, s);
    return 0;
}

// Panels of the allocation-free matmul: one static device array per GPU,
// shared by every call on that GPU.  Host code never passes `scratch`
// itself (that is a host-side stand-in); it passes the device address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device on its first call.
#define H2P_SCRATCH_BYTES H2P_HIP_NOALLOC_BYTES
static __device__ __align__(256) double scratch[H2P_SCRATCH_BYTES / sizeof(double)];
static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;

static int scratch_ptr(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return 0;
    }
    double *p = nullptr;
    HIP_TRY(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(H2P_HIP_EDEVICE, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = p;
    *out = p;
    return 0;
}

extern "C" {

int h2p_hip_matmul_dev(h2p_hip_handle *h, int layout, int n_vec, const double *d_X, int ldx,
                       double *d_Y, int ldy, void *stream) {
    int rc = check_matmul_args(h, layout, n_vec, d_X, ldx, d_Y, ldy);
    if (rc || n_vec == 0) return rc;
    if ((rc = ensure_panels(h, panel_width(n_vec)))) return rc;
    return run_matmul(h, layout, n_vec, d_X, ldx, d_Y, ldy, (hipStream_t)stream, h->d_panels, h->use_graph != 0);
}

int h2p_hip_matmul_dev_noalloc(h2p_hip_handle *h, int layout, int n_vec, const double *d_X, int ldx,
                               double *d_Y, int ldy, void *stream) {
    int rc = check_matmul_args(h, layout, n_vec, d_X, ldx, d_Y, ldy);
    if (rc || n_vec == 0) return rc;
    const size_t need = sizeof(double) * panel_rows(h->N, h->S) * panel_width(n_vec);
    if (need > H2P_SCRATCH_BYTES)
        return fail(H2P_HIP_EINVAL, "matmul panels need " + std::to_string(need) + " bytes, more than the " +
                                        std::to_string(H2P_SCRATCH_BYTES) + "-byte static scratch");
    double *panels = nullptr;
    if ((rc = scratch_ptr(&panels))) return rc;
    // no graph: instantiating one allocates
    return run_matmul(h, layout, n_vec, d_X, ldx, d_Y, ldy, (hipStream_t)stream, panels, false);
}

// Tile-variant choices shared by handles whose matmul phase has the same
// kernel, mode, panel width class and block-size histogram.
static std::mutex g_tune_mu;
static std::map<std::string, int> g_tune_cache;

static std::string tune_signature(const h2p_hip_handle *h, int ph, int cls, const std::vector<MMTask> &tasks,
                                  const std::vector<MMSrc> &srcs) {
    // histogram of (log2 M, log2 K) over the phase's (task, source) products
    std::map<std::pair<int, int>, long long> hist;
    std::vector<Level> lv = mm_phase_levels(h, ph), lv32 = mm_phase_levels32(h, ph);
    lv.insert(lv.end(), lv32.begin(), lv32.end());
    for (const Level &L : lv)
        for (int i = L.off; i < L.off + L.n; i++) {
            const MMTask &t = tasks[i];
            for (int k = t.sbeg; k < t.send; k++) {
                const int bm = t.M > 0 ? (int)std::floor(std::log2((double)t.M)) : -1;
                const int bk = srcs[k].K > 0 ? (int)std::floor(std::log2((double)srcs[k].K)) : -1;
                hist[{bm, bk}]++;
            }
        }
    std::string sig = std::to_string(h->kid) + "/" + std::to_string(h->mode) + "/" + std::to_string(ph) + "/" +
                      std::to_string(cls) + (h->mm_f32 ? "/f32" : "") + ":";
    for (const auto &e : hist) {
        // counts on a log2 scale, so nearby histograms share a choice
        const int c = (int)std::floor(std::log2((double)e.second));
        sig += std::to_string(e.first.first) + "," + std::to_string(e.first.second) + "," + std::to_string(c) + ";";
    }
    return sig;
}

int h2p_hip_matmul(h2p_hip_handle *h, int layout, int n_vec, const double *X, int ldx, double *Y,
                   int ldy) {
    if (!h || (n_vec > 0 && (!X || !Y))) return fail(H2P_HIP_EINVAL, "NULL argument");
    if (n_vec <= 0) return n_vec < 0 ? fail(H2P_HIP_EINVAL, "n_vec < 0") : 0;
    const int N = h->N;
    const bool rm = layout == H2P_HIP_ROW_MAJOR;
    if (!rm && layout != H2P_HIP_COL_MAJOR) return fail(H2P_HIP_EINVAL, "bad layout");
    if (rm ? (ldx < n_vec || ldy < n_vec) : (ldx < N || ldy < N)) return fail(H2P_HIP_EINVAL, "bad leading dimension");
    // pack the caller's panel densely, run on the device, unpack
    double *dX = nullptr, *dY = nullptr;
    const size_t elems = (size_t)N * n_vec;
    if (hipMalloc(&dX, sizeof(double) * elems) != hipSuccess) return fail(H2P_HIP_ENOMEM, "hipMalloc failed");
    if (hipMalloc(&dY, sizeof(double) * elems) != hipSuccess) { (void)hipFree(dX); return fail(H2P_HIP_ENOMEM, "hipMalloc failed"); }
    hipStream_t s = h->own_stream;
    int rc = 0;
    const int rows = rm ? N : n_vec, cols = rm ? n_vec : N;
    if (hipMemcpy2DAsync(dX, sizeof(double) * cols, X, sizeof(double) * ldx, sizeof(double) * cols, rows,
                         hipMemcpyHostToDevice, s) != hipSuccess)
        rc = fail(H2P_HIP_EDEVICE, "upload of X failed");
    if (!rc) rc = h2p_hip_matmul_dev(h, layout, n_vec, dX, cols, dY, cols, s);
    if (!rc && hipMemcpy2DAsync(Y, sizeof(double) * ldy, dY, sizeof(double) * cols, sizeof(double) * cols, rows,
                                hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = fail(H2P_HIP_EDEVICE, "download of Y failed");
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(H2P_HIP_EDEVICE, "stream sync failed");
    (void)hipFree(dX);
    (void)hipFree(dY);
    return rc;
}

// Mean time of the matvec interior over reps calls after one warm-up,
// launched eagerly on the handle's streams: the same kernels, streams and
// events as the captured matvec graph, without instantiating a graph per
// schedule (each instantiate / launch / destroy cycle of a forked graph is
// one more pass through the runtime's graph launch path, where processes
// have taken host faults).
static int time_interior(h2p_hip_handle *h, int reps, double *ms) {
    hipStream_t s = h->own_stream;
    launch_interior(h, s, nullptr);   // warm-up
    HIP_TRY(hipGetLastError());
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float t = 0.f;
    int rc = hipEventRecord(e0, s) == hipSuccess ? 0 : fail(H2P_HIP_EDEVICE, "hipEventRecord failed");
    for (int r = 0; r < reps && !rc; r++) launch_interior(h, s, nullptr);
    if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        rc = fail(H2P_HIP_EDEVICE, "timing the matvec interior failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
    return rc;
}

// n_vec == 0: the matvec's grouped block launch (default or occupancy
// variant; the RPY JIT kernel has one variant), then, on a split handle,
// whether its D blocks run beside the sweeps on the second stream or in
// order (concurrency pays when the D blocks cover the sweeps, and can cost
// time where the sweeps are long and the D blocks few).  Same results.
// Schedule of a split handle's matvec (bit 0: D blocks in order, bit 1:
// level split off); the captured graph is dropped when it changes.
static void set_schedule(h2p_hip_handle *h, int k) {
    const bool ser = h->split && (k & 1);
    const int lvon = (h->lv_possible && !(k & 2)) ? 1 : 0;
    if (h->d_serial != ser || h->lv_on != lvon) drop_matvec_graph(h);
    h->d_serial = ser;
    h->lv_on = lvon;
}

static int autotune_matvec(h2p_hip_handle *h, double *t_default_ms, double *t_tuned_ms) {
    if (t_default_ms) *t_default_ms = 0.0;
    if (t_tuned_ms) *t_tuned_ms = 0.0;
    if (h->n_blk == 0 || (h->mode == H2P_HIP_JIT && h->kid == KR)) return 0;
    std::map<std::pair<int, int>, long long> hist;
    for (const BlkTask &t : h->h_blk) {
        const int a = t.nrow > 0 ? (int)std::floor(std::log2((double)t.nrow)) : -1;
        const int b = t.ncol > 0 ? (int)std::floor(std::log2((double)t.ncol)) : -1;
        hist[{a, b}]++;
    }
    std::string sig = "mv/" + std::to_string(h->kid) + "/" + std::to_string(h->mode) + "/" + std::to_string(h->prec) + ":";
    for (const auto &e : hist)
        sig += std::to_string(e.first.first) + "," + std::to_string(e.first.second) + "," +
               std::to_string((int)std::floor(std::log2((double)e.second))) + ";";
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(sig);
        if (it != g_tune_cache.end()) {
            const int v = it->second & 1;
            const bool ser = h->split && (it->second & 2);
            const int lvon = (h->lv_possible && !(it->second & 4)) ? 1 : 0;
            if (h->mv_var != v || h->d_serial != ser || h->lv_on != lvon) drop_matvec_graph(h);
            h->mv_var = v;
            h->d_serial = ser;
            h->lv_on = lvon;
            return 0;
        }
    }
    hipStream_t s = h->own_stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const int saved = h->mv_var, reps = 3;
    double t[2] = {0.0, 0.0};
    int err = 0;
    for (int v = 0; v < 2 && !err; v++) {
        h->mv_var = v;
        for (int k = 0; k < 4; k++) launch_blocks(h, s, k);   // warm-up
        if (hipEventRecord(e0, s) != hipSuccess) { err = 1; break; }
        for (int r = 0; r < reps; r++)
            for (int k = 0; k < 4; k++) launch_blocks(h, s, k);
        float ms = 0.f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            err = 1;
        t[v] = ms / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->mv_var = saved;
    if (err) return fail(H2P_HIP_EDEVICE, "timing the matvec variants failed");
    const int best = t[1] < t[0] ? 1 : 0;
    if (best != h->mv_var) drop_matvec_graph(h);
    h->mv_var = best;
    if (t_default_ms) *t_default_ms = t[0];
    if (t_tuned_ms) *t_tuned_ms = t[best];
    int code = best;
    if (h->split) {
        // schedules: bit 0 = D blocks in order (else beside the sweeps),
        // bit 1 = level split off (else the shallow levels' chain runs
        // beside the deep B blocks)
        const int nsch = h->lv_possible ? 4 : 2;
        double ti[4] = {0.0, 0.0, 0.0, 0.0};
        int bk = 0;
        for (int k = 0; k < nsch; k++) {
            set_schedule(h, k);
            int rc = time_interior(h, 8, &ti[k]);
            if (rc) return rc;
            if (ti[k] < ti[bk]) bk = k;
        }
        set_schedule(h, bk);
        code |= ((bk & 1) ? 2 : 0) | ((bk & 2) ? 4 : 0);
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[sig] = code;
    return 0;
}

int h2p_hip_autotune(h2p_hip_handle *h, int n_vec, double *t_default_ms, double *t_tuned_ms) {
    if (!h) return fail(H2P_HIP_EINVAL, "NULL handle");
    if (n_vec < 0) return fail(H2P_HIP_EINVAL, "n_vec must not be negative");
    if (n_vec == 0) return autotune_matvec(h, t_default_ms, t_tuned_ms);
    if (!h->mm_ok) return fail(H2P_HIP_ENOSYS, "FP32-mixed matmul is available for the scalar kernels in JIT mode and FP32-stored AOT blocks of them");
    const int nvp = panel_width(n_vec), cls = nv_class(nvp);
    int rc = ensure_panels(h, nvp);
    if (rc) return rc;
    const MMPanels P = panels_of(h, nvp, h->d_panels);
    hipStream_t s = h->own_stream;
    HIP_TRY(hipMemsetAsync(h->d_panels, 0, sizeof(double) * panel_rows(h->N, h->S) * nvp, s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double tdef = 0.0, ttun = 0.0;
    int err = 0;
    for (int ph = 0; ph < 4 && !err; ph++) {
        bool empty = true;
        for (const Level &L : mm_phase_levels(h, ph)) empty = empty && L.n <= 0;
        for (const Level &L : mm_phase_levels32(h, ph)) empty = empty && L.n <= 0;
        if (empty) continue;
        const std::string sig = tune_signature(h, ph, cls, h->h_mmt, h->h_mms);
        double t[H2P_MM_NVARIANT];
        int best = 0;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune_cache.find(sig);
            if (it != g_tune_cache.end()) {
                if (h->mm_var[ph][cls] != it->second) drop_mm_graphs(h);
                h->mm_var[ph][cls] = it->second;
                continue;
            }
        }
        const int reps = 3;
        for (int v = 0; v < H2P_MM_NVARIANT && !err; v++) {
            launch_mm_phase(h, ph, P, s, v);   // warm-up
            if (hipEventRecord(e0, s) != hipSuccess) { err = 1; break; }
            for (int r = 0; r < reps; r++) launch_mm_phase(h, ph, P, s, v);
            float ms = 0.f;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                err = 1;
            t[v] = ms / reps;
            if (t[v] < t[best]) best = v;
        }
        if (err) break;
        if (h->mm_var[ph][cls] != best) drop_mm_graphs(h);
        h->mm_var[ph][cls] = best;
        tdef += t[0];
        ttun += t[best];
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[sig] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err) return fail(H2P_HIP_EDEVICE, "timing the matmul variants failed");
    HIP_TRY(hipGetLastError());
    if (t_default_ms) *t_default_ms = tdef;
    if (t_tuned_ms) *t_tuned_ms = ttun;
    return 0;
}

int h2p_hip_mm_variant(h2p_hip_handle *h, int phase, int n_vec, int variant) {
    if (!h) return fail(H2P_HIP_EINVAL, "NULL handle");
    if (phase == 5) {   // the matvec's schedule on a split handle
        if (variant < -1 || variant > 3) return fail(H2P_HIP_EINVAL, "bad matvec schedule");
        if (variant < 0) return (h->d_serial ? 1 : 0) | (h->lv_on ? 0 : 2);
        set_schedule(h, variant);
        return 0;
    }
    if (phase == 4) {   // the matvec's block launch
        if (variant < -1 || variant > 1) return fail(H2P_HIP_EINVAL, "bad matvec variant");
        if (variant < 0) return h->mv_var;
        if (h->mode == H2P_HIP_JIT && h->kid == KR && variant != 0)
            return fail(H2P_HIP_ENOSYS, "the RPY JIT block launch has one variant");
        if (variant != h->mv_var) drop_matvec_graph(h);
        h->mv_var = variant;
        return 0;
    }
    if (phase < -1 || phase > 3 || n_vec <= 0 || variant < -1 || variant >= H2P_MM_NVARIANT)
        return fail(H2P_HIP_EINVAL, "bad phase, width or variant");
    const int cls = nv_class(panel_width(n_vec));
    if (variant < 0) return h->mm_var[phase < 0 ? 0 : phase][cls];
    for (int ph = 0; ph < 4; ph++)
        if ((phase < 0 || ph == phase) && h->mm_var[ph][cls] != variant) {
            drop_mm_graphs(h);
            h->mm_var[ph][cls] = variant;
        }
    return 0;
}

int h2p_hip_set_profiling(h2p_hip_handle *h, int on) {
    if (!h) return fail(H2P_HIP_EINVAL, "NULL handle");
    h->profiling = on ? 1 : 0;
    h->ev_used = 0;
    if (on && h->ev_cap == 0) {
        const int cap = 512;
        h->ev_pool.resize((size_t)cap * H2P_EV_PER_CALL);
        for (auto &e : h->ev_pool)
            if (hipEventCreate(&e) != hipSuccess) return fail(H2P_HIP_EDEVICE, "hipEventCreate failed");
        h->ev_cap = cap;
    }
    return 0;
}

int h2p_hip_get_info(h2p_hip_handle *h, h2p_hip_info *info) {
    if (!h || !info) return fail(H2P_HIP_EINVAL, "NULL argument");
    if (h->ev_used > 0) {
        // average the phases of the calls profiled since the last read
        // intervals: fwd, B blocks, D blocks, skeleton sums, bwd, leaf outputs
        double acc[H2P_EV_PER_CALL - 1] = {};
        for (int c = 0; c < h->ev_used; c++) {
            hipEvent_t *ev = &h->ev_pool[(size_t)c * H2P_EV_PER_CALL];
            HIP_TRY(hipEventSynchronize(ev[H2P_EV_PER_CALL - 1]));
            for (int i = 0; i < H2P_EV_PER_CALL - 1; i++) {
                float t = 0.f;
                HIP_TRY(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
                acc[i] += t;
            }
        }
        const double n = h->ev_used;
        h->info.t_fwd = acc[0] / n;
        h->info.t_blocks_b = acc[1] / n;
        h->info.t_blocks_d = acc[2] / n;
        h->info.t_blocks = (acc[1] + acc[2]) / n;
        h->info.t_reduce = acc[3] / n;
        h->info.t_bwd = acc[4] / n;
        h->info.t_final = acc[5] / n;
        h->info.t_total = (acc[0] + acc[1] + acc[2] + acc[3] + acc[4] + acc[5]) / n;
        h->info.n_profiled = h->ev_used;
        h->ev_used = 0;
    }
    h->info.mv_variant = h->mv_var;
    h->info.mv_d_concurrent = (h->split && !h->d_serial) ? 1 : 0;
    h->info.mv_level_split = (h->split && h->lv_on && h->lv_possible) ? 1 : 0;
    *info = h->info;
    return 0;
}

void h2p_hip_free(h2p_hip_handle *h) { free_handle(h); }

int h2p_hip_direct_nbody(const h2p_kernel_desc *krnl, int xpt_dim, const double *src, int src_ld,
                         int n_src, const double *x, const double *dst, int dst_ld, int n_dst,
                         double *y) {
    if (!krnl || !src || !x || !dst || !y) return fail(H2P_HIP_EINVAL, "NULL argument");
    if (n_src <= 0 || n_dst <= 0 || src_ld < n_src || dst_ld < n_dst) return fail(H2P_HIP_EINVAL, "bad sizes");
    const int kid = krnl->id;
    if (kid < 0 || kid > 3) return fail(H2P_HIP_EINVAL, "bad kernel id");
    const int kd = kid == KR ? 3 : 1;
    if (kid == KR && xpt_dim < 4) return fail(H2P_HIP_EINVAL, "RPY needs xpt_dim 4");
    KParam kp{};
    kp.p0 = (kid == KC) ? (krnl->n_param > 0 ? krnl->param[0] : 0.0) : krnl->param[0];
    if (kid == KR) kp.p1 = 1.0 / (6.0 * M_PI * krnl->param[0]);
    const int xd = kid == KR ? 4 : 3;
    const int chunk = 8192;
    const int nchunk = (n_src + chunk - 1) / chunk;
    double *ds = nullptr, *dx = nullptr, *dd = nullptr, *dp = nullptr, *dy = nullptr;
    int rc = 0;
    size_t acct = 0;
    std::vector<double> sc((size_t)xd * n_src), dc((size_t)xd * n_dst);
    for (int d = 0; d < xd; d++) {
        std::memcpy(sc.data() + (size_t)d * n_src, src + (size_t)d * src_ld, sizeof(double) * n_src);
        std::memcpy(dc.data() + (size_t)d * n_dst, dst + (size_t)d * dst_ld, sizeof(double) * n_dst);
    }
    std::vector<double> xv(x, x + (size_t)n_src * kd);
    if ((rc = dev_upload(&ds, sc, &acct)) || (rc = dev_upload(&dd, dc, &acct)) || (rc = dev_upload(&dx, xv, &acct)) ||
        (rc = dev_alloc(&dp, (size_t)nchunk * n_dst * kd, &acct)) || (rc = dev_alloc(&dy, (size_t)n_dst * kd, &acct))) {
    } else {
        dim3 grid((n_dst + 255) / 256, nchunk);
        switch (kid) {
            case KC: hipLaunchKernelGGL((k_nbody_partial<KC, 1>), grid, dim3(256), 0, 0, ds, (long long)n_src, n_src, dx, dd, (long long)n_dst, n_dst, kp, chunk, dp); break;
            case KG: hipLaunchKernelGGL((k_nbody_partial<KG, 1>), grid, dim3(256), 0, 0, ds, (long long)n_src, n_src, dx, dd, (long long)n_dst, n_dst, kp, chunk, dp); break;
            case KM: hipLaunchKernelGGL((k_nbody_partial<KM, 1>), grid, dim3(256), 0, 0, ds, (long long)n_src, n_src, dx, dd, (long long)n_dst, n_dst, kp, chunk, dp); break;
            default: hipLaunchKernelGGL((k_nbody_partial<KR, 3>), grid, dim3(256), 0, 0, ds, (long long)n_src, n_src, dx, dd, (long long)n_dst, n_dst, kp, chunk, dp); break;
        }
        int len = n_dst * kd;
        hipLaunchKernelGGL(k_nbody_reduce, dim3((len + 255) / 256), dim3(256), 0, 0, dp, nchunk, len, dy);
        if (hipDeviceSynchronize() != hipSuccess) rc = fail(H2P_HIP_EDEVICE, "n-body kernel failed");
        else if (hipMemcpy(y, dy, sizeof(double) * len, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(H2P_HIP_EDEVICE, "n-body download failed");
    }
    void *ps[] = {ds, dx, dd, dp, dy};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    return rc;
}

}  // extern "C"

// ===========================================================================
// Co-scheduled execution of many handles.
//
// Every call builds its launch tables on the host (a request table and, per
// launch, the items that map workgroups to (request, task slice)), copies
// them in one transfer from pinned staging to the device, and enqueues one
// launch per phase for all requests.  Two staging slots alternate; a slot is
// reused only after the launches that read it have finished.

namespace {

struct TableBuf {
    std::vector<char> bytes;
    template <typename T>
    size_t add(const T *p, size_t n) {
        size_t off = (bytes.size() + 15) & ~(size_t)15;
        bytes.resize(off + sizeof(T) * std::max<size_t>(n, 1));
        if (n) std::memcpy(bytes.data() + off, p, sizeof(T) * n);
        return off;
    }
};

// one launch over items of several requests
struct BLaunch {
    size_t items_off = 0;
    int n_items = 0, grid = 0;
    int mode = 0, prec = 0, kid = 0;   // kernel choice of the block launches
};

struct ItemList {
    std::vector<BatchItem> items;
    int total = 0;
    // count workgroups; block launches also give the item's workgroup-unit
    // tasks and its task count (wave units follow the workgroup units)
    void add(int req, const void *tasks, int count, int nbig = 0, int ntask = 0) {
        if (count <= 0) return;
        BatchItem it{};
        it.begin = total;
        it.req = req;
        it.tasks = tasks;
        it.nbig = nbig;
        it.ntask = ntask;
        items.push_back(it);
        total += count;
    }
};

static BLaunch finish(TableBuf &tb, const ItemList &L) {
    BLaunch b;
    b.n_items = (int)L.items.size();
    b.grid = L.total;
    b.items_off = tb.add(L.items.data(), L.items.size());
    return b;
}

struct MVPlan {
    size_t reqs_off = 0;
    BLaunch gather, red, leaf;
    std::vector<BLaunch> fwd, blocks, bwd;
};

// matvec tables for handles hs with inputs x, outputs y, shifts
static MVPlan plan_matvec(TableBuf &tb, int n, h2p_hip_handle *const *hs, const double *const *x,
                          double *const *y, const double *shift) {
    MVPlan P;
    std::vector<ReqCtx> reqs(n);
    size_t max_fwd = 0, max_bwd = 0;
    for (int i = 0; i < n; i++) {
        const h2p_hip_handle *h = hs[i];
        ReqCtx &c = reqs[i];
        c.coord = h->d_coord;
        c.CT = h->CT;
        c.scale = kscale(h->kid, h->kp.p0, h->prec == H2P_HIP_FP64);
        c.diag = h->kp.p0;
        c.kp = h->kp;
        c.xin = h->d_xin;
        c.y1 = h->d_y1;
        c.yp = h->d_yp;
        c.arena = h->d_arena;
        c.U = h->d_U;
        c.UT = h->d_UT;
        c.blob = h->d_blob;
        c.fwd = h->d_fwd;
        c.contrib = h->d_contrib;
        c.shifts = h->d_shift;
        c.N = h->N;
        c.pad = 0;
        c.x = x[i];
        c.y = y[i];
        c.shift = shift ? shift[i] : 0.0;
        max_fwd = std::max(max_fwd, h->fwd_levels.size());
        max_bwd = std::max(max_bwd, h->bwd_levels.size());
    }
    P.reqs_off = tb.add(reqs.data(), reqs.size());
    ItemList g, r, lf;
    for (int i = 0; i < n; i++) {
        g.add(i, nullptr, (hs[i]->N + 255) / 256);
        r.add(i, hs[i]->d_red, hs[i]->n_red);
        lf.add(i, hs[i]->d_leaf, hs[i]->n_leaf);
    }
    P.gather = finish(tb, g);
    P.red = finish(tb, r);
    P.leaf = finish(tb, lf);
    for (size_t k = 0; k < max_fwd; k++) {
        ItemList L;
        for (int i = 0; i < n; i++)
            if (k < hs[i]->fwd_levels.size()) {
                const Level &lv = hs[i]->fwd_levels[k];
                L.add(i, hs[i]->d_fwdt + lv.off, lv.n);
            }
        P.fwd.push_back(finish(tb, L));
    }
    for (size_t k = 0; k < max_bwd; k++) {
        ItemList L;
        for (int i = 0; i < n; i++)
            if (k < hs[i]->bwd_levels.size()) {
                const Level &lv = hs[i]->bwd_levels[k];
                L.add(i, hs[i]->d_bwdt + lv.off, lv.n);
            }
        P.bwd.push_back(finish(tb, L));
    }
    // block launches grouped by kernel choice, in order of first appearance
    std::vector<int> done(n, 0);
    for (int i = 0; i < n; i++) {
        if (done[i]) continue;
        ItemList L;
        for (int j = i; j < n; j++) {
            const h2p_hip_handle *a = hs[i], *b = hs[j];
            if (done[j] || a->mode != b->mode || a->prec != b->prec || a->kid != b->kid) continue;
            done[j] = 1;
            for (const auto &P : b->part) L.add(j, b->d_blk + P.off, P.units, P.big, P.n);
        }
        BLaunch bl = finish(tb, L);
        bl.mode = hs[i]->mode;
        bl.prec = hs[i]->prec;
        bl.kid = hs[i]->kid;
        if (bl.grid > 0) P.blocks.push_back(bl);
    }
    return P;
}

template <typename K>
static void launch_items(K kernel, const BLaunch &b, const char *dbase, const void *reqs, hipStream_t s) {
    if (b.grid <= 0) return;
    hipLaunchKernelGGL(kernel, dim3(b.grid), dim3(256), 0, s, (const BatchItem *)(dbase + b.items_off), b.n_items,
                       (const ReqCtx *)reqs);
}

static void launch_blocks_b(const BLaunch &b, const char *dbase, const ReqCtx *reqs, hipStream_t s) {
    const BatchItem *it = (const BatchItem *)(dbase + b.items_off);
    const dim3 grid(b.grid), blk(256);
    const bool f32 = b.prec == H2P_HIP_FP32MIX;
    if (b.mode == H2P_HIP_AOT) {
        if (f32) hipLaunchKernelGGL(k_blocks_aot_b<float>, grid, blk, 0, s, it, b.n_items, reqs);
        else hipLaunchKernelGGL(k_blocks_aot_b<double>, grid, blk, 0, s, it, b.n_items, reqs);
        return;
    }
    switch (b.kid) {
        case KC:
            if (f32) hipLaunchKernelGGL((k_blocks_jit2_b<KC, float, 4>), grid, blk, 0, s, it, b.n_items, reqs);
            else hipLaunchKernelGGL((k_blocks_jit2_b<KC, double, 1>), grid, blk, 0, s, it, b.n_items, reqs);
            break;
        case KG:
            if (f32) hipLaunchKernelGGL((k_blocks_jit2_b<KG, float, 4>), grid, blk, 0, s, it, b.n_items, reqs);
            else hipLaunchKernelGGL((k_blocks_jit2_b<KGW, double, 4>), grid, blk, 0, s, it, b.n_items, reqs);
            break;
        case KM:
            if (f32) hipLaunchKernelGGL((k_blocks_jit2_b<KM, float, 4>), grid, blk, 0, s, it, b.n_items, reqs);
            else hipLaunchKernelGGL((k_blocks_jit2_b<KMW, double, 4>), grid, blk, 0, s, it, b.n_items, reqs);
            break;
        default:
            if (f32) hipLaunchKernelGGL((k_blocks_jit_b<KR, 3, float>), grid, blk, 0, s, it, b.n_items, reqs);
            else hipLaunchKernelGGL((k_blocks_jit_b<KR, 3, double>), grid, blk, 0, s, it, b.n_items, reqs);
            break;
    }
}

static void launch_matvec_b(const MVPlan &P, const char *dbase, hipStream_t s) {
    const ReqCtx *reqs = (const ReqCtx *)(dbase + P.reqs_off);
    launch_items(k_gather_b, P.gather, dbase, reqs, s);
    for (const BLaunch &b : P.fwd) launch_items(k_fwd2_b, b, dbase, reqs, s);
    for (const BLaunch &b : P.blocks) launch_blocks_b(b, dbase, reqs, s);
    launch_items(k_reduce_y1_b, P.red, dbase, reqs, s);
    for (const BLaunch &b : P.bwd) launch_items(k_bwd2_b, b, dbase, reqs, s);
    launch_items(k_leaf_final2_b, P.leaf, dbase, reqs, s);
}

struct CGPlan {
    size_t cgs_off = 0, first_off = 0;
    int n = 0, grid = 0;
};

static CGPlan plan_cg(TableBuf &tb, int n, h2p_hip_handle *const *hs) {
    CGPlan P;
    std::vector<CGCtx> cgs(n);
    std::vector<int> first(n);
    int total = 0;
    for (int i = 0; i < n; i++) {
        const h2p_hip_handle *h = hs[i];
        CGCtx &c = cgs[i];
        c.x = h->cg_x;
        c.r = h->cg_r;
        c.p = h->cg_p;
        c.s = h->cg_s;
        c.b = h->cg_b;
        c.part = h->cg_part;
        c.sc = h->cg_sc;
        c.tol = h->cg_tol;
        c.replace = h->cg_replace;
        c.N = h->N;
        c.nblk = h->cg_nblk;
        c.max_iter = h->cg_max_iter;
        c.pad = 0;
        first[i] = total;
        total += h->cg_nblk;
    }
    P.cgs_off = tb.add(cgs.data(), cgs.size());
    P.first_off = tb.add(first.data(), first.size());
    P.n = n;
    P.grid = total;
    return P;
}

}  // namespace

struct h2p_hip_batch {
    static constexpr int NSLOT = 2;
    char *h_stage[NSLOT] = {nullptr, nullptr};
    char *d_stage[NSLOT] = {nullptr, nullptr};
    size_t cap[NSLOT] = {0, 0};
    hipEvent_t ev[NSLOT] = {nullptr, nullptr};
    bool ev_live[NSLOT] = {false, false};
    int slot = 0;
    CGScal *h_poll = nullptr;   // pinned read-back of solver scalars
    int poll_cap = 0;
};

static void free_batch(h2p_hip_batch *b) {
    if (!b) return;
    for (int k = 0; k < h2p_hip_batch::NSLOT; k++) {
        if (b->ev[k]) {
            (void)hipEventSynchronize(b->ev[k]);
            (void)hipEventDestroy(b->ev[k]);
        }
        if (b->h_stage[k]) (void)hipHostFree(b->h_stage[k]);
        if (b->d_stage[k]) (void)hipFree(b->d_stage[k]);
    }
    if (b->h_poll) (void)hipHostFree(b->h_poll);
    delete b;
}

// Copy the tables into the next slot and return the device base of the slot.
static int stage_tables(h2p_hip_batch *B, const TableBuf &tb, hipStream_t s, const char **dbase) {
    const int k = B->slot;
    B->slot = (B->slot + 1) % h2p_hip_batch::NSLOT;
    if (B->ev_live[k]) HIP_TRY(hipEventSynchronize(B->ev[k]));
    const size_t need = std::max<size_t>(tb.bytes.size(), 256);
    if (B->cap[k] < need) {
        const size_t c = std::max(need, 2 * B->cap[k]);
        if (B->h_stage[k]) (void)hipHostFree(B->h_stage[k]);
        if (B->d_stage[k]) (void)hipFree(B->d_stage[k]);
        B->h_stage[k] = nullptr;
        B->d_stage[k] = nullptr;
        B->cap[k] = 0;
        if (hipHostMalloc((void **)&B->h_stage[k], c, hipHostMallocDefault) != hipSuccess)
            return fail(H2P_HIP_ENOMEM, "hipHostMalloc of the batch staging failed");
        if (hipMalloc((void **)&B->d_stage[k], c) != hipSuccess)
            return fail(H2P_HIP_ENOMEM, "hipMalloc of the batch tables failed");
        B->cap[k] = c;
    }
    std::memcpy(B->h_stage[k], tb.bytes.data(), tb.bytes.size());
    HIP_TRY(hipMemcpyAsync(B->d_stage[k], B->h_stage[k], tb.bytes.size(), hipMemcpyHostToDevice, s));
    *dbase = B->d_stage[k];
    return 0;
}

static int release_slot(h2p_hip_batch *B, hipStream_t s) {
    // the slot just used is the one before B->slot
    const int k = (B->slot + h2p_hip_batch::NSLOT - 1) % h2p_hip_batch::NSLOT;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(B->ev[k], s));
    B->ev_live[k] = true;
    return 0;
}

static int check_handles(int n, h2p_hip_handle *const *hs) {
    if (n < 0) return fail(H2P_HIP_EINVAL, "n < 0");
    if (n > 0 && !hs) return fail(H2P_HIP_EINVAL, "NULL handle array");
    for (int i = 0; i < n; i++) {
        if (!hs[i]) return fail(H2P_HIP_EINVAL, "NULL handle");
        if (hs[i]->periodic) return fail(H2P_HIP_ENOSYS, "periodic handles are not batched");
        for (int j = 0; j < i; j++)
            if (hs[j] == hs[i]) return fail(H2P_HIP_EINVAL, "a handle appears twice in one batch");
    }
    return 0;
}

extern "C" {

int h2p_hip_batch_create(h2p_hip_batch **out) {
    if (!out) return fail(H2P_HIP_EINVAL, "out is NULL");
    *out = nullptr;
    h2p_hip_batch *b = new h2p_hip_batch();
    for (int k = 0; k < h2p_hip_batch::NSLOT; k++)
        if (hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming) != hipSuccess) {
            free_batch(b);
            return fail(H2P_HIP_EDEVICE, "hipEventCreate failed");
        }
    *out = b;
    return 0;
}

void h2p_hip_batch_free(h2p_hip_batch *b) { free_batch(b); }

int h2p_hip_matvec_batch(h2p_hip_batch *B, int n, h2p_hip_handle *const *hs, const double *const *d_x,
                         double *const *d_y, const double *shift, void *stream) {
    if (!B) return fail(H2P_HIP_EINVAL, "NULL batch");
    int rc = check_handles(n, hs);
    if (rc) return rc;
    if (n == 0) return 0;
    if (!d_x || !d_y) return fail(H2P_HIP_EINVAL, "NULL vector array");
    for (int i = 0; i < n; i++)
        if (!d_x[i] || !d_y[i]) return fail(H2P_HIP_EINVAL, "NULL vector");
    hipStream_t s = (hipStream_t)stream;
    TableBuf tb;
    const MVPlan P = plan_matvec(tb, n, hs, d_x, d_y, shift);
    const char *dbase = nullptr;
    if ((rc = stage_tables(B, tb, s, &dbase))) return rc;
    launch_matvec_b(P, dbase, s);
    return release_slot(B, s);
}

int h2p_hip_matmul_batch(h2p_hip_batch *B, int n, h2p_hip_handle *const *hs, int layout, const int *n_vec,
                         const double *const *d_X, const int *ldx, double *const *d_Y, const int *ldy,
                         void *stream) {
    if (!B) return fail(H2P_HIP_EINVAL, "NULL batch");
    int rc = check_handles(n, hs);
    if (rc) return rc;
    if (n == 0) return 0;
    if (!n_vec || !d_X || !ldx || !d_Y || !ldy) return fail(H2P_HIP_EINVAL, "NULL argument array");
    const bool rm = layout == H2P_HIP_ROW_MAJOR;
    if (!rm && layout != H2P_HIP_COL_MAJOR) return fail(H2P_HIP_EINVAL, "bad layout");
    int max_nv = 0;
    for (int i = 0; i < n; i++) {
        const h2p_hip_handle *h = hs[i];
        if (n_vec[i] < 0) return fail(H2P_HIP_EINVAL, "n_vec < 0");
        if (n_vec[i] > 0 && (!d_X[i] || !d_Y[i])) return fail(H2P_HIP_EINVAL, "NULL panel");
        if (n_vec[i] > 0 && (rm ? (ldx[i] < n_vec[i] || ldy[i] < n_vec[i]) : (ldx[i] < h->N || ldy[i] < h->N)))
            return fail(H2P_HIP_EINVAL, "bad leading dimension");
        if (n_vec[i] > 0 && !h->mm_ok) return fail(H2P_HIP_ENOSYS, "FP32-mixed matmul is available for the scalar kernels in JIT mode and FP32-stored AOT blocks of them");
        max_nv = std::max(max_nv, n_vec[i]);
    }
    for (int i = 0; i < n; i++)
        if (n_vec[i] > 0 && (rc = ensure_panels(hs[i], panel_width(n_vec[i])))) return rc;
    hipStream_t s = (hipStream_t)stream;
    for (int v0 = 0; v0 < max_nv; v0 += H2P_MM_MAXNVP) {
        // requests with columns left in this chunk
        std::vector<h2p_hip_handle *> act;
        std::vector<MMCtx> ctx;
        for (int i = 0; i < n; i++) {
            if (n_vec[i] <= v0) continue;
            h2p_hip_handle *h = hs[i];
            MMCtx c{};
            c.nv = std::min(H2P_MM_MAXNVP, n_vec[i] - v0);
            c.srcs = h->d_mms;
            c.shifts = h->d_shift;
            c.U = h->d_U;
            c.UT = h->d_UT;
            c.BD = (const double *)h->d_blob;
            c.coord = h->d_coord;
            c.CT = h->CT;
            c.kp = h->kp;
            c.P = panels_of(h, (c.nv + 15) & ~15, h->d_panels);
            c.fwd = h->d_fwd;
            c.N = h->N;
            c.X = d_X[i];
            c.Y = d_Y[i];
            c.ldx = ldx[i];
            c.ldy = ldy[i];
            c.row_major = rm ? 1 : 0;
            c.v0 = v0;
            act.push_back(h);
            ctx.push_back(c);
        }
        const int m = (int)act.size();
        TableBuf tb;
        const size_t ctx_off = tb.add(ctx.data(), ctx.size());
        ItemList g, sc;
        size_t max_fwd = 0, max_bwd = 0;
        for (int i = 0; i < m; i++) {
            const long long ne = (long long)act[i]->N * ctx[i].P.nvp, no = (long long)act[i]->N * ctx[i].nv;
            g.add(i, nullptr, (int)((ne + 255) / 256));
            sc.add(i, nullptr, (int)((no + 255) / 256));
            max_fwd = std::max(max_fwd, act[i]->mm_fwd_levels.size());
            max_bwd = std::max(max_bwd, act[i]->mm_bwd_levels.size());
        }
        const BLaunch lg = finish(tb, g), ls = finish(tb, sc);
        // product phases, grouped by kernel id: (level list, launch)
        // (FP32-mixed handles run their coupling phase and the B/D part of
        // their leaf outputs in the FP32 kernel: grouped apart)
        struct MMLaunch { BLaunch b; int kid; bool f32; };
        std::vector<MMLaunch> phases;
        auto add_phase = [&](auto level_of, bool coupling) {
            std::vector<int> used(m, 0);
            for (int i = 0; i < m; i++) {
                if (used[i]) continue;
                const bool f32 = coupling && act[i]->mm_f32;
                ItemList L;
                for (int j = i; j < m; j++) {
                    if (used[j] || act[j]->kid != act[i]->kid || (coupling && act[j]->mm_f32) != f32) continue;
                    used[j] = 1;
                    const Level *lv = level_of(act[j]);
                    if (lv) L.add(j, act[j]->d_mmt + lv->off, lv->n);
                }
                BLaunch b = finish(tb, L);
                if (b.grid > 0) phases.push_back({b, act[i]->kid, f32});
            }
        };
        for (size_t k = 0; k < max_fwd; k++)
            add_phase([&](h2p_hip_handle *h) { return k < h->mm_fwd_levels.size() ? &h->mm_fwd_levels[k] : (const Level *)nullptr; }, false);
        add_phase([&](h2p_hip_handle *h) { return (const Level *)&h->mm_intmd; }, true);
        for (size_t k = 0; k < max_bwd; k++)
            add_phase([&](h2p_hip_handle *h) { return k < h->mm_bwd_levels.size() ? &h->mm_bwd_levels[k] : (const Level *)nullptr; }, false);
        add_phase([&](h2p_hip_handle *h) { return (const Level *)&h->mm_leaf; }, false);
        add_phase([&](h2p_hip_handle *h) { return h->mm_f32 ? (const Level *)&h->mm_leaf32 : (const Level *)nullptr; }, true);
        const char *dbase = nullptr;
        if ((rc = stage_tables(B, tb, s, &dbase))) return rc;
        const MMCtx *dctx = (const MMCtx *)(dbase + ctx_off);
        if (lg.grid > 0)
            hipLaunchKernelGGL(k_mm_gather_b, dim3(lg.grid), dim3(256), 0, s, (const BatchItem *)(dbase + lg.items_off),
                               lg.n_items, dctx);
        int max_nvp = 16;
        for (const MMCtx &c : ctx) max_nvp = std::max(max_nvp, c.P.nvp);
        const int ncls = nv_class(max_nvp);
        for (const MMLaunch &ph : phases) {
            const BatchItem *it = (const BatchItem *)(dbase + ph.b.items_off);
            const dim3 grid(ph.b.grid), blk(256);
            if (ph.f32) {
                switch (ph.kid * 3 + ncls) {
                    case KC * 3 + 0: hipLaunchKernelGGL((k_mm32_b<KC, MMVar<1, 16, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                    case KC * 3 + 1: hipLaunchKernelGGL((k_mm32_b<KC, MMVar<1, 32, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                    case KC * 3 + 2: hipLaunchKernelGGL((k_mm32_b<KC, MMVar<1, 64, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                    case KG * 3 + 0: hipLaunchKernelGGL((k_mm32_b<KG, MMVar<1, 16, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                    case KG * 3 + 1: hipLaunchKernelGGL((k_mm32_b<KG, MMVar<1, 32, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                    case KG * 3 + 2: hipLaunchKernelGGL((k_mm32_b<KG, MMVar<1, 64, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                    case KM * 3 + 0: hipLaunchKernelGGL((k_mm32_b<KM, MMVar<1, 16, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                    case KM * 3 + 1: hipLaunchKernelGGL((k_mm32_b<KM, MMVar<1, 32, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                    default: hipLaunchKernelGGL((k_mm32_b<KM, MMVar<1, 64, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                }
                continue;
            }
            switch (ph.kid * 3 + ncls) {
                case KC * 3 + 0: hipLaunchKernelGGL((k_mm_b<KC, 1, MMVar<1, 16, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KC * 3 + 1: hipLaunchKernelGGL((k_mm_b<KC, 1, MMVar<1, 32, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KC * 3 + 2: hipLaunchKernelGGL((k_mm_b<KC, 1, MMVar<1, 64, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KG * 3 + 0: hipLaunchKernelGGL((k_mm_b<KG, 1, MMVar<1, 16, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KG * 3 + 1: hipLaunchKernelGGL((k_mm_b<KG, 1, MMVar<1, 32, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KG * 3 + 2: hipLaunchKernelGGL((k_mm_b<KG, 1, MMVar<1, 64, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KM * 3 + 0: hipLaunchKernelGGL((k_mm_b<KM, 1, MMVar<1, 16, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KM * 3 + 1: hipLaunchKernelGGL((k_mm_b<KM, 1, MMVar<1, 32, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KM * 3 + 2: hipLaunchKernelGGL((k_mm_b<KM, 1, MMVar<1, 64, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KR * 3 + 0: hipLaunchKernelGGL((k_mm_b<KR, 3, MMVar<3, 16, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                case KR * 3 + 1: hipLaunchKernelGGL((k_mm_b<KR, 3, MMVar<3, 32, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
                default: hipLaunchKernelGGL((k_mm_b<KR, 3, MMVar<3, 64, 0>>), grid, blk, 0, s, it, ph.b.n_items, dctx); break;
            }
        }
        if (ls.grid > 0)
            hipLaunchKernelGGL(k_mm_scatter_b, dim3(ls.grid), dim3(256), 0, s, (const BatchItem *)(dbase + ls.items_off),
                               ls.n_items, dctx);
        if ((rc = release_slot(B, s))) return rc;
    }
    for (int i = 0; i < n; i++)
        if (n_vec[i] > 0) mark_use(hs[i], s);   // their panels are in use on s
    return 0;
}

int h2p_hip_cg_start(h2p_hip_batch *B, int n, h2p_hip_handle *const *hs, const double *const *d_b,
                     double *const *d_x, const double *shift, const double *tol, const int *max_iter,
                     void *stream) {
    if (!B) return fail(H2P_HIP_EINVAL, "NULL batch");
    int rc = check_handles(n, hs);
    if (rc) return rc;
    if (n == 0) return 0;
    if (!d_b || !d_x || !tol || !max_iter) return fail(H2P_HIP_EINVAL, "NULL argument array");
    for (int i = 0; i < n; i++) {
        if (!d_b[i] || !d_x[i]) return fail(H2P_HIP_EINVAL, "NULL vector");
        if (!(tol[i] >= 0.0)) return fail(H2P_HIP_EINVAL, "tol must be >= 0");
    }
    for (int i = 0; i < n; i++) {
        h2p_hip_handle *h = hs[i];
        if (!h->d_cg) {
            const size_t nd = cg_doubles(h->N);
            if (hipMalloc(&h->d_cg, sizeof(double) * nd) != hipSuccess)
                return fail(H2P_HIP_ENOMEM, "hipMalloc of the CG state failed");
            h->device_bytes += sizeof(double) * nd;
            h->info.device_bytes = h->device_bytes;
            h->cg_nblk = cg_blocks(h->N);
            h->cg_r = h->d_cg;
            h->cg_p = h->cg_r + h->N;
            h->cg_s = h->cg_p + h->N;
            h->cg_part = h->cg_s + h->N;
            h->cg_sc = (CGScal *)(h->cg_part + 2 * (size_t)h->cg_nblk);
        }
        h->cg_x = d_x[i];
        h->cg_b = d_b[i];
        h->cg_shift = shift ? shift[i] : 0.0;
        h->cg_tol = tol[i];
        h->cg_max_iter = max_iter[i];
        h->cg_started = 1;
    }
    hipStream_t s = (hipStream_t)stream;
    TableBuf tb;
    std::vector<const double *> xs(n);
    std::vector<double *> ys(n);
    std::vector<double> sh(n);
    for (int i = 0; i < n; i++) {
        xs[i] = hs[i]->cg_x;
        ys[i] = hs[i]->cg_s;
        sh[i] = hs[i]->cg_shift;
    }
    const MVPlan P = plan_matvec(tb, n, hs, xs.data(), ys.data(), sh.data());
    const CGPlan C = plan_cg(tb, n, hs);
    const char *dbase = nullptr;
    if ((rc = stage_tables(B, tb, s, &dbase))) return rc;
    const int *first = (const int *)(dbase + C.first_off);
    const CGCtx *cgs = (const CGCtx *)(dbase + C.cgs_off);
    launch_matvec_b(P, dbase, s);
    hipLaunchKernelGGL(k_cg_init, dim3(C.grid), dim3(256), 0, s, first, C.n, cgs);
    hipLaunchKernelGGL(k_cg_init_finish, dim3(C.n), dim3(64), 0, s, C.n, cgs);
    return release_slot(B, s);
}

int h2p_hip_cg_step(h2p_hip_batch *B, int n, h2p_hip_handle *const *hs, int n_iter, void *stream) {
    if (!B) return fail(H2P_HIP_EINVAL, "NULL batch");
    int rc = check_handles(n, hs);
    if (rc) return rc;
    if (n_iter < 0) return fail(H2P_HIP_EINVAL, "n_iter < 0");
    if (n == 0 || n_iter == 0) return 0;
    for (int i = 0; i < n; i++)
        if (!hs[i]->cg_started) return fail(H2P_HIP_EINVAL, "CG step on a handle without h2p_hip_cg_start");
    hipStream_t s = (hipStream_t)stream;
    TableBuf tb;
    std::vector<const double *> xs(n);
    std::vector<double *> ys(n);
    std::vector<double> sh(n);
    for (int i = 0; i < n; i++) {
        xs[i] = hs[i]->cg_p;
        ys[i] = hs[i]->cg_s;
        sh[i] = hs[i]->cg_shift;
    }
    const MVPlan P = plan_matvec(tb, n, hs, xs.data(), ys.data(), sh.data());
    const CGPlan C = plan_cg(tb, n, hs);
    const char *dbase = nullptr;
    if ((rc = stage_tables(B, tb, s, &dbase))) return rc;
    const int *first = (const int *)(dbase + C.first_off);
    const CGCtx *cgs = (const CGCtx *)(dbase + C.cgs_off);
    for (int it = 0; it < n_iter; it++) {
        hipLaunchKernelGGL(k_cg_p, dim3(C.grid), dim3(256), 0, s, first, C.n, cgs);
        launch_matvec_b(P, dbase, s);
        hipLaunchKernelGGL(k_cg_dot_ps, dim3(C.grid), dim3(256), 0, s, first, C.n, cgs);
        hipLaunchKernelGGL(k_cg_alpha, dim3(C.n), dim3(64), 0, s, C.n, cgs);
        hipLaunchKernelGGL(k_cg_update, dim3(C.grid), dim3(256), 0, s, first, C.n, cgs);
        hipLaunchKernelGGL(k_cg_finish, dim3(C.n), dim3(64), 0, s, C.n, cgs);
    }
    return release_slot(B, s);
}

int h2p_hip_cg_set_replace(h2p_hip_handle *h, double ratio) {
    if (!h) return fail(H2P_HIP_EINVAL, "NULL handle");
    if (!(ratio >= 0.0 && ratio < 1.0)) return fail(H2P_HIP_EINVAL, "the replacement ratio must be in [0, 1)");
    h->cg_replace = ratio * ratio;
    return 0;
}

int h2p_hip_cg_replace(h2p_hip_batch *B, int n, h2p_hip_handle *const *hs, h2p_hip_handle *const *hs_acc,
                       void *stream) {
    if (!B) return fail(H2P_HIP_EINVAL, "NULL batch");
    int rc = check_handles(n, hs);
    if (rc) return rc;
    if ((rc = check_handles(n, hs_acc))) return rc;
    if (n == 0) return 0;
    for (int i = 0; i < n; i++) {
        if (!hs[i]->cg_started) return fail(H2P_HIP_EINVAL, "CG replacement on a handle without h2p_hip_cg_start");
        if (hs_acc[i]->N != hs[i]->N) return fail(H2P_HIP_EINVAL, "the product's handle has another order");
    }
    hipStream_t s = (hipStream_t)stream;
    TableBuf tb;
    std::vector<const double *> xs(n);
    std::vector<double *> ys(n);
    std::vector<double> sh(n);
    for (int i = 0; i < n; i++) {
        xs[i] = hs[i]->cg_x;
        ys[i] = hs[i]->cg_s;
        sh[i] = hs[i]->cg_shift;
    }
    // s = (A + shift) x by the accurate handles, then r = b - s and the stop test
    const MVPlan P = plan_matvec(tb, n, hs_acc, xs.data(), ys.data(), sh.data());
    const CGPlan C = plan_cg(tb, n, hs);
    const char *dbase = nullptr;
    if ((rc = stage_tables(B, tb, s, &dbase))) return rc;
    const int *first = (const int *)(dbase + C.first_off);
    const CGCtx *cgs = (const CGCtx *)(dbase + C.cgs_off);
    launch_matvec_b(P, dbase, s);
    hipLaunchKernelGGL(k_cg_init, dim3(C.grid), dim3(256), 0, s, first, C.n, cgs);
    hipLaunchKernelGGL(k_cg_replace_finish, dim3(C.n), dim3(64), 0, s, C.n, cgs);
    return release_slot(B, s);
}

int h2p_hip_cg_poll(h2p_hip_batch *B, int n, h2p_hip_handle *const *hs, int *done, int *iter, double *relres,
                    void *stream) {
    if (!B) return fail(H2P_HIP_EINVAL, "NULL batch");
    int rc = check_handles(n, hs);
    if (rc) return rc;
    if (n == 0) return 0;
    for (int i = 0; i < n; i++)
        if (!hs[i]->cg_started) return fail(H2P_HIP_EINVAL, "CG poll on a handle without h2p_hip_cg_start");
    hipStream_t s = (hipStream_t)stream;
    if (B->poll_cap < n) {
        if (B->h_poll) (void)hipHostFree(B->h_poll);
        B->h_poll = nullptr;
        B->poll_cap = 0;
        if (hipHostMalloc((void **)&B->h_poll, sizeof(CGScal) * n, hipHostMallocDefault) != hipSuccess)
            return fail(H2P_HIP_ENOMEM, "hipHostMalloc failed");
        B->poll_cap = n;
    }
    for (int i = 0; i < n; i++)
        HIP_TRY(hipMemcpyAsync(B->h_poll + i, hs[i]->cg_sc, sizeof(CGScal), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) {
        const CGScal &c = B->h_poll[i];
        if (done) done[i] = c.done;
        if (iter) iter[i] = c.iter;
        if (relres) relres[i] = std::sqrt(c.rr) / std::sqrt(c.bb);   // 0 / 0 = NaN for b = 0, as pcg.c:126
    }
    return 0;
}

}  // extern "C"
