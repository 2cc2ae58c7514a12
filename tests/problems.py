"""Seeded H2 problems shared by the tests (built once per session)."""
from __future__ import annotations

import functools

import numpy as np

from h2pack_amd import construct as C

# (kernel id, params, N, density, leaf, reltol, pp kind)
SPECS = {
    "coulomb": (C.KRNL_COULOMB, [], 6000, 1.0, 150, 1e-6, "numerical"),
    "coulomb_diag": (C.KRNL_COULOMB, [2.5], 3000, 1.0, 100, 1e-6, "surface"),
    "gaussian": (C.KRNL_GAUSSIAN, [1.0], 6000, 10.0, 150, 1e-6, "numerical"),
    "matern32": (C.KRNL_MATERN32, [1.0], 6000, 10.0, 150, 1e-6, "numerical"),
    # leaves of up to 300 points: blocks wider than one 128-column pass and
    # taller than one 256-row chunk
    "gaussian_leaf300": (C.KRNL_GAUSSIAN, [1.0], 17000, 10.0, 300, 1e-6, "numerical"),
    "matern32_leaf300": (C.KRNL_MATERN32, [1.0], 17000, 10.0, 300, 1e-6, "numerical"),
    "rpy_leaf300": (C.KRNL_RPY, [1.0], 17000, 0.03, 300, 1e-7, "surface"),
    "rpy": (C.KRNL_RPY, [1.0], 2500, 0.03, 120, 1e-7, "surface"),
    # deeper trees (leaves on two levels, admissible pairs on three or more
    # levels): the matvec's level split has levels above the leaf level
    "gaussian_deep": (C.KRNL_GAUSSIAN, [1.0], 60000, 10.0, 100, 1e-6, "surface"),
    "coulomb_deep": (C.KRNL_COULOMB, [], 60000, 1.0, 100, 1e-6, "surface"),
    # tiny trees for dense assembly
    "tiny_coulomb": (C.KRNL_COULOMB, [], 900, 1.0, 40, 1e-8, "surface"),
    "tiny_gaussian": (C.KRNL_GAUSSIAN, [0.5], 900, 4.0, 40, 1e-8, "numerical"),
    "tiny_rpy": (C.KRNL_RPY, [1.0], 500, 0.03, 30, 1e-8, "surface"),
    # HSS mode (weak admissibility, hybrid bases)
    "hss_coulomb": (C.KRNL_COULOMB, [], 3000, 1.0, 100, 1e-6, "numerical", {"hss": True}),
    "hss_gaussian": (C.KRNL_GAUSSIAN, [1.0], 3000, 10.0, 100, 1e-6, "numerical", {"hss": True}),
    "tiny_hss": (C.KRNL_GAUSSIAN, [0.5], 900, 4.0, 40, 1e-8, "numerical", {"hss": True}),
}


def coords_for(kid, n, density, seed):
    if kid == C.KRNL_RPY:
        # radius 1 particles at `density` volume fraction; cube side from it
        L = (n * (4.0 / 3.0) * np.pi / density) ** (1.0 / 3.0)
        rng = np.random.default_rng(seed)
        c = rng.random((3, n)) * L
        return np.concatenate([c, np.ones((1, n))], axis=0)
    return C.uniform_cube_points(n, density, seed)


@functools.lru_cache(maxsize=None)
def problem(name: str, seed: int = 1234):
    kid, param, n, dens, leaf, tol, ppk, *extra = SPECS[name]
    coord = coords_for(kid, n, dens, seed)
    h2 = C.build_h2(coord, kid, param, tol, leaf, BD_JIT=0, pp_kind=ppk, host_blocks=True,
                    **(extra[0] if extra else {}))
    return h2, coord


@functools.lru_cache(maxsize=None)
def periodic_problem(n: int = 500, L: float = 14.0, seed: int = 2):
    """Periodic RPY (radius 0.5, eta = 1/(6 pi) so that the plain and Ewald
    kernels share their scale) in a cubic cell of side L; Ewald parameters
    xi = 0.3, 6 real-space and reciprocal layers; reltol 1e-6, leaf 60."""
    rng = np.random.default_rng(seed)
    coord = np.concatenate([rng.random((3, n)) * L, np.full((1, n), 0.5)], axis=0)
    ewald = (L, 0.3, 6, 6)
    h2 = C.build_periodic(coord, [0.0, 0.0, 0.0, L, L, L], 1.0 / (6.0 * np.pi), ewald, 1e-6, 60)
    return h2, coord, ewald


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def dense_h2(h2) -> np.ndarray:
    """Assemble the H2 matrix explicitly (original point order) by applying
    the representation to every unit vector through the block formulas:
    A = P^T (D + sum over admissible pairs of  U-nested bases * B * ...) P.
    Built column by column from independent numpy products (no oracle)."""
    N = h2.krnl_mat_size
    kd = h2.krnl_dim
    mc = h2.mat_cluster.reshape(-1, 2)
    nn = h2.n_node
    ch = h2.children.reshape(nn, -1)
    # explicit nested basis V[node]: (points of node * kd) x r, V = U for leaves,
    # V = blockdiag(V[children]) U for non-leaves
    V = [None] * nn
    for lv in range(h2.max_level, h2.min_adm_level - 1, -1):
        for n in h2.level_list(lv):
            U = h2.U[n]
            if U.shape[1] == 0:
                V[n] = np.zeros(((mc[n, 1] - mc[n, 0] + 1), 0))
                continue
            if h2.n_child[n] == 0:
                V[n] = U
            else:
                blocks, off = [], 0
                for c in ch[n, :h2.n_child[n]]:
                    rc = h2.U[c].shape[1]
                    blocks.append(V[c] @ U[off:off + rc])
                    off += rc
                V[n] = np.concatenate(blocks, axis=0)
    A = np.zeros((N, N))
    for i in range(h2.n_r_adm_pair):
        n0, n1 = h2.r_adm_pairs[2 * i], h2.r_adm_pairs[2 * i + 1]
        l0, l1 = h2.node_level[n0], h2.node_level[n1]
        c0, c1 = C.B_block_coords(h2, i)
        B = C.kernel_eval(h2.krnl_id, h2.krnl_param, c0, c1)
        L = V[n0] if l0 >= l1 else np.eye(mc[n0, 1] - mc[n0, 0] + 1)
        R = V[n1] if l0 <= l1 else np.eye(mc[n1, 1] - mc[n1, 0] + 1)
        blk = L @ B @ R.T
        A[mc[n0, 0]:mc[n0, 1] + 1, mc[n1, 0]:mc[n1, 1] + 1] += blk
        A[mc[n1, 0]:mc[n1, 1] + 1, mc[n0, 0]:mc[n0, 1] + 1] += blk.T
    for i in range(h2.n_D):
        c0, c1 = C.D_block_coords(h2, i)
        D = C.kernel_eval(h2.krnl_id, h2.krnl_param, c0, c1)
        if i < h2.n_leaf_node:
            n = h2.leaf_nodes[i]
            A[mc[n, 0]:mc[n, 1] + 1, mc[n, 0]:mc[n, 1] + 1] += D
        else:
            k = i - h2.n_leaf_node
            n0, n1 = h2.r_inadm_pairs[2 * k], h2.r_inadm_pairs[2 * k + 1]
            A[mc[n0, 0]:mc[n0, 1] + 1, mc[n1, 0]:mc[n1, 1] + 1] += D
            A[mc[n1, 0]:mc[n1, 1] + 1, mc[n0, 0]:mc[n0, 1] + 1] += D.T
    # sorted order -> original order: x_sorted = x[fwd], y = y_sorted[bwd]
    fwd = h2.fwd_pmt_idx
    Ao = np.empty_like(A)
    Ao[np.ix_(fwd, fwd)] = A
    return Ao
