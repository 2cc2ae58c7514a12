/*
 * Host-side kernel block evaluation used by the H2 construction
 * (h2pack_amd/construct.py).  It plays the role of the reference's
 * krnl_eval plugins during construction: src/H2Pack_3D_kernels.h
 * Coulomb :52-96, Gaussian :169-209, Matern32 :388-432, RPY :893-959.
 *
 * Layout (as krnl_eval): coordinates are xpt_dim rows of length ld, one point
 * per column; the output block is row-major with kd x kd sub-blocks per pair.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#define KRNL_COULOMB  0
#define KRNL_GAUSSIAN 1
#define KRNL_MATERN32 2
#define KRNL_RPY      3

/* Rotne-Prager-Yamakawa mobility of two spheres of radii ai, aj at centre
 * distance r, as K = f I + g u u^T (u = unit separation) in units of
 * 1/(6 pi eta), for all three configurations (Zuk, Wajnryb, Mizerski,
 * Szymczak, J. Fluid Mech. 741, R5, 2014):
 *   one sphere inside the other (r <= |ai - aj|): the larger sphere's
 *     self-mobility, f = 1 / max(ai, aj), g = 0;
 *   partial overlap (|ai - aj| < r <= ai + aj):
 *     f = (16 r^3 (ai + aj) - ((ai - aj)^2 + 3 r^2)^2) / (32 r^3 ai aj),
 *     g = 3 ((ai - aj)^2 - r^2)^2 / (32 r^3 ai aj);
 *   disjoint spheres (r > ai + aj), with b = (ai^2 + aj^2) / r^2:
 *     f = 3 / (4 r) (1 + b / 3),  g = 3 / (4 r) (1 - b).
 * The reference kernel is src/H2Pack_3D_kernels.h:893-959. */
static void rpy_far(double r, double ai, double aj, double *f, double *g)
{
    const double b = (ai * ai + aj * aj) / (r * r), c = 0.75 / r;
    *f = c * (1.0 + b / 3.0);
    *g = c * (1.0 - b);
}

static void rpy_mobility(double r, double ai, double aj, double *f, double *g)
{
    const double dif = fabs(ai - aj), sum = ai + aj;
    if (r <= dif) {
        *f = 1.0 / (ai > aj ? ai : aj);
        *g = 0.0;
    } else if (r <= sum) {
        const double r2 = r * r, d2 = dif * dif, w = 1.0 / (32.0 * r2 * r * ai * aj);
        const double h = d2 + 3.0 * r2, e = d2 - r2;
        *f = w * (16.0 * r2 * r * sum - h * h);
        *g = w * 3.0 * e * e;
    } else {
        rpy_far(r, ai, aj, f, g);
    }
}

/* out (3x3, row stride ldm) = f I + g u u^T with u = d * inv_r */
static void put_fg(double *out, long ldm, double f, double g, const double *d, double inv_r)
{
    const double u[3] = {d[0] * inv_r, d[1] * inv_r, d[2] * inv_r};
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) out[a * ldm + b] = g * u[a] * u[b] + (a == b ? f : 0.0);
}

static void eval_rows(int kid, const double *param,
                      const double *c0, int ld0, int i0, int i1,
                      const double *c1, int ld1, int n1,
                      double *mat, long ldm)
{
    const double *x0 = c0, *y0 = c0 + ld0, *z0 = c0 + 2 * (long)ld0;
    const double *x1 = c1, *y1 = c1 + ld1, *z1 = c1 + 2 * (long)ld1;
    if (kid == KRNL_COULOMB) {
        const double dv = param ? param[0] : 0.0;
        for (int i = i0; i < i1; i++) {
            double *row = mat + (long)i * ldm;
            const double xi = x0[i], yi = y0[i], zi = z0[i];
            #pragma omp simd
            for (int j = 0; j < n1; j++) {
                double dx = xi - x1[j], dy = yi - y1[j], dz = zi - z1[j];
                double r2 = dx * dx + dy * dy + dz * dz;
                row[j] = (r2 == 0.0) ? dv : 1.0 / sqrt(r2);
            }
        }
    } else if (kid == KRNL_GAUSSIAN) {
        const double nl = -param[0];
        for (int i = i0; i < i1; i++) {
            double *row = mat + (long)i * ldm;
            const double xi = x0[i], yi = y0[i], zi = z0[i];
            #pragma omp simd
            for (int j = 0; j < n1; j++) {
                double dx = xi - x1[j], dy = yi - y1[j], dz = zi - z1[j];
                row[j] = exp(nl * (dx * dx + dy * dy + dz * dz));
            }
        }
    } else if (kid == KRNL_MATERN32) {
        const double s = -1.7320508075688772 * param[0];
        for (int i = i0; i < i1; i++) {
            double *row = mat + (long)i * ldm;
            const double xi = x0[i], yi = y0[i], zi = z0[i];
            #pragma omp simd
            for (int j = 0; j < n1; j++) {
                double dx = xi - x1[j], dy = yi - y1[j], dz = zi - z1[j];
                double t = s * sqrt(dx * dx + dy * dy + dz * dz);
                row[j] = (1.0 - t) * exp(t);
            }
        }
    } else {  /* RPY, 3x3 blocks: C (f I + g u u^T) */
        const double *a0 = c0 + 3 * (long)ld0, *a1 = c1 + 3 * (long)ld1;
        const double C = 1.0 / (6.0 * M_PI * param[0]);
        for (int i = i0; i < i1; i++) {
            double *out = mat + 3L * i * ldm;
            for (int j = 0; j < n1; j++) {
                const double d[3] = {x0[i] - x1[j], y0[i] - y1[j], z0[i] - z1[j]};
                const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                double f, g;
                rpy_mobility(r, a0[i], a1[j], &f, &g);
                put_fg(out + 3 * j, ldm, C * f, C * g, d, r > 0.0 ? 1.0 / r : 0.0);
            }
        }
    }
}

/* mat (n0*kd x n1*kd, leading dimension ldm) = K(coord0, coord1). */
void h2p_host_kernel_eval(int kid, const double *param,
                          const double *c0, int ld0, int n0,
                          const double *c1, int ld1, int n1,
                          double *mat, long ldm, int n_thread)
{
    if (n_thread <= 1 || (long)n0 * n1 < 20000) {
        eval_rows(kid, param, c0, ld0, 0, n0, c1, ld1, n1, mat, ldm);
        return;
    }
    #pragma omp parallel num_threads(n_thread)
    {
        #pragma omp for schedule(static)
        for (int b = 0; b < n0; b += 16) {
            int e = b + 16 < n0 ? b + 16 : n0;
            eval_rows(kid, param, c0, ld0, b, e, c1, ld1, n1, mat, ldm);
        }
    }
}

/* Sketched kernel block for the numerical proxy points (kd = 1): with
 * A^T = K(Y, X) (ny x nx) and a sparse sketch S^T whose row c holds nnz
 * entries (columns rows[c*nnz + t], values vals[c*nnz + t]), returns
 * out = (S^T A^T)^T, nx x nx row-major:
 *     out[j][c] = sum_t vals[c*nnz+t] * K(Y[rows[c*nnz+t]], X[j]),
 * summed in t order.  Computes the sketch without forming K(Y, X). */
#define SKETCH_CHUNK 2048
void h2p_host_sketch(int kid, const double *param,
                     const double *cy, int ldy, const double *cx, int ldx, int nx,
                     const int *rows, const double *vals, int nnz, double *out, int n_thread)
{
    #pragma omp parallel for schedule(static) num_threads(n_thread > 0 ? n_thread : 1)
    for (int c = 0; c < nx; c++) {
        double col[SKETCH_CHUNK], kr[SKETCH_CHUNK];
        for (int j0 = 0; j0 < nx; j0 += SKETCH_CHUNK) {
            const int n1 = (j0 + SKETCH_CHUNK < nx ? j0 + SKETCH_CHUNK : nx) - j0;
            for (int j = 0; j < n1; j++) col[j] = 0.0;
            for (int t = 0; t < nnz; t++) {
                const int r = rows[(long)c * nnz + t];
                const double v = vals[(long)c * nnz + t];
                /* K(Y[r], X[j0 .. j0+n1)) as one row (leading dimension 0) */
                eval_rows(kid, param, cy, ldy, r, r + 1, cx + j0, ldx, n1, kr - 0, 0);
                for (int j = 0; j < n1; j++) col[j] += v * kr[j];
            }
            for (int j = 0; j < n1; j++) out[(long)(j0 + j) * nx + c] = col[j];
        }
    }
}

/* Periodic RPY kernel by Ewald summation: the periodic system kernel of
 * H2P_build_periodic (reference: RPY_Ewald_eval_std, src/H2Pack_3D_kernels.h:
 * 961-1210), here derived from Beenakker, "Ewald sum of the Rotne-Prager
 * tensor", J. Chem. Phys. 85, 1581 (1986), with his a^2 replaced by the
 * pair's a3 = (ai^2 + aj^2) / 2 for unequal radii.  Units of 1/(6 pi eta)
 * (no such factor, as in the reference), cubic cell of side L, splitting
 * parameter xi:
 *
 *   M(d) = sum_{n in Z^3, |n_i| <= nr, d + nL != 0}  Mr(d + nL)
 *        + sum_{h in Z^3, h != 0, |h_i| <= nk}  Mk(k = 2 pi h / L) cos(k . d)
 *        + the self term when the two points coincide (up to a lattice
 *          vector), or else the overlap correction of the nearest image.
 *
 * Real space, with s = xi^2 r^2 and alpha = xi^2 a3 (both dimensionless):
 *   Mr(r) = A I + B u u^T,
 *   A = erfc(xi r) 3/(4r) (1 + 2 a3 / (3 r^2))
 *     + xi e^-s / sqrt(pi) (alpha / s + 14 alpha - 9/2 + (3 - 20 alpha) s + 4 alpha s^2),
 *   B = erfc(xi r) 3/(4r) (1 - 2 a3 / r^2)
 *     + xi e^-s / sqrt(pi) (-3 alpha / s + 3/2 - 2 alpha + (16 alpha - 3) s - 4 alpha s^2).
 * Reciprocal space, y = k^2 / (4 xi^2), V = L^3:
 *   Mk(k) = (6 pi / (V k^2)) (1 + y + 2 y^2) e^-y (1 - a3 k^2 / 3) (I - k^ k^T);
 *   h and -h give the same term, so only the half-space h > 0
 *   (lexicographically) is summed, with weight 2.
 * Self term (coincident points, a = ai = aj):
 *   (1/a - (6 - 40/3 xi^2 a^2) xi / sqrt(pi)) I.
 * Overlap correction: the Ewald sum treats every pair with the disjoint-
 * sphere form; for the nearest image at distance r < ai + aj the full
 * three-regime mobility replaces it: + (f - f_far) I + (g - g_far) u u^T.
 * Two points count as coincident when their nearest-image distance is at
 * most 1e-12 L.  Coordinates: 4 rows (x, y, z, radius) of length ld. */
typedef struct { double kx, ky, kz, w, wk; } ewald_k;   /* w = weight, wk = w k^2 / 3 */

int h2p_host_rpy_ewald(const double *c0, int ld0, int n0, const double *c1, int ld1, int n1,
                       double L, double xi, int nr, int nk, double *mat, long ldm, int n_thread)
{
    if (nr < 0 || nk < 0 || !(L > 0.0) || !(xi > 0.0)) return -1;
    const int side_k = 2 * nk + 1, side_r = 2 * nr + 1;
    const int n_k = (side_k * side_k * side_k - 1) / 2, n_r = side_r * side_r * side_r;
    ewald_k *kv = (ewald_k *)malloc(sizeof(ewald_k) * (n_k > 0 ? n_k : 1));
    double *sh = (double *)malloc(sizeof(double) * 3 * n_r);
    if (!kv || !sh) { free(kv); free(sh); return -2; }
    const double two_pi_L = 2.0 * M_PI / L, V = L * L * L;
    int m = 0;
    for (int hx = 0; hx <= nk; hx++)           /* h > 0: hx > 0, or hx = 0 and (hy, hz) > 0 */
        for (int hy = (hx > 0 ? -nk : 0); hy <= nk; hy++)
            for (int hz = (hx > 0 || hy > 0 ? -nk : 1); hz <= nk; hz++) {
                const double kx = two_pi_L * hx, ky = two_pi_L * hy, kz = two_pi_L * hz;
                const double k2 = kx * kx + ky * ky + kz * kz, y = k2 / (4.0 * xi * xi);
                const double w = 2.0 * (6.0 * M_PI / (V * k2)) * (1.0 + y + 2.0 * y * y) * exp(-y);
                kv[m++] = (ewald_k){kx, ky, kz, w, w * k2 / 3.0};
            }
    m = 0;
    for (int nx = -nr; nx <= nr; nx++)
        for (int ny = -nr; ny <= nr; ny++)
            for (int nz = -nr; nz <= nr; nz++, m++) {
                sh[3 * m] = L * nx;
                sh[3 * m + 1] = L * ny;
                sh[3 * m + 2] = L * nz;
            }
    const double xi2 = xi * xi, xsp = xi / sqrt(M_PI), coinc2 = 1e-24 * L * L;
    const double *x0 = c0, *y0 = c0 + ld0, *z0 = c0 + 2 * (long)ld0, *r0 = c0 + 3 * (long)ld0;
    const double *x1 = c1, *y1 = c1 + ld1, *z1 = c1 + 2 * (long)ld1, *r1 = c1 + 3 * (long)ld1;
    #pragma omp parallel for schedule(dynamic) num_threads(n_thread > 0 ? n_thread : 1)
    for (int i = 0; i < n0; i++) {
        for (int j = 0; j < n1; j++) {
            const double ai = r0[i], aj = r1[j], a3 = 0.5 * (ai * ai + aj * aj), alpha = xi2 * a3;
            const double d[3] = {x0[i] - x1[j], y0[i] - y1[j], z0[i] - z1[j]};
            double M[3][3] = {{0.0}};
            /* nearest image */
            double e[3], e2 = 0.0;
            for (int c = 0; c < 3; c++) {
                e[c] = d[c] - L * rint(d[c] / L);
                e2 += e[c] * e[c];
            }
            const int coincident = e2 <= coinc2;
            for (int t = 0; t < n_r; t++) {
                const double v[3] = {d[0] + sh[3 * t], d[1] + sh[3 * t + 1], d[2] + sh[3 * t + 2]};
                const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                if (r2 == 0.0 || (coincident && r2 <= coinc2)) continue;
                const double r = sqrt(r2), s2 = xi2 * r2, q = 0.75 / r, gs = xsp * exp(-s2), ec = erfc(xi * r);
                const double A = ec * q * (1.0 + 2.0 * a3 / (3.0 * r2)) +
                                 gs * (alpha / s2 + 14.0 * alpha - 4.5 + (3.0 - 20.0 * alpha) * s2 + 4.0 * alpha * s2 * s2);
                const double B = ec * q * (1.0 - 2.0 * a3 / r2) +
                                 gs * (-3.0 * alpha / s2 + 1.5 - 2.0 * alpha + (16.0 * alpha - 3.0) * s2 - 4.0 * alpha * s2 * s2);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) M[a][b] += B * v[a] * v[b] / r2 + (a == b ? A : 0.0);
            }
            for (int t = 0; t < n_k; t++) {
                const ewald_k *k = kv + t;
                const double k2 = 3.0 * k->wk / k->w, kk[3] = {k->kx, k->ky, k->kz};
                const double c = (k->w - a3 * k->wk) * cos(k->kx * d[0] + k->ky * d[1] + k->kz * d[2]);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) M[a][b] += c * ((a == b ? 1.0 : 0.0) - kk[a] * kk[b] / k2);
            }
            if (coincident) {
                const double self = 1.0 / ai - (6.0 - 40.0 / 3.0 * xi2 * ai * ai) * xsp;
                for (int a = 0; a < 3; a++) M[a][a] += self;
            } else {
                const double r = sqrt(e2);
                if (r <= ai + aj) {
                    double f, g, ff, gf;
                    rpy_mobility(r, ai, aj, &f, &g);
                    rpy_far(r, ai, aj, &ff, &gf);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) M[a][b] += (g - gf) * e[a] * e[b] / e2 + (a == b ? f - ff : 0.0);
                }
            }
            double *out = mat + (long)(3 * i) * ldm + 3 * j;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) out[a * ldm + b] = M[a][b];
        }
    }
    free(kv);
    free(sh);
    return 0;
}
