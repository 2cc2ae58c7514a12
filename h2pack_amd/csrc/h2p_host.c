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
    } else {  /* RPY, 3x3 blocks */
        const double *a0 = c0 + 3 * (long)ld0, *a1 = c1 + 3 * (long)ld1;
        const double C = 1.0 / (6.0 * M_PI * param[0]);
        for (int i = i0; i < i1; i++) {
            const double tx = x0[i], ty = y0[i], tz = z0[i], ta = a0[i];
            double *r0 = mat + 3L * i * ldm, *r1 = r0 + ldm, *r2p = r1 + ldm;
            for (int j = 0; j < n1; j++) {
                double dx = tx - x1[j], dy = ty - y1[j], dz = tz - z1[j];
                double sa = a1[j];
                double rr2 = dx * dx + dy * dy + dz * dz;
                double r = sqrt(rr2);
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
                    t1 = tmp0 + 3.0 * rr2;
                    t1 = tmp1 * (16.0 * rr2 * r * tps - t1 * t1);
                    t2 = tmp0 - rr2;
                    t2 = tmp1 * 3.0 * t2 * t2;
                } else {
                    t1 = C / (ta > sa ? ta : sa);
                    t2 = 0.0;
                }
                double *k0 = r0 + 3 * j, *k1 = r1 + 3 * j, *k2 = r2p + 3 * j;
                double u0 = t2 * dx, u1 = t2 * dy, u2 = t2 * dz;
                k0[0] = u0 * dx + t1; k0[1] = u0 * dy;      k0[2] = u0 * dz;
                k1[0] = u1 * dx;      k1[1] = u1 * dy + t1; k1[2] = u1 * dz;
                k2[0] = u2 * dx;      k2[1] = u2 * dy;      k2[2] = u2 * dz + t1;
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

/* Periodic RPY kernel by Ewald summation (the periodic system kernel of
 * H2P_build_periodic), following RPY_Ewald_init_workbuf / RPY_Ewald_eval_std
 * of src/H2Pack_3D_kernels.h:961-1210: for a cubic cell of side L, splitting
 * parameter xi, real-space images |i| <= nr per axis and reciprocal vectors
 * |k| <= nk per axis (half of them, by symmetry), each 3x3 block is the
 * real-space sum (excluding r = 0) + the reciprocal sum + the overlap
 * correction of the minimum image + the self term for coincident points.
 * Coordinates: 4 rows (x, y, z, radius) of length ld.  Like the reference,
 * the result carries no 1/(6 pi eta) factor. */
int h2p_host_rpy_ewald(const double *c0, int ld0, int n0, const double *c1, int ld1, int n1,
                       double L, double xi, int nr, int nk, double *mat, long ldm, int n_thread)
{
    const int r_size = (2 * nr + 1) * (2 * nr + 1) * (2 * nr + 1);
    const int k_size = ((2 * nk + 1) * (2 * nk + 1) * (2 * nk + 1) - 1) / 2;
    if (nr < 0 || nk < 0 || k_size > r_size) return -1;   /* the k vectors come from the nr cube */
    double *rs = (double *)malloc(sizeof(double) * 3 * r_size);
    double *kk = (double *)malloc(sizeof(double) * 6 * (k_size > 0 ? k_size : 1));
    if (!rs || !kk) { free(rs); free(kk); return -2; }
    const double V_inv = 1.0 / (L * L * L), tpl = 2.0 * M_PI / L;
    /* one walk over the nr cube: every (ix, iy, iz) is a real-space shift,
     * and the first k_size of them (lexicographic order, all before the
     * centre when nk = nr) are the reciprocal vectors, as in
     * RPY_Ewald_init_workbuf */
    int idx = 0;
    for (int ix = -nr; ix <= nr; ix++)
        for (int iy = -nr; iy <= nr; iy++)
            for (int iz = -nr; iz <= nr; iz++, idx++) {
                rs[3 * idx] = L * ix; rs[3 * idx + 1] = L * iy; rs[3 * idx + 2] = L * iz;
                if (idx < k_size) {
                    const double k = tpl * sqrt((double)(ix * ix + iy * iy + iz * iz));
                    const double k2 = k * k, q = k2 / (xi * xi);
                    kk[6 * idx] = k;
                    kk[6 * idx + 1] = 2.0 * V_inv * (1.0 + 0.25 * q + 0.125 * q * q) * 6.0 * M_PI / k2 * exp(-0.25 * q);
                    kk[6 * idx + 2] = tpl * ix; kk[6 * idx + 3] = tpl * iy; kk[6 * idx + 4] = tpl * iz;
                    kk[6 * idx + 5] = 1.0 / k;
                }
            }
    const double xi2 = xi * xi, xi3 = xi2 * xi, xi5 = xi3 * xi2, xi7 = xi5 * xi2;
    const double isp = 1.0 / sqrt(M_PI);
    const double *x0 = c0, *y0 = c0 + ld0, *z0 = c0 + 2 * (long)ld0, *a0 = c0 + 3 * (long)ld0;
    const double *x1 = c1, *y1 = c1 + ld1, *z1 = c1 + 2 * (long)ld1, *a1 = c1 + 3 * (long)ld1;
    #pragma omp parallel for schedule(dynamic) num_threads(n_thread > 0 ? n_thread : 1)
    for (int i = 0; i < n0; i++) {
        const double ai = a0[i];
        const double self_t = 1.0 / ai - (6.0 - 40.0 / 3.0 * xi2 * ai * ai) * xi * isp;
        for (int j = 0; j < n1; j++) {
            double m00 = 0, m10 = 0, m20 = 0, m11 = 0, m21 = 0, m22 = 0;
            const double dx = x0[i] - x1[j], dy = y0[i] - y1[j], dz = z0[i] - z1[j], aj = a1[j];
            const double a3 = 0.5 * (ai * ai + aj * aj);
            for (int s = 0; s < r_size; s++) {               /* real-space sum */
                double rx = dx + rs[3 * s], ry = dy + rs[3 * s + 1], rz = dz + rs[3 * s + 2];
                const double r2 = rx * rx + ry * ry + rz * rz;
                if (r2 == 0.0) continue;
                const double r = sqrt(r2), ri = 1.0 / r, ri2 = ri * ri, ri3 = ri * ri2, r4 = r2 * r2;
                const double e = erfc(xi * r), g = isp * exp(-xi2 * r2);
                const double t0 = 0.75 * ri, t1 = 0.5 * ri3 * a3, t2 = 4.0 * xi7 * a3 * r4;
                const double t3 = 3.0 * xi3 * r2, t4 = 4.0 * xi5 * a3 * r2, t5 = 2.0 * xi3 * a3;
                const double t6 = xi * a3 * ri2;
                const double p = (t0 + t1) * e + (t2 + t3 - 5.0 * t4 - 4.5 * xi + 7.0 * t5 + t6) * g;
                const double q = (t0 - 3.0 * t1) * e + (-t2 - t3 + 4.0 * t4 + 1.5 * xi - t5 - 3.0 * t6) * g;
                rx *= ri; ry *= ri; rz *= ri;
                m00 += q * rx * rx + p; m10 += q * rx * ry; m20 += q * rx * rz;
                m11 += q * ry * ry + p; m21 += q * ry * rz; m22 += q * rz * rz + p;
            }
            for (int s = 0; s < k_size; s++) {               /* reciprocal sum */
                const double k = kk[6 * s], m2 = kk[6 * s + 1], ki = kk[6 * s + 5];
                const double t = m2 * cos(kk[6 * s + 2] * dx + kk[6 * s + 3] * dy + kk[6 * s + 4] * dz) *
                                 (1.0 - a3 * k * k / 3.0);
                const double ux = kk[6 * s + 2] * ki, uy = kk[6 * s + 3] * ki, uz = kk[6 * s + 4] * ki;
                m00 += t * (1.0 - ux * ux); m10 -= t * ux * uy; m20 -= t * ux * uz;
                m11 += t * (1.0 - uy * uy); m21 -= t * uy * uz; m22 += t * (1.0 - uz * uz);
            }
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= 1e-30) {                               /* overlap correction, minimum image */
                double rx = fmod(dx + 2 * L, L), ry = fmod(dy + 2 * L, L), rz = fmod(dz + 2 * L, L);
                rx = rx > 0.5 * L ? rx - L : rx;
                ry = ry > 0.5 * L ? ry - L : ry;
                rz = rz > 0.5 * L ? rz - L : rz;
                const double s2 = rx * rx + ry * ry + rz * rz, r = sqrt(s2), r3 = s2 * r;
                const double ri = 1.0 / r, ri3 = ri * ri * ri;
                rx *= ri; ry *= ri; rz *= ri;
                const double u = (ai * ai + aj * aj) / s2;
                const double f1 = 0.75 * ri * (1.0 + u / 3.0), f2 = 0.75 * ri * (1.0 - u);
                const double d = ai - aj;
                double t1, t2;
                if (r > ai + aj) {
                    t1 = f1; t2 = f2;
                } else if (r > fabs(d)) {
                    const double w = ri3 / (32.0 * ai * aj);
                    t1 = d * d + 3.0 * s2;
                    t1 = (16.0 * r3 * (ai + aj) - t1 * t1) * w;
                    t2 = d * d - s2;
                    t2 = 3.0 * t2 * t2 * w;
                } else {
                    t1 = 1.0 / (ai > aj ? ai : aj); t2 = 0.0;
                }
                t1 -= f1; t2 -= f2;
                m00 += t2 * rx * rx + t1; m10 += t2 * rx * ry; m20 += t2 * rx * rz;
                m11 += t2 * ry * ry + t1; m21 += t2 * ry * rz; m22 += t2 * rz * rz + t1;
            } else {                                         /* self term */
                m00 += self_t; m11 += self_t; m22 += self_t;
            }
            double *b = mat + (long)(3 * i) * ldm + 3 * j;
            b[0] = m00;       b[1] = m10;       b[2] = m20;
            b[ldm] = m10;     b[ldm + 1] = m11; b[ldm + 2] = m21;
            b[2 * ldm] = m20; b[2 * ldm + 1] = m21; b[2 * ldm + 2] = m22;
        }
    }
    free(rs);
    free(kk);
    return 0;
}
