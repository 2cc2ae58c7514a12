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
