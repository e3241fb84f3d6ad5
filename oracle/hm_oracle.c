/*
 * hm_oracle.c -- FP64 CPU restatement of hitomi_measurement's periodic-box
 * estimator path.  TEST INFRASTRUCTURE ONLY (see hm_oracle.h): the checker
 * for tests/, smoke() and bench.py's cpu_baseline, never part of the
 * product.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -march, so that the
 * FP64 index arithmetic is evaluated exactly as the reference's plain
 * `g++ -O3` x86-64 build evaluates it: no fused multiply-add).
 *
 * Citations are reference file:line under src/hitomi_measurement/.
 */
#include "hm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* parameters                                                           */
/* ------------------------------------------------------------------ */

void or_param_init(or_param *p, int n0, int n1, int n2, double L0, double L1, double L2, int assign) {
    memset(p, 0, sizeof(*p));
    p->n_mesh[0] = n0;
    p->n_mesh[1] = n1;
    p->n_mesh[2] = n2;
    p->boxsize[0] = L0;
    p->boxsize[1] = L1;
    p->boxsize[2] = L2;
    p->volume = L0 * L1 * L2;                       /* parameter.hpp:232 */
    p->n_mesh_tot = (long long)n0 * n1 * n2;        /* parameter.hpp:233 (int there, Q8) */
    p->assign = assign;
    p->n_kbin = 1;
    p->n_rbin = 1;
}

void or_set_kbin(const or_param *p, double *kbin) {
    double dk = (p->kmax - p->kmin) / (double)(p->n_kbin - 1);
    for (int i = 0; i < p->n_kbin; i++) kbin[i] = p->kmin + dk * (double)i;
}

void or_set_rbin(const or_param *p, double *rbin) {
    double dr = (p->rmax - p->rmin) / (double)(p->n_rbin - 1);
    for (int i = 0; i < p->n_rbin; i++) rbin[i] = p->rmin + dr * (double)i;
}

long long or_coord(const or_param *p, int i, int j, int k) {
    return ((long long)i * p->n_mesh[1] + j) * p->n_mesh[2] + k;
}

/* ------------------------------------------------------------------ */
/* mass assignment, field.hpp:47-184                                    */
/* ------------------------------------------------------------------ */

/* Per axis: the first stencil index and the weights (field.hpp:84-86,
 * 121-126, 162-169). */
static int stencil_axis(int assign, double xp, int *iw0, double *w) {
    if (assign == OR_NGP) {
        *iw0 = (int)(xp + 0.5);
        w[0] = 1.0;
        return 1;
    } else if (assign == OR_CIC) {
        int i0 = (int)(xp);
        double dx = xp - (double)i0;
        w[0] = 1.0 - dx;
        w[1] = dx;
        *iw0 = i0;
        return 2;
    } else {
        int i1 = (int)(xp + 0.5);
        double dx = xp - (double)i1;
        w[0] = 0.5 * (0.5 - dx) * (0.5 - dx);
        w[1] = 0.75 - dx * dx;
        w[2] = 0.5 * (0.5 + dx) * (0.5 + dx);
        *iw0 = i1 - 1;
        return 3;
    }
}

int or_calc_field(const or_param *p, const double *pos, const double *weight, long long np, double *field) {
    for (long long i = 0; i < p->n_mesh_tot; i++) {
        field[2 * i] = 0.0;
        field[2 * i + 1] = 0.0;
    }
    double dV = p->volume / (double)(p->n_mesh_tot);
    double factor = 1.0 / dV;
    for (long long q = 0; q < np; q++) {
        double w[3][3];
        int iw0[3];
        int s = 1;
        for (int axes = 0; axes < 3; axes++) {
            double xp = pos[3 * q + axes] * (double)(p->n_mesh[axes]) / p->boxsize[axes];
            s = stencil_axis(p->assign, xp, &iw0[axes], w[axes]);
        }
        for (int i = 0; i < s; i++)
            for (int j = 0; j < s; j++)
                for (int k = 0; k < s; k++) {
                    /* field.hpp:175: the per-axis index is never wrapped (Q1);
                     * only the linear coord is range-checked. */
                    long long coord = ((long long)(iw0[0] + i) * p->n_mesh[1] + (iw0[1] + j)) * p->n_mesh[2] +
                                      (iw0[2] + k);
                    if ((coord >= 0) && (coord < p->n_mesh_tot)) {
                        field[2 * coord] += weight[2 * q] * factor * w[0][i] * w[1][j] * w[2][k];
                        field[2 * coord + 1] += weight[2 * q + 1] * factor * w[0][i] * w[1][j] * w[2][k];
                    }
                }
    }
    return 0;
}

int or_cell_coords(const or_param *p, const double *pos, long long np, long long *coords) {
    int s = (p->assign == OR_NGP) ? 1 : (p->assign == OR_CIC ? 2 : 3);
    int s3 = s * s * s;
    for (long long q = 0; q < np; q++) {
        double w[3][3];
        int iw0[3];
        for (int axes = 0; axes < 3; axes++) {
            double xp = pos[3 * q + axes] * (double)(p->n_mesh[axes]) / p->boxsize[axes];
            stencil_axis(p->assign, xp, &iw0[axes], w[axes]);
        }
        int t = 0;
        for (int i = 0; i < s; i++)
            for (int j = 0; j < s; j++)
                for (int k = 0; k < s; k++, t++) {
                    long long coord = ((long long)(iw0[0] + i) * p->n_mesh[1] + (iw0[1] + j)) * p->n_mesh[2] +
                                      (iw0[2] + k);
                    coords[q * s3 + t] = ((coord >= 0) && (coord < p->n_mesh_tot)) ? coord : -1;
                }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* FFT (replaces fftw_plan_dft_3d + fftw_execute, FFTW_ESTIMATE)        */
/* Mixed-radix Stockham autosort: radix 4, 2, 3, 5 and a direct DFT     */
/* stage for any other prime factor.                                     */
/* ------------------------------------------------------------------ */

typedef struct {
    int n, sign, nf;
    int f[64];
    double *tw; /* tw[m] = exp(sign * 2 pi i m / n), m < n */
} fft_plan1;

static void plan1_init(fft_plan1 *pl, int n, int sign) {
    pl->n = n;
    pl->sign = sign;
    pl->nf = 0;
    int m = n;
    while (m % 4 == 0) { pl->f[pl->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->f[pl->nf++] = 2; m /= 2; }
    for (int r = 3; (long long)r * r <= m; r += 2)
        while (m % r == 0) { pl->f[pl->nf++] = r; m /= r; }
    if (m > 1) pl->f[pl->nf++] = m;
    pl->tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int t = 0; t < n; t++) {
        double a = 2.0 * M_PI * (double)t / (double)n;
        pl->tw[2 * t] = cos(a);
        pl->tw[2 * t + 1] = (double)sign * sin(a);
    }
}

static void plan1_free(fft_plan1 *pl) { free(pl->tw); pl->tw = NULL; }

static void stockham_stage(const double *x, double *y, const fft_plan1 *pl, int r, int ns, double *a) {
    const int n = pl->n;
    const int m = n / r;
    const long long tstep = n / (ns * r);
    const double *tw = pl->tw;
    const double s = (double)pl->sign;
    for (int j = 0; j < m; j++) {
        int k = j % ns;
        for (int t = 0; t < r; t++) {
            double xr = x[2 * (j + t * m)], xi = x[2 * (j + t * m) + 1];
            long long e = ((long long)t * k * tstep) % n;
            double cr = tw[2 * e], ci = tw[2 * e + 1];
            a[2 * t] = xr * cr - xi * ci;
            a[2 * t + 1] = xr * ci + xi * cr;
        }
        int base = (j / ns) * ns * r + k;
        if (r == 2) {
            y[2 * base] = a[0] + a[2];
            y[2 * base + 1] = a[1] + a[3];
            y[2 * (base + ns)] = a[0] - a[2];
            y[2 * (base + ns) + 1] = a[1] - a[3];
        } else if (r == 4) {
            double t0r = a[0] + a[4], t0i = a[1] + a[5];
            double t1r = a[0] - a[4], t1i = a[1] - a[5];
            double t2r = a[2] + a[6], t2i = a[3] + a[7];
            /* (a1 - a3) * (sign * i) */
            double t3r = -s * (a[3] - a[7]), t3i = s * (a[2] - a[6]);
            y[2 * base] = t0r + t2r;
            y[2 * base + 1] = t0i + t2i;
            y[2 * (base + ns)] = t1r + t3r;
            y[2 * (base + ns) + 1] = t1i + t3i;
            y[2 * (base + 2 * ns)] = t0r - t2r;
            y[2 * (base + 2 * ns) + 1] = t0i - t2i;
            y[2 * (base + 3 * ns)] = t1r - t3r;
            y[2 * (base + 3 * ns) + 1] = t1i - t3i;
        } else {
            const long long ustep = n / r;
            for (int u = 0; u < r; u++) {
                double sr = 0.0, si = 0.0;
                for (int t = 0; t < r; t++) {
                    long long e = ((long long)t * u * ustep) % n;
                    double cr = tw[2 * e], ci = tw[2 * e + 1];
                    sr += a[2 * t] * cr - a[2 * t + 1] * ci;
                    si += a[2 * t] * ci + a[2 * t + 1] * cr;
                }
                y[2 * (base + u * ns)] = sr;
                y[2 * (base + u * ns) + 1] = si;
            }
        }
    }
}

/* x: n complex values; work: 2n doubles; a: 2*maxradix doubles. */
static void fft_line(double *x, const fft_plan1 *pl, double *work, double *a) {
    double *src = x, *dst = work;
    int ns = 1;
    for (int s = 0; s < pl->nf; s++) {
        stockham_stage(src, dst, pl, pl->f[s], ns, a);
        ns *= pl->f[s];
        double *t = src; src = dst; dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)pl->n);
}

static int max_radix(const fft_plan1 *pl) {
    int r = 4;
    for (int s = 0; s < pl->nf; s++) if (pl->f[s] > r) r = pl->f[s];
    return r;
}

int or_fft1d(double *x, int n, int sign) {
    fft_plan1 pl;
    plan1_init(&pl, n, sign);
    double *work = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *a = (double *)malloc(sizeof(double) * 2 * (size_t)max_radix(&pl));
    fft_line(x, &pl, work, a);
    free(a);
    free(work);
    plan1_free(&pl);
    return 0;
}

#define OR_BLK 16

/* Transform along one axis of a row-major [n0][n1][n2] complex grid:
 * lines of length len with point stride `stride`, for all (outer, inner). */
static void fft_axis(double *f, int len, long long stride, long long n_outer, long long outer_stride,
                     long long n_inner, int sign) {
    fft_plan1 pl;
    plan1_init(&pl, len, sign);
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)len * OR_BLK);
    double *work = (double *)malloc(sizeof(double) * 2 * (size_t)len);
    double *a = (double *)malloc(sizeof(double) * 2 * (size_t)max_radix(&pl));
    for (long long o = 0; o < n_outer; o++) {
        for (long long i0 = 0; i0 < n_inner; i0 += OR_BLK) {
            int nb = (int)((n_inner - i0) < OR_BLK ? (n_inner - i0) : OR_BLK);
            double *base = f + 2 * (o * outer_stride + i0);
            for (int t = 0; t < len; t++)
                for (int b = 0; b < nb; b++) {
                    buf[2 * ((size_t)b * len + t)] = base[2 * (t * stride + b)];
                    buf[2 * ((size_t)b * len + t) + 1] = base[2 * (t * stride + b) + 1];
                }
            for (int b = 0; b < nb; b++) fft_line(buf + 2 * (size_t)b * len, &pl, work, a);
            for (int t = 0; t < len; t++)
                for (int b = 0; b < nb; b++) {
                    base[2 * (t * stride + b)] = buf[2 * ((size_t)b * len + t)];
                    base[2 * (t * stride + b) + 1] = buf[2 * ((size_t)b * len + t) + 1];
                }
        }
    }
    free(a);
    free(work);
    free(buf);
    plan1_free(&pl);
}

int or_fft3d(const or_param *p, double *f, int sign) {
    const long long n0 = p->n_mesh[0], n1 = p->n_mesh[1], n2 = p->n_mesh[2];
    /* z: contiguous lines */
    {
        fft_plan1 pl;
        plan1_init(&pl, (int)n2, sign);
        double *work = (double *)malloc(sizeof(double) * 2 * (size_t)n2);
        double *a = (double *)malloc(sizeof(double) * 2 * (size_t)max_radix(&pl));
        for (long long l = 0; l < n0 * n1; l++) fft_line(f + 2 * l * n2, &pl, work, a);
        free(a);
        free(work);
        plan1_free(&pl);
    }
    /* y: stride n2, for each x-plane */
    fft_axis(f, (int)n1, n2, n0, n1 * n2, n2, sign);
    /* x: stride n1*n2 */
    fft_axis(f, (int)n0, n1 * n2, 1, 0, n1 * n2, sign);
    return 0;
}

/* field.hpp:310-326 */
void or_fourier_transform(const or_param *p, double *field) {
    double dV = p->volume / (double)(p->n_mesh_tot);
    for (long long i = 0; i < p->n_mesh_tot; i++) {
        field[2 * i] *= dV;
        field[2 * i + 1] *= dV;
    }
    or_fft3d(p, field, -1);
}

/* field.hpp:328-344 */
void or_inverse_fourier_transform(const or_param *p, double *field) {
    double V = p->volume;
    for (long long i = 0; i < p->n_mesh_tot; i++) {
        field[2 * i] /= V;
        field[2 * i + 1] /= V;
    }
    or_fft3d(p, field, +1);
}

/* ------------------------------------------------------------------ */
/* assignment / shot-noise window functions                             */
/* ------------------------------------------------------------------ */

static void kvec_of(const or_param *p, int i, int j, int k, const double *dk, double *kvec) {
    /* field.hpp:797-799: the Nyquist index maps to the negative frequency (Q15) */
    kvec[0] = (i < p->n_mesh[0] / 2) ? (double)i * dk[0] : (double)(i - p->n_mesh[0]) * dk[0];
    kvec[1] = (j < p->n_mesh[1] / 2) ? (double)j * dk[1] : (double)(j - p->n_mesh[1]) * dk[1];
    kvec[2] = (k < p->n_mesh[2] / 2) ? (double)k * dk[2] : (double)(k - p->n_mesh[2]) * dk[2];
}

/* field.hpp:641-717 (truncation toward zero of k/dk + 1e-5: n<0 -> |n|-1, Q2) */
double or_assignment_function(const or_param *p, const double *kvec) {
    double dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    int i = (int)(kvec[0] / dk[0] + 1.0e-5);
    int j = (int)(kvec[1] / dk[1] + 1.0e-5);
    int k = (int)(kvec[2] / dk[2] + 1.0e-5);
    double x1 = M_PI * (double)i / (double)(p->n_mesh[0]);
    double x2 = M_PI * (double)j / (double)(p->n_mesh[1]);
    double x3 = M_PI * (double)k / (double)(p->n_mesh[2]);
    double f1 = (i != 0) ? sin(x1) / x1 : 1.0;
    double f2 = (j != 0) ? sin(x2) / x2 : 1.0;
    double f3 = (k != 0) ? sin(x3) / x3 : 1.0;
    double ret = f1 * f2 * f3;
    if (p->assign == OR_NGP) return pow(ret, 1);
    if (p->assign == OR_CIC) return pow(ret, 2);
    if (p->assign == OR_TSC) return pow(ret, 3);
    return 1.0;
}

/* field.hpp:1252-1310 */
double or_shotnoise_function(const or_param *p, const double *kvec) {
    if (p->assign == OR_NGP) return 1.0;
    double dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    int i = (int)(kvec[0] / dk[0] + 1.0e-5);
    int j = (int)(kvec[1] / dk[1] + 1.0e-5);
    int k = (int)(kvec[2] / dk[2] + 1.0e-5);
    double x1 = M_PI * (double)i / (double)(p->n_mesh[0]);
    double x2 = M_PI * (double)j / (double)(p->n_mesh[1]);
    double x3 = M_PI * (double)k / (double)(p->n_mesh[2]);
    double f1 = (i != 0) ? sin(x1) : 0.0;
    double f2 = (j != 0) ? sin(x2) : 0.0;
    double f3 = (k != 0) ? sin(x3) : 0.0;
    if (p->assign == OR_CIC) {
        return (1.0 - (2.0 / 3.0) * f1 * f1) * (1.0 - (2.0 / 3.0) * f2 * f2) * (1.0 - (2.0 / 3.0) * f3 * f3);
    }
    if (p->assign == OR_TSC) {
        return (1.0 - f1 * f1 + 2.0 * pow(f1, 4) / 15.0) * (1.0 - f2 * f2 + 2.0 * pow(f2, 4) / 15.0) *
               (1.0 - f3 * f3 + 2.0 * pow(f3, 4) / 15.0);
    }
    return 0.0;
}

/* field.hpp:610-639 */
void or_assignment_correction(const or_param *p, double *field) {
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dk, kvec);
                double wf = or_assignment_function(p, kvec);
                field[2 * coord] /= wf;
                field[2 * coord + 1] /= wf;
            }
}

/* ------------------------------------------------------------------ */
/* special functions                                                    */
/* ------------------------------------------------------------------ */

static double factd(int n) {
    double r = 1.0;
    for (int i = 2; i <= n; i++) r *= (double)i;
    return r;
}

/* gsl_sf_legendre_sphPlm(l, m, x) = sqrt((2l+1)/(4pi)) sqrt((l-m)!/(l+m)!) P_l^m(x),
 * P_l^m with the Condon-Shortley phase (-1)^m, 0 <= m <= l. */
double or_sphPlm(int l, int m, double x) {
    if (m < 0 || m > l) return 0.0;
    double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double pmm = 1.0;
    double fact = 1.0;
    for (int i = 1; i <= m; i++) {
        pmm *= -fact * somx2;
        fact += 2.0;
    }
    double plm;
    if (l == m) {
        plm = pmm;
    } else {
        double pmmp1 = x * (2 * m + 1) * pmm;
        if (l == m + 1) {
            plm = pmmp1;
        } else {
            double pll = 0.0;
            for (int ll = m + 2; ll <= l; ll++) {
                pll = (x * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pmm) / (double)(ll - m);
                pmm = pmmp1;
                pmmp1 = pll;
            }
            plm = pll;
        }
    }
    double norm = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * factd(l - m) / factd(l + m));
    return norm * plm;
}

/* tools.hpp:12-68: conj(Y_lm(v)) * sqrt(4pi/(2l+1)) with the reference's conventions. */
void or_ylm(int ell, int M, const double *pos, double *out) {
    if ((ell == 0) && (M == 0)) {
        out[0] = 1.0;
        out[1] = 0.0;
        return;
    }
    double xmag_xyz2 = 0.0;
    for (int axes = 0; axes < 3; axes++) xmag_xyz2 += pos[axes] * pos[axes];
    double xmag_xyz = sqrt(xmag_xyz2);
    double xmag_xy = sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    double mu = 0.0;
    if (fabs(xmag_xyz) < 1.0e-15) {
        out[0] = 0.0;
        out[1] = 0.0;
        return;
    } else {
        mu = pos[2] / xmag_xyz;
    }
    double phi = 0.0;
    if (fabs(xmag_xy) < 1.0e-15) {
        phi = 0.0;
    } else {
        double cosphi = pos[0] / xmag_xy;
        phi = acos(cosphi);
    }
    if (pos[1] < 0.0) phi = -phi + 2.0 * M_PI;
    int S = abs(M);
    double sgn = pow(-1.0, (M - S) / 2.0);
    double plm = or_sphPlm(ell, S, mu);
    double c = cos((double)M * phi), s = sin((double)M * phi);
    double yr = sgn * plm * c, yi = sgn * plm * s;
    yi = -yi; /* conj */
    double nrm = sqrt(4.0 * M_PI / (2.0 * (double)ell + 1.0));
    out[0] = yr * nrm;
    out[1] = yi * nrm;
}

/* gsl_sf_coupling_3j via the Racah formula. */
double or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0) return 0.0;
    if (j1 < 0 || j2 < 0 || j3 < 0) return 0.0;
    if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
    if (j3 < abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    double tri = factd(j1 + j2 - j3) * factd(j1 - j2 + j3) * factd(-j1 + j2 + j3) / factd(j1 + j2 + j3 + 1);
    double pre = sqrt(tri * factd(j1 + m1) * factd(j1 - m1) * factd(j2 + m2) * factd(j2 - m2) * factd(j3 + m3) *
                      factd(j3 - m3));
    int kmin = 0;
    if (j2 - j3 - m1 > kmin) kmin = j2 - j3 - m1;
    if (j1 - j3 + m2 > kmin) kmin = j1 - j3 + m2;
    int kmax = j1 + j2 - j3;
    if (j1 - m1 < kmax) kmax = j1 - m1;
    if (j2 + m2 < kmax) kmax = j2 + m2;
    double sum = 0.0;
    for (int k = kmin; k <= kmax; k++) {
        double den = factd(k) * factd(j1 + j2 - j3 - k) * factd(j1 - m1 - k) * factd(j2 + m2 - k) *
                     factd(j3 - j2 + m1 + k) * factd(j3 - j1 - m2 + k);
        sum += ((k % 2) ? -1.0 : 1.0) / den;
    }
    double sgn = (((j1 - j2 - m3) % 2) != 0) ? -1.0 : 1.0;
    return sgn * pre * sum;
}

/* Spherical Bessel j_l(x), x >= 0: power series for x < l + 2, upward
 * recurrence from j0, j1 beyond. */
double or_bessel_jl(int l, double x) {
    if (x == 0.0) return (l == 0) ? 1.0 : 0.0;
    if (x < (double)l + 2.0) {
        double dfact = 1.0; /* (2l+1)!! */
        for (int i = 1; i <= 2 * l + 1; i += 2) dfact *= (double)i;
        double pre = pow(x, l) / dfact;
        double term = 1.0, sum = 1.0, y = -0.5 * x * x;
        for (int k = 1; k < 200; k++) {
            term *= y / ((double)k * (double)(2 * l + 2 * k + 1));
            sum += term;
            if (fabs(term) < 1e-18 * fabs(sum)) break;
        }
        return pre * sum;
    }
    double s = sin(x), c = cos(x);
    double j0 = s / x;
    if (l == 0) return j0;
    double j1 = s / (x * x) - c / x;
    for (int n = 1; n < l; n++) {
        double j2 = (2.0 * n + 1.0) / x * j1 - j0;
        j0 = j1;
        j1 = j2;
    }
    return j1;
}

/* bessel.hpp:10-37: natural cubic spline (gsl_interp_cspline) of j_l on
 * 10^6 equidistant nodes in [0, 1e4]. */
struct or_bessel {
    int nn;
    double xmin, xmax, dx;
    double *x, *y, *c; /* c = second derivative / 2 at the nodes */
};

or_bessel *or_bessel_new(int ell) {
    or_bessel *b = (or_bessel *)calloc(1, sizeof(or_bessel));
    const int NN = 1000000;
    b->nn = NN;
    b->xmin = 0;
    b->xmax = 10000;
    b->dx = (b->xmax - b->xmin) / (double)(NN - 1);
    b->x = (double *)malloc(sizeof(double) * NN);
    b->y = (double *)malloc(sizeof(double) * NN);
    b->c = (double *)calloc(NN, sizeof(double));
    for (int i = 0; i < NN; i++) {
        b->x[i] = b->xmin + b->dx * (double)i;
        b->y[i] = or_bessel_jl(ell, b->x[i]);
    }
    /* natural spline: c_0 = c_{N-1} = 0; tridiagonal system for c_1..c_{N-2}
     * h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1} = 3 (g_i - g_{i-1}) */
    int m = NN - 2;
    double *diag = (double *)malloc(sizeof(double) * m);
    double *off = (double *)malloc(sizeof(double) * m);
    double *rhs = (double *)malloc(sizeof(double) * m);
    for (int i = 0; i < m; i++) {
        double h_i = b->x[i + 2] - b->x[i + 1];
        double h_im1 = b->x[i + 1] - b->x[i];
        double g_i = (b->y[i + 2] - b->y[i + 1]) / h_i;
        double g_im1 = (b->y[i + 1] - b->y[i]) / h_im1;
        diag[i] = 2.0 * (h_im1 + h_i);
        off[i] = h_i;
        rhs[i] = 3.0 * (g_i - g_im1);
    }
    /* Thomas algorithm (symmetric tridiagonal) */
    for (int i = 1; i < m; i++) {
        double w = off[i - 1] / diag[i - 1];
        diag[i] -= w * off[i - 1];
        rhs[i] -= w * rhs[i - 1];
    }
    b->c[m] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 2; i >= 0; i--) b->c[i + 1] = (rhs[i] - off[i] * b->c[i + 2]) / diag[i];
    free(diag);
    free(off);
    free(rhs);
    return b;
}

double or_bessel_eval(const or_bessel *b, double x) {
    if (x < b->x[0] || x > b->x[b->nn - 1]) return NAN; /* GSL_EDOM (Q13) */
    /* interval with x_i <= x < x_{i+1} (last interval closed) */
    long lo = 0, hi = b->nn - 1;
    while (hi > lo + 1) {
        long mid = (lo + hi) / 2;
        if (b->x[mid] > x) hi = mid; else lo = mid;
    }
    long i = lo;
    double x_lo = b->x[i], x_hi = b->x[i + 1];
    double y_lo = b->y[i], y_hi = b->y[i + 1];
    double dx = x_hi - x_lo, delx = x - x_lo;
    double c_i = b->c[i], c_ip1 = b->c[i + 1];
    double bb = (y_hi - y_lo) / dx - dx * (c_ip1 + 2.0 * c_i) / 3.0;
    double d = (c_ip1 - c_i) / (3.0 * dx);
    return y_lo + delx * (bb + delx * (c_i + delx * d));
}

void or_bessel_free(or_bessel *b) {
    if (!b) return;
    free(b->x);
    free(b->y);
    free(b->c);
    free(b);
}

/* ------------------------------------------------------------------ */
/* two-point statistics                                                 */
/* ------------------------------------------------------------------ */

/* (ar + i ai)(br + i bi) as std::complex<double> evaluates it */
#define CMUL(ar, ai, br, bi, outr, outi) \
    do {                                  \
        double _r = (ar) * (br) - (ai) * (bi); \
        double _i = (ar) * (bi) + (ai) * (br); \
        outr = _r;                        \
        outi = _i;                        \
    } while (0)

/* field.hpp:765-852 */
/* Shot-noise shape of the binned estimators: the reference's
 * ShotnoiseFunction (0), or W(k)^2 for an interlaced field, whose odd-order
 * aliases cancel (1).  Set only by the interlaced drivers below. */
static int g_sn_shape_w2 = 0;
static double sn_shape(const or_param *p, const double *kvec) {
    return g_sn_shape_w2 ? pow(or_assignment_function(p, kvec), 2) : or_shotnoise_function(p, kvec);
}

int or_calc_power_spectrum(const or_param *p, const double *f1, const double *f2, const double *kbin, double sn_re,
                           double sn_im, int ell, int m, double *pk, int *n_mode) {
    double dk_bin = kbin[1] - kbin[0];
    const int n_temp = 100000;
    const double dk_temp = 0.0001;
    double *pk_temp = (double *)calloc(2 * (size_t)n_temp, sizeof(double));
    int *n_mode_temp = (int *)calloc((size_t)n_temp, sizeof(int));
    for (int i = 0; i < p->n_kbin; i++) {
        pk[2 * i] = 0.0;
        pk[2 * i + 1] = 0.0;
        n_mode[i] = 0;
    }
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dk, kvec);
                double kmag = sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
                int kindex = (int)(kmag / dk_temp + 0.5);
                if (kindex < n_temp) {
                    double pr, pi;
                    CMUL(f1[2 * coord], f1[2 * coord + 1], f2[2 * coord], -f2[2 * coord + 1], pr, pi);
                    double S = sn_shape(p, kvec);
                    pr -= sn_re * S;
                    pi -= sn_im * S;
                    double wf = or_assignment_function(p, kvec);
                    double w2 = pow(wf, 2);
                    pr /= w2;
                    pi /= w2;
                    double y[2];
                    or_ylm(ell, m, kvec, y);
                    CMUL(pr, pi, y[0], y[1], pr, pi);
                    pk_temp[2 * kindex] += pr;
                    pk_temp[2 * kindex + 1] += pi;
                    n_mode_temp[kindex]++;
                }
            }
    for (int j = 0; j < p->n_kbin; j++) {
        double lower = (kbin[j] > dk_bin * 0.5) ? (kbin[j] - dk_bin * 0.5) : 0.0;
        double upper = kbin[j] + dk_bin * 0.5;
        for (int i = 0; i < n_temp; i++) {
            double kmag_temp = (double)i * dk_temp;
            if ((kmag_temp > lower) && (kmag_temp <= upper)) {
                pk[2 * j] += pk_temp[2 * i];
                pk[2 * j + 1] += pk_temp[2 * i + 1];
                n_mode[j] += n_mode_temp[i];
            }
        }
    }
    for (int i = 0; i < p->n_kbin; i++) {
        if (n_mode[i] != 0) {
            pk[2 * i] /= (double)n_mode[i];
            pk[2 * i + 1] /= (double)n_mode[i];
        } else {
            pk[2 * i] = 0.0;
            pk[2 * i + 1] = 0.0;
        }
    }
    free(pk_temp);
    free(n_mode_temp);
    return 0;
}

/* Mode counts alone, in one pass over the mesh: the P_l bins of
 * calcPowerSpectrum (field.hpp:793-837: |k| from the Nyquist-negative
 * frequencies, fine index int(|k|/1e-4 + 0.5) below 1e5, fold over fine
 * centres in (lower, upper]) and the n_mode of every bispectrum shell of
 * calcInverseFourierTransformForBispectrum (field.hpp:361-387: the raw |k| in
 * (max(k_i - dk/2, 0), k_i + dk/2]) for shell centres kbin[i] and width
 * kbin[1] - kbin[0], as calcBiSpectrumForBOX passes them (bispec.hpp:1100-1110).
 * The same FP64 expressions as or_calc_power_spectrum and
 * or_ifft_for_bispectrum, without the field arithmetic, so it runs at 512^3. */
int or_mode_counts(const or_param *p, const double *kbin, long long *pk_nmode, long long *shell_nmode) {
    const int n_temp = 100000;
    const double dk_temp = 0.0001;
    const int nk = p->n_kbin;
    double dk_bin = kbin[1] - kbin[0];
    long long *n_mode_temp = (long long *)calloc((size_t)n_temp, sizeof(long long));
    double *slo = (double *)malloc(sizeof(double) * (size_t)nk), *shi = (double *)malloc(sizeof(double) * (size_t)nk);
    if (!n_mode_temp || !slo || !shi) {
        free(n_mode_temp);
        free(slo);
        free(shi);
        return -1;
    }
    for (int j = 0; j < nk; j++) {
        slo[j] = (kbin[j] > dk_bin * 0.5) ? (kbin[j] - dk_bin * 0.5) : 0.0;
        shi[j] = kbin[j] + dk_bin * 0.5;
        pk_nmode[j] = 0;
        shell_nmode[j] = 0;
    }
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                kvec_of(p, i, j, k, dk, kvec);
                double kmag = sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
                int kindex = (int)(kmag / dk_temp + 0.5);
                if (kindex < n_temp) n_mode_temp[kindex]++;
                for (int b = 0; b < nk; b++)
                    if ((kmag > slo[b]) && (kmag <= shi[b])) shell_nmode[b]++;
            }
    for (int j = 0; j < nk; j++)
        for (int i = 0; i < n_temp; i++) {
            double kmag_temp = (double)i * dk_temp;
            if ((kmag_temp > slo[j]) && (kmag_temp <= shi[j])) pk_nmode[j] += n_mode_temp[i];
        }
    free(n_mode_temp);
    free(slo);
    free(shi);
    return 0;
}

/* field.hpp:854-977 */
int or_calc_correlation_function(const or_param *p, const double *f1, const double *f2, const double *rbin,
                                 double sn_re, double sn_im, int ell, int m, double *xi, int *n_mode) {
    double *xi3d = (double *)calloc(2 * (size_t)p->n_mesh_tot, sizeof(double));
    double factor = 1.0 / p->volume;
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dk, kvec);
                double pr, pi;
                CMUL(f1[2 * coord], f1[2 * coord + 1], f2[2 * coord], -f2[2 * coord + 1], pr, pi);
                double S = sn_shape(p, kvec);
                pr -= sn_re * S;
                pi -= sn_im * S;
                double w2 = pow(or_assignment_function(p, kvec), 2);
                pr /= w2;
                pi /= w2;
                xi3d[2 * coord] = factor * pr;
                xi3d[2 * coord + 1] = factor * pi;
            }
    or_fft3d(p, xi3d, +1);
    const int n_temp = 20000;
    const double dr_temp = 0.5;
    double *xi_temp = (double *)calloc(2 * (size_t)n_temp, sizeof(double));
    int *n_mode_temp = (int *)calloc((size_t)n_temp, sizeof(int));
    for (int i = 0; i < p->n_rbin; i++) {
        xi[2 * i] = 0.0;
        xi[2 * i + 1] = 0.0;
        n_mode[i] = 0;
    }
    double rvec[3], dr[3];
    dr[0] = p->boxsize[0] / (double)(p->n_mesh[0]);
    dr[1] = p->boxsize[1] / (double)(p->n_mesh[1]);
    dr[2] = p->boxsize[2] / (double)(p->n_mesh[2]);
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dr, rvec);
                double rmag = sqrt(rvec[0] * rvec[0] + rvec[1] * rvec[1] + rvec[2] * rvec[2]);
                int rindex = (int)(rmag / dr_temp + 0.5);
                if (rindex < n_temp) {
                    double y[2], xr, xim;
                    or_ylm(ell, m, rvec, y);
                    CMUL(xi3d[2 * coord], xi3d[2 * coord + 1], y[0], y[1], xr, xim);
                    xi_temp[2 * rindex] += xr;
                    xi_temp[2 * rindex + 1] += xim;
                    n_mode_temp[rindex]++;
                }
            }
    double dr_bin = rbin[1] - rbin[0];
    for (int j = 0; j < p->n_rbin; j++) {
        double lower = (rbin[j] > dr_bin * 0.5) ? (rbin[j] - dr_bin * 0.5) : 0.0;
        double upper = rbin[j] + dr_bin * 0.5;
        for (int i = 0; i < n_temp; i++) {
            double rmag_temp = (double)i * dr_temp;
            if ((rmag_temp > lower) && (rmag_temp <= upper)) {
                xi[2 * j] += xi_temp[2 * i];
                xi[2 * j + 1] += xi_temp[2 * i + 1];
                n_mode[j] += n_mode_temp[i];
            }
        }
    }
    for (int i = 0; i < p->n_rbin; i++) {
        if (n_mode[i] != 0) {
            xi[2 * i] /= (double)n_mode[i];
            xi[2 * i + 1] /= (double)n_mode[i];
        } else {
            xi[2 * i] = 0.0;
            xi[2 * i + 1] = 0.0;
        }
    }
    free(xi3d);
    free(xi_temp);
    free(n_mode_temp);
    return 0;
}

/* field.hpp:346-401: shell (max(k-dk/2,0), k+dk/2] on the raw |k| (Q3),
 * value Y * delta / W, backward FFT, divided by the mode count. */
int or_ifft_for_bispectrum(const or_param *p, const double *density, double kmag_in, double dk_in, const double *ylm,
                           double *out) {
    for (long long i = 0; i < p->n_mesh_tot; i++) {
        out[2 * i] = 0.0;
        out[2 * i + 1] = 0.0;
    }
    int n_mode = 0;
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dk, kvec);
                double kmag = sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
                double lower = (kmag_in > dk_in * 0.5) ? (kmag_in - dk_in * 0.5) : 0.0;
                double upper = kmag_in + dk_in * 0.5;
                if ((kmag > lower) && (kmag <= upper)) {
                    double wf = or_assignment_function(p, kvec);
                    double dr = density[2 * coord] / wf, di = density[2 * coord + 1] / wf;
                    double yr = ylm ? ylm[2 * coord] : 1.0, yi = ylm ? ylm[2 * coord + 1] : 0.0;
                    CMUL(yr, yi, dr, di, out[2 * coord], out[2 * coord + 1]);
                    n_mode++;
                } else {
                    out[2 * coord] = 0.0;
                    out[2 * coord + 1] = 0.0;
                }
            }
    or_fft3d(p, out, +1);
    for (long long c = 0; c < p->n_mesh_tot; c++) {
        out[2 * c] /= (double)n_mode;
        out[2 * c + 1] /= (double)n_mode;
    }
    return n_mode;
}

/* field.hpp:1116-1160 */
int or_shotnoise_bispectrum_ijk(const or_param *p, const double *f1, const double *f2, double sn_re, double sn_im,
                                double *xi) {
    double factor = 1.0 / p->volume;
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dk, kvec);
                double pr, pi;
                CMUL(f1[2 * coord], f1[2 * coord + 1], f2[2 * coord], -f2[2 * coord + 1], pr, pi);
                double S = or_shotnoise_function(p, kvec);
                pr -= sn_re * S;
                pi -= sn_im * S;
                double w2 = pow(or_assignment_function(p, kvec), 2);
                pr /= w2;
                pi /= w2;
                xi[2 * coord] = factor * pr;
                xi[2 * coord + 1] = factor * pi;
            }
    or_fft3d(p, xi, +1);
    return 0;
}

/* ------------------------------------------------------------------ */
/* box density fields and drivers                                       */
/* ------------------------------------------------------------------ */

static double *unit_weights(long long np) {
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)(np > 0 ? np : 1));
    for (long long q = 0; q < np; q++) {
        w[2 * q] = 1.0;
        w[2 * q + 1] = 0.0;
    }
    return w;
}

int or_box_density_fluctuation(const or_param *p, const double *pos, long long np, double *field) {
    double *w = unit_weights(np);
    or_calc_field(p, pos, w, np, field);
    free(w);
    for (long long i = 0; i < p->n_mesh_tot; i++) {
        field[2 * i] -= ((double)np / p->volume);
        field[2 * i + 1] -= 0.0;
    }
    return 0;
}

int or_box_density(const or_param *p, const double *pos, long long np, double *field) {
    double *w = unit_weights(np);
    or_calc_field(p, pos, w, np, field);
    free(w);
    return 0;
}

static double *grid_alloc(const or_param *p) {
    return (double *)calloc(2 * (size_t)p->n_mesh_tot, sizeof(double));
}

/* powerspec.hpp:273-332 */
int or_box_pk(const or_param *p, const double *pos, long long np, const double *kbin, double *pk_out, int *n_mode) {
    if (!((p->ELL == p->ell1) && (p->ell2 == 0))) return -1;
    double *dn = grid_alloc(p);
    or_box_density_fluctuation(p, pos, np, dn);
    or_fourier_transform(p, dn);
    double *pk = (double *)malloc(sizeof(double) * 2 * p->n_kbin);
    or_calc_power_spectrum(p, dn, dn, kbin, (double)np, 0.0, p->ELL, 0, pk, n_mode);
    double norm = p->volume / (double)np / (double)np;
    for (int i = 0; i < p->n_kbin; i++) {
        double sr = 0.0 + (double)(2 * p->ELL + 1) * pk[2 * i];
        double si = 0.0 + (double)(2 * p->ELL + 1) * pk[2 * i + 1];
        pk_out[2 * i] = norm * sr;
        pk_out[2 * i + 1] = norm * si;
    }
    free(pk);
    free(dn);
    return 0;
}

/* powerspec.hpp:334-393 */
int or_box_xi(const or_param *p, const double *pos, long long np, const double *rbin, double *xi_out) {
    if (!((p->ELL == p->ell1) && (p->ell2 == 0))) return -1;
    double *dn = grid_alloc(p);
    or_box_density_fluctuation(p, pos, np, dn);
    or_fourier_transform(p, dn);
    double *xi = (double *)malloc(sizeof(double) * 2 * p->n_rbin);
    int *nm = (int *)malloc(sizeof(int) * p->n_rbin);
    or_calc_correlation_function(p, dn, dn, rbin, (double)np, 0.0, p->ELL, 0, xi, nm);
    double norm = p->volume / (double)np / (double)np;
    for (int i = 0; i < p->n_rbin; i++) xi_out[i] = norm * (0.0 + (double)(2 * p->ELL + 1) * xi[2 * i]);
    free(nm);
    free(xi);
    free(dn);
    return 0;
}

static void store_ylm_fourier(const or_param *p, int ell, int m, double *ylm) {
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                kvec_of(p, i, j, k, dk, kvec);
                or_ylm(ell, m, kvec, ylm + 2 * or_coord(p, i, j, k));
            }
}

static void store_ylm_config(const or_param *p, int ell, int m, double *ylm) {
    double rvec[3], dr[3];
    dr[0] = p->boxsize[0] / (double)(p->n_mesh[0]);
    dr[1] = p->boxsize[1] / (double)(p->n_mesh[1]);
    dr[2] = p->boxsize[2] / (double)(p->n_mesh[2]);
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                kvec_of(p, i, j, k, dr, rvec);
                or_ylm(ell, m, rvec, ylm + 2 * or_coord(p, i, j, k));
            }
}

/* i^n, exactly (the reference's pow(complex I, n) agrees to ~1e-16) */
static void ipow(int n, double *r, double *i) {
    switch (((n % 4) + 4) % 4) {
    case 0: *r = 1.0; *i = 0.0; break;
    case 1: *r = 0.0; *i = 1.0; break;
    case 2: *r = -1.0; *i = 0.0; break;
    default: *r = 0.0; *i = -1.0; break;
    }
}

/* The transformed fluctuation dn_k and density N_k the box estimators start
 * from: rho - nbar and rho (field.hpp:505-523, 560-572), or, for the
 * reconstructed box, D - alpha R and D + alpha^2 R with unit weights
 * (field.hpp:528-558, 574-605). */
static void box_fields(const or_param *p, const double *posD, long long nD, const double *posR, long long nR,
                       double alpha, int recon, double *dn_k, double *N_k) {
    if (!recon) {
        or_box_density_fluctuation(p, posD, nD, dn_k);
        or_box_density(p, posD, nD, N_k);
    } else {
        double *fR = grid_alloc(p);
        or_box_density(p, posD, nD, dn_k);
        or_box_density(p, posD, nD, N_k);
        or_box_density(p, posR, nR, fR);
        for (long long i = 0; i < p->n_mesh_tot; i++) {
            dn_k[2 * i] -= alpha * fR[2 * i];
            dn_k[2 * i + 1] -= alpha * fR[2 * i + 1];
            N_k[2 * i] += alpha * alpha * fR[2 * i];
            N_k[2 * i + 1] += alpha * alpha * fR[2 * i + 1];
        }
        free(fR);
    }
    or_fourier_transform(p, dn_k);
    or_fourier_transform(p, N_k);
}

/* bispec.hpp:877-1203 (box) and 1426-1753 (reconstructed box) from the
 * transformed fluctuation dn_k and density N_k: shotnoise is N or
 * N_D - alpha^3 N_R (field.hpp:1242-1250), n_norm N or N_D. */
static int box_bk_core(const or_param *p, const double *dn_k, const double *N_k, double shotnoise, double n_norm,
                       const double *kbin, double *bk_save, double *sn_save, double *out) {
    const int l1 = p->ell1, l2 = p->ell2, L = p->ELL, nk = p->n_kbin, ith = p->ith_kbin;
    if (fabs(or_wigner3j(l1, l2, L, 0, 0, 0)) < 1.0e-10) return -1;
    for (int i = 0; i < nk; i++) {
        sn_save[2 * i] = 0.0;
        sn_save[2 * i + 1] = 0.0;
        bk_save[2 * i] = 0.0;
        bk_save[2 * i + 1] = 0.0;
    }
    const long long ng = p->n_mesh_tot;
    double *pk = (double *)malloc(sizeof(double) * 2 * nk);
    int *nm = (int *)malloc(sizeof(int) * nk);

    /* shot noise: two-point terms (bispec.hpp:912-963) */
    {
        const double *dn_00_sn = dn_k, *N_LM_sn = N_k;
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                int M = 0;
                double w = or_wigner3j(l1, l2, L, 0, 0, 0) * or_wigner3j(l1, l2, L, m1, m2, M);
                w *= (double)(2 * L + 1) * (double)(2 * l1 + 1) * (double)(2 * l2 + 1);
                if (fabs(w) < 1.0e-10) continue;
                if ((l1 == 0) && (l2 == 0))
                    for (int i = 0; i < nk; i++) sn_save[2 * i] += w * shotnoise;
                if (l2 == 0) {
                    or_calc_power_spectrum(p, dn_00_sn, N_LM_sn, kbin, shotnoise, 0.0, l1, m1, pk, nm);
                    for (int i = 0; i < nk; i++) {
                        sn_save[2 * i] += w * pk[2 * ith];
                        sn_save[2 * i + 1] += w * pk[2 * ith + 1];
                    }
                }
                if (l1 == 0) {
                    or_calc_power_spectrum(p, dn_00_sn, N_LM_sn, kbin, shotnoise, 0.0, l2, m2, pk, nm);
                    for (int i = 0; i < nk; i++) {
                        sn_save[2 * i] += w * pk[2 * i];
                        sn_save[2 * i + 1] += w * pk[2 * i + 1];
                    }
                }
            }
    }

    /* shot noise: the ijk term (bispec.hpp:972-1073) */
    {
        const double *N_00_sn = N_k, *dn_LM_sn = dn_k;
        or_bessel *sj1 = or_bessel_new(l1);
        or_bessel *sj2 = or_bessel_new(l2);
        double *Ylm1 = grid_alloc(p), *Ylm2 = grid_alloc(p), *xi = grid_alloc(p);
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                int M = 0;
                double w = or_wigner3j(l1, l2, L, 0, 0, 0) * or_wigner3j(l1, l2, L, m1, m2, M);
                w *= (double)(2 * L + 1) * (double)(2 * l1 + 1) * (double)(2 * l2 + 1);
                if (fabs(w) < 1.0e-10) continue;
                store_ylm_config(p, l1, m1, Ylm1);
                store_ylm_config(p, l2, m2, Ylm2);
                memset(xi, 0, sizeof(double) * 2 * (size_t)ng);
                or_shotnoise_bispectrum_ijk(p, dn_LM_sn, N_00_sn, shotnoise, 0.0, xi);
                double dr[3];
                dr[0] = p->boxsize[0] / (double)(p->n_mesh[0]);
                dr[1] = p->boxsize[1] / (double)(p->n_mesh[1]);
                dr[2] = p->boxsize[2] / (double)(p->n_mesh[2]);
                for (int ik = 0; ik < nk; ik++) {
                    double kmag2 = kbin[ik];
                    double kmag1 = kbin[ith];
                    double sr = 0.0, si = 0.0;
                    double rvec[3];
                    for (int i = 0; i < p->n_mesh[0]; i++)
                        for (int j = 0; j < p->n_mesh[1]; j++)
                            for (int k = 0; k < p->n_mesh[2]; k++) {
                                long long coord = or_coord(p, i, j, k);
                                kvec_of(p, i, j, k, dr, rvec);
                                double rmag = sqrt(rvec[0] * rvec[0] + rvec[1] * rvec[1] + rvec[2] * rvec[2]);
                                double j1 = or_bessel_eval(sj1, kmag1 * rmag);
                                double j2 = or_bessel_eval(sj2, kmag2 * rmag);
                                double jj = j1 * j2;
                                double tr = jj * xi[2 * coord], ti = jj * xi[2 * coord + 1];
                                CMUL(tr, ti, Ylm1[2 * coord], Ylm1[2 * coord + 1], tr, ti);
                                CMUL(tr, ti, Ylm2[2 * coord], Ylm2[2 * coord + 1], tr, ti);
                                sr += tr;
                                si += ti;
                            }
                    double ir, ii;
                    ipow(l1 + l2, &ir, &ii);
                    double fac = p->volume / (double)(p->n_mesh_tot);
                    double cr = ir * fac, ci = ii * fac;
                    CMUL(sr, si, cr, ci, sr, si);
                    sn_save[2 * ik] += w * sr;
                    sn_save[2 * ik + 1] += w * si;
                }
            }
        or_bessel_free(sj1);
        or_bessel_free(sj2);
        free(Ylm1);
        free(Ylm2);
        free(xi);
    }

    /* bispectrum (bispec.hpp:1100-1182) */
    {
        const double *dn_00 = dn_k;
        double *dn_LM = grid_alloc(p);
        memcpy(dn_LM, dn_k, sizeof(double) * 2 * (size_t)ng);
        or_assignment_correction(p, dn_LM);
        or_inverse_fourier_transform(p, dn_LM);
        double *Ylm1 = grid_alloc(p), *Ylm2 = grid_alloc(p);
        double *t1 = grid_alloc(p), *t2 = grid_alloc(p);
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                int M = 0;
                double w = or_wigner3j(l1, l2, L, 0, 0, 0) * or_wigner3j(l1, l2, L, m1, m2, M);
                w *= (double)(2 * L + 1) * (double)(2 * l1 + 1) * (double)(2 * l2 + 1);
                if (fabs(w) < 1.0e-10) continue;
                store_ylm_fourier(p, l1, m1, Ylm1);
                store_ylm_fourier(p, l2, m2, Ylm2);
                double dk = kbin[1] - kbin[0];
                double kmag1 = kbin[ith];
                or_ifft_for_bispectrum(p, dn_00, kmag1, dk, Ylm1, t1);
                for (int ik = 0; ik < nk; ik++) {
                    double kmag2 = kbin[ik];
                    or_ifft_for_bispectrum(p, dn_00, kmag2, dk, Ylm2, t2);
                    double sr = 0.0, si = 0.0;
                    double fac = p->volume / (double)(p->n_mesh_tot);
                    for (long long c = 0; c < ng; c++) {
                        double ar = fac * t1[2 * c], ai = fac * t1[2 * c + 1];
                        CMUL(ar, ai, t2[2 * c], t2[2 * c + 1], ar, ai);
                        CMUL(ar, ai, dn_LM[2 * c], dn_LM[2 * c + 1], ar, ai);
                        sr += ar;
                        si += ai;
                    }
                    bk_save[2 * ik] += w * sr;
                    bk_save[2 * ik + 1] += w * si;
                }
            }
        free(Ylm1);
        free(Ylm2);
        free(t1);
        free(t2);
        free(dn_LM);
    }

    double norm = p->volume / n_norm / n_norm;
    norm *= (p->volume / n_norm);
    for (int i = 0; i < nk; i++) {
        out[2 * i] = norm * (bk_save[2 * i] - sn_save[2 * i]);
        out[2 * i + 1] = norm * (bk_save[2 * i + 1] - sn_save[2 * i + 1]);
    }
    free(pk);
    free(nm);
    return 0;
}

/* bispec.hpp:877-1203 */
int or_box_bk(const or_param *p, const double *pos, long long np, const double *kbin, double *bk_save,
              double *sn_save, double *out) {
    if (fabs(or_wigner3j(p->ell1, p->ell2, p->ELL, 0, 0, 0)) < 1.0e-10) return -1;
    double *dn_k = grid_alloc(p), *N_k = grid_alloc(p);
    box_fields(p, pos, np, NULL, 0, 0.0, 0, dn_k, N_k);
    int rc = box_bk_core(p, dn_k, N_k, (double)np, (double)np, kbin, bk_save, sn_save, out);
    free(dn_k);
    free(N_k);
    return rc;
}

/* bispec.hpp:1426-1753: posD, posR displaced and wrapped; alpha = N_D / N_R
 * as main_box.cpp:153 computes it. */
int or_box_recon_bk(const or_param *p, const double *posD, long long nD, const double *posR, long long nR,
                    double alpha, const double *kbin, double *bk_save, double *sn_save, double *out) {
    if (fabs(or_wigner3j(p->ell1, p->ell2, p->ELL, 0, 0, 0)) < 1.0e-10) return -1;
    double *dn_k = grid_alloc(p), *N_k = grid_alloc(p);
    box_fields(p, posD, nD, posR, nR, alpha, 1, dn_k, N_k);
    int rc = box_bk_core(p, dn_k, N_k, (double)nD - pow(alpha, 3) * (double)nR, (double)nD, kbin, bk_save,
                         sn_save, out);
    free(dn_k);
    free(N_k);
    return rc;
}


/* ------------------------------------------------------------------ */
/* survey (main.cpp:109-166, particle.hpp, field.hpp:186-228, 449-503,  */
/* 1164-1188, powerspec.hpp:4-91)                                       */
/* ------------------------------------------------------------------ */

/* particle.hpp:334-353 */
static double weight_sum(const double *w, long long n) {
    double s = 0.0;
    for (long long p = 0; p < n; p++) s += w[p];
    return s;
}

/* particle.hpp:355-375 */
static void calc_los(const double *pos, long long n, double *los) {
    for (long long p = 0; p < n; p++) {
        double m = sqrt(pos[3 * p] * pos[3 * p] + pos[3 * p + 1] * pos[3 * p + 1] + pos[3 * p + 2] * pos[3 * p + 2]);
        los[3 * p] = pos[3 * p] / m;
        los[3 * p + 1] = pos[3 * p + 1] / m;
        los[3 * p + 2] = pos[3 * p + 2] / m;
    }
}

/* particle.hpp:160-200 */
static void min_pos(const double *pos, long long n, double *mn) {
    mn[0] = pos[0];
    mn[1] = pos[1];
    mn[2] = pos[2];
    for (long long i = 0; i < n; i++)
        for (int a = 0; a < 3; a++)
            if (mn[a] > pos[3 * i + a]) mn[a] = pos[3 * i + a];
}

int or_survey_prepare(const or_param *p, double *posD, const double *wD, long long nD, double *posR,
                      const double *wR, long long nR, double *losD, double *losR, double *info) {
    double swD = weight_sum(wD, nD), swR = weight_sum(wR, nR);
    calc_los(posD, nD, losD);
    calc_los(posR, nR, losR);
    if (swR == 0 || swD == 0) return -1;               /* particle.hpp:274-276 */
    double alpha = swD / swR;                           /* particle.hpp:278 */
    double dP[3];                                       /* particle.hpp:222-265, factor = 1 */
    min_pos(posR, nR, dP);
    for (int a = 0; a < 3; a++) dP[a] -= 1.0 * p->boxsize[a] / (double)(p->n_mesh[a]);
    for (long long q = 0; q < nD; q++)
        for (int a = 0; a < 3; a++) posD[3 * q + a] -= dP[a];
    for (long long q = 0; q < nR; q++)
        for (int a = 0; a < 3; a++) posR[3 * q + a] -= dP[a];
    /* field.hpp:449-503 */
    double *w = (double *)calloc(2 * (size_t)nR, sizeof(double));
    for (long long q = 0; q < nR; q++) w[2 * q] = wR[q];
    double *f = (double *)malloc(sizeof(double) * 2 * (size_t)p->n_mesh_tot);
    or_calc_field(p, posR, w, nR, f);
    double dV = p->volume / (double)(p->n_mesh_tot);
    double norm = 0.0;
    for (long long i = 0; i < p->n_mesh_tot; i++) norm += f[2 * i] * f[2 * i] * dV;
    double vs = weight_sum(wR, nR) * weight_sum(wR, nR) / norm;
    free(f);
    free(w);
    /* particle.hpp:284-299 */
    double sw2 = 0.0;
    for (long long q = 0; q < nD; q++) sw2 += pow(wD[q], 2);
    double nbar = swD * swD / vs / sw2;
    info[0] = swD;
    info[1] = swR;
    info[2] = alpha;
    info[3] = vs;
    info[4] = nbar;
    info[5] = dP[0];
    info[6] = dP[1];
    info[7] = dP[2];
    return 0;
}

/* field.hpp:186-228: D with w Y_LM(los), minus alpha times R with w Y_LM(los) */
int or_ylm_weighted_fluctuation(const or_param *p, const double *posD, const double *losD, const double *wD,
                                long long nD, const double *posR, const double *losR, const double *wR, long long nR,
                                double alpha, int ell, int M, double *field) {
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)(nD > nR ? nD : nR));
    double *fR = (double *)malloc(sizeof(double) * 2 * (size_t)p->n_mesh_tot);
    double y[2];
    for (long long q = 0; q < nD; q++) {
        or_ylm(ell, M, losD + 3 * q, y);
        w[2 * q] = y[0] * wD[q];
        w[2 * q + 1] = y[1] * wD[q];
    }
    or_calc_field(p, posD, w, nD, field);
    for (long long q = 0; q < nR; q++) {
        or_ylm(ell, M, losR + 3 * q, y);
        w[2 * q] = y[0] * wR[q];
        w[2 * q + 1] = y[1] * wR[q];
    }
    or_calc_field(p, posR, w, nR, fR);
    for (long long i = 0; i < p->n_mesh_tot; i++) {
        field[2 * i] -= alpha * fR[2 * i];
        field[2 * i + 1] -= alpha * fR[2 * i + 1];
    }
    free(fR);
    free(w);
    return 0;
}

/* field.hpp:1164-1188 */
void or_shotnoise_pk(const double *losD, const double *wD, long long nD, const double *losR, const double *wR,
                     long long nR, double alpha, int ell, int M, double *out) {
    double dr = 0.0, di = 0.0, rr = 0.0, ri = 0.0, y[2];
    for (long long q = 0; q < nD; q++) {
        or_ylm(ell, M, losD + 3 * q, y);
        dr += pow(wD[q], 2) * y[0];
        di += pow(wD[q], 2) * y[1];
    }
    for (long long q = 0; q < nR; q++) {
        or_ylm(ell, M, losR + 3 * q, y);
        rr += pow(wR[q], 2) * y[0];
        ri += pow(wR[q], 2) * y[1];
    }
    out[0] = dr + pow(alpha, 2) * rr;
    out[1] = di + pow(alpha, 2) * ri;
}

/* powerspec.hpp:4-91: pk_save (2*n_kbin, complex) times norm, as written to pk%d.dat */
int or_survey_pk(const or_param *p, const double *posD, const double *losD, const double *wD, long long nD,
                 const double *posR, const double *losR, const double *wR, long long nR, double alpha, double vs,
                 const double *kbin, double *out, int *n_mode) {
    if (!((p->ELL == p->ell1) && (p->ell2 == 0))) return -1;
    const int nk = p->n_kbin;
    double *d00 = grid_alloc(p), *dLM = grid_alloc(p);
    or_ylm_weighted_fluctuation(p, posD, losD, wD, nD, posR, losR, wR, nR, alpha, 0, 0, d00);
    or_fourier_transform(p, d00);
    double *pk = (double *)calloc(2 * (size_t)nk, sizeof(double));
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)nk);
    for (int i = 0; i < 2 * nk; i++) out[i] = 0.0;
    for (int M = -p->ELL; M <= p->ELL; M++) {
        or_ylm_weighted_fluctuation(p, posD, losD, wD, nD, posR, losR, wR, nR, alpha, p->ELL, M, dLM);
        or_fourier_transform(p, dLM);
        double sn[2];
        or_shotnoise_pk(losD, wD, nD, losR, wR, nR, alpha, p->ELL, M, sn);
        for (int m1 = -p->ell1; m1 <= p->ell1; m1++) {
            double w = or_wigner3j(p->ell1, 0, p->ELL, 0, 0, 0) * or_wigner3j(p->ell1, 0, p->ELL, m1, 0, M);
            w *= (double)(2 * p->ELL + 1) * (double)(2 * p->ell1 + 1);
            if (fabs(w) < 1.0e-10) continue;
            or_calc_power_spectrum(p, dLM, d00, kbin, sn[0], sn[1], p->ell1, m1, tmp, n_mode);
            for (int i = 0; i < nk; i++) {
                pk[2 * i] += w * tmp[2 * i];
                pk[2 * i + 1] += w * tmp[2 * i + 1];
            }
        }
    }
    double norm = vs / weight_sum(wD, nD) / weight_sum(wD, nD); /* particle.hpp:302-313 */
    for (int i = 0; i < 2 * nk; i++) out[i] = norm * pk[i];
    free(tmp);
    free(pk);
    free(dLM);
    free(d00);
    return 0;
}


/* field.hpp:230-247: R alone with w Y_LM(los) */
static void ylm_weighted_density(const or_param *p, const double *posR, const double *losR, const double *wR,
                                 long long nR, int ell, int M, double *field) {
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)(nR > 0 ? nR : 1));
    double y[2];
    for (long long q = 0; q < nR; q++) {
        or_ylm(ell, M, losR + 3 * q, y);
        w[2 * q] = y[0] * wR[q];
        w[2 * q + 1] = y[1] * wR[q];
    }
    or_calc_field(p, posR, w, nR, field);
    free(w);
}

/* powerspec.hpp:93-179 (window = 0) and 182-270 (window = 1): norm * xi_save (2*n_rbin, complex);
 * the files hold the real part. */
int or_survey_xi(const or_param *p, const double *posD, const double *losD, const double *wD, long long nD,
                 const double *posR, const double *losR, const double *wR, long long nR, double alpha, double vs,
                 const double *rbin, int window, double *out, int *n_mode) {
    if (!((p->ELL == p->ell1) && (p->ell2 == 0))) return -1;
    const int nr = p->n_rbin;
    double *d00 = grid_alloc(p), *dLM = grid_alloc(p);
    if (window) ylm_weighted_density(p, posR, losR, wR, nR, 0, 0, d00);
    else or_ylm_weighted_fluctuation(p, posD, losD, wD, nD, posR, losR, wR, nR, alpha, 0, 0, d00);
    or_fourier_transform(p, d00);
    double *xs = (double *)calloc(2 * (size_t)nr, sizeof(double));
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)nr);
    for (int M = -p->ELL; M <= p->ELL; M++) {
        double sn[2];
        if (window) {
            ylm_weighted_density(p, posR, losR, wR, nR, p->ELL, M, dLM);
            /* field.hpp:1202-1218 */
            double y[2];
            sn[0] = sn[1] = 0.0;
            for (long long q = 0; q < nR; q++) {
                or_ylm(p->ELL, M, losR + 3 * q, y);
                sn[0] += pow(wR[q], 2) * y[0];
                sn[1] += pow(wR[q], 2) * y[1];
            }
        } else {
            or_ylm_weighted_fluctuation(p, posD, losD, wD, nD, posR, losR, wR, nR, alpha, p->ELL, M, dLM);
            or_shotnoise_pk(losD, wD, nD, losR, wR, nR, alpha, p->ELL, M, sn);
        }
        or_fourier_transform(p, dLM);
        for (int m1 = -p->ell1; m1 <= p->ell1; m1++) {
            double w = or_wigner3j(p->ell1, 0, p->ELL, 0, 0, 0) * or_wigner3j(p->ell1, 0, p->ELL, m1, 0, M);
            w *= (double)(2 * p->ELL + 1) * (double)(2 * p->ell1 + 1);
            if (fabs(w) < 1.0e-10) continue;
            or_calc_correlation_function(p, dLM, d00, rbin, sn[0], sn[1], p->ell1, m1, tmp, n_mode);
            for (int i = 0; i < nr; i++) {
                xs[2 * i] += w * tmp[2 * i];
                xs[2 * i + 1] += w * tmp[2 * i + 1];
            }
        }
    }
    const double sw = window ? weight_sum(wR, nR) : weight_sum(wD, nD);
    const double norm = vs / sw / sw; /* particle.hpp:302-313 on P_D, or on P_R for the window */
    for (int i = 0; i < 2 * nr; i++) out[i] = norm * xs[i];
    free(tmp);
    free(xs);
    free(dLM);
    free(d00);
    return 0;
}


/* ------------------------------------------------------------------ */
/* reconstruction (particle_reconstruction.hpp)                         */
/* ------------------------------------------------------------------ */

/* particle_reconstruction.hpp:56-112, 252-302: Psi_a(x) for a corrected
 * spectrum dk (delta/W, complex, n_mesh_tot): Psi(k) = -i (G/b) k_a/k^2 dk,
 * G = exp(-(k RG)^2/2), 0 for |k| < 1e-10 (left unwritten there, Q5), / V,
 * backward FFT.  psi: 3 complex grids. */
void or_recon_psi(const or_param *p, const double *dk, double b1, double RG, double *psi) {
    const long long ng = p->n_mesh_tot;
    double kvec[3], dkv[3];
    dkv[0] = 2.0 * M_PI / p->boxsize[0];
    dkv[1] = 2.0 * M_PI / p->boxsize[1];
    dkv[2] = 2.0 * M_PI / p->boxsize[2];
    for (long long i = 0; i < 6 * ng; i++) psi[i] = 0.0;
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dkv, kvec);
                double kmag = sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
                if (kmag < 1.0e-10) continue;
                double kmag2 = kmag * kmag;
                double W = exp(-pow(kmag * RG, 2) / 2.0);
                double fr = dk[2 * coord], fi = dk[2 * coord + 1];
                for (int a = 0; a < 3; a++) {
                    /* (-i) (W/b) (k_a/k^2) f: (c f_im, -c f_re) with c = (W/b)(k_a/k^2) */
                    double c = (W / b1) * (kvec[a] / kmag2);
                    psi[2 * (a * ng + coord)] = c * fi;
                    psi[2 * (a * ng + coord) + 1] = -(c * fr);
                }
            }
    for (long long i = 0; i < 6 * ng; i++) psi[i] /= p->volume;
    for (int a = 0; a < 3; a++) or_fft3d(p, psi + 2 * a * ng, +1);
}

/* particle_reconstruction.hpp:141-205, 334-388: TSC gather of Re Psi * norm
 * added to the positions, linear index range check only (Q1), no wrap. */
void or_recon_gather(const or_param *p, const double *psi, double norm, double *pos, long long np) {
    const long long ng = p->n_mesh_tot;
    for (long long q = 0; q < np; q++) {
        double w[3][3];
        int iw[3][3];
        for (int a = 0; a < 3; a++) {
            double xp = pos[3 * q + a] * (double)(p->n_mesh[a]) / p->boxsize[a];
            iw[1][a] = (int)(xp + 0.5);
            double dx = xp - (double)iw[1][a];
            w[0][a] = 0.5 * (0.5 - dx) * (0.5 - dx);
            w[1][a] = 0.75 - dx * dx;
            w[2][a] = 0.5 * (0.5 + dx) * (0.5 + dx);
            iw[0][a] = iw[1][a] - 1;
            iw[2][a] = iw[1][a] + 1;
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    long long coord = ((long long)iw[i][0] * p->n_mesh[1] + iw[j][1]) * p->n_mesh[2] + iw[k][2];
                    if (coord >= 0 && coord < ng) {
                        double ww = w[i][0] * w[j][1] * w[k][2];
                        for (int a = 0; a < 3; a++) pos[3 * q + a] += ww * (psi[2 * (a * ng + coord)] * norm);
                    }
                }
    }
}

/* particle_reconstruction.hpp:12-208 (weights given) and 210-402 (weights NULL: 1):
 * D - alpha R on p's mesh (the recon mesh), FFT, / W, Psi, gather into both catalogues. */
int or_recon_displace(const or_param *p, double *posD, const double *wD, long long nD, double *posR,
                      const double *wR, long long nR, double alpha, double norm, double b1, double RG) {
    const long long ng = p->n_mesh_tot;
    double *f = grid_alloc(p), *fR = grid_alloc(p);
    double *w = (double *)calloc(2 * (size_t)((nD > nR ? nD : nR) + 1), sizeof(double));
    for (long long q = 0; q < nD; q++) w[2 * q] = wD ? wD[q] : 1.0;
    or_calc_field(p, posD, w, nD, f);
    for (long long q = 0; q < nR; q++) w[2 * q] = wR ? wR[q] : 1.0;
    or_calc_field(p, posR, w, nR, fR);
    for (long long i = 0; i < ng; i++) {
        f[2 * i] -= alpha * fR[2 * i];
        f[2 * i + 1] -= alpha * fR[2 * i + 1];
    }
    or_fourier_transform(p, f);
    or_assignment_correction(p, f);
    double *psi = (double *)malloc(sizeof(double) * 6 * (size_t)ng);
    or_recon_psi(p, f, b1, RG, psi);
    or_recon_gather(p, psi, norm, posD, nD);
    or_recon_gather(p, psi, norm, posR, nR);
    free(psi);
    free(w);
    free(fR);
    free(f);
    return 0;
}

/* powerspec.hpp:396-456 (xi = 0) and 459-519 (xi = 1): D - alpha R with unit
 * weights, SN = N_D + alpha^2 N_R (field.hpp:1190-1197), norm V / N_D^2.
 * out: 2*n_kbin (P) or 2*n_rbin (xi, real part written). */
int or_box_recon_2pt(const or_param *p, const double *posD, long long nD, const double *posR, long long nR,
                     double alpha, const double *bins, int xi, double *out, int *n_mode) {
    if (!((p->ELL == p->ell1) && (p->ell2 == 0))) return -1;
    double *f = grid_alloc(p), *fR = grid_alloc(p);
    double *w = (double *)calloc(2 * (size_t)((nD > nR ? nD : nR) + 1), sizeof(double));
    for (long long q = 0; q < nD; q++) w[2 * q] = 1.0;
    or_calc_field(p, posD, w, nD, f);
    for (long long q = 0; q < nR; q++) w[2 * q] = 1.0;
    or_calc_field(p, posR, w, nR, fR);
    for (long long i = 0; i < p->n_mesh_tot; i++) {
        f[2 * i] -= alpha * fR[2 * i];
        f[2 * i + 1] -= alpha * fR[2 * i + 1];
    }
    or_fourier_transform(p, f);
    double sn = (double)nD + pow(alpha, 2) * (double)nR;
    int nb = xi ? p->n_rbin : p->n_kbin;
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)nb);
    if (xi) or_calc_correlation_function(p, f, f, bins, sn, 0.0, p->ELL, 0, tmp, n_mode);
    else or_calc_power_spectrum(p, f, f, bins, sn, 0.0, p->ELL, 0, tmp, n_mode);
    double norm = p->volume / (double)nD / (double)nD;
    for (int i = 0; i < nb; i++) {
        out[2 * i] = norm * (0.0 + (double)(2 * p->ELL + 1) * tmp[2 * i]);
        out[2 * i + 1] = norm * (0.0 + (double)(2 * p->ELL + 1) * tmp[2 * i + 1]);
    }
    free(tmp);
    free(w);
    free(fR);
    free(f);
    return 0;
}


/* ------------------------------------------------------------------ */
/* three-point correlation function (box)                               */
/* ------------------------------------------------------------------ */

/* field.hpp:403-445: IFFT of j_l(|k| r) Y(k) (d/W) / V, Y a complex grid (or NULL: 1) */
static void ifft_for_3pcf(const or_param *p, const double *density, double rmag, const double *ylm,
                          const or_bessel *sj, double *out) {
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dk, kvec);
                double kmag = sqrt(kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2]);
                double wf = or_assignment_function(p, kvec);
                double dr = density[2 * coord] / wf, di = density[2 * coord + 1] / wf;
                double yr = ylm ? ylm[2 * coord] : 1.0, yi = ylm ? ylm[2 * coord + 1] : 0.0;
                double tr, ti;
                CMUL(yr, yi, dr, di, tr, ti);
                double jl = or_bessel_eval(sj, kmag * rmag);
                out[2 * coord] = jl * tr / p->volume;
                out[2 * coord + 1] = jl * ti / p->volume;
            }
    or_fft3d(p, out, +1);
}

/* field.hpp:979-1113 with Y1 Y2 tables: returns xi[i] (complex) for the
 * 3PCF bin rule and the (-1)^(l1+l2) / n_mode^2 / dV scaling (Q14) */
static void corr_for_3pcf(const or_param *p, const double *f1, const double *f2, const double *rbin, double sn,
                          double sn_im,
                          const double *ylm1, const double *ylm2, double *xi, int *n_mode) {
    const long long ng = p->n_mesh_tot;
    double *x3 = (double *)calloc(2 * (size_t)ng, sizeof(double));
    double kvec[3], dk[3];
    dk[0] = 2.0 * M_PI / p->boxsize[0];
    dk[1] = 2.0 * M_PI / p->boxsize[1];
    dk[2] = 2.0 * M_PI / p->boxsize[2];
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dk, kvec);
                double pr, pi;
                CMUL(f1[2 * coord], f1[2 * coord + 1], f2[2 * coord], -f2[2 * coord + 1], pr, pi);
                double S = or_shotnoise_function(p, kvec);
                pr -= sn * S;
                pi -= sn_im * S;
                double w2 = pow(or_assignment_function(p, kvec), 2);
                pr /= w2;
                pi /= w2;
                x3[2 * coord] = pr / p->volume;
                x3[2 * coord + 1] = pi / p->volume;
            }
    or_fft3d(p, x3, +1);
    const int n_temp = 10000;
    const double dr_temp = 0.5;
    double *xt = (double *)calloc(2 * (size_t)n_temp, sizeof(double));
    int *nt = (int *)calloc((size_t)n_temp, sizeof(int));
    double rvec[3], dr[3];
    dr[0] = p->boxsize[0] / (double)(p->n_mesh[0]);
    dr[1] = p->boxsize[1] / (double)(p->n_mesh[1]);
    dr[2] = p->boxsize[2] / (double)(p->n_mesh[2]);
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dr, rvec);
                double rmag = sqrt(rvec[0] * rvec[0] + rvec[1] * rvec[1] + rvec[2] * rvec[2]);
                int rindex = (int)(rmag / dr_temp + 0.5);
                if (rindex < n_temp) {
                    double yr, yi, tr, ti;
                    CMUL(ylm1[2 * coord], ylm1[2 * coord + 1], ylm2[2 * coord], ylm2[2 * coord + 1], yr, yi);
                    CMUL(x3[2 * coord], x3[2 * coord + 1], yr, yi, tr, ti);
                    xt[2 * rindex] += tr;
                    xt[2 * rindex + 1] += ti;
                    nt[rindex]++;
                }
            }
    const int nr = p->n_rbin;
    double *drb = (double *)malloc(sizeof(double) * (size_t)(nr > 1 ? nr - 1 : 1));
    for (int j = 0; j < nr - 1; j++) drb[j] = rbin[j + 1] - rbin[j];
    for (int j = 0; j < nr; j++) {
        double lower = (j == 0) ? ((rbin[j] > drb[j] * 0.5) ? (rbin[j] - drb[j] * 0.5) : 0.0) : rbin[j] - drb[j - 1] * 0.5;
        double upper = (j == nr - 1) ? (rbin[j] + drb[j - 1] * 0.5) : (rbin[j] + drb[j] * 0.5);
        xi[2 * j] = 0.0;
        xi[2 * j + 1] = 0.0;
        n_mode[j] = 0;
        for (int i = 0; i < n_temp; i++) {
            double rt = (double)i * dr_temp;
            if ((rt > lower) && (rt <= upper)) {
                xi[2 * j] += xt[2 * i];
                xi[2 * j + 1] += xt[2 * i + 1];
                n_mode[j] += nt[i];
            }
        }
    }
    double dV = p->volume / (double)ng;
    for (int j = 0; j < nr; j++) {
        if (n_mode[j] != 0) {
            double f = pow(-1.0, p->ell1 + p->ell2) / (double)n_mode[j] / (double)n_mode[j] / dV;
            xi[2 * j] *= f;
            xi[2 * j + 1] *= f;
        } else {
            xi[2 * j] = 0.0;
            xi[2 * j + 1] = 0.0;
        }
    }
    free(drb);
    free(xt);
    free(nt);
    free(x3);
}

/* bispec.hpp:1205-1424 (box) and 1755-1975 (reconstructed box) for
 * p->ith_rbin from the transformed fluctuation dn_k and density N_k:
 * zeta_save, sn_save (2*n_rbin) and out = norm * (zeta - sn) (real part
 * written to zeta%d%d%d_%02d[_recon].dat). */
static int box_zeta_core(const or_param *p, const double *dn_k, const double *N_k, double shotnoise, double n_norm,
                         const double *rbin, double *zeta_save, double *sn_save, double *out) {
    const int l1 = p->ell1, l2 = p->ell2, L = p->ELL, nr = p->n_rbin, ith = p->ith_rbin;
    if (fabs(or_wigner3j(l1, l2, L, 0, 0, 0)) < 1.0e-10) return -1;
    const long long ng = p->n_mesh_tot;
    for (int i = 0; i < 2 * nr; i++) {
        zeta_save[i] = 0.0;
        sn_save[i] = 0.0;
    }
    double *Y1 = grid_alloc(p), *Y2 = grid_alloc(p);
    int *nm = (int *)malloc(sizeof(int) * nr);
    double *tmp = (double *)malloc(sizeof(double) * 2 * nr);
    /* shot noise (bispec.hpp:1229-1290) */
    {
        const double *N00 = N_k, *dnLM = dn_k;
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                double w = or_wigner3j(l1, l2, L, 0, 0, 0) * or_wigner3j(l1, l2, L, m1, m2, 0);
                w *= (double)(2 * L + 1) * (double)(2 * l1 + 1) * (double)(2 * l2 + 1);
                if (fabs(w) < 1.0e-10) continue;
                store_ylm_config(p, l1, m1, Y1);
                store_ylm_config(p, l2, m2, Y2);
                corr_for_3pcf(p, dnLM, N00, rbin, shotnoise, 0.0, Y1, Y2, tmp, nm);
                sn_save[2 * ith] += w * tmp[2 * ith];
                sn_save[2 * ith + 1] += w * tmp[2 * ith + 1];
            }
    }
    /* three-point function (bispec.hpp:1300-1400) */
    {
        const double *dn00 = dn_k;
        double *dnLM = grid_alloc(p), *t1 = grid_alloc(p), *t2 = grid_alloc(p);
        memcpy(dnLM, dn_k, sizeof(double) * 2 * (size_t)ng);
        or_assignment_correction(p, dnLM);
        or_inverse_fourier_transform(p, dnLM);
        or_bessel *sj1 = or_bessel_new(l1), *sj2 = or_bessel_new(l2);
        const double fac = p->volume / (double)ng;
        double ir, ii;
        ipow(l1 + l2, &ir, &ii);
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                double w = or_wigner3j(l1, l2, L, 0, 0, 0) * or_wigner3j(l1, l2, L, m1, m2, 0);
                w *= (double)(2 * L + 1) * (double)(2 * l1 + 1) * (double)(2 * l2 + 1);
                if (fabs(w) < 1.0e-10) continue;
                store_ylm_fourier(p, l1, m1, Y1);
                store_ylm_fourier(p, l2, m2, Y2);
                ifft_for_3pcf(p, dn00, rbin[ith], Y1, sj1, t1);
                for (int ir2 = 0; ir2 < nr; ir2++) {
                    ifft_for_3pcf(p, dn00, rbin[ir2], Y2, sj2, t2);
                    double sr = 0.0, si = 0.0;
                    for (long long c = 0; c < ng; c++) {
                        /* pow(I, l1+l2) * fac * f1 * f2 * f3, left to right */
                        double ar = ir * fac, ai = ii * fac;
                        CMUL(ar, ai, t1[2 * c], t1[2 * c + 1], ar, ai);
                        CMUL(ar, ai, t2[2 * c], t2[2 * c + 1], ar, ai);
                        CMUL(ar, ai, dnLM[2 * c], dnLM[2 * c + 1], ar, ai);
                        sr += ar;
                        si += ai;
                    }
                    zeta_save[2 * ir2] += w * sr;
                    zeta_save[2 * ir2 + 1] += w * si;
                }
            }
        or_bessel_free(sj1);
        or_bessel_free(sj2);
        free(dnLM);
        free(t1);
        free(t2);
    }
    double norm = p->volume / n_norm / n_norm;
    norm *= (p->volume / n_norm);
    for (int i = 0; i < nr; i++) {
        out[2 * i] = norm * (zeta_save[2 * i] - sn_save[2 * i]);
        out[2 * i + 1] = norm * (zeta_save[2 * i + 1] - sn_save[2 * i + 1]);
    }
    free(Y1);
    free(Y2);
    free(nm);
    free(tmp);
    return 0;
}

/* bispec.hpp:1205-1424 */
int or_box_zeta(const or_param *p, const double *pos, long long np, const double *rbin, double *zeta_save,
                double *sn_save, double *out) {
    if (fabs(or_wigner3j(p->ell1, p->ell2, p->ELL, 0, 0, 0)) < 1.0e-10) return -1;
    double *dn_k = grid_alloc(p), *N_k = grid_alloc(p);
    box_fields(p, pos, np, NULL, 0, 0.0, 0, dn_k, N_k);
    int rc = box_zeta_core(p, dn_k, N_k, (double)np, (double)np, rbin, zeta_save, sn_save, out);
    free(dn_k);
    free(N_k);
    return rc;
}

/* bispec.hpp:1755-1975: posD, posR displaced and wrapped; alpha = N_D / N_R. */
int or_box_recon_zeta(const or_param *p, const double *posD, long long nD, const double *posR, long long nR,
                      double alpha, const double *rbin, double *zeta_save, double *sn_save, double *out) {
    if (fabs(or_wigner3j(p->ell1, p->ell2, p->ELL, 0, 0, 0)) < 1.0e-10) return -1;
    double *dn_k = grid_alloc(p), *N_k = grid_alloc(p);
    box_fields(p, posD, nD, posR, nR, alpha, 1, dn_k, N_k);
    int rc = box_zeta_core(p, dn_k, N_k, (double)nD - pow(alpha, 3) * (double)nR, (double)nD, rbin, zeta_save,
                           sn_save, out);
    free(dn_k);
    free(N_k);
    return rc;
}


/* ------------------------------------------------------------------ */
/* survey bispectrum and three-point functions (bispec.hpp:4-874)       */
/* ------------------------------------------------------------------ */

/* The catalogues of a survey estimator: D and R (shifted positions, lines
 * of sight, weights) with alpha, or R alone for the window functions. */
typedef struct {
    const double *posD, *losD, *wD, *posR, *losR, *wR;
    long long nD, nR;
    double alpha;
    int window;
} svy_src;

/* Y_LM(los) w^k-weighted field of one catalogue, Y conjugated when conj. */
static void svy_weighted(const or_param *p, const double *pos, const double *los, const double *w, long long n,
                         int L, int M, int k, int conj, double *field) {
    double *wt = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    double y[2];
    for (long long q = 0; q < n; q++) {
        or_ylm(L, M, los + 3 * q, y);
        if (conj) y[1] = -y[1];
        wt[2 * q] = y[0] * pow(w[q], k);
        wt[2 * q + 1] = y[1] * pow(w[q], k);
    }
    or_calc_field(p, pos, wt, n, field);
    free(wt);
}

/* field.hpp:186-247: the Y_LM-weighted fluctuation D - alpha R, or R alone */
static void svy_fluct(const or_param *p, const svy_src *c, int L, int M, double *field) {
    if (c->window) svy_weighted(p, c->posR, c->losR, c->wR, c->nR, L, M, 1, 0, field);
    else or_ylm_weighted_fluctuation(p, c->posD, c->losD, c->wD, c->nD, c->posR, c->losR, c->wR, c->nR, c->alpha, L, M,
                                     field);
}

/* field.hpp:249-306: conj(Y_LM) w^2 of D plus alpha^2 that of R, or of R alone */
static void svy_sn_density(const or_param *p, const svy_src *c, int L, int M, double *field) {
    if (c->window) {
        svy_weighted(p, c->posR, c->losR, c->wR, c->nR, L, M, 2, 1, field);
        return;
    }
    double *fR = grid_alloc(p);
    svy_weighted(p, c->posD, c->losD, c->wD, c->nD, L, M, 2, 1, field);
    svy_weighted(p, c->posR, c->losR, c->wR, c->nR, L, M, 2, 1, fR);
    for (long long i = 0; i < p->n_mesh_tot; i++) {
        field[2 * i] += c->alpha * c->alpha * fR[2 * i];
        field[2 * i + 1] += c->alpha * c->alpha * fR[2 * i + 1];
    }
    free(fR);
}

/* field.hpp:1202-1240: sum_D w^3 Y - alpha^3 sum_R w^3 Y, or sum_R w^2 Y for the window */
static void svy_shotnoise(const svy_src *c, int L, int M, double *out) {
    double y[2];
    out[0] = out[1] = 0.0;
    if (c->window) {
        for (long long q = 0; q < c->nR; q++) {
            or_ylm(L, M, c->losR + 3 * q, y);
            out[0] += pow(c->wR[q], 2) * y[0];
            out[1] += pow(c->wR[q], 2) * y[1];
        }
        return;
    }
    double dr = 0.0, di = 0.0, rr = 0.0, ri = 0.0;
    for (long long q = 0; q < c->nD; q++) {
        or_ylm(L, M, c->losD + 3 * q, y);
        dr += pow(c->wD[q], 3) * y[0];
        di += pow(c->wD[q], 3) * y[1];
    }
    for (long long q = 0; q < c->nR; q++) {
        or_ylm(L, M, c->losR + 3 * q, y);
        rr += pow(c->wR[q], 3) * y[0];
        ri += pow(c->wR[q], 3) * y[1];
    }
    out[0] = dr - pow(c->alpha, 3) * rr;
    out[1] = di - pow(c->alpha, 3) * ri;
}

static double w3j_weight(int l1, int l2, int L, int m1, int m2, int M) {
    double w = or_wigner3j(l1, l2, L, 0, 0, 0) * or_wigner3j(l1, l2, L, m1, m2, M);
    return w * (double)(2 * L + 1) * (double)(2 * l1 + 1) * (double)(2 * l2 + 1);
}

/* bispec.hpp:4-353 for p->ith_kbin: bk_save, sn_save (2*n_kbin complex) and
 * out = norm * (bk - sn), norm = Vsurvey^2 / (sum w_D)^3 (particle.hpp:319-332). */
int or_survey_bk(const or_param *p, const double *posD, const double *losD, const double *wD, long long nD,
                 const double *posR, const double *losR, const double *wR, long long nR, double alpha, double vs,
                 const double *kbin, double *bk_save, double *sn_save, double *out) {
    const int l1 = p->ell1, l2 = p->ell2, L = p->ELL, nk = p->n_kbin, ith = p->ith_kbin;
    if (fabs(or_wigner3j(l1, l2, L, 0, 0, 0)) < 1.0e-10) return -1;
    svy_src c = {posD, losD, wD, posR, losR, wR, nD, nR, alpha, 0};
    const long long ng = p->n_mesh_tot;
    for (int i = 0; i < 2 * nk; i++) bk_save[i] = sn_save[i] = 0.0;
    double *pk = (double *)malloc(sizeof(double) * 2 * nk);
    int *nm = (int *)malloc(sizeof(int) * nk);
    double sn[2];
    /* two-point shot-noise terms (bispec.hpp:33-93) */
    {
        double *dn_00_sn = grid_alloc(p), *N_LM_sn = grid_alloc(p);
        svy_fluct(p, &c, 0, 0, dn_00_sn);
        or_fourier_transform(p, dn_00_sn);
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++)
                for (int M = -L; M <= L; M++) {
                    double w = w3j_weight(l1, l2, L, m1, m2, M);
                    if (fabs(w) < 1.0e-10) continue;
                    svy_sn_density(p, &c, L, M, N_LM_sn);
                    or_fourier_transform(p, N_LM_sn);
                    svy_shotnoise(&c, L, M, sn);
                    if ((l1 == 0) && (l2 == 0))
                        for (int i = 0; i < nk; i++) {
                            sn_save[2 * i] += w * sn[0];
                            sn_save[2 * i + 1] += w * sn[1];
                        }
                    if (l2 == 0) {
                        or_calc_power_spectrum(p, dn_00_sn, N_LM_sn, kbin, sn[0], sn[1], l1, m1, pk, nm);
                        for (int i = 0; i < nk; i++) {
                            sn_save[2 * i] += w * pk[2 * ith];
                            sn_save[2 * i + 1] += w * pk[2 * ith + 1];
                        }
                    }
                    if (l1 == 0) {
                        or_calc_power_spectrum(p, dn_00_sn, N_LM_sn, kbin, sn[0], sn[1], l2, m2, pk, nm);
                        for (int i = 0; i < nk; i++) {
                            sn_save[2 * i] += w * pk[2 * i];
                            sn_save[2 * i + 1] += w * pk[2 * i + 1];
                        }
                    }
                }
        free(dn_00_sn);
        free(N_LM_sn);
    }
    /* the ijk term (bispec.hpp:95-211) */
    {
        double *N_00_sn = grid_alloc(p), *dn_LM_sn = grid_alloc(p), *xi = grid_alloc(p);
        double *Y1 = grid_alloc(p), *Y2 = grid_alloc(p);
        svy_sn_density(p, &c, 0, 0, N_00_sn);
        or_fourier_transform(p, N_00_sn);
        or_bessel *sj1 = or_bessel_new(l1), *sj2 = or_bessel_new(l2);
        double dr[3];
        for (int a = 0; a < 3; a++) dr[a] = p->boxsize[a] / (double)(p->n_mesh[a]);
        double ir, ii;
        ipow(l1 + l2, &ir, &ii);
        const double fac = p->volume / (double)ng;
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                store_ylm_config(p, l1, m1, Y1);
                store_ylm_config(p, l2, m2, Y2);
                for (int M = -L; M <= L; M++) {
                    double w = w3j_weight(l1, l2, L, m1, m2, M);
                    if (fabs(w) < 1.0e-10) continue;
                    svy_fluct(p, &c, L, M, dn_LM_sn);
                    or_fourier_transform(p, dn_LM_sn);
                    svy_shotnoise(&c, L, M, sn);
                    or_shotnoise_bispectrum_ijk(p, dn_LM_sn, N_00_sn, sn[0], sn[1], xi);
                    for (int ik = 0; ik < nk; ik++) {
                        double sr = 0.0, si = 0.0, rvec[3];
                        for (int i = 0; i < p->n_mesh[0]; i++)
                            for (int j = 0; j < p->n_mesh[1]; j++)
                                for (int k = 0; k < p->n_mesh[2]; k++) {
                                    long long coord = or_coord(p, i, j, k);
                                    kvec_of(p, i, j, k, dr, rvec);
                                    double rmag = sqrt(rvec[0] * rvec[0] + rvec[1] * rvec[1] + rvec[2] * rvec[2]);
                                    double jj = or_bessel_eval(sj1, kbin[ith] * rmag) * or_bessel_eval(sj2, kbin[ik] * rmag);
                                    double tr = jj * xi[2 * coord], ti = jj * xi[2 * coord + 1];
                                    CMUL(tr, ti, Y1[2 * coord], Y1[2 * coord + 1], tr, ti);
                                    CMUL(tr, ti, Y2[2 * coord], Y2[2 * coord + 1], tr, ti);
                                    sr += tr;
                                    si += ti;
                                }
                        double cr = ir * fac, ci = ii * fac;
                        CMUL(sr, si, cr, ci, sr, si);
                        sn_save[2 * ik] += w * sr;
                        sn_save[2 * ik + 1] += w * si;
                    }
                }
            }
        or_bessel_free(sj1);
        or_bessel_free(sj2);
        free(Y1);
        free(Y2);
        free(xi);
        free(dn_LM_sn);
        free(N_00_sn);
    }
    /* the bispectrum (bispec.hpp:216-323) */
    {
        double *dn_00 = grid_alloc(p), *dn_LM = grid_alloc(p), *Y1 = grid_alloc(p), *Y2 = grid_alloc(p);
        double *t1 = grid_alloc(p), *t2 = grid_alloc(p);
        svy_fluct(p, &c, 0, 0, dn_00);
        or_fourier_transform(p, dn_00);
        const double dk = kbin[1] - kbin[0], fac = p->volume / (double)ng;
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                store_ylm_fourier(p, l1, m1, Y1);
                store_ylm_fourier(p, l2, m2, Y2);
                for (int M = -L; M <= L; M++) {
                    double w = w3j_weight(l1, l2, L, m1, m2, M);
                    if (fabs(w) < 1.0e-10) continue;
                    svy_fluct(p, &c, L, M, dn_LM);
                    or_fourier_transform(p, dn_LM);
                    or_assignment_correction(p, dn_LM);
                    or_inverse_fourier_transform(p, dn_LM);
                    or_ifft_for_bispectrum(p, dn_00, kbin[ith], dk, Y1, t1);
                    for (int ik = 0; ik < nk; ik++) {
                        or_ifft_for_bispectrum(p, dn_00, kbin[ik], dk, Y2, t2);
                        double sr = 0.0, si = 0.0;
                        for (long long q = 0; q < ng; q++) {
                            double ar = fac * t1[2 * q], ai = fac * t1[2 * q + 1];
                            CMUL(ar, ai, t2[2 * q], t2[2 * q + 1], ar, ai);
                            CMUL(ar, ai, dn_LM[2 * q], dn_LM[2 * q + 1], ar, ai);
                            sr += ar;
                            si += ai;
                        }
                        bk_save[2 * ik] += w * sr;
                        bk_save[2 * ik + 1] += w * si;
                    }
                }
            }
        free(dn_00);
        free(dn_LM);
        free(Y1);
        free(Y2);
        free(t1);
        free(t2);
    }
    const double sw = weight_sum(wD, nD);
    const double norm = vs / sw / sw * (vs / sw);
    for (int i = 0; i < nk; i++) {
        out[2 * i] = norm * (bk_save[2 * i] - sn_save[2 * i]);
        out[2 * i + 1] = norm * (bk_save[2 * i + 1] - sn_save[2 * i + 1]);
    }
    free(pk);
    free(nm);
    return 0;
}

/* bispec.hpp:355-614 (window = 0) and 617-874 (window = 1) for p->ith_rbin:
 * zeta_save, sn_save and out = norm * (zeta - sn), norm = Vsurvey^2 / (sum w)^3
 * of D, or of R for the window. */
int or_survey_zeta(const or_param *p, const double *posD, const double *losD, const double *wD, long long nD,
                   const double *posR, const double *losR, const double *wR, long long nR, double alpha, double vs,
                   const double *rbin, int window, double *zeta_save, double *sn_save, double *out) {
    const int l1 = p->ell1, l2 = p->ell2, L = p->ELL, nr = p->n_rbin, ith = p->ith_rbin;
    if (fabs(or_wigner3j(l1, l2, L, 0, 0, 0)) < 1.0e-10) return -1;
    svy_src c = {posD, losD, wD, posR, losR, wR, nD, nR, alpha, window};
    const long long ng = p->n_mesh_tot;
    for (int i = 0; i < 2 * nr; i++) zeta_save[i] = sn_save[i] = 0.0;
    double *Y1 = grid_alloc(p), *Y2 = grid_alloc(p);
    int *nm = (int *)malloc(sizeof(int) * nr);
    double *tmp = (double *)malloc(sizeof(double) * 2 * nr);
    double sn[2];
    /* shot noise (bispec.hpp:388-460, 644-720) */
    {
        double *N00 = grid_alloc(p), *dnLM = grid_alloc(p);
        svy_sn_density(p, &c, 0, 0, N00);
        or_fourier_transform(p, N00);
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                store_ylm_config(p, l1, m1, Y1);
                store_ylm_config(p, l2, m2, Y2);
                for (int M = -L; M <= L; M++) {
                    double w = w3j_weight(l1, l2, L, m1, m2, M);
                    if (fabs(w) < 1.0e-10) continue;
                    svy_fluct(p, &c, L, M, dnLM);
                    or_fourier_transform(p, dnLM);
                    svy_shotnoise(&c, L, M, sn);
                    corr_for_3pcf(p, dnLM, N00, rbin, sn[0], sn[1], Y1, Y2, tmp, nm);
                    sn_save[2 * ith] += w * tmp[2 * ith];
                    sn_save[2 * ith + 1] += w * tmp[2 * ith + 1];
                }
            }
        free(N00);
        free(dnLM);
    }
    /* the three-point function (bispec.hpp:470-590, 726-850) */
    {
        double *dn00 = grid_alloc(p), *dnLM = grid_alloc(p), *t1 = grid_alloc(p), *t2 = grid_alloc(p);
        svy_fluct(p, &c, 0, 0, dn00);
        or_fourier_transform(p, dn00);
        or_bessel *sj1 = or_bessel_new(l1), *sj2 = or_bessel_new(l2);
        const double fac = p->volume / (double)ng;
        double ir, ii;
        ipow(l1 + l2, &ir, &ii);
        for (int m1 = -l1; m1 <= l1; m1++)
            for (int m2 = -l2; m2 <= l2; m2++) {
                store_ylm_fourier(p, l1, m1, Y1);
                store_ylm_fourier(p, l2, m2, Y2);
                for (int M = -L; M <= L; M++) {
                    double w = w3j_weight(l1, l2, L, m1, m2, M);
                    if (fabs(w) < 1.0e-10) continue;
                    svy_fluct(p, &c, L, M, dnLM);
                    or_fourier_transform(p, dnLM);
                    or_assignment_correction(p, dnLM);
                    or_inverse_fourier_transform(p, dnLM);
                    ifft_for_3pcf(p, dn00, rbin[ith], Y1, sj1, t1);
                    for (int ir2 = 0; ir2 < nr; ir2++) {
                        ifft_for_3pcf(p, dn00, rbin[ir2], Y2, sj2, t2);
                        double sr = 0.0, si = 0.0;
                        for (long long q = 0; q < ng; q++) {
                            double ar = ir * fac, ai = ii * fac;
                            CMUL(ar, ai, t1[2 * q], t1[2 * q + 1], ar, ai);
                            CMUL(ar, ai, t2[2 * q], t2[2 * q + 1], ar, ai);
                            CMUL(ar, ai, dnLM[2 * q], dnLM[2 * q + 1], ar, ai);
                            sr += ar;
                            si += ai;
                        }
                        zeta_save[2 * ir2] += w * sr;
                        zeta_save[2 * ir2 + 1] += w * si;
                    }
                }
            }
        or_bessel_free(sj1);
        or_bessel_free(sj2);
        free(dn00);
        free(dnLM);
        free(t1);
        free(t2);
    }
    const double sw = window ? weight_sum(wR, nR) : weight_sum(wD, nD);
    const double norm = vs / sw / sw * (vs / sw);
    for (int i = 0; i < nr; i++) {
        out[2 * i] = norm * (zeta_save[2 * i] - sn_save[2 * i]);
        out[2 * i + 1] = norm * (zeta_save[2 * i + 1] - sn_save[2 * i + 1]);
    }
    free(Y1);
    free(Y2);
    free(nm);
    free(tmp);
    return 0;
}


/* ------------------------------------------------------------------ */
/* interlacing (no counterpart in the reference; off by default)        */
/* ------------------------------------------------------------------ */

/* The transformed box fluctuation of the positions and of the positions
 * shifted by half a cell on every axis (wrapped into the box), combined as
 * (d_a(k) + exp(i k.h) d_b(k)) / 2, h = (L_a / 2 n_a), k with the
 * Nyquist index negative as everywhere else, for k_z index <= n_z / 2; the
 * rest by conjugate symmetry. */
int or_box_interlaced_k(const or_param *p, const double *pos, long long np, double *out) {
    double *fb = grid_alloc(p);
    double *ps = (double *)malloc(sizeof(double) * 3 * (size_t)(np > 0 ? np : 1));
    double h[3], dk[3], kvec[3];
    for (int a = 0; a < 3; a++) {
        h[a] = 0.5 * p->boxsize[a] / (double)p->n_mesh[a];
        dk[a] = 2.0 * M_PI / p->boxsize[a];
    }
    for (long long q = 0; q < np; q++)
        for (int a = 0; a < 3; a++) {
            double x = pos[3 * q + a] + h[a];
            if (x >= p->boxsize[a]) x -= p->boxsize[a];
            ps[3 * q + a] = x;
        }
    or_box_density_fluctuation(p, pos, np, out);
    or_fourier_transform(p, out);
    or_box_density_fluctuation(p, ps, np, fb);
    or_fourier_transform(p, fb);
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long c = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dk, kvec);
                double ph = kvec[0] * h[0] + kvec[1] * h[1] + kvec[2] * h[2];
                double cr = cos(ph), ci = sin(ph), br, bi;
                CMUL(cr, ci, fb[2 * c], fb[2 * c + 1], br, bi);
                out[2 * c] = 0.5 * (out[2 * c] + br);
                out[2 * c + 1] = 0.5 * (out[2 * c + 1] + bi);
            }
    /* the interlaced field is real: modes with k_z index above n_z / 2 are the
     * conjugates of their mirrors, which fixes the phase of the Nyquist-index
     * modes the way a half-spectrum (real-to-complex) layout does */
    const int n0 = p->n_mesh[0], n1 = p->n_mesh[1], n2 = p->n_mesh[2];
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++)
            for (int k = n2 / 2 + 1; k < n2; k++) {
                long long c = or_coord(p, i, j, k), cm = or_coord(p, (n0 - i) % n0, (n1 - j) % n1, n2 - k);
                out[2 * c] = out[2 * cm];
                out[2 * c + 1] = -out[2 * cm + 1];
            }
    free(ps);
    free(fb);
    return 0;
}

/* powerspec.hpp:273-393 on the interlaced fluctuation with the shot noise
 * N W(k)^2: P (xi = 0, 2*n_kbin) or xi (xi = 1, 2*n_rbin) as written to
 * pk%d.dat / xi%d.dat. */
int or_box_interlaced_2pt(const or_param *p, const double *pos, long long np, const double *bins, int xi,
                          double *out, int *n_mode) {
    double *f = grid_alloc(p);
    or_box_interlaced_k(p, pos, np, f);
    int nb = xi ? p->n_rbin : p->n_kbin;
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)nb);
    g_sn_shape_w2 = 1;
    if (xi) or_calc_correlation_function(p, f, f, bins, (double)np, 0.0, p->ELL, 0, tmp, n_mode);
    else or_calc_power_spectrum(p, f, f, bins, (double)np, 0.0, p->ELL, 0, tmp, n_mode);
    g_sn_shape_w2 = 0;
    double norm = p->volume / (double)np / (double)np;
    for (int i = 0; i < nb; i++) {
        out[2 * i] = norm * (0.0 + (double)(2 * p->ELL + 1) * tmp[2 * i]);
        out[2 * i + 1] = norm * (0.0 + (double)(2 * p->ELL + 1) * tmp[2 * i + 1]);
    }
    free(tmp);
    free(f);
    return 0;
}

/* ------------------------------------------------------------------------
 * Mock generator (no counterpart in the reference).
 *
 * Philox4x32-10 as published by Salmon, Moraes, Dror and Shaw, "Parallel
 * random numbers: as easy as 1, 2, 3" (SC 2011), section 3.3: ten rounds of
 *   (L0, R0, L1, R1) -> (hi(R1 * M1) ^ L0' ^ k0, lo(R1 * M1), hi(R0 * M0) ^ L1' ^ k1, lo(R0 * M0))
 * in the Random123 word order, the key bumped by the Weyl constants between
 * rounds.  Pinned by the Random123 known-answer vectors (tests).
 * ------------------------------------------------------------------------ */
#include <stdint.h>

void or_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t a = (uint64_t)0xD2511F53u * (uint64_t)x0;
        const uint64_t b = (uint64_t)0xCD9E8D57u * (uint64_t)x2;
        const uint32_t y0 = (uint32_t)(b >> 32) ^ x1 ^ k0;
        const uint32_t y1 = (uint32_t)b;
        const uint32_t y2 = (uint32_t)(a >> 32) ^ x3 ^ k1;
        const uint32_t y3 = (uint32_t)a;
        x0 = y0;
        x1 = y1;
        x2 = y2;
        x3 = y3;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

/* Standard normal of cell i: counter (i lo, i hi, 0, 0), key (seed lo, seed
 * hi); u1 = (w0 >> 5, w1 >> 6) as a 53-bit fraction plus 2^-54 (never 0),
 * u2 likewise from w2, w3 (no offset); sqrt(-2 ln u1) cos(2 pi u2). */
void or_mock_white_noise(long long n, unsigned long long seed, double *out) {
    const unsigned int key[2] = {(unsigned int)(seed & 0xFFFFFFFFu), (unsigned int)(seed >> 32)};
    for (long long i = 0; i < n; i++) {
        const unsigned long long u = (unsigned long long)i;
        const unsigned int ctr[4] = {(unsigned int)(u & 0xFFFFFFFFu), (unsigned int)(u >> 32), 0u, 0u};
        unsigned int w[4];
        or_philox4x32_10(ctr, key, w);
        const double f1 = (double)(((uint64_t)(w[0] >> 5) << 26) | (uint64_t)(w[1] >> 6));
        const double f2 = (double)(((uint64_t)(w[2] >> 5) << 26) | (uint64_t)(w[3] >> 6));
        const double u1 = f1 / 9007199254740992.0 + 1.0 / 18014398509481984.0;
        const double u2 = f2 / 9007199254740992.0;
        out[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

/* splitmix64 (Steele, Lea and Flood, OOPSLA 2014; Vigna's constants): the
 * state advances by the golden-ratio increment and is then mixed. */
static uint64_t or_splitmix(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t or_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Sites of particles 0..np-1: all sites in order when np == sites, else a
 * balanced Feistel network on 2h bits (2^(2h) >= sites, h minimal), four
 * rounds (l, r) -> (r, l ^ (mix(r ^ key_s) & mask)) with keys from
 * splitmix64 seeded by seed ^ 0x5DEECE66D, cycle-walked until < sites. */
void or_mock_sites(long long np, long long sites, unsigned long long seed, long long *out) {
    if (np == sites) {
        for (long long i = 0; i < np; i++) out[i] = i;
        return;
    }
    int bits = 0;
    while ((1ull << bits) < (unsigned long long)sites) bits++;
    const int h = (bits + 1) / 2;
    const uint64_t mask = (1ull << h) - 1ull;
    uint64_t st = (uint64_t)seed ^ 0x5DEECE66Dull, key[4];
    for (int s = 0; s < 4; s++) key[s] = or_splitmix(&st);
    for (long long i = 0; i < np; i++) {
        uint64_t y = (uint64_t)i;
        do {
            uint64_t l = y >> h, r = y & mask;
            for (int s = 0; s < 4; s++) {
                const uint64_t t = l ^ (or_mix(r ^ key[s]) & mask);
                l = r;
                r = t;
            }
            y = (l << h) | r;
        } while (y >= (uint64_t)sites);
        out[i] = (long long)y;
    }
}

/* Uniforms of the biased-tracer selection: counter (s lo, s hi, 1, 0), key
 * (seed lo, seed hi), (w0 >> 5, w1 >> 6) as a 53-bit fraction. */
void or_mock_select_uniform(long long n, unsigned long long seed, double *out) {
    const unsigned int key[2] = {(unsigned int)(seed & 0xFFFFFFFFu), (unsigned int)(seed >> 32)};
    for (long long i = 0; i < n; i++) {
        const unsigned long long u = (unsigned long long)i;
        const unsigned int ctr[4] = {(unsigned int)(u & 0xFFFFFFFFu), (unsigned int)(u >> 32), 1u, 0u};
        unsigned int w[4];
        or_philox4x32_10(ctr, key, w);
        out[i] = (double)(((uint64_t)(w[0] >> 5) << 26) | (uint64_t)(w[1] >> 6)) / 9007199254740992.0;
    }
}

/* Sky-patch randoms: the restatement of the library's generator (counters
 * (i lo, i hi, 3, 0) and (i lo, i hi, 4, 0), key (seed lo, seed hi); r^3, RA
 * and sin(Dec) uniform; x = r cos(Dec) cos(RA), y = r cos(Dec) sin(RA),
 * z = r sin(Dec) as fits2dat_galaxy_weight.py:130-132). */
void or_sky_randoms(const double patch[6], unsigned long long seed, long long n, double *out) {
    const unsigned int key[2] = {(unsigned int)(seed & 0xFFFFFFFFu), (unsigned int)(seed >> 32)};
    const double deg = M_PI / 180.0;
    const double r3a = patch[0] * patch[0] * patch[0], r3b = patch[1] * patch[1] * patch[1];
    const double sa = sin(patch[4] * deg), sb = sin(patch[5] * deg);
    for (long long i = 0; i < n; i++) {
        const unsigned long long u = (unsigned long long)i;
        const unsigned int c1[4] = {(unsigned int)(u & 0xFFFFFFFFu), (unsigned int)(u >> 32), 3u, 0u};
        const unsigned int c2[4] = {(unsigned int)(u & 0xFFFFFFFFu), (unsigned int)(u >> 32), 4u, 0u};
        unsigned int w[4], v[4];
        or_philox4x32_10(c1, key, w);
        or_philox4x32_10(c2, key, v);
        const double u1 = (double)(((uint64_t)(w[0] >> 5) << 26) | (uint64_t)(w[1] >> 6)) / 9007199254740992.0;
        const double u2 = (double)(((uint64_t)(w[2] >> 5) << 26) | (uint64_t)(w[3] >> 6)) / 9007199254740992.0;
        const double u3 = (double)(((uint64_t)(v[0] >> 5) << 26) | (uint64_t)(v[1] >> 6)) / 9007199254740992.0;
        const double r = cbrt(r3a + u1 * (r3b - r3a));
        const double ra = (patch[2] + u2 * (patch[3] - patch[2])) * deg;
        const double sd = sa + u3 * (sb - sa);
        const double cd = sqrt(1.0 - sd * sd);
        out[3 * i] = r * cd * cos(ra);
        out[3 * i + 1] = r * cd * sin(ra);
        out[3 * i + 2] = r * sd;
    }
}

/* ------------------------------------------------------------------ */
/* single seams (the pieces the drivers above compose), exported for    */
/* the per-seam entry points of the library                             */
/* ------------------------------------------------------------------ */

/* ToolClass::storeReducedSphericalHarmonicsIn{Fourier,Configuration}Space
 * (tools.hpp:70-115): config = 0 k-space, 1 r-space; ylm: 2*ng doubles. */
void or_store_ylm(const or_param *p, int ell, int m, int config, double *ylm) {
    if (config) store_ylm_config(p, ell, m, ylm);
    else store_ylm_fourier(p, ell, m, ylm);
}

/* field.hpp:403-445 with the spline of j_ell_j (bessel.hpp:4-41) */
int or_ifft_for_3pcf(const or_param *p, const double *density, double rmag, const double *ylm, int ell_j,
                     double *out) {
    or_bessel *sj = or_bessel_new(ell_j);
    if (!sj) return -1;
    ifft_for_3pcf(p, density, rmag, ylm, sj, out);
    or_bessel_free(sj);
    return 0;
}

/* field.hpp:979-1113 (p->ell1 + p->ell2 sets the sign, as the reference's param does) */
int or_corr_for_3pcf(const or_param *p, const double *f1, const double *f2, const double *rbin, double sn_re,
                     double sn_im, const double *ylm1, const double *ylm2, double *xi, int *n_mode) {
    corr_for_3pcf(p, f1, f2, rbin, sn_re, sn_im, ylm1, ylm2, xi, n_mode);
    return 0;
}

/* bispec.hpp:1036-1056 for one (kmag1, kmag2): sum_x j_l1(k1 r) j_l2(k2 r)
 * xi Ylm1 Ylm2, times i^(l1 + l2) dV; ylm1 / ylm2 r-space tables. */
int or_sn_bessel_sweep(const or_param *p, const double *xi, const double *ylm1, const double *ylm2, int ell1,
                       int ell2, double kmag1, double kmag2, double *out) {
    or_bessel *sj1 = or_bessel_new(ell1), *sj2 = or_bessel_new(ell2);
    if (!sj1 || !sj2) {
        or_bessel_free(sj1);
        or_bessel_free(sj2);
        return -1;
    }
    double sr = 0.0, si = 0.0, rvec[3], dr[3];
    dr[0] = p->boxsize[0] / (double)(p->n_mesh[0]);
    dr[1] = p->boxsize[1] / (double)(p->n_mesh[1]);
    dr[2] = p->boxsize[2] / (double)(p->n_mesh[2]);
    for (int i = 0; i < p->n_mesh[0]; i++)
        for (int j = 0; j < p->n_mesh[1]; j++)
            for (int k = 0; k < p->n_mesh[2]; k++) {
                long long coord = or_coord(p, i, j, k);
                kvec_of(p, i, j, k, dr, rvec);
                double rmag = sqrt(rvec[0] * rvec[0] + rvec[1] * rvec[1] + rvec[2] * rvec[2]);
                double j1 = or_bessel_eval(sj1, kmag1 * rmag);
                double j2 = or_bessel_eval(sj2, kmag2 * rmag);
                double ar = j1 * j2 * xi[2 * coord], ai = j1 * j2 * xi[2 * coord + 1], br, bi, cr, ci;
                CMUL(ar, ai, ylm1[2 * coord], ylm1[2 * coord + 1], br, bi);
                CMUL(br, bi, ylm2[2 * coord], ylm2[2 * coord + 1], cr, ci);
                sr += cr;
                si += ci;
            }
    double ir, ii;
    ipow(ell1 + ell2, &ir, &ii);
    double fac = p->volume / (double)(p->n_mesh_tot);
    double tr = sr * ir - si * ii, ti = sr * ii + si * ir;
    out[0] = tr * fac;
    out[1] = ti * fac;
    or_bessel_free(sj1);
    or_bessel_free(sj2);
    return 0;
}

/* bispec.hpp:1159-1168: sum_x fac f1 f2 f3, fac = V / N_grid */
void or_triple_sum(const or_param *p, const double *f1, const double *f2, const double *f3, double *out) {
    double fac = p->volume / (double)(p->n_mesh_tot), sr = 0.0, si = 0.0;
    for (long long c = 0; c < p->n_mesh_tot; c++) {
        double ar = fac * f1[2 * c], ai = fac * f1[2 * c + 1], br, bi, cr, ci;
        CMUL(ar, ai, f2[2 * c], f2[2 * c + 1], br, bi);
        CMUL(br, bi, f3[2 * c], f3[2 * c + 1], cr, ci);
        sr += cr;
        si += ci;
    }
    out[0] = sr;
    out[1] = si;
}

/* field.hpp:1164-1240: sum_p pow(w_p, power) Y_red(ell, m)(los_p); w NULL: 1 */
void or_particle_sum(const double *los, const double *w, long long n, int ell, int m, int power, double *out) {
    double sr = 0.0, si = 0.0, y[2];
    for (long long q = 0; q < n; q++) {
        or_ylm(ell, m, los + 3 * q, y);
        double f = pow(w ? w[q] : 1.0, power);
        sr += f * y[0];
        si += f * y[1];
    }
    out[0] = sr;
    out[1] = si;
}

/* calcSurveyVolume (field.hpp:449-503): (sum w)^2 / sum_x rho_R(x)^2 dV */
double or_survey_volume(const or_param *p, const double *pos, const double *w, long long n) {
    double *weight = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    for (long long q = 0; q < n; q++) {
        weight[2 * q] = w ? w[q] : 1.0;
        weight[2 * q + 1] = 0.0;
    }
    double *field = grid_alloc(p);
    or_calc_field(p, pos, weight, n, field);
    free(weight);
    double dV = p->volume / (double)(p->n_mesh_tot), norm = 0.0, nw = 0.0;
    for (long long i = 0; i < p->n_mesh_tot; i++) norm += field[2 * i] * field[2 * i] * dV;
    for (long long q = 0; q < n; q++) nw += w ? w[q] : 1.0;
    free(field);
    return nw * nw / norm;
}
