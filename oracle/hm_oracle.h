/*
 * hm_oracle.h -- FP64 CPU restatement of hitomi_measurement's FFT estimator
 * path (periodic-box programs), used ONLY as test infrastructure: by tests/,
 * by __graft_entry__.smoke() as the checker and by bench.py's cpu_baseline
 * leg.  The product path (libhitomi_amd.so, the hitomi_box CLI) never links,
 * loads or calls anything in this directory.
 *
 * Every function follows a reference file:line (paths relative to the
 * reference repository, src/hitomi_measurement/), keeping its loop order,
 * FP64 expression order and its quirks.  FFTW3 and GSL are not available
 * here, so the three library pieces the reference calls are restated from
 * their published definitions:
 *   - fftw_plan_dft_3d / fftw_execute (fftw-3.3.10): unnormalised in-place
 *     complex DFT, sign -1 forward / +1 backward, row-major (or_fft3d);
 *   - gsl_sf_coupling_3j, gsl_sf_legendre_sphPlm, gsl_sf_bessel_jl and the
 *     natural cubic spline gsl_interp_cspline (gsl-2.7.1): Racah formula,
 *     three-term Legendre recurrence with the Condon-Shortley phase,
 *     series/upward recurrence for j_l, tridiagonal natural spline.
 * How each part is pinned (golden vectors, closed forms, brute force) is
 * listed in tests/test_oracle.py.
 */
#ifndef HM_ORACLE_H
#define HM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_NGP = 0, OR_CIC = 1, OR_TSC = 2 };

/* The subset of ParameterClass (parameter.hpp:8-55) the estimators read. */
typedef struct {
    int n_mesh[3];
    double boxsize[3];
    double volume;
    long long n_mesh_tot;
    int assign;
    int ell1, ell2, ELL;
    int n_kbin, n_rbin, ith_kbin, ith_rbin;
    double kmin, kmax, rmin, rmax;
} or_param;

void or_param_init(or_param *p, int n0, int n1, int n2, double L0, double L1, double L2, int assign);
void or_set_kbin(const or_param *p, double *kbin); /* tools.hpp:118-124 */
void or_set_rbin(const or_param *p, double *rbin); /* tools.hpp:126-132 */

/* field.hpp:47-184: rho = sum_p w_p W(x - x_p) / dV on a complex grid (2*ng doubles). */
int or_calc_field(const or_param *p, const double *pos, const double *weight, long long np, double *field);
/* Per-particle stencil cell indices as the reference computes them
 * (field.hpp:84-93, 121-133, 162-176): writes 1/8/27 coords per particle,
 * -1 where the reference's range check drops the contribution. */
int or_cell_coords(const or_param *p, const double *pos, long long np, long long *coords);
/* P_l bin mode counts (field.hpp:793-837) and bispectrum shell n_mode
 * (field.hpp:361-387) for the bins of tools.hpp:118-124, counts only. */
int or_mode_counts(const or_param *p, const double *kbin, long long *pk_nmode, long long *shell_nmode);

/* Unnormalised in-place 3D complex DFT, sign -1 (FFTW_FORWARD) or +1. */
int or_fft3d(const or_param *p, double *field, int sign);
/* 1D complex DFT of length n, stride 1, in place. */
int or_fft1d(double *x, int n, int sign);
void or_fourier_transform(const or_param *p, double *field);          /* field.hpp:310-326 */
void or_inverse_fourier_transform(const or_param *p, double *field);  /* field.hpp:328-344 */
void or_assignment_correction(const or_param *p, double *field);      /* field.hpp:610-639 */
double or_assignment_function(const or_param *p, const double *kvec); /* field.hpp:641-717 */
double or_shotnoise_function(const or_param *p, const double *kvec);  /* field.hpp:1252-1310 */

/* tools.hpp:12-68 and the GSL pieces it calls. */
void or_ylm(int ell, int m, const double *v, double *out2);
double or_sphPlm(int l, int m, double x);
double or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);
double or_bessel_jl(int l, double x);

/* bessel.hpp:4-41: 10^6-node natural cubic spline of j_l on [0, 1e4]. */
typedef struct or_bessel or_bessel;
or_bessel *or_bessel_new(int ell);
double or_bessel_eval(const or_bessel *b, double x); /* NaN outside [0,1e4] */
void or_bessel_free(or_bessel *b);

/* field.hpp:765-852 */
int or_calc_power_spectrum(const or_param *p, const double *f1, const double *f2, const double *kbin,
                           double sn_re, double sn_im, int ell, int m, double *pk, int *n_mode);
/* field.hpp:854-977 */
int or_calc_correlation_function(const or_param *p, const double *f1, const double *f2, const double *rbin,
                                 double sn_re, double sn_im, int ell, int m, double *xi, int *n_mode);
/* field.hpp:346-401: returns n_mode */
int or_ifft_for_bispectrum(const or_param *p, const double *density, double kmag_in, double dk_in,
                           const double *ylm, double *out);
/* field.hpp:1116-1160 */
int or_shotnoise_bispectrum_ijk(const or_param *p, const double *f1, const double *f2, double sn_re, double sn_im,
                                double *xi);

/* field.hpp:505-523 */
int or_box_density_fluctuation(const or_param *p, const double *pos, long long np, double *field);
/* field.hpp:560-572 */
int or_box_density(const or_param *p, const double *pos, long long np, double *field);

/* powerspec.hpp:273-332: pk (2*n_kbin, already x norm x (2l+1)) as written to pk%d.dat. */
int or_box_pk(const or_param *p, const double *pos, long long np, const double *kbin, double *pk_out, int *n_mode);
/* powerspec.hpp:334-393: xi (n_rbin, real part, normalised) as written to xi%d.dat. */
int or_box_xi(const or_param *p, const double *pos, long long np, const double *rbin, double *xi_out);
/* bispec.hpp:877-1203 for p->ith_kbin: bk_save, sn_save (2*n_kbin each, raw sums) and
 * out (2*n_kbin) = norm * (bk - sn) as written to bk%d%d%d_%02d.dat. */
int or_box_bk(const or_param *p, const double *pos, long long np, const double *kbin,
              double *bk_save, double *sn_save, double *out);

/* Survey (main.cpp:109-166): weight sums, LOS of the unshifted positions, alpha,
 * the FFT-box shift (positions are shifted in place), Vsurvey and nbar.
 * info = {sum w_D, sum w_R, alpha, Vsurvey, nbar, dP[3]}. */
int or_survey_prepare(const or_param *p, double *posD, const double *wD, long long nD, double *posR,
                      const double *wR, long long nR, double *losD, double *losR, double *info);
/* field.hpp:186-228 (complex field, 2*n_mesh_tot). */
int or_ylm_weighted_fluctuation(const or_param *p, const double *posD, const double *losD, const double *wD,
                                long long nD, const double *posR, const double *losR, const double *wR, long long nR,
                                double alpha, int ell, int M, double *field);
/* field.hpp:1164-1188 (out2 complex). */
void or_shotnoise_pk(const double *losD, const double *wD, long long nD, const double *losR, const double *wR,
                     long long nR, double alpha, int ell, int M, double *out2);
/* powerspec.hpp:4-91: norm * pk_save (2*n_kbin) as written to pk%d.dat; positions already shifted. */
int or_survey_pk(const or_param *p, const double *posD, const double *losD, const double *wD, long long nD,
                 const double *posR, const double *losR, const double *wR, long long nR, double alpha, double vs,
                 const double *kbin, double *out, int *n_mode);

/* powerspec.hpp:93-179 (window 0) / 182-270 (window 1): norm * xi_save (2*n_rbin). */
int or_survey_xi(const or_param *p, const double *posD, const double *losD, const double *wD, long long nD,
                 const double *posR, const double *losR, const double *wR, long long nR, double alpha, double vs,
                 const double *rbin, int window, double *out, int *n_mode);

/* Reconstruction (particle_reconstruction.hpp).  psi: 3 complex grids. */
void or_recon_psi(const or_param *p, const double *dk, double b1, double RG, double *psi);
void or_recon_gather(const or_param *p, const double *psi, double norm, double *pos, long long np);
int or_recon_displace(const or_param *p, double *posD, const double *wD, long long nD, double *posR,
                      const double *wR, long long nR, double alpha, double norm, double b1, double RG);
/* powerspec.hpp:396-519: box P_l / xi_l after reconstruction (out complex, times norm). */
int or_box_recon_2pt(const or_param *p, const double *posD, long long nD, const double *posR, long long nR,
                     double alpha, const double *bins, int xi, double *out, int *n_mode);

/* bispec.hpp:1205-1424 for p->ith_rbin: zeta_save, sn_save, out = norm * (zeta - sn) (2*n_rbin each). */
int or_box_zeta(const or_param *p, const double *pos, long long np, const double *rbin, double *zeta_save,
                double *sn_save, double *out);
/* bispec.hpp:1426-1753 and 1755-1975: the reconstructed box (D - alpha R,
 * shot-noise density D + alpha^2 R, N_D - alpha^3 N_R, norm from N_D). */
int or_box_recon_bk(const or_param *p, const double *posD, long long nD, const double *posR, long long nR,
                    double alpha, const double *kbin, double *bk_save, double *sn_save, double *out);
int or_box_recon_zeta(const or_param *p, const double *posD, long long nD, const double *posR, long long nR,
                      double alpha, const double *rbin, double *zeta_save, double *sn_save, double *out);

/* Row-major index helper used by the tests. */
long long or_coord(const or_param *p, int i, int j, int k);

#ifdef __cplusplus
}
#endif
/* bispec.hpp:4-353 (survey B) and 355-874 (survey and window zeta) on the
 * prepared catalogues (shifted positions, lines of sight, weights). */
int or_survey_bk(const or_param *p, const double *posD, const double *losD, const double *wD, long long nD,
                 const double *posR, const double *losR, const double *wR, long long nR, double alpha, double vs,
                 const double *kbin, double *bk_save, double *sn_save, double *out);
int or_survey_zeta(const or_param *p, const double *posD, const double *losD, const double *wD, long long nD,
                   const double *posR, const double *losR, const double *wR, long long nR, double alpha, double vs,
                   const double *rbin, int window, double *zeta_save, double *sn_save, double *out);

/* Single seams exported for the library's per-seam entry points. */
void or_store_ylm(const or_param *p, int ell, int m, int config, double *ylm); /* tools.hpp:70-115 */
int or_ifft_for_3pcf(const or_param *p, const double *density, double rmag, const double *ylm, int ell_j,
                     double *out);                                               /* field.hpp:403-445 */
int or_corr_for_3pcf(const or_param *p, const double *f1, const double *f2, const double *rbin, double sn_re,
                     double sn_im, const double *ylm1, const double *ylm2, double *xi, int *n_mode); /* field.hpp:979-1113 */
int or_sn_bessel_sweep(const or_param *p, const double *xi, const double *ylm1, const double *ylm2, int ell1,
                       int ell2, double kmag1, double kmag2, double *out);     /* bispec.hpp:1036-1056 */
void or_triple_sum(const or_param *p, const double *f1, const double *f2, const double *f3, double *out); /* bispec.hpp:1159-1168 */
void or_particle_sum(const double *los, const double *w, long long n, int ell, int m, int power, double *out); /* field.hpp:1164-1240 */
double or_survey_volume(const or_param *p, const double *pos, const double *w, long long n); /* field.hpp:449-503 */

/* Interlacing (no counterpart in the reference): the half-cell-shifted
 * second grid combined in Fourier space, and P_l / xi_l from it with the
 * shot noise N W(k)^2. */
int or_box_interlaced_k(const or_param *p, const double *pos, long long np, double *out);
int or_box_interlaced_2pt(const or_param *p, const double *pos, long long np, const double *bins, int xi,
                          double *out, int *n_mode);

/* Mock generator (no counterpart in the reference): Philox4x32-10 of Salmon
 * et al. (SC 2011), the per-cell standard normal (FP64 Box-Muller of two
 * 53-bit uniforms) and the seeded choice of lattice sites (a four-round
 * Feistel network with the splitmix64 finaliser as round function,
 * cycle-walked into [0, sites)).  The FFT part is numpy in pyoracle.py. */
void or_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]);
void or_mock_white_noise(long long n, unsigned long long seed, double *out);
void or_mock_sites(long long np, long long sites, unsigned long long seed, long long *out);
void or_mock_select_uniform(long long n, unsigned long long seed, double *out);
/* sky-patch randoms: patch = {rmin, rmax, ra_min, ra_max, dec_min, dec_max} (degrees) */
void or_sky_randoms(const double patch[6], unsigned long long seed, long long n, double *out);

#endif
