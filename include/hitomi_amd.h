/*
 * hitomi_amd.h -- C ABI of libhitomi_amd.so, the MI355X (gfx950) kernels of
 * hitomi_measurement's FFT estimator path.
 *
 * The reference has no FFI: its estimators are header-only C++ templates
 * called from main_box.cpp / main.cpp (reference src/hitomi_measurement/).
 * Each entry point below replaces one of those internal seams; the citation
 * names the reference function (file:line) whose arithmetic it performs.
 *
 * Conventions
 *  - Every call returns 0 on success and -1 on failure; hm_last_error()
 *    then describes the failure.  No C++ exception crosses this boundary.
 *  - Pointers named d_* are device pointers (from hm_dev_alloc); all other
 *    pointers are host memory owned by the caller and only borrowed for the
 *    duration of the call.
 *  - Calls are ordered on the context's stream.  Calls that return host
 *    results (hm_box_plan_execute, hm_memcpy_d2h) synchronise that stream.
 *  - One host thread per context.  Grids are row-major [n0][n1][n2] with
 *    the z index fastest, as DensityFieldClass (field.hpp:364) lays them.
 *  - FFT sign and scale follow FFTW as the reference calls it
 *    (field.hpp:321, field.hpp:339): sign -1 forward, +1 backward,
 *    unnormalised.
 */
#ifndef HITOMI_AMD_H
#define HITOMI_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hm_ctx hm_ctx;
typedef struct hm_box_plan hm_box_plan;

enum { HM_NGP = 0, HM_CIC = 1, HM_TSC = 2 };  /* parameter.hpp:533 "assignment" */
enum { HM_FP32 = 0, HM_FP64 = 1 };            /* precision of the grids and FFTs */

/* The mesh part of ParameterClass (parameter.hpp:13-16, 533). */
typedef struct {
    int n_mesh[3];
    double boxsize[3];
    int assign; /* HM_NGP, HM_CIC or HM_TSC */
} hm_mesh;

/* ---------------------------------------------------------------- */
/* context and device memory                                          */
/* ---------------------------------------------------------------- */
hm_ctx *hm_ctx_create(int device, int flags);
void hm_ctx_destroy(hm_ctx *ctx);
const char *hm_last_error(const hm_ctx *ctx);
int hm_sync(hm_ctx *ctx);
void *hm_dev_alloc(hm_ctx *ctx, size_t bytes);
int hm_dev_free(hm_ctx *ctx, void *d_ptr);
int hm_memcpy_h2d(hm_ctx *ctx, void *d_dst, const void *src, size_t bytes);
int hm_memcpy_d2h(hm_ctx *ctx, void *dst, const void *d_src, size_t bytes);
/* Device-to-device copy with the library's own copy kernel, ordered on the
 * context's stream (no counterpart in the reference: plumbing, and the
 * known-size kernel that scales HBM traffic counters). */
int hm_memcpy_d2d(hm_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
int hm_memset_zero(hm_ctx *ctx, void *d_ptr, size_t bytes);

/* Kernel timing with HIP events recorded on the launch stream around every
 * launch (off by default).  hm_timing_get() synchronises and reports the
 * accumulated time and launch count of the kernels launched under `name`. */
int hm_timing_enable(hm_ctx *ctx, int on);
int hm_timing_get(hm_ctx *ctx, const char *name, double *total_ms, long long *launches);
int hm_timing_reset(hm_ctx *ctx);
/* Busy time of `name` since the last hm_timing_reset(): the length of the
 * union of its launches' [start, end) intervals, so that launches overlapping
 * on several streams (batch lanes, merged by hm_batch_collect_timing) count
 * once.  No counterpart in the reference (a measurement aid). */
int hm_timing_get_busy(hm_ctx *ctx, const char *name, double *busy_ms);
/* Restrict the timed launches to one kernel name (NULL or "" times all). */
int hm_timing_select(hm_ctx *ctx, const char *name);

/* ---------------------------------------------------------------- */
/* mass assignment                                                    */
/* ---------------------------------------------------------------- */

/* DensityFieldClass::calcField (field.hpp:47-184): rho(x) = sum_p w_p W(x-x_p)/dV
 * with NGP/CIC/TSC, the reference's unwrapped stencil and linear range check.
 * d_pos: np x 3 FP64 positions; d_w: np complex FP64 weights (re, im) or NULL
 * for unit weights; wabs_sum >= sum_p (|re w_p| + |im w_p|) (<= 0: np).
 * Output: FP64 grids (d_out_im may be NULL).  Cell sums are accumulated in
 * 64-bit fixed point, so the result is bitwise independent of particle order. */
int hm_assign(hm_ctx *ctx, const hm_mesh *mesh, const double *d_pos, const double *d_w, long long np,
              double wabs_sum, double *d_out_re, double *d_out_im);
/* The same grid from the scatter form of field.hpp:172-180 (one 64-bit
 * fixed-point atomic per stencil cell); bitwise equal to hm_assign, which
 * sorts particles by cell and gathers.  Kept as the cross-check. */
int hm_assign_atomic(hm_ctx *ctx, const hm_mesh *mesh, const double *d_pos, const double *d_w, long long np,
                     double wabs_sum, double *d_out_re, double *d_out_im);

/* hm_assign without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same grids as
 * hm_assign, bit for bit.  The sort work lives in a fixed 128 MiB device array
 * of the library (one per GPU, held until the process ends): about 48 B per
 * particle with weights (32 B without) plus 8 B per (x, y, 64-cell z chunk);
 * larger inputs are rejected with an error.  The int64 cell sums are
 * accumulated in d_out_re / d_out_im themselves and converted in place.
 * Asynchronous on the context's stream (no synchronisation).  Every call on a
 * GPU shares that GPU's scratch: calls that overlap on one GPU (other streams,
 * other contexts, other threads) race and must be ordered by the caller. */
int hm_assign_noalloc(hm_ctx *ctx, const hm_mesh *mesh, const double *d_pos, const double *d_w, long long np,
                      double wabs_sum, double *d_out_re, double *d_out_im);

/* The 1/8/27 stencil cell indices of each particle as field.hpp:84-93,
 * 121-133, 162-176 compute them; -1 where the reference's range check
 * drops the contribution.  d_coords: np * (1|8|27) int64. */
int hm_cell_coords(hm_ctx *ctx, const hm_mesh *mesh, const double *d_pos, long long np, long long *d_coords);

/* ---------------------------------------------------------------- */
/* FFT (replaces fftw_plan_dft_3d/fftw_execute, field.hpp:310-344)    */
/* ---------------------------------------------------------------- */
/* In-place complex 3D DFT of d_data ([n0][n1][n2] complex FP32/FP64). */
int hm_fft_c2c(hm_ctx *ctx, int precision, void *d_data, const int n[3], int sign);
/* Real [n0][n1][n2] -> half spectrum [n0][n1][n2/2+1], forward (sign -1). */
int hm_fft_r2c(hm_ctx *ctx, int precision, const void *d_in, void *d_out, const int n[3]);
/* Half spectrum -> real, backward (sign +1); d_in is overwritten. */
int hm_fft_c2r(hm_ctx *ctx, int precision, void *d_in, void *d_out, const int n[3]);

/* ---------------------------------------------------------------- */
/* single seams on the reference's own grid layout                    */
/* ---------------------------------------------------------------- */
/* A "grid" below is DensityFieldClass's field (field.hpp:20-31): complex
 * FP64 [n0][n1][n2] with (re, im) interleaved like fftw_complex, over the
 * whole mesh, in device memory.  Each entry point performs one member
 * function of the reference on such grids, so a caller can replace one seam
 * at a time; the plans further down fuse whole estimators instead.  Sums
 * are deterministic (fixed-order reductions).  Calls that return host
 * values synchronise the context's stream; the others are ordered on it
 * and also synchronise before returning. */

/* DensityFieldClass constructor (field.hpp:20-31): a zeroed grid; NULL on failure. */
double *hm_grid_alloc(hm_ctx *ctx, const hm_mesh *mesh);
/* finalizeDensityField (field.hpp:35-45). */
int hm_grid_free(hm_ctx *ctx, double *d_grid);
/* calcField (field.hpp:47-184) into a grid: rho = sum_p w_p W(x - x_p) / dV
 * with complex weights d_w (np x (re, im); NULL: 1), as hm_assign computes it
 * (64-bit fixed point, order-independent); wabs_sum as for hm_assign. */
int hm_assign_field(hm_ctx *ctx, const hm_mesh *mesh, const double *d_pos, const double *d_w, long long np,
                    double wabs_sum, double *d_grid);
/* calcFourierTransform (field.hpp:310-326, sign -1: x dV, then FFTW_FORWARD)
 * and calcInverseFourierTransform (field.hpp:328-344, sign +1: / V, then
 * FFTW_BACKWARD), in place. */
int hm_fourier_transform(hm_ctx *ctx, const hm_mesh *mesh, double *d_grid, int sign);
/* calcAssignmentFunctionCorrection (field.hpp:610-639): grid / W(k), with
 * the reference's truncated index for negative k. */
int hm_window_correct(hm_ctx *ctx, const hm_mesh *mesh, double *d_grid);
/* TwoPointStatisticsClass::calcPowerSpectrum (field.hpp:765-852):
 * (f1 conj(f2) - sn S(k)) / W(k)^2 Y_red(ell, m)(k) on the fine 1e-4 grid,
 * folded into the bins around kbin[0..n_kbin-1] (n_kbin >= 2) and averaged.
 * sn: complex shot noise (NULL: 0).  pk: [n_kbin][2]; n_mode: [n_kbin] (may be NULL). */
int hm_bin_pk(hm_ctx *ctx, const hm_mesh *mesh, const double *d_f1, const double *d_f2, const double *kbin,
              int n_kbin, const double sn[2], int ell, int m, double *pk, long long *n_mode);
/* calcCorrelationFunction (field.hpp:854-977): the same product / V,
 * FFTW_BACKWARD, then Y_red(ell, m)(r) on the 0.5 Mpc/h fine grid folded into
 * the bins around rbin (n_rbin >= 2) and averaged.  xi: [n_rbin][2]. */
int hm_bin_xi(hm_ctx *ctx, const hm_mesh *mesh, const double *d_f1, const double *d_f2, const double *rbin,
              int n_rbin, const double sn[2], int ell, int m, double *xi, long long *n_mode);
/* calcCorrelationFunctionForThreePointFunction (field.hpp:979-1113): as
 * hm_bin_xi with the weight Y_red(ell1, m1)(r) Y_red(ell2, m2)(r) (the
 * reference's Ylm1, Ylm2 tables), the 3PCF bin edges and the scaling
 * (-1)^(ell1 + ell2) / n_mode^2 / dV. */
int hm_bin_xi_3pcf(hm_ctx *ctx, const hm_mesh *mesh, const double *d_f1, const double *d_f2, const double *rbin,
                   int n_rbin, const double sn[2], int ell1, int m1, int ell2, int m2, double *xi,
                   long long *n_mode);
/* calcShotNoiseForBispectrum_ijk (field.hpp:1116-1160): d_xi = FFTW_BACKWARD
 * of (f1 conj(f2) - sn S(k)) / W(k)^2 / V. */
int hm_sn_field_ijk(hm_ctx *ctx, const hm_mesh *mesh, const double *d_f1, const double *d_f2, const double sn[2],
                    double *d_xi);
/* calcInverseFourierTransformForBispectrum (field.hpp:346-401): d_out =
 * FFTW_BACKWARD of Y_red(ell, m)(k) d(k) / W(k) on the shell
 * max(kmag - dk/2, 0) < |k| <= kmag + dk/2, divided by the shell's mode
 * count, which is returned in n_mode (may be NULL).  d_out != d_delta. */
int hm_shell_c2r(hm_ctx *ctx, const hm_mesh *mesh, const double *d_delta, double kmag, double dk, int ell, int m,
                 double *d_out, long long *n_mode);
/* calcInverseFourierTransformForThreePointFunction (field.hpp:403-445):
 * d_out = FFTW_BACKWARD of j_ell_j(|k| rmag) Y_red(ell, m)(k) (d(k) / W(k)) / V,
 * j from the 10^6-node spline of SphericalBesselClass (bessel.hpp:4-41). */
int hm_bessel_c2r(hm_ctx *ctx, const hm_mesh *mesh, const double *d_delta, double rmag, int ell_j, int ell, int m,
                  double *d_out);
/* The bispectrum sums of bispec.hpp:1159-1168 for many pairs at once:
 * out[p] = sum_x dV F_a F_b F3 with (a, b) = pairs[2p], pairs[2p + 1]
 * indexing d_fields (host array of n_fields grid pointers).  out: [n_pairs][2]. */
int hm_triple_sums(hm_ctx *ctx, const hm_mesh *mesh, const double *const *d_fields, int n_fields, const int *pairs,
                   int n_pairs, const double *d_f3, double *out);
/* The shot-noise sweep of bispec.hpp:1036-1056 for many (k1, k2) pairs:
 * out[p] = i^(ell1 + ell2) dV sum_x j_ell1(k1[p] |r|) j_ell2(k2[p] |r|)
 * xi(x) Y_red(ell1, m1)(r) Y_red(ell2, m2)(r), xi from hm_sn_field_ijk.
 * out: [n_pairs][2]. */
int hm_sn_bessel_sweep(hm_ctx *ctx, const hm_mesh *mesh, const double *d_xi, int ell1, int m1, int ell2, int m2,
                       const double *k1, const double *k2, int n_pairs, double *out);
/* The particle sums of the shot noise (field.hpp:1164-1240):
 * out = sum_p w_p^power Y_red(ell, m)(los_p); d_los: n x 3 (may be NULL for
 * ell = 0), d_w: n weights (NULL: 1).  power 2 gives calcShotNoiseFor-
 * Powerspectrum's and the window's sums, power 3 calcShotNoiseForBispectrum's. */
int hm_particle_sums(hm_ctx *ctx, const double *d_los, const double *d_w, long long n, int ell, int m, int power,
                     double out[2]);
/* calcSurveyVolume (field.hpp:449-503): the weighted randoms' field,
 * I = sum_x rho(x)^2 dV, survey_volume = (sum w)^2 / I.  d_w: n weights or NULL. */
int hm_survey_volume(hm_ctx *ctx, const hm_mesh *mesh, const double *d_pos, const double *d_w, long long n,
                     double *survey_volume);

/* ---------------------------------------------------------------- */
/* periodic-box estimators (main_box.cpp:171-228: pk, 2PCF, bk, 3PCF) */
/* ---------------------------------------------------------------- */
#define HM_MAX_ELLS 4

typedef struct {
    hm_mesh mesh;
    int precision;             /* HM_FP32: FP32 grids/FFTs, FP64 bin sums; HM_FP64: FP64 throughout */
    int n_kbin;                /* k-bins of ToolClass::setKbin (tools.hpp:118-124) */
    double kmin, kmax;
    int n_pk_ells;             /* P_l for l in pk_ells (calcPowerSpectrumForBOX, powerspec.hpp:273-332) */
    int pk_ells[HM_MAX_ELLS];
    int bk;                    /* 1: B_{ell1 ell2 ELL} for every (ith_kbin, k2) pair (calcBiSpectrumForBOX, bispec.hpp:877-1203) */
    int regrid;                /* 1: evaluate the triple product on the smallest mesh that holds it exactly */
    int prune;                 /* 1: skip the zero lines of band-limited transforms */
    int n_rbin;                /* r-bins of ToolClass::setRbin (tools.hpp:126-132) */
    double rmin, rmax;
    int n_xi_ells;             /* xi_l for l in xi_ells (calcTwoPointCorrelationFunctionForBOX, powerspec.hpp:334-393) */
    int xi_ells[HM_MAX_ELLS];
    int bk_ell[3];             /* (ell1, ell2, ELL) of the bispectrum multipole; {0, 0, 0}: B_000 */
    int zeta;                  /* 1: zeta_{ell1 ell2 ELL}(r1, r2) for every (ith_rbin, r2) pair on the r-bins
                                  (calcThreePointCorrelationFunctionForBOX, bispec.hpp:1205-1424) */
    int zeta_ell[3];
    int interlace;             /* 1: combine the density with a second grid of the positions shifted by half a cell,
                                  d = (d_a + exp(i k.h) d_b) / 2, shot noise N W(k)^2 (no counterpart in the
                                  reference; P_l and xi_l only; 0 gives the reference's estimator) */
} hm_box_config;

/* Host result arrays (any may be NULL). */
typedef struct {
    double *pk;            /* [n_pk_ells][n_kbin][2] norm*(2l+1)*P as written to pk%d.dat */
    double *pk_raw;        /* [n_pk_ells][n_kbin][2] bin averages of calcPowerSpectrum (field.hpp:839-845) */
    long long *pk_nmode;   /* [n_kbin] modes per bin */
    double *bk;            /* [n_kbin][n_kbin][2] norm*(bk - sn), row i = file bk%d%d%d_%02d.dat with ith_kbin = i */
    double *bk_sum;        /* [n_kbin][n_kbin][2] bk_save (bispec.hpp:1168) */
    double *sn_sum;        /* [n_kbin][n_kbin][2] sn_save (bispec.hpp:941-1056) */
    long long *shell_nmode;/* [n_kbin] n_mode of calcInverseFourierTransformForBispectrum (field.hpp:383) */
    double *xi;            /* [n_xi_ells][n_rbin] norm*(2l+1)*Re xi as written to xi%d.dat */
    double *xi_raw;        /* [n_xi_ells][n_rbin] real part of the bin averages of calcCorrelationFunction (field.hpp:961-968) */
    long long *xi_nmode;   /* [n_rbin] mesh points per r-bin */
    double *zeta;          /* [n_rbin][n_rbin][2] norm*(zeta - sn); row i = file zeta%d%d%d_%02d.dat (real part)
                              with ith_rbin = i */
    double *zeta_sum;      /* [n_rbin][n_rbin][2] zeta_save (bispec.hpp:1384-1389) */
    double *zeta_sn;       /* [n_rbin][n_rbin][2] sn_save (bispec.hpp:1279-1285; diagonal only) */
} hm_box_result;

hm_box_plan *hm_box_plan_create(hm_ctx *ctx, const hm_box_config *cfg);
void hm_box_plan_destroy(hm_ctx *ctx, hm_box_plan *plan);
/* One mock: d_pos (np x 3 FP64, already periodic) -> the configured P_l,
 * xi_l, B_{ell1 ell2 ELL} and zeta_{ell1 ell2 ELL}. */
int hm_box_plan_execute(hm_ctx *ctx, hm_box_plan *plan, const double *d_pos, long long np, hm_box_result *res);
/* The reconstructed box (main_box.cpp:199-228, flag_recon = True): P_l and
 * xi_l of D - alpha R with unit weights (field.hpp:528-558), shot noise
 * N_D + alpha^2 N_R (field.hpp:1190-1197), norm V / N_D^2 (powerspec.hpp:396-519);
 * B and zeta of the same field with the shot-noise density D + alpha^2 R
 * (field.hpp:574-605) and shot noise N_D - alpha^3 N_R (field.hpp:1242-1250),
 * norm V^2 / N_D^3 (bispec.hpp:1426-1975).
 * d_pos_d / d_pos_r: displaced and wrapped positions (n x 3 FP64). */
int hm_box_plan_execute_recon(hm_ctx *ctx, hm_box_plan *plan, const double *d_pos_d, long long n_d,
                              const double *d_pos_r, long long n_r, double alpha, hm_box_result *res);
/* Device bytes the plan holds. */
size_t hm_box_plan_bytes(const hm_box_plan *plan);
/* Mesh used for the triple product (n_mesh when regrid is off). */
int hm_box_plan_triple_mesh(const hm_box_plan *plan, int m[3]);
/* Autotuner (no counterpart in the reference): times every combination of
 * the launch shapes it can vary -- the wide (2x) column tile of the FP32
 * column passes of length 256, 270 or 288, and 512 (bit 0, 1, 2 of wide_mask;
 * bit 3 the triple stage's shell x pass of length 270 or 288 on its own; used
 * when the device allows the LDS) and 64, 128 or 256 threads per block of the
 * P_l binning -- over reps executions of d_pos, and keeps the fastest for
 * this plan.  The results do not depend on these shapes.  report (may be
 * NULL) receives the timings.  Get / set read and impose them. */
int hm_box_plan_autotune(hm_ctx *ctx, hm_box_plan *plan, const double *d_pos, long long np, int reps, char *report,
                         size_t report_len);
int hm_box_plan_get_shapes(const hm_box_plan *plan, unsigned *wide_mask, int *bin_block);
int hm_box_plan_set_shapes(hm_box_plan *plan, unsigned wide_mask, int bin_block);

/* ---------------------------------------------------------------- */
/* special functions                                                  */
/* ---------------------------------------------------------------- */
/* gsl_sf_bessel_jl (bessel.hpp:20), the node values of the j_l spline the
 * bispectrum shot noise and the three-point function sample; host code,
 * no GPU needed.  NaN for l < 0 or x < 0. */
double hm_sph_bessel_jl(int l, double x);
/* gsl_sf_coupling_3j with integer arguments (common.hpp:29; the weights of
 * every P, B and zeta multipole), by the Racah formula in long double; host
 * code.  0 outside the triangle and selection rules. */
double hm_wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3);
/* The reduced spherical harmonic of tools.hpp:12-68, conj(Y_lm(v/|v|)) *
 * sqrt(4 pi / (2l + 1)), with its (0, 0) and |v| < 1e-15 cases; the same
 * function the kernels evaluate per mode and per separation, here on the
 * host.  Writes the real and imaginary parts to out[0], out[1]. */
void hm_ylm_reduced(int l, int m, const double v[3], double out[2]);

/* ---------------------------------------------------------------- */
/* reconstruction (particle_reconstruction.hpp)                       */
/* ---------------------------------------------------------------- */
typedef struct {
    hm_mesh mesh;      /* n_mesh_recon and boxsize; assign is the density's scheme (the gather is TSC) */
    int precision;     /* HM_FP32 or HM_FP64 grids and FFTs */
    double b1;         /* b1_fid */
    double RG;         /* Gaussian smoothing radius RG [Mpc/h] */
} hm_recon_config;

/* calcReconstructionParticles (particle_reconstruction.hpp:12-208) and
 * calcReconstructionParticlesForBOX (210-402): the field of D - alpha R with
 * weights d_w_* (NULL: 1), transformed and divided by the assignment window,
 * Psi(k) = -i (G/b1) k/k^2 delta/W / V (0 at k = 0, where the reference leaves
 * the grid unwritten), transformed back, scaled by norm and TSC-gathered
 * (real part) onto every particle of both catalogues, in place, without a
 * wrap.  The box caller passes alpha = double(N_D / N_R) with the reference's
 * integer division and norm = V / N_D; the survey caller alpha = sum w_D /
 * sum w_R and norm = Vsurvey / sum w_D. */
int hm_recon_displace(hm_ctx *ctx, const hm_recon_config *cfg, double *d_pos_d, const double *d_w_d, long long n_d,
                      double *d_pos_r, const double *d_w_r, long long n_r, double alpha, double norm);
/* calcPeriodicBoundary (particleGadget.hpp:565-591): one step into [0, L) per axis. */
int hm_periodic_wrap(hm_ctx *ctx, double *d_pos, long long n, const double boxsize[3]);

/* ---------------------------------------------------------------- */
/* survey estimators (main.cpp:109-274: pk, 2PCF, window2PCF, bk,   */
/* 3PCF, window3PCF)                                                  */
/* ---------------------------------------------------------------- */
typedef struct hm_survey_plan hm_survey_plan;

enum {
    HM_SURVEY_PK = 0,        /* measure = pk: calcPowerSpectrum (powerspec.hpp:4-91) */
    HM_SURVEY_XI = 1,        /* measure = 2PCF: calcTwoPointCorrelationFunction (powerspec.hpp:93-179) */
    HM_SURVEY_XI_WINDOW = 2, /* measure = window2PCF: calcTwoPointWindowCorrelationFunction (powerspec.hpp:182-270) */
    HM_SURVEY_BK = 3,        /* measure = bk: calcBiSpectrum (bispec.hpp:4-353), any (ell1, ell2, ELL) */
    HM_SURVEY_ZETA = 4,      /* measure = 3PCF: calcThreePointCorrelationFunction (bispec.hpp:355-614), any (ell1, ell2, ELL) */
    HM_SURVEY_ZETA_WINDOW = 5 /* measure = window3PCF: calcThreePointWindowCorrelationFunction (bispec.hpp:617-874), any (ell1, ell2, ELL) */
};

typedef struct {
    hm_mesh mesh;      /* FFT box: n_mesh and boxsize from the parameter file */
    int precision;     /* HM_FP32 or HM_FP64 grids and FFTs; FP64 sums */
    int n_kbin;        /* tools.hpp:118-124 */
    double kmin, kmax;
    int ell;           /* ELL; for pk and the 2PCFs ell1 = ELL, ell2 = 0 (calcPowerSpectrum, powerspec.hpp:11-15) */
    int prune;         /* 1: bin only the modes (separations) that can reach a bin */
    int measure;       /* HM_SURVEY_PK, HM_SURVEY_XI or HM_SURVEY_XI_WINDOW */
    int n_rbin;        /* tools.hpp:126-132, for the correlation functions */
    double rmin, rmax;
    int recon;         /* 1: reconstruct first (main.cpp:172-183, flag_recon = True) */
    int n_mesh_recon[3];
    double b1, RG;     /* b1_fid, RG */
    int ell1, ell2;    /* the bispectrum / three-point multipole (ell1, ell2, ELL) */
} hm_survey_config;

/* The survey quantities main.cpp derives before measuring. */
typedef struct {
    double wsum_d, wsum_r; /* n_tot_weight (particle.hpp:334-353) */
    double alpha;          /* sum w_D / sum w_R (particle.hpp:268-282) */
    double vsurvey;        /* (sum w_R)^2 / sum_x rho_R^2 dV (field.hpp:449-503) */
    double nbar;           /* (sum w_D)^2 / Vsurvey / sum w_D^2 (particle.hpp:284-299) */
    double shift[3];       /* min(R) - one cell, subtracted from every position (particle.hpp:222-265) */
    double norm;           /* Vsurvey / (sum w)^2 of D, or of R for the window (particle.hpp:302-313) */
} hm_survey_info;

hm_survey_plan *hm_survey_plan_create(hm_ctx *ctx, const hm_survey_config *cfg);
void hm_survey_plan_destroy(hm_ctx *ctx, hm_survey_plan *plan);
/* The configured measurement after main.cpp's preparation (main.cpp:109-166)
 * and, with recon = 1, the reconstruction (calcReconstructionParticles on the
 * n_mesh_recon mesh with alpha and norm = Vsurvey / sum w_D, then the FFT-box
 * shift again, main.cpp:172-183; the lines of sight and Vsurvey stay those of
 * the unreconstructed catalogues, as there):
 * d_pos_* are the unshifted Cartesian positions (n x 3 FP64; the line of
 * sight is pos/|pos| of these), d_w_* the weights (FP64).
 * out: [n_bins][2] = norm * pk_save (pk%d.dat) or norm * xi_save (whose real
 * part xi%d.dat and xi%d_window.dat hold); nmode: [n_bins] modes or mesh
 * points per bin.  For bk: out [n_kbin][n_kbin][2] = norm * (bk - sn), row i
 * the file bk%d%d%d_%02d.dat of ith_kbin = i, nmode the shell n_mode; for
 * 3PCF / window3PCF: out [n_rbin][n_rbin][2] = norm * (zeta - sn), whose real
 * part row i of zeta%d%d%d_%02d[_window].dat holds, nmode unused (may be NULL).
 * norm (info) = Vsurvey^2 / (sum w)^3 for these (particle.hpp:319-332). */
int hm_survey_plan_execute(hm_ctx *ctx, hm_survey_plan *plan, const double *d_pos_d, const double *d_w_d, long long n_d,
                           const double *d_pos_r, const double *d_w_r, long long n_r, hm_survey_info *info, double *out,
                           long long *nmode);
size_t hm_survey_plan_bytes(const hm_survey_plan *plan);

/* ---------------------------------------------------------------- */
/* Batched periodic-box pipeline and covariance.  No counterpart in the
 * reference, which runs one mock per process and forms the covariance
 * downstream with numpy.cov over the written files
 * (struct/covariance/calc_data_and_cov.py:284-307).  Each mock's data vector
 * is [P_l(k_b)] for the configured multipoles, then, with bk, the real part
 * of B_000(k_a, k_b) - SN for the bin pairs a <= b (row-major upper
 * triangle): the values hm_box_plan_execute returns in pk and bk, bit for
 * bit.  Mocks are pushed without host synchronisation; the vectors and the
 * FP64 sums of the covariance stay on the device. */
typedef struct hm_batch hm_batch;
typedef struct {
    hm_box_config box;     /* P_l and, with bk, B_000 (no xi, zeta, interlacing) */
    long long max_mocks;   /* data vectors kept on the device (further mocks still enter the covariance) */
    int cov_with_bk;       /* 0: covariance of the P_l entries; 1: of the whole vector */
    int use_graph;         /* 1: the mesh stage is captured once into a hipGraph and replayed */
    int lanes;             /* plans on streams of their own that take mocks in turn (0 or 1: one); the vectors and
                              the covariance sums are the same for any number */
} hm_batch_config;

hm_batch *hm_batch_create(hm_ctx *ctx, const hm_batch_config *cfg);
void hm_batch_destroy(hm_ctx *ctx, hm_batch *b);
/* vec_len: entries per data vector; cov_len: entries the covariance covers. */
int hm_batch_dims(const hm_batch *b, int *vec_len, int *cov_len);
/* One mock (d_pos: n x 3 FP64, already periodic; any n > 0).  Asynchronous:
 * d_pos must stay valid until the next hm_sync or result call.  With lanes,
 * mock i goes to lane i mod lanes; the lane waits for work queued on ctx's
 * stream before it reads d_pos, and ctx's stream waits until it has read it,
 * so a caller may refill d_pos with further work on ctx's stream. */
int hm_batch_push(hm_ctx *ctx, hm_batch *b, const double *d_pos, long long n);
/* Mocks pushed since creation or the last reset. */
long long hm_batch_count(const hm_batch *b);
/* Kernel timings recorded on the lanes' streams (hm_timing_enable on ctx)
 * added to ctx's, after the lanes finish. */
int hm_batch_collect_timing(hm_ctx *ctx, hm_batch *b);
/* The first min(count, max_mocks, max_rows) data vectors, [rows][vec_len]. */
int hm_batch_vectors(hm_ctx *ctx, hm_batch *b, double *out, long long max_rows);
/* mean [cov_len] and the unbiased (N - 1) sample covariance [cov_len][cov_len]
 * of numpy.cov (calc_data_and_cov.py:287), from device sums of the vectors
 * minus the first one.  Needs two or more mocks. */
int hm_batch_covariance(hm_ctx *ctx, hm_batch *b, double *mean, double *cov);
int hm_batch_reset(hm_ctx *ctx, hm_batch *b);
/* hm_box_plan_autotune on the batch's plan; a captured graph is dropped and
 * captured again at the next push. */
int hm_batch_autotune(hm_ctx *ctx, hm_batch *b, const double *d_pos, long long n, int reps, char *report,
                      size_t report_len);

/* ---------------------------------------------------------------- */
/* Covariance of window-normalised data vectors, accumulated on the device.
 * Replaces the numpy stage of
 * struct/covariance/boss/modified_gravity/calc_data_and_cov.py: the
 * normalisation xi_l / w_0 and zeta_l1l2L / w_000 (:72-105), and numpy.cov of
 * the stacked multipoles and of each pair of blocks (:284-307), which are the
 * blocks of one covariance of the concatenated vector.  Each push computes
 * x = v / norm (FP64, elementwise) and adds d = x - x_0 (x_0 the first
 * vector) into device sums of d and d d^T; pushes are ordered on ctx's
 * stream and do not synchronise with the host. */
typedef struct hm_cov hm_cov;
/* dim: vector length (1..4096); norm: dim window values on the host (copied;
 * none may be zero) or NULL for no division; keep_rows: normalised vectors
 * kept on the device for hm_cov_vectors. */
hm_cov *hm_cov_create(hm_ctx *ctx, int dim, const double *norm, long long keep_rows);
void hm_cov_destroy(hm_ctx *ctx, hm_cov *c);
/* One vector already in device memory (dim FP64), read in stream order. */
int hm_cov_push(hm_ctx *ctx, hm_cov *c, const double *d_vec);
/* One vector in host memory (copied before return). */
int hm_cov_push_host(hm_ctx *ctx, hm_cov *c, const double *vec);
long long hm_cov_count(const hm_cov *c);
/* The first min(count, keep_rows, max_rows) normalised vectors, [rows][dim]. */
int hm_cov_vectors(hm_ctx *ctx, hm_cov *c, double *out, long long max_rows);
/* mean [dim] and the N - 1 sample covariance [dim][dim] (numpy.cov); needs
 * two or more vectors. */
int hm_cov_result(hm_ctx *ctx, hm_cov *c, double *mean, double *cov);
int hm_cov_reset(hm_ctx *ctx, hm_cov *c);

/* ---------------------------------------------------------------- */
/* Seeded Zel'dovich mocks generated on the device.  No counterpart in the
 * reference (its only synthetic input is an unseeded uniform catalogue,
 * struct/measurement/nbody/pk/calc_random.py:38-49).  White noise from the
 * counter-based Philox4x32-10 generator (Salmon et al. 2011) with FP64
 * Box-Muller; the field delta(k) = FFT(noise) sqrt(P(k) / V_cell) with every
 * mode on a Nyquist plane set to zero; displacement Psi(k) = i k / k^2
 * delta(k), times 1 + f_growth along z; n particles on the lattice sites
 * chosen by a seeded permutation, moved by Psi and wrapped into [0, L). */
typedef struct {
    int n_side;            /* lattice and field mesh per axis: a power of two in [16, 1024] */
    double boxsize;        /* L */
    double amp, k0, slope, kcut;     /* P(k) = amp (k/k0)^slope exp(-(k/kcut)^2), P(0) = 0 */
    double bao_amp, bao_period;      /* times 1 + bao_amp sin(2 pi k / bao_period) */
    double f_growth;       /* linear redshift-space factor along z (0: real space) */
} hm_mock_config;

/* Standard normal noise of the mock with this seed, [n_side^3] FP64. */
int hm_mock_white_noise(hm_ctx *ctx, const hm_mock_config *cfg, unsigned long long seed, double *d_out);
/* n <= n_side^3 particles, d_pos [n][3] FP64.  Asynchronous on the
 * context's stream; the context keeps a workspace of 2 n_side^3 doubles and
 * four half spectra (0.8 GB at n_side = 256) until it is destroyed. */
int hm_mock_zeldovich(hm_ctx *ctx, const hm_mock_config *cfg, unsigned long long seed, long long n, double *d_pos);
/* Biased tracers: site s is kept when its uniform (Philox counter (s, 1, 0, 0),
 * a 53-bit fraction) is below rate max(0, 1 + b_lag delta_L(s)), delta_L the
 * linear field on the lattice and rate = mean_n / sum_s max(0, 1 + b_lag
 * delta_L(s)) (so mean_n is the expected count while rate (1 + b_lag delta_L)
 * stays below 1); kept sites in site order, moved
 * and wrapped as above.  d_shift (may be NULL) [n][3] receives the shift that
 * returns each particle to its site, -Psi (with the z factor), unwrapped.
 * *n_out = the number kept; more than capacity is an error (the first
 * capacity are written).  Synchronises the context's stream. */
int hm_mock_zeldovich_biased(hm_ctx *ctx, const hm_mock_config *cfg, unsigned long long seed, double mean_n,
                             double b_lag, long long capacity, double *d_pos, double *d_shift, long long *n_out);

/* A sky patch of the survey geometry: comoving distance shell [rmin, rmax]
 * (Mpc/h) times RA in [ra_min, ra_max) and Dec in [dec_min, dec_max)
 * (degrees).  Cartesian positions follow fits2dat_galaxy_weight.py:130-132:
 * x = r cos(Dec) cos(RA), y = r cos(Dec) sin(RA), z = r sin(Dec). */
typedef struct {
    double rmin, rmax;
    double ra_min, ra_max;
    double dec_min, dec_max;
} hm_sky_patch;
/* n points uniform in the volume of the patch (the random catalogue of a
 * survey-geometry mock; the reference reads its randoms from files,
 * struct/data/boss/fits2dat_galaxy_random_weight.py): point i from two
 * Philox4x32-10 blocks with counters (i, 3) and (i, 4) and key seed, r^3, RA
 * and sin(Dec) uniform.  d_pos [n][3] FP64; asynchronous on the context's
 * stream. */
int hm_mock_sky_randoms(hm_ctx *ctx, const hm_sky_patch *patch, unsigned long long seed, long long n, double *d_pos);

/* ---------------------------------------------------------------- */
/* Learned reconstruction.  No counterpart in the reference, whose
 * reconstruction is the plain one (particle_reconstruction.hpp:210-402).  A
 * residual network corrects the plain reconstruction's shift field: features
 * [v][4] = (norm Psi_a / psi_scale, a = 0..2; (rho / rho_bar - 1) / 8) of the
 * plain reconstruction (Psi as hm_recon_displace forms it, b1 and RG, norm =
 * V / N, no randoms) on an n_mesh^3 TSC mesh; four 3^3 convolutions with
 * periodic padding (4 -> 16 -> 16 -> 16 -> 3 channels, ReLU after the first
 * three), BF16 operands and FP32 accumulation on the matrix cores; shift
 * S_i = norm Psi(x_i) + psi_scale out(x_i), both gathered with the TSC
 * stencil of the reference's reconstruction gather (particle_reconstruction.hpp:
 * 334-388, so the plain shift is the reference's displacement); loss
 * mean_i |S_i - T_i|^2, whose mesh gradient is the library's mass assignment
 * of the residuals (the gather's adjoint).  Parameters, FP32, in the order W1 b1 W2 b2 W3
 * b3 W4 b4, W_l [27][cin][cout] with tap t = 9 (dx + 1) + 3 (dy + 1) + (dz + 1)
 * (input voxel v + (dx, dy, dz)); the last layer starts at zero, so a new
 * model predicts the plain shifts. */
typedef struct hm_lrecon hm_lrecon;
typedef struct {
    int n_mesh;                   /* per axis: a multiple of 16 and an FFT length of this library */
    double boxsize;
    double b1, RG;                /* the plain reconstruction's bias and Gaussian smoothing (Mpc/h) */
    double psi_scale;             /* unit of the features and the output (Mpc/h) */
    double lr;                    /* Adam step size */
    unsigned long long init_seed; /* He-normal initialisation of the first three layers */
    int precision;                /* of the plain reconstruction's transforms: HM_FP32 or HM_FP64 */
    int conv_cache_taps;          /* 0: convolution input tiles staged in LDS where they fit; 1: every tap read
                                     through the caches (the same results, bit for bit) */
} hm_lrecon_config;

hm_lrecon *hm_lrecon_create(hm_ctx *ctx, const hm_lrecon_config *cfg);
void hm_lrecon_destroy(hm_ctx *ctx, hm_lrecon *model);
long long hm_lrecon_n_params(const hm_lrecon *model);
int hm_lrecon_get_params(hm_ctx *ctx, hm_lrecon *model, float *out);
/* Also resets the Adam moments. */
int hm_lrecon_set_params(hm_ctx *ctx, hm_lrecon *model, const float *in);
/* The features of a catalogue (d_pos n x 3, in [0, L)), d_feat [n_mesh^3][4]. */
int hm_lrecon_features(hm_ctx *ctx, hm_lrecon *model, const double *d_pos, long long n, float *d_feat);
/* The network alone: d_x [n_mesh^3][4] network input (FP32, rounded to BF16),
 * d_out [n_mesh^3][4] (channels 0..2).  Backward: d_gout = d loss / d out,
 * grad (host) [n_params]. */
int hm_lrecon_net_forward(hm_ctx *ctx, hm_lrecon *model, const float *d_x, float *d_out);
int hm_lrecon_net_backward(hm_ctx *ctx, hm_lrecon *model, const float *d_x, const float *d_gout, float *grad);
/* Per-particle shifts [n][3]: plain and corrected (either may be NULL). */
int hm_lrecon_predict(hm_ctx *ctx, hm_lrecon *model, const double *d_pos, long long n, double *d_shift_plain,
                      double *d_shift_pred);
/* Losses of the plain and corrected shifts against d_target [n][3] and the
 * parameter gradient of the corrected one (grad host, may be NULL; d_gout
 * device [n_mesh^3][4], may be NULL: d loss / d out). */
int hm_lrecon_gradient(hm_ctx *ctx, hm_lrecon *model, const double *d_pos, const double *d_target, long long n,
                       double *loss_plain, double *loss_pred, float *grad, float *d_gout);
/* hm_lrecon_gradient, then one Adam step. */
int hm_lrecon_train_step(hm_ctx *ctx, hm_lrecon *model, const double *d_pos, const double *d_target, long long n,
                         double *loss_plain, double *loss_pred);

#ifdef __cplusplus
}
#endif
#endif
