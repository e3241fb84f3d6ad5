// hm_survey.hip -- survey-geometry P_l on one MI355X (main.cpp:109-166 and
// calcPowerSpectrum, powerspec.hpp:4-91).
//
// One execute takes a data and a random catalogue (unshifted Cartesian
// positions and weights) to the normalised P_l(k) the reference writes:
//   1. weight sums, alpha, the minimum of the randoms and the FFT-box shift;
//   2. the survey volume from the squared random density;
//   3. the Y_lm-weighted fluctuations D - alpha R for M = 0..l (the negative
//      M follow from F_{-M}(x) = (-1)^M conj F_M(x), real weights) and the
//      plain one for (0, 0), each assigned in 64-bit fixed point and
//      transformed forward;
//   4. the shot noise sums for M = 0..l;
//   5. one pass over the modes that can reach a k-bin that forms, per mode,
//      sum_M w_M (F_LM conj F_00 - SN_M S)/W^2 Y_{l,-M}(k) with the
//      reference's 3j weights, fine 1e-4 histogram and coarse fold.
#include "hm_assign.hpp"
#include "hm_bins.hpp"
#include "hm_fft_launch.hpp"
#include "hm_recon.hpp"

#include "../../include/hitomi_amd.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hm {

constexpr int SV_MAX_ELL = 8;

// ---------------------------------------------------------------------------
// Wigner 3j symbol (gsl_sf_coupling_3j, common.hpp:29) by the Racah sum with
// long double factorials; the arguments are small integers here.
// ---------------------------------------------------------------------------
static long double lfact(int n) {
    long double r = 1.0L;
    for (int i = 2; i <= n; i++) r *= (long double)i;
    return r;
}

double wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0 || j1 < 0 || j2 < 0 || j3 < 0) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    const long double delta =
        lfact(j1 + j2 - j3) * lfact(j1 - j2 + j3) * lfact(-j1 + j2 + j3) / lfact(j1 + j2 + j3 + 1);
    const long double pre = std::sqrt(delta * lfact(j1 + m1) * lfact(j1 - m1) * lfact(j2 + m2) * lfact(j2 - m2) *
                                      lfact(j3 + m3) * lfact(j3 - m3));
    const int tlo = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
    const int thi = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
    long double sum = 0.0L;
    for (int t = tlo; t <= thi; t++) {
        const long double den = lfact(t) * lfact(j3 - j2 + t + m1) * lfact(j3 - j1 + t - m2) * lfact(j1 + j2 - j3 - t) *
                                lfact(j1 - t - m1) * lfact(j2 - t + m2);
        sum += ((t & 1) ? -1.0L : 1.0L) / den;
    }
    const int ph = j1 - j2 - m3;
    return (double)(((ph % 2) != 0 ? -1.0L : 1.0L) * pre * sum);
}

// ---------------------------------------------------------------------------
// particle passes
// ---------------------------------------------------------------------------
// per-block partials: sum w_D, sum w_D^2, sum w_R, sum |w_D|, sum |w_R|; fixed order.
// w * w is the correctly rounded square that pow(w, 2) returns; the cubes
// (pow(w, 3), only the three-point shot noise uses them) only when asked for.
constexpr int SV_NSUM = 8;  // sum w_D, w_D^2, w_R, |w_D|, |w_R|, w_R^2, w_D^3, w_R^3
__global__ __launch_bounds__(256) void k_survey_wsums(const double* __restrict__ wD, long long nD,
                                                     const double* __restrict__ wR, long long nR, int cubes,
                                                     double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double sm[SV_NSUM][256];
    double v[SV_NSUM] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const long long st = (long long)gridDim.x * 256;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < nD; q += st) {
        const double w = wD[q];
        v[0] += w;
        v[1] += w * w;
        v[3] += fabs(w);
        if (cubes) v[6] += pow(w, 3);
    }
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < nR; q += st) {
        const double w = wR[q];
        v[2] += w;
        v[4] += fabs(w);
        v[5] += w * w;
        if (cubes) v[7] += pow(w, 3);
    }
    for (int k = 0; k < SV_NSUM; k++) sm[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < SV_NSUM; k++) sm[k][threadIdx.x] += sm[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < SV_NSUM) part[blockIdx.x * SV_NSUM + threadIdx.x] = sm[threadIdx.x][0];
}

// per-block minimum of the positions per axis (order independent)
__global__ __launch_bounds__(256) void k_survey_minpos(const double* __restrict__ pos, long long n,
                                                      double* __restrict__ part) {
    __shared__ double sm[3][256];
    double m[3] = {INFINITY, INFINITY, INFINITY};
    const long long st = (long long)gridDim.x * 256;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += st)
        for (int a = 0; a < 3; a++) m[a] = fmin(m[a], pos[3 * q + a]);
    for (int a = 0; a < 3; a++) sm[a][threadIdx.x] = m[a];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int a = 0; a < 3; a++) sm[a][threadIdx.x] = fmin(sm[a][threadIdx.x], sm[a][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = sm[threadIdx.x][0];
}

// Shifted positions of D then R (particle.hpp:204-218) in one array.
__global__ __launch_bounds__(256) void k_survey_shift(const double* __restrict__ pD, long long nD,
                                                     const double* __restrict__ pR, long long nR, double dx, double dy,
                                                     double dz, double* __restrict__ out) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= nD + nR) return;
    const double* s = q < nD ? pD + 3 * q : pR + 3 * (q - nD);
    out[3 * q] = s[0] - dx;
    out[3 * q + 1] = s[1] - dy;
    out[3 * q + 2] = s[2] - dz;
}

__global__ __launch_bounds__(256) void k_shift_inplace(double* __restrict__ pos, long long n, double dx, double dy,
                                                      double dz) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    pos[3 * q] -= dx;
    pos[3 * q + 1] -= dy;
    pos[3 * q + 2] -= dz;
}

// LOS of an unshifted position (particle.hpp:355-375); the lines of sight are
// always taken from the caller's input positions, which nothing modifies, so
// they stay those of the unreconstructed catalogues (main.cpp:156-177)
__device__ inline void los_of(const double* p, double* l) {
#pragma clang fp contract(off)
    const double m = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    l[0] = p[0] / m;
    l[1] = p[1] / m;
    l[2] = p[2] / m;
}

// Complex weights of one field over D then R: w Y(l, M)(los) for D and
// -alpha w Y(l, M)(los) for R (field.hpp:186-228; D - alpha R is formed by
// the fixed-point sum).  mode 1: the random catalogue alone with (w, 0)
// (field.hpp:476-485); mode 2: the random catalogue alone with w Y(l, M);
// mode 3 (l = M = 0): mode 0 in re and mode 1 in im.
__global__ __launch_bounds__(256) void k_survey_weights(const double* __restrict__ pD, const double* __restrict__ wD,
                                                       long long nD, const double* __restrict__ pR,
                                                       const double* __restrict__ wR, long long nR, double alpha, int l,
                                                       int M, int mode, double* __restrict__ out) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (mode == 1) {
        if (q < nR) {
            out[2 * q] = wR[q];
            out[2 * q + 1] = 0.0;
        }
        return;
    }
    if (mode == 3) {  // Y_00 = 1: D - alpha R in re, and the random catalogue alone (w, field.hpp:476-485) in im
        if (q < nD + nR) {
            const bool d = q < nD;
            out[2 * q] = d ? wD[q] : -(alpha * wR[q - nD]);
            out[2 * q + 1] = d ? 0.0 : wR[q - nD];
        }
        return;
    }
    if (mode == 2) {  // the random catalogue alone with w Y(l, M)(los) (field.hpp:230-247)
        if (q < nR) {
            double los[3];
            los_of(pR + 3 * q, los);
            const Cplx y = ylm_red(l, M, los[0], los[1], los[2]);
            out[2 * q] = y.re * wR[q];
            out[2 * q + 1] = y.im * wR[q];
        }
        return;
    }
    if (q >= nD + nR) return;
    const bool d = q < nD;
    const double* p = d ? pD + 3 * q : pR + 3 * (q - nD);
    double los[3];
    los_of(p, los);
    const Cplx y = ylm_red(l, M, los[0], los[1], los[2]);
    const double w = d ? wD[q] : wR[q - nD];
    double re = y.re * w, im = y.im * w;
    if (!d) {
        re = -(alpha * re);
        im = -(alpha * im);
    }
    out[2 * q] = re;
    out[2 * q + 1] = im;
}

// Shot-noise sums of field.hpp:1164-1188 for M = 0..l: per block
// [M][sum_D re, sum_D im, sum_R re, sum_R im], fixed order.
__global__ __launch_bounds__(256) void k_survey_sn(const double* __restrict__ pD, const double* __restrict__ wD,
                                                  long long nD, const double* __restrict__ pR,
                                                  const double* __restrict__ wR, long long nR, int l,
                                                  double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double sm[256];
    const int nm = l + 1;
    double acc[4 * (SV_MAX_ELL + 1)];
    for (int i = 0; i < 4 * nm; i++) acc[i] = 0.0;
    const long long st = (long long)gridDim.x * 256;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < nD + nR; q += st) {
        const bool d = q < nD;
        const double* p = d ? pD + 3 * q : pR + 3 * (q - nD);
        const double w = d ? wD[q] : wR[q - nD];
        const double w2 = pow(w, 2);
        double los[3];
        los_of(p, los);
        for (int M = 0; M < nm; M++) {
            const Cplx y = ylm_red(l, M, los[0], los[1], los[2]);
            acc[4 * M + (d ? 0 : 2)] += w2 * y.re;
            acc[4 * M + (d ? 1 : 3)] += w2 * y.im;
        }
    }
    for (int i = 0; i < 4 * nm; i++) {
        sm[threadIdx.x] = acc[i];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(long long)blockIdx.x * 4 * nm + i] = sm[0];
        __syncthreads();
    }
}

// sum_D w^pw Y(L, M)(los) and sum_R w^pw Y(L, M)(los) for M = -L..L (field.hpp:1202-1240; pw = 3, or 2
// for the window): per block [M + L][D re, D im, R re, R im], fixed order
__global__ __launch_bounds__(256) void k_survey_w3y(const double* __restrict__ pD, const double* __restrict__ wD,
                                                   long long nD, const double* __restrict__ pR,
                                                   const double* __restrict__ wR, long long nR, int L, int pw,
                                                   double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double sm[256];
    const int nm = 2 * L + 1;
    double acc[4 * (2 * SV_MAX_ELL + 1)];
    for (int i = 0; i < 4 * nm; i++) acc[i] = 0.0;
    const long long st = (long long)gridDim.x * 256;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < nD + nR; q += st) {
        const bool d = q < nD;
        const double* p = d ? pD + 3 * q : pR + 3 * (q - nD);
        const double w3 = pow(d ? wD[q] : wR[q - nD], pw);
        double los[3];
        los_of(p, los);
        for (int M = -L; M <= L; M++) {
            const Cplx y = ylm_red(L, M, los[0], los[1], los[2]);
            acc[4 * (M + L) + (d ? 0 : 2)] += w3 * y.re;
            acc[4 * (M + L) + (d ? 1 : 3)] += w3 * y.im;
        }
    }
    for (int i = 0; i < 4 * nm; i++) {
        sm[threadIdx.x] = acc[i];
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(long long)blockIdx.x * 4 * nm + i] = sm[0];
        __syncthreads();
    }
}

// sum over cells of rho^2 dV, rho = (q 2^-s)(1/dV) (field.hpp:487-492); per-block partials
__global__ __launch_bounds__(256) void k_survey_sqsum(const long long* __restrict__ q, long long n, double inv_scale,
                                                     double factor, double dV, double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double sm[256];
    // four cells per step, each into its own sum, so four loads are in flight
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const long long st = (long long)gridDim.x * 256;
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * st < n; i += 4 * st) {
        long long v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = q[i + u * st];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double r = ((double)v[u] * inv_scale) * factor;
            a[u] += r * r * dV;
        }
    }
    for (; i < n; i += st) {
        const double r = ((double)q[i] * inv_scale) * factor;
        a[0] += r * r * dV;
    }
    sm[threadIdx.x] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// ---------------------------------------------------------------------------
// fused binning over M (calcPowerSpectrum, field.hpp:765-852, summed with the
// weights of powerspec.hpp:44-59)
// ---------------------------------------------------------------------------
struct SurveyBinArgs {
    int n[3], s[3], c[3];
    long long nrows, ng;
    AxisTabs t;
    int wp;
    const int2* f2b;
    int nk;
    int l;
    double w[2 * SV_MAX_ELL + 1];       // w_M, M = -l..l
    double sn[2 * (SV_MAX_ELL + 1)];    // SN_M (re, im), M = 0..l
};

template <typename T>
__global__ __launch_bounds__(256) void k_bin_survey(const C<T>* __restrict__ f00, const C<T>* __restrict__ fl,
                                                   SurveyBinArgs a, double* __restrict__ part,
                                                   long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.nrows) return;
    const int i = support_map((int)(row / a.c[1]), a.s[0], a.n[0]);
    const int j = support_map((int)(row % a.c[1]), a.s[1], a.n[1]);
    const int im = (a.n[0] - i) % a.n[0], jm = (a.n[1] - j) % a.n[1];
    double* P = part + row * 2 * (long long)a.nk;
    long long* Cn = cnt + row * (long long)a.nk;
    for (int kc = 0; kc < a.c[2]; kc++) {
        const int k = support_map(kc, a.s[2], a.n[2]);
        const double kx = a.t.kv[0][i], ky = a.t.kv[1][j], kz = a.t.kv[2][k];
        const double kmag = kmag3(kx, ky, kz);
        const int kindex = (int)(kmag / FINE_DK + 0.5);
        if (kindex >= FINE_N) continue;
        const int2 b = a.f2b[kindex];
        if (b.x < 0) continue;
        const long long lin = ((long long)i * a.n[1] + j) * a.n[2] + k;
        const long long mir = ((long long)im * a.n[1] + jm) * a.n[2] + (a.n[2] - k) % a.n[2];
        const double S = (a.t.sf[0][i] * a.t.sf[1][j]) * a.t.sf[2][k];
        const double W = wpow((a.t.wf[0][i] * a.t.wf[1][j]) * a.t.wf[2][k], a.wp);
        const double W2 = pow(W, 2);
        const C<T> g = f00[lin];
        const double gr = (double)g.x, gi = -(double)g.y;  // conj F_00
        double accr = 0.0, acci = 0.0;
        for (int M = -a.l; M <= a.l; M++) {
            const double wM = a.w[M + a.l];
            if (wM == 0.0) continue;
            const int am = M < 0 ? -M : M;
            const double sg = (M < 0 && (am & 1)) ? -1.0 : 1.0;
            double fr, fi, sr, si;
            if (M >= 0) {
                const C<T> f = fl[(long long)am * a.ng + lin];
                fr = (double)f.x;
                fi = (double)f.y;
                sr = a.sn[2 * am];
                si = a.sn[2 * am + 1];
            } else {
                const C<T> f = fl[(long long)am * a.ng + mir];
                fr = sg * (double)f.x;
                fi = -(sg * (double)f.y);
                sr = sg * a.sn[2 * am];
                si = -(sg * a.sn[2 * am + 1]);
            }
            // (F conj F00 - SN S) / W^2
            double pr = fr * gr - fi * gi, pi = fr * gi + fi * gr;
            pr -= sr * S;
            pi -= si * S;
            pr /= W2;
            pi /= W2;
            const Cplx y = ylm_red(a.l, -M, kx, ky, kz);
            const double vr = pr * y.re - pi * y.im, vi = pr * y.im + pi * y.re;
            accr += wM * vr;
            acci += wM * vi;
        }
        for (int bb = b.x; bb <= b.y; bb++) {
            P[2 * bb] += accr;
            P[2 * bb + 1] += acci;
            Cn[bb] += 1;
        }
    }
}

// ---------------------------------------------------------------------------
// correlation functions (calcCorrelationFunction, field.hpp:854-977, summed
// with the weights of powerspec.hpp:134-147): per M, g_M(k) = (F_LM conj F_00
// - SN_M S)/W^2 / V on the full mesh, its inverse transform restricted to the
// separations that reach an r-bin, and the binning of w_M xi_M(r) Y_{l,-M}(r)
// added to the partial sums of the previous M
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_survey_g(const C<T>* __restrict__ f00, const C<T>* __restrict__ fl, int n0,
                                                 int n1, int n2, AxisTabs t, int wp, int M, double snr, double sni,
                                                 double inv_v, C<T>* __restrict__ g) {
#pragma clang fp contract(off)
    const long long ng = (long long)n0 * n1 * n2;
    const long long st = (long long)gridDim.x * 256;
    const int am = M < 0 ? -M : M;
    const double sg = (M < 0 && (am & 1)) ? -1.0 : 1.0;
    for (long long lin = (long long)blockIdx.x * 256 + threadIdx.x; lin < ng; lin += st) {
        const int k = (int)(lin % n2);
        const long long r = lin / n2;
        const int j = (int)(r % n1), i = (int)(r / n1);
        double fr, fi;
        if (M >= 0) {
            const C<T> f = fl[(long long)am * ng + lin];
            fr = (double)f.x;
            fi = (double)f.y;
        } else {
            const long long mir = ((long long)((n0 - i) % n0) * n1 + (n1 - j) % n1) * n2 + (n2 - k) % n2;
            const C<T> f = fl[(long long)am * ng + mir];
            fr = sg * (double)f.x;
            fi = -(sg * (double)f.y);
        }
        const C<T> h = f00[lin];
        const double hr = (double)h.x, hi = -(double)h.y;
        double pr = fr * hr - fi * hi, pi = fr * hi + fi * hr;
        const double S = (t.sf[0][i] * t.sf[1][j]) * t.sf[2][k];
        pr -= snr * S;
        pi -= sni * S;
        const double W = wpow((t.wf[0][i] * t.wf[1][j]) * t.wf[2][k], wp);
        const double W2 = pow(W, 2);
        pr /= W2;
        pi /= W2;
        g[lin] = mk<T>((T)(inv_v * pr), (T)(inv_v * pi));
    }
}

struct SurveyXiArgs {
    int n[3], s[3], c[3];
    double dr[3];
    long long nrows;
    const int2* f2b;
    int nr, l, m1, count;
    double w;
};

template <typename T>
__global__ __launch_bounds__(256) void k_bin_survey_xi(const C<T>* __restrict__ xi, SurveyXiArgs a,
                                                      double* __restrict__ part, long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.nrows) return;
    const int i = support_map((int)(row / a.c[1]), a.s[0], a.n[0]);
    const int j = support_map((int)(row % a.c[1]), a.s[1], a.n[1]);
    const double rx = mesh_vec(i, a.n[0], a.dr[0]), ry = mesh_vec(j, a.n[1], a.dr[1]);
    double* P = part + row * 2 * (long long)a.nr;
    long long* Cn = cnt + row * (long long)a.nr;
    for (int kc = 0; kc < a.c[2]; kc++) {
        const int k = support_map(kc, a.s[2], a.n[2]);
        const double rz = mesh_vec(k, a.n[2], a.dr[2]);
        const double rmag = kmag3(rx, ry, rz);
        const int ri = (int)(rmag / FINE_DR + 0.5);
        if (ri >= FINE_NR) continue;
        const int2 b = a.f2b[ri];
        if (b.x < 0) continue;
        const C<T> v = xi[((long long)i * a.n[1] + j) * a.n[2] + k];
        const Cplx y = ylm_red(a.l, a.m1, rx, ry, rz);
        const double vr = (double)v.x * y.re - (double)v.y * y.im, vi = (double)v.x * y.im + (double)v.y * y.re;
        for (int bb = b.x; bb <= b.y; bb++) {
            P[2 * bb] += a.w * vr;
            P[2 * bb + 1] += a.w * vi;
            if (a.count) Cn[bb] += 1;
        }
    }
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct SurveyPlanImpl {
    hm_survey_config cfg;
    Mesh mesh;
    int nk = 0, l = 0, nf = 0;  // fields: index 0 = (0, 0), 1 + M = (l, M)
    std::vector<double> kbin, lower, upper, wM;
    int s[3];
    long long nrows = 0;
    long long tab_off[3];
    int measure = HM_SURVEY_PK;
    int nr = 0;            // r-bins of the correlation functions
    int rs[3] = {-1, -1, -1};
    long long nrows_r = 0;
    SortWork sort;
    ReconWork rw;
    DevBuf tab, f2b, f2b_r, pos, wts, fre, fim, fields, gwork, part, cnt, red, cnt_red, scratch;
    size_t bytes = 0;
    BoxPlanImpl* box = nullptr;  // bk / 3PCF / window3PCF: the triple-product machinery on weighted catalogues
};

static bool delegated(int measure) {
    return measure == HM_SURVEY_BK || measure == HM_SURVEY_ZETA || measure == HM_SURVEY_ZETA_WINDOW;
}

static int sv_alloc(hm_ctx* ctx, SurveyPlanImpl* p, DevBuf& b, size_t bytes) {
    const size_t before = b.bytes;
    if (dev_alloc(ctx, b, bytes)) return -1;
    p->bytes += b.bytes - before;
    return 0;
}

void survey_destroy(SurveyPlanImpl* p) {
    if (!p) return;
    DevBuf* bufs[] = {&p->tab, &p->f2b, &p->f2b_r, &p->pos, &p->wts, &p->fre, &p->fim, &p->fields,
                      &p->gwork, &p->part, &p->cnt, &p->red, &p->cnt_red, &p->scratch};
    for (DevBuf* b : bufs) dev_free(*b);
    sort_work_free(p->sort);
    recon_work_free(p->rw);
    if (p->box) plan_destroy(p->box);
    delete p;
}

SurveyPlanImpl* survey_create(hm_ctx* ctx, const hm_survey_config* cfg) {
    auto* p = new SurveyPlanImpl();
    p->cfg = *cfg;
    Mesh& m = p->mesh;
    auto fail = [&](const std::string& s) -> SurveyPlanImpl* {
        ctx->set_error(s);
        survey_destroy(p);
        return nullptr;
    };
    for (int a = 0; a < 3; a++) {
        m.n[a] = cfg->mesh.n_mesh[a];
        m.L[a] = cfg->mesh.boxsize[a];
        if (!fft_length_supported(m.n[a])) return fail("survey plan: unsupported n_mesh " + std::to_string(m.n[a]));
        if (!(m.L[a] > 0)) return fail("survey plan: boxsize must be > 0");
    }
    m.assign = cfg->mesh.assign;
    if (m.assign < 0 || m.assign > 2) return fail("survey plan: assign must be NGP, CIC or TSC");
    if (cfg->precision != HM_FP32 && cfg->precision != HM_FP64) return fail("survey plan: precision must be FP32 or FP64");
    if (cfg->n_kbin < 2 || cfg->n_kbin > 256) return fail("survey plan: n_kbin must be in [2, 256]");
    if (cfg->ell < 0 || cfg->ell > SV_MAX_ELL) return fail("survey plan: ell must be in [0, 8]");
    if (cfg->recon) {
        for (int a = 0; a < 3; a++)
            if (!fft_length_supported(cfg->n_mesh_recon[a])) return fail("survey plan: unsupported n_mesh_recon");
        if (!(cfg->b1 != 0.0)) return fail("survey plan: b1_fid must be non-zero");
    }
    p->measure = cfg->measure;
    if (p->measure < HM_SURVEY_PK || p->measure > HM_SURVEY_ZETA_WINDOW) return fail("survey plan: unknown measure");
    if (delegated(p->measure)) {
        // calcBiSpectrum / calcThreePoint(Window)CorrelationFunction (bispec.hpp:4-874) on the box machinery:
        // with ELL = 0 (no Y_LM(los) weight) the box estimators of the weighted catalogues; ELL > 0 adds per
        // M the Y_LM(los)-weighted third field and its shot-noise terms
        hm_box_config bc;
        std::memset(&bc, 0, sizeof(bc));
        bc.mesh = cfg->mesh;
        bc.precision = cfg->precision;
        bc.n_kbin = cfg->n_kbin;
        bc.kmin = cfg->kmin;
        bc.kmax = cfg->kmax;
        bc.regrid = 1;
        bc.prune = cfg->prune;
        bc.n_rbin = cfg->n_rbin;
        bc.rmin = cfg->rmin;
        bc.rmax = cfg->rmax;
        const int ells[3] = {cfg->ell1, cfg->ell2, cfg->ell};
        for (int a = 0; a < 3; a++) bc.bk_ell[a] = bc.zeta_ell[a] = ells[a];
        bc.bk = p->measure == HM_SURVEY_BK ? 1 : 0;
        bc.zeta = p->measure == HM_SURVEY_BK ? 0 : 1;
        p->box = plan_create(ctx, &bc);
        if (!p->box) return fail(std::string(ctx->err));
    }
    p->nk = cfg->n_kbin;
    p->l = cfg->ell;
    p->nf = p->l == 0 ? 1 : p->l + 2;
    bin_edges(cfg->kmin, cfg->kmax, p->nk, p->kbin, p->lower, p->upper);
    std::vector<int2> f2b;
    const std::string e = fine_to_coarse(p->lower, p->upper, FINE_N, FINE_DK, f2b);
    if (!e.empty()) return fail("survey plan: " + e);
    // weights of powerspec.hpp:44-52 with ell1 = ELL = l, ell2 = 0: only m1 = -M survives
    p->wM.assign(2 * p->l + 1, 0.0);
    for (int M = -p->l; M <= p->l; M++)
        for (int m1 = -p->l; m1 <= p->l; m1++) {
            double w = wigner_3j(p->l, 0, p->l, 0, 0, 0) * wigner_3j(p->l, 0, p->l, m1, 0, M);
            w *= (double)(2 * p->l + 1) * (double)(2 * p->l + 1);
            if (fabs(w) < 1.0e-10) continue;
            if (m1 != -M) return fail("survey plan: unexpected 3j selection");
            p->wM[M + p->l] = w;
        }
    double umax = 0;
    for (double u : p->upper) umax = std::max(umax, u);
    for (int a = 0; a < 3; a++) {
        const double dk = 2.0 * M_PI / m.L[a];
        const long long s = (long long)std::floor((umax + FINE_DK) / dk) + 1;
        p->s[a] = (!cfg->prune || 2 * s + 1 >= m.n[a] || s >= m.n[a] / 2) ? -1 : (int)s;
    }
    p->nrows = (long long)support_count(p->s[0], m.n[0]) * support_count(p->s[1], m.n[1]);
    std::vector<int2> f2b_r;
    if (p->measure != HM_SURVEY_PK) {
        if (cfg->n_rbin < 2 || cfg->n_rbin > 256) return fail("survey plan: n_rbin must be in [2, 256]");
        if (!(cfg->rmax > cfg->rmin)) return fail("survey plan: rmax must exceed rmin");
        p->nr = cfg->n_rbin;
        std::vector<double> rb, rlo, rup;
        bin_edges(cfg->rmin, cfg->rmax, p->nr, rb, rlo, rup);
        const std::string er = fine_to_coarse(rlo, rup, FINE_NR, FINE_DR, f2b_r);
        if (!er.empty()) return fail("survey plan: " + er);
        double rumax = 0.0;
        for (double u : rup) rumax = std::max(rumax, u);
        // a separation reaches a bin only if |r| < rumax + 0.25 (fine centre rule)
        for (int a = 0; a < 3; a++) {
            const double dr = m.L[a] / (double)m.n[a];
            const long long sr = (long long)std::floor((rumax + FINE_DR) / dr) + 1;
            p->rs[a] = (!cfg->prune || 2 * sr + 1 >= m.n[a] || sr >= m.n[a] / 2) ? -1 : (int)sr;
        }
        p->nrows_r = (long long)support_count(p->rs[0], m.n[0]) * support_count(p->rs[1], m.n[1]);
    }
    const int nbin = p->measure == HM_SURVEY_PK ? p->nk : p->nr;
    const long long nrow_bins = p->measure == HM_SURVEY_PK ? p->nrows : p->nrows_r;
    std::vector<double> tabs;
    host_axis_tables(m, tabs, p->tab_off);
    const size_t elem = cfg->precision == HM_FP64 ? 16 : 8;
    const long long ng = m.ng();
    if (delegated(p->measure)) {  // the survey volume grid only
        if (sv_alloc(ctx, p, p->tab, sizeof(double) * tabs.size()) || sv_alloc(ctx, p, p->f2b, sizeof(int2) * FINE_N) ||
            sv_alloc(ctx, p, p->fre, sizeof(long long) * (size_t)ng))
            return fail(ctx->err);
        return p;
    }
    if (sv_alloc(ctx, p, p->tab, sizeof(double) * tabs.size()) || sv_alloc(ctx, p, p->f2b, sizeof(int2) * FINE_N) ||
        sv_alloc(ctx, p, p->fre, sizeof(long long) * (size_t)ng) ||
        sv_alloc(ctx, p, p->fim, sizeof(long long) * (size_t)ng) ||
        sv_alloc(ctx, p, p->fields, elem * (size_t)ng * (size_t)p->nf) ||
        sv_alloc(ctx, p, p->part, sizeof(double) * 2 * (size_t)nbin * (size_t)nrow_bins) ||
        sv_alloc(ctx, p, p->cnt, sizeof(long long) * (size_t)nbin * (size_t)nrow_bins) ||
        sv_alloc(ctx, p, p->red, sizeof(double) * 2 * (size_t)nbin) ||
        sv_alloc(ctx, p, p->cnt_red, sizeof(long long) * (size_t)nbin))
        return fail(ctx->err);
    if (p->measure != HM_SURVEY_PK &&
        (sv_alloc(ctx, p, p->gwork, elem * (size_t)ng) || sv_alloc(ctx, p, p->f2b_r, sizeof(int2) * FINE_NR)))
        return fail(ctx->err);
    if (hipMemcpy(p->tab.p, tabs.data(), sizeof(double) * tabs.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->f2b.p, f2b.data(), sizeof(int2) * FINE_N, hipMemcpyHostToDevice) != hipSuccess)
        return fail("survey plan: table upload failed");
    if (p->measure != HM_SURVEY_PK &&
        hipMemcpy(p->f2b_r.p, f2b_r.data(), sizeof(int2) * FINE_NR, hipMemcpyHostToDevice) != hipSuccess)
        return fail("survey plan: table upload failed");
    return p;
}

size_t survey_bytes(const SurveyPlanImpl* p) {
    return p ? p->bytes + p->sort.bytes() + recon_work_bytes(p->rw) + (p->box ? plan_bytes(p->box) : 0) : 0;
}

// Complex source of a z-row transform from fixed-point cell sums: the value of
// k_fixed_to_complex (((q * 2^-s) / dV) * dV, im from the second grid or 0).
template <typename T> struct FixedCplxSrc {
    const long long* re;
    const long long* im;
    int n2;
    double inv_scale, factor, dV;
    __device__ inline C<T> operator()(long long r, int m) const {
#pragma clang fp contract(off)
        const long long i = r * n2 + m;
        const double a = (((double)re[i] * inv_scale) * factor) * dV;
        const double b = im ? (((double)im[i] * inv_scale) * factor) * dV : 0.0;
        return mk<T>((T)a, (T)b);
    }
};

static unsigned red_blocks(long long n) { return (unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 1024)); }

template <typename T>
static int survey_run(hm_ctx* ctx, SurveyPlanImpl* p, const double* pD, const double* wD, long long nD,
                      const double* pR, const double* wR, long long nR, hm_survey_info* info, double* pk_out,
                      long long* nmode_out) {
    const Mesh& m = p->mesh;
    const long long N = nD + nR, ng = m.ng();
    const double dV = m.volume() / (double)ng;
    if (nD <= 0 || nR <= 0) {
        ctx->set_error("survey execute: data and randoms must be non-empty");
        return -1;
    }
    const unsigned rb = red_blocks(std::max(nD, nR));
    const unsigned rbg = red_blocks(ng);
    const int nsn = 4 * (p->l + 1);
    const size_t scratch = sizeof(double) * std::max<size_t>({(size_t)rb * SV_NSUM, (size_t)rb * 3, (size_t)rb * nsn,
                                                             (size_t)rbg, 64});
    if (sv_alloc(ctx, p, p->scratch, scratch + sizeof(double) * 64) || sv_alloc(ctx, p, p->pos, sizeof(double) * 3 * (size_t)N) ||
        sv_alloc(ctx, p, p->wts, sizeof(double) * 2 * (size_t)N))
        return -1;
    double* part = (double*)p->scratch.p;
    double* red = part + scratch / sizeof(double);
    // 1. weight sums (particle.hpp:334-353, 284-299) and the randoms' minimum (particle.hpp:160-200)
    std::vector<double> hp((size_t)rb * 3);
    if (ctx->launch("survey_wsums", [&] {
            hipLaunchKernelGGL(k_survey_wsums, dim3(rb), dim3(256), 0, ctx->stream, wD, nD, wR, nR,
                               delegated(p->measure) ? 1 : 0, part);
        }))
        return -1;
    if (reduce_cols<double>(ctx, "survey_reduce", part, rb, SV_NSUM, red)) return -1;
    double sums[SV_NSUM], mn[3];
    HM_HIP_CHECK(ctx, hipMemcpyAsync(sums, red, sizeof(sums), hipMemcpyDeviceToHost, ctx->stream));
    if (ctx->launch("survey_minpos", [&] {
            hipLaunchKernelGGL(k_survey_minpos, dim3(rb), dim3(256), 0, ctx->stream, pR, nR, part);
        }))
        return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(hp.data(), part, sizeof(double) * 3 * rb, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int a = 0; a < 3; a++) {
        mn[a] = hp[a];
        for (unsigned b = 1; b < rb; b++) mn[a] = std::min(mn[a], hp[3 * b + a]);
    }
    const double swD = sums[0], sw2D = sums[1], swR = sums[2], awD = sums[3], awR = sums[4], sw2R = sums[5],
                 sw3D = sums[6], sw3R = sums[7];
    if (swD == 0.0 || swR == 0.0) {
        ctx->set_error("survey execute: zero total weight (particle.hpp:274-276)");
        return -1;
    }
    const double alpha = swD / swR;  // particle.hpp:278
    double dP[3];
    for (int a = 0; a < 3; a++) dP[a] = mn[a] - 1.0 * m.L[a] / (double)m.n[a];  // particle.hpp:236-239, factor 1
    double* pos = (double*)p->pos.p;
    double* wts = (double*)p->wts.p;
    const unsigned pb = (unsigned)((N + 255) / 256);
    if (ctx->launch("survey_shift", [&] {
            hipLaunchKernelGGL(k_survey_shift, dim3(pb), dim3(256), 0, ctx->stream, pD, nD, pR, nR, dP[0], dP[1], dP[2],
                               pos);
        }))
        return -1;
    AssignGeom ag;
    for (int a = 0; a < 3; a++) {
        ag.n[a] = m.n[a];
        ag.L[a] = m.L[a];
    }
    ag.ng = ng;
    ag.assign = m.assign;
    {
        const size_t before = p->sort.bytes();
        if (sort_work_alloc(ctx, p->sort, ag, N, true)) return -1;
        p->bytes += p->sort.bytes() - before;
    }
    long long* fre = (long long*)p->fre.p;
    long long* fim = (long long*)p->fim.p;
    // 2. survey volume (field.hpp:449-503): Vsurvey = (sum w_R)^2 / sum_x rho_R(x)^2 dV from a fixed-point grid
    // of the randoms with weights w_R.  Without reconstruction the P_l / xi_l fields below share their sort by
    // cell with it: the grid is the imaginary channel of the first field (the window's first field is that grid).
    double vs = 0.0, nbar = 0.0;
    auto volume_from = [&](const long long* q, double scale) -> int {
        if (ctx->launch("survey_sqsum", [&] {
                hipLaunchKernelGGL(k_survey_sqsum, dim3(rbg), dim3(256), 0, ctx->stream, q, ng, 1.0 / scale, 1.0 / dV,
                                   dV, part);
            }))
            return -1;
        if (reduce_cols<double>(ctx, "survey_reduce", part, rbg, 1, red)) return -1;
        double isq = 0.0;
        HM_HIP_CHECK(ctx, hipMemcpyAsync(&isq, red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        vs = swR * swR / isq;           // field.hpp:499
        nbar = swD * swD / vs / sw2D;   // particle.hpp:296
        return 0;
    };
    const bool fold_volume = !p->cfg.recon && !delegated(p->measure);
    if (!fold_volume) {
        if (ctx->launch("survey_weights", [&] {
                hipLaunchKernelGGL(k_survey_weights, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, ctx->stream, pD,
                                   wD, nD, pR, wR, nR, alpha, 0, 0, 1, wts + 2 * nD);
            }))
            return -1;
        const double sv_scale = fixed_scale_for(awR);
        if (assign_sorted(ctx, ag, pos + 3 * nD, wts + 2 * nD, nR, sv_scale, fre, nullptr, p->sort)) return -1;
        if (volume_from(fre, sv_scale)) return -1;
    }
    if (p->cfg.recon) {
        // main.cpp:172-183: reconstruct on the n_mesh_recon mesh, then shift into the FFT box again
        Mesh mr;
        for (int a = 0; a < 3; a++) {
            mr.n[a] = p->cfg.n_mesh_recon[a];
            mr.L[a] = m.L[a];
        }
        mr.assign = m.assign;
        if (recon_displace<T>(ctx, mr, p->cfg.b1, p->cfg.RG, pos, wD, nD, pos + 3 * nD, wR, nR, alpha, vs / swD, p->rw))
            return -1;
        if (ctx->launch("survey_minpos", [&] {
                hipLaunchKernelGGL(k_survey_minpos, dim3(rb), dim3(256), 0, ctx->stream, pos + 3 * nD, nR, part);
            }))
            return -1;
        HM_HIP_CHECK(ctx, hipMemcpyAsync(hp.data(), part, sizeof(double) * 3 * rb, hipMemcpyDeviceToHost, ctx->stream));
        HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        double d2[3];
        for (int a = 0; a < 3; a++) {
            double v = hp[a];
            for (unsigned b = 1; b < rb; b++) v = std::min(v, hp[3 * b + a]);
            d2[a] = v - 1.0 * m.L[a] / (double)m.n[a];
        }
        if (ctx->launch("survey_shift", [&] {
                hipLaunchKernelGGL(k_shift_inplace, dim3(pb), dim3(256), 0, ctx->stream, pos, N, d2[0], d2[1], d2[2]);
            }))
            return -1;
    }
    if (delegated(p->measure)) {
        // 3'. bispectrum / three-point function (bispec.hpp:4-874): w (D - alpha R) and w^2 (D + alpha^2 R),
        // shot noise sum_D w^3 - alpha^3 sum_R w^3, or R alone with sum_R w^2 for the window; norm
        // Vsurvey^2 / (sum w)^3 (particle.hpp:319-332); ELL > 0 weights the third field by Y_LM(los)
        const bool win = p->measure == HM_SURVEY_ZETA_WINDOW;
        BoxWeighted bw;
        const double sw = win ? swR : swD;
        bw.norm = (vs / sw / sw) * (vs / sw);
        if (win) {
            bw.wD = wR;
            bw.wR = nullptr;
            bw.c1 = bw.c2 = 0.0;
            bw.abs1 = awR;
            bw.abs2 = sw2R;
            bw.sn3 = sw2R;
        } else {
            bw.wD = wD;
            bw.wR = wR;
            bw.c1 = -alpha;
            bw.c2 = alpha * alpha;
            bw.abs1 = awD + std::fabs(alpha) * awR;
            bw.abs2 = sw2D + alpha * alpha * sw2R;
            bw.sn3 = sw3D - pow(alpha, 3) * sw3R;
        }
        std::vector<double> snLM;
        if (p->l > 0) {
            // Y_LM(los)-weighted shot noise of every M (field.hpp:1202-1240): sum_D w^3 Y - alpha^3 sum_R w^3 Y,
            // or sum_R w^2 Y for the window; lines of sight from the input (unshifted, unreconstructed) positions
            const int L = p->l, nm = 2 * L + 1;
            const long long n1 = win ? nR : nD, n2 = win ? 0 : nR;
            const double* p1 = win ? pR : pD;
            const double* w1 = win ? wR : wD;
            const double a3 = win ? 0.0 : pow(alpha, 3);
            const unsigned nb = (unsigned)std::max<long long>(1, std::min<long long>((n1 + n2 + 255) / 256, 256));
            DevBuf yp;
            if (dev_alloc(ctx, yp, sizeof(double) * 4 * (size_t)nm * nb)) return -1;
            int rc = ctx->launch("survey_w3y", [&] {
                hipLaunchKernelGGL(k_survey_w3y, dim3(nb), dim3(256), 0, ctx->stream, p1, w1, n1, pR, wR, n2, L,
                                   win ? 2 : 3, (double*)yp.p);
            });
            std::vector<double> h(4 * (size_t)nm * nb);
            if (!rc && (hipMemcpyAsync(h.data(), yp.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream) !=
                            hipSuccess ||
                        hipStreamSynchronize(ctx->stream) != hipSuccess))
                rc = -1;
            dev_free(yp);
            if (rc) return -1;
            snLM.assign(2 * (size_t)nm, 0.0);
            for (int M = 0; M < nm; M++) {
                double dr = 0.0, di = 0.0, rr = 0.0, ri = 0.0;
                for (unsigned b = 0; b < nb; b++) {
                    const double* q = &h[((size_t)b * nm + M) * 4];
                    dr += q[0];
                    di += q[1];
                    rr += q[2];
                    ri += q[3];
                }
                snLM[2 * M] = dr - a3 * rr;
                snLM[2 * M + 1] = di - a3 * ri;
            }
            bw.L = L;
            bw.losD = p1;
            bw.losR = win ? nullptr : pR;
            bw.snLM = snLM.data();
        }
        hm_box_result res;
        std::memset(&res, 0, sizeof(res));
        if (p->measure == HM_SURVEY_BK) {
            res.bk = pk_out;
            res.shell_nmode = nmode_out;
        } else {
            res.zeta = pk_out;
        }
        const int rc = win ? plan_execute_weighted(ctx, p->box, pos + 3 * nD, nR, nullptr, 0, &bw, &res)
                           : plan_execute_weighted(ctx, p->box, pos, nD, pos + 3 * nD, nR, &bw, &res);
        if (rc) return -1;
        if (info) {
            info->wsum_d = swD;
            info->wsum_r = swR;
            info->alpha = alpha;
            info->vsurvey = vs;
            info->nbar = nbar;
            info->norm = bw.norm;
            for (int ax = 0; ax < 3; ax++) info->shift[ax] = dP[ax];
        }
        return 0;
    }
    // 3. Y_lm-weighted fluctuations and their transforms.  |Y_red| <= 1, so every
    // particle's |re| + |im| is at most sqrt(2) |w|: the bound fixes the fixed-point scale.
    // The window function (field.hpp:230-247) uses the random catalogue alone.
    const bool window = p->measure == HM_SURVEY_XI_WINDOW;
    const double wabs = window ? 1.5 * awR : 1.5 * (awD + std::fabs(alpha) * awR);
    C<T>* fields = (C<T>*)p->fields.p;
    const int n3[3] = {m.n[0], m.n[1], m.n[2]};
    for (int f = 0; f < p->nf; f++) {
        const int l = f == 0 ? 0 : p->l, M = f == 0 ? 0 : f - 1;
        // the first field carries the survey volume's grid in its imaginary channel (mode 3)
        const bool with_volume = fold_volume && f == 0 && !window;
        if (ctx->launch("survey_weights", [&] {
                if (window)
                    hipLaunchKernelGGL(k_survey_weights, dim3((unsigned)((nR + 255) / 256)), dim3(256), 0, ctx->stream,
                                       pD, wD, nD, pR, wR, nR, alpha, l, M, 2, wts + 2 * nD);
                else
                    hipLaunchKernelGGL(k_survey_weights, dim3(pb), dim3(256), 0, ctx->stream, pD, wD, nD, pR, wR, nR,
                                       alpha, l, M, with_volume ? 3 : 0, wts);
            }))
            return -1;
        const double scale = fixed_scale_for(with_volume ? wabs + awR : wabs);
        const bool cplx = M != 0;
        // the positions are the same for every field: sort them once, then move only the weights
        const int arc = window ? assign_sorted(ctx, ag, pos + 3 * nD, wts + 2 * nD, nR, scale, fre, cplx ? fim : nullptr,
                                               p->sort, f > 0)
                               : assign_sorted(ctx, ag, pos, wts, N, scale, fre, (cplx || with_volume) ? fim : nullptr,
                                               p->sort, f > 0);
        if (arc) return -1;
        if (fold_volume && f == 0 && volume_from(window ? fre : fim, scale)) return -1;
        C<T>* out = fields + (long long)f * ng;
        if (p->measure == HM_SURVEY_PK) {
            // forward transform (field.hpp:310-326) with the fixed-point cells converted on load; P_l reads only
            // the modes of the binning support (and their mirrors, in the same cube), so each pass stores only
            // those: the y and x passes run on the support's columns alone
            FixedCplxSrc<T> fs{fre, cplx ? fim : nullptr, m.n[2], 1.0 / scale, 1.0 / dV, dV};
            RowGeom rz{(long long)m.n[0] * m.n[1], m.n[2], m.n[2], -1};
            rz.out_s = p->s[2];
            if (launch_rows_f<T, -1>(ctx, "survey_fft_z", m.n[2], fs, out, rz)) return -1;
            const int cz = support_count(p->s[2], m.n[2]), cy = support_count(p->s[1], m.n[1]);
            ColGeom gy;
            gy.OS = (long long)m.n[1] * m.n[2];
            gy.PS = m.n[2];
            gy.n_outer = m.n[0];
            gy.outer_full = m.n[0];
            gy.outer_s = -1;
            gy.n_inner = cz;
            gy.inner_full = m.n[2];
            gy.inner_s = p->s[2];
            gy.point_s = -1;
            gy.out_s = p->s[1];
            if (fft_cols<T>(ctx, "survey_fft_y", m.n[1], -1, out, out, gy)) return -1;
            ColGeom gx;
            gx.OS = m.n[2];
            gx.PS = (long long)m.n[1] * m.n[2];
            gx.n_outer = cy;
            gx.outer_full = m.n[1];
            gx.outer_s = p->s[1];
            gx.n_inner = cz;
            gx.inner_full = m.n[2];
            gx.inner_s = p->s[2];
            gx.point_s = -1;
            gx.out_s = p->s[0];
            if (fft_cols<T>(ctx, "survey_fft_x", m.n[0], -1, out, out, gx)) return -1;
        } else {
            if (ctx->launch("survey_to_complex", [&] {
                    hipLaunchKernelGGL(k_fixed_to_complex<T>, dim3(rbg), dim3(256), 0, ctx->stream, fre,
                                       cplx ? fim : nullptr, ng, 1.0 / scale, 1.0 / dV, dV, out);
                }))
                return -1;
            if (fft3d_c2c<T>(ctx, out, n3, -1)) return -1;
        }
    }
    // 4. shot noise sums for M = 0..l; for l = 0 (Y_00 = 1) they are the sums of w^2 already formed in step 1
    std::vector<double> snr(nsn, 0.0);
    if (p->l == 0) {
        snr[0] = sw2D;
        snr[2] = sw2R;
    } else {
        if (ctx->launch("survey_sn", [&] {
                hipLaunchKernelGGL(k_survey_sn, dim3(rb), dim3(256), 0, ctx->stream, pD, wD, nD, pR, wR, nR, p->l,
                                   part);
            }))
            return -1;
        if (reduce_cols<double>(ctx, "survey_reduce", part, rb, nsn, red)) return -1;
        HM_HIP_CHECK(ctx, hipMemcpyAsync(snr.data(), red, sizeof(double) * nsn, hipMemcpyDeviceToHost, ctx->stream));
        HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    // shot noise per M: sum_D + alpha^2 sum_R (field.hpp:1187), or sum_R for the window (field.hpp:1202-1218)
    std::vector<double> snM(2 * (size_t)(p->l + 1));
    for (int M = 0; M <= p->l; M++) {
        snM[2 * M] = window ? snr[4 * M + 2] : snr[4 * M] + pow(alpha, 2) * snr[4 * M + 2];
        snM[2 * M + 1] = window ? snr[4 * M + 3] : snr[4 * M + 1] + pow(alpha, 2) * snr[4 * M + 3];
    }
    const C<T>* fl = p->l == 0 ? fields : fields + ng;
    const double* base = (const double*)p->tab.p;
    const long long ntot = (long long)m.n[0] + m.n[1] + m.n[2];
    AxisTabs tabs;
    for (int ax = 0; ax < 3; ax++) {
        tabs.kv[ax] = base + p->tab_off[ax];
        tabs.wf[ax] = base + ntot + p->tab_off[ax];
        tabs.sf[ax] = base + 2 * ntot + p->tab_off[ax];
        tabs.iw[ax] = base + 3 * ntot + p->tab_off[ax];
    }
    const int nbin = p->measure == HM_SURVEY_PK ? p->nk : p->nr;
    const long long nrow_bins = p->measure == HM_SURVEY_PK ? p->nrows : p->nrows_r;
    if (fill_zero(ctx, p->part.p, sizeof(double) * 2 * (size_t)nbin * nrow_bins)) return -1;
    if (fill_zero(ctx, p->cnt.p, sizeof(long long) * (size_t)nbin * nrow_bins)) return -1;
    if (p->measure == HM_SURVEY_PK) {
        // 5. binning
        SurveyBinArgs a;
        for (int ax = 0; ax < 3; ax++) {
            a.n[ax] = m.n[ax];
            a.s[ax] = p->s[ax];
            a.c[ax] = support_count(p->s[ax], m.n[ax]);
        }
        a.nrows = p->nrows;
        a.ng = ng;
        a.t = tabs;
        a.wp = wpow_of(m.assign);
        a.f2b = (const int2*)p->f2b.p;
        a.nk = p->nk;
        a.l = p->l;
        for (int i = 0; i < 2 * SV_MAX_ELL + 1; i++) a.w[i] = i < 2 * p->l + 1 ? p->wM[i] : 0.0;
        for (int i = 0; i < 2 * (p->l + 1); i++) a.sn[i] = snM[i];
        if (ctx->launch("bin_survey", [&] {
                hipLaunchKernelGGL(k_bin_survey<T>, dim3((unsigned)((p->nrows + 255) / 256)), dim3(256), 0,
                                   ctx->stream, fields, fl, a, (double*)p->part.p, (long long*)p->cnt.p);
            }))
            return -1;
    } else {
        // 5'. per M: g_M, its inverse transform on the separations that reach a bin, binning
        C<T>* g = (C<T>*)p->gwork.p;
        const int cz = support_count(p->rs[2], m.n[2]), cy = support_count(p->rs[1], m.n[1]);
        SurveyXiArgs xa;
        for (int ax = 0; ax < 3; ax++) {
            xa.n[ax] = m.n[ax];
            xa.s[ax] = p->rs[ax];
            xa.c[ax] = support_count(p->rs[ax], m.n[ax]);
            xa.dr[ax] = m.L[ax] / (double)m.n[ax];
        }
        xa.nrows = p->nrows_r;
        xa.f2b = (const int2*)p->f2b_r.p;
        xa.nr = p->nr;
        xa.l = p->l;
        bool first = true;
        for (int M = -p->l; M <= p->l; M++) {
            const double wM = p->wM[M + p->l];
            if (wM == 0.0) continue;
            const int am = M < 0 ? -M : M;
            const double sg = (M < 0 && (am & 1)) ? -1.0 : 1.0;
            const double snre = M >= 0 ? snM[2 * am] : sg * snM[2 * am];
            const double snim = M >= 0 ? snM[2 * am + 1] : -(sg * snM[2 * am + 1]);
            if (ctx->launch("survey_g", [&] {
                    hipLaunchKernelGGL(k_survey_g<T>, dim3(rbg), dim3(256), 0, ctx->stream, fields, fl, m.n[0], m.n[1],
                                       m.n[2], tabs, wpow_of(m.assign), M, snre, snim, 1.0 / m.volume(), g);
                }))
                return -1;
            RowGeom rg{(long long)m.n[0] * m.n[1], m.n[2], m.n[2], -1};
            if (fft_rows<T>(ctx, "survey_xi_z", m.n[2], +1, g, g, rg)) return -1;
            ColGeom gy;
            gy.OS = (long long)m.n[1] * m.n[2];
            gy.PS = m.n[2];
            gy.n_outer = m.n[0];
            gy.outer_full = m.n[0];
            gy.outer_s = -1;
            gy.n_inner = cz;
            gy.inner_full = m.n[2];
            gy.inner_s = p->rs[2];
            gy.point_s = -1;
            if (fft_cols<T>(ctx, "survey_xi_y", m.n[1], +1, g, g, gy)) return -1;
            ColGeom gx;
            gx.OS = m.n[2];
            gx.PS = (long long)m.n[1] * m.n[2];
            gx.n_outer = cy;
            gx.outer_full = m.n[1];
            gx.outer_s = p->rs[1];
            gx.n_inner = cz;
            gx.inner_full = m.n[2];
            gx.inner_s = p->rs[2];
            gx.point_s = -1;
            if (fft_cols<T>(ctx, "survey_xi_x", m.n[0], +1, g, g, gx)) return -1;
            xa.m1 = -M;
            xa.w = wM;
            xa.count = first ? 1 : 0;
            first = false;
            if (ctx->launch("bin_survey_xi", [&] {
                    hipLaunchKernelGGL(k_bin_survey_xi<T>, dim3((unsigned)((p->nrows_r + 255) / 256)), dim3(256), 0,
                                       ctx->stream, g, xa, (double*)p->part.p, (long long*)p->cnt.p);
                }))
                return -1;
        }
    }
    if (reduce_cols<double>(ctx, "bin_reduce", (const double*)p->part.p, nrow_bins, 2 * nbin, (double*)p->red.p))
        return -1;
    if (reduce_cols<long long>(ctx, "bin_reduce", (const long long*)p->cnt.p, nrow_bins, nbin, (long long*)p->cnt_red.p))
        return -1;
    std::vector<double> pk(2 * (size_t)nbin);
    std::vector<long long> cnt(nbin);
    HM_HIP_CHECK(ctx, hipMemcpyAsync(pk.data(), p->red.p, sizeof(double) * pk.size(), hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipMemcpyAsync(cnt.data(), p->cnt_red.p, sizeof(long long) * nbin, hipMemcpyDeviceToHost,
                                     ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    // particle.hpp:302-313 on P_D, or on P_R for the window (powerspec.hpp:249)
    const double norm = window ? vs / swR / swR : vs / swD / swD;
    if (pk_out)
        for (int b = 0; b < nbin; b++) {
            pk_out[2 * b] = cnt[b] ? norm * (pk[2 * b] / (double)cnt[b]) : 0.0;
            pk_out[2 * b + 1] = cnt[b] ? norm * (pk[2 * b + 1] / (double)cnt[b]) : 0.0;
        }
    if (nmode_out)
        for (int b = 0; b < nbin; b++) nmode_out[b] = cnt[b];
    if (info) {
        info->wsum_d = swD;
        info->wsum_r = swR;
        info->alpha = alpha;
        info->vsurvey = vs;
        info->nbar = nbar;
        info->norm = norm;
        for (int ax = 0; ax < 3; ax++) info->shift[ax] = dP[ax];
    }
    return 0;
}

int survey_execute(hm_ctx* ctx, SurveyPlanImpl* p, const double* pD, const double* wD, long long nD, const double* pR,
                   const double* wR, long long nR, hm_survey_info* info, double* pk, long long* nmode) {
    return p->cfg.precision == HM_FP64 ? survey_run<double>(ctx, p, pD, wD, nD, pR, wR, nR, info, pk, nmode)
                                       : survey_run<float>(ctx, p, pD, wD, nD, pR, wR, nR, info, pk, nmode);
}

}  // namespace hm
