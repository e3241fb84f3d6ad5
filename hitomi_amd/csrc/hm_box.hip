// hm_box.hip -- periodic-box P_l and B_000 on one MI355X.
//
// Pipeline for one mock (reference: calcPowerSpectrumForBOX,
// powerspec.hpp:273-332, and calcBiSpectrumForBOX, bispec.hpp:877-1203,
// for every ith_kbin at once):
//   1. fixed-point TSC/CIC/NGP assignment into an int64 mesh;
//   2. r2c 3D FFT whose first pass converts the fixed-point mesh to
//      (rho - nbar) dV on the fly;
//   3. P_l binning of the half spectrum over the low-k cube only, with the
//      reference's fine (1e-4) histogram and coarse fold, FP64 sums;
//   4. shot noise of B: xi_sn = IFFT[(|d|^2 - N S)/W^2 / V] (even part, the
//      odd part sums to zero against the isotropic Bessel kernel), grouped
//      by exact |n|^2, then sum_n2 j0(k1 r) j0(k2 r) H(n2);
//   5. the 20 shell fields and delta/W inverse-transformed on the smallest
//      mesh M >= 4 s + 1 that holds the triple product without aliasing
//      (s = largest |n_a| in any shell), with pruned passes;
//   6. all 210 triple sums sum_x F_a F_b F_3 in one pass (FP32 products,
//      FP64 accumulation per 64 voxels, fixed-order block reduction).
#include "hm_assign.hpp"
#include "hm_bessel.hpp"
#include "hm_bins.hpp"
#include "hm_fft_launch.hpp"
#include "hm_recon.hpp"

#include "../../include/hitomi_amd.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace hm {

// ---------------------------------------------------------------------------
// P_l binning (calcPowerSpectrum, field.hpp:765-852) on the half spectrum
// ---------------------------------------------------------------------------
constexpr int BIN_MAX_ELLS = 8;  // user multipoles plus those the bispectrum shot noise needs

struct BinArgs {
    int n[3], nzh;
    int s01;          // support on x, y (-1: all)
    int nrow_y;       // compact count along y
    long long nrows;
    int nzmax;        // last kz index visited
    AxisTabs t;
    int wp;           // window power
    const int2* f2b;  // fine index -> [lo, hi] coarse bins
    int nk, ne;
    int ells[BIN_MAX_ELLS];
    double sn;
    int sn_w2 = 0;  // 1: shot-noise shape W(k)^2 (interlaced field) instead of the aliased S(k)
    const double* sn_ptr = nullptr;  // set: the shot noise read from device memory instead of sn
    int kc = 1;                      // kz chunks per row, each with its own partial sums (kcl kz each)
    int kcl = 0;
    long long cm = 0;  // > 0: partials column-major, column c of lane t at c * cm + t (else t * columns + c)
};

template <typename T>
// dk2 set: the cross spectrum dk conj(dk2) (complex; the mirrored mode takes
// its conjugate) with the real parts into part and the imaginary ones into
// part_im.
__global__ __launch_bounds__(256) void k_bin_box(const C<T>* __restrict__ dk, BinArgs a, double* __restrict__ part,
                                                long long* __restrict__ cnt, const C<T>* __restrict__ dk2 = nullptr,
                                                double* __restrict__ part_im = nullptr) {
#pragma clang fp contract(off)
    // one lane per (row, kz chunk); partial sums at [row * kc + chunk]
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.nrows * a.kc) return;
    const long long row = tid / a.kc;
    const int ch = (int)(tid % a.kc);
    const int kcl = a.kc > 1 ? a.kcl : a.nzmax + 1;
    const int kz0 = ch * kcl, kz1 = (kz0 + kcl - 1 < a.nzmax) ? kz0 + kcl - 1 : a.nzmax;
    const double sn = a.sn_ptr ? *a.sn_ptr : a.sn;
    const int i = support_map((int)(row / a.nrow_y), a.s01, a.n[0]);
    const int j = support_map((int)(row % a.nrow_y), a.s01, a.n[1]);
    // element c of this lane's partial sums: strided by cm (column-major) or contiguous
    const long long ps = a.cm ? a.cm : 1;
    const long long pb = a.cm ? tid : tid * (long long)a.nk * a.ne, cb = a.cm ? tid : tid * (long long)a.nk;
    double* P = part + pb;
    double* Pi = part_im ? part_im + pb : nullptr;
    long long* Cn = cnt + cb;
    const long long rbase = ((long long)i * a.n[1] + j) * a.nzh;
    const int im = (a.n[0] - i) % a.n[0], jm = (a.n[1] - j) % a.n[1];
    // the row's x and y table values (the loop's stores to the partial sums could alias the tables, so the
    // compiler would reload them for every kz)
    const double kxv[2] = {a.t.kv[0][i], a.t.kv[0][im]}, kyv[2] = {a.t.kv[1][j], a.t.kv[1][jm]};
    const double wxv[2] = {a.t.wf[0][i], a.t.wf[0][im]}, wyv[2] = {a.t.wf[1][j], a.t.wf[1][jm]};
    const double sxv[2] = {a.t.sf[0][i], a.t.sf[0][im]}, syv[2] = {a.t.sf[1][j], a.t.sf[1][jm]};
    // the sums of the bin the last modes fell in, held in registers while consecutive modes stay in that
    // one bin (read from memory when the bin is entered, written back when it is left: every partial sum
    // sees the same additions in the same order as with a read-modify-write per mode)
    int cb_ = -1;
    double cre[BIN_MAX_ELLS], cim[BIN_MAX_ELLS];
    long long ccn = 0;
    auto flush = [&]() {
        if (cb_ < 0) return;
#pragma unroll
        for (int e = 0; e < BIN_MAX_ELLS; e++) {
            if (e >= a.ne) break;
            P[(cb_ * a.ne + e) * ps] = cre[e];
            if (Pi) Pi[(cb_ * a.ne + e) * ps] = cim[e];
        }
        Cn[cb_ * ps] = ccn;
        cb_ = -1;
    };
    for (int kz = kz0; kz <= kz1; kz++) {
        const C<T> d = dk[rbase + kz];
        const double re = (double)d.x, imv = (double)d.y;
        double p2 = re * re + imv * imv, pim = 0.0;  // f1 conj(f2), f1 = f2
        if (dk2) {
            const C<T> e = dk2[rbase + kz];
            p2 = re * (double)e.x + imv * (double)e.y;
            pim = imv * (double)e.x - re * (double)e.y;
        }
        const int nmir = (kz > 0 && kz < a.n[2] - kz) ? 2 : 1;
        for (int m = 0; m < nmir; m++) {
            const int kk = m ? (a.n[2] - kz) : kz;
            const double kx = kxv[m], ky = kyv[m], kzv = a.t.kv[2][kk];
            const double kmag = kmag3(kx, ky, kzv);
            const int kindex = (int)(kmag / FINE_DK + 0.5);
            if (kindex >= FINE_N) continue;
            const int2 b = a.f2b[kindex];
            if (b.x < 0) continue;
            const double W = wpow((wxv[m] * wyv[m]) * a.t.wf[2][kk], a.wp);
            const double S = a.sn_w2 ? W * W : (sxv[m] * syv[m]) * a.t.sf[2][kk];
            const double val = (p2 - sn * S) / (W * W);
            const double vim = (m ? -pim : pim) / (W * W);
            const double mu = (fabs(kmag) < 1.0e-15) ? 0.0 : kzv / kmag;
            if (b.x == b.y) {
                if (b.x != cb_) {
                    flush();
                    cb_ = b.x;
#pragma unroll
                    for (int e = 0; e < BIN_MAX_ELLS; e++) {
                        if (e >= a.ne) break;
                        cre[e] = P[(cb_ * a.ne + e) * ps];
                        cim[e] = Pi ? Pi[(cb_ * a.ne + e) * ps] : 0.0;
                    }
                    ccn = Cn[cb_ * ps];
                }
#pragma unroll
                for (int e = 0; e < BIN_MAX_ELLS; e++) {
                    if (e >= a.ne) break;
                    const int l = a.ells[e];
                    const double y = (l == 0) ? 1.0 : ((fabs(kmag) < 1.0e-15) ? 0.0 : legendre(l, mu));
                    cre[e] += val * y;
                    if (Pi) cim[e] += vim * y;
                }
                ccn += 1;
                continue;
            }
            flush();
            for (int bb = b.x; bb <= b.y; bb++) {
                for (int e = 0; e < a.ne; e++) {
                    const int l = a.ells[e];
                    const double y = (l == 0) ? 1.0 : ((fabs(kmag) < 1.0e-15) ? 0.0 : legendre(l, mu));
                    P[(bb * a.ne + e) * ps] += val * y;
                    if (Pi) Pi[(bb * a.ne + e) * ps] += vim * y;
                }
                Cn[bb * ps] += 1;
            }
        }
    }
    flush();
}

// ---------------------------------------------------------------------------
// shell mode counts n_mode (field.hpp:354-384) on the full N mesh, exact
// ---------------------------------------------------------------------------
struct CountArgs {
    int n[3];
    int s[3];
    int c[3];  // compact extents
    long long total;
    const double* kv[3];
    const double* lower;
    const double* upper;
    int nk;
};

__global__ __launch_bounds__(256) void k_shell_count(CountArgs a, unsigned long long* __restrict__ out) {
#pragma clang fp contract(off)
    __shared__ unsigned int sc[64];
    if (threadIdx.x < 64) sc[threadIdx.x] = 0;
    __syncthreads();
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.total) {
        const int kc = (int)(t % a.c[2]);
        const long long r = t / a.c[2];
        const int jc = (int)(r % a.c[1]);
        const int ic = (int)(r / a.c[1]);
        const int i = support_map(ic, a.s[0], a.n[0]);
        const int j = support_map(jc, a.s[1], a.n[1]);
        const int k = support_map(kc, a.s[2], a.n[2]);
        const double kmag = kmag3(a.kv[0][i], a.kv[1][j], a.kv[2][k]);
        for (int b = 0; b < a.nk; b++)
            if (kmag > a.lower[b] && kmag <= a.upper[b]) atomicAdd(&sc[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < a.nk && sc[threadIdx.x]) atomicAdd(&out[threadIdx.x], (unsigned long long)sc[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// spectrum source of the first (x) pass of a triple-product field on the
// M mesh: value = [shell] * delta(k) / W(k) (field.hpp:371-382) read from the
// N-mesh half spectrum, W with the reference's truncated index (Q2).
// ---------------------------------------------------------------------------
template <typename T> struct ShellSrc {
    const C<T>* dk;
    int n[3], nzh;
    int M[3];
    AxisTabs t;
    int wp;
    int shell;  // 1: keep modes with lower < |k| <= upper; 0: all modes (delta/W)
    double lower, upper;
    __device__ inline int to_n(int m, int a) const {
        if (M[a] == n[a]) return m;
        const int sgn = (m <= M[a] / 2) ? m : m - M[a];
        return sgn >= 0 ? sgn : sgn + n[a];
    }
    // o = y, b = z, p = x on the M mesh
    __device__ inline C<T> operator()(int, int o, int b, int p, long long) const {
#pragma clang fp contract(off)
        const int i = to_n(p, 0), j = to_n(o, 1), k = to_n(b, 2);
        if (shell) {
            const double kmag = kmag3(t.kv[0][i], t.kv[1][j], t.kv[2][k]);
            if (!(kmag > lower && kmag <= upper)) return mk<T>((T)0, (T)0);
        }
        C<T> d;
        if (k < nzh) {
            d = dk[((long long)i * n[1] + j) * nzh + k];
        } else {
            const int im = (n[0] - i) % n[0], jm = (n[1] - j) % n[1];
            d = cconj<T>(dk[((long long)im * n[1] + jm) * nzh + (n[2] - k)]);
        }
        const double W = wpow((t.wf[0][i] * t.wf[1][j]) * t.wf[2][k], wp);
        return mk<T>((T)((double)d.x / W), (T)((double)d.y / W));
    }
};

// ---------------------------------------------------------------------------
// shot-noise field source: even part of g(k) = (Re d conj(n) - N S)/W^2 / V
// (calcShotNoiseForBispectrum_ijk, field.hpp:1116-1151), n(k) = d(k) + N [k=0]
// ---------------------------------------------------------------------------
template <typename T> struct XiSrc {
    const C<T>* dk;
    int n[3], nzh;
    AxisTabs t;
    int wp;
    double sn, inv_v;
    double dc;  // added to Re d at k = 0 in Re d conj(n): N for the shot-noise field, 0 for xi
    int odd = 0;  // 1: the spectrum whose transform is Im xi instead of Re xi
    const C<T>* dk2 = nullptr;  // set: g from the cross spectrum d conj(d2) (no dc term)
    int sn_w2 = 0;              // 1: shot-noise shape W(k)^2 (interlaced field)
    const double* sn_ptr = nullptr;  // set: sn and dc read from device memory
    const double* dc_ptr = nullptr;
    // per x index i, {1/W_x^2(i), 1/W_x^2(-i), S_x(i)/W_x^2(i), S_x(-i)/W_x^2(-i)} (the x factors
    // of both modes of the pair (k, -k), one 32-byte load; required)
    const double* xq = nullptr;
    // x pass over the half spectrum: o = 0, b = y * nzh + kz, p = x
    __device__ inline C<T> operator()(int, int, int b, int p, long long addr) const {
        return at(p, b / nzh, b % nzh, addr);
    }
    // the mode (i, j, k) has the fast form: 0 < k_z < n_z/2, no second spectrum, the aliased shape
    __device__ inline bool fast_ok(int k) const { return !dk2 && !sn_w2 && k != 0 && 2 * k != n[2]; }
    // inside 0 < k_z < n_z/2 the spectrum at -k is the conjugate, so with p2 = |d|^2 the two sides
    // combine to Re = (p2 (Q1 + Q2) - sn (S1 Q1 + S2 Q2)) / 2V (even part) and
    // Im = -(p2 (Q1 - Q2) - sn (S1 Q1 - S2 Q2)) / 2V (odd part); Q, S separable per axis
    // the y, z factors of the fast form (the same for a whole x line) and the x table entry of a point
    struct FastYZ {
        double qa, qb, sa, sb, snv;
    };
    __device__ inline FastYZ fast_yz(int j, int k) const {
#pragma clang fp contract(off)
        const int jj = j ? n[1] - j : 0, kk = n[2] - k;
        const double qa = t.iw[1][j] * t.iw[2][k], qb = t.iw[1][jj] * t.iw[2][kk];
        const double sa = (t.sf[1][j] * t.sf[2][k]) * qa, sb = (t.sf[1][jj] * t.sf[2][kk]) * qb;
        return FastYZ{qa, qb, sa, sb, sn_ptr ? *sn_ptr : sn};
    }
    __device__ inline double4 fast_x(int i) const { return *reinterpret_cast<const double4*>(xq + 4 * (long long)i); }
    __device__ inline C<T> fast_q(const double4& q, const FastYZ& f, C<T> d) const {
#pragma clang fp contract(off)
        const double re = (double)d.x, im = (double)d.y;
        const double p2 = re * re + im * im;
        const double Q1 = q.x * f.qa, Q2 = q.y * f.qb, SQ1 = q.z * f.sa, SQ2 = q.w * f.sb;
        if (odd) return mk<T>((T)0, (T)(-(0.5 * inv_v) * (p2 * (Q1 - Q2) - f.snv * (SQ1 - SQ2))));
        return mk<T>((T)((0.5 * inv_v) * (p2 * (Q1 + Q2) - f.snv * (SQ1 + SQ2))), (T)0);
    }
    __device__ inline C<T> fast(int i, int j, int k, C<T> d) const { return fast_q(fast_x(i), fast_yz(j, k), d); }
    // the value at mode (i, j, k) stored at addr
    __device__ inline C<T> at(int i, int j, int k, long long addr) const {
#pragma clang fp contract(off)
        const C<T> d = dk[addr];
        if (fast_ok(k)) return fast(i, j, k, d);
        const double re = (double)d.x, im = (double)d.y;
        double p2, pim = 0.0;
        if (dk2) {
            const C<T> e = dk2[addr];
            p2 = re * (double)e.x + im * (double)e.y;
            pim = im * (double)e.x - re * (double)e.y;
        } else if (i == 0 && j == 0 && k == 0) {
            p2 = re * (re + (dc_ptr ? *dc_ptr : dc)) + im * im;
        } else {
            p2 = re * re + im * im;
        }
        // S and 1/W^2 of k and of -k from per-axis factors (the y, z ones are
        // the same along a column and leave the loop)
        const int ii = i ? n[0] - i : 0, jj = j ? n[1] - j : 0, kk = k ? n[2] - k : 0;
        // the spectrum at -k: the conjugate for k_z inside (0, n_z/2); on the
        // self-conjugate planes k_z = 0, n_z/2 the stored mirror element (an
        // interlaced field need not be Hermitian there)
        double p2m = p2, pimm = -pim;
        if ((k == 0 || 2 * k == n[2]) && !(ii == i && jj == j)) {
            const long long am = ((long long)ii * n[1] + jj) * nzh + k;
            const C<T> dm = dk[am];
            const double mr = (double)dm.x, mi = (double)dm.y;
            if (dk2) {
                const C<T> em = dk2[am];
                p2m = mr * (double)em.x + mi * (double)em.y;
                pimm = mi * (double)em.x - mr * (double)em.y;
            } else {
                p2m = mr * mr + mi * mi;
                pimm = 0.0;
            }
        }
        const double Q1 = t.iw[0][i] * (t.iw[1][j] * t.iw[2][k]);
        const double Q2 = t.iw[0][ii] * (t.iw[1][jj] * t.iw[2][kk]);
        const double S1 = sn_w2 ? 1.0 / Q1 : t.sf[0][i] * (t.sf[1][j] * t.sf[2][k]);
        const double S2 = sn_w2 ? 1.0 / Q2 : t.sf[0][ii] * (t.sf[1][jj] * t.sf[2][kk]);
        // g(k) = a + i b at k and at -k (whose spectrum is the conjugate):
        // Re xi transforms (g(k) + conj g(-k)) / 2, Im xi (g(k) - conj g(-k)) / 2i
        const double snv = sn_ptr ? *sn_ptr : sn;
        const double g1 = inv_v * ((p2 - snv * S1) * Q1);
        const double g2 = inv_v * ((p2m - snv * S2) * Q2);
        const double b1 = inv_v * (pim * Q1), b2 = inv_v * (pimm * Q2);
        if (odd) return mk<T>((T)(0.5 * (b1 + b2)), (T)(-(0.5 * (g1 - g2))));
        return mk<T>((T)(0.5 * (g1 + g2)), (T)(0.5 * (b1 - b2)));
    }
};

// XiSrc on the planes k_z = 0 and n_z/2 of the half spectrum in one column
// pass: batch f is the plane k_z = f * kzs (batch stride kzs), the y index the
// inner line index (stride nzh), so a tile holds 8 or 16 y columns of a plane.
template <typename T> struct XiPlaneSrc {
    XiSrc<T> xs;
    int kzs;
    __device__ inline C<T> operator()(int f, int, int b, int p, long long addr) const { return xs.at(p, b, f * kzs, addr); }
};

// The forward x pass of the spectrum fused with the inverse x pass of the
// shot-noise xi (columns b = y * nzh + k_z of the half spectrum, points x
// PS apart): each tile is transformed forward and stored (the spectrum d),
// then the xi spectrum of its points is formed in LDS from d and transformed
// back into xh -- d is read from HBM once for both.  Columns on the planes
// k_z = 0, n_z/2 need the mirror points of other columns; they are left to a
// pass with XiPlaneSrc afterwards.
template <typename T, int L>
__global__ __launch_bounds__(FFT_NT) void k_fwd_x_xi(C<T>* __restrict__ dk, C<T>* __restrict__ xh, XiSrc<T> xs,
                                                    long long PS, int n_inner, const C<T>* __restrict__ tw) {
    constexpr int B = ColTile<T, L>::B;
    constexpr int LP = LdsPad<L>::LP;
    constexpr int PSTEP = FFT_NT / B;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const uint3 wg = xcd_block();
    const int tid = threadIdx.x, bl = tid % B;
    const int bc = (int)wg.x * B + bl;
    const bool bvalid = bc < n_inner;
    const int b = bvalid ? bc : 0;
    // all of the lane's points are loaded before any is written to LDS (loads in flight together)
    constexpr int NI = (L + PSTEP - 1) / PSTEP;
    const int p0 = tid / B;
    // element addresses of the lane's points: a0 + i * ast (the uniform step folds into scalar arithmetic)
    const long long a0 = b + (long long)p0 * PS, ast = (long long)PSTEP * PS;
    C<T>* stw = sm + B * LP;  // the twiddle table, staged once for both transforms
    {
        C<T> buf[NI];
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int p = p0 + i * PSTEP;
            buf[i] = (p < L && bvalid) ? dk[a0 + i * ast] : mk<T>((T)0, (T)0);
        }
        for (int e = tid; e < L; e += FFT_NT) stw[e] = tw[e];
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int p = p0 + i * PSTEP;
            if (p < L) sm[p * B + bl] = buf[i];
        }
    }
    __syncthreads();
    lds_fft_cm<T, -1, L, B>(sm, stw);
    const int j = b / xs.nzh, k = b % xs.nzh;
    const bool form = bvalid && xs.fast_ok(k);
    const typename XiSrc<T>::FastYZ yz = form ? xs.fast_yz(j, k) : typename XiSrc<T>::FastYZ{0.0, 0.0, 0.0, 0.0, 0.0};
    // each lane reads and rewrites only the points it stores; two points at a time: their LDS values and
    // x table entries are read before any store
    constexpr int CH = 2;
#pragma unroll
    for (int i0 = 0; i0 < NI; i0 += CH) {
        C<T> dv[CH];
        double4 qv[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int p = p0 + (i0 + u) * PSTEP;
            const bool ok = i0 + u < NI && p < L;
            dv[u] = ok ? sm[p * B + bl] : mk<T>((T)0, (T)0);
            qv[u] = (ok && form) ? xs.fast_x(p) : make_double4(0.0, 0.0, 0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int p = p0 + (i0 + u) * PSTEP;
            if (i0 + u >= NI || p >= L) continue;
            if (bvalid) dk[a0 + (i0 + u) * ast] = dv[u];
            sm[p * B + bl] = form ? xs.fast_q(qv[u], yz, dv[u]) : mk<T>((T)0, (T)0);
        }
    }
    __syncthreads();
    lds_fft_cm<T, +1, L, B>(sm, stw);
    if (bvalid) {
        C<T> out[NI];
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int p = p0 + i * PSTEP;
            if (p < L) out[i] = sm[p * B + bl];
        }
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int p = p0 + i * PSTEP;
            if (p < L) xh[a0 + i * ast] = out[i];
        }
    }
}

// The fused pass for a compile-time length; 1 when it ran, 0 when L has no
// compile-time plan (the caller runs the two passes separately), -1 on error.
template <typename T>
static int launch_fwd_x_xi(hm_ctx* ctx, int L, C<T>* dk, C<T>* xh, const XiSrc<T>& xs, long long PS, int n_inner) {
    switch (L) {
#define HM_CASE(LL)                                                                                                    \
    case LL: {                                                                                                         \
        const C<T>* tw = ctx->twiddle<T>(LL);                                                                          \
        if (!tw) return -1;                                                                                            \
        constexpr int B = ColTile<T, LL>::B;                                                                           \
        const size_t lds = col_lds<T, LL>();                                                                           \
        const unsigned gx = (unsigned)((n_inner + B - 1) / B);                                                         \
        return ctx->launch("fft_x_xi", [&] {                                                                           \
                   hipLaunchKernelGGL((k_fwd_x_xi<T, LL>), dim3(gx), dim3(FFT_NT), lds, ctx->stream, dk, xh, xs, PS,   \
                                      n_inner, tw);                                                                    \
               })                                                                                                      \
                   ? -1                                                                                                \
                   : 1;                                                                                                \
    }
        HM_POW2_LENGTHS(HM_CASE)
        HM_MIXED_LENGTHS(HM_CASE)
#undef HM_CASE
    default: return 0;
    }
}

// Rows of a half spectrum restricted to |x|, |y| within a support, in the
// compact order of support_map (the z c2r of a pruned inverse transform).
template <typename T> struct SupHalfSrc {
    const C<T>* p;
    int nzh, n0, n1, s0, s1, c1;
    __device__ inline C<T> operator()(long long r, int k) const {
        const int x = support_map((int)(r / c1), s0, n0), y = support_map((int)(r % c1), s1, n1);
        return p[((long long)x * n1 + y) * nzh + k];
    }
};

// ---------------------------------------------------------------------------
// xi_l binning (calcCorrelationFunction, field.hpp:906-945) over the cube of
// separations that can reach an r-bin: fine 0.5 Mpc/h index, coarse fold
// (lower, upper], Y_l0(r) = P_l(r_z / |r|), FP64 sums per compact row.
// ---------------------------------------------------------------------------

struct XiBinArgs {
    int n[3], s[3], c[3];
    double dr[3];
    long long nrows;
    const int2* f2b;
    int nr, ne;
    int ells[HM_MAX_ELLS];
};


template <typename T>
__global__ __launch_bounds__(256) void k_bin_xi(const T* __restrict__ xr, XiBinArgs a, double* __restrict__ part,
                                               long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.nrows) return;
    const int i = support_map((int)(row / a.c[1]), a.s[0], a.n[0]);
    const int j = support_map((int)(row % a.c[1]), a.s[1], a.n[1]);
    const double rx = mesh_vec(i, a.n[0], a.dr[0]), ry = mesh_vec(j, a.n[1], a.dr[1]);
    const T* line = xr + row * (long long)a.n[2];
    double* P = part + row * (long long)a.nr * a.ne;
    long long* Cn = cnt + row * (long long)a.nr;
    for (int kc = 0; kc < a.c[2]; kc++) {
        const int k = support_map(kc, a.s[2], a.n[2]);
        const double rz = mesh_vec(k, a.n[2], a.dr[2]);
        const double rmag = kmag3(rx, ry, rz);
        const int ri = (int)(rmag / FINE_DR + 0.5);
        if (ri >= FINE_NR) continue;
        const int2 b = a.f2b[ri];
        if (b.x < 0) continue;
        const double v = (double)line[k];
        const double mu = (fabs(rmag) < 1.0e-15) ? 0.0 : rz / rmag;
        for (int bb = b.x; bb <= b.y; bb++) {
            for (int e = 0; e < a.ne; e++) {
                const int l = a.ells[e];
                const double y = (l == 0) ? 1.0 : ((fabs(rmag) < 1.0e-15) ? 0.0 : legendre(l, mu));
                P[bb * a.ne + e] += v * y;
            }
            Cn[bb] += 1;
        }
    }
}

// Shot-noise radial sums, step 1 (before the last, z, transform, which is
// linear): G[a][kz] = sum of the half-spectrum rows (i, j, .) whose
// minimum-image i'^2 + j'^2 equals a (rows listed per a in a fixed order),
// one workgroup per distinct a, FP64 sums.  The z transform of G then gives
// R[a][k], the sum of xi over those rows, for ~10x fewer rows than the mesh.
// One workgroup per group; blockDim.x (up to XG_NT) lanes over k_z, launched
// with all nzh = n_z/2 + 1 of them where that fits, so that the Nyquist column
// is summed beside the others rather than in a second pass of one lane.
constexpr int XG_NT = 512;
template <typename T>
__global__ __launch_bounds__(XG_NT) void k_xi_group(const C<T>* __restrict__ xh, int nzh, const int* __restrict__ a_start,
                                                   const int* __restrict__ rows, C<T>* __restrict__ G) {
    const int ai = blockIdx.x;
    const int r0 = a_start[ai], r1 = a_start[ai + 1];
    for (int k = threadIdx.x; k < nzh; k += blockDim.x) {
        double ar = 0.0, aim = 0.0;
        int r = r0;
        // four rows' values loaded before any is added; the sums keep the row order
        for (; r + 4 <= r1; r += 4) {
            C<T> v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = xh[(long long)rows[r + u] * nzh + k];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                ar += (double)v[u].x;
                aim += (double)v[u].y;
            }
        }
        for (; r < r1; r++) {
            const C<T> v = xh[(long long)rows[r] * nzh + k];
            ar += (double)v.x;
            aim += (double)v.y;
        }
        G[(long long)ai * nzh + k] = mk<T>((T)ar, (T)aim);
    }
}

// Keys of the shot-noise radial sums.  Equal cell sizes: key = i'^2 + j'^2 +
// k'^2 (|r| = sqrt(key) dr).  Unequal cell sizes (the reference's anisotropic
// survey boxes): key = (|i'| e1 + |j'|) e2 + |k'|, one key per distinct
// (|i'|, |j'|, |k'|), |r| = sqrt((|i'| d0)^2 + (|j'| d1)^2 + (|k'| d2)^2); the
// rows (i, j) are grouped by a = |i'| e1 + |j'|.  Signed indices follow the
// rvec rule of bispec.hpp:1040-1042 (the Nyquist index negative).
struct SnKeys {
    int aniso = 0;
    int e1 = 0, e2 = 0;  // extents of |j'| and |k'| (+1)
    double d[3] = {0.0, 0.0, 0.0};
};
__device__ __forceinline__ double sn_key_r(const SnKeys& K, long long key, double dr) {
#pragma clang fp contract(off)
    if (!K.aniso) return sqrt((double)key) * dr;
    const long long i = key / ((long long)K.e1 * K.e2), j = (key / K.e2) % K.e1, k = key % K.e2;
    const double x = (double)i * K.d[0], y = (double)j * K.d[1], z = (double)k * K.d[2];
    return sqrt(x * x + y * y + z * z);
}

// Step 2 with unequal cell sizes: H[(a, q)] = R[a][q] + R[a][n2 - q] (one
// term for q = 0 and the Nyquist plane q = n2 / 2).
template <typename T>
__global__ __launch_bounds__(256) void k_xi_hist_aniso(const T* __restrict__ R, int n2, const int* __restrict__ a_idx,
                                                      SnKeys K, long long nkey, double* __restrict__ H) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nkey) return;
    const long long a = n / K.e2;
    const int q = (int)(n % K.e2);
    const int ai = a_idx[a];
    double acc = 0.0;
    if (ai >= 0) {
        const T* row = R + (long long)ai * n2;
        if (q == 0) acc = (double)row[0];
        else if (2 * q == n2) acc = (double)row[q];
        else acc = (double)row[q] + (double)row[n2 - q];
    }
    H[n] = acc;
}

// Step 2: H[n] = sum over k with a = n - k'^2 a row key of R[a][k]: the sum of
// xi over all voxels with minimum-image |n|^2 = n (rvec rule of
// bispec.hpp:1040-1043), in a fixed order.  XH_G lanes share a key: lane l
// sums the terms q = l, l + XH_G, ... and the partials are added in a fixed
// butterfly order (every lane ends with the same total), so a wavefront has
// XH_G times the loads in flight that one lane per key would have (the keys'
// dependent row lookups are latency-bound).
constexpr int XH_G = 8;
template <typename T>
__global__ __launch_bounds__(256) void k_xi_hist(const T* __restrict__ R, int n2, const int* __restrict__ a_idx,
                                                long long amax, long long nkey, double* __restrict__ H) {
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = gt / XH_G;
    const int l = (int)(gt % XH_G);
    double acc = 0.0;
    if (n < nkey) {
        long long m = (long long)sqrt((double)n);
        while (m * m > n) m--;
        while ((m + 1) * (m + 1) <= n) m++;
        if (m > n2 / 2) m = n2 / 2;
        for (long long q = l; q <= m; q += XH_G) {
            const long long a = n - q * q;
            if (a > amax) continue;
            const int ai = a_idx[a];
            if (ai < 0) continue;
            const T* row = R + (long long)ai * n2;
            if (q == 0) acc += (double)row[0];
            else if (q == n2 / 2) acc += (double)row[n2 / 2];
            else acc += (double)row[q] + (double)row[n2 - q];
        }
    }
#pragma unroll
    for (int o = XH_G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, XH_G);
    if (l == 0 && n < nkey) H[n] = acc;
}

// As k_xi_hist for a general multipole: each voxel's xi is weighted by
// A(mu) = sum_m w_m Y(l1, m)(r) Y(l2, -m)(r), which depends on mu = r_z/|r|
// alone (the e^{i m phi} factors cancel) and is a polynomial of degree
// l1 + l2 in mu, evaluated from its Chebyshev coefficients (Clenshaw); A0 is
// its value at r = 0, where the reference's Y vanish for l > 0.
constexpr int ANG_MAX = 17;
struct AngPoly {
    int deg;
    double c[ANG_MAX];
    double a0;
};

__device__ __forceinline__ double ang_eval(const AngPoly& A, double mu) {
    double b1 = 0.0, b2 = 0.0;
    for (int k = A.deg; k >= 1; k--) {
        const double b0 = 2.0 * mu * b1 - b2 + A.c[k];
        b2 = b1;
        b1 = b0;
    }
    return mu * b1 - b2 + A.c[0];
}

template <typename T>
__global__ __launch_bounds__(256) void k_xi_hist_ang(const T* __restrict__ R, int n2, const int* __restrict__ a_idx,
                                                    long long amax, long long nkey, AngPoly A,
                                                    double* __restrict__ H) {
    // XH_G lanes per key as in k_xi_hist (terms q = l, l + XH_G, ...; fixed butterfly sum)
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = gt / XH_G;
    const int l = (int)(gt % XH_G);
    double acc = 0.0;
    if (n < nkey) {
        long long m = (long long)sqrt((double)n);
        while (m * m > n) m--;
        while ((m + 1) * (m + 1) <= n) m++;
        if (m > n2 / 2) m = n2 / 2;
        const double rn = sqrt((double)n);
        for (long long q = l; q <= m; q += XH_G) {
            const long long a = n - q * q;
            if (a > amax) continue;
            const int ai = a_idx[a];
            if (ai < 0) continue;
            const T* row = R + (long long)ai * n2;
            if (n == 0) {
                acc += A.a0 * (double)row[0];
            } else if (q == 0) {
                acc += ang_eval(A, 0.0) * (double)row[0];
            } else if (q == n2 / 2) {  // the Nyquist plane is the negative z (field.hpp:1042)
                acc += ang_eval(A, -(double)q / rn) * (double)row[n2 / 2];
            } else {
                acc += ang_eval(A, (double)q / rn) * (double)row[q] + ang_eval(A, -(double)q / rn) * (double)row[n2 - q];
            }
        }
    }
#pragma unroll
    for (int o = XH_G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, XH_G);
    if (l == 0 && n < nkey) H[n] = acc;
}

// As k_xi_hist_aniso with each voxel's xi weighted by A(mu), mu = r_z / |r|
// (the Nyquist plane negative, field.hpp:1042), A0 at r = 0.
template <typename T>
__global__ __launch_bounds__(256) void k_xi_hist_ang_aniso(const T* __restrict__ R, int n2,
                                                          const int* __restrict__ a_idx, SnKeys K, long long nkey,
                                                          AngPoly A, double* __restrict__ H) {
#pragma clang fp contract(off)
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nkey) return;
    const long long a = n / K.e2;
    const int q = (int)(n % K.e2);
    const int ai = a_idx[a];
    double acc = 0.0;
    if (ai >= 0) {
        const T* row = R + (long long)ai * n2;
        const double rn = sn_key_r(K, n, 0.0);
        const double rz = (double)q * K.d[2];
        if (n == 0) acc = A.a0 * (double)row[0];
        else if (q == 0) acc = ang_eval(A, 0.0) * (double)row[0];
        else if (2 * q == n2) acc = ang_eval(A, -rz / rn) * (double)row[q];
        else acc = ang_eval(A, rz / rn) * (double)row[q] + ang_eval(A, -rz / rn) * (double)row[n2 - q];
    }
    H[n] = acc;
}

// J[a][n2] = j_l(k_a sqrt(n2) dr) from the spline of bessel.hpp:12-25
__global__ void k_jl_table(const double* __restrict__ sy, const double* __restrict__ sc, const double* __restrict__ kb,
                           int nk, long long nkey, double dr, SnKeys K, double* __restrict__ J) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nk * nkey) return;
    const int a = (int)(t / nkey);
    const long long key = t % nkey;
    J[t] = jl_spline_eval(sy, sc, kb[a] * sn_key_r(K, key, dr));
}

// J[a][n2] = j0(k_a sqrt(n2) dr) (bessel.hpp spline of j0 replaced by the
// closed form sin x / x)
__global__ void k_j0_table(const double* __restrict__ kb, int nk, long long nkey, double dr, SnKeys K,
                           double* __restrict__ J) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nk * nkey) return;
    const int a = (int)(t / nkey);
    const long long key = t % nkey;
    const double x = kb[a] * sn_key_r(K, key, dr);
    J[t] = (x == 0.0) ? 1.0 : sin(x) / x;
}

// S[a][b] = dV * sum_n2 J1[a][n2] J2[b][n2] H[n2] for all pairs: each workgroup takes SNP_KPB keys,
// stages J1[a][key] H[key] dV and J2[b][key] for every bin in LDS, SNP_CH keys
// at a time, and every lane accumulates up to SNP_NPT pairs; part[blk][pair]
// is summed over workgroups afterwards.  Reads each table once instead of once
// per pair.
constexpr int SNP_CH = 32;
constexpr int SNP_CHP = SNP_CH + 1;  // row pitch in LDS (bank spread)
constexpr int SNP_KPB = 512;
constexpr int SNP_NPT = 16;
__global__ __launch_bounds__(256) void k_sn_pairs_blk(const double* __restrict__ J1, const double* __restrict__ J2,
                                                     int nk, long long nkey, const double* __restrict__ H,
                                                     const short2* __restrict__ pairs, int npairs, double dV,
                                                     double* __restrict__ part) {
    extern __shared__ double snsm[];
    double* ja = snsm;
    double* jb = snsm + (size_t)nk * SNP_CHP;
    const int t = threadIdx.x;
    const long long k0 = (long long)blockIdx.x * SNP_KPB;
    const long long k1 = (k0 + SNP_KPB < nkey) ? k0 + SNP_KPB : nkey;
    double acc[SNP_NPT];
    short2 pr[SNP_NPT];
#pragma unroll
    for (int i = 0; i < SNP_NPT; i++) {
        acc[i] = 0.0;
        const int q = t + 256 * i;
        pr[i] = q < npairs ? pairs[q] : make_short2(0, 0);
    }
    for (long long c0 = k0; c0 < k1; c0 += SNP_CH) {
        for (int e = t; e < nk * SNP_CH; e += 256) {
            const int a = e / SNP_CH, c = e % SNP_CH;
            const long long key = c0 + c;
            double va = 0.0, vb = 0.0;
            if (key < k1) {
                va = J1[(long long)a * nkey + key] * (H[key] * dV);
                vb = J2[(long long)a * nkey + key];
            }
            ja[a * SNP_CHP + c] = va;
            jb[a * SNP_CHP + c] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SNP_NPT; i++) {
            if (t + 256 * i < npairs) {
                const double* xa = ja + pr[i].x * SNP_CHP;
                const double* xb = jb + pr[i].y * SNP_CHP;
                double s = acc[i];
#pragma unroll 8
                for (int c = 0; c < SNP_CH; c++) s += xa[c] * xb[c];
                acc[i] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < SNP_NPT; i++) {
        const int q = t + 256 * i;
        if (q < npairs) part[(long long)blockIdx.x * npairs + q] = acc[i];
    }
}

// ---------------------------------------------------------------------------
// triple sums G[a][b] = sum_x F_a(x) F_b(x) F_3(x) for all pairs a <= b
// ---------------------------------------------------------------------------
constexpr int GRAM_CH = 64;

template <typename T, int CH = GRAM_CH>
__global__ __launch_bounds__(256) void k_gram_sym(const C<T>* __restrict__ fields, long long fstride, int nf,
                                                 const C<T>* __restrict__ f3, long long nvox, int vpb,
                                                 const short2* __restrict__ pairs, int npairs,
                                                 double2* __restrict__ part) {
    extern __shared__ unsigned char smraw[];
    C<T>* f3s = reinterpret_cast<C<T>*>(smraw);
    C<T>* cs = f3s + CH;                   // [nf][CH + 1]: F_a F_3
    C<T>* gs = cs + (size_t)nf * (CH + 1);  // [nf][CH + 1]: F_b
    const int t = threadIdx.x;
    const long long v0 = (long long)blockIdx.x * vpb;
    const long long v1 = (v0 + vpb < nvox) ? v0 + vpb : nvox;
    short2 pr = make_short2(0, 0);
    if (t < npairs) pr = pairs[t];
    double accr = 0.0, acci = 0.0;
    for (long long c0 = v0; c0 < v1; c0 += CH) {
        if (t < CH) {
            const long long v = c0 + t;
            f3s[t] = (v < v1) ? f3[v] : mk<T>((T)0, (T)0);
        }
        __syncthreads();
        for (int e = t; e < nf * CH; e += 256) {
            const int f = e / CH, x = e % CH;
            const long long v = c0 + x;
            const C<T> g = (v < v1) ? fields[(long long)f * fstride + v] : mk<T>((T)0, (T)0);
            gs[f * (CH + 1) + x] = g;
            cs[f * (CH + 1) + x] = cmul<T>(g, f3s[x]);
        }
        __syncthreads();
        if (t < npairs) {
            T ar = 0, ai = 0;
            const C<T>* ca = cs + pr.x * (CH + 1);
            const C<T>* gb = gs + pr.y * (CH + 1);
#pragma unroll 8
            for (int x = 0; x < CH; x++) {
                const C<T> c = ca[x], g = gb[x];
                ar += c.x * g.x - c.y * g.y;
                ai += c.x * g.y + c.y * g.x;
            }
            accr += (double)ar;
            acci += (double)ai;
        }
        __syncthreads();
    }
    if (t < npairs) part[(long long)blockIdx.x * npairs + t] = make_double2(accr, acci);
}

// ---------------------------------------------------------------------------
// triple-product fields on the M mesh
// ---------------------------------------------------------------------------
// support that actually prunes an axis of length M (-1: whole axis)
static inline int sup_eff(int s, int M) { return (s >= 0 && 2 * s + 1 < M) ? s : -1; }

// compact index of an in-support signed index (inverse of support_map)
__host__ __device__ inline int support_unmap(int m, int s, int n) {
    if (s < 0) return m;
    return (m <= s) ? m : m - (n - (2 * s + 1));
}

// Stage: D = delta(k) / W(k) (field.hpp:378-379, 610-639) on the support cube
// of delta/W, written in place of the first transform's input, and, for the
// modes inside the shell support cube, the first and last k-shell that holds
// the mode (shell rule of field.hpp:371-373 on the N-mesh |k|; two shells only
// where rounding makes neighbouring bin edges overlap).
template <typename T> struct TriStage {
    const C<T>* dk;
    const C<T>* dfull = nullptr;  // set: a full (c2c) spectrum [n0][n1][n2] instead of the half one
    int n[3], nzh, M[3];
    int s3[3], c3[3];  // delta/W support and compact extents
    int s[3], cs[3];   // shell support and compact extents
    AxisTabs t;
    int wp, nk;
    const double* lower;
    const double* upper;
};

template <typename T>
__global__ __launch_bounds__(256) void k_tri_stage(TriStage<T> a, C<T>* __restrict__ D, char2* __restrict__ sid) {
#pragma clang fp contract(off)
    const long long tot = (long long)a.c3[0] * a.c3[1] * a.c3[2];
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= tot) return;
    int cx, cy, cz;
    if (tot <= 0xffffffffLL) {
        // 32-bit quotients (a 64-bit division by a run-time value costs several times more)
        const unsigned eu = (unsigned)e, c2 = (unsigned)a.c3[2], c1 = (unsigned)a.c3[1];
        const unsigned r = eu / c2;
        cz = (int)(eu - r * c2);
        const unsigned rx = r / c1;
        cy = (int)(r - rx * c1);
        cx = (int)rx;
    } else {
        cz = (int)(e % a.c3[2]);
        const long long r = e / a.c3[2];
        cy = (int)(r % a.c3[1]);
        cx = (int)(r / a.c3[1]);
    }
    const int x = support_map(cx, a.s3[0], a.M[0]), y = support_map(cy, a.s3[1], a.M[1]),
              z = support_map(cz, a.s3[2], a.M[2]);
    int idx[3];
    const int mm[3] = {x, y, z};
    for (int q = 0; q < 3; q++) {
        if (a.M[q] == a.n[q]) {
            idx[q] = mm[q];
        } else {
            const int sg = (mm[q] <= a.M[q] / 2) ? mm[q] : mm[q] - a.M[q];
            idx[q] = sg >= 0 ? sg : sg + a.n[q];
        }
    }
    const int i = idx[0], j = idx[1], k = idx[2];
    C<T> d;
    if (a.dfull) {
        d = a.dfull[((long long)i * a.n[1] + j) * a.n[2] + k];
    } else if (k < a.nzh) {
        d = a.dk[((long long)i * a.n[1] + j) * a.nzh + k];
    } else {
        const int im = (a.n[0] - i) % a.n[0], jm = (a.n[1] - j) % a.n[1];
        d = cconj<T>(a.dk[((long long)im * a.n[1] + jm) * a.nzh + (a.n[2] - k)]);
    }
    const double W = wpow((a.t.wf[0][i] * a.t.wf[1][j]) * a.t.wf[2][k], a.wp);
    D[((long long)x * a.M[1] + y) * a.M[2] + z] = mk<T>((T)((double)d.x / W), (T)((double)d.y / W));
    if (sid && in_support(x, a.s[0], a.M[0]) && in_support(y, a.s[1], a.M[1]) && in_support(z, a.s[2], a.M[2])) {
        const double kmag = kmag3(a.t.kv[0][i], a.t.kv[1][j], a.t.kv[2][k]);
        int lo = -1, hi = -2;
        for (int b = 0; b < a.nk; b++)
            if (kmag > a.lower[b] && kmag <= a.upper[b]) {
                if (lo < 0) lo = b;
                hi = b;
            }
        const int sx = support_unmap(x, a.s[0], a.M[0]), sy = support_unmap(y, a.s[1], a.M[1]),
                  sz = support_unmap(z, a.s[2], a.M[2]);
        sid[((long long)sx * a.cs[1] + sy) * a.cs[2] + sz] = make_char2((char)lo, (char)hi);
    }
}

// First (x) pass source of shell f: D where shell f holds the mode, else 0.
// With ylm set, field f < nk is shell f weighted by Y(l1, m1)(k) and field
// nk + f shell f weighted by Y(l2, m2)(k) (calcInverseFourierTransformForBispectrum
// with a Y table, field.hpp:346-401).
template <typename T> struct ShellMaskSrc {
    const C<T>* D;
    const char2* sid;
    int M[3], s[3], cs[3];
    int ylm = 0, nk = 0, l1 = 0, m1 = 0, l2 = 0, m2 = 0;
    int n[3];
    AxisTabs t;
    __device__ inline int to_n(int m, int a) const {
        if (M[a] == n[a]) return m;
        const int sg = (m <= M[a] / 2) ? m : m - M[a];
        return sg >= 0 ? sg : sg + n[a];
    }
    // two phases (the column kernel loads every point's pair before it cooks
    // any): delta/W (D covers the whole mesh) and the shell-id range of the cell
    struct Raw {
        C<T> d;
        unsigned r;  // the char2 as loaded (16 bits, zero-extended), unpacked in cook
    };
    __device__ inline Raw pre(int, int o, int b, int p) const {
        const int sx = support_unmap(p, s[0], M[0]), sy = support_unmap(o, s[1], M[1]),
                  sz = support_unmap(b, s[2], M[2]);
        const unsigned short* sr = reinterpret_cast<const unsigned short*>(sid);
        return Raw{D[((long long)p * M[1] + o) * M[2] + b], (unsigned)sr[((long long)sx * cs[1] + sy) * cs[2] + sz]};
    }
    __device__ inline C<T> operator()(int f, int o, int b, int p, long long) const {
        return cook(pre(f, o, b, p), f, o, b, p);
    }
    __device__ inline C<T> cook(const Raw& w, int f, int o, int b, int p) const {
#pragma clang fp contract(off)
        const C<T> d = w.d;
        const int rlo = (int)(signed char)(w.r & 0xffu), rhi = (int)(signed char)(w.r >> 8);  // char2 x, y
        const int sh = ylm ? f % nk : f;
        if (sh < rlo || sh > rhi) return mk<T>((T)0, (T)0);
        if (!ylm) return d;
        const bool second = f >= nk;
        const Cplx y = ylm_red(second ? l2 : l1, second ? m2 : m1, t.kv[0][to_n(p, 0)], t.kv[1][to_n(o, 1)],
                               t.kv[2][to_n(b, 2)]);
        const double dr = (double)d.x, di = (double)d.y;
        return mk<T>((T)(y.re * dr - y.im * di), (T)(y.re * di + y.im * dr));
    }
};

// Fused last (z) transform + triple sums.  A workgroup takes rows (x, y) of
// the M mesh one at a time: the in-support z coefficients of every shell and
// of delta/W are loaded into LDS (NB lines, delta/W on the last: NB = 21 for
// up to 20 shells, else 24), transformed there, and sum_z F_a F_3 F_b is
// accumulated for 4 x 4 blocks of shell pairs (block (bi, bj), bi <= bj), each
// block split over z among `nsl` lanes.  The fields never return to HBM.
constexpr int TRI_B = 24;  // most lines of a tile (up to 23 shells)
// Rows per workgroup of the fused kernel.  Two workgroups fit a CU, so the
// grid runs in waves of 2 x CUs workgroups, and a partly filled last wave
// costs a full one: take the count in [12, 24] with the fewest rows per CU over
// all waves (18 for 288^2 rows on 256 CUs).  A workgroup's FP32 partial sums
// depend on it, so results are deterministic per device.
static int tri_rows_per_wg(long long nrows, int cus) {
    const long long slots = 2LL * (cus > 0 ? cus : 1);
    int best = 16;
    long long best_cost = -1;
    for (int r = 12; r <= 24; r++) {
        const long long waves = ((nrows + r - 1) / r + slots - 1) / slots;
        if (best_cost < 0 || waves * r < best_cost) {
            best_cost = waves * r;
            best = r;
        }
    }
    return best;
}

template <int L> struct TriFits {
    static constexpr bool value = ((size_t)TRI_B * LdsPad<L, false>::LP + L) * sizeof(float2) <= 65536;
};

constexpr int TRI_NT = 512;  // lanes per workgroup of the fused kernel

template <typename T, int L, int NB>
__global__ __launch_bounds__(TRI_NT) void k_tri_z_gram(const C<T>* __restrict__ F, long long fstride, int nk,
                                                   const int* __restrict__ zsup, int s_f3, long long nrows, int ngrp,
                                                   int nblk, int nsl, int rpw, const C<T>* __restrict__ tw,
                                                   double* __restrict__ part) {
    constexpr int LP = LdsPad<L, false>::LP;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const int t = threadIdx.x;
    const int blk = t / nsl, sl = t % nsl;
    const bool active = blk < nblk;
    int bi = 0, bj = 0;
    if (active) {
        int q = blk;
        while (q >= ngrp - bi) {
            q -= ngrp - bi;
            bi++;
        }
        bj = bi + q;
    }
    // FP32 sums over the workgroup's rows (rpw rows x L / nsl voxels per
    // lane), FP64 from the slice reduction on
    T ar[16], ai[16];
#pragma unroll
    for (int q = 0; q < 16; q++) ar[q] = ai[q] = (T)0;
    const long long r0 = (long long)blockIdx.x * rpw;
    const long long r1 = (r0 + rpw < nrows) ? r0 + rpw : nrows;
    // the next row's in-support coefficients are loaded into registers while
    // the current row is transformed and reduced
    constexpr int NPRE = (NB * L + TRI_NT - 1) / TRI_NT;
    static_assert(NPRE <= 32, "one mask bit per staged element");
    C<T> pre[NPRE];
    // which of this lane's staged elements lie in their field's z support
    // (the same for every row): shell f in zsup[f], delta/W (last line) in s_f3
    unsigned need = 0;
#pragma unroll
    for (int q = 0; q < NPRE; q++) {
        const int e = t + q * TRI_NT;
        const int f = e / L, z = e % L;
        bool in = false;
        if (e < NB * L) {
            if (f < nk) in = in_support(z, zsup[f], L);
            else if (f == NB - 1) in = in_support(z, s_f3, L);
        }
        need |= (unsigned)in << q;
    }
    // tl: the lane index passed through an empty asm statement once per row, so
    // that the compiler recomputes the row-invariant LDS, twiddle and load
    // addresses inside the loop instead of holding them in registers across
    // it (that would cost occupancy).  Staged element q of a lane is
    // e = tl + q TRI_NT: its (line, point) pair is stepped from the previous
    // one (one division per row), and fields are addressed with 32-bit
    // element offsets from the row's base (the host checks NB fstride < 2^31).
    constexpr int DF = TRI_NT / L, DZ = TRI_NT % L;
    const unsigned fs = (unsigned)fstride;
    auto load_row = [&](long long r, int tl) {
        const C<T>* Fr = F + r * L;
        int f = tl / L, z = tl - f * L;
#pragma unroll
        for (int q = 0; q < NPRE; q++) {
            C<T> v = mk<T>((T)0, (T)0);
            if ((need >> q) & 1u) v = Fr[(unsigned)(f < nk ? f : nk) * fs + (unsigned)z];
            pre[q] = v;
            z += DZ;
            f += DF;
            if (z >= L) {
                z -= L;
                f++;
            }
        }
    };
    // the twiddle table in LDS after the tile (read by every stage of every row)
    C<T>* stw = sm + NB * LP;
    for (int e = t; e < L; e += TRI_NT) stw[e] = tw[e];
    if (r0 < r1) load_row(r0, t);
    for (long long r = r0; r < r1; r++) {
        int tl = t;
        asm volatile("" : "+v"(tl));
        {
            int f = tl / L, z = tl - f * L;
#pragma unroll
            for (int q = 0; q < NPRE; q++) {
                if (f < NB) sm[f * LP + z] = pre[q];
                z += DZ;
                f += DF;
                if (z >= L) {
                    z -= L;
                    f++;
                }
            }
        }
        __syncthreads();
        if (r + 1 < r1) load_row(r + 1, tl);
        lds_fft<T, +1, L, NB, TRI_NT, false>(sm, stw, tl);
        if (active) {
            const C<T>* f3 = sm + (NB - 1) * LP;
            const C<T>* fa = sm + 4 * bi * LP;
            const C<T>* fb = sm + 4 * bj * LP;
            for (int z = sl; z < L; z += nsl) {
                const C<T> w = f3[z];
                C<T> c[4], g[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    c[i] = cmul<T>(fa[i * LP + z], w);
                    g[i] = fb[i * LP + z];
                }
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {  // two fused multiply-adds per part into the running sum
                        ar[i * 4 + j] = fma(c[i].x, g[j].x, ar[i * 4 + j]);
                        ar[i * 4 + j] = fma(-c[i].y, g[j].y, ar[i * 4 + j]);
                        ai[i * 4 + j] = fma(c[i].x, g[j].y, ai[i * 4 + j]);
                        ai[i * 4 + j] = fma(c[i].y, g[j].x, ai[i * 4 + j]);
                    }
            }
        }
        __syncthreads();
    }
    // fixed-order reduction over the z slices of each block through LDS, four
    // of the 16 pair sums (re or im) per pass
    double* red = reinterpret_cast<double*>(smraw);
    for (int h = 0; h < 8; h++) {
        const int q0 = (h % 4) * 4;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const T v = (h < 4) ? ar[q0 + q] : ai[q0 + q];
            red[t * 4 + q] = active ? (double)v : 0.0;
        }
        __syncthreads();
        for (int o = t; o < nblk * 4; o += TRI_NT) {
            const int b = o / 4, q = o % 4;
            double sum = 0.0;
            for (int l = 0; l < nsl; l++) sum += red[(b * nsl + l) * 4 + q];
            part[((long long)blockIdx.x * nblk * 16 + b * 16 + q0 + q) * 2 + (h < 4 ? 0 : 1)] = sum;
        }
        __syncthreads();
    }
}

template <typename T, int L>
static int launch_tri_z_gram_L(hm_ctx* ctx, const C<T>* F, long long fstride, int nk, const int* zsup, int s_f3,
                               long long nrows, int ngrp, int nblk, int nsl, int rpw, double* part) {
    const int nb = nk <= 20 ? 21 : TRI_B;
    const size_t fft_lds = ((size_t)nb * LdsPad<L, false>::LP + L) * sizeof(C<T>);
    const size_t red_lds = (size_t)TRI_NT * 4 * sizeof(double);
    const size_t lds = fft_lds > red_lds ? fft_lds : red_lds;
    if ((long long)nb * fstride >= (1LL << 31)) {  // the kernel addresses fields with 32-bit offsets
        ctx->set_error("tri_z_gram: fields too large for 32-bit offsets");
        return -1;
    }
    const C<T>* tw = ctx->twiddle<T>(L);
    if (!tw) return -1;
    const unsigned blocks = (unsigned)((nrows + rpw - 1) / rpw);
    return ctx->launch("tri_z_gram", [&] {
        if (nb == 21)
            hipLaunchKernelGGL((k_tri_z_gram<T, L, 21>), dim3(blocks), dim3(TRI_NT), lds, ctx->stream, F, fstride, nk,
                               zsup, s_f3, nrows, ngrp, nblk, nsl, rpw, tw, part);
        else
            hipLaunchKernelGGL((k_tri_z_gram<T, L, TRI_B>), dim3(blocks), dim3(TRI_NT), lds, ctx->stream, F, fstride,
                               nk, zsup, s_f3, nrows, ngrp, nblk, nsl, rpw, tw, part);
    });
}

// FP32 and a z extent whose 24-line tile fits 64 KB of LDS.
static int tri_fused_ok(int prec, int L) {
    if (prec != HM_FP32) return 0;
    switch (L) {
#define HM_CASE(LL) \
    case LL: return TriFits<LL>::value ? 1 : 0;
        HM_POW2_LENGTHS(HM_CASE)
        HM_MIXED_LENGTHS(HM_CASE)
#undef HM_CASE
    default: return 0;
    }
}

template <typename T>
static int launch_tri_z_gram(hm_ctx* ctx, int L, const C<T>* F, long long fstride, int nk, const int* zsup, int s_f3,
                             long long nrows, int ngrp, int nblk, int nsl, int rpw, double* part) {
    if constexpr (sizeof(T) == sizeof(float)) {
        switch (L) {
#define HM_CASE(LL)                                                                                                    \
    case LL:                                                                                                           \
        if constexpr (TriFits<LL>::value)                                                                              \
            return launch_tri_z_gram_L<T, LL>(ctx, F, fstride, nk, zsup, s_f3, nrows, ngrp, nblk, nsl, rpw, part);     \
        break;
            HM_POW2_LENGTHS(HM_CASE)
            HM_MIXED_LENGTHS(HM_CASE)
#undef HM_CASE
        default: break;
        }
    }
    ctx->set_error("tri_z_gram: unsupported configuration");
    return -1;
}

// ---------------------------------------------------------------------------
// three-point function (calcThreePointCorrelationFunctionForBOX,
// bispec.hpp:1205-1424): every r-bin field j_l(|k| r) Y(l, m)(k) (d/W) / V
// (field.hpp:403-445) formed on the first pass of a full inverse transform
// ---------------------------------------------------------------------------
template <typename T> struct BesselSrc {
    const C<T>* dk;
    int n[3], nzh;
    AxisTabs t;
    int wp, nr;
    double inv_v;
    const double* rb;
    int l[2], m[2];
    const double* sy[2];
    const double* sc[2];
    int plain = 0;  // 1: (d/W) / V alone (calcInverseFourierTransform after the window correction)
    // x pass of a full complex mesh: b = y * n2 + z, p = x, f = field (set f / nr, r-bin f % nr)
    __device__ inline C<T> operator()(int f, int, int b, int p, long long) const {
#pragma clang fp contract(off)
        const int i = p, j = b / n[2], k = b % n[2];
        C<T> d;
        if (k < nzh) {
            d = dk[((long long)i * n[1] + j) * nzh + k];
        } else {
            const int im = (n[0] - i) % n[0], jm = (n[1] - j) % n[1];
            d = cconj<T>(dk[((long long)im * n[1] + jm) * nzh + (n[2] - k)]);
        }
        const double kx = t.kv[0][i], ky = t.kv[1][j], kz = t.kv[2][k];
        const double kmag = kmag3(kx, ky, kz);
        const double wf = wpow((t.wf[0][i] * t.wf[1][j]) * t.wf[2][k], wp);
        const double dr = (double)d.x / wf, di = (double)d.y / wf;
        if (plain) return mk<T>((T)(dr * inv_v), (T)(di * inv_v));
        const int set = f / nr, a = f % nr;
        const Cplx y = ylm_red(l[set], m[set], kx, ky, kz);
        const double tr = y.re * dr - y.im * di, ti = y.re * di + y.im * dr;
        const double jl = jl_spline_eval(sy[set], sc[set], kmag * rb[a]);
        return mk<T>((T)(jl * tr * inv_v), (T)(jl * ti * inv_v));
    }
};

// Shot noise of the three-point function (field.hpp:979-1113): xi(r) weighted
// by Y(l1, m1)(r) Y(l2, m2)(r) summed over m (the polynomial A(mu)), binned
// with the 0.5 Mpc/h fine index below 10^4 fine bins and the 3PCF edges;
// the even part of xi adds to the real sums (and counts), the odd part (imag)
// to the imaginary ones.
constexpr int FINE_NZ = 10000;  // field.hpp:1027
struct ZetaSnArgs {
    int n[3], s[3], c[3];
    double dr[3];
    long long nrows;
    const int2* f2b;
    int nr;
    AngPoly A;
    int imag, count;
};

template <typename T>
__global__ __launch_bounds__(256) void k_bin_xi_ang(const T* __restrict__ xr, ZetaSnArgs a, double* __restrict__ part,
                                                   long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.nrows) return;
    const int i = support_map((int)(row / a.c[1]), a.s[0], a.n[0]);
    const int j = support_map((int)(row % a.c[1]), a.s[1], a.n[1]);
    const double rx = mesh_vec(i, a.n[0], a.dr[0]), ry = mesh_vec(j, a.n[1], a.dr[1]);
    const T* line = xr + row * (long long)a.n[2];
    double* P = part + row * 2 * (long long)a.nr;
    long long* Cn = cnt + row * (long long)a.nr;
    for (int kc = 0; kc < a.c[2]; kc++) {
        const int k = support_map(kc, a.s[2], a.n[2]);
        const double rz = mesh_vec(k, a.n[2], a.dr[2]);
        const double rmag = kmag3(rx, ry, rz);
        const int ri = (int)(rmag / FINE_DR + 0.5);
        if (ri >= FINE_NZ) continue;
        const int2 b = a.f2b[ri];
        if (b.x < 0) continue;
        const double w = (fabs(rmag) < 1.0e-15) ? a.A.a0 : ang_eval(a.A, rz / rmag);
        const double v = w * (double)line[k];
        for (int bb = b.x; bb <= b.y; bb++) {
            P[2 * bb + a.imag] += v;
            if (a.count) Cn[bb] += 1;
        }
    }
}

// A(mu) = sum_m w_m Y(l1, m)(r) Y(l2, -m)(r) as Chebyshev coefficients (exact:
// a polynomial of degree l1 + l2 in mu), and its value a0 at r = 0.
static void make_ang(int l1, int l2, const std::vector<int>& ml, const std::vector<double>& wl, double a0, AngPoly& A) {
    const int deg = l1 + l2;
    A.deg = deg;
    std::vector<double> fv(deg + 1);
    for (int j = 0; j <= deg; j++) {
        const double mu = std::cos(M_PI * (j + 0.5) / (deg + 1));
        const double sx = std::sqrt(std::max(0.0, 1.0 - mu * mu));
        double acc = 0.0;
        for (size_t q = 0; q < ml.size(); q++) {
            const Cplx y1 = ylm_red(l1, ml[q], sx, 0.0, mu);
            const Cplx y2 = ylm_red(l2, -ml[q], sx, 0.0, mu);
            acc += wl[q] * (y1.re * y2.re - y1.im * y2.im);
        }
        fv[j] = acc;
    }
    for (int k = 0; k < ANG_MAX; k++) A.c[k] = 0.0;
    for (int k = 0; k <= deg; k++) {
        double c = 0.0;
        for (int j = 0; j <= deg; j++) c += fv[j] * std::cos(M_PI * k * (j + 0.5) / (deg + 1));
        A.c[k] = (k == 0 ? 1.0 : 2.0) * c / (deg + 1);
    }
    A.a0 = a0;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// survey bispectrum with ELL > 0 (bispec.hpp:4-353): Y_LM(los)-weighted
// fields per M, their cross spectra with d_00 and the shot-noise ijk term
// with a direction-dependent angular weight
// ---------------------------------------------------------------------------
// weights over D then R: w^pw Y(L, M)(los) (conjugated with conj), times cR
// for R; los = p / |p| of the unshifted positions (field.hpp:186-306)
__global__ __launch_bounds__(256) void k_ylm_weights(const double* __restrict__ losD, const double* __restrict__ wD,
                                                    long long nD, const double* __restrict__ losR,
                                                    const double* __restrict__ wR, long long nR, int L, int M, int pw,
                                                    int cj, double cR, double* __restrict__ w) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= nD + nR) return;
    const bool d = q < nD;
    const double* p = d ? losD + 3 * q : losR + 3 * (q - nD);
    const double m = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const Cplx y = ylm_red(L, M, p[0] / m, p[1] / m, p[2] / m);
    const double ww = d ? wD[q] : wR[q - nD];
    const double wp = pw == 1 ? ww : pow(ww, pw);
    double re = y.re * wp, im = (cj ? -y.im : y.im) * wp;
    if (!d) {
        re = cR * re;
        im = cR * im;
    }
    w[2 * q] = re;
    w[2 * q + 1] = im;
}

// P_{l m}(k) sums of d_00 conj(N) with N a full (c2c) spectrum (calcPowerSpectrum,
// field.hpp:765-852): per row partials [nk][2] += w Y(l, m)(k) (d conj N - sn S)/W^2
struct YBinArgs {
    int n[3], nzh, s01, nrow_y, nzmax;
    long long nrows;
    AxisTabs t;
    int wp;
    const int2* f2b;
    int nk, l, m;
    double w, sn_re, sn_im;
};

template <typename T>
__global__ __launch_bounds__(256) void k_bin_ycross(const C<T>* __restrict__ dk, const C<T>* __restrict__ nf,
                                                   YBinArgs a, double* __restrict__ part, long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.nrows) return;
    const int i = support_map((int)(row / a.nrow_y), a.s01, a.n[0]);
    const int j = support_map((int)(row % a.nrow_y), a.s01, a.n[1]);
    double* P = part + row * 2 * (long long)a.nk;
    long long* Cn = cnt ? cnt + row * (long long)a.nk : nullptr;
    const long long rbase = ((long long)i * a.n[1] + j) * a.nzh;
    const int im = (a.n[0] - i) % a.n[0], jm = (a.n[1] - j) % a.n[1];
    for (int kz = 0; kz <= a.nzmax; kz++) {
        const C<T> d = dk[rbase + kz];
        const int nmir = (kz > 0 && kz < a.n[2] - kz) ? 2 : 1;
        for (int mm = 0; mm < nmir; mm++) {
            const int ii = mm ? im : i, jj = mm ? jm : j, kk = mm ? (a.n[2] - kz) : kz;
            const double kx = a.t.kv[0][ii], ky = a.t.kv[1][jj], kzv = a.t.kv[2][kk];
            const double kmag = kmag3(kx, ky, kzv);
            const int kindex = (int)(kmag / FINE_DK + 0.5);
            if (kindex >= FINE_N) continue;
            const int2 b = a.f2b[kindex];
            if (b.x < 0) continue;
            const double dr = (double)d.x, di = mm ? -(double)d.y : (double)d.y;  // d_00(-k) = conj d_00(k)
            const C<T> e = nf[((long long)ii * a.n[1] + jj) * a.n[2] + kk];
            double pr = dr * (double)e.x + di * (double)e.y, pi = di * (double)e.x - dr * (double)e.y;
            const double S = (a.t.sf[0][ii] * a.t.sf[1][jj]) * a.t.sf[2][kk];
            pr -= a.sn_re * S;
            pi -= a.sn_im * S;
            const double W = wpow((a.t.wf[0][ii] * a.t.wf[1][jj]) * a.t.wf[2][kk], a.wp);
            const double W2 = pow(W, 2);
            pr /= W2;
            pi /= W2;
            const Cplx y = ylm_red(a.l, a.m, kx, ky, kzv);
            const double vr = pr * y.re - pi * y.im, vi = pr * y.im + pi * y.re;
            for (int bb = b.x; bb <= b.y; bb++) {
                P[2 * bb] += a.w * vr;
                P[2 * bb + 1] += a.w * vi;
                if (Cn) Cn[bb] += 1;
            }
        }
    }
}

// (dn_LM conj N_00 - sn S)/W^2 / V on the full mesh, N_00 from its half spectrum
// (calcShotNoiseForBispectrum_ijk, field.hpp:1116-1160)
struct YXiArgs {
    int n[3], nzh;
    AxisTabs t;
    int wp;
    double sn_re, sn_im, inv_v;
};

template <typename T>
__global__ __launch_bounds__(256) void k_ycross_xi(const C<T>* __restrict__ dl, const C<T>* __restrict__ n00, YXiArgs a,
                                                  C<T>* __restrict__ out) {
#pragma clang fp contract(off)
    const long long ng = (long long)a.n[0] * a.n[1] * a.n[2];
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ng; e += st) {
        const int k = (int)(e % a.n[2]);
        const long long r = e / a.n[2];
        const int j = (int)(r % a.n[1]), i = (int)(r / a.n[1]);
        C<T> nn;
        if (k < a.nzh) {
            nn = n00[((long long)i * a.n[1] + j) * a.nzh + k];
        } else {
            const int im = (a.n[0] - i) % a.n[0], jm = (a.n[1] - j) % a.n[1];
            nn = cconj<T>(n00[((long long)im * a.n[1] + jm) * a.nzh + (a.n[2] - k)]);
        }
        const C<T> d = dl[e];
        double pr = (double)d.x * (double)nn.x + (double)d.y * (double)nn.y;
        double pi = (double)d.y * (double)nn.x - (double)d.x * (double)nn.y;
        const double S = (a.t.sf[0][i] * a.t.sf[1][j]) * a.t.sf[2][k];
        pr -= a.sn_re * S;
        pi -= a.sn_im * S;
        const double W2 = pow(wpow((a.t.wf[0][i] * a.t.wf[1][j]) * a.t.wf[2][k], a.wp), 2);
        out[e] = mk<T>((T)(a.inv_v * (pr / W2)), (T)(a.inv_v * (pi / W2)));
    }
}

// |n|^2 histogram of A(r) xi(r), A(r) = sum_c w_c Y(l1, m1_c)(r) Y(l2, m2_c)(r)
// (r with the Nyquist index negative, field.hpp:797-799): FP64 atomics (their
// order can change the last bits of a sum)
constexpr int YH_MAXC = 32;
struct YHistArgs {
    int n[3];
    int l1, l2, nc;
    int m1[YH_MAXC], m2[YH_MAXC];
    double w[YH_MAXC];
    long long nkey;
    SnKeys keys;
};

template <typename T>
__global__ __launch_bounds__(256) void k_yhist(const C<T>* __restrict__ xi, YHistArgs a, double* __restrict__ Hre,
                                              double* __restrict__ Him) {
#pragma clang fp contract(off)
    const long long ng = (long long)a.n[0] * a.n[1] * a.n[2];
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ng; e += st) {
        const int k = (int)(e % a.n[2]);
        const long long r = e / a.n[2];
        const int j = (int)(r % a.n[1]), i = (int)(r / a.n[1]);
        const int si = i < a.n[0] / 2 ? i : i - a.n[0], sj = j < a.n[1] / 2 ? j : j - a.n[1],
                  sk = k < a.n[2] / 2 ? k : k - a.n[2];
        const long long key =
            a.keys.aniso ? ((long long)abs(si) * a.keys.e1 + abs(sj)) * a.keys.e2 + abs(sk)
                         : (long long)si * si + (long long)sj * sj + (long long)sk * sk;
        if (key >= a.nkey) continue;
        // the direction of r (in cell units when the cells are cubes)
        const double vx = a.keys.aniso ? (double)si * a.keys.d[0] : (double)si;
        const double vy = a.keys.aniso ? (double)sj * a.keys.d[1] : (double)sj;
        const double vz = a.keys.aniso ? (double)sk * a.keys.d[2] : (double)sk;
        double ar = 0.0, ai = 0.0;
        for (int c = 0; c < a.nc; c++) {
            const Cplx y1 = ylm_red(a.l1, a.m1[c], vx, vy, vz);
            const Cplx y2 = ylm_red(a.l2, a.m2[c], vx, vy, vz);
            ar += a.w[c] * (y1.re * y2.re - y1.im * y2.im);
            ai += a.w[c] * (y1.re * y2.im + y1.im * y2.re);
        }
        const C<T> x = xi[e];
        const double vr = ar * (double)x.x - ai * (double)x.y, vi = ar * (double)x.y + ai * (double)x.x;
        if (vr != 0.0) atomicAdd(&Hre[key], vr);
        if (vi != 0.0) atomicAdd(&Him[key], vi);
    }
}

// Shot noise of the survey three-point function with ELL > 0 (bispec.hpp:388-460,
// 644-720, corr_for_3pcf in field.hpp:979-1113): the complex xi_LM on the full
// mesh, weighted by A(r) = sum_c w_c Y(l1, m1_c)(r) Y(l2, m2_c)(r) (r with the
// Nyquist index negative), binned by |r| like k_bin_xi_ang into per-row
// partials (re, im) and, when count is set, mesh points per bin.
struct ZetaYArgs {
    int n[3], s[3], c[3];
    double dr[3];
    long long nrows;
    const int2* f2b;
    int nr;
    int l1, l2, nc;
    int m1[YH_MAXC], m2[YH_MAXC];
    double w[YH_MAXC];
    int count;
};

template <typename T>
__global__ __launch_bounds__(256) void k_bin_xi_y(const C<T>* __restrict__ xi, ZetaYArgs a, double* __restrict__ part,
                                                 long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= a.nrows) return;
    const int i = support_map((int)(row / a.c[1]), a.s[0], a.n[0]);
    const int j = support_map((int)(row % a.c[1]), a.s[1], a.n[1]);
    const double rx = mesh_vec(i, a.n[0], a.dr[0]), ry = mesh_vec(j, a.n[1], a.dr[1]);
    const C<T>* line = xi + ((long long)i * a.n[1] + j) * (long long)a.n[2];
    double* P = part + row * 2 * (long long)a.nr;
    long long* Cn = cnt + row * (long long)a.nr;
    for (int kc = 0; kc < a.c[2]; kc++) {
        const int k = support_map(kc, a.s[2], a.n[2]);
        const double rz = mesh_vec(k, a.n[2], a.dr[2]);
        const double rmag = kmag3(rx, ry, rz);
        const int ri = (int)(rmag / FINE_DR + 0.5);
        if (ri >= FINE_NZ) continue;
        const int2 b = a.f2b[ri];
        if (b.x < 0) continue;
        double ar = 0.0, ai = 0.0;
        for (int q = 0; q < a.nc; q++) {
            const Cplx y1 = ylm_red(a.l1, a.m1[q], rx, ry, rz);
            const Cplx y2 = ylm_red(a.l2, a.m2[q], rx, ry, rz);
            ar += a.w[q] * (y1.re * y2.re - y1.im * y2.im);
            ai += a.w[q] * (y1.re * y2.im + y1.im * y2.re);
        }
        const C<T> x = line[k];
        const double vr = ar * (double)x.x - ai * (double)x.y, vi = ar * (double)x.y + ai * (double)x.x;
        for (int bb = b.x; bb <= b.y; bb++) {
            P[2 * bb] += vr;
            P[2 * bb + 1] += vi;
            if (a.count) Cn[bb] += 1;
        }
    }
}

// (f / W) / V on the full mesh (the window correction of calcInverseFourierTransform)
template <typename T>
__global__ __launch_bounds__(256) void k_full_over_w(const C<T>* __restrict__ f, int n0, int n1, int n2, AxisTabs t,
                                                    int wp, double inv_v, C<T>* __restrict__ out) {
#pragma clang fp contract(off)
    const long long ng = (long long)n0 * n1 * n2;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ng; e += st) {
        const int k = (int)(e % n2);
        const long long r = e / n2;
        const int j = (int)(r % n1), i = (int)(r / n1);
        const double W = wpow((t.wf[0][i] * t.wf[1][j]) * t.wf[2][k], wp);
        const C<T> v = f[e];
        out[e] = mk<T>((T)(((double)v.x / W) * inv_v), (T)(((double)v.y / W) * inv_v));
    }
}

struct BoxPlanImpl {
    hm_box_config cfg;
    // launch shapes (the autotuner's choice; results do not depend on them)
    unsigned tune_wide = 0;   // wide column tiles, bit per tunable length (256, 270 or 288, 512; 3: the shell x pass)
    int tune_bin = 256;       // threads per block of the P_l binning
    Mesh mesh;
    int nk = 0, ne = 0;
    std::vector<double> kbin, lower, upper;
    std::vector<int2> f2b;
    long long tab_off[3];
    int s[3];   // shell support (signed |n| bound) per axis, -1 = all
    int s3[3];  // support of delta/W on the triple mesh
    int M[3];
    int nzmax = 0, s01 = -1, nrow_y = 0;
    long long nrows = 0;
    long long nkey = 0;   // shot-noise radial keys (SnKeys)
    long long amax = 0;   // largest row key
    SnKeys keys;
    int na = 0;           // distinct i'^2 + j'^2 over the rows of the mesh
    SortWork sort;
    std::vector<short2> pairs;
    std::vector<long long> shell_nmode;
    bool counted = false;
    size_t bytes = 0;
    DevBuf tab, f2b_d, lu_d, kb_d, fixed, dk, pk_part, pk_cnt, pk_red, pk_cnt_red, xi_half, xi_real, hist, jtab,
        pairs_d, a_start, a_rows, a_idx, xi_R, sn_pairs, fields, gram_part, gram_red, shell_cnt;
    long long gram_blocks = 0;
    int gram_vpb = 4096;
    bool fused = false;  // z transform of the triple fields fused with the triple sums
    int ngrp = 0, nblk = 0, nsl = 0;  // 4-shell groups, group pairs, z slices per pair block (fused)
    int tri_rpw = 16;                 // rows per workgroup of the fused kernel
    int bin_kc = 1;                   // kz chunks per row of the P_l binning
    int gram_cols = 0;                // doubles per partial row of the triple sums
    DevBuf sid;                       // char2 first/last shell per mode of the shell support cube
    DevBuf rpos, rw;                  // D then R positions and weights of a reconstructed box
    // xi_l
    int nr = 0, nxe = 0;
    std::vector<double> rbin, rlower, rupper;
    int rs[3] = {-1, -1, -1};  // support of the separations that reach an r-bin
    long long xi_rows = 0;
    DevBuf f2b_r_d, xi_r, xi_part, xi_cnt, xi_red, xi_cnt_red;
    // bispectrum multipole (l1, l2, L); bk_gen: anything but 000
    int l1 = 0, l2 = 0, Lb = 0;
    bool bk_gen = false;
    std::vector<int> mlist;           // m1 with a non-zero 3j weight (m2 = -m1, M = 0)
    std::vector<double> wlist;
    double w_const = 0, w_row = 0, w_col = 0;  // weights of N, P_l1(k1) (l2 = 0) and P_l2(k2) (l1 = 0)
    std::vector<int> bin_ells;        // multipoles binned: the requested ones plus those the shot noise needs
    int e_row = -1, e_col = -1;       // their indices in bin_ells
    AngPoly ang;
    bool sn_odd = false;              // odd l1 + l2: the odd part of xi enters the Bessel sweep too
    DevBuf jtab2, hist2, sn_pairs2, snp_part;
    // reconstructed box, bispectrum / three-point function: the spectrum of D + alpha^2 R and the binned
    // cross spectrum with D - alpha R (re, im)
    DevBuf dk2, pkx_part, pkx_part_im, pkx_red;
    DevBuf dki, spos;  // interlacing: the shifted grid's spectrum and positions
    // per shell field: its own supports on the triple mesh (x, y, z; -1 full)
    // and the z one alone
    std::vector<int4> fsup;
    std::vector<int> fsz;
    DevBuf fsup_d, fsz_d;
    DevBuf gpairs_d;  // (a, nk + b): the cross sums of the two Y-weighted shell sets
    // three-point function
    bool zeta = false, z_gen = false, z_odd = false;
    // an odd mesh length: the reference's separation rule (i < n/2 ? i : i - n, bispec.hpp:1040-1042) gives
    // the planes next to the Nyquist index different |r| at x and -x, so the odd part of xi enters the
    // shot-noise sums even for even l1 + l2
    bool odd_mesh = false;
    int zl[3] = {0, 0, 0};
    std::vector<int> zmlist;
    std::vector<double> zwlist;
    AngPoly zang;
    std::vector<short2> zpairs;       // 000: a <= b on one field set; else (a, nr + b)
    int zrs[3] = {-1, -1, -1};
    long long z_rows = 0;
    DevBuf zf2b_d, zrb_d, zsy[2], zsc[2], zfields, zr, zpart, zcnt, zred, zcnt_red, zpairs_d;
    DevBuf qo_d;  // batched mode: raw triple-sum column of each bin pair
    // survey B with ELL > 0: full spectra of the Y_LM-weighted fields, xi, the
    // imaginary fixed-point mesh, binning partials, the |n|^2 histogram
    DevBuf cf1, cf2, cx, fixed_im, ypart, yred, yh;
};

static int alloc(hm_ctx* ctx, BoxPlanImpl* p, DevBuf& b, size_t bytes) {
    if (dev_alloc(ctx, b, bytes)) return -1;
    p->bytes += bytes;
    return 0;
}

template <typename T> static int plan_alloc(hm_ctx* ctx, BoxPlanImpl* p) {
    const Mesh& m = p->mesh;
    if (alloc(ctx, p, p->fixed, sizeof(long long) * (size_t)m.ng())) return -1;
    if (alloc(ctx, p, p->dk, sizeof(C<T>) * (size_t)m.nh())) return -1;
    const size_t ncols = (size_t)p->nk * (size_t)std::max(p->ne, 1);
    // kz chunks of the P_l binning: enough lanes to fill the device, 16+ kz per chunk
    p->bin_kc = std::max(1, std::min(8, (p->nzmax + 1) / 16));
    const size_t prow = (size_t)p->nrows * (size_t)p->bin_kc;
    if (alloc(ctx, p, p->pk_part, sizeof(double) * ncols * prow)) return -1;
    if (alloc(ctx, p, p->pk_cnt, sizeof(long long) * (size_t)p->nk * prow)) return -1;
    if (alloc(ctx, p, p->pk_red, sizeof(double) * ncols)) return -1;
    if (alloc(ctx, p, p->pk_cnt_red, sizeof(long long) * (size_t)p->nk)) return -1;
    if (p->cfg.bk) {
        if (alloc(ctx, p, p->xi_half, sizeof(C<T>) * (size_t)m.nh())) return -1;
        if (alloc(ctx, p, p->xi_real, sizeof(C<T>) * (size_t)p->na * (size_t)m.nzh())) return -1;
        if (alloc(ctx, p, p->hist, sizeof(double) * (size_t)p->nkey)) return -1;
        if (alloc(ctx, p, p->xi_R, sizeof(T) * (size_t)p->na * (size_t)m.n[2])) return -1;
        if (alloc(ctx, p, p->jtab, sizeof(double) * (size_t)p->nk * (size_t)p->nkey)) return -1;
        if (p->bk_gen && p->l2 != p->l1 &&
            alloc(ctx, p, p->jtab2, sizeof(double) * (size_t)p->nk * (size_t)p->nkey))
            return -1;
        // (the second, odd-part sweep also serves the cross spectrum of a reconstructed box)
        if ((alloc(ctx, p, p->hist2, sizeof(double) * (size_t)p->nkey) ||
             alloc(ctx, p, p->sn_pairs2, sizeof(double) * p->pairs.size())))
            return -1;
        if (alloc(ctx, p, p->pairs_d, sizeof(short2) * p->pairs.size())) return -1;
        if (alloc(ctx, p, p->sn_pairs, sizeof(double) * p->pairs.size())) return -1;
        if (alloc(ctx, p, p->snp_part,
                  sizeof(double) * p->pairs.size() * (size_t)((p->nkey + SNP_KPB - 1) / SNP_KPB)))
            return -1;
        const long long nvox = (long long)p->M[0] * p->M[1] * p->M[2];
        // general multipole: two shell sets, delta/W transformed, and delta/W kept untransformed for the
        // shell loader of every m
        const int nfield = p->bk_gen ? 2 * p->nk + 2 : p->nk + 1;
        if (alloc(ctx, p, p->fields, sizeof(C<T>) * (size_t)nvox * (size_t)nfield)) return -1;
        p->gram_blocks = (nvox + p->gram_vpb - 1) / p->gram_vpb;
        p->fused = !p->bk_gen && tri_fused_ok(p->cfg.precision, p->M[2]) && p->nk < TRI_B;
        p->gram_cols = 2 * std::min((int)p->pairs.size(), 256);
        if (p->fused) {
            p->tri_rpw = tri_rows_per_wg((long long)p->M[0] * p->M[1], ctx->cu_count);
            p->gram_blocks = ((long long)p->M[0] * p->M[1] + p->tri_rpw - 1) / p->tri_rpw;
            p->ngrp = (p->nk + 3) / 4;
            p->nblk = p->ngrp * (p->ngrp + 1) / 2;
            p->nsl = TRI_NT / p->nblk;
            p->gram_cols = 2 * 16 * p->nblk;
        }
        if (alloc(ctx, p, p->gram_part, sizeof(double) * (size_t)p->gram_blocks * p->gram_cols)) return -1;
        if (alloc(ctx, p, p->gram_red, sizeof(double) * (size_t)p->gram_cols)) return -1;
        long long ns = 1;
        for (int a = 0; a < 3; a++) ns *= support_count(sup_eff(p->s[a], p->M[a]), p->M[a]);
        if (alloc(ctx, p, p->sid, sizeof(char2) * (size_t)ns)) return -1;
    }
    if (alloc(ctx, p, p->shell_cnt, sizeof(unsigned long long) * (size_t)p->nk)) return -1;
    if (p->zeta) {
        const int nr = p->cfg.n_rbin;
        const int nf = p->z_gen ? 2 * nr : nr;
        if (!p->xi_half.p && alloc(ctx, p, p->xi_half, sizeof(C<T>) * (size_t)m.nh())) return -1;
        if (alloc(ctx, p, p->zfields, sizeof(C<T>) * (size_t)m.ng() * (size_t)(nf + 1)) ||
            alloc(ctx, p, p->zr, sizeof(T) * (size_t)p->z_rows * (size_t)m.n[2]) ||
            alloc(ctx, p, p->zpart, sizeof(double) * 2 * (size_t)nr * (size_t)std::max(p->z_rows, (long long)1) +
                                        sizeof(double) * 2 * 256 * (size_t)((m.ng() + p->gram_vpb - 1) / p->gram_vpb)) ||
            alloc(ctx, p, p->zcnt, sizeof(long long) * (size_t)nr * (size_t)p->z_rows) ||
            alloc(ctx, p, p->zred, sizeof(double) * 2 * 256) || alloc(ctx, p, p->zcnt_red, sizeof(long long) * nr) ||
            alloc(ctx, p, p->zpairs_d, sizeof(short2) * p->zpairs.size()))
            return -1;
    }
    if (p->nxe > 0) {
        if (!p->cfg.bk && !p->xi_half.p && alloc(ctx, p, p->xi_half, sizeof(C<T>) * (size_t)m.nh())) return -1;
        const size_t cols = (size_t)p->nr * (size_t)p->nxe;
        if (alloc(ctx, p, p->xi_r, sizeof(T) * (size_t)p->xi_rows * (size_t)m.n[2])) return -1;
        if (alloc(ctx, p, p->xi_part, sizeof(double) * cols * (size_t)p->xi_rows)) return -1;
        if (alloc(ctx, p, p->xi_cnt, sizeof(long long) * (size_t)p->nr * (size_t)p->xi_rows)) return -1;
        if (alloc(ctx, p, p->xi_red, sizeof(double) * cols)) return -1;
        if (alloc(ctx, p, p->xi_cnt_red, sizeof(long long) * (size_t)p->nr)) return -1;
    }
    return 0;
}

static AxisTabs axis_tabs(const BoxPlanImpl* p) {
    AxisTabs t;
    const double* base = reinterpret_cast<const double*>(p->tab.p);
    const long long ntot = (long long)p->mesh.n[0] + p->mesh.n[1] + p->mesh.n[2];
    for (int a = 0; a < 3; a++) {
        t.kv[a] = base + p->tab_off[a];
        t.wf[a] = base + ntot + p->tab_off[a];
        t.sf[a] = base + 2 * ntot + p->tab_off[a];
        t.iw[a] = base + 3 * ntot + p->tab_off[a];
    }
    return t;
}


// XiSrc's per-x table, after the four axis tables (32-byte aligned: the buffer
// is, and the offset is 4 (n0 + n1 + n2) doubles)
static const double* xi_xq(const BoxPlanImpl* p) {
    const long long ntot = (long long)p->mesh.n[0] + p->mesh.n[1] + p->mesh.n[2];
    return reinterpret_cast<const double*>(p->tab.p) + 4 * ntot;
}

void plan_destroy(BoxPlanImpl* p);

BoxPlanImpl* plan_create(hm_ctx* ctx, const hm_box_config* cfg) {
    auto* p = new BoxPlanImpl();
    p->cfg = *cfg;
    Mesh& m = p->mesh;
    for (int a = 0; a < 3; a++) {
        m.n[a] = cfg->mesh.n_mesh[a];
        m.L[a] = cfg->mesh.boxsize[a];
    }
    m.assign = cfg->mesh.assign;
    auto fail = [&](const std::string& s) -> BoxPlanImpl* {
        ctx->set_error(s);
        plan_destroy(p);
        return nullptr;
    };
    for (int a = 0; a < 3; a++) {
        if (m.n[a] < 8 || !fft_length_supported(m.n[a]))
            return fail("box plan: n_mesh must be in [8, " + std::to_string(RT_MAX_L) + "]");
        if (!(m.L[a] > 0)) return fail("box plan: boxsize must be > 0");
    }
    if (m.n[2] < 16) return fail("box plan: n_mesh_z must be >= 16");
    if (m.n[2] % 2) return fail("box plan: n_mesh_z must be even (real-to-complex transform of the z rows)");
    p->odd_mesh = (m.n[0] % 2) != 0 || (m.n[1] % 2) != 0;
    if (m.assign < 0 || m.assign > 2) return fail("box plan: assign must be NGP, CIC or TSC");
    if (cfg->n_kbin < 2 || cfg->n_kbin > 64) return fail("box plan: n_kbin must be in [2, 64]");
    if (cfg->n_pk_ells < 0 || cfg->n_pk_ells > HM_MAX_ELLS) return fail("box plan: n_pk_ells out of range");
    p->nk = cfg->n_kbin;
    p->bin_ells.assign(cfg->pk_ells, cfg->pk_ells + cfg->n_pk_ells);
    if (cfg->bk) {
        // calcBiSpectrumForBOX (bispec.hpp:885-891, 926-963): the multipole and its weights
        p->l1 = cfg->bk_ell[0];
        p->l2 = cfg->bk_ell[1];
        p->Lb = cfg->bk_ell[2];
        if (p->l1 < 0 || p->l2 < 0 || p->Lb < 0 || p->l1 > 8 || p->l2 > 8 || p->Lb > 16)
            return fail("box plan: bispectrum multipole out of range");
        if (std::fabs(wigner_3j(p->l1, p->l2, p->Lb, 0, 0, 0)) < 1.0e-10)
            return fail("box plan: this multipole combination is not allowed (wigner_3j(ell1, ell2, ELL, 0,0,0) == 0)");
        p->bk_gen = !(p->l1 == 0 && p->l2 == 0 && p->Lb == 0);
        p->sn_odd = ((p->l1 + p->l2) & 1) != 0;
        for (int m1 = -p->l1; m1 <= p->l1; m1++)
            for (int m2 = -p->l2; m2 <= p->l2; m2++) {
                double w = wigner_3j(p->l1, p->l2, p->Lb, 0, 0, 0) * wigner_3j(p->l1, p->l2, p->Lb, m1, m2, 0);
                w *= (double)(2 * p->Lb + 1) * (double)(2 * p->l1 + 1) * (double)(2 * p->l2 + 1);
                if (std::fabs(w) < 1.0e-10) continue;
                p->mlist.push_back(m1);
                p->wlist.push_back(w);
                if (p->l1 == 0 && p->l2 == 0) p->w_const += w;
                if (p->l2 == 0) p->w_row += w;
                if (p->l1 == 0) p->w_col += w;
            }
        auto need = [&](int l) -> int {
            for (size_t e = 0; e < p->bin_ells.size(); e++)
                if (p->bin_ells[e] == l) return (int)e;
            p->bin_ells.push_back(l);
            return (int)p->bin_ells.size() - 1;
        };
        if (p->l2 == 0) p->e_row = need(p->l1);
        if (p->l1 == 0) p->e_col = need(p->l2);
        if ((int)p->bin_ells.size() > BIN_MAX_ELLS) return fail("box plan: too many multipoles");
        make_ang(p->l1, p->l2, p->mlist, p->wlist, p->w_const, p->ang);
    }
    p->ne = (int)p->bin_ells.size();
    // k bins, ToolClass::setKbin (tools.hpp:118-124)
    p->kbin.resize(p->nk);
    const double dkb = (cfg->kmax - cfg->kmin) / (double)(p->nk - 1);
    for (int i = 0; i < p->nk; i++) p->kbin[i] = cfg->kmin + dkb * (double)i;
    const double dk_bin = p->kbin[1] - p->kbin[0];
    p->lower.resize(p->nk);
    p->upper.resize(p->nk);
    for (int j = 0; j < p->nk; j++) {
        p->lower[j] = (p->kbin[j] > dk_bin * 0.5) ? (p->kbin[j] - dk_bin * 0.5) : 0.0;  // field.hpp:828, 371
        p->upper[j] = p->kbin[j] + dk_bin * 0.5;
    }
    // fine -> coarse fold (field.hpp:827-837)
    p->f2b.assign(FINE_N, make_int2(-1, -2));
    for (int i = 0; i < FINE_N; i++) {
        const double kt = (double)i * FINE_DK;
        int lo = -1, hi = -2;
        for (int j = 0; j < p->nk; j++)
            if ((kt > p->lower[j]) && (kt <= p->upper[j])) {
                if (lo < 0) lo = j;
                else if (j != hi + 1) return fail("box plan: non-contiguous k-bin overlap");
                hi = j;
            }
        p->f2b[i] = make_int2(lo, hi);
    }
    // supports: every mode whose fine centre can land in a bin
    double umax = 0;
    for (int j = 0; j < p->nk; j++) umax = std::max(umax, p->upper[j]);
    for (int a = 0; a < 3; a++) {
        const double dk = 2.0 * M_PI / m.L[a];
        long long s = (long long)std::floor((umax + FINE_DK) / dk) + 1;
        p->s[a] = (!cfg->prune || 2 * s + 1 >= m.n[a] || s >= m.n[a] / 2) ? -1 : (int)s;
    }
    // P binning rows: x, y in the support, kz up to the support
    p->s01 = (p->s[0] < 0 || p->s[1] < 0) ? -1 : std::max(p->s[0], p->s[1]);
    if (p->s01 >= 0 && (2 * p->s01 + 1 >= m.n[0] || 2 * p->s01 + 1 >= m.n[1])) p->s01 = -1;
    const int nrx = support_count(p->s01, m.n[0]);
    p->nrow_y = support_count(p->s01, m.n[1]);
    p->nrows = (long long)nrx * p->nrow_y;
    p->nzmax = (p->s[2] < 0) ? m.n[2] / 2 : std::min(p->s[2], m.n[2] / 2);
    // triple-product mesh
    for (int a = 0; a < 3; a++) {
        p->M[a] = m.n[a];
        p->s3[a] = -1;
        if (cfg->regrid && p->s[a] >= 0) {
            const int want = fft_length_at_least(4 * p->s[a] + 1);
            if (want > 0 && want < m.n[a]) {
                p->M[a] = want;
                p->s3[a] = 2 * p->s[a];
            }
        }
    }
    // pairs a <= b
    for (int a = 0; a < p->nk; a++)
        for (int b = a; b < p->nk; b++) p->pairs.push_back(make_short2((short)a, (short)b));
    if (p->bk_gen) {  // every ordered pair: B_{l1 l2 L}(k1, k2) is not symmetric
        p->pairs.clear();
        for (int a = 0; a < p->nk; a++)
            for (int b = 0; b < p->nk; b++) p->pairs.push_back(make_short2((short)a, (short)b));
    }
    if (cfg->bk && !p->bk_gen && p->pairs.size() > 256) return fail("box plan: at most 22 k-bins for the bispectrum");
    if (cfg->bk) {
        const double dr0 = m.L[0] / m.n[0], dr1 = m.L[1] / m.n[1], dr2 = m.L[2] / m.n[2];
        // signed indices of the rvec rule (bispec.hpp:1040-1042): |i'| <= n - n/2
        const long long h0 = m.n[0] - m.n[0] / 2, h1 = m.n[1] - m.n[1] / 2, h2 = m.n[2] / 2;
        p->keys.aniso = !(dr0 == dr1 && dr1 == dr2);
        if (p->keys.aniso) {
            p->keys.e1 = (int)(h1 + 1);
            p->keys.e2 = (int)(h2 + 1);
            p->keys.d[0] = dr0;
            p->keys.d[1] = dr1;
            p->keys.d[2] = dr2;
            p->nkey = (h0 + 1) * (h1 + 1) * (h2 + 1);
            p->amax = (h0 + 1) * (h1 + 1) - 1;
        } else {
            p->nkey = h0 * h0 + h1 * h1 + h2 * h2 + 1;
            p->amax = h0 * h0 + h1 * h1;
        }
        // rows (i, j) grouped by their row key: the minimum-image i'^2 + j'^2, or (|i'|, |j'|)
        std::vector<int> cnt((size_t)p->amax + 1, 0), key((size_t)m.n[0] * m.n[1]);
        for (int i = 0; i < m.n[0]; i++) {
            long long a = (i < m.n[0] / 2) ? i : i - m.n[0];
            for (int j = 0; j < m.n[1]; j++) {
                long long b = (j < m.n[1] / 2) ? j : j - m.n[1];
                const long long rk = p->keys.aniso ? std::llabs(a) * p->keys.e1 + std::llabs(b) : a * a + b * b;
                key[(size_t)i * m.n[1] + j] = (int)rk;
                cnt[(size_t)rk]++;
            }
        }
        std::vector<int> a_idx((size_t)p->amax + 1, -1), a_start(1, 0);
        for (long long a = 0; a <= p->amax; a++)
            if (cnt[(size_t)a]) {
                a_idx[(size_t)a] = (int)(a_start.size() - 1);
                a_start.push_back(a_start.back() + cnt[(size_t)a]);
            }
        p->na = (int)(a_start.size() - 1);
        std::vector<int> fillp(a_start.begin(), a_start.end() - 1), rows(key.size());
        for (size_t r = 0; r < key.size(); r++) rows[(size_t)fillp[(size_t)a_idx[(size_t)key[r]]]++] = (int)r;
        if (alloc(ctx, p, p->a_start, sizeof(int) * a_start.size()) ||
            alloc(ctx, p, p->a_rows, sizeof(int) * rows.size()) || alloc(ctx, p, p->a_idx, sizeof(int) * a_idx.size())) {
            plan_destroy(p);
            return nullptr;
        }
        if (hipMemcpy(p->a_start.p, a_start.data(), sizeof(int) * a_start.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->a_rows.p, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->a_idx.p, a_idx.data(), sizeof(int) * a_idx.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail("box plan: row table upload failed");
    }
    // xi_l: r bins (tools.hpp:126-132), fine -> coarse fold (field.hpp:948-960)
    if (cfg->n_xi_ells < 0 || cfg->n_xi_ells > HM_MAX_ELLS) return fail("box plan: n_xi_ells out of range");
    p->nxe = cfg->n_xi_ells;
    std::vector<int2> f2b_r;
    if (p->nxe > 0) {
        if (cfg->n_rbin < 2 || cfg->n_rbin > 256) return fail("box plan: n_rbin must be in [2, 256]");
        if (!(cfg->rmax > cfg->rmin)) return fail("box plan: rmax must exceed rmin");
        p->nr = cfg->n_rbin;
        p->rbin.resize(p->nr);
        const double drb = (cfg->rmax - cfg->rmin) / (double)(p->nr - 1);
        for (int i = 0; i < p->nr; i++) p->rbin[i] = cfg->rmin + drb * (double)i;
        const double dr_bin = p->rbin[1] - p->rbin[0];
        p->rlower.resize(p->nr);
        p->rupper.resize(p->nr);
        double rumax = 0.0;
        for (int j = 0; j < p->nr; j++) {
            p->rlower[j] = (p->rbin[j] > dr_bin * 0.5) ? (p->rbin[j] - dr_bin * 0.5) : 0.0;
            p->rupper[j] = p->rbin[j] + dr_bin * 0.5;
            rumax = std::max(rumax, p->rupper[j]);
        }
        f2b_r.assign(FINE_NR, make_int2(-1, -2));
        for (int i = 0; i < FINE_NR; i++) {
            const double rt = (double)i * FINE_DR;
            int lo = -1, hi = -2;
            for (int j = 0; j < p->nr; j++)
                if ((rt > p->rlower[j]) && (rt <= p->rupper[j])) {
                    if (lo < 0) lo = j;
                    else if (j != hi + 1) return fail("box plan: non-contiguous r-bin overlap");
                    hi = j;
                }
            f2b_r[i] = make_int2(lo, hi);
        }
        // a separation reaches a bin only if its fine centre int(|r|/0.5 + 0.5)*0.5 <= rumax,
        // i.e. |r| < rumax + 0.25, so every |r_a| is below that too
        for (int a = 0; a < 3; a++) {
            const double dr = m.L[a] / (double)m.n[a];
            const long long s = (long long)std::floor((rumax + FINE_DR) / dr) + 1;
            p->rs[a] = (!cfg->prune || 2 * s + 1 >= m.n[a] || s >= m.n[a] / 2) ? -1 : (int)s;
        }
        p->xi_rows = (long long)support_count(p->rs[0], m.n[0]) * support_count(p->rs[1], m.n[1]);
    }
    // three-point function: multipole weights, r-bins with the 3PCF edges (field.hpp:1066-1095)
    std::vector<int2> zf2b;
    std::vector<double> zrb;
    if (cfg->interlace && (cfg->bk || cfg->zeta))
        return fail("box plan: interlacing applies to P_l and xi_l (the bispectrum and three-point function use the "
                    "reference's single grid)");
    if (cfg->zeta) {
        p->zeta = true;
        for (int a = 0; a < 3; a++) p->zl[a] = cfg->zeta_ell[a];
        if (p->zl[0] < 0 || p->zl[1] < 0 || p->zl[2] < 0 || p->zl[0] > 8 || p->zl[1] > 8 || p->zl[2] > 16)
            return fail("box plan: three-point multipole out of range");
        if (std::fabs(wigner_3j(p->zl[0], p->zl[1], p->zl[2], 0, 0, 0)) < 1.0e-10)
            return fail("box plan: this multipole combination is not allowed (wigner_3j(ell1, ell2, ELL, 0,0,0) == 0)");
        if (cfg->n_rbin < 2 || cfg->n_rbin > 64 || !(cfg->rmax > cfg->rmin))
            return fail("box plan: the three-point function needs n_rbin in [2, 64] and rmax > rmin");
        const double dr0 = m.L[0] / m.n[0], dr1 = m.L[1] / m.n[1], dr2 = m.L[2] / m.n[2];
        (void)dr0, (void)dr1, (void)dr2;
        p->z_gen = !(p->zl[0] == 0 && p->zl[1] == 0 && p->zl[2] == 0);
        p->z_odd = ((p->zl[0] + p->zl[1]) & 1) != 0;
        double wconst = 0.0;
        for (int m1 = -p->zl[0]; m1 <= p->zl[0]; m1++)
            for (int m2 = -p->zl[1]; m2 <= p->zl[1]; m2++) {
                double w = wigner_3j(p->zl[0], p->zl[1], p->zl[2], 0, 0, 0) *
                           wigner_3j(p->zl[0], p->zl[1], p->zl[2], m1, m2, 0);
                w *= (double)(2 * p->zl[2] + 1) * (double)(2 * p->zl[0] + 1) * (double)(2 * p->zl[1] + 1);
                if (std::fabs(w) < 1.0e-10) continue;
                p->zmlist.push_back(m1);
                p->zwlist.push_back(w);
                if (p->zl[0] == 0 && p->zl[1] == 0) wconst += w;
            }
        make_ang(p->zl[0], p->zl[1], p->zmlist, p->zwlist, wconst, p->zang);
        const int nr = cfg->n_rbin;
        zrb.resize(nr);
        const double drr = (cfg->rmax - cfg->rmin) / (double)(nr - 1);
        for (int i = 0; i < nr; i++) zrb[i] = cfg->rmin + drr * (double)i;  // tools.hpp:126-132
        std::vector<double> drb(nr - 1), lo(nr), up(nr);
        for (int j = 0; j < nr - 1; j++) drb[j] = zrb[j + 1] - zrb[j];
        double rumax = 0.0;
        for (int j = 0; j < nr; j++) {
            lo[j] = (j == 0) ? ((zrb[j] > drb[j] * 0.5) ? (zrb[j] - drb[j] * 0.5) : 0.0) : zrb[j] - drb[j - 1] * 0.5;
            up[j] = (j == nr - 1) ? (zrb[j] + drb[j - 1] * 0.5) : (zrb[j] + drb[j] * 0.5);
            rumax = std::max(rumax, up[j]);
        }
        const std::string ez = fine_to_coarse(lo, up, FINE_NZ, FINE_DR, zf2b);
        if (!ez.empty()) return fail("box plan: " + ez);
        for (int a = 0; a < 3; a++) {
            const double dr = m.L[a] / (double)m.n[a];
            const long long sr = (long long)std::floor((rumax + FINE_DR) / dr) + 1;
            p->zrs[a] = (!cfg->prune || 2 * sr + 1 >= m.n[a] || sr >= m.n[a] / 2) ? -1 : (int)sr;
        }
        p->z_rows = (long long)support_count(p->zrs[0], m.n[0]) * support_count(p->zrs[1], m.n[1]);
        if (p->z_gen) {
            for (int a = 0; a < nr; a++)
                for (int b = 0; b < nr; b++) p->zpairs.push_back(make_short2((short)a, (short)(nr + b)));
        } else {
            for (int a = 0; a < nr; a++)
                for (int b = a; b < nr; b++) p->zpairs.push_back(make_short2((short)a, (short)b));
        }
    }
    // device tables
    std::vector<double> tabs;
    host_axis_tables(m, tabs, p->tab_off);
    {
        // the x factors of XiSrc's pair (k, -k), 4 doubles per x index, after the axis tables
        const long long ntot = (long long)m.n[0] + m.n[1] + m.n[2];
        const double* iw0 = tabs.data() + 3 * ntot + p->tab_off[0];
        const double* sf0 = tabs.data() + 2 * ntot + p->tab_off[0];
        std::vector<double> xq(4 * (size_t)m.n[0]);
        for (int i = 0; i < m.n[0]; i++) {
            const int ii = i ? m.n[0] - i : 0;
            xq[4 * i] = iw0[i];
            xq[4 * i + 1] = iw0[ii];
            xq[4 * i + 2] = sf0[i] * iw0[i];
            xq[4 * i + 3] = sf0[ii] * iw0[ii];
        }
        tabs.insert(tabs.end(), xq.begin(), xq.end());
    }
    if (alloc(ctx, p, p->tab, sizeof(double) * tabs.size())) { plan_destroy(p); return nullptr; }
    if (alloc(ctx, p, p->f2b_d, sizeof(int2) * FINE_N)) { plan_destroy(p); return nullptr; }
    if (alloc(ctx, p, p->lu_d, sizeof(double) * 2 * p->nk)) { plan_destroy(p); return nullptr; }
    if (alloc(ctx, p, p->kb_d, sizeof(double) * p->nk)) { plan_destroy(p); return nullptr; }
    std::vector<double> lu(p->lower);
    lu.insert(lu.end(), p->upper.begin(), p->upper.end());
    hipError_t e = hipSuccess;
    e = (hipError_t)(e | hipMemcpy(p->tab.p, tabs.data(), sizeof(double) * tabs.size(), hipMemcpyHostToDevice));
    e = (hipError_t)(e | hipMemcpy(p->f2b_d.p, p->f2b.data(), sizeof(int2) * FINE_N, hipMemcpyHostToDevice));
    e = (hipError_t)(e | hipMemcpy(p->lu_d.p, lu.data(), sizeof(double) * lu.size(), hipMemcpyHostToDevice));
    e = (hipError_t)(e | hipMemcpy(p->kb_d.p, p->kbin.data(), sizeof(double) * p->nk, hipMemcpyHostToDevice));
    if (e != hipSuccess) return fail("box plan: table upload failed");
    int rc = cfg->precision == HM_FP64 ? plan_alloc<double>(ctx, p) : plan_alloc<float>(ctx, p);
    if (rc) { plan_destroy(p); return nullptr; }
    if (p->zeta) {
        if (alloc(ctx, p, p->zf2b_d, sizeof(int2) * FINE_NZ) || alloc(ctx, p, p->zrb_d, sizeof(double) * zrb.size()) ||
            hipMemcpy(p->zf2b_d.p, zf2b.data(), sizeof(int2) * FINE_NZ, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->zrb_d.p, zrb.data(), sizeof(double) * zrb.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->zpairs_d.p, p->zpairs.data(), sizeof(short2) * p->zpairs.size(), hipMemcpyHostToDevice) !=
                hipSuccess)
            return fail("box plan: three-point table upload failed");
        for (int s2 = 0; s2 < 2; s2++) {
            if (s2 == 1 && p->zl[1] == p->zl[0]) break;
            std::vector<double> sy, sc;
            jl_spline(p->zl[s2], sy, sc);
            if (alloc(ctx, p, p->zsy[s2], sizeof(double) * sy.size()) ||
                alloc(ctx, p, p->zsc[s2], sizeof(double) * sc.size()) ||
                hipMemcpy(p->zsy[s2].p, sy.data(), sizeof(double) * sy.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(p->zsc[s2].p, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice) != hipSuccess)
                return fail("box plan: spline upload failed");
        }
    }
    if (p->nxe > 0) {
        if (alloc(ctx, p, p->f2b_r_d, sizeof(int2) * FINE_NR)) { plan_destroy(p); return nullptr; }
        if (hipMemcpy(p->f2b_r_d.p, f2b_r.data(), sizeof(int2) * FINE_NR, hipMemcpyHostToDevice) != hipSuccess)
            return fail("box plan: r-bin table upload failed");
    }
    if (cfg->bk) {
        if (hipMemcpy(p->pairs_d.p, p->pairs.data(), sizeof(short2) * p->pairs.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail("box plan: pair table upload failed");
        // every mode of shell j has |k| <= upper[j] (field.hpp:371-373), so |n_a| <= upper[j] / dk_a; one
        // index of margin for the rounding of |k|
        const int nsh = p->bk_gen ? 2 * p->nk : p->nk;
        p->fsup.resize(nsh);
        p->fsz.resize(nsh);
        for (int f = 0; f < nsh; f++) {
            const int j = f % p->nk;
            int sv[3];
            for (int a = 0; a < 3; a++) {
                if (p->s[a] < 0) {
                    sv[a] = sup_eff(-1, p->M[a]);
                } else {
                    const double dka = 2.0 * M_PI / m.L[a];
                    const long long sj = (long long)std::floor(p->upper[j] / dka) + 1;
                    sv[a] = sup_eff((int)std::min<long long>(sj, p->s[a]), p->M[a]);
                }
            }
            p->fsup[f] = make_int4(sv[0], sv[1], sv[2], 0);
            p->fsz[f] = sv[2];
        }
        if (alloc(ctx, p, p->fsup_d, sizeof(int4) * nsh) || alloc(ctx, p, p->fsz_d, sizeof(int) * nsh) ||
            hipMemcpy(p->fsup_d.p, p->fsup.data(), sizeof(int4) * nsh, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->fsz_d.p, p->fsz.data(), sizeof(int) * nsh, hipMemcpyHostToDevice) != hipSuccess)
            return fail("box plan: shell support upload failed");
        if (p->bk_gen) {
            std::vector<short2> gp(p->pairs);
            for (auto& q : gp) q.y = (short)(q.y + p->nk);
            if (alloc(ctx, p, p->gpairs_d, sizeof(short2) * gp.size()) ||
                hipMemcpy(p->gpairs_d.p, gp.data(), sizeof(short2) * gp.size(), hipMemcpyHostToDevice) != hipSuccess)
                return fail("box plan: pair table upload failed");
        }
        const double dr = m.L[0] / m.n[0];
        const long long tot = (long long)p->nk * p->nkey;
        // j_l(k r) tables: the closed form for l = 0, else the spline of bessel.hpp:12-25
        auto jtable = [&](int l, DevBuf& out) -> int {
            if (l == 0)
                return ctx->launch("j0_table", [&] {
                    hipLaunchKernelGGL(k_j0_table, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                                       (const double*)p->kb_d.p, p->nk, p->nkey, dr, p->keys, (double*)out.p);
                });
            std::vector<double> sy, sc;
            jl_spline(l, sy, sc);
            DevBuf dy, dc;
            int rc = (dev_alloc(ctx, dy, sizeof(double) * sy.size()) || dev_alloc(ctx, dc, sizeof(double) * sc.size()))
                         ? -1
                         : 0;
            if (!rc && (hipMemcpy(dy.p, sy.data(), sizeof(double) * sy.size(), hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(dc.p, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice) != hipSuccess)) {
                ctx->set_error("box plan: spline upload failed");
                rc = -1;
            }
            if (!rc)
                rc = ctx->launch("jl_table", [&] {
                    hipLaunchKernelGGL(k_jl_table, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                                       (const double*)dy.p, (const double*)dc.p, (const double*)p->kb_d.p, p->nk,
                                       p->nkey, dr, p->keys, (double*)out.p);
                });
            if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = -1;
            dev_free(dy);
            dev_free(dc);
            return rc;
        };
        if (jtable(p->l1, p->jtab)) { plan_destroy(p); return nullptr; }
        if (p->bk_gen && p->l2 != p->l1 && jtable(p->l2, p->jtab2)) { plan_destroy(p); return nullptr; }
    }
    // shell mode counts on the N mesh (data independent)
    {
        CountArgs ca;
        for (int a = 0; a < 3; a++) {
            ca.n[a] = m.n[a];
            ca.s[a] = p->s[a];
            ca.c[a] = support_count(p->s[a], m.n[a]);
        }
        ca.total = (long long)ca.c[0] * ca.c[1] * ca.c[2];
        AxisTabs t = axis_tabs(p);
        for (int a = 0; a < 3; a++) ca.kv[a] = t.kv[a];
        ca.lower = (const double*)p->lu_d.p;
        ca.upper = ca.lower + p->nk;
        ca.nk = p->nk;
        if (fill_zero(ctx, p->shell_cnt.p, sizeof(unsigned long long) * p->nk)) { plan_destroy(p); return nullptr; }
        if (ctx->launch("shell_count", [&] {
                hipLaunchKernelGGL(k_shell_count, dim3((unsigned)((ca.total + 255) / 256)), dim3(256), 0, ctx->stream, ca,
                                   (unsigned long long*)p->shell_cnt.p);
            }))
        { plan_destroy(p); return nullptr; }
        p->shell_nmode.resize(p->nk);
        if (hipMemcpyAsync(p->shell_nmode.data(), p->shell_cnt.p, sizeof(long long) * p->nk, hipMemcpyDeviceToHost,
                           ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail("box plan: shell count readback failed");
    }
    return p;
}

void plan_destroy(BoxPlanImpl* p) {
    if (!p) return;
    DevBuf* bufs[] = {&p->tab, &p->f2b_d, &p->lu_d, &p->kb_d, &p->fixed, &p->dk, &p->pk_part, &p->pk_cnt, &p->pk_red,
                      &p->pk_cnt_red, &p->xi_half, &p->xi_real, &p->hist, &p->jtab, &p->pairs_d, &p->a_start, &p->a_rows, &p->a_idx, &p->xi_R,
                      &p->sn_pairs, &p->fields, &p->gram_part, &p->gram_red, &p->shell_cnt, &p->sid,
                      &p->f2b_r_d, &p->xi_r, &p->xi_part, &p->xi_cnt, &p->xi_red, &p->xi_cnt_red, &p->rpos, &p->rw,
                      &p->jtab2, &p->gpairs_d, &p->hist2, &p->sn_pairs2, &p->snp_part, &p->fsup_d, &p->fsz_d, &p->dk2, &p->pkx_part,
                      &p->pkx_part_im, &p->pkx_red, &p->dki, &p->spos, &p->zf2b_d, &p->zrb_d, &p->zsy[0],
                      &p->zsy[1], &p->zsc[0], &p->zsc[1], &p->zfields, &p->zr, &p->zpart, &p->zcnt, &p->zred,
                      &p->zcnt_red, &p->zpairs_d, &p->qo_d, &p->cf1, &p->cf2, &p->cx, &p->fixed_im,
                      &p->ypart, &p->yred, &p->yh};
    for (DevBuf* b : bufs) dev_free(*b);
    sort_work_free(p->sort);
    delete p;
}

static ColGeom geom(long long OS, long long PS, int n_outer, int outer_full, int outer_s, int n_inner, int inner_full,
                    int inner_s, int point_s) {
    ColGeom g;
    g.OS = OS;
    g.PS = PS;
    g.n_outer = n_outer;
    g.outer_full = outer_full;
    g.outer_s = outer_s;
    g.n_inner = n_inner;
    g.inner_full = inner_full;
    g.inner_s = inner_s;
    g.point_s = point_s;
    return g;
}

// x pass over nf fields (stride FS) of support sup: lines (y, z) in the
// support, points x in the support.
static ColGeom tri_geom_x(const int M[3], const int sup[3], long long FS, int nf, const int4* fsup = nullptr,
                           const int4* fsup_host = nullptr) {
    ColGeom g = geom(M[2], (long long)M[1] * M[2], support_count(sup[1], M[1]), M[1], sup[1],
                     support_count(sup[2], M[2]), M[2], sup[2], sup[0]);
    g.FS = FS;
    g.nf = nf;
    if (fsup) {
        g.fsup = fsup;
        g.fsup_host = fsup_host;
        g.role_outer = 1;
        g.role_inner = 2;
        g.role_point = 0;
    }
    return g;
}
// y pass: lines (x all, z in the support), points y in the support.
static ColGeom tri_geom_y(const int M[3], const int sup[3], long long FS, int nf, const int4* fsup = nullptr,
                           const int4* fsup_host = nullptr) {
    ColGeom g = geom((long long)M[1] * M[2], M[2], M[0], M[0], -1, support_count(sup[2], M[2]), M[2], sup[2], sup[1]);
    g.FS = FS;
    g.nf = nf;
    if (fsup) {
        g.fsup = fsup;
        g.fsup_host = fsup_host;
        g.role_inner = 2;
        g.role_point = 1;
    }
    return g;
}

// Triple sums from the reduced partial columns to pair order (a <= b).
static void gram_to_pairs(const BoxPlanImpl* p, const std::vector<double>& raw, std::vector<double>& out) {
    if (!p->fused) {
        out = raw;
        return;
    }
    std::vector<int> pidx((size_t)p->nk * p->nk, -1);
    for (size_t q = 0; q < p->pairs.size(); q++) pidx[(size_t)p->pairs[q].x * p->nk + p->pairs[q].y] = (int)q;
    int blk = 0;
    for (int bi = 0; bi < p->ngrp; bi++)
        for (int bj = bi; bj < p->ngrp; bj++, blk++)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    const int a = 4 * bi + i, b = 4 * bj + j;
                    if (a >= p->nk || b >= p->nk || a > b) continue;
                    const int q = pidx[(size_t)a * p->nk + b];
                    const int o = blk * 16 + i * 4 + j;
                    out[2 * (size_t)q] = raw[2 * (size_t)o];
                    out[2 * (size_t)q + 1] = raw[2 * (size_t)o + 1];
                }
}

// Inverse transform of one triple-product field on the M mesh:
// x pass from the spectrum source, then y columns, then z rows.
template <typename T>
static int field_transform(hm_ctx* ctx, BoxPlanImpl* p, const ShellSrc<T>& src, C<T>* out, const int sup[3],
                           bool z_pass) {
    const int* M = p->M;
    const long long M12 = (long long)M[1] * M[2];
    ColGeom ga = geom(M[2], M12, support_count(sup[1], M[1]), M[1], sup[1] >= 0 && 2 * sup[1] + 1 < M[1] ? sup[1] : -1,
                      support_count(sup[2], M[2]), M[2], sup[2] >= 0 && 2 * sup[2] + 1 < M[2] ? sup[2] : -1,
                      sup[0] >= 0 && 2 * sup[0] + 1 < M[0] ? sup[0] : -1);
    if (launch_cols<T, +1>(ctx, "tri_pass_x", M[0], src, out, ga)) return -1;
    ColGeom gb = geom(M12, M[2], M[0], M[0], -1, support_count(sup[2], M[2]), M[2],
                      sup[2] >= 0 && 2 * sup[2] + 1 < M[2] ? sup[2] : -1, sup[1] >= 0 && 2 * sup[1] + 1 < M[1] ? sup[1] : -1);
    if (fft_cols<T>(ctx, "tri_pass_y", M[1], +1, out, out, gb)) return -1;
    if (!z_pass) return 0;
    RowGeom rg{(long long)M[0] * M[1], M[2], M[2], sup[2] >= 0 && 2 * sup[2] + 1 < M[2] ? sup[2] : -1};
    return fft_rows<T>(ctx, "tri_pass_z", M[2], +1, out, out, rg);
}

// Interlacing: the positions shifted by h (half a cell per axis), wrapped
// into [0, L), and the combination d = (d_a + exp(i k.h) d_b) / 2 on the
// half spectrum, k with the Nyquist index negative.
__global__ __launch_bounds__(256) void k_shift_wrap(const double* __restrict__ pos, long long n, double h0, double h1,
                                                   double h2, double L0, double L1, double L2,
                                                   double* __restrict__ out) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const double h[3] = {h0, h1, h2}, L[3] = {L0, L1, L2};
    for (int a = 0; a < 3; a++) {
        double x = pos[3 * q + a] + h[a];
        if (x >= L[a]) x -= L[a];
        out[3 * q + a] = x;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_interlace(C<T>* __restrict__ d, const C<T>* __restrict__ db, int n1, int nzh,
                                                  long long nh, AxisTabs t, double h0, double h1, double h2) {
#pragma clang fp contract(off)
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= nh) return;
    const int k = (int)(e % nzh);
    const long long r = e / nzh;
    const int j = (int)(r % n1), i = (int)(r / n1);
    const double ph = (t.kv[0][i] * h0 + t.kv[1][j] * h1) + t.kv[2][k] * h2;
    double sn, cs;
    sincos(ph, &sn, &cs);
    const C<T> a = d[e], b = db[e];
    const double br = cs * (double)b.x - sn * (double)b.y, bi = cs * (double)b.y + sn * (double)b.x;
    d[e] = mk<T>((T)(0.5 * ((double)a.x + br)), (T)(0.5 * ((double)a.y + bi)));
}

// D then R into one array with weights wD^pw (1 when wD is null) and
// cR * wR^pw: the fields D - alpha R (pw = 1, cR = -alpha) and
// D + alpha^2 R (pw = 2, cR = alpha^2) of the reconstructed box and the survey.
__global__ __launch_bounds__(256) void k_pack_pow(const double* __restrict__ pD, const double* __restrict__ wD,
                                                 long long nD, const double* __restrict__ pR,
                                                 const double* __restrict__ wR, long long nR, double cR, int pw,
                                                 double* __restrict__ pos, double* __restrict__ w) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= nD + nR) return;
    const bool d = q < nD;
    const double* src = d ? pD + 3 * q : pR + 3 * (q - nD);
    pos[3 * q] = src[0];
    pos[3 * q + 1] = src[1];
    pos[3 * q + 2] = src[2];
    const double* wt = d ? wD : wR;
    double v = wt ? wt[d ? q : q - nD] : 1.0;
    if (pw == 2) v = v * v;
    if (!d) v = cR * v;
    w[2 * q] = v;
    w[2 * q + 1] = 0.0;
}

static int pack_pow(hm_ctx* ctx, const double* pD, const double* wD, long long nD, const double* pR, const double* wR,
                    long long nR, double cR, int pw, double* pos, double* w) {
    if (nD + nR <= 0) return 0;
    return ctx->launch("pack_two", [&] {
        hipLaunchKernelGGL(k_pack_pow, dim3((unsigned)((nD + nR + 255) / 256)), dim3(256), 0, ctx->stream, pD, wD, nD, pR,
                           wR, nR, cR, pw, pos, w);
    });
}

// Survey B with ELL > 0: the full spectrum of the field with weights
// w^pw Y(L, M)(los) (conjugated with cj) of D and cR times those of R, in out.
template <typename T>
// reuse: p->sort already holds the cell sort of these packed positions (from
// this execution's density assignment), so only the weights move.
static int svg_field(hm_ctx* ctx, BoxPlanImpl* p, const AssignGeom& ag, const BoxWeighted* wt, long long nD,
                     long long nR, int M, int pw, int cj, double cR, double abs, C<T>* out, bool reuse) {
    const Mesh& m = p->mesh;
    const long long ntot = nD + nR, ng = m.ng();
    if (!p->fixed_im.p && alloc(ctx, p, p->fixed_im, sizeof(long long) * (size_t)ng)) return -1;
    if (ctx->launch("svg_weights", [&] {
            hipLaunchKernelGGL(k_ylm_weights, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, ctx->stream, wt->losD,
                               wt->wD, nD, wt->losR, wt->wR, nR, wt->L, M, pw, cj, cR, (double*)p->rw.p);
        }))
        return -1;
    const double scale = fixed_scale_for(1.5 * abs);  // |re| + |im| <= sqrt(2) |w|: |Y_red| <= 1
    if (assign_sorted(ctx, ag, (const double*)p->rpos.p, (const double*)p->rw.p, ntot, scale, (long long*)p->fixed.p,
                      (long long*)p->fixed_im.p, p->sort, reuse))
        return -1;
    const double dV = m.volume() / (double)ng;
    const unsigned gb = (unsigned)std::max<long long>(1, std::min<long long>((ng + 255) / 256, 16384));
    if (ctx->launch("svg_to_complex", [&] {
            hipLaunchKernelGGL(k_fixed_to_complex<T>, dim3(gb), dim3(256), 0, ctx->stream, (const long long*)p->fixed.p,
                               (const long long*)p->fixed_im.p, ng, 1.0 / scale, 1.0 / dV, dV, out);
        }))
        return -1;
    const int n3[3] = {m.n[0], m.n[1], m.n[2]};
    return fft3d_c2c<T>(ctx, out, n3, -1);
}

// The per-catalogue scalars of an execution into device memory (batched mode).
__global__ void k_set_dyn(double* __restrict__ dyn, double inv_scale, double nbar, double sn2, double sn3, double dc,
                          double n) {
    if (threadIdx.x == 0) {
        dyn[DYN_INV_SCALE] = inv_scale;
        dyn[DYN_NBAR] = nbar;
        dyn[DYN_SN2] = sn2;
        dyn[DYN_SN3] = sn3;
        dyn[DYN_DC] = dc;
        dyn[DYN_N] = n;
    }
}

template <typename T>
// d_pos_r / nr / alpha: the reconstructed box (main_box.cpp:199-228): the
// density is D - alpha R with unit weights (field.hpp:528-558) instead of
// rho - nbar, the two-point shot noise N_D + alpha^2 N_R (field.hpp:1190-1197).
// wt: survey-geometry catalogues for B and zeta with ELL = 0 (bispec.hpp:4-874):
// the fields w (D - alpha R) and w^2 (D + alpha^2 R), or R alone for the
// window, with the shot noise and normalisation of the survey estimators.
static int plan_execute(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos, long long np, hm_box_result* res,
                        const double* d_pos_r = nullptr, long long nr = 0, double alpha = 0.0,
                        const BoxWeighted* wt = nullptr, const ExecOpts* xo = nullptr) {
    const Mesh& m = p->mesh;
    ctx->wide_mask = p->tune_wide;
    const double* dyn = xo ? xo->dyn : nullptr;
    const bool run_assign = !xo || xo->assign;
    const AxisTabs tabs = axis_tabs(p);
    const int wp = wpow_of(m.assign);
    if (np <= 0) {
        ctx->set_error("box execute: no particles");
        return -1;
    }
    // 1. assignment (field.hpp:145-184, 505-523)
    AssignGeom ag;
    for (int a = 0; a < 3; a++) {
        ag.n[a] = m.n[a];
        ag.L[a] = m.L[a];
    }
    ag.ng = m.ng();
    ag.assign = m.assign;
    const bool weighted = wt != nullptr;
    const bool recon = weighted || nr > 0;  // two catalogues, no mean subtraction, cross-spectrum shot noise
    const double scale = fixed_scale_for(weighted ? wt->abs1 : (recon ? (double)np + std::fabs(alpha) * (double)nr : (double)np));
    if (run_assign) {
        const size_t before = p->sort.bytes() + p->rpos.bytes + p->rw.bytes;
        if (sort_work_alloc(ctx, p->sort, ag, np + nr, recon)) return -1;
        if (recon && (dev_alloc(ctx, p->rpos, sizeof(double) * 3 * (size_t)(np + nr)) ||
                      dev_alloc(ctx, p->rw, sizeof(double) * 2 * (size_t)(np + nr))))
            return -1;
        p->bytes += p->sort.bytes() + p->rpos.bytes + p->rw.bytes - before;
    }
    if (!run_assign) {
    } else if (recon) {
        if (pack_pow(ctx, d_pos, weighted ? wt->wD : nullptr, np, d_pos_r, weighted ? wt->wR : nullptr, nr,
                     weighted ? wt->c1 : -alpha, 1, (double*)p->rpos.p, (double*)p->rw.p))
            return -1;
        if (assign_sorted(ctx, ag, (const double*)p->rpos.p, (const double*)p->rw.p, np + nr, scale,
                          (long long*)p->fixed.p, nullptr, p->sort))
            return -1;
    } else if (assign_sorted(ctx, ag, d_pos, nullptr, np, scale, (long long*)p->fixed.p, nullptr, p->sort)) {
        return -1;
    }
    // two-point shot noise: N (field.hpp:523) or N_D + alpha^2 N_R (field.hpp:1190-1197); that of the
    // bispectrum and three-point function: N or N_D - alpha^3 N_R (field.hpp:1242-1250)
    const double sn2 = recon ? (double)np + pow(alpha, 2) * (double)nr : (double)np;
    const double sn3 = weighted ? wt->sn3 : (recon ? (double)np - pow(alpha, 3) * (double)nr : (double)np);
    if (dyn && run_assign &&
        ctx->launch("set_dyn", [&] {
            hipLaunchKernelGGL(k_set_dyn, dim3(1), dim3(64), 0, ctx->stream, (double*)dyn, 1.0 / scale,
                               recon ? 0.0 : (double)np / m.volume(), sn2, sn3, recon ? 0.0 : (double)np, (double)np);
        }))
        return -1;
    if (xo && !xo->mesh) return 0;
    // 2. forward r2c with the fluctuation (rho - N/V) dV formed on load
    const double dV = m.volume() / (double)m.ng();
    FixedRealSrc<T> fsrc{(const long long*)p->fixed.p, m.n[2] / 2, 1.0 / scale, 1.0 / dV,
                         recon ? 0.0 : (double)np / m.volume(), dV};
    fsrc.dyn = dyn;
    C<T>* dk = (C<T>*)p->dk.p;
    const long long nzh = m.nzh();
    if (launch_r2c<T>(ctx, "fft_r2c_z", m.n[2] / 2, fsrc, dk, (long long)m.n[0] * m.n[1])) return -1;
    if (fft_cols<T>(ctx, "fft_r2c_y", m.n[1], -1, dk, dk,
                    geom((long long)m.n[1] * nzh, nzh, m.n[0], m.n[0], -1, (int)nzh, (int)nzh, -1, -1)))
        return -1;
    // the source of the shot-noise xi of B (field.hpp:1116-1160): (|d|^2 - N S) / W^2 / V of each mode pair
    auto make_xs = [&](const C<T>* d2) {
        XiSrc<T> xs;
        xs.dk = dk;
        for (int a = 0; a < 3; a++) xs.n[a] = m.n[a];
        xs.nzh = (int)nzh;
        xs.t = tabs;
        xs.xq = xi_xq(p);
        xs.wp = wp;
        xs.sn = sn3;
        xs.inv_v = 1.0 / m.volume();
        xs.dc = recon ? 0.0 : (double)np;
        xs.sn_ptr = dyn ? dyn + DYN_SN3 : nullptr;
        xs.dc_ptr = dyn ? dyn + DYN_DC : nullptr;
        xs.dk2 = d2;
        return xs;
    };
    // with one parity of xi and no second spectrum, xi's inverse x pass rides on the forward x pass (the
    // spectrum is read once for both); its planes k_z = 0, n_z/2 follow in step 4.  Not with xi_l, whose
    // transform (step 3b) works in the same buffer before step 4 reads it.
    const bool fuse_xi = p->cfg.bk && !(weighted && wt->L > 0) && !recon && p->cfg.interlace == 0 && !p->sn_odd &&
                         !p->odd_mesh && (m.n[2] % 2) == 0 && p->cfg.n_xi_ells == 0;
    int fused_xi = 0;
    if (fuse_xi) {
        fused_xi = launch_fwd_x_xi<T>(ctx, m.n[0], dk, (C<T>*)p->xi_half.p, make_xs(nullptr),
                                      (long long)m.n[1] * nzh, (int)(m.n[1] * nzh));
        if (fused_xi < 0) return -1;
    }
    if (!fused_xi && fft_cols<T>(ctx, "fft_r2c_x", m.n[0], -1, dk, dk,
                                 geom(0, (long long)m.n[1] * nzh, 1, 1, -1, (int)(m.n[1] * nzh), (int)(m.n[1] * nzh),
                                      -1, -1)))
        return -1;
    // 2''. interlacing: the half-cell-shifted grid, transformed and combined (no counterpart in the reference)
    const bool inter = p->cfg.interlace != 0;
    if (inter) {
        if (recon) {
            ctx->set_error("box execute: interlacing is not available for two-catalogue fields");
            return -1;
        }
        if ((!p->dki.p && alloc(ctx, p, p->dki, sizeof(C<T>) * (size_t)m.nh())) ||
            dev_alloc(ctx, p->spos, sizeof(double) * 3 * (size_t)np))
            return -1;
        const double h[3] = {0.5 * m.L[0] / m.n[0], 0.5 * m.L[1] / m.n[1], 0.5 * m.L[2] / m.n[2]};
        if (ctx->launch("interlace_shift", [&] {
                hipLaunchKernelGGL(k_shift_wrap, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, d_pos,
                                   np, h[0], h[1], h[2], m.L[0], m.L[1], m.L[2], (double*)p->spos.p);
            }))
            return -1;
        if (assign_sorted(ctx, ag, (const double*)p->spos.p, nullptr, np, scale, (long long*)p->fixed.p, nullptr,
                          p->sort))
            return -1;
        C<T>* dki = (C<T>*)p->dki.p;
        if (launch_r2c<T>(ctx, "fft_r2c_z", m.n[2] / 2, fsrc, dki, (long long)m.n[0] * m.n[1])) return -1;
        if (fft_cols<T>(ctx, "fft_r2c_y", m.n[1], -1, dki, dki,
                        geom((long long)m.n[1] * nzh, nzh, m.n[0], m.n[0], -1, (int)nzh, (int)nzh, -1, -1)))
            return -1;
        if (fft_cols<T>(ctx, "fft_r2c_x", m.n[0], -1, dki, dki,
                        geom(0, (long long)m.n[1] * nzh, 1, 1, -1, (int)(m.n[1] * nzh), (int)(m.n[1] * nzh), -1,
                             -1)))
            return -1;
        const long long nh = m.nh();
        if (ctx->launch("interlace", [&] {
                hipLaunchKernelGGL(k_interlace<T>, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, ctx->stream, dk,
                                   dki, m.n[1], (int)nzh, nh, tabs, h[0], h[1], h[2]);
            }))
            return -1;
    }
    // 2'. reconstructed box with B or zeta: the shot-noise density D + alpha^2 R
    // (calcNormalDensityForBispectrumShotnoiseForBOXForReconstruction, field.hpp:574-605), transformed
    C<T>* dk2 = nullptr;
    if (recon && (p->cfg.bk || p->zeta)) {
        if (!p->dk2.p && alloc(ctx, p, p->dk2, sizeof(C<T>) * (size_t)m.nh())) return -1;
        dk2 = (C<T>*)p->dk2.p;
        if (pack_pow(ctx, d_pos, weighted ? wt->wD : nullptr, np, d_pos_r, weighted ? wt->wR : nullptr, nr,
                     weighted ? wt->c2 : alpha * alpha, 2, (double*)p->rpos.p, (double*)p->rw.p))
            return -1;
        const double scale2 = fixed_scale_for(weighted ? wt->abs2 : (double)np + alpha * alpha * (double)nr);
        // the same packed positions as the density above, sorted on p->sort in this execution (unless
        // the assignment was skipped or the interlacing pass sorted its shifted copy since): only the
        // weights move
        if (assign_sorted(ctx, ag, (const double*)p->rpos.p, (const double*)p->rw.p, np + nr, scale2,
                          (long long*)p->fixed.p, nullptr, p->sort, run_assign && !inter))
            return -1;
        FixedRealSrc<T> fsrc2{(const long long*)p->fixed.p, m.n[2] / 2, 1.0 / scale2, 1.0 / dV, 0.0, dV};
        if (launch_r2c<T>(ctx, "fft_r2c_z", m.n[2] / 2, fsrc2, dk2, (long long)m.n[0] * m.n[1])) return -1;
        if (fft_cols<T>(ctx, "fft_r2c_y", m.n[1], -1, dk2, dk2,
                        geom((long long)m.n[1] * nzh, nzh, m.n[0], m.n[0], -1, (int)nzh, (int)nzh, -1, -1)))
            return -1;
        if (fft_cols<T>(ctx, "fft_r2c_x", m.n[0], -1, dk2, dk2,
                        geom(0, (long long)m.n[1] * nzh, 1, 1, -1, (int)(m.n[1] * nzh), (int)(m.n[1] * nzh), -1,
                             -1)))
            return -1;
    }
    // 3. P_l binning (field.hpp:765-852)
    BinArgs ba;
    for (int a = 0; a < 3; a++) ba.n[a] = m.n[a];
    ba.nzh = (int)nzh;
    ba.s01 = p->s01;
    ba.nrow_y = p->nrow_y;
    ba.nrows = p->nrows;
    ba.nzmax = p->nzmax;
    ba.t = tabs;
    ba.wp = wp;
    ba.f2b = (const int2*)p->f2b_d.p;
    ba.nk = p->nk;
    ba.ne = std::max(p->ne, 1);
    for (int e = 0; e < BIN_MAX_ELLS; e++) ba.ells[e] = (e < p->ne) ? p->bin_ells[e] : 0;
    ba.sn = sn2;
    ba.sn_ptr = dyn ? dyn + DYN_SN2 : nullptr;
    ba.sn_w2 = inter ? 1 : 0;
    const int ncols = p->nk * ba.ne;
    const bool do_pk = p->ne > 0;
    if (do_pk) {
    ba.kc = p->bin_kc;
    ba.kcl = (p->nzmax + 1 + p->bin_kc - 1) / p->bin_kc;
    const long long prow = p->nrows * p->bin_kc;
    ba.cm = prow;
    if (fill_zero(ctx, p->pk_part.p, sizeof(double) * (size_t)ncols * prow)) return -1;
    if (fill_zero(ctx, p->pk_cnt.p, sizeof(long long) * (size_t)p->nk * prow)) return -1;
    if (ctx->launch("bin_pk", [&] {
            const int bb = p->tune_bin;
            hipLaunchKernelGGL(k_bin_box<T>, dim3((unsigned)((prow + bb - 1) / bb)), dim3(bb), 0, ctx->stream, dk, ba,
                               (double*)p->pk_part.p, (long long*)p->pk_cnt.p);
        }))
        return -1;
    if (reduce_cols_cm<double>(ctx, "bin_reduce", (const double*)p->pk_part.p, prow, ncols, (double*)p->pk_red.p))
        return -1;
    if (reduce_cols_cm<long long>(ctx, "bin_reduce", (const long long*)p->pk_cnt.p, prow, p->nk,
                                  (long long*)p->pk_cnt_red.p))
        return -1;
    ba.kc = 1;
    ba.kcl = 0;
    ba.cm = 0;
    }
    // 3'. reconstructed box with B: the two-point shot-noise terms bin the cross spectrum of D - alpha R
    // and D + alpha^2 R with SN N_D - alpha^3 N_R (bispec.hpp:1455-1508), complex
    const bool cross_pk = recon && p->cfg.bk && do_pk;
    if (cross_pk) {
        const size_t nb = sizeof(double) * (size_t)ncols * p->nrows;
        if (!p->pkx_part.p && (alloc(ctx, p, p->pkx_part, nb) || alloc(ctx, p, p->pkx_part_im, nb) ||
                               alloc(ctx, p, p->pkx_red, sizeof(double) * 2 * (size_t)ncols)))
            return -1;
        BinArgs bx = ba;
        bx.sn = sn3;
        bx.sn_ptr = dyn ? dyn + DYN_SN3 : nullptr;
        if (fill_zero(ctx, p->pkx_part.p, nb) || fill_zero(ctx, p->pkx_part_im.p, nb) ||
            fill_zero(ctx, p->pk_cnt.p, sizeof(long long) * (size_t)p->nk * p->nrows))
            return -1;
        if (ctx->launch("bin_pk_cross", [&] {
                hipLaunchKernelGGL(k_bin_box<T>, dim3((unsigned)((p->nrows + 255) / 256)), dim3(256), 0, ctx->stream,
                                   dk, bx, (double*)p->pkx_part.p, (long long*)p->pk_cnt.p, dk2,
                                   (double*)p->pkx_part_im.p);
            }))
            return -1;
        if (reduce_cols<double>(ctx, "bin_reduce", (const double*)p->pkx_part.p, p->nrows, ncols,
                                (double*)p->pkx_red.p) ||
            reduce_cols<double>(ctx, "bin_reduce", (const double*)p->pkx_part_im.p, p->nrows, ncols,
                                (double*)p->pkx_red.p + ncols))
            return -1;
    }
    // 3b. xi_l (calcCorrelationFunction, field.hpp:854-977): the even part of
    // (|d|^2 - N S)/W^2/V inverse-transformed (its real part, which is all the
    // reference writes, powerspec.hpp:385), then binned over the separations
    // that reach an r-bin
    if (p->nxe > 0) {
        XiSrc<T> xs;
        xs.dk = dk;
        for (int a = 0; a < 3; a++) xs.n[a] = m.n[a];
        xs.nzh = (int)nzh;
        xs.t = tabs;
        xs.xq = xi_xq(p);
        xs.wp = wp;
        xs.sn = sn2;
        xs.sn_ptr = dyn ? dyn + DYN_SN2 : nullptr;
        xs.inv_v = 1.0 / m.volume();
        xs.dc = 0.0;
        xs.sn_w2 = inter ? 1 : 0;
        C<T>* xh = (C<T>*)p->xi_half.p;
        const int cx = support_count(p->rs[0], m.n[0]), cy = support_count(p->rs[1], m.n[1]);
        if (launch_cols<T, +1>(ctx, "xi2_pass_x", m.n[0], xs, xh,
                               geom(0, (long long)m.n[1] * nzh, 1, 1, -1, (int)(m.n[1] * nzh), (int)(m.n[1] * nzh), -1, -1)))
            return -1;
        if (fft_cols<T>(ctx, "xi2_pass_y", m.n[1], +1, xh, xh,
                        geom((long long)m.n[1] * nzh, nzh, cx, m.n[0], p->rs[0], (int)nzh, (int)nzh, -1, -1)))
            return -1;
        SupHalfSrc<T> hs{xh, (int)nzh, m.n[0], m.n[1], p->rs[0], p->rs[1], cy};
        T* xr = (T*)p->xi_r.p;
        if (launch_c2r<T>(ctx, "xi2_pass_z", m.n[2] / 2, hs, xr, p->xi_rows, -1)) return -1;
        XiBinArgs xa;
        for (int a = 0; a < 3; a++) {
            xa.n[a] = m.n[a];
            xa.s[a] = p->rs[a];
            xa.c[a] = support_count(p->rs[a], m.n[a]);
            xa.dr[a] = m.L[a] / (double)m.n[a];
        }
        xa.nrows = p->xi_rows;
        xa.f2b = (const int2*)p->f2b_r_d.p;
        xa.nr = p->nr;
        xa.ne = p->nxe;
        for (int e = 0; e < HM_MAX_ELLS; e++) xa.ells[e] = (e < p->nxe) ? p->cfg.xi_ells[e] : 0;
        const int xcols = p->nr * p->nxe;
        if (fill_zero(ctx, p->xi_part.p, sizeof(double) * (size_t)xcols * p->xi_rows)) return -1;
        if (fill_zero(ctx, p->xi_cnt.p, sizeof(long long) * (size_t)p->nr * p->xi_rows)) return -1;
        if (ctx->launch("bin_xi", [&] {
                hipLaunchKernelGGL(k_bin_xi<T>, dim3((unsigned)((p->xi_rows + 255) / 256)), dim3(256), 0, ctx->stream, xr,
                                   xa, (double*)p->xi_part.p, (long long*)p->xi_cnt.p);
            }))
            return -1;
        if (reduce_cols<double>(ctx, "bin_reduce", (const double*)p->xi_part.p, p->xi_rows, xcols, (double*)p->xi_red.p))
            return -1;
        if (reduce_cols<long long>(ctx, "bin_reduce", (const long long*)p->xi_cnt.p, p->xi_rows, p->nr,
                                   (long long*)p->xi_cnt_red.p))
            return -1;
    }
    const int npairs = (int)p->pairs.size();
    std::vector<double> gacc;  // weighted cross sums of a general multipole, per ordered pair
    // survey B with ELL > 0 (bispec.hpp:4-353): per M the Y_LM-weighted field, its ijk shot noise, the two-point
    // shot-noise terms and the triple sums; the raw sums land in sv_row / sv_col (per bin, complex, not yet
    // divided by the mode counts) and sv_ijk (per pair, complex, before i^(l1 + l2))
    const bool svg = weighted && wt->L > 0;
    std::vector<double> sv_row, sv_col, sv_ijk;
    auto svg_stage = [&](C<T>* F3, TriStage<T> st, auto&& f3_passes, auto&& gram_combo) -> int {
        const int L = wt->L, nk = p->nk;
        const long long ng = m.ng();
        if ((!p->cf1.p && alloc(ctx, p, p->cf1, sizeof(C<T>) * (size_t)ng)) ||
            (!p->cx.p && alloc(ctx, p, p->cx, sizeof(C<T>) * (size_t)ng)) ||
            (!p->yh.p && alloc(ctx, p, p->yh, sizeof(double) * 2 * (size_t)p->nkey)) ||
            (!p->ypart.p && alloc(ctx, p, p->ypart, sizeof(double) * 2 * (size_t)nk * (size_t)p->nrows)) ||
            (!p->yred.p && alloc(ctx, p, p->yred, sizeof(double) * 2 * (size_t)nk)))
            return -1;
        if ((p->l1 == 0 || p->l2 == 0) && !p->cf2.p && alloc(ctx, p, p->cf2, sizeof(C<T>) * (size_t)ng)) return -1;
        C<T>* cf1 = (C<T>*)p->cf1.p;
        C<T>* cf2 = (C<T>*)p->cf2.p;
        C<T>* cx = (C<T>*)p->cx.p;
        double* Hre = (double*)p->yh.p;
        double* Him = Hre + p->nkey;
        sv_row.assign(2 * (size_t)nk, 0.0);
        sv_col.assign(2 * (size_t)nk, 0.0);
        sv_ijk.assign(2 * (size_t)npairs, 0.0);
        const int n3[3] = {m.n[0], m.n[1], m.n[2]};
        const unsigned gb = (unsigned)std::max<long long>(1, std::min<long long>((ng + 255) / 256, 16384));
        // P_{l m}(k) of d_00 conj N_LM into acc (bins, complex), weight w (field.hpp:765-852)
        auto ybin = [&](int l, int mm, double w, double snr, double sni, std::vector<double>& acc) -> int {
            YBinArgs ya;
            for (int a = 0; a < 3; a++) ya.n[a] = m.n[a];
            ya.nzh = (int)nzh;
            ya.s01 = p->s01;
            ya.nrow_y = p->nrow_y;
            ya.nzmax = p->nzmax;
            ya.nrows = p->nrows;
            ya.t = tabs;
            ya.wp = wp;
            ya.f2b = (const int2*)p->f2b_d.p;
            ya.nk = nk;
            ya.l = l;
            ya.m = mm;
            ya.w = w;
            ya.sn_re = snr;
            ya.sn_im = sni;
            if (fill_zero(ctx, p->ypart.p, sizeof(double) * 2 * (size_t)nk * (size_t)p->nrows)) return -1;
            if (ctx->launch("svg_bin", [&] {
                    hipLaunchKernelGGL(k_bin_ycross<T>, dim3((unsigned)((p->nrows + 255) / 256)), dim3(256), 0,
                                       ctx->stream, dk, cf2, ya, (double*)p->ypart.p, (long long*)nullptr);
                }))
                return -1;
            if (reduce_cols<double>(ctx, "svg_bin_reduce", (const double*)p->ypart.p, p->nrows, 2 * nk,
                                    (double*)p->yred.p))
                return -1;
            std::vector<double> h(2 * (size_t)nk);
            HM_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), p->yred.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                                             ctx->stream));
            HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
            for (size_t e = 0; e < h.size(); e++) acc[e] += h[e];
            return 0;
        };
        for (int M = -L; M <= L; M++) {
            YHistArgs ha;
            ha.l1 = p->l1;
            ha.l2 = p->l2;
            ha.nc = 0;
            for (int m1 = -p->l1; m1 <= p->l1; m1++) {
                const int m2 = -M - m1;
                if (m2 < -p->l2 || m2 > p->l2) continue;
                double w = wigner_3j(p->l1, p->l2, L, 0, 0, 0) * wigner_3j(p->l1, p->l2, L, m1, m2, M);
                w *= (double)(2 * L + 1) * (double)(2 * p->l1 + 1) * (double)(2 * p->l2 + 1);
                if (std::fabs(w) < 1.0e-10) continue;
                if (ha.nc >= YH_MAXC) {
                    ctx->set_error("survey bispectrum: too many (m1, m2) terms");
                    return -1;
                }
                ha.m1[ha.nc] = m1;
                ha.m2[ha.nc] = m2;
                ha.w[ha.nc] = w;
                ha.nc++;
            }
            if (ha.nc == 0) continue;
            const double snr = wt->snLM[2 * (M + L)], sni = wt->snLM[2 * (M + L) + 1];
            // the Y_LM-weighted fluctuation w (D - alpha R) (field.hpp:186-228), transformed
            if (svg_field<T>(ctx, p, ag, wt, np, nr, M, 1, 0, wt->c1, wt->abs1, cf1, run_assign && !inter)) return -1;
            // ijk: xi_LM = IFFT[(dn_LM conj N_00 - SN_LM S)/W^2/V], histogram of A_M xi by |n|^2, Bessel sweep
            YXiArgs xa;
            for (int a = 0; a < 3; a++) xa.n[a] = m.n[a];
            xa.nzh = (int)nzh;
            xa.t = tabs;
            xa.wp = wp;
            xa.sn_re = snr;
            xa.sn_im = sni;
            xa.inv_v = 1.0 / m.volume();
            if (ctx->launch("svg_xi_src", [&] {
                    hipLaunchKernelGGL(k_ycross_xi<T>, dim3(gb), dim3(256), 0, ctx->stream, (const C<T>*)cf1,
                                       (const C<T>*)dk2, xa, cx);
                }))
                return -1;
            if (fft3d_c2c<T>(ctx, cx, n3, +1)) return -1;
            for (int a = 0; a < 3; a++) ha.n[a] = m.n[a];
            ha.nkey = p->nkey;
            ha.keys = p->keys;
            if (fill_zero(ctx, p->yh.p, sizeof(double) * 2 * (size_t)p->nkey)) return -1;
            if (ctx->launch("svg_hist", [&] {
                    hipLaunchKernelGGL(k_yhist<T>, dim3(gb), dim3(256), 0, ctx->stream, (const C<T>*)cx, ha, Hre, Him);
                }))
                return -1;
            const long long snb = (p->nkey + SNP_KPB - 1) / SNP_KPB;
            for (int par = 0; par < 2; par++) {
                if (ctx->launch("sn_pairs", [&] {
                        hipLaunchKernelGGL(k_sn_pairs_blk, dim3((unsigned)snb), dim3(256),
                                           sizeof(double) * 2 * (size_t)p->nk * SNP_CHP, ctx->stream,
                                           (const double*)p->jtab.p,
                                           (const double*)(p->l2 != p->l1 ? p->jtab2.p : p->jtab.p), p->nk, p->nkey,
                                           (const double*)(par ? Him : Hre), (const short2*)p->pairs_d.p, (int)npairs,
                                           dV, (double*)p->snp_part.p);
                    }))
                    return -1;
                if (reduce_cols<double>(ctx, "sn_pairs_reduce", (const double*)p->snp_part.p, snb, (int)npairs,
                                        (double*)p->sn_pairs.p))
                    return -1;
                std::vector<double> h((size_t)npairs);
                HM_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), p->sn_pairs.p, sizeof(double) * h.size(),
                                                 hipMemcpyDeviceToHost, ctx->stream));
                HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
                for (int q = 0; q < npairs; q++) sv_ijk[2 * (size_t)q + par] += h[q];
            }
            // two-point terms (bispec.hpp:33-93): P_{l1 m1}(k1) for l2 = 0, P_{l2 m2}(k2) for l1 = 0, of d_00
            // against N_LM = conj(Y_LM) w^2 (D + alpha^2 R) (field.hpp:249-306)
            if (p->l1 == 0 || p->l2 == 0) {
                if (svg_field<T>(ctx, p, ag, wt, np, nr, M, 2, 1, wt->c2, wt->abs2, cf2, run_assign && !inter))
                    return -1;
                for (int c = 0; c < ha.nc; c++) {
                    if (p->l2 == 0 && ybin(p->l1, ha.m1[c], ha.w[c], snr, sni, sv_row)) return -1;
                    if (p->l1 == 0 && ybin(p->l2, ha.m2[c], ha.w[c], snr, sni, sv_col)) return -1;
                }
            }
            // the third field dn_LM / W on the triple mesh, then the triple sums of every (m1, m2)
            TriStage<T> s3 = st;
            s3.dfull = cf1;
            const long long nst = (long long)s3.c3[0] * s3.c3[1] * s3.c3[2];
            if (ctx->launch("tri_stage", [&] {
                    hipLaunchKernelGGL(k_tri_stage<T>, dim3((unsigned)((nst + 255) / 256)), dim3(256), 0, ctx->stream,
                                       s3, F3, (char2*)nullptr);
                }))
                return -1;
            if (f3_passes()) return -1;
            for (int c = 0; c < ha.nc; c++)
                if (gram_combo(ha.m1[c], ha.m2[c], ha.w[c])) return -1;
        }
        return 0;
    };
    if (p->cfg.bk) {
        // 4. shot-noise ijk term (field.hpp:1116-1160, bispec.hpp:1023-1056); per M below for a survey ELL > 0
        if (!svg) {
        XiSrc<T> xs = make_xs(dk2);
        // Re xi; for an odd l1 + l2 also Im xi (field.hpp:1116-1160 keeps both: an odd angular weight sees
        // the odd part of xi, and the even part on the z Nyquist plane), and always for the complex cross
        // spectrum of a reconstructed box
        for (int par = 0; par < ((p->sn_odd || p->odd_mesh || recon) ? 2 : 1); par++) {
        xs.odd = par;
        void* hbuf = par ? p->hist2.p : p->hist.p;
        void* sbuf = par ? p->sn_pairs2.p : p->sn_pairs.p;
        C<T>* xh = (C<T>*)p->xi_half.p;
        if (fused_xi) {
            // the planes k_z = 0 and n_z/2 the fused pass left, both in one launch (y columns of stride nzh)
            XiPlaneSrc<T> ps{xs, m.n[2] / 2};
            ColGeom gp = geom(0, (long long)m.n[1] * nzh, 1, 1, -1, m.n[1], m.n[1], -1, -1);
            gp.IS = nzh;
            gp.FS = m.n[2] / 2;
            gp.nf = 2;
            if (launch_cols<T, +1>(ctx, "xi_planes_x", m.n[0], ps, xh, gp)) return -1;
        } else if (launch_cols<T, +1>(ctx, "xi_pass_x", m.n[0], xs, xh,
                                      geom(0, (long long)m.n[1] * nzh, 1, 1, -1, (int)(m.n[1] * nzh),
                                           (int)(m.n[1] * nzh), -1, -1)))
            return -1;
        if (fft_cols<T>(ctx, "xi_pass_y", m.n[1], +1, xh, xh,
                        geom((long long)m.n[1] * nzh, nzh, m.n[0], m.n[0], -1, (int)nzh, (int)nzh, -1, -1)))
            return -1;
        C<T>* xg = (C<T>*)p->xi_real.p;
        T* xr = (T*)p->xi_R.p;
        if (ctx->launch("xi_group", [&] {
                const unsigned nt = (unsigned)std::min<long long>(XG_NT, (nzh + 63) / 64 * 64);
                hipLaunchKernelGGL(k_xi_group<T>, dim3((unsigned)p->na), dim3(nt), 0, ctx->stream, xh, (int)nzh,
                                   (const int*)p->a_start.p, (const int*)p->a_rows.p, xg);
            }))
            return -1;
        if (fft_c2r_rows<T>(ctx, "xi_pass_z", m.n[2] / 2, xg, xr, (long long)p->na, -1)) return -1;
        if (ctx->launch("xi_hist", [&] {
                if (p->keys.aniso && p->bk_gen)
                    hipLaunchKernelGGL(k_xi_hist_ang_aniso<T>, dim3((unsigned)((p->nkey + 255) / 256)), dim3(256), 0,
                                       ctx->stream, xr, m.n[2], (const int*)p->a_idx.p, p->keys, p->nkey, p->ang,
                                       (double*)hbuf);
                else if (p->keys.aniso)
                    hipLaunchKernelGGL(k_xi_hist_aniso<T>, dim3((unsigned)((p->nkey + 255) / 256)), dim3(256), 0,
                                       ctx->stream, xr, m.n[2], (const int*)p->a_idx.p, p->keys, p->nkey,
                                       (double*)hbuf);
                else if (p->bk_gen)
                    hipLaunchKernelGGL(k_xi_hist_ang<T>, dim3((unsigned)((p->nkey * XH_G + 255) / 256)), dim3(256), 0,
                                       ctx->stream, xr, m.n[2], (const int*)p->a_idx.p, p->amax, p->nkey, p->ang,
                                       (double*)hbuf);
                else
                    hipLaunchKernelGGL(k_xi_hist<T>, dim3((unsigned)((p->nkey * XH_G + 255) / 256)), dim3(256), 0,
                                       ctx->stream, xr, m.n[2], (const int*)p->a_idx.p, p->amax, p->nkey,
                                       (double*)hbuf);
            }))
            return -1;
        const long long snb = (p->nkey + SNP_KPB - 1) / SNP_KPB;
        if (ctx->launch("sn_pairs", [&] {
                hipLaunchKernelGGL(k_sn_pairs_blk, dim3((unsigned)snb), dim3(256),
                                   sizeof(double) * 2 * (size_t)p->nk * SNP_CHP, ctx->stream, (const double*)p->jtab.p,
                                   (const double*)(p->bk_gen && p->l2 != p->l1 ? p->jtab2.p : p->jtab.p), p->nk,
                                   p->nkey, (const double*)hbuf, (const short2*)p->pairs_d.p, (int)npairs, dV,
                                   (double*)p->snp_part.p);
            }))
            return -1;
        if (reduce_cols<double>(ctx, "sn_pairs_reduce", (const double*)p->snp_part.p, snb, (int)npairs,
                                (double*)sbuf))
            return -1;
        }
        }
        // 5. shell fields and delta/W on the triple mesh (field.hpp:346-401, 610-639, 328-344)
        const long long nvox = (long long)p->M[0] * p->M[1] * p->M[2];
        C<T>* F = (C<T>*)p->fields.p;
        const int nsh = p->bk_gen ? 2 * p->nk : p->nk;  // shell fields: one set, or two Y-weighted sets
        C<T>* F3 = F + (long long)nsh * nvox;
        C<T>* Draw = p->bk_gen ? F3 + nvox : F3;  // delta/W on the triple mesh before any transform
        TriStage<T> st;
        st.dk = dk;
        st.nzh = (int)nzh;
        for (int a = 0; a < 3; a++) {
            st.n[a] = m.n[a];
            st.M[a] = p->M[a];
            st.s3[a] = sup_eff(p->s3[a], p->M[a]);
            st.c3[a] = support_count(st.s3[a], p->M[a]);
            st.s[a] = sup_eff(p->s[a], p->M[a]);
            st.cs[a] = support_count(st.s[a], p->M[a]);
        }
        st.t = tabs;
        st.wp = wp;
        st.nk = p->nk;
        st.lower = (const double*)p->lu_d.p;
        st.upper = st.lower + p->nk;
        const long long nst = (long long)st.c3[0] * st.c3[1] * st.c3[2];
        if (ctx->launch("tri_stage", [&] {
                hipLaunchKernelGGL(k_tri_stage<T>, dim3((unsigned)((nst + 255) / 256)), dim3(256), 0, ctx->stream, st,
                                   Draw, (char2*)p->sid.p);
            }))
            return -1;
        if (p->bk_gen && !svg)
            HM_HIP_CHECK(ctx, hipMemcpyAsync(F3, Draw, sizeof(C<T>) * (size_t)nvox, hipMemcpyDeviceToDevice, ctx->stream));
        const int4* fsup = (const int4*)p->fsup_d.p;
        const int* fsz = (const int*)p->fsz_d.p;
        ShellMaskSrc<T> ms;
        ms.D = Draw;
        ms.sid = (const char2*)p->sid.p;
        for (int a = 0; a < 3; a++) {
            ms.M[a] = p->M[a];
            ms.s[a] = st.s[a];
            ms.cs[a] = st.cs[a];
        }
        if (p->bk_gen) {
            const long long nrows = (long long)p->M[0] * p->M[1];
            // the third field on the triple mesh, all three passes
            auto f3_passes = [&]() -> int {
                ColGeom gx3 = tri_geom_x(p->M, st.s3, nvox, 1);
                if (fft_cols<T>(ctx, "tri3_pass_x", p->M[0], +1, F3, F3, gx3)) return -1;
                ColGeom gy3 = tri_geom_y(p->M, st.s3, nvox, 1);
                if (fft_cols<T>(ctx, "tri3_pass_y", p->M[1], +1, F3, F3, gy3)) return -1;
                RowGeom rg3{nrows, p->M[2], p->M[2], st.s3[2]};
                return fft_rows<T>(ctx, "tri3_pass_z", p->M[2], +1, F3, F3, rg3);
            };
            // shells weighted by Y(l1, m1) and by Y(l2, m2), all cross sums with the third field, times wq
            ms.ylm = 1;
            ms.nk = p->nk;
            ms.l1 = p->l1;
            ms.l2 = p->l2;
            ms.t = tabs;
            for (int a = 0; a < 3; a++) ms.n[a] = m.n[a];
            gacc.assign(2 * (size_t)npairs, 0.0);
            auto gram_combo = [&](int m1, int m2, double wq) -> int {
                ms.m1 = m1;
                ms.m2 = m2;
                ColGeom gx = tri_geom_x(p->M, st.s, nvox, nsh, fsup, p->fsup.data());
                if (launch_cols<T, +1>(ctx, "tri_pass_x", p->M[0], ms, F, gx)) return -1;
                ColGeom gy = tri_geom_y(p->M, st.s, nvox, nsh, fsup, p->fsup.data());
                if (fft_cols<T>(ctx, "tri_pass_y", p->M[1], +1, F, F, gy)) return -1;
                RowGeom rg{nrows * nsh, p->M[2], p->M[2], st.s[2], fsz, nrows};
                if (fft_rows<T>(ctx, "tri_pass_z", p->M[2], +1, F, F, rg)) return -1;
                for (int c0 = 0; c0 < npairs; c0 += 256) {
                    const int cnt = std::min(256, npairs - c0);
                    const short2* gp = (const short2*)p->gpairs_d.p + c0;
                    size_t lds = sizeof(C<T>) * (GRAM_CH + 2 * (size_t)nsh * (GRAM_CH + 1));
                    const bool narrow = lds > 65536;
                    if (narrow) lds = sizeof(C<T>) * (32 + 2 * (size_t)nsh * 33);
                    if (lds > 65536) {
                        ctx->set_error("box plan: too many k-bins for the general bispectrum multipole");
                        return -1;
                    }
                    if (ctx->launch("gram", [&] {
                            if (narrow)
                                hipLaunchKernelGGL((k_gram_sym<T, 32>), dim3((unsigned)p->gram_blocks), dim3(256), lds,
                                                   ctx->stream, F, nvox, nsh, F3, nvox, p->gram_vpb, gp, cnt,
                                                   (double2*)p->gram_part.p);
                            else
                                hipLaunchKernelGGL((k_gram_sym<T>), dim3((unsigned)p->gram_blocks), dim3(256), lds,
                                                   ctx->stream, F, nvox, nsh, F3, nvox, p->gram_vpb, gp, cnt,
                                                   (double2*)p->gram_part.p);
                        }))
                        return -1;
                    if (reduce_cols<double>(ctx, "gram_reduce", (const double*)p->gram_part.p, p->gram_blocks,
                                            2 * cnt, (double*)p->gram_red.p))
                        return -1;
                    std::vector<double> part(2 * (size_t)cnt);
                    HM_HIP_CHECK(ctx, hipMemcpyAsync(part.data(), p->gram_red.p, sizeof(double) * part.size(),
                                                     hipMemcpyDeviceToHost, ctx->stream));
                    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
                    for (int e = 0; e < 2 * cnt; e++) gacc[2 * (size_t)c0 + e] += wq * part[e];
                }
                return 0;
            };
            if (!svg) {
                // per m1 (bispec.hpp:1107-1180): Y(l1, m1) and Y(l2, -m1), delta/W third
                if (f3_passes()) return -1;
                for (size_t q = 0; q < p->mlist.size(); q++)
                    if (gram_combo(p->mlist[q], -p->mlist[q], p->wlist[q])) return -1;
            } else if (svg_stage(F3, st, f3_passes, gram_combo)) {
                return -1;
            }
        } else {
            // x pass of every shell (one launch), then of delta/W in place
            ColGeom gx = tri_geom_x(p->M, st.s, nvox, p->nk, fsup, p->fsup.data());
            if (launch_cols<T, +1>(ctx, "tri_pass_x", p->M[0], ms, F, gx)) return -1;
            ColGeom gx3 = tri_geom_x(p->M, st.s3, nvox, 1);
            if (fft_cols<T>(ctx, "tri3_pass_x", p->M[0], +1, F3, F3, gx3)) return -1;
            // y pass
            ColGeom gy = tri_geom_y(p->M, st.s, nvox, p->nk, fsup, p->fsup.data());
            if (fft_cols<T>(ctx, "tri_pass_y", p->M[1], +1, F, F, gy)) return -1;
            ColGeom gy3 = tri_geom_y(p->M, st.s3, nvox, 1);
            if (fft_cols<T>(ctx, "tri3_pass_y", p->M[1], +1, F3, F3, gy3)) return -1;
            // 6. z pass and triple sums
            const long long nrows = (long long)p->M[0] * p->M[1];
            if (p->fused) {
                if (launch_tri_z_gram<T>(ctx, p->M[2], F, nvox, p->nk, fsz, st.s3[2], nrows, p->ngrp, p->nblk, p->nsl,
                                         p->tri_rpw, (double*)p->gram_part.p))
                    return -1;
            } else {
                RowGeom rg{nrows * p->nk, p->M[2], p->M[2], st.s[2], fsz, nrows};
                if (fft_rows<T>(ctx, "tri_pass_z", p->M[2], +1, F, F, rg)) return -1;
                RowGeom rg3{nrows, p->M[2], p->M[2], st.s3[2]};
                if (fft_rows<T>(ctx, "tri3_pass_z", p->M[2], +1, F3, F3, rg3)) return -1;
                const size_t lds = sizeof(C<T>) * (GRAM_CH + 2 * (size_t)p->nk * (GRAM_CH + 1));
                if (ctx->launch("gram", [&] {
                        hipLaunchKernelGGL(k_gram_sym<T>, dim3((unsigned)p->gram_blocks), dim3(256), lds, ctx->stream, F,
                                           nvox, p->nk, F3, nvox, p->gram_vpb, (const short2*)p->pairs_d.p, npairs,
                                           (double2*)p->gram_part.p);
                    }))
                    return -1;
            }
            if (reduce_cols<double>(ctx, "gram_reduce", (const double*)p->gram_part.p, p->gram_blocks, p->gram_cols,
                                    (double*)p->gram_red.p))
                return -1;
        }
    }
    // 6'. three-point function (bispec.hpp:1205-1424)
    std::vector<double> zacc, zsum, zcntv;
    if (p->zeta) {
        const long long n_rand = nr;
        const int nr = p->cfg.n_rbin;
        const int nf = p->z_gen ? 2 * nr : nr;
        const long long ng = m.ng();
        C<T>* ZF = (C<T>*)p->zfields.p;
        C<T>* Z3 = ZF + (long long)nf * ng;
        BesselSrc<T> bs;
        bs.dk = dk;
        for (int a = 0; a < 3; a++) bs.n[a] = m.n[a];
        bs.nzh = (int)nzh;
        bs.t = tabs;
        bs.wp = wp;
        bs.nr = nr;
        bs.inv_v = 1.0 / m.volume();
        bs.rb = (const double*)p->zrb_d.p;
        bs.l[0] = p->zl[0];
        bs.l[1] = p->zl[1];
        bs.sy[0] = (const double*)p->zsy[0].p;
        bs.sc[0] = (const double*)p->zsc[0].p;
        bs.sy[1] = (const double*)(p->zl[1] == p->zl[0] ? p->zsy[0].p : p->zsy[1].p);
        bs.sc[1] = (const double*)(p->zl[1] == p->zl[0] ? p->zsc[0].p : p->zsc[1].p);
        const long long plane = (long long)m.n[1] * m.n[2];
        auto full_inverse = [&](const char* nx, const char* ny, const char* nz, C<T>* out, int nfield) -> int {
            ColGeom gx = geom(0, plane, 1, 1, -1, (int)plane, (int)plane, -1, -1);
            gx.FS = ng;
            gx.nf = nfield;
            if (launch_cols<T, +1>(ctx, nx, m.n[0], bs, out, gx)) return -1;
            ColGeom gy = geom(plane, m.n[2], m.n[0], m.n[0], -1, m.n[2], m.n[2], -1, -1);
            gy.FS = ng;
            gy.nf = nfield;
            if (fft_cols<T>(ctx, ny, m.n[1], +1, out, out, gy)) return -1;
            RowGeom rg{(long long)m.n[0] * m.n[1] * nfield, m.n[2], m.n[2], -1};
            return fft_rows<T>(ctx, nz, m.n[2], +1, out, out, rg);
        };
        const int nzp = (int)p->zpairs.size();
        const long long gblocks = (ng + p->gram_vpb - 1) / p->gram_vpb;
        double* gpart = (double*)p->zpart.p + 2 * (size_t)nr * (size_t)std::max(p->z_rows, (long long)1);
        zacc.assign(2 * (size_t)nzp, 0.0);
        // sum_x F1 F2 Z3 over every r-bin pair of the fields in ZF, times wq, into zacc
        auto zeta_gram = [&](double wq) -> int {
            for (int c0 = 0; c0 < nzp; c0 += 256) {
                const int cnt = std::min(256, nzp - c0);
                const short2* gp = (const short2*)p->zpairs_d.p + c0;
                size_t lds = sizeof(C<T>) * (GRAM_CH + 2 * (size_t)nf * (GRAM_CH + 1));
                const bool narrow = lds > 65536;
                if (narrow) lds = sizeof(C<T>) * (32 + 2 * (size_t)nf * 33);
                if (lds > 65536) {
                    ctx->set_error("box plan: too many r-bins for the three-point function");
                    return -1;
                }
                if (ctx->launch("zeta_gram", [&] {
                        if (narrow)
                            hipLaunchKernelGGL((k_gram_sym<T, 32>), dim3((unsigned)gblocks), dim3(256), lds, ctx->stream,
                                               ZF, ng, nf, Z3, ng, p->gram_vpb, gp, cnt, (double2*)gpart);
                        else
                            hipLaunchKernelGGL((k_gram_sym<T>), dim3((unsigned)gblocks), dim3(256), lds, ctx->stream,
                                               ZF, ng, nf, Z3, ng, p->gram_vpb, gp, cnt, (double2*)gpart);
                    }))
                    return -1;
                if (reduce_cols<double>(ctx, "gram_reduce", gpart, gblocks, 2 * cnt, (double*)p->zred.p)) return -1;
                std::vector<double> part(2 * (size_t)cnt);
                HM_HIP_CHECK(ctx, hipMemcpyAsync(part.data(), p->zred.p, sizeof(double) * part.size(),
                                                 hipMemcpyDeviceToHost, ctx->stream));
                HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
                for (int e = 0; e < 2 * cnt; e++) zacc[2 * (size_t)c0 + e] += wq * part[e];
            }
            return 0;
        };
        ZetaSnArgs za;
        for (int a = 0; a < 3; a++) {
            za.n[a] = m.n[a];
            za.s[a] = p->zrs[a];
            za.c[a] = support_count(p->zrs[a], m.n[a]);
            za.dr[a] = m.L[a] / (double)m.n[a];
        }
        za.nrows = p->z_rows;
        za.f2b = (const int2*)p->zf2b_d.p;
        za.nr = nr;
        za.A = p->zang;
        if (fill_zero(ctx, p->zpart.p, sizeof(double) * 2 * (size_t)nr * p->z_rows)) return -1;
        if (fill_zero(ctx, p->zcnt.p, sizeof(long long) * (size_t)nr * p->z_rows)) return -1;
        if (!svg) {
        // (d/W)/V transformed back once (bispec.hpp:1308-1317)
        bs.plain = 1;
        if (full_inverse("zeta3_pass_x", "zeta3_pass_y", "zeta3_pass_z", Z3, 1)) return -1;
        bs.plain = 0;
        for (size_t q = 0; q < p->zmlist.size(); q++) {
            bs.m[0] = p->zmlist[q];
            bs.m[1] = -p->zmlist[q];
            if (full_inverse("zeta_pass_x", "zeta_pass_y", "zeta_pass_z", ZF, nf)) return -1;
            if (zeta_gram(p->zwlist[q])) return -1;
        }
        // shot noise: xi of (d conj(n) - N S)/W^2/V, weighted by A(mu), binned (field.hpp:979-1113)
        XiSrc<T> xs;
        xs.dk = dk;
        for (int a = 0; a < 3; a++) xs.n[a] = m.n[a];
        xs.nzh = (int)nzh;
        xs.t = tabs;
        xs.xq = xi_xq(p);
        xs.wp = wp;
        xs.sn = sn3;
        xs.inv_v = 1.0 / m.volume();
        xs.dc = recon ? 0.0 : (double)np;
        xs.sn_ptr = dyn ? dyn + DYN_SN3 : nullptr;
        xs.dc_ptr = dyn ? dyn + DYN_DC : nullptr;
        xs.dk2 = dk2;
        C<T>* xh = (C<T>*)p->xi_half.p;
        const int cx = support_count(p->zrs[0], m.n[0]), cy = support_count(p->zrs[1], m.n[1]);
        for (int par = 0; par < ((p->z_odd || p->odd_mesh || recon) ? 2 : 1); par++) {
            xs.odd = par;
            if (launch_cols<T, +1>(ctx, "zsn_pass_x", m.n[0], xs, xh,
                                   geom(0, (long long)m.n[1] * nzh, 1, 1, -1, (int)(m.n[1] * nzh), (int)(m.n[1] * nzh),
                                        -1, -1)))
                return -1;
            if (fft_cols<T>(ctx, "zsn_pass_y", m.n[1], +1, xh, xh,
                            geom((long long)m.n[1] * nzh, nzh, cx, m.n[0], p->zrs[0], (int)nzh, (int)nzh, -1, -1)))
                return -1;
            SupHalfSrc<T> hs{xh, (int)nzh, m.n[0], m.n[1], p->zrs[0], p->zrs[1], cy};
            if (launch_c2r<T>(ctx, "zsn_pass_z", m.n[2] / 2, hs, (T*)p->zr.p, p->z_rows, -1)) return -1;
            za.imag = par;
            za.count = par == 0;
            if (ctx->launch("zsn_bin", [&] {
                    hipLaunchKernelGGL(k_bin_xi_ang<T>, dim3((unsigned)((p->z_rows + 255) / 256)), dim3(256), 0,
                                       ctx->stream, (const T*)p->zr.p, za, (double*)p->zpart.p,
                                       (long long*)p->zcnt.p);
                }))
                return -1;
        }
        } else {
            // survey three-point function with ELL > 0 (bispec.hpp:355-874): per M the Y_LM(los)-weighted
            // fluctuation dn_LM; (dn_LM / W) / V transformed back is the third field of every (m1, m2) with
            // m1 + m2 + M = 0, and xi_LM = IFFT[(dn_LM conj N_00 - SN_LM S)/W^2/V] weighted by
            // sum_c w_c Y(l1, m1_c)(r) Y(l2, m2_c)(r) is the shot noise
            const int L = wt->L;
            const int n3[3] = {m.n[0], m.n[1], m.n[2]};
            if ((!p->cf1.p && alloc(ctx, p, p->cf1, sizeof(C<T>) * (size_t)ng)) ||
                (!p->cx.p && alloc(ctx, p, p->cx, sizeof(C<T>) * (size_t)ng)))
                return -1;
            C<T>* cf1 = (C<T>*)p->cf1.p;
            C<T>* cxf = (C<T>*)p->cx.p;
            const unsigned gb = (unsigned)std::max<long long>(1, std::min<long long>((ng + 255) / 256, 16384));
            ZetaYArgs ya;
            for (int a = 0; a < 3; a++) {
                ya.n[a] = za.n[a];
                ya.s[a] = za.s[a];
                ya.c[a] = za.c[a];
                ya.dr[a] = za.dr[a];
            }
            ya.nrows = za.nrows;
            ya.f2b = za.f2b;
            ya.nr = nr;
            ya.l1 = p->zl[0];
            ya.l2 = p->zl[1];
            bool first = true;
            for (int M = -L; M <= L; M++) {
                ya.nc = 0;
                for (int m1 = -p->zl[0]; m1 <= p->zl[0]; m1++) {
                    const int m2 = -M - m1;
                    if (m2 < -p->zl[1] || m2 > p->zl[1]) continue;
                    double w = wigner_3j(p->zl[0], p->zl[1], L, 0, 0, 0) * wigner_3j(p->zl[0], p->zl[1], L, m1, m2, M);
                    w *= (double)(2 * L + 1) * (double)(2 * p->zl[0] + 1) * (double)(2 * p->zl[1] + 1);
                    if (std::fabs(w) < 1.0e-10) continue;
                    if (ya.nc >= YH_MAXC) {
                        ctx->set_error("survey three-point function: too many (m1, m2) terms");
                        return -1;
                    }
                    ya.m1[ya.nc] = m1;
                    ya.m2[ya.nc] = m2;
                    ya.w[ya.nc] = w;
                    ya.nc++;
                }
                if (ya.nc == 0) continue;
                if (svg_field<T>(ctx, p, ag, wt, np, n_rand, M, 1, 0, wt->c1, wt->abs1, cf1, run_assign && !inter))
                    return -1;
                YXiArgs xa;
                for (int a = 0; a < 3; a++) xa.n[a] = m.n[a];
                xa.nzh = (int)nzh;
                xa.t = tabs;
                xa.wp = wp;
                xa.sn_re = wt->snLM[2 * (M + L)];
                xa.sn_im = wt->snLM[2 * (M + L) + 1];
                xa.inv_v = 1.0 / m.volume();
                if (ctx->launch("svz_xi_src", [&] {
                        hipLaunchKernelGGL(k_ycross_xi<T>, dim3(gb), dim3(256), 0, ctx->stream, (const C<T>*)cf1,
                                           (const C<T>*)dk2, xa, cxf);
                    }))
                    return -1;
                if (fft3d_c2c<T>(ctx, cxf, n3, +1)) return -1;
                ya.count = first ? 1 : 0;
                first = false;
                if (ctx->launch("svz_sn_bin", [&] {
                        hipLaunchKernelGGL(k_bin_xi_y<T>, dim3((unsigned)((p->z_rows + 255) / 256)), dim3(256), 0,
                                           ctx->stream, (const C<T>*)cxf, ya, (double*)p->zpart.p,
                                           (long long*)p->zcnt.p);
                    }))
                    return -1;
                if (ctx->launch("svz_third", [&] {
                        hipLaunchKernelGGL(k_full_over_w<T>, dim3(gb), dim3(256), 0, ctx->stream, (const C<T>*)cf1,
                                           m.n[0], m.n[1], m.n[2], tabs, wp, 1.0 / m.volume(), Z3);
                    }))
                    return -1;
                if (fft3d_c2c<T>(ctx, Z3, n3, +1)) return -1;
                for (int c = 0; c < ya.nc; c++) {
                    bs.m[0] = ya.m1[c];
                    bs.m[1] = ya.m2[c];
                    if (full_inverse("zeta_pass_x", "zeta_pass_y", "zeta_pass_z", ZF, nf)) return -1;
                    if (zeta_gram(ya.w[c])) return -1;
                }
            }
        }
        if (reduce_cols<double>(ctx, "bin_reduce", (const double*)p->zpart.p, p->z_rows, 2 * nr, (double*)p->zred.p))
            return -1;
        if (reduce_cols<long long>(ctx, "bin_reduce", (const long long*)p->zcnt.p, p->z_rows, nr,
                                   (long long*)p->zcnt_red.p))
            return -1;
        zsum.resize(2 * (size_t)nr);
        std::vector<long long> zc(nr);
        HM_HIP_CHECK(ctx, hipMemcpyAsync(zsum.data(), p->zred.p, sizeof(double) * zsum.size(), hipMemcpyDeviceToHost,
                                         ctx->stream));
        HM_HIP_CHECK(ctx, hipMemcpyAsync(zc.data(), p->zcnt_red.p, sizeof(long long) * nr, hipMemcpyDeviceToHost,
                                         ctx->stream));
        HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        zcntv.assign(zc.begin(), zc.end());
    }
    if (xo && !xo->readback) return 0;
    // 7. results to host
    std::vector<double> pk(ncols), gram(2 * (size_t)npairs), graw((size_t)p->gram_cols), snp((size_t)npairs),
        snp2((size_t)npairs, 0.0);
    std::vector<long long> cnt(p->nk);
    std::vector<double> xsum((size_t)p->nr * p->nxe);
    std::vector<long long> xcnt(p->nr);
    if (do_pk) {
        HM_HIP_CHECK(ctx, hipMemcpyAsync(pk.data(), p->pk_red.p, sizeof(double) * ncols, hipMemcpyDeviceToHost,
                                         ctx->stream));
        HM_HIP_CHECK(ctx, hipMemcpyAsync(cnt.data(), p->pk_cnt_red.p, sizeof(long long) * p->nk,
                                         hipMemcpyDeviceToHost, ctx->stream));
    }
    if (p->nxe > 0) {
        HM_HIP_CHECK(ctx, hipMemcpyAsync(xsum.data(), p->xi_red.p, sizeof(double) * xsum.size(), hipMemcpyDeviceToHost,
                                         ctx->stream));
        HM_HIP_CHECK(ctx, hipMemcpyAsync(xcnt.data(), p->xi_cnt_red.p, sizeof(long long) * p->nr,
                                         hipMemcpyDeviceToHost, ctx->stream));
    }
    if (p->cfg.bk) {
        if (!p->bk_gen)
            HM_HIP_CHECK(ctx, hipMemcpyAsync(graw.data(), p->gram_red.p, sizeof(double) * p->gram_cols,
                                             hipMemcpyDeviceToHost, ctx->stream));
        HM_HIP_CHECK(ctx, hipMemcpyAsync(snp.data(), p->sn_pairs.p, sizeof(double) * npairs, hipMemcpyDeviceToHost,
                                         ctx->stream));
        if (p->sn_odd || p->odd_mesh || recon)
            HM_HIP_CHECK(ctx, hipMemcpyAsync(snp2.data(), p->sn_pairs2.p, sizeof(double) * npairs,
                                             hipMemcpyDeviceToHost, ctx->stream));
    }
    std::vector<double> pkx(cross_pk ? 2 * (size_t)ncols : 0);
    if (cross_pk)
        HM_HIP_CHECK(ctx, hipMemcpyAsync(pkx.data(), p->pkx_red.p, sizeof(double) * pkx.size(), hipMemcpyDeviceToHost,
                                         ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const int nk = p->nk;
    const double V = m.volume(), N = (double)np;
    if (p->cfg.bk && !p->bk_gen) gram_to_pairs(p, graw, gram);
    if (p->bk_gen) gram = gacc;
    // P_l: average (field.hpp:839-845), (2l+1) and V/N^2 (powerspec.hpp:306-316)
    std::vector<double> praw((size_t)ba.ne * nk);
    for (int e = 0; e < ba.ne; e++)
        for (int b = 0; b < nk; b++) praw[(size_t)e * nk + b] = cnt[b] ? pk[(size_t)b * ba.ne + e] / (double)cnt[b] : 0.0;
    if (res) {
        const double norm = V / N / N;
        for (int e = 0; e < p->cfg.n_pk_ells; e++)  // the requested multipoles lead bin_ells
            for (int b = 0; b < nk; b++) {
                const double v = praw[(size_t)e * nk + b];
                if (res->pk_raw) {
                    res->pk_raw[2 * ((size_t)e * nk + b)] = v;
                    res->pk_raw[2 * ((size_t)e * nk + b) + 1] = 0.0;
                }
                if (res->pk) {
                    res->pk[2 * ((size_t)e * nk + b)] = norm * (0.0 + (double)(2 * p->cfg.pk_ells[e] + 1) * v);
                    res->pk[2 * ((size_t)e * nk + b) + 1] = 0.0;
                }
            }
        if (res->pk_nmode)
            for (int b = 0; b < nk; b++) res->pk_nmode[b] = cnt[b];
        if (res->shell_nmode)
            for (int b = 0; b < nk; b++) res->shell_nmode[b] = p->shell_nmode[b];
        // xi_l: average (field.hpp:961-968), (2l+1) and V/N^2, real part (powerspec.hpp:361-386)
        for (int e = 0; e < p->nxe; e++)
            for (int b = 0; b < p->nr; b++) {
                const double v = xcnt[b] ? xsum[(size_t)b * p->nxe + e] / (double)xcnt[b] : 0.0;
                if (res->xi_raw) res->xi_raw[(size_t)e * p->nr + b] = v;
                if (res->xi) res->xi[(size_t)e * p->nr + b] = (V / N / N) * (0.0 + (double)(2 * p->cfg.xi_ells[e] + 1) * v);
            }
        if (res->xi_nmode)
            for (int b = 0; b < p->nr; b++) res->xi_nmode[b] = xcnt[b];
    }
    if (p->cfg.bk && res) {
        // shot noise (bispec.hpp:926-1073): w N for l1 = l2 = 0, w P_l1(k1) for l2 = 0 (the row bin),
        // w P_l2(k2) for l1 = 0 (the column bin), i^(l1+l2) times the Bessel sweep
        // bin averages of the shot-noise spectrum: |d|^2 - N S, or the cross spectrum (re, im) after
        // reconstruction
        auto praw_of = [&](int e, int b) {
            if (e < 0) return 0.0;
            if (!cross_pk) return praw[(size_t)e * nk + b];
            return cnt[b] ? pkx[(size_t)b * ba.ne + e] / (double)cnt[b] : 0.0;
        };
        auto praw_im = [&](int e, int b) {
            if (e < 0 || !cross_pk) return 0.0;
            return cnt[b] ? pkx[(size_t)ncols + (size_t)b * ba.ne + e] / (double)cnt[b] : 0.0;
        };
        const int ip = ((p->l1 + p->l2) % 4 + 4) % 4;
        const double ir = ip == 0 ? 1.0 : (ip == 2 ? -1.0 : 0.0), ii = ip == 1 ? 1.0 : (ip == 3 ? -1.0 : 0.0);
        const long long M3 = (long long)p->M[0] * p->M[1] * p->M[2];
        const double norm = weighted ? wt->norm : (V / N / N) * (V / N);
        for (int q = 0; svg && q < npairs; q++) {
            // survey ELL > 0 (bispec.hpp:33-211): row and column terms (bin averages), i^(l1+l2) times the ijk sums
            const int a = p->pairs[q].x, b = p->pairs[q].y;
            const double den = (double)M3 * (double)p->shell_nmode[a] * (double)p->shell_nmode[b];
            const double br = gram[2 * q] / den, bi = gram[2 * q + 1] / den;
            double sr = 0.0, si = 0.0;
            if (p->l2 == 0 && cnt[a]) {
                sr += sv_row[2 * (size_t)a] / (double)cnt[a];
                si += sv_row[2 * (size_t)a + 1] / (double)cnt[a];
            }
            if (p->l1 == 0 && cnt[b]) {
                sr += sv_col[2 * (size_t)b] / (double)cnt[b];
                si += sv_col[2 * (size_t)b + 1] / (double)cnt[b];
            }
            const double jr = sv_ijk[2 * (size_t)q], ji = sv_ijk[2 * (size_t)q + 1];
            sr += ir * jr - ii * ji;
            si += ir * ji + ii * jr;
            const size_t o = 2 * ((size_t)a * nk + b);
            if (res->bk_sum) {
                res->bk_sum[o] = br;
                res->bk_sum[o + 1] = bi;
            }
            if (res->sn_sum) {
                res->sn_sum[o] = sr;
                res->sn_sum[o + 1] = si;
            }
            if (res->bk) {
                res->bk[o] = norm * (br - sr);
                res->bk[o + 1] = norm * (bi - si);
            }
        }
        for (int q = 0; !svg && q < npairs; q++) {
            const int a = p->pairs[q].x, b = p->pairs[q].y;
            const double den = (double)M3 * (double)p->shell_nmode[a] * (double)p->shell_nmode[b];
            const double br = gram[2 * q] / den, bi = gram[2 * q + 1] / den;
            const double base = p->w_const * sn3 + p->w_row * praw_of(p->e_row, a) + p->w_col * praw_of(p->e_col, b);
            const double base_i = p->w_row * praw_im(p->e_row, a) + p->w_col * praw_im(p->e_col, b);
            // i^(l1+l2) (S_re + i S_im)
            const double sr = base + (ir * snp[q] - ii * snp2[q]), si = base_i + (ir * snp2[q] + ii * snp[q]);
            for (int sym = 0; sym < (p->bk_gen ? 1 : 2); sym++) {
                const int r = sym ? b : a, c = sym ? a : b;
                const size_t o = 2 * ((size_t)r * nk + c);
                if (res->bk_sum) {
                    res->bk_sum[o] = br;
                    res->bk_sum[o + 1] = bi;
                }
                if (res->sn_sum) {
                    res->sn_sum[o] = sr;
                    res->sn_sum[o + 1] = si;
                }
                if (res->bk) {
                    res->bk[o] = norm * (br - sr);
                    res->bk[o + 1] = norm * (bi - si);
                }
            }
        }
    }
    if (p->zeta && res) {
        const int nr = p->cfg.n_rbin;
        const double dV = V / (double)m.ng();
        const int lz = p->zl[0] + p->zl[1];
        const int ip = lz % 4;
        const double ir = ip == 0 ? 1.0 : (ip == 2 ? -1.0 : 0.0), ii = ip == 1 ? 1.0 : (ip == 3 ? -1.0 : 0.0);
        const double sg = (lz & 1) ? -1.0 : 1.0;  // (-1)^(l1+l2) (field.hpp:1104)
        const double norm = weighted ? wt->norm : (V / N / N) * (V / N);
        auto put = [&](double* dst, int r, int c, double re, double im) {
            if (!dst) return;
            dst[2 * ((size_t)r * nr + c)] = re;
            dst[2 * ((size_t)r * nr + c) + 1] = im;
        };
        for (size_t q = 0; q < p->zpairs.size(); q++) {
            const int a = p->zpairs[q].x, b = p->zpairs[q].y % nr;
            // i^(l1+l2) dV sum_x F1 F2 F3 (bispec.hpp:1381-1387)
            const double gr = zacc[2 * q] * dV, gi = zacc[2 * q + 1] * dV;
            const double zr = ir * gr - ii * gi, zi = ir * gi + ii * gr;
            for (int sym = 0; sym < (p->z_gen ? 1 : 2); sym++) {
                const int r = sym ? b : a, c = sym ? a : b;
                double sr = 0.0, si = 0.0;
                if (r == c && zcntv[r] != 0) {
                    const double f = sg / (double)zcntv[r] / (double)zcntv[r] / dV;
                    sr = zsum[2 * r] * f;
                    si = zsum[2 * r + 1] * f;
                }
                put(res->zeta_sum, r, c, zr, zi);
                put(res->zeta_sn, r, c, sr, si);
                put(res->zeta, r, c, norm * (zr - sr), norm * (zi - si));
            }
        }
    }
    return 0;
}

int plan_execute_any(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos, long long np, hm_box_result* res,
                     const double* d_pos_r, long long nr, double alpha) {
    return p->cfg.precision == HM_FP64 ? plan_execute<double>(ctx, p, d_pos, np, res, d_pos_r, nr, alpha)
                                       : plan_execute<float>(ctx, p, d_pos, np, res, d_pos_r, nr, alpha);
}

int plan_execute_weighted(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos_d, long long n_d, const double* d_pos_r,
                          long long n_r, const BoxWeighted* wt, hm_box_result* res) {
    return p->cfg.precision == HM_FP64 ? plan_execute<double>(ctx, p, d_pos_d, n_d, res, d_pos_r, n_r, 0.0, wt)
                                       : plan_execute<float>(ctx, p, d_pos_d, n_d, res, d_pos_r, n_r, 0.0, wt);
}

int plan_execute_opts(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos, long long np, const ExecOpts& xo) {
    return p->cfg.precision == HM_FP64 ? plan_execute<double>(ctx, p, d_pos, np, nullptr, nullptr, 0, 0.0, nullptr, &xo)
                                       : plan_execute<float>(ctx, p, d_pos, np, nullptr, nullptr, 0, 0.0, nullptr, &xo);
}

std::string plan_batch_check(const BoxPlanImpl* p) {
    if (p->cfg.interlace) return "batched execution: interlacing is not supported";
    if (p->nxe > 0 || p->zeta) return "batched execution: P_l and B only (no xi, no three-point function)";
    if (p->cfg.bk && p->bk_gen) return "batched execution: B_000 only";
    if (p->cfg.n_pk_ells <= 0 && !p->cfg.bk) return "batched execution: nothing to measure";
    return "";
}

void plan_vector_dims(const BoxPlanImpl* p, int* n_pk, int* n_bk) {
    *n_pk = p->cfg.n_pk_ells * p->nk;
    *n_bk = p->cfg.bk ? (int)p->pairs.size() : 0;
}

int plan_batch_prepare(hm_ctx* ctx, BoxPlanImpl* p) {
    if (!p->cfg.bk || p->qo_d.p) return 0;
    // the raw column of each (a <= b) pair, as gram_to_pairs reads it
    const int npairs = (int)p->pairs.size();
    std::vector<int> qo((size_t)npairs);
    if (!p->fused) {
        for (int q = 0; q < npairs; q++) qo[q] = q;
    } else {
        std::vector<int> pidx((size_t)p->nk * p->nk, -1);
        for (int q = 0; q < npairs; q++) pidx[(size_t)p->pairs[q].x * p->nk + p->pairs[q].y] = q;
        int blk = 0;
        for (int bi = 0; bi < p->ngrp; bi++)
            for (int bj = bi; bj < p->ngrp; bj++, blk++)
                for (int i = 0; i < 4; i++)
                    for (int j = 0; j < 4; j++) {
                        const int a = 4 * bi + i, b = 4 * bj + j;
                        if (a >= p->nk || b >= p->nk || a > b) continue;
                        qo[pidx[(size_t)a * p->nk + b]] = blk * 16 + i * 4 + j;
                    }
    }
    if (alloc(ctx, p, p->qo_d, sizeof(int) * qo.size())) return -1;
    HM_HIP_CHECK(ctx, hipMemcpy(p->qo_d.p, qo.data(), sizeof(int) * qo.size(), hipMemcpyHostToDevice));
    return 0;
}

// The data vector of one mock, formed on the device with the operations (and
// their order) of the host finalisation of plan_execute: P_l (field.hpp:839-845,
// powerspec.hpp:306-316) for the requested multipoles, then the real part of
// B_000 - SN (bispec.hpp:926-1073, 1175-1183) for each bin pair a <= b.
struct FinArgs {
    const double* pk;          // [nk][ne] mode sums
    const long long* cnt;      // [nk] mode counts
    int nk, ne, npk_ells;
    int ells[BIN_MAX_ELLS];
    const double* gram;        // raw triple sums (re, im) per column
    const int* qo;             // raw column of each pair
    const short2* pairs;       // (a, b), a <= b
    int npairs;
    const long long* shell;    // shell mode counts
    const double* snp;         // Bessel sweep per pair
    double M3, V, w_const, w_row, w_col;
    int e_row, e_col;
    const double* dyn;         // N and the three-point shot noise
};

__global__ __launch_bounds__(256) void k_batch_finalize(FinArgs a, double* __restrict__ out) {
#pragma clang fp contract(off)
    const double N = a.dyn[DYN_N], sn3 = a.dyn[DYN_SN3];
    const double V = a.V;
    const int npk = a.npk_ells * a.nk;
    for (int t = threadIdx.x; t < npk + a.npairs; t += blockDim.x) {
        if (t < npk) {
            const int e = t / a.nk, b = t % a.nk;
            const double v = a.cnt[b] ? a.pk[(size_t)b * a.ne + e] / (double)a.cnt[b] : 0.0;
            const double norm = V / N / N;
            out[t] = norm * (0.0 + (double)(2 * a.ells[e] + 1) * v);
        } else {
            const int q = t - npk;
            const int pa = a.pairs[q].x, pb = a.pairs[q].y;
            const double den = a.M3 * (double)a.shell[pa] * (double)a.shell[pb];
            const double br = a.gram[2 * (size_t)a.qo[q]] / den;
            const double pr = a.e_row < 0 ? 0.0 : (a.cnt[pa] ? a.pk[(size_t)pa * a.ne + a.e_row] / (double)a.cnt[pa] : 0.0);
            const double pc = a.e_col < 0 ? 0.0 : (a.cnt[pb] ? a.pk[(size_t)pb * a.ne + a.e_col] / (double)a.cnt[pb] : 0.0);
            const double base = a.w_const * sn3 + a.w_row * pr + a.w_col * pc;
            const double sr = base + a.snp[q];  // i^(l1 + l2) = 1
            const double norm = (V / N / N) * (V / N);
            out[t] = norm * (br - sr);
        }
    }
}

int plan_finalize_device(hm_ctx* ctx, BoxPlanImpl* p, const double* dyn, double* out) {
    FinArgs a{};
    a.pk = (const double*)p->pk_red.p;
    a.cnt = (const long long*)p->pk_cnt_red.p;
    a.nk = p->nk;
    a.ne = std::max(p->ne, 1);
    a.npk_ells = p->cfg.n_pk_ells;
    for (int e = 0; e < BIN_MAX_ELLS; e++) a.ells[e] = (e < p->ne) ? p->bin_ells[e] : 0;
    a.gram = (const double*)p->gram_red.p;
    a.qo = (const int*)p->qo_d.p;
    a.pairs = (const short2*)p->pairs_d.p;
    a.npairs = p->cfg.bk ? (int)p->pairs.size() : 0;
    a.shell = (const long long*)p->shell_cnt.p;
    a.snp = (const double*)p->sn_pairs.p;
    a.M3 = (double)((long long)p->M[0] * p->M[1] * p->M[2]);
    a.V = p->mesh.volume();
    a.w_const = p->w_const;
    a.w_row = p->w_row;
    a.w_col = p->w_col;
    a.e_row = p->e_row;
    a.e_col = p->e_col;
    a.dyn = dyn;
    if (a.npairs > 0 && !a.qo) {
        ctx->set_error("batched execution: plan not prepared");
        return -1;
    }
    return ctx->launch("batch_finalize", [&] {
        hipLaunchKernelGGL(k_batch_finalize, dim3(1), dim3(256), 0, ctx->stream, a, out);
    });
}

// Autotuner: every combination of the wide column tile for the tunable
// lengths this plan transforms and of the binning block size, each timed over
// reps executions of the given catalogue; the fastest is kept.  The results
// of a plan do not depend on these shapes (the same arithmetic per line and
// per mode; the tests compare them bit for bit).
int plan_autotune(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos, long long np, int reps, std::string& report) {
    if (reps < 1) reps = 1;
    // bit b: the wide tile of length 256, 270 or 288 (one bit), 512
    auto tunable = [](int L) { return L == 256 ? 0 : ((L == 270 || L == 288) ? 1 : (L == 512 ? 2 : -1)); };
    unsigned avail = 0;
    for (int a = 0; a < 3; a++) {
        const int b0 = tunable(p->mesh.n[a]), b1 = tunable(p->M[a]);
        if (b0 >= 0) avail |= 1u << b0;
        if (b1 >= 0) avail |= 1u << b1;
    }
    if ((avail & 2u) && p->cfg.bk) avail |= 8u;  // the triple stage's shell x pass on its own
    if (p->cfg.precision == HM_FP64) avail = 0;   // the wide tiles exist for FP32 only
    const int bins[3] = {256, 128, 64};
    const unsigned keep_wide = p->tune_wide;
    const int keep_bin = p->tune_bin;
    double best = 1e300;
    unsigned best_wide = 0;
    int best_bin = 256;
    char buf[128];
    report.clear();
    snprintf(buf, sizeof(buf), "lds per workgroup %zu B; ", ctx->lds_max);
    report += buf;
    for (unsigned w = 0; w < 16; w++) {
        if (w & ~avail) continue;
        for (int bb : bins) {
            p->tune_wide = w;
            p->tune_bin = bb;
            if (plan_execute_any(ctx, p, d_pos, np, nullptr, nullptr, 0, 0.0)) {  // warm-up, lazy allocations
                p->tune_wide = keep_wide;
                p->tune_bin = keep_bin;
                return -1;
            }
            const auto t0 = std::chrono::steady_clock::now();
            for (int r = 0; r < reps; r++)
                if (plan_execute_any(ctx, p, d_pos, np, nullptr, nullptr, 0, 0.0)) {
                    p->tune_wide = keep_wide;
                    p->tune_bin = keep_bin;
                    return -1;
                }
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
            snprintf(buf, sizeof(buf), "wide=%u%u%u%u bin=%d: %.3f ms; ", w & 1u, (w >> 1) & 1u, (w >> 2) & 1u,
                     (w >> 3) & 1u, bb, ms);
            report += buf;
            if (ms < best) {
                best = ms;
                best_wide = w;
                best_bin = bb;
            }
        }
    }
    p->tune_wide = best_wide;
    p->tune_bin = best_bin;
    snprintf(buf, sizeof(buf), "chosen: wide=%u%u%u%u (256, 270/288, 512, shell x pass) bin=%d, %.3f ms",
             best_wide & 1u, (best_wide >> 1) & 1u, (best_wide >> 2) & 1u, (best_wide >> 3) & 1u, best_bin, best);
    report += buf;
    return 0;
}

void plan_tune_get(const BoxPlanImpl* p, unsigned* wide, int* bin) {
    *wide = p->tune_wide;
    *bin = p->tune_bin;
}
void plan_tune_set(BoxPlanImpl* p, unsigned wide, int bin) {
    p->tune_wide = wide & 15u;
    p->tune_bin = (bin == 64 || bin == 128) ? bin : 256;
}

size_t plan_bytes(const BoxPlanImpl* p) { return p->bytes; }
void plan_triple_mesh(const BoxPlanImpl* p, int m[3]) {
    for (int a = 0; a < 3; a++) m[a] = p->M[a];
}

}  // namespace hm
