// hm_seams.hip -- the reference's internal seams one at a time, on its own
// grid layout: complex FP64 [n0][n1][n2], (re, im) interleaved as fftw_complex,
// over the whole mesh (the plans work on half spectra, re-gridded and pruned;
// these entry points do what one DensityFieldClass / TwoPointStatisticsClass
// member does, so that a caller can replace a single seam).
//
// Every sum is deterministic: each lane owns a fixed set of mesh points
// (one z index, rows g, g + G, g + 2G, ...), accumulates its partial sums in
// a fixed order and the partials are reduced column by column in a fixed
// order.  Lanes of a wavefront read consecutive z, so loads coalesce.
#include "hm_assign.hpp"
#include "hm_bessel.hpp"
#include "hm_bins.hpp"
#include "hm_internal.hpp"

#include "../../include/hitomi_amd.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace hm {
namespace {

struct Seam {
    Mesh m;
    long long nrows = 0;  // n0 * n1
};

int seam_bad(hm_ctx* ctx, const std::string& msg) {
    if (ctx) ctx->set_error(msg);
    return -1;
}

bool seam_mesh(hm_ctx* ctx, const hm_mesh* mesh, Seam& s, const char* who) {
    if (!mesh) return seam_bad(ctx, std::string(who) + ": mesh is NULL"), false;
    for (int a = 0; a < 3; a++) {
        if (mesh->n_mesh[a] <= 0) return seam_bad(ctx, std::string(who) + ": n_mesh must be > 0"), false;
        if (!(mesh->boxsize[a] > 0.0)) return seam_bad(ctx, std::string(who) + ": boxsize must be > 0"), false;
        s.m.n[a] = mesh->n_mesh[a];
        s.m.L[a] = mesh->boxsize[a];
    }
    if (mesh->assign < HM_NGP || mesh->assign > HM_TSC)
        return seam_bad(ctx, std::string(who) + ": assign must be NGP, CIC or TSC"), false;
    s.m.assign = mesh->assign;
    s.nrows = (long long)s.m.n[0] * s.m.n[1];
    return true;
}

bool seam_fft_ok(hm_ctx* ctx, const Seam& s, const char* who) {
    for (int a = 0; a < 3; a++)
        if (!fft_length_supported(s.m.n[a]))
            return seam_bad(ctx, std::string(who) + ": mesh length " + std::to_string(s.m.n[a]) +
                                     " has no FFT plan (lengths 2..2048)"),
                   false;
    return true;
}

// Per-axis k, window and shot-noise tables (host_axis_tables) in device memory.
struct SeamTabs {
    DevBuf buf;
    AxisTabs t;
};

int seam_tabs(hm_ctx* ctx, const Seam& s, SeamTabs& st) {
    std::vector<double> h;
    long long off[3];
    host_axis_tables(s.m, h, off);
    if (dev_alloc(ctx, st.buf, h.size() * sizeof(double))) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(st.buf.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                                     ctx->stream));
    const long long ntot = (long long)s.m.n[0] + s.m.n[1] + s.m.n[2];
    const double* d = (const double*)st.buf.p;
    for (int a = 0; a < 3; a++) {
        st.t.kv[a] = d + off[a];
        st.t.wf[a] = d + ntot + off[a];
        st.t.sf[a] = d + 2 * ntot + off[a];
        st.t.iw[a] = d + 3 * ntot + off[a];
    }
    return 0;
}

// The j_l spline of SphericalBesselClass (bessel.hpp:4-41) on the device,
// built once per l and context.
int seam_spline(hm_ctx* ctx, int l, const double** y, const double** c) {
    if (l < 0 || l > 32) return seam_bad(ctx, "spherical Bessel order must be in [0, 32]");
    auto it = ctx->jl_splines.find(l);
    if (it == ctx->jl_splines.end()) {
        std::vector<double> hy, hc;
        jl_spline(l, hy, hc);
        void* p = nullptr;
        const size_t bytes = 2 * sizeof(double) * (size_t)JL_NODES;
        HM_HIP_CHECK(ctx, hipMalloc(&p, bytes));
        if (hipMemcpy(p, hy.data(), bytes / 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((double*)p + JL_NODES, hc.data(), bytes / 2, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(p);
            return seam_bad(ctx, "copying the j_l spline to the device failed");
        }
        it = ctx->jl_splines.emplace(l, p).first;
    }
    *y = (const double*)it->second;
    *c = (const double*)it->second + JL_NODES;
    return 0;
}

// Lanes: lane t owns z index t % n2 and rows g, g + G, ... with g = t / n2.
struct Lanes {
    int G;
    long long nl;
};
Lanes seam_lanes(const Seam& s) {
    const long long want = 1 << 18;
    long long G = (want + s.m.n[2] - 1) / s.m.n[2];
    if (G > s.nrows) G = s.nrows;
    if (G < 1) G = 1;
    return {(int)G, G * (long long)s.m.n[2]};
}
unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

// Elementwise grid walk for kernels without sums: rows over blockIdx.y (stride), z over x.
dim3 walk_grid(const Seam& s) {
    const unsigned gy = (unsigned)std::min<long long>(s.nrows, 65535);
    return dim3((unsigned)((s.m.n[2] + 255) / 256), gy);
}

__device__ __forceinline__ double seam_window(const AxisTabs& t, int wp, int i, int j, int k) {
#pragma clang fp contract(off)
    return wpow((t.wf[0][i] * t.wf[1][j]) * t.wf[2][k], wp);  // field.hpp:700-716
}
__device__ __forceinline__ double seam_snshape(const AxisTabs& t, int i, int j, int k) {
#pragma clang fp contract(off)
    return (t.sf[0][i] * t.sf[1][j]) * t.sf[2][k];  // field.hpp:1283-1309
}
__device__ __forceinline__ double2 cm(double2 a, double2 b) {
#pragma clang fp contract(off)
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---------------------------------------------------------------------------
// elementwise passes
// ---------------------------------------------------------------------------
// field *= s (calcFourierTransform's dV, field.hpp:313-318) or field /= s
// (calcInverseFourierTransform's V, field.hpp:331-336)
__global__ __launch_bounds__(256) void k_seam_scale(double2* __restrict__ g, long long nrows, int n2, double s, int div) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n2) return;
    for (long long r = blockIdx.y; r < nrows; r += gridDim.y) {
        double2 v = g[r * n2 + k];
        if (div) {
            v.x /= s;
            v.y /= s;
        } else {
            v.x *= s;
            v.y *= s;
        }
        g[r * n2 + k] = v;
    }
}

// field /= W(k) (calcAssignmentFunctionCorrection, field.hpp:610-639)
__global__ __launch_bounds__(256) void k_seam_window(double2* __restrict__ g, int n1, long long nrows, int n2,
                                                    AxisTabs t, int wp) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n2) return;
    for (long long r = blockIdx.y; r < nrows; r += gridDim.y) {
        const int i = (int)(r / n1), j = (int)(r % n1);
        const double wf = seam_window(t, wp, i, j, k);
        double2 v = g[r * n2 + k];
        v.x /= wf;
        v.y /= wf;
        g[r * n2 + k] = v;
    }
}

// out = factor * (f1 conj(f2) - sn S(k)) / W(k)^2 (field.hpp:862-899, 1116-1150)
__global__ __launch_bounds__(256) void k_seam_cross(const double2* __restrict__ f1, const double2* __restrict__ f2,
                                                   double2* __restrict__ out, int n1, long long nrows, int n2,
                                                   AxisTabs t, int wp, double sn_re, double sn_im, double factor) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n2) return;
    for (long long r = blockIdx.y; r < nrows; r += gridDim.y) {
        const int i = (int)(r / n1), j = (int)(r % n1);
        const long long c = r * n2 + k;
        const double2 a = f1[c], b = f2[c];
        double pr = a.x * b.x - a.y * (-b.y);
        double pi = a.x * (-b.y) + a.y * b.x;
        const double S = seam_snshape(t, i, j, k);
        pr -= sn_re * S;
        pi -= sn_im * S;
        const double wf = seam_window(t, wp, i, j, k);
        const double w2 = wf * wf;
        pr /= w2;
        pi /= w2;
        out[c] = make_double2(factor * pr, factor * pi);
    }
}

// Shell field of calcInverseFourierTransformForBispectrum (field.hpp:346-389):
// Y(k) d(k) / W(k) on lower < |k| <= upper, 0 elsewhere; modes counted
__global__ __launch_bounds__(256) void k_seam_shell(const double2* __restrict__ d, double2* __restrict__ out, int n1,
                                                   long long nrows, int n2, AxisTabs t, int wp, double lower,
                                                   double upper, int ell, int m,
                                                   unsigned long long* __restrict__ n_mode) {
#pragma clang fp contract(off)
    __shared__ unsigned cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int k = blockIdx.x * 256 + threadIdx.x;
    unsigned mine = 0;
    if (k < n2) {
        for (long long r = blockIdx.y; r < nrows; r += gridDim.y) {
            const int i = (int)(r / n1), j = (int)(r % n1);
            const long long c = r * n2 + k;
            const double kx = t.kv[0][i], ky = t.kv[1][j], kz = t.kv[2][k];
            const double kmag = kmag3(kx, ky, kz);
            double2 v = make_double2(0.0, 0.0);
            if ((kmag > lower) && (kmag <= upper)) {
                const double wf = seam_window(t, wp, i, j, k);
                double2 dn = d[c];
                dn.x /= wf;
                dn.y /= wf;
                const Cplx y = ylm_red(ell, m, kx, ky, kz);
                v = cm(make_double2(y.re, y.im), dn);
                mine++;
            }
            out[c] = v;
        }
    }
    if (mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(n_mode, (unsigned long long)cnt);
}

// out /= n_mode (field.hpp:395-398); the count is read from device memory
__global__ __launch_bounds__(256) void k_seam_div_count(double2* __restrict__ g, long long n,
                                                       const unsigned long long* __restrict__ n_mode) {
    const double nm = (double)(long long)*n_mode;
    const long long st = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += st) {
        double2 v = g[i];
        v.x /= nm;
        v.y /= nm;
        g[i] = v;
    }
}

// calcInverseFourierTransformForThreePointFunction's field (field.hpp:410-437):
// j_l(|k| r) (Y(k) d(k) / W(k)) / V
__global__ __launch_bounds__(256) void k_seam_bessel(const double2* __restrict__ d, double2* __restrict__ out, int n1,
                                                    long long nrows, int n2, AxisTabs t, int wp, double rmag,
                                                    int ell, int m, const double* __restrict__ sy,
                                                    const double* __restrict__ sc, double volume) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n2) return;
    for (long long r = blockIdx.y; r < nrows; r += gridDim.y) {
        const int i = (int)(r / n1), j = (int)(r % n1);
        const long long c = r * n2 + k;
        const double kx = t.kv[0][i], ky = t.kv[1][j], kz = t.kv[2][k];
        const double kmag = kmag3(kx, ky, kz);
        const double wf = seam_window(t, wp, i, j, k);
        double2 dn = d[c];
        dn.x /= wf;
        dn.y /= wf;
        const Cplx y = ylm_red(ell, m, kx, ky, kz);
        const double2 v = cm(make_double2(y.re, y.im), dn);
        const double jl = jl_spline_eval(sy, sc, kmag * rmag);
        out[c] = make_double2(jl * v.x / volume, jl * v.y / volume);
    }
}

// ---------------------------------------------------------------------------
// binned sums (lane partials, column-major: column c of lane t at c * nl + t)
// ---------------------------------------------------------------------------
struct BinSeamArgs {
    int n[3];
    long long nrows;
    int G;
    long long nl;
    AxisTabs t;
    int wp;
    const int2* f2b;
    int nb;
    double sn_re, sn_im;
    int ell, m;            // Y(l, m) weight (P_l, xi_l)
    int ell2, m2, two_y;   // two_y: the weight Y(l, m) Y(l2, m2) of the 3PCF xi
    double d[3];           // r spacings (xi)
    double fine;           // fine bin width
    int nfine;
};

// calcPowerSpectrum (field.hpp:765-837): fine index int(|k| / 1e-4 + 0.5),
// (f1 conj(f2) - sn S) / W^2 Y(k) into every coarse bin holding the fine centre
__global__ __launch_bounds__(256) void k_seam_bin_pk(const double2* __restrict__ f1, const double2* __restrict__ f2,
                                                    BinSeamArgs a, double* __restrict__ part,
                                                    long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nl) return;
    const int k = (int)(t % a.n[2]);
    const long long g = t / a.n[2];
    for (long long r = g; r < a.nrows; r += a.G) {
        const int i = (int)(r / a.n[1]), j = (int)(r % a.n[1]);
        const double kx = a.t.kv[0][i], ky = a.t.kv[1][j], kz = a.t.kv[2][k];
        const double kmag = kmag3(kx, ky, kz);
        const int kindex = (int)(kmag / FINE_DK + 0.5);
        if (kindex >= FINE_N) continue;
        const int2 b = a.f2b[kindex];
        if (b.x < 0) continue;
        const long long c = r * a.n[2] + k;
        const double2 p = f1[c], q = f2[c];
        double pr = p.x * q.x - p.y * (-q.y);
        double pi = p.x * (-q.y) + p.y * q.x;
        const double S = seam_snshape(a.t, i, j, k);
        pr -= a.sn_re * S;
        pi -= a.sn_im * S;
        const double wf = seam_window(a.t, a.wp, i, j, k);
        const double w2 = wf * wf;
        pr /= w2;
        pi /= w2;
        const Cplx y = ylm_red(a.ell, a.m, kx, ky, kz);
        const double2 v = cm(make_double2(pr, pi), make_double2(y.re, y.im));
        for (int bb = b.x; bb <= b.y; bb++) {
            part[(2 * bb) * a.nl + t] += v.x;
            part[(2 * bb + 1) * a.nl + t] += v.y;
            cnt[bb * a.nl + t] += 1;
        }
    }
}

// r-space binning of calcCorrelationFunction (field.hpp:921-960) and of its
// 3PCF variant (field.hpp:1041-1095): fine index int(|r| / 0.5 + 0.5)
__global__ __launch_bounds__(256) void k_seam_bin_xi(const double2* __restrict__ x, BinSeamArgs a,
                                                    double* __restrict__ part, long long* __restrict__ cnt) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nl) return;
    const int k = (int)(t % a.n[2]);
    const long long g = t / a.n[2];
    const double rz = mesh_vec(k, a.n[2], a.d[2]);
    for (long long r = g; r < a.nrows; r += a.G) {
        const int i = (int)(r / a.n[1]), j = (int)(r % a.n[1]);
        const double rx = mesh_vec(i, a.n[0], a.d[0]), ry = mesh_vec(j, a.n[1], a.d[1]);
        const double rmag = kmag3(rx, ry, rz);
        const int ri = (int)(rmag / a.fine + 0.5);
        if (ri >= a.nfine) continue;
        const int2 b = a.f2b[ri];
        if (b.x < 0) continue;
        const double2 xv = x[r * a.n[2] + k];
        const Cplx y1 = ylm_red(a.ell, a.m, rx, ry, rz);
        double2 w = make_double2(y1.re, y1.im);
        if (a.two_y) {
            const Cplx y2 = ylm_red(a.ell2, a.m2, rx, ry, rz);
            w = cm(w, make_double2(y2.re, y2.im));
        }
        const double2 v = cm(xv, w);
        for (int bb = b.x; bb <= b.y; bb++) {
            part[(2 * bb) * a.nl + t] += v.x;
            part[(2 * bb + 1) * a.nl + t] += v.y;
            cnt[bb * a.nl + t] += 1;
        }
    }
}

// Sum over the mesh of dV F_a F_b F3 for a chunk of pairs (bispec.hpp:1159-1168)
constexpr int SEAM_PCH = 16;
struct TripleArgs {
    int n2;
    long long nrows;
    int G;
    long long nl;
    const double2* const* f;  // device array of field pointers
    const double2* f3;
    int pa[SEAM_PCH], pb[SEAM_PCH];
    int np;
    double fac;
};
__global__ __launch_bounds__(256) void k_seam_triple(TripleArgs a, double* __restrict__ part) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nl) return;
    const int k = (int)(t % a.n2);
    const long long g = t / a.n2;
    double sr[SEAM_PCH], si[SEAM_PCH];
#pragma unroll
    for (int p = 0; p < SEAM_PCH; p++) sr[p] = si[p] = 0.0;
    for (long long r = g; r < a.nrows; r += a.G) {
        const long long c = r * a.n2 + k;
        const double2 f3 = a.f3[c];
#pragma unroll
        for (int p = 0; p < SEAM_PCH; p++) {
            if (p < a.np) {
                const double2 x = a.f[a.pa[p]][c], y = a.f[a.pb[p]][c];
                // (fac f1) f2 f3 in the order of bispec.hpp:1166
                const double2 u = cm(cm(make_double2(a.fac * x.x, a.fac * x.y), y), f3);
                sr[p] += u.x;
                si[p] += u.y;
            }
        }
    }
#pragma unroll
    for (int p = 0; p < SEAM_PCH; p++) {
        if (p < a.np) {
            part[(2 * p) * a.nl + t] = sr[p];
            part[(2 * p + 1) * a.nl + t] = si[p];
        }
    }
}

// The shot-noise ijk sweep of bispec.hpp:1036-1056 for a chunk of (k1, k2)
// pairs: sum_x j_l1(k1 r) j_l2(k2 r) xi(x) Y1(r) Y2(r)
struct SweepArgs {
    int n[3];
    long long nrows;
    int G;
    long long nl;
    double d[3];
    int l1, m1, l2, m2;
    const double *y1, *c1, *y2, *c2;  // j_l1, j_l2 splines
    double k1[SEAM_PCH], k2[SEAM_PCH];
    int np;
};
__global__ __launch_bounds__(256) void k_seam_sweep(const double2* __restrict__ xi, SweepArgs a,
                                                   double* __restrict__ part) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nl) return;
    const int k = (int)(t % a.n[2]);
    const long long g = t / a.n[2];
    const double rz = mesh_vec(k, a.n[2], a.d[2]);
    double sr[SEAM_PCH], si[SEAM_PCH];
#pragma unroll
    for (int p = 0; p < SEAM_PCH; p++) sr[p] = si[p] = 0.0;
    for (long long r = g; r < a.nrows; r += a.G) {
        const int i = (int)(r / a.n[1]), j = (int)(r % a.n[1]);
        const double rx = mesh_vec(i, a.n[0], a.d[0]), ry = mesh_vec(j, a.n[1], a.d[1]);
        const double rmag = kmag3(rx, ry, rz);
        const Cplx y1 = ylm_red(a.l1, a.m1, rx, ry, rz);
        const Cplx y2 = ylm_red(a.l2, a.m2, rx, ry, rz);
        const double2 w = cm(cm(xi[r * a.n[2] + k], make_double2(y1.re, y1.im)), make_double2(y2.re, y2.im));
#pragma unroll
        for (int p = 0; p < SEAM_PCH; p++) {
            if (p < a.np) {
                const double j12 =
                    jl_spline_eval(a.y1, a.c1, a.k1[p] * rmag) * jl_spline_eval(a.y2, a.c2, a.k2[p] * rmag);
                sr[p] += j12 * w.x;
                si[p] += j12 * w.y;
            }
        }
    }
#pragma unroll
    for (int p = 0; p < SEAM_PCH; p++) {
        if (p < a.np) {
            part[(2 * p) * a.nl + t] = sr[p];
            part[(2 * p + 1) * a.nl + t] = si[p];
        }
    }
}

// Particle sums of field.hpp:1164-1240: sum_p w_p^power Y(l, m)(los_p), and
// sum_p |w_p| (for the fixed-point scale); grid-stride lanes with a fixed
// assignment, then an LDS tree per workgroup (fixed order)
constexpr int PSUM_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_seam_psum(const double* __restrict__ los, const double* __restrict__ w,
                                                  long long n, int ell, int m, int power,
                                                  double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double s[3][256];
    double ar = 0.0, ai = 0.0, aa = 0.0;
    const long long st = (long long)gridDim.x * 256;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += st) {
        const double wp = w ? w[p] : 1.0;
        double f = 1.0;
        for (int e = 0; e < power; e++) f *= wp;
        Cplx y = {1.0, 0.0};
        if (los) y = ylm_red(ell, m, los[3 * p], los[3 * p + 1], los[3 * p + 2]);
        ar += f * y.re;
        ai += f * y.im;
        aa += fabs(wp);
    }
    s[0][threadIdx.x] = ar;
    s[1][threadIdx.x] = ai;
    s[2][threadIdx.x] = aa;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] += s[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[threadIdx.x * PSUM_BLOCKS + blockIdx.x] = s[threadIdx.x][0];
}

// sum_x rho(x)^2 dV of calcSurveyVolume (field.hpp:488-492), per-workgroup partials
__global__ __launch_bounds__(256) void k_seam_sq(const double* __restrict__ rho, long long n, double dV,
                                                double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double s[256];
    double acc = 0.0;
    const long long st = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += st) acc += rho[i] * rho[i] * dV;
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// (w, 0) complex weights
__global__ __launch_bounds__(256) void k_seam_cweights(const double* __restrict__ w, long long n,
                                                      double* __restrict__ out) {
    const long long st = (long long)gridDim.x * 256;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += st) {
        out[2 * p] = w[p];
        out[2 * p + 1] = 0.0;
    }
}

// int64 fixed-point cells -> interleaved complex density (field.hpp:172-180 / dV)
__global__ __launch_bounds__(256) void k_seam_fixed_to_grid(const long long* __restrict__ re,
                                                           const long long* __restrict__ im, long long n,
                                                           double inv_scale, double factor,
                                                           double2* __restrict__ out) {
#pragma clang fp contract(off)
    const long long st = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += st) {
        const double a = ((double)re[i] * inv_scale) * factor;
        const double b = im ? ((double)im[i] * inv_scale) * factor : 0.0;
        out[i] = make_double2(a, b);
    }
}

unsigned stride_blocks(long long n) { return (unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 16384)); }

// Fine -> coarse map in device memory.
int seam_f2b(hm_ctx* ctx, const std::vector<double>& lower, const std::vector<double>& upper, int nfine, double fine,
             DevBuf& buf, const char* who) {
    std::vector<int2> f2b;
    const std::string why = fine_to_coarse(lower, upper, nfine, fine, f2b);
    if (!why.empty()) return seam_bad(ctx, std::string(who) + ": " + why);
    if (dev_alloc(ctx, buf, f2b.size() * sizeof(int2))) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(buf.p, f2b.data(), f2b.size() * sizeof(int2), hipMemcpyHostToDevice,
                                     ctx->stream));
    return 0;
}

// Partials [ncols][nl] doubles + [nb][nl] counts, zeroed; reduced to host sums.
struct Partials {
    DevBuf part, cnt, red;
};
int partials_alloc(hm_ctx* ctx, Partials& P, long long nl, int ncols, int ncnt) {
    if (dev_alloc(ctx, P.part, sizeof(double) * (size_t)nl * ncols)) return -1;
    if (fill_zero(ctx, P.part.p, sizeof(double) * (size_t)nl * ncols)) return -1;
    if (ncnt > 0) {
        if (dev_alloc(ctx, P.cnt, sizeof(long long) * (size_t)nl * ncnt)) return -1;
        if (fill_zero(ctx, P.cnt.p, sizeof(long long) * (size_t)nl * ncnt)) return -1;
    }
    return dev_alloc(ctx, P.red, sizeof(double) * (size_t)(ncols + ncnt));
}
int partials_reduce(hm_ctx* ctx, Partials& P, long long nl, int ncols, int ncnt, double* sums, long long* counts) {
    double* rd = (double*)P.red.p;
    if (reduce_cols_cm<double>(ctx, "seam_reduce", (const double*)P.part.p, nl, ncols, rd)) return -1;
    if (ncnt > 0 &&
        reduce_cols_cm<long long>(ctx, "seam_reduce_cnt", (const long long*)P.cnt.p, nl, ncnt, (long long*)(rd + ncols)))
        return -1;
    std::vector<double> h(ncols + ncnt);
    HM_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), rd, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    std::memcpy(sums, h.data(), sizeof(double) * ncols);
    if (ncnt > 0) std::memcpy(counts, h.data() + ncols, sizeof(long long) * ncnt);
    return 0;
}

int finish(hm_ctx* ctx, int rc) {
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) return seam_bad(ctx, "stream error");
    return rc;
}

// xi binning on a transformed grid x; kind 0: calcCorrelationFunction edges
// and averages (field.hpp:948-968), kind 1: the 3PCF edges and scaling
// (field.hpp:1065-1105)
int seam_bin_xi(hm_ctx* ctx, const Seam& s, const double2* x, const double* rbin, int n_rbin, int ell, int m, int ell2,
                int m2, int kind, double* xi, long long* n_mode, const char* who) {
    std::vector<double> lower(n_rbin), upper(n_rbin);
    if (kind == 0) {
        const double dr_bin = rbin[1] - rbin[0];
        for (int j = 0; j < n_rbin; j++) {
            lower[j] = (rbin[j] > dr_bin * 0.5) ? (rbin[j] - dr_bin * 0.5) : 0.0;
            upper[j] = rbin[j] + dr_bin * 0.5;
        }
    } else {
        std::vector<double> drb(n_rbin - 1);
        for (int j = 0; j < n_rbin - 1; j++) drb[j] = rbin[j + 1] - rbin[j];
        for (int j = 0; j < n_rbin; j++) {
            lower[j] = (j == 0) ? ((rbin[j] > drb[j] * 0.5) ? (rbin[j] - drb[j] * 0.5) : 0.0) : rbin[j] - drb[j - 1] * 0.5;
            upper[j] = (j == n_rbin - 1) ? (rbin[j] + drb[j - 1] * 0.5) : (rbin[j] + drb[j] * 0.5);
        }
    }
    const int nfine = kind == 0 ? FINE_NR : FINE_NR / 2;  // field.hpp:906, 1027
    DevBuf f2b;
    Partials P;
    const Lanes ln = seam_lanes(s);
    int rc = seam_f2b(ctx, lower, upper, nfine, FINE_DR, f2b, who);
    if (!rc) rc = partials_alloc(ctx, P, ln.nl, 2 * n_rbin, n_rbin);
    if (!rc) {
        BinSeamArgs a{};
        for (int q = 0; q < 3; q++) {
            a.n[q] = s.m.n[q];
            a.d[q] = s.m.L[q] / (double)s.m.n[q];
        }
        a.nrows = s.nrows;
        a.G = ln.G;
        a.nl = ln.nl;
        a.f2b = (const int2*)f2b.p;
        a.nb = n_rbin;
        a.ell = ell;
        a.m = m;
        a.ell2 = ell2;
        a.m2 = m2;
        a.two_y = kind;
        a.fine = FINE_DR;
        a.nfine = nfine;
        rc = ctx->launch("seam_bin_xi", [&] {
            hipLaunchKernelGGL(k_seam_bin_xi, dim3(blocks_for(ln.nl)), dim3(256), 0, ctx->stream, x, a,
                               (double*)P.part.p, (long long*)P.cnt.p);
        });
    }
    std::vector<double> sums(2 * n_rbin);
    std::vector<long long> cnt(n_rbin);
    if (!rc) rc = partials_reduce(ctx, P, ln.nl, 2 * n_rbin, n_rbin, sums.data(), cnt.data());
    if (!rc) {
        const double dV = s.m.volume() / (double)s.m.ng();
        const double sgn = ((ell + ell2) % 2) ? -1.0 : 1.0;  // pow(-1.0, ell1 + ell2)
        for (int j = 0; j < n_rbin; j++) {
            double re = sums[2 * j], im = sums[2 * j + 1];
            if (cnt[j] != 0) {
                if (kind == 0) {
                    re /= (double)cnt[j];
                    im /= (double)cnt[j];
                } else {
                    const double f = sgn / (double)cnt[j] / (double)cnt[j] / dV;
                    re *= f;
                    im *= f;
                }
            } else {
                re = im = 0.0;
            }
            xi[2 * j] = re;
            xi[2 * j + 1] = im;
            if (n_mode) n_mode[j] = cnt[j];
        }
    }
    rc = finish(ctx, rc);
    dev_free(f2b);
    dev_free(P.part);
    dev_free(P.cnt);
    dev_free(P.red);
    return rc;
}

int seam_cross_fft(hm_ctx* ctx, const Seam& s, const double2* f1, const double2* f2, const double sn[2], double2* out) {
    SeamTabs tb;
    int rc = seam_tabs(ctx, s, tb);
    if (!rc)
        rc = ctx->launch("seam_cross", [&] {
            hipLaunchKernelGGL(k_seam_cross, walk_grid(s), dim3(256), 0, ctx->stream, f1, f2, out, s.m.n[1], s.nrows,
                               s.m.n[2], tb.t, wpow_of(s.m.assign), sn ? sn[0] : 0.0, sn ? sn[1] : 0.0,
                               1.0 / s.m.volume());
        });
    if (!rc) rc = fft3d_c2c<double>(ctx, out, s.m.n, +1);
    rc = finish(ctx, rc);
    dev_free(tb.buf);
    return rc;
}

}  // namespace
}  // namespace hm

using namespace hm;

extern "C" {

double* hm_grid_alloc(hm_ctx* ctx, const hm_mesh* mesh) {
    if (!ctx) return nullptr;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_grid_alloc")) return nullptr;
    DevBuf b;
    const size_t bytes = sizeof(double2) * (size_t)s.m.ng();
    if (dev_alloc(ctx, b, bytes)) return nullptr;
    if (fill_zero(ctx, b.p, bytes) || hipStreamSynchronize(ctx->stream) != hipSuccess) {
        dev_free(b);
        seam_bad(ctx, "hm_grid_alloc: zeroing the grid failed");
        return nullptr;
    }
    return (double*)b.p;
}

int hm_grid_free(hm_ctx* ctx, double* d_grid) {
    if (!ctx) return -1;
    if (!d_grid) return 0;
    (void)hipStreamSynchronize(ctx->stream);
    HM_HIP_CHECK(ctx, hipFree(d_grid));
    return 0;
}

int hm_assign_field(hm_ctx* ctx, const hm_mesh* mesh, const double* d_pos, const double* d_w, long long np,
                    double wabs_sum, double* d_grid) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_assign_field")) return -1;
    if (np < 0 || (np > 0 && !d_pos) || !d_grid) return seam_bad(ctx, "hm_assign_field: bad arguments");
    AssignGeom g;
    for (int a = 0; a < 3; a++) {
        g.n[a] = s.m.n[a];
        g.L[a] = s.m.L[a];
    }
    g.ng = s.m.ng();
    g.assign = s.m.assign;
    if (!(wabs_sum > 0.0)) wabs_sum = (double)np;
    const double scale = fixed_scale_for(wabs_sum);
    DevBuf q;
    SortWork w;
    int rc = dev_alloc(ctx, q, 2 * sizeof(long long) * (size_t)g.ng);
    long long* qre = (long long*)q.p;
    long long* qim = qre ? qre + g.ng : nullptr;
    if (!rc) rc = sort_work_alloc(ctx, w, g, np, d_w != nullptr);
    if (!rc) rc = assign_sorted(ctx, g, d_pos, d_w, np, scale, qre, d_w ? qim : nullptr, w);
    const double factor = 1.0 / (s.m.volume() / (double)g.ng);
    if (!rc)
        rc = ctx->launch("seam_fixed_to_grid", [&] {
            hipLaunchKernelGGL(k_seam_fixed_to_grid, dim3(stride_blocks(g.ng)), dim3(256), 0, ctx->stream, qre,
                               d_w ? qim : nullptr, g.ng, 1.0 / scale, factor, (double2*)d_grid);
        });
    rc = finish(ctx, rc);
    sort_work_free(w);
    dev_free(q);
    return rc;
}

int hm_fourier_transform(hm_ctx* ctx, const hm_mesh* mesh, double* d_grid, int sign) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_fourier_transform") || !seam_fft_ok(ctx, s, "hm_fourier_transform")) return -1;
    if (!d_grid || (sign != -1 && sign != 1)) return seam_bad(ctx, "hm_fourier_transform: bad arguments");
    const double v = sign < 0 ? s.m.volume() / (double)s.m.ng() : s.m.volume();
    int rc = ctx->launch("seam_scale", [&] {
        hipLaunchKernelGGL(k_seam_scale, walk_grid(s), dim3(256), 0, ctx->stream, (double2*)d_grid, s.nrows, s.m.n[2],
                           v, sign > 0 ? 1 : 0);
    });
    if (!rc) rc = fft3d_c2c<double>(ctx, (double2*)d_grid, s.m.n, sign);
    return finish(ctx, rc);
}

int hm_window_correct(hm_ctx* ctx, const hm_mesh* mesh, double* d_grid) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_window_correct")) return -1;
    if (!d_grid) return seam_bad(ctx, "hm_window_correct: bad arguments");
    SeamTabs tb;
    int rc = seam_tabs(ctx, s, tb);
    if (!rc)
        rc = ctx->launch("seam_window", [&] {
            hipLaunchKernelGGL(k_seam_window, walk_grid(s), dim3(256), 0, ctx->stream, (double2*)d_grid, s.m.n[1],
                               s.nrows, s.m.n[2], tb.t, wpow_of(s.m.assign));
        });
    rc = finish(ctx, rc);
    dev_free(tb.buf);
    return rc;
}

int hm_bin_pk(hm_ctx* ctx, const hm_mesh* mesh, const double* d_f1, const double* d_f2, const double* kbin,
              int n_kbin, const double sn[2], int ell, int m, double* pk, long long* n_mode) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_bin_pk")) return -1;
    if (!d_f1 || !d_f2 || !kbin || n_kbin < 2 || !pk || ell < 0 || ell > 16 || m < -ell || m > ell)
        return seam_bad(ctx, "hm_bin_pk: bad arguments (n_kbin >= 2, 0 <= ell <= 16, |m| <= ell)");
    // bins of field.hpp:826-837
    const double dk_bin = kbin[1] - kbin[0];
    std::vector<double> lower(n_kbin), upper(n_kbin);
    for (int j = 0; j < n_kbin; j++) {
        lower[j] = (kbin[j] > dk_bin * 0.5) ? (kbin[j] - dk_bin * 0.5) : 0.0;
        upper[j] = kbin[j] + dk_bin * 0.5;
    }
    SeamTabs tb;
    DevBuf f2b;
    Partials P;
    const Lanes ln = seam_lanes(s);
    int rc = seam_tabs(ctx, s, tb);
    if (!rc) rc = seam_f2b(ctx, lower, upper, FINE_N, FINE_DK, f2b, "hm_bin_pk");
    if (!rc) rc = partials_alloc(ctx, P, ln.nl, 2 * n_kbin, n_kbin);
    if (!rc) {
        BinSeamArgs a{};
        for (int q = 0; q < 3; q++) a.n[q] = s.m.n[q];
        a.nrows = s.nrows;
        a.G = ln.G;
        a.nl = ln.nl;
        a.t = tb.t;
        a.wp = wpow_of(s.m.assign);
        a.f2b = (const int2*)f2b.p;
        a.nb = n_kbin;
        a.sn_re = sn ? sn[0] : 0.0;
        a.sn_im = sn ? sn[1] : 0.0;
        a.ell = ell;
        a.m = m;
        rc = ctx->launch("seam_bin_pk", [&] {
            hipLaunchKernelGGL(k_seam_bin_pk, dim3(blocks_for(ln.nl)), dim3(256), 0, ctx->stream,
                               (const double2*)d_f1, (const double2*)d_f2, a, (double*)P.part.p,
                               (long long*)P.cnt.p);
        });
    }
    std::vector<double> sums(2 * n_kbin);
    std::vector<long long> cnt(n_kbin);
    if (!rc) rc = partials_reduce(ctx, P, ln.nl, 2 * n_kbin, n_kbin, sums.data(), cnt.data());
    if (!rc)
        for (int j = 0; j < n_kbin; j++) {  // field.hpp:839-845
            const bool z = cnt[j] == 0;
            pk[2 * j] = z ? 0.0 : sums[2 * j] / (double)cnt[j];
            pk[2 * j + 1] = z ? 0.0 : sums[2 * j + 1] / (double)cnt[j];
            if (n_mode) n_mode[j] = cnt[j];
        }
    rc = finish(ctx, rc);
    dev_free(tb.buf);
    dev_free(f2b);
    dev_free(P.part);
    dev_free(P.cnt);
    dev_free(P.red);
    return rc;
}

int hm_sn_field_ijk(hm_ctx* ctx, const hm_mesh* mesh, const double* d_f1, const double* d_f2, const double sn[2],
                    double* d_xi) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_sn_field_ijk") || !seam_fft_ok(ctx, s, "hm_sn_field_ijk")) return -1;
    if (!d_f1 || !d_f2 || !d_xi) return seam_bad(ctx, "hm_sn_field_ijk: bad arguments");
    return seam_cross_fft(ctx, s, (const double2*)d_f1, (const double2*)d_f2, sn, (double2*)d_xi);
}

static int bin_xi_common(hm_ctx* ctx, const hm_mesh* mesh, const double* d_f1, const double* d_f2,
                         const double* rbin, int n_rbin, const double sn[2], int ell, int m, int ell2, int m2,
                         int kind, double* xi, long long* n_mode, const char* who) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, who) || !seam_fft_ok(ctx, s, who)) return -1;
    if (!d_f1 || !d_f2 || !rbin || n_rbin < 2 || !xi || ell < 0 || ell > 16 || m < -ell || m > ell || ell2 < 0 ||
        ell2 > 16 || m2 < -ell2 || m2 > ell2)
        return seam_bad(ctx, std::string(who) + ": bad arguments (n_rbin >= 2, 0 <= ell <= 16, |m| <= ell)");
    DevBuf x;
    int rc = dev_alloc(ctx, x, sizeof(double2) * (size_t)s.m.ng());
    if (!rc) rc = seam_cross_fft(ctx, s, (const double2*)d_f1, (const double2*)d_f2, sn, (double2*)x.p);
    if (!rc) rc = seam_bin_xi(ctx, s, (const double2*)x.p, rbin, n_rbin, ell, m, ell2, m2, kind, xi, n_mode, who);
    rc = finish(ctx, rc);
    dev_free(x);
    return rc;
}

int hm_bin_xi(hm_ctx* ctx, const hm_mesh* mesh, const double* d_f1, const double* d_f2, const double* rbin,
              int n_rbin, const double sn[2], int ell, int m, double* xi, long long* n_mode) {
    return bin_xi_common(ctx, mesh, d_f1, d_f2, rbin, n_rbin, sn, ell, m, 0, 0, 0, xi, n_mode, "hm_bin_xi");
}

int hm_bin_xi_3pcf(hm_ctx* ctx, const hm_mesh* mesh, const double* d_f1, const double* d_f2, const double* rbin,
                   int n_rbin, const double sn[2], int ell1, int m1, int ell2, int m2, double* xi,
                   long long* n_mode) {
    return bin_xi_common(ctx, mesh, d_f1, d_f2, rbin, n_rbin, sn, ell1, m1, ell2, m2, 1, xi, n_mode,
                         "hm_bin_xi_3pcf");
}

int hm_shell_c2r(hm_ctx* ctx, const hm_mesh* mesh, const double* d_delta, double kmag, double dk, int ell, int m,
                 double* d_out, long long* n_mode) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_shell_c2r") || !seam_fft_ok(ctx, s, "hm_shell_c2r")) return -1;
    if (!d_delta || !d_out || d_delta == d_out || ell < 0 || ell > 16 || m < -ell || m > ell)
        return seam_bad(ctx, "hm_shell_c2r: bad arguments (d_out distinct from d_delta, 0 <= ell <= 16, |m| <= ell)");
    const double lower = (kmag > dk * 0.5) ? (kmag - dk * 0.5) : 0.0;  // field.hpp:371-372
    const double upper = kmag + dk * 0.5;
    SeamTabs tb;
    DevBuf cnt;
    int rc = seam_tabs(ctx, s, tb);
    if (!rc) rc = dev_alloc(ctx, cnt, sizeof(unsigned long long));
    if (!rc) rc = fill_zero(ctx, cnt.p, sizeof(unsigned long long));
    if (!rc)
        rc = ctx->launch("seam_shell", [&] {
            hipLaunchKernelGGL(k_seam_shell, walk_grid(s), dim3(256), 0, ctx->stream, (const double2*)d_delta,
                               (double2*)d_out, s.m.n[1], s.nrows, s.m.n[2], tb.t, wpow_of(s.m.assign), lower, upper,
                               ell, m, (unsigned long long*)cnt.p);
        });
    if (!rc) rc = fft3d_c2c<double>(ctx, (double2*)d_out, s.m.n, +1);
    if (!rc)
        rc = ctx->launch("seam_div_count", [&] {
            hipLaunchKernelGGL(k_seam_div_count, dim3(stride_blocks(s.m.ng())), dim3(256), 0, ctx->stream,
                               (double2*)d_out, s.m.ng(), (const unsigned long long*)cnt.p);
        });
    unsigned long long h = 0;
    if (!rc && hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = seam_bad(ctx, "hm_shell_c2r: reading the mode count failed");
    rc = finish(ctx, rc);
    if (!rc && n_mode) *n_mode = (long long)h;
    dev_free(tb.buf);
    dev_free(cnt);
    return rc;
}

int hm_bessel_c2r(hm_ctx* ctx, const hm_mesh* mesh, const double* d_delta, double rmag, int ell_j, int ell, int m,
                  double* d_out) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_bessel_c2r") || !seam_fft_ok(ctx, s, "hm_bessel_c2r")) return -1;
    if (!d_delta || !d_out || d_delta == d_out || !(rmag >= 0.0) || ell < 0 || ell > 16 || m < -ell || m > ell)
        return seam_bad(ctx, "hm_bessel_c2r: bad arguments (d_out distinct from d_delta, rmag >= 0, 0 <= ell <= 16)");
    const double *sy, *sc;
    if (seam_spline(ctx, ell_j, &sy, &sc)) return -1;
    SeamTabs tb;
    int rc = seam_tabs(ctx, s, tb);
    if (!rc)
        rc = ctx->launch("seam_bessel", [&] {
            hipLaunchKernelGGL(k_seam_bessel, walk_grid(s), dim3(256), 0, ctx->stream, (const double2*)d_delta,
                               (double2*)d_out, s.m.n[1], s.nrows, s.m.n[2], tb.t, wpow_of(s.m.assign), rmag, ell, m,
                               sy, sc, s.m.volume());
        });
    if (!rc) rc = fft3d_c2c<double>(ctx, (double2*)d_out, s.m.n, +1);
    rc = finish(ctx, rc);
    dev_free(tb.buf);
    return rc;
}

int hm_triple_sums(hm_ctx* ctx, const hm_mesh* mesh, const double* const* d_fields, int n_fields, const int* pairs,
                   int n_pairs, const double* d_f3, double* out) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_triple_sums")) return -1;
    if (!d_fields || n_fields <= 0 || !pairs || n_pairs < 0 || !d_f3 || (n_pairs > 0 && !out))
        return seam_bad(ctx, "hm_triple_sums: bad arguments");
    for (int f = 0; f < n_fields; f++)
        if (!d_fields[f]) return seam_bad(ctx, "hm_triple_sums: NULL field");
    for (int p = 0; p < 2 * n_pairs; p++)
        if (pairs[p] < 0 || pairs[p] >= n_fields) return seam_bad(ctx, "hm_triple_sums: pair index out of range");
    if (n_pairs == 0) return 0;
    const Lanes ln = seam_lanes(s);
    DevBuf fp;
    Partials P;
    int rc = dev_alloc(ctx, fp, sizeof(void*) * (size_t)n_fields);
    if (!rc)
        rc = hipMemcpyAsync(fp.p, d_fields, sizeof(void*) * (size_t)n_fields, hipMemcpyHostToDevice, ctx->stream) ==
                     hipSuccess
                 ? 0
                 : seam_bad(ctx, "hm_triple_sums: copying the field table failed");
    if (!rc) rc = partials_alloc(ctx, P, ln.nl, 2 * SEAM_PCH, 0);
    const double fac = s.m.volume() / (double)s.m.ng();
    for (int p0 = 0; !rc && p0 < n_pairs; p0 += SEAM_PCH) {
        TripleArgs a{};
        a.n2 = s.m.n[2];
        a.nrows = s.nrows;
        a.G = ln.G;
        a.nl = ln.nl;
        a.f = (const double2* const*)fp.p;
        a.f3 = (const double2*)d_f3;
        a.np = std::min(SEAM_PCH, n_pairs - p0);
        for (int q = 0; q < a.np; q++) {
            a.pa[q] = pairs[2 * (p0 + q)];
            a.pb[q] = pairs[2 * (p0 + q) + 1];
        }
        a.fac = fac;
        rc = ctx->launch("seam_triple", [&] {
            hipLaunchKernelGGL(k_seam_triple, dim3(blocks_for(ln.nl)), dim3(256), 0, ctx->stream, a,
                               (double*)P.part.p);
        });
        std::vector<double> sums(2 * SEAM_PCH);
        if (!rc) rc = partials_reduce(ctx, P, ln.nl, 2 * a.np, 0, sums.data(), nullptr);
        if (!rc) std::memcpy(out + 2 * p0, sums.data(), sizeof(double) * 2 * a.np);
    }
    rc = finish(ctx, rc);
    dev_free(fp);
    dev_free(P.part);
    dev_free(P.red);
    return rc;
}

int hm_sn_bessel_sweep(hm_ctx* ctx, const hm_mesh* mesh, const double* d_xi, int ell1, int m1, int ell2, int m2,
                       const double* k1, const double* k2, int n_pairs, double* out) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_sn_bessel_sweep")) return -1;
    if (!d_xi || n_pairs < 0 || (n_pairs > 0 && (!k1 || !k2 || !out)) || ell1 < 0 || ell1 > 16 || ell2 < 0 ||
        ell2 > 16 || m1 < -ell1 || m1 > ell1 || m2 < -ell2 || m2 > ell2)
        return seam_bad(ctx, "hm_sn_bessel_sweep: bad arguments");
    if (n_pairs == 0) return 0;
    const double *y1, *c1, *y2, *c2;
    if (seam_spline(ctx, ell1, &y1, &c1) || seam_spline(ctx, ell2, &y2, &c2)) return -1;
    const Lanes ln = seam_lanes(s);
    Partials P;
    int rc = partials_alloc(ctx, P, ln.nl, 2 * SEAM_PCH, 0);
    // i^(l1 + l2) dV (bispec.hpp:1052-1054), exact powers of i
    const double fac = s.m.volume() / (double)s.m.ng();
    const int q4 = (ell1 + ell2) % 4;
    const double ir = q4 == 0 ? 1.0 : (q4 == 2 ? -1.0 : 0.0), ii = q4 == 1 ? 1.0 : (q4 == 3 ? -1.0 : 0.0);
    for (int p0 = 0; !rc && p0 < n_pairs; p0 += SEAM_PCH) {
        SweepArgs a{};
        for (int q = 0; q < 3; q++) {
            a.n[q] = s.m.n[q];
            a.d[q] = s.m.L[q] / (double)s.m.n[q];
        }
        a.nrows = s.nrows;
        a.G = ln.G;
        a.nl = ln.nl;
        a.l1 = ell1;
        a.m1 = m1;
        a.l2 = ell2;
        a.m2 = m2;
        a.y1 = y1;
        a.c1 = c1;
        a.y2 = y2;
        a.c2 = c2;
        a.np = std::min(SEAM_PCH, n_pairs - p0);
        for (int q = 0; q < a.np; q++) {
            a.k1[q] = k1[p0 + q];
            a.k2[q] = k2[p0 + q];
        }
        rc = ctx->launch("seam_sweep", [&] {
            hipLaunchKernelGGL(k_seam_sweep, dim3(blocks_for(ln.nl)), dim3(256), 0, ctx->stream, (const double2*)d_xi,
                               a, (double*)P.part.p);
        });
        std::vector<double> sums(2 * SEAM_PCH);
        if (!rc) rc = partials_reduce(ctx, P, ln.nl, 2 * a.np, 0, sums.data(), nullptr);
        if (!rc)
            for (int q = 0; q < a.np; q++) {
                const double sr = sums[2 * q], si = sums[2 * q + 1];
                out[2 * (p0 + q)] = (sr * ir - si * ii) * fac;
                out[2 * (p0 + q) + 1] = (sr * ii + si * ir) * fac;
            }
    }
    rc = finish(ctx, rc);
    dev_free(P.part);
    dev_free(P.red);
    return rc;
}

static int psum(hm_ctx* ctx, const double* d_los, const double* d_w, long long n, int ell, int m, int power,
                double out[3]) {
    DevBuf part, red;
    int rc = dev_alloc(ctx, part, sizeof(double) * 3 * PSUM_BLOCKS);
    if (!rc) rc = dev_alloc(ctx, red, sizeof(double) * 3);
    if (!rc)
        rc = ctx->launch("seam_psum", [&] {
            hipLaunchKernelGGL(k_seam_psum, dim3(PSUM_BLOCKS), dim3(256), 0, ctx->stream, d_los, d_w, n, ell, m,
                               power, (double*)part.p);
        });
    if (!rc) rc = reduce_cols_cm<double>(ctx, "seam_reduce", (const double*)part.p, PSUM_BLOCKS, 3, (double*)red.p);
    if (!rc && hipMemcpyAsync(out, red.p, sizeof(double) * 3, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = seam_bad(ctx, "particle sums: readback failed");
    rc = finish(ctx, rc);
    dev_free(part);
    dev_free(red);
    return rc;
}

int hm_particle_sums(hm_ctx* ctx, const double* d_los, const double* d_w, long long n, int ell, int m, int power,
                     double out[2]) {
    if (!ctx) return -1;
    if (n < 0 || (n > 0 && !d_los && ell != 0) || !out || ell < 0 || ell > 16 || m < -ell || m > ell || power < 0 ||
        power > 8)
        return seam_bad(ctx, "hm_particle_sums: bad arguments (d_los needed for ell > 0, 0 <= power <= 8)");
    double r[3] = {0.0, 0.0, 0.0};
    if (n > 0 && psum(ctx, d_los, d_w, n, ell, m, power, r)) return -1;
    out[0] = r[0];
    out[1] = r[1];
    return 0;
}

int hm_survey_volume(hm_ctx* ctx, const hm_mesh* mesh, const double* d_pos, const double* d_w, long long n,
                     double* survey_volume) {
    if (!ctx) return -1;
    Seam s;
    if (!seam_mesh(ctx, mesh, s, "hm_survey_volume")) return -1;
    if (n <= 0 || !d_pos || !survey_volume) return seam_bad(ctx, "hm_survey_volume: bad arguments");
    double sums[3];
    if (psum(ctx, nullptr, d_w, n, 0, 0, 1, sums)) return -1;  // sum w, sum |w|
    DevBuf cw, rho, part, red;
    int rc = 0;
    if (d_w) {
        rc = dev_alloc(ctx, cw, sizeof(double) * 2 * (size_t)n);
        if (!rc)
            rc = ctx->launch("seam_cweights", [&] {
                hipLaunchKernelGGL(k_seam_cweights, dim3(stride_blocks(n)), dim3(256), 0, ctx->stream, d_w, n,
                                   (double*)cw.p);
            });
    }
    if (!rc) rc = dev_alloc(ctx, rho, sizeof(double) * (size_t)s.m.ng());
    // the R field with weights w (field.hpp:474-485), real part only
    if (!rc) rc = hm_assign(ctx, mesh, d_pos, d_w ? (const double*)cw.p : nullptr, n, sums[2], (double*)rho.p, nullptr);
    const unsigned nb = stride_blocks(s.m.ng());
    if (!rc) rc = dev_alloc(ctx, part, sizeof(double) * nb);
    if (!rc) rc = dev_alloc(ctx, red, sizeof(double));
    const double dV = s.m.volume() / (double)s.m.ng();
    if (!rc)
        rc = ctx->launch("seam_sq", [&] {
            hipLaunchKernelGGL(k_seam_sq, dim3(nb), dim3(256), 0, ctx->stream, (const double*)rho.p, s.m.ng(), dV,
                               (double*)part.p);
        });
    if (!rc) rc = reduce_cols_cm<double>(ctx, "seam_reduce", (const double*)part.p, nb, 1, (double*)red.p);
    double norm = 0.0;
    if (!rc && hipMemcpyAsync(&norm, red.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = seam_bad(ctx, "hm_survey_volume: readback failed");
    rc = finish(ctx, rc);
    if (!rc) *survey_volume = sums[0] * sums[0] / norm;  // field.hpp:499
    dev_free(cw);
    dev_free(rho);
    dev_free(part);
    dev_free(red);
    return rc;
}

}  // extern "C"
