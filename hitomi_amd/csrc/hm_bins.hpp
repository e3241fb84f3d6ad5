// hm_bins.hpp -- pieces shared by the estimators: per-axis k / window /
// shot-noise tables with the reference's index truncation, |k| without
// contraction, Legendre polynomials, the fine -> coarse bin map and a
// fixed-order column reduction.
#pragma once

#include "hm_internal.hpp"

#include <cmath>
#include <string>
#include <vector>

namespace hm {

constexpr int FINE_N = 100000;      // field.hpp:775
constexpr double FINE_DK = 0.0001;  // field.hpp:776

// ---------------------------------------------------------------------------
// per-axis FP64 tables of the N-mesh: k component (field.hpp:797-799),
// assignment-window factor and shot-noise factor with the reference's index
// truncation (field.hpp:700-716, 1293-1309).
// ---------------------------------------------------------------------------
struct AxisTabs {
    const double* kv[3];
    const double* wf[3];
    const double* sf[3];
    const double* iw[3];  // 1 / wf^(2p), p = 1, 2, 3 for NGP, CIC, TSC
};

inline void host_axis_tables(const Mesh& m, std::vector<double>& out, long long off[3]) {
    const long long ntot = (long long)m.n[0] + m.n[1] + m.n[2];
    out.assign(4 * ntot, 0.0);
    const int p = m.assign == 0 ? 1 : (m.assign == 1 ? 2 : 3);
    long long o = 0;
    for (int a = 0; a < 3; a++) {
        off[a] = o;
        const int n = m.n[a];
        const double dk = 2.0 * M_PI / m.L[a];
        for (int i = 0; i < n; i++) {
            double kv = (i < n / 2) ? (double)i * dk : (double)(i - n) * dk;
            int ii = (int)(kv / dk + 1.0e-5);
            double x = M_PI * (double)ii / (double)n;
            double f = (ii != 0) ? std::sin(x) / x : 1.0;
            double s = (ii != 0) ? std::sin(x) : 0.0;
            double sfac;
            if (m.assign == 0) sfac = 1.0;
            else if (m.assign == 1) sfac = 1.0 - (2.0 / 3.0) * s * s;
            else sfac = 1.0 - s * s + 2.0 * std::pow(s, 4) / 15.0;
            out[o + i] = kv;
            out[ntot + o + i] = f;
            out[2 * ntot + o + i] = sfac;
            out[3 * ntot + o + i] = 1.0 / std::pow(f, 2 * p);
        }
        o += n;
    }
}

__device__ __forceinline__ double wpow(double r, int p) { return p == 1 ? r : (p == 2 ? r * r : r * r * r); }

// 1/x for a normal positive x: the hardware estimate and two Newton steps
// (within an ulp or two of the correctly rounded quotient, without the
// scale/fix-up sequence of a full double division)
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// |k| exactly as field.hpp:800 evaluates it (no contraction)
__device__ __forceinline__ double kmag3(double a, double b, double c) {
#pragma clang fp contract(off)
    return sqrt(a * a + b * b + c * c);
}

// Legendre P_l(mu) by the three-term recurrence (the m = 0 reduced
// spherical harmonic of tools.hpp:12-68 is sphPlm(l,0,mu) sqrt(4pi/(2l+1))).
__device__ __forceinline__ double legendre(int l, double mu) {
    if (l == 0) return 1.0;
    double p0 = 1.0, p1 = mu;
    for (int k = 2; k <= l; k++) {
        double p2 = ((2.0 * k - 1.0) * mu * p1 - (k - 1.0) * p0) / (double)k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

constexpr int FINE_NR = 20000;   // field.hpp:906
constexpr double FINE_DR = 0.5;  // field.hpp:907

__device__ __forceinline__ double mesh_vec(int i, int n, double d) {
#pragma clang fp contract(off)
    return (i < n / 2) ? (double)i * d : (double)(i - n) * d;  // field.hpp:929-931
}

// ---------------------------------------------------------------------------
// Reduced spherical harmonic of tools.hpp:12-68:
//   Y(l, M)(v) = conj[(-1)^((M-|M|)/2) N_l^|M|(mu) e^{i M phi}] sqrt(4 pi/(2l+1))
// with 1 for (0, 0), 0 for |v| < 1e-15, mu = v_z/|v|, phi = acos(v_x/|v_xy|)
// (0 when |v_xy| < 1e-15), reflected to 2 pi - phi when v_y < 0.  N_l^m is
// the orthonormal associated Legendre function with the Condon-Shortley
// phase, from the stable normalised recurrence
//   N_m^m = sqrt(1/4pi) prod_{i<=m} (-sqrt(1-x^2) sqrt((2i+1)/(2i))),
//   N_l^m = a_lm (x N_{l-1}^m - N_{l-2}^m / a_{l-1,m}),  a_lm = sqrt((4l^2-1)/(l^2-m^2)).
// ---------------------------------------------------------------------------
__host__ __device__ inline double norm_plm(int l, int m, double x) {
    const double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double ymm = 0.28209479177387814;  // 1 / sqrt(4 pi)
    for (int i = 1; i <= m; i++) ymm *= -somx2 * sqrt((2.0 * i + 1.0) / (2.0 * i));
    if (l == m) return ymm;
    double y1 = x * sqrt(2.0 * m + 3.0) * ymm;  // N_{m+1}^m
    double y0 = ymm;
    double aprev = sqrt(2.0 * m + 3.0);
    for (int ll = m + 2; ll <= l; ll++) {
        const double a = sqrt((4.0 * ll * ll - 1.0) / ((double)ll * ll - (double)m * m));
        const double y2 = a * (x * y1 - y0 / aprev);
        y0 = y1;
        y1 = y2;
        aprev = a;
    }
    return y1;
}

struct Cplx {
    double re, im;
};

__host__ __device__ inline Cplx ylm_red(int l, int M, double vx, double vy, double vz) {
#pragma clang fp contract(off)
    if (l == 0 && M == 0) return {1.0, 0.0};
    const double r = sqrt(vx * vx + vy * vy + vz * vz);
    const double rxy = sqrt(vx * vx + vy * vy);
    if (fabs(r) < 1.0e-15) return {0.0, 0.0};
    const double mu = vz / r;
    double phi = (fabs(rxy) < 1.0e-15) ? 0.0 : acos(vx / rxy);
    if (vy < 0.0) phi = -phi + 2.0 * M_PI;
    const int S = M < 0 ? -M : M;
    const double sgn = (M < 0 && (M & 1)) ? -1.0 : 1.0;
    const double p = sgn * norm_plm(l, S, mu);
    double sn, cs;
    sincos((double)M * phi, &sn, &cs);
    const double nrm = sqrt(4.0 * M_PI / (2.0 * (double)l + 1.0));
    return {p * cs * nrm, -(p * sn) * nrm};
}

// gsl_sf_coupling_3j (common.hpp:29) for small integer arguments
double wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3);

inline int wpow_of(int assign) { return assign == 0 ? 1 : (assign == 1 ? 2 : 3); }

// Fine index -> [first, last] coarse bin whose (lower, upper] holds the fine
// centre i * fine_d (field.hpp:827-837, 948-960); (-1, -2) when none.
// Returns an empty string or the reason the bins cannot be folded.
inline std::string fine_to_coarse(const std::vector<double>& lower, const std::vector<double>& upper, int nfine,
                                  double fine_d, std::vector<int2>& out) {
    out.assign(nfine, make_int2(-1, -2));
    for (int i = 0; i < nfine; i++) {
        const double c = (double)i * fine_d;
        int lo = -1, hi = -2;
        for (size_t j = 0; j < lower.size(); j++)
            if ((c > lower[j]) && (c <= upper[j])) {
                if (lo < 0) lo = (int)j;
                else if ((int)j != hi + 1) return "non-contiguous bin overlap";
                hi = (int)j;
            }
        out[i] = make_int2(lo, hi);
    }
    return "";
}

// Bin edges of ToolClass::setKbin / setRbin (tools.hpp:118-132) as the
// estimators use them: lower = max(c - d/2, 0), upper = c + d/2 with d the
// spacing of the first two centres (field.hpp:773, 826-830).
inline void bin_edges(double vmin, double vmax, int nb, std::vector<double>& centre, std::vector<double>& lower,
                      std::vector<double>& upper) {
    centre.resize(nb);
    const double d = (vmax - vmin) / (double)(nb - 1);
    for (int i = 0; i < nb; i++) centre[i] = vmin + d * (double)i;
    const double db = centre[1] - centre[0];
    lower.resize(nb);
    upper.resize(nb);
    for (int j = 0; j < nb; j++) {
        lower[j] = (centre[j] > db * 0.5) ? (centre[j] - db * 0.5) : 0.0;
        upper[j] = centre[j] + db * 0.5;
    }
}

// Sum rows of a [nrows][ncols] array per column in a fixed order.
template <typename V>
__global__ __launch_bounds__(256) void k_reduce_cols(const V* __restrict__ part, long long nrows, int ncols,
                                                    V* __restrict__ out) {
    __shared__ V sm[256];
    const int c = blockIdx.x;
    V acc = 0;
    for (long long r = threadIdx.x; r < nrows; r += 256) acc += part[r * ncols + c];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sm[0];
}

template <typename V> inline int reduce_cols(hm_ctx* ctx, const char* name, const V* part, long long nrows, int ncols, V* out) {
    if (ncols <= 0) return 0;
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL(k_reduce_cols<V>, dim3((unsigned)ncols), dim3(256), 0, ctx->stream, part, nrows, ncols, out);
    });
}

// The same sums over partials stored column-major (column c at c * nrows):
// one workgroup of 1024 lanes per column, each lane with four independent
// running sums over rows r = lane + 1024 (4 i + u) (fixed order), so that many
// contiguous loads are in flight.
template <typename V>
__global__ __launch_bounds__(1024) void k_reduce_cols_cm(const V* __restrict__ part, long long nrows,
                                                        V* __restrict__ out) {
    __shared__ V sm[1024];
    const int c = blockIdx.x;
    const V* col = part + (long long)c * nrows;
    V acc[4] = {0, 0, 0, 0};
    long long r = threadIdx.x;
    for (; r + 3 * 1024 < nrows; r += 4 * 1024) {
#pragma unroll
        for (int u = 0; u < 4; u++) acc[u] += col[r + u * 1024];
    }
    for (int u = 0; r < nrows; r += 1024, u++) acc[u] += col[r];
    sm[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sm[0];
}

template <typename V>
inline int reduce_cols_cm(hm_ctx* ctx, const char* name, const V* part, long long nrows, int ncols, V* out) {
    if (ncols <= 0) return 0;
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL(k_reduce_cols_cm<V>, dim3((unsigned)ncols), dim3(1024), 0, ctx->stream, part, nrows, out);
    });
}

}  // namespace hm
