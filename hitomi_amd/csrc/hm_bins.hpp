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
};

inline void host_axis_tables(const Mesh& m, std::vector<double>& out, long long off[3]) {
    const long long ntot = (long long)m.n[0] + m.n[1] + m.n[2];
    out.assign(3 * ntot, 0.0);
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
        }
        o += n;
    }
}

__device__ __forceinline__ double wpow(double r, int p) { return p == 1 ? r : (p == 2 ? r * r : r * r * r); }

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

}  // namespace hm
