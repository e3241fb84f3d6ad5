// hm_bessel.hpp -- spherical Bessel functions j_l and the natural cubic
// spline the reference samples them with (SphericalBesselClass,
// bessel.hpp:4-41: 10^6 nodes on [0, 10^4] of gsl_sf_bessel_jl, GSL cspline).
#pragma once

#include "hm_common.hpp"

#include <vector>

namespace hm {

constexpr int JL_NODES = 1000000;      // bessel.hpp:15
constexpr double JL_XMAX = 10000.0;    // bessel.hpp:14

// j_l(x) for x >= 0: power series below x = 1, upward recurrence from j_0,
// j_1 above x = l, and otherwise Miller's downward recurrence normalised by
// sum_n (2n+1) j_n^2 = 1.
double sph_jl(int l, double x);

// Node values y_i = j_l(i dx) and c_i = y''_i / 2 of the natural cubic
// spline through them (second derivatives zero at both ends).
void jl_spline(int l, std::vector<double>& y, std::vector<double>& c);

// The spline at x in [0, 10^4] (NaN outside, where GSL reports a domain
// error): interval i with x_i <= x < x_{i+1}, the last one closed.
__host__ __device__ inline double jl_spline_eval(const double* __restrict__ y, const double* __restrict__ c, double x) {
#pragma clang fp contract(off)
    const double dx = JL_XMAX / (double)(JL_NODES - 1);
    if (!(x >= 0.0) || x > dx * (double)(JL_NODES - 1)) return NAN;
    int i = (int)(x / dx);
    if (i > JL_NODES - 2) i = JL_NODES - 2;
    if (i > 0 && dx * (double)i > x) i--;
    if (i < JL_NODES - 2 && dx * (double)(i + 1) <= x) i++;
    const double x_lo = dx * (double)i, x_hi = dx * (double)(i + 1);
    const double h = x_hi - x_lo, t = x - x_lo;
    const double b = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
    const double d = (c[i + 1] - c[i]) / (3.0 * h);
    return y[i] + t * (b + t * (c[i] + t * d));
}

}  // namespace hm
