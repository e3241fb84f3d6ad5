// hm_bessel.hip -- host side of hm_bessel.hpp: j_l(x) and its spline nodes.
#include "hm_bessel.hpp"

#include <cmath>

namespace hm {

double sph_jl(int l, double x) {
    if (l < 0 || !(x >= 0.0)) return NAN;
    if (x == 0.0) return l == 0 ? 1.0 : 0.0;
    if (x < 1.0) {
        // x^l / (2l+1)!! * sum_k (-x^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1))
        double pre = 1.0;
        for (int i = 1; i <= l; i++) pre *= x / (2.0 * i + 1.0);
        double term = 1.0, sum = 1.0;
        const double h = -0.5 * x * x;
        for (int k = 1; k < 40; k++) {
            term *= h / ((double)k * (2.0 * l + 2.0 * k + 1.0));
            sum += term;
            if (std::fabs(term) < 1e-17 * std::fabs(sum)) break;
        }
        return pre * sum;
    }
    const double s = std::sin(x), c = std::cos(x);
    const double j0 = s / x;
    if (l == 0) return j0;
    const double j1 = s / (x * x) - c / x;
    if (l == 1) return j1;
    if (x > (double)l) {
        double a = j0, b = j1;
        for (int n = 1; n < l; n++) {
            const double nx = (2.0 * n + 1.0) / x * b - a;
            a = b;
            b = nx;
        }
        return b;
    }
    // Miller: j_{n-1} = (2n+1)/x j_n - j_{n+1} from far above l, normalised by
    // sum (2n+1) j_n^2 = 1; rescaled whenever the values grow large
    const int N = l + 30 + (int)x;
    double up = 0.0, cur = 1.0, val = 0.0, norm = 0.0;
    for (int n = N; n >= 0; n--) {
        norm += (2.0 * n + 1.0) * cur * cur;
        if (n == l) val = cur;
        const double down = (2.0 * n + 1.0) / x * cur - up;
        up = cur;
        cur = down;
        if (std::fabs(cur) > 1e100) {
            cur *= 1e-200;
            up *= 1e-200;
            val *= 1e-200;
            norm *= 1e-200;
            norm *= 1e-200;
        }
    }
    return val / std::sqrt(norm);
}

void jl_spline(int l, std::vector<double>& y, std::vector<double>& c) {
    const int n = JL_NODES;
    const double dx = JL_XMAX / (double)(n - 1);
    y.resize(n);
    c.assign(n, 0.0);
    for (int i = 0; i < n; i++) y[i] = sph_jl(l, dx * (double)i);
    // natural spline: h c_{i-1} + 4 h c_i + h c_{i+1} = 3 (y_{i+1} - 2 y_i + y_{i-1}) / h
    // on the uniform grid (c = y''/2), c_0 = c_{n-1} = 0; Thomas algorithm
    const int m = n - 2;
    std::vector<double> cp(m), dp(m);
    for (int k = 0; k < m; k++) {
        const int i = k + 1;
        const double x0 = dx * (double)(i - 1), x1 = dx * (double)i, x2 = dx * (double)(i + 1);
        const double h0 = x1 - x0, h1 = x2 - x1;
        const double diag = 2.0 * (h0 + h1);
        const double rhs = 3.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
        if (k == 0) {
            cp[k] = h1 / diag;
            dp[k] = rhs / diag;
        } else {
            const double den = diag - h0 * cp[k - 1];
            cp[k] = h1 / den;
            dp[k] = (rhs - h0 * dp[k - 1]) / den;
        }
    }
    for (int k = m - 1; k >= 0; k--) c[k + 1] = dp[k] - (k < m - 1 ? cp[k] * c[k + 2] : 0.0);
}

}  // namespace hm
