// hm_assign.hip -- mass assignment kernels (field.hpp:47-184).
#include "hm_assign.hpp"

#include <cmath>

namespace hm {

double fixed_scale_for(double wabs_sum) {
    if (!(wabs_sum > 0.0)) wabs_sum = 1.0;
    int e = 0;
    std::frexp(wabs_sum * (1.0 + 1e-12) + 1.0, &e);  // wabs_sum+1 < 2^e
    return std::ldexp(1.0, 62 - e);
}

// One thread per particle: 1/8/27 stencil contributions, each rounded to the
// nearest multiple of 2^-s and added with a 64-bit integer atomic.
template <int S>
__global__ __launch_bounds__(256) void k_assign_fixed(const double* __restrict__ pos, const double* __restrict__ w,
                                                     long long np, AssignGeom g, double scale,
                                                     unsigned long long* __restrict__ gre,
                                                     unsigned long long* __restrict__ gim) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    double wt[3][3];
    int iw0[3];
#pragma unroll
    for (int a = 0; a < 3; a++) stencil_axis(g.assign, axis_xp(pos[3 * q + a], g.n[a], g.L[a]), &iw0[a], wt[a]);
    const double wre = w ? w[2 * q] : 1.0;
    const double wim = w ? w[2 * q + 1] : 0.0;
#pragma unroll
    for (int i = 0; i < S; i++) {
        const long long ci = (long long)(iw0[0] + i) * g.n[1];
#pragma unroll
        for (int j = 0; j < S; j++) {
            const long long cij = (ci + (iw0[1] + j)) * g.n[2];
            const double wij = wt[0][i] * wt[1][j];
#pragma unroll
            for (int k = 0; k < S; k++) {
                // field.hpp:175-176: unwrapped per-axis index, linear range check only
                const long long coord = cij + (iw0[2] + k);
                if (coord >= 0 && coord < g.ng) {
                    const double wijk = wij * wt[2][k];
                    const long long vr = __double2ll_rn(wre * wijk * scale);
                    atomicAdd(gre + coord, (unsigned long long)vr);
                    if (gim) {
                        const long long vi = __double2ll_rn(wim * wijk * scale);
                        atomicAdd(gim + coord, (unsigned long long)vi);
                    }
                }
            }
        }
    }
}

int assign_fixed(hm_ctx* ctx, const AssignGeom& g, const double* d_pos, const double* d_w, long long np,
                 double scale, long long* d_re, long long* d_im) {
    if (np <= 0) return 0;
    const unsigned blocks = (unsigned)((np + 255) / 256);
    auto* re = reinterpret_cast<unsigned long long*>(d_re);
    auto* im = reinterpret_cast<unsigned long long*>(d_im);
    return ctx->launch("assign_fixed", [&] {
        if (g.assign == 0)
            hipLaunchKernelGGL(k_assign_fixed<1>, dim3(blocks), dim3(256), 0, ctx->stream, d_pos, d_w, np, g, scale, re, im);
        else if (g.assign == 1)
            hipLaunchKernelGGL(k_assign_fixed<2>, dim3(blocks), dim3(256), 0, ctx->stream, d_pos, d_w, np, g, scale, re, im);
        else
            hipLaunchKernelGGL(k_assign_fixed<3>, dim3(blocks), dim3(256), 0, ctx->stream, d_pos, d_w, np, g, scale, re, im);
    });
}

template <int S>
__global__ __launch_bounds__(256) void k_cell_coords(const double* __restrict__ pos, long long np, AssignGeom g,
                                                    long long* __restrict__ out) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    double wt[3][3];
    int iw0[3];
#pragma unroll
    for (int a = 0; a < 3; a++) stencil_axis(g.assign, axis_xp(pos[3 * q + a], g.n[a], g.L[a]), &iw0[a], wt[a]);
    int t = 0;
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++)
            for (int k = 0; k < S; k++, t++) {
                const long long coord = ((long long)(iw0[0] + i) * g.n[1] + (iw0[1] + j)) * g.n[2] + (iw0[2] + k);
                out[q * (S * S * S) + t] = (coord >= 0 && coord < g.ng) ? coord : -1;
            }
}

int cell_coords(hm_ctx* ctx, const AssignGeom& g, const double* d_pos, long long np, long long* d_out) {
    if (np <= 0) return 0;
    const unsigned blocks = (unsigned)((np + 255) / 256);
    return ctx->launch("cell_coords", [&] {
        if (g.assign == 0)
            hipLaunchKernelGGL(k_cell_coords<1>, dim3(blocks), dim3(256), 0, ctx->stream, d_pos, np, g, d_out);
        else if (g.assign == 1)
            hipLaunchKernelGGL(k_cell_coords<2>, dim3(blocks), dim3(256), 0, ctx->stream, d_pos, np, g, d_out);
        else
            hipLaunchKernelGGL(k_cell_coords<3>, dim3(blocks), dim3(256), 0, ctx->stream, d_pos, np, g, d_out);
    });
}

// fixed point -> FP64 density: rho = (q * 2^-s) * (1/dV)   (field.hpp:155-156, 177)
__global__ void k_fixed_to_density(const long long* __restrict__ q, long long n, double inv_scale, double factor,
                                   double* __restrict__ out) {
#pragma clang fp contract(off)
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = i0; i < n; i += st) out[i] = ((double)q[i] * inv_scale) * factor;
}

int fixed_to_density(hm_ctx* ctx, const long long* d_q, long long n, double inv_scale, double factor, double* d_out) {
    long long blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    return ctx->launch("fixed_to_density", [&] {
        hipLaunchKernelGGL(k_fixed_to_density, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d_q, n, inv_scale,
                           factor, d_out);
    });
}

}  // namespace hm
