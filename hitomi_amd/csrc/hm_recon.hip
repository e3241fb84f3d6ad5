// hm_recon.hip -- density-field reconstruction on one MI355X
// (calcReconstructionParticles / calcReconstructionParticlesForBOX,
// particle_reconstruction.hpp:12-402) and the periodic wrap
// (calcPeriodicBoundary, particleGadget.hpp:565-591).
//
// One call: the fixed-point field of D - alpha R on the reconstruction mesh,
// its transform divided by the assignment window, the three Zel'dovich
// displacement components Psi_a(k) = -i (G/b) k_a/k^2 delta/W / V transformed
// back, and a TSC gather of norm * Re Psi added to every particle of both
// catalogues (linear index range check only, as the reference does).
#include "hm_assign.hpp"
#include "hm_bins.hpp"
#include "hm_fft_launch.hpp"
#include "hm_recon.hpp"

#include "../../include/hitomi_amd.h"

#include <algorithm>
#include <cmath>

namespace hm {

// D then R in one array; weights w (1 when absent) for D and -alpha w for R
__global__ __launch_bounds__(256) void k_recon_pack(const double* __restrict__ pD, const double* __restrict__ wD,
                                                   long long nD, const double* __restrict__ pR,
                                                   const double* __restrict__ wR, long long nR, double alpha,
                                                   double* __restrict__ pos, double* __restrict__ w) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= nD + nR) return;
    const bool d = q < nD;
    const double* s = d ? pD + 3 * q : pR + 3 * (q - nD);
    pos[3 * q] = s[0];
    pos[3 * q + 1] = s[1];
    pos[3 * q + 2] = s[2];
    double v = d ? (wD ? wD[q] : 1.0) : (wR ? wR[q - nD] : 1.0);
    if (!d) v = -(alpha * v);
    w[2 * q] = v;
    w[2 * q + 1] = 0.0;
}

int pack_two(hm_ctx* ctx, const double* pD, const double* wD, long long nD, const double* pR, const double* wR,
             long long nR, double alpha, double* pos, double* w) {
    if (nD + nR <= 0) return 0;
    return ctx->launch("pack_two", [&] {
        hipLaunchKernelGGL(k_recon_pack, dim3((unsigned)((nD + nR + 255) / 256)), dim3(256), 0, ctx->stream, pD, wD, nD,
                           pR, wR, nR, alpha, pos, w);
    });
}

// per-block sums of |x[i * stride]| (fixed order)
__global__ __launch_bounds__(256) void k_abs_sum_stride(const double* __restrict__ x, long long n, int stride,
                                                       double* __restrict__ part) {
    __shared__ double sm[256];
    double a = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        a += fabs(x[i * stride]);
    sm[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Psi_a(k) for a = 0, 1, 2 from the forward-transformed field
// (particle_reconstruction.hpp:64-112, 260-308; field.hpp:610-639 for / W)
template <typename T>
__global__ __launch_bounds__(256) void k_recon_psi(const C<T>* __restrict__ dk, int n0, int n1, int n2, AxisTabs t,
                                                  int wp, double b1, double RG, double inv_v, C<T>* __restrict__ psi) {
#pragma clang fp contract(off)
    const long long ng = (long long)n0 * n1 * n2;
    const long long st = (long long)gridDim.x * 256;
    for (long long lin = (long long)blockIdx.x * 256 + threadIdx.x; lin < ng; lin += st) {
        const int k = (int)(lin % n2);
        const long long r = lin / n2;
        const int j = (int)(r % n1), i = (int)(r / n1);
        const double kv[3] = {t.kv[0][i], t.kv[1][j], t.kv[2][k]};
        const double kmag = kmag3(kv[0], kv[1], kv[2]);
        if (kmag < 1.0e-10) {  // left unwritten by the reference (Q5); zero here
            for (int a = 0; a < 3; a++) psi[a * ng + lin] = mk<T>((T)0, (T)0);
            continue;
        }
        const double wf = wpow((t.wf[0][i] * t.wf[1][j]) * t.wf[2][k], wp);
        const C<T> f = dk[lin];
        const double fr = (double)f.x / wf, fi = (double)f.y / wf;
        const double kmag2 = kmag * kmag;
        const double G = exp(-pow(kmag * RG, 2) / 2.0);
        for (int a = 0; a < 3; a++) {
            const double c = (G / b1) * (kv[a] / kmag2);
            psi[a * ng + lin] = mk<T>((T)((c * fi) * inv_v), (T)((-(c * fr)) * inv_v));
        }
    }
}

// TSC gather of norm * Re Psi (particle_reconstruction.hpp:141-205, 334-388)
template <typename T>
__global__ __launch_bounds__(256) void k_recon_gather(const C<T>* __restrict__ psi, AssignGeom g, double norm,
                                                     double* __restrict__ pos, long long np) {
#pragma clang fp contract(off)
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= np) return;
    double w[3][3];
    int i0[3];
    for (int a = 0; a < 3; a++) stencil_axis(2, axis_xp(pos[3 * q + a], g.n[a], g.L[a]), &i0[a], w[a]);
    double p[3] = {pos[3 * q], pos[3 * q + 1], pos[3 * q + 2]};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                const long long coord = ((long long)(i0[0] + i) * g.n[1] + (i0[1] + j)) * g.n[2] + (i0[2] + k);
                if (coord < 0 || coord >= g.ng) continue;
                const double ww = w[0][i] * w[1][j] * w[2][k];  // w[axis][stencil]
                for (int a = 0; a < 3; a++) p[a] += ww * ((double)psi[a * g.ng + coord].x * norm);
            }
    pos[3 * q] = p[0];
    pos[3 * q + 1] = p[1];
    pos[3 * q + 2] = p[2];
}

__global__ __launch_bounds__(256) void k_recon_unpack(const double* __restrict__ pos, long long nD, long long nR,
                                                     double* __restrict__ pD, double* __restrict__ pR) {
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= nD + nR) return;
    double* d = q < nD ? pD + 3 * q : pR + 3 * (q - nD);
    d[0] = pos[3 * q];
    d[1] = pos[3 * q + 1];
    d[2] = pos[3 * q + 2];
}

// calcPeriodicBoundary: one step into [0, L) per axis
__global__ __launch_bounds__(256) void k_periodic_wrap(double* __restrict__ pos, long long n, double L0, double L1,
                                                      double L2) {
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const double L[3] = {L0, L1, L2};
    for (int a = 0; a < 3; a++) {
        double v = pos[3 * q + a];
        if (v >= L[a]) v -= L[a];
        else if (v < 0.0) v += L[a];
        pos[3 * q + a] = v;
    }
}

int periodic_wrap(hm_ctx* ctx, double* pos, long long n, const double* L) {
    if (n <= 0) return 0;
    return ctx->launch("periodic_wrap", [&] {
        hipLaunchKernelGGL(k_periodic_wrap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, pos, n, L[0],
                           L[1], L[2]);
    });
}

void recon_work_free(ReconWork& w) {
    DevBuf* b[] = {&w.pos, &w.wts, &w.fre, &w.grid, &w.psi, &w.tab, &w.part};
    for (DevBuf* x : b) dev_free(*x);
    sort_work_free(w.sort);
}

size_t recon_work_bytes(const ReconWork& w) {
    return w.pos.bytes + w.wts.bytes + w.fre.bytes + w.grid.bytes + w.psi.bytes + w.tab.bytes + w.part.bytes +
           w.sort.bytes();
}

// Psi_a (complex, 3 grids in w.psi) of the field D - alpha R with weights;
// returns the fixed-point scale in *scale_out (w.fre keeps the fixed-point mesh).
template <typename T>
static int recon_psi_stage(hm_ctx* ctx, const Mesh& m, double b1, double RG, const double* pD, const double* wD,
                           long long nD, const double* pR, const double* wR, long long nR, double alpha, ReconWork& w,
                           double* scale_out) {
    const long long N = nD + nR, ng = m.ng();
    if (!(b1 != 0.0)) {
        ctx->set_error("reconstruction: b1_fid must be non-zero");
        return -1;
    }
    const double dV = m.volume() / (double)ng;
    if (dev_alloc(ctx, w.pos, sizeof(double) * 3 * (size_t)N) || dev_alloc(ctx, w.wts, sizeof(double) * 2 * (size_t)N) ||
        dev_alloc(ctx, w.fre, sizeof(long long) * (size_t)ng) || dev_alloc(ctx, w.grid, sizeof(C<T>) * (size_t)ng) ||
        dev_alloc(ctx, w.psi, 3 * sizeof(C<T>) * (size_t)ng))
        return -1;
    std::vector<double> tabs;
    long long off[3];
    host_axis_tables(m, tabs, off);
    if (dev_alloc(ctx, w.tab, sizeof(double) * tabs.size())) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(w.tab.p, tabs.data(), sizeof(double) * tabs.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
    double* pos = (double*)w.pos.p;
    double* wts = (double*)w.wts.p;
    const unsigned pb = (unsigned)((N + 255) / 256);
    if (ctx->launch("recon_pack", [&] {
            hipLaunchKernelGGL(k_recon_pack, dim3(pb), dim3(256), 0, ctx->stream, pD, wD, nD, pR, wR, nR, alpha, pos,
                               wts);
        }))
        return -1;
    // the fixed-point scale needs a bound on sum |weight|
    std::vector<double> hw;
    double wabs = 0.0;
    {
        const unsigned rb = (unsigned)std::max<long long>(1, std::min<long long>((N + 255) / 256, 1024));
        if (dev_alloc(ctx, w.part, sizeof(double) * (rb + 1))) return -1;
        if (ctx->launch("recon_wabs", [&] {
                hipLaunchKernelGGL(k_abs_sum_stride, dim3(rb), dim3(256), 0, ctx->stream, wts, N, 2,
                                   (double*)w.part.p);
            }))
            return -1;
        hw.resize(rb);
        HM_HIP_CHECK(ctx, hipMemcpyAsync(hw.data(), w.part.p, sizeof(double) * rb, hipMemcpyDeviceToHost, ctx->stream));
        HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        for (double v : hw) wabs += v;
    }
    AssignGeom ag;
    for (int a = 0; a < 3; a++) {
        ag.n[a] = m.n[a];
        ag.L[a] = m.L[a];
    }
    ag.ng = ng;
    ag.assign = m.assign;
    if (sort_work_alloc(ctx, w.sort, ag, N, true)) return -1;
    const double scale = fixed_scale_for(1.01 * wabs + 1.0);
    *scale_out = scale;
    long long* fre = (long long*)w.fre.p;
    if (assign_sorted(ctx, ag, pos, wts, N, scale, fre, nullptr, w.sort)) return -1;
    C<T>* grid = (C<T>*)w.grid.p;
    const unsigned gb = (unsigned)std::max<long long>(1, std::min<long long>((ng + 255) / 256, 16384));
    if (ctx->launch("recon_to_complex", [&] {
            hipLaunchKernelGGL(k_fixed_to_complex<T>, dim3(gb), dim3(256), 0, ctx->stream, fre, nullptr, ng, 1.0 / scale,
                               1.0 / dV, dV, grid);
        }))
        return -1;
    const int n3[3] = {m.n[0], m.n[1], m.n[2]};
    if (fft3d_c2c<T>(ctx, grid, n3, -1)) return -1;
    AxisTabs t;
    const double* base = (const double*)w.tab.p;
    const long long ntot = (long long)m.n[0] + m.n[1] + m.n[2];
    for (int a = 0; a < 3; a++) {
        t.kv[a] = base + off[a];
        t.wf[a] = base + ntot + off[a];
        t.sf[a] = base + 2 * ntot + off[a];
    }
    C<T>* psi = (C<T>*)w.psi.p;
    if (ctx->launch("recon_psi", [&] {
            hipLaunchKernelGGL(k_recon_psi<T>, dim3(gb), dim3(256), 0, ctx->stream, grid, m.n[0], m.n[1], m.n[2], t,
                               wpow_of(m.assign), b1, RG, 1.0 / m.volume(), psi);
        }))
        return -1;
    for (int a = 0; a < 3; a++)
        if (fft3d_c2c<T>(ctx, psi + (long long)a * ng, n3, +1)) return -1;
    return 0;
}

template <typename T>
int recon_displace(hm_ctx* ctx, const Mesh& m, double b1, double RG, double* pD, const double* wD, long long nD,
                   double* pR, const double* wR, long long nR, double alpha, double norm, ReconWork& w) {
    const long long N = nD + nR;
    if (N <= 0) return 0;
    double scale = 0.0;
    if (recon_psi_stage<T>(ctx, m, b1, RG, pD, wD, nD, pR, wR, nR, alpha, w, &scale)) return -1;
    AssignGeom tsc;
    for (int a = 0; a < 3; a++) {
        tsc.n[a] = m.n[a];
        tsc.L[a] = m.L[a];
    }
    tsc.ng = m.ng();
    tsc.assign = m.assign;
    double* pos = (double*)w.pos.p;
    const unsigned pb = (unsigned)((N + 255) / 256);
    C<T>* psi = (C<T>*)w.psi.p;
    if (ctx->launch("recon_gather", [&] {
            hipLaunchKernelGGL(k_recon_gather<T>, dim3(pb), dim3(256), 0, ctx->stream, psi, tsc, norm, pos, N);
        }))
        return -1;
    return ctx->launch("recon_unpack", [&] {
        hipLaunchKernelGGL(k_recon_unpack, dim3(pb), dim3(256), 0, ctx->stream, pos, nD, nR, pD, pR);
    });
}

// features [ng][4] FP32: norm * Re Psi_a (a = 0, 1, 2) and rho / rho_bar - 1
template <typename T>
__global__ __launch_bounds__(256) void k_recon_features(const C<T>* __restrict__ psi, const long long* __restrict__ fre,
                                                       long long ng, double norm, double inv_scale, double inv_mean,
                                                       float* __restrict__ out) {
#pragma clang fp contract(off)
    const long long st = (long long)gridDim.x * 256;
    for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < ng; v += st) {
        for (int a = 0; a < 3; a++) out[4 * v + a] = (float)((double)psi[a * ng + v].x * norm);
        out[4 * v + 3] = (float)(((double)fre[v] * inv_scale) * inv_mean - 1.0);
    }
}

template <typename T>
int recon_features(hm_ctx* ctx, const Mesh& m, double b1, double RG, const double* pD, long long nD, double norm,
                   float* out, ReconWork& w) {
    if (nD <= 0) {
        ctx->set_error("reconstruction features: no particles");
        return -1;
    }
    double scale = 0.0;
    if (recon_psi_stage<T>(ctx, m, b1, RG, pD, nullptr, nD, nullptr, nullptr, 0, 0.0, w, &scale)) return -1;
    const long long ng = m.ng();
    const unsigned gb = (unsigned)std::max<long long>(1, std::min<long long>((ng + 255) / 256, 16384));
    const double inv_mean = (double)ng / (double)nD;  // 1 / (mean particles per cell)
    return ctx->launch("recon_features", [&] {
        hipLaunchKernelGGL(k_recon_features<T>, dim3(gb), dim3(256), 0, ctx->stream, (const C<T>*)w.psi.p,
                           (const long long*)w.fre.p, ng, norm, 1.0 / scale, inv_mean, out);
    });
}

template int recon_features<float>(hm_ctx*, const Mesh&, double, double, const double*, long long, double, float*,
                                   ReconWork&);
template int recon_features<double>(hm_ctx*, const Mesh&, double, double, const double*, long long, double, float*,
                                    ReconWork&);
template int recon_displace<float>(hm_ctx*, const Mesh&, double, double, double*, const double*, long long, double*,
                                   const double*, long long, double, double, ReconWork&);
template int recon_displace<double>(hm_ctx*, const Mesh&, double, double, double*, const double*, long long, double*,
                                    const double*, long long, double, double, ReconWork&);

}  // namespace hm
