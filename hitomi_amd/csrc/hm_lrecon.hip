// hm_lrecon.hip -- learned correction of the plain Zel'dovich reconstruction.
//
// No counterpart in the reference (its reconstruction is the plain one,
// particle_reconstruction.hpp:210-402).  A residual network adds a correction
// to the plain reconstruction's shift field:
//
//   features  X [v][4]: norm Psi_a / psi_scale (a = 0, 1, 2) of the plain
//             reconstruction on the mesh and (rho / rho_bar - 1) / 8
//   network   four 3^3 convolutions with periodic padding, 4 -> 16 -> 16 ->
//             16 -> 3 channels, ReLU after the first three; activations in
//             BF16, channels last; MFMA 16x16x32 BF16 with FP32 accumulation
//   shift     S_i = norm Psi(x_i) + psi_scale out(x_i), both gathered with the
//             TSC stencil of the reference's reconstruction gather
//   loss      mean_i |S_i - T_i|^2 against the true shifts T_i
//
// Training is the network's own backward pass: the loss gradient on the
// output mesh is the adjoint of the gather, the sorted 64-bit fixed-point
// mass assignment of the residuals (the same stencil; order-independent), the input
// gradient of a convolution is the same convolution kernel with flipped,
// transposed weights (masked by the ReLU of the layer below), the weight
// gradient an MFMA reduction over voxels; Adam (Kingma and Ba, ICLR 2015)
// updates FP32 master weights.
//
// Convolution as an implicit GEMM on gfx950's v_mfma_f32_16x16x32_bf16: for 16
// consecutive voxels along z (M), 16 output channels (N) and K = 27 taps x cin
// input channels in steps of 32, lane l of a wave holds
// A[voxel l % 16][k = 32 s + 8 (l / 16) + e], e = 0..7 (with 16 channels one
// tap's eight consecutive channels, a 16-byte load; with 4 channels two taps),
// B[k][l % 16] (weights packed per lane once per update), and receives
// D[4 (l / 16) + e][l % 16], e = 0..3.
#include "hm_internal.hpp"
#include "hm_assign.hpp"
#include "hm_recon.hpp"

#include <cmath>
#include <cstdint>

using namespace hm;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// K steps of 32 for the 3^3 taps of cin channels
constexpr int conv_steps(int cin) { return (27 * cin + 31) / 32; }
// packed B operands per layer: at most 14 steps x 64 lanes x 8 values
constexpr size_t WP_LAYER = 16 * 512;

__device__ __forceinline__ f32x4 mfma32(u16x8 a, u16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

namespace {

int bad(hm_ctx* ctx, const std::string& m) {
    if (ctx) ctx->set_error(m);
    return -1;
}

// FP32 -> BF16, round to nearest even (finite values)
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// ---------------------------------------------------------------------------
// weights
// ---------------------------------------------------------------------------
// W [27][cin][cout] (tap t = 9 (dx + 1) + 3 (dy + 1) + (dz + 1)) -> per-lane
// B operands of ks steps ([ks][64][8]).  Forward: B[k][j] = W[t][ci][j],
// k = t cin + ci.  Input gradient (transpose): the convolution of dZ (16
// channels, cout used) with B[k][j] = W[26 - t][j][c], k = 16 t + c, j = ci.
__global__ void k_pack_w(const float* __restrict__ W, int cin, int cout, int ks, int transpose,
                         unsigned short* __restrict__ wp) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ks * 512) return;
    const int s = idx / 512, l = (idx / 8) % 64, e = idx % 8;
    const int k = 32 * s + 8 * (l >> 4) + e, j = l & 15;
    float v = 0.0f;
    if (!transpose) {
        const int t = k / cin, ci = k % cin;
        if (t < 27 && j < cout) v = W[((size_t)t * cin + ci) * cout + j];
    } else {
        const int t = k / 16, c = k % 16;
        if (t < 27 && c < cout && j < cin) v = W[((size_t)(26 - t) * cin + j) * cout + c];
    }
    wp[idx] = f2bf(v);
}

__global__ void k_to_bf16(const float* __restrict__ x, long long n, int cin_src, int cstore,
                          unsigned short* __restrict__ out) {
    // x [nv][cin_src] -> out [nv][cstore] (extra channels zero)
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n * cstore; i += st) {
        const long long v = i / cstore;
        const int c = (int)(i % cstore);
        out[i] = c < cin_src ? f2bf(x[v * cin_src + c]) : (unsigned short)0;
    }
}

// ---------------------------------------------------------------------------
// convolution (forward, and input gradient with transposed weights)
// ---------------------------------------------------------------------------
enum { CONV_RELU_BF16 = 0, CONV_F32 = 1, CONV_MASK_BF16 = 2 };

struct ConvArgs {
    const unsigned short* in;  // [ng][CIN] BF16
    const unsigned short* wp;  // [ks][64][8] BF16
    int ks;
    const float* bias;         // [16] or null
    int n0, n1, n2;
    int mode, cout, ostride;
    void* out;                 // RELU / MASK: [ng][16] BF16; F32: [ng][ostride] FP32
    const unsigned short* mask;  // MASK: [ng][16] BF16, gate > 0
};

// The A operand of K step s for lane group g at voxel z of row base rb (the
// first voxel of the tap's row): 16 channels -> tap 2 s + g / 2, channels
// 8 (g % 2) .. + 7; 4 channels -> taps 8 s + 2 g and 8 s + 2 g + 1, all four.
// Taps past 26 are zero.  row(t) gives the voxel index of (x + dx, y + dy, 0);
// zof(dz) the wrapped z of this lane's voxel.
template <int CIN, typename Row, typename Zof>
__device__ __forceinline__ u16x8 conv_a(const unsigned short* __restrict__ in, int s, int g, Row row, Zof zof) {
    u16x8 av = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    if (CIN == 16) {
        const int t = 2 * s + (g >> 1);
        if (t < 27) av = *reinterpret_cast<const u16x8*>(in + (row(t) + zof(t % 3 - 1)) * 16 + 8 * (g & 1));
    } else {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int t = 8 * s + 2 * g + h;
            if (t < 27) {
                const u16x4 q = *reinterpret_cast<const u16x4*>(in + (row(t) + zof(t % 3 - 1)) * 4);
#pragma unroll
                for (int e = 0; e < 4; e++) av[4 * h + e] = q[e];
            }
        }
    }
    return av;
}

template <int CIN, int TPW>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int segs = a.n2 / (16 * TPW);
    if (wave >= (long long)a.n0 * a.n1 * segs) return;  // uniform per wave
    const int seg = (int)(wave % segs);
    const long long row = wave / segs;
    const int y = (int)(row % a.n1), x = (int)(row / a.n1);
    const int m = lane & 15, g = lane >> 4;
    f32x4 acc[TPW];
#pragma unroll
    for (int tt = 0; tt < TPW; tt++) acc[tt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const u16x8* W = reinterpret_cast<const u16x8*>(a.wp);
    const auto tap_row = [&](int t) -> long long {
        return ((long long)wrapi(x + t / 9 - 1, a.n0) * a.n1 + wrapi(y + (t / 3) % 3 - 1, a.n1)) * a.n2;
    };
    for (int s = 0; s < a.ks; s++) {
        const u16x8 bv = W[s * 64 + lane];
#pragma unroll
        for (int tt = 0; tt < TPW; tt++) {
            const int zb = (seg * TPW + tt) * 16 + m;
            const u16x8 av =
                conv_a<CIN>(a.in, s, g, tap_row, [&](int dz) { return (long long)wrapi(zb + dz, a.n2); });
            acc[tt] = mfma32(av, bv, acc[tt]);
        }
    }
    const int c = lane & 15;
    const float bb = a.bias ? a.bias[c] : 0.0f;
    const long long r0 = ((long long)x * a.n1 + y) * a.n2;
#pragma unroll
    for (int tt = 0; tt < TPW; tt++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const long long v = r0 + (seg * TPW + tt) * 16 + 4 * g + e;
            const float val = acc[tt][e] + bb;
            if (a.mode == CONV_RELU_BF16) {
                static_cast<unsigned short*>(a.out)[v * 16 + c] = f2bf(fmaxf(val, 0.0f));
            } else if (a.mode == CONV_F32) {
                if (c < a.cout) static_cast<float*>(a.out)[v * a.ostride + c] = val;
            } else {
                static_cast<unsigned short*>(a.out)[v * 16 + c] = f2bf(bf2f(a.mask[v * 16 + c]) > 0.0f ? val : 0.0f);
            }
        }
}

// The same convolution with the input tile of a workgroup staged in LDS: a
// workgroup covers 4 x 4 rows (x, y) and 16 TPW voxels along z, loads that
// block with a one-voxel periodic halo once (channels last, 16-byte loads),
// and its four waves (one per row of x) read the A operands of all 27 taps
// from LDS; 16 consecutive voxels x 32 B are contiguous, so a wave's 8-byte
// reads cover 512 B without bank conflicts.
constexpr int CL_RX = 4, CL_RY = 4;

template <int CIN, int TPW>
__global__ __launch_bounds__(256) void k_conv_lds(ConvArgs a) {
    constexpr int ZS = 16 * TPW, HX = CL_RX + 2, HY = CL_RY + 2, HZ = ZS + 2;
    constexpr int HV = HX * HY * HZ;  // halo voxels
    extern __shared__ unsigned char smraw[];
    unsigned short* sm = reinterpret_cast<unsigned short*>(smraw);
    const int nzs = a.n2 / ZS, nys = a.n1 / CL_RY;
    const int zb = (int)(blockIdx.x % nzs);
    const int yb = (int)((blockIdx.x / nzs) % nys);
    const int xb = (int)(blockIdx.x / ((unsigned)nzs * nys));
    const int x0 = xb * CL_RX, y0 = yb * CL_RY, z0 = zb * ZS;
    // stage the halo block: CIN * 2 bytes per voxel, in 16-byte pieces
    constexpr int PPV = CIN * 2 / 16 > 0 ? CIN * 2 / 16 : 1;  // pieces per voxel (CIN = 4: half a piece)
    if (CIN == 16) {
        for (int e = threadIdx.x; e < HV * PPV; e += 256) {
            const int v = e / PPV, pc = e % PPV;
            const int hz = v % HZ, hy = (v / HZ) % HY, hx = v / (HZ * HY);
            const long long g = ((long long)wrapi(x0 + hx - 1, a.n0) * a.n1 + wrapi(y0 + hy - 1, a.n1)) * a.n2 +
                                wrapi(z0 + hz - 1, a.n2);
            reinterpret_cast<uint4*>(sm)[e] = reinterpret_cast<const uint4*>(a.in + g * CIN)[pc];
        }
    } else {
        for (int v = threadIdx.x; v < HV; v += 256) {
            const int hz = v % HZ, hy = (v / HZ) % HY, hx = v / (HZ * HY);
            const long long g = ((long long)wrapi(x0 + hx - 1, a.n0) * a.n1 + wrapi(y0 + hy - 1, a.n1)) * a.n2 +
                                wrapi(z0 + hz - 1, a.n2);
            reinterpret_cast<uint2*>(sm)[v] = *reinterpret_cast<const uint2*>(a.in + g * CIN);
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;  // wave wv: rows x0 + wv, y0 .. y0 + 3
    const int m = lane & 15, g = lane >> 4;
    f32x4 acc[CL_RY][TPW];
#pragma unroll
    for (int ry = 0; ry < CL_RY; ry++)
#pragma unroll
        for (int tt = 0; tt < TPW; tt++) acc[ry][tt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const u16x8* W = reinterpret_cast<const u16x8*>(a.wp);
    for (int s = 0; s < a.ks; s++) {
        const u16x8 bv = W[s * 64 + lane];
#pragma unroll
        for (int ry = 0; ry < CL_RY; ry++) {
            const auto row = [&](int t) -> long long {
                return (long long)((wv + 1 + t / 9 - 1) * HY + (ry + 1 + (t / 3) % 3 - 1)) * HZ;
            };
#pragma unroll
            for (int tt = 0; tt < TPW; tt++) {
                const u16x8 av =
                    conv_a<CIN>(sm, s, g, row, [&](int dz) { return (long long)(tt * 16 + m + 1 + dz); });
                acc[ry][tt] = mfma32(av, bv, acc[ry][tt]);
            }
        }
    }
    const int c = lane & 15;
    const float bb = a.bias ? a.bias[c] : 0.0f;
#pragma unroll
    for (int ry = 0; ry < CL_RY; ry++) {
        const long long r0 = ((long long)(x0 + wv) * a.n1 + (y0 + ry)) * a.n2 + z0;
#pragma unroll
        for (int tt = 0; tt < TPW; tt++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const long long v = r0 + tt * 16 + 4 * g + e;
                const float val = acc[ry][tt][e] + bb;
                if (a.mode == CONV_RELU_BF16) {
                    static_cast<unsigned short*>(a.out)[v * 16 + c] = f2bf(fmaxf(val, 0.0f));
                } else if (a.mode == CONV_F32) {
                    if (c < a.cout) static_cast<float*>(a.out)[v * a.ostride + c] = val;
                } else {
                    static_cast<unsigned short*>(a.out)[v * 16 + c] =
                        f2bf(bf2f(a.mask[v * 16 + c]) > 0.0f ? val : 0.0f);
                }
            }
    }
}

template <int CIN, int TPW> constexpr size_t conv_lds_bytes() {
    return (size_t)(CL_RX + 2) * (CL_RY + 2) * (16 * TPW + 2) * CIN * 2;
}

// ---------------------------------------------------------------------------
// weight gradient: per wave, over its tiles of 32 voxels (the K of one MFMA),
//   G[k][co] += sum_v X[v + off_t][ci] dZ[v][co],  k = t cin + ci,
// plus a bias block (row 0: sum_v dZ[v][co]); partials per wave.  Lane group g
// takes voxels 8 g .. 8 g + 7 of a tile, which lie in one z row (rows are
// multiples of 16 voxels).
// ---------------------------------------------------------------------------
struct DwArgs {
    const unsigned short* x;   // [ng][CIN]
    const unsigned short* dz;  // [ng][16]
    int n0, n1, n2;
    long long tiles, tiles_per_wave;
    float* part;               // [waves][NKB][16][16]
};

// One wave's share of the K blocks [KB0, KB1) (block NKB - 1: the bias).
template <int CIN, int KB0, int KB1>
__device__ __forceinline__ void dw_body(const DwArgs& a, long long wave) {
    constexpr int NKB = (27 * CIN + 15) / 16 + 1;
    constexpr int NB = KB1 - KB0;
    const int lane = threadIdx.x & 63;
    const int m = lane & 15, g = lane >> 4;
    f32x4 acc[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const long long t0 = wave * a.tiles_per_wave;
    const long long t1 = (t0 + a.tiles_per_wave < a.tiles) ? t0 + a.tiles_per_wave : a.tiles;
    const unsigned short o1 = m == 0 ? (unsigned short)0x3F80 : (unsigned short)0;
    const u16x8 one = u16x8{o1, o1, o1, o1, o1, o1, o1, o1};
    for (long long tile = t0; tile < t1; tile++) {
        const long long v0 = tile * 32 + 8 * g;
        const int z0 = (int)(v0 % a.n2);
        const long long row = v0 / a.n2;
        const int y = (int)(row % a.n1), x = (int)(row / a.n1);
        u16x8 bv;
#pragma unroll
        for (int e = 0; e < 8; e++) bv[e] = a.dz[(v0 + e) * 16 + m];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int kb = KB0 + q;
            if (kb == NKB - 1) {
                acc[q] = mfma32(one, bv, acc[q]);
                continue;
            }
            const int k = 16 * kb + m;
            const int t = k / CIN, ci = k % CIN;
            u16x8 av = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if (t < 27) {
                const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
                const long long rb = ((long long)wrapi(x + dx, a.n0) * a.n1 + wrapi(y + dy, a.n1)) * a.n2;
#pragma unroll
                for (int e = 0; e < 8; e++) av[e] = a.x[(rb + wrapi(z0 + e + dz, a.n2)) * CIN + ci];
            }
            acc[q] = mfma32(av, bv, acc[q]);
        }
    }
    float* P = a.part + wave * (long long)NKB * 256;
#pragma unroll
    for (int q = 0; q < NB; q++)
#pragma unroll
        for (int e = 0; e < 4; e++) P[((KB0 + q) * 16 + 4 * g + e) * 16 + m] = acc[q][e];
}

// Two workgroups per group of four waves: the even one takes the first half of
// the K blocks, the odd one the rest (half the accumulators per wave, so
// twice the waves fit a SIMD); each wave's blocks go to its slot of part.
template <int CIN>
__global__ __launch_bounds__(256, 4) void k_conv_dw(DwArgs a) {
    constexpr int NKB = (27 * CIN + 15) / 16 + 1;
    constexpr int NH = NKB / 2;
    const long long wave = (long long)(blockIdx.x >> 1) * 4 + (threadIdx.x >> 6);
    if (blockIdx.x & 1u)
        dw_body<CIN, NH, NKB>(a, wave);
    else
        dw_body<CIN, 0, NH>(a, wave);
}

// fixed-order FP64 sums over the waves in two stages: groups of DW_G waves
// (blockIdx.y = group), then the groups -> dW [27][cin][cout], db [cout]
constexpr int DW_G = 64;

__global__ void k_dw_reduce1(const float* __restrict__ part, long long waves, int nkb, double* __restrict__ gsum) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nkb * 256) return;
    const long long w0 = (long long)blockIdx.y * DW_G;
    const long long w1 = (w0 + DW_G < waves) ? w0 + DW_G : waves;
    double s = 0.0;
    for (long long w = w0; w < w1; w++) s += (double)part[w * (long long)nkb * 256 + idx];
    gsum[(long long)blockIdx.y * nkb * 256 + idx] = s;
}

__global__ void k_dw_reduce(const double* __restrict__ gsum, long long groups, int nkb, int cin, int cout,
                            float* __restrict__ dW, float* __restrict__ db) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nkb * 256) return;
    double s = 0.0;
    for (long long q = 0; q < groups; q++) s += gsum[q * (long long)nkb * 256 + idx];
    const int kb = idx / 256, i = (idx / 16) % 16, j = idx % 16;
    if (j >= cout) return;
    if (kb == nkb - 1) {
        if (i == 0) db[j] = (float)s;
        return;
    }
    const int k = 16 * kb + i;
    if (k < 27 * cin) dW[(size_t)k * cout + j] = (float)s;
}

// ---------------------------------------------------------------------------
// particles: periodic TSC gather and scatter, loss
// ---------------------------------------------------------------------------
struct Mesh3 {
    int n[3];
    double L[3];
};

// S_i = plain (channels 0..2 of feat) [+ psi_scale out]; with target: |S - T|^2
// per block and, with wxy / wz, the scaled residuals as the weights of the
// loss gradient's assignment: (coef r_x, coef r_y) and (coef r_z, 0), r
// clamped to +-L per component
__global__ __launch_bounds__(256) void k_shift_loss(Mesh3 g, const double* __restrict__ pos, long long np,
                                                    const float4* __restrict__ feat, const float4* __restrict__ out,
                                                    double psi_scale, const double* __restrict__ target,
                                                    double* __restrict__ s_plain, double* __restrict__ s_pred,
                                                    double coef, double* __restrict__ wxy, double* __restrict__ wz,
                                                    double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double sm[2][256];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    double l0 = 0.0, l1 = 0.0;
    if (i < np) {
        // the TSC stencil of the reference's reconstruction gather (particle_reconstruction.hpp:141-205,
        // 334-388): the 27 cells in (x, y, z) order, no wrap, a cell whose linear index leaves [0, n^3)
        // skipped (the library's sorted assignment uses the same cells and weights, so it is this gather's
        // exact adjoint); one x plane (9 cells) at a time with the plane's loads issued before any is added
        // (a 27-entry index list would sit in scratch memory)
        double wa[3][3];
        int i0[3];
        for (int a = 0; a < 3; a++) stencil_axis(2, axis_xp(pos[3 * i + a], g.n[a], g.L[a]), &i0[a], wa[a]);
        const long long ng = (long long)g.n[0] * g.n[1] * g.n[2];
        double sp[3] = {0.0, 0.0, 0.0}, so[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int ix = 0; ix < 3; ix++) {
            float4 fv[9], ov[9];
            bool ok[9];
#pragma unroll
            for (int jk = 0; jk < 9; jk++) {
                const int j = jk / 3, k = jk % 3;
                const long long v = ((long long)(i0[0] + ix) * g.n[1] + (i0[1] + j)) * g.n[2] + (i0[2] + k);
                ok[jk] = v >= 0 && v < ng;
                fv[jk] = ok[jk] ? feat[v] : make_float4(0.f, 0.f, 0.f, 0.f);
                ov[jk] = (ok[jk] && out) ? out[v] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int jk = 0; jk < 9; jk++) {
                if (!ok[jk]) continue;
                const double w = (wa[0][ix] * wa[1][jk / 3]) * wa[2][jk % 3];
                sp[0] += w * (double)fv[jk].x;
                sp[1] += w * (double)fv[jk].y;
                sp[2] += w * (double)fv[jk].z;
                if (out) {
                    so[0] += w * (double)ov[jk].x;
                    so[1] += w * (double)ov[jk].y;
                    so[2] += w * (double)ov[jk].z;
                }
            }
        }
        double r[3];
        for (int d = 0; d < 3; d++) {
            const double pl = sp[d], pr = pl + so[d] * psi_scale;  // features hold norm Psi in Mpc/h
            if (s_plain) s_plain[3 * i + d] = pl;
            if (s_pred) s_pred[3 * i + d] = pr;
            r[d] = 0.0;
            if (target) {
                const double t = target[3 * i + d];
                l0 += (pl - t) * (pl - t);
                l1 += (pr - t) * (pr - t);
                r[d] = fmin(fmax(pr - t, -g.L[d]), g.L[d]);
            }
        }
        if (wxy) {
            wxy[2 * i] = r[0] * coef;
            wxy[2 * i + 1] = r[1] * coef;
            wz[2 * i] = r[2] * coef;
            wz[2 * i + 1] = 0.0;
        }
    }
    sm[0][threadIdx.x] = l0;
    sm[1][threadIdx.x] = l1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + s];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && part) {
        part[2 * blockIdx.x] = sm[0][0];
        part[2 * blockIdx.x + 1] = sm[1][0];
    }
}

// fixed point -> BF16 dZ [ng][16] (channels 0..2) and FP32 copy [ng][4]
__global__ void k_grad_to_dz(const unsigned long long* __restrict__ G, long long ng, double inv_scale,
                             unsigned short* __restrict__ dz, float* __restrict__ g32) {
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < ng; v += st) {
        for (int c = 0; c < 16; c++) {
            float val = 0.0f;
            if (c < 3) val = (float)((double)(long long)G[c * ng + v] * inv_scale);
            dz[v * 16 + c] = f2bf(val);
            if (g32 && c < 4) g32[v * 4 + c] = c < 3 ? val : 0.0f;
        }
    }
}

// features [ng][4] FP32 -> network input (psi / psi_scale, delta / 8), BF16 [ng][4]
__global__ void k_feat_in(const float* __restrict__ f, long long ng, float inv_psi, unsigned short* __restrict__ x) {
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < ng; v += st) {
        for (int c = 0; c < 3; c++) x[4 * v + c] = f2bf(f[4 * v + c] * inv_psi);
        x[4 * v + 3] = f2bf(f[4 * v + 3] * 0.125f);
    }
}

__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mi = b1 * m[i] + (1.0f - b1) * g[i];
    const float vi = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

constexpr int NL = 4;
constexpr int CIN_L[NL] = {4, 16, 16, 16};
constexpr int COUT_L[NL] = {16, 16, 16, 3};

int ks_of(int cin) { return conv_steps(cin); }
size_t wcount(int l) { return (size_t)27 * CIN_L[l] * COUT_L[l]; }
size_t param_count() {
    size_t s = 0;
    for (int l = 0; l < NL; l++) s += wcount(l) + COUT_L[l];
    return s;
}

}  // namespace

struct hm_lrecon {
    hm_lrecon_config cfg{};
    Mesh mesh{};
    long long ng = 0;
    int tpw = 1;
    size_t np_ = 0;
    DevBuf params, grads, adam_m, adam_v, wp, wpt, bias16;  // FP32 master, packed BF16 forward / transposed
    DevBuf feat, x, act[3], out, dzA, dzB, G, part, gsum, lpart, wxy, wz, g32;
    long long part_waves = 0, tiles_per_wave = 0;
    ReconWork rw;
    int step = 0;
    size_t woff[NL], boff[NL];
};

namespace {

int ensure(hm_ctx* ctx, DevBuf& b, size_t bytes) { return dev_alloc(ctx, b, bytes); }

int repack(hm_ctx* ctx, hm_lrecon* h) {
    const float* P = (const float*)h->params.p;
    for (int l = 0; l < NL; l++) {
        const int ks = ks_of(CIN_L[l]), kst = ks_of(16);
        unsigned short* wp = (unsigned short*)h->wp.p + (size_t)l * WP_LAYER;
        unsigned short* wpt = (unsigned short*)h->wpt.p + (size_t)l * WP_LAYER;
        if (ctx->launch("lrecon_pack", [&] {
                hipLaunchKernelGGL(k_pack_w, dim3((ks * 512 + 255) / 256), dim3(256), 0, ctx->stream, P + h->woff[l],
                                   CIN_L[l], COUT_L[l], ks, 0, wp);
                if (l > 0)
                    hipLaunchKernelGGL(k_pack_w, dim3((kst * 512 + 255) / 256), dim3(256), 0, ctx->stream,
                                       P + h->woff[l], CIN_L[l], COUT_L[l], kst, 1, wpt);
            }))
            return -1;
        // biases padded to 16
        float* b16 = (float*)h->bias16.p + 16 * l;
        if (fill_zero(ctx, b16, sizeof(float) * 16)) return -1;
        HM_HIP_CHECK(ctx, hipMemcpyAsync(b16, P + h->boff[l], sizeof(float) * COUT_L[l], hipMemcpyDeviceToDevice,
                                         ctx->stream));
    }
    return 0;
}

template <int CIN, int TPW>
int conv_lds_launch(hm_ctx* ctx, const char* name, const ConvArgs& a) {
    const unsigned blocks = (unsigned)((a.n0 / CL_RX) * (a.n1 / CL_RY) * (a.n2 / (16 * TPW)));
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_conv_lds<CIN, TPW>), dim3(blocks), dim3(256), (conv_lds_bytes<CIN, TPW>()), ctx->stream, a);
    });
}

template <int CIN>
int conv_tpw(hm_ctx* ctx, const char* name, const ConvArgs& a, int tpw, long long waves) {
    if (ctx->conv_lds && a.n0 % CL_RX == 0 && a.n1 % CL_RY == 0 && a.n2 % (16 * tpw) == 0) {
        if (tpw == 4 && conv_lds_bytes<CIN, 4>() <= ctx->lds_max) return conv_lds_launch<CIN, 4>(ctx, name, a);
        if (tpw == 2 && conv_lds_bytes<CIN, 2>() <= ctx->lds_max) return conv_lds_launch<CIN, 2>(ctx, name, a);
        if (tpw == 1 && conv_lds_bytes<CIN, 1>() <= ctx->lds_max) return conv_lds_launch<CIN, 1>(ctx, name, a);
    }
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    return ctx->launch(name, [&] {
        if (tpw == 4)
            hipLaunchKernelGGL((k_conv<CIN, 4>), dim3(blocks), dim3(256), 0, ctx->stream, a);
        else if (tpw == 2)
            hipLaunchKernelGGL((k_conv<CIN, 2>), dim3(blocks), dim3(256), 0, ctx->stream, a);
        else
            hipLaunchKernelGGL((k_conv<CIN, 1>), dim3(blocks), dim3(256), 0, ctx->stream, a);
    });
}

int conv(hm_ctx* ctx, hm_lrecon* h, const char* name, int cin, const unsigned short* in, const unsigned short* wp,
         int ks, const float* bias, int mode, int cout, int ostride, void* out, const unsigned short* mask) {
    ConvArgs a;
    a.in = in;
    a.wp = wp;
    a.ks = ks;
    a.bias = bias;
    a.n0 = h->mesh.n[0];
    a.n1 = h->mesh.n[1];
    a.n2 = h->mesh.n[2];
    a.mode = mode;
    a.cout = cout;
    a.ostride = ostride;
    a.out = out;
    a.mask = mask;
    const long long waves = (long long)a.n0 * a.n1 * (a.n2 / (16 * h->tpw));
    ctx->conv_lds = h->cfg.conv_cache_taps == 0;
    return cin == 16 ? conv_tpw<16>(ctx, name, a, h->tpw, waves) : conv_tpw<4>(ctx, name, a, h->tpw, waves);
}

int conv_dw(hm_ctx* ctx, hm_lrecon* h, int l, const unsigned short* x, const unsigned short* dz, float* dW, float* db) {
    DwArgs a;
    a.x = x;
    a.dz = dz;
    a.n0 = h->mesh.n[0];
    a.n1 = h->mesh.n[1];
    a.n2 = h->mesh.n[2];
    a.tiles = h->ng / 32;
    a.tiles_per_wave = h->tiles_per_wave;
    a.part = (float*)h->part.p;
    const int cin = CIN_L[l];
    const int nkb = (27 * cin + 15) / 16 + 1;
    const unsigned blocks = (unsigned)((h->part_waves + 3) / 4);
    if (ctx->launch("lrecon_dw", [&] {
            if (cin == 16)
                hipLaunchKernelGGL(k_conv_dw<16>, dim3(2 * blocks), dim3(256), 0, ctx->stream, a);
            else
                hipLaunchKernelGGL(k_conv_dw<4>, dim3(2 * blocks), dim3(256), 0, ctx->stream, a);
        }))
        return -1;
    const long long waves = (long long)blocks * 4, groups = (waves + DW_G - 1) / DW_G;
    if (ctx->launch("lrecon_dw_reduce", [&] {
            hipLaunchKernelGGL(k_dw_reduce1, dim3((nkb * 256 + 255) / 256, (unsigned)groups), dim3(256), 0, ctx->stream,
                               (const float*)h->part.p, waves, nkb, (double*)h->gsum.p);
        }))
        return -1;
    return ctx->launch("lrecon_dw_reduce", [&] {
        hipLaunchKernelGGL(k_dw_reduce, dim3((nkb * 256 + 255) / 256), dim3(256), 0, ctx->stream,
                           (const double*)h->gsum.p, groups, nkb, cin, COUT_L[l], dW, db);
    });
}

// network forward on h->x (BF16 input) -> act[0..2], out (FP32 [ng][4])
int net_forward(hm_ctx* ctx, hm_lrecon* h) {
    const unsigned short* in = (const unsigned short*)h->x.p;
    const float* b16 = (const float*)h->bias16.p;
    for (int l = 0; l < NL; l++) {
        const unsigned short* wp = (const unsigned short*)h->wp.p + (size_t)l * WP_LAYER;
        if (l < NL - 1) {
            if (conv(ctx, h, "lrecon_conv", CIN_L[l], in, wp, ks_of(CIN_L[l]), b16 + 16 * l, CONV_RELU_BF16, 16, 16,
                     h->act[l].p, nullptr))
                return -1;
            in = (const unsigned short*)h->act[l].p;
        } else {
            if (fill_zero(ctx, h->out.p, sizeof(float) * 4 * (size_t)h->ng)) return -1;
            if (conv(ctx, h, "lrecon_conv", CIN_L[l], in, wp, ks_of(CIN_L[l]), b16 + 16 * l, CONV_F32, COUT_L[l], 4,
                     h->out.p, nullptr))
                return -1;
        }
    }
    return 0;
}

// backward from dZ of the last layer in h->dzA -> h->grads
int net_backward(hm_ctx* ctx, hm_lrecon* h) {
    float* G = (float*)h->grads.p;
    unsigned short* dz = (unsigned short*)h->dzA.p;
    unsigned short* dzn = (unsigned short*)h->dzB.p;
    for (int l = NL - 1; l >= 0; l--) {
        const unsigned short* xin = l == 0 ? (const unsigned short*)h->x.p : (const unsigned short*)h->act[l - 1].p;
        if (conv_dw(ctx, h, l, xin, dz, G + h->woff[l], G + h->boff[l])) return -1;
        if (l == 0) break;
        const unsigned short* wpt = (const unsigned short*)h->wpt.p + (size_t)l * WP_LAYER;
        if (conv(ctx, h, "lrecon_conv_dx", 16, dz, wpt, ks_of(16), nullptr, CONV_MASK_BF16, 16, 16, dzn,
                 (const unsigned short*)h->act[l - 1].p))
            return -1;
        std::swap(dz, dzn);
    }
    return 0;
}

int alloc_model(hm_ctx* ctx, hm_lrecon* h) {
    const long long ng = h->ng;
    const size_t P = param_count();
    if (ensure(ctx, h->params, sizeof(float) * P) || ensure(ctx, h->grads, sizeof(float) * P) ||
        ensure(ctx, h->adam_m, sizeof(float) * P) || ensure(ctx, h->adam_v, sizeof(float) * P) ||
        ensure(ctx, h->wp, sizeof(unsigned short) * NL * WP_LAYER) ||
        ensure(ctx, h->wpt, sizeof(unsigned short) * NL * WP_LAYER) || ensure(ctx, h->bias16, sizeof(float) * 16 * NL) ||
        ensure(ctx, h->feat, sizeof(float) * 4 * (size_t)ng) || ensure(ctx, h->x, sizeof(unsigned short) * 4 * (size_t)ng) ||
        ensure(ctx, h->out, sizeof(float) * 4 * (size_t)ng) || ensure(ctx, h->dzA, sizeof(unsigned short) * 16 * (size_t)ng) ||
        ensure(ctx, h->dzB, sizeof(unsigned short) * 16 * (size_t)ng) ||
        ensure(ctx, h->G, sizeof(unsigned long long) * 3 * (size_t)ng) || ensure(ctx, h->g32, sizeof(float) * 4 * (size_t)ng))
        return -1;
    for (int l = 0; l < 3; l++)
        if (ensure(ctx, h->act[l], sizeof(unsigned short) * 16 * (size_t)ng)) return -1;
    // weight-gradient partials: waves of >= 32 tiles of 32 voxels, at most 4096 waves
    const long long tiles = ng / 32;
    long long tpw = std::max(32LL, (tiles + 4095) / 4096);
    h->tiles_per_wave = tpw;
    h->part_waves = ((tiles + tpw - 1) / tpw + 3) / 4 * 4;
    if (ensure(ctx, h->part, sizeof(float) * (size_t)h->part_waves * 28 * 256)) return -1;
    if (ensure(ctx, h->gsum, sizeof(double) * (size_t)((h->part_waves + DW_G - 1) / DW_G) * 28 * 256)) return -1;
    return 0;
}

double splitmix_u(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * 0x1.0p-53;
}

// He-normal weights (std sqrt(2 / (27 cin))) for the first three layers, zero
// for the last (the network starts as the plain reconstruction), zero biases
std::vector<float> init_params(const hm_lrecon* h) {
    std::vector<float> p(param_count(), 0.0f);
    uint64_t s = h->cfg.init_seed ^ 0xA0761D6478BD642Full;
    for (int l = 0; l < NL - 1; l++) {
        const double sd = std::sqrt(2.0 / (27.0 * CIN_L[l]));
        for (size_t i = 0; i < wcount(l); i++) {
            const double u1 = splitmix_u(s), u2 = splitmix_u(s);
            p[h->woff[l] + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    }
    return p;
}

Mesh3 mesh3(const hm_lrecon* h) {
    Mesh3 g;
    for (int a = 0; a < 3; a++) {
        g.n[a] = h->mesh.n[a];
        g.L[a] = h->mesh.L[a];
    }
    return g;
}

// features and network input for a catalogue
int prepare(hm_ctx* ctx, hm_lrecon* h, const double* d_pos, long long n) {
    const double norm = h->mesh.volume() / (double)n;
    const int rc = h->cfg.precision == HM_FP64
                       ? recon_features<double>(ctx, h->mesh, h->cfg.b1, h->cfg.RG, d_pos, n, norm, (float*)h->feat.p, h->rw)
                       : recon_features<float>(ctx, h->mesh, h->cfg.b1, h->cfg.RG, d_pos, n, norm, (float*)h->feat.p, h->rw);
    if (rc) return -1;
    const unsigned gb = (unsigned)std::min<long long>((h->ng + 255) / 256, 16384);
    return ctx->launch("lrecon_feat", [&] {
        hipLaunchKernelGGL(k_feat_in, dim3(gb), dim3(256), 0, ctx->stream, (const float*)h->feat.p, h->ng,
                           (float)(1.0 / h->cfg.psi_scale), (unsigned short*)h->x.p);
    });
}

// shifts and loss; with grad, the output-mesh gradient into dzA: the residuals
// assigned with the sorted fixed-point assignment (the gather's adjoint)
int shifts(hm_ctx* ctx, hm_lrecon* h, const double* d_pos, long long n, const double* target, double* s_plain,
           double* s_pred, bool use_net, double* loss_plain, double* loss_pred, bool grad) {
    const long long blocks = (n + 255) / 256;
    if (target && (ensure(ctx, h->lpart, sizeof(double) * 2 * (size_t)blocks) ||
                   (grad && (ensure(ctx, h->wxy, sizeof(double) * 2 * (size_t)n) ||
                             ensure(ctx, h->wz, sizeof(double) * 2 * (size_t)n)))))
        return -1;
    const Mesh3 g = mesh3(h);
    // d loss / d out = (2 psi_scale / N) sum_i W r_i; |sum| <= 2 psi_scale max|r| <= 2 psi_scale L
    const double coef = 2.0 * h->cfg.psi_scale / (double)n;
    const double bound = 2.0 * h->cfg.psi_scale * std::max(h->mesh.L[0], std::max(h->mesh.L[1], h->mesh.L[2])) + 1.0;
    const double scale = std::ldexp(1.0, (int)std::floor(std::log2(std::ldexp(1.0, 61) / bound)));
    if (ctx->launch("lrecon_shift", [&] {
            hipLaunchKernelGGL(k_shift_loss, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, g, d_pos, n,
                               (const float4*)h->feat.p, use_net ? (const float4*)h->out.p : nullptr, h->cfg.psi_scale,
                               target, s_plain, s_pred, coef, grad ? (double*)h->wxy.p : nullptr,
                               grad ? (double*)h->wz.p : nullptr, target ? (double*)h->lpart.p : nullptr);
        }))
        return -1;
    if (target) {
        std::vector<double> lp(2 * (size_t)blocks);
        HM_HIP_CHECK(ctx, hipMemcpyAsync(lp.data(), h->lpart.p, sizeof(double) * lp.size(), hipMemcpyDeviceToHost,
                                         ctx->stream));
        HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        double a = 0.0, b = 0.0;
        for (long long i = 0; i < blocks; i++) {
            a += lp[2 * i];
            b += lp[2 * i + 1];
        }
        if (loss_plain) *loss_plain = a / (double)n;
        if (loss_pred) *loss_pred = b / (double)n;
    }
    if (!grad) return 0;
    AssignGeom ag;
    for (int a = 0; a < 3; a++) {
        ag.n[a] = h->mesh.n[a];
        ag.L[a] = h->mesh.L[a];
    }
    ag.ng = h->ng;
    ag.assign = HM_TSC;
    if (sort_work_alloc(ctx, h->rw.sort, ag, n, true)) return -1;
    long long* G = (long long*)h->G.p;
    if (assign_sorted(ctx, ag, d_pos, (const double*)h->wxy.p, n, scale, G, G + h->ng, h->rw.sort) ||
        assign_sorted(ctx, ag, d_pos, (const double*)h->wz.p, n, scale, G + 2 * h->ng, nullptr, h->rw.sort, true))
        return -1;
    const unsigned gb = (unsigned)std::min<long long>((h->ng + 255) / 256, 16384);
    return ctx->launch("lrecon_grad_dz", [&] {
        hipLaunchKernelGGL(k_grad_to_dz, dim3(gb), dim3(256), 0, ctx->stream, (const unsigned long long*)h->G.p, h->ng,
                           1.0 / scale, (unsigned short*)h->dzA.p, (float*)h->g32.p);
    });
}

int adam(hm_ctx* ctx, hm_lrecon* h) {
    h->step++;
    const int P = (int)param_count();
    const float c1 = (float)(1.0 - std::pow(0.9, h->step)), c2 = (float)(1.0 - std::pow(0.999, h->step));
    if (ctx->launch("lrecon_adam", [&] {
            hipLaunchKernelGGL(k_adam, dim3((P + 255) / 256), dim3(256), 0, ctx->stream, P, (float*)h->params.p,
                               (const float*)h->grads.p, (float*)h->adam_m.p, (float*)h->adam_v.p, (float)h->cfg.lr, c1,
                               c2);
        }))
        return -1;
    return repack(ctx, h);
}

void free_model(hm_lrecon* h) {
    DevBuf* b[] = {&h->params, &h->grads, &h->adam_m, &h->adam_v, &h->wp, &h->wpt, &h->bias16, &h->feat, &h->x,
                   &h->out, &h->dzA, &h->dzB, &h->G, &h->part, &h->gsum, &h->lpart, &h->wxy, &h->wz, &h->g32,
                   &h->act[0], &h->act[1], &h->act[2]};
    for (DevBuf* d : b) dev_free(*d);
    recon_work_free(h->rw);
    delete h;
}

}  // namespace

extern "C" {

hm_lrecon* hm_lrecon_create(hm_ctx* ctx, const hm_lrecon_config* cfg) {
    if (!ctx || !cfg) return nullptr;
    const int n = cfg->n_mesh;
    if (n < 16 || n % 16 || !fft_length_supported(n))
        return bad(ctx, "learned reconstruction: n_mesh must be a multiple of 16 and an FFT length of this library"),
               nullptr;
    if (!(cfg->boxsize > 0.0) || !(cfg->psi_scale > 0.0) || !(cfg->b1 != 0.0) || cfg->RG < 0.0 || !(cfg->lr >= 0.0))
        return bad(ctx, "learned reconstruction: boxsize, psi_scale, lr must be positive, b1 non-zero, RG >= 0"),
               nullptr;
    (void)hipSetDevice(ctx->device);
    auto* h = new (std::nothrow) hm_lrecon();
    if (!h) return nullptr;
    h->cfg = *cfg;
    for (int a = 0; a < 3; a++) {
        h->mesh.n[a] = n;
        h->mesh.L[a] = cfg->boxsize;
    }
    h->mesh.assign = HM_TSC;
    h->ng = (long long)n * n * n;
    // 16 z per wave: with LDS staging a 4 x 4 x 16 block (+ halo) takes 21 KB,
    // so seven workgroups fit a CU (4 x 4 x 64 took 76 KB and allowed two: a
    // training step took 21.9 ms against 18.5); the cache-read path takes 64
    h->tpw = cfg->conv_cache_taps ? (n % 64 == 0 ? 4 : (n % 32 == 0 ? 2 : 1)) : 1;
    size_t o = 0;
    for (int l = 0; l < NL; l++) {
        h->woff[l] = o;
        o += wcount(l);
        h->boff[l] = o;
        o += COUT_L[l];
    }
    h->np_ = o;
    if (alloc_model(ctx, h)) {
        free_model(h);
        return nullptr;
    }
    const std::vector<float> p = init_params(h);
    if (hm_lrecon_set_params(ctx, h, p.data())) {
        free_model(h);
        return nullptr;
    }
    return h;
}

void hm_lrecon_destroy(hm_ctx* ctx, hm_lrecon* h) {
    if (!h) return;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    free_model(h);
}

long long hm_lrecon_n_params(const hm_lrecon* h) { return h ? (long long)h->np_ : -1; }

int hm_lrecon_get_params(hm_ctx* ctx, hm_lrecon* h, float* out) {
    if (!ctx || !h || !out) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(out, h->params.p, sizeof(float) * h->np_, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_lrecon_set_params(hm_ctx* ctx, hm_lrecon* h, const float* in) {
    if (!ctx || !h || !in) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(h->params.p, in, sizeof(float) * h->np_, hipMemcpyHostToDevice, ctx->stream));
    if (fill_zero(ctx, h->adam_m.p, sizeof(float) * h->np_) || fill_zero(ctx, h->adam_v.p, sizeof(float) * h->np_))
        return -1;
    h->step = 0;
    if (repack(ctx, h)) return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_lrecon_net_forward(hm_ctx* ctx, hm_lrecon* h, const float* d_x, float* d_out) {
    if (!ctx || !h || !d_x || !d_out) return -1;
    const unsigned gb = (unsigned)std::min<long long>((h->ng * 4 + 255) / 256, 16384);
    if (ctx->launch("lrecon_to_bf16", [&] {
            hipLaunchKernelGGL(k_to_bf16, dim3(gb), dim3(256), 0, ctx->stream, d_x, h->ng, 4, 4,
                               (unsigned short*)h->x.p);
        }))
        return -1;
    if (net_forward(ctx, h)) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(d_out, h->out.p, sizeof(float) * 4 * (size_t)h->ng, hipMemcpyDeviceToDevice,
                                     ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_lrecon_net_backward(hm_ctx* ctx, hm_lrecon* h, const float* d_x, const float* d_gout, float* grad) {
    if (!ctx || !h || !d_x || !d_gout || !grad) return -1;
    if (hm_lrecon_net_forward(ctx, h, d_x, (float*)h->g32.p)) return -1;
    const unsigned gb = (unsigned)std::min<long long>((h->ng * 16 + 255) / 256, 16384);
    if (ctx->launch("lrecon_to_bf16", [&] {
            hipLaunchKernelGGL(k_to_bf16, dim3(gb), dim3(256), 0, ctx->stream, d_gout, h->ng, 4, 16,
                               (unsigned short*)h->dzA.p);
        }))
        return -1;
    if (net_backward(ctx, h)) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(grad, h->grads.p, sizeof(float) * h->np_, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_lrecon_features(hm_ctx* ctx, hm_lrecon* h, const double* d_pos, long long n, float* d_feat) {
    if (!ctx || !h || !d_pos || n <= 0 || !d_feat) return bad(ctx, "learned reconstruction: bad arguments");
    if (prepare(ctx, h, d_pos, n)) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(d_feat, h->feat.p, sizeof(float) * 4 * (size_t)h->ng, hipMemcpyDeviceToDevice,
                                     ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_lrecon_predict(hm_ctx* ctx, hm_lrecon* h, const double* d_pos, long long n, double* d_shift_plain,
                      double* d_shift_pred) {
    if (!ctx || !h || !d_pos || n <= 0) return bad(ctx, "learned reconstruction: bad arguments");
    if (prepare(ctx, h, d_pos, n) || net_forward(ctx, h)) return -1;
    if (shifts(ctx, h, d_pos, n, nullptr, d_shift_plain, d_shift_pred, true, nullptr, nullptr, false)) return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_lrecon_gradient(hm_ctx* ctx, hm_lrecon* h, const double* d_pos, const double* d_target, long long n,
                       double* loss_plain, double* loss_pred, float* grad, float* d_gout) {
    if (!ctx || !h || !d_pos || !d_target || n <= 0) return bad(ctx, "learned reconstruction: bad arguments");
    if (prepare(ctx, h, d_pos, n) || net_forward(ctx, h)) return -1;
    if (shifts(ctx, h, d_pos, n, d_target, nullptr, nullptr, true, loss_plain, loss_pred, true)) return -1;
    if (d_gout)
        HM_HIP_CHECK(ctx, hipMemcpyAsync(d_gout, h->g32.p, sizeof(float) * 4 * (size_t)h->ng, hipMemcpyDeviceToDevice,
                                         ctx->stream));
    if (net_backward(ctx, h)) return -1;
    if (grad)
        HM_HIP_CHECK(ctx, hipMemcpyAsync(grad, h->grads.p, sizeof(float) * h->np_, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_lrecon_train_step(hm_ctx* ctx, hm_lrecon* h, const double* d_pos, const double* d_target, long long n,
                         double* loss_plain, double* loss_pred) {
    if (hm_lrecon_gradient(ctx, h, d_pos, d_target, n, loss_plain, loss_pred, nullptr, nullptr)) return -1;
    if (adam(ctx, h)) return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
