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

// The stencil contributions of one particle, each rounded to the nearest
// multiple of 2^-s and added with a 64-bit integer atomic.
template <int S>
__device__ __forceinline__ void scatter_particle(const double* __restrict__ pos, const double* __restrict__ w,
                                                 long long q, const AssignGeom& g, double scale,
                                                 unsigned long long* __restrict__ gre,
                                                 unsigned long long* __restrict__ gim) {
#pragma clang fp contract(off)
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

template <int S>
__global__ __launch_bounds__(256) void k_assign_fixed(const double* __restrict__ pos, const double* __restrict__ w,
                                                     long long np, AssignGeom g, double scale,
                                                     unsigned long long* __restrict__ gre,
                                                     unsigned long long* __restrict__ gim) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    scatter_particle<S>(pos, w, q, g, scale, gre, gim);
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

// ---------------------------------------------------------------------------
// sort-by-cell assignment
// ---------------------------------------------------------------------------
// Output tile per workgroup: GX x GY x GZ cells.  Particles are sorted by
// the (x, y, z / GZ) chunk of their first stencil cell.
constexpr int GX = 4, GY = 8, GZ = 64;

__host__ __device__ inline int chunks_z(const AssignGeom& g) { return (g.n[2] + GZ - 1) / GZ; }
__host__ __device__ inline long long n_keys(const AssignGeom& g) { return (long long)g.n[0] * g.n[1] * chunks_z(g); }

template <int S>
__device__ __forceinline__ bool first_cell(const double* __restrict__ pos, long long q, const AssignGeom& g,
                                           long long* key) {
#pragma clang fp contract(off)
    double wt[3];
    int iw0[3];
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        stencil_axis(g.assign, axis_xp(pos[3 * q + a], g.n[a], g.L[a]), &iw0[a], wt);
        inside = inside && iw0[a] >= 0 && iw0[a] + S - 1 <= g.n[a] - 1;
    }
    *key = ((long long)iw0[0] * g.n[1] + iw0[1]) * chunks_z(g) + iw0[2] / GZ;
    return inside;
}

// Count particles per first stencil cell; list the border particles
// (one counter atomic per wave for the border list).
template <int S>
__global__ __launch_bounds__(256) void k_sort_count(const double* __restrict__ pos, long long np, AssignGeom g,
                                                   int* __restrict__ count, int* __restrict__ slot,
                                                   int* __restrict__ border) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool is_border = false;
    if (q < np) {
        long long key;
        if (first_cell<S>(pos, q, g, &key)) {
            slot[q] = atomicAdd(count + key, 1);
        } else {
            slot[q] = -1;
            is_border = true;
        }
    }
    const unsigned long long mask = __ballot(is_border);
    if (mask) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((long long)mask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(border, (int)__popcll(mask));
        base = __shfl(base, leader);
        if (is_border) {
            const unsigned long long below = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
            border[1 + base + (int)__popcll(below)] = (int)q;
        }
    }
}

constexpr int SCAN_NT = 256, SCAN_ITEMS = 16, SCAN_BLK = SCAN_NT * SCAN_ITEMS;

__global__ __launch_bounds__(SCAN_NT) void k_scan_reduce(const int* __restrict__ a, long long n,
                                                        int* __restrict__ bsum) {
    __shared__ int sm[SCAN_NT];
    const long long base = (long long)blockIdx.x * SCAN_BLK;
    int s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const long long k = base + i * SCAN_NT + threadIdx.x;
        if (k < n) s += a[k];
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = SCAN_NT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sm[0];
}

// exclusive scan of the block sums in one workgroup
__global__ __launch_bounds__(1024) void k_scan_bsum(int* __restrict__ b, int nb) {
    __shared__ int sm[1024];
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        const int v = (i < nb) ? b[i] : 0;
        sm[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int t = (threadIdx.x >= off) ? sm[threadIdx.x - off] : 0;
            __syncthreads();
            sm[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) b[i] = carry + sm[threadIdx.x] - v;
        carry += sm[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(int* __restrict__ a, long long n,
                                                       const int* __restrict__ bsum) {
    __shared__ int tile[SCAN_BLK];
    __shared__ int tsum[SCAN_NT];
    const long long base = (long long)blockIdx.x * SCAN_BLK;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int e = i * SCAN_NT + threadIdx.x;
        tile[e] = (base + e < n) ? a[base + e] : 0;
    }
    __syncthreads();
    int run = 0;
    int loc[SCAN_ITEMS];
    for (int i = 0; i < SCAN_ITEMS; i++) {
        loc[i] = run;
        run += tile[threadIdx.x * SCAN_ITEMS + i];
    }
    tsum[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < SCAN_NT; off <<= 1) {
        const int t = (threadIdx.x >= off) ? tsum[threadIdx.x - off] : 0;
        __syncthreads();
        tsum[threadIdx.x] += t;
        __syncthreads();
    }
    const int off = bsum[blockIdx.x] + tsum[threadIdx.x] - run;
    for (int i = 0; i < SCAN_ITEMS; i++) tile[threadIdx.x * SCAN_ITEMS + i] = off + loc[i];
    __syncthreads();
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int e = i * SCAN_NT + threadIdx.x;
        if (base + e < n) a[base + e] = tile[e];
    }
}

// Each interior particle to its place in cell order, start[key] + rank; the
// place replaces the rank in slot, so that a later assignment of the same
// positions moves only its weights (k_sort_move_w).
template <int S>
__global__ __launch_bounds__(256) void k_sort_scatter(const double* __restrict__ pos, const double* __restrict__ w,
                                                     long long np, AssignGeom g, const int* __restrict__ start,
                                                     int* __restrict__ slot, double* __restrict__ spos,
                                                     double* __restrict__ sw) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int r = slot[q];
    if (r < 0) return;
    long long key;
    first_cell<S>(pos, q, g, &key);
    const long long d = (long long)start[key] + r;
    slot[q] = (int)d;
    spos[3 * d] = pos[3 * q];
    spos[3 * d + 1] = pos[3 * q + 1];
    spos[3 * d + 2] = pos[3 * q + 2];
    if (w) {
        sw[2 * d] = w[2 * q];
        sw[2 * d + 1] = w[2 * q + 1];
    }
}

__global__ __launch_bounds__(256) void k_sort_move_w(const double* __restrict__ w, long long np,
                                                    const int* __restrict__ place, double* __restrict__ sw) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int d = place[q];
    if (d < 0) return;
    sw[2 * (long long)d] = w[2 * q];
    sw[2 * (long long)d + 1] = w[2 * q + 1];
}

// One workgroup owns one output tile.  The interior particles whose first
// stencil cell lies in the tile extended by S-1 cells towards -x, -y, -z are
// exactly those that can reach the tile; in key order they form one
// contiguous run per extended (x, y) row (the z chunks of the tile and the
// one before it), read from the prefix array.  Each
// such particle adds its in-tile stencil contributions to an LDS tile with
// 64-bit integer atomics (order-independent), and the tile is then stored
// once with plain writes.
template <int S, bool CPLX>
__global__ __launch_bounds__(256) void k_sort_gather(const int* __restrict__ start, const double* __restrict__ spos,
                                                    const double* __restrict__ sw, AssignGeom g, double scale,
                                                    long long* __restrict__ out_re, long long* __restrict__ out_im) {
#pragma clang fp contract(off)
    constexpr int RX = GX + S - 1, RY = GY + S - 1, NR = RX * RY, NC = GX * GY * GZ;
    __shared__ unsigned long long acc[CPLX ? 2 * NC : NC];
    __shared__ int rbeg[NR], roff[NR + 1];
    const int x0 = blockIdx.z * GX, y0 = blockIdx.y * GY, z0 = blockIdx.x * GZ;
    for (int e = threadIdx.x; e < (CPLX ? 2 * NC : NC); e += 256) acc[e] = 0ull;
    if (threadIdx.x < NR) {
        const int r = threadIdx.x;
        const int xx = x0 - (S - 1) + r / RY, yy = y0 - (S - 1) + r % RY;
        const int zlo = (z0 - (S - 1) > 0) ? z0 - (S - 1) : 0;
        const int zhi = (z0 + GZ - 1 < g.n[2] - S) ? z0 + GZ - 1 : g.n[2] - S;
        int b = 0, n = 0;
        if (xx >= 0 && xx <= g.n[0] - S && yy >= 0 && yy <= g.n[1] - S && zlo <= zhi) {
            const long long lo = ((long long)xx * g.n[1] + yy) * chunks_z(g);
            b = start[lo + zlo / GZ];
            n = start[lo + zhi / GZ + 1] - b;
        }
        rbeg[r] = b;
        roff[r + 1] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        roff[0] = 0;
        for (int r = 0; r < NR; r++) roff[r + 1] += roff[r];
    }
    __syncthreads();
    const int total = roff[NR];
    // j only grows, so the row search resumes where the previous particle's ended
    int r = 0;
    for (int j = threadIdx.x; j < total; j += 256) {
        while (roff[r + 1] <= j) r++;
        const long long p = (long long)rbeg[r] + (j - roff[r]);
        double wt[3][3];
        int iw0[3];
        for (int a = 0; a < 3; a++) stencil_axis(g.assign, axis_xp(spos[3 * p + a], g.n[a], g.L[a]), &iw0[a], wt[a]);
        const double wre = sw ? sw[2 * p] : 1.0;
        const double wim = sw ? sw[2 * p + 1] : 0.0;
        for (int i = 0; i < S; i++) {
            const int ox = iw0[0] + i - x0;
            if (ox < 0 || ox >= GX) continue;
            for (int jj = 0; jj < S; jj++) {
                const int oy = iw0[1] + jj - y0;
                if (oy < 0 || oy >= GY) continue;
                const double wij = wt[0][i] * wt[1][jj];
                for (int k = 0; k < S; k++) {
                    const int oz = iw0[2] + k - z0;
                    if (oz < 0 || oz >= GZ) continue;
                    const double wijk = wij * wt[2][k];
                    const int c = (ox * GY + oy) * GZ + oz;
                    atomicAdd(&acc[c], (unsigned long long)__double2ll_rn(wre * wijk * scale));
                    if (CPLX) atomicAdd(&acc[NC + c], (unsigned long long)__double2ll_rn(wim * wijk * scale));
                }
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NC; e += 256) {
        const int x = x0 + e / (GY * GZ), y = y0 + (e / GZ) % GY, z = z0 + e % GZ;
        if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) continue;
        const long long lin = ((long long)x * g.n[1] + y) * g.n[2] + z;
        out_re[lin] = (long long)acc[e];
        if (CPLX) out_im[lin] = (long long)acc[NC + e];
    }
}

template <int S>
__global__ __launch_bounds__(256) void k_assign_border(const double* __restrict__ pos, const double* __restrict__ w,
                                                      const int* __restrict__ border, AssignGeom g, double scale,
                                                      unsigned long long* __restrict__ gre,
                                                      unsigned long long* __restrict__ gim) {
    const int nb = border[0];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x)
        scatter_particle<S>(pos, w, border[1 + i], g, scale, gre, gim);
}

int sort_work_alloc(hm_ctx* ctx, SortWork& w, const AssignGeom& g, long long np, bool weights) {
    w.sig.valid = false;  // the buffers may move: the next assignment sorts afresh
    const long long nk = n_keys(g);
    const long long nb = (nk + 1 + SCAN_BLK - 1) / SCAN_BLK;
    if (dev_alloc(ctx, w.start, sizeof(int) * (size_t)(nk + 1))) return -1;
    if (dev_alloc(ctx, w.slot, sizeof(int) * (size_t)(np > 0 ? np : 1))) return -1;
    if (dev_alloc(ctx, w.spos, sizeof(double) * 3 * (size_t)(np > 0 ? np : 1))) return -1;
    if (weights && dev_alloc(ctx, w.sw, sizeof(double) * 2 * (size_t)(np > 0 ? np : 1))) return -1;
    if (dev_alloc(ctx, w.border, sizeof(int) * (size_t)(np + 1))) return -1;
    if (dev_alloc(ctx, w.bsum, sizeof(int) * (size_t)nb)) return -1;
    return 0;
}

size_t sort_work_carve(const AssignGeom& g, long long np, bool weights, char* base, SortWork* w) {
    const long long nk = n_keys(g);
    const long long nb = (nk + 1 + SCAN_BLK - 1) / SCAN_BLK;
    const size_t npp = (size_t)(np > 0 ? np : 1);
    const size_t sizes[6] = {sizeof(int) * (size_t)(nk + 1), sizeof(int) * npp, sizeof(double) * 3 * npp,
                             weights ? sizeof(double) * 2 * npp : 0, sizeof(int) * (size_t)(np + 1),
                             sizeof(int) * (size_t)nb};
    DevBuf* parts[6] = {&w->start, &w->slot, &w->spos, &w->sw, &w->border, &w->bsum};
    size_t off = 0;
    for (int i = 0; i < 6; i++) {
        if (base) {
            parts[i]->p = sizes[i] ? base + off : nullptr;
            parts[i]->bytes = sizes[i];
        }
        off += (sizes[i] + 255) & ~(size_t)255;
    }
    return off;
}

void sort_work_free(SortWork& w) {
    dev_free(w.start);
    dev_free(w.slot);
    dev_free(w.spos);
    dev_free(w.sw);
    dev_free(w.border);
    dev_free(w.bsum);
}

template <int S>
static int assign_sorted_S(hm_ctx* ctx, const AssignGeom& g, const double* d_pos, const double* d_w, long long np,
                           double scale, long long* d_re, long long* d_im, SortWork& w, bool reuse) {
    int* start = (int*)w.start.p;
    int* slot = (int*)w.slot.p;
    int* border = (int*)w.border.p;
    int* bsum = (int*)w.bsum.p;
    double* spos = (double*)w.spos.p;
    double* sw = d_w ? (double*)w.sw.p : nullptr;
    auto* re = reinterpret_cast<unsigned long long*>(d_re);
    auto* im = reinterpret_cast<unsigned long long*>(d_im);
    const long long n = n_keys(g) + 1;
    const unsigned nb = (unsigned)((n + SCAN_BLK - 1) / SCAN_BLK);
    const unsigned pb = (unsigned)((np + 255) / 256);
    if (!reuse) {
        if (fill_zero(ctx, start, sizeof(int) * (size_t)n)) return -1;
        if (fill_zero(ctx, border, sizeof(int))) return -1;
        if (np > 0 && ctx->launch("assign_count", [&] {
                hipLaunchKernelGGL(k_sort_count<S>, dim3(pb), dim3(256), 0, ctx->stream, d_pos, np, g, start, slot,
                                   border);
            }))
            return -1;
        if (ctx->launch("assign_scan", [&] {
                hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_NT), 0, ctx->stream, start, n, bsum);
                hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(1024), 0, ctx->stream, bsum, (int)nb);
                hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_NT), 0, ctx->stream, start, n, bsum);
            }))
            return -1;
    }
    if (np > 0 && !reuse && ctx->launch("assign_scatter", [&] {
            hipLaunchKernelGGL(k_sort_scatter<S>, dim3(pb), dim3(256), 0, ctx->stream, d_pos, d_w, np, g, start, slot,
                               spos, sw);
        }))
        return -1;
    if (np > 0 && reuse && sw && ctx->launch("assign_scatter", [&] {
            hipLaunchKernelGGL(k_sort_move_w, dim3(pb), dim3(256), 0, ctx->stream, d_w, np, (const int*)slot, sw);
        }))
        return -1;
    const dim3 gg((unsigned)((g.n[2] + GZ - 1) / GZ), (unsigned)((g.n[1] + GY - 1) / GY),
                  (unsigned)((g.n[0] + GX - 1) / GX));
    if (ctx->launch("assign_gather", [&] {
            if (d_im)
                hipLaunchKernelGGL((k_sort_gather<S, true>), gg, dim3(256), 0, ctx->stream, start, spos, sw, g, scale,
                                   d_re, d_im);
            else
                hipLaunchKernelGGL((k_sort_gather<S, false>), gg, dim3(256), 0, ctx->stream, start, spos, sw, g, scale,
                                   d_re, d_im);
        }))
        return -1;
    if (np > 0 && ctx->launch("assign_border", [&] {
            hipLaunchKernelGGL(k_assign_border<S>, dim3(256), dim3(256), 0, ctx->stream, d_pos, d_w, border, g, scale, re,
                               im);
        }))
        return -1;
    return 0;
}

int assign_sorted(hm_ctx* ctx, const AssignGeom& g, const double* d_pos, const double* d_w, long long np,
                  double scale, long long* d_re, long long* d_im, SortWork& w, bool reuse) {
    if (np > 0x7ffffff0LL || g.ng > 0x7ffffff0LL) {
        ctx->set_error("assign: more than 2^31 particles or cells");
        return -1;
    }
    SortWork::Sig& sg = w.sig;
    if (reuse) {
        bool same = sg.valid && sg.pos == d_pos && sg.np == np && sg.assign == g.assign && sg.weights && d_w;
        for (int a = 0; a < 3; a++) same = same && sg.n[a] == g.n[a] && sg.L[a] == g.L[a];
        if (!same) {
            ctx->set_error("assign: a sort is reused for other positions, count, mesh or without weights");
            return -1;
        }
    } else {
        sg.valid = true;
        sg.pos = d_pos;
        sg.np = np;
        sg.assign = g.assign;
        sg.weights = d_w != nullptr;
        for (int a = 0; a < 3; a++) {
            sg.n[a] = g.n[a];
            sg.L[a] = g.L[a];
        }
    }
    if (g.assign == 0) return assign_sorted_S<1>(ctx, g, d_pos, d_w, np, scale, d_re, d_im, w, reuse);
    if (g.assign == 1) return assign_sorted_S<2>(ctx, g, d_pos, d_w, np, scale, d_re, d_im, w, reuse);
    return assign_sorted_S<3>(ctx, g, d_pos, d_w, np, scale, d_re, d_im, w, reuse);
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

// The same conversion where the int64 cells and the FP64 densities share
// memory: each lane reads a cell and overwrites it with its density.
__global__ void k_fixed_to_density_inplace(long long* q, long long n, double inv_scale, double factor) {
#pragma clang fp contract(off)
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long st = (long long)gridDim.x * blockDim.x;
    double* out = reinterpret_cast<double*>(q);
    for (long long i = i0; i < n; i += st) {
        const long long v = q[i];
        out[i] = ((double)v * inv_scale) * factor;
    }
}

int fixed_to_density_inplace(hm_ctx* ctx, long long* d_q, long long n, double inv_scale, double factor) {
    long long blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    return ctx->launch("fixed_to_density", [&] {
        hipLaunchKernelGGL(k_fixed_to_density_inplace, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d_q, n,
                           inv_scale, factor);
    });
}

}  // namespace hm
