// hm_assign.hip -- mass assignment kernels (field.hpp:47-184).
#include "hm_assign.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

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
// the (x, y, z / GZ) chunk of their first stencil cell, and within a chunk
// the last S - 1 first cells (the tail: the only ones whose stencil reaches
// the next chunk) carry a key of their own, right after the chunk's head.
constexpr int GX = 4, GY = 8, GZ = 64;

__host__ __device__ inline int chunks_z(const AssignGeom& g) { return (g.n[2] + GZ - 1) / GZ; }
__host__ __device__ inline long long keys_row(const AssignGeom& g) { return 2LL * chunks_z(g); }
__host__ __device__ inline long long n_keys(const AssignGeom& g) { return (long long)g.n[0] * g.n[1] * keys_row(g); }

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
    const int tail = (S > 1 && iw0[2] % GZ >= GZ - (S - 1)) ? 1 : 0;
    *key = (((long long)iw0[0] * g.n[1] + iw0[1]) * chunks_z(g) + iw0[2] / GZ) * 2 + tail;
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
// contiguous run per extended (x, y) row (the tail key of the chunk before
// the tile, then both keys of the tile's chunk), read from the prefix array.
// Each such particle adds its in-tile stencil contributions to an LDS tile
// with 64-bit integer atomics (order-independent), and the tile is then
// stored once with plain writes.
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
    // the first wave finds each extended row's run and scans the run lengths (NR <= 64)
    static_assert(NR <= 64, "one lane per extended row");
    if (threadIdx.x < 64) {
        const int r = threadIdx.x;
        int b = 0, n = 0;
        if (r < NR) {
            const int xx = x0 - (S - 1) + r / RY, yy = y0 - (S - 1) + r % RY;
            const int zlo = (z0 - (S - 1) > 0) ? z0 - (S - 1) : 0;
            const int zhi = (z0 + GZ - 1 < g.n[2] - S) ? z0 + GZ - 1 : g.n[2] - S;
            if (xx >= 0 && xx <= g.n[0] - S && yy >= 0 && yy <= g.n[1] - S && zlo <= zhi) {
                const long long lo = ((long long)xx * g.n[1] + yy) * chunks_z(g);
                const int cl = zlo / GZ, ch = zhi / GZ;
                b = start[2 * (lo + cl) + (cl < z0 / GZ ? 1 : 0)];
                n = start[2 * (lo + ch) + 2] - b;
            }
        }
        int v = n;  // inclusive scan over the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(v, off);
            if (r >= off) v += u;
        }
        if (r < NR) {
            rbeg[r] = b;
            roff[r + 1] = v;
        }
        if (r == 0) roff[0] = 0;
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
        // a zero weight adds exact zeros: a particle with one non-zero channel adds to that channel alone
        const bool re0 = CPLX && wre == 0.0;
        const int ch0 = re0 ? NC : 0;
        const bool both = CPLX && !re0 && wim != 0.0;
        // the scale is a power of two, so (w * w_ijk) * scale == (w * scale) * w_ijk exactly
        const double s0 = (re0 ? wim : wre) * scale, s1 = wim * scale;
        // |w_ijk| <= 1: below 2^51 every contribution rounds with one add of 1.5 * 2^52 (to nearest, ties to
        // even, as __double2ll_rn); a wave takes that path when all its particles allow it
        const bool small = fabs(s0) < 0x1p51 && fabs(s1) < 0x1p51;
        auto stencil = [&](auto fast) {
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
                        if constexpr (decltype(fast)::value) {
                            atomicAdd(&acc[ch0 + c], (unsigned long long)(__double_as_longlong(s0 * wijk + 0x1.8p52) -
                                                                         0x4338000000000000LL));
                            if (both)
                                atomicAdd(&acc[NC + c], (unsigned long long)(__double_as_longlong(s1 * wijk + 0x1.8p52) -
                                                                            0x4338000000000000LL));
                        } else {
                            atomicAdd(&acc[ch0 + c], (unsigned long long)__double2ll_rn(s0 * wijk));
                            if (both) atomicAdd(&acc[NC + c], (unsigned long long)__double2ll_rn(s1 * wijk));
                        }
                    }
                }
            }
        };
        if (__all(small))
            stencil(std::true_type{});
        else
            stencil(std::false_type{});
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

// ---------------------------------------------------------------------------
// Two-level counting sort by key (x, y, z/64), without global atomics.
// Level 1: chunks of S2_C1 particles count their x-slab buckets (XB planes,
// LK = XB * n1 * nzc keys each) in LDS; an exclusive scan of the chunk counts
// (bucket-major) gives every chunk its run in each bucket, and the chunk
// writes its particles' (index, local key) there, ranked by LDS atomics.
// Level 2: chunks of S2_C2 particles of one bucket count their local keys in
// LDS; one exclusive scan over (bucket, key, chunk) gives every particle its
// final place, and the chunk moves positions and weights there.  Border
// particles (stencil leaving the mesh) go to the end of level 1 and form the
// border list.  Within a key the order depends on LDS atomics; the cell sums
// do not (64-bit integer sums).
// ---------------------------------------------------------------------------
constexpr int S2_C1 = 8192;   // particles per level-1 chunk
constexpr int S2_C2 = 16384;  // particles per level-2 chunk
constexpr int S2_LK = 16384;  // most local keys of a bucket (int32 counters in 64 KB of LDS)
constexpr int S2_U = 4;       // particles per thread loaded before any is placed
constexpr int S2_U2 = 2;      // the same for the level-2 placement (more writes in flight there slow it)

struct TwoLevel {
    int XB, nb, LK;
    long long nch1;  // level-1 chunks
    long long nch2;  // level-2 chunks at most: nch1 * S2_C1 / S2_C2 + nb
};

static bool two_level_geom(const AssignGeom& g, long long np, TwoLevel& t) {
    const long long rowkeys = (long long)g.n[1] * keys_row(g);
    if (rowkeys > S2_LK || np <= 0) return false;
    t.XB = (int)std::min<long long>(S2_LK / rowkeys, g.n[0]);
    t.nb = (g.n[0] + t.XB - 1) / t.XB;
    t.LK = (int)(t.XB * rowkeys);
    t.nch1 = (np + S2_C1 - 1) / S2_C1;
    t.nch2 = (np + S2_C2 - 1) / S2_C2 + t.nb;
    return true;
}

template <int S>
__global__ __launch_bounds__(256) void k_s2_hist1(const double* __restrict__ pos, long long np, AssignGeom g,
                                                 TwoLevel tl, int* __restrict__ key, int* __restrict__ h1) {
    extern __shared__ int hist[];  // nb + 1 (the last: border particles)
    for (int e = threadIdx.x; e <= tl.nb; e += 256) hist[e] = 0;
    __syncthreads();
    const long long c = blockIdx.x, q0 = c * S2_C1, q1 = (q0 + S2_C1 < np) ? q0 + S2_C1 : np;
    for (long long q = q0 + threadIdx.x; q < q1; q += 256) {
        long long K;
        const bool in = first_cell<S>(pos, q, g, &K);
        key[q] = in ? (int)K : -1;
        atomicAdd(&hist[in ? (int)(K / tl.LK) : tl.nb], 1);
    }
    __syncthreads();
    for (int e = threadIdx.x; e <= tl.nb; e += 256) h1[(long long)e * tl.nch1 + c] = hist[e];
}

// Level-1 placement: each chunk's particles in runs per bucket (about
// S2_C1 / buckets records each), as one 32-byte (x, y, z, w_re) and one
// 16-byte (index, local key, w_im) record; level 2 then reads them in order.
__global__ __launch_bounds__(256) void k_s2_scatter1(const int* __restrict__ key, long long np, TwoLevel tl,
                                                    const int* __restrict__ off1, const double* __restrict__ pos,
                                                    const double* __restrict__ w, double* __restrict__ tpos,
                                                    int* __restrict__ tw) {
    extern __shared__ int cur[];  // nb + 1
    const long long c = blockIdx.x, q0 = c * S2_C1, q1 = (q0 + S2_C1 < np) ? q0 + S2_C1 : np;
    for (int e = threadIdx.x; e <= tl.nb; e += 256) cur[e] = off1[(long long)e * tl.nch1 + c];
    __syncthreads();
    // S2_U particles per thread are loaded before any is placed: stores count
    // in the same in-order counter as loads, so one particle at a time would
    // wait for the previous particle's scattered stores on every load
    // without weights the loads read pos (3 np doubles, so 2 q + 1 stays inside) and the values are dropped
    const double* wsrc = w ? w : pos;
    for (long long qb = q0 + threadIdx.x; qb < q1; qb += S2_U * 256) {
        int K[S2_U];
        double x[S2_U], y[S2_U], z[S2_U], wr[S2_U], wi[S2_U];
#pragma unroll
        for (int u = 0; u < S2_U; u++) {
            // past the chunk's end: load the last particle again (unused), so no load waits on a branch
            const long long q = (qb + u * 256 < q1) ? qb + u * 256 : q1 - 1;
            K[u] = key[q];
            x[u] = pos[3 * q];
            y[u] = pos[3 * q + 1];
            z[u] = pos[3 * q + 2];
            wr[u] = wsrc[2 * q];
            wi[u] = wsrc[2 * q + 1];
        }
#pragma unroll
        for (int u = 0; u < S2_U; u++) {
            const long long q = qb + u * 256;
            if (q >= q1) break;
            const int b = K[u] < 0 ? tl.nb : K[u] / tl.LK;
            const long long d = atomicAdd(&cur[b], 1);
            // one 32-byte (x, y, z, w_re) and one 16-byte (q, local key, w_im) record per particle
            reinterpret_cast<double4*>(tpos)[d] = make_double4(x[u], y[u], z[u], w ? wr[u] : 0.0);
            const long long wib = __double_as_longlong(w ? wi[u] : 0.0);
            reinterpret_cast<int4*>(tw)[d] = make_int4((int)q, K[u] < 0 ? -1 : K[u] - b * tl.LK,
                                                       (int)(wib & 0xffffffffLL), (int)(wib >> 32));
        }
    }
}

// Level-2 chunk starts per bucket (one workgroup): cst[b] = sum over b' < b of
// ceil(n_b' / S2_C2), cst[nb] = their total.
__global__ __launch_bounds__(1024) void k_s2_chunks(const int* __restrict__ off1, TwoLevel tl, int* __restrict__ cst) {
    __shared__ int sm[1024];
    int carry = 0;
    for (int c0 = 0; c0 < tl.nb; c0 += 1024) {
        const int b = c0 + threadIdx.x;
        int v = 0;
        if (b < tl.nb) {
            const int n = off1[(long long)(b + 1) * tl.nch1] - off1[(long long)b * tl.nch1];
            v = (n + S2_C2 - 1) / S2_C2;
        }
        sm[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int t = (threadIdx.x >= off) ? sm[threadIdx.x - off] : 0;
            __syncthreads();
            sm[threadIdx.x] += t;
            __syncthreads();
        }
        if (b < tl.nb) cst[b] = carry + sm[threadIdx.x] - v;
        carry += sm[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) cst[tl.nb] = carry;
}

// The bucket of level-2 chunk w: the largest b with cst[b] <= w.
__device__ __forceinline__ int s2_bucket(const int* __restrict__ cst, int nb, int w) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cst[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

constexpr int S2_NT2 = 1024;

// Level-2 key counts, layout per bucket [LK][chunks] (key-major) at LK * cst[b];
// workgroups past the last chunk zero the unused tail of h2 instead. The
// 64 KiB histogram allows two workgroups per CU, so each runs S2_NT2 threads.
__global__ __launch_bounds__(S2_NT2) void k_s2_hist2(const int* __restrict__ off1, const int* __restrict__ cst,
                                                 TwoLevel tl, const int* __restrict__ tk, int* __restrict__ h2) {
    // tk: the level-1 16-byte records (index, local key, w_im)
    extern __shared__ int hist[];  // LK
    const int w = blockIdx.x;
    const int total = cst[tl.nb];
    if (w >= total) {
        for (int e = threadIdx.x; e < tl.LK; e += S2_NT2) h2[(long long)w * tl.LK + e] = 0;
        return;
    }
    const int b = s2_bucket(cst, tl.nb, w);
    const int ch = cst[b + 1] - cst[b], c = w - cst[b];
    const long long bs = off1[(long long)b * tl.nch1], be = off1[(long long)(b + 1) * tl.nch1];
    const long long p0 = bs + (long long)c * S2_C2, p1 = (p0 + S2_C2 < be) ? p0 + S2_C2 : be;
    for (int e = threadIdx.x; e < tl.LK; e += S2_NT2) hist[e] = 0;
    __syncthreads();
    for (long long pb = p0 + threadIdx.x; pb < p1; pb += S2_U * S2_NT2) {
        int k[S2_U];
#pragma unroll
        for (int u = 0; u < S2_U; u++) k[u] = tk[4 * ((pb + u * S2_NT2 < p1) ? pb + u * S2_NT2 : p1 - 1) + 1];
#pragma unroll
        for (int u = 0; u < S2_U; u++)
            if (pb + u * S2_NT2 < p1) atomicAdd(&hist[k[u]], 1);
    }
    __syncthreads();
    int* out = h2 + (long long)tl.LK * cst[b] + c;
    for (int e = threadIdx.x; e < tl.LK; e += S2_NT2) out[(long long)e * ch] = hist[e];
}

// start[K] of every key (the first place of key K; n_interior at K = nk).
__global__ __launch_bounds__(256) void k_s2_start(const int* __restrict__ off1, const int* __restrict__ cst,
                                                 const int* __restrict__ off2, TwoLevel tl, long long nk,
                                                 int* __restrict__ start) {
    const long long K = (long long)blockIdx.x * 256 + threadIdx.x;
    if (K > nk) return;
    if (K == nk) {
        start[nk] = off1[(long long)tl.nb * tl.nch1];
        return;
    }
    const int b = (int)(K / tl.LK), lk = (int)(K - (long long)b * tl.LK);
    const int ch = cst[b + 1] - cst[b];
    start[K] = ch > 0 ? off2[(long long)tl.LK * cst[b] + (long long)lk * ch] : off1[(long long)b * tl.nch1];
}

// Level-2 placement: positions and weights (read in bucket order) to their
// places in key order, all within the bucket's range; slot[p] = place of the
// particle at level-1 position p (a later move of weights reads the particle
// index from the level-1 record).
__global__ __launch_bounds__(S2_NT2) void k_s2_scatter2(const int* __restrict__ off1, const int* __restrict__ cst,
                                                    const int* __restrict__ off2, TwoLevel tl,
                                                    const int* __restrict__ tk, const double* __restrict__ pos,
                                                    bool weights, int* __restrict__ slot, double* __restrict__ spos,
                                                    double* __restrict__ sw) {
    extern __shared__ int cur[];  // LK
    const int wg = blockIdx.x;
    if (wg >= cst[tl.nb]) return;
    const int b = s2_bucket(cst, tl.nb, wg);
    const int ch = cst[b + 1] - cst[b], c = wg - cst[b];
    const long long bs = off1[(long long)b * tl.nch1], be = off1[(long long)(b + 1) * tl.nch1];
    const long long p0 = bs + (long long)c * S2_C2, p1 = (p0 + S2_C2 < be) ? p0 + S2_C2 : be;
    const int* o = off2 + (long long)tl.LK * cst[b] + c;
    for (int e = threadIdx.x; e < tl.LK; e += S2_NT2) cur[e] = o[(long long)e * ch];
    __syncthreads();
    for (long long pb = p0 + threadIdx.x; pb < p1; pb += S2_U2 * S2_NT2) {
        int4 m[S2_U2];
        double4 a[S2_U2];
#pragma unroll
        for (int u = 0; u < S2_U2; u++) {
            const long long p = (pb + u * S2_NT2 < p1) ? pb + u * S2_NT2 : p1 - 1;  // the tail reloads the last
            m[u] = reinterpret_cast<const int4*>(tk)[p];
            a[u] = reinterpret_cast<const double4*>(pos)[p];
        }
#pragma unroll
        for (int u = 0; u < S2_U2; u++) {
            const long long p = pb + u * S2_NT2;
            if (p >= p1) break;
            const long long d = atomicAdd(&cur[m[u].y], 1);
            slot[p] = (int)d;
            spos[3 * d] = a[u].x;
            spos[3 * d + 1] = a[u].y;
            spos[3 * d + 2] = a[u].z;
            if (weights) {
                sw[2 * d] = a[u].w;
                sw[2 * d + 1] = __longlong_as_double(((long long)m[u].w << 32) | (unsigned)m[u].z);
            }
        }
    }
}

// The border particles (the tail of level 1) to the border list; border[0]
// their count, and the interior count (the level-2 positions in use) after
// the list at border[np + 1].
__global__ __launch_bounds__(256) void k_s2_border(const int* __restrict__ off1, TwoLevel tl, long long np,
                                                  const int* __restrict__ tq, int* __restrict__ border) {
    const long long nint = off1[(long long)tl.nb * tl.nch1];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        border[0] = (int)(np - nint);
        border[np + 1] = (int)nint;
    }
    if (i < np - nint) border[1 + i] = tq[4 * (nint + i)];
}

// A later assignment of the same positions after a two-level sort: weights to
// the places of their particles (the reads follow level-1 order).
__global__ __launch_bounds__(256) void k_s2_move_w(const double* __restrict__ w, long long np,
                                                  const int* __restrict__ tq, const int* __restrict__ slot,
                                                  const int* __restrict__ border, double* __restrict__ sw) {
    const long long nint = border[np + 1];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nint) return;
    const long long q = tq[4 * p], d = slot[p];
    sw[2 * d] = w[2 * q];
    sw[2 * d + 1] = w[2 * q + 1];
}

// The same move with the weights read in order: one workgroup per level-1
// chunk stages the chunk's S2_C1 complex weights in LDS (128 KiB), then each
// wave walks the chunk's run in some of the buckets (level-1 order, about
// S2_C1 / buckets records each) and writes every weight to its particle's
// place. The runs hold only the chunk's own particles, so q - q0 stays in
// [0, S2_C1).
constexpr int S2_MW_NT = 1024;
__global__ __launch_bounds__(S2_MW_NT) void k_s2_move_wc(const double* __restrict__ w, long long np, TwoLevel tl,
                                                         const int* __restrict__ off1, const int* __restrict__ tq,
                                                         const int* __restrict__ slot, double* __restrict__ sw) {
    extern __shared__ double2 lw[];  // S2_C1
    const long long c = blockIdx.x, q0 = c * S2_C1, q1 = (q0 + S2_C1 < np) ? q0 + S2_C1 : np;
    for (long long q = q0 + threadIdx.x; q < q1; q += S2_MW_NT) lw[q - q0] = make_double2(w[2 * q], w[2 * q + 1]);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double2* sw2 = reinterpret_cast<double2*>(sw);
    for (int b = wave; b < tl.nb; b += S2_MW_NT / 64) {
        const long long r0 = off1[(long long)b * tl.nch1 + c], r1 = off1[(long long)b * tl.nch1 + c + 1];
        for (long long j = r0 + lane; j < r1; j += 64) sw2[slot[j]] = lw[tq[4 * j] - q0];
    }
}

static int scan_inplace(hm_ctx* ctx, const char* name, int* a, long long n, int* bsum) {
    const unsigned nb = (unsigned)((n + SCAN_BLK - 1) / SCAN_BLK);
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_NT), 0, ctx->stream, a, n, bsum);
        hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(1024), 0, ctx->stream, bsum, (int)nb);
        hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_NT), 0, ctx->stream, a, n, bsum);
    });
}

template <int S>
static int sort_two_level(hm_ctx* ctx, const AssignGeom& g, const TwoLevel& tl, const double* d_pos,
                          const double* d_w, long long np, SortWork& w) {
    int* key = (int*)w.key.p;
    int* h1 = (int*)w.h1.p;
    int* h2 = (int*)w.h2.p;
    int* cst = (int*)w.cst.p;
    int* bsum = (int*)w.bsum.p;
    const size_t lds1 = sizeof(int) * (size_t)(tl.nb + 1), lds2 = sizeof(int) * (size_t)tl.LK;
    const long long nk = n_keys(g);
    if (ctx->launch("assign_count", [&] {
            hipLaunchKernelGGL(k_s2_hist1<S>, dim3((unsigned)tl.nch1), dim3(256), lds1, ctx->stream, d_pos, np, g, tl,
                               key, h1);
        }))
        return -1;
    if (scan_inplace(ctx, "assign_scan", h1, (long long)(tl.nb + 1) * tl.nch1, bsum)) return -1;
    if (ctx->launch("assign_bucket", [&] {
            hipLaunchKernelGGL(k_s2_scatter1, dim3((unsigned)tl.nch1), dim3(256), lds1, ctx->stream, key, np, tl, h1,
                               d_pos, d_w, (double*)w.tpos.p, (int*)w.tw.p);
            hipLaunchKernelGGL(k_s2_chunks, dim3(1), dim3(1024), 0, ctx->stream, h1, tl, cst);
            hipLaunchKernelGGL(k_s2_hist2, dim3((unsigned)tl.nch2), dim3(S2_NT2), lds2, ctx->stream, h1, cst, tl,
                               (const int*)w.tw.p, h2);
        }))
        return -1;
    if (scan_inplace(ctx, "assign_scan", h2, (long long)tl.LK * tl.nch2, bsum)) return -1;
    return ctx->launch("assign_scatter", [&] {
        hipLaunchKernelGGL(k_s2_start, dim3((unsigned)((nk + 1 + 255) / 256)), dim3(256), 0, ctx->stream, h1, cst, h2,
                           tl, nk, (int*)w.start.p);
        hipLaunchKernelGGL(k_s2_scatter2, dim3((unsigned)tl.nch2), dim3(S2_NT2), lds2, ctx->stream, h1, cst, h2, tl,
                           (const int*)w.tw.p, (const double*)w.tpos.p, d_w != nullptr, (int*)w.slot.p,
                           (double*)w.spos.p, (double*)w.sw.p);
        hipLaunchKernelGGL(k_s2_border, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, h1, tl, np,
                           (const int*)w.tw.p, (int*)w.border.p);
    });
}

int sort_work_alloc(hm_ctx* ctx, SortWork& w, const AssignGeom& g, long long np, bool weights) {
    w.sig.valid = false;  // the buffers may move: the next assignment sorts afresh
    const long long nk = n_keys(g);
    long long nscan = nk + 1;
    TwoLevel tl;
    const bool two = two_level_geom(g, np, tl);
    if (two) nscan = std::max<long long>({nscan, (long long)(tl.nb + 1) * tl.nch1, (long long)tl.LK * tl.nch2});
    const long long nb = (nscan + SCAN_BLK - 1) / SCAN_BLK;
    if (dev_alloc(ctx, w.start, sizeof(int) * (size_t)(nk + 1))) return -1;
    if (dev_alloc(ctx, w.slot, sizeof(int) * (size_t)(np > 0 ? np : 1))) return -1;
    if (dev_alloc(ctx, w.spos, sizeof(double) * 3 * (size_t)(np > 0 ? np : 1))) return -1;
    if (weights && dev_alloc(ctx, w.sw, sizeof(double) * 2 * (size_t)(np > 0 ? np : 1))) return -1;
    if (dev_alloc(ctx, w.border, sizeof(int) * (size_t)(np + 2))) return -1;
    if (dev_alloc(ctx, w.bsum, sizeof(int) * (size_t)nb)) return -1;
    if (two) {
        if (dev_alloc(ctx, w.key, sizeof(int) * (size_t)np)) return -1;
        if (dev_alloc(ctx, w.tpos, 32 * (size_t)np)) return -1;
        if (dev_alloc(ctx, w.tw, 16 * (size_t)np)) return -1;
        if (dev_alloc(ctx, w.h1, sizeof(int) * (size_t)(tl.nb + 1) * (size_t)tl.nch1)) return -1;
        if (dev_alloc(ctx, w.h2, sizeof(int) * (size_t)tl.LK * (size_t)tl.nch2)) return -1;
        if (dev_alloc(ctx, w.cst, sizeof(int) * (size_t)(tl.nb + 1))) return -1;
    }
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
    dev_free(w.key);
    dev_free(w.tpos);
    dev_free(w.tw);
    dev_free(w.h1);
    dev_free(w.h2);
    dev_free(w.cst);
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
    TwoLevel tl;
    const bool two = !reuse && two_level_geom(g, np, tl) && w.key.p && w.tpos.bytes >= 32 * (size_t)np &&
                     w.tw.bytes >= 16 * (size_t)np && w.h2.bytes >= sizeof(int) * (size_t)tl.LK * tl.nch2 &&
                     w.h1.bytes >= sizeof(int) * (size_t)(tl.nb + 1) * tl.nch1 &&
                     w.bsum.bytes >= sizeof(int) * (size_t)(((long long)tl.LK * tl.nch2 + SCAN_BLK - 1) / SCAN_BLK);
    if (two) {
        if (sort_two_level<S>(ctx, g, tl, d_pos, d_w, np, w)) return -1;
        w.two_level = true;
    } else if (!reuse) {
        w.two_level = false;
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
    if (np > 0 && !reuse && !two && ctx->launch("assign_scatter", [&] {
            hipLaunchKernelGGL(k_sort_scatter<S>, dim3(pb), dim3(256), 0, ctx->stream, d_pos, d_w, np, g, start, slot,
                               spos, sw);
        }))
        return -1;
    TwoLevel tr;
    const bool chunk_move = reuse && w.two_level && two_level_geom(g, np, tr) &&
                            ctx->lds_max >= sizeof(double2) * (size_t)S2_C1;
    if (np > 0 && reuse && sw && ctx->launch("assign_scatter", [&] {
            if (chunk_move)
                hipLaunchKernelGGL(k_s2_move_wc, dim3((unsigned)tr.nch1), dim3(S2_MW_NT), sizeof(double2) * S2_C1,
                                   ctx->stream, d_w, np, tr, (const int*)w.h1.p, (const int*)w.tw.p, (const int*)slot,
                                   sw);
            else if (w.two_level)
                hipLaunchKernelGGL(k_s2_move_w, dim3(pb), dim3(256), 0, ctx->stream, d_w, np, (const int*)w.tw.p,
                                   (const int*)slot, (const int*)border, sw);
            else
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
    int sexp = 0;
    if (!(scale > 0.0) || std::frexp(scale, &sexp) != 0.5) {  // the gather premultiplies weights by the scale
        ctx->set_error("assign: the fixed-point scale must be a power of two");
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
