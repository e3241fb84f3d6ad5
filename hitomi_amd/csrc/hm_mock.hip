// hm_mock.hip -- seeded Zel'dovich mocks generated on the device.
//
// The reference has no mock generator (its only synthetic input is an
// unseeded uniform random catalogue, struct/measurement/nbody/pk/
// calc_random.py:38-49); configurations with clustered catalogues need
// reproducible ones of a stated shape, generated where they are measured.
//
//   noise(x)   standard normal per lattice cell: Philox4x32-10 (Salmon, Moraes,
//              Dror and Shaw, "Parallel random numbers: as easy as 1, 2, 3",
//              SC 2011) with counter (cell, 0, 0, 0) and key (seed), two 53-bit
//              uniforms, Box-Muller in FP64
//   delta(k)   FFT(noise) sqrt(P(k) / V_cell); modes on a Nyquist plane are zero
//   Psi_a(k)   i k_a / k^2 delta(k) (times 1 + f along z), inverse FFT / n^3
//   particles  n lattice sites chosen by a seeded permutation of the n_side^3
//              sites (a four-round Feistel network on the next even power of
//              two, cycle-walked into range), q + Psi(q), wrapped into [0, L)
#include "hm_internal.hpp"

#include <cstdint>

using namespace hm;

namespace {

int bad(hm_ctx* ctx, const std::string& m) {
    if (ctx) ctx->set_error(m);
    return -1;
}

constexpr uint32_t PH_M0 = 0xD2511F53u, PH_M1 = 0xCD9E8D57u;
constexpr uint32_t PH_W0 = 0x9E3779B9u, PH_W1 = 0xBB67AE85u;

__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)PH_M0 * c[0], p1 = (uint64_t)PH_M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += PH_W0;
        k1 += PH_W1;
    }
}

// 53-bit uniform from two words: (a >> 5) * 2^26 + (b >> 6), times 2^-53
__device__ inline double u53(uint32_t a, uint32_t b) {
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * 0x1.0p-53;
}

__device__ inline double std_normal(uint64_t cell, uint64_t seed) {
#pragma clang fp contract(off)
    uint32_t c[4] = {(uint32_t)cell, (uint32_t)(cell >> 32), 0u, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u1 = u53(c[0], c[1]) + 0x1.0p-54;  // (0, 1)
    const double u2 = u53(c[2], c[3]);              // [0, 1)
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ __launch_bounds__(256) void k_white_noise(long long n, unsigned long long seed, double* __restrict__ out) {
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st)
        out[i] = std_normal((uint64_t)i, seed);
}

struct PkArgs {
    int n;
    double L, amp, k0, slope, kcut, bao_amp, bao_period, fz, inv_vcell;
};

__device__ inline double pk_model(const PkArgs& a, double k) {
#pragma clang fp contract(off)
    if (!(k > 0.0)) return 0.0;
    const double r = k / a.kcut;
    double p = a.amp * pow(k / a.k0, a.slope) * exp(-(r * r));
    if (a.bao_amp != 0.0) p = p * (1.0 + a.bao_amp * sin(6.283185307179586 * k / a.bao_period));
    return p;
}

// Half spectrum of the noise -> the three displacement spectra.
__global__ __launch_bounds__(256) void k_zeldovich_spec(C<double>* __restrict__ w, PkArgs a, int keep_delta,
                                                        C<double>* __restrict__ px, C<double>* __restrict__ py,
                                                        C<double>* __restrict__ pz) {
#pragma clang fp contract(off)
    const int n = a.n, nzh = n / 2 + 1;
    const long long total = (long long)n * n * nzh;
    const long long st = (long long)gridDim.x * blockDim.x;
    const double kf = 6.283185307179586 / a.L;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += st) {
        const int kz = (int)(e % nzh);
        const long long r = e / nzh;
        const int j = (int)(r % n), i = (int)(r / n);
        C<double> zx = mk<double>(0.0, 0.0), zy = zx, zz = zx, dd = zx;
        if (2 * i != n && 2 * j != n && 2 * kz != n && (i | j | kz) != 0) {
            const double kx = kf * (double)(i < n / 2 ? i : i - n);
            const double ky = kf * (double)(j < n / 2 ? j : j - n);
            const double kzv = kf * (double)kz;
            const double k2 = (kx * kx + ky * ky) + kzv * kzv;
            const double amp = sqrt(pk_model(a, sqrt(k2)) * a.inv_vcell);
            const C<double> d = w[e];
            // i k_a / k^2 (amp (dr + i di)) = (k_a / k^2) amp (-di + i dr)
            const double gr = -(amp * d.y) / k2, gi = (amp * d.x) / k2;
            zx = mk<double>(kx * gr, kx * gi);
            zy = mk<double>(ky * gr, ky * gi);
            zz = mk<double>((kzv * a.fz) * gr, (kzv * a.fz) * gi);
            dd = mk<double>(amp * d.x, amp * d.y);
        }
        if (keep_delta) w[e] = dd;
        px[e] = zx;
        py[e] = zy;
        pz[e] = zz;
    }
}

__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Perm {
    uint64_t sites;  // n_side^3
    int h;           // half width in bits: 2^(2h) >= sites
    uint64_t key[4];
};

__device__ inline uint64_t feistel(const Perm& p, uint64_t x) {
    const uint64_t mask = (1ull << p.h) - 1ull;
    uint64_t l = x >> p.h, r = x & mask;
    for (int s = 0; s < 4; s++) {
        const uint64_t t = l ^ (mix64(r ^ p.key[s]) & mask);
        l = r;
        r = t;
    }
    return (l << p.h) | r;
}

// i-th chosen site: the permutation, cycle-walked until it lands below sites
// (it does: the orbit of i < sites under a permutation returns to i)
__device__ inline uint64_t site_of(const Perm& p, uint64_t i) {
    uint64_t y = feistel(p, i);
    while (y >= p.sites) y = feistel(p, y);
    return y;
}

__global__ __launch_bounds__(256) void k_place(long long np, int n, double L, Perm pm, const double* __restrict__ sx,
                                               const double* __restrict__ sy, const double* __restrict__ sz,
                                               double inv_n3, double* __restrict__ pos) {
#pragma clang fp contract(off)
    const long long st = (long long)gridDim.x * blockDim.x;
    const double cell = L / (double)n;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += st) {
        const uint64_t s = pm.sites == (uint64_t)np ? (uint64_t)i : site_of(pm, (uint64_t)i);
        const int iz = (int)(s % (uint64_t)n), iy = (int)((s / (uint64_t)n) % (uint64_t)n),
                  ix = (int)(s / ((uint64_t)n * (uint64_t)n));
        const double q[3] = {((double)ix + 0.5) * cell, ((double)iy + 0.5) * cell, ((double)iz + 0.5) * cell};
        const double d[3] = {sx[s] * inv_n3, sy[s] * inv_n3, sz[s] * inv_n3};
        for (int a = 0; a < 3; a++) {
            const double x = q[a] + d[a];
            double v = x - L * floor(x / L);
            if (v >= L) v -= L;
            pos[3 * i + a] = v;
        }
    }
}

// Lagrangian-biased selection: site s is kept when its uniform (Philox
// counter (s, 1, 0, 0)) falls below p = rate (1 + b delta_L(s)), clipped to
// [0, 1].  Kept sites are numbered in site order: per-block counts, one
// exclusive scan of them, then each block writes its kept sites.
constexpr int SEL_B = 1024;  // sites per block (256 threads x 4)

// per-block sums of max(0, 1 + b delta_L) over SEL_B sites (fixed order)
__global__ __launch_bounds__(256) void k_sel_wsum(long long total, const double* __restrict__ dl, double inv_n3, double b,
                                                  double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double sm[256];
    const long long base = (long long)blockIdx.x * SEL_B;
    double acc = 0.0;
    for (int j = 0; j < 4; j++) {
        const long long s = base + 4 * threadIdx.x + j;
        if (s < total) acc += fmax(0.0, 1.0 + b * (dl[s] * inv_n3));
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

__device__ inline bool keep_site(uint64_t s, uint64_t seed, const double* dl, double inv_n3, double rate, double b) {
#pragma clang fp contract(off)
    uint32_t c[4] = {(uint32_t)s, (uint32_t)(s >> 32), 1u, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u = u53(c[0], c[1]);
    const double p = rate * fmax(0.0, 1.0 + b * (dl[s] * inv_n3));
    return u < p;
}

__global__ __launch_bounds__(256) void k_sel_count(long long total, unsigned long long seed, const double* __restrict__ dl,
                                                   double inv_n3, double rate, double b, long long* __restrict__ cnt) {
    __shared__ int sm[256];
    const long long base = (long long)blockIdx.x * SEL_B;
    int c = 0;
    for (int j = 0; j < 4; j++) {
        const long long s = base + 4 * threadIdx.x + j;
        if (s < total && keep_site((uint64_t)s, seed, dl, inv_n3, rate, b)) c++;
    }
    sm[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[blockIdx.x] = sm[0];
}

// exclusive scan of nb block counts in one workgroup (fixed order); off[nb] = total
__global__ __launch_bounds__(256) void k_sel_scan(const long long* __restrict__ cnt, long long nb, long long* __restrict__ off) {
    __shared__ long long sm[256];
    const long long per = (nb + 255) / 256;
    const long long lo = (long long)threadIdx.x * per, hi = (lo + per < nb) ? lo + per : nb;
    long long s = 0;
    for (long long i = lo; i < hi; i++) s += cnt[i];
    sm[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long run = 0;
        for (int t = 0; t < 256; t++) {
            const long long v = sm[t];
            sm[t] = run;
            run += v;
        }
        off[nb] = run;
    }
    __syncthreads();
    long long run = sm[threadIdx.x];
    for (long long i = lo; i < hi; i++) {
        off[i] = run;
        run += cnt[i];
    }
}

__global__ __launch_bounds__(256) void k_sel_place(long long total, int n, double L, unsigned long long seed,
                                                   const double* __restrict__ dl, double inv_n3, double rate, double b,
                                                   const long long* __restrict__ off, long long capacity,
                                                   const double* __restrict__ sx, const double* __restrict__ sy,
                                                   const double* __restrict__ sz, double* __restrict__ pos,
                                                   double* __restrict__ shift) {
#pragma clang fp contract(off)
    __shared__ int sm[256];
    const long long base = (long long)blockIdx.x * SEL_B;
    int f[4], c = 0;
    for (int j = 0; j < 4; j++) {
        const long long s = base + 4 * threadIdx.x + j;
        f[j] = (s < total && keep_site((uint64_t)s, seed, dl, inv_n3, rate, b)) ? 1 : 0;
        c += f[j];
    }
    // exclusive prefix of c over the workgroup (Hillis-Steele in LDS)
    sm[threadIdx.x] = c;
    __syncthreads();
    for (int w = 1; w < 256; w <<= 1) {
        const int v = (int)threadIdx.x >= w ? sm[threadIdx.x - w] : 0;
        __syncthreads();
        sm[threadIdx.x] += v;
        __syncthreads();
    }
    long long idx = off[blockIdx.x] + sm[threadIdx.x] - c;
    const double cell = L / (double)n;
    for (int j = 0; j < 4; j++) {
        if (!f[j]) continue;
        const long long s = base + 4 * threadIdx.x + j;
        if (idx < capacity) {
            const int iz = (int)(s % n), iy = (int)((s / n) % n), ix = (int)(s / ((long long)n * n));
            const double q[3] = {((double)ix + 0.5) * cell, ((double)iy + 0.5) * cell, ((double)iz + 0.5) * cell};
            const double d[3] = {sx[s] * inv_n3, sy[s] * inv_n3, sz[s] * inv_n3};
            for (int a = 0; a < 3; a++) {
                const double x = q[a] + d[a];
                double v = x - L * floor(x / L);
                if (v >= L) v -= L;
                pos[3 * idx + a] = v;
                if (shift) shift[3 * idx + a] = -d[a];
            }
        }
        idx++;
    }
}

uint64_t splitmix_next(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Points uniform in volume inside a sky patch: point i takes two Philox
// blocks, counters (i lo, i hi, 3, 0) and (i lo, i hi, 4, 0), key (seed):
// u1, u2 from the first, u3 from the second; r^3 uniform in [rmin^3, rmax^3],
// RA uniform, sin(Dec) uniform; Cartesian as fits2dat_galaxy_weight.py:130-132.
__global__ __launch_bounds__(256) void k_sky_randoms(long long n, unsigned long long seed, hm_sky_patch g,
                                                    double* __restrict__ out) {
#pragma clang fp contract(off)
    const long long st = (long long)gridDim.x * blockDim.x;
    const double deg = 3.14159265358979323846 / 180.0;
    const double r3a = g.rmin * g.rmin * g.rmin, r3b = g.rmax * g.rmax * g.rmax;
    const double sa = sin(g.dec_min * deg), sb = sin(g.dec_max * deg);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
        const uint64_t u = (uint64_t)i;
        uint32_t c[4] = {(uint32_t)u, (uint32_t)(u >> 32), 3u, 0u};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        uint32_t e[4] = {(uint32_t)u, (uint32_t)(u >> 32), 4u, 0u};
        philox4x32_10(e, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u1 = u53(c[0], c[1]), u2 = u53(c[2], c[3]), u3 = u53(e[0], e[1]);
        const double r = cbrt(r3a + u1 * (r3b - r3a));
        const double ra = (g.ra_min + u2 * (g.ra_max - g.ra_min)) * deg;
        const double sd = sa + u3 * (sb - sa);
        const double cd = sqrt(1.0 - sd * sd);
        out[3 * i] = r * cd * cos(ra);
        out[3 * i + 1] = r * cd * sin(ra);
        out[3 * i + 2] = r * sd;
    }
}

int check_cfg(hm_ctx* ctx, const hm_mock_config* c) {
    if (!c) return bad(ctx, "mock: no configuration");
    if (c->n_side < 16 || c->n_side > 1024 || (c->n_side & (c->n_side - 1)))
        return bad(ctx, "mock: n_side must be a power of two in [16, 1024]");
    if (!(c->boxsize > 0.0) || !(c->k0 > 0.0) || !(c->kcut > 0.0) || c->amp < 0.0)
        return bad(ctx, "mock: boxsize, k0 and kcut must be positive and amp non-negative");
    if (c->bao_amp != 0.0 && !(c->bao_period > 0.0)) return bad(ctx, "mock: bao_period must be positive");
    return 0;
}

int launch_noise(hm_ctx* ctx, long long total, unsigned long long seed, double* out) {
    const long long blocks = std::min((total + 255) / 256, 65536LL);
    return ctx->launch("mock_noise", [&] {
        hipLaunchKernelGGL(k_white_noise, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, total, seed, out);
    });
}

}  // namespace

extern "C" {

int hm_mock_white_noise(hm_ctx* ctx, const hm_mock_config* cfg, unsigned long long seed, double* d_out) {
    if (!ctx) return -1;
    if (check_cfg(ctx, cfg)) return -1;
    if (!d_out) return bad(ctx, "mock: no output");
    const long long n = cfg->n_side, total = n * n * n;
    if (launch_noise(ctx, total, seed, d_out)) return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"

namespace {

// Noise -> spectrum -> the three displacement grids (x into g[0], y into g[1],
// z into g[2], each n^3 doubles, unnormalised) and, with delta, the linear
// density grid into g[3].  Workspace: noise | spectrum | 3 spectra | 1 grid.
int displacement_grids(hm_ctx* ctx, const hm_mock_config* cfg, unsigned long long seed, bool delta, double* g[4]) {
    const long long n = cfg->n_side, total = n * n * n, half = n * n * (n / 2 + 1);
    const size_t need = sizeof(double) * (size_t)total * 2 + sizeof(C<double>) * (size_t)half * 4;
    if (ctx->mock_ws_bytes < need) {
        HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->mock_ws) (void)hipFree(ctx->mock_ws);
        ctx->mock_ws = nullptr;
        ctx->mock_ws_bytes = 0;
        if (hipMalloc(&ctx->mock_ws, need) != hipSuccess) {
            ctx->mock_ws = nullptr;
            return bad(ctx, "mock: out of device memory");
        }
        ctx->mock_ws_bytes = need;
    }
    char* ws = (char*)ctx->mock_ws;
    double* noise = (double*)ws;
    C<double>* spec = (C<double>*)(ws + sizeof(double) * (size_t)total);
    C<double>* px = spec + half;
    C<double>* py = px + half;
    C<double>* pz = py + half;
    double* extra = (double*)(pz + half);
    const int dims[3] = {(int)n, (int)n, (int)n};
    if (launch_noise(ctx, total, seed, noise)) return -1;
    if (fft3d_r2c<double>(ctx, noise, spec, dims)) return -1;
    PkArgs a;
    a.n = (int)n;
    a.L = cfg->boxsize;
    a.amp = cfg->amp;
    a.k0 = cfg->k0;
    a.slope = cfg->slope;
    a.kcut = cfg->kcut;
    a.bao_amp = cfg->bao_amp;
    a.bao_period = cfg->bao_period;
    a.fz = 1.0 + cfg->f_growth;
    const double cell = cfg->boxsize / (double)n;
    a.inv_vcell = 1.0 / ((cell * cell) * cell);
    if (ctx->launch("mock_spec", [&] {
            const long long blocks = std::min((half + 255) / 256, 65536LL);
            hipLaunchKernelGGL(k_zeldovich_spec, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, spec, a,
                               delta ? 1 : 0, px, py, pz);
        }))
        return -1;
    // a half spectrum (n^2 (n + 2) doubles) holds one real grid once transformed
    g[0] = noise;
    g[1] = (double*)px;
    g[2] = extra;
    g[3] = delta ? (double*)py : nullptr;
    if (fft3d_c2r<double>(ctx, px, g[0], dims)) return -1;
    if (fft3d_c2r<double>(ctx, py, g[1], dims)) return -1;
    if (fft3d_c2r<double>(ctx, pz, g[2], dims)) return -1;
    if (delta && fft3d_c2r<double>(ctx, spec, g[3], dims)) return -1;
    return 0;
}

}  // namespace

extern "C" {

int hm_mock_zeldovich(hm_ctx* ctx, const hm_mock_config* cfg, unsigned long long seed, long long np, double* d_pos) {
    if (!ctx) return -1;
    if (check_cfg(ctx, cfg)) return -1;
    const long long n = cfg->n_side, total = n * n * n;
    if (np <= 0 || np > total) return bad(ctx, "mock: need 0 < n <= n_side^3 particles");
    if (!d_pos) return bad(ctx, "mock: no output");
    double* g[4];
    if (displacement_grids(ctx, cfg, seed, false, g)) return -1;
    Perm pm;
    pm.sites = (uint64_t)total;
    int bits = 0;
    while ((1ull << bits) < (uint64_t)total) bits++;
    pm.h = (bits + 1) / 2;
    uint64_t sm = (uint64_t)seed ^ 0x5DEECE66Dull;
    for (int s = 0; s < 4; s++) pm.key[s] = splitmix_next(sm);
    const double inv_n3 = 1.0 / (double)total;
    return ctx->launch("mock_place", [&] {
        const long long blocks = std::min((np + 255) / 256, 65536LL);
        hipLaunchKernelGGL(k_place, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, np, (int)n, cfg->boxsize, pm,
                           g[0], g[1], g[2], inv_n3, d_pos);
    });
}

int hm_mock_sky_randoms(hm_ctx* ctx, const hm_sky_patch* patch, unsigned long long seed, long long n, double* d_pos) {
    if (!ctx) return -1;
    if (!patch || !(patch->rmin >= 0.0) || !(patch->rmax > patch->rmin) || !(patch->ra_max > patch->ra_min) ||
        !(patch->dec_max > patch->dec_min) || patch->dec_min < -90.0 || patch->dec_max > 90.0)
        return bad(ctx, "mock: the sky patch needs 0 <= rmin < rmax, ra_min < ra_max, -90 <= dec_min < dec_max <= 90");
    if (n < 0 || (n > 0 && !d_pos)) return bad(ctx, "mock: bad arguments");
    if (n == 0) return 0;
    return ctx->launch("mock_sky_randoms", [&] {
        const long long blocks = std::min((n + 255) / 256, 65536LL);
        hipLaunchKernelGGL(k_sky_randoms, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, n, seed, *patch, d_pos);
    });
}

int hm_mock_zeldovich_biased(hm_ctx* ctx, const hm_mock_config* cfg, unsigned long long seed, double mean_n,
                             double b_lag, long long capacity, double* d_pos, double* d_shift, long long* n_out) {
    if (!ctx) return -1;
    if (check_cfg(ctx, cfg)) return -1;
    const long long n = cfg->n_side, total = n * n * n;
    if (!(mean_n > 0.0) || mean_n > (double)total) return bad(ctx, "mock: need 0 < mean_n <= n_side^3");
    if (!d_pos || !n_out || capacity <= 0) return bad(ctx, "mock: no output");
    double* g[4];
    if (displacement_grids(ctx, cfg, seed, true, g)) return -1;
    const long long nb = (total + SEL_B - 1) / SEL_B;
    DevBuf cnt;
    if (dev_alloc(ctx, cnt, sizeof(long long) * (size_t)(2 * nb + 1))) return -1;
    long long* c = (long long*)cnt.p;
    long long* off = c + nb;
    const double inv_n3 = 1.0 / (double)total;
    // rate = mean_n / sum_s max(0, 1 + b delta_L(s)): the expected count is mean_n
    std::vector<double> ws((size_t)nb);
    int rc = ctx->launch("mock_sel_wsum", [&] {
        hipLaunchKernelGGL(k_sel_wsum, dim3((unsigned)nb), dim3(256), 0, ctx->stream, total, (const double*)g[3], inv_n3,
                           b_lag, (double*)c);
    });
    if (!rc && (hipMemcpyAsync(ws.data(), c, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, ctx->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess))
        rc = bad(ctx, "mock: stream error");
    double wsum = 0.0;
    for (double v : ws) wsum += v;
    if (!rc && !(wsum > 0.0)) rc = bad(ctx, "mock: the bias leaves no site with positive weight");
    const double rate = rc ? 0.0 : mean_n / wsum;
    if (!rc)
        rc = ctx->launch("mock_sel_count", [&] {
            hipLaunchKernelGGL(k_sel_count, dim3((unsigned)nb), dim3(256), 0, ctx->stream, total, seed,
                               (const double*)g[3], inv_n3, rate, b_lag, c);
        });
    if (!rc)
        rc = ctx->launch("mock_sel_scan", [&] {
            hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(256), 0, ctx->stream, (const long long*)c, nb, off);
        });
    if (!rc)
        rc = ctx->launch("mock_sel_place", [&] {
            hipLaunchKernelGGL(k_sel_place, dim3((unsigned)nb), dim3(256), 0, ctx->stream, total, (int)n, cfg->boxsize,
                               seed, (const double*)g[3], inv_n3, rate, b_lag, (const long long*)off, capacity, g[0],
                               g[1], g[2], d_pos, d_shift);
        });
    long long count = 0;
    if (!rc && (hipMemcpyAsync(&count, off + nb, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess))
        rc = bad(ctx, "mock: stream error");
    dev_free(cnt);
    if (rc) return rc;
    *n_out = count;
    if (count > capacity) return bad(ctx, "mock: more particles selected than capacity (" + std::to_string(count) + ")");
    return 0;
}

}  // extern "C"
