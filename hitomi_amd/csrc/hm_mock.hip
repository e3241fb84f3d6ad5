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
__global__ __launch_bounds__(256) void k_zeldovich_spec(const C<double>* __restrict__ w, PkArgs a,
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
        C<double> zx = mk<double>(0.0, 0.0), zy = zx, zz = zx;
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
        }
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

uint64_t splitmix_next(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
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

int hm_mock_zeldovich(hm_ctx* ctx, const hm_mock_config* cfg, unsigned long long seed, long long np, double* d_pos) {
    if (!ctx) return -1;
    if (check_cfg(ctx, cfg)) return -1;
    const long long n = cfg->n_side, total = n * n * n, half = n * n * (n / 2 + 1);
    if (np <= 0 || np > total) return bad(ctx, "mock: need 0 < n <= n_side^3 particles");
    if (!d_pos) return bad(ctx, "mock: no output");
    // workspace (kept on the context): noise / x displacement, its spectrum / y displacement, three
    // displacement spectra, z displacement
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
    C<double>* s3 = spec + half;
    double* extra = (double*)(s3 + 3 * half);
    int rc = 0;
    const int dims[3] = {(int)n, (int)n, (int)n};
    C<double>* px = s3;
    C<double>* py = px + half;
    C<double>* pz = py + half;
    if (!rc) rc = launch_noise(ctx, total, seed, noise);
    if (!rc) rc = fft3d_r2c<double>(ctx, noise, spec, dims);
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
    if (!rc)
        rc = ctx->launch("mock_spec", [&] {
            const long long blocks = std::min((half + 255) / 256, 65536LL);
            hipLaunchKernelGGL(k_zeldovich_spec, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                               (const C<double>*)spec, a, px, py, pz);
        });
    // real displacements: x into the noise buffer, y into the noise spectrum's buffer (n^2 (n + 2) doubles
    // hold one real grid), z into a third grid
    double* dx = noise;
    double* dy = (double*)spec;
    double* dz = extra;
    if (!rc) rc = fft3d_c2r<double>(ctx, px, dx, dims);
    if (!rc) rc = fft3d_c2r<double>(ctx, py, dy, dims);
    if (!rc) rc = fft3d_c2r<double>(ctx, pz, dz, dims);
    Perm pm;
    pm.sites = (uint64_t)total;
    int bits = 0;
    while ((1ull << bits) < (uint64_t)total) bits++;
    pm.h = (bits + 1) / 2;
    uint64_t sm = (uint64_t)seed ^ 0x5DEECE66Dull;
    for (int s = 0; s < 4; s++) pm.key[s] = splitmix_next(sm);
    const double inv_n3 = 1.0 / (double)total;
    if (!rc)
        rc = ctx->launch("mock_place", [&] {
            const long long blocks = std::min((np + 255) / 256, 65536LL);
            hipLaunchKernelGGL(k_place, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, np, (int)n, cfg->boxsize, pm,
                               dx, dy, dz, inv_n3, d_pos);
        });
    return rc;
}

}  // extern "C"
