// hm_fft.hip -- context plumbing, twiddle tables and the 3D transforms built
// from the LDS Stockham passes (row-major grids, FFTW sign conventions).
#include "hm_fft_launch.hpp"

#include <cmath>

hipEvent_t hm_ctx::get_event() {
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int hm_ctx::launch(const char* name, const std::function<void()>& fn) {
    const bool t = timing && !capturing && (timing_only.empty() || timing_only == name);
    hipEvent_t a = nullptr, b = nullptr;
    if (t) {
        a = get_event();
        b = get_event();
        if (a) (void)hipEventRecord(a, stream);
    }
    fn();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("launch of ") + name + ": " + hipGetErrorString(e));
        return -1;
    }
    if (t && a && b) {
        (void)hipEventRecord(b, stream);
        pending.push_back({name, a, b});
    }
    return 0;
}

template <typename T> const hm::C<T>* hm_ctx::twiddle(int L) {
    auto key = std::make_pair(L, (int)sizeof(T));
    auto it = twiddles.find(key);
    if (it != twiddles.end()) return reinterpret_cast<const hm::C<T>*>(it->second);
    std::vector<hm::C<T>> h(L);
    for (int e = 0; e < L; e++) {
        // exp(-2 pi i e / L) from FP64 cos/sin of the reduced angle
        double a = 2.0 * M_PI * (double)e / (double)L;
        h[e].x = (T)std::cos(a);
        h[e].y = (T)(-std::sin(a));
    }
    void* d = nullptr;
    if (hipMalloc(&d, sizeof(hm::C<T>) * L) != hipSuccess) {
        set_error("hipMalloc of twiddle table failed");
        return nullptr;
    }
    if (hipMemcpy(d, h.data(), sizeof(hm::C<T>) * L, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("hipMemcpy of twiddle table failed");
        (void)hipFree(d);
        return nullptr;
    }
    twiddles[key] = d;
    return reinterpret_cast<const hm::C<T>*>(d);
}
template const hm::C<float>* hm_ctx::twiddle<float>(int);
template const hm::C<double>* hm_ctx::twiddle<double>(int);

namespace hm {

int dev_alloc(hm_ctx* ctx, DevBuf& b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return 0;
    dev_free(b);
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        ctx->set_error("hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
        return -1;
    }
    b.bytes = bytes;
    return 0;
}

void dev_free(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

__global__ void k_fill_zero(uint4* __restrict__ p, long long n16, unsigned char* __restrict__ tail, int ntail) {
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long st = (long long)gridDim.x * blockDim.x;
    for (long long i = i0; i < n16; i += st) p[i] = make_uint4(0, 0, 0, 0);
    if (i0 < ntail) tail[i0] = 0;
}

int fill_zero(hm_ctx* ctx, void* p, size_t bytes) {
    if (bytes == 0) return 0;
    const long long n16 = (long long)(bytes / 16);
    const int ntail = (int)(bytes % 16);
    long long blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    unsigned char* tail = reinterpret_cast<unsigned char*>(p) + n16 * 16;
    return ctx->launch("fill_zero", [&] {
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                           reinterpret_cast<uint4*>(p), n16, tail, ntail);
    });
}

// Every length in [2, RT_MAX_L]: a compile-time radix plan where one exists
// (kFftLengths), the run-time plan otherwise.
bool fft_length_supported(int L) { return L >= 2 && L <= RT_MAX_L; }

int fft_length_at_least(int n) {
    for (int v : kFftLengths)
        if (v >= n) return v;
    return -1;
}

template <typename T>
int fft_cols(hm_ctx* ctx, const char* name, int L, int dir, const C<T>* src, C<T>* dst, const ColGeom& g) {
    GSrc<T> s{src};
    return dir < 0 ? launch_cols<T, -1>(ctx, name, L, s, dst, g) : launch_cols<T, +1>(ctx, name, L, s, dst, g);
}

template <typename T>
int fft_rows(hm_ctx* ctx, const char* name, int L, int dir, const C<T>* src, C<T>* dst, const RowGeom& g) {
    return dir < 0 ? launch_rows<T, -1>(ctx, name, L, src, dst, g) : launch_rows<T, +1>(ctx, name, L, src, dst, g);
}

template <typename T>
int fft_c2r_rows(hm_ctx* ctx, const char* name, int L2, const C<T>* src, T* dst, long long n_rows, int point_s) {
    HalfSrc<T> s{src, L2};
    return launch_c2r<T>(ctx, name, L2, s, dst, n_rows, point_s);
}

static ColGeom col_geom_full(long long OS, long long PS, int n_outer, int n_inner) {
    ColGeom g;
    g.OS = OS;
    g.PS = PS;
    g.n_outer = n_outer;
    g.outer_full = n_outer;
    g.outer_s = -1;
    g.n_inner = n_inner;
    g.inner_full = n_inner;
    g.inner_s = -1;
    g.point_s = -1;
    return g;
}

template <typename T> int fft3d_c2c(hm_ctx* ctx, C<T>* data, const int n[3], int dir) {
    const long long n0 = n[0], n1 = n[1], n2 = n[2];
    RowGeom rg{n0 * n1, n2, n2, -1};
    if (fft_rows<T>(ctx, "fft_c2c_z", (int)n2, dir, data, data, rg)) return -1;
    if (fft_cols<T>(ctx, "fft_c2c_y", (int)n1, dir, data, data, col_geom_full(n1 * n2, n2, (int)n0, (int)n2)))
        return -1;
    if (n1 * n2 > 0x7fffffffLL) {
        ctx->set_error("fft3d: plane too large");
        return -1;
    }
    return fft_cols<T>(ctx, "fft_c2c_x", (int)n0, dir, data, data, col_geom_full(0, n1 * n2, 1, (int)(n1 * n2)));
}

template <typename T> int fft3d_r2c(hm_ctx* ctx, const T* in, C<T>* out, const int n[3]) {
    const long long n0 = n[0], n1 = n[1], n2 = n[2], nzh = n2 / 2 + 1;
    if (n2 % 2) {
        ctx->set_error("fft r2c: n_z must be even");
        return -1;
    }
    RealSrc<T> rs{in, (int)(n2 / 2)};
    if (launch_r2c<T>(ctx, "fft_r2c_z", (int)(n2 / 2), rs, out, n0 * n1)) return -1;
    if (fft_cols<T>(ctx, "fft_r2c_y", (int)n1, -1, out, out, col_geom_full(n1 * nzh, nzh, (int)n0, (int)nzh)))
        return -1;
    return fft_cols<T>(ctx, "fft_r2c_x", (int)n0, -1, out, out, col_geom_full(0, n1 * nzh, 1, (int)(n1 * nzh)));
}

template <typename T> int fft3d_c2r(hm_ctx* ctx, C<T>* in, T* out, const int n[3]) {
    const long long n0 = n[0], n1 = n[1], n2 = n[2], nzh = n2 / 2 + 1;
    if (n2 % 2) {
        ctx->set_error("fft c2r: n_z must be even");
        return -1;
    }
    if (fft_cols<T>(ctx, "fft_c2r_x", (int)n0, +1, in, in, col_geom_full(0, n1 * nzh, 1, (int)(n1 * nzh)))) return -1;
    if (fft_cols<T>(ctx, "fft_c2r_y", (int)n1, +1, in, in, col_geom_full(n1 * nzh, nzh, (int)n0, (int)nzh))) return -1;
    return fft_c2r_rows<T>(ctx, "fft_c2r_z", (int)(n2 / 2), in, out, n0 * n1, -1);
}

template int fft3d_c2c<float>(hm_ctx*, C<float>*, const int*, int);
template int fft3d_c2c<double>(hm_ctx*, C<double>*, const int*, int);
template int fft3d_r2c<float>(hm_ctx*, const float*, C<float>*, const int*);
template int fft3d_r2c<double>(hm_ctx*, const double*, C<double>*, const int*);
template int fft3d_c2r<float>(hm_ctx*, C<float>*, float*, const int*);
template int fft3d_c2r<double>(hm_ctx*, C<double>*, double*, const int*);
template int fft_cols<float>(hm_ctx*, const char*, int, int, const C<float>*, C<float>*, const ColGeom&);
template int fft_cols<double>(hm_ctx*, const char*, int, int, const C<double>*, C<double>*, const ColGeom&);
template int fft_rows<float>(hm_ctx*, const char*, int, int, const C<float>*, C<float>*, const RowGeom&);
template int fft_rows<double>(hm_ctx*, const char*, int, int, const C<double>*, C<double>*, const RowGeom&);
template int fft_c2r_rows<float>(hm_ctx*, const char*, int, const C<float>*, float*, long long, int);
template int fft_c2r_rows<double>(hm_ctx*, const char*, int, const C<double>*, double*, long long, int);

}  // namespace hm
