// hm_common.hpp -- shared device/host definitions for the gfx950 kernels of
// hitomi_amd (MI355X-native FFT estimator path of hitomi_measurement).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace hm {

// Complex element types: interleaved (re, im), 8 B (FP32) or 16 B (FP64).
template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using C = typename cvec<T>::type;

template <typename T> __host__ __device__ inline C<T> mk(T x, T y) { C<T> r; r.x = x; r.y = y; return r; }
template <typename T> __host__ __device__ inline C<T> cadd(C<T> a, C<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __host__ __device__ inline C<T> csub(C<T> a, C<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __host__ __device__ inline C<T> cmul(C<T> a, C<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __host__ __device__ inline C<T> cconj(C<T> a) { return mk<T>(a.x, -a.y); }
template <typename T> __host__ __device__ inline C<T> cscale(C<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// (s * i) * a for s = +-1
template <typename T> __host__ __device__ inline C<T> cmuli(C<T> a, int s) {
    return s > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

// Error state carried by a context (no exceptions cross the C ABI).
struct Error {
    std::string msg;
};

#define HM_HIP_CHECK(ctx, expr)                                                          \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            (ctx)->set_error(std::string(#expr) + ": " + hipGetErrorString(_e));         \
            return -1;                                                                   \
        }                                                                                \
    } while (0)

// Geometry of a periodic mesh as the reference's ParameterClass holds it
// (parameter.hpp:13-16): n_mesh, boxsize, with derived k and r spacings.
struct Mesh {
    int n[3];
    double L[3];
    int assign;  // 0 NGP, 1 CIC, 2 TSC
    long long ng() const { return (long long)n[0] * n[1] * n[2]; }
    double volume() const { return L[0] * L[1] * L[2]; }
    int nzh() const { return n[2] / 2 + 1; }
    long long nh() const { return (long long)n[0] * n[1] * nzh(); }
};

}  // namespace hm
