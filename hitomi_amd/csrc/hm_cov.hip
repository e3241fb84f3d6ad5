// hm_cov.hip -- covariance of data vectors accumulated on the device in FP64.
//
// The reference forms its covariances downstream, in numpy: every mock's
// 2PCF and 3PCF multipoles are read back from text files, divided bin by bin
// by the monopole of the window (xi_l / w_0, zeta_l1l2L / w_000), and
// numpy.cov is taken of the stacked vectors and of every pair of blocks
// (struct/covariance/boss/modified_gravity/calc_data_and_cov.py:72-105,
// 284-307).  Here the vectors stay on the device: each push divides by the
// window vector, and adds d = x - x_0 (x_0: the first vector, which keeps the
// sums well conditioned) into sum d and sum d d^T, so the N - 1 covariance of
// every block pair is one matrix, read back once at the end.  The pushes are
// ordered on the context's stream; nothing synchronises with the host.
#include "hm_internal.hpp"

#include <cstring>

using namespace hm;

struct hm_cov {
    int dim = 0;
    long long capacity = 0, pushed = 0;
    bool has_norm = false;
    DevBuf norm, shift, s1, s2, count, store, stage;
};

namespace {

int bad(hm_ctx* ctx, const std::string& m) {
    if (ctx) ctx->set_error(m);
    return -1;
}

// x = v / w (elementwise, IEEE FP64 division, as numpy divides), then with
// d = x - x_0: s1 += d, s2 += d d^T.  The first vector sets x_0.  Row c of the
// store keeps x while c < capacity.
__global__ __launch_bounds__(256) void k_cov_push(const double* __restrict__ v, const double* __restrict__ w, int dim,
                                                  const long long* __restrict__ count, double* __restrict__ shift,
                                                  double* __restrict__ s1, double* __restrict__ s2,
                                                  double* __restrict__ store, long long capacity) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)dim * dim) return;
    const long long c = *count;
    const int i = (int)(t / dim), j = (int)(t % dim);
    const double xi = w ? v[i] / w[i] : v[i];
    if (j == 0 && c < capacity) store[c * dim + i] = xi;
    if (c == 0) {
        if (j == 0) shift[i] = xi;
        return;
    }
    const double xj = w ? v[j] / w[j] : v[j];
    const double di = xi - shift[i], dj = xj - shift[j];
    s2[t] += di * dj;
    if (j == 0) s1[i] += di;
}

__global__ void k_cov_count(long long* count) {
    if (threadIdx.x == 0) *count += 1;
}

void cov_free(hm_cov* c) {
    DevBuf* bufs[] = {&c->norm, &c->shift, &c->s1, &c->s2, &c->count, &c->store, &c->stage};
    for (DevBuf* d : bufs) dev_free(*d);
    delete c;
}

}  // namespace

extern "C" {

hm_cov* hm_cov_create(hm_ctx* ctx, int dim, const double* norm, long long keep_rows) {
    if (!ctx) return nullptr;
    if (dim < 1 || dim > 4096) return bad(ctx, "covariance: the vector length must be in [1, 4096]"), nullptr;
    if (keep_rows < 0) return bad(ctx, "covariance: keep_rows < 0"), nullptr;
    (void)hipSetDevice(ctx->device);
    auto* c = new (std::nothrow) hm_cov();
    if (!c) return bad(ctx, "covariance: out of host memory"), nullptr;
    c->dim = dim;
    c->capacity = keep_rows;
    const size_t d = (size_t)dim;
    if (norm) {
        for (int i = 0; i < dim; i++)
            if (!(norm[i] != 0.0)) {
                cov_free(c);
                return bad(ctx, "covariance: a window value is zero or NaN"), nullptr;
            }
        c->has_norm = true;
    }
    if (dev_alloc(ctx, c->shift, sizeof(double) * d) || dev_alloc(ctx, c->s1, sizeof(double) * d) ||
        dev_alloc(ctx, c->s2, sizeof(double) * d * d) || dev_alloc(ctx, c->count, sizeof(long long)) ||
        dev_alloc(ctx, c->store, sizeof(double) * d * (size_t)std::max(keep_rows, 1LL)) ||
        dev_alloc(ctx, c->stage, sizeof(double) * d) || (norm && dev_alloc(ctx, c->norm, sizeof(double) * d))) {
        cov_free(c);
        return nullptr;
    }
    if (norm && hipMemcpy(c->norm.p, norm, sizeof(double) * d, hipMemcpyHostToDevice) != hipSuccess) {
        cov_free(c);
        return bad(ctx, "covariance: cannot copy the window vector"), nullptr;
    }
    if (hm_cov_reset(ctx, c)) {
        cov_free(c);
        return nullptr;
    }
    return c;
}

void hm_cov_destroy(hm_ctx* ctx, hm_cov* c) {
    if (!c) return;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    cov_free(c);
}

int hm_cov_reset(hm_ctx* ctx, hm_cov* c) {
    if (!ctx || !c) return -1;
    const size_t d = (size_t)c->dim;
    if (fill_zero(ctx, c->s1.p, sizeof(double) * d) || fill_zero(ctx, c->s2.p, sizeof(double) * d * d) ||
        fill_zero(ctx, c->shift.p, sizeof(double) * d) || fill_zero(ctx, c->count.p, sizeof(long long)))
        return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    c->pushed = 0;
    return 0;
}

int hm_cov_push(hm_ctx* ctx, hm_cov* c, const double* d_vec) {
    if (!ctx || !c) return -1;
    if (!d_vec) return bad(ctx, "covariance: NULL vector");
    const long long nt = (long long)c->dim * c->dim;
    if (ctx->launch("cov_push", [&] {
            hipLaunchKernelGGL(k_cov_push, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, ctx->stream, d_vec,
                               c->has_norm ? (const double*)c->norm.p : nullptr, c->dim,
                               (const long long*)c->count.p, (double*)c->shift.p, (double*)c->s1.p,
                               (double*)c->s2.p, (double*)c->store.p, c->capacity);
        }))
        return -1;
    if (ctx->launch("count_inc", [&] {
            hipLaunchKernelGGL(k_cov_count, dim3(1), dim3(64), 0, ctx->stream, (long long*)c->count.p);
        }))
        return -1;
    c->pushed++;
    return 0;
}

int hm_cov_push_host(hm_ctx* ctx, hm_cov* c, const double* vec) {
    if (!ctx || !c) return -1;
    if (!vec) return bad(ctx, "covariance: NULL vector");
    HM_HIP_CHECK(ctx, hipMemcpyAsync(c->stage.p, vec, sizeof(double) * (size_t)c->dim, hipMemcpyHostToDevice,
                                     ctx->stream));
    // the staging buffer is reused by the next push: the copy must not be overtaken
    if (hm_cov_push(ctx, c, (const double*)c->stage.p)) return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

long long hm_cov_count(const hm_cov* c) { return c ? c->pushed : -1; }

int hm_cov_vectors(hm_ctx* ctx, hm_cov* c, double* out, long long max_rows) {
    if (!ctx || !c || !out) return -1;
    const long long rows = std::min(std::min(c->pushed, c->capacity), max_rows);
    if (rows <= 0) return 0;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(out, c->store.p, sizeof(double) * (size_t)rows * c->dim, hipMemcpyDeviceToHost,
                                     ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_cov_result(hm_ctx* ctx, hm_cov* c, double* mean, double* cov) {
    if (!ctx || !c) return -1;
    const size_t d = (size_t)c->dim;
    std::vector<double> shift(d), s1(d), s2(d * d);
    long long count = 0;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(shift.data(), c->shift.p, sizeof(double) * d, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipMemcpyAsync(s1.data(), c->s1.p, sizeof(double) * d, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipMemcpyAsync(s2.data(), c->s2.p, sizeof(double) * d * d, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipMemcpyAsync(&count, c->count.p, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (count < 2) return bad(ctx, "covariance: needs two or more vectors");
    const double n = (double)count;
    if (mean)
        for (size_t i = 0; i < d; i++) mean[i] = shift[i] + s1[i] / n;
    if (cov)
        for (size_t i = 0; i < d; i++)
            for (size_t j = 0; j < d; j++) cov[i * d + j] = (s2[i * d + j] - s1[i] * s1[j] / n) / (n - 1.0);
    return 0;
}

}  // extern "C"
