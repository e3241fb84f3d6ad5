// hm_batch.hip -- batched periodic-box pipeline: many mocks of uneven size
// through one box plan, each mock's data vector formed on the device and the
// covariance sums accumulated there in FP64, with no host synchronisation
// between mocks.
//
// Per mock: the assignment stage runs eagerly (its grids depend on the
// particle count) and writes the per-catalogue scalars to device memory; the
// mesh stage (transforms, binning, shot noise, triple sums), the finalisation
// and the covariance update read them from there, so they issue identical
// launches for every mock.  With use_graph they are captured into a hipGraph
// at the second mock (the first runs eagerly and performs every lazy
// allocation) and replayed from then on.
//
// The reference runs one mock per process (README.md:167, main_box.cpp:64-76)
// and forms the covariance downstream with numpy.cov
// (struct/covariance/calc_data_and_cov.py:284-307).
#include "hm_internal.hpp"

#include <cstring>

using namespace hm;

struct hm_batch {
    BoxPlanImpl* plan = nullptr;
    hm_batch_config cfg{};
    int vec_len = 0, cov_len = 0;
    long long capacity = 0, pushed = 0;
    DevBuf dyn, vec, store, shift, s1, s2, count;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int bad(hm_ctx* ctx, const std::string& m) {
    if (ctx) ctx->set_error(m);
    return -1;
}

// Sums of d = x - x_first over the mocks: s1 += d, s2 += d d^T (FP64).  The
// first mock sets the shift (its d is zero).  Also stores the vector.
__global__ __launch_bounds__(256) void k_cov_acc(const double* __restrict__ vec, int vec_len, int cov_len,
                                                 const long long* __restrict__ count, double* __restrict__ shift,
                                                 double* __restrict__ s1, double* __restrict__ s2,
                                                 double* __restrict__ store, long long capacity) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long c = *count;
    if (t < vec_len && c < capacity) store[c * vec_len + t] = vec[t];
    if (t >= (long long)cov_len * cov_len) return;
    const int i = (int)(t / cov_len), j = (int)(t % cov_len);
    if (c == 0) {
        if (j == 0) shift[i] = vec[i];
        return;
    }
    const double di = vec[i] - shift[i], dj = vec[j] - shift[j];
    s2[t] += di * dj;
    if (j == 0) s1[i] += di;
}

__global__ void k_count_inc(long long* count) {
    if (threadIdx.x == 0) *count += 1;
}

// Mesh stage, finalisation and covariance update of the mock whose scalars
// are in b->dyn.
int mesh_tail(hm_ctx* ctx, hm_batch* b) {
    ExecOpts xo;
    xo.dyn = (const double*)b->dyn.p;
    xo.assign = false;
    xo.readback = false;
    // the mesh stage reads the catalogue only through the fixed-point mesh and dyn
    if (plan_execute_opts(ctx, b->plan, nullptr, 1, xo)) return -1;
    if (plan_finalize_device(ctx, b->plan, (const double*)b->dyn.p, (double*)b->vec.p)) return -1;
    const long long nt = std::max((long long)b->cov_len * b->cov_len, (long long)b->vec_len);
    if (ctx->launch("cov_acc", [&] {
            hipLaunchKernelGGL(k_cov_acc, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, ctx->stream,
                               (const double*)b->vec.p, b->vec_len, b->cov_len, (const long long*)b->count.p,
                               (double*)b->shift.p, (double*)b->s1.p, (double*)b->s2.p, (double*)b->store.p,
                               b->capacity);
        }))
        return -1;
    return ctx->launch("count_inc", [&] {
        hipLaunchKernelGGL(k_count_inc, dim3(1), dim3(64), 0, ctx->stream, (long long*)b->count.p);
    });
}

int capture(hm_ctx* ctx, hm_batch* b) {
    HM_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    const int rc = mesh_tail(ctx, b);
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? -1 : bad(ctx, std::string("batch: graph capture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return bad(ctx, std::string("batch: graph instantiate: ") + hipGetErrorString(ei));
    }
    b->graph = g;
    b->exec = x;
    return 0;
}

void batch_free(hm_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    DevBuf* bufs[] = {&b->dyn, &b->vec, &b->store, &b->shift, &b->s1, &b->s2, &b->count};
    for (DevBuf* d : bufs) dev_free(*d);
    plan_destroy(b->plan);
    delete b;
}

}  // namespace

extern "C" {

hm_batch* hm_batch_create(hm_ctx* ctx, const hm_batch_config* cfg) {
    if (!ctx || !cfg) return nullptr;
    if (cfg->max_mocks < 0) return bad(ctx, "batch: max_mocks < 0"), nullptr;
    (void)hipSetDevice(ctx->device);
    auto* b = new (std::nothrow) hm_batch();
    if (!b) return bad(ctx, "batch: out of host memory"), nullptr;
    b->cfg = *cfg;
    b->plan = plan_create(ctx, &cfg->box);
    if (!b->plan) {
        delete b;
        return nullptr;
    }
    const std::string why = plan_batch_check(b->plan);
    if (!why.empty()) {
        batch_free(b);
        return bad(ctx, why), nullptr;
    }
    int npk = 0, nbk = 0;
    plan_vector_dims(b->plan, &npk, &nbk);
    b->vec_len = npk + nbk;
    b->cov_len = (cfg->cov_with_bk || npk == 0) ? b->vec_len : npk;
    b->capacity = cfg->max_mocks;
    const size_t cl = (size_t)b->cov_len;
    if (plan_batch_prepare(ctx, b->plan) || dev_alloc(ctx, b->dyn, sizeof(double) * DYN_LEN) ||
        dev_alloc(ctx, b->vec, sizeof(double) * b->vec_len) ||
        dev_alloc(ctx, b->store, sizeof(double) * (size_t)b->vec_len * (size_t)std::max(b->capacity, 1LL)) ||
        dev_alloc(ctx, b->shift, sizeof(double) * cl) || dev_alloc(ctx, b->s1, sizeof(double) * cl) ||
        dev_alloc(ctx, b->s2, sizeof(double) * cl * cl) || dev_alloc(ctx, b->count, sizeof(long long))) {
        batch_free(b);
        return nullptr;
    }
    if (hm_batch_reset(ctx, b)) {
        batch_free(b);
        return nullptr;
    }
    return b;
}

void hm_batch_destroy(hm_ctx* ctx, hm_batch* b) {
    if (!b) return;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    batch_free(b);
}

int hm_batch_dims(const hm_batch* b, int* vec_len, int* cov_len) {
    if (!b) return -1;
    if (vec_len) *vec_len = b->vec_len;
    if (cov_len) *cov_len = b->cov_len;
    return 0;
}

int hm_batch_reset(hm_ctx* ctx, hm_batch* b) {
    if (!ctx || !b) return -1;
    const size_t cl = (size_t)b->cov_len;
    if (fill_zero(ctx, b->s1.p, sizeof(double) * cl) || fill_zero(ctx, b->s2.p, sizeof(double) * cl * cl) ||
        fill_zero(ctx, b->shift.p, sizeof(double) * cl) || fill_zero(ctx, b->count.p, sizeof(long long)))
        return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    b->pushed = 0;
    return 0;
}

int hm_batch_push(hm_ctx* ctx, hm_batch* b, const double* d_pos, long long n) {
    if (!ctx || !b) return -1;
    if (n <= 0 || !d_pos) return bad(ctx, "batch: a mock needs particles");
    ExecOpts xa;
    xa.dyn = (const double*)b->dyn.p;
    xa.mesh = false;
    xa.readback = false;
    if (plan_execute_opts(ctx, b->plan, d_pos, n, xa)) return -1;
    if (b->cfg.use_graph && b->pushed >= 1) {
        if (!b->exec && capture(ctx, b)) return -1;
        HM_HIP_CHECK(ctx, hipGraphLaunch(b->exec, ctx->stream));
    } else if (mesh_tail(ctx, b)) {
        return -1;
    }
    b->pushed++;
    return 0;
}

int hm_batch_autotune(hm_ctx* ctx, hm_batch* b, const double* d_pos, long long n, int reps, char* report,
                      size_t report_len) {
    if (!ctx || !b) return -1;
    if (!d_pos || n <= 0) return bad(ctx, "autotune: needs a catalogue");
    std::string r;
    const int rc = plan_autotune(ctx, b->plan, d_pos, n, reps, r);
    if (report && report_len > 0) {
        const size_t m = std::min(report_len - 1, r.size());
        memcpy(report, r.data(), m);
        report[m] = 0;
    }
    // a captured graph holds the old shapes: capture again at the next push
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return rc;
}

long long hm_batch_count(const hm_batch* b) { return b ? b->pushed : -1; }

int hm_batch_vectors(hm_ctx* ctx, hm_batch* b, double* out, long long max_rows) {
    if (!ctx || !b || !out) return -1;
    const long long rows = std::min(std::min(b->pushed, b->capacity), max_rows);
    if (rows <= 0) return 0;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(out, b->store.p, sizeof(double) * (size_t)rows * b->vec_len,
                                     hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_batch_covariance(hm_ctx* ctx, hm_batch* b, double* mean, double* cov) {
    if (!ctx || !b) return -1;
    const size_t cl = (size_t)b->cov_len;
    std::vector<double> shift(cl), s1(cl), s2(cl * cl);
    long long count = 0;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(shift.data(), b->shift.p, sizeof(double) * cl, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipMemcpyAsync(s1.data(), b->s1.p, sizeof(double) * cl, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipMemcpyAsync(s2.data(), b->s2.p, sizeof(double) * cl * cl, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipMemcpyAsync(&count, b->count.p, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (count < 2) return bad(ctx, "batch: the covariance needs two or more mocks");
    const double n = (double)count;
    for (size_t i = 0; i < cl; i++)
        if (mean) mean[i] = shift[i] + s1[i] / n;
    if (cov)
        for (size_t i = 0; i < cl; i++)
            for (size_t j = 0; j < cl; j++) cov[i * cl + j] = (s2[i * cl + j] - s1[i] * s1[j] / n) / (n - 1.0);
    return 0;
}

}  // extern "C"
