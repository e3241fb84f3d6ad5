// hm_batch.hip -- batched periodic-box pipeline: many mocks of uneven size
// through box plans, each mock's data vector formed on the device and the
// covariance sums accumulated there in FP64, with no host synchronisation
// between mocks.
//
// Per mock: the assignment stage runs eagerly (its grids depend on the
// particle count) and writes the per-catalogue scalars to device memory; the
// mesh stage (transforms, binning, shot noise, triple sums) and the
// finalisation read them from there, so they issue identical launches for
// every mock.  With use_graph they are captured into a hipGraph at a lane's
// second mock (the first runs eagerly and performs every lazy allocation) and
// replayed from then on.
//
// Lanes: independent plans on streams of their own take mocks in turn, so that
// the latency- and LDS-bound kernels of one mock overlap the bandwidth-bound
// ones of another.  The covariance update of mock i waits for that of mock
// i - 1 (an event), so the sums are accumulated in push order whatever the
// number of lanes.
//
// The reference runs one mock per process (README.md:167, main_box.cpp:64-76)
// and forms the covariance downstream with numpy.cov
// (struct/covariance/calc_data_and_cov.py:284-307).
#include "hm_internal.hpp"

#include <cstring>

using namespace hm;

namespace {

struct Lane {
    hm_ctx* ctx = nullptr;  // the caller's context for lane 0, an own one otherwise
    bool own = false;
    BoxPlanImpl* plan = nullptr;
    DevBuf dyn, vec;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long pushed = 0;
    hipEvent_t ev_read = nullptr, ev_cov = nullptr;
};

}  // namespace

struct hm_batch {
    hm_batch_config cfg{};
    std::vector<Lane> lanes;
    int vec_len = 0, cov_len = 0;
    long long capacity = 0, pushed = 0;
    DevBuf store, shift, s1, s2, count;
    hipEvent_t ev_in = nullptr;      // positions ready on the caller's stream
    hipEvent_t last_cov = nullptr;   // the previous mock's covariance update
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

void drop_graph(Lane& l) {
    if (l.exec) (void)hipGraphExecDestroy(l.exec);
    if (l.graph) (void)hipGraphDestroy(l.graph);
    l.exec = nullptr;
    l.graph = nullptr;
}

// Mesh stage and finalisation of the lane's current mock (scalars in l.dyn).
int mesh_tail(Lane& l) {
    ExecOpts xo;
    xo.dyn = (const double*)l.dyn.p;
    xo.assign = false;
    xo.readback = false;
    // the mesh stage reads the catalogue only through the fixed-point mesh and dyn
    if (plan_execute_opts(l.ctx, l.plan, nullptr, 1, xo)) return -1;
    return plan_finalize_device(l.ctx, l.plan, (const double*)l.dyn.p, (double*)l.vec.p);
}

int capture(Lane& l) {
    hm_ctx* ctx = l.ctx;
    HM_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    const int rc = mesh_tail(l);
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
    l.graph = g;
    l.exec = x;
    return 0;
}

// the covariance update of the lane's vector, after the previous mock's
int cov_tail(hm_batch* b, Lane& l) {
    hm_ctx* ctx = l.ctx;
    if (b->last_cov) HM_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, b->last_cov, 0));
    const long long nt = std::max((long long)b->cov_len * b->cov_len, (long long)b->vec_len);
    if (ctx->launch("cov_acc", [&] {
            hipLaunchKernelGGL(k_cov_acc, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, ctx->stream,
                               (const double*)l.vec.p, b->vec_len, b->cov_len, (const long long*)b->count.p,
                               (double*)b->shift.p, (double*)b->s1.p, (double*)b->s2.p, (double*)b->store.p,
                               b->capacity);
        }))
        return -1;
    if (ctx->launch("count_inc", [&] {
            hipLaunchKernelGGL(k_count_inc, dim3(1), dim3(64), 0, ctx->stream, (long long*)b->count.p);
        }))
        return -1;
    HM_HIP_CHECK(ctx, hipEventRecord(l.ev_cov, ctx->stream));
    b->last_cov = l.ev_cov;
    return 0;
}

int sync_all(hm_ctx* ctx, hm_batch* b) {
    for (Lane& l : b->lanes) HM_HIP_CHECK(ctx, hipStreamSynchronize(l.ctx->stream));
    return 0;
}

void batch_free(hm_batch* b) {
    for (Lane& l : b->lanes) {
        if (l.ctx) (void)hipStreamSynchronize(l.ctx->stream);
        drop_graph(l);
        dev_free(l.dyn);
        dev_free(l.vec);
        plan_destroy(l.plan);
        if (l.ev_read) (void)hipEventDestroy(l.ev_read);
        if (l.ev_cov) (void)hipEventDestroy(l.ev_cov);
        if (l.own) hm_ctx_destroy(l.ctx);
    }
    DevBuf* bufs[] = {&b->store, &b->shift, &b->s1, &b->s2, &b->count};
    for (DevBuf* d : bufs) dev_free(*d);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    delete b;
}

}  // namespace

extern "C" {

hm_batch* hm_batch_create(hm_ctx* ctx, const hm_batch_config* cfg) {
    if (!ctx || !cfg) return nullptr;
    if (cfg->max_mocks < 0) return bad(ctx, "batch: max_mocks < 0"), nullptr;
    const int nl = cfg->lanes < 1 ? 1 : cfg->lanes;
    if (nl > 8) return bad(ctx, "batch: at most 8 lanes"), nullptr;
    (void)hipSetDevice(ctx->device);
    auto* b = new (std::nothrow) hm_batch();
    if (!b) return bad(ctx, "batch: out of host memory"), nullptr;
    b->cfg = *cfg;
    b->lanes.resize(nl);
    for (int i = 0; i < nl; i++) {
        Lane& l = b->lanes[i];
        if (i == 0) {
            l.ctx = ctx;
        } else {
            l.ctx = hm_ctx_create(ctx->device, 0);
            l.own = true;
            if (!l.ctx) {
                batch_free(b);
                return bad(ctx, "batch: cannot create a lane's stream"), nullptr;
            }
        }
        l.plan = plan_create(l.ctx, &cfg->box);
        if (!l.plan) {
            if (l.own) ctx->set_error(l.ctx->err);
            batch_free(b);
            return nullptr;
        }
        if (hipEventCreateWithFlags(&l.ev_read, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&l.ev_cov, hipEventDisableTiming) != hipSuccess) {
            batch_free(b);
            return bad(ctx, "batch: cannot create events"), nullptr;
        }
    }
    if (hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess) {
        batch_free(b);
        return bad(ctx, "batch: cannot create events"), nullptr;
    }
    const std::string why = plan_batch_check(b->lanes[0].plan);
    if (!why.empty()) {
        batch_free(b);
        return bad(ctx, why), nullptr;
    }
    int npk = 0, nbk = 0;
    plan_vector_dims(b->lanes[0].plan, &npk, &nbk);
    b->vec_len = npk + nbk;
    b->cov_len = (cfg->cov_with_bk || npk == 0) ? b->vec_len : npk;
    b->capacity = cfg->max_mocks;
    const size_t cl = (size_t)b->cov_len;
    for (Lane& l : b->lanes)
        if (plan_batch_prepare(l.ctx, l.plan) || dev_alloc(ctx, l.dyn, sizeof(double) * DYN_LEN) ||
            dev_alloc(ctx, l.vec, sizeof(double) * b->vec_len)) {
            if (l.own && !l.ctx->err.empty()) ctx->set_error(l.ctx->err);
            batch_free(b);
            return nullptr;
        }
    if (dev_alloc(ctx, b->store, sizeof(double) * (size_t)b->vec_len * (size_t)std::max(b->capacity, 1LL)) ||
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
    if (sync_all(ctx, b)) return -1;
    const size_t cl = (size_t)b->cov_len;
    if (fill_zero(ctx, b->s1.p, sizeof(double) * cl) || fill_zero(ctx, b->s2.p, sizeof(double) * cl * cl) ||
        fill_zero(ctx, b->shift.p, sizeof(double) * cl) || fill_zero(ctx, b->count.p, sizeof(long long)))
        return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    b->pushed = 0;
    b->last_cov = nullptr;
    for (Lane& l : b->lanes) l.pushed = 0;
    return 0;
}

int hm_batch_push(hm_ctx* ctx, hm_batch* b, const double* d_pos, long long n) {
    if (!ctx || !b) return -1;
    if (n <= 0 || !d_pos) return bad(ctx, "batch: a mock needs particles");
    Lane& l = b->lanes[(size_t)(b->pushed % (long long)b->lanes.size())];
    hm_ctx* lc = l.ctx;
    if (lc != ctx) {
        // the positions are produced on the caller's stream
        HM_HIP_CHECK(ctx, hipEventRecord(b->ev_in, ctx->stream));
        HM_HIP_CHECK(ctx, hipStreamWaitEvent(lc->stream, b->ev_in, 0));
        lc->timing = ctx->timing;
        lc->timing_only = ctx->timing_only;
        lc->wide_mask = ctx->wide_mask;
    }
    ExecOpts xa;
    xa.dyn = (const double*)l.dyn.p;
    xa.mesh = false;
    xa.readback = false;
    const auto fail = [&]() {
        if (lc != ctx) ctx->set_error(lc->err);
        return -1;
    };
    if (plan_execute_opts(lc, l.plan, d_pos, n, xa)) return fail();
    if (lc != ctx) {
        // the caller's later work on its stream waits until the positions are read
        HM_HIP_CHECK(ctx, hipEventRecord(l.ev_read, lc->stream));
        HM_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, l.ev_read, 0));
    }
    if (b->cfg.use_graph && l.pushed >= 1) {
        if (!l.exec && capture(l)) return fail();
        HM_HIP_CHECK(ctx, hipGraphLaunch(l.exec, lc->stream));
    } else if (mesh_tail(l)) {
        return fail();
    }
    if (cov_tail(b, l)) return fail();
    l.pushed++;
    b->pushed++;
    return 0;
}

int hm_batch_autotune(hm_ctx* ctx, hm_batch* b, const double* d_pos, long long n, int reps, char* report,
                      size_t report_len) {
    if (!ctx || !b) return -1;
    if (!d_pos || n <= 0) return bad(ctx, "autotune: needs a catalogue");
    if (sync_all(ctx, b)) return -1;
    std::string r;
    const int rc = plan_autotune(ctx, b->lanes[0].plan, d_pos, n, reps, r);
    if (report && report_len > 0) {
        const size_t m = std::min(report_len - 1, r.size());
        memcpy(report, r.data(), m);
        report[m] = 0;
    }
    unsigned wide = 0;
    int bin = 256;
    plan_tune_get(b->lanes[0].plan, &wide, &bin);
    // the other lanes take the same shapes; captured graphs hold the old ones: capture again
    for (Lane& l : b->lanes) {
        plan_tune_set(l.plan, wide, bin);
        drop_graph(l);
    }
    return rc;
}

long long hm_batch_count(const hm_batch* b) { return b ? b->pushed : -1; }

int hm_batch_collect_timing(hm_ctx* ctx, hm_batch* b) {
    if (!ctx || !b) return -1;
    if (sync_all(ctx, b)) return -1;
    for (Lane& l : b->lanes)
        if (l.own && timing_merge(ctx, l.ctx)) return -1;
    return 0;
}

int hm_batch_vectors(hm_ctx* ctx, hm_batch* b, double* out, long long max_rows) {
    if (!ctx || !b || !out) return -1;
    if (sync_all(ctx, b)) return -1;
    const long long rows = std::min(std::min(b->pushed, b->capacity), max_rows);
    if (rows <= 0) return 0;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(out, b->store.p, sizeof(double) * (size_t)rows * b->vec_len,
                                     hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_batch_covariance(hm_ctx* ctx, hm_batch* b, double* mean, double* cov) {
    if (!ctx || !b) return -1;
    if (sync_all(ctx, b)) return -1;
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
