// hm_api.hip -- the extern "C" boundary of libhitomi_amd.so (declared in
// include/hitomi_amd.h).  Every entry point catches its own failures and
// reports them through hm_last_error().
#include "hm_assign.hpp"
#include "hm_internal.hpp"
#include "hm_bessel.hpp"
#include "hm_bins.hpp"
#include "hm_recon.hpp"

#include "../../include/hitomi_amd.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>

namespace hm {
int plan_execute_any(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos, long long np, hm_box_result* res,
                     const double* d_pos_r, long long nr, double alpha);
void plan_triple_mesh(const BoxPlanImpl* p, int m[3]);
int cell_coords(hm_ctx* ctx, const AssignGeom& g, const double* d_pos, long long np, long long* d_out);
struct SurveyPlanImpl;
SurveyPlanImpl* survey_create(hm_ctx* ctx, const hm_survey_config* cfg);
void survey_destroy(SurveyPlanImpl* p);
int survey_execute(hm_ctx* ctx, SurveyPlanImpl* p, const double* pD, const double* wD, long long nD, const double* pR,
                   const double* wR, long long nR, hm_survey_info* info, double* pk, long long* nmode);
size_t survey_bytes(const SurveyPlanImpl* p);
int fixed_to_density(hm_ctx* ctx, const long long* d_q, long long n, double inv_scale, double factor, double* d_out);
}  // namespace hm

namespace hm {
// Plain device copy: 16 B per lane where both pointers allow it, bytes for the rest.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ s, uint4* __restrict__ d, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) d[i] = s[i];
}
__global__ __launch_bounds__(256) void k_copy1(const unsigned char* __restrict__ s, unsigned char* __restrict__ d,
                                               long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) d[i] = s[i];
}
static unsigned copy_blocks(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 1 << 16); }
}  // namespace hm

struct hm_box_plan {
    hm::BoxPlanImpl* impl;
};
struct hm_survey_plan {
    hm::SurveyPlanImpl* impl;
};

static int bad(hm_ctx* ctx, const char* msg) {
    if (ctx) ctx->set_error(msg);
    return -1;
}

static bool mesh_ok(hm_ctx* ctx, const hm_mesh* m, hm::AssignGeom& g) {
    if (!m) return bad(ctx, "mesh is NULL"), false;
    for (int a = 0; a < 3; a++) {
        if (m->n_mesh[a] <= 0) return bad(ctx, "n_mesh must be > 0"), false;
        if (!(m->boxsize[a] > 0.0)) return bad(ctx, "boxsize must be > 0"), false;
        g.n[a] = m->n_mesh[a];
        g.L[a] = m->boxsize[a];
    }
    if (m->assign < HM_NGP || m->assign > HM_TSC) return bad(ctx, "assign must be NGP, CIC or TSC"), false;
    g.assign = m->assign;
    g.ng = (long long)m->n_mesh[0] * m->n_mesh[1] * m->n_mesh[2];
    return true;
}

extern "C" {

// Scratch of hm_assign_noalloc: the sort work of the sorted mass assignment
// (per-chunk counts and offsets, per-particle slots, positions and weights in
// cell order, the border list, scan sums) for callers that cannot allocate.
// Every process that loads the library holds it on each device it uses, until
// the process ends.  In host code `scratch` is only a stand-in for the device
// symbol: its device address is looked up once per device (scratch_setup) and
// only that address is handed to kernels.
#define HM_NOALLOC_SCRATCH_BYTES (128ull << 20)
#define HM_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[HM_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[HM_NOALLOC_MAX_DEVICES];
static int g_scratch_state[HM_NOALLOC_MAX_DEVICES];  // 0: not looked up, 1: ok, -1: the lookup failed

// With `device` current: the device address of this device's copy of scratch.
static void scratch_setup(int device) {
    if (device < 0 || device >= HM_NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_state[device] != 0) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) {
        g_scratch_ptr[device] = p;
        g_scratch_state[device] = 1;
    } else {
        (void)hipGetLastError();
        g_scratch_state[device] = -1;
    }
}

static void* scratch_ptr(int device) {
    if (device < 0 || device >= HM_NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_state[device] == 1 ? g_scratch_ptr[device] : nullptr;
}

hm_ctx* hm_ctx_create(int device, int flags) {
    (void)flags;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    hm_ctx* ctx = new (std::nothrow) hm_ctx();
    if (!ctx) return nullptr;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        if (prop.sharedMemPerBlock > 0) ctx->lds_max = (size_t)prop.sharedMemPerBlock;
        if (prop.multiProcessorCount > 0) ctx->cu_count = prop.multiProcessorCount;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return nullptr;
    }
    scratch_setup(device);
    return ctx;
}

void hm_ctx_destroy(hm_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->twiddles) (void)hipFree(kv.second);
    if (ctx->mock_ws) (void)hipFree(ctx->mock_ws);
    for (auto& kv : ctx->jl_splines) (void)hipFree(kv.second);
    for (auto& pe : ctx->pending) {
        (void)hipEventDestroy(pe.a);
        (void)hipEventDestroy(pe.b);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->tref) (void)hipEventDestroy(ctx->tref);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* hm_last_error(const hm_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int hm_sync(hm_ctx* ctx) {
    if (!ctx) return -1;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

void* hm_dev_alloc(hm_ctx* ctx, size_t bytes) {
    if (!ctx) return nullptr;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        ctx->set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return nullptr;
    }
    return p;
}

int hm_dev_free(hm_ctx* ctx, void* d_ptr) {
    if (!ctx) return -1;
    if (!d_ptr) return 0;
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    HM_HIP_CHECK(ctx, hipFree(d_ptr));
    return 0;
}

int hm_memcpy_h2d(hm_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
    if (!ctx) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_memcpy_d2d(hm_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    if (!ctx) return -1;
    if (!bytes) return 0;
    if (!d_dst || !d_src) return bad(ctx, "memcpy_d2d: NULL pointer");
    const bool v = (((size_t)d_dst | (size_t)d_src) & 15) == 0;
    const long long n16 = v ? (long long)(bytes / 16) : 0;
    const long long tail = (long long)bytes - 16 * n16;
    return ctx->launch("copy", [&] {
        if (n16)
            hipLaunchKernelGGL(hm::k_copy16, dim3(hm::copy_blocks(n16)), dim3(256), 0, ctx->stream, (const uint4*)d_src,
                               (uint4*)d_dst, n16);
        if (tail)
            hipLaunchKernelGGL(hm::k_copy1, dim3(hm::copy_blocks(tail)), dim3(256), 0, ctx->stream,
                               (const unsigned char*)d_src + 16 * n16, (unsigned char*)d_dst + 16 * n16, tail);
    });
}

int hm_memcpy_d2h(hm_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
    if (!ctx) return -1;
    HM_HIP_CHECK(ctx, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int hm_memset_zero(hm_ctx* ctx, void* d_ptr, size_t bytes) {
    if (!ctx) return -1;
    return hm::fill_zero(ctx, d_ptr, bytes);
}

int hm_timing_enable(hm_ctx* ctx, int on) {
    if (!ctx) return -1;
    ctx->timing = on != 0;
    return 0;
}

}  // extern "C"

// src's finished launches into dst's totals; their spans are placed on dst's
// clock (dst->tref), which works across streams of one device.
static int timing_collect_into(hm_ctx* dst, hm_ctx* src) {
    if (src->pending.empty()) return 0;
    HM_HIP_CHECK(src, hipStreamSynchronize(src->stream));
    const bool clock = dst->tref && hipEventSynchronize(dst->tref) == hipSuccess;
    for (auto& pe : src->pending) {
        float ms = 0.f, t0 = 0.f;
        if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess) {
            auto& s = dst->kstats[pe.name];
            s.first += ms;
            s.second += 1;
            if (clock && hipEventElapsedTime(&t0, dst->tref, pe.a) == hipSuccess)
                dst->kspans[pe.name].push_back({(double)t0, (double)t0 + (double)ms});
        }
        src->event_pool.push_back(pe.a);
        src->event_pool.push_back(pe.b);
    }
    src->pending.clear();
    (void)hipGetLastError();
    return 0;
}

static int timing_collect(hm_ctx* ctx) { return timing_collect_into(ctx, ctx); }

namespace hm {
int timing_merge(hm_ctx* dst, hm_ctx* src) {
    if (!dst || !src || dst == src) return 0;
    if (timing_collect_into(dst, src)) {
        dst->set_error(src->err);
        return -1;
    }
    for (auto& kv : src->kstats) {
        auto& s = dst->kstats[kv.first];
        s.first += kv.second.first;
        s.second += kv.second.second;
    }
    src->kstats.clear();
    src->kspans.clear();
    return 0;
}
}  // namespace hm

extern "C" {

int hm_timing_get(hm_ctx* ctx, const char* name, double* total_ms, long long* launches) {
    if (!ctx || !name) return -1;
    if (timing_collect(ctx)) return -1;
    auto it = ctx->kstats.find(name);
    if (total_ms) *total_ms = it == ctx->kstats.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == ctx->kstats.end() ? 0 : it->second.second;
    return 0;
}

int hm_timing_reset(hm_ctx* ctx) {
    if (!ctx) return -1;
    if (timing_collect(ctx)) return -1;
    ctx->kstats.clear();
    ctx->kspans.clear();
    if (!ctx->tref) HM_HIP_CHECK(ctx, hipEventCreate(&ctx->tref));
    HM_HIP_CHECK(ctx, hipEventRecord(ctx->tref, ctx->stream));
    return 0;
}

int hm_timing_get_busy(hm_ctx* ctx, const char* name, double* busy_ms) {
    if (!ctx || !name || !busy_ms) return -1;
    if (timing_collect(ctx)) return -1;
    *busy_ms = 0.0;
    auto it = ctx->kspans.find(name);
    if (it == ctx->kspans.end()) return 0;
    std::vector<std::pair<double, double>> v = it->second;
    std::sort(v.begin(), v.end());
    double lo = 0.0, hi = -1.0, sum = 0.0;
    for (const auto& s : v) {
        if (s.first > hi) {
            if (hi > lo) sum += hi - lo;
            lo = s.first;
            hi = s.second;
        } else if (s.second > hi) {
            hi = s.second;
        }
    }
    if (hi > lo) sum += hi - lo;
    *busy_ms = sum;
    return 0;
}

int hm_timing_select(hm_ctx* ctx, const char* name) {
    if (!ctx) return -1;
    ctx->timing_only = name ? name : "";
    return 0;
}

static int assign_impl(hm_ctx* ctx, const hm_mesh* mesh, const double* d_pos, const double* d_w, long long np,
                       double wabs_sum, double* d_out_re, double* d_out_im, bool sorted) {
    if (!ctx) return -1;
    hm::AssignGeom g;
    if (!mesh_ok(ctx, mesh, g)) return -1;
    if (np < 0 || (np > 0 && !d_pos) || !d_out_re) return bad(ctx, "hm_assign: bad arguments");
    if (!(wabs_sum > 0.0)) wabs_sum = (double)np;
    const double scale = hm::fixed_scale_for(wabs_sum);
    hm::DevBuf q;
    const size_t bytes = sizeof(long long) * (size_t)g.ng * (d_out_im ? 2 : 1);
    if (hm::dev_alloc(ctx, q, bytes)) return -1;
    long long* qre = (long long*)q.p;
    long long* qim = d_out_im ? qre + g.ng : nullptr;
    int rc = 0;
    hm::SortWork w;
    if (sorted) {
        rc = hm::sort_work_alloc(ctx, w, g, np, d_w != nullptr);
        if (!rc) rc = hm::assign_sorted(ctx, g, d_pos, d_w, np, scale, qre, qim, w);
    } else {
        rc = hm::fill_zero(ctx, q.p, bytes);
        if (!rc) rc = hm::assign_fixed(ctx, g, d_pos, d_w, np, scale, qre, qim);
    }
    const double dV = (g.L[0] * g.L[1] * g.L[2]) / (double)g.ng;
    const double factor = 1.0 / dV;
    if (!rc) rc = hm::fixed_to_density(ctx, qre, g.ng, 1.0 / scale, factor, d_out_re);
    if (!rc && qim) rc = hm::fixed_to_density(ctx, qim, g.ng, 1.0 / scale, factor, d_out_im);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = bad(ctx, "hm_assign: stream error");
    hm::sort_work_free(w);
    hm::dev_free(q);
    return rc;
}

int hm_assign(hm_ctx* ctx, const hm_mesh* mesh, const double* d_pos, const double* d_w, long long np,
              double wabs_sum, double* d_out_re, double* d_out_im) {
    return assign_impl(ctx, mesh, d_pos, d_w, np, wabs_sum, d_out_re, d_out_im, true);
}

int hm_assign_atomic(hm_ctx* ctx, const hm_mesh* mesh, const double* d_pos, const double* d_w, long long np,
                     double wabs_sum, double* d_out_re, double* d_out_im) {
    return assign_impl(ctx, mesh, d_pos, d_w, np, wabs_sum, d_out_re, d_out_im, false);
}

int hm_assign_noalloc(hm_ctx* ctx, const hm_mesh* mesh, const double* d_pos, const double* d_w, long long np,
                      double wabs_sum, double* d_out_re, double* d_out_im) {
    if (!ctx) return -1;
    hm::AssignGeom g;
    if (!mesh_ok(ctx, mesh, g)) return -1;
    if (np < 0 || (np > 0 && !d_pos) || !d_out_re) return bad(ctx, "hm_assign_noalloc: bad arguments");
    if (d_out_im == d_out_re) return bad(ctx, "hm_assign_noalloc: d_out_re and d_out_im must not alias");
    char* base = (char*)scratch_ptr(ctx->device);
    if (!base) return bad(ctx, "hm_assign_noalloc: the device address of the library's scratch is unavailable");
    hm::SortWork w;
    const size_t need = hm::sort_work_carve(g, np, d_w != nullptr, nullptr, &w);
    if (need > HM_NOALLOC_SCRATCH_BYTES)
        return bad(ctx, ("hm_assign_noalloc: " + std::to_string(np) + " particles on this mesh need " +
                         std::to_string(need) + " bytes of scratch, more than the " +
                         std::to_string(HM_NOALLOC_SCRATCH_BYTES) + " the library holds; use hm_assign")
                            .c_str());
    hm::sort_work_carve(g, np, d_w != nullptr, base, &w);
    if (!(wabs_sum > 0.0)) wabs_sum = (double)np;
    const double scale = hm::fixed_scale_for(wabs_sum);
    // the output grids hold the int64 cell sums until they are converted in place
    long long* qre = reinterpret_cast<long long*>(d_out_re);
    long long* qim = reinterpret_cast<long long*>(d_out_im);
    int rc = hm::assign_sorted(ctx, g, d_pos, d_w, np, scale, qre, qim, w);
    const double dV = (g.L[0] * g.L[1] * g.L[2]) / (double)g.ng;
    const double factor = 1.0 / dV;
    if (!rc) rc = hm::fixed_to_density_inplace(ctx, qre, g.ng, 1.0 / scale, factor);
    if (!rc && qim) rc = hm::fixed_to_density_inplace(ctx, qim, g.ng, 1.0 / scale, factor);
    return rc;
}

int hm_cell_coords(hm_ctx* ctx, const hm_mesh* mesh, const double* d_pos, long long np, long long* d_coords) {
    if (!ctx) return -1;
    hm::AssignGeom g;
    if (!mesh_ok(ctx, mesh, g)) return -1;
    if (np < 0 || (np > 0 && (!d_pos || !d_coords))) return bad(ctx, "hm_cell_coords: bad arguments");
    return hm::cell_coords(ctx, g, d_pos, np, d_coords);
}

static bool fft_dims_ok(hm_ctx* ctx, const int n[3], bool real) {
    for (int a = 0; a < 3; a++)
        if (!hm::fft_length_supported(n[a])) return bad(ctx, "fft: unsupported length"), false;
    if (real && !hm::fft_length_supported(n[2] / 2)) return bad(ctx, "fft: unsupported real length"), false;
    return true;
}

int hm_fft_c2c(hm_ctx* ctx, int precision, void* d_data, const int n[3], int sign) {
    if (!ctx) return -1;
    if (!n || !d_data || (sign != -1 && sign != 1)) return bad(ctx, "hm_fft_c2c: bad arguments");
    if (!fft_dims_ok(ctx, n, false)) return -1;
    return precision == HM_FP64 ? hm::fft3d_c2c<double>(ctx, (double2*)d_data, n, sign)
                                : hm::fft3d_c2c<float>(ctx, (float2*)d_data, n, sign);
}

int hm_fft_r2c(hm_ctx* ctx, int precision, const void* d_in, void* d_out, const int n[3]) {
    if (!ctx) return -1;
    if (!n || !d_in || !d_out) return bad(ctx, "hm_fft_r2c: bad arguments");
    if (!fft_dims_ok(ctx, n, true)) return -1;
    return precision == HM_FP64 ? hm::fft3d_r2c<double>(ctx, (const double*)d_in, (double2*)d_out, n)
                                : hm::fft3d_r2c<float>(ctx, (const float*)d_in, (float2*)d_out, n);
}

int hm_fft_c2r(hm_ctx* ctx, int precision, void* d_in, void* d_out, const int n[3]) {
    if (!ctx) return -1;
    if (!n || !d_in || !d_out) return bad(ctx, "hm_fft_c2r: bad arguments");
    if (!fft_dims_ok(ctx, n, true)) return -1;
    return precision == HM_FP64 ? hm::fft3d_c2r<double>(ctx, (double2*)d_in, (double*)d_out, n)
                                : hm::fft3d_c2r<float>(ctx, (float2*)d_in, (float*)d_out, n);
}

hm_box_plan* hm_box_plan_create(hm_ctx* ctx, const hm_box_config* cfg) {
    if (!ctx) return nullptr;
    if (!cfg) {
        ctx->set_error("hm_box_plan_create: cfg is NULL");
        return nullptr;
    }
    hm::BoxPlanImpl* impl = nullptr;
    try {
        impl = hm::plan_create(ctx, cfg);
    } catch (const std::exception& e) {
        ctx->set_error(std::string("hm_box_plan_create: ") + e.what());
        return nullptr;
    }
    if (!impl) return nullptr;
    return new hm_box_plan{impl};
}

void hm_box_plan_destroy(hm_ctx* ctx, hm_box_plan* plan) {
    if (!plan) return;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    hm::plan_destroy(plan->impl);
    delete plan;
}

int hm_box_plan_execute(hm_ctx* ctx, hm_box_plan* plan, const double* d_pos, long long np, hm_box_result* res) {
    if (!ctx) return -1;
    if (!plan || !d_pos) return bad(ctx, "hm_box_plan_execute: bad arguments");
    try {
        return hm::plan_execute_any(ctx, plan->impl, d_pos, np, res, nullptr, 0, 0.0);
    } catch (const std::exception& e) {
        ctx->set_error(std::string("hm_box_plan_execute: ") + e.what());
        return -1;
    }
}

int hm_box_plan_execute_recon(hm_ctx* ctx, hm_box_plan* plan, const double* d_pos_d, long long n_d,
                              const double* d_pos_r, long long n_r, double alpha, hm_box_result* res) {
    if (!ctx) return -1;
    if (!plan || !d_pos_d || !d_pos_r || n_r <= 0) return bad(ctx, "hm_box_plan_execute_recon: bad arguments");
    try {
        return hm::plan_execute_any(ctx, plan->impl, d_pos_d, n_d, res, d_pos_r, n_r, alpha);
    } catch (const std::exception& e) {
        ctx->set_error(std::string("hm_box_plan_execute_recon: ") + e.what());
        return -1;
    }
}

size_t hm_box_plan_bytes(const hm_box_plan* plan) { return plan ? hm::plan_bytes(plan->impl) : 0; }

int hm_box_plan_autotune(hm_ctx* ctx, hm_box_plan* plan, const double* d_pos, long long np, int reps, char* report,
                         size_t report_len) {
    if (!ctx || !plan) return -1;
    if (!d_pos || np <= 0) return bad(ctx, "autotune: needs a catalogue");
    std::string r;
    const int rc = hm::plan_autotune(ctx, plan->impl, d_pos, np, reps, r);
    if (report && report_len > 0) {
        const size_t n = std::min(report_len - 1, r.size());
        memcpy(report, r.data(), n);
        report[n] = 0;
    }
    return rc;
}

int hm_box_plan_get_shapes(const hm_box_plan* plan, unsigned* wide_mask, int* bin_block) {
    if (!plan || !wide_mask || !bin_block) return -1;
    hm::plan_tune_get(plan->impl, wide_mask, bin_block);
    return 0;
}

int hm_box_plan_set_shapes(hm_box_plan* plan, unsigned wide_mask, int bin_block) {
    if (!plan) return -1;
    hm::plan_tune_set(plan->impl, wide_mask, bin_block);
    return 0;
}

int hm_box_plan_triple_mesh(const hm_box_plan* plan, int m[3]) {
    if (!plan || !m) return -1;
    hm::plan_triple_mesh(plan->impl, m);
    return 0;
}

hm_survey_plan* hm_survey_plan_create(hm_ctx* ctx, const hm_survey_config* cfg) {
    if (!ctx) return nullptr;
    if (!cfg) return bad(ctx, "survey plan: config is NULL"), nullptr;
    try {
        hm::SurveyPlanImpl* impl = hm::survey_create(ctx, cfg);
        if (!impl) return nullptr;
        return new hm_survey_plan{impl};
    } catch (const std::exception& e) {
        ctx->set_error(std::string("survey plan: ") + e.what());
        return nullptr;
    }
}

void hm_survey_plan_destroy(hm_ctx* ctx, hm_survey_plan* plan) {
    if (!plan) return;
    if (ctx) (void)hipStreamSynchronize(ctx->stream);
    hm::survey_destroy(plan->impl);
    delete plan;
}

int hm_survey_plan_execute(hm_ctx* ctx, hm_survey_plan* plan, const double* d_pos_d, const double* d_w_d, long long n_d,
                           const double* d_pos_r, const double* d_w_r, long long n_r, hm_survey_info* info, double* pk,
                           long long* nmode) {
    if (!ctx) return -1;
    if (!plan || !d_pos_d || !d_w_d || !d_pos_r || !d_w_r) return bad(ctx, "survey execute: NULL argument");
    try {
        return hm::survey_execute(ctx, plan->impl, d_pos_d, d_w_d, n_d, d_pos_r, d_w_r, n_r, info, pk, nmode);
    } catch (const std::exception& e) {
        ctx->set_error(std::string("survey execute: ") + e.what());
        return -1;
    }
}

size_t hm_survey_plan_bytes(const hm_survey_plan* plan) { return plan ? hm::survey_bytes(plan->impl) : 0; }

int hm_recon_displace(hm_ctx* ctx, const hm_recon_config* cfg, double* d_pos_d, const double* d_w_d, long long n_d,
                      double* d_pos_r, const double* d_w_r, long long n_r, double alpha, double norm) {
    if (!ctx) return -1;
    if (!cfg) return bad(ctx, "reconstruction: config is NULL");
    if ((n_d > 0 && !d_pos_d) || (n_r > 0 && !d_pos_r) || n_d < 0 || n_r < 0)
        return bad(ctx, "reconstruction: bad particle arrays");
    hm::AssignGeom g;
    if (!mesh_ok(ctx, &cfg->mesh, g)) return -1;
    for (int a = 0; a < 3; a++)
        if (!hm::fft_length_supported(g.n[a])) return bad(ctx, "reconstruction: unsupported n_mesh_recon");
    hm::Mesh m;
    for (int a = 0; a < 3; a++) {
        m.n[a] = g.n[a];
        m.L[a] = g.L[a];
    }
    m.assign = g.assign;
    hm::ReconWork w;
    int rc;
    try {
        rc = cfg->precision == HM_FP64
                 ? hm::recon_displace<double>(ctx, m, cfg->b1, cfg->RG, d_pos_d, d_w_d, n_d, d_pos_r, d_w_r, n_r, alpha,
                                              norm, w)
                 : hm::recon_displace<float>(ctx, m, cfg->b1, cfg->RG, d_pos_d, d_w_d, n_d, d_pos_r, d_w_r, n_r, alpha,
                                             norm, w);
        if (rc == 0 && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = bad(ctx, "reconstruction: stream error");
    } catch (const std::exception& e) {
        rc = bad(ctx, e.what());
    }
    (void)hipStreamSynchronize(ctx->stream);
    hm::recon_work_free(w);
    return rc;
}

int hm_periodic_wrap(hm_ctx* ctx, double* d_pos, long long n, const double boxsize[3]) {
    if (!ctx) return -1;
    if (!boxsize || (n > 0 && !d_pos)) return bad(ctx, "periodic wrap: NULL argument");
    return hm::periodic_wrap(ctx, d_pos, n, boxsize);
}

double hm_sph_bessel_jl(int l, double x) { return hm::sph_jl(l, x); }

double hm_wigner_3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    return hm::wigner_3j(j1, j2, j3, m1, m2, m3);
}

void hm_ylm_reduced(int l, int m, const double v[3], double out[2]) {
    const hm::Cplx y = hm::ylm_red(l, m, v[0], v[1], v[2]);
    out[0] = y.re;
    out[1] = y.im;
}

}  // extern "C"
