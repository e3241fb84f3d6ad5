// hm_internal.hpp -- context, workspace and launch plumbing shared by the
// translation units of libhitomi_amd.so.  Nothing here crosses the C ABI.
#pragma once

#include "hm_common.hpp"
#include "hm_fft.hpp"
#include "../../include/hitomi_amd.h"

#include <cstring>
#include <functional>
#include <memory>

struct hm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // twiddle tables exp(-2 pi i e / L), e < L, keyed by (L, sizeof(T))
    std::map<std::pair<int, int>, void*> twiddles;
    // kernel timing (hipEvents on the launch stream), off by default
    bool timing = false;
    std::string timing_only;  // empty: time every launch
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, long long>> kstats;  // name -> (ms total, launches)
    // each timed launch's [start, end) in ms after tref (recorded by
    // hm_timing_reset): launches on other streams that overlap count once in
    // the busy time
    hipEvent_t tref = nullptr;
    std::map<std::string, std::vector<std::pair<double, double>>> kspans;
    bool capturing = false;
    // launch shapes chosen by an autotuner: bit b of wide_mask set = the wide
    // column tile for the b-th tunable length (256; 270 or 288; 512), bit 3 the
    // wide tile of the triple stage's shell x pass (270 or 288) on its own
    unsigned wide_mask = 0;
    size_t lds_max = 65536;  // dynamic LDS per workgroup the device allows
    int cu_count = 256;      // compute units of the device
    bool conv_lds = true;    // convolutions of the learned reconstruction with LDS-staged input tiles
    bool wide_cols(int L, const char* pass = nullptr) const {
        int b = L == 256 ? 0 : ((L == 288 || L == 270) ? 1 : (L == 512 ? 2 : -1));
        if (b == 1 && pass && std::strcmp(pass, "tri_pass_x") == 0) b = 3;
        // the two self-conjugate planes fill 128 narrow tiles, too few to halve
        if (pass && std::strcmp(pass, "xi_planes_x") == 0) return false;
        return b >= 0 && ((wide_mask >> b) & 1u);
    }
    // device workspace of the mock generator, kept between calls
    void* mock_ws = nullptr;
    size_t mock_ws_bytes = 0;
    // j_l spline nodes (y then c, JL_NODES doubles each) of the seam entry
    // points, keyed by l, kept until the context ends
    std::map<int, void*> jl_splines;

    void set_error(const std::string& m) { err = m; }
    hipEvent_t get_event();
    template <typename T> const hm::C<T>* twiddle(int L);
    // Launch wrapper: records HIP events around the launch when timing is on.
    int launch(const char* name, const std::function<void()>& fn);
};

namespace hm {

// Device allocation owned by the library.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};
int dev_alloc(hm_ctx* ctx, DevBuf& b, size_t bytes);
void dev_free(DevBuf& b);

// Zero-fill with a kernel (graph-capture safe; no hipMemsetAsync).
int fill_zero(hm_ctx* ctx, void* p, size_t bytes);

// 3D transforms on row-major grids (FFTW conventions, unnormalised).
template <typename T> int fft3d_c2c(hm_ctx* ctx, C<T>* data, const int n[3], int dir);
template <typename T> int fft3d_r2c(hm_ctx* ctx, const T* in, C<T>* out, const int n[3]);
template <typename T> int fft3d_c2r(hm_ctx* ctx, C<T>* in, T* out, const int n[3]);

// Column / row passes with a plain global source.
template <typename T> int fft_cols(hm_ctx* ctx, const char* name, int L, int dir, const C<T>* src, C<T>* dst,
                                  const ColGeom& g);
template <typename T> int fft_rows(hm_ctx* ctx, const char* name, int L, int dir, const C<T>* src, C<T>* dst,
                                  const RowGeom& g);
template <typename T> int fft_c2r_rows(hm_ctx* ctx, const char* name, int L2, const C<T>* src, T* dst,
                                      long long n_rows, int point_s);

bool fft_length_supported(int L);
int fft_length_at_least(int n);  // smallest supported length >= n, or -1

// Survey-geometry catalogues handed to a box plan for B and zeta with ELL = 0:
// fields with weights wD, c1 * wR and wD^2, c2 * wR^2 (null weights: 1); abs1
// and abs2 bound sum |weight| of the two fields (fixed-point scale); sn3 the
// shot noise (sum_D w^3 - alpha^3 sum_R w^3, or sum_R w^2 for the window);
// norm the estimator normalisation Vsurvey^2 / (sum w)^3.
struct BoxWeighted {
    const double* wD;
    const double* wR;
    double c1, c2, abs1, abs2, sn3, norm;
    // B with ELL = L > 0 (bispec.hpp:4-353): the unshifted positions D then R
    // whose directions are the lines of sight, and per M = -L..L the shot noise
    // sum_D w^3 Y_LM - alpha^3 sum_R w^3 Y_LM (re, im)
    int L = 0;
    const double* losD = nullptr;
    const double* losR = nullptr;
    const double* snLM = nullptr;
};
struct BoxPlanImpl;
BoxPlanImpl* plan_create(hm_ctx* ctx, const hm_box_config* cfg);
void plan_destroy(BoxPlanImpl* p);
size_t plan_bytes(const BoxPlanImpl* p);
int plan_execute_weighted(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos_d, long long n_d, const double* d_pos_r,
                          long long n_r, const BoxWeighted* wt, hm_box_result* res);

// Batched execution.  The per-catalogue scalars live in device memory
// (DYN_LEN doubles), written by the assignment stage and read by the mesh
// stage, so that the mesh stage issues the same launches for every catalogue
// size and can be captured once into a graph and replayed.
enum { DYN_INV_SCALE = 0, DYN_NBAR, DYN_SN2, DYN_SN3, DYN_DC, DYN_N, DYN_LEN = 8 };
struct ExecOpts {
    const double* dyn = nullptr;  // device scalars
    bool assign = true;           // run the assignment (and write dyn)
    bool mesh = true;             // run the transforms, binning and triple sums
    bool readback = true;         // copy the results to the host (plan_execute_opts ignores res)
};
int plan_execute_opts(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos, long long np, const ExecOpts& xo);
std::string plan_batch_check(const BoxPlanImpl* p);  // empty: the plan can run batched
void plan_vector_dims(const BoxPlanImpl* p, int* n_pk, int* n_bk);
int plan_batch_prepare(hm_ctx* ctx, BoxPlanImpl* p);  // device tables of the finalisation (no capture)
// the data vector [P_l(k_b)], then [B_000(k_a, k_b) - SN], a <= b, into out (device)
int plan_finalize_device(hm_ctx* ctx, BoxPlanImpl* p, const double* dyn, double* out);

// Kernel timings of src (another context on the same device) added to dst's.
int timing_merge(hm_ctx* dst, hm_ctx* src);

// Autotuner of the launch shapes (wide column tiles, binning block size).
int plan_autotune(hm_ctx* ctx, BoxPlanImpl* p, const double* d_pos, long long np, int reps, std::string& report);
void plan_tune_get(const BoxPlanImpl* p, unsigned* wide, int* bin);
void plan_tune_set(BoxPlanImpl* p, unsigned wide, int bin);

}  // namespace hm
