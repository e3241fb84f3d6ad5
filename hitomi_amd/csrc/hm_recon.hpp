// hm_recon.hpp -- reconstruction displacement and periodic wrap, shared by
// the box and survey plans and the C ABI.
#pragma once

#include "hm_assign.hpp"

namespace hm {

struct ReconWork {
    DevBuf pos, wts, fre, grid, psi, tab, part;
    SortWork sort;
};
void recon_work_free(ReconWork& w);
size_t recon_work_bytes(const ReconWork& w);

// Displace D and R in place by the TSC-gathered norm * Re Psi of the field
// D - alpha R (weights w, 1 when NULL) on mesh m (particle_reconstruction.hpp).
template <typename T>
int recon_displace(hm_ctx* ctx, const Mesh& m, double b1, double RG, double* pD, const double* wD, long long nD,
                   double* pR, const double* wR, long long nR, double alpha, double norm, ReconWork& w);

// The plain reconstruction's field on mesh m for a catalogue with unit
// weights: out [ng][4] FP32 = norm * Re Psi_a (a = 0, 1, 2) and the assigned
// density contrast rho / rho_bar - 1.
template <typename T>
int recon_features(hm_ctx* ctx, const Mesh& m, double b1, double RG, const double* pD, long long nD, double norm,
                   float* out, ReconWork& w);

// D then R positions in one array with weights w (1 when NULL) for D and
// -alpha w for R (complex, imaginary part 0).
int pack_two(hm_ctx* ctx, const double* pD, const double* wD, long long nD, const double* pR, const double* wR,
             long long nR, double alpha, double* pos, double* w);

// calcPeriodicBoundary (particleGadget.hpp:565-591) on n positions.
int periodic_wrap(hm_ctx* ctx, double* pos, long long n, const double* L);

}  // namespace hm
