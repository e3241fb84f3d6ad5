// hm_assign.hpp -- NGP/CIC/TSC mass assignment with 64-bit fixed-point cell
// sums (DensityFieldClass::calcField, field.hpp:47-184).
//
// Each particle adds round(w * Wx*Wy*Wz * 2^s) to int64 cells with device
// atomics.  Integer addition is associative, so the grid is bitwise
// identical for any particle order or schedule; 2^s is the largest power of
// two with sum|w| * 2^s < 2^62, so no cell can overflow.
#pragma once

#include "hm_internal.hpp"

namespace hm {

struct AssignGeom {
    int n[3];
    double L[3];
    long long ng;
    int assign;
};

// Stencil of one axis, FP64 exactly as field.hpp:84-86, 121-126, 162-169
// (no contraction: the reference is built for x86-64 without FMA).
__device__ __forceinline__ int stencil_axis(int assign, double xp, int* iw0, double* w) {
#pragma clang fp contract(off)
    if (assign == 0) {
        *iw0 = (int)(xp + 0.5);
        w[0] = 1.0;
        return 1;
    } else if (assign == 1) {
        const int i0 = (int)(xp);
        const double dx = xp - (double)i0;
        w[0] = 1.0 - dx;
        w[1] = dx;
        *iw0 = i0;
        return 2;
    } else {
        const int i1 = (int)(xp + 0.5);
        const double dx = xp - (double)i1;
        w[0] = 0.5 * (0.5 - dx) * (0.5 - dx);
        w[1] = 0.75 - dx * dx;
        w[2] = 0.5 * (0.5 + dx) * (0.5 + dx);
        *iw0 = i1 - 1;
        return 3;
    }
}

__device__ __forceinline__ double axis_xp(double pos, int n, double L) {
#pragma clang fp contract(off)
    return pos * (double)n / L;  // field.hpp:162
}

// Fixed-point scale 2^s for a weight bound.
double fixed_scale_for(double wabs_sum);

// Real source for the r2c z-pass reading the fixed-point grid:
// value = ((q * 2^-s) * (1/dV) - nbar) * dV, i.e. the box density fluctuation
// of field.hpp:505-520 times dV as calcFourierTransform applies it
// (field.hpp:313-318).
template <typename T> struct FixedRealSrc {
    const long long* g;
    int L;  // half row length (complex pairs per row)
    double inv_scale, factor, nbar, dV;
    const double* dyn = nullptr;  // set: inv_scale and nbar read from device memory (dyn[0], dyn[1])
    __device__ inline C<T> operator()(long long r, int m) const {
#pragma clang fp contract(off)
        const long long* q = g + r * (2 * (long long)L) + 2 * m;
        const double is = dyn ? dyn[0] : inv_scale, nb = dyn ? dyn[1] : nbar;
        const double a = (((double)q[0] * is) * factor - nb) * dV;
        const double b = (((double)q[1] * is) * factor - nb) * dV;
        return mk<T>((T)a, (T)b);
    }
};

// fixed point -> complex grid times dV (calcFourierTransform, field.hpp:313-318)
template <typename T>
__global__ __launch_bounds__(256) void k_fixed_to_complex(const long long* __restrict__ re,
                                                          const long long* __restrict__ im, long long n,
                                                          double inv_scale, double factor, double dV,
                                                          C<T>* __restrict__ out) {
#pragma clang fp contract(off)
    const long long st = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += st) {
        const double a = (((double)re[i] * inv_scale) * factor) * dV;
        const double b = im ? (((double)im[i] * inv_scale) * factor) * dV : 0.0;
        out[i] = mk<T>((T)a, (T)b);
    }
}

// Accumulate particles into a zeroed fixed-point grid (re and optional im)
// with one 64-bit atomic per stencil cell.
int assign_fixed(hm_ctx* ctx, const AssignGeom& g, const double* d_pos, const double* d_w, long long np,
                 double scale, long long* d_re, long long* d_im);

// Sort-by-cell assignment: particles whose whole stencil lies inside the
// mesh are counting-sorted by the 64-cell z chunk of their first stencil cell
// and each output tile accumulates the particles that reach it in LDS
// (integer atomics there, one plain store per cell in HBM); the few particles whose stencil leaves the mesh (the Q1
// spill/drop cases) are added afterwards with atomics.  Each contribution is
// rounded to the same fixed-point value as in assign_fixed, so the grid is
// bitwise identical to it.
struct SortWork {
    DevBuf start;   // int32 [keys + 1]: particle count, then exclusive prefix, per (x, y, z/64) chunk
    DevBuf slot;    // int32 [np]: rank within the cell, then (after the scatter) the particle's place in cell
                    // order; -1 for a border particle
    DevBuf spos;    // double [3 np] positions in cell order
    DevBuf sw;      // double [2 np] weights in cell order
    DevBuf border;  // int32 [np + 1]: border count, then border particle indices
    DevBuf bsum;    // int32 scan block sums
    // two-level counting sort (no global atomics): per-particle keys, the particles in bucket order
    // (32-byte (x, y, z, w_re) records in tpos, 16-byte (index, local key, w_im) records in tw), the
    // chunk histograms of both levels and the per-bucket chunk starts; left empty (the one-level sort
    // runs instead) by sort_work_carve and for meshes whose rows hold more keys than one bucket's LDS
    // histogram
    DevBuf key, tpos, tw, h1, h2, cst;
    // the last full sort was the two-level one: slot[p] is the place of the particle at level-1
    // position p (the index in its tw record), for p below the interior count
    bool two_level = false;
    // what the last full sort was made for: a reuse must name the same positions, count, mesh and weights
    struct Sig {
        bool valid = false;
        const double* pos = nullptr;
        long long np = 0;
        int n[3] = {0, 0, 0};
        double L[3] = {0.0, 0.0, 0.0};
        int assign = -1;
        bool weights = false;
    } sig;
    size_t bytes() const {
        return start.bytes + slot.bytes + spos.bytes + sw.bytes + border.bytes + bsum.bytes + key.bytes + tpos.bytes +
               tw.bytes + h1.bytes + h2.bytes + cst.bytes;
    }
};
int sort_work_alloc(hm_ctx* ctx, SortWork& w, const AssignGeom& g, long long np, bool weights);
void sort_work_free(SortWork& w);
// Bytes of a SortWork for np particles on g carved out of one region, each part
// 256-byte aligned.  With base set, w's buffers point into [base, base + bytes):
// they are borrowed, never passed to sort_work_free.
size_t sort_work_carve(const AssignGeom& g, long long np, bool weights, char* base, SortWork* w);
// Fixed-point cells -> density where the int64 cells and the FP64 output share memory.
int fixed_to_density_inplace(hm_ctx* ctx, long long* d_q, long long n, double inv_scale, double factor);
// reuse: the previous call on w sorted these same positions (same pointer,
// count and geometry, contents unchanged, weights given both times): only the
// weights are moved into cell order before the gather.  A reuse whose
// positions, count, mesh or weights differ from the last sort's is an error.
int assign_sorted(hm_ctx* ctx, const AssignGeom& g, const double* d_pos, const double* d_w, long long np,
                  double scale, long long* d_re, long long* d_im, SortWork& w, bool reuse = false);

}  // namespace hm
