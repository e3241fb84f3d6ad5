"""The reference's per-field seams, one call each, on device grids.

`DensityField` mirrors DensityFieldClass (field.hpp:4-718) and
`TwoPointStatistics` mirrors TwoPointStatisticsClass (field.hpp:721-1310):
same member names, same argument meaning, the grid is the reference's
complex FP64 [n0][n1][n2] held in device memory.  Each member is one
extern "C" entry point of libhitomi_amd.so (hm_grid_alloc, hm_assign_field,
hm_fourier_transform, hm_window_correct, hm_shell_c2r, hm_bessel_c2r,
hm_bin_pk, hm_bin_xi, hm_bin_xi_3pcf, hm_sn_field_ijk); the bispectrum sums
(bispec.hpp:1036-1056, 1159-1168) and the particle sums (field.hpp:1164-1250,
449-503) are module functions.  Errors raise HitomiError with the library's
message, where the reference prints and returns -1.
"""
import ctypes as ct

import numpy as np

from . import ASSIGN, Context, HitomiError, Mesh, make_mesh


def _cd(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ct.POINTER(ct.c_double))


def _sn(sn):
    sn = complex(sn)
    return (ct.c_double * 2)(sn.real, sn.imag)


class DensityField:
    """DensityFieldClass: owns one zeroed complex FP64 grid on the device (field.hpp:20-31)."""

    def __init__(self, ctx: Context, n_mesh, boxsize, assign="TSC"):
        self.ctx = ctx
        self.mesh: Mesh = make_mesh(n_mesh, boxsize, assign)
        self.shape = tuple(int(self.mesh.n_mesh[a]) for a in range(3))
        self.ptr = ctx.lib.hm_grid_alloc(ctx.h, ct.byref(self.mesh))
        if not self.ptr:
            ctx.raise_error("hm_grid_alloc")

    # field.hpp:35-45
    def finalizeDensityField(self):
        if self.ptr:
            self.ctx.lib.hm_grid_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.finalizeDensityField()
        except Exception:
            pass

    def _live(self):
        if not self.ptr:
            raise HitomiError("the density field has been finalised")
        return self.ptr

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != self.shape:
            raise ValueError(f"grid shape {a.shape} != {self.shape}")
        self.ctx.check(self.ctx.lib.hm_memcpy_h2d(self.ctx.h, self._live(), a.ctypes.data, a.nbytes), "h2d")
        return self

    def download(self):
        out = np.empty(self.shape, dtype=np.complex128)
        self.ctx.check(self.ctx.lib.hm_memcpy_d2h(self.ctx.h, out.ctypes.data, self._live(), out.nbytes), "d2h")
        return out

    # field.hpp:47-184 (weight: complex per particle, or None for 1)
    def calcField(self, pos, weight=None, wabs_sum=0.0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        dpos = self.ctx.array(pos)
        dw = None
        if weight is not None:
            w = np.asarray(weight, dtype=np.complex128).reshape(-1)
            if w.size != pos.shape[0]:
                raise ValueError("one weight per particle")
            dw = self.ctx.array(np.stack([w.real, w.imag], axis=1))
            if not wabs_sum:
                wabs_sum = float(np.abs(w.real).sum() + np.abs(w.imag).sum())
        self.ctx.check(self.ctx.lib.hm_assign_field(self.ctx.h, ct.byref(self.mesh), dpos.ptr, dw.ptr if dw else None,
                                                    pos.shape[0], float(wabs_sum), self._live()), "hm_assign_field")
        return 0

    # field.hpp:310-326
    def calcFourierTransform(self):
        self.ctx.check(self.ctx.lib.hm_fourier_transform(self.ctx.h, ct.byref(self.mesh), self._live(), -1),
                       "hm_fourier_transform")
        return 0

    # field.hpp:328-344
    def calcInverseFourierTransform(self):
        self.ctx.check(self.ctx.lib.hm_fourier_transform(self.ctx.h, ct.byref(self.mesh), self._live(), 1),
                       "hm_fourier_transform")
        return 0

    # field.hpp:610-639
    def calcAssignmentFunctionCorrection(self):
        self.ctx.check(self.ctx.lib.hm_window_correct(self.ctx.h, ct.byref(self.mesh), self._live()),
                       "hm_window_correct")
        return 0

    # field.hpp:346-401; the reference takes a precomputed Y table, here (ell, m) of Y_red(k)
    def calcInverseFourierTransformForBispectrum(self, density: "DensityField", kmag_in, dk_in, ell=0, m=0):
        n = ct.c_longlong()
        self.ctx.check(self.ctx.lib.hm_shell_c2r(self.ctx.h, ct.byref(self.mesh), density._live(), float(kmag_in),
                                                 float(dk_in), int(ell), int(m), self._live(), ct.byref(n)),
                       "hm_shell_c2r")
        return n.value

    # field.hpp:403-445; Y_red(ell, m)(k) and the j_ell_j spline (sj of the reference)
    def calcInverseFourierTransformForThreePointFunction(self, density: "DensityField", rmag_in, ell=0, m=0,
                                                         ell_j=None):
        ell_j = ell if ell_j is None else ell_j
        self.ctx.check(self.ctx.lib.hm_bessel_c2r(self.ctx.h, ct.byref(self.mesh), density._live(), float(rmag_in),
                                                  int(ell_j), int(ell), int(m), self._live()), "hm_bessel_c2r")
        return 0


class TwoPointStatistics:
    """TwoPointStatisticsClass: results in pk / n_mode_pk and xi / n_mode_xi (field.hpp:729-734)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.pk = self.n_mode_pk = self.xi = self.n_mode_xi = None

    def calcPowerSpectrum(self, density1: DensityField, density2: DensityField, kbin, shotnoise, ELL, M):
        kb, kp = _cd(kbin)
        out = np.zeros(2 * kb.size)
        nm = np.zeros(kb.size, dtype=np.int64)
        self.ctx.check(self.ctx.lib.hm_bin_pk(self.ctx.h, ct.byref(density1.mesh), density1._live(), density2._live(),
                                              kp, kb.size, _sn(shotnoise), int(ELL), int(M),
                                              out.ctypes.data_as(ct.POINTER(ct.c_double)),
                                              nm.ctypes.data_as(ct.POINTER(ct.c_longlong))), "hm_bin_pk")
        self.pk = out[0::2] + 1j * out[1::2]
        self.n_mode_pk = nm
        return 0

    def _xi(self, fn, density1, density2, rbin, shotnoise, extra):
        rb, rp = _cd(rbin)
        out = np.zeros(2 * rb.size)
        nm = np.zeros(rb.size, dtype=np.int64)
        self.ctx.check(fn(self.ctx.h, ct.byref(density1.mesh), density1._live(), density2._live(), rp, rb.size,
                          _sn(shotnoise), *extra, out.ctypes.data_as(ct.POINTER(ct.c_double)),
                          nm.ctypes.data_as(ct.POINTER(ct.c_longlong))), fn.__name__)
        self.xi = out[0::2] + 1j * out[1::2]
        self.n_mode_xi = nm
        return 0

    def calcCorrelationFunction(self, density1, density2, rbin, shotnoise, ELL, M):
        return self._xi(self.ctx.lib.hm_bin_xi, density1, density2, rbin, shotnoise, (int(ELL), int(M)))

    # field.hpp:979-1113; (ell1, m1), (ell2, m2) name the reference's Ylm1, Ylm2 tables
    def calcCorrelationFunctionForThreePointFunction(self, density1, density2, rbin, shotnoise, ell1, m1, ell2, m2):
        return self._xi(self.ctx.lib.hm_bin_xi_3pcf, density1, density2, rbin, shotnoise,
                        (int(ell1), int(m1), int(ell2), int(m2)))

    # field.hpp:1116-1160
    def calcShotNoiseForBispectrum_ijk(self, density1, density2, shotnoise, xi: DensityField):
        self.ctx.check(self.ctx.lib.hm_sn_field_ijk(self.ctx.h, ct.byref(density1.mesh), density1._live(),
                                                    density2._live(), _sn(shotnoise), xi._live()), "hm_sn_field_ijk")
        return 0


def triple_sums(ctx: Context, fields, pairs, f3: DensityField):
    """bispec.hpp:1159-1168 for every (a, b) in pairs: sum_x dV F_a F_b F3 (complex array)."""
    ptrs = (ct.c_void_p * len(fields))(*[f._live() for f in fields])
    pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
    out = np.zeros(2 * pr.shape[0])
    ctx.check(ctx.lib.hm_triple_sums(ctx.h, ct.byref(f3.mesh), ptrs, len(fields),
                                     pr.ctypes.data_as(ct.POINTER(ct.c_int)), pr.shape[0], f3._live(),
                                     out.ctypes.data_as(ct.POINTER(ct.c_double))), "hm_triple_sums")
    return out[0::2] + 1j * out[1::2]


def sn_bessel_sweep(ctx: Context, xi: DensityField, ell1, m1, ell2, m2, k1, k2):
    """bispec.hpp:1036-1056 for every (k1[p], k2[p]) (complex array, i^(l1+l2) dV applied)."""
    a1, p1 = _cd(np.broadcast_to(np.asarray(k1, dtype=np.float64), np.shape(k2) or (1,)).copy())
    a2, p2 = _cd(np.broadcast_to(np.asarray(k2, dtype=np.float64), a1.shape).copy())
    out = np.zeros(2 * a1.size)
    ctx.check(ctx.lib.hm_sn_bessel_sweep(ctx.h, ct.byref(xi.mesh), xi._live(), int(ell1), int(m1), int(ell2), int(m2),
                                         p1, p2, a1.size, out.ctypes.data_as(ct.POINTER(ct.c_double))),
              "hm_sn_bessel_sweep")
    return out[0::2] + 1j * out[1::2]


def particle_sums(ctx: Context, los, w, ell, m, power):
    """field.hpp:1164-1240: sum_p w_p^power Y_red(ell, m)(los_p) (complex)."""
    dl = ctx.array(np.ascontiguousarray(los, dtype=np.float64).reshape(-1, 3)) if los is not None else None
    dw = ctx.array(np.ascontiguousarray(w, dtype=np.float64)) if w is not None else None
    n = dl.shape[0] if dl is not None else dw.shape[0]
    out = (ct.c_double * 2)()
    ctx.check(ctx.lib.hm_particle_sums(ctx.h, dl.ptr if dl else None, dw.ptr if dw else None, n, int(ell), int(m),
                                       int(power), out), "hm_particle_sums")
    return complex(out[0], out[1])


def survey_volume(ctx: Context, n_mesh, boxsize, pos, w=None, assign="TSC"):
    """calcSurveyVolume (field.hpp:449-503)."""
    mesh = make_mesh(n_mesh, boxsize, assign)
    dp = ctx.array(np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3))
    dw = ctx.array(np.ascontiguousarray(w, dtype=np.float64)) if w is not None else None
    v = ct.c_double()
    ctx.check(ctx.lib.hm_survey_volume(ctx.h, ct.byref(mesh), dp.ptr, dw.ptr if dw else None, dp.shape[0],
                                       ct.byref(v)), "hm_survey_volume")
    return v.value


__all__ = ["DensityField", "TwoPointStatistics", "triple_sums", "sn_bessel_sweep", "particle_sums", "survey_volume",
           "ASSIGN"]
