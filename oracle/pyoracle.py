"""ctypes view of liboracle.so, the FP64 CPU restatement of the reference
(TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product)."""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
NGP, CIC, TSC = 0, 1, 2


class Param(ct.Structure):
    _fields_ = [
        ("n_mesh", ct.c_int * 3), ("boxsize", ct.c_double * 3), ("volume", ct.c_double),
        ("n_mesh_tot", ct.c_longlong), ("assign", ct.c_int), ("ell1", ct.c_int), ("ell2", ct.c_int),
        ("ELL", ct.c_int), ("n_kbin", ct.c_int), ("n_rbin", ct.c_int), ("ith_kbin", ct.c_int),
        ("ith_rbin", ct.c_int), ("kmin", ct.c_double), ("kmax", ct.c_double), ("rmin", ct.c_double),
        ("rmax", ct.c_double),
    ]


_lib = None
_D = ct.POINTER(ct.c_double)


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        P = ct.POINTER(Param)
        sig = {
            "or_param_init": (None, [P, ct.c_int, ct.c_int, ct.c_int, ct.c_double, ct.c_double, ct.c_double, ct.c_int]),
            "or_set_kbin": (None, [P, _D]),
            "or_set_rbin": (None, [P, _D]),
            "or_calc_field": (ct.c_int, [P, _D, _D, ct.c_longlong, _D]),
            "or_cell_coords": (ct.c_int, [P, _D, ct.c_longlong, ct.POINTER(ct.c_longlong)]),
            "or_mode_counts": (ct.c_int, [P, _D, ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_longlong)]),
            "or_fft3d": (ct.c_int, [P, _D, ct.c_int]),
            "or_fft1d": (ct.c_int, [_D, ct.c_int, ct.c_int]),
            "or_fourier_transform": (None, [P, _D]),
            "or_inverse_fourier_transform": (None, [P, _D]),
            "or_assignment_correction": (None, [P, _D]),
            "or_assignment_function": (ct.c_double, [P, _D]),
            "or_shotnoise_function": (ct.c_double, [P, _D]),
            "or_ylm": (None, [ct.c_int, ct.c_int, _D, _D]),
            "or_sphPlm": (ct.c_double, [ct.c_int, ct.c_int, ct.c_double]),
            "or_wigner3j": (ct.c_double, [ct.c_int] * 6),
            "or_bessel_jl": (ct.c_double, [ct.c_int, ct.c_double]),
            "or_bessel_new": (ct.c_void_p, [ct.c_int]),
            "or_bessel_eval": (ct.c_double, [ct.c_void_p, ct.c_double]),
            "or_bessel_free": (None, [ct.c_void_p]),
            "or_calc_power_spectrum": (ct.c_int, [P, _D, _D, _D, ct.c_double, ct.c_double, ct.c_int, ct.c_int, _D,
                                                  ct.POINTER(ct.c_int)]),
            "or_calc_correlation_function": (ct.c_int, [P, _D, _D, _D, ct.c_double, ct.c_double, ct.c_int, ct.c_int,
                                                        _D, ct.POINTER(ct.c_int)]),
            "or_ifft_for_bispectrum": (ct.c_int, [P, _D, ct.c_double, ct.c_double, _D, _D]),
            "or_box_density_fluctuation": (ct.c_int, [P, _D, ct.c_longlong, _D]),
            "or_box_density": (ct.c_int, [P, _D, ct.c_longlong, _D]),
            "or_box_pk": (ct.c_int, [P, _D, ct.c_longlong, _D, _D, ct.POINTER(ct.c_int)]),
            "or_box_xi": (ct.c_int, [P, _D, ct.c_longlong, _D, _D]),
            "or_box_bk": (ct.c_int, [P, _D, ct.c_longlong, _D, _D, _D, _D]),
            "or_survey_prepare": (ct.c_int, [P, _D, _D, ct.c_longlong, _D, _D, ct.c_longlong, _D, _D, _D]),
            "or_ylm_weighted_fluctuation": (ct.c_int, [P, _D, _D, _D, ct.c_longlong, _D, _D, _D, ct.c_longlong,
                                                       ct.c_double, ct.c_int, ct.c_int, _D]),
            "or_shotnoise_pk": (None, [_D, _D, ct.c_longlong, _D, _D, ct.c_longlong, ct.c_double, ct.c_int, ct.c_int,
                                       _D]),
            "or_survey_pk": (ct.c_int, [P, _D, _D, _D, ct.c_longlong, _D, _D, _D, ct.c_longlong, ct.c_double,
                                        ct.c_double, _D, _D, ct.POINTER(ct.c_int)]),
            "or_recon_psi": (None, [P, _D, ct.c_double, ct.c_double, _D]),
            "or_recon_gather": (None, [P, _D, ct.c_double, _D, ct.c_longlong]),
            "or_recon_displace": (ct.c_int, [P, _D, _D, ct.c_longlong, _D, _D, ct.c_longlong, ct.c_double, ct.c_double,
                                             ct.c_double, ct.c_double]),
            "or_box_recon_2pt": (ct.c_int, [P, _D, ct.c_longlong, _D, ct.c_longlong, ct.c_double, _D, ct.c_int, _D,
                                            ct.POINTER(ct.c_int)]),
            "or_box_zeta": (ct.c_int, [P, _D, ct.c_longlong, _D, _D, _D, _D]),
            "or_box_recon_bk": (ct.c_int, [P, _D, ct.c_longlong, _D, ct.c_longlong, ct.c_double, _D, _D, _D, _D]),
            "or_box_recon_zeta": (ct.c_int, [P, _D, ct.c_longlong, _D, ct.c_longlong, ct.c_double, _D, _D, _D, _D]),
            "or_box_interlaced_k": (ct.c_int, [P, _D, ct.c_longlong, _D]),
            "or_box_interlaced_2pt": (ct.c_int, [P, _D, ct.c_longlong, _D, ct.c_int, _D, ct.POINTER(ct.c_int)]),
            "or_survey_bk": (ct.c_int, [P, _D, _D, _D, ct.c_longlong, _D, _D, _D, ct.c_longlong, ct.c_double,
                                        ct.c_double, _D, _D, _D, _D]),
            "or_survey_zeta": (ct.c_int, [P, _D, _D, _D, ct.c_longlong, _D, _D, _D, ct.c_longlong, ct.c_double,
                                          ct.c_double, _D, ct.c_int, _D, _D, _D]),
            "or_survey_xi": (ct.c_int, [P, _D, _D, _D, ct.c_longlong, _D, _D, _D, ct.c_longlong, ct.c_double,
                                        ct.c_double, _D, ct.c_int, _D, ct.POINTER(ct.c_int)]),
            "or_philox4x32_10": (None, [ct.POINTER(ct.c_uint), ct.POINTER(ct.c_uint), ct.POINTER(ct.c_uint)]),
            "or_mock_white_noise": (None, [ct.c_longlong, ct.c_ulonglong, _D]),
            "or_mock_sites": (None, [ct.c_longlong, ct.c_longlong, ct.c_ulonglong, ct.POINTER(ct.c_longlong)]),
            "or_mock_select_uniform": (None, [ct.c_longlong, ct.c_ulonglong, _D]),
            "or_sky_randoms": (None, [_D, ct.c_ulonglong, ct.c_longlong, _D]),
            "or_shotnoise_bispectrum_ijk": (ct.c_int, [P, _D, _D, ct.c_double, ct.c_double, _D]),
            "or_store_ylm": (None, [P, ct.c_int, ct.c_int, ct.c_int, _D]),
            "or_ifft_for_3pcf": (ct.c_int, [P, _D, ct.c_double, _D, ct.c_int, _D]),
            "or_corr_for_3pcf": (ct.c_int, [P, _D, _D, _D, ct.c_double, ct.c_double, _D, _D, _D,
                                            ct.POINTER(ct.c_int)]),
            "or_sn_bessel_sweep": (ct.c_int, [P, _D, _D, _D, ct.c_int, ct.c_int, ct.c_double, ct.c_double, _D]),
            "or_triple_sum": (None, [P, _D, _D, _D, _D]),
            "or_particle_sum": (None, [_D, _D, ct.c_longlong, ct.c_int, ct.c_int, ct.c_int, _D]),
            "or_survey_volume": (ct.c_double, [P, _D, _D, ct.c_longlong]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def dp(a):
    return a.ctypes.data_as(_D)


def param(n, L, assign=TSC, kmin=0.01, kmax=0.2, n_kbin=20, ell1=0, ell2=0, ELL=0, ith_kbin=0,
          rmin=30.0, rmax=150.0, n_rbin=25, ith_rbin=0) -> Param:
    n = list(n) if np.ndim(n) else [int(n)] * 3
    L = list(L) if np.ndim(L) else [float(L)] * 3
    p = Param()
    lib().or_param_init(ct.byref(p), n[0], n[1], n[2], L[0], L[1], L[2], assign)
    p.kmin, p.kmax, p.n_kbin = kmin, kmax, n_kbin
    p.rmin, p.rmax, p.n_rbin = rmin, rmax, n_rbin
    p.ell1, p.ell2, p.ELL = ell1, ell2, ELL
    p.ith_kbin, p.ith_rbin = ith_kbin, ith_rbin
    return p


def kbin(p: Param):
    out = np.zeros(p.n_kbin)
    lib().or_set_kbin(ct.byref(p), dp(out))
    return out


def rbin(p: Param):
    out = np.zeros(p.n_rbin)
    lib().or_set_rbin(ct.byref(p), dp(out))
    return out


def shape(p: Param):
    return (p.n_mesh[0], p.n_mesh[1], p.n_mesh[2])


def calc_field(p: Param, pos, weight=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    w = np.zeros((pos.shape[0], 2))
    if weight is None:
        w[:, 0] = 1.0
    else:
        w[:] = np.asarray(weight, dtype=np.float64).reshape(-1, 2)
    f = np.zeros(shape(p) + (2,))
    lib().or_calc_field(ct.byref(p), dp(pos), dp(w), pos.shape[0], dp(f))
    return f[..., 0] + 1j * f[..., 1]


def cell_coords(p: Param, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    s = {NGP: 1, CIC: 8, TSC: 27}[p.assign]
    out = np.zeros((pos.shape[0], s), dtype=np.int64)
    lib().or_cell_coords(ct.byref(p), dp(pos), pos.shape[0], out.ctypes.data_as(ct.POINTER(ct.c_longlong)))
    return out


def mode_counts(p: Param):
    """(P_l bin mode counts, bispectrum shell n_mode) for the bins of p, counts only (or_mode_counts)."""
    kb = kbin(p)
    pk = np.zeros(p.n_kbin, dtype=np.int64)
    sh = np.zeros(p.n_kbin, dtype=np.int64)
    LL = ct.POINTER(ct.c_longlong)
    rc = lib().or_mode_counts(ct.byref(p), dp(kb), pk.ctypes.data_as(LL), sh.ctypes.data_as(LL))
    if rc != 0:
        raise MemoryError("or_mode_counts")
    return pk, sh


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.view(np.float64)


def fft3d(p: Param, a, sign):
    a, v = _c(np.array(a, dtype=np.complex128, copy=True))
    lib().or_fft3d(ct.byref(p), dp(v), int(sign))
    return a


def fft1d(a, sign):
    a, v = _c(np.array(a, dtype=np.complex128, copy=True))
    lib().or_fft1d(dp(v), a.shape[0], int(sign))
    return a


def assignment_function(p: Param, kvec):
    k = np.ascontiguousarray(kvec, dtype=np.float64)
    return lib().or_assignment_function(ct.byref(p), dp(k))


def shotnoise_function(p: Param, kvec):
    k = np.ascontiguousarray(kvec, dtype=np.float64)
    return lib().or_shotnoise_function(ct.byref(p), dp(k))


def ylm(ell, m, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(2)
    lib().or_ylm(ell, m, dp(v), dp(out))
    return complex(out[0], out[1])


def sph_plm(l, m, x):
    return lib().or_sphPlm(l, m, x)


def wigner3j(j1, j2, j3, m1, m2, m3):
    return lib().or_wigner3j(j1, j2, j3, m1, m2, m3)


def bessel_jl(l, x):
    return lib().or_bessel_jl(l, float(x))


class BesselSpline:
    def __init__(self, ell):
        self.h = lib().or_bessel_new(ell)

    def __call__(self, x):
        return lib().or_bessel_eval(self.h, float(x))

    def __del__(self):
        try:
            lib().or_bessel_free(self.h)
        except Exception:
            pass


def box_density_fluctuation(p: Param, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    f = np.zeros(shape(p) + (2,))
    lib().or_box_density_fluctuation(ct.byref(p), dp(pos), pos.shape[0], dp(f))
    return f[..., 0] + 1j * f[..., 1]


def fourier_transform(p: Param, field):
    a, v = _c(np.array(field, dtype=np.complex128, copy=True))
    lib().or_fourier_transform(ct.byref(p), dp(v))
    return a


def assignment_correction(p: Param, fk):
    """field.hpp:610-639: the transformed field divided by the assignment window."""
    a, v = _c(np.array(fk, dtype=np.complex128, copy=True))
    lib().or_assignment_correction(ct.byref(p), dp(v))
    return a


def calc_power_spectrum(p: Param, f1, f2, kb, sn, ell, m):
    a1, v1 = _c(f1)
    a2, v2 = _c(f2)
    kb = np.ascontiguousarray(kb, dtype=np.float64)
    pk = np.zeros(2 * p.n_kbin)
    nm = np.zeros(p.n_kbin, dtype=np.int32)
    lib().or_calc_power_spectrum(ct.byref(p), dp(v1), dp(v2), dp(kb), float(np.real(sn)), float(np.imag(sn)), ell, m,
                                 dp(pk), nm.ctypes.data_as(ct.POINTER(ct.c_int)))
    return pk[0::2] + 1j * pk[1::2], nm


def calc_correlation_function(p: Param, f1, f2, rb, sn, ell, m):
    """field.hpp:854-977: (xi[n_rbin] complex, n_mode[n_rbin])."""
    a1, v1 = _c(f1)
    a2, v2 = _c(f2)
    rb = np.ascontiguousarray(rb, dtype=np.float64)
    xi = np.zeros(2 * p.n_rbin)
    nm = np.zeros(p.n_rbin, dtype=np.int32)
    lib().or_calc_correlation_function(ct.byref(p), dp(v1), dp(v2), dp(rb), float(np.real(sn)), float(np.imag(sn)),
                                       ell, m, dp(xi), nm.ctypes.data_as(ct.POINTER(ct.c_int)))
    return xi[0::2] + 1j * xi[1::2], nm


def ifft_for_bispectrum(p: Param, density, kmag, dk):
    a, v = _c(density)
    out = np.zeros(shape(p) + (2,))
    n = lib().or_ifft_for_bispectrum(ct.byref(p), dp(v), float(kmag), float(dk), None, dp(out))
    return out[..., 0] + 1j * out[..., 1], n


def inverse_fourier_transform(p: Param, field):
    """field.hpp:328-344: / V, then the backward transform."""
    a, v = _c(np.array(field, dtype=np.complex128, copy=True))
    lib().or_inverse_fourier_transform(ct.byref(p), dp(v))
    return a


def store_ylm(p: Param, ell, m, config):
    """tools.hpp:70-115: Y_red(ell, m) of every k (config=False) or r (config=True) of the mesh."""
    out = np.zeros(shape(p) + (2,))
    lib().or_store_ylm(ct.byref(p), int(ell), int(m), 1 if config else 0, dp(out))
    return out[..., 0] + 1j * out[..., 1]


def ifft_for_bispectrum_ylm(p: Param, density, kmag, dk, ell, m):
    """field.hpp:346-401 with the k-space Y_red(ell, m) table: (field, n_mode)."""
    a, v = _c(density)
    y, yv = _c(store_ylm(p, ell, m, False))
    out = np.zeros(shape(p) + (2,))
    n = lib().or_ifft_for_bispectrum(ct.byref(p), dp(v), float(kmag), float(dk), dp(yv), dp(out))
    return out[..., 0] + 1j * out[..., 1], n


def ifft_for_3pcf(p: Param, density, rmag, ell_j, ell, m):
    """field.hpp:403-445 with the k-space Y_red(ell, m) table and the j_ell_j spline."""
    a, v = _c(density)
    y, yv = _c(store_ylm(p, ell, m, False))
    out = np.zeros(shape(p) + (2,))
    lib().or_ifft_for_3pcf(ct.byref(p), dp(v), float(rmag), dp(yv), int(ell_j), dp(out))
    return out[..., 0] + 1j * out[..., 1]


def corr_for_3pcf(p: Param, f1, f2, rb, sn, ell1, m1, ell2, m2):
    """field.hpp:979-1113 with r-space Y tables of (ell1, m1), (ell2, m2); p.ell1/p.ell2 set the sign."""
    a1, v1 = _c(f1)
    a2, v2 = _c(f2)
    y1, w1 = _c(store_ylm(p, ell1, m1, True))
    y2, w2 = _c(store_ylm(p, ell2, m2, True))
    rb = np.ascontiguousarray(rb, dtype=np.float64)
    xi = np.zeros(2 * rb.size)
    nm = np.zeros(rb.size, dtype=np.int32)
    p2 = Param.from_buffer_copy(p)
    p2.ell1, p2.ell2, p2.n_rbin = ell1, ell2, rb.size
    lib().or_corr_for_3pcf(ct.byref(p2), dp(v1), dp(v2), dp(rb), float(np.real(sn)), float(np.imag(sn)), dp(w1),
                           dp(w2), dp(xi), nm.ctypes.data_as(ct.POINTER(ct.c_int)))
    return xi[0::2] + 1j * xi[1::2], nm


def shotnoise_bispectrum_ijk(p: Param, f1, f2, sn):
    """field.hpp:1116-1160."""
    a1, v1 = _c(f1)
    a2, v2 = _c(f2)
    out = np.zeros(shape(p) + (2,))
    lib().or_shotnoise_bispectrum_ijk(ct.byref(p), dp(v1), dp(v2), float(np.real(sn)), float(np.imag(sn)), dp(out))
    return out[..., 0] + 1j * out[..., 1]


def sn_bessel_sweep(p: Param, xi, ell1, m1, ell2, m2, k1, k2):
    """bispec.hpp:1036-1056 for one (k1, k2) with r-space Y tables; i^(l1+l2) dV applied."""
    a, v = _c(xi)
    y1, w1 = _c(store_ylm(p, ell1, m1, True))
    y2, w2 = _c(store_ylm(p, ell2, m2, True))
    out = np.zeros(2)
    lib().or_sn_bessel_sweep(ct.byref(p), dp(v), dp(w1), dp(w2), int(ell1), int(ell2), float(k1), float(k2), dp(out))
    return complex(out[0], out[1])


def triple_sum(p: Param, f1, f2, f3):
    """bispec.hpp:1159-1168."""
    a1, v1 = _c(f1)
    a2, v2 = _c(f2)
    a3, v3 = _c(f3)
    out = np.zeros(2)
    lib().or_triple_sum(ct.byref(p), dp(v1), dp(v2), dp(v3), dp(out))
    return complex(out[0], out[1])


def particle_sum(los, w, ell, m, power):
    """field.hpp:1164-1240."""
    los = np.ascontiguousarray(los, dtype=np.float64).reshape(-1, 3)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    out = np.zeros(2)
    lib().or_particle_sum(dp(los), dp(w) if w is not None else None, los.shape[0], int(ell), int(m), int(power),
                          dp(out))
    return complex(out[0], out[1])


def survey_volume(p: Param, pos, w=None):
    """calcSurveyVolume (field.hpp:449-503)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
    return lib().or_survey_volume(ct.byref(p), dp(pos), dp(w) if w is not None else None, pos.shape[0])


def box_pk(p: Param, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    kb = kbin(p)
    out = np.zeros(2 * p.n_kbin)
    nm = np.zeros(p.n_kbin, dtype=np.int32)
    rc = lib().or_box_pk(ct.byref(p), dp(pos), pos.shape[0], dp(kb), dp(out), nm.ctypes.data_as(ct.POINTER(ct.c_int)))
    if rc:
        raise ValueError("or_box_pk rejected the multipole combination")
    return out[0::2] + 1j * out[1::2], nm


def box_xi(p: Param, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    rb = rbin(p)
    out = np.zeros(p.n_rbin)
    rc = lib().or_box_xi(ct.byref(p), dp(pos), pos.shape[0], dp(rb), dp(out))
    if rc:
        raise ValueError("or_box_xi rejected the multipole combination")
    return out


def box_zeta(p: Param, pos):
    """bispec.hpp:1205-1424 for p.ith_rbin: (out, zeta_save, sn_save), complex arrays."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    rb = rbin(p)
    z = np.zeros(2 * p.n_rbin)
    sn = np.zeros(2 * p.n_rbin)
    out = np.zeros(2 * p.n_rbin)
    rc = lib().or_box_zeta(ct.byref(p), dp(pos), pos.shape[0], dp(rb), dp(z), dp(sn), dp(out))
    if rc:
        raise ValueError("or_box_zeta rejected the multipole combination")
    c = lambda a: a[0::2] + 1j * a[1::2]
    return c(out), c(z), c(sn)


def box_bk(p: Param, pos):
    """bispec.hpp:877-1203 for p.ith_kbin: (out, bk_save, sn_save), complex arrays."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    kb = kbin(p)
    bk = np.zeros(2 * p.n_kbin)
    sn = np.zeros(2 * p.n_kbin)
    out = np.zeros(2 * p.n_kbin)
    rc = lib().or_box_bk(ct.byref(p), dp(pos), pos.shape[0], dp(kb), dp(bk), dp(sn), dp(out))
    if rc:
        raise ValueError("or_box_bk rejected the multipole combination")
    c = lambda a: a[0::2] + 1j * a[1::2]
    return c(out), c(bk), c(sn)


def box_recon_bk(p: Param, posD, posR, alpha):
    """bispec.hpp:1426-1753 for p.ith_kbin on displaced, wrapped catalogues: (out, bk_save, sn_save)."""
    posD = np.ascontiguousarray(posD, dtype=np.float64).reshape(-1, 3)
    posR = np.ascontiguousarray(posR, dtype=np.float64).reshape(-1, 3)
    kb = kbin(p)
    bk, sn, out = np.zeros(2 * p.n_kbin), np.zeros(2 * p.n_kbin), np.zeros(2 * p.n_kbin)
    rc = lib().or_box_recon_bk(ct.byref(p), dp(posD), posD.shape[0], dp(posR), posR.shape[0], float(alpha), dp(kb),
                               dp(bk), dp(sn), dp(out))
    if rc:
        raise ValueError("or_box_recon_bk rejected the multipole combination")
    c = lambda a: a[0::2] + 1j * a[1::2]
    return c(out), c(bk), c(sn)


def box_recon_zeta(p: Param, posD, posR, alpha):
    """bispec.hpp:1755-1975 for p.ith_rbin on displaced, wrapped catalogues: (out, zeta_save, sn_save)."""
    posD = np.ascontiguousarray(posD, dtype=np.float64).reshape(-1, 3)
    posR = np.ascontiguousarray(posR, dtype=np.float64).reshape(-1, 3)
    rb = rbin(p)
    z, sn, out = np.zeros(2 * p.n_rbin), np.zeros(2 * p.n_rbin), np.zeros(2 * p.n_rbin)
    rc = lib().or_box_recon_zeta(ct.byref(p), dp(posD), posD.shape[0], dp(posR), posR.shape[0], float(alpha), dp(rb),
                                 dp(z), dp(sn), dp(out))
    if rc:
        raise ValueError("or_box_recon_zeta rejected the multipole combination")
    c = lambda a: a[0::2] + 1j * a[1::2]
    return c(out), c(z), c(sn)


def box_interlaced_k(p: Param, pos):
    """The interlaced transformed box fluctuation (complex grid)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    f = np.zeros(shape(p) + (2,))
    lib().or_box_interlaced_k(ct.byref(p), dp(pos), pos.shape[0], dp(f))
    return f[..., 0] + 1j * f[..., 1]


def box_interlaced_2pt(p: Param, pos, xi=False):
    """P_l (or xi_l) of the interlaced fluctuation: (values complex, n_mode)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    bins = rbin(p) if xi else kbin(p)
    out = np.zeros(2 * len(bins))
    nm = np.zeros(len(bins), dtype=np.int32)
    lib().or_box_interlaced_2pt(ct.byref(p), dp(pos), pos.shape[0], dp(bins), 1 if xi else 0, dp(out),
                                nm.ctypes.data_as(ct.POINTER(ct.c_int)))
    return out[0::2] + 1j * out[1::2], nm


class Survey:
    """main.cpp:109-166 on a data and a random catalogue: positions shifted
    into the FFT box, LOS of the unshifted positions, alpha, Vsurvey, nbar."""

    def __init__(self, p: Param, posD, wD, posR, wR):
        self.posD = np.array(posD, dtype=np.float64).reshape(-1, 3)
        self.posR = np.array(posR, dtype=np.float64).reshape(-1, 3)
        self.wD = np.ascontiguousarray(wD, dtype=np.float64)
        self.wR = np.ascontiguousarray(wR, dtype=np.float64)
        self.losD = np.zeros_like(self.posD)
        self.losR = np.zeros_like(self.posR)
        info = np.zeros(8)
        rc = lib().or_survey_prepare(ct.byref(p), dp(self.posD), dp(self.wD), len(self.wD), dp(self.posR), dp(self.wR),
                                     len(self.wR), dp(self.losD), dp(self.losR), dp(info))
        if rc:
            raise ValueError("or_survey_prepare: empty weights")
        self.wsumD, self.wsumR, self.alpha, self.vsurvey, self.nbar = info[:5]
        self.shift = info[5:8].copy()

    def _args(self):
        return (dp(self.posD), dp(self.losD), dp(self.wD), len(self.wD), dp(self.posR), dp(self.losR), dp(self.wR),
                len(self.wR))

    def field(self, p: Param, ell, M):
        f = np.zeros(shape(p) + (2,))
        lib().or_ylm_weighted_fluctuation(ct.byref(p), *self._args(), float(self.alpha), ell, M, dp(f))
        return f[..., 0] + 1j * f[..., 1]

    def shotnoise(self, ell, M):
        out = np.zeros(2)
        lib().or_shotnoise_pk(dp(self.losD), dp(self.wD), len(self.wD), dp(self.losR), dp(self.wR), len(self.wR),
                              float(self.alpha), ell, M, dp(out))
        return out[0] + 1j * out[1]

    def pk(self, p: Param):
        kb = kbin(p)
        out = np.zeros(2 * p.n_kbin)
        nm = np.zeros(p.n_kbin, dtype=np.int32)
        rc = lib().or_survey_pk(ct.byref(p), *self._args(), float(self.alpha), float(self.vsurvey), dp(kb), dp(out),
                                nm.ctypes.data_as(ct.POINTER(ct.c_int)))
        if rc:
            raise ValueError("or_survey_pk rejected the multipole combination")
        return out[0::2] + 1j * out[1::2], nm

    def xi(self, p: Param, window=False):
        rb = rbin(p)
        out = np.zeros(2 * p.n_rbin)
        nm = np.zeros(p.n_rbin, dtype=np.int32)
        rc = lib().or_survey_xi(ct.byref(p), *self._args(), float(self.alpha), float(self.vsurvey), dp(rb),
                                1 if window else 0, dp(out), nm.ctypes.data_as(ct.POINTER(ct.c_int)))
        if rc:
            raise ValueError("or_survey_xi rejected the multipole combination")
        return out[0::2] + 1j * out[1::2], nm

    def bk(self, p: Param):
        """bispec.hpp:4-353 for p.ith_kbin: (out, bk_save, sn_save), complex."""
        kb = kbin(p)
        bk, sn, out = np.zeros(2 * p.n_kbin), np.zeros(2 * p.n_kbin), np.zeros(2 * p.n_kbin)
        rc = lib().or_survey_bk(ct.byref(p), *self._args(), float(self.alpha), float(self.vsurvey), dp(kb), dp(bk),
                                dp(sn), dp(out))
        if rc:
            raise ValueError("or_survey_bk rejected the multipole combination")
        c = lambda a: a[0::2] + 1j * a[1::2]
        return c(out), c(bk), c(sn)

    def zeta(self, p: Param, window=False):
        """bispec.hpp:355-614 (window: 617-874) for p.ith_rbin: (out, zeta_save, sn_save), complex."""
        rb = rbin(p)
        z, sn, out = np.zeros(2 * p.n_rbin), np.zeros(2 * p.n_rbin), np.zeros(2 * p.n_rbin)
        rc = lib().or_survey_zeta(ct.byref(p), *self._args(), float(self.alpha), float(self.vsurvey), dp(rb),
                                  1 if window else 0, dp(z), dp(sn), dp(out))
        if rc:
            raise ValueError("or_survey_zeta rejected the multipole combination")
        c = lambda a: a[0::2] + 1j * a[1::2]
        return c(out), c(z), c(sn)


def recon_psi(p: Param, dk, b1, RG):
    """particle_reconstruction.hpp:56-112: (3, n0, n1, n2) complex Psi(x)."""
    a, v = _c(dk)
    out = np.zeros((3,) + shape(p) + (2,))
    lib().or_recon_psi(ct.byref(p), dp(v), float(b1), float(RG), dp(out))
    return out[..., 0] + 1j * out[..., 1]


def recon_gather(p: Param, psi, norm, pos):
    psi = np.ascontiguousarray(np.stack([psi.real, psi.imag], -1), dtype=np.float64)
    pos = np.array(pos, dtype=np.float64).reshape(-1, 3)
    lib().or_recon_gather(ct.byref(p), dp(psi), float(norm), dp(pos), pos.shape[0])
    return pos


def recon_displace(p: Param, posD, wD, posR, wR, alpha, norm, b1, RG):
    """particle_reconstruction.hpp:12-402 on p's mesh: (displaced D, displaced R)."""
    posD = np.array(posD, dtype=np.float64).reshape(-1, 3)
    posR = np.array(posR, dtype=np.float64).reshape(-1, 3)
    wd = None if wD is None else np.ascontiguousarray(wD, dtype=np.float64)
    wr = None if wR is None else np.ascontiguousarray(wR, dtype=np.float64)
    lib().or_recon_displace(ct.byref(p), dp(posD), dp(wd) if wd is not None else None, posD.shape[0], dp(posR),
                            dp(wr) if wr is not None else None, posR.shape[0], float(alpha), float(norm), float(b1),
                            float(RG))
    return posD, posR


def box_recon_2pt(p: Param, posD, posR, alpha, xi=False):
    """powerspec.hpp:396-519: norm * (2l+1) * P_l (or xi_l), complex, and modes per bin."""
    posD = np.ascontiguousarray(posD, dtype=np.float64).reshape(-1, 3)
    posR = np.ascontiguousarray(posR, dtype=np.float64).reshape(-1, 3)
    bins = rbin(p) if xi else kbin(p)
    nb = len(bins)
    out = np.zeros(2 * nb)
    nm = np.zeros(nb, dtype=np.int32)
    rc = lib().or_box_recon_2pt(ct.byref(p), dp(posD), posD.shape[0], dp(posR), posR.shape[0], float(alpha), dp(bins),
                                1 if xi else 0, dp(out), nm.ctypes.data_as(ct.POINTER(ct.c_int)))
    if rc:
        raise ValueError("or_box_recon_2pt rejected the multipole combination")
    return out[0::2] + 1j * out[1::2], nm


# ---------------------------------------------------------------------------
# Mock generator (no counterpart in the reference): Philox noise and the site
# choice from the C restatement, the transforms in numpy (FP64).
# ---------------------------------------------------------------------------
def philox4x32_10(ctr, key):
    c = (ct.c_uint * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    k = (ct.c_uint * 2)(*[int(v) & 0xFFFFFFFF for v in key])
    o = (ct.c_uint * 4)()
    lib().or_philox4x32_10(c, k, o)
    return tuple(int(v) for v in o)


def mock_white_noise(n_side, seed):
    out = np.empty(int(n_side) ** 3)
    lib().or_mock_white_noise(out.size, int(seed), dp(out))
    return out.reshape((n_side,) * 3)


def mock_sites(np_, sites, seed):
    out = np.empty(int(np_), dtype=np.int64)
    lib().or_mock_sites(int(np_), int(sites), int(seed), out.ctypes.data_as(ct.POINTER(ct.c_longlong)))
    return out


def mock_pk(k, amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.0, bao_period=0.0628):
    """P(k) = amp (k/k0)^slope exp(-(k/kcut)^2) (1 + bao_amp sin(2 pi k / bao_period)), P(0) = 0."""
    k = np.asarray(k, dtype=np.float64)
    out = np.zeros_like(k)
    m = k > 0
    p = amp * (k[m] / k0) ** slope * np.exp(-((k[m] / kcut) ** 2))
    if bao_amp != 0.0:
        p = p * (1.0 + bao_amp * np.sin(2.0 * np.pi * k[m] / bao_period))
    out[m] = p
    return out


def mock_kgrid(n, L):
    """Wave vectors of the half spectrum (numpy rfftn layout) and the Nyquist mask."""
    f = np.fft.fftfreq(n, d=1.0 / n)
    fz = np.arange(n // 2 + 1)
    kf = 2.0 * np.pi / L
    kx = (kf * f)[:, None, None]
    ky = (kf * f)[None, :, None]
    kz = (kf * fz)[None, None, :]
    nyq = (np.abs(f)[:, None, None] == n // 2) | (np.abs(f)[None, :, None] == n // 2) | (fz[None, None, :] == n // 2)
    return kx, ky, kz, nyq


def zeldovich_displacement(delta_k, n, L, f_growth=0.0):
    """Psi_a(x) = IFFT(i k_a / k^2 delta(k)) (numpy normalisation 1/n^3); modes on
    a Nyquist plane and k = 0 give nothing; Psi_z times 1 + f_growth."""
    kx, ky, kz, nyq = mock_kgrid(n, L)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    keep = (~nyq) & (k2 > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(keep, 1j * delta_k / k2, 0.0)
    comps = [kx, ky, kz * (1.0 + f_growth)]
    return [np.fft.irfftn(c * g, s=(n,) * 3, axes=(0, 1, 2)) for c in comps]


def mock_delta_k(n, L, seed, pk_kw):
    """delta(k) = FFT(noise) sqrt(P(k) / V_cell) on the half spectrum."""
    w = mock_white_noise(n, seed)
    wk = np.fft.rfftn(w)
    kx, ky, kz, _ = mock_kgrid(n, L)
    kmag = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    return wk * np.sqrt(mock_pk(kmag, **pk_kw) / (L / n) ** 3)


def mock_zeldovich(n, L, seed, np_, f_growth=0.0, **pk_kw):
    """(np_, 3) positions: lattice sites (mock_sites) moved by Psi, wrapped as
    x - L floor(x / L), then minus L where that reaches L."""
    psi = zeldovich_displacement(mock_delta_k(n, L, seed, pk_kw), n, L, f_growth)
    sites = mock_sites(np_, n ** 3, seed)
    ix, iy, iz = sites // (n * n), (sites // n) % n, sites % n
    cell = L / n
    pos = np.empty((np_, 3))
    for a, (idx, ps) in enumerate(zip((ix, iy, iz), psi)):
        x = (idx + 0.5) * cell + ps.ravel()[sites]
        v = x - L * np.floor(x / L)
        v[v >= L] -= L
        pos[:, a] = v
    return pos


def mock_zeldovich_biased(n, L, seed, mean_n, b_lag, f_growth=0.0, **pk_kw):
    """Biased tracers: site s kept when u_s < rate max(0, 1 + b_lag delta_L(s)),
    rate = mean_n / sum of those weights;
    returns (positions, shifts -Psi) of the kept sites in site order."""
    dk = mock_delta_k(n, L, seed, pk_kw)
    kx, ky, kz, nyq = mock_kgrid(n, L)
    dk0 = np.where(nyq, 0.0, dk)
    dk0[0, 0, 0] = 0.0
    dl = np.fft.irfftn(dk0, s=(n,) * 3, axes=(0, 1, 2)).ravel()
    u = np.empty(n ** 3)
    lib().or_mock_select_uniform(u.size, int(seed), dp(u))
    wgt = np.maximum(0.0, 1.0 + b_lag * dl)
    keep = np.nonzero(u < (mean_n / wgt.sum()) * wgt)[0]
    psi = zeldovich_displacement(dk, n, L, f_growth)
    ix, iy, iz = keep // (n * n), (keep // n) % n, keep % n
    cell = L / n
    pos = np.empty((keep.size, 3))
    shift = np.empty((keep.size, 3))
    for a, (idx, ps) in enumerate(zip((ix, iy, iz), psi)):
        d = ps.ravel()[keep]
        x = (idx + 0.5) * cell + d
        v = x - L * np.floor(x / L)
        v[v >= L] -= L
        pos[:, a] = v
        shift[:, a] = -d
    return pos, shift, keep


# ---------------------------------------------------------------------------
# Learned reconstruction network (no counterpart in the reference): periodic
# 3^3 convolutions, 4 -> 16 -> 16 -> 16 -> 3 channels, ReLU after the first
# three.  Tap t = 9 (dx + 1) + 3 (dy + 1) + (dz + 1) reads voxel v + (dx, dy, dz).
# With bf16=True every operand the device rounds to BF16 (input, weights,
# stored activations, the incoming gradient and the stored layer gradients)
# is rounded the same way (round to nearest even) and every stored FP32
# accumulator is rounded to FP32 first; sums are FP64.
# ---------------------------------------------------------------------------
LRECON_CIN = (4, 16, 16, 16)
LRECON_COUT = (16, 16, 16, 3)


def bf16_round(x):
    """Round to the nearest BF16 (ties to even), as float32 values."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def lrecon_unpack(params):
    """Flat parameters -> [(W [27][cin][cout], b [cout])] * 4."""
    out, o = [], 0
    for ci, co in zip(LRECON_CIN, LRECON_COUT):
        W = np.asarray(params[o:o + 27 * ci * co], dtype=np.float64).reshape(27, ci, co)
        o += 27 * ci * co
        b = np.asarray(params[o:o + co], dtype=np.float64)
        o += co
        out.append((W, b))
    return out


def lrecon_pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def _taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def conv3_periodic(x, W, b=None):
    """y[v][co] = b[co] + sum_t sum_ci W[t][ci][co] x[v + off_t][ci] (FP64)."""
    y = np.zeros(x.shape[:3] + (W.shape[2],))
    for t, (dx, dy, dz) in enumerate(_taps()):
        y += np.roll(x, shift=(-dx, -dy, -dz), axis=(0, 1, 2)) @ W[t]
    return y if b is None else y + b


def conv3_input_grad(dy, W):
    """d/dx of sum(conv3_periodic(x, W) * dy): sum_t roll(dy @ W[t]^T, +off_t)."""
    dx = np.zeros(dy.shape[:3] + (W.shape[1],))
    for t, (ox, oy, oz) in enumerate(_taps()):
        dx += np.roll(dy @ W[t].T, shift=(ox, oy, oz), axis=(0, 1, 2))
    return dx


def conv3_weight_grad(x, dy):
    """d/dW[t][ci][co] = sum_v x[v + off_t][ci] dy[v][co]; d/db = sum_v dy."""
    ci, co = x.shape[3], dy.shape[3]
    dW = np.zeros((27, ci, co))
    X = dy.reshape(-1, co)
    for t, (dx_, dy_, dz_) in enumerate(_taps()):
        dW[t] = np.roll(x, shift=(-dx_, -dy_, -dz_), axis=(0, 1, 2)).reshape(-1, ci).T @ X
    return dW, X.sum(0)


def lrecon_net(x, params, bf16=True):
    """Forward pass: x [n,n,n,4] -> (out [n,n,n,3], inputs of each layer)."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    f32 = (lambda a: np.asarray(a, dtype=np.float32)) if bf16 else (lambda a: a)
    layers = lrecon_unpack(params)
    a = r(x).astype(np.float64)
    ins = []
    for l, (W, b) in enumerate(layers):
        ins.append(a)
        z = conv3_periodic(a, r(W).astype(np.float64), b)
        if l < 3:
            a = r(f32(np.maximum(z, 0.0))).astype(np.float64)
        else:
            a = f32(z).astype(np.float64)
    return a, ins


def lrecon_net_grad(x, params, gout, bf16=True):
    """Parameter gradient of sum(out * gout) (gout [n,n,n,3])."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    f32 = (lambda a: np.asarray(a, dtype=np.float32)) if bf16 else (lambda a: a)
    layers = lrecon_unpack(params)
    _, ins = lrecon_net(x, params, bf16)
    dz = r(gout).astype(np.float64)
    grads = [None] * 4
    for l in range(3, -1, -1):
        W = r(layers[l][0]).astype(np.float64)
        dW, db = conv3_weight_grad(ins[l], dz)
        grads[l] = (dW, db)
        if l == 0:
            break
        da = conv3_input_grad(dz, W)
        dz = r(f32(np.where(ins[l] > 0.0, da, 0.0))).astype(np.float64)
    return lrecon_pack(grads)


def sky_randoms(patch, seed, n):
    """or_sky_randoms: n points uniform in the sky patch (rmin, rmax, ra_min, ra_max, dec_min, dec_max)."""
    pt = np.ascontiguousarray(patch, dtype=np.float64)
    out = np.zeros((int(n), 3))
    lib().or_sky_randoms(dp(pt), int(seed), int(n), dp(out))
    return out
