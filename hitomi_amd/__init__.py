"""hitomi_amd -- MI355X-native FFT estimator path of hitomi_measurement.

Python view of libhitomi_amd.so (include/hitomi_amd.h) through ctypes.  The
compute happens in the gfx950 kernels of the shared library; this module
only moves arrays and unpacks results.  Importing it does not touch the GPU;
creating a Context does, and fails loudly when the library or a GPU is
missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhitomi_amd.so")

NGP, CIC, TSC = 0, 1, 2
FP32, FP64 = 0, 1
ASSIGN = {"NGP": NGP, "CIC": CIC, "TSC": TSC}
MAX_ELLS = 4

# Launch names the box pipeline reports to hm_timing_get().
BOX_KERNELS = ("fill_zero", "assign_count", "assign_scan", "assign_bucket", "assign_scatter",
               "assign_gather", "assign_border", "fft_r2c_z", "fft_r2c_y", "fft_r2c_x", "fft_x_xi", "bin_pk",
               "bin_reduce", "xi_pass_x", "xi_planes_x", "xi_pass_y", "xi_pass_z", "xi_group", "xi_hist", "sn_pairs", "tri_stage", "tri_pass_x",
               "tri3_pass_x", "tri_pass_y", "tri3_pass_y", "tri_pass_z", "tri3_pass_z", "tri_z_gram", "gram",
               "gram_reduce")


class HitomiError(RuntimeError):
    pass


class Mesh(ct.Structure):
    _fields_ = [("n_mesh", ct.c_int * 3), ("boxsize", ct.c_double * 3), ("assign", ct.c_int)]


class BoxConfig(ct.Structure):
    _fields_ = [
        ("mesh", Mesh),
        ("precision", ct.c_int),
        ("n_kbin", ct.c_int),
        ("kmin", ct.c_double),
        ("kmax", ct.c_double),
        ("n_pk_ells", ct.c_int),
        ("pk_ells", ct.c_int * MAX_ELLS),
        ("bk", ct.c_int),
        ("regrid", ct.c_int),
        ("prune", ct.c_int),
        ("n_rbin", ct.c_int),
        ("rmin", ct.c_double),
        ("rmax", ct.c_double),
        ("n_xi_ells", ct.c_int),
        ("xi_ells", ct.c_int * MAX_ELLS),
        ("bk_ell", ct.c_int * 3),
        ("zeta", ct.c_int),
        ("zeta_ell", ct.c_int * 3),
        ("interlace", ct.c_int),
    ]


class BatchConfig(ct.Structure):
    _fields_ = [("box", BoxConfig), ("max_mocks", ct.c_longlong), ("cov_with_bk", ct.c_int), ("use_graph", ct.c_int),
                ("lanes", ct.c_int)]


class MockConfig(ct.Structure):
    _fields_ = [("n_side", ct.c_int), ("boxsize", ct.c_double), ("amp", ct.c_double), ("k0", ct.c_double),
                ("slope", ct.c_double), ("kcut", ct.c_double), ("bao_amp", ct.c_double), ("bao_period", ct.c_double),
                ("f_growth", ct.c_double)]


class SkyPatch(ct.Structure):
    _fields_ = [("rmin", ct.c_double), ("rmax", ct.c_double), ("ra_min", ct.c_double), ("ra_max", ct.c_double),
                ("dec_min", ct.c_double), ("dec_max", ct.c_double)]


class LReconConfig(ct.Structure):
    _fields_ = [("n_mesh", ct.c_int), ("boxsize", ct.c_double), ("b1", ct.c_double), ("RG", ct.c_double),
                ("psi_scale", ct.c_double), ("lr", ct.c_double), ("init_seed", ct.c_ulonglong), ("precision", ct.c_int),
                ("conv_cache_taps", ct.c_int)]


class BoxResult(ct.Structure):
    _fields_ = [
        ("pk", ct.POINTER(ct.c_double)),
        ("pk_raw", ct.POINTER(ct.c_double)),
        ("pk_nmode", ct.POINTER(ct.c_longlong)),
        ("bk", ct.POINTER(ct.c_double)),
        ("bk_sum", ct.POINTER(ct.c_double)),
        ("sn_sum", ct.POINTER(ct.c_double)),
        ("shell_nmode", ct.POINTER(ct.c_longlong)),
        ("xi", ct.POINTER(ct.c_double)),
        ("xi_raw", ct.POINTER(ct.c_double)),
        ("xi_nmode", ct.POINTER(ct.c_longlong)),
        ("zeta", ct.POINTER(ct.c_double)),
        ("zeta_sum", ct.POINTER(ct.c_double)),
        ("zeta_sn", ct.POINTER(ct.c_double)),
    ]


class SurveyConfig(ct.Structure):
    _fields_ = [
        ("mesh", Mesh),
        ("precision", ct.c_int),
        ("n_kbin", ct.c_int),
        ("kmin", ct.c_double),
        ("kmax", ct.c_double),
        ("ell", ct.c_int),
        ("prune", ct.c_int),
        ("measure", ct.c_int),
        ("n_rbin", ct.c_int),
        ("rmin", ct.c_double),
        ("rmax", ct.c_double),
        ("recon", ct.c_int),
        ("n_mesh_recon", ct.c_int * 3),
        ("b1", ct.c_double),
        ("RG", ct.c_double),
        ("ell1", ct.c_int),
        ("ell2", ct.c_int),
    ]


SURVEY_MEASURES = {"pk": 0, "2PCF": 1, "window2PCF": 2, "bk": 3, "3PCF": 4, "window3PCF": 5}


class SurveyInfo(ct.Structure):
    _fields_ = [
        ("wsum_d", ct.c_double),
        ("wsum_r", ct.c_double),
        ("alpha", ct.c_double),
        ("vsurvey", ct.c_double),
        ("nbar", ct.c_double),
        ("shift", ct.c_double * 3),
        ("norm", ct.c_double),
    ]


class ReconConfig(ct.Structure):
    _fields_ = [
        ("mesh", Mesh),
        ("precision", ct.c_int),
        ("b1", ct.c_double),
        ("RG", ct.c_double),
    ]


# Every symbol include/hitomi_amd.h declares, with (restype, argtypes).
_VP = ct.c_void_p
_SIGS = {
    "hm_ctx_create": (_VP, [ct.c_int, ct.c_int]),
    "hm_ctx_destroy": (None, [_VP]),
    "hm_last_error": (ct.c_char_p, [_VP]),
    "hm_sync": (ct.c_int, [_VP]),
    "hm_dev_alloc": (_VP, [_VP, ct.c_size_t]),
    "hm_dev_free": (ct.c_int, [_VP, _VP]),
    "hm_memcpy_h2d": (ct.c_int, [_VP, _VP, _VP, ct.c_size_t]),
    "hm_memcpy_d2h": (ct.c_int, [_VP, _VP, _VP, ct.c_size_t]),
    "hm_memcpy_d2d": (ct.c_int, [_VP, _VP, _VP, ct.c_size_t]),
    "hm_memset_zero": (ct.c_int, [_VP, _VP, ct.c_size_t]),
    "hm_timing_enable": (ct.c_int, [_VP, ct.c_int]),
    "hm_timing_get": (ct.c_int, [_VP, ct.c_char_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_longlong)]),
    "hm_timing_reset": (ct.c_int, [_VP]),
    "hm_timing_get_busy": (ct.c_int, [_VP, ct.c_char_p, ct.POINTER(ct.c_double)]),
    "hm_timing_select": (ct.c_int, [_VP, ct.c_char_p]),
    "hm_assign": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.c_longlong, ct.c_double, _VP, _VP]),
    "hm_assign_atomic": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.c_longlong, ct.c_double, _VP, _VP]),
    "hm_assign_noalloc": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.c_longlong, ct.c_double, _VP, _VP]),
    "hm_mock_sky_randoms": (ct.c_int, [_VP, ct.POINTER(SkyPatch), ct.c_ulonglong, ct.c_longlong, _VP]),
    "hm_cell_coords": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, ct.c_longlong, _VP]),
    "hm_fft_c2c": (ct.c_int, [_VP, ct.c_int, _VP, ct.POINTER(ct.c_int), ct.c_int]),
    "hm_fft_r2c": (ct.c_int, [_VP, ct.c_int, _VP, _VP, ct.POINTER(ct.c_int)]),
    "hm_fft_c2r": (ct.c_int, [_VP, ct.c_int, _VP, _VP, ct.POINTER(ct.c_int)]),
    "hm_grid_alloc": (_VP, [_VP, ct.POINTER(Mesh)]),
    "hm_grid_free": (ct.c_int, [_VP, _VP]),
    "hm_assign_field": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.c_longlong, ct.c_double, _VP]),
    "hm_fourier_transform": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, ct.c_int]),
    "hm_window_correct": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP]),
    "hm_bin_pk": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.POINTER(ct.c_double), ct.c_int,
                             ct.POINTER(ct.c_double), ct.c_int, ct.c_int, ct.POINTER(ct.c_double),
                             ct.POINTER(ct.c_longlong)]),
    "hm_bin_xi": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.POINTER(ct.c_double), ct.c_int,
                             ct.POINTER(ct.c_double), ct.c_int, ct.c_int, ct.POINTER(ct.c_double),
                             ct.POINTER(ct.c_longlong)]),
    "hm_bin_xi_3pcf": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.POINTER(ct.c_double), ct.c_int,
                                  ct.POINTER(ct.c_double), ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                  ct.POINTER(ct.c_double), ct.POINTER(ct.c_longlong)]),
    "hm_sn_field_ijk": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.POINTER(ct.c_double), _VP]),
    "hm_shell_c2r": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, ct.c_double, ct.c_double, ct.c_int, ct.c_int, _VP,
                                ct.POINTER(ct.c_longlong)]),
    "hm_bessel_c2r": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, ct.c_double, ct.c_int, ct.c_int, ct.c_int, _VP]),
    "hm_triple_sums": (ct.c_int, [_VP, ct.POINTER(Mesh), ct.POINTER(_VP), ct.c_int, ct.POINTER(ct.c_int), ct.c_int,
                                  _VP, ct.POINTER(ct.c_double)]),
    "hm_sn_bessel_sweep": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                      ct.POINTER(ct.c_double), ct.POINTER(ct.c_double), ct.c_int,
                                      ct.POINTER(ct.c_double)]),
    "hm_particle_sums": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong, ct.c_int, ct.c_int, ct.c_int,
                                    ct.POINTER(ct.c_double)]),
    "hm_survey_volume": (ct.c_int, [_VP, ct.POINTER(Mesh), _VP, _VP, ct.c_longlong, ct.POINTER(ct.c_double)]),
    "hm_box_plan_create": (_VP, [_VP, ct.POINTER(BoxConfig)]),
    "hm_box_plan_destroy": (None, [_VP, _VP]),
    "hm_box_plan_execute": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong, ct.POINTER(BoxResult)]),
    "hm_box_plan_execute_recon": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong, _VP, ct.c_longlong, ct.c_double,
                                             ct.POINTER(BoxResult)]),
    "hm_box_plan_bytes": (ct.c_size_t, [_VP]),
    "hm_box_plan_triple_mesh": (ct.c_int, [_VP, ct.POINTER(ct.c_int)]),
    "hm_box_plan_autotune": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong, ct.c_int, ct.c_char_p, ct.c_size_t]),
    "hm_box_plan_get_shapes": (ct.c_int, [_VP, ct.POINTER(ct.c_uint), ct.POINTER(ct.c_int)]),
    "hm_box_plan_set_shapes": (ct.c_int, [_VP, ct.c_uint, ct.c_int]),
    "hm_recon_displace": (ct.c_int, [_VP, ct.POINTER(ReconConfig), _VP, _VP, ct.c_longlong, _VP, _VP, ct.c_longlong,
                                     ct.c_double, ct.c_double]),
    "hm_sph_bessel_jl": (ct.c_double, [ct.c_int, ct.c_double]),
    "hm_wigner_3j": (ct.c_double, [ct.c_int] * 6),
    "hm_ylm_reduced": (None, [ct.c_int, ct.c_int, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "hm_periodic_wrap": (ct.c_int, [_VP, _VP, ct.c_longlong, ct.POINTER(ct.c_double)]),
    "hm_survey_plan_create": (_VP, [_VP, ct.POINTER(SurveyConfig)]),
    "hm_survey_plan_destroy": (None, [_VP, _VP]),
    "hm_survey_plan_execute": (ct.c_int, [_VP, _VP, _VP, _VP, ct.c_longlong, _VP, _VP, ct.c_longlong,
                                          ct.POINTER(SurveyInfo), ct.POINTER(ct.c_double), ct.POINTER(ct.c_longlong)]),
    "hm_survey_plan_bytes": (ct.c_size_t, [_VP]),
    "hm_batch_create": (_VP, [_VP, ct.POINTER(BatchConfig)]),
    "hm_batch_destroy": (None, [_VP, _VP]),
    "hm_batch_dims": (ct.c_int, [_VP, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]),
    "hm_batch_push": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong]),
    "hm_batch_count": (ct.c_longlong, [_VP]),
    "hm_batch_collect_timing": (ct.c_int, [_VP, _VP]),
    "hm_batch_vectors": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong]),
    "hm_batch_covariance": (ct.c_int, [_VP, _VP, _VP, _VP]),
    "hm_batch_reset": (ct.c_int, [_VP, _VP]),
    "hm_batch_autotune": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong, ct.c_int, ct.c_char_p, ct.c_size_t]),
    "hm_cov_create": (_VP, [_VP, ct.c_int, _VP, ct.c_longlong]),
    "hm_cov_destroy": (None, [_VP, _VP]),
    "hm_cov_push": (ct.c_int, [_VP, _VP, _VP]),
    "hm_cov_push_host": (ct.c_int, [_VP, _VP, _VP]),
    "hm_cov_count": (ct.c_longlong, [_VP]),
    "hm_cov_vectors": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong]),
    "hm_cov_result": (ct.c_int, [_VP, _VP, _VP, _VP]),
    "hm_cov_reset": (ct.c_int, [_VP, _VP]),
    "hm_mock_white_noise": (ct.c_int, [_VP, ct.POINTER(MockConfig), ct.c_ulonglong, _VP]),
    "hm_mock_zeldovich": (ct.c_int, [_VP, ct.POINTER(MockConfig), ct.c_ulonglong, ct.c_longlong, _VP]),
    "hm_mock_zeldovich_biased": (ct.c_int, [_VP, ct.POINTER(MockConfig), ct.c_ulonglong, ct.c_double, ct.c_double,
                                            ct.c_longlong, _VP, _VP, ct.POINTER(ct.c_longlong)]),
    "hm_lrecon_create": (_VP, [_VP, ct.POINTER(LReconConfig)]),
    "hm_lrecon_destroy": (None, [_VP, _VP]),
    "hm_lrecon_n_params": (ct.c_longlong, [_VP]),
    "hm_lrecon_get_params": (ct.c_int, [_VP, _VP, _VP]),
    "hm_lrecon_set_params": (ct.c_int, [_VP, _VP, _VP]),
    "hm_lrecon_features": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong, _VP]),
    "hm_lrecon_net_forward": (ct.c_int, [_VP, _VP, _VP, _VP]),
    "hm_lrecon_net_backward": (ct.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "hm_lrecon_predict": (ct.c_int, [_VP, _VP, _VP, ct.c_longlong, _VP, _VP]),
    "hm_lrecon_gradient": (ct.c_int, [_VP, _VP, _VP, _VP, ct.c_longlong, ct.POINTER(ct.c_double),
                                      ct.POINTER(ct.c_double), _VP, _VP]),
    "hm_lrecon_train_step": (ct.c_int, [_VP, _VP, _VP, _VP, ct.c_longlong, ct.POINTER(ct.c_double),
                                        ct.POINTER(ct.c_double)]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> ct.CDLL:
    """Load libhitomi_amd.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HitomiError(f"{path} not found: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ct.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def make_mesh(n_mesh, boxsize, assign="TSC") -> Mesh:
    m = Mesh()
    n = list(n_mesh) if np.ndim(n_mesh) else [int(n_mesh)] * 3
    L = list(boxsize) if np.ndim(boxsize) else [float(boxsize)] * 3
    for a in range(3):
        m.n_mesh[a] = int(n[a])
        m.boxsize[a] = float(L[a])
    m.assign = ASSIGN[assign] if isinstance(assign, str) else int(assign)
    return m


class DeviceArray:
    """A library-owned device allocation with its shape and dtype."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctx.lib.hm_dev_alloc(ctx.h, max(self.nbytes, 1))
        if not self.ptr:
            ctx.raise_error("hm_dev_alloc")

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        self.ctx.check(self.ctx.lib.hm_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "h2d")
        return self

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        self.ctx.check(self.ctx.lib.hm_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes), "d2h")
        return out

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of other's bytes into this array (hm_memcpy_d2d)."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        self.ctx.check(self.ctx.lib.hm_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.nbytes), "d2d")
        return self

    def free(self):
        if self.ptr:
            self.ctx.lib.hm_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device, one stream (hm_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.h = self.lib.hm_ctx_create(device, 0)
        if not self.h:
            raise HitomiError(f"hm_ctx_create({device}) failed: no usable GPU")

    def close(self):
        if self.h:
            self.lib.hm_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def raise_error(self, what):
        msg = self.lib.hm_last_error(self.h)
        raise HitomiError(f"{what}: {msg.decode() if msg else '?'}")

    def check(self, rc, what):
        if rc != 0:
            self.raise_error(what)

    def array(self, a, dtype=None) -> DeviceArray:
        a = np.ascontiguousarray(a, dtype=dtype)
        return DeviceArray(self, a.shape, a.dtype).upload(a)

    def empty(self, shape, dtype) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def sync(self):
        self.check(self.lib.hm_sync(self.h), "hm_sync")

    # timing
    def timing(self, on=True):
        self.check(self.lib.hm_timing_enable(self.h, 1 if on else 0), "hm_timing_enable")

    def kernel_time(self, name):
        ms = ct.c_double()
        n = ct.c_longlong()
        self.check(self.lib.hm_timing_get(self.h, name.encode(), ct.byref(ms), ct.byref(n)), "hm_timing_get")
        return ms.value, n.value

    def kernel_busy(self, name):
        """ms during which at least one timed launch of `name` ran (overlaps count once)."""
        ms = ct.c_double()
        self.check(self.lib.hm_timing_get_busy(self.h, name.encode(), ct.byref(ms)), "hm_timing_get_busy")
        return ms.value

    def timing_select(self, name=None):
        self.check(self.lib.hm_timing_select(self.h, name.encode() if name else None), "hm_timing_select")

    def timing_reset(self):
        self.check(self.lib.hm_timing_reset(self.h), "hm_timing_reset")

    # a6: mass assignment
    def assign(self, mesh: Mesh, pos, weights=None, wabs_sum=0.0, atomic=False):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        npart = pos.shape[0]
        dpos = self.array(pos)
        dw = self.array(np.ascontiguousarray(weights, dtype=np.float64).reshape(-1, 2)) if weights is not None else None
        ng = mesh.n_mesh[0] * mesh.n_mesh[1] * mesh.n_mesh[2]
        shape = (mesh.n_mesh[0], mesh.n_mesh[1], mesh.n_mesh[2])
        re = self.empty(shape, np.float64)
        im = self.empty(shape, np.float64) if weights is not None else None
        fn = self.lib.hm_assign_atomic if atomic else self.lib.hm_assign
        self.check(
            fn(self.h, ct.byref(mesh), dpos.ptr, dw.ptr if dw else None, npart, float(wabs_sum),
                               re.ptr, im.ptr if im else None), "hm_assign")
        out_re = re.download()
        out_im = im.download() if im is not None else np.zeros_like(out_re)
        assert out_re.size == ng
        return out_re, out_im

    # a28: reconstruction
    def recon_displace(self, n_mesh, boxsize, pos_d, pos_r, alpha, norm, b1, RG, w_d=None, w_r=None, assign="TSC",
                       precision=FP64):
        """particle_reconstruction.hpp: (displaced D, displaced R) as host arrays."""
        cfg = ReconConfig()
        cfg.mesh = make_mesh(n_mesh, boxsize, assign)
        cfg.precision = precision
        cfg.b1 = float(b1)
        cfg.RG = float(RG)
        dD = self.array(np.ascontiguousarray(pos_d, dtype=np.float64).reshape(-1, 3))
        dR = self.array(np.ascontiguousarray(pos_r, dtype=np.float64).reshape(-1, 3))
        wd = self.array(np.ascontiguousarray(w_d, dtype=np.float64)) if w_d is not None else None
        wr = self.array(np.ascontiguousarray(w_r, dtype=np.float64)) if w_r is not None else None
        self.check(self.lib.hm_recon_displace(self.h, ct.byref(cfg), dD.ptr, wd.ptr if wd else None, dD.shape[0], dR.ptr,
                                              wr.ptr if wr else None, dR.shape[0], float(alpha), float(norm)),
                   "hm_recon_displace")
        return dD.download(), dR.download()

    def periodic_wrap(self, pos, boxsize):
        d = self.array(np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3))
        L = (ct.c_double * 3)(*(list(boxsize) if np.ndim(boxsize) else [float(boxsize)] * 3))
        self.check(self.lib.hm_periodic_wrap(self.h, d.ptr, d.shape[0], L), "hm_periodic_wrap")
        return d.download()

    def cell_coords(self, mesh: Mesh, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        s = {NGP: 1, CIC: 8, TSC: 27}[mesh.assign]
        dpos = self.array(pos)
        out = self.empty((pos.shape[0], s), np.int64)
        self.check(self.lib.hm_cell_coords(self.h, ct.byref(mesh), dpos.ptr, pos.shape[0], out.ptr), "hm_cell_coords")
        return out.download()

    # a8: FFTs
    def fft_c2c(self, a, sign, precision=FP64):
        dt = np.complex128 if precision == FP64 else np.complex64
        a = np.ascontiguousarray(a, dtype=dt)
        n = (ct.c_int * 3)(*a.shape)
        d = self.array(a)
        self.check(self.lib.hm_fft_c2c(self.h, precision, d.ptr, n, int(sign)), "hm_fft_c2c")
        return d.download()

    def fft_r2c(self, a, precision=FP64):
        dt = np.float64 if precision == FP64 else np.float32
        a = np.ascontiguousarray(a, dtype=dt)
        n = (ct.c_int * 3)(*a.shape)
        d = self.array(a)
        out = self.empty((a.shape[0], a.shape[1], a.shape[2] // 2 + 1),
                         np.complex128 if precision == FP64 else np.complex64)
        self.check(self.lib.hm_fft_r2c(self.h, precision, d.ptr, out.ptr, n), "hm_fft_r2c")
        return out.download()

    def fft_c2r(self, h, nz, precision=FP64):
        dt = np.complex128 if precision == FP64 else np.complex64
        h = np.ascontiguousarray(h, dtype=dt)
        n = (ct.c_int * 3)(h.shape[0], h.shape[1], int(nz))
        d = self.array(h)
        out = self.empty((h.shape[0], h.shape[1], int(nz)), np.float64 if precision == FP64 else np.float32)
        self.check(self.lib.hm_fft_c2r(self.h, precision, d.ptr, out.ptr, n), "hm_fft_c2r")
        return out.download()


class BoxPlan:
    """Periodic-box plan (hm_box_plan): calcPowerSpectrumForBOX
    (powerspec.hpp:273-332), calcTwoPointCorrelationFunctionForBOX
    (powerspec.hpp:334-393) and calcBiSpectrumForBOX (bispec.hpp:877-1203) for
    every ith_kbin at once."""

    def __init__(self, ctx: Context, n_mesh, boxsize, assign="TSC", n_kbin=20, kmin=0.01, kmax=0.2,
                 pk_ells=(0, 2, 4), bk=True, precision=FP32, regrid=True, prune=True, xi_ells=(), n_rbin=25,
                 rmin=30.0, rmax=150.0, bk_ells=(0, 0, 0), zeta=False, zeta_ells=(0, 0, 0), interlace=False):
        self.ctx = ctx
        cfg = BoxConfig()
        cfg.mesh = make_mesh(n_mesh, boxsize, assign)
        cfg.precision = precision
        cfg.n_kbin = int(n_kbin)
        cfg.kmin = float(kmin)
        cfg.kmax = float(kmax)
        cfg.n_pk_ells = len(pk_ells)
        for i, l in enumerate(pk_ells):
            cfg.pk_ells[i] = int(l)
        cfg.bk = 1 if bk else 0
        cfg.regrid = 1 if regrid else 0
        cfg.prune = 1 if prune else 0
        cfg.n_rbin = int(n_rbin)
        cfg.rmin = float(rmin)
        cfg.rmax = float(rmax)
        for i in range(3):
            cfg.bk_ell[i] = int(bk_ells[i])
            cfg.zeta_ell[i] = int(zeta_ells[i])
        cfg.zeta = 1 if zeta else 0
        cfg.interlace = 1 if interlace else 0
        cfg.n_xi_ells = len(xi_ells)
        for i, l in enumerate(xi_ells):
            cfg.xi_ells[i] = int(l)
        self.cfg = cfg
        self.xi_ells = tuple(xi_ells)
        self.n_rbin = int(n_rbin)
        self.n_kbin = int(n_kbin)
        self.pk_ells = tuple(pk_ells)
        self.h = ctx.lib.hm_box_plan_create(ctx.h, ct.byref(cfg))
        if not self.h:
            ctx.raise_error("hm_box_plan_create")

    def close(self):
        if self.h:
            self.ctx.lib.hm_box_plan_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self):
        return self.ctx.lib.hm_box_plan_bytes(self.h)

    @property
    def triple_mesh(self):
        m = (ct.c_int * 3)()
        self.ctx.lib.hm_box_plan_triple_mesh(self.h, m)
        return tuple(m)

    def autotune(self, pos: "DeviceArray", reps=2, npart=None):
        """Choose the launch shapes by timing (results do not depend on them); returns the report."""
        n = pos.shape[0] if npart is None else int(npart)
        buf = ct.create_string_buffer(4096)
        self.ctx.check(self.ctx.lib.hm_box_plan_autotune(self.ctx.h, self.h, pos.ptr, n, int(reps), buf, 4096),
                       "hm_box_plan_autotune")
        return buf.value.decode()

    @property
    def shapes(self):
        w, b = ct.c_uint(), ct.c_int()
        self.ctx.lib.hm_box_plan_get_shapes(self.h, ct.byref(w), ct.byref(b))
        return w.value, b.value

    @shapes.setter
    def shapes(self, wb):
        self.ctx.lib.hm_box_plan_set_shapes(self.h, int(wb[0]), int(wb[1]))

    def execute(self, pos, npart=None, recon=None):
        """pos: host (N,3) float64 array or a DeviceArray of it."""
        if isinstance(pos, DeviceArray):
            dpos = pos
            npart = pos.shape[0] if npart is None else npart
        else:
            pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
            dpos = self.ctx.array(pos)
            npart = pos.shape[0]
        nk, ne = self.n_kbin, len(self.pk_ells)
        out = {
            "pk": np.zeros((ne, nk, 2)), "pk_raw": np.zeros((ne, nk, 2)), "pk_nmode": np.zeros(nk, np.int64),
            "bk": np.zeros((nk, nk, 2)), "bk_sum": np.zeros((nk, nk, 2)), "sn_sum": np.zeros((nk, nk, 2)),
            "shell_nmode": np.zeros(nk, np.int64),
            "xi": np.zeros((len(self.xi_ells), self.n_rbin)), "xi_raw": np.zeros((len(self.xi_ells), self.n_rbin)),
            "xi_nmode": np.zeros(self.n_rbin, np.int64),
            "zeta": np.zeros((self.n_rbin, self.n_rbin, 2)), "zeta_sum": np.zeros((self.n_rbin, self.n_rbin, 2)),
            "zeta_sn": np.zeros((self.n_rbin, self.n_rbin, 2)),
        }
        r = BoxResult()
        dp = ct.POINTER(ct.c_double)
        lp = ct.POINTER(ct.c_longlong)
        r.pk = out["pk"].ctypes.data_as(dp)
        r.pk_raw = out["pk_raw"].ctypes.data_as(dp)
        r.pk_nmode = out["pk_nmode"].ctypes.data_as(lp)
        r.bk = out["bk"].ctypes.data_as(dp)
        r.bk_sum = out["bk_sum"].ctypes.data_as(dp)
        r.sn_sum = out["sn_sum"].ctypes.data_as(dp)
        r.shell_nmode = out["shell_nmode"].ctypes.data_as(lp)
        r.xi = out["xi"].ctypes.data_as(dp)
        r.xi_raw = out["xi_raw"].ctypes.data_as(dp)
        r.xi_nmode = out["xi_nmode"].ctypes.data_as(lp)
        r.zeta = out["zeta"].ctypes.data_as(dp)
        r.zeta_sum = out["zeta_sum"].ctypes.data_as(dp)
        r.zeta_sn = out["zeta_sn"].ctypes.data_as(dp)
        if recon is None:
            self.ctx.check(self.ctx.lib.hm_box_plan_execute(self.ctx.h, self.h, dpos.ptr, int(npart), ct.byref(r)),
                           "hm_box_plan_execute")
        else:
            dr, alpha = recon
            self.ctx.check(self.ctx.lib.hm_box_plan_execute_recon(self.ctx.h, self.h, dpos.ptr, int(npart), dr.ptr,
                                                                  dr.shape[0], float(alpha), ct.byref(r)),
                           "hm_box_plan_execute_recon")
        return out

    def execute_recon(self, pos_d, pos_r, alpha):
        """The reconstructed box: D - alpha R (powerspec.hpp:396-519)."""
        dr = pos_r if isinstance(pos_r, DeviceArray) else self.ctx.array(
            np.ascontiguousarray(pos_r, dtype=np.float64).reshape(-1, 3))
        return self.execute(pos_d, recon=(dr, alpha))


class SurveyPlan:
    """Survey-geometry plan (hm_survey_plan): main.cpp's preparation
    (main.cpp:109-166) and calcPowerSpectrum (powerspec.hpp:4-91),
    calcTwoPointCorrelationFunction (93-179) or its window version (182-270),
    calcBiSpectrum (bispec.hpp:4-353) or the three-point functions (355-874)."""

    def __init__(self, ctx: Context, n_mesh, boxsize, assign="TSC", n_kbin=20, kmin=0.01, kmax=0.2, ell=0,
                 precision=FP32, prune=True, measure="pk", n_rbin=25, rmin=30.0, rmax=150.0, recon=None,
                 ell1=0, ell2=0):
        """recon: None, or (n_mesh_recon, b1_fid, RG) to reconstruct before measuring.
        bk / 3PCF / window3PCF: the multipole (ell1, ell2, ELL = ell)."""
        self.ctx = ctx
        cfg = SurveyConfig()
        if recon is not None:
            nm, b1, RG = recon
            nm = list(nm) if np.ndim(nm) else [int(nm)] * 3
            cfg.recon = 1
            for a in range(3):
                cfg.n_mesh_recon[a] = int(nm[a])
            cfg.b1 = float(b1)
            cfg.RG = float(RG)
        cfg.measure = SURVEY_MEASURES[measure]
        cfg.n_rbin = int(n_rbin)
        cfg.rmin = float(rmin)
        cfg.rmax = float(rmax)
        cfg.mesh = make_mesh(n_mesh, boxsize, assign)
        cfg.precision = precision
        cfg.n_kbin = int(n_kbin)
        cfg.kmin = float(kmin)
        cfg.kmax = float(kmax)
        cfg.ell = int(ell)
        cfg.ell1 = int(ell1)
        cfg.ell2 = int(ell2)
        cfg.prune = 1 if prune else 0
        self.cfg = cfg
        self.measure = measure
        self.n_kbin = int(n_kbin) if measure in ("pk", "bk") else int(n_rbin)
        self.square = measure in ("bk", "3PCF", "window3PCF")
        self.h = ctx.lib.hm_survey_plan_create(ctx.h, ct.byref(cfg))
        if not self.h:
            ctx.raise_error("hm_survey_plan_create")

    def close(self):
        if self.h:
            self.ctx.lib.hm_survey_plan_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self):
        return self.ctx.lib.hm_survey_plan_bytes(self.h)

    def execute(self, pos_d, w_d, pos_r, w_r):
        """Positions (N, 3) float64 in the survey frame (unshifted) and weights (N,),
        host arrays or DeviceArrays.  Returns (norm * pk_save or norm * xi_save, complex per bin;
        modes or mesh points per bin; info dict)."""
        def dev(a, shape):
            if isinstance(a, DeviceArray):
                return a
            return self.ctx.array(np.ascontiguousarray(a, dtype=np.float64).reshape(shape))

        pd = dev(pos_d, (-1, 3))
        pr = dev(pos_r, (-1, 3))
        nd, nr = pd.shape[0], pr.shape[0]
        wd = dev(w_d, (nd,))
        wr = dev(w_r, (nr,))
        pk = np.zeros((self.n_kbin, self.n_kbin, 2) if self.square else (self.n_kbin, 2))
        nm = np.zeros(self.n_kbin, np.int64)
        info = SurveyInfo()
        self.ctx.check(self.ctx.lib.hm_survey_plan_execute(
            self.ctx.h, self.h, pd.ptr, wd.ptr, nd, pr.ptr, wr.ptr, nr, ct.byref(info),
            pk.ctypes.data_as(ct.POINTER(ct.c_double)), nm.ctypes.data_as(ct.POINTER(ct.c_longlong))),
            "hm_survey_plan_execute")
        d = {k: getattr(info, k) for k in ("wsum_d", "wsum_r", "alpha", "vsurvey", "nbar", "norm")}
        d["shift"] = np.array(info.shift[:])
        return pk[..., 0] + 1j * pk[..., 1], nm, d


def box_config(n_mesh, boxsize, assign="TSC", n_kbin=20, kmin=0.01, kmax=0.2, pk_ells=(0, 2, 4), bk=True,
               precision=FP32, regrid=True, prune=True) -> BoxConfig:
    """The hm_box_config of a P_l (+ B_000) plan."""
    cfg = BoxConfig()
    cfg.mesh = make_mesh(n_mesh, boxsize, assign)
    cfg.precision = precision
    cfg.n_kbin = int(n_kbin)
    cfg.kmin = float(kmin)
    cfg.kmax = float(kmax)
    cfg.n_pk_ells = len(pk_ells)
    for i, l in enumerate(pk_ells):
        cfg.pk_ells[i] = int(l)
    cfg.bk = 1 if bk else 0
    cfg.regrid = 1 if regrid else 0
    cfg.prune = 1 if prune else 0
    cfg.n_rbin = 25
    cfg.rmin = 30.0
    cfg.rmax = 150.0
    return cfg


class Batch:
    """Batched P_l (+ B_000) pipeline with device-side covariance sums
    (hm_batch).  No counterpart in the reference: it runs one mock per process
    and forms the covariance with numpy.cov (calc_data_and_cov.py:284-307)."""

    def __init__(self, ctx: Context, n_mesh, boxsize, max_mocks=1024, cov_with_bk=False, use_graph=True, lanes=1,
                 **kw):
        self.ctx = ctx
        cfg = BatchConfig()
        cfg.box = box_config(n_mesh, boxsize, **kw)
        cfg.max_mocks = int(max_mocks)
        cfg.cov_with_bk = 1 if cov_with_bk else 0
        cfg.use_graph = 1 if use_graph else 0
        cfg.lanes = int(lanes)
        self.cfg = cfg
        self.h = ctx.lib.hm_batch_create(ctx.h, ct.byref(cfg))
        if not self.h:
            ctx.raise_error("hm_batch_create")
        v, c = ct.c_int(), ct.c_int()
        ctx.lib.hm_batch_dims(self.h, ct.byref(v), ct.byref(c))
        self.vec_len, self.cov_len = v.value, c.value

    def close(self):
        if self.h:
            self.ctx.lib.hm_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def push(self, pos: DeviceArray, npart=None):
        n = pos.shape[0] if npart is None else int(npart)
        self.ctx.check(self.ctx.lib.hm_batch_push(self.ctx.h, self.h, pos.ptr, n), "hm_batch_push")

    @property
    def count(self):
        return self.ctx.lib.hm_batch_count(self.h)

    def collect_timing(self):
        self.ctx.check(self.ctx.lib.hm_batch_collect_timing(self.ctx.h, self.h), "hm_batch_collect_timing")

    def vectors(self):
        rows = min(self.count, self.cfg.max_mocks)
        out = np.zeros((max(rows, 0), self.vec_len))
        if rows > 0:
            self.ctx.check(self.ctx.lib.hm_batch_vectors(self.ctx.h, self.h, out.ctypes.data, rows),
                           "hm_batch_vectors")
        return out

    def covariance(self):
        mean = np.zeros(self.cov_len)
        cov = np.zeros((self.cov_len, self.cov_len))
        self.ctx.check(self.ctx.lib.hm_batch_covariance(self.ctx.h, self.h, mean.ctypes.data, cov.ctypes.data),
                       "hm_batch_covariance")
        return mean, cov

    def reset(self):
        self.ctx.check(self.ctx.lib.hm_batch_reset(self.ctx.h, self.h), "hm_batch_reset")

    def autotune(self, pos: DeviceArray, reps=2, npart=None):
        n = pos.shape[0] if npart is None else int(npart)
        buf = ct.create_string_buffer(4096)
        self.ctx.check(self.ctx.lib.hm_batch_autotune(self.ctx.h, self.h, pos.ptr, n, int(reps), buf, 4096),
                       "hm_batch_autotune")
        return buf.value.decode()


def mock_config(n_side, boxsize, amp=2.0e4, k0=0.05, slope=-1.5, kcut=0.5, bao_amp=0.0, bao_period=0.0628,
                f_growth=0.0) -> MockConfig:
    c = MockConfig()
    c.n_side = int(n_side)
    c.boxsize = float(boxsize)
    c.amp, c.k0, c.slope, c.kcut = float(amp), float(k0), float(slope), float(kcut)
    c.bao_amp, c.bao_period, c.f_growth = float(bao_amp), float(bao_period), float(f_growth)
    return c


def mock_white_noise(ctx: Context, cfg: MockConfig, seed) -> DeviceArray:
    out = ctx.empty((cfg.n_side,) * 3, np.float64)
    ctx.check(ctx.lib.hm_mock_white_noise(ctx.h, ct.byref(cfg), int(seed), out.ptr), "hm_mock_white_noise")
    return out


def mock_zeldovich(ctx: Context, cfg: MockConfig, seed, npart, out: DeviceArray = None) -> DeviceArray:
    """Seeded Zel'dovich mock generated on the device: (npart, 3) FP64 positions."""
    if out is None:
        out = ctx.empty((int(npart), 3), np.float64)
    ctx.check(ctx.lib.hm_mock_zeldovich(ctx.h, ct.byref(cfg), int(seed), int(npart), out.ptr), "hm_mock_zeldovich")
    return out


def mock_zeldovich_biased(ctx: Context, cfg: MockConfig, seed, mean_n, b_lag, capacity=None):
    """Lagrangian-biased Zel'dovich tracers generated on the device:
    (positions DeviceArray (n, 3), true shifts DeviceArray (n, 3), n)."""
    cap = int(capacity if capacity is not None else 1.1 * mean_n + 10 * mean_n ** 0.5 + 100)
    pos = ctx.empty((cap, 3), np.float64)
    shift = ctx.empty((cap, 3), np.float64)
    n = ct.c_longlong()
    ctx.check(ctx.lib.hm_mock_zeldovich_biased(ctx.h, ct.byref(cfg), int(seed), float(mean_n), float(b_lag), cap,
                                               pos.ptr, shift.ptr, ct.byref(n)), "hm_mock_zeldovich_biased")
    pos.shape = (n.value, 3)
    shift.shape = (n.value, 3)
    pos.nbytes = shift.nbytes = n.value * 24
    return pos, shift, n.value


def sky_patch(rmin, rmax, ra_min, ra_max, dec_min, dec_max) -> SkyPatch:
    return SkyPatch(float(rmin), float(rmax), float(ra_min), float(ra_max), float(dec_min), float(dec_max))


def mock_sky_randoms(ctx: Context, patch: SkyPatch, seed, n, out: DeviceArray = None) -> DeviceArray:
    """n points uniform in the volume of a sky patch (hm_mock_sky_randoms), [n][3] FP64 on the device."""
    out = ctx.empty((int(n), 3), np.float64) if out is None else out
    ctx.check(ctx.lib.hm_mock_sky_randoms(ctx.h, ct.byref(patch), int(seed), int(n), out.ptr), "hm_mock_sky_randoms")
    return out


class LRecon:
    """Learned correction of the plain reconstruction (hm_lrecon)."""

    def __init__(self, ctx: Context, n_mesh, boxsize, b1=2.0, RG=15.0, psi_scale=10.0, lr=1e-3, init_seed=1,
                 precision=FP32, conv_cache_taps=False):
        self.ctx = ctx
        cfg = LReconConfig()
        cfg.n_mesh, cfg.boxsize, cfg.b1, cfg.RG = int(n_mesh), float(boxsize), float(b1), float(RG)
        cfg.psi_scale, cfg.lr, cfg.init_seed, cfg.precision = float(psi_scale), float(lr), int(init_seed), precision
        cfg.conv_cache_taps = 1 if conv_cache_taps else 0
        self.cfg = cfg
        self.n = int(n_mesh)
        self.h = ctx.lib.hm_lrecon_create(ctx.h, ct.byref(cfg))
        if not self.h:
            ctx.raise_error("hm_lrecon_create")
        self.n_params = ctx.lib.hm_lrecon_n_params(self.h)

    def close(self):
        if self.h:
            self.ctx.lib.hm_lrecon_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx.lib.hm_lrecon_get_params(self.ctx.h, self.h, out.ctypes.data), "hm_lrecon_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self.ctx.check(self.ctx.lib.hm_lrecon_set_params(self.ctx.h, self.h, p.ctypes.data), "hm_lrecon_set_params")

    def features(self, pos: DeviceArray, npart=None):
        n = pos.shape[0] if npart is None else int(npart)
        out = self.ctx.empty((self.n,) * 3 + (4,), np.float32)
        self.ctx.check(self.ctx.lib.hm_lrecon_features(self.ctx.h, self.h, pos.ptr, n, out.ptr), "hm_lrecon_features")
        return out

    def net_forward(self, x):
        dx = self.ctx.array(np.ascontiguousarray(x, dtype=np.float32))
        out = self.ctx.empty((self.n,) * 3 + (4,), np.float32)
        self.ctx.check(self.ctx.lib.hm_lrecon_net_forward(self.ctx.h, self.h, dx.ptr, out.ptr), "hm_lrecon_net_forward")
        return out.download()

    def net_backward(self, x, gout):
        dx = self.ctx.array(np.ascontiguousarray(x, dtype=np.float32))
        dg = self.ctx.array(np.ascontiguousarray(gout, dtype=np.float32))
        grad = np.zeros(self.n_params, np.float32)
        self.ctx.check(self.ctx.lib.hm_lrecon_net_backward(self.ctx.h, self.h, dx.ptr, dg.ptr, grad.ctypes.data),
                       "hm_lrecon_net_backward")
        return grad

    def predict(self, pos: DeviceArray, npart=None):
        n = pos.shape[0] if npart is None else int(npart)
        sp = self.ctx.empty((n, 3), np.float64)
        sq = self.ctx.empty((n, 3), np.float64)
        self.ctx.check(self.ctx.lib.hm_lrecon_predict(self.ctx.h, self.h, pos.ptr, n, sp.ptr, sq.ptr),
                       "hm_lrecon_predict")
        return sp, sq

    def gradient(self, pos: DeviceArray, target: DeviceArray, npart=None, want_gout=False):
        n = pos.shape[0] if npart is None else int(npart)
        l0, l1 = ct.c_double(), ct.c_double()
        grad = np.zeros(self.n_params, np.float32)
        gout = self.ctx.empty((self.n,) * 3 + (4,), np.float32) if want_gout else None
        self.ctx.check(self.ctx.lib.hm_lrecon_gradient(self.ctx.h, self.h, pos.ptr, target.ptr, n, ct.byref(l0),
                                                       ct.byref(l1), grad.ctypes.data, gout.ptr if gout else None),
                       "hm_lrecon_gradient")
        return l0.value, l1.value, grad, gout

    def train_step(self, pos: DeviceArray, target: DeviceArray, npart=None):
        n = pos.shape[0] if npart is None else int(npart)
        l0, l1 = ct.c_double(), ct.c_double()
        self.ctx.check(self.ctx.lib.hm_lrecon_train_step(self.ctx.h, self.h, pos.ptr, target.ptr, n, ct.byref(l0),
                                                         ct.byref(l1)), "hm_lrecon_train_step")
        return l0.value, l1.value
