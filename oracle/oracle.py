"""ctypes binding of the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class Params(C.Structure):
    _fields_ = [
        ("rings", C.c_int), ("crystal_radius", C.c_float), ("plane_sep", C.c_float),
        ("bin_size", C.c_float), ("def_elements", C.c_int), ("def_angles", C.c_int),
        ("x_pixels", C.c_int), ("span", C.c_int), ("maxdel", C.c_int), ("subsets", C.c_int),
        ("rel_fov", C.c_float), ("npskip", C.c_int), ("z_min", C.c_int), ("z_max", C.c_int),
        ("epsilon1", C.c_float), ("epsilon2", C.c_float), ("epsilon3", C.c_float),
        ("image_max", C.c_float),
    ]


class FLookup(C.Structure):
    _fields_ = [("zzz0", C.c_float), ("zzi", C.c_int)]


class BLookup(C.Structure):
    _fields_ = [("zsa", C.c_int), ("zea", C.c_int), ("yra", C.c_int), ("yr2a", C.c_float)]


class BlutAcc(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("z_start", "yr_start", "z_end", "z_start_pre", "yr_start_pre", "z_end_pre")] + [
        ("coef", C.c_float), ("coef_pre", C.c_float), ("num", C.c_int)]


_PF = C.POINTER(C.c_float)
_PI = C.POINTER(C.c_int)
_PS = C.POINTER(C.c_short)


class Geom(C.Structure):
    _fields_ = [
        ("p", Params),
        ("x_rebin", C.c_int), ("v_rebin", C.c_int),
        ("sino_sampling", C.c_float), ("z_size", C.c_float), ("ring_spacing", C.c_float),
        ("x_size", C.c_float), ("y_size", C.c_float), ("rfov", C.c_float),
        ("z_pixels", C.c_int), ("z4", C.c_int), ("groupmax", C.c_int), ("th_pixels", C.c_int),
        ("N", C.c_int), ("views", C.c_int), ("sviews", C.c_int), ("segmentsize", C.c_int),
        ("zshift", C.c_int), ("z_min", C.c_int), ("z_max", C.c_int),
        ("yr_bottom", _PI), ("yr_top", _PI), ("seg_offset", _PI), ("segment_offset", _PI),
        ("sin_theta", _PF), ("cos_theta", _PF), ("sin_psi", _PF), ("cos_psi", _PF),
        ("cylwiny", _PS), ("vcos", _PF), ("vsin", _PF), ("tantheta", _PF), ("rescalearray", _PF),
        ("flookup", C.POINTER(FLookup)), ("blookup", C.POINTER(BLookup)),
        ("blutacc", C.POINTER(BlutAcc)), ("blutacc_flag", _PI),
        ("vieworder", _PI), ("order", _PS), ("norm_zero_mask", _PS), ("P", C.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so and liboracle64.so with the committed Makefile (make rebuilds them when a
    source changed)."""
    subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []) + ["all"], check=True)
    return _LIB_PATH


_LIB64_PATH = os.path.join(_HERE, "liboracle64.so")
_libs = {}


def _bind(L, real_ptr_is_double):
    vp = C.c_void_p
    G = C.POINTER(Geom)
    L.og_create.restype = C.POINTER(Geom)
    L.og_create.argtypes = [C.POINTER(Params)]
    L.og_destroy.argtypes = [C.POINTER(Geom)]
    L.og_sequence.argtypes = [C.c_int, _PS]
    L.og_prjbuf_floats.restype = C.c_long
    L.og_prjbuf_floats.argtypes = [C.POINTER(Geom)]
    L.og_forward_proj.argtypes = [G, vp, vp, C.c_int, vp]
    L.og_back_proj_direct.argtypes = [G, vp, vp, C.c_int, vp, vp]
    L.og_back_proj_direct_rotframe.argtypes = [G, vp, vp, C.c_int, vp]
    L.og_back_proj_incremental.argtypes = [G, vp, vp, C.c_int, vp]
    L.og_rotate_add.argtypes = [G, vp, vp, C.c_int, C.c_int, C.c_int]
    L.og_ratio_w3.argtypes = [G, vp, vp, vp]
    L.og_ratio_w012.argtypes = [G, vp, vp]
    L.og_update_image.argtypes = [G, vp, vp, vp]
    L.og_normfac_invert.argtypes = [G, vp, vp]
    L.og_init_image.argtypes = [G, vp]
    L.og_end_iteration.argtypes = [G, vp]
    L.og_gather_pair_f.argtypes = [G, vp, C.c_int, vp]
    L.og_gather_pair_s.argtypes = [G, vp, C.c_int, vp]
    L.og_calculate_normfac.argtypes = [G, vp, vp, C.c_int]
    L.og_calculate_sensitivity.argtypes = [G, vp, vp, C.c_int]
    L.og_osem_subset.argtypes = [G, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int]
    L.og_osem_subset_w012.argtypes = [G, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int]
    L.og_psf_create.restype = vp
    L.og_psf_create.argtypes = [G, C.c_float, C.c_float, C.c_float]
    L.og_psf_destroy.argtypes = [vp]
    L.og_psf_kernel.restype = C.c_int
    L.og_psf_kernel.argtypes = [vp, C.c_int, vp]
    L.og_convolve3d.argtypes = [G, vp, vp, vp, vp]
    L.og_calculate_normfac_psf.argtypes = [G, vp, vp, vp, C.c_int]
    L.og_osem_subset_psf.argtypes = [G, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int]
    if real_ptr_is_double:
        return L
    # staging and flat files: float files, built into the float oracle only
    L.og_read_flat_float.argtypes = [G, vp, vp, C.c_int, C.c_float, C.c_int, C.c_int]
    L.og_read_flat_short.argtypes = [G, vp, vp, C.c_int]
    L.og_prepare_norm.argtypes = [G, vp, vp, vp]
    L.og_prepare_w3.argtypes = [G, vp, vp, vp, vp, vp, vp, vp]
    L.og_prepare_norm_w.argtypes = [G, C.c_int, vp, vp, vp]
    L.og_prepare_w012.restype = C.c_int
    L.og_prepare_w012.argtypes = [G, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp]
    L.og_image_to_flat.argtypes = [G, vp, vp, C.c_int]
    L.og_image_from_flat.argtypes = [G, vp, vp, C.c_int]
    return L


def lib(double: bool = False):
    """The float oracle (liboracle.so), or its FP64 build (liboracle64.so) with double=True."""
    path = _LIB64_PATH if double else _LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        _libs[path] = _bind(C.CDLL(path), double)
    return _libs[path]


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


HRRT = dict(rings=104, crystal_radius=234.5, plane_sep=1.21875, bin_size=1.21875,
            def_elements=256, def_angles=288)
DEFAULTS = dict(subsets=16, rel_fov=0.95, npskip=3, z_min=6, z_max=200,
                epsilon1=5e-2, epsilon2=1e-2, epsilon3=1e-10, image_max=1e4)


def make_params(**kw) -> Params:
    d = dict(HRRT)
    d.update(DEFAULTS)
    d.update(dict(x_pixels=0, span=9, maxdel=67))
    d.update(kw)
    return Params(**d)


class Oracle:
    """Owns one og_geom and exposes the restated reference routines on numpy arrays.

    double=True binds the FP64 build (liboracle64.so): the same routines with double values
    (image, sinograms, buffers, sums) and the same float geometry, positions and factors.
    Value arrays must then be float64; staging and flat files exist in the float build only."""

    def __init__(self, double=False, **kw):
        self.double = bool(double)
        self.real = np.float64 if self.double else np.float32
        self.L = lib(self.double)
        self.params = make_params(**kw)
        self.gp = self.L.og_create(C.byref(self.params))
        if not self.gp:
            raise ValueError("invalid geometry")
        g = self.gp.contents
        self.g = g
        self.N, self.Z, self.Z4, self.S = g.N, g.z_pixels, g.z4, g.segmentsize
        self.views, self.th, self.gm, self.zs = g.views, g.th_pixels, g.groupmax, g.zshift
        self.sviews = g.sviews
        self.subsets = self.params.subsets

    def close(self):
        if getattr(self, "psf", None):
            self.L.og_psf_destroy(self.psf)
            self.psf = None
        if self.gp:
            self.L.og_destroy(self.gp)
            self.gp = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _v(self, a):
        """Pointer to a value array (image, sinogram, buffer) of this build's value type."""
        if a is None:
            return None
        if a.dtype != self.real:
            raise TypeError(f"value arrays of this oracle are {np.dtype(self.real).name}, not {a.dtype}")
        return _ptr(a)

    # ---- tables ----
    def arr(self, name, n, dtype=None):
        p = getattr(self.g, name)
        a = np.ctypeslib.as_array(p, shape=(n,)).copy()
        return a if dtype is None else a.astype(dtype)

    @property
    def yr_bottom(self):
        return self.arr("yr_bottom", self.th)

    @property
    def yr_top(self):
        return self.arr("yr_top", self.th)

    @property
    def segment_offset(self):
        return self.arr("segment_offset", self.th)

    @property
    def seg_offset(self):
        return self.arr("seg_offset", self.th)

    @property
    def cylwiny(self):
        return self.arr("cylwiny", 2 * self.N).reshape(self.N, 2)

    @property
    def sin_psi(self):
        return self.arr("sin_psi", self.views)

    @property
    def cos_psi(self):
        return self.arr("cos_psi", self.views)

    @property
    def sin_theta(self):
        return self.arr("sin_theta", self.th)

    @property
    def cos_theta(self):
        return self.arr("cos_theta", self.th)

    @property
    def tantheta(self):
        return self.arr("tantheta", self.gm * self.Z).reshape(self.gm, self.Z)

    @property
    def rescalearray(self):
        return self.arr("rescalearray", self.gm * self.Z).reshape(self.gm, self.Z)

    @property
    def order(self):
        return self.arr("order", self.subsets)

    @property
    def vieworder(self):
        return self.arr("vieworder", self.subsets * self.sviews).reshape(self.subsets, self.sviews)

    @property
    def norm_zero_mask(self):
        return self.arr("norm_zero_mask", (self.views // 2 + 1) * (self.N // 2 + 1)).reshape(
            self.views // 2 + 1, self.N // 2 + 1)

    def set_norm_zero_mask(self, m):
        m = np.ascontiguousarray(m, dtype=np.int16)
        C.memmove(self.g.norm_zero_mask, m.ctypes.data, m.nbytes)

    def flookup(self):
        n = self.N * self.gm * 2 * self.zs
        a = np.ctypeslib.as_array(C.cast(self.g.flookup, C.POINTER(C.c_int32)), shape=(2 * n,)).copy()
        frac = a[0::2].view(np.float32).reshape(self.N, self.gm, 2 * self.zs)
        zi = a[1::2].reshape(self.N, self.gm, 2 * self.zs)
        return frac, zi

    def blookup(self):
        n = self.N * self.gm
        a = np.ctypeslib.as_array(C.cast(self.g.blookup, C.POINTER(C.c_int32)), shape=(4 * n,)).copy().reshape(n, 4)
        return (a[:, 0].reshape(self.N, self.gm), a[:, 1].reshape(self.N, self.gm),
                a[:, 2].reshape(self.N, self.gm), a[:, 3].view(np.float32).reshape(self.N, self.gm))

    # ---- buffers ----
    def image_zeros(self):
        return np.zeros((self.N, self.N, self.Z4), self.real)

    def prj_zeros(self):
        return np.zeros((2 * self.S, self.N), self.real)

    def prjbuf(self):
        return np.zeros(self.L.og_prjbuf_floats(self.gp), self.real)

    def sino_zeros(self, dtype=np.float32):
        return np.zeros((self.S, self.views, self.N), dtype)

    # ---- projectors ----
    def forward(self, image, view, prj=None, prjbuf=None):
        prj = self.prj_zeros() if prj is None else prj
        prjbuf = self.prjbuf() if prjbuf is None else prjbuf
        self.L.og_forward_proj(self.gp, self._v(image), self._v(prj), view, self._v(prjbuf))
        return prj

    def back_direct(self, prj, view, ima=None):
        prj = prj.copy()
        ima = self.image_zeros() if ima is None else ima
        ib = self.image_zeros()
        self.L.og_back_proj_direct(self.gp, self._v(prj), self._v(ima), view, self._v(ib), self._v(self.prjbuf()))
        return ima

    def back_direct_rotframe(self, prj, view):
        prj = prj.copy()
        ib = self.image_zeros()
        self.L.og_back_proj_direct_rotframe(self.gp, self._v(prj), self._v(ib), view, self._v(self.prjbuf()))
        return ib

    def back_incremental_rotframe(self, prj, view):
        prj = prj.copy()
        ib = self.image_zeros()
        self.L.og_back_proj_incremental(self.gp, self._v(prj), self._v(ib), view, self._v(self.prjbuf()))
        return ib

    def rotate_add(self, imagebuf, view, out=None):
        out = self.image_zeros() if out is None else out
        self.L.og_rotate_add(self.gp, self._v(imagebuf), self._v(out), view, 0, self.N)
        return out

    # ---- elementwise ----
    def ratio_w3(self, est, add, prompt):
        est = np.ascontiguousarray(est, self.real).copy()
        self.L.og_ratio_w3(self.gp, self._v(est), self._v(np.ascontiguousarray(add, self.real)),
                           _ptr(np.ascontiguousarray(prompt, np.int16)))
        return est

    def update_image(self, image, correction, normfac):
        self.L.og_update_image(self.gp, self._v(image), self._v(correction), self._v(normfac))

    def normfac_invert(self, sens):
        out = self.image_zeros()
        self.L.og_normfac_invert(self.gp, self._v(sens), self._v(out))
        return out

    def init_image(self):
        img = self.image_zeros()
        self.L.og_init_image(self.gp, self._v(img))
        return img

    def end_iteration(self, image):
        self.L.og_end_iteration(self.gp, self._v(image))

    # ---- staging ----
    def read_flat_float(self, file, theta, def_val=0.0, norm_flag=0):
        sw = theta + 1 if theta % 2 == 1 else (0 if theta == 0 else theta - 1)
        npl = int(self.yr_top[sw] - self.yr_bottom[sw] + 1)
        out = np.zeros((npl, self.views, self.N), np.float32)
        self.L.og_read_flat_float(self.gp, _ptr(file), _ptr(out), theta, def_val, norm_flag, 0 if file is None else 1)
        return out

    def read_flat_short(self, file, theta):
        sw = theta + 1 if theta % 2 == 1 else (0 if theta == 0 else theta - 1)
        npl = int(self.yr_top[sw] - self.yr_bottom[sw] + 1)
        out = np.zeros((npl, self.views, self.N), np.int16)
        self.L.og_read_flat_short(self.gp, _ptr(file), _ptr(out), theta)
        return out

    def prepare_norm(self, norm_file, atten_file):
        out = self.sino_zeros()
        self.L.og_prepare_norm(self.gp, _ptr(norm_file), _ptr(atten_file), _ptr(out))
        return out

    def prepare_w3(self, prompt_file, delayed_file, norm_file, scatter_file, atten_file):
        prompt = self.sino_zeros(np.int16)
        add = self.sino_zeros()
        self.L.og_prepare_w3(self.gp, _ptr(prompt_file), _ptr(delayed_file), _ptr(norm_file),
                             _ptr(scatter_file), _ptr(atten_file), _ptr(prompt), _ptr(add))
        return prompt, add

    def ratio_w012(self, est, data):
        out = np.ascontiguousarray(est, self.real).copy()
        self.L.og_ratio_w012(self.gp, self._v(out), self._v(np.ascontiguousarray(data, self.real)))
        return out

    def prepare_norm_w(self, weighting, norm_file, atten_file):
        out = self.sino_zeros()
        self.L.og_prepare_norm_w(self.gp, int(weighting), _ptr(norm_file), _ptr(atten_file), _ptr(out))
        return out

    def prepare_w012(self, weighting, data_file, delayed_file=None, norm_file=None, scatter_file=None,
                     atten_file=None, precorrected=False):
        """data_file: int16 (prompts or trues) or float32 (float or precorrected sinogram)."""
        out = self.sino_zeros()
        d16 = data_file if data_file.dtype == np.int16 else None
        df = data_file if data_file.dtype == np.float32 else None
        rc = self.L.og_prepare_w012(self.gp, int(weighting), int(bool(precorrected)), _ptr(d16), _ptr(df),
                                    _ptr(delayed_file), _ptr(norm_file), _ptr(scatter_file), _ptr(atten_file),
                                    _ptr(out))
        if rc != 0:
            raise ValueError("unsupported input combination for prepare_w012")
        return out

    def gather_pair(self, sino, view):
        if sino.dtype == np.int16:
            out = np.zeros((2 * self.S, self.N), np.int16)
            self.L.og_gather_pair_s(self.gp, _ptr(sino), view, _ptr(out))
        else:
            out = self.prj_zeros()
            self.L.og_gather_pair_f(self.gp, self._v(sino), view, self._v(out))
        return out

    # ---- drivers ----
    def calculate_normfac(self, invnorm, nthreads=0):
        nf = np.zeros((self.subsets, self.N, self.N, self.Z4), self.real)
        self.L.og_calculate_normfac(self.gp, self._v(invnorm), self._v(nf), nthreads)
        return nf

    def calculate_sensitivity(self, invnorm, nthreads=0):
        """The sensitivity images before inversion, [subsets][X][Y][Z4] (invnorm is rescaled in place)."""
        sens = np.zeros((self.subsets, self.N, self.N, self.Z4), self.real)
        self.L.og_calculate_sensitivity(self.gp, self._v(invnorm), self._v(sens), nthreads)
        return sens

    def osem_subset(self, image, correction, normfac_subset, add, prompt, isubset, nthreads=0, workers=1):
        self.L.og_osem_subset(self.gp, self._v(image), self._v(correction), self._v(normfac_subset),
                              self._v(add), _ptr(prompt), isubset, nthreads, workers)

    def osem_subset_w012(self, image, correction, normfac_subset, data, isubset, nthreads=0, workers=1):
        self.L.og_osem_subset_w012(self.gp, self._v(image), self._v(correction), self._v(normfac_subset),
                                   self._v(data), isubset, nthreads, workers)

    # ---- flat image files ----
    def image_to_flat(self, image, flip=False):
        flat = np.zeros((self.Z, self.N, self.N), np.float32)
        self.L.og_image_to_flat(self.gp, _ptr(np.ascontiguousarray(image, np.float32)), _ptr(flat), int(flip))
        return flat

    def image_from_flat(self, flat, flip=False):
        image = self.image_zeros()
        self.L.og_image_from_flat(self.gp, _ptr(np.ascontiguousarray(flat, np.float32)), _ptr(image), int(flip))
        return image

    # ---- PSF resolution modelling (-B) ----
    def set_psf(self, fwhm1=0.0, fwhm2=0.0, ratio=0.0):
        """Build the x, y, z kernels (make1d_psf_filter); fwhm1 == 0 selects the HRRT defaults."""
        if getattr(self, "psf", None):
            self.L.og_psf_destroy(self.psf)
        self.psf = self.L.og_psf_create(self.gp, fwhm1, fwhm2, ratio)
        if not self.psf:
            raise ValueError("invalid PSF parameters")

    def psf_kernel(self, axis):
        n = self.L.og_psf_kernel(self.psf, axis, None)
        out = np.zeros(n, np.float32)
        self.L.og_psf_kernel(self.psf, axis, _ptr(out))
        return out

    def convolve3d(self, image, out=None):
        out = self.image_zeros() if out is None else out
        buf = self.image_zeros()
        self.L.og_convolve3d(self.gp, self.psf, self._v(image), self._v(out), self._v(buf))
        return out

    def calculate_normfac_psf(self, invnorm, nthreads=0):
        nf = np.zeros((self.subsets, self.N, self.N, self.Z4), self.real)
        self.L.og_calculate_normfac_psf(self.gp, self.psf, self._v(invnorm), self._v(nf), nthreads)
        return nf

    def osem_subset_psf(self, image, correction, normfac_subset, add, prompt, isubset, nthreads=0, workers=1):
        self.L.og_osem_subset_psf(self.gp, self.psf, self._v(image), self._v(correction),
                                  self._v(normfac_subset), self._v(add), _ptr(prompt), isubset, nthreads, workers)


NCRYS = 8 * 2 * 72 * 104          # crystals, [layer][y][head][x]
NXD = 144                         # crystals of a head row over both layers


class Sol(C.Structure):
    _fields_ = [("nsino", C.c_int), ("d", C.c_float), ("z", C.c_float)]


def _delays_lib():
    L = lib()
    if not getattr(L, "_delays_bound", False):
        vp = C.c_void_p
        L.og_delays_create.restype = vp
        L.og_delays_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]
        L.og_delays_destroy.argtypes = [vp]
        L.og_delays_mpair.restype = C.c_int
        L.og_delays_mpair.argtypes = [C.c_int, _PI, _PI]
        L.og_delays_dims.argtypes = [vp, vp]
        L.og_delays_crystal.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.og_delays_phy_to_pro.argtypes = [vp, vp, vp, C.POINTER(Sol)]
        L.og_delays_sol.argtypes = [vp, C.c_int, vp, vp, vp]
        L.og_delays_tables.argtypes = [vp, vp, vp, vp]
        L.og_delays_drate.argtypes = [vp, C.c_float, vp]
        L.og_delays_singles.restype = C.c_int
        L.og_delays_singles.argtypes = [vp, C.c_float, C.c_float, vp]
        L.og_delays_smooth.restype = C.c_longlong
        L.og_delays_smooth.argtypes = [vp, vp, C.c_float, C.c_float, vp, C.c_int, vp]
        L._delays_bound = True
    return L


class Delays:
    """Smoothed randoms (gen_delays) restated on the CPU: geometry, LUT, singles, smoothing."""

    def __init__(self, span=9, maxrd=67, lr_type=0, pitch=0.0, diam=0.0, thick=0.0):
        self.L = _delays_lib()
        self.g = self.L.og_delays_create(span, maxrd, lr_type, pitch, diam, thick)
        if not self.g:
            raise ValueError("invalid gen_delays parameters")
        d = np.zeros(6, np.int32)
        self.L.og_delays_dims(self.g, _ptr(d))
        self.nprojs, self.nviews, self.nplanes, self.nsegs, self.nsino, self.span = (int(v) for v in d)

    def close(self):
        if self.g:
            self.L.og_delays_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def mpair(mp):
        a, b = C.c_int(), C.c_int()
        if _delays_lib().og_delays_mpair(mp, C.byref(a), C.byref(b)) != 0:
            raise ValueError(mp)
        return a.value, b.value

    def crystal(self, head, layer, x, y):
        pos = np.zeros(3, np.float32)
        self.L.og_delays_crystal(self.g, head, layer, x, y, _ptr(pos))
        return pos

    def phy_to_pro(self, a, b):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        s = Sol()
        self.L.og_delays_phy_to_pro(self.g, _ptr(a), _ptr(b), C.byref(s))
        return s.nsino, np.float32(s.d), np.float32(s.z)

    def sol(self, mp):
        n = np.zeros((NXD, NXD), np.int32)
        d = np.zeros((NXD, NXD), np.float32)
        z = np.zeros((NXD, NXD), np.float32)
        self.L.og_delays_sol(self.g, mp, _ptr(n), _ptr(d), _ptr(z))
        return n, d, z

    def tables(self):
        zpos = np.zeros(104, np.float32)
        zpos2 = np.zeros(104, np.float32)
        segplane = np.zeros((63, 207), np.int16)
        self.L.og_delays_tables(self.g, _ptr(zpos), _ptr(zpos2), _ptr(segplane))
        return zpos, zpos2, segplane

    @staticmethod
    def drate(srate, tau):
        srate = np.ascontiguousarray(srate, np.float32)
        out = np.zeros(NCRYS, np.float32)
        _delays_lib().og_delays_drate(_ptr(srate), tau, _ptr(out))
        return out

    @staticmethod
    def singles(dcoins, tau, dt):
        dcoins = np.ascontiguousarray(dcoins, np.int32)
        assert dcoins.size == NCRYS
        out = np.zeros(NCRYS, np.float32)
        it = _delays_lib().og_delays_singles(_ptr(dcoins), tau, dt, _ptr(out))
        return out, it

    def smooth(self, csings, tau=6e-9, ftime=1.0, mps=None):
        """Returns (delays [nplanes][nviews][nprojs], crystal pairs visited)."""
        csings = np.ascontiguousarray(csings, np.float32)
        assert csings.size == NCRYS
        out = np.zeros((self.nplanes, self.nviews, self.nprojs), np.float32)
        m = None if mps is None else np.ascontiguousarray(mps, np.int32)
        pairs = self.L.og_delays_smooth(self.g, _ptr(csings), tau, ftime, _ptr(m), 0 if m is None else m.size, _ptr(out))
        return out, int(pairs)


LM_PACKET_WORDS = 2 * 1024 * 1024
NBLOCKS = 936


def _lm_lib():
    L = _delays_lib()
    if not getattr(L, "_lm_bound", False):
        vp = C.c_void_p
        L.og_lm_rebin.restype = C.c_longlong
        L.og_lm_rebin.argtypes = [vp] + [C.c_int] * 7
        L.og_lm_histogram.argtypes = [vp, vp, C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      vp, vp, vp, vp, vp, vp, vp, vp]
        L._lm_bound = True
    return L


def lm_histogram(D, words, mode=0, start_ms=0, end_ms=0, ignore_border=False, doi=True, delayed_bytes=True):
    """List-mode histogram of a word stream with the tables of Delays D.
    Returns dict(sino int16 [nplanes][nviews][nprojs] (x2 planes for mode 1 without a byte array),
    delayed uint8 or None, coin_p, coin_d uint32 [crystals], stats int64 [10], blocks)."""
    L = _lm_lib()
    words = np.ascontiguousarray(words, np.uint32)
    S = D.nplanes * D.nsino
    halves = 2 if (mode == 1 and not delayed_bytes) else 1
    sino = np.zeros(halves * S, np.int16)
    delayed = np.zeros(S, np.uint8) if (mode == 1 and delayed_bytes) else None
    coin_p = np.zeros(NCRYS, np.uint32)
    coin_d = np.zeros(NCRYS, np.uint32)
    stats = np.zeros(10, np.int64)
    btot = np.zeros(NBLOCKS, np.int64)
    bns = np.zeros(NBLOCKS, np.int32)
    blast = np.zeros(NBLOCKS, np.int32)
    L.og_lm_histogram(D.g, _ptr(words), words.size, mode, start_ms, end_ms, int(ignore_border), int(doi),
                      _ptr(sino), _ptr(delayed), _ptr(coin_p), _ptr(coin_d), _ptr(stats), _ptr(btot), _ptr(bns),
                      _ptr(blast))
    return dict(sino=sino.reshape(halves * D.nplanes, D.nviews, D.nprojs), delayed=delayed, coin_p=coin_p,
                coin_d=coin_d, stats=stats, block_total=btot, block_nsamples=bns, block_last=blast)


def lm_rebin(D, mp, alayer, ax, ay, blayer, bx, by):
    return int(_lm_lib().og_lm_rebin(D.g, mp, alayer, ax, ay, blayer, bx, by))
