"""Host-side driver of the OSEM hot path on device memory.

Mirrors the reference driver of hrrt_osem3d (hrrt_osem3d/hrrt_osem3d_span3.cpp):
GetScannerParameter/dependencies -> ``Osem3D(...)``, PrepareOsem3dFiles +
PrepareNormfac -> ``stage``, CalculateNormfac -> ``calculate_normfac``,
Init_Image -> ``init_image``, CalculateOsem3d -> ``calculate_osem3d``, and the
projector interface (hrrt_osem3d/hrrt_osem3d.h:197-211) -> ``forward_proj3d`` /
``back_proj3d``.  Device memory is held in torch tensors; every computation
runs in libhrrt_osem.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

HRRT_MODEL = dict(rings=104, crystal_radius=234.5, plane_sep=1.21875, bin_size=1.21875,
                  def_elements=256, def_angles=288)          # scanner_model.cpp:100
RECON_DEFAULTS = dict(subsets=16, rel_fov=0.95, npskip=3, z_min=6, z_max=200,
                      epsilon1=5e-2, epsilon2=1e-2, epsilon3=1e-10, image_max=1e4)  # S3:231-374


def make_params(**kw) -> _lib.Params:
    d = dict(HRRT_MODEL)
    d.update(RECON_DEFAULTS)
    d.update(dict(x_pixels=0, span=9, maxdel=67))
    d.update(kw)
    return _lib.Params(**d)


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Layout:
    """Dimensions, segment tables and subset order of a geometry (host only, no device).

    Computed by the library's own geometry code (hrrt_query_geometry), so the
    values are the ones a context on the device uses.
    """

    def __init__(self, params: _lib.Params):
        self.params = params
        d = _lib.Dims()
        _lib.call("hrrt_query_geometry", C.byref(params), C.byref(d), *([None] * 6))
        self.dims = d
        self.N, self.views, self.Z, self.Z4 = d.N, d.views, d.z_pixels, d.z4
        self.S, self.spp, self.subsets, self.sviews = d.segmentsize, d.spp, d.subsets, d.sviews
        self.th = d.th_pixels
        order = np.zeros(self.subsets, np.int32)
        vo = np.zeros(self.subsets * self.sviews, np.int32)
        seg = [np.zeros(self.th, np.int32) for _ in range(4)]
        ptrs = [a.ctypes.data_as(C.c_void_p) for a in [order, vo] + seg]
        _lib.call("hrrt_query_geometry", C.byref(params), None, *ptrs)
        self.order = order
        self.vieworder = vo.reshape(self.subsets, self.sviews)
        self.yr_bottom, self.yr_top, self.segment_offset, self.seg_data = seg

    def paircol_index(self):
        """(plane index in [S], padded index in a pair-column) for every internal plane."""
        planes, pidx = [], []
        for s in range(len(self.seg_data)):
            n = int(self.yr_top[s] - self.yr_bottom[s] + 1)
            planes.append(np.arange(n) + self.segment_offset[s])
            pidx.append(np.arange(n) + self.seg_data[s])
        return np.concatenate(planes), np.concatenate(pidx)

    def sino_to_paircol(self, sino):
        """[S][views][N] (internal plane order) -> [views/2][2][N][spp] (host numpy)."""
        planes, pidx = self.paircol_index()
        V, N = self.views, self.N
        out = np.zeros((V // 2, 2, N, self.spp), sino.dtype)
        s2 = sino.reshape(self.S, 2, V // 2, N)          # view = h*V/2 + pair
        out[:, :, :, pidx] = np.transpose(s2[planes], (2, 1, 3, 0))
        return out

    def paircol_to_sino(self, pc):
        """Inverse of sino_to_paircol: [views/2][2][N][spp] -> [S][views][N] (host numpy)."""
        planes, pidx = self.paircol_index()
        V, N = self.views, self.N
        out = np.zeros((self.S, 2, V // 2, N), pc.dtype)
        out[planes] = np.transpose(pc[:, :, :, pidx], (3, 1, 0, 2))
        return out.reshape(self.S, V, N)


class Osem3D(Layout):
    """One reconstruction context (geometry tables resident on the device).

    Every tensor handed to the library is checked first: it must live on this context's
    device, be contiguous and have the dtype and element count the C ABI expects, so a
    wrong buffer raises here instead of faulting the GPU."""

    def __init__(self, device: str = "cuda:0", **params):
        super().__init__(make_params(**params))
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        h = C.c_void_p()
        _lib.call("hrrt_create", C.byref(self.params), C.byref(h))
        self.ctx = h
        d = self.dims
        self.image_floats = int(d.image_floats)
        self.paircol_elems = int(d.paircol_elems)
        self.file_bins = int(d.file_bins)

    def close(self):
        if getattr(self, "ctx", None):
            _lib.lib().hrrt_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- argument checks ----
    def _t(self, t, dtype, numel, name, optional=False):
        """Device pointer of tensor t after checking device, dtype, layout and size."""
        if t is None:
            if optional:
                return C.c_void_p(0)
            raise ValueError(f"{name} is required")
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor, not {type(t).__name__}")
        if t.device != self.device:
            raise ValueError(f"{name} is on {t.device}, the context on {self.device}")
        if t.dtype != dtype:
            raise TypeError(f"{name} must be {dtype}, not {t.dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if numel is not None and t.numel() != numel:
            raise ValueError(f"{name} holds {t.numel()} elements, expected {numel}")
        return C.c_void_p(t.data_ptr())

    @staticmethod
    def _h(a, dtype, numel, name, optional=False):
        """Host pointer of a numpy array or CPU tensor after checking dtype, layout and size."""
        if a is None:
            if optional:
                return C.c_void_p(0)
            raise ValueError(f"{name} is required")
        np_dtype = {torch.int16: np.int16, torch.float32: np.float32}[dtype]
        if isinstance(a, torch.Tensor):
            if a.device.type != "cpu":
                raise ValueError(f"{name} must be in host memory, not on {a.device}")
            if a.dtype != dtype or not a.is_contiguous():
                raise TypeError(f"{name} must be a contiguous {dtype} tensor")
            n, ptr = a.numel(), a.data_ptr()
        elif isinstance(a, np.ndarray):
            if a.dtype != np_dtype or not a.flags["C_CONTIGUOUS"]:
                raise TypeError(f"{name} must be a C-contiguous {np.dtype(np_dtype).name} array")
            n, ptr = a.size, a.ctypes.data
        else:
            raise TypeError(f"{name} must be a numpy array or a CPU tensor, not {type(a).__name__}")
        if n != numel:
            raise ValueError(f"{name} holds {n} elements, expected {numel}")
        return C.c_void_p(ptr)

    def _img(self, t, name="image", n=1):
        return self._t(t, torch.float32, n * self.image_floats, name)

    def _nf_all(self, t):
        return self._t(t, torch.float32, self.subsets * self.image_floats, "normfac")

    def _nf_sub(self, t, isubset):
        self._nf_all(t)
        if not 0 <= int(isubset) < self.subsets:
            raise ValueError(f"subset {isubset} out of range")
        return C.c_void_p(t[int(isubset)].data_ptr())

    def _pc(self, t, dtype, name, n=1):
        return self._t(t, dtype, n * self.paircol_elems, name)

    def _est(self, t, name="prj"):
        return self._t(t, torch.float32, 2 * self.S * self.N, name)

    @staticmethod
    def _frames(images):
        if not isinstance(images, torch.Tensor) or images.dim() < 1:
            raise TypeError("images must be a [frames][X][Y][Z4] tensor")
        return int(images.shape[0])

    # ---- buffers ----
    def image_zeros(self):
        return torch.zeros((self.N, self.N, self.Z4), dtype=torch.float32, device=self.device)

    def estimate_zeros(self):
        return torch.zeros((2 * self.S, self.N), dtype=torch.float32, device=self.device)

    def paircol_zeros(self, dtype=torch.float32):
        return torch.zeros((self.views // 2, 2, self.N, self.spp), dtype=dtype, device=self.device)

    # ---- staging (PrepareOsem3dFiles S3:3876 + PrepareNormfac S3:3726) ----
    def stage(self, prompt_file, delayed_file, norm_file, scatter_file, atten_file):
        fb = self.file_bins
        args = [self._t(prompt_file, torch.int16, fb, "prompt_file"),
                self._t(delayed_file, torch.float32, fb, "delayed_file", optional=True),
                self._t(norm_file, torch.float32, fb, "norm_file", optional=True),
                self._t(scatter_file, torch.float32, fb, "scatter_file", optional=True),
                self._t(atten_file, torch.float32, fb, "atten_file", optional=True)]
        prompt = self.paircol_zeros(torch.int16)
        add = self.paircol_zeros()
        invnorm = self.paircol_zeros()
        _lib.call("hrrt_stage_sinograms", self.ctx, *args, _p(prompt), _p(add), _p(invnorm), _stream())
        return prompt, add, invnorm

    def stage_w012(self, weighting: int, data_file, delayed_file=None, norm_file=None, scatter_file=None,
                   atten_file=None, precorrected: bool = False):
        """prepare_osem3dw0/1/2 + prepare_norm for -W 0 (UW), 1 (AW), 2 (ANW): data_file is the
        short prompt (-p) or true (-t) sinogram, or a float one (-Q, or fully precorrected with
        precorrected=True).  Returns (data, invnorm) in pair-column layout."""
        fb = self.file_bins
        if not isinstance(data_file, torch.Tensor):
            raise TypeError("data_file must be a torch tensor")
        short = data_file.dtype == torch.int16
        args = [self._t(data_file, torch.int16, fb, "data_file") if short else C.c_void_p(0),
                C.c_void_p(0) if short else self._t(data_file, torch.float32, fb, "data_file"),
                self._t(delayed_file, torch.float32, fb, "delayed_file", optional=True),
                self._t(norm_file, torch.float32, fb, "norm_file", optional=True),
                self._t(scatter_file, torch.float32, fb, "scatter_file", optional=True),
                self._t(atten_file, torch.float32, fb, "atten_file", optional=True)]
        data = self.paircol_zeros()
        invnorm = self.paircol_zeros()
        _lib.call("hrrt_stage_sinograms_w012", self.ctx, int(weighting), int(bool(precorrected)), *args, _p(data),
                  _p(invnorm), _stream())
        return data, invnorm

    def stage_host(self, prompt_file, delayed_file, norm_file, scatter_file, atten_file):
        """stage() from flat files in host memory (numpy arrays or CPU tensors, pinned or not),
        copied segment by segment with the staging of each segment overlapping later copies
        (read_flat_* per segment, file_io_processor.cpp:110-259).  Returns device buffers as
        stage() does; the host files may be reused when this returns."""
        fb = self.file_bins
        args = [self._h(prompt_file, torch.int16, fb, "prompt_file"),
                self._h(delayed_file, torch.float32, fb, "delayed_file", optional=True),
                self._h(norm_file, torch.float32, fb, "norm_file", optional=True),
                self._h(scatter_file, torch.float32, fb, "scatter_file", optional=True),
                self._h(atten_file, torch.float32, fb, "atten_file", optional=True)]
        prompt = self.paircol_zeros(torch.int16)
        add = self.paircol_zeros()
        invnorm = self.paircol_zeros()
        _lib.call("hrrt_stage_sinograms_host", self.ctx, *args, _p(prompt), _p(add), _p(invnorm), _stream())
        return prompt, add, invnorm

    def stage_w012_host(self, weighting: int, data_file, delayed_file=None, norm_file=None, scatter_file=None,
                        atten_file=None, precorrected: bool = False):
        """stage_w012() from flat files in host memory (see stage_host)."""
        fb = self.file_bins
        short = (data_file.dtype == torch.int16) if isinstance(data_file, torch.Tensor) else \
            (getattr(data_file, "dtype", None) == np.int16)
        args = [self._h(data_file, torch.int16, fb, "data_file") if short else C.c_void_p(0),
                C.c_void_p(0) if short else self._h(data_file, torch.float32, fb, "data_file"),
                self._h(delayed_file, torch.float32, fb, "delayed_file", optional=True),
                self._h(norm_file, torch.float32, fb, "norm_file", optional=True),
                self._h(scatter_file, torch.float32, fb, "scatter_file", optional=True),
                self._h(atten_file, torch.float32, fb, "atten_file", optional=True)]
        data = self.paircol_zeros()
        invnorm = self.paircol_zeros()
        _lib.call("hrrt_stage_sinograms_w012_host", self.ctx, int(weighting), int(bool(precorrected)), *args,
                  _p(data), _p(invnorm), _stream())
        return data, invnorm

    def norm_zero_mask(self):
        m = np.zeros((self.views // 2, self.N // 2 + 1), np.int8)
        _lib.call("hrrt_get_norm_zero_mask", self.ctx, m.ctypes.data_as(C.c_void_p))
        return m

    def set_norm_zero_mask(self, m):
        m = np.ascontiguousarray(m, dtype=np.int8)
        if m.size != (self.views // 2) * (self.N // 2 + 1):
            raise ValueError(f"norm zero mask holds {m.size} entries, expected {(self.views // 2) * (self.N // 2 + 1)}")
        _lib.call("hrrt_set_norm_zero_mask", self.ctx, m.ctypes.data_as(C.c_void_p))

    # ---- sensitivity (CalculateNormfac S3:3760) ----
    def calculate_normfac(self, invnorm):
        src = self._pc(invnorm, torch.float32, "invnorm")
        nf = torch.zeros((self.subsets, self.N, self.N, self.Z4), dtype=torch.float32, device=self.device)
        _lib.call("hrrt_calculate_normfac", self.ctx, src, _p(nf), _stream())
        return nf

    # ---- image and sensitivity files (SaveImage S3:4016, Init_Image S3:3970, FIO:300-427) ----
    def image_to_flat(self, image, flip: bool = False):
        src = self._img(image)
        flat = torch.zeros((self.Z, self.N, self.N), dtype=torch.float32, device=image.device)
        _lib.call("hrrt_image_to_flat", self.ctx, src, _p(flat), int(flip), _stream())
        return flat

    def image_from_flat(self, flat, flip: bool = False):
        src = self._t(flat, torch.float32, self.Z * self.N * self.N, "flat")
        image = torch.empty((self.N, self.N, self.Z4), dtype=torch.float32, device=flat.device)
        _lib.call("hrrt_image_from_flat", self.ctx, src, _p(image), int(flip), _stream())
        return image

    def write_image(self, image, path: str, flip: bool = False, iterations: int = 0, recontype: int = 3,
                    sw_version: str = "", sw_build_id: str = "", **files):
        """Flat image file plus its Interfile header; files: true_file, prompt_file, delayed_file,
        norm_file, atten_file, scatter_file (paths, for the header)."""
        src = self._img(image)
        enc = {k: (v.encode() if v else None) for k, v in files.items()}
        info = _lib.ImageInfo(recontype=recontype, iterations=iterations, sw_version=sw_version.encode(),
                              sw_build_id=sw_build_id.encode(), **enc)
        _lib.call("hrrt_write_image", self.ctx, src, str(path).encode(), int(flip), C.byref(info), _stream())

    def read_image(self, path: str, flip: bool = False):
        image = torch.empty((self.N, self.N, self.Z4), dtype=torch.float32, device=self.device)
        _lib.call("hrrt_read_image", self.ctx, str(path).encode(), _p(image), int(flip), _stream())
        return image

    def write_normfac(self, normfac, path: str):
        _lib.call("hrrt_write_normfac", self.ctx, self._nf_all(normfac), str(path).encode(), _stream())

    def read_normfac(self, path: str):
        nf = torch.zeros((self.subsets, self.N, self.N, self.Z4), dtype=torch.float32, device=self.device)
        _lib.call("hrrt_read_normfac", self.ctx, str(path).encode(), _p(nf), _stream())
        return nf

    # ---- OSEM (CalculateOsem3d S3:4098) ----
    def init_image(self):
        img = torch.empty((self.N, self.N, self.Z4), dtype=torch.float32, device=self.device)
        _lib.call("hrrt_init_image", self.ctx, _p(img), _stream())
        return img

    def osem_subset(self, image, normfac, prompt, add, isubset: int):
        _lib.call("hrrt_osem_subset", self.ctx, self._img(image), self._nf_sub(normfac, isubset),
                  self._pc(prompt, torch.int16, "prompt"), self._pc(add, torch.float32, "add"), int(isubset),
                  _stream())

    def osem_subset_w012(self, image, normfac, data, isubset: int):
        """Subset update with the UW/AW/ANW ratio (update_estimate_w012)."""
        _lib.call("hrrt_osem_subset_w012", self.ctx, self._img(image), self._nf_sub(normfac, isubset),
                  self._pc(data, torch.float32, "data"), int(isubset), _stream())

    def end_iteration(self, image):
        _lib.call("hrrt_end_iteration", self.ctx, self._img(image), _stream())

    # ---- batched dynamic frames ----
    def osem_subset_frames(self, images, normfac, prompts, adds, isubset: int):
        """images [F][X][Y][Z4], prompts/adds [F][views/2][2][N][spp]: one subset for every frame."""
        n = self._frames(images)
        _lib.call("hrrt_osem_subset_frames", self.ctx, n, self._img(images, "images", n),
                  self._nf_sub(normfac, isubset), self._pc(prompts, torch.int16, "prompts", n),
                  self._pc(adds, torch.float32, "adds", n), int(isubset), _stream())

    def osem_iteration_frames(self, images, normfac, prompts, adds):
        """One iteration (all subsets, graph-replayed, then the end-of-iteration threshold) for every frame,
        on the current stream, which must not be the default one.  As for any stream, work that writes
        the inputs on another stream must be ordered before it (torch: stream.wait_stream(...))."""
        s = torch.cuda.current_stream()
        if s.cuda_stream == 0:
            raise ValueError("graph replay needs a non-default stream (use torch.cuda.stream(...))")
        n = self._frames(images)
        _lib.call("hrrt_osem_iteration_frames", self.ctx, n, self._img(images, "images", n), self._nf_all(normfac),
                  self._pc(prompts, torch.int16, "prompts", n), self._pc(adds, torch.float32, "adds", n),
                  C.c_void_p(s.cuda_stream))

    # ---- precision form: FP16 per-bin inputs, Kahan accumulation, FP64 log-likelihood ----
    @staticmethod
    def to_half(t):
        """float32 device tensor -> FP16 (round to nearest even) as an int16 tensor of the FP16 bits."""
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
            raise TypeError("to_half takes a contiguous float32 CUDA tensor")
        out = torch.empty(t.shape, dtype=torch.int16, device=t.device)
        _lib.call("hrrt_to_half", _p(t), _p(out), int(t.numel()), _stream())
        return out

    @staticmethod
    def to_half_range(t):
        """to_half, also returning (overflowed, underflowed): the finite inputs that became +-inf
        and the non-zero inputs that became FP16 subnormals or zero."""
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
            raise TypeError("to_half_range takes a contiguous float32 CUDA tensor")
        out = torch.empty(t.shape, dtype=torch.int16, device=t.device)
        counts = torch.empty(2, dtype=torch.int64, device=t.device)
        _lib.call("hrrt_to_half_range", _p(t), _p(out), int(t.numel()), _p(counts), _stream())
        c = counts.cpu()
        return out, (int(c[0]), int(c[1]))

    def calculate_normfac_mixed(self, invnorm16):
        src = self._pc(invnorm16, torch.int16, "invnorm16")
        nf = torch.zeros((self.subsets, self.N, self.N, self.Z4), dtype=torch.float32, device=self.device)
        _lib.call("hrrt_calculate_normfac_mixed", self.ctx, src, _p(nf), _stream())
        return nf

    def osem_subset_mixed(self, image, normfac, prompt, add16, isubset: int, invnorm16=None, loglik=None):
        """One subset of the precision form; loglik: float64 device tensor of 1 element (or None)."""
        _lib.call("hrrt_osem_subset_mixed", self.ctx, self._img(image), self._nf_sub(normfac, isubset),
                  self._pc(prompt, torch.int16, "prompt"), self._pc(add16, torch.int16, "add16"),
                  self._t(invnorm16, torch.int16, self.paircol_elems, "invnorm16", optional=True), int(isubset),
                  self._t(loglik, torch.float64, 1, "loglik", optional=True), _stream())

    def osem_subset_frames_mixed(self, images, normfac, prompts, adds16, isubset: int, invnorm16=None, loglik=None):
        n = self._frames(images)
        _lib.call("hrrt_osem_subset_frames_mixed", self.ctx, n, self._img(images, "images", n),
                  self._nf_sub(normfac, isubset), self._pc(prompts, torch.int16, "prompts", n),
                  self._pc(adds16, torch.int16, "adds16", n),
                  self._t(invnorm16, torch.int16, self.paircol_elems, "invnorm16", optional=True), int(isubset),
                  self._t(loglik, torch.float64, n, "loglik", optional=True), _stream())

    def osem_iteration_frames_mixed(self, images, normfac, prompts, adds16, invnorm16=None, loglik=None):
        """loglik: float64 device tensor [subsets][frames] (or None)."""
        s = torch.cuda.current_stream()
        if s.cuda_stream == 0:
            raise ValueError("graph replay needs a non-default stream (use torch.cuda.stream(...))")
        n = self._frames(images)
        _lib.call("hrrt_osem_iteration_frames_mixed", self.ctx, n, self._img(images, "images", n),
                  self._nf_all(normfac), self._pc(prompts, torch.int16, "prompts", n),
                  self._pc(adds16, torch.int16, "adds16", n),
                  self._t(invnorm16, torch.int16, self.paircol_elems, "invnorm16", optional=True),
                  self._t(loglik, torch.float64, self.subsets * n, "loglik", optional=True),
                  C.c_void_p(s.cuda_stream))

    def count_moments(self, image, prompt, add, invnorm):
        """FP64 (sum prompts, sum w*add, sum w*est) over the whole sinogram, w = 1/(N*A)."""
        args = [self._img(image), self._pc(prompt, torch.int16, "prompt"), self._pc(add, torch.float32, "add"),
                self._pc(invnorm, torch.float32, "invnorm")]
        out = torch.zeros(3, dtype=torch.float64, device=self.device)
        _lib.call("hrrt_count_moments", self.ctx, *args, _p(out), _stream())
        return out

    def count_scale(self, image, prompt, add, invnorm):
        """Global count normalisation: the factor that makes the expected total counts of image match the data."""
        m = self.count_moments(image, prompt, add, invnorm).cpu().numpy()
        return float((m[0] - m[1]) / m[2])

    def calculate_osem3d(self, image, normfac, prompt, add, iterations: int = 1, weighting: int = 3):
        """CalculateOsem3d (S3:4098-4341): iterations of every subset in sequence order, then the
        end-of-iteration threshold.  weighting 3 (OP) takes prompt and add; 0, 1 and 2 (UW, AW,
        ANW) take the corrected data as `add` and no prompt."""
        for _ in range(iterations):
            for i in range(self.subsets):
                if weighting == 3:
                    self.osem_subset(image, normfac, prompt, add, int(self.order[i]))
                else:
                    self.osem_subset_w012(image, normfac, add, int(self.order[i]))
            self.end_iteration(image)
        return image

    # ---- kernel timing (HIP events on the launch stream) ----
    KERNELS = ("fp_ratio", "backproject", "rotate_update")

    def set_launch_shape(self, tasks_per_thread: int, workgroups_per_ray: int = 0):
        _lib.call("hrrt_set_launch_shape", self.ctx, int(tasks_per_thread), int(workgroups_per_ray))

    def set_backproject_shape(self, waves_per_workgroup: int):
        _lib.call("hrrt_set_backproject_shape", self.ctx, int(waves_per_workgroup))

    def timing_enable(self, on: bool = True):
        _lib.call("hrrt_timing_enable", self.ctx, int(on))

    def timing_read(self):
        ms = np.zeros(3, np.float64)
        cnt = np.zeros(3, np.int32)
        _lib.call("hrrt_timing_read", self.ctx, ms.ctypes.data_as(C.c_void_p), cnt.ctypes.data_as(C.c_void_p))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.KERNELS)}

    # ---- projector interface (H:197-211) ----
    def forward_proj3d(self, image, view: int, prj=None):
        src = self._img(image)
        prj = self.estimate_zeros() if prj is None else prj
        _lib.call("hrrt_forward_proj3d", self.ctx, src, self._est(prj), int(view), _stream())
        return prj

    def forward_proj3d_noalloc(self, image, view: int, prj):
        """forward_proj3d without any device allocation: the pair-column staging is the library's
        static per-GPU scratch (calls that overlap on one GPU race); prj must be given."""
        _lib.call("hrrt_forward_proj3d_noalloc", self.ctx, self._img(image), self._est(prj), int(view), _stream())
        return prj

    def back_proj3d_incremental(self, prj, ima, view: int):
        _lib.call("hrrt_back_proj3d_incremental", self.ctx, self._est(prj), self._img(ima, "ima"), int(view),
                  _stream())
        return ima

    def set_backproject_schedule(self, nthreads: int):
        """The reference's -T nthreads: per subset, the first npairs - npairs % nthreads view
        pairs are back-projected incrementally (back_proj3d_thread1), the rest directly
        (back_proj3d2); 0 = direct for every pair."""
        _lib.call("hrrt_set_backproject_schedule", self.ctx, int(nthreads))

    # ---- PSF resolution model (-B, psf.cpp) ----
    def set_psf(self, fwhm1: float = 0.0, fwhm2: float = 0.0, ratio: float = 0.0, enable: bool = True):
        """The reference's -B fwhm1,fwhm2,ratio (mm); fwhm1 = 0 selects the HRRT defaults."""
        _lib.call("hrrt_set_psf", self.ctx, int(bool(enable)), float(fwhm1), float(fwhm2), float(ratio))

    def psf_kernel(self, axis: int) -> np.ndarray:
        n = C.c_int(0)
        _lib.call("hrrt_get_psf_kernel", self.ctx, int(axis), None, C.byref(n))
        out = np.zeros(n.value, np.float32)
        _lib.call("hrrt_get_psf_kernel", self.ctx, int(axis), out.ctypes.data_as(C.c_void_p), C.byref(n))
        return out

    def convolve3d(self, image, out=None):
        """convolve3d (psf.cpp:441): blur an image [X][Y][Z4]; out may be image."""
        src = self._img(image)
        out = torch.zeros_like(image) if out is None else out
        _lib.call("hrrt_convolve3d", self.ctx, src, self._img(out, "out"), _stream())
        return out

    @property
    def backproject_schedule(self) -> int:
        v = C.c_int(0)
        _lib.call("hrrt_get_backproject_schedule", self.ctx, C.byref(v))
        return v.value

    def back_proj3d(self, prj, ima, view: int):
        _lib.call("hrrt_back_proj3d", self.ctx, self._est(prj), self._img(ima, "ima"), int(view), _stream())
        return ima

    def back_proj3d_rotframe(self, prj, imagebuf, view: int, incremental: bool = True):
        """back_proj3d_thread1 as the reference driver uses it (PJ:1622-1699): rescales prj in
        place and back-projects the pair into the FOV columns of imagebuf, in the rotated frame
        of `view`, without rotating (incremental=False: the direct projector)."""
        _lib.call("hrrt_back_proj3d_rotframe", self.ctx, self._est(prj), self._img(imagebuf, "imagebuf"),
                  int(view), int(bool(incremental)), _stream())
        return imagebuf

    def rotate_add(self, imagebuf, ima, view: int):
        """One buffer of rotate_image_buf (PJ:1701-1725): ima += rotate(imagebuf, view)."""
        _lib.call("hrrt_rotate_add", self.ctx, self._img(imagebuf, "imagebuf"), self._img(ima, "ima"), int(view),
                  _stream())
        return ima

    def update_estimate_w3(self, est, add, prompt):
        _lib.call("hrrt_update_estimate_w3", self.ctx, self._est(est, "est"), self._est(add, "add"),
                  self._t(prompt, torch.int16, 2 * self.S * self.N, "prompt"), _stream())
        return est

    def update_estimate_w012(self, est, data):
        _lib.call("hrrt_update_estimate_w012", self.ctx, self._est(est, "est"), self._est(data, "data"), _stream())
        return est

    def cal_updateimage(self, image, correction, normfac):
        _lib.call("hrrt_cal_updateimage", self.ctx, self._img(image), self._img(correction, "correction"),
                  self._img(normfac, "normfac"), _stream())

    def normfac_calculate(self, sens):
        src = self._img(sens, "sens")
        out = self.image_zeros()
        _lib.call("hrrt_normfac_calculate", self.ctx, src, _p(out), _stream())
        return out
