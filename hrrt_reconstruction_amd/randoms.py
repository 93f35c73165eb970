"""Smoothed randoms from a coincidence histogram on device memory (gen_delays_lib).

Mirrors gen_delays (gen_delays_lib/gen_delays_lib.cpp:337-548): the constructor does
init_geometry_hrrt + init_segment_info + the LUT (read from the rebinner LUT file, or built
as hrrt_rebinner_lut builds it); ``singles`` is compute_csings_from_drates; ``smooth`` is
the 20 compute_delays calls and the plane-major output; ``gen_delays`` chains both
(gen_delays(2, ...)); ``write`` is gen_delays(1, ...)'s output file.  Every computation runs
in libhrrt_osem.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

CRYSTALS = 8 * 2 * 72 * 104      # [layer][ring][head][x]


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class GenDelays:
    def __init__(self, span: int = 9, maxrd: int = 67, lr_type: int = 0, pitch: float = 0.0, diam: float = 0.0,
                 thick: float = 0.0, tau: float = 0.0, lut_file: str | None = None):
        self._h = C.c_void_p()
        params = _lib.DelaysParams(span, maxrd, lr_type, pitch, diam, thick, tau,
                                   None if lut_file is None else lut_file.encode())
        _lib.call("hrrt_delays_create", C.byref(params), C.byref(self._h))
        v = [C.c_int() for _ in range(4)]
        pairs = C.c_longlong()
        _lib.call("hrrt_delays_get_dims", self._h, *[C.byref(x) for x in v], C.byref(pairs))
        self.nprojs, self.nviews, self.nplanes, self.nsegs = (x.value for x in v)
        self.nsino = self.nprojs * self.nviews
        self.pairs_full = pairs.value
        self.span, self.maxrd = span, maxrd

    def close(self):
        if self._h:
            _lib.lib().hrrt_delays_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lut(self, mp: int):
        """(nsino, d, z) of module pair mp (1..20), each [144][144]."""
        n = np.zeros((144, 144), np.int32)
        d = np.zeros((144, 144), np.float32)
        z = np.zeros((144, 144), np.float32)
        _lib.call("hrrt_delays_get_lut", self._h, mp, n.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p),
                  z.ctypes.data_as(C.c_void_p))
        return n, d, z

    def save_lut(self, path: str):
        _lib.call("hrrt_delays_save_lut", self._h, path.encode())

    def singles(self, coins: np.ndarray, scan_duration: float):
        """Crystal singles from a .ch histogram (int32 [2][crystals]); returns (singles, iterations)."""
        coins = np.ascontiguousarray(coins, np.int32).reshape(-1)
        if coins.size != 2 * CRYSTALS:
            raise ValueError(f"coincidence histogram must hold {2 * CRYSTALS} counts")
        out = np.zeros(CRYSTALS, np.float32)
        it = C.c_int()
        _lib.call("hrrt_delays_singles", self._h, coins.ctypes.data_as(C.c_void_p), scan_duration,
                  out.ctypes.data_as(C.c_void_p), C.byref(it))
        return out, it.value

    def empty(self):
        return torch.empty((self.nplanes, self.nviews, self.nprojs), dtype=torch.float32, device="cuda")

    def smooth(self, singles: torch.Tensor, scan_duration: float = 0.0, out: torch.Tensor | None = None):
        """Smoothed randoms [nplanes][nviews][nprojs] from device singles [crystals]."""
        if singles.dtype != torch.float32 or not singles.is_cuda or singles.numel() != CRYSTALS:
            raise ValueError("singles must be a float32 device tensor of 119808 crystals")
        singles = singles.contiguous()
        if out is None:
            out = self.empty()
        _lib.call("hrrt_delays_smooth", self._h, _p(singles), scan_duration, _p(out), _stream())
        return out

    def gen_delays(self, coins: np.ndarray, scan_duration: float, out: torch.Tensor | None = None):
        """gen_delays(2, ...): returns (delays on the device, singles, iterations)."""
        coins = np.ascontiguousarray(coins, np.int32).reshape(-1)
        if coins.size != 2 * CRYSTALS:
            raise ValueError(f"coincidence histogram must hold {2 * CRYSTALS} counts")
        if out is None:
            out = self.empty()
        sg = np.zeros(CRYSTALS, np.float32)
        it = C.c_int()
        _lib.call("hrrt_gen_delays", self._h, coins.ctypes.data_as(C.c_void_p), scan_duration, _p(out),
                  sg.ctypes.data_as(C.c_void_p), C.byref(it), _stream())
        return out, sg, it.value

    def write(self, delays: torch.Tensor, path: str):
        _lib.call("hrrt_delays_write", self._h, _p(delays), path.encode(), _stream())
