"""List-mode histogramming on device memory (lmhistogram_mp).

``histogram`` histograms one frame of a 64-bit list-mode stream (held as 32-bit words in a
device tensor) with the module-pair LUT of a :class:`~hrrt_reconstruction_amd.randoms.GenDelays`
context, as the reference's rebinner and gen_delays share the lor_sinogram tables
(lmhistogram_mp/LM_Rebinner_mp.cpp:138-142).  The coincidence maps it returns are the .ch
halves that ``GenDelays.singles`` / ``GenDelays.gen_delays`` take.  Every computation runs
in libhrrt_osem.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .randoms import CRYSTALS, GenDelays

MODES = {"trues": 0, "prompts_randoms": 1, "prompts": 2}     # HISTOGRAM_MODE histogram_mp.hpp:41-46


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def histogram(g: GenDelays, words: torch.Tensor, mode: int = 0, start_ms: int = 0, end_ms: int = 0,
              ignore_border: bool = False, no_doi: bool = False, delayed_bytes: bool = True,
              coincidence_maps: bool = True):
    """Returns (sino int16 [nplanes][nviews][nprojs] (2x planes in mode 1 without a byte array),
    delayed uint8 or None, coin_p, coin_d uint32 [crystals] or None, stats dict)."""
    if words.dtype not in (torch.int32, torch.uint32) or not words.is_cuda:
        raise ValueError("words must be an int32/uint32 device tensor")
    words = words.contiguous()
    if mode not in (0, 1, 2):
        raise ValueError("mode must be 0 (trues), 1 (prompts and randoms) or 2 (prompts)")
    halves = 2 if (mode == 1 and not delayed_bytes) else 1
    dev = words.device
    sino = torch.empty((halves * g.nplanes, g.nviews, g.nprojs), dtype=torch.int16, device=dev)
    delayed = (torch.empty((g.nplanes, g.nviews, g.nprojs), dtype=torch.uint8, device=dev)
               if (mode == 1 and delayed_bytes) else None)
    coin_p = torch.empty(CRYSTALS, dtype=torch.int32, device=dev) if coincidence_maps else None
    coin_d = torch.empty(CRYSTALS, dtype=torch.int32, device=dev) if coincidence_maps else None
    opt = _lib.LmOptions(mode, start_ms, end_ms, int(ignore_border), int(no_doi))
    st = _lib.LmStats()
    _lib.call("hrrt_lm_histogram", g._h, _p(words), words.numel(), C.byref(opt), _p(sino), _p(delayed),
              _p(coin_p), _p(coin_d), C.byref(st), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    stats = {n: getattr(st, n) for n in ("prompts", "randoms", "events", "tags", "time_tags", "rejected",
                                         "transmission", "first_word", "end_word", "last_time_ms")}
    stats["block_total"] = np.ctypeslib.as_array(st.block_total).copy()
    stats["block_nsamples"] = np.ctypeslib.as_array(st.block_nsamples).copy()
    stats["block_last"] = np.ctypeslib.as_array(st.block_last).copy()
    return sino, delayed, coin_p, coin_d, stats


def coincidence_histogram(coin_p: torch.Tensor, coin_d: torch.Tensor) -> np.ndarray:
    """The .ch contents (prompts then delayed, int32) from the two coincidence maps."""
    return torch.cat([coin_p, coin_d]).cpu().numpy().astype(np.int32)
