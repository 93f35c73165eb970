"""hrrt_forward_proj3d_noalloc: the forward projection with its pair-column staging in the
library's static per-GPU scratch, so that the call allocates no device memory.

Through the C ABI it must give hrrt_forward_proj3d's result bit for bit, on two calls in a
row -- span 3 at 256x256, then a smaller geometry (16 rings, 96 bins) sharing the same
scratch -- and leave the device's free memory unchanged.  A geometry whose pair-column
exceeds the scratch is refused with HRRT_ERR_ARG.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import from_dev, random_image, to_dev

pytestmark = pytest.mark.gpu

SENTINEL = np.float32(-7.25)


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


def _call(g, fn, img, view, prj):
    from hrrt_reconstruction_amd import _lib
    import torch
    return getattr(_lib.lib(), fn)(g.ctx, C.c_void_p(img.data_ptr()), C.c_void_p(prj.data_ptr()), int(view),
                                   C.c_void_p(torch.cuda.current_stream().cuda_stream))


def test_noalloc_forward_matches_allocating_entry_point(torch_dev, oracle_lib):
    from hrrt_reconstruction_amd import _lib
    from hrrt_reconstruction_amd.osem import Osem3D
    torch = torch_dev
    for geom, views in (("c2", (0, 100)), ("tiny", (5, 47))):       # the second call is smaller
        print(f"noalloc forward geom={geom} views={views}", flush=True)
        o = oracle_lib.Oracle(**GEOMS[geom])
        g = Osem3D(**GEOMS[geom])
        img = to_dev(random_image(o, seed=31))
        for v in views:
            ref = to_dev(np.full((2 * g.S, g.N), SENTINEL, np.float32))
            out = to_dev(np.full((2 * g.S, g.N), SENTINEL, np.float32))
            assert _call(g, "hrrt_forward_proj3d", img, v, ref) == _lib.HRRT_OK
            assert _call(g, "hrrt_forward_proj3d_noalloc", img, v, out) == _lib.HRRT_OK
            a, b = from_dev(out), from_dev(ref)
            assert np.array_equal(a, b), f"{geom} view {v}: max |diff| {np.abs(a - b).max()}"
            assert np.count_nonzero(a != SENTINEL) > 0
        # steady state: a further call leaves the device's free memory unchanged
        torch.cuda.synchronize()
        out = to_dev(np.full((2 * g.S, g.N), SENTINEL, np.float32))
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert _call(g, "hrrt_forward_proj3d_noalloc", img, views[0], out) == _lib.HRRT_OK
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free1, "the no-alloc call changed the device's free memory"
        g.close()


def test_noalloc_refuses_geometry_beyond_scratch(torch_dev):
    """Span 3 at 288 bins: a pair-column of 2 x 288 x 6552 floats exceeds the scratch."""
    from hrrt_reconstruction_amd import _lib
    from hrrt_reconstruction_amd.osem import Osem3D
    torch = torch_dev
    g = Osem3D(def_elements=288, def_angles=288, x_pixels=288, span=3, maxdel=67)
    assert 2 * g.N * g.spp > 2 * 256 * 6656
    img = g.image_zeros()
    prj = g.estimate_zeros()
    assert _call(g, "hrrt_forward_proj3d_noalloc", img, 0, prj) == _lib.HRRT_ERR_ARG
    assert b"exceeds" in _lib.lib().hrrt_last_error()
    g.close()
