"""The reference projector's own C++ entry points (libhrrt_osem3d_legacy.so, hrrt_osem3d.h:197-211)
called as the reference's driver calls them -- float*** images whose every [x][y] column and
float** estimates whose every row sit behind their own pointer -- against the oracle, bit for
bit: both forward variants, back_proj3d2 (estimate rescaled in place, rotated into ima),
back_proj3d_thread1 (the rotated frame in imagebuf, outside the FOV untouched) and
rotate_image_buf over several buffers."""
import ctypes as C

import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import random_image
from legacy_util import Image3, Image4, Legacy, Rows2, padded_image, padded_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def legacy():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    L = Legacy()
    yield L
    L.free_dependencies()


def _setup(legacy, oracle_lib, geom):
    from hrrt_reconstruction_amd.osem import make_params
    p = make_params(**GEOMS[geom])
    assert legacy.L.hrrt_legacy_set_params(C.byref(p)) == 1
    o = oracle_lib.Oracle(**GEOMS[geom])
    assert legacy.dependencies(o.N, o.views, 0) == 1
    assert legacy.L.hrrt_legacy_context() is not None
    return o


def _vp(a):
    return a.ctypes.data_as(C.c_void_p)


def _fov(o):
    cyl = o.cylwiny
    y = np.arange(o.N)[None, :]
    return (y >= cyl[:, :1]) & (y < cyl[:, 1:])


CASES = [("tiny", 3), ("tiny9", 11), ("c1", 9)]


@pytest.mark.parametrize("geom,view", CASES)
def test_forward_variants(legacy, oracle_lib, geom, view):
    print(f"legacy forward geom={geom} view={view}", flush=True)
    o = _setup(legacy, oracle_lib, geom)
    rng = np.random.default_rng(31 + view)
    img = random_image(o, seed=view + 3)
    prj0 = rng.uniform(0.0, 2.0, size=(2 * o.S, o.N)).astype(np.float32)
    ref = o.forward(img, view, prj=prj0.copy())
    for fn in (legacy.forward_proj3d2, legacy.forward_proj3d_thread1):
        ima, prj = Image3(padded_image(img)), Rows2(padded_rows(prj0))
        assert fn(ima.ptr, prj.ptr, view, 4, None, None) == 1
        assert np.array_equal(prj.a[:, : o.N], ref)
        assert (prj.a[:, o.N:] == -5.0).all() and np.array_equal(ima.a[:, :, : o.Z4], img)


@pytest.mark.parametrize("geom,view", CASES)
def test_back_proj3d2(legacy, oracle_lib, geom, view):
    print(f"legacy back_proj3d2 geom={geom} view={view}", flush=True)
    o = _setup(legacy, oracle_lib, geom)
    rng = np.random.default_rng(41 + view)
    prj0 = rng.uniform(0.0, 2.0, size=(2 * o.S, o.N)).astype(np.float32)
    ima0 = random_image(o, seed=view + 5)
    ref_prj, ref_ima = prj0.copy(), ima0.copy()
    o.L.og_back_proj_direct(o.gp, _vp(ref_prj), _vp(ref_ima), view, _vp(o.image_zeros()), _vp(o.prjbuf()))
    ima, prj = Image3(padded_image(ima0)), Rows2(padded_rows(prj0))
    assert legacy.back_proj3d2(prj.ptr, ima.ptr, view, 4, None, None) == 1
    assert np.array_equal(prj.a[:, : o.N], ref_prj), "in-place rescale of the estimate differs"
    assert np.array_equal(ima.a[:, :, : o.Z4], ref_ima)
    assert (ima.a[:, :, o.Z4:] == -7.0).all()


@pytest.mark.parametrize("geom", ["tiny", "c1"])
def test_back_proj3d_thread1_and_rotate_image_buf(legacy, oracle_lib, geom):
    """The reference driver's sequence (S3:4174-4196): back_proj3d_thread1 into one imagebuf
    per view pair, then rotate_image_buf of all of them into the correction image."""
    print(f"legacy back_proj3d_thread1 + rotate_image_buf geom={geom}", flush=True)
    o = _setup(legacy, oracle_lib, geom)
    rng = np.random.default_rng(53)
    fov = _fov(o)
    views = [2, 17, o.views // 2 - 1]
    bufs, refs = [], []
    for i, view in enumerate(views):
        prj0 = rng.uniform(0.0, 2.0, size=(2 * o.S, o.N)).astype(np.float32)
        ref_prj, ref_ib = prj0.copy(), o.image_zeros()
        o.L.og_back_proj_incremental(o.gp, _vp(ref_prj), _vp(ref_ib), view, _vp(o.prjbuf()))
        # first buffer: stale values outside the FOV stay; the others start zeroed, as the
        # reference allocates them (alloc_imagexyzf_thread S3:2628-2663)
        start = rng.uniform(-1.0, 1.0, ref_ib.shape).astype(np.float32) if i == 0 else o.image_zeros()
        ib, prj = Image3(padded_image(start)), Rows2(padded_rows(prj0))
        assert legacy.back_proj3d_thread1(prj.ptr, None, view, 4, ib.ptr, None) == 1
        assert np.array_equal(prj.a[:, : o.N], ref_prj), "in-place rescale of the estimate differs"
        expect = np.where(fov[:, :, None], ref_ib, start)
        expect[:, :, o.Z:] = start[:, :, o.Z:]         # pad planes: left as they were
        assert np.array_equal(ib.a[:, :, : o.Z4], expect), f"view {view}"
        bufs.append(ib)
        refs.append(expect)
    ima0 = random_image(o, seed=77)
    ref = ima0.copy()
    for ib, view in zip(refs[1:], views[1:]):
        o.rotate_add(np.ascontiguousarray(ib), view, out=ref)
    ima = Image3(padded_image(ima0))
    view_arr = np.array(views[1:], np.int32)
    assert legacy.rotate_image_buf(Image4(bufs[1:]).ptr, ima.ptr, len(view_arr), _vp(view_arr)) == 1
    assert np.array_equal(ima.a[:, :, : o.Z4], ref)


def test_norm_zero_mask(legacy, oracle_lib):
    """Columns whose normalisation is zero in every plane: the forward projector leaves their
    bins stale and the back projector writes zeros (PJ:1811-1813, 4236-4262); the driver's
    short** mask is handed over as the reference's global is set."""
    print("legacy norm_zero_mask geom=tiny", flush=True)
    o = _setup(legacy, oracle_lib, "tiny")
    rng = np.random.default_rng(61)
    mask = (rng.random((o.views // 2 + 1, o.N // 2 + 1)) > 0.3).astype(np.int8)
    o.set_norm_zero_mask(mask)
    rows = mask[: o.views // 2].astype(np.int16) * rng.integers(1, 5, (o.views // 2, o.N // 2 + 1)).astype(np.int16)
    assert legacy.L.hrrt_legacy_set_norm_zero_mask(Rows2(rows).ptr) == 1
    fov = _fov(o)
    cyl = o.cylwiny
    for view in (2, 30):
        img = random_image(o, seed=view)
        prj0 = rng.uniform(0.0, 2.0, size=(2 * o.S, o.N)).astype(np.float32)
        ref = o.forward(img, view, prj=prj0.copy())
        # masked columns inside the FOV: the reference copies its projection buffer's previous
        # contents there, the oracle (a fresh buffer) zeros; the library leaves the caller's bins
        skip = np.zeros(o.N, bool)
        for m in range(1, o.N // 2 + 1):
            if mask[view, m] == 0 and cyl[m, 0] != 0:
                skip[m] = skip[o.N - m] = True
        assert skip.any() and (ref[:, skip] == 0).all()
        prj = Rows2(padded_rows(prj0))
        assert legacy.forward_proj3d_thread1(Image3(padded_image(img)).ptr, prj.ptr, view, 1, None, None) == 1
        assert np.array_equal(prj.a[:, : o.N], np.where(skip[None, :], prj0, ref)), f"forward view {view}"
        ref_ib = o.back_incremental_rotframe(prj0, view)
        ib = Image3(padded_image(o.image_zeros()))
        assert legacy.back_proj3d_thread1(Rows2(padded_rows(prj0)).ptr, None, view, 1, ib.ptr, None) == 1
        assert np.array_equal(ib.a[:, :, : o.Z4], np.where(fov[:, :, None], ref_ib, 0.0)), f"back view {view}"


@pytest.mark.parametrize("geom,view", [("tiny", 5), ("c1", 33)])
def test_rotframe_and_rotate_add_entry_points(legacy, oracle_lib, geom, view):
    """The C ABI forms under the legacy symbols: hrrt_back_proj3d_rotframe (both projectors)
    and hrrt_rotate_add on device buffers, and their _host variants on host buffers."""
    import torch
    from gpu_util import from_dev, to_dev
    from hrrt_reconstruction_amd import _lib
    from hrrt_reconstruction_amd.osem import Osem3D
    print(f"rotframe / rotate_add geom={geom} view={view}", flush=True)
    o = oracle_lib.Oracle(**GEOMS[geom])
    g = Osem3D("cuda:0", **GEOMS[geom])
    fov = _fov(o)
    rng = np.random.default_rng(71 + view)
    prj0 = rng.uniform(0.0, 2.0, size=(2 * o.S, o.N)).astype(np.float32)
    stale = rng.uniform(-1.0, 1.0, (o.N, o.N, o.Z4)).astype(np.float32)
    for incremental, og in ((True, o.L.og_back_proj_incremental), (False, o.L.og_back_proj_direct_rotframe)):
        ref_prj, ref_ib = prj0.copy(), o.image_zeros()
        og(o.gp, _vp(ref_prj), _vp(ref_ib), view, _vp(o.prjbuf()))
        expect = np.where(fov[:, :, None], ref_ib, stale)
        expect[:, :, o.Z:] = stale[:, :, o.Z:]
        d_prj, d_ib = to_dev(prj0), to_dev(stale)
        g.back_proj3d_rotframe(d_prj, d_ib, view, incremental)
        assert np.array_equal(from_dev(d_prj), ref_prj) and np.array_equal(from_dev(d_ib), expect), incremental
        h_prj, h_ib = prj0.copy(), stale.copy()
        _lib.call("hrrt_back_proj3d_rotframe_host", g.ctx, _vp(h_prj), _vp(h_ib), view, int(incremental))
        assert np.array_equal(h_prj, ref_prj) and np.array_equal(h_ib, expect), incremental
    ib = np.where(fov[:, :, None], o.back_incremental_rotframe(prj0, view), 0.0).astype(np.float32)
    ima0 = random_image(o, seed=view)
    ref = o.rotate_add(ib, view, out=ima0.copy())
    d_ima = to_dev(ima0)
    g.rotate_add(to_dev(ib), d_ima, view)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_ima), ref)
    h_ima = ima0.copy()
    _lib.call("hrrt_rotate_add_host", g.ctx, _vp(ib), _vp(h_ima), view)
    assert np.array_equal(h_ima, ref)


def test_argument_checks(legacy, oracle_lib):
    print("legacy argument checks", flush=True)
    o = _setup(legacy, oracle_lib, "tiny")
    assert legacy.dependencies(o.N + 2, o.views, 0) == 0          # does not match the parameters
    assert legacy.L.hrrt_legacy_context() is None
    o = _setup(legacy, oracle_lib, "tiny")
    img = Image3(padded_image(o.image_zeros()))
    prj = Rows2(padded_rows(o.prj_zeros()))
    assert legacy.forward_proj3d2(img.ptr, prj.ptr, o.views // 2, 1, None, None) == 0
    assert legacy.back_proj3d2(prj.ptr, img.ptr, -1, 1, None, None) == 0
    bad = np.array([o.views], np.int32)
    assert legacy.rotate_image_buf(Image4([img]).ptr, img.ptr, 1, _vp(bad)) == 0
