"""Staging from flat files in host memory (hrrt_stage_sinograms_host / _w012_host): the
segment-by-segment copies with overlapped staging give the device-memory staging's outputs
bit for bit, which are themselves bit-exact against the oracle (test_gpu_parity.py,
test_gpu_weighting.py).

Cases: OP staging (prompts, delayed, norm, scatter, attenuation) from pageable numpy arrays
and from pinned tensors, with 2x2 rebinning of the norm (gap mask per segment), absent files,
repeated calls on one context (the device copies of the files are reused), and the ANW
staging (-W 2) of short prompts and of a float sinogram.
"""
import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import from_dev, to_dev

pytestmark = pytest.mark.gpu

f32 = np.float32


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


def _files(g, seed):
    p = g.params
    shape = (g.S, p.def_angles, p.def_elements)
    rng = np.random.default_rng(seed)
    pr = rng.integers(0, 12, shape).astype(np.int16)
    d = (rng.random(shape) * 3).astype(f32)
    nrm = (rng.random(shape) + 0.5).astype(f32)
    nrm[rng.random(shape) < 0.01] = 0.0
    for v in (3, 3 + p.def_angles // 2):
        for x in (10, p.def_elements - 10):
            nrm[:, v, x] = 0.0
    sc = (rng.random(shape) * 2).astype(f32)
    at = (rng.random(shape) * 3).astype(f32)
    at[rng.random(shape) < 0.01] = 0.0
    return [pr, d, nrm, sc, at]


def _same(a, b, what):
    for x, y, name in zip(a, b, ("first", "second", "third")):
        assert np.array_equal(from_dev(x), from_dev(y)), f"{what}: {name} output differs"


@pytest.mark.parametrize("geom", ["tiny", "tinyrb", "c1"])
def test_stage_host_matches_device(torch_dev, geom):
    from hrrt_reconstruction_amd.osem import Osem3D
    torch = torch_dev
    print(f"stage host geom={geom}", flush=True)
    g = Osem3D(**GEOMS[geom])
    files = _files(g, 91)
    ref = g.stage(*[to_dev(f) for f in files])
    mask = g.norm_zero_mask().copy()
    got = g.stage_host(*[f.ravel() for f in files])
    torch.cuda.synchronize()
    _same(ref, got, "pageable")
    assert np.array_equal(g.norm_zero_mask(), mask)
    pinned = [torch.from_numpy(f.ravel()).pin_memory() for f in files]
    got = g.stage_host(*pinned)
    torch.cuda.synchronize()
    _same(ref, got, "pinned")
    # absent delayed, scatter and attenuation files, on the same context (reused device copies)
    ref2 = g.stage(to_dev(files[0]), None, to_dev(files[2]), None, None)
    got2 = g.stage_host(files[0].ravel(), None, files[2].ravel(), None, None)
    torch.cuda.synchronize()
    _same(ref2, got2, "absent files")


@pytest.mark.parametrize("geom", ["tiny", "c1"])
def test_stage_w012_host_matches_device(torch_dev, geom):
    from hrrt_reconstruction_amd.osem import Osem3D
    torch = torch_dev
    print(f"stage w012 host geom={geom}", flush=True)
    g = Osem3D(**GEOMS[geom])
    pr, d, nrm, sc, at = _files(g, 92)
    ref = g.stage_w012(2, to_dev(pr), to_dev(d), to_dev(nrm), to_dev(sc), to_dev(at))
    got = g.stage_w012_host(2, pr.ravel(), d.ravel(), nrm.ravel(), sc.ravel(), at.ravel())
    torch.cuda.synchronize()
    _same(ref, got, "ANW short prompts")
    fl = (pr.astype(f32) * 0.75).ravel()
    ref = g.stage_w012(2, to_dev(fl), None, to_dev(nrm), None, to_dev(at), precorrected=True)
    got = g.stage_w012_host(2, fl, None, nrm.ravel(), None, at.ravel(), precorrected=True)
    torch.cuda.synchronize()
    _same(ref, got, "ANW precorrected float")


def test_stage_host_rejects_device_tensors(torch_dev):
    from hrrt_reconstruction_amd.osem import Osem3D
    g = Osem3D(**GEOMS["tiny"])
    files = _files(g, 93)
    with pytest.raises(ValueError):
        g.stage_host(to_dev(files[0].ravel()), None, None, None, None)
