"""PSF resolution model (-B) on the GPU, bit for bit against the oracle.

The kernels are built on the host with the reference's double-precision formula and
compared tap by tap; convolve3d, the subset update with -B and the sensitivity images
with -B issue the reference's float operations in its order, so they are compared with
np.array_equal.  C1 has x/y pixels twice the z pixel, which exercises the reference's use
of the y kernel in the z pass.
"""
import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import from_dev, random_image, rel_l2, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


def _pair(geom, oracle_lib, psf=(0.0, 0.0, 0.0)):
    from hrrt_reconstruction_amd.osem import Osem3D
    o = oracle_lib.Oracle(**GEOMS[geom])
    g = Osem3D(**GEOMS[geom])
    o.set_psf(*psf)
    g.set_psf(*psf)
    return o, g


@pytest.mark.parametrize("geom,psf", [("tiny", (0.0, 0.0, 0.0)), ("c1", (0.0, 0.0, 0.0)), ("c2", (0.0, 0.0, 0.0)),
                                      ("tiny", (3.0, 6.0, 0.1)), ("tiny", (2.5, 0.0, 0.0))])
def test_psf_kernels_equal(torch_dev, oracle_lib, geom, psf):
    o, g = _pair(geom, oracle_lib, psf)
    for axis in range(3):
        assert np.array_equal(g.psf_kernel(axis), o.psf_kernel(axis)), axis


@pytest.mark.parametrize("geom", ["tiny", "c1", "c2"])
def test_convolve3d_bitexact(torch_dev, oracle_lib, geom):
    print(f"psf convolve3d geom={geom}", flush=True)
    o, g = _pair(geom, oracle_lib)
    img = random_image(o, seed=21)
    ref = o.convolve3d(img)
    out = from_dev(g.convolve3d(to_dev(img)))
    assert np.array_equal(out, ref), f"max |diff| {np.abs(out - ref).max()} rel {rel_l2(out, ref)}"
    # in place, as the reference blurs the correction and the sensitivity
    d = to_dev(img)
    g.convolve3d(d, out=d)
    assert np.array_equal(from_dev(d)[img[:, :, 10] > 0], ref[img[:, :, 10] > 0])


@pytest.mark.parametrize("geom,subsets,nthreads", [("tiny", (0, 8, 1), 0), ("tiny", (3, 11), 1), ("c1", (0, 8), 1)])
def test_osem_subsets_psf_bitexact(torch_dev, oracle_lib, geom, subsets, nthreads):
    print(f"psf osem geom={geom} subsets={subsets} nthreads={nthreads}", flush=True)
    o, g = _pair(geom, oracle_lib)
    g.set_backproject_schedule(nthreads)
    rng = np.random.default_rng(61)
    shape = (o.S, o.views, o.N)
    prompt = rng.integers(0, 4, shape).astype(np.int16)
    add = rng.uniform(0.0, 0.2, shape).astype(np.float32)
    nf_all = rng.uniform(0.05, 2.0, (o.subsets, o.N, o.N, o.Z4)).astype(np.float32)
    img = o.init_image()
    d_img, d_nf = to_dev(img), to_dev(nf_all)
    d_pr, d_add = to_dev(g.sino_to_paircol(prompt)), to_dev(g.sino_to_paircol(add))
    corr = o.image_zeros()
    for s in subsets:
        o.osem_subset_psf(img, corr, np.ascontiguousarray(nf_all[s]), add, prompt, s, nthreads=nthreads, workers=8)
        g.osem_subset(d_img, d_nf, d_pr, d_add, s)
        out = from_dev(d_img)
        assert np.array_equal(out, img), f"subset {s}: max |diff| {np.abs(out - img).max()} rel {rel_l2(out, img)}"


@pytest.mark.parametrize("geom,nthreads", [("tiny", 0), ("tiny", 2)])
def test_calculate_normfac_psf_bitexact(torch_dev, oracle_lib, geom, nthreads):
    print(f"psf normfac geom={geom} nthreads={nthreads}", flush=True)
    o, g = _pair(geom, oracle_lib)
    g.set_backproject_schedule(nthreads)
    rng = np.random.default_rng(62)
    inv = rng.uniform(0.5, 1.5, (o.S, o.views, o.N)).astype(np.float32)
    ref = o.calculate_normfac_psf(inv.copy(), nthreads=nthreads)
    nf = from_dev(g.calculate_normfac(to_dev(g.sino_to_paircol(inv))))
    assert np.array_equal(nf, ref), f"max |diff| {np.abs(nf - ref).max()}"


def test_psf_off_restores_plain_path_and_frames_refuse(torch_dev, oracle_lib):
    from hrrt_reconstruction_amd import _lib
    o, g = _pair("tiny", oracle_lib)
    d = to_dev(np.stack([o.init_image()]))
    nf = to_dev(np.ones((o.subsets, o.N, o.N, o.Z4), np.float32))
    pr = torch_dev.zeros((1, o.views // 2, 2, o.N, g.spp), dtype=torch_dev.int16, device="cuda")
    ad = torch_dev.zeros((1, o.views // 2, 2, o.N, g.spp), dtype=torch_dev.float32, device="cuda")
    with pytest.raises(_lib.HrrtError):
        g.osem_subset_frames(d, nf, pr, ad, 0)
    g.set_psf(enable=False)
    g.osem_subset_frames(d, nf, pr, ad, 0)
    with pytest.raises(_lib.HrrtError):
        g.convolve3d(d[0])
