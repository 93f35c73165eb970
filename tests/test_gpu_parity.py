"""Parity of the HIP path (through the C ABI) with the oracle, on seeded inputs.

Every kernel issues the reference's float operations in the reference's order,
so the forward projector, the direct back projector, the rotation, the ratio,
the image update, the sensitivity inversion and the staging are compared
bit for bit (np.array_equal) with the oracle's restatement of the reference.
"""
import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import from_dev, random_image, rel_l2, to_dev

pytestmark = pytest.mark.gpu

SENTINEL = np.float32(-7.25)


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


def _pair(geom, oracle_lib):
    from hrrt_reconstruction_amd.osem import Osem3D
    o = oracle_lib.Oracle(**GEOMS[geom])
    g = Osem3D(**GEOMS[geom])
    assert (g.N, g.S, g.views) == (o.N, o.S, o.views)
    return o, g


FP_CASES = [("tiny", 0), ("tiny", 7), ("tiny", 47), ("tiny9", 5), ("c1", 0), ("c1", 37), ("c1", 71),
            ("c2", 0), ("c2", 100)]


@pytest.mark.parametrize("geom,view", FP_CASES)
def test_forward_proj3d_bitexact(torch_dev, oracle_lib, geom, view):
    print(f"forward geom={geom} view={view}", flush=True)
    o, g = _pair(geom, oracle_lib)
    img = random_image(o, seed=11 + view)
    ref = np.full((2 * o.S, o.N), SENTINEL, np.float32)
    o.forward(img, view, prj=ref)
    d_prj = to_dev(np.full((2 * o.S, o.N), SENTINEL, np.float32))
    g.forward_proj3d(to_dev(img), view, prj=d_prj)
    out = from_dev(d_prj)
    assert np.array_equal(out, ref), f"max |diff| {np.abs(out - ref).max()} rel {rel_l2(out, ref)}"


BP_CASES = [("tiny", 3), ("tiny", 40), ("tiny9", 11), ("c1", 9), ("c1", 64), ("c2", 77)]


@pytest.mark.parametrize("geom,view", BP_CASES)
def test_back_proj3d_direct_bitexact(torch_dev, oracle_lib, geom, view):
    print(f"back geom={geom} view={view}", flush=True)
    o, g = _pair(geom, oracle_lib)
    rng = np.random.default_rng(5 + view)
    prj = rng.uniform(0.0, 2.0, size=(2 * o.S, o.N)).astype(np.float32)
    ima0 = random_image(o, seed=view)
    ref_prj = prj.copy()
    ref_ima = ima0.copy()
    ib = o.image_zeros()
    o.L.og_back_proj_direct(o.gp, ref_prj.ctypes.data_as(oracle_lib.C.c_void_p),
                            ref_ima.ctypes.data_as(oracle_lib.C.c_void_p), view,
                            ib.ctypes.data_as(oracle_lib.C.c_void_p),
                            o.prjbuf().ctypes.data_as(oracle_lib.C.c_void_p))
    d_prj, d_ima = to_dev(prj), to_dev(ima0)
    g.back_proj3d(d_prj, d_ima, view)
    out_prj, out_ima = from_dev(d_prj), from_dev(d_ima)
    assert np.array_equal(out_prj, ref_prj), "in-place rescale of the estimate differs"
    assert np.array_equal(out_ima, ref_ima), f"max |diff| {np.abs(out_ima - ref_ima).max()} rel {rel_l2(out_ima, ref_ima)}"


def test_back_proj3d_incremental_within_tolerance(torch_dev, oracle_lib):
    """The reference's default (incremental) back projector differs from the direct one only by
    float rounding; the GPU (direct formulation) must agree with it to 1e-5 relative L2."""
    o, g = _pair("c1", oracle_lib)
    view = 21
    rng = np.random.default_rng(3)
    prj = rng.uniform(0.0, 2.0, size=(2 * o.S, o.N)).astype(np.float32)
    ib = o.back_incremental_rotframe(prj, view)
    ref = o.rotate_add(ib, view)
    d_ima = to_dev(o.image_zeros())
    g.back_proj3d(to_dev(prj), d_ima, view)
    assert rel_l2(from_dev(d_ima), ref) < 1e-5


def test_ratio_update_normfac_bitexact(torch_dev, oracle_lib):
    o, g = _pair("tiny", oracle_lib)
    rng = np.random.default_rng(17)
    n2 = (2 * o.S, o.N)
    est = rng.uniform(0, 3, n2).astype(np.float32)
    est[rng.random(n2) < 0.05] = 0.0
    add = rng.uniform(0, 0.5, n2).astype(np.float32)
    add[rng.random(n2) < 0.05] = 0.0
    prompt = rng.integers(0, 6, n2).astype(np.int16)
    ref = o.ratio_w3(est, add, prompt)
    d_est = to_dev(est)
    g.update_estimate_w3(d_est, to_dev(add), to_dev(prompt))
    assert np.array_equal(from_dev(d_est), ref)

    img = random_image(o, 1)
    corr = rng.uniform(-1, 200, img.shape).astype(np.float32)
    nf = rng.uniform(0, 5, img.shape).astype(np.float32)
    img[0, 0, 0] = 3e4
    ref_img, ref_corr = img.copy(), corr.copy()
    o.update_image(ref_img, ref_corr, nf)
    d_img, d_corr = to_dev(img), to_dev(corr)
    g.cal_updateimage(d_img, d_corr, to_dev(nf))
    assert np.array_equal(from_dev(d_img), ref_img)
    assert np.array_equal(from_dev(d_corr), ref_corr)

    sens = rng.uniform(-0.2, 3, img.shape).astype(np.float32)
    ref_nf = o.normfac_invert(sens)
    assert np.array_equal(from_dev(g.normfac_calculate(to_dev(sens))), ref_nf)


def _study_files(o, rng, scatter=True, zero_cols=False):
    """Random flat files in the reference layout for geometry o (file bins = def_elements x def_angles)."""
    p = o.params
    S = o.S
    shape = (S, p.def_angles, p.def_elements)
    prompt = rng.integers(0, 5, shape).astype(np.int16)
    delayed = rng.uniform(0, 0.3, shape).astype(np.float32)
    norm = np.exp(0.1 * rng.standard_normal(shape)).astype(np.float32)
    norm[rng.random(shape) < 0.01] = 0.0
    if zero_cols:
        norm[:, 3, 10] = 0.0
        norm[:, 3 + p.def_angles // 2, 10] = 0.0
        norm[:, 3, p.def_elements - 10] = 0.0
        norm[:, 3 + p.def_angles // 2, p.def_elements - 10] = 0.0
    atten = np.exp(rng.uniform(0, 1.5, shape)).astype(np.float32)
    atten[rng.random(shape) < 0.01] = 0.5
    sc = rng.uniform(0, 0.2, shape).astype(np.float32) if scatter else None
    return prompt, delayed, norm, sc, atten


@pytest.mark.parametrize("geom", ["tiny", "tinyrb"])
def test_stage_sinograms_bitexact(torch_dev, oracle_lib, geom):
    o, g = _pair(geom, oracle_lib)
    rng = np.random.default_rng(23)
    prompt_f, delayed_f, norm_f, sc_f, atten_f = _study_files(o, rng, zero_cols=(geom == "tiny"))
    ref_prompt, ref_add = o.prepare_w3(prompt_f, delayed_f, norm_f, sc_f, atten_f)
    ref_inv = o.prepare_norm(norm_f, atten_f)
    d = [to_dev(a) for a in (prompt_f, delayed_f, norm_f, sc_f, atten_f)]
    pr, add, inv = g.stage(*d)
    assert np.array_equal(from_dev(pr), g.sino_to_paircol(ref_prompt))
    assert np.array_equal(from_dev(add), g.sino_to_paircol(ref_add))
    assert np.array_equal(from_dev(inv), g.sino_to_paircol(ref_inv))
    assert np.array_equal(g.norm_zero_mask()[:, 1:], o.norm_zero_mask[: o.views // 2, 1:])
    if geom == "tiny":
        assert (g.norm_zero_mask() == 0).sum() >= 1


@pytest.mark.parametrize("geom,nthreads", [("tiny", 0), ("tiny", 2), ("tinyrb", 0), ("tinyrb", 1)])
def test_calculate_normfac_bitexact(torch_dev, oracle_lib, geom, nthreads):
    """Sensitivity images under the reference's thread schedule -T nthreads (0: direct only)."""
    print(f"normfac geom={geom} nthreads={nthreads}", flush=True)
    o, g = _pair(geom, oracle_lib)
    g.set_backproject_schedule(nthreads)
    rng = np.random.default_rng(29)
    _, _, norm_f, _, atten_f = _study_files(o, rng, scatter=False)
    inv = o.prepare_norm(norm_f, atten_f)
    g.set_norm_zero_mask(o.norm_zero_mask[: o.views // 2])
    ref = o.calculate_normfac(inv.copy(), nthreads=nthreads)
    nf = g.calculate_normfac(to_dev(g.sino_to_paircol(inv)))
    assert np.array_equal(from_dev(nf), ref)


@pytest.mark.parametrize("geom,subsets,nthreads", [("tiny", (0, 8, 1), 0), ("tiny", (0, 8), 2), ("tinyrb", (0, 8), 0),
                                                   ("c1", (0, 8), 0), ("c1", (0, 5), 3)])
def test_osem_subsets_bitexact(torch_dev, oracle_lib, geom, subsets, nthreads):
    """Subset updates under the reference's thread schedule -T nthreads: per subset the first
    npairs - npairs % nthreads pairs incremental, the rest direct (0: direct only)."""
    print(f"osem geom={geom} subsets={subsets} nthreads={nthreads}", flush=True)
    o, g = _pair(geom, oracle_lib)
    g.set_backproject_schedule(nthreads)
    rng = np.random.default_rng(31)
    shape = (o.S, o.views, o.N)
    prompt = rng.integers(0, 4, shape).astype(np.int16)
    add = rng.uniform(0.0, 0.2, shape).astype(np.float32)
    nf_all = rng.uniform(0.05, 2.0, (o.subsets, o.N, o.N, o.Z4)).astype(np.float32)
    img = o.init_image()
    d_img = to_dev(img)
    d_nf = to_dev(nf_all)
    d_pr, d_add = to_dev(g.sino_to_paircol(prompt)), to_dev(g.sino_to_paircol(add))
    corr = o.image_zeros()
    for s in subsets:
        o.osem_subset(img, corr, np.ascontiguousarray(nf_all[s]), add, prompt, s, nthreads=nthreads, workers=8)
        g.osem_subset(d_img, d_nf, d_pr, d_add, s)
        out = from_dev(d_img)
        assert np.array_equal(out, img), f"subset {s}: max |diff| {np.abs(out - img).max()} rel {rel_l2(out, img)}"
