"""UW / AW / ANW weighting (-W 0, 1, 2) on the GPU, bit-exact against the oracle.

- Staging (prepare_osem3dw0/1/2 + prepare_norm's weights, hrrt_osem3d_span3.cpp:
  827-1003, 1190-1639): the corrected data, the sensitivity weights and the norm
  zero-column mask, with and without scatter, for short prompts, trues, float and
  fully precorrected sinograms, and under 2x2 rebinning.
- update_estimate_w012 (S3:2436-2447) on one view pair.
- Whole subset updates with the w012 ratio under the reference's thread schedules,
  from staged files through the sensitivity images (the ANW default, -W 2, and the
  others), at the tiny geometries and at span 9, 128x128.
"""
import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import from_dev, rel_l2, to_dev

pytestmark = pytest.mark.gpu

f32 = np.float32


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


def _pair(geom, oracle_lib, **kw):
    from hrrt_reconstruction_amd.osem import Osem3D
    p = dict(GEOMS[geom], **kw)
    return oracle_lib.Oracle(**p), Osem3D(**p)


def _files(o, seed, zero_cols=True):
    p = o.params
    shape = (o.S, p.def_angles, p.def_elements)
    rng = np.random.default_rng(seed)
    pr = rng.integers(0, 12, shape).astype(np.int16)
    d = (rng.random(shape) * 3).astype(f32)
    nrm = (rng.random(shape) + 0.5).astype(f32)
    nrm[rng.random(shape) < 0.01] = 0.0
    if zero_cols:
        for v in (3, 3 + p.def_angles // 2):
            for x in (10, p.def_elements - 10):
                nrm[:, v, x] = 0.0
    sc = (rng.random(shape) * 2).astype(f32)
    at = (rng.random(shape) * 3).astype(f32)
    at[rng.random(shape) < 0.01] = 0.0
    return pr, d, nrm, sc, at


def _check_stage(o, g, out, ref_data, ref_inv):
    data, inv = out
    assert np.array_equal(from_dev(data), g.sino_to_paircol(ref_data)), "corrected data differ"
    assert np.array_equal(from_dev(inv), g.sino_to_paircol(ref_inv)), "sensitivity weights differ"
    assert np.array_equal(g.norm_zero_mask()[:, 1:], o.norm_zero_mask[: o.views // 2, 1:])


@pytest.mark.parametrize("geom", ["tiny", "tinyrb"])
@pytest.mark.parametrize("weighting", [0, 1, 2])
@pytest.mark.parametrize("scatter", [False, True])
def test_stage_w012_bitexact(torch_dev, oracle_lib, geom, weighting, scatter):
    print(f"stage w012 geom={geom} weighting={weighting} scatter={scatter}", flush=True)
    o, g = _pair(geom, oracle_lib)
    pr, d, nrm, sc, at = _files(o, 81, zero_cols=(geom == "tiny"))
    sc_f = sc if scatter else None
    ref_data = o.prepare_w012(weighting, pr, d, nrm, sc_f, at)
    ref_inv = o.prepare_norm_w(weighting, nrm, at)
    out = g.stage_w012(weighting, to_dev(pr), to_dev(d), to_dev(nrm), None if sc_f is None else to_dev(sc), to_dev(at))
    _check_stage(o, g, out, ref_data, ref_inv)
    if geom == "tiny":
        assert (g.norm_zero_mask() == 0).sum() >= 1


@pytest.mark.parametrize("weighting", [0, 1, 2])
def test_stage_w012_other_inputs_bitexact(torch_dev, oracle_lib, weighting):
    """Trues without delayed (-t), a float sinogram (-Q), a fully precorrected sinogram (-P),
    and absent norm / attenuation files."""
    o, g = _pair("tinyrb", oracle_lib)
    pr, d, nrm, sc, at = _files(o, 82, zero_cols=False)
    rng = np.random.default_rng(83)
    fl = rng.uniform(-1.0, 6.0, pr.shape).astype(f32)
    print(f"stage w012 inputs weighting={weighting}: trues", flush=True)
    _check_stage(o, g, g.stage_w012(weighting, to_dev(pr), None, to_dev(nrm), None, to_dev(at)),
                 o.prepare_w012(weighting, pr, None, nrm, None, at), o.prepare_norm_w(weighting, nrm, at))
    print(f"stage w012 inputs weighting={weighting}: float data", flush=True)
    _check_stage(o, g, g.stage_w012(weighting, to_dev(fl), to_dev(d), to_dev(nrm), to_dev(sc), to_dev(at)),
                 o.prepare_w012(weighting, fl, d, nrm, sc, at), o.prepare_norm_w(weighting, nrm, at))
    print(f"stage w012 inputs weighting={weighting}: precorrected", flush=True)
    _check_stage(o, g, g.stage_w012(weighting, to_dev(fl), None, to_dev(nrm), None, to_dev(at), precorrected=True),
                 o.prepare_w012(weighting, fl, None, nrm, None, at, precorrected=True),
                 o.prepare_norm_w(weighting, nrm, at))
    if weighting != 0:
        print(f"stage w012 inputs weighting={weighting}: no norm, no attenuation", flush=True)
        _check_stage(o, g, g.stage_w012(weighting, to_dev(pr), to_dev(d), None, to_dev(sc), None),
                     o.prepare_w012(weighting, pr, d, None, sc, None), o.prepare_norm_w(weighting, None, None))


def test_stage_w012_refuses_unsupported(torch_dev, oracle_lib):
    from hrrt_reconstruction_amd import _lib
    o, g = _pair("tiny", oracle_lib)
    pr, d, nrm, sc, at = _files(o, 84)
    with pytest.raises(_lib.HrrtError) as e:
        g.stage_w012(0, to_dev(pr), to_dev(d), None, to_dev(sc), to_dev(at))
    assert e.value.code == _lib.HRRT_ERR_UNSUPPORTED
    with pytest.raises(_lib.HrrtError) as e:
        g.stage_w012(4, to_dev(pr), to_dev(d), to_dev(nrm), None, to_dev(at))
    assert e.value.code == _lib.HRRT_ERR_ARG


def test_update_estimate_w012_bitexact(torch_dev, oracle_lib):
    o, g = _pair("tiny", oracle_lib)
    rng = np.random.default_rng(85)
    n2 = (2 * o.S, o.N)
    est = rng.uniform(-0.5, 3, n2).astype(f32)
    est[rng.random(n2) < 0.05] = 0.0
    est[0, :3] = [1e-10, 2e-10, -0.0]
    data = rng.uniform(0, 5, n2).astype(f32)
    data[rng.random(n2) < 0.05] = 0.0
    d_est = to_dev(est)
    g.update_estimate_w012(d_est, to_dev(data))
    assert np.array_equal(from_dev(d_est), o.ratio_w012(est, data))


@pytest.mark.parametrize("geom,weighting,nthreads,subsets", [
    ("tiny", 2, 1, (0, 8, 1)), ("tiny", 2, 0, (0, 8)), ("tiny", 0, 2, (0, 8)), ("tiny", 1, 1, (0, 8)),
    ("tinyrb", 2, 1, (0, 8)), ("c1", 2, 1, (0, 8)), ("c1", 1, 0, (0, 8))])
def test_osem_subsets_w012_bitexact(torch_dev, oracle_lib, geom, weighting, nthreads, subsets):
    """From files to image: staging for the weighting, sensitivity images of its weights,
    then subset updates with the w012 ratio, each bit-exact against the oracle."""
    print(f"osem w012 geom={geom} weighting={weighting} nthreads={nthreads} subsets={subsets}", flush=True)
    o, g = _pair(geom, oracle_lib)
    g.set_backproject_schedule(nthreads)
    pr, d, nrm, sc, at = _files(o, 86, zero_cols=(geom == "tiny"))
    sc_f = sc if weighting != 0 else None
    ref_data = o.prepare_w012(weighting, pr, d, nrm, sc_f, at)
    ref_inv = o.prepare_norm_w(weighting, nrm, at)
    d_data, d_inv = g.stage_w012(weighting, to_dev(pr), to_dev(d), to_dev(nrm),
                                 None if sc_f is None else to_dev(sc), to_dev(at))
    ref_nf = o.calculate_normfac(ref_inv.copy(), nthreads=nthreads)
    d_nf = g.calculate_normfac(d_inv)
    assert np.array_equal(from_dev(d_nf), ref_nf), "sensitivity images differ"
    img = o.init_image()
    d_img = to_dev(img)
    corr = o.image_zeros()
    for s in subsets:
        o.osem_subset_w012(img, corr, np.ascontiguousarray(ref_nf[s]), ref_data, s, nthreads=nthreads, workers=16)
        g.osem_subset_w012(d_img, d_nf, d_data, s)
        out = from_dev(d_img)
        assert np.array_equal(out, img), f"subset {s}: max |diff| {np.abs(out - img).max()} rel {rel_l2(out, img)}"
    assert not np.array_equal(img, o.init_image())
