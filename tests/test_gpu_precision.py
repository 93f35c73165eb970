"""Precision form on the GPU: FP16 per-bin inputs, Kahan accumulation, FP64 log-likelihood.

Against an FP64 projector (tests/bruteforce.py, line integrals from first principles):
the forward sinogram, the sensitivity image and the image after one OSEM iteration,
each for the plain FP32 path and the precision form, with the tolerances written
below.  The FP64 log-likelihood and count moments are checked against float64 sums
of the same inputs on the host, and batched frames against single frames.
"""
import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import from_dev, rel_l2, to_dev
import bruteforce as bf

pytestmark = pytest.mark.gpu

# FP16 keeps 11 significant bits: one storage rounding is <= 2^-12 relative.
FP16_REL = 2.0 ** -11


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="module")
def tiny(torch_dev, oracle_lib):
    from hrrt_reconstruction_amd.osem import Osem3D
    kw = GEOMS["tiny"]
    return oracle_lib.Oracle(**kw), Osem3D(**kw), bf.Geometry(**kw)


def half_bits(a):
    with np.errstate(over="ignore"):           # values beyond 65504 round to infinity, as on the GPU
        return np.ascontiguousarray(a, np.float32).astype(np.float16).view(np.int16)


def test_to_half_rounds_like_ieee(torch_dev, tiny):
    o, g, G = tiny
    print("precision to_half", flush=True)
    rng = np.random.default_rng(51)
    x = np.concatenate([rng.normal(0, 1e3, 4096), rng.uniform(-1e-6, 1e-6, 4096), [0.0, -0.0, 65504.0, 7e4, -7e4,
                        6.1e-5, 5.9e-8, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11]]).astype(np.float32)
    got = from_dev(g.to_half(to_dev(x)))
    assert np.array_equal(got, half_bits(x))
    # the range report counts what numpy's own conversion sends to +-inf and to subnormal/zero
    got2, (over, under) = g.to_half_range(to_dev(x))
    assert np.array_equal(from_dev(got2), got)
    h = half_bits(x).view(np.float16).astype(np.float32)
    assert over == int(np.sum(np.isinf(h) & np.isfinite(x)))
    assert under == int(np.sum((x != 0) & (np.abs(h) < 2.0 ** -14)))
    assert over == 2 and under >= 4096          # +-7e4, and every |x| <= 1e-6


def test_forward_vs_fp64(torch_dev, tiny):
    o, g, G = tiny
    print("precision forward", flush=True)
    rng = np.random.default_rng(52)
    img = rng.random((o.N, o.N, o.Z4)).astype(np.float32)
    d = to_dev(img)
    for v in (0, 13, 40):
        est = from_dev(g.forward_proj3d(d, v))
        ref = bf.forward(G, img.astype(np.float64), v)
        assert rel_l2(est, ref) < 2e-6, v


def _sens_setup(o, G, seed):
    rng = np.random.default_rng(seed)
    inv = rng.uniform(0.5, 1.5, (o.S, o.views, o.N)).astype(np.float32)
    return inv


def test_sensitivity_vs_fp64(torch_dev, tiny):
    o, g, G = tiny
    print("precision sensitivity", flush=True)
    inv = _sens_setup(o, G, 53)
    nf32 = from_dev(g.calculate_normfac(to_dev(g.sino_to_paircol(inv))))
    inv16 = g.to_half(to_dev(g.sino_to_paircol(inv)))
    nf16 = from_dev(g.calculate_normfac_mixed(inv16))
    for s in (0, 7):
        ref = bf.invert(G, bf.sensitivity(G, inv.astype(np.float64), s, o.subsets), o.params.epsilon1)
        e32, e16 = rel_l2(nf32[s], ref), rel_l2(nf16[s], ref)
        print(f"subset {s}: sensitivity rel-L2 vs FP64: fp32 {e32:.2e}, fp16+kahan {e16:.2e}", flush=True)
        # both carry the reference's float rotation positions (~1e-4 from exact positions);
        # the FP16 inputs add at most one storage rounding per bin
        assert e32 < 5e-4
        assert e16 < 5e-4 + FP16_REL


def test_osem_iteration_vs_fp64(torch_dev, tiny):
    o, g, G = tiny
    print("precision osem iteration", flush=True)
    rng = np.random.default_rng(54)
    inv = _sens_setup(o, G, 55)
    img_true = np.zeros((o.N, o.N, o.Z4), np.float32)
    xs = np.arange(o.N) - o.N // 2
    img_true[(xs[:, None] ** 2 + xs[None, :] ** 2) <= (0.3 * o.N) ** 2, 6:26] = 1.0
    sino = np.zeros((o.S, o.views, o.N), np.float32)
    for v in range(o.views // 2):
        p = o.forward(img_true, v)
        sino[:, v], sino[:, v + o.views // 2] = p[:o.S], p[o.S:]
    add = np.full_like(sino, 0.3)
    prompt = rng.poisson((sino + add) * 2.0).astype(np.int16)
    nf32 = g.calculate_normfac(to_dev(g.sino_to_paircol(inv)))
    inv16 = g.to_half(to_dev(g.sino_to_paircol(inv)))
    nf16 = g.calculate_normfac_mixed(inv16)
    d_pr = to_dev(g.sino_to_paircol(prompt))
    d_add = to_dev(g.sino_to_paircol(add))
    d_add16 = g.to_half(d_add)
    img0 = o.init_image()
    i32, i16 = to_dev(img0), to_dev(img0)
    ref = img0.astype(np.float64)
    nf64 = [bf.invert(G, bf.sensitivity(G, inv.astype(np.float64), s, o.subsets), o.params.epsilon1)
            for s in range(o.subsets)]
    for s in o.order:
        g.osem_subset(i32, nf32, d_pr, d_add, int(s))
        g.osem_subset_mixed(i16, nf16, d_pr, d_add16, int(s))
        ref = bf.osem_subset(G, ref, nf64[s], add, prompt, int(s), o.subsets)
    e32, e16 = rel_l2(from_dev(i32), ref), rel_l2(from_dev(i16), ref)
    print(f"image after 1 iteration rel-L2 vs FP64: fp32 {e32:.2e}, fp16+kahan {e16:.2e}", flush=True)
    assert e32 < 2e-3 and e16 < 2e-3


def test_loglik_and_moments_fp64(torch_dev, tiny):
    o, g, G = tiny
    torch = torch_dev
    print("precision loglik", flush=True)
    rng = np.random.default_rng(56)
    inv = rng.uniform(0.5, 1.5, (o.S, o.views, o.N)).astype(np.float32)
    add = rng.uniform(0.0, 0.5, (o.S, o.views, o.N)).astype(np.float32)
    prompt = rng.poisson(2.0, (o.S, o.views, o.N)).astype(np.int16)
    img = rng.uniform(0.5, 1.5, (o.N, o.N, o.Z4)).astype(np.float32)
    d_img = to_dev(img)
    d_inv16 = g.to_half(to_dev(g.sino_to_paircol(inv)))
    d_add16 = g.to_half(to_dev(g.sino_to_paircol(add)))
    inv_h = inv.astype(np.float16).astype(np.float32)
    add_h = add.astype(np.float16).astype(np.float32)
    nf = torch.ones((o.subsets, o.N, o.N, o.Z4), dtype=torch.float32, device="cuda")
    s = 5
    ests = {v: from_dev(g.forward_proj3d(d_img, v)) for v in o.vieworder[s][: o.sviews // 2]}
    ll = torch.zeros(1, dtype=torch.float64, device="cuda")
    g.osem_subset_mixed(d_img, nf, to_dev(g.sino_to_paircol(prompt)), d_add16, s, invnorm16=d_inv16, loglik=ll)
    ref = 0.0
    for v, est in ests.items():
        den = est + bf.gather_pair(add_h, v)                  # float32, as on the GPU
        w = bf.gather_pair(inv_h, v).astype(np.float64)
        y = bf.gather_pair(prompt, v).astype(np.float64)
        ok = (w > 0) & (den > o.params.epsilon3)
        # bins outside the FOV or at bin 0 are never written by the forward projector
        ok[:, 0] = False
        ok &= np.repeat(np.array([G.cyl[x][0] != G.cyl[x][1] for x in range(o.N)])[None, :], 2 * o.S, 0)
        mu = den.astype(np.float64) * w
        ref += float((np.where(y > 0, y * np.log(np.where(ok, mu, 1.0)), 0.0) - mu)[ok].sum())
    got = float(ll.cpu().numpy()[0])
    print(f"loglik gpu {got:.9e} host {ref:.9e}", flush=True)
    assert got == pytest.approx(ref, rel=1e-9)
    # count moments over the whole sinogram
    d_img2 = to_dev(img)
    m = g.count_moments(d_img2, to_dev(g.sino_to_paircol(prompt)), to_dev(g.sino_to_paircol(add)),
                        to_dev(g.sino_to_paircol(inv))).cpu().numpy()
    sy = float(prompt.astype(np.float64).sum())
    sa = float((inv.astype(np.float64) * add).sum())
    se = 0.0
    for v in range(o.views // 2):
        est = from_dev(g.forward_proj3d(d_img2, v)).astype(np.float64)
        se += float((bf.gather_pair(inv, v).astype(np.float64) * est).sum())
    assert m[0] == pytest.approx(sy, rel=1e-12)
    assert m[1] == pytest.approx(sa, rel=1e-12)
    assert m[2] == pytest.approx(se, rel=1e-10)


def test_frames_mixed_match_single(torch_dev, tiny):
    o, g, G = tiny
    torch = torch_dev
    print("precision frames", flush=True)
    rng = np.random.default_rng(57)
    F = 3
    shape = (o.S, o.views, o.N)
    prompts = [rng.poisson(1.0 + f, shape).astype(np.int16) for f in range(F)]
    adds = [rng.uniform(0, 0.3, shape).astype(np.float32) for f in range(F)]
    inv = rng.uniform(0.5, 1.5, shape).astype(np.float32)
    d_inv16 = g.to_half(to_dev(g.sino_to_paircol(inv)))
    nf = g.calculate_normfac_mixed(d_inv16)
    d_pr = to_dev(np.stack([g.sino_to_paircol(p) for p in prompts]))
    d_add16 = g.to_half(to_dev(np.stack([g.sino_to_paircol(a) for a in adds])))
    img0 = o.init_image()
    d_imgs = to_dev(np.stack([img0] * F))
    singles = [to_dev(img0) for _ in range(F)]
    llf = torch.zeros(F, dtype=torch.float64, device="cuda")
    lls = torch.zeros(F, dtype=torch.float64, device="cuda")
    for s in (0, 8):
        g.osem_subset_frames_mixed(d_imgs, nf, d_pr, d_add16, s, invnorm16=d_inv16, loglik=llf)
        for f in range(F):
            g.osem_subset_mixed(singles[f], nf, d_pr[f], d_add16[f], s, invnorm16=d_inv16, loglik=lls[f:f + 1])
        got = from_dev(d_imgs)
        for f in range(F):
            assert np.array_equal(got[f], from_dev(singles[f])), (s, f)
        assert np.array_equal(llf.cpu().numpy(), lls.cpu().numpy())
    # graph-replayed iterations of the precision form, with the log-likelihood of every subset
    stream = torch.cuda.Stream()
    ll_it = torch.zeros((o.subsets, F), dtype=torch.float64, device="cuda")
    prev = None
    for it in range(2):
        stream.wait_stream(torch.cuda.current_stream())   # inputs written on the current stream
        with torch.cuda.stream(stream):
            g.osem_iteration_frames_mixed(d_imgs, nf, d_pr, d_add16, invnorm16=d_inv16, loglik=ll_it)
        stream.synchronize()
        cur = ll_it.cpu().numpy().copy()
        assert np.all(np.isfinite(cur))
        if prev is not None:
            # EM raises the likelihood of each frame's data summed over subsets
            assert np.all(cur.sum(0) > prev.sum(0))
        prev = cur
