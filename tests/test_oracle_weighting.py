"""The oracle's UW / AW / ANW weighting (-W 0, 1, 2), pinned on the CPU.

- update_estimate_w012 (hrrt_osem3d_span3.cpp:2422-2472): hand-written answers for
  every branch (estimate above, at and below epsilon3, zero data).
- prepare_norm's sensitivity weights per weighting (S3:914-1003, norm_UW :756,
  norm_AW :774, norm_ANW :796) and prepare_osem3dw0/1/2 (S3:1190-1639): each output
  element is recomputed in numpy float32 -- IEEE single operations, one at a time, in
  the reference's order -- from the flat files, with the norm's end planes killed
  (npskip) as the reference reads it, the numerator clamped at zero
  (setnegative2zero, simd_operation.cpp:78), and the reference's own quirks
  written out where they apply (-W 0 with scatter subtracts the norm).
- Short prompts under 2x2 rebinning: views summed as int16 (wrapping), bins in
  float (sino_rebin_float, file_io_processor.cpp:77-104).
- The subset driver with the w012 ratio equals the composition of the pinned
  pieces: forward projection, the ratio, back projection, rotation, update.
"""
import numpy as np
import pytest

from conftest import GEOMS
import bruteforce as bf

f32 = np.float32


@pytest.fixture(scope="module")
def tiny(oracle_lib):
    return oracle_lib.Oracle(**GEOMS["tiny"]), bf.Geometry(**GEOMS["tiny"])


@pytest.fixture(scope="module")
def tinyrb(oracle_lib):
    return oracle_lib.Oracle(**GEOMS["tinyrb"]), bf.Geometry(**GEOMS["tinyrb"])


def file_shape(o):
    return o.S, o.params.def_angles, o.params.def_elements


def killed_norm(o, G, nrm):
    """The norm as read with norm_flag = 1: the first and last npskip planes of every segment zero."""
    n = nrm.copy()
    for th in range(o.th):
        r = G.seg_rows(th)
        n[r[:o.params.npskip]] = 0
        n[r[-o.params.npskip:]] = 0
    return n


def clamp0(t):
    """_mm_max_ps(t, 0): t where t > 0, +0 elsewhere (NaN and -0 included)."""
    return np.where(t > 0, t, f32(0)).astype(f32)


def study_files(o, seed):
    S, A, E = file_shape(o)
    rng = np.random.default_rng(seed)
    pr = rng.integers(0, 12, (S, A, E)).astype(np.int16)
    d = (rng.random((S, A, E)) * 3).astype(f32)
    nrm = (rng.random((S, A, E)) + 0.5).astype(f32)
    nrm[:, 2, 3] = 0.0
    nrm[5, 4, 4] = -1.5
    sc = (rng.random((S, A, E)) * 2).astype(f32)
    at = (rng.random((S, A, E)) * 3).astype(f32)      # some below 1
    at[:, 6, 6] = 0.0
    return pr, d, nrm, sc, at


# ---------------- update_estimate_w012 ----------------

def test_ratio_w012_known_answers(tiny):
    o, _ = tiny
    n2 = (2 * o.S, o.N)
    est = np.full(n2, f32(2.0))
    data = np.full(n2, f32(3.0))
    est[0, :4] = [1e-10, 2e-10, 0.0, -5.0]        # eps3 = 1e-10: at eps3 -> 0, above -> ratio
    data[1, :2] = [0.0, 7.0]
    est[1, :2] = [4.0, 0.0]
    out = o.ratio_w012(est, data)
    assert out[0, 0] == 0.0 and out[0, 1] == f32(3.0) / f32(2e-10)
    assert out[0, 2] == 0.0 and out[0, 3] == 0.0
    assert out[1, 0] == 0.0 and out[1, 1] == 0.0
    assert out[2, 5] == f32(1.5)
    rng = np.random.default_rng(1)
    est = rng.uniform(-1, 4, n2).astype(f32)
    data = rng.uniform(0, 9, n2).astype(f32)
    exp = np.where(est > f32(1e-10), data / np.where(est == 0, f32(1), est), f32(0))
    np.testing.assert_array_equal(o.ratio_w012(est, data), exp.astype(f32))


# ---------------- prepare_norm per weighting ----------------

@pytest.mark.parametrize("weighting", [0, 1, 2])
def test_prepare_norm_weights(tiny, weighting):
    o, G = tiny
    _, _, nrm, _, at = study_files(o, 11)
    out = o.prepare_norm_w(weighting, nrm, at)
    n = killed_norm(o, G, nrm)
    if weighting == 0:       # norm_UW: 1 where N > 0, N elsewhere
        exp = np.where(n > 0, f32(1), n)
    elif weighting == 1:     # norm_AW: 1/max(A, 1) where N > 0, 0 elsewhere
        exp = np.where(n > 0, f32(1) / np.maximum(at, f32(1)), f32(0))
    else:                    # norm_ANW
        den = n * np.maximum(at, f32(1))
        exp = np.where(n > 0, np.where(den > f32(1e-2), f32(1) / np.where(den == 0, f32(1), den), f32(0)), n)
    np.testing.assert_array_equal(out, exp.astype(f32))
    assert o.norm_zero_mask[2, 3] == 1        # one dead column of four keeps the pair alive
    if weighting == 0:
        assert out[5, 4, 4] == f32(-1.5)      # a negative norm is passed through by norm_UW


def test_prepare_norm_weights_without_files(tiny):
    o, G = tiny
    assert np.all(o.prepare_norm_w(0, None, None) == 1.0)
    _, _, _, _, at = study_files(o, 12)
    np.testing.assert_array_equal(o.prepare_norm_w(1, None, at), (f32(1) / np.maximum(at, f32(1))).astype(f32))


# ---------------- prepare_osem3dw0/1/2 ----------------

@pytest.mark.parametrize("scatter", [False, True])
@pytest.mark.parametrize("weighting", [0, 1, 2])
def test_prepare_w012_known_answers(tiny, weighting, scatter):
    o, G = tiny
    pr, d, nrm, sc, at = study_files(o, 13)
    sc_f = sc if scatter else None
    out = o.prepare_w012(weighting, pr, d, nrm, sc_f, at)
    n = killed_norm(o, G, nrm)
    t = pr.astype(f32) - d                          # readsinogram: prompt - delayed
    if weighting == 0:
        t = t * n
        if scatter:
            t = t - n                               # the reference subtracts the norm buffer
        t = t * at
    elif weighting == 1:
        if scatter:
            t = np.where(n > 0, t - sc / np.where(n > 0, n, f32(1)), f32(0)).astype(f32)
        t = t * n
    else:
        if scatter:
            t = np.where(n > 0, t - sc / np.where(n > 0, n, f32(1)), f32(0)).astype(f32)
    np.testing.assert_array_equal(out, clamp0(t))
    assert np.all(out >= 0)


def test_prepare_w012_trues_floats_and_absent_files(tiny):
    o, G = tiny
    pr, d, nrm, sc, at = study_files(o, 14)
    # trues (-t): no delayed subtraction; ANW without scatter uses the data alone
    np.testing.assert_array_equal(o.prepare_w012(2, pr), clamp0(pr.astype(f32)))
    # float sinogram (-Q): the same arithmetic on float data
    fl = (pr.astype(f32) * f32(0.75)).astype(f32)
    np.testing.assert_array_equal(o.prepare_w012(2, fl, d), clamp0(fl - d))
    # AW with scatter and no norm: t - s
    np.testing.assert_array_equal(o.prepare_w012(1, pr, d, None, sc), clamp0(pr.astype(f32) - d - sc))
    # ANW with scatter and no norm: N = 1, t - s/1
    np.testing.assert_array_equal(o.prepare_w012(2, pr, d, None, sc), clamp0(pr.astype(f32) - d - sc))
    # UW with scatter but no norm reads a stale buffer in the reference: refused
    with pytest.raises(ValueError):
        o.prepare_w012(0, pr, d, None, sc)


@pytest.mark.parametrize("weighting", [0, 1, 2])
def test_prepare_w012_precorrected(tiny, weighting):
    """-P: a fully precorrected float sinogram; AW divides by A (where A > 0), ANW by N then A."""
    o, G = tiny
    pr, d, nrm, sc, at = study_files(o, 15)
    rng = np.random.default_rng(16)
    t = rng.uniform(-1, 5, pr.shape).astype(f32)
    out = o.prepare_w012(weighting, t, d, nrm, sc, at, precorrected=True)
    n = killed_norm(o, G, nrm)
    e = t.copy()
    if weighting == 2:
        e = np.where(n > 0, e / np.where(n > 0, n, f32(1)), f32(0)).astype(f32)
    if weighting >= 1:
        e = np.where(at > 0, e / np.where(at > 0, at, f32(1)), e).astype(f32)
    np.testing.assert_array_equal(out, clamp0(e))


def test_prepare_w012_rebinned_short_data(tinyrb):
    """2x2 rebinning of short prompts into float: the two views are summed as int16 (wrapping
    past 32767), then the two bins are added in float starting from 0."""
    o, G = tinyrb
    S, A, E = file_shape(o)
    rng = np.random.default_rng(17)
    pr = rng.integers(0, 20000, (S, A, E)).astype(np.int16)
    out = o.prepare_w012(2, pr)
    b = pr.reshape(S, A // 2, 2, E // 2, 2).astype(np.int32)
    views = (b[:, :, 0] + b[:, :, 1]).astype(np.int16)          # [S][A/2][E/2][2], wrapped
    exp = (f32(0) + views[..., 0].astype(f32)) + views[..., 1].astype(f32)
    np.testing.assert_array_equal(out, clamp0(exp.astype(f32)))
    assert np.any(views < 0), "the sample must exercise the int16 wrap"


# ---------------- the subset driver ----------------

def test_osem_subset_w012_is_the_composition(tiny):
    """og_osem_subset_w012 = for each pair: forward, w012 ratio, (incremental or direct) back
    projection into the rotated frame; then rotation of every pair and the image update."""
    o, G = tiny
    rng = np.random.default_rng(18)
    data = rng.uniform(0, 4, (o.S, o.views, o.N)).astype(f32)
    nf = rng.uniform(0.05, 2.0, (o.N, o.N, o.Z4)).astype(f32)
    img0 = o.init_image() * rng.uniform(0.5, 2.0, (o.N, o.N, o.Z4)).astype(f32)
    s = 5
    for T in (0, 1):
        img = img0.copy()
        o.osem_subset_w012(img, o.image_zeros(), nf, data, s, nthreads=T, workers=2)
        corr = o.image_zeros()
        for i in range(o.sviews // 2):
            view = int(o.vieworder[s][i])
            est = o.forward(img0, view)
            ratio = o.ratio_w012(est, o.gather_pair(data, view))
            buf = o.back_incremental_rotframe(ratio, view) if T == 1 else o.back_direct_rotframe(ratio, view)
            o.rotate_add(buf, view, out=corr)          # rotate_image_buf: buffers added in pair order
        exp = img0.copy()
        o.update_image(exp, corr, nf)
        np.testing.assert_array_equal(img, exp)
        assert not np.array_equal(img, img0)
