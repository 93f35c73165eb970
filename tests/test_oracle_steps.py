"""The oracle's per-bin / per-voxel steps, sinogram staging and subset drivers (CPU only).

Known answers are written out by hand or computed with numpy float32 (IEEE,
correctly rounded, the same single operations as the reference), and the
drivers are checked against compositions of independently pinned pieces and
against a property of the algorithm (EM raises the Poisson likelihood).
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


# ---------------- ratio (pt_update_est_w3 S3:2493-2505) ----------------

def test_ratio_known_answers(tiny):
    o, _ = tiny
    n = 2 * o.S * o.N
    est = np.zeros(n, f32)
    add = np.zeros(n, f32)
    pr = np.zeros(n, np.int16)
    cases = [  # est, add, prompt, expected
        (2.0, 1.0, 6, 2.0),           # 6 / 3
        (2.0, 1.0, 0, 0.0),           # no prompts -> 0
        (0.0, 0.0, 5, 1.0),           # denominator <= eps3 -> 1
        (1e-11, 0.0, 5, 1.0),
        (-3.0, 1.0, 4, 1.0),          # negative denominator -> 1
        (0.5, 0.0, -2, -4.0),         # int16 prompts are signed
        (3.0, 0.0, 32767, np.float32(32767) / np.float32(3.0)),
        (0.0, 4.0, 1, 0.25),          # additive term alone
    ]
    for i, (e, a, p, _) in enumerate(cases):
        est[i], add[i], pr[i] = e, a, p
    out = o.ratio_w3(est.reshape(2 * o.S, o.N), add.reshape(2 * o.S, o.N), pr.reshape(2 * o.S, o.N)).ravel()
    for i, (_, _, _, exp) in enumerate(cases):
        assert out[i] == f32(exp), (i, out[i], exp)
    rng = np.random.default_rng(0)
    est = rng.random(n).astype(f32) * 5
    add = rng.random(n).astype(f32)
    pr = rng.integers(0, 40, n).astype(np.int16)
    out = o.ratio_w3(est.reshape(2 * o.S, o.N), add.reshape(2 * o.S, o.N), pr.reshape(2 * o.S, o.N)).ravel()
    den = est + add
    exp = np.where(pr == 0, f32(0), pr.astype(f32) / den)
    np.testing.assert_array_equal(out, exp)


# ---------------- update (pt_cal_updateimage S3:2033-2056) ----------------

def test_update_known_answers(tiny):
    o, G = tiny
    shape = (o.N, o.N, o.Z4)
    img = np.ones(shape, f32)
    corr = np.ones(shape, f32)
    nf = np.ones(shape, f32)
    x, y = o.N // 2, o.N // 2
    cases = [  # img, corr, nf, expected
        (2.0, 3.0, 0.5, 3.0),
        (2.0, 0.001, 1.0, 0.02),        # correction clamped below at 0.01
        (2.0, 500.0, 1.0, 200.0),       # ... and above at 100
        (500.0, 90.0, 1.0, 1e4),        # image capped at image_max
        (2.0, 1.0, 0.0, 0.0),           # zero sensitivity inverse
        (2.0, -5.0, 1.0, 0.02),         # negative correction -> lower clamp
    ]
    for z, (i, c, n, _) in enumerate(cases):
        img[x, y, z], corr[x, y, z], nf[x, y, z] = i, c, n
    i0, c0 = img.copy(), corr.copy()
    o.update_image(img, corr, nf)
    for z, (i, c, n, e) in enumerate(cases):
        exp = f32(min(f32(1e4), f32(i) * (f32(min(100.0, max(0.01, c))) * f32(n))))
        assert img[x, y, z] == exp and abs(float(exp) - e) <= 1e-6 * max(1, e), (z, img[x, y, z], e)
    inside = np.zeros(shape, bool)
    for xx in range(o.N):
        lo, hi = G.cyl[xx]
        inside[xx, lo:hi, :] = True
    assert np.all(corr[inside] == 0)                      # correction is consumed
    np.testing.assert_array_equal(img[~inside], i0[~inside])  # outside the FOV untouched
    np.testing.assert_array_equal(corr[~inside], c0[~inside])


# ---------------- sensitivity inverse (pt_normfac_calculate S3:1870-1889) ----------------

def test_normfac_known_answers(tiny):
    o, G = tiny
    shape = (o.N, o.N, o.Z4)
    rng = np.random.default_rng(1)
    s = (rng.random(shape) * 0.2 - 0.1).astype(f32)
    s[o.N // 2, o.N // 2, :4] = [0.05, 0.0500001, -0.06, 0.0]
    out = o.normfac_invert(s)
    with np.errstate(divide="ignore"):
        exp = np.where(np.abs(s) > f32(5e-2), f32(1) / s, f32(0)).astype(f32)
    inside = np.zeros(shape, bool)
    for xx in range(o.N):
        lo, hi = G.cyl[xx]
        inside[xx, lo:hi, :o.Z] = True
    np.testing.assert_array_equal(out[inside], exp[inside])
    assert np.all(out[~inside] == 0)
    assert out[o.N // 2, o.N // 2, 0] == 0 and out[o.N // 2, o.N // 2, 2] == f32(1) / f32(-0.06)


# ---------------- staging (FIO:29-259, S3:796-1056, 1649-1818) ----------------

def file_planes(o):
    return o.S, o.params.def_angles, o.params.def_elements


def test_segment_blocks_follow_file_order(tinyrb):
    """Internal segment s reads file block s (FIO:116-118 swap + seg_offset, PJ:446-457)."""
    o, G = tinyrb
    S, A, E = file_planes(o)
    f = np.broadcast_to(np.arange(S, dtype=f32)[:, None, None], (S, A, E)).copy()
    for th in range(o.th):
        out = o.read_flat_float(f, th)
        planes = out[:, 0, 0] / 4.0                 # 2x2 rebin sums four copies
        np.testing.assert_array_equal(planes, np.arange(G.seg_off[th], G.seg_off[th] + G.nplanes[th]))


def test_rebin_sums_float_and_short(tinyrb):
    o, G = tinyrb
    S, A, E = file_planes(o)
    rng = np.random.default_rng(2)
    f = rng.random((S, A, E)).astype(f32)
    sh = rng.integers(-100, 12000, (S, A, E)).astype(np.int16)
    for th in (0, 1, 4):
        rows = G.seg_rows(th)
        out = o.read_flat_float(f, th)
        blk = f[rows].reshape(len(rows), A // 2, 2, E // 2, 2)
        # view pairs first (in place), then radial pairs, in float32
        exp = (blk[:, :, 0, :, 0] + blk[:, :, 1, :, 0]) + (blk[:, :, 0, :, 1] + blk[:, :, 1, :, 1])
        np.testing.assert_array_equal(out, exp.astype(f32))
        outs = o.read_flat_short(sh, th)
        b = sh[rows].reshape(len(rows), A // 2, 2, E // 2, 2).astype(np.int32)
        exps = (b[:, :, 0, :, 0] + b[:, :, 1, :, 0]) + (b[:, :, 0, :, 1] + b[:, :, 1, :, 1])
        np.testing.assert_array_equal(outs, exps.astype(np.int16))   # int16 wrap-around


def test_rebin_norm_harmonic_gaps_and_skip(tinyrb):
    """Norm rebinning: 1/sum(1/n) over the positive contributors; a rebinned bin with fewer than
    four contributors that are ever positive in the segment is a gap (0 in every plane); the
    first and last npskip planes of every segment are zero."""
    o, G = tinyrb
    S, A, E = file_planes(o)
    rng = np.random.default_rng(3)
    n = (rng.random((S, A, E)) * 2 + 0.5).astype(f32)
    th = 2
    rows = G.seg_rows(th)
    mid = rows[len(rows) // 2]
    n[rows, 10, 20] = 0.0                  # dead crystal pair in every plane: gap at (5, 10)
    n[mid, 30, 40] = 0.0                   # zero in one plane only: three contributors there
    out = o.read_flat_float(n, th, 1.0, 1)
    skip = o.params.npskip
    assert np.all(out[:skip] == 0) and np.all(out[-skip:] == 0)
    assert np.all(out[:, 5, 10] == 0)
    i = mid - rows[0]
    c = n[mid, 30:32, 40:42]
    assert out[i, 15, 20] == pytest.approx(1.0 / np.sum(1.0 / c[c > 0]), rel=1e-6)
    inner = slice(skip, len(rows) - skip)
    blk = n[rows].reshape(len(rows), A // 2, 2, E // 2, 2).astype(np.float64)
    with np.errstate(divide="ignore"):
        h = 1.0 / (1.0 / blk).sum(axis=(2, 4))
    ok = np.ones((A // 2, E // 2), bool)
    ok[5, 10] = ok[15, 20] = False
    np.testing.assert_allclose(out[inner][:, ok], h[inner][:, ok], rtol=1e-6)


def test_prepare_norm_and_mask(tiny):
    """invnorm = N>0 ? (N*max(A,1) > eps2 ? 1/(N*max(A,1)) : 0) : N, and norm_zero_mask marks
    the columns (v, v+views/2, x, N-x) that are zero in every plane."""
    o, G = tiny
    S, A, E = file_planes(o)
    rng = np.random.default_rng(4)
    nrm = (rng.random((S, A, E)) + 0.5).astype(f32)
    att = (rng.random((S, A, E)) * 3).astype(f32)        # some < 1: clamped to 1
    nrm[:, 7, 11] = 0.0
    nrm[:, 7 + A // 2, 11] = 0.0
    nrm[:, 7, E - 11] = 0.0
    nrm[:, 7 + A // 2, E - 11] = 0.0                    # column pair (7, 11) dead
    nrm[:, 8, 12] = 0.0                                 # only one of four: column pair alive
    nrm[40, 3, 3] = -2.0                                # negative norms are passed through
    nrm[41, 3, 3] = 1e-3                                # N*A <= eps2 -> 0
    out = o.prepare_norm(nrm, att)
    mask = o.norm_zero_mask
    assert mask[7, 11] == 0 and mask[8, 12] == 1
    assert np.count_nonzero(mask[:A // 2, 1:] == 0) == 1
    # npskip planes of every segment are zero before weighting
    n_eff = nrm.copy()
    for th in range(o.th):
        r = G.seg_rows(th)
        n_eff[r[:o.params.npskip]] = 0
        n_eff[r[-o.params.npskip:]] = 0
    den = n_eff * np.maximum(att, f32(1))
    exp = np.where(n_eff > 0, np.where(den > f32(1e-2), f32(1) / np.where(den == 0, f32(1), den), f32(0)), n_eff)
    np.testing.assert_array_equal(out, exp.astype(f32))
    assert out[40, 3, 3] == -2.0 and out[41, 3, 3] == 0.0


def test_prepare_w3(tiny):
    """add = ((d * N) + s) * A; prompts are zeroed where the (skip-masked) norm is zero."""
    o, G = tiny
    S, A, E = file_planes(o)
    rng = np.random.default_rng(5)
    pr = rng.integers(0, 30, (S, A, E)).astype(np.int16)
    d = rng.random((S, A, E)).astype(f32)
    nrm = (rng.random((S, A, E)) + 0.5).astype(f32)
    nrm[:, 2, 3] = 0
    sc = rng.random((S, A, E)).astype(f32)
    at = (rng.random((S, A, E)) + 1).astype(f32)
    prompt, add = o.prepare_w3(pr, d, nrm, sc, at)
    n_eff = nrm.copy()
    for th in range(o.th):
        r = G.seg_rows(th)
        n_eff[r[:o.params.npskip]] = 0
        n_eff[r[-o.params.npskip:]] = 0
    np.testing.assert_array_equal(add, ((d * n_eff) + sc) * at)
    np.testing.assert_array_equal(prompt, np.where(n_eff == 0, np.int16(0), pr))
    # absent files: randoms 0, norm 1, atten 1
    prompt2, add2 = o.prepare_w3(pr, None, None, None, None)
    assert np.all(add2 == 0)
    np.testing.assert_array_equal(prompt2, pr)


# ---------------- drivers (CalculateNormfac S3:3760, CalculateOsem3d S3:4098) ----------------

def test_gather_pair(tiny):
    o, _ = tiny
    rng = np.random.default_rng(6)
    s = rng.random((o.S, o.views, o.N)).astype(f32)
    for v in (0, 5, o.views // 2 - 1):
        g = o.gather_pair(s, v)
        np.testing.assert_array_equal(g[:o.S], s[:, v])
        np.testing.assert_array_equal(g[o.S:], s[:, v + o.views // 2])


def test_normfac_is_inverse_of_subset_backprojection(tiny):
    o, _ = tiny
    rng = np.random.default_rng(7)
    inv = (rng.random((o.S, o.views, o.N)) + 0.5).astype(f32)
    work = inv.copy()
    nf = o.calculate_normfac(work)
    # the back projector rescales its input rows in place (PJ:1638-1649), so the reference's
    # sinogram leaves CalculateNormfac scaled by cos^3(theta) span ring_spacing / (2R)
    G = bf.Geometry(**GEOMS["tiny"])
    for th in (0, 1, 4):
        g = (th + 1) // 2
        c = 1.0 if g == 0 else G.cos_nominal[g - 1]
        r = G.seg_rows(th)
        np.testing.assert_allclose(work[r], inv[r] * (c ** 3 * 3 * G.ring_spacing / G.two_r), rtol=1e-6)
    for s in (0, 5):
        sens = o.image_zeros()
        for j in range(o.sviews // 2):
            v = int(o.vieworder[s, j])
            o.back_direct(o.gather_pair(inv, v), v, ima=sens)
        np.testing.assert_array_equal(nf[s], o.normfac_invert(sens))


def poisson_study(o, G, seed):
    rng = np.random.default_rng(seed)
    img = np.zeros((o.N, o.N, o.Z4), f32)
    xs = np.arange(o.N) - o.N // 2
    inside = (xs[:, None] ** 2 + xs[None, :] ** 2) <= (0.35 * o.N) ** 2
    img[inside, 6:26] = 1.0
    img[o.N // 2 - 6:o.N // 2 - 2, o.N // 2:o.N // 2 + 4, 12:16] = 4.0
    sino = np.zeros((o.S, o.views, o.N), f32)
    for v in range(o.views // 2):
        p = o.forward(img, v)
        sino[:, v] = p[:o.S]
        sino[:, v + o.views // 2] = p[o.S:]
    add = np.full_like(sino, 0.2 * sino.mean())
    prompt = rng.poisson(sino + add).astype(np.int16)
    return prompt, add


def loglik(o, image, prompt, add):
    tot = 0.0
    for v in range(o.views // 2):
        m = o.forward(image, v).astype(np.float64) + o.gather_pair(add, v)
        y = o.gather_pair(prompt, v).astype(np.float64)
        ok = m > 0
        tot += float((y[ok] * np.log(m[ok]) - m[ok]).sum())
    return tot


def test_osem_raises_likelihood_and_schedules_agree(tiny):
    """One OSEM iteration from the uniform start raises the Poisson log-likelihood, and the
    reference's thread schedule (incremental back projection for most view pairs) gives the
    same images as the all-direct schedule."""
    o, G = tiny
    prompt, add = poisson_study(o, G, 8)
    inv = np.ones((o.S, o.views, o.N), f32)
    nf = o.calculate_normfac(inv)
    img_a = o.init_image()
    img_b = img_a.copy()
    l0 = loglik(o, img_a, prompt, add)
    corr = o.image_zeros()
    for i, s in enumerate(o.order):
        o.osem_subset(img_a, corr, nf[s], add, prompt, int(s), nthreads=0, workers=2)
        o.osem_subset(img_b, corr, nf[s], add, prompt, int(s), nthreads=2, workers=2)
        if i == 0:
            assert np.linalg.norm(img_a - img_b) / np.linalg.norm(img_a) < 2e-5
    # float differences of ~1e-5 per back projection compound over 16 multiplicative updates
    assert np.linalg.norm(img_a - img_b) / np.linalg.norm(img_a) < 1e-3
    l1 = loglik(o, img_a, prompt, add)
    assert l1 > l0
