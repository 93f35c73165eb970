"""The oracle's projectors against first-principles float64 restatements (CPU only).

tests/bruteforce.py samples the image along each line of response from the
scanner description; agreement to ~1e-5 pins the oracle's lane order,
mirror rows, signs of the axial shifts, end-plane angles and scale factors.
"""
import math

import numpy as np
import pytest

from conftest import GEOMS
import bruteforce as bf

VIEW_FRACTIONS = (0.0, 1.0 / 96, 0.25, 0.49)


@pytest.fixture(scope="module", params=["tiny", "tinyrb", "tiny9", "c1"])
def geo(request, oracle_lib):
    kw = GEOMS[request.param]
    return request.param, oracle_lib.Oracle(**kw), bf.Geometry(**kw)


def views_of(o, name):
    fr = VIEW_FRACTIONS if name != "c1" else (0.0, 0.49)
    return sorted({max(0, min(o.views // 2 - 1, int(round(f * o.views)))) for f in fr} | {1})


def rel_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_forward_matches_line_integrals(geo):
    name, o, G = geo
    rng = np.random.default_rng(11)
    img = rng.random((o.N, o.N, o.Z4)).astype(np.float32)
    for v in views_of(o, name):
        a = o.forward(img, v)
        b = bf.forward(G, img.astype(np.float64), v)
        # float32 accumulation over ~N samples against float64
        assert rel_max(a, b) < 2e-5, (name, v)


def test_back_matches_line_integrals(geo):
    name, o, G = geo
    rng = np.random.default_rng(12)
    prj = rng.random((2 * o.S, o.N)).astype(np.float32)
    for v in views_of(o, name):
        a = o.back_direct_rotframe(prj, v)
        b = bf.back_rotframe(G, prj, v)
        assert rel_max(a, b) < 2e-6, (name, v)


def test_rotation_matches_bilinear(geo):
    name, o, G = geo
    rng = np.random.default_rng(13)
    buf = rng.random((o.N, o.N, o.Z4)).astype(np.float32)
    for v in views_of(o, name):
        a = o.rotate_add(buf, v)
        b = bf.rotate(G, buf.astype(np.float64), v)
        # the reference steps the rotated coordinates incrementally in float along each row
        assert rel_max(a, b) < 1e-3, (name, v)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 3e-4


def test_incremental_back_matches_direct(geo):
    """The incremental back projector (running sums over y) equals the direct one."""
    name, o, G = geo
    rng = np.random.default_rng(14)
    prj = rng.random((2 * o.S, o.N)).astype(np.float32)
    for v in views_of(o, name):
        a = o.back_incremental_rotframe(prj, v)
        b = o.back_direct_rotframe(prj, v)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5, (name, v)
        assert rel_max(a, b) < 1e-4, (name, v)


def test_back_direct_rotates_into_image(oracle_lib):
    """back_direct = rotation of the rotated-frame back projection, added to the image."""
    o = oracle_lib.Oracle(**GEOMS["tiny"])
    rng = np.random.default_rng(15)
    prj = rng.random((2 * o.S, o.N)).astype(np.float32)
    base = rng.random((o.N, o.N, o.Z4)).astype(np.float32)
    for v in (0, 7, 30):
        got = o.back_direct(prj, v, ima=base.copy())
        exp = o.rotate_add(o.back_direct_rotframe(prj, v), v, out=base.copy())
        np.testing.assert_array_equal(got, exp)


def test_forward_uniform_cylinder_chords(oracle_lib):
    """A uniform cylinder of radius r projects to its chord length 2 sqrt(r^2 - s^2) (mm),
    divided by cos(theta) in the oblique segments."""
    kw = GEOMS["c1"]
    o = oracle_lib.Oracle(**kw)
    G = bf.Geometry(**kw)
    r_mm = 60.0
    xs = (np.arange(o.N) - o.N // 2) * G.y_size
    inside = (xs[:, None] ** 2 + xs[None, :] ** 2) <= r_mm ** 2
    img = np.zeros((o.N, o.N, o.Z4), np.float32)
    img[inside, :o.Z] = 1.0
    for v in (0, 20):
        prj = o.forward(img, v)
        s = xs[1:]
        chord = 2 * np.sqrt(np.maximum(r_mm ** 2 - s ** 2, 0))
        central = np.abs(s) < 0.8 * r_mm
        for seg in (0, 1, 2, 9, 14):
            g = (seg + 1) // 2
            cos_t = 1.0 if g == 0 else G.cos_nominal[g - 1]
            mid = G.seg_rows(seg)[len(G.seg_rows(seg)) // 2]
            for half in (0, 1):
                row = prj[half * o.S + mid, 1:]
                err = np.abs(row - chord / cos_t)[central]
                assert err.max() < 2.5 * G.y_size, (v, seg, half, err.max())
                assert abs(err.mean()) < 0.6 * G.y_size


def test_forward_is_linear_and_leaves_bin_zero(oracle_lib):
    o = oracle_lib.Oracle(**GEOMS["tiny9"])
    rng = np.random.default_rng(16)
    x = rng.random((o.N, o.N, o.Z4)).astype(np.float32)
    y = rng.random((o.N, o.N, o.Z4)).astype(np.float32)
    for v in (0, 5):
        px, py = o.forward(x, v), o.forward(y, v)
        pxy = o.forward((2 * x + 3 * y).astype(np.float32), v)
        assert rel_max(pxy, 2 * px + 3 * py) < 1e-5
        sentinel = np.full((2 * o.S, o.N), -7.0, np.float32)
        out = o.forward(x, v, prj=sentinel)
        assert np.all(out[:, 0] == -7.0)            # bin 0 is never written (PJ:1805)
        # bins outside the FOV are not written either; every written bin equals the zero-filled run
        written = out != -7.0
        np.testing.assert_array_equal(out[written], px[written])
        assert np.all(px[~written] == 0)
        G = bf.Geometry(**GEOMS["tiny9"])
        outside = [b for b in range(1, o.N // 2 + 1) if G.cyl[b][0] == G.cyl[b][1]]
        for b in outside:
            assert not written[:, b].any() and not written[:, o.N - b].any()


def test_masked_columns(oracle_lib):
    """norm_zero_mask 0: the forward projector skips the column, so its four bins carry the
    projection buffer's old contents (zero for a fresh buffer; the reference's buffer holds the
    previous view, PJ:1811-1813); the back projector drops them except on the column's centre
    row.  Such columns have zero prompts, so the ratio step zeroes them either way."""
    kw = GEOMS["tiny"]
    o = oracle_lib.Oracle(**kw)
    G = bf.Geometry(**kw)
    m = o.norm_zero_mask.copy()
    v = 9
    dead = [3, 10, 20, o.N // 2]
    m[v, dead] = 0
    o.set_norm_zero_mask(m)
    rng = np.random.default_rng(17)
    img = rng.random((o.N, o.N, o.Z4)).astype(np.float32)
    sentinel = np.full((2 * o.S, o.N), -3.0, np.float32)
    out = o.forward(img, v, prj=sentinel)
    for x in dead:
        assert np.all(out[:, x] == 0) and np.all(out[:, o.N - x] == 0)
    b = bf.forward(G, img.astype(np.float64), v, mask_row=m[v])
    live = out != -3.0
    assert np.all(b[~live] == 0)
    assert rel_max(out[live], b[live]) < 2e-5
    prj = rng.random((2 * o.S, o.N)).astype(np.float32)
    a = o.back_direct_rotframe(prj, v)
    bb = bf.back_rotframe(G, prj, v, mask_row=m[v])
    assert rel_max(a, bb) < 2e-6


def test_adjoint_scale(oracle_lib):
    """<P x, y> and <x, B y> differ only by the documented scale mismatch: forward scales by
    y_size/cos(theta), back by cos^3(theta) span ring_spacing/(2R) (segment 0: y_size vs
    span ring_spacing/(2R)), so on segment-0-only data the ratio is y_size*2R/(span*rs)."""
    kw = GEOMS["tiny"]
    o = oracle_lib.Oracle(**kw)
    G = bf.Geometry(**kw)
    rng = np.random.default_rng(18)
    # image constant along z inside [8, 23) and a sinogram holding only segment 0
    x = np.zeros((o.N, o.N, o.Z4), np.float32)
    x[:, :, 8:23] = rng.random((o.N, o.N, 1)).astype(np.float32)
    yv = np.zeros((2 * o.S, o.N), np.float32)
    rows = G.seg_rows(0)
    yv[rows] = rng.random((len(rows), o.N)).astype(np.float32)
    yv[o.S + rows] = rng.random((len(rows), o.N)).astype(np.float32)
    v = 4
    lhs = float((o.forward(x, v).astype(np.float64) * yv).sum())
    rhs = float((x.astype(np.float64) * o.back_direct(yv, v)).sum())
    ratio = G.y_size * G.two_r / (G.p["span"] * G.ring_spacing)
    # bilinear sampling and bilinear rotation are not exact transposes of each other
    assert lhs / rhs == pytest.approx(ratio, rel=2e-2)
    assert math.isfinite(ratio)
