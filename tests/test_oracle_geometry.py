"""The oracle's geometry tables against the scanner description (CPU only).

Each check derives the expected value from first principles (ring pairs,
circle, angles) or from a known answer, never from the oracle's formulas.
"""
import math

import numpy as np
import pytest

from conftest import GEOMS
import bruteforce as bf


@pytest.fixture(scope="module", params=["tiny", "tinyrb", "tiny9", "c1", "c2"])
def pair(request, oracle_lib):
    kw = GEOMS[request.param]
    return request.param, oracle_lib.Oracle(**kw), bf.Geometry(**kw)


def test_known_sizes(oracle_lib):
    # HRRT (104 rings, maxdel 67): span 9 -> 15 segments / 2209 planes, span 3 -> 45 / 6367
    o = oracle_lib.Oracle(**GEOMS["c1"])
    assert (o.N, o.views, o.Z, o.Z4, o.gm, o.th, o.S, o.zs) == (128, 144, 207, 208, 7, 15, 2209, 8)
    o = oracle_lib.Oracle(**GEOMS["c2"])
    assert (o.N, o.views, o.Z, o.Z4, o.gm, o.th, o.S, o.zs) == (256, 288, 207, 208, 22, 45, 6367, 2)


def test_sequence_known_answer(oracle_lib):
    # subset order of the 16-subset HRRT reconstructions
    o = oracle_lib.Oracle(**GEOMS["c1"])
    assert list(o.order) == [0, 8, 1, 9, 2, 10, 3, 11, 4, 12, 5, 13, 6, 14, 7, 15]
    assert list(o.order) == bf.Geometry(**GEOMS["c1"]).order(16)
    for sets in (2, 3, 4, 6, 8, 12):
        oo = oracle_lib.Oracle(**dict(GEOMS["tiny"], subsets=sets))
        assert list(oo.order) == bf.Geometry().order(sets)
        assert sorted(oo.order) == list(range(sets))


def test_segments(pair):
    name, o, G = pair
    assert o.th == G.th and o.S == G.S and o.Z == G.Z
    assert list(o.yr_bottom) == G.yb and list(o.yr_top) == G.yt
    assert list(o.segment_offset) == G.seg_off
    # file offsets: file order is 0, +1, -1, +2, -2, ...; internal odd segments read the
    # block after the swap (FIO:116-118), so internal segment s lands on file block s
    off = o.seg_offset // (o.N * o.views)       # in internal planes; the reader scales by the rebin
    assert off[0] == 0
    for g in range(1, G.gm + 1):
        assert off[2 * g] == G.seg_off[2 * g - 1]
        assert off[2 * g - 1] == G.seg_off[2 * g]


def test_angles(pair):
    name, o, G = pair
    for g in range(1, G.gm + 1):
        t = G.tan_nominal[g - 1]
        s, c = t / math.sqrt(1 + t * t), 1 / math.sqrt(1 + t * t)
        for k in (2 * g - 1, 2 * g):
            assert o.cos_theta[k] == pytest.approx(c, rel=1e-6)
        assert o.sin_theta[2 * g - 1] == pytest.approx(s, rel=1e-6)
        assert o.sin_theta[2 * g] == pytest.approx(-s, rel=1e-6)
    assert o.sin_theta[0] == 0 and o.cos_theta[0] == 1
    psi = G.psi
    np.testing.assert_allclose(o.sin_psi, np.sin(psi), atol=2e-7)
    # cos is sqrt(1 - sin^2) in float (PJ:476-486): one rounding of 1 - sin^2 (<= 6e-8) grows
    # by 1/(2 cos) towards 90 degrees
    bound = 1e-7 + 6e-8 / np.maximum(np.abs(np.cos(psi)), 1e-3)
    assert np.all(np.abs(o.cos_psi - np.cos(psi)) <= bound)
    assert o.sin_psi[o.views // 2] == 1 and o.cos_psi[o.views // 2] == 0


def test_axial_tangents_are_mean_ring_difference(pair):
    """tantheta[g][yr] = mean ring difference of the LORs in plane yr (ring pairs enumerated)
    times ring spacing / (2R); rescalearray = y_size / cos of that angle."""
    name, o, G = pair
    tt, rs = o.tantheta, o.rescalearray
    for g in range(1, G.gm + 1):
        planes = np.arange(G.yb[2 * g], G.yt[2 * g] + 1)
        np.testing.assert_allclose(tt[g - 1, planes], G.tan[g - 1, planes], rtol=2e-6)
        np.testing.assert_allclose(rs[g - 1, planes], G.y_size * np.sqrt(1 + G.tan[g - 1, planes] ** 2), rtol=2e-6)
        # the planes away from the segment ends hold the nominal angle
        mid = planes[G.zs:-G.zs] if G.zs else planes
        np.testing.assert_allclose(G.tan[g - 1, mid], G.tan_nominal[g - 1], rtol=1e-12)


def test_fov_rows(pair):
    name, o, G = pair
    cyl = o.cylwiny
    r = G.fov_r
    for y in range(o.N):
        d = [abs(math.hypot(x - G.c, y - G.c) - r) for x in range(o.N)]
        if min(d) < 1e-4:
            continue        # a pixel centre on the circle: float rounding decides
        assert tuple(cyl[y]) == tuple(G.cyl[y]), y


def test_forward_end_plane_lookup(pair):
    """flookup = integer and fractional part of yr + (y - c) tan(theta(yr)) for the end planes."""
    name, o, G = pair
    frac, zi = o.flookup()
    for y in (1, G.c // 2, G.c - 1, G.c + 1, G.N - 2):
        for g in range(1, G.gm + 1):
            yb, yt = G.yb[2 * g], G.yt[2 * g]
            planes = list(range(yb, yb + G.zs)) + list(range(yt - G.zs + 1, yt + 1))
            for i, yr in enumerate(planes):
                zz = yr + (y - G.c) * G.tan[g - 1, yr]
                assert zi[y, g - 1, i] == math.trunc(zz) or abs(zz - round(zz)) < 1e-4
                assert zi[y, g - 1, i] + frac[y, g - 1, i] == pytest.approx(zz, abs=2e-5)


def test_back_lookup(pair):
    """blookup: plane z of row y reads segment plane z + (c - y) tan_nominal, z limited to the
    planes whose interpolation touches the segment (with one zero guard plane each side)."""
    name, o, G = pair
    zsa, zea, yra, yr2a = o.blookup()
    for y in (1, G.c // 3, G.c - 1, G.c, G.c + 1, G.N - 2):
        for g in range(1, G.gm + 1):
            sh = (G.c - y) * G.tan_nominal[g - 1]
            fl = math.floor(sh)
            if abs(sh - round(sh)) < 1e-5:
                continue
            assert yr2a[y, g - 1] == pytest.approx(sh - fl, abs=2e-5)
            zs = G.yb[2 * g] - 1 - fl
            ze = G.yt[2 * g] - fl
            assert zsa[y, g - 1] == max(zs, 0)
            assert zea[y, g - 1] == min(ze, G.Z - 1)
            assert yra[y, g - 1] == G.yb[2 * g] - 1 + max(0, -zs)


def test_view_order(pair):
    name, o, G = pair
    vo = o.vieworder
    for s in range(o.subsets):
        assert list(vo[s]) == [s + o.subsets * j for j in range(o.sviews)]


def test_init_image_and_end_iteration(oracle_lib):
    o = oracle_lib.Oracle(**GEOMS["tiny"])
    G = bf.Geometry(**GEOMS["tiny"])
    img = o.init_image()
    exp = np.zeros_like(img)
    for x in range(o.N):
        lo, hi = G.cyl[x]
        exp[x, lo:hi, o.params.z_min:o.params.z_max] = 1
    np.testing.assert_array_equal(img, exp)
    rng = np.random.default_rng(3)
    v = (rng.random(img.shape) * 2e-16).astype(np.float32)
    v[::3] = -1.0
    w = v.copy()
    o.end_iteration(w)
    inside = exp > 0
    keep = ~inside | (v >= 1e-16)
    np.testing.assert_array_equal(w[keep], v[keep])
    assert np.all(w[~keep] == 0)
