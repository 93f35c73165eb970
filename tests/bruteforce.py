"""Float64 restatements from first principles, used to pin the oracle.

Test infrastructure.  Nothing here is derived from the oracle's code: the
geometry comes from the scanner description (ring pairs, circle, angles), the
projectors from the line-integral model they discretise:

* forward, view v, bin x, segment 2k plane yr: sum over the FOV chord rows y'
  of the image sampled (bilinear in x/y after rotating by psi_v, linear in z)
  at z = yr + (y' - c) * tan(theta(yr)); segment 2k-1 uses z = yr - (...);
  scaled by y_size / cos(theta(yr)).  theta(yr) is the angle of the mean
  ring difference of the LORs that fall in plane yr.
* back, view pair v: A(X, Y, z) sums view v bin X at planes z +/- (c-Y) tan,
  B(X, Y, z) sums view v+90deg bin Y at planes z +/- (X-c) tan (nominal
  angles), both after scaling by cos^3(theta) span ring_spacing / (2R);
  the result is rotated back by -psi_v.

The end-plane rules of the reference that are not physics (which rows are
skipped, truncation towards zero) are written out where they apply and named
in the docstring of the function that applies them.
"""
from __future__ import annotations

import math

import numpy as np

HRRT = dict(rings=104, crystal_radius=234.5, plane_sep=1.21875, bin_size=1.21875,
            def_elements=256, def_angles=288)
DEFAULTS = dict(subsets=16, rel_fov=0.95, npskip=3, z_min=6, z_max=200,
                epsilon1=5e-2, epsilon2=1e-2, epsilon3=1e-10, image_max=1e4)


class Geometry:
    """Scanner/sinogram geometry in float64 from the scanner description."""

    def __init__(self, **kw):
        p = dict(HRRT)
        p.update(DEFAULTS)
        p.update(dict(x_pixels=0, span=9, maxdel=67))
        p.update(kw)
        self.p = p
        rings, span = p["rings"], p["span"]
        self.rebin = p["def_angles"] // p["x_pixels"] if 0 < p["x_pixels"] < p["def_elements"] else 1
        self.N = N = p["def_elements"] // self.rebin
        self.views = p["def_angles"] // self.rebin
        self.c = N // 2
        self.Z = 2 * rings - 1
        self.Z4 = 4 * ((self.Z + 3) // 4)
        self.ring_spacing = 2.0 * p["plane_sep"]
        self.two_r = 2.0 * p["crystal_radius"]
        self.y_size = p["bin_size"] * self.rebin
        self.gm = (p["maxdel"] - (span - 1) // 2) // span
        self.zs = span - 1
        self.th = 2 * self.gm + 1
        # plane range of each segment: a ring pair (r1, r2) lands in plane r1 + r2
        self.yb = [0] + [0] * (2 * self.gm)
        self.yt = [self.Z - 1] + [0] * (2 * self.gm)
        for g in range(1, self.gm + 1):
            dmin = span * g - (span - 1) // 2
            self.yb[2 * g - 1] = self.yb[2 * g] = dmin
            self.yt[2 * g - 1] = self.yt[2 * g] = 2 * (rings - 1) - dmin
        self.nplanes = [self.yt[s] - self.yb[s] + 1 for s in range(self.th)]
        self.seg_off = list(np.cumsum([0] + self.nplanes[:-1]))
        self.S = int(sum(self.nplanes))
        # axial tangents: mean ring difference of the LORs in each plane
        self.tan = np.zeros((self.gm, self.Z))
        self.tan_nominal = np.zeros(self.gm)
        for g in range(1, self.gm + 1):
            dmin = span * g - (span - 1) // 2
            ds = np.arange(dmin, dmin + span)
            self.tan_nominal[g - 1] = span * g * self.ring_spacing / self.two_r
            for yr in range(self.yb[2 * g], self.yt[2 * g] + 1):
                ok = [d for d in ds if (yr - d) % 2 == 0 and yr - d >= 0 and (yr + d) // 2 <= rings - 1]
                self.tan[g - 1, yr] = np.mean(ok) * self.ring_spacing / self.two_r
        self.cos_nominal = 1.0 / np.sqrt(1.0 + self.tan_nominal ** 2)
        # FOV: pixel centres within rel_fov * N/2 pixels of the centre
        r = p["rel_fov"] * N / 2.0
        self.cyl = np.zeros((N, 2), np.int64)
        for y in range(N):
            inside = [x for x in range(N) if math.hypot(x - self.c, y - self.c) <= r]
            if inside:
                self.cyl[y] = (inside[0], inside[-1] + 1)
        self.fov_r = r
        self.psi = np.arange(self.views) * math.pi / self.views

    def seg_rows(self, s):
        return np.arange(self.seg_off[s], self.seg_off[s] + self.nplanes[s])

    def order(self, sets):
        """Subsets ordered so that each is as far (circularly) from the previous as possible."""
        out, done = [0], {0}
        for _ in range(1, sets):
            k = out[-1]
            best = max((j for j in range(sets) if j not in done),
                       key=lambda j: (min(abs(j - k), sets - abs(j - k)), -j))
            out.append(best)
            done.add(best)
        return out


def bilinear(img, xx, yy):
    """Columns img[x][y][:] interpolated at (xx, yy) (arrays), bilinear in x and y."""
    xi = np.floor(xx).astype(np.int64)
    yi = np.floor(yy).astype(np.int64)
    fx = (xx - xi)[:, None]
    fy = (yy - yi)[:, None]
    return (img[xi, yi] * (1 - fx) * (1 - fy) + img[xi + 1, yi] * fx * (1 - fy)
            + img[xi, yi + 1] * (1 - fx) * fy + img[xi + 1, yi + 1] * fx * fy)


def lerp_trunc(col, zz, zz_mirror, keep):
    """Rows col[r] (len Z+1, plane Z zero) sampled at positions zz[r, j].

    The reference interpolates a row and its mirror row together: the lower
    plane int(zz) gets the fractional part of the mirror position
    zz_mirror = 2*yr - zz, the upper plane the fractional part of zz, with
    float-to-int conversion truncating towards zero (PJ:1929-1950).  For
    zz >= 0 this is ordinary linear interpolation.  Rows with keep False are
    ignored (their positions may lie outside the column)."""
    Z = col.shape[1] - 1
    z = np.trunc(zz).astype(np.int64)
    zk = z[keep]
    assert zk.size == 0 or (zk.min() >= 0 and zk.max() <= Z - 1), "kept sample outside the image"
    z = np.clip(z, 0, Z - 1)
    w_hi = zz - np.trunc(zz)
    w_lo = zz_mirror - np.trunc(zz_mirror)
    r = np.arange(col.shape[0])[:, None]
    return (col[r, z] * w_lo + col[r, z + 1] * w_hi) * keep[:, None]


def lerp_zero(row, lo, zz):
    """Zero-extended linear interpolation of row (planes lo .. lo+len-1) at plane positions zz."""
    ext = np.concatenate([[0.0], row, [0.0]])
    pos = zz - lo + 1
    out = np.zeros_like(zz, dtype=np.float64)
    ok = (pos > 0) & (pos < len(ext) - 1)
    i = np.floor(pos[ok]).astype(np.int64)
    f = pos[ok] - i
    out[ok] = ext[i] * (1 - f) + ext[i + 1] * f
    return out


def forward(G: Geometry, image, view, mask_row=None):
    """Forward projection of one view pair -> [2S][N] (bins never written stay 0).

    Reference rules applied beyond the line-integral model: bin 0 is never
    written (PJ:1805); for each mirrored row pair (y', N-y'), y' < c, a segment
    pair k is skipped when the first end plane of row y' or the last end plane
    of row N-y' would sample outside planes [0, Z-1] after truncation towards
    zero (PJ:1920-1921); plane Z (beyond the image) reads as zero.
    """
    N, c, Z = G.N, G.c, G.Z
    img = np.zeros((N + 1, N + 1, Z + 1))
    img[:N, :N, :Z] = image[:, :, :Z]
    cs, sn = math.cos(G.psi[view]), math.sin(G.psi[view])
    out = np.zeros((2 * G.S, N))
    for x in range(1, c + 1):
        lo, hi = G.cyl[x]
        if lo == hi or lo == 0:
            continue
        if mask_row is not None and mask_row[x] == 0:
            continue
        ys = np.arange(lo, N - lo + 1)
        pos = [(np.full_like(ys, x), ys), (np.full_like(ys, N - x), ys),
               (N - ys, np.full_like(ys, x)), (N - ys, np.full_like(ys, N - x))]
        cols = []
        for px, py in pos:
            xx = (px - c) * cs - (py - c) * sn + c
            yy = (px - c) * sn + (py - c) * cs + c
            assert xx.min() >= 0 and xx.max() <= N - 2 and yy.min() >= 0 and yy.max() <= N - 2
            cols.append(bilinear(img, xx, yy))        # [len(ys)][Z+1]
        dy = (ys - c).astype(np.float64)
        lanes = [(0, x), (0, N - x), (1, x), (1, N - x)]
        for (half, xr), col in zip(lanes, cols):
            out[half * G.S + G.seg_rows(0), xr] = col[:, :Z].sum(0) * G.y_size
        mirror = np.where(ys < c, ys, N - ys)
        for k in range(1, G.gm + 1):
            yb, yt = G.yb[2 * k], G.yt[2 * k]
            t = G.tan[k - 1]
            # row-pair skip rule (end-plane sample indices after truncation)
            first = np.trunc(yb + (mirror - c) * t[yb])
            last = np.trunc(yt + (c - mirror) * t[yt])
            keep = (mirror != c) & ~((first < 0) | (last > Z - 1))
            centre = int(np.nonzero(ys == c)[0][0])     # sampled without interpolation
            planes = np.arange(yb, yt + 1)
            tt = t[planes]
            sc = G.y_size * np.sqrt(1.0 + tt ** 2)
            zp = planes[None, :] + dy[:, None] * tt[None, :]
            zm = planes[None, :] - dy[:, None] * tt[None, :]
            for (half, xr), col in zip(lanes, cols):
                mid = col[centre, planes]
                out[half * G.S + G.seg_off[2 * k] + planes - yb, xr] = (lerp_trunc(col, zp, zm, keep).sum(0) + mid) * sc
                out[half * G.S + G.seg_off[2 * k - 1] + planes - yb, xr] = (lerp_trunc(col, zm, zp, keep).sum(0) + mid) * sc
    return out


def back_rotframe(G: Geometry, prj, view, mask_row=None):
    """Back projection of one view pair into the rotated frame (A + B), [N][N][Z4].

    Reference rules: a masked column (norm_zero_mask 0) contributes nothing
    except on its centre row (PJ:3048-3083 is not masked)."""
    N, c, Z = G.N, G.c, G.Z
    p = prj.astype(np.float64).copy()
    for s in range(G.th):
        g = (s + 1) // 2
        cth = 1.0 if g == 0 else G.cos_nominal[g - 1]
        p[G.seg_rows(s)] *= cth ** 3 * G.p["span"] * G.ring_spacing / G.two_r
        p[G.S + G.seg_rows(s)] *= cth ** 3 * G.p["span"] * G.ring_spacing / G.two_r
    out = np.zeros((N, N, G.Z4))
    zz = np.arange(Z, dtype=np.float64)[None, :]

    def columns(rows_half, xr, shift):
        """Back-projected columns for bin xr at axial shifts shift[i] (planes per unit tan)."""
        acc = np.repeat(p[rows_half + G.seg_rows(0), xr][None, :], len(shift), 0)
        for k in range(1, G.gm + 1):
            t = G.tan_nominal[k - 1]
            yb = G.yb[2 * k]
            acc += lerp_zero(p[rows_half + G.seg_rows(2 * k), xr], yb, zz + shift[:, None] * t)
            acc += lerp_zero(p[rows_half + G.seg_rows(2 * k - 1), xr], yb, zz - shift[:, None] * t)
        return acc

    for x in range(1, c + 1):
        lo, hi = G.cyl[x]
        if lo == 0 and hi == 0:
            continue
        masked = mask_row is not None and mask_row[x] == 0
        chord = np.arange(lo, N - lo + 1)
        if masked:
            chord = np.array([c])
        for X in sorted({x, N - x}):
            out[X, chord, :Z] += columns(0, X, (c - chord).astype(np.float64))
        for Yp in sorted({x, N - x}):
            out[chord, Yp, :Z] += columns(G.S, Yp, (chord - c).astype(np.float64))
    return out


def rotate(G: Geometry, buf, view):
    """out[x][y] = buf bilinear at R(-psi_v) (x - c, y - c) + c over the FOV rows."""
    N, c = G.N, G.c
    cs, sn = math.cos(G.psi[view]), math.sin(G.psi[view])
    out = np.zeros_like(buf, dtype=np.float64)
    b = np.zeros((N + 1, N + 1, buf.shape[2]))
    b[:N, :N] = buf
    for x in range(N):
        lo, hi = G.cyl[x]
        if lo == hi:
            continue
        ys = np.arange(lo, hi)
        xx = (x - c) * cs + (ys - c) * sn + c
        yy = -(x - c) * sn + (ys - c) * cs + c
        ok = (np.trunc(xx) >= 0) & (np.trunc(xx) <= N - 2) & (np.trunc(yy) >= 0) & (np.trunc(yy) <= N - 2)
        out[x, ys[ok]] = bilinear(b, xx[ok], yy[ok])
    return out


def gather_pair(sino, view):
    """[S][views][N] -> [2S][N] rows of view and view + views/2."""
    V = sino.shape[1]
    return np.concatenate([sino[:, view], sino[:, view + V // 2]], 0)


def sensitivity(G: Geometry, invnorm, subset, subsets):
    """FP64 sum over the subset's view pairs of the rotated back projection of 1/(N*A)."""
    out = np.zeros((G.N, G.N, G.Z4))
    for j in range(G.views // subsets // 2):
        v = subset + subsets * j
        out += rotate(G, back_rotframe(G, gather_pair(invnorm, v).astype(np.float64), v), v)
    return out


def invert(G: Geometry, sens, eps1):
    out = np.zeros_like(sens)
    for x in range(G.N):
        lo, hi = G.cyl[x]
        s = sens[x, lo:hi, :G.Z]
        out[x, lo:hi, :G.Z] = np.where(np.abs(s) > eps1, 1.0 / np.where(s == 0, 1, s), 0.0)
    return out


def osem_subset(G: Geometry, image, normfac, add, prompt, subset, subsets, eps3=1e-10, image_max=1e4):
    """FP64 OSEM subset update with the OP ratio and the clamped multiplicative update."""
    corr = np.zeros((G.N, G.N, G.Z4))
    for j in range(G.views // subsets // 2):
        v = subset + subsets * j
        est = forward(G, image, v)
        a = gather_pair(add, v).astype(np.float64)
        y = gather_pair(prompt, v).astype(np.float64)
        den = est + a
        ratio = np.where(y == 0, 0.0, np.where(den > eps3, y / np.where(den > eps3, den, 1.0), 1.0))
        corr += rotate(G, back_rotframe(G, ratio, v), v)
    out = image.copy()
    for x in range(G.N):
        lo, hi = G.cyl[x]
        c = np.clip(corr[x, lo:hi], 0.01, 100.0)
        out[x, lo:hi] = np.minimum(image_max, image[x, lo:hi] * (c * normfac[x, lo:hi]))
    return out
