"""Pins the oracle's PSF resolution model (-B, psf.cpp) without the reference.

- Kernel: size from the reference's formula worked by hand for the HRRT defaults, symmetry,
  unit sum, and the tap ratios of two Gaussians of FWHM 0.9 and 2.5 mm x sqrt(8 ln 2) (the
  defaults) computed here from the definition of a Gaussian's FWHM.
- convolve3d: an interior point source gives, in closed form, the separable product of the
  kernels with the last tap of each dropped (the reference's loops stop at h-1), divided by
  the sums of the taps used; a constant image stays constant; the map is linear.
- The drivers with -B equal the composition of the oracle's own pinned pieces.
"""
import numpy as np
import pytest

from conftest import GEOMS

FWHM_TO_SIGMA = 1.0 / np.sqrt(8.0 * np.log(2.0))


@pytest.fixture(scope="module")
def tiny(oracle_lib):
    o = oracle_lib.Oracle(**GEOMS["tiny"])
    o.set_psf()
    return o


def test_kernel_size_and_shape(oracle_lib):
    # 2*int(1.5*fwhm2*sqrt(log10/log2)/px/2)+1 with fwhm2 = 2.5*2.3548 mm: 6.60 -> 13 taps at
    # 1.21875 mm, 3.30 -> 7 taps at 2.4375 mm
    for geom, sizes in (("c2", (13, 13, 13)), ("c1", (7, 7, 13)), ("tiny", (13, 13, 13))):
        o = oracle_lib.Oracle(**GEOMS[geom])
        o.set_psf()
        for axis, n in enumerate(sizes):
            k = o.psf_kernel(axis).astype(np.float64)
            assert len(k) == n, (geom, axis)
            assert np.allclose(k, k[::-1], rtol=0, atol=1e-9)
            assert abs(k.sum() - 1.0) < 1e-6
            assert k.argmax() == n // 2


def test_kernel_tap_ratios_from_fwhm(oracle_lib):
    o = oracle_lib.Oracle(**GEOMS["c2"])
    o.set_psf()
    k = o.psf_kernel(0).astype(np.float64)
    h = len(k) // 2
    px = 1.21875
    s1 = 0.9 * np.sqrt(8 * np.log(2)) * FWHM_TO_SIGMA / px      # sigma of the narrow Gaussian, pixels
    s2 = 2.5 * np.sqrt(8 * np.log(2)) * FWHM_TO_SIGMA / px
    for d in range(1, h + 1):
        g = np.exp(-0.5 * (d / s1) ** 2) + 0.05 * np.exp(-0.5 * (d / s2) ** 2)
        assert abs(k[h + d] / k[h] - g / 1.05) < 2e-6 * (g / 1.05) + 1e-9, d
    # explicit parameters: a single Gaussian of FWHM 4 mm has its half maximum 2 mm off centre
    o.set_psf(4.0, 0.0, 0.0)
    k = o.psf_kernel(0).astype(np.float64)
    h = len(k) // 2
    x = np.arange(len(k)) - h
    sigma = 4.0 * FWHM_TO_SIGMA / px
    ref = np.exp(-0.5 * (x / sigma) ** 2)
    assert np.allclose(k, ref / ref.sum(), rtol=2e-6)
    with pytest.raises(ValueError):
        o.set_psf(-1.0, 0.0, 0.0)


def test_point_source_closed_form(tiny):
    o = tiny
    kx, ky, kz = (o.psf_kernel(a).astype(np.float64) for a in range(3))
    h = len(ky) // 2
    x0, y0, z0 = o.N // 2 + 3, o.N // 2 - 5, (o.params.z_min + o.params.z_max) // 2
    img = o.image_zeros()
    img[x0, y0, z0] = 2.5
    out = o.convolve3d(img).astype(np.float64)
    ref = np.zeros_like(out)
    nx, ny, nz = kx[:2 * h].sum(), ky[:2 * h].sum(), kz[:2 * h].sum()   # sums of the taps m = -h .. h-1
    for dx in range(-h + 1, h + 1):
        for dy in range(-h + 1, h + 1):
            for dz in range(-h + 1, h + 1):
                ref[x0 + dx, y0 + dy, z0 + dz] = (2.5 * kx[h - dx] * ky[h - dy] * ky[h - dz]) / (nx * ny * nz)
    assert np.abs(out - ref).max() < 1e-6 * ref.max()
    assert out[x0 - h, y0, z0] == 0.0      # the dropped tap


def test_constant_image_is_kept(tiny):
    o = tiny
    img = o.init_image() * np.float32(3.0)          # 3 inside the FOV and z in [z_min, z_max)
    out = o.convolve3d(img)
    fov = img[:, :, o.params.z_min] > 0
    zin = np.arange(o.Z4)
    h = len(o.psf_kernel(1)) // 2
    reach = (zin + h > o.params.z_min) & (zin - h < o.params.z_max - 1) & (zin < o.Z)
    sel = out[fov][:, reach]
    assert np.allclose(sel, 3.0, rtol=2e-6)
    assert np.all(out[~fov] == 0)


def test_linearity(tiny):
    o = tiny
    rng = np.random.default_rng(3)
    u = (o.init_image() * rng.uniform(0, 2, (o.N, o.N, o.Z4))).astype(np.float32)
    v = (o.init_image() * rng.uniform(0, 2, (o.N, o.N, o.Z4))).astype(np.float32)
    lhs = o.convolve3d((np.float32(0.5) * u + np.float32(2.0) * v).astype(np.float32)).astype(np.float64)
    rhs = 0.5 * o.convolve3d(u).astype(np.float64) + 2.0 * o.convolve3d(v).astype(np.float64)
    assert np.abs(lhs - rhs).max() < 1e-5 * np.abs(rhs).max()


def test_osem_subset_psf_is_the_composition(tiny):
    o = tiny
    rng = np.random.default_rng(9)
    shape = (o.S, o.views, o.N)
    prompt = rng.integers(0, 4, shape).astype(np.int16)
    add = rng.uniform(0.0, 0.2, shape).astype(np.float32)
    nf = rng.uniform(0.05, 2.0, (o.N, o.N, o.Z4)).astype(np.float32)
    s = 5
    img = o.init_image() * rng.uniform(0.5, 1.5, (o.N, o.N, o.Z4)).astype(np.float32)
    img = np.ascontiguousarray(img, np.float32)
    ref = img.copy()
    o.osem_subset_psf(img, o.image_zeros(), nf, add, prompt, s, nthreads=0)
    # composition: blurred image -> forward -> ratio -> direct back projection -> sum, blur, update
    blurred = o.convolve3d(ref)
    corr = o.image_zeros()
    for v in o.vieworder[s][: o.sviews // 2]:
        est = o.forward(blurred, int(v))
        est = o.ratio_w3(est, o.gather_pair(add, int(v)), o.gather_pair(prompt, int(v)))
        o.back_direct(est, int(v), ima=corr)
    corr = o.convolve3d(corr, out=corr)
    o.update_image(ref, corr, nf)
    assert np.array_equal(img, ref)


def test_normfac_psf_is_the_composition(tiny):
    o = tiny
    rng = np.random.default_rng(10)
    inv = rng.uniform(0.5, 1.5, (o.S, o.views, o.N)).astype(np.float32)
    nf = o.calculate_normfac_psf(inv.copy(), nthreads=0)
    inv2 = inv.copy()
    for s in (0, 7):
        sens = o.image_zeros()
        for v in o.vieworder[s][: o.sviews // 2]:
            o.back_direct(o.gather_pair(inv2, int(v)), int(v), ima=sens)
        sens = o.convolve3d(sens, out=sens)
        assert np.array_equal(nf[s], o.normfac_invert(sens)), s
