"""The FP64 oracle (liboracle64.so), pinned on the CPU.

The FP64 build runs the oracle's projectors, steps and drivers with double values while
keeping the reference's float geometry: rotation positions stepped in float, float axial
positions, float scale factors.  It therefore applies the same system matrix as the
reference and differs from the float oracle only by arithmetic rounding.  Pinned by:

- first-principles float64 line integrals (tests/bruteforce.py), to the same bounds the
  float oracle meets (forward 2e-5, back projection 2e-6 max relative): a wrong lane, sign
  or scale in the double build fails these;
- agreement with the float oracle to float rounding (rel-L2 of one forward projection
  <= 5e-7, measured ~1e-7; of one rotation <= 2e-7, measured ~5e-8): a change of
  discretisation in the double build fails this -- rotation positions stepped in double
  instead of float move the rotation by ~3e-5 (checked on a planted copy);
- the incremental and the direct back projectors agreeing to 1e-7 in double (their float
  coefficients aside), several times closer than the float oracle's two (~3e-7): the
  double build really carries double sums;
- whole subsets and the sensitivity build against the float oracle (rel-L2 <= 1e-5).
"""
import numpy as np
import pytest

from conftest import GEOMS
import bruteforce as bf


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_max(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max())


@pytest.fixture(scope="module", params=["tiny", "tiny9", "c1"])
def trio(request, oracle_lib):
    kw = GEOMS[request.param]
    return (request.param, oracle_lib.Oracle(**kw), oracle_lib.Oracle(double=True, **kw), bf.Geometry(**kw))


def views_of(o, name):
    return [0, 1, o.views // 4] if name != "c1" else [0, 37]


def test_value_type_is_checked(oracle_lib):
    o64 = oracle_lib.Oracle(double=True, **GEOMS["tiny"])
    with pytest.raises(TypeError):
        o64.forward(np.zeros((o64.N, o64.N, o64.Z4), np.float32), 0)
    assert o64.image_zeros().dtype == np.float64


def test_forward_fp64(trio):
    name, o, o64, G = trio
    rng = np.random.default_rng(21)
    img = rng.random((o.N, o.N, o.Z4))
    for v in views_of(o, name):
        a64 = o64.forward(img, v)
        a32 = o.forward(img.astype(np.float32), v)
        b = bf.forward(G, img, v)
        assert rel_max(a64, b) < 2e-5, (name, v)
        assert rel_l2(a32, a64) < 5e-7, (name, v)
        assert not np.array_equal(a32.astype(np.float64), a64)   # really computed in double


def test_back_fp64(trio):
    name, o, o64, G = trio
    rng = np.random.default_rng(22)
    prj = rng.random((2 * o.S, o.N))
    for v in views_of(o, name):
        d64 = o64.back_direct_rotframe(prj, v)
        i64 = o64.back_incremental_rotframe(prj, v)
        d32 = o.back_direct_rotframe(prj.astype(np.float32), v)
        i32 = o.back_incremental_rotframe(prj.astype(np.float32), v)
        assert rel_max(d64, bf.back_rotframe(G, prj, v)) < 2e-6, (name, v)
        assert rel_l2(d32, d64) < 2e-6 and rel_l2(i32, d64) < 1e-5, (name, v)
        # the two back projectors differ by rounding (and their float coefficients): in double
        # the gap shrinks from ~3e-7 to ~1e-8
        assert rel_l2(i64, d64) < 1e-7, (name, v)
        assert rel_l2(i32, d32) > 2 * rel_l2(i64, d64), (name, v)


def test_rotation_fp64(trio):
    name, o, o64, G = trio
    rng = np.random.default_rng(23)
    buf = rng.random((o.N, o.N, o.Z4))
    for v in views_of(o, name):
        r64 = o64.rotate_add(buf, v)
        r32 = o.rotate_add(buf.astype(np.float32), v)
        assert rel_l2(r32, r64) < 2e-7, (name, v)
        # same float positions as the float oracle: the exact rotation differs by more
        assert rel_l2(r64, bf.rotate(G, buf, v)) < 3e-4


def test_subsets_and_sensitivity_fp64(oracle_lib):
    kw = GEOMS["tiny"]
    o, o64 = oracle_lib.Oracle(**kw), oracle_lib.Oracle(double=True, **kw)
    rng = np.random.default_rng(24)
    shape = (o.S, o.views, o.N)
    inv = rng.uniform(0.2, 1.5, shape)
    nf32 = o.calculate_normfac(inv.astype(np.float32), nthreads=0)
    nf64 = o64.calculate_normfac(inv.copy(), nthreads=0)
    assert rel_l2(nf32, nf64) < 1e-5
    prompt = rng.poisson(3.0, shape).astype(np.int16)
    add = rng.uniform(0.0, 0.3, shape)
    img32, img64 = o.init_image(), o64.init_image()
    for s in (0, 8, 1):
        o.osem_subset(img32, o.image_zeros(), np.ascontiguousarray(nf32[s]), add.astype(np.float32), prompt, s,
                      nthreads=0, workers=4)
        o64.osem_subset(img64, o64.image_zeros(), np.ascontiguousarray(nf64[s]), add, prompt, s, nthreads=0,
                        workers=4)
    assert rel_l2(img32, img64) < 1e-5
    data = rng.uniform(0.0, 4.0, shape)
    w32, w64 = o.init_image(), o64.init_image()
    o.osem_subset_w012(w32, o.image_zeros(), np.ascontiguousarray(nf32[3]), data.astype(np.float32), 3, workers=4)
    o64.osem_subset_w012(w64, o64.image_zeros(), np.ascontiguousarray(nf64[3]), data, 3, workers=4)
    assert rel_l2(w32, w64) < 1e-5
