"""Host side of the product without a GPU: the C ABI library, its host-only geometry query,
the pair-column layout helpers and the bench contract (CPU only)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import GEOMS, ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("hrrt_osem.h", "hrrt_randoms.h", "hrrt_lmhist.h")]


def declared_functions():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hrrt_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from hrrt_reconstruction_amd import _lib
    return _lib


def test_library_exports_every_declared_function(lib):
    names = declared_functions()
    assert len(names) >= 24
    L = C.CDLL(lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert sorted(lib.SIGNATURES) == names


def test_no_oracle_in_product():
    """The product package never loads the oracle (it is the checker only)."""
    pkg = os.path.join(ROOT, "hrrt_reconstruction_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h", "Makefile")):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"\boracle\b|liboracle|og_[a-z]", src), f


def test_missing_library_fails_loudly(lib, monkeypatch):
    monkeypatch.setattr(lib, "_lib", None)
    monkeypatch.setattr(lib, "LIB_PATH", os.path.join(ROOT, "does_not_exist.so"))
    with pytest.raises(RuntimeError, match="missing"):
        lib.lib()


@pytest.mark.parametrize("name", ["tiny", "tinyrb", "tiny9", "c1", "c2"])
def test_query_geometry_matches_oracle(name, oracle_lib):
    from hrrt_reconstruction_amd.osem import Layout, make_params
    kw = GEOMS[name]
    L = Layout(make_params(**kw))
    o = oracle_lib.Oracle(**kw)
    assert (L.N, L.views, L.Z, L.Z4, L.S, L.th, L.sviews, L.subsets) == \
        (o.N, o.views, o.Z, o.Z4, o.S, o.th, o.sviews, o.subsets)
    assert L.dims.x_rebin == L.dims.v_rebin == o.params.def_angles // o.views
    np.testing.assert_array_equal(L.yr_bottom, o.yr_bottom)
    np.testing.assert_array_equal(L.yr_top, o.yr_top)
    np.testing.assert_array_equal(L.segment_offset, o.segment_offset)
    np.testing.assert_array_equal(L.order, o.order)
    np.testing.assert_array_equal(L.vieworder, o.vieworder)
    # pair-column layout: every segment starts on a multiple of 4 with a zero guard plane on
    # each side, and the column holds them all
    sd = L.seg_data
    n = L.yr_top - L.yr_bottom + 1
    assert np.all(sd % 4 == 0) and sd[0] >= 1
    assert np.all(sd[1:] >= sd[:-1] + n[:-1] + 1)
    assert L.spp % 4 == 0 and L.spp >= sd[-1] + n[-1] + 1
    assert L.dims.image_floats == L.N * L.N * L.Z4
    assert L.dims.sino_bins == L.S * L.views * L.N
    assert L.dims.file_bins == L.S * o.params.def_angles * o.params.def_elements
    assert L.dims.paircol_elems == L.views // 2 * 2 * L.N * L.spp
    cyl = o.cylwiny
    assert L.dims.fov_voxels == int((cyl[:, 1] - cyl[:, 0]).sum())
    assert L.dims.fov_bins == int(sum(1 for x in range(1, L.N) if cyl[x, 1] > cyl[x, 0]))


def test_paircol_layout_round_trip():
    from hrrt_reconstruction_amd.osem import Layout, make_params
    L = Layout(make_params(**GEOMS["tiny"]))
    rng = np.random.default_rng(0)
    s = rng.random((L.S, L.views, L.N)).astype(np.float32)
    pc = L.sino_to_paircol(s)
    assert pc.shape == (L.views // 2, 2, L.N, L.spp)
    np.testing.assert_array_equal(L.paircol_to_sino(pc), s)
    # spot checks of the mapping and of the zero guard planes
    for seg in (0, 1, L.th - 1):
        yb = L.yr_bottom[seg]
        for v in (0, L.views // 2 - 1):
            for h in (0, 1):
                plane = L.segment_offset[seg] + 3
                assert pc[v, h, 7, L.seg_data[seg] + 3] == s[plane, h * L.views // 2 + v, 7]
        assert np.all(pc[:, :, :, L.seg_data[seg] - 1] == 0)
        assert np.all(pc[:, :, :, L.seg_data[seg] + L.yr_top[seg] - yb + 1] == 0)


BAD = [
    dict(subsets=7),                               # views not a multiple of subsets
    dict(def_elements=250, def_angles=288),        # N not a multiple of 4
    dict(span=1),                                  # no end-plane shift
    dict(maxdel=3),                                # no oblique segment group
    dict(x_pixels=100),                            # zoom != 1
    dict(rel_fov=1.0),                             # rotated samples could leave the image
    dict(rings=200),                               # z_pixels > 256
]


@pytest.mark.parametrize("bad", BAD, ids=[next(iter(b)) for b in BAD])
def test_invalid_geometry_is_rejected_without_device(bad, lib):
    from hrrt_reconstruction_amd.osem import make_params
    kw = dict(GEOMS["c1"])
    kw.update(bad)
    p = make_params(**kw)
    d = lib.Dims()
    assert lib.lib().hrrt_query_geometry(C.byref(p), C.byref(d), *([None] * 6)) == lib.HRRT_ERR_ARG
    assert lib.lib().hrrt_last_error()
    h = C.c_void_p()
    assert lib.lib().hrrt_create(C.byref(p), C.byref(h)) == lib.HRRT_ERR_ARG
    assert not h.value
    with pytest.raises(lib.HrrtError):
        lib.call("hrrt_query_geometry", C.byref(p), C.byref(d), *([None] * 6))


@pytest.mark.parametrize("geom", ["tiny", "tinyrb", "tiny9", "c1", "c2"])
def test_context_tables_pass_their_host_checks(geom, lib):
    """hrrt_create builds and checks every table on the host (forward positions against the
    lookup tables, the row-skip threshold, the forward projector's paired task list) before
    the first device allocation: without a device it must get as far as that allocation."""
    from hrrt_reconstruction_amd.osem import make_params
    h = C.c_void_p()
    rc = lib.lib().hrrt_create(C.byref(make_params(**GEOMS[geom])), C.byref(h))
    if rc == lib.HRRT_OK:          # a device is present
        lib.lib().hrrt_destroy(h)
    else:
        assert rc == lib.HRRT_ERR_NOMEM, lib.lib().hrrt_last_error()


def test_null_arguments(lib):
    L = lib.lib()
    assert L.hrrt_create(None, None) == lib.HRRT_ERR_ARG
    assert L.hrrt_get_dims(None, None) == lib.HRRT_ERR_ARG
    assert L.hrrt_query_geometry(None, None, *([None] * 6)) == lib.HRRT_ERR_ARG
    for name, (_, args) in lib.SIGNATURES.items():
        if name in ("hrrt_destroy", "hrrt_delays_destroy", "hrrt_last_error", "hrrt_create", "hrrt_query_geometry"):
            continue
        zero = [0 if a in (C.c_int, C.c_longlong) else (0.0 if a is C.c_float else None) for a in args]
        assert getattr(L, name)(*zero) == lib.HRRT_ERR_ARG, name
    L.hrrt_destroy(None)          # a no-op, like free(NULL)
    L.hrrt_delays_destroy(None)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_bench_algorithmic_counts():
    """The bench's algorithmic work per subset against the figures worked out by hand for the
    two HRRT workloads: 1.92 GFLOP (span 9, 128², 8 of 9 views used) and 45.1 GFLOP (span 3,
    256², 18 views), 46441 FOV pixels per plane at 256²; the forward kernel's bytes are the
    image's FOV columns plus 10 bytes per FOV bin of the subset."""
    sys.path.insert(0, ROOT)
    import bench
    from hrrt_reconstruction_amd.osem import Layout, make_params
    c1 = Layout(make_params(**GEOMS["c1"]))
    c2 = Layout(make_params(**GEOMS["c2"]))
    assert c2.dims.fov_voxels == 46441 and c2.S == 6367 and c2.sviews == 18
    f1 = bench.algorithmic_flops(c1, c1.sviews // 2)
    f2 = bench.algorithmic_flops(c2, c2.sviews // 2)
    assert abs(sum(f1.values()) - 1.9224e9) < 1e6
    assert abs(sum(f2.values()) - 45.1239e9) < 1e6
    b2 = bench.algorithmic_bytes(c2, 9)
    assert b2["fp_ratio"] == 46441 * 208 * 4 + 9 * 2 * 6367 * c2.dims.fov_bins * 10
