"""Batched dynamic frames (hrrt_osem_subset_frames / hrrt_osem_iteration_frames) on the GPU.

Each frame of a batch must equal, bit for bit, the single-study subset update on that
frame, which is itself bit-exact against the oracle (test_gpu_parity.py); one frame is
also checked against the oracle directly.  The graph-replayed iteration is replayed
several times with eager work in between, and every replay is checked.
"""
import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import from_dev, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


def _frames(g, o, nframes, seed):
    rng = np.random.default_rng(seed)
    shape = (o.S, o.views, o.N)
    prompts, adds = [], []
    for f in range(nframes):
        lam = 0.5 + 3.0 * f
        prompts.append(rng.poisson(lam, shape).astype(np.int16))
        adds.append(rng.uniform(0.0, 0.2 * lam, shape).astype(np.float32))
    nf = rng.uniform(0.05, 2.0, (o.subsets, o.N, o.N, o.Z4)).astype(np.float32)
    return prompts, adds, nf


@pytest.mark.parametrize("geom,nframes,subsets", [("tiny", 3, (0, 8, 1)), ("c1", 2, (0, 8))])
def test_frames_subset_matches_single_frames(torch_dev, oracle_lib, geom, nframes, subsets):
    from hrrt_reconstruction_amd.osem import Osem3D
    print(f"frames geom={geom} nframes={nframes} subsets={subsets}", flush=True)
    o = oracle_lib.Oracle(**GEOMS[geom])
    g = Osem3D(**GEOMS[geom])
    prompts, adds, nf = _frames(g, o, nframes, 41)
    d_nf = to_dev(nf)
    img0 = o.init_image()
    d_imgs = to_dev(np.stack([img0] * nframes))
    d_pr = to_dev(np.stack([g.sino_to_paircol(p) for p in prompts]))
    d_add = to_dev(np.stack([g.sino_to_paircol(a) for a in adds]))
    singles = [to_dev(img0) for _ in range(nframes)]
    ref0 = img0.copy()
    corr = o.image_zeros()
    for s in subsets:
        g.osem_subset_frames(d_imgs, d_nf, d_pr, d_add, s)
        for f in range(nframes):
            g.osem_subset(singles[f], d_nf, d_pr[f], d_add[f], s)
        o.osem_subset(ref0, corr, np.ascontiguousarray(nf[s]), adds[0], prompts[0], s,
                      nthreads=g.backproject_schedule, workers=8)
        got = from_dev(d_imgs)
        for f in range(nframes):
            assert np.array_equal(got[f], from_dev(singles[f])), f"subset {s} frame {f}"
        assert np.array_equal(got[0], ref0), f"subset {s}: frame 0 differs from the oracle"


def test_frames_iteration_graph_replays(torch_dev, oracle_lib):
    from hrrt_reconstruction_amd.osem import Osem3D
    torch = torch_dev
    print("frames graph replay geom=tiny nframes=2", flush=True)
    kw = dict(GEOMS["tiny"], subsets=4)
    o = oracle_lib.Oracle(**kw)
    g = Osem3D(**kw)
    nframes = 2
    prompts, adds, nf = _frames(g, o, nframes, 43)
    d_nf = to_dev(nf)
    img0 = o.init_image()
    d_pr = to_dev(np.stack([g.sino_to_paircol(p) for p in prompts]))
    d_add = to_dev(np.stack([g.sino_to_paircol(a) for a in adds]))
    d_imgs = to_dev(np.stack([img0] * nframes))
    eager = [to_dev(img0) for _ in range(nframes)]
    stream = torch.cuda.Stream()
    for it in range(3):
        stream.wait_stream(torch.cuda.current_stream())   # inputs written on the current stream
        with torch.cuda.stream(stream):
            g.osem_iteration_frames(d_imgs, d_nf, d_pr, d_add)
        stream.synchronize()
        # eager reference, with unrelated eager work on the default stream in between
        scratch = torch.zeros(1 << 20, device="cuda")
        scratch += 1.0
        for f in range(nframes):
            g.calculate_osem3d(eager[f], d_nf, d_pr[f], d_add[f], iterations=1)
        torch.cuda.synchronize()
        got = from_dev(d_imgs)
        for f in range(nframes):
            assert np.array_equal(got[f], from_dev(eager[f])), f"replay {it} frame {f}"


def test_frames_iteration_refuses_default_stream(torch_dev, oracle_lib):
    from hrrt_reconstruction_amd.osem import Osem3D
    from hrrt_reconstruction_amd import _lib
    import ctypes as C
    g = Osem3D(**GEOMS["tiny"])
    o = oracle_lib.Oracle(**GEOMS["tiny"])
    d = to_dev(np.stack([o.init_image()]))
    nf = to_dev(np.ones((o.subsets, o.N, o.N, o.Z4), np.float32))
    pr = torch_dev.zeros((1, o.views // 2, 2, o.N, g.spp), dtype=torch_dev.int16, device="cuda")
    ad = torch_dev.zeros((1, o.views // 2, 2, o.N, g.spp), dtype=torch_dev.float32, device="cuda")
    rc = _lib.lib().hrrt_osem_iteration_frames(g.ctx, 1, C.c_void_p(d.data_ptr()), C.c_void_p(nf.data_ptr()),
                                               C.c_void_p(pr.data_ptr()), C.c_void_p(ad.data_ptr()), None)
    assert rc == _lib.HRRT_ERR_ARG


def test_frames_every_autotuner_shape(torch_dev, oracle_lib):
    """The batched path under every forward launch shape the autotuner may pick equals the
    default shape bit for bit, FP32 and FP16-input forms; tune_frames leaves one of them set,
    and an iteration replayed after the change matches one run eagerly subset by subset."""
    from hrrt_reconstruction_amd.autotune import CANDIDATES, tune_frames
    from hrrt_reconstruction_amd.osem import Osem3D
    torch = torch_dev
    print("frames autotuner shapes", flush=True)
    o = oracle_lib.Oracle(**GEOMS["tiny"])
    g = Osem3D(**GEOMS["tiny"])
    prompts, adds, nf = _frames(g, o, 3, 43)
    d_nf = to_dev(nf)
    img0 = o.init_image()
    d_pr = to_dev(np.stack([g.sino_to_paircol(p) for p in prompts]))
    d_add = to_dev(np.stack([g.sino_to_paircol(a) for a in adds]))
    d_add16 = g.to_half(d_add)

    def run(shape, mixed):
        if shape is not None:
            g.set_launch_shape(*shape)
        d = to_dev(np.stack([img0] * 3))
        for s in (0, 8):
            if mixed:
                g.osem_subset_frames_mixed(d, d_nf, d_pr, d_add16, s)
            else:
                g.osem_subset_frames(d, d_nf, d_pr, d_add, s)
        return from_dev(d)

    for mixed in (False, True):
        ref = run((4, 0), mixed)
        for shape in CANDIDATES:
            assert np.array_equal(run(shape, mixed), ref), (shape, mixed)
    best, times = tune_frames(g, to_dev(np.stack([img0] * 3)), d_nf, d_pr, d_add, subsets=(0,), repeats=1)
    assert tuple(best) in CANDIDATES and len(times) == len(CANDIDATES)
    stream = torch.cuda.Stream()
    d = to_dev(np.stack([img0] * 3))
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        g.osem_iteration_frames(d, d_nf, d_pr, d_add)
    stream.synchronize()
    got = from_dev(d)
    for f in range(3):
        e = to_dev(img0)
        g.calculate_osem3d(e, d_nf, d_pr[f], d_add[f], iterations=1)
        assert np.array_equal(got[f], from_dev(e)), f"frame {f} after the tuned shape"
