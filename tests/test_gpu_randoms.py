"""Smoothed randoms (gen_delays) on the GPU against the oracle.

The kernel computes every crystal-pair term with the reference's float operations (so each
term equals the oracle's bit for bit) but sums the terms of one sinogram element exactly, in
64-bit fixed point, and rounds once to float; the oracle adds in float in the reference's
loop order.  The fixed-point unit of a sinogram bin is at most M nlor 2^-50 (M = the
largest term the bin's nlor LORs can produce), so an element of n terms is within
n M nlor 2^-51 of the exact sum; the oracle's float sum is within (n - 1) 2^-24 of it
(positive terms).  The tolerance of 2e-5 relative per element covers both for n up to 300
and terms down to 1e-7 M.  Element totals are compared to 1e-6, the kernel's result must be
bit-identical from run to run, and every element the oracle leaves at zero must be zero.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 2e-5


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


def _singles(seed, keep):
    rng = np.random.default_rng(seed)
    s = rng.uniform(20.0, 200.0, O.NCRYS).astype(np.float32)
    s[rng.random(O.NCRYS) >= keep] = 0.0
    return s


def _compare(out, ref):
    assert out.shape == ref.shape
    total = 0.0
    for p0 in range(0, ref.shape[0], 256):     # plane blocks, to bound host memory
        a = out[p0:p0 + 256].astype(np.float64)
        b = ref[p0:p0 + 256].astype(np.float64)
        assert np.array_equal(a == 0, b == 0), f"zero pattern differs in planes {p0}.."
        err = np.abs(a - b) - RTOL * np.abs(b)
        assert err.max() <= 0, f"planes {p0}..: max rel {np.max(np.abs(a - b) / np.maximum(b, 1e-30))}"
        total += b.sum()
    return total


@pytest.mark.parametrize("span,lr_type,keep,seed", [(9, 0, 0.05, 1), (3, 0, 0.02, 2), (9, 2, 0.25, 3),
                                                    (3, 1, 0.1, 4)])
def test_smooth_matches_oracle(torch_dev, span, lr_type, keep, seed):
    from hrrt_reconstruction_amd.randoms import GenDelays
    print(f"randoms smooth span={span} lr={lr_type} keep={keep}", flush=True)
    s = _singles(seed, keep)
    g = GenDelays(span=span, lr_type=lr_type)
    D = O.Delays(span=span, lr_type=lr_type)
    tau, T = 6e-9, 600.0
    out = g.smooth(torch_dev.from_numpy(s).cuda(), T).cpu().numpy()
    ref, pairs = D.smooth(s, tau, T)
    assert pairs > 0
    total = _compare(out, ref)
    assert abs(out.astype(np.float64).sum() - total) <= 1e-6 * total


def test_smooth_dense_half(torch_dev):
    """Every second A crystal live (the oracle visits half of the 2.8 G crystal pairs)."""
    from hrrt_reconstruction_amd.randoms import GenDelays
    print("randoms smooth dense span=9", flush=True)
    rng = np.random.default_rng(5)
    s = rng.uniform(50.0, 150.0, O.NCRYS).astype(np.float32)
    s[1::2] = 0.0
    g = GenDelays(span=9)
    out = g.smooth(torch_dev.from_numpy(s).cuda(), 1200.0).cpu().numpy()
    ref, pairs = O.Delays(span=9).smooth(s, 6e-9, 1200.0)
    assert pairs > 1e9
    _compare(out, ref)


def test_gen_delays_from_histogram_and_file(torch_dev, tmp_path):
    from hrrt_reconstruction_amd.randoms import GenDelays
    print("randoms gen_delays from .ch", flush=True)
    rng = np.random.default_rng(6)
    coins = rng.integers(0, 3000, 2 * O.NCRYS).astype(np.int32)
    coins[O.NCRYS:][rng.random(O.NCRYS) > 0.03] = 0             # delayed counts on 3% of the crystals
    g = GenDelays(span=9)
    out, sg, it = g.gen_delays(coins, 900.0)
    ref_s, ref_it = O.Delays.singles(coins[O.NCRYS:], 6e-9, 900.0)
    assert it == ref_it and np.array_equal(sg.view(np.int32), ref_s.view(np.int32))
    ref, _ = O.Delays(span=9).smooth(ref_s, 6e-9, 900.0)
    host = out.cpu().numpy()
    _compare(host, ref)
    path = tmp_path / "scan_ra_smo.s"
    g.write(out, str(path))
    assert np.array_equal(np.fromfile(path, np.float32), host.reshape(-1))


def test_repeat_calls_zero_input_and_lut_file(torch_dev, tmp_path):
    from hrrt_reconstruction_amd.randoms import GenDelays
    print("randoms repeat / zero / LUT file", flush=True)
    g = GenDelays(span=9, lr_type=2)
    zero = g.smooth(torch_dev.zeros(O.NCRYS, device="cuda"))
    assert not zero.any().item()
    s = _singles(7, 0.1)
    a = g.smooth(torch_dev.from_numpy(s).cuda(), 60.0)
    b = g.smooth(torch_dev.from_numpy(s).cuda(), 60.0)          # the column scratch is rewritten, not summed
    assert torch_dev.equal(a, b)                                 # exact sums: reproducible bit for bit
    path = tmp_path / "lut"
    g.save_lut(str(path))
    c = GenDelays(span=9, lr_type=2, lut_file=str(path)).smooth(torch_dev.from_numpy(s).cuda(), 60.0)
    ref, _ = O.Delays(span=9, lr_type=2).smooth(s, 6e-9, 60.0)
    for x in (a, b, c):
        _compare(x.cpu().numpy(), ref)
