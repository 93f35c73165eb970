"""Host side of the smoothed randoms library without a GPU: its geometry tables, LUT and singles
iteration against the oracle bit for bit, the LUT file round trip, and argument errors."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def GenDelays():
    from hrrt_reconstruction_amd.randoms import GenDelays
    return GenDelays


@pytest.mark.parametrize("span,lr_type", [(9, 0), (3, 0), (9, 1), (3, 2)])
def test_lut_bitexact(GenDelays, span, lr_type):
    g = GenDelays(span=span, lr_type=lr_type)
    D = O.Delays(span=span, lr_type=lr_type)
    assert (g.nplanes, g.nviews, g.nprojs, g.nsegs) == (D.nplanes, D.nviews, D.nprojs, D.nsegs)
    for mp in range(1, 21):
        for a, b in zip(g.lut(mp), D.sol(mp)):
            assert np.array_equal(a.view(np.int32), b.view(np.int32)), mp


def test_pairs_full_counts_the_ring_window(GenDelays):
    g = GenDelays()
    ring_pairs = sum(min(103, y + 72) - max(0, y - 72) + 1 for y in range(104))
    lors = sum(int((g.lut(mp)[0] >= 0).sum()) for mp in range(1, 21))
    assert g.pairs_full == lors * ring_pairs          # a LOR fixes both crystals and layers


@pytest.mark.parametrize("seed,duration", [(1, 300.0), (2, 1.0), (3, 3600.0)])
def test_singles_bitexact(GenDelays, seed, duration):
    rng = np.random.default_rng(seed)
    coins = rng.integers(0, 40000, 2 * O.NCRYS).astype(np.int32)
    coins[O.NCRYS + rng.integers(0, O.NCRYS, 50)] = 0          # silent crystals
    g = GenDelays()
    s, it = g.singles(coins, duration)
    ref, it_ref = O.Delays.singles(coins[O.NCRYS:], 6e-9, duration)
    assert it == it_ref
    assert np.array_equal(s.view(np.int32), ref.view(np.int32))


def test_lut_file_round_trip(GenDelays, tmp_path):
    g = GenDelays(span=3)
    path = tmp_path / "hrrt_rebinner.lut"
    g.save_lut(str(path))
    assert path.stat().st_size == 20 * 144 * 144 * 12 + 2 * 104 * 4
    g2 = GenDelays(span=3, lut_file=str(path))
    for mp in (1, 7, 20):
        for a, b in zip(g.lut(mp), g2.lut(mp)):
            assert np.array_equal(a, b)
    raw = np.fromfile(path, np.uint8)
    (tmp_path / "short.lut").write_bytes(raw[:1000].tobytes())
    from hrrt_reconstruction_amd import _lib
    with pytest.raises(_lib.HrrtError, match="short LUT"):
        GenDelays(lut_file=str(tmp_path / "short.lut"))
    with pytest.raises(_lib.HrrtError, match="cannot open"):
        GenDelays(lut_file=str(tmp_path / "missing.lut"))


def test_bad_arguments(GenDelays):
    from hrrt_reconstruction_amd import _lib
    for kw in (dict(span=5), dict(maxrd=-1), dict(maxrd=104), dict(lr_type=3), dict(span=9, maxrd=80)):
        with pytest.raises(_lib.HrrtError):
            GenDelays(**kw)
    g = GenDelays()
    with pytest.raises(ValueError):
        g.singles(np.zeros(10, np.int32), 1.0)
    with pytest.raises(_lib.HrrtError):
        g.lut(0)
