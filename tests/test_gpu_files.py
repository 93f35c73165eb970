"""Image and sensitivity files through the GPU path: the flat-plane transposes are bit-exact
against the oracle, the written files hold exactly those planes, the Interfile header has
the reference's lines (write_image_header.cpp:188-258, carried-forward sinogram header
lines filtered as at :266-299), and reading inverts writing."""
import numpy as np
import pytest

from conftest import GEOMS
from gpu_util import from_dev, random_image, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    torch.cuda.set_device(0)
    return torch


def _pair(geom, oracle_lib):
    from hrrt_reconstruction_amd.osem import Osem3D
    return oracle_lib.Oracle(**GEOMS[geom]), Osem3D(**GEOMS[geom])


@pytest.mark.parametrize("geom", ["tiny", "c2"])
def test_flat_transposes_bitexact(torch_dev, oracle_lib, geom):
    print(f"flat image geom={geom}", flush=True)
    o, g = _pair(geom, oracle_lib)
    img = random_image(o, seed=31)
    rng = np.random.default_rng(32)
    flat_in = rng.uniform(0, 3, (o.Z, o.N, o.N)).astype(np.float32)
    for flip in (False, True):
        assert np.array_equal(from_dev(g.image_to_flat(to_dev(img), flip)), o.image_to_flat(img, flip)), flip
        assert np.array_equal(from_dev(g.image_from_flat(to_dev(flat_in), flip)), o.image_from_flat(flat_in, flip))


def test_write_and_read_image_files(torch_dev, oracle_lib, tmp_path):
    o, g = _pair("tiny", oracle_lib)
    img = random_image(o, seed=33)
    prompt = tmp_path / "scan_prompts.s"
    prompt.write_bytes(b"")
    (tmp_path / "scan_prompts.s.hdr").write_text(
        "!INTERFILE\nname of data file := scan_prompts.s\nnumber format := signed integer\n"
        "patient name := phantom\nframe := 0\nmatrix size [1] := 256\nisotope name := F-18\n")
    scatter = tmp_path / "scan_sc.s"
    (tmp_path / "scan_sc.h33").write_text("!INTERFILE\nscatter fraction := 0.1525\n")
    path = tmp_path / "recon.i"
    g.write_image(to_dev(img), str(path), iterations=2, sw_version="1.2", sw_build_id="b42",
                  prompt_file=str(prompt), norm_file="/data/norm.n", scatter_file=str(scatter))
    data = np.fromfile(path, np.float32)
    assert np.array_equal(data, o.image_to_flat(img).reshape(-1))
    hdr = (tmp_path / "recon.i.hdr").read_text()
    expected = (
        "!INTERFILE\n"
        "name of data file := recon.i\n"
        "data format := image\n"
        "number format := float\n"
        "number of bytes per pixel := 4\n"
        "number of dimensions := 3\n"
        f"matrix size [1] := {o.N}\n"
        f"matrix size [2] := {o.N}\n"
        f"matrix size [3] := {o.Z}\n"
        "scaling factor (mm/pixel) [1] := 1.218750\n"
        "scaling factor (mm/pixel) [2] := 1.218750\n"
        "scaling factor (mm/pixel) [3] := 1.218750\n"
        "reconstruction method := inki-2006-may-OSEM3D-OP\n"
        "reconstruction version := 1.2\n"
        "reconstruction build ID := b42\n"
        "prompt file used := scan_prompts.s\n"
        "norm file used := norm.n\n"
        "scatter file used := scan_sc.s\n"
        "scatter fraction := 0.1525\n"
        "number of iterations := 2\n"
        f"number of subsets := {o.subsets}\n"
        "patient name := phantom\n"
        "frame := 0\n"
        "isotope name := F-18\n")
    assert hdr == expected
    back = from_dev(g.read_image(str(path)))
    assert np.array_equal(back, o.image_from_flat(o.image_to_flat(img)))
    g.set_psf()
    g.write_image(to_dev(img), str(path), flip=True)
    assert "inki-2006-may-OSEM3D-OP-PSF" in (tmp_path / "recon.i.hdr").read_text()
    assert np.array_equal(np.fromfile(path, np.float32), o.image_to_flat(img, True).reshape(-1))


def test_normfac_file_round_trip(torch_dev, oracle_lib, tmp_path):
    o, g = _pair("tiny", oracle_lib)
    rng = np.random.default_rng(34)
    nf = rng.uniform(0, 2, (o.subsets, o.N, o.N, o.Z4)).astype(np.float32)
    path = tmp_path / "normfac.i"
    g.write_normfac(to_dev(nf), str(path))
    assert np.array_equal(np.fromfile(path, np.float32).reshape(nf.shape), nf)
    assert np.array_equal(from_dev(g.read_normfac(str(path))), nf)
