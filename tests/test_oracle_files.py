"""Pins the oracle's flat-image restatements (SaveImage S3:4023-4035, Init_Image -I
S3:3970-3993) with known answers: a voxel lands at the plane offset z*N*N + y*N + x of
the file (x and y swapped when flipped), reading inverts writing, and only FOV voxels
are carried."""
import numpy as np

from conftest import GEOMS


def test_voxel_positions_and_round_trip(oracle_lib):
    o = oracle_lib.Oracle(**GEOMS["tiny"])
    x, y, z = 40, 57, 7
    assert o.cylwiny[x, 0] <= y < o.cylwiny[x, 1]
    img = o.image_zeros()
    img[x, y, z] = 5.0
    flat = o.image_to_flat(img)
    assert flat.shape == (o.Z, o.N, o.N)
    assert flat.reshape(-1)[z * o.N * o.N + y * o.N + x] == 5.0 and np.count_nonzero(flat) == 1
    flat_f = o.image_to_flat(img, flip=True)
    assert flat_f[z, x, y] == 5.0 and np.count_nonzero(flat_f) == 1
    for flip in (False, True):
        back = o.image_from_flat(o.image_to_flat(img, flip), flip)
        assert np.array_equal(back, img)


def test_only_fov_is_carried(oracle_lib):
    o = oracle_lib.Oracle(**GEOMS["tiny"])
    rng = np.random.default_rng(4)
    flat = rng.uniform(1, 2, (o.Z, o.N, o.N)).astype(np.float32)
    img = o.image_from_flat(flat)
    fov = np.zeros((o.N, o.N), bool)
    for x in range(o.N):
        fov[x, o.cylwiny[x, 0]:o.cylwiny[x, 1]] = True
    assert np.all(img[~fov] == 0) and np.all(img[:, :, o.Z:] == 0)
    # image (X, Y) holds the file's column X of row Y
    assert np.array_equal(img[fov][:, :o.Z], np.transpose(flat, (2, 1, 0))[fov])
    out = o.image_to_flat(img)
    assert np.array_equal(out, np.where(fov.T[None], flat, 0))
