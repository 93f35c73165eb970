"""Helpers shared by the GPU parity tests (device setup, seeded inputs)."""
import numpy as np
import torch


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def random_image(o, seed, lo=0.5, hi=2.0):
    """Seeded image that is non-zero on the FOV rows for z in [z_min, z_max) (like Init_Image)."""
    rng = np.random.default_rng(seed)
    img = o.init_image()
    img *= rng.uniform(lo, hi, size=img.shape).astype(np.float32)
    return np.ascontiguousarray(img)


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def from_dev(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()
