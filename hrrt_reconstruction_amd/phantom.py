"""Seeded synthetic HRRT studies in the reference's flat file layouts.

Image phantom: a uniform cylinder (radius 100 mm, activity 1) over the initial
image support z in [z_min, z_max), holding six hot spheres (radius 5-15 mm,
activity 4).  Sinograms are produced in the flat file layout
[file plane][def_angles][def_elements] that hrrt_osem3d reads
(hrrt_osem3d/file_io_processor.cpp:110-259):

* norm   N = exp(0.1*eps), eps ~ N(0,1) per bin (log-normal around 1)
* atten  A = exp(0.0096/mm * L) with L the chord of a water cylinder of
           radius 100 mm along the LOR (2*sqrt(100^2 - s^2)/cos(theta))
* trues  t = alpha * (P x) / (N*A), P the forward projector in file geometry
* randoms d uniform, 10 % of the expected counts
* scatter (optional) s = N * blur(t) scaled to 15 % of the expected counts,
           blur = Gaussian along the radial axis (sigma 25 mm)
* prompts y ~ Poisson(t + d + s/N) as int16 (asserted < 32768)

All randomness comes from one torch.Generator seeded by the caller.
"""
from __future__ import annotations

import math

import torch

WATER_MU = 0.0096   # 1/mm
CYL_R = 100.0       # mm


def file_segments(rings: int, span: int, maxdel: int):
    """Plane counts and group of each file segment, in file order (PJ:159-191, 446-457)."""
    groupmax = (maxdel - (span - 1) // 2) // span
    segs = [(2 * rings - 1, 0)]
    for g in range(1, groupmax + 1):
        zmin = (span + 1) // 2 + (g - 1) * span
        n = 2 * rings - 1 - 2 * zmin
        segs += [(n, g), (n, g)]
    return segs


def make_phantom(N, Z4, z_pixels, pixel_mm, z_mm, z_min, z_max, gen, device, activity_scale=1.0,
                 sphere_activity=None):
    """Cylinder + 6 spheres in the [X][Y][Z4] image layout.  Returns (image, spheres)."""
    c = N // 2
    r = torch.rand(6, generator=gen, device="cpu") * 10.0 + 5.0
    spheres = []
    zc_lo, zc_hi = (z_min + 2) * z_mm, (z_max - 3) * z_mm
    for i in range(6):
        while True:
            u = torch.rand(3, generator=gen, device="cpu")
            rad = float(r[i])
            rho = (CYL_R - rad - 3.0) * math.sqrt(float(u[0]))
            ang = 2 * math.pi * float(u[1])
            zc = zc_lo + rad + (zc_hi - zc_lo - 2 * rad) * float(u[2])
            spheres.append((rho * math.cos(ang), rho * math.sin(ang), zc, rad))
            break
    xs = (torch.arange(N, device=device, dtype=torch.float64) - c) * pixel_mm
    zs = torch.arange(Z4, device=device, dtype=torch.float64) * z_mm
    X = xs.view(N, 1, 1)
    Y = xs.view(1, N, 1)
    Zc = zs.view(1, 1, Z4)
    zi = torch.arange(Z4, device=device).view(1, 1, Z4)
    inz = (zi >= z_min) & (zi < z_max)
    img = ((X * X + Y * Y <= CYL_R * CYL_R) & inz).to(torch.float64)
    for k, (sx, sy, sz, sr) in enumerate(spheres):
        a = 4.0 if sphere_activity is None else float(sphere_activity[k])
        inside = ((X - sx) ** 2 + (Y - sy) ** 2 + (Zc - sz) ** 2 <= sr * sr) & inz
        img = torch.where(inside, torch.full_like(img, a), img)
    return (img * activity_scale).to(torch.float32).contiguous(), spheres


def atten_file(def_elements, def_angles, bin_mm, segs, ring_radius, ring_spacing, span, device):
    """exp(mu*L) in file layout (float32)."""
    s = (torch.arange(def_elements, device=device, dtype=torch.float64) - def_elements // 2) * bin_mm
    chord = 2.0 * torch.sqrt(torch.clamp(CYL_R * CYL_R - s * s, min=0.0))
    planes = []
    for n, g in segs:
        t = g * span * ring_spacing
        cos_t = (2 * ring_radius) / math.sqrt(t * t + (2 * ring_radius) ** 2)
        a = torch.exp(WATER_MU * chord / cos_t).to(torch.float32)
        planes.append(a.view(1, 1, def_elements).expand(n, def_angles, def_elements))
    return torch.cat(planes, 0).contiguous()


def gaussian_blur_radial(x, sigma_bins):
    """Blur along the last axis with a normalised Gaussian (zero padding), as a banded matrix product."""
    n = x.shape[-1]
    i = torch.arange(n, device=x.device, dtype=torch.float32)
    k = torch.exp(-0.5 * ((i.view(-1, 1) - i.view(1, -1)) / sigma_bins) ** 2)
    half = int(math.ceil(3 * sigma_bins))
    k = k * ((i.view(-1, 1) - i.view(1, -1)).abs() <= half)
    k = k / torch.exp(-0.5 * (torch.arange(-half, half + 1, device=x.device, dtype=torch.float32) / sigma_bins) ** 2).sum()
    flat = x.reshape(-1, n)
    out = torch.empty_like(flat)
    step = 1 << 20
    for r in range(0, flat.shape[0], step):
        out[r:r + step] = flat[r:r + step] @ k
    return out.view_as(x)


def make_study(proj_file, rings, span, maxdel, def_elements, def_angles, bin_mm, plane_sep, ring_radius,
               total_counts, gen, device, with_scatter=False, norm_sigma=0.1, randoms_frac=0.10,
               scatter_frac=0.15, norm=None, atten=None):
    """Build file-layout sinograms for an already forward-projected phantom.

    proj_file: (P x) in file layout [S][def_angles][def_elements] (float32, device).
    Returns dict(prompt=int16, delayed=f32, norm=f32, atten=f32, scatter=f32 or None, alpha=float).
    norm / atten: reuse the factors of an earlier frame of the same study (a dynamic study has one
    normalisation and one attenuation map).
    """
    segs = file_segments(rings, span, maxdel)
    ring_spacing = 2.0 * plane_sep
    shape = proj_file.shape
    if norm is None:
        norm = torch.exp(norm_sigma * torch.randn(shape, generator=gen, device=device, dtype=torch.float32))
    if atten is None:
        atten = atten_file(def_elements, def_angles, bin_mm, segs, ring_radius, ring_spacing, span, device)
    base = proj_file / (norm * atten)                 # trues per unit activity
    sfrac = scatter_frac if with_scatter else 0.0
    tfrac = 1.0 - randoms_frac - sfrac
    alpha = tfrac * total_counts / float(base.sum(dtype=torch.float64))
    trues = base * alpha
    nbins = proj_file.numel()
    d = torch.full(shape, randoms_frac * total_counts / nbins, device=device, dtype=torch.float32)
    mean = trues + d
    scatter = None
    if with_scatter:
        b = gaussian_blur_radial(trues, 25.0 / bin_mm)
        sc = b * (sfrac * total_counts / float(b.sum(dtype=torch.float64)))
        mean = mean + sc
        scatter = (sc * norm).contiguous()
        del b, sc
    prompt = torch.poisson(mean, generator=gen)
    mx = float(prompt.max())
    assert mx < 32768, f"prompt counts {mx} overflow the reference's int16 sinograms"
    prompt = prompt.to(torch.int16).contiguous()
    return dict(prompt=prompt, delayed=d.contiguous(), norm=norm.contiguous(), atten=atten, scatter=scatter,
                alpha=alpha)


def make_geometry_study(g, total_counts, gen, cpu_gen, with_scatter=False, norm=None, atten=None,
                        sphere_activity=None):
    """Seeded phantom and file-layout study for an Osem3D context g: the phantom is forward
    projected with g's own projector at the reconstruction size and spread back over the
    file's bins when the context rebins (-X 128: each internal bin covers 2 x 2 file bins).
    cpu_gen seeds the phantom (sphere positions), gen the sinogram noise and factors.
    Returns (image [X][Y][Z4], study dict of make_study)."""
    p = g.params
    rb = g.dims.x_rebin
    img, _ = make_phantom(g.N, g.Z4, g.Z, p.bin_size * rb, p.plane_sep, p.z_min, p.z_max, cpu_gen, "cuda",
                          sphere_activity=sphere_activity)
    est = g.estimate_zeros()
    proj = torch.zeros((g.S, g.views, g.N), dtype=torch.float32, device="cuda")
    for v in range(g.views // 2):
        g.forward_proj3d(img, v, prj=est)
        proj[:, v, :] = est[: g.S]
        proj[:, v + g.views // 2, :] = est[g.S:]
    if rb > 1:
        proj = proj.repeat_interleave(g.dims.v_rebin, 1).repeat_interleave(rb, 2) / (rb * g.dims.v_rebin)
    st = make_study(proj, p.rings, p.span, p.maxdel, p.def_elements, p.def_angles, p.bin_size, p.plane_sep,
                    p.crystal_radius, total_counts, gen, "cuda", with_scatter=with_scatter, norm=norm, atten=atten)
    return img, st


def estimates_to_file(estimates_by_view, S, views, N, device):
    """Assemble per-view-pair estimates [2S][N] into the file layout [S][views][N]."""
    out = torch.zeros((S, views, N), dtype=torch.float32, device=device)
    for v, e in estimates_by_view.items():
        out[:, v, :] = e[:S]
        out[:, v + views // 2, :] = e[S:]
    return out


def dynamic_frames(nframes, counts_lo, counts_hi, gen):
    """Frame durations (s), mid times (min), total counts (log-uniform in [counts_lo, counts_hi]) and
    per-sphere washout rates (1/min) of a seeded dynamic study: sphere k has activity
    1 + 3 exp(-rate_k t) at frame mid time t, the background cylinder stays at 1."""
    base = [15, 30, 60, 120, 300, 600]
    per = max(1, nframes // len(base))
    dur = [base[min(i // per, len(base) - 1)] for i in range(nframes)]
    t, mids = 0.0, []
    for d in dur:
        mids.append((t + d / 2) / 60.0)
        t += d
    u = torch.rand(nframes, generator=gen)
    counts = [float(math.exp(math.log(counts_lo) + (math.log(counts_hi) - math.log(counts_lo)) * float(v))) for v in u]
    r = torch.rand(6, generator=gen)
    rates = [float(math.exp(math.log(0.005) + (math.log(0.2) - math.log(0.005)) * float(v))) for v in r]
    return dur, mids, counts, rates


# ---------------------------------------------------------------- list mode

def lm_event_words(mp, ax, ay, bx, by, doia, doib, prompt):
    """64-bit HRRT list-mode events as (ew1, ew2) uint32 words (numpy arrays): the fields
    rebin_packet decodes (lmhistogram_mp/histogram_mp.cpp:1120-1128); the top bits make the
    pair an event (EWTYPES, gen_delays_lib/geometry_info.h:75): ew2 top bits 11 = prompt,
    10 = delayed."""
    import numpy as np
    u = lambda v: np.asarray(v, dtype=np.uint64)
    ew1 = (u(doia) << 22) | ((u(mp) & 7) << 16) | (u(ay) << 8) | u(ax)
    ew2 = np.where(np.asarray(prompt, bool), np.uint64(0xC0000000), np.uint64(0x80000000))
    ew2 = ew2 | (u(doib) << 22) | (((u(mp) >> 3) & 7) << 16) | (u(by) << 8) | u(bx)
    return ew1.astype(np.uint32), ew2.astype(np.uint32)


def lm_tag_words(tag):
    """A 32-bit tag word as an (ew1, ew2) pair: low half in ew1 (top bits 01), high half in ew2
    (top bits 10)."""
    import numpy as np
    tag = np.asarray(tag, dtype=np.uint64)
    return ((np.uint64(0x40000000) | (tag & np.uint64(0xffff))).astype(np.uint32),
            (np.uint64(0x80000000) | (tag >> np.uint64(16))).astype(np.uint32))


def lm_synthetic_stream(nevents, seed, events_per_ms=2000, t0_ms=1000, delayed_fraction=0.1, stray_every=0,
                        bad_fraction=0.0, tx_fraction=0.0, odd_doi_fraction=0.0, singles_every=0):
    """A seeded synthetic list-mode stream (numpy uint32 words): random emission events of the
    20 module pairs (rings within 40 of each other), a time tag (ms) every events_per_ms
    events from t0_ms, optional block-singles tags, stray out-of-sync words, rejected events
    (bad module pair or crystal), transmission (DOI 7) and DOI 2-6 events."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n = int(nevents)
    mp = rng.integers(1, 21, n)
    ax, bx = rng.integers(0, 72, n), rng.integers(0, 72, n)
    ay = rng.integers(0, 104, n)
    by = np.clip(ay + rng.integers(-40, 41, n), 0, 103)
    doia, doib = rng.integers(0, 2, n), rng.integers(0, 2, n)
    prompt = rng.random(n) >= delayed_fraction
    if bad_fraction > 0:
        bad = rng.random(n) < bad_fraction
        kind = rng.integers(0, 3, n)
        mp = np.where(bad & (kind == 0), rng.choice([0, 21, 40, 63], n), mp)
        ax = np.where(bad & (kind == 1), rng.integers(72, 256, n), ax)
        by = np.where(bad & (kind == 2), rng.integers(104, 256, n), by)
    if tx_fraction > 0:
        doia = np.where(rng.random(n) < tx_fraction, 7, doia)
    if odd_doi_fraction > 0:
        doib = np.where(rng.random(n) < odd_doi_fraction, rng.integers(2, 7, n), doib)
    ew1, ew2 = lm_event_words(mp, ax, ay, bx, by, doia, doib, prompt)
    pairs = np.stack([ew1, ew2], axis=1)
    # tags: time every events_per_ms events, singles in between
    chunks = []
    t = t0_ms
    for i0 in range(0, n, events_per_ms):
        tw = np.stack(lm_tag_words(np.array([0x80000000 | t])), axis=1)
        chunks.append(tw)
        if singles_every:
            nb = max(1, events_per_ms // singles_every)
            blocks = rng.integers(0, 936, nb)
            vals = rng.integers(0, 1 << 19, nb)
            chunks.append(np.stack(lm_tag_words((0xA0000000 | (blocks << 19) | vals).astype(np.uint64)), axis=1))
        chunks.append(pairs[i0:i0 + events_per_ms])
        t += 1
    words = np.concatenate(chunks, axis=0).reshape(-1)
    if stray_every:
        pos = np.sort(rng.choice(words.size, max(1, words.size // stray_every), replace=False))
        stray = (0x80000000 | rng.integers(0, 1 << 30, pos.size)).astype(np.uint32)
        words = np.insert(words, pos, stray)
    return words.astype(np.uint32)
