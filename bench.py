#!/usr/bin/env python3
"""OSEM subset-iterations/s on one MI355X: HRRT span 3, 256x256x207, 16 subsets.

Workload (synthetic, seeded): the HRRT geometry (104 rings, 256 bins x 288
views, span 3, maximum ring difference 67 -> 6367 planes), a 256x256x207
image at 1.21875 mm holding a cylinder (r = 100 mm) with six 4:1 hot spheres,
prompts Poisson-sampled to 200M counts from the forward projection, log-normal
norm, water-cylinder attenuation, 10 % uniform randoms and 15 % Gaussian
scatter.  The files are staged on the device and the sensitivity images are
built before the timed region.

A step is one OSEM subset update (forward projection of the subset's 9 view
pairs, OP ratio, back projection, rotation and image update), the reference's
CalculateOsem3d inner body.  Steps cycle through the reference's subset order.
"""
from __future__ import annotations

import argparse
import math
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
_T0 = time.perf_counter()


def log(msg):
    """Progress of the setup phases on stderr (the JSON result line goes to stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)
GEOM = dict(x_pixels=256, span=3, maxdel=67)


def algorithmic_bytes(g, npairs):
    """Bytes each kernel must move per launch (one subset), independent of the implementation.

    Only FOV bins and voxels are read or written (the reference loops over cylwiny).
    """
    d = g.dims
    bins = npairs * 2 * g.S * d.fov_bins        # sinogram bins of the subset that hold a ray
    vol = d.fov_voxels * g.Z4 * 4                # one image volume (FOV columns) in bytes
    return {
        # image read once; int16 prompt + float additive term read, float ratio written per bin
        "fp_ratio": vol + bins * (2 + 4 + 4),
        # ratio read per bin; the two rotated-frame halves of every view pair written
        "backproject": bins * 4 + 2 * npairs * vol,
        # both halves of every pair read; image and 1/sensitivity read, image written
        "rotate_update": 2 * npairs * vol + 3 * vol,
    }


def algorithmic_flops(g, npairs):
    """FP32 operations each kernel must do per launch (one subset), independent of the
    implementation: an axial interpolation and accumulation (4 operations) per sinogram plane
    and FOV pixel of every view, a bilinear rotation (7) per FOV voxel of every view, the ratio
    (3) per bin and the update (4) per voxel."""
    d = g.dims
    views = 2 * npairs
    lerp = views * g.S * d.fov_voxels * 4
    rotate = views * d.fov_voxels * g.Z * 7
    bins = npairs * 2 * g.S * d.fov_bins
    return {
        "fp_ratio": lerp + rotate + 3 * bins,      # the image rotated into each view, then the rays
        "backproject": lerp,
        "rotate_update": rotate + 4 * d.fov_voxels * g.Z,
    }


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_c2.txt")
KERNEL_SOURCES = [os.path.join(ROOT, "hrrt_reconstruction_amd", "csrc", f) for f in ("kernels.hip", "kernels.h")]


def kernel_source_hash():
    """sha256 of the kernel sources the traffic figures were measured on."""
    import hashlib
    h = hashlib.sha256()
    for path in KERNEL_SOURCES:
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def read_traffic(kernel):
    """Measured HBM bytes per launch (fetched + written) from the committed rocprofv3 PMC summary,
    or (None, reason) when it is missing or was measured on other kernel sources."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, "no traffic summary"
    src, val = None, None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 2 and parts[0] == "source-sha256":
                src = parts[1]
            elif len(parts) >= 3 and parts[0] == kernel:
                try:
                    val = float(parts[1]) + float(parts[2])
                except ValueError:
                    val = None
    if src != kernel_source_hash():
        return None, "stale: the traffic summary was measured on other kernel sources"
    return val, None if val is not None else f"no traffic figure for {kernel}"


def write_traffic(fetch_dir, write_dir, launches):
    """Writes profiles/traffic_c2.txt from two rocprofv3 --pmc runs of this bench (FETCH_SIZE,
    then WRITE_SIZE, in KiB per dispatch): the mean of the last `launches` dispatches of each
    subset kernel, scaled by what the same counters read for the bench's 256 MiB device copy,
    with the sha256 of the kernel sources measured."""
    import csv
    import glob

    def per_dispatch(d, counter):
        rows = []
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r["Counter_Name"] == counter:
                        rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
        rows.sort()
        return rows

    kinds = {"fp_ratio": "fp_kernel<4, 1, 1, true>", "backproject": "bp_inc_kernel",
             "rotate_update": "rot_update_kernel<false>"}
    cal_bytes = 256 << 20
    out = {}
    scale = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = per_dispatch(d, counter)
        copies = [v for _, name, v in rows if "copy" in name.lower()]
        if not copies:
            raise SystemExit(f"no copy dispatch in {d}")
        # the calibration copy is the run's first copy dispatch (the bench makes it before
        # anything else); its counter reads cal_bytes / k KiB
        k = cal_bytes / 1024 / copies[0]
        scale[counter] = round(k, 3)
        for kind, pat in kinds.items():
            vals = [v for _, name, v in rows if pat in name][-launches:]
            if not vals:
                raise SystemExit(f"no {pat} dispatch in {d}")
            out.setdefault(kind, []).append(sum(vals) / len(vals) * 1024 * k)
    with open(TRAFFIC_FILE, "w") as f:
        f.write("# HBM bytes per launch (fetched, written) of the OSEM subset kernels, HRRT span 3 256x256x207,\n")
        f.write(f"# from rocprofv3 FETCH_SIZE x {scale['FETCH_SIZE']} / WRITE_SIZE x {scale['WRITE_SIZE']} KiB "
                "(the factors a 256 MiB device copy gives), mean of the last "
                f"{launches} launches; written by bench.py --write-traffic\n")
        f.write(f"source-sha256 {kernel_source_hash()}\n")
        for kind, (fe, wr) in out.items():
            f.write(f"{kind} {int(round(fe))} {int(round(wr))}\n")
    print(open(TRAFFIC_FILE).read())


def build_study(g, torch, seed):
    from hrrt_reconstruction_amd import phantom as ph
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    cpu_gen = torch.Generator()
    cpu_gen.manual_seed(seed)
    p = g.params
    img, spheres = ph.make_phantom(g.N, g.Z4, g.Z, p.bin_size, p.plane_sep, 6, 200, cpu_gen, "cuda")
    est = g.estimate_zeros()
    proj = torch.zeros((g.S, g.views, g.N), dtype=torch.float32, device="cuda")
    for v in range(g.views // 2):
        g.forward_proj3d(img, v, prj=est)
        proj[:, v, :] = est[: g.S]
        proj[:, v + g.views // 2, :] = est[g.S:]
        if v % 48 == 47:
            torch.cuda.synchronize()
            log(f"phantom forward projection: {v + 1}/{g.views // 2} view pairs")
    log("sinogram factors ...")
    study = ph.make_study(proj, p.rings, p.span, p.maxdel, p.def_elements, p.def_angles, p.bin_size,
                          p.plane_sep, p.crystal_radius, 200e6, gen, "cuda", with_scatter=True)
    del proj, est
    torch.cuda.synchronize()
    log("study sinograms generated")
    return img, study


def paircol_to_sino(g, torch, pc):
    planes, pidx = g.paircol_index()
    V = g.views
    sel = pc[:, :, :, torch.as_tensor(pidx, device=pc.device)]          # [V/2][2][N][S]
    return sel.permute(3, 1, 0, 2).reshape(g.S, V, g.N).contiguous()


def cpu_baseline(g, torch, prompt, add, normfac, image, nsub, threads, schedule=None):
    """The oracle's restatement of the reference OSEM subset on `threads` host cores, with the
    reference's thread schedule for that many threads (or the given one)."""
    schedule = threads if schedule is None else schedule
    import numpy as np
    from oracle import oracle
    oracle.build()
    o = oracle.Oracle(**GEOM)
    pr = paircol_to_sino(g, torch, prompt).cpu().numpy()
    ad = paircol_to_sino(g, torch, add).cpu().numpy()
    log("cpu baseline: inputs on the host")
    img = image.cpu().numpy().copy()
    corr = o.image_zeros()
    order = [int(s) for s in g.order]
    # warm-up subset, then the timed ones
    o.osem_subset(img, corr, np.ascontiguousarray(normfac[order[0]].cpu().numpy()), ad, pr, order[0],
                  nthreads=schedule, workers=threads)
    log("cpu baseline: warm-up subset done")
    nfs = [np.ascontiguousarray(normfac[order[1 + i]].cpu().numpy()) for i in range(nsub)]
    t0 = time.perf_counter()
    for i in range(nsub):
        o.osem_subset(img, corr, nfs[i], ad, pr, order[1 + i], nthreads=schedule, workers=threads)
        log(f"cpu baseline: subset {i + 1}/{nsub} done")
    dt = time.perf_counter() - t0
    return nsub / dt, dt


DYN_GEOM = dict(x_pixels=128, span=9, maxdel=67)


def cpu_model():
    """The host CPU's model name (as lscpu reports it), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def c1_leg(torch, seed, iterations, cpu_threads, run_cpu):
    """The reference's CPU-runnable case: span 9, 128x128x207 (the -X 128 rebinning of the
    256x288 sinogram), 16 subsets x 1 iteration, OP weighting, one study of 200M prompts with
    randoms 10 % and scatter 15 %.  GPU: whole iterations (16 subset updates and the
    end-of-iteration threshold) from the initial image, timed; CPU: one whole iteration of the
    oracle with the reference's thread schedule for cpu_threads threads, and the GPU iteration
    under that same schedule compared with it bit for bit."""
    import numpy as np
    from hrrt_reconstruction_amd import phantom as ph
    from hrrt_reconstruction_amd.osem import Osem3D
    g = Osem3D("cuda:0", **DYN_GEOM)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    pg = torch.Generator()
    pg.manual_seed(seed + 1)
    _, st = ph.make_geometry_study(g, 200e6, gen, pg, with_scatter=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prompt, add, invnorm = g.stage(st["prompt"], st["delayed"], st["norm"], st["scatter"], st["atten"])
    torch.cuda.synchronize()
    stage_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    normfac = g.calculate_normfac(invnorm)
    torch.cuda.synchronize()
    normfac_s = time.perf_counter() - t0
    del st, invnorm
    img = g.init_image()
    g.calculate_osem3d(img, normfac, prompt, add, iterations=1)      # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iterations):
        img = g.init_image()
        g.calculate_osem3d(img, normfac, prompt, add, iterations=1)
    torch.cuda.synchronize()
    gpu_it = (time.perf_counter() - t0) / iterations
    log(f"c1: GPU iteration (16 subsets) {gpu_it * 1e3:.2f} ms")
    out = {"workload": "span 9, 128x128x207, 2209-plane 128x144 sinogram (-X 128 of 256x288), 200M prompts, "
                       "randoms 10%, scatter 15%, 16 subsets x 1 iteration",
           "stage_s": round(stage_s, 4), "normfac_s": round(normfac_s, 4),
           "gpu_iteration_ms": round(gpu_it * 1e3, 3), "gpu_subset_iters_per_s": round(g.subsets / gpu_it, 2),
           "gpu_backproject_schedule": g.backproject_schedule}
    if run_cpu:
        from oracle import oracle
        oracle.build()
        o = oracle.Oracle(**DYN_GEOM)
        pr = paircol_to_sino(g, torch, prompt).cpu().numpy()
        ad = paircol_to_sino(g, torch, add).cpu().numpy()
        nfs = [np.ascontiguousarray(normfac[s].cpu().numpy()) for s in range(g.subsets)]
        ref = o.init_image()
        corr = o.image_zeros()
        t0 = time.perf_counter()
        for i in range(g.subsets):
            s_ = int(o.order[i])
            o.osem_subset(ref, corr, nfs[s_], ad, pr, s_, nthreads=cpu_threads, workers=cpu_threads)
        o.end_iteration(ref)
        cpu_it = time.perf_counter() - t0
        log(f"c1: CPU oracle iteration {cpu_it:.2f} s ({cpu_threads} threads)")
        g.set_backproject_schedule(cpu_threads)
        img = g.init_image()
        g.calculate_osem3d(img, normfac, prompt, add, iterations=1)
        same = bool(np.array_equal(img.cpu().numpy(), ref))
        out.update({"cpu_iteration_s": round(cpu_it, 3), "cpu_subset_iters_per_s": round(g.subsets / cpu_it, 3),
                    "cpu_threads": cpu_threads, "cpu_kind": "port", "cpu_model": cpu_model(),
                    "gpu_image_bitexact_vs_cpu_same_schedule": same})
    g.close()
    return out


def psf_leg(torch, g, image0, normfac_psf, prompt, add, order, steps):
    """The same subset updates with the PSF resolution model on (-B with the HRRT defaults):
    blurred image into the forward projector, blurred correction before the update."""
    g.set_psf()
    image = image0.clone()
    tmp = torch.zeros_like(image)
    for i in range(2):
        g.osem_subset(image, normfac_psf, prompt, add, order[i])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        g.osem_subset(image, normfac_psf, prompt, add, order[(2 + i) % g.subsets])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        g.convolve3d(image, out=tmp)
    e1.record()
    torch.cuda.synchronize()
    taps = [len(g.psf_kernel(a)) for a in range(3)]
    g.set_psf(enable=False)
    log(f"PSF: {steps} subset updates in {dt:.3f} s, convolve3d {e0.elapsed_time(e1) / 10:.3f} ms")
    return {"subset_iters_per_s": round(steps / dt, 3), "ms_per_subset": round(1e3 * dt / steps, 4),
            "convolve3d_ms": round(e0.elapsed_time(e1) / 10, 4), "kernel_taps_xyz": taps,
            "note": "two convolve3d per subset (image, correction), sensitivity blurred before inversion"}


def randoms_leg(torch, steps, seed, cpu_mps, span):
    """Smoothed randoms (gen_delays) for the workload's sinogram: singles from a synthetic
    coincidence histogram (delayed counts of the model d = tau s F T) on the host, then the
    smoothing on the GPU.  CPU baseline: the oracle on the listed module pairs."""
    import numpy as np
    from hrrt_reconstruction_amd.randoms import CRYSTALS, GenDelays
    rng = np.random.default_rng(seed)
    tau, T = 6e-9, 600.0
    s = rng.uniform(2e3, 6e3, CRYSTALS)                       # cps; dead-time-free model
    head = s.reshape(2, 104, 8, 72).sum(axis=(0, 1, 3))
    fan = np.array([sum(head[(h + j + 2) % 8] for j in range(5)) for h in range(8)])
    dcoins = rng.poisson(tau * s * np.broadcast_to(fan[None, None, :, None], (2, 104, 8, 72)).reshape(-1) * T)
    coins = np.concatenate([dcoins * 10, dcoins]).astype(np.int32)   # prompts (unused), delayed
    g = GenDelays(span=span, maxrd=GEOM["maxdel"])
    t0 = time.perf_counter()
    singles, niter = g.singles(coins, T)
    singles_ms = 1e3 * (time.perf_counter() - t0)
    d_s = torch.from_numpy(singles).cuda()
    out = g.empty()
    for _ in range(2):
        g.smooth(d_s, T, out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        g.smooth(d_s, T, out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    total = float(out.double().sum().item())
    log(f"randoms: singles {singles_ms:.1f} ms ({niter} iterations), smoothing {ms:.3f} ms")
    res = {"ms_per_smoothing": round(ms, 4), "crystal_pairs_per_s": round(g.pairs_full / (ms * 1e-3), 1),
           "crystal_pairs": g.pairs_full, "singles_ms_host": round(singles_ms, 2), "singles_iterations": niter,
           "span": g.span, "planes": g.nplanes, "sum": total, "cpu_baseline": None}
    del out
    torch.cuda.empty_cache()
    if cpu_mps:
        from oracle import oracle as O
        D = O.Delays(span=span, maxrd=GEOM["maxdel"])
        t0 = time.perf_counter()
        _, pairs = D.smooth(singles, tau, T, mps=cpu_mps)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(pairs / dt, 1), "unit": "crystal pairs/s", "cores": 1, "kind": "port",
                               "sample": f"module pairs {cpu_mps} ({pairs} crystal pairs, {dt:.1f} s)"}
        log(f"randoms CPU oracle: {pairs / dt / 1e6:.1f} M crystal pairs/s")
    return res


def listmode_leg(torch, events, steps, seed, cpu_events, span):
    """List-mode histogramming: a synthetic 64-bit stream (events of the 20 module pairs, a time
    tag per ms, block singles tags, 10% delayed) histogrammed in trues mode into the workload's
    sinogram, with the coincidence maps.  CPU baseline: the oracle on the first cpu_events."""
    import numpy as np
    from hrrt_reconstruction_amd import lmhist
    from hrrt_reconstruction_amd.phantom import lm_synthetic_stream
    from hrrt_reconstruction_amd.randoms import GenDelays
    words = lm_synthetic_stream(events, seed, events_per_ms=20000, singles_every=2000)
    d = torch.from_numpy(words.view(np.int32)).cuda()
    g = GenDelays(span=span, maxrd=GEOM["maxdel"])
    for _ in range(2):
        lmhist.histogram(g, d, mode=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sino, _, cp, cd, st = lmhist.histogram(g, d, mode=0)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    log(f"list mode: {events} events in {ms:.2f} ms ({events / ms / 1e6:.2f} G events/s)")
    res = {"events": events, "words": int(words.size), "ms_per_histogram": round(ms, 3),
           "events_per_s": round(events / (ms * 1e-3), 1), "prompts_binned": st["prompts"], "span": span,
           "note": "whole stream per call: chunk walk, event keys, radix sort by sinogram tile, per-tile counts into int16 bins, statistics", "cpu_baseline": None}
    del sino, cp, cd, d
    torch.cuda.empty_cache()
    if cpu_events:
        from oracle import oracle as O
        D = O.Delays(span=span, maxrd=GEOM["maxdel"])
        n = min(cpu_events, events)
        sub = words[: 2 * n + 2 * (n // 20000 + 1) * 1 + 8]
        t0 = time.perf_counter()
        r = O.lm_histogram(D, sub, mode=0)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(int(r["stats"][2]) / dt, 1), "unit": "events/s", "cores": 1,
                               "kind": "port", "sample": f"first {int(r['stats'][2])} events of the stream ({dt:.1f} s)"}
        log(f"list mode CPU oracle: {int(r['stats'][2]) / dt / 1e6:.1f} M events/s")
    return res


def batched_frames(torch, nframes, iterations, seed, tune=True):
    """Dynamic study in batched mode: every frame's subsets run in one launch per kernel, each subset
    of an iteration replayed from a HIP graph.  FP32 form and FP16-input precision form, each with
    the forward launch shape the batched-path autotuner picks (unless tune is False)."""
    from hrrt_reconstruction_amd import phantom as ph
    from hrrt_reconstruction_amd.osem import Osem3D
    g = Osem3D("cuda:0", **DYN_GEOM)
    p = g.params
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    cg = torch.Generator()
    cg.manual_seed(seed)
    dur, mids, counts, rates = ph.dynamic_frames(nframes, 0.5e6, 200e6, cg)
    pc = (g.views // 2, 2, g.N, g.spp)
    prompts = torch.zeros((nframes,) + pc, dtype=torch.int16, device="cuda")
    adds = torch.zeros((nframes,) + pc, dtype=torch.float32, device="cuda")
    est = g.estimate_zeros()
    norm = atten = invnorm = None
    for f in range(nframes):
        act = [1.0 + 3.0 * math.exp(-r * mids[f]) for r in rates]
        pg = torch.Generator()
        pg.manual_seed(seed + 1)                 # same sphere positions in every frame
        img, _ = ph.make_phantom(g.N, g.Z4, g.Z, p.bin_size * g.dims.x_rebin, p.plane_sep, 6, 200, pg, "cuda",
                                 sphere_activity=act)
        proj = torch.zeros((g.S, g.views, g.N), dtype=torch.float32, device="cuda")
        for v in range(g.views // 2):
            g.forward_proj3d(img, v, prj=est)
            proj[:, v, :] = est[: g.S]
            proj[:, v + g.views // 2, :] = est[g.S:]
        # file-layout sinograms: rebin the projections back to 256 x 288 by repetition
        rb = g.dims.x_rebin
        proj_file = proj.repeat_interleave(rb, 1).repeat_interleave(rb, 2) / (rb * rb)
        st = ph.make_study(proj_file, p.rings, p.span, p.maxdel, p.def_elements, p.def_angles, p.bin_size,
                           p.plane_sep, p.crystal_radius, counts[f], gen, "cuda", with_scatter=False,
                           norm=norm, atten=atten)
        norm, atten = st["norm"], st["atten"]
        pr, ad, inv = g.stage(st["prompt"], st["delayed"], st["norm"], None, st["atten"])
        prompts[f].copy_(pr)
        adds[f].copy_(ad)
        if invnorm is None:
            invnorm = inv
        del proj, proj_file, st, pr, ad, inv
    torch.cuda.synchronize()
    log(f"dynamic study: {nframes} frames generated and staged")
    nf32 = g.calculate_normfac(invnorm)
    inv16 = g.to_half(invnorm)
    nf16 = g.calculate_normfac_mixed(inv16)
    adds16 = g.to_half(adds)
    stream = torch.cuda.Stream()
    res, out, shapes = {}, {}, {}
    init = g.init_image()
    for form in ("fp32", "fp16"):
        imgs = torch.stack([init for _ in range(nframes)])
        ll = torch.zeros((g.subsets, nframes), dtype=torch.float64, device="cuda")
        if tune:
            from hrrt_reconstruction_amd.autotune import tune_frames
            shape, times = (tune_frames(g, imgs, nf32, prompts, adds) if form == "fp32"
                            else tune_frames(g, imgs, nf16, prompts, adds16, mixed=True))
            shapes[form] = list(shape)
            log(f"batched {form} autotune forward " + ", ".join(f"{k}: {v:.3f} ms" for k, v in times.items())
                + f" -> {shape}")

        def iterate(n):
            stream.wait_stream(torch.cuda.current_stream())   # inputs written on the current stream
            with torch.cuda.stream(stream):
                for _ in range(n):
                    if form == "fp32":
                        g.osem_iteration_frames(imgs, nf32, prompts, adds)
                    else:
                        g.osem_iteration_frames_mixed(imgs, nf16, prompts, adds16, invnorm16=inv16, loglik=ll)
            stream.synchronize()
        iterate(1)                            # captures the subsets' graphs
        imgs.copy_(init.expand_as(imgs))      # same buffers: the graphs are replayed
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        iterate(iterations)
        out[form] = time.perf_counter() - t0
        res[form] = (imgs, out[form], ll)
        log(f"batched {form}: {nframes} frames x {g.subsets} subsets x {iterations} iterations in {out[form]:.3f} s")
    # the single-frame FP32 path on one frame, for the per-frame rate without batching
    img = init.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.calculate_osem3d(img, nf32, prompts[0], adds[0], iterations=iterations)
    torch.cuda.synchronize()
    out["single"] = time.perf_counter() - t0
    a, b = res["fp32"][0], res["fp16"][0]
    rel = [float(torch.linalg.vector_norm((b[f] - a[f]).double()) / torch.linalg.vector_norm(a[f].double()))
           for f in range(nframes)]
    return {
        "workload": f"dynamic study, HRRT span 9, 128x128x207 image, {nframes} frames "
                    f"({counts[0] / 1e6:.2f}M..: log-uniform 0.5M-200M prompts, spheres washing out, randoms 10%), "
                    f"{g.subsets} subsets x {iterations} iterations, one launch per kernel for all frames, "
                    "subsets replayed from HIP graphs",
        "frames": nframes, "iterations": iterations,
        "fp32_frames_per_s": round(nframes / out["fp32"], 3),
        "fp16_frames_per_s": round(nframes / out["fp16"], 3),
        "fp16_vs_fp32": round(out["fp32"] / out["fp16"], 3),
        "fp32_single_frame_frames_per_s": round(1.0 / out["single"], 3),
        "fp16_batched_vs_fp32_single": round(out["single"] * nframes / out["fp16"], 3),
        "fp16_image_rel_l2_vs_fp32_max": float(f"{max(rel):.3e}"),
        "fp16_loglik_last_subset": [float(x) for x in res["fp16"][2][-1].cpu().numpy()[:3]],
        "fp_launch_shape": shapes or None,
    }


def c5_leg(torch, nframes, iterations, seed, fp64_frames, fp32_frames, tune=True):
    """Combined form at full scale: a span-3 256x256x207 dynamic study of nframes frames (counts
    log-uniform between 1M and 300M, spheres washing out, randoms 10 %, scatter 15 %), FP16
    additive terms and sensitivity weights, int16 prompts, FP32 images; every subset of an
    iteration replayed for all frames from a HIP graph.  Also timed: the FP32 batched form on
    fp32_frames frames and the FP32 single-frame path on one frame (frames/s of each), and the
    image of fp64_frames frames (lowest and highest counts first) against the FP64 oracle."""
    import numpy as np
    from hrrt_reconstruction_amd import phantom as ph
    from hrrt_reconstruction_amd.osem import Osem3D
    g = Osem3D("cuda:0", **GEOM)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    cg = torch.Generator()
    cg.manual_seed(seed)
    dur, mids, counts, rates = ph.dynamic_frames(nframes, 1e6, 300e6, cg)
    pc = (g.views // 2, 2, g.N, g.spp)
    prompts = torch.zeros((nframes,) + pc, dtype=torch.int16, device="cuda")
    adds16 = torch.zeros((nframes,) + pc, dtype=torch.int16, device="cuda")
    keep = sorted(range(nframes), key=lambda f: counts[f])
    keep = ([keep[0], keep[-1]] + keep[1:-1])[:max(fp64_frames, fp32_frames, 1)]
    adds32 = {}
    half_range = [0, 0]            # FP16 conversions: inputs sent to +-inf, to subnormal or zero
    factors, invnorm = None, None
    for f in range(nframes):
        act = [1.0 + 3.0 * math.exp(-r * mids[f]) for r in rates]
        pg = torch.Generator()
        pg.manual_seed(seed + 1)                 # same sphere positions in every frame
        _, st = ph.make_geometry_study(g, counts[f], gen, pg, with_scatter=True, sphere_activity=act,
                                       norm=factors[0] if factors else None, atten=factors[1] if factors else None)
        factors = (st["norm"], st["atten"])
        pr, ad, inv = g.stage(st["prompt"], st["delayed"], st["norm"], st["scatter"], st["atten"])
        prompts[f].copy_(pr)
        a16, (o_, u_) = g.to_half_range(ad)
        adds16[f].copy_(a16)
        half_range[0] += o_
        half_range[1] += u_
        del a16
        if f in keep:
            adds32[f] = ad
        if invnorm is None:
            invnorm = inv
        del st, pr, inv
        if f % 8 == 7:
            torch.cuda.synchronize()
            log(f"c5: {f + 1}/{nframes} frames generated and staged")
    del factors
    torch.cuda.synchronize()
    inv16, (o_, u_) = g.to_half_range(invnorm)
    log(f"c5: FP16 additive terms: {half_range[0]} overflowed, {half_range[1]} subnormal or zero; "
        f"1/(N*A): {o_} overflowed, {u_} subnormal or zero")
    nf16 = g.calculate_normfac_mixed(inv16)
    nf32 = g.calculate_normfac(invnorm)
    stream = torch.cuda.Stream()
    out = {}

    def run_batched(imgs, fp16, pr, ad, iters):
        stream.wait_stream(torch.cuda.current_stream())   # inputs written on the current stream
        with torch.cuda.stream(stream):
            for _ in range(iters):
                if fp16:
                    g.osem_iteration_frames_mixed(imgs, nf16, pr, ad)
                else:
                    g.osem_iteration_frames(imgs, nf32, pr, ad)
        stream.synchronize()

    # FP16 batched form, all frames: once to capture the graphs, then timed on the same buffers
    init = g.init_image()
    imgs = torch.stack([init for _ in range(nframes)])
    shapes = {}
    if tune:
        from hrrt_reconstruction_amd.autotune import tune_frames
        shape, times = tune_frames(g, imgs, nf16, prompts, adds16, mixed=True)
        shapes["fp16"] = list(shape)
        log("c5: fp16 autotune forward " + ", ".join(f"{k}: {v:.3f} ms" for k, v in times.items()) + f" -> {shape}")
    run_batched(imgs, True, prompts, adds16, 1)
    imgs.copy_(init.expand_as(imgs))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_batched(imgs, True, prompts, adds16, iterations)
    t16 = time.perf_counter() - t0
    log(f"c5: fp16 batched {nframes} frames x {g.subsets} subsets x {iterations} iterations in {t16:.3f} s")
    out["fp16_inputs_overflowed"] = half_range[0] + o_
    out["fp16_inputs_subnormal_or_zero"] = {"additive": half_range[1], "invnorm": u_}
    out["fp16_batched_frames_per_s"] = round(nframes / t16, 3)
    # FP32 batched form on fp32_frames frames (the same work per frame)
    sel = keep[:fp32_frames]
    if sel:
        pr32 = torch.stack([prompts[f] for f in sel])
        ad32 = torch.stack([adds32[f] for f in sel])
        i32 = torch.stack([init for _ in sel])
        if tune:
            shape, times = tune_frames(g, i32, nf32, pr32, ad32)
            shapes["fp32"] = list(shape)
            log("c5: fp32 autotune forward " + ", ".join(f"{k}: {v:.3f} ms" for k, v in times.items()) + f" -> {shape}")
        run_batched(i32, False, pr32, ad32, 1)
        i32.copy_(init.expand_as(i32))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_batched(i32, False, pr32, ad32, iterations)
        t32 = time.perf_counter() - t0
        out["fp32_batched_frames_per_s"] = round(len(sel) / t32, 3)
        out["fp32_batched_frames"] = len(sel)
        # FP32 single-frame path, one frame after another (the plain GPU form)
        img = g.init_image()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.calculate_osem3d(img, nf32, prompts[sel[0]], adds32[sel[0]], iterations=iterations)
        torch.cuda.synchronize()
        t1 = time.perf_counter() - t0
        out["fp32_single_frame_frames_per_s"] = round(1.0 / t1, 3)
        out["fp16_batched_vs_fp32_single"] = round(out["fp16_batched_frames_per_s"] / out["fp32_single_frame_frames_per_s"], 3)
        out["fp16_batched_vs_fp32_batched"] = round(out["fp16_batched_frames_per_s"] / out["fp32_batched_frames_per_s"], 3)
        del pr32, ad32, i32
    # FP64 oracle images of the first fp64_frames kept frames
    if fp64_frames > 0:
        from oracle import oracle
        oracle.build()
        o64 = oracle.Oracle(double=True, **GEOM)
        o64.set_norm_zero_mask(np.pad(g.norm_zero_mask(), ((0, 1), (0, 0))))
        inv64 = paircol_to_sino(g, torch, invnorm).cpu().numpy().astype(np.float64)
        ncores = min(len(os.sched_getaffinity(0)), 16)
        nf64 = o64.calculate_normfac(inv64, nthreads=g.sviews // 2)
        del inv64
        got = imgs.cpu().numpy()
        rel = {}
        for f in keep[:fp64_frames]:
            add64 = paircol_to_sino(g, torch, adds32[f]).cpu().numpy().astype(np.float64)
            pr = paircol_to_sino(g, torch, prompts[f]).cpu().numpy()
            ref = o64.init_image()
            corr = o64.image_zeros()
            for _ in range(iterations):
                for i in range(g.subsets):
                    s = int(g.order[i])
                    o64.osem_subset(ref, corr, np.ascontiguousarray(nf64[s]), add64, pr, s, nthreads=1, workers=ncores)
                o64.end_iteration(ref)
            d = got[f].astype(np.float64) - ref
            rel[f] = float(np.linalg.norm(d) / np.linalg.norm(ref))
            log(f"c5: frame {f} ({counts[f] / 1e6:.1f}M prompts) image rel-L2 vs FP64 oracle {rel[f]:.2e}")
        out["fp16_image_rel_l2_vs_fp64"] = {str(f): float(f"{v:.3e}") for f, v in rel.items()}
        out["fp64_reference"] = f"FP64 oracle, {iterations} iterations, {ncores} host threads"
    out.update({
        "workload": f"dynamic study, HRRT span 3, 256x256x207 image, {nframes} frames (log-uniform 1M-300M prompts, "
                    "spheres washing out, randoms 10%, scatter 15%), FP16 additive terms and sensitivity weights, "
                    f"int16 prompts, FP32 images, {g.subsets} subsets x {iterations} iterations, subsets replayed "
                    "from HIP graphs for all frames",
        "frames": nframes, "iterations": iterations,
        "counts_m": [round(c / 1e6, 2) for c in counts],
        "fp_launch_shape": shapes or None,
    })
    del prompts, adds16, adds32, imgs, nf16, nf32
    g.close()
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-subsets", type=int, default=14, help="timed subsets of the CPU baseline, about 10 s on 8 host threads (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=8)
    ap.add_argument("--no-tune", action="store_true", help="skip the forward-projector launch-shape autotuner")
    ap.add_argument("--frames", type=int, default=24, help="frames of the batched dynamic study (0 = skip)")
    ap.add_argument("--frame-iterations", type=int, default=2)
    ap.add_argument("--psf-steps", type=int, default=16,
                    help="timed subset updates with the PSF resolution model (-B, HRRT defaults; 0 = skip)")
    ap.add_argument("--randoms-steps", type=int, default=10,
                    help="timed smoothed-randoms (gen_delays) computations (0 = skip)")
    ap.add_argument("--randoms-span", type=int, default=GEOM["span"], choices=(3, 9),
                    help="span of the smoothed randoms (default: the workload's)")
    ap.add_argument("--listmode-events", type=int, default=40_000_000,
                    help="events of the synthetic list-mode stream histogrammed per step (0 = skip)")
    ap.add_argument("--listmode-steps", type=int, default=5)
    ap.add_argument("--c5-frames", type=int, default=48,
                    help="frames of the span-3 FP16 batched dynamic study (0 = skip)")
    ap.add_argument("--c5-iterations", type=int, default=3)
    ap.add_argument("--c5-fp32-frames", type=int, default=4,
                    help="frames of that study also run in the FP32 batched and single-frame forms")
    ap.add_argument("--precision", choices=("fp32", "fp16"), default="fp32",
                    help="timed subset path: fp32 (the reference's arithmetic, bit-exact) or fp16 (the "
                         "precision form: FP16 additive term); the reported line is fp32 by default")
    ap.add_argument("--no-host-stage", dest="host_stage", action="store_false",
                    help="skip the staging from host memory")
    ap.add_argument("--c1-iterations", type=int, default=5,
                    help="timed GPU iterations of the span-9 128x128 case (0 = skip)")
    ap.add_argument("--c5-fp64-frames", type=int, default=2,
                    help="frames of that study checked against the FP64 oracle on the host (0 = skip)")
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/traffic_c2.txt from two rocprofv3 --pmc output folders and exit")
    ap.add_argument("--traffic-launches", type=int, default=4)
    args = ap.parse_args()
    if args.write_traffic:
        write_traffic(args.write_traffic[0], args.write_traffic[1], args.traffic_launches)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(frames and view pairs would shard across GPUs)", file=sys.stderr)
        sys.exit(2)

    import torch
    from hrrt_reconstruction_amd.osem import Osem3D

    torch.cuda.set_device(0)
    # plain copy of known size, used to scale the PMC byte counters of the profile runs
    cal = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
    cal2 = torch.empty_like(cal)
    cal2.copy_(cal)
    torch.cuda.synchronize()
    del cal, cal2

    log("torch ready")
    g = Osem3D("cuda:0", **GEOM)
    log("context created")
    npairs = g.sviews // 2
    img_true, study = build_study(g, torch, args.seed)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    prompt, add, invnorm = g.stage(study["prompt"], study["delayed"], study["norm"], study["scatter"], study["atten"])
    torch.cuda.synchronize()
    stage_s = time.perf_counter() - t0
    log(f"staged sinograms in {stage_s:.2f} s")
    stage_host_s = stage_pinned_s = None
    if args.host_stage:
        # the same files in host memory: pageable arrays, then pinned tensors (PCIe included)
        names = ("prompt", "delayed", "norm", "scatter", "atten")
        host = [study[k].cpu().numpy().ravel() for k in names]
        g.stage_host(*host)                       # first call sets up the device file copies
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs = g.stage_host(*host)
        torch.cuda.synchronize()
        stage_host_s = time.perf_counter() - t0
        same = all(torch.equal(a, b) for a, b in zip(outs, (prompt, add, invnorm)))
        del outs
        pinned = [torch.from_numpy(h).pin_memory() for h in host]
        del host
        t0 = time.perf_counter()
        outs = g.stage_host(*pinned)
        torch.cuda.synchronize()
        stage_pinned_s = time.perf_counter() - t0
        same = same and all(torch.equal(a, b) for a, b in zip(outs, (prompt, add, invnorm)))
        del outs, pinned
        if not same:
            raise RuntimeError("staging from host memory differs from staging from device memory")
        log(f"staged from host memory in {stage_host_s:.3f} s (pageable), {stage_pinned_s:.3f} s (pinned)")
    t0 = time.perf_counter()
    normfac = g.calculate_normfac(invnorm)
    torch.cuda.synchronize()
    normfac_s = time.perf_counter() - t0
    log(f"sensitivity images in {normfac_s:.2f} s")
    normfac_psf = None
    if args.psf_steps > 0:
        g.set_psf()
        normfac_psf = g.calculate_normfac(invnorm)
        g.set_psf(enable=False)
    del study, invnorm
    torch.cuda.empty_cache()

    image = g.init_image()
    image0 = image.clone()
    order = [int(s) for s in g.order]
    fp16 = args.precision == "fp16"
    if fp16:
        # the precision form on the same workload: FP16 additive term, FP32 arithmetic
        add16 = g.to_half(add)

        def subset(i):
            g.osem_subset_mixed(image, normfac, prompt, add16, i)
    else:
        def subset(i):
            g.osem_subset(image, normfac, prompt, add, i)
    shape, tune, bp_choice = None, {}, None
    if not args.no_tune:
        from hrrt_reconstruction_amd.autotune import tune_backproject, tune_forward
        shape, tune = tune_forward(g, image, normfac, prompt, add)
        log("autotune forward " + ", ".join(f"{k}: {v:.3f} ms" for k, v in tune.items()) + f" -> {shape}")
        bp_choice, tune_bp = tune_backproject(g, image, normfac, prompt, add)
        log("autotune back " + ", ".join(f"{k}: {v:.3f} ms" for k, v in tune_bp.items()) + f" -> {bp_choice}")
    for i in range(args.warmup):
        subset(order[i % g.subsets])
    torch.cuda.synchronize()

    g.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        subset(order[(args.warmup + i) % g.subsets])
        if (args.warmup + i) % g.subsets == g.subsets - 1:
            g.end_iteration(image)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} subset updates in {elapsed:.3f} s")
    kt = g.timing_read()
    g.timing_enable(False)

    value = args.steps / elapsed
    alg = algorithmic_bytes(g, npairs)
    dom = max(kt, key=lambda k: kt[k][0])
    dom_ms = kt[dom][0] / max(kt[dom][1], 1)
    achieved = alg[dom] / (dom_ms * 1e-3) / 1e9
    traffic, traffic_note = read_traffic(dom)
    flops = algorithmic_flops(g, npairs)
    dom_tflops = flops[dom] / (dom_ms * 1e-3) / 1e12
    subset_tflops = sum(flops.values()) * value / 1e12

    cpu = cpu_strong = None
    if args.cpu_subsets > 0:
        rate, dt = cpu_baseline(g, torch, prompt, add, normfac, image0, args.cpu_subsets, args.cpu_threads)
        cpu = {"value": round(rate, 4), "unit": "subset-iters/s", "cores": args.cpu_threads, "kind": "port",
               "model": cpu_model(),
               "sample": f"{args.cpu_subsets} OSEM subset updates of the same workload after one warm-up "
                         f"subset, reference thread schedule T={args.cpu_threads} ({dt:.1f} s)"}
        # the stronger CPU figure: every host core this process may use (at most 16), each
        # pair back-projected incrementally (the faster projector, schedule T = 1)
        ncores = min(len(os.sched_getaffinity(0)), 16)
        rate2, dt2 = cpu_baseline(g, torch, prompt, add, normfac, image0, args.cpu_subsets, ncores, schedule=1)
        cpu_strong = {"value": round(rate2, 4), "unit": "subset-iters/s", "cores": ncores, "kind": "port",
                      "model": cpu_model(),
                      "sample": f"{args.cpu_subsets} OSEM subset updates after one warm-up subset, {ncores} "
                                f"worker threads, every pair incremental ({dt2:.1f} s)"}

    psf = None
    if args.psf_steps > 0:
        psf = psf_leg(torch, g, image0, normfac_psf, prompt, add, order, args.psf_steps)
    randoms = None
    if args.randoms_steps > 0:
        randoms = randoms_leg(torch, args.randoms_steps, args.seed + 11, [1, 3] if args.cpu_subsets > 0 else None,
                              args.randoms_span)
    listmode = None
    if args.listmode_events > 0 and args.listmode_steps > 0:
        listmode = listmode_leg(torch, args.listmode_events, args.listmode_steps, args.seed + 13,
                                20_000_000 if args.cpu_subsets > 0 else 0, GEOM["span"])
    batched = None
    if args.frames > 0:
        batched = batched_frames(torch, args.frames, args.frame_iterations, args.seed + 7, tune=not args.no_tune)
    c1 = None
    if args.c1_iterations > 0:
        c1 = c1_leg(torch, args.seed + 19, args.c1_iterations, args.cpu_threads, args.cpu_subsets > 0)
    schedule = g.backproject_schedule
    c5 = None
    if args.c5_frames > 0:
        del prompt, add, normfac, image, image0
        g.close()
        torch.cuda.empty_cache()
        c5 = c5_leg(torch, args.c5_frames, args.c5_iterations, args.seed + 17,
                    args.c5_fp64_frames if args.cpu_subsets > 0 else 0, args.c5_fp32_frames, tune=not args.no_tune)

    out = {
        "metric": "OSEM subset-iters/s span3 256²×207; batched frames/s; % HBM roofline",
        "value": round(value, 3),
        "unit": "subset-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if not fp16 else "f32 (FP16 additive term)",
        "data": "synthetic",
        "config": {"workload": "HRRT span 3, 256x256x207 image, 6367-plane 256x288 sinogram, 200M prompts, "
                               "randoms 10%, scatter 15%, 16 subsets (9 view pairs per subset)",
                   "subsets": g.subsets, "view_pairs_per_subset": npairs,
                   "stage_s": round(stage_s, 3), "normfac_s": round(normfac_s, 3),
                   "stage_host_pageable_s": round(stage_host_s, 3) if stage_host_s else None,
                   "stage_host_pinned_s": round(stage_pinned_s, 3) if stage_pinned_s else None,
                   "fp_launch_shape": list(shape) if shape else None, "direct_bp_waves": bp_choice,
                   "back_projector": "incremental" if schedule == 1 else ("direct" if schedule == 0 else "mixed"),
                   "backproject_schedule": schedule},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic, "traffic_note": traffic_note},
        "compute_roofline": {"bound": "fp32", "kernel": dom, "achieved": round(dom_tflops, 3),
                             "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": round(dom_tflops / FP32_PEAK_TFLOPS, 5),
                             "subset_achieved": round(subset_tflops, 3),
                             "subset_frac": round(subset_tflops / FP32_PEAK_TFLOPS, 5)},
        "kernels": {k: {"ms_per_launch": round(v[0] / max(v[1], 1), 4), "launches": v[1],
                        "alg_bytes": alg[k], "alg_flops": flops[k]} for k, v in kt.items()},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_strong,
        "psf": psf,
        "randoms": randoms,
        "listmode": listmode,
        "batched": batched,
        "c1": c1,
        "c5": c5,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
