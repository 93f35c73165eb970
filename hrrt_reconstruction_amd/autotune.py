"""Launch-shape autotuner for the projectors.

The forward projector's workgroup owns planes of one ray column (one x of
one view pair); the shape picks how many planes each thread accumulates in
registers and how many workgroups share one column (each recomputing the
column's rotated samples).  The back
projector's workgroup stages one sinogram column in LDS; its shape is the
number of waves (voxel rows in flight) sharing that column.  Results are
bit-identical for every shape; only speed differs.  The choice between the
incremental and the direct back projector is not a shape: the two differ by
float rounding, as they do in the reference, and that choice stays with the
caller (hrrt_set_backproject_schedule, the reference's -T).  The tuner times the
candidates on the caller's own subset update with HIP events and keeps the
fastest; tune_frames does the same on the batched update of every frame of a
dynamic study, whose forward launches have a frame dimension of their own.
"""
from __future__ import annotations

CANDIDATES = [(1, 0), (1, 6), (2, 0), (2, 3), (2, 4), (3, 0), (3, 3), (4, 0), (4, 2), (4, 3)]


BP_CANDIDATES = [4, 8, 12, 16]


def _time_kernel(g, work, normfac, prompt, add, subsets, repeats, kernel):
    import torch
    g.osem_subset(work, normfac, prompt, add, subsets[0])      # warm-up / first launch of this variant
    torch.cuda.synchronize()
    g.timing_enable(True)
    for _ in range(repeats):
        for s in subsets:
            g.osem_subset(work, normfac, prompt, add, s)
    kt = g.timing_read()
    g.timing_enable(False)
    ms, n = kt[kernel]
    return ms / max(n, 1)


def tune_backproject(g, image, normfac, prompt, add, subsets=(0, 8), repeats=2, candidates=BP_CANDIDATES):
    """Choose the direct back projector's workgroup size (waves per workgroup).

    Workgroup sizes give bit-identical results; only speed differs.  The direct projector
    serves the remainder pairs of a -T schedule, schedule 0 and the precision form, so it
    is timed under schedule 0.  The caller's back-projection schedule (the reference's
    -T, which decides per pair between the incremental and the direct projector and with
    it the float rounding) is restored before returning, and its own back-projection
    time is reported as "schedule T=<n>".  The tuner never changes the algorithm.
    Returns (waves, {label: ms per back-projection launch})."""
    work = image.clone()
    results = {}
    keep = g.backproject_schedule
    try:
        g.set_backproject_schedule(0)
        for w in candidates:
            g.set_backproject_shape(w)
            results[f"direct/{w}w"] = _time_kernel(g, work, normfac, prompt, add, subsets, repeats, "backproject")
        best_w = min(candidates, key=lambda w: results[f"direct/{w}w"])
        g.set_backproject_shape(best_w)
    finally:
        g.set_backproject_schedule(keep)
    if keep != 0:
        results[f"schedule T={keep}"] = _time_kernel(g, work, normfac, prompt, add, subsets, repeats, "backproject")
    del work
    return best_w, results


def tune_forward(g, image, normfac, prompt, add, subsets=(0, 8), repeats=2, candidates=CANDIDATES):
    """Return ((tasks_per_thread, workgroups_per_ray), {shape: ms per forward launch})."""
    import torch
    work = image.clone()
    results = {}
    for shape in candidates:
        try:
            g.set_launch_shape(*shape)
        except Exception:
            continue
        g.osem_subset(work, normfac, prompt, add, subsets[0])      # warm-up / first launch of this variant
        torch.cuda.synchronize()
        g.timing_enable(True)
        for _ in range(repeats):
            for s in subsets:
                g.osem_subset(work, normfac, prompt, add, s)
        kt = g.timing_read()
        g.timing_enable(False)
        ms, n = kt["fp_ratio"]
        results[shape] = ms / max(n, 1)
    best = min(results, key=results.get)
    g.set_launch_shape(*best)
    del work
    return best, results


def tune_frames(g, images, normfac, prompts, adds, mixed=False, subsets=(0, 8), repeats=2, candidates=CANDIDATES):
    """Forward launch shape for the batched path (hrrt_osem_subset_frames, or the FP16-input
    form with mixed=True): every candidate timed on one subset update of all frames, the
    fastest kept on the context.  Shapes give bit-identical results; a new shape drops the
    context's captured iteration graphs, which the next iteration call captures again.
    Returns ((tasks_per_thread, workgroups_per_ray), {shape: ms per forward launch})."""
    import torch
    work = images.clone()
    results = {}

    def run(s):
        if mixed:
            g.osem_subset_frames_mixed(work, normfac, prompts, adds, s)
        else:
            g.osem_subset_frames(work, normfac, prompts, adds, s)

    for shape in candidates:
        try:
            g.set_launch_shape(*shape)
        except Exception:
            continue
        run(subsets[0])                       # warm-up / first launch of this variant
        torch.cuda.synchronize()
        g.timing_enable(True)
        for _ in range(repeats):
            for s in subsets:
                run(s)
        kt = g.timing_read()
        g.timing_enable(False)
        ms, n = kt["fp_ratio"]
        results[shape] = ms / max(n, 1)
    best = min(results, key=results.get)
    g.set_launch_shape(*best)
    del work
    return best, results
