"""Launch-shape autotuner for the projectors.

The forward projector's workgroup owns planes of one ray column (one x of
one view pair); the shape picks how many planes each thread accumulates in
registers and how many workgroups share one column (each recomputing the
column's rotated samples).  The back
projector's workgroup stages one sinogram column in LDS; its shape is the
number of waves (voxel rows in flight) sharing that column.  Results are
bit-identical for every shape; only speed differs.  The back projector itself
(direct or incremental) is chosen the same way; those two differ by float
rounding, as they do in the reference.  The tuner times the
candidates on the caller's own subset update with HIP events and keeps the
fastest.
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
    """Choose the back projector: the direct variant at each workgroup size, and the
    incremental one (every pair, the reference's -T 1 schedule) where the geometry allows it.
    Returns (choice, {label: ms per back-projection launch}); choice is "incremental" or the
    direct variant's waves per workgroup.  The two variants differ by float rounding only."""
    work = image.clone()
    results = {}
    keep = g.backproject_schedule
    g.set_backproject_schedule(0)
    for w in candidates:
        g.set_backproject_shape(w)
        results[f"direct/{w}w"] = _time_kernel(g, work, normfac, prompt, add, subsets, repeats, "backproject")
    best_w = min(candidates, key=lambda w: results[f"direct/{w}w"])
    g.set_backproject_shape(best_w)
    choice = best_w
    try:
        g.set_backproject_schedule(1)
    except Exception:
        g.set_backproject_schedule(keep if keep == 0 else 0)
    else:
        results["incremental"] = _time_kernel(g, work, normfac, prompt, add, subsets, repeats, "backproject")
        if results["incremental"] < results[f"direct/{best_w}w"]:
            choice = "incremental"
        else:
            g.set_backproject_schedule(0)
    del work
    return choice, results


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
