"""Autotuner for the launch parameters behind rm_set_tuning.

Every setting gives the same results (traversal bit for bit; renders up to
the summation order of accumulator sets), so the tuner only times candidates
on the caller's own workload and keeps the fastest:

  tune_trace   traversal variant (stack walk / persistent step walk with
               the stack in scratch or partly in LDS), resident blocks per CU, the idle-lane threshold at which
               a persistent wave fetches new rays, ray reordering, LDS stack
               entries, triangles per step and walk phases, on the caller's
               ray batches;
  tune_render  resident blocks per CU of the persistent bounce kernel (the
               ray-queue consumers) and paths per pixel per round (the depth of
               the path queues), on a copy of the caller's batch with the spp
               scaled down.
"""
from __future__ import annotations

import itertools
import time

from ._lib import RM_ERR_OOM, RmError
from . import (TUNE_FETCH_MIN, TUNE_LDS_STACK_ENTRIES, TUNE_PERSISTENT_BLOCKS_PER_CU, TUNE_RAY_SORT,
               TUNE_NEE_SPLIT, TUNE_RENDER_BLOCKS_PER_CU, TUNE_RENDER_MIN_WAVES, TUNE_TRACE_VARIANT, TUNE_TRIS_PER_STEP,
               TUNE_WALK_PHASES, RmRenderArgs, get_tuning, render_batch, set_tuning)

TRACE_KEYS = (TUNE_TRACE_VARIANT, TUNE_PERSISTENT_BLOCKS_PER_CU, TUNE_FETCH_MIN, TUNE_RAY_SORT, TUNE_LDS_STACK_ENTRIES,
              TUNE_TRIS_PER_STEP, TUNE_WALK_PHASES)


def _time(fn, reps):
    import torch
    best = float("inf")
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune_trace(scene, batches, reps=3, candidates=None, log=None) -> dict:
    """batches: list of device ray tensors [n, 6].  Sets and returns the
    fastest {key: value} for closest-hit traversal of all batches."""
    if candidates is None:
        candidates = [(0, 8, 16, s, 32, 0, 0) for s in (0, 4)]
        candidates += [(2, b, f, s, 32, 0, 0) for b, f, s in itertools.product((4, 8), (32, 48), (0, 4))]
        candidates += [(3, b, f, 5, e, 0, 0) for b, f, e in itertools.product((4, 8), (32, 48), (6, 8, 12, 32))]
        candidates += [(3, b, f, s, e, p, ph) for b, f, e, p, s, ph in
                       itertools.product((8,), (32, 48), (8, 10), (0, 2, 4), (1, 3, 4, 5), (24, 32, 40))]

    def run():
        for r in batches:
            scene.trace_closest(r)

    results = []
    for c in candidates:
        for k, v in zip(TRACE_KEYS, c):
            set_tuning(k, v)
        run()  # first use of a setting (workspace growth)
        t = _time(run, reps)
        results.append((t, c))
        if log:
            log(f"[autotune] trace {c}: {t * 1e3:.2f} ms")
    t, best = min(results)
    for k, v in zip(TRACE_KEYS, best):
        set_tuning(k, v)
    return {"variant": best[0], "blocks_per_cu": best[1], "fetch_min": best[2], "ray_sort": best[3],
            "lds_stack_entries": best[4], "tris_per_step": best[5], "walk_phases": best[6], "seconds": t,
            "tried": len(results)}


def tune_render(jobs, spp_scale=0.125, blocks=(4, 8, 16), waves=(1, 3), paths_per_round=(0, 2, 4, 8),
                walks=((3, 10, 2, 0), (3, 10, 0, 0), (3, 10, 1, 0), (3, 10, 4, 0), (3, 8, 2, 0), (3, 6, 2, 0),
                       (2, 32, 2, 0), (3, 10, 0, 32), (3, 10, 0, 40), (3, 10, 1, 32), (3, 10, 2, 32)),
                full_paths=(4, 8), log=None) -> dict:
    """jobs: list of (Scene, RmRenderArgs, Frame, post_options) as for
    render_batch.  Times the batch (hipGraph launch) with every job's spp
    scaled by spp_scale: first every (blocks per CU, register cap) pair with
    automatic paths per round and the walk settings of walks[0], then the
    other paths-per-round values at the best pair, then the other walk
    settings (trace variant, LDS entries per lane, triangles per step, walk
    phases: variant 2 keeps the walk stacks in scratch, 3 their first entries
    in LDS).  Then paths per round once more, at the full spp, among the
    scaled choice and full_paths: with fewer samples per pixel a round of
    many paths is often the last, partly empty one, so the scaled batch
    prefers fewer paths per round than the full one.  Sets the launch
    parameters and returns the choice; the caller applies paths_per_round
    (0 = automatic) to its jobs' arguments."""
    small = []
    for S, a, fr, post in jobs:
        b = RmRenderArgs.from_buffer_copy(a)
        b.spp = max(4, int(round(a.spp * spp_scale)))
        small.append([S, b, fr, post])

    def trial(bpc, w, k, walk):
        for j in small:
            j[1].paths_per_round = k
        set_tuning(TUNE_RENDER_BLOCKS_PER_CU, bpc)
        set_tuning(TUNE_RENDER_MIN_WAVES, w)
        set_tuning(TUNE_TRACE_VARIANT, walk[0])
        set_tuning(TUNE_LDS_STACK_ENTRIES, walk[1])
        set_tuning(TUNE_TRIS_PER_STEP, walk[2])
        set_tuning(TUNE_WALK_PHASES, walk[3] if len(walk) > 3 else 0)
        try:
            t = _time(lambda: render_batch([tuple(j) for j in small], use_graph=True), 1)
        except RmError as e:
            if e.code != RM_ERR_OOM:
                raise
            t = float("inf")  # the candidate's work buffers do not fit
        if log:
            log(f"[autotune] render blocks/CU {bpc} min waves {w} paths/round {k or 'auto'} walk {walk}: {t:.3f} s")
        return (t, bpc, w, k, walk)

    results = [trial(b, w, paths_per_round[0], walks[0]) for b in blocks for w in waves]
    _, bpc, w, _, _ = min(results)
    results += [trial(bpc, w, k, walks[0]) for k in paths_per_round[1:]]
    _, bpc, w, k, _ = min(results)
    results += [trial(bpc, w, k, walk) for walk in walks[1:]]
    t, bpc, w, k, walk = min(results)
    full = []
    for kk in dict.fromkeys((k,) + tuple(full_paths)):
        for (S, a, fr, post), j in zip(jobs, small):
            j[1].spp = a.spp
        full.append((trial(bpc, w, kk, walk)[0], kk))
    for (S, a, fr, post), j in zip(jobs, small):
        j[1].spp = max(4, int(round(a.spp * spp_scale)))
    if full:
        k = min(full)[1]
    set_tuning(TUNE_RENDER_BLOCKS_PER_CU, bpc)
    set_tuning(TUNE_RENDER_MIN_WAVES, w)
    set_tuning(TUNE_TRACE_VARIANT, walk[0])
    set_tuning(TUNE_LDS_STACK_ENTRIES, walk[1])
    set_tuning(TUNE_TRIS_PER_STEP, walk[2])
    set_tuning(TUNE_WALK_PHASES, walk[3] if len(walk) > 3 else 0)
    return {"blocks_per_cu": bpc, "min_waves": w, "paths_per_round": k, "walk_variant": walk[0],
            "lds_stack_entries": walk[1], "tris_per_step": walk[2], "walk_phases": walk[3] if len(walk) > 3 else 0,
            "seconds_scaled": t, "spp_scale": spp_scale, "tried": len(results),
            "full_spp_seconds": {str(kk): round(tt, 4) for tt, kk in full}}


def tune_adjoint(renderer, params, d_image, splits=(1, 2, 4, 8), log=None) -> dict:
    """Lanes per pixel of the adjoint kernel (each replays a share of the
    pixel's samples): times rm_adjoint of the renderer's last forward for each
    split and keeps the fastest."""
    from . import TUNE_ADJOINT_SPLIT
    renderer.forward_image(*params)
    results = []
    for s in splits:
        set_tuning(TUNE_ADJOINT_SPLIT, s)
        t = _time(lambda: renderer.adjoint(d_image, params[0].numel()), 2)
        results.append((t, s))
        if log:
            log(f"[autotune] adjoint split {s}: {t * 1e3:.1f} ms")
    t, s = min(results)
    set_tuning(TUNE_ADJOINT_SPLIT, s)
    return {"split": s, "seconds": t, "tried": len(results)}


_KEYS = {"trace_variant": TUNE_TRACE_VARIANT, "persistent_blocks_per_cu": TUNE_PERSISTENT_BLOCKS_PER_CU,
         "fetch_min": TUNE_FETCH_MIN, "ray_sort": TUNE_RAY_SORT, "render_blocks_per_cu": TUNE_RENDER_BLOCKS_PER_CU,
         "render_min_waves": TUNE_RENDER_MIN_WAVES, "lds_stack_entries": TUNE_LDS_STACK_ENTRIES,
         "tris_per_step": TUNE_TRIS_PER_STEP, "nee_split": TUNE_NEE_SPLIT, "walk_phases": TUNE_WALK_PHASES}


def current() -> dict:
    return {n: get_tuning(k) for n, k in _KEYS.items()}


def restore(saved: dict) -> None:
    """Sets the launch parameters back to a current() snapshot."""
    for n, v in saved.items():
        set_tuning(_KEYS[n], v)
