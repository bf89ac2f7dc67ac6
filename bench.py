"""Benchmark of the MI355X path-tracing hot path.

Headline (`value`): closest-hit Mrays/s on the raw-traversal workload -- a
1M-triangle random soup (centroids U[0,1]^3, vertices c + N(0,0.01)^3, the
reference median-split BVH) traced with two 8M-ray batches: 2828x2828 coherent
pinhole rays and 8,000,000 incoherent rays (origins U[0,1]^3, directions
uniform on the sphere); tmin 1e-4, tmax inf.  One step = one closest-hit
traversal of both batches, rays resident in HBM.

Also reported: any-hit Mrays/s of the same batches, the traversal kernel's
roofline (algorithmic bytes per ray priced in the reference traversal order:
24 + 8 + 24*box_tests + 8*node_visits + 36*triangle_tests), and the oracle's
CPU restatement of BVH::rayHit timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "closest-hit Mrays/s + path samples/s at 1080p; adjoint:forward time ratio"
HBM_PEAK_GBS = 8000.0
LIBRARY_DEFAULTS = {}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model


def host_threads(O) -> int:
    """Every host core this process may run on (the reference's thread pool
    uses them all, src/render.cpp:593-626)."""
    if hasattr(os, "sched_getaffinity"):
        return max(1, len(os.sched_getaffinity(0)))
    return max(1, O.hw_threads())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-tris", type=int, default=1000000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=4000000, help="rays per batch in the CPU sample")
    ap.add_argument("--no-render", action="store_true", help="skip the 1080p render leg")
    ap.add_argument("--render-frames", type=int, default=2, help="timed 1080p frames (after one warm-up)")
    ap.add_argument("--render-spp", type=int, default=0, help="override the 1080p workload's 256 spp")
    ap.add_argument("--cpu-row-stride", type=int, default=6, help="CPU render sample: every n-th row")
    ap.add_argument("--cpu-col-stride", type=int, default=8, help="CPU render sample: every n-th column")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="accepted for older profile commands; the 1 GiB calibration copy always runs")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch parameters")
    ap.add_argument("--render-tuning", default="",
                    help="fix the 1080p render launch instead of autotuning it: "
                         "'blocks_per_cu,min_waves,paths_per_round,walk_variant,lds_stack_entries,tris_per_step[,walk_phases]'")
    ap.add_argument("--trace", default="", help="fix the traversal launch: 'variant,blocks_per_cu,fetch_min,ray_sort[,lds_stack_entries[,tris_per_step[,walk_phases]]]'")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-serving leg")
    ap.add_argument("--batch-tuning", default="",
                    help="fix the batch's launch instead of autotuning it (fields as --render-tuning)")
    ap.add_argument("--batch-jobs", type=int, default=64, help="jobs in the batched-serving leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rendering leg")
    ap.add_argument("--adam-steps", type=int, default=200, help="inverse-rendering Adam steps (0 skips the loop)")
    ap.add_argument("--no-denoise", action="store_true", help="skip the denoise leg (on the 1080p render)")
    ap.add_argument("--denoise-reps", type=int, default=5, help="timed denoise runs (after one warm-up)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch

    import raym0nade_amd as rm

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    from raym0nade_amd import autotune as _at
    global LIBRARY_DEFAULTS
    LIBRARY_DEFAULTS = _at.current()  # before any tuning: the parameters a plain launch uses
    log(f"[bench] generating soup ({args.n_tris} tris) and rays")
    g = rm.generate(1, 1, n_tris=args.n_tris)
    t0 = time.time()
    S = rm.Scene(g)
    log(f"[bench] scene: build {S.info.build_seconds:.2f}s upload {S.info.upload_seconds:.2f}s "
        f"nodes {S.info.n_nodes} depth {S.info.depth}")
    cam = g.args
    n_coh = cam.width * cam.height
    n_inc = 8000000
    r_coh_h = rm.rays(0, n_coh, seed=1, cam=cam)
    r_inc_h = rm.rays(1, n_inc, seed=1)
    r_coh = torch.from_numpy(r_coh_h).to(dev)
    r_inc = torch.from_numpy(r_inc_h).to(dev)
    bufs = {}
    for name, r in (("coh", r_coh), ("inc", r_inc)):
        n = r.shape[0]
        bufs[name] = (torch.empty(n, dtype=torch.float32, device=dev), torch.empty(n, dtype=torch.int32, device=dev),
                      torch.empty(n, dtype=torch.uint8, device=dev))

    # algorithmic bytes per batch from the instrumented traversal (untimed)
    algo = {}
    for name, r in (("coh", r_coh), ("inc", r_inc)):
        for any_hit in (False, True):
            c = S.trace_count(r, any_hit=any_hit).cpu().numpy().astype(np.float64)
            interior, leaves, boxes, tris = c
            n = r.shape[0]
            out_b = 1 if any_hit else 8
            algo[(name, any_hit)] = n * (24 + out_b) + 24 * boxes + 8 * (interior + leaves) + 36 * tris
    log("[bench] algorithmic bytes/ray:", {f"{k[0]}{'_any' if k[1] else ''}": round(v / (n_coh if k[0] == 'coh' else n_inc), 1)
                                           for k, v in algo.items()})
    tuned = None
    if args.trace:
        fields = [int(x) for x in args.trace.split(",")]
        v, b, f, s, e, p, ph = (fields + [32, 0, 0][len(fields) - 4:])[:7]
        for k, val in ((rm.TUNE_TRACE_VARIANT, v), (rm.TUNE_PERSISTENT_BLOCKS_PER_CU, b), (rm.TUNE_FETCH_MIN, f),
                       (rm.TUNE_RAY_SORT, s), (rm.TUNE_LDS_STACK_ENTRIES, e), (rm.TUNE_TRIS_PER_STEP, p),
                       (rm.TUNE_WALK_PHASES, ph)):
            rm.set_tuning(k, val)
        tuned = {"variant": v, "blocks_per_cu": b, "fetch_min": f, "ray_sort": s, "lds_stack_entries": e,
                 "tris_per_step": p, "walk_phases": ph, "fixed": True}
    elif not args.no_autotune:
        from raym0nade_amd import autotune
        tuned = autotune.tune_trace(S, [r_coh, r_inc], reps=2, log=log)
        log(f"[bench] autotuned traversal: {tuned}")

    def step(times=None):
        for name, r in (("coh", r_coh), ("inc", r_inc)):
            t_out, id_out, _ = bufs[name]
            if times is not None:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
            S.trace_closest(r, t_out=t_out, id_out=id_out, stream=stream)
            if times is not None:
                e1.record(stream)
                times[name].append((e0, e1))

    def any_step(times):
        for name, r in (("coh", r_coh), ("inc", r_inc)):
            occ = bufs[name][2]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            S.trace_any(r, occ=occ, stream=stream)
            e1.record(stream)
            times[name].append((e0, e1))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times = {"coh": [], "inc": []}
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        step(times)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    k_ms = {k: np.array([a.elapsed_time(b) for a, b in v]) for k, v in times.items()}
    any_times = {"coh": [], "inc": []}
    for _ in range(max(3, args.steps // 2)):
        any_step(any_times)
    torch.cuda.synchronize()
    a_ms = {k: np.array([a.elapsed_time(b) for a, b in v]) for k, v in any_times.items()}

    n_rays = n_coh + n_inc
    ms_per_step = wall * 1e3 / args.steps
    kern_ms = k_ms["coh"].mean() + k_ms["inc"].mean()
    value = n_rays / (kern_ms * 1e-3) / 1e6
    mrays = {f"closest_{k}": (n_coh if k == "coh" else n_inc) / (k_ms[k].mean() * 1e-3) / 1e6 for k in k_ms}
    mrays.update({f"any_{k}": (n_coh if k == "coh" else n_inc) / (a_ms[k].mean() * 1e-3) / 1e6 for k in a_ms})
    gbs = {f"closest_{k}": algo[(k, False)] / (k_ms[k].mean() * 1e-3) / 1e9 for k in k_ms}
    gbs.update({f"any_{k}": algo[(k, True)] / (a_ms[k].mean() * 1e-3) / 1e9 for k in a_ms})
    # dominant kernel: the closest-hit traversal kernel over both batches, timed
    # alone by the library's own HIP events on the trace stream (a call also
    # runs the ray-reordering kernels when the batch is incoherent)
    rm.set_tuning(rm.TUNE_KERNEL_TIMING, 1)
    kt = {"coh": [], "inc": []}
    rt = {"coh": [], "inc": []}
    for _ in range(max(3, args.steps // 2)):
        for name, r in (("coh", r_coh), ("inc", r_inc)):
            t_out, id_out, _ = bufs[name]
            S.trace_closest(r, t_out=t_out, id_out=id_out, stream=stream)
            reo, trav = rm.last_trace_ms()
            kt[name].append(trav)
            rt[name].append(reo)
    rm.set_tuning(rm.TUNE_KERNEL_TIMING, 0)
    bytes_per_launch = (algo[("coh", False)] + algo[("inc", False)]) / 2
    dur_per_launch = (np.mean(kt["coh"]) + np.mean(kt["inc"])) / 2 * 1e-3
    achieved = bytes_per_launch / dur_per_launch / 1e9
    kernel_name = {0: "k_trace_closest", 2: "k_trace_persistent",
                   3: "k_trace_persistent_lds"}[rm.get_tuning(rm.TUNE_TRACE_VARIANT)]
    # HBM bytes per launch from the committed PMC passes of this kernel (the
    # launch parameters they were taken at are reported beside them)
    traffic, traffic_launch = None, None
    tf = os.path.join(ROOT, "profiles", f"traffic_{kernel_name}.json")
    if os.path.exists(tf):
        try:
            tj = json.load(open(tf))
            traffic, traffic_launch = tj.get("hbm_bytes_per_launch"), tj.get("launch")
        except (OSError, ValueError):
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_launch": traffic_launch,
                "kernel": kernel_name, "algorithmic_bytes_per_launch": round(bytes_per_launch),
                "avg_launch_ms": round(dur_per_launch * 1e3, 4),
                "launch_ms": {k: round(float(np.mean(v)), 4) for k, v in kt.items()},
                "reorder_ms": {k: round(float(np.mean(v)), 4) for k, v in rt.items()}}

    # lane utilisation of the persistent LDS walk (untimed calls with the
    # instrumented kernel; the results are unchanged)
    if rm.get_tuning(rm.TUNE_TRACE_VARIANT) == 3:
        util = {}
        for name, r in (("coh", r_coh), ("inc", r_inc)):
            t_out, id_out, _ = bufs[name]
            rm.set_tuning(rm.TUNE_TRACE_UTIL, 1)
            S.trace_closest(r, t_out=t_out, id_out=id_out)
            util[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in rm.last_trace_util().items()}
            rm.set_tuning(rm.TUNE_TRACE_UTIL, 0)
        roofline["lane_utilisation"] = util
        log("[bench] lane utilisation:", {k: (v["interior_util"], v["leaf_util"], v["util"]) for k, v in util.items()})

    # measured streaming bandwidth: a plain 1 GiB device copy (1 GiB read + 1 GiB
    # written per launch); also the known-size kernel that scales the
    # FETCH_SIZE / WRITE_SIZE counters
    x = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(5):
        y.copy_(x)
    e1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 2 * (1 << 30) * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del x, y
    roofline["measured_copy_gbs"] = round(copy_gbs, 1)
    roofline["frac_of_measured"] = round(achieved / copy_gbs, 4)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        threads = host_threads(O)
        log(f"[bench] CPU baseline: oracle BVH build + BVH::rayHit on {threads} threads")
        OS = O.OracleScene(g.desc)
        k = max(1, min(args.cpu_sample, n_coh, n_inc))
        sc = r_coh_h[:: max(1, n_coh // k)][:k]
        si = r_inc_h[:k]
        OS.closest(si[:10000], threads=threads)  # warm-up
        reps = []
        for _ in range(3):
            c0 = time.perf_counter()
            OS.closest(sc, threads=threads)
            OS.closest(si, threads=threads)
            reps.append(time.perf_counter() - c0)
        cdt = float(np.median(reps))
        cpu = {"value": round((sc.shape[0] + si.shape[0]) / cdt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
               "kind": "port", "host_cpus": os.cpu_count(),
               "sample": f"{sc.shape[0]} coherent (every {max(1, n_coh // k)}th) + {si.shape[0]} incoherent rays of "
                         f"the same batches, closest hit, median of 3 runs of {cdt:.1f} s; CPU {cpu_info()}; "
                         f"oracle restatement of BVH::rayHit, g++ -O2 -ffp-contract=off"}
        log(f"[bench] CPU: {cpu['value']} Mrays/s")

    render = None
    if not args.no_render:
        render = render_leg(args, torch, rm, dev, stream)
    batch = None
    if not args.no_batch:
        batch = batch_leg(args, torch, rm)
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(args, torch, rm)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "raw traversal: 1M-tri random soup, reference median-split BVH; 2828x2828 coherent + "
                               "8M incoherent rays, closest hit, tmin 1e-4, tmax inf",
                   "n_tris": args.n_tris, "rays_per_step": n_rays, "parallelism": "single-gpu"},
        "mrays_per_s": {k: round(v, 1) for k, v in mrays.items()},
        "algorithmic_gbs": {k: round(v, 1) for k, v in gbs.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "autotune": tuned,
        "render_1080p": render,
        "batch": batch,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


def batch_leg(args, torch, rm):
    """Batched serving: the job set of raym0nade_amd.workloads (seeded Cornell
    and displaced-sphere variants, 128^2..1024^2, spp {16, 64, 256}, depth
    8..16) rendered (a) as one rm_render_batch hipGraph launch and (b) job by
    job with rm_render; both with FXAA post.  Both produce the same samples
    per job (equal RMSE by construction), so samples/s compare directly."""
    from raym0nade_amd import autotune, workloads
    specs = workloads.batch_specs(args.batch_jobs, seed=1)
    log(f"[bench] batch: building {len(specs)} job scenes")
    built = [workloads.build_job(s) for s in specs]
    frames = [rm.Frame(a.width, a.height) for _, _, a in built]
    jobs = [(S, a, fr, 63 | 512) for (_, S, a), fr in zip(built, frames)]
    total = sum(s.samples for s in specs)
    tris = [S.info.n_tris for _, S, _ in built]
    log(f"[bench] batch: {total / 1e6:.0f}M samples, {min(tris)}..{max(tris)} triangles per scene")
    tuned = None
    if args.batch_tuning:
        fields = [int(x) for x in args.batch_tuning.split(",")]
        b, w, k, v, e, p, ph = (fields + [0])[:7]
        for key, val in ((rm.TUNE_RENDER_BLOCKS_PER_CU, b), (rm.TUNE_RENDER_MIN_WAVES, w), (rm.TUNE_TRACE_VARIANT, v),
                         (rm.TUNE_LDS_STACK_ENTRIES, e), (rm.TUNE_TRIS_PER_STEP, p), (rm.TUNE_WALK_PHASES, ph)):
            rm.set_tuning(key, val)
        for _, a, _, _ in jobs:
            a.paths_per_round = k
        tuned = {"blocks_per_cu": b, "min_waves": w, "paths_per_round": k, "walk_variant": v,
                 "lds_stack_entries": e, "tris_per_step": p, "walk_phases": ph, "fixed": True}
        log(f"[bench] batch launch fixed: {tuned}")
    elif not args.no_autotune:
        tuned = autotune.tune_render(jobs, log=log)
        for _, a, _, _ in jobs:
            a.paths_per_round = tuned["paths_per_round"]
        log(f"[bench] batch autotune: {tuned}")
    rm.render_batch(jobs[:2], use_graph=True)  # warm-up

    def size_buffers():
        # every frame's own work buffers exist before a sequential timing, at
        # that timing's launch parameters (a low-spp render sizes them: the
        # sizes depend on the paths per round, not on spp)
        for S, a, fr, post in jobs:
            b = rm.RmRenderArgs.from_buffer_copy(a)
            b.spp = 4
            fr.render(S, b)
        torch.cuda.synchronize()

    size_buffers()
    t0 = time.perf_counter()
    rm.render_batch(jobs, use_graph=True)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    log(f"[bench] batch: one graph launch {tb:.2f} s")
    def sequential(per_job=None):
        t0 = time.perf_counter()
        for S, a, fr, post in jobs:
            j0 = time.perf_counter()
            fr.render(S, a)
            fr.postprocess(post, a.exposure)
            if per_job is not None:
                torch.cuda.synchronize()
                per_job.append(time.perf_counter() - j0)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    ts_tuned = sequential()
    log(f"[bench] batch: sequential rm_render with the batch's launch parameters {ts_tuned:.2f} s")
    # plain launches: rm_render per job with the library's default parameters
    saved = autotune.current()
    autotune.restore(LIBRARY_DEFAULTS)
    for _, a, _, _ in jobs:
        a.paths_per_round = 0
    size_buffers()
    per_job = []
    ts = sequential(per_job)
    autotune.restore(saved)
    log(f"[bench] batch: sequential plain rm_render {ts:.2f} s")
    # where the time goes, job by job (plain launches): each job's samples/s
    # alone, and the share of the sequential time of the jobs under 20M samples
    table = [{"job": i, "config": sp.config, "tris": tris[i], "side": sp.side, "spp": sp.spp, "depth": sp.max_depth,
              "seconds": round(t, 4), "samples_per_s": round(sp.samples / t, 1)}
             for i, (sp, t) in enumerate(zip(specs, per_job))]
    small = [r for r, sp in zip(table, specs) if sp.samples < 20_000_000]
    small_s = sum(r["seconds"] for r in small)
    small_samples = sum(sp.samples for sp in specs if sp.samples < 20_000_000)
    return {"jobs": len(jobs), "samples": total, "batched_samples_per_s": round(total / tb, 1),
            "sequential_samples_per_s": round(total / ts, 1), "batched_over_sequential": round(ts / tb, 3),
            "sequential_tuned_samples_per_s": round(total / ts_tuned, 1),
            "batched_over_sequential_tuned": round(ts_tuned / tb, 3),
            "batched_s": round(tb, 3), "sequential_s": round(ts, 3), "sequential_tuned_s": round(ts_tuned, 3),
            "autotune": tuned, "tris_min": min(tris), "tris_max": max(tris),
            "library_defaults": LIBRARY_DEFAULTS, "per_job": table,
            "small_jobs": {"count": len(small), "seconds": round(small_s, 3), "samples": small_samples,
                           "share_of_sequential_time": round(small_s / max(1e-9, sum(per_job)), 3),
                           "share_of_samples": round(small_samples / total, 3)},
            "note": "sequential = one rm_render + rm_postprocess per job with the library's default launch "
                    "parameters (library_defaults, automatic paths per round); sequential_tuned = the same with "
                    "the batch's autotuned parameters; every frame's buffers exist beforehand"}


def adjoint_leg(args, torch, rm):
    """Differentiable rendering at 512^2, 128 spp on the textured Cornell box
    (five 256^2 fp32 albedo textures, per-wall GGX roughness): time of the
    forward render (rm_render + linear shade) and of the path-replay adjoint
    (rm_adjoint, gradients for every texel, roughness and emission).  The
    renders run at the library's default launch parameters, not at those an
    earlier leg tuned for its own workload."""
    from raym0nade_amd import autotune as _at
    from raym0nade_amd.diff import DiffRenderer
    _at.restore(LIBRARY_DEFAULTS)
    g = rm.generate(4, 1)
    R = DiffRenderer(g)
    tex, rough, emis = R.initial_params()
    img = R.forward_image(tex, rough, emis)
    d = torch.randn_like(img)
    R.adjoint(d, tex.numel())  # warm-up
    tuned = None
    if not args.no_autotune:
        from raym0nade_amd import autotune
        tuned = autotune.tune_adjoint(R, (tex, rough, emis), d, log=log)
    plain = rm.Frame(R.args.width, R.args.height)  # recording off: the plain forward render
    plain.render(R.scene, R.args)
    fw, rd, pr, ad = [], [], [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain.render(R.scene, R.args)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        R.frame.render(R.scene, R.args)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        R.forward_image(tex, rough, emis)  # parameter upload + recording render + linear shade
        t3 = time.perf_counter()
        R.adjoint(d, tex.numel())
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        pr.append(t1 - t0)
        rd.append(t2 - t1)
        fw.append(t3 - t2)
        ad.append(t4 - t3)
    p, r, f, a = (float(np.median(v)) for v in (pr, rd, fw, ad))
    a_args = R.args
    stride = a_args.spp - int(np.float32(a_args.spp) * np.float32(a_args.p_direct))
    out = {"plain_render_s": round(p, 4), "recording_render_s": round(r, 4), "forward_s": round(f, 4),
           "adjoint_s": round(a, 4), "adjoint_over_forward": round(a / p, 3),
           "adjoint_over_recording_forward": round(a / r, 3), "adjoint_over_forward_with_upload": round(a / f, 3),
           "record_bytes": int(a_args.width * a_args.height * stride * 48),
           "recompute_only": "refused by rm_adjoint (the two-replay form is switched off)",
           "width": a_args.width, "height": a_args.height, "spp": a_args.spp, "texels": int(tex.numel()),
           "split": rm.get_tuning(rm.TUNE_ADJOINT_SPLIT), "autotune": tuned,
           "note": "adjoint_over_forward divides by a plain rm_render (recording off); the adjoint needs the "
                   "recording render (48 B per indirect sample) and replays each sample's path once"}
    log(f"[bench] adjoint: plain render {p:.3f} s, recording render {r:.3f} s, forward {f:.3f} s, "
        f"adjoint {a:.3f} s, ratio {a / p:.2f}")
    if args.adam_steps > 0:
        from raym0nade_amd.inverse import inverse_render
        out["adam"] = inverse_render(g, steps=args.adam_steps, log=log)
    return out


# Algorithmic HBM bytes per pixel of the denoise path (four float4 radiance
# buffers; G-buffer position, shape normal, surface normal, material as 3
# floats each):
#   spatialClamp   rows: read rgb+Var 16, write blur 4; cols: read blur 4,
#                  read rgb+Var 16 (writes only for clamped pixels)  x 4 buffers
#   filterVar      read 16, write 16                                x 4 buffers
#   a-trous pass   read G-buffer 48 once + (read 16 + write 16)     x 4 buffers, x 5 passes
DENOISE_BYTES_PER_PX = 4 * (16 + 4 + 4 + 16) + 4 * (16 + 16) + 5 * (48 + 4 * 32)


def denoise_leg(args, torch, rm, fr, stream):
    """rm_spatial_clamp + rm_filter (Photo::spatialClamp + Photo::filter) on
    the rendered 1080p frame, its radiance restored before every run; HIP
    events on the launch stream bracket each call."""
    from raym0nade_amd import render_check as RC
    saved = [fr.download(b) for b in RC.RAD_ORDER]

    def restore():
        for b, r in zip(RC.RAD_ORDER, saved):
            fr.upload(b, r)

    clamp_ms, filt_ms = [], []
    for k in range(args.denoise_reps + 1):
        restore()
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record(stream)
        fr.spatial_clamp(stream=stream)
        e1.record(stream)
        fr.filter(stream=stream)
        e2.record(stream)
        torch.cuda.synchronize()
        if k > 0:
            clamp_ms.append(e0.elapsed_time(e1))
            filt_ms.append(e1.elapsed_time(e2))
    c, f = float(np.median(clamp_ms)), float(np.median(filt_ms))
    n = fr.width * fr.height
    out = {"mpix_per_s": round(n / ((c + f) * 1e-3) / 1e6, 1), "ms": round(c + f, 3), "clamp_ms": round(c, 3),
           "filter_ms": round(f, 3), "bytes_per_px": DENOISE_BYTES_PER_PX,
           "algorithmic_gbs": round(n * DENOISE_BYTES_PER_PX / ((c + f) * 1e-3) / 1e9, 1), "runs": len(filt_ms)}
    log(f"[bench] denoise 1080p: clamp {c:.3f} ms, filter {f:.3f} ms")
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        restore()
        gb, rad = RC.download_all(fr)
        h0, w0 = fr.height // 4, fr.width // 4
        gb, rad = gb[h0:h0 + fr.height // 2, w0:w0 + fr.width // 2], rad[:, h0:h0 + fr.height // 2, w0:w0 + fr.width // 2]
        c0 = time.perf_counter()
        O.denoise(gb, rad)
        cdt = time.perf_counter() - c0
        m = gb.shape[0] * gb.shape[1]
        out["cpu_baseline"] = {
            "value": round(m / cdt / 1e6, 4), "unit": "Mpixel/s", "cores": 4, "kind": "port",
            "sample": f"the centre {gb.shape[1]}x{gb.shape[0]} crop of the same frame, {cdt:.1f} s; oracle "
                      f"restatement of spatialClamp + filter, one std::thread per buffer as the reference runs it"}
        out["gpu_over_cpu"] = round(out["mpix_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


# the reference's export list (src/render.cpp:641-673) as (shade options,
# step run before it); the depth-of-field five run when CoC > 1e-4
EXPORTS = [(1, None), (1 | 512, None), (64, None), (128, None),
           (4 | 16, "clamp"), (4 | 32, None), (8 | 16, None), (8 | 32, None),
           (63, None), (63 | 256, None), (63 | 512, None), (63 | 768, None),
           (4 | 16, "filter"), (4 | 32, None), (8 | 16, None), (8 | 32, None),
           (63, None), (63 | 256, None), (63 | 512, None), (63 | 768, None),
           (1 | 1024, None), (63 | 1024, None), (63 | 1280, None), (63 | 1536, None), (63 | 1792, None)]


def exports_leg(args, torch, rm, fr, a, stream):
    """Bloom, depth of field and the whole export list of the reference's
    render command on the rendered 1080p frame (GPU time of the library
    calls; PNG encoding excluded).  Depth of field with CoC 8 focused at the
    median distance of the frame's hits."""
    from raym0nade_amd import render_check as RC
    saved = [fr.download(b) for b in RC.RAD_ORDER]
    pos = fr.download(rm.RM_BUF_POSITION)
    cam = np.array([a.position[k] for k in range(3)], np.float32)
    d = np.linalg.norm(pos - cam, axis=-1)
    focus = float(np.median(d[np.isfinite(d)]))
    coc = 8.0
    fr.set_dof(focus, coc, cam)

    def timed(fn, reps=3):
        ts = []
        for _ in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts[1:]))

    out = {"focus": round(focus, 3), "coc": coc}
    out["shade_ms"] = round(timed(lambda: fr.postprocess(63, a.exposure, stream=stream)), 3)
    out["bloom_ms"] = round(timed(lambda: fr.postprocess(63 | 256, a.exposure, stream=stream)), 3)
    out["dof_ms"] = round(timed(lambda: fr.postprocess(63 | 1024, a.exposure, stream=stream)), 3)

    def all_exports():
        for opt, step in EXPORTS:
            if step == "clamp":
                fr.spatial_clamp(stream=stream)
            elif step == "filter":
                fr.filter(stream=stream)
            fr.postprocess(opt, a.exposure, stream=stream)

    def restore():
        for b, r in zip(RC.RAD_ORDER, saved):
            fr.upload(b, r)

    ts = []
    for k in range(3):
        restore()
        torch.cuda.synchronize()
        c0 = time.perf_counter()
        all_exports()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - c0)
    out["all_exports_ms"] = round(1e3 * float(np.median(ts[1:])), 2)
    out["n_exports"] = len(EXPORTS)
    log(f"[bench] exports 1080p: bloom {out['bloom_ms']} ms, dof {out['dof_ms']} ms, "
        f"all {len(EXPORTS)} exports {out['all_exports_ms']} ms")
    restore()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        gb, rad = RC.download_all(fr)
        h0, w0 = fr.height // 4, fr.width // 4
        gb, rad = gb[h0:h0 + fr.height // 2, w0:w0 + fr.width // 2], rad[:, h0:h0 + fr.height // 2, w0:w0 + fr.width // 2]
        m = gb.shape[0] * gb.shape[1]
        cpu = {}
        for name, opt in (("bloom", 63 | 256), ("dof", 63 | 1024)):
            c0 = time.perf_counter()
            O.postprocess(gb, rad, opt, a.exposure, focus=focus, coc=coc, cam=cam)
            cpu[name] = time.perf_counter() - c0
        out["cpu_baseline"] = {
            "value": {k: round(m / v / 1e6, 3) for k, v in cpu.items()}, "unit": "Mpixel/s", "cores": 1,
            "kind": "port",
            "sample": f"the centre {gb.shape[1]}x{gb.shape[0]} crop of the same frame: shade + bloom + gamma "
                      f"{cpu['bloom']:.1f} s, shade + depth of field + gamma {cpu['dof']:.1f} s; oracle restatement, "
                      f"one thread as the reference"}
        out["gpu_over_cpu"] = {"bloom": round(n_px(fr) / (out["bloom_ms"] * 1e-3) / 1e6 / out["cpu_baseline"]["value"]["bloom"], 1),
                               "dof": round(n_px(fr) / (out["dof_ms"] * 1e-3) / 1e6 / out["cpu_baseline"]["value"]["dof"], 1)}
    return out


def n_px(fr):
    return fr.width * fr.height


def render_roofline(rm, torch, dev, S, a, st, counted, wall):
    """Total rays/s of a render and its roofline: every traced ray priced as
    the headline prices it (24 B in, 8 B out for closest hits and 1 B for
    shadow rays, 24 B per box test, 8 B per node visit, 36 B per triangle
    test), plus 60 B of vertex data per shading hit and 152 B per pixel
    (G-buffer and radiance).  Extension and shadow walks are counted by the
    counting render; the primary walks (camera rays, closest hit from 1e-4) by
    rm_trace_count on the same camera rays."""
    rays = torch.from_numpy(rm.rays(0, a.width * a.height, cam=a)).to(dev)
    pi, pl, pb, pt = (float(v) for v in S.trace_count(rays).cpu().numpy())
    del rays
    keys = ("primary_rays", "extension_rays", "shadow_rays", "hits", "paths")
    same = all(st[k] == counted[k] for k in keys)
    n_pri, n_ext, n_sh = st["primary_rays"], st["extension_rays"], st["shadow_rays"]
    walk_bytes = (24 * (pb + counted["walk_boxes"]) + 8 * (pi + pl + counted["walk_interior"] + counted["walk_leaves"])
                  + 36 * (pt + counted["walk_tris"]))
    ray_bytes = 32 * (n_pri + n_ext) + 25 * n_sh
    other = 60 * st["hits"] + 152 * a.width * a.height
    total = walk_bytes + ray_bytes + other
    achieved = total / wall / 1e9
    rays_total = n_pri + n_ext + n_sh
    log(f"[bench] 1080p rays: {rays_total / 1e9:.2f} G ({n_pri} primary, {n_ext} extension, {n_sh} shadow), "
        f"{rays_total / wall / 1e6:.0f} Mrays/s; algorithmic {total / 1e9:.1f} GB, {achieved:.0f} GB/s")
    return {"rays_per_s": round(rays_total / wall, 1), "rays": {"primary": n_pri, "extension": n_ext, "shadow": n_sh},
            "stats": st, "counting_render_matches": same,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                         "kernel": "whole render (k_primary ... k_finish, FXAA post)",
                         "algorithmic_bytes_per_frame": round(total),
                         "bytes": {"walks": round(walk_bytes), "rays": round(ray_bytes), "vertices_pixels": round(other)},
                         "walk_counts": {"primary": [pi, pl, pb, pt],
                                         "extension_shadow": [counted["walk_interior"], counted["walk_leaves"],
                                                              counted["walk_boxes"], counted["walk_tris"]]}}}


def render_leg(args, torch, rm, dev, stream):
    """Path samples/s of the 1080p workload: 64 displaced geodesic spheres
    (10,368,002 triangles) under a 2048x1024 fp32 sky, 1920x1080, 256 spp,
    max depth 16, FXAA post.  Wall time of rm_render + rm_postprocess with the
    scene resident in HBM."""
    log("[bench] 1080p workload: generating 10.4M-triangle scene")
    g = rm.generate(2, 1)
    S = rm.Scene(g)
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    if args.render_spp:
        a.spp = args.render_spp
    log(f"[bench] scene: {S.info.n_tris} tris, build {S.info.build_seconds:.1f}s, "
        f"{S.info.device_bytes / 2**30:.2f} GiB in HBM")
    fr = rm.Frame(a.width, a.height)
    tuned = None
    if args.render_tuning:
        fields = [int(x) for x in args.render_tuning.split(",")]
        b, w, k, v, e, p, ph = (fields + [0])[:7]
        for key, val in ((rm.TUNE_RENDER_BLOCKS_PER_CU, b), (rm.TUNE_RENDER_MIN_WAVES, w), (rm.TUNE_TRACE_VARIANT, v),
                         (rm.TUNE_LDS_STACK_ENTRIES, e), (rm.TUNE_TRIS_PER_STEP, p), (rm.TUNE_WALK_PHASES, ph)):
            rm.set_tuning(key, val)
        a.paths_per_round = k
        tuned = {"blocks_per_cu": b, "min_waves": w, "paths_per_round": k, "walk_variant": v,
                 "lds_stack_entries": e, "tris_per_step": p, "walk_phases": ph, "fixed": True}
    elif not args.no_autotune:
        from raym0nade_amd import autotune
        tuned = autotune.tune_render([(S, a, fr, 63 | 512)], spp_scale=1 / 4, blocks=(4, 8), waves=(1, 3),
                                     paths_per_round=(0, 4, 8), full_paths=(8, 16), log=log)
        a.paths_per_round = tuned["paths_per_round"]
    times = []
    counted = None
    for k in range(args.render_frames + 1):
        # the untimed warm-up frame is a counting render (same image): the
        # node visits and tests of its extension and shadow walks price them
        rm.set_tuning(rm.TUNE_RENDER_COUNT, 1 if k == 0 else 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fr.render(S, a, stream=stream)
        fr.postprocess(63 | 512, a.exposure, stream=stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log(f"[bench] frame {k}: {dt:.3f} s" + (" (warm-up, counting walks)" if k == 0 else ""))
        if k == 0:
            counted = fr.stats()
        else:
            times.append(dt)
    rm.set_tuning(rm.TUNE_RENDER_COUNT, 0)
    wall = float(np.median(times))
    samples = a.width * a.height * a.spp
    st = fr.stats()
    out = {"path_samples_per_s": round(samples / wall, 1), "frame_s": round(wall, 4), "frames": len(times),
           "width": a.width, "height": a.height, "spp": a.spp, "max_depth": a.max_depth, "n_tris": S.info.n_tris,
           "autotune": tuned}
    out.update(render_roofline(rm, torch, dev, S, a, st, counted, wall))
    if not args.no_denoise:
        out["denoise"] = denoise_leg(args, torch, rm, fr, stream)
        out["exports"] = exports_leg(args, torch, rm, fr, a, stream)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        threads = host_threads(O)
        log("[bench] CPU baseline: oracle scene build (single thread, reference builder)")
        c0 = time.perf_counter()
        OS = O.OracleScene(g.desc)
        log(f"[bench] oracle scene built in {time.perf_counter() - c0:.1f} s")
        rs, cs = args.cpu_row_stride, args.cpu_col_stride
        pix = [y * a.width + x for y in range(rs // 2, a.height, rs) for x in range(cs // 2, a.width, cs)]
        c0 = time.perf_counter()
        O.render_pixels(OS, a, pix, threads=threads)
        cdt = time.perf_counter() - c0
        cpu_samples = len(pix) * a.spp
        out["cpu_baseline"] = {
            "value": round(cpu_samples / cdt, 1), "unit": "path samples/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(),
            "sample": f"{len(pix)} pixels on a regular grid (every {rs}th row x every {cs}th column) of the same "
                      f"frame at the full {a.spp} spp, {cdt:.1f} s; oracle restatement of renderPixel on a "
                      f"row-pool of std::threads"}
        out["gpu_over_cpu"] = round(out["path_samples_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


if __name__ == "__main__":
    main()
