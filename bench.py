"""Benchmark: wavefront path tracing at 1920x1080 on one MI355X.

Workload (the plain method at scale): 2,097,152-triangle fBm terrain plus 200
subdivision-4 icospheres (3,121,152 triangles), 2048x1024 log-normal HDR
environment with a 0.5-degree sun, importance-sampled via the 2D
(marginal/conditional) CDF; 1920x1080, depth 8, FP32.  One step = one
wavefront pass: S samples of every pixel of the frame (S chosen by the
autotuner, up to 128; the image equals S one-sample frames bit for bit)
through raygen -> {closest hit -> shade -> shadow any-hit} x 8 ->
accumulate, replayed from one captured hipGraph.  value = samples per pixel
per second (frames/s at 1080p); scene, BVH and queues are resident in HBM
before timing.

Also reported: traversal microbenchmark (10M-triangle soup, 16M incoherent
rays, closest and any hit, Mrays/s), the roofline of the dominant kernel
(trace_closest: algorithmic bytes per launch / mean launch time, HIP events),
and the oracle's FP32 CPU restatement timed on a bounded sample of the same
workload as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
RAY_BYTES = 32 + 8          # ray in (o, d) + hit out (id, t)
NODE_BYTES = 128            # one wide node line
TRI_BYTES = 48              # one triangle record


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traversal_bytes(rays, nodes, tris):
    return rays * RAY_BYTES + nodes * NODE_BYTES + tris * TRI_BYTES


TRAFFIC_SUMMARY = "render_c2_1080p_hbm_traffic.md"
COUNTER_SUMMARY = "render_c2_1080p_counters.md"


def summary_row(name, kernel):
    """The row of `kernel` in the first markdown table of profiles/`name`
    whose header starts with "Kernel", as {column header: float or text};
    {} when the summary or the row is absent."""
    path = os.path.join(ROOT, "profiles", name)
    header = None
    try:
        for line in open(path):
            if not line.startswith("|"):
                header = None
                continue
            cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
            if cells and cells[0] == "Kernel":
                header = cells
            elif header and cells[0] == kernel:
                row = {}
                for h, c in zip(header[1:], cells[1:]):
                    try:
                        row[h] = float(c)
                    except ValueError:
                        row[h] = c
                return row
    except OSError:
        return {}
    return {}


def measured_traffic(kernel="k_trace_closest"):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE runs, scaled by a copy kernel of known
    size); None when the summary is absent."""
    v = summary_row(TRAFFIC_SUMMARY, kernel).get("HBM MB/dispatch")
    return v * 1e6 if isinstance(v, float) else None


def measured_counters(kernel):
    """Derived PMC figures of `kernel` (lanes per VALU instruction, L2 hit
    rate, ...) from the committed counter summary."""
    return {k: v for k, v in summary_row(COUNTER_SUMMARY, kernel).items() if isinstance(v, float)}


def apply_shapes(ctx, shapes):
    """Launch shapes {class: {waves_per_cu, refill_min[, stack_lds,
    waves_per_block, leaf_phase16, node_width]}} (an autotune record) onto a
    context."""
    for cls, sh in shapes.items():
        ctx.set_tuning(cls, sh["waves_per_cu"], sh["refill_min"], sh.get("stack_lds"), sh.get("waves_per_block"),
                       sh.get("leaf_phase16"), sh.get("node_width"))


def run_render(args, pike_amd, scenes):
    import numpy as np

    t0 = time.time()
    data = scenes.config2(seed=2, cells=args.cells, instanced=args.instanced)
    log(f"config2 scene: {data.baked_triangles} triangles, {len(data.instances)} instances "
        f"({time.time() - t0:.1f}s)")
    t0 = time.time()
    bvh = pike_amd.BVH.for_scene(data, threads=min(16, os.cpu_count() or 1), spatial_splits=args.bvh == "sbvh",
                                 sah_collapse=args.collapse == "sah")
    log(f"BVH ({args.bvh}{', two-level' if args.instanced else ''}): {bvh.info()} ({time.time() - t0:.1f}s)")
    ctx = pike_amd.Context(0)
    scene = pike_amd.Scene(ctx, data, bvh)
    cam = scenes.normalized_camera(data.camera, bvh.normalizer())
    W, H, depth = args.width, args.height, args.depth
    frame = pike_amd.Frame(ctx, W, H)
    frame.set_camera(cam)
    if args.spp_per_pass > 1:
        frame.set_samples_per_pass(args.spp_per_pass)
    tuned = None
    if args.tuning:
        cfg = json.loads(args.tuning)
        if "samples_per_pass" in cfg:
            args.spp_per_pass = int(cfg["samples_per_pass"])
            frame.set_samples_per_pass(args.spp_per_pass)
        apply_shapes(ctx, {k: v for k, v in cfg.get("shapes", cfg).items() if isinstance(v, dict)})
        if args.scheduler == "auto" and "scheduler" in cfg:
            ctx.set_scheduler(cfg["scheduler"])
        tuned = dict(shapes=ctx.tuning(), samples_per_pass=frame.samples_per_pass, source="--tuning")
    elif not args.no_autotune:
        t0 = time.time()
        s_before, s_after = frame.autotune_samples(scene, max_depth=depth, max_samples=args.max_spp_per_pass)
        args.spp_per_pass = frame.samples_per_pass
        log(f"autotune samples per pass: {s_before:.2f} -> {s_after:.2f} ms/sample at {args.spp_per_pass}")
        before, after = frame.autotune(scene, max_depth=depth, frames=2)
        tuned = dict(ms_before=before, ms_after=after, shapes=ctx.tuning(), samples_per_pass=args.spp_per_pass,
                     ms_per_sample_s1=s_before, ms_per_sample_chosen=s_after, seconds=time.time() - t0)
        log(f"autotune: traversal {before:.2f} -> {after:.2f} ms/pass, {ctx.tuning()} ({tuned['seconds']:.1f}s)")
        if args.scheduler == "auto":
            sc0, sc1 = frame.autotune_scheduler(scene, max_depth=depth, samples=2 * max(1, args.spp_per_pass))
            log(f"autotune scheduler: {ctx.scheduler()} ({sc0:.3f} -> {sc1:.3f} ms/sample)")
    if args.scheduler != "auto":
        ctx.set_scheduler(args.scheduler)
    if tuned is not None:
        tuned["scheduler"] = ctx.scheduler()
    # One step = one whole wavefront pass: S samples of every pixel (S = the
    # frame's samples per pass), so every timed pass is a full one.
    S = max(1, frame.samples_per_pass)
    # warmup (also instantiates the graph)
    frame.render(scene, max_depth=depth, first_sample=0, num_frames=max(1, args.warmup) * S, use_graph=True)
    ctx.sync()
    frame.reset()
    ctx.sync()
    samples = args.steps * S
    t0 = time.perf_counter()
    frame.render(scene, max_depth=depth, first_sample=0, num_frames=samples, use_graph=True)
    ctx.sync()
    dt = time.perf_counter() - t0
    cnt = frame.counters()
    img = frame.read_image()
    assert np.isfinite(img).all(), "non-finite pixels"
    spp_s = samples / dt
    rays_total = cnt["rays_traced"] + cnt["shadow_rays"]
    log(f"render: {args.steps} passes of {S} samples per pixel ({samples} spp) in {dt:.3f}s = {spp_s:.2f} spp/s, "
        f"{rays_total / dt / 1e6:.0f} Mrays/s, overflow={cnt['stack_overflow']}")
    # the first sample alone (a one-sample pass), for the CPU leg's
    # full-size comparison with the oracle's first frame
    first_frame = None
    if not args.no_cpu:
        chk = pike_amd.Frame(ctx, W, H)
        chk.set_camera(cam)
        chk.render(scene, max_depth=depth, first_sample=0, num_frames=1, use_graph=False)
        first_frame = chk.read_image()
        del chk
    # kernel-level profile: counted run for algorithmic bytes, uncounted for
    # time, over whole passes
    pf = args.profile_frames
    if pf > 0:
        sp = max(1, args.spp_per_pass)
        pf = -(-pf // sp) * sp
    if pf <= 0:
        return dict(spp_s=spp_s, ms_per_step=dt / args.steps * 1e3, rays_per_s=rays_total / dt, counters=cnt,
                    kernel_ms=None, profile_frames=0, profile_counters=None,
                    roofline=dict(bound="hbm", achieved=None, peak=HBM_PEAK_GBS, unit="GB/s", frac=None,
                                  traffic=None, bytes_per_launch=None, launch_ms=None),
                    samples_per_step=S, first_frame=first_frame,
                    triangles=data.baked_triangles, autotune=tuned, scheduler=ctx.scheduler()), data, bvh, cam
    frame.reset()
    frame.count_nodes(True)
    frame.render(scene, max_depth=depth, first_sample=0, num_frames=pf, use_graph=False)
    cc = frame.counters()
    frame.count_nodes(False)
    frame.reset()
    ms = frame.profile(scene, max_depth=depth, first_sample=0, num_frames=pf)
    launches = -(-pf // max(1, args.spp_per_pass)) * depth
    bytes_launch = traversal_bytes(cc["rays_traced"], cc["nodes_visited"], cc["tris_tested"]) / launches
    t_launch = ms["trace_closest"] / launches * 1e-3
    achieved = bytes_launch / t_launch / 1e9
    log(f"profile ({pf} frames): {ms}; per trace_closest launch {bytes_launch / 1e6:.1f} MB in "
        f"{t_launch * 1e3:.3f} ms -> {achieved:.0f} GB/s")
    traffic = measured_traffic()
    pmc = measured_counters("k_trace_closest")
    roof = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                traffic=traffic, kernel="k_trace_closest", bytes_per_launch=bytes_launch, launch_ms=t_launch * 1e3,
                # `achieved` prices the kernel's algorithmic node/triangle/ray
                # bytes, most of which L2 and MALL serve; the DRAM figures
                # divide the PMC-measured HBM bytes per launch by this run's
                # launch time.  The kernel's limiter is VALU issue with idle
                # lanes (divergent traversal), read from the PMC summary.
                achieved_basis="algorithmic bytes per launch (L2/MALL hits included)",
                traffic_source=f"profiles/{TRAFFIC_SUMMARY} (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                dram_gbs=traffic / t_launch / 1e9 if traffic else None,
                dram_frac=traffic / t_launch / 1e9 / HBM_PEAK_GBS if traffic else None,
                limiter="valu_issue (divergent lanes), not HBM",
                valu_lanes_active=pmc.get("lanes active per VALU instruction"),
                l2_hit_rate=pmc.get("L2 hit rate"),
                lds_bank_conflict_frac=pmc.get("LDS bank-conflict cycles / LDS instructions"),
                counters_source=f"profiles/{COUNTER_SUMMARY}" if pmc else None)
    result = dict(
        spp_s=spp_s, ms_per_step=dt / args.steps * 1e3, rays_per_s=rays_total / dt, counters=cnt,
        kernel_ms=ms, profile_frames=pf, profile_counters=cc, roofline=roof, samples_per_step=S,
        first_frame=first_frame, triangles=data.baked_triangles, autotune=tuned, scheduler=ctx.scheduler())
    return result, data, bvh, cam


def run_instanced(args, pike_amd, scenes, ctx, spp_per_pass):
    """The 1080p workload with its 200 icospheres as placements of one mesh
    (two-level BVH) next to the baked layout's sizes: build time, layout
    bytes and spp/s with the render leg's launch shapes and pass size."""
    import numpy as np

    data = scenes.config2(seed=2, cells=args.cells, instanced=True)
    threads = min(16, os.cpu_count() or 1)
    t0 = time.time()
    bvh = pike_amd.BVH.for_scene(data, threads=threads, sah_collapse=args.collapse == "sah")
    build_s = time.time() - t0
    info = bvh.info()
    scene = pike_amd.Scene(ctx, data, bvh)
    frame = pike_amd.Frame(ctx, args.width, args.height)
    frame.set_camera(scenes.normalized_camera(data.camera, bvh.normalizer()))
    while True:  # the render leg's pass size, halved while its queues do not fit
        try:
            frame.set_samples_per_pass(spp_per_pass)
            break
        except pike_amd.PikeError as e:
            if e.code != pike_amd._ffi.PK_E_OOM or spp_per_pass == 1:
                raise
            spp_per_pass //= 2
    if args.scheduler == "auto":
        frame.autotune_scheduler(scene, max_depth=args.depth, samples=2 * spp_per_pass)
    else:
        ctx.set_scheduler(args.scheduler)
    frame.render(scene, max_depth=args.depth, first_sample=0, num_frames=spp_per_pass, use_graph=True)
    ctx.sync()
    frame.reset()
    ctx.sync()
    steps = args.steps * spp_per_pass  # whole passes, as the headline leg
    t0 = time.perf_counter()
    frame.render(scene, max_depth=args.depth, first_sample=0, num_frames=steps, use_graph=True)
    ctx.sync()
    dt = time.perf_counter() - t0
    assert np.isfinite(frame.read_image()).all(), "non-finite pixels"
    out = dict(spp_s=steps / dt, steps=steps, build_s=build_s, placements=len(data.instances),
               baked_triangles=data.baked_triangles, wide_nodes=info["wide_nodes"],
               triangle_records=info["leaf_triangles"], scheduler=ctx.scheduler(),
               layout_mb=(info["wide_nodes"] * 128 + info["leaf_triangles"] * 48) / 1e6,
               overflow=frame.counters()["stack_overflow"])
    log(f"instanced c2: {out}")
    for obj in (frame, scene, bvh):
        obj.close()
    return out


def run_batching(args, pike_amd, scenes, ctx):
    """Config 4: 32 scenes x 64 views (aggregate samples/s), plain per-view
    renders (one sample per pass, and multi-sample passes) vs one batched
    graph launch, each under both pass schedulers (per-bounce wavefront
    kernels, path megakernel; the faster is reported); the images are
    identical (checked on every view), so the comparison is at equal error."""
    import numpy as np

    t0 = time.time()
    sc, views = scenes.config4(seed=4)
    threads = min(16, os.cpu_count() or 1)
    bvhs = [pike_amd.BVH(s.positions, s.triangles, threads=threads, spatial_splits=args.bvh == "sbvh",
                         sah_collapse=args.collapse == "sah") for s in sc]
    per_scene = [pike_amd.Scene(ctx, s, b) for s, b in zip(sc, bvhs)]
    merged = pike_amd.BatchScene(ctx, sc, bvhs)
    cams = [scenes.normalized_camera(c, bvhs[k].normalizer()) for k, _, _, _, c in views]
    samples = sum(w * h * spp for _, w, h, spp, _ in views)
    log(f"config4: {len(sc)} scenes ({sum(s.num_triangles for s in sc)} triangles), {len(views)} views, "
        f"{samples / 1e6:.1f} M samples ({time.time() - t0:.1f}s setup)")
    depth = args.batch_depth
    frames = []
    for i, (k, w, h, spp, _) in enumerate(views):
        f = pike_amd.Frame(ctx, w, h)
        f.set_camera(cams[i])
        frames.append(f)

    def plain():
        for i, (k, w, h, spp, _) in enumerate(views):
            frames[i].reset()
            frames[i].render(per_scene[k], max_depth=depth, first_sample=0, num_frames=spp, use_graph=True)
        ctx.sync()

    # the plain method with multi-sample passes: each view traces up to
    # 2^22 paths per pass
    frames_ms = []
    for i, (k, w, h, spp, _) in enumerate(views):
        f = pike_amd.Frame(ctx, w, h)
        f.set_camera(cams[i])
        f.set_samples_per_pass(max(1, min(spp, (1 << 22) // (w * h))))
        frames_ms.append(f)

    def plain_ms():
        for i, (k, w, h, spp, _) in enumerate(views):
            frames_ms[i].reset()
            frames_ms[i].render(per_scene[k], max_depth=depth, first_sample=0, num_frames=spp, use_graph=True)
        ctx.sync()

    batch = pike_amd.Batch(ctx, merged, [(k, w, h, spp, cams[i]) for i, (k, w, h, spp, _) in enumerate(views)])
    total_px = sum(w * h for _, w, h, _, _ in views)

    def batched():
        batch.reset()
        batch.render(max_depth=depth, use_graph=True)
        ctx.sync()

    # each method under each pass scheduler; the images must be identical
    # across all of them
    ref = None
    per = {}
    same = True
    for sched in ("wavefront", "paths"):
        ctx.set_scheduler(sched)
        plain()
        plain_ms()
        batched()  # warm-up: instantiates every graph
        t0 = time.perf_counter()
        plain()
        t_plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        plain_ms()
        t_plain_ms = time.perf_counter() - t0
        # batched rounds of 1, 2, 4, ... samples per view pixel (below 2^30
        # paths per round, the frame's limit); the fastest is the batched figure
        t_round = {}
        for spr in (1, 2, 4, 8, 16, 32):
            if total_px * spr >= (1 << 30):
                break
            batch.set_samples_per_round(spr)
            batched()  # warm-up: records this shape's graph
            t0 = time.perf_counter()
            batched()
            t_round[spr] = time.perf_counter() - t0
            if not all(np.array_equal(batch.read_view(i).view(np.uint32), frames[i].read_image().view(np.uint32))
                       for i in range(len(views))):
                same = False
        best_spr = min(t_round, key=t_round.get)
        t_batch = t_round[best_spr]
        if ref is None:
            ref = [frames[i].read_image() for i in range(len(views))]
        same = same and all(
            np.array_equal(batch.read_view(i).view(np.uint32), ref[i].view(np.uint32)) and
            np.array_equal(frames_ms[i].read_image().view(np.uint32), ref[i].view(np.uint32)) and
            np.array_equal(frames[i].read_image().view(np.uint32), ref[i].view(np.uint32))
            for i in range(len(views)))
        per[sched] = dict(plain_s=t_plain, plain_multisample_s=t_plain_ms, batched_s=t_batch,
                          batched_samples_per_round=best_spr, batched_s_by_samples_per_round=t_round)
        log(f"batching ({sched}): plain {samples / t_plain / 1e6:.0f}, plain multi-sample "
            f"{samples / t_plain_ms / 1e6:.0f}, batched {samples / t_batch / 1e6:.0f} Msamples/s "
            f"({best_spr} samples per round; " +
            ", ".join(f"{k}: {samples / v / 1e6:.0f}" for k, v in t_round.items()) + ")")
    ctx.set_scheduler("wavefront")
    best = {k: min(v[k] for v in per.values()) for k in ("plain_s", "plain_multisample_s", "batched_s")}
    t_plain, t_plain_ms, t_batch = best["plain_s"], best["plain_multisample_s"], best["batched_s"]
    out = dict(samples=samples, plain_s=t_plain, plain_multisample_s=t_plain_ms, batched_s=t_batch,
               plain_msamples_s=samples / t_plain / 1e6, plain_multisample_msamples_s=samples / t_plain_ms / 1e6,
               batched_msamples_s=samples / t_batch / 1e6, speedup=t_plain / t_batch,
               speedup_vs_plain_multisample=t_plain_ms / t_batch, images_identical=same,
               batched_scheduler=min(per, key=lambda k: per[k]["batched_s"]), per_scheduler=per,
               depth=depth, scenes=len(sc), views=len(views))
    log(f"batching: plain {out['plain_msamples_s']:.0f} Msamples/s, plain multi-sample "
        f"{out['plain_multisample_msamples_s']:.0f}, batched {out['batched_msamples_s']:.0f} Msamples/s "
        f"({out['speedup']:.2f}x / {out['speedup_vs_plain_multisample']:.2f}x), identical images: {same}")
    return out


def run_diff(args, pike_amd, scenes):
    """Config 5 (differentiation): 196,000 triangles (terrain + 50 icospheres),
    512x512 RGBA32F atlas whose base color (RGB) and roughness (G) layers and
    the 512x256 environment are the parameters; target = ground truth at 256
    spp; parameters start at ground truth + N(0, 0.1), clamped.  One iteration
    = 16-spp forward render (graph) + L2 loss/gradient kernel + path-replay
    adjoint over 16 independent samples, all device-resident.  Checked on 64
    random texels (base color, roughness, environment) against central
    differences of the GPU forward (eps 1e-3, common seeds) and against the
    oracle's FP64 dual-number gradient of the same estimator."""
    import numpy as np
    from pike_amd import _ffi

    ATLAS, ENV = _ffi.PK_PARAM_ATLAS, _ffi.PK_PARAM_ENV
    t0 = time.time()
    data = scenes.config5(seed=5)
    bvh = pike_amd.BVH(data.positions, data.triangles, threads=min(16, os.cpu_count() or 1),
                       spatial_splits=args.bvh == "sbvh", sah_collapse=args.collapse == "sah")
    cam = scenes.normalized_camera(data.camera, bvh.normalizer())
    ctx = pike_amd.Context(0)
    scene = pike_amd.Scene(ctx, data, bvh)
    W = H = args.diff_res
    spp, depth = args.diff_spp, args.diff_depth
    frame = pike_amd.Frame(ctx, W, H)
    frame.set_camera(cam)
    frame.set_samples_per_pass(spp)  # the 16 samples of an iteration are one pass
    if args.scheduler == "auto":
        frame.autotune_scheduler(scene, max_depth=depth, samples=4 * spp)
    else:
        ctx.set_scheduler(args.scheduler)
    frame.render(scene, max_depth=depth, num_frames=args.diff_target_spp)
    target = pike_amd.DeviceBuffer(ctx, W * H * 16)
    frame.read_image_device(target)
    ctx.sync()
    log(f"config5: {data.num_triangles} triangles, target at {args.diff_target_spp} spp ({time.time() - t0:.1f}s)")
    rng = np.random.default_rng(55)
    tex = data.tex.copy()
    tex[0, ..., :3] = np.clip(tex[0, ..., :3] + rng.normal(0, 0.1, tex[0, ..., :3].shape), 0.02, 1.0)
    tex[1, ..., 1] = np.clip(tex[1, ..., 1] + rng.normal(0, 0.1, tex[1, ..., 1].shape), 0.05, 1.0)
    info = scene.param_info()
    eh, ew = info["env_height"], info["env_width"]
    env = np.zeros((eh, ew, 4), np.float32)
    env[..., :3] = np.maximum(data.env.reshape(eh, ew, 3) * (1.0 + rng.normal(0, 0.1, (eh, ew, 3))), 0.0)
    scene.write_params(ATLAS, tex.astype(np.float32), also_sampling=True)
    scene.write_params(ENV, env)
    na, ne = scene.param_count(ATLAS), scene.param_count(ENV)
    dimg = pike_amd.DeviceBuffer(ctx, W * H * 16)
    ga, ge = pike_amd.DeviceBuffer(ctx, na * 4), pike_amd.DeviceBuffer(ctx, ne * 4)

    def forward():
        frame.reset()
        frame.render(scene, max_depth=depth, num_frames=spp)

    def iteration():
        forward()
        frame.loss_l2(target, dimg, want_loss=False)
        frame.adjoint_device(scene, dimg, ga, ge, max_depth=depth, first_sample=spp, num_frames=spp)

    for _ in range(2):
        iteration()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.diff_iters):
        iteration()
    ctx.sync()
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(args.diff_iters):
        forward()
    ctx.sync()
    t_fwd = (time.perf_counter() - t0) / args.diff_iters
    t0 = time.perf_counter()
    for _ in range(args.diff_iters):
        frame.adjoint_device(scene, dimg, ga, ge, max_depth=depth, first_sample=spp, num_frames=spp)
    ctx.sync()
    t_adj = (time.perf_counter() - t0) / args.diff_iters
    out = dict(iters_per_s=args.diff_iters / dt, ms_per_iter=dt / args.diff_iters * 1e3, forward_ms=t_fwd * 1e3,
               adjoint_ms=t_adj * 1e3, adjoint_mpaths_s=W * H * spp / t_adj / 1e6, resolution=[W, H], spp=spp,
               depth=depth, triangles=data.num_triangles, scheduler=ctx.scheduler(),
               params=int(info["tex_width"] * info["tex_height"] * 4 + eh * ew * 3))
    log(f"differentiation: {out['iters_per_s']:.1f} iters/s (forward {out['forward_ms']:.2f} ms, adjoint "
        f"{out['adjoint_ms']:.2f} ms)")
    if args.no_diff_checks:
        return out
    # gradient checks: J = sum(d * image) with d = dLoss/dImage at the start,
    # differentiated on the forward's own samples
    forward()
    frame.loss_l2(target, dimg, want_loss=False)
    d = dimg.read(np.float32, W * H * 4).reshape(H, W, 4)
    gatl, genv = frame.adjoint(scene, d, max_depth=depth, num_frames=spp)
    prng = np.random.default_rng(66)
    params = []
    for _ in range(24):
        y, x = prng.integers(0, info["tex_height"]), prng.integers(0, info["tex_width"])
        params.append((ATLAS, int(((0 * info["tex_height"] + y) * info["tex_width"] + x) * 4 + prng.integers(0, 3))))
    for _ in range(16):
        y, x = prng.integers(0, info["tex_height"]), prng.integers(0, info["tex_width"])
        params.append((ATLAS, int(((1 * info["tex_height"] + y) * info["tex_width"] + x) * 4 + 1)))
    for _ in range(24):
        y, x = prng.integers(0, eh), prng.integers(0, ew)
        params.append((ENV, int((y * ew + x) * 4 + prng.integers(0, 3))))
    g = np.array([(gatl.ravel() if w == ATLAS else genv.ravel())[i] for w, i in params], np.float64)
    d3 = d[..., :3].astype(np.float64)
    fd, noise = [], []
    flat = {ATLAS: tex.ravel(), ENV: env.ravel()}
    for which, off in params:
        v = float(flat[which][off])
        hi, lo = np.float32(v + 1e-3), np.float32(v - 1e-3)
        imgs = []
        for val in (hi, lo):
            scene.write_params(which, [val], off)
            forward()
            imgs.append(frame.read_image()[..., :3])
        scene.write_params(which, [np.float32(v)], off)
        den = float(hi) - float(lo)
        fd.append(float((d3 * imgs[0]).sum() - (d3 * imgs[1]).sum()) / den)
        ulps = 2 * spp * (np.spacing(np.abs(imgs[0])) + np.spacing(np.abs(imgs[1])))
        noise.append(float((np.abs(d3) * ulps * (imgs[0] != imgs[1])).sum()) / den)
    fd, noise = np.array(fd), np.array(noise)
    nz = fd != 0
    rel = np.abs(g - fd)[nz] / np.abs(fd[nz])
    out["fd_check"] = dict(texels=len(params), nonzero=int(nz.sum()), zero_agree=bool(np.all(g[~nz] == 0)),
                           within_2pct=int((rel <= 0.02).sum()), median_rel=float(np.median(rel)) if rel.size else None,
                           within_2pct_plus_rounding=int((np.abs(g - fd) <= 0.02 * np.abs(fd) + noise).sum()))
    log(f"FD check: {out['fd_check']}")
    if not args.no_cpu:
        from oracle.oracle import OracleScene

        t0 = time.time()
        orc = OracleScene(data, threads=min(16, os.cpu_count() or 1))
        orc.build_bvh()
        orc.write_params(ATLAS, tex.astype(np.float32), also_sampling=True)
        orc.write_params(ENV, env)
        for prec in (64, 32):
            t0 = time.time()
            gr, _ = orc.gradient(cam, W, H, depth, d, params, num_frames=spp, precision=prec)
            nzr = gr != 0
            out[f"fp{prec}_check"] = dict(rel_l1=float(np.abs(g - gr).sum() / np.abs(gr).sum()),
                                          max_rel=float((np.abs(g - gr)[nzr] / np.abs(gr[nzr])).max()),
                                          median_rel=float(np.median(np.abs(g - gr)[nzr] / np.abs(gr[nzr]))),
                                          zero_agree=bool(np.all(g[~nzr] == 0)), cpu_s=time.time() - t0)
            log(f"FP{prec} dual-number check: {out[f'fp{prec}_check']}")
    return out


def run_cpu_baseline(args, data, cam, first_frame=None):
    """The oracle's FP32 restatement timed on whole frames of the render
    workload; its first frame, computed to size the sample, is also compared
    with the GPU's first sample at full size (returned second, or None)."""
    import numpy as np
    from oracle.oracle import OracleScene

    # Every core this process may run on, up to the 16 host cores the GPU
    # host allots to one GPU (the count the machine reports is recorded too).
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, affinity))
    t0 = time.time()
    orc = OracleScene(data, threads=threads)
    orc.build_bvh()
    log(f"oracle BVH ({time.time() - t0:.1f}s)")
    # one whole frame to size the sample, then whole frames up to ~cpu_seconds
    t0 = time.perf_counter()
    acc0, _, _ = orc.render(cam, args.width, args.height, args.depth, 0, 1, precision=32)
    t1 = time.perf_counter() - t0
    parity = None
    if first_frame is not None:
        a, g = acc0[..., :3].astype(np.float64), first_frame[..., :3].astype(np.float64)
        rel = np.abs(a - g) / np.maximum(np.maximum(np.abs(a), np.abs(g)), 1e-6)
        bad = int(np.any(acc0.view(np.uint32) != first_frame.view(np.uint32), axis=-1).sum())
        parity = dict(pixels=int(args.width * args.height), mismatched_pixels=bad, max_rel=float(rel.max()),
                      compared=f"sample 0 at {args.width}x{args.height}, depth {args.depth}: every pixel's "
                               "running mean (bit patterns) against the oracle's FP32 frame")
        log(f"full-size parity: {bad} of {parity['pixels']} pixels differ from the oracle (max rel {rel.max():.2e})")
    frames = int(max(1, round(args.cpu_seconds / max(t1, 1e-3))))
    t0 = time.perf_counter()
    orc.render(cam, args.width, args.height, args.depth, 1, frames, precision=32)
    dt = time.perf_counter() - t0
    v = frames / dt
    log(f"cpu baseline: {frames} frames of {args.width}x{args.height} in {dt:.2f}s on {threads} threads "
        f"-> {v:.3f} spp/s")
    return dict(value=v, unit="spp/s", cores=threads, kind="port", host_cpus=os.cpu_count(),
                affinity_cpus=affinity, core_limit=16,
                sample=f"oracle FP32 restatement (C++, std::thread), {frames} whole {args.width}x{args.height} "
                       f"frames (1 sample per pixel each, depth {args.depth}) of the same scene and camera"), parity


def run_traversal(args, pike_amd, scenes):
    """The 10M-triangle soup traced with 16M incoherent rays, closest hit and
    any hit; per BVH kind (spatial-split SBVH, and the plain SAH BVH when
    asked for) with its build time and traversal counters."""
    import numpy as np

    data = scenes.config3(seed=3, num_tris=args.soup_tris)
    threads = min(16, os.cpu_count() or 1)
    kinds = {"sbvh": [True], "sah": [False], "both": [True, False]}[args.soup_bvh]
    ctx = pike_amd.Context(0)
    out = {}
    # full-size check, outside the timed region: hit ids, t and bary of a
    # ray subset against the oracle's canonical closest hit (argmin over
    # (t, triangle id) with the reference's f32 intersector) and occlusion
    # against its any-hit answer
    orc = None
    if not args.no_cpu and args.soup_check_rays > 0:
        from oracle.oracle import OracleScene

        t0 = time.time()
        orc = OracleScene(data, threads=threads)
        orc.build_bvh()
        log(f"oracle soup BVH ({time.time() - t0:.1f}s)")
    for spatial in kinds:
        kind = "sbvh" if spatial else "sah"
        t0 = time.time()
        bvh = pike_amd.BVH(data.positions, data.triangles, threads=threads, spatial_splits=spatial,
                           sah_collapse=args.soup_collapse == "sah")
        build_s = time.time() - t0
        info = bvh.info()
        log(f"config3 soup {kind}: {data.num_triangles} triangles, BVH {info} ({build_s:.1f}s, {threads} threads)")
        scene = pike_amd.Scene(ctx, data, bvh)
        rays = scenes.incoherent_rays(args.soup_rays, bvh.normalizer(), seed=33)
        n = rays.shape[0]
        d_rays = pike_amd.DeviceBuffer.from_array(ctx, rays)
        check = rays[:min(n, args.soup_check_rays)].copy() if orc is not None else None
        del rays
        ids = pike_amd.DeviceBuffer(ctx, 4 * n)
        t = pike_amd.DeviceBuffer(ctx, 4 * n)
        b = pike_amd.DeviceBuffer(ctx, 4 * n)
        res = dict(build_s=build_s, build_threads=threads, leaf_references=info["leaf_triangles"],
                   wide_nodes=info["wide_nodes"], collapse=args.soup_collapse)
        for name, anyh in (("closest", False), ("any", True)):
            if not args.no_autotune:
                before, after = ctx.autotune_rays(scene, d_rays.ptr, n, anyh, ids.ptr, t.ptr, b.ptr)
                log(f"autotune {kind} {name}: {before:.2f} -> {after:.2f} ms, "
                    f"{ctx.tuning()['rays_any' if anyh else 'rays_closest']}")
            ctx.trace_rays_device(scene, d_rays.ptr, n, anyh, ids.ptr, t.ptr, b.ptr, True)
            nodes, tris = ctx.trace_stats()
            ms = ctx.time_trace_rays(scene, d_rays.ptr, n, anyh, ids.ptr, t.ptr, b.ptr, args.trace_iters)
            byts = traversal_bytes(n, nodes, tris)
            hit_rate = float((ids.read(np.int32, n) >= 0).mean())
            res[name] = dict(mrays_s=n / (ms * 1e-3) / 1e6, ms=ms, nodes_per_ray=nodes / n, tris_per_ray=tris / n,
                             gbs=byts / (ms * 1e-3) / 1e9, hit_rate=hit_rate)
            if check is not None:
                m = check.shape[0]
                t0 = time.time()
                ids_o, t_o, b_o, _ = orc.trace(check, "canonical", any_hit=anyh)
                ids_g = ids.read(np.int32, m)
                if anyh:  # occlusion: the id an any-hit query reports is order-dependent
                    bad = int(((ids_g >= 0) != (ids_o >= 0)).sum())
                else:
                    bad = int(((ids_g != ids_o) | (t.read(np.uint32, m) != t_o.view(np.uint32)) |
                               (b.read(np.uint32, m) != b_o)).sum())
                res[name]["oracle_check"] = dict(rays=m, mismatches=bad, compared="occlusion" if anyh else
                                                 "id, t bits, bary", cpu_s=time.time() - t0)
            log(f"traversal {kind} {name}: {res[name]}")
        out[kind] = res
        for obj in (scene, bvh, d_rays, ids, t, b):
            obj.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed steps; one step = one wavefront pass of the "
                                                          "autotuned samples per pixel (S), i.e. S spp")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes before the timed ones")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--profile-frames", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traversal", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-instanced", action="store_true", help="skip the instanced (two-level BVH) render leg")
    ap.add_argument("--batch-depth", type=int, default=4)
    ap.add_argument("--tuning", default="", help="JSON {class: {waves_per_cu, refill_min}} to apply instead of "
                                                  "autotuning (e.g. the 'extra.autotune.shapes' of a previous run)")
    ap.add_argument("--soup-tris", type=int, default=10_000_000)
    ap.add_argument("--soup-rays", type=int, default=16_000_000)
    ap.add_argument("--instanced", action="store_true",
                    help="render scene with its 200 icospheres as instances of one mesh (two-level BVH)")
    ap.add_argument("--collapse", choices=["greedy", "sah"], default="sah",
                    help="binary -> 8-wide collapse of the render scenes: open the largest child, or the "
                         "SAH-optimal dynamic program")
    ap.add_argument("--soup-collapse", choices=["greedy", "sah"], default="sah",
                    help="the same for the traversal soup (incoherent rays gain from the SAH-optimal collapse)")
    ap.add_argument("--bvh", choices=["sah", "sbvh"], default="sah",
                    help="BVH kind of the render, batching and differentiation scenes")
    ap.add_argument("--soup-bvh", choices=["sbvh", "sah", "both"], default="both",
                    help="BVH kind(s) the soup is traced under: spatial splits (the workload's), plain SAH, or both")
    ap.add_argument("--scheduler", choices=["auto", "wavefront", "paths"], default="auto",
                    help="pass scheduler of the single-view render legs: per-bounce wavefront kernels, the path "
                         "megakernel, or the faster of the two by timing (auto)")
    ap.add_argument("--trace-iters", type=int, default=5)
    ap.add_argument("--soup-check-rays", type=int, default=1_000_000,
                    help="rays of the soup checked against the oracle's canonical hits (0: none)")
    ap.add_argument("--spp-per-pass", type=int, default=1,
                    help="samples per pixel traced per wavefront pass of the 1080p render (without autotuning)")
    ap.add_argument("--max-spp-per-pass", type=int, default=512, help="autotuner's largest samples per pass")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-diff-checks", action="store_true")
    ap.add_argument("--diff-res", type=int, default=512)
    ap.add_argument("--diff-spp", type=int, default=16)
    ap.add_argument("--diff-depth", type=int, default=4)
    ap.add_argument("--diff-target-spp", type=int, default=256)
    ap.add_argument("--diff-iters", type=int, default=20)
    ap.add_argument("--mode", choices=["full", "render-only", "copy", "diff-only"], default="full",
                    help="render-only: the timed frames alone (for rocprofv3 runs); copy: a 1 GiB device copy "
                         "through the library's streaming copy kernel (HBM byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import pike_amd
    from pike_amd import scenes

    if args.mode == "copy":
        ctx = pike_amd.Context(0)
        nbytes = 1 << 30
        a, b = pike_amd.DeviceBuffer(ctx, nbytes), pike_amd.DeviceBuffer(ctx, nbytes)
        for _ in range(3):
            pike_amd.lib().pk_buffer_copy(ctx.h, b.ptr, a.ptr, nbytes)
        ctx.sync()
        print(json.dumps({"mode": "copy", "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes}))
        return
    if args.mode == "diff-only":
        print(json.dumps({"mode": "diff-only", "differentiation_c5": run_diff(args, pike_amd, scenes)}))
        return
    if args.mode == "render-only":
        args.no_instanced = True
        args.no_cpu = True
        args.no_traversal = True
        args.no_batch = True
        args.no_diff = True
        args.profile_frames = 0

    res, data, bvh, cam = run_render(args, pike_amd, scenes)
    first_frame = res.pop("first_frame")
    extra = dict(mrays_per_s_render=res["rays_per_s"] / 1e6, msamples_per_s=res["spp_s"] * args.width * args.height / 1e6,
                 kernel_ms_per_profile=res["kernel_ms"], profile_frames=res["profile_frames"],
                 trace_closest_bytes_per_launch=res["roofline"]["bytes_per_launch"],
                 trace_closest_launch_ms=res["roofline"]["launch_ms"], counters=res["counters"],
                 autotune=res["autotune"])
    cpu, parity = (None, None) if args.no_cpu else run_cpu_baseline(args, data, cam, first_frame)
    if parity:
        extra["full_size_parity"] = parity
    if not args.no_traversal:
        extra["traversal_10M_soup"] = run_traversal(args, pike_amd, scenes)
    if not args.no_instanced:
        ictx = pike_amd.Context(0)
        if res["autotune"]:
            apply_shapes(ictx, res["autotune"]["shapes"])
        extra["instanced_c2"] = run_instanced(args, pike_amd, scenes, ictx, max(1, args.spp_per_pass))
        extra["instanced_c2"]["baked_layout_mb"] = (bvh.info()["wide_nodes"] * 128 +
                                                    bvh.info()["leaf_triangles"] * 48) / 1e6
    if not args.no_batch:
        bctx = pike_amd.Context(0)
        if res["autotune"]:
            apply_shapes(bctx, res["autotune"]["shapes"])
        extra["batching_32_scenes_64_views"] = run_batching(args, pike_amd, scenes, bctx)
    if not args.no_diff:
        extra["differentiation_c5"] = run_diff(args, pike_amd, scenes)
    roof = {k: v for k, v in res["roofline"].items() if k not in ("bytes_per_launch", "launch_ms")}
    line = {
        "metric": "spp/s at 1920x1080 (plain wavefront, depth 8)",
        "value": res["spp_s"],
        "unit": "spp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural terrain + icospheres, log-normal HDR env with sun; seeded PCG)",
        "config": {"workload": f"plain wavefront path tracing, {res['triangles']:,} triangles "
                               f"({args.bvh.upper()} wide BVH{', 200 instanced icospheres' if args.instanced else ''}), "
                               f"2048x1024 env (2D CDF importance sampling), "
                               f"{args.width}x{args.height}, depth {args.depth}; one step = one wavefront pass of "
                               f"{res['samples_per_step']} samples per pixel",
                   "samples_per_step": res["samples_per_step"],
                   "width": args.width, "height": args.height, "depth": args.depth, "triangles": res["triangles"],
                   "bvh": args.bvh, "collapse": args.collapse, "instanced": args.instanced,
                   "scheduler": res["scheduler"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
