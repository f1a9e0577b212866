#!/usr/bin/env python3
"""Benchmark: forward + adjoint + Adam iterations per second of the plain single-view method on the
64^3 x 6 = 1,572,864-tet grid at 1024 x 1024 (one MI355X).

A step is one training iteration fitting the jittered mesh to one target view: render (projection, binning,
radix sort, tile compositing), MSE loss and its image adjoint, adjoint render (gradients of all vertex
positions and colours), fused Adam on colours (lr 1e-3, clamp >= 0) and positions (lr 1e-4), with the gradient
zeroing of the next step fused into Adam.  Inputs are resident in HBM before the timed region.

Prints one JSON line.  ``--gpus N`` with N > 1 is not supported (one GPU per box).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "fwd+adjoint+Adam iters/s, 1.57M tets @1024²; serve p50/p99 ms"
HBM_PEAK_GBS = 8000.0
# VALU issue: one wave64 VALU instruction per SIMD every 4 cycles (packed FP32 off), 256 CUs x 4 SIMDs at the
# 2.4 GHz peak engine clock
VALU_PEAK_INSTR_S = 256 * 4 * 2.4e9 / 4


def algorithmic_bytes(stage, T, V, P):
    """Compulsory HBM bytes of one launch: every input read once, every output written once.
    forward: cells 16 B/tet + positions and colours 28 B/vertex + image 16 B/pixel.
    adjoint: cells 16 B/tet + positions and colours 28 B/vertex + image and its adjoint 32 B/pixel
             + gradient write 28 B/vertex."""
    if stage == "forward":
        return 16 * T + 28 * V + 16 * P
    if stage == "adjoint":
        return 16 * T + 56 * V + 32 * P
    raise ValueError(stage)


def load_traffic(kernel, workload="grid64_1024"):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def valu_roofline(kernel, avg_ms, workload="grid64_1024"):
    """VALU issue rate of the kernel: its wave64 VALU instructions per launch (SQ_INSTS_VALU from a rocprofv3
    --pmc run, kept in profiles/) over the launch duration measured here, against VALU_PEAK_INSTR_S."""
    path = os.path.join(ROOT, "profiles", f"valu_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        instr = json.load(f).get("valu_instructions_per_launch")
    if not instr:
        return None
    achieved = instr / (avg_ms * 1e-3)
    return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": VALU_PEAK_INSTR_S,
            "unit": "wave64 VALU instr/s", "frac": achieved / VALU_PEAK_INSTR_S, "instructions": instr,
            "avg_ms": avg_ms}


def cpu_baseline(w, view, fovy, W, H, row_step):
    """Times the CPU oracle (C, OpenMP) on every row_step-th image row and scales to a full iteration."""
    import numpy as np

    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    uf, _ = oracle.camera_uniforms(view, fovy, 0.01, 100.0, W, H)
    tgt_frags = oracle.Fragments(w["cells"], w["pos_gt"], uf, 0.01, W, H, 0, H, row_step)
    target, _, _ = oracle.forward(tgt_frags, w["pos_gt"], w["col"], uf, 1.0)
    pos = w["pos"].copy()
    col = w["col"].copy()
    t0 = time.perf_counter()
    fr = oracle.Fragments(w["cells"], pos, uf, 0.01, W, H, 0, H, row_step)
    t1 = time.perf_counter()
    img, _, _ = oracle.forward(fr, pos, col, uf, 1.0, use_ert=True)
    loss, dimg = oracle.mse_grad(img, target)
    gp, gc = oracle.adjoint_exact(fr, pos, col, uf, 1.0, dimg, use_ert=True)
    t2 = time.perf_counter()
    mp, vp = np.zeros_like(pos), np.zeros_like(pos)
    mc, vc = np.zeros_like(col), np.zeros_like(col)
    oracle.adam(col, gc, mc, vc, 1e-3, 1, clamp_nonneg=True)
    oracle.adam(pos, gp, mp, vp, 1e-4, 1)
    t3 = time.perf_counter()
    # projection/sort/raster cost of the sampled rows is a lower bound for the full frame: scale the row work
    t_iter = (t1 - t0) + (t2 - t1) * row_step + (t3 - t2)
    return {"value": 1.0 / t_iter, "unit": "iters/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (oracle/oracle.c, FP32, {threads} OpenMP threads): projection+sort of all "
                      f"{w['cells'].shape[0]} tets, raster+forward+exact adjoint on every {row_step}th row of "
                      f"{W}x{H} (scaled x{row_step}), Adam on all parameters; one iteration, "
                      f"{t3 - t0:.2f} s measured"}


def bench_multiview(args, emit=True):
    """Batched multi-view step: the mesh seen from `views` cameras on a Fibonacci sphere of radius 2 (512 x 512
    each) rendered, differentiated (loss = sum of the per-view MSEs) and optimised in one batched step, against
    the plain single-view step (fwd + adjoint + Adam of one view) run for every view."""
    import numpy as np
    import torch

    import difftetvr_amd as d
    from difftetvr_amd import meshgen
    from difftetvr_amd.optim import FusedAdam, mse_loss_grad, mse_loss_grad_views

    W = H = args.res or 512
    nv = args.views
    fovy = math.pi / 4.0
    views = [meshgen.look_at_view_matrix(tuple(e)) for e in meshgen.fibonacci_sphere(nv, 2.0)]
    w = meshgen.grid_workload(args.grid, seed=0)
    T, V, P = w["cells"].shape[0], w["pos"].shape[0], W * H

    def mesh_of(pos, grads):
        m = d.TetMesh()
        m.set_use_gradients(grads)
        m.set_tet_mesh_data(w["cells"], pos, w["col"])
        return m

    def batch_renderer(m):
        br = d.BatchRenderer()
        br.set_attenuation(1.0)
        br.set_camera_fovy(fovy)
        br.set_views(m, views, W, H)
        br.set_tile_launch(getattr(args, "tile_launch", "grid"))
        return br

    targets = batch_renderer(mesh_of(w["pos_gt"], False)).render().detach().clone()
    mesh = mesh_of(w["pos"], True)
    br = batch_renderer(mesh)
    graph_mode = bool(getattr(args, "graph", False))
    opt = FusedAdam([{"params": [mesh.get_vertex_colors()], "lr": 1e-3, "clamp_nonneg": True},
                     {"params": [mesh.get_vertex_positions()], "lr": 1e-4}], zero_grad_after=True,
                    device_step=graph_mode)

    def eager_step():
        imgs = br.render()
        _, dimg = mse_loss_grad_views(imgs, targets)
        br.render_adjoint(dimg, False)
        opt.step()

    batched_step = eager_step
    graph = None
    if graph_mode:
        # planned rendering (record buffers sized from one ordinary render) and the whole step in one graph
        eager_step()
        torch.cuda.synchronize()
        br.set_record_capacity(int(br.get_stats()["num_pairs"] * 1.25) + 4096)
        eager_step()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            eager_step()
        batched_step = graph.replay

    # the plain single-view path on a second copy of the mesh
    mesh1 = mesh_of(w["pos"], True)
    opt1 = FusedAdam([{"params": [mesh1.get_vertex_colors()], "lr": 1e-3, "clamp_nonneg": True},
                      {"params": [mesh1.get_vertex_positions()], "lr": 1e-4}], zero_grad_after=True)
    singles = []
    for v in views:
        r = d.Renderer(d.RendererType.INTERSECTION)
        r.set_tet_mesh(mesh1)
        r.set_attenuation(1.0)
        r.set_viewport_size(W, H)
        r.set_camera_fovy(fovy)
        r.set_view_matrix(v)
        singles.append(r)

    def single_steps():
        for i, r in enumerate(singles):
            img = r.render()
            _, dimg = mse_loss_grad(img, targets[i])
            r.render_adjoint(dimg, False)
            opt1.step()

    for _ in range(args.warmup):
        batched_step()
        single_steps()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nsingle = max(1, args.steps // 2)
    for _ in range(nsingle):
        single_steps()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if graph is not None and br.records_overflowed()[0]:
        raise RuntimeError("bench: record capacity exceeded during the graph replays")
    br.set_timing(True)
    acc = {}
    nt = 3
    for _ in range(nt):
        eager_step()  # stage timers need the eager launches
        for k, v in br.get_timings().items():
            acc[k] = acc.get(k, 0.0) + v / nt
    br.set_timing(False)
    ms_batch = (t1 - t0) * 1000.0 / args.steps
    ms_single_view = (t2 - t1) * 1000.0 / (nsingle * nv)
    stats = br.get_stats()
    dominant = max(("forward", "adjoint"), key=lambda k: acc.get(k, 0.0))
    abytes = algorithmic_bytes(dominant, T, V, nv * P) if dominant == "forward" else \
        16 * T + 56 * V + 32 * nv * P
    achieved = abytes / (acc[dominant] * 1e-3) / 1e9
    out = {
        "metric": f"batched {nv}-view fwd+adjoint+Adam views/s, 1.57M tets @{W}x{H}",
        "value": nv * 1000.0 / ms_batch,
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{nv} views on a Fibonacci sphere of radius 2, {W}x{H} each, one batched "
                               f"render + summed MSE + batched adjoint + Adam per step, {args.grid}^3 grid x6 tets",
                   "tets": T, "vertices": V, "pixels": nv * P, "tile_pairs": stats["num_pairs"],
                   "hip_graph": graph is not None, "tile_launch": getattr(args, "tile_launch", "grid"),
                   "single_view_views_per_s": 1000.0 / ms_single_view,
                   "speedup_vs_single_view": ms_single_view * nv / ms_batch,
                   "stage_ms": {k: round(v, 4) for k, v in acc.items()}},
        "roofline": {"bound": "hbm", "kernel": f"k_render_{dominant}", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic(f"k_render_{dominant}", f"multiview{nv}_grid{args.grid}_{W}"),
                     "algorithmic_bytes": abytes, "avg_ms": acc[dominant]},
        "valu_roofline": valu_roofline(f"k_render_{dominant}", acc[dominant], f"multiview{nv}_grid{args.grid}_{W}"),
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_projection(args, emit=True):
    """The plain single-view step (fwd + adjoint + Adam, as the headline) with the Projection renderer
    (projected tetrahedra) instead of the intersection renderer: same mesh, camera and optimiser."""
    import numpy as np
    import torch

    import difftetvr_amd as d
    from difftetvr_amd import meshgen
    from difftetvr_amd.optim import FusedAdam, mse_loss_grad

    W = H = args.res or 1024
    fovy = math.pi / 4.0
    eye = tuple(2.0 * np.array([0.3, 0.4, 1.0]) / np.linalg.norm([0.3, 0.4, 1.0]))
    view = meshgen.look_at_view_matrix(eye)
    w = meshgen.grid_workload(args.grid, seed=0)

    def make(pos, grads):
        m = d.TetMesh()
        m.set_use_gradients(grads)
        m.set_tet_mesh_data(w["cells"], pos, w["col"])
        r = d.Renderer(d.RendererType.PROJECTION)
        r.set_tet_mesh(m)
        r.set_attenuation(1.0)
        r.set_clear_color(d.vec4(0, 0, 0, 0))
        r.set_viewport_size(W, H)
        r.set_camera_fovy(fovy)
        r.set_view_matrix(view)
        return m, r

    _, r_gt = make(w["pos_gt"], False)
    target = r_gt.render().detach().clone()
    mesh, r = make(w["pos"], True)
    opt = FusedAdam([{"params": [mesh.get_vertex_colors()], "lr": 1e-3, "clamp_nonneg": True},
                     {"params": [mesh.get_vertex_positions()], "lr": 1e-4}], zero_grad_after=True)

    def step():
        img = r.render()
        _, dimg = mse_loss_grad(img, target)
        r.render_adjoint(dimg, False)
        opt.step()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1000.0 / args.steps
    r.set_timing(True)
    acc = {}
    for _ in range(3):
        step()
        for k, v in r.get_timings().items():
            acc[k] = acc.get(k, 0.0) + v / 3
    r.set_timing(False)
    stats = r.get_stats()
    out = {"metric": "Projection renderer fwd+adjoint+Adam iters/s, 1.57M tets @1024²", "value": 1000.0 / ms_per_step,
           "unit": "iters/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"single-view fwd+adjoint+Adam with the Projection renderer, {args.grid}^3 grid x6 "
                                  f"tets, {W}x{H}, the headline's camera and optimiser",
                      "triangles": stats["num_triangles"], "tile_pairs": stats["num_pairs"],
                      "stage_ms": {k: round(v, 4) for k, v in acc.items()}}}
    if emit:
        print(json.dumps(out), flush=True)
    return out


def _delaunay_job(args):
    n, seed = args
    from difftetvr_amd import meshgen
    pts, cells = meshgen.delaunay_ball(n, seed=seed)
    return pts, cells


def bench_uneven(args):
    """Uneven batch: 8 Delaunay meshes of N in {2k, 8k, 32k, 64k, 128k, 256k, 384k, 512k} seeded points in the
    unit ball (3-octave value-noise density in [0, 40]), one 768 x 768 view each, rendered and differentiated as
    one batch (loss = sum of per-view MSEs against renders of a second colour field) with Adam on every mesh;
    against the same work as 8 single-view steps."""
    import multiprocessing

    import numpy as np
    import torch

    import difftetvr_amd as d
    from difftetvr_amd import meshgen
    from difftetvr_amd.optim import FusedAdam, mse_loss_grad, mse_loss_grad_views

    W = H = args.res or 768
    ns = [2_000, 8_000, 32_000, 64_000, 128_000, 256_000, 384_000, 512_000]
    t0 = time.perf_counter()
    with multiprocessing.get_context("spawn").Pool(min(8, os.cpu_count() or 1)) as pool:
        dels = pool.map(_delaunay_job, [(n, 500 + i) for i, n in enumerate(ns)])
    t_setup = time.perf_counter() - t0
    rng = np.random.default_rng(1)
    views, meshes, targets, tets = [], [], [], 0
    for i, (pts, cells) in enumerate(dels):
        e = rng.normal(size=3)
        views.append(meshgen.look_at_view_matrix(tuple(e / np.linalg.norm(e) * 2.5)))
        col = meshgen.blob_colors(pts, seed=600 + i)
        col[:, 3] = meshgen.value_noise(pts, seed=700 + i, vmax=40.0)
        gt = meshgen.blob_colors(pts, seed=800 + i)
        gt[:, 3] = col[:, 3]
        m = d.TetMesh()
        m.set_use_gradients(True)
        m.set_tet_mesh_data(cells, pts, col)
        mt = d.TetMesh()
        mt.set_tet_mesh_data(cells, pts, gt)
        meshes.append(m)
        targets.append(mt)
        tets += int(cells.shape[0])

    def batch_of(ms):
        br = d.BatchRenderer()
        br.set_attenuation(1.0)
        br.set_camera_fovy(math.pi / 4.0)
        br.set_items([(m, v, W, H) for m, v in zip(ms, views)])
        return br

    target_imgs = batch_of(targets).render().detach().clone()
    br = batch_of(meshes)
    opts = [FusedAdam([{"params": [m.get_vertex_colors()], "lr": 1e-3, "clamp_nonneg": True},
                       {"params": [m.get_vertex_positions()], "lr": 1e-4}], zero_grad_after=True) for m in meshes]

    def batched_step():
        imgs = br.render()
        _, dimg = mse_loss_grad_views(imgs, target_imgs)
        br.render_adjoint(dimg, False)
        for o in opts:
            o.step()

    singles = []
    for m, v in zip(meshes, views):
        r = d.Renderer(d.RendererType.INTERSECTION)
        r.set_tet_mesh(m)
        r.set_attenuation(1.0)
        r.set_viewport_size(W, H)
        r.set_camera_fovy(math.pi / 4.0)
        r.set_view_matrix(v)
        singles.append(r)

    def single_steps():
        for i, r in enumerate(singles):
            img = r.render()
            _, dimg = mse_loss_grad(img, target_imgs[i])
            r.render_adjoint(dimg, False)
            opts[i].step()

    for _ in range(args.warmup):
        batched_step()
        single_steps()
    torch.cuda.synchronize()
    ta = time.perf_counter()
    for _ in range(args.steps):
        batched_step()
    torch.cuda.synchronize()
    tb = time.perf_counter()
    for _ in range(args.steps):
        single_steps()
    torch.cuda.synchronize()
    tc = time.perf_counter()
    br.set_timing(True)
    batched_step()
    stage_ms = br.get_timings()
    br.set_timing(False)
    st = br.get_stats()
    ms_b = (tb - ta) * 1000.0 / args.steps
    ms_s = (tc - tb) * 1000.0 / args.steps
    # autotuner: time every tile size / segments-per-pass setting on this batch, then the step with its choice
    tune = None
    if not args.no_autotune:
        br.set_autotune(True, with_adjoint=True, repeats=3)
        batched_step()  # the first render tunes
        choice = list(d.autotune_cache().values())[-1]
        for _ in range(args.warmup):
            batched_step()
        torch.cuda.synchronize()
        td = time.perf_counter()
        for _ in range(args.steps):
            batched_step()
        torch.cuda.synchronize()
        ms_t = (time.perf_counter() - td) * 1000.0 / args.steps
        tune = {"tile_size": choice["tile_size"], "segments_per_pass": choice["segments_per_pass"],
                "ms_per_step_tuned": ms_t, "ms_per_step_8x8_128": ms_b, "gain": ms_b / ms_t,
                "render_adjoint_ms": {f"{t}/{sp}": round(ms, 3) for t, sp, ms in choice["candidates"]}}
    out = {
        "metric": "uneven 8-mesh batch fwd+adjoint+Adam ms/step, Delaunay 2k..512k points @768x768",
        "value": ms_b, "unit": "ms/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": False, "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "8 Delaunay meshes in the unit ball (value-noise density [0,40]), one 768^2 view "
                               "each, one batch per step", "tets": tets, "tile_pairs": st["num_pairs"],
                   "single_view_ms_per_step": ms_s, "speedup_vs_single_view": ms_s / ms_b,
                   "mesh_setup_s": round(t_setup, 1), "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()}},
    }
    if tune is not None:
        out["autotune"] = tune
    print(json.dumps(out), flush=True)


SERVE_GRID_NS = [int(n) for n in __import__("numpy").linspace(16, 80, 16).round()]
SERVE_DELAUNAY_NS = [int(n) for n in __import__("numpy").geomspace(10_000, 500_000, 16).round()]


def start_serve_meshes(procs=8):
    """Starts the 16 Delaunay tetrahedralisations of the serving mix in child processes (before this process
    touches the GPU) and returns (pool, async result)."""
    import multiprocessing
    pool = multiprocessing.get_context("spawn").Pool(min(procs, os.cpu_count() or 1))
    return pool, pool.map_async(_delaunay_job, [(n, 100 + i) for i, n in enumerate(SERVE_DELAUNAY_NS)])


def bench_serve(args, qps_list=None, pending=None, emit=True):
    """Serving: a Poisson stream of render requests at a fixed rate over 32 resident meshes (16 cube grids of
    16^3..80^3 cells x 6 tets, 16 Delaunay meshes of 10k..500k points in the unit ball), each request a seeded
    random camera on a sphere of radius 1.5..3 with a viewport drawn from {256^2, 512^2, 1024^2}; 90% forward
    renders, 10% forward + MSE loss + adjoint.  Reports p50/p99 latency (submit to completion) and frames/s, one
    JSON object per offered rate in qps_list (default: args.qps)."""
    import threading

    import numpy as np
    import torch

    import difftetvr_amd as d
    from difftetvr_amd import meshgen
    from difftetvr_amd.serve import Server

    t0 = time.perf_counter()
    if pending is None:
        pending = start_serve_meshes()
    pool, res = pending
    dels = res.get()
    pool.close()
    pool.join()
    meshes, tets = [], []
    for i, n in enumerate(SERVE_GRID_NS):
        w = meshgen.grid_workload(int(n), seed=i)
        m = d.TetMesh()
        m.set_use_gradients(True)
        m.set_tet_mesh_data(w["cells"], w["pos"], w["col"])
        meshes.append(m)
        tets.append(int(w["cells"].shape[0]))
    for i, (pts, cells) in enumerate(dels):
        col = meshgen.blob_colors(pts, seed=200 + i)
        col[:, 3] = meshgen.value_noise(pts, seed=300 + i, vmax=40.0)
        m = d.TetMesh()
        m.set_use_gradients(True)
        m.set_tet_mesh_data(cells, pts, col)
        meshes.append(m)
        tets.append(int(cells.shape[0]))
    t_setup = time.perf_counter() - t0
    sizes = [256, 512, 1024]
    # output and target buffers allocated up front (the server writes them from its own streams)
    outs = {sz: [torch.empty((sz, sz, 4), device="cuda") for _ in range(32)] for sz in sizes}
    targets = {sz: torch.zeros((sz, sz, 4), device="cuda") for sz in sizes}
    torch.cuda.synchronize()
    results_all = []
    for qps in (qps_list or [args.qps]):
        rng = np.random.default_rng(0)
        nreq = int(qps * args.duration)
        plan = []
        t = 0.0
        for _ in range(nreq):
            t += rng.exponential(1.0 / qps)
            mi = int(rng.integers(len(meshes)))
            e = rng.normal(size=3)
            e = e / np.linalg.norm(e) * rng.uniform(1.5, 3.0)
            sz = sizes[int(rng.integers(3))]
            plan.append((t, mi, meshgen.look_at_view_matrix(tuple(e)), sz, bool(rng.random() < 0.1)))
        # early ray termination on, the reference renderer's default (TetMeshVolumeRenderer.hpp:243-244)
        srv = Server(num_workers=args.workers, max_batch=args.max_batch, memory_budget=16 << 30, attenuation=1.0,
                     fovy=math.pi / 4.0, use_early_ray_termination=True)
        ids = [srv.add_mesh(m) for m in meshes]
        # warm-up: every mesh once at every size
        for mi in range(len(meshes)):
            for sz in sizes:
                srv.wait(srv.submit(ids[mi], plan[0][2], sz, sz, out=outs[sz][0]))
        tickets = []
        lock = threading.Lock()
        results = []

        def collector():
            k = 0
            while k < nreq:
                with lock:
                    have = len(tickets)
                if k >= have:
                    time.sleep(0.0002)
                    continue
                tk, sz, adj = tickets[k]
                _, lat, q = srv.wait(tk)
                results.append((lat, q, sz, adj))
                k += 1

        th = threading.Thread(target=collector)
        th.start()
        start = time.perf_counter()
        for i, (ta, mi, vm, sz, adj) in enumerate(plan):
            now = time.perf_counter() - start
            if ta > now:
                time.sleep(ta - now)
            tk = srv.submit(ids[mi], vm, sz, sz, target=targets[sz] if adj else None, out=outs[sz][i % 32])
            with lock:
                tickets.append((tk, sz, adj))
        th.join()
        elapsed = time.perf_counter() - start
        st = srv.stats()
        srv.close()
        lat = np.array([r[0] for r in results])
        lat1k = np.array([r[0] for r in results if r[2] == 1024])

        def pct(sel):
            v = np.array([r[0] for r in results if sel(r)])
            return {"n": int(v.size), "p50_ms": float(np.percentile(v, 50)) if v.size else None,
                    "p99_ms": float(np.percentile(v, 99)) if v.size else None}

        classes = {f"render_{sz}": pct(lambda r, sz=sz: r[2] == sz and not r[3]) for sz in sizes}
        classes.update({f"train_{sz}": pct(lambda r, sz=sz: r[2] == sz and r[3]) for sz in sizes})
        out = {
            "metric": "serve p50/p99 ms (Poisson stream, 32 resident meshes)",
            "value": float(np.percentile(lat, 99)),
            "unit": "ms (p99 latency)",
            "n_gpus": 1,
            "higher_is_better": False,
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic",
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "p99_ms_1024": float(np.percentile(lat1k, 99)) if lat1k.size else None,
            "p50_ms_1024": float(np.percentile(lat1k, 50)) if lat1k.size else None,
            "p99_ms_1024_render": classes["render_1024"]["p99_ms"],
            "p50_ms_1024_render": classes["render_1024"]["p50_ms"],
            "by_class": classes,
            "mean_queue_ms": float(np.mean([r[1] for r in results])),
            "frames_per_s": len(results) / elapsed,
            "offered_qps": qps,
            "config": {"workload": "Poisson render requests over 16 grid meshes (16^3..80^3 x6) and 16 Delaunay "
                                   "meshes (10k..500k points), viewports {256,512,1024}^2, 10% with loss + adjoint, "
                                   f"{args.workers} workers, up to {args.max_batch} requests per batch, "
                                   "16 GiB in-flight budget, early ray termination on",
                       "requests": nreq, "duration_s": args.duration, "meshes": len(meshes),
                       "tets_total": int(sum(tets)), "mesh_setup_s": round(t_setup, 1), "batches": st["batches"],
                       "max_batch_seen": st["max_batch_seen"]},
        }
        if emit:
            print(json.dumps(out), flush=True)
        results_all.append(out)
    return results_all


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--res", type=int, default=None, help="viewport edge (default 1024; 512 for multiview)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="uneven workload: skip the autotuner comparison (tile size, segments per pass)")
    ap.add_argument("--workload", choices=["single", "multiview", "uneven", "serve", "projection"], default="single",
                    help="single: the plain single-view step (default); multiview: 16 views per batched step; "
                         "uneven: 8 Delaunay meshes in one batch; serve: Poisson request stream over 32 meshes; "
                         "projection: the single-view step with the Projection renderer")
    ap.add_argument("--qps", type=float, default=150.0, help="serve: offered requests per second")
    ap.add_argument("--duration", type=float, default=8.0, help="serve: seconds of request arrivals")
    ap.add_argument("--views", type=int, default=16)
    ap.add_argument("--workers", type=int, default=4, help="serve: worker threads (one stream and view each)")
    ap.add_argument("--max-batch", type=int, default=16, help="serve: most requests per batch")
    ap.add_argument("--graph", action="store_true",
                    help="single, multiview: capture the step in a HIP graph (planned rendering) and replay it")
    ap.add_argument("--checkpoint-interval", type=int, default=64,
                    help="single: segments between the adjoint's stored states (K)")
    ap.add_argument("--tile-order", choices=["cost", "row-major"], default="cost",
                    help="single: launch order of the screen tiles (heaviest first, or row-major)")
    ap.add_argument("--tile-launch", choices=["grid", "queue"], default="grid",
                    help="single, multiview: one workgroup per tile, or persistent workgroups taking tiles from a "
                         "device work queue")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="single: skip the multi-view and serving legs the default line also reports")
    ap.add_argument("--extra-steps", type=int, default=10, help="single: timed batched steps of the multi-view leg")
    ap.add_argument("--serve-qps", type=str, default="150,250",
                    help="single: offered rates of the serving legs (comma-separated requests/s)")
    ap.add_argument("--cpu-row-step", type=int, default=32)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device-to-device copies of 256 MiB (scale reference for HBM PMC counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 256 * 1024 * 1024 // 4
        src = torch.arange(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this benchmark runs on one GPU (views and "
              "requests would shard across GPUs, each rendering its share).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "multiview":
        bench_multiview(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "uneven":
        bench_uneven(args)
        return
    if args.workload == "projection":
        bench_projection(args)
        return

    # the serving mix's Delaunay meshes are built in child processes while the GPU legs run (started before
    # this process touches the GPU)
    serve_pending = None if args.no_extras else start_serve_meshes()

    import numpy as np
    import torch

    import difftetvr_amd as d
    from difftetvr_amd import meshgen
    from difftetvr_amd.optim import FusedAdam, mse_loss_grad

    W = H = args.res or 1024
    fovy = math.pi / 4.0
    eye = tuple(2.0 * np.array([0.3, 0.4, 1.0]) / np.linalg.norm([0.3, 0.4, 1.0]))
    view = meshgen.look_at_view_matrix(eye)
    w = meshgen.grid_workload(args.grid, seed=0)
    T, V, P = w["cells"].shape[0], w["pos"].shape[0], W * H

    def make(pos, grads):
        m = d.TetMesh()
        m.set_use_gradients(grads)
        m.set_tet_mesh_data(w["cells"], pos, w["col"])
        r = d.Renderer(d.RendererType.INTERSECTION)
        r.set_tet_mesh(m)
        r.set_attenuation(1.0)
        r.set_clear_color(d.vec4(0, 0, 0, 0))
        r.set_viewport_size(W, H)
        r.set_camera_fovy(fovy)
        r.set_view_matrix(view)
        return m, r

    _, r_gt = make(w["pos_gt"], False)
    target = r_gt.render().detach().clone()
    mesh, r = make(w["pos"], True)
    r.set_checkpoint_interval(args.checkpoint_interval)
    r.set_tile_order(args.tile_order)
    r.set_tile_launch(args.tile_launch)
    opt = FusedAdam([{"params": [mesh.get_vertex_colors()], "lr": 1e-3, "clamp_nonneg": True},
                     {"params": [mesh.get_vertex_positions()], "lr": 1e-4}], zero_grad_after=True,
                    device_step=args.graph)
    if args.graph:
        # planned rendering: record buffers sized from an ordinary render, no host synchronisation per step
        r.render()
        r.set_record_capacity(int(r.get_stats()["num_pairs"] * 1.25) + 4096)

    # HIP events around each stage on the stream the kernels run on (torch's current stream), read after the
    # timed region: no synchronisation inside it
    ev = []

    def step(record=False):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)] if record else None
        if record:
            e[0].record()
        img = r.render()
        if record:
            e[1].record()
        _, dimg = mse_loss_grad(img, target)
        if record:
            e[2].record()
        r.render_adjoint(dimg, False)
        if record:
            e[3].record()
        opt.step()
        if record:
            ev.append(e)

    for _ in range(args.warmup):
        step()
    graph = None
    if args.graph:
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if graph is not None:
            graph.replay()
        else:
            step(record=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    # per-stage breakdown of the pipeline (library timers) on a few further steps
    r.set_timing(True)
    acc = {}
    for _ in range(3):
        step(record=graph is not None)
        for k, v in r.get_timings().items():
            acc[k] = acc.get(k, 0.0) + v / 3
    r.set_timing(False)
    stage_ms = dict(acc)
    if ev:  # the timed region's own events (the graph path measures its eager steps above)
        stage_ms["render"] = sum(e[0].elapsed_time(e[1]) for e in ev) / len(ev)
        stage_ms["adjoint"] = sum(e[2].elapsed_time(e[3]) for e in ev) / len(ev)
    dominant = max(("forward", "adjoint"), key=lambda k: stage_ms.get(k, 0.0))
    abytes = algorithmic_bytes(dominant, T, V, P)
    achieved = abytes / (stage_ms[dominant] * 1e-3) / 1e9
    stats = r.get_stats()
    if args.graph and r.records_overflowed()[0]:
        raise RuntimeError("bench: record capacity exceeded during the graph replays")

    out = {
        "metric": METRIC,
        "value": 1000.0 / ms_per_step,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain single-view fwd+adjoint+Adam, {args.grid}^3 grid x6 tets over [-0.5,0.5]^3 "
                               f"(jitter 0.2 cell), {W}x{H}, fovy 45 deg, eye at distance 2 "
                               f"dir (0.3,0.4,1), attenuation 1, ERT on",
                   "tets": T, "vertices": V, "pixels": P, "tile_pairs": stats["num_pairs"],
                   "front_faces": stats["num_triangles"], "sort_passes": stats["sort_passes"],
                   "segments": int(r.get_fragment_counts().to(torch.int64).sum()),
                   "gradient_atomics": stats["num_gradient_atomics"],
                   "hip_graph": bool(args.graph), "checkpoint_interval": args.checkpoint_interval,
                   "tile_order": args.tile_order, "tile_launch": args.tile_launch,
                   "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()}},
        "roofline": {"bound": "hbm", "kernel": f"k_render_{dominant}", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic(f"k_render_{dominant}"),
                     "algorithmic_bytes": abytes, "avg_ms": stage_ms[dominant]},
        # the render kernels are bound by VALU issue, not HBM: the fraction that measures them
        "valu_roofline": valu_roofline(f"k_render_{dominant}", stage_ms[dominant]),
    }
    if not args.no_extras:
        # the rest of the headline metric, measured in the same run: the batched multi-view step (16 views at
        # 512^2 on the same mesh) against the plain single-view path, and serving latency at two offered rates
        del r, r_gt, mesh, opt
        torch.cuda.empty_cache()
        mv_args = argparse.Namespace(**vars(args))
        mv_args.steps, mv_args.warmup, mv_args.res = args.extra_steps, 2, 512
        mv = bench_multiview(mv_args, emit=False)
        out["multiview"] = {"views_per_s": mv["value"], "ms_per_step": mv["ms_per_step"],
                            "single_view_views_per_s": mv["config"]["single_view_views_per_s"],
                            "speedup_vs_single_view": mv["config"]["speedup_vs_single_view"],
                            "views": args.views, "res": 512, "steps": args.extra_steps,
                            "stage_ms": mv["config"]["stage_ms"],
                            "valu_frac": (mv["valu_roofline"] or {}).get("frac")}
        torch.cuda.empty_cache()
        sv = bench_serve(args, qps_list=[float(q) for q in args.serve_qps.split(",")], pending=serve_pending,
                         emit=False)
        out["serve"] = [{"offered_qps": o["offered_qps"], "frames_per_s": o["frames_per_s"], "p50_ms": o["p50_ms"],
                         "p99_ms": o["p99_ms"], "p50_ms_1024_render": o["p50_ms_1024_render"],
                         "p99_ms_1024_render": o["p99_ms_1024_render"], "p99_ms_1024": o["p99_ms_1024"],
                         "requests": o["config"]["requests"], "max_batch_seen": o["config"]["max_batch_seen"]}
                        for o in sv]
        out["serve_config"] = sv[0]["config"]["workload"]
        torch.cuda.empty_cache()
        pr_args = argparse.Namespace(**vars(args))
        pr_args.steps, pr_args.warmup = args.extra_steps, 2
        pr = bench_projection(pr_args, emit=False)
        out["projection"] = {"iters_per_s": pr["value"], "ms_per_step": pr["ms_per_step"], "steps": args.extra_steps,
                             "stage_ms": pr["config"]["stage_ms"]}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, view, fovy, W, H, args.cpu_row_step)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
