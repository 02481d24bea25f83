"""Walk-on-Stars benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config holes|star|ico]

A step is one fused u + grad u solve over the whole query set of the chosen
workload (default: 2D star polygon with 4096 Dirichlet segments and eight
64-segment Neumann holes, 65,536 lattice queries, 1024 walks per point,
eps 1e-4, at most 1024 steps per walk, g = x^2 - y^2).  Inputs are resident in
HBM before timing starts.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic bytes per walk step, counted by the instrumented CPU restatement
# in the reference traversal order with the reference SoA record sizes
# (2D BVH node 24 B, 2D SNCH node 48 B, 3D SNCH node 60 B, segment 20 B,
# triangle 40 B, silhouette vertex 40 B, silhouette edge 68 B, CDF gathers
# 4 B per binary-search halving + 8 B) on 64 seeded points of each workload.
B_ALG = {"star": 3350.8, "holes": 4641.3, "ico": 39510.1}
HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "star": dict(workload="star polygon 256 segments (mixed D/N), g=e^x cos y", n_query=1024, nwalks=256, eps=1e-3),
    "holes": dict(workload="star polygon 4096 segments + 8 Neumann holes x 64, g=x^2-y^2, 256x256-lattice queries",
                  n_query=65536, nwalks=1024, eps=1e-4),
    "ico": dict(workload="noisy icosphere 81920 triangles, y>0 Dirichlet, g=e^x cos y", n_query=262144, nwalks=512,
                eps=1e-3),
}


def make_spec(name: str):
    from dwost_amd import scenes

    c = CONFIGS[name]
    if name == "star":
        return scenes.star_scene_2d(n_query=c["n_query"], eps=c["eps"])
    if name == "holes":
        return scenes.holes_scene_2d(n_query=c["n_query"], eps=c["eps"])
    return scenes.icosphere_scene(level=6, n_query=c["n_query"], eps=c["eps"])


def profile_figures(name: str) -> dict:
    """Figures of the committed rocprofv3 summary of this workload's kernel
    (profiles/k_fused_pers_<name>.md): HBM traffic per walk step from
    FETCH_SIZE / WRITE_SIZE (copy-calibrated), L2 hit rate, VALU busy."""
    path = os.path.join(ROOT, "profiles", f"k_fused_pers_{name}.md")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            k, sep, v = line.strip().partition(": ")
            if sep and k in ("traffic_bytes_per_walk_step", "l2_hit_rate", "valu_busy"):
                out[k] = float(v)
    return out


def cpu_baseline(spec, cfg, target_s: float = 12.0):
    """The CPU restatement (same estimator, same seeds) on a bounded sample of
    the workload's points, all host cores (OpenMP)."""
    import numpy as np

    from oracle import oracle as O

    O.build()
    sc = O.OracleScene(spec.vertices, spec.indices, spec.types)
    rng = np.random.default_rng(0)
    n = 16
    total_t, total_steps, npts = 0.0, 0, 0
    while total_t < target_s and npts < len(spec.queries):
        idx = rng.choice(len(spec.queries), min(n, len(spec.queries)), replace=False)
        t0 = time.perf_counter()
        out = sc.solve_fused(spec.queries[idx], spec.bc_code, cfg["nwalks"], nsteps=1024, eps=cfg["eps"], seed=0)
        dt = time.perf_counter() - t0
        total_t += dt
        total_steps += out["steps"]
        npts += len(idx)
        n = min(4 * n, max(16, int(n * target_s / max(dt, 1e-3) / 2)))
    return dict(value=total_steps / total_t, unit="walk-steps/s", cores=O.num_threads(), kind="port",
                sample=f"{npts} seeded points x {cfg['nwalks']} walks of the same workload "
                       f"({total_steps} walk steps, {total_t:.1f} s), fused u+grad estimator, "
                       f"C++ restatement of the reference algorithm, OpenMP",
                points_per_s=npts / total_t)


def adjoint_ratio(points: int, reps: int = 2):
    """fwd+bwd / fwd time on the 3D adjoint workload (noisy icosphere, 40,962
    vertices, trainable per-vertex g and h, loss = MSE(u) + MSE(grad u) vs
    e^x cos y, 256 walks per point, backward replays the walks from seeds)."""
    import torch

    import dwost_amd
    from dwost_amd import scenes

    prob = scenes.adjoint_problem_3d(level=6, n_query=points)
    spec = prob.spec
    g = torch.as_tensor(prob.g0).cuda().requires_grad_(True)
    h = torch.as_tensor(prob.h0).cuda().requires_grad_(True)
    scene = dwost_amd.Scene3D(spec.vertices, spec.indices, spec.types, dwost_amd.BoundaryCondition.per_vertex(g, h))
    solver = dwost_amd.WoSt3D(eps=1e-3, nwalks=256, nsteps=1024)
    cs = dwost_amd.CustomSolver(solver)
    q = torch.as_tensor(spec.queries).cuda()
    ut, gt = torch.as_tensor(prob.u_target).cuda(), torch.as_tensor(prob.grad_target).cuda()

    def fwd_bwd(seed):
        g.grad = None
        h.grad = None
        u, gr = cs.solve_grad(q, scene, seed=seed)
        loss = torch.mean((u - ut) ** 2) + torch.mean(torch.sum((gr - gt) ** 2, 1))
        loss.backward()

    fwd_bwd(100)  # warm-up
    solver.solve_grad(q, scene, seed=101)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(reps):
        solver.solve_grad(q, scene, seed=k)
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for k in range(reps):
        fwd_bwd(k)
    torch.cuda.synchronize()
    t_fb = (time.perf_counter() - t0) / reps
    return dict(workload=f"noisy icosphere 81920 triangles, per-vertex g/h, {points} points x 256 walks, "
                         "loss MSE(u)+MSE(grad u)", fwd_ms=round(1000 * t_f, 3), fwd_bwd_ms=round(1000 * t_fb, 3),
                ratio=round(t_fb / t_f, 3))


def batch_speedup(max_queries: int, reps: int = 3):
    """One persistent launch over the 64-scene uneven batch (segments
    log-U[64, 16384], 0-4 holes, walks {64, 256, 1024}, queries log-U[100, 1e5]
    capped at max_queries) against the same solves as per-scene launches."""
    import torch

    import dwost_amd
    from dwost_amd import scenes

    specs = scenes.batch_scenes_2d(n_scenes=64, seed=5, max_queries=max_queries)
    items = []
    for k, spec in enumerate(specs):
        sc = dwost_amd.Scene2D(spec.vertices, spec.indices, spec.types, dwost_amd.BoundaryCondition(spec.bc_code))
        sol = dwost_amd.WoSt(eps=spec.meta["eps"], nwalks=spec.meta["nwalks"], nsteps=1024)
        items.append(dwost_amd.BatchItem(sol, sc, torch.as_tensor(spec.queries).cuda(), seed=k))

    def per_scene():
        for it in items:
            it.solver.solve_grad(it.points, it.scene, seed=it.seed)

    b = dwost_amd.BatchSolver(items, graph=True)
    per_scene()
    b.solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        per_scene()
    torch.cuda.synchronize()
    t_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        b.solve()
    torch.cuda.synchronize()
    t_b = (time.perf_counter() - t0) / reps
    steps = int(b.steps[0])
    b.close()
    return dict(workload=f"64 uneven 2D scenes, {sum(len(s.queries) for s in specs)} points "
                         f"(queries capped at {max_queries}), walks per point from {{64, 256, 1024}}",
                per_scene_ms=round(1000 * t_s, 3), batched_graph_ms=round(1000 * t_b, 3),
                speedup=round(t_s / t_b, 3), walk_steps_per_s=round(steps / t_b, 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="holes", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--points", type=int, default=0, help="use only the first N query points (0: all)")
    ap.add_argument("--progress", action="store_true", help="print a line after each timed step")
    ap.add_argument("--batch-queries", type=int, default=4000,
                    help="query cap per scene of the 64-scene batched-vs-per-scene measurement (0: skip)")
    ap.add_argument("--adjoint-points", type=int, default=2048,
                    help="points of the fwd+bwd/fwd measurement on the 3D adjoint workload (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 256 MiB device-to-device copies (scales the FETCH_SIZE / WRITE_SIZE "
                         "counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.ones(64 << 20, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("[bench] copied 3 x 256 MiB", file=sys.stderr, flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} asked): the walks shard by point "
              "range across GPUs, but that path is not implemented or measured here", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import dwost_amd
    from dwost_amd import scenes

    cfg = CONFIGS[args.config]
    spec = make_spec(args.config)
    Scene = dwost_amd.Scene2D if spec.dim == 2 else dwost_amd.Scene3D
    scene = Scene(spec.vertices, spec.indices, spec.types, dwost_amd.BoundaryCondition(spec.bc_code))
    Solver = dwost_amd.WoSt if spec.dim == 2 else dwost_amd.WoSt3D
    solver = Solver(eps=cfg["eps"], nwalks=cfg["nwalks"], nsteps=1024)
    if args.points:
        spec.queries = spec.queries[: args.points]
    q = torch.as_tensor(spec.queries).cuda()
    torch.cuda.synchronize()
    print(f"[bench] scene ready: {args.config}, {len(q)} points", file=sys.stderr, flush=True)

    for w in range(args.warmup):
        solver.solve_grad(q, scene, seed=1000 + w)
    torch.cuda.synchronize()
    print(f"[bench] warmup done ({args.warmup})", file=sys.stderr, flush=True)

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    outs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        outs.append(solver.solve_grad(q, scene, seed=k, count_steps=True))
        ends[k].record()
        if args.progress:
            ends[k].synchronize()
            print(f"[bench] step {k} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    total_steps = sum(int(o["steps"][0]) for o in outs)
    value = total_steps / wall
    ms_per_step = 1000.0 * wall / args.steps
    print(f"[bench] {args.steps} steps in {wall:.3f}s", file=sys.stderr, flush=True)

    # accuracy of the last step against the FP64 analytic field
    u = outs[-1]["u"].cpu().numpy()
    g = outs[-1]["grad"].cpu().numpy()
    ua = scenes.analytic_u(spec.bc_code, spec.queries)
    ga = scenes.analytic_grad(spec.bc_code, spec.queries)
    R = scene.distance(q).cpu().numpy()
    far = R > 0.05
    acc = dict(rmse_u=float(np.sqrt(np.mean((u - ua) ** 2))), rmse_grad=float(np.sqrt(np.mean((g - ga) ** 2))),
               rmse_grad_far=float(np.sqrt(np.mean((g[far] - ga[far]) ** 2))),
               mean_steps_per_walk=total_steps / (args.steps * len(q) * 2 * ((cfg["nwalks"] + 1) // 2)))

    steps_per_launch = total_steps / args.steps
    avg_kernel_s = float(np.mean(kernel_ms)) / 1000.0
    achieved = B_ALG[args.config] * steps_per_launch / avg_kernel_s / 1e9
    prof = profile_figures(args.config)
    traffic = prof.get("traffic_bytes_per_walk_step")
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=None if traffic is None else round(traffic * steps_per_launch),
                kernel="k_fused_pers (HIP events bracket point prep + kernel + reduction; the kernel is >99.6%)",
                b_alg_bytes_per_step=B_ALG[args.config], avg_launch_ms=round(1000 * avg_kernel_s, 3),
                l2_hit_rate=prof.get("l2_hit_rate"), valu_busy=prof.get("valu_busy"),
                note="achieved = algorithmic bytes (reference record sizes) / launch time; traffic = measured HBM "
                     "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE per walk step x steps); the scene is "
                     "cache-resident, so VALU issue or fetch latency binds, not HBM")

    adj = None
    if args.adjoint_points:
        adj = adjoint_ratio(args.adjoint_points)
        print(f"[bench] adjoint ratio {adj['ratio']}", file=sys.stderr, flush=True)

    bat = None
    if args.batch_queries:
        bat = batch_speedup(args.batch_queries)
        print(f"[bench] batch speedup {bat['speedup']}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(spec, cfg)

    line = {
        "metric": "walk-steps/s (fused u + grad u) at fixed RMSE vs analytic",
        "value": round(value, 1),
        "unit": "walk-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene + queries)",
        "config": dict(workload=cfg["workload"], config=args.config, n_points=len(q), nwalks=cfg["nwalks"],
                       eps=cfg["eps"], nsteps=1024, grad_points_per_s=round(len(q) * args.steps / wall, 1),
                       **acc),
        "roofline": roof,
        "cpu_baseline": cpu,
        "adjoint": adj,
        "batch": bat,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
