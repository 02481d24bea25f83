"""Walk-on-Stars benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config ico|holes|star]
    python bench.py --alg-counts [--config ...]  # recount the algorithmic bytes / flops per walk step (CPU)
    python bench.py --copy-calibration           # the copy that scales the FETCH_SIZE / WRITE_SIZE counters

Headline workload (default ``ico``): the noisy icosphere (81,920 triangles,
faces with centroid y > 0 Dirichlet, g = e^x cos y, exact Neumann data on the
rest), 262,144 interior query points, 512 walks per point, eps 1e-3, at most
1024 steps per walk, fp32, fused u + grad u estimator.  A step is one solve of
one quarter of the query set (65,536 points, passed with point_base = its
offset, so their walk streams are those of the whole-set solve); the seed
changes every 4 steps, so each 4 consecutive steps are one complete solve of
all 262,144 points, bit-identical to a single launch over all of them.
Inputs are resident in HBM before timing starts.  Prints one JSON line; its
extra keys measure the 2D holes workload (4096-segment star with eight
64-segment Neumann holes, 65,536 lattice points x 1024 walks, eps 1e-4), the
3D adjoint (fwd+bwd / fwd), the 64-scene batch and the WoStGrad / WoStGrad3D
du/dx estimators (walks/s beside the CPU restatement on the same walks).
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic bytes and fp32 operations per walk step of the fused estimator,
# counted by the instrumented CPU restatement in the reference traversal order
# with the reference SoA record sizes (2D BVH node 24 B, 2D SNCH node 48 B, 3D
# SNCH node 60 B, segment 20 B, triangle 40 B, silhouette vertex 40 B,
# silhouette edge 68 B, CDF gathers 4 B per binary-search halving + 8 B) on
# 1024 seeded points of each workload at its walk count:
#     python bench.py --alg-counts --config <name>
# Steps from a NaN position (3D walks that left the domain) fetch nothing and
# count with 0 bytes; ff_share is their share of all steps.
ALG = {
    "star": dict(b_alg=3397.8, f_alg=2199.8, steps_per_walk=84.46, ff_share=0.0),
    "holes": dict(b_alg=4650.5, f_alg=2920.8, steps_per_walk=41.61, ff_share=0.0),
    "ico": dict(b_alg=40593.4, f_alg=28821.9, steps_per_walk=26.69, ff_share=5.5e-05),
}
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

CONFIGS = {
    "star": dict(workload="star polygon 256 segments (mixed D/N), g=e^x cos y, 1024 interior points", n_query=1024,
                 nwalks=256, eps=1e-3, shards=1),
    "holes": dict(workload="star polygon 4096 segments + 8 Neumann holes x 64, g=x^2-y^2, 256x256-lattice points",
                  n_query=65536, nwalks=1024, eps=1e-4, shards=1),
    "ico": dict(workload="noisy icosphere 81920 triangles (level 6), centroid y>0 Dirichlet, g=e^x cos y, "
                         "262144 interior points", n_query=262144, nwalks=512, eps=1e-3, shards=4),
}


def make_spec(name: str, n_query: int | None = None):
    from dwost_amd import scenes

    c = CONFIGS[name]
    n = n_query or c["n_query"]
    if name == "star":
        return scenes.star_scene_2d(n_query=n, eps=c["eps"])
    if name == "holes":
        return scenes.holes_scene_2d(n_query=n, eps=c["eps"])
    return scenes.icosphere_scene(level=6, n_query=n, eps=c["eps"])


def profile_figures(name: str) -> dict:
    """Figures of the committed rocprofv3 summary of this workload's kernel
    (profiles/k_fused_pers_<name>.md): HBM traffic per walk step from
    FETCH_SIZE / WRITE_SIZE (copy-calibrated), L2 hit rate, the share of wave
    time issuing VALU instructions, the active-lane fraction of issued VALU
    instructions and the VALU issue rate as a fraction of one wave64
    instruction per SIMD every 4 cycles at 2.4 GHz."""
    path = os.path.join(ROOT, "profiles", f"k_fused_pers_{name}.md")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            k, sep, v = line.strip().partition(": ")
            if sep and k in ("traffic_bytes_per_walk_step", "l2_hit_rate", "valu_issue_share", "valu_lane_activity",
                             "valu_issue_rate_frac"):
                out[k] = float(v)
    return out


def alg_counts(name: str, n_points: int = 1024) -> dict:
    """B_alg / F_alg of the fused estimator: the oracle's walk counters on
    n_points seeded points of the workload at its walk count (CPU only)."""
    import numpy as np

    from oracle import oracle as O

    O.build()
    c = CONFIGS[name]
    spec = make_spec(name)
    sc = O.OracleScene(spec.vertices, spec.indices, spec.types)
    idx = np.sort(np.random.default_rng(0).choice(len(spec.queries), min(n_points, len(spec.queries)), replace=False))
    t0 = time.perf_counter()
    o = sc.solve_fused(spec.queries[idx], spec.bc_code, c["nwalks"], nsteps=1024, eps=c["eps"], seed=0, count=True)
    dt = time.perf_counter() - t0
    npairs = (c["nwalks"] + 1) // 2
    st = o["steps"]
    return dict(config=name, points=len(idx), nwalks=c["nwalks"], steps=st, b_alg=round(o["bytes"] / st, 1),
                f_alg=round(o["flops"] / st, 1), steps_per_walk=round(st / (len(idx) * 2 * npairs), 2),
                ff_share=round(o["ff"] / st, 6), nodes_per_step=round(o["nodes"] / st, 1),
                prims_per_step=round(o["prims"] / st, 2), sils_per_step=round(o["sils"] / st, 2),
                closest_point_query=closest_iters(o), cpu_seconds=round(dt, 1), threads=O.num_threads())


def closest_iters(o: dict) -> dict:
    """Per closest-point query: the GPU loop's iterations and interior-record
    fetches at node width 2 and 4 (replayed from the oracle's traversal)."""
    q = max(o["closest_queries"], 1)
    return {f"width{w}": dict(iterations=round(o[f"iters{w}"] / q, 2), record_fetches=round(o[f"fetch{w}"] / q, 2))
            for w in (2, 4)}


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


def timed_solves(solver, scene, q, steps: int, shards: int, seed0: int, count: bool, progress: bool, tag: str):
    """`steps` fused solves, step k over shard k % shards with seed seed0 + k // shards.
    Returns (outputs, per-launch HIP-event ms, wall seconds)."""
    import torch

    n = q.shape[0]
    m = (n + shards - 1) // shards
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    outs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        s = k % shards
        lo, hi = s * m, min(n, (s + 1) * m)
        starts[k].record()
        o = solver.solve_grad(q[lo:hi], scene, seed=seed0 + k // shards, count_steps=count, point_base=lo)
        ends[k].record()
        outs.append((lo, hi, o))
        if progress:
            ends[k].synchronize()
            print(f"[bench] {tag} step {k} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return outs, [a.elapsed_time(b) for a, b in zip(starts, ends)], wall


def accuracy(spec, scene, q, outs, shards: int) -> dict:
    """RMSE of u and grad u against the FP64 analytic field over the first
    complete solve (steps 0 .. shards-1: every point once, one seed)."""
    import numpy as np

    from dwost_amd import scenes

    n = q.shape[0]
    u = np.full(n, np.nan)
    g = np.full((n, spec.dim), np.nan)
    for lo, hi, o in outs[:shards]:
        u[lo:hi] = o["u"].cpu().numpy()
        g[lo:hi] = o["grad"].cpu().numpy()
    cov = np.isfinite(u)
    p = spec.queries[cov]
    ua = scenes.analytic_u(spec.bc_code, p)
    ga = scenes.analytic_grad(spec.bc_code, p)
    R = scene.distance(q).cpu().numpy()[cov]
    far = R > 0.05
    return dict(rmse_points=int(cov.sum()), rmse_u=float(np.sqrt(np.mean((u[cov] - ua) ** 2))),
                rmse_grad=float(np.sqrt(np.mean((g[cov] - ga) ** 2))),
                rmse_grad_far=float(np.sqrt(np.mean((g[cov][far] - ga[far]) ** 2))))


def secondary_2d(progress: bool, reps: int = 2) -> dict:
    """The 2D holes workload (65,536 lattice points x 1024 walks, eps 1e-4):
    `reps` full fused solves after one warm-up."""
    import numpy as np

    import dwost_amd

    c = CONFIGS["holes"]
    spec = make_spec("holes")
    import torch

    scene = dwost_amd.Scene2D(spec.vertices, spec.indices, spec.types, dwost_amd.BoundaryCondition(spec.bc_code))
    solver = dwost_amd.WoSt(eps=c["eps"], nwalks=c["nwalks"], nsteps=1024)
    q = torch.as_tensor(spec.queries).cuda()
    solver.solve_grad(q, scene, seed=999)
    outs, ms, wall = timed_solves(solver, scene, q, reps, 1, 0, True, progress, "holes")
    steps = sum(int(o["steps"][0]) for _, _, o in outs)
    avg_s = float(np.mean(ms)) / 1000.0
    prof = profile_figures("holes")
    ach = ALG["holes"]["b_alg"] * steps / reps / avg_s / 1e9
    acc = accuracy(spec, scene, q, outs, 1)
    return dict(workload=c["workload"], n_points=len(q), nwalks=c["nwalks"], eps=c["eps"],
                walk_steps_per_s=round(steps / wall, 1), grad_points_per_s=round(reps * len(q) / wall, 1),
                ms_per_solve=round(1000 * wall / reps, 3), avg_launch_ms=round(1000 * avg_s, 3),
                roofline_frac=round(ach / HBM_PEAK_GBS, 4), achieved_gbs=round(ach, 1),
                b_alg_bytes_per_step=ALG["holes"]["b_alg"], l2_hit_rate=prof.get("l2_hit_rate"),
                valu_issue_share=prof.get("valu_issue_share"), valu_lane_activity=prof.get("valu_lane_activity"),
                valu_issue_rate_frac=prof.get("valu_issue_rate_frac"),
                mean_steps_per_walk=round(steps / (reps * len(q) * c["nwalks"]), 3), **acc)


def wostgrad_measure(points_2d: int = 65536, points_3d: int = 4096, nwalks: int = 64, cpu_s: float = 6.0) -> dict:
    """WoStGrad on the 256-segment star polygon (mixed Dirichlet / Neumann,
    u = e^x cos y) and WoStGrad3D on the 20,480-triangle icosphere (all
    Neumann, u = x): du/dx at interior points x `nwalks` walks, nsteps 1024,
    one timed solve each after a short warm-up; the CPU restatement (same
    walks, same seeds) on a sample of the same points, all host cores."""
    import numpy as np
    import torch

    import dwost_amd
    from dwost_amd import scenes
    from oracle import oracle as O

    O.build()
    out = {}
    for dim in (2, 3):
        if dim == 2:
            spec = scenes.star_scene_2d(n_query=points_2d)
            types, code, coeffs = spec.types, spec.bc_code, (0.0, 0.0, 0.0, 0.0)
            scene = dwost_amd.Scene2D(spec.vertices, spec.indices, types, dwost_amd.BoundaryCondition(code))
            solver = dwost_amd.WoStGrad(eps=1e-3, nwalks=nwalks, nsteps=1024)
            workload = f"star polygon 256 segments, mixed, u = e^x cos y, {points_2d} points x {nwalks} walks"
        else:
            spec = scenes.icosphere_scene(level=5, n_query=points_3d)
            types, code, coeffs = np.zeros_like(spec.types), 2, (0.0, 1.0, 0.0, 0.0)
            scene = dwost_amd.Scene3D(spec.vertices, spec.indices, types, dwost_amd.BoundaryCondition(code, coeffs))
            solver = dwost_amd.WoStGrad3D(eps=1e-3, nwalks=nwalks, nsteps=1024)
            workload = f"icosphere 20480 triangles, all Neumann, u = x, {points_3d} points x {nwalks} walks"
        pts = np.ascontiguousarray(spec.queries.astype(np.float32))
        e = np.zeros_like(pts)
        e[:, 0] = 1.0
        q, qe = torch.as_tensor(pts).cuda(), torch.as_tensor(e).cuda()
        warm = dataclasses.replace(solver, nsteps=4, nwalks=1)
        warm.solve_dude(q[:64], qe[:64], scene, seed=7)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = solver.solve_dude(q, qe, scene, seed=1, count_steps=True, per_walk=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ws = r["walk_steps"].cpu().numpy()
        steps = int(r["steps"][0])
        dude = r["dude"].cpu().numpy().astype(np.float64)
        x = pts.astype(np.float64)
        exact = np.exp(x[:, 0]) * np.cos(x[:, 1]) if dim == 2 else np.ones(len(x))
        # the CPU restatement on the first k points, k grown until ~cpu_s seconds
        sc = O.OracleScene(spec.vertices, spec.indices, types)
        k, ct, cw, same = 16, 0.0, 0, []
        while ct < cpu_s and cw < len(pts):
            lo, hi = cw, min(len(pts), cw + k)
            t1 = time.perf_counter()
            c = sc.grad_dude(pts[lo:hi], e[lo:hi], code, nwalks, nsteps=1024, eps=1e-3, min_R=solver.min_R,
                             rho_max=solver.rho_max, coeffs=coeffs, seed=1, base=lo, per_walk=True)
            ct += time.perf_counter() - t1
            gw = r["walk_dude"][lo * nwalks:hi * nwalks].cpu().numpy()
            same.append(c["walk_dude"].view(np.uint32) == gw.view(np.uint32))
            cw = hi
            k *= 2
        cpu_walks = cw * nwalks / ct
        out[f"{dim}d"] = dict(
            workload=workload, walks_per_s=round(len(pts) * nwalks / wall, 1), walk_steps_per_s=round(steps / wall, 1),
            ms=round(1000 * wall, 1), mean_steps_per_walk=round(float(ws.mean()), 2), max_steps=int(ws.max()),
            walks_at_step_cap=int((ws >= 1024).sum()), rmse_vs_analytic=float(np.sqrt(np.mean((dude - exact) ** 2))),
            seed=1, note="walk lengths are heavy-tailed: a launch lasts at least as long as its longest walk, "
                         "which one lane runs alone (3D: 3.8-4.2 s on most seeds, ~14 s on seeds whose longest "
                         "walk spends 1,024 steps traversing most of the tree)",
            cpu_baseline=dict(walks_per_s=round(cpu_walks, 1), cores=O.num_threads(), kind="port",
                              sample=f"first {cw} points x {nwalks} walks ({ct:.1f} s), same seeds, OpenMP"),
            cpu_walks_bit_identical=float(np.concatenate(same).mean()),
            gpu_over_cpu=round(len(pts) * nwalks / wall / cpu_walks, 1))
        print(f"[bench] wostgrad {dim}d {out[f'{dim}d']['walks_per_s']:.3e} walks/s "
              f"({out[f'{dim}d']['gpu_over_cpu']}x CPU)", file=sys.stderr, flush=True)
    return out


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
    return dict(workload=f"noisy icosphere 81920 triangles, 40962 vertices with per-vertex g/h, {points} points x "
                         "256 walks, loss MSE(u)+MSE(grad u)", fwd_ms=round(1000 * t_f, 3),
                fwd_bwd_ms=round(1000 * t_fb, 3), ratio=round(t_fb / t_f, 3))


def batch_speedup(max_queries: int | None, reps: int = 1):
    """One persistent launch over the 64-scene uneven batch (segments
    log-U[64, 16384], 0-4 holes, walks {64, 256, 1024}, queries log-U[100, 1e5],
    optionally capped) against the same solves as per-scene launches."""
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
    cap = "uncapped" if max_queries is None else f"capped at {max_queries}"
    return dict(workload=f"64 uneven 2D scenes, {sum(len(s.queries) for s in specs)} points (queries {cap}), "
                         "walks per point from {64, 256, 1024}",
                per_scene_ms=round(1000 * t_s, 3), batched_graph_ms=round(1000 * t_b, 3),
                speedup=round(t_s / t_b, 3), walk_steps_per_s=round(steps / t_b, 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12, help="timed steps (default 12: three complete icosphere solves, about a minute)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="ico", choices=sorted(CONFIGS))
    ap.add_argument("--shards", type=int, default=0, help="steps per complete solve (0: the workload's default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--points", type=int, default=0, help="use only the first N query points (0: all)")
    ap.add_argument("--progress", action="store_true", help="print a line after each timed step")
    ap.add_argument("--no-2d", action="store_true", help="skip the 2D holes measurement")
    ap.add_argument("--no-wostgrad", action="store_true", help="skip the WoStGrad / WoStGrad3D measurement")
    ap.add_argument("--batch-queries", type=int, default=-1,
                    help="query cap per scene of the 64-scene batched-vs-per-scene measurement "
                         "(-1: uncapped, 0: skip)")
    ap.add_argument("--adjoint-points", type=int, default=16384,
                    help="points of the fwd+bwd/fwd measurement on the 3D adjoint workload (0: skip)")
    ap.add_argument("--alg-counts", action="store_true",
                    help="count algorithmic bytes / flops per walk step with the oracle on 1024 points (CPU), print "
                         "them and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 256 MiB device-to-device copies (scales the FETCH_SIZE / WRITE_SIZE "
                         "counters) and exit")
    args = ap.parse_args()
    if args.alg_counts:
        print(json.dumps(alg_counts(args.config)), flush=True)
        return
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

    cfg = CONFIGS[args.config]
    shards = args.shards or cfg["shards"]
    spec = make_spec(args.config)
    if args.points:
        spec.queries = spec.queries[: args.points]
    Scene = dwost_amd.Scene2D if spec.dim == 2 else dwost_amd.Scene3D
    scene = Scene(spec.vertices, spec.indices, spec.types, dwost_amd.BoundaryCondition(spec.bc_code))
    Solver = dwost_amd.WoSt if spec.dim == 2 else dwost_amd.WoSt3D
    solver = Solver(eps=cfg["eps"], nwalks=cfg["nwalks"], nsteps=1024)
    q = torch.as_tensor(spec.queries).cuda()
    torch.cuda.synchronize()
    print(f"[bench] scene ready: {args.config}, {len(q)} points, {shards} shards", file=sys.stderr, flush=True)

    timed_solves(solver, scene, q, args.warmup, shards, 1000, False, args.progress, "warmup")
    print(f"[bench] warmup done ({args.warmup})", file=sys.stderr, flush=True)
    outs, kernel_ms, wall = timed_solves(solver, scene, q, args.steps, shards, 0, True, args.progress, "timed")
    total_steps = sum(int(o["steps"][0]) for _, _, o in outs)
    value = total_steps / wall
    ms_per_step = 1000.0 * wall / args.steps
    pts_per_step = sum(hi - lo for lo, hi, _ in outs) / args.steps
    print(f"[bench] {args.steps} steps in {wall:.3f}s", file=sys.stderr, flush=True)

    acc = accuracy(spec, scene, q, outs, shards)
    npairs = (cfg["nwalks"] + 1) // 2
    acc["mean_steps_per_walk"] = total_steps / (args.steps * pts_per_step * 2 * npairs)

    alg = ALG[args.config]
    steps_per_launch = total_steps / args.steps
    avg_kernel_s = float(np.mean(kernel_ms)) / 1000.0
    achieved = alg["b_alg"] * steps_per_launch / avg_kernel_s / 1e9
    prof = profile_figures(args.config)
    traffic = prof.get("traffic_bytes_per_walk_step")
    kname = f"k_fused_pers<{spec.dim}, false, false, {'true' if spec.dim == 3 else 'false'}>"
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=None if traffic is None else round(traffic * steps_per_launch),
                kernel=f"{kname} (HIP events bracket point prep + kernel + reduction; the kernel is >99% of it)",
                b_alg_bytes_per_step=alg["b_alg"], avg_launch_ms=round(1000 * avg_kernel_s, 3),
                f_alg_flops_per_step=alg["f_alg"],
                achieved_tflops=round(alg["f_alg"] * steps_per_launch / avg_kernel_s / 1e12, 3),
                fp32_peak_tflops=FP32_PEAK_TFLOPS, ff_share=alg["ff_share"],
                l2_hit_rate=prof.get("l2_hit_rate"), valu_issue_share=prof.get("valu_issue_share"),
                valu_lane_activity=prof.get("valu_lane_activity"),
                valu_issue_rate_frac=prof.get("valu_issue_rate_frac"),
                limiter="VALU issue with divergent lanes: the walk kernel issues VALU instructions at about the "
                        "chip's wave64 rate (valu_issue_rate_frac) with few lanes active per instruction "
                        "(valu_lane_activity); the scene is L2/MALL resident and HBM bandwidth does not bind",
                note="achieved = algorithmic bytes (reference record sizes, reference traversal order) / launch "
                     "time; traffic = measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE per walk "
                     "step x steps)")

    holes = None
    if not args.no_2d and args.config != "holes":
        holes = secondary_2d(args.progress)
        print(f"[bench] holes {holes['walk_steps_per_s']:.3e} walk-steps/s", file=sys.stderr, flush=True)

    adj = None
    if args.adjoint_points:
        adj = adjoint_ratio(args.adjoint_points)
        print(f"[bench] adjoint ratio {adj['ratio']}", file=sys.stderr, flush=True)

    bat = None
    if args.batch_queries:
        bat = batch_speedup(None if args.batch_queries < 0 else args.batch_queries)
        print(f"[bench] batch speedup {bat['speedup']}", file=sys.stderr, flush=True)

    grad = None
    if not args.no_wostgrad:
        grad = wostgrad_measure()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(spec, cfg)
        cpu["gpu_over_cpu"] = round(value / cpu["value"], 1)

    line = {
        "metric": "walk-steps/s + grad-u points/s at fixed RMSE vs analytic; fwd+bwd/fwd time ratio",
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
                       eps=cfg["eps"], nsteps=1024, points_per_step=int(pts_per_step), steps_per_solve=shards,
                       grad_points_per_s=round(pts_per_step * args.steps / wall, 1),
                       rmse_note="every walk equals the reference estimator's walk bit for bit (GPU parity "
                                 "tests), so at the same walks per point the RMSE is the reference algorithm's",
                       **acc),
        "roofline": roof,
        "cpu_baseline": cpu,
        "holes_2d": holes,
        "adjoint": adj,
        "batch": bat,
        "wostgrad": grad,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
