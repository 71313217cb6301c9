"""Frame-rate benchmarks of the batched and served forms on one GPU.

    python -m pycpet_amd.framebench batch --frames 64
    python -m pycpet_amd.framebench serve --rate 200 --duration 10
    python -m pycpet_amd.framebench trajectory --frames 2000

batch: the mixed MD frame set (pycpet_amd.workloads.md_frames: half volume
grids of edge 21/41/61/81, half 20,000-start streamline jobs, 3k-40k atoms
per frame), inputs resident in HBM, timed three ways:
  plain     one launch per frame, the simple kernel shape: one lane per grid
            point / streamline, one wave per workgroup, no charge split, no
            slot refill;
  per-frame one launch per frame with the tuned single-frame kernels;
  batched   all volume frames in one launch and all streamline frames in one
            persistent launch, on two streams;
  batched_graph  the same two batched launches captured as parallel
            branches of one HIP graph (pycpet_amd.graphs) and replayed.
trajectory: many small frames of one shape (2,000 charges, 21^3 grid), each
computed eagerly or by replaying one captured HIP graph.
serve: Poisson arrivals (pycpet_amd.serve.poisson_requests) served with
continuous batching vs one request at a time (stopped after --plain-max-s,
unserved requests counted with their latency lower bound); p50/p99 latency
and frames/s.
Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def _frames_to_device(frames, torch, dev):
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = []
    for f in frames:
        g = {"kind": f["kind"], "x": T(f["x"]), "q": T(f["q"])}
        if f["kind"] == "volume":
            g["axes"] = tuple(T(a) for a in f["axes"])
        else:
            g.update(starts=T(f["starts"]), n_iter=T(f["n_iter"]), h=f["h"], dims=f["dims"])
        out.append(g)
    return out


def run_batch(n_frames: int, topo_starts: int, seed: int = 3,
              forms=("batched", "batched_graph", "per_frame", "plain"),
              autotune: bool = True) -> dict:
    """Frames/s of the mixed MD frame set in the chosen forms: batched (one
    volume and one persistent topo launch on two streams), the same as one
    HIP graph, the per-frame calls at their tuned shapes, and the de-tuned
    plain per-frame shape; the batched launch shapes come from the
    autotuner, or from the built-in defaults with autotune=False."""
    import torch

    from . import _native as nat, batch as B, ops, workloads as W

    dev = torch.device("cuda:0")
    frames = W.md_frames(n_frames, seed=seed, topo_starts_per_frame=topo_starts)
    dframes = _frames_to_device(frames, torch, dev)
    vol = [f for f in dframes if f["kind"] == "volume"]
    top = [f for f in dframes if f["kind"] == "topo"]
    ctx = nat.Context(0)

    def per_frame(plain: bool):
        t0 = time.perf_counter()
        for i, f in enumerate(dframes):
            if i and i % 64 == 0:  # progress for long runs (512 frames plain: minutes)
                print(f"{'plain' if plain else 'per_frame'}: {i}/{len(dframes)} frames, "
                      f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            if f["kind"] == "volume":
                if plain:
                    ctx.set_tuning(0, 1, 1)
                ops.field_grid(*f["axes"], f["x"], f["q"], ctx=ctx)
            else:
                if plain:
                    ctx.set_tuning(2, 1, (f["starts"].shape[0] + 63) // 64)
                ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"], ctx=ctx)
            torch.cuda.synchronize()
        ctx.set_tuning(0, 0, 0)
        ctx.set_tuning(2, 0, 0)

    ctx2 = nat.Context(0)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    results = {}

    def batched():
        with torch.cuda.stream(s1):
            results["vol"] = B.batch_volume(vol, ctx=ctx)
        with torch.cuda.stream(s2):
            results["top"] = B.batch_topo(top, ctx=ctx2)
        torch.cuda.synchronize()

    # warm-up (allocations, code objects)
    small = [f for f in dframes[:2]]
    B.batch_volume([f for f in small if f["kind"] == "volume"], ctx=ctx)
    B.batch_topo([f for f in small if f["kind"] == "topo"], ctx=ctx2)
    torch.cuda.synchronize()

    # launch shapes of the batched calls from the autotuner, timed on a subset
    from . import autotune as AT

    kv, kt = (0, 0), (0, 0)
    if autotune:
        kv = AT.tune_batch_volume(ctx, vol[:8])
        kt = AT.tune_batch_topo(ctx2, top[:4])
        ctx.set_profiling(False)
        ctx2.set_profiling(False)
        print(f"autotuned: batch volume K={kv[0]}, batch topo S,G={kt}", file=sys.stderr,
              flush=True)

    gs = None
    if "batched_graph" in forms:
        # both batched launches as parallel branches of one HIP graph
        from . import graphs as G

        gs = G.CapturedStep([lambda c: B.batch_volume(vol, ctx=c),
                             lambda c: B.batch_topo(top, ctx=c)],
                            tuning=[{4: kv}, {3: kt}])

    def batched_graph():
        gs.replay()
        torch.cuda.synchronize()

    timings = {}
    for name, fn in (("batched", batched), ("batched_graph", batched_graph),
                     ("per_frame", lambda: per_frame(False)),
                     ("plain", lambda: per_frame(True))):
        if name not in forms:
            continue
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        timings[name] = time.perf_counter() - t0
        print(f"{name}: {timings[name]:.3f} s", file=sys.stderr, flush=True)
    vpairs = sum(int(np.prod([len(a) for a in f["axes"]])) * f["q"].numel() for f in vol)
    tpairs = sum((int(r.steps.sum().item()) + 4 * r.steps.numel()) * f["q"].numel()
                 for r, f in zip(results["top"], top))
    out = {"workload": "md_frames", "frames": n_frames, "volume_frames": len(vol),
           "autotuned": ({"batch_volume_K": kv[0], "batch_topo_S_G": list(kt)}
                         if autotune else None),
           "topo_frames": len(top), "topo_starts_per_frame": topo_starts,
           "pair_evals": vpairs + tpairs}
    for k, v in timings.items():
        out[f"{k}_s"] = v
        out[f"{k}_frames_per_s"] = n_frames / v
    def maxdiff(a_list, b_list, get):
        return max(float((get(a) - get(b)).abs().max()) for a, b in zip(a_list, b_list))

    if gs is not None:
        # the graph replay recomputes the eager batched results exactly; a
        # second eager run separates run-to-run differences from graph ones
        first = results.copy()
        batched_graph()
        batched()
        out["eager_rerun_maxdiff"] = max(maxdiff(first["vol"], results["vol"], lambda r: r[0]),
                                         maxdiff(first["top"], results["top"], lambda r: r.feats))
        gv_out, gt_out = gs.outputs
        out["graph_vs_eager_maxdiff"] = max(maxdiff(gv_out, results["vol"], lambda r: r[0]),
                                            maxdiff(gt_out, results["top"], lambda r: r.feats))
    if "plain" in timings:
        out["batched_vs_plain"] = timings["plain"] / timings["batched"]
    if "plain" in timings and "batched_graph" in timings:
        out["graph_vs_plain"] = timings["plain"] / timings["batched_graph"]
    if "per_frame" in timings:
        out["batched_vs_per_frame"] = timings["per_frame"] / timings["batched"]
    out["batched_pair_evals_per_s"] = (vpairs + tpairs) / timings["batched"]
    return out


def run_trajectory(n_frames: int, seed: int = 0) -> dict:
    """A trajectory of small frames of one shape (2,000 charges, 21^3 grid,
    field and potential): each frame's coordinates are copied into the same
    device tensors and the frame is computed eagerly (ops.field_grid) or by
    replaying one captured HIP graph (pycpet_amd.graphs.CapturedStep).  At
    this size the launches, not the arithmetic, set the frame time."""
    import torch

    from . import _native as nat, graphs as G, ops, workloads as W

    c = W.volume_2k()
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ax = [T(a) for a in c.axes]
    x, q = T(c.x), T(c.q)
    rng = np.random.default_rng(seed)
    traj = T(np.stack([c.x + rng.normal(0, 0.05, c.x.shape).astype(np.float32)
                       for _ in range(min(n_frames, 64))]))
    ctx = nat.Context(0)
    ctx.set_profiling(False)
    step = G.CapturedStep(lambda cx: ops.field_grid(*ax, x, q, want_phi=True, ctx=cx))
    out = {"workload": "trajectory_2k_charges_21cubed", "frames": n_frames}
    for name in ("eager", "graph"):
        for _ in range(3):  # warm-up
            if name == "eager":
                ops.field_grid(*ax, x, q, want_phi=True, ctx=ctx)
            else:
                step.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(n_frames):
            x.copy_(traj[i % traj.shape[0]])
            if name == "eager":
                ops.field_grid(*ax, x, q, want_phi=True, ctx=ctx)
            else:
                step.replay()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[f"{name}_us_per_frame"] = 1e6 * dt / n_frames
        out[f"{name}_frames_per_s"] = n_frames / dt
        print(f"{name}: {1e6 * dt / n_frames:.1f} us per frame", file=sys.stderr, flush=True)
    # the last replay equals an eager run on the same coordinates
    E, phi = step.replay()
    Ee, pe = ops.field_grid(*ax, x, q, want_phi=True, ctx=ctx)
    torch.cuda.synchronize()
    out["graph_equals_eager"] = bool(torch.equal(E, Ee) and torch.equal(phi, pe))
    out["graph_speedup"] = out["eager_us_per_frame"] / out["graph_us_per_frame"]
    step.close()
    return out


def run_serve(rate: float, duration: float, lanes: int, seed: int = 0,
              plain_max_s: float = 120.0, policy: str = "fifo-small") -> dict:
    import torch

    from . import serve

    reqs = serve.poisson_requests(rate, duration, seed=seed)
    warm = serve.poisson_requests(50.0, 0.2, seed=seed + 1, max_requests=6)
    srv = serve.Server(n_lanes=lanes, hbm_budget=64 << 30, policy=policy)
    srv.run(warm)
    st = srv.run(reqs, progress=lambda m: print(m, file=sys.stderr, flush=True),
                 keep_results=False).summary()
    print(f"continuous batching: {st}", file=sys.stderr, flush=True)
    for r in reqs:
        r.result = None
    torch.cuda.synchronize()
    pl = serve.serve_plain(reqs, max_wall_s=plain_max_s,
                           progress=lambda m: print(m, file=sys.stderr, flush=True),
                           keep_results=False).summary()
    print(f"plain: {pl}", file=sys.stderr, flush=True)
    kinds = {k: sum(r.kind == k for r in reqs) for k in ("volume", "volume_ESP", "topo")}
    return {"workload": "poisson_serving", "rate_req_per_s": rate, "duration_s": duration,
            "requests": len(reqs), "mix": kinds, "lanes": lanes, "policy": policy,
            "offered_frames_per_s": len(reqs) / duration,
            "continuous_batching": st, "plain_one_at_a_time": pl,
            "throughput_ratio": st["frames_per_s"] / max(pl["frames_per_s"], 1e-9),
            "p99_ratio_plain_over_batched": pl["p99_ms"] / max(st["p99_ms"], 1e-9)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("batch")
    b.add_argument("--frames", type=int, default=64)
    b.add_argument("--topo-starts", type=int, default=20000)
    s = sub.add_parser("serve")
    s.add_argument("--rate", type=float, default=200.0)
    s.add_argument("--duration", type=float, default=10.0)
    s.add_argument("--lanes", type=int, default=4)
    s.add_argument("--policy", default="fifo-small", choices=("fifo-small", "greedy"))
    s.add_argument("--plain-max-s", type=float, default=120.0,
                   help="stop the one-at-a-time baseline after this wall time")
    t = sub.add_parser("trajectory")
    t.add_argument("--frames", type=int, default=2000)
    a = ap.parse_args(argv)
    if a.cmd == "batch":
        out = run_batch(a.frames, a.topo_starts)
    elif a.cmd == "trajectory":
        out = run_trajectory(a.frames)
    else:
        out = run_serve(a.rate, a.duration, a.lanes, plain_max_s=a.plain_max_s, policy=a.policy)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
