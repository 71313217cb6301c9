#!/usr/bin/env python
"""Benchmark: streamline topology over a 5,000-charge frame on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One "step" is one full pass of the streamline hot path (order, propagate,
FP64 epilogue with the 20 x 20 feature histogram) over 100,000 starts with
inputs already resident in HBM: 5,000 charges (35 A ball, 3 A exclusion
around the +-1.5 A box), h = 0.001 A, Euler, n_iter = randint(1, max_steps)
(seed 42) capped at 2,000.  The histogram's range is the min / max of the
FP64 features of the same streamlines (one untimed FP64 run first), as the
reference's histogram step takes it from the data (make_histograms,
CPET/utils/calculator.py:702-707).

Prints ONE JSON line.  `value` is charge-point pair-evals/s, counted as
(Euler steps + 4 curvature steps) x charges per streamline, as the reference
does its work (thread_operation, CPET/utils/math_module.c:523-591).
`roofline` is for the dominant kernel (topo_propagate) from HIP events on
the stream it runs on: 18 FLOP per charge-point pair x charges x Euler steps
per launch / the kernel's mean duration, against the FP32 vector peak.
`cpu_baseline` times the reference's own thread_operation (built from its
source under oracle/_ref) on a bounded seeded sample of the same streamlines
on the host cores, or the oracle's FP32 port when that build is absent; its
`volume` and `esp` entries time the reference's compute_looped_field and
calc_esp_base on one core (the 2,000-charge 21^3 grid in full and a 21^3
subset of the 30,000-charge 101^3 grid; ESP with and without the per-point
Python loop).
Secondary lines: `volume_30k` (the 101^3 grid over 30,000 clustered
charges) and `volume_esp_30k` (the same grid, potential only; both
--no-volume), `topo_f64` (the FP64 run: its speed and the FP32 path's match
fraction and feature errors against it; --no-f64 drops the line) and
`topo_rk4` (the same streamlines integrated with RK4, four field
evaluations per step; --no-rk4).  `aux` (--no-aux): point-dipole
streamlines, the residue breakdown, frame I/O and the cluster step, each
beside its CPU counterpart.  `forms` (--no-forms): frames/s of 32 mixed MD
frames batched against per-frame calls, and p50/p99 latency of 10 s of
Poisson serving at 200 req/s against one request at a time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "charge-point pair-evals/s + streamlines/s; frames/s and p50/p99 latency"
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 vector, AMD-published
PEAK_FP32_UNPACKED_TFLOPS = 78.6  # without packed FP32 (disabled in this build)
FLOP_PER_PAIR = 18


TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_topo5k.txt"


def measured_traffic():
    """HBM bytes per topo_propagate launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the same
    counter's reading for a plain copy of known size), or None."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(case, gpu_steps, seconds: float, cores: int):
    """Reference thread_operation on a seeded sample of the same streamlines."""
    from oracle import oracle as O

    rng = np.random.default_rng(123)
    order = rng.permutation(case.L)
    if O.Ref.available():
        ref = O.Ref()
        kind = "reference"
        run_one = lambda i: ref.thread_operation(case.starts[i], case.n_iter[i], case.h,
                                                 case.dims, case.x, case.q)
    else:
        kind = "port"
        run_one = lambda i: O.topo_f32(case.starts[i:i + 1], case.n_iter[i:i + 1], case.h,
                                       case.dims, case.x, case.q)
    from concurrent.futures import ThreadPoolExecutor

    done = []
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        pos = 0
        while time.perf_counter() - t0 < seconds and pos < case.L:
            batch = order[pos:pos + 4 * cores]
            list(ex.map(run_one, batch))
            done.extend(batch.tolist())
            pos += len(batch)
    dt = time.perf_counter() - t0
    done = np.array(done)
    pairs = float((gpu_steps[done].astype(np.int64) + 4).sum()) * case.N
    return {
        "value": pairs / dt,
        "unit": "pair-evals/s",
        "cores": cores,
        "kind": kind,
        "sample": f"{len(done)} of {case.L} streamlines (seeded), {dt:.1f} s wall; "
                  "pair-evals counted with the GPU's per-streamline step counts; "
                  "one reference call per streamline from a thread pool (ctypes releases "
                  "the GIL; the reference fans the same calls over a process Pool)",
        "streamlines_per_s": len(done) / dt,
    }


def cpu_baseline_grids():
    """Reference compute_looped_field and calc_esp_base on one core: the
    2,000-charge 21^3 grid in full, and the 21^3 subset (every 5th point per
    axis) of the 101^3 grid over 30,000 clustered charges.  ESP is timed with
    the reference's per-point Python loop and with the same calls from C."""
    from oracle import oracle as O
    from pycpet_amd import workloads as W

    if not O.Ref.available():
        return None
    ref = O.Ref()
    c0 = W.volume_2k()
    c2 = W.volume_30k()
    sub = [a[::5] for a in c2.axes]
    cases = {"volume_2k_21^3": (c0.mesh_points(), c0.x, c0.q),
             "volume_30k_21^3_of_101^3": (
                 np.stack(np.meshgrid(*sub, indexing="ij"), -1).reshape(-1, 3).astype(np.float32),
                 c2.x, c2.q)}
    out = {"volume": {}, "esp_python_loop": {}, "esp_c_loop": {}}
    for name, (pts, x, q) in cases.items():
        pairs = float(pts.shape[0]) * q.size
        for key, fn in (("volume", ref.compute_looped_field),
                        ("esp_python_loop", ref.esp_grid_python_loop),
                        ("esp_c_loop", ref.esp_grid_c_loop)):
            t0 = time.perf_counter()
            fn(pts, x, q)
            dt = time.perf_counter() - t0
            out[key][name] = {"pair_evals_per_s": pairs / dt, "seconds": dt,
                              "points": int(pts.shape[0]), "charges": int(q.size), "cores": 1}
    return out


def _write_pdb(path, x, q):
    """A frame in the fixed-column PDB layout the reference parser reads
    (coordinates in columns 31-54, charge in 56-64; CPET/utils/io.py:574-636)."""
    with open(path, "w") as fh:
        for i in range(len(q)):
            fh.write(f"ATOM  {(i % 99999) + 1:5d} C{i % 5:<3d} ALA A{1 + (i // 5) % 9999:4d}    "
                     f"{x[i, 0]:8.3f}{x[i, 1]:8.3f}{x[i, 2]:8.3f} {q[i]:8.4f}  1.8240       C  \n")
        fh.write("END\n")


def _best_of(fn, k=3):
    """(shortest wall time of k calls, the last call's result)."""
    best, res = None, None
    for _ in range(k):
        t1 = time.perf_counter()
        res = fn()
        dt = time.perf_counter() - t1
        best = dt if best is None else min(best, dt)
    return best, res


def aux_rows(ctx, T, case, cpu: bool, cores: int):
    """The rows that feed and follow the two hot paths, each timed on the GPU
    (or native host code) beside its CPU counterpart:
      dipole_topo_5k  point-dipole streamlines (thread_operation_dipole,
                      CPET/utils/math_module.c:593-659): the 5,000 positions of
                      the main workload as dipoles, the same 100,000 starts;
      residue_breakdown_200x20k
                      200 residues of 8-24 charges, each alone over the same
                      20,000 starts, batched with 20x20 histograms
                      (scripts/residue_breakdown_analysis.py:160-188);
      frame_io        PDB parse (CPET/utils/io.py:574-636) of a 200,000-atom
                      frame, .top writing (np.savetxt, CPET/source/CPET.py:123)
                      and reading of 1,000,000 rows;
      driver_32_frames
                      the options-file driver (CPET/source/CPET.py:96-170)
                      end to end over 32 frame files, topo and volume;
      cluster_step    make_histograms + construct_distance_matrix
                      (CPET/utils/calculator.py:596-718,1000-1015) over 256
                      frames of 20,000 features."""
    import tempfile

    import torch

    from pycpet_amd import clusterhist as CH, io as cio, ops, workloads as W

    out = {}
    # -- dipole streamlines
    rng = np.random.default_rng(11)
    mu = (0.3 * rng.standard_normal((case.N, 3))).astype(np.float32)
    st, it, dx, dmu = T(case.starts), T(case.n_iter), T(case.x), T(mu)
    ops.topo_dipole(st, it, case.h, case.dims, dx, dmu, ctx=ctx)
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(2):
        r = ops.topo_dipole(st, it, case.h, case.dims, dx, dmu, ctx=ctx)
        kms.append(ctx.phase_times(3)[1])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / 2
    dsteps = r.steps.cpu().numpy()
    pairs = float((dsteps.astype(np.int64) + 4).sum()) * case.N
    row = {"pair_evals_per_s": pairs / wall, "ms_per_step": 1e3 * wall,
           "kernel_ms": float(np.mean(kms)), "streamlines_per_s": case.L / wall,
           "euler_steps": int(dsteps.astype(np.int64).sum())}
    if cpu:
        from concurrent.futures import ThreadPoolExecutor

        from oracle import oracle as O

        if O.Ref.available():
            ref = O.Ref()
            order = np.random.default_rng(124).permutation(case.L)
            done, pos = [], 0
            t1 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=cores) as ex:
                while time.perf_counter() - t1 < 4.0 and pos < case.L:
                    batch = order[pos:pos + 4 * cores]
                    list(ex.map(lambda i: ref.thread_operation_dipole(
                        case.starts[i], case.n_iter[i], case.h, case.dims, case.x, mu), batch))
                    done.extend(batch.tolist())
                    pos += len(batch)
            dt = time.perf_counter() - t1
            cp = float((dsteps[done].astype(np.int64) + 4).sum()) * case.N
            row["cpu_reference"] = {"pair_evals_per_s": cp / dt, "cores": cores,
                                    "sample": f"{len(done)} streamlines, {dt:.1f} s"}
    out["dipole_topo_5k"] = row

    # -- residue breakdown: each residue's few charges alone drive the same
    # starts (scripts/residue_breakdown_analysis.py:160-188), all residues of
    # a frame in batched launches with their 20x20 histograms
    from pycpet_amd import residue_breakdown as RB

    R, Lr = 200, 20_000
    rst, rit, _ = W.topo_starts(Lr, case.dims, case.h, 2000, seed=61, start_seed=62)
    rg = np.random.default_rng(31)
    frames = []
    for _ in range(R):
        n = int(rg.integers(8, 25))
        c = rg.standard_normal(3)
        c *= rg.uniform(5.0, 14.0) / np.linalg.norm(c)
        frames.append({"x": (c + rg.normal(0.0, 1.2, (n, 3))).astype(np.float32),
                       "q": rg.normal(0.0, 0.4, n).astype(np.float32), "starts": rst,
                       "n_iter": rit, "h": case.h, "dims": case.dims})
    edges = (np.linspace(0.0, 3.2, 21), np.linspace(0.0, 40.0, 21))
    xs, qs = [f["x"] for f in frames], [f["q"] for f in frames]
    count = lambda xs_, qs_: RB.residue_counts(xs_, qs_, rst, rit, case.h, case.dims, edges,
                                               ctx, return_steps=True)
    count(xs[:4], qs[:4])
    t_b, (_, rsum) = _best_of(lambda: count(xs, qs), k=2)
    rpairs = float(sum((int(s_) + 4 * Lr) * len(q) for s_, q in zip(rsum, qs)))
    row = {"residues": R, "starts_per_residue": Lr, "charges_per_residue": "8-24",
           "wall_s": t_b, "pair_evals_per_s": rpairs / t_b,
           "streamlines_per_s": R * Lr / t_b,
           "note": "residue_breakdown.residue_counts: host charge sets and starts in, the "
                   "histogram counts out; the starts uploaded once"}
    if cpu:
        from concurrent.futures import ThreadPoolExecutor

        from oracle import oracle as O

        if O.Ref.available():
            ref = O.Ref()
            pick = np.random.default_rng(125).integers(0, [R, Lr], (200_000, 2))
            done, pos = [], 0
            t1 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=cores) as ex:
                while time.perf_counter() - t1 < 4.0 and pos < len(pick):
                    chunk = pick[pos:pos + 16 * cores]
                    list(ex.map(lambda fi: ref.thread_operation(
                        rst[fi[1]], rit[fi[1]], case.h, case.dims, frames[fi[0]]["x"],
                        frames[fi[0]]["q"]), chunk))
                    done.extend(chunk.tolist())
                    pos += len(chunk)
            dt = time.perf_counter() - t1
            # per-streamline steps of the sampled residues, from one host-array batch
            sampled = sorted({f for f, _ in done})
            from pycpet_amd import batch as B

            sres = B.batch_topo([frames[f] for f in sampled], ctx=ctx)
            ssteps = {f: r.steps.astype(np.int64) for f, r in zip(sampled, sres)}
            cp = float(sum((ssteps[f][i] + 4) * len(frames[f]["q"]) for f, i in done))
            row["cpu_reference"] = {"pair_evals_per_s": cp / dt,
                                    "streamlines_per_s": len(done) / dt, "cores": cores,
                                    "sample": f"{len(done)} (residue, start) pairs, {dt:.1f} s"}
    out["residue_breakdown_200x20k"] = row

    with tempfile.TemporaryDirectory() as tmp:
        # -- frame I/O
        n_atoms = 200_000
        xa = np.round(rng.uniform(-60, 60, (n_atoms, 3)), 3)
        qa = np.round(rng.normal(0, 0.4, n_atoms), 4)
        pdb = os.path.join(tmp, "frame.pdb")
        _write_pdb(pdb, xa, qa)
        mb = os.path.getsize(pdb) / 1e6
        # each leg: the best of three runs (page-cache and scheduling noise)
        t_nat, nat_out = _best_of(lambda: cio.parse_pdb(pdb, get_charges=True))
        t_py, py_out = _best_of(lambda: cio._parse_pdb_text(pdb, True, "float32"))
        same = bool(np.array_equal(nat_out[0], py_out[0]) and np.array_equal(nat_out[1], py_out[1]))
        rows = rng.standard_normal((1_000_000, 2))
        top = os.path.join(tmp, "a.top")
        ref_top = os.path.join(tmp, "b.top")
        t_w, _ = _best_of(lambda: cio.write_rows(top, rows, fmt="%.18e"))
        t_ws, _ = _best_of(lambda: np.savetxt(ref_top, rows))
        with open(top, "rb") as fa, open(ref_top, "rb") as fb:
            same_top = fa.read() == fb.read()
        t_r, back = _best_of(lambda: CH.read_topo_file(top))
        t_rl, back_np = _best_of(lambda: np.loadtxt(top))
        out["frame_io"] = {
            "pdb_atoms": n_atoms, "pdb_mb": mb,
            "pdb_parse_native_s": t_nat, "pdb_parse_python_s": t_py,
            "pdb_parse_speedup": t_py / t_nat, "pdb_same_arrays": same,
            "top_rows": len(rows), "top_write_native_s": t_w, "top_write_savetxt_s": t_ws,
            "top_write_speedup": t_ws / t_w, "top_bytes_identical": bool(same_top),
            "top_read_native_s": t_r, "top_read_loadtxt_s": t_rl,
            "top_read_speedup": t_rl / t_r,
            "top_read_equal": bool(np.array_equal(back, back_np)),
            "python_side": "pdb: the line-by-line restatement of the reference parser in "
                           "pycpet_amd.io; .top: numpy.savetxt / numpy.loadtxt"}

        # -- the options-file frame driver (CPET.run): parse, filters and box
        # frame, compute, and the output file of each of 32 frame files
        from pycpet_amd.CPET import CPET

        ind = os.path.join(tmp, "frames")
        os.mkdir(ind)
        nd_atoms = 10_000
        for k in range(32):
            xk, qk = W.ball_exclusion(nd_atoms, 35.0, 1.5, 3.0, seed=900 + k)
            _write_pdb(os.path.join(ind, f"frame{k:02d}.pdb"), xk.astype(np.float64), qk)
        base = {"inputpath": ind, "center": [0.0, 0.0, 0.0], "x": [1.0, 0.0, 0.0],
                "y": [0.0, 1.0, 0.0]}
        drv = {}
        for method, extra in (("topo", {"dimensions": [1.5, 1.5, 1.5], "step_size": 0.005,
                                        "n_samples": 10_000, "max_streamline_init": "fixed_rand"}),
                              ("volume", {"dimensions": [1.0, 1.0, 1.0], "step_size": 0.1}),
                              ("point_field", {"step_size": 0.1})):
            opts = dict(base, CPET_method=method, outputpath=os.path.join(tmp, f"out_{method}"),
                        **extra)
            t1 = time.perf_counter()
            CPET(opts).run()
            dt = time.perf_counter() - t1
            drv[method] = {"frames": 32, "atoms_per_frame": nd_atoms, "wall_s": dt,
                           "frames_per_s": 32 / dt, **extra}
        drv["note"] = ("CPET(options).run() end to end: 32 PDB files parsed, box frame, "
                       "topo in batched launches of 16 frames (10,648 starts each), a 21^3 "
                       "field grid per frame, or the field at the box centre; each frame's "
                       ".top / _efield.dat written, or one point_field.dat line per frame")
        out["driver_32_frames"] = drv

    # -- cluster step: per-frame features from a mixture, like .top files
    F, Lf = 256, 20_000
    feats = []
    for f in range(F):
        g = np.random.default_rng(500 + f)
        feats.append(np.column_stack([np.abs(g.normal(1.0 + 0.002 * f, 0.3, Lf)),
                                      np.abs(g.normal(5.0, 2.0, Lf))]))
    CH.histograms_from_features(feats[:2], ctx=ctx)
    # make_histograms' work on features in memory: the shared bins (min, max
    # and IQR from order statistics selected on the GPU) and the histograms
    t_h, hist = _best_of(lambda: CH.histograms_from_features(feats, ctx=ctx))
    t_e, edges = _best_of(lambda: CH.histogram_edges(feats, ctx=ctx))
    t_d, _ = _best_of(lambda: CH.construct_distance_matrix(hist, ctx=ctx))
    row = {"frames": F, "features_per_frame": Lf, "bins": [len(edges[0]) - 1, len(edges[1]) - 1],
           "histograms_s": t_h, "edges_alone_s": t_e, "chi2_matrix_s": t_d}
    # the same step from .top files, as make_histograms takes it (rows written
    # as the topo driver writes them)
    topdir = tempfile.mkdtemp()
    files = []
    for k, f in enumerate(feats):
        path = os.path.join(topdir, f"frame{k:03d}.top")
        cio.write_rows(path, f, fmt="%.18e")
        files.append(path)
    CH.make_histograms(files[:2], ctx=ctx)
    t_hf, hist_f = _best_of(lambda: CH.make_histograms(files, ctx=ctx), k=2)
    row["histograms_from_files_s"] = t_hf
    row["from_files_equal_in_memory"] = bool(np.array_equal(hist_f, hist))
    if cpu:
        import scipy.stats  # noqa: F401  imported at module load by the reference

        from oracle import oracle as O

        # the reference steps are slow (about 80 ms per frame for the
        # histograms, a Python call per frame pair for the matrix): time them
        # on the first 32 / 16 frames and scale linearly / by frame pairs
        nh, nd_ = 32, 16
        t1 = time.perf_counter()
        oh, _ = O.cluster_histograms(feats[:nh])
        t_oh = time.perf_counter() - t1
        t1 = time.perf_counter()
        O.chi2_matrix(oh[:nd_])
        t_od = time.perf_counter() - t1
        row["cpu_reference"] = {
            "histograms_s_scaled_to_all": t_oh * F / nh,
            "chi2_matrix_s_scaled_to_all": t_od * (F * (F - 1)) / (nd_ * (nd_ - 1)),
            "sample": f"histograms of {nh} frames in {t_oh:.2f} s, the matrix of {nd_} "
                      f"frames in {t_od:.2f} s, scaled to {F} frames",
            "how": "the reference's steps with numpy.histogram2d / scipy iqr and sklearn "
                   "pairwise_distances over the chi2 function, one core"}
        nf = 16
        t1 = time.perf_counter()
        O.make_histograms_files(files[:nf])
        t_of = time.perf_counter() - t1
        row["cpu_reference"]["histograms_from_files_s_scaled_to_all"] = t_of * F / nf
        row["cpu_reference"]["from_files_sample"] = (
            f"make_histograms' three passes over {nf} .top files in {t_of:.2f} s, scaled "
            f"to {F}")
    import shutil

    shutil.rmtree(topdir, ignore_errors=True)
    out["cluster_step"] = row
    return out


def forms_rows() -> dict:
    """Short samples of the two north-star forms, on the driver's own run:
    batching (32 mixed MD frames, 16 volume and 16 streamline frames of
    3k-40k charges: one volume and one persistent streamline launch on two
    streams, default shapes, against the per-frame calls at their tuned
    shapes; and 200 residue-sized streamline frames of 8-24 charges, batched
    against per-frame calls) and serving (Poisson arrivals at 200 req/s for 10 s of the
    120 s configuration, continuous batching against one request at a
    time).  The full-size runs are `python -m pycpet_amd.framebench batch
    --frames 512` and `serve --duration 120`."""
    from pycpet_amd import framebench as FB

    out = {}
    b = FB.run_batch(32, 20_000, forms=("batched", "per_frame"), autotune=False)
    out["batching_32_frames"] = {k: b[k] for k in (
        "frames", "volume_frames", "topo_frames", "pair_evals", "batched_s",
        "batched_frames_per_s", "per_frame_s", "per_frame_frames_per_s",
        "batched_vs_per_frame", "batched_pair_evals_per_s")}
    # small frames: 200 residue-sized frames (8-24 charges) over the same
    # 20,000 starts, batched against the per-frame calls at their tuned shapes
    import torch

    from pycpet_amd import _native as nat, batch as B, ops, workloads as W

    dev = torch.device("cuda:0")
    dims = np.array([1.5, 1.5, 1.5], np.float32)
    rst, rit, _ = W.topo_starts(20_000, dims, 0.001, 2000, seed=61, start_seed=62)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    st, it = T(rst), T(rit)
    rg = np.random.default_rng(31)
    frames = []
    for _ in range(200):
        n = int(rg.integers(8, 25))
        c = rg.standard_normal(3)
        c *= rg.uniform(5.0, 14.0) / np.linalg.norm(c)
        frames.append({"x": T((c + rg.normal(0.0, 1.2, (n, 3))).astype(np.float32)),
                       "q": T(rg.normal(0.0, 0.4, n).astype(np.float32)), "starts": st,
                       "n_iter": it, "h": 0.001, "dims": dims})
    ctx = nat.Context(0)
    edges = (np.linspace(0.0, 3.2, 21), np.linspace(0.0, 40.0, 21))

    def per_frame():
        r = [ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"],
                      edges=edges, ctx=ctx) for f in frames]
        torch.cuda.synchronize()
        return r

    def batched():
        r = B.batch_topo(frames, edges=edges, ctx=ctx)
        torch.cuda.synchronize()
        return r

    per_frame()
    batched()
    t_pf, a = _best_of(per_frame, k=2)
    t_b, b = _best_of(batched, k=2)
    same = all(torch.equal(x.feats, y.feats) and torch.equal(x.hist, y.hist) for x, y in zip(a, b))
    out["batching_small_frames"] = {
        "frames": len(frames), "charges_per_frame": "8-24", "starts_per_frame": 20_000,
        "per_frame_s": t_pf, "batched_s": t_b, "batched_vs_per_frame": t_pf / t_b,
        "batched_frames_per_s": len(frames) / t_b, "results_bitwise_equal": bool(same),
        "note": "device tensors in, 20x20 histograms counted in the launches"}
    sv = FB.run_serve(200.0, 10.0, 4, plain_max_s=20.0)
    keep = ("completed", "rejected", "frames_per_s", "p50_ms", "p99_ms", "wall_s")
    out["serving_200rps_10s"] = {
        "requests": sv["requests"], "mix": sv["mix"], "lanes": sv["lanes"],
        "continuous_batching": {k: v for k, v in sv["continuous_batching"].items() if k in keep},
        "one_at_a_time": {k: v for k, v in sv["plain_one_at_a_time"].items() if k in keep}}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-volume", action="store_true")
    ap.add_argument("--no-rk4", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--no-aux", action="store_true")
    ap.add_argument("--no-forms", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        return 2

    import torch

    from pycpet_amd import _native as nat, ops, workloads as W

    dev = torch.device("cuda:0")
    case = W.topo_5k()
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    starts, n_iter, x, q = T(case.starts), T(case.n_iter), T(case.x), T(case.q)
    dims = case.dims  # a host parameter, like h
    ctx = nat.Context(0)
    ctx.set_profiling(True)

    # the FP64 run (untimed): the histogram range, and the topo_f64 line
    ops.topo(starts, n_iter, case.h, dims, x, q, prec="f64", ctx=ctx)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    r64 = ops.topo(starts, n_iter, case.h, dims, x, q, prec="f64", ctx=ctx)
    k64 = ctx.phase_times(3)[1]
    torch.cuda.synchronize()
    w64 = time.perf_counter() - t3
    f64 = r64.feats.cpu().numpy()
    edges = (torch.from_numpy(np.linspace(f64[:, 0].min(), f64[:, 0].max(), 21)).to(dev),
             torch.from_numpy(np.linspace(f64[:, 1].min(), f64[:, 1].max(), 21)).to(dev))

    for _ in range(args.warmup):
        r = ops.topo(starts, n_iter, case.h, dims, x, q, edges=edges, ctx=ctx)
    torch.cuda.synchronize()

    prop_ms = []
    epi_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = ops.topo(starts, n_iter, case.h, dims, x, q, edges=edges, ctx=ctx)
        ph = ctx.phase_times(3)  # waits for this step's last event
        prop_ms.append(ph[1])
        epi_ms.append(ph[2])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    hist = r.hist.cpu().numpy()
    fr = r.feats.cpu().numpy().astype(np.float32).astype(np.float64)
    hn, _, _ = np.histogram2d(fr[:, 0], fr[:, 1], bins=[e.cpu().numpy() for e in edges])
    if not np.array_equal(hist, hn):
        print("histogram differs from numpy.histogram2d of the features", file=sys.stderr)
        return 1

    steps_np = r.steps.cpu().numpy()
    euler_steps = int(steps_np.astype(np.int64).sum())
    pairs_per_step = float(euler_steps + 4 * case.L) * case.N
    ms_per_step = 1e3 * wall / args.steps
    value = pairs_per_step * args.steps / wall
    prop = float(np.mean(prop_ms))
    achieved = FLOP_PER_PAIR * case.N * euler_steps / (prop * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 positions and curvature steps)",
        "data": "synthetic",
        "config": {"workload": "topo_5k_charges_100k_streamlines", "charges": case.N,
                   "streamlines": case.L, "h": case.h, "box_half_width": 1.5,
                   "n_iter": "randint(1, 5196) seed 42, capped at 2000",
                   "integrator": "euler (reference)", "euler_steps": euler_steps,
                   "histogram": "20x20 (distance, mean curvature), range = min/max of the FP64 "
                                "features, counted in the epilogue"},
        "streamlines_per_s": case.L * args.steps / wall,
        "roofline": {
            "bound": "valu",
            "kernel": "topo_propagate",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "frac_of_unpacked_peak": achieved / PEAK_FP32_UNPACKED_TFLOPS,
            "traffic": measured_traffic(),
            "traffic_unit": "bytes per launch (rocprofv3 PMC, copy-calibrated)",
            "kernel_ms": prop,
            "epilogue_ms": float(np.mean(epi_ms)),
            "flop_per_pair": FLOP_PER_PAIR,
        },
    }

    if not args.no_volume:
        vc = W.volume_30k()
        ax = [T(a) for a in vc.axes]
        vx, vq = T(vc.x), T(vc.q)
        for _ in range(2):
            ops.field_grid(*ax, vx, vq, ctx=ctx)
        torch.cuda.synchronize()
        kms, fms = [], []
        t1 = time.perf_counter()
        for _ in range(args.steps):
            ops.field_grid(*ax, vx, vq, ctx=ctx)
            ph = ctx.phase_times(3)  # field kernel, split reduction, FP64 recompute
            kms.append(ph[0] + ph[1])
            fms.append(ph[2])
        torch.cuda.synchronize()
        vwall = (time.perf_counter() - t1) / args.steps
        vpairs = float(vc.M) * vc.N
        out["volume_30k"] = {
            "pair_evals_per_s": vpairs / vwall,
            "ms_per_frame": 1e3 * vwall,
            "grid_points": vc.M, "charges": vc.N,
            "kernel_ms": float(np.mean(kms)),  # field kernel + charge-split reduction
            "recompute_ms": float(np.mean(fms)),
            "recomputed_points": ctx.refined_points(),
            "achieved_tflops": FLOP_PER_PAIR * vpairs / (np.mean(kms) * 1e-3) / 1e12,
        }
        # the same grid, potential only (volume_ESP): mixed FP32/FP64 pair
        # term, 11 algorithmic FLOP per pair
        for _ in range(2):
            ops.field_grid(*ax, vx, vq, want_E=False, want_phi=True, ctx=ctx)
        torch.cuda.synchronize()
        kms = []
        t1 = time.perf_counter()
        for _ in range(args.steps):
            ops.field_grid(*ax, vx, vq, want_E=False, want_phi=True, ctx=ctx)
            ph = ctx.phase_times(3)
            kms.append(ph[0] + ph[1])
        torch.cuda.synchronize()
        pwall = (time.perf_counter() - t1) / args.steps
        out["volume_esp_30k"] = {
            "pair_evals_per_s": vpairs / pwall,
            "ms_per_frame": 1e3 * pwall,
            "kernel_ms": float(np.mean(kms)),
            "achieved_tflops": 11 * vpairs / (np.mean(kms) * 1e-3) / 1e12,
        }

    if not args.no_f64:
        # the same streamlines with FP64 field sums: its speed, and the FP32
        # path's error against it over all 100,000 streamlines (the FP64 mode
        # is within 4e-15 of the CPU FP64 oracle on the parity tests)
        s64 = r64.steps.cpu().numpy()
        f32 = r.feats.cpu().numpy()
        m = (steps_np == s64) & (r.exited.cpu().numpy() == r64.exited.cpu().numpy())
        dd = np.abs(f32[m, 0] - f64[m, 0]) / np.maximum(np.abs(f64[m, 0]), 1e-12)
        dk = np.abs(f32[m, 1] - f64[m, 1]) / np.maximum(np.abs(f64[m, 1]), 1e-6)
        out["topo_f64"] = {
            "pair_evals_per_s": float(int(s64.astype(np.int64).sum()) + 4 * case.L) * case.N / w64,
            "kernel_ms": k64,
            "f32_vs_f64": {"match_fraction": float(m.mean()),
                           "distance_rel_max": float(dd.max()),
                           "distance_rel_rms": float(np.sqrt(np.mean(dd ** 2))),
                           "curvature_rel_max": float(dk.max()),
                           "curvature_rel_rms": float(np.sqrt(np.mean(dk ** 2)))},
        }

    if not args.no_rk4:
        # the same streamline workload with the RK4 integrator (an extension:
        # four field evaluations per step, the curvature steps RK4 as well)
        ops.topo(starts, n_iter, case.h, dims, x, q, integrator="rk4", ctx=ctx)
        torch.cuda.synchronize()
        kms = []
        t2 = time.perf_counter()
        for _ in range(2):
            rk = ops.topo(starts, n_iter, case.h, dims, x, q, integrator="rk4", ctx=ctx)
            kms.append(ctx.phase_times(3)[1])
        torch.cuda.synchronize()
        rwall = (time.perf_counter() - t2) / 2
        rsteps = int(rk.steps.cpu().numpy().astype(np.int64).sum())
        rpairs = float(4 * rsteps + 16 * case.L) * case.N
        out["topo_rk4"] = {
            "pair_evals_per_s": rpairs / rwall,
            "ms_per_step": 1e3 * rwall,
            "streamlines_per_s": case.L / rwall,
            "rk4_steps": rsteps,
            "kernel_ms": float(np.mean(kms)),
            "achieved_tflops": FLOP_PER_PAIR * 4 * rsteps * case.N / (np.mean(kms) * 1e-3) / 1e12,
        }

    avail = len(os.sched_getaffinity(0))
    cores = min(16, avail)
    if not args.no_aux:
        out["aux"] = aux_rows(ctx, T, case, not args.no_cpu_baseline, cores)
    if not args.no_forms:
        out["forms"] = forms_rows()

    if not args.no_cpu_baseline and int(os.environ.get("RANK", "0")) == 0:
        out["cpu_baseline"] = cpu_baseline(case, steps_np, args.cpu_seconds, cores)
        out["cpu_baseline"]["host_nproc"] = os.cpu_count()
        out["cpu_baseline"]["host_cpus_usable"] = avail
        grids = cpu_baseline_grids()
        if grids:
            out["cpu_baseline"].update(grids)
    else:
        out["cpu_baseline"] = None

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
