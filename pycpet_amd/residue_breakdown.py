"""Per-residue breakdown of a frame's field topology, the command of the
reference's scripts/residue_breakdown_analysis.py (CPET/source/scripts/
residue_breakdown_analysis.py:13-210):

    python -m pycpet_amd.residue_breakdown -t <tops dir> -p <pdb dir>
        -d DMIN,DMAX -dx DRES -c CMIN,CMAX -cx CRES [-o options.json]

For every frame (the sorted PDB files; the frame's full topology is
<tops>/<frame>.top), each residue's own charges alone drive the frame's
streamline starts, and the chi-square distance (distance_numpy) between that
topology's histogram and the full one's is recorded.  The histograms are
numpy.histogram2d over the explicit bounds with int((max - min) / res) bins
per axis, divided by their total (get_hist_grid).  Residues are the
distinct (resname, residue number) pairs of the first frame in order, water
("WAT") left out; per residue, the mean and standard deviation over frames
go to chi_squared_avg.txt sorted by the mean, and the residue list to
residue_list.txt, in the working directory.

On the GPU, a frame's residues run as batched streamline launches
(cpet_batch_topo, one frame per residue, up to a bounded number of
streamlines per launch) on device tensors, the shared starts uploaded once,
and each residue's histogram is counted in the launch's epilogue, exactly
(the float32-rounded features, numpy's bin rule).  Where the reference
constructs a ValueError for a missing options file or malformed bounds
without raising it, this raises it.
"""
from __future__ import annotations

import argparse
import json
import os
from typing import List, Optional, Sequence

import numpy as np

# at most this many streamlines (residues x starts) per batched launch:
# bounds the replicated starts and the features held at once
_MAX_STREAMLINES_PER_LAUNCH = 8_000_000


def load_hist(pdb: str, tops_dir: str) -> np.ndarray:
    """(L, 2) rows of <tops_dir>/<pdb without its 4-character extension>.top
    (CPET/source/scripts/residue_breakdown_analysis.py:13-25), read by the
    native reader (whitespace-separated numbers; "#" lines are skipped)."""
    from .clusterhist import read_topo_file

    return read_topo_file(os.path.join(tops_dir, pdb[:-4] + ".top"))


def get_hist_grid(hist, distbound, distance_nbins, curvbound, curvature_nbins) -> np.ndarray:
    """numpy.histogram2d of the (distance, curvature) rows over the explicit
    bounds, divided by its total, flattened (residue_breakdown_analysis.py:
    28-37)."""
    hist = np.asarray(hist)
    a, _, _ = np.histogram2d(hist[:, 0], hist[:, 1], bins=[distance_nbins, curvature_nbins],
                             range=[distbound, curvbound])
    with np.errstate(invalid="ignore"):
        return (a / np.sum(a)).flatten()


def residue_keys(resnames: Sequence, resnums: Sequence) -> List[dict]:
    """Distinct {"resname", "resid"} pairs in first-appearance order, water
    ("WAT") left out (residue_breakdown_analysis.py:110-125)."""
    seen, out = set(), []
    for rn, ri in zip(resnames, resnums):
        if rn == "WAT":
            continue
        key = (rn, ri)
        if key not in seen:
            seen.add(key)
            out.append({"resname": rn, "resid": ri})
    return out


def residue_counts(xs: Sequence[np.ndarray], qs: Sequence[np.ndarray], starts, n_iter, h,
                   dims, edges, ctx=None, return_steps: bool = False):
    """(R, nd, nc) exact histogram counts of the streamlines that each charge
    set (xs[r], qs[r]) drives from the same starts and step limits: batched
    launches (cpet_batch_topo) on device tensors, the starts uploaded once
    and repeated on the GPU, each launch up to a bounded number of
    streamlines, only the counts copied back.  With return_steps, also each
    set's total Euler steps (R,) int64."""
    import torch

    from . import _native as nat, batch

    ctx = ctx or nat.default_context()
    dev = torch.device("cuda", ctx.device)
    st = torch.from_numpy(np.ascontiguousarray(starts, np.float32).reshape(-1, 3)).to(dev)
    it = torch.from_numpy(np.ascontiguousarray(n_iter, np.int32).ravel()).to(dev)
    sizes = [len(q) for q in qs]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    X = torch.from_numpy(np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros((0, 3)),
                                              np.float32).reshape(-1, 3)).to(dev)
    Q = torch.from_numpy(np.ascontiguousarray(np.concatenate(qs) if qs else np.zeros(0),
                                              np.float32).ravel()).to(dev)
    dims = np.ascontiguousarray(dims, np.float32).ravel()
    nd, nc = len(edges[0]) - 1, len(edges[1]) - 1
    out = np.zeros((len(qs), nd, nc), np.float64)
    steps = np.zeros(len(qs), np.int64)
    per_launch = max(1, _MAX_STREAMLINES_PER_LAUNCH // max(1, st.shape[0]))
    for c0 in range(0, len(qs), per_launch):
        rs = range(c0, min(len(qs), c0 + per_launch))
        frames = [{"x": X[offs[r]:offs[r + 1]], "q": Q[offs[r]:offs[r + 1]], "starts": st,
                   "n_iter": it, "h": h, "dims": dims} for r in rs]
        res = batch.batch_topo(frames, edges=edges, ctx=ctx)
        out[c0:c0 + len(rs)] = torch.stack([r.hist for r in res]).cpu().numpy()
        if return_steps:
            steps[c0:c0 + len(rs)] = torch.stack(
                [r.steps.to(torch.int64).sum() for r in res]).cpu().numpy()
    return (out, steps) if return_steps else out


def residue_histograms(calc, residues: Sequence[dict], edges, ctx=None) -> List[Optional[np.ndarray]]:
    """For each residue, the normalised (flattened) histogram of the
    topology driven by that residue's charges alone over calc's starts, or
    None for a residue with no charges in this frame."""
    names = list(calc.resids)
    nums = list(calc.residue_number)
    # the atoms of each (name, number) pair, in order, in one pass (the
    # reference scans all atoms per residue; dict lookup compares the same way)
    members = {}
    for i, key in enumerate(zip(names, nums)):
        members.setdefault(key, []).append(i)
    ks, xs, qs = [], [], []
    Qall = np.asarray(calc.Q).reshape(-1)
    for k, res in enumerate(residues):
        idx = members.get((res["resname"], res["resid"]), [])
        if idx:
            ks.append(k)
            xs.append(np.asarray(calc.x)[idx])
            qs.append(Qall[idx])
    out: List[Optional[np.ndarray]] = [None] * len(residues)
    if not ks:
        return out
    counts = residue_counts(xs, qs, calc.random_start_points, calc.random_max_samples,
                            calc.step_size, calc.dimensions, edges, ctx)
    for k, a in zip(ks, counts):
        with np.errstate(invalid="ignore"):
            out[k] = (a / np.sum(a)).flatten()
    return out


def _plain(v):
    """Python scalars for the text files (numpy scalars print as plain
    numbers there, as under the numpy the reference pins)."""
    if isinstance(v, dict):
        return {k: _plain(x) for k, x in v.items()}
    if isinstance(v, np.generic):
        return v.item()
    return v


def main(argv=None) -> int:
    from .calculator import calculator
    from .clusterhist import distance_numpy
    from .field_ops import histogram_edges

    ap = argparse.ArgumentParser(description="Residue breakdown analysis")
    ap.add_argument("-t", "--tops", help="Topology file directory", required=True)
    ap.add_argument("-p", "--pdbs", help="PDB file directory", required=True)
    ap.add_argument("-c", "--curvbound", help="Histogram curvature bounds", required=True)
    ap.add_argument("-cx", "--curvres", help="Histogram curvature resolution", required=True)
    ap.add_argument("-d", "--distbound", help="Histogram Euclidean distance bounds",
                    required=True)
    ap.add_argument("-dx", "--distres", help="Histogram Euclidean distance resolution",
                    required=True)
    ap.add_argument("-o", type=str, help="Options for CPET", default="./options/options.json")
    args = ap.parse_args(argv)
    if not os.path.exists(args.o):
        raise ValueError("Error: Options file not found")
    with open(args.o) as fh:
        options = json.load(fh)
    curvbound = [float(v) for v in args.curvbound.split(",")]
    distbound = [float(v) for v in args.distbound.split(",")]
    if len(curvbound) != 2 or len(distbound) != 2:
        raise ValueError("Error: Curvature and distance bounds must be lists of 2 numbers")
    if curvbound[1] <= curvbound[0] or distbound[1] <= distbound[0]:
        raise ValueError("Error: Upper bound (right) must be greater than lower bound (left)")
    distance_nbins = int((distbound[1] - distbound[0]) / float(args.distres))
    curvature_nbins = int((curvbound[1] - curvbound[0]) / float(args.curvres))
    edges = (histogram_edges(distbound[0], distbound[1], distance_nbins),
             histogram_edges(curvbound[0], curvbound[1], curvature_nbins))

    tops = sorted(os.listdir(args.tops))
    pdbs = sorted(os.listdir(args.pdbs))
    for t in tops:
        if t[:-4] + ".pdb" not in pdbs:
            raise ValueError("Error: Topology " + t + " does not have a corresponding pdb")

    first = calculator(options, path_to_pdb=os.path.join(args.pdbs, pdbs[0]))
    residues = residue_keys(first.resids, first.residue_number)
    with open("residue_list.txt", "w") as fh:
        for item in residues:
            fh.write("%s\n" % _plain(item))

    breakdown = []
    for pdb in pdbs:
        calc = calculator(options, path_to_pdb=os.path.join(args.pdbs, pdb))
        original = get_hist_grid(load_hist(pdb, args.tops), distbound, distance_nbins,
                                 curvbound, curvature_nbins)
        grids = residue_histograms(calc, residues, edges)
        per_res = []
        for res, grid in zip(residues, grids):
            with np.errstate(invalid="ignore"):
                chi2 = (np.float64(np.nan) if grid is None
                        else distance_numpy(grid, original))
            print(pdb, _plain(res), chi2)
            per_res.append([res, chi2])
        breakdown.append([pdb, per_res])

    avg = []
    for res in residues:
        vals = [c for _, per_res in breakdown for r, c in per_res if r == res]
        avg.append([res, np.mean(vals), np.std(vals)])
    avg = sorted(avg, key=lambda item: item[1])
    with open("chi_squared_avg.txt", "w") as fh:
        for item in avg:
            fh.write("%s\n" % [_plain(v) for v in item])
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
