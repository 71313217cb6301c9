"""The clustering step after "topo": per-frame feature histograms and their
chi-square distance matrix, with the reference's function names, arguments
and return values.

  make_histograms(topo_files)         CPET/utils/calculator.py:596-718
  make_histograms_mem(files, outdir)  CPET/utils/calculator.py:721-845
  distance_numpy(hist1, hist2)        CPET/utils/calculator.py:975-978
  construct_distance_matrix_mem(files) CPET/utils/calculator.py:981-1000
  construct_distance_matrix(hists)    CPET/utils/calculator.py:1003-1015

The host side keeps the reference's bin rule: global min / max over every
frame's features, bin width 2 * IQR / n^(1/3) (scipy.stats.iqr, n the row
count, or its mean when frames differ in length), int((max - min) / width)
bins, and the edges numpy.histogram2d builds from (bins, range).  The min,
max and the two quartiles come from exact order statistics selected on the
GPU (cpet_order_stats), combined on the host with numpy.percentile's linear
interpolation step for step.  The .top files are read by cpet_read_rows; the
histograms (cpet_cluster_histograms) and the distance matrix
(cpet_chi2_matrix) run on the GPU.
"""
from __future__ import annotations

import ctypes
import os
import warnings
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _native as nat


def read_topo_file(path) -> np.ndarray:
    """(L, 2) float64 [distance, curvature] rows of a .top file."""
    lib = nat.load()
    p = str(path).encode()
    n = ctypes.c_int64(0)
    nat.check(lib.cpet_read_rows(p, 2, None, 0, ctypes.byref(n)), lib)
    out = np.empty((n.value, 2), np.float64)
    nat.check(lib.cpet_read_rows(p, 2, out.ctypes.data if n.value else None, n.value,
                                 ctypes.byref(n)), lib)
    return out


def _read_all(paths, reader=None) -> list:
    """reader(path) for every path, in order, from up to 16 threads: the
    native reader runs outside the GIL, so small files read side by side."""
    from concurrent.futures import ThreadPoolExecutor

    reader = reader or read_topo_file
    paths = list(paths)
    if len(paths) <= 1:
        return [reader(p) for p in paths]
    nat.load()
    workers = min(16, len(paths), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(reader, paths))


def _outer_edges(lo, hi):
    """numpy.histogramdd's range handling (numpy/lib/_histograms_impl.py)."""
    lo, hi = float(lo), float(hi)
    if lo > hi:
        raise ValueError("max must be larger than min in range parameter.")
    if not (np.isfinite(lo) and np.isfinite(hi)):
        raise ValueError(f"supplied range of [{lo}, {hi}] is not finite")
    if lo == hi:
        lo, hi = lo - 0.5, hi + 0.5
    return lo, hi


# scipy.stats.iqr's default rng=(25, 75), divided by 100 as numpy.percentile
# divides it (both exact in binary)
_IQR_Q = (0.25, 0.75)


def _binning_count(features):
    """The n of the bin width: the common row count, or (with the
    reference's warning) the mean one (CPET/utils/calculator.py:654-662)."""
    lens = np.array([len(f) for f in features], dtype=int)
    if lens.size and not np.all(lens == lens[0]):
        warnings.warn(
            f"Topologies provided are of different sizes, using the mean value of "
            f"{np.mean(lens)} to represent binning for {lens}")
        return np.mean(lens)
    return lens[0]


def _flatten(features):
    feats = [np.ascontiguousarray(np.asarray(f, np.float64).reshape(-1, 2)) for f in features]
    off = np.zeros(len(feats) + 1, np.int64)
    off[1:] = np.cumsum([len(f) for f in feats])
    allf = np.concatenate(feats) if off[-1] else np.zeros((0, 2))
    return feats, off, allf


def _lerp_points(count, q):
    """numpy.percentile's 'linear' method (numpy/lib/_function_base_impl.py:
    _quantile, _get_indexes, _get_gamma) for `count` values: the two ranks
    it reads and the weight gamma.  At or past the last index both ranks are
    the last one and gamma = virtual index + 1, as numpy computes it."""
    v = np.float64(count - 1) * np.float64(q)
    if v >= count - 1:
        return count - 1, count - 1, v - np.float64(-1.0)
    lo = np.floor(v)
    return int(lo), int(lo) + 1, v - lo


def _percentile_from_order(count, q, kth):
    """numpy.percentile(x, 100 q) (linear) from kth(r), the r-th smallest
    value of x: numpy's _lerp, a + (b - a) t, or b - (b - a)(1 - t) for
    t >= 0.5, in float64."""
    i, j, g = _lerp_points(count, q)
    a, b = np.float64(kth(i)), np.float64(kth(j))
    diff = b - a
    return b - diff * (np.float64(1.0) - g) if g >= 0.5 else a + diff * g


def _stats_from_order(count, kth, nan_count=0):
    """(min, max, iqr) of one column as np.min, np.max and scipy.stats.iqr
    give them (CPET/utils/calculator.py:664-670): from its order statistics
    kth(r), with numpy's rule that any NaN makes all three NaN."""
    if nan_count:
        nan = np.float64(np.nan)
        return nan, nan, nan
    p25, p75 = (_percentile_from_order(count, q, kth) for q in _IQR_Q)
    return np.float64(kth(0)), np.float64(kth(count - 1)), np.subtract(p75, p25)


def _order_stat_ranks(count):
    ranks = {0, count - 1}
    for q in _IQR_Q:
        i, j, _ = _lerp_points(count, q)
        ranks |= {i, j}
    return sorted(ranks)


def _column_stats(allf, ctx):
    """[(min, max, iqr)] of the two feature columns, from order statistics
    selected on the GPU (cpet_order_stats: exact elements of the column)."""
    P = allf.shape[0]
    ranks = _order_stat_ranks(P)
    plan = [(c, r) for c in range(2) for r in ranks]
    cols = np.array([c for c, _ in plan], np.int32)
    rk = np.array([r for _, r in plan], np.int64)
    vals = np.empty(len(plan), np.float64)
    nans = np.zeros(2, np.int64)
    x = np.ascontiguousarray(allf, np.float64)
    nat.check(ctx.lib.cpet_order_stats(ctx.handle, P, 2, x.ctypes.data, len(plan),
                                       cols.ctypes.data, rk.ctypes.data, vals.ctypes.data,
                                       nans.ctypes.data, 0), ctx.lib)
    at = {key: k for k, key in enumerate(plan)}
    return [_stats_from_order(P, lambda r, c=c: vals[at[(c, r)]], int(nans[c]))
            for c in range(2)]


def _edges_from_stats(stats, n):
    """Shared edges from [(min, max, iqr)] of the distance and curvature
    columns and the binning count n (CPET/utils/calculator.py:669-670,
    684-685, 702-707):
    width 2 iqr / n^(1/3), int((max - min) / width) bins, numpy.histogram2d's
    edges over [min, max]."""
    (dmin, dmax, d_iqr), (cmin, cmax, c_iqr) = stats
    d_res = 2 * d_iqr / (n ** (1 / 3))
    c_res = 2 * c_iqr / (n ** (1 / 3))
    nd = int((dmax - dmin) / d_res)
    nc = int((cmax - cmin) / c_res)
    for k, nb in ((0, nd), (1, nc)):
        if nb < 1:
            raise ValueError(f"`bins[{k}]` must be positive, when an integer")
    ed = np.linspace(*_outer_edges(dmin, dmax), nd + 1)
    ec = np.linspace(*_outer_edges(cmin, cmax), nc + 1)
    return ed, ec


def _edges(allf, n, ctx):
    if allf.shape[0] == 0:
        np.max(allf[:, 0])  # numpy's error for an empty reduction, as the reference
    return _edges_from_stats(_column_stats(allf, ctx), n)


def histogram_edges(features: Sequence[np.ndarray],
                    ctx: Optional[nat.Context] = None) -> Tuple[np.ndarray, np.ndarray]:
    """The shared (distance, curvature) bin edges of make_histograms."""
    ctx = ctx or nat.default_context()
    n = _binning_count(features)
    return _edges(_flatten(features)[2], n, ctx)


def histograms_from_features(features: Sequence[np.ndarray], edges=None,
                             ctx: Optional[nat.Context] = None) -> np.ndarray:
    """(F, nd*nc) normalised histograms of per-frame (L_f, 2) features."""
    ctx = ctx or nat.default_context()
    n = _binning_count(features) if edges is None else None
    feats, off, allf = _flatten(features)
    ed, ec = edges if edges is not None else _edges(allf, n, ctx)
    ed = np.ascontiguousarray(ed, np.float64)
    ec = np.ascontiguousarray(ec, np.float64)
    if not off[-1]:
        allf = np.zeros((1, 2))
    nd, nc = len(ed) - 1, len(ec) - 1
    out = np.empty((len(feats), nd * nc), np.float64)
    nat.check(ctx.lib.cpet_cluster_histograms(
        ctx.handle, len(feats), off.ctypes.data, allf.ctypes.data, nd, ed.ctypes.data, nc,
        ec.ctypes.data, out.ctypes.data, 0), ctx.lib)
    return out


def _histogram_counts(feats, off, allf, ed, ec, ctx) -> np.ndarray:
    """(F, nd, nc) float64 bin counts over the shared edges (cpet_cluster_histograms
    with CPET_F_COUNTS)."""
    ed = np.ascontiguousarray(ed, np.float64)
    ec = np.ascontiguousarray(ec, np.float64)
    if not off[-1]:
        allf = np.zeros((1, 2))
    nd, nc = len(ed) - 1, len(ec) - 1
    out = np.empty((len(feats), nd, nc), np.float64)
    nat.check(ctx.lib.cpet_cluster_histograms(
        ctx.handle, len(feats), off.ctypes.data, allf.ctypes.data, nd, ed.ctypes.data, nc,
        ec.ctypes.data, out.ctypes.data, nat.F_COUNTS), ctx.lib)
    return out


def _read_topo_any_sep(path) -> np.ndarray:
    """make_histograms_mem's reader (CPET/utils/calculator.py:735-748): a
    line that holds a comma is split on commas, any other on whitespace; the
    first two fields of every line not starting with "#".  Files without a
    comma go through the native reader, which follows the whitespace rule."""
    with open(path, "rb") as fh:
        has_comma = b"," in fh.read()
    if not has_comma:
        return read_topo_file(path)
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            line = line.strip()
            parts = line.split(",") if "," in line else line.split()
            rows.append((float(parts[0]), float(parts[1])))
    return np.array(rows, np.float64).reshape(-1, 2)


def _density_normalised(counts, ed, ec) -> np.ndarray:
    """make_histograms_mem's values from one frame's (nd, nc) counts:
    numpy.histogramdd's density=True steps (numpy/lib/_histograms_impl.py:
    count / dx_i / dy_j / total, via plt.hist2d), then the reference divides
    every value by their sum taken row by row in order
    (CPET/utils/calculator.py:813-834)."""
    hist = np.asarray(counts, np.float64)
    s = hist.sum()
    with np.errstate(invalid="ignore", divide="ignore"):
        hist = hist / np.diff(ed).reshape(-1, 1)
        hist = hist / np.diff(ec).reshape(1, -1)
        hist /= s
        flat = hist.ravel()
        norm = np.cumsum(flat)[-1] if flat.size else 0.0  # the sequential sum
        return flat / norm


def make_histograms_mem(topo_files, output_dir, plot=False,
                        ctx: Optional[nat.Context] = None):
    """The reference's file-by-file variant (CPET/utils/calculator.py:
    721-845): the same shared bins as make_histograms, each frame's
    histogram saved as <output_dir>/<file name without its 4-character
    extension>_hist.npy (the density values, normalised to sum 1), and the
    list of those paths returned.  The counts run on the GPU.  `plot` is
    accepted and ignored (the reference shows each histogram on screen)."""
    import os

    ctx = ctx or nat.default_context()
    feats = _read_all(topo_files, _read_topo_any_sep)
    lens = [len(f) for f in feats]
    if lens and not all(x == lens[0] for x in lens):
        warnings.warn(
            f"Topologies provided are of different sizes, using the mean value of "
            f"{np.mean(lens)} to represent binning for {lens}")
        n = np.mean(lens)
    else:
        n = lens[0]
    feats, off, allf = _flatten(feats)
    ed, ec = _edges(allf, n, ctx)
    counts = _histogram_counts(feats, off, allf, ed, ec, ctx)
    hist_list = []
    for f, c in zip(topo_files, counts):
        outfile = os.path.join(output_dir, f"{str(f).split('/')[-1][:-4]}_hist.npy")
        np.save(outfile, _density_normalised(c, ed, ec))
        hist_list.append(outfile)
    return hist_list


def make_histograms(topo_files, plot=False, ctx: Optional[nat.Context] = None) -> np.ndarray:
    """Normalised, flattened 2-D histograms of every .top file's features on
    shared bins; one row per file (CPET/utils/calculator.py:596-718).
    `plot` is accepted and ignored, as the reference ignores it there."""
    feats = _read_all(topo_files)
    return histograms_from_features(feats, ctx=ctx)


def distance_numpy(hist1, hist2):
    """Chi-square distance of two histograms (CPET/utils/calculator.py:
    975-978), on the host for single pairs."""
    a = (hist1 - hist2) ** 2
    b = hist1 + hist2
    return np.sum(np.divide(a, b, out=np.zeros_like(a), where=b != 0)) / 2.0


def construct_distance_matrix_mem(hist_file_list, ctx: Optional[nat.Context] = None) -> np.ndarray:
    """(F, F) chi-square distances of histograms saved as .npy files
    (CPET/utils/calculator.py:981-1000): the reference loads two files per
    pair; here every file is loaded once and the matrix runs on the GPU (F
    histograms held at once: F x bins doubles)."""
    hists = [np.load(p) for p in hist_file_list]
    if not hists:
        return np.zeros((0, 0), np.float64)
    return construct_distance_matrix(np.stack([np.ravel(h) for h in hists]), ctx=ctx)


def construct_distance_matrix(histograms, ctx: Optional[nat.Context] = None) -> np.ndarray:
    """(F, F) chi-square distances of all histogram pairs, zero diagonal
    (CPET/utils/calculator.py:1000-1015), computed on the GPU."""
    ctx = ctx or nat.default_context()
    h = np.ascontiguousarray(np.asarray(histograms, np.float64))
    if h.ndim != 2:
        raise ValueError("histograms must be a 2-D array (frames x bins)")
    F, B = h.shape
    out = np.empty((F, F), np.float64)
    nat.check(ctx.lib.cpet_chi2_matrix(ctx.handle, F, B, h.ctypes.data, out.ctypes.data, 0),
              ctx.lib)
    return out
