"""The clustering step after "topo": per-frame feature histograms and their
chi-square distance matrix, with the reference's function names, arguments
and return values.

  make_histograms(topo_files)         CPET/utils/calculator.py:596-718
  distance_numpy(hist1, hist2)        CPET/utils/calculator.py:975-978
  construct_distance_matrix(hists)    CPET/utils/calculator.py:1000-1015

The host side keeps the reference's bin rule: global min / max over every
frame's features, bin width 2 * IQR / n^(1/3) (scipy.stats.iqr, n the row
count, or its mean when frames differ in length), int((max - min) / width)
bins, and the edges numpy.histogram2d builds from (bins, range).  The .top
files are read by cpet_read_rows; the histograms (cpet_cluster_histograms)
and the distance matrix (cpet_chi2_matrix) run on the GPU.
"""
from __future__ import annotations

import ctypes
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


def histogram_edges(features: Sequence[np.ndarray]) -> Tuple[np.ndarray, np.ndarray]:
    """The shared (distance, curvature) bin edges of make_histograms."""
    from scipy.stats import iqr

    lens = np.array([len(f) for f in features], dtype=int)
    if lens.size and not np.all(lens == lens[0]):
        warnings.warn(
            f"Topologies provided are of different sizes, using the mean value of "
            f"{np.mean(lens)} to represent binning for {lens}")
        n = np.mean(lens)
    else:
        n = lens[0]
    allf = np.concatenate([np.asarray(f, np.float64).reshape(-1, 2) for f in features])
    d, c = allf[:, 0], allf[:, 1]
    dmax, cmax, dmin, cmin = np.max(d), np.max(c), np.min(d), np.min(c)
    d_res = 2 * iqr(d) / (n ** (1 / 3))
    c_res = 2 * iqr(c) / (n ** (1 / 3))
    nd = int((dmax - dmin) / d_res)
    nc = int((cmax - cmin) / c_res)
    for k, nb in ((0, nd), (1, nc)):
        if nb < 1:
            raise ValueError(f"`bins[{k}]` must be positive, when an integer")
    ed = np.linspace(*_outer_edges(dmin, dmax), nd + 1)
    ec = np.linspace(*_outer_edges(cmin, cmax), nc + 1)
    return ed, ec


def histograms_from_features(features: Sequence[np.ndarray], edges=None,
                             ctx: Optional[nat.Context] = None) -> np.ndarray:
    """(F, nd*nc) normalised histograms of per-frame (L_f, 2) features."""
    ctx = ctx or nat.default_context()
    ed, ec = edges if edges is not None else histogram_edges(features)
    ed = np.ascontiguousarray(ed, np.float64)
    ec = np.ascontiguousarray(ec, np.float64)
    feats = [np.ascontiguousarray(np.asarray(f, np.float64).reshape(-1, 2)) for f in features]
    off = np.zeros(len(feats) + 1, np.int64)
    off[1:] = np.cumsum([len(f) for f in feats])
    allf = np.concatenate(feats) if off[-1] else np.zeros((1, 2))
    nd, nc = len(ed) - 1, len(ec) - 1
    out = np.empty((len(feats), nd * nc), np.float64)
    nat.check(ctx.lib.cpet_cluster_histograms(
        ctx.handle, len(feats), off.ctypes.data, allf.ctypes.data, nd, ed.ctypes.data, nc,
        ec.ctypes.data, out.ctypes.data, 0), ctx.lib)
    return out


def make_histograms(topo_files, plot=False, ctx: Optional[nat.Context] = None) -> np.ndarray:
    """Normalised, flattened 2-D histograms of every .top file's features on
    shared bins; one row per file (CPET/utils/calculator.py:596-718)."""
    feats = [read_topo_file(f) for f in topo_files]
    return histograms_from_features(feats, ctx=ctx)


def distance_numpy(hist1, hist2):
    """Chi-square distance of two histograms (CPET/utils/calculator.py:
    975-978), on the host for single pairs."""
    a = (hist1 - hist2) ** 2
    b = hist1 + hist2
    return np.sum(np.divide(a, b, out=np.zeros_like(a), where=b != 0)) / 2.0


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
