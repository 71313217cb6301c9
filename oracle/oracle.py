"""ORACLE — test infrastructure, not product code.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the reference results are compared with or
as the CPU baseline.  The product (pycpet_amd/) never imports it.

Contents
  field_f64 / topo_f64      ctypes front-ends of liboracle.so (cpet_oracle.c),
                            the FP64 restatement the GPU is measured against.
  field_f32 / topo_f32      FP32 port of the reference arithmetic (restated
                            CPU baseline).
  Ref                       ctypes front-end of oracle/_ref/libref_math.so,
                            the reference's own math_module.c compiled from its
                            source (present only where the reference exists at
                            build time; the built .so may travel with the tree).
  grid_axes, histogram2d, filter_* , box_basis
                            numpy restatements of the host-side pieces.
  cluster_histograms, chi2_matrix
                            the clustering step exactly as the reference runs
                            it: numpy.histogram2d + scipy.stats.iqr, and
                            sklearn pairwise_distances over distance_numpy.

Parity pins (where each function is checked) are listed per function; a
function with no pin would say "parity unpinned" here.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libref_math.so"

_P = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_lib = None


def lib():
    """liboracle.so; built by `make -C oracle` (or __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
        L = ctypes.CDLL(str(ORACLE_SO))
        L.oracle_field_f64.argtypes = [_i, _i, _P, _P, _P, _i, _P, _P]
        L.oracle_field_f32.argtypes = [_i, _i, _P, _P, _P, _i, _P, _P]
        L.oracle_topo_f64.argtypes = [_i, _i, _P, _P, _f, _P, _P, _P, _P, _P, _P, _P]
        L.oracle_topo_f32.argtypes = [_i, _i, _P, _P, _f, _P, _P, _P, _P, _P]
        L.oracle_topo_dipole_f64.argtypes = [_i, _i, _P, _P, _f, _P, _P, _P, _i, _P, _P, _P]
        L.oracle_topo_rk4_f64.argtypes = [_i, _i, _P, _P, _f, _P, _P, _P, _P, _P, _P, _P]
        for fn in (L.oracle_field_f64, L.oracle_field_f32, L.oracle_topo_f64, L.oracle_topo_f32,
                   L.oracle_topo_dipole_f64, L.oracle_topo_rk4_f64):
            fn.restype = None
        _lib = L
    return _lib


def _f32(a, last=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(-1, last) if last else a.ravel()


def _p(a):
    return a.ctypes.data if a is not None else None


def field_f64(pts, x, q, *, want_E=True, want_phi=False, soften=True):
    """FP64 E (softened as compute_looped_field, math_module.c:405-451) and/or
    phi (calc_esp_base, math_module.c:453-486).
    Pinned by: tests/test_oracle_field.py (reference-C fixtures, single-charge
    closed form, superposition, symmetric-pair cancellation, floor case)."""
    pts, x, q = _f32(pts, 3), _f32(x, 3), _f32(q)
    M, N = pts.shape[0], q.size
    E = np.zeros((M, 3)) if want_E else None
    phi = np.zeros(M) if want_phi else None
    lib().oracle_field_f64(M, N, _p(pts), _p(x), _p(q), 1 if soften else 0, _p(E), _p(phi))
    return E, phi


def field_f32(pts, x, q, *, want_E=True, want_phi=False, soften=True):
    """FP32 port (restated CPU baseline).  Pinned against field_f64 within
    FP32 rounding in tests/test_oracle_field.py."""
    pts, x, q = _f32(pts, 3), _f32(x, 3), _f32(q)
    M, N = pts.shape[0], q.size
    E = np.zeros((M, 3), np.float32) if want_E else None
    phi = np.zeros(M, np.float32) if want_phi else None
    lib().oracle_field_f32(M, N, _p(pts), _p(x), _p(q), 1 if soften else 0, _p(E), _p(phi))
    return E, phi


def topo_f64(starts, n_iter, h, dims, x, q, *, want_points=False):
    """FP64 streamlines (thread_operation, math_module.c:523-591).
    Returns feats (L,2), steps (L,), exited (L,) [, pts6 (L,6,3)].
    Pinned by: tests/test_oracle_topo.py (reference-C fixtures on steps and
    distance, straight-line closed form with exact step counts, radial-field
    kappa = 0, curvature identity |u0 x u1|/h, dipole field-line curvature)."""
    starts, x, q, dims = _f32(starts, 3), _f32(x, 3), _f32(q), _f32(dims)
    n_iter = np.ascontiguousarray(n_iter, dtype=np.int32).ravel()
    L, N = starts.shape[0], q.size
    feats = np.zeros((L, 2))
    steps = np.zeros(L, np.int32)
    exited = np.zeros(L, np.int32)
    pts6 = np.zeros((L, 6, 3)) if want_points else None
    lib().oracle_topo_f64(N, L, _p(starts), _p(n_iter), float(np.float32(h)), _p(dims),
                          _p(x), _p(q), _p(feats), _p(steps), _p(exited), _p(pts6))
    if want_points:
        return feats, steps, exited, pts6
    return feats, steps, exited


def topo_rk4_f64(starts, n_iter, h, dims, x, q, *, want_points=False):
    """FP64 streamlines of topo_f64 with classical RK4 steps on the unit
    field direction (an extension; the reference integrates with Euler).
    Pinned by: tests/test_oracle_topo.py (straight radial lines equal to the
    Euler walk, fourth-order convergence of the end point in a curved field,
    agreement with topo_f64 as h -> 0)."""
    starts, x, q, dims = _f32(starts, 3), _f32(x, 3), _f32(q), _f32(dims)
    n_iter = np.ascontiguousarray(n_iter, dtype=np.int32).ravel()
    L, N = starts.shape[0], q.size
    feats = np.zeros((L, 2))
    steps = np.zeros(L, np.int32)
    exited = np.zeros(L, np.int32)
    pts6 = np.zeros((L, 6, 3)) if want_points else None
    lib().oracle_topo_rk4_f64(N, L, _p(starts), _p(n_iter), float(np.float32(h)), _p(dims),
                              _p(x), _p(q), _p(feats), _p(steps), _p(exited), _p(pts6))
    if want_points:
        return feats, steps, exited, pts6
    return feats, steps, exited


def topo_dipole_f64(starts, n_iter, h, dims, x, mu, *, physical=False):
    """FP64 dipole streamlines (thread_operation_dipole, math_module.c:
    593-659): the walk of topo_f64 with the dipole field of
    calc_field_base_dipole (math_module.c:335-371), in the reference's form
    c = 3 mu_x R_x + mu_y R_y + mu_z R_z, or the physical c = 3 mu.R.
    Returns feats (L,2), steps (L,), exited (L,).  Pinned by
    tests/test_oracle_topo.py (reference-C thread_operation_dipole vectors,
    the on-axis closed form E = 2 k mu / r^3 with exact exit steps, and the
    two forms agreeing for dipoles along x)."""
    starts, x, mu, dims = _f32(starts, 3), _f32(x, 3), _f32(mu, 3), _f32(dims)
    n_iter = np.ascontiguousarray(n_iter, dtype=np.int32).ravel()
    L, N = starts.shape[0], x.shape[0]
    feats = np.zeros((L, 2))
    steps = np.zeros(L, np.int32)
    exited = np.zeros(L, np.int32)
    lib().oracle_topo_dipole_f64(N, L, _p(starts), _p(n_iter), float(np.float32(h)), _p(dims),
                                 _p(x), _p(mu), 1 if physical else 0, _p(feats), _p(steps),
                                 _p(exited))
    return feats, steps, exited


def topo_f32(starts, n_iter, h, dims, x, q):
    """FP32 port of thread_operation (restated CPU baseline).  Pinned by
    tests/test_oracle_topo.py (the reference C's FP32 outputs in
    tests/golden/ref_topo_200.npz, the live source-built reference where
    present, radial lines and zero steps)."""
    starts, x, q, dims = _f32(starts, 3), _f32(x, 3), _f32(q), _f32(dims)
    n_iter = np.ascontiguousarray(n_iter, dtype=np.int32).ravel()
    L, N = starts.shape[0], q.size
    feats = np.zeros((L, 2), np.float32)
    steps = np.zeros(L, np.int32)
    lib().oracle_topo_f32(N, L, _p(starts), _p(n_iter), float(np.float32(h)), _p(dims),
                          _p(x), _p(q), _p(feats), _p(steps))
    return feats, steps


class Ref:
    """The reference's own math_module.c (compiled under oracle/_ref).

    Calls follow the reference binding CPET/utils/c_ops.py (argument order,
    float32 C-order arrays, caller-allocated outputs)."""

    def __init__(self, path: Path = REF_SO):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} not built (needs the reference source)")
        L = ctypes.CDLL(str(path))
        L.compute_looped_field.argtypes = [_i, _i, _P, _P, _P, _P]
        L.calc_esp_base.argtypes = [_P, _P, _i, _P, _P]
        L.calc_field.argtypes = [_P, _P, _i, _P, _P]
        L.calc_field_base.argtypes = [_P, _P, _i, _P, _P]
        L.thread_operation.argtypes = [_i, _i, _f, _P, _P, _P, _P, _P]
        L.thread_operation_dipole.argtypes = [_i, _i, _f, _P, _P, _P, _P, _P]
        for fn in (L.compute_looped_field, L.calc_esp_base, L.calc_field,
                   L.calc_field_base, L.thread_operation, L.thread_operation_dipole):
            fn.restype = None
        self.L = L

    @staticmethod
    def available() -> bool:
        return REF_SO.exists()

    def compute_looped_field(self, pts, x, q):
        pts, x, q = _f32(pts, 3), _f32(x, 3), _f32(q)
        E = np.zeros_like(pts)
        self.L.compute_looped_field(pts.shape[0], q.size, _p(pts), _p(x), _p(q), _p(E))
        return E

    def esp_points(self, pts, x, q):
        """One calc_esp_base call per point, as compute_ESP_on_grid does
        (CPET/utils/calculator.py:468-471)."""
        pts, x, q = _f32(pts, 3), _f32(x, 3), _f32(q)
        out = np.zeros(pts.shape[0], np.float32)
        one = np.zeros(1, np.float32)
        for i in range(pts.shape[0]):
            one[0] = 0.0
            p = np.ascontiguousarray(pts[i])
            self.L.calc_esp_base(_p(one), _p(p), q.size, _p(x), _p(q))
            out[i] = one[0]
        return out

    def esp_grid_python_loop(self, pts, x, q):
        """compute_ESP_on_grid's loop (CPET/utils/calculator.py:466-475) with
        the reference wrapper's calling pattern: per point a fresh (1,)
        float32 result, an ndpointer-checked calc_esp_base call
        (CPET/utils/c_ops.py:326-341), the centre test and the float16 cast
        at the end.  The CPU baseline of the ESP path with its FFI loop."""
        import numpy.ctypeslib as npct

        fn = self.L.calc_esp_base
        f1 = npct.ndpointer(dtype=np.float32, ndim=1, flags="C")
        f2 = npct.ndpointer(dtype=np.float32, ndim=2, flags="C")
        fn.argtypes = [f1, f1, _i, f2, f1]
        try:
            pts, x, q = _f32(pts, 3), _f32(x, 3), _f32(q)
            out = np.zeros((pts.shape[0], 1), dtype=float)
            qr = q.reshape(len(q))
            for i, x_0 in enumerate(pts):
                res = np.zeros(1, dtype="float32")
                fn(res, x_0, len(q), x, qr)
                out[i] = res
                if x_0[0] == 0 and x_0[1] == 0 and x_0[2] == 0:
                    pass  # the reference logs the centre value here
            return np.concatenate((pts, out), axis=1).astype(np.half)
        finally:
            fn.argtypes = [_P, _P, _i, _P, _P]

    def esp_grid_c_loop(self, pts, x, q):
        """The same per-point calc_esp_base calls from a C loop
        (oracle/ref_loops.c): the ESP baseline without the Python loop."""
        pts, x, q = _f32(pts, 3), _f32(x, 3), _f32(q)
        fn = self.L.ref_esp_points
        fn.argtypes = [_i, _P, _i, _P, _P, _P]
        fn.restype = None
        out = np.zeros(pts.shape[0], np.float32)
        fn(pts.shape[0], _p(pts), q.size, _p(x), _p(q), _p(out))
        return out

    def calc_field(self, pt, x, q):
        pt, x, q = _f32(pt), _f32(x, 3), _f32(q)
        E = np.zeros(3, np.float32)
        self.L.calc_field(_p(E), _p(pt), q.size, _p(x), _p(q))
        return E

    def thread_operation(self, start, n_iter, h, dims, x, q):
        start, x, q, dims = _f32(start), _f32(x, 3), _f32(q), _f32(dims)
        ret = np.zeros(2, np.float32)
        self.L.thread_operation(q.size, int(n_iter), float(np.float32(h)), _p(start), _p(dims),
                                _p(x), _p(q), _p(ret))
        return ret

    def thread_operation_dipole(self, start, n_iter, h, dims, x, mu):
        """Argument order of the C function (math_module.c:593): n_dipoles,
        n_iter, step_size, x_0, dimensions, x, mu, ret."""
        start, x, mu, dims = _f32(start), _f32(x, 3), _f32(mu, 3), _f32(dims)
        ret = np.zeros(2, np.float32)
        self.L.thread_operation_dipole(x.shape[0], int(n_iter), float(np.float32(h)), _p(start),
                                       _p(dims), _p(x), _p(mu), _p(ret))
        return ret

    def topo(self, starts, n_iter, h, dims, x, q):
        starts = _f32(starts, 3)
        return np.stack([self.thread_operation(starts[i], n_iter[i], h, dims, x, q)
                         for i in range(starts.shape[0])])


# ---------------------------------------------------------------- numpy parts

def grid_axes(dims, step):
    """Restates the grid rule: N_cr = int(2 dims / step) in float64
    (CPET/source/calculator.py:258-259) and inclusive linspace per axis cast
    to float32 (CPET/utils/calculator.py:218-233).  Pinned by the coordinates
    of examples/2A_3D-field/outdir/*_efield.dat (11^3, -0.5..0.5 step 0.1) and
    the int-truncation KAT 2*0.7/0.1 -> 13 (tests/test_oracle_host.py)."""
    d = np.asarray(dims, dtype=np.float64)
    n_cr = 2 * d / step
    return [np.linspace(-d[i], d[i], int(n_cr[i]) + 1).astype(np.float32) for i in range(3)]


def histogram2d(a, b, edges_a, edges_b):
    """Brute-force restatement of numpy.histogramdd binning with explicit
    edges: bin i holds edges[i] <= v < edges[i+1], the last bin also holds
    v == edges[-1]; values outside (and NaN) are dropped.  Pinned against
    numpy.histogram2d on random and edge-valued inputs
    (tests/test_oracle_host.py)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ea = np.asarray(edges_a, dtype=np.float64)
    eb = np.asarray(edges_b, dtype=np.float64)
    H = np.zeros((ea.size - 1, eb.size - 1), dtype=np.int64)

    def bin_of(v, e):
        n = e.size - 1
        if not (e[0] <= v <= e[-1]):
            return -1
        if v == e[-1]:
            return n - 1
        i = 0
        while not (e[i] <= v < e[i + 1]):
            i += 1
        return i

    for va, vb in zip(a, b):
        ia, ib = bin_of(va, ea), bin_of(vb, eb)
        if ia >= 0 and ib >= 0:
            H[ia, ib] += 1
    return H


def box_basis(center, xpt, ypt):
    """Box frame T = [x^; y^; z^] with z^ = x^ x y^, y^ = z^ x x^
    (CPET/utils/calculator.py:200-215).  Pinned by
    tests/test_oracle_host.py::test_box_basis_closed_forms (identity and
    cyclic-permutation frames, the y point's x^ component removed,
    right-handedness)."""
    x = np.asarray(xpt) - np.asarray(center)
    y = np.asarray(ypt) - np.asarray(center)
    xu = x / np.linalg.norm(x)
    yu = y / np.linalg.norm(y)
    zu = np.cross(xu, yu)
    zu = zu / np.linalg.norm(zu)
    yu = np.cross(zu, xu)
    yu = yu / np.linalg.norm(yu)
    return np.column_stack([xu, yu, zu]).T


def filter_radius_whole_residue_keep(x, center, radius, resids, resnums):
    """Keep-mask of filter_radius_whole_residue (CPET/utils/io.py:283-338):
    runs of consecutive rows with equal (resid, resnum) are kept whole iff any
    row has |x - center| < radius.  Pinned by the reference test KAT
    tests/test_filter.py:97-128 (keep WAT#10 only) and hand-built run cases."""
    xr = np.asarray(x) - np.asarray(center)
    inside = np.linalg.norm(xr, axis=1) < radius
    keep = np.zeros(len(inside), dtype=bool)
    i = 0
    n = len(inside)
    while i < n:
        j = i
        while j + 1 < n and resids[j + 1] == resids[i] and resnums[j + 1] == resnums[i]:
            j += 1
        keep[i:j + 1] = inside[i:j + 1].any()
        i = j + 1
    return keep


def filter_in_box_keep(x, center, dims):
    """Keep-mask of filter_in_box (CPET/utils/io.py:422-475): drop rows with
    every |(x - center)_i| strictly inside dims_i.  Pinned by the reference
    test KAT tests/test_filter.py:185-210."""
    xr = np.asarray(x) - np.asarray(center)
    d = np.asarray(dims)
    inside = np.ones(xr.shape[0], dtype=bool)
    for i in range(3):
        inside &= (xr[:, i] > -d[i]) & (xr[:, i] < d[i])
    return ~inside


def filter_atomspec_drop(ID, filter_dict, intersect):
    """Indices removed by filter_atomspec (CPET/utils/io.py:341-419).
    Pinned by the reference test KAT tests/test_filter.py:130-183."""
    keys = ["atom_number", "atom_type", "resid", "resnum", "chain"]
    drop = set()
    if not intersect:
        for k in filter_dict:
            vals = filter_dict[k]
            for j, row in enumerate(ID):
                if row[keys.index(k)] in vals:
                    drop.add(j)
    else:
        n = len(next(iter(filter_dict.values())))
        for i in range(n):
            for j, row in enumerate(ID):
                ok = True
                for k in filter_dict:
                    v = filter_dict[k][i]
                    if v != "*" and v != "" and row[keys.index(k)] != v:
                        ok = False
                        break
                if ok:
                    drop.add(j)
    return sorted(drop)


# ----------------------------------------------------------- cluster step
def cluster_histograms(features):
    """make_histograms (CPET/utils/calculator.py:596-718) on in-memory
    per-frame (L, 2) features: the reference's own steps with the libraries
    it calls (scipy.stats.iqr, numpy.histogram2d).  Returns (hists (F, B),
    (xedges, yedges)).  Pinned by tests/test_oracle_cluster.py (brute-force
    counts, normalisation, edge placement)."""
    from scipy.stats import iqr

    lens = np.array([len(f) for f in features], dtype=int)
    n = lens[0] if np.all(lens == lens[0]) else np.mean(lens)
    d = np.concatenate([np.asarray(f, float)[:, 0] for f in features])
    c = np.concatenate([np.asarray(f, float)[:, 1] for f in features])
    dres = 2 * iqr(d) / (n ** (1 / 3))
    cres = 2 * iqr(c) / (n ** (1 / 3))
    nd = int((np.max(d) - np.min(d)) / dres)
    nc = int((np.max(c) - np.min(c)) / cres)
    out, edges = [], None
    for f in features:
        f = np.asarray(f, float)
        a, xe, ye = np.histogram2d(f[:, 0], f[:, 1], bins=[nd, nc],
                                   range=[[np.min(d), np.max(d)], [np.min(c), np.max(c)]])
        out.append((a / np.sum(a)).flatten())
        edges = (xe, ye)
    return np.array(out), edges


def make_histograms_files(topo_files):
    """make_histograms (CPET/utils/calculator.py:596-718) from .top files,
    with the reference's file passes: a line count per file (:603-606), a
    parse of every row into the global arrays (:618-649), then per file a
    second parse for its histogram (:690-707); bins from scipy iqr and the
    global min / max.  Returns (F, B).  Timed as the CPU counterpart of the
    cluster step from files; its arithmetic is cluster_histograms'.  Pinned
    by tests/test_oracle_cluster.py (test_oracle_files_path_equals_in_memory:
    equal to cluster_histograms on the same rows, whose bins and counts are
    pinned to numpy and scipy there)."""
    from scipy.stats import iqr

    lens = np.zeros(len(topo_files), dtype=int)
    for k, path in enumerate(topo_files):
        with open(path) as fh:
            lens[k] = sum(1 for line in fh if not line.startswith("#"))
    d_all = np.zeros(int(lens.sum()))
    c_all = np.zeros(int(lens.sum()))
    pos = 0
    for k, path in enumerate(topo_files):
        with open(path) as fh:
            for line in fh:
                if line.startswith("#"):
                    continue
                parts = line.strip().split()
                d_all[pos] = float(parts[0])
                c_all[pos] = float(parts[1])
                pos += 1
    n = lens[0] if np.all(lens == lens[0]) else np.mean(lens)
    dmax, cmax, dmin, cmin = np.max(d_all), np.max(c_all), np.min(d_all), np.min(c_all)
    nd = int((dmax - dmin) / (2 * iqr(d_all) / (n ** (1 / 3))))
    nc = int((cmax - cmin) / (2 * iqr(c_all) / (n ** (1 / 3))))
    out = []
    for path in topo_files:
        dist, curv = [], []
        with open(path) as fh:
            for line in fh:
                if line.startswith("#"):
                    continue
                parts = line.split()
                dist.append(float(parts[0]))
                curv.append(float(parts[1]))
        a, _, _ = np.histogram2d(dist, curv, bins=[nd, nc],
                                 range=[[dmin, dmax], [cmin, cmax]])
        out.append((a / np.sum(a)).flatten())
    return np.array(out)


def cluster_histograms_mem(features):
    """make_histograms_mem (CPET/utils/calculator.py:721-845) on in-memory
    per-frame (L, 2) features: Python min / max over the rows, scipy iqr,
    matplotlib's hist2d with density=True, then each value divided by the
    row-by-row sum, as the reference does.  Returns (F, B) float64.  Pinned
    by tests/test_oracle_cluster.py (brute-force counts / total, the edges
    of cluster_histograms)."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from scipy.stats import iqr

    dist_list, curv_list, lens = [], [], []
    for f in features:
        f = np.asarray(f, float)
        dist_list.extend(f[:, 0].tolist())
        curv_list.extend(f[:, 1].tolist())
        lens.append(len(f))
    n = lens[0] if all(x == lens[0] for x in lens) else np.mean(lens)
    dmax, cmax, dmin, cmin = max(dist_list), max(curv_list), min(dist_list), min(curv_list)
    dres = 2 * iqr(dist_list) / (n ** (1 / 3))
    cres = 2 * iqr(curv_list) / (n ** (1 / 3))
    nd = int((dmax - dmin) / dres)
    nc = int((cmax - cmin) / cres)
    out = []
    fig = plt.figure()
    try:
        for f in features:
            f = np.asarray(f, float)
            a, _, _, _ = plt.hist2d(f[:, 0].tolist(), f[:, 1].tolist(), bins=(nd, nc),
                                    range=[[dmin, dmax], [cmin, cmax]], density=True)
            norm = 0
            for row in a:
                for m in row:
                    norm += m
            out.append(np.array([[m / norm for m in row] for row in a]).flatten())
    finally:
        plt.close(fig)
    return np.array(out)


def chi2_matrix_mem(hists):
    """construct_distance_matrix_mem (CPET/utils/calculator.py:981-1000) on
    in-memory histograms: distance_numpy over every pair i < j, mirrored,
    zero diagonal.  Pinned by tests/test_oracle_cluster.py
    (test_oracle_chi2_mem_equals_pairwise_path: equal to chi2_matrix, whose
    closed forms pin the pair term)."""
    n = len(hists)
    m = np.zeros((n, n), dtype=np.float64)
    for i in range(n):
        for j in range(i + 1, n):
            m[i][j] = chi2_distance(np.asarray(hists[i]), np.asarray(hists[j]))
            m[j][i] = m[i][j]
    return m


def chi2_distance(h1, h2):
    """distance_numpy, CPET/utils/calculator.py:975-978.  Pinned through
    chi2_matrix, which applies it to every pair: its closed forms in
    tests/test_oracle_cluster.py (identical -> 0, disjoint supports -> 1, a
    hand-computed pair, symmetry)."""
    a = (h1 - h2) ** 2
    b = h1 + h2
    return np.sum(np.divide(a, b, out=np.zeros_like(a), where=b != 0)) / 2.0


def chi2_matrix(hists):
    """construct_distance_matrix (CPET/utils/calculator.py:1000-1015):
    sklearn pairwise_distances with distance_numpy, zero diagonal.  Pinned
    by tests/test_oracle_cluster.py (closed forms: identical -> 0, disjoint
    supports -> 1, a hand-computed pair, symmetry)."""
    from sklearn.metrics.pairwise import pairwise_distances

    m = pairwise_distances(np.asarray(hists, float), metric=chi2_distance, n_jobs=1)
    np.fill_diagonal(m, 0.0)
    return m
