"""Cluster step on the GPU against the oracle (the reference's own
numpy.histogram2d / scipy iqr / sklearn-over-distance_numpy path):
histograms bit-exact (exact counts, correctly rounded count / total), the
chi-square matrix to 1e-12 relative (FP64; bins summed in order where numpy
sums pairwise), end to end from .top files."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(seed, F, L, ragged=False):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(F):
        n = L - (37 * k) % 200 if ragged else L
        d = rng.gamma(2.0, 0.3, n)
        c = rng.lognormal(0.0, 0.5, n)
        out.append(np.stack([d, c], 1).astype(np.float32).astype(np.float64))
    return out


@pytest.mark.parametrize("ragged", [False, True])
def test_histograms_bit_exact(gpu_ctx, oracle_mod, ragged):
    from pycpet_amd import clusterhist as CH

    feats = _frames(1 + ragged, 24, 2000, ragged)
    # put values exactly on interior and outer edges
    want, (xe, ye) = oracle_mod.cluster_histograms(feats)
    feats[3][:5, 0] = xe[[0, 1, 7, -2, -1]]
    feats[3][:5, 1] = ye[[-1, 3, 0, 2, -2]]
    want, (xe, ye) = oracle_mod.cluster_histograms(feats)
    if ragged:  # the reference warns that the frames differ in size
        with pytest.warns(UserWarning, match="different sizes"):
            got = CH.histograms_from_features(feats, ctx=gpu_ctx)
    else:
        got = CH.histograms_from_features(feats, ctx=gpu_ctx)
    print(f"F={len(feats)} bins={len(xe) - 1}x{len(ye) - 1}", flush=True)
    assert got.shape == want.shape
    np.testing.assert_array_equal(got, want)


def test_histograms_with_empty_frames(gpu_ctx, oracle_mod):
    """Frames with no rows (empty .top files) first, in the middle and last:
    their histograms are 0 / 0 = NaN in every bin, as numpy's normalisation
    gives, and every other frame, the shared bins and the reference's
    mean-length rule for unequal frames are as the oracle has them."""
    import warnings

    from pycpet_amd import clusterhist as CH

    feats = _frames(5, 6, 1500, ragged=True)
    feats.insert(0, np.zeros((0, 2)))
    feats.insert(3, np.zeros((0, 2)))
    feats.append(np.zeros((0, 2)))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # the unequal-sizes warning; 0 / 0
        want, _ = oracle_mod.cluster_histograms(feats)
        got = CH.histograms_from_features(feats, ctx=gpu_ctx)
    empty = [k for k, f in enumerate(feats) if len(f) == 0]
    print(f"F={len(feats)} empty frames {empty}", flush=True)
    assert empty == [0, 3, len(feats) - 1]
    assert np.isnan(got[empty]).all() and not np.isnan(np.delete(got, empty, 0)).any()
    np.testing.assert_array_equal(got, want)


def test_chi2_matrix_vs_sklearn(gpu_ctx, oracle_mod):
    from pycpet_amd import clusterhist as CH

    feats = _frames(7, 70, 1500)
    h, _ = oracle_mod.cluster_histograms(feats)
    want = oracle_mod.chi2_matrix(h)
    got = CH.construct_distance_matrix(h, ctx=gpu_ctx)
    assert got.shape == (70, 70)
    assert np.all(np.diag(got) == 0.0)
    np.testing.assert_array_equal(got, got.T)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
    # closed forms through the kernel: identical rows, disjoint supports
    t = np.array([[0.5, 0.5, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.5, 0.5, 0.0]])
    m = CH.construct_distance_matrix(t, ctx=gpu_ctx)
    assert m[0, 3] == 0.0 and m[1, 2] == 1.0 and m[0, 2] == 1.0


def test_make_histograms_from_top_files(gpu_ctx, oracle_mod, tmp_path):
    from pycpet_amd import clusterhist as CH
    from pycpet_amd.io import write_rows

    feats = _frames(11, 9, 800, ragged=True)
    files = []
    for k, f in enumerate(feats):
        p = tmp_path / f"frame{k}.top"
        write_rows(p, f.astype(np.float32), "%.18e")  # as CPET.run_topo writes them
        files.append(str(p))
    with pytest.warns(UserWarning, match="different sizes"):
        got = CH.make_histograms(files, ctx=gpu_ctx)
    want, _ = oracle_mod.cluster_histograms([np.loadtxt(p) for p in files])
    np.testing.assert_array_equal(got, want)
    d = CH.construct_distance_matrix(got, ctx=gpu_ctx)
    np.testing.assert_allclose(d, oracle_mod.chi2_matrix(want), rtol=1e-12, atol=1e-15)


def test_large_matrix_properties(gpu_ctx):
    """1,000 frames x 4,096 bins (1 M pairs): symmetry, zero diagonal, the
    range [0, 1] of the chi-square distance of normalised histograms, and a
    random sample of entries against the host formula."""
    from pycpet_amd import clusterhist as CH

    rng = np.random.default_rng(3)
    h = rng.random((1000, 4096)) ** 4
    h[rng.random(h.shape) < 0.6] = 0.0
    h /= h.sum(1, keepdims=True)
    m = CH.construct_distance_matrix(h, ctx=gpu_ctx)
    np.testing.assert_array_equal(m, m.T)
    assert np.all(np.diag(m) == 0.0) and m.min() >= 0.0 and m.max() <= 1.0 + 1e-12
    for i, j in rng.integers(0, 1000, (40, 2)):
        if i != j:
            assert abs(m[i, j] - CH.distance_numpy(h[i], h[j])) <= 1e-12 * max(m[i, j], 1e-300)


def _order_stats(ctx, x, sel):
    """cpet_order_stats on the (n, cols) array x for [(col, rank)]."""
    from pycpet_amd import _native as nat

    x = np.ascontiguousarray(x, np.float64)
    n, cols = x.shape
    sc = np.array([c for c, _ in sel], np.int32)
    sr = np.array([r for _, r in sel], np.int64)
    vals = np.empty(len(sel))
    nans = np.zeros(cols, np.int64)
    nat.check(ctx.lib.cpet_order_stats(ctx.handle, n, cols, x.ctypes.data, len(sel),
                                       sc.ctypes.data, sr.ctypes.data, vals.ctypes.data,
                                       nans.ctypes.data, 0), ctx.lib)
    return vals, nans


def _sorted_cols(x):
    """numpy's order (NaNs last) with -0.0 read as +0.0."""
    s = np.sort(x, axis=0)
    s[s == 0.0] = 0.0
    return s


def test_order_stats_exact_on_awkward_columns(gpu_ctx):
    """Every requested k-th smallest value is the element numpy's sort puts at
    k: ties, signed zeros, subnormals, infinities, NaNs (counted, ordered
    last), one to eight columns, 1 to 16 selections, sizes 1 to 300k."""
    rng = np.random.default_rng(4)
    cases = []
    for n in (1, 2, 3, 7, 64, 255, 256, 257, 1000, 300_001):
        for cols in (1, 2, 3, 8):
            x = rng.standard_normal((n, cols)) * 10.0 ** rng.integers(-300, 300, (n, cols))
            kind = rng.integers(0, 4)
            if kind == 1:
                x = rng.integers(-3, 4, (n, cols)).astype(np.float64)  # heavy ties
            elif kind == 2:
                pick = rng.random((n, cols))
                x[pick < 0.1] = -0.0
                x[(pick >= 0.1) & (pick < 0.2)] = 0.0
                x[(pick >= 0.2) & (pick < 0.25)] = 5e-324
                x[(pick >= 0.25) & (pick < 0.28)] = -np.inf
                x[(pick >= 0.28) & (pick < 0.3)] = np.inf
            elif kind == 3:
                x[rng.random((n, cols)) < 0.05] = np.nan
            cases.append(x)
    for x in cases:
        n, cols = x.shape
        m = int(rng.integers(1, 17))
        sel = [(int(rng.integers(0, cols)), int(r)) for r in rng.integers(0, n, m)]
        sel[0] = (sel[0][0], 0)
        sel[-1] = (sel[-1][0], n - 1)
        print(f"order stats n={n} cols={cols} sel={len(sel)}", flush=True)
        vals, nans = _order_stats(gpu_ctx, x, sel)
        srt = _sorted_cols(x)
        want = np.array([srt[r, c] for c, r in sel])
        np.testing.assert_array_equal(vals, want)
        assert not np.any(np.signbit(vals[vals == 0.0]))
        np.testing.assert_array_equal(nans, np.isnan(x).sum(0))


def test_order_stats_bench_size(gpu_ctx):
    """The cluster workload's 256 frames x 20,000 rows in one call: the six
    ranks per column that min, max and the quartiles read."""
    from pycpet_amd import clusterhist as CH

    rng = np.random.default_rng(8)
    x = np.stack([rng.gamma(2.0, 0.3, 5_120_000), rng.lognormal(0.0, 0.5, 5_120_000)], 1)
    x = x.astype(np.float32).astype(np.float64)  # float32 .top values: many ties
    ranks = CH._order_stat_ranks(len(x))
    sel = [(c, r) for c in range(2) for r in ranks]
    vals, nans = _order_stats(gpu_ctx, x, sel)
    for c in range(2):
        part = np.partition(x[:, c], ranks)
        for k, r in enumerate(ranks):
            assert vals[c * len(ranks) + k] == part[r]
    assert not nans.any()


def test_histogram_edges_on_gpu_match_oracle(gpu_ctx, oracle_mod):
    """histogram_edges (GPU order statistics + numpy's interpolation on the
    host) equals the oracle's scipy.stats.iqr / np.min / np.max edges bit for
    bit, for equal and ragged frames, and fails as the reference does on a
    NaN feature (int() of a NaN bin count)."""
    from pycpet_amd import clusterhist as CH

    for seed, ragged in ((0, False), (1, True), (2, False)):
        feats = _frames(20 + seed, 40, 3000, ragged)
        if seed == 2:  # float32 rows with exact duplicates of the quartiles
            feats[5][:100] = feats[0][:100]
        _, (xe, ye) = oracle_mod.cluster_histograms(feats)
        if ragged:
            with pytest.warns(UserWarning, match="different sizes"):
                ed, ec = CH.histogram_edges(feats, ctx=gpu_ctx)
        else:
            ed, ec = CH.histogram_edges(feats, ctx=gpu_ctx)
        np.testing.assert_array_equal(ed, xe)
        np.testing.assert_array_equal(ec, ye)
    feats = _frames(30, 4, 500)
    feats[2][7, 1] = np.nan
    with pytest.raises(ValueError, match="NaN"):
        CH.histogram_edges(feats, ctx=gpu_ctx)


def test_order_stats_rejects_bad_selections(gpu_ctx):
    from pycpet_amd._native import CpetError

    x = np.zeros((10, 2))
    for sel in ([(0, 10)], [(2, 0)], [(0, -1)], [(0, 0)] * 17):
        with pytest.raises(CpetError, match="cpet_order_stats"):
            _order_stats(gpu_ctx, x, sel)
    with pytest.raises(CpetError, match="cpet_order_stats"):
        _order_stats(gpu_ctx, np.zeros((4, 9)), [(0, 0)])


def test_make_histograms_mem_files(gpu_ctx, oracle_mod, tmp_path):
    """The file-by-file variant: per-frame .npy histograms bit-equal to the
    oracle's (hist2d density, then the reference's normalisation), frames
    written with commas or whitespace, and the distance matrix from those
    files against the pairwise oracle."""
    from pycpet_amd import clusterhist as CH
    from pycpet_amd.io import write_rows

    feats = _frames(40, 7, 900, ragged=True)
    files = []
    for k, f in enumerate(feats):
        p = tmp_path / f"frame{k}.top"
        if k % 3 == 1:
            p.write_text("# comma separated\n" + "".join(f"{a!r},{b!r}\n" for a, b in f.tolist()))
        else:
            write_rows(p, f, "%.18e")
        files.append(str(p))
    out = tmp_path / "hists"
    out.mkdir()
    with pytest.warns(UserWarning, match="different sizes"):
        paths = CH.make_histograms_mem(files, str(out), ctx=gpu_ctx)
    assert paths == [str(out / f"frame{k}_hist.npy") for k in range(len(files))]
    want = oracle_mod.cluster_histograms_mem(feats)
    got = np.stack([np.load(p) for p in paths])
    np.testing.assert_array_equal(got, want)
    m = CH.construct_distance_matrix_mem(paths, ctx=gpu_ctx)
    np.testing.assert_allclose(m, oracle_mod.chi2_matrix_mem(list(want)), rtol=1e-12, atol=1e-15)
    assert np.all(np.diag(m) == 0.0)
