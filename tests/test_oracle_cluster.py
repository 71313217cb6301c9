"""The cluster-step oracle (numpy.histogram2d / scipy iqr / sklearn over the
reference's distance_numpy) pinned against brute force and closed forms,
and the host-side pieces of pycpet_amd.clusterhist that need no GPU: the
shared bin edges and the .top reader."""
import numpy as np
import pytest


def _brute_counts(f, xe, ye):
    nd, nc = len(xe) - 1, len(ye) - 1
    h = np.zeros((nd, nc))
    for d, c in f:
        i = next((k for k in range(nd) if xe[k] <= d < xe[k + 1]), nd - 1 if d == xe[-1] else -1)
        j = next((k for k in range(nc) if ye[k] <= c < ye[k + 1]), nc - 1 if c == ye[-1] else -1)
        if i >= 0 and j >= 0:
            h[i, j] += 1
    return h


def _frames(seed, F=5, L=300):
    rng = np.random.default_rng(seed)
    return [np.stack([rng.gamma(2.0, 0.3, L), rng.lognormal(0.0, 0.5, L)], 1).astype(np.float32)
            .astype(np.float64) for _ in range(F)]


def test_oracle_histograms_brute_force(oracle_mod):
    feats = _frames(0)
    hists, (xe, ye) = oracle_mod.cluster_histograms(feats)
    for f, h in zip(feats, hists):
        cnt = _brute_counts(f, xe, ye)
        np.testing.assert_array_equal(h, (cnt / cnt.sum()).ravel())
        assert abs(h.sum() - 1.0) < 1e-12
    # the global extremes sit on the outer edges and are counted
    allf = np.concatenate(feats)
    assert xe[0] == allf[:, 0].min() and xe[-1] == allf[:, 0].max()


def test_oracle_chi2_closed_forms(oracle_mod):
    h = np.array([[0.5, 0.5, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.5, 0.5, 0.0]])
    m = oracle_mod.chi2_matrix(h)
    assert m[0, 3] == 0.0 and np.all(np.diag(m) == 0.0)
    assert m[1, 2] == 1.0 and m[0, 2] == 1.0          # disjoint supports
    assert abs(m[0, 1] - (0.25 / 1.5 + 0.25 / 0.5) / 2) < 1e-15  # hand value
    np.testing.assert_array_equal(m, m.T)


def _np_kth(col):
    """r -> the r-th smallest of col, NaNs last (numpy's sort order)."""
    srt = np.sort(np.asarray(col, np.float64))
    return lambda r: srt[r]


def _edges_numpy_stats(feats):
    """histogram_edges' host steps with numpy's own order statistics in place
    of the GPU selection."""
    from pycpet_amd import clusterhist as CH

    n = CH._binning_count(feats)
    allf = CH._flatten(feats)[2]
    P = allf.shape[0]
    stats = [CH._stats_from_order(P, _np_kth(allf[:, c]), int(np.isnan(allf[:, c]).sum()))
             for c in range(2)]
    return CH._edges_from_stats(stats, n)


def test_bin_edges_match_numpy_histogram2d(oracle_mod):
    for seed in range(3):
        feats = _frames(seed, F=4, L=200 + 50 * seed)
        if seed:  # ragged frames: binning uses the mean row count
            feats = [f[:len(f) - 17 * k] for k, f in enumerate(feats)]
        if seed == 0:
            ed, ec = _edges_numpy_stats(feats)
        else:
            with pytest.warns(UserWarning, match="different sizes"):
                ed, ec = _edges_numpy_stats(feats)
        _, (xe, ye) = oracle_mod.cluster_histograms(feats)
        np.testing.assert_array_equal(ed, xe)
        np.testing.assert_array_equal(ec, ye)


def _adversarial_columns():
    rng = np.random.default_rng(21)
    cols = [np.array([v]) for v in (2.5, -0.0, np.inf, -np.inf)]
    for n in (2, 3, 4, 5, 6, 7, 8, 9, 13):
        cols.append(rng.standard_normal(n))
        cols.append(rng.integers(-2, 3, n).astype(np.float64))  # ties
    cols.append(np.array([1.0, np.inf]))
    cols.append(np.array([-np.inf, 0.0, 1.0, np.inf]))
    cols.append(np.array([0.0, -0.0, 0.0, -0.0, 1e-310, -5e-324]))
    cols.append(np.array([1.0, np.nan, 2.0]))
    cols.append(np.array([1e308, -1e308, 1.7976931348623157e308, 3.0]))
    cols.append(rng.standard_normal(1001) * 1e-300)
    cols.append(np.repeat(rng.standard_normal(7), 143))
    return cols


def test_column_stats_match_numpy_and_scipy():
    """min, max and the IQR from order statistics equal np.min, np.max and
    scipy.stats.iqr bit for bit, with numpy's NaN and infinity results
    (inf - inf in the interpolation gives NaN there too)."""
    from scipy.stats import iqr

    from pycpet_amd import clusterhist as CH

    for col in _adversarial_columns():
        with np.errstate(invalid="ignore"):
            got = CH._stats_from_order(len(col), _np_kth(col), int(np.isnan(col).sum()))
            want = (np.min(col), np.max(col), iqr(col))
        np.testing.assert_array_equal(np.array(got), np.array(want), err_msg=repr(col))


def test_order_stat_ranks_cover_the_interpolation():
    from pycpet_amd import clusterhist as CH

    for n in (1, 2, 3, 4, 5, 100, 5_120_000):
        ranks = CH._order_stat_ranks(n)
        assert ranks[0] == 0 and ranks[-1] == n - 1 and len(ranks) <= 6
        for q in CH._IQR_Q:
            i, j, g = CH._lerp_points(n, q)
            assert i in ranks and j in ranks and 0 <= i <= j < n
            assert g == (n - 1) * q - i or n == 1


def test_topo_reader_matches_loadtxt(tmp_path):
    from pycpet_amd.clusterhist import read_topo_file

    rng = np.random.default_rng(5)
    f = rng.standard_normal((1000, 2)).astype(np.float32)
    p = tmp_path / "a.top"
    with open(p, "w") as fh:
        fh.write("# header line\n")
        np.savetxt(fh, f)
        fh.write("#trailer\n")
    got = read_topo_file(p)
    np.testing.assert_array_equal(got, np.loadtxt(p))
    assert got.dtype == np.float64 and got.shape == (1000, 2)
    p.write_text("1.0 2.0\n\n3.0 4.0\n")
    from pycpet_amd._native import CpetError
    with pytest.raises(CpetError, match="line 2"):
        read_topo_file(p)


def test_oracle_mem_histograms_brute_force(oracle_mod):
    """make_histograms_mem's oracle (matplotlib hist2d with density=True,
    then the reference's row-by-row normalisation) against brute-force
    counts: the values are count / total to rounding, sum to 1, and use the
    same edges as make_histograms.  The host step that turns counts into
    those values (clusterhist._density_normalised) reproduces them bit for
    bit from the brute-force counts."""
    from pycpet_amd import clusterhist as CH

    for seed in range(3):
        feats = _frames(10 + seed, F=4, L=250)
        if seed == 2:
            feats = [f[:len(f) - 23 * k] for k, f in enumerate(feats)]
        mem = oracle_mod.cluster_histograms_mem(feats)
        _, (xe, ye) = oracle_mod.cluster_histograms(feats)
        for f, h in zip(feats, mem):
            cnt = _brute_counts(f, xe, ye)
            np.testing.assert_allclose(h, (cnt / cnt.sum()).ravel(), rtol=1e-14, atol=1e-17)
            assert abs(h.sum() - 1.0) < 1e-12
            np.testing.assert_array_equal(CH._density_normalised(cnt, xe, ye), h)


def test_oracle_chi2_mem_equals_pairwise_path(oracle_mod):
    rng = np.random.default_rng(2)
    h = rng.random((9, 40)) ** 3
    h[rng.random(h.shape) < 0.3] = 0.0
    h /= h.sum(1, keepdims=True)
    m = oracle_mod.chi2_matrix_mem(list(h))
    np.testing.assert_allclose(m, oracle_mod.chi2_matrix(h), rtol=1e-14, atol=1e-17)
    assert np.all(np.diag(m) == 0.0)


def test_mem_reader_commas_and_whitespace(tmp_path):
    from pycpet_amd import clusterhist as CH

    p = tmp_path / "c.top"
    p.write_text("# header\n1.5,2.25\n 3.0 , 4.0 \n5e-1\t6\n#x\n7,8,9\n")
    np.testing.assert_array_equal(CH._read_topo_any_sep(p),
                                  [[1.5, 2.25], [3.0, 4.0], [0.5, 6.0], [7.0, 8.0]])
    q = tmp_path / "w.top"
    q.write_text("# header\n1.5 2.25\n3 4 99\n")
    np.testing.assert_array_equal(CH._read_topo_any_sep(q), [[1.5, 2.25], [3.0, 4.0]])


def test_oracle_files_path_equals_in_memory(oracle_mod, tmp_path):
    """The from-files restatement (three passes over each .top file) gives
    the in-memory oracle's histograms bit for bit on savetxt-written rows."""
    feats = _frames(4, F=3, L=400)
    files = []
    for k, f in enumerate(feats):
        p = tmp_path / f"f{k}.top"
        np.savetxt(p, f)
        files.append(str(p))
    want, _ = oracle_mod.cluster_histograms(feats)
    np.testing.assert_array_equal(oracle_mod.make_histograms_files(files), want)
