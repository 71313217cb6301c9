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


def test_bin_edges_match_numpy_histogram2d(oracle_mod):
    from pycpet_amd.clusterhist import histogram_edges

    for seed in range(3):
        feats = _frames(seed, F=4, L=200 + 50 * seed)
        if seed:  # ragged frames: binning uses the mean row count
            feats = [f[:len(f) - 17 * k] for k, f in enumerate(feats)]
        if seed == 0:
            ed, ec = histogram_edges(feats)
        else:
            with pytest.warns(UserWarning, match="different sizes"):
                ed, ec = histogram_edges(feats)
        _, (xe, ye) = oracle_mod.cluster_histograms(feats)
        np.testing.assert_array_equal(ed, xe)
        np.testing.assert_array_equal(ec, ye)


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
