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
    got = CH.histograms_from_features(feats, ctx=gpu_ctx)
    print(f"F={len(feats)} bins={len(xe) - 1}x{len(ye) - 1}", flush=True)
    assert got.shape == want.shape
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
