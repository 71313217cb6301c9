"""GPU parity of the streamline kernels against the FP64 oracle.

Matching rule (stated once): a streamline "matches" when its step count and
stop kind (box exit / step limit) equal the oracle's — indexing and counting
are bit-exact.  A one-ulp difference in a field sum can move a box crossing
across a step boundary, so at least 99% of streamlines must match; on the
matched set the features must agree within 1e-4 relative:
    |d_gpu - d_64| <= 1e-4 |d_64|,  |k_gpu - k_64| <= 1e-4 max(|k_64|, 1e-6).
Histogram counts must equal numpy.histogram2d of the GPU's own (float32-
rounded) features exactly, and equal the oracle's histogram wherever the
binned features match.
"""
import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FEAT_TOL = 1e-4
MATCH_MIN = 0.99


def compare(r, feats64, steps64, exited64, label):
    m = (r.steps == steps64) & (r.exited == exited64)
    frac = float(m.mean())
    dd = np.abs(r.feats[m, 0] - feats64[m, 0]) / np.maximum(np.abs(feats64[m, 0]), 1e-12)
    dk = np.abs(r.feats[m, 1] - feats64[m, 1]) / np.maximum(np.abs(feats64[m, 1]), 1e-6)
    print(f"{label}: match {frac:.5f} dist max {dd.max():.3e} curv max {dk.max():.3e}", flush=True)
    assert frac >= MATCH_MIN, frac
    assert dd.max() <= FEAT_TOL and dk.max() <= FEAT_TOL
    return m


def test_small_vs_oracle(gpu_ctx, oracle_mod):
    from pycpet_amd import ops, workloads as W

    c = W.small_topo(N=300, L=512, cap=400, seed=5)
    print(f"case {c.name}", flush=True)
    r = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, ctx=gpu_ctx)
    f64, s64, e64 = oracle_mod.topo_f64(c.starts, c.n_iter, c.h, c.dims, c.x, c.q)
    compare(r, f64, s64, e64, c.name)


def test_h1e3_vs_oracle(gpu_ctx, oracle_mod):
    """h = 1e-3 A as in the 100k-streamline workload, on a case the CPU oracle
    finishes in seconds."""
    from pycpet_amd import ops, workloads as W

    c = W.small_topo(N=800, L=384, cap=1500, seed=6, dims=(1.5, 1.5, 1.5), h=0.001, R=30.0)
    print(f"case {c.name} h=1e-3", flush=True)
    r = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, ctx=gpu_ctx)
    f64, s64, e64 = oracle_mod.topo_f64(c.starts, c.n_iter, c.h, c.dims, c.x, c.q)
    compare(r, f64, s64, e64, c.name)


def test_radial_lines_exact_steps(gpu_ctx):
    from pycpet_amd import ops
    from test_oracle_topo import _radial_expect

    rng = np.random.default_rng(7)
    dims = np.array([0.6, 0.5, 0.4], np.float32)
    h = 0.01
    c = np.array([3.0, 0.5, -0.4])
    for sign in (+1.0, -1.0):
        starts = rng.uniform(-0.35, 0.35, (300, 3)).astype(np.float32)
        n_iter = rng.integers(1, 200, 300).astype(np.int32)
        r = ops.topo(starts, n_iter, h, dims, c[None].astype(np.float32),
                     np.array([sign * 0.8], np.float32), ctx=gpu_ctx)
        exp_steps, clean, exp_exit = _radial_expect(
            starts, c.astype(np.float32).astype(float), sign, dims.astype(float),
            float(np.float32(h)), n_iter)
        np.testing.assert_array_equal(r.steps[clean], exp_steps[clean])
        np.testing.assert_array_equal(r.exited[clean].astype(bool), exp_exit[clean])
        np.testing.assert_allclose(r.feats[:, 0], r.steps * float(np.float32(h)), rtol=1e-6)
        assert np.abs(r.feats[:, 1]).max() < 1e-3


def test_histogram_exact(gpu_ctx, oracle_mod):
    from pycpet_amd import ops, workloads as W

    c = W.small_topo(N=300, L=1024, cap=400, seed=8)
    f64, s64, e64 = oracle_mod.topo_f64(c.starts, c.n_iter, c.h, c.dims, c.x, c.q)
    ed = np.linspace(f64[:, 0].min(), f64[:, 0].max(), 21)
    ec = np.linspace(f64[:, 1].min(), f64[:, 1].max(), 21)
    r = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, edges=(ed, ec), ctx=gpu_ctx)
    fr = r.feats.astype(np.float32).astype(np.float64)
    Hn, _, _ = np.histogram2d(fr[:, 0], fr[:, 1], bins=[ed, ec])
    np.testing.assert_array_equal(r.hist, Hn.astype(np.uint32))
    # vs the oracle: identical counts over streamlines whose binned features agree
    fo = f64.astype(np.float32).astype(np.float64)
    same_bin = (np.searchsorted(ed, fr[:, 0], "right") == np.searchsorted(ed, fo[:, 0], "right")) & \
               (np.searchsorted(ec, fr[:, 1], "right") == np.searchsorted(ec, fo[:, 1], "right"))
    assert same_bin.mean() >= MATCH_MIN
    Ha, _, _ = np.histogram2d(fr[same_bin, 0], fr[same_bin, 1], bins=[ed, ec])
    Hb = oracle_mod.histogram2d(fo[same_bin, 0], fo[same_bin, 1], ed, ec)
    np.testing.assert_array_equal(Ha.astype(np.int64), Hb)


def test_schedule_invariance(gpu_ctx):
    """Which slot or workgroup runs a streamline does not change its result;
    the charge-slice count only regroups FP64 partial sums."""
    from pycpet_amd import ops, workloads as W

    c = W.small_topo(N=700, L=900, cap=300, seed=9)
    runs = {}
    for S, G in ((8, 1), (8, 3), (8, 0), (1, 0), (4, 2), (16, 0)):
        gpu_ctx.set_tuning(2, S, G)
        runs[(S, G)] = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, ctx=gpu_ctx)
    gpu_ctx.set_tuning(2, 0, 0)
    base = runs[(8, 0)]
    for key in ((8, 1), (8, 3)):
        np.testing.assert_array_equal(runs[key].feats, base.feats)
        np.testing.assert_array_equal(runs[key].steps, base.steps)
    for key in ((1, 0), (4, 2), (16, 0)):
        m = runs[key].steps == base.steps
        assert m.mean() >= MATCH_MIN
        np.testing.assert_allclose(runs[key].feats[m], base.feats[m], rtol=1e-6, atol=1e-9)


def test_reference_golden_vectors(gpu_ctx):
    from pycpet_amd import ops

    g = np.load(GOLDEN / "ref_topo_200.npz")
    r = ops.topo(g["starts"], g["n_iter"], float(g["h"]), g["dims"], g["x"], g["q"], ctx=gpu_ctx)
    ref = g["ret"].astype(float)
    rd = np.abs(ref[:, 0] - r.feats[:, 0]) / np.maximum(r.feats[:, 0], 1e-12)
    assert np.mean(rd < 1e-4) >= 0.95
    ok = rd < 1e-4
    assert np.median(np.abs(ref[ok, 1] - r.feats[ok, 1]) / np.abs(r.feats[ok, 1])) < 1e-2


def test_edge_cases(gpu_ctx, oracle_mod):
    from pycpet_amd import ops
    from pycpet_amd._native import CpetError
    from pycpet_amd.c_ops import Math_ops

    x = np.array([[2.0, 1.0, -1.5], [-1.7, 0.4, 2.2]], np.float32)
    q = np.array([0.3, -0.6], np.float32)
    dims = np.full(3, 1.0, np.float32)
    starts = np.array([[0.1, 0.0, 0.2], [0.0, 0.1, 0.0], [0.99, 0.0, 0.0], [0.2, 0.3, -0.1]],
                      np.float32)
    n_iter = np.array([1, 0, 5, 2000], np.int32)
    r = ops.topo(starts, n_iter, 0.01, dims, x, q, ctx=gpu_ctx)
    f64, s64, e64 = oracle_mod.topo_f64(starts, n_iter, 0.01, dims, x, q)
    np.testing.assert_array_equal(r.steps, s64)
    np.testing.assert_array_equal(r.exited, e64)
    np.testing.assert_allclose(r.feats, f64, rtol=1e-5, atol=1e-9)
    assert r.steps[1] == 0 and r.feats[1, 0] == 0.0
    # single streamline through the reference-named entry point
    ret = Math_ops().thread_operation(starts[3], 2000, x, q, 0.01, dims)
    assert ret.dtype == np.float32 and ret.shape == (2,)
    np.testing.assert_allclose(ret, f64[3].astype(np.float32), rtol=1e-5)
    with pytest.raises(CpetError):
        ops.topo(starts, n_iter, float("nan"), dims, x, q, ctx=gpu_ctx)
    with pytest.raises(CpetError):
        ops.topo(starts, n_iter, 0.01, np.array([1.0, -1.0, 1.0], np.float32), x, q, ctx=gpu_ctx)


def test_boundary_starts(gpu_ctx, oracle_mod):
    """Starts outside the box (they leave after one step), exactly on faces,
    edges and a corner (inside: the exit test is strict), and just inside,
    against the FP64 oracle (itself pinned on these starts to the
    reference C): step counts and stop kinds exact, features to 1e-5."""
    from pycpet_amd import ops

    x = np.array([[2.0, 1.0, -1.5], [-1.7, 0.4, 2.2], [0.3, -2.5, 0.8]], np.float32)
    q = np.array([0.3, -0.6, 0.45], np.float32)
    dims = np.full(3, 1.0, np.float32)
    starts = np.array([[1.5, 0.0, 0.0], [0.0, -1.2, 0.3], [1.0, 0.0, 0.0], [-1.0, 0.2, 0.1],
                       [0.3, 1.0, -1.0], [-1.0, -1.0, -1.0], [0.9999, 0.0, 0.0],
                       [0.0, 0.0, 0.0]], np.float32)
    n_iter = np.full(len(starts), 400, np.int32)
    f64, s64, e64 = oracle_mod.topo_f64(starts, n_iter, 0.01, dims, x, q)
    for prec in ("auto", "f64"):
        r = ops.topo(starts, n_iter, 0.01, dims, x, q, prec=prec, ctx=gpu_ctx)
        np.testing.assert_array_equal(r.steps, s64)
        np.testing.assert_array_equal(r.exited, e64)
        np.testing.assert_allclose(r.feats, f64, rtol=1e-5, atol=1e-9)
    assert r.steps[0] == 1 and r.exited[0] == 1


def test_start_on_a_charge(gpu_ctx, oracle_mod):
    """A start exactly on a charge: the reference's field there is 0 * inf =
    NaN (math_module.c:296-333), so its walk never passes the box test and
    runs all n_iter steps with NaN features; the GPU does the same, the
    histogram drops the NaN row, and the other streamlines are unaffected."""
    from pycpet_amd import ops

    x = np.array([[0.2, 0.1, -0.3], [-1.7, 0.4, 2.2]], np.float32)
    q = np.array([0.3, -0.6], np.float32)
    dims = np.full(3, 1.0, np.float32)
    starts = np.array([[0.2, 0.1, -0.3], [0.0, 0.0, 0.0]], np.float32)
    n_iter = np.array([50, 50], np.int32)
    f64, s64, e64 = oracle_mod.topo_f64(starts, n_iter, 0.01, dims, x, q)
    assert np.isnan(f64[0]).all() and s64[0] == 50 and e64[0] == 0
    edges = (np.linspace(0.0, 1.0, 5), np.linspace(0.0, 1.0, 5))
    r = ops.topo(starts, n_iter, 0.01, dims, x, q, edges=edges, ctx=gpu_ctx)
    assert np.isnan(r.feats[0]).all()
    np.testing.assert_array_equal(r.steps, s64)
    np.testing.assert_array_equal(r.exited, e64)
    np.testing.assert_allclose(r.feats[1], f64[1], rtol=1e-5)
    assert int(r.hist.sum()) == 1


def _random_configurations():
    """Twelve seeded frames across the shapes the kernels branch on: charge
    counts from 1 to 2,100 (the four-charge slice split, 64-charge tiles,
    the padding, one-wave group streaming up to 64 records), anisotropic
    boxes of half-widths 0.4-1.6 A, h of 0.002-0.02 A, step limits 0-600,
    charges 0.3-3 A outside the box; 200 starts each."""
    from pycpet_amd import workloads as W

    rng = np.random.default_rng(2026)
    frames = []
    for k, N in enumerate((1, 2, 5, 17, 63, 64, 65, 127, 255, 401, 1000, 2100)):
        dims = rng.uniform(0.4, 1.6, 3).astype(np.float32)
        h = float(rng.choice([0.002, 0.005, 0.01, 0.02]))
        excl = float(rng.uniform(0.3, 3.0))
        box = float(dims.max())
        x, q = W.ball_exclusion(N, box + excl + 12.0, box, excl, seed=500 + k)
        if N == 1:  # the mean-free charge of one would be 0
            q = np.array([0.7 if k % 2 else -0.7], np.float32)
        starts = rng.uniform(-dims, dims, (200, 3)).astype(np.float32)
        n_iter = rng.integers(0, 601, 200).astype(np.int32)
        frames.append({"x": x, "q": q, "starts": starts, "n_iter": n_iter, "h": h, "dims": dims})
    return frames


def test_random_configurations_vs_oracle(gpu_ctx, oracle_mod):
    """Each of twelve random frames alone at the default launch shape, and
    all twelve as one batch at the default shape and with one-wave
    workgroups forced (every frame, 2,100 charges included), against the
    FP64 oracle under the matching rule above; the batched results equal
    the single-frame ones bit for bit; FP64 sums (prec="f64") give the
    oracle's step counts and stop kinds exactly and its features within
    1e-8 relative, 1e-10 absolute (a straight line's curvature is FP64
    rounding noise on both sides, up to about 1e-16 |x| / h^2 = 2.5e-11 at
    h = 0.002)."""
    from pycpet_amd import batch, ops

    frames = _random_configurations()
    single = []
    for i, f in enumerate(frames):
        label = f"random frame {i}: N={len(f['q'])} h={f['h']} dims={np.round(f['dims'], 2)}"
        print(label, flush=True)
        f64, s64, e64 = oracle_mod.topo_f64(f["starts"], f["n_iter"], f["h"], f["dims"],
                                            f["x"], f["q"])
        r = ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"], ctx=gpu_ctx)
        compare(r, f64, s64, e64, label)
        single.append(r)
        r64 = ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"], prec="f64",
                       ctx=gpu_ctx)
        np.testing.assert_array_equal(r64.steps, s64)
        np.testing.assert_array_equal(r64.exited, e64)
        np.testing.assert_allclose(r64.feats, f64, rtol=1e-8, atol=1e-10)
    try:
        for S in (0, 1):
            gpu_ctx.set_tuning(3, S, 0)
            print(f"random frames batched, waves per workgroup {S or 'default'}", flush=True)
            outs = batch.batch_topo(frames, ctx=gpu_ctx)
            for a, b in zip(single, outs):
                np.testing.assert_array_equal(a.feats, b.feats)
                np.testing.assert_array_equal(a.steps, b.steps)
                np.testing.assert_array_equal(a.exited, b.exited)
    finally:
        gpu_ctx.set_tuning(3, 0, 0)


def test_full_size_100k(gpu_ctx, oracle_mod):
    """The 5k-charge / 100k-streamline workload at full size: invariants over
    all streamlines plus a seeded 48-streamline sample against the oracle."""
    from pycpet_amd import ops, workloads as W

    c = W.topo_5k()
    print(f"case {c.name} N={c.N} L={c.L}", flush=True)
    r = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, ctx=gpu_ctx)
    assert np.all(r.steps >= 1) and np.all(r.steps <= c.n_iter)
    assert np.all(r.steps[r.exited == 0] == c.n_iter[r.exited == 0])
    h = float(np.float32(c.h))
    assert np.all(r.feats[:, 0] <= r.steps * h * (1 + 1e-9))  # chord <= arc
    assert np.all(np.isfinite(r.feats))
    idx = np.random.default_rng(0).choice(c.L, 48, replace=False)
    f64, s64, e64 = oracle_mod.topo_f64(c.starts[idx], c.n_iter[idx], c.h, c.dims, c.x, c.q)
    sub = ops.TopoResult(r.feats[idx], r.steps[idx], r.exited[idx])
    m = (sub.steps == s64) & (sub.exited == e64)
    assert m.mean() >= 0.95  # 48 samples: allow one boundary tie
    dd = np.abs(sub.feats[m, 0] - f64[m, 0]) / f64[m, 0]
    dk = np.abs(sub.feats[m, 1] - f64[m, 1]) / np.maximum(np.abs(f64[m, 1]), 1e-6)
    assert dd.max() <= FEAT_TOL and dk.max() <= FEAT_TOL


@pytest.mark.parametrize("N,L", [(300, 512), (800, 384)])
def test_rk4_vs_oracle(gpu_ctx, oracle_mod, N, L):
    """integrator="rk4" against the FP64 RK4 oracle, with the same matching
    rule as the Euler kernel."""
    from pycpet_amd import ops, workloads as W

    c = W.small_topo(N=N, L=L)
    r = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, integrator="rk4", ctx=gpu_ctx)
    f64, s64, e64 = oracle_mod.topo_rk4_f64(c.starts, c.n_iter, c.h, c.dims, c.x, c.q)
    compare(r, f64, s64, e64, f"rk4_{N}x{L}")


def test_rk4_random_configurations_vs_oracle(gpu_ctx, oracle_mod):
    """integrator="rk4" on the twelve random frames (1 to 2,100 charges,
    anisotropic boxes, h 0.002-0.02 A) against the FP64 RK4 oracle, with the
    same matching rule."""
    from pycpet_amd import ops

    for i, f in enumerate(_random_configurations()):
        label = f"rk4 random frame {i}: N={len(f['q'])} h={f['h']}"
        print(label, flush=True)
        r = ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"],
                     integrator="rk4", ctx=gpu_ctx)
        f64, s64, e64 = oracle_mod.topo_rk4_f64(f["starts"], f["n_iter"], f["h"], f["dims"],
                                                f["x"], f["q"])
        compare(r, f64, s64, e64, label)


def test_rk4_radial_lines_equal_euler(gpu_ctx):
    """A single charge: straight radial lines, so RK4 and Euler take the same
    steps (the RK4 stage bookkeeping must not change the step count)."""
    from pycpet_amd import ops

    x = np.array([[0.0, 0.0, 0.0]], np.float32)
    q = np.array([1.0], np.float32)
    rng = np.random.default_rng(4)
    d = rng.standard_normal((300, 3))
    starts = (0.3 * d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    n_iter = rng.integers(0, 300, 300).astype(np.int32)
    dims = np.full(3, 1.0, np.float32)
    e = ops.topo(starts, n_iter, 0.01, dims, x, q, ctx=gpu_ctx)
    k = ops.topo(starts, n_iter, 0.01, dims, x, q, integrator="rk4", ctx=gpu_ctx)
    np.testing.assert_array_equal(k.steps, e.steps)
    np.testing.assert_array_equal(k.exited, e.exited)
    np.testing.assert_allclose(k.feats[:, 0], e.feats[:, 0], rtol=1e-9, atol=1e-12)
    with pytest.raises(ValueError):
        ops.topo(starts, n_iter, 0.01, dims, x, q, integrator="rk2", ctx=gpu_ctx)


def test_f64_mode_vs_oracle(gpu_ctx, oracle_mod):
    """prec="f64": FP64 field sums in the walk.  Against the FP64 oracle the
    features agree far inside the FP32 path's errors (Euler and RK4)."""
    from pycpet_amd import ops, workloads as W

    c = W.small_topo(N=500, L=256)
    for integ, ref in (("euler", oracle_mod.topo_f64), ("rk4", oracle_mod.topo_rk4_f64)):
        f64, s64, e64 = ref(c.starts, c.n_iter, c.h, c.dims, c.x, c.q)
        r = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, integrator=integ, prec="f64",
                     ctx=gpu_ctx)
        m = compare(r, f64, s64, e64, f"f64_{integ}")
        assert m.all()  # FP64 sums: every step count and stop kind equal
        dd = np.abs(r.feats[:, 0] - f64[:, 0]) / np.maximum(np.abs(f64[:, 0]), 1e-12)
        assert dd.max() <= 1e-10, dd.max()
    with pytest.raises(ValueError):
        ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, prec="kahan", ctx=gpu_ctx)


def _far_frame(N, L, cap, seed, r_in=6.0, r_out=30.0):
    """Charges in a shell r_in..r_out A around a +-1.5 A box, h = 1e-3:
    with r_in >= 6 every charge passes the far-field guard
    (|dims + h| + |c| <= 3 dist(c, box))."""
    from pycpet_amd import workloads as W

    rng = np.random.default_rng(seed)
    u = rng.standard_normal((N, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x = (u * rng.uniform(r_in, r_out, (N, 1))).astype(np.float32)
    q = rng.uniform(-1, 1, N).astype(np.float32)
    dims = np.array([1.5, 1.5, 1.5], np.float32)
    starts, n_iter, _ = W.topo_starts(L, dims, 0.001, cap, seed=seed + 42, start_seed=seed)
    return dict(x=x, q=q, starts=starts, n_iter=n_iter, h=0.001, dims=dims)


def test_far_field_form_opt_in(gpu_ctx, oracle_mod):
    """The expanded far-field pair term (tuning kind 5, a = 2), single frame
    and batched: on frames whose charges all pass its guard it meets the same
    oracle bar as the direct form (and is a different sum: not bit-identical
    to it); a frame with charges next to the box keeps the direct form
    (bit-identical)."""
    from pycpet_amd import batch, ops

    fa = _far_frame(600, 384, 1500, 21)
    fb = _far_frame(500, 256, 300, 22, r_in=3.0)  # charges from 3 A: some fail the guard
    f64, s64, e64 = oracle_mod.topo_f64(fa["starts"], fa["n_iter"], fa["h"], fa["dims"], fa["x"],
                                        fa["q"])
    run = lambda f: ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"],
                             ctx=gpu_ctx)
    direct_a, direct_b = run(fa), run(fb)
    direct_batch = batch.batch_topo([fa, fb], ctx=gpu_ctx)
    gpu_ctx.set_tuning(5, 2, 0)
    try:
        far_a, far_b = run(fa), run(fb)
        far_batch = batch.batch_topo([fa, fb], ctx=gpu_ctx)
    finally:
        gpu_ctx.set_tuning(5, 0, 0)
    print("case far-field shells 600x384 (r >= 6) and 500x256 (r >= 3), h=1e-3", flush=True)
    for r in (far_a, far_batch[0]):
        compare(r, f64, s64, e64, "far")
        assert not np.array_equal(r.feats, direct_a.feats)
    compare(direct_batch[0], f64, s64, e64, "direct batch")
    for r, d in ((far_b, direct_b), (far_batch[1], direct_batch[1])):
        np.testing.assert_array_equal(r.feats, d.feats)
        np.testing.assert_array_equal(r.steps, d.steps)
        np.testing.assert_array_equal(r.exited, d.exited)
