"""The streamline, batching and serving configurations at their full sizes.

  c1  5,000 charges, 100,000 starts, h = 1e-3 A, n_iter capped at 2,000,
      Euler (the reference's walk) and RK4 (the configuration's),
      with the 20 x 20 (distance, curvature) histogram whose range comes
      from the FP64 features' min / max (the range rule of make_histograms,
      CPET/utils/calculator.py:702-707, and get_hist_grid,
      CPET/source/scripts/residue_breakdown_analysis.py:28-37).
  c3  MD-like frames of the batching set (3k-40k charges, 21^3-81^3 grids,
      20,000-start streamline frames): every frame of a batch equals the
      frame run alone, bitwise, plus seeded samples against the FP64 oracle.
  c4  serving requests drawn from the serving ranges (1k-60k charges,
      21^3-101^3 grids, 10k-200k starts): each result is the same alone and
      in the mix.
Tolerances are those of the single-frame tests: fields 1e-5 pointwise
(floor 1e-3 RMS), streamline features 1e-4 on streamlines whose step count
and stop kind match, histogram counts exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5
FEAT_TOL = 1e-4


def _field_err(a, b):
    a = np.asarray(a, float).reshape(len(b), -1)
    b = np.asarray(b, float).reshape(len(b), -1)
    nb = np.linalg.norm(b, axis=1)
    d = np.linalg.norm(a - b, axis=1)
    return float((d / np.maximum(nb, 1e-3 * np.sqrt(np.mean(nb ** 2)))).max())


def _bins(v, e):
    """numpy.histogramdd's bin of each value (-1 outside), as the kernel
    bins float32-rounded features."""
    i = np.searchsorted(e, v, side="right")
    i[v == e[-1]] -= 1
    out = i - 1
    out[(i < 1) | (i > e.size - 1)] = -1
    return out


def test_c1_full_size_with_histogram(gpu_ctx, oracle_mod):
    from pycpet_amd import ops, workloads as W

    c = W.topo_5k()
    print(f"case c1 topo_5k N={c.N} L={c.L} with 20x20 histogram", flush=True)
    # FP64 mode over all streamlines: the range source; pinned on a sample
    r64 = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, prec="f64", ctx=gpu_ctx)
    idx = np.random.default_rng(5).choice(c.L, 24, replace=False)
    o64, os64, oe64 = oracle_mod.topo_f64(c.starts[idx], c.n_iter[idx], c.h, c.dims, c.x, c.q)
    np.testing.assert_array_equal(r64.steps[idx], os64)
    np.testing.assert_array_equal(r64.exited[idx], oe64)
    np.testing.assert_allclose(r64.feats[idx], o64, rtol=1e-9, atol=1e-12)
    f64 = r64.feats
    ed = np.linspace(f64[:, 0].min(), f64[:, 0].max(), 21)
    ec = np.linspace(f64[:, 1].min(), f64[:, 1].max(), 21)

    r = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, edges=(ed, ec), ctx=gpu_ctx)
    fr = r.feats.astype(np.float32).astype(np.float64)
    Hn, _, _ = np.histogram2d(fr[:, 0], fr[:, 1], bins=[ed, ec])
    np.testing.assert_array_equal(r.hist, Hn.astype(np.uint32))
    assert int(r.hist.sum()) >= 0.999 * c.L  # the float32 rounding of an extreme may drop it

    # the FP32 path against the FP64 mode over all 100,000 streamlines
    m = (r.steps == r64.steps) & (r.exited == r64.exited)
    dd = np.abs(r.feats[m, 0] - f64[m, 0]) / np.maximum(np.abs(f64[m, 0]), 1e-12)
    dk = np.abs(r.feats[m, 1] - f64[m, 1]) / np.maximum(np.abs(f64[m, 1]), 1e-6)
    print(f"c1: match {m.mean():.5f} dist {dd.max():.2e} curv {dk.max():.2e}", flush=True)
    assert m.mean() >= 0.99 and dd.max() <= FEAT_TOL and dk.max() <= FEAT_TOL

    # counts agree with the FP64 run's wherever the binned features agree
    f64r = f64.astype(np.float32).astype(np.float64)
    H64, _, _ = np.histogram2d(f64r[:, 0], f64r[:, 1], bins=[ed, ec])
    same = m & (_bins(fr[:, 0], ed) == _bins(f64r[:, 0], ed)) & \
        (_bins(fr[:, 1], ec) == _bins(f64r[:, 1], ec))
    Hs32, _, _ = np.histogram2d(fr[same, 0], fr[same, 1], bins=[ed, ec])
    Hs64, _, _ = np.histogram2d(f64r[same, 0], f64r[same, 1], bins=[ed, ec])
    np.testing.assert_array_equal(Hs32, Hs64)
    moved = int((~same).sum())
    print(f"c1: {moved} streamlines change bin; |H32 - H64|_1 = {np.abs(Hn - H64).sum():.0f}",
          flush=True)
    assert same.mean() >= 0.99
    assert np.abs(Hn - H64).sum() <= 2 * moved


def test_c1_rk4_full_size_with_histogram(gpu_ctx, oracle_mod):
    """c1 as the configuration names it, with RK4 steps: the FP64-sum RK4
    walk pinned to the RK4 oracle on a seeded sample, its features giving
    the histogram range, and the FP32 RK4 walk with its histogram against it
    over all 100,000 streamlines."""
    from pycpet_amd import ops, workloads as W

    c = W.topo_5k()
    print(f"case c1 rk4 topo_5k N={c.N} L={c.L} with 20x20 histogram", flush=True)
    r64 = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, prec="f64", integrator="rk4",
                   ctx=gpu_ctx)
    idx = np.random.default_rng(6).choice(c.L, 24, replace=False)
    o64, os64, oe64 = oracle_mod.topo_rk4_f64(c.starts[idx], c.n_iter[idx], c.h, c.dims, c.x,
                                              c.q)
    np.testing.assert_array_equal(r64.steps[idx], os64)
    np.testing.assert_array_equal(r64.exited[idx], oe64)
    np.testing.assert_allclose(r64.feats[idx], o64, rtol=1e-9, atol=1e-12)
    f64 = r64.feats
    ed = np.linspace(f64[:, 0].min(), f64[:, 0].max(), 21)
    ec = np.linspace(f64[:, 1].min(), f64[:, 1].max(), 21)
    r = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, edges=(ed, ec), integrator="rk4",
                 ctx=gpu_ctx)
    fr = r.feats.astype(np.float32).astype(np.float64)
    Hn, _, _ = np.histogram2d(fr[:, 0], fr[:, 1], bins=[ed, ec])
    np.testing.assert_array_equal(r.hist, Hn.astype(np.uint32))
    m = (r.steps == r64.steps) & (r.exited == r64.exited)
    dd = np.abs(r.feats[m, 0] - f64[m, 0]) / np.maximum(np.abs(f64[m, 0]), 1e-12)
    dk = np.abs(r.feats[m, 1] - f64[m, 1]) / np.maximum(np.abs(f64[m, 1]), 1e-6)
    print(f"c1 rk4: match {m.mean():.5f} dist {dd.max():.2e} curv {dk.max():.2e}", flush=True)
    assert m.mean() >= 0.99 and dd.max() <= FEAT_TOL and dk.max() <= FEAT_TOL
    f64r = f64.astype(np.float32).astype(np.float64)
    same = m & (_bins(fr[:, 0], ed) == _bins(f64r[:, 0], ed)) & \
        (_bins(fr[:, 1], ec) == _bins(f64r[:, 1], ec))
    Hs32, _, _ = np.histogram2d(fr[same, 0], fr[same, 1], bins=[ed, ec])
    Hs64, _, _ = np.histogram2d(f64r[same, 0], f64r[same, 1], bins=[ed, ec])
    np.testing.assert_array_equal(Hs32, Hs64)
    assert same.mean() >= 0.99


def test_c3_md_frames_at_config_size(gpu_ctx, oracle_mod):
    from pycpet_amd import batch, ops, workloads as W

    base = W.md_base()
    spec = [("volume", 40000, 81), ("volume", 3000, 21), ("volume", 21500, 61),
            ("topo", 40000, 0), ("topo", 3000, 0), ("topo", 21500, 0)]
    frames = [W.md_frame(base, i, k, n, edge=e or None) for i, (k, n, e) in enumerate(spec)]
    vol = [f for f in frames if f["kind"] == "volume"]
    top = [f for f in frames if f["kind"] == "topo"]
    print("case c3: volume frames " + ", ".join(f"{f['q'].size}x{len(f['axes'][0])}^3" for f in vol)
          + "; topo frames " + ", ".join(f"{f['q'].size}x{f['starts'].shape[0]}" for f in top),
          flush=True)

    # grids: the batch equals each frame alone (same points per lane, no
    # charge split in either), bitwise
    gpu_ctx.set_tuning(4, 2, 0)
    gpu_ctx.set_tuning(0, 2, 1)
    try:
        outs = batch.batch_volume(vol, out_f64=True, ctx=gpu_ctx)
        for f, (E, _) in zip(vol, outs):
            E1, _ = ops.field_grid(*f["axes"], f["x"], f["q"], out_f64=True, ctx=gpu_ctx)
            np.testing.assert_array_equal(E, E1)
    finally:
        gpu_ctx.set_tuning(4, 0, 0)
        gpu_ctx.set_tuning(0, 0, 0)
    f = vol[0]  # 40k charges on the 81^3 grid: a seeded sample vs FP64
    pts = np.stack(np.meshgrid(*f["axes"], indexing="ij"), -1).reshape(-1, 3)
    idx = np.random.default_rng(1).choice(pts.shape[0], 400, replace=False)
    E64, _ = oracle_mod.field_f64(pts[idx], f["x"], f["q"])
    assert _field_err(outs[0][0][idx], E64) <= TOL

    # streamlines: the persistent batch launch equals each frame alone
    outs = batch.batch_topo(top, ctx=gpu_ctx)
    for f, r in zip(top, outs):
        r1 = ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"], ctx=gpu_ctx)
        np.testing.assert_array_equal(r.feats, r1.feats)
        np.testing.assert_array_equal(r.steps, r1.steps)
        np.testing.assert_array_equal(r.exited, r1.exited)
    f, r = top[0], outs[0]  # 40k charges, 20k starts: a seeded sample vs FP64
    idx = np.random.default_rng(2).choice(f["starts"].shape[0], 12, replace=False)
    o64, s64, e64 = oracle_mod.topo_f64(f["starts"][idx], f["n_iter"][idx], f["h"], f["dims"],
                                        f["x"], f["q"])
    m = (r.steps[idx] == s64) & (r.exited[idx] == e64)
    assert m.mean() >= 11 / 12  # a box-exit tie may move one step
    dd = np.abs(r.feats[idx][m, 0] - o64[m, 0]) / np.maximum(o64[m, 0], 1e-12)
    dk = np.abs(r.feats[idx][m, 1] - o64[m, 1]) / np.maximum(np.abs(o64[m, 1]), 1e-6)
    assert dd.max() <= FEAT_TOL and dk.max() <= FEAT_TOL


def test_c4_serving_requests_at_config_size(gpu_ctx):
    from pycpet_amd import ops, serve

    reqs = serve.poisson_requests(rate=200.0, duration=1.0, seed=13, max_requests=20)
    kinds = {r.kind for r in reqs}
    sizes = [(r.kind, r.data["q"].size,
              len(r.data["axes"][0]) if r.kind != "topo" else r.data["starts"].shape[0])
             for r in reqs]
    print(f"case c4: {len(reqs)} requests {sizes}", flush=True)
    assert kinds == {"volume", "volume_ESP", "topo"}
    assert all(1000 <= n <= 60000 for _, n, _ in sizes)
    assert max(n for _, n, _ in sizes) > 50000
    assert max(e for k, _, e in sizes if k != "topo") >= 85
    assert max(e for k, _, e in sizes if k == "topo") >= 100000
    srv = serve.Server(n_lanes=4)
    stats = srv.run(reqs)
    assert stats.completed == len(reqs) and stats.rejected == 0
    for r in reqs:
        d = r.data
        if r.kind == "topo":
            alone = ops.topo(d["starts"], d["n_iter"], d["h"], d["dims"], d["x"], d["q"],
                             ctx=gpu_ctx).feats
            np.testing.assert_array_equal(r.result, alone)
        else:
            E, phi = ops.field_grid(*d["axes"], d["x"], d["q"], want_E=r.kind == "volume",
                                    want_phi=r.kind == "volume_ESP", ctx=gpu_ctx)
            assert _field_err(r.result, E if r.kind == "volume" else phi) <= TOL
    s = stats.summary()
    print(f"c4: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms batches {s['batches']}",
          flush=True)
