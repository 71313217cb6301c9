"""Many frames per launch: every frame's result equals the same frame run
alone (bitwise — the per-point / per-streamline arithmetic is independent of
which workgroup or launch runs it when the charge tiling is the same), and
matches the FP64 oracle with the single-frame tolerances."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_frames(n=6, seed=0):
    from pycpet_amd import workloads as W

    rng = np.random.default_rng(seed)
    frames = []
    for i in range(n):
        N = int(rng.integers(40, 900))
        x, q = W.ball_exclusion(N, 12.0, 0.6, 1.5, seed=100 + i)
        if i % 2 == 0:
            e = int(rng.choice([5, 9, 13, 21]))
            ax = np.linspace(-0.5, 0.5, e).astype(np.float32)
            ay = np.linspace(-0.3, 0.4, e + 2).astype(np.float32)
            frames.append({"kind": "volume", "x": x, "q": q, "axes": (ax, ay, ax)})
        else:
            d = np.array([0.6, 0.5, 0.55], np.float32)
            st, it, _ = W.topo_starts(int(rng.integers(30, 400)), d, 0.01, 300, seed=7 + i,
                                      start_seed=8 + i)
            frames.append({"kind": "topo", "x": x, "q": q, "starts": st, "n_iter": it,
                           "h": 0.01, "dims": d})
    return frames


def test_batch_volume_equals_single_frames(gpu_ctx, oracle_mod):
    from pycpet_amd import batch, ops

    frames = [f for f in _small_frames(12) if f["kind"] == "volume"]
    for K in (4, 2):
        gpu_ctx.set_tuning(4, K, 0)
        gpu_ctx.set_tuning(0, K, 1)  # single-frame: same K, no charge split
        outs = batch.batch_volume(frames, want_phi=True, out_f64=True, ctx=gpu_ctx)
        for f, (E, phi) in zip(frames, outs):
            E1, p1 = ops.field_grid(*f["axes"], f["x"], f["q"], want_phi=True, out_f64=True,
                                    ctx=gpu_ctx)
            np.testing.assert_array_equal(E, E1)
            np.testing.assert_array_equal(phi, p1)
    gpu_ctx.set_tuning(4, 0, 0)
    gpu_ctx.set_tuning(0, 0, 0)
    f = frames[0]
    pts = np.stack(np.meshgrid(*f["axes"], indexing="ij"), -1).reshape(-1, 3)
    E64, p64 = oracle_mod.field_f64(pts, f["x"], f["q"], want_phi=True)
    E, phi = batch.batch_volume(frames, want_phi=True, ctx=gpu_ctx)[0]
    assert gpu_ctx.refined_points() > 0  # this frame has strongly cancelling points
    nE = np.linalg.norm(E64, axis=1)
    err = np.linalg.norm(E - E64, axis=1) / np.maximum(nE, 1e-3 * np.sqrt(np.mean(nE ** 2)))
    assert err.max() <= 1e-5


def test_batch_volume_floor_near_grid(gpu_ctx, oracle_mod):
    """E alone (the floored volume field): frames with a charge on or next to
    a grid point keep the r^2 >= 1e-6 floor, the others skip it; each frame
    equals the frame run alone, and the near frames match the oracle."""
    from pycpet_amd import batch, ops

    frames = [f for f in _small_frames(12, seed=3) if f["kind"] == "volume"]
    for i, f in enumerate(frames[:3]):  # frames 0-2: a charge near a grid point
        g = np.array([f["axes"][0][1], f["axes"][1][2], f["axes"][2][0]], np.float32)
        near = g + np.float32((0.0, 4e-4, 9e-4)[i]) * np.array([0, 0, 1], np.float32)
        f["x"] = np.concatenate([f["x"], near[None]]).astype(np.float32)
        f["q"] = np.concatenate([f["q"], np.float32([0.4])]).astype(np.float32)
    gpu_ctx.set_tuning(4, 2, 0)
    gpu_ctx.set_tuning(0, 2, 1)  # single-frame: same K, no charge split
    outs = batch.batch_volume(frames, out_f64=True, ctx=gpu_ctx)
    for f, (E, _) in zip(frames, outs):
        E1, _ = ops.field_grid(*f["axes"], f["x"], f["q"], out_f64=True, ctx=gpu_ctx)
        np.testing.assert_array_equal(E, E1)
    gpu_ctx.set_tuning(4, 0, 0)
    gpu_ctx.set_tuning(0, 0, 0)
    for f, (E, _) in list(zip(frames, outs))[:3]:
        pts = np.stack(np.meshgrid(*f["axes"], indexing="ij"), -1).reshape(-1, 3)
        E64, _ = oracle_mod.field_f64(pts, f["x"], f["q"])
        nE = np.linalg.norm(E64, axis=1)
        err = np.linalg.norm(E - E64, axis=1) / np.maximum(nE, 1e-3 * np.sqrt(np.mean(nE ** 2)))
        assert np.all(np.isfinite(E)) and err.max() <= 1e-5


def test_batch_topo_equals_single_frames(gpu_ctx, oracle_mod):
    from pycpet_amd import batch, ops

    frames = [f for f in _small_frames(12, seed=1) if f["kind"] == "topo"]
    ed = np.linspace(0.0, 1.5, 11)
    ec = np.linspace(0.0, 30.0, 13)
    for S, G in ((8, 0), (4, 3), (16, 1)):
        gpu_ctx.set_tuning(3, S, G)
        gpu_ctx.set_tuning(2, S, 0)
        outs = batch.batch_topo(frames, edges=(ed, ec), ctx=gpu_ctx)
        for f, r in zip(frames, outs):
            r1 = ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"],
                          ctx=gpu_ctx)
            np.testing.assert_array_equal(r.feats, r1.feats)
            np.testing.assert_array_equal(r.steps, r1.steps)
            np.testing.assert_array_equal(r.exited, r1.exited)
            fr = r.feats.astype(np.float32).astype(np.float64)
            Hn, _, _ = np.histogram2d(fr[:, 0], fr[:, 1], bins=[ed, ec])
            np.testing.assert_array_equal(r.hist, Hn.astype(np.uint32))
    gpu_ctx.set_tuning(3, 0, 0)
    gpu_ctx.set_tuning(2, 0, 0)
    f = frames[1]
    f64, s64, e64 = oracle_mod.topo_f64(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"],
                                        f["q"])
    r = batch.batch_topo(frames, ctx=gpu_ctx)[1]
    m = (r.steps == s64) & (r.exited == e64)
    assert m.mean() >= 0.99
    np.testing.assert_allclose(r.feats[m], f64[m], rtol=1e-4, atol=1e-9)


def test_batch_device_tensors_match_host(gpu_ctx):
    import torch

    from pycpet_amd import batch

    frames = _small_frames(8, seed=2)
    vol = [f for f in frames if f["kind"] == "volume"]
    top = [f for f in frames if f["kind"] == "topo"]
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    vol_t = [{"x": T(f["x"]), "q": T(f["q"]), "axes": tuple(T(a) for a in f["axes"])} for f in vol]
    top_t = [{"x": T(f["x"]), "q": T(f["q"]), "starts": T(f["starts"]), "n_iter": T(f["n_iter"]),
              "h": f["h"], "dims": f["dims"]} for f in top]
    a = batch.batch_volume(vol, ctx=gpu_ctx)
    b = batch.batch_volume(vol_t, ctx=gpu_ctx)
    ra = batch.batch_topo(top, ctx=gpu_ctx)
    rb = batch.batch_topo(top_t, ctx=gpu_ctx)
    torch.cuda.synchronize()
    for (E, _), (Et, _) in zip(a, b):
        np.testing.assert_array_equal(E, Et.cpu().numpy())
    for r, rt in zip(ra, rb):
        np.testing.assert_array_equal(r.feats, rt.feats.cpu().numpy())
        np.testing.assert_array_equal(r.steps, rt.steps.cpu().numpy())


def test_one_wave_walk_equals_sixteen_waves(gpu_ctx):
    """One-wave workgroups walk the 16 waves' charge slices in order, so
    streamlines come out the same bit for bit as with 16 waves: single
    frames of 1 to 1,000 charges (group- and tile-ragged counts), charges
    and dipoles, FP64 sums and RK4; and the batch default for small frames
    (one wave) against the single-frame default (16 waves)."""
    from pycpet_amd import batch, ops
    from pycpet_amd import workloads as W

    d = np.array([0.6, 0.5, 0.55], np.float32)
    st, it, _ = W.topo_starts(700, d, 0.01, 300, seed=3, start_seed=4)
    try:
        for N in (1, 3, 16, 61, 64, 65, 128, 131, 300, 1000):
            x, q = W.ball_exclusion(N, 12.0, 0.6, 1.5, seed=200 + N)
            if N == 1:
                q = np.array([0.5], np.float32)  # the mean-free charge of one would be 0
            mu = np.random.default_rng(N).standard_normal((N, 3)).astype(np.float32)
            print(f"one-wave walk N={N}", flush=True)
            runs = {}
            for S in (16, 1):
                gpu_ctx.set_tuning(2, S, 0)
                runs[S] = [ops.topo(st, it, 0.01, d, x, q, ctx=gpu_ctx),
                           ops.topo(st, it, 0.01, d, x, q, prec="f64", ctx=gpu_ctx),
                           ops.topo(st, it, 0.01, d, x, q, integrator="rk4", ctx=gpu_ctx),
                           ops.topo_dipole(st, it, 0.01, d, x, mu, ctx=gpu_ctx)]
            for a, b in zip(runs[16], runs[1]):
                np.testing.assert_array_equal(a.feats, b.feats)
                np.testing.assert_array_equal(a.steps, b.steps)
                np.testing.assert_array_equal(a.exited, b.exited)
    finally:
        gpu_ctx.set_tuning(2, 0, 0)
    frames = []
    for i, N in enumerate((5, 12, 40, 77, 128, 9)):
        x, q = W.ball_exclusion(N, 12.0, 0.6, 1.5, seed=300 + i)
        frames.append({"x": x, "q": q, "starts": st, "n_iter": it, "h": 0.01, "dims": d})
    outs = batch.batch_topo(frames, ctx=gpu_ctx)  # every frame <= 128 charges: one wave
    for f, r in zip(frames, outs):
        r1 = ops.topo(f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], f["q"], ctx=gpu_ctx)
        np.testing.assert_array_equal(r.feats, r1.feats)
        np.testing.assert_array_equal(r.steps, r1.steps)


def test_autotuned_batch_shape(gpu_ctx):
    """tune_batch_topo times the candidates on real frames and leaves the
    fastest applied; the tuned launch gives the default launch's results
    (bit for bit with 1 or 16 waves, to FP32 rounding otherwise), and a
    cached choice is applied without timing again."""
    from pycpet_amd import autotune as AT
    from pycpet_amd import batch
    from pycpet_amd import workloads as W

    d = np.array([0.6, 0.5, 0.55], np.float32)
    st, it, _ = W.topo_starts(3000, d, 0.01, 300, seed=5, start_seed=6)
    frames = []
    for i, N in enumerate((9, 20, 33, 48)):
        x, q = W.ball_exclusion(N, 12.0, 0.6, 1.5, seed=400 + i)
        frames.append({"x": x, "q": q, "starts": st, "n_iter": it, "h": 0.01, "dims": d})
    base = batch.batch_topo(frames, ctx=gpu_ctx)
    cache = {}
    try:
        S, G = AT.tune_batch_topo(gpu_ctx, frames, cache=cache)
        print(f"autotuned S={S} G={G}", flush=True)
        assert (S, G) == (1, 0) or S in (4, 8, 16)
        tuned = batch.batch_topo(frames, ctx=gpu_ctx)
        for a, b in zip(base, tuned):
            if S in (1, 16):
                np.testing.assert_array_equal(a.feats, b.feats)
            else:
                same = a.steps == b.steps
                assert same.mean() >= 0.99
                np.testing.assert_allclose(a.feats[same], b.feats[same], rtol=1e-4, atol=1e-9)
        assert list(cache.values()) == [[S, G]]
        gpu_ctx.set_tuning(3, 0, 0)
        assert AT.tune_batch_topo(gpu_ctx, frames, cache=cache) == (S, G)
    finally:
        gpu_ctx.set_tuning(3, 0, 0)
