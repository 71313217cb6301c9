"""Dipole streamlines on the GPU (cpet_topo_dipole, thread_operation_dipole)
against the FP64 oracle: step counts and stop kind identical on >= 99 % of
starts, distance and curvature within 1e-4 there (the charge-streamline
bar), in the reference's and the physical field form; the on-axis closed
form exactly; and the reference C's own outputs through the drop-in."""
import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _case(seed=2, N=300, L=512):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((N, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x = (d * rng.uniform(2.0, 9.0, (N, 1))).astype(np.float32)
    mu = rng.normal(0.0, 0.3, (N, 3)).astype(np.float32)
    dims = np.array([0.5, 0.5, 0.5], np.float32)
    starts = rng.uniform(-0.45, 0.45, (L, 3)).astype(np.float32)
    n_iter = rng.integers(1, 300, L).astype(np.int32)
    return x, mu, dims, starts, n_iter


@pytest.mark.parametrize("physical", [False, True])
def test_dipole_vs_oracle(gpu_ctx, oracle_mod, physical):
    from pycpet_amd import ops

    x, mu, dims, starts, n_iter = _case()
    r = ops.topo_dipole(starts, n_iter, 0.01, dims, x, mu, physical=physical, ctx=gpu_ctx)
    f64, s64, e64 = oracle_mod.topo_dipole_f64(starts, n_iter, 0.01, dims, x, mu,
                                               physical=physical)
    match = (r.steps == s64) & (r.exited == e64)
    rel = np.abs(r.feats - f64) / np.maximum(np.abs(f64), 1e-6)
    print(f"physical={physical} match {match.mean():.5f} dist max {rel[match, 0].max():.3e} "
          f"curv max {rel[match, 1].max():.3e}", flush=True)
    assert match.mean() >= 0.99
    assert rel[match, 0].max() <= 1e-4 and rel[match, 1].max() <= 1e-4


def test_random_dipole_frames_vs_oracle(gpu_ctx, oracle_mod):
    """The twelve random streamline frames of the charge tests (1 to 2,100
    sources, anisotropic boxes, h 0.002-0.02 A) with random moments, in both
    field forms, against the FP64 oracle under the bar above; sixteen-wave
    and one-wave workgroups give the same results bit for bit."""
    from pycpet_amd import ops
    from test_gpu_topo import _random_configurations

    try:
        for i, f in enumerate(_random_configurations()):
            mu = np.random.default_rng(3000 + i).standard_normal((len(f["q"]), 3))
            mu = (0.5 * mu).astype(np.float32)
            for physical in (False, True):
                label = f"random dipole frame {i}: N={len(mu)} h={f['h']} physical={physical}"
                print(label, flush=True)
                f64, s64, e64 = oracle_mod.topo_dipole_f64(
                    f["starts"], f["n_iter"], f["h"], f["dims"], f["x"], mu, physical=physical)
                runs = []
                for S in (16, 1):
                    gpu_ctx.set_tuning(2, S, 0)
                    runs.append(ops.topo_dipole(f["starts"], f["n_iter"], f["h"], f["dims"],
                                                f["x"], mu, physical=physical, ctx=gpu_ctx))
                r = runs[0]
                for a in ("feats", "steps", "exited"):
                    np.testing.assert_array_equal(getattr(runs[1], a), getattr(r, a))
                match = (r.steps == s64) & (r.exited == e64)
                rel = np.abs(r.feats - f64) / np.maximum(np.abs(f64), 1e-6)
                print(f"{label}: match {match.mean():.5f} dist max {rel[match, 0].max():.3e} "
                      f"curv max {rel[match, 1].max():.3e}", flush=True)
                assert match.mean() >= 0.99
                assert rel[match, 0].max() <= 1e-4 and rel[match, 1].max() <= 1e-4
    finally:
        gpu_ctx.set_tuning(2, 0, 0)


def test_dipole_on_axis_exact(gpu_ctx, oracle_mod):
    from pycpet_amd import ops

    dims = np.array([0.6, 0.5, 0.4], np.float32)
    x = np.array([[-3.0, 0.0, 0.0]], np.float32)
    mu = np.array([[0.7, 0.0, 0.0]], np.float32)
    a = np.array([[-0.31, 0, 0], [0.12, 0, 0], [0.55, 0, 0]], np.float32)
    n_iter = np.array([500, 500, 3], np.int32)
    r = ops.topo_dipole(a, n_iter, 0.01, dims, x, mu, ctx=gpu_ctx)
    f64, s64, e64 = oracle_mod.topo_dipole_f64(a, n_iter, 0.01, dims, x, mu)
    np.testing.assert_array_equal(r.steps, s64)
    np.testing.assert_array_equal(r.exited, e64)
    assert np.all(r.feats[:, 1] < 1e-9)


def test_dropin_thread_operation_dipole(gpu_ctx):
    from pycpet_amd import ops
    from pycpet_amd.c_ops import Math_ops

    g = np.load(GOLDEN / "ref_topo_dipole.npz")
    m = Math_ops()
    got = np.stack([m.thread_operation_dipole(g["starts"][i], g["n_iter"][i], g["x"], g["mu"],
                                              float(g["h"]), g["dims"])
                    for i in range(len(g["starts"]))])
    assert got.dtype == np.float32 and got.shape == g["feats"].shape
    want = g["feats"].astype(float)
    close = np.abs(got[:, 0] - want[:, 0]) <= 1e-3 * np.maximum(want[:, 0], 1e-2)
    assert close.mean() >= 0.9, close.mean()
    batch = ops.topo_dipole(g["starts"], g["n_iter"], float(g["h"]), g["dims"], g["x"], g["mu"],
                            ctx=gpu_ctx).feats.astype(np.float32)
    np.testing.assert_array_equal(got, batch)


def test_calculator_dipole_method_as_the_script_builds_it(gpu_ctx):
    """scripts/field_topology_dipole.py makes the calculator with __new__,
    sets x, mu, the starts and the step limits, and calls
    compute_topo_complete_c_shared_dipole: (L, 2) float32 rows in start
    order, the same numbers as the batched call on the reference golden."""
    import logging

    from pycpet_amd import ops
    from pycpet_amd.calculator import calculator

    g = np.load(GOLDEN / "ref_topo_dipole.npz")
    c = calculator.__new__(calculator)
    c.n_samples = len(g["starts"])
    c.step_size = float(g["h"])
    c.x = g["x"].astype(np.float32)
    c.mu = g["mu"].astype(np.float32)
    c.dimensions = g["dims"].astype(np.float32)
    c.random_start_points = g["starts"]
    c.random_max_samples = g["n_iter"]
    c.concur_slip = 4
    c.log = logging.getLogger("dipole-script")
    hist = c.compute_topo_complete_c_shared_dipole()
    assert hist.dtype == np.float32 and hist.shape == (len(g["starts"]), 2)
    batch = ops.topo_dipole(g["starts"], g["n_iter"], float(g["h"]), g["dims"], g["x"], g["mu"],
                            ctx=gpu_ctx).feats.astype(np.float32)
    np.testing.assert_array_equal(hist, batch)
    np.testing.assert_array_equal(c.hist, hist)


def test_field_topology_dipole_command(gpu_ctx, tmp_path, monkeypatch):
    """The dipole-topology command on one .dip frame: the .top file holds
    one row per start of the cube of starts, equal to the dipole kernel on
    the same starts and step limits (regenerated from the same global numpy
    seed), rotated into the box frame."""
    from pycpet_amd import field_topology_dipole as FT
    from pycpet_amd import ops
    from pycpet_amd.field_ops import initialize_box_points_uniform

    rng = np.random.default_rng(6)
    D = 40
    pos = rng.uniform(-6, 6, (D, 3))
    pos = pos[np.linalg.norm(pos, axis=1) > 2.5]
    mom = rng.standard_normal((len(pos), 3))
    table = np.column_stack([np.arange(len(pos)), pos, mom])
    (tmp_path / "in").mkdir()
    np.savetxt(tmp_path / "in" / "frame_a.v1.dip", table, header="idx x y z mx my mz")
    monkeypatch.chdir(tmp_path)
    np.random.seed(7)
    argv = ["-d", str(tmp_path / "in"), "-c", "0.5", "0", "0", "-x", "1.5", "0", "0",
            "-y", "0.5", "1", "0", "-dim", "1.5", "1.5", "1.5", "-n", "60", "-s", "0.05"]
    assert FT.main(argv) == 0
    got = np.loadtxt(tmp_path / "frame_a.top")
    assert got.shape == (64, 2)  # 60 starts round up to 4^3

    np.random.seed(7)
    dims = np.array([1.5, 1.5, 1.5], np.float32)
    center = np.array([0.5, 0.0, 0.0])
    starts, limits, T = initialize_box_points_uniform(
        center=center, x=np.array([1.5, 0, 0.0]), y=np.array([0.5, 1, 0.0]), dimensions=dims,
        N_cr=[4, 4, 4], dtype="float32", max_steps=round(2 * np.linalg.norm(dims) / 0.05),
        ret_rand_max=True, inclusive=False, seed=None)
    d = np.loadtxt(tmp_path / "in" / "frame_a.v1.dip", skiprows=1, dtype=np.float32)
    Tinv = np.linalg.inv(T)
    r = ops.topo_dipole(starts.reshape(-1, 3), limits, 0.05, dims,
                        ((d[:, 1:4] - center) @ Tinv).astype(np.float32),
                        (d[:, 4:7] @ Tinv).astype(np.float32), ctx=gpu_ctx)
    np.testing.assert_array_equal(got, r.feats.astype(np.float32).astype(np.float64))
