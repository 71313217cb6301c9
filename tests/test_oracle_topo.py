"""Pins the FP64 streamline oracle (oracle/cpet_oracle.c, thread_operation
semantics of CPET/utils/math_module.c:523-591) before it is trusted.

Checks: the reference's compiled thread_operation on seeded inputs (golden
vectors), radial straight lines from a single charge (exact exit steps,
distance = steps * h, zero curvature), the continuous field-line curvature
|T x (J T)| / |E| from the analytic Coulomb Jacobian (the Euler finite
difference converges to it as h -> 0), and the n_iter = 0 / 1 edge cases.
"""
import numpy as np
import pytest

from conftest import GOLDEN

K = 14.3996451


def test_oracle_matches_reference_topo(oracle_mod):
    g = np.load(GOLDEN / "ref_topo_200.npz")
    feats, steps, exited = oracle_mod.topo_f64(g["starts"], g["n_iter"], float(g["h"]),
                                               g["dims"], g["x"], g["q"])
    ref = g["ret"].astype(float)
    rd = np.abs(ref[:, 0] - feats[:, 0]) / np.maximum(feats[:, 0], 1e-12)
    # the FP32 reference drifts off the FP64 path only near box-exit ties
    assert np.mean(rd < 1e-4) >= 0.95, np.mean(rd < 1e-4)
    ok = rd < 1e-4
    rc = np.abs(ref[ok, 1] - feats[ok, 1]) / np.abs(feats[ok, 1])
    assert np.median(rc) < 1e-2, np.median(rc)  # FP32 finite-difference noise
    assert exited.sum() > 0 and (exited == 0).sum() > 0  # both stop kinds occur


def _radial_expect(starts, c, sign, dims, h, n_iter):
    u = sign * (starts.astype(float) - c)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    t = np.full(len(starts), np.inf)
    for i in range(3):
        with np.errstate(divide="ignore", invalid="ignore"):
            ti = np.where(u[:, i] > 0, (dims[i] - starts[:, i]) / u[:, i],
                          np.where(u[:, i] < 0, (-dims[i] - starts[:, i]) / u[:, i], np.inf))
        t = np.minimum(t, ti)
    frac = t / h - np.floor(t / h)
    n_exit = np.floor(t / h).astype(np.int64) + 1
    return np.minimum(n_iter, n_exit), (frac > 1e-6) & (frac < 1 - 1e-6), n_exit <= n_iter


def test_radial_lines_exact(oracle_mod):
    rng = np.random.default_rng(7)
    dims = np.array([0.6, 0.5, 0.4], np.float32)
    h = 0.01
    c = np.array([3.0, 0.5, -0.4])
    for sign in (+1.0, -1.0):
        starts = rng.uniform(-0.35, 0.35, (64, 3)).astype(np.float32)
        n_iter = rng.integers(1, 200, 64).astype(np.int32)
        feats, steps, exited = oracle_mod.topo_f64(starts, n_iter, h, dims,
                                                   c[None].astype(np.float32),
                                                   np.array([sign * 0.8], np.float32))
        exp_steps, clean, exp_exit = _radial_expect(starts, c.astype(np.float32).astype(float),
                                                    sign, dims.astype(float), float(np.float32(h)),
                                                    n_iter)
        assert clean.sum() > 50
        np.testing.assert_array_equal(steps[clean], exp_steps[clean])
        np.testing.assert_array_equal(exited[clean].astype(bool), exp_exit[clean])
        np.testing.assert_allclose(feats[:, 0], steps * float(np.float32(h)), rtol=1e-9)
        assert np.abs(feats[:, 1]).max() < 1e-7


def _continuous_curvature(p, x, q):
    d = p[None, :] - x.astype(float)
    r = np.linalg.norm(d, axis=1)
    qq = q.astype(float)
    E = K * (qq[:, None] * d / r[:, None] ** 3).sum(0)
    J = K * (qq[:, None, None] * (np.eye(3)[None] / r[:, None, None] ** 3
                                  - 3 * d[:, :, None] * d[:, None, :] / r[:, None, None] ** 5)).sum(0)
    T = E / np.linalg.norm(E)
    return np.linalg.norm(np.cross(T, J @ T)) / np.linalg.norm(E)


def test_curvature_matches_continuous_limit(oracle_mod):
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((12, 3)) * 2.5).astype(np.float32)
    x = x[np.linalg.norm(x, axis=1) > 1.2]
    q = rng.normal(0, 0.6, x.shape[0]).astype(np.float32)
    starts = rng.uniform(-0.3, 0.3, (20, 3)).astype(np.float32)
    n_iter = np.zeros(20, np.int32)  # xf = x0, so the feature is kappa(x0)
    feats, steps, _ = oracle_mod.topo_f64(starts, n_iter, 1e-5, np.full(3, 0.5, np.float32), x, q)
    assert np.all(steps == 0) and np.all(feats[:, 0] == 0)
    kc = np.array([_continuous_curvature(s.astype(float), x, q) for s in starts])
    np.testing.assert_allclose(feats[:, 1], kc, rtol=2e-3)


def test_curvature_identity(oracle_mod):
    """|v' x v''|/|v'|^3 with Euler unit steps equals |u0 x u1| / h."""
    rng = np.random.default_rng(9)
    x = (rng.standard_normal((30, 3)) * 3).astype(np.float32)
    x = x[np.linalg.norm(x, axis=1) > 1.5]
    q = rng.normal(0, 0.5, x.shape[0]).astype(np.float32)
    starts = rng.uniform(-0.4, 0.4, (16, 3)).astype(np.float32)
    h = float(np.float32(0.01))
    feats, steps, exited, p6 = oracle_mod.topo_f64(starts, np.full(16, 30, np.int32), h,
                                                    np.full(3, 0.5, np.float32), x, q,
                                                    want_points=True)
    def kappa(a, b, c):
        u0 = (b - a) / h
        u1 = (c - b) / h
        return np.linalg.norm(np.cross(u0, u1), axis=1) / h
    k = 0.5 * (kappa(p6[:, 0], p6[:, 1], p6[:, 2]) + kappa(p6[:, 3], p6[:, 4], p6[:, 5]))
    np.testing.assert_allclose(feats[:, 1], k, rtol=1e-6)
    np.testing.assert_allclose(feats[:, 0], np.linalg.norm(p6[:, 0] - p6[:, 3], axis=1), rtol=1e-12)


def test_one_step_and_zero_steps(oracle_mod):
    x = np.array([[2.0, 1.0, -1.5]], np.float32)
    q = np.array([0.3], np.float32)
    starts = np.array([[0.1, 0.0, 0.2], [0.0, 0.1, 0.0]], np.float32)
    feats, steps, exited = oracle_mod.topo_f64(starts, np.array([1, 0], np.int32), 0.01,
                                               np.full(3, 1.0, np.float32), x, q)
    assert list(steps) == [1, 0] and list(exited) == [0, 0]
    np.testing.assert_allclose(feats[0, 0], float(np.float32(0.01)), rtol=1e-12)
    assert feats[1, 0] == 0.0


def _dipoles(seed, N=60, R=6.0):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((N, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x = (d * rng.uniform(2.0, R, (N, 1))).astype(np.float32)
    mu = rng.normal(0.0, 0.3, (N, 3)).astype(np.float32)
    return x, mu


def test_dipole_oracle_matches_reference_c(oracle_mod):
    """FP64 dipole walk vs the reference's own thread_operation_dipole (FP32,
    -ffast-math): steps and stop kind on >= 90 % of starts, distances there
    within 1e-3 (box-exit ties and FP32 drift account for the rest)."""
    if not oracle_mod.Ref.available():
        pytest.skip("reference C not built here")
    ref = oracle_mod.Ref()
    x, mu = _dipoles(3)
    rng = np.random.default_rng(4)
    dims = np.array([0.5, 0.5, 0.5], np.float32)
    starts = rng.uniform(-0.45, 0.45, (60, 3)).astype(np.float32)
    n_iter = rng.integers(1, 120, 60).astype(np.int32)
    feats, steps, exited = oracle_mod.topo_dipole_f64(starts, n_iter, 0.01, dims, x, mu)
    want = np.stack([ref.thread_operation_dipole(starts[i], n_iter[i], 0.01, dims, x, mu)
                     for i in range(60)])
    close = np.abs(feats[:, 0] - want[:, 0]) <= 1e-3 * np.maximum(want[:, 0], 1e-2)
    assert close.mean() >= 0.9, close.mean()
    # curvature is an FP32 finite difference in the reference: order of magnitude
    ok = close & (want[:, 1] > 0)
    assert np.median(np.abs(np.log10(feats[ok, 1] / want[ok, 1]))) < 0.5


def test_dipole_on_axis_exact(oracle_mod):
    """A dipole mu x-hat at the origin: on the x axis both field forms give
    E = 2 k mu / r^3 x-hat, so a start (a, 0, 0) walks straight out along x
    and stops on the first step past the box face, curvature 0."""
    dims = np.array([0.6, 0.5, 0.4], np.float32)
    h = 0.01
    x = np.array([[-3.0, 0.0, 0.0]], np.float32)
    mu = np.array([[0.7, 0.0, 0.0]], np.float32)
    a = np.array([[-0.31, 0, 0], [0.12, 0, 0], [0.55, 0, 0]], np.float32)
    n_iter = np.array([500, 500, 3], np.int32)
    for physical in (False, True):
        feats, steps, exited = oracle_mod.topo_dipole_f64(a, n_iter, h, dims, x, mu,
                                                          physical=physical)
        exp_steps, clean, exp_exit = _radial_expect(a, x[0].astype(float), 1.0, dims, h, n_iter)
        assert np.all(steps[clean] == exp_steps[clean]) and np.all(exited[clean] == exp_exit[clean])
        np.testing.assert_allclose(feats[clean, 0], steps[clean] * np.float64(np.float32(h)),
                                   rtol=1e-12)
        assert np.all(feats[:, 1] < 1e-9)


def test_dipole_forms_agree_for_x_moments(oracle_mod):
    """With every moment along x the reference's coefficient 3 mu_x R_x +
    mu_y R_y + mu_z R_z equals the physical 3 mu . R, so both walks agree;
    with general moments they differ."""
    x, mu = _dipoles(8)
    rng = np.random.default_rng(9)
    dims = np.array([0.5, 0.5, 0.5], np.float32)
    starts = rng.uniform(-0.4, 0.4, (30, 3)).astype(np.float32)
    n_iter = np.full(30, 80, np.int32)
    mux = mu.copy()
    mux[:, 1:] = 0.0
    a = oracle_mod.topo_dipole_f64(starts, n_iter, 0.01, dims, x, mux)
    b = oracle_mod.topo_dipole_f64(starts, n_iter, 0.01, dims, x, mux, physical=True)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_allclose(a[0], b[0], rtol=1e-12, atol=1e-15)
    c = oracle_mod.topo_dipole_f64(starts, n_iter, 0.01, dims, x, mu)
    d = oracle_mod.topo_dipole_f64(starts, n_iter, 0.01, dims, x, mu, physical=True)
    assert np.abs(c[0] - d[0]).max() > 1e-3


def test_dipole_oracle_matches_golden(oracle_mod):
    """The FP64 dipole walk against the reference C's own outputs kept in
    tests/golden/ref_topo_dipole.npz (FP32 -ffast-math): distance within
    1e-3 on >= 90 % of the 160 starts."""
    g = np.load(GOLDEN / "ref_topo_dipole.npz")
    feats, _, _ = oracle_mod.topo_dipole_f64(g["starts"], g["n_iter"], float(g["h"]), g["dims"],
                                             g["x"], g["mu"])
    want = g["feats"].astype(float)
    close = np.abs(feats[:, 0] - want[:, 0]) <= 1e-3 * np.maximum(want[:, 0], 1e-2)
    assert close.mean() >= 0.9, close.mean()


def test_rk4_radial_lines_equal_euler(oracle_mod):
    """A single charge's field lines are straight radial lines: every RK4
    stage sees the same direction, so the RK4 walk equals the Euler walk
    (same exit steps, distance = steps * h, zero curvature)."""
    x = np.array([[0.0, 0.0, 0.0]], np.float32)
    q = np.array([1.0], np.float32)
    rng = np.random.default_rng(2)
    d = rng.standard_normal((40, 3))
    starts = (0.3 * d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    n_iter = rng.integers(1, 400, 40).astype(np.int32)
    dims = np.full(3, 1.0, np.float32)
    fe, se, ee = oracle_mod.topo_f64(starts, n_iter, 0.01, dims, x, q)
    fr, sr, er = oracle_mod.topo_rk4_f64(starts, n_iter, 0.01, dims, x, q)
    np.testing.assert_array_equal(sr, se)
    np.testing.assert_array_equal(er, ee)
    np.testing.assert_allclose(fr[:, 0], fe[:, 0], rtol=1e-12)
    np.testing.assert_allclose(fr[:, 0], sr * float(np.float32(0.01)), rtol=1e-9)
    assert np.abs(fr[:, 1]).max() < 1e-6


def _two_charges():
    x = np.array([[1.2, 0.3, -0.2], [-1.1, -0.4, 0.5]], np.float32)
    q = np.array([0.8, -0.6], np.float32)
    return x, q


def test_rk4_fourth_order(oracle_mod):
    """End point after a fixed arc length in a curved two-charge field: the
    error against a much finer RK4 walk falls 16x per halving of h (Euler:
    2x).  A wrong stage weight or stage point lowers the order."""
    x, q = _two_charges()
    start = np.array([[0.05, 0.02, 0.0]], np.float32)
    dims = np.full(3, 50.0, np.float32)  # no exit
    S = 0.64  # arc length, a multiple of every h below (exact in float32)
    end = {}
    for h in (0.08, 0.04, 0.02, 0.0025):
        n = np.array([int(round(S / h))], np.int32)
        *_, pts = oracle_mod.topo_rk4_f64(start, n, h, dims, x, q, want_points=True)
        end[h] = pts[0, 3]
    e1 = np.linalg.norm(end[0.08] - end[0.0025])
    e2 = np.linalg.norm(end[0.04] - end[0.0025])
    e3 = np.linalg.norm(end[0.02] - end[0.0025])
    assert 12.0 < e1 / e2 < 20.0, e1 / e2
    assert 12.0 < e2 / e3 < 20.0, e2 / e3
    # the Euler walk of the same arc is first order and far less accurate
    *_, pe = oracle_mod.topo_f64(start, np.array([8], np.int32), 0.08, dims, x, q,
                                 want_points=True)
    assert np.linalg.norm(pe[0, 3] - end[0.0025]) > 100 * e1


def test_rk4_agrees_with_euler_as_h_shrinks(oracle_mod):
    x, q = _two_charges()
    start = np.array([[0.05, 0.02, 0.0]], np.float32)
    dims = np.full(3, 50.0, np.float32)
    *_, pr = oracle_mod.topo_rk4_f64(start, np.array([20], np.int32), 0.01, dims, x, q,
                                     want_points=True)
    *_, pe = oracle_mod.topo_f64(start, np.array([20000], np.int32), 0.00001, dims, x, q,
                                 want_points=True)
    assert np.linalg.norm(pr[0, 3] - pe[0, 3]) < 1e-5


def test_f32_port_matches_reference_golden(oracle_mod):
    """The FP32 port of thread_operation (oracle_topo_f32, the restated CPU
    baseline) against the reference C's own FP32 outputs kept in
    tests/golden/ref_topo_200.npz: both walk in FP32, so distances agree to
    a few float32 ulps and curvatures to FP32 finite-difference noise."""
    g = np.load(GOLDEN / "ref_topo_200.npz")
    f32, s32 = oracle_mod.topo_f32(g["starts"], g["n_iter"], float(g["h"]), g["dims"], g["x"],
                                   g["q"])
    ref = g["ret"].astype(float)
    rd = np.abs(ref[:, 0] - f32[:, 0]) / np.maximum(ref[:, 0], 1e-12)
    assert rd.max() < 1e-5, rd.max()
    rc = np.abs(ref[:, 1] - f32[:, 1]) / np.abs(ref[:, 1])
    assert np.median(rc) < 1e-2 and np.percentile(rc, 90) < 3e-2, (np.median(rc),)
    # the same walk as the FP64 oracle: identical step counts on this case
    _, s64, _ = oracle_mod.topo_f64(g["starts"], g["n_iter"], float(g["h"]), g["dims"], g["x"],
                                    g["q"])
    np.testing.assert_array_equal(s32, s64)


def test_f32_port_matches_live_reference(oracle_mod):
    """A fresh seeded case at h = 1e-3 through the reference's compiled
    thread_operation (oracle/_ref), one call per streamline."""
    if not oracle_mod.Ref.available():
        pytest.skip("reference C not built here")
    rng = np.random.default_rng(21)
    d = rng.standard_normal((300, 3))
    x = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(3.0, 12.0, (300, 1)))
    x = x.astype(np.float32)
    q = rng.normal(0, 0.4, 300).astype(np.float32)
    starts = rng.uniform(-0.9, 0.9, (48, 3)).astype(np.float32)
    n_iter = rng.integers(1, 1500, 48).astype(np.int32)
    dims = np.full(3, 1.0, np.float32)
    f32, _ = oracle_mod.topo_f32(starts, n_iter, 1e-3, dims, x, q)
    want = oracle_mod.Ref().topo(starts, n_iter, 1e-3, dims, x, q).astype(float)
    rd = np.abs(want[:, 0] - f32[:, 0]) / np.maximum(want[:, 0], 1e-12)
    assert np.mean(rd < 1e-4) >= 0.95, np.mean(rd < 1e-4)


def test_f32_port_radial_lines_and_zero_steps(oracle_mod):
    """Closed forms for the FP32 port: radial straight lines from one charge
    stop at the exact exit step away from ties, with distance = steps * h to
    float32 rounding and curvature ~ 0; n_iter = 0 gives distance 0."""
    rng = np.random.default_rng(7)
    dims = np.array([0.6, 0.5, 0.4], np.float32)
    h = 0.01
    c = np.array([3.0, 0.5, -0.4])
    starts = rng.uniform(-0.35, 0.35, (64, 3)).astype(np.float32)
    n_iter = rng.integers(1, 200, 64).astype(np.int32)
    n_iter[:4] = 0
    f32, s32 = oracle_mod.topo_f32(starts, n_iter, h, dims, c[None].astype(np.float32),
                                   np.array([0.8], np.float32))
    exp_steps, clean, _ = _radial_expect(starts, c.astype(np.float32).astype(float), 1.0,
                                         dims.astype(float), float(np.float32(h)), n_iter)
    clean &= n_iter > 0
    assert clean.sum() > 50
    np.testing.assert_array_equal(s32[clean], exp_steps[clean])
    np.testing.assert_allclose(f32[clean, 0], s32[clean] * float(np.float32(h)), rtol=1e-5)
    assert np.all(f32[:4, 0] == 0.0) and np.all(s32[:4] == 0)
    assert np.abs(f32[:, 1]).max() < 0.5  # FP32 finite differences of straight lines


def _boundary_starts():
    """Starts outside the box, exactly on faces, edges and a corner, and
    just inside; the reference steps once before its strict box test
    (math_module.c:534-556)."""
    return np.array([[1.5, 0.0, 0.0], [0.0, -1.2, 0.3], [1.0, 0.0, 0.0], [-1.0, 0.2, 0.1],
                     [0.3, 1.0, -1.0], [-1.0, -1.0, -1.0], [0.9999, 0.0, 0.0],
                     [0.0, 0.0, 0.0]], np.float32)


def test_boundary_starts_vs_live_reference(oracle_mod):
    """Outside / on-face / corner starts: the FP64 oracle's distance and
    curvature against the reference's compiled thread_operation, and its
    exit rule (a start outside leaves after one step)."""
    if not oracle_mod.Ref.available():
        pytest.skip("reference C not built here")
    x = np.array([[2.0, 1.0, -1.5], [-1.7, 0.4, 2.2], [0.3, -2.5, 0.8]], np.float32)
    q = np.array([0.3, -0.6, 0.45], np.float32)
    dims = np.full(3, 1.0, np.float32)
    starts = _boundary_starts()
    n_iter = np.full(len(starts), 400, np.int32)
    f64, s64, e64 = oracle_mod.topo_f64(starts, n_iter, 0.01, dims, x, q)
    assert s64[0] == 1 and e64[0] == 1 and s64[1] == 1 and e64[1] == 1
    want = oracle_mod.Ref().topo(starts, n_iter, 0.01, dims, x, q).astype(float)
    np.testing.assert_allclose(f64[:, 0], want[:, 0], rtol=2e-4, atol=1e-6)
    np.testing.assert_allclose(f64[:, 1], want[:, 1], rtol=5e-2, atol=1e-3)
