"""Pins the FP64 streamline oracle (oracle/cpet_oracle.c, thread_operation
semantics of CPET/utils/math_module.c:523-591) before it is trusted.

Checks: the reference's compiled thread_operation on seeded inputs (golden
vectors), radial straight lines from a single charge (exact exit steps,
distance = steps * h, zero curvature), the continuous field-line curvature
|T x (J T)| / |E| from the analytic Coulomb Jacobian (the Euler finite
difference converges to it as h -> 0), and the n_iter = 0 / 1 edge cases.
"""
import numpy as np

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
