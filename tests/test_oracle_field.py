"""Pins the FP64 field/ESP oracle (oracle/cpet_oracle.c) before it is trusted.

Checks: the reference's own compiled kernels on seeded inputs (golden
vectors from tests/golden/make_golden.py), closed forms (single charge),
superposition, symmetric cancellation, the sign convention and the
r^2 >= 1e-6 floor of compute_looped_field (CPET/utils/math_module.c:407,433).
"""
import numpy as np
import pytest

from conftest import GOLDEN

K = 14.3996451


def rel_err(a, b, floor_frac=1e-3):
    """Pointwise relative error with a denominator floor of floor_frac x RMS
    of the reference magnitude, and normwise RMS relative error."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if b.ndim == 1:
        a, b = a[:, None], b[:, None]
    nb = np.linalg.norm(b, axis=1)
    d = np.linalg.norm(a - b, axis=1)
    rms = np.sqrt(np.mean(nb ** 2))
    return float((d / np.maximum(nb, floor_frac * rms)).max()), float(np.sqrt(np.mean(d ** 2)) / rms)


def mesh(ax, ay, az):
    return np.stack(np.meshgrid(ax, ay, az, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


def test_oracle_matches_reference_volume(oracle_mod):
    g = np.load(GOLDEN / "ref_volume_500.npz")
    pts = mesh(g["ax"], g["ay"], g["az"])
    E, phi = oracle_mod.field_f64(pts, g["x"], g["q"], want_E=True, want_phi=True)
    mx, rms = rel_err(g["E"], E)
    assert mx < 5e-5 and rms < 3e-6, (mx, rms)  # FP32 -ffast-math reference
    mx, rms = rel_err(g["phi"], phi)
    assert mx < 5e-5 and rms < 3e-6, (mx, rms)


def test_oracle_matches_reference_point_field(oracle_mod):
    g = np.load(GOLDEN / "ref_volume_500.npz")
    E, _ = oracle_mod.field_f64(np.zeros((1, 3), np.float32), g["x"], g["q"], soften=False)
    np.testing.assert_allclose(E[0], g["point_field"], rtol=2e-5, atol=1e-6)


def test_oracle_floor_matches_reference(oracle_mod):
    g = np.load(GOLDEN / "ref_volume_floor.npz")
    pts = mesh(g["ax"], g["ay"], g["az"])
    E, _ = oracle_mod.field_f64(pts, g["x"], g["q"])
    mx, _ = rel_err(g["E"], E)
    assert mx < 5e-5, mx
    # the near-singular point is dominated by the floored term
    i = np.argmax(np.linalg.norm(g["E"], axis=1))
    Eh, _ = oracle_mod.field_f64(pts[i:i + 1], g["x"], g["q"], soften=False)
    assert np.linalg.norm(Eh[0]) > 2 * np.linalg.norm(E[i])  # dropping the floor fails


def test_single_charge_closed_form(oracle_mod):
    rng = np.random.default_rng(3)
    x = np.array([[0.3, -0.2, 0.1]], np.float32)
    q = np.array([0.55], np.float32)
    pts = rng.uniform(-2, 2, (50, 3)).astype(np.float32)
    E, phi = oracle_mod.field_f64(pts, x, q, want_phi=True)
    d = pts.astype(float) - x.astype(float)
    r = np.linalg.norm(d, axis=1)
    np.testing.assert_allclose(E, K * float(q[0]) * d / r[:, None] ** 3, rtol=1e-12)
    np.testing.assert_allclose(phi, K * float(q[0]) / r, rtol=1e-12)
    # positive charge: field points away from the charge
    assert np.all(np.einsum("ij,ij->i", E, d) > 0)


def test_floor_closed_form(oracle_mod):
    x = np.array([[0.0, 0.0, 0.0]], np.float32)
    q = np.array([1.0], np.float32)
    p = np.array([[1e-4, 0.0, 0.0]], np.float32)
    E, _ = oracle_mod.field_f64(p, x, q)
    r2f = float(np.float32(1e-6))
    np.testing.assert_allclose(E[0, 0], K * float(p[0, 0]) / r2f ** 1.5, rtol=1e-12)


def test_superposition_and_cancellation(oracle_mod):
    rng = np.random.default_rng(4)
    x1 = rng.uniform(-5, 5, (40, 3)).astype(np.float32)
    x2 = rng.uniform(-5, 5, (30, 3)).astype(np.float32)
    q1 = rng.normal(0, 0.5, 40).astype(np.float32)
    q2 = rng.normal(0, 0.5, 30).astype(np.float32)
    pts = rng.uniform(-1, 1, (20, 3)).astype(np.float32)
    Ea, pa = oracle_mod.field_f64(pts, x1, q1, want_phi=True)
    Eb, pb = oracle_mod.field_f64(pts, x2, q2, want_phi=True)
    Ec, pc = oracle_mod.field_f64(pts, np.concatenate([x1, x2]), np.concatenate([q1, q2]),
                                  want_phi=True)
    np.testing.assert_allclose(Ec, Ea + Eb, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pc, pa + pb, rtol=1e-12, atol=1e-12)
    # equal charges at +-a on z: no z-field on the z = 0 plane; +q/-q: phi = 0 there
    xs = np.array([[0, 0, 1.5], [0, 0, -1.5]], np.float32)
    plane = np.column_stack([rng.uniform(-1, 1, (10, 2)), np.zeros(10)]).astype(np.float32)
    E, _ = oracle_mod.field_f64(plane, xs, np.array([0.4, 0.4], np.float32))
    assert np.abs(E[:, 2]).max() < 1e-12
    _, phi = oracle_mod.field_f64(plane, xs, np.array([0.4, -0.4], np.float32),
                                  want_E=False, want_phi=True)
    assert np.abs(phi).max() < 1e-12


def test_fp32_port_within_fp32_rounding(oracle_mod):
    g = np.load(GOLDEN / "ref_volume_500.npz")
    pts = mesh(g["ax"], g["ay"], g["az"])
    E64, p64 = oracle_mod.field_f64(pts, g["x"], g["q"], want_phi=True)
    E32, p32 = oracle_mod.field_f32(pts, g["x"], g["q"], want_phi=True)
    assert rel_err(E32, E64)[0] < 5e-5
    assert rel_err(p32, p64)[0] < 5e-5
