"""Pins the oracle's numpy restatements: grid rule, histogram binning, filters.

KATs come from the reference's own tests (tests/test_filter.py) and its
example output (examples/2A_3D-field: an 11^3 grid from -0.5 to 0.5); the
histogram restatement is checked against numpy.histogram2d itself, which is
what the reference calls (CPET/utils/calculator.py:702-707).
"""
import numpy as np
import pytest


def test_grid_rule_example_2a(oracle_mod):
    # examples/2A_3D-field/options: dimensions 0.5, step 0.1 -> "Sample Density: 11 11 11"
    ax = oracle_mod.grid_axes((0.5, 0.5, 0.5), 0.1)
    assert [a.size for a in ax] == [11, 11, 11]
    g = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    # first rows of 1_alcdehydro_run1_efield.dat (coordinates, %.3f)
    expect = ["-0.500 -0.500 -0.500", "-0.500 -0.500 -0.400", "-0.500 -0.500 -0.300",
              "-0.500 -0.500 -0.200", "-0.500 -0.500 -0.100"]
    assert [" ".join(f"{v:.3f}" for v in row) for row in g[:5]] == expect
    assert " ".join(f"{v:.3f}" for v in g[-1]) == "0.500 0.500 0.500"
    assert g.shape[0] == 1331


def test_grid_rule_truncation(oracle_mod):
    # 2*0.7/0.1 = 13.999999999999998 in float64 -> int() -> 13 -> 14 points
    ax = oracle_mod.grid_axes((0.7, 0.5, 0.5), 0.1)
    assert ax[0].size == 14 and ax[1].size == 11
    assert ax[0][0] == np.float32(-0.7) and ax[0][-1] == np.float32(0.7)


def test_histogram_restatement_matches_numpy(oracle_mod):
    rng = np.random.default_rng(0)
    a = rng.normal(1.0, 0.5, 3000)
    b = rng.gamma(2.0, 0.3, 3000)
    ea = np.linspace(0.0, 2.0, 21)
    eb = np.linspace(0.1, 1.5, 21)
    # plant values exactly on edges, on the last edge, outside, and NaN
    a[:21] = ea
    b[:21] = eb
    a[21:25] = [ea[-1], -1.0, 5.0, np.nan]
    b[25:29] = [eb[-1], eb[0], np.nan, 100.0]
    H = oracle_mod.histogram2d(a, b, ea, eb)
    Hn, _, _ = np.histogram2d(a, b, bins=[ea, eb])
    np.testing.assert_array_equal(H, Hn.astype(np.int64))
    # and with the (bins, range) form the reference uses
    Hr, _, _ = np.histogram2d(a, b, bins=[20, 20], range=[[0.0, 2.0], [0.1, 1.5]])
    np.testing.assert_array_equal(H, Hr.astype(np.int64))


def _synthetic():
    # tests/test_filter.py:15-48 of the reference
    x = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [2.1, 1.0, 0.0], [5.0, 5.0, 5.0]])
    Q = np.array([1.0, -0.5, 0.3, 1.2])
    resids = np.array(["WAT", "WAT", "FE", "NI"])
    resnums = np.array([10, 10, 651, 652])
    atom_numbers = np.array([1001, 1002, 2001, 3001])
    atom_types = np.array(["O", "H", "FE", "NI"])
    return x, Q, resids, resnums, atom_numbers, atom_types


def test_filter_radius_whole_residue_kat(oracle_mod):
    x, Q, resids, resnums, *_ = _synthetic()
    keep = oracle_mod.filter_radius_whole_residue_keep(x, np.zeros(3), 2.0, resids, resnums)
    assert keep.tolist() == [True, True, False, False]  # reference KAT: WAT#10 kept
    # a residue straddling the radius is kept whole; non-adjacent equal ids split
    res = np.array(["A", "A", "B", "A"])
    num = np.array([1, 1, 2, 1])
    xx = np.array([[0.5, 0, 0], [9, 0, 0], [9, 0, 0], [9, 0, 0]])
    keep = oracle_mod.filter_radius_whole_residue_keep(xx, np.zeros(3), 1.0, res, num)
    assert keep.tolist() == [True, True, False, False]


def test_filter_in_box_kat(oracle_mod):
    x, Q, *_ = _synthetic()
    keep = oracle_mod.filter_in_box_keep(x, np.array([1.0, 0.0, 0.0]), [1.5, 1.5, 1.5])
    assert keep.tolist() == [False, False, False, True]
    assert Q[keep].tolist() == [1.2]


def test_filter_atomspec_kat(oracle_mod):
    x, Q, resids, resnums, atom_numbers, atom_types = _synthetic()
    ID = [(atom_numbers[i], atom_types[i], resids[i], resnums[i], "A") for i in range(4)]
    # reference test_filter_IDs: remove atom_type FE (wildcards elsewhere) -> 3 remain
    fd = {"atom_type": ["FE"], "resid": [""], "resnum": [""], "chain": [""]}
    drop = oracle_mod.filter_atomspec_drop(ID, fd, intersect=True)
    assert drop == [2]
    # union form: any listed attribute value matches
    drop = oracle_mod.filter_atomspec_drop(ID, {"resid": ["WAT"], "atom_type": ["NI"]}, False)
    assert drop == [0, 1, 3]


def test_box_basis_closed_forms(oracle_mod):
    """box_basis (CPET/utils/calculator.py:200-215) on frames with known
    answers: rows are x^, y^, z^ with z^ = x^ x y^ and y^ = z^ x x^, so the
    y point's component along x^ is removed and the frame is right-handed."""
    c = np.array([1.0, 2.0, 3.0])
    # axis-aligned points: the identity, for any centre and distances
    np.testing.assert_allclose(oracle_mod.box_basis(c, c + [2.0, 0, 0], c + [0, 3.0, 0]),
                               np.eye(3), atol=1e-15)
    # a y point leaning along x: only its part orthogonal to x^ counts
    np.testing.assert_allclose(oracle_mod.box_basis(c, c + [1.0, 0, 0], c + [5.0, 1.0, 0]),
                               np.eye(3), atol=1e-15)
    # x along +y, y along +z: z^ = e_y x e_z = e_x, rows (e_y, e_z, e_x)
    want = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
    np.testing.assert_allclose(oracle_mod.box_basis(c, c + [0, 0.7, 0], c + [0, 0.2, 1.3]),
                               want, atol=1e-15)
    # random frames: orthonormal rows, determinant +1, x^ along the x point
    rng = np.random.default_rng(11)
    for _ in range(20):
        xp, yp = c + rng.standard_normal(3), c + rng.standard_normal(3)
        T = oracle_mod.box_basis(c, xp, yp)
        np.testing.assert_allclose(T @ T.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(T) - 1.0) < 1e-12
        np.testing.assert_allclose(T[0], (xp - c) / np.linalg.norm(xp - c), atol=1e-15)
        assert T[1] @ (yp - c) > 0 and abs(T[2] @ (yp - c)) < 1e-12
