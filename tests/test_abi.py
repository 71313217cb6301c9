"""The C-ABI library loads and exports every entry point include/cpet_hip.h
declares (no compute calls; this runs without a GPU)."""
import ctypes
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "cpet_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|char\s*\*|const\s+char\s*\*)\s*\*?\s*(\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_declares_reference_names():
    names = declared_functions()
    for n in ["compute_looped_field", "calc_esp_base", "calc_field", "calc_field_base",
              "thread_operation", "cpet_field_grid", "cpet_field_points", "cpet_topo",
              "cpet_ctx_create", "cpet_last_error"]:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    from pycpet_amd import _native as nat

    if not nat.LIB_PATH.exists():
        pytest.fail("libcpet_hip.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", str(nat.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    lib = nat.load()
    for n in declared_functions():
        assert hasattr(lib, n)
    assert set(nat.SIGNATURES) | {"cpet_last_error"} >= set(declared_functions())


def test_gpu_code_object_targets_gfx950():
    from pycpet_amd import _native as nat

    # the offload bundle of our own build names its target triple
    assert b"amdgcn-amd-amdhsa--gfx950" in nat.LIB_PATH.read_bytes()


# ---- the reference binding (CPET/utils/c_ops.py:26-159) against this library

BINDINGS = ROOT / "tests" / "golden" / "ref_c_ops_bindings.json"


def _reference_argtypes(spec):
    """ctypes argtypes from the committed list of the reference binding."""
    import numpy as np
    import numpy.ctypeslib as npct

    scalars = {"int": ctypes.c_int, "float": ctypes.c_float}
    out = []
    for s in spec:
        if s in scalars:
            out.append(scalars[s])
            continue
        dt, nd = s.split("_")
        out.append(npct.ndpointer(dtype={"f32": np.float32, "i32": np.int32, "f64": np.double}[dt],
                                  ndim=int(nd[0]), flags="C"))
    return out


def _bind_like_reference(path):
    """What Math_ops.__init__ does with its shared library: CDLL, then
    argtypes (and restype None) on every symbol of the list."""
    import json

    spec = json.loads(BINDINGS.read_text())["symbols"]
    lib = ctypes.CDLL(str(path))
    for name, args in spec.items():
        fn = getattr(lib, name)  # AttributeError here = the constructor would fail
        fn.argtypes = _reference_argtypes(args)
        fn.restype = None
    return lib, spec


def test_reference_binding_binds_every_symbol():
    from pycpet_amd import _native as nat

    lib, spec = _bind_like_reference(nat.LIB_PATH)
    assert len(spec) == 14
    declared = set(declared_functions())
    missing = [n for n in spec if n not in declared]
    assert not missing, missing  # every bound name is a declared entry point too
    assert set(spec) <= set(nat.SIGNATURES)


def _ref_lib():
    from oracle import oracle as O

    if not O.REF_SO.exists():
        return None
    return ctypes.CDLL(str(O.REF_SO))


def test_host_helpers_through_reference_binding():
    """The float32 helpers called exactly as the reference binding calls them,
    against numpy and (where oracle/_ref is built) the reference's own C."""
    import numpy as np

    from pycpet_amd import _native as nat

    lib, _ = _bind_like_reference(nat.LIB_PATH)
    rng = np.random.default_rng(11)
    A = rng.standard_normal(257).astype(np.float32)
    B = rng.standard_normal(257).astype(np.float32)
    res = np.zeros(257, np.float32)
    lib.vecaddn(res, A, B, 257)
    np.testing.assert_array_equal(res, A + B)

    M = rng.standard_normal((13, 70)).astype(np.float32)
    rs = np.zeros(13, np.float32)
    lib.einsum_ij_i(13, 70, M, rs)
    seq = np.zeros(13, np.float32)
    for j in range(70):
        seq += M[:, j]
    np.testing.assert_array_equal(rs, seq)  # float32, row order
    np.testing.assert_allclose(rs, M.astype(float).sum(1), rtol=0, atol=2e-5)

    M3 = rng.standard_normal((4, 9, 33)).astype(np.float32)
    rb = np.zeros((4, 9), np.float32)
    lib.einsum_ij_i_batch(4, 9, 33, M3, rb)
    np.testing.assert_allclose(rb, M3.astype(float).sum(2), rtol=0, atol=2e-5)

    n = 500
    R = rng.standard_normal((3, n, 3)).astype(np.float32)
    rmag = rng.uniform(0.01, 1.0, (3, n)).astype(np.float32)
    Q = rng.standard_normal(n).astype(np.float32)
    eb = np.zeros((3, 3), np.float32)
    lib.einsum_operation_batch(3, n, rmag, Q, R, eb)
    exact = 14.3996451 * np.einsum("ki,i,kid->kd", rmag.astype(float), Q.astype(float),
                                   R.astype(float))
    scale = 14.3996451 * np.einsum("ki,i,kid->kd", rmag.astype(float), np.abs(Q.astype(float)),
                                   np.abs(R.astype(float)))
    assert np.all(np.abs(eb - exact) <= 1e-5 * scale)
    e1 = np.zeros(3, np.float32)
    lib.einsum_operation(n, np.ascontiguousarray(rmag[1]), Q, np.ascontiguousarray(R[1]), e1)
    np.testing.assert_array_equal(e1, eb[1])  # same loop for one point

    ref = _ref_lib()
    if ref is not None:
        for name in ("vecaddn", "einsum_ij_i", "einsum_ij_i_batch", "einsum_operation_batch"):
            getattr(ref, name).restype = None
        r2 = np.zeros(257, np.float32)
        ref.vecaddn(r2.ctypes.data_as(ctypes.c_void_p), A.ctypes.data_as(ctypes.c_void_p),
                    B.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(257))
        np.testing.assert_array_equal(res, r2)
        rr = np.zeros(13, np.float32)
        ref.einsum_ij_i(ctypes.c_int(13), ctypes.c_int(70), M.ctypes.data_as(ctypes.c_void_p),
                        rr.ctypes.data_as(ctypes.c_void_p))
        np.testing.assert_allclose(rs, rr, rtol=0, atol=2e-5)  # -ffast-math may reorder
        re = np.zeros((3, 3), np.float32)
        ref.einsum_operation_batch(ctypes.c_int(3), ctypes.c_int(n),
                                   rmag.ctypes.data_as(ctypes.c_void_p),
                                   Q.ctypes.data_as(ctypes.c_void_p),
                                   R.ctypes.data_as(ctypes.c_void_p),
                                   re.ctypes.data_as(ctypes.c_void_p))
        assert np.all(np.abs(eb - re) <= 2e-5 * scale)


def test_fp64_helpers_match_their_definition():
    """sparse_dot and dot read FP64 arrays (math_module.c:15,49); checked
    against scipy.sparse / numpy and the reference's own C."""
    import numpy as np
    import scipy.sparse as sp

    from pycpet_amd.c_ops import Math_ops

    m = Math_ops()
    rng = np.random.default_rng(12)
    S = sp.random(40, 25, density=0.2, random_state=3, format="csr")
    S.data = rng.standard_normal(S.nnz)
    v = rng.standard_normal(25)
    np.testing.assert_allclose(m.sparse_dot(S, v), S @ v, rtol=1e-13, atol=1e-13)
    D = rng.standard_normal((17, 9))
    w = rng.standard_normal(9)
    np.testing.assert_allclose(m.dot(D, w), D @ w, rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(m.vecaddn(np.ones(4), np.arange(4)), np.arange(1, 5))
    Z = sp.csr_matrix((3, 4))
    np.testing.assert_array_equal(m.sparse_dot(Z, np.ones(4)), np.zeros(3))

    ref = _ref_lib()
    if ref is not None:
        out = np.zeros(40)
        ip = S.indptr.astype(np.int32)
        ix = S.indices.astype(np.int32)
        ref.sparse_dot.restype = None
        P = ctypes.c_void_p
        ref.sparse_dot(out.ctypes.data_as(P), ip.ctypes.data_as(P), ctypes.c_int(41),
                       ix.ctypes.data_as(P), ctypes.c_int(ix.size),
                       S.data.ctypes.data_as(P), ctypes.c_int(S.nnz), v.ctypes.data_as(P),
                       ctypes.c_int(25))
        np.testing.assert_allclose(m.sparse_dot(S, v), out, rtol=1e-13, atol=1e-13)
        o2 = np.zeros(17)
        ref.dot.restype = None
        ref.dot(o2.ctypes.data_as(P), D.ctypes.data_as(P), w.ctypes.data_as(P),
                ctypes.c_int(17), ctypes.c_int(9))
        np.testing.assert_allclose(m.dot(D, w), o2, rtol=1e-13, atol=1e-13)


def test_host_helpers_reject_bad_arguments():
    import numpy as np

    from pycpet_amd import _native as nat

    lib = nat.load()
    ip = np.array([0, 2, 1], np.int32)  # decreasing row pointer
    ix = np.array([0, 1], np.int32)
    a = np.ones(2)
    out = np.zeros(2)
    rc = lib.sparse_dot(out.ctypes.data, ip.ctypes.data, 3, ix.ctypes.data, 2, a.ctypes.data, 2,
                        a.ctypes.data, 2)
    assert rc == nat.CPET_ERR_ARG and "out of range" in nat.last_error(lib)
    ip = np.array([0, 2], np.int32)
    ix = np.array([0, 5], np.int32)  # column beyond B
    rc = lib.sparse_dot(out.ctypes.data, ip.ctypes.data, 2, ix.ctypes.data, 2, a.ctypes.data, 2,
                        a.ctypes.data, 2)
    assert rc == nat.CPET_ERR_ARG
    assert lib.vecaddn(None, None, None, -1) == nat.CPET_ERR_ARG
    assert lib.einsum_ij_i(-1, 2, None, None) == nat.CPET_ERR_ARG
