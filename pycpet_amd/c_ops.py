"""Drop-in for the reference's ctypes wrapper class (CPET/utils/c_ops.py).

`Math_ops` keeps the reference's constructor (`shared_loc`) and the method
names, argument order and return shapes/dtypes of the hot-path methods; the
native calls go to libcpet_hip.so (gfx950) instead of math_module.so.  Unlike
the reference, every call checks the returned status and raises CpetError.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from . import ops


def _f32(a, last=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(-1, last) if last else a.reshape(-1)


class Math_ops:
    def __init__(self, shared_loc=None):
        # the reference passes the path of its own .so; this class always
        # binds libcpet_hip.so (shared_loc may name a different build of it)
        self.math = nat.load(shared_loc) if shared_loc and str(shared_loc).endswith(
            "libcpet_hip.so") else nat.load()

    def _check(self, rc):
        nat.check(rc, self.math)

    # CPET/utils/c_ops.py:250-263 -> math_module.c:405-451
    def compute_looped_field(self, x_0, x, Q):
        x_0 = _f32(x_0, 3)
        x = _f32(x, 3)
        Q = _f32(Q)
        res = np.zeros_like(x_0, dtype="float32")
        self._check(self.math.compute_looped_field(
            int(x_0.shape[0]), len(Q), x_0.ctypes.data, x.ctypes.data, Q.ctypes.data,
            res.ctypes.data))
        return res

    # CPET/utils/c_ops.py:326-341 -> math_module.c:453-486
    def calc_esp_base(self, x_0, x, Q):
        x_0 = _f32(x_0)
        x = _f32(x, 3)
        Q = _f32(Q)
        res = np.zeros(1, dtype="float32")
        self._check(self.math.calc_esp_base(res.ctypes.data, x_0.ctypes.data, len(Q),
                                            x.ctypes.data, Q.ctypes.data))
        return res

    # CPET/utils/c_ops.py:343-358 -> math_module.c:296-333
    def calc_field_base(self, x_0, x, Q):
        x_0 = _f32(x_0)
        x = _f32(x, 3)
        Q = _f32(Q)
        res = np.zeros(3, dtype="float32")
        self._check(self.math.calc_field_base(res.ctypes.data, x_0.ctypes.data, len(Q),
                                              x.ctypes.data, Q.ctypes.data))
        return res

    # CPET/utils/c_ops.py:360-375 -> math_module.c:255-293
    def calc_field(self, x_0, x, Q):
        x_0 = _f32(x_0)
        x = _f32(x, 3)
        Q = _f32(Q)
        res = np.zeros(3, dtype="float32")
        self._check(self.math.calc_field(res.ctypes.data, x_0.ctypes.data, len(Q),
                                         x.ctypes.data, Q.ctypes.data))
        return res

    # CPET/utils/c_ops.py:281-302 -> math_module.c:523-591
    def thread_operation(self, x_0, n_iter, x, Q, step_size, dimensions):
        x_0 = _f32(x_0)
        x = _f32(x, 3)
        Q = _f32(Q)
        dims = _f32(dimensions)
        res = np.zeros(2, dtype="float32")
        self._check(self.math.thread_operation(
            len(Q), int(n_iter), ctypes.c_float(float(step_size)), x_0.ctypes.data,
            dims.ctypes.data, x.ctypes.data, Q.ctypes.data, res.ctypes.data))
        return res

    def thread_operation_dipole(self, x_0, n_iter, x, mu, step_size, dimensions):
        """(2,) float32 [distance, mean curvature] of one streamline in the
        field of point dipoles (CPET/utils/c_ops.py:304-325, argument order
        kept)."""
        x_0 = _f32(x_0)
        x = _f32(x, 3)
        mu = _f32(mu, 3)
        dims = _f32(dimensions)
        res = np.zeros(2, dtype="float32")
        self._check(self.math.thread_operation_dipole(
            len(mu), int(n_iter), ctypes.c_float(float(step_size)), x_0.ctypes.data,
            dims.ctypes.data, x.ctypes.data, mu.ctypes.data, res.ctypes.data))
        return res

    # CPET/utils/c_ops.py:265-279 -> math_module.c:374-403
    def compute_batch_field(self, x_0, x, Q, batch_size):
        x_0 = _f32(x_0, 3)
        x = _f32(x, 3)
        Q = _f32(Q)
        res = np.zeros_like(x_0, dtype="float32")
        self._check(self.math.compute_batched_field(
            int(x_0.shape[0]), int(batch_size), len(Q), x_0.ctypes.data, x.ctypes.data,
            Q.ctypes.data, res.ctypes.data))
        return res

    # Host helpers (CPET/utils/c_ops.py:161-248 -> math_module.c:15-252).
    # The reference wrappers hand float32 buffers to the FP64 C arguments of
    # sparse_dot and dot; here the arrays are converted to the types the C
    # functions read, and the row results come back as float64.
    def sparse_dot(self, A, B):
        """CSR matrix A (scipy.sparse) times vector B: one value per row."""
        indptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
        indices = np.ascontiguousarray(A.indices, dtype=np.int32)
        data = np.ascontiguousarray(A.data, dtype=np.float64)
        b = np.ascontiguousarray(B, dtype=np.float64).ravel()
        res = np.zeros(max(len(indptr) - 1, 0), dtype=np.float64)
        self._check(self.math.sparse_dot(res.ctypes.data, indptr.ctypes.data, len(indptr),
                                         indices.ctypes.data, len(indices), data.ctypes.data,
                                         len(data), b.ctypes.data, len(b)))
        return res

    def dot(self, A, B):
        """Dense matrix A (rows, cols) times vector B (cols)."""
        a = np.ascontiguousarray(A, dtype=np.float64)
        a = a.reshape(a.shape[0], -1)
        b = np.ascontiguousarray(B, dtype=np.float64).ravel()
        if a.shape[1] != b.size:
            raise ValueError(f"A has {a.shape[1]} columns but B has {b.size} entries")
        res = np.zeros(a.shape[0], dtype=np.float64)
        self._check(self.math.dot(res.ctypes.data, a.ctypes.data, b.ctypes.data, a.shape[0],
                                  a.shape[1]))
        return res

    def vecaddn(self, A, B):
        a, b = _f32(A), _f32(B)
        if a.size != b.size:
            raise ValueError("A and B must have the same length")
        res = np.zeros(a.size, dtype="float32")
        self._check(self.math.vecaddn(res.ctypes.data, a.ctypes.data, b.ctypes.data, a.size))
        return res

    def einsum_ij_i(self, A):
        a = np.ascontiguousarray(A, dtype=np.float32)
        res = np.zeros((a.shape[0]), dtype="float32")
        self._check(self.math.einsum_ij_i(a.shape[0], a.shape[1], a.ctypes.data,
                                          res.ctypes.data))
        return res

    def einsum_ij_i_batch(self, A):
        a = np.ascontiguousarray(A, dtype=np.float32)
        res = np.zeros((a.shape[0], a.shape[1]), dtype="float32")
        self._check(self.math.einsum_ij_i_batch(a.shape[0], a.shape[1], a.shape[2],
                                                a.ctypes.data, res.ctypes.data))
        # the reference reshapes (not transposes) the (batch, rows) result
        return res.reshape(res.shape[1], res.shape[0])

    def einsum_operation(self, R, r_mag, Q):
        r_mag = _f32(r_mag)
        R = _f32(R, 3)
        Q = _f32(Q)
        if not (R.shape[0] == r_mag.size == Q.size):
            raise ValueError("R, r_mag and Q must describe the same charges")
        res = np.zeros(3, dtype="float32")
        self._check(self.math.einsum_operation(Q.size, r_mag.ctypes.data, Q.ctypes.data,
                                               R.ctypes.data, res.ctypes.data))
        return res

    def einsum_operation_batch(self, R, r_mag, Q, batch_size):
        Q = _f32(Q)
        r_mag = np.ascontiguousarray(r_mag, dtype=np.float32).reshape(batch_size, Q.size)
        R = np.ascontiguousarray(R, dtype=np.float32).reshape(batch_size, Q.size, 3)
        res = np.zeros((batch_size, 3), dtype="float32")
        self._check(self.math.einsum_operation_batch(batch_size, Q.size, r_mag.ctypes.data,
                                                     Q.ctypes.data, R.ctypes.data,
                                                     res.ctypes.data))
        return res

    # Batched forms (no per-point / per-streamline FFI loop).
    def esp_points(self, x_0, x, Q):
        """phi at every row of x_0 in one launch (float32)."""
        _, phi = ops.field_points(_f32(x_0, 3), x, Q, want_E=False, want_phi=True)
        return phi

    def thread_operation_batch(self, starts, n_iter, x, Q, step_size, dimensions):
        """thread_operation for every start in one launch; (L,2) float32."""
        r = ops.topo(starts, n_iter, step_size, dimensions, x, Q)
        return r.feats.astype(np.float32)
