"""GPU frame preparation: residue radius filter, box-frame transform,
in-box filter and stable compaction (csrc/frame.hip).

Each step's arithmetic type follows numpy's promotion of the operands the
reference combines (CPET/source/calculator.py:238-371), e.g. float32
coordinates minus a float64 centre are subtracted in float64, so the
decisions match numpy's on identical inputs.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _native as nat


def _code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float32:
        return 0
    if dt.kind in "fiu":
        return 1
    raise TypeError(f"unsupported dtype {dt}")


def _as_store(x) -> np.ndarray:
    x = np.ascontiguousarray(x)
    if x.dtype not in (np.float32, np.float64):
        x = x.astype(np.float64)
    return x.reshape(-1, 3)


def _d3(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def radius_residue_keep(x, center, radius: float, runs,
                        ctx: Optional[nat.Context] = None) -> np.ndarray:
    """Keep-mask of filter_radius_whole_residue (CPET/utils/io.py:283-338)."""
    ctx = ctx or nat.default_context()
    xs = _as_store(x)
    c = np.asarray(center)
    sub = np.result_type(xs.dtype, c.dtype)
    runs = np.ascontiguousarray(runs, dtype=np.int32)
    n_runs = int(runs.max()) + 1 if runs.size else 0
    keep = np.empty(xs.shape[0], dtype=np.uint8)
    cd = _d3(c.astype(sub))
    # the reference compares against a Python float: in the norm's own type
    r = float(np.asarray(radius, dtype=sub))
    nat.check(ctx.lib.cpet_filter_radius_residue(
        ctx.handle, xs.shape[0], xs.ctypes.data, _code(xs.dtype), cd.ctypes.data, _code(sub),
        ctypes.c_double(r), runs.ctypes.data, n_runs, keep.ctypes.data, 0), ctx.lib)
    return keep.astype(bool)


def frame_transform(x, center, Tinv=None, box_shift=None,
                    ctx: Optional[nat.Context] = None) -> np.ndarray:
    """(x - center) @ Tinv - box_shift with numpy's dtype at each step."""
    ctx = ctx or nat.default_context()
    xs = _as_store(x)
    c = np.asarray(center)
    sub = np.result_type(xs.dtype, c.dtype)
    dot = sub
    T = None
    if Tinv is not None:
        T = np.asarray(Tinv)
        dot = np.result_type(sub, T.dtype)
    shift = dot
    s = None
    if box_shift is not None:
        s = np.asarray(box_shift)
        shift = np.result_type(dot, s.dtype)
    out = np.empty(xs.shape, dtype=shift if shift in (np.float32, np.float64) else np.float64)
    cd = _d3(c.astype(sub))
    Td = np.ascontiguousarray(T.astype(dot).astype(np.float64).reshape(9)) if T is not None else None
    sd = _d3(s.astype(shift)) if s is not None else None
    nat.check(ctx.lib.cpet_frame_transform(
        ctx.handle, xs.shape[0], xs.ctypes.data, _code(xs.dtype), cd.ctypes.data, _code(sub),
        Td.ctypes.data if Td is not None else None, _code(dot),
        sd.ctypes.data if sd is not None else None, _code(shift),
        out.ctypes.data, _code(out.dtype), 0), ctx.lib)
    return out


def in_box_keep(x, center, dims, ctx: Optional[nat.Context] = None) -> np.ndarray:
    """Keep-mask of filter_in_box (CPET/utils/io.py:422-475)."""
    ctx = ctx or nat.default_context()
    xs = _as_store(x)
    c = np.asarray(center)
    sub = np.result_type(xs.dtype, c.dtype)
    keep = np.empty(xs.shape[0], dtype=np.uint8)
    cd = _d3(c.astype(sub))
    dd = _d3(dims)
    nat.check(ctx.lib.cpet_filter_in_box(
        ctx.handle, xs.shape[0], xs.ctypes.data, _code(xs.dtype), cd.ctypes.data, _code(sub),
        dd.ctypes.data, keep.ctypes.data, 0), ctx.lib)
    return keep.astype(bool)


def compact(keep, x, q=None, ctx: Optional[nat.Context] = None):
    """Stable compaction of kept rows: (x_out, q_out, original indices)."""
    ctx = ctx or nat.default_context()
    k = np.ascontiguousarray(keep, dtype=np.uint8)
    xs = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    qs = np.ascontiguousarray(q, dtype=np.float32).ravel() if q is not None else None
    N = xs.shape[0]
    xo = np.empty((N, 3), np.float32)
    qo = np.empty(N, np.float32) if qs is not None else None
    io = np.empty(N, np.int32)
    cnt = ctypes.c_int(0)
    nat.check(ctx.lib.cpet_compact_charges(
        ctx.handle, N, k.ctypes.data, xs.ctypes.data, qs.ctypes.data if qs is not None else None,
        xo.ctypes.data, qo.ctypes.data if qo is not None else None, io.ctypes.data,
        ctypes.byref(cnt), 0), ctx.lib)
    n = cnt.value
    return xo[:n], (qo[:n] if qo is not None else None), io[:n]
