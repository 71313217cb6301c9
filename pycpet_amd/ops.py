"""Array-level entry points over libcpet_hip.

Each function accepts either numpy arrays (host; the library stages them to
the device and copies results back) or torch tensors already on the GPU
(passed as device pointers and run on torch's current stream, so inputs stay
resident in HBM and results come back as device tensors).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np

from . import _native as nat


def _is_torch(a: Any) -> bool:
    return type(a).__module__.split(".")[0] == "torch"


def _host_f32(a, shape_last=None) -> np.ndarray:
    arr = np.ascontiguousarray(a, dtype=np.float32)
    if shape_last is not None:
        arr = arr.reshape(-1, shape_last)
    return arr


def _ptr(a) -> int:
    if _is_torch(a):
        return a.data_ptr()
    return a.ctypes.data


def dev_tensor(t, kind: str, name: str):
    """Checks a torch tensor handed to the library as a raw device pointer.

    kind "f32": coordinates, charges, axes, starts, moments — must be
    float32 (a float64 or float16 tensor would be read as the wrong type, or
    past its end), else ValueError.  kind "i32": step limits — any integer
    dtype, converted to int32 (an int64 tensor from np.asarray of Python
    ints is common); a floating dtype is a ValueError.  The tensor must live
    on a GPU and be contiguous.  Returns the tensor to pass (a converted
    copy for integer step limits of another width)."""
    import torch

    if kind == "f32":
        if t.dtype != torch.float32:
            raise ValueError(f"{name} must be a float32 tensor (got {t.dtype}); "
                             f"convert it with .float()")
    elif kind == "i32":
        if t.dtype.is_floating_point or t.dtype.is_complex or t.dtype == torch.bool:
            raise ValueError(f"{name} must be an integer tensor (got {t.dtype})")
    else:
        raise ValueError(f"unknown tensor kind {kind!r}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor (got one on {t.device})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if kind == "i32" and t.dtype != torch.int32:
        if t.numel() and (int(t.max()) > 2**31 - 1 or int(t.min()) < -2**31):
            raise ValueError(f"{name} holds values outside int32")
        t = t.to(torch.int32)
    return t


def _rows3(a, name: str) -> int:
    """Rows of an (n,3) array or tensor (a flat one of 3n values counts as
    n rows); anything else is a ValueError, so no kernel reads past it."""
    n = a.numel() if _is_torch(a) else a.size
    if n % 3 or (a.ndim == 2 and a.shape[1] != 3) or a.ndim > 2:
        raise ValueError(f"{name} must hold rows of 3 values (got shape {tuple(a.shape)})")
    return n // 3


def _torch_dev_args(ctx: nat.Context, *tensors):
    """Validates float32 device tensors (see dev_tensor) and points the
    context at torch's current stream on their device."""
    import torch

    for i, t in enumerate(tensors):
        if t is not None:
            dev_tensor(t, "f32", f"tensor argument {i}")
    dev = tensors[0].device
    for t in tensors:
        if t is not None and t.device != dev:
            raise ValueError("all tensor arguments must be on one device")
    ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    return dev


def _prec_flag(prec: str) -> int:
    try:
        return nat.PREC[prec]
    except KeyError as exc:
        raise ValueError(f"precision must be one of {sorted(nat.PREC)}") from exc


def field_grid(ax, ay, az, x, q, *, want_E: bool = True, want_phi: bool = False,
               soften: bool = True, prec: str = "auto", out_f64: bool = False,
               ctx: Optional[nat.Context] = None):
    """Field and/or potential on the mesh ax x ay x az ('ij' order, z fastest).

    Returns (E (M,3) or None, phi (M,) or None) with M = nx*ny*nz.
    """
    ctx = ctx or nat.default_context()
    flags = _prec_flag(prec) | (nat.F_OUT_F64 if out_f64 else 0)
    if _is_torch(x):
        import torch

        dev = _torch_dev_args(ctx, ax, ay, az, x, q)
        nx, ny, nz = ax.numel(), ay.numel(), az.numel()
        M = nx * ny * nz
        odt = torch.float64 if out_f64 else torch.float32
        E = torch.empty((M, 3), dtype=odt, device=dev) if want_E else None
        phi = torch.empty((M,), dtype=odt, device=dev) if want_phi else None
        flags |= nat.F_DEVICE | nat.F_NO_SYNC
        N = q.numel()
    else:
        ax, ay, az = (_host_f32(v).ravel() for v in (ax, ay, az))
        x = _host_f32(x, 3)
        q = _host_f32(q).ravel()
        nx, ny, nz = ax.size, ay.size, az.size
        M = nx * ny * nz
        odt = np.float64 if out_f64 else np.float32
        E = np.empty((M, 3), dtype=odt) if want_E else None
        phi = np.empty((M,), dtype=odt) if want_phi else None
        N = q.size
    if _rows3(x, "x") != N:
        raise ValueError(f"x has {_rows3(x, 'x')} rows but q has {N} charges")
    nat.check(ctx.lib.cpet_field_grid(
        ctx.handle, nx, ny, nz, _ptr(ax), _ptr(ay), _ptr(az), N, _ptr(x), _ptr(q),
        _ptr(E) if E is not None else None, _ptr(phi) if phi is not None else None,
        1 if soften else 0, flags), ctx.lib)
    return E, phi


def field_points(pts, x, q, *, want_E: bool = True, want_phi: bool = False,
                 soften: bool = True, prec: str = "auto", out_f64: bool = False,
                 ctx: Optional[nat.Context] = None):
    """Field and/or potential at arbitrary points pts (M,3)."""
    ctx = ctx or nat.default_context()
    flags = _prec_flag(prec) | (nat.F_OUT_F64 if out_f64 else 0)
    if _is_torch(x):
        import torch

        dev = _torch_dev_args(ctx, pts, x, q)
        M = _rows3(pts, "pts")
        odt = torch.float64 if out_f64 else torch.float32
        E = torch.empty((M, 3), dtype=odt, device=dev) if want_E else None
        phi = torch.empty((M,), dtype=odt, device=dev) if want_phi else None
        flags |= nat.F_DEVICE | nat.F_NO_SYNC
        N = q.numel()
    else:
        pts = _host_f32(pts, 3)
        x = _host_f32(x, 3)
        q = _host_f32(q).ravel()
        M = pts.shape[0]
        odt = np.float64 if out_f64 else np.float32
        E = np.empty((M, 3), dtype=odt) if want_E else None
        phi = np.empty((M,), dtype=odt) if want_phi else None
        N = q.size
    if _rows3(x, "x") != N:
        raise ValueError(f"x has {_rows3(x, 'x')} rows but q has {N} charges")
    nat.check(ctx.lib.cpet_field_points(
        ctx.handle, M, N, _ptr(pts), _ptr(x), _ptr(q),
        _ptr(E) if E is not None else None, _ptr(phi) if phi is not None else None,
        1 if soften else 0, flags), ctx.lib)
    return E, phi


@dataclass
class TopoResult:
    feats: Any            # (L,2) float64: distance, mean curvature (input order)
    steps: Any            # (L,) int32 steps taken (Euler or RK4)
    exited: Any           # (L,) int32: 1 = left the box, 0 = hit n_iter
    hist: Any = None      # (nb_d, nb_c) uint32 or None


def topo(starts, n_iter, h: float, dims, x, q, *, edges=None, integrator: str = "euler",
         prec: str = "auto", ctx: Optional[nat.Context] = None) -> TopoResult:
    """Streamlines from every start (thread_operation semantics).

    edges: optional (edges_d, edges_c) increasing float64 bin edges; the
    histogram counts the float32-rounded features with numpy.histogramdd
    rules.  dims is read on the host in every mode (pass a numpy array, or
    a tensor, which is copied to the host).  integrator: "euler" (the
    reference's) or "rk4" (classical RK4 on E/|E| for the walk and the
    curvature steps, an extension).  prec: "auto" (FP32 tiled field sums,
    FP64 positions) or "f64" (FP64 field sums, a precision reference).
    """
    ctx = ctx or nat.default_context()
    if integrator not in ("euler", "rk4"):
        raise ValueError(f"unknown integrator {integrator!r}")
    if prec not in ("auto", "f64"):
        raise ValueError(f"streamline precision must be 'auto' or 'f64' (got {prec!r})")
    flags = nat.F_TOPO_RK4 if integrator == "rk4" else 0
    if prec == "f64":
        flags |= nat.PREC["f64"]
    h = float(np.float32(h))
    if _is_torch(dims):
        dims = dims.detach().cpu().numpy()
    dims = _host_f32(dims).ravel()
    if dims.size != 3:
        raise ValueError("dims must hold 3 half-widths")
    if _is_torch(x):
        import torch

        dev = _torch_dev_args(ctx, starts, x, q)
        n_iter = dev_tensor(n_iter, "i32", "n_iter").reshape(-1)
        L = _rows3(starts, "starts")
        feats = torch.empty((L, 2), dtype=torch.float64, device=dev)
        steps = torch.empty((L,), dtype=torch.int32, device=dev)
        exited = torch.empty((L,), dtype=torch.int32, device=dev)
        hist = None
        ed = ec = None
        if edges is not None:
            ed = torch.as_tensor(edges[0], dtype=torch.float64, device=dev).contiguous()
            ec = torch.as_tensor(edges[1], dtype=torch.float64, device=dev).contiguous()
            hist = torch.empty((ed.numel() - 1, ec.numel() - 1), dtype=torch.int32, device=dev)
        flags |= nat.F_DEVICE | nat.F_NO_SYNC
        N = q.numel()
    else:
        starts = _host_f32(starts, 3)
        n_iter = np.ascontiguousarray(n_iter, dtype=np.int32).ravel()
        x = _host_f32(x, 3)
        q = _host_f32(q).ravel()
        L = starts.shape[0]
        feats = np.empty((L, 2), dtype=np.float64)
        steps = np.empty((L,), dtype=np.int32)
        exited = np.empty((L,), dtype=np.int32)
        hist = None
        ed = ec = None
        if edges is not None:
            ed = np.ascontiguousarray(edges[0], dtype=np.float64).ravel()
            ec = np.ascontiguousarray(edges[1], dtype=np.float64).ravel()
            hist = np.empty((ed.size - 1, ec.size - 1), dtype=np.uint32)
        N = q.size
    if n_iter.shape[0] != L:
        raise ValueError("n_iter must have one entry per start point")
    if _rows3(x, "x") != N:
        raise ValueError(f"x has {_rows3(x, 'x')} rows but q has {N} charges")
    nb_d = (ed.shape[0] - 1) if ed is not None else 0
    nb_c = (ec.shape[0] - 1) if ec is not None else 0
    nat.check(ctx.lib.cpet_topo(
        ctx.handle, N, L, _ptr(starts), _ptr(n_iter), ctypes.c_float(h), _ptr(dims),
        _ptr(x), _ptr(q), _ptr(feats), _ptr(steps), _ptr(exited),
        nb_d, _ptr(ed) if ed is not None else None,
        nb_c, _ptr(ec) if ec is not None else None,
        _ptr(hist) if hist is not None else None, flags), ctx.lib)
    return TopoResult(feats, steps, exited, hist)


def topo_dipole(starts, n_iter, h: float, dims, x, mu, *, physical: bool = False,
                ctx: Optional[nat.Context] = None) -> TopoResult:
    """Streamlines in the field of point dipoles x (N,3) with moments mu
    (N,3): thread_operation_dipole semantics (CPET/utils/math_module.c:
    593-659) with the reference's pair term, or the physical one."""
    ctx = ctx or nat.default_context()
    flags = nat.F_DIPOLE_PHYSICAL if physical else 0
    h = float(np.float32(h))
    if _is_torch(dims):
        dims = dims.detach().cpu().numpy()
    dims = _host_f32(dims).ravel()
    if _is_torch(x):
        import torch

        dev = _torch_dev_args(ctx, starts, x, mu)
        n_iter = dev_tensor(n_iter, "i32", "n_iter").reshape(-1)
        L = _rows3(starts, "starts")
        feats = torch.empty((L, 2), dtype=torch.float64, device=dev)
        steps = torch.empty((L,), dtype=torch.int32, device=dev)
        exited = torch.empty((L,), dtype=torch.int32, device=dev)
        flags |= nat.F_DEVICE | nat.F_NO_SYNC
        N = _rows3(x, "x")
    else:
        starts = _host_f32(starts, 3)
        n_iter = np.ascontiguousarray(n_iter, dtype=np.int32).ravel()
        x = _host_f32(x, 3)
        mu = _host_f32(mu, 3)
        L = starts.shape[0]
        feats = np.empty((L, 2), dtype=np.float64)
        steps = np.empty((L,), dtype=np.int32)
        exited = np.empty((L,), dtype=np.int32)
        N = x.shape[0]
    if n_iter.shape[0] != L:
        raise ValueError("n_iter must have one entry per start point")
    if _rows3(mu, "mu") != N:
        raise ValueError("mu must be (N, 3) like x")
    nat.check(ctx.lib.cpet_topo_dipole(
        ctx.handle, N, L, _ptr(starts), _ptr(n_iter), ctypes.c_float(h), _ptr(dims), _ptr(x),
        _ptr(mu), _ptr(feats), _ptr(steps), _ptr(exited), flags), ctx.lib)
    return TopoResult(feats, steps, exited, None)
