"""Many frames per launch (csrc/batch.hip).

    batch_volume(frames)  -> [(E, phi), ...]   one launch for all grids
    batch_topo(frames)    -> [TopoResult, ...]  one persistent launch

A frame is a dict: volume frames {"x", "q", "axes": (ax, ay, az)}, topo
frames {"x", "q", "starts", "n_iter", "h", "dims"}.  Arrays may be numpy
(host; staged by the library) or torch device tensors (kept in HBM, the
call is queued on torch's current stream).
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np

from . import _native as nat
from .ops import TopoResult, _is_torch, _prec_flag, _ptr, dev_tensor


def _cat(parts, dtype, torch_mode, name):
    """Concatenates the frames' arrays flat.  Device tensors are checked
    like the single-frame calls' (float32 data; integer step limits become
    int32), so none is read as another type."""
    if torch_mode:
        import torch

        kind = "i32" if dtype == np.int32 else "f32"
        for p in parts:
            if not _is_torch(p):
                raise ValueError(f"{name}: mix of numpy arrays and torch tensors")
        parts = [dev_tensor(p.contiguous(), kind, name) for p in parts]
        if len({p.device for p in parts}) != 1:
            raise ValueError(f"{name}: frames on different devices")
        return torch.cat([p.reshape(-1) for p in parts]).contiguous()
    return np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=dtype).reshape(-1)
                                                for p in parts]))


def batch_volume(frames: Sequence[dict], *, want_E: bool = True, want_phi: bool = False,
                 soften: bool = True, prec: str = "auto", out_f64: bool = False,
                 ctx: Optional[nat.Context] = None) -> List[tuple]:
    ctx = ctx or nat.default_context()
    F = len(frames)
    tm = _is_torch(frames[0]["x"])
    nch = np.array([int(np.prod(f["q"].shape)) for f in frames], dtype=np.int32)
    dims = np.array([[len(a) for a in f["axes"]] for f in frames], dtype=np.int32).reshape(-1)
    M = [int(np.prod(dims[3 * i:3 * i + 3])) for i in range(F)]
    x = _cat([f["x"] for f in frames], np.float32, tm, "x")
    q = _cat([f["q"] for f in frames], np.float32, tm, "q")
    axes = _cat([a for f in frames for a in f["axes"]], np.float32, tm, "axes")
    if x.shape[0] != 3 * int(nch.sum()):
        raise ValueError("x must hold 3 coordinates per charge of every frame")
    flags = _prec_flag(prec) | (nat.F_OUT_F64 if out_f64 else 0)
    tot = sum(M)
    if tm:
        import torch

        dev = x.device
        ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        odt = torch.float64 if out_f64 else torch.float32
        E = torch.empty((tot, 3), dtype=odt, device=dev) if want_E else None
        phi = torch.empty((tot,), dtype=odt, device=dev) if want_phi else None
        flags |= nat.F_DEVICE | nat.F_NO_SYNC
    else:
        odt = np.float64 if out_f64 else np.float32
        E = np.empty((tot, 3), dtype=odt) if want_E else None
        phi = np.empty((tot,), dtype=odt) if want_phi else None
    nat.check(ctx.lib.cpet_batch_volume(
        ctx.handle, F, nch.ctypes.data, _ptr(x), _ptr(q), dims.ctypes.data, _ptr(axes),
        _ptr(E) if E is not None else None, _ptr(phi) if phi is not None else None,
        1 if soften else 0, flags), ctx.lib)
    out, o = [], 0
    for m in M:
        out.append((E[o:o + m] if E is not None else None, phi[o:o + m] if phi is not None else None))
        o += m
    return out


def batch_topo(frames: Sequence[dict], *, edges=None,
               ctx: Optional[nat.Context] = None) -> List[TopoResult]:
    ctx = ctx or nat.default_context()
    F = len(frames)
    tm = _is_torch(frames[0]["x"])
    nch = np.array([int(np.prod(f["q"].shape)) for f in frames], dtype=np.int32)
    ns = np.array([int(f["starts"].shape[0]) for f in frames], dtype=np.int32)
    h = np.array([float(np.float32(f["h"])) for f in frames], dtype=np.float32)
    dims = np.ascontiguousarray(np.array([np.asarray(f["dims"], dtype=np.float32).reshape(3)
                                          if not _is_torch(f["dims"]) else
                                          f["dims"].detach().cpu().numpy().reshape(3)
                                          for f in frames], dtype=np.float32).reshape(-1))
    x = _cat([f["x"] for f in frames], np.float32, tm, "x")
    q = _cat([f["q"] for f in frames], np.float32, tm, "q")
    starts = _cat([f["starts"] for f in frames], np.float32, tm, "starts")
    n_iter = _cat([f["n_iter"] for f in frames], np.int32, tm, "n_iter")
    Lt = int(ns.sum())
    if x.shape[0] != 3 * int(nch.sum()) or starts.shape[0] != 3 * Lt or n_iter.shape[0] != Lt:
        raise ValueError("frame arrays do not match their charge / start counts")
    flags = 0
    if tm:
        import torch

        dev = x.device
        ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        feats = torch.empty((Lt, 2), dtype=torch.float64, device=dev)
        steps = torch.empty((Lt,), dtype=torch.int32, device=dev)
        exited = torch.empty((Lt,), dtype=torch.int32, device=dev)
        flags |= nat.F_DEVICE | nat.F_NO_SYNC
        ed = ec = hist = None
        if edges is not None:
            ed = torch.as_tensor(edges[0], dtype=torch.float64, device=dev).contiguous()
            ec = torch.as_tensor(edges[1], dtype=torch.float64, device=dev).contiguous()
            hist = torch.empty((F, ed.numel() - 1, ec.numel() - 1), dtype=torch.int32, device=dev)
    else:
        feats = np.empty((Lt, 2), np.float64)
        steps = np.empty(Lt, np.int32)
        exited = np.empty(Lt, np.int32)
        ed = ec = hist = None
        if edges is not None:
            ed = np.ascontiguousarray(edges[0], dtype=np.float64)
            ec = np.ascontiguousarray(edges[1], dtype=np.float64)
            hist = np.empty((F, ed.size - 1, ec.size - 1), np.uint32)
    nb_d = (ed.shape[0] - 1) if ed is not None else 0
    nb_c = (ec.shape[0] - 1) if ec is not None else 0
    nat.check(ctx.lib.cpet_batch_topo(
        ctx.handle, F, nch.ctypes.data, _ptr(x), _ptr(q), ns.ctypes.data, _ptr(starts),
        _ptr(n_iter), h.ctypes.data, dims.ctypes.data, _ptr(feats), _ptr(steps), _ptr(exited),
        nb_d, _ptr(ed) if ed is not None else None, nb_c, _ptr(ec) if ec is not None else None,
        _ptr(hist) if hist is not None else None, flags), ctx.lib)
    out, o = [], 0
    for i, n in enumerate(ns):
        out.append(TopoResult(feats[o:o + n], steps[o:o + n], exited[o:o + n],
                              hist[i] if hist is not None else None))
        o += int(n)
    return out
