"""Device-tensor calls on torch's default stream are ordered with torch's
own work: back-to-back requests whose input tensors are dropped (and their
memory reused by torch) right after each call, with no device-wide
synchronisation, return the same results as calls that synchronise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_default_stream_calls_without_sync(gpu_ctx):
    import torch

    from pycpet_amd import ops, workloads as W

    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    c = W.small_topo(400, 3000, cap=200, seed=3)
    v = W.volume_2k()
    want_t = ops.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q, ctx=gpu_ctx).feats
    want_v = ops.field_grid(*v.axes, v.x, v.q, ctx=gpu_ctx)[0]
    outs = []
    for k in range(12):  # no torch.cuda.synchronize() in this loop
        if k % 2 == 0:
            r = ops.topo(T(c.starts), T(c.n_iter), c.h, c.dims, T(c.x), T(c.q), ctx=gpu_ctx)
            outs.append(("t", r.feats.cpu().numpy()))
        else:
            E, _ = ops.field_grid(*(T(a) for a in v.axes), T(v.x), T(v.q), ctx=gpu_ctx)
            outs.append(("v", E.cpu().numpy()))
    print(f"{len(outs)} calls", flush=True)
    for kind, o in outs:
        np.testing.assert_array_equal(o, want_t if kind == "t" else want_v)
