"""cpet_topo_noalloc (the allocation-free streamline entry point) against
cpet_topo on the same device tensors: bitwise equal features, step counts,
stop kinds and histograms, on two calls in a row, the second smaller; a
call that needs more scratch than the fixed array is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(lib, stream, c, T, edges):
    import torch

    dev = torch.device("cuda:0")
    L = c.L
    feats = torch.empty((L, 2), dtype=torch.float64, device=dev)
    steps = torch.empty(L, dtype=torch.int32, device=dev)
    exited = torch.empty(L, dtype=torch.int32, device=dev)
    ed, ec = (torch.as_tensor(e, dtype=torch.float64, device=dev) for e in edges)
    hist = torch.empty((ed.numel() - 1, ec.numel() - 1), dtype=torch.int32, device=dev)
    st, it, x, q = T(c.starts), T(c.n_iter), T(c.x), T(c.q)
    dims = np.ascontiguousarray(c.dims, np.float32)
    rc = lib.cpet_topo_noalloc(stream, c.N, L, st.data_ptr(), it.data_ptr(),
                               ctypes.c_float(float(np.float32(c.h))), dims.ctypes.data,
                               x.data_ptr(), q.data_ptr(), feats.data_ptr(), steps.data_ptr(),
                               exited.data_ptr(), ed.numel() - 1, ed.data_ptr(), ec.numel() - 1,
                               ec.data_ptr(), hist.data_ptr(), 0)
    return rc, feats, steps, exited, hist


def test_noalloc_equals_cpet_topo(gpu_ctx):
    import torch

    from pycpet_amd import _native as nat, ops, workloads as W

    lib = nat.load()
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    cases = [W.small_topo(N=4000, L=20000, cap=300, seed=11, dims=(1.0, 1.0, 1.0), R=20.0),
             W.small_topo(N=700, L=3000, cap=200, seed=12),
             W.small_topo(N=12, L=60_000, cap=200, seed=15)]  # one-wave workgroups
    for c in cases:
        print(f"case noalloc {c.name}", flush=True)
        edges = (np.linspace(0.0, 2.0, 11), np.linspace(0.0, 40.0, 9))
        rc, feats, steps, exited, hist = _noalloc(lib, stream, c, T, edges)
        assert rc == 0, nat.last_error(lib)
        ref = ops.topo(T(c.starts), T(c.n_iter), c.h, c.dims, T(c.x), T(c.q), edges=edges,
                       ctx=gpu_ctx)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(feats.cpu().numpy(), ref.feats.cpu().numpy())
        np.testing.assert_array_equal(steps.cpu().numpy(), ref.steps.cpu().numpy())
        np.testing.assert_array_equal(exited.cpu().numpy(), ref.exited.cpu().numpy())
        np.testing.assert_array_equal(hist.cpu().numpy(), ref.hist.cpu().numpy())
        assert int(hist.sum()) > 0


def test_noalloc_refuses_oversized_calls(gpu_ctx):
    import torch

    from pycpet_amd import _native as nat, workloads as W

    lib = nat.load()
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    c = W.small_topo(N=64, L=700_000, cap=1, seed=13)  # ~42 MB of scratch
    rc, *_ = _noalloc(lib, torch.cuda.current_stream(dev).cuda_stream, c, T,
                      (np.linspace(0, 1, 3), np.linspace(0, 1, 3)))
    assert rc == nat.CPET_ERR_BUDGET
    assert "scratch" in nat.last_error(lib)


def test_noalloc_under_graph_capture(gpu_ctx):
    """The call captured in a HIP graph (CPET_F_NO_SYNC: it only queues
    kernels and allocates nothing) and replayed three times, with eager
    library calls between the replays: every replay equals the eager
    result bitwise."""
    import torch

    from pycpet_amd import _native as nat, ops, workloads as W

    lib = nat.load()
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    c = W.small_topo(N=900, L=5000, cap=200, seed=14)
    edges = (np.linspace(0.0, 1.5, 9), np.linspace(0.0, 30.0, 7))
    print(f"case noalloc graph {c.name}", flush=True)
    st, it, x, q = T(c.starts), T(c.n_iter), T(c.x), T(c.q)
    ed, ec = (torch.as_tensor(e, dtype=torch.float64, device=dev) for e in edges)
    feats = torch.empty((c.L, 2), dtype=torch.float64, device=dev)
    steps = torch.empty(c.L, dtype=torch.int32, device=dev)
    exited = torch.empty(c.L, dtype=torch.int32, device=dev)
    hist = torch.empty((ed.numel() - 1, ec.numel() - 1), dtype=torch.int32, device=dev)
    dims = np.ascontiguousarray(c.dims, np.float32)

    def call(stream, flags):
        return lib.cpet_topo_noalloc(stream, c.N, c.L, st.data_ptr(), it.data_ptr(),
                                     ctypes.c_float(float(np.float32(c.h))), dims.ctypes.data,
                                     x.data_ptr(), q.data_ptr(), feats.data_ptr(),
                                     steps.data_ptr(), exited.data_ptr(), ed.numel() - 1,
                                     ed.data_ptr(), ec.numel() - 1, ec.data_ptr(),
                                     hist.data_ptr(), flags)

    assert call(torch.cuda.current_stream(dev).cuda_stream, 0) == 0, nat.last_error(lib)
    want = [t.clone() for t in (feats, steps, exited, hist)]
    side = torch.cuda.Stream(dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        rc = call(side.cuda_stream, nat.F_NO_SYNC)
    assert rc == 0, nat.last_error(lib)
    for rep in range(3):
        for t in (feats, steps, exited, hist):
            t.zero_()
        other = ops.topo(c.starts[:64], c.n_iter[:64], c.h, c.dims, c.x, c.q, ctx=gpu_ctx)
        assert other.steps.shape == (64,)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for got, exp in zip((feats, steps, exited, hist), want):
            np.testing.assert_array_equal(got.cpu().numpy(), exp.cpu().numpy())
        print(f"replay {rep} equal", flush=True)
