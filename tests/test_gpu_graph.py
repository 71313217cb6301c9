"""HIP-graph captured steps (pycpet_amd.graphs): each replay equals an eager
run on the same inputs bitwise, with eager work and synchronisations between
replays, and after the inputs are replaced in place."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_captured_field_grid_replays(gpu_ctx):
    import torch

    from pycpet_amd import graphs, ops, workloads as W

    c = W.volume_2k()
    ax = [_dev(a) for a in c.axes]
    x, q = _dev(c.x), _dev(c.q)
    step = graphs.CapturedStep(lambda ctx: ops.field_grid(*ax, x, q, want_phi=True, ctx=ctx))
    rng = np.random.default_rng(0)
    for k in range(4):
        if k >= 2:  # new coordinates, same shapes
            x.copy_(_dev(c.x + rng.normal(0, 0.05, c.x.shape).astype(np.float32)))
        E, phi = step.replay()
        torch.cuda.synchronize()
        Ee, pe = ops.field_grid(*ax, x, q, want_phi=True, ctx=gpu_ctx)  # eager, other context
        torch.cuda.synchronize()
        print(f"replay {k}", flush=True)
        assert torch.equal(E, Ee) and torch.equal(phi, pe)
    step.close()


def test_captured_topo_replays(gpu_ctx):
    import torch

    from pycpet_amd import graphs, ops, workloads as W

    c = W.small_topo(300, 512, seed=5)
    st, it, x, q = _dev(c.starts), _dev(c.n_iter), _dev(c.x), _dev(c.q)
    ed = np.linspace(0, 1.0, 11)
    ec = np.linspace(0, 50.0, 9)
    ed_t, ec_t = _dev(ed), _dev(ec)  # device edges: no copy inside the capture
    step = graphs.CapturedStep(
        lambda ctx: ops.topo(st, it, c.h, c.dims, x, q, edges=(ed_t, ec_t), ctx=ctx))
    for k in range(3):
        if k == 2:
            st.copy_(_dev(c.starts[::-1].copy()))
        r = step.replay()
        torch.cuda.synchronize()
        e = ops.topo(st, it, c.h, c.dims, x, q, edges=(ed, ec), ctx=gpu_ctx)
        torch.cuda.synchronize()
        print(f"replay {k}", flush=True)
        assert torch.equal(r.feats, e.feats) and torch.equal(r.steps, e.steps)
        assert torch.equal(r.hist, e.hist)
    step.close()


def test_captured_batched_step(gpu_ctx):
    """Many frames per launch inside one graph: the batched volume and
    streamline calls of a mixed frame set."""
    import torch

    from pycpet_amd import batch, graphs, workloads as W

    frames = W.md_frames(6, seed=9, n_min=500, n_max=3000, edges=(9, 13),
                         topo_starts_per_frame=700, cap=150)
    dev = []
    for f in frames:
        g = {k: (_dev(v) if isinstance(v, np.ndarray) and k != "dims" else v)
             for k, v in f.items() if k != "axes"}
        if "axes" in f:
            g["axes"] = tuple(_dev(a) for a in f["axes"])
        dev.append(g)
    vol = [f for f in dev if f["kind"] == "volume"]
    top = [f for f in dev if f["kind"] == "topo"]

    def step_fn(ctx):
        return batch.batch_volume(vol, want_phi=True, ctx=ctx), batch.batch_topo(top, ctx=ctx)

    step = graphs.CapturedStep(step_fn)
    for k in range(3):
        if k == 2:
            vol[0]["x"].add_(0.01)
            top[1]["x"].add_(-0.01)
        outs_v, outs_t = step.replay()
        torch.cuda.synchronize()
        ev = batch.batch_volume(vol, want_phi=True, ctx=gpu_ctx)
        et = batch.batch_topo(top, ctx=gpu_ctx)
        torch.cuda.synchronize()
        print(f"replay {k}", flush=True)
        for (E, p), (E2, p2) in zip(outs_v, ev):
            assert torch.equal(E, E2) and torch.equal(p, p2)
        for r, r2 in zip(outs_t, et):
            assert torch.equal(r.feats, r2.feats) and torch.equal(r.steps, r2.steps)
    step.close()


def test_parallel_branches(gpu_ctx):
    """A list of step functions becomes parallel branches of one graph; each
    branch's outputs equal its eager run."""
    import torch

    from pycpet_amd import batch, graphs, workloads as W

    frames = W.md_frames(4, seed=12, n_min=500, n_max=2000, edges=(9,),
                         topo_starts_per_frame=500, cap=120)
    vol = [{"x": _dev(f["x"]), "q": _dev(f["q"]), "axes": tuple(_dev(a) for a in f["axes"])}
           for f in frames if f["kind"] == "volume"]
    top = [{"x": _dev(f["x"]), "q": _dev(f["q"]), "starts": _dev(f["starts"]),
            "n_iter": _dev(f["n_iter"]), "h": f["h"], "dims": f["dims"]}
           for f in frames if f["kind"] == "topo"]
    step = graphs.CapturedStep([lambda c: batch.batch_volume(vol, ctx=c),
                                lambda c: batch.batch_topo(top, ctx=c)])
    for k in range(3):
        ov, ot = step.replay()
        torch.cuda.synchronize()
        ev = batch.batch_volume(vol, ctx=gpu_ctx)
        et = batch.batch_topo(top, ctx=gpu_ctx)
        torch.cuda.synchronize()
        print(f"replay {k}", flush=True)
        assert all(torch.equal(a[0], b[0]) for a, b in zip(ov, ev))
        assert all(torch.equal(a.feats, b.feats) for a, b in zip(ot, et))
    step.close()
