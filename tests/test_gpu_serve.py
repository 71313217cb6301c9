"""Serving with continuous batching: every request's result is the same when
served alone (per-frame entry point) and in the mix (batched with other
requests on several overlapping streams): bitwise for streamlines, within
the field tolerance (1e-5 pointwise, 1e-3 RMS floor) for grids, whose batch
kernel does not split the charge sum across workgroups."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, float).reshape(len(b), -1)
    b = np.asarray(b, float).reshape(len(b), -1)
    nb = np.linalg.norm(b, axis=1)
    d = np.linalg.norm(a - b, axis=1)
    return float((d / np.maximum(nb, 1e-3 * np.sqrt(np.mean(nb ** 2)))).max())


@pytest.mark.parametrize("policy,small_work", [("fifo-small", 2e9), ("greedy", 2e9),
                                               ("fifo-small", 0.0)])
def test_alone_equals_in_the_mix(gpu_ctx, policy, small_work):
    """fifo-small with the default threshold coalesces these small requests;
    with small_work = 0 every request is large and runs alone, one at a time."""
    from pycpet_amd import ops, serve

    reqs = serve.poisson_requests(rate=400.0, duration=0.2, seed=3, n_range=(500, 6000),
                                  edge_range=(9, 31), starts_range=(200, 3000), max_requests=40)
    assert {r.kind for r in reqs} == {"volume", "volume_ESP", "topo"}
    srv = serve.Server(n_lanes=3, hbm_budget=1 << 30, max_batch=8, policy=policy,
                       small_work=small_work)
    stats = srv.run(reqs)
    assert stats.completed == len(reqs) and stats.rejected == 0
    if small_work > 0:
        assert stats.batches < len(reqs)  # requests were batched together
    else:
        assert stats.batches == len(reqs)  # one large request per launch
    for r in reqs:
        d = r.data
        if r.kind == "topo":
            alone = ops.topo(d["starts"], d["n_iter"], d["h"], d["dims"], d["x"], d["q"],
                             ctx=gpu_ctx).feats
            np.testing.assert_array_equal(r.result, alone)
        else:
            E, phi = ops.field_grid(*d["axes"], d["x"], d["q"], want_E=r.kind == "volume",
                                    want_phi=r.kind == "volume_ESP", ctx=gpu_ctx)
            alone = E if r.kind == "volume" else phi
            assert _rel(r.result, alone) <= 1e-5
    s = stats.summary()
    assert s["p99_ms"] >= s["p50_ms"] > 0


def test_budget_rejects_oversized_request(gpu_ctx):
    from pycpet_amd import serve

    reqs = serve.poisson_requests(rate=100.0, duration=0.05, seed=4, n_range=(500, 800),
                                  edge_range=(9, 11), starts_range=(100, 200), max_requests=3)
    srv = serve.Server(n_lanes=1, hbm_budget=min(r.nbytes() for r in reqs) - 1)
    stats = srv.run(reqs)
    assert stats.rejected == len(reqs) and stats.completed == 0
    assert all(r.error for r in reqs)
