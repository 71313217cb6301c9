"""The serving admission policy on the host, without a GPU: which queued
requests a free lane takes under "fifo-small" and "greedy", the HBM
budget, and the work estimate that sorts requests into small and large."""
import collections

import numpy as np
import pytest

from pycpet_amd import serve


class _FakeLane:
    def __init__(self):
        self.groups = collections.deque()

    @property
    def busy(self):
        return bool(self.groups)


def _server(policy="fifo-small", n_lanes=3, budget=1 << 40, max_batch=8, small_work=1e6):
    s = object.__new__(serve.Server)
    s.lanes = [_FakeLane() for _ in range(n_lanes)]
    s.budget = budget
    s.max_batch = max_batch
    s.inflight = 0
    s.policy = policy
    s.small_work = small_work
    return s


def _vol(rid, edge, n):
    ax = np.linspace(-1, 1, edge).astype(np.float32)
    data = {"x": np.zeros((n, 3), np.float32), "q": np.zeros(n, np.float32), "axes": (ax, ax, ax)}
    return serve.Request(rid, "volume", data)


def _topo(rid, L, n, steps):
    data = {"x": np.zeros((n, 3), np.float32), "q": np.zeros(n, np.float32),
            "starts": np.zeros((L, 3), np.float32), "n_iter": np.full(L, steps, np.int32),
            "h": 0.1, "dims": np.ones(3, np.float32)}
    return serve.Request(rid, "topo", data)


def test_work_estimate():
    assert _vol(0, 11, 100).work() == 100 * 11 ** 3
    # streamlines: every start runs its n_iter steps, plus the 4 epilogue steps
    assert _topo(1, 50, 20, 7).work() == 20 * (50 * 7 + 4 * 50)


def test_fifo_small_routes_large_to_lane0_in_order():
    s = _server()
    small = collections.deque([_vol(i, 5, 10) for i in range(3)])      # 1,250 pairs each
    large = collections.deque([_vol(10 + i, 41, 1000) for i in range(3)])  # 6.9e7 pairs
    # lane 0 takes the oldest large request, and a second one queued behind it
    take, is_large = s._take(0, small, large)
    assert is_large and [r.rid for r in take] == [10]
    s.lanes[0].groups.append("running")
    take, is_large = s._take(0, small, large)
    assert is_large and [r.rid for r in take] == [11]
    s.lanes[0].groups.append("queued")
    # lane 0 holds two: it takes nothing more
    assert s._take(0, small, large) == ([], False)
    # the other lanes coalesce the small requests
    take, is_large = s._take(1, small, large)
    assert not is_large and [r.rid for r in take] == [0, 1, 2]
    assert [r.rid for r in large] == [12] and not small


def test_fifo_small_single_lane_and_max_batch():
    s = _server(n_lanes=1, max_batch=2)
    small = collections.deque([_vol(i, 5, 10) for i in range(3)])
    large = collections.deque([_vol(9, 41, 1000)])
    take, is_large = s._take(0, small, large)
    assert is_large and take[0].rid == 9
    s.lanes[0].groups.append("running")
    assert s._take(0, small, large) == ([], False)  # busy lane takes nothing
    s.lanes[0].groups.clear()
    take, _ = s._take(0, small, large)
    assert [r.rid for r in take] == [0, 1]  # max_batch


def test_greedy_and_budget():
    s = _server(policy="greedy")
    small = collections.deque([_vol(0, 41, 1000), _vol(1, 5, 10), _vol(2, 5, 10)])
    take, is_large = s._take(0, small, collections.deque())
    assert not is_large and [r.rid for r in take] == [0, 1, 2]
    # the budget stops a batch before the request that does not fit
    reqs = [_vol(i, 5, 10) for i in range(3)]
    s = _server(policy="greedy", budget=2 * reqs[0].nbytes())
    q = collections.deque(reqs)
    take, _ = s._take(0, q, collections.deque())
    assert [r.rid for r in take] == [0, 1] and [r.rid for r in q] == [2]


def test_unknown_policy_rejected():
    with pytest.raises(ValueError):
        serve.Server.__init__(object.__new__(serve.Server), policy="lifo")
