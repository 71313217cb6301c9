"""Request serving with continuous batching on one GPU.

Requests (volume / volume_ESP / topo frames) arrive over time.  The server
keeps `n_lanes` HIP streams (each with its own libcpet_hip context).  A lane
that is free takes admitted requests that fit the HBM budget, copies their
inputs to the device, and launches one batch call per request kind
(cpet_batch_volume / cpet_batch_topo), each followed by asynchronous copies
of its results into pinned host memory and an event: a kind's requests
complete when their own event fires.  Latency is measured from a request's
arrival to its results being in host memory.

Policy "fifo-small" (the default) sorts requests by estimated work:
  * a large request (>= small_work pair evaluations) fills the GPU by
    itself.  Large requests run in arrival order on lane 0 through the
    per-frame entry points (cpet_field_grid / cpet_topo, whose charge splits
    suit one big frame), one at a time on the device; the next one is
    enqueued while the current one runs, so its input copies overlap.
    Batching them would only make the first of a batch wait for the rest;
  * small requests cannot fill the GPU alone: a free lane among the others
    coalesces every queued small request (up to max_batch) into one batch
    call per kind, and these run beside the large request in flight instead
    of queueing behind it.
Policy "greedy" lets every free lane take every queued request (up to
max_batch) into batch calls, whatever its size.

The "plain" mode serves the same arrivals one request at a time on one
stream through the per-frame entry points (cpet_field_grid / cpet_topo),
which is the baseline the batched form is compared against.

No counterpart in the reference (one process handles one frame at a time,
CPET/source/CPET.py:96-229).
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _native as nat
from . import batch as B
from . import ops


@dataclass
class Request:
    rid: int
    kind: str                 # "volume" | "volume_ESP" | "topo"
    data: dict                # frame arrays (numpy)
    arrival: float = 0.0      # seconds from the start of the run
    t_admit: float = math.nan
    t_done: float = math.nan
    result: object = None
    error: Optional[str] = None
    _work: Optional[float] = None
    _nbytes: Optional[int] = None

    def work(self) -> float:
        """Pair evaluations, an upper bound for streamlines (every start
        running its n_iter steps); computed once."""
        if self._work is None:
            d = self.data
            n = float(d["q"].size)
            if self.kind in ("volume", "volume_ESP"):
                self._work = n * float(np.prod([len(a) for a in d["axes"]]))
            else:
                self._work = n * float(np.sum(d["n_iter"]) + 4 * d["starts"].shape[0])
        return self._work

    def nbytes(self) -> int:
        """Device footprint: inputs, outputs and per-request scratch;
        computed once."""
        if self._nbytes is None:
            self._nbytes = self._footprint()
        return self._nbytes

    def _footprint(self) -> int:
        d = self.data
        n = int(d["q"].size)
        b = 16 * n + 32 * n  # raw + packed charges (float4 and double4)
        if self.kind in ("volume", "volume_ESP"):
            m = int(np.prod([len(a) for a in d["axes"]]))
            b += 4 * sum(len(a) for a in d["axes"]) + m * (12 if self.kind == "volume" else 4)
        else:
            L = int(d["starts"].shape[0])
            b += L * (12 + 4 + 16 + 24 + 8)
        return b


@dataclass
class ServeStats:
    completed: int = 0
    rejected: int = 0
    wall: float = 0.0
    latencies: List[float] = field(default_factory=list)
    batches: int = 0
    max_inflight_bytes: int = 0

    def summary(self) -> Dict[str, float]:
        lat = np.array(self.latencies) if self.latencies else np.array([math.nan])
        return {
            "completed": self.completed,
            "rejected": self.rejected,
            "frames_per_s": self.completed / self.wall if self.wall > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3),
            "batches": self.batches,
            "wall_s": self.wall,
            "max_inflight_bytes": self.max_inflight_bytes,
        }


class _Lane:
    def __init__(self, device: int):
        import torch

        self.torch = torch
        self.stream = torch.cuda.Stream(device=device)  # default priority
        self.ctx = nat.Context(device)
        self.groups = collections.deque()  # (event, [(request, (pinned view, pinned buffer))])

    @property
    def busy(self) -> bool:
        return bool(self.groups)


class _PinnedPool:
    """Reusable pinned host staging buffers in power-of-two size classes.
    A result is copied device -> pinned asynchronously, then into pageable
    memory when its request completes, and the pinned buffer is reused; so
    the pinned memory held is bounded by the results in flight, not by how
    many requests a long run has served."""

    def __init__(self, torch):
        self.torch = torch
        self.free = collections.defaultdict(list)

    def get(self, nbytes: int):
        size = max(4096, 1 << max(0, int(nbytes - 1).bit_length()))
        if self.free[size]:
            return self.free[size].pop()
        return self.torch.empty((size,), dtype=self.torch.uint8, pin_memory=True)

    def put(self, buf) -> None:
        self.free[buf.numel()].append(buf)


class Server:
    def __init__(self, device: int = 0, n_lanes: int = 4, hbm_budget: int = 8 << 30,
                 max_batch: int = 64, policy: str = "fifo-small", small_work: float = 2e9):
        if policy not in ("fifo-small", "greedy"):
            raise ValueError(f"unknown policy {policy!r}")
        import torch

        self.torch = torch
        self.dev = torch.device(f"cuda:{device}")
        self.lanes = [_Lane(device) for _ in range(n_lanes)]
        self.budget = hbm_budget
        self.max_batch = max_batch
        self.inflight = 0
        self.pool = _PinnedPool(torch)
        self.keep_results = True
        self.policy = policy
        self.small_work = small_work

    def _take(self, lane_idx: int, small: collections.deque, large: collections.deque) -> tuple:
        """The requests lane `lane_idx` takes next from the FIFO queues of
        small and large requests, and whether they are one large request
        (see the module docstring).  Under "greedy" every request is queued
        as small."""
        lane = self.lanes[lane_idx]
        if self.policy == "fifo-small" and lane_idx == 0 and len(self.lanes) > 1:
            if large and len(lane.groups) < 2 and \
                    self.inflight + large[0].nbytes() <= self.budget:
                return [large.popleft()], True
            return [], False
        if lane.busy:
            return [], False
        if self.policy == "fifo-small" and len(self.lanes) == 1 and large and \
                self.inflight + large[0].nbytes() <= self.budget:
            return [large.popleft()], True
        take, acc = [], 0
        while small and len(take) < self.max_batch and \
                self.inflight + acc + small[0].nbytes() <= self.budget:
            acc += small[0].nbytes()
            take.append(small.popleft())
        return take, False

    # -------------------------------------------------------------- batching
    def _to_dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to(self.dev, non_blocking=True)

    def _pinned(self, t):
        """Asynchronous copy of device tensor t into a pooled pinned buffer."""
        nbytes = t.numel() * t.element_size()
        buf = self.pool.get(nbytes)
        h = buf[:nbytes].view(t.dtype).view(t.shape)
        h.copy_(t, non_blocking=True)
        return h, buf

    def _close_group(self, lane: _Lane, outputs) -> None:
        ev = self.torch.cuda.Event()
        ev.record(lane.stream)
        lane.groups.append((ev, outputs))

    def _launch_one(self, lane: _Lane, r: Request) -> None:
        """One large request through the per-frame entry point on the lane's
        stream, then its result copy and event."""
        torch, T, d = self.torch, self._to_dev, r.data
        with torch.cuda.stream(lane.stream):
            if r.kind == "topo":
                out = ops.topo(T(d["starts"]), T(d["n_iter"]), d["h"], d["dims"], T(d["x"]),
                               T(d["q"]), ctx=lane.ctx).feats
            else:
                E, phi = ops.field_grid(*(T(a) for a in d["axes"]), T(d["x"]), T(d["q"]),
                                        want_E=r.kind == "volume",
                                        want_phi=r.kind == "volume_ESP", ctx=lane.ctx)
                out = E if r.kind == "volume" else phi
            self._close_group(lane, [(r, self._pinned(out))])

    def _launch(self, lane: _Lane, reqs: List[Request]) -> None:
        """Queues one batch call per request kind on the lane's stream, each
        followed by its results' copies into pinned memory and an event, so
        a kind's requests complete as soon as their own work is done."""
        torch, T = self.torch, self._to_dev
        pinned = self._pinned
        close_group = lambda outputs: self._close_group(lane, outputs)

        with torch.cuda.stream(lane.stream):
            for kind in ("volume", "volume_ESP", "topo"):
                grp = [r for r in reqs if r.kind == kind]
                if not grp:
                    continue
                if kind == "topo":
                    outs = B.batch_topo([{"x": T(r.data["x"]), "q": T(r.data["q"]),
                                          "starts": T(r.data["starts"]),
                                          "n_iter": T(r.data["n_iter"]), "h": r.data["h"],
                                          "dims": r.data["dims"]} for r in grp], ctx=lane.ctx)
                    close_group([(r, pinned(res.feats)) for r, res in zip(grp, outs)])
                else:
                    outs = B.batch_volume([{"x": T(r.data["x"]), "q": T(r.data["q"]),
                                            "axes": tuple(T(a) for a in r.data["axes"])}
                                           for r in grp], want_E=kind == "volume",
                                          want_phi=kind == "volume_ESP", ctx=lane.ctx)
                    close_group([(r, pinned(E if kind == "volume" else phi))
                                 for r, (E, phi) in zip(grp, outs)])

    def _finish_ready(self, lane: _Lane, t0: float, stats: ServeStats) -> int:
        """Completes the lane's groups whose events have fired, in order."""
        n = 0
        while lane.groups and lane.groups[0][0].query():
            _, outputs = lane.groups.popleft()
            now = time.perf_counter() - t0
            for r, (h, buf) in outputs:
                res = h.numpy().copy()
                r.result = res if self.keep_results else None
                self.pool.put(buf)
                r.t_done = now
                stats.latencies.append(r.t_done - r.arrival)
                stats.completed += 1
                self.inflight -= r.nbytes()
                n += 1
        return n

    def run(self, requests: List[Request], stats: Optional[ServeStats] = None,
            poll_s: float = 2e-4, progress=None, keep_results: bool = True) -> ServeStats:
        """Serves `requests` (sorted by arrival); returns when all are done.
        `progress`, if given, receives a status line every 500 completions.
        keep_results=False copies each result to the host and drops it (a
        load test of many requests), instead of storing it in r.result."""
        self.keep_results = keep_results
        stats = stats or ServeStats()
        next_report = 500
        arrivals = collections.deque(sorted(requests, key=lambda r: r.arrival))
        small: collections.deque = collections.deque()
        large: collections.deque = collections.deque()
        t0 = time.perf_counter()
        remaining = len(requests)
        while remaining > 0:
            now = time.perf_counter() - t0
            while arrivals and arrivals[0].arrival <= now:
                r = arrivals.popleft()
                if r.nbytes() > self.budget:
                    r.error = "request exceeds the HBM budget"
                    stats.rejected += 1
                    remaining -= 1
                    continue
                if self.policy == "fifo-small" and r.work() >= self.small_work:
                    large.append(r)
                else:
                    small.append(r)
            for lane in self.lanes:
                remaining -= self._finish_ready(lane, t0, stats)
            if progress is not None and stats.completed >= next_report:
                progress(f"batched: {stats.completed} done at {now:.1f} s, "
                         f"{len(small) + len(large)} queued")
                next_report += 500
            for li, lane in enumerate(self.lanes):
                if not (small or large):
                    break
                take, is_large = self._take(li, small, large)
                for r in take:
                    r.t_admit = time.perf_counter() - t0
                    self.inflight += r.nbytes()
                if take:
                    stats.max_inflight_bytes = max(stats.max_inflight_bytes, self.inflight)
                    if is_large:
                        self._launch_one(lane, take[0])
                    else:
                        self._launch(lane, take)
                    stats.batches += 1
            if not any(l.busy for l in self.lanes) and not (small or large) and arrivals:
                time.sleep(max(0.0, min(arrivals[0].arrival - now, 0.01)))
            else:
                time.sleep(poll_s)
        stats.wall = time.perf_counter() - t0
        return stats


def serve_plain(requests: List[Request], device: int = 0,
                stats: Optional[ServeStats] = None, max_wall_s: Optional[float] = None,
                progress=None, keep_results: bool = True) -> ServeStats:
    """One request at a time, per-frame entry points, one stream.

    With max_wall_s, serving stops at that wall time; requests not served
    by then are counted in `stats.rejected` and their latency enters the
    percentiles as the lower bound (stop time - arrival).  `progress`, if
    given, is called with a status string every 100 requests.
    keep_results=False copies each result to the host and drops it."""
    import torch

    stats = stats or ServeStats()
    dev = torch.device(f"cuda:{device}")
    ctx = nat.Context(device)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    queue = collections.deque(sorted(requests, key=lambda r: r.arrival))
    busy = collections.Counter()
    t0 = time.perf_counter()
    while queue:
        now = time.perf_counter() - t0
        if max_wall_s is not None and now > max_wall_s:
            for r in queue:
                stats.latencies.append(max(now - r.arrival, 0.0))
                stats.rejected += 1
            queue.clear()
            break
        r = queue.popleft()
        if r.arrival > now:
            time.sleep(r.arrival - now)
        r.t_admit = time.perf_counter() - t0
        d = r.data
        if r.kind in ("volume", "volume_ESP"):
            E, phi = ops.field_grid(*(T(a) for a in d["axes"]), T(d["x"]), T(d["q"]),
                                    want_E=r.kind == "volume", want_phi=r.kind == "volume_ESP",
                                    ctx=ctx)
            out = E if r.kind == "volume" else phi
        else:
            res = ops.topo(T(d["starts"]), T(d["n_iter"]), d["h"], d["dims"], T(d["x"]),
                           T(d["q"]), ctx=ctx)
            out = res.feats
        res = out.cpu().numpy()
        r.result = res if keep_results else None
        r.t_done = time.perf_counter() - t0
        busy[r.kind] += r.t_done - r.t_admit
        stats.latencies.append(r.t_done - r.arrival)
        stats.completed += 1
        if progress is not None and stats.completed % 100 == 0:
            progress(f"plain: {stats.completed} done at {r.t_done:.1f} s, service s by kind "
                     f"{dict((k, round(v, 2)) for k, v in busy.items())}")
    stats.wall = time.perf_counter() - t0
    return stats


def poisson_requests(rate: float, duration: float, seed: int = 0, *, mix=(0.5, 0.2, 0.3),
                     n_range=(1000, 60000), edge_range=(21, 101), starts_range=(10000, 200000),
                     topo_h: float = 0.1, topo_box: float = 0.5, topo_cap: Optional[int] = None,
                     max_requests: Optional[int] = None) -> List[Request]:
    """Poisson arrivals with the serving mix: volume / volume_ESP / topo, atom
    counts log-uniform in n_range, odd grid edges uniform in edge_range
    (step 0.1 A), topo starts uniform in starts_range.  Topo requests use
    h = 0.1 A in a +-0.5 A box (reference step-limit rule: at most 16 steps)
    so the offered load stays below one GPU's capacity.  Request arrays are
    views into shared seeded base arrays (no per-request copies)."""
    from . import workloads as W

    rng = np.random.default_rng(seed)
    base_x, base_q = W.clustered(n_range[1] // 10 + 1, 10, 45.0, 1.5, -3.0, seed)
    dims = np.full(3, topo_box, np.float32)
    base_st, base_it, _ = W.topo_starts(starts_range[1], dims, topo_h, topo_cap, seed=seed + 42,
                                        start_seed=seed + 1)
    t, reqs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_requests and len(reqs) >= max_requests):
            break
        u = rng.random()
        kind = "volume" if u < mix[0] else ("volume_ESP" if u < mix[0] + mix[1] else "topo")
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        off = int(rng.integers(0, base_q.size - n))
        x, q = base_x[off:off + n], base_q[off:off + n]
        if kind == "topo":
            L = int(rng.integers(starts_range[0], starts_range[1] + 1))
            so = int(rng.integers(0, starts_range[1] - L + 1))
            data = {"x": x, "q": q, "starts": base_st[so:so + L], "n_iter": base_it[so:so + L],
                    "h": topo_h, "dims": dims}
        else:
            e = int(rng.integers(edge_range[0] // 2, edge_range[1] // 2 + 1)) * 2 + 1
            d = (e - 1) * 0.05
            ax = np.linspace(-d, d, e).astype(np.float32)
            data = {"x": x, "q": q, "axes": (ax, ax, ax)}
        reqs.append(Request(len(reqs), kind, data, arrival=t))
    return reqs
