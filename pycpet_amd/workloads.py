"""Seeded synthetic inputs for the measured and tested workloads.

No datasets or checkpoints are available, so every workload is generated:

  ball_exclusion(n, R, box, excl, seed)
      n point charges uniform in a radius-R ball (r = R u^(1/3), direction =
      normalised N(0, I)), rejecting charges within `excl` A of the box
      [-box, box]^3; q ~ N(0, 0.4) e shifted to zero net charge (FP64), then
      cast to float32.
  clustered(n_clusters, per, R, sigma, net, seed)
      residue-like Gaussian clusters (centres uniform in the ball), charges a
      bimodal mix of +-0.8 e and small |q| < 0.2 e, adjusted so the FP64 sum
      is `net`.

  volume_2k   2,000 charges, 25 A ball, 3 A exclusion, grid +-2 A step 0.2
              (21^3 points), field and ESP.
  topo_5k     5,000 charges, 35 A ball, 3 A exclusion, 100,000 starts
              uniform in +-1.5 A, h = 0.001 A, n_iter = randint(1,
              max_steps) (seed 42) capped at 2,000 steps, Euler.
  volume_30k  30,000 clustered charges (3,000 clusters of 10, sigma 1.5 A,
              45 A ball, net -3 e), grid +-5 A step 0.1 (101^3 points).

Grids follow the reference rule N_cr = int(2 dims / step) on float64 and
linspace(-d, d, N_cr + 1) cast to float32 (CPET/source/calculator.py:258-259,
CPET/utils/calculator.py:218-233); topology step limits follow
max_steps = round(2 |dims| / h) and randint(1, max_steps)
(CPET/source/calculator.py:272, CPET/utils/calculator.py:167).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np


def _ball(rng: np.random.Generator, n: int, R: float) -> np.ndarray:
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = R * rng.random(n) ** (1.0 / 3.0)
    return d * r[:, None]


def _dist_to_box(p: np.ndarray, box: float) -> np.ndarray:
    over = np.maximum(np.abs(p) - box, 0.0)
    return np.sqrt((over * over).sum(axis=1))


def ball_exclusion(n: int, R: float, box: float, excl: float, seed: int):
    rng = np.random.default_rng(seed)
    pts = np.empty((0, 3))
    while pts.shape[0] < n:
        cand = _ball(rng, 2 * n, R)
        cand = cand[_dist_to_box(cand, box) >= excl]
        pts = np.concatenate([pts, cand])
    pts = pts[:n]
    q = rng.normal(0.0, 0.4, n)
    q -= q.mean()
    return pts.astype(np.float32), q.astype(np.float32)


def clustered(n_clusters: int, per: int, R: float, sigma: float, net: float,
              seed: int, jitter: float = 0.0):
    rng = np.random.default_rng(seed)
    centres = _ball(rng, n_clusters, R)
    pts = (centres[:, None, :] + sigma * rng.standard_normal((n_clusters, per, 3))).reshape(-1, 3)
    n = pts.shape[0]
    big = rng.random(n) < 0.5
    q = np.where(big, np.where(rng.random(n) < 0.5, 0.8, -0.8), rng.uniform(-0.19, 0.19, n))
    # adjust the small charges so the FP64 sum is exactly `net`
    small = ~big
    q[small] -= (q.sum() - net) / small.sum()
    q[small] = np.clip(q[small], -0.1999, 0.1999)
    q[np.flatnonzero(small)[0]] -= q.sum() - net
    if jitter > 0:
        pts = pts + jitter * rng.standard_normal(pts.shape)
    return pts.astype(np.float32), q.astype(np.float32)


def grid_axes(dims, step):
    """Per-axis float32 coordinates of the reference's inclusive box grid."""
    d = np.asarray(dims, dtype=np.float64)
    n_cr = 2 * d / step
    n_cr = [int(n_cr[0]), int(n_cr[1]), int(n_cr[2])]
    return [np.linspace(-d[i], d[i], n_cr[i] + 1).astype(np.float32) for i in range(3)]


def max_steps_for(dims, h) -> int:
    return round(2 * np.linalg.norm(np.asarray(dims, dtype=np.float64)) / h)


def topo_starts(L: int, dims, h: float, cap: Optional[int], seed: int = 42,
                start_seed: int = 1):
    """Uniform starts in +-dims and reference-rule step limits (seeded)."""
    dims = np.asarray(dims, dtype=np.float64)
    rs = np.random.RandomState(start_seed)
    starts = np.column_stack([rs.uniform(-dims[i], dims[i], L) for i in range(3)]).astype(np.float32)
    ms = max_steps_for(dims, h)
    n_iter = np.random.RandomState(seed).randint(1, ms, L)
    if cap is not None:
        n_iter = np.minimum(n_iter, cap)
    return starts, n_iter.astype(np.int32), ms


@dataclass
class VolumeCase:
    name: str
    x: np.ndarray
    q: np.ndarray
    axes: list
    dims: tuple
    step: float

    @property
    def M(self) -> int:
        return int(np.prod([a.size for a in self.axes]))

    @property
    def N(self) -> int:
        return int(self.q.size)

    def mesh_points(self) -> np.ndarray:
        g = np.meshgrid(*self.axes, indexing="ij")
        return np.stack(g, axis=-1).reshape(-1, 3).astype(np.float32)


@dataclass
class TopoCase:
    name: str
    x: np.ndarray
    q: np.ndarray
    starts: np.ndarray
    n_iter: np.ndarray
    h: float
    dims: np.ndarray
    max_steps: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.q.size)

    @property
    def L(self) -> int:
        return int(self.starts.shape[0])


def volume_2k(seed: int = 0) -> VolumeCase:
    x, q = ball_exclusion(2000, 25.0, 2.0, 3.0, seed)
    return VolumeCase("volume_2k", x, q, grid_axes((2.0, 2.0, 2.0), 0.2), (2.0, 2.0, 2.0), 0.2)


def volume_30k(seed: int = 2) -> VolumeCase:
    x, q = clustered(3000, 10, 45.0, 1.5, -3.0, seed)
    return VolumeCase("volume_30k", x, q, grid_axes((5.0, 5.0, 5.0), 0.1), (5.0, 5.0, 5.0), 0.1)


def topo_5k(L: int = 100_000, cap: int = 2000, seed: int = 1) -> TopoCase:
    x, q = ball_exclusion(5000, 35.0, 1.5, 3.0, seed)
    dims = np.array([1.5, 1.5, 1.5], dtype=np.float32)
    h = 0.001
    starts, n_iter, ms = topo_starts(L, dims, h, cap)
    return TopoCase("topo_5k", x, q, starts, n_iter, h, dims, ms)


def small_topo(N: int = 300, L: int = 256, cap: int = 400, seed: int = 5,
               dims=(0.6, 0.6, 0.6), h: float = 0.01, R: float = 12.0) -> TopoCase:
    """A small streamline case the CPU oracle finishes in seconds."""
    x, q = ball_exclusion(N, R, max(dims), 2.0, seed)
    d = np.asarray(dims, dtype=np.float32)
    starts, n_iter, ms = topo_starts(L, d, h, cap, seed=seed + 42, start_seed=seed)
    return TopoCase(f"small_topo_{N}x{L}", x, q, starts, n_iter, h, d, ms)


def md_base(n_max: int = 40000, seed: int = 3):
    """The shared clustered charge set MD-like frames are cut from."""
    return clustered(n_max // 10, 10, 45.0, 1.5, -3.0, seed)


def md_frame(base, i: int, kind: str, n: int, *, edge: Optional[int] = None,
             edges=(21, 41, 61, 81), seed: int = 3, topo_starts_per_frame: int = 20000,
             h: float = 0.001, cap: int = 2000, topo_box: float = 1.5, jitter: float = 0.3,
             rng=None):
    """Frame i of the mixed MD-like set: the first n charges of `base` with
    Gaussian position jitter, as a volume job on an edge^3 grid (step
    0.1 A, half-width (edge-1)*0.05 A; edge drawn from `edges` after the
    jitter when not given) or a streamline job with `topo_starts_per_frame`
    starts in +-topo_box."""
    base_x, base_q = base
    if n > base_q.size:
        raise ValueError(f"frame of {n} charges from a base of {base_q.size}")
    rng = rng if rng is not None else np.random.default_rng(seed * 7919 + i)
    x = (base_x[:n] + jitter * rng.standard_normal((n, 3))).astype(np.float32)
    q = base_q[:n].copy()
    if kind == "volume":
        if edge is None:
            edge = int(rng.choice(edges))
        d = (edge - 1) * 0.05
        ax = np.linspace(-d, d, edge).astype(np.float32)
        return {"kind": "volume", "x": x, "q": q, "axes": (ax, ax, ax)}
    dims = np.full(3, topo_box, np.float32)
    st, it, _ = topo_starts(topo_starts_per_frame, dims, h, cap, seed=seed * 1000 + i,
                            start_seed=seed * 1000 + i + 1)
    return {"kind": "topo", "x": x, "q": q, "starts": st, "n_iter": it, "h": h, "dims": dims}


def md_frames(n_frames: int = 512, seed: int = 3, n_min: int = 3000, n_max: int = 40000,
              edges=(21, 41, 61, 81), topo_starts_per_frame: int = 20000, h: float = 0.001,
              cap: int = 2000, topo_box: float = 1.5, jitter: float = 0.3):
    """Mixed MD-like frame set: atom counts uniform in [n_min, n_max], the
    clustered recipe of volume_30k with per-frame Gaussian jitter; even
    frames are volume jobs on a grid of edge drawn from `edges` (step 0.1 A,
    so the box half-width is (edge-1)*0.05 A), odd frames are streamline
    jobs with `topo_starts_per_frame` starts in +-topo_box."""
    rng = np.random.default_rng(seed)
    base = md_base(n_max, seed)
    frames = []
    for i in range(n_frames):
        n = int(rng.integers(n_min, n_max + 1)) // 10 * 10
        kind = "volume" if i % 2 == 0 else "topo"
        frames.append(md_frame(base, i, kind, n, edges=edges, seed=seed,
                               topo_starts_per_frame=topo_starts_per_frame, h=h, cap=cap,
                               topo_box=topo_box, jitter=jitter, rng=rng))
    return frames
