"""Autotuner for the launch shapes of the field and streamline kernels.

For a workload shape it times each candidate configuration of the runtime
knobs (cpet_set_tuning) with HIP events on the kernel's stream, keeps the
fastest, and caches the choice per shape bucket in a JSON file so later
runs apply it without searching:

  grid field     kind 0: points per lane along z (1,2,3,4,6,8) x charge splits
  streamlines    kind 2: waves per workgroup (charge slices) x workgroups
  batch topo     kind 3: waves per workgroup x persistent workgroups
  batch volume   kind 4: points per lane (1,2,4)

Buckets are powers of two of the charge count and of the point / streamline
count, so frames of similar size share a choice.  Grid candidates compute the
same per-pair arithmetic (charge splits regroup FP64 partial sums).
Streamline candidates with 1 or 16 waves per workgroup give the same results
bit for bit (the one wave walks the 16 waves' charge slices in order); 2, 4
and 8 waves slice the charges differently, and agree to FP32 rounding.
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path
from typing import Callable, Dict, Iterable, Optional, Tuple

import numpy as np

from . import _native as nat

CACHE = Path(os.environ.get("CPET_TUNE_CACHE", Path.home() / ".cache" / "pycpet_amd_tune.json"))


def _bucket(n: int) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(1, n)))))


def key(kind: int, n_charges: int, n_work: int) -> str:
    return f"{kind}:{_bucket(n_charges)}:{_bucket(n_work)}"


def load_cache(path: Path = CACHE) -> Dict[str, list]:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def save_cache(cache: Dict[str, list], path: Path = CACHE) -> None:
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    tmp = p.with_suffix(".tmp")
    tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
    os.replace(tmp, p)


def sweep(ctx: nat.Context, kind: int, candidates: Iterable[Tuple[int, int]],
          run: Callable[[], None], phase: int, repeats: int = 2) -> Tuple[Tuple[int, int], dict]:
    """Times run() for each (a, b) candidate; returns the best and all times
    (ms of the kernel phase `phase` reported by cpet_ctx_phase_times)."""
    ctx.set_profiling(True)
    times = {}
    for a, b in candidates:
        ctx.set_tuning(kind, a, b)
        try:
            run()  # warm-up (code object load, scratch growth)
            ms = []
            for _ in range(repeats):
                run()
                ms.append(ctx.phase_times(4)[phase])
            times[(a, b)] = float(np.median(ms))
        except nat.CpetError:
            continue
    ctx.set_tuning(kind, 0, 0)
    best = min(times, key=times.get)
    return best, {f"{a},{b}": t for (a, b), t in times.items()}


def tune_topo(ctx: nat.Context, starts, n_iter, h, dims, x, q, *, cache: Optional[dict] = None,
              waves=(4, 8, 16), groups=(0, 256, 512, 1024)) -> Tuple[int, int]:
    from . import ops

    n = int(np.prod(q.shape))
    L = int(starts.shape[0])
    k = key(2, n, L)
    if cache is not None and k in cache:
        a, b = cache[k]
        ctx.set_tuning(2, a, b)
        return a, b
    # one-wave workgroups (same results as 16 waves) at their default count
    cands = [(1, 0)] + [(s, g) for s in waves if s != 1 for g in groups]
    best, _ = sweep(ctx, 2, cands, lambda: ops.topo(starts, n_iter, h, dims, x, q, ctx=ctx),
                    phase=1)
    if cache is not None:
        cache[k] = list(best)
    ctx.set_tuning(2, *best)
    return best


def tune_grid(ctx: nat.Context, axes, x, q, *, want_phi=False, cache: Optional[dict] = None,
              ks=(2, 3, 4, 6, 8), splits=(0,)) -> Tuple[int, int]:
    from . import ops

    n = int(np.prod(q.shape))
    m = int(np.prod([len(a) for a in axes]))
    k = key(0, n, m)
    if cache is not None and k in cache:
        a, b = cache[k]
        ctx.set_tuning(0, a, b)
        return a, b
    cands = [(kk, s) for kk in ks for s in splits]
    best, _ = sweep(ctx, 0, cands,
                    lambda: ops.field_grid(*axes, x, q, want_phi=want_phi, ctx=ctx), phase=0)
    if cache is not None:
        cache[k] = list(best)
    ctx.set_tuning(0, *best)
    return best


def tune_batch_topo(ctx: nat.Context, frames, *, cache: Optional[dict] = None,
                    waves=(4, 8, 16), groups=(0, 256, 512, 1024)) -> Tuple[int, int]:
    """(waves per workgroup, persistent workgroups) of the batched
    streamline launch, timed on `frames` (a representative subset)."""
    from . import batch as B

    n = int(np.mean([int(np.prod(f["q"].shape)) for f in frames]))
    L = int(sum(int(f["starts"].shape[0]) for f in frames))
    k = key(3, n, L)
    if cache is not None and k in cache:
        a, b = cache[k]
        ctx.set_tuning(3, a, b)
        return a, b
    # one-wave workgroups (same results as 16 waves) at their default count
    cands = [(1, 0)] + [(w, g) for w in waves if w != 1 for g in groups]
    best, _ = sweep(ctx, 3, cands, lambda: B.batch_topo(frames, ctx=ctx), phase=0)
    if cache is not None:
        cache[k] = list(best)
    ctx.set_tuning(3, *best)
    return best


def tune_batch_volume(ctx: nat.Context, frames, *, cache: Optional[dict] = None,
                      ks=(1, 2, 4)) -> Tuple[int, int]:
    """Points per lane of the batched grid launch, timed on `frames`."""
    from . import batch as B

    n = int(np.mean([int(np.prod(f["q"].shape)) for f in frames]))
    m = int(sum(int(np.prod([len(a) for a in f["axes"]])) for f in frames))
    k = key(4, n, m)
    if cache is not None and k in cache:
        a, b = cache[k]
        ctx.set_tuning(4, a, b)
        return a, b
    best, _ = sweep(ctx, 4, [(kk, 0) for kk in ks], lambda: B.batch_volume(frames, ctx=ctx),
                    phase=0)
    if cache is not None:
        cache[k] = list(best)
    ctx.set_tuning(4, *best)
    return best
