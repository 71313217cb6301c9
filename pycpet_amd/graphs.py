"""HIP-graph captured steps.

A step is a function of a libcpet_hip context that launches device-resident
work only: ops.field_grid / ops.topo / batch.batch_volume / batch.batch_topo
on torch device tensors of fixed shapes.  `CapturedStep(fn)` gives it a
context of its own, runs it eagerly (scratch buffers grow, the batch tables
are uploaded), then captures one run into a graph with torch.cuda.graph.
`replay()` re-launches every kernel of the step on the current contents of
the input tensors and rewrites the same output tensors.

Why this is safe to capture: with device pointers and CPET_F_NO_SYNC the
library neither waits on the stream nor reads it back; scratch buffers only
grow (the eager run sized them); the batch tables are uploaded only when
they change (upload_table), so the captured step contains kernels only, and
buffers are cleared by fill kernels, never by hipMemsetAsync.  The context
is the step's own: another call on it with different shapes would replace
the tables the graph reads.

Replace input data in place (tensor.copy_) between replays; shapes, charge
counts and grid / streamline counts are fixed by the capture.
"""
from __future__ import annotations

from typing import Any, Callable, Dict, Optional, Tuple

from . import _native as nat


class CapturedStep:
    def __init__(self, fn: Callable[[nat.Context], Any], device: int = 0, warmup: int = 1,
                 tuning: Optional[Dict[int, Tuple[int, int]]] = None):
        import torch

        self.torch = torch
        self.ctx = nat.Context(device)
        self.ctx.set_profiling(False)
        for kind, (a, b) in (tuning or {}).items():  # launch shapes are fixed at capture
            self.ctx.set_tuning(kind, a, b)
        side = torch.cuda.Stream(device=device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            for _ in range(max(1, warmup)):
                fn(self.ctx)
        torch.cuda.current_stream(device).wait_stream(side)
        torch.cuda.synchronize(device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = fn(self.ctx)
        torch.cuda.synchronize(device)

    def replay(self):
        """Runs the captured step on torch's current stream; returns the
        step's output tensors (the same objects every time)."""
        self.graph.replay()
        return self.outputs

    def close(self) -> None:
        self.graph = None
        self.outputs = None
        self.ctx.close()
