"""HIP-graph captured steps.

A step is a function of a libcpet_hip context that launches device-resident
work only: ops.field_grid / ops.topo / batch.batch_volume / batch.batch_topo
on torch device tensors of fixed shapes.  `CapturedStep(fn)` gives it a
context of its own, runs it eagerly (scratch buffers grow, the batch tables
are uploaded), then captures one run into a graph with torch.cuda.graph.
`replay()` re-launches every kernel of the step on the current contents of
the input tensors and rewrites the same output tensors.  Several step
functions become parallel branches of one graph.

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

from . import _native as nat


class CapturedStep:
    """fn: a step function of a context, or a list of them.  A list is
    captured as parallel branches of one graph (each branch on its own
    context and side stream, forked from and joined back to the capture
    stream), so one replay runs e.g. the batched volume and streamline
    launches concurrently.  tuning: {kind: (a, b)} for every branch, or a
    list with one dict per branch."""

    def __init__(self, fn, device: int = 0, warmup: int = 1, tuning=None):
        import torch

        self.torch = torch
        fns = list(fn) if isinstance(fn, (list, tuple)) else [fn]
        tunings = list(tuning) if isinstance(tuning, (list, tuple)) else [tuning] * len(fns)
        if len(tunings) != len(fns):
            raise ValueError("one tuning dict per branch")
        self.ctxs = []
        for t in tunings:
            c = nat.Context(device)
            c.set_profiling(False)
            for kind, (a, b) in (t or {}).items():  # launch shapes are fixed at capture
                c.set_tuning(kind, a, b)
            self.ctxs.append(c)
        self.ctx = self.ctxs[0]
        main = torch.cuda.current_stream(device)
        sides = [torch.cuda.Stream(device=device) for _ in fns]
        for s in sides:
            s.wait_stream(main)
        for f, c, s in zip(fns, self.ctxs, sides):
            with torch.cuda.stream(s):
                for _ in range(max(1, warmup)):
                    f(c)
        for s in sides:
            main.wait_stream(s)
        torch.cuda.synchronize(device)
        self.graph = torch.cuda.CUDAGraph()
        outs = []
        with torch.cuda.graph(self.graph):
            cap = torch.cuda.current_stream(device)
            for f, c, s in zip(fns, self.ctxs, sides):
                s.wait_stream(cap)
                with torch.cuda.stream(s):
                    outs.append(f(c))
            for s in sides:
                cap.wait_stream(s)
        torch.cuda.synchronize(device)
        self.outputs = outs if isinstance(fn, (list, tuple)) else outs[0]

    def replay(self):
        """Runs the captured step on torch's current stream; returns the
        step's output tensors (the same objects every time)."""
        self.graph.replay()
        return self.outputs

    def close(self) -> None:
        self.graph = None
        self.outputs = None
        for c in self.ctxs:
            c.close()
