"""Per-pass GPU timing with HIP events recorded on the stream the passes are
launched on (torch's current stream, which is also the stream handed to the
C-ABI).  Inactive unless a PassTimer is installed, so the step pays nothing
by default."""
from __future__ import annotations

import contextlib
from collections import defaultdict

import torch

_active = None


class PassTimer:
    def __init__(self):
        self.records = []

    def __enter__(self):
        global _active
        self._prev = _active
        _active = self
        return self

    def __exit__(self, *exc):
        global _active
        _active = self._prev
        return False

    @contextlib.contextmanager
    def scope(self, name):
        s = torch.cuda.current_stream()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        try:
            yield
        finally:
            b.record(s)
            self.records.append((name, a, b))

    def summary(self):
        """{name: (calls, total_ms, mean_ms)} (synchronizes)."""
        torch.cuda.synchronize()
        acc = defaultdict(lambda: [0, 0.0])
        for name, a, b in self.records:
            acc[name][0] += 1
            acc[name][1] += a.elapsed_time(b)
        return {k: (v[0], v[1], v[1] / v[0]) for k, v in acc.items()}


def scope(name):
    if _active is None:
        return contextlib.nullcontext()
    return _active.scope(name)
