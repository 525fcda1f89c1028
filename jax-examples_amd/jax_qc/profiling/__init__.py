"""Hierarchical stage timer with the reference's stage names
(jax_qc/profiling/timer.py:40-172): ``integrals/{overlap,kinetic,nuclear,
h_core,eri,nuclear_repulsion}``, ``scf/{orthogonalize,initial_guess,
fock_build,damp,diis,diag,density,energy}``.  With ``sync_device`` the GPU
is synchronised around each stage so wall times are device times."""

from __future__ import annotations

import contextlib
import time
from collections import OrderedDict

import torch


class StageTimer:
    def __init__(self, sync_device: bool = True):
        self.sync = sync_device
        self.totals = OrderedDict()
        self.counts = OrderedDict()
        self._stack = []

    @contextlib.contextmanager
    def stage(self, name: str, fp: str = ""):
        path = "/".join(self._stack + [name])
        self._stack.append(name)
        if self.sync and torch.cuda.is_available():
            torch.cuda.synchronize()
        t = time.perf_counter()
        try:
            yield
        finally:
            if self.sync and torch.cuda.is_available():
                torch.cuda.synchronize()
            dt = time.perf_counter() - t
            self._stack.pop()
            self.totals[path] = self.totals.get(path, 0.0) + dt
            self.counts[path] = self.counts.get(path, 0) + 1

    def report(self) -> str:
        lines = []
        for k, v in self.totals.items():
            depth = k.count("/")
            lines.append(f"{'  ' * depth}{k.split('/')[-1]:<24s} {v * 1e3:10.3f} ms  x{self.counts[k]}")
        return "\n".join(lines)


def optional_stage(timer, name: str, fp: str = ""):
    if timer is None:
        return contextlib.nullcontext()
    return timer.stage(name, fp)
