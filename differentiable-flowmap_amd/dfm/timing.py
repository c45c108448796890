"""Optional HIP-event timing of named regions on the current stream.

Disabled by default (one boolean test per call).  When enabled, every region
records a start/stop event pair on torch's current stream -- the stream every
dfm launch uses -- so the elapsed times are the device durations of the
enclosed launches.
"""
from __future__ import annotations

import contextlib
from collections import defaultdict

import torch

_enabled = False
_events = defaultdict(list)
_counts = defaultdict(int)


def enable(on: bool = True):
    global _enabled
    _enabled = on


def reset():
    _events.clear()
    _counts.clear()


@contextlib.contextmanager
def region(name, calls=1):
    """Time the enclosed launches as `calls` calls of `name` (a region that
    covers several overlapped operations counts each of them)."""
    if not _enabled:
        yield
        return
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    try:
        yield
    finally:
        e.record()
        _events[name].append((s, e, calls))


def add_count(name, k=1):
    if _enabled:
        _counts[name] += k


def summary():
    """{name: (calls, total_ms)} after synchronising."""
    torch.cuda.synchronize()
    return {k: (sum(c for _, _, c in v), sum(s.elapsed_time(e) for s, e, _ in v)) for k, v in _events.items()}


def counts():
    return dict(_counts)
