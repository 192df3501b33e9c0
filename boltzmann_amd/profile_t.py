"""Accumulating wall-clock timers (semantics of BESolver/python/profile_t.py:1-33)."""
from time import perf_counter


class profile_t:
    def __init__(self, name):
        self.name = name
        self.seconds = 0.0
        self.snap = 0.0
        self._t_start = 0.0
        self.iter = 0

    def start(self):
        self._t_start = perf_counter()

    def stop(self):
        self.snap = perf_counter() - self._t_start
        self.seconds += self.snap
        self.iter += 1

    def reset(self):
        self.seconds = 0.0
        self.snap = 0.0
        self._t_start = 0.0
        self.iter = 0
