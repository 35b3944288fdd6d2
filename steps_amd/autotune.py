"""Launch-shape autotuner for the period kernel.

Times the fused period kernel under each candidate launch shape (threads per
block, graph replay of a run's launch sequence) on the simulation's own state
and keeps the fastest. Every random draw is keyed by (seed, tet, period), so
the launch shape changes speed only, never results; the state is checkpointed
before the trials and restored after them.
"""
from __future__ import annotations

import numpy as np

from .engine import Engine


def autotune(engine: Engine, trial_periods: int = 6, blocks=(64, 128, 256, 512),
             graphs=(False, True)) -> dict:
    state = engine.checkpoint()
    t0 = engine.time()
    dts = [engine.update_period(r) for r in range(engine.n_real)]
    finite = [d for d in dts if np.isfinite(d)]
    span = trial_periods * (max(finite) if finite else 1e-3)
    timings = {}
    for b in blocks:
        for g in graphs:
            engine.restore(state)
            engine.set_launch(b, g)
            engine.run(t0 + span)            # warm-up (captures the graph)
            engine.restore(state)
            engine.run(t0 + span)
            ms, n = engine.last_run_kernel_ms()
            timings[(b, g)] = ms
    best = min(timings, key=timings.get)
    engine.restore(state)
    engine.set_launch(*best)
    return {"block": best[0], "graph": best[1], "ms_per_launch": timings[best],
            "timings": {f"{b}/{'graph' if g else 'stream'}": v for (b, g), v in timings.items()}}
