"""Launch-shape autotuner for the period kernel.

Times a few periods of the simulation's own state under each candidate launch
shape (threads per block, graph replay of a run's launch sequence, the generic
or the model-specialised kernel) and keeps the fastest by wall time, which
includes the launch overhead that graph replay removes. Every random draw is
keyed by (seed, tet, period) and the specialised kernel performs the generic
kernel's operations, so the shape changes speed only, never results; the state
is checkpointed before the trials and restored after them.
"""
from __future__ import annotations

import time

import numpy as np

from .engine import Engine


def autotune(engine: Engine, trial_periods: int = 6, blocks=(64, 128, 256, 512),
             graphs=(False, True), repeats: int = 2, specialize=(False, True),
             spec_waves=(0,)) -> dict:
    """Each shape runs trial_periods periods once to warm up and then `repeats`
    times timed; its figure is the fastest of those. The specialised kernel is
    tried for blocks of up to 256 when the model allows it, compiled with each
    waves-per-SIMD bound in spec_waves (0: the compiler's choice)."""
    from ._native import StgError

    spec_opts = [None] if False in specialize else []   # None: the generic kernel
    if True in specialize:
        try:
            for w in spec_waves:
                engine.set_specialize(True, w)
                spec_opts.append(w)
        except StgError:
            pass
        engine.set_specialize(False)
    if not spec_opts:
        spec_opts = [None]
    state = engine.checkpoint()
    t0 = engine.time()
    dts = [engine.update_period(r) for r in range(engine.n_real)]
    finite = [d for d in dts if np.isfinite(d)]
    span = trial_periods * (max(finite) if finite else 1e-3)
    timings = {}
    cands = [(b, g, sp) for sp in spec_opts for b in blocks for g in graphs if not (sp is not None and b > 256)]
    for b, g, sp in cands:
        engine.set_specialize(sp is not None, sp or 0)
        engine.restore(state)
        engine.set_launch(b, g)
        engine.run(t0 + span)            # warm-up (captures the graph)
        best_t = None
        for _ in range(max(1, repeats)):
            engine.restore(state)
            tic = time.perf_counter()
            engine.run(t0 + span)        # returns once the stream has drained
            wall = time.perf_counter() - tic
            ms, n = engine.last_run_kernel_ms()
            cand = (wall * 1e3 / max(n, 1), ms)
            if best_t is None or cand[0] < best_t[0]:
                best_t = cand
        timings[(b, g, sp)] = best_t
    best = min(timings, key=lambda k: timings[k][0])
    engine.restore(state)
    engine.set_specialize(best[2] is not None, best[2] or 0)
    engine.set_launch(best[0], best[1])

    def label(b, g, sp):
        kern = "" if sp is None else f"/spec{sp or ''}"
        return f"{b}/{'graph' if g else 'stream'}{kern}"

    return {"block": best[0], "graph": best[1], "specialized": best[2] is not None,
            "spec_min_waves": best[2], "wall_ms_per_launch": timings[best][0],
            "kernel_ms_per_launch": timings[best][1],
            "timings": {label(*k): round(v[0], 5) for k, v in timings.items()}}
