"""Launch-shape autotuner for the period kernel.

Times a few periods of the simulation's own state under each candidate launch
shape (threads per block, lanes per tet, face slot width, graph replay of a
run's launch sequence, the generic or the model-specialised kernel) and keeps
the fastest by wall time, which includes the launch overhead that graph replay
removes. 8-bit face slots move fewer bytes than 16-bit ones as long as few faces
pass more than 127 molecules in a period (those take a 32-bit side entry).
Lanes per tet > 1 (the tile schedule) are tried for volume-only models: on a
mesh too small to fill the GPU a period lasts one wavefront's critical path,
which the tile schedule shortens; on a large mesh one lane per tet keeps every
lane busy. Every random draw is
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
             spec_waves=(0,), lanes=(1, 2, 4), slot_widths=(2, 1)) -> dict:
    """Each shape runs trial_periods periods once to warm up and then `repeats`
    times timed; its figure is the fastest of those. The specialised kernel is
    tried for blocks of up to 256 when the model allows it, compiled with each
    waves-per-SIMD bound in spec_waves (0: the compiler's choice). Lane counts
    above 1 are tried when the model allows them (volume-only models), with
    stream launches; graph replay is tried with one lane per tet. Every shape
    is tried with each face slot width in slot_widths."""
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
    lane_opts = [1]
    for g in lanes:
        if g == 1:
            continue
        try:
            engine.set_lanes(g)
            lane_opts.append(g)
        except StgError:       # the model needs the membrane variant
            break
    engine.set_lanes(1)
    state = engine.checkpoint()
    t0 = engine.time()
    dts = [engine.update_period(r) for r in range(engine.n_real)]
    finite = [d for d in dts if np.isfinite(d)]
    span = trial_periods * (max(finite) if finite else 1e-3)
    timings = {}
    cands = [(b, g, sp, ln, sw) for sw in slot_widths for ln in lane_opts for sp in spec_opts for b in blocks
             for g in graphs if not (sp is not None and b > 256) and not (ln > 1 and g)]
    for b, g, sp, ln, sw in cands:
        engine.set_slot_width(sw)
        engine.set_lanes(ln)
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
        timings[(b, g, sp, ln, sw)] = best_t
    best = min(timings, key=lambda k: timings[k][0])
    engine.restore(state)
    engine.set_slot_width(best[4])
    engine.set_lanes(best[3])
    engine.set_specialize(best[2] is not None, best[2] or 0)
    engine.set_launch(best[0], best[1])

    def label(b, g, sp, ln, sw):
        kern = "" if sp is None else f"/spec{sp or ''}"
        tile = "" if ln == 1 else f"/lanes{ln}"
        slot = "" if sw == 2 else "/slot8"
        return f"{b}/{'graph' if g else 'stream'}{kern}{tile}{slot}"

    return {"block": best[0], "graph": best[1], "specialized": best[2] is not None,
            "spec_min_waves": best[2], "lanes": best[3], "slot_width": best[4],
            "wall_ms_per_launch": timings[best][0],
            "kernel_ms_per_launch": timings[best][1],
            "timings": {label(*k): round(v[0], 5) for k, v in timings.items()}}
