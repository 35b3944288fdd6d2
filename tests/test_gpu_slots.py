"""8-bit face slots (stg_set_slot_width) against the oracle and against 16-bit
slots: counts above 127 per face and period take the 32-bit side entry, so
every count, extent and period must be identical.

Cases: the four-species model (both schedule modes, generic and specialised,
one and four lanes per tet), molecule counts that escape on most faces, the
calcium membrane model (membrane variant, surface reactions), a membrane
between two compartments (SSA groups), and switching the width between runs.
"""
import numpy as np
import pytest

from oracle import bind
from steps_amd import workloads as W
from steps_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _engine(w, seed, mode=0):
    eng = Engine(w.fm, [w.mt], [seed], mode=mode)
    for s in range(len(w.fm.specs)):
        eng.set_tet_counts(s, w.counts[s])
    for s in range(len(w.fm.sspecs)):
        eng.set_tri_counts(s, w.tri_counts[s])
    return eng


def _oracle(w, seed):
    o = bind.OracleSim(w.fm, w.mt, seed=seed)
    o.counts[:] = w.counts
    if len(w.fm.sspecs):
        o.tri_counts[:] = w.tri_counts
    return o


def _counts(eng, w):
    return np.stack([eng.get_tet_counts(s) for s in range(len(w.fm.specs))])


@pytest.mark.parametrize("lanes", [1, 4])
@pytest.mark.parametrize("spec", [False, True])
@pytest.mark.parametrize("mode", [0, 1])
def test_four_species_8bit_slots_bit_exact(mode, spec, lanes):
    w = W.four_species_cube(n=5, side=5e-6, seed=2, end_time=0.05)
    print(f"case four_species slot8 mode={mode} spec={spec} lanes={lanes}", flush=True)
    eng = _engine(w, 9, mode)
    eng.set_slot_width(1)
    eng.set_lanes(lanes)
    if spec:
        eng.set_specialize(True)
        assert eng.specialized
    assert eng.slot_width == 1
    o = _oracle(w, 9)
    for t in (0.004, 0.02, w.end_time):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_counts(eng, w), o.counts), f"t={t}"
    assert eng.nsteps() == o.nsteps and eng.extents() == (o.events.value, o.moved.value)


@pytest.mark.parametrize("mode", [0, 1])
def test_escapes_on_most_faces(mode):
    """About 400,000 molecules per tet (no reactions): almost every face passes
    far more than 127 molecules per period, a few pass fewer."""
    w = W.ab_binding_cube(n=2, side=2e-6, n_a=10_000_000, n_b=10_000_000, seed=6,
                          end_time=0.15, kf=0.0, kb=0.0)
    print(f"case slot8 escapes mode={mode}", flush=True)
    eng = _engine(w, 31, mode)
    eng.set_slot_width(1)
    o = _oracle(w, 31)
    for t in (0.05, 0.15):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_counts(eng, w), o.counts), f"t={t}"
    assert eng.extents()[1] == o.moved.value and o.moved.value > 100 * 32767


def test_counts_near_the_8bit_limit():
    """Poisson(300) per tet: faces pass molecule counts on both sides of 127."""
    w = W.four_species_cube(n=4, side=4e-6, seed=3, end_time=0.02, mean_count=300)
    print("case slot8 near the limit", flush=True)
    wide = _engine(w, 12)
    narrow = _engine(w, 12)
    narrow.set_slot_width(1)
    narrow.set_specialize(True)
    o = _oracle(w, 12)
    for t in (0.006, w.end_time):
        wide.run(t)
        narrow.run(t)
        o.run(t, 0)
        assert np.array_equal(_counts(narrow, w), o.counts), f"t={t}"
        assert np.array_equal(_counts(wide, w), o.counts), f"t={t}"


def test_membrane_models_with_8bit_slots():
    w = W.calcium_membrane_cube(n=6, side=2e-6, seed=4, end_time=2e-4, buf_um=100.0, ca_um=50.0)
    print("case slot8 calcium membrane", flush=True)
    for spec in (False, True):
        eng = _engine(w, 123)
        eng.set_slot_width(1)
        if spec:
            eng.set_specialize(True)
        o = _oracle(w, 123)
        eng.run(w.end_time)
        o.run(w.end_time, 0)
        assert np.array_equal(_counts(eng, w), o.counts), f"spec={spec}"
        tri = np.stack([eng.get_tri_counts(s) for s in range(len(w.fm.sspecs))])
        assert np.array_equal(tri, o.tri_counts[: len(w.fm.sspecs)]), f"spec={spec}"


def test_switching_the_width_between_runs():
    w = W.four_species_cube(n=4, side=4e-6, seed=5, end_time=0.04)
    print("case slot width switched between runs", flush=True)
    ref = _engine(w, 6)
    eng = _engine(w, 6)
    eng.set_specialize(True)
    for k, width in enumerate([1, 2, 1, 1, 2]):
        eng.set_slot_width(width)
        t = (k + 1) * w.end_time / 5
        ref.run(t)
        eng.run(t)
        assert np.array_equal(_counts(eng, w), _counts(ref, w)), f"width={width}"
        assert eng.extents() == ref.extents() and eng.specialized
    with pytest.raises(Exception):
        eng.set_slot_width(4)
