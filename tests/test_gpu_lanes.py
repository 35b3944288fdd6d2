"""The tile schedule (several lanes per tet, stg_set_lanes) against the oracle
and against one lane per tet: every count, extent and period bit for bit.

The lanes of a tet split its species for the entering gather, the leaving pass
and the stores while one of them runs the tet's SSA; each species keeps its
own Philox channel, so the draws are those of one lane per tet. The cases
cover the generic and the model-specialised kernel, both schedule modes,
species counts that do not divide by the lane count (3 and 4 species with 2,
4 and 8 lanes), batches of unequal realizations, escaped 16-bit face slots,
switching the lane count between runs, and the refusal for membrane models.
"""
import numpy as np
import pytest

from oracle import bind
from steps_amd import workloads as W
from steps_amd.engine import Engine
from steps_amd._native import NotImplErr

pytestmark = pytest.mark.gpu


def _engine(w, seed, mode=0):
    eng = Engine(w.fm, [w.mt], [seed], mode=mode)
    for s in range(len(w.fm.specs)):
        eng.set_tet_counts(s, w.counts[s])
    for s in range(len(w.fm.sspecs)):
        eng.set_tri_counts(s, w.tri_counts[s])
    return eng


def _counts(eng, w, real=0):
    return np.stack([eng.get_tet_counts(s, real) for s in range(len(w.fm.specs))])


@pytest.mark.parametrize("spec", [False, True])
@pytest.mark.parametrize("lanes", [2, 4, 8])
@pytest.mark.parametrize("mode", [0, 1])
def test_four_species_lanes_bit_exact(mode, lanes, spec):
    w = W.four_species_cube(n=5, side=5e-6, seed=2, end_time=0.05)
    print(f"case four_species lanes={lanes} spec={spec} mode={mode}", flush=True)
    eng = _engine(w, 9, mode)
    eng.set_lanes(lanes)
    if spec:
        eng.set_specialize(True)
        assert eng.specialized
    eng.set_launch(128, False)
    assert eng.lanes == lanes
    o = bind.OracleSim(w.fm, w.mt, seed=9)
    o.counts[:] = w.counts
    for t in (0.004, 0.02, w.end_time):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_counts(eng, w), o.counts), f"t={t}"
    assert eng.nsteps() == o.nsteps
    assert eng.extents() == (o.events.value, o.moved.value)
    assert o.events.value > 0 and o.moved.value > 0


@pytest.mark.parametrize("lanes", [2, 4])
def test_batch_of_unequal_realizations(lanes):
    fm, meshes, counts, reac_k, info = W.ab_binding_ensemble(n_real=6, n_min=3, n_max=7,
                                                             side=4e-6, n_a=900, n_b=700, seed=4)
    print(f"case batch of 6 lanes={lanes}", flush=True)
    eng = Engine(fm, meshes, [21] * 6, reac_k=reac_k)
    for r in range(6):
        for s in range(3):
            eng.set_tet_counts(s, counts[r][s], real=r)
    eng.set_lanes(lanes)
    eng.set_launch(64, False)
    eng.run(0.03)
    eng.run(0.08)
    for r in range(6):
        o = bind.OracleSim(fm, meshes[r], seed=21, stream=r, reac_k=reac_k[r])
        o.counts[:] = counts[r]
        o.run(0.03, 0)
        o.run(0.08, 0)
        got = np.stack([eng.get_tet_counts(s, r) for s in range(3)])
        assert np.array_equal(got, o.counts), f"realization {r}"
        assert eng.nsteps(r) == o.nsteps and eng.extents(r) == (o.events.value, o.moved.value)


def test_escaped_face_slots_with_lanes():
    """Counts above 32,767 through one face take the 32-bit escape table."""
    w = W.ab_binding_cube(n=2, side=2e-6, n_a=10_000_000, n_b=10_000_000, seed=6,
                          end_time=0.15, kf=0.0, kb=0.0)
    print("case escaped slots lanes=2", flush=True)
    eng = _engine(w, 31)
    eng.set_lanes(2)
    o = bind.OracleSim(w.fm, w.mt, seed=31)
    o.counts[:] = w.counts
    for t in (0.05, w.end_time):
        eng.run(t)
        o.run(t, 0)
        assert np.array_equal(_counts(eng, w), o.counts), f"t={t}"
    assert eng.extents()[1] == o.moved.value and o.moved.value > 100 * 32767


def test_switching_lanes_between_runs_keeps_results_and_extents():
    w = W.four_species_cube(n=4, side=4e-6, seed=5, end_time=0.04)
    print("case lanes switched between runs", flush=True)
    ref = _engine(w, 6)
    eng = _engine(w, 6)
    for k, (lanes, block) in enumerate([(4, 256), (1, 64), (8, 512), (2, 128)]):
        eng.set_lanes(lanes)
        eng.set_launch(block, False)
        t = (k + 1) * w.end_time / 4
        ref.run(t)
        eng.run(t)
        assert np.array_equal(_counts(eng, w), _counts(ref, w)), f"lanes={lanes}"
        assert eng.extents() == ref.extents() and eng.nsteps() == ref.nsteps()
    # specialisation on, then a lane count whose variant is not compiled yet
    eng.set_specialize(True)
    ref.set_specialize(True)
    eng.set_lanes(4)
    assert eng.specialized
    ref.run(w.end_time + 0.01)
    eng.run(w.end_time + 0.01)
    assert np.array_equal(_counts(eng, w), _counts(ref, w))


def test_membrane_models_refuse_several_lanes():
    w = W.calcium_membrane_cube(n=4, side=4e-6, seed=0)
    eng = _engine(w, 1)
    with pytest.raises(NotImplErr):
        eng.set_lanes(4)
    assert eng.lanes == 1
    with pytest.raises(Exception):
        eng.set_lanes(3)
