"""Host-side solver interface: every public method of the reference's solver
classes on this path exists (tests/golden/solver_method_names.json, read from
pysteps/cysteps_mpi.pyx, cysteps_solver.pyx and cysteps_dist.pyx by
tests/golden/make_solver_method_names.py), the ones this engine does not
provide raise NotImplErr, and the per-element state of steps_amd/elements.py
follows the reference's kproc rules (Diff::setDcst / setDirectionDcst /
setDiffBndActive, src/steps/mpi/tetopsplit/diff.cpp:340-455)."""
import json
from pathlib import Path

import numpy as np
import pytest

from steps_amd import _native as N
from steps_amd.elements import TriElements, VolElements
from steps_amd.solver import NOT_PROVIDED, DistTetOpSplit, TetOpSplit

NAMES = json.loads((Path(__file__).parent / "golden" / "solver_method_names.json").read_text())


def _all_names():
    out = {}
    for cls, v in NAMES.items():
        for name, line in v["methods"].items():
            out.setdefault(name, f"{v['file']}:{line}")
    return out


def test_fixture_covers_the_four_classes():
    assert {k: len(v["methods"]) for k, v in NAMES.items()} == {
        "_py_API": 76, "_py_TetAPI": 118, "_py_TetOpSplitP": 76, "_py_DistTetOpSplitP": 40}


@pytest.mark.parametrize("cls", [TetOpSplit, DistTetOpSplit])
def test_every_reference_method_exists(cls):
    missing = [f"{n} ({where})" for n, where in _all_names().items() if not callable(getattr(cls, n, None))]
    assert not missing, missing


def test_not_provided_methods_raise_notimplerr():
    names = _all_names()
    assert set(NOT_PROVIDED) <= set(names)
    for n in NOT_PROVIDED:
        with pytest.raises(N.NotImplErr):
            getattr(TetOpSplit, n)(object(), "a", "b")
    # the rest are real methods, not stubs
    implemented = [n for n in names if n not in NOT_PROVIDED]
    assert len(implemented) >= 230
    for n in implemented:
        assert "raises NotImplErr" not in (getattr(TetOpSplit, n).__doc__ or ""), n


def _line_tets(n=3):
    """n tets in a row: tet i shares face 0 with tet i+1 and face 1 with tet
    i-1; tet 0's face 2 is on diffusion-boundary direction 0."""
    nbr_same = np.zeros((n, 4), dtype=bool)
    bdir = -np.ones((n, 4), dtype=np.int64)
    for i in range(n):
        if i + 1 < n:
            nbr_same[i, 0] = True
        if i > 0:
            nbr_same[i, 1] = True
    bdir[0, 2] = 0
    return nbr_same, bdir


def test_diffusion_rules_follow_the_kproc():
    nbr_same, bdir = _line_tets()
    vel = VolElements(3, 1, 2, nbr_same, bdir)
    vel.diff_table_on(np.array([1.0, 2.0]), np.array([True, True]),
                      lambda d, s: s == 1, lambda d, s: None)
    t = vel.diff_table()
    # same-compartment faces carry the rule's constant; the boundary face only
    # for the species whose boundary is open; other faces are closed
    assert t[0, 0].tolist() == [1.0, 0.0, 0.0, 0.0]
    assert t[1, 0].tolist() == [2.0, 0.0, 2.0, 0.0]
    assert t[1, 1].tolist() == [2.0, 2.0, 0.0, 0.0]
    # setDirectionDcst on a boundary face opens it (diff.cpp:394-420)
    vel.set_direction_dcst(0, 0, 2, 7.0)
    assert vel.diff_table()[0, 0].tolist() == [1.0, 0.0, 7.0, 0.0]
    # setDcst clears the directional constants (diff.cpp:340-343)
    vel.set_dcst(0, 0, 3.0)
    assert vel.diff_table()[0, 0].tolist() == [3.0, 0.0, 3.0, 0.0]
    # setDiffBndActive(false) closes the face and resets the directional constants
    vel.set_direction_dcst(1, 0, 0, 9.0)
    vel.set_bnd_active(1, 0, False)
    assert vel.diff_table()[1, 0].tolist() == [2.0, 0.0, 0.0, 0.0]
    # an inactive rule moves nothing
    vel.dAct[0, 1] = False
    assert not vel.diff_table()[0, 1].any()


def test_reaction_table_and_clamps():
    nbr_same, bdir = _line_tets()
    vel = VolElements(3, 2, 3, nbr_same, bdir)
    vel.reac_table_on(np.array([5.0, 6.0]), np.array([True, False]))
    vel.k[0, 1] = 8.0
    vel.kact[1, 2] = True
    assert vel.reac_table().tolist() == [[5.0, 8.0, 5.0], [0.0, 0.0, 6.0]]
    vel.set_clamped(2, [0, 2], True)
    vel.set_clamped(0, 2, True)
    vel.set_clamped(2, 2, False)
    assert vel.clamp.tolist() == [4, 0, 1]
    assert vel.clamped(2, np.arange(3)).tolist() == [True, False, False]


def test_surface_table_holds_only_the_patch_entries():
    # two surface reactions (the second voltage dependent) and one GHK current;
    # triangle 1 belongs to a patch without the second reaction and the current
    held = np.array([[1, 1], [1, 0], [1, 0]], dtype=bool)
    tel = TriElements(2, held, np.array([False, True]), np.array([[3.0, 3.0], [0.0, 0.0]]), 1)
    assert tel.sreac_table().tolist() == [[3.0, 3.0], [1.0, 0.0], [1.0, 0.0]]
    tel.act[0, 0] = False
    tel.k[0, 1] = 4.0
    assert tel.sreac_table().tolist() == [[0.0, 4.0], [1.0, 0.0], [1.0, 0.0]]
