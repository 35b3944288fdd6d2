"""Compartments and diffusion boundaries on the GPU: the engine created with
stg_create_comps against the oracle's compartment layout, bit-for-bit, and
the reference's diffusion-boundary calls through the solver interface."""
import numpy as np
import pytest

from oracle import bind
from steps_amd import _native as N, geom, model as M, workloads as W
from steps_amd import rng as R
from steps_amd.engine import Engine
from steps_amd.solver import TetOpSplit, DistTetOpSplit

pytestmark = pytest.mark.gpu


def _setup(n=4, side=4e-6):
    mdl, mesh, comps, db = W.two_comp_cube(n=n, side=side)
    ct = M.flatten_comps(mdl, [["vl"], ["vr"]])
    mts, dirs = geom.comps_mesh_tables(mesh, [comps["left"], comps["right"]], {}, [db])
    return ct, mts, dirs


def _gpu_counts(eng, S, n_comp):
    return np.concatenate([np.stack([eng.get_tet_counts(s, real=c) for s in range(S)], axis=0)
                           for c in range(n_comp)], axis=1)


@pytest.mark.parametrize("spec", [False, True])
@pytest.mark.parametrize("mode", [0, 1])
def test_two_compartments_bit_exact(mode, spec):
    """spec: the period kernel compiled at run time for the model."""
    ct, mts, dirs = _setup()
    print(f"case two compartments mode={mode} specialised={spec}", flush=True)
    S = len(ct.fm.specs)
    eng = Engine(ct.fm, mts, [21, 21], streams=[0, 1], reac_k=ct.comp_k, spec_dcst=ct.comp_dcst,
                 mode=mode, comps=True)
    if spec:
        eng.set_specialize(True)
    o = bind.OracleSim.comps(ct.fm, mts, 21, ct.comp_k, ct.comp_dcst, n_bdir=len(dirs))
    a, b, c = (ct.fm.specs.index(x) for x in "ABC")
    # A open both ways at its own constant; C one-way right -> left at 1e-12
    for d, (_, src, dst) in enumerate(dirs):
        eng.set_boundary_dcst(d, a, ct.comp_dcst[src, a])
        o.bnd_dcst[d, a] = ct.comp_dcst[src, a]
        if src == 1:
            eng.set_boundary_dcst(d, c, 1e-12)
            o.bnd_dcst[d, c] = 1e-12
    rng = np.random.default_rng(2)
    init = rng.poisson(25, o.counts.shape).astype(np.int32)
    init[b, :o.comp_off[1]] = 0
    o.counts[:] = init
    for ci in range(2):
        for s in range(S):
            eng.set_tet_counts(s, init[s, o.comp_off[ci]:o.comp_off[ci + 1]], real=ci)
    assert abs(eng.update_period() - o.update_period()) == 0
    for t in (0.02, 0.07, 0.2):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(eng, S, 2), o.counts), t
    assert eng.nsteps(0) == eng.nsteps(1) == o.nsteps
    ev = sum(eng.extents(r)[0] for r in range(2))
    mv = sum(eng.extents(r)[1] for r in range(2))
    assert (ev, mv) == (o.events.value, o.moved.value)
    assert eng.boundary_dcst(0, a) == ct.comp_dcst[0, a]


def test_diffusion_boundary_through_the_solver():
    mdl, mesh, comps, db = W.two_comp_cube(n=4, side=4e-6)
    r = R.create("r123", 512)
    r.initialize(5)
    print("case solver diffusion boundary", flush=True)
    sim = TetOpSplit(mdl, mesh, r)
    with pytest.raises(N.ArgErr):
        sim.setCompSpecCount("left", "D_unknown", 10)
    sim.setCompSpecCount("left", "A", 4000)
    sim.setCompSpecCount("right", "C", 0)
    assert sim.getCompSpecCount("left", "A") == 4000
    assert not sim.getDiffBoundarySpecDiffusionActive("mid", "A")
    sim.run(0.5)
    assert sim.getCompSpecCount("right", "A") == 0          # closed boundary
    sim.setDiffBoundarySpecDiffusionActive("mid", "A", True)
    assert sim.getDiffBoundarySpecDiffusionActive("mid", "A")
    sim.run(1.0)
    left, right = sim.getCompSpecCount("left", "A"), sim.getCompSpecCount("right", "A")
    assert right > 0 and left + right + 2 * sim.getCompSpecCount("left", "C") <= 4000
    # one-way: only into the left compartment
    sim.setDiffBoundarySpecDiffusionActive("mid", "A", False)
    sim.setDiffBoundarySpecDcst("mid", "A", 1e-12, direction_comp="left")
    before = sim.getCompSpecCount("right", "A")
    sim.run(1.5)
    assert sim.getCompSpecCount("right", "A") <= before
    # a species not defined on both sides cannot cross (R is only on the right)
    with pytest.raises(N.ArgErr):
        sim.setDiffBoundarySpecDiffusionActive("mid", "R", True)
    with pytest.raises(N.ArgErr):
        sim.setCompSpecCount("left", "R", 5)
    # per-compartment reaction constants
    sim.setCompReacK("right", "split", 4.0)
    assert sim.getCompReacK("right", "split") == 4.0
    with pytest.raises(N.ArgErr):
        sim.getCompReacK("left", "split")
    assert sim.getTetVol(0) > 0


def test_dist_solver_with_compartments_matches_oracle():
    mdl, mesh, comps, db = W.two_comp_cube(n=4, side=4e-6)
    r = R.create("r123", 512)
    r.initialize(17)
    print("case dist solver compartments vs oracle", flush=True)
    sim = DistTetOpSplit(mdl, mesh, r)
    sim.setDiffBoundarySpecDiffusionActive("mid", "A", True)
    sim.setDiffBoundarySpecDcst("mid", "C", 3e-13, direction_comp="right")
    o = bind.OracleSim.comps(sim.fm, sim.mts, 17, sim._ck, sim._cd, n_bdir=len(sim._bdirs))
    for d in range(len(sim._bdirs)):
        for s in range(len(sim.fm.specs)):
            o.bnd_dcst[d, s] = sim.eng.boundary_dcst(d, s)
    sim.setCompSpecCount("left", "A", 3000)
    sim.setCompSpecCount("right", "C", 2000)
    for ci, name in enumerate(["left", "right"]):
        for s in range(len(sim.fm.specs)):
            o.counts[s, o.comp_off[ci]:o.comp_off[ci + 1]] = sim.eng.get_tet_counts(s, real=ci)
    sim.run(0.3)
    o.run(0.3, 1)
    S = len(sim.fm.specs)
    got = np.concatenate([np.stack([sim.eng.get_tet_counts(s, real=c) for s in range(S)])
                          for c in range(2)], axis=1)
    assert np.array_equal(got, o.counts)


def test_roi_totals_on_device():
    """ROIs across both compartments: device-reduced totals equal the sum of
    the per-tet counts; setROISpecCount places exactly n molecules in the ROI."""
    from steps_amd.geom import ELEM_TET, ELEM_TRI

    mdl, mesh, comps, db = W.two_comp_cube(n=4, side=4e-6)
    rng = np.random.default_rng(0)
    roi_tets = rng.choice(mesh.countTets(), 100, replace=False)
    mesh.addROI("blob", ELEM_TET, roi_tets)
    mesh.addROI("faces", ELEM_TRI, mesh.getSurfTris()[:10])
    r = R.create("r123", 512)
    r.initialize(3)
    print("case roi", flush=True)
    sim = TetOpSplit(mdl, mesh, r)
    sim.setCompSpecCount("left", "A", 5000)
    sim.setCompSpecCount("right", "A", 3000)
    counts = np.array([sim.getTetSpecCount(int(t), "A") for t in sorted(roi_tets)])
    assert sim.getROISpecCount("blob", "A") == counts.sum()
    c = np.zeros(100)
    sim.getROITetSpecCountsNP("blob", "A", c)
    assert np.array_equal(c, counts)
    assert abs(sim.getROIVol("blob") - mesh.tet_vol[roi_tets].sum()) < 1e-30
    sim.setROISpecCount("blob", "C", 777)
    assert sim.getROISpecCount("blob", "C") == 777
    with pytest.raises(N.ArgErr):
        sim.getROISpecCount("faces", "A")      # a triangle ROI has no volume species
    with pytest.raises(N.ArgErr):
        sim.getROIVol("nope")
    assert abs(sim.getROIArea("faces") - mesh.tri_area[mesh.getSurfTris()[:10]].sum()) < 1e-30


def test_solver_specialises_large_meshes(monkeypatch):
    """The solver compiles the model's kernel for meshes at or above its size
    threshold; lowered here, the results equal the generic kernel's."""
    import steps_amd.solver as S

    out = []
    for threshold in (10**9, 0):
        monkeypatch.setattr(S, "SPECIALIZE_MIN_TETS", threshold)
        mdl, mesh, comps, db = W.two_comp_cube(n=4, side=4e-6)
        r = R.create("r123", 512)
        r.initialize(23)
        sim = DistTetOpSplit(mdl, mesh, r)
        assert sim.eng.specialized == (threshold == 0)
        sim.setDiffBoundarySpecDiffusionActive("mid", "A", True)
        sim.setCompSpecCount("left", "A", 3000)
        sim.setCompSpecCount("right", "C", 2000)
        sim.run(0.2)
        out.append([sim.getCompSpecCount(c, s) for c in ("left", "right") for s in ("A", "B", "C")])
    print("case solver specialisation threshold", out, flush=True)
    assert out[0] == out[1]


def _membrane_engine_and_oracle(seed, mode, binding=True, ghk=True):
    mdl, mesh, comps, patch = W.two_comp_membrane(n=4, side=4e-6, binding=binding, ghk=ghk)
    cl = [comps["inner"], comps["outer"]]
    ct = M.flatten_comps(mdl, [c.volsys for c in cl], [(patch.surfsys, 0, 1)])
    mts, _ = geom.comps_mesh_tables(mesh, cl, [(patch.tris, 0, 1)], [])
    S, SS = len(ct.fm.specs), len(ct.fm.sspecs)
    eng = Engine(ct.fm, mts, [seed, seed], streams=[0, 1], reac_k=ct.comp_k, spec_dcst=ct.comp_dcst,
                 mode=mode, comps=True)
    o = bind.OracleSim.comps(ct.fm, mts, seed, ct.comp_k, ct.comp_dcst)
    rng = np.random.default_rng(seed)
    init = np.zeros((S, o.comp_off[2]), dtype=np.int32)
    a, b, ca = (ct.fm.specs.index(x) for x in ("A", "B", "Ca"))
    init[a, o.comp_off[1]:] = rng.poisson(20, o.comp_off[2] - o.comp_off[1])
    init[b, :o.comp_off[1]] = rng.poisson(10, o.comp_off[1])
    init[ca, :] = rng.poisson(15, o.comp_off[2])
    tri = np.zeros((SS, len(patch.tris)), dtype=np.int32)
    tri[ct.fm.sspecs.index("T")] = rng.integers(1, 4, len(patch.tris))
    tri[ct.fm.sspecs.index("Ch")] = rng.integers(0, 3, len(patch.tris))
    o.counts[:] = init
    o.tri_counts[:SS, :len(patch.tris)] = tri
    for ci in range(2):
        for s in range(S):
            eng.set_tet_counts(s, init[s, o.comp_off[ci]:o.comp_off[ci + 1]], real=ci)
    for q in range(SS):
        eng.set_tri_counts(q, tri[q], real=0)
    v = np.linspace(-0.08, 0.02, len(patch.tris))
    eng.set_tri_v(v, real=0)
    o.tri_v[:len(patch.tris)] = v
    return eng, o, ct, S, SS, len(patch.tris)


@pytest.mark.parametrize("mode", [0, 1])
def test_membrane_between_compartments_bit_exact(mode):
    """Transport reactions with outer-compartment reactants and products and a
    GHK current reading the outer tet, on a membrane between two compartments
    (every cut triangle joins an inner and an outer tet into one SSA group),
    with a binding reaction on the inner side: GPU against the oracle."""
    eng, o, ct, S, SS, T = _membrane_engine_and_oracle(7, mode)
    print(f"case membrane between compartments mode={mode}", flush=True)
    for t in (0.01, 0.04, 0.1):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(eng, S, 2), o.counts), t
        assert np.array_equal(np.stack([eng.get_tri_counts(q, 0) for q in range(SS)]),
                              o.tri_counts[:SS, :T]), t
    ev = sum(eng.extents(r)[0] for r in range(2))
    mv = sum(eng.extents(r)[1] for r in range(2))
    assert (ev, mv) == (o.events.value, o.moved.value) and ev > 0
    ca = ct.fm.specs.index("Ca")
    assert o.counts[ca].sum() == _gpu_counts(eng, S, 2)[ca].sum()


def test_membrane_groups_with_tables_and_graph():
    """The group path with per-element tables (clamped calcium in a few tets,
    per-tet binding constants) and graph replay at block 128."""
    eng, o, ct, S, SS, T = _membrane_engine_and_oracle(11, 0)
    print("case membrane groups with tables and graph", flush=True)
    n0 = o.comp_off[1]
    rng = np.random.default_rng(1)
    clamp = np.where(rng.random(n0) < 0.2, 1 << ct.fm.specs.index("Ca"), 0).astype(np.uint32)
    k = ct.comp_k[0][:, None] * rng.uniform(0.5, 2.0, size=(len(ct.fm.reacs), n0))
    eng.set_tet_clamp(clamp, real=0)
    eng.set_tet_reac_table(k, real=0)
    eng.set_launch(128, True)
    full_clamp = np.zeros(o.comp_off[2], dtype=np.uint32)
    full_clamp[:n0] = clamp
    full_k = np.concatenate([k, np.repeat(ct.comp_k[1][:, None], o.comp_off[2] - n0, axis=1)], axis=1)
    o.set_tables(tet_clamp=full_clamp, tet_k=full_k)
    for t in (0.02, 0.05):
        eng.run(t)
        o.run(t, 0)
        assert np.array_equal(_gpu_counts(eng, S, 2), o.counts), t


def test_membrane_between_compartments_with_8bit_slots():
    """The SSA-group path with 8-bit face slots: identical to the oracle."""
    eng, o, ct, S, SS, T = _membrane_engine_and_oracle(7, 0)
    eng.set_slot_width(1)
    print("case membrane between compartments, 8-bit slots", flush=True)
    for t in (0.01, 0.04):
        eng.run(t)
        o.run(t, 0)
        assert np.array_equal(_gpu_counts(eng, S, 2), o.counts), t
        assert np.array_equal(np.stack([eng.get_tri_counts(q, 0) for q in range(SS)]),
                              o.tri_counts[:SS, :T]), t
