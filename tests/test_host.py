"""Host-side logic that needs no GPU: model flattening, mesh tables,
initial placement and solver name mapping."""
import numpy as np
import pytest

from steps_amd import geom, model as M, rng as R, solver as S, workloads as W
from steps_amd import _native as N


def test_flatten_stoichiometry_and_order():
    mdl = M.Model()
    A, B, C, P = (M.Spec(x, mdl) for x in "ABCP")
    vs = M.Volsys("v", mdl)
    M.Reac("r1", vs, lhs=[A, A, B], rhs=[C], kcst=2.0)
    M.Reac("r2", vs, lhs=[C], rhs=[A, B, B], kcst=3.0)
    M.Diff("dA", vs, A, 1e-12)
    ss = M.Surfsys("s", mdl)
    M.SReac("sr", ss, ilhs=[A], slhs=[P], srhs=[P], irhs=[C], kcst=5.0)
    fm = M.flatten(mdl, ["v"], ["s"])
    assert fm.specs == ["A", "B", "C"] and fm.sspecs == ["P"]
    assert fm.reac_lhs.tolist() == [[2, 1, 0], [0, 0, 1]]
    assert fm.reac_upd.tolist() == [[-2, -1, 1], [1, 2, -1]]
    assert fm.spec_dcst.tolist() == [1e-12, 0.0, 0.0]
    assert fm.sreac_lhs_i.tolist() == [[1, 0, 0]] and fm.sreac_upd_i.tolist() == [[-1, 0, 1]]
    assert fm.sreac_lhs_s.tolist() == [[1]] and fm.sreac_upd_s.tolist() == [[0]]


def test_flatten_needs_an_outer_compartment_and_rejects_high_order():
    from steps_amd import _native as N

    mdl = M.Model()
    A = M.Spec("A", mdl)
    ss = M.Surfsys("s", mdl)
    M.SReac("sr", ss, olhs=[A], kcst=1.0)
    with pytest.raises(N.ArgErr, match="outer compartment"):
        M.flatten(mdl, [], ["s"])                      # a patch without an outer compartment
    with pytest.raises(N.ArgErr, match="outer compartment"):
        M.flatten_comps(mdl, [[], []], [(["s"], 0, None)])
    ct = M.flatten_comps(mdl, [[], []], [(["s"], 0, 1)])
    assert ct.fm.sreac_lhs_o.tolist() == [[1]] and ct.fm.sreac_upd_o.tolist() == [[-1]]
    assert ct.fm.outer() and ct.fm.patch_sreacs == [{0}]
    with pytest.raises(ValueError):
        M.SReac("both", ss, ilhs=[A], olhs=[A], kcst=1.0)
    vs = M.Volsys("v", mdl)
    with pytest.raises(ValueError):
        M.Reac("r", vs, lhs=[A] * 5, rhs=[], kcst=1.0)


def test_mesh_tables_close_faces_to_other_compartments():
    mesh = geom.kuhn_cube(2, 1e-6)
    half = np.nonzero(mesh.tet_bary[:, 0] < 0.5e-6)[0]
    comp = geom.TmComp("left", mesh, half)
    mt = geom.mesh_tables(mesh, comp)
    assert len(mt.vol) == len(half)
    assert np.all((mt.nbr >= 0) == (mt.face_g > 0))
    for t in range(len(half)):
        for f in range(4):
            j = mt.nbr[t, f]
            if j >= 0:
                assert mt.nbr[j, mt.nbr_face[t, f]] == t


def test_patch_tables_map_triangles_to_inner_tets():
    mesh = geom.kuhn_cube(3, 1e-6)
    comp = geom.TmComp("c", mesh)
    tris = mesh.getSurfTris()
    mt = geom.mesh_tables(mesh, comp, tris)
    assert len(tris) == 6 * 9 * 2
    for i, tri in enumerate(tris):
        assert tri in mesh.tet_tri[mt.tri_tet[i]]
    assert np.isclose(mt.tri_area.sum(), 6e-12)


def test_distribute_counts_is_fair_and_exact():
    rng = np.random.default_rng(0)
    vol = rng.uniform(1, 3, 1000)
    for n in (0, 1, 999, 1000.0, 12345.6):
        c = S.distribute_counts(n, vol, rng)
        assert c.sum() in (int(np.floor(n)), int(np.floor(n)) + 1)
        share = np.floor(c.sum() * vol / vol.sum())
        assert np.all((c == share) | (c == share + 1))
    with pytest.raises(ValueError):
        S.distribute_counts(-1, vol, rng)


def test_solver_name_mapping_and_rng_factory():
    assert S.get_solver("TetOpSplit") is S.TetOpSplit
    assert S.get_solver("DistTetOpSplit") is S.DistTetOpSplit
    with pytest.raises(N.ArgErr):
        S.get_solver("Tetexact")
    r = R.create("r123", 512)
    r.initialize(23412)
    assert r.seed == 23412
    with pytest.raises(ValueError):
        R.create("nonsense", 10)


def test_workloads_shapes():
    w = W.ab_binding_cube(n=4, side=2e-6, n_a=100, n_b=50, seed=1)
    assert w.counts.shape == (3, 384) and w.counts[0].sum() == 100 and w.counts[1].sum() == 50
    w2 = W.calcium_membrane_cube(n=4, side=1e-6, seed=1)
    assert len(w2.mt.tri_tet) == 6 * 16 * 2
    assert w2.tri_counts[w2.fm.sspec_index("Chan")].sum() == 32
    fm, meshes, counts, k, info = W.ab_binding_ensemble(n_real=6, n_min=2, n_max=3, n_a=10, n_b=10)
    assert len(meshes) == 6 and k.shape == (6, 2)
    assert all(c[0].sum() == 10 for c in counts)


def test_memb_collects_patch_triangles_and_conduction_volume():
    mesh = geom.kuhn_cube(3, 3e-6)
    surf = mesh.getSurfTris()
    inner = geom.TmComp("inner", mesh, np.arange(60))
    other = geom.TmComp("other", mesh, np.arange(60, mesh.countTets()))
    p1 = geom.TmPatch("a", mesh, surf[:10], inner)
    p2 = geom.TmPatch("b", mesh, surf[10:25], inner)
    m = geom.Memb("memb", mesh, [p1, p2])
    assert np.array_equal(m.getAllTriIndices(), surf[:25])
    assert np.array_equal(m.getAllVolTetIndices(), np.arange(60))
    m2 = geom.Memb("memb2", mesh, [p1], supplementary_comps=[other])
    assert len(m2.getAllVolTetIndices()) == mesh.countTets()
    assert mesh.membs["memb"] is m
    with pytest.raises(ValueError):
        geom.Memb("dup", mesh, [p1, p1])


def test_solver_rejects_unknown_membrane_solver():
    mesh = geom.kuhn_cube(2, 1e-6)
    geom.TmComp("c", mesh)
    mdl = M.Model()
    M.Volsys("vsys", mdl)
    with pytest.raises(N.ArgErr):
        S.TetOpSplit(mdl, mesh, R.create("r123", 16), calcMembPot=9)


def test_channel_and_current_tables():
    mdl = M.Model()
    ch = M.Chan("K", mdl)
    n0, n1 = M.ChanState("n0", mdl, ch), M.ChanState("n1", mdl, ch)
    k = M.Spec("Kion", mdl, valence=1)
    vs = M.Volsys("vsys", mdl)
    M.Diff("dK", vs, k, 1e-10)
    ss = M.Surfsys("ssys", mdl)
    M.VDepSReac("open", ss, slhs=[n0], srhs=[n1], k=lambda v: 100.0)
    M.OhmicCurr("leak", ss, n0, -0.07, 1e-12)
    g = M.GHKcurr("ghk", ss, n1, k, virtual_oconc=4e-3, vshift=0.001)
    with pytest.raises(ValueError):
        M.flatten(mdl, ["vsys"], ["ssys"])          # no permeability yet
    g.setP(1e-19)
    fm = M.flatten(mdl, ["vsys"], ["ssys"])
    assert fm.sspecs == ["n0", "n1"] and fm.specs == ["Kion"]
    assert fm.ohmic_chan.tolist() == [0] and fm.ohmic_g.tolist() == [1e-12]
    assert fm.ghk_chan.tolist() == [1] and fm.ghk_ion.tolist() == [0]
    assert fm.ghk_valence.tolist() == [1] and fm.ghk_vshift.tolist() == [0.001]
    assert ch.states == [n0, n1] and n1.getChan() is ch
    with pytest.raises(ValueError):
        M.GHKcurr("bad", ss, n1, M.Spec("neutral", mdl))   # valence zero
    g2 = M.GHKcurr("nooc", ss, n1, k)
    g2.setP(1e-19)
    from steps_amd import _native as N
    with pytest.raises(N.ArgErr, match="outer compartment"):
        M.flatten(mdl, ["vsys"], ["ssys"])          # reads the outer compartment: the patch has none
