"""The solver interface on the GPU: several patches on one compartment with an
outer compartment, per-element setters, host-computed propensities and the
GHK / ohmic current getters.

Engine results are compared bit for bit with the oracle given the element
tables the solver derived; the propensity getters are checked against the
textbook formulas evaluated here from counts and mesh geometry (Reac::rate,
SReac::rate, Diff::rate), and the current getters against the ions that
crossed (every GHK event moves one ion and z elementary charges).
"""
import numpy as np
import pytest

from oracle import bind
from steps_amd import geom, model as M, rng as R
from steps_amd import workloads as W
from steps_amd.solver import AVOGADRO, EF_DEFAULT, TetOpSplit

pytestmark = pytest.mark.gpu

E_CHARGE = 1.602176487e-19


def _rng(seed):
    r = R.create("mt19937", 512)
    r.initialize(seed)
    return r


def _two_patch_sim(seed=5):
    mdl, mesh, comps, patches = W.two_patch_membrane(n=4, side=4e-6)
    sim = TetOpSplit(mdl, mesh, _rng(seed))
    sim.setCompSpecCount("outer", "A", 3000)
    sim.setCompSpecCount("inner", "A", 500)
    sim.setCompSpecCount("inner", "B", 800)
    sim.setCompSpecCount("inner", "Ca", 600)
    sim.setCompSpecCount("outer", "Ca", 900)
    pa, pb = patches["pa"], patches["pb"]
    sim.setPatchSpecCount("pa", "T", 2 * len(pa.tris))
    sim.setPatchSpecCount("pa", "Ch", len(pa.tris))
    sim.setPatchSpecCount("pb", "T2", 3 * len(pb.tris))
    return sim, mesh, comps, patches


def _oracle(sim, seed):
    """The oracle of the solver's engine state: same counts, and the tables
    the solver's per-element state stands for in every compartment."""
    fm, mts = sim.fm, sim.mts
    nc = len(mts)
    o = bind.OracleSim.comps(fm, mts, seed, sim._ck * sim._cract, sim._cd * sim._cdact,
                             n_bdir=len(sim._bdirs))
    off = o.comp_off
    toff = np.concatenate([[0], np.cumsum([len(m.tri_tet) for m in mts])])
    S, SS, R_, RS, NG = (len(fm.specs), len(fm.sspecs), len(fm.reacs), len(fm.sreacs), len(fm.ghk_chan))
    for c in range(nc):
        for s in range(S):
            o.counts[s, off[c]:off[c + 1]] = sim.eng.get_tet_counts(s, real=c)
        for q in range(SS):
            o.tri_counts[q, toff[c]:toff[c + 1]] = sim.eng.get_tri_counts(q, real=c)
    k, fD, clamp, sk, eD, tclamp = [], [], [], [], [], []
    for c in range(nc):
        n, m = len(mts[c].vol), len(mts[c].tri_tet)
        vel, tel = sim._vel[c], sim._tel[c]
        k.append(vel.reac_table() if vel is not None and vel.k is not None
                 else np.repeat((sim._ck[c] * sim._cract[c])[:, None], n, axis=1))
        if vel is not None and vel.dD is not None:
            fD.append(vel.diff_table())
        else:
            same = (mts[c].nbr >= 0) & (mts[c].nbr_comp == c)
            fD.append(np.where(same[None], (sim._cd[c] * sim._cdact[c])[:, None, None], 0.0))
        clamp.append(vel.clamp if vel is not None and vel.clamp is not None else np.zeros(n, np.uint32))
        if m:
            assert tel is not None      # several patches on one compartment: table mode
            sk.append(tel.sreac_table())
            eD.append(sim._tri_eD(c))
            tclamp.append(tel.clamp if tel.clamp is not None else np.zeros(m, np.uint32))
    o.set_tables(tet_k=np.concatenate(k, axis=1), tet_fD=np.concatenate(fD, axis=1),
                 tet_clamp=np.concatenate(clamp), tri_sk=np.concatenate(sk, axis=1),
                 tri_eD=np.concatenate(eD, axis=1), tri_clamp=np.concatenate(tclamp))
    return o, off, toff


def _check(sim, o, off, toff, t):
    sim.run(t)
    o.run(t, 0)
    for c in range(len(sim.mts)):
        for s in range(len(sim.fm.specs)):
            assert np.array_equal(sim.eng.get_tet_counts(s, real=c), o.counts[s, off[c]:off[c + 1]]), (t, c, s)
        for q in range(len(sim.fm.sspecs)):
            assert np.array_equal(sim.eng.get_tri_counts(q, real=c), o.tri_counts[q, toff[c]:toff[c + 1]]), \
                (t, c, q)
    assert sim.getReacExtent() == o.events.value and sim.getDiffExtent() == o.moved.value


def test_two_patches_with_an_outer_compartment_bit_exact():
    """Patches 'pa' and 'pb' with different surface systems on the same inner
    compartment, both reaching the outer compartment: each triangle runs only
    its patch's reactions and GHK current, surface diffusion stays within
    'pb', and the counts equal the oracle's."""
    sim, mesh, comps, patches = _two_patch_sim(5)
    print("case two patches with an outer compartment", flush=True)
    assert sim.getNPatches() == 2 and sim.getNComps() == 2
    o, off, toff = _oracle(sim, 5)
    for t in (0.01, 0.05):
        _check(sim, o, off, toff, t)
    # the patches' own species: pa never sees T2 / T3, pb never T / Ch
    assert sim.getPatchSpecCount("pb", "T2") + sim.getPatchSpecCount("pb", "T3") == 3 * len(patches["pb"].tris)
    with pytest.raises(ValueError):
        sim.getPatchSpecCount("pa", "T2")
    tot = sim.getCompSpecCount("inner", "A") + sim.getCompSpecCount("outer", "A") \
        + sim.getCompSpecCount("inner", "C")
    assert tot == 3500                              # A is conserved (C holds bound A)


def test_element_setters_bit_exact():
    """Per-tet and per-triangle constants, activation flags, directional
    diffusion constants and clamps set through the solver reach the engine as
    the tables the oracle runs with."""
    sim, mesh, comps, patches = _two_patch_sim(9)
    inner, outer = comps["inner"].tets, comps["outer"].tets
    pa, pb = patches["pa"].tris, patches["pb"].tris
    rng = np.random.default_rng(2)
    print("case element setters", flush=True)
    for t in inner[rng.random(len(inner)) < 0.3]:
        sim.setTetReacK(int(t), "bind", 3e6)
    for t in inner[rng.random(len(inner)) < 0.2]:
        sim.setTetReacActive(int(t), "unbind", False)
    t0 = int(inner[3])
    nb = [int(x) for x in mesh.tet_nbr[t0] if x >= 0 and x in set(inner.tolist())]
    sim.setTetDiffD(t0, "dA", 5e-11, nb[0])
    assert sim.getTetDiffD(t0, "dA", nb[0]) == 5e-11 and sim.getTetDiffD(t0, "dA") == 1e-11
    for t in outer[:10]:
        sim.setTetDiffActive(int(t), "dCa", False)
    for t in inner[rng.random(len(inner)) < 0.1]:
        sim.setTetSpecClamped(int(t), "Ca", True)
    for t in pa[rng.random(len(pa)) < 0.3]:
        sim.setTriSReacK(int(t), "in", 6e9)
    for t in pb[rng.random(len(pb)) < 0.3]:
        sim.setTriSReacActive(int(t), "flip", False)
    sim.setTriSDiffD(int(pb[0]), "dT2", 1e-12)
    sim.setPatchSpecClamped("pa", "Ch", True)
    o, off, toff = _oracle(sim, 9)
    ca_clamped = [int(t) for t in inner if sim.getTetSpecClamped(int(t), "Ca")]
    before = {t: sim.getTetSpecCount(t, "Ca") for t in ca_clamped}
    for t in (0.01, 0.03):
        _check(sim, o, off, toff, t)
    assert all(sim.getTetSpecCount(t, "Ca") == n for t, n in before.items())
    assert sim.getPatchSpecClamped("pa", "Ch") and not sim.getPatchSpecClamped("pb", "T2")
    # compartment-level calls overwrite every element (_setCompReacK and the like)
    sim.setCompReacK("inner", "bind", 2e6)
    assert all(sim.getTetReacK(int(t), "bind") == 2e6 for t in inner[:20])
    sim.setCompReacActive("inner", "unbind", True)
    assert all(sim.getTetReacActive(int(t), "unbind") for t in inner)


def test_propensity_getters_follow_the_formulas():
    """getTetReacA / getCompReacA / getTriSReacA / getPatchSReacA /
    getTetDiffA against h c computed here: for A + B -> C, k nA nB /
    (1e3 V N_A); for T + A(outer), k nT nA(outer) / (1e3 V_outer N_A); for a
    surface-only reaction, k nT2; diffusion, n sum_f D A_f / (V d_f)."""
    sim, mesh, comps, patches = _two_patch_sim(3)
    print("case propensity getters", flush=True)
    sim.run(0.005)
    inner = comps["inner"].tets
    tot = 0.0
    for t in inner:
        t = int(t)
        na, nb = sim.getTetSpecCount(t, "A"), sim.getTetSpecCount(t, "B")
        v = mesh.tet_vol[t]
        want = 1e6 * na * nb / (1e3 * v * AVOGADRO)
        assert sim.getTetReacA(t, "bind") == pytest.approx(want, rel=1e-12, abs=0)
        assert sim.getTetReacH(t, "bind") == na * nb
        tot += want
    assert sim.getCompReacA("inner", "bind") == pytest.approx(tot, rel=1e-10)
    assert sim.getCompReacH("inner", "bind") == sum(sim.getTetSpecCount(int(t), "A")
                                                    * sim.getTetSpecCount(int(t), "B") for t in inner)
    otet = {}
    for t in patches["pa"].tris:
        a, b = mesh.tri_tets[t]
        otet[int(t)] = int(b) if int(a) in set(inner.tolist()) else int(a)
    ptot = 0.0
    for t, o_ in otet.items():
        want = 2e9 * sim.getTriSpecCount(t, "T") * sim.getTetSpecCount(o_, "A") / (1e3 * mesh.tet_vol[o_] * AVOGADRO)
        assert sim.getTriSReacA(t, "in") == pytest.approx(want, rel=1e-12, abs=0)
        ptot += want
    assert sim.getPatchSReacA("pa", "in") == pytest.approx(ptot, rel=1e-10)
    t = int(patches["pb"].tris[0])
    assert sim.getTriSReacA(t, "flip") == pytest.approx(5.0 * sim.getTriSpecCount(t, "T2"), rel=1e-14)
    assert sim.getTriSReacC(t, "flip") == 5.0
    for t in inner[:30]:
        t = int(t)
        rate = 0.0
        for f in range(4):
            nbr = mesh.tet_nbr[t, f]
            if nbr >= 0 and nbr in set(inner.tolist()):
                rate += 1e-11 * mesh.tet_face_area[t, f] / (mesh.tet_vol[t] * mesh.tet_face_dist[t, f])
        want = rate * sim.getTetSpecCount(t, "A")
        assert sim.getTetDiffA(t, "dA") == pytest.approx(want, rel=1e-12, abs=1e-300)


def _ghk_cell(n=4, side=2e-6):
    mesh = geom.kuhn_cube(n, side)
    cyto = geom.TmComp("cyto", mesh)
    cyto.addVolsys("vsys")
    surf = mesh.getSurfTris()
    patch = geom.TmPatch("memb", mesh, surf, cyto)
    patch.addSurfsys("ssys")
    geom.Memb("membrane", mesh, [patch])
    mdl = M.Model()
    ch = M.Chan("CaChan", mdl)
    op = M.ChanState("Open", mdl, ch)
    leak = M.ChanState("Leak", mdl, M.Chan("LChan", mdl))
    ca = M.Spec("Ca", mdl, valence=2)
    vs = M.Volsys("vsys", mdl)
    M.Diff("dCa", vs, ca, 2e-10)
    ss = M.Surfsys("ssys", mdl)
    g = M.GHKcurr("ghk", ss, op, ca, computeflux=True, virtual_oconc=2e-3)
    g.setPInfo(2e-12, -65e-3, 293.15, 2e-3, 1e-7)
    M.OhmicCurr("oc", ss, leak, -70e-3, 5e-12)
    return mdl, mesh, surf


def test_ghk_and_ohmic_current_getters():
    """Over each EField step the GHK currents of all triangles times the step,
    over z e, equal the calcium ions that entered (every event moves one ion);
    the batch getters agree with the per-triangle ones; getTriI is GHK plus
    ohmic, and the ohmic current is n g (V - erev) with a per-triangle erev."""
    mdl, mesh, surf = _ghk_cell()
    sim = TetOpSplit(mdl, mesh, _rng(4), EF_DEFAULT)
    efdt = 2e-5
    sim.setEfieldDT(efdt)
    sim.setCompSpecCount("cyto", "Ca", 4000)
    for t in surf:
        sim.setTriSpecCount(int(t), "Open", 3)
        sim.setTriSpecCount(int(t), "Leak", 2)
    print("case ghk and ohmic current getters", flush=True)
    assert sim.getTriGHKI(int(surf[0])) == 0.0          # no EField step yet
    buf = np.zeros(len(surf))
    for step in range(1, 4):
        before = sim.getCompSpecCount("cyto", "Ca")
        sim.run(step * efdt)
        entered = sim.getCompSpecCount("cyto", "Ca") - before
        sim.getBatchTriGHKIsNP(surf, "ghk", buf)
        charge = buf.sum() * efdt / E_CHARGE
        assert round(charge) == -2 * entered and abs(charge - round(charge)) < 1e-6, step
        assert sim.sumBatchTriGHKIsNP(surf, "ghk") == pytest.approx(buf.sum(), rel=1e-14, abs=1e-30)
        assert buf[5] == sim.getTriGHKI(int(surf[5]), "ghk") == sim.getTriGHKI(int(surf[5]))
    both = np.zeros(2 * len(surf))
    sim.getBatchTriBatchGHKIsNP(surf, ["ghk", "ghk"], both)
    assert np.array_equal(both[0::2], buf) and np.array_equal(both[1::2], buf)
    t = int(surf[7])
    v = sim.getTriV(t)
    assert sim.getTriOhmicI(t, "oc") == pytest.approx(2 * 5e-12 * (v + 70e-3), rel=1e-12)
    sim.setTriOhmicErev(t, "oc", -50e-3)
    assert sim.getTriOhmicErev(t, "oc") == -50e-3 and sim.getTriOhmicErev(int(surf[8]), "oc") == -70e-3
    assert sim.getTriOhmicI(t, "oc") == pytest.approx(2 * 5e-12 * (v + 50e-3), rel=1e-12)
    assert sim.getTriI(t) == pytest.approx(sim.getTriGHKI(t) + sim.getTriOhmicI(t), rel=1e-14)
    oh = np.zeros(len(surf))
    sim.getBatchTriOhmicIsNP(surf, "oc", oh)
    assert oh[7] == sim.getTriOhmicI(t, "oc")


def test_per_triangle_erev_and_capacitance_match_membrane_wide_values():
    """Setting every triangle's ohmic reversal potential (Tri::setOCerev) or
    capacitance (applyTriCapacitance) to one value gives the run the
    membrane-wide setting gives."""
    runs = []
    for per_tri in (False, True):
        mdl, mesh, surf = _ghk_cell()
        sim = TetOpSplit(mdl, mesh, _rng(6), EF_DEFAULT)
        sim.setEfieldDT(2e-5)
        sim.setCompSpecCount("cyto", "Ca", 2000)
        for t in surf:
            sim.setTriSpecCount(int(t), "Leak", 4)
        if per_tri:
            for t in surf:
                sim.setTriCapac(int(t), 0.02)
        else:
            sim.setMembCapac("membrane", 0.02)
        print(f"case per-triangle parameters per_tri={per_tri}", flush=True)
        sim.run(1e-4)
        v = np.zeros(mesh.countVertices())
        sim.getBatchVertVsNP(np.unique(mesh.tris[surf]), v[:len(np.unique(mesh.tris[surf]))])
        runs.append(v.copy())
    assert np.allclose(runs[0], runs[1], rtol=1e-9, atol=1e-15)
    assert np.any(runs[0] != -65e-3)


@pytest.mark.parametrize("split", [False, True])
def test_sdiff_boundary_bit_exact(split):
    """The set-up of the reference's test_sdiff_bugfix.py on a cube (one
    compartment, or the two patches on two compartments): the boundary opened
    for X, 0.1 x the constant toward patchA; counts equal the oracle's, X is
    conserved, patchA never holds A, and closing the boundary again stops the
    crossing."""
    mdl, mesh, comps, patches, sdb = W.sdiff_boundary_patches(split=split)
    sim = TetOpSplit(mdl, mesh, _rng(11))
    print(f"case sdiff boundary split={split}", flush=True)
    pa, pb = patches["patchA"].tris, patches["patchB"].tris
    sim.setPatchSpecCount("patchB", "X", 40 * len(pb))
    assert not sim.getSDiffBoundarySpecDiffusionActive("sdiffb", "X")
    sim.setSDiffBoundarySpecDiffusionActive("sdiffb", "X", True)
    assert sim.getSDiffBoundarySpecDiffusionActive("sdiffb", "X")
    sim.setSDiffBoundarySpecDcst("sdiffb", "X", 1e-13, "patchA")
    ta, tb = (int(t) for t in sdb.bar_tris[0])
    assert sim.getTriSDiffD(tb, "diffX2", ta) == 1e-13 and sim.getTriSDiffD(tb, "diffX2") == 1e-12
    assert sim.getTriSDiffD(ta, "diffX1", tb) == 1e-12
    with pytest.raises(ValueError):
        sim.getTriSDiffD(ta, "diffX2")                      # the rule of the other patch
    with pytest.raises(ValueError):
        sim.setSDiffBoundarySpecDiffusionActive("sdiffb", "V", True)   # not a species of the patches
    o, off, toff = _oracle(sim, 11)
    for t in (0.5, 2.0, 5.0):
        _check(sim, o, off, toff, t)
    n = 40 * len(pb)
    assert sim.getPatchSpecCount("patchA", "X") + sim.getPatchSpecCount("patchB", "X") == n
    assert sim.getPatchSpecCount("patchA", "X") > 0 and sim.getPatchSpecCount("patchA", "A") == 0
    # the deprecated spelling closes it again; nothing crosses afterwards
    sim.setSDiffBoundaryDiffusionActive("sdiffb", "X", False)
    assert not sim.getSDiffBoundaryDiffusionActive("sdiffb", "X")
    na = sim.getPatchSpecCount("patchA", "X")
    sim.run(8.0)
    assert sim.getPatchSpecCount("patchA", "X") == na


def test_reset_restores_the_model_constants():
    """TetOpSplitP::reset runs every kproc's reset: constants back to the
    model's, reactions and diffusion active, clamps released, boundaries
    closed. After reset the run equals an oracle built from the model."""
    sim, mesh, comps, patches = _two_patch_sim(13)
    inner = comps["inner"].tets
    pb = patches["pb"].tris
    sim.setCompReacK("inner", "bind", 3e6)
    sim.setTetReacActive(int(inner[0]), "unbind", False)
    sim.setCompDiffD("outer", "dA", 4e-11)
    sim.setTetSpecClamped(int(inner[1]), "Ca", True)
    sim.setPatchSReacK("pb", "leak", 1e9)
    sim.setTriSDiffD(int(pb[0]), "dT2", 5e-13)
    sim.setPatchSpecClamped("pa", "Ch", True)
    print("case reset restores constants", flush=True)
    sim.reset()
    assert sim.getCompReacK("inner", "bind") == 1e6
    assert sim.getTetReacActive(int(inner[0]), "unbind")
    assert sim.getCompDiffD("outer", "dA") == 1e-11
    assert not sim.getTetSpecClamped(int(inner[1]), "Ca")
    assert sim.getPatchSReacK("pb", "leak") == 3e9
    assert sim.getTriSDiffD(int(pb[0]), "dT2") == 2e-13
    assert not sim.getPatchSpecClamped("pa", "Ch")
    sim.setCompSpecCount("outer", "A", 3000)
    sim.setCompSpecCount("inner", "B", 800)
    sim.setPatchSpecCount("pb", "T2", 3 * len(pb))
    sim.setPatchSpecCount("pa", "T", 2 * len(patches["pa"].tris))
    o, off, toff = _oracle(sim, 13)
    # the oracle's tables are the model's
    ref, *_ = _two_patch_sim(13)
    assert np.array_equal(sim._ck, ref._ck) and np.array_equal(sim._cd, ref._cd)
    assert all(np.array_equal(sim._tri_eD(c), ref._tri_eD(c)) for c in range(len(sim.mts)))
    for t in (0.01, 0.03):
        _check(sim, o, off, toff, t)
