"""GPU EField against the FP64 oracle (oracle/efield_ref.py) and Rallpack1.

Tolerances: the mixed-precision solve (FP32 CG, FP64 dots and refinement)
must match the FP64 sparse-LU solution to 1e-6 relative in every step's dV;
the FP64 CG mode to 1e-9. Rallpack1: RMS error below 1 mV against the golden
traces, as test/validation/API_2/validation_efield/test_rallpack1.py requires.
"""
import numpy as np
import pytest
from scipy.sparse.linalg import spsolve

from oracle import bind
from oracle.efield_ref import EFieldOracle
from steps_amd import geom, model as M
from steps_amd.efield import EField, FP64, MIXED
from steps_amd.engine import Engine
from test_efield_oracle import RALL, rallpack1_setup, read_trace, sample_verts

pytestmark = pytest.mark.gpu


def _cube_case(n=6, side=3e-6):
    mesh = geom.kuhn_cube(n, side)
    tris = mesh.tris[mesh.getSurfTris()]
    return mesh, tris


@pytest.mark.parametrize("precision,tol", [(MIXED, 1e-6), (FP64, 1e-9)])
def test_steps_match_fp64_oracle(precision, tol):
    mesh, tris = _cube_case()
    print(f"case efield cube n=6 precision={precision}", flush=True)
    ef = EField(mesh.verts, mesh.tets, tris)
    ef.set_solver(precision, 1e-11)
    ora = EFieldOracle(mesh.verts, mesh.tets, tris)
    for obj in (ef, ora):
        obj.set_potential(-0.065)
        obj.set_capac(0.01)
        obj.set_volres(1.0)
        obj.set_memb_res(1.0, -0.065)
    face = np.nonzero(np.all(mesh.verts[tris][:, :, 0] < 1e-12, axis=1))[0]
    for t in face:
        ef.set_tri_iclamp(int(t), 1e-9 / len(face))
        ora.set_tri_iclamp(int(t), 1e-9 / len(face))
    v_prev = ora.V.copy() * 1e-3
    for step in range(5):
        ef.advance(1e-5)
        ora.advance(1e-5)
        gpu = ef.verts_v()
        ref = ora.V * 1e-3
        dv_ref = ref - v_prev
        err = np.abs((gpu - v_prev) - dv_ref).max() / np.abs(dv_ref).max()
        assert err < tol, (step, err, ef.stats())
        v_prev = ref
        ef.set_verts_v(ref)      # compare step by step from the same state
    st = ef.stats()
    assert st["rel_residual"] <= 1e-9


def test_clamped_vertices_hold():
    mesh, tris = _cube_case(n=4, side=2e-6)
    ef = EField(mesh.verts, mesh.tets, tris)
    ora = EFieldOracle(mesh.verts, mesh.tets, tris)
    for v in (0, 5):
        ef.set_vert_clamped(v)
        ora.clamped[v] = True
    for obj in (ef, ora):
        obj.set_vert_iclamp(10, 5e-10)
    for _ in range(3):
        ef.advance(2e-5)
        ora.advance(2e-5)
    assert np.allclose(ef.verts_v()[[0, 5]], -0.065, atol=0, rtol=0)
    assert np.allclose(ef.verts_v(), ora.V * 1e-3, rtol=1e-9, atol=1e-12)


def test_rallpack1_gpu_coupled_with_ohmic_leak():
    """Rallpack1 through the coupled loop: the leak is an ohmic current of one
    'Leak' channel per membrane triangle held by the reaction-diffusion engine."""
    p = RALL
    mesh, memb, zmin_tris, zmax_tris = rallpack1_setup()
    comp = geom.TmComp("cyto", mesh)
    comp.addVolsys("vsys")
    patch = geom.TmPatch("memb", mesh, memb, comp)
    patch.addSurfsys("ssys")
    mdl = M.Model()
    leak = M.Spec("Leak", mdl)
    M.Volsys("vsys", mdl)
    ss = M.Surfsys("ssys", mdl)
    M.SReac("hold", ss, slhs=[leak], srhs=[leak], kcst=0.0)
    fm = M.flatten(mdl, ["vsys"], ["ssys"])
    mt = geom.mesh_tables(mesh, comp, memb)
    eng = Engine(fm, [mt], [7])
    eng.set_tri_counts(0, np.ones(len(memb), dtype=np.int32))
    ef = EField(mesh.verts, mesh.tets, mesh.tris[memb])
    area_cyl = np.pi * p["diameter"] * p["length"]
    tri_area = mt.tri_area.sum()
    ef.set_ohmic([0], [area_cyl / p["R_M"] / len(memb)], [p["E_M"]])
    ef.attach(eng)
    ef.set_potential(p["E_M"])
    ef.set_volres(p["R_A"])
    ef.set_capac(p["C_M"] / (tri_area / area_cyl))
    v_zmin = np.unique(mesh.tris[zmin_tris])
    for v in v_zmin:
        ef.set_vert_iclamp(int(v), p["Iinj"] / len(v_zmin))
    s0, s1 = sample_verts(mesh, zmin_tris), sample_verts(mesh, zmax_tris)
    out0, out1 = [], []
    n_samples = 1001          # the first 50 ms of the 250 ms trace (2,000 EField steps)
    for k in range(n_samples):
        t = k * 5e-5
        if k > 0:
            ef.run(eng, t, p["EF_dt"])
        v = ef.verts_v()
        out0.append(v[s0].mean())
        out1.append(v[s1].mean())
        if k % 100 == 0:
            print(f"rallpack1 t={t:.4f} s v0={out0[-1]:.6f} vx={out1[-1]:.6f} {ef.stats()}", flush=True)
    v0, vx = read_trace("v0"), read_trace("vx")
    sim0, sim1 = np.array(out0)[1:], np.array(out1)[1:]
    n = min(len(sim0), len(v0))
    rms0 = np.linalg.norm(v0[:n] - sim0[:n]) / np.sqrt(n)
    rms1 = np.linalg.norm(vx[:n] - sim1[:n]) / np.sqrt(n)
    print(f"rallpack1 rms {rms0:.3e} {rms1:.3e}", flush=True)
    assert rms0 < 1e-3 and rms1 < 1e-3


def test_ohmic_current_uses_time_averaged_channel_count():
    """Channels C.s <-> O.s gate many times within one EField step; the ohmic
    current must use the open count averaged over the step (Tri::computeI,
    src/steps/mpi/tetopsplit/tri.cpp:664-693), not the count at its end. The
    opening rate is voltage dependent, so the potentials the EField publishes
    to the triangles feed back into the next step. GPU (FP64 CG) vs oracle
    (FP64 LU) with bit-exact counts: dV within 1e-9 relative per step."""
    mesh, tris_v = _cube_case(n=4, side=2e-6)
    comp = geom.TmComp("cyto", mesh)
    memb = mesh.getSurfTris()
    mdl = M.Model()
    c, o_ = M.Spec("C", mdl), M.Spec("O", mdl)
    M.Volsys("vsys", mdl)
    ss = M.Surfsys("ssys", mdl)
    M.VDepSReac("open", ss, slhs=[c], srhs=[o_], k=lambda v: 2e4 * np.exp((v + 0.065) / 0.02))
    M.SReac("close", ss, slhs=[o_], srhs=[c], kcst=3e4)
    fm = M.flatten(mdl, ["vsys"], ["ssys"])
    mt = geom.mesh_tables(mesh, comp, memb)
    qc, qo = fm.sspec_index("C"), fm.sspec_index("O")
    n0 = np.full(len(memb), 5, dtype=np.int32)
    g, erev, T = 1e-12, -0.077, 1e-4
    print("case ohmic time average", flush=True)
    eng = Engine(fm, [mt], [5])
    eng.set_tri_counts(qc, n0)
    ef = EField(mesh.verts, mesh.tets, mesh.tris[memb])
    ef.set_solver(FP64, 1e-12)
    ora = EFieldOracle(mesh.verts, mesh.tets, mesh.tris[memb])
    for obj in (ef, ora):
        obj.set_potential(-0.065)
        obj.set_capac(0.01)
        obj.set_volres(1.0)
    ef.set_ohmic([qo], [g], [erev])
    ef.attach(eng)
    sim = bind.OracleSim(fm, mt, seed=5)
    sim.tri_counts[qc] = n0
    sim.integrate_tri_species([qo])
    for step in range(1, 4):
        # the potentials the EField published are the ones k(V) reads
        sim.tri_v[:] = eng.get_tri_v()
        assert np.allclose(sim.tri_v, ora.tri_v(), rtol=1e-12, atol=0)
        v_prev = ora.V * 1e-3
        ef.run(eng, step * T, T)
        sim.run(step * T, 0)
        assert np.array_equal(eng.get_tri_counts(qo), sim.tri_counts[qo]), step
        vt = ora.tri_v()
        i_avg = g * (sim.tri_intg[qo] / T) * (vt - erev)
        i_end = g * sim.tri_counts[qo] * (vt - erev)
        sim.tri_intg[:] = 0.0
        dv_end = _oracle_dv(ora, T, i_end)
        dv = ora.advance(T, i_avg) * 1e-3
        got = ef.verts_v() - v_prev
        err = np.abs(got - dv).max() / np.abs(dv).max()
        assert err < 1e-9, (step, err, ef.stats())
        # the end-of-step count gives a visibly different answer
        assert np.abs(dv_end - dv).max() / np.abs(dv).max() > 1e-4
        ef.set_verts_v(ora.V * 1e-3)


def _oracle_dv(ora, dt, i_amps):
    """dV (volts) of one oracle step, leaving the oracle's state unchanged."""
    A, b = ora.system(dt, i_amps)
    dv = spsolve(A.tocsc(), b)
    dv[ora.clamped] = 0.0
    return dv * 1e-3


def test_ghk_charge_drives_the_potential():
    """GHK calcium channels on the membrane: the charge each triangle moved
    over an EField step enters as its mean current (Tri::computeI); the GHK
    rate depends on the published triangle potentials. GPU (FP64 CG) vs
    oracle (FP64 LU): bit-exact counts, dV within 1e-9 relative per step."""
    mesh, _ = _cube_case(n=4, side=2e-6)
    comp = geom.TmComp("cyto", mesh)
    memb = mesh.getSurfTris()
    mdl = M.Model()
    ch = M.Chan("CaChan", mdl)
    op = M.ChanState("Open", mdl, ch)
    ca = M.Spec("Ca", mdl, valence=2)
    vs = M.Volsys("vsys", mdl)
    M.Diff("dCa", vs, ca, 2e-10)
    ss = M.Surfsys("ssys", mdl)
    cur = M.GHKcurr("ghk", ss, op, ca, computeflux=True, virtual_oconc=2e-3)
    cur.setPInfo(2e-12, -65e-3, 293.15, 2e-3, 1e-7)
    fm = M.flatten(mdl, ["vsys"], ["ssys"])
    mt = geom.mesh_tables(mesh, comp, memb)
    qo, sca = fm.sspec_index("Open"), fm.spec_index("Ca")
    n_open = np.full(len(memb), 3, dtype=np.int32)
    ca0 = np.full(len(mt.vol), 40, dtype=np.int32)
    T = 5e-5
    print("case ghk charge", flush=True)
    eng = Engine(fm, [mt], [9])
    eng.set_tri_counts(qo, n_open)
    eng.set_tet_counts(sca, ca0)
    ef = EField(mesh.verts, mesh.tets, mesh.tris[memb])
    ef.set_solver(FP64, 1e-12)
    ora = EFieldOracle(mesh.verts, mesh.tets, mesh.tris[memb])
    for obj in (ef, ora):
        obj.set_potential(-0.065)
        obj.set_capac(0.01)
        obj.set_volres(1.0)
    ef.attach(eng)
    assert np.all(ef.tri_ghk_currents(-1) == 0.0)   # no step yet
    sim = bind.OracleSim(fm, mt, seed=9)
    sim.tri_counts[qo] = n_open
    sim.counts[sca] = ca0
    for step in range(1, 4):
        sim.tri_v[:] = eng.get_tri_v()
        assert np.allclose(sim.tri_v, ora.tri_v(), rtol=1e-12, atol=0)
        v_prev = ora.V * 1e-3
        ef.run(eng, step * T, T)
        sim.run(step * T, 0)
        assert np.array_equal(eng.get_tet_counts(sca), sim.counts[sca]), step
        i_tri = sim.tri_charge[:len(memb)] * 1.602176487e-19 / T
        sim.clear_charge()
        assert i_tri.sum() < 0   # calcium flows in
        # Tri::getGHKI: the last step's charge over the step, per current and in total
        assert np.allclose(ef.tri_ghk_currents(0), i_tri, rtol=1e-12, atol=0)
        assert np.allclose(ef.tri_ghk_currents(-1), i_tri, rtol=1e-12, atol=0)
        dv = ora.advance(T, i_tri) * 1e-3
        got = ef.verts_v() - v_prev
        err = np.abs(got - dv).max() / np.abs(dv).max()
        assert err < 1e-9, (step, err, ef.stats())
        ef.set_verts_v(ora.V * 1e-3)


def test_rallpack1_through_the_solver_interface():
    """Rallpack1 as the reference script sets it up (rallpack1.py:39-110):
    a Memb over the cable's patch, one 'Leak' channel state per triangle with
    an OhmicCurr, TetOpSplit(calcMembPot=EF_DEFAULT), membrane setters and a
    vertex current clamp. First 20 ms against the golden traces, RMS < 1 mV."""
    from steps_amd import rng as R
    from steps_amd.solver import EF_DEFAULT, TetOpSplit

    p = RALL
    mesh, memb_tris, zmin_tris, zmax_tris = rallpack1_setup()
    cyto = geom.TmComp("cyto", mesh)
    patch = geom.TmPatch("memb", mesh, memb_tris, cyto)
    patch.addSurfsys("ssys")
    geom.Memb("membrane", mesh, [patch])
    mdl = M.Model()
    ssys = M.Surfsys("ssys", mdl)
    chan = M.Chan("L", mdl)
    leak = M.ChanState("Leak", mdl, chan)
    area_cyl = np.pi * p["diameter"] * p["length"]
    M.OhmicCurr("OC_L", ssys, leak, p["E_M"], area_cyl / p["R_M"] / len(memb_tris))
    r = R.create("mt19937", 512)
    r.initialize(7)
    print("case rallpack1 solver interface", flush=True)
    sim = TetOpSplit(mdl, mesh, r, EF_DEFAULT)
    sim.setEfieldDT(p["EF_dt"])
    for t in memb_tris:
        sim.setTriSpecCount(int(t), "Leak", 1)
    area_factor = patch.getArea() / area_cyl
    sim.setMembPotential("membrane", p["E_M"])
    sim.setMembVolRes("membrane", p["R_A"])
    sim.setMembCapac("membrane", p["C_M"] / area_factor)
    v_zmin = np.unique(mesh.tris[zmin_tris])
    for v in v_zmin:
        sim.setVertIClamp(int(v), p["Iinj"] / len(v_zmin))
    s0, s1 = sample_verts(mesh, zmin_tris), sample_verts(mesh, zmax_tris)
    out0, out1 = [], []
    buf = np.zeros(2)
    for k in range(1, 401):
        sim.run(k * 5e-5)
        sim.getBatchVertVsNP(np.array(s0), buf)
        out0.append(buf.mean())
        sim.getBatchVertVsNP(np.array(s1), buf)
        out1.append(buf.mean())
        if k % 100 == 0:
            print(f"rallpack1 solver t={k * 5e-5:.4f} v0={out0[-1]:.6f}", flush=True)
    v0, vx = read_trace("v0"), read_trace("vx")
    n = len(out0)
    rms0 = np.linalg.norm(v0[:n] - np.array(out0)) / np.sqrt(n)
    rms1 = np.linalg.norm(vx[:n] - np.array(out1)) / np.sqrt(n)
    print(f"rallpack1 solver rms {rms0:.3e} {rms1:.3e}", flush=True)
    assert rms0 < 1e-3 and rms1 < 1e-3
    assert abs(sim.getTriV(int(memb_tris[0])) - p["E_M"]) < 0.05
    assert sim.getEfieldDT() == p["EF_dt"]


def _large_case(n=28, side=14e-6):
    """29^3 = 24,389 vertices: above the one-workgroup solver, so the
    multigrid-preconditioned CG runs."""
    mesh = geom.kuhn_cube(n, side)
    tris = mesh.tris[mesh.getSurfTris()]
    return mesh, tris


def _drive(mesh, tris, objs, clamp=()):
    for obj in objs:
        obj.set_potential(-0.065)
        obj.set_capac(0.01)
        obj.set_volres(1.0)
        obj.set_memb_res(1.0, -0.065)
    face = np.nonzero(np.all(mesh.verts[tris][:, :, 0] < 1e-12, axis=1))[0]
    for t in face:
        for obj in objs:
            obj.set_tri_iclamp(int(t), 1e-9 / len(face))
    for v in clamp:
        objs[0].set_vert_clamped(v)
        if len(objs) > 1:
            objs[1].clamped[v] = True


@pytest.mark.parametrize("precision,tol", [(MIXED, 1e-6), (FP64, 1e-9)])
def test_multigrid_steps_match_fp64_oracle(precision, tol):
    from steps_amd.efield import MULTIGRID

    mesh, tris = _large_case()
    print(f"case efield multigrid cube n=28 precision={precision}", flush=True)
    ef = EField(mesh.verts, mesh.tets, tris)
    ef.set_solver(precision, 1e-11)
    ef.set_precond(MULTIGRID)
    ora = EFieldOracle(mesh.verts, mesh.tets, tris)
    _drive(mesh, tris, [ef, ora], clamp=(7, 4000))
    v_prev = ora.V.copy() * 1e-3
    for step in range(3):
        ef.advance(1e-5)
        ora.advance(1e-5)
        gpu = ef.verts_v()
        ref = ora.V * 1e-3
        dv_ref = ref - v_prev
        err = np.abs((gpu - v_prev) - dv_ref).max() / np.abs(dv_ref).max()
        assert err < tol, (step, err, ef.stats())
        assert gpu[7] == -0.065 and gpu[4000] == -0.065
        v_prev = ref
        ef.set_verts_v(ref)
    info = ef.precond_info()
    assert info["levels"] >= 2 and 0 < info["coarsest"] <= 2000, info
    assert ef.stats()["rel_residual"] <= 1e-9


def test_multigrid_cuts_iterations():
    """The same step with Jacobi and with the multigrid V-cycle: both solve it
    to the FP64 tolerance; the V-cycle needs far fewer CG iterations."""
    from steps_amd.efield import JACOBI, MULTIGRID

    mesh, tris = _large_case()
    print("case efield multigrid vs jacobi iterations", flush=True)
    out = {}
    for kind in (JACOBI, MULTIGRID):
        ef = EField(mesh.verts, mesh.tets, tris)
        ef.set_precond(kind)
        _drive(mesh, tris, [ef])
        ef.advance(1e-5)
        out[kind] = (ef.verts_v(), ef.stats()["iterations"], ef.precond_info())
    v_j, it_j, info_j = out[JACOBI]
    v_m, it_m, info_m = out[MULTIGRID]
    dv = np.abs(v_j + 0.065).max()
    assert np.abs(v_m - v_j).max() <= 1e-6 * dv
    assert info_j["levels"] == 0 and info_m["levels"] >= 2
    assert it_m * 4 <= it_j, (it_m, it_j)
