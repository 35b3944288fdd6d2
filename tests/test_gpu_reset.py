"""reset() and checkpoint/restore semantics on the GPU.

TetOpSplitP::reset (tetopsplit.cpp:1436-1503) clears counts, time and step
counts but leaves the RNG running, so the reference's own validation loops
(for i in range(NITER): sim.reset(); set counts; run ...,
test/validation/API_2/validation_rd_mpi/parallel_spatial_tetopsplit.py:276-277)
collect independent runs. TetOpSplitP::checkpoint / restore save and restore
the EField with the reaction-diffusion state (tetopsplit.cpp:309, 446).
"""
import numpy as np
import pytest
from scipy import stats

from oracle import bind
from steps_amd import geom, model as M, rng as R
from steps_amd import workloads as W
from steps_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _counts(eng, S, real=0):
    return np.stack([eng.get_tet_counts(s, real) for s in range(S)])


@pytest.mark.parametrize("mode", [0, 1])
def test_reset_continues_the_streams_bit_exact(mode):
    w = W.ab_binding_cube(n=4, side=5e-6, n_a=3000, n_b=2000, seed=4, end_time=0.06)
    print(f"case reset streams mode={mode}", flush=True)
    eng = Engine(w.fm, [w.mt], [13], mode=mode)
    o = bind.OracleSim(w.fm, w.mt, seed=13)
    runs = []
    for it in range(3):
        eng.reset()
        o.reset()
        for s in range(3):
            eng.set_tet_counts(s, w.counts[s])
        o.counts[:] = w.counts
        eng.run(w.end_time)
        o.run(w.end_time, mode)
        got = _counts(eng, 3)
        assert np.array_equal(got, o.counts), f"iteration {it}"
        assert eng.nsteps() == o.nsteps and eng.time() == w.end_time
        assert eng.extents() == (o.events.value, o.moved.value)
        runs.append(got)
    assert not np.array_equal(runs[0], runs[1]) and not np.array_equal(runs[1], runs[2])
    # re-keying restarts the streams: the first run is reproduced
    eng.set_seed(13)
    eng.reset()
    for s in range(3):
        eng.set_tet_counts(s, w.counts[s])
    eng.run(w.end_time)
    assert np.array_equal(_counts(eng, 3), runs[0])


def test_reset_loop_gives_the_binomial_decay_ensemble():
    """First-order decay A -> 0 (k = 2 /s) with diffusion, through the solver
    interface as the reference's validation scripts drive it: NITER iterations
    of reset(), setCompSpecCount, run(t). The survivors of N0 = 100 molecules at
    t = 0.35 s are exactly Binomial(100, exp(-k t)) across iterations."""
    from steps_amd.solver import TetOpSplit

    mdl = M.Model()
    A = M.Spec("A", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Reac("decay", vs, lhs=[A], rhs=[], kcst=2.0)
    M.Diff("dA", vs, A, 1e-12)
    mesh = geom.kuhn_cube(3, 3e-6)
    comp = geom.TmComp("cyto", mesh)
    comp.addVolsys("vsys")
    r = R.create("r123", 1024)
    r.initialize(1234)
    sim = TetOpSplit(mdl, mesh, r)
    niter, n0, t = 400, 100, 0.35
    print(f"case reset loop decay niter={niter}", flush=True)
    out = np.empty(niter)
    for i in range(niter):
        sim.reset()
        sim.setCompSpecCount("cyto", "A", n0)
        sim.run(t)
        out[i] = sim.getCompSpecCount("cyto", "A")
    p = np.exp(-2.0 * t)
    mean, var = n0 * p, n0 * p * (1 - p)
    assert len(np.unique(out)) > 10
    z = (out.mean() - mean) / np.sqrt(var / niter)
    assert abs(z) < 2.576, (out.mean(), mean)
    # sample variance against the binomial variance (chi-square, 99%)
    chi = (niter - 1) * out.var(ddof=1) / var
    lo, hi = stats.chi2.ppf([0.005, 0.995], niter - 1)
    assert lo < chi < hi, (out.var(ddof=1), var)


def _channel_membrane_sim(seed):
    """A 2 um cube whose whole surface (24 um^2, 0.24 pF) is membrane: 40
    channels open with a voltage-dependent rate and pass an ohmic current and a
    GHK calcium current, 6 vertices take 0.1 pA each; calcium diffuses inside.
    The potential moves by a few mV per ms, inside the k(V) table."""
    from steps_amd.solver import EF_DEFAULT, TetOpSplit

    mesh = geom.kuhn_cube(4, 2e-6)
    cyto = geom.TmComp("cyto", mesh)
    cyto.addVolsys("vsys")
    patch = geom.TmPatch("memb", mesh, mesh.getSurfTris(), cyto)
    patch.addSurfsys("ssys")
    geom.Memb("membrane", mesh, [patch])
    mdl = M.Model()
    ca = M.Spec("Ca", mdl, valence=2)
    vs = M.Volsys("vsys", mdl)
    M.Diff("dCa", vs, ca, 2e-10)
    ss = M.Surfsys("ssys", mdl)
    ch = M.Chan("K", mdl)
    c, o = M.ChanState("C", mdl, ch), M.ChanState("O", mdl, ch)
    M.VDepSReac("open", ss, slhs=[c], srhs=[o], k=lambda v: 400.0 / (1.0 + np.exp(-(v + 0.03) / 0.01)))
    M.SReac("close", ss, slhs=[o], srhs=[c], kcst=200.0)
    M.OhmicCurr("oc", ss, o, 0.0, 2e-12)
    g = M.GHKcurr("ghk", ss, o, ca, computeflux=True, virtual_oconc=2e-3)
    g.setPInfo(1e-12, -65e-3, 293.15, 2e-3, 1e-7)
    r = R.create("r123", 1024)
    r.initialize(seed)
    sim = TetOpSplit(mdl, mesh, r, EF_DEFAULT)
    sim.setEfieldDT(5e-5)
    sim.setPatchSpecCount("memb", "C", 40)
    sim.setCompSpecConc("cyto", "Ca", 5e-6)
    v = np.unique(mesh.tris[mesh.getSurfTris()])[:6]
    for x in v:
        sim.setVertIClamp(int(x), 1e-13)
    return sim, mesh


def _state(sim, mesh):
    tets = np.arange(mesh.countTets(), dtype=np.int64)
    ca = np.zeros(len(tets))
    sim.getBatchTetSpecCountsNP(tets, "Ca", ca)
    tris = mesh.getSurfTris()
    o = np.zeros(len(tris))
    sim.getBatchTriSpecCountsNP(tris, "O", o)
    vv = np.zeros(mesh.countVertices())
    sim.getBatchVertVsNP(np.arange(mesh.countVertices()), vv)
    return ca, o, vv


def test_checkpoint_restore_with_efield_continues_exactly(tmp_path):
    print("case checkpoint with efield", flush=True)
    sim, mesh = _channel_membrane_sim(21)
    sim.run(1e-3)
    f = tmp_path / "cp.bin"
    sim.checkpoint(str(f))
    sim.run(2e-3)
    ref = _state(sim, mesh)
    assert ref[1].sum() > 0 and np.ptp(ref[2]) > 0
    # the same solver, restored
    sim.restore(str(f))
    assert sim.getTime() == 1e-3
    sim.run(2e-3)
    for a, b in zip(_state(sim, mesh), ref):
        assert np.array_equal(a, b)
    # a fresh solver, restored
    sim2, _ = _channel_membrane_sim(21)
    sim2.restore(str(f))
    sim2.run(2e-3)
    for a, b in zip(_state(sim2, mesh), ref):
        assert np.array_equal(a, b)


def test_reset_puts_the_membrane_back_to_its_default_potential():
    print("case reset efield", flush=True)
    sim, mesh = _channel_membrane_sim(5)
    sim.run(1e-3)
    assert np.abs(_state(sim, mesh)[2] + 0.065).max() > 1e-6
    sim.reset()
    assert sim.getTime() == 0.0 and sim.getNSteps() == 0
    assert np.all(_state(sim, mesh)[2] == -0.065)
    sim.setPatchSpecCount("memb", "C", 40)
    sim.run(1e-3)
    assert sim.getTime() == 1e-3
