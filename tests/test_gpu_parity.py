"""GPU engine vs CPU oracle: integer counts must agree bit-for-bit.

Both sides draw from the same keyed Philox4x32-8 streams with the same
floating-point operation sequence, so any difference in a molecule count is a
bug. Runs through the C ABI (steps_amd.engine -> include/stg.h).
"""
import numpy as np
import pytest

from oracle import bind
from steps_amd import workloads as W
from steps_amd.engine import Engine
from steps_amd import _native as N

pytestmark = pytest.mark.gpu


def _engine(w, seed, mode=0, **kw):
    eng = Engine(w.fm, [w.mt], [seed], mode=mode, **kw)
    for s in range(len(w.fm.specs)):
        eng.set_tet_counts(s, w.counts[s])
    for s in range(len(w.fm.sspecs)):
        eng.set_tri_counts(s, w.tri_counts[s])
    return eng


def _oracle(w, seed, **kw):
    o = bind.OracleSim(w.fm, w.mt, seed=seed, **kw)
    o.counts[:] = w.counts
    if len(w.fm.sspecs):
        o.tri_counts[:] = w.tri_counts
    return o


def _gpu_counts(eng, w, real=0):
    return np.stack([eng.get_tet_counts(s, real) for s in range(len(w.fm.specs))])


def _gpu_tri_counts(eng, w, real=0):
    return np.stack([eng.get_tri_counts(s, real) for s in range(len(w.fm.sspecs))])


@pytest.mark.parametrize("seed", [1, 2, 7])
@pytest.mark.parametrize("mode", [0, 1])
def test_ab_binding_bit_exact(seed, mode):
    w = W.ab_binding_cube(n=4, side=5e-6, n_a=3000, n_b=2000, seed=seed, end_time=0.2)
    print(f"case ab_binding seed={seed} mode={mode}", flush=True)
    eng = _engine(w, seed, mode)
    o = _oracle(w, seed)
    for t in (0.013, 0.05, 0.2):  # several run() calls, as result saving does
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(eng, w), o.counts), f"t={t}"
    assert eng.nsteps() == o.nsteps
    ev, mv = eng.extents()
    assert ev == o.events.value and mv == o.moved.value
    tot = o.counts.sum(axis=1)
    assert tot[0] + tot[2] == 3000 and tot[1] + tot[2] == 2000


@pytest.mark.parametrize("mode", [0, 1])
def test_escaped_face_slots_bit_exact(mode):
    """Face slots are 16-bit, with counts above 32,767 escaped to a 32-bit side
    table. About 400,000 molecules per tet (no reactions) send more than 32,767
    through most faces in every period, so most gathers take the escape; a few
    faces stay below it. Counts must equal the oracle's, and the total is kept."""
    w = W.ab_binding_cube(n=2, side=2e-6, n_a=10_000_000, n_b=10_000_000, seed=6,
                          end_time=0.15, kf=0.0, kb=0.0)
    print(f"case escaped face slots mode={mode}", flush=True)
    eng = _engine(w, 31, mode)
    o = _oracle(w, 31)
    for t in (0.05, 0.15):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(eng, w), o.counts), f"t={t}"
    assert eng.nsteps() == o.nsteps and eng.nsteps() >= 4
    tot = o.counts.sum(axis=1)
    assert tot[0] == 10_000_000 and tot[1] == 10_000_000 and tot[2] == 0
    assert eng.extents()[1] == o.moved.value and o.moved.value > 100 * 32767


@pytest.mark.parametrize("mode", [0, 1])
def test_irregular_mesh_bit_exact(mode):
    """A Kuhn cube with jittered interior vertices: irregular geometry, every tet
    with its own volume and face factors (the other cases use regular cubes,
    whose tets fall into a few shape classes)."""
    from steps_amd import geom, model as M

    n, side = 5, 5e-6
    mesh = geom.kuhn_cube(n, side)
    v = mesh.verts.copy()
    h = side / n
    inner = np.all((v > 0.5 * h) & (v < side - 0.5 * h), axis=1)
    v[inner] += np.random.default_rng(8).uniform(-0.2 * h, 0.2 * h, size=(inner.sum(), 3))
    mesh = geom.Tetmesh(v, mesh.tets)
    comp = geom.TmComp("cyto", mesh)
    comp.addVolsys("vsys")
    fm = M.flatten(W.ab_binding_model(), ["vsys"])
    mt = geom.mesh_tables(mesh, comp)
    assert len(np.unique(mt.vol)) > len(mt.vol) // 2
    rng = np.random.default_rng(3)
    counts = np.zeros((3, len(mt.vol)), dtype=np.int32)
    counts[0] = W.place_uniform(4000, mt.vol, rng)
    counts[1] = W.place_uniform(3000, mt.vol, rng)
    w = W.Workload("jittered", fm, mt, counts, np.zeros((0, 0), np.int32), 0.2)
    print(f"case irregular mesh mode={mode}", flush=True)
    eng = _engine(w, 17, mode)
    o = _oracle(w, 17)
    for t in (0.05, 0.2):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(eng, w), o.counts), f"t={t}"
    assert eng.extents() == (o.events.value, o.moved.value)


@pytest.mark.parametrize("mode", [0, 1])
def test_four_species_bit_exact(mode):
    w = W.four_species_cube(n=6, side=6e-6, seed=5, end_time=0.05)
    print(f"case four_species n=6 mode={mode}", flush=True)
    eng = _engine(w, 99, mode)
    o = _oracle(w, 99)
    eng.run(w.end_time)
    o.run(w.end_time, mode)
    assert np.array_equal(_gpu_counts(eng, w), o.counts)
    assert eng.extents() == (o.events.value, o.moved.value)


def test_calcium_membrane_bit_exact():
    w = W.calcium_membrane_cube(n=6, side=2e-6, seed=4, end_time=2e-4, buf_um=100.0, ca_um=50.0)
    print("case calcium_membrane n=6", flush=True)
    eng = _engine(w, 123)
    o = _oracle(w, 123)
    eng.run(w.end_time)
    o.run(w.end_time, 0)
    assert np.array_equal(_gpu_counts(eng, w), o.counts)
    assert np.array_equal(_gpu_tri_counts(eng, w), o.tri_counts[: len(w.fm.sspecs)])
    # calcium is conserved: free + bound + extruded - admitted through the channels
    assert o.events.value > 0


def test_batch_matches_independent_runs():
    fm, meshes, counts, reac_k, info = W.ab_binding_ensemble(n_real=5, n_min=3, n_max=6,
                                                             side=4e-6, n_a=800, n_b=600, seed=2)
    print("case batch of 5", flush=True)
    eng = Engine(fm, meshes, [17] * 5, reac_k=reac_k)
    for r in range(5):
        for s in range(3):
            eng.set_tet_counts(s, counts[r][s], real=r)
    eng.run(0.08)
    for r in range(5):
        o = bind.OracleSim(fm, meshes[r], seed=17, stream=r, reac_k=reac_k[r])
        o.counts[:] = counts[r]
        o.run(0.08, 0)
        got = np.stack([eng.get_tet_counts(s, r) for s in range(3)])
        assert np.array_equal(got, o.counts), f"realization {r}"
        assert eng.nsteps(r) == o.nsteps


def test_batch_repeated_and_changing_run_lengths():
    """Runs of equal length reuse the uploaded block order and the blocks of
    each launch; a run of another length rebuilds them; a long run grows the
    run tables (dt per period and realization). Every realization stays
    bit-exact against its own oracle run through the same saves."""
    fm, meshes, counts, reac_k, info = W.ab_binding_ensemble(n_real=4, n_min=3, n_max=6,
                                                             side=4e-6, n_a=800, n_b=600, seed=5)
    print("case batch of 4, saves of changing length", flush=True)
    eng = Engine(fm, meshes, [23] * 4, reac_k=reac_k)
    oracles = []
    for r in range(4):
        for s in range(3):
            eng.set_tet_counts(s, counts[r][s], real=r)
        o = bind.OracleSim(fm, meshes[r], seed=23, stream=r, reac_k=reac_k[r])
        o.counts[:] = counts[r]
        oracles.append(o)
    # one period per save up to 12 ms (periods of 13-53 ms are clipped at the
    # save), 1-4 periods to 60 ms, and up to 26 periods to 400 ms, beyond the
    # run tables' first allocation (16 periods per realization)
    for t in (0.001, 0.002, 0.003, 0.010, 0.011, 0.012, 0.060, 0.061, 0.400, 0.401):
        eng.run(t)
        for r, o in enumerate(oracles):
            o.run(t, 0)
            got = np.stack([eng.get_tet_counts(s, r) for s in range(3)])
            assert np.array_equal(got, o.counts), f"t={t} realization {r}"
            assert eng.nsteps(r) == o.nsteps, f"t={t} realization {r}"


@pytest.mark.parametrize("block,graph", [(64, False), (1024, False), (256, True)])
def test_launch_shape_invariance(block, graph):
    w = W.four_species_cube(n=5, side=5e-6, seed=8, end_time=0.03)
    print(f"case launch block={block} graph={graph}", flush=True)
    ref = _engine(w, 3)
    ref.run(w.end_time / 2)
    ref.run(w.end_time)
    eng = _engine(w, 3)
    eng.set_launch(block, graph)
    eng.run(w.end_time / 2)
    eng.run(w.end_time)
    assert np.array_equal(_gpu_counts(eng, w), _gpu_counts(ref, w))


def test_conservation_full_size_diffusion_only():
    """Exact conservation at the 384,000-tet size with reactions switched off."""
    w = W.four_species_cube(n=40, side=40e-6, seed=0)
    print("case conservation n=40", flush=True)
    eng = _engine(w, 5)
    for j in range(len(w.fm.reacs)):
        eng.set_reac_k(j, 0.0)
    before = w.counts.sum(axis=1)
    eng.run(20 * eng.update_period())
    after = np.array([eng.sum_tet_counts(s) for s in range(4)])
    assert np.array_equal(before, after)
    assert all(np.all(eng.get_tet_counts(s) >= 0) for s in range(4))
    idx = np.array([0, 5, 383999, 1234])
    assert np.array_equal(eng.batch_tet_counts(2, idx), eng.get_tet_counts(2)[idx].astype(float))


def test_errors_are_reported():
    w = W.ab_binding_cube(n=2, side=2e-6, n_a=10, n_b=10)
    eng = _engine(w, 1)
    eng.run(0.01)
    with pytest.raises(N.ArgErr):
        eng.run(0.005)  # end time before current time
    with pytest.raises(N.ArgErr):
        eng.set_tet_counts(0, -np.ones(w.n_tets))
    with pytest.raises(N.ArgErr):
        eng.set_launch(100, False)


@pytest.mark.parametrize("threshold", [0, 3, 1000])
def test_diff_apply_threshold_bit_exact(threshold):
    """Per-molecule vs conditional-binomial face split, at any threshold."""
    w = W.four_species_cube(n=6, side=6e-6, seed=11, end_time=0.02)
    print(f"case threshold={threshold}", flush=True)
    eng = _engine(w, 5)
    eng.set_diff_apply_threshold(threshold)
    o = _oracle(w, 5, diff_threshold=threshold)
    eng.run(w.end_time)
    o.run(w.end_time, 0)
    assert np.array_equal(_gpu_counts(eng, w), o.counts)


def test_checkpoint_restore_continues_bit_exact(tmp_path):
    w = W.four_species_cube(n=5, side=5e-6, seed=21, end_time=0.04)
    print("case checkpoint/restore", flush=True)
    eng = _engine(w, 9)
    eng.run(0.013)
    state = eng.checkpoint()
    eng.run(0.04)
    ref = _gpu_counts(eng, w)
    ev = eng.extents()
    eng.restore(state)
    assert eng.time() == 0.013
    eng.run(0.04)
    assert np.array_equal(_gpu_counts(eng, w), ref)
    assert eng.extents() == ev
    # through the solver-level file interface of a fresh engine
    eng2 = _engine(w, 9)
    p = tmp_path / "cp.bin"
    p.write_bytes(state)
    eng2.restore(p.read_bytes())
    eng2.run(0.04)
    assert np.array_equal(_gpu_counts(eng2, w), ref)


def test_autotune_keeps_results():
    from steps_amd.autotune import autotune

    w = W.four_species_cube(n=6, side=6e-6, seed=4, end_time=0.03)
    print("case autotune", flush=True)
    ref = _engine(w, 2)
    ref.run(w.end_time)
    eng = _engine(w, 2)
    res = autotune(eng, trial_periods=3, blocks=(64, 256), graphs=(False, True), spec_waves=(0, 8))
    # generic and model-specialised kernels (two register bounds), both blocks:
    # stream and graph with one lane per tet (12 shapes), stream with 2 and 4
    # lanes per tet (6 shapes each), each with 16- and 8-bit face slots
    assert res["block"] in (64, 256) and res["lanes"] in (1, 2, 4) and len(res["timings"]) == 48
    assert sum(k.endswith("/lanes4") for k in res["timings"]) == 6
    assert sum(k.endswith("/slot8") for k in res["timings"]) == 24
    assert eng.lanes == res["lanes"] and eng.slot_width == res["slot_width"]
    assert eng.time() == 0.0
    eng.run(w.end_time)
    assert np.array_equal(_gpu_counts(eng, w), _gpu_counts(ref, w))


@pytest.mark.parametrize("mode", [0, 1])
def test_vdep_membrane_bit_exact(mode):
    """Voltage-gated influx channels at per-triangle potentials spread over
    -90..+30 mV (most between table points), GPU vs oracle."""
    w = W.calcium_membrane_cube(n=6, side=2e-6, seed=8, end_time=3e-4, buf_um=100.0,
                                ca_um=20.0, influx=3e6, vdep_influx=True)
    print(f"case vdep membrane mode={mode}", flush=True)
    assert (w.fm.sreac_vdep >= 0).sum() == 1
    v = np.random.default_rng(3).uniform(-90e-3, 30e-3, len(w.mt.tri_tet))
    eng = _engine(w, 77, mode)
    eng.set_tri_v(v)
    assert np.array_equal(eng.get_tri_v(), v)
    o = _oracle(w, 77)
    o.tri_v[:] = v
    for t in (1e-4, w.end_time):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(eng, w), o.counts), f"t={t}"
        assert np.array_equal(_gpu_tri_counts(eng, w), o.tri_counts[:len(w.fm.sspecs)])
    assert eng.extents() == (o.events.value, o.moved.value)


def test_vdep_voltage_outside_table_is_reported():
    w = W.calcium_membrane_cube(n=3, side=1e-6, seed=1, end_time=1e-4, vdep_influx=True)
    print("case vdep out of range", flush=True)
    eng = _engine(w, 1)
    v = np.full(len(w.mt.tri_tet), -65e-3)
    v[0] = 0.2  # above the 100 mV end of the table
    eng.set_tri_v(v)
    with pytest.raises(N.StgError, match="outside"):
        eng.run(w.end_time)


@pytest.mark.parametrize("mode", [0, 1])
def test_ghk_membrane_bit_exact(mode):
    """GHK calcium influx (virtual outer 2 mM) at per-triangle potentials,
    GPU vs oracle, with the temperature changed from its default."""
    w = W.calcium_membrane_cube(n=6, side=2e-6, seed=8, end_time=3e-4, buf_um=100.0,
                                ca_um=20.0, ghk_influx=True)
    print(f"case ghk membrane mode={mode}", flush=True)
    assert len(w.fm.ghk) == 1
    v = np.random.default_rng(4).uniform(-90e-3, 30e-3, len(w.mt.tri_tet))
    eng = _engine(w, 55, mode)
    eng.set_tri_v(v)
    eng.set_temp(308.15)
    assert eng.temp() == 308.15
    o = _oracle(w, 55)
    o.tri_v[:] = v
    o.set_temp(308.15)
    for t in (1e-4, w.end_time):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(eng, w), o.counts), f"t={t}"
        assert np.array_equal(_gpu_tri_counts(eng, w), o.tri_counts[:len(w.fm.sspecs)])
    assert eng.extents() == (o.events.value, o.moved.value)
    assert o.tri_charge.sum() < 0   # net calcium influx


@pytest.mark.parametrize("mode", [0, 1])
def test_surface_diffusion_bit_exact(mode):
    """Pumps and extruded ions diffusing on the membrane (CaOut is produced by
    the pump reaction, so its occupancy is tracked), GPU vs oracle."""
    w = W.calcium_membrane_cube(n=6, side=2e-6, seed=12, end_time=3e-4, buf_um=100.0,
                                ca_um=30.0, pump_density=300.0, pump_dcst=1e-12)
    print(f"case surface diffusion mode={mode}", flush=True)
    assert w.fm.sspec_dcst.max() > 0
    eng = _engine(w, 31, mode)
    o = _oracle(w, 31)
    assert eng.update_period() == o.update_period()
    tot0 = w.tri_counts.sum(axis=1)
    for t in (1e-4, w.end_time):
        eng.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(eng, w), o.counts), f"t={t}"
        assert np.array_equal(_gpu_tri_counts(eng, w), o.tri_counts[:len(w.fm.sspecs)]), f"t={t}"
    assert eng.extents() == (o.events.value, o.moved.value)
    p = w.fm.sspec_index("Pump")
    assert _gpu_tri_counts(eng, w)[p].sum() == tot0[p]          # pumps are conserved
    assert not np.array_equal(_gpu_tri_counts(eng, w)[p], w.tri_counts[p])   # and moved


def _higher_order_workload(seed=3):
    """Reactions of orders 0-3 with repeated reactants (2A, A + 2B, 3C), a
    species that does not diffuse (E) and one that is only produced."""
    from steps_amd import geom, model as M

    mdl = M.Model()
    A, B, C, D, E = (M.Spec(x, mdl) for x in "ABCDE")
    vs = M.Volsys("vsys", mdl)
    M.Reac("dim", vs, lhs=[A, A], rhs=[B], kcst=2e6)
    M.Reac("undim", vs, lhs=[B], rhs=[A, A], kcst=3.0)
    M.Reac("tri", vs, lhs=[A, B, B], rhs=[C], kcst=1e12)
    M.Reac("tri3", vs, lhs=[C, C, C], rhs=[D, E], kcst=1e11)
    M.Reac("prod", vs, lhs=[], rhs=[A], kcst=1e-7)
    M.Reac("deg", vs, lhs=[D], rhs=[], kcst=0.7)
    M.Reac("conv", vs, lhs=[E], rhs=[C], kcst=0.3)
    for sp, d in zip((A, B, C, D), (1e-12, 4e-13, 2e-13, 1e-12)):
        M.Diff("d" + sp.id, vs, sp, d)
    fm = M.flatten(mdl, ["vsys"])
    mesh = geom.kuhn_cube(5, 5e-6)
    comp = geom.TmComp("c", mesh)
    comp.addVolsys("vsys")
    mt = geom.mesh_tables(mesh, comp)
    rng = np.random.default_rng(seed)
    counts = rng.poisson([[60], [40], [30], [5], [2]], size=(5, len(mt.vol))).astype(np.int32)
    return W.Workload("higher_order_cube", fm, mt, counts, np.zeros((0, 0), np.int32), 0.05, {}, mesh)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("which", ["higher_order", "four_species"])
def test_specialised_kernel_matches_generic_and_oracle(mode, which):
    """The kernel compiled at run time with the model's tables as constants
    reproduces the generic kernel and the oracle bit for bit."""
    w = (_higher_order_workload() if which == "higher_order"
         else W.four_species_cube(n=5, side=5e-6, seed=2, end_time=0.05))
    print(f"case specialised {which} mode={mode}", flush=True)
    gen = _engine(w, 9, mode)
    spec = _engine(w, 9, mode)
    spec.set_specialize(True, 4 if which == "four_species" else 0)
    assert spec.specialized and not gen.specialized
    o = _oracle(w, 9)
    for t in (0.004, 0.02, w.end_time):
        gen.run(t)
        spec.run(t)
        o.run(t, mode)
        assert np.array_equal(_gpu_counts(spec, w), _gpu_counts(gen, w)), f"t={t}"
        assert np.array_equal(_gpu_counts(spec, w), o.counts), f"t={t}"
    assert spec.extents() == gen.extents() and spec.nsteps() == gen.nsteps()
    # switching back and forth keeps the results
    spec.set_specialize(False)
    gen.run(w.end_time + 0.01)
    spec.run(w.end_time + 0.01)
    assert np.array_equal(_gpu_counts(spec, w), _gpu_counts(gen, w))


def test_specialised_kernel_batch_and_graph():
    fm, meshes, counts, reac_k, info = W.ab_binding_ensemble(n_real=6, n_min=3, n_max=7,
                                                             side=4e-6, n_a=900, n_b=700, seed=4)
    print("case specialised batch of 6 with graph replay", flush=True)
    engs = []
    for spec in (False, True):
        eng = Engine(fm, meshes, [21] * 6, reac_k=reac_k)
        for r in range(6):
            for s in range(3):
                eng.set_tet_counts(s, counts[r][s], real=r)
        if spec:
            eng.set_specialize(True)
            eng.set_launch(128, True)
        eng.run(0.03)
        eng.run(0.08)
        engs.append(eng)
    for r in range(6):
        a = np.stack([engs[0].get_tet_counts(s, r) for s in range(3)])
        b = np.stack([engs[1].get_tet_counts(s, r) for s in range(3)])
        assert np.array_equal(a, b), f"realization {r}"
        assert engs[0].nsteps(r) == engs[1].nsteps(r)


@pytest.mark.parametrize("kind", ["calcium", "vdep", "ghk", "sdiff"])
def test_specialised_membrane_kernel_matches_generic(kind):
    """Membrane models: the volume reactions as constants, surface reactions,
    GHK currents, k(V) and surface diffusion from run-time tables."""
    kw = dict(calcium=dict(ca_um=20.0, influx=3e6), vdep=dict(ca_um=20.0, influx=3e6, vdep_influx=True),
              ghk=dict(ca_um=20.0, ghk_influx=True),
              sdiff=dict(ca_um=30.0, pump_density=300.0, pump_dcst=1e-12))[kind]
    w = W.calcium_membrane_cube(n=6, side=2e-6, seed=8, end_time=3e-4, buf_um=100.0, **kw)
    print(f"case specialised membrane {kind}", flush=True)
    gen = _engine(w, 5)
    spec = _engine(w, 5)
    spec.set_specialize(True)
    assert spec.specialized
    if kind in ("vdep", "ghk"):
        v = np.linspace(-0.09, 0.03, len(w.mt.tri_tet))
        gen.set_tri_v(v)
        spec.set_tri_v(v)
    for t in (1e-4, w.end_time):
        gen.run(t)
        spec.run(t)
        assert np.array_equal(_gpu_counts(spec, w), _gpu_counts(gen, w)), f"t={t}"
        assert np.array_equal(_gpu_tri_counts(spec, w), _gpu_tri_counts(gen, w)), f"t={t}"
    assert spec.extents() == gen.extents() and spec.extents()[0] > 0


def test_all_realization_totals_in_one_call():
    """stg_sum_all_tet_counts / stg_sum_all_tri_counts equal the per-call
    device totals and the host sums of the count arrays, for a batch of
    unequal realizations (volume) and for a membrane model (surface)."""
    fm, meshes, counts, reac_k, info = W.ab_binding_ensemble(n_real=7, n_min=3, n_max=8,
                                                             side=4e-6, n_a=900, n_b=700, seed=5)
    print("case totals of 7 realizations", flush=True)
    eng = Engine(fm, meshes, [3] * 7, reac_k=reac_k)
    for r in range(7):
        for s in range(3):
            eng.set_tet_counts(s, counts[r][s], real=r)
    for t in (0.0, 0.05):
        if t:
            eng.run(t)
        tot = eng.sum_all_tet_counts()
        assert tot.shape == (7, 3)
        for r in range(7):
            for s in range(3):
                assert tot[r, s] == eng.sum_tet_counts(s, r) == int(eng.get_tet_counts(s, r).sum())
    w = W.calcium_membrane_cube(n=4, side=2e-6, seed=1, end_time=1e-4, ca_um=50.0)
    meng = _engine(w, 5)
    meng.run(w.end_time)
    tt = meng.sum_all_tri_counts()
    assert tt.shape == (1, len(w.fm.sspecs))
    assert all(tt[0, q] == int(meng.get_tri_counts(q).sum()) for q in range(len(w.fm.sspecs)))
    assert tt.sum() > 0
