"""Per-element constants, activation and clamping on the GPU, bit for bit
against the CPU oracle with the same tables.

The tables are the device form of TetOpSplitP's per-tet and per-triangle
setters (_setTetReacK / _setTetReacActive, tetopsplit.cpp:3996-4100;
_setTetDiffD with a direction / _setTetDiffActive, :4103-4188;
_setTetSpecClamped, :3951; _setTriSReacK, :4464; the patch and surface
diffusion counterparts).
"""
import numpy as np
import pytest

from oracle import bind
from steps_amd import workloads as W
from steps_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _pair(w, seed, mode=0):
    eng = Engine(w.fm, [w.mt], [seed], mode=mode)
    for s in range(len(w.fm.specs)):
        eng.set_tet_counts(s, w.counts[s])
    for s in range(len(w.fm.sspecs)):
        eng.set_tri_counts(s, w.tri_counts[s])
    o = bind.OracleSim(w.fm, w.mt, seed=seed)
    o.counts[:] = w.counts
    if len(w.fm.sspecs):
        o.tri_counts[:] = w.tri_counts
    return eng, o


def _check(eng, o, w, t, mode):
    eng.run(t)
    o.run(t, mode)
    got = np.stack([eng.get_tet_counts(s) for s in range(len(w.fm.specs))])
    assert np.array_equal(got, o.counts), f"t={t}"
    if len(w.fm.sspecs):
        tri = np.stack([eng.get_tri_counts(s) for s in range(len(w.fm.sspecs))])
        assert np.array_equal(tri, o.tri_counts[:len(w.fm.sspecs)]), f"triangles t={t}"
    assert eng.extents() == (o.events.value, o.moved.value)
    assert eng.update_period() == o.update_period()


@pytest.mark.parametrize("mode", [0, 1])
def test_volume_tables_bit_exact(mode):
    w = W.four_species_cube(n=5, side=5e-6, seed=3, end_time=0.03)
    S, R, N = len(w.fm.specs), len(w.fm.reacs), w.n_tets
    rng = np.random.default_rng(7)
    k = w.fm.reac_k[:, None] * rng.uniform(0.3, 3.0, size=(R, N))
    k[rng.random((R, N)) < 0.2] = 0.0                      # inactive in some tets
    D = w.fm.spec_dcst[:, None, None] * rng.uniform(0.5, 1.5, size=(S, N, 4))
    D[rng.random((S, N, 4)) < 0.1] = 0.0                   # closed directions
    clamp = np.where(rng.random(N) < 0.1, 1 << 1, 0).astype(np.uint32)   # B clamped
    clamp[rng.random(N) < 0.05] |= 1 << 3                                # D clamped
    print(f"case volume tables mode={mode}", flush=True)
    eng, o = _pair(w, 17, mode)
    eng.set_tet_reac_table(k)
    eng.set_tet_diff_table(D)
    eng.set_tet_clamp(clamp)
    o.set_tables(tet_k=k, tet_fD=D, tet_clamp=clamp)
    _check(eng, o, w, 0.013, mode)
    _check(eng, o, w, w.end_time, mode)
    b = eng.get_tet_counts(1)
    assert np.array_equal(b[(clamp >> 1) & 1 == 1], w.counts[1][(clamp >> 1) & 1 == 1])
    # tables removed: the realization's constants apply again
    eng.set_tet_reac_table(None)
    eng.set_tet_diff_table(None)
    eng.set_tet_clamp(None)
    o.set_tables()
    _check(eng, o, w, w.end_time + 0.01, mode)


def test_volume_tables_with_graph_and_launch_shapes():
    w = W.four_species_cube(n=5, side=5e-6, seed=9, end_time=0.02)
    S, R, N = len(w.fm.specs), len(w.fm.reacs), w.n_tets
    rng = np.random.default_rng(1)
    k = w.fm.reac_k[:, None] * rng.uniform(0.5, 2.0, size=(R, N))
    clamp = np.where(rng.random(N) < 0.2, 1, 0).astype(np.uint32)
    print("case volume tables, graph replay and block 1024", flush=True)
    res = []
    for block, graph in ((256, False), (1024, False), (128, True)):
        eng, o = _pair(w, 5)
        eng.set_specialize(True)            # the tables take the generic membrane kernel
        eng.set_launch(block, graph)
        eng.set_tet_reac_table(k)
        eng.set_tet_clamp(clamp)
        eng.run(w.end_time / 2)
        eng.run(w.end_time)
        res.append(np.stack([eng.get_tet_counts(s) for s in range(S)]))
    o.set_tables(tet_k=k, tet_clamp=clamp)
    o.run(w.end_time / 2, 0)
    o.run(w.end_time, 0)
    for r in res:
        assert np.array_equal(r, o.counts)


@pytest.mark.parametrize("kind", ["ghk", "vdep", "sdiff"])
def test_surface_tables_bit_exact(kind):
    kw = dict(ghk=dict(ca_um=20.0, ghk_influx=True), vdep=dict(ca_um=20.0, influx=3e6, vdep_influx=True),
              sdiff=dict(ca_um=30.0, pump_density=300.0, pump_dcst=1e-12))[kind]
    w = W.calcium_membrane_cube(n=6, side=2e-6, seed=8, end_time=3e-4, buf_um=100.0, **kw)
    T, RS, NG, SS = len(w.mt.tri_tet), len(w.fm.sreacs), len(w.fm.ghk_chan), len(w.fm.sspecs)
    rng = np.random.default_rng(11)
    sk = np.empty((RS + NG, T))
    for q in range(RS):
        if w.fm.sreac_vdep[q] >= 0:
            sk[q] = (rng.random(T) < 0.7).astype(float)          # k(V) reaction on / off
        else:
            sk[q] = w.fm.sreac_k[q] * rng.uniform(0.2, 5.0, T)
            sk[q, rng.random(T) < 0.2] = 0.0
    for g in range(NG):
        sk[RS + g] = (rng.random(T) < 0.6).astype(float)          # GHK current on / off
    tclamp = np.where(rng.random(T) < 0.15, 1 << w.fm.sspec_index("Pump"), 0).astype(np.uint32)
    vclamp = np.where(rng.random(w.n_tets) < 0.1, 1 << w.fm.spec_index("Ca"), 0).astype(np.uint32)
    print(f"case surface tables {kind}", flush=True)
    eng, o = _pair(w, 23)
    v = np.linspace(-0.09, 0.03, T)
    if kind in ("ghk", "vdep"):
        eng.set_tri_v(v)
        o.tri_v[:] = v
    eng.set_tri_sreac_table(sk)
    eng.set_tri_clamp(tclamp)
    eng.set_tet_clamp(vclamp)
    tabs = dict(tri_sk=sk, tri_clamp=tclamp, tet_clamp=vclamp)
    if kind == "sdiff":
        eD = w.fm.sspec_dcst[:, None, None] * rng.uniform(0.5, 2.0, size=(SS, T, 3))
        eD[rng.random((SS, T, 3)) < 0.1] = 0.0
        eng.set_tri_sdiff_table(eD)
        tabs["tri_eD"] = eD
    o.set_tables(**tabs)
    _check(eng, o, w, 1e-4, 0)
    _check(eng, o, w, w.end_time, 0)
    pump = eng.get_tri_counts(w.fm.sspec_index("Pump"))
    held = (tclamp >> w.fm.sspec_index("Pump")) & 1 == 1
    assert np.array_equal(pump[held], w.tri_counts[w.fm.sspec_index("Pump")][held])


def test_clamped_source_bit_exact_and_constant():
    """The oracle's closed-form case (A clamped at 100 per tet feeding B, which
    decays) on the GPU: A never changes and every count equals the oracle's."""
    from test_oracle import _clamp_decay_case

    o, fm, mt = _clamp_decay_case(3, t_end=1.0)
    print("case clamped source", flush=True)
    eng = Engine(fm, [mt], [3])
    eng.set_tet_counts(0, np.full(len(mt.vol), 100))
    eng.set_tet_clamp(np.ones(len(mt.vol), dtype=np.uint32))
    eng.run(1.0)
    o.run(1.0, 0)
    assert np.all(eng.get_tet_counts(0) == 100)
    assert np.array_equal(eng.get_tet_counts(1), o.counts[1])
    assert eng.extents() == (o.events.value, o.moved.value)
