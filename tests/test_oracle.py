"""Pins the CPU oracle (oracle/stg_oracle.c) before it is trusted as the checker.

- Philox4x32-8 against known-answer vectors produced by the reference's own
  vendored Random123 header (tests/golden/philox4x32_8_kat.json, and live
  against oracle/_ref/philox_kat when it has been built from /root/reference).
- log / sqrt / log-factorial against exact decimal arithmetic.
- Binomial sampler against the exact pmf (chi-square, both algorithms, both
  flips and the degenerate cases), with a power check that a wrong p fails.
- Period schedules against the reference's loop structure worked by hand.
- Split period against the closed-form Kuhn-cube value.
- SSA kinetics against the closed forms of the reference's validation suite
  (test/validation/API_2/validation_rd_mpi/parallel_spatial_tetopsplit.py,
  same constants and `tolerable` criterion), on many isolated tets.
- Diffusion: exact conservation, leaving counts ~ Binomial, face split ~
  r_f / sum r, time-averaged occupancy, point-source profile vs the 3-D
  Green's function, equilibrium uniformity.
- Voltage-dependent surface reactions: the k(V) table as the reference fills
  it, bit-exact agreement with a constant-rate reaction at exactly
  representable interpolation points, first-order decay at a nonlinear k(V),
  and the out-of-range error.
- Channel-count time integral against the first-order closed form.
- GHK currents: exp against exact decimal arithmetic, setPInfo's permeability
  against the slope conductance it must reproduce, event rate and charge
  against the textbook GHK current for cations, anions and both directions,
  the computeflux switch, vshift and temperature.
"""
import decimal
import json
import math
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

from oracle import bind
from steps_amd import geom, model as M, workloads as W

GOLDEN = Path(__file__).resolve().parent / "golden"
AV = 6.02214179e23


def tolerable(a, b, tol):
    # test/validation/API_2/validation_rd_mpi/tol_funcs.py
    return a == b == 0 or (a + b != 0 and abs(2 * (a - b) / (a + b)) <= tol)


# ------------------------------------------------------------------ Philox

def test_philox_matches_reference_kat_fixture():
    kat = json.loads((GOLDEN / "philox4x32_8_kat.json").read_text())
    assert len(kat["cases"]) >= 60
    for c in kat["cases"]:
        assert bind.philox(c["ctr"], c["key"]) == c["out"]


def test_philox_matches_live_reference_header():
    rnd = np.random.default_rng(7)
    lines = [tuple(int(x) for x in rnd.integers(0, 2**32, 6)) for _ in range(200)]
    ref = bind.ref_philox_kat(lines)
    if ref is None:
        pytest.skip("oracle/_ref/philox_kat not built (needs the reference tree)")
    for ln, r in zip(lines, ref):
        assert tuple(bind.philox(ln[:4], ln[4:6])) == r[6:]


def test_uniforms_open_interval_and_uniform():
    u = bind.uniforms(5, 1, 2, 3, 4, 200_000)
    assert u.min() > 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3


# ------------------------------------------------------------- scalar math

def _ulps(a, b):
    return abs(a - b) / math.ulp(b) if b != 0 else abs(a) / 5e-324


def test_log_within_two_ulp_of_exact():
    lib = bind.load()
    decimal.getcontext().prec = 40
    rng = np.random.default_rng(1)
    xs = np.concatenate([
        (rng.integers(0, 2**32, 2000) + 0.5) * 2.0 ** -32,   # the uniforms the engine feeds it
        np.exp(rng.uniform(-700, 700, 2000)),
        1.0 + rng.uniform(-1e-3, 1e-3, 500),
        [1.0, 2.0, 0.5, 1e-300, 5e-324, 1.7976931348623157e308],
    ])
    worst = 0.0
    for x in xs:
        exact = float(decimal.Decimal(float(x)).ln())
        got = lib.stgo_log(float(x))
        if exact == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, _ulps(got, exact))
    assert worst <= 2.0, worst


def test_exp_within_one_ulp_of_exact():
    lib = bind.load()
    decimal.getcontext().prec = 40
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-700, 700, 3000), rng.uniform(-1, 1, 2000),
                         [0.0, 1e-300, -1e-300, 0.34657359027997264, -0.34657359027997264, 709.7]])
    worst = 0.0
    for x in xs:
        exact = float(decimal.Decimal(float(x)).exp())
        worst = max(worst, _ulps(lib.stgo_exp(float(x)), exact))
    assert worst <= 1.0, worst
    assert lib.stgo_exp(710.0) == math.inf and lib.stgo_exp(-746.0) == 0.0


def test_sqrt_is_correctly_rounded():
    lib = bind.load()
    rng = np.random.default_rng(2)
    for x in np.exp(rng.uniform(-300, 300, 5000)):
        assert lib.stgo_sqrt(float(x)) == math.sqrt(float(x))


def test_log_factorial_table_exact_and_tail():
    lib = bind.load()
    decimal.getcontext().prec = 50
    acc = decimal.Decimal(0)
    for k in range(0, 1024):
        if k > 0:
            acc += decimal.Decimal(k).ln()
        assert lib.stgo_lfact(k) == float(acc), k
    for k in (1024, 2000, 10**5, 10**7):
        assert abs(lib.stgo_lfact(k) - math.lgamma(k + 1)) <= 1e-13 * math.lgamma(k + 1)


# --------------------------------------------------------------- binomial

def _chi2_pvalue(sample, n, p):
    """Chi-square of a sample against Binomial(n, p), bins merged to >= 5 expected."""
    k = np.arange(n + 1)
    pmf = stats.binom.pmf(k, n, p)
    obs = np.bincount(sample, minlength=n + 1).astype(float)
    exp = pmf * len(sample)
    # merge from both tails toward the mode
    edges, cur_o, cur_e = [], 0.0, 0.0
    for o, e in zip(obs, exp):
        cur_o += o
        cur_e += e
        if cur_e >= 5:
            edges.append((cur_o, cur_e))
            cur_o = cur_e = 0.0
    if edges and (cur_e > 0 or cur_o > 0):
        o, e = edges.pop()
        edges.append((o + cur_o, e + cur_e))
    o = np.array([x[0] for x in edges])
    e = np.array([x[1] for x in edges])
    if len(o) < 2:
        return 1.0
    return stats.chisquare(o, e * o.sum() / e.sum()).pvalue


@pytest.mark.parametrize("n,p", [
    (1, 0.3), (10, 0.5), (20, 0.04), (50, 0.19), (50, 0.21), (100, 0.5), (1000, 0.003),
    (1000, 0.7), (30, 0.97), (200, 0.999), (100_000, 0.3), (5000, 0.0015), (11, 0.95),
    # either side of the inversion / rejection switch at n*min(p, 1-p) = 30, and
    # inversion with a long search from a large n
    (100, 0.29), (100, 0.31), (60, 0.51), (200_000, 0.000149),
])
def test_binomial_matches_exact_pmf(n, p):
    x = bind.binomials(42, n, 7, 3, 9, n, p, 200_000)
    assert x.min() >= 0 and x.max() <= n
    assert _chi2_pvalue(x, n, p) > 1e-3
    z = (x.mean() - n * p) / math.sqrt(n * p * (1 - p) / len(x))
    assert abs(z) < 4.5


def test_binomial_check_has_power():
    """The chi-square check rejects samples drawn with a slightly wrong p."""
    x = bind.binomials(42, 1, 2, 3, 4, 100, 0.51, 200_000)
    assert _chi2_pvalue(x, 100, 0.50) < 1e-6
    y = bind.binomials(42, 1, 2, 3, 4, 40, 0.105, 200_000)
    assert _chi2_pvalue(y, 40, 0.10) < 1e-6


@pytest.mark.parametrize("n,p,expect", [(0, 0.5, 0), (7, 0.0, 0), (7, 1.0, 7), (7, 1.5, 7)])
def test_binomial_degenerate(n, p, expect):
    assert np.all(bind.binomials(1, 0, 0, 0, 0, n, p, 100) == expect)


# --------------------------------------------------------------- schedule

def test_schedule_tetopsplit_clips_last_period():
    # tetopsplit.cpp:1545-1558
    assert np.allclose(bind.schedule(0.0, 0.1, 0.0462, 0), [0.0462, 0.0462, 0.0076])
    assert np.allclose(bind.schedule(0.0, 0.1, 1.0, 0), [0.1])
    assert len(bind.schedule(0.3, 0.3, 0.01, 0)) == 0
    assert np.allclose(bind.schedule(0.0, 0.1, np.inf, 0), [0.1])


def test_schedule_dist_standard_then_remainders():
    # simulation.cpp:1358-1384: floor((T-t)/dt)-1 standard periods, then
    # either dt + remainder or a single remainder
    assert np.allclose(bind.schedule(0.0, 0.1, 0.0462, 1), [0.0462, 0.0462, 0.0076])
    d = bind.schedule(0.0, 0.3, 0.1, 1)
    assert len(d) == 3 and np.allclose(d, 0.1)
    d = bind.schedule(0.0, 0.35, 0.1, 1)
    assert np.allclose(d, [0.1, 0.1, 0.1, 0.05])
    assert np.allclose(bind.schedule(0.0, 0.05, 0.1, 1), [0.05])
    with pytest.raises(ValueError):
        bind.schedule(1.0, 0.5, 0.1, 1)


# --------------------------------------------------------- geometry / dt

@pytest.mark.parametrize("n,side,D,rate", [
    (8, 10e-6, 1e-12, 21.6307), (40, 40e-6, 1e-11, 337.98), (4, 0.8e-6, 2e-10, 1.68990e5)])
def test_update_period_kuhn_closed_form(n, side, D, rate):
    """max_t sum_f D A_f/(V d_f) = 33.798 D / h^2 on a Kuhn-split cube (h = side/n),
    the value of the independent geometry computation for these cubes."""
    mesh = geom.kuhn_cube(n, side)
    mt = geom.mesh_tables(mesh, geom.TmComp("c", mesh))
    fm = M.flatten(W.ab_binding_model(dA=D, dB=D, dC=D / 2), ["vsys"])
    o = bind.OracleSim(fm, mt, seed=0)
    assert abs(1.0 / o.update_period() / rate - 1) < 2e-5
    h = side / n
    assert abs(1.0 / o.update_period() - 33.798 * D / h**2) < 1e-4 * 33.798 * D / h**2


def test_kuhn_cube_topology():
    mesh = geom.kuhn_cube(8, 10e-6)
    assert mesh.countTets() == 3072 and mesh.countVertices() == 729
    assert len(mesh.getSurfTris()) == 768
    assert abs(mesh.getMeshVolume() - 1e-15) < 1e-27
    assert np.allclose(mesh.tet_vol, 1e-15 / 3072)


def test_face_factor_reciprocity_on_irregular_mesh():
    """A_f/d_f is shared by both sides of a face: g_ij V_i = g_ji V_j."""
    mesh0 = geom.kuhn_cube(4, 1e-6)
    rng = np.random.default_rng(3)
    v = mesh0.verts + rng.uniform(-0.08e-6, 0.08e-6, mesh0.verts.shape)
    mesh = geom.Tetmesh(v, mesh0.tets)
    mt = geom.mesh_tables(mesh, geom.TmComp("c", mesh))
    for t in range(mesh.countTets()):
        for f in range(4):
            j = mt.nbr[t, f]
            if j < 0:
                assert mt.face_g[t, f] == 0
                continue
            jf = mt.nbr_face[t, f]
            assert mt.nbr[j, jf] == t
            assert math.isclose(mt.face_g[t, f] * mt.vol[t], mt.face_g[j, jf] * mt.vol[j],
                                rel_tol=1e-12)


# ------------------------------------------------------- SSA kinetics

def _isolated(n_tets, vol, with_tris=False, area=1e-12):
    z4 = -np.ones((n_tets, 4), dtype=np.int32)
    return geom.MeshTables(
        vol=np.full(n_tets, vol), nbr=z4, nbr_face=z4.copy(), face_g=np.zeros((n_tets, 4)),
        tri_tet=np.arange(n_tets, dtype=np.int32) if with_tris else np.zeros(0, np.int32),
        tri_area=np.full(n_tets, area) if with_tris else np.zeros(0))


def _single_reaction_model(lhs, rhs, k, names):
    mdl = M.Model()
    sp = {x: M.Spec(x, mdl) for x in names}
    vs = M.Volsys("vsys", mdl)
    M.Reac("r", vs, lhs=[sp[x] for x in lhs], rhs=[sp[x] for x in rhs], kcst=k)
    return M.flatten(mdl, ["vsys"])


def _trajectory(fm, mt, init, tpnts, seed=1):
    o = bind.OracleSim(fm, mt, seed=seed)
    for s, c in init.items():
        o.counts[fm.spec_index(s)] = c
    out = []
    for t in tpnts:
        o.run(t, 0)
        out.append(o.counts.copy())
    return np.array(out), o


TP = np.arange(1, 12) * 0.1   # DT = 0.1, INT = 1.1 (parallel_spatial_tetopsplit.py:160-161)
VOL = 4.18879e-18             # volume of the reference's unit-radius validation sphere


def test_first_order_irreversible():
    # KCST 5, N 50, tolerance 2% on mean and std (parallel_spatial_tetopsplit.py:44-55, 306-314)
    fm = _single_reaction_model(["A"], [], 5.0, ["A"])
    traj, _ = _trajectory(fm, _isolated(100_000, VOL), {"A": 50}, TP)
    for i, t in enumerate(TP):
        a = traj[i, 0]
        analy = 50 * math.exp(-5 * t)
        std = math.sqrt(50 * math.exp(-5 * t) * (1 - math.exp(-5 * t)))
        assert tolerable(analy, a.mean(), 0.02), (t, a.mean(), analy)
        assert tolerable(std, a.std(), 0.02), (t, a.std(), std)


def test_first_order_reversible_equilibrium():
    # kf 10, kb 2; equilibrium A/B = kb/kf (parallel_spatial_tetopsplit.py:57-66, 316-323)
    fm = _single_reaction_model(["A"], ["B"], 10.0, ["A", "B"])
    fm.reac_lhs = np.array([[1, 0], [0, 1]], np.int32)
    fm.reac_upd = np.array([[-1, 1], [1, -1]], np.int32)
    fm.reac_k = np.array([10.0, 2.0])
    fm.reacs = ["f", "b"]
    traj, _ = _trajectory(fm, _isolated(500, VOL), {"A": 1000}, TP)
    aeq = 1000 * (2.0 / 10.0) / (1 + 2.0 / 10.0)
    for i in range(7, len(TP)):
        assert tolerable(traj[i, 0].mean(), aeq, 0.01)
        assert tolerable(traj[i, 1].mean(), 1000 - aeq, 0.01)


def _conc(counts, vol=VOL):
    return counts / (1e3 * vol * AV)


def test_second_order_2A():
    # 2A -> C, k 10e6, [A0] 10 uM: 1/[A] = 1/[A0] + 2kt within 3%
    fm = _single_reaction_model(["A", "A"], ["C"], 10.0e6, ["A", "C"])
    n0 = int(round(10e-6 * 1e3 * VOL * AV))
    traj, _ = _trajectory(fm, _isolated(200, VOL), {"A": n0}, TP)
    a0 = _conc(n0)
    for i, t in enumerate(TP):
        assert tolerable(1.0 / _conc(traj[i, 0].mean()), 1.0 / a0 + 2 * 10.0e6 * t, 0.03)


def test_second_order_AB_equal():
    # A + B -> C, k 5e6, A0 = B0 = 20 uM: 1/[A] = 1/[A0] + kt within 2%
    fm = _single_reaction_model(["A", "B"], ["C"], 5.0e6, ["A", "B", "C"])
    n0 = int(round(20e-6 * 1e3 * VOL * AV))
    traj, _ = _trajectory(fm, _isolated(200, VOL), {"A": n0, "B": n0}, TP)
    a0 = _conc(n0)
    for i, t in enumerate(TP):
        assert tolerable(1.0 / _conc(traj[i, 0].mean()), 1.0 / a0 + 5.0e6 * t, 0.02)
        assert tolerable(1.0 / _conc(traj[i, 1].mean()), 1.0 / a0 + 5.0e6 * t, 0.02)


def test_second_order_AB_unequal():
    # A0 = 1 uM, B0 = 0.5 uM: ln(B/A) = ln(B0/A0) - (A0 - B0) k t within 1%
    fm = _single_reaction_model(["A", "B"], ["C"], 5.0e6, ["A", "B", "C"])
    na = int(round(1e-6 * 1e3 * VOL * AV))
    nb = na // 2
    traj, _ = _trajectory(fm, _isolated(4000, VOL), {"A": na, "B": nb}, TP)
    a0, b0 = _conc(na), _conc(nb)
    for i, t in enumerate(TP):
        a, b = _conc(traj[i, 0].mean()), _conc(traj[i, 1].mean())
        assert tolerable(math.log(b / a), math.log(b0 / a0) - (a0 - b0) * 5.0e6 * t, 0.01)


def test_third_order_3A():
    # 3A -> C, k 1e12, [A0] 10 uM: 1/[A]^2 = 1/[A0]^2 + 6kt within 3%
    fm = _single_reaction_model(["A", "A", "A"], ["C"], 1.0e12, ["A", "C"])
    n0 = int(round(10e-6 * 1e3 * VOL * AV))
    traj, _ = _trajectory(fm, _isolated(200, VOL), {"A": n0}, TP)
    a0 = _conc(n0)
    for i, t in enumerate(TP):
        assert tolerable(1.0 / _conc(traj[i, 0].mean()) ** 2, 1.0 / a0**2 + 6 * 1.0e12 * t, 0.03)


def test_third_order_2A_B():
    # 2A + B -> C, k 0.1e12, A0 30 uM, B0 20 uM (parallel_spatial_tetopsplit.py:128-138, 340-357)
    fm = _single_reaction_model(["A", "A", "B"], ["C"], 0.1e12, ["A", "B", "C"])
    na = int(round(30e-6 * 1e3 * VOL * AV))
    nb = int(round(20e-6 * 1e3 * VOL * AV))
    traj, _ = _trajectory(fm, _isolated(300, VOL), {"A": na, "B": nb}, TP)
    A0, B0 = _conc(na), _conc(nb)
    d_ab, d_ba = A0 - 2 * B0, 2 * B0 - A0
    for i, t in enumerate(TP):
        A, B = _conc(traj[i, 0].mean()), _conc(traj[i, 1].mean())
        line = (-1.0 / d_ab) * ((-1.0 / d_ba) * math.log((B / A) / (B0 / A0)) + 1.0 / A - 1.0 / A0)
        assert tolerable(t * 0.1e12, line, 0.01)


def test_zero_order_production():
    k = 2e-9  # M/s -> rate k * 1e3 * V * NA per tet
    fm = _single_reaction_model([], ["A"], k, ["A"])
    traj, _ = _trajectory(fm, _isolated(5000, VOL), {}, [0.5])
    expect = k * 1e3 * VOL * AV * 0.5
    a = traj[0, 0]
    assert abs(a.mean() - expect) < 4 * math.sqrt(expect / len(a))
    kmax = int(a.max())
    obs = np.bincount(a, minlength=kmax + 1)[: kmax + 1].astype(float)
    pmf = stats.poisson(expect).pmf(np.arange(kmax + 1))
    pmf[-1] += stats.poisson(expect).sf(kmax)
    keep = pmf * len(a) >= 5
    o = np.append(obs[keep], obs[~keep].sum())
    e = np.append(pmf[keep], pmf[~keep].sum()) * len(a)
    assert stats.chisquare(o, e * o.sum() / e.sum()).pvalue > 1e-3


def _surface_model(slhs, srhs, ilhs, irhs, k):
    mdl = M.Model()
    names = sorted(set(slhs + srhs + ilhs + irhs + ["V"]))
    sp = {x: M.Spec(x, mdl) for x in names}
    vs = M.Volsys("vsys", mdl)
    M.Diff("dV", vs, sp["V"], 0.0)
    ss = M.Surfsys("ssys", mdl)
    M.SReac("sr", ss, slhs=[sp[x] for x in slhs], srhs=[sp[x] for x in srhs],
            ilhs=[sp[x] for x in ilhs], irhs=[sp[x] for x in irhs], kcst=k)
    return M.flatten(mdl, ["vsys"], ["ssys"])


def test_surface_second_order_2d():
    # A.s + B.s -> C.s, K 10e10, counts 100 / 50, area 10 um^2: ln(B/A) line within 2%
    # (parallel_spatial_tetopsplit.py:140-150, 400-410), CCST = K / (NA * area)
    area = 10.0e-12
    fm = _surface_model(["A", "B"], ["C"], [], [], 10.0e10)
    mt = _isolated(2000, VOL, with_tris=True, area=area)
    o = bind.OracleSim(fm, mt, seed=3)
    o.tri_counts[fm.sspec_index("A")] = 100
    o.tri_counts[fm.sspec_index("B")] = 50
    ccst = 10.0e10 / (AV * area)
    for t in TP:
        o.run(t, 0)
        a = o.tri_counts[fm.sspec_index("A")].mean()
        b = o.tri_counts[fm.sspec_index("B")].mean()
        assert tolerable(math.log(b / a), math.log(50 / 100) - 50 * ccst * t, 0.02)


def test_surface_volume_reaction_uses_inner_volume_scale():
    # P.s + Ca.i -> P.s: rate k/(1e3 V NA) * P * Ca, first order in Ca at fixed P
    k = 1e8
    fm = _surface_model(["P"], ["P"], ["Ca"], [], k)
    mt = _isolated(3000, VOL, with_tris=True)
    o = bind.OracleSim(fm, mt, seed=4)
    o.tri_counts[fm.sspec_index("P")] = 3
    o.counts[fm.spec_index("Ca")] = 200
    o.run(0.2, 0)
    rate = k / (1e3 * VOL * AV) * 3
    assert tolerable(o.counts[fm.spec_index("Ca")].mean(), 200 * math.exp(-rate * 0.2), 0.01)


# ------------------------------------------------ voltage-dependent surface reactions

def _gate_model(kfun=None, vrange=None, kcst=None):
    """C.s -> O.s, either voltage dependent (kfun over vrange) or constant kcst."""
    mdl = M.Model()
    c, o_, vol = M.Spec("C", mdl), M.Spec("O", mdl), M.Spec("V", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Diff("dV", vs, vol, 0.0)
    ss = M.Surfsys("ssys", mdl)
    if kfun is None:
        M.SReac("open", ss, slhs=[c], srhs=[o_], kcst=kcst)
    else:
        M.VDepSReac("open", ss, slhs=[c], srhs=[o_], k=kfun, vrange=vrange)
    return M.flatten(mdl, ["vsys"], ["ssys"])


# Exactly representable grid: V = -0.125 + i * 2^-10, k linear in V
VR_EXACT = (-0.125, 0.125, 2.0 ** -10)


def _k_lin(v):
    return 1000.0 + 8000.0 * (v + 0.125)


def test_vdep_table_matches_reference_fill():
    # int((vmax - vmin)/dv) + 1 entries at V accumulated by repeated addition
    # (pysteps/steps/API_1/model.py:510-517); default range -150..100 mV step 0.1 mV
    fm = _gate_model(lambda v: v + 1.0, (-150.0e-3, 100.0e-3, 1.0e-4))
    assert fm.vdep_off.tolist() == [0, 2501]   # 2500 intervals of 0.1 mV plus the end point
    v, want = -150.0e-3, []
    for _ in range(2501):
        want.append(v + 1.0)
        v += 1.0e-4
    assert np.array_equal(fm.vdep_k, np.array(want))
    assert fm.sreac_vdep.tolist() == [0]


@pytest.mark.parametrize("v,k_expect", [
    (-0.0625, 1500.0),                        # on grid point 64: r = 0 picks K[64] alone
    (-0.0625 + 2.0 ** -12, 1501.953125),      # v2 = 64.25: 0.75 K[64] + 0.25 K[65]
])
def test_vdep_equals_constant_rate_at_interpolated_k(v, k_expect):
    # every value here is exact in binary, so k(V) must reproduce the constant
    # surface reaction with the interpolated k bit for bit; a swapped weight
    # (0.25 K[64] + 0.75 K[65] = 1505.859375) or an off-by-one index fails
    area = 1e-13
    mt = _isolated(400, VOL, with_tris=True, area=area)
    fv = _gate_model(_k_lin, VR_EXACT)
    fc = _gate_model(kcst=k_expect)
    ov, oc = bind.OracleSim(fv, mt, seed=9), bind.OracleSim(fc, mt, seed=9)
    for o, fm in ((ov, fv), (oc, fc)):
        o.tri_counts[fm.sspec_index("C")] = 50
    ov.tri_v[:] = v
    for t in (1e-4, 3e-4):
        ov.run(t, 0)
        oc.run(t, 0)
        assert np.array_equal(ov.tri_counts, oc.tri_counts)
    assert 0 < ov.tri_counts[fv.sspec_index("O")].sum() < 400 * 50


def test_vdep_first_order_follows_k_of_v():
    # a nonlinear k(V) (Hodgkin-Huxley alpha_n, per second), triangles at several
    # potentials: surviving fraction exp(-k(V) t) with k read off the table by
    # numpy's piecewise-linear interpolation
    def alpha_n(v):
        vm = v * 1e3
        return 1e3 * 0.01 * (10 - (vm + 65)) / (math.exp((10 - (vm + 65)) / 10) - 1)

    vr = (-100.0e-3, 50.0e-3, 1.0e-4)
    fm = _gate_model(alpha_n, vr)
    vs = np.array([-80.0e-3, -65.0e-3, -40.33e-3, 12.345e-3])
    n_per = 1500
    mt = _isolated(n_per * len(vs), VOL, with_tris=True)
    o = bind.OracleSim(fm, mt, seed=11)
    o.tri_counts[fm.sspec_index("C")] = 20
    o.tri_v[:] = np.repeat(vs, n_per)
    t = 2e-3
    o.run(t, 0)
    grid = vr[0] + np.arange(len(fm.vdep_k)) * vr[2]
    left = o.tri_counts[fm.sspec_index("C")].reshape(len(vs), n_per).mean(axis=1)
    for v, m in zip(vs, left):
        k = np.interp(v, grid, fm.vdep_k)
        assert tolerable(m, 20 * math.exp(-k * t), 0.03), (v, m, k)


def test_channel_count_time_integral():
    # open channels O.s -> C.s at k from n0: E[int_0^T n dt] = n0 (1 - e^{-kT}) / k;
    # T spans five periods (one per run call, nothing diffuses), so a correction
    # taken from the absolute time instead of the period start, or with the
    # wrong sign, misses by far more than the 1% tolerance
    k, n0, T = 400.0, 30, 5e-3
    mdl = M.Model()
    c, o_, vol = M.Spec("C", mdl), M.Spec("O", mdl), M.Spec("V", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Diff("dV", vs, vol, 0.0)
    ss = M.Surfsys("ssys", mdl)
    M.SReac("close", ss, slhs=[o_], srhs=[c], kcst=k)
    fm = M.flatten(mdl, ["vsys"], ["ssys"])
    mt = _isolated(4000, VOL, with_tris=True)
    o = bind.OracleSim(fm, mt, seed=21)
    qo, qc = fm.sspec_index("O"), fm.sspec_index("C")
    o.tri_counts[qo] = n0
    o.integrate_tri_species([qo, qc])
    for i in range(1, 6):
        o.run(i * T / 5, 0)
    assert tolerable(o.tri_intg[qo].mean(), n0 * (1 - math.exp(-k * T)) / k, 0.01)
    # O + C is constant per triangle, so the two integrals add up to n0 * T
    assert np.allclose(o.tri_intg[qo] + o.tri_intg[qc], n0 * T, rtol=1e-12)
    # a species no event touches integrates to n * T exactly up to rounding
    o2 = bind.OracleSim(fm, mt, seed=21)
    o2.tri_counts[qc] = 7
    o2.integrate_tri_species([qc])
    for i in range(1, 6):
        o2.run(i * T / 5, 0)
    assert np.allclose(o2.tri_intg[qc], 7 * T, rtol=1e-14)


FARADAY, GAS_R = 96485.3399, 8.314472


def _ghk_current(P, V, z, T, ci, co):
    """Textbook GHK current (A) of one channel, concentrations in M."""
    ci, co = ci * 1e3, co * 1e3
    x = z * V * FARADAY / (GAS_R * T)
    return P * z * z * FARADAY ** 2 * V / (GAS_R * T) * (ci - co * math.exp(-x)) / (1 - math.exp(-x))


def test_ghk_permeability_gives_slope_conductance():
    # setPInfo(g, V, T, oconc, iconc) picks P so that dI/dV = g at V
    for g, V, z, ci, co in [(20e-12, -0.05, 2, 1e-7, 2e-3), (5e-12, 0.03, 1, 0.14, 0.005),
                            (1e-12, -0.08, -1, 0.01, 0.12)]:
        P = M.ghk_permeability(g, V, z, 293.15, ci, co)
        h = 1e-7
        slope = (_ghk_current(P, V + h, z, 293.15, ci, co)
                 - _ghk_current(P, V - h, z, 293.15, ci, co)) / (2 * h)
        assert abs(slope / g - 1) < 1e-6


def _ghk_model(z, perm, voconc, realflux=True, vshift=0.0):
    mdl = M.Model()
    chan = M.Chan("ch", mdl)
    op = M.ChanState("Open", mdl, chan)
    ion = M.Spec("X", mdl, valence=z)
    vs = M.Volsys("vsys", mdl)
    M.Diff("dX", vs, ion, 0.0)
    ss = M.Surfsys("ssys", mdl)
    cur = M.GHKcurr("ghk", ss, op, ion, computeflux=realflux, virtual_oconc=voconc, vshift=vshift)
    cur.setP(perm)
    return M.flatten(mdl, ["vsys"], ["ssys"])


@pytest.mark.parametrize("z,V,ci_n,voconc", [
    (2, -0.065, 30, 2e-3),     # Ca influx against a large outward concentration ratio
    (1, 0.02, 5000, 0.005),    # K efflux at a depolarised potential
    (-1, -0.07, 3000, 0.12),   # Cl
])
def test_ghk_event_rate_and_charge(z, V, ci_n, voconc):
    # over a short run (inner count nearly constant) the mean charge moved out
    # per triangle is z * rt * channels * t with rt the textbook current / (z e)
    perm, nch, t = 2e-19, 3, 2e-4
    fm = _ghk_model(z, perm, voconc)
    mt = _isolated(3000, VOL, with_tris=True)
    o = bind.OracleSim(fm, mt, seed=31)
    o.counts[fm.spec_index("X")] = ci_n
    o.tri_counts[fm.sspec_index("Open")] = nch
    o.tri_v[:] = V
    o.run(t, 0)
    ci = ci_n / (1e3 * VOL * AV)
    rt = _ghk_current(perm, V, z, 293.15, ci, voconc) / (z * 1.602176487e-19)
    q = o.tri_charge.astype(float)
    assert np.sign(q.sum()) == np.sign(z * rt)   # charge counted outward positive
    assert tolerable(q.mean(), z * rt * nch * t, 0.03), (q.mean(), z * rt * nch * t)
    # every event moves one ion: ion change = -charge / z exactly, per triangle
    assert np.array_equal(o.counts[fm.spec_index("X")] - ci_n, -(o.tri_charge[:3000] // z))


def test_ghk_without_flux_moves_charge_only():
    fm = _ghk_model(1, 2e-19, 0.005, realflux=False)
    mt = _isolated(500, VOL, with_tris=True)
    o = bind.OracleSim(fm, mt, seed=3)
    o.counts[fm.spec_index("X")] = 4000
    o.tri_counts[fm.sspec_index("Open")] = 3
    o.tri_v[:] = 0.02
    o.run(2e-4, 0)
    assert np.all(o.counts[fm.spec_index("X")] == 4000) and o.tri_charge.sum() != 0


def test_ghk_vshift_and_temperature_enter_the_exponent():
    # vshift adds to V; T scales the exponent: (V=-0.0625, shift 0.015625)
    # must match (V=-0.046875, no shift) bit for bit (all exact in binary)
    mt = _isolated(200, VOL, with_tris=True)
    runs = []
    for V, shift in ((-0.0625, 0.015625), (-0.046875, 0.0)):
        fm = _ghk_model(2, 2e-19, 2e-3, vshift=shift)
        o = bind.OracleSim(fm, mt, seed=5)
        o.counts[fm.spec_index("X")] = 100
        o.tri_counts[fm.sspec_index("Open")] = 2
        o.tri_v[:] = V
        o.run(1e-4, 0)
        runs.append(o.tri_charge.copy())
    assert np.array_equal(runs[0], runs[1]) and runs[0].sum() != 0
    fm = _ghk_model(2, 2e-19, 2e-3)
    hot, cold = (bind.OracleSim(fm, mt, seed=5) for _ in range(2))
    for o, T in ((hot, 310.0), (cold, 273.0)):
        o.set_temp(T)
        o.counts[fm.spec_index("X")] = 100
        o.tri_counts[fm.sspec_index("Open")] = 2
        o.tri_v[:] = -0.05
        o.run(1e-4, 0)
    # influx grows in magnitude as |zFV/RT| grows
    assert hot.tri_charge.sum() > cold.tri_charge.sum()


def test_vdep_voltage_outside_table_is_an_error():
    # getVDepK raises for V above vmax or below vmin (vdepsreacdef.cpp:196-207)
    mt = _isolated(4, VOL, with_tris=True)
    fm = _gate_model(_k_lin, VR_EXACT)
    o = bind.OracleSim(fm, mt, seed=1)
    o.tri_counts[fm.sspec_index("C")] = 5
    o.tri_v[:] = 0.2
    with pytest.raises(RuntimeError, match="-2"):
        o.run(1e-4, 0)


# ---------------------------------------------------------------- diffusion

def _pairs(n_pairs, vol=1e-19, g=1e12):
    """Disjoint two-tet components joined through face 0 of each."""
    n = 2 * n_pairs
    nbr = -np.ones((n, 4), np.int32)
    nf = -np.ones((n, 4), np.int32)
    G = np.zeros((n, 4))
    a, b = np.arange(0, n, 2), np.arange(1, n, 2)
    nbr[a, 0], nbr[b, 0] = b, a
    nf[a, 0], nf[b, 0] = 0, 0
    G[a, 0] = G[b, 0] = g
    return geom.MeshTables(np.full(n, vol), nbr, nf, G, np.zeros(0, np.int32), np.zeros(0))


def _diff_model(D):
    mdl = M.Model()
    A = M.Spec("A", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Diff("dA", vs, A, D)
    return M.flatten(mdl, ["vsys"])


def test_leaving_count_is_binomial():
    D, g = 1e-12, 1e12            # sum of rates 1 /s, stability period 1 s
    fm = _diff_model(D)
    o = bind.OracleSim(fm, _pairs(20_000, g=g), seed=9)
    o.counts[0, 0::2] = 40
    o.run(0.3, 0)                 # one clipped period: p = 0.3
    moved = o.counts[0, 1::2]
    assert np.all(o.counts[0, 0::2] + moved == 40)
    assert _chi2_pvalue(moved, 40, 0.3) > 1e-3


@pytest.mark.parametrize("threshold", [0, 10, 1000])
def test_face_split_follows_rates(threshold):
    """Star of a centre tet and four neighbours with rates 1:2:3:4: each
    neighbour's count is Binomial(N, p r_f / sum r)."""
    n_star = 4000
    n = 5 * n_star
    nbr = -np.ones((n, 4), np.int32)
    nf = -np.ones((n, 4), np.int32)
    G = np.zeros((n, 4))
    gs = np.array([1.0, 2.0, 3.0, 4.0]) * 1e11
    c = np.arange(0, n, 5)
    for f in range(4):
        leaf = c + 1 + f
        nbr[c, f], nf[c, f], G[c, f] = leaf, 0, gs[f]
        nbr[leaf, 0], nf[leaf, 0], G[leaf, 0] = c, f, 1e9
    mt = geom.MeshTables(np.full(n, 1e-19), nbr, nf, G, np.zeros(0, np.int32), np.zeros(0))
    fm = _diff_model(1e-12)
    o = bind.OracleSim(fm, mt, seed=12, diff_threshold=threshold)
    N = 30
    o.counts[0, c] = N
    o.run(0.8, 0)                 # sum of rates 1 /s: p = 0.8
    assert np.all(o.counts[0].reshape(-1, 5).sum(axis=1) == N)
    for f in range(4):
        k = o.counts[0, c + 1 + f]
        assert _chi2_pvalue(k, N, 0.8 * gs[f] / gs.sum()) > 1e-3, f


def test_occupancy_is_time_average():
    """Molecules produced during a period leave according to their time-averaged
    count (mol_state.hpp:87-200): with p = 1, E[moved] = lambda dt / 2, not lambda dt."""
    D, g, vol = 1e-12, 1e12, 1e-19
    lam_dt = 20.0
    mdl = M.Model()
    A = M.Spec("A", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Reac("prod", vs, lhs=[], rhs=[A], kcst=lam_dt / (1e3 * vol * AV))
    M.Diff("dA", vs, A, D)
    fm = M.flatten(mdl, ["vsys"])
    o = bind.OracleSim(fm, _pairs(5000, vol=vol, g=g), seed=2)
    dt = o.update_period()
    assert math.isclose(dt, 1.0)
    o.run(dt, 0)
    per_tet = o.moved.value / o.n_tets
    assert abs(per_tet - lam_dt / 2) < 0.15, per_tet
    assert o.counts.sum() == o.events.value


def test_point_source_matches_greens_function():
    """10,000 molecules released in the centre tet of a 20^3 x 6 cube (side
    10 um), D = 1e-12: radial shell counts after 0.6 s follow
    N (4 pi D t)^{-3/2} exp(-r^2 / 4Dt) (the unbounded-diffusion oracle of
    test/validation/API_2/validation_rd/test_unbdiff2D.py, in 3-D)."""
    D, t, side, N = 1e-12, 0.6, 10e-6, 10_000
    mesh = geom.kuhn_cube(20, side)
    mt = geom.mesh_tables(mesh, geom.TmComp("c", mesh))
    fm = _diff_model(D)
    centre = mesh.findTetByPoint([side / 2 * 1.001, side / 2 * 0.999, side / 2 * 1.0005])
    r = np.linalg.norm(mesh.tet_bary - mesh.tet_bary[centre], axis=1)
    edges = np.linspace(0, 3.0e-6, 7)
    obs = np.zeros(len(edges) - 1)
    exp = np.zeros(len(edges) - 1)
    for seed in range(4):
        o = bind.OracleSim(fm, mt, seed=seed)
        o.counts[0, centre] = N
        o.run(t, 0)
        assert o.counts.sum() == N
        for i in range(len(edges) - 1):
            sel = (r >= edges[i]) & (r < edges[i + 1])
            obs[i] += o.counts[0, sel].sum()
            dens = N / (4 * math.pi * D * t) ** 1.5 * np.exp(-r[sel] ** 2 / (4 * D * t))
            exp[i] += (dens * mt.vol[sel]).sum()
    for i in range(1, len(obs)):
        assert tolerable(obs[i], exp[i], 0.08), (i, obs[i], exp[i])


def test_equilibrium_is_uniform_by_volume():
    mesh = geom.kuhn_cube(4, 2e-6)
    mt = geom.mesh_tables(mesh, geom.TmComp("c", mesh))
    fm = _diff_model(1e-12)
    o = bind.OracleSim(fm, mt, seed=8)
    o.counts[0, 0] = 38_400
    o.run(20.0, 0)
    c = o.counts[0]
    assert c.sum() == 38_400
    assert stats.chisquare(c).pvalue > 1e-3


def test_reaction_diffusion_conserves_exactly():
    w = W.ab_binding_cube(n=5, side=5e-6, n_a=4000, n_b=3000, seed=1, end_time=0.3)
    o = bind.OracleSim(w.fm, w.mt, seed=4)
    o.counts[:] = w.counts
    for t in (0.05, 0.17, 0.3):
        o.run(t, 1)
        tot = o.counts.sum(axis=1)
        assert tot[0] + tot[2] == 4000 and tot[1] + tot[2] == 3000
        assert o.counts.min() >= 0


def test_oracle_reset_keeps_the_philox_counter():
    """OracleSim.reset() clears state and step count but not the Philox period
    counter, as the engine's stg_reset does (the reference's RNG runs on
    through TetOpSplitP::reset, tetopsplit.cpp:1436-1503)."""
    w = W.ab_binding_cube(n=3, side=3e-6, n_a=500, n_b=500, seed=1, end_time=0.05)
    o = bind.OracleSim(w.fm, w.mt, seed=3)
    o.counts[:] = w.counts
    o.run(0.05, 0)
    first = o.counts.copy()
    n = o.nsteps
    o.reset()
    assert o.nsteps == 0 and o.time == 0.0 and o.rng_steps == n and o.counts.sum() == 0
    o.counts[:] = w.counts
    o.run(0.05, 0)
    assert not np.array_equal(o.counts, first)
    # a fresh oracle started at the same counter reproduces the second run
    o2 = bind.OracleSim(w.fm, w.mt, seed=3)
    o2.rng_steps = n
    o2.counts[:] = w.counts
    o2.run(0.05, 0)
    assert np.array_equal(o2.counts, o.counts)


def _clamp_decay_case(seed, t_end=3.0):
    """A (clamped at 100 per tet) -> A + B at k = 0.5 /s, B -> 0 at k2 = 2 /s, A
    and B diffusing, on a 2^3 x 6 cube (48 tets)."""
    mdl = M.Model()
    A, B = M.Spec("A", mdl), M.Spec("B", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Reac("make", vs, lhs=[A], rhs=[B], kcst=0.5)
    M.Reac("decay", vs, lhs=[B], rhs=[], kcst=2.0)
    M.Diff("dA", vs, A, 1e-12)
    M.Diff("dB", vs, B, 1e-12)
    mesh = geom.kuhn_cube(2, 2e-6)
    comp = geom.TmComp("c", mesh)
    comp.addVolsys("vsys")
    fm = M.flatten(mdl, ["vsys"])
    mt = geom.mesh_tables(mesh, comp)
    o = bind.OracleSim(fm, mt, seed=seed)
    o.counts[0] = 100
    o.set_tables(tet_clamp=np.ones(len(mt.vol), dtype=np.uint32))   # species 0 = A
    return o, fm, mt


def test_oracle_clamped_source_feeds_decay_to_its_steady_state():
    """A clamped species never changes (reactions consume it and diffusion moves
    it out, neither reduces it: reac.cpp:240-243, diff.cpp:476-577), so B is an
    immigration-death process with E[B(t)] = k N_A / k2 (1 - exp(-k2 t)) and
    Poisson fluctuations. Dropping the clamp (a plausible mistake) drains A
    and B falls far below."""
    t, niter = 3.0, 16
    mean = 0.5 * 4800 / 2.0 * (1 - np.exp(-2.0 * t))
    out = []
    for seed in range(niter):
        o, fm, mt = _clamp_decay_case(seed)
        o.run(t, 0)
        assert np.all(o.counts[0] == 100)
        out.append(o.counts[1].sum())
    z = (np.mean(out) - mean) / np.sqrt(mean / niter)
    assert abs(z) < 3.0, (np.mean(out), mean)
    o, fm, mt = _clamp_decay_case(99)
    o.set_tables()
    o.run(t, 0)
    assert o.counts[0].sum() < 4800 and o.counts[1].sum() < 0.5 * mean


def test_oracle_face_table_closes_and_opens_faces():
    """A per-face diffusion table replaces the constants: zero on every face of
    the x < side/2 tets and on their neighbours' faces toward them freezes that
    half (no reactions), while the other half keeps diffusing and its total is
    conserved; a table equal to the constants reproduces the plain run."""
    w = W.ab_binding_cube(n=4, side=4e-6, n_a=4000, n_b=0, seed=2, kf=0.0, kb=0.0)
    left = w.mesh.tet_bary[:, 0] < 2e-6
    S, N = 3, w.n_tets
    D = np.broadcast_to(w.fm.spec_dcst[:, None, None], (S, N, 4)).copy()
    plain = bind.OracleSim(w.fm, w.mt, seed=4)
    plain.counts[:] = w.counts
    same = bind.OracleSim(w.fm, w.mt, seed=4)
    same.counts[:] = w.counts
    same.set_tables(tet_fD=D)
    plain.run(0.2, 0)
    same.run(0.2, 0)
    assert np.array_equal(plain.counts, same.counts)
    D[:, left, :] = 0.0
    nb = w.mt.nbr.reshape(N, 4)
    for f in range(4):
        to_left = (nb[:, f] >= 0) & left[np.maximum(nb[:, f], 0)]
        D[:, to_left, f] = 0.0
    o = bind.OracleSim(w.fm, w.mt, seed=4)
    o.counts[:] = w.counts
    o.set_tables(tet_fD=D)
    o.run(0.2, 0)
    assert np.array_equal(o.counts[:, left], w.counts[:, left])
    assert o.counts[0, ~left].sum() == w.counts[0, ~left].sum()
    assert not np.array_equal(o.counts[:, ~left], w.counts[:, ~left])
