"""Statistics of batched ensembles on the GPU.

The parity tests show the engine equal to the oracle bit for bit on the same
streams; these show what an ensemble run in one batched launch samples, the way
the reference's validation suite judges its solvers (ensemble means and
variances against closed forms at 99% confidence, and a two-sample KS test
against an independent ensemble):

- first-order decay A -> 0 with diffusion: every molecule survives to t with
  probability exp(-k t) whatever its path, so the surviving count of each
  realization is exactly Binomial(N0, exp(-k t));
- A + B <-> C with diffusion: the GPU ensemble and an ensemble of the CPU
  oracle on other Philox keys must sample the same law of the final C total.
"""
import numpy as np
import pytest
from scipy import stats

from oracle import bind
from steps_amd import geom, model as M, workloads as W
from steps_amd.engine import Engine

pytestmark = pytest.mark.gpu

Z99 = 2.5758293035489004  # two-sided 99% normal quantile


def _cube_tables(n, side, volsys="vsys"):
    mesh = geom.kuhn_cube(n, side)
    comp = geom.TmComp("cyto", mesh)
    comp.addVolsys(volsys)
    return geom.mesh_tables(mesh, comp)


def test_decay_ensemble_is_binomial():
    """1,024 realizations of A -> 0 (k = 10 /s) with A diffusing on a 162-tet cube,
    200 molecules each, run to t = 0.1 s in one batched engine: surviving counts
    ~ Binomial(200, e^-1). Mean and variance within their 99% intervals, and the
    counts pass a chi-square goodness-of-fit test against the binomial pmf."""
    mdl = M.Model()
    A = M.Spec("A", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Reac("decay", vs, lhs=[A], rhs=[], kcst=10.0)
    M.Diff("dA", vs, A, 1e-12)
    fm = M.flatten(mdl, ["vsys"])
    mt = _cube_tables(3, 3e-6)
    R, N0, t_end, k = 1024, 200, 0.1, 10.0
    rng = np.random.default_rng(11)
    print("case decay ensemble 1024 x 162 tets", flush=True)
    eng = Engine(fm, [mt] * R, [4242] * R)
    for r in range(R):
        eng.set_tet_counts(0, W.place_uniform(N0, mt.vol, rng), real=r)
    eng.run(t_end)
    n = np.array([eng.sum_tet_counts(0, r) for r in range(R)], dtype=float)
    p = np.exp(-k * t_end)
    mu, var = N0 * p, N0 * p * (1 - p)
    assert abs(n.mean() - mu) < Z99 * np.sqrt(var / R), (n.mean(), mu)
    s2 = n.var(ddof=1)
    lo, hi = stats.chi2.ppf([0.005, 0.995], R - 1) * var / (R - 1)
    assert lo < s2 < hi, (s2, var, lo, hi)
    # chi-square goodness of fit over bins holding >= 5 expected counts each
    edges = np.arange(0, N0 + 2)
    expected = R * stats.binom.pmf(edges[:-1], N0, p)
    observed = np.bincount(n.astype(int), minlength=N0 + 1)[: N0 + 1].astype(float)
    obs_b, exp_b, acc_o, acc_e = [], [], 0.0, 0.0
    for o, e in zip(observed, expected):
        acc_o += o
        acc_e += e
        if acc_e >= 5.0:
            obs_b.append(acc_o)
            exp_b.append(acc_e)
            acc_o = acc_e = 0.0
    obs_b[-1] += acc_o
    exp_b[-1] += acc_e
    chi = float(np.sum((np.array(obs_b) - np.array(exp_b)) ** 2 / np.array(exp_b)))
    assert chi < stats.chi2.ppf(0.99, len(obs_b) - 1), (chi, len(obs_b))


def test_binding_ensemble_matches_independent_oracle_ensemble():
    """A + B <-> C on a 162-tet cube: 400 GPU realizations in one batched engine
    and 400 CPU-oracle runs on other Philox keys. The final C totals of the two
    ensembles pass a two-sample KS test at 1%, and their means agree within the
    99% interval of the difference."""
    fm = M.flatten(W.ab_binding_model(kf=1e9, kb=5.0), ["vsys"])
    mt = _cube_tables(3, 3e-6)
    R, t_end = 400, 0.05
    rng = np.random.default_rng(12)
    init = []
    for _ in range(R):
        c = np.zeros((3, len(mt.vol)), dtype=np.int32)
        c[0] = W.place_uniform(150, mt.vol, rng)
        c[1] = W.place_uniform(120, mt.vol, rng)
        init.append(c)
    print("case binding ensemble GPU 400 vs oracle 400", flush=True)
    eng = Engine(fm, [mt] * R, [777] * R)
    for r in range(R):
        for s in range(3):
            eng.set_tet_counts(s, init[r][s], real=r)
    eng.run(t_end)
    gpu = np.array([eng.sum_tet_counts(2, r) for r in range(R)], dtype=float)
    cpu = np.empty(R)
    for r in range(R):
        o = bind.OracleSim(fm, mt, seed=90210, stream=r)
        o.counts[:] = init[r]
        o.run(t_end, 0)
        cpu[r] = o.counts[2].sum()
        assert o.counts[0].sum() + cpu[r] == 150 and o.counts[1].sum() + cpu[r] == 120
    assert gpu.std() > 0 and cpu.std() > 0
    assert stats.ks_2samp(gpu, cpu).pvalue > 0.01
    se = np.sqrt(gpu.var(ddof=1) / R + cpu.var(ddof=1) / R)
    assert abs(gpu.mean() - cpu.mean()) < Z99 * se, (gpu.mean(), cpu.mean(), se)
