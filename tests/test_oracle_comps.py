"""Pins the oracle's compartments and diffusion boundaries (CPU).

Compartments are laid out as stg_create_comps lays them out (one Philox
stream and local tet numbering per compartment); a face between compartments
is closed unless it lies on a diffusion boundary, whose per-direction,
per-species constant opens it (TetOpSplitP::_setDiffBoundarySpecDiffusionActive
and _setDiffBoundarySpecDcst, src/steps/mpi/tetopsplit/tetopsplit.cpp:2745-2931;
Diff::setDcst/setDirectionDcst, diff.cpp:340-460).
"""
import numpy as np
import pytest
from scipy import stats

from oracle import bind
from steps_amd import geom, model as M, workloads as W


def _two_comp(no_reactions=False):
    mdl, mesh, comps, db = W.two_comp_cube(n=4, side=4e-6)
    ct = M.flatten_comps(mdl, [["vl"], ["vr"]])
    mts, dirs = geom.comps_mesh_tables(mesh, [comps["left"], comps["right"]], {}, [db])
    k = np.zeros_like(ct.comp_k) if no_reactions else ct.comp_k
    return ct, mts, dirs, k


def _oracle(ct, mts, k, seed, n_bdir):
    return bind.OracleSim.comps(ct.fm, mts, seed, k, ct.comp_dcst, n_bdir=n_bdir)


def _comp_totals(o, sp):
    off = o.comp_off
    return [int(o.counts[sp, off[c]:off[c + 1]].sum()) for c in range(len(off) - 1)]


def test_closed_boundary_keeps_each_compartment():
    ct, mts, dirs, k = _two_comp(no_reactions=True)
    o = _oracle(ct, mts, k, 3, len(dirs))
    rng = np.random.default_rng(0)
    o.counts[:] = rng.poisson(30, o.counts.shape)
    o.counts[1, :o.comp_off[1]] = 0          # B is not diffusing on the left
    before = [_comp_totals(o, s) for s in range(3)]
    dt = o.update_period()
    o.run_periods([dt] * 40)
    assert [_comp_totals(o, s) for s in range(3)] == before
    assert o.moved.value > 0


def test_open_boundary_equilibrates_by_volume():
    # A opened across the boundary (same D both sides): all molecules start
    # on the left and spread to the volume share (both halves equal)
    ct, mts, dirs, k = _two_comp(no_reactions=True)
    o = _oracle(ct, mts, k, 5, len(dirs))
    a = ct.fm.specs.index("A")
    o.bnd_dcst[:, a] = ct.comp_dcst[0, a]
    o.counts[a, :o.comp_off[1]] = 200
    total = int(o.counts[a].sum())
    dt = o.update_period()
    o.run_periods([dt] * 4000)          # about 4 mixing times L^2/D of the cube
    left, right = _comp_totals(o, a)
    assert left + right == total
    # binomial(total, 1/2) at equilibrium
    assert abs(left - total / 2) < 5 * np.sqrt(total / 4)


def test_one_way_boundary_only_fills_the_destination():
    # a directional constant from left to right only (setDiffBoundarySpecDcst
    # with direction_comp = right): the right total never decreases
    ct, mts, dirs, k = _two_comp(no_reactions=True)
    o = _oracle(ct, mts, k, 7, len(dirs))
    a = ct.fm.specs.index("A")
    d_lr = next(d for d, (_, src, dst) in enumerate(dirs) if src == 0 and dst == 1)
    o.bnd_dcst[d_lr, a] = 2e-12
    o.counts[a] = 50
    dt = o.update_period()
    prev = _comp_totals(o, a)[1]
    for _ in range(60):
        o.run_periods([dt])
        now = _comp_totals(o, a)[1]
        assert now >= prev
        prev = now
    assert prev > 50 * (o.comp_off[2] - o.comp_off[1])


def test_reactions_stay_in_their_compartment():
    # C -> A + B runs only on the right: with the boundary closed the left
    # total of C is exact and the right one decays at 2 /s
    ct, mts, dirs, k = _two_comp()
    o = _oracle(ct, mts, k, 11, len(dirs))
    a, b, c = (ct.fm.specs.index(x) for x in "ABC")
    o.counts[c] = 40
    o.counts[a] = 0
    o.counts[b] = 0
    left0 = _comp_totals(o, c)[0]
    nr = o.comp_off[2] - o.comp_off[1]
    o.run(0.2, 0)
    # bind on the left needs A and B, which never reach it through a closed boundary
    assert _comp_totals(o, c)[0] == left0
    assert _comp_totals(o, a)[0] == 0 and _comp_totals(o, b)[0] == 0
    right = _comp_totals(o, c)[1]
    expect = 40 * nr * np.exp(-2.0 * 0.2)
    assert abs(right - expect) < 5 * np.sqrt(expect)


def test_boundary_leaving_count_is_binomial():
    # isolated pairs: tet 2i in the left compartment, its neighbour across a
    # boundary face in the right one; species 1 crosses at a quarter of the
    # stability rate, so each period ~Binomial(n, 1/4) molecules cross
    n_pairs, g = 4000, 1e12
    mts = []
    for c in range(2):
        nbr = -np.ones((n_pairs, 4), dtype=np.int32)
        nbr[:, 0] = np.arange(n_pairs)
        nf = -np.ones((n_pairs, 4), dtype=np.int32)
        nf[:, 0] = 0
        fg = np.zeros((n_pairs, 4))
        fg[:, 0] = g
        bd = -np.ones((n_pairs, 4), dtype=np.int32)
        bd[:, 0] = c
        mts.append(geom.MeshTables(vol=np.full(n_pairs, 1e-19), nbr=nbr, nbr_face=nf, face_g=fg,
                                   tri_tet=np.zeros(0, np.int32), tri_area=np.zeros(0),
                                   nbr_comp=np.full((n_pairs, 4), 1 - c, dtype=np.int32),
                                   bdir=bd))
    mdl = M.Model()
    X, Y = M.Spec("X", mdl), M.Spec("Y", mdl)
    vs = M.Volsys("v", mdl)
    M.Diff("dX", vs, X, 1e-12)
    M.Diff("dY", vs, Y, 0.25e-12)
    ct = M.flatten_comps(mdl, [["v"], ["v"]])
    o = bind.OracleSim.comps(ct.fm, mts, 13, np.zeros((2, 0)), ct.comp_dcst, n_bdir=2)
    o.bnd_dcst[:, :] = ct.comp_dcst[0]
    assert abs(o.update_period() - 1.0) < 1e-12     # 1 / (D_X g)
    y = ct.fm.specs.index("Y")
    o.counts[y, :n_pairs] = 40
    o.run_periods([1.0])
    crossed = 40 - o.counts[y, :n_pairs]
    assert np.array_equal(o.counts[y, n_pairs:], crossed)
    k = np.arange(41)
    exp = stats.binom.pmf(k, 40, 0.25) * n_pairs
    obs = np.bincount(crossed, minlength=41)
    keep = exp >= 5
    e = np.append(exp[keep], exp[~keep].sum())
    ob = np.append(obs[keep], obs[~keep].sum())
    assert stats.chisquare(ob, e * ob.sum() / e.sum()).pvalue > 1e-3


# ---------------------------------------------------------- surface diffusion

def _sdiff_model(D_fast, D_slow=None):
    mdl = M.Model()
    X = M.Spec("X", mdl)
    ss = M.Surfsys("ssys", mdl)
    M.Diff("sdX", ss, X, D_fast)
    if D_slow is not None:
        M.Diff("sdY", ss, M.Spec("Y", mdl), D_slow)
    M.Volsys("vsys", mdl)
    return mdl


def test_surface_leaving_count_is_binomial_and_conserved():
    # isolated triangle pairs joined across edge 0 (SDiff with one open edge);
    # Y diffuses at a quarter of the stability rate: Binomial(n, 1/4) crosses
    n_pairs, g = 3000, 2e12
    n_tri = 2 * n_pairs
    tri_nbr = -np.ones((n_tri, 3), dtype=np.int32)
    tri_nbr[0::2, 0] = np.arange(1, n_tri, 2)
    tri_nbr[1::2, 0] = np.arange(0, n_tri, 2)
    tri_edge = np.where(tri_nbr >= 0, 0, -1).astype(np.int32)
    tg = np.where(tri_nbr >= 0, g, 0.0)
    z4 = -np.ones((n_tri, 4), dtype=np.int32)
    mt = geom.MeshTables(vol=np.full(n_tri, 1e-19), nbr=z4, nbr_face=z4.copy(),
                         face_g=np.zeros((n_tri, 4)), tri_tet=np.arange(n_tri, dtype=np.int32),
                         tri_area=np.full(n_tri, 1e-13), tri_nbr=tri_nbr, tri_nbr_edge=tri_edge,
                         tri_edge_g=tg)
    fm = M.flatten(_sdiff_model(1e-12, 0.25e-12), ["vsys"], ["ssys"])
    o = bind.OracleSim(fm, mt, seed=19)
    assert abs(o.update_period() - 1.0 / (1e-12 * g)) < 1e-12 * o.update_period()
    y = fm.sspec_index("Y")
    o.tri_counts[y, 0::2] = 40
    o.run_periods([o.update_period()])
    crossed = 40 - o.tri_counts[y, 0::2]
    assert np.array_equal(o.tri_counts[y, 1::2], crossed)
    k = np.arange(41)
    exp = stats.binom.pmf(k, 40, 0.25) * n_pairs
    obs = np.bincount(crossed, minlength=41)
    keep = exp >= 5
    e = np.append(exp[keep], exp[~keep].sum())
    ob = np.append(obs[keep], obs[~keep].sum())
    assert stats.chisquare(ob, e * ob.sum() / e.sum()).pvalue > 1e-3


def test_surface_point_source_spreads_as_2d_diffusion():
    # molecules released at the centre of the flat z = 0 face of a cube: the
    # mean squared displacement on the plane grows as 4 D t (2-D diffusion),
    # before the spread reaches the face's edges
    n, side, D = 24, 24e-6, 1e-12
    mesh = geom.kuhn_cube(n, side)
    comp = geom.TmComp("c", mesh)
    surf = mesh.getSurfTris()
    z = mesh.verts[mesh.tris[surf]][:, :, 2]
    face = surf[np.all(np.abs(z) < 1e-12, axis=1)]
    mt = geom.mesh_tables(mesh, comp, face)
    fm = M.flatten(_sdiff_model(D), ["vsys"], ["ssys"])
    o = bind.OracleSim(fm, mt, seed=23)
    bary = mesh.tri_bary[face][:, :2]
    centre = np.array([side / 2, side / 2])
    src = int(np.argmin(np.linalg.norm(bary - centre, axis=1)))
    o.tri_counts[0, src] = 20000
    t = 8.0
    dt = o.update_period()
    o.run_periods([dt] * int(round(t / dt)))
    c = o.tri_counts[0, :len(face)]
    assert c.sum() == 20000
    r2 = (np.linalg.norm(bary - bary[src], axis=1) ** 2 * c).sum() / c.sum()
    assert abs(r2 / (4 * D * t) - 1) < 0.1, r2 / (4 * D * t)
