"""Synthetic reaction-diffusion workloads on Kuhn-split cubes.

Each builder returns a ``Workload``: the flattened model, the mesh tables of
one realization and its initial counts. The same workloads feed the GPU
engine, the CPU oracle, the tests and bench.py.

  ab_binding_cube          A + B <-> C on an 8^3 x 6 cube of side 10 um
  four_species_cube        4 species, 6 reactions on a 40^3 x 6 cube of side 40 um
  calcium_membrane_cube    Ca buffering with membrane pumps and influx, 100^3 x 6 cube, 20 um
  ab_binding_ensemble      1024 realizations of A + B <-> C on cubes n = 6..20
  two_comp_membrane        two compartments joined by a membrane with transport
                           reactions and a GHK current between them
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import geom, model as M

AVOGADRO = 6.02214179e23


@dataclass
class Workload:
    name: str
    fm: M.FlatModel
    mt: geom.MeshTables
    counts: np.ndarray                 # (S, n_tets) int32
    tri_counts: np.ndarray             # (SS, n_tris) int32
    end_time: float
    params: dict = field(default_factory=dict)
    mesh: geom.Tetmesh | None = None

    @property
    def n_tets(self) -> int:
        return len(self.mt.vol)


def _cube(n, side, with_patch, device=False):
    mesh = geom.kuhn_cube(n, side, device=device)
    comp = geom.TmComp("cyto", mesh)
    comp.addVolsys("vsys")
    tris = mesh.getSurfTris() if with_patch else None
    if with_patch:
        patch = geom.TmPatch("memb", mesh, tris, comp)
        patch.addSurfsys("ssys")
    return mesh, comp, tris


def place_uniform(n_mol: int, vol: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """Molecules placed independently, each tet chosen with probability V_t / V."""
    return rng.multinomial(n_mol, vol / vol.sum()).astype(np.int32)


def ab_binding_model(kf=1e6, kb=1.0, dA=1e-12, dB=1e-12, dC=5e-13):
    mdl = M.Model()
    A, B, C = M.Spec("A", mdl), M.Spec("B", mdl), M.Spec("C", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Reac("fwd", vs, lhs=[A, B], rhs=[C], kcst=kf)
    M.Reac("bwd", vs, lhs=[C], rhs=[A, B], kcst=kb)
    M.Diff("dA", vs, A, dA)
    M.Diff("dB", vs, B, dB)
    M.Diff("dC", vs, C, dC)
    return mdl


def ab_binding_cube(n=8, side=10e-6, n_a=10_000, n_b=10_000, seed=0, end_time=0.1,
                    kf=1e6, kb=1.0) -> Workload:
    mdl = ab_binding_model(kf, kb)
    mesh, comp, _ = _cube(n, side, False)
    fm = M.flatten(mdl, ["vsys"])
    mt = geom.mesh_tables(mesh, comp)
    rng = np.random.default_rng(seed)
    counts = np.zeros((3, len(mt.vol)), dtype=np.int32)
    counts[0] = place_uniform(n_a, mt.vol, rng)
    counts[1] = place_uniform(n_b, mt.vol, rng)
    return Workload("ab_binding_cube", fm, mt, counts, np.zeros((0, 0), np.int32), end_time,
                    dict(n=n, side=side, n_a=n_a, n_b=n_b, seed=seed, kf=kf, kb=kb), mesh)


def four_species_cube(n=40, side=40e-6, seed=0, end_time=1.0, mean_count=50, device=False) -> Workload:
    """A + B <-> C, C + B <-> D (kf 1e6 /M/s, kb 1 /s), 0 -> A at 10 /s per tet,
    D -> 0 at 0.5 /s; D log-uniform in [1e-13, 1e-11] m^2/s per species;
    initial counts Poisson(mean_count) per (tet, species)."""
    rng = np.random.default_rng(seed)
    dcst = 10.0 ** rng.uniform(-13, -11, size=4)
    dcst[int(np.argmax(dcst))] = 1e-11  # the fastest species sets dt at the stability bound
    mesh, comp, _ = _cube(n, side, False, device)
    vol_t = float(mesh.tet_vol[0])
    mdl = M.Model()
    A, B, Cs, D = (M.Spec(x, mdl) for x in "ABCD")
    vs = M.Volsys("vsys", mdl)
    M.Reac("bind1_f", vs, lhs=[A, B], rhs=[Cs], kcst=1e6)
    M.Reac("bind1_b", vs, lhs=[Cs], rhs=[A, B], kcst=1.0)
    M.Reac("bind2_f", vs, lhs=[Cs, B], rhs=[D], kcst=1e6)
    M.Reac("bind2_b", vs, lhs=[D], rhs=[Cs, B], kcst=1.0)
    M.Reac("prod", vs, lhs=[], rhs=[A], kcst=10.0 / (1e3 * vol_t * AVOGADRO))
    M.Reac("deg", vs, lhs=[D], rhs=[], kcst=0.5)
    for sp, d in zip((A, B, Cs, D), dcst):
        M.Diff("d" + sp.id, vs, sp, float(d))
    fm = M.flatten(mdl, ["vsys"])
    mt = geom.mesh_tables(mesh, comp)
    counts = rng.poisson(mean_count, size=(4, len(mt.vol))).astype(np.int32)
    return Workload("four_species_cube", fm, mt, counts, np.zeros((0, 0), np.int32), end_time,
                    dict(n=n, side=side, seed=seed, dcst=dcst.tolist()), mesh)


def calcium_membrane_cube(n=100, side=20e-6, seed=0, end_time=0.05, buf_um=100.0,
                          ca_um=0.1, pump_density=10.0, influx=1e5,
                          vdep_influx=False, ghk_influx=False, pump_dcst=0.0, device=False) -> Workload:
    """Ca + Buf <-> CaB (kf 1e8 /M/s, kb 10 /s), D_Ca 2e-10 m^2/s (Buf, CaB
    immobile); every boundary triangle is membrane. Pumps P (pump_density per
    um^2) run P + Ca_in -> P + Ca_out with k = kcat/Km = 50 /s / 1 uM; Ca_out
    is a surface species counting extruded ions. The x = 0 face carries one
    influx channel per triangle, Chan -> Chan + Ca_in, with k set so the face
    admits ``influx`` ions/s in total. With ``vdep_influx`` the channel is
    voltage gated: k(V) = k_max / (1 + exp(-(V + 20 mV) / 6 mV)) over -150..100 mV.
    With ``ghk_influx`` the channel instead passes a GHK current of Ca (valence
    2, virtual outer concentration 2 mM, 20 pS slope conductance at -65 mV).
    pump_dcst > 0 lets the pumps diffuse on the membrane (surface diffusion)."""
    rng = np.random.default_rng(seed)
    mesh, comp, tris = _cube(n, side, True, device)
    mdl = M.Model()
    Ca, Buf, CaB = M.Spec("Ca", mdl, valence=2), M.Spec("Buf", mdl), M.Spec("CaB", mdl)
    P, CaOut, Chan = M.Spec("Pump", mdl), M.Spec("CaOut", mdl), M.Spec("Chan", mdl)
    vs = M.Volsys("vsys", mdl)
    M.Reac("bind", vs, lhs=[Ca, Buf], rhs=[CaB], kcst=1e8)
    M.Reac("unbind", vs, lhs=[CaB], rhs=[Ca, Buf], kcst=10.0)
    M.Diff("dCa", vs, Ca, 2e-10)
    ss = M.Surfsys("ssys", mdl)
    M.SReac("pump", ss, ilhs=[Ca], slhs=[P], srhs=[P, CaOut], kcst=50.0 / 1e-6)
    if pump_dcst > 0:
        M.Diff("dPump", ss, P, pump_dcst)
        M.Diff("dCaOut", ss, CaOut, pump_dcst)
    face_tris = tris[np.all(np.abs(mesh.verts[mesh.tris[tris]][:, :, 0]) < 1e-3 * side / n, axis=1)]
    n_face = len(face_tris)
    k_in = influx / max(n_face, 1)
    if ghk_influx:
        cur = M.GHKcurr("influx", ss, Chan, Ca, computeflux=True, virtual_oconc=2e-3)
        cur.setPInfo(20e-12, -65e-3, 293.15, 2e-3, 1e-7)
    elif vdep_influx:
        M.VDepSReac("influx", ss, slhs=[Chan], srhs=[Chan], irhs=[Ca],
                    k=lambda v: k_in / (1.0 + np.exp(-(v + 20e-3) / 6e-3)))
    else:
        M.SReac("influx", ss, slhs=[Chan], srhs=[Chan], irhs=[Ca], kcst=k_in)
    fm = M.flatten(mdl, ["vsys"], ["ssys"])
    mt = geom.mesh_tables(mesh, comp, tris)
    vol_l = mt.vol * 1e3
    counts = np.zeros((len(fm.specs), len(mt.vol)), dtype=np.int32)
    counts[fm.spec_index("Buf")] = rng.poisson(buf_um * 1e-6 * AVOGADRO * vol_l)
    counts[fm.spec_index("Ca")] = rng.poisson(ca_um * 1e-6 * AVOGADRO * vol_l)
    tri_counts = np.zeros((len(fm.sspecs), len(tris)), dtype=np.int32)
    area_um2 = mt.tri_area * 1e12
    tri_counts[fm.sspec_index("Pump")] = place_uniform(int(round(pump_density * area_um2.sum())),
                                                       mt.tri_area, rng)
    is_face = np.isin(tris, face_tris)
    tri_counts[fm.sspec_index("Chan")] = is_face.astype(np.int32)
    return Workload("calcium_membrane_cube", fm, mt, counts, tri_counts, end_time,
                    dict(n=n, side=side, seed=seed, n_influx_tris=int(n_face)), mesh)


def ab_binding_ensemble(n_real=1024, n_min=6, n_max=20, side=10e-6, seed=0, end_time=0.1,
                        n_a=10_000, n_b=10_000, kf=1e6, kb=1.0):
    """Independent A + B <-> C realizations: cube resolution n ~ U{n_min..n_max},
    kf and kb log-uniform over +-1 decade, molecules placed by the realization's
    own seed. Returns (flat model, [mesh tables], [counts], reac_k (R, 2))."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=n_real)
    kfs = kf * 10.0 ** rng.uniform(-1, 1, size=n_real)
    kbs = kb * 10.0 ** rng.uniform(-1, 1, size=n_real)
    fm = M.flatten(ab_binding_model(kf, kb), ["vsys"])
    cache = {}
    meshes, counts = [], []
    for r in range(n_real):
        n = int(sizes[r])
        if n not in cache:
            mesh = geom.kuhn_cube(n, side)
            cache[n] = geom.mesh_tables(mesh, geom.TmComp("cyto", mesh))
        mt = cache[n]
        meshes.append(mt)
        c = np.zeros((3, len(mt.vol)), dtype=np.int32)
        rr = np.random.default_rng([seed, r])
        c[0] = place_uniform(n_a, mt.vol, rr)
        c[1] = place_uniform(n_b, mt.vol, rr)
        counts.append(c)
    reac_k = np.stack([kfs, kbs], axis=1)
    return fm, meshes, counts, reac_k, dict(sizes=sizes.tolist(), seed=seed, end_time=end_time)


def two_comp_cube(n=6, side=6e-6, dA=1e-12, dC_left=5e-13, dC_right=2e-12):
    """Kuhn cube cut at x = side/2 into compartments 'left' and 'right' with a
    diffusion boundary 'mid' on the cut. left: A + B -> C (1e6 /M/s), A and C
    diffuse (B is immobile there); right: C -> A + B (2 /s), A, B and C
    diffuse, and species R exists only on the right. Returns (model, mesh,
    {'left', 'right'} comps, boundary)."""
    mesh = geom.kuhn_cube(n, side)
    left = np.nonzero(mesh.tet_bary[:, 0] < side / 2)[0]
    right = np.nonzero(mesh.tet_bary[:, 0] >= side / 2)[0]
    cl = geom.TmComp("left", mesh, left)
    cr = geom.TmComp("right", mesh, right)
    cl.addVolsys("vl")
    cr.addVolsys("vr")
    tt = mesh.tri_tets
    both = (tt[:, 1] >= 0)
    side_of = np.zeros(mesh.countTets(), dtype=np.int64)
    side_of[right] = 1
    cut = np.nonzero(both & (side_of[tt[:, 0]] != side_of[np.maximum(tt[:, 1], 0)]))[0]
    db = geom.DiffBoundary("mid", mesh, cut)
    mdl = M.Model()
    A, B, C = M.Spec("A", mdl), M.Spec("B", mdl), M.Spec("C", mdl)
    vl, vr = M.Volsys("vl", mdl), M.Volsys("vr", mdl)
    M.Reac("bind", vl, lhs=[A, B], rhs=[C], kcst=1e6)
    M.Diff("dA", vl, A, dA)
    M.Diff("dC", vl, C, dC_left)
    M.Reac("split", vr, lhs=[C], rhs=[A, B], kcst=2.0)
    M.Diff("dA", vr, A, dA)
    M.Diff("dB", vr, B, dA / 2)
    M.Diff("dC", vr, C, dC_right)
    M.Diff("dR", vr, M.Spec("R", mdl), dA)
    return mdl, mesh, {"left": cl, "right": cr}, db


def two_comp_membrane(n=4, side=4e-6, k_in=2e9, k_out=1e9, n_a=4000, n_t=1, binding=True,
                      ghk=True, dA=1e-12, seed=0):
    """A Kuhn cube cut at x = side/2: 'inner' (x < side/2) and 'outer'
    compartments, both with A (and Ca) diffusing at dA; the triangles of the
    cut form patch 'memb' (inner 'inner', outer 'outer') with n_t transporters
    T per triangle: T + A(outer) -> T + A(inner) at k_in and T + A(inner) ->
    T + A(outer) at k_out (/M/s). With ``binding`` the inner compartment also
    holds A + B <-> C (1e6 /M/s, 1 /s; B and C immobile); with ``ghk`` every
    triangle has one open channel Ch passing a GHK calcium current (20 pS at
    -65 mV) between the two compartments. A and Ca start in the outer
    compartment. Returns (model, mesh, comps {'inner', 'outer'}, patch)."""
    mesh = geom.kuhn_cube(n, side)
    left = np.nonzero(mesh.tet_bary[:, 0] < side / 2)[0]
    right = np.nonzero(mesh.tet_bary[:, 0] >= side / 2)[0]
    ci = geom.TmComp("inner", mesh, left)
    co = geom.TmComp("outer", mesh, right)
    ci.addVolsys("vin")
    co.addVolsys("vout")
    tt = mesh.tri_tets
    side_of = np.zeros(mesh.countTets(), dtype=np.int64)
    side_of[right] = 1
    both = tt[:, 1] >= 0
    cut = np.nonzero(both & (side_of[tt[:, 0]] != side_of[np.maximum(tt[:, 1], 0)]))[0]
    patch = geom.TmPatch("memb", mesh, cut, ci, co)
    patch.addSurfsys("ssys")
    mdl = M.Model()
    A, B, C = M.Spec("A", mdl), M.Spec("B", mdl), M.Spec("C", mdl)
    Ca = M.Spec("Ca", mdl, valence=2)
    T, Ch = M.Spec("T", mdl), M.Spec("Ch", mdl)
    vin, vout = M.Volsys("vin", mdl), M.Volsys("vout", mdl)
    for vs in (vin, vout):
        M.Diff("dA", vs, A, dA)
        M.Diff("dCa", vs, Ca, dA)
    if binding:
        M.Reac("bind", vin, lhs=[A, B], rhs=[C], kcst=1e6)
        M.Reac("unbind", vin, lhs=[C], rhs=[A, B], kcst=1.0)
    ss = M.Surfsys("ssys", mdl)
    M.SReac("in", ss, slhs=[T], olhs=[A], srhs=[T], irhs=[A], kcst=k_in)
    M.SReac("out", ss, slhs=[T], ilhs=[A], srhs=[T], orhs=[A], kcst=k_out)
    if ghk:
        g = M.GHKcurr("gca", ss, Ch, Ca, computeflux=True)
        g.setPInfo(20e-12, -65e-3, 293.15, 2e-3, 1e-7)
    return mdl, mesh, {"inner": ci, "outer": co}, patch


def two_patch_membrane(n=4, side=4e-6, k_in=2e9, k_out=1e9, k_leak=3e9, dA=1e-11, dT2=2e-13):
    """The cut of two_comp_membrane split into two patches with different
    surface systems, both between 'inner' and 'outer': patch 'pa' (cut
    triangles with barycentre y < side/2, surface system 'ssa') moves A in and
    out through transporter T (T + A(outer) -> T + A(inner) at k_in, the
    reverse at k_out) and passes the GHK calcium current of channel Ch; patch
    'pb' (surface system 'ssb') leaks A outward through T2 (T2 + A(inner) ->
    T2 + A(outer) at k_leak), flips T2 <-> T3 (5 /s, 2 /s) and lets T2
    diffuse at dT2. The inner compartment holds A + B <-> C (1e6 /M/s, 1 /s).
    Returns (model, mesh, comps {'inner', 'outer'}, patches {'pa', 'pb'})."""
    mesh = geom.kuhn_cube(n, side)
    left = np.nonzero(mesh.tet_bary[:, 0] < side / 2)[0]
    right = np.nonzero(mesh.tet_bary[:, 0] >= side / 2)[0]
    ci = geom.TmComp("inner", mesh, left)
    co = geom.TmComp("outer", mesh, right)
    ci.addVolsys("vin")
    co.addVolsys("vout")
    tt = mesh.tri_tets
    side_of = np.zeros(mesh.countTets(), dtype=np.int64)
    side_of[right] = 1
    both = tt[:, 1] >= 0
    cut = np.nonzero(both & (side_of[tt[:, 0]] != side_of[np.maximum(tt[:, 1], 0)]))[0]
    low = mesh.tri_bary[cut, 1] < side / 2
    pa = geom.TmPatch("pa", mesh, cut[low], ci, co)
    pb = geom.TmPatch("pb", mesh, cut[~low], ci, co)
    pa.addSurfsys("ssa")
    pb.addSurfsys("ssb")
    mdl = M.Model()
    A, B, C = M.Spec("A", mdl), M.Spec("B", mdl), M.Spec("C", mdl)
    Ca = M.Spec("Ca", mdl, valence=2)
    T, T2, T3, Ch = M.Spec("T", mdl), M.Spec("T2", mdl), M.Spec("T3", mdl), M.Spec("Ch", mdl)
    vin, vout = M.Volsys("vin", mdl), M.Volsys("vout", mdl)
    for vs in (vin, vout):
        M.Diff("dA", vs, A, dA)
        M.Diff("dCa", vs, Ca, dA)
    M.Reac("bind", vin, lhs=[A, B], rhs=[C], kcst=1e6)
    M.Reac("unbind", vin, lhs=[C], rhs=[A, B], kcst=1.0)
    ssa, ssb = M.Surfsys("ssa", mdl), M.Surfsys("ssb", mdl)
    M.SReac("in", ssa, slhs=[T], olhs=[A], srhs=[T], irhs=[A], kcst=k_in)
    M.SReac("out", ssa, slhs=[T], ilhs=[A], srhs=[T], orhs=[A], kcst=k_out)
    g = M.GHKcurr("gca", ssa, Ch, Ca, computeflux=True)
    g.setPInfo(20e-12, -65e-3, 293.15, 2e-3, 1e-7)
    M.SReac("leak", ssb, slhs=[T2], ilhs=[A], srhs=[T2], orhs=[A], kcst=k_leak)
    M.SReac("flip", ssb, slhs=[T2], srhs=[T3], kcst=5.0)
    M.SReac("flop", ssb, slhs=[T3], srhs=[T2], kcst=2.0)
    M.Diff("dT2", ssb, T2, dT2)
    return mdl, mesh, {"inner": ci, "outer": co}, {"pa": pa, "pb": pb}


def sdiff_boundary_patches(n=4, side=4e-6, dX=1e-12, split=False):
    """The top face (z = side) of a Kuhn cube split at x = side/2 into patch
    'patchA' (x > side/2, surface system 'ssys1': A -> A at 1e5 /s and X
    diffusing by rule 'diffX1') and patch 'patchB' (surface system 'ssys2': X
    diffusing by rule 'diffX2'), both at dX, joined by the surface diffusion
    boundary 'sdiffb' on the bars they share: the set-up of the reference's
    test_sdiff_bugfix.py on a cube. split=False: one compartment 'cyto' holds
    both patches; split=True: compartments 'left' and 'right' (tets by
    barycentre x) hold patchB and patchA, so the boundary joins triangles of
    two compartments. The compartments hold a volume species V diffusing at
    1e-12. Returns (model, mesh, comps, patches, sdiffb)."""
    mesh = geom.kuhn_cube(n, side)
    if split:
        right = np.nonzero(mesh.tet_bary[:, 0] >= side / 2)[0]
        left = np.nonzero(mesh.tet_bary[:, 0] < side / 2)[0]
        comps = {"left": geom.TmComp("left", mesh, left), "right": geom.TmComp("right", mesh, right)}
        ca, cb = comps["right"], comps["left"]
    else:
        comps = {"cyto": geom.TmComp("cyto", mesh)}
        ca = cb = comps["cyto"]
    for c in comps.values():
        c.addVolsys("vsys")
    top = np.nonzero(np.all(np.isclose(mesh.verts[mesh.tris][:, :, 2], side), axis=1))[0]
    xa = mesh.tri_bary[top, 0] > side / 2
    pa = geom.TmPatch("patchA", mesh, top[xa], ca)
    pb = geom.TmPatch("patchB", mesh, top[~xa], cb)
    pa.addSurfsys("ssys1")
    pb.addSurfsys("ssys2")
    bars = sorted(set(mesh.tri_bars[pa.tris].ravel().tolist()) & set(mesh.tri_bars[pb.tris].ravel().tolist()))
    sdb = geom.SDiffBoundary("sdiffb", mesh, bars, [pa, pb])
    mdl = M.Model()
    A, X, V = M.Spec("A", mdl), M.Spec("X", mdl), M.Spec("V", mdl)
    vsys = M.Volsys("vsys", mdl)
    M.Diff("dV", vsys, V, 1e-12)
    s1, s2 = M.Surfsys("ssys1", mdl), M.Surfsys("ssys2", mdl)
    M.SReac("sreac", s1, slhs=[A], srhs=[A], kcst=1e5)
    M.Diff("diffX1", s1, X, dX)
    M.Diff("diffX2", s2, X, dX)
    return mdl, mesh, comps, {"patchA": pa, "patchB": pb}, sdb
