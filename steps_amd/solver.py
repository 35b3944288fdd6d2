"""Solver classes with the interface of the reference's operator-split solvers.

``TetOpSplit`` mirrors ``_py_TetOpSplitP`` (pysteps/cysteps_mpi.pyx:74-1546,
C++ TetOpSplitP in src/steps/mpi/tetopsplit/tetopsplit.hpp) and
``DistTetOpSplit`` mirrors ``_py_DistTetOpSplitP`` (pysteps/cysteps_dist.pyx:
1574-2932). Both run on the HIP engine behind include/stg.h with the whole mesh
resident on one GPU; the host partition arguments are accepted and ignored.
Method names, argument meaning and error types follow the reference: invalid
arguments raise ``ArgErr`` (a ValueError), unsupported features ``NotImplErr``.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .engine import Engine
from .geom import Tetmesh, comps_mesh_tables, mesh_tables
from .model import Model, flatten, flatten_comps

EF_NONE, EF_DEFAULT, EF_DV_BDSYS, EF_DV_PETSC = 0, 1, 2, 3  # src/steps/solver/api.hpp:61-66
DEFAULT_MEMB_POT = -65.0e-3                                  # src/steps/util/common.hpp:101
AVOGADRO = 6.02214179e23
# meshes of at least this many tets run the period kernel compiled for their model
SPECIALIZE_MIN_TETS = 200_000


def _argerr(msg):
    return N.ArgErr(N.STG_ERR_ARG, msg)


def _notimpl(msg):
    return N.NotImplErr(N.STG_ERR_NOTIMPL, msg)


def distribute_counts(n: float, vol: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """Volume-weighted fair distribution of n molecules over tets, the scheme of
    util::distribute_quantity (src/steps/util/distribute.hpp): stochastic
    rounding of n, floor of each tet's share, and the remainder placed on
    distinct tets sampled with probability proportional to the fractional
    parts (Pareto pi-ps sampling)."""
    if n < 0:
        raise _argerr("negative number of molecules")
    x = np.floor(n) + (1.0 if rng.random() < n - np.floor(n) else 0.0)
    share = x * vol / vol.sum()
    base = np.floor(share)
    frac = share - base
    out = base.astype(np.int64)
    rem = int(round(x - base.sum()))
    if rem > 0:
        pi = frac * rem / frac.sum()
        pi = np.clip(pi, 1e-300, 1 - 1e-12)
        u = rng.random(len(vol))
        q = u * (1 - pi) / ((1 - u) * pi)
        out[np.argsort(q, kind="stable")[:rem]] += 1
    if np.any(out > np.iinfo(np.int32).max):
        raise _argerr("count exceeds the int32 range of a tet")
    return out.astype(np.int32)


class TetOpSplit:
    """Operator-split stochastic reaction-diffusion on one GPU (STEPS 'TetOpSplit')."""

    _mode = N.MODE_TETOPSPLIT
    _name = "TetOpSplitP"

    def __init__(self, model: Model, geom: Tetmesh, rng, calcMembPot: int = 0, tet_hosts=(),
                 tri_hosts=None, wm_hosts=()):
        if calcMembPot not in (EF_NONE, EF_DEFAULT, EF_DV_BDSYS, EF_DV_PETSC):
            raise _argerr(f"unknown membrane potential solver {calcMembPot}")
        self._setup(model, geom, rng)
        if calcMembPot != EF_NONE:
            # every EField solver choice maps to the GPU preconditioned CG
            self._setup_efield(geom)

    # --- construction ---
    def _setup(self, model, geom, rng):
        if not isinstance(geom, Tetmesh):
            raise _argerr("geometry must be a Tetmesh")
        self.model, self.geom, self.rng = model, geom, rng
        self.comps = list(geom.comps.values())
        if not self.comps:
            raise _argerr("the geometry has no compartment")
        self._ci = {c.name: i for i, c in enumerate(self.comps)}
        patches = [p for p in geom.patches.values() if p.surfsys]
        if len(patches) > 1:
            raise _notimpl("at most one patch with surface systems is supported")
        self.patch = patches[0] if patches else None
        if self.patch is not None and self.patch.icomp.name not in self._ci:
            raise _argerr("the patch's inner compartment is not in the geometry")
        self._preal = self._ci[self.patch.icomp.name] if self.patch is not None else 0
        ptris = self.patch.tris if self.patch is not None else None
        ssys = self.patch.surfsys if self.patch else ()
        nc = len(self.comps)
        self.diffbs = list(geom.diffbs.values())
        if nc == 1:
            self.fm = flatten(model, self.comps[0].volsys, ssys)
            self.mts = [mesh_tables(geom, self.comps[0], ptris)]
            self._bdirs = []
            self._comp_reacs = [{r: i for i, r in enumerate(self.fm.reacs)}]
            self._comp_diffs = [dict(self.fm.diffs)]
            self._comp_specs = [set(self.fm.specs)]
            self._ck = self.fm.reac_k.copy()[None, :]
            self._cd = self.fm.spec_dcst.copy()[None, :]
        else:
            ct = flatten_comps(model, [c.volsys for c in self.comps], ssys)
            self.fm = ct.fm
            self.mts, self._bdirs = comps_mesh_tables(
                geom, self.comps, {self._preal: ptris} if ptris is not None else {}, self.diffbs)
            self._comp_reacs, self._comp_diffs = ct.comp_reacs, ct.comp_diffs
            self._comp_specs = ct.comp_specs
            self._ck, self._cd = ct.comp_k.copy(), ct.comp_dcst.copy()
        if self.patch is not None:
            self._comp_specs[self._preal] |= {self.fm.specs[j] for j in
                                              np.nonzero(np.abs(self.fm.sreac_lhs_i).sum(0)
                                                         + np.abs(self.fm.sreac_upd_i).sum(0))[0]}
        self.comp = self.comps[0]
        self.mt = self.mts[self._preal]
        # global tet -> (compartment, local index)
        self._tet_real = -np.ones(geom.countTets(), dtype=np.int64)
        self._tet_local = -np.ones(geom.countTets(), dtype=np.int64)
        for ci, c in enumerate(self.comps):
            self._tet_real[c.tets] = ci
            self._tet_local[c.tets] = np.arange(len(c.tets))
        self._tri_local = -np.ones(geom.countTris(), dtype=np.int64)
        if self.patch is not None:
            self._tri_local[self.patch.tris] = np.arange(len(self.patch.tris))
        self._seed = int(getattr(rng, "seed", 0))
        if nc == 1:
            self.eng = Engine(self.fm, [self.mt], [self._seed], mode=self._mode)
        else:
            self.eng = Engine(self.fm, self.mts, [self._seed] * nc, streams=np.arange(nc),
                              reac_k=self._ck, spec_dcst=self._cd, mode=self._mode, comps=True)
        # Large meshes get the period kernel compiled for this model (identical
        # results, about a second to compile); without run-time compilation
        # the generic kernel stays.
        if geom.countTets() >= SPECIALIZE_MIN_TETS:
            try:
                self.eng.set_specialize(True)
            except N.StgError:
                pass
        # diffusion boundaries: open directions and directional constants
        self._db_active = set()        # (boundary, species, source comp)
        self._db_dcst = {}             # (boundary, species, source comp) -> constant
        self._host_rng = np.random.default_rng([self._seed, 0x5EED])
        self.ef = None

    def _setup_efield(self, geom):
        """Membrane potential on the membrane's conduction volume
        (TetOpSplitP::_setupEField, tetopsplit.cpp:1080-1170)."""
        from .efield import EField

        membs = list(geom.membs.values())
        if len(membs) != 1:
            raise _argerr("the membrane potential solve needs exactly one Memb (got %d)"
                          % len(membs))
        memb = self.memb = membs[0]
        ptris = self.patch.tris if self.patch is not None else np.zeros(0, np.int64)
        if len(ptris) and set(ptris.tolist()) != set(memb.tris.tolist()):
            raise _notimpl("the membrane must consist of the patch that holds the surface systems")
        # EField triangle k is the solver's patch triangle k
        tris = ptris if len(ptris) else memb.tris
        tets = geom.tets[memb.tets]
        vids, local = np.unique(tets, return_inverse=True)
        self._ef_vert_local = -np.ones(geom.countVertices(), dtype=np.int64)
        self._ef_vert_local[vids] = np.arange(len(vids))
        tri_v = self._ef_vert_local[geom.tris[tris]]
        if np.any(tri_v < 0):
            raise _argerr("a membrane triangle is not on the conduction volume")
        self._ef_tris = tris
        self._ef_tri_local = -np.ones(geom.countTris(), dtype=np.int64)
        self._ef_tri_local[tris] = np.arange(len(tris))
        self.ef = EField(geom.verts[vids], local.reshape(-1, 4), tri_v)
        self._efdt = 1.0e-5                                  # tetopsplit.hpp:938
        self._vert_iclamp = np.zeros(len(vids))
        self._tri_iclamp = np.zeros(len(tris))
        self._vert_clamped = np.zeros(len(vids), dtype=bool)
        if len(self.fm.ohmic):
            if len(self.fm.ohmic) > 8:
                raise _notimpl("at most 8 ohmic currents are supported")
            self.ef.set_ohmic(self.fm.ohmic_chan, self.fm.ohmic_g, self.fm.ohmic_erev)
        if self.patch is not None:
            self.ef.attach(self.eng, self._preal)

    # --- identity ---
    def getSolverName(self):
        return self._name

    def getSolverDesc(self):
        return "Operator-split spatial SSA and multinomial diffusion on one MI355X"

    def getSolverAuthors(self):
        return "steps_amd"

    def getSolverEmail(self):
        return ""

    # --- time stepping ---
    def reset(self):
        """TetOpSplitP::reset (tetopsplit.cpp:1436-1503): counts, time, step
        count and extents to zero and, with an EField, the membrane potential
        back to DEFAULT_MEMB_POT. The random streams keep running, as the
        reference's RNG does, so reset-and-run loops give independent runs;
        re-initialising the RNG with a new seed before reset() re-keys every
        compartment and restarts the streams from that seed."""
        seed = int(getattr(self.rng, "seed", self._seed))
        if seed != self._seed:
            self._seed = seed
            for r in range(self.eng.n_real):
                self.eng.set_seed(seed, r)
            self._host_rng = np.random.default_rng([self._seed, 0x5EED])
        self.eng.reset()
        if self.ef is not None:
            self.ef.set_potential(DEFAULT_MEMB_POT)
            if self.patch is not None:
                self.ef.attach(self.eng, self._preal)   # channel integrals restart at t = 0

    def run(self, endtime: float):
        if self.ef is not None:
            self.ef.run(self.eng, endtime, self._efdt)   # _runWithEField
        else:
            self.eng.run(endtime)

    def advance(self, adv: float):
        if adv < 0.0:
            raise _argerr("Time to advance cannot be negative")
        self.run(self.getTime() + adv)

    def step(self):
        raise _notimpl("step() is not implemented for operator splitting")  # tetopsplit.cpp:1969

    def getTime(self) -> float:
        return self.eng.time()

    def setTime(self, t: float):
        self.eng.set_time(t)

    def getNSteps(self) -> int:
        return self.eng.nsteps()

    getNIteration = getNSteps

    def getUpdPeriod(self) -> float:
        return self.eng.update_period()

    def getReacExtent(self, local: bool = False) -> int:
        return sum(self.eng.extents(r)[0] for r in range(self.eng.n_real))

    def getDiffExtent(self, local: bool = False) -> int:
        return sum(self.eng.extents(r)[1] for r in range(self.eng.n_real))

    def setDiffApplyThreshold(self, threshold: int):
        self.eng.set_diff_apply_threshold(threshold)

    # --- lookups ---
    def _spec(self, s: str) -> int:
        try:
            return self.fm.spec_index(s)
        except KeyError as e:
            raise _argerr(str(e)) from None

    def _sspec(self, s: str) -> int:
        try:
            return self.fm.sspec_index(s)
        except KeyError as e:
            raise _argerr(str(e)) from None

    def _check_comp(self, c: str) -> int:
        if c not in self._ci:
            raise _argerr(f"unknown compartment {c!r}")
        return self._ci[c]

    def _comp_spec(self, ci: int, s: str) -> int:
        sp = self._spec(s)
        if s not in self._comp_specs[ci]:
            raise _argerr(f"species {s!r} is undefined in compartment {self.comps[ci].name!r}")
        return sp

    def _check_patch(self, p: str):
        if self.patch is None or p != self.patch.name:
            raise _argerr(f"unknown patch {p!r}")

    def _tet(self, t: int):
        """(compartment, local index) of a mesh tetrahedron."""
        if t < 0 or t >= len(self._tet_local) or self._tet_local[t] < 0:
            raise _argerr(f"tetrahedron {t} is not in a compartment")
        return int(self._tet_real[t]), int(self._tet_local[t])

    def _tri(self, t: int) -> int:
        if t < 0 or t >= len(self._tri_local) or self._tri_local[t] < 0:
            raise _argerr(f"triangle {t} is not in the patch")
        return int(self._tri_local[t])

    # --- compartments ---
    def getCompVol(self, c: str) -> float:
        return float(self.mts[self._check_comp(c)].vol.sum())

    def getCompSpecCount(self, c: str, s: str) -> float:
        ci = self._check_comp(c)
        return float(self.eng.sum_tet_counts(self._comp_spec(ci, s), real=ci))

    def setCompSpecCount(self, c: str, s: str, n: float):
        ci = self._check_comp(c)
        self.eng.set_tet_counts(self._comp_spec(ci, s),
                                distribute_counts(n, self.mts[ci].vol, self._host_rng), real=ci)

    def getCompSpecAmount(self, c: str, s: str) -> float:
        return self.getCompSpecCount(c, s) / AVOGADRO

    def setCompSpecAmount(self, c: str, s: str, a: float):
        self.setCompSpecCount(c, s, a * AVOGADRO)

    def getCompSpecConc(self, c: str, s: str) -> float:
        return self.getCompSpecCount(c, s) / (1e3 * self.getCompVol(c) * AVOGADRO)

    def setCompSpecConc(self, c: str, s: str, conc: float):
        self.setCompSpecCount(c, s, conc * 1e3 * self.getCompVol(c) * AVOGADRO)

    def _reac(self, ci: int, r: str) -> int:
        if r not in self._comp_reacs[ci]:
            raise _argerr(f"unknown reaction {r!r} in compartment {self.comps[ci].name!r}")
        return self._comp_reacs[ci][r]

    def getCompReacK(self, c: str, r: str) -> float:
        ci = self._check_comp(c)
        return float(self._ck[ci, self._reac(ci, r)])

    def setCompReacK(self, c: str, r: str, k: float):
        ci = self._check_comp(c)
        j = self._reac(ci, r)
        self.eng.set_reac_k(j, k, real=ci)
        self._ck[ci, j] = k

    def _diff(self, ci: int, d: str) -> int:
        if d not in self._comp_diffs[ci]:
            raise _argerr(f"unknown diffusion rule {d!r} in compartment {self.comps[ci].name!r}")
        return self._comp_diffs[ci][d]

    def getCompDiffD(self, c: str, d: str) -> float:
        ci = self._check_comp(c)
        return float(self._cd[ci, self._diff(ci, d)])

    def setCompDiffD(self, c: str, d: str, dcst: float, direction_comp=None):
        ci = self._check_comp(c)
        s = self._diff(ci, d)
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        self.eng.set_dcst(s, dcst, real=ci)
        self._cd[ci, s] = dcst
        # Diff::setDcst also resets the rule's boundary directions to the new
        # constant (diff.cpp:340-390)
        for key in [k for k in self._db_dcst if k[1] == s and k[2] == ci]:
            del self._db_dcst[key]
        self._push_boundaries(s)

    # --- diffusion boundaries (tetopsplit.cpp:2745-2931) ---
    def _db(self, db: str) -> int:
        for i, b in enumerate(self.diffbs):
            if b.id == db:
                return i
        raise _argerr(f"unknown diffusion boundary {db!r}")

    def _db_spec(self, bi: int, s: str) -> int:
        sp = self._spec(s)
        b = self.diffbs[bi]
        for c in (b.compA, b.compB):
            if s not in self._comp_specs[self._ci[c.name]]:
                raise _argerr("Species undefined in compartments connected by diffusion boundary.")
        return sp

    def _push_boundaries(self, sp: int):
        """Boundary constants of species sp from the activation state: a
        direction from compartment c uses c's diffusion constant for the
        species, or the directional constant set for it; 0 when inactive or
        when c has no diffusion rule for it (no Diff kproc there)."""
        for d, (bi, src, _dst) in enumerate(self._bdirs):
            has_rule = sp in self._comp_diffs[src].values()
            on = (bi, sp, src) in self._db_active and has_rule
            val = self._db_dcst.get((bi, sp, src), self._cd[src, sp]) if on else 0.0
            self.eng.set_boundary_dcst(d, sp, val)

    def setDiffBoundarySpecDiffusionActive(self, db: str, s: str, act: bool):
        bi = self._db(db)
        sp = self._db_spec(bi, s)
        b = self.diffbs[bi]
        for c in (b.compA, b.compB):
            key = (bi, sp, self._ci[c.name])
            if act:
                self._db_active.add(key)
            else:
                self._db_active.discard(key)
            self._db_dcst.pop(key, None)    # setDiffBndActive -> setDcst clears overrides
        self._push_boundaries(sp)

    def getDiffBoundarySpecDiffusionActive(self, db: str, s: str) -> bool:
        """True when the species crosses the boundary in both directions."""
        bi = self._db(db)
        sp = self._db_spec(bi, s)
        b = self.diffbs[bi]
        return all((bi, sp, self._ci[c.name]) in self._db_active for c in (b.compA, b.compB))

    def setDiffBoundarySpecDcst(self, db: str, s: str, dcst: float, direction_comp=None):
        """Activate the boundary for s with constant dcst from both sides, or
        only into direction_comp (tetopsplit.cpp:2860-2917)."""
        bi = self._db(db)
        sp = self._db_spec(bi, s)
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        b = self.diffbs[bi]
        dest = None
        if direction_comp is not None:
            name = direction_comp if isinstance(direction_comp, str) else direction_comp.name
            if name not in (b.compA.name, b.compB.name):
                raise _argerr("direction compartment is not on the diffusion boundary")
            dest = self._ci[name]
        for c in (b.compA, b.compB):
            src = self._ci[c.name]
            if src != dest:
                self._db_active.add((bi, sp, src))
                self._db_dcst[(bi, sp, src)] = float(dcst)
        self._push_boundaries(sp)

    # --- tetrahedra ---
    def getTetVol(self, t: int) -> float:
        ci, i = self._tet(t)
        return float(self.mts[ci].vol[i])

    def getTetSpecCount(self, t: int, s: str) -> float:
        ci, i = self._tet(t)
        return float(self.eng.batch_tet_counts(self._comp_spec(ci, s), [i], real=ci)[0])

    def setTetSpecCount(self, t: int, s: str, n: float):
        if n < 0:
            raise _argerr("negative number of molecules")
        ci, i = self._tet(t)
        sp = self._comp_spec(ci, s)
        c = self.eng.get_tet_counts(sp, real=ci)
        c[i] = int(np.floor(n) + (1 if self._host_rng.random() < n - np.floor(n) else 0))
        self.eng.set_tet_counts(sp, c, real=ci)

    def getTetSpecConc(self, t: int, s: str) -> float:
        return self.getTetSpecCount(t, s) / (1e3 * self.getTetVol(t) * AVOGADRO)

    def setTetSpecConc(self, t: int, s: str, conc: float):
        self.setTetSpecCount(t, s, conc * 1e3 * self.getTetVol(t) * AVOGADRO)

    def getBatchTetSpecCounts(self, tets, s: str):
        c = np.zeros(len(tets))
        self.getBatchTetSpecCountsNP(np.asarray(tets, dtype=np.int64), s, c)
        return c.tolist()

    def getBatchTetSpecCountsNP(self, index_array, s: str, counts):
        idx = np.asarray(index_array, dtype=np.int64)
        if len(counts) != len(idx):
            raise _argerr("counts and index arrays differ in length")
        if len(idx) and (np.any(idx < 0) or np.any(idx >= len(self._tet_local))
                         or np.any(self._tet_local[idx] < 0)):
            raise _argerr("tetrahedron index out of range")
        sp = self._spec(s)
        real, loc = self._tet_real[idx], self._tet_local[idx]
        for ci in np.unique(real):
            m = real == ci
            counts[m] = self.eng.batch_tet_counts(sp, loc[m], real=int(ci))

    def getBatchTetSpecConcsNP(self, index_array, s: str, concs):
        self.getBatchTetSpecCountsNP(index_array, s, concs)
        idx = np.asarray(index_array, dtype=np.int64)
        vol = np.array([self.mts[r].vol[i] for r, i in zip(self._tet_real[idx], self._tet_local[idx])])
        concs[:] = concs / (1e3 * vol * AVOGADRO)

    def sumBatchTetCountsNP(self, tet_array, s: str) -> float:
        c = np.zeros(len(tet_array))
        self.getBatchTetSpecCountsNP(tet_array, s, c)
        return float(c.sum())

    # --- regions of interest (tetopsplit.cpp getROI*/setROI*) ---
    def _roi(self, roi: str, want: int):
        from .geom import ELEM_TET
        if roi not in self.geom.rois:
            raise _argerr(f"ROI {roi!r} does not exist")
        typ, idx = self.geom.rois[roi]
        if typ != want:
            raise _argerr(f"ROI {roi!r} holds {'tetrahedra' if typ == ELEM_TET else 'other elements'}, "
                          f"not {'tetrahedra' if want == ELEM_TET else 'triangles'}")
        return idx

    def _roi_tets(self, roi: str):
        from .geom import ELEM_TET
        idx = self._roi(roi, ELEM_TET)
        if np.any(self._tet_local[idx] < 0):
            raise _argerr(f"ROI {roi!r} holds a tetrahedron outside every compartment")
        return idx

    def getROIVol(self, roi: str) -> float:
        return float(self.geom.tet_vol[self._roi_tets(roi)].sum())

    def getROIArea(self, roi: str) -> float:
        from .geom import ELEM_TRI
        return float(self.geom.tri_area[self._roi(roi, ELEM_TRI)].sum())

    def getROITetSpecCounts(self, roi: str, s: str):
        idx = self._roi_tets(roi)
        c = np.zeros(len(idx))
        self.getBatchTetSpecCountsNP(idx, s, c)
        return c.tolist()

    def getROITetSpecCountsNP(self, roi: str, s: str, counts):
        self.getBatchTetSpecCountsNP(self._roi_tets(roi), s, counts)

    def getROITriSpecCounts(self, roi: str, s: str):
        from .geom import ELEM_TRI
        idx = self._roi(roi, ELEM_TRI)
        c = np.zeros(len(idx))
        self.getBatchTriSpecCountsNP(idx, s, c)
        return c.tolist()

    def getROITriSpecCountsNP(self, roi: str, s: str, counts):
        from .geom import ELEM_TRI
        self.getBatchTriSpecCountsNP(self._roi(roi, ELEM_TRI), s, counts)

    def getROISpecCount(self, roi: str, s: str) -> float:
        """Total over the ROI, reduced on the device per compartment."""
        from .geom import ELEM_TET, ELEM_TRI
        typ = self.geom.rois[roi][0] if roi in self.geom.rois else None
        if typ == ELEM_TRI:
            return float(sum(self.getROITriSpecCounts(roi, s)))
        idx = self._roi_tets(roi)
        sp = self._spec(s)
        real, loc = self._tet_real[idx], self._tet_local[idx]
        return float(sum(self.eng.sum_batch_tet_counts(sp, loc[real == ci], real=int(ci))
                         for ci in np.unique(real)))

    def setROISpecCount(self, roi: str, s: str, n: float):
        """Distribute n molecules over the ROI's tets by volume (setROISpecCount)."""
        idx = self._roi_tets(roi)
        sp = self._spec(s)
        real, loc = self._tet_real[idx], self._tet_local[idx]
        vol = self.geom.tet_vol[idx]
        add = distribute_counts(n, vol, self._host_rng)
        for ci in np.unique(real):
            m = real == ci
            c = self.eng.get_tet_counts(sp, real=int(ci))
            c[loc[m]] = add[m]
            self.eng.set_tet_counts(sp, c, real=int(ci))

    def getROISpecAmount(self, roi: str, s: str) -> float:
        return self.getROISpecCount(roi, s) / AVOGADRO

    def setROISpecAmount(self, roi: str, s: str, a: float):
        self.setROISpecCount(roi, s, a * AVOGADRO)

    def getROISpecConc(self, roi: str, s: str) -> float:
        return self.getROISpecCount(roi, s) / (1e3 * self.getROIVol(roi) * AVOGADRO)

    def setROISpecConc(self, roi: str, s: str, conc: float):
        self.setROISpecCount(roi, s, conc * 1e3 * self.getROIVol(roi) * AVOGADRO)

    # --- patches and triangles ---
    def getPatchSpecCount(self, p: str, s: str) -> float:
        self._check_patch(p)
        return float(self.eng.sum_tri_counts(self._sspec(s), real=self._preal))

    def setPatchSpecCount(self, p: str, s: str, n: float):
        self._check_patch(p)
        self.eng.set_tri_counts(self._sspec(s),
                                distribute_counts(n, self.mt.tri_area, self._host_rng), real=self._preal)

    def getPatchSReacK(self, p: str, r: str) -> float:
        self._check_patch(p)
        return float(self.fm.sreac_k[self._sreac(r)])

    def setPatchSReacK(self, p: str, r: str, k: float):
        self._check_patch(p)
        j = self._sreac(r)
        self.eng.set_sreac_k(j, k, real=self._preal)
        self.fm.sreac_k[j] = k

    def _sdiff(self, d: str) -> int:
        if d not in self.fm.sdiffs:
            raise _argerr(f"unknown surface diffusion rule {d!r}")
        return self.fm.sdiffs[d]

    def getPatchSDiffD(self, p: str, d: str) -> float:
        self._check_patch(p)
        return float(self.fm.sspec_dcst[self._sdiff(d)])

    def setPatchSDiffD(self, p: str, d: str, dcst: float, direction_patch=None):
        """Surface diffusion constant of a rule on the patch (setPatchSDiffD;
        directional constants need surface diffusion boundaries, which this
        solver does not model)."""
        self._check_patch(p)
        if direction_patch is not None:
            raise _notimpl("surface diffusion boundaries are not supported")
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        q = self._sdiff(d)
        self.eng.set_sdcst(q, dcst, real=self._preal)
        self.fm.sspec_dcst[q] = dcst

    def _sreac(self, r):
        try:
            return self.fm.sreacs.index(r)
        except ValueError:
            raise _argerr(f"unknown surface reaction {r!r}") from None

    def getTriSpecCount(self, tri: int, s: str) -> float:
        return float(self.eng.get_tri_counts(self._sspec(s), real=self._preal)[self._tri(tri)])

    def setTriSpecCount(self, tri: int, s: str, n: float):
        if n < 0:
            raise _argerr("negative number of molecules")
        sp = self._sspec(s)
        c = self.eng.get_tri_counts(sp, real=self._preal)
        c[self._tri(tri)] = int(round(n))
        self.eng.set_tri_counts(sp, c, real=self._preal)

    def getBatchTriSpecCounts(self, tris, s: str):
        c = self.eng.get_tri_counts(self._sspec(s), real=self._preal)
        return [float(c[self._tri(int(t))]) for t in tris]

    def sumBatchTriCountsNP(self, tri_array, s: str) -> float:
        c = self.eng.get_tri_counts(self._sspec(s), real=self._preal)
        return float(c[[self._tri(int(t)) for t in tri_array]].sum())

    def getBatchTriSpecCountsNP(self, index_array, s: str, counts):
        c = self.eng.get_tri_counts(self._sspec(s), real=self._preal)
        counts[:] = c[[self._tri(int(t)) for t in index_array]]

    # --- membrane potential (TetOpSplitP EField methods, tetopsplit.cpp:4751-5000) ---
    def _need_ef(self):
        if self.ef is None:
            raise _argerr("Method not available: EField calculation not included in simulation.")

    def _vert(self, v: int) -> int:
        self._need_ef()
        if not 0 <= v < len(self._ef_vert_local) or self._ef_vert_local[v] < 0:
            raise _argerr(f"Vertex index {v} not assigned to the conduction volume.")
        return int(self._ef_vert_local[v])

    def _eftri(self, t: int) -> int:
        self._need_ef()
        if not 0 <= t < len(self._ef_tri_local) or self._ef_tri_local[t] < 0:
            raise _argerr(f"Triangle index {t} not assigned to a membrane.")
        return int(self._ef_tri_local[t])

    def _memb(self, m: str):
        self._need_ef()
        if m != self.memb.id:
            raise _argerr(f"unknown membrane {m!r}")

    def setEfieldDT(self, efdt: float):
        if not efdt > 0.0:
            raise _argerr("EField dt must be positive")
        self._need_ef()
        self._efdt = float(efdt)

    def getEfieldDT(self) -> float:
        self._need_ef()
        return self._efdt

    def setTemp(self, t: float):
        if t < 0.0:
            raise _argerr("temperature must be >= 0 K")
        self.eng.set_temp(t)

    def getTemp(self) -> float:
        return self.eng.temp()

    def setMembPotential(self, m: str, v: float):
        self._memb(m)
        self.ef.set_potential(v)

    def setMembCapac(self, m: str, cm: float):
        self._memb(m)
        if cm < 0.0:
            raise _argerr("Capacitance must be non-negative")
        self.ef.set_capac(cm)

    def setMembVolRes(self, m: str, ro: float):
        self._memb(m)
        if ro < 0.0:
            raise _argerr("Resistivity must be non-negative")
        self.ef.set_volres(ro)

    def setMembRes(self, m: str, ro: float, vrev: float):
        self._memb(m)
        if ro <= 0.0:
            raise _argerr("Resistivity must be greater than zero")
        self.ef.set_memb_res(ro, vrev)

    def getVertV(self, v: int) -> float:
        return float(self.ef.verts_v()[self._vert(v)])

    def setVertV(self, v: int, val: float):
        i = self._vert(v)
        vv = self.ef.verts_v()
        vv[i] = val
        self.ef.set_verts_v(vv)

    def getVertVClamped(self, v: int) -> bool:
        return bool(self._vert_clamped[self._vert(v)])

    def setVertVClamped(self, v: int, cl: bool):
        i = self._vert(v)
        self.ef.set_vert_clamped(i, cl)
        self._vert_clamped[i] = bool(cl)

    def getVertIClamp(self, v: int) -> float:
        return float(self._vert_iclamp[self._vert(v)])

    def setVertIClamp(self, v: int, i: float):
        k = self._vert(v)
        self.ef.set_vert_iclamp(k, i)
        self._vert_iclamp[k] = i

    def getTetV(self, t: int) -> float:
        self._need_ef()
        vv = self.ef.verts_v()
        return float(vv[[self._vert(int(x)) for x in self.geom.tets[t]]].mean())

    def setTetV(self, t: int, v: float):
        self._need_ef()
        vv = self.ef.verts_v()
        vv[[self._vert(int(x)) for x in self.geom.tets[t]]] = v
        self.ef.set_verts_v(vv)

    def getTetVClamped(self, t: int) -> bool:
        self._need_ef()
        return bool(all(self._vert_clamped[self._vert(int(x))] for x in self.geom.tets[t]))

    def setTetVClamped(self, t: int, cl: bool):
        self._need_ef()
        for x in self.geom.tets[t]:
            self.setVertVClamped(int(x), cl)

    def getTriV(self, tri: int) -> float:
        """Potential (V) of a membrane triangle, the mean of its vertices (EField::getTriV,
        efield.cpp:290-302). Without an EField, the value voltage-dependent reactions read."""
        if self.ef is None:
            return float(self.eng.get_tri_v(self._preal)[self._tri(tri)])
        return float(self.ef.tris_v()[self._eftri(tri)])

    def setTriV(self, tri: int, v: float):
        """Set the potential of a membrane triangle: with an EField, of its three
        vertices (EField::setTriV, efield.cpp:306-317; TetOpSplitP::_setTriV,
        tetopsplit.cpp:4871-4893). Without one, only the value voltage-dependent
        reactions read (no counterpart in the reference, which raises)."""
        if self.ef is None:
            vv = self.eng.get_tri_v(self._preal)
            vv[self._tri(tri)] = v
            self.eng.set_tri_v(vv, self._preal)
            return
        self._eftri(tri)
        vv = self.ef.verts_v()
        vv[[self._vert(int(x)) for x in self.geom.tris[tri]]] = v
        self.ef.set_verts_v(vv)

    def getTriVClamped(self, tri: int) -> bool:
        self._eftri(tri)
        return bool(all(self._vert_clamped[self._vert(int(x))] for x in self.geom.tris[tri]))

    def setTriVClamped(self, tri: int, cl: bool):
        self._eftri(tri)
        for x in self.geom.tris[tri]:
            self.setVertVClamped(int(x), cl)

    def getTriIClamp(self, tri: int) -> float:
        return float(self._tri_iclamp[self._eftri(tri)])

    def setTriIClamp(self, tri: int, i: float):
        k = self._eftri(tri)
        self.ef.set_tri_iclamp(k, i)
        self._tri_iclamp[k] = i

    def getTriOhmicI(self, tri: int, oc: str) -> float:
        """Present ohmic current (A) of one current on a triangle: open channels
        times g times (V - erev) (Tri::getOhmicI, tri.cpp:709-718)."""
        k = self._eftri(tri)
        try:
            j = self.fm.ohmic.index(oc)
        except ValueError:
            raise _argerr(f"unknown ohmic current {oc!r}") from None
        n = self.eng.get_tri_counts(int(self.fm.ohmic_chan[j]), real=self._preal)[self._tri(tri)]
        v = self.ef.tris_v()[k]
        return float(n * self.fm.ohmic_g[j] * (v - self.fm.ohmic_erev[j]))

    def getBatchTriVsNP(self, index_array, vs):
        if self.ef is None:
            vv = self.eng.get_tri_v(self._preal)
            vs[:] = vv[[self._tri(int(t)) for t in index_array]]
        else:
            vv = self.ef.tris_v()
            vs[:] = vv[[self._eftri(int(t)) for t in index_array]]

    def getBatchTetVsNP(self, index_array, vs):
        self._need_ef()
        vv = self.ef.verts_v()
        for k, t in enumerate(index_array):
            vs[k] = vv[[self._vert(int(x)) for x in self.geom.tets[int(t)]]].mean()

    def getBatchTriOhmicIsNP(self, index_array, oc: str, cur):
        for k, t in enumerate(index_array):
            cur[k] = self.getTriOhmicI(int(t), oc)

    def sumBatchTriOhmicIsNP(self, tri_array, oc: str) -> float:
        return float(sum(self.getTriOhmicI(int(t), oc) for t in tri_array))

    def getBatchVertVsNP(self, index_array, vs):
        self._need_ef()
        vv = self.ef.verts_v()
        vs[:] = vv[[self._vert(int(v)) for v in index_array]]

    def setBatchTriVsNP(self, index_array, vs):
        for t, v in zip(index_array, vs):
            self.setTriV(int(t), float(v))

    # --- checkpointing (own binary format: stg_checkpoint and stg_efield_checkpoint in include/stg.h) ---
    _CP_MAGIC = b"STGCKPT1"

    def checkpoint(self, file_name: str):
        """TetOpSplitP::checkpoint (tetopsplit.cpp:217-320): the engine state and,
        with an EField, the EField state (pEField->checkpoint, tetopsplit.cpp:309)."""
        eng = self.eng.checkpoint()
        ef = self.ef.checkpoint() if self.ef is not None else b""
        with open(file_name, "wb") as f:
            f.write(self._CP_MAGIC)
            f.write(np.array([len(eng), len(ef)], dtype=np.uint64).tobytes())
            f.write(eng)
            f.write(ef)

    def restore(self, file_name: str):
        """TetOpSplitP::restore (tetopsplit.cpp:353-453), EField included (:446)."""
        with open(file_name, "rb") as f:
            data = f.read()
        if data[:8] != self._CP_MAGIC or len(data) < 24:
            raise _argerr(f"{file_name!r} is not a checkpoint of this solver")
        n_eng, n_ef = (int(x) for x in np.frombuffer(data[8:24], dtype=np.uint64))
        if len(data) < 24 + n_eng + n_ef:
            raise _argerr("checkpoint file is truncated")
        if (n_ef > 0) != (self.ef is not None):
            raise _argerr("checkpoint and solver differ in whether they hold an EField")
        self.eng.restore(data[24:24 + n_eng])
        if self.ef is not None:
            self.ef.restore(data[24 + n_eng:24 + n_eng + n_ef])


class DistTetOpSplit(TetOpSplit):
    """Distributed-mesh operator splitting (STEPS 'DistTetOpSplit'), whole mesh on one GPU.
    The per-tet SSA always uses independent groups; the SSA method and search
    method arguments are accepted for compatibility (cysteps_dist.pyx:1581-1633)."""

    _mode = N.MODE_DISTTETOPSPLIT
    _name = "DistTetOpSplitP"
    UNIFORM, MULTINOMIAL = 0, 1

    def __init__(self, model: Model, mesh: Tetmesh, rng, SSAMethod=0, searchMethod=0,
                 indepKProcs: bool = False, isEfield: bool = True):
        self._setup(model, mesh, rng)
        if isEfield and mesh.membs:
            self._setup_efield(mesh)

    def setCompSpecCount(self, c: str, s: str, n: float, distribution: int = 0):
        if distribution == self.MULTINOMIAL:
            ci = self._check_comp(c)
            vol = self.mts[ci].vol
            cnt = self._host_rng.multinomial(int(round(n)), vol / vol.sum())
            self.eng.set_tet_counts(self._comp_spec(ci, s), cnt.astype(np.int32), real=ci)
        else:
            super().setCompSpecCount(c, s, n)


_SOLVERS = {"TetOpSplit": TetOpSplit, "DistTetOpSplit": DistTetOpSplit}


def get_solver(name: str):
    """Solver class for a reference solver name, as MPI._getSolver does for
    _solverNameMapping (pysteps/steps/API_2/sim.py:1502-1505, 1564-1579)."""
    try:
        return _SOLVERS[name]
    except KeyError:
        raise _argerr(f"solver {name!r} is not provided by this engine") from None
