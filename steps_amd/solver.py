"""Solver classes with the interface of the reference's operator-split solvers.

``TetOpSplit`` mirrors ``_py_TetOpSplitP`` (pysteps/cysteps_mpi.pyx:74-1546,
C++ TetOpSplitP in src/steps/mpi/tetopsplit/tetopsplit.hpp) together with the
``_py_API`` / ``_py_TetAPI`` methods it inherits (pysteps/cysteps_solver.pyx);
``DistTetOpSplit`` mirrors ``_py_DistTetOpSplitP`` (pysteps/cysteps_dist.pyx:
1574-2932). Both run on the HIP engine behind include/stg.h with the whole mesh
resident on one GPU; the host partition arguments are accepted and ignored.
Method names, argument meaning and error types follow the reference: invalid
arguments raise ``ArgErr`` (a ValueError), unsupported features ``NotImplErr``;
every reference method this engine does not provide exists and raises
``NotImplErr`` (REFERENCE_METHODS below).

Per-element constants, activation flags and clamps follow the reference's
per-kproc state (steps_amd/elements.py) and reach the engine as per-element
tables; until an element-level setter is used the engine runs on
compartment-wide constants.
"""
from __future__ import annotations

import time

import numpy as np

from . import _native as N
from .elements import TriElements, VolElements
from .engine import Engine
from .geom import ELEM_TET, ELEM_TRI, Tetmesh, comps_mesh_tables, mesh_tables
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


def _falling_h(counts: np.ndarray, lhs: np.ndarray) -> np.ndarray:
    """Combinatorial factor per element: the product over reactants of
    n (n-1) ... (n - lhs + 1), 0 when a reactant is short (Reac::rate,
    src/steps/mpi/tetopsplit/reac.cpp; SReac::rate, sreac.cpp:305-430).
    counts: [species][elements]."""
    h = np.ones(counts.shape[1])
    for s in np.nonzero(lhs)[0]:
        n = counts[s].astype(np.float64)
        for i in range(int(lhs[s])):
            h = h * (n - i)
        h = np.where(counts[s] < lhs[s], 0.0, h)
    return h


def _round_count(n: float, rng) -> int:
    """Stochastic rounding of a molecule count (Tet::setCount callers)."""
    return int(np.floor(n) + (1 if rng.random() < n - np.floor(n) else 0))


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

    # ------------------------------------------------------------ construction
    def _setup(self, model, geom, rng):
        """TetOpSplitP::_setup (tetopsplit.cpp:800-1075): compartments,
        patches (each with its inner and optional outer compartment),
        diffusion boundaries."""
        if not isinstance(geom, Tetmesh):
            raise _argerr("geometry must be a Tetmesh")
        self.model, self.geom, self.rng = model, geom, rng
        self.comps = list(geom.comps.values())
        if not self.comps:
            raise _argerr("the geometry has no compartment")
        self._ci = {c.name: i for i, c in enumerate(self.comps)}
        self.patches = [p for p in geom.patches.values() if p.surfsys]
        self._pi = {p.name: i for i, p in enumerate(self.patches)}
        for p in self.patches:
            if p.icomp.name not in self._ci or (p.ocomp is not None and p.ocomp.name not in self._ci):
                raise _argerr(f"a compartment of patch {p.name!r} is not in the geometry")
        nc = len(self.comps)
        self.diffbs = list(geom.diffbs.values())
        self.sdiffbs = list(geom.sdiffbs.values())
        self._sdbi = {b.id: i for i, b in enumerate(self.sdiffbs)}
        for b in self.sdiffbs:
            if b.patchA.name not in self._pi or b.patchB.name not in self._pi:
                raise _argerr(f"a patch of surface diffusion boundary {b.id!r} has no surface system")
        self.patch = self.patches[0] if len(self.patches) == 1 else None
        simple = nc == 1 and len(self.patches) <= 1 and all(p.ocomp is None for p in self.patches) \
            and not self.sdiffbs
        if simple:
            ssys = self.patches[0].surfsys if self.patches else ()
            ptris = self.patches[0].tris if self.patches else None
            self.fm = flatten(model, self.comps[0].volsys, ssys)
            self.mts = [mesh_tables(geom, self.comps[0], ptris)]
            self._bdirs = []
            self._comp_reacs = [{r: i for i, r in enumerate(self.fm.reacs)}]
            self._comp_diffs = [dict(self.fm.diffs)]
            self._comp_specs = [set(self.fm.specs)]
            self._ck = self.fm.reac_k.copy()[None, :]
            self._cd = self.fm.spec_dcst.copy()[None, :]
        else:
            pl = [(p.surfsys, self._ci[p.icomp.name], None if p.ocomp is None else self._ci[p.ocomp.name])
                  for p in self.patches]
            ct = flatten_comps(model, [c.volsys for c in self.comps], pl)
            self.fm = ct.fm
            self.mts, self._bdirs = comps_mesh_tables(
                geom, self.comps, [(p.tris, pl[i][1], pl[i][2]) for i, p in enumerate(self.patches)],
                self.diffbs, self.sdiffbs)
            self._comp_reacs, self._comp_diffs = ct.comp_reacs, ct.comp_diffs
            self._comp_specs = ct.comp_specs
            self._ck, self._cd = ct.comp_k.copy(), ct.comp_dcst.copy()
        self._simple = simple
        fm = self.fm
        RS, NG = len(fm.sreacs), len(fm.ghk_chan)
        # a patch's surface reactions and GHK currents define their volume
        # species in the inner (and outer) compartment (Patchdef::setup_indices)
        for pi, p in enumerate(self.patches):
            inner, outer = set(), set()
            for j in fm.patch_sreacs[pi]:
                inner |= {fm.specs[x] for x in np.nonzero(fm.sreac_lhs_i[j] | fm.sreac_upd_i[j])[0]}
                outer |= {fm.specs[x] for x in np.nonzero(fm.sreac_lhs_o[j] | fm.sreac_upd_o[j])[0]}
            for g in fm.patch_ghk[pi]:
                inner.add(fm.specs[fm.ghk_ion[g]])
                if fm.ghk_voconc[g] < 0:
                    outer.add(fm.specs[fm.ghk_ion[g]])
            self._comp_specs[self._ci[p.icomp.name]] |= inner
            if p.ocomp is not None:
                self._comp_specs[self._ci[p.ocomp.name]] |= outer
        self.comp = self.comps[0]
        # global tet -> (compartment, local index)
        self._tet_real = -np.ones(geom.countTets(), dtype=np.int64)
        self._tet_local = -np.ones(geom.countTets(), dtype=np.int64)
        for ci, c in enumerate(self.comps):
            self._tet_real[c.tets] = ci
            self._tet_local[c.tets] = np.arange(len(c.tets))
        # global triangle -> (compartment holding it, index there, patch); a
        # compartment's triangles are those of its patches in patch order
        self._tri_real = -np.ones(geom.countTris(), dtype=np.int64)
        self._tri_local = -np.ones(geom.countTris(), dtype=np.int64)
        self._tri_patch = -np.ones(geom.countTris(), dtype=np.int64)
        fill = np.zeros(nc, dtype=np.int64)
        for pi, p in enumerate(self.patches):
            r = self._ci[p.icomp.name]
            self._tri_real[p.tris] = r
            self._tri_local[p.tris] = fill[r] + np.arange(len(p.tris))
            self._tri_patch[p.tris] = pi
            fill[r] += len(p.tris)
        self._rtri_patch = [self._tri_patch[self._real_tris(r)] for r in range(nc)]
        self.mt = self.mts[self._ci[self.patch.icomp.name]] if self.patch is not None else self.mts[0]
        self._ck0, self._cd0 = self._ck.copy(), self._cd.copy()
        self._patch_defaults()
        self._seed = int(getattr(rng, "seed", 0))
        if simple:
            self.eng = Engine(fm, [self.mts[0]], [self._seed], mode=self._mode)
        else:
            sk = np.zeros((nc, RS))
            for r in range(nc):
                pats = np.unique(self._rtri_patch[r])
                if len(pats) == 1:
                    sk[r] = self._psk[pats[0]]
            self.eng = Engine(fm, self.mts, [self._seed] * nc, streams=np.arange(nc),
                              reac_k=self._ck, sreac_k=sk, spec_dcst=self._cd, mode=self._mode,
                              comps=True)
        # Large meshes get the period kernel compiled for this model (identical
        # results, about a second to compile); without run-time compilation
        # the generic kernel stays.
        if geom.countTets() >= SPECIALIZE_MIN_TETS:
            try:
                self.eng.set_specialize(True)
            except N.StgError:
                pass
        # compartment-level switches (Reac / Diff active flags of every tet)
        self._cract = np.ones((nc, len(fm.reacs)), dtype=bool)
        self._cdact = np.ones((nc, len(fm.specs)), dtype=bool)
        # diffusion boundaries: open directions and directional constants
        self._db_active = set()        # (boundary, species, source comp)
        self._db_dcst = {}             # (boundary, species, source comp) -> constant
        # per-element state, created on first use
        self._vel = [None] * nc
        self._tel = [None] * nc
        self._init_tri_state()
        self._host_rng = np.random.default_rng([self._seed, 0x5EED])
        self.ef = None
        self._efreal = None
        self._rd_time = 0.0
        self._ef_time = 0.0

    def _patch_defaults(self):
        """Patch-level surface constants and switches from the model: surface
        reaction constants and activation, diffusion rule constants and, per
        species, whether the patch has a rule for it (Patchdef::reset,
        src/steps/solver/patchdef.cpp:1065-1100)."""
        fm = self.fm
        RS, NG, SS, NP = len(fm.sreacs), len(fm.ghk_chan), len(fm.sspecs), len(self.patches)
        self._psk = np.zeros((NP, RS))
        self._psact = np.ones((NP, RS + NG), dtype=bool)
        self._psd = np.zeros((NP, SS))
        self._psdh = np.zeros((NP, SS), dtype=bool)
        for pi in range(NP):
            for j in fm.patch_sreacs[pi]:
                self._psk[pi, j] = fm.sreac_k[j]
            for q, dcst in fm.patch_sdiffs[pi].values():
                self._psd[pi, q], self._psdh[pi, q] = dcst, True

    def _restore_defaults(self):
        """The kproc resets of TetOpSplitP::reset (tetopsplit.cpp:1436-1503):
        rate and diffusion constants back to the model's (Compdef::reset,
        compdef.cpp:479-504; Patchdef::reset), every reaction, diffusion rule
        and current active (Reac::reset, reac.cpp:96-115; Diff::reset,
        diff.cpp:284-303; SDiff::reset, sdiff.cpp:295-315), directional
        constants cleared, diffusion boundaries closed and clamps released
        (Tet::reset and Tri::reset clear the pool flags)."""
        nc = len(self.comps)
        self._ck[:], self._cd[:] = self._ck0, self._cd0
        self._cract[:] = True
        self._cdact[:] = True
        self._db_active.clear()
        self._db_dcst.clear()
        self._patch_defaults()
        for ci in range(nc):
            if self._vel[ci] is not None:
                self.eng.set_tet_reac_table(None, real=ci)
                self.eng.set_tet_diff_table(None, real=ci)
                self.eng.set_tet_clamp(None, real=ci)
                self._vel[ci] = None
            self._push_reac(ci)
            for sp in range(len(self.fm.specs)):
                self.eng.set_dcst(sp, self._cd[ci, sp], real=ci)
        for sp in range(len(self.fm.specs)):
            self._push_boundaries(sp)
        for r in range(nc):
            if self._tel[r] is not None:
                self.eng.set_tri_sreac_table(None, real=r)
                self.eng.set_tri_sdiff_table(None, real=r)
                self.eng.set_tri_clamp(None, real=r)
                self._tel[r] = None
            if len(self._rtri_patch[r]):
                self._push_tri(r)
        self._init_tri_state()

    def _real_tris(self, r):
        """Global indices of compartment r's triangles in the engine's order."""
        tris = np.nonzero(self._tri_real == r)[0]
        return tris[np.argsort(self._tri_local[tris])]

    def _needs_tri_table(self, r) -> bool:
        """Triangles of several patches in one compartment, or a GHK current
        its patch lacks, need per-triangle constants."""
        pats = np.unique(self._rtri_patch[r])
        if len(pats) > 1:
            return True
        return len(pats) == 1 and len(self.fm.patch_ghk[pats[0]]) < len(self.fm.ghk_chan)

    def _needs_sdiff_table(self, r) -> bool:
        """Surface diffusion needs per-edge constants where a compartment's
        triangles come from patches with different diffusion rules, or where
        a surface diffusion boundary runs (its edges open per species)."""
        pats = np.unique(self._rtri_patch[r])
        if len(pats) == 0 or not self._psdh[pats].any():
            return False
        if self.mts[r].tri_bnd is not None and np.any(self.mts[r].tri_bnd >= 0):
            return True
        return len(pats) > 1 and (np.any(self._psdh[pats] != self._psdh[pats[0]])
                                  or np.any(self._psd[pats] != self._psd[pats[0]]))

    def _real_sd(self, r) -> np.ndarray:
        """Realization-wide surface diffusion constants (no per-edge table):
        its patch's rule constants, 0 for species without a rule."""
        pats = np.unique(self._rtri_patch[r])
        out = np.zeros(len(self.fm.sspecs))
        for pi in pats:
            out = np.where(self._psdh[pi], self._psd[pi], out)
        return out

    def _init_tri_state(self):
        """Patch-level surface constants into the engine, and the per-triangle
        tables of the compartments that need them from the start."""
        for r in range(len(self.comps)):
            if len(self._rtri_patch[r]) == 0:
                continue
            sd = self._real_sd(r)
            for q in range(len(self.fm.sspecs)):
                self.eng.set_sdcst(q, sd[q], real=r)
            if self._needs_tri_table(r):
                self._tri_elements(r)
                self._push_tri(r)
            if self._needs_sdiff_table(r):
                self._sdiff_table_on(r)
                self._push_sdiff(r)

    def _setup_efield(self, geom):
        """Membrane potential on the membrane's conduction volume
        (TetOpSplitP::_setupEField, tetopsplit.cpp:1080-1170)."""
        from .efield import EField

        membs = list(geom.membs.values())
        if len(membs) != 1:
            raise _argerr("the membrane potential solve needs exactly one Memb (got %d)"
                          % len(membs))
        memb = self.memb = membs[0]
        held = [p for p in memb.patches if p.surfsys]
        tris = memb.tris
        if held:
            inner = {p.icomp.name for p in held}
            if len(inner) != 1:
                raise _notimpl("the patches of the membrane must share one inner compartment")
            r = self._ci[inner.pop()]
            rt = self._real_tris(r)
            if set(rt.tolist()) != set(memb.tris.tolist()):
                raise _notimpl("the membrane must consist of the patches that hold the surface "
                               "systems of its inner compartment")
            ohm = [frozenset(oc for s in p.surfsys for oc in self.model.surfsys[s].ohmiccurrs) for p in held]
            if len(set(ohm)) > 1:
                raise _notimpl("every patch of the membrane must hold the same ohmic currents")
            tris = rt                   # EField triangle k is the compartment's triangle k
            self._efreal = r
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
        self._memb_res = (0.0, 0.0)
        if len(self.fm.ohmic):
            if len(self.fm.ohmic) > 8:
                raise _notimpl("at most 8 ohmic currents are supported")
            self.ef.set_ohmic(self.fm.ohmic_chan, self.fm.ohmic_g, self.fm.ohmic_erev)
        if self._efreal is not None:
            self.ef.attach(self.eng, self._efreal)

    # ---------------------------------------------------------------- identity
    def getSolverName(self):
        return self._name

    def getSolverDesc(self):
        return "Operator-split spatial SSA and multinomial diffusion on one MI355X"

    def getSolverAuthors(self):
        return "steps_amd"

    def getSolverEmail(self):
        return ""

    # ----------------------------------------------------------- time stepping
    def reset(self):
        """TetOpSplitP::reset (tetopsplit.cpp:1436-1503): counts, time, step
        count and extents to zero and, with an EField, the membrane potential
        back to DEFAULT_MEMB_POT. The random streams keep running, as the
        reference's RNG does, so reset-and-run loops give independent runs;
        re-initialising the RNG with a new seed before reset() re-keys every
        compartment and restarts the streams from that seed. Rate and
        diffusion constants return to the model's values, every reaction,
        diffusion rule and current is active again, diffusion boundaries
        close and clamps are released, as the reference's kproc resets do."""
        self._restore_defaults()
        seed = int(getattr(self.rng, "seed", self._seed))
        if seed != self._seed:
            self._seed = seed
            for r in range(self.eng.n_real):
                self.eng.set_seed(seed, r)
            self._host_rng = np.random.default_rng([self._seed, 0x5EED])
        self.eng.reset()
        if self.ef is not None:
            self.ef.set_potential(DEFAULT_MEMB_POT)
            if self._efreal is not None:
                self.ef.attach(self.eng, self._efreal)   # channel integrals restart at t = 0

    def repartitionAndReset(self, tet_hosts=(), tri_hosts=None, wm_hosts=()):
        """TetOpSplitP::repartitionAndReset: the whole mesh lives on one GPU, so
        the partition is ignored and the simulation is reset."""
        self.reset()

    def run(self, endtime: float):
        tic = time.perf_counter()
        if self.ef is not None:
            self.ef.run(self.eng, endtime, self._efdt)   # _runWithEField
            self._ef_time += time.perf_counter() - tic
        else:
            self.eng.run(endtime)
            self._rd_time += time.perf_counter() - tic

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

    def setNSteps(self, nsteps: int):
        """TetOpSplitP::setNSteps (tetopsplit.cpp:1995)."""
        if nsteps < 0:
            raise _argerr("the step count must be >= 0")
        for r in range(self.eng.n_real):
            self.eng.set_nsteps(int(nsteps), r)

    getNIteration = getNSteps

    def getUpdPeriod(self) -> float:
        return self.eng.update_period()

    def getReacExtent(self, local: bool = False) -> int:
        return sum(self.eng.extents(r)[0] for r in range(self.eng.n_real))

    def getDiffExtent(self, local: bool = False) -> int:
        return sum(self.eng.extents(r)[1] for r in range(self.eng.n_real))

    def getNReacExtent(self, local: bool = False) -> int:
        return self.getReacExtent(local)

    def getNDiffExtent(self, local: bool = False) -> int:
        return self.getDiffExtent(local)

    def setDiffApplyThreshold(self, threshold: int):
        self.eng.set_diff_apply_threshold(threshold)

    # Wall-clock accounting (TetOpSplitP::getCompTime and the like). One GPU
    # holds the whole mesh: there is no synchronisation, idle or exchange time.
    def getCompTime(self) -> float:
        return self._rd_time + self._ef_time

    def getRDTime(self) -> float:
        return self._rd_time

    def getEFieldTime(self) -> float:
        return self._ef_time

    def getSyncTime(self) -> float:
        return 0.0

    def getIdleTime(self) -> float:
        return 0.0

    def getDataExchangeTime(self) -> float:
        return 0.0

    # ------------------------------------------------------------------ lookups
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

    def _check_patch(self, p: str) -> int:
        if p not in self._pi:
            raise _argerr(f"unknown patch {p!r}")
        return self._pi[p]

    def _patch_sspec(self, pi: int, s: str) -> int:
        q = self._sspec(s)
        if s not in self.fm.patch_sspecs[pi]:
            raise _argerr(f"species {s!r} is undefined in patch {self.patches[pi].name!r}")
        return q

    def _tet(self, t: int):
        """(compartment, local index) of a mesh tetrahedron."""
        if t < 0 or t >= len(self._tet_local) or self._tet_local[t] < 0:
            raise _argerr(f"tetrahedron {t} is not in a compartment")
        return int(self._tet_real[t]), int(self._tet_local[t])

    def _tri(self, t: int):
        """(compartment holding it, local index there, patch) of a patch triangle."""
        if t < 0 or t >= len(self._tri_local) or self._tri_local[t] < 0:
            raise _argerr(f"triangle {t} is not in a patch")
        return int(self._tri_real[t]), int(self._tri_local[t]), int(self._tri_patch[t])

    def _tri_sspec(self, tri: int, s: str):
        r, i, pi = self._tri(tri)
        return r, i, pi, self._patch_sspec(pi, s)

    # ---------------------------------------------------- per-element state
    def _vol_elements(self, ci: int) -> VolElements:
        if self._vel[ci] is None:
            mt = self.mts[ci]
            nb = np.asarray(mt.nbr)
            same = nb >= 0
            if mt.nbr_comp is not None:
                same &= np.asarray(mt.nbr_comp) == ci
            bdir = np.asarray(mt.bdir) if mt.bdir is not None else -np.ones_like(nb)
            self._vel[ci] = VolElements(len(mt.vol), len(self.fm.reacs), len(self.fm.specs),
                                        same & (bdir < 0), bdir)
        return self._vel[ci]

    def _reac_table_on(self, ci: int) -> VolElements:
        vel = self._vol_elements(ci)
        vel.reac_table_on(self._ck[ci], self._cract[ci])
        return vel

    def _push_reac(self, ci: int):
        vel = self._vel[ci]
        if vel is not None and vel.k is not None:
            self.eng.set_tet_reac_table(vel.reac_table(), real=ci)
        else:
            for j in range(len(self.fm.reacs)):
                self.eng.set_reac_k(j, self._ck[ci, j] if self._cract[ci, j] else 0.0, real=ci)

    def _diff_table_on(self, ci: int) -> VolElements:
        vel = self._vol_elements(ci)
        if vel.dD is None:
            dirs = self._bdirs

            def act(d, s):
                return (dirs[d][0], s, ci) in self._db_active

            def dcst(d, s):
                return self._db_dcst.get((dirs[d][0], s, ci))

            vel.diff_table_on(self._cd[ci], self._cdact[ci], act, dcst)
        return vel

    def _push_diff(self, ci: int):
        vel = self._vel[ci]
        if vel is not None and vel.dD is not None:
            self.eng.set_tet_diff_table(vel.diff_table(), real=ci)

    def _push_clamp(self, ci: int):
        vel = self._vel[ci]
        m = None if vel is None or vel.clamp is None or not vel.clamp.any() else vel.clamp
        self.eng.set_tet_clamp(m, real=ci)

    def _tri_elements(self, r: int) -> TriElements:
        if self._tel[r] is None:
            pats = self._rtri_patch[r]
            m = len(pats)
            RS, NG = len(self.fm.sreacs), len(self.fm.ghk_chan)
            held = np.zeros((RS + NG, m), dtype=bool)
            k = np.zeros((RS, m))
            for pi in np.unique(pats):
                sel = pats == pi
                for j in self.fm.patch_sreacs[pi]:
                    held[j, sel] = True
                for g in self.fm.patch_ghk[pi]:
                    held[RS + g, sel] = True
                k[:, sel] = self._psk[pi][:, None]
            tel = TriElements(m, held, self.fm.sreac_vdep >= 0, k, len(self.fm.sspecs))
            for pi in np.unique(pats):
                tel.act[:, pats == pi] = self._psact[pi][:, None]
            self._tel[r] = tel
        return self._tel[r]

    def _push_tri(self, r: int):
        tel = self._tel[r]
        if tel is not None:
            self.eng.set_tri_sreac_table(tel.sreac_table(), real=r)
        else:
            pats = np.unique(self._rtri_patch[r])
            if len(pats) == 1:
                pi = pats[0]
                for j in range(len(self.fm.sreacs)):
                    self.eng.set_sreac_k(j, self._psk[pi, j] if self._psact[pi, j] else 0.0, real=r)

    def _push_tri_clamp(self, r: int):
        tel = self._tel[r]
        m = None if tel is None or tel.clamp is None or not tel.clamp.any() else tel.clamp
        self.eng.set_tri_clamp(m, real=r)

    def _sdiff_table_on(self, r: int) -> TriElements:
        tel = self._tri_elements(r)
        if tel.sD is None:
            mt = self.mts[r]
            pats = self._rtri_patch[r]
            g = np.asarray(mt.tri_edge_g)
            bnd = np.asarray(mt.tri_bnd) if mt.tri_bnd is not None else -np.ones_like(g, dtype=np.int64)
            tel.sdiff_table_on(self._psdh[pats].T, self._psd[pats].T, (g > 0) & (bnd < 0),
                               np.where(g > 0, bnd, -1))
        return tel

    def _push_sdiff(self, r: int):
        tel = self._tel[r]
        if tel is not None and tel.sD is not None:
            self.eng.set_tri_sdiff_table(tel.sdiff_table(), real=r)
            self._push_tri(r)

    def _tri_eD(self, r: int) -> np.ndarray:
        """[SS][m][3] surface diffusion constant through each edge of
        compartment r's triangles, as the engine runs it."""
        tel = self._tel[r]
        if tel is not None and tel.sD is not None:
            return tel.sdiff_table()
        if self.mts[r].tri_edge_g is None:
            return np.zeros((len(self.fm.sspecs), len(self.mts[r].tri_tet), 3))
        g = np.asarray(self.mts[r].tri_edge_g)
        return np.where(g[None] > 0, self._real_sd(r)[:, None, None], 0.0)

    # ------------------------------------------------------------- compartments
    def getNComps(self) -> int:
        return len(self.comps)

    def getCompName(self, c_idx: int) -> str:
        if not 0 <= c_idx < len(self.comps):
            raise _argerr("compartment index out of range")
        return self.comps[c_idx].name

    def getNCompSpecs(self, c: str) -> int:
        return len(self._comp_specs[self._check_comp(c)])

    def getCompSpecName(self, c: str, s_idx: int) -> str:
        names = [s for s in self.fm.specs if s in self._comp_specs[self._check_comp(c)]]
        if not 0 <= s_idx < len(names):
            raise _argerr("species index out of range")
        return names[s_idx]

    def getCompVol(self, c: str) -> float:
        return float(self.mts[self._check_comp(c)].vol.sum())

    def getCompSpecCount(self, c: str, s: str) -> float:
        ci = self._check_comp(c)
        return float(self.eng.sum_tet_counts(self._comp_spec(ci, s), real=ci))

    def setCompSpecCount(self, c: str, s: str, n: float):
        ci = self._check_comp(c)
        sp = self._comp_spec(ci, s)
        new = distribute_counts(n, self.mts[ci].vol, self._host_rng)
        self.eng.set_tet_counts(sp, new, real=ci)

    def getCompSpecAmount(self, c: str, s: str) -> float:
        return self.getCompSpecCount(c, s) / AVOGADRO

    def setCompSpecAmount(self, c: str, s: str, a: float):
        self.setCompSpecCount(c, s, a * AVOGADRO)

    def getCompSpecConc(self, c: str, s: str) -> float:
        return self.getCompSpecCount(c, s) / (1e3 * self.getCompVol(c) * AVOGADRO)

    def setCompSpecConc(self, c: str, s: str, conc: float):
        self.setCompSpecCount(c, s, conc * 1e3 * self.getCompVol(c) * AVOGADRO)

    def getCompSpecClamped(self, c: str, s: str) -> bool:
        """True when the species is clamped in every tet of the compartment
        (_getCompSpecClamped)."""
        ci = self._check_comp(c)
        sp = self._comp_spec(ci, s)
        vel = self._vel[ci]
        return bool(vel is not None and vel.clamped(sp, np.arange(vel.n)).all())

    def setCompSpecClamped(self, c: str, s: str, b: bool):
        """_setCompSpecClamped (tetopsplit.cpp:2189): every tet of the compartment."""
        ci = self._check_comp(c)
        sp = self._comp_spec(ci, s)
        vel = self._vol_elements(ci)
        vel.set_clamped(sp, np.arange(vel.n), bool(b))
        self._push_clamp(ci)

    def _reac(self, ci: int, r: str) -> int:
        if r not in self._comp_reacs[ci]:
            raise _argerr(f"unknown reaction {r!r} in compartment {self.comps[ci].name!r}")
        return self._comp_reacs[ci][r]

    def getCompReacK(self, c: str, r: str) -> float:
        ci = self._check_comp(c)
        return float(self._ck[ci, self._reac(ci, r)])

    def setCompReacK(self, c: str, r: str, k: float):
        """_setCompReacK: the constant of the reaction in every tet."""
        ci = self._check_comp(c)
        j = self._reac(ci, r)
        if k < 0:
            raise _argerr("rate constant must be >= 0")
        self._ck[ci, j] = k
        vel = self._vel[ci]
        if vel is not None and vel.k is not None:
            vel.k[j, :] = k
        self._push_reac(ci)

    def getCompReacActive(self, c: str, r: str) -> bool:
        """True unless the reaction is inactive in every tet (_getCompReacActive)."""
        ci = self._check_comp(c)
        j = self._reac(ci, r)
        vel = self._vel[ci]
        if vel is not None and vel.kact is not None:
            return bool(vel.kact[j].any())
        return bool(self._cract[ci, j])

    def setCompReacActive(self, c: str, r: str, a: bool):
        """_setCompReacActive (tetopsplit.cpp:2309): the flag of every tet."""
        ci = self._check_comp(c)
        j = self._reac(ci, r)
        self._cract[ci, j] = bool(a)
        vel = self._vel[ci]
        if vel is not None and vel.kact is not None:
            vel.kact[j, :] = bool(a)
        self._push_reac(ci)

    def _reac_hc(self, ci: int, j: int, tets=None):
        """(h, c) per tet of compartment ci (or the given local tets) for
        reaction j at the present counts (Reac::rate, Reac::resetCcst)."""
        vel = self._vel[ci]
        idx = np.arange(len(self.mts[ci].vol)) if tets is None else np.asarray(tets)
        lhs = self.fm.reac_lhs[j]
        counts = np.zeros((len(lhs), len(idx)), dtype=np.int64)
        for s in np.nonzero(lhs)[0]:
            counts[s] = self.eng.batch_tet_counts(int(s), idx, real=ci)
        h = _falling_h(counts, lhs)
        if vel is not None and vel.k is not None:
            k = vel.reac_table()[j, idx]
        else:
            k = np.full(len(idx), self._ck[ci, j] if self._cract[ci, j] else 0.0)
        order = int(lhs.sum())
        vscale = 1e3 * self.mts[ci].vol[idx] * AVOGADRO
        return h, k * vscale ** (1 - order)

    def getCompReacH(self, c: str, r: str) -> float:
        ci = self._check_comp(c)
        return float(self._reac_hc(ci, self._reac(ci, r))[0].sum())

    def getCompReacC(self, c: str, r: str) -> float:
        """Volume-weighted mean of the tets' c (_getCompReacC)."""
        ci = self._check_comp(c)
        _, cc = self._reac_hc(ci, self._reac(ci, r))
        v = self.mts[ci].vol
        return float((cc * v).sum() / v.sum())

    def getCompReacA(self, c: str, r: str) -> float:
        ci = self._check_comp(c)
        h, cc = self._reac_hc(ci, self._reac(ci, r))
        return float((h * cc).sum())

    def _diff(self, ci: int, d: str) -> int:
        if d not in self._comp_diffs[ci]:
            raise _argerr(f"unknown diffusion rule {d!r} in compartment {self.comps[ci].name!r}")
        return self._comp_diffs[ci][d]

    def getCompDiffD(self, c: str, d: str) -> float:
        ci = self._check_comp(c)
        return float(self._cd[ci, self._diff(ci, d)])

    def setCompDiffD(self, c: str, d: str, dcst: float):
        """_setCompDiffD: Diff::setDcst on every tet, which also clears the
        rule's directional constants (diff.cpp:340-390)."""
        ci = self._check_comp(c)
        s = self._diff(ci, d)
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        self._cd[ci, s] = dcst
        for key in [k for k in self._db_dcst if k[1] == s and k[2] == ci]:
            del self._db_dcst[key]
        vel = self._vel[ci]
        if vel is not None and vel.dD is not None:
            vel.set_dcst(s, slice(None), dcst)
            self._push_diff(ci)
        self.eng.set_dcst(s, dcst if self._cdact[ci, s] else 0.0, real=ci)
        self._push_boundaries(s)

    def getCompDiffActive(self, c: str, d: str) -> bool:
        ci = self._check_comp(c)
        s = self._diff(ci, d)
        vel = self._vel[ci]
        if vel is not None and vel.dAct is not None:
            return bool(vel.dAct[s].any())
        return bool(self._cdact[ci, s])

    def setCompDiffActive(self, c: str, d: str, act: bool):
        """_setCompDiffActive (tetopsplit.cpp:2437): the flag of every tet; an
        inactive rule moves nothing, across diffusion boundaries either."""
        ci = self._check_comp(c)
        s = self._diff(ci, d)
        self._cdact[ci, s] = bool(act)
        vel = self._vel[ci]
        if vel is not None and vel.dAct is not None:
            vel.dAct[s, :] = bool(act)
            self._push_diff(ci)
        self.eng.set_dcst(s, self._cd[ci, s] if act else 0.0, real=ci)
        self._push_boundaries(s)

    # ----------------------------------------- diffusion boundaries (2745-2931)
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
        """Realization-level boundary constants of species sp: a direction from
        compartment c carries c's diffusion constant for the species, or the
        directional constant set for it; 0 when the boundary is closed for it,
        when c has no diffusion rule for it (no Diff kproc there) or the rule is
        inactive. Compartments on per-face tables carry their own values."""
        for d, (bi, src, _dst) in enumerate(self._bdirs):
            has_rule = sp in self._comp_diffs[src].values()
            on = (bi, sp, src) in self._db_active and has_rule and self._cdact[src, sp]
            val = self._db_dcst.get((bi, sp, src), self._cd[src, sp]) if on else 0.0
            self.eng.set_boundary_dcst(d, sp, val)

    def _db_dirs(self, bi: int, src: int):
        return [d for d, (b, s_, _) in enumerate(self._bdirs) if b == bi and s_ == src]

    def setDiffBoundarySpecDiffusionActive(self, db: str, s: str, act: bool):
        bi = self._db(db)
        sp = self._db_spec(bi, s)
        b = self.diffbs[bi]
        for c in (b.compA, b.compB):
            ci = self._ci[c.name]
            key = (bi, sp, ci)
            if act:
                self._db_active.add(key)
            else:
                self._db_active.discard(key)
            self._db_dcst.pop(key, None)    # setDiffBndActive -> setDcst clears overrides
            vel = self._vel[ci]
            if vel is not None and vel.dD is not None:
                for d in self._db_dirs(bi, ci):
                    vel.set_bnd_active(sp, d, bool(act))
                self._push_diff(ci)
        self._push_boundaries(sp)

    def getDiffBoundarySpecDiffusionActive(self, db: str, s: str) -> bool:
        """True when the species crosses the boundary in both directions."""
        bi = self._db(db)
        sp = self._db_spec(bi, s)
        b = self.diffbs[bi]
        return all((bi, sp, self._ci[c.name]) in self._db_active for c in (b.compA, b.compB))

    def setDiffBoundarySpecDcst(self, db: str, s: str, dcst: float, direction_comp=None):
        """Open the boundary for s with constant dcst from both sides, or only
        into direction_comp (tetopsplit.cpp:2860-2917)."""
        bi = self._db(db)
        sp = self._db_spec(bi, s)
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        b = self.diffbs[bi]
        dest = None
        if direction_comp not in (None, ""):
            name = direction_comp if isinstance(direction_comp, str) else direction_comp.name
            if name not in (b.compA.name, b.compB.name):
                raise _argerr("direction compartment is not on the diffusion boundary")
            dest = self._ci[name]
        for c in (b.compA, b.compB):
            src = self._ci[c.name]
            if src == dest:
                continue
            self._db_active.add((bi, sp, src))
            self._db_dcst[(bi, sp, src)] = float(dcst)
            vel = self._vel[src]
            if vel is not None and vel.dD is not None:
                for d in self._db_dirs(bi, src):
                    for t, f in zip(*np.nonzero(vel.bdir == d)):
                        vel.set_direction_dcst(sp, t, f, float(dcst))
                self._push_diff(src)
        self._push_boundaries(sp)

    # ----------------------------- surface diffusion boundaries (3023-3200)
    def _sdb(self, sdb: str) -> int:
        if sdb not in self._sdbi:
            raise _argerr(f"unknown surface diffusion boundary {sdb!r}")
        return self._sdbi[sdb]

    def _sdb_spec(self, bi: int, s: str) -> int:
        b = self.sdiffbs[bi]
        for p in (b.patchA, b.patchB):
            if s not in self.fm.patch_sspecs[self._pi[p.name]]:
                raise _argerr("Species undefined in patches connected by surface diffusion boundary.")
        return self.fm.sspecs.index(s)

    def _sdb_edges(self, bi: int):
        """(global triangle, compartment, index there, patch, edge slot) of
        every boundary triangle, per bar the patchA triangle then the patchB
        one (SDiffBoundary::getTris / getTriDirection order)."""
        b = self.sdiffbs[bi]
        for bar, pair in zip(b.bars, b.bar_tris):
            for t in pair:
                r, i, pi = self._tri(int(t))
                yield int(t), r, i, pi, self.geom.tri_edge_of_bar(int(t), int(bar))

    def setSDiffBoundarySpecDiffusionActive(self, sdb: str, s: str, act: bool):
        """_setSDiffBoundarySpecDiffusionActive (tetopsplit.cpp:3023-3078):
        SDiff::setSDiffBndActive on the boundary edge of every boundary
        triangle whose patch has a diffusion rule for s."""
        bi = self._sdb(sdb)
        q = self._sdb_spec(bi, s)
        touched = set()
        for _, r, i, pi, e in self._sdb_edges(bi):
            if not self._psdh[pi, q]:
                continue
            self._sdiff_table_on(r).set_sbnd_active(q, i, e, bool(act))
            touched.add(r)
        for r in touched:
            self._push_sdiff(r)

    def getSDiffBoundarySpecDiffusionActive(self, sdb: str, s: str) -> bool:
        """_getSDiffBoundarySpecDiffusionActive (tetopsplit.cpp:3081-3140): the
        flag of the last boundary triangle (in boundary order) whose patch has
        a diffusion rule for s; true when there is none."""
        bi = self._sdb(sdb)
        q = self._sdb_spec(bi, s)
        active = True
        for _, r, i, pi, e in self._sdb_edges(bi):
            if not self._psdh[pi, q]:
                continue
            tel = self._tel[r]
            active = bool(tel is not None and tel.sBnd is not None and tel.sBnd[q, i, e])
        return active

    def setSDiffBoundarySpecDcst(self, sdb: str, s: str, dcst: float, direction_patch=""):
        """_setSDiffBoundarySpecDcst (tetopsplit.cpp:3143-3200):
        SDiff::setDirectionDcst on the boundary edge of every boundary
        triangle, which opens it; with direction_patch, only on the triangles
        of the other patch (diffusion toward direction_patch)."""
        bi = self._sdb(sdb)
        q = self._sdb_spec(bi, s)
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        skip = -1
        if direction_patch not in (None, ""):
            name = direction_patch if isinstance(direction_patch, str) else direction_patch.name
            skip = self._check_patch(name)
        touched = set()
        for _, r, i, pi, e in self._sdb_edges(bi):
            if pi == skip or not self._psdh[pi, q]:
                continue
            self._sdiff_table_on(r).set_sdirection_dcst(q, i, e, float(dcst))
            touched.add(r)
        for r in touched:
            self._push_sdiff(r)

    # deprecated names (src/steps/solver/api_deprecated.cpp:560-590)
    def setSDiffBoundaryDiffusionActive(self, sdb: str, s: str, act: bool):
        self.setSDiffBoundarySpecDiffusionActive(sdb, s, act)

    def getSDiffBoundaryDiffusionActive(self, sdb: str, s: str) -> bool:
        return self.getSDiffBoundarySpecDiffusionActive(sdb, s)

    def setSDiffBoundaryDcst(self, sdb: str, s: str, dcst: float, direction_patch=""):
        self.setSDiffBoundarySpecDcst(sdb, s, dcst, direction_patch)

    # deprecated spellings (cysteps_solver.pyx)
    def setDiffBoundaryDiffusionActive(self, db: str, s: str, act: bool):
        self.setDiffBoundarySpecDiffusionActive(db, s, act)

    def getDiffBoundaryDiffusionActive(self, db: str, s: str) -> bool:
        return self.getDiffBoundarySpecDiffusionActive(db, s)

    def setDiffBoundaryDcst(self, db: str, s: str, dcst: float, direction_comp=""):
        self.setDiffBoundarySpecDcst(db, s, dcst, direction_comp or None)

    # -------------------------------------------------------------- tetrahedra
    def getTetVol(self, t: int) -> float:
        ci, i = self._tet(t)
        return float(self.mts[ci].vol[i])

    def getTetSpecDefined(self, t: int, s: str) -> bool:
        ci, _ = self._tet(t)
        self._spec(s)
        return s in self._comp_specs[ci]

    def getTetSpecCount(self, t: int, s: str) -> float:
        ci, i = self._tet(t)
        return float(self.eng.batch_tet_counts(self._comp_spec(ci, s), [i], real=ci)[0])

    def setTetSpecCount(self, t: int, s: str, n: float):
        if n < 0:
            raise _argerr("negative number of molecules")
        ci, i = self._tet(t)
        sp = self._comp_spec(ci, s)
        c = self.eng.get_tet_counts(sp, real=ci)
        c[i] = _round_count(n, self._host_rng)
        self.eng.set_tet_counts(sp, c, real=ci)

    def getTetSpecAmount(self, t: int, s: str) -> float:
        return self.getTetSpecCount(t, s) / AVOGADRO

    def setTetSpecAmount(self, t: int, s: str, a: float):
        self.setTetSpecCount(t, s, a * AVOGADRO)

    def getTetSpecConc(self, t: int, s: str) -> float:
        return self.getTetSpecCount(t, s) / (1e3 * self.getTetVol(t) * AVOGADRO)

    def setTetSpecConc(self, t: int, s: str, conc: float):
        self.setTetSpecCount(t, s, conc * 1e3 * self.getTetVol(t) * AVOGADRO)

    def getTetSpecClamped(self, t: int, s: str) -> bool:
        ci, i = self._tet(t)
        sp = self._comp_spec(ci, s)
        vel = self._vel[ci]
        return bool(vel is not None and vel.clamped(sp, i))

    def setTetSpecClamped(self, t: int, s: str, buf: bool):
        """_setTetSpecClamped (tetopsplit.cpp:3951)."""
        ci, i = self._tet(t)
        sp = self._comp_spec(ci, s)
        self._vol_elements(ci).set_clamped(sp, i, bool(buf))
        self._push_clamp(ci)

    def getTetReacK(self, t: int, r: str) -> float:
        ci, i = self._tet(t)
        j = self._reac(ci, r)
        vel = self._vel[ci]
        return float(vel.k[j, i]) if vel is not None and vel.k is not None else float(self._ck[ci, j])

    def setTetReacK(self, t: int, r: str, kf: float):
        """_setTetReacK (tetopsplit.cpp:3996)."""
        ci, i = self._tet(t)
        j = self._reac(ci, r)
        if kf < 0:
            raise _argerr("rate constant must be >= 0")
        self._reac_table_on(ci).k[j, i] = kf
        self._push_reac(ci)

    def getTetReacActive(self, t: int, r: str) -> bool:
        ci, i = self._tet(t)
        j = self._reac(ci, r)
        vel = self._vel[ci]
        if vel is not None and vel.kact is not None:
            return bool(vel.kact[j, i])
        return bool(self._cract[ci, j])

    def setTetReacActive(self, t: int, r: str, act: bool):
        """_setTetReacActive (tetopsplit.cpp:4045)."""
        ci, i = self._tet(t)
        j = self._reac(ci, r)
        self._reac_table_on(ci).kact[j, i] = bool(act)
        self._push_reac(ci)

    def _tet_face_toward(self, t: int, other: int) -> int:
        f = np.nonzero(self.geom.tet_nbr[t] == other)[0]
        if len(f) == 0:
            raise _argerr(f"Tetrahedron {other} is not a neighbor of tetrahedron {t}.")
        return int(f[0])

    def getTetDiffD(self, t: int, d: str, direction_tet=None) -> float:
        ci, i = self._tet(t)
        s = self._diff(ci, d)
        vel = self._vel[ci]
        if vel is None or vel.dD is None:
            if direction_tet is not None and direction_tet >= 0:
                bi = self._tet_face_toward(t, int(direction_tet))
                bd = self.mts[ci].bdir[i, bi] if self.mts[ci].bdir is not None else -1
                if bd >= 0 and (self._bdirs[bd][0], s, ci) in self._db_dcst:
                    return float(self._db_dcst[(self._bdirs[bd][0], s, ci)])
            return float(self._cd[ci, s])
        if direction_tet is None or direction_tet < 0:
            return float(vel.dD[s, i])
        v = vel.dDir[s, i, self._tet_face_toward(t, int(direction_tet))]
        return float(vel.dD[s, i] if np.isnan(v) else v)

    def setTetDiffD(self, t: int, d: str, dk: float, direction_tet=None):
        """_setTetDiffD (tetopsplit.cpp:4103): the rule's constant in one tet,
        or through the face toward direction_tet (Diff::setDirectionDcst)."""
        ci, i = self._tet(t)
        s = self._diff(ci, d)
        if dk < 0:
            raise _argerr("diffusion constant must be >= 0")
        vel = self._diff_table_on(ci)
        if direction_tet is None or direction_tet < 0:
            vel.set_dcst(s, i, float(dk))
        else:
            vel.set_direction_dcst(s, i, self._tet_face_toward(t, int(direction_tet)), float(dk))
        self._push_diff(ci)

    def getTetDiffActive(self, t: int, d: str) -> bool:
        ci, i = self._tet(t)
        s = self._diff(ci, d)
        vel = self._vel[ci]
        if vel is not None and vel.dAct is not None:
            return bool(vel.dAct[s, i])
        return bool(self._cdact[ci, s])

    def setTetDiffActive(self, t: int, d: str, act: bool):
        """_setTetDiffActive (tetopsplit.cpp:4166)."""
        ci, i = self._tet(t)
        s = self._diff(ci, d)
        self._diff_table_on(ci).dAct[s, i] = bool(act)
        self._push_diff(ci)

    def getTetReacH(self, t: int, r: str) -> float:
        ci, i = self._tet(t)
        return float(self._reac_hc(ci, self._reac(ci, r), [i])[0][0])

    def getTetReacC(self, t: int, r: str) -> float:
        ci, i = self._tet(t)
        return float(self._reac_hc(ci, self._reac(ci, r), [i])[1][0])

    def getTetReacA(self, t: int, r: str) -> float:
        ci, i = self._tet(t)
        h, c = self._reac_hc(ci, self._reac(ci, r), [i])
        return float(h[0] * c[0])

    def getTetDiffA(self, t: int, d: str) -> float:
        """Diff::rate (diff.cpp:458-470): the count times sum_f D_f A_f / (V d_f)."""
        ci, i = self._tet(t)
        s = self._diff(ci, d)
        g = self.mts[ci].face_g[i]
        vel = self._vel[ci]
        if vel is not None and vel.dD is not None:
            D = vel.diff_table()[s, i]
        else:
            D = np.full(4, self._cd[ci, s] if self._cdact[ci, s] else 0.0)
            bd = self.mts[ci].bdir[i] if self.mts[ci].bdir is not None else -np.ones(4, dtype=int)
            for f in range(4):
                if bd[f] >= 0:
                    D[f] = self.eng.boundary_dcst(int(bd[f]), s)
        rate = float(((D[0] * g[0] + D[1] * g[1]) + D[2] * g[2]) + D[3] * g[3])
        return rate * float(self.eng.batch_tet_counts(s, [i], real=ci)[0])

    # batch access
    def _tet_idx(self, index_array):
        idx = np.asarray(index_array, dtype=np.int64)
        if len(idx) and (np.any(idx < 0) or np.any(idx >= len(self._tet_local))
                         or np.any(self._tet_local[idx] < 0)):
            raise _argerr("tetrahedron index out of range")
        return idx

    def getBatchTetSpecCounts(self, tets, s: str):
        c = np.zeros(len(tets))
        self.getBatchTetSpecCountsNP(np.asarray(tets, dtype=np.int64), s, c)
        return c.tolist()

    def getBatchTetSpecCountsNP(self, index_array, s: str, counts):
        idx = self._tet_idx(index_array)
        if len(counts) != len(idx):
            raise _argerr("counts and index arrays differ in length")
        sp = self._spec(s)
        real, loc = self._tet_real[idx], self._tet_local[idx]
        for ci in np.unique(real):
            m = real == ci
            counts[m] = self.eng.batch_tet_counts(sp, loc[m], real=int(ci))

    def getBatchTetSpecConcs(self, tets, s: str):
        c = np.zeros(len(tets))
        self.getBatchTetSpecConcsNP(np.asarray(tets, dtype=np.int64), s, c)
        return c.tolist()

    def getBatchTetSpecConcsNP(self, index_array, s: str, concs):
        self.getBatchTetSpecCountsNP(index_array, s, concs)
        idx = np.asarray(index_array, dtype=np.int64)
        vol = self.geom.tet_vol[idx]
        concs[:] = concs / (1e3 * vol * AVOGADRO)

    def setBatchTetSpecCountsNP(self, index_array, s: str, counts):
        """Counts of the listed tets, rounded stochastically
        (TetOpSplitP::setBatchTetSpecCounts)."""
        idx = self._tet_idx(index_array)
        vals = np.asarray(counts, dtype=np.float64)
        if len(vals) != len(idx):
            raise _argerr("counts and index arrays differ in length")
        if np.any(vals < 0):
            raise _argerr("negative number of molecules")
        sp = self._spec(s)
        real, loc = self._tet_real[idx], self._tet_local[idx]
        fl = np.floor(vals)
        rounded = (fl + (self._host_rng.random(len(vals)) < vals - fl)).astype(np.int64)
        for ci in np.unique(real):
            m = real == ci
            self._comp_spec(int(ci), s)
            c = self.eng.get_tet_counts(sp, real=int(ci))
            c[loc[m]] = rounded[m]
            self.eng.set_tet_counts(sp, c, real=int(ci))

    def setBatchTetSpecCounts(self, tets, s: str, counts):
        self.setBatchTetSpecCountsNP(np.asarray(tets, dtype=np.int64), s, np.asarray(counts, dtype=float))

    def setBatchTetSpecConcsNP(self, index_array, s: str, concs):
        idx = self._tet_idx(index_array)
        self.setBatchTetSpecCountsNP(idx, s, np.asarray(concs, dtype=float) * 1e3 * self.geom.tet_vol[idx]
                                     * AVOGADRO)

    def setBatchTetSpecConcs(self, tets, s: str, concs):
        self.setBatchTetSpecConcsNP(np.asarray(tets, dtype=np.int64), s, concs)

    def sumBatchTetSpecCountsNP(self, tet_array, s: str) -> float:
        c = np.zeros(len(tet_array))
        self.getBatchTetSpecCountsNP(tet_array, s, c)
        return float(c.sum())

    sumBatchTetCountsNP = sumBatchTetSpecCountsNP

    # -------------------------------------------------------- regions of interest
    def _roi(self, roi: str, want: int):
        if roi not in self.geom.rois:
            raise _argerr(f"ROI {roi!r} does not exist")
        typ, idx = self.geom.rois[roi]
        if typ != want:
            raise _argerr(f"ROI {roi!r} holds {'tetrahedra' if typ == ELEM_TET else 'other elements'}, "
                          f"not {'tetrahedra' if want == ELEM_TET else 'triangles'}")
        return idx

    def _roi_tets(self, roi: str):
        idx = self._roi(roi, ELEM_TET)
        if np.any(self._tet_local[idx] < 0):
            raise _argerr(f"ROI {roi!r} holds a tetrahedron outside every compartment")
        return idx

    def _roi_tris(self, roi: str):
        idx = self._roi(roi, ELEM_TRI)
        if np.any(self._tri_local[idx] < 0):
            raise _argerr(f"ROI {roi!r} holds a triangle outside every patch")
        return idx

    def _roi_is_tri(self, roi: str) -> bool:
        return roi in self.geom.rois and self.geom.rois[roi][0] == ELEM_TRI

    def getROIVol(self, roi: str) -> float:
        return float(self.geom.tet_vol[self._roi_tets(roi)].sum())

    def getROIArea(self, roi: str) -> float:
        return float(self.geom.tri_area[self._roi(roi, ELEM_TRI)].sum())

    def getROITetSpecCounts(self, roi: str, s: str):
        idx = self._roi_tets(roi)
        c = np.zeros(len(idx))
        self.getBatchTetSpecCountsNP(idx, s, c)
        return c.tolist()

    def getROITetSpecCountsNP(self, roi: str, s: str, counts):
        self.getBatchTetSpecCountsNP(self._roi_tets(roi), s, counts)

    def getROITriSpecCounts(self, roi: str, s: str):
        idx = self._roi(roi, ELEM_TRI)
        c = np.zeros(len(idx))
        self.getBatchTriSpecCountsNP(idx, s, c)
        return c.tolist()

    def getROITriSpecCountsNP(self, roi: str, s: str, counts):
        self.getBatchTriSpecCountsNP(self._roi(roi, ELEM_TRI), s, counts)

    def getROISpecCount(self, roi: str, s: str) -> float:
        """Total over the ROI, reduced on the device per compartment."""
        if self._roi_is_tri(roi):
            return float(sum(self.getROITriSpecCounts(roi, s)))
        idx = self._roi_tets(roi)
        sp = self._spec(s)
        real, loc = self._tet_real[idx], self._tet_local[idx]
        return float(sum(self.eng.sum_batch_tet_counts(sp, loc[real == ci], real=int(ci))
                         for ci in np.unique(real)))

    def setROISpecCount(self, roi: str, s: str, n: float):
        """Distribute n molecules over the ROI's tets by volume, or its
        triangles by area (setROISpecCount)."""
        if self._roi_is_tri(roi):
            idx = self._roi_tris(roi)
            self.setBatchTriSpecCountsNP(idx, s, distribute_counts(n, self.geom.tri_area[idx], self._host_rng))
            return
        idx = self._roi_tets(roi)
        sp = self._spec(s)
        real, loc = self._tet_real[idx], self._tet_local[idx]
        add = distribute_counts(n, self.geom.tet_vol[idx], self._host_rng)
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

    def _roi_by_comp(self, roi: str):
        idx = self._roi_tets(roi)
        real, loc = self._tet_real[idx], self._tet_local[idx]
        return [(int(ci), loc[real == ci]) for ci in np.unique(real)]

    def setROISpecClamped(self, roi: str, s: str, clamped: bool):
        """Clamp the species on every tet (or triangle) of the ROI."""
        if self._roi_is_tri(roi):
            for t in self._roi_tris(roi):
                self.setTriSpecClamped(int(t), s, clamped)
            return
        for ci, loc in self._roi_by_comp(roi):
            sp = self._comp_spec(ci, s)
            self._vol_elements(ci).set_clamped(sp, loc, bool(clamped))
            self._push_clamp(ci)

    def setROIReacK(self, roi: str, r: str, kf: float):
        if kf < 0:
            raise _argerr("rate constant must be >= 0")
        for ci, loc in self._roi_by_comp(roi):
            j = self._reac(ci, r)
            self._reac_table_on(ci).k[j, loc] = kf
            self._push_reac(ci)

    def setROIReacActive(self, roi: str, r: str, a: bool):
        for ci, loc in self._roi_by_comp(roi):
            j = self._reac(ci, r)
            self._reac_table_on(ci).kact[j, loc] = bool(a)
            self._push_reac(ci)

    def setROIDiffD(self, roi: str, d: str, dcst: float):
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        for ci, loc in self._roi_by_comp(roi):
            s = self._diff(ci, d)
            self._diff_table_on(ci).set_dcst(s, loc, float(dcst))
            self._push_diff(ci)

    def setROIDiffActive(self, roi: str, d: str, act: bool):
        for ci, loc in self._roi_by_comp(roi):
            s = self._diff(ci, d)
            self._diff_table_on(ci).dAct[s, loc] = bool(act)
            self._push_diff(ci)

    def setROISReacK(self, roi: str, sr: str, kf: float):
        for t in self._roi_tris(roi):
            self.setTriSReacK(int(t), sr, kf)

    def setROISReacActive(self, roi: str, sr: str, a: bool):
        for t in self._roi_tris(roi):
            self.setTriSReacActive(int(t), sr, a)

    def setROIVDepSReacActive(self, roi: str, vsr: str, a: bool):
        for t in self._roi_tris(roi):
            self.setTriVDepSReacActive(int(t), vsr, a)

    # deprecated spellings (cysteps_solver.pyx)
    def getROICount(self, roi: str, s: str) -> float:
        return self.getROISpecCount(roi, s)

    def setROICount(self, roi: str, s: str, n: float):
        self.setROISpecCount(roi, s, n)

    def getROIAmount(self, roi: str, s: str) -> float:
        return self.getROISpecAmount(roi, s)

    def setROIAmount(self, roi: str, s: str, a: float):
        self.setROISpecAmount(roi, s, a)

    def getROIConc(self, roi: str, s: str) -> float:
        return self.getROISpecConc(roi, s)

    def setROIConc(self, roi: str, s: str, c: float):
        self.setROISpecConc(roi, s, c)

    def setROIClamped(self, roi: str, s: str, b: bool):
        self.setROISpecClamped(roi, s, b)

    def getROITetCounts(self, roi: str, s: str):
        return self.getROITetSpecCounts(roi, s)

    def getROITetCountsNP(self, roi: str, s: str, counts):
        self.getROITetSpecCountsNP(roi, s, counts)

    def getROITriCounts(self, roi: str, s: str):
        return self.getROITriSpecCounts(roi, s)

    def getROITriCountsNP(self, roi: str, s: str, counts):
        self.getROITriSpecCountsNP(roi, s, counts)

    # ------------------------------------------------------ patches and triangles
    def getNPatches(self) -> int:
        return len(self.patches)

    def getPatchName(self, p_idx: int) -> str:
        if not 0 <= p_idx < len(self.patches):
            raise _argerr("patch index out of range")
        return self.patches[p_idx].name

    def getNPatchSpecs(self, p: str) -> int:
        return len(self.fm.patch_sspecs[self._check_patch(p)])

    def getPatchSpecName(self, p: str, s_idx: int) -> str:
        pi = self._check_patch(p)
        names = [s for s in self.fm.sspecs if s in self.fm.patch_sspecs[pi]]
        if not 0 <= s_idx < len(names):
            raise _argerr("species index out of range")
        return names[s_idx]

    def getPatchArea(self, p: str) -> float:
        return float(self.geom.tri_area[self.patches[self._check_patch(p)].tris].sum())

    def _patch_local(self, pi: int):
        """(compartment holding the patch's triangles, their local indices)."""
        p = self.patches[pi]
        return self._ci[p.icomp.name], self._tri_local[p.tris]

    def getPatchSpecCount(self, p: str, s: str) -> float:
        pi = self._check_patch(p)
        q = self._patch_sspec(pi, s)
        r, loc = self._patch_local(pi)
        if len(self._rtri_patch[r]) == len(loc):
            return float(self.eng.sum_tri_counts(q, real=r))
        return float(self.eng.get_tri_counts(q, real=r)[loc].sum())

    def setPatchSpecCount(self, p: str, s: str, n: float):
        pi = self._check_patch(p)
        q = self._patch_sspec(pi, s)
        r, loc = self._patch_local(pi)
        c = self.eng.get_tri_counts(q, real=r)
        c[loc] = distribute_counts(n, self.geom.tri_area[self.patches[pi].tris], self._host_rng)
        self.eng.set_tri_counts(q, c, real=r)

    def getPatchSpecAmount(self, p: str, s: str) -> float:
        return self.getPatchSpecCount(p, s) / AVOGADRO

    def setPatchSpecAmount(self, p: str, s: str, a: float):
        self.setPatchSpecCount(p, s, a * AVOGADRO)

    def getPatchSpecClamped(self, p: str, s: str) -> bool:
        pi = self._check_patch(p)
        q = self._patch_sspec(pi, s)
        r, loc = self._patch_local(pi)
        tel = self._tel[r]
        return bool(tel is not None and tel.clamped(q, loc).all())

    def setPatchSpecClamped(self, p: str, s: str, b: bool):
        """_setPatchSpecClamped: every triangle of the patch."""
        pi = self._check_patch(p)
        q = self._patch_sspec(pi, s)
        r, loc = self._patch_local(pi)
        self._tri_elements(r).set_clamped(q, loc, bool(b))
        self._push_tri(r)
        self._push_tri_clamp(r)

    def _sreac(self, r, pi=None, vdep=None) -> int:
        try:
            j = self.fm.sreacs.index(r)
        except ValueError:
            raise _argerr(f"unknown surface reaction {r!r}") from None
        if pi is not None and j not in self.fm.patch_sreacs[pi]:
            raise _argerr(f"surface reaction {r!r} is undefined in patch {self.patches[pi].name!r}")
        if vdep is not None and bool(self.fm.sreac_vdep[j] >= 0) != vdep:
            raise _argerr(f"{r!r} is {'not ' if vdep else ''}a voltage-dependent surface reaction")
        return j

    def getPatchSReacK(self, p: str, r: str) -> float:
        pi = self._check_patch(p)
        return float(self._psk[pi, self._sreac(r, pi, vdep=False)])

    def setPatchSReacK(self, p: str, r: str, k: float):
        """_setPatchSReacK: the constant of the reaction on every triangle of the patch."""
        pi = self._check_patch(p)
        j = self._sreac(r, pi, vdep=False)
        if k < 0:
            raise _argerr("rate constant must be >= 0")
        self._psk[pi, j] = k
        rr, loc = self._patch_local(pi)
        if self._tel[rr] is not None:
            self._tel[rr].k[j, loc] = k
        self._push_tri(rr)

    def _set_patch_active(self, p: str, r: str, a: bool, vdep: bool):
        pi = self._check_patch(p)
        j = self._sreac(r, pi, vdep=vdep)
        self._psact[pi, j] = bool(a)
        rr, loc = self._patch_local(pi)
        self._tri_elements(rr).act[j, loc] = bool(a)
        self._push_tri(rr)

    def _patch_active(self, p: str, r: str, vdep: bool) -> bool:
        pi = self._check_patch(p)
        j = self._sreac(r, pi, vdep=vdep)
        rr, loc = self._patch_local(pi)
        tel = self._tel[rr]
        return bool(tel.act[j, loc].any()) if tel is not None else bool(self._psact[pi, j])

    def getPatchSReacActive(self, p: str, r: str) -> bool:
        return self._patch_active(p, r, False)

    def setPatchSReacActive(self, p: str, r: str, a: bool):
        self._set_patch_active(p, r, a, False)

    def getPatchVDepSReacActive(self, p: str, r: str) -> bool:
        return self._patch_active(p, r, True)

    def setPatchVDepSReacActive(self, p: str, r: str, a: bool):
        self._set_patch_active(p, r, a, True)

    def _tri_v_of(self, r, loc):
        """Potential (V) of triangles loc of compartment r, as reactions read it."""
        if self.ef is not None and self._efreal == r:
            return self.ef.tris_v()[loc]
        return self.eng.get_tri_v(r)[loc]

    def _vdep_k(self, j: int, v: np.ndarray) -> np.ndarray:
        """k(V) by linear interpolation in the reaction's table
        (VDepSReacdef::getVDepK); outside the table is an error, as there."""
        t = int(self.fm.sreac_vdep[j])
        tab = self.fm.vdep_k[self.fm.vdep_off[t]:self.fm.vdep_off[t + 1]]
        x = (np.asarray(v, dtype=np.float64) - self.fm.vdep_vmin[t]) / self.fm.vdep_dv[t]
        if np.any(~((x >= 0.0) & (x <= len(tab) - 1))):
            raise N.StgError(N.STG_ERR_PROG, "triangle potential outside a voltage-dependent rate table")
        lo, hi = np.floor(x).astype(int), np.ceil(x).astype(int)
        fr = x - np.floor(x)
        return (1.0 - fr) * tab[lo] + fr * tab[hi]

    def _sreac_hc(self, pi: int, j: int, sel=None):
        """(h, c) per triangle of patch pi (or its triangles sel, positions in
        the patch) for surface reaction j at the present counts (SReac::rate,
        sreac.cpp:305-430; constants as SReac::resetCcst, sreac.cpp:80-100:
        scaled by the inner or outer tet's volume when volume species react,
        else by the triangle's area)."""
        fm = self.fm
        p = self.patches[pi]
        r, loc_all = self._patch_local(pi)
        pos = np.arange(len(p.tris)) if sel is None else np.asarray(sel)
        tris, loc = p.tris[pos], loc_all[pos]
        h = np.ones(len(pos))
        if len(fm.sspecs):
            lhs = fm.sreac_lhs_s[j]
            cnt = np.zeros((len(lhs), len(pos)), dtype=np.int64)
            for q in np.nonzero(lhs)[0]:
                cnt[q] = self.eng.get_tri_counts(int(q), real=r)[loc]
            h = h * _falling_h(cnt, lhs)
        tt = self.geom.tri_tets[tris]
        inner = np.where(self._tet_real[tt[:, 0]] == r, tt[:, 0], tt[:, 1])
        outer = np.where(tt[:, 0] == inner, tt[:, 1], tt[:, 0])
        side = None
        for lhs, tets, on in ((fm.sreac_lhs_i[j], inner, True), (fm.sreac_lhs_o[j], outer, False)):
            if not lhs.any():
                continue
            side = tets
            rc = self._tet_real[tets]
            cnt = np.zeros((len(lhs), len(pos)), dtype=np.int64)
            for s in np.nonzero(lhs)[0]:
                for cc in np.unique(rc):
                    m = rc == cc
                    cnt[s, m] = self.eng.batch_tet_counts(int(s), self._tet_local[tets[m]], real=int(cc))
            h = h * _falling_h(cnt, lhs)
        tel = self._tel[r]
        vd = fm.sreac_vdep[j] >= 0
        if tel is not None:
            k = tel.sreac_table()[j, loc]
        else:
            k = np.full(len(pos), (1.0 if vd else self._psk[pi, j]) if self._psact[pi, j] else 0.0)
        if vd:
            k = k * self._vdep_k(j, self._tri_v_of(r, loc))
        order = int(fm.sreac_lhs_s[j].sum() + fm.sreac_lhs_i[j].sum() + fm.sreac_lhs_o[j].sum())
        if side is None:
            scale = self.geom.tri_area[tris] * AVOGADRO
        else:
            scale = 1e3 * self.geom.tet_vol[side] * AVOGADRO
        return h, k * scale ** (1 - order)

    def getPatchSReacH(self, p: str, r: str) -> float:
        pi = self._check_patch(p)
        return float(self._sreac_hc(pi, self._sreac(r, pi))[0].sum())

    def getPatchSReacC(self, p: str, r: str) -> float:
        """Area-weighted mean of the triangles' c (_getPatchSReacC)."""
        pi = self._check_patch(p)
        _, c = self._sreac_hc(pi, self._sreac(r, pi))
        area = self.geom.tri_area[self.patches[pi].tris]
        return float((c * area).sum() / area.sum())

    def getPatchSReacA(self, p: str, r: str) -> float:
        pi = self._check_patch(p)
        h, c = self._sreac_hc(pi, self._sreac(r, pi))
        return float((h * c).sum())

    def _sdiff(self, d: str, pi: int | None = None) -> int:
        """Surface species of diffusion rule d; with pi, the rule must be one of
        patch pi's (Patchdef::surfdiffG2L)."""
        if d not in self.fm.sdiffs:
            raise _argerr(f"unknown surface diffusion rule {d!r}")
        if pi is not None and d not in self.fm.patch_sdiffs[pi]:
            raise _argerr(f"Diffusion rule {d!r} undefined in patch {self.patches[pi].name!r}.")
        return self.fm.sdiffs[d]

    def _tri_edge_nbr_patch(self, r: int, i: int, e: int) -> int:
        """Patch of the triangle across edge e of compartment r's triangle i, or -1."""
        mt = self.mts[r]
        if mt.tri_nbr is None or mt.tri_nbr[i, e] < 0:
            return -1
        rn = int(mt.tri_nbr_comp[i, e]) if mt.tri_nbr_comp is not None else r
        return int(self._rtri_patch[rn][mt.tri_nbr[i, e]])

    def getPatchSDiffD(self, p: str, d: str) -> float:
        """The rule's constant in the patch: its model value or the last
        setPatchSDiffD."""
        pi = self._check_patch(p)
        return float(self._psd[pi, self._sdiff(d, pi)])

    def setPatchSDiffD(self, p: str, d: str, dcst: float, direction_patch=None):
        """SDiff::setDcst on every triangle of the patch; with direction_patch,
        SDiff::setDirectionDcst on the patch's edges toward that patch's
        triangles (which opens surface diffusion boundary edges)."""
        pi = self._check_patch(p)
        q = self._sdiff(d, pi)
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        r, loc = self._patch_local(pi)
        if direction_patch in (None, ""):
            self._psd[pi, q] = dcst
            tel = self._tel[r]
            if (tel is not None and tel.sD is not None) or self._needs_sdiff_table(r):
                tel = self._sdiff_table_on(r)
                tel.set_sdcst(q, loc, dcst)
                self._push_sdiff(r)
            else:
                self.eng.set_sdcst(q, dcst, real=r)
            return
        dp = self._check_patch(direction_patch if isinstance(direction_patch, str) else direction_patch.name)
        tel = self._sdiff_table_on(r)
        for i in loc:
            for e in range(3):
                if self._tri_edge_nbr_patch(r, int(i), e) == dp:
                    tel.set_sdirection_dcst(q, int(i), e, float(dcst))
        self._push_sdiff(r)

    def getTriArea(self, tri: int) -> float:
        if not 0 <= tri < self.geom.countTris():
            raise _argerr(f"triangle {tri} out of range")
        return float(self.geom.tri_area[tri])

    def getTriSpecDefined(self, tri: int, s: str) -> bool:
        r, i, pi = self._tri(tri)
        self._sspec(s)
        return s in self.fm.patch_sspecs[pi]

    def getTriSpecCount(self, tri: int, s: str) -> float:
        r, i, pi, q = self._tri_sspec(tri, s)
        return float(self.eng.get_tri_counts(q, real=r)[i])

    def setTriSpecCount(self, tri: int, s: str, n: float):
        if n < 0:
            raise _argerr("negative number of molecules")
        r, i, pi, q = self._tri_sspec(tri, s)
        c = self.eng.get_tri_counts(q, real=r)
        c[i] = _round_count(n, self._host_rng)
        self.eng.set_tri_counts(q, c, real=r)

    def getTriSpecAmount(self, tri: int, s: str) -> float:
        return self.getTriSpecCount(tri, s) / AVOGADRO

    def setTriSpecAmount(self, tri: int, s: str, a: float):
        self.setTriSpecCount(tri, s, a * AVOGADRO)

    def getTriSpecClamped(self, tri: int, s: str) -> bool:
        r, i, pi, q = self._tri_sspec(tri, s)
        tel = self._tel[r]
        return bool(tel is not None and tel.clamped(q, i))

    def setTriSpecClamped(self, tri: int, s: str, buf: bool):
        """_setTriSpecClamped."""
        r, i, pi, q = self._tri_sspec(tri, s)
        self._tri_elements(r).set_clamped(q, i, bool(buf))
        self._push_tri(r)
        self._push_tri_clamp(r)

    def getTriSReacK(self, tri: int, r: str) -> float:
        rr, i, pi = self._tri(tri)
        j = self._sreac(r, pi, vdep=False)
        tel = self._tel[rr]
        return float(tel.k[j, i]) if tel is not None else float(self._psk[pi, j])

    def setTriSReacK(self, tri: int, r: str, kf: float):
        """_setTriSReacK (tetopsplit.cpp:4464)."""
        rr, i, pi = self._tri(tri)
        j = self._sreac(r, pi, vdep=False)
        if kf < 0:
            raise _argerr("rate constant must be >= 0")
        self._tri_elements(rr).k[j, i] = kf
        self._push_tri(rr)

    def _tri_active(self, tri: int, r: str, vdep: bool) -> bool:
        rr, i, pi = self._tri(tri)
        j = self._sreac(r, pi, vdep=vdep)
        tel = self._tel[rr]
        return bool(tel.act[j, i]) if tel is not None else bool(self._psact[pi, j])

    def _set_tri_active(self, tri: int, r: str, act: bool, vdep: bool):
        rr, i, pi = self._tri(tri)
        j = self._sreac(r, pi, vdep=vdep)
        self._tri_elements(rr).act[j, i] = bool(act)
        self._push_tri(rr)

    def getTriSReacActive(self, tri: int, r: str) -> bool:
        return self._tri_active(tri, r, False)

    def setTriSReacActive(self, tri: int, r: str, act: bool):
        """_setTriSReacActive."""
        self._set_tri_active(tri, r, act, False)

    def getTriVDepSReacActive(self, tri: int, r: str) -> bool:
        return self._tri_active(tri, r, True)

    def setTriVDepSReacActive(self, tri: int, r: str, act: bool):
        """_setTriVDepSReacActive."""
        self._set_tri_active(tri, r, act, True)

    def getTriVDepSReacK(self, tri: int, r: str) -> float:
        """k(V) at the triangle's present potential (_getTriVDepSReacK)."""
        rr, i, pi = self._tri(tri)
        j = self._sreac(r, pi, vdep=True)
        return float(self._vdep_k(j, self._tri_v_of(rr, np.array([i])))[0])

    def _tri_sreac_hc(self, tri: int, r: str):
        rr, i, pi = self._tri(tri)
        j = self._sreac(r, pi)
        pos = int(np.nonzero(self.patches[pi].tris == tri)[0][0])
        h, c = self._sreac_hc(pi, j, [pos])
        return float(h[0]), float(c[0])

    def getTriSReacH(self, tri: int, r: str) -> float:
        return self._tri_sreac_hc(tri, r)[0]

    def getTriSReacC(self, tri: int, r: str) -> float:
        return self._tri_sreac_hc(tri, r)[1]

    def getTriSReacA(self, tri: int, r: str) -> float:
        h, c = self._tri_sreac_hc(tri, r)
        return h * c

    def _tri_edge_toward(self, tri: int, other: int) -> int:
        """Edge slot of triangle tri toward patch triangle other, the two
        sharing a bar (Tri::getTriDirection over the neighbours found through
        the bars, tetopsplit.cpp:575-600)."""
        self._tri(tri)
        g = self.geom
        ok = 0 <= other < g.countTris() and other != tri and self._tri_patch[other] >= 0
        shared = set(g.tri_bars[tri].tolist()) & set(g.tri_bars[other].tolist()) if ok else set()
        if not shared:
            raise _argerr(f"Triangle {other} is not a neighbor of triangle {tri}.")
        return g.tri_edge_of_bar(tri, shared.pop())

    def getTriSDiffD(self, tri: int, d: str, direction_tri=None) -> float:
        """_getTriSDiffD (tetopsplit.cpp:4534-4566): the rule's constant on the
        triangle, or the directional constant toward direction_tri if one is
        set (SDiff::dcst)."""
        rr, i, pi = self._tri(tri)
        q = self._sdiff(d, pi)
        e = None if direction_tri is None or direction_tri < 0 else self._tri_edge_toward(tri, int(direction_tri))
        tel = self._tel[rr]
        if tel is None or tel.sD is None:
            return float(self._psd[pi, q])
        return tel.sdcst(q, i, e)

    def setTriSDiffD(self, tri: int, d: str, dcst: float, direction_tri=None):
        """_setTriSDiffD (tetopsplit.cpp:4570-4600): the rule's constant on one
        triangle (directional constants cleared), or through the edge toward
        direction_tri (a boundary edge opens)."""
        rr, i, pi = self._tri(tri)
        q = self._sdiff(d, pi)
        if dcst < 0:
            raise _argerr("diffusion constant must be >= 0")
        e = None if direction_tri is None or direction_tri < 0 else self._tri_edge_toward(tri, int(direction_tri))
        tel = self._sdiff_table_on(rr)
        if e is None:
            tel.set_sdcst(q, [i], float(dcst))
        else:
            tel.set_sdirection_dcst(q, i, e, float(dcst))
        self._push_sdiff(rr)

    # batch access to triangles
    def _tri_idx(self, index_array):
        idx = np.asarray(index_array, dtype=np.int64)
        for t in idx:
            self._tri(int(t))
        return idx

    def getBatchTriSpecCounts(self, tris, s: str):
        c = np.zeros(len(tris))
        self.getBatchTriSpecCountsNP(np.asarray(tris, dtype=np.int64), s, c)
        return c.tolist()

    def getBatchTriSpecCountsNP(self, index_array, s: str, counts):
        idx = self._tri_idx(index_array)
        if len(counts) != len(idx):
            raise _argerr("counts and index arrays differ in length")
        q = self._sspec(s)
        real, loc = self._tri_real[idx], self._tri_local[idx]
        for r in np.unique(real):
            m = real == r
            counts[m] = self.eng.get_tri_counts(q, real=int(r))[loc[m]]

    def setBatchTriSpecCountsNP(self, index_array, s: str, counts):
        idx = self._tri_idx(index_array)
        vals = np.asarray(counts, dtype=np.float64)
        if len(vals) != len(idx):
            raise _argerr("counts and index arrays differ in length")
        if np.any(vals < 0):
            raise _argerr("negative number of molecules")
        q = self._sspec(s)
        real, loc = self._tri_real[idx], self._tri_local[idx]
        fl = np.floor(vals)
        rounded = (fl + (self._host_rng.random(len(vals)) < vals - fl)).astype(np.int64)
        for r in np.unique(real):
            m = real == r
            c = self.eng.get_tri_counts(q, real=int(r))
            c[loc[m]] = rounded[m]
            self.eng.set_tri_counts(q, c, real=int(r))

    def setBatchTriSpecCounts(self, tris, s: str, counts):
        self.setBatchTriSpecCountsNP(np.asarray(tris, dtype=np.int64), s, np.asarray(counts, dtype=float))

    def sumBatchTriSpecCountsNP(self, tri_array, s: str) -> float:
        c = np.zeros(len(tri_array))
        self.getBatchTriSpecCountsNP(tri_array, s, c)
        return float(c.sum())

    sumBatchTriCountsNP = sumBatchTriSpecCountsNP

    # --------------------------------------------- membrane potential (EField)
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
        self._memb_res = (float(ro), float(vrev))

    def getMembRes(self, m: str):
        """(resistivity, reversal potential) of the membrane's leak
        (TetOpSplitP::_getMembRes, tetopsplit.cpp:5305)."""
        self._memb(m)
        return self._memb_res

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
        """Potential (V) of a membrane triangle, the mean of its vertices
        (EField::getTriV, efield.cpp:290-302). Without an EField, the value
        voltage-dependent reactions read."""
        if self.ef is None:
            r, i, _ = self._tri(tri)
            return float(self.eng.get_tri_v(r)[i])
        return float(self.ef.tris_v()[self._eftri(tri)])

    def setTriV(self, tri: int, v: float):
        """Set the potential of a membrane triangle: with an EField, of its three
        vertices (EField::setTriV, efield.cpp:306-317; TetOpSplitP::_setTriV,
        tetopsplit.cpp:4871-4893). Without one, only the value voltage-dependent
        reactions read (no counterpart in the reference, which raises)."""
        if self.ef is None:
            r, i, _ = self._tri(tri)
            vv = self.eng.get_tri_v(r)
            vv[i] = v
            self.eng.set_tri_v(vv, r)
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

    def setTriCapac(self, tri: int, cm: float):
        """_setTriCapac (tetopsplit.cpp:5187): specific capacitance (F/m^2) of
        one membrane triangle."""
        k = self._eftri(tri)
        if cm < 0.0:
            raise _argerr("Capacitance must be non-negative")
        self.ef.set_tri_capac(k, cm)

    def _ohmic(self, oc: str) -> int:
        try:
            return self.fm.ohmic.index(oc)
        except ValueError:
            raise _argerr(f"unknown ohmic current {oc!r}") from None

    def getTriOhmicErev(self, tri: int, oc: str) -> float:
        """_getTriOhmicErev (tetopsplit.cpp:4963)."""
        k = self._eftri(tri)
        return self.ef.tri_ohmic_erev(self._ohmic(oc), k)

    def setTriOhmicErev(self, tri: int, oc: str, erev: float):
        """_setTriOhmicErev (tetopsplit.cpp:4933, Tri::setOCerev)."""
        k = self._eftri(tri)
        self.ef.set_tri_ohmic_erev(self._ohmic(oc), k, erev)

    def getTriOhmicI(self, tri: int, oc: str = "") -> float:
        """Present ohmic current (A) of one current on a triangle, or of all of
        them: open channels times g times (V - erev) (Tri::getOhmicI,
        tri.cpp:697-718)."""
        k = self._eftri(tri)
        r, i, _ = self._tri(tri)
        v = self.ef.tris_v()[k]
        js = [self._ohmic(oc)] if oc else range(len(self.fm.ohmic))
        tot = 0.0
        for j in js:
            n = float(self.eng.get_tri_counts(int(self.fm.ohmic_chan[j]), real=r)[i])
            tot += (n * float(self.fm.ohmic_g[j])) * (v - self.ef.tri_ohmic_erev(j, k))
        return float(tot)

    def _ghk(self, ghk: str) -> int:
        try:
            return self.fm.ghk.index(ghk)
        except ValueError:
            raise _argerr(f"unknown GHK current {ghk!r}") from None

    def getTriGHKI(self, tri: int, ghk: str = "") -> float:
        """GHK current (A, outward positive) of one current on a triangle, or
        of all of them: the charge its ions carried over the last EField step
        over the step (Tri::getGHKI, tri.cpp:634-660)."""
        k = self._eftri(tri)
        g = -1
        if ghk:
            g = self._ghk(ghk)
            if g not in self.fm.patch_ghk[self._tri(tri)[2]]:
                raise _argerr("GHK current undefined in triangle.")
        return float(self.ef.tri_ghk_currents(g)[k])

    def getTriI(self, tri: int) -> float:
        """_getTriI (tetopsplit.cpp:5101): GHK plus ohmic current."""
        return self.getTriGHKI(tri) + self.getTriOhmicI(tri)

    def _batch_ef(self, index_array, n_out, counts):
        """EField positions of the listed triangles (-1 for triangles outside
        every patch, which read 0 as in the reference)."""
        self._need_ef()
        idx = np.asarray(index_array, dtype=np.int64)
        if len(counts) != n_out:
            raise _argerr("Error: output array (counts) size does not match the input")
        if np.any(idx < 0) or np.any(idx >= self.geom.countTris()):
            raise _argerr("Error (Index Overbound): there is no such triangle")
        pos = np.where(self._tri_local[idx] >= 0, self._ef_tri_local[idx], -1)
        if np.any((self._tri_local[idx] >= 0) & (pos < 0)):
            raise _argerr("a triangle is not assigned to the membrane")
        return idx, pos

    def getBatchTriGHKIsNP(self, index_array, ghk: str, counts):
        """TetOpSplitP::getBatchTriGHKIsNP (tetopsplit.cpp:6244-6305)."""
        idx, pos = self._batch_ef(index_array, len(index_array), counts)
        g = self._ghk(ghk)
        cur = self.ef.tri_ghk_currents(g)
        held = np.array([pos[k] >= 0 and g in self.fm.patch_ghk[self._tri_patch[idx[k]]]
                         for k in range(len(idx))], dtype=bool)
        counts[:] = np.where(held, cur[np.maximum(pos, 0)], 0.0)

    def sumBatchTriGHKIsNP(self, tri_array, ghk: str) -> float:
        c = np.zeros(len(tri_array))
        self.getBatchTriGHKIsNP(tri_array, ghk, c)
        return float(c.sum())

    def getBatchTriBatchGHKIsNP(self, index_array, ghks, counts):
        """counts[t * len(ghks) + k] = current ghks[k] on triangle index_array[t]
        (TetOpSplitP::getBatchTriBatchGHKIsNP, tetopsplit.cpp:6481-6550)."""
        n = len(ghks)
        idx, pos = self._batch_ef(index_array, len(index_array) * n, counts)
        for k, ghk in enumerate(ghks):
            col = np.zeros(len(idx))
            self.getBatchTriGHKIsNP(idx, ghk, col)
            counts[k::n] = col

    def getBatchTriOhmicIsNP(self, index_array, oc: str, counts):
        idx, pos = self._batch_ef(index_array, len(index_array), counts)
        j = self._ohmic(oc)
        for k, t in enumerate(idx):
            counts[k] = self.getTriOhmicI(int(t), oc) if pos[k] >= 0 else 0.0

    def sumBatchTriOhmicIsNP(self, tri_array, oc: str) -> float:
        c = np.zeros(len(tri_array))
        self.getBatchTriOhmicIsNP(tri_array, oc, c)
        return float(c.sum())

    def getBatchTriBatchOhmicIsNP(self, index_array, ocs, counts):
        """counts[t * len(ocs) + k] = current ocs[k] on triangle index_array[t]
        (TetOpSplitP::getBatchTriBatchOhmicIsNP, tetopsplit.cpp:6406-6478)."""
        n = len(ocs)
        self._batch_ef(index_array, len(index_array) * n, counts)
        for k, oc in enumerate(ocs):
            col = np.zeros(len(index_array))
            self.getBatchTriOhmicIsNP(index_array, oc, col)
            counts[k::n] = col

    def getBatchTriVsNP(self, index_array, vs):
        if self.ef is None:
            for k, t in enumerate(index_array):
                vs[k] = self.getTriV(int(t))
        else:
            vv = self.ef.tris_v()
            vs[:] = vv[[self._eftri(int(t)) for t in index_array]]

    def getBatchTetVsNP(self, index_array, vs):
        self._need_ef()
        vv = self.ef.verts_v()
        for k, t in enumerate(index_array):
            vs[k] = vv[[self._vert(int(x)) for x in self.geom.tets[int(t)]]].mean()

    def getBatchVertVsNP(self, index_array, vs):
        self._need_ef()
        vv = self.ef.verts_v()
        vs[:] = vv[[self._vert(int(v)) for v in index_array]]

    def setBatchTriVsNP(self, index_array, vs):
        for t, v in zip(index_array, vs):
            self.setTriV(int(t), float(v))

    # ----------------------- checkpointing (stg_checkpoint / stg_efield_checkpoint)
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

    # --------------------------------------------------- deprecated spellings
    # The reference keeps these older names beside the *Spec* methods
    # (cysteps_solver.pyx, marked DEPRECATED there).
    getCompCount, setCompCount = getCompSpecCount, setCompSpecCount
    getCompAmount, setCompAmount = getCompSpecAmount, setCompSpecAmount
    getCompConc, setCompConc = getCompSpecConc, setCompSpecConc
    getCompClamped, setCompClamped = getCompSpecClamped, setCompSpecClamped
    getPatchCount, setPatchCount = getPatchSpecCount, setPatchSpecCount
    getPatchAmount, setPatchAmount = getPatchSpecAmount, setPatchSpecAmount
    getPatchClamped, setPatchClamped = getPatchSpecClamped, setPatchSpecClamped
    getTetCount, setTetCount = getTetSpecCount, setTetSpecCount
    getTetAmount, setTetAmount = getTetSpecAmount, setTetSpecAmount
    getTetConc, setTetConc = getTetSpecConc, setTetSpecConc
    getTetClamped, setTetClamped = getTetSpecClamped, setTetSpecClamped
    getTriCount, setTriCount = getTriSpecCount, setTriSpecCount
    getTriAmount, setTriAmount = getTriSpecAmount, setTriSpecAmount
    getTriClamped, setTriClamped = getTriSpecClamped, setTriSpecClamped
    getBatchTetCounts, getBatchTetCountsNP = getBatchTetSpecCounts, getBatchTetSpecCountsNP
    setBatchTetCounts, setBatchTetCountsNP = setBatchTetSpecCounts, setBatchTetSpecCountsNP
    getBatchTetConcs, getBatchTetConcsNP = getBatchTetSpecConcs, getBatchTetSpecConcsNP
    setBatchTetConcs, setBatchTetConcsNP = setBatchTetSpecConcs, setBatchTetSpecConcsNP
    getBatchTriCounts, getBatchTriCountsNP = getBatchTriSpecCounts, getBatchTriSpecCountsNP
    setBatchTriCounts, setBatchTriCountsNP = setBatchTriSpecCounts, setBatchTriSpecCountsNP


# Reference methods this engine does not provide, each with the reason. They
# raise NotImplErr, as the reference's API base does for methods a solver
# lacks (src/steps/solver/api_*.cpp: NotImplErrLog).
NOT_PROVIDED = {
    # complexes (multi-state species) belong to the reference's newer solvers only
    **{m: "complexes are not part of the operator-split path" for m in (
        "getCompComplexAmount", "getCompComplexConc", "getCompComplexCount", "getCompComplexReacExtent",
        "getCompComplexSUSAmount", "getCompComplexSUSConc", "getCompComplexSUSCount",
        "setCompComplexAmount", "setCompComplexConc", "setCompComplexCount",
        "getPatchComplexAmount", "getPatchComplexCount", "getPatchComplexSReacExtent",
        "getPatchComplexSUSCount", "setPatchComplexAmount", "setPatchComplexCount")},
    # geometry is fixed by the mesh (TetOpSplitP::_setCompVol and
    # _setPatchArea throw NotImplErr too)
    **{m: "volumes and areas are fixed by the mesh" for m in ("setCompVol", "setPatchArea")},
    # per-reaction extents: the engine counts events per realization
    **{m: "per-reaction extents are not counted; getReacExtent gives the total" for m in (
        "getCompReacExtent", "resetCompReacExtent", "getPatchSReacExtent", "resetPatchSReacExtent")},
    # ROI extents (TetOpSplitP throws NotImplErr for these, tetopsplit.cpp:7605-7879)
    **{m: "ROI extents are not implemented for this solver" for m in (
        "getROIReacExtent", "resetROIReacExtent", "getROISReacExtent", "resetROISReacExtent",
        "getROIDiffExtent", "resetROIDiffExtent")},
    # the whole-state total propensity (API::getA0; TetOpSplitP does not provide it)
    "getA0": "the total propensity of the state is not provided by operator splitting",
    # DistTetOpSplit's Gibson-Bruck dependency graph and its mesh-bandwidth file
    "dumpDepGraphToFile": "there is no global dependency graph: each tet runs its own SSA",
    "saveMembOpt": "the GPU EField is a sparse CG solve: there is no bandwidth ordering to save",
}


def _make_notimpl(name, why):
    def method(self, *args, **kwargs):
        raise _notimpl(f"{name}: {why}")
    method.__name__ = name
    method.__doc__ = f"Not provided by this engine ({why}); raises NotImplErr."
    return method


for _n, _why in NOT_PROVIDED.items():
    if not hasattr(TetOpSplit, _n):
        setattr(TetOpSplit, _n, _make_notimpl(_n, _why))


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
