"""Kinetic model objects with the reference's legacy (API_1) constructors.

``Model``, ``Spec``, ``Volsys``, ``Surfsys``, ``Reac``, ``SReac``, ``VDepSReac``,
``Diff``, ``Chan``, ``ChanState``, ``OhmicCurr`` and ``GHKcurr`` take the same
arguments as steps.model (pysteps/steps/API_1/model.py, pysteps/cysteps_model.pyx,
backed by src/steps/model/*.cpp). ``flatten`` turns a model bound to one compartment and
at most one patch into the stoichiometry tables the engine reads, the role of
Statedef/Compdef/Patchdef (src/steps/solver/compdef.hpp:252-258, 420).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import math

import numpy as np

FARADAY = 96485.3399        # src/steps/math/constants.hpp:49
GAS_CONSTANT = 8.314472     # src/steps/math/constants.hpp:53


def _argerr(msg):
    from ._native import STG_ERR_ARG, ArgErr
    return ArgErr(STG_ERR_ARG, msg)


def _notimpl(msg):
    from ._native import STG_ERR_NOTIMPL, NotImplErr
    return NotImplErr(STG_ERR_NOTIMPL, msg)


class Model:
    def __init__(self):
        self.specs: dict[str, Spec] = {}
        self.volsys: dict[str, Volsys] = {}
        self.surfsys: dict[str, Surfsys] = {}

    def getSpec(self, name: str) -> "Spec":
        return self.specs[name]

    def getAllSpecs(self):
        return list(self.specs.values())


class Spec:
    def __init__(self, id: str, model: Model, valence: int = 0):
        if id in model.specs:
            raise ValueError(f"species {id!r} already defined")
        self.id = id
        self.model = model
        self.valence = valence
        model.specs[id] = self

    def getID(self) -> str:
        return self.id

    def __repr__(self):
        return f"Spec({self.id})"


class Volsys:
    def __init__(self, id: str, model: Model):
        self.id = id
        self.model = model
        self.reacs: dict[str, Reac] = {}
        self.diffs: dict[str, Diff] = {}
        model.volsys[id] = self


class Chan:
    """A channel: the parent of its channel states (cysteps_model.pyx:500-520)."""

    def __init__(self, id: str, model: Model):
        self.id = id
        self.model = model
        self.states: list[ChanState] = []


class ChanState(Spec):
    """A state of a channel; a surface species (cysteps_model.pyx:638-655)."""

    def __init__(self, id: str, model: Model, chan: Chan):
        super().__init__(id, model)
        self.chan = chan
        chan.states.append(self)

    def getChan(self) -> Chan:
        return self.chan


class Surfsys:
    def __init__(self, id: str, model: Model):
        self.id = id
        self.model = model
        self.sreacs: dict[str, SReac] = {}
        self.diffs: dict[str, Diff] = {}
        self.ohmiccurrs: dict[str, OhmicCurr] = {}
        self.ghkcurrs: dict[str, GHKcurr] = {}
        model.surfsys[id] = self


class Reac:
    """Volume reaction lhs -> rhs with macroscopic constant kcst (M^(1-order)/s)."""

    def __init__(self, id: str, volsys: Volsys, lhs=(), rhs=(), kcst: float = 0.0):
        self.id = id
        self.volsys = volsys
        self.lhs = list(lhs)
        self.rhs = list(rhs)
        self.kcst = float(kcst)
        if len(self.lhs) > 4:
            raise ValueError("reaction order above 4 is not supported")
        volsys.reacs[id] = self

    def getOrder(self) -> int:
        return len(self.lhs)

    def setKcst(self, k: float) -> None:
        self.kcst = float(k)


class SReac:
    """Surface reaction over patch species (slhs/srhs), inner-compartment
    species (ilhs/irhs) and outer-compartment species (olhs/orhs); volume
    reactants on one side only (src/steps/model/sreac.cpp:55). On a patch
    with an outer compartment the triangle's inner and outer tets react in one
    SSA group."""

    def __init__(self, id: str, surfsys: Surfsys, ilhs=(), slhs=(), olhs=(), irhs=(), srhs=(),
                 orhs=(), kcst: float = 0.0):
        self.id = id
        self.surfsys = surfsys
        self.ilhs, self.slhs, self.olhs = list(ilhs), list(slhs), list(olhs)
        self.irhs, self.srhs, self.orhs = list(irhs), list(srhs), list(orhs)
        self.kcst = float(kcst)
        if self.ilhs and self.olhs:
            raise ValueError("Volume lhs species must belong to either inner or outer compartment, "
                             "not both.")
        if len(self.ilhs) + len(self.slhs) + len(self.olhs) > 4:
            raise ValueError("reaction order above 4 is not supported")
        surfsys.sreacs[id] = self

    def reqOutside(self) -> bool:
        """True when the reaction reads or changes outer-compartment species."""
        return bool(self.olhs or self.orhs)

    def getOrder(self) -> int:
        return len(self.ilhs) + len(self.slhs) + len(self.olhs)


class VDepSReac(SReac):
    """Voltage-dependent surface reaction: the rate constant is k(V) of the
    triangle's membrane potential, tabulated over vrange = [vmin, vmax, dv] and
    linearly interpolated. The table holds int((vmax-vmin)/dv)+1 values of k at
    V = vmin, vmin+dv, ... with V accumulated by repeated addition, as
    _VoltageTable builds it (pysteps/steps/API_1/model.py:496-521); lookup is
    VDepSReacdef::getVDepK (src/steps/solver/vdepsreacdef.cpp:194-216)."""

    def __init__(self, id: str, surfsys: Surfsys, ilhs=(), slhs=(), olhs=(), irhs=(), srhs=(),
                 orhs=(), k=None, vrange=(-150.0e-3, 100.0e-3, 1.0e-4)):
        if k is None or not callable(k):
            raise ValueError("VDepSReac needs a callable k(V)")
        vmin, vmax, dv = (float(x) for x in vrange)
        if not (vmax > vmin and dv > 0):
            raise ValueError("vrange must be [vmin, vmax, dv] with vmax > vmin and dv > 0")
        n = int((vmax - vmin) / dv) + 1
        self.vmin, self.vmax, self.dv = vmin, vmax, dv
        ktab, v = [], vmin
        for _ in range(n):
            ktab.append(float(k(v)))
            v += dv
        self.ktab = np.array(ktab)
        if np.any(self.ktab < 0) or not np.all(np.isfinite(self.ktab)):
            raise ValueError("k(V) must be finite and >= 0 over vrange")
        super().__init__(id, surfsys, ilhs, slhs, olhs, irhs, srhs, orhs, kcst=0.0)


class OhmicCurr:
    """Ohmic current I = g (V - erev) per channel in state chanstate
    (src/steps/model/ohmiccurr.cpp; cysteps_model.pyx:3115-3141)."""

    def __init__(self, id: str, surfsys: Surfsys, chanstate: Spec, erev: float, g: float):
        if g < 0:
            raise ValueError("conductance must be >= 0")
        self.id = id
        self.surfsys = surfsys
        self.chanstate = chanstate
        self.erev = float(erev)
        self.g = float(g)
        surfsys.ohmiccurrs[id] = self

    def getERev(self) -> float:
        return self.erev

    def setERev(self, erev: float) -> None:
        self.erev = float(erev)

    def getG(self) -> float:
        return self.g

    def setG(self, g: float) -> None:
        if g < 0:
            raise ValueError("conductance must be >= 0")
        self.g = float(g)


def ghk_permeability(g, v, z, temp, iconc, oconc):
    """Single-channel permeability (m^3/s) giving conductance g (S) at
    potential v (V), temperature temp (K) and concentrations in M: the
    inversion of the GHK flux equation's slope, math::permeability
    (src/steps/math/ghk.cpp:34-60)."""
    b, c = iconc * 1000.0, oconc * 1000.0
    d = ((0.0 - z) * FARADAY) / (GAS_CONSTANT * temp)
    den = c * math.exp(2.0 * d * v) + ((b - c) * d * v - c - b) * math.exp(d * v) + b
    num = math.exp(2.0 * d * v) - 2.0 * math.exp(d * v) + 1
    eps = np.finfo(float).eps
    a = g * (num / den) if (abs(num) > eps or abs(den) > eps) else 2.0 * g / (c + b)
    p = (a * GAS_CONSTANT * temp) / (z ** 2.0 * FARADAY ** 2.0)
    if math.isnan(p):
        raise ValueError("Failed to find permeability, check parameters")
    return p


class GHKcurr:
    """GHK current of ``ion`` through channels in state chanstate
    (src/steps/model/ghkcurr.cpp:41-170; cysteps_model.pyx:3336-3562). The
    permeability comes from setP or setPInfo. computeflux=False moves charge
    but no ions. virtual_oconc (M) stands in for the outer compartment; when
    it is negative (the default) the patch's outer compartment is read and
    exchanges the ions."""

    def __init__(self, id: str, surfsys: Surfsys, chanstate: Spec, ion: Spec,
                 computeflux: bool = True, virtual_oconc: float = -1.0, vshift: float = 0.0):
        if ion.valence == 0:
            raise ValueError("Ion provided to GHKcurr initializer function has valence zero")
        self.id = id
        self.surfsys = surfsys
        self.chanstate = chanstate
        self.ion = ion
        self.computeflux = bool(computeflux)
        self.virtual_oconc = float(virtual_oconc)
        self.vshift = float(vshift)
        self.P = 0.0
        surfsys.ghkcurrs[id] = self

    def setP(self, p: float) -> None:
        if not p > 0.0:
            raise ValueError("Permeability provided to GHKcurr::setP function can't be "
                             "negative or zero")
        self.P = float(p)

    def getP(self) -> float:
        return self.P

    def setPInfo(self, g: float, V: float, T: float, oconc: float, iconc: float) -> None:
        if not g > 0.0:
            raise ValueError("Conductance provided to GHKcurr::setPInfo function can't be "
                             "negative or zero")
        if V == 0.0:
            raise ValueError("Potential provided to GHKcurr::setPInfo function can't be zero.")
        if T < 0.0 or oconc < 0.0 or iconc < 0.0:
            raise ValueError("temperature and concentrations must be >= 0")
        self.P = ghk_permeability(g, V, self.ion.valence, T, iconc, oconc)


class Diff:
    """Diffusion rule of ``lig`` with constant dcst (m^2/s)."""

    def __init__(self, id: str, sys, lig: Spec, dcst: float = 0.0):
        self.id = id
        self.sys = sys
        self.lig = lig
        self.dcst = float(dcst)
        sys.diffs[id] = self


@dataclass
class FlatModel:
    """Engine tables (stg_model in include/stg.h) plus name maps."""

    specs: list[str]
    sspecs: list[str]
    reacs: list[str]
    sreacs: list[str]
    diffs: dict[str, int] = field(default_factory=dict)   # diff id -> species index
    reac_lhs: np.ndarray = None
    reac_upd: np.ndarray = None
    reac_k: np.ndarray = None
    spec_dcst: np.ndarray = None
    sreac_lhs_s: np.ndarray = None
    sreac_upd_s: np.ndarray = None
    sreac_lhs_i: np.ndarray = None
    sreac_upd_i: np.ndarray = None
    sreac_k: np.ndarray = None
    sreac_vdep: np.ndarray = None     # [n_sreacs] k(V) table index, -1 = constant k
    vdep_vmin: np.ndarray = None
    vdep_dv: np.ndarray = None
    vdep_off: np.ndarray = None
    vdep_k: np.ndarray = None
    ghk: list = field(default_factory=list)          # GHKcurr ids
    ghk_chan: np.ndarray = None
    ghk_ion: np.ndarray = None
    ghk_valence: np.ndarray = None
    ghk_perm: np.ndarray = None
    ghk_voconc: np.ndarray = None
    ghk_vshift: np.ndarray = None
    ghk_realflux: np.ndarray = None
    ohmic: list = field(default_factory=list)        # OhmicCurr ids
    ohmic_chan: np.ndarray = None                    # surface species index
    ohmic_g: np.ndarray = None                       # S per channel
    ohmic_erev: np.ndarray = None                    # V
    sspec_dcst: np.ndarray = None                    # [n_sspecs] surface diffusion m^2/s
    sdiffs: dict = field(default_factory=dict)       # surface diff id -> surface species index
    sreac_lhs_o: np.ndarray = None                   # [n_sreacs][S] outer-compartment reactants
    sreac_upd_o: np.ndarray = None                   # [n_sreacs][S] outer-compartment changes
    # several patches (flatten_comps): per patch the surface reactions, GHK
    # currents and surface species it holds (index sets), its compartments
    patch_sreacs: list = field(default_factory=list)
    patch_ghk: list = field(default_factory=list)
    patch_sspecs: list = field(default_factory=list)
    # per patch its surface diffusion rules: rule id -> (surface species, constant)
    patch_sdiffs: list = field(default_factory=list)

    def outer(self) -> bool:
        """The model reaches the outer side of a membrane."""
        return bool((self.sreac_lhs_o is not None and np.any(self.sreac_lhs_o))
                    or (self.sreac_upd_o is not None and np.any(self.sreac_upd_o))
                    or (self.ghk_voconc is not None and np.any(self.ghk_voconc < 0)))

    def spec_index(self, name: str) -> int:
        try:
            return self.specs.index(name)
        except ValueError:
            raise KeyError(f"species {name!r} is not defined in the compartment") from None

    def sspec_index(self, name: str) -> int:
        try:
            return self.sspecs.index(name)
        except ValueError:
            raise KeyError(f"species {name!r} is not defined in the patch") from None


def _count(lst, names):
    v = np.zeros(len(names), dtype=np.int32)
    for s in lst:
        v[names.index(s.id)] += 1
    return v


def _patch_sdiffs(ssys, sspecs) -> dict:
    """Surface diffusion rules of one patch (its surface systems): rule id ->
    (surface species index, constant). Each rule is one SDiff kproc per
    triangle (Patchdef, src/steps/solver/patchdef.cpp); one species with two
    rules in a patch would diffuse at the sum of their rates, which this
    engine does not model."""
    out, seen = {}, set()
    for ss in ssys:
        for d in ss.diffs.values():
            q = sspecs.index(d.lig.id)
            if q in seen:
                raise _notimpl(f"species {d.lig.id!r} has more than one surface diffusion rule in a patch")
            seen.add(q)
            out[d.id] = (q, float(d.dcst))
    return out


def flatten(model: Model, volsys_ids, surfsys_ids=()) -> FlatModel:
    """Species, reactions and diffusion constants of one compartment (its volume
    systems) and one patch (its surface systems). Species are indexed in model
    definition order, as specG2L orders them (src/steps/solver/compdef.cpp)."""
    vsys = [model.volsys[v] for v in volsys_ids]
    ssys = [model.surfsys[s] for s in surfsys_ids]
    used_v, used_s = set(), set()
    for vs in vsys:
        for r in vs.reacs.values():
            used_v.update(s.id for s in r.lhs + r.rhs)
        for d in vs.diffs.values():
            used_v.add(d.lig.id)
    for ss in ssys:
        for r in ss.sreacs.values():
            if r.olhs or r.orhs:
                raise _argerr(f"surface reaction {r.id!r} involves the outer compartment, but the "
                              "patch has none")
            used_v.update(s.id for s in r.ilhs + r.irhs)
            used_s.update(s.id for s in r.slhs + r.srhs)
        for c in ss.ohmiccurrs.values():
            used_s.add(c.chanstate.id)
        for c in ss.ghkcurrs.values():
            if c.virtual_oconc < 0:
                raise _argerr(f"GHK current {c.id!r} reads the outer compartment, but the patch has "
                              "none; give it a virtual_oconc >= 0")
            if not c.P > 0.0:
                raise ValueError(f"Permeability not defined for GHK current object {c.id!r}")
            used_s.add(c.chanstate.id)
            used_v.add(c.ion.id)
        for d in ss.diffs.values():
            used_s.add(d.lig.id)
    order = list(model.specs)
    specs = [s for s in order if s in used_v]
    sspecs = [s for s in order if s in used_s]
    reacs = [r for vs in vsys for r in vs.reacs.values()]
    sreacs = [r for ss in ssys for r in ss.sreacs.values()]
    S, SS = len(specs), len(sspecs)
    fm = FlatModel(specs=specs, sspecs=sspecs, reacs=[r.id for r in reacs],
                   sreacs=[r.id for r in sreacs])
    fm.reac_lhs = np.zeros((len(reacs), S), dtype=np.int32)
    fm.reac_upd = np.zeros((len(reacs), S), dtype=np.int32)
    for i, r in enumerate(reacs):
        lhs = _count(r.lhs, specs)
        fm.reac_lhs[i] = lhs
        fm.reac_upd[i] = _count(r.rhs, specs) - lhs
    fm.reac_k = np.array([r.kcst for r in reacs], dtype=np.float64)
    fm.spec_dcst = np.zeros(S)
    for vs in vsys:
        for d in vs.diffs.values():
            fm.spec_dcst[specs.index(d.lig.id)] = d.dcst
            fm.diffs[d.id] = specs.index(d.lig.id)
    n = len(sreacs)
    fm.sreac_lhs_s = np.zeros((n, SS), dtype=np.int32)
    fm.sreac_upd_s = np.zeros((n, SS), dtype=np.int32)
    fm.sreac_lhs_i = np.zeros((n, S), dtype=np.int32)
    fm.sreac_upd_i = np.zeros((n, S), dtype=np.int32)
    for i, r in enumerate(sreacs):
        ls = _count(r.slhs, sspecs)
        li = _count(r.ilhs, specs)
        fm.sreac_lhs_s[i] = ls
        fm.sreac_lhs_i[i] = li
        fm.sreac_upd_s[i] = _count(r.srhs, sspecs) - ls
        fm.sreac_upd_i[i] = _count(r.irhs, specs) - li
    fm.sreac_k = np.array([r.kcst for r in sreacs], dtype=np.float64)
    vd = [r for r in sreacs if isinstance(r, VDepSReac)]
    fm.sreac_vdep = np.array([vd.index(r) if isinstance(r, VDepSReac) else -1 for r in sreacs],
                             dtype=np.int32)
    fm.vdep_vmin = np.array([r.vmin for r in vd], dtype=np.float64)
    fm.vdep_dv = np.array([r.dv for r in vd], dtype=np.float64)
    fm.vdep_off = np.concatenate([[0], np.cumsum([len(r.ktab) for r in vd])]).astype(np.int32)
    fm.vdep_k = (np.concatenate([r.ktab for r in vd]) if vd else np.zeros(0)).astype(np.float64)
    ghk = [c for ss in ssys for c in ss.ghkcurrs.values()]
    if len(ghk) > 4:
        raise _notimpl("at most 4 GHK currents are supported")
    fm.ghk = [c.id for c in ghk]
    fm.ghk_chan = np.array([sspecs.index(c.chanstate.id) for c in ghk], dtype=np.int32)
    fm.ghk_ion = np.array([specs.index(c.ion.id) for c in ghk], dtype=np.int32)
    fm.ghk_valence = np.array([c.ion.valence for c in ghk], dtype=np.int32)
    fm.ghk_perm = np.array([c.P for c in ghk], dtype=np.float64)
    fm.ghk_voconc = np.array([c.virtual_oconc for c in ghk], dtype=np.float64)
    fm.ghk_vshift = np.array([c.vshift for c in ghk], dtype=np.float64)
    fm.ghk_realflux = np.array([int(c.computeflux) for c in ghk], dtype=np.int32)
    fm.sspec_dcst = np.zeros(SS)
    for ss in ssys:
        for d in ss.diffs.values():
            q = sspecs.index(d.lig.id)
            fm.sspec_dcst[q] = max(fm.sspec_dcst[q], d.dcst)   # mobile if any patch's rule moves it
            fm.sdiffs[d.id] = q
    ohm = [c for ss in ssys for c in ss.ohmiccurrs.values()]
    fm.ohmic = [c.id for c in ohm]
    fm.ohmic_chan = np.array([sspecs.index(c.chanstate.id) for c in ohm], dtype=np.int32)
    fm.ohmic_g = np.array([c.g for c in ohm], dtype=np.float64)
    fm.ohmic_erev = np.array([c.erev for c in ohm], dtype=np.float64)
    fm.sreac_lhs_o = np.zeros((n, S), dtype=np.int32)
    fm.sreac_upd_o = np.zeros((n, S), dtype=np.int32)
    if ssys:
        fm.patch_sreacs = [set(range(n))]
        fm.patch_ghk = [set(range(len(ghk)))]
        fm.patch_sspecs = [set(sspecs)]
        fm.patch_sdiffs = [_patch_sdiffs(ssys, sspecs)]
    return fm


@dataclass
class CompTables:
    """Several compartments and patches of one simulation: one flat model over
    the union of their species and reactions (reaction i belongs to
    compartment reac_comp[i]), with per-compartment constants (zero where a
    compartment lacks a reaction or a diffusion rule), as Statedef holds one
    Compdef per compartment and one Patchdef per patch (src/steps/solver/
    statedef.cpp, compdef.cpp, patchdef.cpp). The surface reactions, GHK
    currents and surface species are the union over the patches; fm.patch_*
    says which of them each patch holds."""

    fm: FlatModel
    comp_reacs: list          # [comp] {reaction id: global index}
    comp_diffs: list          # [comp] {diff id: species index}
    comp_specs: list          # [comp] set of species names defined there
    comp_k: np.ndarray        # [comp][reaction]
    comp_dcst: np.ndarray     # [comp][species]


def flatten_comps(model: Model, comps_volsys, patches=()) -> CompTables:
    """Compartments with the given volume systems (one list per compartment)
    and patches given as (surface system ids, inner compartment index, outer
    compartment index or None). Surface reactions with outer-compartment
    species and GHK currents reading the outer compartment need a patch with an
    outer compartment."""
    per = [flatten(model, vs, ()) for vs in comps_volsys]
    ssys_ids = []
    for sids, _, _ in patches:
        for sid in sids:
            if sid not in ssys_ids:
                ssys_ids.append(sid)
    ssys = [model.surfsys[x] for x in ssys_ids]
    for sids, _, oc in patches:
        for sid in sids:
            ss = model.surfsys[sid]
            for r in ss.sreacs.values():
                if r.reqOutside() and oc is None:
                    raise _argerr(f"surface reaction {r.id!r} involves the outer compartment, but "
                                  "the patch has none")
            for c in ss.ghkcurrs.values():
                if c.virtual_oconc < 0 and oc is None:
                    raise _argerr(f"GHK current {c.id!r} reads the outer compartment, but the patch "
                                  "has none; give it a virtual_oconc >= 0")
                if not c.P > 0.0:
                    raise ValueError(f"Permeability not defined for GHK current object {c.id!r}")
    used = set()
    for f in per:
        used.update(f.specs)
    used_s = set()
    for ss in ssys:
        for r in ss.sreacs.values():
            used.update(x.id for x in r.ilhs + r.irhs + r.olhs + r.orhs)
            used_s.update(x.id for x in r.slhs + r.srhs)
        for c in ss.ohmiccurrs.values():
            used_s.add(c.chanstate.id)
        for c in ss.ghkcurrs.values():
            used_s.add(c.chanstate.id)
            used.add(c.ion.id)
        for d in ss.diffs.values():
            used_s.add(d.lig.id)
    order = list(model.specs)
    specs = [x for x in order if x in used]
    sspecs = [x for x in order if x in used_s]
    S, SS = len(specs), len(sspecs)
    reacs = [(ci, model.volsys[v].reacs[r]) for ci, vs in enumerate(comps_volsys)
             for v in vs for r in model.volsys[v].reacs]
    R = len(reacs)
    sreacs = [r for ss in ssys for r in ss.sreacs.values()]
    fm = FlatModel(specs=specs, sspecs=sspecs, reacs=[f"{ci}:{r.id}" for ci, r in reacs],
                   sreacs=[r.id for r in sreacs])
    fm.reac_lhs = np.zeros((R, S), dtype=np.int32)
    fm.reac_upd = np.zeros((R, S), dtype=np.int32)
    for i, (_, r) in enumerate(reacs):
        lhs = _count(r.lhs, specs)
        fm.reac_lhs[i] = lhs
        fm.reac_upd[i] = _count(r.rhs, specs) - lhs
    fm.reac_k = np.zeros(R)
    comp_k = np.zeros((len(per), R))
    comp_reacs = [dict() for _ in per]
    for i, (ci, r) in enumerate(reacs):
        comp_k[ci, i] = r.kcst
        comp_reacs[ci][r.id] = i
    comp_dcst = np.zeros((len(per), S))
    comp_diffs = []
    for ci, f in enumerate(per):
        d = {}
        for did, sj in f.diffs.items():
            sg = specs.index(f.specs[sj])
            comp_dcst[ci, sg] = f.spec_dcst[sj]
            d[did] = sg
        comp_diffs.append(d)
    fm.spec_dcst = comp_dcst.max(axis=0) if len(per) else np.zeros(S)
    for did in set().union(*[d.keys() for d in comp_diffs]) if comp_diffs else ():
        fm.diffs[did] = next(d[did] for d in comp_diffs if did in d)
    n = len(sreacs)
    fm.sreac_lhs_s = np.zeros((n, SS), dtype=np.int32)
    fm.sreac_upd_s = np.zeros((n, SS), dtype=np.int32)
    fm.sreac_lhs_i = np.zeros((n, S), dtype=np.int32)
    fm.sreac_upd_i = np.zeros((n, S), dtype=np.int32)
    fm.sreac_lhs_o = np.zeros((n, S), dtype=np.int32)
    fm.sreac_upd_o = np.zeros((n, S), dtype=np.int32)
    for i, r in enumerate(sreacs):
        ls, li, lo = _count(r.slhs, sspecs), _count(r.ilhs, specs), _count(r.olhs, specs)
        fm.sreac_lhs_s[i], fm.sreac_lhs_i[i], fm.sreac_lhs_o[i] = ls, li, lo
        fm.sreac_upd_s[i] = _count(r.srhs, sspecs) - ls
        fm.sreac_upd_i[i] = _count(r.irhs, specs) - li
        fm.sreac_upd_o[i] = _count(r.orhs, specs) - lo
    fm.sreac_k = np.array([r.kcst for r in sreacs], dtype=np.float64)
    vd = [r for r in sreacs if isinstance(r, VDepSReac)]
    fm.sreac_vdep = np.array([vd.index(r) if isinstance(r, VDepSReac) else -1 for r in sreacs],
                             dtype=np.int32)
    fm.vdep_vmin = np.array([r.vmin for r in vd], dtype=np.float64)
    fm.vdep_dv = np.array([r.dv for r in vd], dtype=np.float64)
    fm.vdep_off = np.concatenate([[0], np.cumsum([len(r.ktab) for r in vd])]).astype(np.int32)
    fm.vdep_k = (np.concatenate([r.ktab for r in vd]) if vd else np.zeros(0)).astype(np.float64)
    ghk = [c for ss in ssys for c in ss.ghkcurrs.values()]
    if len(ghk) > 4:
        raise _notimpl("at most 4 GHK currents are supported")
    fm.ghk = [c.id for c in ghk]
    fm.ghk_chan = np.array([sspecs.index(c.chanstate.id) for c in ghk], dtype=np.int32)
    fm.ghk_ion = np.array([specs.index(c.ion.id) for c in ghk], dtype=np.int32)
    fm.ghk_valence = np.array([c.ion.valence for c in ghk], dtype=np.int32)
    fm.ghk_perm = np.array([c.P for c in ghk], dtype=np.float64)
    fm.ghk_voconc = np.array([c.virtual_oconc for c in ghk], dtype=np.float64)
    fm.ghk_vshift = np.array([c.vshift for c in ghk], dtype=np.float64)
    fm.ghk_realflux = np.array([int(c.computeflux) for c in ghk], dtype=np.int32)
    fm.sspec_dcst = np.zeros(SS)
    for ss in ssys:
        for d in ss.diffs.values():
            q = sspecs.index(d.lig.id)
            fm.sspec_dcst[q] = max(fm.sspec_dcst[q], d.dcst)   # mobile if any patch's rule moves it
            fm.sdiffs[d.id] = q
    ohm = [c for ss in ssys for c in ss.ohmiccurrs.values()]
    fm.ohmic = [c.id for c in ohm]
    fm.ohmic_chan = np.array([sspecs.index(c.chanstate.id) for c in ohm], dtype=np.int32)
    fm.ohmic_g = np.array([c.g for c in ohm], dtype=np.float64)
    fm.ohmic_erev = np.array([c.erev for c in ohm], dtype=np.float64)
    for sids, _, _ in patches:
        held = [model.surfsys[x] for x in sids]
        fm.patch_sreacs.append({sreacs.index(r) for ss in held for r in ss.sreacs.values()})
        fm.patch_ghk.append({ghk.index(c) for ss in held for c in ss.ghkcurrs.values()})
        ps = set()
        for ss in held:
            for r in ss.sreacs.values():
                ps.update(x.id for x in r.slhs + r.srhs)
            ps.update(c.chanstate.id for c in ss.ohmiccurrs.values())
            ps.update(c.chanstate.id for c in ss.ghkcurrs.values())
            ps.update(d.lig.id for d in ss.diffs.values())
        fm.patch_sspecs.append(ps)
        fm.patch_sdiffs.append(_patch_sdiffs(held, sspecs))
    comp_specs = [set(f.specs) for f in per]
    return CompTables(fm, comp_reacs, comp_diffs, comp_specs, comp_k, comp_dcst)
