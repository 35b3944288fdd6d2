"""Kinetic model objects with the reference's legacy (API_1) constructors.

``Model``, ``Spec``, ``Volsys``, ``Surfsys``, ``Reac``, ``SReac`` and ``Diff``
take the same arguments as steps.model (pysteps/steps/API_1/model.py, backed by
src/steps/model/*.cpp). ``flatten`` turns a model bound to one compartment and
at most one patch into the stoichiometry tables the engine reads, the role of
Statedef/Compdef/Patchdef (src/steps/solver/compdef.hpp:252-258, 420).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


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


class Surfsys:
    def __init__(self, id: str, model: Model):
        self.id = id
        self.model = model
        self.sreacs: dict[str, SReac] = {}
        self.diffs: dict[str, Diff] = {}
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
    """Surface reaction over patch species (slhs/srhs) and inner-compartment
    species (ilhs/irhs). Outer-compartment species (olhs/orhs) need a tet on
    both sides of the membrane and are rejected by the solver."""

    def __init__(self, id: str, surfsys: Surfsys, ilhs=(), slhs=(), olhs=(), irhs=(), srhs=(),
                 orhs=(), kcst: float = 0.0):
        self.id = id
        self.surfsys = surfsys
        self.ilhs, self.slhs, self.olhs = list(ilhs), list(slhs), list(olhs)
        self.irhs, self.srhs, self.orhs = list(irhs), list(srhs), list(orhs)
        self.kcst = float(kcst)
        if len(self.ilhs) + len(self.slhs) + len(self.olhs) > 4:
            raise ValueError("reaction order above 4 is not supported")
        surfsys.sreacs[id] = self

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
                raise NotImplementedError(
                    f"surface reaction {r.id!r} involves the outer compartment, which needs a "
                    "tet on both sides of the membrane")
            used_v.update(s.id for s in r.ilhs + r.irhs)
            used_s.update(s.id for s in r.slhs + r.srhs)
        if ss.diffs:
            raise NotImplementedError("surface diffusion is not supported by this solver")
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
    return fm
