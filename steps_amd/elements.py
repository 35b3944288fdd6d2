"""Per-element constants and switches of one compartment or patch: the state
that TetOpSplitP keeps in every per-tet and per-triangle kproc, composed into
the engine's per-element tables (stg_set_tet_reac_table and the like in
include/stg.h).

The reference holds, per tet and reaction, a constant and an active flag
(Reac::setKcst / setActive, src/steps/mpi/tetopsplit/reac.cpp); per tet and
diffusion rule a constant, per-direction constants, the activation of each
diffusion-boundary direction and an active flag (Diff::setDcst /
setDirectionDcst / setDiffBndActive / setActive, diff.cpp:300-455); per tet
and species a clamp flag (Tet::setClamped); and the same per triangle for
surface reactions, surface diffusion and surface species. The compartment- and
patch-level setters of TetOpSplitP write every element (e.g. _setCompReacK,
tetopsplit.cpp:2253-2300), so a per-element value lasts until the next
compartment-level call. These classes keep that state only once an
element-level setter (or a switch the realization-wide constants cannot
express) is used; until then the engine runs on realization-wide constants.
"""
from __future__ import annotations

import numpy as np


class VolElements:
    """Reactions, diffusion rules and clamps of the n tets of one compartment.

    nbr_same[t, f]: face f of tet t leads to a tet of the same compartment;
    bdir[t, f]: the diffusion-boundary direction of the face or -1."""

    def __init__(self, n: int, R: int, S: int, nbr_same: np.ndarray, bdir: np.ndarray):
        self.n, self.R, self.S = n, R, S
        self.nbr_same = np.asarray(nbr_same, dtype=bool)
        self.bdir = np.asarray(bdir)
        self.k = None        # [R][n] constants      (reaction table in use when not None)
        self.kact = None     # [R][n] active flags
        self.dD = None       # [S][n] Diff constant  (diffusion table in use when not None)
        self.dDir = None     # [S][n][4] per-direction constant, NaN = none
        self.dBnd = None     # [S][n][4] boundary direction active
        self.dAct = None     # [S][n] active flags
        self.clamp = None    # [n] uint32 bit s = species s clamped

    # --- reactions (Reac::setKcst / setActive) ---
    def reac_table_on(self, k_comp, act_comp) -> None:
        if self.k is None:
            self.k = np.repeat(np.asarray(k_comp, dtype=np.float64)[:, None], self.n, axis=1)
            self.kact = np.repeat(np.asarray(act_comp, dtype=bool)[:, None], self.n, axis=1)

    def reac_table(self) -> np.ndarray:
        return np.where(self.kact, self.k, 0.0)

    # --- diffusion rules (Diff, diff.cpp:300-455) ---
    def diff_table_on(self, d_comp, act_comp, bnd_active, bnd_dcst) -> None:
        """Start from the compartment's state: constant d_comp[s], active
        act_comp[s], boundary direction d of species s active when
        bnd_active(d, s) and with the directional constant bnd_dcst(d, s)
        (None: the rule's own)."""
        if self.dD is not None:
            return
        S, n = self.S, self.n
        self.dD = np.repeat(np.asarray(d_comp, dtype=np.float64)[:, None], n, axis=1)
        self.dAct = np.repeat(np.asarray(act_comp, dtype=bool)[:, None], n, axis=1)
        self.dDir = np.full((S, n, 4), np.nan)
        self.dBnd = np.zeros((S, n, 4), dtype=bool)
        for t, f in zip(*np.nonzero(self.bdir >= 0)):
            d = int(self.bdir[t, f])
            for s in range(S):
                self.dBnd[s, t, f] = bool(bnd_active(d, s))
                v = bnd_dcst(d, s)
                if v is not None:
                    self.dDir[s, t, f] = v

    def set_dcst(self, s: int, tets, dcst: float) -> None:
        """Diff::setDcst: the rule's constant, every directional constant cleared."""
        self.dD[s, tets] = dcst
        self.dDir[s, tets, :] = np.nan

    def set_direction_dcst(self, s: int, t: int, f: int, dcst: float) -> None:
        """Diff::setDirectionDcst: one face's constant; a boundary face opens."""
        self.dDir[s, t, f] = dcst
        if self.bdir[t, f] >= 0:
            self.dBnd[s, t, f] = True

    def set_bnd_active(self, s: int, d: int, act: bool) -> None:
        """Diff::setDiffBndActive on every face with boundary direction d: a
        change of the flag resets the rule with setDcst(its constant), which
        clears that tet's directional constants."""
        for t, f in zip(*np.nonzero(self.bdir == d)):
            if self.dBnd[s, t, f] != act:
                self.dBnd[s, t, f] = act
                self.dDir[s, t, :] = np.nan

    def diff_table(self) -> np.ndarray:
        """[S][n][4] diffusion constant through each face: the directional
        constant or the rule's, on faces to the same compartment and on active
        boundary faces; 0 for inactive rules and closed faces."""
        base = np.where(np.isnan(self.dDir), self.dD[:, :, None], self.dDir)
        open_ = self.nbr_same[None, :, :] | ((self.bdir >= 0)[None, :, :] & self.dBnd)
        return np.where(open_ & self.dAct[:, :, None], base, 0.0)

    def face_dcst(self, s: int, t: int) -> np.ndarray:
        return self.diff_table()[s, t] if self.dD is not None else None

    # --- clamps (Tet::setClamped) ---
    def set_clamped(self, s: int, tets, on: bool) -> None:
        if self.clamp is None:
            self.clamp = np.zeros(self.n, dtype=np.uint32)
        if on:
            self.clamp[tets] |= np.uint32(1 << s)
        else:
            self.clamp[tets] &= np.uint32(~(1 << s) & 0xFFFFFFFF)

    def clamped(self, s: int, tets) -> np.ndarray:
        if self.clamp is None:
            return np.zeros(np.size(tets), dtype=bool)
        return ((self.clamp[tets] >> np.uint32(s)) & 1).astype(bool)


class TriElements:
    """Surface reactions, GHK currents, surface diffusion and clamps of the m
    triangles one compartment holds (its patches' triangles, in the engine's
    order for that compartment).

    held[q, i]: entry q (surface reaction, then GHK current) exists on
    triangle i (its patch's surface systems hold it); vdep[q]: surface
    reaction q is voltage dependent (its table value is a factor on k(V))."""

    def __init__(self, m: int, held: np.ndarray, vdep: np.ndarray, k_patch: np.ndarray, SS: int):
        self.m, self.SS = m, SS
        self.held = np.asarray(held, dtype=bool)          # [RS+NG][m]
        self.vdep = np.asarray(vdep, dtype=bool)          # [RS]
        self.k = np.asarray(k_patch, dtype=np.float64).copy()     # [RS][m] surface constants
        self.act = np.ones_like(self.held)                 # [RS+NG][m]
        self.clamp = None                                  # [m] uint32
        # surface diffusion (diffusion table in use when sD is not None)
        self.sHeld = None    # [SS][m] the triangle's patch has a rule for the species (an SDiff)
        self.sD = None       # [SS][m] the rule's constant
        self.sDir = None     # [SS][m][3] per-edge constant, NaN = none
        self.sBnd = None     # [SS][m][3] boundary edge active
        self.edge_same = None    # [m][3] edge leads to a triangle of the same patch
        self.ebnd = None         # [m][3] surface diffusion boundary of the edge, -1 none

    def sreac_table(self) -> np.ndarray:
        """[RS+NG][m]: the constant (or 1 for a k(V) reaction and a GHK current)
        where the entry exists and is active, else 0."""
        RS = len(self.vdep)
        val = np.ones_like(self.held, dtype=np.float64)
        val[:RS] = np.where(self.vdep[:, None], 1.0, self.k)
        return np.where(self.held & self.act, val, 0.0)

    # --- surface diffusion rules (SDiff, sdiff.cpp:42-90, 295-455, 599-616) ---
    def sdiff_table_on(self, s_held, s_dcst, edge_same, ebnd) -> None:
        """Start from the patches' state: s_held[q, i] the triangle's patch has
        a diffusion rule for q, s_dcst[q, i] its constant; edge_same[i, e] the
        edge leads to a triangle of the same patch; ebnd[i, e] the surface
        diffusion boundary on the edge or -1. Boundary edges start closed."""
        if self.sD is not None:
            return
        SS, m = self.SS, self.m
        self.sHeld = np.asarray(s_held, dtype=bool).reshape(SS, m).copy()
        self.sD = np.asarray(s_dcst, dtype=np.float64).reshape(SS, m).copy()
        self.sDir = np.full((SS, m, 3), np.nan)
        self.sBnd = np.zeros((SS, m, 3), dtype=bool)
        self.edge_same = np.asarray(edge_same, dtype=bool).reshape(m, 3)
        self.ebnd = np.asarray(ebnd).reshape(m, 3)

    def set_sdcst(self, q: int, tris, dcst: float) -> None:
        """SDiff::setDcst: the rule's constant, every directional constant cleared."""
        self.sD[q, tris] = dcst
        self.sDir[q, tris, :] = np.nan

    def set_sdirection_dcst(self, q: int, i: int, e: int, dcst: float) -> None:
        """SDiff::setDirectionDcst: one edge's constant; a boundary edge opens."""
        self.sDir[q, i, e] = dcst
        if self.ebnd[i, e] >= 0:
            self.sBnd[q, i, e] = True

    def set_sbnd_active(self, q: int, i: int, e: int, act: bool) -> None:
        """SDiff::setSDiffBndActive on a boundary edge: a change of the flag
        resets the rule with setDcst(its constant), which clears the
        triangle's directional constants."""
        if self.sBnd[q, i, e] != act:
            self.sBnd[q, i, e] = act
            self.sDir[q, i, :] = np.nan

    def sdcst(self, q: int, i: int, e: int | None = None) -> float:
        """SDiff::dcst(direction): the edge's directional constant if set,
        else the rule's."""
        if e is not None and not np.isnan(self.sDir[q, i, e]):
            return float(self.sDir[q, i, e])
        return float(self.sD[q, i])

    def sdiff_table(self) -> np.ndarray:
        """[SS][m][3] surface diffusion constant through each edge: the
        directional constant or the rule's, on edges to the same patch and on
        active boundary edges of triangles that hold the rule; else 0."""
        base = np.where(np.isnan(self.sDir), self.sD[:, :, None], self.sDir)
        open_ = self.edge_same[None, :, :] | ((self.ebnd >= 0)[None, :, :] & self.sBnd)
        return np.where(open_ & self.sHeld[:, :, None], base, 0.0)

    # --- clamps (Tri::setClamped) ---
    def set_clamped(self, q: int, tris, on: bool) -> None:
        if self.clamp is None:
            self.clamp = np.zeros(self.m, dtype=np.uint32)
        if on:
            self.clamp[tris] |= np.uint32(1 << q)
        else:
            self.clamp[tris] &= np.uint32(~(1 << q) & 0xFFFFFFFF)

    def clamped(self, q: int, tris) -> np.ndarray:
        if self.clamp is None:
            return np.zeros(np.size(tris), dtype=bool)
        return ((self.clamp[tris] >> np.uint32(q)) & 1).astype(bool)
