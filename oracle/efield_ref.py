"""CPU oracle of the membrane-potential (EField) step, in FP64.

TEST INFRASTRUCTURE ONLY (the checker of the HIP EField path and the FP64 CG
reference its error is stated against). Restates, in numpy/scipy:

  - the vertex-vertex geometric coupling of a linear potential inside each tet,
    integrated over the barycentric dual cell of every vertex
    (TetCoupler::fluxCoeficients / coupleMesh,
    src/steps/solver/efield/tetcoupler.cpp:48-180, 193-290, Matrix::lvprod
    src/steps/solver/efield/matrix.cpp:121-134);
  - vertex surface areas, one third of each membrane triangle
    (TetMesh::allocateSurface, src/steps/solver/efield/tetmesh.cpp:231-260);
  - units: microns, ms, mV, pF, nS, pA (src/steps/solver/efield/efield.cpp:33-44,
    84-90, 178-190, 358-381);
  - the implicit step of dVSolverBase::_advance
    (src/steps/solver/efield/dVsolver.hpp:126-171), with the membrane-current
    signs of dVsolver.hpp:94-120 (membrane currents are outward-positive and
    subtracted, clamp currents depolarising and added);
  - solved by sparse LU (exact to FP64 rounding) or FP64 conjugate gradients.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def flux_coefficients(ve, ves):
    """Coefficients w_k of the flux into the dual cell of vertex ve from one tet
    (ve plus three neighbours ves), flux = sum_k w_k (V_k - V_ve); geometric
    units (length)."""
    m = np.array([ves[i] - ve for i in range(3)], dtype=np.float64)
    swap = np.linalg.det(m) < 0.0
    if swap:
        m[[1, 2]] = m[[2, 1]]
    vmi = np.linalg.inv(m)
    ret = np.zeros(3)
    f, g = 1.0 / 12.0, 2.0 / 12.0
    for iv in range(3):
        ip, iq, ir = iv, (iv + 1) % 3, (iv + 2) % 3
        c1 = np.cross(m[ip], m[iq])
        c2 = np.cross(m[iq], m[ir])
        c3 = np.cross(m[ir], m[ip])
        vec = f * c1 + g * c2 + f * c3
        ret += 0.5 * (vec @ vmi)
    if swap:
        ret[[1, 2]] = ret[[2, 1]]
    return ret


def coupling_matrix(verts, tets):
    """Symmetric geometric coupling w_ij (sparse, zero diagonal) from the
    per-vertex flux coefficients summed over the tets around each vertex."""
    nv = len(verts)
    rows, cols, vals = [], [], []
    for tet in tets:
        for a in range(4):
            others = [tet[b] for b in range(4) if b != a]
            w = flux_coefficients(verts[tet[a]], [verts[o] for o in others])
            for o, wo in zip(others, w):
                rows.append(tet[a])
                cols.append(o)
                vals.append(wo)
    W = sp.coo_matrix((vals, (rows, cols)), shape=(nv, nv)).tocsr()
    W.sum_duplicates()
    return W


class EFieldOracle:
    """One membrane (the given triangles) on a tet mesh; SI in, internal units inside."""

    def __init__(self, verts_m, tets, memb_tris):
        self.verts = np.asarray(verts_m, dtype=np.float64) * 1e6   # microns
        self.tets = np.asarray(tets, dtype=np.int64)
        self.tris = np.asarray(memb_tris, dtype=np.int64)
        nv = len(self.verts)
        self.Wgeo = coupling_matrix(self.verts, self.tets)
        p = self.verts[self.tris]
        self.tri_area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
        self.vert_area = np.zeros(nv)
        for k in range(3):
            np.add.at(self.vert_area, self.tris[:, k], self.tri_area / 3.0)
        self.V = np.full(nv, -65.0)                 # mV
        self.capac = 0.01 * self.vert_area          # pF (1 uF/cm^2 = 0.01 pF/um^2)
        self.sigma = 1.0 / 1.0e-3                   # nS/um (1 ohm.m)
        self.gext = np.zeros(nv)                    # nS
        self.vext = 0.0                             # mV
        self.vert_clamp_i = np.zeros(nv)            # pA, depolarising positive
        self.tri_clamp_i = np.zeros(len(self.tris))
        self.clamped = np.zeros(nv, dtype=bool)

    # setters in SI, as EField (efield.cpp:151-216)
    def set_potential(self, v):
        self.V[:] = v * 1e3

    def set_capac(self, cm):
        self.capac = cm * self.vert_area

    def set_volres(self, ro):
        self.sigma = 1.0 / (ro * 1.0e-3)

    def set_memb_res(self, rspec, vrev):
        rs = 1.0e3 * rspec
        self.gext = self.vert_area / rs
        self.vext = vrev * 1e3

    def set_vert_iclamp(self, i, amps):
        self.vert_clamp_i[i] = amps * 1e12

    def set_tri_iclamp(self, i, amps):
        self.tri_clamp_i[i] = amps * 1e12

    def tri_v(self):
        return self.V[self.tris].mean(axis=1) * 1e-3   # volts

    def system(self, dt_s, tri_memb_i_amps=None):
        """Matrix A and right-hand side b of one step (internal units)."""
        nv = len(self.V)
        cur = self.vert_clamp_i.copy()
        c = self.tri_clamp_i.copy()
        if tri_memb_i_amps is not None:
            c = c - np.asarray(tri_memb_i_amps) * 1e12
        for k in range(3):
            np.add.at(cur, self.tris[:, k], c / 3.0)
        W = self.Wgeo * self.sigma
        rowsum = np.asarray(W.sum(axis=1)).ravel()
        oodt = 1.0 / (dt_s * 1e3)
        diag = self.capac * oodt + self.gext + rowsum
        rhs = cur + self.gext * (self.vext - self.V) + (W @ self.V - rowsum * self.V)
        A = sp.diags(diag) - W
        if self.clamped.any():
            A = A.tolil()
            for i in np.nonzero(self.clamped)[0]:
                A.rows[i] = [i]
                A.data[i] = [1.0]
            A = A.tocsr()
            rhs[self.clamped] = 0.0
        return A.tocsr(), rhs

    def advance(self, dt_s, tri_memb_i_amps=None, method="lu", tol=1e-13):
        A, b = self.system(dt_s, tri_memb_i_amps)
        if method == "lu":
            dv = spla.spsolve(A.tocsc(), b)
        else:
            dv, info = spla.cg(A, b, rtol=tol, atol=0.0, maxiter=100_000)
            if info != 0:
                raise RuntimeError("FP64 CG did not converge")
        self.V[~self.clamped] += dv[~self.clamped]
        return dv
