"""Pins the FP64 EField oracle (oracle/efield_ref.py).

- The restated flux coefficients (tetcoupler.cpp:193-290) equal the P1
  finite-element stiffness -V grad(phi_i).grad(phi_j) on random tets (the
  closed form the engine assembles).
- Rallpack1 (test/validation/API_2/validation_efield/rallpack1.py,
  test_rallpack1.py): passive 1 mm cable, RMS error below 1 mV against the
  golden traces data/rallpack1_correct/{v0,vx}.
- A passive cable held by a steady current reaches V = E + I R_in.
"""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle.efield_ref import EFieldOracle, flux_coefficients
from steps_amd import geom

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_flux_coefficients_are_p1_stiffness():
    rng = np.random.default_rng(0)
    for _ in range(50):
        P = rng.normal(size=(4, 3))
        T = np.hstack([np.ones((4, 1)), P])
        grads = np.linalg.inv(T)[1:, :].T
        vol = abs(np.linalg.det(T)) / 6
        K = vol * grads @ grads.T
        for a in range(4):
            others = [b for b in range(4) if b != a]
            w = flux_coefficients(P[a], [P[b] for b in others])
            assert np.allclose(w, -K[a, others], rtol=1e-10, atol=1e-12)


RALL = dict(R_A=1.0, R_M=4.0, C_M=0.01, E_M=-0.065, Iinj=0.1e-9, diameter=1.0e-6,
            length=1.0e-3, sim_end=0.25, EF_dt=1.0e-5)


def rallpack1_setup():
    mesh, _ = geom.load_gmsh(GOLDEN / "meshes" / "axon_cube_L1000um_D866nm_1135tets.msh", 1.0)
    surf = mesh.getSurfTris()
    z = mesh.verts[:, 2]
    zt = z[mesh.tris[surf]]
    zmin_tris = surf[np.all(zt <= z.min(), axis=1)]
    zmax_tris = surf[np.all(zt >= z.max(), axis=1)]
    memb = np.setdiff1d(surf, np.concatenate([zmin_tris, zmax_tris]))
    return mesh, memb, zmin_tris, zmax_tris


def sample_verts(mesh, tris):
    vs = np.unique(mesh.tris[tris])
    d = np.linalg.norm(mesh.verts[vs, :2], axis=1)
    return [vs[np.argmin(d)], vs[np.argmax(d)]]


def read_trace(name):
    return np.loadtxt(GOLDEN / "rallpack1" / name)[:, 1]


def test_rallpack1_oracle_matches_golden_traces():
    p = RALL
    mesh, memb, zmin_tris, zmax_tris = rallpack1_setup()
    ef = EFieldOracle(mesh.verts, mesh.tets, mesh.tris[memb])
    area_cyl = np.pi * p["diameter"] * p["length"]
    area_factor = ef.tri_area.sum() * 1e-12 / area_cyl
    g_leak = area_cyl / p["R_M"] / len(memb)          # S per triangle, one open channel
    ef.set_potential(p["E_M"])
    ef.set_volres(p["R_A"])
    ef.set_capac(p["C_M"] / area_factor)
    v_zmin = np.unique(mesh.tris[zmin_tris])
    for v in v_zmin:
        ef.set_vert_iclamp(v, p["Iinj"] / len(v_zmin))
    s0, s1 = sample_verts(mesh, zmin_tris), sample_verts(mesh, zmax_tris)
    dt, nsteps, every = p["EF_dt"], int(round(p["sim_end"] / p["EF_dt"])), 5
    A, _ = ef.system(dt)
    lu = spla.splu(A.tocsc())
    out0, out1 = [ef.V[s0].mean() * 1e-3], [ef.V[s1].mean() * 1e-3]
    for k in range(1, nsteps + 1):
        i_mem = g_leak * (ef.tri_v() - p["E_M"])      # Tri::computeI (tri.cpp:664-693)
        _, b = ef.system(dt, i_mem)
        ef.V += lu.solve(b)
        if k % every == 0:
            out0.append(ef.V[s0].mean() * 1e-3)
            out1.append(ef.V[s1].mean() * 1e-3)
    v0, vx = read_trace("v0"), read_trace("vx")
    sim0, sim1 = np.array(out0)[1:], np.array(out1)[1:]    # rallpack1.py:177-188
    n = min(len(sim0), len(v0))
    rms0 = np.linalg.norm(v0[:n] - sim0[:n]) / np.sqrt(n)
    rms1 = np.linalg.norm(vx[:n] - sim1[:n]) / np.sqrt(n)
    assert rms0 < 1e-3 and rms1 < 1e-3, (rms0, rms1)


def test_passive_cable_steady_state():
    """Steady state of a sealed passive cable: V(0) = E + I R_in with
    R_in = r_a lambda coth(L / lambda)."""
    p = RALL
    mesh, memb, zmin_tris, _ = rallpack1_setup()
    ef = EFieldOracle(mesh.verts, mesh.tets, mesh.tris[memb])
    ef.set_potential(p["E_M"])
    ef.set_volres(p["R_A"])
    area_cyl = np.pi * p["diameter"] * p["length"]
    rm_mesh = p["R_M"] * ef.tri_area.sum() * 1e-12 / area_cyl
    ef.set_memb_res(rm_mesh, p["E_M"])
    v_zmin = np.unique(mesh.tris[zmin_tris])
    for v in v_zmin:
        ef.set_vert_iclamp(v, p["Iinj"] / len(v_zmin))
    for _ in range(60):
        ef.advance(0.02)
    xs_area = (mesh.getMeshVolume() / p["length"])
    perim_area = ef.tri_area.sum() * 1e-12 / p["length"]
    ra = p["R_A"] / xs_area                       # ohm / m
    rm = p["R_M"] / (np.pi * p["diameter"])       # ohm m, of the modelled cylinder
    lam = np.sqrt(rm / ra)
    r_in = ra * lam / np.tanh(p["length"] / lam)
    v0 = ef.V[v_zmin].mean() * 1e-3
    assert abs((v0 - p["E_M"]) / (p["Iinj"] * r_in) - 1) < 0.03, (v0, p["E_M"] + p["Iinj"] * r_in)
    assert perim_area > 0
