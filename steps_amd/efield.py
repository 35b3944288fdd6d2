"""Membrane potential (EField) on the GPU: object wrapper over the stg_efield
C ABI (include/stg.h), with the reference's EField setters
(src/steps/solver/efield/efield.cpp) and its coupling loop with the
reaction-diffusion solver (TetOpSplitP::_runWithEField,
src/steps/mpi/tetopsplit/tetopsplit.cpp:1851-1954).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

MIXED, FP64 = 0, 1
JACOBI, MULTIGRID = 0, 1


class EField:
    """One membrane (the given triangles) on a tet mesh; SI arguments."""

    def __init__(self, verts, tets, memb_tris):
        self.lib = N.load()
        v = N.as_f64(np.asarray(verts).reshape(-1, 3))
        t = N.as_i32(np.asarray(tets).reshape(-1, 4))
        tr = N.as_i32(np.asarray(memb_tris).reshape(-1, 3))
        m = N.StgEfieldMesh(len(v), N.ptr(v, C.c_double), len(t), N.ptr(t, C.c_int32), len(tr),
                            N.ptr(tr, C.c_int32))
        h = C.c_void_p()
        N.check(self.lib.stg_efield_create(C.byref(m), C.byref(h)), None, efield=True)
        self.h = h
        self.n_verts, self.n_tris = len(v), len(tr)

    def close(self):
        if getattr(self, "h", None):
            self.lib.stg_efield_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        N.check(rc, self.h, efield=True)

    def set_potential(self, v):
        self._chk(self.lib.stg_efield_set_potential(self.h, float(v)))

    def set_capac(self, cm):
        self._chk(self.lib.stg_efield_set_capac(self.h, float(cm)))

    def set_volres(self, ro):
        self._chk(self.lib.stg_efield_set_volres(self.h, float(ro)))

    def set_memb_res(self, rspec, vrev):
        self._chk(self.lib.stg_efield_set_memb_res(self.h, float(rspec), float(vrev)))

    def set_vert_iclamp(self, v, amps):
        self._chk(self.lib.stg_efield_set_vert_iclamp(self.h, int(v), float(amps)))

    def set_tri_iclamp(self, t, amps):
        self._chk(self.lib.stg_efield_set_tri_iclamp(self.h, int(t), float(amps)))

    def set_vert_clamped(self, v, clamped=True):
        self._chk(self.lib.stg_efield_set_vert_clamped(self.h, int(v), 1 if clamped else 0))

    def verts_v(self) -> np.ndarray:
        a = np.empty(self.n_verts)
        self._chk(self.lib.stg_efield_get_verts_v(self.h, N.ptr(a, C.c_double)))
        return a

    def set_verts_v(self, v):
        a = N.as_f64(v)
        self._chk(self.lib.stg_efield_set_verts_v(self.h, N.ptr(a, C.c_double)))

    def tris_v(self) -> np.ndarray:
        a = np.empty(self.n_tris)
        self._chk(self.lib.stg_efield_get_tris_v(self.h, N.ptr(a, C.c_double)))
        return a

    def set_ohmic(self, sspecs, g, erev):
        s, gg, e = N.as_i32(sspecs), N.as_f64(g), N.as_f64(erev)
        self._chk(self.lib.stg_efield_set_ohmic(self.h, len(s), N.ptr(s, C.c_int32),
                                                N.ptr(gg, C.c_double), N.ptr(e, C.c_double)))

    def set_tri_ohmic_erev(self, k, t, erev):
        self._chk(self.lib.stg_efield_set_tri_ohmic_erev(self.h, int(k), int(t), float(erev)))

    def tri_ohmic_erev(self, k, t) -> float:
        x = C.c_double()
        self._chk(self.lib.stg_efield_get_tri_ohmic_erev(self.h, int(k), int(t), C.byref(x)))
        return x.value

    def set_tri_capac(self, t, cm):
        self._chk(self.lib.stg_efield_set_tri_capac(self.h, int(t), float(cm)))

    def tri_capac(self, t) -> float:
        x = C.c_double()
        self._chk(self.lib.stg_efield_get_tri_capac(self.h, int(t), C.byref(x)))
        return x.value

    def tri_ghk_currents(self, g=-1) -> np.ndarray:
        """Per membrane triangle, the mean current (A, outward positive) GHK
        current g (every current for -1) carried over the last step."""
        a = np.empty(self.n_tris)
        self._chk(self.lib.stg_efield_get_tri_ghk_currents(self.h, int(g), N.ptr(a, C.c_double)))
        return a

    def attach(self, engine, real=0):
        self._chk(self.lib.stg_efield_attach(self.h, engine.h if engine else None, int(real)))
        self._engine = engine

    def advance(self, dt):
        self._chk(self.lib.stg_efield_advance(self.h, float(dt)))

    def run(self, engine, endtime, efdt):
        self._chk(self.lib.stg_efield_run(self.h, engine.h, float(endtime), float(efdt)))

    def set_solver(self, precision=MIXED, tol=1e-8, max_iter=20000):
        self._chk(self.lib.stg_efield_set_solver(self.h, int(precision), float(tol), int(max_iter)))

    def set_precond(self, kind=MULTIGRID):
        """MULTIGRID (default): smoothed-aggregation multigrid V-cycle for
        systems above 16,384 vertices; JACOBI: diagonal scaling."""
        self._chk(self.lib.stg_efield_set_precond(self.h, int(kind)))

    def precond_info(self):
        lv, nc, t = C.c_int32(), C.c_int32(), C.c_double()
        self._chk(self.lib.stg_efield_precond_info(self.h, C.byref(lv), C.byref(nc), C.byref(t)))
        return dict(levels=lv.value, coarsest=nc.value, setup_s=t.value)

    def stats(self):
        i, r, x = C.c_int32(), C.c_int32(), C.c_double()
        self._chk(self.lib.stg_efield_stats(self.h, C.byref(i), C.byref(r), C.byref(x)))
        return dict(iterations=i.value, refinements=r.value, rel_residual=x.value)

    def checkpoint(self) -> bytes:
        n = C.c_size_t()
        self._chk(self.lib.stg_efield_state_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._chk(self.lib.stg_efield_checkpoint(self.h, buf, n.value))
        return buf.raw

    def restore(self, state: bytes):
        buf = C.create_string_buffer(state, len(state))
        self._chk(self.lib.stg_efield_restore(self.h, buf, len(state)))

    def sizes(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.lib.stg_efield_sizes(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value
