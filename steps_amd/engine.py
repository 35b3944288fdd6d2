"""Thin object wrapper over the C ABI: one ``Engine`` owns one ``stg_sim`` handle.

Used by the reference-shaped solver classes (steps_amd.solver), the batched
ensemble (steps_amd.ensemble), the tests and bench.py. All arrays crossing the
boundary are numpy arrays; the engine copies them at create time.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .geom import MeshTables
from .model import FlatModel


def _mesh_struct(mt: MeshTables, keep: list) -> N.StgMesh:
    arrs = [N.as_f64(mt.vol), N.as_i32(mt.nbr), N.as_i32(mt.nbr_face), N.as_f64(mt.face_g),
            N.as_i32(mt.tri_tet), N.as_f64(mt.tri_area)]
    keep.extend(arrs)
    vol, nbr, nf, g, tt, ta = arrs
    m = N.StgMesh(len(vol), N.ptr(vol, C.c_double), N.ptr(nbr, C.c_int32), N.ptr(nf, C.c_int32),
                  N.ptr(g, C.c_double), len(tt), N.ptr(tt, C.c_int32), N.ptr(ta, C.c_double))
    if mt.nbr_comp is not None:
        nc, bd = N.as_i32(mt.nbr_comp), N.as_i32(mt.bdir)
        keep.extend([nc, bd])
        m.tet_nbr_comp = N.ptr(nc, C.c_int32)
        m.tet_face_bdir = N.ptr(bd, C.c_int32)
    if mt.tri_nbr is not None and len(mt.tri_tet):
        tn, te, tg = N.as_i32(mt.tri_nbr), N.as_i32(mt.tri_nbr_edge), N.as_f64(mt.tri_edge_g)
        keep.extend([tn, te, tg])
        m.tri_nbr = N.ptr(tn, C.c_int32)
        m.tri_nbr_edge = N.ptr(te, C.c_int32)
        m.tri_edge_g = N.ptr(tg, C.c_double)
        if getattr(mt, "tri_nbr_comp", None) is not None:
            tc = N.as_i32(mt.tri_nbr_comp)
            keep.append(tc)
            m.tri_nbr_comp = N.ptr(tc, C.c_int32)
    if getattr(mt, "tri_otet", None) is not None and len(mt.tri_tet):
        ot = N.as_i32(mt.tri_otet)
        keep.append(ot)
        m.tri_otet = N.ptr(ot, C.c_int32)
        if mt.tri_ocomp is not None:
            oc = N.as_i32(mt.tri_ocomp)
            keep.append(oc)
            m.tri_ocomp = N.ptr(oc, C.c_int32)
    return m


def _model_struct(fm: FlatModel, keep: list) -> N.StgModel:
    arrs = [N.as_i32(fm.reac_lhs), N.as_i32(fm.reac_upd), N.as_f64(fm.reac_k), N.as_f64(fm.spec_dcst),
            N.as_i32(fm.sreac_lhs_s), N.as_i32(fm.sreac_upd_s), N.as_i32(fm.sreac_lhs_i),
            N.as_i32(fm.sreac_upd_i), N.as_f64(fm.sreac_k), N.as_i32(fm.sreac_vdep),
            N.as_f64(fm.vdep_vmin), N.as_f64(fm.vdep_dv), N.as_i32(fm.vdep_off), N.as_f64(fm.vdep_k),
            N.as_i32(fm.ghk_chan), N.as_i32(fm.ghk_ion), N.as_i32(fm.ghk_valence),
            N.as_f64(fm.ghk_perm), N.as_f64(fm.ghk_voconc), N.as_f64(fm.ghk_vshift),
            N.as_i32(fm.ghk_realflux),
            N.as_f64(fm.sspec_dcst if fm.sspec_dcst is not None else np.zeros(len(fm.sspecs)))]
    keep.extend(arrs)
    (rl, ru, rk, dc, sls, sus, sli, sui, sk, svd, vmin, vdv, voff, vk,
     gc, gi, gz, gp, go, gv, gr, sdc) = arrs
    mdl = N.StgModel(len(fm.specs), len(fm.reacs), N.ptr(rl, C.c_int32), N.ptr(ru, C.c_int32),
                      N.ptr(rk, C.c_double), N.ptr(dc, C.c_double), len(fm.sspecs), len(fm.sreacs),
                      N.ptr(sls, C.c_int32), N.ptr(sus, C.c_int32), N.ptr(sli, C.c_int32),
                      N.ptr(sui, C.c_int32), N.ptr(sk, C.c_double), N.ptr(svd, C.c_int32),
                      len(fm.vdep_vmin), N.ptr(vmin, C.c_double), N.ptr(vdv, C.c_double),
                      N.ptr(voff, C.c_int32), N.ptr(vk, C.c_double), len(fm.ghk_chan),
                      N.ptr(gc, C.c_int32), N.ptr(gi, C.c_int32), N.ptr(gz, C.c_int32),
                      N.ptr(gp, C.c_double), N.ptr(go, C.c_double), N.ptr(gv, C.c_double),
                      N.ptr(gr, C.c_int32), N.ptr(sdc, C.c_double))
    if getattr(fm, "sreac_lhs_o", None) is not None and len(fm.sreacs):
        lo, uo = N.as_i32(fm.sreac_lhs_o), N.as_i32(fm.sreac_upd_o)
        keep.extend([lo, uo])
        mdl.sreac_lhs_o = N.ptr(lo, C.c_int32)
        mdl.sreac_upd_o = N.ptr(uo, C.c_int32)
    return mdl


class Engine:
    """One simulation handle: a single realization or a batch of them."""

    def __init__(self, fm: FlatModel, meshes, seeds, streams=None, reac_k=None, sreac_k=None,
                 spec_dcst=None, mode: int = N.MODE_TETOPSPLIT, comps: bool = False):
        """comps=True: the meshes are the compartments of one simulation
        (stg_create_comps), on one period schedule."""
        self.lib = N.load()
        self.fm = fm
        if isinstance(meshes, MeshTables):
            meshes = [meshes]
        n = len(meshes)
        seeds = np.broadcast_to(np.asarray(seeds, dtype=np.uint64), (n,))
        streams = np.arange(n, dtype=np.uint32) if streams is None else np.asarray(streams, dtype=np.uint32)
        keep: list = []
        self._keep = keep
        model = _model_struct(fm, keep)
        mesh_arr = (N.StgMesh * n)(*[_mesh_struct(m, keep) for m in meshes])
        reals = (N.StgRealization * n)()
        for r in range(n):
            reals[r].seed = int(seeds[r])
            reals[r].stream = int(streams[r])
            for name, src in (("reac_k", reac_k), ("sreac_k", sreac_k), ("spec_dcst", spec_dcst)):
                if src is not None:
                    a = N.as_f64(src[r])
                    keep.append(a)
                    setattr(reals[r], name, N.ptr(a, C.c_double))
        h = C.c_void_p()
        create = self.lib.stg_create_comps if comps else self.lib.stg_create_batch
        rc = create(C.byref(model), n, mesh_arr, reals, mode, C.byref(h))
        N.check(rc, None)
        self.h = h
        self.n_real = n
        self.n_specs = len(fm.specs)
        self.n_sspecs = len(fm.sspecs)
        self.n_tets = [len(m.vol) for m in meshes]
        self.n_tris = [len(m.tri_tet) for m in meshes]
        self._keep = []  # the engine copied everything

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.stg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        N.check(rc, self.h)

    # --- state ---
    def reset(self):
        self._chk(self.lib.stg_reset(self.h))

    def run(self, endtime: float):
        self._chk(self.lib.stg_run(self.h, float(endtime)))

    def advance(self, adv: float):
        self._chk(self.lib.stg_advance(self.h, float(adv)))

    def set_seed(self, seed: int, real: int = 0):
        self._chk(self.lib.stg_set_seed(self.h, real, int(seed)))

    def time(self) -> float:
        t = C.c_double()
        self._chk(self.lib.stg_get_time(self.h, C.byref(t)))
        return t.value

    def set_time(self, t: float):
        self._chk(self.lib.stg_set_time(self.h, float(t)))

    def nsteps(self, real: int = 0) -> int:
        n = C.c_uint64()
        self._chk(self.lib.stg_get_nsteps(self.h, real, C.byref(n)))
        return n.value

    def set_nsteps(self, n: int, real: int = 0):
        self._chk(self.lib.stg_set_nsteps(self.h, real, int(n)))

    def update_period(self, real: int = 0) -> float:
        d = C.c_double()
        self._chk(self.lib.stg_get_update_period(self.h, real, C.byref(d)))
        return d.value

    def extents(self, real: int = 0):
        a, b = C.c_uint64(), C.c_uint64()
        self._chk(self.lib.stg_get_extents(self.h, real, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_tet_counts(self, spec: int, counts, real: int = 0):
        a = N.as_i32(counts)
        if a.shape != (self.n_tets[real],):
            raise ValueError(f"expected {self.n_tets[real]} counts")
        self._chk(self.lib.stg_set_tet_counts(self.h, real, spec, N.ptr(a, C.c_int32)))

    def get_tet_counts(self, spec: int, real: int = 0) -> np.ndarray:
        a = np.empty(self.n_tets[real], dtype=np.int32)
        self._chk(self.lib.stg_get_tet_counts(self.h, real, spec, N.ptr(a, C.c_int32)))
        return a

    def set_tri_counts(self, sspec: int, counts, real: int = 0):
        a = N.as_i32(counts)
        if a.shape != (self.n_tris[real],):
            raise ValueError(f"expected {self.n_tris[real]} counts")
        self._chk(self.lib.stg_set_tri_counts(self.h, real, sspec, N.ptr(a, C.c_int32)))

    def get_tri_counts(self, sspec: int, real: int = 0) -> np.ndarray:
        a = np.empty(self.n_tris[real], dtype=np.int32)
        self._chk(self.lib.stg_get_tri_counts(self.h, real, sspec, N.ptr(a, C.c_int32)))
        return a

    def batch_tet_counts(self, spec: int, tets, real: int = 0) -> np.ndarray:
        idx = np.ascontiguousarray(np.asarray(tets, dtype=np.int64))
        out = np.empty(len(idx), dtype=np.float64)
        self._chk(self.lib.stg_get_batch_tet_counts(self.h, real, spec, N.ptr(idx, C.c_int64), len(idx),
                                                    N.ptr(out, C.c_double)))
        return out

    def sum_batch_tet_counts(self, spec: int, tets, real: int = 0) -> int:
        idx = np.ascontiguousarray(np.asarray(tets, dtype=np.int64))
        t = C.c_int64()
        self._chk(self.lib.stg_sum_batch_tet_counts(self.h, real, spec, N.ptr(idx, C.c_int64), len(idx),
                                                    C.byref(t)))
        return t.value

    def sum_tet_counts(self, spec: int, real: int = 0) -> int:
        t = C.c_int64()
        self._chk(self.lib.stg_sum_tet_counts(self.h, real, spec, C.byref(t)))
        return t.value

    def sum_tri_counts(self, sspec: int, real: int = 0) -> int:
        t = C.c_int64()
        self._chk(self.lib.stg_sum_tri_counts(self.h, real, sspec, C.byref(t)))
        return t.value


    def sum_all_tet_counts(self) -> np.ndarray:
        """Totals of every species in every realization, [n_real, n_specs], in one device call."""
        out = np.zeros((self.n_real, self.n_specs), dtype=np.int64)
        self._chk(self.lib.stg_sum_all_tet_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def sum_all_tri_counts(self) -> np.ndarray:
        """Totals of every surface species in every realization, [n_real, n_sspecs]."""
        out = np.zeros((self.n_real, self.n_sspecs), dtype=np.int64)
        if out.size:
            self._chk(self.lib.stg_sum_all_tri_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out
    def set_reac_k(self, reac: int, k: float, real: int = 0):
        self._chk(self.lib.stg_set_reac_k(self.h, real, reac, float(k)))

    def set_sreac_k(self, sreac: int, k: float, real: int = 0):
        self._chk(self.lib.stg_set_sreac_k(self.h, real, sreac, float(k)))

    def set_dcst(self, spec: int, d: float, real: int = 0):
        self._chk(self.lib.stg_set_dcst(self.h, real, spec, float(d)))

    def set_launch(self, block: int = 256, use_graph: bool = False):
        self._chk(self.lib.stg_set_launch(self.h, block, 1 if use_graph else 0))

    def set_lanes(self, lanes: int = 1):
        """Lanes per tet of the period kernel: 1, or 2/4/8 (the tile schedule
        for volume-only models on small meshes; identical results)."""
        self._chk(self.lib.stg_set_lanes(self.h, int(lanes)))

    @property
    def lanes(self) -> int:
        v = C.c_int32()
        self._chk(self.lib.stg_get_lanes(self.h, C.byref(v)))
        return v.value

    def set_slot_width(self, nbytes: int = 2):
        """Face slot width: 2 (16-bit) or 1 (8-bit) bytes; identical results."""
        self._chk(self.lib.stg_set_slot_width(self.h, int(nbytes)))

    @property
    def slot_width(self) -> int:
        v = C.c_int32()
        self._chk(self.lib.stg_get_slot_width(self.h, C.byref(v)))
        return v.value

    def set_specialize(self, on: bool = True, min_waves: int = 0):
        """Use a period kernel compiled at run time for this model's volume
        reaction network (identical results; blocks up to 256).
        min_waves: wavefronts per SIMD the compiler must allow (0: its choice)."""
        self._chk(self.lib.stg_set_specialize(self.h, 1 if on else 0, int(min_waves)))

    @property
    def specialized(self) -> bool:
        v = C.c_int32()
        self._chk(self.lib.stg_get_specialize(self.h, C.byref(v)))
        return bool(v.value)

    def set_tri_v(self, v, real: int = 0):
        a = N.as_f64(np.broadcast_to(np.asarray(v, dtype=np.float64), (self.n_tris[real],)))
        self._chk(self.lib.stg_set_tri_v(self.h, real, N.ptr(a, C.c_double)))

    def get_tri_v(self, real: int = 0) -> np.ndarray:
        a = np.empty(self.n_tris[real])
        self._chk(self.lib.stg_get_tri_v(self.h, real, N.ptr(a, C.c_double)))
        return a

    def set_sdcst(self, sspec: int, d: float, real: int = 0):
        self._chk(self.lib.stg_set_sdcst(self.h, real, sspec, float(d)))

    def set_boundary_dcst(self, bdir: int, spec: int, dcst: float):
        self._chk(self.lib.stg_set_boundary_dcst(self.h, int(bdir), int(spec), float(dcst)))

    def boundary_dcst(self, bdir: int, spec: int) -> float:
        x = C.c_double()
        self._chk(self.lib.stg_get_boundary_dcst(self.h, int(bdir), int(spec), C.byref(x)))
        return x.value

    def set_temp(self, kelvin: float):
        self._chk(self.lib.stg_set_temp(self.h, float(kelvin)))

    def temp(self) -> float:
        x = C.c_double()
        self._chk(self.lib.stg_get_temp(self.h, C.byref(x)))
        return x.value

    # --- per-element tables (stg_set_tet_reac_table and the like); None removes ---
    def _table(self, fn, real, arr, shape, dtype, ctype):
        if arr is None:
            self._chk(fn(self.h, int(real), None))
            return
        a = np.ascontiguousarray(np.asarray(arr, dtype=dtype))
        if a.shape != shape:
            raise ValueError(f"expected a table of shape {shape}, got {a.shape}")
        self._chk(fn(self.h, int(real), a.ctypes.data_as(C.POINTER(ctype))))

    def set_tet_reac_table(self, k, real: int = 0):
        """[n_reacs][n_tets] rate constant of each reaction in each tet (0: inactive)."""
        self._table(self.lib.stg_set_tet_reac_table, real, k, (len(self.fm.reacs), self.n_tets[real]),
                    np.float64, C.c_double)

    def set_tet_diff_table(self, d, real: int = 0):
        """[n_specs][n_tets][4] diffusion constant through each face."""
        self._table(self.lib.stg_set_tet_diff_table, real, d, (len(self.fm.specs), self.n_tets[real], 4),
                    np.float64, C.c_double)

    def set_tet_clamp(self, mask, real: int = 0):
        """[n_tets] bit s set: species s clamped in that tet."""
        self._table(self.lib.stg_set_tet_clamp, real, mask, (self.n_tets[real],), np.uint32, C.c_uint32)

    def set_tri_sreac_table(self, k, real: int = 0):
        """[n_sreacs + n_ghk][n_tris] surface constants (factors on k(V) and GHK)."""
        self._table(self.lib.stg_set_tri_sreac_table, real, k,
                    (len(self.fm.sreacs) + len(self.fm.ghk_chan), self.n_tris[real]), np.float64, C.c_double)

    def set_tri_sdiff_table(self, d, real: int = 0):
        """[n_sspecs][n_tris][3] surface diffusion constant through each edge."""
        self._table(self.lib.stg_set_tri_sdiff_table, real, d, (len(self.fm.sspecs), self.n_tris[real], 3),
                    np.float64, C.c_double)

    def set_tri_clamp(self, mask, real: int = 0):
        """[n_tris] bit q set: surface species q clamped on that triangle."""
        self._table(self.lib.stg_set_tri_clamp, real, mask, (self.n_tris[real],), np.uint32, C.c_uint32)

    def set_diff_apply_threshold(self, threshold: int):
        self._chk(self.lib.stg_set_diff_apply_threshold(self.h, int(threshold)))

    def checkpoint(self) -> bytes:
        n = C.c_size_t()
        self._chk(self.lib.stg_state_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._chk(self.lib.stg_checkpoint(self.h, buf, n.value))
        return buf.raw

    def restore(self, state: bytes):
        buf = C.create_string_buffer(state, len(state))
        self._chk(self.lib.stg_restore(self.h, buf, len(state)))

    def last_run_kernel_ms(self):
        ms, n = C.c_double(), C.c_int32()
        self._chk(self.lib.stg_last_run_kernel_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value
