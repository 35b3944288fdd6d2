"""ctypes binding of the CPU oracle (oracle/stg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py for the CPU baseline. The product never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libstg_oracle.so"
KAT_BIN = HERE / "_ref" / "philox_kat"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class StgoSystem(C.Structure):
    _fields_ = [
        ("n_tets", C.c_int32), ("tet_vol", _dp), ("tet_nbr", _ip), ("tet_nbr_face", _ip),
        ("tet_face_g", _dp), ("n_tris", C.c_int32), ("tri_tet", _ip), ("tri_area", _dp),
        ("n_specs", C.c_int32), ("n_reacs", C.c_int32), ("n_sspecs", C.c_int32),
        ("n_sreacs", C.c_int32), ("reac_lhs", _ip), ("reac_upd", _ip), ("reac_k", _dp),
        ("spec_dcst", _dp), ("sreac_lhs_s", _ip), ("sreac_upd_s", _ip), ("sreac_lhs_i", _ip),
        ("sreac_upd_i", _ip), ("sreac_k", _dp), ("counts", _ip), ("tri_counts", _ip),
        ("seed", C.c_uint64), ("stream", C.c_uint32), ("diff_threshold", C.c_int32),
        ("sreac_vdep", _ip), ("vdep_vmin", _dp), ("vdep_dv", _dp), ("vdep_off", _ip),
        ("vdep_k", _dp), ("tri_v", _dp), ("tri_intg_mask", C.c_uint32), ("tri_intg", _dp),
        ("n_ghk", C.c_int32), ("ghk_chan", _ip), ("ghk_ion", _ip), ("ghk_valence", _ip),
        ("ghk_perm", _dp), ("ghk_voconc", _dp), ("ghk_vshift", _dp), ("ghk_realflux", _ip),
        ("temp", C.c_double), ("tri_charge", _ip),
        ("n_comps", C.c_int32), ("tet_comp", _ip), ("tet_ltet", _ip),
        ("comp_stream", C.POINTER(C.c_uint32)), ("comp_reac_k", _dp), ("comp_dcst", _dp),
        ("tet_bdir", _ip), ("bnd_dcst", _dp),
        ("tri_nbr", _ip), ("tri_nbr_edge", _ip), ("tri_edge_g", _dp), ("sspec_dcst", _dp),
        ("tri_ltri", _ip),
        ("tet_k", _dp), ("tet_fD", _dp), ("tet_clamp", C.POINTER(C.c_uint32)),
        ("tri_sk", _dp), ("tri_eD", _dp), ("tri_clamp", C.POINTER(C.c_uint32)),
        ("sreac_lhs_o", _ip), ("sreac_upd_o", _ip), ("tri_otet", _ip),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    u32 = C.c_uint32
    lib.stgo_philox4x32_8.argtypes = [C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
    lib.stgo_philox4x32_8.restype = None
    lib.stgo_log.argtypes = [C.c_double]
    lib.stgo_log.restype = C.c_double
    lib.stgo_exp.argtypes = [C.c_double]
    lib.stgo_exp.restype = C.c_double
    lib.stgo_sqrt.argtypes = [C.c_double]
    lib.stgo_sqrt.restype = C.c_double
    lib.stgo_lfact.argtypes = [C.c_int32]
    lib.stgo_lfact.restype = C.c_double
    lib.stgo_binom_stream.argtypes = [C.c_uint64, u32, u32, u32, u32, C.c_int32, C.c_double,
                                      C.c_int32, _ip]
    lib.stgo_binom_stream.restype = C.c_int32
    lib.stgo_uniforms.argtypes = [C.c_uint64, u32, u32, u32, u32, C.c_int32, _dp]
    lib.stgo_uniforms.restype = None
    lib.stgo_schedule.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int32, _dp, C.c_int32]
    lib.stgo_schedule.restype = C.c_int32
    lib.stgo_update_period.argtypes = [C.POINTER(StgoSystem)]
    lib.stgo_update_period.restype = C.c_double
    lib.stgo_run.argtypes = [C.POINTER(StgoSystem), _dp, C.c_int32, u32, _ip,
                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.stgo_run.restype = C.c_int32
    lib.stgo_threads.argtypes = []
    lib.stgo_threads.restype = C.c_int32
    _lib = lib
    return lib


def threads() -> int:
    """OpenMP threads the oracle's period loops run on."""
    return int(load().stgo_threads())


def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib.stgo_philox4x32_8(c, k, o)
    return [o[i] for i in range(4)]


def binomials(seed, c0, c1, c2, channel, n, p, reps) -> np.ndarray:
    out = np.empty(reps, dtype=np.int32)
    load().stgo_binom_stream(seed, c0, c1, c2, channel, n, p, reps,
                             out.ctypes.data_as(_ip))
    return out


def uniforms(seed, c0, c1, c2, channel, n) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    load().stgo_uniforms(seed, c0, c1, c2, channel, n, out.ctypes.data_as(_dp))
    return out


def schedule(t0, t_end, dt_std, mode, cap=10_000_000) -> np.ndarray:
    buf = np.empty(cap, dtype=np.float64)
    n = load().stgo_schedule(t0, t_end, dt_std, mode, buf.ctypes.data_as(_dp), cap)
    if n < 0:
        raise ValueError(f"schedule failed ({n})")
    return buf[:n].copy()


def ref_philox_kat(lines: list[tuple[int, ...]]):
    """Run the Random123-header generator compiled from the reference sources
    (oracle/_ref/philox_kat) on (c0..c3, k0, k1) tuples; None if not built."""
    if not KAT_BIN.exists():
        return None
    inp = "\n".join(" ".join(str(int(x)) for x in t) for t in lines) + "\n"
    out = subprocess.run([str(KAT_BIN)], input=inp, capture_output=True, text=True, check=True)
    return [tuple(int(x) for x in ln.split()) for ln in out.stdout.strip().splitlines()]


def _or(obj, name, default):
    v = getattr(obj, name, None)
    return default if v is None else v


class OracleSim:
    """One realization on the CPU oracle, built from the same tables as the engine."""

    def __init__(self, fm, mt, seed: int, stream: int = 0, reac_k=None, sreac_k=None,
                 spec_dcst=None, diff_threshold: int = 10):
        self.lib = load()
        f64 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))  # noqa: E731
        i32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int32))  # noqa: E731
        S, SS = len(fm.specs), len(fm.sspecs)
        self.n_tets, self.n_tris = len(mt.vol), len(mt.tri_tet)
        self.a = dict(
            vol=f64(mt.vol), nbr=i32(mt.nbr), nf=i32(mt.nbr_face), g=f64(mt.face_g),
            tt=i32(mt.tri_tet), ta=f64(mt.tri_area), rl=i32(fm.reac_lhs), ru=i32(fm.reac_upd),
            rk=f64(fm.reac_k if reac_k is None else reac_k),
            dc=f64(fm.spec_dcst if spec_dcst is None else spec_dcst),
            sls=i32(fm.sreac_lhs_s), sus=i32(fm.sreac_upd_s), sli=i32(fm.sreac_lhs_i),
            sui=i32(fm.sreac_upd_i), sk=f64(fm.sreac_k if sreac_k is None else sreac_k),
            counts=np.zeros((S, self.n_tets), dtype=np.int32),
            tcounts=np.zeros((max(SS, 1), max(self.n_tris, 1)), dtype=np.int32),
            svd=i32(fm.sreac_vdep if fm.sreac_vdep is not None else np.full(len(fm.sreacs), -1)),
            vmin=f64(fm.vdep_vmin if fm.vdep_vmin is not None else []),
            vdv=f64(fm.vdep_dv if fm.vdep_dv is not None else []),
            voff=i32(fm.vdep_off if fm.vdep_off is not None else [0]),
            vk=f64(fm.vdep_k if fm.vdep_k is not None else []),
            triv=np.full(max(self.n_tris, 1), -65.0e-3),
            intg=np.zeros((max(SS, 1), max(self.n_tris, 1))),
            gc=i32(_or(fm, "ghk_chan", [])), gi=i32(_or(fm, "ghk_ion", [])),
            gz=i32(_or(fm, "ghk_valence", [])), gp=f64(_or(fm, "ghk_perm", [])),
            go=f64(_or(fm, "ghk_voconc", [])), gv=f64(_or(fm, "ghk_vshift", [])),
            gr=i32(_or(fm, "ghk_realflux", [])),
            charge=np.zeros((max(len(_or(fm, "ghk_chan", [])), 1), max(self.n_tris, 1)), dtype=np.int32),
            sdc=f64(_or(fm, "sspec_dcst", np.zeros(SS))),
        )
        if getattr(mt, "tri_nbr", None) is not None and self.n_tris:
            self.a.update(tn=i32(mt.tri_nbr), te=i32(mt.tri_nbr_edge), tg=f64(mt.tri_edge_g))
        a = self.a
        P = lambda x, t: x.ctypes.data_as(C.POINTER(t))  # noqa: E731
        self.sys = StgoSystem(
            self.n_tets, P(a["vol"], C.c_double), P(a["nbr"], C.c_int32), P(a["nf"], C.c_int32),
            P(a["g"], C.c_double), self.n_tris, P(a["tt"], C.c_int32), P(a["ta"], C.c_double),
            S, len(fm.reacs), SS, len(fm.sreacs), P(a["rl"], C.c_int32), P(a["ru"], C.c_int32),
            P(a["rk"], C.c_double), P(a["dc"], C.c_double), P(a["sls"], C.c_int32),
            P(a["sus"], C.c_int32), P(a["sli"], C.c_int32), P(a["sui"], C.c_int32),
            P(a["sk"], C.c_double), P(a["counts"], C.c_int32), P(a["tcounts"], C.c_int32),
            seed, stream, diff_threshold, P(a["svd"], C.c_int32), P(a["vmin"], C.c_double),
            P(a["vdv"], C.c_double), P(a["voff"], C.c_int32), P(a["vk"], C.c_double),
            P(a["triv"], C.c_double), 0, P(a["intg"], C.c_double), len(a["gc"]),
            P(a["gc"], C.c_int32), P(a["gi"], C.c_int32), P(a["gz"], C.c_int32),
            P(a["gp"], C.c_double), P(a["go"], C.c_double), P(a["gv"], C.c_double),
            P(a["gr"], C.c_int32), 293.15, P(a["charge"], C.c_int32))
        if "tn" in a:
            self.sys.tri_nbr = P(a["tn"], C.c_int32)
            self.sys.tri_nbr_edge = P(a["te"], C.c_int32)
            self.sys.tri_edge_g = P(a["tg"], C.c_double)
        self.sys.sspec_dcst = P(a["sdc"], C.c_double)
        if getattr(fm, "sreac_lhs_o", None) is not None and len(fm.sreacs):
            a["slo"], a["suo"] = i32(fm.sreac_lhs_o), i32(fm.sreac_upd_o)
            self.sys.sreac_lhs_o = P(a["slo"], C.c_int32)
            self.sys.sreac_upd_o = P(a["suo"], C.c_int32)
        if getattr(mt, "tri_otet", None) is not None and self.n_tris:
            a["tot"] = i32(mt.tri_otet)
            self.sys.tri_otet = P(a["tot"], C.c_int32)
        self.scratch = np.zeros(S * self.n_tets * 4, dtype=np.int32)
        self.nsteps = 0
        # Philox period counter: advances with nsteps, survives reset()
        self.rng_steps = 0
        self.time = 0.0
        self.events = C.c_uint64(0)
        self.moved = C.c_uint64(0)

    @property
    def counts(self) -> np.ndarray:
        return self.a["counts"]

    def integrate_tri_species(self, sspecs) -> None:
        """Accumulate the time integral of these surface species' counts into
        tri_intg from now on (the ohmic-current channel average)."""
        for q in sspecs:
            self.sys.tri_intg_mask |= 1 << int(q)

    @classmethod
    def comps(cls, fm, comp_meshes, seed: int, comp_k, comp_dcst, streams=None, n_bdir: int = 0,
              **kw):
        """Several compartments of one simulation, laid out as the engine lays
        out stg_create_comps: compartment-major tets and triangles, compartment c
        on Philox stream streams[c] (default c) with its local tet numbering."""
        from steps_amd.geom import MeshTables

        nc = len(comp_meshes)
        off = np.concatenate([[0], np.cumsum([len(m.vol) for m in comp_meshes])])
        nbr, nf, bdir = [], [], []
        for c, m in enumerate(comp_meshes):
            ncmp = m.nbr_comp if m.nbr_comp is not None else np.full_like(m.nbr, c)
            nbr.append(np.where(m.nbr >= 0, off[ncmp] + m.nbr, -1))
            nf.append(m.nbr_face)
            bdir.append(m.bdir if m.bdir is not None else -np.ones_like(m.nbr))
        tri_tet = np.concatenate([m.tri_tet + off[c] for c, m in enumerate(comp_meshes)])
        toff = np.concatenate([[0], np.cumsum([len(m.tri_tet) for m in comp_meshes])])
        tn = [m for m in comp_meshes if m.tri_nbr is not None and len(m.tri_tet)]
        extra = {}
        if tn:
            extra = dict(
                tri_nbr=np.concatenate([np.where(m.tri_nbr >= 0, m.tri_nbr + toff[
                    m.tri_nbr_comp if getattr(m, "tri_nbr_comp", None) is not None else c], -1)
                    for c, m in enumerate(comp_meshes) if len(m.tri_tet)]),
                tri_nbr_edge=np.concatenate([m.tri_nbr_edge for m in comp_meshes if len(m.tri_tet)]),
                tri_edge_g=np.concatenate([m.tri_edge_g for m in comp_meshes if len(m.tri_tet)]))
        if any(m.tri_otet is not None for m in comp_meshes):
            # outer tets in the concatenated (global) tet numbering
            extra["tri_otet"] = np.concatenate([
                np.where(m.tri_otet >= 0, m.tri_otet + off[m.tri_ocomp if m.tri_ocomp is not None else c], -1)
                if m.tri_otet is not None else -np.ones(len(m.tri_tet), dtype=np.int64)
                for c, m in enumerate(comp_meshes)]).astype(np.int32)
        mt = MeshTables(vol=np.concatenate([m.vol for m in comp_meshes]), nbr=np.concatenate(nbr),
                        nbr_face=np.concatenate(nf), face_g=np.concatenate([m.face_g for m in comp_meshes]),
                        tri_tet=tri_tet.astype(np.int32),
                        tri_area=np.concatenate([m.tri_area for m in comp_meshes]), **extra)
        o = cls(fm, mt, seed, **kw)
        a = o.a
        a["tcomp"] = np.ascontiguousarray(np.repeat(np.arange(nc), np.diff(off)), dtype=np.int32)
        a["tlt"] = np.ascontiguousarray(np.concatenate([np.arange(len(m.vol)) for m in comp_meshes]),
                                        dtype=np.int32)
        a["cstr"] = np.ascontiguousarray(np.arange(nc) if streams is None else streams, dtype=np.uint32)
        a["ck"] = np.ascontiguousarray(np.asarray(comp_k, dtype=np.float64).reshape(nc, -1))
        a["cd"] = np.ascontiguousarray(np.asarray(comp_dcst, dtype=np.float64).reshape(nc, -1))
        a["bdir"] = np.ascontiguousarray(np.concatenate(bdir), dtype=np.int32)
        a["bnd"] = np.zeros((max(n_bdir, 1), len(fm.specs)))
        a["trl"] = np.ascontiguousarray(np.concatenate(
            [np.arange(len(m.tri_tet)) for m in comp_meshes] + [np.zeros(1)]), dtype=np.int32)
        P = lambda x, t: x.ctypes.data_as(C.POINTER(t))  # noqa: E731
        o.sys.n_comps = nc
        o.sys.tet_comp = P(a["tcomp"], C.c_int32)
        o.sys.tet_ltet = P(a["tlt"], C.c_int32)
        o.sys.comp_stream = P(a["cstr"], C.c_uint32)
        o.sys.comp_reac_k = P(a["ck"], C.c_double)
        o.sys.comp_dcst = P(a["cd"], C.c_double)
        o.sys.tet_bdir = P(a["bdir"], C.c_int32)
        o.sys.bnd_dcst = P(a["bnd"], C.c_double)
        o.sys.tri_ltri = P(a["trl"], C.c_int32)
        o.comp_off = off
        return o

    def set_tables(self, tet_k=None, tet_fD=None, tet_clamp=None, tri_sk=None, tri_eD=None,
                   tri_clamp=None) -> None:
        """Per-element tables in the oracle's element order (None: unused):
        tet_k [R][N], tet_fD [S][N][4], tet_clamp [N] (bit s), tri_sk
        [RS + NG][NTRI], tri_eD [SS][NTRI][3], tri_clamp [NTRI] (bit q)."""
        P = lambda x, t: x.ctypes.data_as(C.POINTER(t))  # noqa: E731
        for name, arr, ct, dt in (("tet_k", tet_k, C.c_double, np.float64),
                                  ("tet_fD", tet_fD, C.c_double, np.float64),
                                  ("tet_clamp", tet_clamp, C.c_uint32, np.uint32),
                                  ("tri_sk", tri_sk, C.c_double, np.float64),
                                  ("tri_eD", tri_eD, C.c_double, np.float64),
                                  ("tri_clamp", tri_clamp, C.c_uint32, np.uint32)):
            if arr is None:
                self.a.pop("tab_" + name, None)
                setattr(self.sys, name, None)
            else:
                a = np.ascontiguousarray(np.asarray(arr, dtype=dt))
                self.a["tab_" + name] = a
                setattr(self.sys, name, P(a, ct))

    @property
    def bnd_dcst(self) -> np.ndarray:
        """[direction][species] diffusion constant across each boundary direction."""
        return self.a["bnd"]

    @property
    def tri_charge(self) -> np.ndarray:
        """GHK charge moved out through each triangle by all currents
        (elementary charges; a copy)."""
        return self.a["charge"].sum(axis=0, dtype=np.int64).astype(np.int32)

    @property
    def tri_charge_per_current(self) -> np.ndarray:
        """[current][triangle] GHK charge moved out (elementary charges)."""
        return self.a["charge"]

    def clear_charge(self) -> None:
        self.a["charge"][:] = 0

    def set_temp(self, kelvin: float) -> None:
        self.sys.temp = kelvin

    @property
    def tri_intg(self) -> np.ndarray:
        return self.a["intg"]

    @property
    def tri_v(self) -> np.ndarray:
        """Membrane potential of each triangle (V), read by k(V) reactions."""
        return self.a["triv"]

    @property
    def tri_counts(self) -> np.ndarray:
        return self.a["tcounts"]

    def update_period(self) -> float:
        return self.lib.stgo_update_period(C.byref(self.sys))

    def run_periods(self, dts) -> None:
        d = np.ascontiguousarray(np.asarray(dts, dtype=np.float64))
        if len(d) == 0:
            return
        rc = self.lib.stgo_run(C.byref(self.sys), d.ctypes.data_as(_dp), len(d), self.rng_steps,
                               self.scratch.ctypes.data_as(_ip), C.byref(self.events),
                               C.byref(self.moved))
        if rc != 0:
            raise RuntimeError(f"oracle run failed ({rc})")
        self.nsteps += len(d)
        self.rng_steps += len(d)

    def reset(self) -> None:
        """Counts, time, step count, extents, channel integrals and GHK charges
        to zero (TetOpSplitP::reset, tetopsplit.cpp:1436-1503); the Philox
        period counter carries on, as the reference's RNG does."""
        self.a["counts"][:] = 0
        self.a["tcounts"][:] = 0
        self.a["intg"][:] = 0.0
        self.a["charge"][:] = 0
        self.nsteps = 0
        self.time = 0.0
        self.events.value = 0
        self.moved.value = 0

    def run(self, endtime: float, mode: int = 0) -> None:
        dts = schedule(self.time, endtime, self.update_period(), mode)
        self.run_periods(dts)
        self.time = endtime
