"""Tetrahedral mesh, compartments and patches, and the flat tables the engine reads.

Mirrors the parts of the reference's geometry layer the operator-split solver
consumes:
  - ``Tetmesh(verts, tets)`` with tet/triangle neighbour relations, volumes,
    areas and barycentres (src/steps/geom/tetmesh.hpp, used by
    TetOpSplitP::_setup at src/steps/mpi/tetopsplit/tetopsplit.cpp:645-690);
  - ``TmComp`` / ``TmPatch`` (src/steps/geom/tmcomp.hpp, tmpatch.hpp);
  - the per-face diffusion factor A_f / (V d_f) with d_f the barycentre
    distance (src/steps/mpi/tetopsplit/diff.cpp:74-79,
    src/steps/mpi/dist/tetopsplit/simulation.cpp:1306-1331).

Face f of a tetrahedron is the face opposite its f-th vertex.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def _face_keys(faces: np.ndarray, nverts: int) -> np.ndarray:
    f = np.sort(faces, axis=1).astype(np.int64)
    if nverts < (1 << 21):
        return (f[:, 0] * nverts + f[:, 1]) * nverts + f[:, 2]
    # wide meshes: rank the sorted triples
    order = np.lexsort((f[:, 2], f[:, 1], f[:, 0]))
    fs = f[order]
    new = np.ones(len(fs), dtype=bool)
    new[1:] = np.any(fs[1:] != fs[:-1], axis=1)
    rank = np.cumsum(new) - 1
    keys = np.empty(len(f), dtype=np.int64)
    keys[order] = rank
    return keys


ELEM_VERTEX, ELEM_TRI, ELEM_TET, ELEM_UNDEFINED = 0, 1, 2, 99   # src/steps/geom/tetmesh.hpp:46


# The faces of tet t in the order the reference enumerates them when it
# numbers triangles (tetmesh.cpp:118-139): (v0,v1,v2), (v0,v1,v3), (v0,v2,v3),
# (v1,v2,v3). The j-th of these is the face opposite vertex 3 - j.
_REF_FACES = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])


def _host_topology(verts: np.ndarray, tets: np.ndarray, tris: np.ndarray) -> dict:
    """Triangle numbering and tet/triangle adjacency of Tetmesh::Tetmesh
    (src/steps/geom/tetmesh.cpp:78-190): the supplied triangles first, in their
    order (repeats dropped), then every other tet face in order of first
    appearance, tet by tet; a triangle's vertices sorted; its first tet the
    lowest-numbered tet having it."""
    nt, nv = len(tets), len(verts)
    ns = len(tris)
    faces = np.concatenate([tris.reshape(-1, 3), tets[:, _REF_FACES].reshape(-1, 3)]).astype(np.int64)
    keys = _face_keys(faces, nv)
    _, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(first), dtype=np.int64)
    rank[order] = np.arange(len(first))
    tri_of = rank[inv]
    ntri = len(first)
    out = dict(tris=np.sort(faces[first[order]], axis=1))
    ft = tri_of[ns:]                          # triangle of tet face 4 t + j
    out["tet_tri"] = ft.reshape(nt, 4)[:, ::-1].copy()
    counts = np.bincount(ft, minlength=ntri)
    if counts.max() > 2:
        raise ValueError("non-manifold mesh: a triangle is shared by more than two tets")
    if np.any(counts == 0):
        raise ValueError("a supplied triangle is not a face of any tetrahedron")
    occ = np.argsort(ft, kind="stable")
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    a = occ[starts]
    b = np.where(counts == 2, occ[np.minimum(starts + 1, len(occ) - 1)], -1)
    out["tri_tets"] = np.stack([a // 4, np.where(b >= 0, b // 4, -1)], axis=1)
    nbr = -np.ones((nt, 4), dtype=np.int64)
    nbr_face = -np.ones((nt, 4), dtype=np.int64)
    two = b >= 0
    ta, fa = a[two] // 4, 3 - a[two] % 4
    tb, fb = b[two] // 4, 3 - b[two] % 4
    nbr[ta, fa] = tb
    nbr_face[ta, fa] = fb
    nbr[tb, fb] = ta
    nbr_face[tb, fb] = fa
    out["tet_nbr"], out["tet_nbr_face"] = nbr, nbr_face
    return out


def _device_topology(verts: np.ndarray, tets: np.ndarray, tris: np.ndarray) -> dict:
    """The face topology and measures of Tetmesh, built by the engine library on
    the GPU (stg_mesh_build): same arrays, same triangle numbering."""
    import ctypes as C

    from . import _native as N

    lib = N.load()
    v = np.ascontiguousarray(verts, dtype=np.float64)
    t = np.ascontiguousarray(tets, dtype=np.int32)
    s = np.ascontiguousarray(np.asarray(tris).reshape(-1, 3), dtype=np.int32)
    h = C.c_void_p()
    rc = lib.stg_mesh_build(N.ptr(v, C.c_double), len(v), N.ptr(t, C.c_int32), len(t), N.ptr(s, C.c_int32),
                            len(s), C.byref(h))
    if rc != N.STG_OK:
        msg = lib.stg_mesh_build_last_error()
        if rc == N.STG_ERR_ARG:
            raise ValueError(msg.decode() if msg else "bad mesh")
        raise N.StgError(rc, msg.decode() if msg else "mesh build failed")
    try:
        nt, ntri = C.c_int32(), C.c_int32()
        N.check(lib.stg_mesh_build_sizes(h, C.byref(nt), C.byref(ntri)))
        nt, ntri = nt.value, ntri.value
        out = dict(tet_vol=np.empty(nt), tet_bary=np.empty((nt, 3)), tet_tri=np.empty((nt, 4), np.int32),
                   tet_nbr=np.empty((nt, 4), np.int32), tet_nbr_face=np.empty((nt, 4), np.int32),
                   tet_face_area=np.empty((nt, 4)), tet_face_dist=np.empty((nt, 4)),
                   tris=np.empty((ntri, 3), np.int32), tri_tets=np.empty((ntri, 2), np.int32),
                   tri_area=np.empty(ntri), tri_bary=np.empty((ntri, 3)))
        cd, ci = C.c_double, C.c_int32
        rc = lib.stg_mesh_build_fetch(
            h, N.ptr(out["tet_vol"], cd), N.ptr(out["tet_bary"], cd), N.ptr(out["tet_tri"], ci),
            N.ptr(out["tet_nbr"], ci), N.ptr(out["tet_nbr_face"], ci), N.ptr(out["tet_face_area"], cd),
            N.ptr(out["tet_face_dist"], cd), N.ptr(out["tris"], ci), N.ptr(out["tri_tets"], ci),
            N.ptr(out["tri_area"], cd), N.ptr(out["tri_bary"], cd))
        if rc != N.STG_OK:
            msg = lib.stg_mesh_build_last_error()
            raise N.StgError(rc, msg.decode() if msg else "mesh fetch failed")
    finally:
        lib.stg_mesh_build_destroy(h)
    for k in ("tet_tri", "tet_nbr", "tet_nbr_face", "tris", "tri_tets"):
        out[k] = out[k].astype(np.int64)
    return out


class Tetmesh:
    """Unstructured tetrahedral mesh with its face topology and measures (SI units),
    Tetmesh(verts, tets, tris) of the reference (src/steps/geom/tetmesh.cpp:78-233):
    the optional triangles take the first triangle indices, in their order, and
    the remaining tet faces follow in order of first appearance, so triangle
    indices agree with the reference's for the same input.

    device=True builds the topology and measures on the GPU (the engine
    library's stg_mesh_build) instead of with numpy on the host: the same
    arrays, for meshes of millions of tets."""

    def __init__(self, verts, tets, tris=None, device: bool = False):
        self.verts = np.ascontiguousarray(np.asarray(verts, dtype=np.float64))
        self.tets = np.ascontiguousarray(np.asarray(tets, dtype=np.int64))
        if self.verts.ndim != 2 or self.verts.shape[1] != 3:
            raise ValueError("verts must be (n, 3)")
        if self.tets.ndim != 2 or self.tets.shape[1] != 4:
            raise ValueError("tets must be (n, 4)")
        nt = len(self.tets)
        nv = len(self.verts)
        if nt == 0:
            raise ValueError("mesh has no tetrahedra")
        if self.tets.min() < 0 or self.tets.max() >= nv:
            raise ValueError("tet vertex index out of range")
        sup = np.zeros((0, 3), dtype=np.int64) if tris is None else np.asarray(tris, dtype=np.int64).reshape(-1, 3)
        if len(sup) and (sup.min() < 0 or sup.max() >= nv):
            raise ValueError("triangle vertex index out of range")
        self.supplied_tris = sup
        self.comps: dict[str, TmComp] = {}
        self.patches: dict[str, TmPatch] = {}
        self.membs: dict[str, Memb] = {}
        self.diffbs: dict[str, DiffBoundary] = {}
        self.sdiffbs: dict[str, SDiffBoundary] = {}
        self.rois: dict[str, tuple[int, np.ndarray]] = {}
        if device:
            for k, v in _device_topology(self.verts, self.tets, sup).items():
                setattr(self, k, v)
            return
        p = self.verts[self.tets]  # (nt, 4, 3)
        e1, e2, e3 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]
        self.tet_vol = np.abs(np.einsum("ij,ij->i", e1, np.cross(e2, e3))) / 6.0
        if np.any(self.tet_vol <= 0):
            raise ValueError("degenerate tetrahedron")
        self.tet_bary = p.mean(axis=1)
        for k, v in _host_topology(self.verts, self.tets, sup).items():
            setattr(self, k, v)
        tp = self.verts[self.tris]
        self.tri_area = 0.5 * np.linalg.norm(np.cross(tp[:, 1] - tp[:, 0], tp[:, 2] - tp[:, 0]), axis=1)
        self.tri_bary = tp.mean(axis=1)
        self.tet_face_area = self.tri_area[self.tet_tri]
        d = np.zeros((nt, 4))
        has = self.tet_nbr >= 0
        ti = np.nonzero(has)
        d[ti] = np.linalg.norm(self.tet_bary[ti[0]] - self.tet_bary[self.tet_nbr[ti]], axis=1)
        self.tet_face_dist = d

    # --- loaders (TetMesh.Load / LoadAbaqus / LoadGmsh / LoadVTK / LoadTetGen,
    # pysteps/steps/API_2/geom.py:1466-1600): the mesh, with the file's blocks
    # and groups as vertGroups / triGroups / tetGroups {name: index array} ---
    @classmethod
    def _with_groups(cls, loaded):
        from .meshio import element_groups

        mesh, nodes, tets, tris = loaded
        mesh.vertGroups = element_groups(nodes)
        mesh.triGroups = element_groups(tris)
        mesh.tetGroups = element_groups(tets)
        return mesh

    @classmethod
    def LoadAbaqus(cls, filename, scale=1.0, ebs=None, device: bool = False):
        """An Abaqus file, or a (tetrahedron file, triangle file) pair."""
        from . import meshio

        if isinstance(filename, tuple) and len(filename) == 2:
            return cls._with_groups(meshio.importAbaqus2(filename[0], filename[1], scale, device=device))
        if not isinstance(filename, str):
            raise TypeError(f"expected a string or a 2-tuple of strings, got {filename!r}")
        return cls._with_groups(meshio.importAbaqus(filename, scale, ebs=ebs, device=device))

    @classmethod
    def LoadGmsh(cls, filename, scale=1.0, device: bool = False):
        from . import meshio

        return cls._with_groups(meshio.importGmsh(filename, scale, device=device))

    @classmethod
    def LoadVTK(cls, filename, scale=1.0, device: bool = False):
        from . import meshio

        return cls._with_groups(meshio.importVTK(filename, scale, device=device))

    @classmethod
    def LoadTetGen(cls, pathroot, scale, device: bool = False):
        from . import meshio

        return cls._with_groups(meshio.importTetGen(pathroot, scale, device=device))

    @classmethod
    def Load(cls, path, scale=1.0, strict: bool = False, device: bool = False):
        """A mesh saved with Save (or the reference's TetMesh.Save / saveMesh),
        its compartments and patches included."""
        from . import meshio

        return meshio.loadMesh(path, scale, strict, device=device)[0]

    def Save(self, path) -> None:
        from . import meshio

        meshio.saveMesh(path, self)

    # --- regions of interest (Tetmesh::addROI, src/steps/geom/tetmesh.cpp) ---
    def addROI(self, id: str, type: int, indices) -> None:
        if id in self.rois:
            raise ValueError(f"ROI {id!r} already exists")
        if type not in (ELEM_VERTEX, ELEM_TRI, ELEM_TET):
            raise ValueError("unknown ROI element type")
        idx = np.unique(np.asarray(list(indices), dtype=np.int64))
        limit = {ELEM_VERTEX: len(self.verts), ELEM_TRI: len(self.tris), ELEM_TET: len(self.tets)}[type]
        if len(idx) and (idx[0] < 0 or idx[-1] >= limit):
            raise ValueError("ROI element index out of range")
        self.rois[id] = (type, idx)

    def removeROI(self, id: str) -> None:
        self._roi(id)
        del self.rois[id]

    def _roi(self, id: str):
        if id not in self.rois:
            raise ValueError(f"ROI {id!r} does not exist")
        return self.rois[id]

    def getROIType(self, id: str) -> int:
        return self._roi(id)[0]

    def getROIData(self, id: str):
        return self._roi(id)[1].tolist()

    def getROIDataSize(self, id: str) -> int:
        return len(self._roi(id)[1])

    def getAllROINames(self):
        return list(self.rois)

    # --- reference-style accessors (src/steps/geom/tetmesh.hpp) ---
    def countTets(self) -> int:
        return len(self.tets)

    def countTris(self) -> int:
        return len(self.tris)

    def countVertices(self) -> int:
        return len(self.verts)

    def getTetVol(self, t: int) -> float:
        return float(self.tet_vol[t])

    def getTriArea(self, t: int) -> float:
        return float(self.tri_area[t])

    def getTetTetNeighb(self, t: int):
        return [int(x) if x >= 0 else -1 for x in self.tet_nbr[t]]

    def getTetTriNeighb(self, t: int):
        return [int(x) for x in self.tet_tri[t]]

    def getTriTetNeighb(self, tri: int):
        return [int(x) for x in self.tri_tets[tri]]

    def getSurfTris(self) -> np.ndarray:
        return np.nonzero(self.tri_tets[:, 1] < 0)[0]

    def getTetBarycenter(self, t: int):
        return self.tet_bary[t].tolist()

    def getMeshVolume(self) -> float:
        return float(self.tet_vol.sum())

    def findTetByPoint(self, pt) -> int:
        """Index of a tet containing pt (barycentric test), or -1."""
        pt = np.asarray(pt, dtype=np.float64)
        p = self.verts[self.tets]
        T = np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=2)
        lam = np.linalg.solve(T, (pt - p[:, 0])[:, :, None])[:, :, 0]
        inside = np.all(lam >= -1e-12, axis=1) & (lam.sum(axis=1) <= 1 + 1e-12)
        idx = np.nonzero(inside)[0]
        return int(idx[0]) if len(idx) else -1

    # --- bars (Tetmesh::buildBarData, src/steps/geom/tetmesh.cpp:233-257): the
    # edges of the triangles, numbered in order of first appearance over the
    # triangles' (v0,v1), (v0,v2), (v1,v2), each stored with sorted vertices ---
    def _bar_data(self):
        if getattr(self, "_bars", None) is None:
            t = self.tris
            pairs = np.stack([t[:, [0, 1]], t[:, [0, 2]], t[:, [1, 2]]], axis=1).reshape(-1, 2)
            keys = pairs[:, 0] * np.int64(len(self.verts)) + pairs[:, 1]
            _, first, inv = np.unique(keys, return_index=True, return_inverse=True)
            order = np.argsort(first, kind="stable")
            rank = np.empty(len(first), dtype=np.int64)
            rank[order] = np.arange(len(first))
            self._bars = pairs[first[order]]
            self._tri_bars = rank[inv.reshape(-1)].reshape(-1, 3)
        return self._bars, self._tri_bars

    @property
    def bars(self) -> np.ndarray:
        return self._bar_data()[0]

    @property
    def tri_bars(self) -> np.ndarray:
        return self._bar_data()[1]

    def countBars(self) -> int:
        return len(self.bars)

    def getBar(self, b: int):
        if not 0 <= b < self.countBars():
            raise ValueError("Bar index is out of range.")
        return [int(x) for x in self.bars[b]]

    def getTriBars(self, t: int):
        if not 0 <= t < self.countTris():
            raise ValueError("Triangle index is out of range.")
        return [int(x) for x in self.tri_bars[t]]

    def getBarTriNeighbs(self, b: int):
        """Triangles having bar b, in increasing order (Tetmesh::getBarTriNeighbs)."""
        if not 0 <= b < self.countBars():
            raise ValueError("Bar index is out of range.")
        return np.nonzero(np.any(self.tri_bars == b, axis=1))[0].tolist()

    def tri_edge_of_bar(self, t: int, b: int) -> int:
        """Edge slot of bar b in triangle t in the engine's edge order, edge e
        joining the triangle's vertices e and (e + 1) mod 3; -1 if t lacks b.
        The reference's bar j of a triangle is the engine's edge (0, 2, 1)[j]."""
        j = np.nonzero(self.tri_bars[t] == b)[0]
        return int((0, 2, 1)[j[0]]) if len(j) else -1

    def _tri_patch_index(self):
        """(patch of each triangle as an index into the patch list or -1, the list)."""
        pats = list(self.patches.values())
        out = -np.ones(self.countTris(), dtype=np.int64)
        for i, p in enumerate(pats):
            out[p.tris] = i
        return out, pats


class TmComp:
    """Compartment: a set of tets with volume systems (src/steps/geom/tmcomp.hpp)."""

    def __init__(self, name: str, mesh: Tetmesh, tets=None):
        self.name = name
        self.mesh = mesh
        self.tets = (np.arange(mesh.countTets()) if tets is None
                     else np.unique(np.asarray(tets, dtype=np.int64)))
        self.volsys: list[str] = []
        self._tetset = set(self.tets.tolist())
        mesh.comps[name] = self

    def addVolsys(self, name: str) -> None:
        self.volsys.append(name)

    def getVol(self) -> float:
        return float(self.mesh.tet_vol[self.tets].sum())


class TmPatch:
    """Membrane patch of triangles on the surface of ``icomp`` (src/steps/geom/tmpatch.hpp)."""

    def __init__(self, name: str, mesh: Tetmesh, tris, icomp: TmComp, ocomp: TmComp | None = None):
        self.name = name
        self.mesh = mesh
        self.tris = np.asarray(tris, dtype=np.int64)
        self.icomp = icomp
        self.ocomp = ocomp
        self.surfsys: list[str] = []
        mesh.patches[name] = self

    def addSurfsys(self, name: str) -> None:
        self.surfsys.append(name)

    def getArea(self) -> float:
        return float(self.mesh.tri_area[self.tris].sum())


class DiffBoundary:
    """Triangles between two compartments across which chosen species may
    diffuse once activated (src/steps/geom/diffboundary.cpp; API_1
    steps.geom.DiffBoundary(id, container, tris))."""

    def __init__(self, id: str, container: Tetmesh, tris):
        self.id = id
        self.mesh = container
        self.tris = np.asarray(tris, dtype=np.int64)
        tt = container.tri_tets[self.tris]
        if np.any(tt[:, 1] < 0):
            raise ValueError("a diffusion-boundary triangle must lie between two tetrahedra")
        pairs = set()
        for a, b in tt:
            ca = [c for c in container.comps.values() if a in c._tetset]
            cb = [c for c in container.comps.values() if b in c._tetset]
            if len(ca) != 1 or len(cb) != 1 or ca[0] is cb[0]:
                raise ValueError("a diffusion-boundary triangle must separate two compartments")
            pairs.add(tuple(sorted((ca[0].name, cb[0].name))))
        if len(pairs) != 1:
            raise ValueError("a diffusion boundary must lie between exactly two compartments")
        na, nb_ = pairs.pop()
        self.compA, self.compB = container.comps[na], container.comps[nb_]
        container.diffbs[id] = self

    def getID(self) -> str:
        return self.id

    def getAllTriIndices(self) -> np.ndarray:
        return self.tris.copy()

    def getComps(self):
        return [self.compA, self.compB]


class SDiffBoundary:
    """Bars between two patches across which chosen surface species may
    diffuse once activated (src/steps/geom/sdiffboundary.cpp:37-115; API_1
    steps.geom.SDiffBoundary(id, container, bars, patches)). Every bar must
    carry exactly one triangle of each patch; repeated bars count once.
    bar_tris[i] = (triangle of patches[0], triangle of patches[1]) on bars[i]."""

    def __init__(self, id: str, container: Tetmesh, bars, patches):
        bars = [int(b) for b in bars]
        if not bars:
            raise ValueError("No triangles provided to Surface Diffusion Boundary initializer function.")
        patches = list(patches)
        if len(patches) != 2:
            raise ValueError("The number of patches provided to Surface Diffusion initializer "
                             "function must be length 2.")
        if id in container.sdiffbs:
            raise ValueError(f"surface diffusion boundary {id!r} already exists")
        pa, pb = patches
        nbars = container.countBars()
        owned = {b for s in container.sdiffbs.values() for b in s.bars.tolist()}
        tri_patch, plist = container._tri_patch_index()
        ia = next((i for i, p in enumerate(plist) if p is pa), -2)
        ib = next((i for i, p in enumerate(plist) if p is pb), -2)
        keep, pairs, seen = [], [], set()
        for b in bars:
            if b in seen:
                continue
            seen.add(b)
            if not 0 <= b < nbars:
                raise ValueError(f"Invalid bar index {b}.")
            if b in owned:
                raise ValueError(f"Bar with index {b} already belongs to a surface diffusion boundary.")
            tris = container.getBarTriNeighbs(b)
            side = []
            for idx, p in ((ia, pa), (ib, pb)):
                mine = [t for t in tris if tri_patch[t] == idx]
                if len(mine) > 1:
                    raise ValueError(f"Duplicate copy of patch{p.name} in connected triangles to bar {b}")
                if not mine:
                    raise ValueError(f"Patch{p.name} is not connected to bar {b}")
                side.append(mine[0])
            keep.append(b)
            pairs.append(side)
        self.id = id
        self.mesh = container
        self.bars = np.asarray(keep, dtype=np.int64)
        self.bar_tris = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        self.patchA, self.patchB = pa, pb
        container.sdiffbs[id] = self

    def getID(self) -> str:
        return self.id

    def getAllBarIndices(self) -> np.ndarray:
        return self.bars.copy()

    def getPatches(self):
        return [self.patchA, self.patchB]

    def isBarInside(self, bars):
        s = set(self.bars.tolist())
        return [int(b) in s for b in bars]


class Memb:
    """Membrane for the potential solve: the triangles of ``patches`` and the
    conduction volume of their inner compartments plus ``supplementary_comps``
    (src/steps/geom/memb.cpp; cysteps_geom.pyx:2905-2952). The bandwidth
    ordering options are accepted and ignored: the GPU solve is a sparse CG."""

    def __init__(self, id: str, container: Tetmesh, patches, verify: bool = False,
                 opt_method: int = 1, search_percent: float = 100.0, opt_file_name: str = "",
                 supplementary_comps=()):
        if not patches:
            raise ValueError("a membrane needs at least one patch")
        self.id = id
        self.mesh = container
        self.patches = list(patches)
        self.tris = np.concatenate([np.asarray(p.tris, dtype=np.int64) for p in self.patches])
        if len(np.unique(self.tris)) != len(self.tris):
            raise ValueError("a triangle belongs to more than one patch of the membrane")
        comps = []
        for c in [p.icomp for p in self.patches] + list(supplementary_comps):
            if all(c is not x for x in comps):
                comps.append(c)
        self.comps = comps
        self.tets = np.unique(np.concatenate([c.tets for c in comps]))
        container.membs[id] = self

    def getID(self) -> str:
        return self.id

    def getAllTriIndices(self) -> np.ndarray:
        return self.tris.copy()

    def getAllVolTetIndices(self) -> np.ndarray:
        return self.tets.copy()

    def countVertices(self) -> int:
        return len(np.unique(self.mesh.tets[self.tets]))


def kuhn_cube(n: int, side: float, origin=(0.0, 0.0, 0.0), device: bool = False) -> Tetmesh:
    """Cube of edge ``side`` (m) cut into n^3 sub-cubes, each split into the six
    Kuhn (Freudenthal) tetrahedra sharing the main diagonal. The split is
    conforming across sub-cubes, every tet has volume (side/n)^3/6, and the six
    tets of a sub-cube are consecutive (cube-major, x fastest). device=True
    builds the topology on the GPU."""
    if n < 1:
        raise ValueError("n must be >= 1")
    h = side / n
    g = np.arange(n + 1)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    verts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) * h + np.asarray(origin)

    def vid(i, j, k):
        return (k * (n + 1) + j) * (n + 1) + i

    c = np.arange(n)
    ck, cj, ci = np.meshgrid(c, c, c, indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = np.empty((len(ci), 6, 4), dtype=np.int64)
    for pi, perm in enumerate(perms):
        cur = [ci.copy(), cj.copy(), ck.copy()]
        tets[:, pi, 0] = vid(*cur)
        for s, ax in enumerate(perm):
            cur[ax] = cur[ax] + 1
            tets[:, pi, s + 1] = vid(*cur)
    return Tetmesh(verts, tets.reshape(-1, 4), device=device)


def load_gmsh(path, scale: float = 1.0, device: bool = False):
    """Read a Gmsh 2.2 ASCII mesh (the format of the reference's test/mesh fixtures,
    read there by TetMesh.LoadGmsh, pysteps/steps/API_2/geom.py). Returns the
    Tetmesh and {physical name: (dim, element indices)}: for dim 3 the indices
    are tets, for dim 2 they are the mesh's triangle indices (the file's
    triangles are the mesh's first triangles, in file order)."""
    from .meshio import importGmsh

    mesh, _, tets, tris = importGmsh(path, scale, device=device)
    groups: dict[str, tuple[int, np.ndarray]] = {}
    for proxy, dim in ((tets, 3), (tris, 2)):
        for key, ids in proxy.getGroups().items():
            if key[0] == 0 and isinstance(key[1], str):
                groups[key[1]] = (dim, np.asarray(ids, dtype=np.int64))
    return mesh, groups


@dataclass
class MeshTables:
    """The engine's per-realization mesh arrays (stg_mesh in include/stg.h)."""

    vol: np.ndarray        # (n,) float64
    nbr: np.ndarray        # (n,4) int32
    nbr_face: np.ndarray   # (n,4) int32
    face_g: np.ndarray     # (n,4) float64
    tri_tet: np.ndarray    # (m,) int32
    tri_area: np.ndarray   # (m,) float64
    nbr_comp: np.ndarray | None = None   # (n,4) int32 compartment of the neighbour
    bdir: np.ndarray | None = None       # (n,4) int32 diffusion-boundary direction, -1 none
    tri_nbr: np.ndarray | None = None    # (m,3) int32 neighbour triangle across each edge
    tri_nbr_edge: np.ndarray | None = None   # (m,3) int32 edge slot in the neighbour
    tri_edge_g: np.ndarray | None = None     # (m,3) float64 l_e / (A d_e), 0 = closed
    tri_otet: np.ndarray | None = None   # (m,) int32 outer tet (local in its compartment), -1 none
    tri_ocomp: np.ndarray | None = None  # (m,) int32 compartment of the outer tet
    tri_nbr_comp: np.ndarray | None = None   # (m,3) int32 compartment of each edge neighbour
    tri_bnd: np.ndarray | None = None    # (m,3) int32 surface diffusion boundary of each edge, -1 none


def patch_tri_tables(mesh: Tetmesh, tris):
    """Edge neighbours of the patch triangles ``tris`` (local numbering) for
    surface diffusion: across edge e = (v_e, v_(e+1)) the other patch triangle
    holding that edge, its slot for the edge, and l_e / (A d_e) with l_e the
    edge length, A the area and d_e the distance between the barycentres
    (SDiff::SDiff, src/steps/mpi/tetopsplit/sdiff.cpp:42-90)."""
    tris = np.asarray(tris, dtype=np.int64)
    m = len(tris)
    tv = mesh.tris[tris]
    nbr = -np.ones((m, 3), dtype=np.int32)
    nbe = -np.ones((m, 3), dtype=np.int32)
    g = np.zeros((m, 3))
    owner = {}
    for i in range(m):
        for e in range(3):
            key = tuple(sorted((int(tv[i, e]), int(tv[i, (e + 1) % 3]))))
            owner.setdefault(key, []).append((i, e))
    bary = mesh.tri_bary[tris]
    area = mesh.tri_area[tris]
    for key, occ in owner.items():
        if len(occ) != 2:
            continue            # a patch border, or an edge shared by more than two
        (i, ei), (j, ej) = occ
        length = float(np.linalg.norm(mesh.verts[key[0]] - mesh.verts[key[1]]))
        d = float(np.linalg.norm(bary[i] - bary[j]))
        if d <= 0.0:
            continue
        nbr[i, ei], nbe[i, ei], g[i, ei] = j, ej, length / (area[i] * d)
        nbr[j, ej], nbe[j, ej], g[j, ej] = i, ei, length / (area[j] * d)
    return nbr, nbe, g


def comps_mesh_tables(mesh: Tetmesh, comps, patches, diffbs, sdiffbs=()):
    """Tables of several compartments of one mesh (one MeshTables per
    compartment, local tet numbering in mesh order). A face between two
    compartments stays open only on a diffusion boundary: it then keeps its
    neighbour (nbr_comp, nbr) and carries direction 2 b + side, side 0 when
    the tet is in boundary b's compA (tetopsplit.cpp:2745-2931). Returns the
    tables and, per direction, (boundary index, source comp, destination comp).

    patches: a list of (triangles, inner compartment index, outer compartment
    index or None), or {inner compartment index: triangles}. A compartment's
    triangles are those of its patches in patch order; each triangle gets its
    outer tet when the patch has an outer compartment; surface diffusion joins
    triangles of the same patch, and across each bar of a surface diffusion
    boundary in ``sdiffbs`` the triangles of its two patches (in either
    compartment: tri_nbr_comp names the neighbour's), the edge carrying
    tri_bnd = the boundary's index (tetopsplit.cpp:540-628, 945-1018)."""
    if isinstance(patches, dict):
        patches = [(t, ci, None) for ci, t in patches.items()]
    ncomp = len(comps)
    comp_of = -np.ones(mesh.countTets(), dtype=np.int64)
    local = -np.ones(mesh.countTets(), dtype=np.int64)
    for ci, c in enumerate(comps):
        if np.any(comp_of[c.tets] >= 0):
            raise ValueError("a tetrahedron belongs to more than one compartment")
        comp_of[c.tets] = ci
        local[c.tets] = np.arange(len(c.tets))
    # (tet, face) -> boundary direction
    face_dir = {}
    dirs = []
    for bi, db in enumerate(diffbs):
        ia = comps.index(db.compA)
        for side in (0, 1):
            dirs.append((bi, [ia, comps.index(db.compB)][side], [comps.index(db.compB), ia][side]))
        for tri in db.tris:
            for tet in mesh.tri_tets[tri]:
                f = int(np.nonzero(mesh.tet_tri[tet] == tri)[0][0])
                side = 0 if comp_of[tet] == ia else 1
                face_dir[(int(tet), f)] = 2 * bi + side
    out = []
    for ci, c in enumerate(comps):
        tets = c.tets
        gnbr = mesh.tet_nbr[tets]
        ncomp_arr = np.where(gnbr >= 0, comp_of[np.maximum(gnbr, 0)], -1)
        bdir = -np.ones((len(tets), 4), dtype=np.int64)
        for (tet, f), d in face_dir.items():
            if comp_of[tet] == ci:
                bdir[local[tet], f] = d
        keep = (gnbr >= 0) & (ncomp_arr >= 0) & ((ncomp_arr == ci) | (bdir >= 0))
        lnbr = np.where(keep, local[np.maximum(gnbr, 0)], -1)
        nface = np.where(keep, mesh.tet_nbr_face[tets], -1)
        area = mesh.tet_face_area[tets]
        dist = mesh.tet_face_dist[tets]
        vol = mesh.tet_vol[tets]
        with np.errstate(divide="ignore", invalid="ignore"):
            g = np.where(keep, area / (vol[:, None] * dist), 0.0)
        g = np.where(np.isfinite(g), g, 0.0)
        mine = [(np.asarray(t, dtype=np.int64), oc) for t, ic, oc in patches if ic == ci]
        tri_otet = tri_ocomp = None
        if not mine:
            tri_tet, tri_area = np.zeros(0, dtype=np.int32), np.zeros(0)
            tn = (None, None, None)
        else:
            tris = np.concatenate([t for t, _ in mine])
            tt = mesh.tri_tets[tris]
            inner = np.where(comp_of[np.maximum(tt[:, 0], 0)] == ci, tt[:, 0], tt[:, 1])
            if np.any(inner < 0) or np.any(comp_of[inner] != ci):
                raise ValueError("patch triangle without an inner tet in its compartment")
            tri_tet, tri_area = local[inner].astype(np.int32), mesh.tri_area[tris]
            outer = np.where(tt[:, 0] == inner, tt[:, 1], tt[:, 0])
            if any(oc is not None for _, oc in mine):
                ocs = np.concatenate([np.full(len(t), -1 if oc is None else oc) for t, oc in mine])
                ok = (ocs >= 0) & (outer >= 0)
                if np.any(ok & (comp_of[np.maximum(outer, 0)] != ocs)):
                    raise ValueError("patch triangle whose outer tet is not in the outer compartment")
                tri_otet = np.where(ok, local[np.maximum(outer, 0)], -1).astype(np.int32)
                tri_ocomp = np.where(ok, ocs, ci).astype(np.int32)
            # surface diffusion within each patch; edges between patches closed
            parts = [patch_tri_tables(mesh, t) for t, _ in mine]
            off = np.concatenate([[0], np.cumsum([len(t) for t, _ in mine])])
            tn = (np.concatenate([np.where(p[0] >= 0, p[0] + off[i], -1) for i, p in enumerate(parts)]).astype(np.int32),
                  np.concatenate([p[1] for p in parts]).astype(np.int32),
                  np.concatenate([p[2] for p in parts]))
        out.append(MeshTables(
            vol=np.ascontiguousarray(vol, dtype=np.float64),
            nbr=np.ascontiguousarray(lnbr, dtype=np.int32),
            nbr_face=np.ascontiguousarray(nface, dtype=np.int32),
            face_g=np.ascontiguousarray(g, dtype=np.float64),
            tri_tet=np.ascontiguousarray(tri_tet, dtype=np.int32),
            tri_area=np.ascontiguousarray(tri_area, dtype=np.float64),
            nbr_comp=np.ascontiguousarray(np.where(keep, ncomp_arr, ci), dtype=np.int32),
            bdir=np.ascontiguousarray(bdir, dtype=np.int32),
            tri_nbr=tn[0], tri_nbr_edge=tn[1], tri_edge_g=tn[2],
            tri_otet=tri_otet, tri_ocomp=tri_ocomp))
    if sdiffbs:
        _link_sdiff_boundaries(mesh, out, patches, sdiffbs)
    return out, dirs


def _link_sdiff_boundaries(mesh: Tetmesh, out, patches, sdiffbs) -> None:
    """Edge links across surface diffusion boundary bars, both ways: the
    neighbour, its edge slot and compartment, and l / (A d) with l the bar
    length and d the distance between the two barycentres (SDiff::SDiff,
    sdiff.cpp:42-90). The triangle of a boundary patch holds no other link
    across such a bar (one triangle of each patch per bar)."""
    pos_c = -np.ones(mesh.countTris(), dtype=np.int64)
    pos_l = -np.ones(mesh.countTris(), dtype=np.int64)
    fill = {}
    for t, ic, _ in patches:
        t = np.asarray(t, dtype=np.int64)
        pos_c[t] = ic
        pos_l[t] = fill.get(ic, 0) + np.arange(len(t))
        fill[ic] = fill.get(ic, 0) + len(t)
    for ci, mt in enumerate(out):
        if len(mt.tri_tet):
            mt.tri_nbr_comp = np.full((len(mt.tri_tet), 3), ci, dtype=np.int32)
            mt.tri_bnd = -np.ones((len(mt.tri_tet), 3), dtype=np.int32)
    for bi, sdb in enumerate(sdiffbs):
        for b, (ta, tb) in zip(sdb.bars, sdb.bar_tris):
            if pos_c[ta] < 0 or pos_c[tb] < 0:
                raise ValueError("a surface diffusion boundary triangle is not in a simulated patch")
            ea, eb = mesh.tri_edge_of_bar(ta, b), mesh.tri_edge_of_bar(tb, b)
            v = mesh.bars[b]
            length = float(np.linalg.norm(mesh.verts[v[0]] - mesh.verts[v[1]]))
            d = float(np.linalg.norm(mesh.tri_bary[ta] - mesh.tri_bary[tb]))
            for (t, e, c, lo), (e2, c2, lo2) in (((ta, ea, pos_c[ta], pos_l[ta]), (eb, pos_c[tb], pos_l[tb])),
                                                 ((tb, eb, pos_c[tb], pos_l[tb]), (ea, pos_c[ta], pos_l[ta]))):
                mt = out[c]
                mt.tri_bnd[lo, e] = bi
                if d <= 0.0:
                    continue
                mt.tri_nbr[lo, e] = lo2
                mt.tri_nbr_edge[lo, e] = e2
                mt.tri_nbr_comp[lo, e] = c2
                mt.tri_edge_g[lo, e] = length / (mesh.tri_area[t] * d)


def mesh_tables(mesh: Tetmesh, comp: TmComp, patch_tris=None) -> MeshTables:
    """Tables for the tets of ``comp`` (renumbered 0..n-1 in mesh order).
    Faces to tets outside the compartment are closed (no diffusion), as in
    Diff::Diff (src/steps/mpi/tetopsplit/diff.cpp:66-79)."""
    tets = comp.tets
    n = len(tets)
    local = -np.ones(mesh.countTets(), dtype=np.int64)
    local[tets] = np.arange(n)
    gnbr = mesh.tet_nbr[tets]
    lnbr = np.where(gnbr >= 0, local[np.maximum(gnbr, 0)], -1)
    nface = np.where(lnbr >= 0, mesh.tet_nbr_face[tets], -1)
    area = mesh.tet_face_area[tets]
    dist = mesh.tet_face_dist[tets]
    vol = mesh.tet_vol[tets]
    with np.errstate(divide="ignore", invalid="ignore"):
        g = np.where(lnbr >= 0, area / (vol[:, None] * dist), 0.0)
    g = np.where(np.isfinite(g), g, 0.0)
    if patch_tris is None:
        tri_tet = np.zeros(0, dtype=np.int32)
        tri_area = np.zeros(0)
    else:
        tris = np.asarray(patch_tris, dtype=np.int64)
        tt = mesh.tri_tets[tris]
        inner = np.where(local[np.maximum(tt[:, 0], 0)] >= 0, tt[:, 0], tt[:, 1])
        if np.any(inner < 0) or np.any(local[inner] < 0):
            raise ValueError("patch triangle without an inner tet in the compartment")
        tri_tet = local[inner].astype(np.int32)
        tri_area = mesh.tri_area[tris]
    tn = patch_tri_tables(mesh, patch_tris) if patch_tris is not None else (None, None, None)
    return MeshTables(
        vol=np.ascontiguousarray(vol, dtype=np.float64),
        nbr=np.ascontiguousarray(lnbr, dtype=np.int32),
        nbr_face=np.ascontiguousarray(nface, dtype=np.int32),
        face_g=np.ascontiguousarray(g, dtype=np.float64),
        tri_tet=np.ascontiguousarray(tri_tet, dtype=np.int32),
        tri_area=np.ascontiguousarray(tri_area, dtype=np.float64),
        tri_nbr=tn[0], tri_nbr_edge=tn[1], tri_edge_g=tn[2],
    )
