"""Mesh import and export: the reference's mesh file formats read into a Tetmesh.

The functions mirror pysteps/steps/API_1/utilities/meshio.py (importAbaqus,
importAbaqus2, importTetGen, importGmsh, importVTK, saveMesh, loadMesh) and
return the same things: the Tetmesh and three ElementProxy objects (nodes,
tets, triangles) mapping the file's element ids to mesh indices and holding
the file's blocks and groups. The triangles a file lists are handed to the
Tetmesh first, so they keep their order as triangle indices 0, 1, ...
(Tetmesh(verts, tets, tris), src/steps/geom/tetmesh.cpp:78-190); every other
tet face follows. Parsing runs on the host; the topology of large meshes can
be built on the GPU (device=True, stg_mesh_build).
"""
from __future__ import annotations

import os
import re
import warnings

import numpy as np

from .geom import ELEM_TET, ELEM_TRI, ELEM_VERTEX, Tetmesh, TmComp, TmPatch


class ElementProxy:
    """Import-id <-> mesh-index map of one element kind, with the file's blocks
    (named contiguous index ranges) and groups (named index lists)
    (meshio.py ElementProxy)."""

    def __init__(self, type: str, unitlength: int):
        self.type = type
        self.unitlength = unitlength
        self._data: list = []
        self._importid: list = []
        self._stepsid: dict = {}
        self.blocks: dict = {}
        self.groups: dict = {}
        self._block = None        # (name, first index) of the open block

    def insert(self, import_id, import_data) -> None:
        if len(import_data) != self.unitlength:
            raise ValueError(f"a {self.type} takes {self.unitlength} values")
        self._stepsid[import_id] = len(self._importid)
        self._importid.append(import_id)
        self._data.extend(import_data)

    def getType(self) -> str:
        return self.type

    def getSize(self) -> int:
        return len(self._importid)

    def getDataFromSTEPSID(self, steps_id: int):
        u = self.unitlength
        return self._data[steps_id * u:(steps_id + 1) * u]

    def getDataFromImportID(self, import_id):
        return self.getDataFromSTEPSID(self._stepsid[import_id])

    def getAllData(self) -> list:
        return self._data

    def getSTEPSID(self, import_id) -> int:
        try:
            return self._stepsid[import_id]
        except KeyError:
            raise ValueError(f"no {self.type} with id {import_id} in the file") from None

    def getImportID(self, steps_id: int):
        return self._importid[steps_id]

    def addGroup(self, group_name, group_ids) -> None:
        self.groups[group_name] = list(group_ids)

    def addToGroup(self, group_name, steps_id: int) -> None:
        self.groups.setdefault(group_name, []).append(steps_id)

    def getGroups(self) -> dict:
        return self.groups

    def getGroup(self, group_key) -> list:
        return self.groups.get(group_key, [])

    def blockBegin(self, block_name) -> None:
        self.blockEnd()
        self._block = (block_name, self.getSize())

    def blockEnd(self) -> None:
        if self._block is not None:
            name, first = self._block
            self.blocks[name] = [first, self.getSize() - 1]
            self._block = None

    def getBlocks(self) -> dict:
        return self.blocks

    def blocksToGroups(self) -> dict:
        return {n: list(range(b[0], b[1] + 1)) for n, b in self.blocks.items()}


def _mesh(nodes: ElementProxy, tets: ElementProxy, tris: ElementProxy, device: bool) -> Tetmesh:
    if tets.getSize() == 0:
        raise ValueError("the file holds no tetrahedra")
    return Tetmesh(np.array(nodes.getAllData(), dtype=np.float64).reshape(-1, 3),
                   np.array(tets.getAllData(), dtype=np.int64).reshape(-1, 4),
                   np.array(tris.getAllData(), dtype=np.int64).reshape(-1, 3), device=device)


# ------------------------------------------------------------------ Abaqus
def _abaqus_line(line: str):
    """('comment' | 'keyword' | 'data' | 'blank', fields) of one line of an
    Abaqus input file: a keyword line gives {'keyword': NAME, PARAM: value},
    a data line its comma-separated values."""
    s = line.strip()
    if not s:
        return "blank", None
    if s.startswith("**"):
        return "comment", None
    if s.startswith("*"):
        parts = [p.strip() for p in s[1:].split(",")]
        out = {"keyword": parts[0].upper()}
        for p in parts[1:]:
            if "=" in p:
                k, v = p.split("=", 1)
                out[k.strip().upper()] = v.strip()
            elif p:
                out[p.upper()] = None
        return "keyword", out
    return "data", [x.strip() for x in s.split(",") if x.strip()]


def _read_abaqus(path, scale, nodes, tets, tris, ebs=None, allow_tets=True, allow_tris=True,
                 node_check=None):
    """Nodes (C3D4 tets, STRI3 triangles) of one Abaqus file into the proxies;
    a *NODE or *ELEMENT section with ELSET=name is a block of that name.
    ebs: element blocks to read, None for all. node_check: instead of
    inserting nodes, check that each node the file lists exists with those
    coordinates (a triangle file sharing the tet file's nodes)."""
    seen = set()
    cur = None          # nodes, tets, tris or None: where data lines go

    def close():
        if cur is not None and not (cur is nodes and node_check is not None):
            cur.blockEnd()

    with open(path, "r") as f:
        for line in f:
            kind, r = _abaqus_line(line)
            if kind in ("blank", "comment"):
                continue
            if kind == "keyword":
                close()
                cur = None
                if r["keyword"] == "NODE":
                    cur = nodes
                    if node_check is None:
                        nodes.blockBegin(r["ELSET"] if "ELSET" in r else "AllNodes")
                elif r["keyword"] == "ELEMENT":
                    name = r.get("ELSET", "AllElements")
                    etype = (r.get("TYPE") or "").upper()
                    if ebs is not None and name not in ebs:
                        continue
                    if etype == "C3D4":
                        if not allow_tets:
                            raise ValueError(f"{path}: a triangle file holds tetrahedra")
                        cur = tets
                        tets.blockBegin(name)
                    elif etype == "STRI3":
                        if not allow_tris:
                            raise ValueError(f"{path}: a tetrahedron file holds triangles; use importAbaqus")
                        cur = tris
                        tris.blockBegin(name)
                continue
            if cur is nodes:
                nid = int(r[0])
                xyz = [float(r[1]) * scale, float(r[2]) * scale, float(r[3]) * scale]
                if node_check is None:
                    nodes.insert(nid, xyz)
                elif node_check.get(nid) != tuple(xyz):
                    raise ValueError(f"{path}: node {nid} is not a node of the tetrahedron file")
            elif cur is tets:
                tets.insert(int(r[0]), [nodes.getSTEPSID(int(x)) for x in r[1:5]])
            elif cur is tris:
                v = tuple(nodes.getSTEPSID(int(x)) for x in r[1:4])
                if v in seen:
                    continue          # a triangle listed twice is read once
                seen.add(v)
                tris.insert(int(r[0]), list(v))
    close()


def importAbaqus(filename, scale, ebs=None, shadow_mesh=None, verbose=False, device=False):
    """Mesh, node, tet and triangle proxies from an Abaqus (CUBIT / Netgen)
    input file: C3D4 tetrahedra and STRI3 triangles, the element blocks named
    by ELSET (meshio.py importAbaqus). ebs: read only these element blocks."""
    if not isinstance(filename, (str, os.PathLike)):
        raise TypeError(f"expected a file name, got {filename!r}")
    if shadow_mesh is not None:
        raise NotImplementedError("shadow meshes of the STEPS-CUBIT toolkit are not supported")
    nodes, tets, tris = ElementProxy("node", 3), ElementProxy("tet", 4), ElementProxy("tri", 3)
    _read_abaqus(filename, scale, nodes, tets, tris, ebs=ebs)
    return _mesh(nodes, tets, tris, device), nodes, tets, tris


def importAbaqus2(tetfilename, trifilename, scale, shadow_mesh=None, verbose=False, device=False):
    """As importAbaqus, with the tetrahedra and the triangles in two files that
    share one node numbering (meshio.py importAbaqus2)."""
    if shadow_mesh is not None:
        raise NotImplementedError("shadow meshes of the STEPS-CUBIT toolkit are not supported")
    nodes, tets, tris = ElementProxy("node", 3), ElementProxy("tet", 4), ElementProxy("tri", 3)
    _read_abaqus(tetfilename, scale, nodes, tets, tris, allow_tris=False)
    coords = {nodes.getImportID(i): tuple(nodes.getDataFromSTEPSID(i)) for i in range(nodes.getSize())}
    _read_abaqus(trifilename, scale, nodes, tets, tris, allow_tets=False, node_check=coords)
    return _mesh(nodes, tets, tris, device), nodes, tets, tris


# ------------------------------------------------------------------ TetGen
def _tetgen_records(path):
    """Data lines of a TetGen file, comments ('#' to end of line) removed."""
    with open(path, "r") as f:
        for line in f:
            s = line.split("#", 1)[0].split()
            if s:
                yield s


def importTetGen(pathroot, scale, verbose=False, device=False):
    """Mesh and proxies from TetGen's <root>.node, <root>.ele and the optional
    <root>.face (meshio.py importTetGen). A region attribute of the
    tetrahedra, and a boundary marker of the faces, become groups named by
    their value."""
    nodes, tets, tris = ElementProxy("node", 3), ElementProxy("tet", 4), ElementProxy("tri", 3)
    for ext in (".node", ".ele"):
        if not os.path.isfile(str(pathroot) + ext):
            raise FileNotFoundError(str(pathroot) + ext)
    rec = _tetgen_records(str(pathroot) + ".node")
    head = next(rec)
    nn, dim = int(head[0]), int(head[1])
    if dim != 3:
        raise ValueError("TetGen node file: only 3-dimensional nodes")
    for _ in range(nn):
        r = next(rec)
        nodes.insert(int(r[0]), [float(r[1]) * scale, float(r[2]) * scale, float(r[3]) * scale])
    rec = _tetgen_records(str(pathroot) + ".ele")
    head = next(rec)
    ne, npe, nattr = int(head[0]), int(head[1]), int(head[2])
    if npe not in (4, 10):
        raise ValueError("TetGen element file: 4- or 10-node tetrahedra")
    for _ in range(ne):
        r = next(rec)
        tets.insert(int(r[0]), [nodes.getSTEPSID(int(x)) for x in r[1:5]])
        if nattr >= 1:
            tets.addToGroup(r[1 + npe], tets.getSize() - 1)
    face = str(pathroot) + ".face"
    if os.path.isfile(face):
        rec = _tetgen_records(face)
        head = next(rec)
        nf, nmark = int(head[0]), int(head[1])
        for _ in range(nf):
            r = next(rec)
            tris.insert(int(r[0]), [nodes.getSTEPSID(int(x)) for x in r[1:4]])
            if nmark >= 1:
                tris.addToGroup(r[4], tris.getSize() - 1)
    return _mesh(nodes, tets, tris, device), nodes, tets, tris


# ------------------------------------------------------------------ Gmsh 2.2
def importGmsh(filename, scale, verbose=False, device=False):
    """Mesh and proxies from a Gmsh 2.2 ASCII file (meshio.py importGmsh):
    4-node tetrahedra (type 4) and triangles (type 2); an element with tags
    t0, t1, ... is in groups (0, t0), (1, t1), ..., and (0, name) for a
    physical name of t0."""
    nodes, tets, tris = ElementProxy("node", 3), ElementProxy("tet", 4), ElementProxy("tri", 3)
    names = {}
    with open(filename, "r") as f:
        lines = iter(f.read().splitlines())
    for ln in lines:
        ln = ln.strip()
        if ln == "$MeshFormat":
            fmt = next(lines).split()
            if fmt[0] != "2.2" or fmt[1] != "0":
                raise ValueError("only the Gmsh 2.2 ASCII format is read")
        elif ln == "$PhysicalNames":
            for _ in range(int(next(lines))):
                d, tag, name = next(lines).split(maxsplit=2)
                names[(int(d), int(tag))] = name.strip().strip('"')
        elif ln == "$Nodes":
            for _ in range(int(next(lines))):
                p = next(lines).split()
                nodes.insert(int(p[0]), [float(x) * scale for x in p[1:4]])
        elif ln == "$Elements":
            for _ in range(int(next(lines))):
                p = [int(x) for x in next(lines).split()]
                eid, etype, ntag = p[0], p[1], p[2]
                tags = p[3:3 + ntag]
                if etype == 2:
                    tris.insert(eid, [nodes.getSTEPSID(v) for v in p[-3:]])
                    proxy = tris
                elif etype == 4:
                    tets.insert(eid, [nodes.getSTEPSID(v) for v in p[-4:]])
                    proxy = tets
                else:
                    continue
                for k, tag in enumerate(tags):
                    proxy.addToGroup((k, tag), proxy.getSize() - 1)
    for (dim, tag), name in names.items():
        proxy = {0: nodes, 2: tris, 3: tets}.get(dim)
        if proxy is not None:
            proxy.addGroup((0, name), proxy.getGroup((0, tag)))
    return _mesh(nodes, tets, tris, device), nodes, tets, tris


# ------------------------------------------------------------------ VTK
def importVTK(filename, scale, verbose=False, device=False):
    """Mesh and proxies from a legacy ASCII VTK unstructured grid
    (meshio.py importVTK): cells of type 10 (tetrahedron) and 5 (triangle);
    other cells are skipped with a warning."""
    nodes, tets, tris = ElementProxy("node", 3), ElementProxy("tet", 4), ElementProxy("tri", 3)
    with open(filename, "r") as f:
        toks = [ln.split() for ln in f.read().splitlines()]
    cells = []
    i = 0
    while i < len(toks):
        t = toks[i]
        i += 1
        if not t:
            continue
        key = t[0].upper()
        if key == "DATASET" and t[1].upper() != "UNSTRUCTURED_GRID":
            raise ValueError("VTK file: not an unstructured grid")
        if key == "POINTS":
            n = int(t[1])
            vals = []
            while len(vals) < 3 * n:
                vals.extend(float(x) for x in toks[i])
                i += 1
            for k in range(n):
                nodes.insert(k, [vals[3 * k] * scale, vals[3 * k + 1] * scale, vals[3 * k + 2] * scale])
        elif key == "CELLS":
            for _ in range(int(t[1])):
                cells.append([int(x) for x in toks[i]])
                i += 1
        elif key == "CELL_TYPES":
            skipped = False
            for k in range(int(t[1])):
                ct = int(toks[i][0])
                i += 1
                c = cells[k]
                if ct == 10 and c[0] == 4:
                    tets.insert(k, c[1:5])
                elif ct == 5 and c[0] == 3:
                    tris.insert(k, c[1:4])
                else:
                    skipped = True
            if skipped:
                warnings.warn("VTK file: cells that are neither triangles nor tetrahedra were skipped")
            break
    return _mesh(nodes, tets, tris, device), nodes, tets, tris


# ------------------------------------------------------------------ STEPS XML
def saveMesh(pathname, tetmesh: Tetmesh) -> None:
    """<pathname>.xml in the reference's format (meshio.py saveMesh): vertices,
    every triangle and tet, compartments (id, volume systems, tets) and patches
    (id, surface systems, inner and outer compartment, triangles)."""
    if not isinstance(tetmesh, Tetmesh):
        raise TypeError(f"expected a Tetmesh, got {tetmesh!r}")

    def join(xs):
        return ",".join(str(int(x)) for x in xs)

    out = ['<?xml version="1.0" encoding="ISO-8859-1"?>', "<tetmesh>",
           f'\t<nodes size="{len(tetmesh.verts)}">']
    for i, v in enumerate(tetmesh.verts):
        out += [f'\t\t<node idx = "{i}">', f"\t\t\t<coords>{float(v[0])!r}, {float(v[1])!r}, {float(v[2])!r}</coords>",
                "\t\t</node>"]
    out += ["\t</nodes>", f'\t<triangles size="{len(tetmesh.tris)}">']
    for i, t in enumerate(tetmesh.tris):
        out += [f'\t\t<tri idx="{i}">', f"\t\t\t<nodes>{t[0]}, {t[1]}, {t[2]}</nodes>", "\t\t</tri>"]
    out += ["\t</triangles>", f'\t<tetrahedrons size="{len(tetmesh.tets)}">']
    for i, t in enumerate(tetmesh.tets):
        out += [f'\t\t<tet idx="{i}">', f"\t\t\t<nodes>{t[0]}, {t[1]}, {t[2]}, {t[3]}</nodes>", "\t\t</tet>"]
    out += ["\t</tetrahedrons>", f'\t<compartments size="{len(tetmesh.comps)}">']
    for i, c in enumerate(tetmesh.comps.values()):
        out += [f'\t\t<comp idx="{i}">', f"\t\t\t<id>{c.name}</id>", f"\t\t\t<volsys>{','.join(c.volsys)}</volsys>",
                f"\t\t\t<tets>{join(np.sort(c.tets))}</tets>", "\t\t</comp>"]
    out += ["\t</compartments>", f'\t<patches size="{len(tetmesh.patches)}">']
    for i, p in enumerate(tetmesh.patches.values()):
        out += [f'\t\t<patch idx = "{i}">', f"\t\t\t<id>{p.name}</id>",
                f"\t\t\t<surfsys>{','.join(p.surfsys)}</surfsys>", f"\t\t\t<icomp>{p.icomp.name}</icomp>",
                f"\t\t\t<ocomp>{p.ocomp.name if p.ocomp is not None else 'null'}</ocomp>",
                f"\t\t\t<tris>{join(np.sort(p.tris))}</tris>", "\t\t</patch>"]
    out += ["\t</patches>", "</tetmesh>"]
    with open(str(pathname) + ".xml", "w") as f:
        f.write("\n".join(out) + "\n")


_TAG = re.compile(r"<\s*(/?)(\w+)((?:\s+\w+\s*=\s*(?:\"[^\"]*\"|'[^']*'|\w+))*)\s*/?\s*>")
_ATTR = re.compile(r"(\w+)\s*=\s*(?:\"([^\"]*)\"|'([^']*)'|(\w+))")


def _xml_items(text):
    """(closing, tag, attrs, body) for each element of the reference's
    line-oriented mesh XML, where a value element is <tag>value</tag>."""
    for ln in text.splitlines():
        s = ln.strip()
        if not s or s.startswith("<?"):
            continue
        m = _TAG.match(s)
        if not m:
            raise ValueError(f"unreadable mesh XML line: {s[:60]!r}")
        closing, tag, attrs = m.group(1) == "/", m.group(2), m.group(3)
        a = {k: x or y or z for k, x, y, z in _ATTR.findall(attrs)}
        body = None
        rest = s[m.end():]
        if rest:
            end = f"</{tag}>"
            if not rest.endswith(end):
                raise ValueError(f"unreadable mesh XML line: {s[:60]!r}")
            body = rest[:-len(end)]
        yield closing, tag, a, body


def loadMesh(pathname, scale=1, strict=False, device=False):
    """(mesh, compartments, patches) from <pathname>.xml written by saveMesh
    here or by the reference (meshio.py loadMesh). The file's triangles are
    every triangle of the mesh, so triangle indices are the file's."""
    with open(str(pathname) + ".xml", "r") as f:
        text = f.read()
    verts, tris, tets, comps, patches = [], [], [], [], []
    cur = None
    stack = []
    for closing, tag, attrs, body in _xml_items(text):
        if closing:
            if stack and stack[-1] == tag:
                stack.pop()
            continue
        if body is None:
            stack.append(tag)
            if tag in ("comp", "patch"):
                cur = {}
                (comps if tag == "comp" else patches).append(cur)
            if strict and tag in ("node", "tri", "tet") and "idx" in attrs:
                lst = {"node": verts, "tri": tris, "tet": tets}[tag]
                if int(attrs["idx"]) != len(lst):
                    raise ValueError(f"mesh XML: {tag} {attrs['idx']} out of order")
            continue
        parent = stack[-1] if stack else None
        if tag == "coords":
            verts.append([float(x) * scale for x in body.split(",")])
        elif tag == "nodes" and parent == "tri":
            tris.append([int(x) for x in body.split(",")])
        elif tag == "nodes" and parent == "tet":
            tets.append([int(x) for x in body.split(",")])
        elif cur is not None and tag in ("id", "volsys", "surfsys", "icomp", "ocomp", "tets", "tris"):
            cur[tag] = body.strip()
    mesh = Tetmesh(np.array(verts), np.array(tets, dtype=np.int64), np.array(tris, dtype=np.int64).reshape(-1, 3),
                   device=device)
    if len(tris) != len(mesh.tris):
        raise ValueError("mesh XML: the triangle list is not the mesh's full triangle list")
    cobj = {}
    for c in comps:
        ids = [int(x) for x in c.get("tets", "").split(",") if x.strip()]
        tc = TmComp(c["id"], mesh, ids)
        for v in (x for x in c.get("volsys", "").split(",") if x.strip()):
            tc.addVolsys(v)
        cobj[c["id"]] = tc
    pobj = []
    for p in patches:
        ids = [int(x) for x in p.get("tris", "").split(",") if x.strip()]
        oc = p.get("ocomp", "null")
        tp = TmPatch(p["id"], mesh, ids, cobj[p["icomp"]], None if oc in ("null", "") else cobj[oc])
        for s_ in (x for x in p.get("surfsys", "").split(",") if x.strip()):
            tp.addSurfsys(s_)
        pobj.append(tp)
    return mesh, list(cobj.values()), pobj


def element_groups(proxy: ElementProxy) -> dict:
    """Blocks and groups of one proxy merged into {name: index array}, a group
    taking precedence over a block of the same name (TetMesh._loadElementProxys,
    pysteps/steps/API_2/geom.py)."""
    out = {n: np.arange(b[0], b[1] + 1) for n, b in proxy.getBlocks().items()}
    for n, lst in proxy.getGroups().items():
        if n in out:
            warnings.warn(f"{n!r} names a {proxy.type} block and a group; the group is taken")
        out[n] = np.asarray(lst, dtype=np.int64)
    return out


__all__ = ["ElementProxy", "importAbaqus", "importAbaqus2", "importTetGen", "importGmsh", "importVTK",
           "saveMesh", "loadMesh", "element_groups", "ELEM_VERTEX", "ELEM_TRI", "ELEM_TET"]
