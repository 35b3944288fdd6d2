"""Mesh loaders and the reference's triangle numbering, on the host.

Fixtures are the reference's own loader test meshes
(test/unit/py/API_2/geom_test/meshes, copied to tests/golden/meshes): the
same 1,400-tet brick in Abaqus, Gmsh 2.2, VTK and TetGen formats, a
two-file Abaqus mesh, an Abaqus mesh with node / triangle / tet groups, and a
mesh saved by the reference in its XML format (cyl_len10_diam1.xml, whose
triangle list is the reference's triangle numbering of that mesh). The
checks follow test_tetMesh.py: identical tet volumes and triangle areas across
formats, 1,400 and 535 tets, the group names and their relations, the
reference's findTetByPoint result 376 on the cylinder.
"""
from pathlib import Path

import numpy as np
import pytest

from steps_amd import geom, meshio as M

G = Path(__file__).parent / "golden" / "meshes"


def _reference_numbering(tets, tris):
    """Triangle numbering restated as a plain loop: supplied triangles first
    (repeats dropped), then tet faces (v0,v1,v2), (v0,v1,v3), (v0,v2,v3),
    (v1,v2,v3) of each tet in turn, each new sorted triple taking the next
    index (tetmesh.cpp:118-139)."""
    index, order = {}, []
    for t in tris:
        k = tuple(sorted(int(x) for x in t))
        if k not in index:
            index[k] = len(order)
            order.append(k)
    face_tri = []
    for t in tets:
        row = []
        for a, b, c in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            k = tuple(sorted((int(t[a]), int(t[b]), int(t[c]))))
            if k not in index:
                index[k] = len(order)
                order.append(k)
            row.append(index[k])
        face_tri.append(row)
    return np.array(order), np.array(face_tri)


def test_triangle_numbering_is_the_references():
    mesh = geom.kuhn_cube(3, 3e-6)
    rng = np.random.default_rng(4)
    surf = mesh.tris[np.nonzero(mesh.tri_tets[:, 1] < 0)[0]]
    sup = surf[rng.permutation(len(surf))[:20]][:, ::-1]         # some boundary faces, reversed
    sup = np.concatenate([sup, sup[:3]])                          # with repeats
    m = geom.Tetmesh(mesh.verts, mesh.tets, sup)
    tris, face_tri = _reference_numbering(mesh.tets, sup)
    np.testing.assert_array_equal(m.tris, tris)
    # face j of the reference is the face opposite vertex 3 - j
    np.testing.assert_array_equal(m.tet_tri, face_tri[:, ::-1])
    # every tet face is found from both sides; tri_tets lists the lowest tet first
    for tri in range(len(m.tris)):
        a, b = m.tri_tets[tri]
        assert tri in m.tet_tri[a] and (b < 0 or (tri in m.tet_tri[b] and a < b))
    with pytest.raises(ValueError, match="not a face"):
        geom.Tetmesh(mesh.verts, mesh.tets, [[0, 1, 30]])


def test_xml_mesh_saved_by_the_reference():
    mesh, comps, patches = M.loadMesh(G / "cyl_len10_diam1")
    assert (len(mesh.verts), len(mesh.tets), len(mesh.tris)) == (len(mesh.verts), 420, 990)
    # the file lists every triangle in the reference's numbering; building the
    # mesh from its tets alone gives the same list
    np.testing.assert_array_equal(geom.Tetmesh(mesh.verts, mesh.tets).tris, mesh.tris)
    assert mesh.findTetByPoint([0.01e-6] * 3) == 376          # test_tetMesh.py:179


def test_all_formats_give_the_same_brick():
    ref = geom.Tetmesh.LoadAbaqus(str(G / "brick_40_4_4_1400tets.inp"), 1e-6)
    assert len(ref.tets) == 1400
    for m in (geom.Tetmesh.LoadVTK(str(G / "brick_40_4_4_1400tets.vtk"), 1e-6),
              geom.Tetmesh.LoadGmsh(str(G / "brick_40_4_4_1400tets.msh"), 1e-6),
              geom.Tetmesh.LoadTetGen(str(G / "brick_40_4_4_1400tets"), 1e-6)):
        np.testing.assert_array_equal(m.tets, ref.tets)
        np.testing.assert_array_equal(m.tris, ref.tris)
        assert list(m.tet_vol) == list(ref.tet_vol)
        assert list(m.tri_area) == list(ref.tri_area)
    assert ref.tetGroups.keys() == {"EB1"} and len(ref.tetGroups["EB1"]) == 1400


def test_abaqus_two_files_and_type_check():
    with pytest.raises(TypeError):
        geom.Tetmesh.LoadAbaqus(42, 1e-6)
    m = geom.Tetmesh.LoadAbaqus((str(G / "mesh_tet.inp"), str(G / "mesh_tri.inp")), 1e-6)
    assert len(m.tets) == 535                                   # test_tetMesh.py:103
    n = len(m.supplied_tris)
    assert n > 0 and set(m.triGroups["EB1"].tolist()) == set(range(n))
    # the listed triangles are the first triangles, each a face of a tet
    np.testing.assert_array_equal(m.tris[:n], np.sort(m.supplied_tris, axis=1))
    assert np.all(m.tri_tets[:n, 0] >= 0)


def test_abaqus_groups():
    m = geom.Tetmesh.LoadAbaqus(str(G / "sphere_mesh.inp"), 1e-6)
    assert set(m.vertGroups) == {"NODEGROUP1", "NODEGROUP2"}
    assert set(m.triGroups) == {"TRIGROUP1", "TRIGROUP2"}
    assert set(m.tetGroups) == {"TETGROUP1", "TETGROUP2"}
    tets = set(m.tetGroups["TETGROUP1"].tolist()) | set(m.tetGroups["TETGROUP2"].tolist())
    assert tets == set(range(len(m.tets)))
    # the triangle groups are the surface of TETGROUP1 (test_tetMesh.py:114-116)
    g1 = m.tetGroups["TETGROUP1"]
    faces = m.tet_tri[g1].ravel()
    surf = {int(f) for f, c in zip(*np.unique(faces, return_counts=True)) if c == 1}
    tris = set(m.triGroups["TRIGROUP1"].tolist()) | set(m.triGroups["TRIGROUP2"].tolist())
    assert tris == surf
    nodes = set(m.vertGroups["NODEGROUP1"].tolist()) | set(m.vertGroups["NODEGROUP2"].tolist())
    assert nodes == set(np.unique(m.tets).tolist())


def test_gmsh_groups_by_tag_and_name():
    mesh, nodes, tets, tris = M.importGmsh(G / "sphere_rad1_37tets.msh", 1e-6)
    groups = tets.getGroups()
    assert sum(len(v) for k, v in groups.items() if k[0] == 0 and isinstance(k[1], int)) == len(mesh.tets)
    named = {k[1] for k in groups if k[0] == 0 and isinstance(k[1], str)}
    m2, g = geom.load_gmsh(G / "sphere_rad1_37tets.msh", 1e-6)
    assert {n for n, (d, _) in g.items() if d == 3} == named


def test_save_and_load_round_trip(tmp_path):
    mesh = geom.Tetmesh.LoadAbaqus(str(G / "brick_40_4_4_1400tets.inp"), 1e-6)
    n = len(mesh.tets)
    c1 = geom.TmComp("comp1", mesh, np.arange(n // 2))
    c2 = geom.TmComp("comp2", mesh, np.arange(n // 2, n))
    c1.addVolsys("vsys")
    f1, f2 = set(mesh.tet_tri[c1.tets].ravel().tolist()), set(mesh.tet_tri[c2.tets].ravel().tolist())
    p1 = geom.TmPatch("patch1", mesh, sorted(f1 & f2), c1, c2)
    p1.addSurfsys("ssys")
    mesh.Save(tmp_path / "brick")
    back, comps, patches = M.loadMesh(tmp_path / "brick")
    np.testing.assert_array_equal(back.verts, mesh.verts)
    np.testing.assert_array_equal(back.tets, mesh.tets)
    np.testing.assert_array_equal(back.tris, mesh.tris)
    assert [c.name for c in comps] == ["comp1", "comp2"] and comps[0].volsys == ["vsys"]
    np.testing.assert_array_equal(comps[1].tets, c2.tets)
    assert patches[0].name == "patch1" and patches[0].surfsys == ["ssys"]
    assert patches[0].icomp.name == "comp1" and patches[0].ocomp.name == "comp2"
    np.testing.assert_array_equal(np.sort(patches[0].tris), np.sort(p1.tris))
