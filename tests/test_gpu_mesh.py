"""Mesh topology built on the device (stg_mesh_build) against the host
Tetmesh construction of the same vertex and tet arrays.

The integer topology (triangle list and numbering, tet-triangle map,
neighbours and their face slots, triangle-tet map) must be identical; the
measures are formed with the same formulas in the same order of operations,
so they must agree to 1e-13 relative. The meshes are the reference's Gmsh
test fixtures, a Kuhn cube with jittered interior vertices (irregular
geometry), and a 384,000-tet cube, plus the error cases of the host builder.
"""
from pathlib import Path

import numpy as np
import pytest

from steps_amd import geom

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden" / "meshes"
INTS = ("tet_tri", "tet_nbr", "tet_nbr_face", "tris", "tri_tets")
FLOATS = ("tet_vol", "tet_bary", "tet_face_area", "tet_face_dist", "tri_area", "tri_bary")


def _same(host, dev):
    for k in INTS:
        np.testing.assert_array_equal(getattr(dev, k), getattr(host, k), err_msg=k)
        assert getattr(dev, k).dtype == getattr(host, k).dtype, k
    for k in FLOATS:
        np.testing.assert_allclose(getattr(dev, k), getattr(host, k), rtol=1e-13, atol=0, err_msg=k)


@pytest.mark.parametrize("name", ["cube.msh", "sphere_rad1_37tets.msh", "diamond.msh",
                                  "axon_cube_L1000um_D866nm_1135tets.msh"])
def test_gmsh_fixtures(name):
    print(f"case mesh build {name}", flush=True)
    host, _ = geom.load_gmsh(GOLDEN / name, 1e-6)
    dev = geom.Tetmesh(host.verts, host.tets, host.supplied_tris, device=True)
    _same(host, dev)
    dev2, _ = geom.load_gmsh(GOLDEN / name, 1e-6, device=True)
    _same(host, dev2)


def test_jittered_cube():
    print("case mesh build jittered kuhn cube n=9", flush=True)
    n, side = 9, 9e-6
    m = geom.kuhn_cube(n, side)
    v = m.verts.copy()
    h = side / n
    inner = np.all((v > 0.5 * h) & (v < side - 0.5 * h), axis=1)
    rng = np.random.default_rng(5)
    v[inner] += rng.uniform(-0.2 * h, 0.2 * h, size=(inner.sum(), 3))
    host = geom.Tetmesh(v, m.tets)
    dev = geom.Tetmesh(v, m.tets, device=True)
    _same(host, dev)
    # the reference's definition of the diffusion geometry follows from these
    comp_h, comp_d = geom.TmComp("c", host), geom.TmComp("c", dev)
    th, td = geom.mesh_tables(host, comp_h), geom.mesh_tables(dev, comp_d)
    np.testing.assert_array_equal(td.nbr, th.nbr)
    np.testing.assert_allclose(td.face_g, th.face_g, rtol=1e-12, atol=0)


def test_large_cube():
    print("case mesh build kuhn cube n=40 (384,000 tets)", flush=True)
    host = geom.kuhn_cube(40, 40e-6)
    dev = geom.kuhn_cube(40, 40e-6, device=True)
    _same(host, dev)
    assert len(dev.tris) == len(host.tris)


def test_errors_match_the_host_builder():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [2, 0, 0], [0, 0, -1]],
                 dtype=float) * 1e-6
    with pytest.raises(ValueError, match="degenerate"):
        geom.Tetmesh(v, [[0, 1, 2, 5]], device=True)        # four coplanar vertices
    with pytest.raises(ValueError, match="non-manifold"):    # face (0, 1, 2) in three tets
        geom.Tetmesh(v, [[0, 1, 2, 3], [0, 1, 2, 4], [0, 1, 2, 6]], device=True)
    with pytest.raises(ValueError, match="out of range"):
        geom.Tetmesh(v, [[0, 1, 2, 9]], device=True)
