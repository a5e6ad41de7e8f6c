"""Surface meshes on the host: the oracle's advectSurfaceMesh
(DEFOR_SOLID_3D.cpp:1125-1180) pinned by closed forms, its BVH broad phase
(TRIANGLE_BVH.cpp:179-192) against brute force over triangle boxes, and
libesf's OBJ reader/writer (OBJ.cpp:97-375)."""
import os

import numpy as np
import pytest

from eulersf_amd import _abi
from oracle import oracle as O

N = 17
DH = 1.0 / 16
GRID = (N, N, N, DH, (0.0, 0.0, 0.0))


def node_positions():
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return np.stack([x, y, z], axis=-1).reshape(-1, 3) * DH


def sphere_points(rng, n, c, r):
    d = rng.standard_normal((n, 3))
    return c + r * d / np.linalg.norm(d, axis=1, keepdims=True)


def test_constant_displacement_lands_on_rest_plus_shift():
    """u = c everywhere: F = I, one Newton step puts x at rest + c"""
    rng = np.random.default_rng(1)
    c = np.array([0.013, -0.021, 0.007])
    u = np.tile(c, (N ** 3, 1)).reshape(-1)
    vel = rng.standard_normal(3 * N ** 3) * 0.3
    rest = sphere_points(rng, 50, np.array([0.5, 0.5, 0.5]), 0.2)
    start = rest + c + rng.standard_normal(rest.shape) * 0.01
    out = O.advect_surface_mesh(GRID, vel, u, 1.0 / 240, rest, start)
    np.testing.assert_allclose(out, rest + c, rtol=0, atol=1e-14)


def test_linear_displacement_solves_the_fixed_point():
    """u = A x + c: trilinear interpolation and the cell's F are exact, so
    the Newton steps reach x = (I - A)^-1 (rest + c)"""
    rng = np.random.default_rng(2)
    A = rng.standard_normal((3, 3)) * 0.05
    c = np.array([0.01, 0.02, -0.015])
    u = (node_positions() @ A.T + c).reshape(-1)
    vel = np.zeros(3 * N ** 3)
    rest = sphere_points(rng, 80, np.array([0.5, 0.45, 0.52]), 0.25)
    exact = np.linalg.solve(np.eye(3) - A, (rest + c).T).T
    start = exact + rng.standard_normal(rest.shape) * 0.02
    out = O.advect_surface_mesh(GRID, vel, u, 1.0 / 240, rest, start)
    np.testing.assert_allclose(out, exact, rtol=0, atol=1e-12)
    # a wrong sign of the Newton step would walk away instead
    assert np.max(np.abs(start - exact)) > 1e-3


def test_velocity_moves_vertices_at_the_fixed_point():
    """u = 0 and rest at the vertices: the velocity move is undone by the
    projection back to x - u = rest"""
    rng = np.random.default_rng(3)
    rest = sphere_points(rng, 30, np.array([0.5, 0.5, 0.5]), 0.2)
    vel = np.tile([0.5, -0.2, 0.1], (N ** 3, 1)).reshape(-1)
    out = O.advect_surface_mesh(GRID, vel, np.zeros(3 * N ** 3), 1.0 / 240, rest, rest)
    np.testing.assert_allclose(out, rest, rtol=0, atol=1e-14)


def test_vertices_clamped_into_the_domain():
    rest = np.array([[1.2, 0.5, 0.5], [-0.3, 0.5, 0.5], [0.5, 0.5, 2.0]])
    out = O.advect_surface_mesh(GRID, np.zeros(3 * N ** 3), np.zeros(3 * N ** 3), 0.01, rest, rest)
    np.testing.assert_array_equal(out, [[1.0, 0.5, 0.5], [0.0, 0.5, 0.5], [0.5, 0.5, 1.0]])


def tri_boxes(v, t):
    p = v[t]  # (m, 3, 3)
    return p.min(axis=1), p.max(axis=1)


def brute_force(meshes, grid):
    X, Y, Z, dh, o = grid
    o = np.array(o)
    res = np.array([X, Y, Z])
    lo, hi = o + 2.0 * dh, o + (res - 3) * dh
    for v, t in meshes:
        mn, mx = tri_boxes(v, t)
        bmin, bmax = mn.min(axis=0), mx.max(axis=0)
        if np.any(bmin < lo) or np.any(bmin > hi) or np.any(bmax < lo) or np.any(bmax > hi):
            return 0.6
    for a in range(len(meshes)):
        for b in range(a + 1, len(meshes)):
            amn, amx = tri_boxes(*meshes[a])
            bmn, bmx = tri_boxes(*meshes[b])
            ov = np.all((amn[:, None, :] <= bmx[None, :, :]) & (amx[:, None, :] >= bmn[None, :, :]),
                        axis=2)
            if ov.any():
                return 0.6
    return 1.8


def soup(rng, centre, scale, m):
    v = centre + rng.standard_normal((3 * m, 3)) * scale
    return v, np.arange(3 * m).reshape(m, 3)


@pytest.mark.parametrize("seed", range(40))
def test_bvh_broad_phase_equals_all_pairs(seed):
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, 4))
    meshes = [soup(rng, rng.uniform(0.25, 0.75, 3), rng.uniform(0.005, 0.05),
                   int(rng.integers(1, 60))) for _ in range(k)]
    assert O.mesh_broad_phase(GRID, meshes) == brute_force(meshes, GRID)


def test_broad_phase_cases():
    rng = np.random.default_rng(7)
    a = soup(rng, np.array([0.3, 0.5, 0.5]), 0.01, 20)
    b = soup(rng, np.array([0.7, 0.5, 0.5]), 0.01, 20)
    assert O.mesh_broad_phase(GRID, [a, b]) == 1.8
    c = (a[0] + [0.4, 0.0, 0.0], a[1])  # on top of b
    assert O.mesh_broad_phase(GRID, [c, b]) == 0.6
    # a vertex 1.9 cells from the x = 0 wall
    w = (np.vstack([a[0], [[1.9 * DH, 0.5, 0.5]]]), np.vstack([a[1], [[0, 1, 60]]]))
    assert O.mesh_broad_phase(GRID, [w]) == 0.6
    w2 = (np.vstack([a[0], [[2.0 * DH, 0.5, 0.5]]]), np.vstack([a[1], [[0, 1, 60]]]))
    assert O.mesh_broad_phase(GRID, [w2]) == 1.8


def test_obj_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    v = rng.standard_normal((40, 3))
    t = rng.integers(0, 40, (70, 3)).astype(np.int32)
    p = str(tmp_path / "m.obj")
    _abi.obj_write(p, v, t)
    v2, t2 = _abi.obj_read(p)
    np.testing.assert_array_equal(v2.view(np.uint64), v.view(np.uint64))
    np.testing.assert_array_equal(t2, t)


def test_obj_reader_forms(tmp_path):
    """v/t/n index forms, negative (relative) indices, faces of more than
    three vertices (their first three), comments and other records"""
    p = tmp_path / "m.obj"
    p.write_text("# comment\nv 0 0 0\nv 1 0 0\nvt 0 0\nv 0 1 0\nvn 0 0 1\nv 1 1 0\n"
                 "g grp\nf 1/1/1 2/1/1 3/1/1\nf 2//1 4//1 3//1\nf -4 -3 -2 -1\n")
    v, t = _abi.obj_read(str(p))
    assert v.shape == (4, 3)
    np.testing.assert_array_equal(t, [[0, 1, 2], [1, 3, 2], [0, 1, 2]])
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(_abi.EsfError):
        _abi.obj_read(str(bad))
    with pytest.raises(_abi.EsfError):
        _abi.obj_read(str(tmp_path / "absent.obj"))


SPHERE_OBJ = "/root/reference/meshes/sphere/sphere.obj"


@pytest.mark.skipif(not os.path.exists(SPHERE_OBJ), reason="reference mesh not present")
def test_reads_the_reference_sphere_mesh():
    v, t = _abi.obj_read(SPHERE_OBJ)
    nv = sum(1 for l in open(SPHERE_OBJ) if l.startswith("v "))
    nf = sum(1 for l in open(SPHERE_OBJ) if l.startswith("f "))
    assert v.shape == (nv, 3) and t.shape == (nf, 3)
    assert t.min() >= 0 and t.max() < nv
