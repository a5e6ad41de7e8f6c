"""Surface meshes on the GPU path against the oracle: advectSurfaceMesh
(DEFOR_SOLID_3D.cpp:1125-1180) bit-exact on the same fields, the broad
phase (COUPLED_INTEGRATOR_3D.cpp:384-411) equal to the oracle's BVH
collide, and the mesh the scene loader places."""
import os

import numpy as np
import pytest

from eulersf_amd.cfg import scene_cfg
from oracle import oracle as O
from tests._scenes import ball_cfg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def esf():
    import eulersf_amd
    eulersf_amd.load()
    return eulersf_amd


def grid_of(n):
    return (n + 1, n + 1, n + 1, 1.0 / n, (0.0, 0.0, 0.0))


def test_loader_places_sphere_mesh(esf):
    g = esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(16))
    v, rest, t = g.surfaceMesh(0)
    assert v.shape == (162, 3) and t.shape == (320, 3)  # icosahedron split twice
    np.testing.assert_allclose(np.linalg.norm(rest, axis=1), 0.15, rtol=1e-14)
    np.testing.assert_array_equal(v, rest + np.array([0.5, 0.65, 0.5]))
    g.close()


@pytest.mark.parametrize("perturb", [0.0, 0.02])
def test_advect_surface_mesh_matches_oracle(esf, perturb):
    n = 16
    g = esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(n, velocity=(0.3, -1.0, 0.2)))
    A = esf._abi
    for _ in range(3):
        g.stepImplicit(1.8)
    v, rest, t = g.surfaceMesh(0)
    if perturb:
        v = v + np.random.default_rng(1).standard_normal(v.shape) * perturb
        g.setSurfaceMesh(0, rest, v, t)
    u = g.download(A.S_U, 0)
    vel = g.download(A.S_VELOCITY, 0)
    dt = g.dt()
    expect = O.advect_surface_mesh(grid_of(n), vel, u, dt, rest, v)
    g.advectSurfaceMesh()
    got = g.surfaceMesh(0)[0]
    np.testing.assert_array_equal(got.view(np.uint64), expect.view(np.uint64))
    assert np.max(np.abs(got - v)) > 0
    g.close()


SCENES = {
    "far": [{"center": (0.3, 0.5, 0.5), "radius": 0.1}, {"center": (0.7, 0.5, 0.5), "radius": 0.1}],
    "touching": [{"center": (0.42, 0.5, 0.5), "radius": 0.1},
                 {"center": (0.58, 0.5, 0.5), "radius": 0.1}],
    "near_wall": [{"center": (0.5, 0.5, 0.5), "radius": 0.1},
                  {"center": (0.5, 0.5, 0.86), "radius": 0.1}],
    "one": [{"center": (0.5, 0.4, 0.5), "radius": 0.15}],
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_broad_phase_matches_oracle(esf, name):
    n = 24
    g = esf.CoupledIntegrator3D.from_cfg_text(scene_cfg(n, SCENES[name]))
    meshes = []
    for s in range(len(SCENES[name])):
        v, _, t = g.surfaceMesh(s)
        meshes.append((v, t))
    expect = O.mesh_broad_phase(grid_of(n), meshes)
    assert g.broadPhase() == expect
    assert expect == (1.8 if name in ("far", "one") else 0.6)
    g.close()


def test_frame_round_trip_keeps_mesh(esf, tmp_path):
    text = ball_cfg(16, velocity=(0.3, -1.0, 0.2))
    g = esf.CoupledIntegrator3D.from_cfg_text(text)
    for _ in range(2):
        g.stepImplicit(1.8)
    path = str(tmp_path) + os.sep
    g.writeFrame(path, 20)
    assert os.path.exists(path + "frame.0020.solid.0.obj")
    h = esf.CoupledIntegrator3D.from_cfg_text(text)
    h.readFrame(path, 20)
    a, b = g.surfaceMesh(0), h.surfaceMesh(0)
    np.testing.assert_array_equal(a[0].view(np.uint64), b[0].view(np.uint64))
    np.testing.assert_array_equal(a[2], b[2])
    g.close()
    h.close()
