"""Frame files on the GPU path (COUPLED_INTEGRATOR_3D::writeFrame /
readFrame, COUPLED_INTEGRATOR_3D.cpp:976-1011; DEFOR_SOLID_3D.cpp:1299-1371):
what esf_write_frame puts on disk is the device state, and a fresh context
that reads the frame back continues bitwise as the original does."""
import os

import numpy as np
import pytest

from tests._scenes import ball_cfg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def esf():
    import eulersf_amd
    eulersf_amd.load()
    return eulersf_amd


def bits(a):
    """float64 arrays as their bit patterns (exact comparison, -0 != +0)"""
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def test_write_read_resume(esf, tmp_path):
    A = esf._abi
    text = ball_cfg(16, velocity=(0.3, -1.0, 0.2))
    g = esf.CoupledIntegrator3D.from_cfg_text(text)
    for _ in range(3):
        g.stepImplicit(1.8)
    path = str(tmp_path) + os.sep
    g.writeFrame(path, 10)
    g.writeFrame(path, 3)
    names = sorted(os.listdir(tmp_path))
    assert names == ["frame.0003.intensity", "frame.0003.solid.0.obj", "frame.0010.intensity",
                     "frame.0010.solid.0.displacement", "frame.0010.solid.0.obj",
                     "frame.0010.solid.0.phi", "frame.0010.solid.0.velocity",
                     "frame.0010.velocity"]
    grid, v = A.field_read_gz(path + "frame.0010.velocity", 3)
    assert grid[:3] == (17, 17, 17) and grid[3] == 1.0 / 16
    np.testing.assert_array_equal(bits(v), bits(g.download(A.F_VELOCITY)))
    _, u = A.field_read_gz(path + "frame.0010.solid.0.displacement", 3)
    np.testing.assert_array_equal(bits(u), bits(g.download(A.S_U, 0)))
    _, phi = A.field_read_gz(path + "frame.0010.solid.0.phi", 1)
    np.testing.assert_array_equal(bits(phi), bits(g.download(A.S_PHI, 0)))

    h = esf.CoupledIntegrator3D.from_cfg_text(text)
    h.readFrame(path, 10)
    for fid, solid in ((A.F_VELOCITY, -1), (A.F_INTENSITY, -1), (A.S_U, 0), (A.S_VELOCITY, 0),
                       (A.S_VEL_WORKSPACE, 0), (A.S_PHI, 0)):
        np.testing.assert_array_equal(bits(h.download(fid, solid)), bits(g.download(fid, solid)),
                                      err_msg=f"field {fid}")
    # readFrame's boxes come from the surface mesh's bounding box, and the
    # mass from computeSolidMass(true) on the read state
    # (DEFOR_SOLID_3D.cpp:1318-1357)
    verts = h.surfaceMesh(0)[0]
    lo = np.trunc(verts.min(axis=0) * 16).astype(int)
    hi = np.trunc(verts.max(axis=0) * 16).astype(int)
    assert h.boxes(0) == ([max(v - 5, 0) for v in lo] + [min(v + 5, 17) for v in hi] +
                          [max(v - 20, 0) for v in lo] + [min(v + 20, 17) for v in hi])
    # the original, reading its own frame, reaches the same state
    g.readFrame(path, 10)
    for fid, solid in ((A.F_SOLID_MASS, -1), (A.F_NODE_FRACTION, -1), (A.F_SOLID_INDEX, -1),
                       (A.S_MASS, 0), (A.S_VALID, 0), (A.S_QFRAC, 0)):
        np.testing.assert_array_equal(bits(h.download(fid, solid)), bits(g.download(fid, solid)),
                                      err_msg=f"field {fid}")
    # X as readFrame forms it, (x dh + origin) - u (DEFOR_SOLID_3D.cpp:1307-1313)
    n = 17
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pos = np.stack([x, y, z], axis=-1).reshape(-1, 3) * (1.0 / 16) + 0.0
    uu = g.download(A.S_U, 0).reshape(-1, 3)
    np.testing.assert_array_equal(bits(h.download(A.S_X, 0)), bits((pos - uu).reshape(-1)))
    np.testing.assert_array_equal(bits(h.download(A.S_X, 0)), bits(g.download(A.S_X, 0)))
    # the pressure warm start is not part of a frame (the reference keeps it
    # in a function-local static); hand it over, then both continue alike
    h.upload(A.F_LAMBDA, g.download(A.F_LAMBDA))
    for _ in range(2):
        rg = g.stepImplicit(1.8)
        rh = h.stepImplicit(1.8)
        assert rg["iterations"] == rh["iterations"]
    for fid, solid in ((A.F_VELOCITY, -1), (A.F_INTENSITY, -1), (A.S_U, 0), (A.S_X, 0)):
        np.testing.assert_array_equal(bits(h.download(fid, solid)), bits(g.download(fid, solid)))
    g.close()
    h.close()


def test_read_frame_errors(esf, tmp_path):
    g = esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(16))
    with pytest.raises(esf.EsfError):
        g.readFrame(str(tmp_path) + os.sep, 10)
    # a frame of another grid is refused
    other = esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(12))
    other.writeFrame(str(tmp_path) + os.sep, 20)
    with pytest.raises(esf.EsfError):
        g.readFrame(str(tmp_path) + os.sep, 20)
    g.close()
    other.close()
