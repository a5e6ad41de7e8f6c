"""The FLIP material particles on the GPU against the oracle
(DEFOR_SOLID_3D.cpp:333-367, 891-1107): the particles a scene starts with,
gridToParticle and particleToGrid from the same particles and fields, and
whole substeps with the particle round trip, the velocity extension and the
swept phi (the reference's full finalizeSolution), teacher-forced."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._scenes import rel_err, teacher_force
from tests.test_gpu_configs import compare_step

pytestmark = pytest.mark.gpu

FLIP = {"flip particles": 1, "number of particles per material cell": 8, "flip alpha": 0.99,
        "wall thickness": 0.0625}


@pytest.fixture(scope="module")
def esf():
    import eulersf_amd
    eulersf_amd.load()
    return eulersf_amd


def pair(esf, n=24, **kw):
    text = esf.unit_cube_ball_cfg(n, eps=1e-13, max_iterations=20000, extra=FLIP, **kw)
    return esf.CoupledIntegrator3D.from_cfg_text(text), O.OracleScene.from_cfg_text(text)


def test_initial_particles_bitwise(esf):
    g, o = pair(esf, velocity=(0.3, -1.0, 0.2))
    pg, vg = g.particles(0)
    po, vo = o.particles(0)
    print(f"flip: {len(pg)} particles", flush=True)
    assert len(pg) == len(po) > 0
    np.testing.assert_array_equal(pg, po)
    np.testing.assert_array_equal(vg, vo)
    g.close()
    o.close()


def test_grid_to_particle_and_back_match_oracle(esf):
    """G2P and P2G from the same particles and fields: particle positions and
    velocities, the solid velocity and the 0/1 mass within rounding (the
    B-spline's pow and the gather's summation order)"""
    A = esf._abi
    g, o = pair(esf, velocity=(0.3, -1.0, 0.2))
    rng = np.random.default_rng(30)
    X = 25
    vel = rng.standard_normal((X ** 3, 3)) * 0.5
    ws = rng.standard_normal((X ** 3, 3)) * 0.5
    for f, of, v in ((A.S_VELOCITY, O.S_VELOCITY, vel), (A.S_VEL_WORKSPACE, O.S_VEL_WORKSPACE, ws)):
        g.upload(f, v.reshape(-1), 0)
        o.set_field(of, v.reshape(-1), 0)
    # particles near the floor too (wall impulses)
    pos, pv = o.particles(0)
    pos = pos.copy()
    pos[:50, 1] = rng.uniform(0.0, 0.06, 50)
    o.set_particles(pos, pv)
    g.setParticles(pos, pv)
    g.gridToParticle(0)
    o.L.orc_grid_to_particle(o.h, 0)
    pg, vg = g.particles(0)
    po, vo = o.particles(0)
    assert rel_err(pg, po) < 1e-14 and rel_err(vg, vo) < 1e-13, (rel_err(pg, po), rel_err(vg, vo))
    g.particleToGrid(0)
    o.L.orc_particle_to_grid(o.h, 0)
    np.testing.assert_array_equal(g.download(A.S_MASS, 0), o.field(O.S_MASS, 0))
    assert rel_err(g.download(A.S_VELOCITY, 0), o.field(O.S_VELOCITY, 0)) < 1e-13
    # and the extension of that velocity over the box by phi, bitwise from
    # the same inputs
    g.upload(A.S_VELOCITY, o.field(O.S_VELOCITY, 0), 0)
    g.extendSolidVelocity(0)
    o.L.orc_extend_solid_velocity(o.h, 0)
    np.testing.assert_array_equal(g.download(A.S_VELOCITY, 0), o.field(O.S_VELOCITY, 0))
    g.close()
    o.close()


def test_flip_substeps_match_oracle(esf):
    """whole substeps with the particle round trip (24^3 ball, thrown):
    three steps, each from the oracle's state, every field within 1e-10,
    the new particles bitwise"""
    g, o = pair(esf, velocity=(0.3, -1.0, 0.2))
    for k in range(3):
        if k:
            teacher_force(g, o, particles=True)
        gs = g.stepImplicit(1.8)
        rc, st = o.step(1.8)
        assert gs["status"] == 0 and rc == 0, (k, gs, rc)
        assert gs["dt"] == st.dt
        compare_step(esf, g, o, k, 1e-10)
        pg, vg = g.particles(0)
        po, vo = o.particles(0)
        assert len(pg) == len(po)
        np.testing.assert_array_equal(pg, po)
        assert rel_err(vg, vo) < 1e-10
        np.testing.assert_array_equal(g.download(esf._abi.S_PHI, 0), o.field(O.S_PHI, 0))
    g.close()
    o.close()


def test_flip_refuses_several_solids(esf):
    from eulersf_amd.cfg import scene_cfg
    spheres = [{"center": (0.3, 0.6, 0.5), "radius": 0.1}, {"center": (0.7, 0.6, 0.5), "radius": 0.1}]
    text = scene_cfg(24, spheres, extra=FLIP)
    with pytest.raises(esf.EsfError):
        esf.CoupledIntegrator3D.from_cfg_text(text)
