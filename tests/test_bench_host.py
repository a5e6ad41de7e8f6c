"""Host-side pieces of bench.py that feed the measured workloads: the
random divergence field of the isolated-system benchmark is J u, checked
against the oracle's J (COUPLED_INTEGRATOR_3D.cpp:239-249)."""
import numpy as np

import bench
from eulersf_amd.cfg import random_spheres, scene_cfg
from oracle import oracle as O
from tests._scenes import solution_order


def test_divergence_field_is_the_oracle_J():
    n = 8
    sc = O.OracleScene.from_cfg_text(scene_cfg(n, [{"center": (0.5, 0.5, 0.5), "radius": 0.2}]))
    rng = np.random.default_rng(5)
    X = n + 1
    u = np.zeros((X ** 3, 3))
    inner = np.zeros((X, X, X), dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    u[inner.reshape(-1)] = rng.standard_normal((int(inner.sum()), 3))
    b = bench.divergence_field(u, n)
    ref = sc.apply_J(solution_order(u.reshape(-1), X, X, X))
    np.testing.assert_allclose(b, ref, rtol=0, atol=1e-12)
    sc.close()


def test_random_spheres_are_disjoint_and_inside():
    rng = np.random.default_rng(2026)
    sp = random_spheres(rng, 8, (0.05, 0.12), (3.0, 6.0), center_range=(0.15, 0.85))
    assert len(sp) == 8
    for i, a in enumerate(sp):
        c = np.asarray(a["center"])
        assert np.all(c - a["radius"] > 0) and np.all(c + a["radius"] < 1)
        assert 1e3 <= a["E"] <= 1e6
        for bsp in sp[i + 1:]:
            assert np.linalg.norm(c - np.asarray(bsp["center"])) > a["radius"] + bsp["radius"]
    # the cfg carries every sphere
    text = scene_cfg(32, sp, fill=0.6)
    assert "number of solids = 8" in text and "solid translation 7" in text
