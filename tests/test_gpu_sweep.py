"""The GPU fast sweeping (esf_compute_sdf, esf_extend_vector_field) against
the oracle's restatement of the reference's ordered sweeps: bitwise, on two
solids whose distance field has a ridge between them (so sweepOnePass's
solvable flags clear part-way through passes)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._scenes import ball_cfg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def esf():
    import eulersf_amd
    eulersf_amd.load()
    return eulersf_amd


def two_balls(n, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m = np.zeros((n, n, n))
    for c, r in (((0.3, 0.45, 0.5), 0.14), ((0.7, 0.55, 0.45), 0.11)):
        d = np.sqrt((x / (n - 1) - c[0]) ** 2 + (y / (n - 1) - c[1]) ** 2 + (z / (n - 1) - c[2]) ** 2)
        m[d < r] = rng.uniform(0.5, 2.0, int((d < r).sum()))
    return m


@pytest.mark.parametrize("n,seed", [(24, 1), (33, 2)])
def test_compute_sdf_bitwise(esf, n, seed):
    g = esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(n - 1))
    m = two_balls(n, seed)
    phi, box, tight = g.computeSDF(m)
    # the scene cfg's "tight border" and "extrapolation cells"
    phi_o, box_o, tight_o = O.compute_sdf(m, 5, 20)
    assert list(box) == list(box_o) and list(tight) == list(tight_o)
    assert np.array_equal(phi.view(np.int64), phi_o.view(np.int64))
    g.close()


@pytest.mark.parametrize("n,seed,extrap", [(24, 3, 20), (33, 4, 4)])
def test_extend_vector_field_bitwise(esf, n, seed, extrap):
    g = esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(n - 1))
    m = two_balls(n, seed)
    phi, box, _ = O.compute_sdf(m, 3, extrap)
    f = np.random.default_rng(seed).normal(size=(n, n, n, 3))
    out = g.extendVectorField(box, m, phi, f)
    out_o = O.extend_vector_field(m, phi, box, f)
    assert np.array_equal(out.view(np.int64), out_o.view(np.int64))
    assert not np.array_equal(out, f)
    g.close()


def test_extend_refuses_a_thin_box(esf):
    n = 16
    g = esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(n - 1))
    m = np.zeros((n, n, n))
    with pytest.raises(esf._abi.EsfError):
        g.extendVectorField(np.array([2, 2, 2, 4, 10, 10]), m, np.zeros_like(m),
                            np.zeros((n, n, n, 3)))
    g.close()
