"""esf_compute_sdf_noalloc, the allocation-free variant of esf_compute_sdf
(its staging and counters in the library's static per-GPU scratch): the
same phi and boxes as esf_compute_sdf, bitwise, on two calls in a row, the
second on a smaller grid; a grid larger than the scratch is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def esf():
    import eulersf_amd
    eulersf_amd.load()
    return eulersf_amd


def ball_mass(n, center, radius):
    N = n + 1
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    d = np.sqrt((x / n - center[0]) ** 2 + (y / n - center[1]) ** 2 + (z / n - center[2]) ** 2)
    return (d < radius).astype(np.float64) * 1000.0


def test_noalloc_sdf_equals_allocating_entry_point(esf):
    print("noalloc: 40^3 then 20^3 grid", flush=True)
    for n, center, radius in ((40, (0.5, 0.6, 0.45), 0.2), (20, (0.4, 0.5, 0.55), 0.25)):
        g = esf.CoupledIntegrator3D.from_cfg_text(esf.unit_cube_ball_cfg(n))
        m = ball_mass(n, center, radius)
        phi_a, box_a, tight_a = g.computeSDF(m)
        phi_b, box_b, tight_b = g.computeSDFNoAlloc(m)
        assert np.array_equal(phi_a.view(np.int64), phi_b.view(np.int64)), n
        assert list(box_a) == list(box_b) and list(tight_a) == list(tight_b), n
        # twice on the same context: the counters start afresh every call
        phi_c, box_c, tight_c = g.computeSDFNoAlloc(m)
        assert np.array_equal(phi_a.view(np.int64), phi_c.view(np.int64)), n
        assert list(box_a) == list(box_c) and list(tight_a) == list(tight_c), n
        g.close()


def test_noalloc_sdf_refuses_a_grid_beyond_its_scratch(esf):
    import ctypes
    from eulersf_amd import _abi
    print("noalloc: 130 x 129 x 129 nodes refused", flush=True)
    grid = {"xres": 130, "yres": 129, "zres": 129, "dh": 1.0 / 128, "origin": (0.0, 0.0, 0.0)}
    from eulersf_amd.cfg import Cfg, params_from_cfg
    params = params_from_cfg(Cfg(text=esf.unit_cube_ball_cfg(128)), 1.0 / 128)
    g = esf.CoupledIntegrator3D(grid, params)
    m = np.zeros(130 * 129 * 129)
    phi = np.empty_like(m)
    box = np.zeros(6, dtype=np.int32)
    tight = np.zeros(6, dtype=np.int32)
    st = g.lib.esf_compute_sdf_noalloc(g.ctx, _abi.dptr(m), _abi.dptr(phi),
                                       box.ctypes.data_as(_abi._I), tight.ctypes.data_as(_abi._I))
    assert st == _abi.ESF_ERR_INVALID
    assert b"scratch" in g.lib.esf_last_error(g.ctx)
    g.close()
