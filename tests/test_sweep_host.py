"""The oracle's fast sweeping (computeSDF / sweepPhi / solveDistance and
extendVectorField, DEFOR_SOLID_3D.cpp:216-267, 369-574) pinned by what the
mathematics fixes: the grid Eikonal update's closed forms in one, two and
three dimensions, distances along the axes from a single massive node, the
box arithmetic, and for the extension: massive nodes untouched, constants
kept, and a field that varies only across a massive plane carried along the
axis that phi grows on."""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("a", [0.0, 0.5, 2.25])
def test_solve_distance_closed_forms(a):
    big = 1e6
    # one upwind neighbour: a + 1; two equal: a + 1/sqrt(2); three equal: a + 1/sqrt(3)
    assert O.solve_distance(a, big, big, 1e9) == pytest.approx(a + 1, abs=1e-14)
    assert O.solve_distance(a, a, big, 1e9) == pytest.approx(a + 1 / math.sqrt(2), abs=1e-14)
    assert O.solve_distance(a, a, a, 1e9) == pytest.approx(a + 1 / math.sqrt(3), abs=1e-14)
    # two neighbours a full step apart: the 1D value wins (the other is not upwind)
    assert O.solve_distance(a, a + 1, big, 1e9) == pytest.approx(a + 1, abs=1e-14)


def test_solve_distance_is_symmetric_and_only_lowers():
    rng = np.random.default_rng(7)
    for _ in range(200):
        p, q, r = rng.uniform(0, 5, 3)
        v = O.solve_distance(p, q, r, 1e9)
        # symmetric up to the rounding of p + q + r (the middle value is formed
        # from that sum, in argument order)
        assert O.solve_distance(q, r, p, 1e9) == pytest.approx(v, abs=1e-14)
        assert O.solve_distance(r, p, q, 1e9) == pytest.approx(v, abs=1e-14)
        assert O.solve_distance(p, q, r, v / 2) == v / 2
        # the Eikonal solution satisfies sum over upwind (v - d)^2 = 1
        ds = [d for d in (p, q, r) if d < v]
        assert sum((v - d) ** 2 for d in ds) == pytest.approx(1.0, abs=1e-12)


def test_compute_sdf_single_node():
    n, c, tb, ex = 20, (9, 8, 10), 3, 5
    m = np.zeros((n, n, n))
    m[c[2], c[1], c[0]] = 1.0
    phi, box, tight = O.compute_sdf(m, tb, ex)
    assert list(box) == [c[0] - ex, c[1] - ex, c[2] - ex, c[0] + ex, c[1] + ex, c[2] + ex]
    assert list(tight) == [c[0] - tb, c[1] - tb, c[2] - tb, c[0] + tb, c[1] + tb, c[2] + tb]
    x, y, z = c
    assert phi[z, y, x] == -0.5
    for k in range(1, ex):  # inside the box along each axis: k - 0.5
        assert phi[z, y, x + k] == pytest.approx(k - 0.5, abs=1e-14)
        assert phi[z, y - k, x] == pytest.approx(k - 0.5, abs=1e-14)
        assert phi[z + k, y, x] == pytest.approx(k - 0.5, abs=1e-14)
    # the diagonal neighbours by the 2D and 3D updates
    d2 = 0.5 + 1 / math.sqrt(2)
    assert phi[z, y + 1, x + 1] == pytest.approx(d2, abs=1e-14)
    assert phi[z + 1, y + 1, x + 1] == pytest.approx(d2 + 1 / math.sqrt(3), abs=1e-14)
    # outside the swept box the large initial distance stays
    assert phi[z, y, x + ex] == 3 * n + 2


def test_compute_sdf_box_clamps_to_the_grid():
    n = 16
    m = np.zeros((n, n, n))
    m[1, 14, 2] = 1.0
    _, box, tight = O.compute_sdf(m, 3, 20)
    assert list(box) == [0, 0, 0, n, n, n]
    assert list(tight) == [0, 11, 0, 5, n, 4]


def _slab_scene(n=14, plane=3):
    m = np.zeros((n, n, n))
    m[:, :, : plane + 1] = 1.0
    phi, box, _ = O.compute_sdf(m, 3, 20)
    return m, phi, box


def test_extend_keeps_massive_nodes_and_constants():
    m, phi, box = _slab_scene()
    rng = np.random.default_rng(3)
    f = rng.normal(size=m.shape + (3,))
    out = O.extend_vector_field(m, phi, box, f)
    assert np.array_equal(out[m > 0], f[m > 0])
    c = np.empty_like(f)
    c[...] = (0.3, -1.25, 2.0)
    oc = O.extend_vector_field(m, phi, box, c)
    assert np.abs(oc - c).max() <= 4e-16 * 2.0


def test_extend_carries_the_plane_along_x():
    """phi grows along x only, so every pass takes alpha = 1 (the x
    neighbour): the field on the last massive plane is carried out along x"""
    n, plane = 14, 3
    m, phi, box = _slab_scene(n, plane)
    yy, zz = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    f = np.zeros((n, n, n, 3))
    f[:, :, plane, 0] = yy + 10 * zz
    f[:, :, plane, 1] = -zz
    f[:, :, plane, 2] = 0.5 * yy
    out = O.extend_vector_field(m, phi, box, f)
    inner = slice(1, n - 1)
    for x in range(plane + 1, n - 1):
        assert np.array_equal(out[inner, inner, x], f[inner, inner, plane]), x


def test_whole_box_jacobi_settles_the_set_the_mass_pass_reads():
    """The device's phi is 8 whole-box Jacobi relaxations of the upwind
    update (bitwise, checked on the GPU).  Against the oracle's two rounds of
    ordered sweeps on a ball deformed by steps: the phi > 2 set over the
    tight box -- all that computeSolidMass(true) reads -- is equal from 5
    relaxations on, values below 3 agree to 1e-4 at 8, and further out the
    relaxation is an upper bound of the swept value.  Scalar solveDistance
    and the numpy restatement the check uses agree bitwise."""
    from eulersf_amd.cfg import unit_cube_ball_cfg
    from tests._scenes import _solve_distance, jacobi_phi, tight_box_phi_above_2
    rng = np.random.default_rng(11)
    p, q, r, v = rng.uniform(-0.5, 4.0, (4, 300))
    v[::3] = 1e9
    ref = np.array([O.solve_distance(*t) for t in zip(p, q, r, v)])
    np.testing.assert_array_equal(_solve_distance(p, q, r, v).view(np.int64), ref.view(np.int64))
    n = 24
    o = O.OracleScene.from_cfg_text(unit_cube_ball_cfg(n, eps=1e-8, velocity=(0.4, -1.2, 0.3)))
    try:
        for _ in range(3):
            o.step(1.8)
        X = Y = Z = n + 1
        b = o.boxes(0)
        ext = b[6:]
        po = o.field(O.S_PHI, 0).reshape(Z, Y, X)
        box_po = po[ext[2]:ext[5], ext[1]:ext[4], ext[0]:ext[3]]
        mass = o.field(O.S_MASS, 0)
        for k in (5, 6, 8):
            pj = np.full((Z, Y, X), float(X + Y + Z + 2))
            pj[ext[2]:ext[5], ext[1]:ext[4], ext[0]:ext[3]] = jacobi_phi(mass, ext, X, Y, Z, k)
            np.testing.assert_array_equal(tight_box_phi_above_2(pj, b, X, Y, Z),
                                          tight_box_phi_above_2(po, b, X, Y, Z))
        got = pj[ext[2]:ext[5], ext[1]:ext[4], ext[0]:ext[3]]
        near = box_po <= 3.0
        assert near.sum() > 100
        assert np.abs(got[near] - box_po[near]).max() < 1e-4
        assert np.all(got >= box_po - 1e-12)
    finally:
        o.close()


def test_relaxation_near_the_tight_box_equals_the_whole_box_there():
    """A step runs relaxation k (of R) only at the nodes within R - 1 - k of
    the tight box, on two buffers as the device does (a node it skips keeps
    the buffer's older value).  Inside the tight box the result is the R
    whole-box relaxations' bit for bit, since a node's value after the
    remaining relaxations reads no node further out; one node less of margin
    already changes it."""
    from tests._scenes import _solve_distance, jacobi_phi
    rng = np.random.default_rng(4)
    X = Y = Z = 40
    mass = np.zeros((Z, Y, X))
    zz, yy, xx = np.mgrid[0:Z, 0:Y, 0:X]
    mass[(xx - 19.3) ** 2 + (yy - 21.1) ** 2 + (zz - 18.7) ** 2 < 30] = 1.0
    mass[rng.random(mass.shape) < 0.002] = 1.0
    ext = [2, 3, 1, 38, 39, 37]
    tight = [12, 13, 11, 27, 29, 26]
    R = 8
    want = jacobi_phi(mass.reshape(-1), ext, X, Y, Z, R)
    large = float(X + Y + Z + 2)
    m = mass[ext[2]:ext[5], ext[1]:ext[4], ext[0]:ext[3]]
    free = m == 0
    lo = [tight[k] - ext[k] for k in range(3)]
    hi = [tight[3 + k] - ext[k] for k in range(3)]
    z, y, x = np.mgrid[0:m.shape[0], 0:m.shape[1], 0:m.shape[2]]

    def run(extra):
        buf = [np.where(m > 0, -0.5, large), np.where(m > 0, -0.5, large)]
        for k in range(R):
            mg = R - 1 - k - extra
            src, dst = buf[k & 1], buf[1 - (k & 1)]
            pad = np.pad(src, 1, constant_values=large)
            p = np.fmin(pad[1:-1, 1:-1, :-2], pad[1:-1, 1:-1, 2:])
            q = np.fmin(pad[1:-1, :-2, 1:-1], pad[1:-1, 2:, 1:-1])
            r = np.fmin(pad[:-2, 1:-1, 1:-1], pad[2:, 1:-1, 1:-1])
            near = ((x >= lo[0] - mg) & (x < hi[0] + mg) & (y >= lo[1] - mg) & (y < hi[1] + mg) &
                    (z >= lo[2] - mg) & (z < hi[2] + mg))
            upd = free & near
            dst[upd] = _solve_distance(p, q, r, src)[upd]
        return buf[R & 1]

    t = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
    got = run(0)
    np.testing.assert_array_equal(got[t].view(np.int64), want[t].view(np.int64))
    assert np.any(got != want)
    assert np.any(run(1)[t] != want[t])
