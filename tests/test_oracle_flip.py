"""Known answers for the oracle's FLIP material particles
(DEFOR_SOLID_3D.cpp:333-367 initMaterialParticles, :891-1034
gridToParticle, :1036-1107 particleToGrid): the MT19937 draws (against
numpy's MT19937, itself checked against the C++ standard's known value), the
particle layout, the interpolated velocity, the FLIP/PIC blend under
uniform and linear fields (the cubic B-spline's partition of unity and
linear reproduction), the wall impulse, and the cutoff of the gather."""
import numpy as np

from oracle import oracle as O
from tests.test_oracle_pins import LAM, MU, block_scene, node_pos


def mt19937_draws(seed, k):
    """the first k outputs of MT19937 (Matsumoto-Nishimura init_genrand)"""
    return np.random.RandomState(seed).randint(0, 2 ** 32, size=k, dtype=np.uint32)


def test_mt19937_reference_values():
    """numpy's MT19937 is the standard generator: the 10000th output of the
    default seed 5489 is 4123659995 (the C++ standard's value for
    std::mt19937)"""
    assert mt19937_draws(5489, 10000)[-1] == 4123659995


def bspline(d):
    a = np.abs(d)
    return np.where(a < 1, 0.5 * a ** 3 - d * d + 2.0 / 3,
                    np.where(a < 2, -1.0 / 6 * a ** 3 + d * d - 2 * a + 4.0 / 3, 0.0))


def test_particles_layout_offsets_and_velocity():
    """8 particles per valid cell, cells in z, y, x order; particle k sits at
    the cell's min corner plus dh * (r_3k, r_3k+1, r_3k+2) with r = MT19937(123456)
    output / (2^32 - 1); its velocity is the solid velocity field there
    (linear field: exact)"""
    n = 8
    dh = 1.0 / n
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    X = n + 1
    rng = np.random.default_rng(20)
    A = rng.standard_normal((3, 3))
    b = rng.standard_normal(3)
    s.set_field(O.S_VELOCITY, (node_pos(n) @ A.T + b).reshape(-1), 0)
    s.L.orc_init_particles(s.h, 0)
    pos, vel = s.particles(0)
    valid = s.field(O.S_VALID, 0).reshape(n, n, n)
    cells = np.argwhere(valid == 1)  # (z, y, x) rows in z, y, x order
    assert len(pos) == 8 * len(cells) == 8 * 27
    r = mt19937_draws(123456, 3 * len(pos)).astype(np.float64) * (1.0 / 4294967295.0)
    corner = np.repeat(cells[:, ::-1] * dh, 8, axis=0)
    np.testing.assert_array_equal(pos, corner + r.reshape(-1, 3) * dh)
    np.testing.assert_allclose(vel, pos @ A.T + b, rtol=0, atol=1e-13)
    s.close()


def test_grid_to_particle_blend_and_move():
    """deep inside the solid (all 64 stencil nodes massive): uniform grid
    velocity V and last-step solid velocity W give (1 - a) V + a (V - W + v_p)
    (the B-spline weights sum to 1); a linear grid velocity with W = 0 gives
    (1 - a) v(x_p) + a (v(x_p) + v_p) (linear functions reproduced); then
    x_p += v_p dt"""
    n = 16
    dh = 1.0 / n
    alpha, dt = 0.95, 1.0 / 240
    s = block_scene(n, (3 * dh, 3 * dh, 3 * dh), (13 * dh, 13 * dh, 13 * dh), flip_alpha=alpha)
    X = n + 1
    s.L.orc_set_dt(s.h, dt)
    s.set_field(O.VELOCITY, np.zeros(3 * X ** 3))
    s.L.orc_solid_copy_velocity(s.h)  # the solid's dt
    rng = np.random.default_rng(21)
    k = 200
    p0 = rng.uniform(6 * dh, 10 * dh, (k, 3))
    v0 = rng.standard_normal((k, 3))
    V, W = np.array([0.4, -0.3, 0.2]), np.array([-0.1, 0.05, 0.3])
    s.set_particles(p0, v0)
    s.set_field(O.S_VELOCITY, np.tile(V, X ** 3), 0)
    s.set_field(O.S_VEL_WORKSPACE, np.tile(W, X ** 3), 0)
    s.L.orc_grid_to_particle(s.h, 0)
    pos, vel = s.particles(0)
    want = (1 - alpha) * V + alpha * (V - W + v0)
    np.testing.assert_allclose(vel, want, rtol=0, atol=1e-13)
    np.testing.assert_allclose(pos, p0 + want * dt, rtol=0, atol=1e-15)
    A = rng.standard_normal((3, 3))
    s.set_particles(p0, v0)
    s.set_field(O.S_VELOCITY, (node_pos(n) @ A.T).reshape(-1), 0)
    s.set_field(O.S_VEL_WORKSPACE, np.zeros(3 * X ** 3), 0)
    s.L.orc_grid_to_particle(s.h, 0)
    pos, vel = s.particles(0)
    vx = p0 @ A.T
    np.testing.assert_allclose(vel, (1 - alpha) * vx + alpha * (vx + v0), rtol=0, atol=1e-12)
    s.close()


def test_grid_to_particle_skips_massless_nodes():
    """a node without solid mass drops out of the stencil, and the weights
    are not renormalised: a particle beside the block's face sees only the
    massive nodes' share of V"""
    n = 16
    dh = 1.0 / n
    s = block_scene(n, (3 * dh, 3 * dh, 3 * dh), (13 * dh, 13 * dh, 13 * dh), flip_alpha=0.0)
    X = n + 1
    s.L.orc_set_dt(s.h, 1.0 / 240)
    s.set_field(O.VELOCITY, np.zeros(3 * X ** 3))
    s.L.orc_solid_copy_velocity(s.h)
    mass = s.field(O.S_MASS, 0).reshape(X, X, X)
    p = np.array([[3.3 * dh, 8.4 * dh, 8.7 * dh]])
    V = np.array([1.0, 2.0, -1.0])
    s.set_particles(p, np.zeros((1, 3)))
    s.set_field(O.S_VELOCITY, np.tile(V, X ** 3), 0)
    s.L.orc_grid_to_particle(s.h, 0)
    _, vel = s.particles(0)
    base = np.floor(p[0] / dh).astype(int) - 1
    wsum = 0.0
    for r in range(base[2], base[2] + 4):
        for t in range(base[1], base[1] + 4):
            for q in range(base[0], base[0] + 4):
                if mass[r, t, q] == 0:
                    continue
                d = p[0] / dh - np.array([q, t, r])
                wsum += bspline(d[0]) * bspline(d[1]) * bspline(d[2])
    assert 0.05 < wsum < 0.95  # part of the stencil lies outside the block
    np.testing.assert_allclose(vel[0], wsum * V, rtol=1e-13)
    s.close()


def test_wall_impulse():
    """a particle within the wall thickness of the floor moving into it
    (against the solid's velocity at the floor node beside it) gets
    dv_n = min(dt k overlap, m (0.1 overlap / dt - v_n)) / m along the
    normal and its tangential velocity scaled by max(0.1 - friction dv_n /
    |v_t|, 0) (DEFOR_SOLID_3D.cpp:996-1031); m = rho dh^3 / 8 per particle"""
    n = 16
    dh = 1.0 / n
    dt, ks, fr, rho = 1.0 / 240, 1e4, 0.5, 1000.0
    thick = 1.5 * dh
    s = block_scene(n, (3 * dh, 3 * dh, 3 * dh), (13 * dh, 13 * dh, 13 * dh), density=rho,
                    flip_alpha=1.0)
    X = n + 1
    s.L.orc_set_dt(s.h, dt)
    s.set_field(O.VELOCITY, np.zeros(3 * X ** 3))
    s.L.orc_solid_copy_velocity(s.h)
    vwall = np.array([0.2, 0.0, -0.1])  # the solid's velocity at the floor nodes
    sv = np.zeros((X, X, X, 3))
    sv[:, 0, :] = vwall
    s.set_field(O.S_VELOCITY, sv.reshape(-1), 0)
    s.set_field(O.S_VEL_WORKSPACE, sv.reshape(-1), 0)
    m = rho * dh ** 3 / 8
    # a fast approach (the tangential part dies: the scale clamps at 0) and a
    # grazing one (a slow approach beside a fast slide: the scale stays
    # positive, 0.1 - friction dv_n / |v_t|)
    for y, v in ((0.6 * dh, np.array([0.5, -2.0, 0.3])), (1.4 * dh, np.array([5.0, -0.01, 0.0]))):
        s.set_particles(np.array([[8.3 * dh, y, 7.2 * dh]]), v[None])
        s.L.orc_grid_to_particle(s.h, 0)
        pos, vel = s.particles(0)
        vrel = v - vwall
        vn = vrel[1]
        overlap = thick - y
        imp = -min(dt * ks * overlap, m * (0.1 * overlap / dt - vn))
        dvn = -imp / m
        vt = vrel - vn * np.array([0, 1.0, 0])
        sc = max(0.1 - fr * dvn / np.linalg.norm(vt), 0.0)
        want = (vn + dvn) * np.array([0, 1.0, 0]) + sc * vt + vwall
        np.testing.assert_allclose(vel[0], want, rtol=1e-13)
        np.testing.assert_allclose(pos[0], [8.3 * dh, y, 7.2 * dh] + want * dt, rtol=1e-15)
        if y > dh:
            assert 0 < sc < 0.1
    s.close()


def test_particle_to_grid_average_and_cutoff():
    """particles of one velocity V: every node whose weight sum (over the
    particles whose 4^3 stencils cover it) reaches the cutoff takes V and
    _mass 1; the others _mass 0 and keep their velocity"""
    n = 16
    dh = 1.0 / n
    s = block_scene(n, (3 * dh, 3 * dh, 3 * dh), (13 * dh, 13 * dh, 13 * dh))
    X = n + 1
    rng = np.random.default_rng(22)
    p = rng.uniform(6 * dh, 9 * dh, (400, 3))
    V = np.array([0.7, -0.2, 0.4])
    s.set_particles(p, np.tile(V, (400, 1)))
    old = rng.standard_normal((X ** 3, 3))
    s.set_field(O.S_VELOCITY, old.reshape(-1), 0)
    s.L.orc_particle_to_grid(s.h, 0)
    w = np.zeros((X, X, X))
    for q in p:
        base = np.floor(q / dh).astype(int) - 1
        for r in range(base[2], base[2] + 4):
            for t in range(base[1], base[1] + 4):
                for x in range(base[0], base[0] + 4):
                    d = q / dh - np.array([x, t, r])
                    w[r, t, x] += bspline(d[0]) * bspline(d[1]) * bspline(d[2])
    full = (w >= 3).reshape(-1)
    assert 0 < full.sum() < (w > 0).sum()  # the cutoff splits the covered nodes
    np.testing.assert_array_equal(s.field(O.S_MASS, 0) == 1, full)
    np.testing.assert_array_equal(s.field(O.S_MASS, 0)[~full], 0)
    vel = s.field(O.S_VELOCITY, 0).reshape(-1, 3)
    np.testing.assert_allclose(vel[full], np.broadcast_to(V, vel[full].shape), rtol=1e-14)
    np.testing.assert_array_equal(vel[~full], old[~full])
    s.close()
