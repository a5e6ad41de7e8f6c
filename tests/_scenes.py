"""Shared scene builders: the same cfg text drives libesf (the product,
through esf_scene_load_cfg) and the CPU oracle."""
import numpy as np

from eulersf_amd.cfg import unit_cube_ball_cfg


def ball_cfg(n, **kw):
    return unit_cube_ball_cfg(n, **kw)


def interior_mask(X, Y, Z):
    m = np.zeros((Z, Y, X), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.reshape(-1)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def solution_order(v_nodes, X, Y, Z):
    """AoS node field (3*nodes) -> the reference's solution order over
    interior nodes (3*(X-2)(Y-2)(Z-2))."""
    v = np.asarray(v_nodes).reshape(Z, Y, X, 3)[1:-1, 1:-1, 1:-1]
    return v.reshape(-1)


def teacher_force(g, o, n_solids=1, particles=False):
    """Copy the oracle scene's state into the GPU context: grid velocity,
    intensity and the pressure warm start; per solid its boxes, u, X,
    velocity, velocity workspace, mass, quarter fractions, valid flags and
    phi; then gatherSolidMasses.  The next step of both then starts from the
    same state (teacher forcing)."""
    from eulersf_amd import _abi as A
    from oracle import oracle as O
    g.upload(A.F_VELOCITY, o.field(O.VELOCITY))
    g.upload(A.F_INTENSITY, o.field(O.INTENSITY))
    g.upload(A.F_LAMBDA, o.field(O.LAMBDA))
    for b in range(n_solids):
        g.set_boxes(b, o.boxes(b))
        for gf, of in ((A.S_U, O.S_U), (A.S_X, O.S_X), (A.S_VELOCITY, O.S_VELOCITY),
                       (A.S_VEL_WORKSPACE, O.S_VEL_WORKSPACE), (A.S_MASS, O.S_MASS),
                       (A.S_QFRAC, O.S_QFRAC), (A.S_VALID, O.S_VALID), (A.S_PHI, O.S_PHI)):
            g.upload(gf, o.field(of, b), b)
        if particles:
            pos, vel = o.particles(b)
            g.setParticles(pos, vel, b)
    g.gatherSolidMasses()


def tight_box_phi_above_2(phi, boxes, X, Y, Z):
    """the set computeSolidMass(true) skips: phi > 2 at the min corners of
    the tight box's cells"""
    t = boxes[:6]
    p = np.asarray(phi).reshape(Z, Y, X)
    return p[t[2]:t[5] - 1, t[1]:t[4] - 1, t[0]:t[3] - 1] > 2


def jt_lambda(o, lam):
    """-J^T lambda over the interior DOFs in the reference's solution order
    (the pressure compared modulo the null space of S)"""
    return o.apply_negJT(lam)


def _solve_distance(p, q, r, v):
    """solveDistance (DEFOR_SOLID_3D.cpp:369-395) elementwise, in the same
    operation order as the device's, so results are bitwise comparable"""
    with np.errstate(invalid="ignore"):
        d0 = np.fmin(p, np.fmin(q, r))
        d2 = np.fmax(p, np.fmax(q, r))
        d1 = p + q + r - d0 - d2
        b0, c0 = d0, d0 * d0 - 1.0
        b1, c1 = d1 + b0, d1 * d1 + c0
        b2, c2 = d2 + b1, d2 * d2 + c1
        disc = b2 * b2 - 3.0 * c2
        nd = (b2 + np.sqrt(disc)) / 3.0
        go = (disc < 0.0) | (nd < d2)
        disc = b1 * b1 - 2.0 * c1
        nd = np.where(go & (disc > 0.0), (b1 + np.sqrt(disc)) / 2.0, nd)
        go = go & ((disc < 0.0) | (nd < d1))
        disc = b0 * b0 - 1.0 * c0
        nd = np.where(go & (disc > 0.0), (b0 + np.sqrt(disc)) / 1.0, nd)
    return np.where(nd < v, nd, v)


def jacobi_phi(mass, ext, X, Y, Z, relaxations):
    """the box's phi after `relaxations` whole-box Jacobi relaxations of the
    upwind update from phi = -0.5 on nodes with mass, `large` elsewhere;
    neighbours outside the extended box `ext` ([lo, hi) per axis, x first)
    read as `large`, nodes with mass keep their value.  Returns the box's
    values, (z, y, x) ordered."""
    large = float(X + Y + Z + 2)
    m = np.asarray(mass).reshape(Z, Y, X)[ext[2]:ext[5], ext[1]:ext[4], ext[0]:ext[3]]
    cur = np.where(m > 0, -0.5, large)
    free = m == 0
    for _ in range(relaxations):
        pad = np.pad(cur, 1, constant_values=large)
        p = np.fmin(pad[1:-1, 1:-1, :-2], pad[1:-1, 1:-1, 2:])
        q = np.fmin(pad[1:-1, :-2, 1:-1], pad[1:-1, 2:, 1:-1])
        r = np.fmin(pad[:-2, 1:-1, 1:-1], pad[2:, 1:-1, 1:-1])
        cur = np.where(free, _solve_distance(p, q, r, cur), cur)
    return cur
