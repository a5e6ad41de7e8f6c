"""Known answers that pin the oracle's restatements of the reference's
per-step functions which no golden vector covers: the deformation gradient
and its degenerate fallback, the element force and stiffness, the
vorticity and its confinement force, the RK2 velocity advection with its
back-off out of the solid, the displacement update's face extrapolation,
the subsampled solid mass, the solid nodes' body forces and damped mass,
and the CFL.  Each expected value comes from the mathematics of the case
(linear fields, rigid motions, uniform stretches, cell-aligned blocks),
chosen so that a plausible slip -- a wrong sign, index, scale or a dropped
term -- changes the answer."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

LAM, MU = 5769.2307692307695, 3846.153846153846


def params(n, **kw):
    dh = 1.0 / n
    p = dict(dt=1.0 / 240.0, vorticity_eps=0.0, gravity=0.0, buoyancy=0.0, rayleigh_alpha=0.0,
             rayleigh_beta=0.0, fluid_density=1000.0, implicit_eps=1e-12, divfree_eps=1e-12,
             cr_max_iterations=20000, mass_sample_dh=dh / 4, tight_border=3,
             extrapolation_cells=20, particles_per_cell=8, flip_alpha=0.95, flip_cutoff=3,
             wall_thickness=1.5 * dh, spring_stiffness=1e4, friction=0.5, flip=0)
    p.update(kw)
    return p


def block_rest_mass(lo, hi, dh, density):
    """a rest-space density `density` on the box [lo, hi] (3-vectors), zero
    outside, on a rest grid of spacing dh / 16 with two empty layers around"""
    rdh = dh / 16
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    org = lo - 3 * rdh
    dims = tuple(int(round((hi[k] - lo[k]) / rdh)) + 7 for k in range(3))
    z, y, x = np.meshgrid(*(np.arange(dims[k]) for k in (2, 1, 0)), indexing="ij")
    pos = np.stack([x, y, z], -1) * rdh + org
    inside = np.all((pos >= lo - 1e-12 * dh) & (pos <= hi + 1e-12 * dh), axis=-1)
    return dims, rdh, org, (inside * density).astype(np.float64).reshape(-1)


def block_scene(n, lo, hi, density=1000.0, velocity=(0, 0, 0), lam=LAM, mu=MU, **kw):
    """an n^3-cell unit cube holding one solid whose rest mass is the box
    [lo, hi] (rest space = world space at rest), after COUPLED_INTEGRATOR_3D::init"""
    grid = {"xres": n + 1, "yres": n + 1, "zres": n + 1, "dh": 1.0 / n, "origin": (0, 0, 0)}
    s = O.OracleScene(grid, params(n, **kw))
    dims, rdh, org, rest = block_rest_mass(lo, hi, 1.0 / n, density)
    s.add_solid(lam, mu, density, dims, rdh, org, rest, np.eye(3).reshape(-1), np.zeros(3),
                np.ones(3), np.asarray(velocity, float))
    s.L.orc_init(s.h, 0.0)
    return s


def node_pos(n):
    X = n + 1
    z, y, x = np.meshgrid(np.arange(X), np.arange(X), np.arange(X), indexing="ij")
    return np.stack([x, y, z], -1).reshape(-1, 3) / n


def set_deformation(s, n, C, solid=0):
    """u = C x at every node (rest X = x - u), the whole grid as tight box"""
    x = node_pos(n)
    u = x @ np.asarray(C, float).T
    s.set_field(O.S_U, u.reshape(-1), solid)
    s.set_field(O.S_X, (x - u).reshape(-1), solid)
    return u


def full_boxes(s, n, solid=0):
    b = (ctypes.c_int * 12)(*([0, 0, 0, n + 1, n + 1, n + 1] * 2))
    s.L.orc_set_boxes(s.h, solid, b)


# ------------------------------------------------------------------ computeF

def test_computeF_of_linear_displacement_and_its_fallback():
    """u = C x gives F^-1 = I - C exactly (the Q1 gradient of a linear field
    at the cell centre), so F = (I - C)^-1 and det F = 1 / det(I - C); when
    |det F^-1| <= 1e-4 the reference substitutes F = I and det F = 1
    (DEFOR_SOLID_3D.cpp:671-682)"""
    n = 6
    s = block_scene(n, (0.25, 0.25, 0.25), (0.75, 0.75, 0.75))
    rng = np.random.default_rng(1)
    F = np.zeros(9)
    idx = 2 + 3 * 7 + 3 * 49
    for _ in range(5):
        C = 0.3 * rng.standard_normal((3, 3))
        set_deformation(s, n, C)
        det = s.L.orc_compute_F(s.h, 0, idx, O.dp(F))
        want = np.linalg.inv(np.eye(3) - C)
        np.testing.assert_allclose(F.reshape(3, 3), want, rtol=0, atol=1e-12 * np.abs(want).max())
        assert abs(det - 1.0 / np.linalg.det(np.eye(3) - C)) < 1e-12 * abs(det)
    # near-degenerate F^-1 (det 1e-5): F = I, det F = 1
    set_deformation(s, n, np.eye(3) - np.diag([1e-5, 1.0, 1.0]))
    det = s.L.orc_compute_F(s.h, 0, idx, O.dp(F))
    np.testing.assert_array_equal(F.reshape(3, 3), np.eye(3))
    assert det == 1.0
    # just above the threshold (det 2e-4): the true inverse
    set_deformation(s, n, np.eye(3) - np.diag([2e-4, 1.0, 1.0]))
    det = s.L.orc_compute_F(s.h, 0, idx, O.dp(F))
    assert abs(F[0] - 5000.0) < 1e-6 and abs(det - 5000.0) < 1e-6
    s.close()


# ------------------------------------------------------------- solid mass

def test_solid_mass_of_a_cell_aligned_block():
    """rest mass on the box [2, 5.5] x [2, 5] x [2, 5] cells at rest: cells
    wholly inside carry all 8 quarter fractions = 1 and _valid = 1; the
    half-covered cells (sum 0.5, not above the 0.5 threshold) _valid = 2
    with fractions 0; node mass = (adjacent valid cells) * rho dh^3 / 8.
    Stretched by S = diag(1.5, 1, 1) (u = (I - S^-1) x) the block covers 1.5x
    the cells in x and every node's mass carries |det F^-1| = 1/1.5"""
    n = 8
    dh = 1.0 / n
    rho = 1000.0
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5.5 * dh, 5 * dh, 5 * dh), density=rho)
    X, C = n + 1, n
    valid = s.field(O.S_VALID, 0).reshape(C, C, C)
    want = np.zeros((C, C, C), dtype=np.int8)
    want[2:5, 2:5, 2:5] = 1
    want[2:5, 2:5, 5] = 2
    np.testing.assert_array_equal(valid, want)
    q = s.field(O.S_QFRAC, 0).reshape(X, X, X, 8)
    qwant = np.zeros((X, X, X, 8))
    qwant[2:5, 2:5, 2:5] = 1.0
    np.testing.assert_array_equal(q, qwant)
    cnt = np.zeros((X, X, X))
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                cnt[oz:oz + C, oy:oy + C, ox:ox + C] += want == 1
    qcm = rho * dh ** 3 * 0.125
    np.testing.assert_allclose(s.field(O.S_MASS, 0).reshape(X, X, X), cnt * qcm, rtol=1e-15)

    # the stretch: rest x in [2, 4] dh maps to world [3, 6] dh
    s2 = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (4 * dh, 5 * dh, 5 * dh), density=rho)
    S = np.diag([1.5, 1.0, 1.0])
    set_deformation(s2, n, np.eye(3) - np.linalg.inv(S))
    full_boxes(s2, n)
    s2.L.orc_compute_solid_mass(s2.h, 0)
    v2 = s2.field(O.S_VALID, 0).reshape(C, C, C)
    w2 = np.zeros((C, C, C), dtype=np.int8)
    w2[2:5, 2:5, 3:6] = 1
    np.testing.assert_array_equal(v2, w2)
    det_inv = s2.field(O.S_DETFINV, 0).reshape(X, X, X)
    np.testing.assert_allclose(det_inv[:C, :C, :C], 1 / 1.5, rtol=1e-14)
    cnt2 = np.zeros((X, X, X))
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                cnt2[oz:oz + C, oy:oy + C, ox:ox + C] += w2 == 1
    np.testing.assert_allclose(s2.field(O.S_MASS, 0).reshape(X, X, X), cnt2 * qcm / 1.5,
                               rtol=1e-14)
    s.close()
    s2.close()


def test_solid_mass_skips_cells_far_from_the_solid():
    """computeSolidMass(true) takes no samples in a tight-box cell whose min
    corner has phi > 2 (DEFOR_SOLID_3D.cpp:620): with phi forced above 2 on
    one layer of the block, that layer's cells lose their mass and flags"""
    n = 8
    dh = 1.0 / n
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    X, C = n + 1, n
    before = s.field(O.S_VALID, 0).reshape(C, C, C).copy()
    phi = s.field(O.S_PHI, 0).reshape(X, X, X)
    assert np.all(phi[2:6, 2:6, 2:6] == -0.5)  # massive nodes
    phi[4, :, :] = 3.0  # the cells with min corner z = 4
    s.set_field(O.S_PHI, phi.reshape(-1), 0)
    s.L.orc_compute_solid_mass_sdf(s.h, 0)
    after = s.field(O.S_VALID, 0).reshape(C, C, C)
    want = before.copy()
    want[4] = 0
    np.testing.assert_array_equal(after, want)
    s.close()


# --------------------------------------------------- element force, stiffness

def solid_force_by_node(s, n):
    """the solid DOF force vector as a node field (X, X, X, 3), zero off solid nodes"""
    X = n + 1
    sidx = s.field(O.SOLID_INDEX)
    f = s.field(O.SOLID_FORCE).reshape(-1, 3)
    out = np.zeros((X ** 3, 3))
    m = sidx >= 0
    out[m] = f[sidx[m]]
    return out.reshape(X, X, X, 3)


def test_element_force_rigid_motions_and_uniform_stretch():
    """gravity off, so the solid force is the elastic one: zero under a rigid
    translation and a rigid rotation (corotational P(R) = 0), and under a
    uniform rotated stretch F = R S the nodal forces -sigma G^T dh^3 summed
    over the valid cells at each node: f(n) = -(dh^2 / 4) sigma sum_c s_c(n),
    with sigma = det(F)^-1 P F^T, P = R (2 mu (S - I) + lambda tr(S - I) I)
    and s_c(n) the sign vector of n's corner in cell c (interior nodes
    cancel)"""
    n = 8
    dh = 1.0 / n
    X, C = n + 1, n
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    s.L.orc_set_dt(s.h, 1.0 / 240)
    valid = s.field(O.S_VALID, 0).reshape(C, C, C) == 1
    assert valid.sum() == 27
    # translation
    x = node_pos(n)
    s.set_field(O.S_U, np.tile([0.01, -0.02, 0.03], X ** 3), 0)
    s.L.orc_forces(s.h)
    # (the gradient's running sums of +-u g leave rounding-level entries)
    assert np.abs(s.field(O.SOLID_FORCE)).max() < 1e-12 * (LAM + MU) * dh ** 2
    # rotation about the z axis by 0.3 rad: rest X = R^T x
    c, sn = np.cos(0.3), np.sin(0.3)
    R = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1]])
    u = x - x @ R  # x - R^T x per row
    s.set_field(O.S_U, u.reshape(-1), 0)
    s.L.orc_forces(s.h)
    assert np.abs(s.field(O.SOLID_FORCE)).max() < 1e-9 * (LAM + MU) * dh ** 2
    # uniform stretch, rotated: F = R S (not symmetric), P = R (2 mu (S - I)
    # + lambda tr(S - I) I), sigma = det(F)^-1 P F^T
    S = np.diag([1.1, 0.95, 1.03])
    F = R @ S
    set_deformation(s, n, np.eye(3) - np.linalg.inv(F))
    s.L.orc_forces(s.h)
    E = S - np.eye(3)
    P = R @ (2 * MU * E + LAM * np.trace(E) * np.eye(3))
    sigma = P @ F.T / np.linalg.det(F)
    want = np.zeros((X, X, X, 3))
    for cz in range(C):
        for cy in range(C):
            for cx in range(C):
                if not valid[cz, cy, cx]:
                    continue
                for i in range(8):
                    ox, oy, oz = i & 1, (i >> 1) & 1, i >> 2
                    sg = np.array([2 * ox - 1, 2 * oy - 1, 2 * oz - 1], float)
                    want[cz + oz, cy + oy, cx + ox] -= dh ** 2 / 4 * sigma @ sg
    got = solid_force_by_node(s, n)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
    assert np.abs(want).max() > 1.0  # the case is not trivial
    s.close()


def test_injected_deformation_per_cell():
    """per-cell deformation gradients injected in place of the displacement's:
    a different F_c = R_c S_c in each valid cell gives the nodal forces
    -(dh^2 / 4) sum_c sigma_c s_c(n) with sigma_c = det(F_c)^-1 P(F_c) F_c^T,
    injecting computeF's own values reproduces the displacement's forces
    to rounding, and the solid CR's answer on the assembled system leaves
    |A x - b| <= eps |b|"""
    from eulersf_amd.cfg import random_deformations
    n = 8
    dh = 1.0 / n
    X, C = n + 1, n
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    s.L.orc_set_dt(s.h, 1.0 / 240)
    valid = s.field(O.S_VALID, 0).reshape(C, C, C) == 1
    nc = s.L.orc_solid_num_cells(s.h, 0)
    assert nc == valid.sum() == 27
    rng = np.random.default_rng(5)
    Fs = np.ascontiguousarray(random_deformations(rng, nc, 0.08))
    s.L.orc_solid_set_deformation(s.h, 0, nc, O.dp(Fs))
    s.L.orc_forces(s.h)
    want = np.zeros((X, X, X, 3))
    e = 0
    for cz in range(C):
        for cy in range(C):
            for cx in range(C):
                if not valid[cz, cy, cx]:
                    continue
                F = Fs[e]
                e += 1
                U, sv, Vt = np.linalg.svd(F)
                R, S = U @ Vt, Vt.T @ np.diag(sv) @ Vt
                E = S - np.eye(3)
                sigma = R @ (2 * MU * E + LAM * np.trace(E) * np.eye(3)) @ F.T / np.linalg.det(F)
                for i in range(8):
                    ox, oy, oz = i & 1, (i >> 1) & 1, i >> 2
                    sg = np.array([2 * ox - 1, 2 * oy - 1, 2 * oz - 1], float)
                    want[cz + oz, cy + oy, cx + ox] -= dh ** 2 / 4 * sigma @ sg
    got = solid_force_by_node(s, n)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
    assert np.abs(want).max() > 1.0
    # the solid CR on the assembled system
    s.L.orc_assemble(s.h)
    ndof = 3 * s.L.orc_num_solid_dofs(s.h)
    b = rng.standard_normal(ndof)
    x = np.zeros(ndof)
    it = ctypes.c_int()
    assert s.L.orc_solid_cr(s.h, O.dp(b), O.dp(x), 1e-12, 20000, ctypes.byref(it)) == 0
    Ax = np.zeros(ndof)
    s.L.orc_solid_apply(s.h, O.dp(x), O.dp(Ax))
    assert 0 < it.value and np.linalg.norm(Ax - b) <= 1e-12 * np.linalg.norm(b)
    # computeF's values injected: the displacement's forces, bitwise
    set_deformation(s, n, 0.05 * rng.standard_normal((3, 3)))
    s.L.orc_solid_set_deformation(s.h, 0, 0, O.dp(Fs))
    s.L.orc_forces(s.h)
    base = s.field(O.SOLID_FORCE).copy()
    Fc = np.zeros((nc, 9))
    e = 0
    for cz in range(C):
        for cy in range(C):
            for cx in range(C):
                if valid[cz, cy, cx]:
                    s.L.orc_compute_F(s.h, 0, cx + cy * X + cz * X * X, O.dp(Fc[e]))
                    e += 1
    s.L.orc_solid_set_deformation(s.h, 0, nc, O.dp(Fc))
    s.L.orc_forces(s.h)
    # (det F recomputed from the injected F, not as 1 / det F^-1: rounding)
    np.testing.assert_allclose(s.field(O.SOLID_FORCE), base, rtol=0, atol=1e-13 * np.abs(base).max())
    assert np.abs(base).max() > 1.0
    s.close()


def test_stiffness_symmetric_and_free_of_rigid_modes():
    """A = dt^2 K + diag(M~) (beta = 0): symmetric (the corotational tangent
    with the doubled lambda trace is still symmetric), A t = M~ t for a
    translation t at any deformation, and at rest (F = I) K also annihilates
    infinitesimal rotations"""
    n = 8
    dh = 1.0 / n
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    dt = 1.0 / 240
    s.L.orc_set_dt(s.h, dt)
    s.L.orc_forces(s.h)
    s.L.orc_assemble(s.h)
    ns = s.num_solid_nodes()
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(3 * ns), rng.standard_normal(3 * ns)
    Aa, Ab = s.solid_apply(a), s.solid_apply(b)
    assert abs(a @ Ab - b @ Aa) < 1e-12 * np.abs(a) @ np.abs(Ab)
    damped = s.field(O.DAMPED_MASS)
    t = np.tile([0.3, -1.1, 0.7], ns)
    np.testing.assert_allclose(s.solid_apply(t), damped * t, rtol=1e-12, atol=0)
    # infinitesimal rotation w x (p - p0) about the block centre
    sidx = s.field(O.SOLID_INDEX)
    pos = node_pos(n)[sidx >= 0][np.argsort(sidx[sidx >= 0])]
    w = np.array([0.2, -0.5, 0.9])
    rot = np.cross(w, pos - 0.4375).reshape(-1)
    Krot = (s.solid_apply(rot) - damped * rot) / dt ** 2
    Kscale = np.abs((s.solid_apply(a) - damped * a) / dt ** 2).max()
    assert np.abs(Krot).max() < 1e-9 * Kscale
    # a deformed block: symmetry and translations still hold
    set_deformation(s, n, 0.05 * rng.standard_normal((3, 3)))
    s.L.orc_forces(s.h)
    s.L.orc_assemble(s.h)
    Aa, Ab = s.solid_apply(a), s.solid_apply(b)
    assert abs(a @ Ab - b @ Aa) < 1e-12 * np.abs(a) @ np.abs(Ab)
    np.testing.assert_allclose(s.solid_apply(t), s.field(O.DAMPED_MASS) * t, rtol=1e-12, atol=0)
    s.close()


def test_stiffness_action_on_a_linear_field():
    """under a uniform deformation F = R S, K applied to a linear nodal field
    du = B x: every cell sees dF = sum_c du_c (G_c F)^T = B F (the Q1 gradient
    of a linear field is exact), so K_e du_e = det(F)^-1 dh^3 dP(B F) (G F)^T
    and at a node (K du)(n) = det(F)^-1 (dh^2 / 4) dP(B F) F^T sum_c s_c(n),
    with dP the corotational differential (golden-pinned on its own)"""
    n = 8
    dh = 1.0 / n
    X = n + 1
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    dt = 1.0 / 240
    s.L.orc_set_dt(s.h, dt)
    c, sn = np.cos(0.4), np.sin(0.4)
    R = np.array([[1, 0, 0], [0, c, -sn], [0, sn, c]])
    F = R @ np.diag([1.08, 0.93, 1.02])
    set_deformation(s, n, np.eye(3) - np.linalg.inv(F))
    s.L.orc_forces(s.h)
    s.L.orc_assemble(s.h)
    rng = np.random.default_rng(13)
    B = rng.standard_normal((3, 3))
    sidx = s.field(O.SOLID_INDEX)
    order = np.flatnonzero(sidx >= 0)[np.argsort(sidx[sidx >= 0])]
    pos = node_pos(n)
    du = (pos[order] @ B.T).reshape(-1)
    damped = s.field(O.DAMPED_MASS)
    Kdu = ((s.solid_apply(du) - damped * du) / dt ** 2).reshape(-1, 3)
    *_, dP = O.corotation(LAM, MU, F, B @ F)
    M = dP @ F.T / np.linalg.det(F) * dh ** 2 / 4
    valid = s.field(O.S_VALID, 0).reshape(n, n, n) == 1
    want = np.zeros((X, X, X, 3))
    for cz, cy, cx in zip(*np.nonzero(valid)):
        for i in range(8):
            ox, oy, oz = i & 1, (i >> 1) & 1, i >> 2
            sg = np.array([2 * ox - 1, 2 * oy - 1, 2 * oz - 1], float)
            want[cz + oz, cy + oy, cx + ox] += M @ sg
    want = want.reshape(-1, 3)[order]
    np.testing.assert_allclose(Kdu, want, rtol=0, atol=1e-8 * np.abs(want).max())
    assert np.abs(want).max() > 10.0
    s.close()


# -------------------------------------------------------- body forces, CFL

def test_body_forces_damped_mass_and_buoyancy():
    """gravity g, buoyancy b, Rayleigh alpha: at solid nodes M~ = m_s (1 + dt
    alpha) + m_f and F = -(m_s + m_f) g y + b m_f I y (u = 0: no elastic
    force); at fluid nodes f = -m g y + b m_f I y with m = m_f; 1/m
    everywhere, m_f = (1 - phi_s) rho_f dh^3"""
    n = 8
    dh = 1.0 / n
    g, b, alpha, dt, rho_f = 9.8, 0.7, 3.0, 1.0 / 240, 1000.0
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh), gravity=g,
                    buoyancy=b, rayleigh_alpha=alpha)
    X = n + 1
    rng = np.random.default_rng(4)
    inten = rng.uniform(0, 1, X ** 3)
    s.set_field(O.INTENSITY, inten)
    s.L.orc_set_dt(s.h, dt)
    s.L.orc_forces(s.h)
    nf = s.field(O.NODE_FRACTION)
    ms = s.field(O.SOLID_MASS)
    mf = np.maximum(1 - nf, 0) * rho_f * dh ** 3
    sidx = s.field(O.SOLID_INDEX)
    solid = sidx >= 0
    order = np.argsort(sidx[solid])
    nodes = np.flatnonzero(solid)[order]
    damped = s.field(O.DAMPED_MASS).reshape(-1, 3)
    np.testing.assert_allclose(damped, np.repeat((ms[nodes] * (1 + dt * alpha) + mf[nodes])[:, None],
                                                 3, 1), rtol=1e-15)
    fs = s.field(O.SOLID_FORCE).reshape(-1, 3)
    np.testing.assert_allclose(fs[:, 1], -(ms[nodes] + mf[nodes]) * g + b * mf[nodes] * inten[nodes],
                               rtol=1e-13)
    assert np.all(fs[:, [0, 2]] == 0)
    inner = np.zeros((X, X, X), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    fu = s.field(O.UNIT_FORCE).reshape(-1, 3)  # interior nodes, z, y, x order
    fl = ~solid[inner]
    m = (mf + ms)[inner]
    np.testing.assert_allclose(fu[fl, 1], -m[fl] * g + b * mf[inner][fl] * inten[inner][fl],
                               rtol=1e-13)
    np.testing.assert_allclose(s.field(O.MINV).reshape(-1, 3)[:, 0], 1.0 / m, rtol=1e-15)
    s.close()


def test_cfl_reads_solid_nodes_only():
    """dt = min(max dt, alpha dh / max_d |v_d|) over nodes with solid mass;
    fluid velocities, however large, do not enter (COUPLED_INTEGRATOR_3D.cpp:88-102)"""
    n = 8
    dh = 1.0 / n
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    X = n + 1
    v = np.zeros((X ** 3, 3))
    solid = s.field(O.SOLID_MASS) > 0
    v[~solid] = 50.0
    rng = np.random.default_rng(5)
    v[solid] = rng.uniform(-2, 2, (solid.sum(), 3))
    s.set_field(O.VELOCITY, v.reshape(-1))
    vmax = np.abs(v[solid]).max()
    assert s.L.orc_cfl(s.h, 1.8) == min(1.0 / 240, 1.8 / (vmax / dh))
    assert s.L.orc_cfl(s.h, 0.6) == min(1.0 / 240, 0.6 / (vmax / dh))
    s.close()


# ---------------------------------------------------------- vorticity

def fluid_scene(n, **kw):
    grid = {"xres": n + 1, "yres": n + 1, "zres": n + 1, "dh": 1.0 / n, "origin": (0, 0, 0)}
    s = O.OracleScene(grid, params(n, **kw))
    s.L.orc_init(s.h, 0.0)
    return s


def test_vorticity_of_a_linear_field_and_its_one_sided_stencils():
    """v = (a y + e z, b x, d x): the reference forms each vorticity
    component from two differences and scales the pair by one axis's
    spacing (omega_y by the x spacing, omega_z by the y spacing,
    COUPLED_INTEGRATOR_3D.cpp:486-494), and takes a one-sided difference
    (width dh, spacing 1/dh) instead of a central one (2 dh, 1/(2 dh)) beside
    a node whose solid fraction is 1 and on the first interior layer.  For
    this field omega = (0, (e w_z - d w_x) / w_x, (b w_x - a w_y) / w_y) with w
    the stencil widths: the curl (0, e - d, b - a) where all are central,
    (0, 2e - d, b/2 - a) beside a full node in x, (0, e - d, 2b - a) beside
    one in y, (0, e/2 - d, b - a) beside one in z; |omega| in the workspace"""
    n = 64  # vorticity eps is scaled by max(64 / maxRes, 1): 1 here
    a, b, d, e = 0.7, -1.3, 0.4, 0.9
    s = fluid_scene(n, vorticity_eps=1.5)
    X = n + 1
    pos = node_pos(n)
    v = np.zeros((X ** 3, 3))
    v[:, 0] = a * pos[:, 1] + e * pos[:, 2]
    v[:, 1] = b * pos[:, 0]
    v[:, 2] = d * pos[:, 0]
    s.set_field(O.VELOCITY, v.reshape(-1))
    # a full-solid node at (30, 30, 30): the gathered solid fraction there is
    # 1 (a fluid scene has no solid to gather it from, so it is set directly)
    nf = np.zeros(X ** 3)
    nf[30 + 30 * X + 30 * X * X] = 1.0
    s.set_field(O.NODE_FRACTION, nf)
    s.L.orc_set_dt(s.h, 1.0 / 240)
    s.L.orc_forces(s.h)
    w = s.field(O.VORTICITY).reshape(X, X, X, 3)
    at = lambda x, y, z: w[z, y, x]
    np.testing.assert_allclose(at(10, 40, 20), [0, e - d, b - a], atol=1e-12)
    np.testing.assert_allclose(at(31, 30, 30), [0, 2 * e - d, b / 2 - a], atol=1e-12)  # left full
    np.testing.assert_allclose(at(29, 30, 30), [0, 2 * e - d, b / 2 - a], atol=1e-12)  # right full
    np.testing.assert_allclose(at(30, 31, 30), [0, e - d, 2 * b - a], atol=1e-12)  # below full
    np.testing.assert_allclose(at(30, 30, 29), [0, e / 2 - d, b - a], atol=1e-12)  # out full
    np.testing.assert_allclose(at(1, 40, 20), [0, 2 * e - d, b / 2 - a], atol=1e-12)  # wall x = 0
    np.testing.assert_allclose(at(10, 40, 63), [0, e / 2 - d, b - a], atol=1e-12)  # wall z = 64
    assert np.all(at(30, 30, 30) == 0)  # the full node itself is skipped
    ws = s.field(O.WORKSPACE).reshape(X, X, X)
    assert abs(ws[20, 40, 10] - np.hypot(e - d, b - a)) < 1e-12
    s.close()


def test_vorticity_confinement_force_closed_form():
    """v = (0, c x^2, 0): omega = (0, 0, 2 c x) exactly (central differences
    of a quadratic), |omega| = 2 c x has the unit gradient N = (1, 0, 0), so
    the confinement force eps dh m_f (N x omega) = (0, -eps dh m_f 2 c x, 0)
    at fluid nodes whose six neighbours are all fluid"""
    n = 64
    c, eps = 0.8, 1.5
    s = fluid_scene(n, vorticity_eps=eps)
    X = n + 1
    dh = 1.0 / n
    pos = node_pos(n)
    v = np.zeros((X ** 3, 3))
    v[:, 1] = c * pos[:, 0] ** 2
    s.set_field(O.VELOCITY, v.reshape(-1))
    s.L.orc_set_dt(s.h, 1.0 / 240)
    s.L.orc_forces(s.h)
    fu = s.field(O.UNIT_FORCE).reshape(n - 1, n - 1, n - 1, 3)  # interior z, y, x
    mf = 1000.0 * dh ** 3
    xs = (np.arange(1, n) * dh)[None, None, :]
    want_y = -eps * dh * mf * 2 * c * xs
    inner = fu[2:-2, 2:-2, 2:-2]
    np.testing.assert_allclose(inner[..., 1], np.broadcast_to(want_y[:, :, 2:-2], inner[..., 1].shape),
                               rtol=1e-9)
    assert np.abs(inner[..., [0, 2]]).max() < 1e-12 * np.abs(want_y).max()
    s.close()


# ---------------------------------------------------------- advection

def test_velocity_advection_is_rk2_midpoint():
    """for a linear velocity v(x) = A x + b (trilinear interpolation is exact
    on it), the RK2 departure point y = x - dt v(x - dt/2 v(x)) gives the new
    velocity A y + b; first order, A (x - dt v(x)) + b, differs by dt^2/2 A^2 v"""
    n = 16
    s = fluid_scene(n)
    X = n + 1
    dt = 1.0 / 30
    rng = np.random.default_rng(7)
    A = 2.0 * rng.standard_normal((3, 3))
    b = np.array([0.3, -0.2, 0.1])
    pos = node_pos(n)
    v = pos @ A.T + b
    s.set_field(O.VELOCITY, v.reshape(-1))
    s.L.orc_set_dt(s.h, dt)
    s.L.orc_advect_velocity(s.h)
    got = s.field(O.VELOCITY).reshape(X, X, X, 3)
    mid = (pos - 0.5 * dt * v) @ A.T + b
    y = pos - dt * mid
    want = (y @ A.T + b).reshape(X, X, X, 3)
    first = ((pos - dt * v) @ A.T + b).reshape(X, X, X, 3)
    sl = (slice(4, -4),) * 3
    np.testing.assert_allclose(got[sl], want[sl], rtol=0, atol=1e-12)
    assert np.abs(first[sl] - want[sl]).max() > 1e-4  # the two orders are distinguishable
    s.close()


def test_velocity_advection_backs_off_the_solid():
    """a departure point where the interpolated solid mass is positive is
    moved back along the midpoint velocity by dt/2, dt/4, dt/8 while it stays
    there (k > dt/16), COUPLED_INTEGRATOR_3D.cpp:808-820.  Uniform v = (V, 0, 0)
    toward a slab of solid nodes at x = 4 dh: the nodes whose departure
    point lands within a cell of the slab take v at the backed-off point; with
    v(x) = (V + a (y-y0), 0, 0) the sampled velocity shows no x dependence,
    so the uniform V with a linear z-component marks the point:
    v = (V, 0, c x) gives the new v_z = c y_x"""
    n = 16
    dh = 1.0 / n
    s = fluid_scene(n)
    X = n + 1
    V, c = 0.5, 2.0
    dt = 1.5 * dh / V  # departure 1.5 cells upstream
    pos = node_pos(n)
    v = np.zeros((X ** 3, 3))
    v[:, 0] = V
    v[:, 2] = c * pos[:, 0]
    s.set_field(O.VELOCITY, v.reshape(-1))
    sm = np.zeros((X, X, X))
    sm[:, :, 4] = 1.0  # solid mass on the plane x = 4 (interior part used)
    s.set_field(O.SOLID_MASS, sm.reshape(-1))
    s.L.orc_set_dt(s.h, dt)
    s.L.orc_advect_velocity(s.h)
    got = s.field(O.VELOCITY).reshape(X, X, X, 3)

    def departure(x0):
        # midpoint velocity: v_x = V everywhere, so y_x = x0 - dt V, then the back-off
        vm = V
        y = x0 - dt * vm
        k = dt / 2
        while abs(y - 4 * dh) < dh and k > dt / 16:
            y += vm * k
            k /= 2
        return y

    for xi in (5, 6, 7, 9):
        x0 = xi * dh
        want_vz = c * departure(x0)
        assert abs(got[8, 8, xi, 2] - want_vz) < 1e-12, (xi, got[8, 8, xi, 2], want_vz)
    # x = 5: departure at 3.5 dh (inside), backed off to 4.25 dh (inside),
    # then 4.625 dh (inside), then k = dt/16 stops: 4.8125 dh
    assert abs(departure(5 * dh) - 4.8125 * dh) < 1e-15
    s.close()


def test_displacement_face_extrapolation_on_the_extended_box():
    """with the solid at rest (v = 0) the displacement is unchanged inside the
    extended box, and every face of the box is rebuilt as 2 u(face + 1) -
    u(face + 2) (DEFOR_SOLID_3D.cpp:1204-1221): garbage written on the faces
    comes back as the linear field's exact values"""
    n = 12
    dh = 1.0 / n
    s = block_scene(n, (4 * dh, 4 * dh, 4 * dh), (8 * dh, 8 * dh, 8 * dh))
    X = n + 1
    rng = np.random.default_rng(8)
    C = 0.05 * rng.standard_normal((3, 3))
    d = np.array([0.01, 0.02, -0.01])
    pos = node_pos(n)
    u = pos @ C.T + d
    box = [2, 3, 1, 10, 11, 12]  # extended box [lo, hi) in x, y, z
    uf = u.reshape(X, X, X, 3).copy()
    lo, hi = box[:3], box[3:]
    for ax in range(3):
        for f in (lo[ax], hi[ax] - 1):
            sl = [slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0])]
            sl[2 - ax] = f
            uf[tuple(sl)] = rng.standard_normal(uf[tuple(sl)].shape)
    s.set_field(O.S_U, uf.reshape(-1), 0)
    s.set_field(O.S_VELOCITY, np.zeros(3 * X ** 3), 0)
    s.set_field(O.VELOCITY, np.zeros(3 * X ** 3))
    b = (ctypes.c_int * 12)(*([3, 3, 3, 9, 9, 9] + box))
    s.L.orc_set_boxes(s.h, 0, b)
    s.L.orc_set_dt(s.h, 1.0 / 240)
    s.L.orc_solid_copy_velocity(s.h)
    s.L.orc_solid_finalize(s.h)
    got = s.field(O.S_U, 0).reshape(X, X, X, 3)
    want = u.reshape(X, X, X, 3)
    inner = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
    np.testing.assert_allclose(got[inner], want[inner], rtol=0, atol=1e-14)
    Xm = s.field(O.S_X, 0).reshape(X, X, X, 3)
    np.testing.assert_allclose(Xm, pos.reshape(X, X, X, 3) - got, atol=1e-15)
    s.close()


# ------------------------------------------------ momentum, RHS, projection

def test_solid_momentum_sums_solid_and_fluid_parts():
    """M v at solid DOFs: m_s v_s from the solid's own velocity field
    (computeMV, DEFOR_SOLID_3D.cpp:873-889) plus m_f v from the grid
    velocity (COUPLED_INTEGRATOR_3D.cpp:592-593)"""
    n = 8
    dh = 1.0 / n
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    X = n + 1
    rng = np.random.default_rng(14)
    vs = rng.standard_normal((X ** 3, 3))
    vg = rng.standard_normal((X ** 3, 3))
    s.set_field(O.S_VELOCITY, vs.reshape(-1), 0)
    s.set_field(O.VELOCITY, vg.reshape(-1))
    s.L.orc_set_dt(s.h, 1.0 / 240)
    s.L.orc_forces(s.h)
    sidx = s.field(O.SOLID_INDEX)
    nodes = np.flatnonzero(sidx >= 0)[np.argsort(sidx[sidx >= 0])]
    ms = s.field(O.S_MASS, 0)[nodes]
    mf = np.maximum(1 - s.field(O.NODE_FRACTION)[nodes], 0) * 1000.0 * dh ** 3
    want = ms[:, None] * vs[nodes] + mf[:, None] * vg[nodes]
    np.testing.assert_allclose(s.field(O.SOLID_MV).reshape(-1, 3), want, rtol=1e-14, atol=1e-18)
    s.close()


def test_divergence_rhs_completes_the_cell_divergence():
    """c carries the boundary nodes' share of the Q1 cell divergence, so for a
    linear velocity v = A x + b on every node, J v_interior - c = 4 dh tr(A)
    in every cell, the boundary cells included (COUPLED_INTEGRATOR_3D.cpp:
    286-325 with :230-266)"""
    n = 7
    dh = 1.0 / n
    s = fluid_scene(n)
    X = n + 1
    rng = np.random.default_rng(15)
    A = rng.standard_normal((3, 3))
    b = rng.standard_normal(3)
    v = node_pos(n) @ A.T + b
    s.set_field(O.VELOCITY, v.reshape(-1))
    s.L.orc_set_dt(s.h, 1.0 / 240)
    s.L.orc_forces(s.h)
    s.L.orc_assemble(s.h)
    inner = v.reshape(X, X, X, 3)[1:-1, 1:-1, 1:-1].reshape(-1)
    got = s.apply_J(inner) - s.field(O.DIV_RHS)
    np.testing.assert_allclose(got, 4 * dh * np.trace(A), rtol=1e-12, atol=1e-13)
    s.close()


def test_fluid_projection_leaves_the_velocity_divergence_free():
    """without solids the Schur solve is a pressure projection: v1 = v0 -
    M^-1 J^T lambda with J M^-1 J^T lambda = J v0 - c, so J v1 = c up to the
    solve's tolerance (the discrete divergence, boundary share included,
    vanishes); the boundary nodes carry velocity, so c is not zero"""
    n = 10
    s = fluid_scene(n, gravity=9.8, divfree_eps=1e-12)
    X = n + 1
    rng = np.random.default_rng(16)
    # a trace-free linear field on every node (so the boundary nodes give c
    # != 0 and the system stays consistent) plus noise at interior nodes
    A = rng.standard_normal((3, 3))
    A -= np.trace(A) / 3 * np.eye(3)
    v = (node_pos(n) @ A.T).reshape(X, X, X, 3)
    v[1:-1, 1:-1, 1:-1] += rng.standard_normal((X - 2, X - 2, X - 2, 3))
    s.set_field(O.VELOCITY, v.reshape(-1))
    s.L.orc_set_dt(s.h, 1.0 / 240)
    s.L.orc_forces(s.h)
    s.L.orc_assemble(s.h)
    st = O.Stats()
    s.L.orc_solve(s.h, st)
    assert st.converged[1] == 1
    v1 = s.field(O.VELOCITY).reshape(X, X, X, 3)[1:-1, 1:-1, 1:-1].reshape(-1)
    v0 = s.field(O.V0)
    r = s.apply_J(v1) - s.field(O.DIV_RHS)
    assert np.linalg.norm(r) < 1e-10 * np.linalg.norm(s.apply_J(v0) - s.field(O.DIV_RHS))
    s.close()


def test_recover_and_dirichlet_faces():
    """the solid velocity replaces the grid velocity where the solid has mass
    inside its extended box (recoverVelocityField, DEFOR_SOLID_3D.cpp:
    1235-1241), then every face node's velocity is zero and the intensity on
    the y = 0 face is zero (COUPLED_INTEGRATOR_3D.cpp:900-916)"""
    n = 8
    dh = 1.0 / n
    s = block_scene(n, (2 * dh, 2 * dh, 2 * dh), (5 * dh, 5 * dh, 5 * dh))
    X = n + 1
    rng = np.random.default_rng(17)
    vg = rng.standard_normal((X ** 3, 3))
    vs = rng.standard_normal((X ** 3, 3))
    inten = rng.uniform(0.1, 1, X ** 3)
    s.set_field(O.VELOCITY, vg.reshape(-1))
    s.set_field(O.S_VELOCITY, vs.reshape(-1), 0)
    s.set_field(O.INTENSITY, inten)
    s.L.orc_recover_and_bc(s.h)
    got = s.field(O.VELOCITY).reshape(X, X, X, 3)
    m = s.field(O.S_MASS, 0).reshape(X, X, X)
    bx = s.boxes(0)[6:]
    want = vg.reshape(X, X, X, 3).copy()
    for z in range(bx[2], bx[5] - 1):
        for y in range(bx[1], bx[4] - 1):
            for x in range(bx[0], bx[3] - 1):
                if m[z, y, x] > 0:
                    want[z, y, x] = vs.reshape(X, X, X, 3)[z, y, x]
    for sl in ((0,), (X - 1,)):
        want[sl] = 0
        want[:, sl[0]] = 0
        want[:, :, sl[0]] = 0
    np.testing.assert_array_equal(got, want)
    iw = inten.reshape(X, X, X).copy()
    iw[:, 0, :] = 0
    np.testing.assert_array_equal(s.field(O.INTENSITY).reshape(X, X, X), iw)
    s.close()


def test_cfl_minimum_taken_as_alpha_over_the_largest_speed():
    """CFL (COUPLED_INTEGRATOR_3D.cpp:88-102) takes dt = min(1e9, min over
    solid nodes of alpha / vmax); the device reduces the largest vmax and the
    host divides once.  Division and its rounding are monotone, so the two
    agree bit for bit: checked on random speeds over many magnitudes, ties,
    zero speeds (alpha / 0 = inf, so dt = 1e9) and NaN speeds, which both
    forms skip (std::min keeps dt; fmax keeps the maximum)."""
    rng = np.random.default_rng(5)
    for trial in range(300):
        n = int(rng.integers(1, 400))
        v = np.exp(rng.uniform(-30, 30, n)) * rng.choice([1.0, 1.0 + 2 ** -52, 3.0], n)
        if trial % 7 == 0:
            v[rng.integers(0, n)] = 0.0
        if trial % 11 == 0:
            v[rng.integers(0, n)] = np.nan
        if trial % 13 == 0:
            v[:] = 0.0
        alpha = [0.6, 1.8, float(rng.uniform(0.1, 3))][trial % 3]
        dt_ref = 1e9
        for x in v:
            q = alpha / x if x != 0 else np.inf
            dt_ref = q if q < dt_ref else dt_ref  # std::min(dt, q), NaN never smaller
        vm = 0.0
        for x in v:
            vm = np.fmax(vm, x)
        dt_max = min(1e9, alpha / vm if vm != 0 else np.inf)
        assert dt_max == dt_ref, (trial, dt_max, dt_ref)


def test_scaling_by_a_power_of_two_reciprocal_equals_the_division():
    """The trilinear lookups (FIELD_3D.h:368-382 divide the offset by dh)
    multiply by 1 / dh where dh is a power of two.  The reciprocal is then
    exact and both operations round the same real quotient, so they agree
    bit for bit, including subnormal, huge, signed-zero and non-finite
    offsets; for a dh that is not a power of two (1/80, 1/24) they do not,
    which is why the other grids keep the division."""
    rng = np.random.default_rng(9)
    x = np.concatenate([
        rng.uniform(-2, 2, 20000),
        np.exp(rng.uniform(-700, 700, 20000)) * rng.choice([-1.0, 1.0], 20000),
        np.array([0.0, -0.0, 5e-324, -5e-324, 1.7e308, np.inf, -np.inf, np.nan]),
    ])
    with np.errstate(over="ignore", under="ignore", invalid="ignore"):
        for k in (1, 4, 7, 10, 16, 30):
            dh = 2.0 ** -k
            np.testing.assert_array_equal((x / dh).view(np.int64), (x * (1.0 / dh)).view(np.int64))
        for dh in (1.0 / 80, 1.0 / 24):
            assert np.any((x / dh) != (x * (1.0 / dh)))
