"""GPU (libesf.so) against the CPU oracle on the same scenes: per stage and
per step.  Integer/byte work and reordered-but-identical arithmetic must be
bit-exact; Krylov outputs must agree to 1e-10 relative (the parity bar of
the FP64 form), both sides solving to 1e-13."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._scenes import ball_cfg, rel_err, solution_order

pytestmark = pytest.mark.gpu

TIGHT = dict(eps=1e-13, max_iterations=20000)


@pytest.fixture(scope="module")
def esf():
    import eulersf_amd
    eulersf_amd.load()
    return eulersf_amd


def make_pair(esf, n=16, **kw):
    text = ball_cfg(n, **kw)
    g = esf.CoupledIntegrator3D.from_cfg_text(text)
    o = O.OracleScene.from_cfg_text(text)
    return g, o


def test_corotation_batch_matches_oracle(esf):
    rng = np.random.default_rng(1)
    F = np.concatenate([
        np.eye(3) + 0.2 * rng.standard_normal((200, 3, 3)),
        rng.standard_normal((200, 3, 3)),
        np.eye(3)[None].repeat(4, 0),
    ])
    dF = rng.standard_normal(F.shape)
    lam, mu = 5769.2307692307695, 3846.153846153846
    R, S, L, P, e, dP = esf.corotation_batch(F, dF, lam, mu)
    scale = lam + mu
    for i in range(F.shape[0]):
        r = O.corotation(lam, mu, F[i], dF[i])
        assert np.max(np.abs(R[i] - r[0])) < 1e-12
        assert np.max(np.abs(S[i] - r[1])) < 1e-12 * max(1, np.abs(r[1]).max())
        assert np.max(np.abs(L[i] - r[2])) < 1e-12 * scale
        assert np.max(np.abs(P[i] - r[3])) < 1e-12 * scale * max(1, np.abs(F[i]).max())
        assert abs(e[i] - r[4]) <= 1e-12 * scale * max(1, np.abs(F[i]).max() ** 2)
        assert np.max(np.abs(dP[i] - r[5])) < 1e-11 * scale * max(1, np.abs(dF[i]).max())


def test_scene_init_bit_exact(esf):
    g, o = make_pair(esf, 16)
    assert g.num_solid_nodes() == o.num_solid_nodes()
    assert g.boxes(0) == o.boxes(0)
    for gf, of in [(esf._abi.S_MASS, O.S_MASS), (esf._abi.S_QFRAC, O.S_QFRAC),
                   (esf._abi.S_X, O.S_X), (esf._abi.S_U, O.S_U)]:
        np.testing.assert_array_equal(g.download(gf, 0), o.field(of, 0))
    np.testing.assert_array_equal(g.download(esf._abi.S_VALID, 0), o.field(O.S_VALID, 0))
    np.testing.assert_array_equal(g.download(esf._abi.F_SOLID_MASS), o.field(O.SOLID_MASS))
    np.testing.assert_array_equal(g.download(esf._abi.F_NODE_FRACTION), o.field(O.NODE_FRACTION))
    np.testing.assert_array_equal(g.download(esf._abi.F_SOLID_INDEX), o.field(O.SOLID_INDEX))
    np.testing.assert_array_equal(g.download(esf._abi.F_VELOCITY), o.field(O.VELOCITY))
    np.testing.assert_array_equal(g.download(esf._abi.F_INTENSITY), o.field(O.INTENSITY))


def test_forces_and_pressure_operator(esf):
    g, o = make_pair(esf, 16, velocity=(0.3, -1.0, 0.2))
    X = Y = Z = 17
    dt_g = g.CFL(1.8)
    dt_o = o.L.orc_cfl(o.h, 1.8)
    assert dt_g == dt_o
    g.forces()
    o.L.orc_forces(o.h)
    o.L.orc_assemble(o.h)
    # solid force, momentum, damped mass: same sums in the same order
    np.testing.assert_array_equal(g.download(esf._abi.F_SOLID_FORCE), o.field(O.SOLID_FORCE))
    np.testing.assert_array_equal(g.download(esf._abi.F_SOLID_MV), o.field(O.SOLID_MV))
    np.testing.assert_array_equal(g.download(esf._abi.F_DAMPED_MASS), o.field(O.DAMPED_MASS))
    minv_g = solution_order(np.repeat(g.download(esf._abi.F_MINV), 3), X, Y, Z)
    np.testing.assert_array_equal(minv_g, o.field(O.MINV))
    np.testing.assert_array_equal(g.download(esf._abi.F_WORKSPACE), o.field(O.WORKSPACE))
    # the pressure operator: bitwise the reference's sparse order
    rng = np.random.default_rng(3)
    x = rng.standard_normal(g.cells)
    np.testing.assert_array_equal(g.getMatVecMult(x), o.pressure_apply(x))
    np.testing.assert_array_equal(g.pressureDiag(), o.pressure_diag())


@pytest.mark.parametrize("n", [9, 13, 21])
def test_pressure_operator_uneven_tiles(esf, n):
    """S x on grids whose cell counts leave partial 32x8 tiles in x and y
    and a block of one cell layer at the top of the z-march: still bitwise
    the reference's sparse order."""
    g, o = make_pair(esf, n, velocity=(0.3, -1.0, 0.2))
    g.CFL(1.8)
    o.L.orc_cfl(o.h, 1.8)
    g.forces()
    o.L.orc_forces(o.h)
    o.L.orc_assemble(o.h)
    rng = np.random.default_rng(n)
    for _ in range(2):
        x = rng.standard_normal(g.cells)
        np.testing.assert_array_equal(g.getMatVecMult(x), o.pressure_apply(x))


@pytest.mark.parametrize("layers,depth", [(4, 1), (8, 2), (16, 2), (16, 3), (16, 1), (24, 2)])
def test_pressure_launch_shapes(esf, monkeypatch, layers, depth):
    """every z-march shape the tuning knobs select (cell layers per block,
    load sets in flight; read when a context is created): S x stays bitwise
    the reference's order, and the coupled solve, whose CR update and
    checks run in the same shape, stays within 1e-10 of the oracle.  24
    cell layers leave a partial block at 16."""
    monkeypatch.setenv("ESF_SPMV_LAYERS", str(layers))
    monkeypatch.setenv("ESF_SPMV_DEPTH", str(depth))
    g, o = make_pair(esf, 24, velocity=(0.3, -1.0, 0.2), **TIGHT)
    g.CFL(1.8)
    o.L.orc_cfl(o.h, 1.8)
    g.forces()
    o.L.orc_forces(o.h)
    o.L.orc_assemble(o.h)
    x = np.random.default_rng(layers + depth).standard_normal(g.cells)
    np.testing.assert_array_equal(g.getMatVecMult(x), o.pressure_apply(x))
    gs = g.pcrImplicitSolve()
    o.L.orc_solve(o.h, O.Stats())
    assert gs["status"] == 0
    assert rel_err(g.download(esf._abi.F_VELOCITY), o.field(O.VELOCITY)) < 1e-10


def test_coupled_solve_matches_oracle(esf):
    g, o = make_pair(esf, 16, velocity=(0.3, -1.0, 0.2), **TIGHT)
    g.CFL(1.8)
    o.L.orc_cfl(o.h, 1.8)
    g.forces()
    o.L.orc_forces(o.h)
    o.L.orc_assemble(o.h)
    gs = g.pcrImplicitSolve()
    st = O.Stats()
    o.L.orc_solve(o.h, st)
    assert gs["status"] == 0
    v_g = g.download(esf._abi.F_VELOCITY)
    v_o = o.field(O.VELOCITY)
    assert rel_err(v_g, v_o) < 1e-10
    np.testing.assert_allclose(g.download(esf._abi.F_SYS_DIAG), o.field(O.SYS_DIAG), rtol=1e-12)
    np.testing.assert_allclose(g.download(esf._abi.F_DIV_RHS), o.field(O.DIV_RHS), rtol=0, atol=0)


def test_steps_match_oracle(esf):
    g, o = make_pair(esf, 16, velocity=(0.3, -1.0, 0.2), **TIGHT)
    for k in range(3):
        gs = g.stepImplicit(1.8)
        rc, st = o.step(1.8)
        assert gs["dt"] == st.dt
        v_g, v_o = g.download(esf._abi.F_VELOCITY), o.field(O.VELOCITY)
        assert rel_err(v_g, v_o) < 1e-10, k
        u_g, u_o = g.download(esf._abi.S_U, 0), o.field(O.S_U, 0)
        assert rel_err(u_g, u_o) < 1e-10, k
        i_g, i_o = g.download(esf._abi.F_INTENSITY), o.field(O.INTENSITY)
        assert rel_err(i_g, i_o) < 1e-10, k


def test_material_positions_and_velocity_workspace_follow_the_reference(esf):
    """free-running steps write _X and _vecWorkspace only inside the
    extended box (the rest of _X when a caller reads it): downloaded after
    each step both equal the oracle's, and an upload of u leaves _X at x - u
    of the u it replaced, as in the reference"""
    A = esf._abi
    g, o = make_pair(esf, 16, velocity=(0.3, -1.0, 0.2), **TIGHT)
    fields = ((A.S_X, O.S_X), (A.S_VEL_WORKSPACE, O.S_VEL_WORKSPACE),
              (A.S_VELOCITY, O.S_VELOCITY), (A.S_MASS, O.S_MASS), (A.S_U, O.S_U))

    def same(k):
        for gf, of in fields:
            assert rel_err(g.download(gf, 0), o.field(of, 0)) < 1e-10, (k, gf)

    for k in range(2):
        g.stepImplicit(1.8)
        o.step(1.8)
        same(k)
    x_before = g.download(A.S_X, 0)
    u2 = o.field(O.S_U, 0) * 1.5
    g.upload(A.S_U, u2, 0)
    o.set_field(O.S_U, u2, 0)
    np.testing.assert_array_equal(g.download(A.S_X, 0), x_before)
    same("upload")
    for k in range(2, 4):
        g.stepImplicit(1.8)
        o.step(1.8)
        same(k)
    g.close()
    o.close()


def test_material_positions_outside_the_box_written_before_a_wider_mass_pass(esf):
    """A step writes _X = x - u inside the extended box only.  A block of
    nodes outside it is displaced so that x - u lands inside the ball's rest
    shape; a mass pass over a tight box widened by the caller to that block
    then finds mass there, equal bit for bit to a context whose _X was
    downloaded (written out at every node) first."""
    A = esf._abi
    n = 80
    N = n + 1
    dh = 1.0 / n
    ctxs = [esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(n, velocity=(0.3, -1.0, 0.2)))
            for _ in range(2)]
    ext = ctxs[0].boxes(0)[6:]
    assert ext[0] >= 6, ext  # the block x < 6 lies outside the extended box
    pos = np.stack(np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")[::-1],
                   axis=-1) * dh
    u = ctxs[0].download(A.S_U, 0).reshape(N, N, N, 3).copy()
    X0 = ctxs[0].download(A.S_X, 0).reshape(N, N, N, 3)
    # the block takes the material positions of the nodes 37 to its right,
    # around the ball's middle
    z, y, x = np.meshgrid(np.arange(36, 45), np.arange(46, 55), np.arange(1, 6), indexing="ij")
    u[z, y, x] = pos[z, y, x] - X0[z, y, x + 37]
    for g in ctxs:
        g.upload(A.S_U, u.reshape(-1), 0)
        st = g.stepImplicit(1.8)
        assert st["status"] == 0
    ctxs[1].download(A.S_X, 0)
    wide = [1, 30, 30, 60, 60, 60, 0, 0, 0, N, N, N]
    for g in ctxs:
        g.set_boxes(0, wide)
        g.computeSolidMass(0)
    m0, m1 = (g.download(A.S_MASS, 0).reshape(N, N, N) for g in ctxs)
    np.testing.assert_array_equal(m0, m1)
    assert m0[36:45, 46:55, 1:6].sum() > 0
    uu = ctxs[0].download(A.S_U, 0).reshape(N, N, N, 3)
    np.testing.assert_array_equal(ctxs[0].download(A.S_X, 0).reshape(N, N, N, 3),
                                  pos + (0.0 - uu))
    for g in ctxs:
        g.close()


def test_mixed_form_steps_within_1e6_of_fp64_oracle(esf):
    """the mixed form (FP32 CR inside FP64 refinement) solving to 1e-8
    against the FP64 oracle solving to 1e-13: after three steps every
    field within 1e-6 relative (the bound set for the mixed form)"""
    kw = dict(velocity=(0.3, -1.0, 0.2))
    g = esf.CoupledIntegrator3D.from_cfg_text(ball_cfg(24, eps=1e-8, max_iterations=20000, **kw),
                                              esf.ESF_MIXED)
    o = O.OracleScene.from_cfg_text(ball_cfg(24, **TIGHT, **kw))
    for k in range(3):
        gs = g.stepImplicit(1.8)
        rc, st = o.step(1.8)
        assert gs["status"] == 0 and rc == 0
        assert abs(gs["dt"] - st.dt) <= 1e-6 * st.dt
    for gf, of, solid in ((esf._abi.F_VELOCITY, O.VELOCITY, -1),
                          (esf._abi.F_INTENSITY, O.INTENSITY, -1), (esf._abi.S_U, O.S_U, 0)):
        err = rel_err(g.download(gf, solid), o.field(of, solid))
        assert err < 1e-6, (gf, err)
    g.close()
    o.close()


def three_sphere_cfg(n=24):
    """three spheres of different stiffness, two of them one cell apart so
    their extended boxes (and possibly nodes) are shared"""
    from eulersf_amd.cfg import scene_cfg
    spheres = [
        {"center": (0.3, 0.62, 0.5), "radius": 0.14, "E": 1e4, "velocity": (0.2, -0.6, 0.0)},
        {"center": (0.62, 0.62, 0.5), "radius": 0.14, "E": 4e4, "velocity": (-0.3, -0.2, 0.1)},
        {"center": (0.5, 0.3, 0.7), "radius": 0.1, "E": 2e3},
    ]
    return scene_cfg(n, spheres, fill=0.35, **TIGHT)


def test_three_solids_match_oracle(esf):
    """several solids in one context: mass gather, per-solid stiffness rows
    and masks, extended boxes -- two teacher-free steps against the oracle"""
    text = three_sphere_cfg()
    g = esf.CoupledIntegrator3D.from_cfg_text(text)
    o = O.OracleScene.from_cfg_text(text)
    assert g.num_solid_nodes() == o.num_solid_nodes()
    np.testing.assert_array_equal(g.download(esf._abi.F_SOLID_MASS), o.field(O.SOLID_MASS))
    for k in range(2):
        gs = g.stepImplicit(1.8)
        rc, st = o.step(1.8)
        assert gs["dt"] == st.dt
        assert rel_err(g.download(esf._abi.F_VELOCITY), o.field(O.VELOCITY)) < 1e-10, k
        for b in range(3):
            assert rel_err(g.download(esf._abi.S_U, b), o.field(O.S_U, b)) < 1e-10, (k, b)


def test_isolated_pressure_system_matches_oracle(esf):
    """the isolated system of the 256^3 benchmark, at 24^3 cells: 8 seeded
    spheres, b = J u.  FP64: the reference's CR from x = 0 to 1e-13, same
    iteration count (one check interval) and S x as the oracle's CR.  Mixed:
    FP32 CR in FP64 refinement reaches 1e-8 with S x within 1e-6."""
    import ctypes
    import bench
    from eulersf_amd.cfg import random_spheres, scene_cfg
    n = 24
    rng = np.random.default_rng(2026)
    spheres = random_spheres(rng, 8, (0.05, 0.12), (3.0, 6.0), center_range=(0.15, 0.85))
    text = scene_cfg(n, spheres, fill=0.6, **TIGHT)
    X = n + 1
    u = np.zeros((X ** 3, 3))
    inner = np.zeros((X, X, X), dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    u[inner.reshape(-1)] = rng.standard_normal((int(inner.sum()), 3))
    b = bench.divergence_field(u, n)
    o = O.OracleScene.from_cfg_text(text)
    o.L.orc_forces(o.h)
    xo = np.zeros(n ** 3)
    it_o = ctypes.c_int()
    assert o.L.orc_pressure_cr(o.h, O.dp(b), O.dp(xo), 1e-13, 20000, ctypes.byref(it_o)) == 0
    Sxo = o.pressure_apply(xo)
    for prec, eps, tol in ((esf.ESF_PLAIN_FP64, 1e-13, 1e-10), (esf.ESF_MIXED, 1e-8, 1e-6)):
        g = esf.CoupledIntegrator3D.from_cfg_text(text, prec)
        g.forces()
        g.pressureDiag()
        g.upload(esf._abi.F_PRESSURE_RHS, b)
        g.upload(esf._abi.F_LAMBDA, np.zeros(n ** 3))
        it, ref, rr, ms, st = g.pressureSolveResident(eps, 20000)
        assert st == 0 and rr < eps, (prec, rr)
        x = g.download(esf._abi.F_LAMBDA)
        assert rel_err(g.getMatVecMult(x), Sxo) < tol, prec
        if prec == esf.ESF_PLAIN_FP64:
            assert abs(it - it_o.value) <= 5, (it, it_o.value)
            assert ref == 0
        else:
            assert ref >= 1
        g.close()
    o.close()


@pytest.mark.parametrize("eps,max_it", [(1e-4, 20000), (1e-13, 6)])
def test_solid_check_returns_checked_iterate(esf, eps, max_it):
    """the true-residual check of iteration i runs inside iteration i + 1's
    passes; the solve must still return iteration i's x and count.  The
    first solid solve of the coupled solve is compared with the oracle's
    (same right-hand side): at 1e-4 one extra iteration would move x by
    about 1e-4 relative; with 6 iterations at most, the check after the
    last iteration runs alone and the solve stops unconverged, as the
    oracle's does"""
    g, o = make_pair(esf, 16, velocity=(0.3, -1.0, 0.2), eps=eps, max_iterations=max_it)
    g.CFL(1.8)
    o.L.orc_cfl(o.h, 1.8)
    g.forces()
    o.L.orc_forces(o.h)
    o.L.orc_assemble(o.h)
    gs = g.pcrImplicitSolve()
    st = O.Stats()
    o.L.orc_solve(o.h, st)
    assert gs["iterations"][0] == st.iterations[0], (gs, list(st.iterations))
    assert gs["converged"][0] == st.converged[0] == (1 if max_it > 6 else 0)
    assert rel_err(g.download(esf._abi.F_SOLID_V0), o.field(O.SOLID_V0)) < 1e-12
    g.close()
    o.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_step_overlap_bitwise_equals_staged_sequence(esf, mixed):
    """esf_step runs the fluid's vorticity and body forces on a second stream
    beside the first solid solve and queues the stress stage ahead of the dt
    read-back; the stage entry points called one after another on one stream
    must give bitwise the same fields (no race, no reordering of
    arithmetic), step after step, in both forms"""
    text = ball_cfg(24, velocity=(0.3, -1.0, 0.2), eps=1e-10, max_iterations=20000)
    prec = esf.ESF_MIXED if mixed else esf.ESF_PLAIN_FP64
    a = esf.CoupledIntegrator3D.from_cfg_text(text, prec)
    b = esf.CoupledIntegrator3D.from_cfg_text(text, prec)
    fields = [(esf._abi.F_VELOCITY, -1), (esf._abi.F_INTENSITY, -1), (esf._abi.S_U, 0),
              (esf._abi.S_VELOCITY, 0), (esf._abi.S_MASS, 0), (esf._abi.S_PHI, 0)]
    for k in range(3):
        sa = a.stepImplicit(1.8)
        dt = b.CFL(1.8)
        b.forces()
        sb = b.pcrImplicitSolve()
        b.solidCopyVelocity()
        b.solidFinalize()
        b.computeSolidMassSdf(0)
        b.advectSurfaceMesh()
        b.solidSDF(0)
        b.gatherSolidMasses()
        b.advectFluidVelocityField()
        b.advectIntensity()
        b.recoverAndBoundary()
        assert sa["dt"] == dt, k
        for f, s in fields:
            va, vb = a.download(f, s), b.download(f, s)
            assert np.array_equal(va.view(np.int64), vb.view(np.int64)), (k, f)
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fluid_only_scene_steps_match_oracle(esf, precision):
    """A scene without solids (the pool under gravity, a box of fluid): no
    solid DOFs, no boxes, no mass pass; dt from the default, the broad phase
    without meshes.  Three free-running steps of esf_step against the
    oracle's: velocity, intensity and pressure within 1e-10 (FP64) / 1e-6
    (mixed), dt equal."""
    from eulersf_amd.cfg import scene_cfg
    text = scene_cfg(16, [], fill=0.3, eps=1e-12)
    prec = esf.ESF_PLAIN_FP64 if precision == "fp64" else esf.ESF_MIXED
    g = esf.CoupledIntegrator3D.from_cfg_text(text, prec)
    o = O.OracleScene.from_cfg_text(text)
    tol = 1e-10 if precision == "fp64" else 1e-6
    A = esf._abi
    try:
        assert g.num_solid_nodes() == 0
        for step in range(3):
            r = g.stepImplicit(-1.0)
            rc, st = o.step(-1.0)
            assert r["dt"] == st.dt, (step, r["dt"], st.dt)
            assert r["n_solid_nodes"] == 0
            # the pool stays at rest: the velocity is round-off, measured
            # against what one step of gravity would give
            dv = np.abs(g.download(A.F_VELOCITY) - o.field(O.VELOCITY)).max()
            assert dv <= tol * 9.8 * st.dt, (step, dv)
            assert rel_err(g.download(A.F_INTENSITY), o.field(O.INTENSITY)) < tol, step
            # the pressure up to the constant null mode of S: compare J^T lambda
            lam_g, lam_o = g.download(A.F_LAMBDA), o.field(O.LAMBDA)
            assert rel_err(o.apply_negJT(lam_g), o.apply_negJT(lam_o)) < tol, step
    finally:
        g.close()
        o.close()
