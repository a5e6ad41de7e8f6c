"""Python mirror of COUPLED_INTEGRATOR_3D (include/integrator/
COUPLED_INTEGRATOR_3D.h:30-76) over the C ABI of libesf.so.

Method names follow the reference's member functions; every call goes to
the HIP library and raises EsfError when it reports an error.
"""
import ctypes
import os
import tempfile

import numpy as np

from . import _abi
from ._abi import check, dptr


class CoupledIntegrator3D:
    """One scene resident on the GPU.  Create it from a reference cfg file
    (``from_cfg``) or from explicit grid/params (``__init__``)."""

    def __init__(self, grid, params, precision=_abi.ESF_PLAIN_FP64, _ctx=None):
        self.lib = _abi.load()
        self.grid = grid
        if _ctx is None:
            g = _abi.Grid(grid["xres"], grid["yres"], grid["zres"], grid["dh"],
                          (ctypes.c_double * 3)(*grid["origin"]))
            p = _abi.Params(**{k: v for k, v in params.items() if k in dict(_abi.Params._fields_)})
            p.precision = precision
            ctx = ctypes.c_void_p()
            check(None, self.lib.esf_ctx_create(ctypes.byref(ctx), ctypes.byref(g), ctypes.byref(p)),
                  "esf_ctx_create")
            _ctx = ctx
        self.ctx = _ctx
        X, Y, Z = grid["xres"], grid["yres"], grid["zres"]
        self.nodes = X * Y * Z
        self.cells = (X - 1) * (Y - 1) * (Z - 1)
        self.info = None

    @classmethod
    def from_cfg(cls, path, precision=_abi.ESF_PLAIN_FP64):
        """APPLICATION::init (projects/rubberBall/main.cpp:63-97) from a cfg."""
        from .cfg import Cfg
        lib = _abi.load()
        ctx = ctypes.c_void_p()
        info = _abi.SceneInfo()
        st = lib.esf_scene_load_cfg(path.encode(), precision, ctypes.byref(ctx), ctypes.byref(info))
        if st < 0:
            raise _abi.EsfError(st, f"esf_scene_load_cfg({path})")
        c = Cfg(path=path)
        grid = {"xres": c.get_int("xRes", 100), "yres": c.get_int("yRes", 100),
                "zres": c.get_int("zRes", 100), "dh": c.get_float("dh", 0.01),
                "origin": c.get_vec3("origin", (0, 0, 0))}
        obj = cls(grid, None, precision, _ctx=ctx)
        obj.info = info
        return obj

    @classmethod
    def from_cfg_text(cls, text, precision=_abi.ESF_PLAIN_FP64):
        fd, path = tempfile.mkstemp(suffix=".cfg")
        with os.fdopen(fd, "w") as f:
            f.write(text)
        try:
            return cls.from_cfg(path, precision)
        finally:
            os.unlink(path)

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if self.ctx:
            self.lib.esf_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self):
        """device memory this context holds"""
        v = ctypes.c_size_t()
        check(self.ctx, self.lib.esf_ctx_device_bytes(self.ctx, ctypes.byref(v)))
        return v.value

    # -- fields -------------------------------------------------------------
    def _shape(self, field):
        n, c = self.nodes, self.cells
        ns = self.num_solid_nodes()
        table = {
            _abi.F_VELOCITY: (3 * n, np.float64), _abi.F_INTENSITY: (n, np.float64),
            _abi.F_SOLID_MASS: (n, np.float64), _abi.F_NODE_FRACTION: (n, np.float64),
            _abi.F_LAMBDA: (c, np.float64), _abi.F_SOLID_INDEX: (n, np.int32),
            _abi.F_V0: (3 * n, np.float64), _abi.F_MINV: (n, np.float64),
            _abi.F_PRESSURE_RHS: (c, np.float64), _abi.F_DIV_RHS: (c, np.float64),
            _abi.F_SOLID_FORCE: (3 * ns, np.float64), _abi.F_SOLID_MV: (3 * ns, np.float64),
            _abi.F_DAMPED_MASS: (3 * ns, np.float64), _abi.F_SYS_DIAG: (3 * ns, np.float64),
            _abi.F_PRESSURE_DIAG: (c, np.float64), _abi.F_SOLID_V0: (3 * ns, np.float64),
            _abi.F_SOLID_DV: (3 * ns, np.float64), _abi.F_WORKSPACE: (n, np.float64),
            _abi.F_SOLID_RHS: (3 * ns, np.float64),
            _abi.S_U: (3 * n, np.float64), _abi.S_X: (3 * n, np.float64),
            _abi.S_VELOCITY: (3 * n, np.float64), _abi.S_MASS: (n, np.float64),
            _abi.S_VALID: (c, np.int8), _abi.S_QFRAC: (8 * n, np.float64),
            _abi.S_VEL_WORKSPACE: (3 * n, np.float64), _abi.S_PHI: (n, np.float64),
        }
        return table[field]

    def download(self, field, solid=-1):
        count, dtype = self._shape(field)
        out = np.empty(count, dtype=dtype)
        check(self.ctx, self.lib.esf_download(self.ctx, field, solid, out.ctypes.data_as(ctypes.c_void_p),
                                              out.nbytes), f"download {field}")
        return out

    def upload(self, field, arr, solid=-1):
        count, dtype = self._shape(field)
        a = np.ascontiguousarray(arr, dtype=dtype).reshape(-1)
        if a.size != count:
            raise ValueError(f"field {field}: expected {count} values, got {a.size}")
        check(self.ctx, self.lib.esf_upload(self.ctx, field, solid, a.ctypes.data_as(ctypes.c_void_p),
                                            a.nbytes), f"upload {field}")

    def synchronize(self):
        check(self.ctx, self.lib.esf_synchronize(self.ctx))

    def num_solid_nodes(self):
        n = ctypes.c_int()
        check(self.ctx, self.lib.esf_num_solid_nodes(self.ctx, ctypes.byref(n)))
        return n.value

    def boxes(self, solid=0):
        b = (ctypes.c_int * 12)()
        check(self.ctx, self.lib.esf_solid_get_boxes(self.ctx, solid, b))
        return list(b)

    def set_boxes(self, solid, boxes):
        b = (ctypes.c_int * 12)(*boxes)
        check(self.ctx, self.lib.esf_solid_set_boxes(self.ctx, solid, b))

    def _node_array(self, a, comps, name, copy=False):
        """a C-ordered float64 copy (or view) of a node field: (Z, Y, X) for
        comps == 1, (Z, Y, X, 3) for comps == 3, or its flat form; any other
        shape, or a size that is not the grid's, is refused"""
        X, Y, Z = self.grid["xres"], self.grid["yres"], self.grid["zres"]
        a = np.asarray(a)
        shape = (Z, Y, X) if comps == 1 else (Z, Y, X, 3)
        if a.shape not in (shape, (comps * self.nodes,)):
            raise ValueError(f"{name}: expected shape {shape} or ({comps * self.nodes},), "
                             f"got {a.shape}")
        return np.array(a, dtype=np.float64, order="C", copy=True) if copy else \
            np.ascontiguousarray(a, dtype=np.float64)

    @staticmethod
    def _vector(a, n, name, copy=False):
        a = np.asarray(a)
        if a.shape != (n,):
            raise ValueError(f"{name}: expected shape ({n},), got {a.shape}")
        return np.array(a, dtype=np.float64, order="C", copy=True) if copy else \
            np.ascontiguousarray(a, dtype=np.float64)

    # -- stages (COUPLED_INTEGRATOR_3D / DEFOR_SOLID_3D members) -------------
    def setMaxDt(self, dt):
        check(self.ctx, self.lib.esf_set_max_dt(self.ctx, dt))

    def setDt(self, dt):
        check(self.ctx, self.lib.esf_set_dt(self.ctx, dt))

    def dt(self):
        v = ctypes.c_double()
        check(self.ctx, self.lib.esf_get_dt(self.ctx, ctypes.byref(v)))
        return v.value

    def CFL(self, alpha):
        v = ctypes.c_double()
        check(self.ctx, self.lib.esf_cfl(self.ctx, alpha, ctypes.byref(v)), "CFL")
        return v.value

    def computeSolidMass(self, solid=0):
        check(self.ctx, self.lib.esf_compute_solid_mass(self.ctx, solid), "computeSolidMass")

    def computeSolidMassSdf(self, solid=0):
        """DEFOR_SOLID_3D::computeSolidMass(true): cells with phi > 2 skipped"""
        check(self.ctx, self.lib.esf_compute_solid_mass_sdf(self.ctx, solid), "computeSolidMass(true)")

    def updateBoxes(self, solid=0):
        check(self.ctx, self.lib.esf_update_boxes(self.ctx, solid), "updateBoxes")

    def solidSDF(self, solid=0):
        """DEFOR_SOLID_3D::computeSDF on the device: boxes and phi"""
        check(self.ctx, self.lib.esf_solid_sdf(self.ctx, solid), "solidSDF")

    def gatherSolidMasses(self):
        check(self.ctx, self.lib.esf_gather_solid_masses(self.ctx), "gatherSolidMasses")

    def recoverVelocityField(self):
        check(self.ctx, self.lib.esf_recover_velocity(self.ctx), "recoverVelocityField")

    def forces(self):
        e = ctypes.c_double()
        check(self.ctx, self.lib.esf_forces(self.ctx, ctypes.byref(e)), "forces")
        return e.value

    def pcrImplicitSolve(self):
        s = _abi.SolveStats()
        st = check(self.ctx, self.lib.esf_coupled_solve(self.ctx, ctypes.byref(s)), "coupled solve")
        d = s.as_dict()
        d["status"] = st
        return d

    def solidCopyVelocity(self):
        check(self.ctx, self.lib.esf_solid_copy_velocity(self.ctx))

    def solidFinalize(self):
        check(self.ctx, self.lib.esf_solid_finalize(self.ctx))

    def advectFluidVelocityField(self):
        check(self.ctx, self.lib.esf_advect_velocity(self.ctx))

    def advectIntensity(self):
        check(self.ctx, self.lib.esf_advect_intensity(self.ctx))

    def recoverAndBoundary(self):
        check(self.ctx, self.lib.esf_recover_and_bc(self.ctx))

    def stepImplicit(self, alpha=-1.0):
        """One drop-in substep (esf_step); returns the solve report."""
        s = _abi.SolveStats()
        st = check(self.ctx, self.lib.esf_step(self.ctx, alpha, ctypes.byref(s)), "stepImplicit")
        d = s.as_dict()
        d["status"] = st
        return d

    def frame(self, frame_dt, alpha=-1.0):
        n = ctypes.c_int()
        check(self.ctx, self.lib.esf_frame(self.ctx, frame_dt, alpha, ctypes.byref(n)), "frame")
        return n.value

    # -- operators ----------------------------------------------------------
    def setSurfaceMesh(self, solid, rest, vertices, triangles):
        """DEFOR_SOLID_3D::loadSurfaceMesh with the placed vertices"""
        r = np.ascontiguousarray(rest, dtype=np.float64).reshape(-1)
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1)
        check(self.ctx, self.lib.esf_solid_set_mesh(self.ctx, solid, r.size // 3, _abi.dptr(r),
                                                    _abi.dptr(v), t.size // 3,
                                                    t.ctypes.data_as(_abi._I)), "setSurfaceMesh")

    def surfaceMesh(self, solid=0):
        """(vertices (n, 3), rest pose (n, 3), triangles (m, 3)) or None"""
        nv, nt = ctypes.c_int(), ctypes.c_int()
        check(self.ctx, self.lib.esf_solid_mesh_size(self.ctx, solid, ctypes.byref(nv),
                                                     ctypes.byref(nt)), "surfaceMesh")
        if nv.value == 0:
            return None
        v = np.empty(3 * nv.value)
        r = np.empty(3 * nv.value)
        t = np.empty(3 * nt.value, dtype=np.int32)
        check(self.ctx, self.lib.esf_solid_get_mesh(self.ctx, solid, _abi.dptr(v), _abi.dptr(r),
                                                    t.ctypes.data_as(_abi._I)), "surfaceMesh")
        return v.reshape(-1, 3), r.reshape(-1, 3), t.reshape(-1, 3)

    def advectSurfaceMesh(self):
        """DEFOR_SOLID_3D::advectSurfaceMesh for every solid with a mesh"""
        check(self.ctx, self.lib.esf_advect_surface_mesh(self.ctx), "advectSurfaceMesh")

    def broadPhase(self):
        """CFL alpha of the stepImplicit broad phase: 0.6 or 1.8"""
        a = ctypes.c_double()
        check(self.ctx, self.lib.esf_broad_phase(self.ctx, ctypes.byref(a)), "broadPhase")
        return a.value

    def computeSDF(self, mass):
        """DEFOR_SOLID_3D::computeSDF on a node mass field (Z, Y, X) of this
        grid: (phi, box, tight), boxes as (xmin, ymin, zmin, xmax, ymax, zmax)"""
        m = self._node_array(mass, 1, "mass")
        phi = np.empty_like(m)
        box = np.zeros(6, dtype=np.int32)
        tight = np.zeros(6, dtype=np.int32)
        check(self.ctx, self.lib.esf_compute_sdf(self.ctx, dptr(m), dptr(phi), box.ctypes.data_as(_abi._I),
                                                 tight.ctypes.data_as(_abi._I)), "computeSDF")
        return phi, box, tight

    def computeSDFNoAlloc(self, mass):
        """computeSDF as computeSDF() does, through esf_compute_sdf_noalloc
        (no device allocation: the library's static scratch)"""
        m = self._node_array(mass, 1, "mass")
        phi = np.empty_like(m)
        box = np.zeros(6, dtype=np.int32)
        tight = np.zeros(6, dtype=np.int32)
        check(self.ctx, self.lib.esf_compute_sdf_noalloc(
            self.ctx, dptr(m), dptr(phi), box.ctypes.data_as(_abi._I),
            tight.ctypes.data_as(_abi._I)), "computeSDFNoAlloc")
        return phi, box, tight

    def extendVectorField(self, box, mass, phi, field):
        """DEFOR_SOLID_3D::extendVectorField of a (Z, Y, X, 3) node field where
        mass == 0, over box; returns the extended field"""
        m = self._node_array(mass, 1, "mass")
        ph = self._node_array(phi, 1, "phi")
        out = self._node_array(field, 3, "field", copy=True)
        b = np.ascontiguousarray(box, dtype=np.int32)
        if b.shape != (6,):
            raise ValueError(f"box: expected 6 ints, got shape {b.shape}")
        check(self.ctx, self.lib.esf_extend_vector_field(self.ctx, b.ctypes.data_as(_abi._I), dptr(m), dptr(ph),
                                                         dptr(out)), "extendVectorField")
        return out

    # -- the isolated system's solids -----------------------------------------
    def solidNumCells(self, solid=0):
        n = ctypes.c_int()
        check(self.ctx, self.lib.esf_solid_num_cells(self.ctx, solid, ctypes.byref(n)))
        return n.value

    def setDeformation(self, F, solid=0):
        """one 3x3 F per valid cell (None: back to the displacement)"""
        if F is None:
            check(self.ctx, self.lib.esf_solid_set_deformation(self.ctx, solid, 0, None))
            return
        f = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
        check(self.ctx, self.lib.esf_solid_set_deformation(self.ctx, solid, f.size // 9, dptr(f)),
              "setDeformation")

    def solidSolveResident(self, eps, max_it):
        """A x = b (b in F_SOLID_RHS) from x = 0 after forces(); returns
        (iterations, relative residual, device ms, status); x in F_SOLID_V0"""
        it, rr, ms = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        st = check(self.ctx, self.lib.esf_solid_solve_resident(
            self.ctx, eps, max_it, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(ms)),
            "solid solve")
        return it.value, rr.value, ms.value, st

    # -- FLIP particles (DEFOR_SOLID_3D) --------------------------------------
    def initMaterialParticles(self, solid=0):
        check(self.ctx, self.lib.esf_init_particles(self.ctx, solid), "initMaterialParticles")

    def gridToParticle(self, solid=0):
        check(self.ctx, self.lib.esf_grid_to_particle(self.ctx, solid), "gridToParticle")

    def particleToGrid(self, solid=0):
        check(self.ctx, self.lib.esf_particle_to_grid(self.ctx, solid), "particleToGrid")

    def extendSolidVelocity(self, solid=0):
        check(self.ctx, self.lib.esf_extend_solid_velocity(self.ctx, solid), "extendVectorField")

    def particles(self, solid=0):
        """positions and velocities of the solid's particles, (n, 3) each"""
        n = ctypes.c_long()
        check(self.ctx, self.lib.esf_num_particles(self.ctx, solid, ctypes.byref(n)))
        pos = np.empty(3 * n.value)
        vel = np.empty(3 * n.value)
        check(self.ctx, self.lib.esf_get_particles(self.ctx, solid, dptr(pos), dptr(vel)))
        return pos.reshape(-1, 3), vel.reshape(-1, 3)

    def setParticles(self, pos, vel, solid=0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1)
        if pos.size != vel.size or pos.size % 3:
            raise ValueError("positions and velocities: 3 doubles per particle each")
        check(self.ctx, self.lib.esf_set_particles(self.ctx, solid, pos.size // 3, dptr(pos),
                                                   dptr(vel)), "setParticles")

    def writeFrame(self, data_path, frame, debug=False):
        """COUPLED_INTEGRATOR_3D::writeFrame (COUPLED_INTEGRATOR_3D.cpp:999-1011)"""
        check(self.ctx, self.lib.esf_write_frame(self.ctx, data_path.encode(), frame,
                                                 1 if debug else 0), "writeFrame")

    def readFrame(self, data_path, frame, debug=False):
        """COUPLED_INTEGRATOR_3D::readFrame (COUPLED_INTEGRATOR_3D.cpp:976-997)"""
        check(self.ctx, self.lib.esf_read_frame(self.ctx, data_path.encode(), frame,
                                                1 if debug else 0), "readFrame")

    def getMatVecMult(self, x):
        x = self._vector(x, self.cells, "x")
        y = np.empty_like(x)
        check(self.ctx, self.lib.esf_pressure_apply(self.ctx, dptr(x), dptr(y)), "pressure apply")
        return y

    def pressureDiag(self):
        d = np.empty(self.cells)
        check(self.ctx, self.lib.esf_pressure_diag(self.ctx, dptr(d)))
        return d

    def solidApply(self, x):
        x = self._vector(x, 3 * self.num_solid_nodes(), "x")
        y = np.empty_like(x)
        check(self.ctx, self.lib.esf_solid_apply(self.ctx, dptr(x), dptr(y)))
        return y

    def pressureSolve(self, b, x0, eps, max_it):
        b = self._vector(b, self.cells, "b")
        x = self._vector(x0, self.cells, "x0", copy=True)
        it = ctypes.c_int()
        rr = ctypes.c_double()
        st = check(self.ctx, self.lib.esf_pressure_solve(self.ctx, dptr(b), dptr(x), eps, max_it,
                                                         ctypes.byref(it), ctypes.byref(rr)))
        return x, it.value, rr.value, st

    def pressureSolveResident(self, eps, max_it):
        """S x = b with b (F_PRESSURE_RHS) and the start x (F_LAMBDA) already
        uploaded, in the context's precision; returns (iterations,
        refinements, relative residual, device ms, status)"""
        it, ref = ctypes.c_int(), ctypes.c_int()
        rr, ms = ctypes.c_double(), ctypes.c_double()
        st = check(self.ctx, self.lib.esf_pressure_solve_resident(
            self.ctx, eps, max_it, ctypes.byref(it), ctypes.byref(ref), ctypes.byref(rr),
            ctypes.byref(ms)), "pressure solve")
        return it.value, ref.value, rr.value, ms.value, st

    # -- timing -------------------------------------------------------------
    def bench_pressure_apply(self, fp32=False, reps=50):
        v = ctypes.c_double()
        check(self.ctx, self.lib.esf_bench_pressure_apply(self.ctx, 1 if fp32 else 0, reps,
                                                          ctypes.byref(v)), "bench apply")
        return v.value

    def set_stage_timing(self, on=True):
        check(self.ctx, self.lib.esf_set_stage_timing(self.ctx, 1 if on else 0))

    def stage_times(self, reset=False):
        a = np.zeros(16)
        check(self.ctx, self.lib.esf_stage_times(self.ctx, dptr(a), 16, 1 if reset else 0))
        return dict(zip(_abi.STAGES, a[:len(_abi.STAGES)]))

    def bench_copy(self, nbytes, reps=20):
        v = ctypes.c_double()
        check(self.ctx, self.lib.esf_bench_copy(self.ctx, nbytes, reps, ctypes.byref(v)), "bench copy")
        return v.value

    def set_kernel_timing(self, on=True):
        check(self.ctx, self.lib.esf_set_kernel_timing(self.ctx, 1 if on else 0))

    def kernel_timing(self, reset=False):
        """{class: (total_ms, launches)} for the Krylov kernels timed so far."""
        nc = len(_abi.KERNEL_CLASSES)
        ms = np.zeros(nc)
        n = (ctypes.c_long * nc)()
        check(self.ctx, self.lib.esf_kernel_timing(self.ctx, dptr(ms), n, 1 if reset else 0))
        return {c: (float(ms[i]), int(n[i])) for i, c in enumerate(_abi.KERNEL_CLASSES)}


def corotation_batch(F, dF, lam, mu):
    """COROTATION on a batch of F (n,3,3) on the GPU: returns R, S, L, P,
    energy, dP(dF)."""
    lib = _abi.load()
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 9)
    dF = np.ascontiguousarray(dF, dtype=np.float64).reshape(-1, 9)
    n = F.shape[0]
    out = np.empty((n, 36))
    en = np.empty(n)
    dP = np.empty((n, 9))
    st = lib.esf_corotation_batch(n, lam, mu, dptr(F), dptr(dF), dptr(out), dptr(en), dptr(dP))
    if st < 0:
        raise _abi.EsfError(st, "esf_corotation_batch")
    o = out.reshape(n, 4, 3, 3)
    return o[:, 0], o[:, 1], o[:, 2], o[:, 3], en, dP.reshape(n, 3, 3)


def step_batch(scenes, alpha=-1.0):
    """One substep of every scene in `scenes` (CoupledIntegrator3D objects)
    through esf_step_batch: each kernel of the step launched once over the
    scenes that reach it together, each scene's results bitwise those of
    its own stepImplicit.  Returns the per-scene solve reports (with their
    "status"); raises EsfError when a scene's step fails."""
    lib = _abi.load()
    n = len(scenes)
    ctxs = (ctypes.c_void_p * n)(*[g.ctx.value for g in scenes])
    stats = (_abi.SolveStats * n)()
    status = (ctypes.c_int * n)()
    st = lib.esf_step_batch(ctxs, n, alpha, stats, status)
    out = []
    for i in range(n):
        d = stats[i].as_dict()
        d["status"] = status[i]
        out.append(d)
    if st < 0:
        bad = [i for i in range(n) if status[i] < 0]
        check(scenes[bad[0]].ctx if bad else scenes[0].ctx, st, "esf_step_batch")
    return out

