/*
 * esf.h -- C ABI of libesf.so, the MI355X (gfx950) drop-in for the
 * stress, coupled-solve and advection stages of the reference's
 * COUPLED_INTEGRATOR_3D::stepImplicit().
 *
 * The reference has no C ABI; these entry points replace C++ member calls
 * (cited per entry point, paths relative to the reference root).  Its only
 * C-style extension point is the matrix-free callback of the CR solver
 *   bool solvePCR(const VECTOR& b, VECTOR& x,
 *                 void (*getMatVecMult)(void*, const VECTOR&, VECTOR&), void*)
 *   (include/linearalgebra/CONJUGATE_RESIDUAL.h:38-40), served by
 *   COUPLED_INTEGRATOR_3D::staticGetMatVecMult
 *   (include/integrator/COUPLED_INTEGRATOR_3D.h:69-73);
 * esf_pressure_apply() has exactly that callback's semantics.
 *
 * Conventions
 *  - Every call returns esf_status: 0 = ok, 1 = a Krylov solve did not reach
 *    its tolerance (reported, not fatal, as in CONJUGATE_RESIDUAL.inl:186-189),
 *    negative = error; esf_last_error() describes the last error.
 *  - Host buffers are owned by the caller and never retained.  Device state
 *    (grids, solids, Krylov vectors) is owned by the context and stays
 *    resident in HBM between calls.
 *  - One host thread per context; each context owns one HIP stream.
 *    Contexts are independent (no global state), unlike the reference's
 *    function-local statics (COUPLED_INTEGRATOR_3D.cpp:639, :712).
 *  - Layouts are the reference's: node n = x + y*X + z*X*Y (FIELD_3D.h:72);
 *    vector fields are AoS, 3 doubles per node (VEC3F_FIELD_3D.h:64-72);
 *    pressure cell r = x + y*(X-1) + z*(X-1)*(Y-1)
 *    (COUPLED_INTEGRATOR_3D.cpp:302-306); quarter fractions [8*n + corner];
 *    solid DOFs 3*s + c with s numbered in z,y,x order over interior nodes
 *    carrying solid mass (COUPLED_INTEGRATOR_3D.cpp:361-372).
 *  - The pressure solve warm-starts from the previous call's lambda, as the
 *    reference's static lambda1 does (COUPLED_INTEGRATOR_3D.cpp:712).
 */
#ifndef ESF_H
#define ESF_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct esf_ctx esf_ctx;
typedef int esf_status;

#define ESF_OK 0
#define ESF_NOT_CONVERGED 1
#define ESF_ERR_INVALID (-1)
#define ESF_ERR_OOM (-2)
#define ESF_ERR_DEVICE (-3)
#define ESF_ERR_NAN (-4)
#define ESF_ERR_STATE (-5)
#define ESF_ERR_IO (-6)

/* precision of the pressure solve */
#define ESF_PLAIN_FP64 0  /* CR in FP64, the reference's algorithm */
#define ESF_MIXED 1       /* FP32 CR inside FP64 iterative refinement */

typedef struct {
  int xres, yres, zres; /* node counts ("xRes","yRes","zRes") */
  double dh;            /* "dh" */
  double origin[3];     /* "origin" */
} esf_grid;

/* cfg-derived parameters; keys and defaults are the reference's
 * (COUPLED_INTEGRATOR_3D.cpp:43-74, :445-446, :499-501;
 *  CONJUGATE_RESIDUAL.inl:26-27; DEFOR_SOLID_3D.cpp:62, :243-244). */
typedef struct {
  double dt;                 /* "dt" */
  double vorticity_eps;      /* "vorticity eps" as written (scaled inside) */
  double gravity;            /* "gravity" */
  double buoyancy;           /* "buoyancy" */
  double rayleigh_alpha;     /* "rayleigh alpha" */
  double rayleigh_beta;      /* "rayleigh beta" */
  double fluid_density;      /* "fluid density" */
  double implicit_eps;       /* "implicit solver eps" (solid solves) */
  double divfree_eps;        /* "div free solver eps" (pressure solve) */
  int cr_max_iterations;     /* "cr max iterations" */
  double mass_sample_dh;     /* "mass sample dh" */
  int tight_border;          /* "tight border" */
  int extrapolation_cells;   /* "extrapolation cells" */
  int precision;             /* ESF_PLAIN_FP64 or ESF_MIXED */
  double inner_eps;          /* mixed: FP32 inner solve relative tolerance */
  /* FLIP material particles (DEFOR_SOLID_3D.cpp:336-337, 901, 948-952,
   * 1101); esf_step runs the particle round trip only when flip != 0 */
  int particles_per_cell;    /* "number of particles per material cell" (8) */
  double flip_alpha;         /* "flip alpha" (0.95) */
  int flip_cutoff;           /* "flip cutoff" (3) */
  double wall_thickness;     /* "wall thickness" (1.5 dh) */
  double spring_stiffness;   /* "contact spring stiffness" (1e4) */
  double friction;           /* "friction" (0.5) */
  int flip;                  /* "flip particles" (0; not a reference key) */
} esf_params;

/* per-step solve report */
typedef struct {
  int iterations[3];      /* solid solve 1, pressure solve, solid solve 2 */
  int converged[3];
  double rel_residual[3]; /* true |Ax-b|/|b| at exit, FP64 */
  int refinements;        /* mixed: outer FP64 refinement steps */
  int inner_iterations;   /* mixed: FP32 iterations summed over refinements */
  double dt;              /* the substep's dt */
  int n_solid_nodes;      /* solid DOF nodes this step */
  int n_solid_cells;      /* valid solid cells (all solids) this step */
} esf_solve_stats;

/* field ids for esf_upload / esf_download */
enum {
  ESF_F_VELOCITY = 0,      /* double[3*nodes]  _velocity                 */
  ESF_F_INTENSITY = 1,     /* double[nodes]    _intensity                */
  ESF_F_SOLID_MASS = 2,    /* double[nodes]    _solidMass (download)     */
  ESF_F_NODE_FRACTION = 3, /* double[nodes]    _nodeSolidVolumeFraction  */
  ESF_F_LAMBDA = 4,        /* double[cells]    pressure (warm start)     */
  ESF_F_SOLID_INDEX = 6,   /* int32[nodes]     _gridIdxToSolidIdx        */
  ESF_F_MINV = 9,          /* double[nodes]    1/m at interior nodes     */
  ESF_F_PRESSURE_RHS = 10, /* double[cells]    b = J v0 - c (writable)   */
  ESF_F_DIV_RHS = 11,      /* double[cells]    c                         */
  ESF_F_V0 = 7,            /* double[3*nodes]  v0 on interior nodes      */
  ESF_F_SOLID_FORCE = 13,  /* double[3*nSolid] elastic+body force (pre-dt) */
  ESF_F_SOLID_MV = 14,     /* double[3*nSolid] */
  ESF_F_DAMPED_MASS = 15,  /* double[3*nSolid] */
  ESF_F_SYS_DIAG = 20,     /* double[3*nSolid] diag(A) */
  ESF_F_PRESSURE_DIAG = 21,/* double[cells]    diag(J M^-1 J^T) */
  ESF_F_SOLID_RHS = 16,    /* double[3*nSolid] dt F + M v (writable)   */
  ESF_F_SOLID_V0 = 18,     /* double[3*nSolid] */
  ESF_F_SOLID_DV = 19,     /* double[3*nSolid] */
  ESF_F_WORKSPACE = 23,    /* double[nodes]    |omega| */
  /* per solid (pass the solid id) -- DEFOR_SOLID_3D members */
  ESF_S_U = 100,           /* double[3*nodes]  _u            */
  ESF_S_X = 101,           /* double[3*nodes]  _X            */
  ESF_S_VELOCITY = 102,    /* double[3*nodes]  _velocity     */
  ESF_S_MASS = 103,        /* double[nodes]    _mass         */
  ESF_S_VALID = 104,       /* int8[cells]      _valid        */
  ESF_S_QFRAC = 105,       /* double[8*nodes]  _quarterVolumeFractions */
  ESF_S_VEL_WORKSPACE = 106,/* double[3*nodes] _vecWorkspace */
  ESF_S_PHI = 108          /* double[nodes]    _phi (see esf_solid_sdf) */
};

/* --- context --------------------------------------------------------- */
/* COUPLED_INTEGRATOR_3D ctor (COUPLED_INTEGRATOR_3D.cpp:36-78) */
esf_status esf_ctx_create(esf_ctx** out, const esf_grid* grid,
                          const esf_params* params);
void esf_ctx_destroy(esf_ctx* ctx);
const char* esf_last_error(const esf_ctx* ctx);
esf_status esf_set_params(esf_ctx* ctx, const esf_params* params);
/* COUPLED_INTEGRATOR_3D::setMaxDt / dt (COUPLED_INTEGRATOR_3D.h:45-46) */
esf_status esf_set_max_dt(esf_ctx* ctx, double max_dt);
esf_status esf_set_dt(esf_ctx* ctx, double dt);
esf_status esf_get_dt(esf_ctx* ctx, double* dt);
esf_status esf_num_solid_nodes(esf_ctx* ctx, int* n);
/* wait for all work queued on the context's stream */
esf_status esf_synchronize(esf_ctx* ctx);
/* device memory the context holds (bytes); no reference counterpart: the
 * serving loop admits scenes against an HBM budget with it */
esf_status esf_ctx_device_bytes(const esf_ctx* ctx, size_t* bytes);

/* --- solids -------------------------------------------------------------*/
/* DEFOR_SOLID_3D ctor + setMaterial + setMassDensity
 * (DEFOR_SOLID_3D.cpp:43-84, :113-124; DEFOR_SOLID_3D.h:64-66); rest mass
 * is the FIELD_3Df the ctor takes (read from "<name>.mass" by
 * rubberBall/main.cpp:108-111). */
esf_status esf_solid_add(esf_ctx* ctx, double lambda, double mu,
                         double density, int rest_x, int rest_y, int rest_z,
                         double rest_dh, const double rest_origin[3],
                         const double* rest_mass, int* solid_id);
/* tight box and extended box, as tightMin,tightMax,min,max (12 ints) */
esf_status esf_solid_set_boxes(esf_ctx* ctx, int solid, const int boxes[12]);
esf_status esf_solid_get_boxes(esf_ctx* ctx, int solid, int boxes[12]);

/* Surface mesh of a solid (DEFOR_SOLID_3D::loadSurfaceMesh,
 * DEFOR_SOLID_3D.cpp:157-164): nv vertices as the rest pose (the OBJ's own
 * coordinates, OBJ.cpp:255) and as placed in the scene (scaled, rotated,
 * translated, rubberBall/main.cpp:148-152), 3 doubles each; nt triangles,
 * 3 zero-based vertex indices each (OBJ.cpp:256-277).  nv = 0 removes the
 * mesh.  esf_step advects the mesh of every solid that has one. */
esf_status esf_solid_set_mesh(esf_ctx* ctx, int solid, int nv, const double* rest,
                              const double* vertices, int nt, const int* tris);
/* vertex and triangle counts of a solid's mesh (0, 0 without one) */
esf_status esf_solid_mesh_size(esf_ctx* ctx, int solid, int* nv, int* nt);
/* current vertices, rest pose and triangles of a solid's mesh; any output
 * may be NULL.  ESF_ERR_STATE when the solid has no mesh. */
esf_status esf_solid_get_mesh(esf_ctx* ctx, int solid, double* vertices,
                              double* rest, int* tris);

/* --- host <-> device ----------------------------------------------------*/
esf_status esf_upload(esf_ctx* ctx, int field, int solid, const void* src,
                      size_t bytes);
esf_status esf_download(esf_ctx* ctx, int field, int solid, void* dst,
                        size_t bytes);

/* --- stages -------------------------------------------------------------*/
/* DEFOR_SOLID_3D::computeSolidMass(false) (DEFOR_SOLID_3D.cpp:576-669) */
esf_status esf_compute_solid_mass(esf_ctx* ctx, int solid);
/* DEFOR_SOLID_3D::computeSolidMass(true) (DEFOR_SOLID_3D.cpp:576-669): as
 * esf_compute_solid_mass, but a tight-box cell whose min corner has
 * _phi > 2 (the solid's phi from the last esf_solid_sdf) takes no samples
 * and stays invalid (:620). */
esf_status esf_compute_solid_mass_sdf(esf_ctx* ctx, int solid);
/* box part of DEFOR_SOLID_3D::computeSDF (DEFOR_SOLID_3D.cpp:216-266) */
esf_status esf_update_boxes(esf_ctx* ctx, int solid);
/* DEFOR_SOLID_3D::computeSDF (DEFOR_SOLID_3D.cpp:216-267) on the solid's
 * own mass, on the device: the tight and extended boxes, then _phi = -0.5
 * on massive nodes and the grid Eikonal distance (in cells) around them.
 * The reference sweeps the whole extended box; here phi is relaxed to the
 * same upwind fixed point ring by ring from the massive nodes, which
 * settles it where computeSolidMass(true) reads it (phi <= 2 against the
 * reference's ordered sweeps: the same set, values within 1e-4); further
 * out it is an upper bound.  For the swept field itself use
 * esf_compute_sdf. */
esf_status esf_solid_sdf(esf_ctx* ctx, int solid);
/* COUPLED_INTEGRATOR_3D::gatherSolidMasses (COUPLED_INTEGRATOR_3D.cpp:327-375) */
esf_status esf_gather_solid_masses(esf_ctx* ctx);
/* DEFOR_SOLID_3D::recoverVelocityField for every solid (:1235-1241) */
esf_status esf_recover_velocity(esf_ctx* ctx);
/* COUPLED_INTEGRATOR_3D::CFL (COUPLED_INTEGRATOR_3D.cpp:88-102); sets dt */
esf_status esf_cfl(esf_ctx* ctx, double alpha, double* dt);
/* stepImplicit :413-607: corotational stress and element stiffness
 * (DEFOR_SOLID_3D.cpp:697-871), computeMV (:873-889), vorticity and body
 * forces.  energy may be NULL. */
esf_status esf_forces(esf_ctx* ctx, double* energy);
/* stepImplicit :609-632 + pcrImplicitSolve (:662-764): system diagonal,
 * divergence RHS, solid solve, pressure solve, solid correction solve,
 * velocity write-back.  Returns ESF_NOT_CONVERGED if a solve stopped at
 * "cr max iterations". */
esf_status esf_coupled_solve(esf_ctx* ctx, esf_solve_stats* stats);
/* DEFOR_SOLID_3D::finalizeSolution(velocity, dt) (:1182-1197) */
esf_status esf_solid_copy_velocity(esf_ctx* ctx);
/* DEFOR_SOLID_3D::finalizeSolution() (:1198-1233) */
esf_status esf_solid_finalize(esf_ctx* ctx);
/* advectFluidVelocityField (COUPLED_INTEGRATOR_3D.cpp:797-826) */
esf_status esf_advect_velocity(esf_ctx* ctx);
/* advectField(_intensity) (COUPLED_INTEGRATOR_3D.cpp:766-796) */
esf_status esf_advect_intensity(esf_ctx* ctx);
/* DEFOR_SOLID_3D::advectSurfaceMesh (DEFOR_SOLID_3D.cpp:1125-1180) for every
 * solid with a mesh, one GPU thread per vertex: move with the solid
 * velocity for dt, then up to 10 Newton steps towards x - u(x) = rest,
 * clamped into the domain */
esf_status esf_advect_surface_mesh(esf_ctx* ctx);
/* The broad phase of stepImplicit (COUPLED_INTEGRATOR_3D.cpp:384-411): *alpha
 * = 0.6 when a solid is within 2 cells of a wall (DEFOR_SOLID_3D.h:85) or
 * two solids collide (DEFOR_SOLID_3D.h:84, TRIANGLE_BVH.cpp:179-192),
 * else 1.8.  With a mesh on every solid the test runs on the meshes' triangle
 * boxes (on the GPU; a BVH collision is exactly a pair of overlapping
 * triangle boxes); otherwise each solid's mass bounding box stands in. */
esf_status esf_broad_phase(esf_ctx* ctx, double* alpha);
/* DEFOR_SOLID_3D::computeSDF (DEFOR_SOLID_3D.cpp:216-267) on a node mass
 * field of the context's grid (host arrays of X*Y*Z doubles, x fastest):
 * phi = -0.5 where mass > 0, X + Y + Z + 2 elsewhere, then two rounds of
 * sweepPhi (:397-454, solveDistance :369-395) over the box of the massive
 * nodes widened by "extrapolation cells".  box and tight receive
 * {xmin, ymin, zmin, xmax, ymax, zmax} as _xmin.._zmax and _tightMin /
 * _tightMax ("tight border").  Bitwise the reference's sweeps, run on the
 * GPU plane by plane. */
esf_status esf_compute_sdf(esf_ctx* ctx, const double* mass, double* phi, int box[6],
                           int tight[6]);
/* esf_compute_sdf without allocating: no reference counterpart.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same results
 * as esf_compute_sdf, bitwise; its device staging (mass, phi) and extent
 * counters live in a static device array of the library, one per GPU, held
 * while the process lives: 34.3 MB, enough for grids of at most 129^3 =
 * 2,146,689 nodes (larger grids: ESF_ERR_INVALID).  Every call on a GPU
 * uses that GPU's one array, so calls that overlap on one GPU (from
 * several threads or contexts) race; calls in sequence are safe, each
 * starts its counters afresh. */
esf_status esf_compute_sdf_noalloc(esf_ctx* ctx, const double* mass, double* phi,
                                   int box[6], int tight[6]);
/* DEFOR_SOLID_3D::extendVectorField (DEFOR_SOLID_3D.cpp:566-574): four
 * rounds of sweepVectorField (:521-564; eight sweepOnePass :457-519, then
 * the box faces) on a node field of 3 doubles per node, in place, where
 * mass == 0; phi and box as computeSDF leaves them.  The box must lie in
 * the grid and span at least 3 nodes per axis.  Bitwise the reference's
 * sweeps, run on the GPU plane by plane. */
esf_status esf_extend_vector_field(esf_ctx* ctx, const int box[6], const double* mass,
                                   const double* phi, double* field);
/* --- FLIP material particles (one solid per scene) ---------------------*/
/* DEFOR_SOLID_3D::initMaterialParticles (DEFOR_SOLID_3D.cpp:333-367) on the
 * solid's valid cells (after esf_gather_solid_masses): "particles per cell"
 * particles per cell at offsets drawn from MT19937 seeded with 123456 (the
 * reference reseeds every call; the three draws of an offset taken x, y,
 * z), velocity interpolated from the solid's velocity field. */
esf_status esf_init_particles(esf_ctx* ctx, int solid);
/* DEFOR_SOLID_3D::gridToParticle (:891-1034): the FLIP/PIC blend over the
 * 4^3 cubic B-spline stencil of the solid's nodes with mass, the walls'
 * spring/friction impulse, the move by dt.  Contact between solids
 * (:969-995) reads other solids' rest-space SDFs and is not done here. */
esf_status esf_grid_to_particle(esf_ctx* ctx, int solid);
/* DEFOR_SOLID_3D::particleToGrid (:1036-1107): per node the weighted
 * velocity of the particles whose stencils cover it; where the weight
 * reaches "flip cutoff" the solid's velocity becomes that average and its
 * _mass 1, elsewhere _mass is 0.  Sums in a fixed per-node order (binned
 * particles), not the reference's particle order: equal to rounding. */
esf_status esf_particle_to_grid(esf_ctx* ctx, int solid);
/* DEFOR_SOLID_3D::extendVectorField(_velocity) (:566-574) on the solid's own
 * velocity, _mass and _phi, over its extended box, on the device */
esf_status esf_extend_solid_velocity(esf_ctx* ctx, int solid);
esf_status esf_num_particles(esf_ctx* ctx, int solid, long* n);
/* positions and velocities, 3 doubles per particle; either may be NULL */
esf_status esf_get_particles(esf_ctx* ctx, int solid, double* pos, double* vel);
esf_status esf_set_particles(esf_ctx* ctx, int solid, long n, const double* pos,
                             const double* vel);
/* recoverVelocityField + Dirichlet faces (COUPLED_INTEGRATOR_3D.cpp:894-916) */
esf_status esf_recover_and_bc(esf_ctx* ctx);
/* One substep of the drop-in: CFL(alpha) -> forces -> coupled solve ->
 * solid velocity copy -> solid finalize -> computeSolidMass(true) of every
 * solid -> surface meshes (solids that have one) -> computeSDF boxes and
 * phi (esf_solid_sdf) -> gather solid masses -> advect velocity -> advect
 * intensity -> recover + Dirichlet (COUPLED_INTEGRATOR_3D.cpp:377-917).
 * alpha < 0 takes the CFL alpha from the broad phase (esf_broad_phase), as
 * stepImplicit does.  With params.flip set (scenes of one solid) the step
 * is the reference's whole substep: after the velocity copy, gridToParticle,
 * particleToGrid and the velocity extension (:833-854), phi from the
 * ordered sweeps, and new particles after the mass gather; without it the
 * solid velocity is the grid velocity copied into its box.  A step
 * that runs alone in the process computes the fluid's vorticity and body
 * forces on a second stream of the context, beside the first solid solve;
 * results are the same either way.  It returns once the report, the boxes
 * and the solid counts have landed; the advection and recovery may still
 * run on the context's stream, ahead of anything queued on the context
 * later (reads through this interface wait for them; esf_synchronize waits
 * for all). */
esf_status esf_step(esf_ctx* ctx, double alpha, esf_solve_stats* stats);
/* esf_step of n contexts at once, each kernel of the step launched once
 * over the scenes that reach it together (the reference steps its one scene
 * per process; this batches independent scenes on one GPU).  stats[i] and
 * status[i] (either may be NULL) receive context i's report and
 * esf_step's return value; each context's results are bitwise those of
 * esf_step on it alone.  The work runs on ctxs[0]'s stream, ordered after
 * the work already queued on every context's stream and before any queued
 * on them later.  Contexts must be distinct, on the current device and in
 * no other batch; several threads may each run a batch of their own.
 * A batch of one runs esf_step on its context.  Returns the most severe
 * status. */
esf_status esf_step_batch(esf_ctx** ctxs, int n, double alpha, esf_solve_stats* stats,
                          esf_status* status);

/* --- operators (unit level) ---------------------------------------------*/
/* getMatVecMult (COUPLED_INTEGRATOR_3D.cpp:638-660): y = J M^-1 J^T x on
 * the current masses.  x, y are host arrays of (X-1)(Y-1)(Z-1) doubles. */
esf_status esf_pressure_apply(esf_ctx* ctx, const double* x, double* y);
/* _JMInvJTDiag (COUPLED_INTEGRATOR_3D.cpp:694-705) */
esf_status esf_pressure_diag(esf_ctx* ctx, double* d);
/* A x with A = (dt^2 + beta dt) K + diag(M~) (COUPLED_INTEGRATOR_3D.cpp:
 * 609-619), matrix-free, after esf_forces.  x, y: 3*nSolid doubles. */
esf_status esf_solid_apply(esf_ctx* ctx, const double* x, double* y);
/* solvePCR(b, x, getMatVecMult, app) on the pressure system
 * (CONJUGATE_RESIDUAL.inl:109-192); x in/out (warm start). */
esf_status esf_pressure_solve(esf_ctx* ctx, const double* b, double* x,
                              double eps, int max_it, int* iterations,
                              double* rel_residual);
/* Measurement entry point, no reference counterpart (the isolated-system
 * benchmark): solvePCR on S x = b with b (ESF_F_PRESSURE_RHS) and the start
 * x (ESF_F_LAMBDA) already resident, D from the last esf_pressure_diag /
 * esf_forces, in the context's precision (ESF_MIXED: FP32 CR inside FP64
 * refinement).  iterations: CR iterations (mixed: FP32 ones summed);
 * refinements: outer FP64 rounds (0 in FP64); device_ms: the solve's device
 * time between HIP events on the context stream.  The solution stays in
 * ESF_F_LAMBDA. */
esf_status esf_pressure_solve_resident(esf_ctx* ctx, double eps, int max_it,
                                       int* iterations, int* refinements,
                                       double* rel_residual, double* device_ms);
/* valid (stress-carrying) cells of a solid, after esf_gather_solid_masses */
esf_status esf_solid_num_cells(esf_ctx* ctx, int solid, int* n);
/* Measurement entry point, no reference counterpart (the isolated coupled
 * system): the stress stage takes F from here, one row-major 3x3 per valid
 * cell in the cells' linear order (n = esf_solid_num_cells), with
 * det F^-1 = 1 / det F, instead of forming it from the displacement
 * (DEFOR_SOLID_3D.cpp:758-760).  n = 0 returns to the displacement. */
esf_status esf_solid_set_deformation(esf_ctx* ctx, int solid, long n, const double* F);
/* Measurement entry point, no reference counterpart: the first solid solve
 * of pcrImplicitSolve (COUPLED_INTEGRATOR_3D.cpp:664-675), A x = b with
 * A = (dt^2 + beta dt) K + diag(M~) after esf_forces, b in ESF_F_SOLID_RHS
 * (resident), from x = 0 in FP64 (CONJUGATE_RESIDUAL.inl:194-276); x in
 * ESF_F_SOLID_V0; device_ms between HIP events on the context stream. */
esf_status esf_solid_solve_resident(esf_ctx* ctx, double eps, int max_it, int* iterations,
                                    double* rel_residual, double* device_ms);
/* COROTATION::init + firstPiolaKirchhoff + strainEnergy +
 * firstPKDifferential (COROTATION.cpp:37-107) on n deformation gradients,
 * one GPU thread each.  F, dF: n*9 row-major; out: n*36 (R,S,L,P);
 * energy: n; dP: n*9 (dF applied). */
esf_status esf_corotation_batch(int n, double lambda, double mu,
                                const double* F, const double* dF,
                                double* out, double* energy, double* dP);

/* --- timing ---------------------------------------------------------------*/
/* enable per-stage HIP-event timing (synchronises at every stage end) */
esf_status esf_set_stage_timing(esf_ctx* ctx, int on);
/* milliseconds spent per stage since the last reset, measured with HIP
 * events on the context's stream (when enabled): 0 cfl, 1 forces, 2 solid solves,
 * 3 pressure solve, 4 solid finalize, 5 gather, 6 advect velocity,
 * 7 advect intensity, 8 recover+bc */
esf_status esf_stage_times(esf_ctx* ctx, double* ms, int n, int reset);
/* measurement entry point, no reference counterpart: `reps` back-to-back
 * launches of the pressure SpMV (FP64, or FP32 on a mixed context) on the
 * current masses, bracketed by HIP events on the context stream; *avg_ms is
 * the average launch duration */
esf_status esf_bench_pressure_apply(esf_ctx* ctx, int fp32, int reps,
                                    double* avg_ms);
/* measurement entry point, no reference counterpart: a streaming copy of
 * `bytes` between two device buffers, `reps` launches timed with HIP events
 * (reads + writes = 2 * bytes per launch) */
esf_status esf_bench_copy(esf_ctx* ctx, long bytes, int reps, double* avg_ms);
/* Record HIP events around every Krylov kernel launch (no synchronisation;
 * events are read back by esf_kernel_timing). */
esf_status esf_set_kernel_timing(esf_ctx* ctx, int on);
/* Accumulated device time (ms) and launch count per kernel class since the
 * last reset: 0-5 pressure SpMV FP64 (apply, setup-r, setup-s, CR update,
 * residual check, refinement residual), 6-11 the same in FP32, 12/13 unused,
 * 14 the solid CR cell pass (K_e r per cell), 15 the solid CR node pass
 * (gather + update), 16/17 the same two passes when they also carry the
 * true-residual check of the previous iteration.  Arrays of
 * ESF_KERNEL_CLASSES. */
#define ESF_KERNEL_CLASSES 18
esf_status esf_kernel_timing(esf_ctx* ctx, double* total_ms, long* counts,
                             int reset);

/* --- host scene setup (cfg-driven, rubberBall/main.cpp:63-160) ------------*/
typedef struct {
  int n_solids;
  int frame_start, frame_end;
  double frame_dt;
  double carpet_thickness;
} esf_scene_info;
/* Read a reference cfg (SIMPLE_PARSER semantics, src/util/SIMPLE_PARSER.cpp:
 * 64-236) into the grid and parameters the context would get, with the
 * reference's defaults for absent keys.  Host only: needs no GPU. */
esf_status esf_cfg_read(const char* cfg_path, esf_grid* grid,
                        esf_params* params, esf_scene_info* info);
/* Parse a reference cfg (SIMPLE_PARSER semantics), create the context,
 * add the solids, run COUPLED_INTEGRATOR_3D::init and lay the smoke carpet.
 * A solid's rest mass is read from "<solid path i><solid name i>.mass"
 * (FIELD_3D::readGz format) when that file exists, otherwise rasterized
 * from the analytic sphere "solid sphere radius i" in rest space. */
esf_status esf_scene_load_cfg(const char* cfg_path, int precision,
                              esf_ctx** out, esf_scene_info* info);
/* APPLICATION::step (rubberBall/main.cpp:174-205): substeps until the
 * frame's dt is covered.  alpha < 0 chooses the reference's broad-phase
 * rule from the solid mass bounding box. */
esf_status esf_frame(esf_ctx* ctx, double frame_dt, double alpha,
                     int* substeps);

/* --- frame files (host; the reference's gz field format) -------------------*/
/* FIELD_3D::writeGz / VEC3F_FIELD_3D::writeGz (FIELD_3D.h:286-321,
 * VEC3F_FIELD_3D.h:199-229): gzip level 1, header int32 X, Y, Z; double dh,
 * origin[3]; then components * X * Y * Z doubles, x fastest, vector
 * components interleaved.  components: 1 or 3.  Host only: needs no GPU. */
esf_status esf_field_write_gz(const char* path, const esf_grid* grid,
                              int components, const double* data);
/* FIELD_3D::readGz / VEC3F_FIELD_3D::readGz (FIELD_3D.h:323-365,
 * VEC3F_FIELD_3D.h:231-267): the header into *grid, then, when data is not
 * NULL, the payload (ESF_ERR_INVALID if it exceeds `capacity` doubles).
 * ESF_ERR_IO for a missing or short file (the reference exits there).
 * Host only. */
esf_status esf_field_read_gz(const char* path, esf_grid* grid, int components,
                             double* data, size_t capacity);
/* OBJ::Load (src/geometry/OBJ.cpp:97-295) as the solver uses it: "v" lines,
 * and one triangle per "f" line from its first three vertex indices.  Counts
 * into *nv, *nt; vertices (3 per vertex) and tris (3 zero-based indices per
 * triangle) are filled when not NULL and cap_v >= nv, cap_t >= nt.
 * ESF_ERR_IO for a missing or malformed file.  Host only. */
esf_status esf_obj_read(const char* path, int* nv, int* nt, double* vertices,
                        int cap_v, int* tris, int cap_t);
/* OBJ::Save (OBJ.cpp:336-375): "v x y z" lines (17 significant digits
 * where the reference streams 6) and "f a b c" lines.  Host only. */
esf_status esf_obj_write(const char* path, int nv, const double* vertices,
                         int nt, const int* tris);
/* COUPLED_INTEGRATOR_3D::writeFrame (COUPLED_INTEGRATOR_3D.cpp:999-1011) with
 * DEFOR_SOLID_3D::writeFrame (DEFOR_SOLID_3D.cpp:1361-1371): under
 * "<data_path>frame.NNNN" (at least four digits, IO.h:27-37) the intensity
 * every frame; the grid velocity and, per solid i, ".solid.i.displacement",
 * ".solid.i.velocity" and ".solid.i.phi" when debug or frame % 10 == 0, and
 * every frame ".solid.i.obj" for a solid with a surface mesh.  The ".pbrt"
 * render file is not written. */
esf_status esf_write_frame(esf_ctx* ctx, const char* data_path, int frame,
                           int debug);
/* COUPLED_INTEGRATOR_3D::readFrame (COUPLED_INTEGRATOR_3D.cpp:976-997) with
 * DEFOR_SOLID_3D::readFrame (DEFOR_SOLID_3D.cpp:1299-1317): the same files
 * back (a grid that differs from the context's, or a mesh of other
 * topology, is ESF_ERR_INVALID), each solid's velocity workspace set from
 * its velocity and X = x - u, mesh vertices from ".solid.i.obj", the boxes
 * from the mesh's bounding box (:1318-1355; a solid without a mesh keeps
 * its boxes) and computeSolidMass(true) (:1357), then gatherSolidMasses and
 * recoverVelocityField. */
esf_status esf_read_frame(esf_ctx* ctx, const char* data_path, int frame,
                          int debug);

#ifdef __cplusplus
}
#endif
#endif /* ESF_H */
