/*
 * esf_oracle.h -- C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++17 (+OpenMP)
 * restatement of the reference's coupled solid-fluid substep
 * (src/integrator/COUPLED_INTEGRATOR_3D.cpp, src/integrator/DEFOR_SOLID_3D.cpp,
 * src/material/COROTATION.cpp, include/linearalgebra/CONJUGATE_RESIDUAL.inl).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the reported CPU baseline.  The
 * product (eulersf_amd/, libesf.so) never links or calls it.
 *
 * Array layouts follow the reference exactly:
 *   node n = x + y*X + z*X*Y, vector fields AoS (3 doubles per node),
 *   cell r = x + y*(X-1) + z*(X-1)*(Y-1), quarter fractions [8*n + corner].
 */
#ifndef ESF_ORACLE_H
#define ESF_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double dt;              /* "dt": frame step and default max dt            */
  double vorticity_eps;   /* "vorticity eps" as written in the cfg          */
  double gravity;         /* "gravity"                                      */
  double buoyancy;        /* "buoyancy"                                     */
  double rayleigh_alpha;  /* "rayleigh alpha"                               */
  double rayleigh_beta;   /* "rayleigh beta"                                */
  double fluid_density;   /* "fluid density"                                */
  double implicit_eps;    /* "implicit solver eps" (solid solves)           */
  double divfree_eps;     /* "div free solver eps" (pressure solve)         */
  int cr_max_iterations;  /* "cr max iterations"                            */
  double mass_sample_dh;  /* "mass sample dh"                               */
  int tight_border;       /* "tight border"                                 */
  int extrapolation_cells;/* "extrapolation cells"                          */
  /* the FLIP particle round trip of finalizeSolution (DEFOR_SOLID_3D.cpp:
   * 333-367, 891-1107), run by orc_step when flip != 0 */
  int particles_per_cell; /* "number of particles per material cell" (8)    */
  double flip_alpha;      /* "flip alpha" (0.95)                            */
  int flip_cutoff;        /* "flip cutoff" (3)                              */
  double wall_thickness;  /* "wall thickness" (1.5 dh)                      */
  double spring_stiffness;/* "contact spring stiffness" (1e4)               */
  double friction;        /* "friction" (0.5)                               */
  int flip;
} orc_params;

/* field ids for orc_field() */
enum {
  ORC_VELOCITY = 0,       /* 3*nodes */
  ORC_INTENSITY = 1,      /* nodes */
  ORC_SOLID_MASS = 2,     /* nodes, sum over solids */
  ORC_NODE_FRACTION = 3,  /* nodes */
  ORC_LAMBDA = 4,         /* cells (warm-started pressure) */
  ORC_VEL_WORKSPACE = 5,  /* 3*nodes */
  ORC_SOLID_INDEX = 6,    /* int32 nodes (gridIdxToSolidIdx) */
  ORC_V0 = 7,             /* 3*interior nodes, solution order */
  ORC_V1 = 8,             /* 3*interior nodes */
  ORC_MINV = 9,           /* 3*interior nodes (nzMassVecInv) */
  ORC_PRESSURE_RHS = 10,  /* cells: b = J v0 - c */
  ORC_DIV_RHS = 11,       /* cells: c */
  ORC_JT_LAMBDA = 12,     /* 3*interior nodes: -J^T lambda */
  ORC_SOLID_FORCE = 13,   /* 3*nSolid (_solidUnitForce before *dt) */
  ORC_SOLID_MV = 14,      /* 3*nSolid */
  ORC_DAMPED_MASS = 15,   /* 3*nSolid */
  ORC_UNIT_FORCE = 16,    /* 3*interior (_unitForce before *dt) */
  ORC_MV = 17,            /* 3*interior */
  ORC_SOLID_V0 = 18,      /* 3*nSolid */
  ORC_SOLID_DV = 19,      /* 3*nSolid */
  ORC_SYS_DIAG = 20,      /* 3*nSolid: diag(A) */
  ORC_PRESSURE_DIAG = 21, /* cells */
  ORC_VORTICITY = 22,     /* 3*nodes */
  ORC_WORKSPACE = 23,     /* nodes (|omega| after body forces) */
  /* per-solid fields, select the solid with the `solid` argument */
  ORC_S_U = 100,          /* 3*nodes displacement */
  ORC_S_X = 101,          /* 3*nodes material coordinates */
  ORC_S_VELOCITY = 102,   /* 3*nodes */
  ORC_S_MASS = 103,       /* nodes */
  ORC_S_VALID = 104,      /* int8 cells (X-1)(Y-1)(Z-1) */
  ORC_S_QFRAC = 105,      /* 8*nodes */
  ORC_S_VEL_WORKSPACE = 106, /* 3*nodes */
  ORC_S_DETFINV = 107,    /* nodes */
  ORC_S_PHI = 108         /* nodes */
};

typedef struct {
  int iterations[3];      /* solid 1, pressure, solid 2 */
  int converged[3];
  double rel_residual[3]; /* |Ax-b|/|b| at exit (true residual) */
  double dt;
  double energy;
} orc_stats;

void* orc_create(int X, int Y, int Z, double dh, const double origin[3],
                 const orc_params* p);
void orc_destroy(void* h);
/* DEFOR_SOLID_3D ctor + setMaterial + setMassDensity + setInitialState +
 * setInitialVelocity (rubberBall/main.cpp:99-160). rot is row-major. */
int orc_add_solid(void* h, double lambda, double mu, double density,
                  int rX, int rY, int rZ, double rdh, const double rorigin[3],
                  const double* rest_mass, const double rot[9],
                  const double trans[3], const double scale[3],
                  const double v0[3]);
/* COUPLED_INTEGRATOR_3D::init + smoke carpet (main.cpp:78-90) */
void orc_init(void* h, double carpet_thickness);
/* raw pointer into oracle storage; *n receives the element count */
void* orc_field(void* h, int id, int solid, size_t* n);
int orc_num_solid_dofs(void* h);
void orc_get_boxes(void* h, int solid, int out[12]); /* tightMin,tightMax,min,max */
void orc_set_boxes(void* h, int solid, const int in[12]);

/* stages of stepImplicit (COUPLED_INTEGRATOR_3D.cpp:377-636) */
double orc_cfl(void* h, double alpha);           /* sets and returns _dt */
void orc_set_dt(void* h, double dt);
void orc_set_max_dt(void* h, double dt);         /* setMaxDt */
double orc_forces(void* h);        /* :413-607 stress, computeMV, body forces */
void orc_assemble(void* h);        /* :609-629 A and c */
int orc_solve(void* h, orc_stats* st);           /* :662-764 */
void orc_solid_copy_velocity(void* h);           /* DEFOR :1182-1197 */
void orc_solid_finalize(void* h);                /* DEFOR :1198-1233 */
void orc_compute_solid_mass(void* h, int solid); /* DEFOR :576-669 (usesdf=false) */
void orc_update_boxes(void* h, int solid);       /* DEFOR :216-266 (box part) */
void orc_compute_solid_mass_sdf(void* h, int solid); /* DEFOR :576-669 (usesdf=true) */
/* FLIP particles of a solid (DEFOR_SOLID_3D.cpp:333-367, 891-1107):
 * initMaterialParticles, gridToParticle, particleToGrid; the count, and
 * the particles' positions and velocities (3 doubles each) */
void orc_init_particles(void* h, int solid);
void orc_grid_to_particle(void* h, int solid);
void orc_particle_to_grid(void* h, int solid);
long orc_num_particles(void* h, int solid);
void orc_get_particles(void* h, int solid, double* pos, double* vel);
void orc_set_particles(void* h, int solid, long n, const double* pos, const double* vel);
/* the isolated system: valid cells of a solid, a deformation gradient per
 * valid cell for the stress stage, the solid CR from x = 0 on the
 * assembled A */
long orc_solid_num_cells(void* h, int solid);
void orc_solid_set_deformation(void* h, int solid, long n, const double* F);
int orc_solid_cr(void* h, const double* b, double* x, double eps, int max_it, int* iters);
/* DEFOR :566-574 on the solid's own velocity, mass and phi */
void orc_extend_solid_velocity(void* h, int solid);
/* DEFOR :671-682: F of the cell with min-corner node `index` (row-major), returns det F */
double orc_compute_F(void* h, int solid, long index, double* F);
void orc_solid_sdf(void* h, int solid);          /* DEFOR :216-267 (boxes + phi) */
void orc_gather(void* h);                        /* :327-375 */
void orc_advect_velocity(void* h);               /* :797-826 */
void orc_advect_intensity(void* h);              /* :766-796 */
void orc_recover_and_bc(void* h);                /* :894-916 */
/* the drop-in step: CFL -> forces -> assemble -> solve -> solid velocity
 * copy -> solid finalize -> gather -> advect velocity -> advect intensity
 * -> recover + Dirichlet.  alpha < 0 picks 1.8 (no contact). */
int orc_step(void* h, double alpha, orc_stats* st);

/* operators on the current state (for unit tests) */
void orc_pressure_apply(void* h, const double* x, double* y);   /* :638-660 */
void orc_pressure_diag(void* h, double* d);                     /* :694-705 */
void orc_solid_apply(void* h, const double* x, double* y);      /* A x (CSR) */
/* CR on the pressure system from x (warm start), reference cadence */
int orc_pressure_cr(void* h, const double* b, double* x, double eps,
                    int max_it, int* iters);

/* J v (v in solution order, 3*interior) -> cells, and -J^T lam -> 3*interior */
void orc_apply_J(void* h, const double* v, double* out);
void orc_apply_negJT(void* h, const double* lam, double* out);
/* the oracle's CR on a dense operator with an explicit Jacobi diagonal;
 * x is the start vector (warm start) */
int orc_pcr_dense(int n, const double* A, const double* diag, const double* b,
                  double* x, double eps, int max_it, int* iters);

/* corotational material on one F (COROTATION.cpp).  out: R,S,L,P (36),
 * energy; dP applied to dF written to dP_out (9).  Row-major 3x3. */
void orc_corotation(double lambda, double mu, const double F[9],
                    const double dF[9], double out[36], double* energy,
                    double dP_out[9]);
/* FIELD_3D::interpolate_value over an arbitrary scalar / AoS field */
void orc_advect_surface_mesh(int X, int Y, int Z, double dh, const double o[3],
                             const double* vel, const double* u, double dt, int nv,
                             const double* rest, double* vert);
double orc_mesh_broad_phase(int n, const int* nt, const long* vo, const long* to,
                            const double* v, const int* tri, int X, int Y, int Z, double dh,
                            const double o[3]);
double orc_interp_scalar(const double* f, int X, int Y, int Z, double dh,
                         const double o[3], const double p[3]);
void orc_interp_vec3(const double* f, int X, int Y, int Z, double dh,
                     const double o[3], const double p[3], double out[3]);
double orc_last_stage_seconds(void* h, int stage);
int orc_num_threads(void);
/* rest-space mass field of a sphere of `radius` centred at the origin,
 * rasterized as projects/rasterizeSolid/main.cpp:42-72 does with the SDF
 * |p| - radius: dims and origin first, then the data (X*Y*Z doubles) */
void orc_sphere_rest_dims(double radius, double mass_dh, int dims[3], double origin[3]);
void orc_sphere_rest_fill(double radius, double mass_dh, double density, double* out);
/* ---- fast sweeping (TEST INFRASTRUCTURE) ----------------------------------
 * computeSDF (DEFOR_SOLID_3D.cpp:216-267) with sweepPhi / solveDistance
 * (:369-454), and extendVectorField (:566-574) with sweepVectorField
 * (:521-564) and sweepOnePass (:457-519), restated on plain node arrays of
 * X*Y*Z (x fastest).  box = {xmin, ymin, zmin, xmax, ymax, zmax} as the
 * reference's _xmin.._zmax (the max ends exclusive in its loops), tight the
 * same for _tightMin/_tightMax. */
void orc_compute_sdf(int X, int Y, int Z, const double* mass, int tight_border, int extrap,
                     double* phi, int box[6], int tight[6]);
void orc_sweep_phi(int X, int Y, int Z, const int box[6], const double* mass, double* phi);
void orc_extend_vector_field(int X, int Y, int Z, const int box[6], const double* mass,
                             const double* phi, double* field);
double orc_solve_distance(double p, double q, double r, double v);

#ifdef __cplusplus
}
#endif
#endif
