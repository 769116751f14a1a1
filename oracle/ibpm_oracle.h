/*
 * FP64 CPU oracle for the IBPM time step of ImmersedBodies.jl.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP library is
 * compared against (parity tests, the smoke check and the CPU-baseline leg of the
 * benchmark). No product code path links or calls it.
 *
 * It restates, in plain C, the algorithm of
 *   src/fluid.jl      (all but the unused 3-argument curl! at :69-99)
 *   src/coupling.jl   (all)
 *   src/solver.jl:78-190 (CnabSolver construction and advance!)
 *   src/bodies.jl:1-51
 *   src/problem.jl:363-429 (grid layout), :566-596 (CNAB), :626-708 (State, time index),
 *                  :676-685 (update_vars!), :853-856 (total force)
 * Arrays use the reference layout: column-major float64, one column per level.
 *   q[:, lev]  = [u ((nx+1) x ny) ; v (nx x (ny+1))]
 *   G[:, lev]  = circulation on the (nx-1) x (ny-1) interior vertices
 *   bc vector  = left (ny+1), right (ny+1), bottom (nx+1), top (nx+1)
 * The DST-I (FFTW RODFT00 in the reference) is computed with an FFT written here.
 */
#ifndef IBPM_ORACLE_H
#define IBPM_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ibo_grid {
  int nx, ny, nlevel;
  double x0, y0, h;
} ibo_grid;

/* ---- operators on one level (reference file:line in each definition) ---- */
void ibo_dst2d(const ibo_grid* g, double* x);                 /* in place, unnormalised RODFT00 */
void ibo_dst1d_batch(int n, int count, double* x);           /* count contiguous vectors */
void ibo_lap_eigs(const ibo_grid* g, double* lam);
void ibo_lap_inv(const ibo_grid* g, const double* diag, double* x, const double* b);
void ibo_diag_poisson(const ibo_grid* g, double* d);         /* scale / lambda */
void ibo_diag_A(const ibo_grid* g, double Re, double dt, int lev, int implicit, double* d);
void ibo_rot(const ibo_grid* g, double* G, const double* q);
void ibo_curl(const ibo_grid* g, double* q, const double* psi, const double* bc);
void ibo_coarsify(const ibo_grid* g, double* Gc, const double* G);
void ibo_get_bc(const ibo_grid* g, double* bc, const double* r);
void ibo_apply_bc(const ibo_grid* g, double* r, const double* bc, double fac);
void ibo_nonlinear(const ibo_grid* g, double* nonlin, const double* q, const double* q0,
                   const double* G, const double* Gbc, int lev);
void ibo_vort2flux(const ibo_grid* g, double* psi, double* q, double* G);
double ibo_delta_h(double rf, double rb, double dr);
int ibo_reg_update(const ibo_grid* g, int npanel, const double* xb, int* coords, double* w);
void ibo_regT(const ibo_grid* g, int npanel, const int* coords, const double* w,
              double* fb, const double* q1);
void ibo_reg(const ibo_grid* g, int npanel, const int* coords, const double* w,
             double* q1, const double* fb);

/* ---- the solver (CnabSolver) ---- */
typedef struct ibo_solver ibo_solver;

typedef struct ibo_config {
  ibo_grid grid;
  double Re, dt, t0;
  int npanel;
  const double* xb;       /* npanel x 2 column-major: reference panel positions */
  const double* ds;       /* npanel */
  int nbody;
  const int* body_start;  /* first panel of each body */
  const int* body_count;
  const int* body_moving; /* 0 static, 1 MovingRigidBody */
  int cg_maxiter;         /* reference default 5000 */
  double cg_reltol;       /* reference default 1e-12 */
  double cg_abstol;       /* reference default 0 */
  int cg_tol_mode;        /* 0: reltol*||r0|| (IterativeSolvers 0.9); 1: reltol*||b|| */
  int timing_only;        /* 1: skip the dense B^-1 assembly and use the identity (same
                             per-step cost; results meaningless) for CPU step timing */
  int grid_moving;        /* 0 StaticGrid, 1 MovingGrid: ibo_set_grid_motion before each
                             ibo_advance gives the GridVelocity at the new time */
} ibo_config;

ibo_solver* ibo_create(const ibo_config* c, char* err, size_t errlen);
void ibo_destroy(ibo_solver* s);
/* motion: per body (pos x, pos y, vel x, vel y, angle, angular_vel) at the new time */
int ibo_advance(ibo_solver* s, const double fs_vel[2], const double* motion, char* err,
                size_t errlen);
/* GridVelocity of a MovingGrid at the next step's time: centre x, y, velocity x, y, angle,
   angular_vel (problem.jl:476-490) */
void ibo_set_grid_motion(ibo_solver* s, const double gm[6]);
/* field access; names as in ibpm.h: 0 q,1 q0,2 Gamma,3 psi,4 nonlin1,5 nonlin2,6 xb,7 ub,
   8 Ftilde,9 fb,10 total force, 11 ds */
double* ibo_field(ibo_solver* s, int which, size_t* n);
int ibo_index(const ibo_solver* s);
double ibo_time(const ibo_solver* s);
void ibo_set_index(ibo_solver* s, int index);
int ibo_last_cg_iters(const ibo_solver* s);
double ibo_last_cg_residual(const ibo_solver* s);
/* apply the B operator (coupling.jl:150-177) of the current state */
void ibo_apply_B(ibo_solver* s, double* x, const double* z);
/* dense B^-1 of a static problem (npanel2 x npanel2), NULL when bodies move */
const double* ibo_Binv(const ibo_solver* s);
void ibo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
