// State of one solver (ibpm_solver of include/ibpm.h), shared by the translation units
// that drive it (ibpm.hip: the step; batch.hip: many solvers captured in one graph).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/ibpm.h"
#include "ibpm_internal.h"

constexpr int kForceChunk = 1024;  // steps of total force held in pinned memory (h_force)

struct ibpm_solver {
  ibpm::GridDims g{};
  int L = 0;
  double x0 = 0, y0 = 0, h = 0, Re = 0, dt = 0, t0 = 0;
  int index = 0;
  int n = 0, nbody = 0, moving = 0;
  int precision = IBPM_FP64, refine = 1;
  int dst_tiles = -1;  // sine-pass tile variant, -1: autotuned
  int cg_maxiter = 5000, cg_tol_mode = 0;
  double cg_reltol = 1e-12, cg_abstol = 0.0;
  // preconditioned CG: inverse of B at an earlier configuration (pc_P), the bodies'
  // angles then (pc_theta0, device), scratch vectors and GEMV partials
  int cg_precond = 0, cg_precond_refresh = 0;
  int helmholtz = 0;  // 0: Gauss-Seidel where it converges fast, else sine transforms; 1: sine
  bool pc_ready = false, pc_due = false;
  double *pc_P = nullptr, *pc_theta0 = nullptr, *pc_t = nullptr, *pc_z = nullptr;
  std::vector<double> cur_theta;  // host: each body's angle at the current step
  int device = 0;
  hipStream_t st = nullptr;
  double fs[2] = {0.0, 0.0};
  int grid_moving = 0;  // MovingGrid: gm = centre x, y, velocity x, y, angle, angular_vel
  double gm[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  std::vector<double> hc;  // grid step per level
  double scale = 0.0;      // 1 / (4 nx ny): the two unnormalised DST-I scale factors

  // fields
  double *q = nullptr, *G = nullptr, *psi = nullptr, *nl1 = nullptr, *nl2 = nullptr;
  double *Gs = nullptr, *rhs = nullptr;
  double *T = nullptr, *U = nullptr, *R = nullptr;  // kMaxPlanes scratch planes each
  double *cosx = nullptr, *cosy = nullptr;  // cos(pi k / n): spectral eigenvalues
  void *SoX = nullptr, *SeX = nullptr, *SoY = nullptr, *SeY = nullptr;  // FP64 tables
  void *SqX = nullptr, *SqY = nullptr;  // FP64 double-fold tables
  void *BqX = nullptr, *BqY = nullptr;  // split bf16 double-fold tables (mixed, N + 1 = 4Q)
  double *F1 = nullptr, *G1 = nullptr;  // FP64: level-1 spectral planes of the step
  // box update of the final vort2flux (FP64, static bodies): the change of Gamma after the
  // trial vort2flux lies in box [bi0, bi0 + bw) x [bj0, bj0 + bh) of each level (0-based
  // plane indices; bw = 0: none), and its forward transform is added to the trial's
  int box_opt = 0;  // 0: when it pays, 1: off, 2: wherever valid
  bool box_on = false;
  int bi0[16] = {}, bw[16] = {}, bj0[16] = {}, bh[16] = {}, bxfirst[16] = {};
  // K-slices of a level's first box product (1: one product) and their partials' offsets
  int bns[16] = {};
  size_t bpoff[16] = {};
  size_t btoff[16] = {};
  double *Sfx = nullptr, *Sfy = nullptr;  // full sine matrices 2 sin(pi k i / n)
  double *Ftr = nullptr;                   // L spectral planes of the trial vort2flux
  double *Tbox = nullptr;                  // per-level intermediate products
  double* Tpart = nullptr;                 // K-slice partials of the long ones
  // trial psi on level 1 only where couple_surface's E reads it (a box around the panels
  // clear of the level's edge): S_x[I, :] E_1 S_y[:, J] by K-split products
  bool pbox_on = false;
  int pi0 = 0, pbw = 0, pj0 = 0, pbh = 0;
  double *Pa = nullptr, *Pb = nullptr, *T2 = nullptr;
  // moving bodies: the boxes follow the panels, recomputed every step from the range of
  // their cells (bbox_d on the device); buffer capacities grow as needed
  bool box_dyn = false;
  double* bbox_d = nullptr;
  size_t tbox_cap = 0, pa_cap = 0, pb_cap = 0, t2_cap = 0, tpart_cap = 0;
  int gtmp_box[4] = {-1, -1, -1, -1};  // level-1 box B last wrote into Gtmp (-1: none)
  int gtmp_cols = 0;
  void *BfX = nullptr, *BfY = nullptr;                                  // bf16 tables
  // multi-level ring chain: sine rows/columns, per-plane factors, per-chain scratch
  double *edge = nullptr, *pw = nullptr;
  double *fac = nullptr, *rvec = nullptr, *rpart = nullptr;
  double *Mx = nullptr, *My = nullptr;  // ring operators (My aliases Mx when nx == ny)
  // bodies
  double *xb = nullptr, *ub = nullptr, *xb_ref = nullptr, *Ft = nullptr, *fb = nullptr;
  double *Ftmp = nullptr, *total = nullptr, *motion = nullptr;
  int *body_of = nullptr, *moving_d = nullptr, *coords = nullptr, *perm = nullptr;
  // static panels: each vertex's rot(E^T f) terms (ScatterArgs rowptr, ep, ew)
  int *sc_rowptr = nullptr, *sc_ep = nullptr;
  double* sc_ew = nullptr;
  int *xstart = nullptr, *err = nullptr, *ywin = nullptr;
  double* gemv_part = nullptr;
  double* w = nullptr;
  double* Binv = nullptr;  // static bodies: B^-1, row-major once assembled
  // B operator scratch (ncol columns)
  int bcols = 0;
  double *Gtmp = nullptr, *psitmp = nullptr, *Sc = nullptr;
  // CG
  double *cg_r = nullptr, *cg_u = nullptr, *cg_c = nullptr, *cg_scal = nullptr;
  double* h_scal = nullptr;  // pinned: [0..5] CG scalars, [8..10] warm-start scalars
  double* h_force = nullptr;  // mapped pinned: total force of the last kForceChunk steps
  double* d_force = nullptr;  // h_force's device address
  double* force_slot = nullptr;  // where this step's forces kernel also writes the total (or null)
  // a step captured into a batch's graph: no host reads or decisions; the CG runs
  // cg_maxiter iterations whose launches exit once it has converged, the learned guess
  // is kept or dropped on the device, and the step's freestream / grid motion come from
  // dyn_d (14 values, BaseFlux::dyn) and its body motion from `motion`, both filled on the
  // device before the step
  bool graph_mode = false;
  // q is curl(psi) of the current psi but not written: steps outside a graph leave the
  // final curl to the next step's nonlinear kernel, which forms the fluxes from psi;
  // reads and writes of the state write q first (materialize_q)
  bool q_stale = false;
  double* dyn_d = nullptr;
  int graph_cg_budget = 0;  // CG iterations captured per step (0: cg_maxiter)
  // instrumentation (ibpm_time_kernel): HIP events around each launch of kernel class
  // `probe` (0: sine passes over all levels, 1: the nonlinear kernel) inside real steps
  int probe = -1;
  std::vector<hipEvent_t> probe_ev;  // pairs
  int probe_n = 0;
  // CG warm start: extrapolation history F~b of the two steps before the previous one,
  // the extrapolated guess x_e, B x_e, and (rms(r_e), ||r_e||, ||b||)
  int ws_mode = IBPM_WS_PREVIOUS, ws_record = 0, nhist = 0;
  ibpm_net* net = nullptr;
  double *ws_h0 = nullptr, *ws_h1 = nullptr, *ws_xe = nullptr, *ws_Bxe = nullptr;
  double* ws_scal = nullptr;
  // host copies
  std::vector<double> ds;
  std::vector<double> q0_shadow;
  bool q0_shadow_valid = false;
  std::vector<int> body_start, body_count, body_moving;
  std::vector<double> staged_motion;  // ibpm_set_body_motion: 6 per body, and its time
  std::vector<double> staged_t;
  std::vector<char> staged_set;

  size_t plane() const { return (size_t)g.nG; }
};

namespace ibpm {
// base_flux! of the solver's current freestream and grid motion (fluid.jl:22-67).
BaseFlux base_flux(const ibpm_solver* s);
// One advance! of a solver (solver.jl:155-190) queued on s->st: no host synchronisation
// for static bodies. fsv: freestream at the new time; motion: 6 x nbody or null.
// phase 0: the whole step; 1: up to and including couple_surface (with a captured step's
// CG budget); 2: the rest (forces, project_circ, final vort2flux).
int solver_step(ibpm_solver* s, const double* fsv, const double* motion, ibpm_step_stats* stats,
                int phase = 0);
// Host-driven CG iterations from where a captured step's budget left them, to the stopping
// test, on the solver's stream.
int solver_cg_continue(ibpm_solver* s);
// Writes q = curl(psi) where the last step left it to the next nonlinear kernel (q_stale),
// on the solver's stream.
int materialize_q(ibpm_solver* s);
// BaseFlux::dyn values of a step from its freestream and GridVelocity fs8 = (ux, uy,
// centre x, y, velocity x, y, angle, angular_vel): the numbers base_flux() puts in the
// kernel arguments, bit for bit.
void base_flux_dyn(const ibpm_solver* s, const double fs8[8], double out[14]);
}  // namespace ibpm
