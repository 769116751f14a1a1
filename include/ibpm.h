/*
 * libibpm: MI355X (gfx950) implementation of the IBPM time step of ImmersedBodies.jl.
 *
 * C ABI: plain pointers and sizes, column-major float64 arrays in the reference's layout
 * (problem.jl:363-389, :626-665). The library owns device memory; callers own every host
 * buffer they pass. A solver is not thread-safe (the reference's CnabSolver is not
 * either: its closures share temporaries, solver.jl:343-358).
 *
 * Every entry point returns IBPM_OK (0) or an error code; ibpm_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Reference interface replaced (paths relative to the reference repository):
 *   ibpm_create       <- create_solver(state; fftw, cg) / CnabSolver(state), solver.jl:84-152
 *                        (includes _init_body_motion!, bodies.jl:33-41, and Binv_operator,
 *                        coupling.jl:179-202)
 *   ibpm_advance      <- advance!(solver::CnabSolver), solver.jl:155-190, nsteps times;
 *   ibpm_set_body_motion <- _update_body_motion!'s motion(t) (bodies.jl:43-51), staged;
 *                        freestream_vel(t) and body motion(t) are evaluated by the caller
 *                        (problem.jl:546, :259) and passed in per step
 *   ibpm_read/_write  <- the State fields read by getstate/savers and written by
 *                        load_state! (problem.jl:626-665, io.jl:42-71)
 *   ibpm_get_index/_set_index <- state.index / set_time_index! (problem.jl:703-708)
 *   ibpm_last_error   <- the exceptions of coupling.jl:40 ("outside allowed region")
 *   ibpm_destroy      <- garbage collection of the solver
 */
#ifndef IBPM_H
#define IBPM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ibpm_solver ibpm_solver;

enum ibpm_status {
  IBPM_OK = 0,
  IBPM_E_ARG = 1,            /* invalid argument or size */
  IBPM_E_OUTSIDE_REGION = 2, /* a panel too close to the level-1 edge (coupling.jl:39-41) */
  IBPM_E_HIP = 3,            /* HIP runtime failure */
  IBPM_E_SINGULAR = 4,       /* dense B is singular */
  IBPM_E_NO_DEVICE = 5       /* no gfx950 device */
};

enum ibpm_precision {
  IBPM_FP64 = 0,  /* sine contractions on FP64 MFMA */
  IBPM_MIXED = 1  /* 3xBF16 split products, FP32 accumulation, FP64 refinement */
};

enum ibpm_field {
  IBPM_Q = 0,           /* nq x nlevel   velocity flux            state.q  */
  IBPM_Q0 = 1,          /* nq x nlevel   base flux                state.q0 */
  IBPM_GAMMA = 2,       /* nG x nlevel   circulation              state.Γ  */
  IBPM_PSI = 3,         /* nG x nlevel   streamfunction           state.ψ  */
  IBPM_NONLIN1 = 4,     /* nG x nlevel   state.nonlin[1]                   */
  IBPM_NONLIN2 = 5,     /* nG x nlevel   state.nonlin[2]                   */
  IBPM_XB = 6,          /* npanel x 2    panel positions          state.panels.xb */
  IBPM_UB = 7,          /* npanel x 2    panel velocities         state.panels.ub */
  IBPM_FTILDE = 8,      /* 2 npanel      state.F̃b                          */
  IBPM_FB = 9,          /* npanel x 2    state.fb                          */
  IBPM_TOTAL_FORCE = 10,/* 2             boundary_total_force(state)       */
  IBPM_DS = 11,         /* npanel        state.panels.ds                   */
  IBPM_FREESTREAM = 12  /* 2             state.freestream_vel              */
};

/* Problem description (Problem + State(prob; t0) of problem.jl). */
typedef struct ibpm_problem {
  int nx, ny, nlevel;       /* cells per level (>= 8), levels                */
  double x0, y0, h;         /* lower-left corner and step of level 1          */
  double Re, dt;            /* Reynolds number, CNAB time step                */
  double t0;                /* time at index 1                                */
  int npanel;
  const double* xb;         /* npanel x 2 column-major reference positions    */
  const double* ds;         /* npanel panel lengths (stored, not used)        */
  int nbody;
  const int* body_start;    /* first panel of each body                       */
  const int* body_count;
  const int* body_moving;   /* 0: StaticBody, 1: MovingRigidBody              */
  int grid_moving;          /* 0: StaticGrid; 1: MovingGrid (problem.jl:508-530), the
                               grid's GridVelocity follows the freestream of every step
                               in ibpm_advance                                  */
} ibpm_problem;

typedef struct ibpm_options {
  int precision;      /* enum ibpm_precision                                     */
  int refine_sweeps;  /* IBPM_MIXED: FP64 residual refinement sweeps per solve   */
  int cg_maxiter;     /* moving bodies: default 5000 (solver.jl:137)             */
  double cg_reltol;   /* default 1e-12 (solver.jl:137)                           */
  double cg_abstol;   /* default 0                                               */
  int cg_tol_mode;    /* 0: reltol*||r0|| (IterativeSolvers), 1: reltol*||b||    */
  int device;         /* HIP device ordinal                                      */
  int dst_tiles;      /* sine-pass tiles: -1 (default) autotuned at creation: the
                         variants of the deepest fold a launch shape allows (FP64:
                         the third fold, 20..23, else the double fold, 8..19; mixed:
                         the double fold) are timed on each launch shape the step
                         uses and the fastest kept (process-wide, per shape);
                         0..23 fixed (mixed 0..19): for v < 16, v % 4 picks
                         (TB, TM) = (64,64), (64,32), (32,64), (32,32), FP64 loads
                         run 1 (v % 8 < 4) or 2 chunks ahead, v >= 8 folds the sine
                         matrix twice instead of once; 16..19 fold twice and copy
                         the sine tables straight into LDS with (TB, TM, K chunk) =
                         (32,64,16), (32,64,8), (32,32,16), (32,32,8) where (N+1)/4
                         is a multiple of 16 (else they run as 10 or 11); 20..23
                         (FP64) fold three times with the same (TB, TM, K chunk),
                         TM counting the (N+1)/8 row pairs, where (N+1)/4 is a
                         multiple of 16 and (N+1)/8 >= TM (a 64-row variant takes
                         32 rows where only those fit; else they run as 16..19);
                         mixed uses v % 8 for one fold (TB, TM and the prefetch
                         depth as above). Variants of one fold give identical
                         results (same summation order); the folds agree to
                         rounding.                                                  */
  int cg_precond;     /* moving bodies: 0 (default) the reference's unpreconditioned
                         cg!; 1 preconditions it with the inverse of B assembled at
                         an earlier body configuration, applied in each body's frame
                         (its rotation since the assembly). CG still stops on the
                         true residual of the same tolerance.                       */
  int cg_precond_refresh; /* with cg_precond: reassemble at the next step when a
                         step needed more iterations than this (0: never)          */
  int helmholtz;      /* implicit Helmholtz solves (da I - k L) x = b: 0 (default)
                         red-black Gauss-Seidel in FP64 where 64 half-sweeps bring
                         the error below 2^-56 max|x| (Jacobi contraction
                         rho = 4k/(da+4k) <= 0.54; the sine transforms elsewhere);
                         1 always the sine transforms (the reference's method)     */
  int box_update;     /* FP64 with bodies whose level-1 Helmholtz solve is Gauss-
                         Seidel: 0 (default) the final vort2flux of a step adds the
                         forward transform of the change of Gamma since the trial one
                         (zero outside a box around the bodies) to the trial's instead
                         of transforming Gamma again, the trial forms level 1's psi
                         only where the panels read it, and B runs on the same boxes,
                         when that is cheaper (the boxes follow moving bodies every
                         step); 1 off; 2 wherever valid, cheaper or not (tests)    */
} ibpm_options;

typedef struct ibpm_step_stats {
  int cg_iters;        /* CG iterations of the step (0 for static bodies)        */
  int cg_converged;    /* 1 when the CG stopping test was met                    */
  double cg_residual;  /* final CG residual norm                                 */
  double force[2];     /* boundary_total_force after the step                    */
  int b_applies;       /* applications of the B operator in the step (CG initial
                          residual + iterations + warm-start residual)            */
  int ws_used;         /* IBPM_WS_LEARNED: 1 when the network's guess was kept    */
  double ws_base_residual;  /* ||b - B x_e|| / ||b|| of the extrapolated guess x_e
                               (0 when not computed)                              */
  double ws_residual;  /* ||b - B x0|| / ||b|| of the guess CG started from       */
  int pc_assembled;    /* cg_precond: 1 when the preconditioner was assembled (2n
                          applications of B) in this step                         */
} ibpm_step_stats;

/* Defaults of the reference: FP64, cg maxiter 5000, reltol 1e-12. */
void ibpm_default_options(ibpm_options* o);

int ibpm_create(const ibpm_problem* p, const ibpm_options* o, ibpm_solver** out);

/* Advance nsteps steps. freestream: 2 x nsteps (ux, uy at each new time), or 8 x nsteps
 * for a moving grid (ux, uy, then the grid's GridVelocity at that time: centre x, y,
 * velocity x, y, angle, angular_vel). motion:
 * 6 x nbody x nsteps (pos x, pos y, vel x, vel y, angle, angular_vel of every body at
 * each new time; entries of static bodies are ignored), or NULL: when no body moves, or
 * for one step (nsteps = 1) whose motion every moving body received through
 * ibpm_set_body_motion. stats: nsteps entries or NULL. */
int ibpm_advance(ibpm_solver* s, int nsteps, const double* freestream, const double* motion,
                 ibpm_step_stats* stats);

/* Stage the motion of moving body `body` (0-based, in Bodies order) for the next step:
 * its RigidBodyTransform evaluated by the caller at time t (motion(t), problem.jl:241-261,
 * bodies.jl:43-51). t must be the time of the next step (ibpm_get_time after it); the
 * next ibpm_advance with motion = NULL uses the staged values and clears them. */
int ibpm_set_body_motion(ibpm_solver* s, int body, double t, const double pos[2],
                         const double vel[2], double angle, double angular_vel);

/* Field copies to and from host memory (state.q, .psi, ... as the table above). The step
 * forms the fluxes q only where it needs them; ibpm_read(IBPM_Q) first writes q = curl psi
 * (one launch) when the last step left it unformed, so it always returns the state's q. */
int ibpm_read(ibpm_solver* s, int field, double* dst, size_t n);
int ibpm_write(ibpm_solver* s, int field, const double* src, size_t n);
size_t ibpm_field_size(const ibpm_solver* s, int field);

int ibpm_get_index(const ibpm_solver* s);
int ibpm_set_index(ibpm_solver* s, int index);
double ibpm_get_time(const ibpm_solver* s);

/* Wait for queued work of the solver's stream. */
int ibpm_synchronize(ibpm_solver* s);

/* Operator access for checking the step's building blocks one at a time (used by the
 * parity tests), in the solver's precision:
 *   IBPM_OP_POISSON     Delta_inv_operator, one plane (fluid.jl:117-119, :145-166)
 *   IBPM_OP_HELMHOLTZ   Ainv[level] of A_operators (fluid.jl:121-143), level is 1-based
 *   IBPM_OP_VORT2FLUX   streamfunction of vort2flux (fluid.jl:175-206): in = Gamma
 *                       (nG x nlevel), out = psi (nG x nlevel); in is not modified
 * n is the number of doubles of in and out. */
enum ibpm_op { IBPM_OP_POISSON = 0, IBPM_OP_HELMHOLTZ = 1, IBPM_OP_VORT2FLUX = 2 };
int ibpm_apply(ibpm_solver* s, int op, int level, const double* in, double* out, size_t n);

/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * operators of ibpm_apply, same kernels and results, on DEVICE buffers `in` and `out`
 * (non-overlapping), queued on the solver's stream without waiting (ibpm_synchronize
 * waits). It calls no hipMalloc, hipFree or other allocating routine: its intermediate
 * planes live in one static device buffer of the library, IBPM_NOALLOC_SCRATCH_BYTES per
 * GPU (less a 64 KiB tail it leaves unused), held until the process ends. An operator that
 * needs more fails with IBPM_E_ARG and the size in the message: about (2 + mixed) planes
 * of (nx-1)(ny-1) doubles for the two solves, (3 + mixed) nlevel planes for vort2flux.
 * Every call on one GPU shares that buffer: calls that overlap in time on one GPU (from
 * different streams or threads) race, and must be serialised by the caller. Nothing in the
 * buffer needs to start at zero; it keeps the last call's contents. */
#define IBPM_NOALLOC_SCRATCH_BYTES (32u << 20)
int ibpm_apply_noalloc(ibpm_solver* s, int op, int level, const double* in, double* out,
                       size_t n);

/* Instrumentation, no counterpart in the reference: the mean duration of one kernel of the
 * step, measured with HIP events on the solver's stream.
 *   kernel 0: the sine-contraction passes over all levels, inside `reps` real steps
 *   kernel 1: the fused nonlinear term + CNAB right-hand side, inside `reps` real steps
 *             (kernels 0 and 1 advance the solver by reps steps at its current freestream;
 *             static bodies or none)
 *   kernel 2: `reps` launches of a register-only loop of the MFMA instruction of the
 *             solver's precision (f64 16x16x4 or bf16 16x16x32): the matrix-core rate the
 *             device sustains
 * avg_ms: mean launch duration; flops: MFMA flops executed per launch (0 for kernel 1);
 * bytes: minimum HBM bytes per launch (each input read once, each output written once). */
int ibpm_time_kernel(ibpm_solver* s, int kernel, int reps, double* avg_ms, double* flops,
                     double* bytes);

/* Instrumentation: tile shape (TB values of the batch index x TM output rows per block,
 * loads issued PF K chunks ahead, PF = 0: sine tables copied straight into LDS), fold (2:
 * one even/odd fold, 4: two, 8: three) and K chunk KC that the solver's sine passes of
 * `nplanes` planes along x (xdir != 0) or y use. fold and KC may be NULL. */
int ibpm_dst_tiles(const ibpm_solver* s, int nplanes, int xdir, int* TB, int* TM, int* PF,
                   int* fold, int* KC);

/* Instrumentation: 1 when the solver's steps use the box update of the final
 * vort2flux (ibpm_options.box_update), with box[0..3] = first column, columns, first row,
 * rows (0-based interior-vertex indices) of the change of Gamma on `level` (1-based; 0
 * columns: none on that level); 0 when they do not. level 0: 1 when the trial vort2flux
 * computes level 1's psi only on the box that the interpolation to the panels reads
 * (given in box), 0 when it runs the inverse transforms. */
int ibpm_box_update(const ibpm_solver* s, int level, int box[4]);

const char* ibpm_last_error(void);
void ibpm_destroy(ibpm_solver* s);

/* ---- Batched ensembles ----
 *
 * No counterpart in the reference, where each CnabSolver advances alone
 * (solver.jl:155-190). A batch advances many independent solvers together: any grid
 * sizes, level counts, body-point counts and Reynolds numbers, static or moving bodies,
 * static or moving grids. One step of every solver is captured into a hipGraph (two, one
 * per parity of the nonlinear-term swap) spread over concurrent streams; without moving
 * bodies each batch step is one graph replay, with no host synchronisation. Every solver
 * ends in the same state,
 * bit for bit, as when advanced alone with the same inputs.
 *
 * Each step's freestream, grid motion and body motion are read by the graph from device
 * tables that ibpm_batch_advance fills from its arguments. With moving bodies a step is
 * two graph replays: the first runs up to and including the CG, captured as a budget of
 * iterations whose launches return at once after the stopping test is met; a CG that has
 * not met it then continues on the host's side, up to cg_maxiter; the second replay runs
 * the rest. The budget grows to the iterations the problems need (at most 32; the graphs
 * are captured again). The iterates are those of the solver alone. Moving-body solvers in
 * a batch need: box_update = 1 (off); no warm-start recording; three steps of history for
 * the extrapolated or learned guess; an assembled preconditioner (one step) and
 * cg_precond_refresh = 0 when preconditioned. ibpm_batch_create says which is missing.
 *
 * freestream: 2 x nprob, the velocity held for steps that ibpm_batch_advance gives none
 * (NULL: each solver's last). The batch must be destroyed before its solvers. Between
 * batch steps the solvers may be read, written or advanced alone, as long as all of them
 * advance by the same parity of steps (otherwise ibpm_batch_advance fails). */
typedef struct ibpm_batch ibpm_batch;
int ibpm_batch_create(int nprob, ibpm_solver* const* solvers, const double* freestream,
                      ibpm_batch** out);
/* freestream: nsteps x nprob x 8 (at each new time: ux, uy, then the grid's GridVelocity:
 *             centre x, y, velocity x, y, angle, angular_vel; the last six are ignored for
 *             static grids), or NULL (the held freestream, grids at rest);
 * motion:     nsteps x (6 x nbody of problem 0, 6 x nbody of problem 1, ...), as in
 *             ibpm_advance (entries of static bodies ignored); NULL only without moving
 *             bodies;
 * forces:     nsteps x nprob x 2 (total force of every problem after every step) or NULL;
 * stats:      nsteps x nprob entries or NULL (pc_assembled is always 0).
 * A panel that leaves the allowed region fails the call with IBPM_E_OUTSIDE_REGION after
 * the steps of the call have run (the message names the problem and the step). */
int ibpm_batch_advance(ibpm_batch* b, int nsteps, const double* freestream, const double* motion,
                       double* forces, ibpm_step_stats* stats);
void ibpm_batch_destroy(ibpm_batch* b);

/* ---- Warm start of the CG surface-force solve of moving bodies ----
 *
 * No counterpart in the reference, which starts cg! from the previous step's F̃b
 * (coupling.jl:194-202, :246-247). The modes below replace only the initial guess: CG
 * still runs to the same stopping test, so forces agree with IBPM_WS_PREVIOUS to the CG
 * tolerance whatever the guess.
 *
 *   IBPM_WS_PREVIOUS     x0 = F̃b of the previous step (the reference)
 *   IBPM_WS_EXTRAPOLATE  x0 = x_e, the quadratic extrapolation 3F1 - 3F2 + F3 of the last
 *                        three steps' F̃b (linear, then previous, while fewer exist)
 *   IBPM_WS_LEARNED      x0 = x_e + s * net(features), where s = rms(r_e), r_e = b - B x_e
 *                        (one extra application of B), and the network (ibpm_net) maps
 *                        per-point features to the correction (F̃b - x_e) / s. When
 *                        ||b - B x0|| > ||r_e|| (or is not finite) CG starts from x_e
 *                        instead, without another application of B.
 *
 * Per-point features, in this channel order (each body is one sequence, in panel order):
 *   0,1  r_e (x, y) / rms(r_e)
 *   2,3  sin, cos of 2 pi (x - x0) / h       4,5  sin, cos of 2 pi (y - y0) / h
 *   6,7  unit tangent of the body at the point (central differences, one-sided at ends)
 *   8,9  ub (x, y) / rms(ub)  (0 when the body is at rest)
 */
enum ibpm_warm_start { IBPM_WS_PREVIOUS = 0, IBPM_WS_EXTRAPOLATE = 1, IBPM_WS_LEARNED = 2 };
#define IBPM_NET_FEATURES 10
#define IBPM_NET_OUTPUTS 2

typedef struct ibpm_net ibpm_net;

/* 1-D convolutional network over body points: `layers` convolutions of odd width
 * `kernel` with zero padding, IBPM_NET_FEATURES -> channels -> ... -> IBPM_NET_OUTPUTS,
 * ReLU between layers. BF16 weights and activations, FP32 accumulation on MFMA, FP32
 * master weights trained with Adam. */
typedef struct ibpm_net_config {
  int channels;        /* hidden channels, a multiple of 64 (default 128)       */
  int layers;          /* >= 2 (default 4)                                      */
  int kernel;          /* 3, 5 or 7 (default 5)                                 */
  int max_points;      /* largest body-point count of one sample (default 1024) */
  int capacity;        /* samples the training set holds (default 20000); once
                          full, new samples overwrite the oldest                */
  int max_batch;       /* largest training minibatch in samples (default 64)    */
  unsigned long long seed;  /* weight initialisation and minibatch sampling      */
  int device;
} ibpm_net_config;

void ibpm_net_default_config(ibpm_net_config* c);
int ibpm_net_create(const ibpm_net_config* c, ibpm_net** out);
void ibpm_net_destroy(ibpm_net* net);

/* Select the warm start of a solver. net: required for IBPM_WS_LEARNED or record != 0,
 * and must outlive its use by the solver. record != 0 appends one training sample
 * (features, (F̃b - x_e) / rms(r_e)) per step with moving bodies to the net's training
 * set; it computes r_e (one extra application of B) in every mode. */
/* Switching between the extrapolated and learned modes (or recording) keeps the F~b history
 * of the last steps; switching from IBPM_WS_PREVIOUS starts it anew. */
int ibpm_set_warm_start(ibpm_solver* s, int mode, ibpm_net* net, int record);

int ibpm_net_dataset_size(const ibpm_net* net);  /* samples held */
int ibpm_net_clear_dataset(ibpm_net* net);

/* `iters` Adam steps (betas 0.9, 0.999, eps 1e-8) on minibatches of `batch` samples drawn
 * uniformly from the training set, learning rate decaying from lr to lr/10 along a
 * cosine; loss of the mean squared error of the normalised output. The output scale
 * (rms of the training targets) is fixed at the first call. loss: iters entries or NULL. */
int ibpm_net_train(ibpm_net* net, int iters, int batch, double lr, double* loss);

/* Parameters as float32, per layer the weight [cout][cin][kernel] then the bias [cout],
 * followed by the output scale (one value). */
size_t ibpm_net_param_count(const ibpm_net* net);
int ibpm_net_get_params(const ibpm_net* net, float* p, size_t n);
/* Also resets the Adam state (moments and step count). */
int ibpm_net_set_params(ibpm_net* net, const float* p, size_t n);

/* Network on host rows (at most max_batch * max_points): x is rows x IBPM_NET_FEATURES,
 * seg gives each row's sequence id (rows of one sequence are consecutive; ids >= 0),
 * y is rows x IBPM_NET_OUTPUTS in normalised
 * units (multiply by the output scale for (F̃b - x_e) / rms(r_e)). */
int ibpm_net_forward(ibpm_net* net, const float* x, const int* seg, int rows, float* y);
/* Loss mean((y - target)^2) over rows x IBPM_NET_OUTPUTS and its gradient with respect to
 * the parameters (layout of ibpm_net_get_params; the output-scale entry is 0). */
int ibpm_net_gradients(ibpm_net* net, const float* x, const int* seg, const float* target,
                       int rows, float* grad, size_t n, double* loss);
/* One Adam step (count of steps taken so far + 1) at learning rate lr on the given rows. */
int ibpm_net_step(ibpm_net* net, const float* x, const int* seg, const float* target, int rows,
                  double lr, double* loss);

#ifdef __cplusplus
}
#endif
#endif /* IBPM_H */
