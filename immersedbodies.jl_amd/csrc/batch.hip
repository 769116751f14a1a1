// Batched ensembles: many independent solvers advanced by one hipGraph per step.
//
// One step of every solver of the batch (advance! of each CnabSolver, solver.jl:155-190)
// is captured into one graph: a fork from an origin stream onto kBranches branch
// streams, the solvers' steps spread round-robin over the branches (independent
// problems run concurrently on the chip), and a join. Around each solver's step sit two
// tiny kernels: one copies that step's inputs from device tables indexed by the batch's
// device step counter (the freestream and grid motion, as BaseFlux::dyn, and the body
// motion), the other files the step's total force and CG statistics into a device
// history ring. The _cycle! of the nonlinear terms (fluid.jl:510) and the warm start's
// history swap two buffers every step, so two graphs are captured, one per parity of the
// step, and replayed alternately. A step of the batch is then one hipGraphLaunch,
// whatever the number of problems or kernels.
//
// Moving bodies: a step becomes two graphs. Graph A runs each solver's step up to and
// including its CG (cg!, coupling.jl:194-202), captured as a budget of iterations whose
// launches return at once after the stopping test is met (the solver's graph mode,
// ibpm.hip); graph B runs the rest (forces, project_circ, the final vort2flux). Between
// them the host reads each moving problem's CG state (one small copy): a problem whose CG
// has not met its stopping test within the budget continues on its own stream, with the
// solver's host-driven loop, up to its own cg_maxiter. The iterates are those of the
// solver advanced alone. The budget follows the iterations the problems need (grown and
// the graphs captured again when a step needed more).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/ibpm.h"
#include "ibpm_internal.h"
#include "solver.h"

using namespace ibpm;

namespace {

constexpr int kBranches = 4;      // concurrent streams of the graph (hardware queues)
constexpr int kHistSteps = 1024;  // device tables and history ring, in steps
constexpr int kDyn = 14;          // BaseFlux::dyn values per problem and step
constexpr int kRec = 10;          // history values per problem and step (see record kernel)
constexpr int kCgBudget0 = 8;     // CG iterations captured per step at first
constexpr int kCgBudgetMax = 32;  // most CG iterations captured per step

// This step's inputs of one problem: dyn row -> the solver's dyn buffer, motion row
// (6 per body) -> its motion buffer.
__global__ void batch_inputs_kernel(const double* dyn_tab, const double* mot_tab, const int* counter,
                                    int nprob, int slot, int mrow, int moff, int nmot,
                                    double* dyn, double* motion) {
  const int c = *counter;
  for (int t = threadIdx.x; t < kDyn; t += blockDim.x)
    dyn[t] = dyn_tab[((size_t)c * nprob + slot) * kDyn + t];
  if (motion)
    for (int t = threadIdx.x; t < nmot; t += blockDim.x)
      motion[t] = mot_tab[(size_t)c * mrow + moff + t];
}

// History of one problem after its step: [0..1] total force, [2] final CG residual,
// [3] iterations, [4] tolerance, [5] ||b||, [6] learned guess kept, [7] residual of the
// initial guess, [8] ||r_e|| of the extrapolated guess, [9] panel error (first offending
// panel + 1, 0 if none).
__global__ void batch_record_kernel(const double* total, const double* scal, const double* ws_scal,
                                    const int* err, double* hist, const int* counter, int slot,
                                    int nprob) {
  if (threadIdx.x != 0) return;
  double* r = hist + ((size_t)(*counter) * nprob + slot) * kRec;
  r[0] = total ? total[0] : 0.0;
  r[1] = total ? total[1] : 0.0;
  r[2] = scal ? scal[0] : 0.0;
  r[3] = scal ? scal[3] : 0.0;
  r[4] = scal ? scal[2] : 0.0;
  r[5] = scal ? scal[5] : 0.0;
  r[6] = scal ? scal[7] : 0.0;
  r[7] = scal ? scal[8] : 0.0;
  r[8] = ws_scal ? ws_scal[1] : 0.0;
  r[9] = err ? (double)err[0] : 0.0;
}

// After graph A: a moving problem's CG iterations and converged flag.
__global__ void batch_cgstate_kernel(const double* scal, double* cgst, int slot) {
  if (threadIdx.x == 0) {
    cgst[2 * slot] = scal[3];
    cgst[2 * slot + 1] = scal[4];
  }
}

__global__ void batch_tick_kernel(int* counter) {
  if (threadIdx.x == 0) *counter += 1;
}

__global__ void batch_zero_kernel(int* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

}  // namespace

struct ibpm_batch {
  int nprob = 0;
  int device = 0;
  std::vector<ibpm_solver*> s;
  std::vector<double> fs;  // 8 x nprob: freestream (2) and GridVelocity (6) held when none given
  std::vector<int> moff;   // offset of each problem's motion in a step's motion row
  int mrow = 0;            // motion values per step (6 per body of every problem)
  bool any_moving = false;
  std::vector<int> bapp;   // B applications outside the CG iterations, per problem step
  std::vector<int> br_of;  // branch stream of each problem
  hipStream_t origin = nullptr;
  hipStream_t br[kBranches] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kBranches] = {};
  hipGraphExec_t exec[2] = {nullptr, nullptr};  // per parity: the step (A with moving bodies)
  hipGraphExec_t execB = nullptr;               // moving bodies: the rest of the step
  bool split = false;                           // moving bodies: graphs A and B
  int budget = kCgBudget0;                      // CG iterations captured per step
  double* cgst = nullptr;                       // nprob x 2: CG iterations, converged
  std::vector<double*> nl1_at[2];  // each solver's nl1 when graph g starts
  std::vector<double*> wsh_at[2];  // and its warm-start history head
  double* hist = nullptr;
  int* counter = nullptr;
  double* dyn_tab = nullptr;  // kHistSteps x nprob x kDyn
  double* mot_tab = nullptr;  // kHistSteps x mrow
};

namespace {

#define BCHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_error(IBPM_E_HIP, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)

void batch_free(ibpm_batch* b) {
  for (hipGraphExec_t& e : b->exec)
    if (e) (void)hipGraphExecDestroy(e);
  if (b->execB) (void)hipGraphExecDestroy(b->execB);
  for (void* p : {(void*)b->hist, (void*)b->counter, (void*)b->dyn_tab, (void*)b->mot_tab,
                  (void*)b->cgst})
    if (p) (void)hipFree(p);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (int i = 0; i < kBranches; ++i) {
    if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    if (b->br[i]) (void)hipStreamDestroy(b->br[i]);
  }
  if (b->origin) (void)hipStreamDestroy(b->origin);
}

bool uses_ws(const ibpm_solver* s) { return s->ws_mode != IBPM_WS_PREVIOUS || s->ws_record; }

// Capture into *out one step of every solver (phase 0), or with moving bodies its first
// part (phase 1, graph A) or the rest (phase 2, graph B). Host-side effects of the step on
// the solvers (time index, freestream, the buffer swaps) happen once per capture of a
// step's first part; the caller restores the index.
int capture(ibpm_batch* b, int phase, hipGraphExec_t* out) {
  BCHK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed));
  int rc = IBPM_OK;
  hipError_t e = hipEventRecord(b->fork, b->origin);
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->br[i], b->fork, 0);
  for (int p = 0; p < b->nprob && e == hipSuccess && rc == IBPM_OK; ++p) {
    ibpm_solver* s = b->s[p];
    hipStream_t own = s->st;
    s->st = b->br[b->br_of[p]];
    const bool mov = s->moving && s->n > 0;
    if (phase != 2) {
      hipLaunchKernelGGL(batch_inputs_kernel, dim3(1), dim3(64), 0, s->st, (const double*)b->dyn_tab,
                         (const double*)b->mot_tab, (const int*)b->counter, b->nprob, p, b->mrow,
                         b->moff[p], mov ? 6 * s->nbody : 0, s->dyn_d, mov ? s->motion : nullptr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      s->graph_mode = true;
      s->graph_cg_budget = b->split ? b->budget : 0;
      rc = solver_step(s, &b->fs[8 * p], nullptr, nullptr, phase);
      s->graph_mode = false;
      s->graph_cg_budget = 0;
    }
    if (e == hipSuccess && rc == IBPM_OK && phase == 1 && mov) {
      hipLaunchKernelGGL(batch_cgstate_kernel, dim3(1), dim3(64), 0, s->st,
                         (const double*)s->cg_scal, b->cgst, p);
      e = hipGetLastError();
    }
    if (e == hipSuccess && rc == IBPM_OK && phase != 1) {
      const bool need_res = s->ws_mode == IBPM_WS_LEARNED;
      hipLaunchKernelGGL(batch_record_kernel, dim3(1), dim3(64), 0, s->st,
                         (const double*)(s->n > 0 ? s->total : nullptr),
                         (const double*)(mov ? s->cg_scal : nullptr),
                         (const double*)(mov && need_res ? s->ws_scal : nullptr),
                         (const int*)(s->n > 0 ? s->err : nullptr), b->hist, (const int*)b->counter,
                         p, b->nprob);
      e = hipGetLastError();
    }
    s->st = own;
  }
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) {
    e = hipEventRecord(b->join[i], b->br[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join[i], 0);
  }
  if (e == hipSuccess && phase != 1) {
    hipLaunchKernelGGL(batch_tick_kernel, dim3(1), dim3(64), 0, b->origin, b->counter);
    e = hipGetLastError();
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->origin, &graph);
  if (rc != IBPM_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return set_error(IBPM_E_HIP, (std::string("batch graph capture: ") +
                                  hipGetErrorString(e != hipSuccess ? e : ec)).c_str());
  }
  e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess)
    return set_error(IBPM_E_HIP, (std::string("hipGraphInstantiate: ") + hipGetErrorString(e)).c_str());
  return IBPM_OK;
}

// (Re)capture the batch's graphs from the solvers' current state: exec[0] the step of the
// current parity, exec[1] of the next (and execB with moving bodies); the solvers' host
// state is left as it was.
int capture_all(ibpm_batch* b) {
  const int np = b->nprob;
  // the captured steps read the fluxes from q (eager curl): write q first where a step
  // outside the batch left it to the next nonlinear kernel
  for (int p = 0; p < np; ++p) {
    ibpm_solver* s = b->s[p];
    if (!s->q_stale) continue;
    int rc0 = materialize_q(s);
    if (rc0 == IBPM_OK && hipStreamSynchronize(s->st) != hipSuccess)
      rc0 = set_error(IBPM_E_HIP, "batch capture: stream synchronize failed");
    if (rc0) return rc0;
  }
  for (hipGraphExec_t& e : b->exec)
    if (e) (void)hipGraphExecDestroy(e), e = nullptr;
  if (b->execB) (void)hipGraphExecDestroy(b->execB), b->execB = nullptr;
  std::vector<int> index(np);
  std::vector<double> fs0(2 * (size_t)np), gm0(6 * (size_t)np);
  for (int p = 0; p < np; ++p) {
    index[p] = b->s[p]->index;
    for (int k = 0; k < 2; ++k) fs0[2 * p + k] = b->s[p]->fs[k];
    for (int k = 0; k < 6; ++k) gm0[6 * p + k] = b->s[p]->gm[k];
  }
  int rc = IBPM_OK;
  for (int g = 0; g < 2 && rc == IBPM_OK; ++g) {
    b->nl1_at[g].resize(np);
    b->wsh_at[g].resize(np);
    for (int p = 0; p < np; ++p) {
      b->nl1_at[g][p] = b->s[p]->nl1;
      b->wsh_at[g][p] = b->s[p]->ws_h0;
    }
    rc = capture(b, b->split ? 1 : 0, &b->exec[g]);
    if (rc == IBPM_OK && g == 0 && b->split) rc = capture(b, 2, &b->execB);
  }
  for (int p = 0; p < np; ++p) {
    ibpm_solver* s = b->s[p];
    s->index = index[p];
    for (int k = 0; k < 2; ++k) s->fs[k] = fs0[2 * p + k];
    for (int k = 0; k < 6; ++k) s->gm[k] = gm0[6 * p + k];
    // two captured steps swap the buffers back; a failed capture may not have
    if (b->nl1_at[0].size() == (size_t)np && s->nl1 != b->nl1_at[0][p]) std::swap(s->nl1, s->nl2);
    if (b->wsh_at[0].size() == (size_t)np && uses_ws(s) && s->ws_h0 != b->wsh_at[0][p])
      std::swap(s->ws_h0, s->ws_h1);
  }
  return rc;
}

// Why solver s cannot be captured, or null.
const char* batch_refusal(const ibpm_solver* s) {
  if (!s->moving) return nullptr;
  if (s->box_dyn)
    return "moving bodies in a batch need box_update = 1 (their boxes follow the bodies on the host)";
  if (s->ws_record) return "recording warm-start samples is not available in a batch";
  if (s->ws_mode != IBPM_WS_PREVIOUS && s->nhist < 3)
    return "the extrapolated or learned warm start needs three steps of history before batching";
  if (s->cg_precond && (!s->pc_ready || s->pc_due))
    return "a preconditioned solver needs its preconditioner assembled (one step) before batching";
  if (s->cg_precond && s->cg_precond_refresh > 0)
    return "preconditioner refreshes are host decisions: cg_precond_refresh must be 0 in a batch";
  return nullptr;
}

}  // namespace

extern "C" {

int ibpm_batch_create(int nprob, ibpm_solver* const* solvers, const double* freestream,
                      ibpm_batch** out) {
  if (!out || nprob < 1 || !solvers) return set_error(IBPM_E_ARG, "ibpm_batch_create: bad arguments");
  *out = nullptr;
  const int dev = solvers[0] ? solvers[0]->device : 0;
  for (int p = 0; p < nprob; ++p) {
    const ibpm_solver* s = solvers[p];
    if (!s) return set_error(IBPM_E_ARG, "ibpm_batch_create: null solver");
    if (const char* why = batch_refusal(s))
      return set_error(IBPM_E_ARG, (std::string("ibpm_batch_create: problem ") + std::to_string(p) +
                                    ": " + why).c_str());
    if (s->device != dev) return set_error(IBPM_E_ARG, "ibpm_batch_create: solvers on different devices");
    for (int q = 0; q < p; ++q)
      if (solvers[q] == s) return set_error(IBPM_E_ARG, "ibpm_batch_create: a solver appears twice");
  }
  BCHK(hipSetDevice(dev));
  ibpm_batch* b = new ibpm_batch;
  b->nprob = nprob;
  b->device = dev;
  b->s.assign(solvers, solvers + nprob);
  b->fs.assign(8 * (size_t)nprob, 0.0);
  b->moff.assign(nprob, 0);
  b->bapp.assign(nprob, 0);
  // round-robin over the branches, except that solvers sharing a warm-start network (its
  // activation buffers) run on one branch, in order
  b->br_of.assign(nprob, 0);
  for (int p = 0, next = 0; p < nprob; ++p) {
    const ibpm_solver* s = b->s[p];
    int q = -1;
    if (s->net && s->ws_mode == IBPM_WS_LEARNED)
      for (int r = 0; r < p && q < 0; ++r)
        if (b->s[r]->net == s->net && b->s[r]->ws_mode == IBPM_WS_LEARNED) q = r;
    b->br_of[p] = q >= 0 ? b->br_of[q] : (next++ % kBranches);
  }
  for (int p = 0; p < nprob; ++p) {
    ibpm_solver* s = b->s[p];
    if (freestream) {
      b->fs[8 * p] = freestream[2 * p];
      b->fs[8 * p + 1] = freestream[2 * p + 1];
    } else {
      b->fs[8 * p] = s->fs[0];
      b->fs[8 * p + 1] = s->fs[1];
    }
    for (int k = 0; k < 6; ++k) b->fs[8 * p + 2 + k] = s->gm[k];
    b->moff[p] = b->mrow;
    b->mrow += 6 * s->nbody;
    b->any_moving = b->any_moving || (s->moving && s->n > 0);
    b->bapp[p] = s->moving && s->n > 0 ? 1 + (s->ws_mode == IBPM_WS_LEARNED ? 1 : 0) : 0;
  }
  hipError_t e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) {
    e = hipStreamCreateWithFlags(&b->br[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc((void**)&b->hist, (size_t)kHistSteps * nprob * kRec * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&b->counter, sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void**)&b->dyn_tab, (size_t)kHistSteps * nprob * kDyn * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc((void**)&b->mot_tab, (size_t)kHistSteps * std::max(b->mrow, 1) * sizeof(double));
  b->split = b->any_moving;
  if (e == hipSuccess) e = hipMalloc((void**)&b->cgst, 2 * (size_t)nprob * sizeof(double));
  for (int p = 0; p < nprob && e == hipSuccess; ++p) e = hipStreamSynchronize(b->s[p]->st);
  if (e != hipSuccess) {
    batch_free(b);
    delete b;
    return set_error(IBPM_E_HIP, (std::string("ibpm_batch_create: ") + hipGetErrorString(e)).c_str());
  }
  const int rc = capture_all(b);
  if (rc != IBPM_OK) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return IBPM_OK;
}

int ibpm_batch_advance(ibpm_batch* b, int nsteps, const double* freestream, const double* motion,
                       double* forces, ibpm_step_stats* stats) {
  if (!b || nsteps < 0) return set_error(IBPM_E_ARG, "ibpm_batch_advance: bad arguments");
  if (b->any_moving && nsteps > 0 && !motion)
    return set_error(IBPM_E_ARG, "ibpm_batch_advance: moving bodies need the motion array");
  BCHK(hipSetDevice(b->device));
  int g = -1;
  for (int c = 0; c < 2 && g < 0; ++c) {
    bool all = true;
    for (int p = 0; p < b->nprob; ++p)
      all = all && b->s[p]->nl1 == b->nl1_at[c][p] &&
            (!uses_ws(b->s[p]) || b->s[p]->ws_h0 == b->wsh_at[c][p]);
    if (all) g = c;
  }
  if (g < 0)
    return set_error(IBPM_E_ARG, "ibpm_batch_advance: the solvers were advanced out of step "
                                 "with each other; create a new batch");
  // the graphs read the fluxes from q: write q where a step outside the batch left it to
  // the next nonlinear kernel
  for (int p = 0; p < b->nprob; ++p) {
    const int rc0 = materialize_q(b->s[p]);
    if (rc0) return rc0;
  }
  for (int p = 0; p < b->nprob; ++p) BCHK(hipStreamSynchronize(b->s[p]->st));
  const int np = b->nprob;
  std::vector<double> dyn((size_t)std::min(nsteps, kHistSteps) * np * kDyn);
  std::vector<double> rec((size_t)std::min(nsteps, kHistSteps) * np * kRec);
  std::vector<double> cgst(2 * (size_t)np);
  const double* last_fs = b->fs.data();  // 8 per problem: the values of the last step
  int done = 0;
  int bad_step = -1, bad_prob = -1, bad_panel = 0;
  while (done < nsteps) {
    const int k = std::min(kHistSteps, nsteps - done);
    for (int j = 0; j < k; ++j)
      for (int p = 0; p < np; ++p) {
        const double* f8 = freestream ? freestream + ((size_t)(done + j) * np + p) * 8 : &b->fs[8 * p];
        base_flux_dyn(b->s[p], f8, &dyn[((size_t)j * np + p) * kDyn]);
      }
    BCHK(hipMemcpyAsync(b->dyn_tab, dyn.data(), (size_t)k * np * kDyn * sizeof(double),
                        hipMemcpyHostToDevice, b->origin));
    if (motion && b->mrow > 0)
      BCHK(hipMemcpyAsync(b->mot_tab, motion + (size_t)done * b->mrow,
                          (size_t)k * b->mrow * sizeof(double), hipMemcpyHostToDevice, b->origin));
    hipLaunchKernelGGL(batch_zero_kernel, dim3(1), dim3(64), 0, b->origin, b->counter);
    BCHK(hipGetLastError());
    for (int i = 0; i < k; ++i) {
      BCHK(hipGraphLaunch(b->exec[g], b->origin));
      g ^= 1;
      if (!b->split) continue;
      // moving bodies: continue the CG of problems the budget left short of the stopping
      // test, then run the rest of the step
      BCHK(hipStreamSynchronize(b->origin));
      BCHK(hipMemcpy(cgst.data(), b->cgst, cgst.size() * sizeof(double), hipMemcpyDeviceToHost));
      int needed = 0;
      for (int p = 0; p < np; ++p) {
        ibpm_solver* s = b->s[p];
        if (!(s->moving && s->n > 0) || cgst[2 * p + 1] != 0.0) continue;
        const int rc = solver_cg_continue(s);
        if (rc) return rc;
        BCHK(hipStreamSynchronize(s->st));
        needed = std::max(needed, (int)s->h_scal[3]);
      }
      BCHK(hipGraphLaunch(b->execB, b->origin));
      if (needed > b->budget && b->budget < kCgBudgetMax) {
        // capture more iterations from the next step on (from the state after this one)
        BCHK(hipStreamSynchronize(b->origin));
        b->budget = std::min(kCgBudgetMax, needed + 2);
        // the solvers' host pointers are those after the step just run only once the
        // swaps it made are applied; apply them, capture, and undo
        const int steps_run = done + i + 1;
        for (int p = 0; p < np; ++p) {
          ibpm_solver* s = b->s[p];
          if (steps_run & 1) {
            std::swap(s->nl1, s->nl2);
            if (uses_ws(s)) std::swap(s->ws_h0, s->ws_h1);
          }
        }
        const int rc = capture_all(b);
        for (int p = 0; p < np; ++p) {
          ibpm_solver* s = b->s[p];
          if (steps_run & 1) {
            std::swap(s->nl1, s->nl2);
            if (uses_ws(s)) std::swap(s->ws_h0, s->ws_h1);
          }
        }
        if (rc) return rc;
        g = 0;
      }
    }
    BCHK(hipMemcpyAsync(rec.data(), b->hist, (size_t)k * np * kRec * sizeof(double),
                        hipMemcpyDeviceToHost, b->origin));
    BCHK(hipStreamSynchronize(b->origin));
    for (int j = 0; j < k; ++j)
      for (int p = 0; p < np; ++p) {
        const double* r = &rec[((size_t)j * np + p) * kRec];
        const size_t o = (size_t)(done + j) * np + p;
        if (forces) {
          forces[2 * o] = r[0];
          forces[2 * o + 1] = r[1];
        }
        if (stats) {
          ibpm_step_stats& st = stats[o];
          memset(&st, 0, sizeof st);
          st.force[0] = r[0];
          st.force[1] = r[1];
          st.cg_iters = (int)r[3];
          st.cg_residual = r[2];
          st.cg_converged = b->s[p]->moving && b->s[p]->n > 0 ? (r[2] <= r[4] ? 1 : 0) : 1;
          st.b_applies = b->bapp[p] + st.cg_iters;
          st.ws_used = b->s[p]->ws_mode == IBPM_WS_LEARNED ? (int)r[6] : 0;
          st.ws_base_residual = r[5] > 0.0 ? r[8] / r[5] : 0.0;
          st.ws_residual = r[5] > 0.0 ? r[7] / r[5] : 0.0;
        }
        if (r[9] != 0.0 && bad_step < 0) {
          bad_step = done + j;
          bad_prob = p;
          bad_panel = (int)r[9];
        }
      }
    if (freestream) last_fs = freestream + (size_t)(done + k - 1) * np * 8;
    done += k;
  }
  for (int p = 0; p < np; ++p) {
    ibpm_solver* s = b->s[p];
    s->index += nsteps;
    if (nsteps > 0) {
      s->fs[0] = last_fs[8 * p];
      s->fs[1] = last_fs[8 * p + 1];
      for (int k = 0; k < 6; ++k) s->gm[k] = last_fs[8 * p + 2 + k];
      if (motion && s->moving)
        for (int q = 0; q < s->nbody; ++q)
          s->cur_theta[q] = motion[(size_t)(nsteps - 1) * b->mrow + b->moff[p] + 6 * q + 4];
    }
    s->q0_shadow_valid = false;
    if (nsteps & 1) {
      std::swap(s->nl1, s->nl2);
      if (uses_ws(s)) std::swap(s->ws_h0, s->ws_h1);
    }
  }
  if (bad_step >= 0) {
    ibpm_solver* s = b->s[bad_prob];
    const int pan = s->n - bad_panel;  // 0-based panel (the kernel files n - panel)
    int zero = 0;
    (void)hipMemcpy(s->err, &zero, sizeof(int), hipMemcpyHostToDevice);
    char msg[256];
    snprintf(msg, sizeof msg,
             "problem %d, step %d of the call: Panel index=%d outside allowed region", bad_prob,
             bad_step + 1, pan + 1);
    return set_error(IBPM_E_OUTSIDE_REGION, msg);
  }
  return IBPM_OK;
}

void ibpm_batch_destroy(ibpm_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->origin);
  batch_free(b);
  delete b;
}

}  // extern "C"
