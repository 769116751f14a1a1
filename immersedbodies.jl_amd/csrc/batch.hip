// Batched ensembles: many independent solvers advanced by one hipGraph per step.
//
// One step of every solver of the batch (advance! of each CnabSolver, solver.jl:155-190)
// is captured into one graph: a fork from an origin stream onto kBranches branch
// streams, the solvers' steps spread round-robin over the branches (independent
// problems run concurrently on the chip), a tiny kernel per solver that files its total
// force into a device history ring, and a join. The _cycle! of the nonlinear terms
// (fluid.jl:510) swaps two buffers every step, so two graphs are captured, one per
// parity of the step, and replayed alternately. A step of the batch is then one
// hipGraphLaunch, whatever the number of problems or kernels.
//
// Only static bodies can be batched: a moving body's CG synchronises with the host every
// iteration (the stopping test), which a graph cannot hold. Each problem's freestream is
// fixed for the batch (it is a kernel argument of the captured launches).
#include <hip/hip_runtime.h>
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
constexpr int kHistSteps = 1024;  // force-history ring, in steps

__global__ void batch_record_kernel(const double* total, double* hist, const int* counter,
                                    int slot, int nprob, int cap) {
  if (threadIdx.x == 0) {
    const size_t c = (size_t)(*counter % cap);
    hist[(c * nprob + slot) * 2] = total ? total[0] : 0.0;
    hist[(c * nprob + slot) * 2 + 1] = total ? total[1] : 0.0;
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
  std::vector<double> fs;  // 2 x nprob
  hipStream_t origin = nullptr;
  hipStream_t br[kBranches] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kBranches] = {};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<double*> nl1_at[2];  // each solver's nl1 when graph g starts
  double* hist = nullptr;
  int* counter = nullptr;
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
  if (b->hist) (void)hipFree(b->hist);
  if (b->counter) (void)hipFree(b->counter);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (int i = 0; i < kBranches; ++i) {
    if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    if (b->br[i]) (void)hipStreamDestroy(b->br[i]);
  }
  if (b->origin) (void)hipStreamDestroy(b->origin);
}

// Capture one step of every solver into exec[g]. Host-side effects of the step on the
// solvers (time index, freestream, the nonlinear-term swap) happen once per capture; the
// caller restores the index.
int capture(ibpm_batch* b, int g) {
  BCHK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed));
  int rc = IBPM_OK;
  hipError_t e = hipEventRecord(b->fork, b->origin);
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->br[i], b->fork, 0);
  for (int p = 0; p < b->nprob && e == hipSuccess && rc == IBPM_OK; ++p) {
    ibpm_solver* s = b->s[p];
    hipStream_t own = s->st;
    s->st = b->br[p % kBranches];
    rc = solver_step(s, &b->fs[2 * p], nullptr, nullptr);
    if (rc == IBPM_OK) {
      hipLaunchKernelGGL(batch_record_kernel, dim3(1), dim3(64), 0, s->st,
                         (const double*)(s->n > 0 ? s->total : nullptr), b->hist,
                         (const int*)b->counter, p, b->nprob, kHistSteps);
      e = hipGetLastError();
    }
    s->st = own;
  }
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) {
    e = hipEventRecord(b->join[i], b->br[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->join[i], 0);
  }
  if (e == hipSuccess) {
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
  e = hipGraphInstantiate(&b->exec[g], graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess)
    return set_error(IBPM_E_HIP, (std::string("hipGraphInstantiate: ") + hipGetErrorString(e)).c_str());
  return IBPM_OK;
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
    if (s->moving)
      return set_error(IBPM_E_ARG, "ibpm_batch_create: moving bodies cannot be batched "
                                   "(their CG synchronises with the host)");
    if (s->grid_moving)
      return set_error(IBPM_E_ARG, "ibpm_batch_create: a moving grid's base flux changes every "
                                   "step; batch static grids only");
    if (s->device != dev) return set_error(IBPM_E_ARG, "ibpm_batch_create: solvers on different devices");
    for (int q = 0; q < p; ++q)
      if (solvers[q] == s) return set_error(IBPM_E_ARG, "ibpm_batch_create: a solver appears twice");
  }
  BCHK(hipSetDevice(dev));
  ibpm_batch* b = new ibpm_batch;
  b->nprob = nprob;
  b->device = dev;
  b->s.assign(solvers, solvers + nprob);
  b->fs.assign(2 * (size_t)nprob, 0.0);
  if (freestream) b->fs.assign(freestream, freestream + 2 * (size_t)nprob);
  hipError_t e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) {
    e = hipStreamCreateWithFlags(&b->br[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc((void**)&b->hist, (size_t)kHistSteps * nprob * 2 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&b->counter, sizeof(int));
  for (int p = 0; p < nprob && e == hipSuccess; ++p) e = hipStreamSynchronize(b->s[p]->st);
  if (e != hipSuccess) {
    batch_free(b);
    delete b;
    return set_error(IBPM_E_HIP, (std::string("ibpm_batch_create: ") + hipGetErrorString(e)).c_str());
  }
  std::vector<int> index(nprob);
  for (int p = 0; p < nprob; ++p) index[p] = b->s[p]->index;
  int rc = IBPM_OK;
  for (int g = 0; g < 2 && rc == IBPM_OK; ++g) {
    b->nl1_at[g].resize(nprob);
    for (int p = 0; p < nprob; ++p) b->nl1_at[g][p] = b->s[p]->nl1;
    rc = capture(b, g);
  }
  for (int p = 0; p < nprob; ++p) {
    b->s[p]->index = index[p];
    // two captured steps swap the nonlinear terms back; a failed capture may not have
    if (b->nl1_at[0].size() == (size_t)nprob && b->s[p]->nl1 != b->nl1_at[0][p])
      std::swap(b->s[p]->nl1, b->s[p]->nl2);
  }
  if (rc != IBPM_OK) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return IBPM_OK;
}

int ibpm_batch_advance(ibpm_batch* b, int nsteps, double* forces) {
  if (!b || nsteps < 0) return set_error(IBPM_E_ARG, "ibpm_batch_advance: bad arguments");
  BCHK(hipSetDevice(b->device));
  int g = -1;
  for (int c = 0; c < 2 && g < 0; ++c) {
    bool all = true;
    for (int p = 0; p < b->nprob; ++p) all = all && b->s[p]->nl1 == b->nl1_at[c][p];
    if (all) g = c;
  }
  if (g < 0)
    return set_error(IBPM_E_ARG, "ibpm_batch_advance: the solvers were advanced out of step "
                                 "with each other; create a new batch");
  for (int p = 0; p < b->nprob; ++p) BCHK(hipStreamSynchronize(b->s[p]->st));
  int done = 0;
  while (done < nsteps) {
    const int k = std::min(kHistSteps, nsteps - done);
    hipLaunchKernelGGL(batch_zero_kernel, dim3(1), dim3(64), 0, b->origin, b->counter);
    BCHK(hipGetLastError());
    for (int i = 0; i < k; ++i) {
      BCHK(hipGraphLaunch(b->exec[g], b->origin));
      g ^= 1;
    }
    if (forces)
      BCHK(hipMemcpyAsync(forces + (size_t)done * b->nprob * 2, b->hist,
                          (size_t)k * b->nprob * 2 * sizeof(double), hipMemcpyDeviceToHost,
                          b->origin));
    BCHK(hipStreamSynchronize(b->origin));
    done += k;
  }
  for (int p = 0; p < b->nprob; ++p) {
    ibpm_solver* s = b->s[p];
    s->index += nsteps;
    s->fs[0] = b->fs[2 * p];
    s->fs[1] = b->fs[2 * p + 1];
    s->q0_shadow_valid = false;
    if (nsteps & 1) std::swap(s->nl1, s->nl2);
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
