// libibpm: the IBPM time step (advance! of ImmersedBodies.jl's CnabSolver) on one MI355X.
//
// Host orchestration of the kernels in dst_kernels.hip, stencil_kernels.hip and
// coupling_kernels.hip behind the C ABI of include/ibpm.h. Everything the step touches
// lives in HBM in the reference's layout; the host only launches kernels, evaluates
// nothing per grid point, and synchronises once per CG iteration (moving bodies) or
// once per ibpm_advance call (static bodies).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ibpm.h"
#include "ibpm_device.h"
#include "ibpm_internal.h"
#include "solver.h"

using namespace ibpm;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace

namespace ibpm {
thread_local const double* g_skip = nullptr;

int set_error(int code, const char* msg) {
  g_err = msg;
  return code;
}
}  // namespace ibpm

namespace {

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(IBPM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  return hipMalloc((void**)p, n * sizeof(T));
}

// One plane of a DST-I solve  x = S_x S_y (D . (S_x b S_y)) with the spectral diagonal
// D = scale / (da + lambda dnum / dden), optionally x = base - solve; the system operator
// is a I + c K (for the refinement residual).
struct SolveJob {
  const double* b;
  double da, dnum, dden;
  double* x;
  int epi;             // EPI_STORE, EPI_SUB (x = base - solve)
  const double* base;
  double a, c;
  double* spec = nullptr;  // FP64: the solve's spectral plane, unscaled (PassPlane::spec)
  // Gauss-Seidel only: solve on vertices [ri0, ri0 + rw) x [rj0, rj0 + rh) with x = 0
  // around them, writing only there (rw = 0: the whole plane)
  int ri0 = 0, rj0 = 0, rw = 0, rh = 0;
};

}  // namespace


namespace {

void free_all(ibpm_solver* s) {
  void* ptrs[] = {s->q, s->G, s->psi, s->nl1, s->nl2, s->Gs, s->rhs, s->T, s->U, s->R,
                  s->cosx, s->cosy, s->SoX, s->SeX, s->SoY, s->SeY, s->SqX, s->SqY, s->BfX, s->BfY, s->BqX, s->BqY,
                  s->xb, s->ub, s->xb_ref, s->Ft, s->fb, s->Ftmp, s->total, s->motion,
                  s->body_of, s->moving_d, s->coords, s->perm, s->xstart, s->err, s->w,
                  s->sc_rowptr, s->sc_ep, s->sc_ew,
                  s->ywin, s->gemv_part, s->pc_P, s->pc_theta0, s->pc_t, s->pc_z,
                  s->Binv, s->Gtmp, s->psitmp, s->Sc, s->cg_r, s->cg_u, s->cg_c, s->cg_scal, s->dyn_d,
                  s->F1, s->G1, s->Sfx, s->Sfy != s->Sfx ? s->Sfy : nullptr, s->Ftr, s->Tbox, s->Tpart, s->Pa, s->Pb, s->T2, s->bbox_d,
                  s->edge, s->pw, s->fac, s->rvec, s->rpart, s->Mx,
                  s->My != s->Mx ? s->My : nullptr,
                  s->ws_h0, s->ws_h1, s->ws_xe,
                  s->ws_Bxe, s->ws_scal};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->h_scal) (void)hipHostFree(s->h_scal);
  if (s->h_force) (void)hipHostFree(s->h_force);
  if (s->st) (void)hipStreamDestroy(s->st);
}

const void* tab_odd(const ibpm_solver* s, bool xdir) {
  if (s->precision == IBPM_MIXED) return xdir ? s->BfX : s->BfY;
  return xdir ? s->SoX : s->SoY;
}
const void* tab_even(const ibpm_solver* s, bool xdir) {
  if (s->precision == IBPM_MIXED) {
    const int N = xdir ? s->g.Nx : s->g.Ny;
    const uint16_t* base = (const uint16_t*)(xdir ? s->BfX : s->BfY);
    return base + sine_table_bf16_odd_count(N);
  }
  return xdir ? s->SeX : s->SeY;
}

// Instrumentation: events around one launch of the probed kernel class (the pool holds
// the pairs of one step; ibpm_time_kernel reads them after each step).
void probe_begin(ibpm_solver* s) {
  if (2 * s->probe_n + 1 < (int)s->probe_ev.size()) (void)hipEventRecord(s->probe_ev[2 * s->probe_n], s->st);
}
void probe_end(ibpm_solver* s) {
  if (2 * s->probe_n + 1 < (int)s->probe_ev.size()) {
    (void)hipEventRecord(s->probe_ev[2 * s->probe_n + 1], s->st);
    ++s->probe_n;
  }
}

// One pass over up to kMaxPlanes planes.
int run_pass(ibpm_solver* s, PassBatch& pb, int np, bool xdir, int epi, bool in_f32,
             bool out_f32) {
  const int Nk = xdir ? s->g.Nx : s->g.Ny;
  const int Nb = xdir ? s->g.Ny : s->g.Nx;
  for (int i = 0; i < np; ++i) {
    pb.p[i].Nk = Nk;
    pb.p[i].Nb = Nb;
    pb.p[i].So = tab_odd(s, xdir);
    pb.p[i].Se = tab_even(s, xdir);
    pb.p[i].Sq = xdir ? s->SqX : s->SqY;
    pb.p[i].Sqm = xdir ? s->BqX : s->BqY;
    pb.p[i].cosb = xdir ? s->cosy : s->cosx;
    pb.p[i].cosj = xdir ? s->cosx : s->cosy;
    pb.p[i].dscale = s->scale;
  }
  const int H = (Nk + 1) / 2;
  const bool probed = s->probe == 0 && np == s->L;
  if (probed) probe_begin(s);
  if (s->precision == IBPM_MIXED)
    HIPCHK(launch_dst_pass_mixed(pb, np, Nb, H, epi, in_f32, out_f32, s->st, s->dst_tiles));
  else
    HIPCHK(launch_dst_pass_f64(pb, np, Nb, H, epi, s->st, s->dst_tiles));
  if (probed) probe_end(s);
  return IBPM_OK;
}

// Autotune the sine-pass tile variant of every launch shape the step uses: the batched
// level solves (L planes), single-plane solves, and the column groups of the B operator.
int autotune_passes(ibpm_solver* s) {
  const bool mixed = s->precision == IBPM_MIXED;
  const size_t P = s->plane();
  HIPCHK(hipMemsetAsync(s->T, 0, kMaxPlanes * P * sizeof(double), s->st));
  // single planes, the levels of one column, the column groups of the B operator, and the
  // two finest levels of one column or of a B-operator group (inverse passes of vort2flux
  // when only they are read)
  const int g = std::max(1, kMaxPlanes / s->L), two = std::min(2, s->L);
  const int counts[5] = {1, s->L, g * s->L, two, g * two};
  for (int ci = 0; ci < 5; ++ci) {
    const int np = counts[ci];
    bool seen = false;
    for (int cj = 0; cj < ci; ++cj) seen = seen || counts[cj] == np;
    if (seen || np < 1 || np > kMaxPlanes) continue;
    for (int xd = 0; xd < 2; ++xd) {
      const bool xdir = xd == 0;
      if (!xdir && s->g.Nx == s->g.Ny) continue;  // same shape as the x pass
      PassBatch pb;
      memset(&pb, 0, sizeof pb);
      const int Nk = xdir ? s->g.Nx : s->g.Ny;
      const int Nb = xdir ? s->g.Ny : s->g.Nx;
      for (int i = 0; i < np; ++i) {
        pb.p[i].in = s->T + i * P;
        pb.p[i].out = s->U + i * P;
        pb.p[i].Nk = Nk;
        pb.p[i].Nb = Nb;
        pb.p[i].So = tab_odd(s, xdir);
        pb.p[i].Se = tab_even(s, xdir);
        pb.p[i].Sq = xdir ? s->SqX : s->SqY;
        pb.p[i].Sqm = xdir ? s->BqX : s->BqY;
        pb.p[i].cosb = xdir ? s->cosy : s->cosx;
        pb.p[i].cosj = xdir ? s->cosx : s->cosy;
        pb.p[i].dscale = s->scale;
      }
      HIPCHK(tune_dst_pass(pb, np, Nb, (Nk + 1) / 2, mixed, mixed, mixed, s->st, nullptr, nullptr));
    }
  }
  return IBPM_OK;
}

// DST-I solves x = S (D . (S b S)) S of up to kMaxPlanes independent planes (the
// implicit Helmholtz solves); scratch planes T[i], U[i], R[i]. In mixed precision every
// solve is followed by `refine` sweeps x += solve(b - (a I + c K) x) with an FP64 residual.
int solve_chunk(ibpm_solver* s, const SolveJob* jobs, int nj) {
  const size_t P = s->plane();
  const bool mixed = s->precision == IBPM_MIXED;
  PassBatch pb;
  auto four_passes = [&](bool from_residual, bool correction) -> int {
    memset(&pb, 0, sizeof pb);
    for (int i = 0; i < nj; ++i) {
      pb.p[i].in = from_residual ? (const void*)(s->R + i * P) : (const void*)jobs[i].b;
      pb.p[i].out = s->T + i * P;
    }
    int rc = run_pass(s, pb, nj, true, EPI_STORE, from_residual && mixed, mixed);
    if (rc) return rc;
    memset(&pb, 0, sizeof pb);
    for (int i = 0; i < nj; ++i) {
      pb.p[i].in = s->T + i * P;
      pb.p[i].out = s->U + i * P;
      pb.p[i].da = jobs[i].da;
      pb.p[i].dnum = jobs[i].dnum;
      pb.p[i].dden = jobs[i].dden;
      pb.p[i].spec = correction || mixed ? nullptr : jobs[i].spec;
    }
    rc = run_pass(s, pb, nj, false, EPI_DIAG, mixed, mixed);
    if (rc) return rc;
    memset(&pb, 0, sizeof pb);
    for (int i = 0; i < nj; ++i) {
      pb.p[i].in = s->U + i * P;
      pb.p[i].out = s->T + i * P;
    }
    rc = run_pass(s, pb, nj, true, EPI_STORE, mixed, mixed);
    if (rc) return rc;
    memset(&pb, 0, sizeof pb);
    int epi = EPI_STORE;
    for (int i = 0; i < nj; ++i) {
      pb.p[i].in = s->T + i * P;
      pb.p[i].out = jobs[i].x;
      if (correction) {
        epi = jobs[i].epi == EPI_SUB ? EPI_SUB : EPI_ADD;
        pb.p[i].base = jobs[i].x;
      } else {
        epi = jobs[i].epi;
        pb.p[i].base = jobs[i].base;
      }
    }
    return run_pass(s, pb, nj, false, epi, mixed, false);
  };
  int rc = four_passes(false, false);
  if (rc) return rc;
  if (mixed) {
    for (int sweep = 0; sweep < s->refine; ++sweep) {
      ResidualArgs ra;
      memset(&ra, 0, sizeof ra);
      ra.g = s->g;
      ra.nplanes = nj;
      ra.r_f32 = 1;
      for (int i = 0; i < nj; ++i) {
        ra.b[i] = jobs[i].b;
        ra.x[i] = jobs[i].x;
        ra.r[i] = s->R + i * P;
        ra.a[i] = jobs[i].a;
        ra.c[i] = jobs[i].c;
        ra.ring[i] = nullptr;
        ra.xbase[i] = jobs[i].epi == EPI_SUB ? jobs[i].base : nullptr;
      }
      HIPCHK(launch_residual(ra, s->st));
      rc = four_passes(true, true);
      if (rc) return rc;
    }
  }
  return IBPM_OK;
}

// ---- Helmholtz solves by red-black Gauss-Seidel (gs_kernels.hip) ----
constexpr int kGsMaxSweeps = 64;
constexpr int kPsiSlicesA = 8, kPsiSlicesB = 16;  // K slices of the trial psi box products
constexpr int kPsiFoldSlices = 4;  // K slices of the folded first product (<= kPsiSlicesA)
constexpr int kBoxSliceK = 128;  // K per slice of a long first box-update product

// Half-sweeps m with rho^(m-1) <= 2^-56 for (da I + k (-L)) x = b, k = dnum / dden, where
// rho = 4k / (da + 4k) bounds the error contraction per half-sweep in the max norm; 0 when
// the job is no Helmholtz solve (da = 0: the Poisson solve) or m would exceed kGsMaxSweeps.
int gs_sweeps(const SolveJob& j) {
  if (!(j.da > 0.0) || !(j.dnum > 0.0) || !(j.dden > 0.0)) return 0;
  const double k = j.dnum / j.dden;
  const double rho = 4.0 * k / (j.da + 4.0 * k);
  if (!(rho > 0.0)) return 1;
  const double m = 1.0 + std::ceil(56.0 * std::log(2.0) / -std::log(rho));
  return m > kGsMaxSweeps ? 0 : (int)m;
}

bool use_gs(const ibpm_solver* s, const SolveJob& j) {
  return s->helmholtz == 0 && (j.epi == EPI_STORE || j.epi == EPI_SUB) && gs_sweeps(j) > 0;
}

// Up to kMaxPlanes Gauss-Seidel solves: launches of at most kGsHalo half-sweeps each,
// the iterate passed between launches through the scratch planes T[i] and U[i].
int solve_gs(ibpm_solver* s, const SolveJob* jobs, int nj) {
  const size_t P = s->plane();
  int m[kMaxPlanes], done[kMaxPlanes];
  for (int i = 0; i < nj; ++i) {
    m[i] = gs_sweeps(jobs[i]);
    done[i] = 0;
    if (jobs[i].rw != jobs[0].rw || jobs[i].rh != jobs[0].rh || jobs[i].ri0 != jobs[0].ri0 ||
        jobs[i].rj0 != jobs[0].rj0)
      return fail(IBPM_E_ARG, "Gauss-Seidel solves of one launch share their region");
  }
  // a region: plane pointers offset to its first vertex, colours kept those of the plane
  const bool region = jobs[0].rw > 0;
  const size_t o = region ? (size_t)jobs[0].ri0 + (size_t)jobs[0].rj0 * s->g.Nx : 0;
  const int coff = region ? (jobs[0].ri0 + jobs[0].rj0) & 1 : 0;
  for (int round = 0;; ++round) {
    GsBatch bt;
    memset(&bt, 0, sizeof bt);
    bt.Nx = region ? jobs[0].rw : s->g.Nx;
    bt.Ny = region ? jobs[0].rh : s->g.Ny;
    bt.ld = s->g.Nx;
    int np = 0;
    for (int i = 0; i < nj; ++i) {
      if (done[i] >= m[i]) continue;
      const SolveJob& j = jobs[i];
      const double k = j.dnum / j.dden, d = j.da + 4.0 * k;
      GsPlane& q = bt.p[np++];
      q.b = j.b + o;
      q.c = k / d;
      q.inv_d = 1.0 / d;
      q.nsweep = std::min(kGsHalo, m[i] - done[i]);
      q.color0 = (done[i] & 1) ^ coff;
      q.xin = done[i] ? ((round - 1) & 1 ? s->U : s->T) + i * P + o : nullptr;
      const bool last = done[i] + q.nsweep == m[i];
      q.out = (last ? j.x : ((round & 1) ? s->U : s->T) + i * P) + o;
      q.base = last && j.epi == EPI_SUB ? j.base + o : nullptr;
      done[i] += q.nsweep;
    }
    if (!np) break;
    HIPCHK(launch_gs_helmholtz(bt, np, s->st));
  }
  return IBPM_OK;
}

// Helmholtz solves that Gauss-Seidel brings to rounding go to solve_gs, the others (and
// the Poisson solves) to the sine transforms; a chunk of the latter shares one epilogue.
int solve(ibpm_solver* s, const SolveJob* jobs, int nj) {
  std::vector<SolveJob> gs, rest;
  for (int i = 0; i < nj; ++i) (use_gs(s, jobs[i]) ? gs : rest).push_back(jobs[i]);
  for (size_t i = 0; i < gs.size(); i += kMaxPlanes) {
    int rc = solve_gs(s, gs.data() + i, (int)std::min<size_t>(kMaxPlanes, gs.size() - i));
    if (rc) return rc;
  }
  const int nr = (int)rest.size();
  int i = 0;
  while (i < nr) {
    int k = i + 1;
    while (k < nr && k - i < kMaxPlanes && rest[k].epi == rest[i].epi) ++k;
    int rc = solve_chunk(s, rest.data() + i, k - i);
    if (rc) return rc;
    i = k;
  }
  return IBPM_OK;
}

// Multi-level Poisson solve of vort2flux for ng columns of L planes (ng*L <= kMaxPlanes):
// one batched forward pass pair, the FP64 ring chain level L..2, one batched inverse pass
// pair whose first pass adds the ring terms. in[p], out[p]: plane p = column*L + level.
// spec_in[p] (FP64, may be null): plane p's spectral plane S Gamma S is known, and its
// forward passes are skipped; spec_out[p] (may be null): where the forward passes put it
// (else the scratch plane U[p]).
int psi_box_product(ibpm_solver* s, const double* E, double* psi);

// psi_box (FP64): instead of the inverse passes, only level 1's psi of every column on the
// psi box (psi_box_product); the rest of out[] is left as it was.
int v2f_group(ibpm_solver* s, int ng, const void* const* in, bool in_f32, double* const* out,
              int epi, int nout, const double* const* spec_in = nullptr,
              double* const* spec_out = nullptr, bool psi_box = false) {
  const size_t P = s->plane();
  const int L = s->L, np = ng * L;
  const bool mixed = s->precision == IBPM_MIXED;
  const size_t nf = 2 * (size_t)s->g.Nx + 2 * (size_t)s->g.Ny;
  const void* F[kMaxPlanes];  // spectral plane of each plane
  int fw[kMaxPlanes], nfw = 0;  // planes that need the forward passes
  for (int p = 0; p < np; ++p) {
    if (spec_in && spec_in[p]) {
      F[p] = spec_in[p];
    } else {
      F[p] = spec_out && spec_out[p] ? spec_out[p] : s->U + p * P;
      fw[nfw++] = p;
    }
  }
  PassBatch pb;
  memset(&pb, 0, sizeof pb);
  for (int i = 0; i < nfw; ++i) {
    pb.p[i].in = in[fw[i]];
    pb.p[i].out = s->T + fw[i] * P;
  }
  int rc = nfw ? run_pass(s, pb, nfw, true, EPI_STORE, in_f32, mixed) : IBPM_OK;
  if (rc) return rc;
  memset(&pb, 0, sizeof pb);
  for (int i = 0; i < nfw; ++i) {
    pb.p[i].in = s->T + fw[i] * P;
    pb.p[i].out = (void*)F[fw[i]];
  }
  rc = nfw ? run_pass(s, pb, nfw, false, EPI_STORE, mixed, mixed) : IBPM_OK;
  if (rc) return rc;
  // ring chain, coarse to fine: E_l = D . (F_l + R_l) into T and, for l >= 1, the ring
  // factors of level l-1 from the projections of E_l
  RingChainArgs ra;
  memset(&ra, 0, sizeof ra);
  ra.nx = s->g.nx;
  ra.ny = s->g.ny;
  ra.Nx = s->g.Nx;
  ra.Ny = s->g.Ny;
  ra.F_f32 = mixed ? 1 : 0;
  ra.cosx = s->cosx;
  ra.cosy = s->cosy;
  ra.scale = s->scale;
  ra.edge = s->edge;
  ra.pw = s->pw;
  ra.Mx = s->Mx;
  ra.My = s->My;
  const size_t nv = ring_vector_count(s->g.Nx, s->g.Ny);
  const size_t npart = ring_partial_count(s->g.Nx, s->g.Ny);
  const size_t npa = ring_part_a_count(s->g.Nx, s->g.Ny);
  for (int l = L - 1; l >= 0; --l) {
    for (int c = 0; c < ng; ++c) {
      ra.F[c] = F[c * L + l];
      ra.E[c] = s->T + (c * L + l) * P;
      ra.fac_in[c] = l == L - 1 ? nullptr : s->fac + (c * L + l) * nf;
      ra.fac_out[c] = l > 0 ? s->fac + (c * L + l - 1) * nf : nullptr;
      ra.vec[c] = s->rvec + c * nv;
      ra.part_a[c] = s->rpart + c * npart;
      ra.part_c[c] = s->rpart + c * npart + npa;
    }
    // E is read by the inverse passes of the levels below nout, or for level 1 only by
    // the psi box products
    ra.write_e = psi_box ? l == 0 : l < nout;
    HIPCHK(launch_ring_sweep(ra, ng, l > 0, s->st));
    if (l > 0) HIPCHK(launch_ring_factor(ra, ng, s->st));
  }
  if (psi_box) {
    for (int c = 0; c < ng; ++c) {
      rc = psi_box_product(s, s->T + (size_t)c * L * P, out[c * L]);
      if (rc) return rc;
    }
    return IBPM_OK;
  }
  // inverse passes of the levels the caller reads (l < nout of every column)
  memset(&pb, 0, sizeof pb);
  int no = 0;
  for (int c = 0; c < ng; ++c)
    for (int l = 0; l < nout; ++l) {
      pb.p[no].in = s->T + (c * L + l) * P;
      pb.p[no].out = s->U + (c * L + l) * P;
      ++no;
    }
  rc = run_pass(s, pb, no, true, EPI_STORE, mixed, mixed);
  if (rc) return rc;
  memset(&pb, 0, sizeof pb);
  no = 0;
  for (int c = 0; c < ng; ++c)
    for (int l = 0; l < nout; ++l) {
      const int p = c * L + l;
      pb.p[no].in = s->U + p * P;
      pb.p[no].out = out[p];
      pb.p[no].base = out[p];
      ++no;
    }
  return run_pass(s, pb, no, false, epi, mixed, false);
}

// vort2flux (fluid.jl:175-206) for ncol independent columns of L planes each. coarsify!
// writes the centre regions in place; with fwd_given (the final call of a step with box
// updates) column 0 changed since the trial call only on level 1's box, so coarsify! is
// re-evaluated on each level's box only, the trial values it overwrites saved into
// saved[0] (which holds level 1's already). q: curl all levels.
// nout: psi is written on the nout finest levels only (the others are only needed for
// their boundary values, which the spectral ring chain carries); the mixed path's
// refinement reads every level and always writes all of them. spec0 (FP64): the forward
// transform S Gamma_1 S of column 0's finest level, known from the solve that produced
// it (S S = 2 n I makes a sine solve's spectral plane that transform), so its two
// forward passes are skipped.
int box_update(ibpm_solver* s, const double* G, const double* Gs, double* F, bool accumulate);

// fwd (FP64, column 0): the L spectral planes of column 0; with fwd_given they hold those
// of the trial vort2flux of the step, which box_update turns into this call's (Gc[0] minus
// saved[0] on the boxes, zero elsewhere), else the forward passes write them there.
int vort2flux(ibpm_solver* s, int ncol, double* const* Gc, double* const* psic,
              double* const* saved, double* const* qc, int nout = 0,
              const double* spec0 = nullptr, double* fwd = nullptr, bool fwd_given = false,
              bool psi_box = false) {
  const size_t P = s->plane();
  const int L = s->L;
  const bool mixed = s->precision == IBPM_MIXED;
  if (nout <= 0 || nout > L || mixed || qc) nout = L;
  if (mixed) spec0 = nullptr, fwd = nullptr;
  if (!fwd) fwd_given = false;
  if (fwd_given && !(saved && saved[0])) return fail(IBPM_E_ARG, "box update without saved planes");
  if (mixed || ncol != 1 || qc || !s->pbox_on) psi_box = false;
  for (int c = 0; c < ncol; ++c) {
    if (fwd_given && c == 0) {
      BoxSet bs;
      memset(&bs, 0, sizeof bs);
      for (int l = 0; l < L; ++l)
        bs.i0[l] = s->bi0[l], bs.j0[l] = s->bj0[l], bs.w[l] = s->bw[l], bs.h[l] = s->bh[l];
      HIPCHK(launch_coarsify_boxes(s->g, L, Gc[0], saved[0], bs, s->st));
      continue;
    }
    for (int l = 1; l < L; ++l)
      HIPCHK(launch_coarsify(s->g, Gc[c] + l * P, Gc[c] + (l - 1) * P, nullptr, s->st));
  }
  if (fwd_given) {
    int rc = box_update(s, Gc[0], saved[0], fwd, true);
    if (rc) return rc;
  }
  const double* sin0[kMaxPlanes] = {};
  double* sout0[kMaxPlanes] = {};
  for (int l = 0; l < L; ++l) {
    if (fwd_given) sin0[l] = fwd + l * P;
    else if (fwd) sout0[l] = fwd + l * P;
  }
  if (spec0) sin0[0] = spec0;
  const int gmax = std::max(1, kMaxPlanes / L);
  for (int c0 = 0; c0 < ncol; c0 += gmax) {
    const int ng = std::min(gmax, ncol - c0);
    const void* in[kMaxPlanes];
    double* out[kMaxPlanes];
    for (int c = 0; c < ng; ++c)
      for (int l = 0; l < L; ++l) {
        in[c * L + l] = Gc[c0 + c] + l * P;
        out[c * L + l] = psic[c0 + c] + l * P;
      }
    int rc = v2f_group(s, ng, in, false, out, EPI_STORE, nout, c0 == 0 ? sin0 : nullptr,
                       c0 == 0 ? sout0 : nullptr, psi_box);
    if (rc) return rc;
    if (mixed) {
      // refinement of the coupled multi-level system: r_l = Gamma_l + apply_bc(get_bc(
      // psi_{l+1})) - K psi_l, corrected through the same solve
      for (int sweep = 0; sweep < s->refine; ++sweep) {
        ResidualArgs ra;
        memset(&ra, 0, sizeof ra);
        ra.g = s->g;
        ra.nplanes = ng * L;
        ra.r_f32 = 1;
        const void* rin[kMaxPlanes];
        for (int c = 0; c < ng; ++c)
          for (int l = 0; l < L; ++l) {
            const int p = c * L + l;
            ra.b[p] = Gc[c0 + c] + l * P;
            ra.x[p] = psic[c0 + c] + l * P;
            ra.r[p] = s->R + p * P;
            ra.a[p] = 0.0;
            ra.c[p] = 1.0;
            ra.ring[p] = l + 1 < L ? psic[c0 + c] + (l + 1) * P : nullptr;
            ra.xbase[p] = nullptr;
            rin[p] = s->R + p * P;
          }
        HIPCHK(launch_residual(ra, s->st));
        rc = v2f_group(s, ng, rin, true, out, EPI_ADD, L);
        if (rc) return rc;
      }
    }
  }
  if (qc)
    for (int c = 0; c < ncol; ++c) HIPCHK(launch_curl_all(s->g, L, qc[c], psic[c], s->st));
  return IBPM_OK;
}

double cA(const ibpm_solver* s, int l) { return s->dt / (2.0 * s->Re * (s->hc[l] * s->hc[l])); }

// Forward transform of the change D_l = G_l - Gs_l of each level (of G_l when Gs is null),
// which vanishes outside box l, added to F_l (or stored there when not accumulate):
// F_l += S_x[:, I] D_l[I, J] S_y[J, :], as two products with the sine
// columns of the box, in the cheaper order (the x contraction first: an Nx x |J|
// intermediate; else a |I| x Ny one). All levels' first products in one launch, then all
// second ones.
int box_update(ibpm_solver* s, const double* G, const double* Gs, double* F, bool accumulate) {
  const size_t P = s->plane();
  const int Nx = s->g.Nx, Ny = s->g.Ny;
  GemmBatch g1, g2, g2m;  // g2m: second products of the x-first levels (mirror jobs)
  SliceSumBatch sb;
  memset(&g1, 0, sizeof g1);
  memset(&g2, 0, sizeof g2);
  memset(&g2m, 0, sizeof g2m);
  memset(&sb, 0, sizeof sb);
  int n1 = 0, n2 = 0, n2m = 0, nsum = 0;
  for (int l = 0; l < s->L; ++l) {
    if (s->bw[l] <= 0 || s->bh[l] <= 0) continue;
    const size_t o = (size_t)s->bi0[l] + (size_t)s->bj0[l] * Nx;
    const double *Gl = G + l * P + o, *Gsl = Gs ? Gs + l * P + o : nullptr;
    double* T = s->Tbox + s->btoff[l];
    const bool xf = s->bxfirst[l] != 0;
    // T (Nx x bh) = S_x[:, I] D (bw x bh);  F += T S_y[J, :]   (x first), or
    // T (bw x Ny) = D S_y[J, :];            F += S_x[:, I] T   (y first)
    const int M1 = xf ? Nx : s->bw[l], N1 = xf ? s->bh[l] : Ny, K1 = xf ? s->bw[l] : s->bh[l];
    const int ldt = xf ? Nx : s->bw[l];
    const size_t tsize = (size_t)M1 * N1;
    const int ns = std::max(1, s->bns[l]);
    const int kk = ns > 1 ? ((K1 + ns - 1) / ns + 31) / 32 * 32 : K1;
    int made = 0;
    for (int k0 = 0; k0 < K1; k0 += kk, ++made) {
      GemmJob& a = g1.j[n1++];
      if (xf) {
        a.A = s->Sfx + (size_t)(s->bi0[l] + k0) * Nx, a.lda = Nx;
        a.B = Gl + k0, a.B2 = Gsl ? Gsl + k0 : nullptr, a.ldb = Nx;
      } else {
        a.A = Gl + (size_t)k0 * Nx, a.A2 = Gsl ? Gsl + (size_t)k0 * Nx : nullptr, a.lda = Nx;
        a.B = s->Sfy + s->bj0[l] + k0, a.ldb = Ny;
      }
      a.C = ns > 1 ? s->Tpart + s->bpoff[l] + (size_t)made * tsize : T, a.ldc = ldt;
      a.M = M1, a.N = N1, a.K = std::min(kk, K1 - k0);
    }
    if (ns > 1) sb.j[nsum++] = SliceSumJob{s->Tpart + s->bpoff[l], tsize, made, ldt, M1, N1, T, ldt};
    GemmJob& b = xf ? g2m.j[n2m++] : g2.j[n2++];
    b.M = Nx, b.N = Ny;
    if (xf) {
      b.A = T, b.lda = Nx;
      b.B = s->Sfy + s->bj0[l], b.ldb = Ny;
      b.K = s->bh[l];
      // S_y[j, Ny-1-n] = (-1)^j S_y[j, n]: columns n and Ny-1-n from the sums over even and
      // odd rows j = bj0 + k, half the products
      b.mirror = 1, b.mref = Ny - 1, b.kpar = s->bj0[l] & 1;
      b.N = (Ny + 1) / 2;
    } else {
      b.A = s->Sfx + (size_t)s->bi0[l] * Nx, b.lda = Nx;
      b.B = T, b.ldb = s->bw[l];
      b.K = s->bw[l];
    }
    b.C = F + l * P, b.ldc = Nx;
    b.accumulate = accumulate ? 1 : 0;
  }
  HIPCHK(launch_gemm_f64(g1, n1, s->st));
  if (nsum) HIPCHK(launch_sum_slices_batch(sb, nsum, s->st));
  if (n2m) HIPCHK(launch_gemm_f64(g2m, n2m, s->st));
  if (n2) HIPCHK(launch_gemm_f64(g2, n2, s->st));
  return IBPM_OK;
}

// The boxes of the box updates (FP64, Gauss-Seidel on level 1). A panel in flux cell
// (ix, iy) feeds rot(E^T f) on vertices [ix - 4, ix + 3] x [iy - 4, iy + 3]
// (scatter_rot_kernel); m Gauss-Seidel half-sweeps from zero move the support of
// Ainv1 rot(E^T f), hence of Gamma_1 - Gamma_s1 and of B's input, by at most m cells;
// coarsify! maps fine vertices 2 + 2a + {-1, 0, 1} to the coarse vertex ic0 + a of the
// centre, and outside the centre Gamma_l = Gamma_s,l (B's input: zero). Each box gets
// one more cell on every side.

// the level-1 Helmholtz job of project_circ and B; its Gauss-Seidel sweep count (0: the
// box updates do not apply)
int box_sweeps(const ibpm_solver* s) {
  if (s->box_opt == 1 || s->precision == IBPM_MIXED || s->n == 0) return 0;
  SolveJob j;
  j.b = nullptr;
  j.x = nullptr;
  j.da = 1.0;
  j.dnum = s->dt;
  j.dden = 2.0 * s->Re * (s->hc[0] * s->hc[0]);
  j.epi = EPI_SUB;
  j.base = nullptr;
  j.a = 1.0;
  j.c = cA(s, 0);
  return use_gs(s, j) ? gs_sweeps(j) : 0;
}

// grow-only device buffer (callers are outside graph capture: creation, moving-body steps)
hipError_t ensure_buffer(double** p, size_t* cap, size_t n) {
  if (n <= *cap && *p) return hipSuccess;
  if (*p) {
    hipError_t e = hipFree(*p);
    if (e != hipSuccess) return e;
    *p = nullptr;
  }
  *cap = 0;
  hipError_t e = dalloc(p, std::max<size_t>(n, 1));
  if (e == hipSuccess) *cap = std::max<size_t>(n, 1);
  return e;
}

// Boxes for panel cells whose rot reaches vertices [ilo, ihi] x [jlo, jhi]; sets box_on
// and pbox_on by the cost rule (or box_opt 2) and sizes the scratch.
int compute_boxes(ibpm_solver* s, int ilo, int ihi, int jlo, int jhi) {
  s->box_on = s->pbox_on = false;
  const int m = box_sweeps(s);
  if (m == 0) return IBPM_OK;
  const int Nx = s->g.Nx, Ny = s->g.Ny, L = s->L;
  // 1-based vertex boxes per level
  int lo[2] = {ilo - m - 1, jlo - m - 1}, hi[2] = {ihi + m + 1, jhi + m + 1};
  const CoarsifyRange cr = coarsify_range(s->g.nx, s->g.ny);
  const int c0[2] = {cr.ic0, cr.jc0}, cn[2] = {cr.ni, cr.nj};
  auto floor_div2 = [](int v) { return v >= 0 ? v / 2 : -((1 - v) / 2); };
  double cost = 0.0;
  size_t toff = 0;
  for (int l = 0; l < L; ++l) {
    if (l > 0) {
      for (int d = 0; d < 2; ++d) {
        int alo = floor_div2(lo[d] - 3), ahi = floor_div2(hi[d] - 1) + 1;
        alo = std::max(alo, 0);
        ahi = std::min(ahi, cn[d] - 1);
        if (alo > ahi) {
          lo[d] = 1, hi[d] = 0;  // empty
        } else {
          lo[d] = c0[d] + alo - 1;
          hi[d] = c0[d] + ahi + 1;
        }
      }
    }
    const int i0 = std::max(lo[0], 1), i1 = std::min(hi[0], Nx);
    const int j0 = std::max(lo[1], 1), j1 = std::min(hi[1], Ny);
    s->bi0[l] = i0 - 1;
    s->bj0[l] = j0 - 1;
    s->bw[l] = std::max(0, i1 - i0 + 1);
    s->bh[l] = std::max(0, j1 - j0 + 1);
    if (s->bw[l] == 0 || s->bh[l] == 0) {
      s->bw[l] = s->bh[l] = 0;
      continue;
    }
    const double bw = s->bw[l], bh = s->bh[l];
    const double cx = (double)Nx * bw * bh + (double)Nx * Ny * bh;  // x contraction first
    const double cy = bw * (double)Ny * bh + (double)Nx * Ny * bw;
    s->bxfirst[l] = cx <= cy ? 1 : 0;
    cost += std::min(cx, cy);
    s->btoff[l] = toff;
    toff += s->bxfirst[l] ? (size_t)Nx * s->bh[l] : (size_t)s->bw[l] * Ny;
  }
  // A first product whose contraction (bw, or bh) exceeds 2 kBoxSliceK runs as K-slices
  // of about kBoxSliceK in separate work-groups, summed in slice order: a long K in few
  // tiles (a wide box of moving bodies) is otherwise a chain of sequential chunks.
  // At most kMaxPlanes slice jobs per launch: the longest levels give up slices first.
  int ns_total = 0;
  for (int l = 0; l < L; ++l) {
    const int K1 = s->bxfirst[l] ? s->bw[l] : s->bh[l];
    s->bns[l] = K1 > 2 * kBoxSliceK ? std::min(16, (K1 + kBoxSliceK - 1) / kBoxSliceK) : (K1 > 0 ? 1 : 0);
    ns_total += s->bns[l];
  }
  while (ns_total > kMaxPlanes) {
    int lm = 0;
    for (int l = 1; l < L; ++l)
      if (s->bns[l] > s->bns[lm]) lm = l;
    if (s->bns[lm] <= 1) break;
    --s->bns[lm];
    --ns_total;
  }
  size_t poff = 0;
  for (int l = 0; l < L; ++l) {
    s->bpoff[l] = poff;
    if (s->bns[l] > 1)
      poff += (size_t)s->bns[l] * (s->bxfirst[l] ? (size_t)Nx * s->bh[l] : (size_t)s->bw[l] * Ny);
  }
  // the forward pass pair it replaces: about Nx Ny (Nx + Ny) / 4 multiply-adds per plane
  // with the twice-folded sine matrices; take the box update when it needs at most half
  const double passes = (double)L * Nx * Ny * (double)(Nx + Ny) / 4.0;
  if (s->box_opt != 2 && !(cost <= 0.5 * passes)) return IBPM_OK;
  const size_t P = s->plane();
  hipError_t e = s->Ftr ? hipSuccess : dalloc(&s->Ftr, (size_t)L * P);
  if (!e) e = ensure_buffer(&s->Tbox, &s->tbox_cap, toff);
  if (!e && poff) e = ensure_buffer(&s->Tpart, &s->tpart_cap, poff);
  auto sine_matrix = [](int Nn, double** out) -> hipError_t {
    const long long n = Nn + 1;
    std::vector<double> tab(2 * n), S((size_t)Nn * Nn);
    for (long long a = 0; a < 2 * n; ++a) tab[a] = 2.0 * sin(M_PI * (double)a / (double)n);
    for (long long c = 1; c <= Nn; ++c)
      for (long long r = 1; r <= Nn; ++r) S[(size_t)(r - 1) + (size_t)(c - 1) * Nn] = tab[(r * c) % (2 * n)];
    hipError_t r = dalloc(out, S.size());
    if (!r) r = hipMemcpy(*out, S.data(), S.size() * 8, hipMemcpyHostToDevice);
    return r;
  };
  if (!e && !s->Sfx) e = sine_matrix(Nx, &s->Sfx);
  if (!e && !s->Sfy) {
    if (Ny == Nx) s->Sfy = s->Sfx;
    else e = sine_matrix(Ny, &s->Sfy);
  }
  if (e) return fail(IBPM_E_HIP, hipGetErrorString(e));
  s->box_on = true;
  // the psi box: gather_kernel reads faces [ix - 3, ix + 3] x [iy - 3, iy + 3] of each
  // panel, whose curl reads psi on vertices [ix - 4, ix + 3] x [iy - 4, iy + 3], i.e.
  // [ilo, ihi] x [jlo, jhi]; with those and one more vertex inside the level (and so no
  // face on its first row or column) no boundary value, hence no coarser level, is read.
  const int nx = s->g.nx, ny = s->g.ny;
  if (ilo >= 2 && ihi <= nx - 2 && jlo >= 2 && jhi <= ny - 2) {
    const int ia = ilo - 1, ib = ihi + 1, ja = jlo - 1, jb = jhi + 1;  // vertices read +- 1
    s->pi0 = ia - 1;
    s->pbw = ib - ia + 1;
    s->pj0 = ja - 1;
    s->pbh = jb - ja + 1;
    // E S_y[:, J] and S_x[I, :] (.) against the inverse pass pairs over levels 1 and 2
    const double pcost = (double)Nx * Ny * s->pbh + (double)s->pbw * s->pbh * Nx;
    if (s->box_opt == 2 || pcost <= 0.5 * std::min(2, L) * (double)Nx * Ny * (double)(Nx + Ny) / 4.0) {
      e = ensure_buffer(&s->Pa, &s->pa_cap, (size_t)kPsiSlicesA * Nx * s->pbh);
      if (!e) e = ensure_buffer(&s->T2, &s->t2_cap, (size_t)Nx * s->pbh);
      if (!e) e = ensure_buffer(&s->Pb, &s->pb_cap, (size_t)kPsiSlicesB * s->pbw * s->pbh);
      if (e) return fail(IBPM_E_HIP, hipGetErrorString(e));
      s->pbox_on = true;
    }
  }
  return IBPM_OK;
}

// Static bodies: the boxes once, from the panel positions (the cells as body_update_kernel
// computes them). Moving bodies: box_dyn, boxes every step (update_moving_boxes).
int setup_box_update(ibpm_solver* s, const ibpm_problem* p) {
  s->box_on = s->pbox_on = s->box_dyn = false;
  if (box_sweeps(s) == 0) return IBPM_OK;
  if (s->moving) {
    HIPCHK(dalloc(&s->bbox_d, 4));
    s->box_dyn = true;
    return IBPM_OK;
  }
  int ilo = 1 << 30, ihi = -(1 << 30), jlo = 1 << 30, jhi = -(1 << 30);
  for (int k = 0; k < s->n; ++k) {
    const int ix = (int)rint((p->xb[k] - s->x0) / s->h + 1.0 - 0.25);
    const int iy = (int)rint((p->xb[s->n + k] - s->y0) / s->h + 1.0 - 0.25);
    ilo = std::min(ilo, ix - 4);
    ihi = std::max(ihi, ix + 3);
    jlo = std::min(jlo, iy - 4);
    jhi = std::max(jhi, iy + 3);
  }
  return compute_boxes(s, ilo, ihi, jlo, jhi);
}

// Moving bodies: the boxes of this step from the range of the panels' cells (after
// body_update_kernel); one small device-to-host read.
int update_moving_boxes(ibpm_solver* s) {
  if (!s->box_dyn) return IBPM_OK;
  HIPCHK(launch_body_bbox(s->coords, s->n, s->bbox_d, s->st));
  HIPCHK(hipMemcpyAsync(s->h_scal + 12, s->bbox_d, 4 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  const int ixmin = (int)s->h_scal[12], ixmax = (int)s->h_scal[13];
  const int iymin = (int)s->h_scal[14], iymax = (int)s->h_scal[15];
  return compute_boxes(s, ixmin - 4, ixmax + 3, iymin - 4, iymax + 3);
}

// Trial psi of level 1 on the box [pi0, pi0 + pbw) x [pj0, pj0 + pbh) from the level's
// ring-chain spectral plane E: psi[I, J] = S_x[I, :] (E S_y[:, J]). Both products contract
// a full length (Ny, then Nx) into a small output, so K is split into slices computed by
// separate work-groups and summed in slice order (deterministic).
int psi_box_product(ibpm_solver* s, const double* E, double* psi) {
  const int Nx = s->g.Nx, Ny = s->g.Ny, bw = s->pbw, bh = s->pbh;
  GemmBatch g;
  memset(&g, 0, sizeof g);
  // the first product folded over its contraction: S_y[Ny-1-k, j] = (-1)^j S_y[k, j], so
  // column j of E S_y[:, J] is sum over k < (Ny+1)/2 of (E[:, k] + (-1)^j E[:, Ny-1-k])
  // S_y[k, j] (the middle column of an odd Ny once): half the contraction, one job per
  // K slice and parity of j (every other column of B and of the slice's output)
  const int Kh = (Ny + 1) / 2, mid = (Ny & 1) ? (Ny - 1) / 2 : -1;
  const int ka = ((Kh + kPsiFoldSlices - 1) / kPsiFoldSlices + 31) / 32 * 32;
  int na = 0, nj = 0;
  for (int k0 = 0; k0 < Kh && na < kPsiFoldSlices; k0 += ka, ++na) {
    const int K = std::min(ka, Kh - k0);
    for (int p = 0; p < 2; ++p) {
      const int np0 = ((p - s->pj0) % 2 + 2) % 2;  // first box column of parity p
      if (np0 >= bh) continue;
      GemmJob& j = g.j[nj++];
      j.A = E + (size_t)k0 * Nx, j.lda = Nx;
      j.A2 = E + (size_t)(Ny - 1 - k0) * Nx;
      j.amirror = 1, j.asign = p == 0 ? 1.0 : -1.0;
      j.amid = mid >= k0 && mid < k0 + K ? mid - k0 : -1;
      j.B = s->Sfy + k0 + (size_t)(s->pj0 + np0) * Ny, j.ldb = 2 * Ny;
      j.C = s->Pa + (size_t)na * Nx * bh + (size_t)np0 * Nx, j.ldc = 2 * Nx;
      j.M = Nx, j.N = (bh - np0 + 1) / 2, j.K = K;
    }
  }
  HIPCHK(launch_gemm_f64(g, nj, s->st));
  HIPCHK(launch_sum_slices(s->Pa, (size_t)Nx * bh, na, Nx, Nx, bh, s->T2, Nx, s->st));
  memset(&g, 0, sizeof g);
  const int kb = ((Nx + kPsiSlicesB - 1) / kPsiSlicesB + 31) / 32 * 32;
  int nb = 0;
  for (int k0 = 0; k0 < Nx && nb < kPsiSlicesB; k0 += kb, ++nb) {
    GemmJob& j = g.j[nb];
    j.A = s->Sfx + s->pi0 + (size_t)k0 * Nx, j.lda = Nx;
    j.B = s->T2 + k0, j.ldb = Nx;
    j.C = s->Pb + (size_t)nb * bw * bh, j.ldc = bw;
    j.M = bw, j.N = bh, j.K = std::min(kb, Nx - k0);
  }
  HIPCHK(launch_gemm_f64(g, nb, s->st));
  HIPCHK(launch_sum_slices(s->Pb, (size_t)bw * bh, nb, bw, bw, bh,
                           psi + s->pi0 + (size_t)s->pj0 * Nx, Nx, s->st));
  return IBPM_OK;
}

// B operator (coupling.jl:150-177) on ncol columns: x_c = E vort2flux(Ainv1 rot(E^T z_c))[1]
int apply_B_box(ibpm_solver* s, int ncol, const double* const* z, double* const* x);

int apply_B(ibpm_solver* s, int ncol, const double* const* z, double* const* x) {
  if (s->box_on && s->pbox_on) return apply_B_box(s, ncol, z, x);
  const size_t P = s->plane();
  const int L = s->L;
  for (int c0 = 0; c0 < ncol; c0 += 64) {
    const int nc = std::min(64, ncol - c0);
    ScatterArgs sa;
    memset(&sa, 0, sizeof sa);
    sa.g = s->g;
    sa.npanel = s->n;
    sa.nrhs = nc;
    sa.coords = s->coords;
    sa.w = s->w;
    sa.perm = s->perm;
    sa.xstart = s->xstart;
    sa.ywin = s->ywin;
    sa.rowptr = s->sc_rowptr;
    sa.ep = s->sc_ep;
    sa.ew = s->sc_ew;
    for (int c = 0; c < nc; ++c) {
      sa.f[c] = z[c0 + c];
      sa.out[c] = s->Sc + (size_t)(c0 + c) * P;
    }
    HIPCHK(launch_scatter_rot(sa, s->st));
  }
  std::vector<SolveJob> jobs(ncol);
  for (int c = 0; c < ncol; ++c) {
    SolveJob& j = jobs[c];
    j.b = s->Sc + (size_t)c * P;
    j.da = 1.0;
    j.dnum = s->dt;
    j.dden = 2.0 * s->Re * (s->hc[0] * s->hc[0]);
    j.x = s->Gtmp + (size_t)c * L * P;
    j.epi = EPI_STORE;
    j.base = nullptr;
    j.a = 1.0;
    j.c = cA(s, 0);
  }
  int rc = solve(s, jobs.data(), ncol);
  if (rc) return rc;
  std::vector<double*> Gc(ncol), pc(ncol);
  for (int c = 0; c < ncol; ++c) {
    Gc[c] = s->Gtmp + (size_t)c * L * P;
    pc[c] = s->psitmp + (size_t)c * L * P;
  }
  rc = vort2flux(s, ncol, Gc.data(), pc.data(), nullptr, nullptr, std::min(2, L));  // E reads levels 1, 2
  if (rc) return rc;
  for (int c0 = 0; c0 < ncol; c0 += 64) {
    const int nc = std::min(64, ncol - c0);
    GatherArgs ga;
    memset(&ga, 0, sizeof ga);
    ga.g = s->g;
    ga.npanel = s->n;
    ga.nrhs = nc;
    ga.coords = s->coords;
    ga.w = s->w;
    for (int c = 0; c < nc; ++c) {
      ga.psi1[c] = pc[c0 + c];
      ga.psi2[c] = L > 1 ? pc[c0 + c] + P : nullptr;
      ga.out[c] = x[c0 + c];
    }
    memset(&ga.bf, 0, sizeof ga.bf);  // B has no base flux
    ga.ub = nullptr;
    ga.h = s->h;
    HIPCHK(launch_gather(ga, s->st));
  }
  return IBPM_OK;
}

// B by the box updates: Ainv1 rot(E^T z) vanishes outside the level-1 box and coarsify!
// carries it to the coarser boxes (the rest of every level is zero), so rot and the
// Gauss-Seidel solve run on the box, each level's forward transform is a box product, the
// ring chain runs as always, and E reads psi_1 on the psi box only.
int apply_B_box(ibpm_solver* s, int ncol, const double* const* z, double* const* x) {
  const size_t P = s->plane();
  const int L = s->L;
  const int r[4] = {s->bi0[0], s->bw[0], s->bj0[0], s->bh[0]};
  if (r[0] != s->gtmp_box[0] || r[1] != s->gtmp_box[1] || r[2] != s->gtmp_box[2] ||
      r[3] != s->gtmp_box[3] || ncol > s->gtmp_cols) {
    // level 1 of the columns is written on the box only: clear what an earlier box left
    for (int c = 0; c < s->bcols; ++c) HIPCHK(launch_fill(s->Gtmp + (size_t)c * L * P, P, 0.0, s->st));
    for (int k = 0; k < 4; ++k) s->gtmp_box[k] = r[k];
    s->gtmp_cols = s->bcols;
  }
  for (int c0 = 0; c0 < ncol; c0 += 64) {
    const int nc = std::min(64, ncol - c0);
    ScatterArgs sa;
    memset(&sa, 0, sizeof sa);
    sa.g = s->g;
    sa.npanel = s->n;
    sa.nrhs = nc;
    sa.coords = s->coords;
    sa.w = s->w;
    sa.perm = s->perm;
    sa.xstart = s->xstart;
    sa.ywin = s->ywin;
    sa.rowptr = s->sc_rowptr;
    sa.ep = s->sc_ep;
    sa.ew = s->sc_ew;
    sa.ri0 = r[0], sa.rw = r[1], sa.rj0 = r[2], sa.rh = r[3];
    for (int c = 0; c < nc; ++c) {
      sa.f[c] = z[c0 + c];
      sa.out[c] = s->Sc + (size_t)(c0 + c) * P;
    }
    HIPCHK(launch_scatter_rot(sa, s->st));
  }
  std::vector<SolveJob> jobs(ncol);
  for (int c = 0; c < ncol; ++c) {
    SolveJob& j = jobs[c];
    j.b = s->Sc + (size_t)c * P;
    j.da = 1.0;
    j.dnum = s->dt;
    j.dden = 2.0 * s->Re * (s->hc[0] * s->hc[0]);
    j.x = s->Gtmp + (size_t)c * L * P;
    j.epi = EPI_STORE;
    j.base = nullptr;
    j.a = 1.0;
    j.c = cA(s, 0);
    j.ri0 = r[0], j.rw = r[1], j.rj0 = r[2], j.rh = r[3];
  }
  int rc = solve(s, jobs.data(), ncol);
  if (rc) return rc;
  for (int c = 0; c < ncol; ++c)
    for (int l = 1; l < L; ++l)
      HIPCHK(launch_coarsify(s->g, s->Gtmp + (size_t)c * L * P + l * P,
                             s->Gtmp + (size_t)c * L * P + (l - 1) * P, nullptr, s->st));
  const int gmax = std::max(1, kMaxPlanes / L);
  for (int c0 = 0; c0 < ncol; c0 += gmax) {
    const int ng = std::min(gmax, ncol - c0);
    const double* sin[kMaxPlanes] = {};
    double* out[kMaxPlanes];
    const void* in[kMaxPlanes];
    for (int c = 0; c < ng; ++c) {
      // forward transforms of the column's levels into U: box products from zero
      rc = box_update(s, s->Gtmp + (size_t)(c0 + c) * L * P, nullptr, s->U + (size_t)c * L * P, false);
      if (rc) return rc;
      for (int l = 0; l < L; ++l) {
        sin[c * L + l] = s->U + (size_t)(c * L + l) * P;
        in[c * L + l] = nullptr;
        out[c * L + l] = s->psitmp + (size_t)(c0 + c) * L * P + l * P;
      }
    }
    rc = v2f_group(s, ng, in, false, out, EPI_STORE, 1, sin, nullptr, true);
    if (rc) return rc;
  }
  for (int c0 = 0; c0 < ncol; c0 += 64) {
    const int nc = std::min(64, ncol - c0);
    GatherArgs ga;
    memset(&ga, 0, sizeof ga);
    ga.g = s->g;
    ga.npanel = s->n;
    ga.nrhs = nc;
    ga.coords = s->coords;
    ga.w = s->w;
    for (int c = 0; c < nc; ++c) {
      ga.psi1[c] = s->psitmp + (size_t)(c0 + c) * L * P;
      ga.psi2[c] = nullptr;  // the psi box keeps every read inside level 1
      ga.out[c] = x[c0 + c];
    }
    memset(&ga.bf, 0, sizeof ga.bf);  // B has no base flux
    ga.ub = nullptr;
    ga.h = s->h;
    HIPCHK(launch_gather(ga, s->st));
  }
  return IBPM_OK;
}

int alloc_bscratch(ibpm_solver* s, int ncol) {
  const size_t P = s->plane();
  if (s->Gtmp) (void)hipFree(s->Gtmp);
  if (s->psitmp) (void)hipFree(s->psitmp);
  if (s->Sc) (void)hipFree(s->Sc);
  s->Gtmp = s->psitmp = s->Sc = nullptr;
  s->bcols = ncol;
  if (ncol == 0) return IBPM_OK;
  HIPCHK(dalloc(&s->Gtmp, (size_t)ncol * s->L * P));
  HIPCHK(dalloc(&s->psitmp, (size_t)ncol * s->L * P));
  HIPCHK(dalloc(&s->Sc, (size_t)ncol * P));
  // coarse levels of Gtmp stay zero outside the centre region (Gamma_tmp .= 0)
  HIPCHK(hipMemsetAsync(s->Gtmp, 0, (size_t)ncol * s->L * P * sizeof(double), s->st));
  s->gtmp_box[0] = s->gtmp_box[1] = s->gtmp_box[2] = s->gtmp_box[3] = -1;
  s->gtmp_cols = 0;
  return IBPM_OK;
}

// B of the current body configuration, column by column (apply_B on unit vectors in
// groups of 32 columns), then its inverse on the device (Gauss-Jordan with partial
// pivoting, the role of inv() in coupling.jl:190) into `inv` (2n x 2n, column-major,
// allocated here when null). Leaves the B-operator scratch sized for one column.
int assemble_B_inverse(ibpm_solver* s, double*& inv) {
  const int m = 2 * s->n;
  const int ncol = std::min(m, 32);
  int rc = alloc_bscratch(s, ncol);
  if (rc) return rc;
  double *Bm = nullptr, *E = nullptr, *rowk = nullptr, *pval = nullptr;
  int* ip = nullptr;
  hipError_t e = hipMalloc((void**)&Bm, (size_t)m * m * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&E, (size_t)m * ncol * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&rowk, 2 * (size_t)m * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&pval, sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&ip, 2 * sizeof(int));
  if (e == hipSuccess && !inv) e = hipMalloc((void**)&inv, (size_t)m * m * sizeof(double));
  std::vector<double> eh((size_t)m * ncol);
  for (int c0 = 0; c0 < m && e == hipSuccess && rc == IBPM_OK; c0 += ncol) {
    const int nc = std::min(ncol, m - c0);
    std::fill(eh.begin(), eh.end(), 0.0);
    for (int c = 0; c < nc; ++c) eh[(size_t)c * m + c0 + c] = 1.0;
    e = hipMemcpyAsync(E, eh.data(), eh.size() * 8, hipMemcpyHostToDevice, s->st);
    std::vector<const double*> z(nc);
    std::vector<double*> x(nc);
    for (int c = 0; c < nc; ++c) {
      z[c] = E + (size_t)c * m;
      x[c] = Bm + (size_t)(c0 + c) * m;
    }
    if (e == hipSuccess) rc = apply_B(s, nc, z.data(), x.data());
    if (e == hipSuccess && rc == IBPM_OK) e = hipStreamSynchronize(s->st);  // eh is reused
  }
  int singular = 0;
  if (e == hipSuccess && rc == IBPM_OK)
    e = launch_gauss_jordan_inverse(m, Bm, inv, rowk, ip, pval, ip + 1, s->st);
  if (e == hipSuccess && rc == IBPM_OK)
    e = hipMemcpyAsync(&singular, ip + 1, sizeof(int), hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess && rc == IBPM_OK) e = hipStreamSynchronize(s->st);
  for (void* p : {(void*)Bm, (void*)E, (void*)rowk, (void*)pval, (void*)ip})
    if (p) (void)hipFree(p);
  if (rc) return rc;
  if (e != hipSuccess) return fail(IBPM_E_HIP, std::string("assemble_B_inverse: ") + hipGetErrorString(e));
  if (singular) return fail(IBPM_E_SINGULAR, "B is singular");
  return alloc_bscratch(s, 1);
}

// Preconditioner of the moving-body CG: the inverse of B at the current configuration and
// each body's angle now (the CG applies it rotated by the angle changes since).
int assemble_precond(ibpm_solver* s) {
  int rc = assemble_B_inverse(s, s->pc_P);
  if (rc) return rc;
  HIPCHK(hipMemcpy(s->pc_theta0, s->cur_theta.data(), s->nbody * sizeof(double), hipMemcpyHostToDevice));
  s->pc_ready = true;
  s->pc_due = false;
  return IBPM_OK;
}

int check_body_error(ibpm_solver* s) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, s->err, sizeof(int), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  if (e) {
    const int p = s->n - e;  // 0-based panel
    double px = 0, py = 0;
    (void)hipMemcpy(&px, s->xb + p, sizeof(double), hipMemcpyDeviceToHost);
    (void)hipMemcpy(&py, s->xb + s->n + p, sizeof(double), hipMemcpyDeviceToHost);
    int zero = 0;
    (void)hipMemcpy(s->err, &zero, sizeof(int), hipMemcpyHostToDevice);
    char msg[256];
    snprintf(msg, sizeof msg, "Panel index=%d at (%.17g, %.17g) outside allowed region", p + 1, px, py);
    return fail(IBPM_E_OUTSIDE_REGION, msg);
  }
  return IBPM_OK;
}

BodyDev body_dev(ibpm_solver* s) {
  BodyDev b;
  b.npanel = s->n;
  b.nbody = s->nbody;
  b.x0 = s->x0;
  b.y0 = s->y0;
  b.h = s->h;
  b.nx = s->g.nx;
  b.ny = s->g.ny;
  b.xb = s->xb;
  b.ub = s->ub;
  b.xb_ref = s->xb_ref;
  b.body_of = s->body_of;
  b.moving = s->moving_d;
  b.motion = s->motion;
  b.coords = s->coords;
  b.w = s->w;
  b.perm = s->perm;
  b.xstart = s->xstart;
  b.ywin = s->ywin;
  b.err = s->err;
  return b;
}

CgDev cg_dev(ibpm_solver* s) {
  CgDev c;
  memset(&c, 0, sizeof c);
  c.n = 2 * s->n;
  c.x = s->Ft;
  c.r = s->cg_r;
  c.u = s->cg_u;
  c.c = s->cg_c;
  c.b = s->Ftmp;
  c.scal = s->cg_scal;
  c.maxiter = s->cg_maxiter;
  c.reltol = s->cg_reltol;
  c.abstol = s->cg_abstol;
  c.tol_mode = s->cg_tol_mode;
  return c;
}

PcDev pc_dev(ibpm_solver* s) {
  PcDev pd;
  memset(&pd, 0, sizeof pd);
  pd.n = s->n;
  pd.P = s->pc_P;
  pd.body_of = s->body_of;
  pd.motion = s->motion;
  pd.theta0 = s->pc_theta0;
  return pd;
}

// Host-driven CG iterations (cg!, IterativeSolvers' CGIterable / PCGIterable order) from
// the state in cg_* with its scalars in h_scal, until the stopping test (or maxiter);
// `start` begins the preconditioned search direction first.
int cg_iterate(ibpm_solver* s, bool start) {
  const CgDev c = cg_dev(s);
  double* xc[1] = {s->cg_c};
  const double* zu[1] = {s->cg_u};
  const PcDev pd = pc_dev(s);
  if (s->cg_precond && start && s->h_scal[4] == 0.0) HIPCHK(launch_pcg_start(c, s->st));
  while (s->h_scal[4] == 0.0) {
    if (s->cg_precond) HIPCHK(launch_pcg_precond_dir(c, pd, s->pc_t, s->pc_z, s->gemv_part, s->st));
    int rc = apply_B(s, 1, zu, xc);
    if (rc) return rc;
    HIPCHK(s->cg_precond ? launch_pcg_step(c, s->st) : launch_cg_step(c, s->st));
    HIPCHK(hipMemcpyAsync(s->h_scal, s->cg_scal, 6 * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
  }
  return IBPM_OK;
}

// Continue a CG whose first iterations a captured step ran (a batch's CG budget) until its
// stopping test: the iterates the uncaptured step would have reached.
int cg_continue(ibpm_solver* s) {
  HIPCHK(hipMemcpyAsync(s->h_scal, s->cg_scal, 6 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return cg_iterate(s, false);
}

// couple_surface_rigid (coupling.jl:233-251) incl. the CG of moving bodies
int couple_surface(ibpm_solver* s, ibpm_step_stats* stats) {
  const size_t P = s->plane();
  GatherArgs ga;
  memset(&ga, 0, sizeof ga);
  ga.g = s->g;
  ga.npanel = s->n;
  ga.nrhs = 1;
  ga.coords = s->coords;
  ga.w = s->w;
  ga.psi1[0] = s->psi;
  ga.psi2[0] = s->L > 1 ? s->psi + P : nullptr;
  ga.out[0] = s->Ftmp;
  ga.bf = base_flux(s);
  ga.ub = s->ub;
  ga.h = s->h;
  HIPCHK(launch_gather(ga, s->st));
  if (!s->moving) {
    HIPCHK(launch_gemv_rows(2 * s->n, s->Binv, s->Ftmp, s->Ft, s->st));
    if (stats) {
      stats->cg_iters = 0;
      stats->cg_converged = 1;
      stats->cg_residual = 0.0;
    }
    return IBPM_OK;
  }
  // cg! (coupling.jl:194-202, solver.jl:137). The reference starts from the previous F~b;
  // the warm-start modes only change that initial guess (ibpm_set_warm_start).
  const int n2 = 2 * s->n;
  const CgDev c = cg_dev(s);
  double* xc[1] = {s->cg_c};
  int rc = IBPM_OK;
  int bapp = 0;
  const bool use_ws = s->ws_mode != IBPM_WS_PREVIOUS || s->ws_record;
  const bool need_res = s->ws_mode == IBPM_WS_LEARNED || s->ws_record;
  WsArgs wa;
  memset(&wa, 0, sizeof wa);
  if (use_ws) {
    // x_e from F~b of the last three steps; then shift the history
    HIPCHK(launch_ws_extrap(n2, s->Ft, s->ws_h0, s->ws_h1, s->nhist, s->ws_xe, s->st));
    std::swap(s->ws_h0, s->ws_h1);
    HIPCHK(launch_copy(s->ws_h0, s->Ft, n2, s->st));
    s->nhist = std::min(s->nhist + 1, 3);
    if (need_res) {
      const double* ze[1] = {s->ws_xe};
      double* xe[1] = {s->ws_Bxe};
      rc = apply_B(s, 1, ze, xe);
      if (rc) return rc;
      ++bapp;
      wa.n = s->n;
      wa.b = s->Ftmp;
      wa.Bxe = s->ws_Bxe;
      wa.xb = s->xb;
      wa.ub = s->ub;
      wa.body_of = s->body_of;
      wa.x0 = s->x0;
      wa.y0 = s->y0;
      wa.h = s->h;
      wa.xe = s->ws_xe;
      wa.F = s->Ft;
      wa.scal = s->ws_scal;
      rc = net_features(s->net, wa, s->st);
      if (rc) return rc;
    }
    if (s->ws_mode == IBPM_WS_LEARNED) {
      rc = net_predict(s->net, wa, s->Ft, s->st);
      if (rc) return rc;
    } else if (s->ws_mode == IBPM_WS_EXTRAPOLATE) {
      HIPCHK(launch_copy(s->Ft, s->ws_xe, n2, s->st));
    }
  }
  int pc_assembled = 0;
  if (s->cg_precond && (!s->pc_ready || s->pc_due)) {
    rc = assemble_precond(s);
    if (rc) return rc;
    pc_assembled = 1;
  }
  const double* zc[1] = {s->Ft};
  rc = apply_B(s, 1, zc, xc);
  if (rc) return rc;
  ++bapp;
  HIPCHK(launch_cg_init(c, s->st));
  if (s->graph_mode) {
    // A step captured into a batch graph: the same iterations, decided on the device.
    // The learned guess is kept when it beats x_e (cg_scal[7] = 1), else CG restarts from
    // x_e with the B x_e already known; then cg_maxiter iterations are queued, each of
    // whose launches returns at once after the stopping test has been met (cg_scal[4]),
    // so the iterates are those of the host-driven loop below, maxiter included.
    HIPCHK(launch_cg_note(s->cg_scal, s->ws_mode == IBPM_WS_LEARNED ? s->ws_scal : nullptr, s->st));
    if (s->ws_mode == IBPM_WS_LEARNED) {
      SkipScope keep(s->cg_scal + 7);
      HIPCHK(launch_copy(s->Ft, s->ws_xe, n2, s->st));
      HIPCHK(launch_copy(s->cg_c, s->ws_Bxe, n2, s->st));
      HIPCHK(launch_cg_init(c, s->st));
    }
    SkipScope converged(s->cg_scal + 4);
    const PcDev pd = pc_dev(s);
    if (s->cg_precond) HIPCHK(launch_pcg_start(c, s->st));
    const double* zu[1] = {s->cg_u};
    const int K = s->graph_cg_budget > 0 ? std::min(s->graph_cg_budget, s->cg_maxiter)
                                         : s->cg_maxiter;
    for (int it = 0; it < K; ++it) {
      if (s->cg_precond) HIPCHK(launch_pcg_precond_dir(c, pd, s->pc_t, s->pc_z, s->gemv_part, s->st));
      rc = apply_B(s, 1, zu, xc);
      if (rc) return rc;
      HIPCHK(s->cg_precond ? launch_pcg_step(c, s->st) : launch_cg_step(c, s->st));
    }
    return IBPM_OK;
  }
  HIPCHK(hipMemcpyAsync(s->h_scal, s->cg_scal, 6 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  if (need_res)
    HIPCHK(hipMemcpyAsync(s->h_scal + 8, s->ws_scal, 3 * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  rc = check_body_error(s);
  if (rc) return rc;
  int ws_used = 0;
  double base_res = need_res ? s->h_scal[9] : 0.0;
  if (s->ws_mode == IBPM_WS_LEARNED) {
    // keep the network's guess only when it beats x_e; otherwise restart from x_e, whose
    // B x_e is already known
    ws_used = s->h_scal[0] <= base_res ? 1 : 0;
    if (!ws_used) {
      HIPCHK(launch_copy(s->Ft, s->ws_xe, n2, s->st));
      HIPCHK(launch_copy(s->cg_c, s->ws_Bxe, n2, s->st));
      HIPCHK(launch_cg_init(c, s->st));
      HIPCHK(hipMemcpyAsync(s->h_scal, s->cg_scal, 6 * sizeof(double), hipMemcpyDeviceToHost, s->st));
      HIPCHK(hipStreamSynchronize(s->st));
    }
  }
  const double bnorm = s->h_scal[5];
  const double res0 = s->h_scal[0];
  rc = cg_iterate(s, true);
  if (rc) return rc;
  if (s->cg_precond && s->cg_precond_refresh > 0 && s->h_scal[3] > s->cg_precond_refresh)
    s->pc_due = true;
  if (s->ws_record) {
    rc = net_record(s->net, wa, s->st);
    if (rc) return rc;
  }
  if (stats) {
    stats->cg_iters = (int)s->h_scal[3];
    stats->cg_residual = s->h_scal[0];
    stats->cg_converged = s->h_scal[0] <= s->h_scal[2] ? 1 : 0;
    stats->b_applies = bapp + stats->cg_iters;
    stats->ws_used = ws_used;
    stats->ws_base_residual = bnorm > 0.0 ? base_res / bnorm : 0.0;
    stats->ws_residual = bnorm > 0.0 ? res0 / bnorm : 0.0;
    stats->pc_assembled = pc_assembled;
  }
  return IBPM_OK;
}

// Level 1's spectral plane S Gamma_s S is kept when the sine transforms solve level 1
// (FP64); both vort2flux calls of the step then skip its forward passes.
bool keeps_spec1(const ibpm_solver* s) {
  SolveJob j;
  j.b = nullptr;
  j.da = 1.0;
  j.dnum = s->dt;
  j.dden = 2.0 * s->Re * (s->hc[0] * s->hc[0]);
  j.x = nullptr;
  j.epi = EPI_STORE;
  j.base = nullptr;
  j.a = 1.0;
  j.c = cA(s, 0);
  return s->F1 && !use_gs(s, j);
}

// advance! up to and including couple_surface (its CG); step_tail does the rest.
int step_head(ibpm_solver* s, const double* fsv, const double* motion, ibpm_step_stats* stats) {
  const size_t P = s->plane();
  const int L = s->L;
  // inc_time_index!, freestream (solver.jl:164-167)
  s->index += 1;
  s->fs[0] = fsv ? fsv[0] : 0.0;
  s->fs[1] = fsv ? fsv[1] : 0.0;
  if (s->grid_moving)  // MovingGrid: the grid's GridVelocity at the new time
    for (int k = 0; k < 6; ++k) s->gm[k] = fsv ? fsv[2 + k] : 0.0;
  s->q0_shadow_valid = false;
  // _update_body_motion! (bodies.jl:43-51)
  if (s->moving) {
    // a step captured into a batch graph finds this step's motion on the device already
    if (motion) {
      for (int b = 0; b < s->nbody; ++b) s->cur_theta[b] = motion[6 * b + 4];
      HIPCHK(hipMemcpyAsync(s->motion, motion, 6 * s->nbody * sizeof(double), hipMemcpyHostToDevice, s->st));
    } else if (!s->graph_mode) {
      return fail(IBPM_E_ARG, "moving bodies need this step's motion");
    }
    BodyDev b = body_dev(s);
    HIPCHK(launch_body_update(b, 1, s->st));
    HIPCHK(launch_body_sort(b, s->st));
    int rc = update_moving_boxes(s);
    if (rc) return rc;
  }
  // get_trial_state (fluid.jl:467-506): nonlinear term and rhs of every level, then the
  // implicit solves of every level in one batch (levels only read the old state)
  NonlinArgs na;
  memset(&na, 0, sizeof na);
  na.g = s->g;
  na.nlevel = L;
  na.q = s->q;
  if (s->q_stale) {
    // a graph replays the step with fluxes from q: q must be written before a capture
    if (s->graph_mode) return fail(IBPM_E_ARG, "internal: q not written before a graph capture");
    na.psi = s->psi;
  }
  na.G = s->G;
  na.nl_new = s->nl1;
  na.nl_old = s->nl2;
  na.rhs = s->rhs;
  na.bf = base_flux(s);
  for (int l = 0; l < L; ++l) {
    const double hc = s->hc[l];
    na.inv_hc2[l] = 1.0 / (hc * hc);
    na.cA[l] = cA(s, l);
    na.fac_bc[l] = 0.25 * s->dt / (s->Re * (hc * hc));
  }
  na.dt = s->dt;
  na.beta1 = 1.5;
  na.beta2 = -0.5;
  if (s->probe == 1) probe_begin(s);
  HIPCHK(launch_nonlin_rhs(na, s->st));
  if (s->probe == 1) probe_end(s);
  std::vector<SolveJob> jobs(L);
  for (int l = 0; l < L; ++l) {
    SolveJob& j = jobs[l];
    j.b = s->rhs + l * P;
    j.da = 1.0;
    j.dnum = s->dt;
    j.dden = 2.0 * s->Re * (s->hc[l] * s->hc[l]);
    j.x = s->G + l * P;  // the trial state Gamma_s in place: Gamma^n is not read after rhs
    j.epi = EPI_STORE;
    j.base = nullptr;
    j.a = 1.0;
    j.c = cA(s, l);
  }
  const bool spec1 = keeps_spec1(s);
  jobs[0].spec = spec1 ? s->F1 : nullptr;
  int rc = solve(s, jobs.data(), L);
  if (rc) return rc;
  std::swap(s->nl1, s->nl2);  // _cycle!
  if (s->n > 0) {
    // trial psi: only couple_surface's E reads it (levels 1 and 2); the final vort2flux
    // below writes every level. Without bodies that final call is the same computation
    // (coarsify! is idempotent), so the trial one is skipped.
    double* Gc[1] = {s->G};
    double* pc[1] = {s->psi};
    rc = vort2flux(s, 1, Gc, pc, nullptr, nullptr, std::min(2, L), jobs[0].spec,
                   s->box_on && !spec1 ? s->Ftr : nullptr, false, s->box_on && !spec1);
    if (rc) return rc;
  }
  // couple_surface! (coupling.jl:233-251)
  if (s->n > 0) return couple_surface(s, stats);
  return IBPM_OK;
}

// advance! after couple_surface: the forces, project_circ! (coupling.jl:217-231):
// Gamma = Gamma_s, Gamma[:,1] -= Ainv1 rot(E^T Ft), and the final vort2flux. Gamma_s is
// the state G itself (the Helmholtz solves wrote it there): level 1 is saved into Gs where
// the solve changes it (its box, or the whole plane), the solve writes Gamma_1 = saved -
// correction, and the final vort2flux re-evaluates coarsify! where that reaches.
int step_tail(ibpm_solver* s, ibpm_step_stats* stats) {
  const size_t P = s->plane();
  const bool spec1 = keeps_spec1(s);
  int rc = IBPM_OK;
  if (s->n > 0) {
    ScatterArgs sa;
    memset(&sa, 0, sizeof sa);
    sa.g = s->g;
    sa.npanel = s->n;
    sa.nrhs = 1;
    sa.coords = s->coords;
    sa.w = s->w;
    sa.perm = s->perm;
    sa.xstart = s->xstart;
    sa.ywin = s->ywin;
    sa.rowptr = s->sc_rowptr;
    sa.ep = s->sc_ep;
    sa.ew = s->sc_ew;
    sa.f[0] = s->Ft;
    sa.out[0] = s->rhs;
    // with the box update, Ainv1 rot(E^T Ft) from zero vanishes outside the level-1 box
    // (the Gauss-Seidel support argument): rot and the solve run on the box only, and
    // Gamma_1 = Gamma_s1 elsewhere (left in place)
    const bool pbox = s->box_on && !spec1;
    if (pbox) sa.ri0 = s->bi0[0], sa.rj0 = s->bj0[0], sa.rw = s->bw[0], sa.rh = s->bh[0];
    if (pbox && sa.rowptr && sa.rw > 0 && sa.rh > 0) {
      // the forces, rot(E^T Ft) on the box and Gamma_s1 saved on the box in one launch
      HIPCHK(launch_project_prologue(sa, s->G, s->Gs, s->fb, s->total, s->h, s->dt,
                                     s->force_slot, s->st));
    } else {
      HIPCHK(launch_forces(s->n, s->Ft, s->fb, s->total, s->h, s->dt, s->force_slot, s->st));
      HIPCHK(launch_copy(s->Gs, s->G, P, s->st));  // Gamma_s1
      HIPCHK(launch_scatter_rot(sa, s->st));
    }
    SolveJob j;
    j.b = s->rhs;
    j.da = 1.0;
    j.dnum = s->dt;
    j.dden = 2.0 * s->Re * (s->hc[0] * s->hc[0]);
    j.x = s->G;
    j.epi = EPI_SUB;
    j.base = s->Gs;
    j.a = 1.0;
    j.c = cA(s, 0);
    j.spec = spec1 ? s->G1 : nullptr;  // the same operator as level 1's solve above
    if (pbox) j.ri0 = s->bi0[0], j.rj0 = s->bj0[0], j.rw = s->bw[0], j.rh = s->bh[0];
    rc = solve(s, &j, 1);
    if (rc) return rc;
    // S Gamma_1 S of the projected circulation: that of Gamma_s minus the correction's
    if (spec1) HIPCHK(launch_sub(s->F1, s->F1, s->G1, P, s->st));
    double* Gc[1] = {s->G};
    double* pc[1] = {s->psi};
    double* sc[1] = {s->Gs};
    double* qc[1] = {s->q};
    // with the box update, only the change of Gamma since the trial vort2flux is
    // transformed (it vanishes outside the boxes; Gs holds Gamma_s there); outside a graph
    // the final curl! is left to the next step's nonlinear kernel (q_stale)
    rc = vort2flux(s, 1, Gc, pc, pbox ? sc : nullptr, s->graph_mode ? qc : nullptr, 0,
                   spec1 ? s->F1 : nullptr, pbox ? s->Ftr : nullptr, pbox);
    if (rc) return rc;
    s->q_stale = !s->graph_mode;
  } else {
    double* Gc[1] = {s->G};  // Gamma = Gamma_s, already in place
    double* pc[1] = {s->psi};
    double* qc[1] = {s->q};
    rc = vort2flux(s, 1, Gc, pc, nullptr, s->graph_mode ? qc : nullptr, 0, spec1 ? s->F1 : nullptr);
    if (rc) return rc;
    s->q_stale = !s->graph_mode;
    if (stats) {
      stats->cg_iters = 0;
      stats->cg_converged = 1;
      stats->cg_residual = 0.0;
    }
  }
  return IBPM_OK;
}

int step(ibpm_solver* s, const double* fsv, const double* motion, ibpm_step_stats* stats) {
  const int rc = step_head(s, fsv, motion, stats);
  return rc ? rc : step_tail(s, stats);
}

}  // namespace

namespace ibpm {
BaseFlux base_flux(const ibpm_solver* s) {
  BaseFlux b;
  memset(&b, 0, sizeof b);
  b.moving = s->grid_moving;
  b.dyn = s->graph_mode ? s->dyn_d : nullptr;
  for (int l = 0; l < s->L && l < 16; ++l) {
    const double hc = s->hc[l];
    b.q0u[l] = s->fs[0] * hc;
    b.q0v[l] = s->fs[1] * hc;
    b.hc[l] = hc;
    // subdomain (problem.jl:424-429): x0 + n (h - hc) / 2
    b.x0[l] = s->x0 + (s->g.nx * (s->h - hc)) / 2;
    b.y0[l] = s->y0 + (s->g.ny * (s->h - hc)) / 2;
  }
  b.ux = s->fs[0];
  b.uy = s->fs[1];
  b.cx = s->gm[0];
  b.cy = s->gm[1];
  b.vx = s->gm[2];
  b.vy = s->gm[3];
  const double c = cos(s->gm[4]), sn = sin(s->gm[4]), om = s->gm[5];
  b.rx11 = c;
  b.rx21 = sn;
  b.rx12 = -sn;
  b.rx22 = c;
  b.rv11 = om * -sn;
  b.rv12 = om * -c;
  b.rv21 = om * c;
  b.rv22 = om * -sn;
  return b;
}

int solver_step(ibpm_solver* s, const double* fsv, const double* motion, ibpm_step_stats* stats,
                int phase) {
  if (s->graph_mode && s->box_dyn)
    return fail(IBPM_E_ARG, "a captured step cannot follow moving bodies with box updates");
  if (phase == 1) return step_head(s, fsv, motion, stats);
  if (phase == 2) return step_tail(s, stats);
  return step(s, fsv, motion, stats);
}

int solver_cg_continue(ibpm_solver* s) { return cg_continue(s); }

void base_flux_dyn(const ibpm_solver* s, const double fs8[8], double out[14]) {
  // the scalars base_flux() derives from the freestream and GridVelocity, same expressions
  const double c = cos(fs8[6]), sn = sin(fs8[6]), om = fs8[7];
  const double v[14] = {fs8[0], fs8[1], fs8[2], fs8[3], fs8[4], fs8[5], c, sn, -sn, c,
                        om * -sn, om * -c, om * c, om * -sn};
  (void)s;
  for (int k = 0; k < 14; ++k) out[k] = v[k];
}

// curl! (fluid.jl:230-276) of the current psi into q where the last step left it to the
// next nonlinear kernel; the same kernel and bits the step would have written.
int materialize_q(ibpm_solver* s) {
  if (!s->q_stale) return IBPM_OK;
  HIPCHK(launch_curl_all(s->g, s->L, s->q, s->psi, s->st));
  s->q_stale = false;
  return IBPM_OK;
}
}  // namespace ibpm

// ---- Operator access that allocates nothing (ibpm_apply_noalloc) ----
//
// Its intermediate planes live in this file-scope device array: one copy per GPU, at its
// own address, held until the process ends. On the host the name `scratch` only stands
// for the symbol; each device's address is looked up once (hipGetSymbolAddress, at the
// first ibpm_create on that device) and every launch takes that address.
static __device__ __align__(256) double scratch[IBPM_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
std::atomic<double*> g_noalloc[kNoallocMaxDevices];
hipError_t g_noalloc_err[kNoallocMaxDevices];

// With `device` current: that device's address of `scratch`, once per device.
void noalloc_setup(int device) {
  if (device < 0 || device >= kNoallocMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc[device].load()) return;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_noalloc_err[device] = e;
  if (e == hipSuccess && p) g_noalloc[device].store((double*)p);
  else (void)hipGetLastError();
}

constexpr size_t round256(size_t doubles) { return (doubles + 31) & ~(size_t)31; }
// carve-outs end 64 KiB before the array's end, so no clamped or vector load of the last
// plane can reach past it
constexpr size_t kNoallocUsable = IBPM_NOALLOC_SCRATCH_BYTES - 65536;

// Points the solver's scratch (sine-pass planes T, U, R and the ring chain's factors and
// partial projections) at a carve-out of `scratch` for one call, and back afterwards.
struct ScratchSwap {
  ibpm_solver* s;
  double *T, *U, *R, *fac, *rvec, *rpart;
  ScratchSwap(ibpm_solver* s_, double* T_, double* U_, double* R_, double* fac_,
              double* rvec_, double* rpart_)
      : s(s_), T(s_->T), U(s_->U), R(s_->R), fac(s_->fac), rvec(s_->rvec), rpart(s_->rpart) {
    s->T = T_;
    s->U = U_;
    s->R = R_;
    s->fac = fac_;
    s->rvec = rvec_;
    s->rpart = rpart_;
  }
  ~ScratchSwap() {
    s->T = T;
    s->U = U;
    s->R = R;
    s->fac = fac;
    s->rvec = rvec;
    s->rpart = rpart;
  }
};

}  // namespace

extern "C" {

void ibpm_default_options(ibpm_options* o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->precision = IBPM_FP64;
  o->refine_sweeps = 2;
  o->cg_maxiter = 5000;
  o->cg_reltol = 1e-12;
  o->cg_abstol = 0.0;
  o->cg_tol_mode = 0;
  o->device = 0;
  o->dst_tiles = -1;
  o->cg_precond = 0;
  o->cg_precond_refresh = 0;
  o->helmholtz = 0;
  o->box_update = 0;
}

const char* ibpm_last_error(void) { return g_err.c_str(); }

int ibpm_box_update(const ibpm_solver* s, int level, int box[4]) {
  if (!s || !box || level < 0 || level > s->L) return fail(IBPM_E_ARG, "bad arguments");
  box[0] = box[1] = box[2] = box[3] = 0;
  if (level == 0) {
    if (!s->pbox_on) return 0;
    box[0] = s->pi0;
    box[1] = s->pbw;
    box[2] = s->pj0;
    box[3] = s->pbh;
    return 1;
  }
  if (!s->box_on) return 0;
  box[0] = s->bi0[level - 1];
  box[1] = s->bw[level - 1];
  box[2] = s->bj0[level - 1];
  box[3] = s->bh[level - 1];
  return 1;
}

int ibpm_dst_tiles(const ibpm_solver* s, int nplanes, int xdir, int* TB, int* TM, int* PF,
                   int* fold, int* KC) {
  if (!s || !TB || !TM || !PF) return fail(IBPM_E_ARG, "null argument");
  int v = s->dst_tiles;
  if (v < 0) {
    const int Nk = xdir ? s->g.Nx : s->g.Ny, Nb = xdir ? s->g.Ny : s->g.Nx;
    v = dst_tuned_variant(s->precision == IBPM_MIXED, nplanes, Nb, (Nk + 1) / 2);
    if (v < 0) v = dst_default_variant(s->precision == IBPM_MIXED);
  }
  v = dst_effective_variant(v, xdir ? s->g.Nx : s->g.Ny, s->precision == IBPM_MIXED);
  int f = 2, kc = 16;
  dst_variant_tiles(v, s->precision == IBPM_MIXED, TB, TM, PF, &f, &kc);
  if (fold) *fold = f;
  if (KC) *KC = kc;
  return IBPM_OK;
}

int ibpm_create(const ibpm_problem* p, const ibpm_options* o, ibpm_solver** out) {
  if (!p || !out) return fail(IBPM_E_ARG, "null argument");
  *out = nullptr;
  ibpm_options opt;
  ibpm_default_options(&opt);
  if (o) opt = *o;
  if (p->nx < 8 || p->ny < 8) return fail(IBPM_E_ARG, "nx and ny must be at least 8");
  if (p->nlevel < 1 || p->nlevel > 16) return fail(IBPM_E_ARG, "nlevel must be in 1..16");
  if (p->npanel < 0 || p->npanel > 8191) return fail(IBPM_E_ARG, "npanel must be in 0..8191");
  if (p->npanel > 0 && (!p->xb || p->nbody < 1 || !p->body_start || !p->body_count || !p->body_moving))
    return fail(IBPM_E_ARG, "panels need xb and a body partition");
  if (opt.precision != IBPM_FP64 && opt.precision != IBPM_MIXED)
    return fail(IBPM_E_ARG, "unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= opt.device)
    return fail(IBPM_E_NO_DEVICE, "no HIP device");
  HIPCHK(hipSetDevice(opt.device));
  noalloc_setup(opt.device);

  ibpm_solver* s = new ibpm_solver();
  s->g = make_dims(p->nx, p->ny);
  s->L = p->nlevel;
  s->x0 = p->x0;
  s->y0 = p->y0;
  s->h = p->h;
  s->Re = p->Re;
  s->dt = p->dt;
  s->t0 = p->t0;
  s->n = p->npanel;
  s->nbody = p->npanel > 0 ? p->nbody : 0;
  s->precision = opt.precision;
  s->refine = std::max(0, opt.refine_sweeps);
  s->cg_maxiter = opt.cg_maxiter > 0 ? opt.cg_maxiter : 5000;
  s->cg_reltol = opt.cg_reltol > 0 ? opt.cg_reltol : 1e-12;
  s->cg_abstol = opt.cg_abstol;
  s->cg_tol_mode = opt.cg_tol_mode;
  s->cg_precond = opt.cg_precond ? 1 : 0;
  s->cg_precond_refresh = std::max(0, opt.cg_precond_refresh);
  s->helmholtz = opt.helmholtz == 1 ? 1 : 0;
  s->box_opt = opt.box_update == 1 || opt.box_update == 2 ? opt.box_update : 0;
  s->device = opt.device;
  s->grid_moving = p->grid_moving ? 1 : 0;
  s->hc.resize(s->L);
  for (int l = 0; l < s->L; ++l) s->hc[l] = pow(2.0, l) * s->h;

  auto bail = [&](int rc) {
    std::string m = g_err;
    free_all(s);
    delete s;
    g_err = m;
    return rc;
  };
#define CK(expr)                                                                        \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bail(fail(IBPM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
  } while (0)

  CK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
  const size_t P = s->plane(), Q = (size_t)s->g.nq, L = s->L;
  CK(dalloc(&s->q, Q * L));
  CK(dalloc(&s->G, P * L));
  CK(dalloc(&s->psi, P * L));
  CK(dalloc(&s->nl1, P * L));
  CK(dalloc(&s->nl2, P * L));
  CK(dalloc(&s->Gs, P * L));
  CK(dalloc(&s->rhs, P * L));
  CK(dalloc(&s->T, P * kMaxPlanes));
  CK(dalloc(&s->U, P * kMaxPlanes));
  if (s->precision == IBPM_MIXED) CK(dalloc(&s->R, P * kMaxPlanes));
  CK(hipMemsetAsync(s->q, 0, Q * L * 8, s->st));
  CK(hipMemsetAsync(s->G, 0, P * L * 8, s->st));
  CK(hipMemsetAsync(s->psi, 0, P * L * 8, s->st));
  CK(hipMemsetAsync(s->nl1, 0, P * L * 8, s->st));
  CK(hipMemsetAsync(s->nl2, 0, P * L * 8, s->st));
  CK(hipMemsetAsync(s->Gs, 0, P * L * 8, s->st));

  // spectral eigenvalues lambda = -2((cos(pi i/nx) + cos(pi j/ny)) - 2) and the DST-I
  // scale 1/(4 nx ny) (fluid.jl:117-134, :155, :168-173): the diagonals are evaluated in
  // the kernels from the two cosine vectors
  {
    const int nx = s->g.nx, ny = s->g.ny, Nx = s->g.Nx, Ny = s->g.Ny;
    s->scale = 1.0 / (4.0 * nx * ny);
    std::vector<double> cx(Nx), cy(Ny);
    for (int i = 1; i <= Nx; ++i) cx[i - 1] = cos(M_PI * i / nx);
    for (int j = 1; j <= Ny; ++j) cy[j - 1] = cos(M_PI * j / ny);
    CK(dalloc(&s->cosx, (size_t)Nx));
    CK(dalloc(&s->cosy, (size_t)Ny));
    CK(hipMemcpy(s->cosx, cx.data(), Nx * 8, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->cosy, cy.data(), Ny * 8, hipMemcpyHostToDevice));
  }
  // sine tables
  {
    auto build = [&](int N, void** So, void** Se, void** Sq, void** Bf, void** Bq) -> hipError_t {
      std::vector<double> o(sine_table_odd_count(N)), e(std::max<size_t>(1, sine_table_even_count(N)));
      build_sine_tables_f64(N, o.data(), e.data());
      std::vector<double> q(sine_table_quarter_count(N));
      build_sine_tables_f64q(N, q.data());
      hipError_t r = dalloc((double**)Sq, q.size());
      if (r) return r;
      r = hipMemcpy(*Sq, q.data(), q.size() * 8, hipMemcpyHostToDevice);
      if (r) return r;
      r = dalloc((double**)So, o.size());
      if (r) return r;
      r = dalloc((double**)Se, e.size());
      if (r) return r;
      r = hipMemcpy(*So, o.data(), o.size() * 8, hipMemcpyHostToDevice);
      if (r) return r;
      r = hipMemcpy(*Se, e.data(), e.size() * 8, hipMemcpyHostToDevice);
      if (r) return r;
      if (s->precision == IBPM_MIXED) {
        const size_t no = sine_table_bf16_odd_count(N), ne = sine_table_bf16_even_count(N);
        std::vector<uint16_t> bf(no + ne);
        const size_t half_o = no / 2, half_e = ne / 2;
        build_sine_tables_bf16(N, bf.data(), bf.data() + half_o, bf.data() + no, bf.data() + no + half_e);
        r = dalloc((uint16_t**)Bf, bf.size());
        if (r) return r;
        r = hipMemcpy(*Bf, bf.data(), bf.size() * 2, hipMemcpyHostToDevice);
        if (r) return r;
        if (mixed_quarter_ok(N)) {
          std::vector<uint16_t> bq(sine_table_bf16q_count(N));
          build_sine_tables_bf16q(N, bq.data());
          r = dalloc((uint16_t**)Bq, bq.size());
          if (r) return r;
          r = hipMemcpy(*Bq, bq.data(), bq.size() * 2, hipMemcpyHostToDevice);
        }
      }
      return r;
    };
    CK(build(s->g.Nx, &s->SoX, &s->SeX, &s->SqX, &s->BfX, &s->BqX));
    CK(build(s->g.Ny, &s->SoY, &s->SeY, &s->SqY, &s->BfY, &s->BqY));
    // ring chain tables: the first and last rows of S (apply_bc! edges) and the ring-line
    // projection weights, with exact argument reduction
    const int nx = s->g.nx, ny = s->g.ny, Nx = s->g.Nx, Ny = s->g.Ny;
    auto s2 = [](long long num, long long n) {
      return 2.0 * sin(M_PI * (double)(num % (2 * n)) / (double)n);
    };
    std::vector<double> ed(2 * Nx + 2 * Ny), pw(2 * Ny + 2 * Nx);
    for (int k = 1; k <= Nx; ++k) {
      ed[k - 1] = s2(k, nx);                       // S_x[k, 1]
      ed[Nx + k - 1] = s2((long long)k * Nx, nx);  // S_x[k, Nx]
      pw[2 * Ny + k - 1] = s2((long long)k * (nx / 4), nx);           // S_x[nx/4, k]
      pw[2 * Ny + Nx + k - 1] = s2((long long)k * (3 * nx / 4), nx);  // S_x[3nx/4, k]
    }
    for (int k = 1; k <= Ny; ++k) {
      ed[2 * Nx + k - 1] = s2(k, ny);
      ed[2 * Nx + Ny + k - 1] = s2((long long)k * Ny, ny);
      pw[k - 1] = s2((long long)k * (ny / 4), ny);             // S_y[k, ny/4]
      pw[Ny + k - 1] = s2((long long)k * (3 * ny / 4), ny);    // S_y[k, 3ny/4]
    }
    CK(dalloc(&s->edge, ed.size()));
    CK(dalloc(&s->pw, pw.size()));
    CK(hipMemcpy(s->edge, ed.data(), ed.size() * 8, hipMemcpyHostToDevice));
    CK(hipMemcpy(s->pw, pw.data(), pw.size() * 8, hipMemcpyHostToDevice));
    const size_t nf = ed.size(), nbc = 2 * (size_t)(nx + 1) + 2 * (size_t)(ny + 1);
    CK(dalloc(&s->fac, kMaxPlanes * nf));
    if (s->precision != IBPM_MIXED) {  // level-1 spectral planes of the step (FP64)
      CK(dalloc(&s->F1, s->plane()));
      CK(dalloc(&s->G1, s->plane()));
    }
    CK(dalloc(&s->rvec, kMaxChains * ring_vector_count(Nx, Ny)));
    CK(dalloc(&s->rpart, kMaxChains * ring_partial_count(Nx, Ny)));
    (void)nbc;
    // ring operators M_n = S_n Interp Ring S_n, one FP64 sine pass each (ring_kernels.hip)
    for (int dir = 0; dir < 2; ++dir) {
      const int N = dir == 0 ? Nx : Ny;
      if (dir == 1 && Ny == Nx) {
        s->My = s->Mx;
        break;
      }
      std::vector<double> Bh((size_t)N * N);
      build_ring_operator_input(N + 1, Bh.data());
      double* Bd = nullptr;
      double*& Md = dir == 0 ? s->Mx : s->My;
      CK(dalloc(&Md, (size_t)N * N));
      CK(dalloc(&Bd, (size_t)N * N));
      CK(hipMemcpy(Bd, Bh.data(), Bh.size() * 8, hipMemcpyHostToDevice));
      PassBatch pb;
      memset(&pb, 0, sizeof pb);
      pb.p[0].in = Bd;
      pb.p[0].out = Md;
      pb.p[0].Nk = N;
      pb.p[0].Nb = N;
      pb.p[0].So = dir == 0 ? s->SoX : s->SoY;
      pb.p[0].Se = dir == 0 ? s->SeX : s->SeY;
      pb.p[0].Sq = dir == 0 ? s->SqX : s->SqY;
      pb.p[0].cosb = dir == 0 ? s->cosx : s->cosy;
      pb.p[0].cosj = pb.p[0].cosb;
      CK(launch_dst_pass_f64(pb, 1, N, (N + 1) / 2, EPI_STORE, s->st, dst_default_variant(false)));
      CK(hipStreamSynchronize(s->st));
      (void)hipFree(Bd);
    }
  }
  s->dst_tiles = opt.dst_tiles >= 0 && opt.dst_tiles < kDstVariants ? opt.dst_tiles : -1;
  if (s->dst_tiles < 0) {
    int rc = autotune_passes(s);
    if (rc) return bail(rc);
  }
  // bodies
  const int n = s->n;
  CK(dalloc(&s->xb, 2 * (size_t)n));
  CK(dalloc(&s->ub, 2 * (size_t)n));
  CK(dalloc(&s->xb_ref, 2 * (size_t)n));
  CK(dalloc(&s->Ft, 2 * (size_t)n));
  CK(dalloc(&s->fb, 2 * (size_t)n));
  CK(dalloc(&s->Ftmp, 2 * (size_t)n));
  CK(dalloc(&s->total, 2));
  CK(dalloc(&s->motion, 6 * (size_t)std::max(1, s->nbody)));
  CK(dalloc(&s->body_of, (size_t)n));
  CK(dalloc(&s->moving_d, (size_t)std::max(1, s->nbody)));
  CK(dalloc(&s->coords, 2 * (size_t)n));
  CK(dalloc(&s->perm, (size_t)n));
  CK(dalloc(&s->xstart, (size_t)s->g.nx + 3));
  CK(dalloc(&s->ywin, 2 * ((size_t)s->g.nx + 1)));
  CK(dalloc(&s->gemv_part, gemv_partial_count(2 * n)));
  CK(dalloc(&s->err, 1));
  CK(dalloc(&s->w, 98 * (size_t)n));
  CK(dalloc(&s->cg_scal, 16));
  CK(dalloc(&s->dyn_d, 16));
  CK(hipHostMalloc((void**)&s->h_scal, 16 * sizeof(double), hipHostMallocDefault));
  CK(hipHostMalloc((void**)&s->h_force, 2 * kForceChunk * sizeof(double), hipHostMallocMapped));
  CK(hipHostGetDevicePointer((void**)&s->d_force, s->h_force, 0));
  // allocations hold max(count, 1) elements (dalloc)
  auto zero_bytes = [](size_t count, size_t elem) { return std::max<size_t>(count, 1) * elem; };
  CK(hipMemsetAsync(s->Ft, 0, zero_bytes(2 * (size_t)n, 8), s->st));
  CK(hipMemsetAsync(s->fb, 0, zero_bytes(2 * (size_t)n, 8), s->st));
  CK(hipMemsetAsync(s->total, 0, 16, s->st));
  CK(hipMemsetAsync(s->ub, 0, zero_bytes(2 * (size_t)n, 8), s->st));
  CK(hipMemsetAsync(s->xb, 0, zero_bytes(2 * (size_t)n, 8), s->st));
  CK(hipMemsetAsync(s->err, 0, sizeof(int), s->st));
  CK(hipMemsetAsync(s->coords, 0, zero_bytes(2 * (size_t)n, sizeof(int)), s->st));
  CK(hipMemsetAsync(s->w, 0, zero_bytes(98 * (size_t)n, 8), s->st));
  s->ds.assign(n, 0.0);
  if (n > 0) {
    std::vector<int> bof(n, 0), mv(s->nbody, 0);
    s->body_start.assign(p->body_start, p->body_start + s->nbody);
    s->body_count.assign(p->body_count, p->body_count + s->nbody);
    s->body_moving.assign(p->body_moving, p->body_moving + s->nbody);
    s->staged_motion.assign(6 * (size_t)s->nbody, 0.0);
    s->staged_t.assign(s->nbody, 0.0);
    s->staged_set.assign(s->nbody, 0);
    for (int b = 0; b < s->nbody; ++b) {
      const int st0 = p->body_start[b], cnt = p->body_count[b];
      if (st0 < 0 || cnt < 0 || st0 + cnt > n) return bail(fail(IBPM_E_ARG, "body partition out of range"));
      for (int k = st0; k < st0 + cnt; ++k) bof[k] = b;
      mv[b] = p->body_moving[b] ? 1 : 0;
      if (mv[b]) s->moving = 1;
    }
    CK(hipMemcpy(s->body_of, bof.data(), n * sizeof(int), hipMemcpyHostToDevice));
    CK(hipMemcpy(s->moving_d, mv.data(), s->nbody * sizeof(int), hipMemcpyHostToDevice));
    CK(hipMemcpy(s->xb_ref, p->xb, 2 * n * 8, hipMemcpyHostToDevice));
    // _init!(StaticBody): xb, ds taken now, ub = 0; moving bodies keep zeros until the
    // first step (bodies.jl:1-12, :33-41)
    std::vector<double> xb0(2 * n, 0.0);
    for (int b = 0; b < s->nbody; ++b) {
      if (mv[b]) continue;
      const int st0 = p->body_start[b], cnt = p->body_count[b];
      for (int k = st0; k < st0 + cnt; ++k) {
        xb0[k] = p->xb[k];
        xb0[n + k] = p->xb[n + k];
        s->ds[k] = p->ds ? p->ds[k] : 0.0;
      }
    }
    CK(hipMemcpy(s->xb, xb0.data(), 2 * n * 8, hipMemcpyHostToDevice));
    BodyDev bd = body_dev(s);
    CK(launch_body_update(bd, 0, s->st));  // static panels: weights now
    // moving panels are skipped by only_moving=0? no: static branch reads xb (zeros for
    // moving bodies) -- recompute moving ones at the first step; clear their error.
    CK(hipStreamSynchronize(s->st));
    if (!s->moving) {
      int rc = check_body_error(s);
      if (rc) return bail(rc);
    } else {
      int e = 0;
      CK(hipMemcpy(&e, s->err, sizeof(int), hipMemcpyDeviceToHost));
      if (e) {
        // an error from a static panel is real; one from a moving panel at zero is not
        int rc = check_body_error(s);
        const int pbad = n - e;
        if (!mv[bof[pbad]]) return bail(rc);
        g_err.clear();
      }
    }
    CK(launch_body_sort(bd, s->st));
    if (!s->moving) {
      // static panels: rot(E^T f) from per-vertex term lists (build_scatter_lists)
      std::vector<int> hc(2 * (size_t)n), hp(n), rowptr, ep;
      std::vector<double> hw(98 * (size_t)n), ew;
      CK(hipMemcpyAsync(hc.data(), s->coords, hc.size() * sizeof(int), hipMemcpyDeviceToHost, s->st));
      CK(hipMemcpyAsync(hp.data(), s->perm, hp.size() * sizeof(int), hipMemcpyDeviceToHost, s->st));
      CK(hipMemcpyAsync(hw.data(), s->w, hw.size() * sizeof(double), hipMemcpyDeviceToHost, s->st));
      CK(hipStreamSynchronize(s->st));
      build_scatter_lists(n, s->g.Nx, s->g.Ny, hc.data(), hp.data(), hw.data(), rowptr, ep, ew);
      CK(dalloc(&s->sc_rowptr, rowptr.size()));
      CK(dalloc(&s->sc_ep, ep.size()));
      CK(dalloc(&s->sc_ew, ew.size()));
      CK(hipMemcpy(s->sc_rowptr, rowptr.data(), rowptr.size() * sizeof(int), hipMemcpyHostToDevice));
      if (!ep.empty()) {
        CK(hipMemcpy(s->sc_ep, ep.data(), ep.size() * sizeof(int), hipMemcpyHostToDevice));
        CK(hipMemcpy(s->sc_ew, ew.data(), ew.size() * sizeof(double), hipMemcpyHostToDevice));
      }
    }
  }
  // B^-1 (static) or CG buffers (moving): coupling.jl:179-202
  if (n > 0) {
    int rc = setup_box_update(s, p);  // the boxes (static bodies) also serve B's assembly
    if (rc) return bail(rc);
    if (!s->moving) {
      rc = assemble_B_inverse(s, s->Binv);
      if (rc) return bail(rc);
      {  // stored row-major for the per-step product (launch_gemv_rows)
        const size_t m = 2 * (size_t)n;
        std::vector<double> cm(m * m), rm(m * m);
        CK(hipStreamSynchronize(s->st));  // the assembly runs on the solver's stream
        CK(hipMemcpy(cm.data(), s->Binv, m * m * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t j = 0; j < m; ++j)
          for (size_t i = 0; i < m; ++i) rm[i * m + j] = cm[i + j * m];
        CK(hipMemcpy(s->Binv, rm.data(), m * m * sizeof(double), hipMemcpyHostToDevice));
      }
      rc = alloc_bscratch(s, 0);  // static bodies never apply B again
      if (rc) return bail(rc);
    } else {
      rc = alloc_bscratch(s, 1);
      if (rc) return bail(rc);
      CK(dalloc(&s->cg_r, 2 * (size_t)n));
      CK(dalloc(&s->cg_u, 2 * (size_t)n));
      CK(dalloc(&s->cg_c, 2 * (size_t)n));
      CK(dalloc(&s->ws_h0, 2 * (size_t)n));
      CK(dalloc(&s->ws_h1, 2 * (size_t)n));
      CK(dalloc(&s->ws_xe, 2 * (size_t)n));
      CK(dalloc(&s->ws_Bxe, 2 * (size_t)n));
      CK(dalloc(&s->ws_scal, 4));
      s->cur_theta.assign(std::max(1, s->nbody), 0.0);
      if (s->cg_precond) {
        CK(dalloc(&s->pc_theta0, (size_t)std::max(1, s->nbody)));
        CK(dalloc(&s->pc_t, 2 * (size_t)n));
        CK(dalloc(&s->pc_z, 2 * (size_t)n));
      }
    }
  }
  CK(hipStreamSynchronize(s->st));
#undef CK
  *out = s;
  return IBPM_OK;
}

int ibpm_advance(ibpm_solver* s, int nsteps, const double* freestream, const double* motion,
                 ibpm_step_stats* stats) {
  if (!s || nsteps < 0) return fail(IBPM_E_ARG, "bad arguments");
  if (s->moving && !motion && nsteps > 0) {
    // staged motion (ibpm_set_body_motion): one step, every moving body at its time
    if (nsteps != 1)
      return fail(IBPM_E_ARG, "moving bodies need the motion array (staged motion covers one step)");
    const double t_next = s->t0 + s->dt * s->index;  // time after inc_time_index!
    for (int b = 0; b < s->nbody; ++b) {
      if (!s->body_moving[b]) continue;
      if (!s->staged_set[b])
        return fail(IBPM_E_ARG, "moving body " + std::to_string(b) +
                                    " has no motion for the next step (ibpm_set_body_motion)");
      if (fabs(s->staged_t[b] - t_next) > 1e-9 * std::max(1.0, fabs(t_next)))
        return fail(IBPM_E_ARG, "motion of body " + std::to_string(b) + " was evaluated at t = " +
                                    std::to_string(s->staged_t[b]) + ", the next step is at t = " +
                                    std::to_string(t_next));
    }
    motion = s->staged_motion.data();
  }
  HIPCHK(hipSetDevice(s->device));
  // The forces kernel writes each step's total force into the mapped pinned h_force
  // (a copy into the caller's pageable stats would make the host wait for the stream
  // every step); stats take them after a synchronisation every kForceChunk steps.
  for (int k = 0; k < nsteps; ++k) {
    ibpm_step_stats* sk = stats ? stats + k : nullptr;
    if (sk) memset(sk, 0, sizeof *sk);
    s->force_slot = sk && s->n > 0 ? s->d_force + 2 * (k % kForceChunk) : nullptr;
    int rc = step(s, freestream ? freestream + (s->grid_moving ? 8 : 2) * k : nullptr,
                  motion ? motion + (size_t)6 * s->nbody * k : nullptr, sk);
    s->force_slot = nullptr;
    if (rc) return rc;
    if (!sk) continue;
    if (s->n == 0) {
      sk->force[0] = sk->force[1] = 0.0;
      continue;
    }
    if ((k + 1) % kForceChunk == 0 || k + 1 == nsteps) {
      HIPCHK(hipStreamSynchronize(s->st));
      for (int j = k - k % kForceChunk; j <= k; ++j) {
        stats[j].force[0] = s->h_force[2 * (j % kForceChunk)];
        stats[j].force[1] = s->h_force[2 * (j % kForceChunk) + 1];
      }
    }
  }
  std::fill(s->staged_set.begin(), s->staged_set.end(), 0);
  HIPCHK(hipStreamSynchronize(s->st));
  return IBPM_OK;
}

int ibpm_set_body_motion(ibpm_solver* s, int body, double t, const double pos[2],
                         const double vel[2], double angle, double angular_vel) {
  if (!s || !pos || !vel) return fail(IBPM_E_ARG, "null argument");
  if (body < 0 || body >= s->nbody) return fail(IBPM_E_ARG, "body index out of range");
  if (!s->body_moving[body]) return fail(IBPM_E_ARG, "body " + std::to_string(body) + " is static");
  double* m = &s->staged_motion[6 * (size_t)body];
  m[0] = pos[0];
  m[1] = pos[1];
  m[2] = vel[0];
  m[3] = vel[1];
  m[4] = angle;
  m[5] = angular_vel;
  s->staged_t[body] = t;
  s->staged_set[body] = 1;
  return IBPM_OK;
}

size_t ibpm_field_size(const ibpm_solver* s, int field) {
  if (!s) return 0;
  const size_t P = s->plane(), Q = s->g.nq, L = s->L, n = s->n;
  switch (field) {
    case IBPM_Q: case IBPM_Q0: return Q * L;
    case IBPM_GAMMA: case IBPM_PSI: case IBPM_NONLIN1: case IBPM_NONLIN2: return P * L;
    case IBPM_XB: case IBPM_UB: case IBPM_FTILDE: case IBPM_FB: return 2 * n;
    case IBPM_TOTAL_FORCE: case IBPM_FREESTREAM: return 2;
    case IBPM_DS: return n;
  }
  return 0;
}

static double* field_ptr(ibpm_solver* s, int field) {
  switch (field) {
    case IBPM_Q: return s->q;
    case IBPM_GAMMA: return s->G;
    case IBPM_PSI: return s->psi;
    case IBPM_NONLIN1: return s->nl1;
    case IBPM_NONLIN2: return s->nl2;
    case IBPM_XB: return s->xb;
    case IBPM_UB: return s->ub;
    case IBPM_FTILDE: return s->Ft;
    case IBPM_FB: return s->fb;
    case IBPM_TOTAL_FORCE: return s->total;
  }
  return nullptr;
}

int ibpm_read(ibpm_solver* s, int field, double* dst, size_t n) {
  if (!s || !dst) return fail(IBPM_E_ARG, "null argument");
  const size_t need = ibpm_field_size(s, field);
  if (field < IBPM_Q || field > IBPM_FREESTREAM) return fail(IBPM_E_ARG, "unknown field");
  if (n != need) return fail(IBPM_E_ARG, "size mismatch: expected " + std::to_string(need));
  if (need == 0) return IBPM_OK;
  HIPCHK(hipSetDevice(s->device));
  if (field == IBPM_Q) {
    int rc = materialize_q(s);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(s->st));
  if (field == IBPM_Q0) {
    if (s->q0_shadow_valid) {
      memcpy(dst, s->q0_shadow.data(), need * 8);
      return IBPM_OK;
    }
    // base_flux! (fluid.jl:22-67) of the last step; zero before the first step
    const size_t Q = s->g.nq;
    const int nx = s->g.nx, ny = s->g.ny;
    const BaseFlux bf = base_flux(s);
    for (int l = 0; l < s->L; ++l) {
      double* qx = dst + l * Q;
      double* qy = qx + s->g.nu;
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx + 1; ++i)
          qx[(i - 1) + (size_t)(j - 1) * (nx + 1)] = s->index > 0 ? base_flux_u(bf, l, i, j) : 0.0;
      for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx; ++i)
          qy[(i - 1) + (size_t)(j - 1) * nx] = s->index > 0 ? base_flux_v(bf, l, i, j) : 0.0;
    }
    return IBPM_OK;
  }
  if (field == IBPM_DS) {
    memcpy(dst, s->ds.data(), need * 8);
    return IBPM_OK;
  }
  if (field == IBPM_FREESTREAM) {
    dst[0] = s->fs[0];
    dst[1] = s->fs[1];
    return IBPM_OK;
  }
  HIPCHK(hipMemcpy(dst, field_ptr(s, field), need * 8, hipMemcpyDeviceToHost));
  return IBPM_OK;
}

int ibpm_write(ibpm_solver* s, int field, const double* src, size_t n) {
  if (!s || !src) return fail(IBPM_E_ARG, "null argument");
  const size_t need = ibpm_field_size(s, field);
  if (field < IBPM_Q || field > IBPM_FREESTREAM) return fail(IBPM_E_ARG, "unknown field");
  if (n != need) return fail(IBPM_E_ARG, "size mismatch: expected " + std::to_string(need));
  if (need == 0) return IBPM_OK;
  HIPCHK(hipSetDevice(s->device));
  if (field == IBPM_PSI) {  // q stays the curl of the psi it came from
    int rc = materialize_q(s);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(s->st));
  switch (field) {
    case IBPM_Q0:
      s->q0_shadow.assign(src, src + need);
      s->q0_shadow_valid = true;
      return IBPM_OK;
    case IBPM_DS:
      s->ds.assign(src, src + need);
      return IBPM_OK;
    case IBPM_FREESTREAM:
      s->fs[0] = src[0];
      s->fs[1] = src[1];
      return IBPM_OK;
    case IBPM_TOTAL_FORCE:
      return fail(IBPM_E_ARG, "total force is derived from fb");
  }
  HIPCHK(hipMemcpy(field_ptr(s, field), src, need * 8, hipMemcpyHostToDevice));
  if (field == IBPM_FTILDE) s->nhist = 0;  // the extrapolation history no longer applies
  if (field == IBPM_Q) s->q_stale = false;
  return IBPM_OK;
}

int ibpm_get_index(const ibpm_solver* s) { return s ? s->index : 0; }

int ibpm_set_index(ibpm_solver* s, int index) {
  if (!s) return fail(IBPM_E_ARG, "null solver");
  s->index = index;
  s->nhist = 0;
  return IBPM_OK;
}

int ibpm_set_warm_start(ibpm_solver* s, int mode, ibpm_net* net, int record) {
  if (!s) return fail(IBPM_E_ARG, "null solver");
  if (mode < IBPM_WS_PREVIOUS || mode > IBPM_WS_LEARNED) return fail(IBPM_E_ARG, "unknown warm-start mode");
  if ((mode == IBPM_WS_LEARNED || record) && !net)
    return fail(IBPM_E_ARG, "IBPM_WS_LEARNED and recording need a network");
  if (net && net_device(net) != s->device)
    return fail(IBPM_E_ARG, "the network lives on another device than the solver");
  if (net && s->n > net_max_points(net))
    return fail(IBPM_E_ARG, "the solver has more body points than the network's max_points");
  // the F~b history is kept when both the old and the new setting extend it every step
  const bool had = s->ws_mode != IBPM_WS_PREVIOUS || s->ws_record;
  s->ws_mode = mode;
  s->ws_record = record ? 1 : 0;
  s->net = net;
  if (!(had && (mode != IBPM_WS_PREVIOUS || record))) s->nhist = 0;
  return IBPM_OK;
}

double ibpm_get_time(const ibpm_solver* s) {
  // set_time_index!: t = t0 + dt (index - 1) (problem.jl:703-708)
  return s ? s->t0 + s->dt * (double)(s->index - 1) : 0.0;
}

int ibpm_synchronize(ibpm_solver* s) {
  if (!s) return fail(IBPM_E_ARG, "null solver");
  HIPCHK(hipStreamSynchronize(s->st));
  return IBPM_OK;
}

int ibpm_apply(ibpm_solver* s, int op, int level, const double* in, double* out, size_t n) {
  if (!s || !in || !out) return fail(IBPM_E_ARG, "null argument");
  const size_t P = s->plane();
  const int L = s->L;
  const size_t need = op == IBPM_OP_VORT2FLUX ? P * L : P;
  if (n != need) return fail(IBPM_E_ARG, "size mismatch: expected " + std::to_string(need));
  if (op == IBPM_OP_HELMHOLTZ && (level < 1 || level > L)) return fail(IBPM_E_ARG, "bad level");
  if (op < IBPM_OP_POISSON || op > IBPM_OP_VORT2FLUX) return fail(IBPM_E_ARG, "unknown op");
  HIPCHK(hipSetDevice(s->device));
  // staging: rhs planes hold the input, Gs planes (vort2flux) the output; both are step
  // scratch, overwritten by the next step. The solver's stream does not synchronise with
  // the null stream, so every host copy waits for it explicitly.
  HIPCHK(hipStreamSynchronize(s->st));
  HIPCHK(hipMemcpy(s->rhs, in, need * 8, hipMemcpyHostToDevice));
  if (op == IBPM_OP_VORT2FLUX) {
    double* Gc[1] = {s->rhs};
    double* pc[1] = {s->Gs};
    int rc = vort2flux(s, 1, Gc, pc, nullptr, nullptr);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipMemcpy(out, s->Gs, need * 8, hipMemcpyDeviceToHost));
    return IBPM_OK;
  }
  SolveJob j;
  j.b = s->rhs;
  j.x = s->rhs + P;
  j.epi = EPI_STORE;
  j.base = nullptr;
  if (op == IBPM_OP_POISSON) {
    j.da = 0.0;
    j.dnum = 1.0;
    j.dden = 1.0;
    j.a = 0.0;
    j.c = 1.0;
  } else {
    const double hc = s->hc[level - 1];
    j.da = 1.0;
    j.dnum = s->dt;
    j.dden = 2.0 * s->Re * (hc * hc);
    j.a = 1.0;
    j.c = cA(s, level - 1);
  }
  if (L < 2) return fail(IBPM_E_ARG, "operator access needs nlevel >= 2 (scratch planes)");
  int rc = solve(s, &j, 1);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->st));
  HIPCHK(hipMemcpy(out, s->rhs + P, need * 8, hipMemcpyDeviceToHost));
  return IBPM_OK;
}

int ibpm_apply_noalloc(ibpm_solver* s, int op, int level, const double* in, double* out,
                       size_t n) {
  if (!s || !in || !out) return fail(IBPM_E_ARG, "null argument");
  const size_t P = s->plane();
  const int L = s->L;
  if (op < IBPM_OP_POISSON || op > IBPM_OP_VORT2FLUX) return fail(IBPM_E_ARG, "unknown op");
  const bool v2f = op == IBPM_OP_VORT2FLUX;
  const size_t need = v2f ? P * L : P;
  if (n != need) return fail(IBPM_E_ARG, "size mismatch: expected " + std::to_string(need));
  if (op == IBPM_OP_HELMHOLTZ && (level < 1 || level > L)) return fail(IBPM_E_ARG, "bad level");
  const uintptr_t a0 = (uintptr_t)in, b0 = (uintptr_t)out, len = need * sizeof(double);
  if (a0 < b0 + len && b0 < a0 + len) return fail(IBPM_E_ARG, "in and out must not overlap");
  if (s->device < 0 || s->device >= kNoallocMaxDevices)
    return fail(IBPM_E_ARG, "device ordinal beyond the scratch table");
  HIPCHK(hipSetDevice(s->device));
  double* base = g_noalloc[s->device].load();
  if (!base)
    return fail(IBPM_E_HIP, std::string("no scratch address on this device: ") +
                                hipGetErrorString(g_noalloc_err[s->device]));
  // carve-outs, each starting on a 256-byte boundary: vort2flux stages its input (it
  // coarsifies in place) and uses L planes of T, U (and R in mixed precision), the ring
  // factors and the partial projections; a single solve one plane of each
  const bool mixed = s->precision == IBPM_MIXED;
  const size_t np = v2f ? (size_t)L : 1;
  const size_t nf = 2 * (size_t)s->g.Nx + 2 * (size_t)s->g.Ny;
  size_t off = 0;
  auto take = [&](size_t k) { double* p = base + off; off += round256(k); return p; };
  double* stage = v2f ? take(P * L) : nullptr;
  double* T = take(P * np);
  double* U = take(P * np);
  double* R = mixed ? take(P * np) : nullptr;
  double* fac = v2f ? take((size_t)L * nf) : nullptr;
  double* rvec = v2f ? take(ring_vector_count(s->g.Nx, s->g.Ny)) : nullptr;
  double* rpart = v2f ? take(ring_partial_count(s->g.Nx, s->g.Ny)) : nullptr;
  if (off * sizeof(double) > kNoallocUsable)
    return fail(IBPM_E_ARG, "operator needs " + std::to_string(off * sizeof(double)) +
                                " bytes of scratch; ibpm_apply_noalloc has " +
                                std::to_string(kNoallocUsable));
  if (!v2f && L < 2) return fail(IBPM_E_ARG, "operator access needs nlevel >= 2");
  ScratchSwap swap(s, T, U, R, fac, rvec, rpart);
  if (v2f) {
    HIPCHK(hipMemcpyAsync(stage, in, need * sizeof(double), hipMemcpyDeviceToDevice, s->st));
    double* Gc[1] = {stage};
    double* pc[1] = {out};
    return vort2flux(s, 1, Gc, pc, nullptr, nullptr);
  }
  SolveJob j;
  j.b = in;
  j.x = out;
  j.epi = EPI_STORE;
  j.base = nullptr;
  if (op == IBPM_OP_POISSON) {
    j.da = 0.0;
    j.dnum = 1.0;
    j.dden = 1.0;
    j.a = 0.0;
    j.c = 1.0;
  } else {
    const double hc = s->hc[level - 1];
    j.da = 1.0;
    j.dnum = s->dt;
    j.dden = 2.0 * s->Re * (hc * hc);
    j.a = 1.0;
    j.c = cA(s, level - 1);
  }
  // one job: solve()'s routing without its host-side job lists
  return use_gs(s, j) ? solve_gs(s, &j, 1) : solve_chunk(s, &j, 1);
}

int ibpm_time_kernel(ibpm_solver* s, int kernel, int reps, double* avg_ms, double* flops,
                     double* bytes) {
  if (!s || reps < 1 || !avg_ms || !flops || !bytes) return fail(IBPM_E_ARG, "bad arguments");
  if (kernel < 0 || kernel > 2) return fail(IBPM_E_ARG, "unknown kernel");
  HIPCHK(hipSetDevice(s->device));
  const size_t P = s->plane();
  const bool mixed = s->precision == IBPM_MIXED;
  if (kernel == 2) {
    // register-only MFMA loop, back-to-back launches
    const int probe_blocks = 8 * 256;  // needs probe_blocks x 256 doubles of scratch
    if ((size_t)probe_blocks * 256 > kMaxPlanes * P) return fail(IBPM_E_ARG, "grid too small for the probe");
    double probe_flops = 0.0;
    HIPCHK(launch_mfma_probe(mixed, 1000, probe_blocks, s->T, s->st, &probe_flops));  // warm-up
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipError_t e = hipEventRecord(e0, s->st);
    for (int r = 0; r < reps && e == hipSuccess; ++r)
      e = launch_mfma_probe(mixed, 1000, probe_blocks, s->T, s->st, &probe_flops);
    if (e == hipSuccess) e = hipEventRecord(e1, s->st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(e);
    *avg_ms = ms / reps;
    *flops = probe_flops;
    *bytes = 0.0;
    return IBPM_OK;
  }
  // kernels 0 and 1: `reps` real steps (static bodies or none; the solver advances) with
  // events around each launch of the kernel class
  if (s->moving) return fail(IBPM_E_ARG, "kernel timing runs real steps: static bodies only");
  HIPCHK(hipStreamSynchronize(s->st));
  s->probe_ev.assign(64, nullptr);
  hipError_t e = hipSuccess;
  for (hipEvent_t& ev : s->probe_ev)
    if (e == hipSuccess) e = hipEventCreate(&ev);
  double fs8[8] = {s->fs[0], s->fs[1], s->gm[0], s->gm[1], s->gm[2], s->gm[3], s->gm[4], s->gm[5]};
  double total = 0.0;
  long count = 0;
  int rc = IBPM_OK;
  s->probe = kernel;
  for (int r = 0; r < reps && e == hipSuccess && rc == IBPM_OK; ++r) {
    s->probe_n = 0;
    rc = step(s, fs8, nullptr, nullptr);
    if (rc == IBPM_OK) e = hipStreamSynchronize(s->st);
    for (int k = 0; k < s->probe_n && e == hipSuccess; ++k) {
      float one = 0.f;
      e = hipEventElapsedTime(&one, s->probe_ev[2 * k], s->probe_ev[2 * k + 1]);
      total += one;
      ++count;
    }
  }
  s->probe = -1;
  for (hipEvent_t ev : s->probe_ev)
    if (ev) (void)hipEventDestroy(ev);
  s->probe_ev.clear();
  if (rc) return rc;
  HIPCHK(e);
  if (!count) return fail(IBPM_E_ARG, "the step launched no kernel of that class");
  *avg_ms = total / count;
  if (kernel == 0) {
    // matrix-core flops of the contraction one launch (all levels) runs: the single fold
    // multiplies H x H and (H-1) x (H-1) tables, the double fold three Q x Q and one
    // (Q-1) x (Q-1), the third fold per block and pair of rows (m, Q-1-m) two Q-long and
    // two Q/2-long contractions (3 Q^2 per block)
    const int L = s->L;
    const double Nk = s->g.Nx, Nb = s->g.Ny;
    const int Hi = (s->g.Nx + 1) / 2;
    const double H = Hi, Q = (s->g.Nx + 1) / 4, E = ((s->g.Nx + 1) & 1) ? Hi : Hi - 1;
    int v = s->dst_tiles >= 0 ? s->dst_tiles : dst_tuned_variant(mixed, L, s->g.Ny, (s->g.Nx + 1) / 2);
    if (v < 0) v = dst_default_variant(mixed);
    int TB = 0, TM = 0, PF = 0, fold = 2;
    dst_variant_tiles(dst_effective_variant(v, s->g.Nx, mixed), mixed, &TB, &TM, &PF, &fold);
    const double per_product = fold == 8   ? 2.0 * Nb * 6.0 * Q * Q / 2.0
                               : fold == 4 ? 2.0 * Nb * (3.0 * Q * Q + (Q - 1) * (Q - 1))
                                           : 2.0 * Nb * (H * H + E * E);
    *flops = L * per_product * (mixed ? 3.0 : 1.0);
    *bytes = L * Nk * Nb * (8.0 + (mixed ? 4.0 : 8.0));  // fp64 plane in, plane out
  } else {
    *flops = 0.0;
    // q (nq), or psi (nG) where the kernel forms the fluxes, + Gamma + nl_old in,
    // nl_new + rhs out, per level
    *bytes = (double)s->L * ((s->q_stale ? (double)P : (double)s->g.nq) + 4.0 * (double)P) * 8.0;
  }
  return IBPM_OK;
}

void ibpm_destroy(ibpm_solver* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->st) (void)hipStreamSynchronize(s->st);
  free_all(s);
  delete s;
}

}  // extern "C"
