// Grid-stencil kernels of the IBPM step: the fused nonlinear term + CNAB right-hand side,
// coarsify!, curl! and the residual used by the mixed-precision refinement.
// All are HBM-bound sweeps. In the nonlinear kernel a lane takes its neighbours' columns
// from the lanes beside it; in the others one thread owns one output value and reads its
// neighbours through L2 (neighbouring threads share every load).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ibpm_device.h"
#include "ibpm_internal.h"

// The stencils are HBM-bound; keeping each product and sum separately rounded, in the
// reference's order, makes the nonlinear term and the coarsening bitwise equal to the
// FP64 restatement of the reference.
#pragma clang fp contract(off)

namespace ibpm {

// ---------------------------------------------------------------------------------
// Nonlinear term and CNAB right-hand side (fluid.jl:352-455 and :467-497), every level
// in one launch (blockIdx.z = level). Per interior vertex (I, J):
//   nonlin = (fqx(I+1,J) - fqx(I+1,J+1)) + (fqy(I+1,J+1) - fqy(I,J+1))) / hc^2
//   fqx(i,j) = (Qy(i,j+1) Ge(i-1,j) + Qy(i,j) Ge(i-1,j-1)) / 2
//   fqy(i,j) = -(Qx(i+1,j) Ge(i,j-1) + Qx(i,j) Ge(i-1,j-1)) / 2
//   Qx, Qy   = face averages of q + q0 (avg_flux!)
// where Ge is the circulation on the level's vertices with the boundary ring taken from
// the next coarser level times 0.25 (get_bc!, then bc .*= 0.25 at fluid.jl:480).
//   rhs = A Gamma + dt (beta1 nonlin_new + beta2 nonlin_old) + rhsbc
// with A = I - dt/(2 Re hc^2) K applied as the 5-point stencil (K = rot curl with zero
// Dirichlet values, the operator the DST diagonalises) and rhsbc = apply_bc! of the
// unscaled ring with fac = 0.25 dt / (Re hc^2).
// ---------------------------------------------------------------------------------
// Work layout: a wave owns 62 consecutive columns I and a window of NL_RJ rows; its 64
// lanes hold columns bx*62 + lane, so every lane loads one value per row of each input
// (Gamma on the vertices, Qy at v-face column I+1, Qx at u-face column I+1) and takes
// its neighbours' columns from the lanes beside it (DPP wave shifts); lanes 0 and 63
// only feed their neighbours. That is about 3 loads per row instead of the 8 a thread
// reading all its neighbours issues; the values, and their sums, are the same bits.
constexpr int NL_RJ = 2;   // rows J per wave window
constexpr int NL_W = 62;   // output columns per wave

// Value of lane L-1 (lane 0 keeps its own) and of lane L+1 (lane 63 keeps its own).
__device__ __forceinline__ double lane_prev(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_next(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false),
                          __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false));
}

// PSI: the fluxes come from the streamfunction (a.psi, with the coarser level's ring) by
// the curl! of curl_all_kernel, evaluated where they are needed: the same expressions,
// so the same bits as reading the q that curl_all_kernel writes; a lane reads its psi
// column once per row and takes the next column from the lane beside it.
template <bool DYN, bool PSI>
__global__ __launch_bounds__(256) void nonlin_rhs_kernel(const NonlinArgs a) {
  const GridDims g = a.g;
  const int lev = blockIdx.z;
  const int lane = threadIdx.x & 63;
  const int I = blockIdx.x * NL_W + lane;  // lanes 0 and 63: neighbour columns only
  const int J0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * NL_RJ + 1;
  if (J0 > g.Ny) return;  // uniform per wave: the lane shifts below see every lane
  const int nx = g.nx, ny = g.ny;
  const int nrow = min(NL_RJ, g.Ny - J0 + 1);
  const double* q = a.q + (size_t)lev * g.nq;
  const double* qx = q;
  const double* qy = q + g.nu;
  const double* G = a.G + (size_t)lev * g.nG;
  const double* Gc = lev + 1 < a.nlevel ? a.G + (size_t)(lev + 1) * g.nG : nullptr;

  // Gamma at column I, rows J0-1 .. J0+NL_RJ, with the 0.25-scaled coarse ring outside
  // the interior: interior values by unconditional loads (index clamped into the plane),
  // the ring (ext_value) only where the window leaves the interior; rows past the top
  // edge (j > ny) and columns past nx are never used
  const int Ic = min(max(I, 1), g.Nx);
  double gv[NL_RJ + 2];
#pragma unroll
  for (int r = 0; r < NL_RJ + 2; ++r) {
    const int jc = min(max(J0 - 1 + r, 1), g.Ny);
    gv[r] = G[(size_t)(Ic - 1) + (size_t)(jc - 1) * g.Nx];
  }
  // psi(I, j), j = J0-1 .. J0+NL_RJ (PSI; index clamped into the plane) and the previous
  // nonlinear term at the window's output rows (only the output lanes' values are used),
  // loaded with Gamma before any boundary rule: one memory round trip for all of them
  const double* p = PSI ? a.psi + (size_t)lev * g.nG : nullptr;
  const double* pc = PSI && lev + 1 < a.nlevel ? a.psi + (size_t)(lev + 1) * g.nG : nullptr;
  double pv[NL_RJ + 2];
  if (PSI) {
#pragma unroll
    for (int r = 0; r < NL_RJ + 2; ++r) {
      const int jc = min(max(J0 - 1 + r, 1), g.Ny);
      pv[r] = p[(size_t)(Ic - 1) + (size_t)(jc - 1) * g.Nx];
    }
  }
  double nlo[NL_RJ];
#pragma unroll
  for (int r = 0; r < NL_RJ; ++r) {
    const int jc = min(J0 + r, g.Ny);
    nlo[r] = a.nl_old[(size_t)(Ic - 1) + (size_t)(jc - 1) * g.Nx + (size_t)lev * g.nG];
  }
#pragma unroll
  for (int r = 0; r < NL_RJ + 2; ++r) {
    const int j = J0 - 1 + r;
    const bool interior = I > 0 && I < nx && j > 0 && j < ny;
    if (!interior) gv[r] = (r <= nrow + 1 && j <= ny && I <= nx) ? ext_value(G, Gc, nx, ny, I, j, 0.25) : 0.0;
  }
  // Qy(I+1, j), j = J0 .. J0+NL_RJ+1 (v-faces up to ny+1); Qy(I, j) is the previous lane's
  const int Iy = min(I + 1, nx);
  double qa[NL_RJ + 2];
  // Qx(I+1, j), j = J0 .. J0+NL_RJ (u-faces up to ny)
  const int Ix = min(I + 1, nx + 1);
  double qxr[NL_RJ + 1];
  if (PSI) {
    // psi(I+1, j) from the next lane; faces with a boundary vertex, or whose column is
    // clamped, go through the boundary rules (curl_u / curl_v). Lane 63's v-faces feed no
    // other lane.
#pragma unroll
    for (int r = 0; r < NL_RJ + 2; ++r) {
      const double pn = lane_next(pv[r]);  // psi(I + 1, J0 - 1 + r)
      const int jv = min(J0 + r, ny + 1);
      const bool in = Iy == I + 1 && I + 1 >= 2 && I + 1 <= nx - 1 && jv >= 2 && jv <= ny;
      qa[r] = in ? pv[r] - pn : (lane == 63 ? 0.0 : curl_v(p, pc, nx, ny, Iy, jv));
    }
#pragma unroll
    for (int r = 0; r < NL_RJ + 1; ++r) {
      const int ju = min(J0 + r, ny);
      const bool in = Ix == I + 1 && I + 1 >= 2 && I + 1 <= nx && ju >= 2 && ju <= ny - 1;
      qxr[r] = in ? pv[r + 1] - pv[r] : curl_u(p, pc, nx, ny, Ix, ju);
    }
  } else {
#pragma unroll
    for (int r = 0; r < NL_RJ + 2; ++r) qa[r] = qy[(Iy - 1) + (size_t)(min(J0 + r, ny + 1) - 1) * nx];
#pragma unroll
    for (int r = 0; r < NL_RJ + 1; ++r) qxr[r] = qx[(Ix - 1) + (size_t)(min(J0 + r, ny) - 1) * (nx + 1)];
  }
  double qyv[NL_RJ + 2];
#pragma unroll
  for (int r = 0; r < NL_RJ + 2; ++r) {
    const int j = J0 + r;
    const double qb = lane_prev(qa[r]);
    qyv[r] = r <= nrow + 1
                 ? (((qa[r] + qb) + base_flux_v<DYN>(a.bf, lev, I + 1, j)) + base_flux_v<DYN>(a.bf, lev, I, j)) / 2
                 : 0.0;
  }
#pragma unroll
  for (int r = 0; r < NL_RJ + 1; ++r) qxr[r] = r <= nrow ? qxr[r] : 0.0;

  const double fac = Gc ? a.fac_bc[lev] : 0.0;
  const bool out_lane = lane > 0 && lane < 63 && I <= g.Nx;
#pragma unroll
  for (int r = 0; r < NL_RJ; ++r) {
    if (r >= nrow) break;
    const int J = J0 + r;
    // Qx at u-face column I+1 (this lane), I and I+2 (the lanes beside it)
    const double qxb = (((qxr[r + 1] + qxr[r]) + base_flux_u<DYN>(a.bf, lev, I + 1, J + 1)) + base_flux_u<DYN>(a.bf, lev, I + 1, J)) / 2;
    const double qxa = lane_prev(qxb), qxc = lane_next(qxb);
    const double g00 = gv[r + 1];
    const double gm0 = lane_prev(g00);
    const double gp0 = lane_next(g00);
    const double g0m = gv[r];
    const double g0p = gv[r + 2];
    const double qy1 = qyv[r], qy2 = qyv[r + 1], qy3 = qyv[r + 2];

    const double fqx_a = (qy2 * g00 + qy1 * g0m) / 2;   // fqx(I+1, J)
    const double fqx_b = (qy3 * g0p + qy2 * g00) / 2;   // fqx(I+1, J+1)
    const double fqy_a = -(qxc * gp0 + qxb * g00) / 2;  // fqy(I+1, J+1)
    const double fqy_b = -(qxb * g00 + qxa * gm0) / 2;  // fqy(I, J+1)
    const double nl = ((fqx_a - fqx_b) + (fqy_a - fqy_b)) * a.inv_hc2[lev];
    if (!out_lane) continue;

    const size_t idx = (size_t)(I - 1) + (size_t)(J - 1) * g.Nx + (size_t)lev * g.nG;
    a.nl_new[idx] = nl;

    // A Gamma with zero Dirichlet neighbours: the window's values inside the plane
    const double gl = I > 1 ? gm0 : 0.0, gr = I < g.Nx ? gp0 : 0.0;
    const double gb = J > 1 ? g0m : 0.0, gt = J < g.Ny ? g0p : 0.0;
    const double Kg = 4.0 * g00 - gl - gr - gb - gt;
    double v = g00 - a.cA[lev] * Kg;
    v = v + (a.dt * a.beta1) * nl;
    v = v + (a.dt * a.beta2) * nlo[r];
    if (Gc) {
      double rb = 0.0;
      if (J == 1) rb += fac * ring_value(Gc, nx, ny, 2, I + 1);
      if (J == g.Ny) rb += fac * ring_value(Gc, nx, ny, 3, I + 1);
      if (I == 1) rb += fac * ring_value(Gc, nx, ny, 0, J + 1);
      if (I == g.Nx) rb += fac * ring_value(Gc, nx, ny, 1, J + 1);
      v = v + rb;
    }
    a.rhs[idx] = v;
  }
}

hipError_t launch_nonlin_rhs(const NonlinArgs& a, hipStream_t st) {
  if (g_skip) return no_skip();
  dim3 grid((a.g.Nx + NL_W - 1) / NL_W, (a.g.Ny + 4 * NL_RJ - 1) / (4 * NL_RJ), a.nlevel);
  if (a.psi) {
    if (a.bf.dyn)
      hipLaunchKernelGGL((nonlin_rhs_kernel<true, true>), grid, dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((nonlin_rhs_kernel<false, true>), grid, dim3(256), 0, st, a);
  } else {
    if (a.bf.dyn)
      hipLaunchKernelGGL((nonlin_rhs_kernel<true, false>), grid, dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((nonlin_rhs_kernel<false, false>), grid, dim3(256), 0, st, a);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// coarsify! (fluid.jl:208-228): coarse vertex ic0 + a, jc0 + c of the centre region takes
// the 1-1/2-1/4 full weighting of fine vertex (2 + 2a, 2 + 2c), with the reference's
// ranges: ic0 = nx÷2 - nx÷4 + 1 and a < min(nx÷2 - 1, 2 (nx÷4) - 1) (its zip of the
// fine and coarse index ranges stops at the shorter), likewise in y. For nx divisible by 4
// that is the region [nx/4+1, 3nx/4-1]. With `outer`, the rest of the coarse plane is
// copied from it (project_circ's state.Gamma .= Gamma_s fused into the final pass).
// ---------------------------------------------------------------------------------
// (CoarsifyRange, coarsify_range: ibpm_internal.h)

// the full weighting of fine vertex (i, j) (1-based), in the reference's order
__device__ __forceinline__ double coarsify_value(const double* __restrict__ fine, int Nx, int i, int j) {
#define GF(ii, jj) fine[((ii)-1) + (size_t)((jj)-1) * Nx]
  return GF(i, j) + 0.5 * (GF(i + 1, j) + GF(i, j + 1) + GF(i - 1, j) + GF(i, j - 1)) +
         0.25 * (GF(i + 1, j + 1) + GF(i + 1, j - 1) + GF(i - 1, j - 1) + GF(i - 1, j + 1));
#undef GF
}

__global__ __launch_bounds__(256) void coarsify_kernel(GridDims g, double* __restrict__ coarse,
                                                       const double* __restrict__ fine,
                                                       const double* __restrict__ outer,
                                                       const double* skip) {
  if (skipped(skip)) return;
  const CoarsifyRange cr = coarsify_range(g.nx, g.ny);
  int Ic, Jc;
  if (outer) {
    Ic = blockIdx.x * 64 + (threadIdx.x & 63) + 1;
    Jc = blockIdx.y * 4 + (threadIdx.x >> 6) + 1;
  } else {
    Ic = cr.ic0 + blockIdx.x * 64 + (threadIdx.x & 63);
    Jc = cr.jc0 + blockIdx.y * 4 + (threadIdx.x >> 6);
  }
  if (Ic > g.Nx || Jc > g.Ny) return;
  const size_t idx = (size_t)(Ic - 1) + (size_t)(Jc - 1) * g.Nx;
  const int a = Ic - cr.ic0, c = Jc - cr.jc0;
  const bool centre = a >= 0 && a < cr.ni && c >= 0 && c < cr.nj;
  if (!centre) {
    if (outer) coarse[idx] = outer[idx];
    return;
  }
  coarse[idx] = coarsify_value(fine, g.Nx, 2 + 2 * a, 2 + 2 * c);
}

hipError_t launch_coarsify(GridDims g, double* coarse, const double* fine, const double* outer,
                           hipStream_t st) {
  dim3 grid;
  if (outer) {
    grid = dim3((g.Nx + 63) / 64, (g.Ny + 3) / 4, 1);
  } else {
    const CoarsifyRange cr = coarsify_range(g.nx, g.ny);
    if (cr.ni == 0 || cr.nj == 0) return hipSuccess;
    grid = dim3((cr.ni + 63) / 64, (cr.nj + 3) / 4, 1);
  }
  hipLaunchKernelGGL(coarsify_kernel, grid, dim3(256), 0, st, g, coarse, fine, outer, g_skip);
  return hipGetLastError();
}

// coarsify! re-evaluated on box l of level l only (intersected with the centre region),
// where a change of the finer level can reach it (the boxes of the box updates: box l
// holds every coarse vertex whose weighting touches box l-1); elsewhere the coarse values
// are already those of the finer level's unchanged values. The values it overwrites are
// first saved at the same offsets of `saved`. One vertex per thread, one launch per level
// (level l reads what level l-1's launch wrote).
__global__ __launch_bounds__(256) void coarsify_box_kernel(GridDims g, double* __restrict__ coarse,
                                                           const double* __restrict__ fine,
                                                           double* __restrict__ saved, int i0,
                                                           int w, int j0, int n, const double* skip) {
  if (skipped(skip)) return;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const CoarsifyRange cr = coarsify_range(g.nx, g.ny);
  const int Ic = i0 + e % w, Jc = j0 + e / w;
  const size_t idx = (size_t)(Ic - 1) + (size_t)(Jc - 1) * g.Nx;
  const double v = coarsify_value(fine, g.Nx, 2 + 2 * (Ic - cr.ic0), 2 + 2 * (Jc - cr.jc0));
  saved[idx] = coarse[idx];
  coarse[idx] = v;
}

hipError_t launch_coarsify_boxes(GridDims g, int L, double* G, double* saved, const BoxSet& b,
                                 hipStream_t st) {
  const CoarsifyRange cr = coarsify_range(g.nx, g.ny);
  if (cr.ni == 0 || cr.nj == 0) return hipSuccess;
  for (int l = 1; l < L; ++l) {
    if (b.w[l] <= 0 || b.h[l] <= 0) continue;
    // box l (1-based vertices) within the centre
    const int i0 = std::max(b.i0[l] + 1, cr.ic0), i1 = std::min(b.i0[l] + b.w[l], cr.ic0 + cr.ni - 1);
    const int j0 = std::max(b.j0[l] + 1, cr.jc0), j1 = std::min(b.j0[l] + b.h[l], cr.jc0 + cr.nj - 1);
    if (i1 < i0 || j1 < j0) continue;
    const int w = i1 - i0 + 1, n = w * (j1 - j0 + 1);
    hipLaunchKernelGGL(coarsify_box_kernel, dim3((n + 255) / 256), dim3(256), 0, st, g,
                       G + (size_t)l * g.nG, G + (size_t)(l - 1) * g.nG, saved + (size_t)l * g.nG,
                       i0, w, j0, n, g_skip);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------------
// curl! with boundary streamfunction (fluid.jl:230-276) for all levels at once:
//   u(i,j) = psi_e(i-1, j) - psi_e(i-1, j-1),  v(i,j) = psi_e(i-1, j-1) - psi_e(i, j-1)
// psi_e = psi on interior vertices, get_bc! ring of the next coarser level on the
// boundary (zero on the coarsest level, fluid.jl:187-190), with the reference's choice of
// side at the corners (curl_u / curl_v).
// ---------------------------------------------------------------------------------
// Work layout: a wave holds psi columns c = bx*63 + lane (0 and nx: the boundary) for a
// window of CURL_R face rows; each lane loads its column once per row and writes the
// u-face and v-face of column c + 1, the v-face's second vertex coming from the next lane;
// lane 63 only feeds lane 62.
constexpr int CURL_R = 4;

__global__ __launch_bounds__(256) void curl_all_kernel(GridDims g, int nlevel, double* __restrict__ q,
                                                       const double* __restrict__ psi) {
  const int lev = blockIdx.z;
  const int lane = threadIdx.x & 63;
  const int nx = g.nx, ny = g.ny;
  const int c = blockIdx.x * 63 + lane;
  const int J0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * CURL_R + 1;
  if (J0 > ny + 1) return;  // uniform per wave: the lane shifts below see every lane
  const double* p = psi + (size_t)lev * g.nG;
  const double* pc = lev + 1 < nlevel ? psi + (size_t)(lev + 1) * g.nG : nullptr;
  double* ql = q + (size_t)lev * g.nq;
  // psi(c, J0-1+r) by unconditional loads (index clamped into the plane); faces with a
  // boundary vertex go through the boundary rules (curl_u / curl_v) instead
  const int cc = min(max(c, 1), g.Nx);
  double pv[CURL_R + 1];
#pragma unroll
  for (int r = 0; r < CURL_R + 1; ++r) {
    const int jc = min(max(J0 - 1 + r, 1), g.Ny);
    pv[r] = p[(size_t)(cc - 1) + (size_t)(jc - 1) * g.Nx];
  }
#pragma unroll
  for (int r = 0; r < CURL_R; ++r) {
    const int j = J0 + r;
    const double pn = lane_next(pv[r]);  // psi(c + 1, j - 1)
    if (lane == 63) continue;
    const int i = c + 1;
    if (j <= ny && i <= nx + 1) {
      const bool in = i >= 2 && i <= nx && j >= 2 && j <= ny - 1;
      const double v = in ? pv[r + 1] - pv[r] : curl_u(p, pc, nx, ny, i, j);
      ql[(size_t)(i - 1) + (size_t)(j - 1) * (nx + 1)] = v;
    }
    if (j <= ny + 1 && i <= nx) {
      const bool in = i >= 2 && i <= nx - 1 && j >= 2 && j <= ny;
      const double v = in ? pv[r] - pn : curl_v(p, pc, nx, ny, i, j);
      ql[(size_t)g.nu + (size_t)(i - 1) + (size_t)(j - 1) * nx] = v;
    }
  }
}

hipError_t launch_curl_all(GridDims g, int nlevel, double* q, const double* psi, hipStream_t st) {
  if (g_skip) return no_skip();
  dim3 grid((g.nx + 1 + 62) / 63, (g.ny + 1 + 4 * CURL_R - 1) / (4 * CURL_R), nlevel);
  hipLaunchKernelGGL(curl_all_kernel, grid, dim3(256), 0, st, g, nlevel, q, psi);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Residual of the Poisson / Helmholtz system in FP64: r = b_e - (a x + c K x), K the
// 5-point rot curl with zero Dirichlet values, b_e = b plus the apply_bc! ring terms
// when a ring plane is given. Feeds the mixed-precision refinement sweeps.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void residual_kernel(const ResidualArgs a) {
  if (skipped(a.skip)) return;
  const GridDims g = a.g;
  const int pl = blockIdx.z;
  const int I = blockIdx.x * 64 + (threadIdx.x & 63) + 1;
  const int J = blockIdx.y * 4 + (threadIdx.x >> 6) + 1;
  if (I > g.Nx || J > g.Ny) return;
  const double* x = a.x[pl];
  const double* xb = a.xbase[pl];
  auto X = [&](int i, int j) -> double {
    if (!(i >= 1 && i <= g.Nx && j >= 1 && j <= g.Ny)) return 0.0;
    const size_t k = (i - 1) + (size_t)(j - 1) * g.Nx;
    return xb ? xb[k] - x[k] : x[k];
  };
  const size_t idx = (size_t)(I - 1) + (size_t)(J - 1) * g.Nx;
  double b = a.b[pl][idx];
  const double* rc = a.ring[pl];
  if (rc) {
    if (J == 1) b += ring_value(rc, g.nx, g.ny, 2, I + 1);
    if (J == g.Ny) b += ring_value(rc, g.nx, g.ny, 3, I + 1);
    if (I == 1) b += ring_value(rc, g.nx, g.ny, 0, J + 1);
    if (I == g.Nx) b += ring_value(rc, g.nx, g.ny, 1, J + 1);
  }
  const double xc = X(I, J);
  const double Kx = 4.0 * xc - X(I - 1, J) - X(I + 1, J) - X(I, J - 1) - X(I, J + 1);
  const double r = b - (a.a[pl] * xc + a.c[pl] * Kx);
  if (a.r_f32)
    ((float*)a.r[pl])[idx] = (float)r;
  else
    ((double*)a.r[pl])[idx] = r;
}

hipError_t launch_residual(const ResidualArgs& a0, hipStream_t st) {
  ResidualArgs a = a0;
  a.skip = g_skip;
  dim3 grid((a.g.Nx + 63) / 64, (a.g.Ny + 3) / 4, a.nplanes);
  hipLaunchKernelGGL(residual_kernel, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------------
__global__ void fill_kernel(double* p, size_t n, double v, const double* skip) {
  if (skipped(skip)) return;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}
// 4 elements per thread, 256 apart (coalesced), all four loads issued before the stores
__global__ void copy_kernel(double* __restrict__ d, const double* __restrict__ s, size_t n,
                            const double* skip) {
  if (skipped(skip)) return;
  const size_t i0 = (size_t)blockIdx.x * 1024 + threadIdx.x;
  double v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = s[i0 + 256 * k < n ? i0 + 256 * k : n - 1];  // clamped
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (i0 + 256 * k < n) d[i0 + 256 * k] = v[k];
}
hipError_t launch_fill(double* p, size_t n, double v, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v, g_skip);
  return hipGetLastError();
}
__global__ void sub_kernel(double* __restrict__ d, const double* __restrict__ a,
                           const double* __restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) d[i] = a[i] - b[i];
}

hipError_t launch_sub(double* d, const double* a, const double* b, size_t n, hipStream_t st) {
  if (g_skip) return no_skip();
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(sub_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, a, b, n);
  return hipGetLastError();
}

hipError_t launch_copy(double* d, const double* s, size_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, st, d, s, n, g_skip);
  return hipGetLastError();
}

}  // namespace ibpm
