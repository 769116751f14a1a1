// Device helpers shared by the HIP translation units of libibpm.
#pragma once

#include "ibpm_internal.h"
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace ibpm {

// The launch context's flag (g_skip of the launching thread): a kernel launched while it
// was set returns at once when the flag is nonzero.
__device__ __forceinline__ bool skipped(const double* flag) { return flag && *flag != 0.0; }

// ---------------------------------------------------------------------------------
// get_bc! entry (fluid.jl:278-306): value on `side` (0 left, 1 right, 2 bottom, 3 top)
// at 1-based position k of the boundary vector, interpolated from the coarse plane r.
// ---------------------------------------------------------------------------------
__host__ __device__ __forceinline__ double ring_value(const double* __restrict__ r, int nx, int ny,
                                             int side, int k) {
  const int Nx = nx - 1;
  if (side < 2) {
    const int I0 = side == 0 ? nx / 4 : 3 * nx / 4;
    if (k & 1) return r[(I0 - 1) + (size_t)(ny / 4 + (k - 1) / 2 - 1) * Nx];
    const int j = ny / 4 + k / 2;
    return 0.5 * (r[(I0 - 1) + (size_t)(j - 1) * Nx] + r[(I0 - 1) + (size_t)(j - 2) * Nx]);
  }
  const int J0 = side == 2 ? ny / 4 : 3 * ny / 4;
  if (k & 1) return r[(nx / 4 + (k - 1) / 2 - 1) + (size_t)(J0 - 1) * Nx];
  const int i = nx / 4 + k / 2;
  return 0.5 * (r[(i - 1) + (size_t)(J0 - 1) * Nx] + r[(i - 2) + (size_t)(J0 - 1) * Nx]);
}

// Extended-plane value of a level: interior vertices (1..nx-1, 1..ny-1) come from the
// plane p, boundary vertices (i = 0, nx or j = 0, ny) from the get_bc! ring of the
// coarser plane rc (zero when rc is null), times scale. A corner vertex belongs to two
// sides of the boundary vector; for even nx and ny both entries interpolate the same
// coarse vertex. The callers that read corners (the curl's boundary rows and columns)
// choose the side the reference reads instead (curl_all_kernel).
__host__ __device__ __forceinline__ double ext_value(const double* __restrict__ p,
                                            const double* __restrict__ rc, int nx, int ny,
                                            int i, int j, double scale) {
  if (i > 0 && i < nx && j > 0 && j < ny) return p[(i - 1) + (size_t)(j - 1) * (nx - 1)];
  if (!rc) return 0.0;
  double v;
  if (j == 0) v = ring_value(rc, nx, ny, 2, i + 1);
  else if (j == ny) v = ring_value(rc, nx, ny, 3, i + 1);
  else if (i == 0) v = ring_value(rc, nx, ny, 0, j + 1);
  else v = ring_value(rc, nx, ny, 1, j + 1);
  return scale * v;
}

// Entry k (1-based) of side `side` (0 left, 1 right, 2 bottom, 3 top) of the get_bc!
// boundary vector of the coarser plane rc: zero when rc is null, and zero at k = n + 1 of
// a side of odd length n, which get_bc! never writes (fluid.jl:278-306) and vort2flux
// zeroes before use (fluid.jl:188).
__host__ __device__ __forceinline__ double bc_entry(const double* __restrict__ rc, int nx, int ny,
                                                    int side, int k) {
  if (!rc) return 0.0;
  if (side < 2 ? (k == ny + 1 && (ny & 1)) : (k == nx + 1 && (nx & 1))) return 0.0;
  return ring_value(rc, nx, ny, side, k);
}

// curl! (fluid.jl:230-276) on u-face (i, j) and v-face (i, j): the u-faces of the left and
// right boundary columns take both values from the left / right side of the boundary
// vector, the v-faces of the bottom and top rows from the bottom / top side (corners
// included); elsewhere the extended plane.
__host__ __device__ __forceinline__ double curl_u(const double* __restrict__ p,
                                                  const double* __restrict__ rc, int nx, int ny,
                                                  int i, int j) {
  if (i == 1 || i == nx + 1) {
    const int side = i == 1 ? 0 : 1;
    return bc_entry(rc, nx, ny, side, j + 1) - bc_entry(rc, nx, ny, side, j);
  }
  return ext_value(p, rc, nx, ny, i - 1, j, 1.0) - ext_value(p, rc, nx, ny, i - 1, j - 1, 1.0);
}

__host__ __device__ __forceinline__ double curl_v(const double* __restrict__ p,
                                                  const double* __restrict__ rc, int nx, int ny,
                                                  int i, int j) {
  if (j == 1 || j == ny + 1) {
    const int side = j == 1 ? 2 : 3;
    return bc_entry(rc, nx, ny, side, i) - bc_entry(rc, nx, ny, side, i + 1);
  }
  return ext_value(p, rc, nx, ny, i - 1, j - 1, 1.0) - ext_value(p, rc, nx, ny, i, j - 1, 1.0);
}

// Base flux on u-face (i, j) and v-face (i, j) of level l (0-based level, 1-based faces),
// in the order of base_flux! (fluid.jl:34-67): u-faces sit at (x0 + h (i-1),
// y0 + h/2 + h (j-1)), v-faces at (x0 + h/2 + h (i-1), y0 + h (j-1)) (problem.jl:318-334).
// With DYN (b.dyn set), the step's freestream and grid motion come from device memory
// (the same values base_flux() puts in the arguments, so results are the same bits); the
// kernels are instantiated for both, so a solver's own steps read only their arguments.
template <bool DYN = true>
__host__ __device__ inline double base_flux_u(const BaseFlux& b, int l, int i, int j) {
#pragma clang fp contract(off)
  const double* d = DYN ? b.dyn : nullptr;
  if (!b.moving) return d ? d[0] * b.hc[l] : b.q0u[l];
  const double hc = b.hc[l];
  const double x = b.x0[l] + hc * (double)(i - 1);
  const double y = (b.y0[l] + hc / 2) + hc * (double)(j - 1);
  if (d) {
    const double dx = x - d[2], dy = y - d[3];
    const double wx = (d[0] - d[4]) - (d[10] * dx + d[11] * dy);
    const double wy = (d[1] - d[5]) - (d[12] * dx + d[13] * dy);
    return hc * (d[6] * wx + d[7] * wy);
  }
  const double dx = x - b.cx, dy = y - b.cy;
  const double wx = (b.ux - b.vx) - (b.rv11 * dx + b.rv12 * dy);
  const double wy = (b.uy - b.vy) - (b.rv21 * dx + b.rv22 * dy);
  return hc * (b.rx11 * wx + b.rx21 * wy);
}

template <bool DYN = true>
__host__ __device__ inline double base_flux_v(const BaseFlux& b, int l, int i, int j) {
#pragma clang fp contract(off)
  const double* d = DYN ? b.dyn : nullptr;
  if (!b.moving) return d ? d[1] * b.hc[l] : b.q0v[l];
  const double hc = b.hc[l];
  const double x = (b.x0[l] + hc / 2) + hc * (double)(i - 1);
  const double y = b.y0[l] + hc * (double)(j - 1);
  if (d) {
    const double dx = x - d[2], dy = y - d[3];
    const double wx = (d[0] - d[4]) - (d[10] * dx + d[11] * dy);
    const double wy = (d[1] - d[5]) - (d[12] * dx + d[13] * dy);
    return hc * (d[8] * wx + d[9] * wy);
  }
  const double dx = x - b.cx, dy = y - b.cy;
  const double wx = (b.ux - b.vx) - (b.rv11 * dx + b.rv12 * dy);
  const double wy = (b.uy - b.vy) - (b.rv21 * dx + b.rv22 * dy);
  return hc * (b.rx12 * wx + b.rx22 * wy);
}

}  // namespace ibpm
