// Device helpers shared by the HIP translation units of libibpm.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace ibpm {

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
// sides of the boundary vector; both entries interpolate the same coarse vertex, so
// the side chosen here does not change the value.
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

}  // namespace ibpm
