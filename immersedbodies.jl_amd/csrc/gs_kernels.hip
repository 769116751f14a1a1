// Red-black Gauss-Seidel solves of the implicit Helmholtz systems of the CN step,
//   (da I + cA (-L)) x = b,  L the 5-point Laplacian on the interior vertices, x = 0 on
// the level's boundary (the boundary values of Gamma are already in b),
// the operator that ImmersedBodies.jl inverts with its sine transforms (fluid.jl:141-172,
// Ainv in solver.jl:201-210). The Jacobi iteration matrix of this M-matrix has infinity
// norm rho = 4 cA / (da + 4 cA); a red half-sweep followed by a black one multiplies the
// error of the colour just updated by at most rho per half-sweep, so m half-sweeps from
// x = 0 leave an error of at most rho^(m-1) max|x|. The host picks m from rho (a bound of
// 2^-56, below FP64 rounding) and uses the sine transforms where m would be large.
//
// Temporal blocking: a work-group owns a 128 x 128 tile of one plane (8 waves of 16 rows;
// lane l holds columns 2l, 2l+1 of its wave's rows in registers) and runs up to kGsHalo
// half-sweeps on it without leaving registers. Values on the tile's edge use wrong
// neighbours for those outside the tile; a half-sweep moves that error one cell inward, so
// the centre 96 x 96, kGsHalo cells from the edge, is exact and is what the tile writes.
// Horizontal neighbours come from the adjacent lane by DPP wave shifts, vertical ones
// from the thread's own registers, and across waves through a small LDS row exchange.
// The colour of vertex (i, j) is (i + j) & 1 in plane indices; tile origins are even,
// so within a tile it is (column + row) & 1 and each row updates one column per lane.
#include "ibpm_device.h"

namespace ibpm {
namespace {

constexpr int kRows = 16;                 // rows per wave
constexpr int kWaves = 8;                 // waves per work-group
constexpr int kTile = 128;                // tile edge: 64 lanes x 2 columns, 8 x 16 rows
constexpr int kOut = kTile - 2 * kGsHalo; // exact centre written per tile
static_assert(kRows * kWaves == kTile, "tile rows");
static_assert((kOut % 2) == 0 && (kGsHalo % 2) == 0, "tile origins must stay even");

// value of the lane below (lane l reads lane l - 1); lane 0 keeps its own (a tile edge)
__device__ __forceinline__ double from_left(double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// value of the lane above (lane l reads lane l + 1); lane 63 keeps its own (a tile edge)
__device__ __forceinline__ double from_right(double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), 0x130, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// One half-sweep over the points of colour C held by this thread. rok: bit r set when
// row r of the wave lies in the plane; cok[cp]: column cp of the lane does.
template <int C>
__device__ __forceinline__ void half_sweep(double (&x)[kRows][2], const double (&bb)[kRows][2],
                                           const double above[2], const double below[2],
                                           double c, unsigned rok, const bool cok[2]) {
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int cp = (C + r) & 1;
    const double up = r > 0 ? x[r > 0 ? r - 1 : 0][cp] : above[cp];
    const double dn = r < kRows - 1 ? x[r < kRows - 1 ? r + 1 : 0][cp] : below[cp];
    double left, right;
    if (cp == 0) {
      left = from_left(x[r][1]);
      right = x[r][1];
    } else {
      left = x[r][0];
      right = from_right(x[r][0]);
    }
    const double v = bb[r][cp] + c * ((up + dn) + (left + right));
    x[r][cp] = (((rok >> r) & 1u) && cok[cp]) ? v : 0.0;
  }
}

__global__ __launch_bounds__(kWaves * 64) void gs_helmholtz_kernel(const GsBatch bt) {
  const GsPlane& pl = bt.p[blockIdx.y];
  const int Nx = bt.Nx, Ny = bt.Ny;
  const int tx = blockIdx.x % bt.ntx, ty = blockIdx.x / bt.ntx;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i0 = tx * kOut - kGsHalo + 2 * lane;  // this lane's first column
  const int j0 = ty * kOut - kGsHalo + kRows * w;  // this wave's first row
  __shared__ double2 edge[2][kWaves][2][64];       // [buffer][wave][first/last row][lane]

  const bool cok[2] = {i0 >= 0 && i0 < Nx, i0 + 1 >= 0 && i0 + 1 < Nx};
  unsigned rok = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r)
    if (j0 + r >= 0 && j0 + r < Ny) rok |= 1u << r;

  const double c = pl.c, inv_d = pl.inv_d;
  double x[kRows][2], bb[kRows][2];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
#pragma unroll
    for (int cp = 0; cp < 2; ++cp) {
      const bool ok = ((rok >> r) & 1u) && cok[cp];
      const size_t idx = (size_t)(i0 + cp) + (size_t)(j0 + r) * Nx;
      bb[r][cp] = ok ? inv_d * pl.b[idx] : 0.0;
      x[r][cp] = ok && pl.xin ? pl.xin[idx] : 0.0;
    }
  }
  edge[0][w][0][lane] = make_double2(x[0][0], x[0][1]);
  edge[0][w][1][lane] = make_double2(x[kRows - 1][0], x[kRows - 1][1]);
  __syncthreads();
  for (int s = 0; s < pl.nsweep; ++s) {
    const int rb = s & 1;
    double above[2] = {0.0, 0.0}, below[2] = {0.0, 0.0};
    if (w > 0) {
      const double2 a = edge[rb][w - 1][1][lane];
      above[0] = a.x;
      above[1] = a.y;
    }
    if (w < kWaves - 1) {
      const double2 b = edge[rb][w + 1][0][lane];
      below[0] = b.x;
      below[1] = b.y;
    }
    if (((pl.color0 + s) & 1) == 0)
      half_sweep<0>(x, bb, above, below, c, rok, cok);
    else
      half_sweep<1>(x, bb, above, below, c, rok, cok);
    edge[rb ^ 1][w][0][lane] = make_double2(x[0][0], x[0][1]);
    edge[rb ^ 1][w][1][lane] = make_double2(x[kRows - 1][0], x[kRows - 1][1]);
    __syncthreads();
  }
  // the exact centre: tile rows and columns kGsHalo .. kTile - kGsHalo - 1
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int lr = kRows * w + r;
    if (lr < kGsHalo || lr >= kTile - kGsHalo || !((rok >> r) & 1u)) continue;
#pragma unroll
    for (int cp = 0; cp < 2; ++cp) {
      const int lc = 2 * lane + cp;
      if (lc < kGsHalo || lc >= kTile - kGsHalo || !cok[cp]) continue;
      const size_t idx = (size_t)(i0 + cp) + (size_t)(j0 + r) * Nx;
      pl.out[idx] = pl.base ? pl.base[idx] - x[r][cp] : x[r][cp];
    }
  }
}

}  // namespace

int gs_tiles(int n) { return (n + kOut - 1) / kOut; }

hipError_t launch_gs_helmholtz(GsBatch& bt, int nplanes, hipStream_t st) {
  bt.ntx = gs_tiles(bt.Nx);
  bt.nty = gs_tiles(bt.Ny);
  for (int p = 0; p < nplanes; ++p)
    if (bt.p[p].nsweep < 1 || bt.p[p].nsweep > kGsHalo) return hipErrorInvalidValue;
  hipLaunchKernelGGL(gs_helmholtz_kernel, dim3(bt.ntx * bt.nty, nplanes), dim3(kWaves * 64), 0, st,
                     bt);
  return hipGetLastError();
}

}  // namespace ibpm
