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
// Temporal blocking: a work-group owns a 128-wide tile of one plane (8 waves of R rows,
// R = 16 or 8; lane l holds columns 2l, 2l+1 of its wave's rows in registers) and runs up
// to H <= kGsHalo half-sweeps on it without leaving registers. Values on the tile's edge
// use wrong neighbours for those outside the tile; a half-sweep moves that error one cell
// inward, so the centre (128 - 2H) x (8R - 2H), H cells from the edge, is exact and is
// what the tile writes (H: the plane's half-sweep count in this launch, rounded up to
// even; each plane of a launch has its own tile grid). The exact centre is the global
// iterate whatever the tiles, so R changes no result: launches whose 128 x 128 tiles
// would not fill the chip (small regions) take R = 8, half the sequential work per wave.
// Horizontal neighbours come from the adjacent lane by DPP wave shifts, vertical ones
// from the thread's own registers, and across waves through a small LDS row exchange.
// The colour of vertex (i, j) is (i + j) & 1 in the indices of the region solved; tile
// origins are even, so within a tile it is (column + row) & 1 and each row updates one
// column per lane. A region smaller than the plane (row pitch ld) is solved with x = 0
// around it.
#include "ibpm_device.h"

namespace ibpm {
namespace {

constexpr int kWaves = 8;   // waves per work-group
constexpr int kTile = 128;  // tile width: 64 lanes x 2 columns
static_assert(kGsHalo % 2 == 0 && 2 * kGsHalo < 8 * kWaves, "halo within the short tiles");

// value of the lane below (lane l reads lane l - 1; lane 0 reads 0, a tile edge)
__device__ __forceinline__ double from_left(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// value of the lane above (lane l reads lane l + 1; lane 63 reads 0, a tile edge)
__device__ __forceinline__ double from_right(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// One half-sweep over the points of colour C held by this thread. With ROWS, rok
// (wave-uniform) has bit r set when row r of the wave lies in the plane, and rows outside
// keep x = 0 (waves on the plane's lower or upper edge); without, every row does.
// cl[cp]: c for a column in the plane, 0 outside, where bb = 0 too, so x stays 0 there.
template <int C, bool ROWS, int kRows>
__device__ __forceinline__ void half_sweep(double (&x)[kRows][2], const double (&bb)[kRows][2],
                                           const double above[2], const double below[2],
                                           const double cl[2], unsigned rok) {
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (ROWS && !((rok >> r) & 1u)) continue;
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
    x[r][cp] = bb[r][cp] + cl[cp] * ((up + dn) + (left + right));
  }
}

template <int kRows>
__global__ __launch_bounds__(kWaves * 64) void gs_helmholtz_kernel(const GsBatch bt) {
  constexpr int kTileH = kRows * kWaves;  // tile height
  const GsPlane& pl = bt.p[blockIdx.y];
  const int Nx = bt.Nx, Ny = bt.Ny, ld = bt.ld, H = pl.H, out_w = kTile - 2 * H, out_h = kTileH - 2 * H;
  if ((int)blockIdx.x >= pl.ntx * pl.nty) return;  // whole work-group: barriers stay uniform
  if (skipped(bt.skip)) return;
  const int tx = blockIdx.x % pl.ntx, ty = blockIdx.x / pl.ntx;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = tx * out_w - H + 2 * lane;  // this lane's first column (even)
  const int j0 = ty * out_h - H + kRows * w;  // this wave's first row (even)
  __shared__ double2 edge[2][kWaves][2][64];  // [buffer][wave][first/last row][lane]

  const bool cok[2] = {i0 >= 0 && i0 < Nx, i0 + 1 >= 0 && i0 + 1 < Nx};
  unsigned rok = 0;
#pragma unroll
  for (int r = 0; r < kRows; ++r)
    if (j0 + r >= 0 && j0 + r < Ny) rok |= 1u << r;

  const double inv_d = pl.inv_d;
  const double cl[2] = {cok[0] ? pl.c : 0.0, cok[1] ? pl.c : 0.0};
  // clamped indices: every load is in bounds and issues unconditionally (points outside
  // the plane are zeroed by the selects)
  const int ic0 = min(max(i0, 0), Nx - 1), ic1 = min(max(i0 + 1, 0), Nx - 1);
  double x[kRows][2], bb[kRows][2];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const size_t row = (size_t)min(max(j0 + r, 0), Ny - 1) * ld;
    const double v0 = pl.b[row + ic0], v1 = pl.b[row + ic1];
    const bool rk = (rok >> r) & 1u;
    bb[r][0] = rk && cok[0] ? inv_d * v0 : 0.0;
    bb[r][1] = rk && cok[1] ? inv_d * v1 : 0.0;
  }
  if (pl.xin) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const size_t row = (size_t)min(max(j0 + r, 0), Ny - 1) * ld;
      const double v0 = pl.xin[row + ic0], v1 = pl.xin[row + ic1];
      const bool rk = (rok >> r) & 1u;
      x[r][0] = rk && cok[0] ? v0 : 0.0;
      x[r][1] = rk && cok[1] ? v1 : 0.0;
    }
  } else {
#pragma unroll
    for (int r = 0; r < kRows; ++r) x[r][0] = x[r][1] = 0.0;
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
    const bool all_rows = rok == (1u << kRows) - 1;
    if (((pl.color0 + s) & 1) == 0) {
      if (all_rows)
        half_sweep<0, false, kRows>(x, bb, above, below, cl, rok);
      else
        half_sweep<0, true, kRows>(x, bb, above, below, cl, rok);
    } else {
      if (all_rows)
        half_sweep<1, false, kRows>(x, bb, above, below, cl, rok);
      else
        half_sweep<1, true, kRows>(x, bb, above, below, cl, rok);
    }
    edge[rb ^ 1][w][0][lane] = make_double2(x[0][0], x[0][1]);
    edge[rb ^ 1][w][1][lane] = make_double2(x[kRows - 1][0], x[kRows - 1][1]);
    __syncthreads();
  }
  // the exact centre: tile rows H .. kTileH - H - 1, columns H .. kTile - H - 1
  if (kRows * (w + 1) <= H || kRows * w >= kTileH - H) return;
  if (pl.base) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const size_t row = (size_t)min(max(j0 + r, 0), Ny - 1) * ld;
      x[r][0] = pl.base[row + ic0] - x[r][0];
      x[r][1] = pl.base[row + ic1] - x[r][1];
    }
  }
  const bool cw[2] = {cok[0] && 2 * lane >= H && 2 * lane < kTile - H,
                      cok[1] && 2 * lane + 1 >= H && 2 * lane + 1 < kTile - H};
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int lr = kRows * w + r;
    if (!((rok >> r) & 1u) || lr < H || lr >= kTileH - H) continue;
    const size_t row = (size_t)(j0 + r) * ld;
    if (cw[0]) pl.out[row + i0] = x[r][0];
    if (cw[1]) pl.out[row + i0 + 1] = x[r][1];
  }
}

}  // namespace

hipError_t launch_gs_helmholtz(GsBatch& bt, int nplanes, hipStream_t st) {
  if (bt.ld == 0) bt.ld = bt.Nx;
  if (bt.ld < bt.Nx || bt.Nx < 1 || bt.Ny < 1 || nplanes < 1) return hipErrorInvalidValue;
  // tiles of one plane: 128 wide, rows tall
  const auto tiles_of = [&](const GsPlane& q, int rows, int& ntx, int& nty) {
    ntx = (bt.Nx + kTile - 2 * q.H - 1) / (kTile - 2 * q.H);
    nty = (bt.Ny + rows - 2 * q.H - 1) / (rows - 2 * q.H);
  };
  int total16 = 0;
  for (int p = 0; p < nplanes; ++p) {
    GsPlane& q = bt.p[p];
    if (q.nsweep < 1 || q.nsweep > kGsHalo) return hipErrorInvalidValue;
    // even: tile origins stay even, so colours are (column + row) & 1 in a tile
    q.H = q.nsweep + (q.nsweep & 1);
    int tx, ty;
    tiles_of(q, 16 * kWaves, tx, ty);
    total16 += tx * ty;
  }
  // fewer 128 x 128 tiles than CUs (a small region): 128 x 64 tiles
  const int rows = total16 < 256 ? 8 : 16;
  int tiles = 0;
  for (int p = 0; p < nplanes; ++p) {
    GsPlane& q = bt.p[p];
    tiles_of(q, rows * kWaves, q.ntx, q.nty);
    tiles = q.ntx * q.nty > tiles ? q.ntx * q.nty : tiles;
  }
  bt.ntx = bt.nty = 0;
  bt.skip = g_skip;
  if (rows == 8)
    hipLaunchKernelGGL(gs_helmholtz_kernel<8>, dim3(tiles, nplanes), dim3(kWaves * 64), 0, st, bt);
  else
    hipLaunchKernelGGL(gs_helmholtz_kernel<16>, dim3(tiles, nplanes), dim3(kWaves * 64), 0, st, bt);
  return hipGetLastError();
}

}  // namespace ibpm
