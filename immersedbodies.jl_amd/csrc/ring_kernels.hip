// Multi-level Poisson chain of vort2flux (fluid.jl:175-206) in spectral form.
//
// The reference solves level L, reads its streamfunction on the ring where the next finer
// level's boundary lies (get_bc!), adds those values to the finer right-hand side
// (apply_bc!) and solves again, L times in sequence. By linearity the same triangular
// system is solved here as
//   F_l      = S_x Gamma_l S_y                        (all levels in one batched pass pair)
//   psi_l    = S_x (D . (F_l + R_l)) S_y              (all levels in one batched pass pair)
//   R_l      = S_x (apply_bc of ring r_l) S_y = uB sy1^T + uT syN^T + sx1 vL^T + sxN vR^T
// where only the rank-4 ring terms travel level to level:
//   a_b, a_t = E_l^T S_y[:, ny/4], E_l^T S_y[:, 3ny/4]   (projections onto the ring rows)
//   c_l, c_r = S_x[nx/4, :] E_l,  S_x[3nx/4, :] E_l      (and ring columns)
//   uB, uT   = M_x a_b, M_x a_t;   vL, vR = M_y c_l, M_y c_r   (factors of level l-1)
// M_n = S_n . Interp . Ring . S_n maps a projection to the next finer level's factor: the
// ring values are the DST-I of the projection at the ring positions n/4..3n/4, get_bc!'s
// interpolation (fluid.jl:285-303) turns them into the apply_bc! edge vector
// (fluid.jl:319-335), and the factor is that edge's DST-I. M_n is a fixed N x N matrix,
// built once per solver with one FP64 sine pass (build_ring_operator).
//
// Per level the chain is three short launches: a tiled sweep of the spectral plane
// (E = D . (F + R) written to a separate plane, and per-tile partial projections), the
// reduction of the partials, and one matrix-vector product (a wave per row). Every sum runs
// in a fixed order: results do not depend on scheduling.
//
// D = scale / lambda with lambda = -2 ((cos(pi kx/nx) + cos(pi ky/ny)) - 2), evaluated
// from the two cosine vectors exactly as the reference's _lap_eigs (fluid.jl:168-173).
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <vector>

#include "ibpm_device.h"
#include "ibpm_internal.h"

namespace ibpm {

constexpr int RT = 64;   // tile width (kx) of the sweep
constexpr int RTY = 16;  // tile height (ky): 4 rows per wave

// One 64 x 16 tile of the spectral plane per block (lanes along kx, each wave 4 rows):
// E = D . (F + R) written to E and, with do_proj, the tile's partial projections
//   part_a[by][side][kx] = sum over the tile's ky of E S_y[ky, ny/4 | 3ny/4]
//   part_c[bx][side][ky] = sum over the tile's kx of S_x[nx/4 | 3nx/4, kx] E
// (completed by ring_reduce_kernel; every sum runs in a fixed order). Short tiles give
// four blocks per CU on a 1023^2 plane, enough loads in flight to cover the latency.
// F32: the spectral planes F (and E) in FP32 (the mixed path).
template <bool F32>
__global__ __launch_bounds__(256) void ring_sweep_kernel(const RingChainArgs a, int do_proj) {
  __shared__ double yv[7][RTY];          // cosy, sy1, syN, vL, vR, wb, wt of the tile's rows
  __shared__ double rs[2][RTY][RT + 1];  // S_x-weighted E of the tile, per row, for the row sums
  __shared__ double pa[2][4][RT];        // per-wave partial a_b, a_t
  __shared__ double part[2][8][RTY];     // row sums over eighths of the tile's kx
  if (skipped(a.skip)) return;
  const int c = blockIdx.z;
  const int Nx = a.Nx, Ny = a.Ny;
  const int tid = threadIdx.x, tx = tid & 63, w = tid >> 6;
  const int bx = blockIdx.x, by = blockIdx.y;
  const int kx = bx * RT + tx + 1;
  const int ky0 = by * RTY + 1;
  const bool okx = kx <= Nx;
  const double* __restrict__ fac = a.fac_in[c];
  // Loads in the order their values are needed, all issued before the first wait: the
  // row tables of row tid % RTY (stored by the first RTY threads), then the plane and
  // column values. Every index is clamped into range (out-of-range values unused), and
  // every load is unconditional: without ring factors (the coarsest level) they are read
  // from the edge vectors, of the same layout, and replaced by zeros (a branch around
  // the loads would make the compiler wait for the earlier ones at its end).
  const bool hf = fac != nullptr;
  const double* __restrict__ fsrc = hf ? fac : a.edge;
  const int yr = tid & (RTY - 1);
  const int kyr = min(ky0 + yr, Ny) - 1;
  const double y0 = a.cosy[kyr], y5 = a.pw[kyr], y6 = a.pw[Ny + kyr];
  const double y1 = a.edge[2 * Nx + kyr], y2 = a.edge[2 * Nx + Ny + kyr];
  const double y3 = fsrc[2 * Nx + kyr], y4 = fsrc[2 * Nx + Ny + kyr];
  const int kc = (okx ? kx : Nx) - 1;
  void* __restrict__ E = a.E[c];
  double fv[RTY / 4];
#pragma unroll
  for (int m = 0; m < RTY / 4; ++m) {
    const size_t o = (size_t)kc + (size_t)(min(ky0 + w * (RTY / 4) + m, Ny) - 1) * Nx;
    fv[m] = F32 ? (double)((const float*)a.F[c])[o] : ((const double*)a.F[c])[o];
  }
  const double cx_k = a.cosx[kc], sxl_k = a.pw[2 * Ny + kc], sxr_k = a.pw[2 * Ny + Nx + kc];
  const double fuB = hf ? fsrc[kc] : 0.0, fuT = hf ? fsrc[Nx + kc] : 0.0;
  const double e1 = hf ? a.edge[kc] : 0.0, eN = hf ? a.edge[Nx + kc] : 0.0;
  if (tid < RTY) {
    const bool ok = ky0 + tid <= Ny;
    yv[0][tid] = ok ? y0 : 0.0;
    yv[1][tid] = ok && hf ? y1 : 0.0;
    yv[2][tid] = ok && hf ? y2 : 0.0;
    yv[3][tid] = ok && hf ? y3 : 0.0;
    yv[4][tid] = ok && hf ? y4 : 0.0;
    yv[5][tid] = ok ? y5 : 0.0;
    yv[6][tid] = ok ? y6 : 0.0;
  }
  __syncthreads();
  const double cxk = okx ? cx_k : 0.0, sxl = okx ? sxl_k : 0.0, sxr = okx ? sxr_k : 0.0;
  double ab = 0.0, at = 0.0;
#pragma unroll
  for (int m = 0; m < RTY / 4; ++m) {
    const int yy = w * (RTY / 4) + m;
    const int ky = ky0 + yy;
    double e = 0.0;
    if (okx && ky <= Ny) {
      const size_t idx = (size_t)(kx - 1) + (size_t)(ky - 1) * Nx;
      double f = fv[m];
      if (fac) f += fuB * yv[1][yy] + fuT * yv[2][yy] + e1 * yv[3][yy] + eN * yv[4][yy];
      const double lam = -2.0 * ((cxk + yv[0][yy]) - 2.0);
      e = a.scale / lam * f;
      if (a.write_e) {
        if (F32) ((float*)E)[idx] = (float)e;
        else ((double*)E)[idx] = e;
      }
      ab += e * yv[5][yy];
      at += e * yv[6][yy];
    }
    rs[0][yy][tx] = sxl * e;
    rs[1][yy][tx] = sxr * e;
  }
  if (!do_proj) return;
  pa[0][w][tx] = ab;
  pa[1][w][tx] = at;
  __syncthreads();
  if (w < 2 && okx)
    a.part_a[c][((size_t)by * 2 + w) * Nx + (kx - 1)] = ((pa[w][0][tx] + pa[w][1][tx]) + pa[w][2][tx]) + pa[w][3][tx];
  {  // row sums: thread -> (row, side, eighth of the kx range)
    const int row = tid & (RTY - 1), side = (tid >> 4) & 1, e8 = tid >> 5;
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < RT / 8; ++q) sum += rs[side][row][e8 * (RT / 8) + q];
    part[side][e8][row] = sum;
  }
  __syncthreads();
  if (tid < 2 * RTY) {
    const int row = tid & (RTY - 1), side = tid >> 4;
    const int ky = ky0 + row;
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) sum += part[side][q][row];
    if (ky <= Ny) a.part_c[c][((size_t)bx * 2 + side) * Ny + (ky - 1)] = sum;
  }
}

// Complete projections from the per-tile partials. Four lanes share one projection entry,
// each summing a quarter of the tiles in order (all its loads in flight together), and
// the quarters are added in order: a fixed summation order.
__global__ __launch_bounds__(256) void ring_reduce_kernel(const RingChainArgs a, int ntx, int nty) {
  if (skipped(a.skip)) return;
  const int c = blockIdx.y;
  const int Nx = a.Nx, Ny = a.Ny;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int i = t >> 2, qt = t & 3;  // entry, quarter of the tiles
  const bool ok = i < 2 * Nx + 2 * Ny;
  const double* src;
  size_t stride;
  int nt;
  if (!ok || i < 2 * Nx) {
    const int ii = ok ? i : 0, side = ii / Nx, x = ii % Nx;
    src = a.part_a[c] + (size_t)side * Nx + x;
    stride = 2 * (size_t)Nx;
    nt = nty;
  } else {
    const int j = i - 2 * Nx, side = j / Ny, y = j % Ny;
    src = a.part_c[c] + (size_t)side * Ny + y;
    stride = 2 * (size_t)Ny;
    nt = ntx;
  }
  const int per = (nt + 3) / 4, q0 = qt * per, q1 = min(nt, q0 + per);
  double sum = 0.0;
  for (int b = q0; b < q1; b += 16) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = src[(size_t)min(b + q, q1 - 1) * stride];  // clamped: unconditional
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (b + q < q1) sum += v[q];
  }
  // quarters 0..3 of the entry sit in consecutive lanes: ((s0 + s1) + s2) + s3
  const double s1 = __shfl_down(sum, 1, 64), s2 = __shfl_down(sum, 2, 64), s3 = __shfl_down(sum, 3, 64);
  if (ok && qt == 0) a.vec[c][i] = ((sum + s1) + s2) + s3;
}

// Factors of the next finer level: uB = M_x a_b, uT = M_x a_t, vL = M_y c_l, vR = M_y c_r.
// One wave per row j of M (row-major, j contiguous over the input index); 2 right-hand
// sides per wave.
__global__ __launch_bounds__(256) void ring_factor_kernel(const RingChainArgs a) {
  if (skipped(a.skip)) return;
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const bool ydir = blockIdx.y == 1;
  const int N = ydir ? a.Ny : a.Nx;
  if (j >= N) return;  // whole waves
  const int c = blockIdx.z;
  const double* M = (ydir ? a.My : a.Mx) + (size_t)j * N;
  const double* v0 = a.vec[c] + (ydir ? 2 * (size_t)a.Nx : 0);
  const double* v1 = v0 + N;
  double s0 = 0.0, s1 = 0.0;
  // 16 elements per lane per chunk, every load of the chunk issued before the products
  for (int k0 = 0; k0 < N; k0 += 1024) {
    double m[16], x0[16], x1[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = k0 + lane + 64 * q;
      const int kc = k < N ? k : 0;
      m[q] = M[kc];
      x0[q] = v0[kc];
      x1[q] = v1[kc];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      if (k0 + lane + 64 * q < N) {
        s0 += m[q] * x0[q];
        s1 += m[q] * x1[q];
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, 64);
    s1 += __shfl_down(s1, off, 64);
  }
  if (lane == 0) {
    double* out = a.fac_out[c] + (ydir ? 2 * (size_t)a.Nx : 0);
    out[j] = s0;
    out[N + j] = s1;
  }
}

hipError_t launch_ring_sweep(const RingChainArgs& a0, int nchain, bool do_proj, hipStream_t st) {
  RingChainArgs a = a0;
  a.skip = g_skip;
  const int ntx = (a.Nx + RT - 1) / RT, nty = (a.Ny + RTY - 1) / RTY;
  if (a.F_f32)
    hipLaunchKernelGGL(ring_sweep_kernel<true>, dim3(ntx, nty, nchain), dim3(256), 0, st, a, do_proj ? 1 : 0);
  else
    hipLaunchKernelGGL(ring_sweep_kernel<false>, dim3(ntx, nty, nchain), dim3(256), 0, st, a, do_proj ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !do_proj) return e;
  const int nv = 2 * a.Nx + 2 * a.Ny;
  hipLaunchKernelGGL(ring_reduce_kernel, dim3((4 * nv + 255) / 256, nchain), dim3(256), 0, st, a, ntx, nty);
  return hipGetLastError();
}

size_t ring_part_a_count(int Nx, int Ny) { return (size_t)2 * ((Ny + RTY - 1) / RTY) * Nx; }

size_t ring_partial_count(int Nx, int Ny) {
  const int ntx = (Nx + RT - 1) / RT;
  return ring_part_a_count(Nx, Ny) + (size_t)2 * ntx * Ny;
}

hipError_t launch_ring_factor(const RingChainArgs& a0, int nchain, hipStream_t st) {
  RingChainArgs a = a0;
  a.skip = g_skip;
  const int rows = std::max(a.Nx, a.Ny);
  hipLaunchKernelGGL(ring_factor_kernel, dim3((rows + 3) / 4, 2, nchain), dim3(256), 0, st, a);
  return hipGetLastError();
}

size_t ring_vector_count(int Nx, int Ny) { return 2 * (size_t)Nx + 2 * (size_t)Ny; }

// B[k][c] = row k of Interp . Ring . S_n applied to unit vector c, k = edge element, stored
// k-contiguous per column c, so that one sine pass along k gives M = S_n B row-major.
void build_ring_operator_input(int n, double* B) {
  const int N = n - 1;
  auto s2 = [n](long long j, long long m) {
    return 2.0 * sin(M_PI * (double)((j * m) % (2LL * n)) / (double)n);
  };
  for (int cc = 0; cc < N; ++cc) {
    for (int k = 0; k < N; ++k) {
      const int bc = k + 2;  // 1-based boundary position of edge element k+1
      double v;
      if (bc & 1) {
        v = s2(n / 4 + (bc - 1) / 2, cc + 1);
      } else {
        v = 0.5 * (s2(n / 4 + bc / 2, cc + 1) + s2(n / 4 + bc / 2 - 1, cc + 1));
      }
      B[(size_t)k + (size_t)cc * N] = v;
    }
  }
}

}  // namespace ibpm
