// Multi-level Poisson chain of vort2flux (fluid.jl:175-206) in spectral form.
//
// The reference solves level L, reads its streamfunction on the ring where the next finer
// level's boundary lies (get_bc!), adds those values to the finer right-hand side
// (apply_bc!) and solves again, L times in sequence. By linearity the same triangular
// system is solved here as
//   F_l      = S_x Gamma_l S_y                        (all levels in one batched pass pair)
//   psi_l    = S_x (D . (F_l + R_l)) S_y              (all levels in one batched pass pair)
//   R_l      = S_x (apply_bc of ring r_l) S_y = uB sy1^T + uT syN^T + sx1 vL^T + sxN vR^T
//   r_{l-1}  = get_bc(psi_l), read from the spectral plane by projection
// where only the rank-4 ring terms travel level to level. Per level the chain reads the
// spectral plane once (and writes E = D . (F + R) back in place for the inverse passes)
// and evaluates a few thousand length-N sine sums, all in FP64.
//
// D = scale / lambda with lambda = -2 ((cos(pi kx/nx) + cos(pi ky/ny)) - 2), evaluated
// from the two cosine vectors exactly as the reference's _lap_eigs (fluid.jl:168-173).
// Length-N sine sums use the folded FP64 tables of the sine-contraction passes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ibpm_internal.h"

namespace ibpm {

constexpr int RP_TX = 64;  // kx per projection block
constexpr int RP_TY = 64;  // ky per projection block

// E = D . (F + R) in place on the spectral plane; with do_proj also the partial
// projections of E onto the four ring lines:
//   part_a[0|1][by][kx] = sum_ky E[kx,ky] S_y[ky, ny/4 | 3ny/4]
//   part_c[0|1][bx][ky] = sum_kx S_x[nx/4 | 3nx/4, kx] E[kx,ky]
// with the projection weights pw = [S_y[:, ny/4], S_y[:, 3ny/4], S_x[nx/4, :], S_x[3nx/4, :]].
__global__ __launch_bounds__(256) void ring_project_kernel(const RingChainArgs a, int do_proj) {
  __shared__ double cl[RP_TY][RP_TX + 1];
  __shared__ double cr[RP_TY][RP_TX + 1];
  __shared__ double pa[2][4][RP_TX];
  const int c = blockIdx.z;
  const int Nx = a.Nx, Ny = a.Ny;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int kx = blockIdx.x * RP_TX + tx + 1;
  const bool okx = kx <= Nx;
  const double* wb = a.pw;
  const double* wt = a.pw + Ny;
  const double sxl = okx ? a.pw[2 * Ny + kx - 1] : 0.0;
  const double sxr = okx ? a.pw[2 * Ny + Nx + kx - 1] : 0.0;
  const double cxk = okx ? a.cosx[kx - 1] : 0.0;
  const double* fac = a.fac_in[c];
  double fuB = 0, fuT = 0, e1 = 0, eN = 0;
  if (fac && okx) {
    fuB = fac[kx - 1];
    fuT = fac[Nx + kx - 1];
    e1 = a.edge[kx - 1];
    eN = a.edge[Nx + kx - 1];
  }
  double ab = 0.0, at = 0.0;
  for (int m = 0; m < RP_TY / 4; ++m) {
    const int yy = ty + 4 * m;
    const int ky = blockIdx.y * RP_TY + yy + 1;
    double e = 0.0;
    if (okx && ky <= Ny) {
      const size_t idx = (size_t)(kx - 1) + (size_t)(ky - 1) * Nx;
      double f = a.F_f32 ? (double)((const float*)a.F[c])[idx] : ((const double*)a.F[c])[idx];
      if (fac)
        f += fuB * a.edge[2 * Nx + ky - 1] + fuT * a.edge[2 * Nx + Ny + ky - 1] +
             e1 * fac[2 * Nx + ky - 1] + eN * fac[2 * Nx + Ny + ky - 1];
      const double lam = -2.0 * ((cxk + a.cosy[ky - 1]) - 2.0);
      e = a.scale / lam * f;
      if (a.F_f32)
        ((float*)a.F[c])[idx] = (float)e;
      else
        ((double*)a.F[c])[idx] = e;
      ab += e * wb[ky - 1];
      at += e * wt[ky - 1];
    }
    if (do_proj) {
      cl[yy][tx] = sxl * e;
      cr[yy][tx] = sxr * e;
    }
  }
  if (!do_proj) return;
  pa[0][ty][tx] = ab;
  pa[1][ty][tx] = at;
  __syncthreads();
  if (ty < 2 && okx) {
    const double s = pa[ty][0][tx] + pa[ty][1][tx] + pa[ty][2][tx] + pa[ty][3][tx];
    a.part_a[c][((size_t)ty * gridDim.y + blockIdx.y) * Nx + (kx - 1)] = s;
  }
  // column sums over the block's kx for each of its ky: thread -> (side, yy)
  if (threadIdx.x < 2 * RP_TY) {
    const int side = threadIdx.x / RP_TY, yy = threadIdx.x % RP_TY;
    const int ky = blockIdx.y * RP_TY + yy + 1;
    double s = 0.0;
    for (int t = 0; t < RP_TX; ++t) s += side ? cr[yy][t] : cl[yy][t];
    if (ky <= Ny) a.part_c[c][((size_t)side * gridDim.x + blockIdx.x) * Ny + (ky - 1)] = s;
  }
}

// Slice of the DST-I output j (1-based) over k = k0, k0 + step, ... of the folded sum, from
// the folded vector (P, Q) in LDS and the folded FP64 tables (odd rows So[k][m], even rows
// Se[k][m], m contiguous): lanes with consecutive m read consecutive table entries.
__device__ __forceinline__ double dst1_slice(const double* __restrict__ So,
                                             const double* __restrict__ Se, int H,
                                             const double* P, const double* Q, int j, int k0,
                                             int step) {
  double s = 0.0;
  if (j & 1) {
    const int m = (j - 1) / 2;
    for (int k = k0; k <= H; k += step) s += So[(size_t)(k - 1) * H + m] * P[k - 1];
  } else {
    const int m = j / 2 - 1;
    for (int k = k0; k <= H - 1; k += step) s += Se[(size_t)(k - 1) * (H - 1) + m] * Q[k - 1];
  }
  return s;
}

// Fold x (length N = 2H-1, in LDS) into P (H) and Q (H-1): P_k = x_k + x_{M-k},
// Q_k = x_k - x_{M-k}, P_H = x_H.
__device__ __forceinline__ void fold_lds(const double* x, int N, double* P, double* Q) {
  const int H = (N + 1) / 2;
  for (int k = threadIdx.x + 1; k <= H; k += blockDim.x) {
    if (k == H) {
      P[k - 1] = x[k - 1];
    } else {
      P[k - 1] = x[k - 1] + x[N - k];
      Q[k - 1] = x[k - 1] - x[N - k];
    }
  }
}

// The ring lines are rows of a 1-D DST evaluated at 1-based positions j = n/4 .. 3n/4; the
// slot order puts the odd positions first and the even ones after, so a wave reads one
// table.
__device__ __forceinline__ int ring_position(int t, int n) {
  const int lo = n / 4, hi = 3 * n / 4;
  const int first_odd = lo | 1;
  const int nodd = first_odd > hi ? 0 : (hi - first_odd) / 2 + 1;
  if (t < nodd) return first_odd + 2 * t;
  const int first_even = (lo & 1) ? lo + 1 : lo;
  return first_even + 2 * (t - nodd);
}

constexpr int RS_OUT = 64;     // outputs per block (one per lane)
constexpr int RS_SLICES = 16;  // waves per block, each summing one slice of k

// Streamfunction on the four ring lines of level l from the projections:
//   line 0/1 (bottom/top): psi(i, ny/4 | 3ny/4) = (S_x a_{b|t})_i, i = nx/4..3nx/4
//   line 2/3 (left/right): psi(nx/4 | 3nx/4, j) = (S_y c_{l|r})_j, j = ny/4..3ny/4
__global__ __launch_bounds__(1024) void ring_values_kernel(const RingChainArgs a, int nchx, int nchy) {
  extern __shared__ double sh[];
  const int line = blockIdx.y, c = blockIdx.z;
  const bool xl = line < 2;
  const int n = xl ? a.nx : a.ny, N = n - 1, H = n / 2;
  double* vec = sh;                  // N
  double* P = vec + N;               // H
  double* Q = P + H;                 // H - 1
  double* red = Q + H;               // RS_SLICES x RS_OUT
  const int side = line & 1;
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    double s = 0.0;
    if (xl)
      for (int ch = 0; ch < nchy; ++ch) s += a.part_a[c][((size_t)side * nchy + ch) * N + k];
    else
      for (int ch = 0; ch < nchx; ++ch) s += a.part_c[c][((size_t)side * nchx + ch) * N + k];
    vec[k] = s;
  }
  __syncthreads();
  fold_lds(vec, N, P, Q);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = blockIdx.x * RS_OUT + lane;
  const int nout = n / 2 + 1;
  const int j = t < nout ? ring_position(t, n) : 1;
  red[w * RS_OUT + lane] =
      t < nout ? dst1_slice(xl ? a.SoX : a.SoY, xl ? a.SeX : a.SeY, H, P, Q, j, w + 1, RS_SLICES) : 0.0;
  __syncthreads();
  if (w == 0 && t < nout) {
    double s = 0.0;
    for (int q = 0; q < RS_SLICES; ++q) s += red[q * RS_OUT + lane];
    a.lv[c][line * a.lvstride + (j - n / 4)] = s;
  }
}

// Ring factors of the next finer level: the apply_bc! edge vectors (fluid.jl:319-335)
// eB[i] = rbc[(B+1)+i] etc. assembled from the ring lines by get_bc!'s interpolation
// (fluid.jl:285-303), then uB = S_x eB, uT = S_x eT, vL = S_y eL, vR = S_y eR.
__global__ __launch_bounds__(1024) void ring_factor_kernel(const RingChainArgs a) {
  extern __shared__ double sh[];
  const int v = blockIdx.y, c = blockIdx.z;  // v: 0 uB, 1 uT, 2 vL, 3 vR
  const bool xv = v < 2;
  const int n = xv ? a.nx : a.ny, N = n - 1, H = n / 2;
  double* e = sh;
  double* P = e + N;
  double* Q = P + H;
  double* red = Q + H;
  const double* lv = a.lv[c] + v * a.lvstride;  // line v: bottom, top, left, right
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    const int bc = k + 2;  // 1-based boundary position of edge element k+1
    e[k] = (bc & 1) ? lv[(bc - 1) / 2] : 0.5 * (lv[bc / 2] + lv[bc / 2 - 1]);
  }
  __syncthreads();
  fold_lds(e, N, P, Q);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = blockIdx.x * RS_OUT + lane;
  // odd outputs first: j = 2t+1 for t < H, then even j = 2(t-H+1)
  const int j = t < H ? 2 * t + 1 : 2 * (t - H + 1);
  red[w * RS_OUT + lane] =
      t < N ? dst1_slice(xv ? a.SoX : a.SoY, xv ? a.SeX : a.SeY, H, P, Q, j, w + 1, RS_SLICES) : 0.0;
  __syncthreads();
  if (w == 0 && t < N) {
    double s = 0.0;
    for (int q = 0; q < RS_SLICES; ++q) s += red[q * RS_OUT + lane];
    const int off = v == 0 ? 0 : v == 1 ? a.Nx : v == 2 ? 2 * a.Nx : 2 * a.Nx + a.Ny;
    a.fac_out[c][off + j - 1] = s;
  }
}

hipError_t launch_ring_project(const RingChainArgs& a, int nchain, bool do_proj, hipStream_t st) {
  const int nchx = (a.Nx + RP_TX - 1) / RP_TX, nchy = (a.Ny + RP_TY - 1) / RP_TY;
  hipLaunchKernelGGL(ring_project_kernel, dim3(nchx, nchy, nchain), dim3(256), 0, st, a,
                     do_proj ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_ring_step(const RingChainArgs& a, int nchain, hipStream_t st) {
  const int nchx = (a.Nx + RP_TX - 1) / RP_TX, nchy = (a.Ny + RP_TY - 1) / RP_TY;
  hipError_t e = launch_ring_project(a, nchain, true, st);
  if (e) return e;
  const int nmax = std::max(a.nx, a.ny);
  const size_t shv = (size_t)(2 * nmax + RS_SLICES * RS_OUT) * sizeof(double);
  const int nout = nmax / 2 + 1;
  hipLaunchKernelGGL(ring_values_kernel, dim3((nout + RS_OUT - 1) / RS_OUT, 4, nchain), dim3(1024),
                     shv, st, a, nchx, nchy);
  e = hipGetLastError();
  if (e) return e;
  hipLaunchKernelGGL(ring_factor_kernel, dim3((nmax - 1 + RS_OUT - 1) / RS_OUT, 4, nchain),
                     dim3(1024), shv, st, a);
  return hipGetLastError();
}

size_t ring_partial_count(int Nx, int Ny) {
  const int nchx = (Nx + RP_TX - 1) / RP_TX, nchy = (Ny + RP_TY - 1) / RP_TY;
  return (size_t)2 * nchy * Nx + (size_t)2 * nchx * Ny;
}

size_t ring_partial_a_count(int Nx, int Ny) {
  const int nchy = (Ny + RP_TY - 1) / RP_TY;
  return (size_t)2 * nchy * Nx;
}

}  // namespace ibpm
