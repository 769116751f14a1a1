// FP64 matrix-core products C (+)= A B of the box-restricted spectral updates: the
// forward sine transform of a change of Gamma that vanishes outside a box of the plane,
// S_x[:, I] D[I, J] S_y[J, :], as two dense products of the box with the sine-matrix
// columns it touches (ibpm.hip, box_update). Column-major operands; A or B may be given
// as the difference of two matrices of one leading dimension (X - X2, the change of
// Gamma read from the two planes it lies between).
//
// Tile: 64 x 64 of C per work-group, 4 waves of 32 x 32 (2 x 2 v_mfma_f64_16x16x4), K in
// chunks of 16 staged through LDS with the next chunk's loads in flight. The matrix core
// takes B^T as its first operand so that a lane's accumulator entries run along the
// column of C (16 consecutive rows per store). Each C entry sums its K terms in one fixed
// order, so results do not depend on scheduling.
#include "ibpm_internal.h"

namespace ibpm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int GT = 64;  // C tile edge
constexpr int GK = 16;  // K chunk

__global__ __launch_bounds__(256) void gemm_f64_kernel(const GemmBatch gb) {
  const GemmJob& jb = gb.j[blockIdx.y];
  const int ntm = (jb.M + GT - 1) / GT, ntn = (jb.N + GT - 1) / GT;
  if ((int)blockIdx.x >= ntm * ntn) return;  // whole work-group: barriers stay uniform
  const int m0 = (blockIdx.x % ntm) * GT, n0 = (blockIdx.x / ntm) * GT;
  __shared__ double As[GK][GT + 1];  // As[k][m]
  __shared__ double Bs[GK][GT + 1];  // Bs[k][n]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int col = lane & 15, kq = lane >> 4;

  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

  // A chunk: 64 m x 16 k, thread -> m = tid % 64, k = tid / 64 + 4r (m contiguous)
  // B chunk: 16 k x 64 n, thread -> k = tid % 16, n = tid / 16 + 16r (k contiguous)
  double ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + (tid & 63), k = k0 + (tid >> 6) + 4 * r;
      double v = 0.0;
      if (m < jb.M && k < jb.K) {
        const size_t o = (size_t)m + (size_t)k * jb.lda;
        v = jb.A[o];
        if (jb.A2) v -= jb.A2[o];
      }
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = k0 + (tid & 15), n = n0 + (tid >> 4) + 16 * r;
      double v = 0.0;
      if (k < jb.K && n < jb.N) {
        const size_t o = (size_t)k + (size_t)n * jb.ldb;
        v = jb.B[o];
        if (jb.B2) v -= jb.B2[o];
      }
      rb[r] = v;
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < jb.K; k0 += GK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      As[(tid >> 6) + 4 * r][tid & 63] = ra[r];
      Bs[tid & 15][(tid >> 4) + 16 * r] = rb[r];
    }
    __syncthreads();
    if (k0 + GK < jb.K) fetch(k0 + GK);
#pragma unroll
    for (int ks = 0; ks < GK / 4; ++ks) {
      const int kr = ks * 4 + kq;
      double av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = Bs[kr][wn * 32 + i * 16 + col];  // (n = col, k)
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = As[kr][wm * 32 + j * 16 + col];  // (k, m = col)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // acc[i][j][r] at lane: C(m = m0 + wm*32 + j*16 + col, n = n0 + wn*32 + i*16 + kq + 4r)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int m = m0 + wm * 32 + j * 16 + col;
      if (m >= jb.M) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * 32 + i * 16 + kq + 4 * r;
        if (n >= jb.N) continue;
        double* c = jb.C + (size_t)m + (size_t)n * jb.ldc;
        *c = jb.accumulate ? *c + acc[i][j][r] : acc[i][j][r];
      }
    }
}

}  // namespace

hipError_t launch_gemm_f64(const GemmBatch& gb, int njobs, hipStream_t st) {
  int tiles = 0;
  for (int i = 0; i < njobs; ++i) {
    const GemmJob& j = gb.j[i];
    if (j.M <= 0 || j.N <= 0 || j.K <= 0) return hipErrorInvalidValue;
    const int t = ((j.M + GT - 1) / GT) * ((j.N + GT - 1) / GT);
    tiles = t > tiles ? t : tiles;
  }
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(gemm_f64_kernel, dim3(tiles, njobs), dim3(256), 0, st, gb);
  return hipGetLastError();
}

}  // namespace ibpm
