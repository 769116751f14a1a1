// FP64 matrix-core products C (+)= A B of the box-restricted spectral updates: the
// forward sine transform of a change of Gamma that vanishes outside a box of the plane,
// S_x[:, I] D[I, J] S_y[J, :], as two dense products of the box with the sine-matrix
// columns it touches (ibpm.hip, box_update). Column-major operands; A or B may be given
// as the difference of two matrices of one leading dimension (X - X2, the change of
// Gamma read from the two planes it lies between).
//
// Tile: 64 x 64 of C per work-group, 4 waves of 32 x 32 (2 x 2 v_mfma_f64_16x16x4), K in
// chunks of 32 staged through LDS with the next chunk's loads in flight. The matrix core
// takes B^T as its first operand so that a lane's accumulator entries run along the
// column of C (16 consecutive rows per store). Each C entry sums its K terms in one fixed
// order, so results do not depend on scheduling.
#include <string.h>

#include <algorithm>

#include "ibpm_device.h"
#include "ibpm_internal.h"

namespace ibpm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int GT = 64;  // C tile edge
constexpr int GK = 32;  // K chunk (the next chunk's loads fly during this one's products)
constexpr int GR = GT * GK / 256;  // elements of A (and of B) each thread stages per chunk

// DA, DB: some job of the launch has a difference term A2 (B2); without, the operand is
// loaded once.
template <bool DA, bool DB>
__global__ __launch_bounds__(256) void gemm_f64_kernel(const GemmBatch gb) {
  if (skipped(gb.skip)) return;
  // the job's fields in registers (uniform), read once; a missing difference term reads
  // the matrix itself with weight 0 (a - 0 * a = a exactly), so no load is conditional
  const GemmJob& jb = gb.j[blockIdx.y];
  const double *A = jb.A, *B = jb.B;
  const double* A2 = jb.A2 ? jb.A2 : jb.A;
  const double* B2 = jb.B2 ? jb.B2 : jb.B;
  const double wa = jb.A2 ? 1.0 : 0.0, wb = jb.B2 ? 1.0 : 0.0;
  double* C = jb.C;
  const int lda = jb.lda, ldb = jb.ldb, ldc = jb.ldc, M = jb.M, N = jb.N, K = jb.K;
  const bool accumulate = jb.accumulate != 0;
  const int ntm = (M + GT - 1) / GT, ntn = (N + GT - 1) / GT;
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

  // A chunk: 64 m x GK k, thread -> m = tid % 64, k = tid / 64 + 4r (m contiguous)
  // B chunk: GK k x 64 n, thread -> k = tid % GK, n = tid / GK + (256 / GK) r (k contiguous)
  // Every load is unconditional, its indices clamped into the matrix (the value is zeroed
  // when staged), so the loads issue back to back; the two terms of a difference are
  // subtracted when staged, after the current chunk's products.
  constexpr int NS = 256 / GK;
  static_assert(256 % GK == 0 && GT * GK == 256 * GR && GT == 64, "chunk mapping");
  const int tm = tid & 63, tq = tid >> 6, tkb = tid % GK, tnb = tid / GK;
  const size_t am = (size_t)min(m0 + tm, M - 1);
  const bool mok = m0 + tm < M;
  double ra[GR], ra2[GR], rb[GR], rb2[GR];
#define GEMM_FETCH(k0)                                                            \
  {                                                                               \
    const size_t bk_ = (size_t)min((k0) + tkb, K - 1);                            \
    _Pragma("unroll") for (int r = 0; r < GR; ++r) {                              \
      const size_t o = am + (size_t)min((k0) + tq + 4 * r, K - 1) * lda;          \
      ra[r] = A[o];                                                               \
      if (DA) ra2[r] = A2[o];                                                     \
    }                                                                             \
    _Pragma("unroll") for (int r = 0; r < GR; ++r) {                              \
      const size_t o = bk_ + (size_t)min(n0 + tnb + NS * r, N - 1) * ldb;         \
      rb[r] = B[o];                                                               \
      if (DB) rb2[r] = B2[o];                                                     \
    }                                                                             \
  }
  GEMM_FETCH(0)
  for (int k0 = 0; k0 < K; k0 += GK) {
    const bool kbok = k0 + tkb < K;
#pragma unroll
    for (int r = 0; r < GR; ++r) {
      const bool aok = mok && k0 + tq + 4 * r < K, bok = kbok && n0 + tnb + NS * r < N;
      const double a = DA ? fma(-wa, ra2[r], ra[r]) : ra[r];
      const double b = DB ? fma(-wb, rb2[r], rb[r]) : rb[r];
      As[tq + 4 * r][tm] = aok ? a : 0.0;
      Bs[tkb][tnb + NS * r] = bok ? b : 0.0;
    }
    __syncthreads();
    if (k0 + GK < K) GEMM_FETCH(k0 + GK)
    const int ksteps = (min(K - k0, GK) + 3) / 4;  // zero-padded past K: skip those steps
    for (int ks = 0; ks < ksteps; ++ks) {
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
#undef GEMM_FETCH
  // C entries of this thread (accumulate), loaded after the products: not held in
  // registers through them (other work-groups on the CU cover the wait)
  double cin[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = min(m0 + wm * 32 + j * 16 + col, M - 1);
        const int n = min(n0 + wn * 32 + i * 16 + kq + 4 * r, N - 1);
        cin[i][j][r] = accumulate ? C[(size_t)m + (size_t)n * ldc] : 0.0;
      }
  // acc[i][j][r] at lane: C(m = m0 + wm*32 + j*16 + col, n = n0 + wn*32 + i*16 + kq + 4r)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int m = m0 + wm * 32 + j * 16 + col;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * 32 + i * 16 + kq + 4 * r;
        if (m < M && n < N) C[(size_t)m + (size_t)n * ldc] = cin[i][j][r] + acc[i][j][r];
      }
    }
}

// out[m + n ldo] = sum over s of P[s stride + m + n ldp], the slices in order; one job
// per blockIdx.z
__global__ __launch_bounds__(256) void sum_slices_kernel(const SliceSumBatch sb) {
  if (skipped(sb.skip)) return;
  const SliceSumJob& jb = sb.j[blockIdx.z];
  const int m = blockIdx.x * 64 + (threadIdx.x & 63);
  const int n = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (m >= jb.M || n >= jb.N) return;
  const double* p = jb.P + (size_t)m + (size_t)n * jb.ldp;
  const int nslice = jb.nslice;
  double v[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) v[s] = s < nslice ? p[(size_t)s * jb.stride] : 0.0;
  double sum = v[0];
#pragma unroll
  for (int s = 1; s < 16; ++s)
    if (s < nslice) sum += v[s];
  jb.out[(size_t)m + (size_t)n * jb.ldo] = sum;
}

}  // namespace

hipError_t launch_sum_slices_batch(const SliceSumBatch& sb0, int njobs, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  if (njobs > kMaxPlanes) return hipErrorInvalidValue;
  SliceSumBatch sb = sb0;
  sb.skip = g_skip;
  int mx = 0, nx = 0;
  for (int i = 0; i < njobs; ++i) {
    const SliceSumJob& j = sb.j[i];
    if (j.nslice < 1 || j.nslice > 16 || j.M <= 0 || j.N <= 0) return hipErrorInvalidValue;
    mx = std::max(mx, j.M);
    nx = std::max(nx, j.N);
  }
  hipLaunchKernelGGL(sum_slices_kernel, dim3((mx + 63) / 64, (nx + 3) / 4, njobs), dim3(256), 0, st, sb);
  return hipGetLastError();
}

hipError_t launch_sum_slices(const double* P, size_t stride, int nslice, int ldp, int M, int N,
                             double* out, int ldo, hipStream_t st) {
  SliceSumBatch sb;
  memset(&sb, 0, sizeof sb);
  sb.j[0] = SliceSumJob{P, stride, nslice, ldp, M, N, out, ldo};
  return launch_sum_slices_batch(sb, 1, st);
}

hipError_t launch_gemm_f64(const GemmBatch& gb0, int njobs, hipStream_t st) {
  GemmBatch gb = gb0;
  gb.skip = g_skip;
  int tiles = 0;
  bool da = false, db = false;
  for (int i = 0; i < njobs; ++i) {
    const GemmJob& j = gb.j[i];
    if (j.M <= 0 || j.N <= 0 || j.K <= 0) return hipErrorInvalidValue;
    const int t = ((j.M + GT - 1) / GT) * ((j.N + GT - 1) / GT);
    tiles = t > tiles ? t : tiles;
    da = da || j.A2;
    db = db || j.B2;
  }
  if (njobs <= 0) return hipSuccess;
  const dim3 grid(tiles, njobs);
  if (da && db)
    hipLaunchKernelGGL((gemm_f64_kernel<true, true>), grid, dim3(256), 0, st, gb);
  else if (da)
    hipLaunchKernelGGL((gemm_f64_kernel<true, false>), grid, dim3(256), 0, st, gb);
  else if (db)
    hipLaunchKernelGGL((gemm_f64_kernel<false, true>), grid, dim3(256), 0, st, gb);
  else
    hipLaunchKernelGGL((gemm_f64_kernel<false, false>), grid, dim3(256), 0, st, gb);
  return hipGetLastError();
}

}  // namespace ibpm
