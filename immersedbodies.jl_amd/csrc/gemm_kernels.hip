// FP64 matrix-core products C (+)= A B of the box-restricted spectral updates: the
// forward sine transform of a change of Gamma that vanishes outside a box of the plane,
// S_x[:, I] D[I, J] S_y[J, :], as two dense products of the box with the sine-matrix
// columns it touches (ibpm.hip, box_update), and the trial psi on the panels' box.
// Column-major operands; A or B may be given as the difference of two matrices of one
// leading dimension (X - X2, the change of Gamma read from the two planes it lies
// between). The matrix core takes B^T as its first operand so that a lane's accumulator
// entries run along the column of C. Each C entry sums its K terms in one fixed order, so
// results do not depend on scheduling.
//
// Two kernels, picked per launch by its size:
//  - gemm_f64_kernel, for launches of many tiles (a full plane of C per job): 64 x 64 of
//    C per work-group, 4 waves of 32 x 32 (2 x 2 v_mfma_f64_16x16x4), K in chunks of 32
//    staged through LDS with the next chunk's loads in flight;
//  - gemm_f64_reg_kernel, for the small products (a few hundred tiles, K of tens to a few
//    hundred), which are latency-bound: 32 x 32 of C per work-group, one memory round trip
//    per K chunk of 16 KR values: the chunk is dealt to the 16 (wave, lane quarter) pairs,
//    each lane loads all its values of A and B straight into registers (no LDS staging),
//    and the four waves' partial tiles are summed in wave order through LDS.
#include <string.h>

#include <algorithm>

#include "ibpm_device.h"
#include "ibpm_internal.h"

namespace ibpm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// Tile order of a launch: each job's T tiles, m-tile major with the n-tiles fastest
// (consecutive tiles share rows of A), are cut into 8 contiguous ranges of near-equal
// size, [x T / 8, (x + 1) T / 8) for XCD x (each XCD has its own L2); work-groups are dealt
// round-robin to the XCDs, so work-group w walks the ranges of XCD w % 8 over the jobs in
// order, ceil(T / 8) slots per job (gb.lbase: the first slot of each job, filled by the
// launcher). Each XCD brings its rows of A into its L2 once, and every job, short or long
// in K, is spread over all XCDs.
template <int TM, int TN>
__device__ __forceinline__ bool gemm_tile(const GemmBatch& gb, int& job, int& m0, int& n0) {
  const int x = blockIdx.x & 7, l = blockIdx.x >> 3;
  job = 0;
#pragma unroll
  for (int i = 1; i < kMaxPlanes; ++i) job = (i < gb.njobs && l >= gb.lbase[i]) ? i : job;
  const GemmJob& jb = gb.j[job];
  const int ntn = (jb.N + TN - 1) / TN, T = ((jb.M + TM - 1) / TM) * ntn;
  const int u = x * T / 8 + (l - gb.lbase[job]);
  if (u >= (x + 1) * T / 8) return false;
  m0 = (u / ntn) * TM;
  n0 = (u % ntn) * TN;
  return true;
}

constexpr int GT = 64;  // C tile edge of gemm_f64_kernel
constexpr int GK = 32;  // K chunk (the next chunk's loads fly during this one's products)
constexpr int GR = GT * GK / 256;  // elements of A (and of B) each thread stages per chunk

// DA, DB: some job of the launch has a difference term A2 (B2); without, the operand is
// loaded once.
// MIR: a launch of mirror jobs (GemmJob::mirror): the two half-length sums over even and
// odd sine rows, written to the mirrored pairs of output columns.
template <bool DA, bool DB, bool MIR>
__global__ __launch_bounds__(256) void gemm_f64_kernel(const GemmBatch gb) {
  if (skipped(gb.skip)) return;
  int job, m0, n0;
  if (!gemm_tile<GT, GT>(gb, job, m0, n0)) return;  // whole work-group: barriers stay uniform
  // the job's fields in registers (uniform), read once; a missing difference term reads
  // the matrix itself with weight 0 (a - 0 * a = a exactly), so no load is conditional
  const GemmJob& jb = gb.j[job];
  const double *A = jb.A, *B = jb.B;
  const double* A2 = jb.A2 ? jb.A2 : jb.A;
  const double* B2 = jb.B2 ? jb.B2 : jb.B;
  const double wa = jb.A2 ? 1.0 : 0.0, wb = jb.B2 ? 1.0 : 0.0;
  double* C = jb.C;
  const int lda = jb.lda, ldb = jb.ldb, ldc = jb.ldc, M = jb.M, N = jb.N, K = jb.K;
  const bool accumulate = jb.accumulate != 0;
  constexpr bool mir = MIR;
  const int pu = jb.kpar & 1;  // mirror: local row parity of the even sine rows (U)
  __shared__ double As[GK][GT + 1];  // As[k][m]
  __shared__ double Bs[GK][GT + 1];  // Bs[k][n]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int col = lane & 15, kq = lane >> 4;

  d4 acc[2][2], accv[MIR ? 2 : 1][MIR ? 2 : 1];  // accv: mirror jobs' odd-row sums V
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < (MIR ? 2 : 1); ++i)
#pragma unroll
    for (int j = 0; j < (MIR ? 2 : 1); ++j) accv[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

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
    if (MIR && mir) {
      // per step, rows ku = 2 (4 s + kq) + pu (even sine rows) into U (acc) and ku ^ 1 (odd
      // ones) into V; k0 is a multiple of GK, so local and global row parities agree
      const int ks2 = (min(K - k0, GK) + 7) / 8;
      for (int ks = 0; ks < ks2; ++ks) {
        const int ku = 2 * (ks * 4 + kq) + pu, kv = ku ^ 1;
        double au[2], bu[2], av[2], bv[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          au[i] = Bs[ku][wn * 32 + i * 16 + col];
          av[i] = Bs[kv][wn * 32 + i * 16 + col];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          bu[j] = As[ku][wm * 32 + j * 16 + col];
          bv[j] = As[kv][wm * 32 + j * 16 + col];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(au[i], bu[j], acc[i][j], 0, 0, 0);
            accv[i & (MIR ? 1 : 0)][j & (MIR ? 1 : 0)] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                av[i], bv[j], accv[i & (MIR ? 1 : 0)][j & (MIR ? 1 : 0)], 0, 0, 0);
          }
      }
      __syncthreads();
      continue;
    }
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
  if (MIR && mir) {
    // columns n and mref - n: C (+)= U + V and U - V, one 16-column subtile at a time (its
    // column pair's loads first)
    const int mref = jb.mref;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      double c0[2][4], c1[2][4];
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = min(m0 + wm * 32 + j * 16 + col, M - 1);
          const int n = min(n0 + wn * 32 + i * 16 + kq + 4 * r, N - 1);
          c0[j][r] = accumulate ? C[(size_t)m + (size_t)n * ldc] : 0.0;
          c1[j][r] = accumulate ? C[(size_t)m + (size_t)(mref - n) * ldc] : 0.0;
        }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int m = m0 + wm * 32 + j * 16 + col;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn * 32 + i * 16 + kq + 4 * r;
          if (m >= M || n >= N) continue;
          const double u = acc[i][j][r], v = accv[i & (MIR ? 1 : 0)][j & (MIR ? 1 : 0)][r];
          C[(size_t)m + (size_t)n * ldc] = c0[j][r] + (u + v);
          if (mref - n != n) C[(size_t)m + (size_t)(mref - n) * ldc] = c1[j][r] + (u - v);
        }
      }
    }
    return;
  }
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

constexpr int GM = 2, GN = 2;                // 16-wide subtiles of the C tile along m, n
constexpr int GTM = 16 * GM, GTN = 16 * GN;  // C tile of gemm_f64_reg_kernel
constexpr int GCE = GTM * GTN / 256;         // C entries each thread sums and stores

// DA, DB: some job of the launch has a difference term A2 (B2); without, the operand is
// loaded once. KR: the longest run a lane loads per chunk (chunks of 16 KR values of K).
// MK: the launch's jobs have a mirrored A (GemmJob::amirror; DA false).
template <bool DA, bool DB, int KR, bool MK = false>
__global__ __launch_bounds__(256) void gemm_f64_reg_kernel(const GemmBatch gb) {
  if (skipped(gb.skip)) return;
  int job, m0, n0;
  if (!gemm_tile<GTM, GTN>(gb, job, m0, n0)) return;  // whole work-group: barriers stay uniform
  // the job's fields in registers (uniform), read once; a missing difference term reads
  // the matrix itself with weight 0 (a - 0 * a = a exactly), so no load is conditional
  const GemmJob& jb = gb.j[job];
  const double *A = jb.A, *B = jb.B;
  const double* A2 = jb.A2 ? jb.A2 : jb.A;
  const double* B2 = jb.B2 ? jb.B2 : jb.B;
  const double wa = jb.A2 ? 1.0 : 0.0, wb = jb.B2 ? 1.0 : 0.0;
  double* C = jb.C;
  const int lda = jb.lda, ldb = jb.ldb, ldc = jb.ldc, M = jb.M, N = jb.N, K = jb.K;
  const int amid = jb.amid;
  const double asign = jb.asign;
  const bool accumulate = jb.accumulate != 0;
  __shared__ double red[4][GTN][GTM + 1];  // the waves' partial tiles, [n][m]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int col = lane & 15, kq = lane >> 4, run = 4 * w + kq;

  // the C entries this thread finishes, loaded first so that they arrive with the operands
  double cin[GCE];
#pragma unroll
  for (int e = 0; e < GCE; ++e) {
    const int idx = tid + 256 * e, m = idx % GTM, n = idx / GTM;
    const size_t o = (size_t)min(m0 + m, M - 1) + (size_t)min(n0 + n, N - 1) * ldc;
    cin[e] = accumulate ? C[o] : 0.0;
  }
  size_t ao[GM], bo[GN];
#pragma unroll
  for (int i = 0; i < GM; ++i) ao[i] = (size_t)min(m0 + 16 * i + col, M - 1);
#pragma unroll
  for (int j = 0; j < GN; ++j) bo[j] = (size_t)min(n0 + 16 * j + col, N - 1) * ldb;

  d4 acc[GN][GM];
#pragma unroll
  for (int j = 0; j < GN; ++j)
#pragma unroll
    for (int i = 0; i < GM; ++i) acc[j][i] = (d4){0.0, 0.0, 0.0, 0.0};

  // one chunk: lane quarter kq of wave w takes k = kc0 + run + 16 r, r < KR; every load is
  // unconditional with its index clamped into the matrix (zeroed below), so all of them
  // issue back to back; past K they repeat the last column's address
  auto chunk = [&](const int kc0) __attribute__((always_inline)) {
    double a[GM][KR], b[GN][KR];
#pragma unroll
    for (int r = 0; r < KR; ++r) {
      const size_t k = (size_t)min(kc0 + run + 16 * r, K - 1);
#pragma unroll
      for (int i = 0; i < GM; ++i) {
        const size_t o = ao[i] + k * lda;
        if (MK)  // A(m, k) + asign A(m, mirror of k): fma with weight +-1 or 0 rounds once
          a[i][r] = fma((int)k == amid ? 0.0 : asign, A2[(long long)ao[i] - (long long)k * lda], A[o]);
        else
          a[i][r] = DA ? fma(-wa, A2[o], A[o]) : A[o];
      }
#pragma unroll
      for (int j = 0; j < GN; ++j) {
        const size_t o = k + bo[j];
        b[j][r] = DB ? fma(-wb, B2[o], B[o]) : B[o];
      }
    }
    const int nr = (min(16 * KR, K - kc0) + 15) / 16;  // steps holding some k < K (uniform)
#pragma unroll
    for (int r = 0; r < KR; ++r) {
      if (r < nr) {
        const bool ok = kc0 + run + 16 * r < K;
#pragma unroll
        for (int j = 0; j < GN; ++j)
#pragma unroll
          for (int i = 0; i < GM; ++i)
            acc[j][i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ok ? b[j][r] : 0.0, ok ? a[i][r] : 0.0,
                                                             acc[j][i], 0, 0, 0);
      }
    }
  };
  // the first chunk outside the loop: its loads do not wait for those of the C entries
  chunk(0);
  for (int kc0 = 16 * KR; kc0 < K; kc0 += 16 * KR) chunk(kc0);
  // acc[j][i][r] at lane: the wave's partial C(m = m0 + 16 i + col, n = n0 + 16 j + kq + 4 r)
#pragma unroll
  for (int j = 0; j < GN; ++j)
#pragma unroll
    for (int i = 0; i < GM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[w][16 * j + kq + 4 * r][16 * i + col] = acc[j][i][r];
  __syncthreads();
#pragma unroll
  for (int e = 0; e < GCE; ++e) {
    const int idx = tid + 256 * e, m = idx % GTM, n = idx / GTM;
    const double v = ((red[0][n][m] + red[1][n][m]) + red[2][n][m]) + red[3][n][m];
    if (m0 + m < M && n0 + n < N) C[(size_t)(m0 + m) + (size_t)(n0 + n) * ldc] = cin[e] + v;
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

// slots of the jobs in the order of gemm_tile (gb.lbase) and the number of TM x TN tiles;
// returns the number of work-groups
template <int TM, int TN>
int gemm_groups(GemmBatch& gb, int njobs, int* tiles) {
  int l = 0, t = 0;
  for (int i = 0; i < njobs; ++i) {
    const int T = ((gb.j[i].M + TM - 1) / TM) * ((gb.j[i].N + TN - 1) / TN);
    gb.lbase[i] = l;
    l += (T + 7) / 8;
    t += T;
  }
  *tiles = t;
  return 8 * l;
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
  if (njobs <= 0) return hipSuccess;
  if (njobs > kMaxPlanes) return hipErrorInvalidValue;
  GemmBatch gb = gb0;
  gb.skip = g_skip;
  gb.njobs = njobs;
  bool da = false, db = false, mir = false, mk = false;
  for (int i = 0; i < njobs; ++i) {
    const GemmJob& j = gb.j[i];
    if (j.M <= 0 || j.N <= 0 || j.K <= 0) return hipErrorInvalidValue;
    if (j.mirror && (j.mref < j.N - 1 || j.mref - (j.N - 1) > j.N)) return hipErrorInvalidValue;
    if ((j.mirror != 0) != (gb.j[0].mirror != 0)) return hipErrorInvalidValue;  // all or none
    da = da || (j.A2 && !j.amirror);  // a mirrored A's A2 is its mirror, not a difference
    db = db || j.B2;
    mir = mir || j.mirror;
    mk = mk || j.amirror;
    if ((j.amirror != 0) != (gb.j[0].amirror != 0)) return hipErrorInvalidValue;  // all or none
    if (j.amirror && (!j.A2 || j.amid < -1 || j.amid >= j.K)) return hipErrorInvalidValue;
  }
  // the 64 x 64 kernel once the launch holds two of its tiles per CU; smaller launches
  // take the register-staged 32 x 32 one
  int tiles = 0;
  const int wide_groups = gemm_groups<GT, GT>(gb, njobs, &tiles);
  if (mir) {  // mirror jobs only (never with a difference term): the 64 x 64 kernel
    if (da || db) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm_f64_kernel<false, false, true>), dim3(wide_groups), dim3(256), 0, st, gb);
    return hipGetLastError();
  }
  if (mk) {  // mirrored A (no difference terms): the register-staged kernel, K chunks of 64
    if (da || db || mir) return hipErrorInvalidValue;
    const dim3 grid(gemm_groups<GTM, GTN>(gb, njobs, &tiles));
    hipLaunchKernelGGL((gemm_f64_reg_kernel<false, false, 4, true>), grid, dim3(256), 0, st, gb);
    return hipGetLastError();
  }
  if (tiles >= 512) {
    const dim3 grid(wide_groups);
    if (da && db)
      hipLaunchKernelGGL((gemm_f64_kernel<true, true, false>), grid, dim3(256), 0, st, gb);
    else if (da)
      hipLaunchKernelGGL((gemm_f64_kernel<true, false, false>), grid, dim3(256), 0, st, gb);
    else if (db)
      hipLaunchKernelGGL((gemm_f64_kernel<false, true, false>), grid, dim3(256), 0, st, gb);
    else
      hipLaunchKernelGGL((gemm_f64_kernel<false, false, false>), grid, dim3(256), 0, st, gb);
    return hipGetLastError();
  }
  const dim3 grid(gemm_groups<GTM, GTN>(gb, njobs, &tiles));
  // K chunks of 64 (KR 4): their registers leave room for 4 waves per SIMD (the most the
  // reduction buffer allows), and more round trips at that occupancy measured faster than
  // one chunk over the whole K at 2 or 3 waves (the box update's first product 15.0 ->
  // 13.0 us at K = 128)
#define GEMM_LAUNCH(DA_, DB_, KR_) \
  hipLaunchKernelGGL((gemm_f64_reg_kernel<DA_, DB_, KR_>), grid, dim3(256), 0, st, gb)
  if (da && db) {
    GEMM_LAUNCH(true, true, 4);
  } else if (da) {
    GEMM_LAUNCH(true, false, 4);
  } else if (db) {
    GEMM_LAUNCH(false, true, 4);
  } else {
    GEMM_LAUNCH(false, false, 4);
  }
#undef GEMM_LAUNCH
  return hipGetLastError();
}

}  // namespace ibpm
