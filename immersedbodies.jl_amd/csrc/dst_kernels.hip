// Sine-transform passes of the DST-I Poisson / Helmholtz inverses as dense sine-matrix
// contractions on the CDNA4 matrix cores.
//
// Reference: the 2-D RODFT00 of FFTW used by _lap_inv_operator (fluid.jl:113-115,
// :145-166). A 2-D transform is two passes; each pass contracts the contiguous index
// of its input with the sine matrix and writes the result transposed, so the second
// pass again contracts a contiguous index and the plane returns to its original layout.
//
// Even/odd folding. With M = N+1 and H = floor(M/2), row j of S (S_jk = 2 sin(pi j k/M))
// satisfies S_{j,M-k} = (-1)^{j+1} S_{j,k}. Folding the input as
//   P_k = x_k + x_{M-k} (k <= H; for even M, P_H = x_H)  -> odd rows  j = 2m+1, m < H
//   Q_k = x_k - x_{M-k} (k < H, or k <= H for odd M)     -> even rows j = 2m+2, m < E
// halves the contraction: odd rows use So[m][k] = 2 sin(pi (2m+1) k / M) (H x H) and
// even rows use Se[m][k] = 2 sin(pi 2(m+1) k / M) (E x E, E = H - 1 for even M and H for
// odd M).
//
// The fold and the transpose are staged through LDS; the products run on
//   FP64 path : v_mfma_f64_16x16x4_f64
//   mixed path: v_mfma_f32_16x16x32_bf16, three products hi*hi + hi*lo + lo*hi of the
//               bf16 split of fp32 operands, fp32 accumulation.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <map>
#include <mutex>

#include "ibpm_device.h"
#include "ibpm_internal.h"

namespace ibpm {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <typename T>
__device__ __forceinline__ double ld(const T* p, size_t i) {
  return (double)p[i];
}

// Input element (k0, b) of a pass (k0 along the contracted index).
template <typename InT>
__device__ __forceinline__ double load_in(const PassPlane& d, int k0, int b) {
  return ld((const InT*)d.in, (size_t)k0 + (size_t)b * d.Nk);
}

// Output element (b, jc) with the pass epilogue. EPI_DIAG applies the spectral diagonal
// D = dscale / (da + lambda dnum / dden), lambda = -2((cos_b + cos_j) - 2): the
// reference's Delta_inv (da = 0, dnum = dden = 1) and Ainv (da = 1, dnum = dt,
// dden = 2 Re hc^2) eigenvalues (fluid.jl:117-134, :168-173), in its evaluation order.
template <typename OutT, int EPI>
__device__ __forceinline__ void store_out(const PassPlane& d, int b, int jc, double val) {
  const size_t idx = (size_t)b + (size_t)jc * d.Nb;
  OutT* out = (OutT*)d.out;
  if (EPI == EPI_DIAG) {
    const double lam = -2.0 * ((d.cosb[b] + d.cosj[jc]) - 2.0);
    const double den = d.da + lam * d.dnum / d.dden;
    if (d.spec) d.spec[idx] = val / den;
    val *= d.dscale / den;
  }
  if (EPI == EPI_SUB) val = d.base[idx] - val;
  if (EPI == EPI_ADD) val = d.base[idx] + val;
  out[idx] = (OutT)val;
}

// XCD-aware tile order. Work-groups are dealt round-robin to the 8 XCDs (each with its own
// L2); logical tiles are numbered with the m-tile fastest, so the tiles that share a
// plane's rows of input (same plane, same b-tile) are consecutive, and work-group w takes
// logical tile (w % 8) * ceil(T/8) + w / 8: each XCD walks a contiguous range and reads
// its rows of input from its own L2.
__device__ __forceinline__ bool xcd_tile(const PassBatch& b, int& z, int& bt, int& mt) {
  if (skipped(b.skip)) return false;
  const int T = b.nplanes * b.nbt * b.nmt;
  const int per = (T + 7) / 8;
  const int w = blockIdx.x;
  const int t = (w & 7) * per + (w >> 3);
  if (t >= T) return false;
  mt = t % b.nmt;
  bt = (t / b.nmt) % b.nbt;
  z = t / (b.nmt * b.nbt);
  return true;
}

// ---------------------------------------------------------------------------------
// FP64 pass. Work-group tile: TB values of b x (TM odd + TM even) output rows, K chunk
// KC = 16, loads issued PF chunks ahead of their use; TB, TM in {32, 64} and PF in {1, 2}
// are picked per launch shape by the autotuner (a deeper prefetch hides more load latency
// at the cost of registers, hence occupancy). 4 waves: wave w
// computes parity (w>>1) on b-half (w&1), a TB/2 x TM block as (TB/32) x (TM/16) MFMA
// 16x16x4 tiles. Input chunk and sine operands are staged through LDS; the global loads
// of the next chunk are issued before the matrix-core work of the current one.
// ---------------------------------------------------------------------------------
template <typename InT, typename OutT, int EPI, int TB, int TM, int KC, int PF>
__global__ __launch_bounds__(256) void dst_pass_f64_kernel(const PassBatch batch) {
  constexpr int BS = TB / 32;            // 16-wide b subtiles per wave
  constexpr int MS = TM / 16;            // 16-wide m subtiles per wave
  constexpr int RA = KC * TB / 256;      // input elements per thread and chunk
  constexpr int RB = KC * TM / 256;      // sine elements per thread and chunk
  int z, bt, mt;
  if (!xcd_tile(batch, z, bt, mt)) return;
  const PassPlane& d = batch.p[z];
  const int Nk = d.Nk, Nb = d.Nb;
  const int H = (Nk + 1) / 2;
  const bool modd = (Nk + 1) & 1;  // odd M: every k <= H has a partner, H even rows
  const int E = modd ? H : H - 1;  // even output rows (and their contraction length)
  const int b0 = bt * TB;
  const int m0 = mt * TM;
  if (b0 >= Nb || m0 >= H) return;

  __shared__ double Ps[KC][TB + 1];
  __shared__ double Qs[KC][TB + 1];
  __shared__ double So_s[KC][TM + 1];
  __shared__ double Se_s[KC][TM + 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int parity = wave >> 1;  // 0: odd rows, 1: even rows
  const int bh = wave & 1;
  const double* SoT = (const double*)d.So;
  const double* SeT = (const double*)d.Se;

  d4 acc[BS][MS];
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int j = 0; j < MS; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

  // PF register stages: the loads of chunk c + PF are issued while chunk c is multiplied
  double ra[PF][RA], rm[PF][RA], rso[PF][RB], rse[PF][RB];
  // loads unconditional (indices clamped into range), zeroed when staged
  auto fetch = [&](const int st, int kc) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int e = tid + 256 * r;
      const int k1c = min(kc + e % KC + 1, H);  // 1-based k; KC lanes read KC consecutive k
      const int bc = min(b0 + e / KC, Nb - 1);
      ra[st][r] = load_in<InT>(d, k1c - 1, bc);
      rm[st][r] = load_in<InT>(d, Nk - k1c, bc);
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int e = tid + 256 * r;
      const int k0 = kc + e / TM;
      const int m = m0 + e % TM;
      rso[st][r] = SoT[(size_t)min(k0, H - 1) * H + min(m, H - 1)];
      rse[st][r] = E > 0 ? SeT[(size_t)min(k0, E - 1) * E + min(m, E - 1)] : 0.0;
    }
  };
  const int col = lane & 15;
  const int kq = lane >> 4;
  // fold into LDS: P = x_k + x_{M-k}, Q = x_k - x_{M-k}; for even M, k = H has no
  // partner (Q = 0)
  auto stage = [&](const int st, int kc) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int e = tid + 256 * r;
      const int kk = e % KC;
      const int k1 = kc + kk + 1;
      const bool ok = k1 <= H && b0 + e / KC < Nb;
      const double va = ok ? ra[st][r] : 0.0;
      const double vm = ok && (k1 < H || modd) ? rm[st][r] : 0.0;
      const bool mid = !modd && k1 == H;
      Ps[kk][e / KC] = va + vm;
      Qs[kk][e / KC] = mid ? 0.0 : va - vm;
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int e = tid + 256 * r;
      const int k1 = kc + e / TM + 1;
      const int m = m0 + e % TM;
      So_s[e / TM][e % TM] = (k1 <= H && m < H) ? rso[st][r] : 0.0;
      Se_s[e / TM][e % TM] = (k1 <= E && m < E) ? rse[st][r] : 0.0;
    }
  };
  auto multiply = [&]() {
    const double(*A)[TB + 1] = parity == 0 ? Ps : Qs;
    const double(*Bs)[TM + 1] = parity == 0 ? So_s : Se_s;
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int krow = ks * 4 + kq;
      double av[BS];
#pragma unroll
      for (int bs = 0; bs < BS; ++bs) av[bs] = A[krow][bh * (TB / 2) + bs * 16 + col];
#pragma unroll
      for (int ms = 0; ms < MS; ++ms) {
        const double bv = Bs[krow][ms * 16 + col];
#pragma unroll
        for (int bs = 0; bs < BS; ++bs)
          acc[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[bs], bv, acc[bs][ms], 0, 0, 0);
      }
    }
  };
#pragma unroll
  for (int st = 0; st < PF; ++st)
    if (st * KC < H) fetch(st, st * KC);
  for (int kc = 0; kc < H; kc += PF * KC) {
#pragma unroll
    for (int st = 0; st < PF; ++st) {
      const int k = kc + st * KC;  // the same for every thread: barriers stay uniform
      if (k < H) {
        stage(st, k);
        __syncthreads();
        if (k + PF * KC < H) fetch(st, k + PF * KC);
        multiply();
        __syncthreads();
      }
    }
  }

  // epilogue: the f64 16x16 accumulator holds D[row = lane/16 + 4r][col = lane%16]
  const int rq = lane >> 4;
  const int mlim = parity == 0 ? H : E;
#pragma unroll
  for (int bs = 0; bs < BS; ++bs)
#pragma unroll
    for (int ms = 0; ms < MS; ++ms) {
      const int m = m0 + ms * 16 + col;
      if (m >= mlim) continue;
      const int jc = parity == 0 ? 2 * m : 2 * m + 1;  // 0-based output row index j-1
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = b0 + bh * (TB / 2) + bs * 16 + rq + 4 * r;
        if (b < Nb) store_out<OutT, EPI>(d, b, jc, acc[bs][ms][r]);
      }
    }
}

// ---------------------------------------------------------------------------------
// FP64 pass with a second fold. M = N+1 = 4Q (nx and ny are multiples of 4), H = 2Q.
// Odd rows: rows j = 2m+1 and M-j = 2(H-1-m)+1 satisfy S_{M-j,k} = (-1)^{k+1} S_{j,k},
// so with (m < Q)
//   A_m = sum_{i<Q} T1[i][m] P_{2i+1},  B_m = sum_{i<Q} T2[i][m] P_{2i+2}
// the outputs are Y_{2m+1} = A_m + B_m and Y_{M-1-2m} = A_m - B_m.
// Even rows: the block is a DST-I of modulus H of Q_k and folds the same way once more:
//   Y_{4m+2} = sum_{i<Q} T2[i][m] P'_{i+1},   P'_k = Q_k + Q_{H-k} (k < Q), P'_Q = Q_Q
//   Y_{4m+4} = sum_{i<Q-1} T4[i][m] Q'_{i+1}, Q'_k = Q_k - Q_{H-k}
// with T1[i][m] = 2 sin(pi (2m+1)(2i+1)/M), T2[i][m] = 2 sin(pi (2m+1)(2i+2)/M) (Q x Q)
// and T4[i][m] = 2 sin(pi 4 (m+1)(i+1)/M) ((Q-1) x (Q-1)), stored [T1 | T2 | T4], i-major,
// each in a Qp x Qp block (Qp = Q rounded up to 64) whose rows and columns beyond the
// table are zero.
// Four Q-long contractions replace two H-long ones: half the matrix-core work of the
// single fold. Work-group: one block (odd: A and B, even: the two quarter outputs), TB
// values of b x TM quarter rows, K chunk KC, PF register stages; wave w computes the
// TB/2 x TM/2 block (b-half w&1, m-half w>>1) of both of the block's products, so the
// odd block's butterfly A +- B stays in registers.
// ---------------------------------------------------------------------------------
// K-chunk order of the double-fold passes. The fold index i runs in blocks of BLK: the
// even-row group (grp 1) takes them in order, the odd-row group (grp 0) as 0, nb-1, 1,
// nb-2, ... Block b of the even group reads the input near elements BLK b, H-BLK b,
// H+BLK b and M-BLK b of each line (x at i+1, H-i-1, H+i+1, M-i-1), block b of the odd
// group near 2 BLK b and M-2 BLK b (x at 2i+1, 2i+2 and their mirrors), so blocks b and
// nb-1-b of the odd group read, over the same two blocks' time, the stretches that blocks
// 2b and 2b+1 of the even group read: the work-groups of a b-tile walk its lines together
// and share them in L2, instead of each group fetching the whole lines at its own time. A
// block is BLK / KC chunks taken in order, so every K chunk size sums in the same order
// and all FP64 double-fold variants (BLK 16) give the same bits.
template <int BLK = 16>
__device__ __forceinline__ int fold_chunk(int grp, int t, int KC, int nb) {
  const int per = BLK / KC, tb = t / per, h = t - tb * per;
  const int blk = grp ? tb : ((tb & 1) ? nb - 1 - (tb >> 1) : (tb >> 1));
  return BLK * blk + KC * h;
}

// The 32-row tiles with one register stage run at 3 waves per SIMD or more: the epilogues
// that scale or combine (EPI_DIAG, EPI_SUB, EPI_ADD) would otherwise take enough
// registers to drop to 2.
template <typename InT, typename OutT, int EPI, int TB, int TM, int KC, int PF>
__global__ __launch_bounds__(256, (TB == 32 && PF == 1) ? 3 : 1) void dst_pass_f64q_kernel(
    const PassBatch batch) {
  constexpr int BS = TB / 32;        // 16-wide b subtiles per wave
  constexpr int MS = TM / 32;        // 16-wide m subtiles per wave
  constexpr int RA = KC * TB / 256;  // (i, b) input slots per thread and chunk
  constexpr int RB = KC * TM / 256;  // sine elements per thread, chunk and table
  static_assert(BS >= 1 && MS >= 1 && RA >= 1 && RB >= 1, "tile too small");
  int z, bt, mt2;
  if (!xcd_tile(batch, z, bt, mt2)) return;
  const int grp = mt2 & 1;  // 0: odd rows, 1: even rows
  const int mt = mt2 >> 1;
  const PassPlane& d = batch.p[z];
  const int Nk = d.Nk, Nb = d.Nb;
  const int M = Nk + 1, H = M / 2, Q = M / 4;
  const int b0 = bt * TB;
  const int m0 = mt * TM;
  if (b0 >= Nb || m0 >= Q) return;

  __shared__ double X0s[KC][TB + 1];
  __shared__ double X1s[KC][TB + 1];
  __shared__ double Sa_s[KC][TM + 1];
  __shared__ double Sb_s[KC][TM + 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int bh = wave & 1;
  const int mh = wave >> 1;
  const int Qp = quarter_pitch(Q);
  const double* T1 = (const double*)d.Sq;
  const double* T2 = T1 + (size_t)Qp * Qp;
  const double* T4 = T2 + (size_t)Qp * Qp;
  const double* SA = grp == 0 ? T1 : T2;
  const double* SB = grp == 0 ? T2 : T4;
  const int nb_lim = grp == 0 ? Q : Q - 1;  // rows and columns of SB
  const InT* xin = (const InT*)d.in;

  d4 acc0[BS][MS], acc1[BS][MS];
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int j = 0; j < MS; ++j) {
      acc0[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
      acc1[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
    }

  // Loads are unconditional: indices clamped into range (always valid addresses), the
  // values of out-of-range (i, b) and of the partner-less k = H zeroed when staged, so
  // a chunk's loads issue back to back with no wait between them.
  // element offsets of the four inputs x_k of fold index i in a line (grp-uniform)
  const int o0m = grp == 0 ? 2 : 1, o1c = grp == 0 ? M - 2 : M - 2, o1m = grp == 0 ? 2 : 1;
  double r0[PF][RA], r1[PF][RA], r2[PF][RA], r3[PF][RA], rsa[PF][RB], rsb[PF][RB];
  auto fetch = [&](const int st, int kc) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int e = tid + 256 * r;
      const int i = min(kc + e % KC, Q - 1);  // KC lanes read KC consecutive i
      const int b = min(b0 + e / KC, Nb - 1);
      const InT* x = xin + (size_t)b * Nk;  // x_k at x[k-1]
      // grp 0: x_{2i+1}, x_{M-2i-1}, x_{2i+2}, x_{M-2i-2}
      // grp 1: x_{i+1},  x_{M-i-1},  x_{H-i-1}, x_{H+i+1}
      const int k0 = o0m * i, k1 = o1c - o1m * i;
      const int k2 = grp == 0 ? 2 * i + 1 : H - i - 2;
      const int k3 = grp == 0 ? M - 2 * i - 3 : H + i;
      r0[st][r] = (double)x[k0];
      r1[st][r] = (double)x[k1];
      r2[st][r] = (double)x[k2];
      r3[st][r] = (double)x[k3];
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int e = tid + 256 * r;
      const int i = kc + e / TM;
      const int m = m0 + e % TM;
      rsa[st][r] = SA[(size_t)min(i, Q - 1) * Qp + min(m, Q - 1)];
      rsb[st][r] = SB[(size_t)min(i, nb_lim - 1) * Qp + min(m, nb_lim - 1)];
    }
  };
  const int col = lane & 15;
  const int kq = lane >> 4;
  auto stage = [&](const int st, int kc) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int e = tid + 256 * r;
      const int ii = e % KC;
      const int i = kc + ii;
      const bool ok = i < Q && b0 + e / KC < Nb;
      double a, b;
      if (grp == 0) {
        const double v3 = 2 * i + 2 < H ? r3[st][r] : 0.0;  // k = H has no partner
        a = r0[st][r] + r1[st][r];  // P_{2i+1}
        b = r2[st][r] + v3;         // P_{2i+2}
      } else {
        const double qk = r0[st][r] - r1[st][r];  // Q_{i+1}
        const double qh = r2[st][r] - r3[st][r];  // Q_{H-i-1}
        const bool last = i == Q - 1;             // k = Q is its own partner
        a = last ? qk : qk + qh;
        b = last ? 0.0 : qk - qh;
      }
      X0s[ii][e / KC] = ok ? a : 0.0;
      X1s[ii][e / KC] = ok ? b : 0.0;
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int e = tid + 256 * r;
      const int i = kc + e / TM;
      const int m = m0 + e % TM;
      Sa_s[e / TM][e % TM] = (i < Q && m < Q) ? rsa[st][r] : 0.0;
      Sb_s[e / TM][e % TM] = (i < nb_lim && m < nb_lim) ? rsb[st][r] : 0.0;
    }
  };
  auto multiply = [&]() {
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int krow = ks * 4 + kq;
      double a0[BS], a1[BS];
#pragma unroll
      for (int bs = 0; bs < BS; ++bs) {
        a0[bs] = X0s[krow][bh * (TB / 2) + bs * 16 + col];
        a1[bs] = X1s[krow][bh * (TB / 2) + bs * 16 + col];
      }
#pragma unroll
      for (int ms = 0; ms < MS; ++ms) {
        const double ba = Sa_s[krow][mh * (TM / 2) + ms * 16 + col];
        const double bb = Sb_s[krow][mh * (TM / 2) + ms * 16 + col];
#pragma unroll
        for (int bs = 0; bs < BS; ++bs) {
          acc0[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[bs], ba, acc0[bs][ms], 0, 0, 0);
          acc1[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[bs], bb, acc1[bs][ms], 0, 0, 0);
        }
      }
    }
  };
  static_assert(KC == 16, "one chunk per block of fold_chunk");
  const int nb = (Q + 15) / 16;  // chunks, taken in fold_chunk order
#pragma unroll
  for (int st = 0; st < PF; ++st)
    if (st < nb) fetch(st, fold_chunk(grp, st, KC, nb));
  for (int t0 = 0; t0 < nb; t0 += PF) {
#pragma unroll
    for (int st = 0; st < PF; ++st) {
      const int t = t0 + st;  // the same for every thread: barriers stay uniform
      if (t < nb) {
        stage(st, fold_chunk(grp, t, KC, nb));
        __syncthreads();
        if (t + PF < nb) fetch(st, fold_chunk(grp, t + PF, KC, nb));
        multiply();
        __syncthreads();
      }
    }
  }

  // epilogue: the f64 16x16 accumulator holds D[row = lane/16 + 4r][col = lane%16]
  const int rq = lane >> 4;
#pragma unroll
  for (int bs = 0; bs < BS; ++bs)
#pragma unroll
    for (int ms = 0; ms < MS; ++ms) {
      const int m = m0 + mh * (TM / 2) + ms * 16 + col;
      if (m >= Q) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = b0 + bh * (TB / 2) + bs * 16 + rq + 4 * r;
        if (b >= Nb) continue;
        const double u = acc0[bs][ms][r], v = acc1[bs][ms][r];
        if (grp == 0) {
          store_out<OutT, EPI>(d, b, 2 * m, u + v);          // row 2m+1
          store_out<OutT, EPI>(d, b, M - 2 - 2 * m, u - v);  // row M-1-2m
        } else {
          store_out<OutT, EPI>(d, b, 4 * m + 1, u);                  // row 4m+2
          if (m < Q - 1) store_out<OutT, EPI>(d, b, 4 * m + 3, v);  // row 4m+4
        }
      }
    }
}

// ---------------------------------------------------------------------------------
// FP64 double-fold pass with the sine tables copied straight into LDS. The same
// contraction, work-group tile and wave layout as dst_pass_f64q_kernel (and the same sums
// in the same order: bitwise the same result), for launches whose every plane has Q a
// multiple of KC (no partial K chunk):
//   - the table tiles (KC rows x TM columns of the Qp-pitched, zero-padded tables) go
//     from global memory to LDS by global_load_lds_dwordx4 (16 B per lane), no registers,
//     masks or clamps; a wave instruction fills 1 KB, 128 / TM rows;
//   - each thread keeps 32-bit byte offsets of its four inputs per slot and steps them by
//     a uniform amount per chunk (no 64-bit address arithmetic in the loop);
//   - the folded input tile is XOR-swizzled (column b ^ k) instead of padded.
// The smaller register and LDS footprint (TB = 32, TM = 64: 96 VGPRs, 24 KB) runs five
// work-groups per CU: the 1280 tiles of a 5-plane 1023 x 1023 pass all run at once.
// Loop per chunk: stage the inputs fetched one chunk ahead, wait for the table copy,
// barrier, fetch the next chunk's inputs, multiply, barrier, start the next table copy.
// ---------------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) const void global_void;

template <typename InT, typename OutT, int EPI, int TB, int TM, int KC, int MINW>
__global__ __launch_bounds__(256, MINW) void dst_pass_f64d_kernel(const PassBatch batch) {
  constexpr int BS = TB / 32, MS = TM / 32, RA = KC * TB / 256;
  constexpr int RPI = 128 / TM;      // table rows per wave instruction (1 KB)
  constexpr int NT = KC * TM / 128;  // wave instructions per table and chunk
  constexpr int NI = (NT + 3) / 4;   // per wave (waves past NT / NI issue none)
  static_assert(BS >= 1 && MS >= 1 && RA >= 1 && NT >= 1 && KC <= 16, "tile");
  int z, bt, mt2;
  if (!xcd_tile(batch, z, bt, mt2)) return;
  const int grp = mt2 & 1;  // 0: odd rows, 1: even rows
  const int mt = mt2 >> 1;
  const PassPlane& d = batch.p[z];
  const int Nk = d.Nk, Nb = d.Nb;
  const int M = Nk + 1, H = M / 2, Q = M / 4, Qp = quarter_pitch(Q);
  const int b0 = bt * TB;
  const int m0 = mt * TM;
  if (b0 >= Nb || m0 >= Q) return;

  __shared__ double X0s[KC][TB];
  __shared__ double X1s[KC][TB];
  __shared__ __attribute__((aligned(16))) double Sa_s[KC][TM];
  __shared__ __attribute__((aligned(16))) double Sb_s[KC][TM];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int bh = wave & 1;
  const int mh = wave >> 1;
  const size_t tbytes = (size_t)Qp * Qp * 8;
  const char* T1 = (const char*)d.Sq;
  const char* SA = grp == 0 ? T1 : T1 + tbytes;
  const char* SB = grp == 0 ? T1 + tbytes : T1 + 2 * tbytes;
  const char* xb = (const char*)d.in;

  // input offsets: grp 0 reads x_{2i+1}, x_{M-2i-1}, x_{2i+2}, x_{M-2i-2}; grp 1 x_{i+1},
  // x_{M-i-1}, x_{H-i-1}, x_{H+i+1} (x_k at element k-1 of the line); rows past Nb read
  // the last row and stage zeros
  unsigned o0[RA], o1[RA], o2[RA], o3[RA];
  bool rowok[RA];
#pragma unroll
  for (int r = 0; r < RA; ++r) {
    const int e = tid + 256 * r;
    const int ii = e % KC;
    const int b = b0 + e / KC;
    rowok[r] = b < Nb;
    const unsigned line = (unsigned)min(b, Nb - 1) * (unsigned)Nk;
    const int k0 = grp == 0 ? 2 * ii : ii;
    const int k1 = grp == 0 ? M - 2 - 2 * ii : M - 2 - ii;
    const int k2 = grp == 0 ? 2 * ii + 1 : H - 2 - ii;
    const int k3 = grp == 0 ? M - 3 - 2 * ii : H + ii;
    o0[r] = (line + k0) * 8u;
    o1[r] = (line + k1) * 8u;
    o2[r] = (line + k2) * 8u;
    o3[r] = (line + k3) * 8u;
  }
  // per chunk starting at fold index kc: x_{2i+1}, x_{2i+2} (grp 0) or x_{i+1} (grp 1)
  // move up by a = (2 or 1) 8 kc bytes and their mirrors down; x_{H-i-1}, x_{H+i+1}
  // (grp 1) move down and up by 8 kc
  const unsigned sa = grp == 0 ? 16u : 8u;
  // table copy: wave instruction j fills rows (wave NI + j) RPI + lane / (TM/2), columns
  // m0 + 2 (lane % (TM/2)) and the next, of the chunk's KC = NT RPI rows
  const int rr = lane / (TM / 2), cp = lane % (TM / 2);
  const unsigned ot = (unsigned)((wave * NI * RPI + rr) * Qp + m0 + 2 * cp) * 8u;
  const unsigned ostep = (unsigned)(RPI * Qp * 8), tstep = (unsigned)(Qp * 8);
  auto copy_tables = [&](int kc) {
    const unsigned o = ot + (unsigned)kc * tstep;
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      const int w = wave * NI + j;
      if (NT % 4 == 0 || w < NT) {
        const int r0 = w * RPI;
        __builtin_amdgcn_global_load_lds((global_void*)(SA + o + j * ostep), (lds_void*)&Sa_s[r0][0], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((global_void*)(SB + o + j * ostep), (lds_void*)&Sb_s[r0][0], 16, 0, 0);
      }
    }
  };

  d4 acc0[BS][MS], acc1[BS][MS];
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int j = 0; j < MS; ++j) {
      acc0[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
      acc1[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
    }
  double r0[RA], r1[RA], r2[RA], r3[RA];
  auto at = [](const char* base, unsigned off) { return (double)*(const InT*)(base + off); };
  auto fetch = [&](int kc) {
    const unsigned a = (unsigned)kc * sa, c = grp == 0 ? a : 0u - (unsigned)kc * 8u;
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      r0[r] = at(xb, o0[r] + a);
      r1[r] = at(xb, o1[r] - a);
      r2[r] = at(xb, o2[r] + c);
      r3[r] = at(xb, o3[r] - c);
    }
  };
  const int col = lane & 15;
  const int kq = lane >> 4;
  auto stage = [&](int kc) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int e = tid + 256 * r;
      const int ii = e % KC;
      const int i = kc + ii;
      const int bl = e / KC;
      double a, b;
      if (grp == 0) {
        const double v3 = 2 * i + 2 < H ? r3[r] : 0.0;  // k = H has no partner
        a = r0[r] + r1[r];                               // P_{2i+1}
        b = r2[r] + v3;                                  // P_{2i+2}
      } else {
        const double qk = r0[r] - r1[r];  // Q_{i+1}
        const double qh = r2[r] - r3[r];  // Q_{H-i-1}
        const bool last = i == Q - 1;     // k = Q is its own partner
        a = last ? qk : qk + qh;
        b = last ? 0.0 : qk - qh;
      }
      X0s[ii][bl ^ ii] = rowok[r] ? a : 0.0;
      X1s[ii][bl ^ ii] = rowok[r] ? b : 0.0;
    }
  };
  auto multiply = [&]() {
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int krow = ks * 4 + kq;
      double a0[BS], a1[BS];
#pragma unroll
      for (int bs = 0; bs < BS; ++bs) {
        const int c = (bh * (TB / 2) + bs * 16 + col) ^ krow;
        a0[bs] = X0s[krow][c];
        a1[bs] = X1s[krow][c];
      }
#pragma unroll
      for (int ms = 0; ms < MS; ++ms) {
        const double ba = Sa_s[krow][mh * (TM / 2) + ms * 16 + col];
        const double bb = Sb_s[krow][mh * (TM / 2) + ms * 16 + col];
#pragma unroll
        for (int bs = 0; bs < BS; ++bs) {
          acc0[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[bs], ba, acc0[bs][ms], 0, 0, 0);
          acc1[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[bs], bb, acc1[bs][ms], 0, 0, 0);
        }
      }
    }
  };
  // chunks in fold_chunk order (Q a multiple of 16: every chunk is whole)
  const int nb = Q / 16, nsteps = Q / KC;
  int kc = fold_chunk(grp, 0, KC, nb);
  fetch(kc);
  copy_tables(kc);
  for (int t = 0; t < nsteps; ++t) {  // the same for every thread: barriers stay uniform
    const int kn = t + 1 < nsteps ? fold_chunk(grp, t + 1, KC, nb) : 0;
    stage(kc);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this chunk's table copy has landed
    __syncthreads();
    if (t + 1 < nsteps) fetch(kn);
    multiply();
    __syncthreads();
    if (t + 1 < nsteps) copy_tables(kn);
    kc = kn;
  }

  // epilogue: the f64 16x16 accumulator holds D[row = lane/16 + 4r][col = lane%16]
  const int rq = lane >> 4;
#pragma unroll
  for (int bs = 0; bs < BS; ++bs)
#pragma unroll
    for (int ms = 0; ms < MS; ++ms) {
      const int m = m0 + mh * (TM / 2) + ms * 16 + col;
      if (m >= Q) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = b0 + bh * (TB / 2) + bs * 16 + rq + 4 * r;
        if (b >= Nb) continue;
        const double u = acc0[bs][ms][r], v = acc1[bs][ms][r];
        if (grp == 0) {
          store_out<OutT, EPI>(d, b, 2 * m, u + v);          // row 2m+1
          store_out<OutT, EPI>(d, b, M - 2 - 2 * m, u - v);  // row M-1-2m
        } else {
          store_out<OutT, EPI>(d, b, 4 * m + 1, u);                  // row 4m+2
          if (m < Q - 1) store_out<OutT, EPI>(d, b, 4 * m + 3, v);  // row 4m+4
        }
      }
    }
}

// ---------------------------------------------------------------------------------
// FP64 pass with a third fold (of the T2 blocks). T2[i][m] = 2 sin(pi (2m+1) k / (2Q)),
// k = i+1, satisfies T2[i][Q-1-m] = (-1)^(k+1) T2[i][m], so for m < Q/2 the rows m and
// m~ = Q-1-m of a T2 product come from two half-length sums:
//   O_m = sum_{i even} T2[i][m] X_{i},  E_m = sum_{i odd} T2[i][m] X_{i}
//   row m = O_m + E_m,  row m~ = O_m - E_m
// Both blocks of the double fold hold one T2 product: the odd block's B (X = P_{2i+2})
// and the even block's rows 4m+2 (X = P'_{i+1}). A work-group takes m in [m0, m0 + TM)
// of [0, Q/2) and the mirror rows m~, so it also forms the block's other product on both
// (T1 on P_{2i+1} for the odd block, T4 on Q'_{i+1} for the even block, at columns m and
// m~), and writes four output rows per m:
//   odd block:  Y_{2m+1}, Y_{M-1-2m} = A_m +- (O_m + E_m),
//               Y_{2m~+1}, Y_{M-1-2m~} = A_m~ +- (O_m - E_m)
//   even block: Y_{4m+2} = O_m + E_m, Y_{4m~+2} = O_m - E_m, Y_{4m+4} = C_m,
//               Y_{4m~+4} = C_m~ (m~ < Q-1)
// Three Q-long contractions per pair of m instead of four: three quarters of the double
// fold's matrix-core work, with the same inputs, the same tables (T1, T4 read at columns
// m and m~, T2 at columns m) and the same chunk order (fold_chunk) as
// dst_pass_f64d_kernel, whose table copies and staging it keeps; half as many
// work-groups read each input line. Every K chunk size sums in the same order: all
// triple-fold variants give the same bits.
// ---------------------------------------------------------------------------------
template <typename InT, typename OutT, int EPI, int TB, int TM, int KC, int MINW>
__global__ __launch_bounds__(256, MINW) void dst_pass_f64t_kernel(const PassBatch batch) {
  constexpr int BS = TB / 32, MS = TM / 32, RA = KC * TB / 256;
  constexpr int RPI = 128 / TM;      // table rows per wave instruction (1 KB)
  constexpr int NT = KC * TM / 128;  // wave instructions per table and chunk
  constexpr int NI = (NT + 3) / 4;   // per wave (waves past NT / NI issue none)
  static_assert(BS >= 1 && MS >= 1 && RA >= 1 && NT >= 1 && KC <= 16 && KC % 8 == 0, "tile");
  int z, bt, mt2;
  if (!xcd_tile(batch, z, bt, mt2)) return;
  const int grp = mt2 & 1;  // 0: odd rows, 1: even rows
  const int mt = mt2 >> 1;
  const PassPlane& d = batch.p[z];
  const int Nk = d.Nk, Nb = d.Nb;
  const int M = Nk + 1, H = M / 2, Q = M / 4, Qp = quarter_pitch(Q), Q2 = Q / 2;
  const int b0 = bt * TB;
  const int m0 = mt * TM;
  if (b0 >= Nb || m0 >= Q2) return;

  __shared__ double Xs[2][KC][TB];  // [0]: P_{2i+1} or P'_{i+1}, [1]: P_{2i+2} or Q'_{i+1}
  __shared__ __attribute__((aligned(16))) double Sf_s[KC][TM];  // full product, columns m
  __shared__ __attribute__((aligned(16))) double Sg_s[KC][TM];  // full product, columns m~
  __shared__ __attribute__((aligned(16))) double Sh_s[KC][TM];  // T2, columns m

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int bh = wave & 1;
  const int mh = wave >> 1;
  const size_t tbytes = (size_t)Qp * Qp * 8;
  const char* T1 = (const char*)d.Sq;
  const char* SF = grp == 0 ? T1 : T1 + 2 * tbytes;  // T1 (odd block), T4 (even block)
  const char* SH = T1 + tbytes;                      // T2
  const char* xb = (const char*)d.in;

  // inputs as in dst_pass_f64d_kernel
  unsigned o0[RA], o1[RA], o2[RA], o3[RA];
  bool rowok[RA];
#pragma unroll
  for (int r = 0; r < RA; ++r) {
    const int e = tid + 256 * r;
    const int ii = e % KC;
    const int b = b0 + e / KC;
    rowok[r] = b < Nb;
    const unsigned line = (unsigned)min(b, Nb - 1) * (unsigned)Nk;
    const int k0 = grp == 0 ? 2 * ii : ii;
    const int k1 = grp == 0 ? M - 2 - 2 * ii : M - 2 - ii;
    const int k2 = grp == 0 ? 2 * ii + 1 : H - 2 - ii;
    const int k3 = grp == 0 ? M - 3 - 2 * ii : H + ii;
    o0[r] = (line + k0) * 8u;
    o1[r] = (line + k1) * 8u;
    o2[r] = (line + k2) * 8u;
    o3[r] = (line + k3) * 8u;
  }
  const unsigned sa = grp == 0 ? 16u : 8u;
  // table copy: wave instruction j fills rows (wave NI + j) RPI + lane / (TM/2), columns
  // c0 + 2 (lane % (TM/2)) and the next, c0 = m0 (columns m) or Q - m0 - TM (columns m~,
  // in increasing order: local column TM-1-c holds m~ of local column c)
  const int rr = lane / (TM / 2), cp = lane % (TM / 2);
  const unsigned orow = (unsigned)((wave * NI * RPI + rr) * Qp) * 8u;
  const unsigned ot = orow + (unsigned)(m0 + 2 * cp) * 8u;
  const unsigned og = orow + (unsigned)(Q - m0 - TM + 2 * cp) * 8u;
  const unsigned ostep = (unsigned)(RPI * Qp * 8), tstep = (unsigned)(Qp * 8);
  auto copy_tables = [&](int kc) {
    const unsigned k = (unsigned)kc * tstep;
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      const int w = wave * NI + j;
      if (NT % 4 == 0 || w < NT) {
        const int r0 = w * RPI;
        __builtin_amdgcn_global_load_lds((global_void*)(SF + ot + k + j * ostep), (lds_void*)&Sf_s[r0][0], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((global_void*)(SF + og + k + j * ostep), (lds_void*)&Sg_s[r0][0], 16, 0, 0);
        __builtin_amdgcn_global_load_lds((global_void*)(SH + ot + k + j * ostep), (lds_void*)&Sh_s[r0][0], 16, 0, 0);
      }
    }
  };

  d4 accF[BS][MS], accG[BS][MS], accO[BS][MS], accE[BS][MS];
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int j = 0; j < MS; ++j) {
      accF[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
      accG[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
      accO[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
      accE[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
    }
  double r0[RA], r1[RA], r2[RA], r3[RA];
  auto at = [](const char* base, unsigned off) { return (double)*(const InT*)(base + off); };
  auto fetch = [&](int kc) {
    const unsigned a = (unsigned)kc * sa, c = grp == 0 ? a : 0u - (unsigned)kc * 8u;
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      r0[r] = at(xb, o0[r] + a);
      r1[r] = at(xb, o1[r] - a);
      r2[r] = at(xb, o2[r] + c);
      r3[r] = at(xb, o3[r] - c);
    }
  };
  const int col = lane & 15;
  const int kq = lane >> 4;
  auto stage = [&](int kc) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int e = tid + 256 * r;
      const int ii = e % KC;
      const int i = kc + ii;
      const int bl = e / KC;
      double a, b;
      if (grp == 0) {
        const double v3 = 2 * i + 2 < H ? r3[r] : 0.0;  // k = H has no partner
        a = r0[r] + r1[r];                               // P_{2i+1}
        b = r2[r] + v3;                                  // P_{2i+2}
      } else {
        const double qk = r0[r] - r1[r];  // Q_{i+1}
        const double qh = r2[r] - r3[r];  // Q_{H-i-1}
        const bool last = i == Q - 1;     // k = Q is its own partner
        a = last ? qk : qk + qh;
        b = last ? 0.0 : qk - qh;
      }
      Xs[0][ii][bl ^ ii] = rowok[r] ? a : 0.0;
      Xs[1][ii][bl ^ ii] = rowok[r] ? b : 0.0;
    }
  };
  // the full product's input: Xs[0] (odd block) or Xs[1] (even block); T2's the other
  const int xf = grp, xh = 1 - grp;
  auto multiply = [&]() {
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int krow = ks * 4 + kq;
      double a[BS];
#pragma unroll
      for (int bs = 0; bs < BS; ++bs) a[bs] = Xs[xf][krow][(bh * (TB / 2) + bs * 16 + col) ^ krow];
#pragma unroll
      for (int ms = 0; ms < MS; ++ms) {
        const int mc = mh * (TM / 2) + ms * 16 + col;
        const double bf = Sf_s[krow][mc];
        const double bg = Sg_s[krow][TM - 1 - mc];
#pragma unroll
        for (int bs = 0; bs < BS; ++bs) {
          accF[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bs], bf, accF[bs][ms], 0, 0, 0);
          accG[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bs], bg, accG[bs][ms], 0, 0, 0);
        }
      }
    }
    // T2: even local rows (i even: k odd) into O, odd ones into E
#pragma unroll
    for (int ks = 0; ks < KC / 8; ++ks) {
      const int ro = 2 * (ks * 4 + kq), re = ro + 1;
      double ao[BS], ae[BS];
#pragma unroll
      for (int bs = 0; bs < BS; ++bs) {
        const int c = bh * (TB / 2) + bs * 16 + col;
        ao[bs] = Xs[xh][ro][c ^ ro];
        ae[bs] = Xs[xh][re][c ^ re];
      }
#pragma unroll
      for (int ms = 0; ms < MS; ++ms) {
        const int mc = mh * (TM / 2) + ms * 16 + col;
        const double bo = Sh_s[ro][mc];
        const double be = Sh_s[re][mc];
#pragma unroll
        for (int bs = 0; bs < BS; ++bs) {
          accO[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(ao[bs], bo, accO[bs][ms], 0, 0, 0);
          accE[bs][ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(ae[bs], be, accE[bs][ms], 0, 0, 0);
        }
      }
    }
  };
  // chunks in fold_chunk order (Q a multiple of 16: every chunk is whole)
  const int nb = Q / 16, nsteps = Q / KC;
  int kc = fold_chunk(grp, 0, KC, nb);
  fetch(kc);
  copy_tables(kc);
  for (int t = 0; t < nsteps; ++t) {  // the same for every thread: barriers stay uniform
    const int kn = t + 1 < nsteps ? fold_chunk(grp, t + 1, KC, nb) : 0;
    stage(kc);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this chunk's table copy has landed
    __syncthreads();
    if (t + 1 < nsteps) fetch(kn);
    multiply();
    __syncthreads();
    if (t + 1 < nsteps) copy_tables(kn);
    kc = kn;
  }

  // epilogue: the f64 16x16 accumulator holds D[row = lane/16 + 4r][col = lane%16]
  const int rq = lane >> 4;
#pragma unroll
  for (int bs = 0; bs < BS; ++bs)
#pragma unroll
    for (int ms = 0; ms < MS; ++ms) {
      const int m = m0 + mh * (TM / 2) + ms * 16 + col;
      if (m >= Q2) continue;
      const int mb = Q - 1 - m;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = b0 + bh * (TB / 2) + bs * 16 + rq + 4 * r;
        if (b >= Nb) continue;
        const double f = accF[bs][ms][r], g = accG[bs][ms][r];
        const double tp = accO[bs][ms][r] + accE[bs][ms][r];  // T2 product, row m
        const double tm = accO[bs][ms][r] - accE[bs][ms][r];  // row m~
        if (grp == 0) {
          store_out<OutT, EPI>(d, b, 2 * m, f + tp);            // row 2m+1
          store_out<OutT, EPI>(d, b, M - 2 - 2 * m, f - tp);    // row M-1-2m
          store_out<OutT, EPI>(d, b, 2 * mb, g + tm);           // row 2m~+1
          store_out<OutT, EPI>(d, b, M - 2 - 2 * mb, g - tm);   // row M-1-2m~
        } else {
          store_out<OutT, EPI>(d, b, 4 * m + 1, tp);                   // row 4m+2
          store_out<OutT, EPI>(d, b, 4 * mb + 1, tm);                  // row 4m~+2
          store_out<OutT, EPI>(d, b, 4 * m + 3, f);                    // row 4m+4
          if (mb < Q - 1) store_out<OutT, EPI>(d, b, 4 * mb + 3, g);  // row 4m~+4
        }
      }
    }
}

// ---------------------------------------------------------------------------------
// Mixed pass: fp32 operands split into bf16 hi + lo, three MFMA products, fp32 sums.
// Work-group tile: TB b x (TM odd + TM even) rows, K chunk 32 (one MFMA k-step), TB, TM
// in {32, 64} picked by the autotuner. The sine tables are stored per output row with k
// contiguous and zero padding to a multiple of 32 in both k and m; a chunk's TM rows of
// both tables are staged in LDS.
// ---------------------------------------------------------------------------------
constexpr int MX_KC = 32;
constexpr int MX_LDK = MX_KC + 8;  // padded LDS row (bf16 elements)

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// PF: register stages of prefetch (the loads of the next PF chunks in flight while one
// runs on the matrix cores). Inputs are held in their own type until folded: the FP32 sum
// or difference of two FP32 values is the same correctly rounded number however formed.
template <typename InT, typename OutT, int EPI, int TB, int TM, int PF>
__global__ __launch_bounds__(256) void dst_pass_mixed_kernel(const PassBatch batch) {
  constexpr int BS = TB / 32;  // 16-wide b subtiles per wave
  constexpr int MS = TM / 16;  // 16-wide m subtiles per wave
  constexpr int RA = TB / 16;  // input element pairs per thread and chunk
  int z, bt, mt;
  if (!xcd_tile(batch, z, bt, mt)) return;
  const PassPlane& d = batch.p[z];
  const int Nk = d.Nk, Nb = d.Nb;
  const int H = (Nk + 1) / 2;
  const bool modd = (Nk + 1) & 1;  // odd M: every k <= H has a partner, H even rows
  const int E = modd ? H : H - 1;
  const int Hp = (H + 31) & ~31;
  const int b0 = bt * TB;
  const int m0 = mt * TM;
  if (b0 >= Nb || m0 >= H) return;

  // A operand (folded input, split) and B operand (sine rows, split) of one chunk
  __shared__ __attribute__((aligned(16))) __bf16 Ah[2][TB][MX_LDK];  // [P|Q] hi
  __shared__ __attribute__((aligned(16))) __bf16 Al[2][TB][MX_LDK];  // [P|Q] lo
  __shared__ __attribute__((aligned(16))) __bf16 Bh[2][TM][MX_LDK];  // [odd|even] hi
  __shared__ __attribute__((aligned(16))) __bf16 Bl[2][TM][MX_LDK];  // [odd|even] lo

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int parity = wave >> 1;
  const int bh = wave & 1;
  const int col = lane & 15;
  const int kq = lane >> 4;

  // hi and lo tables are stored back to back: [hi (Mp x Hp)][lo (Mp x Hp)]
  const int Mp = (H + 31) / 32 * 32;
  const size_t tab = (size_t)Mp * Hp;
  const __bf16* So = (const __bf16*)d.So;
  const __bf16* Se = (const __bf16*)d.Se;

  f4 acc[BS][MS];
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int j = 0; j < MS; ++j) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};

  // loader mapping for the input: element pair (2kp, 2kp+1) of row bb, 16 lanes per row
  const InT* xin = (const InT*)d.in;
  InT ra[PF][RA][2], rm[PF][RA][2];
  // loader mapping for the sine rows: 16-byte piece (row sm, k8) of table st (4 tables)
  const int sm = tid >> 2, sk = (tid & 3) * 8;
  bf8 rs[PF][4];
  // input loads unconditional (k and b clamped into range), zeroed when staged, so the
  // chunk's loads issue back to back
  auto fetch = [&](const int st, int kc) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int e = tid + 256 * r;
      const int kp = e & 15;
      const int bc = min(b0 + (e >> 4), Nb - 1);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int k1c = min(kc + 2 * kp + j + 1, H);
        ra[st][r][j] = xin[(size_t)(k1c - 1) + (size_t)bc * Nk];
        rm[st][r][j] = xin[(size_t)(Nk - k1c) + (size_t)bc * Nk];
      }
    }
    const size_t off = (size_t)(m0 + sm) * Hp + kc + sk;  // m0 + sm < Mp by padding
    if (sm < TM && m0 + sm < Mp) {
      rs[st][0] = *(const bf8*)(So + off);
      rs[st][1] = *(const bf8*)(So + tab + off);
      rs[st][2] = *(const bf8*)(Se + off);
      rs[st][3] = *(const bf8*)(Se + tab + off);
    } else {
      for (int t = 0; t < 4; ++t)
        for (int i = 0; i < 8; ++i) rs[st][t][i] = (__bf16)0.f;
    }
  };
  auto stage = [&](const int st, int kc) {
#pragma unroll
    for (int r = 0; r < RA; ++r) {
      const int e = tid + 256 * r;
      const int kp = e & 15, bb = e >> 4;
      __bf16 ph[2], pl[2], qh[2], ql[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int k1 = kc + 2 * kp + j + 1;
        const bool ok = k1 <= H && b0 + bb < Nb;
        const InT va = ok ? ra[st][r][j] : (InT)0;
        const InT vm = ok && (k1 < H || modd) ? rm[st][r][j] : (InT)0;
        const bool mid = !modd && k1 == H;  // even M: k = H has no partner
        split_bf16((float)(va + vm), ph[j], pl[j]);
        split_bf16(mid ? 0.f : (float)(va - vm), qh[j], ql[j]);
      }
      typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
      *(bf2*)&Ah[0][bb][2 * kp] = (bf2){ph[0], ph[1]};
      *(bf2*)&Al[0][bb][2 * kp] = (bf2){pl[0], pl[1]};
      *(bf2*)&Ah[1][bb][2 * kp] = (bf2){qh[0], qh[1]};
      *(bf2*)&Al[1][bb][2 * kp] = (bf2){ql[0], ql[1]};
    }
    if (sm < TM) {
      *(bf8*)&Bh[0][sm][sk] = rs[st][0];
      *(bf8*)&Bl[0][sm][sk] = rs[st][1];
      *(bf8*)&Bh[1][sm][sk] = rs[st][2];
      *(bf8*)&Bl[1][sm][sk] = rs[st][3];
    }
  };
  auto multiply = [&]() {
    bf8 ah[BS], al[BS];
#pragma unroll
    for (int bs = 0; bs < BS; ++bs) {
      const int row = bh * (TB / 2) + bs * 16 + col;
      ah[bs] = *(const bf8*)&Ah[parity][row][kq * 8];
      al[bs] = *(const bf8*)&Al[parity][row][kq * 8];
    }
#pragma unroll
    for (int ms = 0; ms < MS; ++ms) {
      const bf8 bhv = *(const bf8*)&Bh[parity][ms * 16 + col][kq * 8];
      const bf8 blv = *(const bf8*)&Bl[parity][ms * 16 + col][kq * 8];
#pragma unroll
      for (int bs = 0; bs < BS; ++bs) {
        acc[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bs], bhv, acc[bs][ms], 0, 0, 0);
        acc[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bs], blv, acc[bs][ms], 0, 0, 0);
        acc[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bs], bhv, acc[bs][ms], 0, 0, 0);
      }
    }
  };
#pragma unroll
  for (int st = 0; st < PF; ++st)
    if (st * MX_KC < H) fetch(st, st * MX_KC);
  for (int kc = 0; kc < H; kc += PF * MX_KC) {
#pragma unroll
    for (int st = 0; st < PF; ++st) {
      const int k = kc + st * MX_KC;  // the same for every thread: barriers stay uniform
      if (k < H) {
        stage(st, k);
        __syncthreads();
        if (k + PF * MX_KC < H) fetch(st, k + PF * MX_KC);
        multiply();
        __syncthreads();
      }
    }
  }

  // the fp32 16x16 accumulator of the bf16 MFMA holds D[row = 4*(lane/16) + r][col = lane%16]
  const int rq = lane >> 4;
  const int mlim = parity == 0 ? H : E;
#pragma unroll
  for (int bs = 0; bs < BS; ++bs)
#pragma unroll
    for (int ms = 0; ms < MS; ++ms) {
      const int m = m0 + ms * 16 + col;
      if (m >= mlim) continue;
      const int jc = parity == 0 ? 2 * m : 2 * m + 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = b0 + bh * (TB / 2) + bs * 16 + 4 * rq + r;
        if (b < Nb) store_out<OutT, EPI>(d, b, jc, (double)acc[bs][ms][r]);
      }
    }
}

// ---------------------------------------------------------------------------------
// Mixed pass with the second fold (M = N+1 = 4Q): the contraction of dst_pass_f64q_kernel
// (odd rows Y_{2m+1}, Y_{M-1-2m} = A_m +- B_m from T1 . P_{2i+1} and T2 . P_{2i+2}; even
// rows Y_{4m+2}, Y_{4m+4} from T2 . P'_{i+1} and T4 . Q'_{i+1}) with the folded inputs and
// the tables split into bf16 hi + lo and three MFMA products per pair (hi*hi + hi*lo +
// lo*hi), fp32 sums: half the products of the single fold, and half its LDS reads of
// table fragments, which bound that pass. Folding runs in the input's type, rounded once
// to fp32 before the split. Tables: [T1 | T2 | T4], each hi then lo, Qm x Qm bf16 with k
// contiguous per row m (Qm = Q rounded up to 32; zero beyond each table). Work-group: one
// block (grp 0 odd rows, grp 1 even rows), TB values of b x TM quarter rows, K chunk 32 in
// fold_chunk<32> order, the loads of the next PF chunks in flight; wave w computes the
// (b-half w&1, m-half w>>1) quarter of both of the block's products, so the odd block's
// butterfly A +- B stays in registers. PF = 0: one stage of inputs, and the table tiles go
// from global memory straight to LDS (global_load_lds_dwordx4, no registers), 64-byte rows
// whose 16-byte pieces are XOR-swizzled by row (piece c of row r at c ^ ((r >> 2) & 3)) so
// that the 16 rows an MFMA operand read touches fall in distinct banks. The 32-row tiles
// with one stage of fp32 inputs are held to five work-groups per CU: the 1280 tiles of a
// 5-plane 1023 x 1023 pass then run in one round (four per CU leave a second round a
// quarter full).
// ---------------------------------------------------------------------------------
template <typename InT, typename OutT, int EPI, int TB, int TM, int PF>
__global__ __launch_bounds__(256, (TB == 32 && PF <= 1 && sizeof(InT) == 4) ? 5 : 1) void dst_pass_mixedq_kernel(
    const PassBatch batch) {
  constexpr bool DT = PF == 0;       // tables straight into LDS
  constexpr int NS = DT ? 1 : PF;    // register stages of prefetch
  constexpr int BS = TB / 32;  // 16-wide b subtiles per wave
  constexpr int MS = TM / 32;  // 16-wide m subtiles per wave
  constexpr int RA = TB / 16;  // (b, i-pair) input slots per thread and chunk
  static_assert(BS >= 1 && MS >= 1 && RA >= 1, "tile");
  int z, bt, mt2;
  if (!xcd_tile(batch, z, bt, mt2)) return;
  const int grp = mt2 & 1;  // 0: odd rows, 1: even rows
  const int mt = mt2 >> 1;
  const PassPlane& d = batch.p[z];
  const int Nk = d.Nk, Nb = d.Nb;
  const int M = Nk + 1, H = M / 2, Q = M / 4, Qm = (Q + 31) & ~31;
  const int b0 = bt * TB;
  const int m0 = mt * TM;
  if (b0 >= Nb || m0 >= Q) return;

  __shared__ __attribute__((aligned(16))) __bf16 Xh[2][TB][MX_LDK];  // [X0|X1] hi
  __shared__ __attribute__((aligned(16))) __bf16 Xl[2][TB][MX_LDK];  // [X0|X1] lo
  __shared__ __attribute__((aligned(16))) __bf16 Th[2][DT ? 1 : TM][MX_LDK];  // [SA|SB] hi
  __shared__ __attribute__((aligned(16))) __bf16 Tl[2][DT ? 1 : TM][MX_LDK];  // [SA|SB] lo
  __shared__ __attribute__((aligned(16))) __bf16 Tq[DT ? 4 : 1][DT ? TM : 1][MX_KC];  // DT: SA hi, lo, SB hi, lo

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int bh = wave & 1;
  const int mh = wave >> 1;
  const int col = lane & 15;
  const int kq = lane >> 4;

  const size_t tq = (size_t)Qm * Qm;  // one table half (hi or lo)
  const __bf16* T1 = (const __bf16*)d.Sqm;
  const __bf16* SA = grp == 0 ? T1 : T1 + 2 * tq;
  const __bf16* SB = grp == 0 ? T1 + 2 * tq : T1 + 4 * tq;

  f4 acc0[BS][MS], acc1[BS][MS];
#pragma unroll
  for (int i = 0; i < BS; ++i)
#pragma unroll
    for (int j = 0; j < MS; ++j) {
      acc0[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
      acc1[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
    }

  // input quad (line bl, fold indices i0..i0+3 with i0 = kc + 4 (tid % 8)): grp 0 reads the
  // runs x_{2i0+1..2i0+8} and x_{M-2i0-8..M-2i0-1}; grp 1 x_{i0+1..i0+4}, x_{M-i0-4..M-i0-1},
  // x_{H-i0-4..H-i0-1} and x_{H+i0+1..H+i0+4}: four contiguous runs, each read as 16-byte
  // vector loads from 32-bit element offsets (dword-aligned; planes below 2^31 bytes). With
  // Q a multiple of 4 a quad is wholly inside or outside the range; loads are
  // unconditional (i0 and b clamped into range), values of out-of-range quads zeroed when
  // staged.
  constexpr int RQ = TB / 32;                // quads per thread and chunk
  constexpr int VE = 16 / (int)sizeof(InT);  // elements per 16-byte load
  typedef InT vec16 __attribute__((ext_vector_type(VE), aligned(sizeof(InT))));
  const char* xb = (const char*)d.in;
  auto run = [&](unsigned e, InT* out, int n) {  // n consecutive elements from element e
#pragma unroll
    for (int v = 0; v < n; v += VE) {
      const vec16 w = *(const vec16*)(xb + (size_t)((e + v) * (unsigned)sizeof(InT)));
#pragma unroll
      for (int u = 0; u < VE; ++u) out[v + u] = w[u];
    }
  };
  unsigned line[RQ];
#pragma unroll
  for (int r = 0; r < RQ; ++r) line[r] = (unsigned)min(b0 + (tid >> 3) + 32 * r, Nb - 1) * (unsigned)Nk;
  const int q4 = 4 * (tid & 7);
  InT rx[NS][RQ][16];
  // table loader: 16-byte piece (row sm, k8) of SA hi, SA lo, SB hi, SB lo
  const int sm = tid >> 2, sk = (tid & 3) * 8;
  const char* SAc = (const char*)SA;
  const char* SBc = (const char*)SB;
  const unsigned toff = (unsigned)((m0 + sm) * Qm + sk) * 2u, tqb = (unsigned)tq * 2u;
  bf8 rs[NS][4];
  // DT: wave instruction j copies 16 rows of part j / (TM / 16) (SA hi, SA lo, SB hi, SB lo);
  // lane L fills row 16 (j % (TM / 16)) + L / 4, piece L % 4, with logical piece
  // (L % 4) ^ ((row >> 2) & 3) of that row (rows past the table read its last row: their
  // outputs are never stored)
  constexpr int NTI = 4 * TM / 16;  // wave instructions per chunk
  auto copy_tables = [&](int kc) {
#pragma unroll
    for (int jj = 0; jj < (NTI + 3) / 4; ++jj) {
      const int j = wave * ((NTI + 3) / 4) + jj;
      if (j < NTI) {
        const int part = j / (TM / 16), row = 16 * (j % (TM / 16)) + (lane >> 2);
        const int piece = (lane & 3) ^ ((row >> 2) & 3);
        const char* src = (const char*)(part < 2 ? SA : SB) + (part & 1) * tqb +
                          (size_t)((unsigned)(min(m0 + row, Qm - 1) * Qm + kc + 8 * piece) * 2u);
        __builtin_amdgcn_global_load_lds((global_void*)src, (lds_void*)&Tq[part][16 * (j % (TM / 16))][0], 16, 0, 0);
      }
    }
  };
  auto fetch = [&](const int st, int kc) {
    const unsigned i0 = (unsigned)min(kc + q4, Q - 4);
#pragma unroll
    for (int r = 0; r < RQ; ++r) {
      const unsigned l = line[r];
      InT* o = rx[st][r];
      if (grp == 0) {
        run(l + 2 * i0, o, 8);                               // x_{2i0+1..2i0+8}
        run(l + (unsigned)(M - 9) - 2 * i0, o + 8, 8);       // x_{M-2i0-8..M-2i0-1}
      } else {
        run(l + i0, o, 4);                                   // x_{i0+1..i0+4}
        run(l + (unsigned)(M - 5) - i0, o + 4, 4);           // x_{M-i0-4..M-i0-1}
        run(l + (unsigned)(H - 5) - i0, o + 8, 4);           // x_{H-i0-4..H-i0-1}
        run(l + (unsigned)H + i0, o + 12, 4);                // x_{H+i0+1..H+i0+4}
      }
    }
    if (DT) return;
    const unsigned off = toff + (unsigned)kc * 2u;
    if (sm < TM && m0 + sm < Qm) {
      rs[st][0] = *(const bf8*)(SAc + off);
      rs[st][1] = *(const bf8*)(SAc + tqb + off);
      rs[st][2] = *(const bf8*)(SBc + off);
      rs[st][3] = *(const bf8*)(SBc + tqb + off);
    } else {
      for (int t = 0; t < 4; ++t)
        for (int i = 0; i < 8; ++i) rs[st][t][i] = (__bf16)0.f;
    }
  };
  auto stage = [&](const int st, int kc) {
    typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
    const int i0 = kc + q4;
#pragma unroll
    for (int r = 0; r < RQ; ++r) {
      const int bl = (tid >> 3) + 32 * r;
      const bool ok = i0 < Q && b0 + bl < Nb;
      const InT* v = rx[st][r];
      bf4 ah, al, ch, cl;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = i0 + j;
        InT a, c;
        if (grp == 0) {
          a = v[2 * j] + v[15 - 2 * j];                              // P_{2i+1}
          c = v[2 * j + 1] + (2 * i + 2 < H ? v[14 - 2 * j] : (InT)0);  // P_{2i+2} (k = H: no partner)
        } else {
          const InT qk = v[j] - v[7 - j];      // Q_{i+1}
          const InT qh = v[11 - j] - v[12 + j];  // Q_{H-i-1}
          const bool last = i == Q - 1;        // k = Q is its own partner
          a = last ? qk : qk + qh;
          c = last ? (InT)0 : qk - qh;
        }
        __bf16 h, lo;
        split_bf16(ok ? (float)a : 0.f, h, lo);
        ah[j] = h;
        al[j] = lo;
        split_bf16(ok ? (float)c : 0.f, h, lo);
        ch[j] = h;
        cl[j] = lo;
      }
      *(bf4*)&Xh[0][bl][q4] = ah;
      *(bf4*)&Xl[0][bl][q4] = al;
      *(bf4*)&Xh[1][bl][q4] = ch;
      *(bf4*)&Xl[1][bl][q4] = cl;
    }
    if (!DT && sm < TM) {
      *(bf8*)&Th[0][sm][sk] = rs[st][0];
      *(bf8*)&Tl[0][sm][sk] = rs[st][1];
      *(bf8*)&Th[1][sm][sk] = rs[st][2];
      *(bf8*)&Tl[1][sm][sk] = rs[st][3];
    }
  };
  auto multiply = [&]() {
    bf8 x0h[BS], x0l[BS], x1h[BS], x1l[BS];
#pragma unroll
    for (int bs = 0; bs < BS; ++bs) {
      const int row = bh * (TB / 2) + bs * 16 + col;
      x0h[bs] = *(const bf8*)&Xh[0][row][kq * 8];
      x0l[bs] = *(const bf8*)&Xl[0][row][kq * 8];
      x1h[bs] = *(const bf8*)&Xh[1][row][kq * 8];
      x1l[bs] = *(const bf8*)&Xl[1][row][kq * 8];
    }
#pragma unroll
    for (int ms = 0; ms < MS; ++ms) {
      const int row = mh * (TM / 2) + ms * 16 + col;
      bf8 sah, sal, sbh, sbl;
      if (DT) {
        const int pc = 8 * (kq ^ ((row >> 2) & 3));
        sah = *(const bf8*)&Tq[0][row][pc];
        sal = *(const bf8*)&Tq[1][row][pc];
        sbh = *(const bf8*)&Tq[2][row][pc];
        sbl = *(const bf8*)&Tq[3][row][pc];
      } else {
        sah = *(const bf8*)&Th[0][row][kq * 8];
        sal = *(const bf8*)&Tl[0][row][kq * 8];
        sbh = *(const bf8*)&Th[1][row][kq * 8];
        sbl = *(const bf8*)&Tl[1][row][kq * 8];
      }
#pragma unroll
      for (int bs = 0; bs < BS; ++bs) {
        acc0[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x0l[bs], sah, acc0[bs][ms], 0, 0, 0);
        acc0[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x0h[bs], sal, acc0[bs][ms], 0, 0, 0);
        acc0[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x0h[bs], sah, acc0[bs][ms], 0, 0, 0);
        acc1[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x1l[bs], sbh, acc1[bs][ms], 0, 0, 0);
        acc1[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x1h[bs], sbl, acc1[bs][ms], 0, 0, 0);
        acc1[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x1h[bs], sbh, acc1[bs][ms], 0, 0, 0);
      }
    }
  };
  const int nb = (Q + MX_KC - 1) / MX_KC;  // chunks, taken in fold_chunk order
  if (DT) {
    // per chunk: stage the inputs fetched one chunk ahead, wait for the table copy,
    // barrier, fetch the next inputs, multiply, barrier, start the next table copy
    int kc = fold_chunk<MX_KC>(grp, 0, MX_KC, nb);
    fetch(0, kc);
    copy_tables(kc);
    for (int t = 0; t < nb; ++t) {  // the same for every thread: barriers stay uniform
      const int kn = t + 1 < nb ? fold_chunk<MX_KC>(grp, t + 1, MX_KC, nb) : 0;
      stage(0, kc);
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this chunk's table copy has landed
      __syncthreads();
      if (t + 1 < nb) fetch(0, kn);
      multiply();
      __syncthreads();
      if (t + 1 < nb) copy_tables(kn);
      kc = kn;
    }
  } else {
#pragma unroll
    for (int st = 0; st < NS; ++st)
      if (st < nb) fetch(st, fold_chunk<MX_KC>(grp, st, MX_KC, nb));
    for (int t0 = 0; t0 < nb; t0 += NS) {
#pragma unroll
      for (int st = 0; st < NS; ++st) {
        const int t = t0 + st;  // the same for every thread: barriers stay uniform
        if (t < nb) {
          stage(st, fold_chunk<MX_KC>(grp, t, MX_KC, nb));
          __syncthreads();
          if (t + NS < nb) fetch(st, fold_chunk<MX_KC>(grp, t + NS, MX_KC, nb));
          multiply();
          __syncthreads();
        }
      }
    }
  }

  // the fp32 16x16 accumulator of the bf16 MFMA holds D[row = 4*(lane/16) + r][col = lane%16]
  const int rq = lane >> 4;
#pragma unroll
  for (int bs = 0; bs < BS; ++bs)
#pragma unroll
    for (int ms = 0; ms < MS; ++ms) {
      const int m = m0 + mh * (TM / 2) + ms * 16 + col;
      if (m >= Q) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = b0 + bh * (TB / 2) + bs * 16 + 4 * rq + r;
        if (b >= Nb) continue;
        const double u = acc0[bs][ms][r], v = acc1[bs][ms][r];
        if (grp == 0) {
          store_out<OutT, EPI>(d, b, 2 * m, u + v);          // row 2m+1
          store_out<OutT, EPI>(d, b, M - 2 - 2 * m, u - v);  // row M-1-2m
        } else {
          store_out<OutT, EPI>(d, b, 4 * m + 1, u);                  // row 4m+2
          if (m < Q - 1) store_out<OutT, EPI>(d, b, 4 * m + 3, v);  // row 4m+4
        }
      }
    }
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
// Tile variants (fold, TB, TM, prefetch depth, K chunk); the autotuner times them on a
// launch shape and keeps the best. FP64: 0..7 single fold, 8..15 double fold (TM counts
// quarter rows there), 16..19 double fold with the tables copied straight into LDS
// (dst_pass_f64d_kernel; where Q is not a multiple of 16 they run as variant 10 or 11),
// 20..23 the third fold (dst_pass_f64t_kernel, TM of the Q/2 rows m; where it does not
// apply they run as the direct double-fold variant of the same TM and K chunk).
// The mixed pass has its own 20: 0..7 one fold (4..7 two prefetch stages), 8..19 the
// double fold (dst_pass_mixedq_kernel) with 1 or 2 prefetch stages (8..15) or with the
// tables copied straight into LDS (16..19); v % 4 picks the tiles as for 0..15.
constexpr int kVariants = kDstVariants;
constexpr int kMixedVariants = 20;
// the single-fold mixed variant of the same tiles (two prefetch stages for 12..19)
static int mixed_single(int v) { return v < 8 ? v : (v < 12 ? v % 4 : 4 + v % 4); }
constexpr int kVarTB[kVariants] = {64, 64, 32, 32, 64, 64, 32, 32, 64, 64, 32, 32,
                                   64, 64, 32, 32, 32, 32, 32, 32, 32, 32, 32, 32};
constexpr int kVarTM[kVariants] = {64, 32, 64, 32, 64, 32, 64, 32, 64, 32, 64, 32,
                                   64, 32, 64, 32, 64, 64, 32, 32, 64, 64, 32, 32};
constexpr int kVarPF[kVariants] = {1, 1, 1, 1, 2, 2, 2, 2, 1, 1, 1, 1,
                                   2, 2, 2, 2, 0, 0, 0, 0, 0, 0, 0, 0};
constexpr int kVarFold[kVariants] = {2, 2, 2, 2, 2, 2, 2, 2, 4, 4, 4, 4,
                                     4, 4, 4, 4, 4, 4, 4, 4, 8, 8, 8, 8};
constexpr int kVarKC[kVariants] = {16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16, 16,
                                   16, 16, 16, 16, 16, 8, 16, 8, 16, 8, 16, 8};
constexpr int kFirstDirect = 16, kFirstTriple = 20;
static_assert(kVariants == 24, "variant tables");

static dim3 tile_grid(PassBatch& b, int nplanes, int maxNb, int maxH, int TB, int TM) {
  b.nbt = (maxNb + TB - 1) / TB;
  b.nmt = (maxH + TM - 1) / TM;
  b.nplanes = nplanes;
  const int T = nplanes * b.nbt * b.nmt;
  return dim3(8 * ((T + 7) / 8));
}

template <int KC, int PF>
struct F64KC {
  template <typename InT, typename OutT, int EPI, int TB, int TM>
  struct K {
    static void go(dim3 g, hipStream_t st, const PassBatch& b) {
      hipLaunchKernelGGL((dst_pass_f64_kernel<InT, OutT, EPI, TB, TM, KC, PF>), g, dim3(256), 0, st, b);
    }
  };
};
template <int PF>
struct MixKC {
  template <typename InT, typename OutT, int EPI, int TB, int TM>
  struct K {
    static void go(dim3 g, hipStream_t st, const PassBatch& b) {
      hipLaunchKernelGGL((dst_pass_mixed_kernel<InT, OutT, EPI, TB, TM, PF>), g, dim3(256), 0, st, b);
    }
  };
};

template <int PF>
struct MixQ {
  template <typename InT, typename OutT, int EPI, int TB, int TM>
  struct K {
    static constexpr bool quarter = true;
    static void go(dim3 g, hipStream_t st, const PassBatch& b) {
      hipLaunchKernelGGL((dst_pass_mixedq_kernel<InT, OutT, EPI, TB, TM, PF>), g, dim3(256), 0, st, b);
    }
  };
};

template <int KC, int PF>
struct F64QKC {
  template <typename InT, typename OutT, int EPI, int TB, int TM>
  struct K {
    static constexpr bool quarter = true;
    static void go(dim3 g, hipStream_t st, const PassBatch& b) {
      hipLaunchKernelGGL((dst_pass_f64q_kernel<InT, OutT, EPI, TB, TM, KC, PF>), g, dim3(256), 0, st, b);
    }
  };
};

// The third fold: tiles of the Q/2 rows m (the launch grid counts those).
template <int KC>
struct F64TKC {
  template <typename InT, typename OutT, int EPI, int TB, int TM>
  struct K {
    static constexpr bool quarter = true, eighth = true;
    static void go(dim3 g, hipStream_t st, const PassBatch& b) {
      hipLaunchKernelGGL((dst_pass_f64t_kernel<InT, OutT, EPI, TB, TM, KC, TM == 64 ? 3 : 5>), g,
                         dim3(256), 0, st, b);
    }
  };
};

// Direct table copies: five work-groups per CU for 64-wide tiles, six for 32-wide ones.
template <int KC>
struct F64DKC {
  template <typename InT, typename OutT, int EPI, int TB, int TM>
  struct K {
    static constexpr bool quarter = true;
    static void go(dim3 g, hipStream_t st, const PassBatch& b) {
      hipLaunchKernelGGL((dst_pass_f64d_kernel<InT, OutT, EPI, TB, TM, KC, TM == 64 ? 5 : 6>), g,
                         dim3(256), 0, st, b);
    }
  };
};

template <class T, class = void>
struct is_quarter {
  static constexpr bool value = false;
};
template <class T>
struct is_quarter<T, decltype((void)T::quarter)> {
  static constexpr bool value = T::quarter;
};
template <class T, class = void>
struct is_eighth {
  static constexpr bool value = false;
};
template <class T>
struct is_eighth<T, decltype((void)T::eighth)> {
  static constexpr bool value = T::eighth;
};

template <template <typename, typename, int, int, int> class K, typename InT, typename OutT,
          int TB, int TM>
static void launch_epi(const PassBatch& b_in, int nplanes, int maxNb, int maxH, int epi,
                       hipStream_t st) {
  PassBatch b = b_in;
  b.skip = g_skip;
  // double fold: TM quarter rows of each of the two blocks (odd, even) per tile
  constexpr bool quarter = is_quarter<K<InT, OutT, EPI_STORE, TB, TM>>::value;
  constexpr bool eighth = is_eighth<K<InT, OutT, EPI_STORE, TB, TM>>::value;
  dim3 g = tile_grid(b, nplanes, maxNb, eighth ? maxH / 4 : (quarter ? maxH / 2 : maxH), TB, TM);
  if (quarter) {
    b.nmt *= 2;
    g = dim3(8 * ((nplanes * b.nbt * b.nmt + 7) / 8));
  }
  switch (epi) {
    case EPI_STORE: K<InT, OutT, EPI_STORE, TB, TM>::go(g, st, b); break;
    case EPI_DIAG: K<InT, OutT, EPI_DIAG, TB, TM>::go(g, st, b); break;
    case EPI_SUB: K<InT, OutT, EPI_SUB, TB, TM>::go(g, st, b); break;
    default: K<InT, OutT, EPI_ADD, TB, TM>::go(g, st, b); break;
  }
}

template <template <typename, typename, int, int, int> class K, typename InT, typename OutT>
static void launch_variant(int v, const PassBatch& b, int nplanes, int maxNb, int maxH, int epi,
                           hipStream_t st) {
  switch (v) {
    case 1: launch_epi<K, InT, OutT, 64, 32>(b, nplanes, maxNb, maxH, epi, st); break;
    case 2: launch_epi<K, InT, OutT, 32, 64>(b, nplanes, maxNb, maxH, epi, st); break;
    case 3: launch_epi<K, InT, OutT, 32, 32>(b, nplanes, maxNb, maxH, epi, st); break;
    default: launch_epi<K, InT, OutT, 64, 64>(b, nplanes, maxNb, maxH, epi, st); break;
  }
}

static void launch_f64_variant(int v, const PassBatch& b, int nplanes, int maxNb, int maxH, int epi,
                               hipStream_t st) {
  if (v >= kFirstTriple) {
    const bool tm64 = kVarTM[v] == 64;
    if (kVarKC[v] == 8) {
      if (tm64) launch_epi<F64TKC<8>::K, double, double, 32, 64>(b, nplanes, maxNb, maxH, epi, st);
      else launch_epi<F64TKC<8>::K, double, double, 32, 32>(b, nplanes, maxNb, maxH, epi, st);
    } else {
      if (tm64) launch_epi<F64TKC<16>::K, double, double, 32, 64>(b, nplanes, maxNb, maxH, epi, st);
      else launch_epi<F64TKC<16>::K, double, double, 32, 32>(b, nplanes, maxNb, maxH, epi, st);
    }
  } else if (v >= kFirstDirect) {
    const bool tm64 = kVarTM[v] == 64;  // TB is 32 for all of them
    if (kVarKC[v] == 8) {
      if (tm64) launch_epi<F64DKC<8>::K, double, double, 32, 64>(b, nplanes, maxNb, maxH, epi, st);
      else launch_epi<F64DKC<8>::K, double, double, 32, 32>(b, nplanes, maxNb, maxH, epi, st);
    } else {
      if (tm64) launch_epi<F64DKC<16>::K, double, double, 32, 64>(b, nplanes, maxNb, maxH, epi, st);
      else launch_epi<F64DKC<16>::K, double, double, 32, 32>(b, nplanes, maxNb, maxH, epi, st);
    }
  } else if (v >= 12) launch_variant<F64QKC<16, 2>::K, double, double>(v - 12, b, nplanes, maxNb, maxH, epi, st);
  else if (v >= 8) launch_variant<F64QKC<16, 1>::K, double, double>(v - 8, b, nplanes, maxNb, maxH, epi, st);
  else if (v >= 4) launch_variant<F64KC<16, 2>::K, double, double>(v - 4, b, nplanes, maxNb, maxH, epi, st);
  else launch_variant<F64KC<16, 1>::K, double, double>(v, b, nplanes, maxNb, maxH, epi, st);
}

template <template <typename, typename, int, int, int> class K>
static void launch_mixed_types(int v, const PassBatch& b, int nplanes, int maxNb, int maxH, int epi,
                               bool in_f32, bool out_f32, hipStream_t st) {
  if (!in_f32 && out_f32) launch_variant<K, double, float>(v, b, nplanes, maxNb, maxH, epi, st);
  else if (in_f32 && out_f32) launch_variant<K, float, float>(v, b, nplanes, maxNb, maxH, epi, st);
  else if (in_f32 && !out_f32) launch_variant<K, float, double>(v, b, nplanes, maxNb, maxH, epi, st);
  else launch_variant<K, double, double>(v, b, nplanes, maxNb, maxH, epi, st);
}

static void launch_mixed_variant(int v, const PassBatch& b, int nplanes, int maxNb, int maxH, int epi,
                                 bool in_f32, bool out_f32, hipStream_t st) {
  if (v >= 16) launch_mixed_types<MixQ<0>::K>(v - 16, b, nplanes, maxNb, maxH, epi, in_f32, out_f32, st);
  else if (v >= 12) launch_mixed_types<MixQ<2>::K>(v - 12, b, nplanes, maxNb, maxH, epi, in_f32, out_f32, st);
  else if (v >= 8) launch_mixed_types<MixQ<1>::K>(v - 8, b, nplanes, maxNb, maxH, epi, in_f32, out_f32, st);
  else if (v >= 4) launch_mixed_types<MixKC<2>::K>(v - 4, b, nplanes, maxNb, maxH, epi, in_f32, out_f32, st);
  else launch_mixed_types<MixKC<1>::K>(v, b, nplanes, maxNb, maxH, epi, in_f32, out_f32, st);
}

// ---- autotuner ----
// Process-wide table: launch shape (precision, planes, Nb and H of the largest plane) ->
// the fastest variant measured on it. Each path is tuned over the variants of the
// deepest fold the lengths allow (FP64: the third fold, else the double, else the single
// one; mixed: the double, else the single), so every launch of a shape sums in the same
// order whichever tiles it gets; shapes never tuned take the first variant of that fold
// (FP64 20, mixed 10), which falls back like any variant where its fold does not apply.
int dst_default_variant(bool mixed) { return mixed ? 10 : kFirstTriple; }
namespace {
struct TuneKey {
  int mixed, nplanes, maxNb, maxH;
  bool operator<(const TuneKey& o) const {
    if (mixed != o.mixed) return mixed < o.mixed;
    if (nplanes != o.nplanes) return nplanes < o.nplanes;
    if (maxNb != o.maxNb) return maxNb < o.maxNb;
    return maxH < o.maxH;
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;

int tuned_variant(int mixed, int nplanes, int maxNb, int maxH) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(TuneKey{mixed, nplanes, maxNb, maxH});
  return it == g_tune.end() ? dst_default_variant(mixed != 0) : it->second;
}
}  // namespace

// The double fold needs M = N + 1 to be a multiple of 4 on every plane of the launch (in
// mixed precision a multiple of 16, whole quads of fold indices, and the split double-fold
// tables); elsewhere a double-fold variant runs as the single-fold variant of the same
// tiles.
static bool quarter_ok(const PassBatch& b, int nplanes, bool mixed = false) {
  for (int i = 0; i < nplanes; ++i) {
    if (mixed ? !(mixed_quarter_ok(b.p[i].Nk) && b.p[i].Sqm) : (b.p[i].Nk + 1) % 4 != 0) return false;
  }
  return true;
}

// The direct-copy variants need Q = (N + 1) / 4 a multiple of 16 (whole blocks of
// fold_chunk, whatever their K chunk) on every plane, and planes small enough for 32-bit
// byte offsets; elsewhere they run as the double-fold variant of the same tiles with
// register staging.
static bool direct_ok(const PassBatch& b, int nplanes, int KC) {
  (void)KC;
  for (int i = 0; i < nplanes; ++i) {
    const int Q = (b.p[i].Nk + 1) / 4;
    if (Q % 16 || (size_t)b.p[i].Nk * b.p[i].Nb * 8 >= ((size_t)1 << 31)) return false;
  }
  return true;
}
static int direct_fallback(int v) { return v < kFirstDirect + 2 ? 10 : 11; }

// The third fold needs, besides what the direct variants need, TM <= Q/2 on every plane
// (a tile's mirror columns m~ stay inside the tables); elsewhere it runs as the direct
// double-fold variant of the same TM and K chunk.
static bool triple_ok(const PassBatch& b, int nplanes, int TM) {
  for (int i = 0; i < nplanes; ++i)
    if ((b.p[i].Nk + 1) / 4 / 2 < TM) return false;
  return true;
}
static int triple_fallback(int v) { return kFirstDirect + (v - kFirstTriple); }

int dst_effective_variant(int v, int N, bool mixed) {
  if (mixed) return v >= 8 && !mixed_quarter_ok(N) ? mixed_single(v) : v;
  const int Q = (N + 1) / 4;
  // a 64-row third-fold variant takes 32-row tiles where only those fit (the same bits)
  if (v >= kFirstTriple && kVarTM[v] == 64 && Q / 2 < 64 && Q / 2 >= 32) v += 2;
  if (v >= kFirstTriple && ((N + 1) % 4 || Q % 16 || Q / 2 < kVarTM[v])) v = triple_fallback(v);
  if (v >= kFirstDirect && ((N + 1) % 4 || Q % 16)) v = direct_fallback(v);
  return v >= 8 && (N + 1) % 4 ? v - 8 : v;
}

hipError_t launch_dst_pass_f64(const PassBatch& batch, int nplanes, int maxNb, int maxH, int epi,
                               hipStream_t st, int variant) {
  int v = variant >= 0 && variant < kVariants ? variant : tuned_variant(0, nplanes, maxNb, maxH);
  // a 64-row third-fold variant takes 32-row tiles where only those fit (the same bits)
  if (v >= kFirstTriple && kVarTM[v] == 64 && !triple_ok(batch, nplanes, 64) &&
      triple_ok(batch, nplanes, 32))
    v += 2;
  if (v >= kFirstTriple && !(quarter_ok(batch, nplanes) && direct_ok(batch, nplanes, kVarKC[v]) &&
                             triple_ok(batch, nplanes, kVarTM[v])))
    v = triple_fallback(v);
  if (v >= kFirstDirect && !(quarter_ok(batch, nplanes) && direct_ok(batch, nplanes, kVarKC[v])))
    v = direct_fallback(v);
  if (v >= 8 && !quarter_ok(batch, nplanes)) v -= 8;
  launch_f64_variant(v, batch, nplanes, maxNb, maxH, epi, st);
  return hipGetLastError();
}

hipError_t launch_dst_pass_mixed(const PassBatch& batch, int nplanes, int maxNb, int maxH, int epi,
                                 bool in_f32, bool out_f32, hipStream_t st, int variant) {
  int v = variant >= 0 && variant < kMixedVariants ? variant : tuned_variant(1, nplanes, maxNb, maxH);
  if (v >= 8 && !quarter_ok(batch, nplanes, true)) v = mixed_single(v);
  launch_mixed_variant(v, batch, nplanes, maxNb, maxH, epi, in_f32, out_f32, st);
  return hipGetLastError();
}

hipError_t tune_dst_pass(const PassBatch& batch, int nplanes, int maxNb, int maxH, bool mixed,
                         bool in_f32, bool out_f32, hipStream_t st, int* chosen, float* best_ms) {
  const TuneKey key{mixed ? 1 : 0, nplanes, maxNb, maxH};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      if (chosen) *chosen = it->second;
      if (best_ms) *best_ms = -1.f;
      return hipSuccess;
    }
  }
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return e;
  }
  // three rounds over the variants, best mean of 4 launches per round (interleaving the
  // rounds spreads clock and power-state drift over all variants)
  constexpr int reps = 4, rounds = 3;
  // the variants of the deepest fold the lengths allow: FP64 the third fold (20..23), else
  // the double fold (8..19), else the single fold (0..7); mixed the double fold, else the
  // single one
  const bool q = quarter_ok(batch, nplanes, mixed);
  const bool tq = !mixed && q && direct_ok(batch, nplanes, 8) && triple_ok(batch, nplanes, 32);
  const int v0 = tq ? kFirstTriple : (q ? 8 : 0);
  const int v1 = tq ? kVariants : (q ? (mixed ? kMixedVariants : kFirstTriple) : 8);
  float t[kVariants];
  for (int v = 0; v < kVariants; ++v) t[v] = 1e30f;
  for (int round = 0; round < rounds && e == hipSuccess; ++round) {
    for (int v = v0; v < v1 && e == hipSuccess; ++v) {
      if (v >= kFirstDirect && !direct_ok(batch, nplanes, kVarKC[v])) continue;
      if (v >= kFirstTriple && !triple_ok(batch, nplanes, kVarTM[v])) continue;
      auto run = [&]() {
        if (mixed) launch_mixed_variant(v, batch, nplanes, maxNb, maxH, EPI_STORE, in_f32, out_f32, st);
        else launch_f64_variant(v, batch, nplanes, maxNb, maxH, EPI_STORE, st);
      };
      run();  // warm-up
      e = hipEventRecord(e0, st);
      for (int r = 0; r < reps && e == hipSuccess; ++r) run();
      if (e == hipSuccess) e = hipGetLastError();
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e == hipSuccess && ms / reps < t[v]) t[v] = ms / reps;
    }
  }
  int best = v0;
  float best_t = t[v0];
  for (int v = v0 + 1; v < v1; ++v)
    if (t[v] < best_t) {
      best_t = t[v];
      best = v;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return e;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best;
  }
  if (chosen) *chosen = best;
  if (best_ms) *best_ms = best_t;
  return hipSuccess;
}

int dst_tuned_variant(int mixed, int nplanes, int maxNb, int maxH) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(TuneKey{mixed ? 1 : 0, nplanes, maxNb, maxH});
  return it == g_tune.end() ? -1 : it->second;
}

int dst_variant_tiles(int v, bool mixed, int* TB, int* TM, int* PF, int* fold, int* KC) {
  if (v < 0 || v >= (mixed ? kMixedVariants : kVariants)) return 0;
  if (mixed) {  // (TB, TM) by v % 4 as for 0..15; PF 1, 2 (one fold), 1, 2, 0 (double fold)
    *TB = kVarTB[v % 4];
    *TM = kVarTM[v % 4];
    *PF = v < 8 ? 1 + v / 4 : (v < 16 ? 1 + (v - 8) / 4 : 0);
    if (fold) *fold = v < 8 ? 2 : 4;
    if (KC) *KC = MX_KC;
    return 1;
  }
  *TB = kVarTB[v];
  *TM = kVarTM[v];
  *PF = kVarPF[v];
  if (fold) *fold = kVarFold[v];
  if (KC) *KC = kVarKC[v];
  return 1;
}

// ---- matrix-core ceiling ----
// Register-only loops of the two MFMA instructions the passes use (8 independent
// accumulators per wave, 8 work-groups of 4 waves per CU): the rate the matrix cores
// sustain on this device, the ceiling the passes are measured against beside the
// published peak.
__global__ __launch_bounds__(256) void mfma_f64_probe_kernel(double* out, int iters) {
  d4 acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) acc[i] = (d4){0.0, 0.0, 0.0, 0.0};
  const double a = 1e-3 * threadIdx.x, b = 1e-3 * blockIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  }
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) sum += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = sum;
}

__global__ __launch_bounds__(256) void mfma_bf16_probe_kernel(double* out, int iters) {
  f4 acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) acc[i] = (f4){0.f, 0.f, 0.f, 0.f};
  bf8 a, b;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = (__bf16)(1e-3f * threadIdx.x);
    b[i] = (__bf16)(1e-3f * blockIdx.x);
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
  }
  float sum = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) sum += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = sum;
}

hipError_t launch_mfma_probe(bool bf16, int iters, int blocks, double* out, hipStream_t st,
                             double* flops) {
  if (bf16) hipLaunchKernelGGL(mfma_bf16_probe_kernel, dim3(blocks), dim3(256), 0, st, out, iters);
  else hipLaunchKernelGGL(mfma_f64_probe_kernel, dim3(blocks), dim3(256), 0, st, out, iters);
  // 4 waves per block, 8 instructions per iteration: 16x16x4 f64 = 2048 flops,
  // 16x16x32 bf16 = 16384 flops
  *flops = (double)blocks * 4.0 * iters * 8.0 * (bf16 ? 16384.0 : 2048.0);
  return hipGetLastError();
}

// ---- sine tables ----
static inline double sine_entry(long long num, long long M) {
  // 2 sin(pi num / M) with exact integer reduction of the argument modulo 2M
  long long a = num % (2 * M);
  return 2.0 * sin(M_PI * (double)a / (double)M);
}

size_t sine_table_odd_count(int N) {
  const size_t H = (size_t)(N + 1) / 2;
  return H * H;
}
size_t sine_table_even_count(int N) {
  const size_t H = (size_t)(N + 1) / 2, E = ((N + 1) & 1) ? H : H - 1;
  return E * E;
}

void build_sine_tables_f64(int N, double* So, double* Se) {
  const long long M = N + 1, H = M / 2;
  for (long long k = 1; k <= H; ++k)
    for (long long m = 0; m < H; ++m) So[(k - 1) * H + m] = sine_entry((2 * m + 1) * k, M);
  const long long E = (M & 1) ? H : H - 1;
  for (long long k = 1; k <= E; ++k)
    for (long long m = 0; m < E; ++m) Se[(k - 1) * E + m] = sine_entry(2 * (m + 1) * k, M);
}

size_t sine_table_quarter_count(int N) {
  const size_t Qp = quarter_pitch((N + 1) / 4);
  return 3 * Qp * Qp;
}

void build_sine_tables_f64q(int N, double* Sq) {
  const long long M = N + 1, Q = M / 4, Qp = quarter_pitch((int)Q);
  for (long long k = 0; k < 3 * Qp * Qp; ++k) Sq[k] = 0.0;
  double* T1 = Sq;
  double* T2 = T1 + Qp * Qp;
  double* T4 = T2 + Qp * Qp;
  for (long long i = 0; i < Q; ++i)
    for (long long m = 0; m < Q; ++m) {
      T1[i * Qp + m] = sine_entry((2 * m + 1) * (2 * i + 1), M);
      T2[i * Qp + m] = sine_entry((2 * m + 1) * (2 * i + 2), M);
    }
  for (long long i = 0; i < Q - 1; ++i)
    for (long long m = 0; m < Q - 1; ++m) T4[i * Qp + m] = sine_entry(4 * (m + 1) * (i + 1), M);
}

static inline uint16_t bf16_bits_rne(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}
static inline float bf16_to_float(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static size_t bf16_tab(int N) {
  const size_t H = (size_t)(N + 1) / 2;
  const size_t Hp = (H + 31) & ~(size_t)31;
  const size_t Mp = (H + 31) / 32 * 32;
  return Mp * Hp;
}
size_t sine_table_bf16q_count(int N) {
  const size_t Qm = ((size_t)(N + 1) / 4 + 31) & ~(size_t)31;
  return 6 * Qm * Qm;
}

void build_sine_tables_bf16q(int N, uint16_t* T) {
  const long long M = N + 1, Q = M / 4, Qm = (Q + 31) & ~31LL, tq = Qm * Qm;
  memset(T, 0, 6 * tq * 2);
  auto put = [&](int t, long long m, long long i, double v) {
    const uint16_t hi = bf16_bits_rne((float)v);
    const uint16_t lo = bf16_bits_rne((float)(v - (double)bf16_to_float(hi)));
    T[(2 * t) * tq + m * Qm + i] = hi;
    T[(2 * t + 1) * tq + m * Qm + i] = lo;
  };
  for (long long m = 0; m < Q; ++m)
    for (long long i = 0; i < Q; ++i) {
      put(0, m, i, sine_entry((2 * m + 1) * (2 * i + 1), M));  // T1
      put(1, m, i, sine_entry((2 * m + 1) * (2 * i + 2), M));  // T2
    }
  for (long long m = 0; m < Q - 1; ++m)
    for (long long i = 0; i < Q - 1; ++i) put(2, m, i, sine_entry(4 * (m + 1) * (i + 1), M));  // T4
}

size_t sine_table_bf16_odd_count(int N) { return 2 * bf16_tab(N); }
size_t sine_table_bf16_even_count(int N) { return 2 * bf16_tab(N); }

// layout: [hi | lo], each Mp x Hp with k contiguous per output row m (0-based k0 = k-1)
void build_sine_tables_bf16(int N, uint16_t* So_hi, uint16_t* So_lo, uint16_t* Se_hi,
                            uint16_t* Se_lo) {
  const long long M = N + 1, H = M / 2;
  const long long Hp = (H + 31) & ~31LL;
  const size_t tab = bf16_tab(N);
  memset(So_hi, 0, tab * 2);
  memset(So_lo, 0, tab * 2);
  memset(Se_hi, 0, tab * 2);
  memset(Se_lo, 0, tab * 2);
  for (long long m = 0; m < H; ++m)
    for (long long k = 1; k <= H; ++k) {
      const double v = sine_entry((2 * m + 1) * k, M);
      uint16_t hi = bf16_bits_rne((float)v);
      uint16_t lo = bf16_bits_rne((float)(v - (double)bf16_to_float(hi)));
      So_hi[m * Hp + (k - 1)] = hi;
      So_lo[m * Hp + (k - 1)] = lo;
    }
  const long long E = (M & 1) ? H : H - 1;
  for (long long m = 0; m < E; ++m)
    for (long long k = 1; k <= E; ++k) {
      const double v = sine_entry(2 * (m + 1) * k, M);
      uint16_t hi = bf16_bits_rne((float)v);
      uint16_t lo = bf16_bits_rne((float)(v - (double)bf16_to_float(hi)));
      Se_hi[m * Hp + (k - 1)] = hi;
      Se_lo[m * Hp + (k - 1)] = lo;
    }
}

}  // namespace ibpm
