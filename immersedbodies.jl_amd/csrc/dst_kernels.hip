// Sine-transform passes of the DST-I Poisson / Helmholtz inverses as dense sine-matrix
// contractions on the CDNA4 matrix cores.
//
// Reference: the 2-D RODFT00 of FFTW used by _lap_inv_operator (fluid.jl:113-115,
// :145-166). A 2-D transform is two passes; each pass contracts the contiguous index
// of its input with the sine matrix and writes the result transposed, so the second
// pass again contracts a contiguous index and the plane returns to its original layout.
//
// Even/odd folding. With M = N+1 (even) and H = M/2, row j of S (S_jk = 2 sin(pi j k/M))
// satisfies S_{j,M-k} = (-1)^{j+1} S_{j,k}. Folding the input as
//   P_k = x_k + x_{M-k} (k < H), P_H = x_H        -> odd rows  j = 2m+1, m = 0..H-1
//   Q_k = x_k - x_{M-k} (k < H)                    -> even rows j = 2m,   m = 1..H-1
// halves the contraction: odd rows use So[m][k] = 2 sin(pi (2m+1) k / M) (H x H) and
// even rows use Se[m][k] = 2 sin(pi m k / H) ((H-1) x (H-1)).
//
// The fold and the transpose are staged through LDS; the products run on
//   FP64 path : v_mfma_f64_16x16x4_f64
//   mixed path: v_mfma_f32_16x16x32_bf16, three products hi*hi + hi*lo + lo*hi of the
//               bf16 split of fp32 operands, fp32 accumulation.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

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

// Input element (k0, b) of the first pass, with the apply_bc! edge terms (fac = 1,
// fluid.jl:308-338) when a ring plane is given.
template <typename InT, bool RING>
__device__ __forceinline__ double load_in(const PassPlane& d, int k0, int b) {
  const InT* in = (const InT*)d.in;
  double v = ld(in, (size_t)k0 + (size_t)b * d.Nk);
  if (RING) {
    const int I = k0 + 1, J = b + 1;
    const int Nx = d.Nk, Ny = d.Nb;
    if (J == 1) v += ring_value(d.ring, d.nx, d.ny, 2, I + 1);
    if (J == Ny) v += ring_value(d.ring, d.nx, d.ny, 3, I + 1);
    if (I == 1) v += ring_value(d.ring, d.nx, d.ny, 0, J + 1);
    if (I == Nx) v += ring_value(d.ring, d.nx, d.ny, 1, J + 1);
  }
  return v;
}

template <typename OutT, int EPI>
__device__ __forceinline__ void store_out(const PassPlane& d, size_t idx, double val) {
  OutT* out = (OutT*)d.out;
  if (EPI == EPI_DIAG) val *= d.diag[idx];
  if (EPI == EPI_SUB) val = d.base[idx] - val;
  if (EPI == EPI_ADD) val = d.base[idx] + val;
  out[idx] = (OutT)val;
}

// ---------------------------------------------------------------------------------
// FP64 pass. Work-group tile: 64 values of b x 32 odd rows + 32 even rows; K chunk 16.
// 4 waves: wave w computes parity (w>>1) on b-half (w&1), a 32 x 32 block as 2 x 2
// MFMA 16x16x4 tiles.
// ---------------------------------------------------------------------------------
constexpr int F64_TB = 64;
constexpr int F64_TM = 32;
constexpr int F64_KC = 16;

template <typename InT, typename OutT, int EPI, bool RING>
__global__ __launch_bounds__(256) void dst_pass_f64_kernel(const PassBatch batch) {
  const PassPlane& d = batch.p[blockIdx.z];
  const int Nk = d.Nk, Nb = d.Nb;
  const int H = (Nk + 1) / 2;
  const int b0 = blockIdx.x * F64_TB;
  const int m0 = blockIdx.y * F64_TM;
  if (b0 >= Nb || m0 >= H) return;

  __shared__ double Ps[F64_KC][F64_TB + 1];
  __shared__ double Qs[F64_KC][F64_TB + 1];
  __shared__ double So_s[F64_KC][F64_TM + 1];
  __shared__ double Se_s[F64_KC][F64_TM + 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int parity = wave >> 1;  // 0: odd rows, 1: even rows
  const int bh = wave & 1;
  const double* SoT = (const double*)d.So;
  const double* SeT = (const double*)d.Se;

  d4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

  for (int kc = 0; kc < H; kc += F64_KC) {
    // fold-and-transpose of the input chunk into LDS
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + 256 * r;
      const int kk = e & 15;
      const int bb = e >> 4;
      const int k1 = kc + kk + 1;  // 1-based k
      const int b = b0 + bb;
      double p = 0.0, qv = 0.0;
      if (k1 <= H && b < Nb) {
        const double va = load_in<InT, RING>(d, k1 - 1, b);
        if (k1 == H) {
          p = va;
        } else {
          const double vm = load_in<InT, RING>(d, Nk - k1, b);
          p = va + vm;
          qv = va - vm;
        }
      }
      Ps[kk][bb] = p;
      Qs[kk][bb] = qv;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int e = tid + 256 * r;
      const int mm = e & 31;
      const int kk = e >> 5;
      const int k1 = kc + kk + 1;
      const int m = m0 + mm;
      So_s[kk][mm] = (k1 <= H && m < H) ? SoT[(size_t)(k1 - 1) * H + m] : 0.0;
      Se_s[kk][mm] = (k1 <= H - 1 && m < H - 1) ? SeT[(size_t)(k1 - 1) * (H - 1) + m] : 0.0;
    }
    __syncthreads();
    const int col = lane & 15;
    const int kq = lane >> 4;
#pragma unroll
    for (int ks = 0; ks < F64_KC / 4; ++ks) {
      const int krow = ks * 4 + kq;
      double a0, a1, bv0, bv1;
      if (parity == 0) {
        a0 = Ps[krow][bh * 32 + col];
        a1 = Ps[krow][bh * 32 + 16 + col];
        bv0 = So_s[krow][col];
        bv1 = So_s[krow][16 + col];
      } else {
        a0 = Qs[krow][bh * 32 + col];
        a1 = Qs[krow][bh * 32 + 16 + col];
        bv0 = Se_s[krow][col];
        bv1 = Se_s[krow][16 + col];
      }
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }

  // epilogue: the f64 16x16 accumulator holds D[row = lane/16 + 4r][col = lane%16]
  const int col = lane & 15;
  const int rq = lane >> 4;
  const int mlim = parity == 0 ? H : H - 1;
#pragma unroll
  for (int bs = 0; bs < 2; ++bs)
#pragma unroll
    for (int ms = 0; ms < 2; ++ms) {
      const int m = m0 + ms * 16 + col;
      if (m >= mlim) continue;
      const int jc = parity == 0 ? 2 * m : 2 * m + 1;  // 0-based output row index j-1
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = b0 + bh * 32 + bs * 16 + rq + 4 * r;
        if (b < Nb) store_out<OutT, EPI>(d, (size_t)b + (size_t)jc * Nb, acc[bs][ms][r]);
      }
    }
}

// ---------------------------------------------------------------------------------
// Mixed pass: fp32 operands split into bf16 hi + lo, three MFMA products, fp32 sums.
// Work-group tile: 64 b x (32 odd + 32 even) rows, K chunk 32 (one MFMA k-step).
// The sine tables are stored per output row with k contiguous and zero padding to a
// multiple of 32 in both k and m, so the B operand is one 16-byte load per lane.
// ---------------------------------------------------------------------------------
constexpr int MX_TB = 64;
constexpr int MX_TM = 32;
constexpr int MX_KC = 32;
constexpr int MX_LDK = MX_KC + 8;  // padded LDS row (bf16 elements)

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <typename InT, typename OutT, int EPI, bool RING>
__global__ __launch_bounds__(256) void dst_pass_mixed_kernel(const PassBatch batch) {
  const PassPlane& d = batch.p[blockIdx.z];
  const int Nk = d.Nk, Nb = d.Nb;
  const int H = (Nk + 1) / 2;
  const int Hp = (H + 31) & ~31;
  const int b0 = blockIdx.x * MX_TB;
  const int m0 = blockIdx.y * MX_TM;
  if (b0 >= Nb || m0 >= H) return;

  __shared__ __attribute__((aligned(16))) __bf16 Ph[MX_TB][MX_LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Pl[MX_TB][MX_LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Qh[MX_TB][MX_LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Ql[MX_TB][MX_LDK];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int parity = wave >> 1;
  const int bh = wave & 1;
  const int col = lane & 15;
  const int kq = lane >> 4;

  const __bf16* Sh = (const __bf16*)(parity == 0 ? d.So : d.Se);
  // hi and lo tables are stored back to back: [hi (Mp x Hp)][lo (Mp x Hp)]
  const int Mp = (H + MX_TM - 1) / MX_TM * MX_TM;
  const size_t tab = (size_t)Mp * Hp;
  const __bf16* Sl = Sh + tab;

  f4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};

  // loader mapping: thread -> row bb = tid/4, k sub-block ks8 = (tid%4)*8 .. +7
  const int lb = tid >> 2;
  const int lk = (tid & 3) * 8;

  for (int kc = 0; kc < H; kc += MX_KC) {
    {
      const int b = b0 + lb;
      __bf16 ph[8], pl[8], qh[8], ql[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k1 = kc + lk + i + 1;
        double p = 0.0, qv = 0.0;
        if (k1 <= H && b < Nb) {
          const double va = load_in<InT, RING>(d, k1 - 1, b);
          if (k1 == H) {
            p = va;
          } else {
            const double vm = load_in<InT, RING>(d, Nk - k1, b);
            p = va + vm;
            qv = va - vm;
          }
        }
        split_bf16((float)p, ph[i], pl[i]);
        split_bf16((float)qv, qh[i], ql[i]);
      }
      *(bf8*)&Ph[lb][lk] = *(bf8*)ph;
      *(bf8*)&Pl[lb][lk] = *(bf8*)pl;
      *(bf8*)&Qh[lb][lk] = *(bf8*)qh;
      *(bf8*)&Ql[lb][lk] = *(bf8*)ql;
    }
    __syncthreads();
    bf8 ah[2], al[2], bhv[2], blv[2];
#pragma unroll
    for (int bs = 0; bs < 2; ++bs) {
      const int row = bh * 32 + bs * 16 + col;
      if (parity == 0) {
        ah[bs] = *(const bf8*)&Ph[row][kq * 8];
        al[bs] = *(const bf8*)&Pl[row][kq * 8];
      } else {
        ah[bs] = *(const bf8*)&Qh[row][kq * 8];
        al[bs] = *(const bf8*)&Ql[row][kq * 8];
      }
    }
#pragma unroll
    for (int ms = 0; ms < 2; ++ms) {
      const int m = m0 + ms * 16 + col;
      const size_t off = (size_t)m * Hp + kc + kq * 8;
      bhv[ms] = *(const bf8*)(Sh + off);
      blv[ms] = *(const bf8*)(Sl + off);
    }
#pragma unroll
    for (int bs = 0; bs < 2; ++bs)
#pragma unroll
      for (int ms = 0; ms < 2; ++ms) {
        acc[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bs], bhv[ms], acc[bs][ms], 0, 0, 0);
        acc[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bs], blv[ms], acc[bs][ms], 0, 0, 0);
        acc[bs][ms] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bs], bhv[ms], acc[bs][ms], 0, 0, 0);
      }
    __syncthreads();
  }

  // the fp32 16x16 accumulator of the bf16 MFMA holds D[row = 4*(lane/16) + r][col = lane%16]
  const int rq = lane >> 4;
  const int mlim = parity == 0 ? H : H - 1;
#pragma unroll
  for (int bs = 0; bs < 2; ++bs)
#pragma unroll
    for (int ms = 0; ms < 2; ++ms) {
      const int m = m0 + ms * 16 + col;
      if (m >= mlim) continue;
      const int jc = parity == 0 ? 2 * m : 2 * m + 1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = b0 + bh * 32 + bs * 16 + 4 * rq + r;
        if (b < Nb) store_out<OutT, EPI>(d, (size_t)b + (size_t)jc * Nb, (double)acc[bs][ms][r]);
      }
    }
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
#define IBPM_DISPATCH_EPI(KERNEL, InT, OutT, RING)                                         \
  switch (epi) {                                                                           \
    case EPI_STORE: hipLaunchKernelGGL((KERNEL<InT, OutT, EPI_STORE, RING>), grid, block, 0, st, batch); break; \
    case EPI_DIAG: hipLaunchKernelGGL((KERNEL<InT, OutT, EPI_DIAG, RING>), grid, block, 0, st, batch); break;   \
    case EPI_SUB: hipLaunchKernelGGL((KERNEL<InT, OutT, EPI_SUB, RING>), grid, block, 0, st, batch); break;     \
    default: hipLaunchKernelGGL((KERNEL<InT, OutT, EPI_ADD, RING>), grid, block, 0, st, batch); break;          \
  }

hipError_t launch_dst_pass_f64(const PassBatch& batch, int nplanes, int maxNb, int maxH, int epi,
                               bool ring, hipStream_t st) {
  dim3 grid((maxNb + F64_TB - 1) / F64_TB, (maxH + F64_TM - 1) / F64_TM, nplanes);
  dim3 block(256);
  if (ring) {
    IBPM_DISPATCH_EPI(dst_pass_f64_kernel, double, double, true)
  } else {
    IBPM_DISPATCH_EPI(dst_pass_f64_kernel, double, double, false)
  }
  return hipGetLastError();
}

hipError_t launch_dst_pass_mixed(const PassBatch& batch, int nplanes, int maxNb, int maxH, int epi,
                                 bool ring, bool in_f32, bool out_f32, hipStream_t st) {
  dim3 grid((maxNb + MX_TB - 1) / MX_TB, (maxH + MX_TM - 1) / MX_TM, nplanes);
  dim3 block(256);
  if (!in_f32 && out_f32) {
    if (ring) {
      IBPM_DISPATCH_EPI(dst_pass_mixed_kernel, double, float, true)
    } else {
      IBPM_DISPATCH_EPI(dst_pass_mixed_kernel, double, float, false)
    }
  } else if (in_f32 && out_f32) {
    IBPM_DISPATCH_EPI(dst_pass_mixed_kernel, float, float, false)
  } else if (in_f32 && !out_f32) {
    IBPM_DISPATCH_EPI(dst_pass_mixed_kernel, float, double, false)
  } else {
    if (ring) {
      IBPM_DISPATCH_EPI(dst_pass_mixed_kernel, double, double, true)
    } else {
      IBPM_DISPATCH_EPI(dst_pass_mixed_kernel, double, double, false)
    }
  }
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
  const size_t H = (size_t)(N + 1) / 2;
  return (H - 1) * (H - 1);
}

void build_sine_tables_f64(int N, double* So, double* Se) {
  const long long M = N + 1, H = M / 2;
  for (long long k = 1; k <= H; ++k)
    for (long long m = 0; m < H; ++m) So[(k - 1) * H + m] = sine_entry((2 * m + 1) * k, M);
  for (long long k = 1; k <= H - 1; ++k)
    for (long long m = 0; m < H - 1; ++m) Se[(k - 1) * (H - 1) + m] = sine_entry(2 * (m + 1) * k, M);
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
  const size_t Mp = (H + MX_TM - 1) / MX_TM * MX_TM;
  return Mp * Hp;
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
  for (long long m = 0; m < H - 1; ++m)
    for (long long k = 1; k <= H - 1; ++k) {
      const double v = sine_entry(2 * (m + 1) * k, M);
      uint16_t hi = bf16_bits_rne((float)v);
      uint16_t lo = bf16_bits_rne((float)(v - (double)bf16_to_float(hi)));
      Se_hi[m * Hp + (k - 1)] = hi;
      Se_lo[m * Hp + (k - 1)] = lo;
    }
}

}  // namespace ibpm
