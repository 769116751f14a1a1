// Kernels of the CG warm-start network and of the solver-side guess.
//
// The network is a stack of 1-D convolutions over the points of each body. A convolution
// of width K over rows r (one row per body point, channels contiguous) is
//     Y[r][co] = sum_{k, ci} X[r + k - p][ci] * W[co][k][ci],   p = K / 2,
// where a shifted row only counts when it lies in the same sequence (body) as r. It runs
// as a BF16 matrix product on v_mfma_f32_16x16x32_bf16 with FP32 accumulation: M = rows,
// N = output channels, the contraction walks the K taps and 32-channel chunks, and the
// shifted rows are gathered straight from global memory (the "im2col" operand is never
// materialised).
//
// Backward:
//   data    dX[r][ci] = sum_{k, co} dY[r + k - p][co] * Wt[ci][k][co] with the flipped,
//           transposed weights Wt[ci][k][co] = W[co][K-1-k][ci]: the same kernel.
//   weights dW[co][k][ci] = sum_r dY[r][co] * X[r + k - p][ci]: rows are the contraction,
//           both operands are staged transposed in LDS; chunks of rows reduce in a second
//           pass (deterministic, no atomics).
//
// Lane layouts of v_mfma_f32_16x16x32_bf16 (64 lanes, l = lane):
//   A: lane l holds A[l % 16][8 (l / 16) + i], i = 0..7
//   B: lane l holds B[8 (l / 16) + i][l % 16]
//   D: lane l holds D[4 (l / 16) + j][l % 16], j = 0..3
#include <math.h>
#include <stdint.h>

#include "ibpm_internal.h"

namespace ibpm {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

namespace {

__device__ __forceinline__ bf8 load_bf8(const uint16_t* p) {
  const uint4 u = *(const uint4*)p;
  return __builtin_bit_cast(bf8, u);
}
__device__ __forceinline__ bf8 zero_bf8() { return __builtin_bit_cast(bf8, make_uint4(0, 0, 0, 0)); }
__device__ __forceinline__ uint16_t to_bf16(float x) {
  const __bf16 b = (__bf16)x;  // round to nearest even
  return __builtin_bit_cast(uint16_t, b);
}
__device__ __forceinline__ float from_bf16(uint16_t u) {
  return __uint_as_float(((uint32_t)u) << 16);
}

// ---- convolution (forward and backward-data) ----
// One work-group: 64 rows x 32 output channels; each wave 16 rows x 2 column fragments.
template <int EPI>
__global__ __launch_bounds__(256) void conv_kernel(const ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const int r0 = blockIdx.x * 64 + wave * 16;
  const int co0 = blockIdx.y * 32;
  const int arow = r0 + m;
  const int p = a.K >> 1;
  const bool rowok = arow < a.R;
  const int sr = rowok ? a.seg[arow] : 0;
  f4 acc[2];
  acc[0] = f4{0.f, 0.f, 0.f, 0.f};
  acc[1] = acc[0];
  const size_t wstride = (size_t)a.K * a.Cin;
  for (int k = 0; k < a.K; ++k) {
    const int rs = arow + k - p;
    const bool ok = rowok && rs >= 0 && rs < a.R && a.seg[rs] == sr;
    const uint16_t* xrow = a.X + (size_t)(ok ? rs : 0) * a.Cin + 8 * kg;
    const uint16_t* w0 = a.W + (size_t)(co0 + m) * wstride + (size_t)k * a.Cin + 8 * kg;
    const uint16_t* w1 = w0 + 16 * wstride;
    for (int c0 = 0; c0 < a.Cin; c0 += 32) {
      const bf8 av = ok ? load_bf8(xrow + c0) : zero_bf8();
      const bf8 b0 = load_bf8(w0 + c0);
      const bf8 b1 = load_bf8(w1 + c0);
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b0, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b1, acc[1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int f = 0; f < 2; ++f) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + 4 * kg + j;
      const int co = co0 + 16 * f + m;
      if (r >= a.Rpad) continue;
      const size_t o = (size_t)r * a.Cout + co;
      float v = acc[f][j];
      if (r >= a.R) v = 0.f;
      else if (EPI == CONV_RELU) v = fmaxf(v + a.bias[co], 0.f);
      else if (EPI == CONV_LINEAR_F32) v = v + a.bias[co];
      else if (from_bf16(a.mask[o]) <= 0.f) v = 0.f;
      if (EPI == CONV_LINEAR_F32) ((float*)a.Y)[o] = v;
      else ((uint16_t*)a.Y)[o] = to_bf16(v);
    }
  }
}

// ---- weight gradients ----
// Work-group: a chunk of rows x 32 output channels x 32 input channels, all K taps.
// (co-fragment, ci-fragment, tap) triples are spread over the 4 waves.
constexpr int WG_MAXT = 7;  // ceil(4 * 7 / 4) triples per wave for K <= 7
constexpr int WG_LDA = 40;  // LDS row pitch (bf16) of the staged dY^T
constexpr int WG_LDB = 48;  // LDS row pitch of the staged X^T (32 rows + halo)

__global__ __launch_bounds__(256) void wgrad_kernel(const WgradArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[32][WG_LDA];  // [co][row]
  __shared__ __attribute__((aligned(16))) uint16_t sB[32][WG_LDB];  // [ci][row + halo]
  __shared__ int sSeg[WG_LDB];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int m = lane & 15, kg = lane >> 4;
  const int co0 = blockIdx.y * 32, ci0 = blockIdx.z * 32;
  const int K = a.K, p = K >> 1;
  const int rbeg = blockIdx.x * a.rows_per_chunk;
  const int rend = min(rbeg + a.rows_per_chunk, a.R);
  const int ntrip = 4 * K;
  const bool do_bias = blockIdx.z == 0;
  float bsum = 0.f;
  f4 acc[WG_MAXT];
#pragma unroll
  for (int i = 0; i < WG_MAXT; ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};

  for (int r0 = rbeg; r0 < rend; r0 += 32) {
    // dY rows r0..r0+31 (zero past rend), transposed into sA
    if (t < 128) {
      const int row = t >> 2, g = t & 3;
      const int r = r0 + row;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (r < rend) u = *(const uint4*)(a.dY + (size_t)r * a.Cout + co0 + 8 * g);
      const uint16_t* e = (const uint16_t*)&u;
#pragma unroll
      for (int i = 0; i < 8; ++i) sA[8 * g + i][row] = e[i];
    }
    // X rows r0-p .. r0+31+p (zero outside [0, R)), transposed into sB
    const int nrow = 32 + 2 * p;
    for (int it = t; it < nrow * 4; it += 256) {
      const int row = it >> 2, g = it & 3;
      const int r = r0 - p + row;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (r >= 0 && r < a.R) u = *(const uint4*)(a.X + (size_t)r * a.Cin + ci0 + 8 * g);
      const uint16_t* e = (const uint16_t*)&u;
#pragma unroll
      for (int i = 0; i < 8; ++i) sB[8 * g + i][row] = e[i];
    }
    if (t < nrow) {
      const int r = r0 - p + t;
      sSeg[t] = (r >= 0 && r < a.R) ? a.seg[r] : -0x7fffffff;
    }
    __syncthreads();
    if (do_bias && t < 32) {
      for (int row = 0; row < 32; ++row) bsum += from_bf16(sA[t][row]);
    }
#pragma unroll
    for (int i = 0; i < WG_MAXT; ++i) {
      const int trip = wave + 4 * i;
      if (trip < ntrip) {
        const int cf = trip & 1, cif = (trip >> 1) & 1, k = trip >> 2;
        const bf8 av = load_bf8(&sA[16 * cf + m][8 * kg]);
        bf8 bv;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int kk = 8 * kg + e;
          const bool ok = sSeg[kk + k] == sSeg[kk + p];
          bv[e] = ok ? __builtin_bit_cast(__bf16, sB[16 * cif + m][kk + k]) : (__bf16)0.f;
        }
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const size_t nw = (size_t)a.Cout * K * a.Cin;
  float* part = a.part + (size_t)blockIdx.x * (nw + a.Cout);
#pragma unroll
  for (int i = 0; i < WG_MAXT; ++i) {
    const int trip = wave + 4 * i;
    if (trip < ntrip) {
      const int cf = trip & 1, cif = (trip >> 1) & 1, k = trip >> 2;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int co = co0 + 16 * cf + 4 * kg + j;
        const int ci = ci0 + 16 * cif + m;
        part[((size_t)co * K + k) * a.Cin + ci] = acc[i][j];
      }
    }
  }
  if (do_bias && t < 32) part[nw + co0 + t] = bsum;
}

__global__ void reduce_chunks_kernel(const float* part, int nchunk, size_t stride, float* out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < stride;
       i += (size_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * stride + i];
    out[i] = s;
  }
}

// ---- loss ----
// Mean squared error over rows x 2 outputs; dY = 2 (y - t) / (2R) as bf16 (padded
// channels and rows zero), per-block partial sums of (y - t)^2 in FP64.
__global__ __launch_bounds__(256) void mse_kernel(const float* y, int Cp, const float* tgt, int R,
                                                  int Rpad, uint16_t* dY, double* partial) {
  __shared__ double sh[256];
  const size_t i = blockIdx.x * (size_t)256 + threadIdx.x;
  double e2 = 0.0;
  if (i < (size_t)Rpad * Cp) {
    const int r = (int)(i / Cp), c = (int)(i % Cp);
    float g = 0.f;
    if (r < R && c < 2) {
      const float d = y[i] - tgt[2 * r + c];
      e2 = (double)d * (double)d;
      g = d / (float)R;  // 2 d / (2R)
    }
    dY[i] = to_bf16(g);
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void loss_reduce_kernel(const double* partial, int nb, int R,
                                                          double* out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0] / (2.0 * (double)R);
}

// ---- optimiser ----
__global__ void adam_kernel(float* p, const float* g, float* m, float* v, size_t n, float lr,
                            float b1, float b2, float eps, float bc1, float bc2) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

// bf16 copies of one layer's weights: W[co][k][ci] and Wt[ci][k][co] = W[co][K-1-k][ci]
__global__ void cast_weights_kernel(const float* w, int Cout, int K, int Cin, uint16_t* wb,
                                    uint16_t* wt) {
  const size_t n = (size_t)Cout * K * Cin;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int ci = (int)(i % Cin);
    const int k = (int)((i / Cin) % K);
    const int co = (int)(i / ((size_t)Cin * K));
    const uint16_t b = to_bf16(w[i]);
    wb[i] = b;
    wt[((size_t)ci * K + (K - 1 - k)) * Cout + co] = b;
  }
}

// ---- input rows ----
// Minibatch rows from the training set (fixed stride of max_points rows per sample).
__global__ __launch_bounds__(256) void gather_batch_kernel(const GatherBatchArgs a) {
  const int s = blockIdx.x;
  if (s == a.S) {  // padding rows
    for (int r = a.R + threadIdx.x; r < a.Rpad; r += 256) {
      a.seg[r] = -1;
      for (int c = 0; c < kNetInPad; ++c) a.A0[(size_t)r * kNetInPad + c] = 0;
    }
    return;
  }
  const int idx = a.idx[s];
  const int cnt = a.counts[idx];
  const int rb = a.off[s];
  for (int j = blockIdx.y * 256 + threadIdx.x; j < cnt; j += gridDim.y * 256) {
    const size_t src = (size_t)idx * a.stride + j;
    const int r = rb + j;
    for (int c = 0; c < kNetInPad; ++c)
      a.A0[(size_t)r * kNetInPad + c] = c < kNetFeatures ? to_bf16(a.feat[src * kNetFeatures + c]) : 0;
    a.tgt_out[2 * r] = a.tgt[2 * src] * a.inv_scale;
    a.tgt_out[2 * r + 1] = a.tgt[2 * src + 1] * a.inv_scale;
    a.seg[r] = s * 65536 + a.body[src];
  }
}

// Host-given rows: features (R x kNetFeatures) to bf16 input rows.
__global__ void pack_rows_kernel(const float* x, const int* seg_in, int R, int Rpad, uint16_t* A0,
                                 int* seg) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < Rpad; r += gridDim.x * blockDim.x) {
    for (int c = 0; c < kNetInPad; ++c)
      A0[(size_t)r * kNetInPad + c] =
          (r < R && c < kNetFeatures) ? to_bf16(x[(size_t)r * kNetFeatures + c]) : 0;
    seg[r] = r < R ? seg_in[r] : -1;
  }
}

__global__ void sumsq_kernel(const float* x, size_t n, double* partial) {
  __shared__ double sh[256];
  double s = 0.0;
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    s += (double)x[i] * (double)x[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// ---- solver-side warm start ----
__global__ void ws_extrap_kernel(int n2, const double* F1, const double* F2, const double* F3,
                                 int nhist, double* xe) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += gridDim.x * blockDim.x) {
    double v;
    if (nhist >= 3) v = 3.0 * F1[i] - 3.0 * F2[i] + F3[i];
    else if (nhist == 2) v = 2.0 * F1[i] - F2[i];
    else v = F1[i];
    xe[i] = v;
  }
}

__device__ double block_sum_1024_ws(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Features of one step (single work-group): r_e = b - B x_e, its rms and norm, ||b||.
__global__ __launch_bounds__(1024) void ws_features_kernel(const WsArgs a) {
  __shared__ double sh[1024];
  const int n = a.n;
  double rr = 0.0, bb = 0.0, uu = 0.0;
  for (int i = threadIdx.x; i < 2 * n; i += 1024) {
    const double r = a.b[i] - a.Bxe[i];
    rr += r * r;
    bb += a.b[i] * a.b[i];
    uu += a.ub[i] * a.ub[i];
  }
  rr = block_sum_1024_ws(rr, sh);
  bb = block_sum_1024_ws(bb, sh);
  uu = block_sum_1024_ws(uu, sh);
  const double rms = sqrt(rr / (2.0 * n));
  const double urms = sqrt(uu / (2.0 * n));
  const double ir = rms > 0.0 ? 1.0 / rms : 0.0;
  const double iu = urms > 0.0 ? 1.0 / urms : 0.0;
  const double two_pi_h = 2.0 * M_PI / a.h;
  for (int i = threadIdx.x; i < a.Rpad; i += 1024) {
    float f[kNetFeatures];
    for (int c = 0; c < kNetFeatures; ++c) f[c] = 0.f;
    if (i < n) {
      const double px = a.xb[i], py = a.xb[n + i];
      f[0] = (float)((a.b[i] - a.Bxe[i]) * ir);
      f[1] = (float)((a.b[n + i] - a.Bxe[n + i]) * ir);
      const double ax = two_pi_h * (px - a.x0), ay = two_pi_h * (py - a.y0);
      f[2] = (float)sin(ax);
      f[3] = (float)cos(ax);
      f[4] = (float)sin(ay);
      f[5] = (float)cos(ay);
      const int body = a.body_of[i];
      const int ip = (i > 0 && a.body_of[i - 1] == body) ? i - 1 : i;
      const int in = (i + 1 < n && a.body_of[i + 1] == body) ? i + 1 : i;
      double tx = a.xb[in] - a.xb[ip], ty = a.xb[n + in] - a.xb[n + ip];
      const double tn = sqrt(tx * tx + ty * ty);
      if (tn > 0.0) {
        tx /= tn;
        ty /= tn;
      } else {
        tx = 1.0;
        ty = 0.0;
      }
      f[6] = (float)tx;
      f[7] = (float)ty;
      f[8] = (float)(a.ub[i] * iu);
      f[9] = (float)(a.ub[n + i] * iu);
      for (int c = 0; c < kNetFeatures; ++c) a.feat[(size_t)i * kNetFeatures + c] = f[c];
      a.seg[i] = body;
    } else {
      a.seg[i] = -1;
    }
    for (int c = 0; c < kNetInPad; ++c)
      a.A0[(size_t)i * kNetInPad + c] = c < kNetFeatures ? to_bf16(f[c]) : 0;
  }
  if (threadIdx.x == 0) {
    a.scal[0] = rms;
    a.scal[1] = sqrt(rr);
    a.scal[2] = sqrt(bb);
  }
}

// x0 = x_e + rms(r_e) * out_scale * y
__global__ void ws_apply_kernel(int n, const double* xe, const float* y, int Cp, const double* scal,
                                float out_scale, double* x0) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double s = scal[0] * (double)out_scale;
    x0[i] = xe[i] + s * (double)y[(size_t)i * Cp];
    x0[n + i] = xe[n + i] + s * (double)y[(size_t)i * Cp + 1];
  }
}

// One training sample: features, target (F - x_e) / rms(r_e), body ids; rows past n zero.
__global__ void ws_record_kernel(const WsArgs a, int stride, float* ds_feat, float* ds_tgt,
                                 int* ds_body) {
  const double rms = a.scal[0];
  const double ir = rms > 0.0 ? 1.0 / rms : 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < stride; i += gridDim.x * blockDim.x) {
    const bool ok = i < a.n;
    for (int c = 0; c < kNetFeatures; ++c)
      ds_feat[(size_t)i * kNetFeatures + c] = ok ? a.feat[(size_t)i * kNetFeatures + c] : 0.f;
    ds_tgt[2 * (size_t)i] = ok ? (float)((a.F[i] - a.xe[i]) * ir) : 0.f;
    ds_tgt[2 * (size_t)i + 1] = ok ? (float)((a.F[a.n + i] - a.xe[a.n + i]) * ir) : 0.f;
    ds_body[i] = ok ? a.body_of[i] : -1;
  }
}

}  // namespace

hipError_t launch_conv(const ConvArgs& a, int epi, hipStream_t st) {
  if (g_skip) return no_skip();
  if (a.Cin % 32 || a.Cout % 32 || a.Rpad % 64 || a.K < 1 || a.K > 7 || !(a.K & 1))
    return hipErrorInvalidValue;
  const dim3 grid(a.Rpad / 64, a.Cout / 32);
  if (epi == CONV_RELU) hipLaunchKernelGGL(conv_kernel<CONV_RELU>, grid, dim3(256), 0, st, a);
  else if (epi == CONV_LINEAR_F32)
    hipLaunchKernelGGL(conv_kernel<CONV_LINEAR_F32>, grid, dim3(256), 0, st, a);
  else hipLaunchKernelGGL(conv_kernel<CONV_MASK>, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

int wgrad_chunks(int R, int rows_per_chunk) { return (R + rows_per_chunk - 1) / rows_per_chunk; }

hipError_t launch_wgrad(const WgradArgs& a, float* out, hipStream_t st) {
  if (a.Cin % 32 || a.Cout % 32 || a.rows_per_chunk % 32 || a.K > 7 || !(a.K & 1) || a.R < 1)
    return hipErrorInvalidValue;
  const int nchunk = wgrad_chunks(a.R, a.rows_per_chunk);
  hipLaunchKernelGGL(wgrad_kernel, dim3(nchunk, a.Cout / 32, a.Cin / 32), dim3(256), 0, st, a);
  const size_t stride = (size_t)a.Cout * a.K * a.Cin + a.Cout;
  const int nb = (int)((stride + 255) / 256);
  hipLaunchKernelGGL(reduce_chunks_kernel, dim3(nb < 1024 ? nb : 1024), dim3(256), 0, st, a.part,
                     nchunk, stride, out);
  return hipGetLastError();
}

int mse_blocks(int Rpad, int Cp) { return (int)(((size_t)Rpad * Cp + 255) / 256); }

hipError_t launch_mse(const float* y, int Cp, const float* tgt, int R, int Rpad, uint16_t* dY,
                      double* partial, double* loss_out, hipStream_t st) {
  const int nb = mse_blocks(Rpad, Cp);
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(256), 0, st, y, Cp, tgt, R, Rpad, dY, partial);
  hipLaunchKernelGGL(loss_reduce_kernel, dim3(1), dim3(256), 0, st, partial, nb, R, loss_out);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, size_t n, float lr, float b1,
                       float b2, float eps, int t, hipStream_t st) {
  const float bc1 = 1.f - powf(b1, (float)t), bc2 = 1.f - powf(b2, (float)t);
  const size_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(adam_kernel, dim3(nb < 2048 ? (unsigned)nb : 2048u), dim3(256), 0, st, p, g,
                     m, v, n, lr, b1, b2, eps, bc1, bc2);
  return hipGetLastError();
}

hipError_t launch_cast_weights(const float* w, int Cout, int K, int Cin, uint16_t* wb,
                               uint16_t* wt, hipStream_t st) {
  const size_t n = (size_t)Cout * K * Cin;
  const size_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(cast_weights_kernel, dim3(nb < 2048 ? (unsigned)nb : 2048u), dim3(256), 0, st,
                     w, Cout, K, Cin, wb, wt);
  return hipGetLastError();
}

hipError_t launch_gather_batch(const GatherBatchArgs& a, int max_points, hipStream_t st) {
  const int gy = (max_points + 255) / 256;
  hipLaunchKernelGGL(gather_batch_kernel, dim3(a.S + 1, gy), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_pack_rows(const float* x, const int* seg_in, int R, int Rpad, uint16_t* A0,
                            int* seg, hipStream_t st) {
  const int nb = (Rpad + 255) / 256;
  hipLaunchKernelGGL(pack_rows_kernel, dim3(nb), dim3(256), 0, st, x, seg_in, R, Rpad, A0, seg);
  return hipGetLastError();
}

hipError_t launch_sumsq(const float* x, size_t n, double* partial, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(sumsq_kernel, dim3(nblocks), dim3(256), 0, st, x, n, partial);
  return hipGetLastError();
}

hipError_t launch_ws_extrap(int n2, const double* F1, const double* F2, const double* F3, int nhist,
                            double* xe, hipStream_t st) {
  if (g_skip) return no_skip();
  hipLaunchKernelGGL(ws_extrap_kernel, dim3((n2 + 255) / 256), dim3(256), 0, st, n2, F1, F2, F3,
                     nhist, xe);
  return hipGetLastError();
}

hipError_t launch_ws_features(const WsArgs& a, hipStream_t st) {
  if (g_skip) return no_skip();
  hipLaunchKernelGGL(ws_features_kernel, dim3(1), dim3(1024), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_ws_apply(int n, const double* xe, const float* y, int Cp, const double* scal,
                           float out_scale, double* x0, hipStream_t st) {
  if (g_skip) return no_skip();
  hipLaunchKernelGGL(ws_apply_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, xe, y, Cp, scal,
                     out_scale, x0);
  return hipGetLastError();
}

hipError_t launch_ws_record(const WsArgs& a, int stride, float* ds_feat, float* ds_tgt, int* ds_body,
                            hipStream_t st) {
  if (g_skip) return no_skip();
  hipLaunchKernelGGL(ws_record_kernel, dim3((stride + 255) / 256), dim3(256), 0, st, a, stride,
                     ds_feat, ds_tgt, ds_body);
  return hipGetLastError();
}

}  // namespace ibpm
