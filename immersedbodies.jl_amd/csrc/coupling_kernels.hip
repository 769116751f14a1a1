// Fluid-body coupling kernels: regularized-delta weights, E (interpolation onto panels),
// rot(E^T f) (regularisation onto the fluid, scattered without atomics by a gather over
// panels sorted by column), the dense B^-1 product with the force output, and the
// conjugate-gradient reductions kept in FP64.
#include <hip/hip_runtime.h>
#include <limits.h>

#include <algorithm>
#include <vector>

#include "ibpm_device.h"
#include "ibpm_internal.h"

// No fused multiply-add in this file: the flux-point coordinates x0 + h*k feed the branch
// tests of delta_h (r/h <= 1, 2, 3). Panels that sit exactly on grid lines put r/h on a
// branch point, and the reference rounds the product and the sum separately; fusing them
// can move r/h across the branch and change a weight in its 10th digit.
#pragma clang fp contract(off)

namespace ibpm {

// delta_h of coupling.jl:91-148: the Yang3 smoothed 3-point kernel times h. The
// polynomial coefficients are the reference's literal constants (they define the
// kernel); asin and sqrt are the device library's.
__device__ double delta_h(double rf, double rb, double dr) {
  const double r = fabs(rf - rb);
  const double r1 = r / dr;
  const double r2 = r1 * r1;
  const double r3 = r2 * r1;
  const double r4 = r3 * r1;
  if (r1 <= 1.0) {
    const double a5 = asin((1.0 / 2.0) * sqrt(3.0) * (2.0 * r1 - 1.0));
    const double a8 = sqrt(1.0 - 12.0 * r2 + 12.0 * r1);
    return 4.166666667e-2 * r4 + (-0.1388888889 + 3.472222222e-2 * a8) * r3 +
           (-7.121664902e-2 - 5.208333333e-2 * a8 + 0.2405626122 * a5) * r2 +
           (-0.2405626122 * a5 - 0.3792313933 + 0.1012731481 * a8) * r1 +
           8.0187537413e-2 * a5 - 4.195601852e-2 * a8 + 0.6485698427;
  } else if (r1 <= 2.0) {
    const double a6 = asin((1.0 / 2.0) * sqrt(3.0) * (-3.0 + 2.0 * r1));
    const double a9 = sqrt(-23.0 + 36.0 * r1 - 12.0 * r2);
    return -6.250000000e-2 * r4 + (0.4861111111 - 1.736111111e-2 * a9) * r3 +
           (-1.143175026 + 7.812500000e-2 * a9 - 0.1202813061 * a6) * r2 +
           (0.8751991178 + 0.3608439183 * a6 - 0.1548032407 * a9) * r1 - 0.2806563809 * a6 +
           8.22848104e-3 + 0.1150173611 * a9;
  } else if (r1 <= 3.0) {
    const double a1 = asin((1.0 / 2.0 * (2.0 * r1 - 5.0)) * sqrt(3.0));
    const double a7 = sqrt(-71.0 - 12.0 * r2 + 60.0 * r1);
    return 2.083333333e-2 * r4 + (3.472222222e-3 * a7 - 0.2638888889) * r3 +
           (1.214391675 - 2.604166667e-2 * a7 + 2.405626122e-2 * a1) * r2 +
           (-0.1202813061 * a1 - 2.449273192 + 7.262731481e-2 * a7) * r1 + 0.1523563211 * a1 +
           1.843201677 - 7.306134259e-2 * a7;
  }
  return 0.0;
}

// ---------------------------------------------------------------------------------
// Panel kinematics and weights. For moving bodies (bodies.jl:12-31):
//   xb = r0 + R(theta) rb,  ub = v0 + Omega R'(theta) rb,  R' = [-s -c; c -s]
// then _update!(reg, ...) (coupling.jl:27-54): nearest flux cell
//   (ix, iy) = round((xb - x0)/h + 0.75), 3 < ix <= nx-3, 3 < iy <= ny-3
// and the 7 x 7 weights at the u and v flux points around it.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void body_update_kernel(const BodyDev b, int only_moving) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= b.npanel) return;
  const int n = b.npanel;
  const int body = b.body_of[p];
  const bool mv = b.moving[body] != 0;
  if (only_moving && !mv) return;
  double px, py;
  if (mv) {
    const double* m = b.motion + 6 * body;
    const double r0x = m[0], r0y = m[1], v0x = m[2], v0y = m[3], ang = m[4], om = m[5];
    const double c = cos(ang), s = sin(ang);
    const double rx = b.xb_ref[p], ry = b.xb_ref[n + p];
    px = r0x + (c * rx + -s * ry);
    py = r0y + (s * rx + c * ry);
    const double Rv11 = om * -s, Rv12 = om * -c, Rv21 = om * c, Rv22 = om * -s;
    b.xb[p] = px;
    b.xb[n + p] = py;
    b.ub[p] = v0x + (Rv11 * rx + Rv12 * ry);
    b.ub[n + p] = v0y + (Rv21 * rx + Rv22 * ry);
  } else {
    px = b.xb[p];
    py = b.xb[n + p];
  }
  const double h = b.h;
  const int ix = (int)rint((px - b.x0) / h + 1.0 - 0.25);
  const int iy = (int)rint((py - b.y0) / h + 1.0 - 0.25);
  const bool ok = (3 < ix && ix <= b.nx - 3) && (3 < iy && iy <= b.ny - 3);
  if (!ok) {
    atomicMax(b.err, n - p);  // records the smallest offending panel index
    b.coords[2 * p] = 4;      // keep later gathers inside the plane
    b.coords[2 * p + 1] = 4;
    return;
  }
  b.coords[2 * p] = ix;
  b.coords[2 * p + 1] = iy;
  double dxu[7], dyu[7], dxv[7], dyv[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const double X = b.x0 + h * (double)(ix + k - 4);
    const double Y = b.y0 + h * (double)(iy + k - 4);
    dxu[k] = delta_h(X, px, h);
    dyu[k] = delta_h(Y + h / 2, py, h);
    dxv[k] = delta_h(X + h / 2, px, h);
    dyv[k] = delta_h(Y, py, h);
  }
  double* w = b.w + (size_t)p * 98;
#pragma unroll
  for (int bb = 0; bb < 7; ++bb)
#pragma unroll
    for (int a = 0; a < 7; ++a) {
      w[bb * 7 + a] = dxu[a] * dyu[bb];
      w[49 + bb * 7 + a] = dxv[a] * dyv[bb];
    }
}

hipError_t launch_body_update(const BodyDev& b, int only_moving, hipStream_t st) {
  if (g_skip) return no_skip();
  if (b.npanel == 0) return hipSuccess;
  hipLaunchKernelGGL(body_update_kernel, dim3((b.npanel + 255) / 256), dim3(256), 0, st, b,
                     only_moving);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Sort panels by flux column ix (bitonic sort of (ix, p) keys in LDS, one work-group)
// and build xstart[c] = number of panels with ix < c, c = 0 .. nx+2. rot(E^T f) then
// gathers, for each vertex, only the panels of the 8 columns that can reach it.
// ---------------------------------------------------------------------------------
constexpr int SORT_MAX = 8192;

__global__ __launch_bounds__(1024) void body_sort_kernel(const BodyDev b) {
  __shared__ unsigned keys[SORT_MAX];
  const int n = b.npanel;
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  for (int i = threadIdx.x; i < np2; i += 1024)
    keys[i] = i < n ? ((unsigned)b.coords[2 * i] << 13) | (unsigned)i : 0xffffffffu;
  __syncthreads();
  for (int k = 2; k <= np2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < np2; i += 1024) {
        const int l = i ^ j;
        if (l > i) {
          const unsigned a = keys[i], c = keys[l];
          const bool up = (i & k) == 0;
          if ((a > c) == up) {
            keys[i] = c;
            keys[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < n; i += 1024) b.perm[i] = (int)(keys[i] & 8191u);
  for (int c = threadIdx.x; c <= b.nx + 2; c += 1024) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((int)(keys[mid] >> 13) < c) lo = mid + 1;
      else hi = mid;
    }
    b.xstart[c] = lo;
  }
  __syncthreads();
  // rows that panels can reach from vertex column I: a panel at (ix, iy) feeds the rot of
  // vertex rows J in [iy - 4, iy + 3] of columns I in [ix - 4, ix + 3] (scatter_rot_kernel)
  for (int I = threadIdx.x; I <= b.nx; I += 1024) {
    const int c0 = max(I - 3, 0), c1 = min(I + 5, b.nx + 2);
    int lo = 1 << 30, hi = -(1 << 30);
    for (int s = b.xstart[c0]; s < b.xstart[c1]; ++s) {
      const int iy = b.coords[2 * (int)(keys[s] & 8191u) + 1];
      lo = min(lo, iy - 4);
      hi = max(hi, iy + 3);
    }
    b.ywin[2 * I] = lo;
    b.ywin[2 * I + 1] = hi;
  }
}

// Range of the panels' flux cells: out = (min ix, max ix, min iy, max iy), one work-group.
__global__ __launch_bounds__(256) void body_bbox_kernel(const int* __restrict__ coords, int n,
                                                        double* __restrict__ out) {
  __shared__ int red[4][256];
  int v[4] = {1 << 30, -(1 << 30), 1 << 30, -(1 << 30)};
  for (int p = threadIdx.x; p < n; p += 256) {
    const int ix = coords[2 * p], iy = coords[2 * p + 1];
    v[0] = min(v[0], ix);
    v[1] = max(v[1], ix);
    v[2] = min(v[2], iy);
    v[3] = max(v[3], iy);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const int o = threadIdx.x + w;
      red[0][threadIdx.x] = min(red[0][threadIdx.x], red[0][o]);
      red[1][threadIdx.x] = max(red[1][threadIdx.x], red[1][o]);
      red[2][threadIdx.x] = min(red[2][threadIdx.x], red[2][o]);
      red[3][threadIdx.x] = max(red[3][threadIdx.x], red[3][o]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = (double)red[threadIdx.x][0];
}

hipError_t launch_body_bbox(const int* coords, int n, double* out, hipStream_t st) {
  if (g_skip) return no_skip();
  if (n <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(body_bbox_kernel, dim3(1), dim3(256), 0, st, coords, n, out);
  return hipGetLastError();
}

hipError_t launch_body_sort(const BodyDev& b, hipStream_t st) {
  if (g_skip) return no_skip();
  if (b.npanel == 0) return hipSuccess;
  hipLaunchKernelGGL(body_sort_kernel, dim3(1), dim3(1024), 0, st, b);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// E (regT!, coupling.jl:57-69) of the level-1 flux curl(psi1) + q0, minus ub*h
// (couple_surface_rigid, coupling.jl:239-244). One wave per (panel, component, rhs): lane
// l < 49 forms the product of weight (a, b) = (l % 7, l / 7) with the flux there, and the
// 49 products are summed by a fixed shuffle tree (deterministic).
// ---------------------------------------------------------------------------------
template <bool DYN>
__global__ __launch_bounds__(256) void gather_kernel(const GatherArgs a) {
  if (skipped(a.skip)) return;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int n = a.npanel;
  if (t >= 2 * n) return;  // whole waves leave together
  const int rhs = blockIdx.y;
  const int comp = t >= n;
  const int p = comp ? t - n : t;
  const int nx = a.g.nx, ny = a.g.ny;
  const double* p1 = a.psi1[rhs];
  const double* p2 = a.psi2[rhs];
  double s = 0.0;
  if (lane < 49) {
    const int ix = a.coords[2 * p], iy = a.coords[2 * p + 1];
    const int bb = lane / 7, aa = lane % 7;
    const int j = iy + bb - 3, i = ix + aa - 3;
    double qv;
    if (!comp)  // u(i,j) = psi_e(i-1, j) - psi_e(i-1, j-1)
      qv = curl_u(p1, p2, nx, ny, i, j) + base_flux_u<DYN>(a.bf, 0, i, j);
    else  // v(i,j) = psi_e(i-1, j-1) - psi_e(i, j-1)
      qv = curl_v(p1, p2, nx, ny, i, j) + base_flux_v<DYN>(a.bf, 0, i, j);
    s = qv * a.w[(size_t)p * 98 + comp * 49 + lane];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) {
    if (a.ub) s -= a.ub[t] * a.h;
    a.out[rhs][t] = s;
  }
}

hipError_t launch_gather(const GatherArgs& a0, hipStream_t st) {
  if (a0.npanel == 0) return hipSuccess;
  GatherArgs a = a0;
  a.skip = g_skip;
  const dim3 grid((2 * a.npanel + 3) / 4, a.nrhs);
  if (a.bf.dyn)
    hipLaunchKernelGGL(gather_kernel<true>, grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(gather_kernel<false>, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// rot(E^T f) on the level-1 plane (reg! coupling.jl:72-85, then rot! fluid.jl:101-111):
//   Gamma(I,J) = (u(I+1,J) - u(I+1,J+1)) + (v(I+1,J+1) - v(I,J+1)),
//   u(i,j) = sum over panels p with |i-ix_p| <= 3, |j-iy_p| <= 3 of w_u[p] f_x[p].
// Every vertex writes its own value (zero away from the bodies): no atomics.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void scatter_rot_kernel(const ScatterArgs a) {
  if (skipped(a.skip)) return;
  const GridDims g = a.g;
  const int rhs = blockIdx.z;
  const int I = a.ri0 + blockIdx.x * 64 + (threadIdx.x & 63) + 1;
  const int J = a.rj0 + blockIdx.y * 4 + (threadIdx.x >> 6) + 1;
  if (I > a.ri0 + a.rw || J > a.rj0 + a.rh) return;
  const int n = a.npanel;
  const int c0 = max(I - 3, 0);
  const int c1 = min(I + 5, g.nx + 2);  // columns ix in [I-3, I+4]
  // rows out of every panel's reach skip the panel loop (the sum is zero there)
  const bool reach = J >= a.ywin[2 * I] && J <= a.ywin[2 * I + 1];
  const int s0 = a.xstart[c0], s1 = reach ? a.xstart[c1] : s0;
  double ua = 0.0, ub = 0.0, va = 0.0, vb = 0.0;
  const double* f = a.f[rhs];
  // panels in batches of 8: each batch's loads (sort order, cells, forces, weights) are
  // issued together, then the terms are added in panel order as one at a time would
  constexpr int PB = 8;
  for (int sb = s0; sb < s1; sb += PB) {
    int pp[PB];
#pragma unroll
    for (int q = 0; q < PB; ++q) pp[q] = a.perm[sb + q < s1 ? sb + q : s0];
    int ix[PB], iy[PB];
    double fx[PB], fy[PB];
#pragma unroll
    for (int q = 0; q < PB; ++q) {
      ix[q] = a.coords[2 * pp[q]];
      iy[q] = a.coords[2 * pp[q] + 1];
      fx[q] = f[pp[q]];
      fy[q] = f[n + pp[q]];
    }
    double wua[PB], wub[PB], wva[PB], wvb[PB];
    bool tua[PB], tub[PB], tvb[PB];
#pragma unroll
    for (int q = 0; q < PB; ++q) {
      const int db = J - iy[q] + 3;   // b index of row J; row J+1 has db+1
      const int au = I + 1 - ix[q] + 3;  // a index of column I+1
      const int av = I - ix[q] + 3;      // a index of column I
      const bool in = sb + q < s1 && db >= -1 && db <= 6;
      const bool cu = in && au >= 0 && au <= 6;
      tua[q] = cu && db >= 0;
      tub[q] = cu && db + 1 <= 6;  // also the va term
      tvb[q] = in && av >= 0 && av <= 6 && db + 1 <= 6;
      const double* w = a.w + (size_t)pp[q] * 98;
      wua[q] = w[tua[q] ? db * 7 + au : 0];
      wub[q] = w[tub[q] ? (db + 1) * 7 + au : 0];
      wva[q] = w[tub[q] ? 49 + (db + 1) * 7 + au : 0];
      wvb[q] = w[tvb[q] ? 49 + (db + 1) * 7 + av : 0];
    }
#pragma unroll
    for (int q = 0; q < PB; ++q) {
      if (tua[q]) ua += wua[q] * fx[q];
      if (tub[q]) {
        ub += wub[q] * fx[q];
        va += wva[q] * fy[q];
      }
      if (tvb[q]) vb += wvb[q] * fy[q];
    }
  }
  const double v = (ua - ub) + (va - vb);
  const size_t idx = (size_t)(I - 1) + (size_t)(J - 1) * g.Nx;
  if (a.out_f32)
    ((float*)a.out[rhs])[idx] = (float)v;
  else
    ((double*)a.out[rhs])[idx] = v;
}

// The same sums from the vertex lists of static panels: a vertex reads its entry range,
// then its entries (panel, 4 weights), then the panels' forces, and adds the terms in the
// panel order of scatter_rot_kernel (a term a panel does not reach has weight 0, which
// leaves the sum unchanged), so both kernels give the same result.
__device__ __forceinline__ void scatter_rot_lists(const ScatterArgs& a, int bx, int by, int rhs) {
  const GridDims g = a.g;
  const int I = a.ri0 + bx * 64 + (threadIdx.x & 63) + 1;
  const int J = a.rj0 + by * 4 + (threadIdx.x >> 6) + 1;
  if (I > a.ri0 + a.rw || J > a.rj0 + a.rh) return;
  const int n = a.npanel;
  const size_t idx = (size_t)(I - 1) + (size_t)(J - 1) * g.Nx;
  const int e0 = a.rowptr[idx], e1 = a.rowptr[idx + 1];
  const double* f = a.f[rhs];
  const double2* ew = (const double2*)a.ew;
  double ua = 0.0, ub = 0.0, va = 0.0, vb = 0.0;
  constexpr int EB = 8;  // entries per batch, their loads issued together
  for (int eb = e0; eb < e1; eb += EB) {
    int pp[EB];
    double2 wu[EB], wv[EB];
#pragma unroll
    for (int q = 0; q < EB; ++q) {
      const int e = min(eb + q, e1 - 1);
      pp[q] = a.ep[e];
      wu[q] = ew[2 * (size_t)e];
      wv[q] = ew[2 * (size_t)e + 1];
    }
    double fx[EB], fy[EB];
#pragma unroll
    for (int q = 0; q < EB; ++q) {
      fx[q] = f[pp[q]];
      fy[q] = f[n + pp[q]];
    }
#pragma unroll
    for (int q = 0; q < EB; ++q) {
      if (eb + q < e1) {
        ua += wu[q].x * fx[q];
        ub += wu[q].y * fx[q];
        va += wv[q].x * fy[q];
        vb += wv[q].y * fy[q];
      }
    }
  }
  const double v = (ua - ub) + (va - vb);
  if (a.out_f32)
    ((float*)a.out[rhs])[idx] = (float)v;
  else
    ((double*)a.out[rhs])[idx] = v;
}

__global__ __launch_bounds__(256) void scatter_rot_lists_kernel(const ScatterArgs a) {
  if (skipped(a.skip)) return;
  scatter_rot_lists(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

hipError_t launch_scatter_rot(const ScatterArgs& a0, hipStream_t st) {
  ScatterArgs a = a0;
  a.skip = g_skip;
  if (a.rw <= 0) a.ri0 = 0, a.rj0 = 0, a.rw = a.g.Nx, a.rh = a.g.Ny;
  if (a.ri0 < 0 || a.rj0 < 0 || a.ri0 + a.rw > a.g.Nx || a.rj0 + a.rh > a.g.Ny || a.rh <= 0)
    return hipErrorInvalidValue;
  dim3 grid((a.rw + 63) / 64, (a.rh + 3) / 4, a.nrhs);
  if (a.rowptr && a.ep && a.ew)
    hipLaunchKernelGGL(scatter_rot_lists_kernel, grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(scatter_rot_kernel, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

void build_scatter_lists(int n, int Nx, int Ny, const int* coords, const int* perm, const double* w,
                         std::vector<int>& rowptr, std::vector<int>& ep, std::vector<double>& ew) {
  // the terms scatter_rot_kernel adds at vertex (I, J) from the panel at cell (ix, iy):
  // I in [ix - 4, ix + 3], J in [iy - 4, iy + 3], with its index conditions
  const auto visit = [&](auto&& emit) {
    for (int k = 0; k < n; ++k) {
      const int p = perm[k], ix = coords[2 * p], iy = coords[2 * p + 1];
      const double* wp = w + (size_t)p * 98;
      for (int J = std::max(iy - 4, 1); J <= std::min(iy + 3, Ny); ++J)
        for (int I = std::max(ix - 4, 1); I <= std::min(ix + 3, Nx); ++I) {
          const int db = J - iy + 3, au = I + 1 - ix + 3, av = I - ix + 3;
          const bool cu = au >= 0 && au <= 6;
          const bool tua = cu && db >= 0, tub = cu && db + 1 <= 6;
          const bool tvb = av >= 0 && av <= 6 && db + 1 <= 6;
          if (!tua && !tub && !tvb) continue;
          emit((size_t)(I - 1) + (size_t)(J - 1) * Nx, p, tua ? wp[db * 7 + au] : 0.0,
               tub ? wp[(db + 1) * 7 + au] : 0.0, tub ? wp[49 + (db + 1) * 7 + au] : 0.0,
               tvb ? wp[49 + (db + 1) * 7 + av] : 0.0);
        }
    }
  };
  const size_t nv = (size_t)Nx * Ny;
  rowptr.assign(nv + 1, 0);
  visit([&](size_t v, int, double, double, double, double) { ++rowptr[v + 1]; });
  for (size_t v = 0; v < nv; ++v) rowptr[v + 1] += rowptr[v];
  ep.assign(rowptr[nv], 0);
  ew.assign(4 * (size_t)rowptr[nv], 0.0);
  std::vector<int> fill(rowptr.begin(), rowptr.end() - 1);
  visit([&](size_t v, int p, double a, double b, double c, double d) {
    const int e = fill[v]++;
    ep[e] = p;
    ew[4 * (size_t)e] = a;
    ew[4 * (size_t)e + 1] = b;
    ew[4 * (size_t)e + 2] = c;
    ew[4 * (size_t)e + 3] = d;
  });
}

// ---------------------------------------------------------------------------------
// Dense B^-1 product (coupling.jl:191) and update_vars! (problem.jl:676-685):
//   Ft = Binv * b;  fb = Ft * h / dt;  total = column sums of fb (FP64).
// ---------------------------------------------------------------------------------
// Work-group: 64 rows (one per lane, coalesced column-major reads) x one range of
// columns split over its 4 waves; the 4 waves reduce through LDS into a partial per column
// range, and a second kernel sums the ranges in order (deterministic).
constexpr int GEMV_SPLIT = 16;

__global__ __launch_bounds__(256) void gemv_kernel(int m, const double* __restrict__ M,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ part,
                                                   const double* skip) {
  __shared__ double red[4][64];
  if (skipped(skip)) return;
  const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + r;
  const int per = (m + GEMV_SPLIT - 1) / GEMV_SPLIT;
  const int j0 = blockIdx.y * per, j1 = min(m, j0 + per);
  double acc = 0.0;
  if (i < m)
    for (int j = j0 + g; j < j1; j += 4) acc += M[i + (size_t)j * m] * b[j];
  red[g][r] = acc;
  __syncthreads();
  if (g == 0 && i < m)
    part[(size_t)blockIdx.y * m + i] = ((red[0][r] + red[1][r]) + red[2][r]) + red[3][r];
}

__global__ void gemv_sum_kernel(int m, const double* __restrict__ part, double* __restrict__ x,
                                const double* skip) {
  if (skipped(skip)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  double v[GEMV_SPLIT];
#pragma unroll
  for (int k = 0; k < GEMV_SPLIT; ++k) v[k] = part[(size_t)k * m + i];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < GEMV_SPLIT; ++k) s += v[k];
  x[i] = s;
}

// The force sums of one work-group of NT threads (1024 or 256): NT threads play the 1024
// slots of the reduction, so the sums are the same bits for either size.
template <int NT>
__device__ __forceinline__ void forces_sum(int n, const double* __restrict__ Ft,
                                           double* __restrict__ fb, double* __restrict__ total,
                                           double h, double dt, double* __restrict__ host_total) {
  __shared__ double sx[1024], sy[1024];
#pragma unroll
  for (int k = 0; k < 1024 / NT; ++k) {
    const int t = threadIdx.x + NT * k;
    double ax = 0.0, ay = 0.0;
    for (int p = t; p < n; p += 1024) {
      const double fx = Ft[p] * h / dt, fy = Ft[n + p] * h / dt;
      fb[p] = fx;
      fb[n + p] = fy;
      ax += fx;
      ay += fy;
    }
    sx[t] = ax;
    sy[t] = ay;
  }
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    for (int t = threadIdx.x; t < s; t += NT) {
      sx[t] += sx[t + s];
      sy[t] += sy[t + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    total[0] = sx[0];
    total[1] = sy[0];
    if (host_total) {
      host_total[0] = sx[0];
      host_total[1] = sy[0];
    }
  }
}

__global__ __launch_bounds__(1024) void forces_kernel(int n, const double* __restrict__ Ft,
                                                      double* __restrict__ fb,
                                                      double* __restrict__ total, double h,
                                                      double dt, double* __restrict__ host_total) {
  forces_sum<1024>(n, Ft, fb, total, h, dt, host_total);
}

// project_circ's first launch for static bodies (three independent jobs by block range):
// rot(E^T Ft) on the box from the vertex lists, the trial Gamma_1 on that box saved (the
// solve then writes Gamma_1 = saved - correction there, and the final box update reads
// the change against it), and the forces of update_vars!.
__global__ __launch_bounds__(256) void project_prologue_kernel(const ScatterArgs a, int gx, int nsc,
                                                               const double* __restrict__ G,
                                                               double* __restrict__ saved,
                                                               double* __restrict__ fb,
                                                               double* __restrict__ total, double h,
                                                               double dt, double* __restrict__ host_total) {
  const int b = blockIdx.x;
  if (b < nsc) {
    scatter_rot_lists(a, b % gx, b / gx, 0);
  } else if (b == nsc) {
    forces_sum<256>(a.npanel, a.f[0], fb, total, h, dt, host_total);
  } else {  // 1024 values of the box per block, all four loads issued before the stores
    const int n = a.rw * a.rh;
    const int e0 = (b - nsc - 1) * 1024 + threadIdx.x;
    double v[4];
    size_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = min(e0 + 256 * k, n - 1);
      o[k] = (size_t)(a.ri0 + e % a.rw) + (size_t)(a.rj0 + e / a.rw) * a.g.Nx;
      v[k] = G[o[k]];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (e0 + 256 * k < n) saved[o[k]] = v[k];
  }
}

hipError_t launch_project_prologue(const ScatterArgs& a0, const double* G, double* saved,
                                   double* fb, double* total, double h, double dt,
                                   double* host_total, hipStream_t st) {
  if (g_skip) return no_skip();
  ScatterArgs a = a0;
  a.skip = nullptr;
  if (!a.rowptr || a.nrhs != 1 || a.rw <= 0 || a.rh <= 0 || a.ri0 < 0 || a.rj0 < 0 ||
      a.ri0 + a.rw > a.g.Nx || a.rj0 + a.rh > a.g.Ny)
    return hipErrorInvalidValue;
  const int gx = (a.rw + 63) / 64, nsc = gx * ((a.rh + 3) / 4);
  const int ncp = (a.rw * a.rh + 1023) / 1024;
  hipLaunchKernelGGL(project_prologue_kernel, dim3(nsc + 1 + ncp), dim3(256), 0, st, a, gx, nsc, G,
                     saved, fb, total, h, dt, host_total);
  return hipGetLastError();
}

size_t gemv_partial_count(int m) { return (size_t)GEMV_SPLIT * (size_t)(m > 0 ? m : 1); }

hipError_t launch_gemv(int m, const double* M, const double* b, double* x, double* part,
                       hipStream_t st) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(gemv_kernel, dim3((m + 63) / 64, GEMV_SPLIT), dim3(256), 0, st, m, M, b, part, g_skip);
  hipLaunchKernelGGL(gemv_sum_kernel, dim3((m + 255) / 256), dim3(256), 0, st, m, part, x, g_skip);
  return hipGetLastError();
}

// x = M b with M row-major (m x m): one wave per row, lane l summing columns l, l + 64,
// ... in order (all its loads of a 1024-column chunk in flight together), the lane sums
// added by a fixed shuffle tree: one launch, a fixed summation order.
__global__ __launch_bounds__(256) void gemv_rows_kernel(int m, const double* __restrict__ M,
                                                        const double* __restrict__ b,
                                                        double* __restrict__ x) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= m) return;  // whole waves
  const double* row = M + (size_t)i * m;
  double s = 0.0;
  for (int j0 = 0; j0 < m; j0 += 1024) {
    double mv[16], bv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int j = j0 + lane + 64 * q;
      const int jc = j < m ? j : 0;  // clamped: unconditional
      mv[q] = row[jc];
      bv[q] = b[jc];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (j0 + lane + 64 * q < m) s += mv[q] * bv[q];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) x[i] = s;
}

hipError_t launch_gemv_rows(int m, const double* M, const double* b, double* x, hipStream_t st) {
  if (g_skip) return no_skip();
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(gemv_rows_kernel, dim3((m + 3) / 4), dim3(256), 0, st, m, M, b, x);
  return hipGetLastError();
}

hipError_t launch_forces(int n, const double* Ft, double* fb, double* total, double h, double dt,
                         double* host_total, hipStream_t st) {
  if (g_skip) return no_skip();
  hipLaunchKernelGGL(forces_kernel, dim3(1), dim3(1024), 0, st, n, Ft, fb, total, h, dt, host_total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Conjugate gradient of IterativeSolvers.jl's cg! (unpreconditioned CGIterable), as
// called by _Binv_iterative (coupling.jl:194-202), with FP64 reductions in one
// work-group. scal: [0] residual, [1] prev residual, [2] tolerance, [3] iterations,
// [4] converged flag, [5] ||b||.
// ---------------------------------------------------------------------------------
__device__ double block_sum_1024(double v, double* sh) {
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

__global__ __launch_bounds__(1024) void cg_init_kernel(const CgDev c) {
  __shared__ double sh[1024];
  if (skipped(c.skip)) return;
  double rr = 0.0, bb = 0.0;
  for (int i = threadIdx.x; i < c.n; i += 1024) {
    const double r = c.b[i] - c.c[i];
    c.r[i] = r;
    rr += r * r;
    bb += c.b[i] * c.b[i];
  }
  const double res = sqrt(block_sum_1024(rr, sh));
  const double bn = sqrt(block_sum_1024(bb, sh));
  double tol = c.reltol * (c.tol_mode == 1 ? bn : res);
  if (c.abstol > tol) tol = c.abstol;
  const bool conv = (0 >= c.maxiter) || res <= tol;
  // first search direction: u = r + beta*0 with beta = res^2 / 1^2
  for (int i = threadIdx.x; i < c.n; i += 1024) c.u[i] = conv ? 0.0 : c.r[i];
  if (threadIdx.x == 0) {
    c.scal[0] = res;
    c.scal[1] = 1.0;
    c.scal[2] = tol;
    c.scal[3] = 0.0;
    c.scal[4] = conv ? 1.0 : 0.0;
    c.scal[5] = bn;
  }
}

__global__ __launch_bounds__(1024) void cg_step_kernel(const CgDev c) {
  __shared__ double sh[1024];
  if (c.scal[4] != 0.0) return;
  const double res = c.scal[0];
  double uc = 0.0;
  for (int i = threadIdx.x; i < c.n; i += 1024) uc += c.u[i] * c.c[i];
  const double alpha = res * res / block_sum_1024(uc, sh);
  double rr = 0.0;
  for (int i = threadIdx.x; i < c.n; i += 1024) {
    c.x[i] += alpha * c.u[i];
    const double r = c.r[i] - alpha * c.c[i];
    c.r[i] = r;
    rr += r * r;
  }
  const double nres = sqrt(block_sum_1024(rr, sh));
  const double iters = c.scal[3] + 1.0;
  const bool conv = iters >= (double)c.maxiter || nres <= c.scal[2];
  if (!conv) {
    const double beta = nres * nres / (res * res);
    for (int i = threadIdx.x; i < c.n; i += 1024) c.u[i] = c.r[i] + beta * c.u[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    c.scal[1] = res;
    c.scal[0] = nres;
    c.scal[3] = iters;
    c.scal[4] = conv ? 1.0 : 0.0;
  }
}

// ---------------------------------------------------------------------------------
// Preconditioned CG (IterativeSolvers.jl's PCGIterable order: c = Pl \ r, rho = c.r,
// u = c + (rho / rho_prev) u, c = A u, alpha = rho / u.c, x += alpha u, r -= alpha c,
// stop on the true ||r||). The preconditioner is a dense inverse of B assembled at an
// earlier body configuration, applied in each body's frame: z = R P R^T r, with R the
// rotation of every point's (x, y) pair by its body's angle change since the assembly.
// scal additionally holds [6] rho.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ void pc_rotation(const PcDev& p, int i, double& cs, double& sn) {
  const int b = p.body_of[i];
  const double d = p.motion[6 * b + 4] - p.theta0[b];
  cs = cos(d);
  sn = sin(d);
}

// t = R^T r (a point's pair is (i, n + i))
__global__ __launch_bounds__(256) void pcg_rot_in_kernel(const PcDev p, const double* r, double* t) {
  if (skipped(p.skip)) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n) return;
  double cs, sn;
  pc_rotation(p, i, cs, sn);
  const double rx = r[i], ry = r[p.n + i];
  t[i] = cs * rx + sn * ry;
  t[p.n + i] = -sn * rx + cs * ry;
}

// z = R zr, rho = r.z, u = z + (rho / rho_prev) u
__global__ __launch_bounds__(1024) void pcg_dir_kernel(const CgDev c, const PcDev p, const double* zr) {
  __shared__ double sh[1024];
  if (c.scal[4] != 0.0) return;
  const int n = p.n;
  double rz = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    double cs, sn;
    pc_rotation(p, i, cs, sn);
    const double zx = cs * zr[i] - sn * zr[n + i];
    const double zy = sn * zr[i] + cs * zr[n + i];
    rz += c.r[i] * zx + c.r[n + i] * zy;
  }
  const double rho = block_sum_1024(rz, sh);
  const double beta = rho / c.scal[6];
  for (int i = threadIdx.x; i < n; i += 1024) {
    double cs, sn;
    pc_rotation(p, i, cs, sn);
    const double zx = cs * zr[i] - sn * zr[n + i];
    const double zy = sn * zr[i] + cs * zr[n + i];
    c.u[i] = zx + beta * c.u[i];
    c.u[n + i] = zy + beta * c.u[n + i];
  }
  __syncthreads();
  if (threadIdx.x == 0) c.scal[6] = rho;
}

// alpha = rho / u.c, x += alpha u, r -= alpha c, residual and stopping test
__global__ __launch_bounds__(1024) void pcg_step_kernel(const CgDev c) {
  __shared__ double sh[1024];
  if (c.scal[4] != 0.0) return;
  double uc = 0.0;
  for (int i = threadIdx.x; i < c.n; i += 1024) uc += c.u[i] * c.c[i];
  const double alpha = c.scal[6] / block_sum_1024(uc, sh);
  double rr = 0.0;
  for (int i = threadIdx.x; i < c.n; i += 1024) {
    c.x[i] += alpha * c.u[i];
    const double r = c.r[i] - alpha * c.c[i];
    c.r[i] = r;
    rr += r * r;
  }
  const double nres = sqrt(block_sum_1024(rr, sh));
  const double iters = c.scal[3] + 1.0;
  const bool conv = iters >= (double)c.maxiter || nres <= c.scal[2];
  __syncthreads();
  if (threadIdx.x == 0) {
    c.scal[1] = c.scal[0];
    c.scal[0] = nres;
    c.scal[3] = iters;
    c.scal[4] = conv ? 1.0 : 0.0;
  }
}

// the search direction starts at zero with rho = 1 (so the first beta multiplies zero)
__global__ void pcg_start_kernel(const CgDev c) {
  if (skipped(c.skip)) return;
  for (int i = threadIdx.x; i < c.n; i += blockDim.x) c.u[i] = 0.0;
  if (threadIdx.x == 0) c.scal[6] = 1.0;
}

hipError_t launch_pcg_start(const CgDev& c0, hipStream_t st) {
  CgDev c = c0;
  c.skip = g_skip;
  hipLaunchKernelGGL(pcg_start_kernel, dim3(1), dim3(1024), 0, st, c);
  return hipGetLastError();
}

hipError_t launch_pcg_precond_dir(const CgDev& c, const PcDev& p0, double* t, double* zr,
                                  double* part, hipStream_t st) {
  PcDev p = p0;
  p.skip = g_skip;
  hipLaunchKernelGGL(pcg_rot_in_kernel, dim3((p.n + 255) / 256), dim3(256), 0, st, p,
                     (const double*)c.r, t);
  hipError_t e = launch_gemv(2 * p.n, p.P, t, zr, part, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(pcg_dir_kernel, dim3(1), dim3(1024), 0, st, c, p, (const double*)zr);
  return hipGetLastError();
}

hipError_t launch_pcg_step(const CgDev& c, hipStream_t st) {
  hipLaunchKernelGGL(pcg_step_kernel, dim3(1), dim3(1024), 0, st, c);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Dense inverse by Gauss-Jordan elimination with partial pivoting on the augmented
// [A | X], X = I at the start (the role of LAPACK getrf/getri in inv(), coupling.jl:190).
// Column-major m x m. Step k: the pivot is the first row of largest |A[i,k]|, i >= k
// (idamax); rows k and p are swapped in A (columns >= k: the earlier columns are already
// unit vectors) and in X, the pivot row is divided by the pivot into `rowk`, and every
// other row i subtracts A[i,k] times it. At the end A = I and X = A^-1.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void gj_pivot_kernel(int m, const double* A, int k, int* piv,
                                                        double* pval, int* singular) {
  __shared__ double sv[1024];
  __shared__ int si[1024];
  double best = -1.0;
  int bi = m;
  const double* col = A + (size_t)k * m;
  for (int i = k + threadIdx.x; i < m; i += 1024) {
    const double v = fabs(col[i]);
    if (v > best) {  // strided scan in increasing i: keeps the first of equal values
      best = v;
      bi = i;
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double v2 = sv[threadIdx.x + w];
      const int i2 = si[threadIdx.x + w];
      if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
        sv[threadIdx.x] = v2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int p = si[0] < m ? si[0] : k;
    piv[0] = p;
    pval[0] = col[p];
    if (col[p] == 0.0) singular[0] = 1;
  }
}

// swap rows k and piv in A (columns >= k) and X; rowk = the new row k / pivot
__global__ __launch_bounds__(256) void gj_swap_kernel(int m, double* A, double* X, int k,
                                                      const int* piv, const double* pval,
                                                      double* rowk) {
  const int t = blockIdx.x * 256 + threadIdx.x;  // over (m - k) columns of A, then m of X
  const int na = m - k;
  if (t >= na + m) return;
  const int p = piv[0];
  const double d = pval[0];
  double* col = t < na ? A + (size_t)(k + t) * m : X + (size_t)(t - na) * m;
  const double a = col[k], b = col[p];
  col[k] = b;
  col[p] = a;
  rowk[t] = d != 0.0 ? b / d : 0.0;
}

// M[i, j] -= A[i,k] * rowk[j] (i != k), M[k, j] = rowk[j], over A's columns > k and X
__global__ __launch_bounds__(256) void gj_elim_kernel(int m, double* A, double* X, int k,
                                                      const double* rowk) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int t = blockIdx.y * 4 + (threadIdx.x >> 6) + 1;  // skips A's column k (t = 0)
  const int na = m - k;
  if (i >= m || t >= na + m) return;
  double* col = t < na ? A + (size_t)(k + t) * m : X + (size_t)(t - na) * m;
  const double r = rowk[t];
  if (i == k) col[i] = r;
  else col[i] -= A[(size_t)k * m + i] * r;
}

__global__ void gj_identity_kernel(int m, double* X, int* singular) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx == 0) singular[0] = 0;
  if (idx >= (size_t)m * m) return;
  X[idx] = (idx % (m + 1) == 0) ? 1.0 : 0.0;
}

hipError_t launch_gauss_jordan_inverse(int m, double* A, double* X, double* rowk, int* ipiv,
                                       double* pval, int* singular, hipStream_t st) {
  if (g_skip) return no_skip();
  hipLaunchKernelGGL(gj_identity_kernel, dim3((unsigned)(((size_t)m * m + 255) / 256)), dim3(256), 0,
                     st, m, X, singular);
  for (int k = 0; k < m; ++k) {
    hipLaunchKernelGGL(gj_pivot_kernel, dim3(1), dim3(1024), 0, st, m, (const double*)A, k, ipiv,
                       pval, singular);
    const int ncol = (m - k) + m;
    hipLaunchKernelGGL(gj_swap_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, m, A, X, k,
                       (const int*)ipiv, (const double*)pval, rowk);
    hipLaunchKernelGGL(gj_elim_kernel, dim3((m + 63) / 64, (ncol - 1 + 3) / 4), dim3(256), 0, st,
                       m, A, X, k, (const double*)rowk);
  }
  return hipGetLastError();
}

__global__ void cg_note_kernel(double* scal, const double* ws_scal) {
  if (threadIdx.x == 0) {
    scal[8] = scal[0];
    scal[7] = ws_scal && scal[0] <= ws_scal[1] ? 1.0 : 0.0;
  }
}

hipError_t launch_cg_note(double* scal, const double* ws_scal, hipStream_t st) {
  if (g_skip) return no_skip();
  hipLaunchKernelGGL(cg_note_kernel, dim3(1), dim3(64), 0, st, scal, ws_scal);
  return hipGetLastError();
}

hipError_t launch_cg_init(const CgDev& c0, hipStream_t st) {
  CgDev c = c0;
  c.skip = g_skip;
  hipLaunchKernelGGL(cg_init_kernel, dim3(1), dim3(1024), 0, st, c);
  return hipGetLastError();
}
hipError_t launch_cg_step(const CgDev& c, hipStream_t st) {
  hipLaunchKernelGGL(cg_step_kernel, dim3(1), dim3(1024), 0, st, c);
  return hipGetLastError();
}

}  // namespace ibpm
