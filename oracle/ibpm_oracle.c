/*
 * FP64 CPU oracle of the ImmersedBodies.jl IBPM time step.
 *
 * TEST INFRASTRUCTURE: only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark use this file, as the checker. The shipped HIP library never calls it.
 *
 * Each function cites the reference definition it restates (paths relative to the
 * reference repository). Indices in comments are the reference's 1-based ones; the
 * macros below keep that convention so the loops read like the reference loops.
 * Build with -ffp-contract=off so no multiply-add is fused behind the source's back.
 */
#include "ibpm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REG_SUPPORT 3
#define NSUPP 7

/* Planted mistakes. The oracle's math checks (tests/test_oracle_pins.py) build this file
 * again with -DIBO_MUTATION=k to show that each check fails on the slip it guards against.
 * The default build (IBO_MUTATION 0) contains none of them.
 *   1  get_trial_state: boundary-source factor 0.5 dt/(Re hc^2) instead of 0.25 (fluid.jl:475)
 *   2  get_trial_state: coarse ring scaled by 0.5 instead of 0.25 (fluid.jl:480)
 *   3  vort2flux: boundary values from the level being solved instead of level+1 (fluid.jl:195)
 *   4  vort2flux: apply_bc with fac -1 instead of 1 (fluid.jl:196)
 *   5  project_circ: Gamma_1 += instead of -= (coupling.jl:227)
 *   6  couple_surface_rigid: + ub h instead of - ub h (coupling.jl:245)
 *   7  get_trial_state: boundary-source factor with the step of level+1 (fluid.jl:475)
 *   8  vort2flux: coarsify levels from coarse to fine, from stale finer planes (fluid.jl:182-184)
 */
#ifndef IBO_MUTATION
#define IBO_MUTATION 0
#endif
#define MUT(k) (IBO_MUTATION == (k))

/* ------------------------------------------------------------------------------------
 * Complex FFT (mixed radix, recursive decimation in time) used for the DST-I.
 * The reference calls FFTW's RODFT00 (fluid.jl:113-115); that library is not present,
 * so the same transform is formed here from the odd extension of the input.
 * ---------------------------------------------------------------------------------- */
typedef struct { double re, im; } cpx;

typedef struct fft_plan {
  int n;          /* complex length */
  int nfac;
  int fac[64];
  cpx* tw;        /* exp(-2 pi i k / n), k = 0..n-1 */
} fft_plan;

static void plan_init(fft_plan* p, int n) {
  p->n = n;
  p->nfac = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
    if (f * f > m && m > 1) { p->fac[p->nfac++] = m; m = 1; }
  }
  p->tw = (cpx*)malloc(sizeof(cpx) * (size_t)n);
  for (int k = 0; k < n; ++k) {
    /* exact argument reduction: angle = -2 pi k / n */
    double a = -2.0 * M_PI * (double)k / (double)n;
    p->tw[k].re = cos(a);
    p->tw[k].im = sin(a);
  }
}

static void plan_free(fft_plan* p) { free(p->tw); p->tw = NULL; }

static inline cpx cmul(cpx a, cpx b) {
  cpx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}

/* out[k] = sum_j in[j*stride] w^(jk), w = exp(-2 pi i / n); n = N / twstride */
static void fft_rec(const fft_plan* p, const cpx* in, cpx* out, int n, int stride,
                    int fi, int twstride, cpx* scratch) {
  if (n == 1) { out[0] = in[0]; return; }
  int r = p->fac[fi];
  int m = n / r;
  for (int q = 0; q < r; ++q)
    fft_rec(p, in + (size_t)q * stride, out + (size_t)q * m, m, stride * r, fi + 1,
            twstride * r, scratch);
  const cpx* tw = p->tw;
  int N = p->n;
  if (r == 2) {
    for (int k = 0; k < m; ++k) {
      cpx a = out[k];
      cpx b = cmul(out[k + m], tw[(size_t)k * twstride]);
      out[k].re = a.re + b.re; out[k].im = a.im + b.im;
      out[k + m].re = a.re - b.re; out[k + m].im = a.im - b.im;
    }
  } else if (r == 4) {
    for (int k = 0; k < m; ++k) {
      cpx a0 = out[k];
      cpx a1 = cmul(out[k + m], tw[(size_t)k * twstride]);
      cpx a2 = cmul(out[k + 2 * m], tw[(size_t)2 * k * twstride]);
      cpx a3 = cmul(out[k + 3 * m], tw[(size_t)3 * k * twstride]);
      cpx s02 = {a0.re + a2.re, a0.im + a2.im}, d02 = {a0.re - a2.re, a0.im - a2.im};
      cpx s13 = {a1.re + a3.re, a1.im + a3.im}, d13 = {a1.re - a3.re, a1.im - a3.im};
      /* -i * d13 */
      cpx md = {d13.im, -d13.re};
      out[k].re = s02.re + s13.re; out[k].im = s02.im + s13.im;
      out[k + m].re = d02.re + md.re; out[k + m].im = d02.im + md.im;
      out[k + 2 * m].re = s02.re - s13.re; out[k + 2 * m].im = s02.im - s13.im;
      out[k + 3 * m].re = d02.re - md.re; out[k + 3 * m].im = d02.im - md.im;
    }
  } else {
    /* generic radix: DFT of length r on twiddled inputs */
    for (int k = 0; k < m; ++k) {
      for (int q = 0; q < r; ++q)
        scratch[q] = cmul(out[k + (size_t)q * m], tw[((size_t)q * k * twstride) % N]);
      for (int s = 0; s < r; ++s) {
        cpx acc = {0.0, 0.0};
        for (int q = 0; q < r; ++q) {
          cpx w = tw[((size_t)((q * s) % r) * (N / r))];
          cpx t = cmul(scratch[q], w);
          acc.re += t.re; acc.im += t.im;
        }
        out[k + (size_t)s * m] = acc;
      }
    }
  }
}

/* Plan cache keyed by the DST length (a handful of sizes per run). */
typedef struct plan_entry { int n; fft_plan p; } plan_entry;
static plan_entry g_plans[32];
static int g_nplans = 0;

static const fft_plan* get_plan(int n) {
  const fft_plan* found = NULL;
#pragma omp critical(ibo_plan_cache)
  {
    for (int i = 0; i < g_nplans; ++i)
      if (g_plans[i].n == n) found = &g_plans[i].p;
    if (!found) {
      if (g_nplans == 32) {
        for (int i = 0; i < g_nplans; ++i) plan_free(&g_plans[i].p);
        g_nplans = 0;
      }
      g_plans[g_nplans].n = n;
      plan_init(&g_plans[g_nplans].p, n);
      found = &g_plans[g_nplans].p;
      g_nplans++;
    }
  }
  return found;
}

/*
 * DST-I (FFTW RODFT00, unnormalised) of two real vectors a, b of length N at once:
 *   Y_k = 2 sum_n x_n sin(pi (n+1)(k+1)/(N+1)).
 * The odd extension z of length M = 2(N+1) of (a + i b) has FFT Z with
 *   Y_a[k] = -Im Z[k+1],  Y_b[k] = Re Z[k+1].
 */
static void dst_pair(const fft_plan* p, int N, double* a, double* b, cpx* z, cpx* Z,
                     cpx* scratch) {
  int M = 2 * (N + 1);
  z[0].re = 0.0; z[0].im = 0.0;
  z[N + 1].re = 0.0; z[N + 1].im = 0.0;
  for (int n = 0; n < N; ++n) {
    double br = b ? b[n] : 0.0;
    z[n + 1].re = a[n]; z[n + 1].im = br;
    z[M - 1 - n].re = -a[n]; z[M - 1 - n].im = -br;
  }
  fft_rec(p, z, Z, M, 1, 0, 1, scratch);
  for (int k = 0; k < N; ++k) {
    a[k] = -Z[k + 1].im;
    if (b) b[k] = Z[k + 1].re;
  }
}

/* count contiguous vectors of length n, transformed in place */
void ibo_dst1d_batch(int n, int count, double* x) {
  const fft_plan* p = get_plan(2 * (n + 1));
  int npair = (count + 1) / 2;
#pragma omp parallel
  {
    int M = 2 * (n + 1);
    cpx* z = (cpx*)malloc(sizeof(cpx) * (size_t)M);
    cpx* Z = (cpx*)malloc(sizeof(cpx) * (size_t)M);
    cpx* scr = (cpx*)malloc(sizeof(cpx) * (size_t)M);
#pragma omp for schedule(static)
    for (int pr = 0; pr < npair; ++pr) {
      double* a = x + (size_t)(2 * pr) * n;
      double* b = (2 * pr + 1 < count) ? x + (size_t)(2 * pr + 1) * n : NULL;
      dst_pair(p, n, a, b, z, Z, scr);
    }
    free(z); free(Z); free(scr);
  }
}

/* 2-D RODFT00 over both dimensions of an Nx x Ny column-major array (fluid.jl:113-115) */
void ibo_dst2d(const ibo_grid* g, double* x) {
  int Nx = g->nx - 1, Ny = g->ny - 1;
  ibo_dst1d_batch(Nx, Ny, x); /* along x: columns are contiguous */
  const fft_plan* p = get_plan(2 * (Ny + 1));
  int npair = (Nx + 1) / 2;
#pragma omp parallel
  {
    int M = 2 * (Ny + 1);
    cpx* z = (cpx*)malloc(sizeof(cpx) * (size_t)M);
    cpx* Z = (cpx*)malloc(sizeof(cpx) * (size_t)M);
    cpx* scr = (cpx*)malloc(sizeof(cpx) * (size_t)M);
    double* a = (double*)malloc(sizeof(double) * (size_t)Ny);
    double* b = (double*)malloc(sizeof(double) * (size_t)Ny);
#pragma omp for schedule(static)
    for (int pr = 0; pr < npair; ++pr) {
      int i0 = 2 * pr, i1 = 2 * pr + 1;
      int has_b = i1 < Nx;
      for (int j = 0; j < Ny; ++j) {
        a[j] = x[i0 + (size_t)j * Nx];
        if (has_b) b[j] = x[i1 + (size_t)j * Nx];
      }
      dst_pair(p, Ny, a, has_b ? b : NULL, z, Z, scr);
      for (int j = 0; j < Ny; ++j) {
        x[i0 + (size_t)j * Nx] = a[j];
        if (has_b) x[i1 + (size_t)j * Nx] = b[j];
      }
    }
    free(z); free(Z); free(scr); free(a); free(b);
  }
}

/* ------------------------------------------------------------------------------------
 * Grid operators (fluid.jl)
 * ---------------------------------------------------------------------------------- */
#define NXg (g->nx)
#define NYg (g->ny)
/* 1-based accessors in the reference's index convention */
#define GAM(A, i, j) (A)[((i)-1) + (size_t)((j)-1) * (size_t)(NXg - 1)]
#define QXA(A, i, j) (A)[((i)-1) + (size_t)((j)-1) * (size_t)(NXg + 1)]
#define QYA(A, i, j) (A)[((i)-1) + (size_t)((j)-1) * (size_t)(NXg)]
#define BCA(A, k) (A)[(k)-1]

static inline size_t n_gamma(const ibo_grid* g) { return (size_t)(g->nx - 1) * (size_t)(g->ny - 1); }
static inline size_t n_u(const ibo_grid* g) { return (size_t)(g->nx + 1) * (size_t)g->ny; }
static inline size_t n_q(const ibo_grid* g) { return n_u(g) + (size_t)g->nx * (size_t)(g->ny + 1); }
static inline size_t n_bc(const ibo_grid* g) { return 2 * (size_t)(g->nx + 1) + 2 * (size_t)(g->ny + 1); }
/* boundary offsets, problem.jl:380-383 */
#define OFF_L 0
#define OFF_R (NYg + 1)
#define OFF_B (2 * (NYg + 1))
#define OFF_T (2 * (NYg + 1) + NXg + 1)

/* gridstep(grid, level) = 2^(level-1) h, problem.jl:417 */
static inline double gridstep(const ibo_grid* g, int lev) { return pow(2.0, lev - 1) * g->h; }

/* _lap_eigs, fluid.jl:168-173 */
void ibo_lap_eigs(const ibo_grid* g, double* lam) {
  int Nx = g->nx - 1, Ny = g->ny - 1;
  for (int j = 1; j <= Ny; ++j)
    for (int i = 1; i <= Nx; ++i)
      lam[(i - 1) + (size_t)(j - 1) * Nx] =
          -2.0 * (cos(M_PI * i / g->nx) + cos(M_PI * j / g->ny) - 2.0);
}

/* scale / Lambda of Delta_inv_operator, fluid.jl:117-119 with :155,:161 */
void ibo_diag_poisson(const ibo_grid* g, double* d) {
  size_t n = n_gamma(g);
  double scale = 1.0 / (4.0 * g->nx * g->ny);
  ibo_lap_eigs(g, d);
  for (size_t k = 0; k < n; ++k) d[k] = scale / d[k];
}

/* A_operator, fluid.jl:121-134: implicit=0 -> A (explicit half), 1 -> Ainv */
void ibo_diag_A(const ibo_grid* g, double Re, double dt, int lev, int implicit, double* d) {
  size_t n = n_gamma(g);
  double scale = 1.0 / (4.0 * g->nx * g->ny);
  double hc = gridstep(g, lev);
  ibo_lap_eigs(g, d);
  for (size_t k = 0; k < n; ++k) {
    double lam = d[k];
    double Lam;
    if (!implicit)
      Lam = 1.0 / (1.0 - lam * dt / (2.0 * Re * (hc * hc)));
    else
      Lam = 1.0 + lam * dt / (2.0 * Re * (hc * hc));
    d[k] = scale / Lam;
  }
}

/* _lap_inv_operator closure, fluid.jl:145-166; diag holds scale / Lambda */
void ibo_lap_inv(const ibo_grid* g, const double* diag, double* x, const double* b) {
  size_t n = n_gamma(g);
  double* t = (double*)malloc(sizeof(double) * n);
  memcpy(t, b, sizeof(double) * n);
  ibo_dst2d(g, t);
  for (size_t k = 0; k < n; ++k) t[k] *= diag[k];
  ibo_dst2d(g, t);
  memcpy(x, t, sizeof(double) * n);
  free(t);
}

/* rot!, fluid.jl:101-111 */
void ibo_rot(const ibo_grid* g, double* G, const double* q) {
  const double* qx = q;
  const double* qy = q + n_u(g);
  int nx = g->nx, ny = g->ny;
  for (int j = 2; j <= ny; ++j)
    for (int i = 2; i <= nx; ++i)
      GAM(G, i - 1, j - 1) = (QXA(qx, i, j - 1) - QXA(qx, i, j)) + (QYA(qy, i, j) - QYA(qy, i - 1, j));
}

/* curl! with boundary streamfunction, fluid.jl:230-276 */
void ibo_curl(const ibo_grid* g, double* q, const double* psi, const double* bc) {
  double* qx = q;
  double* qy = q + n_u(g);
  int nx = g->nx, ny = g->ny;
  const int L = OFF_L, R = OFF_R, B = OFF_B, T = OFF_T;
  for (int j = 2; j <= ny - 1; ++j)
    for (int i = 2; i <= nx; ++i) QXA(qx, i, j) = GAM(psi, i - 1, j) - GAM(psi, i - 1, j - 1);
  for (int i = 2; i <= nx; ++i) QXA(qx, i, 1) = GAM(psi, i - 1, 1) - BCA(bc, B + i);
  for (int i = 2; i <= nx; ++i) QXA(qx, i, ny) = BCA(bc, i + T) - GAM(psi, i - 1, ny - 1);
  for (int j = 1; j <= ny; ++j) QXA(qx, 1, j) = BCA(bc, (L + 1) + j) - BCA(bc, L + j);
  for (int j = 1; j <= ny; ++j) QXA(qx, nx + 1, j) = BCA(bc, (R + 1) + j) - BCA(bc, R + j);

  for (int j = 2; j <= ny; ++j)
    for (int i = 2; i <= nx - 1; ++i) QYA(qy, i, j) = GAM(psi, i - 1, j - 1) - GAM(psi, i, j - 1);
  for (int j = 2; j <= ny; ++j) QYA(qy, 1, j) = BCA(bc, L + j) - GAM(psi, 1, j - 1);
  for (int j = 2; j <= ny; ++j) QYA(qy, nx, j) = GAM(psi, nx - 1, j - 1) - BCA(bc, R + j);
  for (int i = 1; i <= nx; ++i) QYA(qy, i, 1) = BCA(bc, B + i) - BCA(bc, (B + 1) + i);
  for (int i = 1; i <= nx; ++i) QYA(qy, i, ny + 1) = BCA(bc, T + i) - BCA(bc, (T + 1) + i);
}

/* coarsify!, fluid.jl:208-228 (the ranges use Julia's integer division; their zip stops at
 * the shorter one, which matters when nx or ny is not divisible by 4) */
void ibo_coarsify(const ibo_grid* g, double* Gc, const double* G) {
  int nx = g->nx, ny = g->ny;
  /* is = nx/2 + (-nx/2+2 : 2 : nx/2-2), ics = nx/2 + (-nx/4+1 : nx/4-1) */
  int ni = 0, nj = 0;
  for (int a = -nx / 2 + 2; a <= nx / 2 - 2; a += 2) ni++;
  for (int a = -ny / 2 + 2; a <= ny / 2 - 2; a += 2) nj++;
  int nic = nx / 4 - 1 - (-nx / 4 + 1) + 1;
  int njc = ny / 4 - 1 - (-ny / 4 + 1) + 1;
  if (nic < ni) ni = nic; /* zip stops at the shorter range */
  if (njc < nj) nj = njc;
  for (int bj = 0; bj < nj; ++bj) {
    int j = ny / 2 + (-ny / 2 + 2) + 2 * bj;
    int jc = ny / 2 + (-ny / 4 + 1) + bj;
    for (int bi = 0; bi < ni; ++bi) {
      int i = nx / 2 + (-nx / 2 + 2) + 2 * bi;
      int ic = nx / 2 + (-nx / 4 + 1) + bi;
      GAM(Gc, ic, jc) =
          GAM(G, i, j) +
          0.5 * (GAM(G, i + 1, j) + GAM(G, i, j + 1) + GAM(G, i - 1, j) + GAM(G, i, j - 1)) +
          0.25 * (GAM(G, i + 1, j + 1) + GAM(G, i + 1, j - 1) + GAM(G, i - 1, j - 1) +
                  GAM(G, i - 1, j + 1));
    }
  }
}

/* get_bc!, fluid.jl:278-306 */
void ibo_get_bc(const ibo_grid* g, double* rbc, const double* r) {
  int nx = g->nx, ny = g->ny;
  const int L = OFF_L, R = OFF_R, B = OFF_B, T = OFF_T;
  for (int a = 0; a <= nx / 2; ++a) {
    int i = nx / 4 + a, ibc = 1 + 2 * a;
    BCA(rbc, B + ibc) = GAM(r, i, ny / 4);
    BCA(rbc, T + ibc) = GAM(r, i, 3 * ny / 4);
  }
  for (int a = 1; a <= nx / 2; ++a) {
    int i = nx / 4 + a, ibc = 2 * a;
    BCA(rbc, B + ibc) = 0.5 * (GAM(r, i, ny / 4) + GAM(r, i - 1, ny / 4));
    BCA(rbc, T + ibc) = 0.5 * (GAM(r, i, 3 * ny / 4) + GAM(r, i - 1, 3 * ny / 4));
  }
  for (int a = 0; a <= ny / 2; ++a) {
    int j = ny / 4 + a, jbc = 1 + 2 * a;
    BCA(rbc, L + jbc) = GAM(r, nx / 4, j);
    BCA(rbc, R + jbc) = GAM(r, 3 * nx / 4, j);
  }
  for (int a = 1; a <= ny / 2; ++a) {
    int j = ny / 4 + a, jbc = 2 * a;
    BCA(rbc, L + jbc) = 0.5 * (GAM(r, nx / 4, j) + GAM(r, nx / 4, j - 1));
    BCA(rbc, R + jbc) = 0.5 * (GAM(r, 3 * nx / 4, j) + GAM(r, 3 * nx / 4, j - 1));
  }
}

/* apply_bc!, fluid.jl:308-338 */
void ibo_apply_bc(const ibo_grid* g, double* r, const double* rbc, double fac) {
  int nx = g->nx, ny = g->ny;
  const int L = OFF_L, R = OFF_R, B = OFF_B, T = OFF_T;
  for (int i = 1; i <= nx - 1; ++i) {
    GAM(r, i, 1) += fac * BCA(rbc, (B + 1) + i);
    GAM(r, i, ny - 1) += fac * BCA(rbc, (T + 1) + i);
  }
  for (int j = 1; j <= ny - 1; ++j) {
    GAM(r, 1, j) += fac * BCA(rbc, (L + 1) + j);
    GAM(r, nx - 1, j) += fac * BCA(rbc, (R + 1) + j);
  }
}

/* avg_flux!, fluid.jl:352-370 */
static void avg_flux(const ibo_grid* g, double* Q, const double* q, const double* q0) {
  int nx = g->nx, ny = g->ny;
  const double *qx = q, *qy = q + n_u(g);
  const double *q0x = q0, *q0y = q0 + n_u(g);
  double *Qx = Q, *Qy = Q + n_u(g);
  memset(Q, 0, sizeof(double) * n_q(g));
  for (int j = 2; j <= ny; ++j)
    for (int i = 1; i <= nx + 1; ++i)
      QXA(Qx, i, j) = (QXA(qx, i, j) + QXA(qx, i, j - 1) + QXA(q0x, i, j) + QXA(q0x, i, j - 1)) / 2;
  for (int j = 1; j <= ny + 1; ++j)
    for (int i = 2; i <= nx; ++i)
      QYA(Qy, i, j) = (QYA(qy, i, j) + QYA(qy, i - 1, j) + QYA(q0y, i, j) + QYA(q0y, i - 1, j)) / 2;
}

/* direct_product!, fluid.jl:372-439 */
static void direct_product(const ibo_grid* g, double* fq, const double* Q, const double* G,
                           const double* Gbc) {
  int nx = g->nx, ny = g->ny;
  const int L = OFF_L, R = OFF_R, B = OFF_B, T = OFF_T;
  const double *Qx = Q, *Qy = Q + n_u(g);
  double *fqx = fq, *fqy = fq + n_u(g);
  memset(fq, 0, sizeof(double) * n_q(g));
  for (int j = 2; j <= ny - 1; ++j)
    for (int i = 2; i <= nx; ++i)
      QXA(fqx, i, j) = (QYA(Qy, i, j + 1) * GAM(G, i - 1, j) + QYA(Qy, i, j) * GAM(G, i - 1, j - 1)) / 2;
  for (int i = 2; i <= nx; ++i) {
    int j = 1;
    QXA(fqx, i, j) = (QYA(Qy, i, j + 1) * GAM(G, i - 1, j) + QYA(Qy, i, j) * BCA(Gbc, B + i)) / 2;
  }
  for (int i = 2; i <= nx; ++i) {
    int j = ny;
    QXA(fqx, i, j) = (QYA(Qy, i, j) * GAM(G, i - 1, j - 1) + QYA(Qy, i, j + 1) * BCA(Gbc, T + i)) / 2;
  }
  for (int j = 2; j <= ny; ++j)
    for (int i = 2; i <= nx - 1; ++i)
      QYA(fqy, i, j) = -(QXA(Qx, i + 1, j) * GAM(G, i, j - 1) + QXA(Qx, i, j) * GAM(G, i - 1, j - 1)) / 2;
  for (int j = 2; j <= ny; ++j) {
    int i = 1;
    QYA(fqy, i, j) = -(QXA(Qx, i + 1, j) * GAM(G, i, j - 1) + QXA(Qx, i, j) * BCA(Gbc, L + j)) / 2;
  }
  for (int j = 2; j <= ny; ++j) {
    int i = nx;
    QYA(fqy, i, j) = -(QXA(Qx, i, j) * GAM(G, i - 1, j - 1) + QXA(Qx, i + 1, j) * BCA(Gbc, R + j)) / 2;
  }
}

/* nonlinear closure, fluid.jl:441-455 (rhs_force :340-350) */
void ibo_nonlinear(const ibo_grid* g, double* nonlin, const double* q, const double* q0,
                   const double* G, const double* Gbc, int lev) {
  size_t nq = n_q(g);
  double* Q = (double*)malloc(sizeof(double) * nq);
  double* fq = (double*)malloc(sizeof(double) * nq);
  avg_flux(g, Q, q, q0);
  direct_product(g, fq, Q, G, Gbc);
  ibo_rot(g, nonlin, fq);
  double hc = gridstep(g, lev);
  double s = 1.0 / (hc * hc);
  size_t nG = n_gamma(g);
  for (size_t k = 0; k < nG; ++k) nonlin[k] *= s;
  free(Q);
  free(fq);
}

/* vort2flux closure, fluid.jl:175-206; mutates G (coarse-level centres) */
static void vort2flux_diag(const ibo_grid* g, const double* dpois, double* psi, double* q,
                           double* G) {
  int L = g->nlevel;
  size_t nG = n_gamma(g), nq = n_q(g), nbc = n_bc(g);
  if (MUT(8)) {
    for (int lev = L; lev >= 2; --lev)
      ibo_coarsify(g, G + (size_t)(lev - 1) * nG, G + (size_t)(lev - 2) * nG);
  } else {
    for (int lev = 2; lev <= L; ++lev)
      ibo_coarsify(g, G + (size_t)(lev - 1) * nG, G + (size_t)(lev - 2) * nG);
  }
  memset(psi, 0, sizeof(double) * nG * L);
  double* bc = (double*)calloc(nbc, sizeof(double));
  double* Gt = (double*)malloc(sizeof(double) * nG);
  ibo_lap_inv(g, dpois, psi + (size_t)(L - 1) * nG, G + (size_t)(L - 1) * nG);
  ibo_curl(g, q + (size_t)(L - 1) * nq, psi + (size_t)(L - 1) * nG, bc);
  for (int lev = L - 1; lev >= 1; --lev) {
    memcpy(Gt, G + (size_t)(lev - 1) * nG, sizeof(double) * nG);
    ibo_get_bc(g, bc, psi + (size_t)(MUT(3) ? lev - 1 : lev) * nG);
    ibo_apply_bc(g, Gt, bc, MUT(4) ? -1.0 : 1.0);
    ibo_lap_inv(g, dpois, psi + (size_t)(lev - 1) * nG, Gt);
    ibo_curl(g, q + (size_t)(lev - 1) * nq, psi + (size_t)(lev - 1) * nG, bc);
  }
  free(bc);
  free(Gt);
}

void ibo_vort2flux(const ibo_grid* g, double* psi, double* q, double* G) {
  double* d = (double*)malloc(sizeof(double) * n_gamma(g));
  ibo_diag_poisson(g, d);
  vort2flux_diag(g, d, psi, q, G);
  free(d);
}

/* ------------------------------------------------------------------------------------
 * Regularization (coupling.jl)
 * ---------------------------------------------------------------------------------- */

/* delta_h: Yang3 smoothed 3-point kernel times h, coupling.jl:91-148. The polynomial
 * coefficients are the reference's literal constants; the evaluation keeps its
 * left-to-right order. */
double ibo_delta_h(double rf, double rb, double dr) {
  double r = fabs(rf - rb);
  double r1 = r / dr;
  double r2 = r1 * r1;
  double r3 = r2 * r1;
  double r4 = r3 * r1;
  if (r1 <= 1.0) {
    double a5 = asin((1.0 / 2.0) * sqrt(3.0) * (2.0 * r1 - 1.0));
    double a8 = sqrt(1.0 - 12.0 * r2 + 12.0 * r1);
    return 4.166666667e-2 * r4 + (-0.1388888889 + 3.472222222e-2 * a8) * r3 +
           (-7.121664902e-2 - 5.208333333e-2 * a8 + 0.2405626122 * a5) * r2 +
           (-0.2405626122 * a5 - 0.3792313933 + 0.1012731481 * a8) * r1 +
           8.0187537413e-2 * a5 - 4.195601852e-2 * a8 + 0.6485698427;
  } else if (r1 <= 2.0) {
    double a6 = asin((1.0 / 2.0) * sqrt(3.0) * (-3.0 + 2.0 * r1));
    double a9 = sqrt(-23.0 + 36.0 * r1 - 12.0 * r2);
    return -6.250000000e-2 * r4 + (0.4861111111 - 1.736111111e-2 * a9) * r3 +
           (-1.143175026 + 7.812500000e-2 * a9 - 0.1202813061 * a6) * r2 +
           (0.8751991178 + 0.3608439183 * a6 - 0.1548032407 * a9) * r1 - 0.2806563809 * a6 +
           8.22848104e-3 + 0.1150173611 * a9;
  } else if (r1 <= 3.0) {
    double a1 = asin((1.0 / 2.0 * (2.0 * r1 - 5.0)) * sqrt(3.0));
    double a7 = sqrt(-71.0 - 12.0 * r2 + 60.0 * r1);
    return 2.083333333e-2 * r4 + (3.472222222e-3 * a7 - 0.2638888889) * r3 +
           (1.214391675 - 2.604166667e-2 * a7 + 2.405626122e-2 * a1) * r2 +
           (-0.1202813061 * a1 - 2.449273192 + 7.262731481e-2 * a7) * r1 + 0.1523563211 * a1 +
           1.843201677 - 7.306134259e-2 * a7;
  }
  return 0.0;
}

/* _update!(reg, index, point), coupling.jl:27-54.
 * coords[2p] = ix, coords[2p+1] = iy; w[(p*2 + c)*49 + b*7 + a] = weights[p, c, a, b].
 * Returns -1 - p for the first panel outside the allowed region, 0 on success. */
int ibo_reg_update(const ibo_grid* g, int npanel, const double* xb, int* coords, double* w) {
  double h = g->h;
  int nx = g->nx, ny = g->ny;
  for (int p = 0; p < npanel; ++p) {
    double px = xb[p], py = xb[npanel + p];
    /* round(Int, (pb - p0)/h + 1.0 - 0.25): round half to even as Julia's round */
    int ix = (int)nearbyint((px - g->x0) / h + 1.0 - 0.25);
    int iy = (int)nearbyint((py - g->y0) / h + 1.0 - 0.25);
    if (!(REG_SUPPORT < ix && ix <= nx - REG_SUPPORT)) return -1 - p;
    if (!(REG_SUPPORT < iy && iy <= ny - REG_SUPPORT)) return -1 - p;
    coords[2 * p] = ix;
    coords[2 * p + 1] = iy;
    double xs[NSUPP], ys[NSUPP];
    for (int k = -REG_SUPPORT; k <= REG_SUPPORT; ++k) {
      xs[k + REG_SUPPORT] = g->x0 + h * (double)(ix + k - 1);
      ys[k + REG_SUPPORT] = g->y0 + h * (double)(iy + k - 1);
    }
    for (int b = 0; b < NSUPP; ++b)
      for (int a = 0; a < NSUPP; ++a) {
        w[((size_t)p * 2 + 0) * 49 + b * 7 + a] =
            ibo_delta_h(xs[a], px, h) * ibo_delta_h(ys[b] + h / 2, py, h);
        w[((size_t)p * 2 + 1) * 49 + b * 7 + a] =
            ibo_delta_h(xs[a] + h / 2, px, h) * ibo_delta_h(ys[b], py, h);
      }
  }
  return 0;
}

/* regT! (E, interpolation onto panels), coupling.jl:57-69 */
void ibo_regT(const ibo_grid* g, int npanel, const int* coords, const double* w, double* fb,
              const double* q1) {
  const double* qx = q1;
  const double* qy = q1 + n_u(g);
  for (int p = 0; p < npanel; ++p) {
    int ix = coords[2 * p], iy = coords[2 * p + 1];
    double sx = 0.0, sy = 0.0;
    for (int b = 0; b < NSUPP; ++b)
      for (int a = 0; a < NSUPP; ++a) {
        int i = ix + a - REG_SUPPORT, j = iy + b - REG_SUPPORT;
        sx += QXA(qx, i, j) * w[((size_t)p * 2 + 0) * 49 + b * 7 + a];
        sy += QYA(qy, i, j) * w[((size_t)p * 2 + 1) * 49 + b * 7 + a];
      }
    fb[p] = sx;
    fb[npanel + p] = sy;
  }
}

/* reg! (E^T, regularisation onto the fluid), coupling.jl:72-85 */
void ibo_reg(const ibo_grid* g, int npanel, const int* coords, const double* w, double* q1,
             const double* fb) {
  double* qx = q1;
  double* qy = q1 + n_u(g);
  memset(q1, 0, sizeof(double) * n_q(g));
  for (int p = 0; p < npanel; ++p) {
    int ix = coords[2 * p], iy = coords[2 * p + 1];
    for (int b = 0; b < NSUPP; ++b)
      for (int a = 0; a < NSUPP; ++a) {
        int i = ix + a - REG_SUPPORT, j = iy + b - REG_SUPPORT;
        QXA(qx, i, j) += w[((size_t)p * 2 + 0) * 49 + b * 7 + a] * fb[p];
        QYA(qy, i, j) += w[((size_t)p * 2 + 1) * 49 + b * 7 + a] * fb[npanel + p];
      }
  }
}

/* ------------------------------------------------------------------------------------
 * Solver (solver.jl:78-190)
 * ---------------------------------------------------------------------------------- */
struct ibo_solver {
  ibo_grid g;
  double Re, dt, t0, t;
  int index;
  int npanel, nbody, moving;
  size_t nG, nq, nbc;
  /* state (problem.jl:626-665) */
  double *q, *q0, *G, *psi, *nonlin[2], *xb, *ds, *ub, *Ft, *fb, total[2], fs[2];
  int grid_moving;
  double gm[6];
  /* reference body description */
  double* xb_ref;
  int *body_start, *body_count, *body_moving;
  /* solver temporaries */
  double *qs, *Gs, *Gtmp, *psitmp, *qtmp, *Ftmp, *rhs, *rhsbc, *Gt1, *q1, *q2, *bc;
  double *dpois, *dA, *dAinv; /* nlevel planes each for A / Ainv */
  int* coords;
  double* w;
  double* Binv; /* static: dense inverse */
  /* CG */
  int cg_maxiter, cg_tol_mode;
  double cg_reltol, cg_abstol;
  int last_cg_iters;
  double last_cg_res;
  double *cg_u, *cg_r, *cg_c;
};

void ibo_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static void lap_inv_level(ibo_solver* s, const double* d, double* x, const double* b) {
  ibo_lap_inv(&s->g, d, x, b);
}

/* B_operator closure, coupling.jl:150-177 */
void ibo_apply_B(ibo_solver* s, double* x, const double* z) {
  const ibo_grid* g = &s->g;
  memset(s->Gtmp, 0, sizeof(double) * s->nG * g->nlevel);
  ibo_reg(g, s->npanel, s->coords, s->w, s->q2, z);
  ibo_rot(g, s->rhsbc, s->q2);
  lap_inv_level(s, s->dAinv, s->Gtmp, s->rhsbc);
  vort2flux_diag(g, s->dpois, s->psitmp, s->qtmp, s->Gtmp);
  ibo_regT(g, s->npanel, s->coords, s->w, x, s->qtmp);
}

/* Julia's LinearAlgebra.norm for a vector of finite values */
static double vnorm(const double* v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += v[i] * v[i];
  return sqrt(s);
}
static double vdot(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* cg! of IterativeSolvers.jl (0.9.x, unpreconditioned CGIterable), as called from
 * coupling.jl:194-202 with the keywords of solver.jl:137: x is the initial guess. */
static void cg_solve(ibo_solver* s, double* x, const double* b) {
  int n = 2 * s->npanel;
  double *u = s->cg_u, *r = s->cg_r, *c = s->cg_c;
  for (int i = 0; i < n; ++i) u[i] = 0.0;
  memcpy(r, b, sizeof(double) * n);
  ibo_apply_B(s, c, x);
  for (int i = 0; i < n; ++i) r[i] -= c[i];
  double residual = vnorm(r, n);
  double ref = s->cg_tol_mode == 1 ? vnorm(b, n) : residual;
  double tol = s->cg_reltol * ref;
  if (s->cg_abstol > tol) tol = s->cg_abstol;
  double prev = 1.0;
  int it = 0;
  while (!(it >= s->cg_maxiter || residual <= tol)) {
    double beta = residual * residual / (prev * prev);
    for (int i = 0; i < n; ++i) u[i] = r[i] + beta * u[i];
    ibo_apply_B(s, c, u);
    double alpha = residual * residual / vdot(u, c, n);
    for (int i = 0; i < n; ++i) x[i] += alpha * u[i];
    for (int i = 0; i < n; ++i) r[i] -= alpha * c[i];
    prev = residual;
    residual = vnorm(r, n);
    it++;
  }
  s->last_cg_iters = it;
  s->last_cg_res = residual;
}

/* LU with partial pivoting, then the inverse column by column (LAPACK getrf/getri in
 * the reference's inv, coupling.jl:190). Returns nonzero for a singular matrix. */
static int dense_inverse(int n, const double* Ain, double* Ainv) {
  double* A = (double*)malloc(sizeof(double) * (size_t)n * n);
  int* piv = (int*)malloc(sizeof(int) * n);
  memcpy(A, Ain, sizeof(double) * (size_t)n * n);
#define AE(i, j) A[(i) + (size_t)(j) * n]
  for (int k = 0; k < n; ++k) {
    int p = k;
    double mx = fabs(AE(k, k));
    for (int i = k + 1; i < n; ++i)
      if (fabs(AE(i, k)) > mx) { mx = fabs(AE(i, k)); p = i; }
    if (mx == 0.0) { free(A); free(piv); return 1; }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) { double t = AE(k, j); AE(k, j) = AE(p, j); AE(p, j) = t; }
    double inv = 1.0 / AE(k, k);
    for (int i = k + 1; i < n; ++i) AE(i, k) *= inv;
#pragma omp parallel for schedule(static)
    for (int j = k + 1; j < n; ++j) {
      double akj = AE(k, j);
      if (akj != 0.0)
        for (int i = k + 1; i < n; ++i) AE(i, j) -= AE(i, k) * akj;
    }
  }
#pragma omp parallel for schedule(static)
  for (int col = 0; col < n; ++col) {
    double* x = Ainv + (size_t)col * n;
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    x[col] = 1.0;
    for (int k = 0; k < n; ++k)
      if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < i; ++k) x[i] -= AE(i, k) * x[k];
    for (int i = n - 1; i >= 0; --i) {
      for (int k = i + 1; k < n; ++k) x[i] -= AE(i, k) * x[k];
      x[i] /= AE(i, i);
    }
  }
#undef AE
  free(A);
  free(piv);
  return 0;
}

static void seterr(char* err, size_t len, const char* msg) {
  if (err && len) { strncpy(err, msg, len - 1); err[len - 1] = 0; }
}

static int reg_update_range(ibo_solver* s, int start, int count, char* err, size_t errlen) {
  /* update a contiguous panel range (bodies.jl:33-51 -> coupling.jl:20-25) */
  int n = s->npanel;
  double* xbr = (double*)malloc(sizeof(double) * 2 * count);
  for (int k = 0; k < count; ++k) {
    xbr[k] = s->xb[start + k];
    xbr[count + k] = s->xb[n + start + k];
  }
  int rc = ibo_reg_update(&s->g, count, xbr, s->coords + 2 * start, s->w + (size_t)start * 98);
  if (rc != 0) {
    int p = -rc - 1 + start;
    char msg[256];
    snprintf(msg, sizeof msg, "Panel index=%d at (%.17g, %.17g) outside allowed region", p + 1,
             s->xb[p], s->xb[n + p]);
    seterr(err, errlen, msg);
  }
  free(xbr);
  return rc;
}

ibo_solver* ibo_create(const ibo_config* c, char* err, size_t errlen) {
  const ibo_grid* g = &c->grid;
  if (g->nx < 8 || g->ny < 8 || g->nlevel < 1) {
    seterr(err, errlen, "nx and ny must be at least 8 and nlevel >= 1");
    return NULL;
  }
  ibo_solver* s = (ibo_solver*)calloc(1, sizeof(ibo_solver));
  s->g = *g;
  s->Re = c->Re; s->dt = c->dt; s->t0 = c->t0;
  s->npanel = c->npanel; s->nbody = c->nbody;
  s->nG = n_gamma(g); s->nq = n_q(g); s->nbc = n_bc(g);
  int L = g->nlevel, n = c->npanel;
  size_t nG = s->nG, nq = s->nq;
  s->q = calloc(nq * L, 8); s->q0 = calloc(nq * L, 8);
  s->G = calloc(nG * L, 8); s->psi = calloc(nG * L, 8);
  s->nonlin[0] = calloc(nG * L, 8); s->nonlin[1] = calloc(nG * L, 8);
  s->xb = calloc(2 * n + 1, 8); s->ds = calloc(n + 1, 8); s->ub = calloc(2 * n + 1, 8);
  s->Ft = calloc(2 * n + 1, 8); s->fb = calloc(2 * n + 1, 8);
  s->xb_ref = calloc(2 * n + 1, 8);
  s->body_start = calloc(c->nbody + 1, sizeof(int));
  s->body_count = calloc(c->nbody + 1, sizeof(int));
  s->body_moving = calloc(c->nbody + 1, sizeof(int));
  memcpy(s->xb_ref, c->xb, sizeof(double) * 2 * n);
  s->moving = 0;
  s->grid_moving = c->grid_moving ? 1 : 0;
  for (int k = 0; k < 6; ++k) s->gm[k] = 0.0;
  for (int b = 0; b < c->nbody; ++b) {
    s->body_start[b] = c->body_start[b];
    s->body_count[b] = c->body_count[b];
    s->body_moving[b] = c->body_moving[b];
    if (c->body_moving[b]) s->moving = 1;
  }
  s->qs = calloc(nq * L, 8); s->Gs = calloc(nG * L, 8); s->Gtmp = calloc(nG * L, 8);
  s->psitmp = calloc(nG * L, 8); s->qtmp = calloc(nq * L, 8); s->Ftmp = calloc(2 * n + 1, 8);
  s->rhs = calloc(nG, 8); s->rhsbc = calloc(nG, 8); s->Gt1 = calloc(nG, 8);
  s->q1 = calloc(nq, 8); s->q2 = calloc(nq, 8); s->bc = calloc(s->nbc, 8);
  s->dpois = malloc(nG * 8); s->dA = malloc(nG * L * 8); s->dAinv = malloc(nG * L * 8);
  s->coords = calloc(2 * n + 2, sizeof(int)); s->w = calloc((size_t)n * 98 + 1, 8);
  s->cg_u = calloc(2 * n + 1, 8); s->cg_r = calloc(2 * n + 1, 8); s->cg_c = calloc(2 * n + 1, 8);
  s->cg_maxiter = c->cg_maxiter > 0 ? c->cg_maxiter : 5000;
  s->cg_reltol = c->cg_reltol > 0 ? c->cg_reltol : 1e-12;
  s->cg_abstol = c->cg_abstol;
  s->cg_tol_mode = c->cg_tol_mode;

  ibo_diag_poisson(g, s->dpois);
  for (int lev = 1; lev <= L; ++lev) {
    ibo_diag_A(g, c->Re, c->dt, lev, 0, s->dA + (size_t)(lev - 1) * nG);
    ibo_diag_A(g, c->Re, c->dt, lev, 1, s->dAinv + (size_t)(lev - 1) * nG);
  }
  /* State(prob): index 0, t = t0 - dt (problem.jl:661,703-708) */
  ibo_set_index(s, 0);
  /* _init_body_motion!, bodies.jl:1-6,33-41: static bodies take their panels now */
  for (int b = 0; b < c->nbody; ++b) {
    int st = c->body_start[b], cnt = c->body_count[b];
    if (!c->body_moving[b]) {
      for (int k = 0; k < cnt; ++k) {
        s->xb[st + k] = c->xb[st + k];
        s->xb[n + st + k] = c->xb[n + st + k];
        s->ds[st + k] = c->ds ? c->ds[st + k] : 0.0;
        s->ub[st + k] = 0.0;
        s->ub[n + st + k] = 0.0;
      }
      if (reg_update_range(s, st, cnt, err, errlen)) { ibo_destroy(s); return NULL; }
    }
  }
  /* Binv_operator, coupling.jl:179-192 */
  if (!s->moving && n > 0 && c->timing_only) {
    int m = 2 * n;
    s->Binv = calloc((size_t)m * m, sizeof(double));
    for (int i = 0; i < m; ++i) s->Binv[i + (size_t)i * m] = 1.0;
  } else if (!s->moving && n > 0) {
    int m = 2 * n;
    double* M = malloc(sizeof(double) * (size_t)m * m);
    double* e = calloc(m, 8);
    for (int i = 0; i < m; ++i) {
      e[i] = 1.0;
      if (i > 0) e[i - 1] = 0.0;
      ibo_apply_B(s, M + (size_t)i * m, e);
    }
    s->Binv = malloc(sizeof(double) * (size_t)m * m);
    if (dense_inverse(m, M, s->Binv)) {
      seterr(err, errlen, "B is singular");
      free(M); free(e); ibo_destroy(s); return NULL;
    }
    free(M);
    free(e);
  }
  return s;
}

void ibo_destroy(ibo_solver* s) {
  if (!s) return;
  void* ptrs[] = {s->q, s->q0, s->G, s->psi, s->nonlin[0], s->nonlin[1], s->xb, s->ds, s->ub,
                  s->Ft, s->fb, s->xb_ref, s->body_start, s->body_count, s->body_moving, s->qs,
                  s->Gs, s->Gtmp, s->psitmp, s->qtmp, s->Ftmp, s->rhs, s->rhsbc, s->Gt1, s->q1,
                  s->q2, s->bc, s->dpois, s->dA, s->dAinv, s->coords, s->w, s->Binv, s->cg_u,
                  s->cg_r, s->cg_c};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
  free(s);
}

void ibo_set_grid_motion(ibo_solver* s, const double gm[6]) {
  for (int k = 0; k < 6; ++k) s->gm[k] = gm[k];
}

void ibo_set_index(ibo_solver* s, int index) {
  s->index = index;
  s->t = s->t0 + s->dt * (double)(index - 1);
}
int ibo_index(const ibo_solver* s) { return s->index; }
double ibo_time(const ibo_solver* s) { return s->t; }
int ibo_last_cg_iters(const ibo_solver* s) { return s->last_cg_iters; }
double ibo_last_cg_residual(const ibo_solver* s) { return s->last_cg_res; }
const double* ibo_Binv(const ibo_solver* s) { return s->Binv; }

double* ibo_field(ibo_solver* s, int which, size_t* n) {
  int L = s->g.nlevel, np = s->npanel;
  switch (which) {
    case 0: *n = s->nq * L; return s->q;
    case 1: *n = s->nq * L; return s->q0;
    case 2: *n = s->nG * L; return s->G;
    case 3: *n = s->nG * L; return s->psi;
    case 4: *n = s->nG * L; return s->nonlin[0];
    case 5: *n = s->nG * L; return s->nonlin[1];
    case 6: *n = 2 * (size_t)np; return s->xb;
    case 7: *n = 2 * (size_t)np; return s->ub;
    case 8: *n = 2 * (size_t)np; return s->Ft;
    case 9: *n = 2 * (size_t)np; return s->fb;
    case 10: {
      double fx = 0.0, fy = 0.0;
      for (int p = 0; p < np; ++p) { fx += s->fb[p]; fy += s->fb[np + p]; }
      s->total[0] = fx; s->total[1] = fy;
      *n = 2; return s->total;
    }
    case 11: *n = (size_t)np; return s->ds;
    case 12: *n = 2; return s->fs;
  }
  *n = 0;
  return NULL;
}

/* get_trial_state closure, fluid.jl:457-507 */
static void get_trial_state(ibo_solver* s) {
  const ibo_grid* g = &s->g;
  int L = g->nlevel;
  size_t nG = s->nG, nq = s->nq;
  const double beta[2] = {1.5, -0.5}; /* CNAB, problem.jl:566-577 */
  for (int lev = L; lev >= 1; --lev) {
    memset(s->bc, 0, sizeof(double) * s->nbc);
    memset(s->rhsbc, 0, sizeof(double) * nG);
    double hc = gridstep(g, lev);
    if (lev < L) {
      double hf = MUT(7) ? gridstep(g, lev + 1) : hc;
      ibo_get_bc(g, s->bc, s->G + (size_t)lev * nG);
      ibo_apply_bc(g, s->rhsbc, s->bc, (MUT(1) ? 0.5 : 0.25) * s->dt / (s->Re * (hf * hf)));
    }
    for (size_t k = 0; k < s->nbc; ++k) s->bc[k] *= MUT(2) ? 0.5 : 0.25;
    ibo_nonlinear(g, s->nonlin[0] + (size_t)(lev - 1) * nG, s->q + (size_t)(lev - 1) * nq,
                  s->q0 + (size_t)(lev - 1) * nq, s->G + (size_t)(lev - 1) * nG, s->bc, lev);
    lap_inv_level(s, s->dA + (size_t)(lev - 1) * nG, s->rhs, s->G + (size_t)(lev - 1) * nG);
    for (int m = 0; m < 2; ++m) {
      const double* nl = s->nonlin[m] + (size_t)(lev - 1) * nG;
      for (size_t k = 0; k < nG; ++k) s->rhs[k] += s->dt * beta[m] * nl[k];
    }
    for (size_t k = 0; k < nG; ++k) s->rhs[k] += s->rhsbc[k];
    lap_inv_level(s, s->dAinv + (size_t)(lev - 1) * nG, s->Gs + (size_t)(lev - 1) * nG, s->rhs);
  }
  /* _cycle!: [n1, n2] -> [n2, n1] (fluid.jl:510) */
  double* t = s->nonlin[1];
  s->nonlin[1] = s->nonlin[0];
  s->nonlin[0] = t;
  vort2flux_diag(g, s->dpois, s->psi, s->qs, s->Gs);
}

int ibo_advance(ibo_solver* s, const double fs_vel[2], const double* motion, char* err,
                size_t errlen) {
  const ibo_grid* g = &s->g;
  int L = g->nlevel, n = s->npanel;
  size_t nG = s->nG, nq = s->nq;
  double h = g->h;
  /* inc_time_index! (problem.jl:701-708) */
  ibo_set_index(s, s->index + 1);
  s->fs[0] = fs_vel[0];
  s->fs[1] = fs_vel[1];
  /* _update_body_motion!, bodies.jl:12-31,43-51 */
  for (int b = 0; b < s->nbody; ++b) {
    if (!s->body_moving[b]) continue;
    const double* m = motion + 6 * b;
    double r0x = m[0], r0y = m[1], v0x = m[2], v0y = m[3], ang = m[4], om = m[5];
    double c = cos(ang), sn = sin(ang);
    double Rv11 = om * -sn, Rv12 = om * -c, Rv21 = om * c, Rv22 = om * -sn;
    int st = s->body_start[b], cnt = s->body_count[b];
    for (int k = st; k < st + cnt; ++k) {
      double rx = s->xb_ref[k], ry = s->xb_ref[n + k];
      s->xb[k] = r0x + (c * rx + -sn * ry);
      s->xb[n + k] = r0y + (sn * rx + c * ry);
      s->ub[k] = v0x + (Rv11 * rx + Rv12 * ry);
      s->ub[n + k] = v0y + (Rv21 * rx + Rv22 * ry);
    }
    if (reg_update_range(s, st, cnt, err, errlen)) return 1;
  }
  if (!s->grid_moving) {
    /* base_flux! for a static grid, fluid.jl:22-32 */
    for (int lev = 1; lev <= L; ++lev) {
      double hc = gridstep(g, lev);
      double* q0 = s->q0 + (size_t)(lev - 1) * nq;
      size_t nu = n_u(g);
      for (size_t k = 0; k < nu; ++k) q0[k] = s->fs[0] * hc;
      for (size_t k = nu; k < nq; ++k) q0[k] = s->fs[1] * hc;
    }
  } else {
    /* base_flux! for a moving grid, fluid.jl:34-67: per face,
       q0 = hc Rx[:, c] . ((uinf - u0) - Rv ((x, y) - centre)) */
    const double cx = s->gm[0], cy = s->gm[1], u0x = s->gm[2], u0y = s->gm[3];
    const double ct = cos(s->gm[4]), st = sin(s->gm[4]), om = s->gm[5];
    const double Rx11 = ct, Rx21 = st, Rx12 = -st, Rx22 = ct;
    const double Rv11 = om * -st, Rv12 = om * -ct, Rv21 = om * ct, Rv22 = om * -st;
    const int nx = g->nx, ny = g->ny;
    for (int lev = 1; lev <= L; ++lev) {
      double hc = gridstep(g, lev);
      /* subdomain (problem.jl:424-429) */
      double sx0 = g->x0 + (nx * (g->h - hc)) / 2, sy0 = g->y0 + (ny * (g->h - hc)) / 2;
      double* qx = s->q0 + (size_t)(lev - 1) * nq;
      double* qy = qx + n_u(g);
      /* GridU: (x0 + hc (i-1), y0 + hc/2 + hc (j-1)) (problem.jl:318-328) */
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx + 1; ++i) {
          double x = sx0 + hc * (double)(i - 1), y = (sy0 + hc / 2) + hc * (double)(j - 1);
          double dx = x - cx, dy = y - cy;
          double wx = (s->fs[0] - u0x) - (Rv11 * dx + Rv12 * dy);
          double wy = (s->fs[1] - u0y) - (Rv21 * dx + Rv22 * dy);
          qx[(i - 1) + (size_t)(j - 1) * (nx + 1)] = hc * (Rx11 * wx + Rx21 * wy);
        }
      /* GridV: (x0 + hc/2 + hc (i-1), y0 + hc (j-1)) (problem.jl:330-334) */
      for (int j = 1; j <= ny + 1; ++j)
        for (int i = 1; i <= nx; ++i) {
          double x = (sx0 + hc / 2) + hc * (double)(i - 1), y = sy0 + hc * (double)(j - 1);
          double dx = x - cx, dy = y - cy;
          double wx = (s->fs[0] - u0x) - (Rv11 * dx + Rv12 * dy);
          double wy = (s->fs[1] - u0y) - (Rv21 * dx + Rv22 * dy);
          qy[(i - 1) + (size_t)(j - 1) * nx] = hc * (Rx12 * wx + Rx22 * wy);
        }
    }
  }
  get_trial_state(s);
  /* couple_surface_rigid, coupling.jl:233-251 */
  if (n > 0) {
    for (size_t k = 0; k < nq; ++k) s->q1[k] = s->qs[k] + s->q0[k];
    ibo_regT(g, n, s->coords, s->w, s->Ftmp, s->q1);
    for (int k = 0; k < 2 * n; ++k) s->Ftmp[k] -= (MUT(6) ? -1.0 : 1.0) * (s->ub[k] * h);
    if (!s->moving) {
      int m = 2 * n;
      for (int i = 0; i < m; ++i) {
        double acc = 0.0;
        for (int j = 0; j < m; ++j) acc += s->Binv[i + (size_t)j * m] * s->Ftmp[j];
        s->Ft[i] = acc;
      }
      s->last_cg_iters = 0;
    } else {
      cg_solve(s, s->Ft, s->Ftmp);
    }
  }
  /* project_circ, coupling.jl:217-231 */
  memcpy(s->G, s->Gs, sizeof(double) * nG * L);
  if (n > 0) {
    ibo_reg(g, n, s->coords, s->w, s->q1, s->Ft);
    ibo_rot(g, s->rhsbc, s->q1);
    lap_inv_level(s, s->dAinv, s->Gs, s->rhsbc);
    if (MUT(5))
      for (size_t k = 0; k < nG; ++k) s->G[k] += s->Gs[k];
    else
      for (size_t k = 0; k < nG; ++k) s->G[k] -= s->Gs[k];
  }
  /* final vort2flux, solver.jl:184 */
  vort2flux_diag(g, s->dpois, s->psi, s->q, s->G);
  /* update_vars!, problem.jl:676-685 */
  for (int k = 0; k < 2 * n; ++k) s->fb[k] = s->Ft[k] * h / s->dt;
  return 0;
}
