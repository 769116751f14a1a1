// Internal declarations shared by the HIP translation units of libibpm.
//
// Layout of every field in HBM is the reference's (ImmersedBodies.jl, problem.jl:363-389):
// column-major float64, one plane per multi-domain level.
//   Gamma / psi plane : Nx x Ny, Nx = nx-1, Ny = ny-1, x fastest
//   q plane           : u ((nx+1) x ny) followed by v (nx x (ny+1))
//   bc vector         : left (ny+1), right (ny+1), bottom (nx+1), top (nx+1)
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

struct ibpm_net;  // warm-start network (include/ibpm.h)

namespace ibpm {

// Sets the thread-local message returned by ibpm_last_error() and returns `code`.
int set_error(int code, const char* msg);

// Launch context of the calling thread. While `g_skip` points at a device flag, every
// kernel launched by the launchers below returns at once when that flag is nonzero; the
// CG iterations of a captured step run under the CG's converged flag, so a replay does
// the work of the iterations CG needs and exits the rest. Launchers of kernels that do not
// take the flag refuse to launch while it is set (hipErrorNotSupported).
extern thread_local const double* g_skip;
struct SkipScope {
  const double* saved;
  explicit SkipScope(const double* flag) : saved(g_skip) { g_skip = flag; }
  ~SkipScope() { g_skip = saved; }
};
inline hipError_t no_skip() { return g_skip ? hipErrorNotSupported : hipSuccess; }

// Geometry of one multi-domain grid (same for all its levels).
struct GridDims {
  int nx, ny;   // cells per level
  int Nx, Ny;   // interior vertices per direction
  int nu;       // (nx+1)*ny
  int nq;       // nu + nx*(ny+1)
  int nG;       // Nx*Ny
};

inline GridDims make_dims(int nx, int ny) {
  GridDims d;
  d.nx = nx; d.ny = ny; d.Nx = nx - 1; d.Ny = ny - 1;
  d.nu = (nx + 1) * ny;
  d.nq = d.nu + nx * (ny + 1);
  d.nG = d.Nx * d.Ny;
  return d;
}

// One plane of one sine-contraction pass.
//   in  : Nk x Nb, k contiguous (in[k + b*Nk])
//   out : Nb x Nk, b contiguous (out[b + j*Nb])
// EPI_DIAG multiplies by the spectral diagonal D = dscale / (da + lambda*dnum/dden) with
// lambda = -2((cos_b[b] + cos_j[j]) - 2) (the reference's _lap_eigs, fluid.jl:168-173);
// EPI_SUB / EPI_ADD combine with `base`.
struct PassPlane {
  const void* in;
  void* out;
  const double* base;
  const double* cosb;  // cos(pi k / n) along the output's b direction
  const double* cosj;  // along the transformed direction
  const void* So;      // odd-row sine table of the length-Nk transform
  const void* Se;      // even-row sine table
  const void* Sq;      // FP64 double-fold tables [T1 | T2 | T4]
  const void* Sqm;     // mixed double-fold tables, split bf16 (null: single fold only)
  double* spec;        // EPI_DIAG, FP64: also store val / (da + lambda dnum / dden) here
                       // (the plane a forward transform of the solve's result would give)
  double dscale, da, dnum, dden;
  int Nk, Nb;
};

enum Epilogue : int {
  EPI_STORE = 0,   // out = val
  EPI_DIAG = 1,    // out = val * diag
  EPI_SUB = 2,     // out = base - val
  EPI_ADD = 3,     // out = base + val
};

enum Precision : int { PREC_F64 = 0, PREC_MIXED = 1 };

// Up to kMaxPlanes planes of one pass travel by value in the kernel arguments.
constexpr int kMaxPlanes = 16;
struct PassBatch {
  PassPlane p[kMaxPlanes];
  int nbt, nmt, nplanes;  // tile grid (filled by the launcher)
  const double* skip;     // launch context flag (filled by the launcher)
};

// Host launchers (dst_kernels.hip).
// variant: tile variant 0..19: v % 4 picks (TB, TM) = (64,64), (64,32), (32,64), (32,32)
// for v < 16, loads are issued 1 (v % 8 < 4) or 2 K chunks ahead, and v >= 8 folds the
// sine matrix twice (four quarter-size contractions) instead of once; 16..19 (FP64) copy
// the tables straight into LDS. The mixed pass has 0..7 (one fold) and 8..11 (double
// fold, one prefetch stage). -1 takes the autotuned variant of the launch shape
// (dst_default_variant when the shape was never tuned).
hipError_t launch_dst_pass_f64(const PassBatch& b, int nplanes, int maxNb, int maxH, int epi,
                               hipStream_t st, int variant = -1);
hipError_t launch_dst_pass_mixed(const PassBatch& b, int nplanes, int maxNb, int maxH, int epi,
                                 bool in_f32, bool out_f32, hipStream_t st, int variant = -1);
// Autotuner: time every tile variant of a pass launch shape (planes, largest Nb and H)
// once, remember the fastest for later launches of that shape (process-wide). Runs real
// passes on the given planes (EPI_STORE); not to be called under graph capture.
hipError_t tune_dst_pass(const PassBatch& b, int nplanes, int maxNb, int maxH, bool mixed,
                         bool in_f32, bool out_f32, hipStream_t st, int* chosen, float* best_ms);
// (TB, TM, PF, fold, KC) of variant v; PF 0: sine tables copied straight into LDS
int dst_variant_tiles(int v, bool mixed, int* TB, int* TM, int* PF, int* fold, int* KC = nullptr);
constexpr int kDstVariants = 24;  // FP64 pass variants (the mixed pass has 20)
// Register-only MFMA loop (f64 16x16x4 or bf16 16x16x32) on `blocks` x 256 threads; out
// needs blocks x 256 doubles. flops: matrix flops of the launch.
hipError_t launch_mfma_probe(bool bf16, int iters, int blocks, double* out, hipStream_t st,
                             double* flops);
// Tuned variant of a launch shape, -1 when never tuned.
int dst_tuned_variant(int mixed, int nplanes, int maxNb, int maxH);
// Variant of the shapes never tuned: double fold (32 x 64) in both precisions.
int dst_default_variant(bool mixed);
// The variant a launch along a length-N transform runs: a double-fold variant becomes its
// single-fold counterpart unless N + 1 is a multiple of 4 (mixed: of 16, where the solver
// builds the split double-fold tables).
int dst_effective_variant(int v, int N, bool mixed = false);

// ---- multi-level ring chain (ring_kernels.hip) ----
constexpr int kMaxChains = 16;
struct RingChainArgs {
  int nx, ny, Nx, Ny;
  int write_e;                      // 0: the level's E is not read (projections only)
  const void* F[kMaxChains];        // spectral plane of the level, per chain
  void* E[kMaxChains];              // E = D . (F + R), written (same precision as F)
  int F_f32;
  const double* cosx;               // cos(pi kx / nx), kx = 1..Nx
  const double* cosy;
  double scale;                     // 1 / (4 nx ny)
  const double* fac_in[kMaxChains]; // ring factors of this level (null: none)
  double* fac_out[kMaxChains];      // ring factors of the next finer level
  double* vec[kMaxChains];          // projections a_b, a_t (Nx each), c_l, c_r (Ny each)
  double* part_a[kMaxChains];       // per-tile partial projections (ring_partial_count)
  double* part_c[kMaxChains];
  const double* edge;               // sx1, sxN, sy1, syN
  const double* pw;                 // S_y[:, ny/4], S_y[:, 3ny/4], S_x[nx/4, :], S_x[3nx/4, :]
  const double* Mx;                 // Nx x Nx ring operator (row-major)
  const double* My;                 // Ny x Ny
  const double* skip;               // launch context flag (filled by the launchers)
};
// E = D . (F + R) of every chain and, with do_proj, the four projections (vec): a sweep
// and a reduction launch.
hipError_t launch_ring_sweep(const RingChainArgs& a, int nchain, bool do_proj, hipStream_t st);
// fac_out = (M_x a_b, M_x a_t, M_y c_l, M_y c_r).
hipError_t launch_ring_factor(const RingChainArgs& a, int nchain, hipStream_t st);
size_t ring_vector_count(int Nx, int Ny);
size_t ring_partial_count(int Nx, int Ny);
// leading part_a share of the partials (part_c follows it)
size_t ring_part_a_count(int Nx, int Ny);
// Input plane (N x N, k contiguous per column) whose sine pass along k is the ring operator
// M_n (row-major): projection at level l -> factor at level l-1, for an axis of n cells.
void build_ring_operator_input(int n, double* B);

// Sine tables for a DST-I of length N (N+1 even). Host builds, returns element counts.
void build_sine_tables_f64(int N, double* So, double* Se);
// double-fold tables of the length-N transform (N + 1 a multiple of 4): three Qp x Qp
// blocks, Qp = Q = (N + 1) / 4 rounded up to a multiple of 64, zero beyond each table
__host__ __device__ inline int quarter_pitch(int Q) { return (Q + 63) / 64 * 64; }
size_t sine_table_quarter_count(int N);
void build_sine_tables_f64q(int N, double* Sq);
// Split bf16 hi/lo tables, laid out for the MFMA B operand (k contiguous per column).
void build_sine_tables_bf16(int N, uint16_t* So_hi, uint16_t* So_lo, uint16_t* Se_hi,
                            uint16_t* Se_lo);
size_t sine_table_odd_count(int N);
size_t sine_table_even_count(int N);
size_t sine_table_bf16_odd_count(int N);
size_t sine_table_bf16_even_count(int N);
// Split bf16 double-fold tables of the length-N transform (N + 1 a multiple of 4):
// [T1 | T2 | T4], each hi then lo, Qm x Qm with k contiguous per output row m (Qm = Q
// rounded up to 32, zero beyond each table); returns element counts.
size_t sine_table_bf16q_count(int N);
// the mixed double fold runs where M = N + 1 is a multiple of 16 (Q a multiple of 4)
inline bool mixed_quarter_ok(int N) { return (N + 1) % 16 == 0; }
void build_sine_tables_bf16q(int N, uint16_t* T);

// base_flux! of one step (fluid.jl:22-67): the flux of the freestream seen in the grid's
// frame. Static grid: uniform (ux, uy) h_l per level. Moving grid: per face,
//   q0 = h_l Rx[:, c] . ((uinf - u0) - Rv ((x, y) - center))
// with Rx the grid's orientation and Rv = angular_vel [-sin -cos; cos -sin]
// (evaluated by base_flux_u / base_flux_v in ibpm_device.h).
struct BaseFlux {
  int moving;
  // when set (batched steps): this step's values on the device, in the order ux, uy, cx,
  // cy, vx, vy, rx11, rx21, rx12, rx22, rv11, rv12, rv21, rv22; q0u = ux hc, q0v = uy hc
  const double* dyn;
  double q0u[16], q0v[16];         // static grid
  double hc[16], x0[16], y0[16];   // moving grid: step and lower-left corner per level
  double ux, uy;                   // freestream velocity (global frame)
  double cx, cy, vx, vy;           // GridVelocity: centre and velocity
  double rx11, rx21, rx12, rx22;   // Rx = [rx11 rx12; rx21 rx22]
  double rv11, rv12, rv21, rv22;   // Rv
};

// ---- stencil kernels (stencil_kernels.hip) ----
// Nonlinear term and CNAB right-hand side for every level in one launch
// (fluid.jl:441-455 and :467-497).
struct NonlinArgs {
  GridDims g;
  int nlevel;
  const double* q;       // nq x L
  const double* psi;     // nG x L; when set, the fluxes are curl(psi) formed in the kernel
                         // (q is not read)
  const double* G;       // nG x L (state circulation)
  double* nl_new;        // nG x L, written
  const double* nl_old;  // nG x L
  double* rhs;           // nG x L, written
  BaseFlux bf;           // base flux of the step
  double inv_hc2[16];    // 1/hc^2 per level
  double cA[16];         // dt / (2 Re hc^2): A = I - cA*K
  double fac_bc[16];     // 0.25*dt/(Re hc^2)
  double dt, beta1, beta2;
};
hipError_t launch_nonlin_rhs(const NonlinArgs& a, hipStream_t st);

// coarsify! (fluid.jl:208-228) from level `fine` into `coarse`; when `outer` is not null
// the points outside the centre region are copied from it (fused state copy).
// boxes of the box updates, per level (0-based vertex offsets and sizes)
struct BoxSet {
  int i0[kMaxPlanes], j0[kMaxPlanes], w[kMaxPlanes], h[kMaxPlanes];
};
// coarsify! on box l of every level l >= 1 only, in place in G (L planes), the values it
// overwrites saved at the same offsets of `saved`
hipError_t launch_coarsify_boxes(GridDims g, int L, double* G, double* saved, const BoxSet& b,
                                 hipStream_t st);
hipError_t launch_coarsify(GridDims g, double* coarse, const double* fine, const double* outer,
                           hipStream_t st);

// curl! with boundary streamfunction (fluid.jl:230-276) for every level in one launch.
hipError_t launch_curl_all(GridDims g, int nlevel, double* q, const double* psi, hipStream_t st);

// r = b - (a*I + c*K) x  (5-point K = rot(curl), zero Dirichlet) for nplanes planes.
struct ResidualArgs {
  GridDims g;
  int nplanes;
  const double* b[16];
  const double* x[16];
  void* r[16];
  double a[16], c[16];
  const double* ring[16];  // vort2flux boundary values (added to b), may be null
  const double* xbase[16]; // when set, x stands for xbase - x (solution stored as base - x)
  int r_f32;
  const double* skip;      // launch context flag (filled by the launcher)
};
hipError_t launch_residual(const ResidualArgs& a, hipStream_t st);

// ---- coupling kernels (coupling_kernels.hip) ----
struct BodyDev {
  int npanel;
  int nbody;
  double x0, y0, h;
  int nx, ny;
  double* xb;         // 2n
  double* ub;         // 2n
  const double* xb_ref;  // 2n reference positions
  const int* body_of;    // n
  const int* moving;     // nbody
  const double* motion;  // 6 x nbody
  int* coords;        // 2n (ix, iy), 1-based
  double* w;          // n x 98
  int* perm;          // n, panels sorted by ix
  int* xstart;        // nx + 3
  int* ywin;          // 2 x (nx + 1): rows of vertex column I that panels can reach
  int* err;           // first offending panel + 1, 0 if none
};
hipError_t launch_body_update(const BodyDev& b, int only_moving, hipStream_t st);
hipError_t launch_body_sort(const BodyDev& b, hipStream_t st);

// F = E * (curl(psi1) + q0) - ub*h (couple_surface_rigid's right-hand side), and E of an
// explicit flux plane when q1 is given instead of psi.
struct GatherArgs {
  GridDims g;
  int npanel;
  int nrhs;
  const int* coords;
  const double* w;
  const double* psi1[64];   // level-1 streamfunction per rhs (curl on the fly)
  const double* psi2[64];   // level-2 streamfunction for boundary values (null: zero)
  double* out[64];          // 2n per rhs
  BaseFlux bf;              // base flux (level 1 used; moving = 0 and zeros for none)
  const double* ub;         // subtract ub*h if not null
  double h;
  const double* skip;       // launch context flag (filled by the launcher)
};
hipError_t launch_gather(const GatherArgs& a, hipStream_t st);

// Gamma1 = rot(E^T f) on the full level-1 plane (coupling.jl:72-85 then fluid.jl:101-111).
struct ScatterArgs {
  GridDims g;
  int npanel;
  int nrhs;
  const int* coords;
  const double* w;
  const int* perm;
  const int* xstart;
  const int* ywin;    // [2I], [2I+1]: first and last row J of vertex column I with panels
  const double* f[64];
  void* out[64];
  int out_f32;
  int ri0, rj0, rw, rh;  // vertices written: [ri0, ri0 + rw) x [rj0, rj0 + rh) (0-based);
                         // rw = 0: the whole plane
  // static panels: each vertex's terms listed once (build_scatter_lists), vertex v =
  // (I-1) + (J-1) Nx owning entries rowptr[v] .. rowptr[v+1]-1, each a panel (ep) and its
  // weights (ew: w_u of rows J and J+1 at column I+1, w_v of row J+1 at columns I+1 and
  // I; zero for a term the panel does not reach), in the panel order of perm. Null: the
  // panels are searched by column (moving bodies).
  const int* rowptr;
  const int* ep;
  const double* ew;
  const double* skip;  // launch context flag (filled by the launcher)
};
// the vertex lists of ScatterArgs from the panels' cells, weights and sort order (host
// copies); rowptr: Nx Ny + 1 entries
void build_scatter_lists(int n, int Nx, int Ny, const int* coords, const int* perm, const double* w,
                         std::vector<int>& rowptr, std::vector<int>& ep, std::vector<double>& ew);
hipError_t launch_scatter_rot(const ScatterArgs& a, hipStream_t st);
// project_circ's first launch for static bodies, three independent jobs in one launch:
// rot(E^T f[0]) from the vertex lists on the box (nrhs 1), saved = G on that box (plane
// pitch Nx), and the forces (launch_forces with Ft = f[0]).
hipError_t launch_project_prologue(const ScatterArgs& a, const double* G, double* saved,
                                   double* fb, double* total, double h, double dt,
                                   double* host_total, hipStream_t st);
// (min ix, max ix, min iy, max iy) of the panels' flux cells into out[0..3]
hipError_t launch_body_bbox(const int* coords, int n, double* out, hipStream_t st);

// x = M b for a dense column-major m x m matrix; part: gemv_partial_count(m) doubles.
size_t gemv_partial_count(int m);
hipError_t launch_gemv(int m, const double* M, const double* b, double* x, double* part,
                       hipStream_t st);
// x = M b, M row-major m x m (one launch)
hipError_t launch_gemv_rows(int m, const double* M, const double* b, double* x, hipStream_t st);
// fb = Ft*h/dt and the total force (column sums), FP64; the total also to host_total
// (device address of mapped pinned host memory) when that is not null.
hipError_t launch_forces(int n, const double* Ft, double* fb, double* total, double h, double dt,
                         double* host_total, hipStream_t st);

// Conjugate gradient helpers (single work-group kernels, FP64 reductions).
struct CgDev {
  int n;
  double* x;
  double* r;
  double* u;
  double* c;
  const double* b;
  double* scal;  // [0] residual, [1] prev_residual, [2] tol, [3] iters, [4] converged,
                 // [5] ||b||, [6] rho (preconditioned)
  int maxiter;
  double reltol, abstol;
  int tol_mode;
  const double* skip;  // launch context flag (filled by the launchers)
};
hipError_t launch_cg_init(const CgDev& c, hipStream_t st);
// After cg_init of a captured step: scal[8] = scal[0] (the initial residual) and, with
// ws_scal, scal[7] = 1 when that residual is at most ws_scal[1] = ||r_e|| (the learned
// guess beats the extrapolated one), else 0
hipError_t launch_cg_note(double* scal, const double* ws_scal, hipStream_t st);
hipError_t launch_cg_step(const CgDev& c, hipStream_t st);
// Preconditioned CG: the dense inverse P (2n x 2n, column-major) of B at an earlier body
// configuration, applied in each body's frame (angle change motion[6b+4] - theta0[b]).
struct PcDev {
  int n;                 // points (vectors hold x then y: 2n)
  const double* P;
  const int* body_of;
  const double* motion;  // 6 per body: pos, vel, angle, angular velocity (this step)
  const double* theta0;  // per body: angle when P was assembled
  const double* skip;    // launch context flag (filled by the launcher)
};
// X = A^-1 by Gauss-Jordan with partial pivoting (A is destroyed); column-major m x m.
// rowk: 2m doubles; ipiv, pval, singular: one element each (singular = 1 on a zero pivot).
hipError_t launch_gauss_jordan_inverse(int m, double* A, double* X, double* rowk, int* ipiv,
                                       double* pval, int* singular, hipStream_t st);
hipError_t launch_pcg_start(const CgDev& c, hipStream_t st);
// u = z + beta u with z = R P R^T r; t, zr: 2n scratch; part: gemv_partial_count(2n)
hipError_t launch_pcg_precond_dir(const CgDev& c, const PcDev& p, double* t, double* zr,
                                  double* part, hipStream_t st);
hipError_t launch_pcg_step(const CgDev& c, hipStream_t st);

hipError_t launch_fill(double* p, size_t n, double v, hipStream_t st);
hipError_t launch_copy(double* dst, const double* src, size_t n, hipStream_t st);
// ---- red-black Gauss-Seidel Helmholtz solves (gs_kernels.hip) ----
// One launch runs nsweep <= kGsHalo half-sweeps of x = inv_d b + c (sum of the 4
// neighbours) on each plane, starting from xin (zero when null) with colour color0
// ((i + j) & 1 of the points updated first), and writes x, or base - x when base is set,
// to out (which must not alias b, xin or base).
constexpr int kGsHalo = 24;
struct GsPlane {
  const double* b;
  const double* xin;
  double* out;
  const double* base;
  double c, inv_d;
  int nsweep, color0;
  int H, ntx, nty;  // halo (nsweep rounded up to even) and tile grid (by the launcher)
};
struct GsBatch {
  GsPlane p[kMaxPlanes];
  int Nx, Ny, ntx, nty;  // region solved (x = 0 outside it); ntx, nty unused
  int ld;                // row pitch of every plane pointer (0: Nx)
  const double* skip;    // launch context flag (filled by the launcher)
};
hipError_t launch_gs_helmholtz(GsBatch& bt, int nplanes, hipStream_t st);

// coarsify! centre region (stencil_kernels.hip): coarse vertices ic0 + a, jc0 + c (1-based,
// a < ni, c < nj) take the full weighting of fine vertex (2 + 2a, 2 + 2c)
struct CoarsifyRange {
  int ic0, ni, jc0, nj;
};
__host__ __device__ inline CoarsifyRange coarsify_range(int nx, int ny) {
  CoarsifyRange r;
  r.ic0 = nx / 2 - nx / 4 + 1;
  r.jc0 = ny / 2 - ny / 4 + 1;
  const int fi = nx / 2 - 1, ci = 2 * (nx / 4) - 1, fj = ny / 2 - 1, cj = 2 * (ny / 4) - 1;
  r.ni = fi < ci ? fi : ci;
  r.nj = fj < cj ? fj : cj;
  if (r.ni < 0) r.ni = 0;
  if (r.nj < 0) r.nj = 0;
  return r;
}

// ---- FP64 matrix-core products (gemm_kernels.hip) ----
// C (+)= (A - A2) (B - B2), column-major, M x K times K x N; A2 / B2 may be null.
struct GemmJob {
  const double *A, *A2;
  const double *B, *B2;
  double* C;
  int lda, ldb, ldc;
  int M, N, K;
  int accumulate;  // 1: C += A B, 0: C = A B
  // mirror (B rows k of sine-matrix rows j = k + kpar, modulo 2, and columns n whose
  // mirror column mref - n has B(k, mref - n) = (-1)^j B(k, n)): only columns n < N are
  // formed, from the sums U over even j and V over odd j, C(:, n) (+)= U + V and
  // C(:, mref - n) (+)= U - V (once where the two columns are one)
  int mirror, mref, kpar;
  // mirrored A (amirror): A(m, k) = A[m + k lda] + asign A2[m - k lda], the second term left
  // out at k = amid (the middle column, its own mirror; -1: none), A2 pointing at the mirror
  // of column 0
  int amirror, amid;
  double asign;
};
struct GemmBatch {
  GemmJob j[kMaxPlanes];
  const double* skip;  // launch context flag (filled by the launcher)
  int njobs;                // filled by the launcher
  int lbase[kMaxPlanes];    // first tile slot of each job per XCD (filled by the launcher)
};
hipError_t launch_gemm_f64(const GemmBatch& gb, int njobs, hipStream_t st);
// out[m + n ldo] = sum over s < nslice (<= 16) of P[s stride + m + n ldp], in order of s
hipError_t launch_sum_slices(const double* P, size_t stride, int nslice, int ldp, int M, int N,
                             double* out, int ldo, hipStream_t st);
// the same for up to kMaxPlanes matrices in one launch
struct SliceSumJob {
  const double* P;
  size_t stride;
  int nslice, ldp, M, N;
  double* out;
  int ldo;
};
struct SliceSumBatch {
  SliceSumJob j[kMaxPlanes];
  const double* skip;  // launch context flag (filled by the launcher)
};
hipError_t launch_sum_slices_batch(const SliceSumBatch& sb, int njobs, hipStream_t st);

// d = a - b elementwise
hipError_t launch_sub(double* d, const double* a, const double* b, size_t n, hipStream_t st);
hipError_t launch_axpy_f32(double* x, const float* y, size_t n, hipStream_t st);

// ---- warm-start network (net_kernels.hip, net.hip) ----
constexpr int kNetFeatures = 10;  // per-point input channels
constexpr int kNetInPad = 32;     // input channels as stored (MFMA K granularity)
constexpr int kNetOutPad = 32;    // output channels as computed (2 used)

// Y (Rpad x Cout) = conv(X (Rpad x Cin), W (Cout x K x Cin)) over rows of equal seg.
struct ConvArgs {
  const uint16_t* X;  // bf16
  int Cin;
  const uint16_t* W;  // bf16
  int Cout;
  int K;
  const float* bias;      // Cout (CONV_RELU, CONV_LINEAR_F32)
  const int* seg;         // R sequence ids
  int R, Rpad;            // rows, rows rounded up to 64 (pad rows are written as 0)
  void* Y;                // bf16, or float for CONV_LINEAR_F32
  const uint16_t* mask;   // CONV_MASK: Rpad x Cout bf16, keep where > 0
};
enum ConvEpi : int { CONV_RELU = 0, CONV_LINEAR_F32 = 1, CONV_MASK = 2 };
hipError_t launch_conv(const ConvArgs& a, int epi, hipStream_t st);

// dW (Cout x K x Cin) and db (Cout), stored contiguously in `out`, from dY and X.
struct WgradArgs {
  const uint16_t* dY;
  int Cout;
  const uint16_t* X;
  int Cin;
  const int* seg;
  int R, K;
  int rows_per_chunk;  // multiple of 32
  float* part;         // chunks x (Cout*K*Cin + Cout)
};
int wgrad_chunks(int R, int rows_per_chunk);
hipError_t launch_wgrad(const WgradArgs& a, float* out, hipStream_t st);

int mse_blocks(int Rpad, int Cp);
hipError_t launch_mse(const float* y, int Cp, const float* tgt, int R, int Rpad, uint16_t* dY,
                      double* partial, double* loss_out, hipStream_t st);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, size_t n, float lr, float b1,
                       float b2, float eps, int t, hipStream_t st);
hipError_t launch_cast_weights(const float* w, int Cout, int K, int Cin, uint16_t* wb,
                               uint16_t* wt, hipStream_t st);

struct GatherBatchArgs {
  const float* feat;   // training set: capacity x stride x kNetFeatures
  const float* tgt;    // capacity x stride x 2
  const int* body;     // capacity x stride
  const int* counts;   // rows per sample
  int stride;          // rows reserved per sample (max_points)
  const int* idx;      // S sample indices
  const int* off;      // S+1 row offsets
  int S, R, Rpad;
  float inv_scale;     // 1 / output scale
  uint16_t* A0;        // Rpad x kNetInPad
  float* tgt_out;      // R x 2
  int* seg;            // Rpad
};
hipError_t launch_gather_batch(const GatherBatchArgs& a, int max_points, hipStream_t st);
hipError_t launch_pack_rows(const float* x, const int* seg_in, int R, int Rpad, uint16_t* A0,
                            int* seg, hipStream_t st);
hipError_t launch_sumsq(const float* x, size_t n, double* partial, int nblocks, hipStream_t st);

// Solver side of the warm start.
struct WsArgs {
  int n;                  // body points
  const double* b;        // CG right-hand side (2n)
  const double* Bxe;      // B x_e (2n)
  const double* xb;       // positions (2n)
  const double* ub;       // velocities (2n)
  const int* body_of;     // n
  double x0, y0, h;       // level-1 origin and step
  const double* xe;       // extrapolated guess
  const double* F;        // converged F~b (record)
  double* scal;           // [0] rms(r_e), [1] ||r_e||, [2] ||b||
  float* feat;            // n x kNetFeatures (written)
  uint16_t* A0;           // Rpad x kNetInPad (written)
  int* seg;               // Rpad (written)
  int Rpad;
};
hipError_t launch_ws_extrap(int n2, const double* F1, const double* F2, const double* F3, int nhist,
                            double* xe, hipStream_t st);
hipError_t launch_ws_features(const WsArgs& a, hipStream_t st);
hipError_t launch_ws_apply(int n, const double* xe, const float* y, int Cp, const double* scal,
                           float out_scale, double* x0, hipStream_t st);
hipError_t launch_ws_record(const WsArgs& a, int stride, float* ds_feat, float* ds_tgt, int* ds_body,
                            hipStream_t st);

// Network entry points used by a solver step (net.hip); all run on the solver's stream.
int net_device(const ::ibpm_net* net);
int net_max_points(const ::ibpm_net* net);
// Features of the step into the network's input rows (fills a.feat, a.A0, a.seg, a.Rpad).
int net_features(::ibpm_net* net, WsArgs& a, hipStream_t st);
// x0 = x_e + rms(r_e) * scale * net(features) after net_features.
int net_predict(::ibpm_net* net, const WsArgs& a, double* x0, hipStream_t st);
// Append (features, (F - x_e) / rms(r_e)) to the training set after net_features.
int net_record(::ibpm_net* net, const WsArgs& a, hipStream_t st);

}  // namespace ibpm
