// aph_poisson.hip -- Poisson self-gravity (src/gravity/gsolvers/poisson_gravity.hpp) on a
// 3D mesh whose outer faces are all physical boundaries: a uniform lattice (below) or a
// multilevel mesh (the "multilevel meshes" section: the composite operator as CSR,
// BiCGSTAB with a multigrid V-cycle over coarser meshes of the tree).
//
// Per call of the solver's tasks:
//   k_pmom_block, k_pmom_final  ComputeMPCoeff + ScaleMPCoeff (:430-500): the nine
//                               multipole moments of rho V, per block then over blocks in
//                               a fixed order (the same sums every run)
//   k_prhs                      ComputeRhs (:502-589): 4 pi G rho minus 2 phi_b / dx^2 on
//                               every domain face, gathered into one lattice array
//   the solve                   the discrete problem of PoissonEquation (:31-388) with
//                               D = 1: conjugate gradients on -A (symmetric positive
//                               definite: the 7-point Laplacian with phi = 0 on the domain
//                               faces, plus alpha), preconditioned by one geometric
//                               multigrid V-cycle (cell averages down, constant values up,
//                               two weighted-Jacobi sweeps each way), to the relative
//                               residual tolerance.  The reference reaches the same
//                               discrete solution with Parthenon's MG / BiCGSTAB.
//   k_pfield                    ComputeGravityVector (:591-625) with MultipoleGravityBC's
//                               ghost values (gravity.cpp:37-104): g = -grad phi.
// Scalars of the iteration live on the device; once converged every later kernel of the
// iteration returns at once, and the host reads the flag every few iterations.
#include <climits>
#include <type_traits>

#include "aph_device.hpp"
#include "aph_poisson_tree.hpp"
#include "aph_runtime.hpp"

namespace {
constexpr int kMaxLevels = 10;
constexpr int kRedBlocks = 1024;  // workgroups of a reduction's first pass
}  // namespace

struct aph_poisson {
  const aph_blockset *bs = nullptr;
  int n[3] = {0, 0, 0};  // lattice cells along x, y, z
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, dx[3] = {0, 0, 0};
  double G = 0.0, alpha = 0.0, tol = 1e-8;
  int max_it = 0;
  int nlev = 0;
  int ln[kMaxLevels][3] = {};
  int32_t *d_boff = nullptr;  // per block: its lattice offset (x, y, z)
  double *x = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *e[kMaxLevels] = {}, *rl[kMaxLevels] = {}, *t[kMaxLevels] = {};
  double *part = nullptr;  // reduction partials
  double *sc = nullptr;    // device scalars (see Sc)
  double *coef = nullptr;  // nine multipole coefficients
  double *bmom = nullptr;  // per-block moments
  double *h_sc = nullptr;  // pinned host copy of the scalars
  // multilevel mesh (aph_poisson_create_tree): the operator as CSR over the leaf cells,
  // BiCGSTAB preconditioned by the multigrid V-cycle below (d_dinv: the inverse diagonal,
  // which level 0's smoother uses, and the Jacobi preconditioner when no coarser level
  // exists)
  bool tree = false;
  int64_t nt = 0, ncb = 0;                  // leaf cells, cells per block
  int64_t *d_ptr = nullptr, *d_nbptr = nullptr;
  int32_t *d_col = nullptr, *d_nbcol = nullptr, *d_wall = nullptr;
  double *d_val = nullptr, *d_nbval = nullptr, *d_dinv = nullptr;
  double *rh = nullptr, *v = nullptr, *y = nullptr, *s = nullptr, *tv = nullptr;
  // multigrid preconditioner of the multilevel solve: level 0 is the operator above,
  // levels 1.. the coarser meshes (aph_poisson_tree.hpp); empty = the Jacobi preconditioner
  struct MgLevel {
    int64_t n = 0;
    int64_t *ptr = nullptr, *rptr = nullptr;
    int32_t *col = nullptr, *cmap = nullptr, *rcol = nullptr;
    double *val = nullptr, *dinv = nullptr, *rval = nullptr;
    double *r = nullptr, *x = nullptr, *t = nullptr;
  };
  std::vector<MgLevel> mg;
  double *d_cinv = nullptr;  // the coarsest level's inverse (dense, row-major), or NULL
};

namespace aph {
namespace {

constexpr int kThreads = 256;
constexpr double kOmega = 0.8;  // weighted Jacobi (two sweeps before and after)
constexpr int kCoarseSweeps = 40;

// device scalars (conjugate gradients; BiCGSTAB reuses RR, BB, ALPHA, DONE and ITERS and
// keeps its own in the other slots)
enum Sc { RZ = 0, PQ, RR, BB, ALPHA, BETA, RZN, DONE, ITERS, NSC };
enum ScBi { RV = PQ, OMEGA = BETA, TS = NSC, TT, FAILED, NSC_ALL };  // rho: RZ / RZN

struct Grid {
  int nx, ny, nz;
  double ix2, iy2, iz2;  // 1/dx^2 per axis
  double alpha;
};

__device__ __forceinline__ bool finished(const double *sc) { return sc[DONE] != 0.0; }

__device__ __forceinline__ double multipole(const double *c, double x, double y, double z) {
  const double r2 = x * x + y * y + z * z;
  if (r2 == 0.0) return 0.0;
  const double r = sqrt(r2);
  const double r3 = r * r2;
  const double r5 = r3 * r2;
  return c[0] / r + (c[1] * y + c[2] * z + c[3] * x) / r3 +
         (c[4] * x * y + c[5] * y * z + c[6] * (3.0 * z * z - r2) + c[7] * z * x +
          c[8] * 0.5 * (x * x - y * y)) /
             r5;
}

// L e = -(A e): sum over axes of (2 e - e_- - e_+) / dx^2 + alpha e, e_ghost = -e.
__device__ __forceinline__ double apply_at(const Grid &g, const double *e, int i, int j, int k,
                                           int64_t c) {
  const double ec = e[c];
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  const double xm = i > 0 ? e[c - 1] : -ec, xp = i < g.nx - 1 ? e[c + 1] : -ec;
  const double ym = j > 0 ? e[c - sy] : -ec, yp = j < g.ny - 1 ? e[c + sy] : -ec;
  const double zm = k > 0 ? e[c - sz] : -ec, zp = k < g.nz - 1 ? e[c + sz] : -ec;
  return (2.0 * ec - xm - xp) * g.ix2 + (2.0 * ec - ym - yp) * g.iy2 +
         (2.0 * ec - zm - zp) * g.iz2 + g.alpha * ec;
}

__device__ __forceinline__ double diag_at(const Grid &g, int i, int j, int k) {
  return (2.0 + (i == 0) + (i == g.nx - 1)) * g.ix2 + (2.0 + (j == 0) + (j == g.ny - 1)) * g.iy2 +
         (2.0 + (k == 0) + (k == g.nz - 1)) * g.iz2 + g.alpha;
}

__device__ __forceinline__ void split3(const Grid &g, int64_t c, int &i, int &j, int &k) {
  i = (int)(c % g.nx);
  const int64_t r = c / g.nx;
  j = (int)(r % g.ny);
  k = (int)(r / g.ny);
}

#define APH_GRID_LOOP(g, c)                                                              \
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x,                      \
               n_ = (int64_t)(g).nx * (g).ny * (g).nz;                                  \
       c < n_; c += (int64_t)gridDim.x * blockDim.x)

__global__ void __launch_bounds__(kThreads) k_apply(Grid g, const double *sc, const double *e,
                                                    double *out) {
  if (finished(sc)) return;
  APH_GRID_LOOP(g, c) {
    int i, j, k;
    split3(g, c, i, j, k);
    out[c] = apply_at(g, e, i, j, k, c);
  }
}

// first smoothing sweep from e = 0: e = omega r / diag
__global__ void __launch_bounds__(kThreads) k_jacobi0(Grid g, const double *sc, const double *r,
                                                      double *e) {
  if (finished(sc)) return;
  APH_GRID_LOOP(g, c) {
    int i, j, k;
    split3(g, c, i, j, k);
    e[c] = kOmega * r[c] / diag_at(g, i, j, k);
  }
}

__global__ void __launch_bounds__(kThreads) k_jacobi(Grid g, const double *sc, const double *r,
                                                     const double *e, double *out) {
  if (finished(sc)) return;
  APH_GRID_LOOP(g, c) {
    int i, j, k;
    split3(g, c, i, j, k);
    out[c] = e[c] + kOmega * (r[c] - apply_at(g, e, i, j, k, c)) / diag_at(g, i, j, k);
  }
}

// coarse residual: the mean of r - L e over each 2x2x2 family
__global__ void __launch_bounds__(kThreads) k_restrict(Grid fine, Grid coarse, const double *sc,
                                                       const double *r, const double *e,
                                                       double *rc) {
  if (finished(sc)) return;
  APH_GRID_LOOP(coarse, c) {
    int i, j, k;
    split3(coarse, c, i, j, k);
    double s = 0.0;
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
          const int64_t f = ((int64_t)fk * fine.ny + fj) * fine.nx + fi;
          s += r[f] - apply_at(fine, e, fi, fj, fk, f);
        }
    rc[c] = 0.125 * s;
  }
}

__global__ void __launch_bounds__(kThreads) k_prolong_add(Grid fine, Grid coarse,
                                                          const double *sc, const double *ec,
                                                          double *e) {
  if (finished(sc)) return;
  APH_GRID_LOOP(fine, c) {
    int i, j, k;
    split3(fine, c, i, j, k);
    e[c] += ec[((int64_t)(k >> 1) * coarse.ny + (j >> 1)) * coarse.nx + (i >> 1)];
  }
}

// first pass of a dot product: one partial per workgroup, each over a fixed index set
__global__ void __launch_bounds__(kThreads) k_dot(int64_t n, const double *sc, const double *a,
                                                  const double *b, double *part) {
  __shared__ double s[kThreads];
  if (finished(sc)) return;
  double v = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x)
    v += a[c] * b[c];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// second pass: the partials in a fixed order into scalar `slot`
__global__ void __launch_bounds__(kThreads) k_dot_final(int nparts, double *sc, int slot,
                                                        const double *part) {
  __shared__ double s[kThreads];
  if (finished(sc)) return;
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += part[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sc[slot] = s[0];
}

__global__ void k_cg_init(double *sc) {
  for (int i = 0; i < NSC; ++i) sc[i] = 0.0;
}

__global__ void k_cg_start(double *sc) {  // after bb: a zero right-hand side is solved
  if (sc[BB] == 0.0) sc[DONE] = 1.0;
}

__global__ void k_cg_alpha(double *sc) {
  if (finished(sc)) return;
  sc[ALPHA] = sc[RZ] / sc[PQ];
}

__global__ void __launch_bounds__(kThreads) k_cg_xr(int64_t n, const double *sc, double *x,
                                                    double *r, const double *p,
                                                    const double *q) {
  if (finished(sc)) return;
  const double a = sc[ALPHA];
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    x[c] += a * p[c];
    r[c] -= a * q[c];
  }
}

__global__ void k_cg_check(double *sc, double tol2) {
  if (finished(sc)) return;
  sc[ITERS] += 1.0;
  if (sc[RR] <= tol2 * sc[BB]) sc[DONE] = 1.0;
}

__global__ void k_cg_beta(double *sc) {
  if (finished(sc)) return;
  sc[BETA] = sc[RZN] / sc[RZ];
  sc[RZ] = sc[RZN];
}

__global__ void __launch_bounds__(kThreads) k_cg_p(int64_t n, const double *sc, const double *z,
                                                   double *p, int first) {
  if (finished(sc)) return;
  const double bta = first ? 0.0 : sc[BETA];
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x)
    p[c] = first ? z[c] : z[c] + bta * p[c];
}

// r = b, x = 0
__global__ void __launch_bounds__(kThreads) k_cg_reset(int64_t n, const double *a, double *out,
                                                       double *zero) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    out[c] = a[c];
    zero[c] = 0.0;
  }
}

// ComputeMPCoeff for one block: the nine moments of its interior (fixed order)
__global__ void __launch_bounds__(kThreads)
    k_pmom_block(const aph_block *__restrict__ blocks, int nvar, const double *field,
                 double dfloor, double *bmom) {
  __shared__ double s[9][kThreads];
  const aph_block &bk = blocks[blockIdx.x];
  const int N0 = bk.nx[0] + 2 * bk.ng, N1 = bk.nx[1] + 2 * bk.ng;
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t t = threadIdx.x; t < nint; t += blockDim.x) {
    const int i = (int)(t % bk.nx[0]), j = (int)((t / bk.nx[0]) % bk.nx[1]),
              k = (int)(t / ((int64_t)bk.nx[0] * bk.nx[1]));
    const int64_t cell = ((int64_t)(k + bk.ng) * N1 + (j + bk.ng)) * N0 + (i + bk.ng);
    double rho = field[(int64_t)nvar * bk.cell_base + cell];
    rho = (rho > dfloor) ? rho : dfloor;
    const double x = bk.xmin[0] + (i + 0.5) * bk.dx[0];
    const double y = bk.xmin[1] + (j + 0.5) * bk.dx[1];
    const double z = bk.xmin[2] + (k + 0.5) * bk.dx[2];
    const double r2 = x * x + y * y + z * z;
    const double sm = rho * bk.vol;
    m[0] += sm;
    m[1] += sm * y;
    m[2] += sm * z;
    m[3] += sm * x;
    m[4] += sm * x * y;
    m[5] += sm * y * z;
    m[6] += sm * (3.0 * z * z - r2);
    m[7] += sm * z * x;
    m[8] += sm * 0.5 * (x * x - y * y);
  }
  for (int n = 0; n < 9; ++n) s[n][threadIdx.x] = m[n];
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int n = 0; n < 9; ++n) s[n][threadIdx.x] += s[n][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 9) bmom[(int64_t)blockIdx.x * 9 + threadIdx.x] = s[threadIdx.x][0];
}

// the blocks' moments in block order, scaled (ScaleMPCoeff, :479-495)
__global__ void k_pmom_final(int nblocks, const double *bmom, double G, double *coef) {
  const int n = threadIdx.x;
  if (n >= 9) return;
  const double scale[9] = {-1.0, -1.0, -1.0, -1.0, -3.0, -3.0, -0.25, -3.0, -3.0};
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += bmom[(int64_t)b * 9 + n];
  coef[n] = s * scale[n] * G;
}

// ComputeRhs into the lattice, negated for the positive definite form -A
__global__ void __launch_bounds__(kThreads)
    k_prhs(const aph_block *__restrict__ blocks, int nvar, const double *field, double dfloor,
           const int32_t *__restrict__ boff, Grid g, double lox, double loy, double loz,
           double hix, double hiy, double hiz, double four_pi_g, const double *coef,
           double *b) {
  const aph_block &bk = blocks[blockIdx.y];
  const int N0 = bk.nx[0] + 2 * bk.ng, N1 = bk.nx[1] + 2 * bk.ng;
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % bk.nx[0]), j = (int)((t / bk.nx[0]) % bk.nx[1]),
              k = (int)(t / ((int64_t)bk.nx[0] * bk.nx[1]));
    const int64_t cell = ((int64_t)(k + bk.ng) * N1 + (j + bk.ng)) * N0 + (i + bk.ng);
    double rho = field[(int64_t)nvar * bk.cell_base + cell];
    rho = (rho > dfloor) ? rho : dfloor;
    const int gi = boff[3 * blockIdx.y] + i, gj = boff[3 * blockIdx.y + 1] + j,
              gk = boff[3 * blockIdx.y + 2] + k;
    const double x = bk.xmin[0] + (i + 0.5) * bk.dx[0];
    const double y = bk.xmin[1] + (j + 0.5) * bk.dx[1];
    const double z = bk.xmin[2] + (k + 0.5) * bk.dx[2];
    double v = four_pi_g * rho;
    if (gi == 0) v -= 2.0 * multipole(coef, lox, y, z) * g.ix2;
    if (gi == g.nx - 1) v -= 2.0 * multipole(coef, hix, y, z) * g.ix2;
    if (gj == 0) v -= 2.0 * multipole(coef, x, loy, z) * g.iy2;
    if (gj == g.ny - 1) v -= 2.0 * multipole(coef, x, hiy, z) * g.iy2;
    if (gk == 0) v -= 2.0 * multipole(coef, x, y, loz) * g.iz2;
    if (gk == g.nz - 1) v -= 2.0 * multipole(coef, x, y, hiz) * g.iz2;
    b[((int64_t)gk * g.ny + gj) * g.nx + gi] = -v;
  }
}

// g = -grad phi, centred, with the ghost value 2 phi_0(face) - phi(mirror) at the domain
// faces; phi and g back into the pack layout
__global__ void __launch_bounds__(kThreads)
    k_pfield(const aph_block *__restrict__ blocks, const int32_t *__restrict__ boff, Grid g,
             double lox, double loy, double loz, double hix, double hiy, double hiz,
             const double *coef, const double *phi_l, double *grav, double *phi) {
  const aph_block &bk = blocks[blockIdx.y];
  const int N0 = bk.nx[0] + 2 * bk.ng, N1 = bk.nx[1] + 2 * bk.ng, N2 = bk.nx[2] + 2 * bk.ng;
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % bk.nx[0]), j = (int)((t / bk.nx[0]) % bk.nx[1]),
              k = (int)(t / ((int64_t)bk.nx[0] * bk.nx[1]));
    const int gi = boff[3 * blockIdx.y] + i, gj = boff[3 * blockIdx.y + 1] + j,
              gk = boff[3 * blockIdx.y + 2] + k;
    const double x = bk.xmin[0] + (i + 0.5) * bk.dx[0];
    const double y = bk.xmin[1] + (j + 0.5) * bk.dx[1];
    const double z = bk.xmin[2] + (k + 0.5) * bk.dx[2];
    const int64_t c = ((int64_t)gk * g.ny + gj) * g.nx + gi;
    const double pc = phi_l[c];
    const double xm = gi > 0 ? phi_l[c - 1] : 2.0 * multipole(coef, lox, y, z) - pc;
    const double xp = gi < g.nx - 1 ? phi_l[c + 1] : 2.0 * multipole(coef, hix, y, z) - pc;
    const double ym = gj > 0 ? phi_l[c - sy] : 2.0 * multipole(coef, x, loy, z) - pc;
    const double yp = gj < g.ny - 1 ? phi_l[c + sy] : 2.0 * multipole(coef, x, hiy, z) - pc;
    const double zm = gk > 0 ? phi_l[c - sz] : 2.0 * multipole(coef, x, y, loz) - pc;
    const double zp = gk < g.nz - 1 ? phi_l[c + sz] : 2.0 * multipole(coef, x, y, hiz) - pc;
    const int64_t cell = ((int64_t)(k + bk.ng) * N1 + (j + bk.ng)) * N0 + (i + bk.ng);
    double *gp = grav + 3 * bk.cell_base + cell;
    gp[0] = -((xp - xm) / (2.0 * bk.dx[0]));
    gp[nc] = -((yp - ym) / (2.0 * bk.dx[1]));
    gp[2 * nc] = -((zp - zm) / (2.0 * bk.dx[2]));
    if (phi) phi[bk.cell_base + cell] = pc;
  }
}

Grid level_grid(const aph_poisson *ps, int l) {
  Grid g;
  g.nx = ps->ln[l][0];
  g.ny = ps->ln[l][1];
  g.nz = ps->ln[l][2];
  const double f = (double)(1 << l);
  g.ix2 = 1.0 / ((ps->dx[0] * f) * (ps->dx[0] * f));
  g.iy2 = 1.0 / ((ps->dx[1] * f) * (ps->dx[1] * f));
  g.iz2 = 1.0 / ((ps->dx[2] * f) * (ps->dx[2] * f));
  g.alpha = ps->alpha;
  return g;
}

int64_t cells(const Grid &g) { return (int64_t)g.nx * g.ny * g.nz; }

dim3 vec_grid(int64_t n) {
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads,
                                                               4096)));
}

void dot(aph_context *ctx, aph_poisson *ps, int64_t n, const double *a, const double *b,
         int slot) {
  const unsigned parts =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, kRedBlocks));
  hipLaunchKernelGGL(k_dot, dim3(parts), dim3(kThreads), 0, ctx->stream, n, ps->sc, a, b,
                     ps->part);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kThreads), 0, ctx->stream, (int)parts, ps->sc,
                     slot, ps->part);
}

// e_l = V-cycle(r_l) on level l
void vcycle(aph_context *ctx, aph_poisson *ps, int l, const double *r, double *e) {
  const Grid g = level_grid(ps, l);
  const int64_t n = cells(g);
  const dim3 gr = vec_grid(n);
  double *t = ps->t[l];
  if (l == ps->nlev - 1) {  // coarsest: Jacobi sweeps
    hipLaunchKernelGGL(k_jacobi0, gr, dim3(kThreads), 0, ctx->stream, g, ps->sc, r, e);
    for (int s = 1; s < kCoarseSweeps; s += 2) {
      hipLaunchKernelGGL(k_jacobi, gr, dim3(kThreads), 0, ctx->stream, g, ps->sc, r, e, t);
      hipLaunchKernelGGL(k_jacobi, gr, dim3(kThreads), 0, ctx->stream, g, ps->sc, r, t, e);
    }
    return;
  }
  // pre-smoothing (kSweeps = 2: e -> t -> e)
  hipLaunchKernelGGL(k_jacobi0, gr, dim3(kThreads), 0, ctx->stream, g, ps->sc, r, t);
  hipLaunchKernelGGL(k_jacobi, gr, dim3(kThreads), 0, ctx->stream, g, ps->sc, r, t, e);
  const Grid gc = level_grid(ps, l + 1);
  hipLaunchKernelGGL(k_restrict, vec_grid(cells(gc)), dim3(kThreads), 0, ctx->stream, g, gc,
                     ps->sc, r, e, ps->rl[l + 1]);
  vcycle(ctx, ps, l + 1, ps->rl[l + 1], ps->e[l + 1]);
  hipLaunchKernelGGL(k_prolong_add, gr, dim3(kThreads), 0, ctx->stream, g, gc, ps->sc,
                     ps->e[l + 1], e);
  // post-smoothing (the adjoint order of the pre-smoothing: e -> t -> e)
  hipLaunchKernelGGL(k_jacobi, gr, dim3(kThreads), 0, ctx->stream, g, ps->sc, r, e, t);
  hipLaunchKernelGGL(k_jacobi, gr, dim3(kThreads), 0, ctx->stream, g, ps->sc, r, t, e);
}

// ------------------------------------------------------------ multilevel meshes
// The composite operator of a multilevel mesh as CSR over the leaf cells (block order,
// cell (i, j, k) of block b at b ncb + (k ny + j) nx + i), assembled on the host from the
// tree.  BiCGSTAB (the operator is not symmetric once prolongation enters it) with x0 = 0,
// right-preconditioned by one multigrid V-cycle over coarser meshes of the same problem
// (mg_vcycle below; the reference's MGBiCGSTAB), or by the inverse diagonal where the mesh
// has no coarser level.  Five launches per iteration besides the two cycles, one per
// global dependency:
//   k_bi_p   the residual test and rho = (rh, r) from k_bi_x's partials; p (y = M^-1 p)
//   k_bi_av  v = A y, partials of (rh, v)
//   k_bi_s   alpha = rho / (rh, v); s = r - alpha v, x += alpha y (z = M^-1 s)
//   k_bi_at  t = A z, partials of (t, s), (t, t)
//   k_bi_x   omega = (t, s) / (t, t); x += omega z, r = s - omega t, partials of (r, r),
//            (rh, r)
// Every workgroup of a kernel sums the previous kernel's partials itself, in one fixed
// order (so all get the same bits), instead of a one-workgroup reduction launch; a scalar
// a kernel needs comes from its own sum or from an earlier kernel (workgroup 0 writes the
// scalars it forms, and rho alternates between two slots by iteration parity, so no kernel
// reads a slot it writes).  A matrix row takes kLanes lanes (rows hold ~8 entries), summed
// by a fixed shuffle tree.
struct Csr {
  const int64_t *ptr;
  const int32_t *col;
  const double *val;
};

constexpr int kLanes = 8;  // lanes per matrix row in the products

// the slot of rho for iterations of this parity
__device__ __forceinline__ int rho_slot(int it) { return (it & 1) ? RZN : RZ; }

__device__ __forceinline__ double row_dot(const Csr &A, int64_t r, const double *x) {
  double s = 0.0;
  for (int64_t q = A.ptr[r], e = A.ptr[r + 1]; q < e; ++q) s += A.val[q] * x[A.col[q]];
  return s;
}

// the sum over the workgroup of each of K per-thread values, in every thread: a butterfly
// within each wave (partners add the same two values, so every lane holds the same bits),
// then the waves' sums in wave order
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
  constexpr int kWaves = kThreads / 64;
  __shared__ double sh[K][kWaves];
  for (int k = 0; k < K; ++k) {
    double w = v[k];
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
    if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  for (int k = 0; k < K; ++k) {
    double t = sh[k][0];
    for (int i = 1; i < kWaves; ++i) t += sh[k][i];
    v[k] = t;
  }
  __syncthreads();  // sh may be written again by the next call in the kernel
}

// K per-thread values into part[k * kRedBlocks + blockIdx.x]
template <int K>
__device__ __forceinline__ void block_partials(double (&v)[K], double *part) {
  block_sum<K>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) part[k * kRedBlocks + blockIdx.x] = v[k];
}

// the sums of K partial arrays (nparts each), the same order in every workgroup; every
// thread gets them
template <int K>
__device__ __forceinline__ void sum_partials(const double *part, int nparts, double (&out)[K]) {
  for (int k = 0; k < K; ++k) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += part[k * kRedBlocks + i];
    out[k] = v;
  }
  block_sum<K>(out);
}

__device__ __forceinline__ bool lead() { return blockIdx.x == 0 && threadIdx.x == 0; }

#define APH_VEC_LOOP(n, c)                                                                 \
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (n);               \
       c += (int64_t)gridDim.x * blockDim.x)

// y = A x over kLanes lanes per row; calls f(row, value) on the row's first lane
template <class F>
__device__ __forceinline__ void spmv(int64_t n, const Csr &A, const double *x, F f) {
  const int lane = threadIdx.x & (kLanes - 1);
  const int64_t step = (int64_t)gridDim.x * blockDim.x / kLanes;
  for (int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanes; c < n;
       c += step) {
    double s = 0.0;
    for (int64_t q = A.ptr[c] + lane, e = A.ptr[c + 1]; q < e; q += kLanes)
      s += A.val[q] * x[A.col[q]];
    s += __shfl_xor(s, 4, kLanes);
    s += __shfl_xor(s, 2, kLanes);
    s += __shfl_xor(s, 1, kLanes);
    if (lane == 0) f(c, s);
  }
}

// x = 0, r = rh = b, p = v = 0; partials (b, b) twice (the first iteration's (r, r) and
// (rh, r)); the scalars' starting values
__global__ void __launch_bounds__(kThreads) k_bi_init(int64_t n, const double *b, double *x,
                                                      double *r, double *rh, double *p,
                                                      double *v, double *sc,
                                                      double *part_x) {
  double a[2] = {0.0, 0.0};
  APH_VEC_LOOP(n, c) {
    const double bc = b[c];
    x[c] = 0.0;
    r[c] = bc;
    rh[c] = bc;
    p[c] = 0.0;
    v[c] = 0.0;
    a[0] += bc * bc;
  }
  a[1] = a[0];
  if (lead()) {
    for (int i = 0; i < NSC_ALL; ++i) sc[i] = 0.0;
    sc[ALPHA] = sc[OMEGA] = 1.0;
    sc[rho_slot(1)] = 1.0;  // rho before the first iteration
  }
  block_partials<2>(a, part_x);
}

// iteration it: the residual test on r, rho = (rh, r), beta; p = r + beta (p - omega v),
// y = p / diag.  check_only: the test alone (after the last iteration).
__global__ void __launch_bounds__(kThreads)
    k_bi_p(int it, int check_only, int64_t n, int nparts, double tol2, double *sc,
           const double *part_x, const double *r, const double *v, const double *dinv,
           double *p, double *y) {
  if (finished(sc)) return;
  double s[2];
  sum_partials<2>(part_x, nparts, s);
  const double rr = s[0], rho = s[1];
  const double bb = it == 0 ? rr : sc[BB];
  if (lead()) {
    sc[RR] = rr;
    if (it == 0) sc[BB] = rr;
  }
  if (rr <= tol2 * bb) {  // converged (a zero right-hand side at once)
    if (lead()) sc[DONE] = 1.0;
    return;
  }
  if (check_only) return;
  const double om = sc[OMEGA];
  if (rho == 0.0 || om == 0.0) {  // breakdown
    if (lead()) {
      sc[FAILED] = 1.0;
      sc[DONE] = 1.0;
    }
    return;
  }
  const double beta = (rho / sc[rho_slot(it + 1)]) * (sc[ALPHA] / om);
  if (lead()) sc[rho_slot(it)] = rho;
  APH_VEC_LOOP(n, c) {
    const double pc = r[c] + beta * (p[c] - om * v[c]);
    p[c] = pc;
    if (dinv) y[c] = pc * dinv[c];  // else the multigrid cycle forms y
  }
}

// v = A y; partials (rh, v)
__global__ void __launch_bounds__(kThreads) k_bi_av(int64_t n, Csr A, const double *sc,
                                                    const double *y, const double *rh,
                                                    double *v, double *part_v) {
  if (finished(sc)) return;
  double a[1] = {0.0};
  spmv(n, A, y, [&](int64_t c, double vc) {
    v[c] = vc;
    a[0] += rh[c] * vc;
  });
  block_partials<1>(a, part_v);
}

// alpha = rho / (rh, v); s = r - alpha v, x += alpha y, z = s / diag
__global__ void __launch_bounds__(kThreads)
    k_bi_s(int it, int64_t n, int nparts, double *sc, const double *part_v, const double *r,
           const double *v, const double *y, const double *dinv, double *x, double *s,
           double *z) {
  if (finished(sc)) return;
  double rv[1];
  sum_partials<1>(part_v, nparts, rv);
  if (rv[0] == 0.0) {  // breakdown
    if (lead()) {
      sc[FAILED] = 1.0;
      sc[DONE] = 1.0;
    }
    return;
  }
  const double al = sc[rho_slot(it)] / rv[0];
  if (lead()) {
    sc[RV] = rv[0];
    sc[ALPHA] = al;
  }
  APH_VEC_LOOP(n, c) {
    const double sv = r[c] - al * v[c];
    s[c] = sv;
    x[c] += al * y[c];
    if (dinv) z[c] = sv * dinv[c];  // else the multigrid cycle forms z
  }
}

// t = A z; partials (t, s), (t, t)
__global__ void __launch_bounds__(kThreads) k_bi_at(int64_t n, Csr A, const double *sc,
                                                    const double *z, const double *s,
                                                    double *t, double *part_t) {
  if (finished(sc)) return;
  double a[2] = {0.0, 0.0};
  spmv(n, A, z, [&](int64_t c, double tc) {
    t[c] = tc;
    a[0] += tc * s[c];
    a[1] += tc * tc;
  });
  block_partials<2>(a, part_t);
}

// omega = (t, s) / (t, t); x += omega z, r = s - omega t; partials (r, r), (rh, r)
__global__ void __launch_bounds__(kThreads)
    k_bi_x(int64_t n, int nparts, double *sc, const double *part_t, const double *z,
           const double *s, const double *t, const double *rh, double *x, double *r,
           double *part_x) {
  if (finished(sc)) return;
  double tt[2];
  sum_partials<2>(part_t, nparts, tt);
  if (tt[1] == 0.0) {  // t = A z = 0: s = 0, x is the solution
    if (lead()) {
      sc[RR] = 0.0;
      sc[ITERS] += 1.0;
      sc[DONE] = 1.0;
    }
    return;
  }
  const double om = tt[0] / tt[1];
  if (lead()) {
    sc[TS] = tt[0];
    sc[TT] = tt[1];
    sc[OMEGA] = om;
    sc[ITERS] += 1.0;
  }
  double a[2] = {0.0, 0.0};
  APH_VEC_LOOP(n, c) {
    x[c] += om * z[c];
    const double rc = s[c] - om * t[c];
    r[c] = rc;
    a[0] += rc * rc;
    a[1] += rh[c] * rc;
  }
  block_partials<2>(a, part_x);
}

// ---- the multigrid V-cycle over the coarser meshes (the tree cut one level at a time,
// then the base blocks with half the cells per axis; each the same problem re-assembled):
// two weighted-Jacobi sweeps before and after (w = 0.8), the residual's mean over each
// coarse cell down, the coarse correction added back cell by cell, the coarsest level
// solved with its inverse (or 40 sweeps when it is large).  Four launches per level: both
// sweeps from zero in one (the first is pointwise), the residual with its restriction,
// the correction with the first sweep after (also pointwise), the last sweep.  The cycle
// is a fixed linear map, as BiCGSTAB's preconditioner must be.

// sum_j A_cj f(j) over kLanes lanes per row; g(row, value) on the row's first lane
template <class F, class G>
__device__ __forceinline__ void spmv_f(int64_t n, const Csr &A, F f, G g) {
  const int lane = threadIdx.x & (kLanes - 1);
  const int64_t step = (int64_t)gridDim.x * blockDim.x / kLanes;
  for (int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanes; c < n;
       c += step) {
    double s = 0.0;
    for (int64_t q = A.ptr[c] + lane, e = A.ptr[c + 1]; q < e; q += kLanes)
      s += A.val[q] * f(A.col[q]);
    s += __shfl_xor(s, 4, kLanes);
    s += __shfl_xor(s, 2, kLanes);
    s += __shfl_xor(s, 1, kLanes);
    if (lane == 0) g(c, s);
  }
}

// two sweeps from zero: x1 = w D^-1 r, x = x1 + w D^-1 (r - A x1)
__global__ void __launch_bounds__(kThreads) k_mg_pre(int64_t n, Csr A, const double *dinv,
                                                     const double *sc, const double *r,
                                                     double *x) {
  if (finished(sc)) return;
  spmv_f(n, A, [&](int32_t j) { return kOmega * dinv[j] * r[j]; },
         [&](int64_t c, double ax) {
           const double x1 = kOmega * dinv[c] * r[c];
           x[c] = x1 + kOmega * dinv[c] * (r[c] - ax);
         });
}

// the coarse right-hand side: rc = mean over each coarse cell's fine cells of r - A x
__global__ void __launch_bounds__(kThreads) k_mg_restrict(int64_t nc, Csr R, Csr A,
                                                          const double *sc, const double *r,
                                                          const double *x, double *rc) {
  if (finished(sc)) return;
  const int lane = threadIdx.x & (kLanes - 1);
  const int64_t step = (int64_t)gridDim.x * blockDim.x / kLanes;
  for (int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kLanes; c < nc;
       c += step) {
    double v = 0.0;
    const int64_t q = R.ptr[c] + lane;
    if (q < R.ptr[c + 1]) {  // at most 2^3 fine cells
      const int32_t f = R.col[q];
      v = R.val[q] * (r[f] - row_dot(A, f, x));
    }
    v += __shfl_xor(v, 4, kLanes);
    v += __shfl_xor(v, 2, kLanes);
    v += __shfl_xor(v, 1, kLanes);
    if (lane == 0) rc[c] = v;
  }
}

// the coarse correction and the first sweep after: xn = x + ec[cmap],
// out = xn + w D^-1 (r - A xn)
__global__ void __launch_bounds__(kThreads)
    k_mg_post1(int64_t n, Csr A, const double *dinv, const double *sc, const double *r,
               const double *x, const int32_t *cmap, const double *ec, double *out) {
  if (finished(sc)) return;
  spmv_f(n, A, [&](int32_t j) { return x[j] + ec[cmap[j]]; },
         [&](int64_t c, double ax) {
           const double xn = x[c] + ec[cmap[c]];
           out[c] = xn + kOmega * dinv[c] * (r[c] - ax);
         });
}

// one sweep: out = t + w D^-1 (r - A t)
__global__ void __launch_bounds__(kThreads) k_mg_sweep(int64_t n, Csr A, const double *dinv,
                                                       const double *sc, const double *r,
                                                       const double *t, double *out) {
  if (finished(sc)) return;
  spmv_f(n, A, [&](int32_t j) { return t[j]; },
         [&](int64_t c, double ax) { out[c] = t[c] + kOmega * dinv[c] * (r[c] - ax); });
}

// the coarsest level solved: x = A^-1 r with the dense inverse, one wave per row
__global__ void __launch_bounds__(kThreads) k_mg_dense(int64_t n, const double *inv,
                                                       const double *sc, const double *r,
                                                       double *x) {
  if (finished(sc)) return;
  const int lane = threadIdx.x & 63;
  const int64_t step = (int64_t)gridDim.x * blockDim.x / 64;
  for (int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; c < n; c += step) {
    double v = 0.0;
    for (int64_t j = lane; j < n; j += 64) v += inv[c * n + j] * r[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) x[c] = v;
  }
}

constexpr int kCoarseDense = 512;   // largest coarsest level solved with its inverse
constexpr int kCoarseJacobi = 40;   // sweeps on a larger coarsest level

dim3 lane_grid(int64_t n, int lanes) {
  return dim3((unsigned)std::max<int64_t>(
      1, std::min<int64_t>((n * lanes + kThreads - 1) / kThreads, 4096)));
}

// x = M^-1 r on level l (level 0: the operator of the solve)
void mg_vcycle(aph_poisson *ps, hipStream_t st, size_t l, const double *r, double *x) {
  const aph_poisson::MgLevel &L = ps->mg[l];
  const Csr A{L.ptr, L.col, L.val};
  const dim3 bl(kThreads);
  if (l + 1 == ps->mg.size()) {  // the coarsest level
    if (ps->d_cinv) {
      hipLaunchKernelGGL(k_mg_dense, lane_grid(L.n, 64), bl, 0, st, L.n, ps->d_cinv, ps->sc, r,
                         x);
      return;
    }
    hipLaunchKernelGGL(k_mg_pre, lane_grid(L.n, kLanes), bl, 0, st, L.n, A, L.dinv, ps->sc, r,
                       x);
    for (int k = 2; k < kCoarseJacobi; k += 2) {
      hipLaunchKernelGGL(k_mg_sweep, lane_grid(L.n, kLanes), bl, 0, st, L.n, A, L.dinv,
                         ps->sc, r, x, L.t);
      hipLaunchKernelGGL(k_mg_sweep, lane_grid(L.n, kLanes), bl, 0, st, L.n, A, L.dinv,
                         ps->sc, r, L.t, x);
    }
    return;
  }
  const aph_poisson::MgLevel &C = ps->mg[l + 1];
  hipLaunchKernelGGL(k_mg_pre, lane_grid(L.n, kLanes), bl, 0, st, L.n, A, L.dinv, ps->sc, r, x);
  hipLaunchKernelGGL(k_mg_restrict, lane_grid(C.n, kLanes), bl, 0, st, C.n,
                     Csr{C.rptr, C.rcol, C.rval}, A, ps->sc, r, x, C.r);
  mg_vcycle(ps, st, l + 1, C.r, C.x);
  hipLaunchKernelGGL(k_mg_post1, lane_grid(L.n, kLanes), bl, 0, st, L.n, A, L.dinv, ps->sc, r,
                     x, C.cmap, C.x, L.t);
  hipLaunchKernelGGL(k_mg_sweep, lane_grid(L.n, kLanes), bl, 0, st, L.n, A, L.dinv, ps->sc, r,
                     L.t, x);
}

__device__ __forceinline__ double face_potential(const double *coef, const aph_block &bk,
                                                 int d, int side, double x, double y,
                                                 double z) {
  double p[3] = {x, y, z};
  p[d] = bk.xmin[d] + (side ? bk.nx[d] : 0) * bk.dx[d];
  return multipole(coef, p[0], p[1], p[2]);
}

// ComputeRhs over the leaf cells: 4 pi G rho minus 2 phi_b A_f / (dx V) on every face on
// the domain boundary
__global__ void __launch_bounds__(kThreads)
    k_trhs(const aph_block *__restrict__ blocks, int nvar, const double *field, double dfloor,
           const int32_t *__restrict__ wall, double four_pi_g, const double *coef,
           int64_t ncb, double *b) {
  const aph_block &bk = blocks[blockIdx.y];
  const int N0 = bk.nx[0] + 2 * bk.ng, N1 = bk.nx[1] + 2 * bk.ng;
  const int wm = wall[blockIdx.y];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncb;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % bk.nx[0]), j = (int)((t / bk.nx[0]) % bk.nx[1]),
              k = (int)(t / ((int64_t)bk.nx[0] * bk.nx[1]));
    const int64_t cell = ((int64_t)(k + bk.ng) * N1 + (j + bk.ng)) * N0 + (i + bk.ng);
    double rho = field[(int64_t)nvar * bk.cell_base + cell];
    rho = (rho > dfloor) ? rho : dfloor;
    double v = four_pi_g * rho;
    if (wm) {
      const double x = bk.xmin[0] + (i + 0.5) * bk.dx[0];
      const double y = bk.xmin[1] + (j + 0.5) * bk.dx[1];
      const double z = bk.xmin[2] + (k + 0.5) * bk.dx[2];
      const int ijk[3] = {i, j, k};
      for (int d = 0; d < 3; ++d)
        for (int side = 0; side < 2; ++side)
          if ((wm >> (2 * d + side) & 1) && ijk[d] == (side ? bk.nx[d] - 1 : 0))
            v -= 2.0 * face_potential(coef, bk, d, side, x, y, z) *
                 (bk.area[d] / (bk.dx[d] * bk.vol));
    }
    b[(int64_t)blockIdx.y * ncb + t] = v;
  }
}

// ComputeGravityVector: g = -grad phi, centred, the neighbours as the operator sees them,
// 2 phi_b - phi on the domain faces; phi and g into the pack layout
__global__ void __launch_bounds__(kThreads)
    k_tfield(const aph_block *__restrict__ blocks, const int32_t *__restrict__ wall, Csr NB,
             int64_t n, int64_t ncb, const double *coef, const double *x, double *grav,
             double *phi) {
  const aph_block &bk = blocks[blockIdx.y];
  const int N0 = bk.nx[0] + 2 * bk.ng, N1 = bk.nx[1] + 2 * bk.ng, N2 = bk.nx[2] + 2 * bk.ng;
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int wm = wall[blockIdx.y];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncb;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % bk.nx[0]), j = (int)((t / bk.nx[0]) % bk.nx[1]),
              k = (int)(t / ((int64_t)bk.nx[0] * bk.nx[1]));
    const int ijk[3] = {i, j, k};
    const int64_t c = (int64_t)blockIdx.y * ncb + t;
    const double pc = x[c];
    const double cx = bk.xmin[0] + (i + 0.5) * bk.dx[0];
    const double cy = bk.xmin[1] + (j + 0.5) * bk.dx[1];
    const double cz = bk.xmin[2] + (k + 0.5) * bk.dx[2];
    const int64_t cell = ((int64_t)(k + bk.ng) * N1 + (j + bk.ng)) * N0 + (i + bk.ng);
    double *gp = grav + 3 * bk.cell_base + cell;
    for (int d = 0; d < 3; ++d) {
      double nbv[2];
      for (int side = 0; side < 2; ++side)
        nbv[side] = ((wm >> (2 * d + side) & 1) && ijk[d] == (side ? bk.nx[d] - 1 : 0))
                        ? 2.0 * face_potential(coef, bk, d, side, cx, cy, cz) - pc
                        : row_dot(NB, (int64_t)(2 * d + side) * n + c, x);
      gp[d * nc] = -((nbv[1] - nbv[0]) / (2.0 * bk.dx[d]));
    }
    if (phi) phi[bk.cell_base + cell] = pc;
  }
}

int tree_pass(aph_context *ctx, aph_poisson *ps, const double *field, double dfloor,
              double *grav, double *phi) {
  const aph_blockset *bs = ps->bs;
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_pmom_block, dim3((unsigned)bs->nblocks), dim3(kThreads), 0, st, bs->d,
                     bs->nvar, field, dfloor, ps->bmom);
  hipLaunchKernelGGL(k_pmom_final, dim3(1), dim3(64), 0, st, bs->nblocks, ps->bmom, ps->G,
                     ps->coef);
  const dim3 cg((unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (ps->ncb + kThreads - 1) /
                                                                         kThreads)),
                (unsigned)bs->nblocks);
  hipLaunchKernelGGL(k_trhs, cg, dim3(kThreads), 0, st, bs->d, bs->nvar, field, dfloor,
                     ps->d_wall, 4.0 * M_PI * ps->G, ps->coef, ps->ncb, ps->b);
  const int64_t n = ps->nt;
  const int parts =
      (int)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, kRedBlocks));
  const int mparts = (int)std::max<int64_t>(
      1, std::min<int64_t>((n * kLanes + kThreads - 1) / kThreads, kRedBlocks));
  const dim3 gr((unsigned)parts), gm((unsigned)mparts), bl(kThreads);
  const Csr A{ps->d_ptr, ps->d_col, ps->d_val};
  const double tol2 = ps->tol * ps->tol;
  double *part_x = ps->part, *part_v = ps->part + 2 * kRedBlocks,
         *part_t = ps->part + 3 * kRedBlocks;
  hipLaunchKernelGGL(k_bi_init, gr, bl, 0, st, n, ps->b, ps->x, ps->r, ps->rh, ps->p, ps->v,
                     ps->sc, part_x);
  bool done = false;
  auto poll = [&]() -> int {  // later kernels are no-ops once the flag is set
    APH_HIP_CHECK(ctx, hipMemcpyAsync(ps->h_sc, ps->sc, sizeof(double) * NSC_ALL,
                                      hipMemcpyDeviceToHost, st));
    APH_HIP_CHECK(ctx, hipStreamSynchronize(st));
    done = ps->h_sc[DONE] != 0.0;
    return APH_OK;
  };
  const bool mg = ps->mg.size() > 1;
  const double *jac = mg ? nullptr : ps->d_dinv;  // the Jacobi preconditioner without levels
  const int poll_every = mg ? 2 : 4;
  for (int it = 0; it < ps->max_it && !done; ++it) {
    hipLaunchKernelGGL(k_bi_p, gr, bl, 0, st, it, 0, n, parts, tol2, ps->sc, part_x, ps->r,
                       ps->v, jac, ps->p, ps->y);
    if (mg) mg_vcycle(ps, st, 0, ps->p, ps->y);
    hipLaunchKernelGGL(k_bi_av, gm, bl, 0, st, n, A, ps->sc, ps->y, ps->rh, ps->v, part_v);
    hipLaunchKernelGGL(k_bi_s, gr, bl, 0, st, it, n, mparts, ps->sc, part_v, ps->r, ps->v,
                       ps->y, jac, ps->x, ps->s, ps->z);
    if (mg) mg_vcycle(ps, st, 0, ps->s, ps->z);
    hipLaunchKernelGGL(k_bi_at, gm, bl, 0, st, n, A, ps->sc, ps->z, ps->s, ps->tv, part_t);
    hipLaunchKernelGGL(k_bi_x, gr, bl, 0, st, n, mparts, ps->sc, part_t, ps->z, ps->s, ps->tv,
                       ps->rh, ps->x, ps->r, part_x);
    if (it % poll_every == poll_every - 1) {
      int rc = poll();
      if (rc) return rc;
    }
  }
  if (!done) {  // the residual test of the last iteration
    hipLaunchKernelGGL(k_bi_p, gr, bl, 0, st, ps->max_it, 1, n, parts, tol2, ps->sc, part_x,
                       ps->r, ps->v, ps->d_dinv, ps->p, ps->y);
    int rc = poll();
    if (rc) return rc;
  }
  const bool converged = done && ps->h_sc[FAILED] == 0.0;
  if (!converged)
    return set_error(ctx, APH_E_STATE,
                     "poisson gravity: %s (tolerance %g, %d iterations, relative residual %g)",
                     ps->h_sc[FAILED] != 0.0 ? "BiCGSTAB breakdown" : "no convergence",
                     ps->tol, (int)ps->h_sc[ITERS],
                     ps->h_sc[BB] > 0.0 ? sqrt(ps->h_sc[RR] / ps->h_sc[BB]) : 0.0);
  const Csr NB{ps->d_nbptr, ps->d_nbcol, ps->d_nbval};
  hipLaunchKernelGGL(k_tfield, cg, bl, 0, st, bs->d, ps->d_wall, NB, n, ps->ncb, ps->coef,
                     ps->x, grav, phi);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int check_lattice(aph_context *ctx, const aph_blockset *bs, aph_poisson *ps,
                  std::vector<int32_t> &boff) {
  if (bs->ndim != 3)
    return set_error(ctx, APH_E_INVALID, "aph_poisson_create: 3D block sets only");
  if (bs->nblocks < 1) return set_error(ctx, APH_E_INVALID, "aph_poisson_create: no blocks");
  const aph_block &b0 = bs->h[0];
  for (int d = 0; d < 3; ++d) {
    ps->dx[d] = b0.dx[d];
    ps->lo[d] = b0.xmin[d];
    ps->hi[d] = b0.xmin[d] + b0.nx[d] * b0.dx[d];
  }
  for (const aph_block &b : bs->h)
    for (int d = 0; d < 3; ++d) {
      if (b.nx[d] != b0.nx[d] || b.dx[d] != b0.dx[d])
        return set_error(ctx, APH_E_INVALID,
                         "aph_poisson_create: blocks of one size and spacing required");
      ps->lo[d] = std::min(ps->lo[d], b.xmin[d]);
      ps->hi[d] = std::max(ps->hi[d], b.xmin[d] + b.nx[d] * b.dx[d]);
    }
  int64_t tot = 1;
  for (int d = 0; d < 3; ++d) {
    ps->n[d] = (int)llround((ps->hi[d] - ps->lo[d]) / ps->dx[d]);
    tot *= ps->n[d];
  }
  if (tot != (int64_t)bs->nblocks * b0.nx[0] * b0.nx[1] * b0.nx[2])
    return set_error(ctx, APH_E_INVALID, "aph_poisson_create: the blocks must tile a box");
  boff.assign(3 * (size_t)bs->nblocks, 0);
  for (int b = 0; b < bs->nblocks; ++b) {
    const aph_block &k = bs->h[b];
    for (int d = 0; d < 3; ++d) {
      const int o = (int)llround((k.xmin[d] - ps->lo[d]) / ps->dx[d]);
      boff[3 * (size_t)b + d] = o;
      const int sdn = (d == 0) ? 1 : (d == 1 ? 3 : 9);
      // a block on a face of the box must have no neighbour beyond it (not periodic)
      if ((o == 0 && k.nbr[13 - sdn] >= 0) || (o + k.nx[d] == ps->n[d] && k.nbr[13 + sdn] >= 0))
        return set_error(ctx, APH_E_INVALID,
                         "aph_poisson_create: every face of the domain must be a physical "
                         "boundary (the periodic problem needs a mean density subtracted)");
    }
  }
  return APH_OK;
}

}  // namespace

// The solver's whole pass (PreComputeTasks then AddTasks) for a density field (conserved
// or primitive, floored at dfloor): grav (3 components) and, if given, phi.
const aph_blockset *poisson_blockset(const aph_poisson *ps) { return ps->bs; }

int poisson_pass(aph_context *ctx, aph_poisson *ps, const double *field, double dfloor,
                 double *grav, double *phi) {
  if (ps->tree) return tree_pass(ctx, ps, field, dfloor, grav, phi);
  const aph_blockset *bs = ps->bs;
  hipLaunchKernelGGL(k_pmom_block, dim3((unsigned)bs->nblocks), dim3(kThreads), 0,
                     ctx->stream, bs->d, bs->nvar, field, dfloor, ps->bmom);
  hipLaunchKernelGGL(k_pmom_final, dim3(1), dim3(64), 0, ctx->stream, bs->nblocks, ps->bmom,
                     ps->G, ps->coef);
  const Grid g = level_grid(ps, 0);
  const int64_t n = cells(g);
  const dim3 cg((unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (bs->max_interior + kThreads - 1) / kThreads)),
                (unsigned)bs->nblocks);
  hipLaunchKernelGGL(k_prhs, cg, dim3(kThreads), 0, ctx->stream, bs->d, bs->nvar, field,
                     dfloor, ps->d_boff, g, ps->lo[0], ps->lo[1], ps->lo[2], ps->hi[0],
                     ps->hi[1], ps->hi[2], 4.0 * M_PI * ps->G, ps->coef, ps->b);
  const dim3 gr = vec_grid(n);
  hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1), 0, ctx->stream, ps->sc);
  hipLaunchKernelGGL(k_cg_reset, gr, dim3(kThreads), 0, ctx->stream, n, ps->b, ps->r, ps->x);
  dot(ctx, ps, n, ps->b, ps->b, BB);
  hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(1), 0, ctx->stream, ps->sc);
  vcycle(ctx, ps, 0, ps->r, ps->z);
  dot(ctx, ps, n, ps->r, ps->z, RZ);
  hipLaunchKernelGGL(k_cg_p, gr, dim3(kThreads), 0, ctx->stream, n, ps->sc, ps->z, ps->p, 1);
  const double tol2 = ps->tol * ps->tol;
  bool converged = false;
  for (int it = 0; it < ps->max_it; ++it) {
    hipLaunchKernelGGL(k_apply, gr, dim3(kThreads), 0, ctx->stream, g, ps->sc, ps->p, ps->q);
    dot(ctx, ps, n, ps->p, ps->q, PQ);
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1), 0, ctx->stream, ps->sc);
    hipLaunchKernelGGL(k_cg_xr, gr, dim3(kThreads), 0, ctx->stream, n, ps->sc, ps->x, ps->r,
                       ps->p, ps->q);
    dot(ctx, ps, n, ps->r, ps->r, RR);
    hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(1), 0, ctx->stream, ps->sc, tol2);
    if (it % 4 == 3 || it == ps->max_it - 1) {  // read the flag now and then; later
                                                // kernels are no-ops once it is set
      APH_HIP_CHECK(ctx, hipMemcpyAsync(ps->h_sc, ps->sc, sizeof(double) * NSC,
                                        hipMemcpyDeviceToHost, ctx->stream));
      APH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      if (ps->h_sc[DONE] != 0.0) {
        converged = true;
        break;
      }
    }
    vcycle(ctx, ps, 0, ps->r, ps->z);
    dot(ctx, ps, n, ps->r, ps->z, RZN);
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1), 0, ctx->stream, ps->sc);
    hipLaunchKernelGGL(k_cg_p, gr, dim3(kThreads), 0, ctx->stream, n, ps->sc, ps->z, ps->p, 0);
  }
  if (!converged)
    return set_error(ctx, APH_E_STATE,
                     "poisson gravity: no convergence to %g in %d iterations (relative "
                     "residual %g)", ps->tol, ps->max_it,
                     ps->h_sc[BB] > 0.0 ? sqrt(ps->h_sc[RR] / ps->h_sc[BB]) : 0.0);
  hipLaunchKernelGGL(k_pfield, cg, dim3(kThreads), 0, ctx->stream, bs->d, ps->d_boff, g,
                     ps->lo[0], ps->lo[1], ps->lo[2], ps->hi[0], ps->hi[1], ps->hi[2],
                     ps->coef, ps->x, grav, phi);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // namespace aph

using namespace aph;

extern "C" {

int aph_poisson_create(aph_context *ctx, const aph_blockset *bs, double gravity_constant,
                       double diagonal_alpha, double tolerance, int32_t max_iterations,
                       aph_poisson **out) {
  if (!ctx) return APH_E_INVALID;
  if (!bs || !out || !(tolerance > 0.0) || max_iterations < 1 || diagonal_alpha < 0.0)
    return set_error(ctx, APH_E_INVALID,
                     "aph_poisson_create: tolerance > 0, max_iterations >= 1, alpha >= 0");
  aph_poisson *ps = new aph_poisson;
  std::vector<int32_t> boff;
  int rc = check_lattice(ctx, bs, ps, boff);
  if (rc) {
    delete ps;
    return rc;
  }
  ps->bs = bs;
  ps->G = gravity_constant;
  ps->alpha = diagonal_alpha;
  ps->tol = tolerance;
  ps->max_it = max_iterations;
  // levels: halve while every extent is even and at least 4
  int m[3] = {ps->n[0], ps->n[1], ps->n[2]};
  ps->nlev = 0;
  while (ps->nlev < kMaxLevels) {
    for (int d = 0; d < 3; ++d) ps->ln[ps->nlev][d] = m[d];
    ++ps->nlev;
    if (m[0] % 2 || m[1] % 2 || m[2] % 2 || m[0] < 4 || m[1] < 4 || m[2] < 4) break;
    for (int d = 0; d < 3; ++d) m[d] /= 2;
  }
  const int64_t n = (int64_t)ps->n[0] * ps->n[1] * ps->n[2];
  hipError_t e = hipMalloc(&ps->d_boff, sizeof(int32_t) * boff.size());
  for (double **v : {&ps->x, &ps->b, &ps->r, &ps->z, &ps->p, &ps->q})
    if (e == hipSuccess) e = hipMalloc(v, sizeof(double) * n);
  for (int l = 0; l < ps->nlev && e == hipSuccess; ++l) {
    const int64_t nl = (int64_t)ps->ln[l][0] * ps->ln[l][1] * ps->ln[l][2];
    if (l > 0) e = hipMalloc(&ps->e[l], sizeof(double) * nl);
    if (l > 0 && e == hipSuccess) e = hipMalloc(&ps->rl[l], sizeof(double) * nl);
    if (e == hipSuccess) e = hipMalloc(&ps->t[l], sizeof(double) * nl);
  }
  if (e == hipSuccess) e = hipMalloc(&ps->part, sizeof(double) * kRedBlocks);
  if (e == hipSuccess) e = hipMalloc(&ps->sc, sizeof(double) * NSC_ALL);
  if (e == hipSuccess) e = hipMalloc(&ps->coef, sizeof(double) * 9);
  if (e == hipSuccess) e = hipMalloc(&ps->bmom, sizeof(double) * 9 * bs->nblocks);
  if (e == hipSuccess) e = hipHostMalloc(&ps->h_sc, sizeof(double) * NSC_ALL);
  if (e == hipSuccess)
    e = hipMemcpy(ps->d_boff, boff.data(), sizeof(int32_t) * boff.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(ps->sc, 0, sizeof(double) * NSC_ALL);
  if (e == hipSuccess) e = hipMemset(ps->coef, 0, sizeof(double) * 9);
  if (e != hipSuccess) {
    aph_poisson_destroy(ctx, ps);
    return set_error(ctx, APH_E_HIP, "aph_poisson_create: %s", hipGetErrorString(e));
  }
  *out = ps;
  return APH_OK;
}

int aph_poisson_create_tree(aph_context *ctx, const aph_blockset *bs, const int32_t nbase[3],
                            const int32_t *level, const int64_t *loc, int32_t flux_correct,
                            int32_t prolongation, double gravity_constant,
                            double diagonal_alpha, double tolerance, int32_t max_iterations,
                            aph_poisson **out) {
  if (!ctx) return APH_E_INVALID;
  if (!bs || !out || !nbase || !level || !loc || !(tolerance > 0.0) || max_iterations < 1 ||
      diagonal_alpha < 0.0)
    return set_error(ctx, APH_E_INVALID,
                     "aph_poisson_create_tree: null argument, or not tolerance > 0, "
                     "max_iterations >= 1, alpha >= 0");
  if (bs->ndim != 3)
    return set_error(ctx, APH_E_INVALID, "aph_poisson_create_tree: 3D block sets only");
  PoissonTreeCsr m;
  char msg[256];
  int rc = poisson_tree_csr(bs->h.data(), bs->nblocks, nbase, level, loc, flux_correct,
                            prolongation, diagonal_alpha, m, msg, sizeof msg);
  if (rc) return set_error(ctx, rc, "aph_poisson_create_tree: %s", msg);
  std::vector<double> dinv((size_t)m.n);
  for (int64_t r = 0; r < m.n; ++r) {
    double dg = 0.0;
    for (int64_t q = m.rowptr[r]; q < m.rowptr[r + 1]; ++q)
      if (m.col[q] == r) dg = m.val[q];
    if (dg == 0.0)
      return set_error(ctx, APH_E_INVALID, "aph_poisson_create_tree: zero diagonal at cell %lld",
                       (long long)r);
    dinv[r] = 1.0 / dg;
  }
  aph_poisson *ps = new aph_poisson;
  ps->bs = bs;
  ps->G = gravity_constant;
  ps->alpha = diagonal_alpha;
  ps->tol = tolerance;
  ps->max_it = max_iterations;
  ps->tree = true;
  ps->nt = m.n;
  ps->ncb = m.n / bs->nblocks;
  const size_t n = (size_t)m.n;
  hipError_t e = hipSuccess;
  auto put = [&](auto **dst, const auto &src) {
    using T = typename std::decay_t<decltype(src)>::value_type;
    if (e == hipSuccess) e = hipMalloc(dst, sizeof(T) * std::max<size_t>(1, src.size()));
    if (e == hipSuccess && !src.empty())
      e = hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
  };
  put(&ps->d_ptr, m.rowptr);
  put(&ps->d_col, m.col);
  put(&ps->d_val, m.val);
  put(&ps->d_nbptr, m.nb_rowptr);
  put(&ps->d_nbcol, m.nb_col);
  put(&ps->d_nbval, m.nb_val);
  put(&ps->d_wall, m.wall);
  put(&ps->d_dinv, dinv);
  for (double **v : {&ps->x, &ps->b, &ps->r, &ps->rh, &ps->p, &ps->v, &ps->y, &ps->s, &ps->z,
                     &ps->tv})
    if (e == hipSuccess) e = hipMalloc(v, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(&ps->part, sizeof(double) * 5 * kRedBlocks);
  if (e == hipSuccess) e = hipMalloc(&ps->sc, sizeof(double) * NSC_ALL);
  if (e == hipSuccess) e = hipMalloc(&ps->coef, sizeof(double) * 9);
  if (e == hipSuccess) e = hipMalloc(&ps->bmom, sizeof(double) * 9 * bs->nblocks);
  if (e == hipSuccess) e = hipHostMalloc(&ps->h_sc, sizeof(double) * NSC_ALL);
  if (e == hipSuccess) e = hipMemset(ps->sc, 0, sizeof(double) * NSC_ALL);
  if (e == hipSuccess) e = hipMemset(ps->coef, 0, sizeof(double) * 9);
  // the multigrid levels: level 0 is the operator above, then the coarser meshes
  std::vector<PoissonMgLevel> lv;
  rc = poisson_tree_levels(bs->h.data(), bs->nblocks, nbase, level, loc, flux_correct,
                           prolongation, diagonal_alpha, kCoarseDense, lv, msg, sizeof msg);
  if (rc) {
    aph_poisson_destroy(ctx, ps);
    return set_error(ctx, rc, "aph_poisson_create_tree: %s", msg);
  }
  if (!lv.empty() && e == hipSuccess) {
    ps->mg.resize(lv.size() + 1);
    aph_poisson::MgLevel &m0 = ps->mg[0];
    m0.n = m.n;
    m0.ptr = ps->d_ptr;
    m0.col = ps->d_col;
    m0.val = ps->d_val;
    m0.dinv = ps->d_dinv;
    e = hipMalloc(&m0.t, sizeof(double) * n);
    for (size_t l = 0; l < lv.size() && e == hipSuccess; ++l) {
      const PoissonMgLevel &h = lv[l];
      aph_poisson::MgLevel &d = ps->mg[l + 1];
      d.n = h.n;
      std::vector<double> di((size_t)h.n);
      for (int64_t r = 0; r < h.n; ++r) {
        double dg = 0.0;
        for (int64_t q = h.rowptr[r]; q < h.rowptr[r + 1]; ++q)
          if (h.col[q] == r) dg = h.val[q];
        di[r] = dg != 0.0 ? 1.0 / dg : 0.0;
      }
      put(&d.ptr, h.rowptr);
      put(&d.col, h.col);
      put(&d.val, h.val);
      put(&d.dinv, di);
      put(&d.cmap, h.cmap);
      put(&d.rptr, h.rptr);
      put(&d.rcol, h.rcol);
      put(&d.rval, h.rval);
      for (double **v : {&d.r, &d.x, &d.t})
        if (e == hipSuccess) e = hipMalloc(v, sizeof(double) * (size_t)h.n);
    }
    const PoissonMgLevel &c = lv.back();
    if (e == hipSuccess && c.n <= kCoarseDense) {  // the coarsest level's inverse
      const int64_t nc = c.n;
      std::vector<double> a((size_t)(nc * nc), 0.0), inv((size_t)(nc * nc), 0.0);
      for (int64_t r = 0; r < nc; ++r) {
        for (int64_t q = c.rowptr[r]; q < c.rowptr[r + 1]; ++q) a[r * nc + c.col[q]] += c.val[q];
        inv[r * nc + r] = 1.0;
      }
      for (int64_t k = 0; k < nc; ++k) {  // Gauss-Jordan with partial pivoting
        int64_t pr = k;
        for (int64_t r = k + 1; r < nc; ++r)
          if (std::fabs(a[r * nc + k]) > std::fabs(a[pr * nc + k])) pr = r;
        if (a[pr * nc + k] == 0.0) {
          aph_poisson_destroy(ctx, ps);
          return set_error(ctx, APH_E_INVALID,
                           "aph_poisson_create_tree: singular coarsest multigrid level");
        }
        if (pr != k)
          for (int64_t j = 0; j < nc; ++j) {
            std::swap(a[k * nc + j], a[pr * nc + j]);
            std::swap(inv[k * nc + j], inv[pr * nc + j]);
          }
        const double piv = 1.0 / a[k * nc + k];
        for (int64_t j = 0; j < nc; ++j) {
          a[k * nc + j] *= piv;
          inv[k * nc + j] *= piv;
        }
        for (int64_t r = 0; r < nc; ++r) {
          if (r == k) continue;
          const double f = a[r * nc + k];
          if (f == 0.0) continue;
          for (int64_t j = 0; j < nc; ++j) {
            a[r * nc + j] -= f * a[k * nc + j];
            inv[r * nc + j] -= f * inv[k * nc + j];
          }
        }
      }
      put(&ps->d_cinv, inv);
    }
  }
  if (e != hipSuccess) {
    aph_poisson_destroy(ctx, ps);
    return set_error(ctx, APH_E_HIP, "aph_poisson_create_tree: %s", hipGetErrorString(e));
  }
  *out = ps;
  return APH_OK;
}

int aph_poisson_destroy(aph_context *ctx, aph_poisson *ps) {
  (void)ctx;
  if (!ps) return APH_OK;
  for (void *v : {(void *)ps->d_boff, (void *)ps->x, (void *)ps->b, (void *)ps->r,
                  (void *)ps->z, (void *)ps->p, (void *)ps->q, (void *)ps->part,
                  (void *)ps->sc, (void *)ps->coef, (void *)ps->bmom, (void *)ps->d_ptr,
                  (void *)ps->d_nbptr, (void *)ps->d_col, (void *)ps->d_nbcol,
                  (void *)ps->d_wall, (void *)ps->d_val, (void *)ps->d_nbval,
                  (void *)ps->d_dinv, (void *)ps->rh, (void *)ps->v, (void *)ps->y,
                  (void *)ps->s, (void *)ps->tv})
    if (v) (void)hipFree(v);
  for (int l = 0; l < kMaxLevels; ++l)
    for (void *v : {(void *)ps->e[l], (void *)ps->rl[l], (void *)ps->t[l]})
      if (v) (void)hipFree(v);
  for (size_t l = 0; l < ps->mg.size(); ++l) {  // level 0 borrows the solve's operator
    aph_poisson::MgLevel &d = ps->mg[l];
    if (l > 0)
      for (void *v : {(void *)d.ptr, (void *)d.col, (void *)d.val, (void *)d.dinv,
                      (void *)d.cmap, (void *)d.rptr, (void *)d.rcol, (void *)d.rval,
                      (void *)d.r, (void *)d.x})
        if (v) (void)hipFree(v);
    if (d.t) (void)hipFree(d.t);
  }
  if (ps->d_cinv) (void)hipFree(ps->d_cinv);
  if (ps->h_sc) (void)hipHostFree(ps->h_sc);
  delete ps;
  return APH_OK;
}

int aph_poisson_solve(aph_context *ctx, aph_poisson *ps, const double *prim, double *grav,
                      double *phi) {
  if (!ctx) return APH_E_INVALID;
  if (!ps || !prim || !grav) return set_error(ctx, APH_E_INVALID, "aph_poisson_solve: null");
  // primitive densities are floored already: a floor of -inf leaves them as they are
  return poisson_pass(ctx, ps, prim, -HUGE_VAL, grav, phi);
}

int aph_poisson_get_info(aph_context *ctx, aph_poisson *ps, double coeffs[9],
                         int32_t *iterations, double *rel_residual) {
  if (!ctx) return APH_E_INVALID;
  if (!ps) return set_error(ctx, APH_E_INVALID, "aph_poisson_get_info: null");
  double c[9];
  APH_HIP_CHECK(ctx, hipMemcpyAsync(c, ps->coef, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
  APH_HIP_CHECK(ctx, hipMemcpyAsync(ps->h_sc, ps->sc, sizeof(double) * NSC_ALL,
                                    hipMemcpyDeviceToHost, ctx->stream));
  APH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (coeffs)
    for (int i = 0; i < 9; ++i) coeffs[i] = c[i];
  if (iterations) *iterations = (int32_t)ps->h_sc[ITERS];
  if (rel_residual)
    *rel_residual = ps->h_sc[BB] > 0.0 ? sqrt(ps->h_sc[RR] / ps->h_sc[BB]) : 0.0;
  return APH_OK;
}

}  // extern "C"
