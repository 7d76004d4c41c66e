// aph_gravity.hip -- monopole self-gravity (src/gravity/gsolvers/monopole_gravity.hpp)
// and ApplyGravity with a per-cell field (src/gravity/gravity.cpp:374-410).
//
// Per call of the solver's tasks:
//   k_mono_zero   ZeroRadialBins: clear the 128-bit bin accumulators and the exponent bound
//   k_mono_bound  the largest binary exponent of a cell mass rho*V (r < rmax), which fixes
//                 the fixed-point scale s of the sums
//   k_mono_accum  AccumulateMass: each cell's mass, rounded to a multiple of 2^-s, added as
//                 a 128-bit integer to its radial bin (workgroup partial sums in LDS)
//   k_mono_tables BuildRadialTables: bin masses, m_enc (prefix) and phi_rad (suffix) in
//                 the reference's sequential order, one workgroup
//   k_mono_field  ComputeGravityVectorDirect: g (3 components) and phi per interior cell
//
// The reference adds a bin's masses with floating-point atomics, whose order (and so the
// result's last bits) varies from run to run.  Integer addition does not depend on order:
// the bin masses here are the same on every run and equal the correctly rounded exact sum
// of the rounded cell masses (about 100 bits below the largest mass).
#include <climits>

#include "aph_device.hpp"
#include "aph_runtime.hpp"

struct aph_monopole {
  int32_t ndim = 0, nrbin = 0;
  double rmax = 0.0, G = 0.0;
  unsigned long long *acc = nullptr;  // [nrbin][2]: low, high 64 bits (two's complement)
  int *emax = nullptr;                // largest exponent of a cell mass of the last pass
  double *mbin = nullptr;             // bin masses, then the potential increments
  double *m_enc = nullptr, *phi_rad = nullptr;  // [nrbin + 1] each
  const aph_blockset *bs = nullptr;   // the block set of the last aph_monopole_zero_radial_bins
  int log2_cells = 0;                 // ceil(log2(its interior cells))
};

namespace aph {
namespace {

constexpr int kThreads = 256;
constexpr int kLdsMaxBins = 4080;  // LDS partial sums within 64 KiB per workgroup
using u128 = unsigned __int128;

__device__ __forceinline__ int mono_exponent(double dm) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(dm);
  const int ex = (int)((bits >> 52) & 0x7ff);
  return (ex == 0 ? 1 : ex) - 1023;
}

// The sum of the block set's cell masses, each below 2^(emax + 1), stays below 2^125.
__device__ __forceinline__ int mono_scale(int emax, int log2_cells) {
  return emax == INT_MIN ? 0 : 125 - log2_cells - (emax + 1);
}

// dm * 2^s rounded to nearest (ties to even), two's complement.
__device__ __forceinline__ u128 mono_quantize(double dm, int s) {
  if (dm == 0.0) return 0;
  const unsigned long long bits = (unsigned long long)__double_as_longlong(dm);
  int ex = (int)((bits >> 52) & 0x7ff);
  unsigned long long man = bits & ((1ull << 52) - 1);
  if (ex == 0)
    ex = 1;
  else
    man |= 1ull << 52;
  const int k = ex - 1075 + s;
  u128 q;
  if (k >= 0) {
    q = (u128)man << k;
  } else if (-k >= 64) {
    q = 0;
  } else {
    const int sh = -k;
    unsigned long long m = man >> sh;
    const unsigned long long rem = man & ((1ull << sh) - 1), half = 1ull << (sh - 1);
    if (rem > half || (rem == half && (m & 1))) ++m;
    q = m;
  }
  return (bits >> 63) ? (u128)0 - q : q;
}

// q * 2^-s rounded once to nearest even.
__device__ double mono_to_double(u128 q, int s) {
  const bool neg = (q >> 127) != 0;
  if (neg) q = (u128)0 - q;
  if (q == 0) return 0.0;
  const unsigned long long hi = (unsigned long long)(q >> 64), lo = (unsigned long long)q;
  const int t = hi ? 127 - __clzll((long long)hi) : 63 - __clzll((long long)lo);
  unsigned long long m;
  int e = 0;
  if (t <= 52) {
    m = lo;
  } else {
    const int sh = t - 52;
    m = (unsigned long long)(q >> sh);
    const u128 rem = q & (((u128)1 << sh) - 1), half = (u128)1 << (sh - 1);
    if (rem > half || (rem == half && (m & 1))) ++m;
    if (m == (1ull << 53)) {
      m >>= 1;
      ++e;
    }
    e += sh;
  }
  const double r = ldexp((double)m, e - s);
  return neg ? -r : r;
}

// 128-bit add into acc[2*bin], acc[2*bin+1]: the low word's atomic returns the value it
// added to, which tells whether this add carried into the high word.
__device__ __forceinline__ void add128(unsigned long long *acc, int bin, unsigned long long lo,
                                       unsigned long long hi) {
  const unsigned long long old = atomicAdd(acc + 2 * bin, lo);
  const unsigned long long h = hi + (old + lo < old ? 1ull : 0ull);
  if (h) atomicAdd(acc + 2 * bin + 1, h);
}

__device__ __forceinline__ void decompose(int64_t t, int nx0, int nx1, int &i, int &j, int &k) {
  i = (int)(t % nx0);
  const int64_t r = t / nx0;
  j = (int)(r % nx1);
  k = (int)(r / nx1);
}

// Interior cell (i, j, k): its centre on every axis (an inactive axis' single cell is
// centred on the domain), r, and its mass rho * V with ConsToPrim's density floor
// (idempotent on primitive densities).
struct CellMass {
  double x, y, z, r, dm;
};
__device__ __forceinline__ CellMass cell_mass(const aph_block &bk, int ndim, int nvar,
                                              const double *field, double dfloor, int i,
                                              int j, int k) {
  CellMass c;
  c.x = bk.xmin[0] + (i + 0.5) * bk.dx[0];
  c.y = bk.xmin[1] + (j + 0.5) * bk.dx[1];
  c.z = bk.xmin[2] + (k + 0.5) * bk.dx[2];
  c.r = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim);
  const int64_t cell = ((int64_t)(k + gst(bk, 2, ndim)) * N1 + (j + gst(bk, 1, ndim))) * N0 +
                       (i + gst(bk, 0, ndim));
  double rho = field[(int64_t)nvar * bk.cell_base + cell];
  rho = (rho > dfloor) ? rho : dfloor;
  c.dm = rho * bk.vol;
  return c;
}

__global__ void k_mono_zero(unsigned long long *acc, int n, int *emax) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    acc[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) *emax = INT_MIN;
}

// One workgroup takes up to kChunk consecutive interior cells of one block (a whole 32^3
// block), so that it touches few radial bins and adds each of them to HBM once.
constexpr int kChunk = 32768;
constexpr int kPassThreads = 1024;  // bound / accumulate: 16 waves per workgroup

__global__ void __launch_bounds__(kPassThreads)
    k_mono_bound(const aph_block *__restrict__ blocks, int first, int ndim, int nvar,
                 const double *__restrict__ field, double dfloor, double rmax, int *emax) {
  __shared__ int wave_max[kPassThreads / 64];
  const aph_block &bk = blocks[first + blockIdx.y];
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const int64_t c0 = (int64_t)blockIdx.x * kChunk;
  const int64_t c1 = c0 + kChunk < nint ? c0 + kChunk : nint;
  int e = INT_MIN;
  for (int64_t t = c0 + threadIdx.x; t < c1; t += blockDim.x) {
    int i, j, k;
    decompose(t, bk.nx[0], bk.nx[1], i, j, k);
    const CellMass c = cell_mass(bk, ndim, nvar, field, dfloor, i, j, k);
    if (c.r < rmax && c.dm != 0.0) e = max(e, mono_exponent(c.dm));
  }
  for (int o = 32; o > 0; o >>= 1) e = max(e, __shfl_xor(e, o));
  if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kPassThreads / 64; ++w) e = max(e, wave_max[w]);
    // one atomic per workgroup, none when the bound is already as large
    if (e != INT_MIN && e > __hip_atomic_load(emax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(emax, e);
  }
}

template <bool LDS>
__global__ void __launch_bounds__(kPassThreads)
    k_mono_accum(const aph_block *__restrict__ blocks, int first, int ndim, int nvar,
                 const double *__restrict__ field, double dfloor, double rmax, int nrbin,
                 int log2_cells, const int *__restrict__ emax, unsigned long long *acc) {
  extern __shared__ unsigned long long part[];  // [nrbin][2] when LDS
  __shared__ int range[2];                      // bins this workgroup touched
  const aph_block &bk = blocks[first + blockIdx.y];
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const int64_t c0 = (int64_t)blockIdx.x * kChunk;
  const int64_t c1 = c0 + kChunk < nint ? c0 + kChunk : nint;
  const double dr = rmax / nrbin;
  const int s = mono_scale(*emax, log2_cells);
  if constexpr (LDS) {
    for (int i = threadIdx.x; i < 2 * nrbin; i += blockDim.x) part[i] = 0ull;
    if (threadIdx.x == 0) {
      range[0] = nrbin;
      range[1] = -1;
    }
    __syncthreads();
  }
  int blo = nrbin, bhi = -1;
  for (int64_t t = c0 + threadIdx.x; t < c1; t += blockDim.x) {
    int i, j, k;
    decompose(t, bk.nx[0], bk.nx[1], i, j, k);
    const CellMass c = cell_mass(bk, ndim, nvar, field, dfloor, i, j, k);
    if (c.r >= rmax) continue;
    int bin = (int)(c.r / dr);
    bin = bin < 0 ? 0 : (bin >= nrbin ? nrbin - 1 : bin);
    blo = min(blo, bin);
    bhi = max(bhi, bin);
    const u128 q = mono_quantize(c.dm, s);
    add128(LDS ? part : acc, bin, (unsigned long long)q, (unsigned long long)(q >> 64));
  }
  if constexpr (LDS) {
    if (bhi >= 0) {
      atomicMin(&range[0], blo);
      atomicMax(&range[1], bhi);
    }
    __syncthreads();
    for (int b = range[0] + (int)threadIdx.x; b <= range[1]; b += blockDim.x) {
      const unsigned long long lo = part[2 * b], hi = part[2 * b + 1];
      if (lo | hi) add128(acc, b, lo, hi);
    }
  }
}

// BuildRadialTables (monopole_gravity.hpp:155-213).  One workgroup: the bin masses and
// the potential increments in parallel, the two recurrences by one lane in the
// reference's order, reading their operands from LDS (nrbin <= kLdsTableBins).
constexpr int kLdsTableBins = 2048;

template <bool LDS>
__global__ void __launch_bounds__(kThreads)
    k_mono_tables(int nrbin, int ndim, double rmax, double G, int log2_cells,
                  const int *__restrict__ emax, const unsigned long long *__restrict__ acc,
                  double *mbin, double *m_enc, double *phi_rad) {
  extern __shared__ double tab[];  // LDS: [nrbin] bin masses / increments, [nrbin+1] m_enc
  double *a = LDS ? tab : mbin;
  double *e = LDS ? tab + nrbin : m_enc;
  const int s = mono_scale(*emax, log2_cells);
  const double dr = rmax / nrbin;
  for (int i = threadIdx.x; i < nrbin; i += blockDim.x)
    a[i] = mono_to_double(((u128)acc[2 * i + 1] << 64) | acc[2 * i], s);
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < nrbin; ++i) {
      run = run + a[i];
      e[i + 1] = run;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nrbin; i += blockDim.x) {
    const double r_inner = i * dr, r_outer = (i + 1) * dr;
    double dphi;
    if (ndim == 3) {
      const double m_inner = e[i];
      const double shell_mass = e[i + 1] - m_inner;
      if (r_inner == 0.0) {
        dphi = -G * shell_mass / (2.0 * r_outer);
      } else {
        const double denom = r_outer * r_outer * r_outer - r_inner * r_inner * r_inner;
        dphi = -G * (m_inner * (1.0 / r_inner - 1.0 / r_outer) +
                     shell_mass / denom *
                         (0.5 * (r_outer * r_outer - r_inner * r_inner) +
                          r_inner * r_inner * r_inner * (1.0 / r_outer - 1.0 / r_inner)));
      }
    } else {
      const double r_mid = 0.5 * (r_inner + r_outer);
      const double m_mid = 0.5 * (e[i] + e[i + 1]);
      double gr = 0.0;
      if (r_mid > 0.0) gr = (ndim == 1) ? -G * m_mid : -2.0 * G * m_mid / r_mid;
      dphi = gr * dr;
    }
    a[i] = dphi;
    if (LDS) m_enc[i] = e[i];
  }
  if (LDS && threadIdx.x == 0) m_enc[nrbin] = e[nrbin];
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = (rmax > 0.0 && ndim == 3) ? -G * e[nrbin] / rmax : 0.0;
    phi_rad[nrbin] = run;
    for (int i = nrbin - 1; i >= 0; --i) {
      run = run + a[i];
      phi_rad[i] = run;
    }
  }
}

// ComputeGravityVectorDirect (monopole_gravity.hpp:215-302) for every interior cell:
// grav (3 components, pack layout), phi (1 component, may be null).
__global__ void __launch_bounds__(kThreads)
    k_mono_field(const aph_block *__restrict__ blocks, int first, int ndim, int nrbin,
                 double rmax, double G, const double *__restrict__ m_enc,
                 const double *__restrict__ phi_rad, double *__restrict__ grav,
                 double *__restrict__ phi) {
  const aph_block &bk = blocks[first + blockIdx.y];
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  const double dr = rmax / nrbin;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(t, bk.nx[0], bk.nx[1], i, j, k);
    const double x = bk.xmin[0] + (i + 0.5) * bk.dx[0];
    const double y = ndim >= 2 ? bk.xmin[1] + (j + 0.5) * bk.dx[1] : 0.0;
    const double z = ndim >= 3 ? bk.xmin[2] + (k + 0.5) * bk.dx[2] : 0.0;
    const double r2 = x * x + y * y + z * z;
    const double r = sqrt(r2);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, phi_val;
    if (r < 1e-10) {
      phi_val = phi_rad[0];
    } else {
      double m_enclosed;
      if (r >= rmax) {
        m_enclosed = m_enc[nrbin];
        phi_val = (ndim == 3) ? -G * m_enclosed / r : phi_rad[nrbin];
      } else {
        const double rf = r / dr;
        int bin = (int)rf;
        bin = bin < 0 ? 0 : (bin >= nrbin ? nrbin - 1 : bin);
        const double r_inner = bin * dr, r_outer = (bin + 1) * dr;
        const double shell_mass = m_enc[bin + 1] - m_enc[bin];
        double mass_frac;
        if (ndim == 1) {
          mass_frac = (r - r_inner) / dr;
        } else if (ndim == 2) {
          const double denom = r_outer * r_outer - r_inner * r_inner;
          mass_frac = (r * r - r_inner * r_inner) / denom;
        } else {
          const double r_inner3 = r_inner * r_inner * r_inner;
          const double r_outer3 = r_outer * r_outer * r_outer;
          mass_frac = (r * r * r - r_inner3) / (r_outer3 - r_inner3);
        }
        if (mass_frac < 0.0) mass_frac = 0.0;
        if (mass_frac > 1.0) mass_frac = 1.0;
        m_enclosed = m_enc[bin] + shell_mass * mass_frac;
        const double frac = rf - bin;
        phi_val = (1.0 - frac) * phi_rad[bin] + frac * phi_rad[bin + 1];
      }
      double g_r;
      if (ndim == 1)
        g_r = -G * m_enclosed;
      else if (ndim == 2)
        g_r = -2.0 * G * m_enclosed / r;
      else
        g_r = -G * m_enclosed / r2;
      const double rinv = 1.0 / r;
      g0 = g_r * x * rinv;
      g1 = (ndim >= 2) ? g_r * y * rinv : 0.0;
      g2 = (ndim >= 3) ? g_r * z * rinv : 0.0;
    }
    const int64_t c = ((int64_t)(k + gst(bk, 2, ndim)) * N1 + (j + gst(bk, 1, ndim))) * N0 +
                      (i + gst(bk, 0, ndim));
    double *gp = grav + 3 * bk.cell_base + c;
    gp[0] = g0;
    gp[nc] = g1;
    gp[2 * nc] = g2;
    if (phi) phi[bk.cell_base + c] = phi_val;
  }
}

// ApplyGravity (gravity.cpp:374-410) with the field: momenta += g rho dt, energy +=
// (g . v) rho dt, with the stage-start primitives.
template <int NV>
__global__ void __launch_bounds__(kThreads)
    k_gravity_field(const aph_block *__restrict__ blocks, int first, int ndim,
                    double *__restrict__ cons, const double *__restrict__ prim,
                    const double *__restrict__ grav, double dt) {
  const aph_block &bk = blocks[first + blockIdx.y];
  const int N0 = ext(bk, 0, ndim), N1 = ext(bk, 1, ndim), N2 = ext(bk, 2, ndim);
  const int64_t nc = (int64_t)N0 * N1 * N2;
  const int64_t nint = (int64_t)bk.nx[0] * bk.nx[1] * bk.nx[2];
  double *u = cons + (int64_t)NV * bk.cell_base;
  const double *w = prim + (int64_t)NV * bk.cell_base;
  const double *gf = grav + 3 * bk.cell_base;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nint;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    decompose(t, bk.nx[0], bk.nx[1], i, j, k);
    const int64_t c = ((int64_t)(k + gst(bk, 2, ndim)) * N1 + (j + gst(bk, 1, ndim))) * N0 +
                      (i + gst(bk, 0, ndim));
    const double g0 = gf[c], g1 = gf[nc + c], g2 = gf[2 * nc + c];
    u[nc + c] += g0 * w[c] * dt;
    double de = g0 * w[nc + c];
    if (ndim >= 2) {
      u[2 * nc + c] += g1 * w[c] * dt;
      de += g1 * w[2 * nc + c];
    }
    if (ndim >= 3) {
      u[3 * nc + c] += g2 * w[c] * dt;
      de += g2 * w[3 * nc + c];
    }
    u[4 * nc + c] += de * w[c] * dt;
  }
}

dim3 chunk_grid(const aph_blockset *bs, int cnt) {
  return dim3((unsigned)std::max<int64_t>(1, (bs->max_interior + kChunk - 1) / kChunk),
              (unsigned)cnt);
}

dim3 cell_grid(const aph_blockset *bs, int cnt) {
  const int64_t per = (bs->max_interior + kThreads - 1) / kThreads;
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(per, 64)), (unsigned)cnt);
}

int check_mono(aph_context *ctx, const aph_monopole *m, const aph_blockset *bs,
               const char *who) {
  if (!m || !bs) return set_error(ctx, APH_E_INVALID, "%s: null", who);
  if (bs->ndim != m->ndim)
    return set_error(ctx, APH_E_INVALID, "%s: block set is %dD, solver %dD", who, bs->ndim,
                     m->ndim);
  return APH_OK;
}

int log2_ceil(int64_t n) {
  int L = 0;
  while ((int64_t(1) << L) < n) ++L;
  return L;
}

// ZeroRadialBins plus the exponent bound over every block of bs (density from `field`,
// floored at dfloor), and AccumulateMass over blocks [lo, lo + cnt).
int mono_begin(aph_context *ctx, aph_monopole *m, const aph_blockset *bs, const double *field,
               double dfloor) {
  int64_t cells = 0;
  for (const aph_block &b : bs->h) cells += (int64_t)b.nx[0] * b.nx[1] * b.nx[2];
  m->bs = bs;
  m->log2_cells = log2_ceil(std::max<int64_t>(cells, 1));
  const int n = 2 * m->nrbin;
  hipLaunchKernelGGL(k_mono_zero, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     ctx->stream, m->acc, n, m->emax);
  APH_LAUNCH_CHECK(ctx);
  if (bs->nblocks > 0) {
    hipLaunchKernelGGL(k_mono_bound, chunk_grid(bs, bs->nblocks), dim3(kPassThreads), 0,
                       ctx->stream, bs->d, 0, bs->ndim, bs->nvar, field, dfloor, m->rmax,
                       m->emax);
    APH_LAUNCH_CHECK(ctx);
  }
  return APH_OK;
}

int mono_accumulate(aph_context *ctx, aph_monopole *m, const aph_blockset *bs, int lo, int cnt,
                    const double *field, double dfloor) {
  if (m->nrbin <= kLdsMaxBins) {
    hipLaunchKernelGGL(k_mono_accum<true>, chunk_grid(bs, cnt), dim3(kPassThreads),
                       sizeof(unsigned long long) * 2 * m->nrbin, ctx->stream, bs->d, lo,
                       bs->ndim, bs->nvar, field, dfloor, m->rmax, m->nrbin, m->log2_cells,
                       m->emax, m->acc);
  } else {
    hipLaunchKernelGGL(k_mono_accum<false>, chunk_grid(bs, cnt), dim3(kPassThreads), 0,
                       ctx->stream, bs->d, lo, bs->ndim, bs->nvar, field, dfloor, m->rmax,
                       m->nrbin, m->log2_cells, m->emax, m->acc);
  }
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // namespace

// The solver's whole pass over every block of bs, as the fused plan runs it before each
// stage: bins from the stage input's density (`field` conserved, floored at dfloor as
// ConsToPrim does), the tables, and the field (grav, 3 components).
int monopole_pass(aph_context *ctx, aph_monopole *m, const aph_blockset *bs,
                  const double *field, double dfloor, double *grav) {
  int rc = mono_begin(ctx, m, bs, field, dfloor);
  if (!rc && bs->nblocks > 0) rc = mono_accumulate(ctx, m, bs, 0, bs->nblocks, field, dfloor);
  if (!rc) rc = aph_monopole_build_radial_tables(ctx, m);
  if (!rc) rc = aph_monopole_gravity_field(ctx, m, bs, 0, bs->nblocks, grav, nullptr);
  return rc;
}

}  // namespace aph

using namespace aph;

extern "C" {

int aph_monopole_create(aph_context *ctx, int32_t ndim, int32_t nrbin, double rmax,
                        double gravity_constant, aph_monopole **out) {
  if (!ctx) return APH_E_INVALID;
  if (!out || ndim < 1 || ndim > 3 || nrbin < 1 || !(rmax > 0.0))
    return set_error(ctx, APH_E_INVALID,
                     "aph_monopole_create: ndim 1..3, nrbin >= 1 and rmax > 0 required");
  aph_monopole *m = new aph_monopole;
  m->ndim = ndim;
  m->nrbin = nrbin;
  m->rmax = rmax;
  m->G = gravity_constant;
  const size_t nb = (size_t)nrbin;
  hipError_t e = hipMalloc(&m->acc, sizeof(unsigned long long) * 2 * nb);
  if (e == hipSuccess) e = hipMalloc(&m->emax, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&m->mbin, sizeof(double) * nb);
  if (e == hipSuccess) e = hipMalloc(&m->m_enc, sizeof(double) * (nb + 1));
  if (e == hipSuccess) e = hipMalloc(&m->phi_rad, sizeof(double) * (nb + 1));
  if (e == hipSuccess) e = hipMemset(m->m_enc, 0, sizeof(double) * (nb + 1));
  if (e == hipSuccess) e = hipMemset(m->phi_rad, 0, sizeof(double) * (nb + 1));
  if (e != hipSuccess) {
    aph_monopole_destroy(ctx, m);
    return set_error(ctx, APH_E_HIP, "aph_monopole_create: %s", hipGetErrorString(e));
  }
  *out = m;
  return APH_OK;
}

int aph_monopole_destroy(aph_context *ctx, aph_monopole *m) {
  (void)ctx;
  if (!m) return APH_OK;
  for (void *p : {(void *)m->acc, (void *)m->emax, (void *)m->mbin, (void *)m->m_enc,
                  (void *)m->phi_rad})
    if (p) (void)hipFree(p);
  delete m;
  return APH_OK;
}

int aph_monopole_zero_radial_bins(aph_context *ctx, aph_monopole *m, const aph_blockset *bs,
                                  const double *prim) {
  if (!ctx) return APH_E_INVALID;
  int rc = check_mono(ctx, m, bs, "aph_monopole_zero_radial_bins");
  if (rc) return rc;
  if (!prim && bs->nblocks > 0)
    return set_error(ctx, APH_E_INVALID, "aph_monopole_zero_radial_bins: null field");
  // primitive densities are floored already: a floor of -inf leaves them as they are
  return mono_begin(ctx, m, bs, prim, -HUGE_VAL);
}

int aph_monopole_accumulate_mass(aph_context *ctx, aph_monopole *m, const aph_blockset *bs,
                                 int32_t first_block, int32_t nblocks, const double *prim) {
  if (!ctx) return APH_E_INVALID;
  int rc = check_mono(ctx, m, bs, "aph_monopole_accumulate_mass");
  if (rc) return rc;
  if (m->bs != bs)
    return set_error(ctx, APH_E_STATE,
                     "aph_monopole_accumulate_mass: call aph_monopole_zero_radial_bins on "
                     "this block set first");
  int lo, cnt;
  rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (cnt == 0) return APH_OK;
  if (!prim) return set_error(ctx, APH_E_INVALID, "aph_monopole_accumulate_mass: null");
  return mono_accumulate(ctx, m, bs, lo, cnt, prim, -HUGE_VAL);
}

int aph_monopole_build_radial_tables(aph_context *ctx, aph_monopole *m) {
  if (!ctx) return APH_E_INVALID;
  if (!m) return set_error(ctx, APH_E_INVALID, "aph_monopole_build_radial_tables: null");
  if (m->nrbin <= kLdsTableBins)
    hipLaunchKernelGGL(k_mono_tables<true>, dim3(1), dim3(kThreads),
                       sizeof(double) * (2 * (size_t)m->nrbin + 1), ctx->stream, m->nrbin,
                       m->ndim, m->rmax, m->G, m->log2_cells, m->emax, m->acc, m->mbin,
                       m->m_enc, m->phi_rad);
  else
    hipLaunchKernelGGL(k_mono_tables<false>, dim3(1), dim3(kThreads), 0, ctx->stream,
                       m->nrbin, m->ndim, m->rmax, m->G, m->log2_cells, m->emax, m->acc,
                       m->mbin, m->m_enc, m->phi_rad);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int aph_monopole_gravity_field(aph_context *ctx, aph_monopole *m, const aph_blockset *bs,
                               int32_t first_block, int32_t nblocks, double *grav,
                               double *phi) {
  if (!ctx) return APH_E_INVALID;
  int rc = check_mono(ctx, m, bs, "aph_monopole_gravity_field");
  if (rc) return rc;
  int lo, cnt;
  rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (cnt == 0) return APH_OK;
  if (!grav) return set_error(ctx, APH_E_INVALID, "aph_monopole_gravity_field: null");
  hipLaunchKernelGGL(k_mono_field, cell_grid(bs, cnt), dim3(kThreads), 0, ctx->stream, bs->d,
                     lo, bs->ndim, m->nrbin, m->rmax, m->G, m->m_enc, m->phi_rad, grav, phi);
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

int aph_monopole_get_tables(aph_context *ctx, aph_monopole *m, double *m_enc,
                            double *phi_rad) {
  if (!ctx) return APH_E_INVALID;
  if (!m || !m_enc || !phi_rad)
    return set_error(ctx, APH_E_INVALID, "aph_monopole_get_tables: null");
  const size_t n = sizeof(double) * ((size_t)m->nrbin + 1);
  APH_HIP_CHECK(ctx, hipMemcpyAsync(m_enc, m->m_enc, n, hipMemcpyDeviceToHost, ctx->stream));
  APH_HIP_CHECK(ctx,
                hipMemcpyAsync(phi_rad, m->phi_rad, n, hipMemcpyDeviceToHost, ctx->stream));
  APH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return APH_OK;
}

int aph_apply_gravity_field(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                            int32_t nblocks, double *cons, const double *prim,
                            const double *grav, double dt) {
  if (!ctx) return APH_E_INVALID;
  int lo, cnt;
  int rc = block_range(ctx, bs, first_block, nblocks, &lo, &cnt);
  if (rc) return rc;
  if (cnt == 0) return APH_OK;
  if (!cons || !prim || !grav)
    return set_error(ctx, APH_E_INVALID, "aph_apply_gravity_field: null");
  const dim3 gr = cell_grid(bs, cnt);
  switch (bs->nvar) {
    case 5:
      hipLaunchKernelGGL(k_gravity_field<5>, gr, dim3(kThreads), 0, ctx->stream, bs->d, lo,
                         bs->ndim, cons, prim, grav, dt);
      break;
    case 12:
      hipLaunchKernelGGL(k_gravity_field<12>, gr, dim3(kThreads), 0, ctx->stream, bs->d, lo,
                         bs->ndim, cons, prim, grav, dt);
      break;
    case 14:
      hipLaunchKernelGGL(k_gravity_field<14>, gr, dim3(kThreads), 0, ctx->stream, bs->d, lo,
                         bs->ndim, cons, prim, grav, dt);
      break;
    default:
      return set_error(ctx, APH_E_INVALID, "aph_apply_gravity_field: nvar %d", bs->nvar);
  }
  APH_LAUNCH_CHECK(ctx);
  return APH_OK;
}

}  // extern "C"
