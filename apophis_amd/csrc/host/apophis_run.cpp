// apophis_run.cpp -- a C++ host program driving the flux path through the adapter
// (apophis_hydro_package.hpp) the way the reference's main/driver do:
//   main.cpp:10-68                 parse the deck + overrides, pick the problem
//   hydro.cpp:33-349               InitializeHydro -> flux-function registry
//   gravity.cpp:220-370            InitializeGravity -> the stage's gravity solver
//   pgen/sod.cpp:11-46             problem generator (1D shock tube)
//   pgen/evrards_collapse.cpp:15-49  problem generator (Evrard's collapse, 3D)
//   apophis_driver.cpp:27-130      stage task list (gravity solve, fluxes, update,
//                                  ApplyGravity, ghosts, ConsToPrim), upstream dt policy
// The mesh is one uniform lattice of blocks (parthenon/mesh refinement = none).
// Usage: apophis_run <deck> [block/key=value ...] [--out file.bin]
// Writes the final conserved field (5 components, ghosts included, pack layout) when
// --out is given.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "apophis_hydro_package.hpp"

using namespace Apophis;

namespace {

void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) throw ApophisError(std::string(what) + ": " + hipGetErrorString(e));
}

int bc_code(const std::string &s) {
  if (s == "outflow") return APH_BC_OUTFLOW;
  if (s == "reflecting" || s == "reflect") return APH_BC_REFLECT;
  if (s == "periodic") return APH_BC_PERIODIC;
  throw ApophisError("unknown boundary condition " + s);
}

// One uniform lattice from parthenon/mesh and parthenon/meshblock: nb[d] blocks of bn[d]
// cells per axis (one block of one cell on inactive axes, whose extent sets the
// coordinate width, [-1/2, 1/2] by default as in Parthenon), block b at lattice position
// (lx, ly, lz) = b in x-fastest order, neighbours through the periodic faces.
struct Lattice {
  int ndim = 1, ng = 0;
  int nb[3] = {1, 1, 1}, bn[3] = {1, 1, 1};
  std::vector<aph_block> blocks;
  int64_t ncell = 0;  // entire cells of the mesh (one component)
};

Lattice make_lattice(parthenon::ParameterInput &pin) {
  Lattice L;
  if (pin.GetOrAddString("parthenon/mesh", "refinement", "none") != "none")
    throw ApophisError("this driver runs one uniform lattice (parthenon/mesh refinement = "
                       "none); multilevel meshes run from the Python driver");
  L.ng = pin.GetInteger("parthenon/mesh", "nghost");
  int nx[3];
  Real lo[3], hi[3];
  int code[6];
  for (int d = 0; d < 3; ++d) {
    const std::string n = std::to_string(d + 1);
    nx[d] = pin.GetOrAddInteger("parthenon/mesh", "nx" + n, 1);
    lo[d] = pin.GetOrAddReal("parthenon/mesh", "x" + n + "min", d == 0 ? 0.0 : -0.5);
    hi[d] = pin.GetOrAddReal("parthenon/mesh", "x" + n + "max", d == 0 ? 1.0 : 0.5);
    L.bn[d] = pin.GetOrAddInteger("parthenon/meshblock", "nx" + n, nx[d]);
    if (nx[d] < 1 || L.bn[d] < 1)
      throw ApophisError("mesh and meshblock nx" + n + " must be positive");
    if (nx[d] % L.bn[d]) throw ApophisError("meshblock nx" + n + " must divide mesh nx" + n);
    L.nb[d] = nx[d] / L.bn[d];
    code[2 * d] = bc_code(pin.GetOrAddString("parthenon/mesh", "ix" + n + "_bc", "outflow"));
    code[2 * d + 1] =
        bc_code(pin.GetOrAddString("parthenon/mesh", "ox" + n + "_bc", "outflow"));
    if (nx[d] > 1) L.ndim = d + 1;
  }
  const Real dx[3] = {(hi[0] - lo[0]) / (L.nb[0] * L.bn[0]), (hi[1] - lo[1]) / (L.nb[1] * L.bn[1]),
                      (hi[2] - lo[2]) / (L.nb[2] * L.bn[2])};
  const int nbt = L.nb[0] * L.nb[1] * L.nb[2];
  int64_t per = 1;
  for (int d = 0; d < 3; ++d) per *= L.bn[d] + (d < L.ndim ? 2 * L.ng : 0);
  L.blocks.resize(nbt);
  for (int lz = 0; lz < L.nb[2]; ++lz)
    for (int ly = 0; ly < L.nb[1]; ++ly)
      for (int lx = 0; lx < L.nb[0]; ++lx) {
        const int b = (lz * L.nb[1] + ly) * L.nb[0] + lx;
        aph_block &k = L.blocks[b];
        std::memset(&k, 0, sizeof(k));
        const int loc[3] = {lx, ly, lz};
        for (int d = 0; d < 3; ++d) {
          k.nx[d] = L.bn[d];
          k.dx[d] = dx[d];
          k.xmin[d] = lo[d] + (loc[d] * L.bn[d]) * dx[d];
        }
        k.ng = L.ng;
        k.cell_base = (int64_t)b * per;
        k.area[0] = dx[1] * dx[2];
        k.area[1] = dx[0] * dx[2];
        k.area[2] = dx[0] * dx[1];
        k.vol = dx[0] * dx[1] * dx[2];
        for (int f = 0; f < 6; ++f) k.bc[f] = code[f] == APH_BC_PERIODIC ? APH_BC_OUTFLOW : code[f];
        for (int oz = -1; oz <= 1; ++oz)
          for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
              const int o[3] = {ox, oy, oz};
              int p[3] = {loc[0], loc[1], loc[2]}, nbr = 0;
              for (int d = 0; d < 3 && nbr >= 0; ++d) {
                if (o[d] == 0) continue;
                if (d >= L.ndim) {  // no neighbour along an inactive axis
                  nbr = -1;
                  break;
                }
                p[d] += o[d];
                if (p[d] >= 0 && p[d] < L.nb[d]) continue;
                // the crossed face: periodic wraps, a physical boundary has no neighbour
                if (code[2 * d + (o[d] > 0 ? 1 : 0)] == APH_BC_PERIODIC)
                  p[d] = (p[d] + L.nb[d]) % L.nb[d];
                else
                  nbr = -1;
              }
              k.nbr[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)] =
                  nbr < 0 ? -1 : (p[2] * L.nb[1] + p[1]) * L.nb[0] + p[0];
            }
        k.mesh_id = 0;
      }
  L.ncell = (int64_t)nbt * per;
  return L;
}

// interior cell (i, j, k) of block b: its offset within one component
int64_t cell_at(const Lattice &L, const aph_block &k, int i, int j, int kk) {
  const int g0 = L.ng, g1 = L.ndim > 1 ? L.ng : 0, g2 = L.ndim > 2 ? L.ng : 0;
  const int64_t N0 = k.nx[0] + 2 * g0, N1 = k.nx[1] + 2 * g1;
  return ((int64_t)(kk + g2) * N1 + (j + g1)) * N0 + (i + g0);
}

// pgen/sod.cpp:11-46
void sod(parthenon::ParameterInput &pin, const Lattice &L, std::vector<Real> &u) {
  const Real rl = pin.GetOrAddReal("problem/sod", "rho_l", 1.0);
  const Real pl = pin.GetOrAddReal("problem/sod", "pres_l", 1.0);
  const Real ul = pin.GetOrAddReal("problem/sod", "u_l", 0.0);
  const Real rr = pin.GetOrAddReal("problem/sod", "rho_r", 0.125);
  const Real pr = pin.GetOrAddReal("problem/sod", "pres_r", 0.1);
  const Real ur = pin.GetOrAddReal("problem/sod", "u_r", 0.0);
  const Real xd = pin.GetOrAddReal("problem/sod", "x_discont", 0.5);
  const Real gamma = pin.GetOrAddReal("hydro", "gamma", 1.4);
  const int64_t per = L.ncell / (int64_t)L.blocks.size();
  for (const aph_block &k : L.blocks) {
    Real *ub = u.data() + 5 * k.cell_base;
    for (int kk = 0; kk < k.nx[2]; ++kk)
      for (int j = 0; j < k.nx[1]; ++j)
        for (int i = 0; i < k.nx[0]; ++i) {
          const int64_t c = cell_at(L, k, i, j, kk);
          const Real xc = k.xmin[0] + (i + 0.5) * k.dx[0];
          const bool left = xc < xd;
          const Real rho = left ? rl : rr, vel = left ? ul : ur, p = left ? pl : pr;
          ub[0 * per + c] = rho;
          ub[1 * per + c] = rho * vel;
          ub[2 * per + c] = 0.0;
          ub[3 * per + c] = 0.0;
          ub[4 * per + c] = 0.5 * rho * vel * vel + p / (gamma - 1.0);
        }
  }
}

// pgen/evrards_collapse.cpp:15-49: rho = M / (2 pi R^2 r) inside R, 1e-5 outside, at rest,
// E = u_therm rho
void evrard(parthenon::ParameterInput &pin, const Lattice &L, std::vector<Real> &u) {
  const Real mass = std::stod(pin.GetString("problem", "mass"));
  const Real radius = std::stod(pin.GetString("problem", "radius"));
  const Real u_therm = std::stod(pin.GetString("problem", "u_therm"));
  if (L.ndim != 3) throw ApophisError("evrards_collapse is a 3D problem");
  const int64_t per = L.ncell / (int64_t)L.blocks.size();
  for (const aph_block &k : L.blocks) {
    Real *ub = u.data() + 5 * k.cell_base;
    for (int kk = 0; kk < k.nx[2]; ++kk)
      for (int j = 0; j < k.nx[1]; ++j)
        for (int i = 0; i < k.nx[0]; ++i) {
          const int64_t c = cell_at(L, k, i, j, kk);
          const Real x = k.xmin[0] + (i + 0.5) * k.dx[0];
          const Real y = k.xmin[1] + (j + 0.5) * k.dx[1];
          const Real z = k.xmin[2] + (kk + 0.5) * k.dx[2];
          const Real r = std::sqrt(x * x + y * y + z * z);
          const Real rho = r < radius ? mass / (2.0 * M_PI * (radius * radius) * r) : 1e-5;
          ub[0 * per + c] = rho;
          ub[1 * per + c] = 0.0;
          ub[2 * per + c] = 0.0;
          ub[3 * per + c] = 0.0;
          ub[4 * per + c] = u_therm * rho;
        }
  }
}

}  // namespace

int main(int argc, char **argv) {
  try {
    if (argc < 2) {
      std::fprintf(stderr, "usage: %s <deck> [block/key=value ...] [--out file]\n", argv[0]);
      return 2;
    }
    parthenon::ParameterInput pin;
    pin.LoadDeck(argv[1]);
    std::string out;
    for (int a = 2; a < argc; ++a) {
      if (!std::strcmp(argv[a], "--out") && a + 1 < argc)
        out = argv[++a];
      else
        pin.ApplyOverride(argv[a]);
    }
    const std::string problem = pin.GetString("job", "problem_id");
    if (problem != "sod" && problem != "evrards_collapse")
      throw ApophisError("this driver implements problem_id = sod and evrards_collapse");
    auto pkg = InitializeHydro(&pin);
    auto gpkg = InitializeGravity(&pin);
    if (gpkg->Param<Gravity>("gravity") == Gravity::constant)
      throw ApophisError("gravity/type = constant: no problem generator here sets the field");
    const aph_eos eos = pkg->Param<aph_eos>("eos");
    const Real cfl = pkg->Param<Real>("cfl");
    // upstream low-storage stage tables: {gam0, gam1, beta} per stage
    struct Stage { Real gam0, gam1, beta; };
    std::vector<Stage> stages;
    switch (pkg->Param<Integrator>("integrator")) {
      case Integrator::rk1: stages = {{0.0, 1.0, 1.0}}; break;
      case Integrator::rk2: stages = {{0.0, 1.0, 1.0}, {0.5, 0.5, 0.5}}; break;
      case Integrator::rk3:
        stages = {{0.0, 1.0, 1.0}, {0.25, 0.75, 0.25}, {2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0}};
        break;
      case Integrator::vl2: stages = {{0.0, 1.0, 0.5}, {0.0, 1.0, 1.0}}; break;
      default: throw ApophisError("[Apophis]: Integrator not recognized. Exiting.");
    }

    const Lattice L = make_lattice(pin);
    const int nb = static_cast<int>(L.blocks.size());
    const int64_t ncell = L.ncell;
    const int64_t nfield = 5 * ncell;
    std::vector<Real> u(nfield, 0.0);
    if (problem == "sod")
      sod(pin, L, u);
    else
      evrard(pin, L, u);

    // device state: cons, prim, three flux arrays, u1, gamma (2), dt scratch, and the
    // gravity field (3) and potential (1) when a solver writes them
    hipStream_t stream;
    hip_ok(hipStreamCreate(&stream), "hipStreamCreate");
    aph_context *ctx = nullptr;
    if (aph_context_create(stream, &ctx) != APH_OK) throw ApophisError("no HIP device");
    aph_blockset *bs = nullptr;
    check(ctx, aph_blockset_create(ctx, L.blocks.data(), nb, L.ndim, 0, &bs));
    auto gsolver = MakeGravitySolver(*gpkg, ctx, bs, L.ndim);
    const bool gfield = gsolver->Solves();
    Real *d = nullptr;
    const int64_t total = 6 * nfield + 2 * ncell + 1 + (gfield ? 4 * ncell : 0);
    hip_ok(hipMalloc(&d, sizeof(Real) * total), "hipMalloc");
    hip_ok(hipMemset(d, 0, sizeof(Real) * total), "hipMemset");
    auto md = std::make_shared<MeshData<Real>>();
    md->ctx = ctx;
    md->blocks = bs;
    md->ndim = L.ndim;
    md->cons = d;
    md->prim = d + nfield;
    for (int a = 0; a < 3; ++a) md->flux[a] = d + (2 + a) * nfield;
    Real *u1 = d + 5 * nfield;
    md->gamma = d + 6 * nfield;
    Real *dt_dev = d + 6 * nfield + 2 * ncell;
    md->grav = gfield ? dt_dev + 1 : nullptr;
    md->phi = gfield ? md->grav + 3 * ncell : nullptr;
    const std::vector<std::shared_ptr<MeshData<Real>>> parts{md};
    hip_ok(hipMemcpy(md->cons, u.data(), sizeof(Real) * nfield, hipMemcpyHostToDevice),
           "upload");

    // PostInitialization: ghosts + FillDerived, then the first dt
    FillGhosts(md);
    ConsToPrim(md.get(), eos);
    Real est = EstimateTimestep(md.get(), eos, cfl, dt_dev);
    const Real tlim = pin.GetOrAddReal("parthenon/time", "tlim", 0.2);
    const int nlim = pin.GetOrAddInteger("parthenon/time", "nlim", 100000);
    Real t = 0.0, dt = std::numeric_limits<Real>::infinity();
    int n = 0;
    while (t < tlim * (1 - 1e-14) && n < nlim) {
      dt = std::min(2.0 * dt, est);
      if (t < tlim && tlim - t < dt) dt = tlim - t;
      hip_ok(hipMemcpyAsync(u1, md->cons, sizeof(Real) * nfield, hipMemcpyDeviceToDevice,
                            stream), "u1 copy");
      for (size_t s = 0; s < stages.size(); ++s) {
        if (gfield) gsolver->AddTasks(parts, md->grav, md->phi);  // apophis_driver.cpp:64-66
        auto *flux = pkg->Param<FluxFun_t *>(s == 0 ? "flux_first_stage" : "flux_other_stage");
        flux(md);
        UpdateWithFluxDivergence(md.get(), u1, stages[s].gam0, stages[s].gam1,
                                 stages[s].beta * dt);
        if (gfield) ApplyGravity(md, stages[s].beta * dt);  // :84-85
        FillGhosts(md);
        ConsToPrim(md.get(), eos);
      }
      est = EstimateTimestep(md.get(), eos, cfl, dt_dev);
      t += dt;
      ++n;
    }
    check(ctx, aph_context_synchronize(ctx));
    hip_ok(hipMemcpy(u.data(), md->cons, sizeof(Real) * nfield, hipMemcpyDeviceToHost),
           "download");
    std::printf("apophis_run: problem=%s steps=%d t=%.17g", problem.c_str(), n, t);
    if (auto *ps = dynamic_cast<PoissonGravitySolver *>(gsolver.get())) {
      int it = 0;
      Real res = 0;
      ps->Info(&it, &res);
      std::printf(" poisson_iterations=%d residual=%.3e", it, res);
    }
    std::printf("\n");
    if (!out.empty()) {
      FILE *f = std::fopen(out.c_str(), "wb");
      if (!f) throw ApophisError("cannot write " + out);
      std::fwrite(u.data(), sizeof(Real), nfield, f);
      std::fclose(f);
    }
    gsolver.reset();
    (void)hipFree(d);
    aph_blockset_destroy(ctx, bs);
    aph_context_destroy(ctx);
    (void)hipStreamDestroy(stream);
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "apophis_run: %s\n", e.what());
    return 1;
  }
}
