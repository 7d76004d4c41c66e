// apophis_run.cpp -- a C++ host program driving the flux path through the adapter
// (apophis_hydro_package.hpp) the way the reference's main/driver do:
//   main.cpp:10-68                 parse the deck + overrides, pick the problem
//   hydro.cpp:33-349               InitializeHydro -> flux-function registry
//   pgen/sod.cpp:11-46             problem generator (1D shock tube)
//   apophis_driver.cpp:27-130      RK2 stage task list, upstream dt policy
// Usage: apophis_run <deck> [block/key=value ...] [--out file.bin]
// Writes the final interior conserved state (5 x nx1 doubles) when --out is given.
#include <hip/hip_runtime.h>

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
  return 0;  // periodic
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
    if (pin.GetString("job", "problem_id") != "sod")
      throw ApophisError("this driver implements problem_id = sod");
    auto pkg = InitializeHydro(&pin);
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

    // mesh: 1D lattice of nx1/mb_nx1 blocks
    const int ng = pin.GetInteger("parthenon/mesh", "nghost");
    const int nx = pin.GetInteger("parthenon/mesh", "nx1");
    const int mbx = pin.GetOrAddInteger("parthenon/meshblock", "nx1", nx);
    const Real x0 = pin.GetOrAddReal("parthenon/mesh", "x1min", 0.0);
    const Real x1 = pin.GetOrAddReal("parthenon/mesh", "x1max", 1.0);
    const int ibc = bc_code(pin.GetOrAddString("parthenon/mesh", "ix1_bc", "outflow"));
    const int obc = bc_code(pin.GetOrAddString("parthenon/mesh", "ox1_bc", "outflow"));
    const int nb = nx / mbx;
    const Real dx = (x1 - x0) / nx;
    std::vector<aph_block> blocks(nb);
    for (int b = 0; b < nb; ++b) {
      aph_block &k = blocks[b];
      std::memset(&k, 0, sizeof(k));
      k.nx[0] = mbx;
      k.nx[1] = k.nx[2] = 1;
      k.ng = ng;
      k.cell_base = (int64_t)b * (mbx + 2 * ng);
      k.dx[0] = dx;
      k.dx[1] = k.dx[2] = 1.0;  // inactive axes span [-1/2, 1/2]
      k.area[0] = k.dx[1] * k.dx[2];
      k.area[1] = k.dx[0] * k.dx[2];
      k.area[2] = k.dx[0] * k.dx[1];
      k.vol = k.dx[0] * k.dx[1] * k.dx[2];
      k.xmin[0] = x0 + b * mbx * dx;
      for (int o = 0; o < 27; ++o) k.nbr[o] = -1;
      k.nbr[13] = b;
      const bool per = (ibc == 0 && obc == 0);
      k.nbr[12] = b > 0 ? b - 1 : (per ? nb - 1 : -1);
      k.nbr[14] = b < nb - 1 ? b + 1 : (per ? 0 : -1);
      k.bc[0] = ibc ? ibc : APH_BC_OUTFLOW;
      k.bc[1] = obc ? obc : APH_BC_OUTFLOW;
      k.bc[2] = k.bc[3] = k.bc[4] = k.bc[5] = APH_BC_OUTFLOW;
      k.mesh_id = 0;
    }
    const int64_t ncell = (int64_t)nb * (mbx + 2 * ng);
    const int64_t nfield = 5 * ncell;

    // problem generator (sod.cpp:11-46)
    const Real rl = pin.GetOrAddReal("problem/sod", "rho_l", 1.0);
    const Real pl = pin.GetOrAddReal("problem/sod", "pres_l", 1.0);
    const Real ul = pin.GetOrAddReal("problem/sod", "u_l", 0.0);
    const Real rr = pin.GetOrAddReal("problem/sod", "rho_r", 0.125);
    const Real pr = pin.GetOrAddReal("problem/sod", "pres_r", 0.1);
    const Real ur = pin.GetOrAddReal("problem/sod", "u_r", 0.0);
    const Real xd = pin.GetOrAddReal("problem/sod", "x_discont", 0.5);
    const Real gamma = pin.GetOrAddReal("hydro", "gamma", 1.4);
    std::vector<Real> u(nfield, 0.0);
    for (int b = 0; b < nb; ++b) {
      const int64_t nc = mbx + 2 * ng;
      Real *ub = u.data() + 5 * blocks[b].cell_base;
      for (int i = 0; i < mbx; ++i) {
        const Real xc = blocks[b].xmin[0] + (i + 0.5) * dx;
        const bool left = xc < xd;
        const Real rho = left ? rl : rr, vel = left ? ul : ur, p = left ? pl : pr;
        ub[0 * nc + ng + i] = rho;
        ub[1 * nc + ng + i] = rho * vel;
        ub[2 * nc + ng + i] = 0.0;
        ub[3 * nc + ng + i] = 0.0;
        ub[4 * nc + ng + i] = 0.5 * rho * vel * vel + p / (gamma - 1.0);
      }
    }

    // device state
    hipStream_t stream;
    hip_ok(hipStreamCreate(&stream), "hipStreamCreate");
    aph_context *ctx = nullptr;
    if (aph_context_create(stream, &ctx) != APH_OK) throw ApophisError("no HIP device");
    aph_blockset *bs = nullptr;
    check(ctx, aph_blockset_create(ctx, blocks.data(), nb, 1, 0, &bs));
    Real *d = nullptr;
    hip_ok(hipMalloc(&d, sizeof(Real) * (4 * nfield + 2 * ncell + 1)), "hipMalloc");
    auto md = std::make_shared<MeshData<Real>>();
    md->ctx = ctx;
    md->blocks = bs;
    md->cons = d;
    md->prim = d + nfield;
    md->flux[0] = d + 2 * nfield;
    Real *u1 = d + 3 * nfield;
    md->gamma = d + 4 * nfield;
    Real *dt_dev = d + 4 * nfield + 2 * ncell;
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
        auto *flux = pkg->Param<FluxFun_t *>(s == 0 ? "flux_first_stage" : "flux_other_stage");
        flux(md);
        UpdateWithFluxDivergence(md.get(), u1, stages[s].gam0, stages[s].gam1,
                                 stages[s].beta * dt);
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
    std::printf("apophis_run: problem=sod steps=%d t=%.17g\n", n, t);
    if (!out.empty()) {
      FILE *f = std::fopen(out.c_str(), "wb");
      if (!f) throw ApophisError("cannot write " + out);
      std::fwrite(u.data(), sizeof(Real), nfield, f);
      std::fclose(f);
    }
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
