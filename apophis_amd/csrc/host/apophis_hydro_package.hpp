// apophis_hydro_package.hpp -- C++ host adapter with the reference's package/task
// interface, calling the C ABI of include/apophis_hydro.h.
//
// The reference's flux task has the shape
//     TaskStatus CalculateFluxes<Fluid, Reconstruction, RiemannSolver>(
//         std::shared_ptr<MeshData<Real>> &md)                    (src/hydro/hydro.hpp:26-30)
// and is selected from a (fluid, recon, riemann)-keyed registry filled by add_flux_fun
// (hydro.hpp:32-37, hydro.cpp:100-139).  Parthenon is not part of this build, so this
// header carries the minimal subset of its types the task signatures need (Real,
// TaskStatus, MeshData, StateDescriptor params, ParameterInput); inside a real Parthenon
// build these would be the Parthenon types and MeshData would hand out the packs'
// device pointers instead of owning them.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <any>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/apophis_hydro.h"

namespace parthenon {

using Real = double;
enum class TaskStatus { complete, incomplete, fail };

// Minimal ParameterInput: <block>/<key> = value strings with typed getters, read from
// an input deck in the reference's format ("<block>" lines, "key = value  # comment")
// plus "block/key=value" command-line overrides.
class ParameterInput {
 public:
  void LoadDeck(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open input deck " + path);
    std::string line, block;
    while (std::getline(f, line)) {
      auto h = line.find('#');
      if (h != std::string::npos) line = line.substr(0, h);
      auto trim = [](std::string s) {
        const char *ws = " \t\r\n";
        s.erase(0, s.find_first_not_of(ws));
        auto e = s.find_last_not_of(ws);
        return e == std::string::npos ? std::string() : s.substr(0, e + 1);
      };
      line = trim(line);
      if (line.empty()) continue;
      if (line.front() == '<' && line.back() == '>') {
        block = line.substr(1, line.size() - 2);
        continue;
      }
      auto eq = line.find('=');
      if (eq == std::string::npos) continue;
      Set(block, trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
    }
  }
  void ApplyOverride(const std::string &arg) {  // block/sub/key=value
    auto eq = arg.find('='), sl = arg.rfind('/', eq);
    if (eq == std::string::npos || sl == std::string::npos)
      throw std::runtime_error("bad override " + arg);
    Set(arg.substr(0, sl), arg.substr(sl + 1, eq - sl - 1), arg.substr(eq + 1));
  }
  void Set(const std::string &block, const std::string &key, const std::string &v) {
    kv_[block + "/" + key] = v;
  }
  bool Has(const std::string &block, const std::string &key) const {
    return kv_.count(block + "/" + key) > 0;
  }
  std::string GetString(const std::string &block, const std::string &key) const {
    auto it = kv_.find(block + "/" + key);
    if (it == kv_.end()) throw std::runtime_error("missing required " + block + "/" + key);
    return it->second;
  }
  std::string GetOrAddString(const std::string &b, const std::string &k, const std::string &d) {
    if (!Has(b, k)) Set(b, k, d);
    return GetString(b, k);
  }
  Real GetOrAddReal(const std::string &b, const std::string &k, Real d) {
    if (!Has(b, k)) {  // %.17g reads back as the same double (to_string keeps 6 decimals)
      char txt[32];
      std::snprintf(txt, sizeof txt, "%.17g", d);
      Set(b, k, txt);
    }
    return std::stod(GetString(b, k));
  }
  int GetInteger(const std::string &b, const std::string &k) const {
    return std::stoi(GetString(b, k));
  }
  int GetOrAddInteger(const std::string &b, const std::string &k, int d) {
    if (!Has(b, k)) Set(b, k, std::to_string(d));
    return GetInteger(b, k);
  }

 private:
  std::map<std::string, std::string> kv_;
};

// Device state of one partition (the reference's MeshData<Real>): a block set plus the
// field and flux arrays in pack layout.  Pointers are device pointers owned by the
// caller; first/count select the partition's blocks.
template <typename T>
struct MeshData {
  aph_context *ctx = nullptr;
  aph_blockset *blocks = nullptr;
  T *cons = nullptr, *prim = nullptr, *gamma = nullptr;
  T *flux[3] = {nullptr, nullptr, nullptr};
  T *grav = nullptr, *phi = nullptr;  // "gravity" (3 components) and "gravity.phi" (1)
  int32_t first = 0, count = -1;
  int32_t ndim = 1;
};

class StateDescriptor {
 public:
  explicit StateDescriptor(std::string name) : name_(std::move(name)) {}
  template <typename V>
  void AddParam(const std::string &k, V v) { params_[k] = v; }
  template <typename V>
  const V &Param(const std::string &k) const {
    auto it = params_.find(k);
    if (it == params_.end()) throw std::runtime_error("no param " + k);
    return *std::any_cast<V>(&it->second);
  }

 private:
  std::string name_;
  std::map<std::string, std::any> params_;
};

}  // namespace parthenon

// reference enums (src/main.hpp:27-61)
enum class Fluid { undefined, euler };
enum class Reconstruction { undefined, dc, plm, ppm };
enum class RiemannSolver { undefined, none, hlle, hllc, lhllc };
enum class Integrator { undefined, rk1, rk2, rk3, vl2 };
enum class Gravity { undefined, none, constant, monopole, poisson };

namespace Apophis {

using parthenon::MeshData;
using parthenon::Real;
using parthenon::StateDescriptor;
using parthenon::TaskStatus;

// Error raised where the reference calls PARTHENON_FAIL (configuration errors) and
// for a failing C ABI call (message from aph_last_error).
struct ApophisError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(aph_context *ctx, int rc) {
  if (rc != APH_OK) throw ApophisError(aph_last_error(ctx));
}

// hydro.cpp:535-744 -- the flux task; fills cons.flux of the partition.
template <Fluid fluid, Reconstruction recon, RiemannSolver rsolver>
TaskStatus CalculateFluxes(std::shared_ptr<MeshData<Real>> &md) {
  static_assert(fluid == Fluid::euler, "only the euler fluid exists in the reference");
  check(md->ctx, aph_calculate_fluxes(md->ctx, md->blocks, md->first, md->count,
                                      static_cast<int32_t>(recon),
                                      static_cast<int32_t>(rsolver), md->prim, md->gamma,
                                      md->flux[0], md->flux[1], md->flux[2]));
  return TaskStatus::complete;
}

using FluxFun_t =
    decltype(CalculateFluxes<Fluid::euler, Reconstruction::plm, RiemannSolver::hllc>);
using FluxFunKey_t = std::tuple<Fluid, Reconstruction, RiemannSolver>;

template <Fluid fluid, Reconstruction recon, RiemannSolver rsolver>
void add_flux_fun(std::map<FluxFunKey_t, FluxFun_t *> &flux_functions) {
  flux_functions[std::make_tuple(fluid, recon, rsolver)] =
      Apophis::CalculateFluxes<fluid, recon, rsolver>;
}

// hydro.cpp:422-533 (FillDerived)
inline void ConsToPrim(MeshData<Real> *md, const aph_eos &eos) {
  check(md->ctx, aph_cons_to_prim(md->ctx, md->blocks, md->first, md->count, &eos, md->cons,
                                  md->prim, md->gamma));
}

// hydro.cpp:351-420; dt_dev is device scratch for the reduction (1 double).
inline Real EstimateTimestep(MeshData<Real> *md, const aph_eos &eos, Real cfl, Real *dt_dev) {
  check(md->ctx, aph_estimate_timestep(md->ctx, md->blocks, md->first, md->count, &eos,
                                       md->prim, md->cons, cfl, dt_dev));
  Real dt = 0;
  check(md->ctx, aph_context_synchronize(md->ctx));
  if (hipMemcpy(&dt, dt_dev, sizeof(Real), hipMemcpyDeviceToHost) != hipSuccess)
    throw ApophisError("dt readback failed");
  return dt;
}

// upstream Parthenon Update::UpdateWithFluxDivergence (apophis_driver.cpp:79-82)
inline TaskStatus UpdateWithFluxDivergence(MeshData<Real> *md0, const Real *u1, Real gam0,
                                           Real gam1, Real beta_dt) {
  check(md0->ctx, aph_update_with_flux_divergence(md0->ctx, md0->blocks, md0->first,
                                                  md0->count, md0->cons, u1, md0->flux[0],
                                                  md0->flux[1], md0->flux[2], gam0, gam1,
                                                  beta_dt));
  return TaskStatus::complete;
}

// gravity.cpp:374-410 for a uniform acceleration
inline TaskStatus ApplyGravity(std::shared_ptr<MeshData<Real>> &md, const Real g[3],
                               const Real dt) {
  check(md->ctx, aph_apply_gravity(md->ctx, md->blocks, md->first, md->count, md->cons,
                                   md->prim, g, dt));
  return TaskStatus::complete;
}

// gravity.cpp:374-410, the reference's signature: the pack's "gravity" field (md->grav, 3
// components, pack layout), filled by the stage's gravity solver; prim is the stage-start
// primitive state
inline TaskStatus ApplyGravity(std::shared_ptr<MeshData<Real>> &md, const Real dt) {
  if (!md->grav) throw ApophisError("ApplyGravity: the pack has no gravity field");
  check(md->ctx, aph_apply_gravity_field(md->ctx, md->blocks, md->first, md->count,
                                         md->cons, md->prim, md->grav, dt));
  return TaskStatus::complete;
}

// boundary exchange + ApplyBoundaryConditions (apophis_driver.cpp:94, :103)
inline TaskStatus FillGhosts(std::shared_ptr<MeshData<Real>> &md) {
  check(md->ctx, aph_fill_ghosts(md->ctx, md->blocks, md->first, md->count, md->cons));
  return TaskStatus::complete;
}

// hydro.cpp:33-349, the options that concern the flux path.  Throws ApophisError with
// the reference's messages where it calls PARTHENON_FAIL.
inline std::shared_ptr<StateDescriptor> InitializeHydro(parthenon::ParameterInput *pin) {
  auto pkg = std::make_shared<StateDescriptor>("Hydro");
  pkg->AddParam<Real>("cfl", pin->GetOrAddReal("parthenon/time", "cfl", 0.3));
  if (pin->GetOrAddString("hydro", "fluid", "euler") != "euler")
    throw ApophisError("[Apophis]: Fluid not recognized. Exiting.");
  pkg->AddParam<Fluid>("fluid", Fluid::euler);
  pkg->AddParam<int>("nhydro", 5);
  const auto recon_str = pin->GetString("hydro", "reconstruction");
  int need = 3;
  Reconstruction recon = Reconstruction::undefined;
  if (recon_str == "dc") {
    recon = Reconstruction::dc;
    need = 1;
  } else if (recon_str == "plm") {
    recon = Reconstruction::plm;
    need = 2;
  } else if (recon_str == "ppm") {
    recon = Reconstruction::ppm;
    need = 3;
  } else {
    throw ApophisError("[Apophis]: Reconstruction not recognized. Exiting.");
  }
  if (pin->GetInteger("parthenon/mesh", "nghost") < need)
    throw ApophisError("[Apophis]: Not enough ghost zones for reconstruction. Exiting.");
  pkg->AddParam<Reconstruction>("reconstruction", recon);
  const auto riemann_str = pin->GetString("hydro", "riemann");
  RiemannSolver riemann = RiemannSolver::undefined;
  if (riemann_str == "hllc") {
    riemann = RiemannSolver::hllc;
  } else if (riemann_str == "hlle") {
    riemann = RiemannSolver::hlle;
  } else if (riemann_str == "lhllc") {
    riemann = RiemannSolver::lhllc;
  } else if (riemann_str == "none") {
    if (recon_str != "dc")
      throw ApophisError(
          "[Apophis]: 'none' Riemann solver only works with dc reconstruction. Exiting.");
    riemann = RiemannSolver::none;
  } else {
    throw ApophisError("[Apophis]: Riemann solver not recognized. Exiting.");
  }
  pkg->AddParam<RiemannSolver>("riemann_solver", riemann);
  std::map<FluxFunKey_t, FluxFun_t *> flux_functions{};
  add_flux_fun<Fluid::euler, Reconstruction::dc, RiemannSolver::hllc>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::plm, RiemannSolver::hllc>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::ppm, RiemannSolver::hllc>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::dc, RiemannSolver::hlle>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::plm, RiemannSolver::hlle>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::ppm, RiemannSolver::hlle>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::dc, RiemannSolver::lhllc>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::plm, RiemannSolver::lhllc>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::ppm, RiemannSolver::lhllc>(flux_functions);
  add_flux_fun<Fluid::euler, Reconstruction::dc, RiemannSolver::none>(flux_functions);
  FluxFun_t *flux_other = flux_functions.at(std::make_tuple(Fluid::euler, recon, riemann));
  FluxFun_t *flux_first = flux_other;
  const auto integ = pin->GetString("parthenon/time", "integrator");
  Integrator integrator = Integrator::undefined;
  if (integ == "rk1") {
    integrator = Integrator::rk1;
  } else if (integ == "rk2") {
    integrator = Integrator::rk2;
  } else if (integ == "rk3") {
    integrator = Integrator::rk3;
  } else if (integ == "vl2") {
    integrator = Integrator::vl2;
    flux_first = flux_functions.at(std::make_tuple(Fluid::euler, Reconstruction::dc, riemann));
  } else {
    throw ApophisError("[Apophis]: Integrator not recognized. Exiting.");
  }
  pkg->AddParam<Integrator>("integrator", integrator);
  pkg->AddParam<FluxFun_t *>("flux_first_stage", flux_first);
  pkg->AddParam<FluxFun_t *>("flux_other_stage", flux_other);
  aph_eos eos{};
  eos.gm1 = pin->GetOrAddReal("eos", "gm1", 0.6666667);
  eos.cv = pin->GetOrAddReal("eos", "cv", 1.5);
  const Real fl = std::sqrt(1024 * 1.17549435082228750797e-38);  // sqrt(1024*FLT_MIN)
  eos.dfloor = pin->GetOrAddReal("hydro", "dfloor", fl);
  eos.pfloor = pin->GetOrAddReal("hydro", "pfloor", fl);
  pkg->AddParam<aph_eos>("eos", eos);
  return pkg;
}

// ----------------------------------------------------------------------- gravity package
constexpr Real GRAVITY_G = 6.673e-8;  // constants.hpp:5

inline bool deck_true(const std::string &v) {
  std::string s;
  for (char c : v) s += static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s == "true" || s == "1";
}

// gravity.cpp:220-370 for the solvers this library runs.  Throws ApophisError with the
// reference's messages where it calls PARTHENON_FAIL.  Parameters:
//   "gravity_constant", "gravity" (the Gravity enum);
//   monopole: "nrbin" (gravity/nrbin, else the mesh's nx1: monopole_gravity.hpp:83-88),
//             "rmax" (gravity/rmax, else the distance to the farthest mesh corner);
//   poisson:  "max_iterations", "diagonal_alpha", "solver", "boundary_prolongation",
//             "flux_correct", "err_tol" as gravity.cpp:297-330 stores them, and the
//             solver's own limits from its gravity/solver_params block (Parthenon's MG /
//             BiCGSTAB read "residual_tolerance" and "max_iterations" there; the
//             reference's gravity/max_iterations is stored but read by nothing):
//             "residual_tolerance" (else err_tol) and "solver_max_iterations" (else
//             gravity/max_iterations).
inline std::shared_ptr<StateDescriptor> InitializeGravity(parthenon::ParameterInput *pin) {
  auto pkg = std::make_shared<StateDescriptor>("Gravity");
  const auto type = pin->GetOrAddString("gravity", "type", "none");
  pkg->AddParam<Real>("gravity_constant",
                      pin->GetOrAddReal("gravity", "gravity_constant", GRAVITY_G));
  Gravity gravity = Gravity::undefined;
  if (type == "none") {
    gravity = Gravity::none;
  } else if (type == "constant") {
    gravity = Gravity::constant;
  } else if (type == "monopole") {
    gravity = Gravity::monopole;
  } else if (type == "poisson") {
    gravity = Gravity::poisson;
  } else {
    throw ApophisError("[Apophis]: Gravity not recognized. Exiting.");
  }
  pkg->AddParam<Gravity>("gravity", gravity);
  if (gravity == Gravity::monopole) {
    int nrbin = pin->Has("gravity", "nrbin") ? pin->GetInteger("gravity", "nrbin") : -1;
    if (nrbin <= 0) nrbin = pin->GetInteger("parthenon/mesh", "nx1");
    Real edge[3];
    for (int a = 0; a < 3; ++a) {
      const std::string n = std::to_string(a + 1);
      const Real lo = a == 0 ? std::stod(pin->GetString("parthenon/mesh", "x1min"))
                             : pin->GetOrAddReal("parthenon/mesh", "x" + n + "min", 0.0);
      const Real hi = a == 0 ? std::stod(pin->GetString("parthenon/mesh", "x1max"))
                             : pin->GetOrAddReal("parthenon/mesh", "x" + n + "max", 0.0);
      edge[a] = std::max(std::abs(lo), std::abs(hi));
    }
    const Real default_rmax =
        std::sqrt(edge[0] * edge[0] + edge[1] * edge[1] + edge[2] * edge[2]);
    const Real rmax = pin->Has("gravity", "rmax") ? std::stod(pin->GetString("gravity", "rmax"))
                                                  : default_rmax;
    if (rmax <= 0.0)
      throw ApophisError("[Apophis]: gravity/rmax must be positive for monopole gravity.");
    pkg->AddParam<int>("nrbin", nrbin);
    pkg->AddParam<Real>("rmax", rmax);
  }
  if (gravity == Gravity::poisson) {
    const int max_it = pin->GetOrAddInteger("gravity", "max_iterations", 10000);
    pkg->AddParam<int>("max_iterations", max_it);
    pkg->AddParam<Real>("diagonal_alpha", pin->GetOrAddReal("gravity", "diagonal_alpha", 0.0));
    const auto solver = pin->GetOrAddString("gravity", "solver", "MG");
    if (solver != "MG" && solver != "MGBiCGSTAB")
      throw ApophisError("Unknown solver type " + solver + ".");
    pkg->AddParam<std::string>("solver", solver);
    const auto prolong = pin->GetOrAddString("gravity", "boundary_prolongation", "Linear");
    if (prolong != "Linear" && prolong != "Quadratic" && prolong != "Constant")
      throw ApophisError("Unknown prolongation method for gravity boundaries.");
    pkg->AddParam<std::string>("boundary_prolongation", prolong);
    pkg->AddParam<bool>("flux_correct",
                        deck_true(pin->GetOrAddString("gravity", "flux_correct", "false")));
    if (deck_true(pin->GetOrAddString("gravity", "set_flux_boundary", "false")))
      throw ApophisError("gravity/set_flux_boundary = true is not built (the fine/coarse flux "
                         "boundary of PoissonEquation)");
    const Real err_tol = pin->GetOrAddReal("gravity", "err_tol", 1.0e-8);
    pkg->AddParam<Real>("err_tol", err_tol);
    pkg->AddParam<Real>("residual_tolerance",
                        pin->GetOrAddReal("gravity/solver_params", "residual_tolerance", err_tol));
    pkg->AddParam<int>("solver_max_iterations",
                       pin->GetOrAddInteger("gravity/solver_params", "max_iterations", max_it));
  }
  return pkg;
}

// A gravity solver (gravity.hpp:13-20) fills the "gravity" field (3 components, pack
// layout) and, where it has one, "gravity.phi" (1 component), once per stage before the
// flux tasks (apophis_driver.cpp:64-66).  The reference adds its PreComputeTasks and
// AddTasks to the stage's task list per partition; AddTasks here runs both, in order,
// over the mesh's partitions.
struct GravitySolver {
  virtual ~GravitySolver() = default;
  virtual TaskStatus AddTasks(const std::vector<std::shared_ptr<MeshData<Real>>> &parts,
                              Real *grav, Real *phi) = 0;
  // whether the solver writes the field (the constant solver's field is the problem
  // generator's, and "none" has none)
  virtual bool Solves() const { return true; }
};

struct NoneGravitySolver : GravitySolver {  // none_gravity.hpp
  TaskStatus AddTasks(const std::vector<std::shared_ptr<MeshData<Real>>> &, Real *,
                      Real *) override {
    return TaskStatus::complete;
  }
  bool Solves() const override { return false; }
};

struct ConstantGravitySolver : NoneGravitySolver {};  // constant_gravity.hpp

// monopole_gravity.hpp:15-305 on the device: the tasks of aph_monopole_*.
class MonopoleGravitySolver : public GravitySolver {
 public:
  MonopoleGravitySolver(aph_context *ctx, int ndim, int nrbin, Real rmax, Real G) : ctx_(ctx) {
    check(ctx, aph_monopole_create(ctx, ndim, nrbin, rmax, G, &m_));
  }
  ~MonopoleGravitySolver() override { aph_monopole_destroy(ctx_, m_); }
  MonopoleGravitySolver(const MonopoleGravitySolver &) = delete;
  MonopoleGravitySolver &operator=(const MonopoleGravitySolver &) = delete;
  TaskStatus AddTasks(const std::vector<std::shared_ptr<MeshData<Real>>> &parts, Real *grav,
                      Real *phi) override {
    const auto &all = parts.front();
    check(ctx_, aph_monopole_zero_radial_bins(ctx_, m_, all->blocks, all->prim));  // :107-112
    for (const auto &p : parts)                                                    // :114-153
      check(ctx_, aph_monopole_accumulate_mass(ctx_, m_, p->blocks, p->first, p->count,
                                               p->prim));
    check(ctx_, aph_monopole_build_radial_tables(ctx_, m_));                        // :155-213
    for (const auto &p : parts)                                                    // :215-302
      check(ctx_, aph_monopole_gravity_field(ctx_, m_, p->blocks, p->first, p->count, grav,
                                             phi));
    return TaskStatus::complete;
  }

 private:
  aph_context *ctx_;
  aph_monopole *m_ = nullptr;
};

// poisson_gravity.hpp on the device, on one uniform lattice (aph_poisson_create): one
// solve over the whole mesh per stage.
class PoissonGravitySolver : public GravitySolver {
 public:
  PoissonGravitySolver(aph_context *ctx, const aph_blockset *bs, Real G, Real alpha, Real tol,
                       int max_it)
      : ctx_(ctx) {
    check(ctx, aph_poisson_create(ctx, bs, G, alpha, tol, max_it, &ps_));
  }
  ~PoissonGravitySolver() override { aph_poisson_destroy(ctx_, ps_); }
  PoissonGravitySolver(const PoissonGravitySolver &) = delete;
  PoissonGravitySolver &operator=(const PoissonGravitySolver &) = delete;
  TaskStatus AddTasks(const std::vector<std::shared_ptr<MeshData<Real>>> &parts, Real *grav,
                      Real *phi) override {
    check(ctx_, aph_poisson_solve(ctx_, ps_, parts.front()->prim, grav, phi));
    return TaskStatus::complete;
  }
  // iterations and relative residual of the last solve
  void Info(int *iterations, Real *residual) const {
    int32_t it = 0;
    check(ctx_, aph_poisson_get_info(ctx_, ps_, nullptr, &it, residual));
    *iterations = it;
  }

 private:
  aph_context *ctx_;
  aph_poisson *ps_ = nullptr;
};

// The package's solver bound to a device block set (the reference keeps it in the
// package as "gravity_solver", built from the deck; here it needs the device state).
// Multilevel meshes take aph_poisson_create_tree, which this uniform-lattice adapter does
// not build.
inline std::shared_ptr<GravitySolver> MakeGravitySolver(const StateDescriptor &pkg,
                                                        aph_context *ctx,
                                                        const aph_blockset *bs, int ndim) {
  switch (pkg.Param<Gravity>("gravity")) {
    case Gravity::monopole:
      return std::make_shared<MonopoleGravitySolver>(ctx, ndim, pkg.Param<int>("nrbin"),
                                                     pkg.Param<Real>("rmax"),
                                                     pkg.Param<Real>("gravity_constant"));
    case Gravity::poisson:
      if (pkg.Param<std::string>("boundary_prolongation") == "Quadratic")
        throw ApophisError("gravity/boundary_prolongation = Quadratic: the quadratic form is "
                           "not built");
      return std::make_shared<PoissonGravitySolver>(
          ctx, bs, pkg.Param<Real>("gravity_constant"), pkg.Param<Real>("diagonal_alpha"),
          pkg.Param<Real>("residual_tolerance"), pkg.Param<int>("solver_max_iterations"));
    case Gravity::constant:
      return std::make_shared<ConstantGravitySolver>();
    default:
      return std::make_shared<NoneGravitySolver>();
  }
}

}  // namespace Apophis
