"""Host-side mirror of the reference's Hydro package and stage driver.

This is the Python face of the reference interface around the flux path:

* enums with the reference's values (``src/main.hpp:32-53``);
* ``initialize_hydro(pin)`` -- option parsing, ghost check and the flux-function
  registry keyed by (fluid, reconstruction, riemann) (``src/hydro/hydro.cpp:33-349``,
  ``src/hydro/hydro.hpp:26-37``); invalid options raise :class:`ApophisError`
  where the reference calls ``PARTHENON_FAIL``;
* task functions with the reference's shapes, each returning ``TaskStatus`` or a
  value like the reference: ``calculate_fluxes`` (``hydro.cpp:535-744``),
  ``cons_to_prim`` (``:422-533``), ``estimate_timestep`` (``:351-420``),
  ``update_with_flux_divergence`` (upstream Parthenon), ``apply_gravity``
  (``src/gravity/gravity.cpp:374-410``), ``fill_ghosts`` (boundary exchange + BCs),
  ``global_quant_hst`` (``hydro.cpp:746-771``);
* ``ApophisDriver`` -- the RK2 stage sequence of ``ApophisDriver::MakeTaskCollection``
  (``src/apophis_driver.cpp:27-130``);
* ``FusedDriver`` -- the batched form: one fused launch per stage, device-resident
  dt, optionally replayed from a captured hipGraph.

All compute goes through the C ABI (``_lib``); torch only provides device memory and
the stream.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import enum

import numpy as np
import torch

from . import _lib
from .mesh import BlockSet, EOS_DTYPE, make_eos


class Fluid(enum.IntEnum):
    undefined = 0
    euler = 1


class Reconstruction(enum.IntEnum):
    undefined = 0
    dc = 1
    plm = 2
    ppm = 3


class RiemannSolver(enum.IntEnum):
    undefined = 0
    none = 1
    hlle = 2
    hllc = 3
    lhllc = 4


class Integrator(enum.IntEnum):
    undefined = 0
    rk1 = 1
    rk2 = 2
    rk3 = 3
    vl2 = 4


class TaskStatus(enum.Enum):
    complete = 0
    incomplete = 1
    fail = 2


class ApophisError(RuntimeError):
    """Raised where the reference aborts with PARTHENON_FAIL."""


# upstream "rk2" low-storage coefficients
# Upstream Parthenon low-storage integrators (stage k: u0 <- gam0 u0 + gam1 u1 +
# beta dt div F(u0), with u1 the step's starting state).
INTEGRATORS = {
    "rk1": {"gam0": (0.0,), "gam1": (1.0,), "beta": (1.0,), "nstages": 1},
    "rk2": {"gam0": (0.0, 0.5), "gam1": (1.0, 0.5), "beta": (1.0, 0.5), "nstages": 2},
    "rk3": {"gam0": (0.0, 0.25, 2.0 / 3.0), "gam1": (1.0, 0.75, 1.0 / 3.0),
            "beta": (1.0, 0.25, 2.0 / 3.0), "nstages": 3},
    "vl2": {"gam0": (0.0, 0.0), "gam1": (1.0, 1.0), "beta": (0.5, 1.0), "nstages": 2},
}
RK2 = INTEGRATORS["rk2"]


def _ptr(t):
    if t is None:
        return None
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    return ctypes.c_void_p(t.data_ptr())


# ------------------------------------------------------------------ package
class StateDescriptor:
    def __init__(self, name):
        self.name = name
        self.params = {}

    def add_param(self, key, value):
        self.params[key] = value

    def param(self, key):
        return self.params[key]


def _get(pin, block, key, default=None, required=False):
    blk = pin.get(block, {})
    if key in blk:
        return blk[key]
    if required:
        raise ApophisError(f"[Apophis]: missing required parameter {block}/{key}")
    return default


FLUX_FUNCTIONS = {}


def add_flux_fun(fluid, recon, rsolver):
    """Register the task function for one (fluid, recon, riemann) key (hydro.hpp:33-37)."""
    FLUX_FUNCTIONS[(fluid, recon, rsolver)] = _make_flux_task(recon, rsolver)


def _make_flux_task(recon, rsolver):
    def calculate_fluxes(md):
        md.ctx.check(_lib.lib().aph_calculate_fluxes(
            md.ctx.h, md.dbs, md.first, md.count, int(recon), int(rsolver), _ptr(md.prim),
            _ptr(md.gamma), _ptr(md.flux[0]), _ptr(md.flux[1]), _ptr(md.flux[2])))
        return TaskStatus.complete

    calculate_fluxes.__name__ = f"CalculateFluxes<euler,{recon.name},{rsolver.name}>"
    calculate_fluxes.key = (Fluid.euler, recon, rsolver)
    return calculate_fluxes


# hydro.cpp:103-112 registers these ten combinations.
for _r in (Reconstruction.dc, Reconstruction.plm, Reconstruction.ppm):
    for _s in (RiemannSolver.hllc, RiemannSolver.hlle, RiemannSolver.lhllc):
        add_flux_fun(Fluid.euler, _r, _s)
add_flux_fun(Fluid.euler, Reconstruction.dc, RiemannSolver.none)


def initialize_hydro(pin):
    """Mirror of Apophis::InitializeHydro (hydro.cpp:33-349) for the flux path.

    pin: nested dict {block: {key: value}} using the reference's input-deck names."""
    pkg = StateDescriptor("Hydro")
    pkg.add_param("cfl", float(_get(pin, "parthenon/time", "cfl", 0.3)))
    fluid_str = _get(pin, "hydro", "fluid", "euler")
    if fluid_str != "euler":
        raise ApophisError("[Apophis]: Fluid not recognized. Exiting.")
    pkg.add_param("fluid", Fluid.euler)
    pkg.add_param("nhydro", 5)
    recon_str = _get(pin, "hydro", "reconstruction", required=True)
    need = {"dc": 1, "plm": 2, "ppm": 3}
    if recon_str not in need:
        raise ApophisError("[Apophis]: Reconstruction not recognized. Exiting.")
    recon = Reconstruction[recon_str]
    nghost = int(_get(pin, "parthenon/mesh", "nghost", required=True))
    if nghost < need[recon_str]:
        raise ApophisError("[Apophis]: Not enough ghost zones for reconstruction. Exiting.")
    pkg.add_param("reconstruction", recon)
    riemann_str = _get(pin, "hydro", "riemann", required=True)
    if riemann_str == "none" and recon_str != "dc":
        raise ApophisError(
            "[Apophis]: 'none' Riemann solver only works with dc reconstruction. Exiting.")
    if riemann_str not in ("hllc", "hlle", "lhllc", "none"):
        raise ApophisError("[Apophis]: Riemann solver not recognized. Exiting.")
    rsolver = RiemannSolver[riemann_str]
    pkg.add_param("riemann_solver", rsolver)
    flux_other = FLUX_FUNCTIONS[(Fluid.euler, recon, rsolver)]
    integ_str = _get(pin, "parthenon/time", "integrator", required=True)
    if integ_str not in ("rk1", "rk2", "rk3", "vl2"):
        raise ApophisError("[Apophis]: Integrator not recognized. Exiting.")
    flux_first = flux_other
    if integ_str == "vl2":
        flux_first = FLUX_FUNCTIONS[(Fluid.euler, Reconstruction.dc, rsolver)]
    pkg.add_param("integrator", Integrator[integ_str])
    pkg.add_param("flux_first_stage", flux_first)
    pkg.add_param("flux_other_stage", flux_other)
    # passive scalars (hydro.cpp:141-177): the ONe composition (6 species densities and
    # Ye), then two fields per level set; "nscalars" adds generic scalars (an extension:
    # the reference has no such key, the KH workload's marker scalar needs it)
    nscalars = 0
    comp_str = _get(pin, "hydro", "composition", "unset")
    ncomp, labels, abar, zbar, ebind = 0, [], [], [], []
    if comp_str == "ONe":
        nscalars += 7
        ncomp = 6
        labels = ["alpha", "C12", "O16", "Ne20", "IME", "IGE"]
        abar = [4.0, 12.0, 16.0, 20.0, 30.0, 56.0]
        zbar = [2.0, 6.0, 8.0, 10.0, 15.0, 28.0]
        ebind = [6.8266e+18, 7.41121e+18, 7.69691e+18, 7.74994e+18, 8.17906e+18, 8.34e+18]
    pkg.add_param("ncomp", ncomp)
    pkg.add_param("comp_labels", labels)
    pkg.add_param("comp_abar", abar)
    pkg.add_param("comp_zbar", zbar)
    pkg.add_param("comp_ebind", ebind)
    nlset = int(_get(pin, "hydro", "nlset", 0))
    if nlset > 0 and ncomp < 1:
        raise ApophisError(
            "[Apophis]: Levelset is enabled but no composition is set. Exiting.")
    nscalars += 2 * nlset
    pkg.add_param("nlset", nlset)
    nscalars += int(_get(pin, "hydro", "nscalars", 0))
    pkg.add_param("nscalars", nscalars)
    cons_labels = ["Density", "MomentumDensity1", "MomentumDensity2", "MomentumDensity3",
                   "TotalEnergyDensity"]
    prim_labels = ["Density", "Velocity1", "Velocity2", "Velocity3", "Pressure"]
    extra = labels + (["ye"] if ncomp > 0 else [])
    for i in range(nlset):
        extra += [f"lset{i}", f"xfuel{i}"]
    cons_labels += ["scalar_density_" + x for x in extra]
    prim_labels += ["scalar_" + x for x in extra]
    pkg.add_param("cons_labels", cons_labels)
    pkg.add_param("prim_labels", prim_labels)
    # refinement criterion (hydro.cpp:303-333)
    refine_str = _get(pin, "refinement", "type", "unset")
    pkg.add_param("refinement/type", refine_str)
    if refine_str == "pressure_gradient":
        thr = float(_get(pin, "refinement", "threshold_pressure_gradient", 0.0))
        if not thr > 0.0:
            raise ApophisError("Make sure to set refinement/threshold_pressure_gradient >0.")
        pkg.add_param("refinement/threshold_pressure_gradient", thr)
        pkg.add_param("refinement/criterion", (REFINE_CRITERIA[refine_str], thr, 0.0))
    elif refine_str == "xyvelocity_gradient":
        thr = float(_get(pin, "refinement", "threshold_xyvelocity_gradient", 0.0))
        if not thr > 0.0:
            raise ApophisError("Make sure to set refinement/threshold_xyvelocity_gradient >0.")
        pkg.add_param("refinement/threshold_xyvelocity_gradient", thr)
        pkg.add_param("refinement/criterion", (REFINE_CRITERIA[refine_str], thr, 0.0))
    elif refine_str == "maxdensity":
        lo = float(_get(pin, "refinement", "maxdensity_deref_below", 0.0))
        hi = float(_get(pin, "refinement", "maxdensity_refine_above", 0.0))
        if not lo > 0.0:
            raise ApophisError("Make sure to set refinement/maxdensity_deref_below > 0.")
        if not hi > 0.0:
            raise ApophisError("Make sure to set refinement/maxdensity_refine_above > 0.")
        if not lo < hi:
            raise ApophisError("Make sure to set refinement/maxdensity_deref_below < "
                               "refinement/maxdensity_refine_above")
        pkg.add_param("refinement/maxdensity_deref_below", lo)
        pkg.add_param("refinement/maxdensity_refine_above", hi)
        pkg.add_param("refinement/criterion", (REFINE_CRITERIA[refine_str], lo, hi))
    eos_type = _get(pin, "eos", "type", "ideal")
    if eos_type == "helm":
        if ncomp < 1:
            raise ApophisError("[Apophis]: Helmholtz EOS is enabled but no composition is "
                               "set. Exiting.")
        raise ApophisError("[Apophis]: Helmholtz EOS (singularity-eos tables) is not "
                           "available in this build.")
    if eos_type != "ideal":
        raise ApophisError("[Apophis]: EOS not recognized. Exiting.")
    floor = float(np.sqrt(1024.0 * np.finfo(np.float32).tiny))
    pkg.add_param("eos", make_eos(float(_get(pin, "eos", "gm1", 0.6666667)),
                                  float(_get(pin, "eos", "cv", 1.5)),
                                  float(_get(pin, "hydro", "dfloor", floor)),
                                  float(_get(pin, "hydro", "pfloor", floor))))
    return pkg


# ------------------------------------------------------------------ mesh data
class DeviceBlockSet:
    """A BlockSet uploaded through aph_blockset_create."""

    def __init__(self, ctx, bs: BlockSet):
        self.ctx = ctx
        self.bs = bs
        self.h = ctypes.c_void_p()
        ctx.check(_lib.lib().aph_blockset_create(
            ctx.h, bs.blocks.ctypes.data_as(ctypes.c_void_p), bs.nblocks, bs.ndim,
            bs.nscalars, ctypes.byref(self.h)))

    def __del__(self):
        try:
            if self.h:
                _lib.lib().aph_blockset_destroy(self.ctx.h, self.h)
        except Exception:
            pass


class MeshData:
    """Device state of one partition: cons (u0), prim, gamma and face fluxes.

    Field tensors use the pack layout of include/apophis_hydro.h."""

    def __init__(self, ctx, bs: BlockSet, cons=None, device="cuda", first=0, count=-1,
                 dbs=None):
        self.ctx = ctx
        self.bs = bs
        self.device_bs = dbs or DeviceBlockSet(ctx, bs)
        self.dbs = self.device_bs.h
        self.first, self.count = first, count
        n = bs.field_size(bs.nvar)
        self.cons = torch.zeros(n, dtype=torch.float64, device=device)
        if cons is not None:
            self.cons.copy_(torch.from_numpy(np.ascontiguousarray(cons)))
        self.prim = torch.zeros_like(self.cons)
        self.gamma = torch.zeros(bs.field_size(2), dtype=torch.float64, device=device)
        self.flux = [torch.zeros_like(self.cons) if d < bs.ndim else None for d in range(3)]

    def pack(self, first, count):
        """A view of a contiguous range of blocks (the reference's pack_size)."""
        md = MeshData.__new__(MeshData)
        md.__dict__.update(self.__dict__)
        md.first, md.count = first, count
        return md


# ------------------------------------------------------------------ task functions
def cons_to_prim(md, eos):
    md.ctx.check(_lib.lib().aph_cons_to_prim(md.ctx.h, md.dbs, md.first, md.count,
                                             eos.ctypes.data_as(ctypes.c_void_p),
                                             _ptr(md.cons), _ptr(md.prim), _ptr(md.gamma)))
    return TaskStatus.complete


def update_with_flux_divergence(md0, u1, gam0, gam1, beta_dt):
    md0.ctx.check(_lib.lib().aph_update_with_flux_divergence(
        md0.ctx.h, md0.dbs, md0.first, md0.count, _ptr(md0.cons), _ptr(u1),
        _ptr(md0.flux[0]), _ptr(md0.flux[1]), _ptr(md0.flux[2]), gam0, gam1, beta_dt))
    return TaskStatus.complete


def apply_gravity(md, g, dt):
    garr = (ctypes.c_double * 3)(*g)
    md.ctx.check(_lib.lib().aph_apply_gravity(md.ctx.h, md.dbs, md.first, md.count,
                                              _ptr(md.cons), _ptr(md.prim), garr, dt))
    return TaskStatus.complete


def fill_ghosts(md):
    md.ctx.check(_lib.lib().aph_fill_ghosts(md.ctx.h, md.dbs, md.first, md.count,
                                            _ptr(md.cons)))
    return TaskStatus.complete


def estimate_timestep(md, eos, cfl, out=None):
    """Returns cfl * min(dx/(|v|+c)) as a 1-element device tensor."""
    out = out if out is not None else torch.empty(1, dtype=torch.float64, device=md.cons.device)
    md.ctx.check(_lib.lib().aph_estimate_timestep(
        md.ctx.h, md.dbs, md.first, md.count, eos.ctypes.data_as(ctypes.c_void_p),
        _ptr(md.prim), _ptr(md.cons), cfl, _ptr(out)))
    return out


# aph_refine_criterion codes, keyed by the reference's refinement/type names
REFINE_CRITERIA = {"pressure_gradient": 1, "xyvelocity_gradient": 2, "maxdensity": 3}


def refinement_tag(md, pkg):
    """The package's CheckRefinementBlock over every block of md (Refinement::Tag,
    apophis_driver.cpp:120-127): (tags, values) device tensors, tags -1 / 0 / 1 =
    derefine / same / refine.  prim must be current (after ConsToPrim)."""
    crit, a, b = pkg.param("refinement/criterion")
    n = md.bs.nblocks
    tags = torch.zeros(n, dtype=torch.int32, device=md.cons.device)
    vals = torch.zeros(n, dtype=torch.float64, device=md.cons.device)
    params = (ctypes.c_double * 2)(a, b)
    assert tags.is_cuda and tags.is_contiguous()
    md.ctx.check(_lib.lib().aph_refinement_tag(md.ctx.h, md.dbs, crit, params, _ptr(md.prim),
                                               ctypes.c_void_p(tags.data_ptr()), _ptr(vals)))
    return tags, vals


def global_quant_hst(md, var, out=None):
    out = out if out is not None else torch.empty(1, dtype=torch.float64, device=md.cons.device)
    md.ctx.check(_lib.lib().aph_global_sum(md.ctx.h, md.dbs, _ptr(md.cons), var, _ptr(out)))
    return out


# ------------------------------------------------------------------ gravity
GRAVITY_G = 6.673e-8  # constants.hpp:5 (cm^3 g^-1 s^-2)


class Gravity(enum.IntEnum):  # main.hpp:55-61
    undefined = 0
    none = 1
    constant = 2
    monopole = 3
    poisson = 4


def initialize_gravity(pin):
    """InitializeGravity (gravity.cpp:220-370) for the solvers this library runs: "none",
    "constant" (the field is the driver's g), "monopole" (nrbin: gravity/nrbin, else
    the mesh's nx1, monopole_gravity.hpp:83-88; rmax: gravity/rmax, else the distance to
    the farthest mesh corner, gravity.cpp:262-276) and "poisson" (gravity/max_iterations,
    diagonal_alpha, solver, boundary_prolongation, err_tol, flux_correct, and the
    solver_params block's residual_tolerance and max_iterations, the limits the solver
    runs with; PoissonGravitySolver runs it)."""
    pkg = StateDescriptor("Gravity")
    gtype = _get(pin, "gravity", "type", "none")
    pkg.add_param("gravity_constant", float(_get(pin, "gravity", "gravity_constant",
                                                 GRAVITY_G)))
    if gtype not in ("none", "constant", "monopole", "poisson"):
        raise ApophisError("[Apophis]: Gravity not recognized. Exiting.")
    pkg.add_param("gravity", Gravity[gtype])
    if gtype == "poisson":  # gravity.cpp:297-330 and the solver block's tolerance
        pkg.add_param("max_iterations", int(_get(pin, "gravity", "max_iterations", 10000)))
        pkg.add_param("diagonal_alpha", float(_get(pin, "gravity", "diagonal_alpha", 0.0)))
        solver = _get(pin, "gravity", "solver", "MG")
        if solver not in ("MG", "MGBiCGSTAB"):
            raise ApophisError(f"Unknown solver type {solver}.")
        pkg.add_param("solver", solver)
        prolong = _get(pin, "gravity", "boundary_prolongation", "Linear")
        if prolong not in ("Linear", "Quadratic", "Constant"):
            raise ApophisError("Unknown prolongation method for gravity boundaries.")
        pkg.add_param("boundary_prolongation", prolong)
        fc = _get(pin, "gravity", "flux_correct", False)  # poisson_gravity.hpp:41
        if isinstance(fc, str):
            fc = fc.strip().lower() == "true"
        pkg.add_param("flux_correct", bool(fc))
        sfb = _get(pin, "gravity", "set_flux_boundary", False)  # :42, SetFluxBoundaries
        if (sfb.strip().lower() == "true") if isinstance(sfb, str) else bool(sfb):
            raise ApophisError("gravity/set_flux_boundary = true is not built (the fine/"
                               "coarse flux boundary of PoissonEquation)")
        err_tol = float(_get(pin, "gravity", "err_tol", 1.0e-8))
        pkg.add_param("err_tol", err_tol)
        # the solver's own limits: Parthenon's MG / BiCGSTAB read them from the
        # gravity/solver_params block (the reference stores gravity/max_iterations, :305-306,
        # but nothing reads it), else gravity/err_tol and gravity/max_iterations
        pkg.add_param("residual_tolerance",
                      float(_get(pin, "gravity/solver_params", "residual_tolerance", err_tol)))
        pkg.add_param("solver_max_iterations",
                      int(_get(pin, "gravity/solver_params", "max_iterations",
                               pkg.param("max_iterations"))))
    if gtype == "monopole":
        nrbin = int(_get(pin, "gravity", "nrbin", -1))
        if nrbin <= 0:
            nrbin = int(_get(pin, "parthenon/mesh", "nx1", required=True))
        edge = []
        for a in "123":
            lo = float(_get(pin, "parthenon/mesh", f"x{a}min", 0.0, required=(a == "1")))
            hi = float(_get(pin, "parthenon/mesh", f"x{a}max", 0.0, required=(a == "1")))
            edge.append(max(abs(lo), abs(hi)))
        default_rmax = float(np.sqrt(edge[0] * edge[0] + edge[1] * edge[1] + edge[2] * edge[2]))
        rmax = float(_get(pin, "gravity", "rmax", default_rmax))
        if rmax <= 0.0:
            raise ApophisError("[Apophis]: gravity/rmax must be positive for monopole gravity.")
        pkg.add_param("nrbin", nrbin)
        pkg.add_param("rmax", rmax)
    return pkg


class MonopoleGravitySolver:
    """MonopoleGravitySolver (monopole_gravity.hpp:15-305) on the device: AddTasks'
    sequence (:47-80) over the partitions of a mesh, writing the "gravity" field (3
    components) and, if given, "gravity.phi" (1 component), both in the pack layout."""

    def __init__(self, ctx, ndim, nrbin, rmax, gravity_constant):
        self.ctx = ctx
        self.ndim, self.nrbin = int(ndim), int(nrbin)
        self.rmax, self.G = float(rmax), float(gravity_constant)
        self.h = ctypes.c_void_p()
        ctx.check(_lib.lib().aph_monopole_create(ctx.h, self.ndim, self.nrbin, self.rmax,
                                                 self.G, ctypes.byref(self.h)))

    @classmethod
    def from_package(cls, ctx, pkg, ndim):
        return cls(ctx, ndim, pkg.param("nrbin"), pkg.param("rmax"),
                   pkg.param("gravity_constant"))

    def __del__(self):
        h = getattr(self, "h", None)
        try:
            if h and h.value:
                _lib.lib().aph_monopole_destroy(self.ctx.h, h)
                self.h = ctypes.c_void_p()
        except Exception:  # interpreter shutdown: the module may be gone already
            pass

    def zero_radial_bins(self, md):
        """ZeroRadialBins (:107-112); also fixes the scale of the sums from every block's
        primitive density (md.prim)."""
        self.ctx.check(_lib.lib().aph_monopole_zero_radial_bins(self.ctx.h, self.h, md.dbs,
                                                                _ptr(md.prim)))
        return TaskStatus.complete

    def accumulate_mass(self, md):
        """AccumulateMass (:114-153) over md's blocks."""
        self.ctx.check(_lib.lib().aph_monopole_accumulate_mass(
            self.ctx.h, self.h, md.dbs, md.first, md.count, _ptr(md.prim)))
        return TaskStatus.complete

    def build_radial_tables(self):
        """BuildRadialTables (:155-213)."""
        self.ctx.check(_lib.lib().aph_monopole_build_radial_tables(self.ctx.h, self.h))
        return TaskStatus.complete

    def compute_gravity_vector_direct(self, md, grav, phi=None):
        """ComputeGravityVectorDirect (:215-302) over md's blocks."""
        self.ctx.check(_lib.lib().aph_monopole_gravity_field(
            self.ctx.h, self.h, md.dbs, md.first, md.count, _ptr(grav),
            _ptr(phi) if phi is not None else None))
        return TaskStatus.complete

    def add_tasks(self, parts, grav, phi=None):
        """The task list of one stage: bins zeroed once, masses per partition, the
        tables once (after the all-reduce, here a single device's sums), the field per
        partition."""
        self.zero_radial_bins(parts[0])
        for p in parts:
            self.accumulate_mass(p)
        self.build_radial_tables()
        for p in parts:
            self.compute_gravity_vector_direct(p, grav, phi)
        return TaskStatus.complete

    def tables(self):
        """(m_enc, phi_rad) of the last pass, each nrbin + 1 values (synchronous)."""
        m_enc, phi_rad = np.zeros(self.nrbin + 1), np.zeros(self.nrbin + 1)
        self.ctx.check(_lib.lib().aph_monopole_get_tables(
            self.ctx.h, self.h, m_enc.ctypes.data_as(ctypes.c_void_p),
            phi_rad.ctypes.data_as(ctypes.c_void_p)))
        return m_enc, phi_rad


POISSON_PROLONGATION = {"Linear": 0, "Constant": 1}  # aph_poisson_prolongation


class PoissonGravitySolver:
    """PoissonGravitySolver (poisson_gravity.hpp:390-625) on the device, for a 3D block set
    with physical boundaries on every face: the multipole coefficients, the right-hand
    side, the solve (to the relative residual tolerance) and the "gravity" / "gravity.phi"
    fields.  mesh None: the blocks tile a box in one lattice (aph_poisson_create); mesh an
    amr.AmrMesh whose block set is dbs's: the composite operator of the multilevel mesh
    (aph_poisson_create_tree: BiCGSTAB with a multigrid V-cycle over coarser meshes of
    the tree) with the deck's flux_correct and boundary_prolongation ("Linear" or
    "Constant"; "Quadratic" is not built and is refused)."""

    def __init__(self, ctx, dbs, gravity_constant, diagonal_alpha=0.0, tolerance=1e-8,
                 max_iterations=10000, flux_correct=False, prolongation="Linear", mesh=None):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        if mesh is None:
            ctx.check(_lib.lib().aph_poisson_create(ctx.h, dbs, float(gravity_constant),
                                                    float(diagonal_alpha), float(tolerance),
                                                    int(max_iterations),
                                                    ctypes.byref(self.h)))
            return
        if prolongation not in POISSON_PROLONGATION:
            raise ApophisError(f"gravity/boundary_prolongation {prolongation}: Linear or "
                               "Constant on a multilevel mesh (the quadratic form is not "
                               "built)")
        vp = ctypes.c_void_p
        self._tree = (np.array(mesh.nbase, dtype=np.int32),
                      np.ascontiguousarray(mesh.level, dtype=np.int32),
                      np.ascontiguousarray(mesh.loc, dtype=np.int64))
        nbase, level, loc = self._tree
        ctx.check(_lib.lib().aph_poisson_create_tree(
            ctx.h, dbs, nbase.ctypes.data_as(vp), level.ctypes.data_as(vp),
            loc.ctypes.data_as(vp), int(bool(flux_correct)),
            POISSON_PROLONGATION[prolongation], float(gravity_constant),
            float(diagonal_alpha), float(tolerance), int(max_iterations),
            ctypes.byref(self.h)))

    @classmethod
    def from_package(cls, ctx, pkg, dbs, mesh=None):
        return cls(ctx, dbs, pkg.param("gravity_constant"), pkg.param("diagonal_alpha"),
                   pkg.param("residual_tolerance"), pkg.param("solver_max_iterations"),
                   pkg.param("flux_correct"), pkg.param("boundary_prolongation"), mesh=mesh)

    def __del__(self):
        h = getattr(self, "h", None)
        try:
            if h and h.value:
                _lib.lib().aph_poisson_destroy(self.ctx.h, h)
                self.h = ctypes.c_void_p()
        except Exception:  # interpreter shutdown: the module may be gone already
            pass

    def solve(self, md, grav, phi=None):
        """The stage's tasks on md.prim (every block): grav and phi filled."""
        self.ctx.check(_lib.lib().aph_poisson_solve(self.ctx.h, self.h, _ptr(md.prim),
                                                    _ptr(grav),
                                                    _ptr(phi) if phi is not None else None))
        return TaskStatus.complete

    def info(self):
        """(nine multipole coefficients, iterations, relative residual) of the last solve."""
        c = (ctypes.c_double * 9)()
        it = ctypes.c_int32()
        res = ctypes.c_double()
        self.ctx.check(_lib.lib().aph_poisson_get_info(self.ctx.h, self.h, c,
                                                       ctypes.byref(it), ctypes.byref(res)))
        return np.array(c[:]), int(it.value), float(res.value)


def apply_gravity_field(md, grav, dt):
    """ApplyGravity (gravity.cpp:374-410) with a per-cell field (3 components)."""
    md.ctx.check(_lib.lib().aph_apply_gravity_field(md.ctx.h, md.dbs, md.first, md.count,
                                                    _ptr(md.cons), _ptr(md.prim),
                                                    _ptr(grav), dt))
    return TaskStatus.complete


# ------------------------------------------------------------------ drivers
class ApophisDriver:
    """Stage sequence of ApophisDriver::MakeTaskCollection, plain form, for the
    package's integrator (rk1 / rk2 / rk3 / vl2).

    pack_size: blocks per partition (None = all blocks in one pack, the default
    Parthenon partitioning); each task is issued once per partition.  partitions: an
    explicit list of (first_block, nblocks) ranges instead (e.g. one pack per block size,
    the reference's packs needing blocks of one shape, hydro.cpp:537-549)."""

    def __init__(self, pkg, md, g=(0.0, 0.0, 0.0), pack_size=None, amr=None, partitions=None,
                 gravity=None):
        self.pkg = pkg
        self.md = md
        self.amr = amr  # amr.AmrPlan of a multilevel mesh, or None
        self.eos = pkg.param("eos")
        self.cfl = pkg.param("cfl")
        self.g = tuple(float(v) for v in g)
        self.u1 = torch.zeros_like(md.cons)
        # gravity: a package from initialize_gravity; "monopole" runs its solver each stage
        # and keeps the "gravity" and "gravity.phi" fields on md (md.grav, md.phi)
        self.mono = None
        self.poisson = None
        kind = gravity.param("gravity") if gravity is not None else Gravity.none
        if kind in (Gravity.monopole, Gravity.poisson):
            if kind == Gravity.monopole:
                self.mono = MonopoleGravitySolver.from_package(md.ctx, gravity, md.bs.ndim)
            else:
                self.poisson = PoissonGravitySolver.from_package(
                    md.ctx, gravity, md.dbs, mesh=amr.mesh if amr is not None else None)
            md.grav = torch.zeros(md.bs.field_size(3), dtype=torch.float64,
                                  device=md.cons.device)
            md.phi = torch.zeros(md.bs.field_size(1), dtype=torch.float64,
                                 device=md.cons.device)
        nb = md.bs.nblocks
        if partitions is not None:
            cover = sorted((int(f), int(c)) for f, c in partitions)
            at = 0
            for f, c in cover:
                if f != at or c < 1:
                    raise ValueError("partitions must cover the blocks once, in ranges")
                at = f + c
            if at != nb:
                raise ValueError("partitions must cover the blocks once, in ranges")
            self.parts = [md.pack(f, c) for f, c in partitions] or [md.pack(0, 0)]
        else:
            ps = pack_size or max(nb, 1)
            # an empty block set is one empty partition
            self.parts = ([md.pack(s, min(ps, nb - s)) for s in range(0, nb, ps)] or
                          [md.pack(0, 0)])
        self.dt_dev = torch.empty(1, dtype=torch.float64, device=md.cons.device)
        self.rk = INTEGRATORS[pkg.param("integrator").name]

    def initialize(self):
        """Ghost fill + FillDerived after the problem generator (PostInitialization)."""
        for p in self.parts:
            fill_ghosts(p)
        if self.amr is not None:
            self.amr.fill(self.md.cons)
        for p in self.parts:
            cons_to_prim(p, self.eos)
        return self.new_dt()

    def new_dt(self):
        est = []
        for p in self.parts:
            est.append(float(estimate_timestep(p, self.eos, self.cfl, self.dt_dev).item()))
        return min(est)

    def make_task_collection(self, stage, dt):
        """Run the tasks of one stage (apophis_driver.cpp:27-130) in dependency order."""
        s = stage - 1
        if stage == 1:  # :44-55
            self.u1.copy_(self.md.cons)
        flux_fun = self.pkg.param("flux_first_stage" if stage == 1 else "flux_other_stage")
        grav = any(v != 0.0 for v in self.g)
        rk = self.rk
        if self.mono is not None:
            self.mono.add_tasks(self.parts, self.md.grav, self.md.phi)      # :64-66
        if self.poisson is not None:  # one solve over the whole mesh       # :64-66
            self.poisson.solve(self.md, self.md.grav, self.md.phi)
        for p in self.parts:
            flux_fun(p)                                                      # :70-72
        if self.amr is not None:
            self.amr.flux_correct(self.md.flux)                              # :74-76
        for p in self.parts:
            update_with_flux_divergence(p, self.u1, rk["gam0"][s], rk["gam1"][s],
                                        rk["beta"][s] * dt)                  # :79-82
            if self.mono is not None or self.poisson is not None:
                apply_gravity_field(p, self.md.grav, rk["beta"][s] * dt)     # :84-85
            elif grav:
                apply_gravity(p, self.g, rk["beta"][s] * dt)                 # :84-85
        for p in self.parts:
            fill_ghosts(p)                                                   # :94, :103
        if self.amr is not None:
            self.amr.fill(self.md.cons)                                      # :94
        for p in self.parts:
            cons_to_prim(p, self.eos)                                        # :110-111
        if stage == rk["nstages"]:
            return self.new_dt()                                             # :113-116
        return None

    def step(self, dt):
        est = None
        for stage in range(1, self.rk["nstages"] + 1):
            est = self.make_task_collection(stage, dt)
        return est

    def history_sums(self):
        """The hydro package's history variables (hydro.cpp:336-346): total_mass and
        total_energy, GlobalQuantHst sums of rho V and E V over every block."""
        return [float(global_quant_hst(self.md, v).item()) for v in (0, 4)]

    def evolve(self, tlim=None, nlim=None, fixed_dt=None, hst=None):
        """EvolutionDriver loop with the upstream dt policy (<= 2x growth, clip to tlim).
        hst: an output.HstOutput, written at the initial time and then whenever the time
        reaches its next output time (its dt column: the step's dt, the initial estimate at
        t = 0)."""
        t, n = 0.0, 0
        est = self.initialize()
        dt = np.inf
        nb = self.md.bs.nblocks
        if hst is not None:
            hst.maybe_write(t, est, n, nb, self.history_sums)
        while (nlim is None or n < nlim) and (tlim is None or t < tlim * (1 - 1e-14)):
            if fixed_dt is not None:
                dt = fixed_dt
            else:
                dt = min(2.0 * dt, est)
                if tlim is not None and t < tlim and tlim - t < dt:
                    dt = tlim - t
            est = self.step(dt)
            t += dt
            n += 1
            if hst is not None:
                hst.maybe_write(t, dt, n, nb, self.history_sums)
        return t, n


class FusedDriver:
    """Batched form: one fused launch per RK stage, device-resident dt.

    u0 is the state (raw conserved variables; floors are applied on the fly, once per
    stage); u1 holds the stage-1 result.  Ghost cells are never read."""

    def __init__(self, ctx, bs: BlockSet, cons, recon, riemann, eos, cfl, g=(0.0, 0.0, 0.0),
                 tlim=np.inf, device="cuda", dbs=None, integrator="rk2", amr_mesh=None,
                 monopole=None, poisson=None):
        self.ctx = ctx
        self.bs = bs
        self.device_bs = dbs or DeviceBlockSet(ctx, bs)
        self.eos = eos
        self.cfl = float(cfl)
        self.tlim = float(tlim)
        self.u0 = torch.as_tensor(np.ascontiguousarray(cons), dtype=torch.float64).to(device)
        self.u1 = self.u0.clone()
        self.dt = torch.zeros(1, dtype=torch.float64, device=device)
        self.time = torch.zeros(1, dtype=torch.float64, device=device)
        self.plan = ctypes.c_void_p()
        garr = (ctypes.c_double * 3)(*[float(v) for v in g])
        r = recon if isinstance(recon, int) else int(Reconstruction[recon])
        s = riemann if isinstance(riemann, int) else int(RiemannSolver[riemann])
        ctx.check(_lib.lib().aph_plan_create(ctx.h, self.device_bs.h, r, s,
                                             eos.ctypes.data_as(ctypes.c_void_p), garr,
                                             ctypes.byref(self.plan)))
        self.graph_steps = 0
        if integrator != "rk2":
            self.set_integrator(integrator)
        self.mono = None
        if monopole is not None:  # (nrbin, rmax, gravity_constant): monopole gravity
            self.mono = MonopoleGravitySolver(ctx, bs.ndim, *monopole)
            ctx.check(_lib.lib().aph_plan_set_monopole(ctx.h, self.plan, self.mono.h))
        self.poisson = None
        if poisson is not None:  # (gravity_constant, diagonal_alpha, tolerance, max_it
            # [, flux_correct, prolongation]); on a multilevel mesh the composite operator
            self.poisson = PoissonGravitySolver(ctx, self.device_bs.h, *poisson,
                                                mesh=amr_mesh)
            ctx.check(_lib.lib().aph_plan_set_poisson(ctx.h, self.plan, self.poisson.h))
        self.amr = None
        if amr_mesh is not None:  # a multilevel mesh (amr.AmrMesh) of this block set
            from . import amr as _amr
            self.amr = _amr.AmrPlan(ctx, self.device_bs.h, amr_mesh, bs)
            faces, level, nbuf = _amr.fused_faces(amr_mesh, bs)
            self._amr_faces, self._amr_level = faces, level
            ctx.check(_lib.lib().aph_plan_set_amr(
                ctx.h, self.plan, self.amr.h, faces.ctypes.data_as(ctypes.c_void_p),
                level.ctypes.data_as(ctypes.c_void_p), nbuf))

    def set_integrator(self, name):
        """rk1 / rk2 / rk3 / vl2 (aph_plan_set_integrator)."""
        self.ctx.check(_lib.lib().aph_plan_set_integrator(self.ctx.h, self.plan,
                                                          int(Integrator[name])))
        self.integrator = name
        self.graph_steps = 0

    def set_tile(self, n, tx, ty, tz):
        t = _lib.Tile(n, tx, ty, tz)
        self.ctx.check(_lib.lib().aph_plan_set_tile(self.ctx.h, self.plan, ctypes.byref(t)))
        self.graph_steps = 0

    def autotune(self, reps=3):
        """Time the candidate tiles of every block shape on the current state and keep
        the fastest (aph_plan_autotune; u1 is used as scratch).  Returns the timing
        report."""
        buf = ctypes.create_string_buffer(1 << 18)
        self.ctx.check(_lib.lib().aph_plan_autotune(self.ctx.h, self.plan, _ptr(self.u0),
                                                    _ptr(self.u1), _ptr(self.dt), int(reps),
                                                    buf, len(buf)))
        self.graph_steps = 0
        return buf.value.decode()

    def tiles(self):
        """{block shape (nx, ny, nz): tile (tx, ty, tz)} in use."""
        out = {}
        for b in range(self.bs.nblocks):
            shape = tuple(int(v) for v in self.bs.blocks["nx"][b])
            if shape not in out:
                t = _lib.Tile()
                self.ctx.check(_lib.lib().aph_plan_get_tile(self.ctx.h, self.plan, b,
                                                            ctypes.byref(t)))
                out[shape] = (t.tx, t.ty, t.tz)
        return out

    def set_workgroup(self, lanes):
        """Lanes per workgroup of the stage launches (64, 128 or 256)."""
        self.ctx.check(_lib.lib().aph_plan_set_workgroup(self.ctx.h, self.plan, int(lanes)))
        self.graph_steps = 0

    def set_lds_reserve(self, nbytes):
        """At least nbytes of LDS per workgroup (occupancy studies; same results)."""
        self.ctx.check(_lib.lib().aph_plan_set_lds_reserve(self.ctx.h, self.plan, int(nbytes)))
        self.graph_steps = 0

    def launch(self):
        """{block shape (nx, ny, nz): (lanes per workgroup, segment x, y, z)} in use."""
        out = {}
        for b in range(self.bs.nblocks):
            shape = tuple(int(v) for v in self.bs.blocks["nx"][b])
            if shape not in out:
                v = (ctypes.c_int32 * 4)()
                self.ctx.check(_lib.lib().aph_plan_get_launch(self.ctx.h, self.plan, b, v))
                out[shape] = tuple(int(x) for x in v)
        return out

    def initial_dt(self):
        """dt for the first step: the estimate of the initial state, clipped to tlim."""
        self.ctx.check(_lib.lib().aph_plan_estimate_timestep(self.ctx.h, self.plan,
                                                             _ptr(self.u0), self.cfl,
                                                             _ptr(self.dt)))
        if np.isfinite(self.tlim):
            self.dt.clamp_(max=self.tlim)
        return self.dt

    def set_dt(self, dt):
        self.dt.fill_(float(dt))

    def stage(self, stage, u_in, u_n, u_out):
        self.ctx.check(_lib.lib().aph_fused_stage(self.ctx.h, self.plan, stage, _ptr(u_in),
                                                  _ptr(u_n), _ptr(u_out), _ptr(self.dt)))

    def step(self, fixed_dt=False):
        self.ctx.check(_lib.lib().aph_plan_step(self.ctx.h, self.plan, _ptr(self.u0),
                                                _ptr(self.u1), _ptr(self.dt), _ptr(self.time),
                                                self.cfl, self.tlim, int(bool(fixed_dt))))

    def capture(self, nsteps, fixed_dt=False):
        self.ctx.check(_lib.lib().aph_plan_graph_capture(
            self.ctx.h, self.plan, int(nsteps), _ptr(self.u0), _ptr(self.u1), _ptr(self.dt),
            _ptr(self.time), self.cfl, self.tlim, int(bool(fixed_dt))))
        self.graph_steps = int(nsteps)

    def replay(self):
        self.ctx.check(_lib.lib().aph_plan_graph_launch(self.ctx.h, self.plan))

    def status(self):
        """Synchronises and raises if a multilevel wait of any launch since the last call
        ran out (aph_plan_status; the flag is cleared once reported)."""
        self.ctx.check(_lib.lib().aph_plan_status(self.ctx.h, self.plan))

    def set_wait_limit(self, polls):
        """Polls of a multilevel wait before it gives up (0: default; negative: give up
        at once, which forces the error path)."""
        self.ctx.check(_lib.lib().aph_plan_set_wait_limit(self.ctx.h, self.plan, int(polls)))
        self.graph_steps = 0

    def close(self):
        if self.plan:
            _lib.lib().aph_plan_destroy(self.ctx.h, self.plan)
            self.plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
