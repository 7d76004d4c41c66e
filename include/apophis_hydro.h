/*
 * apophis_hydro.h -- C ABI of the MI355X (gfx950) hydro flux path.
 *
 * This library replaces the per-meshblock hydro flux path of AlexHls/apophis:
 *   reconstruction      src/recon/plm_simple.hpp, src/recon/ppm_simple.hpp, dc_simple.hpp
 *   Riemann solvers     src/hydro/rsolvers/hydro_hlle.hpp, hydro_hllc.hpp
 *   flux task           Apophis::CalculateFluxes<F,R,S>      src/hydro/hydro.cpp:535-744
 *   derived fields      Apophis::ConsToPrim<EOS_t>           src/hydro/hydro.cpp:422-533
 *   time step           Apophis::EstimateTimestep<euler>     src/hydro/hydro.cpp:351-420
 *   history sums        Apophis::GlobalQuantHst<T,idx>       src/hydro/hydro.cpp:746-771
 *   gravity source      Apophis::ApplyGravity                src/gravity/gravity.cpp:374-410
 *   monopole gravity    MonopoleGravitySolver     src/gravity/gsolvers/monopole_gravity.hpp
 *   stage sequence      ApophisDriver::MakeTaskCollection    src/apophis_driver.cpp:27-130
 * and the upstream Parthenon pieces that task graph calls (flux-divergence update,
 * RK2 register copy, ghost exchange, outflow/periodic boundaries).
 *
 * Conventions (all entry points):
 *  - every array argument is a DEVICE pointer borrowed from the caller;
 *  - every function returns 0 on success or a negative APH_E* code, and stores a
 *    message retrievable with aph_last_error(ctx); no C++ exception crosses the ABI;
 *  - work is enqueued on the context's stream; nothing synchronises implicitly,
 *    except the functions documented as "synchronous";
 *  - one context per stream; a context must not be used from two threads at once.
 *
 * Field storage ("pack" layout, same as the reference's VariablePack): a field with
 * NV components over a block set stores block b, component v, cell (k, j, i) at
 *     NV * blocks[b].cell_base + v * ncell(b) + (k * Ny(b) + j) * Nx(b) + i
 * where Nx = nx[0] + 2*ng on an active axis and nx[0] on an inactive one (same for
 * Ny, Nz), and ncell(b) = Nx * Ny * Nz ("entire" cells, ghosts included).
 * Face arrays (fluxes) use the same layout: face i of axis x1 lies between cells
 * i-1 and i.  Variable indices follow src/main.hpp:6-20 (IDN=0, IM1..IM3=1..3,
 * IEN=4; primitive IV1..IV3=1..3, IPR=4); passive scalars follow at 5.. .
 */
#ifndef APOPHIS_HYDRO_H_
#define APOPHIS_HYDRO_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APH_ABI_VERSION 1

/* error codes */
#define APH_OK 0
#define APH_E_INVALID (-1)   /* bad argument / unsupported combination */
#define APH_E_HIP (-2)       /* a HIP runtime call failed */
#define APH_E_NOMEM (-3)     /* device allocation failed */
#define APH_E_STATE (-4)     /* call out of order (e.g. graph launch before capture) */

/* Enumerations carry the values of the reference enums (src/main.hpp:32-53). */
enum aph_reconstruction { APH_RECON_DC = 1, APH_RECON_PLM = 2, APH_RECON_PPM = 3 };
enum aph_riemann { APH_RIEMANN_NONE = 1, APH_RIEMANN_HLLE = 2, APH_RIEMANN_HLLC = 3,
                   APH_RIEMANN_LHLLC = 4 };
enum aph_integrator { APH_INTEGRATOR_RK1 = 1, APH_INTEGRATOR_RK2 = 2, APH_INTEGRATOR_RK3 = 3,
                      APH_INTEGRATOR_VL2 = 4 };
/* Physical boundary applied on a face whose neighbour entry is -1.
   Periodicity is expressed through the neighbour table instead. */
/* APH_BC_AMR: the region beyond this face belongs to blocks of another refinement level;
   aph_fill_ghosts leaves its ghost cells to aph_amr_fill (multilevel meshes below).
   APH_BC_PERIODIC names a periodic domain face where a mesh's six face codes are given
   (the multilevel tree below); a block's own bc[] never holds it. */
enum aph_bc { APH_BC_PERIODIC = 0, APH_BC_OUTFLOW = 1, APH_BC_REFLECT = 2, APH_BC_AMR = 3 };

/* One meshblock.  240 bytes, no implicit padding (numpy mirrors it field by field). */
typedef struct aph_block {
  int32_t nx[3];     /* interior cells per axis, 1 on inactive axes                  */
  int32_t ng;        /* ghost layers on every active axis                           */
  int64_t cell_base; /* sum of ncell() of all preceding blocks                      */
  double dx[3];      /* cell widths (inactive axes: full extent of the domain)      */
  double area[3];    /* face areas: dx1*dx2 (x1 faces), dx0*dx2 (x2), dx0*dx1 (x3)  */
  double vol;        /* cell volume dx0*dx1*dx2                                     */
  double xmin[3];    /* coordinate of the lower interior corner                     */
  int32_t nbr[27];   /* neighbour block at offset (ox,oy,oz) in {-1,0,1}^3, stored at
                        (oz+1)*9 + (oy+1)*3 + (ox+1); -1 = physical boundary         */
  int32_t bc[6];     /* aph_bc for faces -x1,+x1,-x2,+x2,-x3,+x3 when nbr is -1      */
  int32_t mesh_id;   /* independent mesh (instance) this block belongs to           */
} aph_block;

/* Ideal-gas EOS and floors (hydro.cpp:260-299). */
typedef struct aph_eos {
  double gm1;    /* gamma - 1 (eos/gm1)                     */
  double cv;     /* specific heat (eos/cv), carried for API parity, unused by ideal gas */
  double dfloor; /* hydro/dfloor                            */
  double pfloor; /* hydro/pfloor                            */
} aph_eos;

typedef struct aph_context aph_context;
typedef struct aph_blockset aph_blockset;
typedef struct aph_plan aph_plan;
typedef struct aph_monopole aph_monopole;
typedef struct aph_poisson aph_poisson;

/* ---------------------------------------------------------------- context */
int aph_abi_version(void);
/* stream: a hipStream_t (NULL = the default stream). */
int aph_context_create(void *stream, aph_context **out);
int aph_context_destroy(aph_context *ctx);
const char *aph_last_error(const aph_context *ctx);
/* Synchronous: waits for all work enqueued on the context's stream. */
int aph_context_synchronize(aph_context *ctx);

/* ------------------------------------------------------------- block set
 * Copies the block table to the device.  ndim in {1,2,3}; nscalars passive scalars
 * follow the 5 hydro variables (hydro.cpp:141-202): 0, 1 (a generic scalar), 7 (the
 * ONe composition: 6 species densities and Ye, hydro.cpp:154-162) or 9 (ONe plus one
 * level set's two fields, hydro.cpp:170-174); other counts fail with APH_E_INVALID.
 * Synchronous. */
int aph_blockset_create(aph_context *ctx, const aph_block *blocks, int32_t nblocks,
                        int32_t ndim, int32_t nscalars, aph_blockset **out);
int aph_blockset_destroy(aph_context *ctx, aph_blockset *bs);
/* Total entire cells (sum of ncell) -- one component of a field has this many doubles. */
int64_t aph_blockset_ncells(const aph_blockset *bs);

/* ----------------------------------------------------- plain form (per pack)
 * These mirror the reference task functions one to one and run on every block of
 * the block set (the reference's MeshData partition).  `first_block`/`nblocks`
 * select a contiguous sub-range of blocks (the reference's pack_size); pass 0 and
 * -1 for all blocks. */

/* ConsToPrim (hydro.cpp:422-533): over all cells incl. ghosts, applies the density
   and pressure floors to `cons` in place, writes prim (5+nscalars) and gamma (2). */
int aph_cons_to_prim(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                     int32_t nblocks, const aph_eos *eos, double *cons, double *prim,
                     double *gamma);

/* CalculateFluxes<euler,recon,rsolver> (hydro.cpp:535-744): writes cons.flux for
   every face the reference loops visit, including the one transverse ghost row.
   flux_x2 / flux_x3 may be NULL when ndim < 2 / < 3. */
int aph_calculate_fluxes(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                         int32_t nblocks, int32_t recon, int32_t rsolver,
                         const double *prim, const double *gamma, double *flux_x1,
                         double *flux_x2, double *flux_x3);

/* Parthenon Update::UpdateWithFluxDivergence (call site apophis_driver.cpp:79-82):
   interior u0 <- gam0*u0 + gam1*u1 + beta_dt * (-div F). */
int aph_update_with_flux_divergence(aph_context *ctx, const aph_blockset *bs,
                                    int32_t first_block, int32_t nblocks, double *u0,
                                    const double *u1, const double *flux_x1,
                                    const double *flux_x2, const double *flux_x3,
                                    double gam0, double gam1, double beta_dt);

/* ApplyGravity (gravity.cpp:374-410) for a constant acceleration g[3]. */
int aph_apply_gravity(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                      int32_t nblocks, double *cons, const double *prim,
                      const double g[3], double dt);

/* ApplyGravity (gravity.cpp:374-410) with a per-cell acceleration field: grav holds 3
   components in the pack layout (3 * cell_base + v * ncell(b) + cell), e.g. from
   aph_monopole_gravity_field; prim is the stage-start primitive state. */
int aph_apply_gravity_field(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                            int32_t nblocks, double *cons, const double *prim,
                            const double *grav, double dt);

/* Monopole self-gravity: MonopoleGravitySolver (src/gravity/gsolvers/monopole_gravity.hpp,
   gravity type "monopole", gravity.cpp:236-238,262-272).  nrbin radial bins of width
   rmax / nrbin (the reference's defaults: nrbin = mesh nx1, rmax = distance to the
   farthest mesh corner); gravity_constant is G.  The per-stage tasks, in order:
     aph_monopole_zero_radial_bins    ZeroRadialBins (:107-112) over the whole block set
                                      (it also fixes the scale of the sums from prim)
     aph_monopole_accumulate_mass     AccumulateMass (:114-153), per partition
     aph_monopole_build_radial_tables BuildRadialTables (:155-213), once
     aph_monopole_gravity_field       ComputeGravityVectorDirect (:215-302), per partition
   Bin masses are summed exactly (each cell mass rho*V rounded to a multiple of 2^-s, s
   about 100 bits below the largest, summed as 128-bit integers), so they do not depend on
   partitions or order; the reference's floating-point atomics do, in the last bits. */
int aph_monopole_create(aph_context *ctx, int32_t ndim, int32_t nrbin, double rmax,
                        double gravity_constant, aph_monopole **out);
int aph_monopole_destroy(aph_context *ctx, aph_monopole *m);
int aph_monopole_zero_radial_bins(aph_context *ctx, aph_monopole *m, const aph_blockset *bs,
                                  const double *prim);
int aph_monopole_accumulate_mass(aph_context *ctx, aph_monopole *m, const aph_blockset *bs,
                                 int32_t first_block, int32_t nblocks, const double *prim);
int aph_monopole_build_radial_tables(aph_context *ctx, aph_monopole *m);
/* grav: 3 components (pack layout); phi ("gravity.phi", 1 component) may be NULL. */
int aph_monopole_gravity_field(aph_context *ctx, aph_monopole *m, const aph_blockset *bs,
                               int32_t first_block, int32_t nblocks, double *grav,
                               double *phi);
/* Copy m_enc[0..nrbin] and phi_rad[0..nrbin] to host arrays (synchronous). */
int aph_monopole_get_tables(aph_context *ctx, aph_monopole *m, double *m_enc,
                            double *phi_rad);

/* Poisson self-gravity: PoissonGravitySolver (src/gravity/gsolvers/poisson_gravity.hpp,
   gravity type "poisson", gravity.cpp:239-241,279-366) on a 3D block set whose blocks tile
   a box in one lattice, every face of the box a physical boundary (the reference's
   gravity/solver_params apply: tolerance = residual_tolerance, relative to the right-hand
   side's norm; diagonal_alpha).  The discrete problem is PoissonEquation's with D = 1 (the
   7-point Laplacian over cell centres minus alpha, phi = 0 on the domain faces, the face
   values of the multipole expansion moved to the right-hand side, ComputeRhs :502-589);
   it is solved by conjugate gradients preconditioned with a geometric multigrid V-cycle.
   The reference solves the same discrete problem with Parthenon's MG or BiCGSTAB, so the
   two agree to the tolerance.  Multilevel meshes: aph_poisson_create_tree; periodic faces,
   1D and 2D: refused. */
int aph_poisson_create(aph_context *ctx, const aph_blockset *bs, double gravity_constant,
                       double diagonal_alpha, double tolerance, int32_t max_iterations,
                       aph_poisson **out);
int aph_poisson_destroy(aph_context *ctx, aph_poisson *ps);
/* One stage's solver tasks on the primitive state: PreComputeTasks (ComputeMPCoeff,
   ScaleMPCoeff, :394-500) and AddTasks (ComputeRhs, the solve, the boundary values of phi,
   ComputeGravityVector, :415-625).  grav: 3 components (pack layout); phi ("gravity.phi",
   1 component) may be NULL.  Synchronous (the iteration's convergence test reads the
   device every few iterations); fails with APH_E_STATE when max_iterations pass without
   convergence. */
int aph_poisson_solve(aph_context *ctx, aph_poisson *ps, const double *prim, double *grav,
                      double *phi);
/* The last solve's nine multipole coefficients (scaled), iterations and relative residual.
   Synchronous; any output may be NULL. */
int aph_poisson_get_info(aph_context *ctx, aph_poisson *ps, double coeffs[9],
                         int32_t *iterations, double *rel_residual);

/* The phi field's prolongation at fine/coarse block faces (gravity/boundary_prolongation,
   gravity.cpp:312-363): Parthenon's ProlongateSharedLinear (c + sum_d (+-1/4) (c_+d -
   c_-d) / 2, unlimited) or ProlongatePiecewiseConstant.  "Quadratic" is not built. */
enum aph_poisson_prolongation { APH_POISSON_PROLONG_LINEAR = 0, APH_POISSON_PROLONG_CONSTANT = 1 };

/* Poisson self-gravity on a multilevel 3D mesh (the reference's static-refinement decks,
   inputs/grav_test.in and inputs/evrards_collapse.in: PoissonEquation on the mesh's leaf
   cells, poisson_gravity.hpp:31-388).  bs: the leaves' block set; level / loc (3 per
   block) / nbase: the tree as for aph_amr_build, every domain face physical.  The
   operator: per block face the flux (phi_low - phi_high) / dx, the value beyond a block
   face from the same level (a leaf cell or the mean of a refined cell's children) or
   prolongated from the coarser leaf beyond, minus the cell inside beyond a domain face;
   flux_correct (gravity/flux_correct, :89-99) replaces a coarse face's flux over a finer
   region by the mean of its fine faces.  Assembled once on the host from the tree, with
   the same problem on coarser meshes (the tree cut one level at a time, then the base
   blocks with half the cells per axis); each solve runs BiCGSTAB from phi = 0 to the
   relative residual tolerance, preconditioned by one multigrid V-cycle over those
   meshes (the reference's MGBiCGSTAB: the same discrete solution, a like number of
   iterations; the decks' max_iterations = 15 suffices).
   aph_poisson_solve, aph_poisson_get_info, aph_poisson_destroy and aph_plan_set_poisson
   take the solver as they take the lattice one; the right-hand side, the multipole
   coefficients and the field are the lattice solver's, over the leaves.  Refused: 1D,
   2D, periodic faces, blocks whose spacing or corner differs from their level and
   position, prolongation other than the two above. */
int aph_poisson_create_tree(aph_context *ctx, const aph_blockset *bs, const int32_t nbase[3],
                            const int32_t *level, const int64_t *loc, int32_t flux_correct,
                            int32_t prolongation, double gravity_constant,
                            double diagonal_alpha, double tolerance, int32_t max_iterations,
                            aph_poisson **out);
/* Host only (no device work): the operator aph_poisson_create_tree assembles, as CSR over
   the leaf cells (cell (i, j, k) of block b at b nx0 nx1 nx2 + (k nx1 + j) nx0 + i; rowptr
   n + 1 entries), and the neighbour rows of the field (row (2 d + side) n + c: the value
   beyond cell c's face on that side of axis d as a map of the cell values, empty on a
   domain face; nb_rowptr 6 n + 1 entries).  Call with NULL arrays for nnz and nb_nnz. */
int aph_poisson_tree_matrix(const aph_block *blocks, int32_t nblocks, const int32_t nbase[3],
                            const int32_t *level, const int64_t *loc, int32_t flux_correct,
                            int32_t prolongation, double diagonal_alpha, int64_t *rowptr,
                            int32_t *col, double *val, int64_t *nnz, int64_t *nb_rowptr,
                            int32_t *nb_col, double *nb_val, int64_t *nb_nnz, char *err,
                            size_t err_len);

/* Ghost exchange between neighbouring blocks followed by the physical boundary
   conditions (outflow / reflect, apophis_boundaries.cpp:17-92), axis by axis. */
int aph_fill_ghosts(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                    int32_t nblocks, double *cons);

/* EstimateTimestep<euler> (hydro.cpp:351-420): *dt_out (device) = cfl * min over
   interior cells and active axes of dx/(|v|+lambda). */
int aph_estimate_timestep(aph_context *ctx, const aph_blockset *bs, int32_t first_block,
                          int32_t nblocks, const aph_eos *eos, const double *prim,
                          const double *cons, double cfl, double *dt_out);

/* GlobalQuantHst<Sum, var> (hydro.cpp:746-771): *out (device) = sum(cons[var]*vol). */
int aph_global_sum(aph_context *ctx, const aph_blockset *bs, const double *cons,
                   int32_t var, double *out);
/* The same sum without any allocation.  No counterpart in the reference: it is for
   callers that cannot allocate device memory (code under HIP graph capture, tight
   serving loops).  aph_global_sum allocates its per-block partial sums on first use;
   this variant keeps them in a fixed device array inside the library (one per GPU,
   65536 blocks = 512 KiB, held until the process ends), so it fails with APH_E_INVALID
   for block sets of more than 65536 blocks.  Every call on one GPU shares that array:
   calls that overlap on one GPU (e.g. on two streams) race and must be ordered by the
   caller.  Same kernels and summation order as aph_global_sum, so the same bits. */
int aph_global_sum_noalloc(aph_context *ctx, const aph_blockset *bs, const double *cons,
                           int32_t var, double *out);

/* Refinement criteria of the hydro package (hydro.cpp:303-333, src/refinement/), one
 * batched launch over the blocks: per block, the reference's reduction over its own index
 * ranges of prim (which must hold valid ghost cells, as after aph_cons_to_prim):
 *   APH_REFINE_PRESSURE_GRADIENT (gradient.cpp:11-53): eps = max sqrt(sum_d (dp_d/2)^2)/p
 *     over the interior grown by one cell on every active axis (1D: tag 0);
 *     refine if eps > params[0], derefine if eps < params[0]/4;
 *   APH_REFINE_VELOCITY_GRADIENT (gradient.cpp:56-84, ndim >= 2): vg = max sqrt(vgx^2 +
 *     vgy^2), vgy = |v2(i+1) - v2(i-1)|/2, vgx = |v1(j+1) - v1(j-1)|/2, over i and j grown
 *     by one, k interior; refine if vg > params[0], derefine if vg < params[0]/2;
 *   APH_REFINE_MAX_DENSITY (other.cpp:11-34): max rho over k, j interior and i from the
 *     first interior cell to one past the last; refine if > params[1], derefine if <
 *     params[0].
 * tags (device, nblocks): -1 derefine, 0 same, 1 refine (Parthenon's AmrTag); values
 * (device, nblocks, may be NULL): the reduced quantity. */
enum aph_refine_criterion {
  APH_REFINE_PRESSURE_GRADIENT = 1,
  APH_REFINE_VELOCITY_GRADIENT = 2,
  APH_REFINE_MAX_DENSITY = 3
};
int aph_refinement_tag(aph_context *ctx, const aph_blockset *bs, int32_t criterion,
                       const double params[2], const double *prim, int32_t *tags,
                       double *values);

/* ------------------------------------------- multilevel meshes (plain form)
 * Blocks of several refinement levels in one mesh, 2:1 balanced (leaves that touch differ
 * by at most one level).  nbr[] links blocks of the same level only; a face whose
 * neighbour region belongs to another level has nbr = -1 and bc = APH_BC_AMR.  The ghost
 * cells there, and the coarse fluxes on faces shared with finer blocks, come from a plan
 * the host builds from the mesh tree.  This replaces
 * Parthenon's multilevel boundary exchange (RestrictAverage / ProlongateSharedMinMod
 * for the cons field) and its flux correction (LoadAndSendFluxCorrections /
 * SetFluxCorrections), the tasks at apophis_driver.cpp:68,74-76 and :94.
 *
 * A cell is an aph_amr_ref: its component v is at element offset + v * stride (stride =
 * its block's ncell).  A value is an aph_amr_val: the mean of refs[first .. first+count)
 * (count 1: a copy; 2^ndim: the restriction of a coarse cell's children), summed in
 * order and scaled by 1/count. */
typedef struct aph_amr_ref {
  int64_t offset;
  int64_t stride;
} aph_amr_ref;
typedef struct aph_amr_val {
  int32_t first;
  int32_t count;
} aph_amr_val;
enum aph_amr_kind { APH_AMR_MEAN = 0, APH_AMR_PROLONG = 1 };
/* kind's low byte is an aph_amr_kind; bit APH_AMR_MIRROR_BIT + k - 1 (k = 1..6) marks value
   slot k of a PROLONG op as the mirror image of a coarse cell across a reflecting domain
   face normal to its axis d = (k - 1) / 2: its momentum component 1 + d is negated. */
#define APH_AMR_MIRROR_BIT 8
/* One ghost cell of cons (all components):
 *   APH_AMR_MEAN:    value(val[0]);
 *   APH_AMR_PROLONG: c = value(val[0]); for each active axis d in order,
 *                    g_d = minmod(c - value(val[1+2d]), value(val[2+2d]) - c),
 *                    minmod(a, b) = (SIGN(a) + SIGN(b))/2 * min(|a|, |b|), SIGN(0) = 1;
 *                    result = (((c + sign[0]/4 g_0) + sign[1]/4 g_1) + sign[2]/4 g_2),
 *                    sign[d] = -1 / +1 for the lower / upper child of the coarse cell.
 *                    A slope neighbour beyond a physical domain face is the coarse level's
 *                    boundary image (Parthenon applies the boundary conditions to the
 *                    coarse buffer before it prolongates): outflow repeats the coarse cell,
 *                    reflect mirrors it (the mirror bit above). */
typedef struct aph_amr_op {
  int64_t dst;        /* component-0 element offset of the ghost cell   */
  int64_t dst_stride; /* its block's ncell                               */
  int32_t kind;
  int32_t sign[3];
  aph_amr_val val[7];
} aph_amr_op;
/* One coarse face of flux array `dir` (0: flux_x1, ...): all components set to the mean
   of the fine faces val refers to (offsets into the same flux array). */
typedef struct aph_flux_fix {
  int64_t dst;
  int64_t dst_stride;
  int32_t dir;
  int32_t reserved;
  aph_amr_val val;
} aph_flux_fix;
typedef struct aph_amr_plan aph_amr_plan;

/* Validates every op, fix and ref against the block set (offsets inside the fields,
   value ranges inside refs, at most 8 cells per value) and copies them to the device.
   Synchronous. */
int aph_amr_plan_create(aph_context *ctx, const aph_blockset *bs, const aph_amr_op *ops,
                        int64_t nops, const aph_flux_fix *fixes, int64_t nfixes,
                        const aph_amr_ref *refs, int64_t nrefs, aph_amr_plan **out);
int aph_amr_plan_destroy(aph_context *ctx, aph_amr_plan *plan);
/* The op lists of a multilevel mesh from its leaves (the Parthenon mesh tree's neighbour
   finding across levels; host only, no context or device needed).  blocks: the block
   set's table (every block the same nx, even, and ng; nbr / bc as described above);
   level[b] and loc[3*b .. 3*b+2]: each block's level and logical position at its level in
   a base lattice of nbase[0..ndim-1] blocks; the tree must be 2:1 balanced.  bc: the six
   domain face codes (-x1, +x1, -x2, ...; APH_BC_PERIODIC on both faces of a periodic
   axis, else APH_BC_OUTFLOW or APH_BC_REFLECT), NULL: every axis periodic.  Ops: one per
   ghost cell inside the domain with no same-level neighbour in its direction (MEAN over a
   refined node, PROLONG from a coarser leaf); ghost cells beyond a physical domain face
   are left to the boundary conditions.  Fixes: one per coarse face cell on a face to a
   refined node; in block order, ghost cells k, j, i ascending, children (cz, cy, cx) x
   fastest.  On input *nops, *nfixes, *nrefs are the capacities of the arrays (an array may
   be NULL: then only its length is returned); on output the lengths.  On failure the
   message goes to err (err_len bytes, may be NULL). */
/* The mesh tree (host only, like aph_amr_build).  A tree is given by its leaves:
   level[b] and loc[3*b .. 3*b+2] in a base lattice of nbase[0..ndim-1] nodes; the leaves
   must cover the domain exactly once.  aph_amr_regrid_tree and aph_amr_remap take every
   axis periodic (the adaptive path); aph_amr_block_table takes the domain face codes.
   aph_amr_regrid_tree: one regrid from per-leaf tags (-1 / 0 / 1), Parthenon's rules --
   leaves tagged 1 below max_level split; every leaf touching a leaf two levels finer splits
   (2:1 balance); sibling groups all tagged -1 (none split) merge, finest level first, then
   by position (x, then y, then z), unless the parent would touch a leaf two levels finer.
   Writes the new leaves in depth-first order (base lattice z, y, x; children z, y, x) to
   out_level / out_loc (cap entries; NULL: only *out_n is set).
   aph_amr_block_table: the aph_block table of the leaves (in their order) after checking
   2:1 balance: extents, ghosts, dx = (xmax - xmin) / (nbase nx 2^level) per active axis,
   areas, volume, xmin, nbr to same-level leaves (wrapping on periodic axes only), bc =
   the domain code on faces on a physical domain face, APH_BC_AMR on other faces without a
   same-level neighbour (APH_BC_OUTFLOW elsewhere), cell_base 0 (the block set fills it).
   bc: the six domain face codes as for aph_amr_build (NULL: periodic). */
int aph_amr_regrid_tree(int32_t ndim, const int32_t nbase[3], int32_t max_level, int32_t n,
                        const int32_t *level, const int64_t *loc, const int32_t *tags,
                        int32_t *out_level, int64_t *out_loc, int32_t *out_n, int32_t cap,
                        char *err, size_t err_len);
int aph_amr_block_table(int32_t ndim, const int32_t nbase[3], const int32_t nx[3], int32_t ng,
                        const double xmin[3], const double xmax[3], const int32_t *bc,
                        int32_t mesh_id, int32_t n, const int32_t *level, const int64_t *loc,
                        aph_block *out, char *err, size_t err_len);
/* The interiors of a regridded mesh (host arrays in the pack layout, every block nx^ndim
   cells with ng ghosts, nvar components, blocks in leaf order): a kept leaf copies its
   cells; any other new leaf takes its box of the old mesh's level values -- a leaf's cells,
   the mean of a refined node's children ((dz, dy, dx) order, scaled by 2^-ndim), or under
   a coarser leaf c + sum_d (sign_d / 4) minmod(c - c_-d, c_+d - c) of the level above, with
   the stencil's values taken the same way (Parthenon's refinement / derefinement data
   transfer).  Ghost cells of new_u are left untouched. */
int aph_amr_remap(int32_t ndim, const int32_t nbase[3], const int32_t nx[3], int32_t ng,
                  int32_t nvar, int32_t n_old, const int32_t *old_level, const int64_t *old_loc,
                  const double *old_u, int32_t n_new, const int32_t *new_level,
                  const int64_t *new_loc, double *new_u, char *err, size_t err_len);
int aph_amr_build(const aph_block *blocks, int32_t nblocks, int32_t ndim, int32_t nvar,
                  const int32_t nbase[3], const int32_t *level, const int64_t *loc,
                  const int32_t *bc, aph_amr_op *ops, int64_t *nops, aph_flux_fix *fixes,
                  int64_t *nfixes, aph_amr_ref *refs, int64_t *nrefs, char *err,
                  size_t err_len);
/* Cross-level ghost values of cons; run after aph_fill_ghosts.  The ops read interior
   cells only and write ghost cells only.  On a mesh with physical domain faces the
   boundary conditions of aph_fill_ghosts then run again over every block on one, so that
   the edge and corner ghost cells beyond the domain take their values from the cross-level
   ghost cells beside them (Parthenon fills the neighbour regions first and the physical
   boundaries last).  The block set given to aph_amr_plan_create must outlive the plan. */
int aph_amr_fill(aph_context *ctx, const aph_amr_plan *plan, double *cons);
/* Flux correction before the update: every coarse face shared with finer blocks takes
   the mean of the fine fluxes across it.  flux_x2 / flux_x3 may be NULL when unused. */
int aph_amr_flux_correct(aph_context *ctx, const aph_amr_plan *plan, double *flux_x1,
                         double *flux_x2, double *flux_x3);

/* ----------------------------------------------------- batched fused form
 * A plan binds a block set to one (recon, rsolver) pair and a tile shape per block
 * size.  aph_fused_stage runs ONE persistent launch that, for every block, does
 * cons->prim (with floors), reconstruction, the Riemann solve and the flux-
 * divergence update of one RK2 stage, reading ghost values straight from the
 * neighbouring blocks (or through the physical boundary rule) instead of from
 * ghost storage:
 *   stage 1:  u_out = 0*u_in + 1*u_in + dt   * L(u_in)    (u_n unused)
 *   stage 2:  u_out = .5*u_in + .5*u_n + dt/2 * L(u_in)   (u_out may alias u_n)
 * Stage 1 also writes the floored interior state back to u_in.  dt is read from
 * device memory (*dt_dev).  After stage 2 the plan's time-step estimate for the
 * new state is min-reduced on the device (see aph_plan_step). */
typedef struct aph_tile {
  int32_t n;      /* block edge this entry applies to (8, 16, 32, ...)      */
  int32_t tx, ty, tz; /* tile shape in cells; must divide the block edge    */
} aph_tile;

int aph_plan_create(aph_context *ctx, const aph_blockset *bs, int32_t recon,
                    int32_t rsolver, const aph_eos *eos, const double g[3],
                    aph_plan **out);
int aph_plan_destroy(aph_context *ctx, aph_plan *plan);
/* Override the tile shape for blocks of edge t->n (autotuner hook).  Synchronous. */
int aph_plan_set_tile(aph_context *ctx, aph_plan *plan, const aph_tile *t);
/* Autotuner (no counterpart in the reference, whose flux kernels take one fixed team
 * shape, hydro.cpp:580-583).  For each workgroup size (64, 128, 256 lanes) and every block
 * shape of the plan it times `reps` stage-1 launches (u_in -> u_scratch, *dt_dev on the
 * device) over the blocks of that shape for each candidate tile (divisors of the block
 * edges, one to sixteen cells per lane, within LDS) and segment length (the shortest that
 * keeps every lane busy, or twice that), keeps each shape's fastest, and takes the
 * workgroup size whose shapes sum to the least time (a larger size only when it is more
 * than 2% faster than the best smaller one); tiles set with aph_plan_set_tile are
 * kept as they are (their segments are not tuned).  Synchronous; u_in
 * is only read, u_scratch is overwritten.  If `report` is non-null, a text table of the
 * candidates and their average launch times (ms) is written to it (at most report_len
 * bytes, NUL-terminated).  Fails with APH_E_STATE while the context's stream is
 * capturing a graph. */
int aph_plan_autotune(aph_context *ctx, aph_plan *plan, const double *u_in, double *u_scratch,
                      const double *dt_dev, int32_t reps, char *report, size_t report_len);
/* Tile currently used for block `block` of the plan's block set. */
int aph_plan_get_tile(aph_context *ctx, aph_plan *plan, int32_t block, aph_tile *out);
/* Lanes per workgroup of the plan's stage launches: 64, 128 or 256 (default 64; the
 * autotuner sets it).  Results do not depend on it bit for bit.  Synchronous.  No
 * reference counterpart. */
int aph_plan_set_workgroup(aph_context *ctx, aph_plan *plan, int32_t lanes);
/* Reserve at least `bytes` (0..160 KiB) of LDS per workgroup of the plan's stage launches.
 * Results are unchanged; fewer workgroups fit on a CU.  For occupancy studies (what a
 * variant needing that much LDS would pay).  Synchronous.  No reference counterpart. */
int aph_plan_set_lds_reserve(aph_context *ctx, aph_plan *plan, int64_t bytes);
/* Launch shape for block `block`: out[0] = lanes per workgroup (one per plan: 64, 128 or
 * 256), out[1..3] = the segment length one lane marches along x, y, z.  No reference
 * counterpart (the reference's flux kernels take one fixed team shape). */
int aph_plan_get_launch(aph_context *ctx, aph_plan *plan, int32_t block, int32_t out[4]);
/* Time integrator of the plan (default rk2): the upstream Parthenon low-storage
 * schemes rk1 / rk2 / rk3 / vl2 that ApophisDriver uses (apophis_driver.cpp:78-85,
 * hydro.cpp:114-139; vl2 predicts with DC reconstruction).  rk3 allocates one more
 * state register inside the plan.  Synchronous.  The adjoint entry points follow the
 * plan's integrator. */
int aph_plan_set_integrator(aph_context *ctx, aph_plan *plan, int32_t integrator);
/* Monopole gravity for every stage of the plan (replaces the constant g): before each
   stage the solver bins the stage input's density (ConsToPrim's floor), builds its tables
   and the field, and the stage adds the source with that field.  m must outlive the plan
   (or a later call with m = NULL, which returns to the plan's constant g).  The adjoint
   refuses such plans.  Synchronous (allocates the field). */
int aph_plan_set_monopole(aph_context *ctx, aph_plan *plan, aph_monopole *m);
/* Poisson gravity for every stage of the plan (as aph_plan_set_monopole; NULL returns to
   the constant g).  A plan with it cannot be captured in a graph (the solve reads its
   convergence on the host). */
int aph_plan_set_poisson(aph_context *ctx, aph_plan *plan, aph_poisson *ps);
/* Stage `stage` (1..number of stages of the plan's integrator):
 * u_out = gam0*u_in + gam1*u_n + beta*dt*(-div F(u_in)) [+ gravity]; u_n (the step's
 * start state) is unused in stage 1. */
int aph_fused_stage(aph_context *ctx, aph_plan *plan, int32_t stage, double *u_in,
                    const double *u_n, double *u_out, const double *dt_dev);

/* Device-resident time state used by aph_plan_step / the captured graph:
   dt_dev[0] = current dt, time_dev[0] = simulation time.  One step of the plan's
   integrator -- rk2 / vl2: stage 1 (U0 -> U1), stage 2 (U1, U0 -> U0); rk3: U0 -> U1,
   (U1, U0) -> U2 (inside the plan), (U2, U0) -> U0; rk1: U0 -> U1 copied to U0 -- then
     dt <- min(2*dt, cfl * min_est), clipped so that time + dt <= tlim, time += dt_old.
   When fixed_dt != 0 the dt is left unchanged (frozen-dt rollouts). */
int aph_plan_step(aph_context *ctx, aph_plan *plan, double *u0, double *u1,
                  double *dt_dev, double *time_dev, double cfl, double tlim,
                  int32_t fixed_dt);
/* Capture `nsteps` calls of aph_plan_step into a hipGraph (stream capture) and keep
   it in the plan; aph_plan_graph_launch replays it on the context stream. */
int aph_plan_graph_capture(aph_context *ctx, aph_plan *plan, int32_t nsteps, double *u0,
                           double *u1, double *dt_dev, double *time_dev, double cfl,
                           double tlim, int32_t fixed_dt);
int aph_plan_graph_launch(aph_context *ctx, aph_plan *plan);
/* Initial dt estimate for the plan's block set from cons (device), written to *dt_out:
   cfl * min(dx/(|v|+lambda)) with the ConsToPrim floors applied on the fly. */
int aph_plan_estimate_timestep(aph_context *ctx, aph_plan *plan, const double *cons,
                               double cfl, double *dt_out);

/* Multilevel meshes in the fused form.  Per block face, for faces with bc = APH_BC_AMR
   (zeros elsewhere): the fine side of every level jump writes the flux of each of its
   boundary faces to an interface buffer, ordered per coarse face cell (transverse axes in
   ascending order, the first slowest) and child (the same order); the coarse side
   replaces each boundary face flux by the mean of its children there, in that order --
   the flux correction of aph_amr_flux_correct, bitwise.  A block between two levels can
   hold faces of both roles.  A stage runs aph_amr_fill on its input, then one persistent
   launch over every level whose items run the finest level first: a tile with a face to
   a finer region waits (bounded) until the fine tiles that write it have published.  If
   that wait runs out (the device could not keep the whole grid resident), the stage
   writes NaN to the first element of u_out, aph_plan_step also sets dt and time to NaN,
   and the plan's error flag stays set until aph_plan_status reports it. */
typedef struct aph_amr_face {
  int64_t base;  /* element offset of the coarse face region in the interface buffer */
  int32_t role;  /* 0 none, 1 fine side (writes), 2 coarse side (reads)               */
  int32_t qa, qb;/* fine side: offset of this block's face inside the coarse face
                    region, in fine cells along the two transverse axes              */
  int32_t na;    /* coarse face cells along the first transverse axis               */
} aph_amr_face;
/* faces: nblocks*6 entries (face order -x1,+x1,-x2,+x2,-x3,+x3); level: nblocks entries
   (0 the base level, up to 30); buffer_doubles: interface buffer length.  Checked: each
   role-2 face owns [base, base + its face cells x 2^(ndim-1) x nvar) inside the buffer,
   regions do not overlap, and each role-1 face lies inside the region of the opposite
   face of a block one level coarser.  The amr plan must outlive the plan.  The adjoint
   entry points run on such a plan as well (they fill the cross-level ghost cells of their
   state inputs first, as a forward stage does).  Synchronous. */
int aph_plan_set_amr(aph_context *ctx, aph_plan *plan, const aph_amr_plan *amr,
                     const aph_amr_face *faces, const int32_t *level,
                     int64_t buffer_doubles);
/* Synchronous: waits for the context's stream, then returns APH_E_STATE (with a message)
   if a multilevel wait of any launch since the last call ran out, and clears the flag;
   APH_OK otherwise (always, for a single-level plan).  No reference counterpart: the
   reference's multilevel exchange is host MPI, which fails by aborting. */
int aph_plan_status(aph_context *ctx, aph_plan *plan);
/* Polls a multilevel wait makes before it gives up (0: the default, 2^20).  A negative
   value makes every wait give up at once, which forces the error path above from the
   host (used by its tests).  Drops a captured graph. */
int aph_plan_set_wait_limit(aph_context *ctx, aph_plan *plan, int32_t polls);

/* ----------------------------------------------------- adjoint (frozen dt)
 * Reverse-mode counterparts of the fused stage (no reference counterpart: the
 * reference has no adjoint).  Stage s of the plan's integrator computes
 * u_out = gam0 u_in + gam1 u_n + beta dt L(u_in) (u_n = u_in in stage 1; vl2's stage 1
 * reconstructs with DC).  Given the cotangent lam_out of a stage's output:
 *   stage 1:  lam_in = (gam0 + gam1) * lam_out + beta dt J^T lam_out [+ lam_acc]
 *   stage s > 1:  lam_in = gam0 * lam_out + beta dt J^T lam_out [+ lam_acc],
 *                 lam_n = gam1 * lam_out
 * where J is the Jacobian of the flux-divergence (+ gravity) term with respect to the
 * stage input u_in (ConsToPrim, reconstruction through the limiter's active branch, the
 * Riemann solver's active wave branches; HLLE, HLLC and LHLLC; nvar 5, 6, 12 or 14; on a
 * multilevel plan also the cross-level ghost values and the flux correction).  Only
 * interior cells of the cotangent arrays are read or written; on a multilevel plan the
 * cross-level ghost cells of the state inputs are filled.  lam_acc and lam_n may be
 * NULL. */
int aph_adjoint_stage(aph_context *ctx, aph_plan *plan, int32_t stage, const double *u_in,
                      const double *dt_dev, const double *lam_out, const double *lam_acc,
                      double *lam_in, double *lam_n);
/* One step of the plan's integrator backwards (rk1, rk2, rk3, vl2): u0 is the step's
 * input state (checkpoint), u1 scratch of the same size (receives the recomputed
 * stage-1 state), lam holds the cotangent of the step's output on entry and that of u0
 * on return.  rk3 recomputes its stage-2 state into the plan's third register. */
int aph_adjoint_step(aph_context *ctx, aph_plan *plan, double *u0, double *u1,
                     const double *dt_dev, double *lam);
/* The same with the step's stage-1 state u1 = stage1(u0) supplied by the caller (kept
 * when the forward step was taken or recomputed), so the first stage is not recomputed
 * (rk1 ignores u1; rk3 still recomputes its stage-2 state).  u0 and u1 are only read. */
int aph_adjoint_step_stored(aph_context *ctx, aph_plan *plan, const double *u0,
                            const double *u1, const double *dt_dev, double *lam);

#ifdef __cplusplus
}
#endif
#endif /* APOPHIS_HYDRO_H_ */
