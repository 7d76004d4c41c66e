/*
 * hydro_oracle.h -- CPU restatement of the AlexHls/apophis hydro flux path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is compared
 * against (and the CPU baseline that bench.py times beside it).  Product code must
 * never call it: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.
 *
 * Every function restates the reference's arithmetic in the reference's operation
 * order (build with -ffp-contract=off, no fast-math).  Layout and block descriptors
 * are those of include/apophis_hydro.h.
 */
#ifndef APOPHIS_HYDRO_ORACLE_H_
#define APOPHIS_HYDRO_ORACLE_H_

#include "../include/apophis_hydro.h"

#ifdef __cplusplus
extern "C" {
#endif

/* point functions (one interface / one face) */
void orc_plm(double q_im1, double q_i, double q_ip1, double *ql_ip1, double *qr_i);
void orc_ppm(double q_im2, double q_im1, double q_i, double q_ip1, double q_ip2,
             double *ql_ip1, double *qr_i);
/* w* = (rho, v_normal, v_t1, v_t2, p) already permuted; if* = (gamma_c, gamma_e). */
void orc_hlle(const double *wl, const double *wr, const double *ifl, const double *ifr,
              double *flx);
void orc_lhllc(const double *wl, const double *wr, const double *ifl, const double *ifr,
               double *flx);
void orc_hllc(const double *wl, const double *wr, const double *ifl, const double *ifr,
              double *flx);

/* mesh-level functions; nthreads <= 0 means "OpenMP default" */
void orc_set_threads(int nthreads);
void orc_cons_to_prim(const aph_block *blocks, int nb, int ndim, int nvar,
                      const aph_eos *eos, double *cons, double *prim, double *gamma);
int orc_calculate_fluxes(const aph_block *blocks, int nb, int ndim, int nvar, int recon,
                         int rsolver, const double *prim, const double *gamma,
                         double *fx1, double *fx2, double *fx3);
void orc_update(const aph_block *blocks, int nb, int ndim, int nvar, double *u0,
                const double *u1, const double *fx1, const double *fx2, const double *fx3,
                double gam0, double gam1, double beta_dt);
void orc_apply_gravity(const aph_block *blocks, int nb, int ndim, int nvar, double *cons,
                       const double *prim, const double *g, double dt);
void orc_fill_ghosts(const aph_block *blocks, int nb, int ndim, int nvar, double *cons);
double orc_estimate_timestep(const aph_block *blocks, int nb, int ndim, int nvar,
                             const aph_eos *eos, const double *prim, const double *cons,
                             double cfl);
double orc_global_sum(const aph_block *blocks, int nb, int ndim, int nvar,
                      const double *cons, int var);

/* One RK2 step with the stage sequence of ApophisDriver::MakeTaskCollection.
 * u0/prim/gamma must hold a consistent state (ghosts filled, ConsToPrim applied).
 * u1 and the flux arrays are scratch of the same size as u0.  Returns the new
 * EstimateTimestep value (cfl * min ...). */
double orc_rk2_step(const aph_block *blocks, int nb, int ndim, int nvar, int recon,
                    int rsolver, const aph_eos *eos, const double *g, double cfl,
                    double dt, double *u0, double *u1, double *prim, double *gamma,
                    double *fx1, double *fx2, double *fx3);
/* One step of `integrator` (1 rk1, 2 rk2, 3 rk3, 4 vl2 -- main.hpp values); the same
 * contract as orc_rk2_step.  orc_integrator_stages fills the per-stage coefficients and
 * returns the stage count (0 for an unknown integrator). */
double orc_step(const aph_block *blocks, int nb, int ndim, int nvar, int integrator,
                int recon, int rsolver, const aph_eos *eos, const double *g, double cfl,
                double dt, double *u0, double *u1, double *prim, double *gamma, double *fx1,
                double *fx2, double *fx3);
int orc_integrator_stages(int integrator, double *gam0, double *gam1, double *beta);

/* Monopole gravity (monopole_gravity.hpp): the per-call fixed-point scale, the bin
 * masses (returns the scale s), the radial tables, one point's g/phi, the field of every
 * interior cell (grav: 3 components, phi: 1 or NULL, pack layout), ApplyGravity with
 * that field, and one integrator step with it. */
int orc_mono_scale(int emax, long ncells);
int orc_mono_bin_masses(const aph_block *blocks, int nb, int ndim, int nvar,
                        const double *prim, int nrbin, double rmax, double *mbin);
void orc_mono_tables(int nrbin, int ndim, double rmax, double G, const double *mbin,
                     double *m_enc, double *phi_rad);
void orc_mono_point(int nrbin, int ndim, double rmax, double G, const double *m_enc,
                    const double *phi_rad, double x, double y, double z, double *g,
                    double *phi);
void orc_mono_field(const aph_block *blocks, int nb, int ndim, int nrbin, double rmax,
                    double G, const double *m_enc, const double *phi_rad, double *grav,
                    double *phi);
void orc_apply_gravity_field(const aph_block *blocks, int nb, int ndim, int nvar,
                             double *cons, const double *prim, const double *grav,
                             double dt);
double orc_step_monopole(const aph_block *blocks, int nb, int ndim, int nvar, int integrator,
                         int recon, int rsolver, const aph_eos *eos, int nrbin, double rmax,
                         double G, double cfl, double dt, double *u0, double *u1,
                         double *prim, double *gamma, double *fx1, double *fx2, double *fx3,
                         double *grav, double *m_enc, double *phi_rad);

#ifdef __cplusplus
}
#endif
#endif
