/*
 * oracle.h — CPU FP64 restatement of Macaw's CMF2 propagation path.
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker for the HIP engine.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the engine never links or calls it.
 *
 * It consumes the same mg_problem descriptor as the engine
 * (include/macaw_gpu.h) and follows the reference routine by routine; each
 * function in oracle.c cites the file:line it restates.
 */
#ifndef MACAW_ORACLE_H
#define MACAW_ORACLE_H

#include <stdint.h>
#include "macaw_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc *orc_create(const mg_problem *p);
void orc_destroy(orc *o);
const char *orc_error(const orc *o);
int64_t orc_nconf(const orc *o);
int64_t orc_dimpsi(const orc *o);
int64_t orc_nspfmax(const orc *o);
int64_t orc_nopmax(const orc *o);
/* operator basis: nOper[nmodes], iOper[nmodes, nham] (1-based), IsIdentity[nmodes, nopmax] */
void orc_basis(const orc *o, int64_t *nOper, int64_t *iOper, int64_t *isIdentity);

void orc_set_state(orc *o, const double *A, const double *psi);
void orc_get_state(const orc *o, double *A, double *psi);
void orc_set_start(orc *o, const double *A, const double *psi);

void orc_calc_h1mode(orc *o, double *h1_out);
void orc_calc_rho(orc *o, double *rho_out);
int orc_calc_meanfield(orc *o, double *mf_out, double *rmf_out, double *energy_out);
void orc_apply_h(orc *o, const double *B, double *HB);
/* VMF (derivatives.f90:39-243, integrator.f90:36-221, propagator.f90:108-140):
 * derivatives of the current point; whole propagation (method 1 = Cash-Karp,
 * 2 = Dormand-Prince). */
int orc_calc_derivs(orc *o, const double *A, const double *psi, double *dA, double *dpsi);
int orc_propagate_vmf(orc *o, int64_t method, double t0, double tfinal, mg_output *outs, double *pops,
                      int64_t max_out, int64_t *n_out, mg_attempt *atts, int64_t max_att, int64_t *n_att);
/* Test hook: one Cash-Karp / Dormand-Prince step of y' = lam y from y = 1. */
void orc_rk_scalar_step(int64_t method, double lam_re, double lam_im, double dt, double *y5, double *y4);
/* Test hook: chunk count of the Gram reductions (summation order only). */
void orc_set_reduction_chunks(int64_t n);
void orc_der_spf(orc *o, int64_t im, const double *psi_in, double *dpsi_out);
double orc_norm_wf(orc *o, const double *A, const double *psi);
void orc_wf_ovl(orc *o, const double *A1, const double *psi1, const double *A2,
                const double *psi2, double *ovl);
/* Test hooks: DeltaPsi (gmctdh_objects.f90:206-227), and Calc_Err_SPF
 * (cmf2_derivatives.f90:239-334) of mode im (0-based) with the density
 * matrices of the last orc_calc_rho. */
double orc_delta_psi(orc *o, const double *A1, const double *psi1, const double *A2, const double *psi2);
double orc_err_spf(orc *o, int64_t im, const double *psi1, const double *psi2);
/* Lanczos (cmf2_integration.f90:30-103); returns nKrylov or -1 on overflow. */
int orc_lanczos(orc *o, const double *B0, double *BT, double dt, int build);
/* propagate_SPF (cmf2_base.f90:210-260) for DVR mode index i (1-based) on psi. */
int orc_propagate_spf(orc *o, double dt, int64_t i, double *psi);
int orc_cmf2_step(orc *o, double t, double htry, double *hdid, double *hnext,
                  mg_attempt *atts, int64_t max_att, int64_t *n_att);
int orc_observables(orc *o, double t, mg_output *out, double *pops);
int orc_propagate(orc *o, double t0, double tfinal, mg_output *outs, double *pops,
                  int64_t max_out, int64_t *n_out, mg_attempt *atts, int64_t max_att,
                  int64_t *n_att);
/* inorm_psi of propagate_wf (globals.f90:23-26): 0 never, 1 first output
 * interval, 2 every interval, 3 every step; used by orc_propagate and
 * orc_propagate_vmf. */
void orc_set_normalize(orc *o, int64_t mode);
/* Test hook: a planted mistake in cmf2 (0 none, 1 mean fields of t0 for the
 * second SPF half step, 2 A propagated with the t0 Hamiltonian). */
void orc_set_cmf2_fault(orc *o, int64_t fault);

/* Small dense kernels exposed for pinning tests. */
void orc_std_reg_inv(int64_t n, const double *A, double eps, int64_t type, double *B);
void orc_herm_eig(int64_t n, const double *A, double *w, double *U);
void orc_tridiag_eig(int64_t n, const double *alpha, const double *beta, double *w, double *U);

/* Gaussian (G-MCTDH) pieces: GauOvlMultiD, GauInt1D / Gau10Int1D (deriv = 1),
 * StdRegSol 'tikh', InverseSquareRootU, S00 and D0 of the current psi,
 * Calc_Der_SPF of a Gaussian mode, A_to_B / B_to_A. */
void orc_gau_ovl(int64_t nmd, const double *a1, const double *a2, const double *xi1, const double *xi2, double *out);
int orc_gau_int(double A1, double A2, const double *xi1, const double *xi2, int64_t op, int64_t deriv, double *out);
void orc_std_reg_sol_tikh(int64_t n, const double *A, const double *y, double eps, double *x);
int orc_inverse_sqrt_u(int64_t n, const double *S, double *D);
int orc_gwp_overlaps(orc *o, double *s00_out, double *dd_out);
int orc_der_spf_gwp(orc *o, int64_t im, const double *psi_in, const double *dd_in, double *dpsi_out,
                    double *ddd_out);
void orc_a_to_b(orc *o, const double *A, double *B, const double *dd);
int orc_prepare_gwp(orc *o);
void orc_b_to_a(orc *o, const double *B, double *A, const double *dd);

#ifdef __cplusplus
}
#endif
#endif
