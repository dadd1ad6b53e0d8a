/*
 * macaw_gpu.h — C-ABI of the MI355X G-MCTDH propagation engine.
 *
 * The library replaces the CMF2 hot path of Macaw (DavidPicconi/Macaw,
 * Fortran 90).  Every entry point takes plain pointers and sizes, returns 0
 * on success and a non-zero status on error (message via mg_last_error), so
 * a Fortran driver can bind it through ISO_C_BINDING.  Integers are int64
 * (Macaw is built with -fdefault-integer-8, Makefile:18); complex numbers are
 * interleaved (re, im) doubles, i.e. Fortran `double complex`.  Arrays are
 * column-major with the Fortran index meaning given per field.
 *
 * One context = one HIP stream on the current device; a context is not
 * thread-safe.  State (A-vector, SPFs) stays resident on the device between
 * calls; host copies are made only by mg_set_state / mg_get_state and by the
 * host-pointer forms of the kernels.
 */
#ifndef MACAW_GPU_H
#define MACAW_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Problem descriptor: Macaw's module data after read_input and init_calc
 * (psidef.f90:11-55, dvr.f90:14-27, hamilton.f90:24-36, globals.f90:20-49).
 * Grid (DVR) modes and Gaussian (G-MCTDH, "gwp") modes; the Gaussian fields
 * sit at the end so a grid-only caller may leave them zero. */
typedef struct mg_problem {
    int64_t nstates, nmodes, ndof;
    int64_t maxdof;               /* leading dimension of iDOF            */
    const int64_t *nmodeDOF;      /* [nmodes]                              */
    const int64_t *iDOF;          /* [maxdof, nmodes], 1-based DOF ids     */
    const int64_t *MultiSet;      /* [nmodes], 1 = multi-set               */
    const int64_t *nSPF;          /* [nmodes, nstates]                     */
    const int64_t *ngp;           /* [ndof] grid points per DOF            */
    int64_t ngpmax;               /* leading dimension of the grid arrays  */
    const double *grid;           /* [ngpmax, ndof]                        */
    const double *d1;             /* [ngpmax, ngpmax, ndof] d/dx           */
    const double *d2;             /* [ngpmax, ngpmax, ndof] d^2/dx^2       */
    const double *xd1;            /* [ngpmax, ngpmax, ndof] x d/dx + d/dx x*/
    int64_t nham;
    const double *ae;             /* [2*nham] complex coefficients         */
    const int64_t *lhs_el;        /* [nham] 1-based                        */
    const int64_t *rhs_el;        /* [nham] 1-based                        */
    const int64_t *iOpDof;        /* [ndof, nham] primitive operator codes */
    double eps_int, eps_psi, eps_rho;
    double writetime;             /* output interval, a.u.                 */
    double min_stepsize;          /* a.u.                                  */
    int64_t int_type_psi;         /* 1 = Cash-Karp, 2 = Dormand-Prince     */
    /* Gaussian modes (psidef.f90:23-34,41; storage.f90; globals.f90:34-35).
     * dof_gwp[ndof]: 1 for a DOF of grid type "gwp" (all DOFs of a mode share
     * the type), or NULL when there is none.  gwp_a[maxdof, nspfmax, nmodes]:
     * the fixed width parameter A < 0 of DOF k of Gaussian j of mode m
     * (GWPa(k, j, i) of the i-th Gaussian mode; nspfmax = max nSPF); the
     * SPF block of a Gaussian mode holds the xi(k, j) (complex, nmodeDOF per
     * Gaussian) in psi.  eps_C, eps_ovl: Tikhonov parameters of the C-matrix
     * solve and of the Gaussian overlap inverse. */
    const int64_t *dof_gwp;
    const double *gwp_a;
    double eps_C, eps_ovl;
    /* Parametrised one-DOF operators (hamilton.f90:212-300; input.f90:950-
     * 1090): operator code 200 + p multiplies by gridoper(:, p, k) on the
     * grid of DVR DOF k, code -200 - p applies the projector phi phi^T with
     * phi = gridoper(:, p, k) (MultTensDiag / MultTensProj,
     * hamilton.f90:598-661).  gridoper[ngpmax, nopparam, ndof] is DOF-indexed
     * (Macaw's gridOper is indexed by DVR-dof: pass gridOper(:, p, indDVR(k))
     * for DOF k); nopparam = 0 and NULL when there is none.  Not supported on
     * Gaussian DOFs. */
    int64_t nopparam;
    const double *gridoper;
} mg_problem;

/* Observables written at every output time (output.f90:35-87,143-150). */
typedef struct mg_output {
    double time;                  /* a.u. */
    double auto_re, auto_im;      /* C(t) = <Psi(0)|Psi(t)>                */
    double norm;                  /* sum of state populations              */
    double energy;                /* a.u., energyMF / norm                 */
} mg_output;

/* Per-attempt CMF2 record (the reference keeps rejected attempts internal,
 * cmf2_base.f90:180-184; they are logged here). */
typedef struct mg_attempt {
    double time, h, err_A, err_psi;
    int64_t accepted, nkrylov1, nkrylov2;
} mg_attempt;

typedef struct mg_ctx mg_ctx;

/* ~ init_calc.f90:1-226 (tables) + alloc_h1mode/alloc_IsIdentity
 * (hamilton.f90:91-373) + SetupCMF2 (cmf2_base.f90:15-43).  precision:
 * 0 = complex FP64 everywhere; 1 = mixed: the H.A chains of every Krylov
 * product, the Krylov vectors after the start vector and the mean-field
 * chains in complex FP32, the A-vector, SPFs, density matrices, mean-field
 * matrices, reductions and DeltaPsi in FP64. */
int mg_create(const mg_problem *prob, int64_t precision, mg_ctx **ctx);
int mg_destroy(mg_ctx *ctx);
const char *mg_last_error(const mg_ctx *ctx);

/* Sizes: nAConftot (psidef.f90:14) and dimpsi (init_calc.f90:160-181);
 * nSPFmax (psidef.f90:15) and nOpMax (hamilton.f90:25), the leading
 * dimensions of the matrix outputs below. */
int64_t mg_nconf(const mg_ctx *ctx);
int64_t mg_dimpsi(const mg_ctx *ctx);
int mg_dims(const mg_ctx *ctx, int64_t *nspfmax, int64_t *nopmax);

/* Upload / download AVector and psi (psidef.f90:44-45); host pointers. */
int mg_set_state(mg_ctx *ctx, const double *A, const double *psi);
int mg_get_state(mg_ctx *ctx, double *A, double *psi);
/* AVector_Start / psi_Start for C(t) (init_calc.f90:222-223). */
int mg_set_start(mg_ctx *ctx, const double *A, const double *psi);

/* Calc_h1mode for all modes from the resident psi (hamilton.f90:416-531).
 * h1mode_out (host, optional): [nspfmax, nspfmax, nopmax, nmodes]. */
int mg_calc_h1mode(mg_ctx *ctx, double *h1mode_out);
/* Calc_rho1mode for all modes from the resident A (gmctdh_objects.f90:408-477).
 * rho_out (host, optional): [nspfmax, nspfmax, nmodes, nstates], full
 * Hermitian matrices. */
int mg_calc_rho(mg_ctx *ctx, double *rho_out);
/* Calc_MeanField, Hamiltonian part, + Calc_rhom1MF
 * (gmctdh_objects.f90:481-605, 1147-1194).  Requires mg_calc_h1mode and
 * mg_calc_rho first.  Outputs optional: [nspfmax, nspfmax, nopmax, nmodes]. */
int mg_calc_meanfield(mg_ctx *ctx, double *mf_out, double *rhom1mf_out);
/* Calc_DHD_B (cmf2_derivatives.f90:1-80): HB = H B with the current h1mode;
 * host pointers, complex [nconf]. */
int mg_apply_h(mg_ctx *ctx, const double *B, double *HB);
/* Same on device pointers, ordered on the context stream. */
int mg_apply_h_device(mg_ctx *ctx, const void *dB, void *dHB);
/* Calc_Der_SPF, DVR branch (cmf2_derivatives.f90:84-173), mode im (1-based)
 * with the current rho^-1 * MeanField; host pointers, complex [dimpsi]. */
int mg_der_spf(mg_ctx *ctx, int64_t im, const double *psi_in, double *dpsi_out);
/* One CMF2 step (cmf2_base.f90:47-206) on the resident state. */
int mg_cmf2_step(mg_ctx *ctx, double t, double htry, double *hdid, double *hnext);
/* WriteOutput observables (output.f90:17-150) for the resident state;
 * pops: [nstates]. */
int mg_observables(mg_ctx *ctx, double t, mg_output *out, double *pops);
/* propagate_wf, CMF2 branch (propagator.f90:87-106, 257-306): propagate the
 * resident state from t0 to tfinal, filling up to max_out output records,
 * pops [nstates, max_out] and up to max_att attempt records.  Counts are
 * returned in n_out / n_att. */
int mg_propagate(mg_ctx *ctx, double t0, double tfinal, mg_output *outs, double *pops,
                 int64_t max_out, int64_t *n_out, mg_attempt *atts, int64_t max_att,
                 int64_t *n_att);
/* inorm_psi of propagate_wf (globals.f90:23-26, propagator.f90:129-144,
 * 288-305; input keyword `normalize`): 0 never, 1 at the end of the first
 * output interval, 2 at the end of every output interval, 3 after every
 * integration step as well.  Renormalisation divides A by the square root of
 * the wavefunction norm (Trace_wf).  Applies to mg_propagate and
 * mg_propagate_vmf; mode 1 resets to 0 after its renormalisation. */
int mg_set_normalize(mg_ctx *ctx, int64_t mode);
/* Tile autotuner: times H.A and the mean fields on the resident state for
 * the LDS brick sizes (mode grouping) that fit, keeps the best (score =
 * 8 x H.A + 2 x mean fields, their weight in a CMF2 attempt) and returns
 * it.  reps <= 0: 2 timings per candidate.  No reference counterpart. */
int mg_autotune(mg_ctx *ctx, int64_t reps, int64_t *best_brick, double *best_score_ms);
/* Calc_Derivs (derivatives.f90:39-243), DVR modes: the VMF time derivatives
 * of the resident state, dA = -i H A (complex [nconf]) and the SPF equations
 * of motion of every mode (complex [dimpsi]); host pointers (null skips). */
int mg_calc_derivs(mg_ctx *ctx, double *dA_out, double *dpsi_out);
/* propagate_wf, Runge-Kutta (VMF) branch (propagator.f90:108-140) with
 * rkqs + rkck (method 1) or rkdp (method 2) (integrator.f90:36-221): the
 * resident state from t0 to tfinal, outputs and attempts as in mg_propagate
 * (an attempt's err_A holds its DeltaPsi error, err_psi is 0).  The
 * Dormand-Prince derivatives of the first step are evaluated at t0 (the
 * reference enters that step with arrays it never filled). */
int mg_propagate_vmf(mg_ctx *ctx, int64_t method, double t0, double tfinal, mg_output *outs, double *pops,
                     int64_t max_out, int64_t *n_out, mg_attempt *atts, int64_t max_att,
                     int64_t *n_att);
/* Algorithmic traffic (read B + write HB) and flops (8 per complex
 * multiply-add over all mode products) of one mg_apply_h. */
int mg_apply_h_stats(mg_ctx *ctx, double *bytes_alg, double *flops_alg);
/* One mg_apply_h_device with a HIP event pair around every kernel launch:
 * returns the summed kernel time (ms) and the number of launches.  Null
 * pointers select the context's own buffers (the last mg_apply_h input). */
int mg_apply_h_profile(mg_ctx *ctx, const void *dB, void *dHB, double *kernel_ms,
                       int64_t *nlaunch);
/* Replay: the next n CMF2 attempts use the step sizes and accept/reject
 * decisions of `log` (e.g. the attempt log of another run) instead of the
 * error controller, so two runs can be compared arithmetic-only.  The errors
 * are still computed and logged.  n = 0 clears.  No reference counterpart
 * (a test/measurement aid). */
int mg_set_replay(mg_ctx *ctx, const mg_attempt *log, int64_t n);
/* Current CMF2 step size carried between mg_propagate calls (a.u.). */
int mg_stepsize(const mg_ctx *ctx, double *h);
/* Sets that step size: the first attempt of the next mg_propagate /
 * mg_propagate_vmf call (propagate_wf's `stepsize`, propagator.f90:103,
 * which Macaw starts at tout * eps_int and carries through the run).  Lets a
 * run resumed from a saved state continue with the step it had reached.
 * h must be finite and > 0. */
int mg_set_stepsize(mg_ctx *ctx, double h);
/* Plain device copy of n complex elements of the context's work type
 * (complex FP32 for precision 1, FP64 otherwise) between two internal
 * buffers, event-timed; n <= nconf.  A traffic/bandwidth reference for
 * profiling (bytes moved = 2 n element size).  No reference counterpart. */
int mg_profile_copy(mg_ctx *ctx, int64_t n, double *kernel_ms);

/* Gaussian (G-MCTDH) modes.  mg_prepare_gwp: S00, h1mode, D0 =
 * InverseSquareRootU(S00) (maths.f90:621-644) and D^H h D (Calc_DHD,
 * hamilton.f90:535-574) of the current state, after which mg_apply_h is
 * Calc_DHD_B in the orthonormalised basis (cmf2_derivatives.f90:1-80).
 * mg_gwp_overlaps: Calc_S00 (gmctdh_objects.f90:323-381) as
 * S00[nspfmax, nspfmax, nmodes, nstates, nstates] and D0[nspfmax, nspfmax,
 * nmodes, nstates] of the current psi.  mg_der_spf_gwp: Calc_Der_SPF of
 * Gaussian mode im (1-based) at (psi_in, D = dd_in) with rho and the mean
 * fields of the current state (cmf2_derivatives.f90:175-232); dpsi_out gets
 * that mode's blocks, ddd_out dD.  mg_a_to_b: A_to_B (cmf2_base.f90:264-305),
 * B = D^H S A with the overlaps of the current psi. */
int mg_prepare_gwp(mg_ctx *ctx);
int mg_gwp_overlaps(mg_ctx *ctx, double *s00_out, double *dd_out);
int mg_der_spf_gwp(mg_ctx *ctx, int64_t im, const double *psi_in, const double *dd_in, double *dpsi_out,
                   double *ddd_out);
int mg_a_to_b(mg_ctx *ctx, const double *A, const double *dd, double *B);

/* Learned CMF2 step controller (no counterpart in the reference; the
 * reference's controller is cmf2_base.f90:178-199).  A 12-64-64-1 perceptron
 * (SiLU, BF16 weights, FP32 accumulation) maps 12 step features to
 * y = log(h_next / h) after each accepted step; the CMF2 error estimate still
 * accepts or rejects every attempt and rejections use the classical rule.
 * Parameters: MG_MLP_NPARAM floats (feature means and deviations, W1 [12x64],
 * b1, W2 [64x64], b2, w3 [64], b3).  mg_mlp_train trains on n rows X[n x 12]
 * -> t[n] (mean squared error, Adam, minibatches of `batch`, seeded
 * shuffles) on the device; mg_mlp_forward / mg_mlp_grad evaluate the model
 * and its loss gradient; mg_set_step_model(ctx, NULL) restores the classical
 * controller.  mg_log_step_features records the features of every accepted
 * step of the following propagations (mg_step_features reads them with the
 * accepted step sizes). */
/* Ensemble of independent wavepackets (no counterpart in the reference):
 * member i is the problem probs[i] with initial state (A0[i], psi0[i])
 * (complex interleaved, as mg_set_state).  `nthreads` host threads each
 * create a member's context, propagate it over [t0, tfinal] and destroy it;
 * members start largest first and are admitted while their estimated device
 * bytes fit in mem_fraction of the free device memory.  flags bit 0: run
 * mg_autotune (1 repetition) on each member first; bit 1: a progress line
 * on stderr as each member finishes.  Per member i:
 * outs[i * max_out + k], pops[(i * max_out + k) * 8 + s] (states beyond 8
 * not reported; pops may be NULL), n_out[i], status[i] (0 = success) and,
 * if member_wall is not NULL, member_wall[i] = its wall seconds.  The first
 * failure's message goes to errbuf.  Returns 0 if every member succeeded,
 * 1 if any failed, -1 for bad arguments. */
int mg_ensemble_propagate(const mg_problem *const *probs, const double *const *A0, const double *const *psi0,
                          int64_t n, int64_t precision, int64_t nthreads, double mem_fraction, int64_t flags,
                          double t0, double tfinal, mg_output *outs, double *pops, int64_t max_out, int64_t *n_out,
                          int64_t *status, double *member_wall, char *errbuf, int64_t errlen);

/* HIP graphs (no counterpart in the reference): every cached launch plan
 * (H.A schedule, mean-field, density-matrix and Kronecker-chain brick
 * programs) is captured into its own graph on its second use and replayed
 * from then on; enable = 0 launches kernel by kernel (default 1; the
 * environment variable MG_GRAPHS=0 sets the default off).  With
 * MG_ATTEMPT_GRAPH=1 a whole CMF2 attempt (every decision on the device) is
 * captured into one graph instead; off by default because its replay
 * serialises the SPF half steps that otherwise overlap the Krylov builds on
 * a second stream, which made it 11% slower on a small problem and no faster
 * on the others measured.  Results are bit-identical either way. */
int mg_set_graphs(mg_ctx *ctx, int64_t enable);

#define MG_MLP_NFEAT 12
#define MG_MLP_NPARAM 5081
/* Largest batch of mg_mlp_forward_noalloc. */
#define MG_MLP_NOALLOC_MAX_ROWS 4096
int mg_mlp_train(mg_ctx *ctx, const float *X, const float *t, int64_t n, int64_t epochs, int64_t batch, double lr,
                 uint64_t seed, float *params, double *loss);
int mg_mlp_forward(mg_ctx *ctx, const float *params, const float *X, int64_t n, float *y);
int mg_mlp_grad(mg_ctx *ctx, const float *params, const float *X, const float *t, int64_t n, float *grad,
                double *loss);
int mg_set_step_model(mg_ctx *ctx, const float *params);
/* mg_mlp_forward without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments and results as mg_mlp_forward for 1 <= n <=
 * MG_MLP_NOALLOC_MAX_ROWS (larger n is refused with an error); its device
 * memory is a fixed 6.6 MB array held by every GPU the library runs on until
 * the process ends.  All calls on one GPU share that array: calls that
 * overlap on one GPU (several threads or contexts at once) race. */
int mg_mlp_forward_noalloc(mg_ctx *ctx, const float *params, const float *X, int64_t n, float *y);
int mg_log_step_features(mg_ctx *ctx, int64_t enable);
int mg_step_features(mg_ctx *ctx, float *features, double *h, int64_t max, int64_t *n);
int mg_controller_stats(mg_ctx *ctx, int64_t *proposals, int64_t *fallbacks);

#ifdef __cplusplus
}
#endif

#endif /* MACAW_GPU_H */
