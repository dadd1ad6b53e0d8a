/*
 * oracle.c — CPU FP64 restatement of Macaw's CMF2 propagation path
 * (DavidPicconi/Macaw, source/ *.f90): DVR (grid) modes and frozen-width
 * Gaussian (G-MCTDH) modes, multi-set and single-set; CMF2 and the
 * variable-mean-field Runge-Kutta branch.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Nothing in the
 * engine links or calls it.
 *
 * Complex arithmetic is C99 `double complex`.  The reference works on
 * conj(A) and right-multiplies (e.g. cmf2_derivatives.f90:20-21,77); that is
 * exactly the complex conjugate of left-multiplying A, which is what this
 * file does (conjugation is exact in IEEE arithmetic).  Every mode product
 * below is  out[l, i, r] = sum_j M[i, j] in[l, j, r]  over the mode index.
 *
 * Parity pins: tests/test_oracle_pins.py (brute-force Kronecker H, einsum
 * mean fields, eigh-based regularised inverses, expm Lanczos, DeltaPsi and
 * the SPF error norm against grid-space vectors, Runge-Kutta orders,
 * full-basis CMF2 and VMF == exact propagation) and tests/test_gmctdh_pins.py
 * (Gaussian integrals and overlaps against quadrature, the Tikhonov solve and
 * Cholesky inverse against numpy, coherent-state and separable closed forms).
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;

#define NKRYLOV_MAX 20            /* cmf2_integration.f90:12 */
static const double EPS_B = 1e-8; /* cmf2_integration.f90:13 */

typedef struct {
    int64_t lhs, rhs;             /* electronic labels (1-based)      */
    int64_t codes[8];             /* primitive codes per DOF of mode  */
} oper_t;

struct orc {
    mg_problem p;
    /* owned copies of the descriptor arrays */
    int64_t *nmodeDOF, *iDOF, *MultiSet, *nSPF, *ngp, *lhs, *rhs, *iOpDof;
    double *grid, *d1, *d2, *xd1;
    int64_t nopparam;
    double *gridoper;             /* [ngpmax, nopparam, ndof] parametrised operators on the grids */
    cplx *ae;
    int64_t ns, nm, nham, nspfmax, nopmax, ngpmax;
    int64_t *nAConf, *AStateOffset, nconf, dimpsi;
    int64_t *ipsiSPF;             /* [nm, ns] 1-based */
    int64_t *ngpm;                /* [nm] */
    int64_t *nOper;               /* [nm] */
    oper_t *Oper;                 /* [nm * nopmax] */
    int64_t *iOper;               /* [nm * nham] 1-based op index */
    int64_t *IsIdentity;          /* [nm * nopmax] */
    cplx *A, *psi, *Astart, *psistart;
    cplx *h1;                     /* [nspfmax^2 * nopmax * nm] */
    cplx *rho;                    /* [nspfmax^2 * nm * ns] full Hermitian */
    cplx *rho_ave;                /* [nspfmax^2 * nm] */
    cplx *mf, *rmf;               /* [nspfmax^2 * nopmax * nm] */
    double energyMF;
    int64_t inorm_psi;            /* propagate_wf renormalisation (globals.f90:23-26) */
    int64_t cmf2_fault;           /* test hook: a planted mistake in cmf2 (orc_set_cmf2_fault) */
    /* Lanczos storage */
    cplx *BK;                     /* [nconf * NKRYLOV_MAX] */
    double lan_alpha[NKRYLOV_MAX], lan_U[NKRYLOV_MAX * NKRYLOV_MAX];
    int nkrylov;
    cplx der_factor;
    /* Gaussian (G-MCTDH) modes: psidef.f90:23-34,41, storage.f90:10-42,
     * hamilton.f90:32 (Dh1modeD), cmf2_base.f90:28-29 (DD). */
    int64_t ngwp;                 /* number of Gaussian modes */
    int64_t *gwp;                 /* [nm] 1 = Gaussian mode */
    int64_t *npar;                /* [nm] parameters per SPF: grid points or nmodeDOF */
    double *gwpa;                 /* [maxdof * nspfmax * nm] widths A < 0 */
    cplx *S00;                    /* [nspfmax^2 * nm * ns * ns] overlaps, S00(j1,j2,m,s1,s2) */
    cplx *S00m1;                  /* [nspfmax^2 * nm * ns] regularised inverses */
    cplx *DD;                     /* [nspfmax^2 * nm * ns] D matrices of the current state */
    cplx *DHD;                    /* [nspfmax^2 * nopmax * nm] D^H h D */
    char err[256];
};

#define H1(o, m, op) ((o)->h1 + (size_t)(o)->nspfmax * (o)->nspfmax * ((op) + (o)->nopmax * (m)))
#define MF(o, m, op) ((o)->mf + (size_t)(o)->nspfmax * (o)->nspfmax * ((op) + (o)->nopmax * (m)))
#define RMF(o, m, op) ((o)->rmf + (size_t)(o)->nspfmax * (o)->nspfmax * ((op) + (o)->nopmax * (m)))
#define RHO(o, m, s) ((o)->rho + (size_t)(o)->nspfmax * (o)->nspfmax * ((m) + (o)->nm * (s)))
#define RHOAVE(o, m) ((o)->rho_ave + (size_t)(o)->nspfmax * (o)->nspfmax * (m))
#define NSPF(o, m, s) ((o)->nSPF[(m) + (o)->nm * (s)])
#define NN(o) ((size_t)(o)->nspfmax * (o)->nspfmax)
#define S00P(o, m, s1, s2) ((o)->S00 + NN(o) * ((m) + (o)->nm * ((s1) + (o)->ns * (s2))))
#define S00M1(o, m, s) ((o)->S00m1 + NN(o) * ((m) + (o)->nm * (s)))
#define DDP(o, dd, m, s) ((dd) + NN(o) * ((m) + (o)->nm * (s)))
#define DHDP(o, m, op) ((o)->DHD + NN(o) * ((op) + (o)->nopmax * (m)))
#define GWPA(o, k, j, m) ((o)->gwpa[(k) + (o)->p.maxdof * ((j) + (o)->nspfmax * (m))])
#define NSETS(o, m) ((o)->MultiSet[m] ? (o)->ns : 1)
#define IPSI(o, m, s) ((o)->ipsiSPF[(m) + (o)->nm * (s)])

static void *xcalloc(size_t n, size_t sz) {
    void *v = calloc(n ? n : 1, sz);
    if (!v) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return v;
}

static int64_t *dup_i64(const int64_t *src, size_t n) {
    int64_t *d = xcalloc(n, sizeof(int64_t));
    memcpy(d, src, n * sizeof(int64_t));
    return d;
}

static double *dup_f64(const double *src, size_t n) {
    double *d = xcalloc(n, sizeof(double));
    memcpy(d, src, n * sizeof(double));
    return d;
}

/* ---------------------------------------------------------------------- */
/* Mode product on a tensor with per-mode sizes `sz` (mode 0 fastest):
 * out[l, i, r] = sum_j M[i + ldm*j] in[l, j, r]; i < nout, j < sz[k].
 * Restates Mult_HeM_MatV / Mult_GM_MatV (gmctdh_objects.f90:1407-1468) in
 * left-multiplication form. */
static void mode_product(int64_t nm, const int64_t *sz, int64_t k, const cplx *M, int64_t ldm,
                         int64_t nout, const cplx *in, cplx *out) {
    int64_t ll = 1, nl = 1;
    for (int64_t q = 0; q < k; q++) ll *= sz[q];
    for (int64_t q = k + 1; q < nm; q++) nl *= sz[q];
    const int64_t nin = sz[k];
    const int64_t chunk = 256;
    const int64_t nch = (ll + chunk - 1) / chunk;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t r = 0; r < nl; r++)
        for (int64_t c = 0; c < nch; c++) {
            const int64_t l0 = c * chunk, l1 = (l0 + chunk < ll) ? l0 + chunk : ll;
            const cplx *ib = in + r * ll * nin;
            cplx *ob = out + r * ll * nout;
            for (int64_t i = 0; i < nout; i++) {
                cplx *o = ob + i * ll;
                for (int64_t l = l0; l < l1; l++) o[l] = 0;
                for (int64_t j = 0; j < nin; j++) {
                    const cplx mij = M[i + ldm * j];
                    const cplx *x = ib + j * ll;
                    for (int64_t l = l0; l < l1; l++) o[l] += mij * x[l];
                }
            }
        }
}

/* Hole contraction over mode k: G[j1 + ldg*j2] += alpha * sum_{l,r}
 * conj(X[l, j1, r]) Y[l, j2, r].  X has n1 values on mode k, Y has n2, both
 * share ll (modes < k) and nl (modes > k).  Restates the zgemm('T','N') /
 * zher2k blocks of gmctdh_objects.f90:452-457, 592-600.
 * The r range is cut into a fixed number of chunks (reduction_chunks) whose
 * partial sums are added in chunk order, so the result does not depend on the
 * thread count or scheduling; changing the chunk count changes only the
 * summation order (a test of how rounding alone moves a trajectory). */
static int64_t reduction_chunks = 64;

static void hole_gram(int64_t ll, int64_t nl, int64_t n1, int64_t n2, const cplx *X,
                      const cplx *Y, cplx alpha, cplx *G, int64_t ldg) {
    const int64_t nc = nl < reduction_chunks ? nl : reduction_chunks;
    cplx *part = (cplx *)malloc(sizeof(cplx) * (size_t)(nc > 0 ? nc : 1));
    for (int64_t j2 = 0; j2 < n2; j2++) {
        for (int64_t j1 = 0; j1 < n1; j1++) {
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < nc; c++) {
                const int64_t r0 = nl * c / nc, r1 = nl * (c + 1) / nc;
                cplx acc = 0;
                for (int64_t r = r0; r < r1; r++) {
                    const cplx *x = X + (r * n1 + j1) * ll;
                    const cplx *y = Y + (r * n2 + j2) * ll;
                    cplx a = 0;
                    for (int64_t l = 0; l < ll; l++) a += conj(x[l]) * y[l];
                    acc += a;
                }
                part[c] = acc;
            }
            cplx acc = 0;
            for (int64_t c = 0; c < nc; c++) acc += part[c];
            G[j1 + ldg * j2] += alpha * acc;
        }
    }
    free(part);
}

/* ---------------------------------------------------------------------- */
/* Small dense linear algebra (replaces zheev / dstev / zposv). */

/* Cyclic complex Jacobi eigen-decomposition of the Hermitian n x n matrix A
 * (column-major, full).  w ascending, U columns = eigenvectors. */
static void herm_eig(int64_t n, const cplx *Ain, double *w, cplx *U) {
    cplx *A = xcalloc((size_t)n * n, sizeof(cplx));
    memcpy(A, Ain, sizeof(cplx) * n * n);
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) U[i + n * j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < n; i++) {
                double a2 = creal(A[i + n * j] * conj(A[i + n * j]));
                if (i != j) off += a2; else diag += a2;
            }
        if (off <= 1e-32 * (diag + 1e-300) || off == 0.0) break;
        for (int64_t p = 0; p < n - 1; p++)
            for (int64_t q = p + 1; q < n; q++) {
                cplx apq = A[p + n * q];
                double aabs = cabs(apq);
                if (aabs == 0.0) continue;
                cplx ph = apq / aabs;                /* e^{i phi} */
                double app = creal(A[p + n * p]), aqq = creal(A[q + n * q]);
                double theta = (aqq - app) / (2.0 * aabs);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                /* J: J[p][p]=c, J[q][q]=c*conj(ph), J[p][q]=s, J[q][p]=-s*conj(ph) */
                cplx jpp = c, jqq = c * conj(ph), jpq = s, jqp = -s * conj(ph);
                /* A <- A J (columns p, q) */
                for (int64_t k = 0; k < n; k++) {
                    cplx akp = A[k + n * p], akq = A[k + n * q];
                    A[k + n * p] = akp * jpp + akq * jqp;
                    A[k + n * q] = akp * jpq + akq * jqq;
                }
                /* A <- J^H A (rows p, q) */
                for (int64_t k = 0; k < n; k++) {
                    cplx apk = A[p + n * k], aqk = A[q + n * k];
                    A[p + n * k] = conj(jpp) * apk + conj(jqp) * aqk;
                    A[q + n * k] = conj(jpq) * apk + conj(jqq) * aqk;
                }
                A[p + n * q] = 0;
                A[q + n * p] = 0;
                A[p + n * p] = creal(A[p + n * p]);
                A[q + n * q] = creal(A[q + n * q]);
                for (int64_t k = 0; k < n; k++) {
                    cplx ukp = U[k + n * p], ukq = U[k + n * q];
                    U[k + n * p] = ukp * jpp + ukq * jqp;
                    U[k + n * q] = ukp * jpq + ukq * jqq;
                }
            }
    }
    for (int64_t i = 0; i < n; i++) w[i] = creal(A[i + n * i]);
    /* sort ascending */
    for (int64_t i = 0; i < n; i++) {
        int64_t k = i;
        for (int64_t j = i + 1; j < n; j++) if (w[j] < w[k]) k = j;
        if (k != i) {
            double tw = w[i]; w[i] = w[k]; w[k] = tw;
            for (int64_t r = 0; r < n; r++) {
                cplx tu = U[r + n * i]; U[r + n * i] = U[r + n * k]; U[r + n * k] = tu;
            }
        }
    }
    free(A);
}

/* StdRegInv (maths.f90:8-83): B = U f(W) U^H with f from the regularisation
 * type: 1 'tikh' w/(w^2+eps^2); 2 'poly' (w<eps -> (eps+w^2/eps)/2, then 1/w). */
static void std_reg_inv(int64_t n, const cplx *A, double eps, int type, cplx *B) {
    double *w = xcalloc(n, sizeof(double));
    cplx *U = xcalloc((size_t)n * n, sizeof(cplx));
    herm_eig(n, A, w, U);
    for (int64_t j = 0; j < n; j++) {
        if (type == 1) {
            w[j] = w[j] / (w[j] * w[j] + eps * eps);
        } else {
            if (w[j] < eps) w[j] = 0.5 * (eps + w[j] * w[j] / eps);
            w[j] = 1.0 / w[j];
        }
    }
    for (int64_t j2 = 0; j2 < n; j2++)
        for (int64_t j1 = 0; j1 < n; j1++) {
            cplx acc = 0;
            for (int64_t k = 0; k < n; k++) acc += w[k] * U[j1 + n * k] * conj(U[j2 + n * k]);
            B[j1 + n * j2] = acc;
        }
    for (int64_t j = 0; j < n; j++) B[j + n * j] = creal(B[j + n * j]);
    for (int64_t j2 = 0; j2 < n; j2++)
        for (int64_t j1 = j2 + 1; j1 < n; j1++) B[j1 + n * j2] = conj(B[j2 + n * j1]);
    free(w);
    free(U);
}

/* Cholesky solve S X = Y for Hermitian positive definite S (zposv 'U',
 * cmf2_derivatives.f90:170).  S: n x n (full), Y: n x m overwritten by X. */
static int chol_solve(int64_t n, const cplx *S, int64_t lds, cplx *Y, int64_t ldy, int64_t m) {
    cplx *U = xcalloc((size_t)n * n, sizeof(cplx));
    for (int64_t j = 0; j < n; j++) {
        double d = creal(S[j + lds * j]);
        for (int64_t k = 0; k < j; k++) d -= creal(conj(U[k + n * j]) * U[k + n * j]);
        if (!(d > 0.0)) { free(U); return -1; }
        double ujj = sqrt(d);
        U[j + n * j] = ujj;
        for (int64_t i = j + 1; i < n; i++) {
            cplx v = S[j + lds * i];
            for (int64_t k = 0; k < j; k++) v -= conj(U[k + n * j]) * U[k + n * i];
            U[j + n * i] = v / ujj;
        }
    }
    for (int64_t c = 0; c < m; c++) {
        cplx *y = Y + ldy * c;
        /* U^H z = y */
        for (int64_t i = 0; i < n; i++) {
            cplx v = y[i];
            for (int64_t k = 0; k < i; k++) v -= conj(U[k + n * i]) * y[k];
            y[i] = v / U[i + n * i];
        }
        /* U x = z */
        for (int64_t i = n - 1; i >= 0; i--) {
            cplx v = y[i];
            for (int64_t k = i + 1; k < n; k++) v -= U[i + n * k] * y[k];
            y[i] = v / U[i + n * i];
        }
    }
    free(U);
    return 0;
}

/* Real symmetric tridiagonal eigenproblem (replaces dstev, cmf2_integration.f90:85)
 * by cyclic Jacobi on the dense matrix; w ascending, U column eigenvectors. */
static void tridiag_eig(int64_t n, const double *alpha, const double *beta, double *w, double *U) {
    double A[NKRYLOV_MAX * NKRYLOV_MAX];
    for (int64_t i = 0; i < n * n; i++) A[i] = 0.0;
    for (int64_t i = 0; i < n; i++) {
        A[i + n * i] = alpha[i];
        if (i + 1 < n) { A[i + n * (i + 1)] = beta[i]; A[(i + 1) + n * i] = beta[i]; }
    }
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) U[i + n * j] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, dg = 0.0;
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < n; i++) {
                if (i != j) off += A[i + n * j] * A[i + n * j]; else dg += A[i + n * j] * A[i + n * j];
            }
        if (off <= 1e-32 * (dg + 1e-300) || off == 0.0) break;
        for (int64_t p = 0; p < n - 1; p++)
            for (int64_t q = p + 1; q < n; q++) {
                double apq = A[p + n * q];
                if (apq == 0.0) continue;
                double theta = (A[q + n * q] - A[p + n * p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int64_t k = 0; k < n; k++) {
                    double akp = A[k + n * p], akq = A[k + n * q];
                    A[k + n * p] = c * akp - s * akq;
                    A[k + n * q] = s * akp + c * akq;
                }
                for (int64_t k = 0; k < n; k++) {
                    double apk = A[p + n * k], aqk = A[q + n * k];
                    A[p + n * k] = c * apk - s * aqk;
                    A[q + n * k] = s * apk + c * aqk;
                }
                A[p + n * q] = A[q + n * p] = 0.0;
                for (int64_t k = 0; k < n; k++) {
                    double ukp = U[k + n * p], ukq = U[k + n * q];
                    U[k + n * p] = c * ukp - s * ukq;
                    U[k + n * q] = s * ukp + c * ukq;
                }
            }
    }
    for (int64_t i = 0; i < n; i++) w[i] = A[i + n * i];
    for (int64_t i = 0; i < n; i++) {
        int64_t k = i;
        for (int64_t j = i + 1; j < n; j++) if (w[j] < w[k]) k = j;
        if (k != i) {
            double tw = w[i]; w[i] = w[k]; w[k] = tw;
            for (int64_t r = 0; r < n; r++) { double tu = U[r + n * i]; U[r + n * i] = U[r + n * k]; U[r + n * k] = tu; }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Gaussian integrals.  g = exp(A x^2 + xi x), A < 0 real, xi complex; the
 * normalised g has the real factor (-2A/pi)^(1/4) exp(Re(xi)^2 / (4A)). */

/* <g1|g2> of normalised nmd-dimensional Gaussians (GauOvlMultiD,
 * gmctdh_objects.f90:1497-1519, with its equal-width branch :1509-1513). */
static cplx gau_ovl(const double *a1, const double *a2, const cplx *xi1, const cplx *xi2, int64_t nmd,
                    int64_t sa1, int64_t sa2) {
    int equal = 1;
    for (int64_t k = 0; k < nmd; k++) if (!(fabs(a1[k * sa1] - a2[k * sa2]) < 2.220446049250313e-16)) equal = 0;
    cplx c = 0;
    if (equal) {
        for (int64_t k = 0; k < nmd; k++) {
            const cplx x = conj(xi1[k]) + xi2[k];
            c += (creal(xi1[k]) * creal(xi1[k]) + creal(xi2[k]) * creal(xi2[k]) - 0.5 * x * x) / a1[k * sa1];
        }
        return cexp(0.25 * c);
    }
    double pr = 1.0;
    for (int64_t k = 0; k < nmd; k++) {
        const double A1 = a1[k * sa1], A2 = a2[k * sa2];
        const cplx x = conj(xi1[k]) + xi2[k];
        c += creal(xi1[k]) * creal(xi1[k]) / A1 + creal(xi2[k]) * creal(xi2[k]) / A2 - x * x / (A1 + A2);
        pr *= 4.0 * A1 * A2 / ((A1 + A2) * (A1 + A2));
    }
    return cexp(0.25 * c) * sqrt(sqrt(pr));
}

/* Number of pairings: binomial(n, 2k) * (2k-1)!!, the moment coefficients
 * of GauInt1D (hamilton.f90:761-785 tabulates them for n <= 20). */
static double moment_coeff(int k, int n) {
    double b = 1.0;
    for (int i = 0; i < 2 * k; i++) b = b * (double)(n - i) / (double)(i + 1);
    double df = 1.0;
    for (int i = 2 * k - 1; i > 1; i -= 2) df *= (double)i;
    return b * df;
}

static cplx cpow_int(cplx z, int n) {
    cplx r = 1.0;
    for (int i = 0; i < n; i++) r *= z;
    return r;
}

/* <g1|O|g2> / <g1|g2> for the primitive operator code op (GauInt1D,
 * hamilton.f90:745-849): 0 identity, -1 p, -2 d^2/dq^2, -3 qp + pq,
 * 1..20 q^n.  Parametrised operators (codes > 200) are not supported:
 * *bad is set. */
static cplx gau_int(double A1, double A2, cplx xi1, cplx xi2, int64_t op, int *bad) {
    const cplx x1c = conj(xi1);
    if (op == 0) return 1.0;
    if (op == -1) return -I * (A1 * xi2 - A2 * x1c) / (A1 + A2);
    if (op == -2) {
        const double s = A1 + A2;
        const cplx t = (A2 * x1c - A1 * xi2) / s;
        return t * t + 2.0 * A1 * A2 / s;
    }
    if (op == -3) {
        const cplx g = A1 * A1 - A2 * A2 + (A2 * x1c - A1 * xi2) * (x1c + xi2);
        return -I * g / ((A1 + A2) * (A1 + A2));
    }
    if (op == 1) return (x1c + xi2) / (-2.0 * (A1 + A2));
    if (op == 2) {
        const double s2 = 2.0 * (A1 + A2);
        const cplx t = (x1c + xi2) / s2;
        return t * t - 1.0 / s2;
    }
    if (op >= 3 && op <= 20) {
        const int n = (int)op;
        const cplx c = x1c + xi2;
        const double a = -2.0 * (A1 + A2);
        cplx g = 0;
        if (n % 2) {
            for (int k = 0; k <= n / 2; k++)
                g += moment_coeff(k, n) / pow(a, (double)(n - k)) * cpow_int(c, n - 2 * k);
        } else {
            const cplx c2 = c * c / a;
            cplx c1 = 1.0;
            for (int k = n / 2; k >= 0; k--) {
                g += moment_coeff(k, n) * c1;
                c1 *= c2;
            }
            g /= pow(a, (double)(n / 2));
        }
        return g;
    }
    if (bad) *bad = 1;
    return 0;
}

/* <dg1/dxi1|O|g2> / <g1|g2> (Gau10Int1D, hamilton.f90:853-902); the
 * derivative is taken of the unnormalised g1 = exp(A1 x^2 + xi1 x). */
static cplx gau10_int(double A1, double A2, cplx xi1, cplx xi2, int64_t op, int *bad) {
    if (op == -1)
        return -I * xi2 * gau_int(A1, A2, xi1, xi2, 1, bad) - I * (2.0 * A2) * gau_int(A1, A2, xi1, xi2, 2, bad);
    if (op == -2)
        return (A2 + A2 + xi2 * xi2) * gau_int(A1, A2, xi1, xi2, 1, bad) +
               4.0 * A2 * xi2 * gau_int(A1, A2, xi1, xi2, 2, bad) +
               4.0 * A2 * A2 * gau_int(A1, A2, xi1, xi2, 3, bad);
    if (op == -3)
        return -I * (gau_int(A1, A2, xi1, xi2, 1, bad) + 2.0 * xi2 * gau_int(A1, A2, xi1, xi2, 2, bad) +
                     4.0 * A2 * gau_int(A1, A2, xi1, xi2, 3, bad));
    if (op >= 0 && op <= 19) return gau_int(A1, A2, xi1, xi2, op + 1, bad);
    if (bad) *bad = 1;
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Setup: layout (init_calc.f90:25-34, 160-181) and operator basis
 * (alloc_h1mode hamilton.f90:91-147, alloc_IsIdentity :353-373). */

orc *orc_create(const mg_problem *p) {
    orc *o = xcalloc(1, sizeof(orc));
    o->p = *p;
    o->ns = p->nstates;
    o->nm = p->nmodes;
    o->nham = p->nham;
    o->ngpmax = p->ngpmax;
    const int64_t nm = o->nm, ns = o->ns, nd = p->ndof;
    o->nmodeDOF = dup_i64(p->nmodeDOF, nm);
    o->iDOF = dup_i64(p->iDOF, (size_t)p->maxdof * nm);
    o->MultiSet = dup_i64(p->MultiSet, nm);
    o->nSPF = dup_i64(p->nSPF, (size_t)nm * ns);
    o->ngp = dup_i64(p->ngp, nd);
    o->lhs = dup_i64(p->lhs_el, p->nham);
    o->rhs = dup_i64(p->rhs_el, p->nham);
    o->iOpDof = dup_i64(p->iOpDof, (size_t)nd * p->nham);
    o->grid = dup_f64(p->grid, (size_t)p->ngpmax * nd);
    o->d1 = dup_f64(p->d1, (size_t)p->ngpmax * p->ngpmax * nd);
    o->d2 = dup_f64(p->d2, (size_t)p->ngpmax * p->ngpmax * nd);
    o->xd1 = dup_f64(p->xd1, (size_t)p->ngpmax * p->ngpmax * nd);
    o->nopparam = p->nopparam > 0 ? p->nopparam : 0;
    o->gridoper = o->nopparam ? dup_f64(p->gridoper, (size_t)p->ngpmax * o->nopparam * nd) : NULL;
    o->ae = xcalloc(p->nham, sizeof(cplx));
    for (int64_t t = 0; t < p->nham; t++) o->ae[t] = p->ae[2 * t] + I * p->ae[2 * t + 1];
    o->der_factor = -I;
    o->nspfmax = 0;
    for (int64_t i = 0; i < nm * ns; i++) if (o->nSPF[i] > o->nspfmax) o->nspfmax = o->nSPF[i];
    o->nAConf = xcalloc(ns, sizeof(int64_t));
    o->AStateOffset = xcalloc(ns, sizeof(int64_t));
    int64_t ic = 0;
    for (int64_t s = 0; s < ns; s++) {
        int64_t n = 1;
        for (int64_t m = 0; m < nm; m++) n *= NSPF(o, m, s);
        o->nAConf[s] = n;
        o->AStateOffset[s] = ic;
        ic += n;
    }
    o->nconf = ic;
    o->ngpm = xcalloc(nm, sizeof(int64_t));
    for (int64_t m = 0; m < nm; m++) {
        int64_t g = 1;
        for (int64_t k = 0; k < o->nmodeDOF[m]; k++) g *= o->ngp[o->iDOF[k + p->maxdof * m] - 1];
        o->ngpm[m] = g;
    }
    /* Gaussian modes: a mode is Gaussian when its first DOF is (init_calc.f90:56-75);
     * its SPF block holds nmodeDOF complex xi per Gaussian (init_calc.f90:169-171) */
    o->gwp = xcalloc(nm, sizeof(int64_t));
    o->npar = xcalloc(nm, sizeof(int64_t));
    for (int64_t m = 0; m < nm; m++) {
        o->gwp[m] = p->dof_gwp ? (p->dof_gwp[o->iDOF[p->maxdof * m] - 1] != 0) : 0;
        o->npar[m] = o->gwp[m] ? o->nmodeDOF[m] : o->ngpm[m];
        o->ngwp += o->gwp[m];
    }
    if (o->ngwp && !p->gwp_a) snprintf(o->err, sizeof o->err, "Gaussian modes need gwp_a");
    o->ipsiSPF = xcalloc((size_t)nm * ns, sizeof(int64_t));
    int64_t dim = 0;
    for (int64_t s = 0; s < ns; s++)
        for (int64_t m = 0; m < nm; m++) {
            if (!o->MultiSet[m] && s > 0) { IPSI(o, m, s) = IPSI(o, m, 0); continue; }
            IPSI(o, m, s) = dim + 1;
            dim += o->npar[m] * NSPF(o, m, s);
        }
    o->dimpsi = dim;
    /* operator basis */
    oper_t *tmp = xcalloc((size_t)nm * o->nham, sizeof(oper_t));
    o->nOper = xcalloc(nm, sizeof(int64_t));
    o->iOper = xcalloc((size_t)nm * o->nham, sizeof(int64_t));
    for (int64_t m = 0; m < nm; m++) {
        int64_t nop = 0, nmd = o->nmodeDOF[m];
        for (int64_t t = 0; t < o->nham; t++) {
            oper_t v;
            memset(&v, 0, sizeof v);
            if (o->MultiSet[m]) { v.lhs = o->lhs[t]; v.rhs = o->rhs[t]; }
            else { v.lhs = 1; v.rhs = 1; }
            for (int64_t k = 0; k < nmd; k++)
                v.codes[k] = o->iOpDof[(o->iDOF[k + p->maxdof * m] - 1) + nd * t];
            int64_t found = -1;
            for (int64_t i = 0; i < nop; i++) {
                oper_t *u = &tmp[m * o->nham + i];
                int same = (u->lhs == v.lhs) && (u->rhs == v.rhs);
                for (int64_t k = 0; k < nmd && same; k++) same = (u->codes[k] == v.codes[k]);
                if (same) { found = i; break; }
            }
            if (found < 0) { tmp[m * o->nham + nop] = v; found = nop++; }
            o->iOper[m + nm * t] = found + 1;
        }
        o->nOper[m] = nop;
        if (nop > o->nopmax) o->nopmax = nop;
    }
    o->Oper = xcalloc((size_t)nm * o->nopmax, sizeof(oper_t));
    o->IsIdentity = xcalloc((size_t)nm * o->nopmax, sizeof(int64_t));
    for (int64_t m = 0; m < nm; m++)
        for (int64_t i = 0; i < o->nOper[m]; i++) {
            oper_t v = tmp[m * o->nham + i];
            o->Oper[m * o->nopmax + i] = v;
            int ident = (v.lhs == v.rhs);
            for (int64_t k = 0; k < o->nmodeDOF[m] && ident; k++) ident = (v.codes[k] == 0);
            o->IsIdentity[m + nm * i] = ident;
        }
    free(tmp);
    const size_t nn = (size_t)o->nspfmax * o->nspfmax;
    o->h1 = xcalloc(nn * o->nopmax * nm, sizeof(cplx));
    o->mf = xcalloc(nn * o->nopmax * nm, sizeof(cplx));
    o->rmf = xcalloc(nn * o->nopmax * nm, sizeof(cplx));
    o->rho = xcalloc(nn * nm * ns, sizeof(cplx));
    o->rho_ave = xcalloc(nn * nm, sizeof(cplx));
    /* identity h1mode for same-state identity operators (hamilton.f90:296-306) */
    for (int64_t m = 0; m < nm; m++)
        for (int64_t i = 0; i < o->nOper[m]; i++)
            if (o->IsIdentity[m + nm * i]) {
                int64_t s = o->Oper[m * o->nopmax + i].lhs - 1;
                for (int64_t j = 0; j < NSPF(o, m, s); j++) H1(o, m, i)[j + o->nspfmax * j] = 1.0;
            }
    o->A = xcalloc(o->nconf, sizeof(cplx));
    o->Astart = xcalloc(o->nconf, sizeof(cplx));
    o->psi = xcalloc(o->dimpsi, sizeof(cplx));
    o->psistart = xcalloc(o->dimpsi, sizeof(cplx));
    o->BK = xcalloc((size_t)o->nconf * NKRYLOV_MAX, sizeof(cplx));
    o->gwpa = xcalloc((size_t)p->maxdof * o->nspfmax * nm, sizeof(double));
    if (o->ngwp && p->gwp_a) memcpy(o->gwpa, p->gwp_a, sizeof(double) * p->maxdof * o->nspfmax * nm);
    o->S00 = xcalloc(nn * nm * ns * ns, sizeof(cplx));
    o->S00m1 = xcalloc(nn * nm * ns, sizeof(cplx));
    o->DD = xcalloc(nn * nm * ns, sizeof(cplx));
    o->DHD = xcalloc(nn * o->nopmax * nm, sizeof(cplx));
    /* diagonal of the intra-state overlaps stays 1 (storage.f90:55-63) */
    for (int64_t m = 0; m < nm; m++)
        if (o->gwp[m])
            for (int64_t s = 0; s < ns; s++)
                for (int64_t j = 0; j < NSPF(o, m, s); j++) S00P(o, m, s, s)[j + o->nspfmax * j] = 1.0;
    for (int64_t m = 0; m < nm; m++)
        if (o->nmodeDOF[m] > 8) snprintf(o->err, sizeof o->err, "more than 8 DOFs in a mode");
    return o;
}

void orc_destroy(orc *o) {
    if (!o) return;
    void *ptrs[] = {o->nmodeDOF, o->iDOF, o->MultiSet, o->nSPF, o->ngp, o->lhs, o->rhs,
                    o->iOpDof, o->grid, o->d1, o->d2, o->xd1, o->ae, o->nAConf,
                    o->AStateOffset, o->ipsiSPF, o->ngpm, o->nOper, o->Oper, o->iOper,
                    o->IsIdentity, o->A, o->psi, o->Astart, o->psistart, o->h1, o->rho,
                    o->rho_ave, o->mf, o->rmf, o->BK, o->gwp, o->npar, o->gwpa, o->S00,
                    o->S00m1, o->DD, o->DHD, o->gridoper};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
    free(o);
}

const char *orc_error(const orc *o) { return o->err; }
int64_t orc_nconf(const orc *o) { return o->nconf; }
int64_t orc_dimpsi(const orc *o) { return o->dimpsi; }
int64_t orc_nspfmax(const orc *o) { return o->nspfmax; }
int64_t orc_nopmax(const orc *o) { return o->nopmax; }

void orc_basis(const orc *o, int64_t *nOper, int64_t *iOper, int64_t *isIdentity) {
    memcpy(nOper, o->nOper, sizeof(int64_t) * o->nm);
    memcpy(iOper, o->iOper, sizeof(int64_t) * o->nm * o->nham);
    memcpy(isIdentity, o->IsIdentity, sizeof(int64_t) * o->nm * o->nopmax);
}

void orc_set_state(orc *o, const double *A, const double *psi) {
    memcpy(o->A, A, sizeof(cplx) * o->nconf);
    memcpy(o->psi, psi, sizeof(cplx) * o->dimpsi);
}
void orc_get_state(const orc *o, double *A, double *psi) {
    memcpy(A, o->A, sizeof(cplx) * o->nconf);
    memcpy(psi, o->psi, sizeof(cplx) * o->dimpsi);
}
void orc_set_start(orc *o, const double *A, const double *psi) {
    memcpy(o->Astart, A, sizeof(cplx) * o->nconf);
    memcpy(o->psistart, psi, sizeof(cplx) * o->dimpsi);
}

/* ---------------------------------------------------------------------- */
/* DVR primitive operators on an SPF block (MultTens / MultTensDiag,
 * hamilton.f90:577-629, and the case list of Calc_h1mode :452-470).
 * blk: ngpm x nS (column-major), the DOF k of mode m is acted on. */
static void apply_prim(orc *o, int64_t m, int64_t k, int64_t code, int64_t nS, cplx *blk) {
    const int64_t maxdof = o->p.maxdof, gmax = o->ngpmax;
    const int64_t dof = o->iDOF[k + maxdof * m] - 1;
    int64_t ll = 1, nl = 1;
    for (int64_t q = 0; q < k; q++) ll *= o->ngp[o->iDOF[q + maxdof * m] - 1];
    for (int64_t q = k + 1; q < o->nmodeDOF[m]; q++) nl *= o->ngp[o->iDOF[q + maxdof * m] - 1];
    const int64_t ng = o->ngp[dof];
    const int64_t nblk = nS * nl;
    if (code > 200) {   /* MultTensDiag with gridOper (hamilton.f90:467-468) */
        const double *f = o->gridoper + gmax * ((code - 201) + o->nopparam * dof);
        for (int64_t b = 0; b < nblk; b++)
            for (int64_t g = 0; g < ng; g++) {
                cplx *c = blk + (b * ng + g) * ll;
                for (int64_t l = 0; l < ll; l++) c[l] *= f[g];
            }
        return;
    }
    if (code <= -200) { /* MultTensProj: x_k -> phi phi^T (hamilton.f90:632-661) */
        const double *phi = o->gridoper + gmax * ((-code - 201) + o->nopparam * dof);
        cplx *tmp = xcalloc(ll, sizeof(cplx));
        for (int64_t b = 0; b < nblk; b++) {
            cplx *c = blk + b * ng * ll;
            for (int64_t l = 0; l < ll; l++) {
                cplx acc = 0;
                for (int64_t g = 0; g < ng; g++) acc += c[l + ll * g] * phi[g];
                tmp[l] = acc;
            }
            for (int64_t g = 0; g < ng; g++)
                for (int64_t l = 0; l < ll; l++) c[l + ll * g] = tmp[l] * phi[g];
        }
        free(tmp);
        return;
    }
    if (code >= 1 && code <= 20) {
        const double *x = o->grid + gmax * dof;
        for (int64_t b = 0; b < nblk; b++)
            for (int64_t g = 0; g < ng; g++) {
                double v = x[g];
                for (int64_t e = 1; e < code; e++) v *= x[g];
                cplx *c = blk + (b * ng + g) * ll;
                for (int64_t l = 0; l < ll; l++) c[l] *= v;
            }
        return;
    }
    const double *Mat;
    cplx fac = 1.0;
    switch (code) {
    case -1: Mat = o->d1 + gmax * gmax * dof; fac = I; break;
    case -2: Mat = o->d2 + gmax * gmax * dof; break;
    case -3: Mat = o->xd1 + gmax * gmax * dof; fac = I; break;
    default:
        snprintf(o->err, sizeof o->err, "primitive operator %lld not supported", (long long)code);
        return;
    }
    cplx *tmp = xcalloc(ng * ll, sizeof(cplx));
    for (int64_t b = 0; b < nblk; b++) {
        cplx *c = blk + b * ng * ll;
        /* zlacrm: C(l, g') = sum_g A(l, g) Mat(g, g') */
        for (int64_t g2 = 0; g2 < ng; g2++)
            for (int64_t l = 0; l < ll; l++) {
                cplx acc = 0;
                for (int64_t g = 0; g < ng; g++) acc += c[l + ll * g] * Mat[g + gmax * g2];
                tmp[l + ll * g2] = acc;
            }
        for (int64_t i = 0; i < ng * ll; i++) c[i] = fac * tmp[i];
    }
    free(tmp);
}

/* Calc_h1mode, DVR branch (hamilton.f90:433-485). */
static void calc_h1mode_mode(orc *o, const cplx *psi0, int64_t m) {
    const int64_t nmax = o->nspfmax, nG = o->ngpm[m];
    for (int64_t i = 0; i < o->nOper[m]; i++) {
        if (o->IsIdentity[m + o->nm * i]) continue;
        oper_t *op = &o->Oper[m * o->nopmax + i];
        int64_t s1 = op->lhs - 1, s2 = op->rhs - 1;
        int64_t nS1 = NSPF(o, m, s1), nS2 = NSPF(o, m, s2);
        const cplx *phi2 = psi0 + IPSI(o, m, s2) - 1;
        const cplx *phi1 = psi0 + IPSI(o, m, s1) - 1;
        cplx *sect = xcalloc(nG * nS2, sizeof(cplx));
        memcpy(sect, phi2, sizeof(cplx) * nG * nS2);
        for (int64_t k = 0; k < o->nmodeDOF[m]; k++)
            if (op->codes[k] != 0) apply_prim(o, m, k, op->codes[k], nS2, sect);
        cplx *h = H1(o, m, i);
        if (s1 == s2) {
            /* zher2k('U','C', 1/2): h = (phi^H sect + sect^H phi)/2, Hermitian */
            for (int64_t b = 0; b < nS2; b++)
                for (int64_t a = 0; a <= b; a++) {
                    cplx x = 0, y = 0;
                    for (int64_t g = 0; g < nG; g++) {
                        x += conj(phi2[g + nG * a]) * sect[g + nG * b];
                        y += conj(sect[g + nG * a]) * phi2[g + nG * b];
                    }
                    cplx v = 0.5 * x + 0.5 * y;
                    if (a == b) v = creal(v);
                    h[a + nmax * b] = v;
                    h[b + nmax * a] = conj(v);
                }
        } else {
            for (int64_t b = 0; b < nS2; b++)
                for (int64_t a = 0; a < nS1; a++) {
                    cplx x = 0;
                    for (int64_t g = 0; g < nG; g++) x += conj(phi1[g + nG * a]) * sect[g + nG * b];
                    h[a + nmax * b] = x;
                }
        }
        free(sect);
    }
}

/* Xi of Gaussian j (0-based) of mode m in the SPF block of set s. */
static const cplx *gxi(const orc *o, const cplx *psi, int64_t m, int64_t s, int64_t j) {
    return psi + IPSI(o, m, s) - 1 + o->nmodeDOF[m] * j;
}

/* Calc_S00 (gmctdh_objects.f90:323-381) for Gaussian mode m: intra-state
 * overlaps (diagonal kept at 1) and, for multi-set modes, the inter-state
 * blocks and their Hermitian conjugates. */
static void calc_s00(orc *o, const cplx *psi0, int64_t m) {
    const int64_t nmax = o->nspfmax, nmd = o->nmodeDOF[m];
    for (int64_t s = 0; s < NSETS(o, m); s++) {
        const int64_t nS = NSPF(o, m, s);
        cplx *S = S00P(o, m, s, s);
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t j1 = 0; j1 < j2; j1++) {
                S[j1 + nmax * j2] = gau_ovl(&GWPA(o, 0, j1, m), &GWPA(o, 0, j2, m), gxi(o, psi0, m, s, j1),
                                            gxi(o, psi0, m, s, j2), nmd, 1, 1);
                S[j2 + nmax * j1] = conj(S[j1 + nmax * j2]);
            }
    }
    if (!o->MultiSet[m]) return;
    for (int64_t s2 = 0; s2 < o->ns; s2++)
        for (int64_t s1 = 0; s1 < s2; s1++) {
            const int64_t n1 = NSPF(o, m, s1), n2 = NSPF(o, m, s2);
            cplx *S = S00P(o, m, s1, s2), *T = S00P(o, m, s2, s1);
            for (int64_t j2 = 0; j2 < n2; j2++)
                for (int64_t j1 = 0; j1 < n1; j1++) {
                    S[j1 + nmax * j2] = gau_ovl(&GWPA(o, 0, j1, m), &GWPA(o, 0, j2, m), gxi(o, psi0, m, s1, j1),
                                                gxi(o, psi0, m, s2, j2), nmd, 1, 1);
                    T[j2 + nmax * j1] = conj(S[j1 + nmax * j2]);
                }
        }
}

/* Calc_S00m1 (gmctdh_objects.f90:385-404): Tikhonov-regularised inverse
 * of the intra-state overlap, eps_ovl. */
static void calc_s00m1(orc *o, int64_t m) {
    const int64_t nmax = o->nspfmax;
    for (int64_t s = 0; s < NSETS(o, m); s++) {
        const int64_t nS = NSPF(o, m, s);
        cplx *a = xcalloc((size_t)nS * nS, sizeof(cplx)), *b = xcalloc((size_t)nS * nS, sizeof(cplx));
        const cplx *S = S00P(o, m, s, s);
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t j1 = 0; j1 < nS; j1++)
                a[j1 + nS * j2] = (j1 <= j2) ? S[j1 + nmax * j2] : conj(S[j2 + nmax * j1]);
        for (int64_t j = 0; j < nS; j++) a[j + nS * j] = creal(a[j + nS * j]);
        std_reg_inv(nS, a, o->p.eps_ovl, 1, b);
        cplx *R = S00M1(o, m, s);
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t j1 = 0; j1 < nS; j1++) R[j1 + nmax * j2] = b[j1 + nS * j2];
        free(a);
        free(b);
    }
}

/* Calc_h1mode, Gaussian branch (hamilton.f90:486-525): h = S00 prod_k
 * GauInt1D; the identity operator copies S00.  Same-state matrices are
 * Hermitian (the upper triangle is computed, the lower mirrored). */
static int calc_h1mode_gwp(orc *o, const cplx *psi0, int64_t m) {
    const int64_t nmax = o->nspfmax, nmd = o->nmodeDOF[m];
    int bad = 0;
    for (int64_t i = 0; i < o->nOper[m]; i++) {
        const oper_t *op = &o->Oper[m * o->nopmax + i];
        const int64_t s1 = op->lhs - 1, s2 = op->rhs - 1;
        const int64_t n1 = NSPF(o, m, s1), n2 = NSPF(o, m, s2);
        const cplx *S = S00P(o, m, s1, s2);
        cplx *h = H1(o, m, i);
        const int same = (s1 == s2);
        for (int64_t j2 = 0; j2 < n2; j2++)
            for (int64_t j1 = 0; j1 < (same ? j2 + 1 : n1); j1++) {
                cplx v = S[j1 + nmax * j2];
                if (!o->IsIdentity[m + o->nm * i]) {
                    const cplx *x1 = gxi(o, psi0, m, s1, j1), *x2 = gxi(o, psi0, m, s2, j2);
                    cplx pr = 1.0;
                    for (int64_t k = 0; k < nmd; k++)
                        pr *= gau_int(GWPA(o, k, j1, m), GWPA(o, k, j2, m), x1[k], x2[k], op->codes[k], &bad);
                    v = v * pr;
                }
                h[j1 + nmax * j2] = v;
                if (same) h[j2 + nmax * j1] = conj(v);
            }
    }
    if (bad) snprintf(o->err, sizeof o->err, "operator not supported on a Gaussian mode");
    return bad ? -1 : 0;
}

static void calc_h1mode_any(orc *o, const cplx *psi0, int64_t m) {
    if (o->gwp[m]) calc_h1mode_gwp(o, psi0, m);
    else calc_h1mode_mode(o, psi0, m);
}

/* Calc_S00 for every Gaussian mode, then Calc_h1mode for every mode (the
 * order of cmf2_base.f90:83-90). */
static void calc_operators(orc *o, const cplx *psi0) {
    for (int64_t m = 0; m < o->nm; m++) if (o->gwp[m]) calc_s00(o, psi0, m);
    for (int64_t m = 0; m < o->nm; m++) calc_h1mode_any(o, psi0, m);
}

/* Calc_DHD (hamilton.f90:535-574): D^H h D per Gaussian mode and operator,
 * 1 for the identity (D^H S D = 1), entries with |x| <= eps_machine set to 0. */
static void calc_dhd(orc *o, const cplx *DD) {
    const int64_t nmax = o->nspfmax;
    cplx *aux = xcalloc(NN(o), sizeof(cplx));
    for (int64_t m = 0; m < o->nm; m++) {
        if (!o->gwp[m]) continue;
        for (int64_t i = 0; i < o->nOper[m]; i++) {
            const oper_t *op = &o->Oper[m * o->nopmax + i];
            const int64_t s1 = op->lhs - 1, s2 = op->rhs - 1;
            const int64_t n1 = NSPF(o, m, s1), n2 = NSPF(o, m, s2);
            cplx *R = DHDP(o, m, i);
            memset(R, 0, sizeof(cplx) * NN(o));
            if (o->IsIdentity[m + o->nm * i]) {
                for (int64_t j = 0; j < (n1 < n2 ? n1 : n2); j++) R[j + nmax * j] = 1.0;
                continue;
            }
            const cplx *h = H1(o, m, i), *D1 = DDP(o, DD, m, s1), *D2 = DDP(o, DD, m, s2);
            if (s1 == s2) {
                /* aux = h D; R = (D^H aux + aux^H D) / 2, upper triangle, real diagonal */
                for (int64_t b = 0; b < n2; b++)
                    for (int64_t a = 0; a < n1; a++) {
                        cplx acc = 0;
                        for (int64_t k = 0; k < n1; k++) acc += h[a + nmax * k] * D2[k + nmax * b];
                        aux[a + nmax * b] = acc;
                    }
                for (int64_t b = 0; b < n1; b++)
                    for (int64_t a = 0; a <= b; a++) {
                        cplx x = 0, y = 0;
                        for (int64_t k = 0; k < n1; k++) {
                            x += conj(D1[k + nmax * a]) * aux[k + nmax * b];
                            y += conj(aux[k + nmax * a]) * D1[k + nmax * b];
                        }
                        cplx v = 0.5 * x + 0.5 * y;
                        if (a == b) v = creal(v);
                        R[a + nmax * b] = v;
                    }
            } else {
                /* aux = D1^H h; R = aux D2 */
                for (int64_t b = 0; b < n2; b++)
                    for (int64_t a = 0; a < n1; a++) {
                        cplx acc = 0;
                        for (int64_t k = 0; k < n1; k++) acc += conj(D1[k + nmax * a]) * h[k + nmax * b];
                        aux[a + nmax * b] = acc;
                    }
                for (int64_t b = 0; b < n2; b++)
                    for (int64_t a = 0; a < n1; a++) {
                        cplx acc = 0;
                        for (int64_t k = 0; k < n2; k++) acc += aux[a + nmax * k] * D2[k + nmax * b];
                        R[a + nmax * b] = acc;
                    }
            }
            const double eps = 2.220446049250313e-16;
            for (int64_t b = 0; b < n2; b++)
                for (int64_t a = 0; a < (s1 == s2 ? b + 1 : n1); a++)
                    if (!(cabs(R[a + nmax * b]) > eps)) R[a + nmax * b] = 0;
            if (s1 == s2)
                for (int64_t b = 0; b < n1; b++)
                    for (int64_t a = b + 1; a < n1; a++) R[a + nmax * b] = conj(R[b + nmax * a]);
        }
    }
    free(aux);
}

/* InverseSquareRootU (maths.f90:621-644): D = U^-1 with S = U^H U
 * (Cholesky), so D^H S D = 1 and D upper triangular. */
static int inverse_sqrt_u(int64_t n, const cplx *S, int64_t lds, cplx *D, int64_t ldd) {
    cplx *U = xcalloc((size_t)n * n, sizeof(cplx));
    for (int64_t j = 0; j < n; j++) {
        double d = creal(S[j + lds * j]);
        for (int64_t k = 0; k < j; k++) d -= creal(conj(U[k + n * j]) * U[k + n * j]);
        if (!(d > 0.0)) { free(U); return -1; }
        const double ujj = sqrt(d);
        U[j + n * j] = ujj;
        for (int64_t i = j + 1; i < n; i++) {
            cplx v = S[j + lds * i];
            for (int64_t k = 0; k < j; k++) v -= conj(U[k + n * j]) * U[k + n * i];
            U[j + n * i] = v / ujj;
        }
    }
    /* upper-triangular inverse, column by column */
    for (int64_t c = 0; c < n; c++) {
        for (int64_t r = 0; r < n; r++) D[r + ldd * c] = 0;
        D[c + ldd * c] = 1.0 / U[c + n * c];
        for (int64_t r = c - 1; r >= 0; r--) {
            cplx acc = 0;
            for (int64_t k = r + 1; k <= c; k++) acc += U[r + n * k] * D[k + ldd * c];
            D[r + ldd * c] = -acc / U[r + n * r];
        }
    }
    free(U);
    return 0;
}

/* D0 = InverseSquareRootU(S00(psi0)) for every Gaussian mode and set
 * (cmf2_base.f90:72-80). */
static int init_dd(orc *o, const cplx *psi0, cplx *DD) {
    int rc = 0;
    for (int64_t m = 0; m < o->nm; m++) {
        if (!o->gwp[m]) continue;
        calc_s00(o, psi0, m);
        for (int64_t s = 0; s < NSETS(o, m); s++) {
            cplx *S = S00P(o, m, s, s);
            if (inverse_sqrt_u(NSPF(o, m, s), S, o->nspfmax, DDP(o, DD, m, s), o->nspfmax) != 0) {
                snprintf(o->err, sizeof o->err, "Gaussian overlap not positive definite (mode %lld)",
                         (long long)m + 1);
                rc = -1;
            }
        }
    }
    return rc;
}

/* Mode products over the Gaussian modes only: Y_s = (x_m M_(m,set(s))) X_s,
 * M from mat(o, m, set, scratch). */
typedef const cplx *(*gmat_fn)(orc *o, int64_t m, int64_t set, const cplx *DD, cplx *scratch);

static void gwp_chain(orc *o, const cplx *X, cplx *Y, gmat_fn mat, const cplx *DD) {
    int64_t big = 0;
    for (int64_t s = 0; s < o->ns; s++) if (o->nAConf[s] > big) big = o->nAConf[s];
    cplx *p1 = xcalloc(big, sizeof(cplx)), *p2 = xcalloc(big, sizeof(cplx)), *sc = xcalloc(NN(o), sizeof(cplx));
    int64_t sz[64];
    for (int64_t s = 0; s < o->ns; s++) {
        for (int64_t m = 0; m < o->nm; m++) sz[m] = NSPF(o, m, s);
        memcpy(p1, X + o->AStateOffset[s], sizeof(cplx) * o->nAConf[s]);
        for (int64_t m = 0; m < o->nm; m++) {
            if (!o->gwp[m]) continue;
            const cplx *M = mat(o, m, o->MultiSet[m] ? s : 0, DD, sc);
            mode_product(o->nm, sz, m, M, o->nspfmax, sz[m], p1, p2);
            cplx *sw = p1; p1 = p2; p2 = sw;
        }
        memcpy(Y + o->AStateOffset[s], p1, sizeof(cplx) * o->nAConf[s]);
    }
    free(p1); free(p2); free(sc);
}

/* (S D)^H = D^H S for A_to_B (cmf2_base.f90:264-305: B = D^H S A). */
static const cplx *mat_dhs(orc *o, int64_t m, int64_t set, const cplx *DD, cplx *sc) {
    const int64_t nmax = o->nspfmax, n = NSPF(o, m, set);
    const cplx *S = S00P(o, m, set, set), *D = DDP(o, DD, m, set);
    for (int64_t b = 0; b < n; b++)
        for (int64_t a = 0; a < n; a++) {
            cplx acc = 0;       /* (S D)[b, a] */
            for (int64_t k = 0; k < n; k++) acc += S[b + nmax * k] * D[k + nmax * a];
            sc[a + nmax * b] = conj(acc);
        }
    return sc;
}

static const cplx *mat_d(orc *o, int64_t m, int64_t set, const cplx *DD, cplx *sc) {
    (void)sc;
    return DDP(o, DD, m, set);
}

static void a_to_b(orc *o, const cplx *A, cplx *B, const cplx *DD) {
    if (!o->ngwp) { if (A != B) memcpy(B, A, sizeof(cplx) * o->nconf); return; }
    gwp_chain(o, A, B, mat_dhs, DD);
}

/* B_to_A (cmf2_base.f90:309-345): A = D B. */
static void b_to_a(orc *o, const cplx *B, cplx *A, const cplx *DD) {
    if (!o->ngwp) { if (A != B) memcpy(A, B, sizeof(cplx) * o->nconf); return; }
    gwp_chain(o, B, A, mat_d, DD);
}

void orc_calc_h1mode(orc *o, double *h1_out) {
    calc_operators(o, o->psi);
    if (h1_out) memcpy(h1_out, o->h1, sizeof(cplx) * o->nspfmax * o->nspfmax * o->nopmax * o->nm);
}

/* Calc_rho1mode (gmctdh_objects.f90:408-477).  With Gaussian modes the
 * overlaps S00 of every other Gaussian mode are applied first (:439-447)
 * and rho = (X^H A + A^H X) / 2 on the hole matrices (zher2k, :452-457);
 * S00 must hold the overlaps of the current SPFs. */
static void calc_rho(orc *o, const cplx *A0) {
    const int64_t nm = o->nm, nmax = o->nspfmax;
    int64_t big = 0;
    for (int64_t s = 0; s < o->ns; s++) if (o->nAConf[s] > big) big = o->nAConf[s];
    cplx *X = o->ngwp ? xcalloc(big, sizeof(cplx)) : NULL, *T = o->ngwp ? xcalloc(big, sizeof(cplx)) : NULL;
    for (int64_t m = 0; m < nm; m++) {
        for (int64_t s = 0; s < o->ns; s++) {
            cplx *r = RHO(o, m, s);
            for (int64_t i = 0; i < nmax * nmax; i++) r[i] = 0;
            int64_t ll = 1, nl = 1, nS = NSPF(o, m, s);
            for (int64_t q = 0; q < m; q++) ll *= NSPF(o, q, s);
            for (int64_t q = m + 1; q < nm; q++) nl *= NSPF(o, q, s);
            const cplx *As = A0 + o->AStateOffset[s];
            int other = 0;
            for (int64_t q = 0; q < nm; q++) if (o->gwp[q] && q != m) other = 1;
            if (!other) {
                hole_gram(ll, nl, nS, nS, As, As, 1.0, r, nmax);
            } else {
                int64_t sz[64];
                for (int64_t q = 0; q < nm; q++) sz[q] = NSPF(o, q, s);
                memcpy(X, As, sizeof(cplx) * o->nAConf[s]);
                for (int64_t q = 0; q < nm; q++) {
                    if (!o->gwp[q] || q == m) continue;
                    const int64_t jel = o->MultiSet[q] ? s : 0;
                    mode_product(nm, sz, q, S00P(o, q, jel, jel), nmax, sz[q], X, T);
                    cplx *sw = X; X = T; T = sw;
                }
                hole_gram(ll, nl, nS, nS, X, As, 0.5, r, nmax);
                hole_gram(ll, nl, nS, nS, As, X, 0.5, r, nmax);
            }
            for (int64_t j = 0; j < nS; j++) r[j + nmax * j] = creal(r[j + nmax * j]);
            for (int64_t b = 0; b < nS; b++)
                for (int64_t a = b + 1; a < nS; a++) r[a + nmax * b] = conj(r[b + nmax * a]);
        }
        if (!o->MultiSet[m]) {
            cplx *ra = RHOAVE(o, m);
            for (int64_t i = 0; i < nmax * nmax; i++) {
                cplx acc = 0;
                for (int64_t s = 0; s < o->ns; s++) acc += RHO(o, m, s)[i];
                ra[i] = acc;
            }
        }
    }
    free(X);
    free(T);
}

void orc_calc_rho(orc *o, double *rho_out) {
    calc_rho(o, o->A);
    if (rho_out) memcpy(rho_out, o->rho, sizeof(cplx) * o->nspfmax * o->nspfmax * o->nm * o->ns);
}

/* Largest tensor a chain over mixed states can produce: prod_m max_s n(m,s). */
static int64_t mixed_size(const orc *o) {
    int64_t big = 1;
    for (int64_t m = 0; m < o->nm; m++) {
        int64_t mx = 0;
        for (int64_t s = 0; s < o->ns; s++) if (NSPF(o, m, s) > mx) mx = NSPF(o, m, s);
        big *= mx;
    }
    return big;
}

/* Calc_MeanField, Hamiltonian part (gmctdh_objects.f90:518-605), then
 * Calc_rhom1MF (:1147-1194). */
static int calc_meanfield(orc *o, const cplx *A0) {
    const int64_t nm = o->nm, nmax = o->nspfmax, nn = nmax * nmax;
    memset(o->mf, 0, sizeof(cplx) * nn * o->nopmax * nm);
    o->energyMF = 0.0;
    /* mode products can change the size of a state block (n_s1 vs n_s2) */
    const int64_t big = mixed_size(o);
    cplx *AL = xcalloc((size_t)big * nm, sizeof(cplx));
    cplx *AR = xcalloc((size_t)big * (nm + 1), sizeof(cplx));
    int64_t sz[64];
    for (int64_t t = 0; t < o->nham; t++) {
        const int64_t s1 = o->lhs[t] - 1, s2 = o->rhs[t] - 1;
        const cplx ae = o->ae[t];
        /* right chain: R_0 = A_s2, modes f..1 */
        for (int64_t m = 0; m < nm; m++) sz[m] = NSPF(o, m, s2);
        int64_t cur = o->nAConf[s2];
        memcpy(AR, A0 + o->AStateOffset[s2], sizeof(cplx) * cur);
        for (int64_t im = 1; im <= nm; im++) {
            int64_t jm = nm - im;
            int64_t op = o->iOper[jm + nm * t] - 1;
            cplx *src = AR + (size_t)big * (im - 1), *dst = AR + (size_t)big * im;
            if (o->IsIdentity[jm + nm * op] && !o->gwp[jm]) { memcpy(dst, src, sizeof(cplx) * cur); continue; }
            int64_t nS1 = NSPF(o, jm, s1);
            mode_product(nm, sz, jm, H1(o, jm, op), nmax, nS1, src, dst);
            cur = cur / sz[jm] * nS1;
            sz[jm] = nS1;
        }
        /* left chain: L_0 = A_s1, modes 1..f-1, with h^H */
        for (int64_t m = 0; m < nm; m++) sz[m] = NSPF(o, m, s1);
        cur = o->nAConf[s1];
        memcpy(AL, A0 + o->AStateOffset[s1], sizeof(cplx) * cur);
        cplx *hh = xcalloc(nn, sizeof(cplx));
        for (int64_t im = 1; im <= nm - 1; im++) {
            int64_t m = im - 1;
            int64_t op = o->iOper[m + nm * t] - 1;
            cplx *src = AL + (size_t)big * (im - 1), *dst = AL + (size_t)big * im;
            if (o->IsIdentity[m + nm * op] && !o->gwp[m]) { memcpy(dst, src, sizeof(cplx) * cur); continue; }
            int64_t nS1 = NSPF(o, m, s1), nS2 = NSPF(o, m, s2);
            for (int64_t a = 0; a < nS2; a++)
                for (int64_t b = 0; b < nS1; b++) hh[a + nmax * b] = conj(H1(o, m, op)[b + nmax * a]);
            mode_product(nm, sz, m, hh, nmax, nS2, src, dst);
            cur = cur / sz[m] * nS2;
            sz[m] = nS2;
        }
        free(hh);
        /* energy: Re(ae * <A_s1 | R_f>) */
        {
            const cplx *a1 = A0 + o->AStateOffset[s1];
            const cplx *rf = AR + (size_t)big * nm;
            cplx acc = 0;
            for (int64_t i = 0; i < o->nAConf[s1]; i++) acc += conj(a1[i]) * rf[i];
            o->energyMF += creal(ae * acc);
        }
        /* mean fields */
        for (int64_t im = 1; im <= nm; im++) {
            int64_t m = im - 1;
            int64_t op = o->iOper[m + nm * t] - 1;
            if (o->IsIdentity[m + nm * op]) continue;
            int64_t jm = nm - im;
            int64_t nS1 = NSPF(o, m, s1), nS2 = NSPF(o, m, s2);
            int64_t ll = 1, nl = 1;
            for (int64_t q = 0; q < m; q++) ll *= NSPF(o, q, s2);
            for (int64_t q = m + 1; q < nm; q++) nl *= NSPF(o, q, s1);
            hole_gram(ll, nl, nS1, nS2, AL + (size_t)big * (im - 1), AR + (size_t)big * jm, ae,
                      MF(o, m, op), nmax);
        }
    }
    free(AL);
    free(AR);
    /* Calc_rhom1MF (grid modes only, gmctdh_objects.f90:1159) */
    cplx *rinv = xcalloc(nn, sizeof(cplx)), *rsq = xcalloc(nn, sizeof(cplx));
    for (int64_t m = 0; m < nm; m++) {
        if (o->gwp[m]) continue;
        int64_t nsets = o->MultiSet[m] ? o->ns : 1;
        for (int64_t s = 0; s < nsets; s++) {
            int64_t nS = NSPF(o, m, s);
            const cplx *r = o->MultiSet[m] ? RHO(o, m, s) : RHOAVE(o, m);
            for (int64_t b = 0; b < nS; b++)
                for (int64_t a = 0; a < nS; a++) rsq[a + nS * b] = r[a + nmax * b];
            std_reg_inv(nS, rsq, o->p.eps_rho, 2, rinv);
            for (int64_t i = 0; i < o->nOper[m]; i++) {
                oper_t *op = &o->Oper[m * o->nopmax + i];
                if (o->IsIdentity[m + nm * i]) continue;
                if (o->MultiSet[m] && op->lhs - 1 != s) continue;
                int64_t nS2 = NSPF(o, m, op->rhs - 1);
                cplx *dst = RMF(o, m, i);
                const cplx *src = MF(o, m, i);
                for (int64_t b = 0; b < nS2; b++)
                    for (int64_t a = 0; a < nS; a++) {
                        cplx acc = 0;
                        for (int64_t k = 0; k < nS; k++) acc += rinv[a + nS * k] * src[k + nmax * b];
                        dst[a + nmax * b] = acc;
                    }
            }
        }
    }
    free(rinv);
    free(rsq);
    return 0;
}

int orc_calc_meanfield(orc *o, double *mf_out, double *rmf_out, double *energy_out) {
    calc_meanfield(o, o->A);
    size_t n = (size_t)o->nspfmax * o->nspfmax * o->nopmax * o->nm;
    if (mf_out) memcpy(mf_out, o->mf, sizeof(cplx) * n);
    if (rmf_out) memcpy(rmf_out, o->rmf, sizeof(cplx) * n);
    if (energy_out) *energy_out = o->energyMF;
    return 0;
}

/* Calc_DHD_B (cmf2_derivatives.f90:1-80): h1mode on grid modes, D^H h D on
 * Gaussian modes (DHD must be current), identities skipped on both. */
static void apply_h(orc *o, const cplx *B, cplx *HB) {
    const int64_t nm = o->nm, nmax = o->nspfmax;
    memset(HB, 0, sizeof(cplx) * o->nconf);
    const int64_t big = mixed_size(o);
    cplx *b1 = xcalloc(big, sizeof(cplx)), *b2 = xcalloc(big, sizeof(cplx));
    int64_t sz[64];
    for (int64_t t = 0; t < o->nham; t++) {
        const int64_t s1 = o->lhs[t] - 1, s2 = o->rhs[t] - 1;
        for (int64_t m = 0; m < nm; m++) sz[m] = NSPF(o, m, s2);
        memcpy(b2, B + o->AStateOffset[s2], sizeof(cplx) * o->nAConf[s2]);
        for (int64_t m = 0; m < nm; m++) {
            int64_t op = o->iOper[m + nm * t] - 1;
            if (o->IsIdentity[m + nm * op]) continue;
            int64_t nS1 = NSPF(o, m, s1);
            mode_product(nm, sz, m, o->gwp[m] ? DHDP(o, m, op) : H1(o, m, op), nmax, nS1, b2, b1);
            cplx *sw = b1; b1 = b2; b2 = sw;
            sz[m] = nS1;
        }
        cplx ae = o->ae[t];
        cplx *out = HB + o->AStateOffset[s1];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < o->nAConf[s1]; i++) out[i] += ae * b2[i];
    }
    free(b1);
    free(b2);
}

void orc_apply_h(orc *o, const double *B, double *HB) { apply_h(o, (const cplx *)B, (cplx *)HB); }

/* H A in the Gaussian basis: every term's h1mode on every Gaussian mode (the
 * identity there is S00) and on the non-identity grid modes -- the vector
 * conj(Adot_tmp) / der_factor of Calc_MeanField (gmctdh_objects.f90:550-578),
 * used by the VMF derivatives (derivatives.f90:160-176). */
static void apply_hg(orc *o, const cplx *B, cplx *HB) {
    const int64_t nm = o->nm, nmax = o->nspfmax;
    memset(HB, 0, sizeof(cplx) * o->nconf);
    const int64_t big = mixed_size(o);
    cplx *b1 = xcalloc(big, sizeof(cplx)), *b2 = xcalloc(big, sizeof(cplx));
    int64_t sz[64];
    for (int64_t t = 0; t < o->nham; t++) {
        const int64_t s1 = o->lhs[t] - 1, s2 = o->rhs[t] - 1;
        for (int64_t m = 0; m < nm; m++) sz[m] = NSPF(o, m, s2);
        memcpy(b2, B + o->AStateOffset[s2], sizeof(cplx) * o->nAConf[s2]);
        for (int64_t m = 0; m < nm; m++) {
            int64_t op = o->iOper[m + nm * t] - 1;
            if (o->IsIdentity[m + nm * op] && !o->gwp[m]) continue;
            int64_t nS1 = NSPF(o, m, s1);
            mode_product(nm, sz, m, H1(o, m, op), nmax, nS1, b2, b1);
            cplx *sw = b1; b1 = b2; b2 = sw;
            sz[m] = nS1;
        }
        cplx ae = o->ae[t];
        cplx *out = HB + o->AStateOffset[s1];
        for (int64_t i = 0; i < o->nAConf[s1]; i++) out[i] += ae * b2[i];
    }
    free(b1);
    free(b2);
}

/* Calc_Der_SPF, DVR branch (cmf2_derivatives.f90:118-173); mode m is 0-based. */
static int der_spf(orc *o, int64_t m, const cplx *psi0, cplx *dpsi) {
    const int64_t nmax = o->nspfmax, nG = o->ngpm[m], nm = o->nm;
    int64_t nsets = o->MultiSet[m] ? o->ns : 1;
    for (int64_t s = 0; s < nsets; s++)
        memset(dpsi + IPSI(o, m, s) - 1, 0, sizeof(cplx) * nG * NSPF(o, m, s));
    cplx *sect = xcalloc(nG * nmax, sizeof(cplx));
    for (int64_t i = 0; i < o->nOper[m]; i++) {
        if (o->IsIdentity[m + nm * i]) continue;
        oper_t *op = &o->Oper[m * o->nopmax + i];
        int64_t s1 = op->lhs - 1, s2 = op->rhs - 1;
        int64_t nS1 = NSPF(o, m, s1), nS2 = NSPF(o, m, s2);
        memcpy(sect, psi0 + IPSI(o, m, s2) - 1, sizeof(cplx) * nG * nS2);
        for (int64_t k = 0; k < o->nmodeDOF[m]; k++)
            if (op->codes[k] != 0) apply_prim(o, m, k, op->codes[k], nS2, sect);
        cplx *d = dpsi + IPSI(o, m, s1) - 1;
        const cplx *r = RMF(o, m, i);
        /* zgemm('N','T'): d += der_factor * sect * rmf^T */
        for (int64_t j1 = 0; j1 < nS1; j1++)
            for (int64_t g = 0; g < nG; g++) {
                cplx acc = 0;
                for (int64_t j2 = 0; j2 < nS2; j2++) acc += sect[g + nG * j2] * r[j1 + nmax * j2];
                d[g + nG * j1] += o->der_factor * acc;
            }
    }
    free(sect);
    /* projector (1 - phi (phi^H phi)^-1 phi^H) */
    cplx *ovl = xcalloc(nmax * nmax, sizeof(cplx)), *od = xcalloc(nmax * nmax, sizeof(cplx));
    int rc = 0;
    for (int64_t s = 0; s < nsets; s++) {
        int64_t nS = NSPF(o, m, s);
        const cplx *phi = psi0 + IPSI(o, m, s) - 1;
        cplx *d = dpsi + IPSI(o, m, s) - 1;
        for (int64_t b = 0; b < nS; b++)
            for (int64_t a = 0; a < nS; a++) {
                cplx x = 0, y = 0;
                for (int64_t g = 0; g < nG; g++) {
                    x += conj(phi[g + nG * a]) * phi[g + nG * b];
                    y += conj(phi[g + nG * a]) * d[g + nG * b];
                }
                ovl[a + nS * b] = x;
                od[a + nS * b] = y;
            }
        if (chol_solve(nS, ovl, nS, od, nS, nS) != 0) {
            snprintf(o->err, sizeof o->err, "SPF overlap not positive definite (mode %lld)", (long long)m + 1);
            rc = -1;
        }
        for (int64_t b = 0; b < nS; b++)
            for (int64_t g = 0; g < nG; g++) {
                cplx acc = 0;
                for (int64_t k = 0; k < nS; k++) acc += phi[g + nG * k] * od[k + nS * b];
                d[g + nG * b] -= acc;
            }
    }
    free(ovl);
    free(od);
    return rc;
}

void orc_der_spf(orc *o, int64_t im, const double *psi_in, double *dpsi_out) {
    der_spf(o, im - 1, (const cplx *)psi_in, (cplx *)dpsi_out);
}

/* StdRegSol 'tikh' (maths.f90:87-162): x = U diag(w / (w^2 + eps^2)) U^H y
 * for the Hermitian matrix whose upper triangle is in A (ld lda). */
static void std_reg_sol_tikh(int64_t n, const cplx *A, int64_t lda, const cplx *y, double eps, cplx *x) {
    cplx *F = xcalloc((size_t)n * n, sizeof(cplx)), *U = xcalloc((size_t)n * n, sizeof(cplx));
    cplx *uy = xcalloc(n, sizeof(cplx));
    double *w = xcalloc(n, sizeof(double));
    for (int64_t j2 = 0; j2 < n; j2++)
        for (int64_t j1 = 0; j1 < n; j1++)
            F[j1 + n * j2] = (j1 < j2) ? A[j1 + lda * j2] : (j1 == j2 ? creal(A[j1 + lda * j1]) : conj(A[j2 + lda * j1]));
    herm_eig(n, F, w, U);
    for (int64_t k = 0; k < n; k++) {
        cplx acc = 0;
        for (int64_t i = 0; i < n; i++) acc += conj(U[i + n * k]) * y[i];
        uy[k] = acc * (w[k] / (w[k] * w[k] + eps * eps));
    }
    for (int64_t i = 0; i < n; i++) {
        cplx acc = 0;
        for (int64_t k = 0; k < n; k++) acc += U[i + n * k] * uy[k];
        x[i] = acc;
    }
    free(F); free(U); free(uy); free(w);
}

/* Calc_Der_SPF, Gaussian branch (cmf2_derivatives.f90:175-232) with
 * Calc_CY (gmctdh_objects.f90:1198-1333, the form without DD0):
 * S00, S00^-1 and h1mode of mode m are recomputed from psi0 (as the
 * reference does, leaving them in the shared arrays); rho and the mean
 * fields are those of the last calc_meanfield.  dpsi receives der_factor
 * times the regularised solution of C dxi = Y; dDD = -tau D. */
static int der_spf_gwp(orc *o, int64_t m, const cplx *psi0, const cplx *DD0, cplx *dpsi, cplx *dDD,
                       cplx *tau_out) {
    const int64_t nmax = o->nspfmax, nmd = o->nmodeDOF[m], nm = o->nm;
    int bad = 0;
    calc_s00(o, psi0, m);
    calc_s00m1(o, m);
    if (calc_h1mode_gwp(o, psi0, m) != 0) return -1;
    int64_t ncy = 0;
    for (int64_t s = 0; s < o->ns; s++) if (NSPF(o, m, s) * nmd > ncy) ncy = NSPF(o, m, s) * nmd;
    cplx *gx = xcalloc((size_t)nmd * nmax * nmax, sizeof(cplx));      /* gIntx(k, j1, j2) */
    cplx *Sa0 = xcalloc((size_t)ncy * nmax, sizeof(cplx));
    cplx *SaSm = xcalloc((size_t)ncy * nmax, sizeof(cplx));            /* Sa0 S00^-1 */
    cplx *Sab = xcalloc((size_t)ncy * ncy, sizeof(cplx));
    cplx *C = xcalloc((size_t)ncy * ncy, sizeof(cplx));
    cplx *Y = xcalloc(ncy, sizeof(cplx));
    cplx *Ha0 = xcalloc((size_t)ncy * nmax, sizeof(cplx));
    cplx *gI = xcalloc(nmd, sizeof(cplx));
    cplx *tau = xcalloc(NN(o), sizeof(cplx));
#define GX(k, a, b) gx[(k) + nmd * ((a) + nmax * (b))]
    for (int64_t s = 0; s < NSETS(o, m); s++) {
        const int64_t nS = NSPF(o, m, s), nG = nS * nmd;
        const cplx *S = S00P(o, m, s, s), *Sm = S00M1(o, m, s);
        memset(Y, 0, sizeof(cplx) * ncy);
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t j1 = 0; j1 <= j2; j1++) {
                const cplx *x1 = gxi(o, psi0, m, s, j1), *x2 = gxi(o, psi0, m, s, j2);
                for (int64_t k = 0; k < nmd; k++)
                    GX(k, j1, j2) = gau_int(GWPA(o, k, j1, m), GWPA(o, k, j2, m), x1[k], x2[k], 1, &bad);
            }
        /* Sa0((j1 nmd + k), j2) = <x_k g_j1 | g_j2> */
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t j1 = 0; j1 < nS; j1++)
                for (int64_t k = 0; k < nmd; k++) {
                    if (j1 <= j2) Sa0[(j1 * nmd + k) + ncy * j2] = S[j1 + nmax * j2] * GX(k, j1, j2);
                }
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t j1 = j2 + 1; j1 < nS; j1++)
                for (int64_t k = 0; k < nmd; k++)
                    Sa0[(j1 * nmd + k) + ncy * j2] = conj(Sa0[(j2 * nmd + k) + ncy * j1]);
        /* Sa0 S00^-1 (zhemm 'R','U') */
        for (int64_t c = 0; c < nS; c++)
            for (int64_t r = 0; r < nG; r++) {
                cplx acc = 0;
                for (int64_t q = 0; q < nS; q++) acc += Sa0[r + ncy * q] * Sm[q + nmax * c];
                SaSm[r + ncy * c] = acc;
            }
        /* Sab for the upper block triangle (j1 <= j2) */
        for (int64_t i = 0; i < ncy * ncy; i++) Sab[i] = 0;
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t k2 = 0; k2 < nmd; k2++)
                for (int64_t j1 = 0; j1 <= j2; j1++)
                    for (int64_t k1 = 0; k1 < nmd; k1++) {
                        cplx v;
                        if (k1 == k2)
                            v = gau_int(GWPA(o, k1, j1, m), GWPA(o, k2, j2, m), gxi(o, psi0, m, s, j1)[k1],
                                        gxi(o, psi0, m, s, j2)[k2], 2, &bad);
                        else
                            v = GX(k1, j1, j2) * GX(k2, j1, j2);
                        Sab[(j1 * nmd + k1) + ncy * (j2 * nmd + k2)] = S[j1 + nmax * j2] * v;
                    }
        /* Sab -= (SaSm Sa0^H + Sa0 SaSm^H) / 2 on the upper triangle (zher2k) */
        for (int64_t c = 0; c < nG; c++)
            for (int64_t r = 0; r <= c; r++) {
                cplx acc = 0;
                for (int64_t q = 0; q < nS; q++)
                    acc += SaSm[r + ncy * q] * conj(Sa0[c + ncy * q]) + Sa0[r + ncy * q] * conj(SaSm[c + ncy * q]);
                cplx v = Sab[r + ncy * c] - 0.5 * acc;
                if (r == c) v = creal(v);
                Sab[r + ncy * c] = v;
            }
        /* C = rho o Sab (upper), rho of the set (multi-set) or rho_ave */
        const cplx *rh = o->MultiSet[m] ? RHO(o, m, s) : RHOAVE(o, m);
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t k2 = 0; k2 < nmd; k2++)
                for (int64_t j1 = 0; j1 <= j2; j1++)
                    for (int64_t k1 = 0; k1 < nmd; k1++) {
                        const int64_t r = j1 * nmd + k1, c = j2 * nmd + k2;
                        C[r + ncy * c] = rh[j1 + nmax * j2] * Sab[r + ncy * c];
                    }
        /* Y = sum_op MF(j1, j2) (H_a0 - Sa0 S^-1 h)(j1 k, j2) */
        for (int64_t i = 0; i < o->nOper[m]; i++) {
            if (o->IsIdentity[m + nm * i]) continue;
            const oper_t *op = &o->Oper[m * o->nopmax + i];
            if (op->lhs - 1 != s && o->MultiSet[m]) continue;
            const int64_t s2 = op->rhs - 1, n2 = NSPF(o, m, s2);
            const cplx *S12 = S00P(o, m, s, s2);
            for (int64_t j2 = 0; j2 < n2; j2++)
                for (int64_t j1 = 0; j1 < nS; j1++) {
                    const cplx *x1 = gxi(o, psi0, m, s, j1), *x2 = gxi(o, psi0, m, s2, j2);
                    for (int64_t k = 0; k < nmd; k++)
                        gI[k] = gau_int(GWPA(o, k, j1, m), GWPA(o, k, j2, m), x1[k], x2[k], op->codes[k], &bad);
                    for (int64_t k = 0; k < nmd; k++) {
                        cplx pr = S12[j1 + nmax * j2];
                        for (int64_t q = 0; q < nmd; q++)
                            if (q != k) pr *= gI[q];
                        Ha0[(j1 * nmd + k) + ncy * j2] =
                            pr * gau10_int(GWPA(o, k, j1, m), GWPA(o, k, j2, m), x1[k], x2[k], op->codes[k], &bad);
                    }
                }
            const cplx *h = H1(o, m, i);
            for (int64_t j2 = 0; j2 < n2; j2++)
                for (int64_t r = 0; r < nG; r++) {
                    cplx acc = 0;
                    for (int64_t q = 0; q < nS; q++) acc += SaSm[r + ncy * q] * h[q + nmax * j2];
                    Ha0[r + ncy * j2] -= acc;
                }
            const cplx *mf = MF(o, m, i);
            for (int64_t j2 = 0; j2 < n2; j2++)
                for (int64_t j1 = 0; j1 < nS; j1++)
                    for (int64_t k = 0; k < nmd; k++)
                        Y[j1 * nmd + k] += mf[j1 + nmax * j2] * Ha0[(j1 * nmd + k) + ncy * j2];
        }
        /* dxi = der_factor * StdRegSol(C, Y, eps_C, 'tikh') */
        cplx *dx = dpsi + IPSI(o, m, s) - 1;
        std_reg_sol_tikh(nG, C, ncy, Y, o->p.eps_C, dx);
        for (int64_t i = 0; i < nG; i++) dx[i] *= o->der_factor;
        /* tau(j1, j2) = sum_k conj(SaSm(j2 nmd + k, j1)) dxi(j2, k);
         * tau(j, j) += (1/2) sum_k Re(xi_jk) Re(dxi_jk) / A_jk */
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t j1 = 0; j1 < nS; j1++) {
                cplx acc = 0;
                for (int64_t k = 0; k < nmd; k++) acc += conj(SaSm[(j2 * nmd + k) + ncy * j1]) * dx[j2 * nmd + k];
                tau[j1 + nmax * j2] = acc;
            }
        for (int64_t j = 0; j < nS; j++) {
            double d = 0.0;
            const cplx *x = gxi(o, psi0, m, s, j);
            for (int64_t k = 0; k < nmd; k++) d += creal(x[k]) * creal(dx[j * nmd + k]) / GWPA(o, k, j, m);
            tau[j + nmax * j] += d * 0.5;
        }
        if (tau_out) memcpy(DDP(o, tau_out, m, s), tau, sizeof(cplx) * NN(o));
        const cplx *D0 = DDP(o, DD0, m, s);
        cplx *dD = DDP(o, dDD, m, s);
        for (int64_t b = 0; b < nS; b++)
            for (int64_t a = 0; a < nS; a++) {
                cplx acc = 0;
                for (int64_t q = 0; q < nS; q++) acc += tau[a + nmax * q] * D0[q + nmax * b];
                dD[a + nmax * b] = -acc;
            }
    }
#undef GX
    free(gx); free(Sa0); free(SaSm); free(Sab); free(C); free(Y); free(Ha0); free(gI); free(tau);
    if (bad) {
        snprintf(o->err, sizeof o->err, "operator not supported on a Gaussian mode");
        return -1;
    }
    return 0;
}

/* Calc_Err_SPF, Gaussian branch (cmf2_derivatives.f90:293-334):
 * Tr(rho (S11 + S22 - S12 - S21)) from Gaussian overlaps. */
static double err_spf_gwp(orc *o, int64_t m, const cplx *p1, const cplx *p2) {
    const int64_t nmax = o->nspfmax, nmd = o->nmodeDOF[m];
    double error = 0.0;
    cplx *ov = xcalloc(NN(o), sizeof(cplx));
    for (int64_t s = 0; s < NSETS(o, m); s++) {
        const int64_t nS = NSPF(o, m, s);
        for (int64_t j = 0; j < nS; j++)
            ov[j + nmax * j] = 2.0 - 2.0 * creal(gau_ovl(&GWPA(o, 0, j, m), &GWPA(o, 0, j, m), gxi(o, p1, m, s, j),
                                                         gxi(o, p2, m, s, j), nmd, 1, 1));
        for (int64_t j2 = 0; j2 < nS; j2++)
            for (int64_t j1 = 0; j1 < j2; j1++) {
                const double *a1 = &GWPA(o, 0, j1, m), *a2 = &GWPA(o, 0, j2, m);
                ov[j1 + nmax * j2] = gau_ovl(a1, a2, gxi(o, p1, m, s, j1), gxi(o, p1, m, s, j2), nmd, 1, 1) +
                                     gau_ovl(a1, a2, gxi(o, p2, m, s, j1), gxi(o, p2, m, s, j2), nmd, 1, 1) -
                                     gau_ovl(a1, a2, gxi(o, p1, m, s, j1), gxi(o, p2, m, s, j2), nmd, 1, 1) -
                                     gau_ovl(a1, a2, gxi(o, p2, m, s, j1), gxi(o, p1, m, s, j2), nmd, 1, 1);
            }
        const cplx *r = o->MultiSet[m] ? RHO(o, m, s) : RHOAVE(o, m);
        for (int64_t j2 = 0; j2 < nS; j2++) {
            for (int64_t j1 = 0; j1 < j2; j1++) error += 2.0 * creal(r[j1 + nmax * j2] * conj(ov[j1 + nmax * j2]));
            error += creal(r[j2 + nmax * j2] * ov[j2 + nmax * j2]);
        }
    }
    free(ov);
    return sqrt(fabs(error));
}

/* Calc_Err_SPF, DVR branch (cmf2_derivatives.f90:239-291).  As in the
 * reference, rho(j1, j2) = sum conj(A_j1) A_j2 is paired with conj(ov(j1, j2))
 * = <d_j2|d_j1>, so the result is the norm of the change of the wavefunction
 * with conjugated coefficients (the norm of the change itself when rho is
 * real); kept, since it decides the step sequence. */
static double err_spf(orc *o, int64_t m, const cplx *p1, const cplx *p2) {
    if (o->gwp[m]) return err_spf_gwp(o, m, p1, p2);
    const int64_t nmax = o->nspfmax, nG = o->ngpm[m];
    int64_t nsets = o->MultiSet[m] ? o->ns : 1;
    double error = 0.0;
    cplx *dl = xcalloc(nG * nmax, sizeof(cplx));
    for (int64_t s = 0; s < nsets; s++) {
        int64_t nS = NSPF(o, m, s), off = IPSI(o, m, s) - 1;
        for (int64_t i = 0; i < nG * nS; i++) dl[i] = p1[off + i] - p2[off + i];
        const cplx *r = o->MultiSet[m] ? RHO(o, m, s) : RHOAVE(o, m);
        for (int64_t j2 = 0; j2 < nS; j2++) {
            for (int64_t j1 = 0; j1 < j2; j1++) {
                cplx ov = 0;
                for (int64_t g = 0; g < nG; g++) ov += conj(dl[g + nG * j1]) * dl[g + nG * j2];
                error += 2.0 * creal(r[j1 + nmax * j2] * conj(ov));
            }
            double dd = 0.0;
            for (int64_t g = 0; g < nG; g++) dd += creal(conj(dl[g + nG * j2]) * dl[g + nG * j2]);
            error += creal(r[j2 + nmax * j2] * dd);
        }
    }
    free(dl);
    return sqrt(fabs(error));
}

/* Cash-Karp (1990) and Dormand-Prince (1980) tableaux, shared by the SPF
 * integrators (cmf2_integration.f90:107-412) and the VMF ones
 * (integrator.f90:95-221). */
static const double ckB[6][5] = {
    {0, 0, 0, 0, 0}, {0.2, 0, 0, 0, 0}, {0.075, 0.225, 0, 0, 0}, {0.3, -0.9, 1.2, 0, 0},
    {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
static const double ckC[6] = {37.0 / 378.0, 0, 250.0 / 621.0, 125.0 / 594.0, 0, 512.0 / 1771.0};
static const double ckD[6] = {2825.0 / 27648.0, 0, 18575.0 / 48384.0, 13525.0 / 55296.0,
                              277.0 / 14336.0, 0.25};
static const double dpB[6][5] = {
    {0, 0, 0, 0, 0}, {0.2, 0, 0, 0, 0}, {0.075, 0.225, 0, 0, 0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0}};
static const double dpC[6] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0,
                              -2187.0 / 6784.0, 11.0 / 84.0};
static const double dpD[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0,
                              -92097.0 / 339200.0, 187.0 / 2100.0, 0.025};

/* Derivatives of the SPF block of mode m (and of its D matrices for a
 * Gaussian mode). */
static int der_any(orc *o, int64_t m, const cplx *psi, const cplx *DD, cplx *dpsi, cplx *dDD) {
    if (o->gwp[m]) return der_spf_gwp(o, m, psi, DD, dpsi, dDD, NULL);
    return der_spf(o, m, psi, dpsi);
}

/* out = y0 + dt * sum_q w[q] k[q] on the SPF block of mode m and, for a
 * Gaussian mode, on its D blocks. */
static void rk_stage(orc *o, int64_t m, const cplx *y0, const cplx *D0, double dt, const double *w, int nq,
                     cplx *const *k, cplx *const *kD, cplx *out, cplx *outD, int first) {
    const int64_t nmax = o->nspfmax;
    for (int64_t s = 0; s < NSETS(o, m); s++) {
        const int64_t i1 = IPSI(o, m, s) - 1, n = o->npar[m] * NSPF(o, m, s);
        for (int64_t i = i1; i < i1 + n; i++) {
            if (first) { out[i] = y0[i] + w[0] * dt * k[0][i]; continue; }
            cplx acc = 0;
            for (int q = 0; q < nq; q++) if (w[q] != 0.0) acc += w[q] * k[q][i];
            out[i] = y0[i] + dt * acc;
        }
        if (!o->gwp[m]) continue;
        const int64_t nS = NSPF(o, m, s);
        const size_t off = NN(o) * (m + o->nm * s);
        for (int64_t b = 0; b < nS; b++)
            for (int64_t a = 0; a < nS; a++) {
                const size_t i = off + a + nmax * b;
                if (first) { outD[i] = D0[i] + w[0] * dt * kD[0][i]; continue; }
                cplx acc = 0;
                for (int q = 0; q < nq; q++) if (w[q] != 0.0) acc += w[q] * kD[q][i];
                outD[i] = D0[i] + dt * acc;
            }
    }
}

/* rk_SPF_ck / rk_SPF_dp (cmf2_integration.f90:107-412) on the blocks of
 * mode m; for a Gaussian mode the D matrices are integrated alongside. */
static int rk_spf(orc *o, int64_t m, cplx *psi0, cplx *DD0, double htry, double *hdid, double *hnext) {
    const double safety = 0.9, pgrow = -0.2, pshrnk = -0.25, minscal = 0.1, maxscal = 5.0;
    const double errcon = pow(5.0 / safety, 1.0 / pgrow);
    const int dp = (o->p.int_type_psi == 2);
    const int64_t D = o->dimpsi;
    const size_t nD = NN(o) * o->nm * o->ns;
    cplx *k[7], *kD[7], *tmp = xcalloc(D, sizeof(cplx)), *perr = xcalloc(D, sizeof(cplx));
    cplx *tmpD = xcalloc(nD, sizeof(cplx)), *errD = xcalloc(nD, sizeof(cplx));
    for (int i = 0; i < 7; i++) { k[i] = xcalloc(D, sizeof(cplx)); kD[i] = xcalloc(nD, sizeof(cplx)); }
    memcpy(tmpD, DD0, sizeof(cplx) * nD);
    double dt = htry, error = 0.0;
    int rc = 0;
    for (;;) {
        rc |= der_any(o, m, psi0, DD0, k[0], kD[0]);
        for (int st = 1; st < 6; st++) {
            const double *b = dp ? dpB[st] : ckB[st];
            rk_stage(o, m, psi0, DD0, dt, b, st, k, kD, tmp, tmpD, st == 1);
            rc |= der_any(o, m, tmp, tmpD, k[st], kD[st]);
        }
        rk_stage(o, m, psi0, DD0, dt, dp ? dpC : ckC, 6, k, kD, tmp, tmpD, 0);
        if (!dp) {
            rk_stage(o, m, psi0, DD0, dt, ckD, 6, k, kD, perr, errD, 0);
        } else {
            rc |= der_any(o, m, tmp, tmpD, k[6], kD[6]);
            rk_stage(o, m, psi0, DD0, dt, dpD, 7, k, kD, perr, errD, 0);
        }
        error = err_spf(o, m, tmp, perr) / o->p.eps_psi;
        if (error < 1.0) break;
        double a = minscal * dt, b2 = safety * dt * pow(error, pshrnk);
        dt = a > b2 ? a : b2;
        if (rc) break;
    }
    *hnext = (error > errcon) ? safety * dt * pow(error, pgrow) : dt * maxscal;
    *hdid = dt;
    for (int64_t s = 0; s < NSETS(o, m); s++) {
        int64_t i1 = IPSI(o, m, s) - 1, n = o->npar[m] * NSPF(o, m, s);
        memcpy(psi0 + i1, tmp + i1, sizeof(cplx) * n);
        if (o->gwp[m]) {
            const size_t off = NN(o) * (m + o->nm * s);
            memcpy(DD0 + off, tmpD + off, sizeof(cplx) * NN(o));
        }
    }
    for (int i = 0; i < 7; i++) { free(k[i]); free(kD[i]); }
    free(tmp); free(perr); free(tmpD); free(errD);
    return rc;
}

/* propagate_SPF (cmf2_base.f90:210-260). */
static int propagate_spf(orc *o, double dt, int64_t m, cplx *psi, cplx *DD) {
    double dt_done = 0.0, stepsize = dt, step_done, step_next;
    const double minstep = o->p.min_stepsize;
    int rc = 0;
    for (;;) {
        rc |= rk_spf(o, m, psi, DD, stepsize, &step_done, &step_next);
        dt_done = dt_done + step_done;
        if (dt_done >= dt || rc) break;
        if (step_next < minstep * 2) step_next = minstep * 2;
        stepsize = (dt - dt_done) / (double)((int64_t)((dt - dt_done) / step_next) + 1);
        if (stepsize > dt - dt_done) stepsize = dt - dt_done;
    }
    return rc;
}

int orc_propagate_spf(orc *o, double dt, int64_t i, double *psi) {
    return propagate_spf(o, dt, i - 1, (cplx *)psi, o->DD);
}

/* norm_wf and WF_Ovl, DVR modes (gmctdh_objects.f90:21-97, 101-180). */
static double norm_wf(orc *o, const cplx *A0, const cplx *psi0) {
    const int64_t nm = o->nm, nmax = o->nspfmax;
    int64_t big = 0;
    for (int64_t s = 0; s < o->ns; s++) if (o->nAConf[s] > big) big = o->nAConf[s];
    cplx *p1 = xcalloc(big, sizeof(cplx)), *p2 = xcalloc(big, sizeof(cplx));
    cplx *S = xcalloc(nmax * nmax, sizeof(cplx));
    double wfnorm = 0.0;
    int64_t sz[64];
    for (int64_t s = 0; s < o->ns; s++) {
        for (int64_t m = 0; m < nm; m++) sz[m] = NSPF(o, m, s);
        memcpy(p1, A0 + o->AStateOffset[s], sizeof(cplx) * o->nAConf[s]);
        /* grid modes first, then Gaussian modes (gmctdh_objects.f90:56-90) */
        for (int64_t pass = 0; pass < 2; pass++)
        for (int64_t m = 0; m < nm; m++) {
            if (o->gwp[m] != pass) continue;
            int64_t nS = sz[m], nG = o->ngpm[m], nmd = o->nmodeDOF[m];
            const cplx *phi = psi0 + IPSI(o, m, s) - 1;
            for (int64_t b = 0; b < nS; b++)
                for (int64_t a = 0; a <= b; a++) {
                    cplx x = 0;
                    if (o->gwp[m]) {
                        x = (a == b) ? 1.0
                                     : gau_ovl(&GWPA(o, 0, a, m), &GWPA(o, 0, b, m), phi + nmd * a, phi + nmd * b,
                                               nmd, 1, 1);
                    } else {
                        for (int64_t g = 0; g < nG; g++) x += conj(phi[g + nG * a]) * phi[g + nG * b];
                        if (a == b) x = creal(x);
                    }
                    S[a + nmax * b] = x;
                    S[b + nmax * a] = conj(x);
                }
            mode_product(nm, sz, m, S, nmax, nS, p1, p2);
            cplx *sw = p1; p1 = p2; p2 = sw;
        }
        const cplx *a = A0 + o->AStateOffset[s];
        cplx acc = 0;
        for (int64_t i = 0; i < o->nAConf[s]; i++) acc += conj(p1[i]) * a[i];
        wfnorm += creal(acc);
    }
    free(p1); free(p2); free(S);
    return wfnorm;
}

static cplx wf_ovl(orc *o, const cplx *A1, const cplx *psi1, const cplx *A2, const cplx *psi2) {
    const int64_t nm = o->nm, nmax = o->nspfmax;
    int64_t big = 0;
    for (int64_t s = 0; s < o->ns; s++) if (o->nAConf[s] > big) big = o->nAConf[s];
    cplx *p1 = xcalloc(big, sizeof(cplx)), *p2 = xcalloc(big, sizeof(cplx));
    cplx *S = xcalloc(nmax * nmax, sizeof(cplx));
    cplx overlap = 0;
    int64_t sz[64];
    for (int64_t s = 0; s < o->ns; s++) {
        for (int64_t m = 0; m < nm; m++) sz[m] = NSPF(o, m, s);
        memcpy(p1, A2 + o->AStateOffset[s], sizeof(cplx) * o->nAConf[s]);
        for (int64_t pass = 0; pass < 2; pass++)
        for (int64_t m = 0; m < nm; m++) {
            if (o->gwp[m] != pass) continue;
            int64_t nS = sz[m], nG = o->ngpm[m], nmd = o->nmodeDOF[m];
            const cplx *f1 = psi1 + IPSI(o, m, s) - 1, *f2 = psi2 + IPSI(o, m, s) - 1;
            for (int64_t b = 0; b < nS; b++)
                for (int64_t a = 0; a < nS; a++) {
                    cplx x = 0;
                    if (o->gwp[m]) {
                        x = gau_ovl(&GWPA(o, 0, a, m), &GWPA(o, 0, b, m), f1 + nmd * a, f2 + nmd * b, nmd, 1, 1);
                    } else {
                        for (int64_t g = 0; g < nG; g++) x += conj(f1[g + nG * a]) * f2[g + nG * b];
                    }
                    S[a + nmax * b] = x;
                }
            mode_product(nm, sz, m, S, nmax, nS, p1, p2);
            cplx *sw = p1; p1 = p2; p2 = sw;
        }
        const cplx *a = A1 + o->AStateOffset[s];
        for (int64_t i = 0; i < o->nAConf[s]; i++) overlap += conj(a[i]) * p1[i];
    }
    free(p1); free(p2); free(S);
    return overlap;
}

double orc_norm_wf(orc *o, const double *A, const double *psi) {
    return norm_wf(o, (const cplx *)A, (const cplx *)psi);
}
void orc_wf_ovl(orc *o, const double *A1, const double *psi1, const double *A2,
                const double *psi2, double *ovl) {
    cplx v = wf_ovl(o, (const cplx *)A1, (const cplx *)psi1, (const cplx *)A2, (const cplx *)psi2);
    ovl[0] = creal(v);
    ovl[1] = cimag(v);
}

/* DeltaPsi, npackets = 1 (gmctdh_objects.f90:206-227). */
static double delta_psi(orc *o, const cplx *A1, const cplx *psi1, const cplx *A2, const cplx *psi2) {
    double o11 = norm_wf(o, A1, psi1), o22 = norm_wf(o, A2, psi2);
    cplx o12 = wf_ovl(o, A1, psi1, A2, psi2);
    return sqrt(fabs(2.0 * (1.0 - creal(o12) / sqrt(o11 * o22))));
}

double orc_delta_psi(orc *o, const double *A1, const double *psi1, const double *A2, const double *psi2) {
    return delta_psi(o, (const cplx *)A1, (const cplx *)psi1, (const cplx *)A2, (const cplx *)psi2);
}
double orc_err_spf(orc *o, int64_t im, const double *psi1, const double *psi2) {
    return err_spf(o, im, (const cplx *)psi1, (const cplx *)psi2);
}

static double vnorm(const cplx *x, int64_t n) {
    double acc = 0.0;
    for (int64_t i = 0; i < n; i++) acc += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
    return sqrt(acc);
}

/* Lanczos (cmf2_integration.f90:30-103). Returns nKrylov, or -1 when the
 * maximum Krylov size is reached (the reference stops there). */
static int lanczos(orc *o, const cplx *B0, cplx *BT, double dt, int build) {
    const int64_t N = o->nconf;
    cplx *BK = o->BK;
    if (build) {
        double beta[NKRYLOV_MAX], alpha[NKRYLOV_MAX];
        memcpy(BK, B0, sizeof(cplx) * N);
        double d0 = vnorm(BK, N);
        for (int64_t i = 0; i < N; i++) BK[i] /= d0;
        apply_h(o, BK, BK + N);
        cplx acc = 0;
        for (int64_t i = 0; i < N; i++) acc += conj(BK[i]) * BK[N + i];
        alpha[0] = creal(acc);
        for (int64_t i = 0; i < N; i++) BK[N + i] -= alpha[0] * BK[i];
        beta[0] = vnorm(BK + N, N);
        for (int64_t i = 0; i < N; i++) BK[N + i] /= beta[0];
        double error = beta[0] * fabs(dt);
        int j = 2;
        for (;;) {
            if (j == NKRYLOV_MAX) {
                snprintf(o->err, sizeof o->err, "Maximum number of Krylov vectors reached");
                return -1;
            }
            cplx *vj = BK + (size_t)N * (j - 1), *vn = BK + (size_t)N * j, *vp = BK + (size_t)N * (j - 2);
            apply_h(o, vj, vn);
            acc = 0;
            for (int64_t i = 0; i < N; i++) acc += conj(vj[i]) * vn[i];
            alpha[j - 1] = creal(acc);
            for (int64_t i = 0; i < N; i++) vn[i] -= beta[j - 2] * vp[i];
            for (int64_t i = 0; i < N; i++) vn[i] -= alpha[j - 1] * vj[i];
            beta[j - 1] = vnorm(vn, N);
            for (int64_t i = 0; i < N; i++) vn[i] /= beta[j - 1];
            error = error * beta[j - 1] / (double)j * fabs(dt);
            if (error < EPS_B) break;
            j++;
        }
        o->nkrylov = j;
        double w[NKRYLOV_MAX];
        tridiag_eig(j, alpha, beta, w, o->lan_U);
        for (int r = 0; r < j; r++) o->lan_alpha[r] = w[r];
    }
    const int n = o->nkrylov;
    cplx coef[NKRYLOV_MAX];
    for (int i = 0; i < n; i++) {
        cplx acc = 0;
        for (int r = 0; r < n; r++)
            acc += o->lan_U[i + n * r] * o->lan_U[0 + n * r] * cexp(o->der_factor * o->lan_alpha[r] * dt);
        coef[i] = acc;
    }
    for (int64_t q = 0; q < N; q++) {
        cplx acc = 0;
        for (int i = 0; i < n; i++) acc += BK[(size_t)N * i + q] * coef[i];
        BT[q] = acc;
    }
    double dn = vnorm(BT, N);
    for (int64_t q = 0; q < N; q++) BT[q] /= dn;
    return n;
}

int orc_lanczos(orc *o, const double *B0, double *BT, double dt, int build) {
    return lanczos(o, (const cplx *)B0, (cplx *)BT, dt, build);
}

/* cmf2 (cmf2_base.f90:47-206).  With Gaussian modes the coefficients are
 * propagated as B = D^H S A in the basis orthonormalised by D (D0 from the
 * Cholesky factor of S00 at the start of the step), the SPF propagation
 * carries D along, and A = D B at the half step and at the end; without
 * them A_to_B / B_to_A are copies. */
static int cmf2(orc *o, double tt, double htry, double *hdid_out, double *hnext_out,
                mg_attempt *atts, int64_t max_att, int64_t *n_att) {
    const int64_t N = o->nconf, D = o->dimpsi;
    const size_t nD = NN(o) * o->nm * o->ns;
    const double safe = 0.95, maxscal = 1.5;
    cplx *B0 = xcalloc(N, sizeof(cplx)), *BH = xcalloc(N, sizeof(cplx)), *B02 = xcalloc(N, sizeof(cplx));
    cplx *AH = o->ngwp ? xcalloc(N, sizeof(cplx)) : BH;
    cplx *pH = xcalloc(D, sizeof(cplx)), *pH2 = xcalloc(D, sizeof(cplx));
    cplx *DH = xcalloc(nD, sizeof(cplx)), *DH2 = xcalloc(nD, sizeof(cplx));
    double hdid = htry;
    int rc = 0;
    if (o->ngwp) rc |= init_dd(o, o->psi, o->DD);
    for (;;) {
        if (rc) break;
        for (int64_t m = 0; m < o->nm; m++) if (o->gwp[m]) calc_s00(o, o->psi, m);
        calc_rho(o, o->A);
        for (int64_t m = 0; m < o->nm; m++) calc_h1mode_any(o, o->psi, m);
        calc_meanfield(o, o->A);
        if (o->ngwp) calc_dhd(o, o->DD);
        a_to_b(o, o->A, B0, o->DD);
        int nk1 = lanczos(o, B0, BH, 0.5 * hdid, 1);
        if (nk1 < 0) { rc = -1; break; }
        memcpy(pH, o->psi, sizeof(cplx) * D);
        memcpy(DH, o->DD, sizeof(cplx) * nD);
        for (int64_t pass = 0; pass < 2; pass++)
            for (int64_t m = 0; m < o->nm; m++)
                if (o->gwp[m] == pass) rc |= propagate_spf(o, 0.5 * hdid, m, pH, DH);
        b_to_a(o, BH, AH, DH);
        if (o->cmf2_fault != 1) {   /* fault 1: keep the t0 mean fields */
            for (int64_t m = 0; m < o->nm; m++) if (o->gwp[m]) calc_s00(o, pH, m);
            calc_rho(o, AH);
            for (int64_t m = 0; m < o->nm; m++) calc_h1mode_any(o, pH, m);
            calc_meanfield(o, AH);
        }
        memcpy(pH2, o->psi, sizeof(cplx) * D);
        memcpy(DH2, o->DD, sizeof(cplx) * nD);
        for (int64_t pass = 0; pass < 2; pass++)
            for (int64_t m = 0; m < o->nm; m++)
                if (o->gwp[m] == pass) rc |= propagate_spf(o, 0.5 * hdid, m, pH2, DH2);
        double err_psi = delta_psi(o, AH, pH, AH, pH2);
        for (int64_t pass = 0; pass < 2; pass++)
            for (int64_t m = 0; m < o->nm; m++)
                if (o->gwp[m] == pass) rc |= propagate_spf(o, 0.5 * hdid, m, pH2, DH2);
        for (int64_t m = 0; m < o->nm; m++) if (o->gwp[m]) calc_s00(o, pH2, m);
        for (int64_t m = 0; m < o->nm; m++) calc_h1mode_any(o, pH2, m);
        if (o->ngwp) calc_dhd(o, DH2);
        int nk2 = lanczos(o, BH, B02, -0.5 * hdid, 1);
        if (nk2 < 0) { rc = -1; break; }
        for (int64_t i = 0; i < N; i++) B02[i] -= B0[i];
        double err_A = vnorm(B02, N) * 0.5;
        double error = (err_A > err_psi ? err_A : err_psi) / o->p.eps_int;
        if (atts && n_att && *n_att < max_att) {
            mg_attempt *a = &atts[(*n_att)++];
            a->time = tt; a->h = hdid; a->err_A = err_A; a->err_psi = err_psi;
            a->accepted = !(error > 1.0); a->nkrylov1 = nk1; a->nkrylov2 = nk2;
        }
        if (rc) break;
        if (error > 1.0) {
            hdid = hdid / sqrt(error) * safe;
            /* The reference has no lower bound here (cmf2_base.f90:181-184) and
             * loops forever once err_psi sits on its rounding floor
             * (~sqrt(4 eps) = 2.1e-8 > eps_int); stop with an error instead. */
            if (hdid < o->p.min_stepsize) {
                snprintf(o->err, sizeof o->err, "CMF2 step size fell below minstep");
                rc = -2;
                break;
            }
            continue;
        }
        if (o->cmf2_fault == 2) {
            /* fault 2: A over the whole step with the t0 Hamiltonian */
            for (int64_t m = 0; m < o->nm; m++) calc_h1mode_any(o, o->psi, m);
            if (o->ngwp) calc_dhd(o, o->DD);
            a_to_b(o, o->A, B0, o->DD);
            memcpy(BH, B0, sizeof(cplx) * N);
            if (lanczos(o, BH, B0, hdid, 1) < 0) { rc = -1; break; }
        } else {
            lanczos(o, BH, B0, 0.5 * hdid, 0);
        }
        b_to_a(o, B0, o->A, DH2);
        memcpy(o->psi, pH2, sizeof(cplx) * D);
        memcpy(o->DD, DH2, sizeof(cplx) * nD);
        double hn = hdid / sqrt(error), h2 = hdid * maxscal;
        hn = hn < h2 ? hn : h2;
        hn = hn < o->p.writetime ? hn : o->p.writetime;
        *hnext_out = hn;
        break;
    }
    *hdid_out = hdid;
    free(B0); free(BH); free(B02); free(pH); free(pH2); free(DH); free(DH2);
    if (o->ngwp) free(AH);
    return rc;
}

/* Test hook: plant a mistake in the CMF2 step, so that the order test can
 * show it fails.  1: the SPFs' second half step uses the mean fields of t0
 * (no re-evaluation at the half step); 2: the accepted A is propagated over
 * the whole step with the t0 Hamiltonian.  0 restores the scheme. */
void orc_set_cmf2_fault(orc *o, int64_t fault) { o->cmf2_fault = fault; }

int orc_cmf2_step(orc *o, double t, double htry, double *hdid, double *hnext,
                  mg_attempt *atts, int64_t max_att, int64_t *n_att) {
    return cmf2(o, t, htry, hdid, hnext, atts, max_att, n_att);
}

/* WriteOutput observables (output.f90:33-87, 143-150). */
int orc_observables(orc *o, double t, mg_output *out, double *pops) {
    for (int64_t m = 0; m < o->nm; m++) if (o->gwp[m]) calc_s00(o, o->psi, m);
    calc_rho(o, o->A);
    for (int64_t m = 0; m < o->nm; m++) calc_h1mode_any(o, o->psi, m);
    calc_meanfield(o, o->A);
    double tot = 0.0;
    const int64_t nmax = o->nspfmax;
    for (int64_t s = 0; s < o->ns; s++) {
        int64_t jel = o->MultiSet[0] ? s : 0;
        int64_t nS = NSPF(o, 0, jel);
        double pop = 0.0;
        if (o->gwp[0]) {
            /* Gaussian first mode: Re sum rho o S00 (output.f90:64-74) */
            const cplx *r = RHO(o, 0, s), *S = S00P(o, 0, jel, jel);
            for (int64_t j2 = 0; j2 < nS; j2++) {
                for (int64_t j1 = 0; j1 < j2; j1++) pop += 2.0 * creal(r[j1 + nmax * j2] * S[j1 + nmax * j2]);
                pop += creal(r[j2 + nmax * j2] * S[j2 + nmax * j2]);
            }
        } else
        for (int64_t i = 0; i < nS; i++) pop += creal(RHO(o, 0, s)[i + o->nspfmax * i]);
        if (pops) pops[s] = pop;
        tot += pop;
    }
    cplx c = wf_ovl(o, o->Astart, o->psistart, o->A, o->psi);
    out->time = t;
    out->auto_re = creal(c);
    out->auto_im = cimag(c);
    out->norm = tot;
    out->energy = o->energyMF / tot;
    return 0;
}

/* propagate_wf, CMF2 branch (propagator.f90:87-106, 257-306). */
/* Renormalisation of propagate_wf (propagator.f90:129-144, 288-305):
 * A /= sqrt(Trace_wf(A, psi)) (gmctdh_objects.f90:184-202, one packet).
 * Mode 3 after every step that does not end an output interval; every
 * nonzero mode at the end of an interval, mode 1 only once. */
static void renormalise(orc *o) {
    const double n = norm_wf(o, o->A, o->psi);
    const double f = 1.0 / sqrt(n);
    for (int64_t i = 0; i < o->nconf; i++) o->A[i] *= f;
}
static void renormalise_interval(orc *o) {
    if (o->inorm_psi == 0) return;
    renormalise(o);
    if (o->inorm_psi == 1) o->inorm_psi = 0;
}
void orc_set_normalize(orc *o, int64_t mode) { o->inorm_psi = mode; }

int orc_propagate(orc *o, double t0, double tfinal, mg_output *outs, double *pops,
                  int64_t max_out, int64_t *n_out, mg_attempt *atts, int64_t max_att,
                  int64_t *n_att) {
    const double wt = o->p.writetime, minstep = o->p.min_stepsize;
    double maxtime = tfinal + wt * 0.1;
    double time = t0, stepsize = wt * o->p.eps_int;
    *n_out = 0;
    if (n_att) *n_att = 0;
    while (time < maxtime) {
        if (*n_out < max_out) {
            orc_observables(o, time, &outs[*n_out], pops ? pops + o->ns * *n_out : NULL);
            (*n_out)++;
        }
        double tnext = time + wt;
        for (;;) {
            double done, next;
            if (cmf2(o, time, stepsize, &done, &next, atts, max_att, n_att) != 0) return -1;
            time = time + done;
            if (time >= tnext - minstep) break;
            int64_t k = (int64_t)((tnext - time) / next);
            stepsize = (k == 0) ? tnext - time : (tnext - time) / (double)(k + 1);
            if (o->inorm_psi == 3) renormalise(o);
        }
        renormalise_interval(o);
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Variable mean field (VMF) integration, DVR modes, npackets = 1.          */

/* Calc_Derivs (derivatives.f90:39-243): mean fields from (A, psi);
 * dA = conj(Adot_tmp) = der_factor * H A (gmctdh_objects.f90:578,
 * derivatives.f90:156, 237); dpsi from the DVR formula of every mode
 * (derivatives.f90:81-143, the same as Calc_Der_SPF's). */
/* Calc_Derivs with Gaussian modes (derivatives.f90:61-89, 155-238):
 * Gaussian-parameter derivatives from the regularised C/Y solve; dA =
 * S^-1 (der_factor H A) - sum_m tau_m A with the Tikhonov S00^-1 on every
 * Gaussian mode and tau_m of Calc_Der_SPF (the Gaussian-basis H includes
 * S00 for the identity on Gaussian modes). */
static int calc_derivs_gwp(orc *o, const cplx *A0, const cplx *psi0, cplx *dA, cplx *dpsi) {
    const int64_t nm = o->nm, N = o->nconf;
    const size_t nD = NN(o) * nm * o->ns;
    for (int64_t m = 0; m < nm; m++)
        if (o->gwp[m]) { calc_s00(o, psi0, m); calc_s00m1(o, m); }
    calc_rho(o, A0);
    for (int64_t m = 0; m < nm; m++) calc_h1mode_any(o, psi0, m);
    if (calc_meanfield(o, A0) != 0) return -1;
    memset(dpsi, 0, sizeof(cplx) * o->dimpsi);
    int rc = 0;
    for (int64_t m = 0; m < nm; m++) if (!o->gwp[m]) rc |= der_spf(o, m, psi0, dpsi);
    cplx *Did = xcalloc(nD, sizeof(cplx)), *dD = xcalloc(nD, sizeof(cplx)), *tau = xcalloc(nD, sizeof(cplx));
    for (int64_t m = 0; m < nm; m++)
        for (int64_t s = 0; s < o->ns; s++)
            for (int64_t j = 0; j < o->nspfmax; j++) DDP(o, Did, m, s)[j + o->nspfmax * j] = 1.0;
    for (int64_t m = 0; m < nm; m++) {
        if (!o->gwp[m]) continue;
        /* der_spf_gwp recomputes S00, S00^-1 and h1mode of m from psi0: unchanged */
        rc |= der_spf_gwp(o, m, psi0, Did, dpsi, dD, tau);
    }
    /* dA = S^-1 (der_factor H A) over the Gaussian modes */
    cplx *HA = xcalloc(N, sizeof(cplx)), *p1 = xcalloc(mixed_size(o), sizeof(cplx)), *p2 = xcalloc(mixed_size(o), sizeof(cplx));
    apply_hg(o, A0, HA);
    int64_t sz[64];
    for (int64_t s = 0; s < o->ns; s++) {
        for (int64_t m = 0; m < nm; m++) sz[m] = NSPF(o, m, s);
        const int64_t n = o->nAConf[s];
        for (int64_t i = 0; i < n; i++) p1[i] = o->der_factor * HA[o->AStateOffset[s] + i];
        for (int64_t m = 0; m < nm; m++) {
            if (!o->gwp[m]) continue;
            const int64_t jel = o->MultiSet[m] ? s : 0;
            mode_product(nm, sz, m, S00M1(o, m, jel), o->nspfmax, sz[m], p1, p2);
            cplx *sw = p1; p1 = p2; p2 = sw;
        }
        memcpy(dA + o->AStateOffset[s], p1, sizeof(cplx) * n);
        /* minus tau_m A for every Gaussian mode */
        for (int64_t m = 0; m < nm; m++) {
            if (!o->gwp[m]) continue;
            const int64_t jel = o->MultiSet[m] ? s : 0;
            mode_product(nm, sz, m, DDP(o, tau, m, jel), o->nspfmax, sz[m], A0 + o->AStateOffset[s], p2);
            for (int64_t i = 0; i < n; i++) dA[o->AStateOffset[s] + i] -= p2[i];
        }
    }
    free(Did); free(dD); free(tau); free(HA); free(p1); free(p2);
    return rc;
}

static int calc_derivs(orc *o, const cplx *A0, const cplx *psi0, cplx *dA, cplx *dpsi) {
    if (o->ngwp) return calc_derivs_gwp(o, A0, psi0, dA, dpsi);
    calc_rho(o, A0);
    for (int64_t m = 0; m < o->nm; m++) calc_h1mode_mode(o, psi0, m);
    if (calc_meanfield(o, A0) != 0) return -1;
    apply_h(o, A0, dA);
    for (int64_t i = 0; i < o->nconf; i++) dA[i] *= o->der_factor;
    memset(dpsi, 0, sizeof(cplx) * o->dimpsi);
    int rc = 0;
    for (int64_t m = 0; m < o->nm; m++) rc |= der_spf(o, m, psi0, dpsi);
    return rc;
}

int orc_calc_derivs(orc *o, const double *A, const double *psi, double *dA, double *dpsi) {
    return calc_derivs(o, (const cplx *)A, (const cplx *)psi, (cplx *)dA, (cplx *)dpsi);
}

/* y_out = y0 + dt * sum_q w[q] k[q] over n entries (zero weights skipped). */
static void rk_combine(int64_t n, const cplx *y0, double dt, const double *w, int nq, cplx *const *k, cplx *out) {
    for (int64_t i = 0; i < n; i++) {
        cplx acc = 0;
        for (int q = 0; q < nq; q++) if (w[q] != 0.0) acc += w[q] * k[q][i];
        out[i] = y0[i] + dt * acc;
    }
}

/* rkck / rkdp (integrator.f90:95-221) on (A, psi).  kA[0], kp[0] hold the
 * derivatives at (A0, psi0); for Dormand-Prince kA[6], kp[6] receive the
 * derivatives at the 5th-order result (reused by the next step). */
static int rk_vmf(orc *o, int dp, const cplx *A0, const cplx *p0, cplx **kA, cplx **kp, double dt, cplx *At,
                  cplx *pt, cplx *Ae, cplx *pe) {
    const int64_t N = o->nconf, D = o->dimpsi;
    int rc = 0;
    for (int st = 1; st < 6; st++) {
        const double *b = dp ? dpB[st] : ckB[st];
        rk_combine(N, A0, dt, b, st, kA, At);
        rk_combine(D, p0, dt, b, st, kp, pt);
        rc |= calc_derivs(o, At, pt, kA[st], kp[st]);
    }
    rk_combine(N, A0, dt, dp ? dpC : ckC, 6, kA, At);
    rk_combine(D, p0, dt, dp ? dpC : ckC, 6, kp, pt);
    if (dp) {
        rc |= calc_derivs(o, At, pt, kA[6], kp[6]);
        rk_combine(N, A0, dt, dpD, 7, kA, Ae);
        rk_combine(D, p0, dt, dpD, 7, kp, pe);
    } else {
        rk_combine(N, A0, dt, ckD, 6, kA, Ae);
        rk_combine(D, p0, dt, ckD, 6, kp, pe);
    }
    return rc;
}

/* rkqs (integrator.f90:36-91): error = DeltaPsi(5th, 4th order) / eps_int. */
static int rkqs_vmf(orc *o, int dp, cplx **kA, cplx **kp, double htry, double tt, double *hdid, double *hnext,
                    mg_attempt *atts, int64_t max_att, int64_t *n_att) {
    const double safety = 0.9, pgrow = -0.2, pshrnk = -0.25, minscal = 0.1, maxscal = 5.0;
    const double errcon = pow(5.0 / safety, 1.0 / pgrow);
    const int64_t N = o->nconf, D = o->dimpsi;
    const double minstep = o->p.min_stepsize;
    cplx *At = xcalloc(N, sizeof(cplx)), *Ae = xcalloc(N, sizeof(cplx));
    cplx *pt = xcalloc(D, sizeof(cplx)), *pe = xcalloc(D, sizeof(cplx));
    double dt = htry, errmax = 0.0;
    int rc = 0;
    for (;;) {
        rc |= rk_vmf(o, dp, o->A, o->psi, kA, kp, dt, At, pt, Ae, pe);
        errmax = delta_psi(o, At, pt, Ae, pe) / o->p.eps_int;
        const int accept = (errmax < 1.0 || dt <= minstep);
        if (atts && n_att && *n_att < max_att)
            atts[(*n_att)++] = (mg_attempt){tt, dt, errmax * o->p.eps_int, 0.0, accept, 0, 0};
        if (accept || rc) break;
        double a = minscal * dt, b = safety * dt * pow(errmax, pshrnk);
        dt = a > b ? a : b;
        if (dt < minstep) dt = minstep;
    }
    double hn = (errmax > errcon) ? safety * dt * pow(errmax, pgrow) : dt * maxscal;
    if (hn < minstep) hn = minstep;
    if (hn > o->p.writetime) hn = o->p.writetime;
    *hnext = hn;
    *hdid = dt;
    memcpy(o->A, At, sizeof(cplx) * N);
    memcpy(o->psi, pt, sizeof(cplx) * D);
    if (dp) {   /* first-same-as-last: derivatives of the new point */
        memcpy(kA[0], kA[6], sizeof(cplx) * N);
        memcpy(kp[0], kp[6], sizeof(cplx) * D);
    }
    free(At); free(Ae); free(pt); free(pe);
    return rc;
}

/* propagate_wf, Runge-Kutta branch (propagator.f90:108-140); method 1 =
 * Cash-Karp, 2 = Dormand-Prince.  The reference enters its first
 * Dormand-Prince step with derivative arrays it never filled
 * (propagator.f90:97-100, 113); here they are computed at t0. */
int orc_propagate_vmf(orc *o, int64_t method, double t0, double tfinal, mg_output *outs, double *pops,
                      int64_t max_out, int64_t *n_out, mg_attempt *atts, int64_t max_att, int64_t *n_att) {
    const int dp = (method == 2);
    const double wt = o->p.writetime, minstep = o->p.min_stepsize;
    const int64_t N = o->nconf, D = o->dimpsi;
    double maxtime = tfinal + wt * 0.1;
    double time = t0, stepsize = wt * o->p.eps_int;
    cplx *kA[7], *kp[7];
    for (int q = 0; q < 7; q++) { kA[q] = xcalloc(N, sizeof(cplx)); kp[q] = xcalloc(D, sizeof(cplx)); }
    int rc = 0;
    *n_out = 0;
    if (n_att) *n_att = 0;
    if (dp) rc |= calc_derivs(o, o->A, o->psi, kA[0], kp[0]);
    while (time < maxtime && !rc) {
        if (*n_out < max_out) {
            orc_observables(o, time, &outs[*n_out], pops ? pops + o->ns * *n_out : NULL);
            (*n_out)++;
        }
        double tnext = time + wt;
        for (;;) {
            double done, next;
            if (!dp) rc |= calc_derivs(o, o->A, o->psi, kA[0], kp[0]);
            rc |= rkqs_vmf(o, dp, kA, kp, stepsize, time, &done, &next, atts, max_att, n_att);
            if (rc) break;
            time = time + done;
            if (time >= tnext - minstep) break;
            int64_t k = (int64_t)((tnext - time) / next);
            stepsize = (k == 0) ? tnext - time : (tnext - time) / (double)(k + 1);
            if (o->inorm_psi == 3) renormalise(o);
        }
        if (!rc) renormalise_interval(o);
    }
    for (int q = 0; q < 7; q++) { free(kA[q]); free(kp[q]); }
    return rc ? -1 : 0;
}

/* Test hook: one explicit step of the Cash-Karp (1) or Dormand-Prince (2)
 * pair on y' = lam y, y(0) = 1: y5 (propagated) and y4 (embedded). */
void orc_rk_scalar_step(int64_t method, double lam_re, double lam_im, double dt, double *y5, double *y4) {
    const int dp = (method == 2);
    const cplx lam = lam_re + I * lam_im;
    cplx k[7], y = 1.0;
    k[0] = lam * y;
    for (int st = 1; st < 6; st++) {
        const double *b = dp ? dpB[st] : ckB[st];
        cplx acc = 0;
        for (int q = 0; q < st; q++) acc += b[q] * k[q];
        k[st] = lam * (y + dt * acc);
    }
    cplx a5 = 0, a4 = 0;
    for (int q = 0; q < 6; q++) a5 += (dp ? dpC : ckC)[q] * k[q];
    const cplx r5 = y + dt * a5;
    if (dp) {
        k[6] = lam * r5;
        for (int q = 0; q < 7; q++) a4 += dpD[q] * k[q];
    } else {
        for (int q = 0; q < 6; q++) a4 += ckD[q] * k[q];
    }
    const cplx r4 = y + dt * a4;
    y5[0] = creal(r5); y5[1] = cimag(r5);
    y4[0] = creal(r4); y4[1] = cimag(r4);
}

/* test hooks */
void orc_set_reduction_chunks(int64_t n) { reduction_chunks = n < 1 ? 1 : n; }
void orc_std_reg_inv(int64_t n, const double *A, double eps, int64_t type, double *B) {
    std_reg_inv(n, (const cplx *)A, eps, (int)type, (cplx *)B);
}
void orc_herm_eig(int64_t n, const double *A, double *w, double *U) {
    herm_eig(n, (const cplx *)A, w, (cplx *)U);
}
void orc_tridiag_eig(int64_t n, const double *alpha, const double *beta, double *w, double *U) {
    tridiag_eig(n, alpha, beta, w, U);
}

/* Gaussian-mode hooks for the pinning tests. */
void orc_gau_ovl(int64_t nmd, const double *a1, const double *a2, const double *xi1, const double *xi2, double *out) {
    cplx v = gau_ovl(a1, a2, (const cplx *)xi1, (const cplx *)xi2, nmd, 1, 1);
    out[0] = creal(v);
    out[1] = cimag(v);
}
int orc_gau_int(double A1, double A2, const double *xi1, const double *xi2, int64_t op, int64_t deriv, double *out) {
    int bad = 0;
    const cplx x1 = xi1[0] + I * xi1[1], x2 = xi2[0] + I * xi2[1];
    cplx v = deriv ? gau10_int(A1, A2, x1, x2, op, &bad) : gau_int(A1, A2, x1, x2, op, &bad);
    out[0] = creal(v);
    out[1] = cimag(v);
    return bad ? -1 : 0;
}
void orc_std_reg_sol_tikh(int64_t n, const double *A, const double *y, double eps, double *x) {
    std_reg_sol_tikh(n, (const cplx *)A, n, (const cplx *)y, eps, (cplx *)x);
}
int orc_inverse_sqrt_u(int64_t n, const double *S, double *D) {
    return inverse_sqrt_u(n, (const cplx *)S, n, (cplx *)D, n);
}
/* S00, then D0 = InverseSquareRootU(S00) of the current psi into dd_out
 * ([nspfmax, nspfmax, nmodes, nstates]) and the overlaps into s00_out
 * ([nspfmax, nspfmax, nmodes, nstates, nstates]); either may be NULL. */
int orc_gwp_overlaps(orc *o, double *s00_out, double *dd_out) {
    int rc = init_dd(o, o->psi, o->DD);
    if (s00_out) memcpy(s00_out, o->S00, sizeof(cplx) * NN(o) * o->nm * o->ns * o->ns);
    if (dd_out) memcpy(dd_out, o->DD, sizeof(cplx) * NN(o) * o->nm * o->ns);
    return rc;
}
/* Calc_Der_SPF of Gaussian mode im (1-based) at psi_in with D = dd_in,
 * using rho and the mean fields of the current state (A, psi). */
int orc_der_spf_gwp(orc *o, int64_t im, const double *psi_in, const double *dd_in, double *dpsi_out,
                    double *ddd_out) {
    for (int64_t m = 0; m < o->nm; m++) if (o->gwp[m]) calc_s00(o, o->psi, m);
    calc_rho(o, o->A);
    for (int64_t m = 0; m < o->nm; m++) calc_h1mode_any(o, o->psi, m);
    calc_meanfield(o, o->A);
    return der_spf_gwp(o, im - 1, (const cplx *)psi_in, (const cplx *)dd_in, (cplx *)dpsi_out, (cplx *)ddd_out,
                       NULL);
}
/* A_to_B / B_to_A with the D matrices dd ([nspfmax, nspfmax, nmodes,
 * nstates]) and the overlaps of the current psi. */
void orc_a_to_b(orc *o, const double *A, double *B, const double *dd) {
    for (int64_t m = 0; m < o->nm; m++) if (o->gwp[m]) calc_s00(o, o->psi, m);
    a_to_b(o, (const cplx *)A, (cplx *)B, (const cplx *)dd);
}
void orc_b_to_a(orc *o, const double *B, double *A, const double *dd) {
    b_to_a(o, (const cplx *)B, (cplx *)A, (const cplx *)dd);
}
/* S00, h1mode, D0 and D^H h D of the current psi, so that orc_apply_h is
 * Calc_DHD_B in the orthonormalised basis (cmf2_base.f90:72-96). */
int orc_prepare_gwp(orc *o) {
    int rc = init_dd(o, o->psi, o->DD);
    calc_operators(o, o->psi);
    calc_dhd(o, o->DD);
    return rc;
}

