// Internal declarations shared by the device kernels (mg_kernels.hip) and the
// host engine (mg_engine.hip).  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mg {

// Modes fused into one LDS brick.  The mean-field Gram program of a term
// with an operator on every mode of a q-mode group holds q(q-1)/2 + q - 1 op
// matrices (multi-set inter-state terms have an overlap on every mode); q = 6
// keeps that within kMaxBrickMats.
constexpr int kMaxGroupModes = 6;
constexpr int kMaxDofPerMode = 4;
constexpr int kNKrylovMax = 20;     // cmf2_integration.f90:12

// Complex helpers on double2 / float2 (interleaved re, im).
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cconjmul(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// Division by a loop-invariant divisor d >= 1 with a multiply-high and a
// shift (round-up method, valid for n < 2^31): q = umulhi(n, m) >> s, d > 1.
// Division by a run-time constant d >= 1 (Granlund & Montgomery 1994, the
// round-up multiplier with an added dividend): with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, n / d = (umulhi(n, m) + n) >> l for every
// n < 2^31, d = 1 included (m = 1), without a branch.
struct FastDiv {
    uint32_t d, m, s;
};
inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    FastDiv f{d, (uint32_t)(((((1ull << l) - d) << 32) / d) + 1), l};
    return f;
}
__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv &f) {
#ifdef __HIP_DEVICE_COMPILE__
    return (__umulhi(n, f.m) + n) >> f.s;
#else
    return (uint32_t)(((((uint64_t)n * f.m) >> 32) + n) >> f.s);
#endif
}

// Mode-group bricks.  A tensor over a group of consecutive modes [m0, m1) is
// viewed as X[l, g, r]: l over the modes before the group (size L), g the
// group multi-index (first mode fastest, size M), r over the modes after the
// group (size R).  A workgroup holds an Lt x M x Rt brick in LDS (Rt > 1 only
// when Lt == L, so that the brick is contiguous), stored as
// lt + Lt * (g + M * rt).
struct ModeTask {
    const void *X;               // storage type
    void *Y;
    int64_t ll, nl;
    int nin, nout;
    const void *M;               // storage type
    int ld;
    int adj;
    FastDiv fd_ll;               // set by launch_mode_product
};

// Hole contraction: G[j1 + ld*j2] (+)= alpha * sum_{l,r} conj(X[l,j1,r]) Y[l,j2,r].
struct GramTask {
    const void *X, *Y;           // storage type; G accumulates in FP64
    int64_t ll, nl;
    int n1, n2;
    double2 *G;
    int ld;
    double2 alpha;
    int accumulate;
    int hermitian;               // 1: X == Y, result made exactly Hermitian
    FastDiv fd_ll;               // set by launch_gram
};

// Brick programs.  A task owns the bricks (Lt x M x Rt, as above) of one
// mode group of tensors that share L and R; each workgroup runs the task's
// instruction list on one brick position, with up to kMaxBrickBufs LDS
// buffers and the task's op matrices in LDS.  Instructions (buffers a, b):
//   kBLoad     a <- X            (M: group multi-index size of X)
//   kBLoadAdd  a += alpha X
//   kBOp       b <- Mat x_p a    (acc = 0) or b += Mat x_p a (acc = 1), mode p
//              of the group; ni/no: op sizes, stride/nfib: fiber geometry of a
//   kBAxpy     b += alpha a
//   kBStore    Y <- alpha a      (acc = 0) or Y += alpha a (acc = 1)
//   kBZero     a <- 0
constexpr int kMaxBrickInstr = 40;
constexpr int kMaxBrickMats = 24;
static_assert(kMaxGroupModes * (kMaxGroupModes - 1) / 2 + kMaxGroupModes - 1 <= kMaxBrickMats,
              "a mean-field Gram program over a full group must fit the matrix table");
constexpr int kMaxBrickBufs = 3;
//   kBGram     slot (mat) += sum over the brick's fibers of mode p of
//              conj(a[f, i]) b[f, j], i < ni, j < no (ni, no <= 8): partial
//              Gram/hole contractions, FP64 across bricks, one partial per
//              workgroup in the launch's partial array (see launch_brick_tasks)
constexpr int kMaxGrams = 8;
enum BrickCode { kBLoad = 0, kBLoadAdd = 1, kBOp = 2, kBAxpy = 3, kBStore = 4, kBZero = 5, kBGram = 6 };
struct BrickInstr {
    int code;
    int a, b;
    int acc;
    int mat;
    int ni, no;
    int stride, nfib;
    int M;
    FastDiv fd_stride, fd_M;
    const void *ptr;             // kBLoad / kBLoadAdd / kBStore: tensor (work type of the plan)
    double2 alpha;
};
struct BrickTask {
    int64_t L, Lt, R, Rt;
    FastDiv fd_Lt;
    int ninstr, nmat, ld;
    int ngram, gram_base;        // Gram slots used, index of the first in the launch's partials
    const void *mat[kMaxBrickMats];
    int matadj[kMaxBrickMats];
    int matnin[kMaxBrickMats];   // columns of the effective matrix
    int matnout[kMaxBrickMats];  // rows
    double2 matscale[kMaxBrickMats];
    BrickInstr ins[kMaxBrickInstr];
};


// Matrix elements between SPF blocks of one mode:
//   out[a + ld*b] = <phiA_a | O | phiB_b>,  O = product of primitive DOF ops.
// sym = 1: same-state Hermitian symmetrisation (zher2k with 1/2).
struct SpfTask {
    const double2 *phiA, *phiB;
    int nG, na, nb;
    int mode;                    // 0-based logical mode (for DOF tables)
    int ncode;
    int codes[kMaxDofPerMode];   // primitive operator per DOF of the mode
    int sym;
    double2 *out;
    int ld;
    double2 *scratch;            // 2 * nG * nb
};

// Per-DOF primitive data on the device (grid values and DVR matrices).
struct DofTables {
    const double *grid;          // [ngpmax, ndof]
    const double *d1, *d2, *xd1; // [ngpmax, ngpmax, ndof]
    int ngpmax;
    const int *ngp;              // [ndof]
    const int *mode_dofs;        // [nmodes, kMaxDofPerMode] 0-based DOF ids
    const int *mode_ndof;        // [nmodes]
    const double *oper;          // [ngpmax, nparam, ndof] parametrised operators on the grids
    int nparam;
};

// SPF propagation of one logical mode over dt (propagate_SPF + rk_SPF_ck/dp).
struct SpfModeDesc {
    int mode;
    int nG;
    int nsets;
    int off[8];                  // offsets (complex elements) of the set blocks in psi
    int loc_off[8];              // offsets of the set blocks in the mode-local block
    int nspf[8];                 // SPFs per set
    int nop;                     // operators of the mode (non-identity)
    int op_lhs_set[32], op_rhs_set[32];
    int op_codes[32][kMaxDofPerMode];
    int op_index[32];            // index in rmf
    int rho_index[8];            // index of the density matrix per set (into rho array)
    int blk;                     // complex elements of all set blocks of the mode
    int scr_off;                 // offset into the scratch pool (complex elements)
    // the distinct (primitive codes, rhs set) products of the operators,
    // each formed once per derivative; op_uq: an operator's product
    int nuq;
    int uq_set[32];
    int uq_ident[32];            // all codes zero: the SPFs themselves
    int uq_codes[32][kMaxDofPerMode];
    int op_uq[32];
    // one-DOF modes whose products are powers of x or DVR matrices: formed in
    // one pass over the nprod non-identity products (product k: uq index
    // prod_uq[k], elements [prod_off[k], prod_off[k + 1]) of the pass)
    int prod_direct;
    int nprod;
    int prod_uq[32];
    int prod_off[33];
};

// Gaussian (G-MCTDH) mode: Gaussians exp(A x^2 + xi x) per DOF with fixed
// widths A; the SPF block of a set holds xi(k, j), nmd per Gaussian.
constexpr int kMaxSets = 8;
constexpr int kMaxGwpOps = 32;
struct GwpDesc {
    int mode;                    // logical mode (0-based)
    int nmd;                     // DOFs per Gaussian
    int nsets;                   // states (multi-set) or 1
    int nspf[kMaxSets];
    int off[kMaxSets];           // psi offsets (complex elements) of the set blocks
    int rho_index[kMaxSets];     // density matrix per set (into the rho array)
    int blk;                     // complex elements of all set blocks
    int nop;                     // operators of the mode (identity included)
    int op_lhs[kMaxGwpOps], op_rhs[kMaxGwpOps];   // sets
    int op_ident[kMaxGwpOps];
    int op_codes[kMaxGwpOps][kMaxDofPerMode];
    int op_index[kMaxGwpOps];    // operator index i of h1(m, i) / mf(m, i)
    int scr_off;                 // scratch offset (complex elements)
};

struct GwpParams {
    const double *gwpa;          // [maxdof, nspfmax, nmodes]
    int maxdof;
    const double2 *rho;          // density matrices, index via rho_index
    const double2 *mf;           // [nspfmax^2 * nopmax * nmodes]
    int nspfmax, nopmax, nm, ns;
    double eps_psi, eps_C, eps_ovl, min_stepsize, dt;
    int dormand_prince;
    double2 der_factor;
    int max_attempts;
    const double *dt_src;        // as in SpfParams
    double dt_scale;
};

struct SpfParams {
    const double2 *rmf;          // [nspfmax^2 * nopmax * nmodes]
    const double2 *rho;          // density matrices, index via rho_index
    int nspfmax, nopmax;
    double eps_psi, min_stepsize, dt;
    int dormand_prince;
    double2 der_factor;
    int max_attempts;
    // when set, the interval is dt_src[0] * dt_scale, read on the device (a
    // CMF2 attempt replayed from a graph takes its step size from memory)
    const double *dt_src;
    double dt_scale;
};

#ifdef __HIPCC__
// ---------------------------------------------------------------------------
// Small dense FP64 linear algebra in one thread (n <= 16).
constexpr int kSmall = 16;

// Cyclic complex Jacobi: A (n x n, ld kSmall, Hermitian) -> eigenvalues w,
// eigenvectors U (columns).  Same rotation as the oracle's herm_eig.
__device__ inline void dev_herm_eig(int n, double2 *A, double *w, double2 *U) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) U[i + kSmall * j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                double2 a = A[i + kSmall * j];
                double a2 = a.x * a.x + a.y * a.y;
                if (i != j) off += a2; else diag += a2;
            }
        if (off <= 1e-32 * (diag + 1e-300) || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double2 apq = A[p + kSmall * q];
                double aabs = hypot(apq.x, apq.y);
                if (aabs == 0.0) continue;
                double2 ph = make_double2(apq.x / aabs, apq.y / aabs);
                double app = A[p + kSmall * p].x, aqq = A[q + kSmall * q].x;
                double theta = (aqq - app) / (2.0 * aabs);
                double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(tt * tt + 1.0), s = tt * c;
                double2 jpp = make_double2(c, 0.0), jqq = cscale(c, cconj(ph)), jpq = make_double2(s, 0.0),
                        jqp = cscale(-s, cconj(ph));
                for (int k = 0; k < n; k++) {
                    double2 akp = A[k + kSmall * p], akq = A[k + kSmall * q];
                    A[k + kSmall * p] = cadd(cmul(akp, jpp), cmul(akq, jqp));
                    A[k + kSmall * q] = cadd(cmul(akp, jpq), cmul(akq, jqq));
                }
                for (int k = 0; k < n; k++) {
                    double2 apk = A[p + kSmall * k], aqk = A[q + kSmall * k];
                    A[p + kSmall * k] = cadd(cmul(cconj(jpp), apk), cmul(cconj(jqp), aqk));
                    A[q + kSmall * k] = cadd(cmul(cconj(jpq), apk), cmul(cconj(jqq), aqk));
                }
                A[p + kSmall * q] = make_double2(0.0, 0.0);
                A[q + kSmall * p] = make_double2(0.0, 0.0);
                A[p + kSmall * p].y = 0.0;
                A[q + kSmall * q].y = 0.0;
                for (int k = 0; k < n; k++) {
                    double2 ukp = U[k + kSmall * p], ukq = U[k + kSmall * q];
                    U[k + kSmall * p] = cadd(cmul(ukp, jpp), cmul(ukq, jqp));
                    U[k + kSmall * q] = cadd(cmul(ukp, jpq), cmul(ukq, jqq));
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = A[i + kSmall * i].x;
}
#endif

// Learned step controller (mg_mlp.hip): 12 -> 64 -> 64 -> 1 perceptron.
constexpr int kMlpNFeat = 12;
constexpr int kMlpHidden = 64;
constexpr int kMlpNParam = 2 * kMlpNFeat + kMlpNFeat * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpHidden + 1;
constexpr int kMlpNWeights = kMlpNFeat * kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden;
struct MlpWork {                 // per-sample activations and gradients, batch-sized
    float *z1, *h1, *z2, *h2, *dz1, *dz2, *dy, *err2;
};
void launch_mlp_bf16(const float *p, uint16_t *wb, hipStream_t s);
void launch_mlp_forward(const float *X, const float *p, const uint16_t *wb, MlpWork w, float *y, const float *t,
                        int B, hipStream_t s);
void launch_mlp_backward_grad(const float *X, const float *p, const uint16_t *wb, MlpWork w, float *grad, int B,
                              hipStream_t s);
void launch_mlp_adam(float *p, const float *grad, float *m, float *v, float lr, int step, hipStream_t s);

// Gaussian-mode launchers (mg_gwp.hip).  dd: D matrices [nspfmax^2, nmodes, nstates].
size_t gwp_scratch_elems(const GwpDesc &d, int nspfmax);
void launch_gwp_s00(const GwpDesc *d_descs, int ndesc, const double2 *psi, GwpParams p, double2 *s00,
                    double2 *sd, hipStream_t s);
void launch_gwp_h1(const GwpDesc *d_descs, int ndesc, const double2 *psi, GwpParams p, const double2 *s00,
                   double2 *h1, hipStream_t s);
void launch_gwp_overlaps(const GwpDesc *d_descs, int ndesc, const double2 *psi1, const double2 *psi2, int sym,
                         GwpParams p, double2 *out, hipStream_t s);
void launch_gwp_dd_init(const GwpDesc *d_descs, int ndesc, const double2 *sd, GwpParams p, double2 *dd,
                        int *status, hipStream_t s);
void launch_gwp_dhd(const GwpDesc *d_descs, int ndesc, const double2 *h1, const double2 *dd, GwpParams p,
                    double2 *dhd, hipStream_t s);
void launch_gwp_ab_mats(const GwpDesc *d_descs, int ndesc, const double2 *sd, const double2 *dd, GwpParams p,
                        double2 *mats, hipStream_t s);
void launch_gwp_propagate(const GwpDesc *d_descs, int ndesc, double2 *psi, double2 *dd, double2 *scratch,
                          GwpParams p, int *status, hipStream_t s);
void launch_gwp_der(const GwpDesc *d_descs, int idesc, const double2 *psi, const double2 *dd, double2 *dpsi,
                    double2 *ddd, double2 *scratch, GwpParams p, int *status, hipStream_t s);
void launch_gwp_s00m1(const GwpDesc *d_descs, int ndesc, const double2 *sd, GwpParams p, double2 *out,
                      hipStream_t s);
void launch_gwp_state_mats(const GwpDesc *d_descs, int ndesc, const double2 *src, double scale, GwpParams p,
                           double2 *dst, hipStream_t s);
void launch_cscale(double2 *out, const double2 *x, double2 alpha, int64_t n, hipStream_t s);

// Host-side launchers (mg_kernels.hip).  prec: 0 = complex FP64 storage,
// 1 = complex FP32 storage (reductions always FP64).
// Gram launches (ngram_total > 0) run a persistent grid of `gram_blocks`
// workgroups over the bricks and write partials[(slot * 64 + e) * gram_blocks
// + block] (e = i + 8 j); launch_gram_slots_final sums them into G[slot].
// skip: when non-null and *skip is set on the device, every workgroup
// returns at once (an iteration of a device-controlled Krylov build past its
// convergence)
void launch_brick_tasks(const BrickTask *d_tasks, int ntasks, int64_t max_blocks_per_task, int nmax,
                        int brick, int nbuf, int nmat, int prec, hipStream_t s, int ngram_total = 0,
                        double2 *partials = nullptr, int gram_blocks = 0, const int *skip = nullptr,
                        int gram_task_max = kMaxGrams);
void launch_gram_slots_final(const double2 *partials, int nslots, int nblocks, double2 *const *targets,
                             int ld, hipStream_t s);
void launch_herm_batch(double2 *const *targets, const int *n, int count, int ld, hipStream_t s);
void launch_mat_combine(double2 *const *dst, const int *rowptr, const double2 *const *src, const double2 *coef,
                        const int *trans, int ndst, int ld, hipStream_t s);
void launch_mode_product(const ModeTask &t, int nmax, int prec, hipStream_t s);
void launch_gram(const GramTask &t, double2 *partials, int nblocks, int nmax, int prec, hipStream_t s);
void launch_spf_tasks(const SpfTask *d_tasks, int ntasks, DofTables dt, int max_ng, hipStream_t s);
void launch_fill(double2 *p, int64_t n, double2 v, hipStream_t s);
void launch_axpby(void *y, const void *x, double a, double b, int64_t n, int prec, hipStream_t s);
void launch_dot(const void *x, const void *y, int64_t n, double *partials, int nblocks,
                double2 *out, int prec, hipStream_t s, const int *skip = nullptr);
void launch_scale(void *x, const double *norm2, int64_t n, int prec, hipStream_t s);
void launch_combine(void *out, const void *basis, int64_t stride, const double2 *coef, int k,
                    int64_t n, int prec, hipStream_t s);
void launch_lanczos_step(void *w, const void *uj, const void *uprev, const double *dot_re, const double *nsq, int j,
                         int64_t n, double *partials, int nblocks, double *nsq_next, int prec, hipStream_t s,
                         const int *skip = nullptr);
// kdev: when non-null, the number of basis vectors is read on the device
void launch_combine_v0(double2 *out, const double2 *v0, const void *basis, int64_t stride, const double2 *coef, int k,
                       int64_t n, int prec, hipStream_t s, const int *kdev = nullptr);
void launch_norm2(const void *x, int64_t n, double *partials, int nblocks, double *out, int prec,
                  hipStream_t s);
// out = |y - x|^2 (complex FP64 vectors)
void launch_diff_norm2(const double2 *y, const double2 *x, int64_t n, double *partials, int nblocks, double *out,
                       hipStream_t s, const double *ynorm2 = nullptr);
void launch_convert(void *dst, int pdst, const void *src, int psrc, int64_t n, hipStream_t s);
void launch_copy(void *dst, const void *src, int64_t n, int prec, hipStream_t s);
// out = y0 + sum_q w[q] k[q]; y0/out complex FP64, k[q] in the storage type.
void launch_rk_combine(double2 *out, const double2 *y0, const void *const *k, const double2 *w, int nq, int64_t n,
                       int prec, hipStream_t s);
void launch_hsum(double2 *hsum, const double2 *h1, const int *rowptr, const int *src,
                 const double2 *coef, int nmat, int nmax, hipStream_t s);
int group_capacity(int brick);
void launch_rmf(const double2 *rho, const double2 *mf, double2 *rmf, const int *jobs, int njobs, int nspfmax,
                int nopmax, double eps, hipStream_t s, double *minw = nullptr);
void launch_spf_propagate(const SpfModeDesc *d_modes, int nmodes, double2 *psi, double2 *scratch,
                          SpfParams p, DofTables dt, int *d_status, hipStream_t s);
// mode < 0: every mode, -mode workgroups (pass -nmodes)
void launch_der_spf(const SpfModeDesc *d_modes, int mode, const double2 *psi, double2 *dpsi,
                    double2 *scratch, SpfParams p, DofTables dt, int *d_status, hipStream_t s);
void launch_mf_accum(double2 *dst1, double2 *dst2, const double2 *G, int n1, int n2, int ld,
                     hipStream_t s);
void launch_mat_axpy(double2 *dst, const double2 *src, double2 alpha, int n, int ld, hipStream_t s);

// ---------------------------------------------------------------------------
// Device-controlled CMF2 attempt (every decision of cmf2_base.f90:76-203 on
// the device, so that an attempt is one HIP graph).
// One Krylov build of the Lanczos integrator (cmf2_integration.f90:30-103).
struct LanczosDev {
    double alpha[kNKrylovMax], beta[kNKrylovMax], nrm[kNKrylovMax + 1];
    double error;
    int done;       // set once the error test passed (later iterations skip)
    int nk;         // Krylov size
    int overflow;   // 1: kNKrylovMax reached; 2: the enqueued iterations ran out
    double w[kNKrylovMax], U[kNKrylovMax * kNKrylovMax], ln[kNKrylovMax];
};
// Inputs (h, forced decision) and results of one attempt.
struct AttemptDev {
    double h;       // attempted step (written by the host before the launch)
    int forced;     // -1: the error test decides; 0 / 1: replayed reject / accept
    double err_A, err_psi, error;
    int accept, nk1, nk2, overflow, spf_status;
};
void launch_lan_begin(LanczosDev *L, hipStream_t s);
void launch_lan_ctrl(LanczosDev *L, const AttemptDev *at, const double2 *dot, const double *nsq, int j,
                     double dtscale, int kmax, hipStream_t s);
void launch_lan_coef(LanczosDev *L, const AttemptDev *at, double dtscale, int build, double2 *coef, hipStream_t s);
void launch_dpsi_traces(const double2 *G, int ns, int ld, double2 *o, hipStream_t s);
void launch_attempt_psi(AttemptDev *at, const double2 *o11, const double2 *o22, const double2 *o12, hipStream_t s);
void launch_attempt_decide(AttemptDev *at, const LanczosDev *L1, const LanczosDev *L2, const double *n2,
                           double eps_int, const int *spf_status, hipStream_t s);
void launch_copy_if(double2 *dst, const double2 *src, int64_t n, const int *flag, hipStream_t s,
                    const double *srcnorm2 = nullptr);

}  // namespace mg
