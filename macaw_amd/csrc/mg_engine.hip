// Host side of the MI355X G-MCTDH engine: problem layout, operator basis,
// term classification, device-resident CMF2 orchestration and the C-ABI of
// include/macaw_gpu.h.
//
// The state (A-vector, SPFs) and every work array live in HBM for the whole
// propagation; the host only issues launches and reads back the few scalars
// the control flow needs (step errors and decisions; with Gaussian modes also
// the Lanczos alpha/beta).  A grid-mode CMF2 attempt takes every decision on
// the device and runs as one HIP graph.
#include "macaw_gpu.h"
#include "mg_internal.h"

#include <algorithm>
#include <cmath>
#include <array>
#include <complex>
#include <map>
#include <random>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <memory>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

using namespace mg;
using cplx = std::complex<double>;

namespace {

struct HipError {
    std::string msg;
};

#define HIP_CHECK(x)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess)                                                                 \
            throw HipError{std::string(#x) + ": " + hipGetErrorString(e_)};                   \
    } while (0)

struct Oper {
    int lhs, rhs;                 // 1-based labels (1,1 for single-set modes)
    int codes[kMaxDofPerMode];
};

struct ChainOp {
    int mode, op;
};

struct Chain {
    int lhs, rhs;                 // 0-based states
    cplx ae;
    std::vector<ChainOp> ops;     // ascending modes, non-identity only
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    template <class T> T *as() const { return static_cast<T *>(p); }
    void alloc(size_t b) {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = b;
        if (b) HIP_CHECK(hipMalloc(&p, b));
    }
};

// A cached launch schedule: task tables uploaded once, replayed per call.
struct Launch {
    size_t off;
    int n;
    int64_t maxblocks;
    int brick;          // largest Lt * M * Rt over the launch's tasks (LDS sizing)
    int nbuf;           // LDS buffers the launch's programs use
    int nmat;           // op matrices held in LDS
    int ngram = 0;      // Gram slots (persistent grid, partials, final reduction)
    int slot_off = 0;   // first slot's entry in Plan::targets
    int gram_blocks = 0;
    int gram_task_max = 0;  // most Gram slots of one task (kernel accumulator count)
};
struct Plan {
    DevBuf buf;
    std::vector<Launch> launches;
    int prec = 0;       // storage type of the tensors the tasks touch
    DevBuf targets;     // double2 * per Gram slot
    DevBuf herm_ptrs, herm_n;   // matrices made exactly Hermitian after the plan
    int nherm = 0;
    // HIP graph of the plan's launches, captured on the second run (the
    // first one sizes the shared Gram partial buffer); `epoch` is the
    // buffer generation it was captured against
    mutable hipGraphExec_t gexec = nullptr;
    mutable int runs = 0;
    mutable int epoch = -1;
    ~Plan() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
    }
};
using PlanKey = std::array<const void *, 4>;
struct SpfPlan {
    DevBuf buf;
    int n = 0;
};

}  // namespace

struct mg_ctx {
    std::string err;
    hipStream_t stream = nullptr;
    // second stream: the SPF propagation runs beside the Lanczos / DeltaPsi
    // work it does not depend on (default priority, like the first)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int *spf_status_host = nullptr;   // pinned: status of the last SPF propagation
    // 0: complex FP64 throughout.  1: mixed form -- the A-vector, SPFs, rho
    // (Grams of the FP64 A), mean-field matrices and every reduction stay
    // FP64; the H.A chains of every Krylov product, the Krylov vectors beyond
    // the start vector and the mean-field mode-product chains are stored and
    // multiplied in complex FP32.  MG_LAN64=1 / MG_RHO64=0 / MG_MF64=1 move
    // the first Krylov product / rho / the mean fields to the other side.
    int prec = 0;
    size_t es = sizeof(double2);   // bytes per complex element of the FP32/FP64 work tensors
    // ---- problem ----
    int ns = 0, nm = 0, ndof = 0, nham = 0, maxdof = 0, ngpmax = 0;
    std::vector<int64_t> nmodeDOF, iDOF, MultiSet, nSPF, ngp, lhs, rhs, iOpDof;
    std::vector<double> grid, d1, d2, xd1;
    int nopparam = 0;
    std::vector<double> gridoper;  // [ngpmax, nopparam, ndof] parametrised operators on the grids
    std::vector<cplx> ae;
    double eps_int = 1e-8, eps_psi = 1e-8, eps_rho = 1e-8, writetime = 1.0, min_stepsize = 0.0;
    int int_type_psi = 1;
    // ---- layout ----
    int nspfmax = 0, nopmax = 0;
    std::vector<int64_t> nAConf, off, ngpm;
    std::vector<int64_t> ipsi;    // 0-based (mode + nm*state)
    int64_t nconf = 0, dimpsi = 0, mixed = 0;
    // ---- operator basis ----
    std::vector<int> nOper;
    std::vector<Oper> opers;      // [m * nopmax + i]
    std::vector<int> iOper;       // [m + nm * t], 0-based
    std::vector<char> isIdent;    // [m + nm * i]
    // ---- terms ----
    std::vector<int> hs_rowptr, hs_src;     // diagonal sums per (s, m)
    std::vector<cplx> hs_coef;
    std::vector<Chain> chains;              // H.A chain terms
    std::vector<Chain> chains_g;            // with Gaussian modes: every term, S00 for identities on them
    std::vector<int> mf1_rowptr, mf1_src;   // one-mode same-state MF = ae * rho
    std::vector<cplx> mf1_coef;
    std::vector<int> mf_chain_terms;        // terms needing AL/AR chains
    std::vector<int> mf_partner;            // per chain term: mirror term index, -1 none, -2 derived
    std::vector<int> rmf_jobs;
    int n_rmf_jobs = 0;
    size_t spf_scr_elems = 0;      // complex elements of the SPF propagation scratch (all grid modes)
    // ---- grouping ----
    int brick = 1024;
    int run_bytes = 32;            // shortest row of a strided brick (make_groups)
    std::vector<std::pair<int, int>> groups;
    // ---- device buffers ----
    DevBuf dA, dpsi, dAstart, dpsistart, dBK, dB0, dBH, dB02, dpH, dpH2, dHtmp, dIn, dOut;
    DevBuf dh1, dhsum, drho, dmf, drmf, dS, dS2, dS3, dGtmp;
    DevBuf ddpsi;                  // DeltaPsi overlap Grams: 3 sets x nstates x nspfmax^2
    DevBuf dchain;                 // 2 * nchains * mixed
    // mixed form, MG_LAN64=1: the first Krylov product H u_0 runs on the
    // FP64 start vector with FP64 matrices (chain temporaries and result in
    // FP64), so that u_1 = H u_0 - alpha u_0 carries no FP32 rounding.  Off
    // by default: over 0-71.5 fs of the 4 x 8^8 workload it moves C(t) by
    // 5.7e-8 instead of 1.9e-8 from the FP64 engine and costs 9% of the rate
    DevBuf dchain64, dW64;
    bool lan64 = false;
    // mixed form: mean fields from the FP64 A-vector with FP64 chains
    // (MG_MF64=1; off by default: FP32 chains do not move the step sequence)
    DevBuf dmfbuf64;
    bool mf64 = false;
    // mixed form: density matrices from the FP64 A-vector.  Their small
    // eigenvalues enter rho^-1 MF; Grams of an FP32 copy of A put enough
    // noise into the SPF error estimate (DeltaPsi) to flip one CMF2
    // accept/reject decision in six against the FP64 run
    bool rho64 = true;
    // MG_PHASES=1: host wall time per CMF2 phase (stream synchronised at
    // every phase boundary), printed to stderr by mg_destroy
    bool phases = false;
    double phase_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::chrono::steady_clock::time_point phase_t0;
    DevBuf dmfbuf;                 // 2 * nm * mixed (AL / AR)
    DevBuf dkron;                  // 2 * nconf-sized temps for overlap chains
    DevBuf dspfscr, dpartials, dscal;
    DevBuf dhs_rowptr, dhs_src, dhs_coef, dmf1_rowptr, dmf1_src, dmf1_coef, drmf_jobs;
    DevBuf dgrid, dd1, dd2, dxd1, dngp, dmode_dofs, dmode_ndof, dspfmodes, dstatus, dgridoper;
    // mixed form only: FP64 first Krylov vector, FP32 copy of the A-vector,
    // FP32 in/out of the C-ABI H.A, FP32 copies of h1mode and the diagonal sums
    DevBuf dAf, dXf, dYf, dh1f, dhsumf;
    // Gram-slot partial sums of brick programs; per-slot mean-field Grams;
    // combination tables turning them (and ae * rho) into the mean fields
    DevBuf dgpart, dgslots, dmf_dst, dmf_rowptr, dmf_src, dmf_coef, dmf_trans, dident;
    // VMF integration (allocated on first use): stage derivatives (H.A in the
    // work type, dpsi in FP64), 5th / 4th order candidates
    DevBuf dvA[7], dvP[7], dvAt, dvAe, dvPt, dvPe;
    bool vmf_ready = false;
    bool mf_bricks = false;        // mean fields through brick programs (square chain ops, <= 8 SPFs)
    std::vector<const void *> mf_L, mf_R;   // per term: L tensor, R_g tensors (work type)
    int nblocks_red = 2048;      // blocks of the streaming reductions and Gram partial sums
    // timing of apply_h_device
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0, bytes_alg = 0.0, flops_alg = 0.0;
    double stepsize = 0.0;         // carried between mg_propagate calls
    int inorm_psi = 0;             // propagate_wf's renormalisation mode (globals.f90:23-26)
    std::vector<mg_attempt> replay;
    size_t replay_pos = 0;
    bool prof = false;             // event pair around every group launch
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
    std::vector<hipEvent_t> prof_ev;
    // cached launch plans keyed by the buffers they read and write
    std::map<PlanKey, std::unique_ptr<Plan>> plans;
    std::map<PlanKey, std::unique_ptr<SpfPlan>> spf_plans;
    // ---- Gaussian (G-MCTDH) modes ----
    int ngwp = 0;
    std::vector<int> gwp;                  // [nm] 1 = Gaussian mode
    std::vector<int64_t> npar;             // [nm] SPF parameters per SPF (grid points or nmodeDOF)
    std::vector<char> mfskip;              // [m + nm * i] mean-field chains skip this operator
    std::vector<int> dvr_desc;             // [nm] index into dspfmodes, -1 for a Gaussian mode
    int ndvr_desc = 0, ngwp_desc = 0;
    unsigned gwp_mask = 0;                 // bit m set for a Gaussian mode
    double eps_C = 1e-8, eps_ovl = 1e-8;
    std::vector<double> gwpa;
    // S00 [nn, nm, ns, ns]; intra-state S00 per state [nn, nm, ns]; D matrices
    // of the state and of the two half-step SPF sets and of the observables;
    // the H.A operator table (h1mode, D^H h D on Gaussian modes) and its FP32
    // copy; A_to_B / B_to_A matrices; the A vector at the half step
    DevBuf dgwpdesc, dgwpa, dS00, dSd, dDD, dDH, dDH2, dDobs, dhA, dhAf, dgwpscr, dABmats, dAH;
    // VMF with Gaussian modes: identity D, Tikhonov S00^-1 and tau per (mode, state), -tau per set
    DevBuf dDid, dSm1, dTauS, dTauSet;
    // ---- HIP graphs of the cached launch plans ----
    bool use_graphs = true;
    int graph_epoch = 0;
    // ---- device-controlled CMF2 attempt (grid modes) and its HIP graph ----
    DevBuf datt;                       // AttemptDev, then two LanczosDev
    AttemptDev *att_host = nullptr;    // pinned: inputs written before, results read after an attempt
    hipGraphExec_t att_exec = nullptr; // one whole attempt, captured on the second attempt
    int att_epoch = -1, att_kmax = 0, att_runs = 0;
    int nk_seen = 2;                   // largest Krylov size of the graph-run attempts so far
    bool att_graphs = false;           // MG_ATTEMPT_GRAPH=1: the attempt as one graph (slower, see mg_set_graphs)
    int att_kmax_cap = kNKrylovMax;    // MG_ATTEMPT_KMAX (test aid): fewer iterations per graph-run build
    bool att_fork = true;              // MG_ATTEMPT_FORK=0: the SPF half steps on the engine stream
    bool dps_fused = true;             // MG_DPSI_FUSED=0: DeltaPsi as three chains and dots
    bool capturing = false;            // an attempt is being captured: plans launch their kernels
    const int *plan_skip = nullptr;    // skip flag handed to the brick launches being enqueued
    // ---- learned step controller ----
    DevBuf drmf_minw;                      // smallest rho eigenvalue per rho^-1 MF job
    DevBuf dmf_prev, dmf_tmp;              // mean fields of the previous accepted step
    bool mf_prev_valid = false;
    DevBuf dmlp_p, dmlp_wb, dmlp_x, dmlp_y, dmlp_work;
    bool mlp_on = false, log_features = false;
    std::vector<double> hist_h, hist_e;    // last attempts, most recent first
    std::vector<float> feat_log;           // 12 features per accepted step
    std::vector<double> feat_h;            // the accepted step size of each entry
    int64_t mlp_proposals = 0, mlp_fallbacks = 0;

    int nsets(int m) const { return MultiSet[m] ? ns : 1; }
    int64_t NS(int m, int s) const { return nSPF[m + (int64_t)nm * s]; }
    int64_t IP(int m, int s) const { return ipsi[m + (int64_t)nm * s]; }
    double2 *h1(int m, int op) const {
        return dh1.as<double2>() + (size_t)nspfmax * nspfmax * (op + (size_t)nopmax * m);
    }
    double2 *hsum(int s, int m) const { return dhsum.as<double2>() + (size_t)nspfmax * nspfmax * (m + (size_t)nm * s); }
    double2 *rho(int m, int s) const { return drho.as<double2>() + (size_t)nspfmax * nspfmax * (m + (size_t)nm * s); }
    double2 *rho_ave(int m) const { return drho.as<double2>() + (size_t)nspfmax * nspfmax * (nm * ns + m); }
    double2 *mf(int m, int op) const { return dmf.as<double2>() + (size_t)nspfmax * nspfmax * (op + (size_t)nopmax * m); }
    double2 *S(int m, int s) const { return dS.as<double2>() + (size_t)nspfmax * nspfmax * (m + (size_t)nm * s); }
    // H.A operator matrices: h1mode, or D^H h D on Gaussian modes
    const void *haop(int m, int op) const {
        if (!ngwp) return h1op(m, op);
        const size_t o = (size_t)nspfmax * nspfmax * (op + (size_t)nopmax * m);
        if (!prec) return dhA.as<double2>() + o;
        return dhAf.as<float2>() + o;
    }
    // operator matrices in the work-tensor storage type
    const void *h1op(int m, int op) const {
        if (!prec) return h1(m, op);
        return dh1f.as<float2>() + (size_t)nspfmax * nspfmax * (op + (size_t)nopmax * m);
    }
    const void *hsumop(int s, int m) const {
        if (!prec) return hsum(s, m);
        return dhsumf.as<float2>() + (size_t)nspfmax * nspfmax * (m + (size_t)nm * s);
    }
    // p + n elements of the work-tensor storage type
    void *at(void *p, int64_t n) const { return static_cast<char *>(p) + (size_t)n * es; }
    const void *at(const void *p, int64_t n) const { return static_cast<const char *>(p) + (size_t)n * es; }
};

namespace {

double2 d2(cplx z) { return make_double2(z.real(), z.imag()); }

// Context of the C-ABI call running on this thread.  Host<->device copies
// and fills are ordered on that context's stream and waited for, never put
// on the legacy null stream: ensemble members run on several host threads,
// and a null-stream operation would synchronise with (and disturb) the
// graph capture of another member's stream.
thread_local mg_ctx *tl_ctx = nullptr;

void hcopy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    const hipStream_t s = tl_ctx ? tl_ctx->stream : nullptr;
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

void hzero(void *dst, size_t bytes) {
    const hipStream_t s = tl_ctx ? tl_ctx->stream : nullptr;
    HIP_CHECK(hipMemsetAsync(dst, 0, bytes, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

// Kernel launches report configuration errors asynchronously; surface them.
void check_launch() { HIP_CHECK(hipGetLastError()); }

int64_t largest_divisor_le(int64_t L, int64_t cap) {
    if (cap >= L) return L;
    for (int64_t d = cap; d >= 1; d--)
        if (L % d == 0) return d;
    return 1;
}

// Mode groups for the current brick capacity (sizes: max over states): the
// first group takes leading modes while their product fits the brick; a
// later group is strided (its brick rows are Lt consecutive elements), so it
// takes at most brick / Lt_min modes' worth, Lt_min = run_bytes / element
// size (rows of at least run_bytes).
void make_groups(mg_ctx *c) {
    const int nm = c->nm, ns = c->ns;
    c->groups.clear();
    std::vector<int64_t> mx(nm, 0);
    for (int m = 0; m < nm; m++)
        for (int s = 0; s < ns; s++) mx[m] = std::max(mx[m], c->NS(m, s));
    const int64_t lt_min = std::max<int64_t>(1, c->run_bytes / (int64_t)c->es);
    int m = 0;
    while (m < nm) {
        const int64_t limit = (m == 0) ? c->brick : std::max<int64_t>(c->brick / lt_min, mx[m]);
        int64_t prod = mx[m];
        int e = m + 1;
        while (e < nm && e - m < kMaxGroupModes && prod * mx[e] <= limit) prod *= mx[e++];
        c->groups.push_back({m, e});
        m = e;
    }
}

bool brick_fits_lds(const mg_ctx *c, int b) {
    const int nmx = c->nspfmax <= 4 ? 4 : (c->nspfmax <= 8 ? 8 : 16);
    const size_t lds = c->es * ((size_t)kMaxBrickMats * nmx * nmx + kMaxBrickBufs * (size_t)group_capacity(b));
    return b >= 64 && lds <= 160 * 1024;
}

// ----------------------------------------------------------------------------
// Setup: layout (init_calc.f90:25-34,160-181), operator basis
// (hamilton.f90:91-147, 353-373) and the classification of Hamiltonian terms.
// Every size and index table of the descriptor is range-checked here, on
// the host, before anything is copied or addressed: the tables below become
// kernel addressing (state offsets, operator tables, DOF ids), so a malformed
// descriptor must end in an error message, never in a device access.
void check_descriptor(const mg_problem *p) {
    auto need = [](bool ok, const std::string &what) {
        if (!ok) throw HipError{"invalid problem descriptor: " + what};
    };
    need(p->nstates >= 1 && p->nmodes >= 1 && p->ndof >= 1 && p->nham >= 1,
         "nstates, nmodes, ndof and nham must be >= 1");
    need(p->nmodes <= 64 && p->ndof <= 256 && p->nham <= 100000, "nmodes <= 64, ndof <= 256, nham <= 100000");
    need(p->maxdof >= 1 && p->ngpmax >= 1, "maxdof and ngpmax must be >= 1");
    need(p->nmodeDOF && p->iDOF && p->MultiSet && p->nSPF && p->ngp && p->grid && p->d1 && p->d2 && p->xd1 &&
             p->ae && p->lhs_el && p->rhs_el && p->iOpDof,
         "a table pointer is null");
    need(p->nopparam >= 0 && p->nopparam <= 10000 && (p->nopparam == 0 || p->gridoper),
         "nopparam must be 0..10000, with gridoper when it is not 0");
    need(p->int_type_psi == 1 || p->int_type_psi == 2,
         "int_type_psi must be 1 (Cash-Karp) or 2 (Dormand-Prince); integrator not yet implemented for the SPFs");
    need(p->eps_int > 0 && p->eps_psi > 0 && p->eps_rho > 0 && p->writetime > 0 && p->min_stepsize >= 0,
         "eps_int, eps_psi, eps_rho and writetime must be > 0, min_stepsize >= 0");
    for (int64_t k = 0; k < p->ndof; k++)
        need(p->ngp[k] >= 1 && p->ngp[k] <= p->ngpmax, "ngp(" + std::to_string(k + 1) + ") outside 1..ngpmax");
    std::vector<int> used(p->ndof, 0);
    for (int64_t m = 0; m < p->nmodes; m++) {
        const int64_t nd = p->nmodeDOF[m];
        need(nd >= 1 && nd <= p->maxdof, "nmodeDOF(" + std::to_string(m + 1) + ") outside 1..maxdof");
        for (int64_t k = 0; k < nd; k++) {
            const int64_t d = p->iDOF[k + p->maxdof * m];
            need(d >= 1 && d <= p->ndof, "iDOF(" + std::to_string(k + 1) + "," + std::to_string(m + 1) +
                                             ") outside 1..ndof");
            need(!used[d - 1]++, "DOF " + std::to_string(d) + " belongs to two modes");
        }
        for (int64_t s = 0; s < p->nstates; s++)
            need(p->nSPF[m + p->nmodes * s] >= 1,
                 "nSPF(" + std::to_string(m + 1) + "," + std::to_string(s + 1) + ") must be >= 1");
    }
    for (int64_t t = 0; t < p->nham; t++) {
        need(p->lhs_el[t] >= 1 && p->lhs_el[t] <= p->nstates && p->rhs_el[t] >= 1 && p->rhs_el[t] <= p->nstates,
             "lhs_el / rhs_el of term " + std::to_string(t + 1) + " outside 1..nstates");
        need(std::isfinite(p->ae[2 * t]) && std::isfinite(p->ae[2 * t + 1]),
             "coefficient of term " + std::to_string(t + 1) + " is not finite");
    }
}

void setup(mg_ctx *c, const mg_problem *p) {
    check_descriptor(p);
    c->ns = (int)p->nstates;
    c->nm = (int)p->nmodes;
    c->ndof = (int)p->ndof;
    c->nham = (int)p->nham;
    c->maxdof = (int)p->maxdof;
    c->ngpmax = (int)p->ngpmax;
    const int nm = c->nm, ns = c->ns, nd = c->ndof;
    c->nmodeDOF.assign(p->nmodeDOF, p->nmodeDOF + nm);
    c->iDOF.assign(p->iDOF, p->iDOF + (size_t)c->maxdof * nm);
    c->MultiSet.assign(p->MultiSet, p->MultiSet + nm);
    c->nSPF.assign(p->nSPF, p->nSPF + (size_t)nm * ns);
    c->ngp.assign(p->ngp, p->ngp + nd);
    c->lhs.assign(p->lhs_el, p->lhs_el + c->nham);
    c->rhs.assign(p->rhs_el, p->rhs_el + c->nham);
    c->iOpDof.assign(p->iOpDof, p->iOpDof + (size_t)nd * c->nham);
    const size_t gm = (size_t)c->ngpmax;
    c->grid.assign(p->grid, p->grid + gm * nd);
    c->d1.assign(p->d1, p->d1 + gm * gm * nd);
    c->d2.assign(p->d2, p->d2 + gm * gm * nd);
    c->xd1.assign(p->xd1, p->xd1 + gm * gm * nd);
    c->ae.resize(c->nham);
    for (int t = 0; t < c->nham; t++) c->ae[t] = cplx(p->ae[2 * t], p->ae[2 * t + 1]);
    c->eps_int = p->eps_int;
    c->eps_psi = p->eps_psi;
    c->eps_rho = p->eps_rho;
    c->writetime = p->writetime;
    c->min_stepsize = p->min_stepsize;
    c->int_type_psi = (int)p->int_type_psi;
    for (int m = 0; m < nm; m++) {
        if (c->nmodeDOF[m] > kMaxDofPerMode) throw HipError{"too many DOFs in a logical mode"};
    }
    if (ns > 8) throw HipError{"at most 8 electronic states are supported"};
    // primitive operator codes: -3..20, and the parametrised operators 200 + p
    // (diagonal) and -200 - p (projector) of the descriptor's table
    c->nopparam = (int)p->nopparam;
    if (c->nopparam > 0) c->gridoper.assign(p->gridoper, p->gridoper + gm * (size_t)c->nopparam * nd);
    for (int64_t v : c->iOpDof)
        if (!((v >= -3 && v <= 20) || (v >= 201 && v <= 200 + c->nopparam) ||
              (v <= -201 && v >= -200 - c->nopparam)))
            throw HipError{"primitive operator code " + std::to_string(v) + " not supported by the GPU engine"};
    // layout
    c->nspfmax = 0;
    for (int64_t v : c->nSPF) c->nspfmax = std::max<int>(c->nspfmax, (int)v);
    if (c->nspfmax > 16) throw HipError{"more than 16 SPFs per mode are not supported"};
    c->nAConf.assign(ns, 1);
    c->off.assign(ns, 0);
    int64_t ic = 0;
    for (int s = 0; s < ns; s++) {
        for (int m = 0; m < nm; m++) c->nAConf[s] *= c->NS(m, s);
        c->off[s] = ic;
        ic += c->nAConf[s];
    }
    c->nconf = ic;
    for (int s = 0; s < ns; s++)
        if (c->nAConf[s] >= (int64_t(1) << 31)) throw HipError{"more than 2^31 configurations in one state"};
    c->mixed = 1;
    for (int m = 0; m < nm; m++) {
        int64_t mx = 0;
        for (int s = 0; s < ns; s++) mx = std::max(mx, c->NS(m, s));
        c->mixed *= mx;
    }
    c->ngpm.assign(nm, 1);
    for (int m = 0; m < nm; m++)
        for (int k = 0; k < c->nmodeDOF[m]; k++) c->ngpm[m] *= c->ngp[c->iDOF[k + (size_t)c->maxdof * m] - 1];
    // Gaussian modes (init_calc.f90:56-75): the SPF block holds nmodeDOF xi per Gaussian
    c->gwp.assign(nm, 0);
    c->npar.assign(nm, 0);
    c->ngwp = 0;
    c->gwp_mask = 0;
    for (int m = 0; m < nm; m++) {
        if (p->dof_gwp) {
            int kinds = 0;
            for (int k = 0; k < c->nmodeDOF[m]; k++) kinds += p->dof_gwp[c->iDOF[k + (size_t)c->maxdof * m] - 1] ? 1 : 0;
            if (kinds != 0 && kinds != c->nmodeDOF[m])
                throw HipError{"A mode cannot have simultaneously dvr and gwp representation"};
            c->gwp[m] = kinds ? 1 : 0;
        }
        c->npar[m] = c->gwp[m] ? c->nmodeDOF[m] : c->ngpm[m];
        if (c->gwp[m]) {
            c->ngwp++;
            if (m >= 32) throw HipError{"Gaussian modes must be among the first 32 modes"};
            c->gwp_mask |= 1u << m;
        }
    }
    for (int m = 0; m < nm; m++)
        for (int k = 0; k < c->nmodeDOF[m] && c->gwp[m]; k++)
            for (int t = 0; t < c->nham; t++) {
                const int64_t v = c->iOpDof[(c->iDOF[k + (size_t)c->maxdof * m] - 1) + (size_t)nd * t];
                if (v > 200 || v <= -200)
                    throw HipError{"parametrised operator on a Gaussian DOF: not supported by the GPU engine"};
            }
    c->eps_C = p->eps_C > 0 ? p->eps_C : 1e-8;
    c->eps_ovl = p->eps_ovl > 0 ? p->eps_ovl : 1e-8;
    if (c->ngwp) {
        if (!p->gwp_a) throw HipError{"Gaussian modes need gwp_a"};
        c->gwpa.assign(p->gwp_a, p->gwp_a + (size_t)c->maxdof * c->nspfmax * nm);
    }
    c->ipsi.assign((size_t)nm * ns, 0);
    int64_t dim = 0;
    for (int s = 0; s < ns; s++)
        for (int m = 0; m < nm; m++) {
            if (!c->MultiSet[m] && s > 0) { c->ipsi[m + (size_t)nm * s] = c->ipsi[m]; continue; }
            c->ipsi[m + (size_t)nm * s] = dim;
            dim += c->npar[m] * c->NS(m, s);
        }
    c->dimpsi = dim;
    // operator basis
    c->nOper.assign(nm, 0);
    c->iOper.assign((size_t)nm * c->nham, 0);
    std::vector<std::vector<Oper>> per(nm);
    for (int m = 0; m < nm; m++) {
        for (int t = 0; t < c->nham; t++) {
            Oper v{};
            if (c->MultiSet[m]) { v.lhs = (int)c->lhs[t]; v.rhs = (int)c->rhs[t]; }
            else { v.lhs = 1; v.rhs = 1; }
            for (int k = 0; k < c->nmodeDOF[m]; k++)
                v.codes[k] = (int)c->iOpDof[(c->iDOF[k + (size_t)c->maxdof * m] - 1) + (size_t)nd * t];
            int found = -1;
            for (size_t i = 0; i < per[m].size(); i++) {
                const Oper &u = per[m][i];
                bool same = u.lhs == v.lhs && u.rhs == v.rhs;
                for (int k = 0; k < c->nmodeDOF[m] && same; k++) same = u.codes[k] == v.codes[k];
                if (same) { found = (int)i; break; }
            }
            if (found < 0) { per[m].push_back(v); found = (int)per[m].size() - 1; }
            c->iOper[m + (size_t)nm * t] = found;
        }
        c->nOper[m] = (int)per[m].size();
        c->nopmax = std::max(c->nopmax, c->nOper[m]);
    }
    c->opers.assign((size_t)nm * c->nopmax, Oper{});
    c->isIdent.assign((size_t)nm * c->nopmax, 0);
    for (int m = 0; m < nm; m++)
        for (int i = 0; i < c->nOper[m]; i++) {
            const Oper &v = per[m][i];
            c->opers[(size_t)m * c->nopmax + i] = v;
            bool id = v.lhs == v.rhs;
            for (int k = 0; k < c->nmodeDOF[m] && id; k++) id = v.codes[k] == 0;
            c->isIdent[m + (size_t)nm * i] = id;
        }
    // the mean-field chains skip identities on grid modes only (SkipTerm,
    // hamilton.f90:386-393); on a Gaussian mode the identity is S00
    c->mfskip.assign((size_t)nm * c->nopmax, 0);
    for (int m = 0; m < nm; m++)
        for (int i = 0; i < c->nOper[m]; i++) c->mfskip[m + (size_t)nm * i] = c->isIdent[m + (size_t)nm * i] && !c->gwp[m];
    // H in the Gaussian basis (VMF derivatives): every term as a chain, the
    // identity on a Gaussian mode being S00
    c->chains_g.clear();
    if (c->ngwp)
        for (int t = 0; t < c->nham; t++) {
            std::vector<ChainOp> ops;
            for (int m = 0; m < nm; m++) {
                const int op = c->iOper[m + (size_t)nm * t];
                if (!c->mfskip[m + (size_t)nm * op]) ops.push_back({m, op});
            }
            c->chains_g.push_back({(int)c->lhs[t] - 1, (int)c->rhs[t] - 1, c->ae[t], ops});
        }
    // term classification
    std::vector<std::vector<std::pair<int, cplx>>> hs((size_t)ns * nm), mf1((size_t)nm * c->nopmax);
    c->chains.clear();
    c->mf_chain_terms.clear();
    for (int t = 0; t < c->nham; t++) {
        const int a = (int)c->lhs[t] - 1, b = (int)c->rhs[t] - 1;
        std::vector<ChainOp> ops;
        for (int m = 0; m < nm; m++) {
            const int op = c->iOper[m + (size_t)nm * t];
            if (!c->isIdent[m + (size_t)nm * op]) ops.push_back({m, op});
        }
        if (a == b && ops.size() <= 1) {
            if (ops.empty()) hs[(size_t)a * nm + 0].push_back({-1, c->ae[t]});
            else hs[(size_t)a * nm + ops[0].mode].push_back({ops[0].op + c->nopmax * ops[0].mode, c->ae[t]});
        } else {
            c->chains.push_back({a, b, c->ae[t], ops});
        }
        // mean fields
        if (ops.empty()) continue;
        if (a == b && ops.size() == 1) {
            const int m = ops[0].mode;
            mf1[(size_t)m * c->nopmax + ops[0].op].push_back({m + nm * a, c->ae[t]});
        } else {
            c->mf_chain_terms.push_back(t);
        }
    }
    // Mirror terms: t' = (b <- a) with conj(ae) and the same primitive codes
    // as t = (a <- b).  All supported primitive operators are Hermitian, so
    // h^(ba) = (h^(ab))^H on every mode and the mean-field contribution of t'
    // is the conjugate transpose of that of t (gmctdh_objects.f90:582-601
    // evaluates both; here t' is derived from t).
    c->mf_partner.assign(c->mf_chain_terms.size(), -1);
    for (size_t i = 0; i < c->mf_chain_terms.size(); i++) {
        if (c->mf_partner[i] != -1) continue;
        const int t = c->mf_chain_terms[i];
        for (size_t j = i + 1; j < c->mf_chain_terms.size(); j++) {
            if (c->mf_partner[j] != -1) continue;
            const int u = c->mf_chain_terms[j];
            if (c->lhs[u] != c->rhs[t] || c->rhs[u] != c->lhs[t] || c->lhs[t] == c->rhs[t]) continue;
            if (std::abs(c->ae[u] - std::conj(c->ae[t])) > 0.0) continue;
            bool same = true;
            for (int d = 0; d < nd && same; d++) same = c->iOpDof[d + (size_t)nd * t] == c->iOpDof[d + (size_t)nd * u];
            if (!same) continue;
            c->mf_partner[i] = u;
            c->mf_partner[j] = -2;
            break;
        }
    }
    c->hs_rowptr.assign(1, 0);
    c->hs_src.clear();
    c->hs_coef.clear();
    for (auto &row : hs) {
        for (auto &e : row) { c->hs_src.push_back(e.first); c->hs_coef.push_back(e.second); }
        c->hs_rowptr.push_back((int)c->hs_src.size());
    }
    c->mf1_rowptr.assign(1, 0);
    c->mf1_src.clear();
    c->mf1_coef.clear();
    for (auto &row : mf1) {
        for (auto &e : row) { c->mf1_src.push_back(e.first); c->mf1_coef.push_back(e.second); }
        c->mf1_rowptr.push_back((int)c->mf1_src.size());
    }
    // rho^-1 MF jobs (Calc_rhom1MF): per (mode, set), the ops with lhs == set
    c->rmf_jobs.clear();
    c->n_rmf_jobs = 0;
    for (int m = 0; m < nm; m++) {
        if (c->gwp[m]) continue;      // Calc_rhom1MF runs over the grid modes
        for (int s = 0; s < c->nsets(m); s++) {
            std::vector<int> job(36, 0);
            job[0] = m;
            job[1] = c->MultiSet[m] ? (m + nm * s) : (nm * ns + m);
            job[2] = (int)c->NS(m, s);
            int nops = 0;
            for (int i = 0; i < c->nOper[m]; i++) {
                const Oper &v = c->opers[(size_t)m * c->nopmax + i];
                if (c->isIdent[m + (size_t)nm * i]) continue;
                if (c->MultiSet[m] && v.lhs - 1 != s) continue;
                if (nops >= 16) throw HipError{"too many operators per mode and state"};
                job[4 + 2 * nops] = i;
                job[5 + 2 * nops] = (int)c->NS(m, v.rhs - 1);
                nops++;
            }
            job[3] = nops;
            c->rmf_jobs.insert(c->rmf_jobs.end(), job.begin(), job.end());
            c->n_rmf_jobs++;
        }
    }
    // mode groups (sizes: max over states); MG_BRICK overrides the LDS brick
    // capacity in complex elements (tuning).  Three bricks plus the op
    // matrices must fit the 160 KB of LDS a workgroup may address.
    c->brick = c->prec ? 2048 : 1024;
    if (const char *e = std::getenv("MG_RUN_BYTES")) {
        const int r = std::atoi(e);
        if (r >= 8 && r <= 512) c->run_bytes = r;
    }
    if (const char *e = std::getenv("MG_BRICK")) {
        // tuning only: a brick that leaves room for two buffers is accepted;
        // a program that needs three is then refused by run_plan's LDS check
        const int b = std::atoi(e);
        const int nmx = c->nspfmax <= 4 ? 4 : (c->nspfmax <= 8 ? 8 : 16);
        if (b >= 64 && c->es * ((size_t)kMaxBrickMats * nmx * nmx + 2 * (size_t)group_capacity(b)) <= 160 * 1024)
            c->brick = b;
    }
    make_groups(c);
    // mean fields through brick programs when every chain op is square and
    // the SPF counts fit one 8-wide Gram slot
    c->mf_bricks = c->nspfmax <= 8 && std::getenv("MG_MF_LEGACY") == nullptr && c->ngwp == 0;
    for (int t : c->mf_chain_terms)
        for (int m = 0; m < nm; m++)
            if (c->NS(m, (int)c->lhs[t] - 1) != c->NS(m, (int)c->rhs[t] - 1)) c->mf_bricks = false;
}

size_t mf_nterms(const mg_ctx *c) {
    size_t n = 0;
    for (size_t ti = 0; ti < c->mf_chain_terms.size(); ti++)
        if (c->mf_partner[ti] != -2) n++;
    return n;
}

// Mean-field work tensors: the legacy chains' AL/AR, or the brick path's
// R_g / L tensors (groups x chain terms).  Grows only.
void alloc_mf_tensors(mg_ctx *c) {
    size_t mfb = 2 * (size_t)c->nm;
    if (c->mf_bricks) mfb = std::max(mfb, c->groups.size() * std::max<size_t>(1, mf_nterms(c)));
    const size_t need = c->es * mfb * (size_t)c->mixed;
    if (c->dmfbuf.bytes < need) c->dmfbuf.alloc(need);
}

// ----------------------------------------------------------------------------
// Device allocation and constant tables.
void allocate(mg_ctx *c) {
    const size_t z = sizeof(double2);
    const size_t N = (size_t)c->nconf, D = (size_t)c->dimpsi;
    const size_t nn = (size_t)c->nspfmax * c->nspfmax;
    c->dA.alloc(z * N);
    c->dpsi.alloc(z * D);
    c->dAstart.alloc(z * N);
    c->dpsistart.alloc(z * D);
    c->dBK.alloc(c->es * N * kNKrylovMax);
    if (c->prec) {
        if (const char *e = std::getenv("MG_LAN64")) c->lan64 = std::atoi(e) != 0;
        if (c->lan64) c->dW64.alloc(z * N);
        if (const char *e = std::getenv("MG_MF64")) c->mf64 = std::atoi(e) != 0;
        if (const char *e = std::getenv("MG_RHO64")) c->rho64 = std::atoi(e) != 0;
        const size_t zf = sizeof(float2);
        c->dAf.alloc(zf * N);
        c->dXf.alloc(zf * N);
        c->dYf.alloc(zf * N);
        c->dh1f.alloc(zf * nn * c->nopmax * c->nm);
        c->dhsumf.alloc(zf * nn * c->nm * c->ns);
    }
    c->dB0.alloc(z * N);
    c->dBH.alloc(z * N);
    c->dB02.alloc(z * N);
    c->dHtmp.alloc(z * N);
    c->dIn.alloc(z * N);
    c->dOut.alloc(z * N);
    c->dpH.alloc(z * D);
    c->dpH2.alloc(z * D);
    c->dh1.alloc(z * nn * c->nopmax * c->nm);
    c->dhsum.alloc(z * nn * c->nm * c->ns);
    c->drho.alloc(z * nn * c->nm * (c->ns + 1));
    c->dmf.alloc(z * nn * c->nopmax * c->nm);
    c->drmf.alloc(z * nn * c->nopmax * c->nm);
    c->dS.alloc(z * nn * c->nm * c->ns);
    c->dS2.alloc(z * nn * c->nm * c->ns);
    c->dS3.alloc(z * nn * c->nm * c->ns);
    c->dGtmp.alloc(z * nn);
    c->dchain.alloc(c->es * 2 * std::max<size_t>(1, std::max(c->chains.size(), c->chains_g.size())) * (size_t)c->mixed);
    alloc_mf_tensors(c);
    if (c->mf_bricks) {
        c->dgslots.alloc(z * nn * c->nm * std::max<size_t>(1, mf_nterms(c)));
        hzero(c->dgslots.p, c->dgslots.bytes);
    }
    c->dkron.alloc(z * 2 * N);
    c->dpartials.alloc(sizeof(double2) * 64 * (size_t)c->nblocks_red);
    c->dscal.alloc(sizeof(double) * 96);   // scalars; n_j^2 at 24..44; coefficients at 48..
    HIP_CHECK(hipMemsetAsync(c->dh1.p, 0, c->dh1.bytes, c->stream));
    HIP_CHECK(hipMemsetAsync(c->dmf.p, 0, c->dmf.bytes, c->stream));
    HIP_CHECK(hipMemsetAsync(c->drmf.p, 0, c->drmf.bytes, c->stream));
    HIP_CHECK(hipMemsetAsync(c->drho.p, 0, c->drho.bytes, c->stream));
    HIP_CHECK(hipMemsetAsync(c->dS.p, 0, c->dS.bytes, c->stream));
    HIP_CHECK(hipMemsetAsync(c->dS2.p, 0, c->dS2.bytes, c->stream));
    HIP_CHECK(hipMemsetAsync(c->dS3.p, 0, c->dS3.bytes, c->stream));
    // identity h1mode for same-state identity operators (hamilton.f90:296-306)
    {
        std::vector<double2> h((size_t)nn * c->nopmax * c->nm, make_double2(0, 0));
        for (int m = 0; m < c->nm; m++)
            for (int i = 0; i < c->nOper[m]; i++)
                if (c->isIdent[m + (size_t)c->nm * i]) {
                    const int s = c->opers[(size_t)m * c->nopmax + i].lhs - 1;
                    for (int j = 0; j < c->NS(m, s); j++)
                        h[nn * (i + (size_t)c->nopmax * m) + j + (size_t)c->nspfmax * j] = make_double2(1, 0);
                }
        hcopy(c->dh1.p, h.data(), z * h.size(), hipMemcpyHostToDevice);
        if (c->prec) launch_convert(c->dh1f.p, 1, c->dh1.p, 0, (int64_t)h.size(), c->stream);
    }
    auto upload_i = [&](DevBuf &b, const std::vector<int> &v) {
        b.alloc(sizeof(int) * std::max<size_t>(1, v.size()));
        if (!v.empty()) hcopy(b.p, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice);
    };
    auto upload_c = [&](DevBuf &b, const std::vector<cplx> &v) {
        b.alloc(sizeof(double2) * std::max<size_t>(1, v.size()));
        std::vector<double2> t(v.size());
        for (size_t i = 0; i < v.size(); i++) t[i] = d2(v[i]);
        if (!t.empty()) hcopy(b.p, t.data(), sizeof(double2) * t.size(), hipMemcpyHostToDevice);
    };
    upload_i(c->dhs_rowptr, c->hs_rowptr);
    upload_i(c->dhs_src, c->hs_src);
    upload_c(c->dhs_coef, c->hs_coef);
    upload_i(c->dmf1_rowptr, c->mf1_rowptr);
    upload_i(c->dmf1_src, c->mf1_src);
    upload_c(c->dmf1_coef, c->mf1_coef);
    upload_i(c->drmf_jobs, c->rmf_jobs);
    // DOF tables
    const size_t gm = (size_t)c->ngpmax;
    c->dgrid.alloc(sizeof(double) * gm * c->ndof);
    c->dd1.alloc(sizeof(double) * gm * gm * c->ndof);
    c->dd2.alloc(sizeof(double) * gm * gm * c->ndof);
    c->dxd1.alloc(sizeof(double) * gm * gm * c->ndof);
    hcopy(c->dgrid.p, c->grid.data(), c->dgrid.bytes, hipMemcpyHostToDevice);
    hcopy(c->dd1.p, c->d1.data(), c->dd1.bytes, hipMemcpyHostToDevice);
    hcopy(c->dd2.p, c->d2.data(), c->dd2.bytes, hipMemcpyHostToDevice);
    hcopy(c->dxd1.p, c->xd1.data(), c->dxd1.bytes, hipMemcpyHostToDevice);
    c->dgridoper.alloc(sizeof(double) * std::max<size_t>(1, c->gridoper.size()));
    if (!c->gridoper.empty())
        hcopy(c->dgridoper.p, c->gridoper.data(), sizeof(double) * c->gridoper.size(), hipMemcpyHostToDevice);
    std::vector<int> ngp(c->ndof), mdofs((size_t)c->nm * kMaxDofPerMode, 0), mnd(c->nm);
    for (int k = 0; k < c->ndof; k++) ngp[k] = (int)c->ngp[k];
    for (int m = 0; m < c->nm; m++) {
        mnd[m] = (int)c->nmodeDOF[m];
        for (int k = 0; k < mnd[m]; k++) mdofs[(size_t)m * kMaxDofPerMode + k] = (int)c->iDOF[k + (size_t)c->maxdof * m] - 1;
    }
    upload_i(c->dngp, ngp);
    upload_i(c->dmode_dofs, mdofs);
    upload_i(c->dmode_ndof, mnd);
    // SPF propagation descriptors and scratch (grid modes)
    std::vector<SpfModeDesc> md;
    c->dvr_desc.assign(c->nm, -1);
    size_t scr = 0;
    for (int m = 0; m < c->nm; m++) {
        if (c->gwp[m]) continue;
        c->dvr_desc[m] = (int)md.size();
        md.emplace_back();
        SpfModeDesc &d = md.back();
        std::memset(&d, 0, sizeof d);
        d.mode = m;
        d.nG = (int)c->ngpm[m];
        d.nsets = c->nsets(m);
        if (d.nsets > kMaxSets) throw HipError{"too many electronic states for a multi-set mode (at most 8)"};
        int blk = 0;
        for (int s = 0; s < d.nsets; s++) {
            d.off[s] = (int)c->IP(m, s);
            d.loc_off[s] = blk;
            d.nspf[s] = (int)c->NS(m, s);
            d.rho_index[s] = c->MultiSet[m] ? (m + c->nm * s) : (c->nm * c->ns + m);
            blk += d.nG * d.nspf[s];
        }
        d.blk = blk;
        int nop = 0;
        for (int i = 0; i < c->nOper[m]; i++) {
            if (c->isIdent[m + (size_t)c->nm * i]) continue;
            if (nop >= 32) throw HipError{"too many operators for one mode"};
            const Oper &v = c->opers[(size_t)m * c->nopmax + i];
            d.op_lhs_set[nop] = c->MultiSet[m] ? v.lhs - 1 : 0;
            d.op_rhs_set[nop] = c->MultiSet[m] ? v.rhs - 1 : 0;
            for (int k = 0; k < kMaxDofPerMode; k++) d.op_codes[nop][k] = v.codes[k];
            d.op_index[nop] = i + c->nopmax * m;
            nop++;
        }
        d.nop = nop;
        d.nuq = 0;
        for (int q = 0; q < nop; q++) {
            int u = 0;
            for (; u < d.nuq; u++) {
                bool same = d.uq_set[u] == d.op_rhs_set[q];
                for (int k = 0; k < kMaxDofPerMode; k++) same = same && d.uq_codes[u][k] == d.op_codes[q][k];
                if (same) break;
            }
            if (u == d.nuq) {
                d.uq_set[u] = d.op_rhs_set[q];
                d.uq_ident[u] = 1;
                for (int k = 0; k < kMaxDofPerMode; k++) {
                    d.uq_codes[u][k] = d.op_codes[q][k];
                    if (d.op_codes[q][k] != 0) d.uq_ident[u] = 0;
                }
                d.nuq++;
            }
            d.op_uq[q] = u;
        }
        d.prod_direct = c->nmodeDOF[m] == 1 ? 1 : 0;
        d.nprod = 0;
        d.prod_off[0] = 0;
        for (int u = 0; u < d.nuq; u++) {
            if (d.uq_ident[u]) continue;
            const int code = d.uq_codes[u][0];
            if (!((code >= 1 && code <= 200) || code == -1 || code == -2 || code == -3)) d.prod_direct = 0;
            d.prod_uq[d.nprod] = u;
            d.prod_off[d.nprod + 1] = d.prod_off[d.nprod] + d.nG * d.nspf[d.uq_set[u]];
            d.nprod++;
        }
        d.scr_off = (int)scr;
        scr += (size_t)10 * blk + (size_t)2 * d.nG * c->nspfmax + (size_t)2 * kMaxSets * nn +
               (size_t)d.nuq * d.nG * c->nspfmax;
    }
    c->spf_scr_elems = scr;
    c->ndvr_desc = (int)md.size();
    c->dspfmodes.alloc(sizeof(SpfModeDesc) * std::max<size_t>(1, md.size()));
    if (!md.empty()) hcopy(c->dspfmodes.p, md.data(), sizeof(SpfModeDesc) * md.size(), hipMemcpyHostToDevice);
    c->dspfscr.alloc(z * scr);
    c->dstatus.alloc(sizeof(double2) * 2);
    // Gaussian modes: descriptors, widths, overlaps, D matrices, scratch
    if (c->ngwp) {
        std::vector<GwpDesc> gd;
        size_t gscr = 0;
        for (int m = 0; m < c->nm; m++) {
            if (!c->gwp[m]) continue;
            gd.emplace_back();
            GwpDesc &d = gd.back();
            std::memset(&d, 0, sizeof d);
            d.mode = m;
            d.nmd = (int)c->nmodeDOF[m];
            d.nsets = c->nsets(m);
            if (d.nsets > kMaxSets) throw HipError{"too many electronic states for a Gaussian mode"};
            int blk = 0;
            for (int st = 0; st < d.nsets; st++) {
                d.nspf[st] = (int)c->NS(m, st);
                d.off[st] = (int)c->IP(m, st);
                d.rho_index[st] = c->MultiSet[m] ? (m + c->nm * st) : (c->nm * c->ns + m);
                if (d.nspf[st] * d.nmd > kSmall) throw HipError{"Gaussians x DOFs of a mode exceed 16"};
                blk += d.nmd * d.nspf[st];
            }
            d.blk = blk;
            if (c->nOper[m] > kMaxGwpOps) throw HipError{"too many operators for a Gaussian mode"};
            d.nop = c->nOper[m];
            for (int i = 0; i < c->nOper[m]; i++) {
                const Oper &v = c->opers[(size_t)m * c->nopmax + i];
                d.op_lhs[i] = c->MultiSet[m] ? v.lhs - 1 : 0;
                d.op_rhs[i] = c->MultiSet[m] ? v.rhs - 1 : 0;
                d.op_ident[i] = c->isIdent[m + (size_t)c->nm * i] ? 1 : 0;
                for (int k = 0; k < kMaxDofPerMode; k++) d.op_codes[i][k] = v.codes[k];
                d.op_index[i] = i + c->nopmax * m;
            }
            d.scr_off = (int)gscr;
            gscr += gwp_scratch_elems(d, c->nspfmax);
        }
        c->ngwp_desc = (int)gd.size();
        c->dgwpdesc.alloc(sizeof(GwpDesc) * gd.size());
        hcopy(c->dgwpdesc.p, gd.data(), sizeof(GwpDesc) * gd.size(), hipMemcpyHostToDevice);
        c->dgwpscr.alloc(z * gscr);
        c->dgwpa.alloc(sizeof(double) * c->gwpa.size());
        hcopy(c->dgwpa.p, c->gwpa.data(), sizeof(double) * c->gwpa.size(), hipMemcpyHostToDevice);
        const size_t nms = nn * c->nm * c->ns;
        c->dS00.alloc(z * nms * c->ns);
        c->dSd.alloc(z * nms);
        c->dDD.alloc(z * nms);
        c->dDH.alloc(z * nms);
        c->dDH2.alloc(z * nms);
        c->dDobs.alloc(z * nms);
        c->dABmats.alloc(z * 2 * nms);
        hzero(c->dS00.p, c->dS00.bytes);
        hzero(c->dSd.p, c->dSd.bytes);
        hzero(c->dDD.p, c->dDD.bytes);
        c->dSm1.alloc(z * nms);
        c->dTauS.alloc(z * nms);
        c->dTauSet.alloc(z * nms);
        {
            std::vector<double2> I(nms, make_double2(0, 0));
            for (int m = 0; m < c->nm; m++)
                for (int st = 0; st < c->ns; st++)
                    for (int j = 0; j < c->nspfmax; j++)
                        I[nn * (m + (size_t)c->nm * st) + j + (size_t)c->nspfmax * j] = make_double2(1, 0);
            c->dDid.alloc(z * nms);
            hcopy(c->dDid.p, I.data(), z * nms, hipMemcpyHostToDevice);
        }
        c->dhA.alloc(z * nn * c->nopmax * c->nm);
        if (c->prec) c->dhAf.alloc(sizeof(float2) * nn * c->nopmax * c->nm);
        c->dAH.alloc(z * (size_t)c->nconf);
    }
    c->drmf_minw.alloc(sizeof(double) * std::max(1, c->n_rmf_jobs));
    c->dmf_prev.alloc(c->dmf.bytes);
    c->dmf_tmp.alloc(c->dmf.bytes);
    // SPF product task scratch: 2 * nG * nspfmax per task
    int64_t maxng = 0;
    for (int m = 0; m < c->nm; m++)
        if (!c->gwp[m]) maxng = std::max(maxng, c->ngpm[m]);
    HIP_CHECK(hipEventCreate(&c->ev0));
    HIP_CHECK(hipEventCreate(&c->ev1));
    // one scratch region per possible SPF task
    const size_t ntask_max = (size_t)c->nm * std::max(c->nopmax, c->ns) * 2 + 1;
    c->dspfscr.alloc(z * std::max(scr, (size_t)0) + z * ntask_max * 2 * (size_t)maxng * c->nspfmax);
    HIP_CHECK(hipStreamSynchronize(c->stream));
}

DofTables dof_tables(const mg_ctx *c) {
    DofTables t;
    t.grid = c->dgrid.as<double>();
    t.d1 = c->dd1.as<double>();
    t.d2 = c->dd2.as<double>();
    t.xd1 = c->dxd1.as<double>();
    t.ngpmax = c->ngpmax;
    t.ngp = c->dngp.as<int>();
    t.mode_dofs = c->dmode_dofs.as<int>();
    t.mode_ndof = c->dmode_ndof.as<int>();
    t.oper = c->dgridoper.as<double>();
    t.nparam = c->nopparam;
    return t;
}

size_t spf_scratch_base(const mg_ctx *c) {
    // SPF propagation scratch comes first (every grid mode's region, as the
    // descriptors lay it out); SPF-product task scratch after it
    return c->spf_scr_elems;
}

// ----------------------------------------------------------------------------
// Brick programs (host side).
int nmax_template(const mg_ctx *c) { return c->nspfmax <= 4 ? 4 : (c->nspfmax <= 8 ? 8 : 16); }

// Builds one BrickTask over mode group g for tensors sharing L (modes before
// the group) and R (modes after).  Each buffer carries the group-mode sizes
// of the data it holds; ops check and update them.
struct TaskBuilder {
    const mg_ctx *c;
    int m0, m1;
    BrickTask t;
    std::vector<int64_t> shape[kMaxBrickBufs];
    int nbuf = 0;
    int64_t maxM = 1;
    std::vector<double2 *> targets;   // per Gram slot

    TaskBuilder(const mg_ctx *cc, int g, int64_t L, int64_t R) : c(cc) {
        std::memset(&t, 0, sizeof t);
        m0 = c->groups[g].first;
        m1 = c->groups[g].second;
        // Mmax: the largest group multi-index any state's tensor can have
        int64_t Mmax = 1;
        for (int m = m0; m < m1; m++) {
            int64_t mx = 0;
            for (int s = 0; s < c->ns; s++) mx = std::max(mx, c->NS(m, s));
            Mmax *= mx;
        }
        t.L = L;
        t.R = R;
        t.Lt = largest_divisor_le(L, std::max<int64_t>(1, c->brick / Mmax));
        if (t.Lt * Mmax > c->brick) throw HipError{"brick too small for a mode group"};
        // when the whole leading block fits, give each workgroup Rt consecutive
        // outer indices (contiguous too)
        t.Rt = (t.Lt == L) ? largest_divisor_le(R, std::max<int64_t>(1, c->brick / (L * Mmax))) : 1;
        t.fd_Lt = make_fastdiv((uint32_t)t.Lt);
        t.ld = c->nspfmax;
    }
    static int64_t prod(const std::vector<int64_t> &v) {
        int64_t p = 1;
        for (auto x : v) p *= x;
        return p;
    }
    BrickInstr &emit(int code) {
        if (t.ninstr >= kMaxBrickInstr) throw HipError{"brick program too long"};
        BrickInstr &I = t.ins[t.ninstr++];
        std::memset(&I, 0, sizeof I);
        I.code = code;
        I.alpha = make_double2(1, 0);
        return I;
    }
    void use(int buf) {
        if (buf >= kMaxBrickBufs) throw HipError{"too many brick buffers"};
        nbuf = std::max(nbuf, buf + 1);
    }
    void set_shape(int buf, const std::vector<int64_t> &sh) {
        use(buf);
        shape[buf] = sh;
        maxM = std::max(maxM, prod(sh));
    }
    // group part of full mode sizes
    std::vector<int64_t> group_sizes(const std::vector<int64_t> &sz) const {
        return std::vector<int64_t>(sz.begin() + m0, sz.begin() + m1);
    }
    void load(int buf, const void *X, const std::vector<int64_t> &sz, int add = 0, cplx alpha = 1.0) {
        const std::vector<int64_t> sh = group_sizes(sz);
        if (add && shape[buf] != sh) throw HipError{"brick load-add shape mismatch"};
        set_shape(buf, sh);
        BrickInstr &I = emit(add ? kBLoadAdd : kBLoad);
        I.a = buf;
        I.ptr = X;
        I.M = (int)prod(sh);
        I.fd_M = make_fastdiv((uint32_t)I.M);
        I.alpha = d2(alpha);
    }
    void zero(int buf, const std::vector<int64_t> &sz) {
        set_shape(buf, group_sizes(sz));
        BrickInstr &I = emit(kBZero);
        I.a = buf;
        I.M = (int)prod(shape[buf]);
    }
    // dst (<-|+=) Mat x_mode src; Mat: nout x nin, stored with leading
    // dimension nspfmax (adj: use its conjugate transpose), times scale
    void op(int src, int dst, int mode, const void *mat, int adj, int nout, int acc, cplx scale = 1.0) {
        if (mode < m0 || mode >= m1) throw HipError{"op outside the brick's mode group"};
        if (t.nmat >= kMaxBrickMats)
            throw HipError{"too many op matrices in a brick program (modes " + std::to_string(m0) + ".." +
                           std::to_string(m1 - 1) + ", op on mode " + std::to_string(mode) + ", " +
                           std::to_string(t.ninstr) + " instructions so far)"};
        const int p = mode - m0;
        std::vector<int64_t> sh = shape[src];
        const int ni = (int)sh[p];
        int64_t st = t.Lt;
        for (int q = 0; q < p; q++) st *= sh[q];
        const int64_t nfib = t.Lt * prod(sh) * t.Rt / ni;
        if (src == dst && ni != nout) throw HipError{"in-place op must be square"};
        sh[p] = nout;
        if (acc) {
            if (shape[dst] != sh) throw HipError{"accumulating op shape mismatch"};
        } else {
            set_shape(dst, sh);
        }
        const int q = t.nmat++;
        t.mat[q] = mat;
        t.matadj[q] = adj;
        t.matnin[q] = ni;
        t.matnout[q] = nout;
        t.matscale[q] = d2(scale);
        BrickInstr &I = emit(kBOp);
        I.a = src;
        I.b = dst;
        I.acc = acc;
        I.mat = q;
        I.ni = ni;
        I.no = nout;
        I.stride = (int)st;
        I.fd_stride = make_fastdiv((uint32_t)st);
        I.nfib = (int)nfib;
    }
    void axpy(int src, int dst, cplx alpha) {
        if (shape[src] != shape[dst]) throw HipError{"brick axpy shape mismatch"};
        BrickInstr &I = emit(kBAxpy);
        I.a = src;
        I.b = dst;
        I.M = (int)prod(shape[src]);
        I.alpha = d2(alpha);
    }
    void store(int buf, void *Y, int acc, cplx alpha = 1.0) {
        BrickInstr &I = emit(kBStore);
        I.a = buf;
        I.ptr = Y;
        I.acc = acc;
        I.M = (int)prod(shape[buf]);
        I.fd_M = make_fastdiv((uint32_t)I.M);
        I.alpha = d2(alpha);
    }
    // Gram slot += sum over fibers of `mode` of conj(buffer la) x (buffer rb);
    // the buffers agree on every group mode but `mode` (sizes <= 8 there)
    void gram(int la, int rb, int mode, double2 *target) {
        if (t.ngram >= kMaxGrams) throw HipError{"too many Gram slots in a brick program"};
        const int p = mode - m0;
        const std::vector<int64_t> &sa = shape[la], &sb = shape[rb];
        for (size_t q = 0; q < sa.size(); q++)
            if ((int)q != p && sa[q] != sb[q]) throw HipError{"Gram operand shapes differ"};
        if (sa[p] > 8 || sb[p] > 8) throw HipError{"Gram slot wider than 8"};
        int64_t st = t.Lt;
        for (int q = 0; q < p; q++) st *= sa[q];
        BrickInstr &I = emit(kBGram);
        I.a = la;
        I.b = rb;
        I.mat = t.ngram++;
        I.ni = (int)sa[p];
        I.no = (int)sb[p];
        I.stride = (int)st;
        I.fd_stride = make_fastdiv((uint32_t)st);
        I.nfib = (int)(t.Lt * prod(sa) * t.Rt / sa[p]);
        targets.push_back(target);
    }
    int brick() const { return (int)(t.Lt * maxM * t.Rt); }
};

// Collects the launches of one operation; phases run in order on the stream.
struct PlanBuilder {
    std::vector<BrickTask> tasks;
    std::vector<Launch> launches;
    std::vector<double2 *> targets;
    std::vector<double2 *> herm;
    std::vector<int> herm_n;
    void phase(std::vector<TaskBuilder> ts) {
        if (ts.empty()) return;
        int64_t mx = 0;
        int br = 1, nbuf = 1, nmat = 0, ngram = 0, gtask = 0;
        const int slot_off = (int)targets.size();
        for (auto &tb : ts) {
            mx = std::max(mx, (tb.t.L / tb.t.Lt) * (tb.t.R / tb.t.Rt));
            br = std::max(br, tb.brick());
            nbuf = std::max(nbuf, tb.nbuf);
            nmat = std::max(nmat, tb.t.nmat);
            tb.t.gram_base = ngram;
            ngram += tb.t.ngram;
            gtask = std::max(gtask, tb.t.ngram);
            targets.insert(targets.end(), tb.targets.begin(), tb.targets.end());
        }
        Launch l{tasks.size(), (int)ts.size(), mx, br, nbuf, nmat};
        l.ngram = ngram;
        l.slot_off = slot_off;
        l.gram_blocks = (int)std::min<int64_t>(512, mx);
        l.gram_task_max = gtask;
        launches.push_back(l);
        for (auto &tb : ts) tasks.push_back(tb.t);
    }
    void hermitian(double2 *G, int n) {
        herm.push_back(G);
        herm_n.push_back(n);
    }
    std::unique_ptr<Plan> finish(int prec) {
        auto p = std::make_unique<Plan>();
        p->prec = prec;
        p->buf.alloc(sizeof(BrickTask) * std::max<size_t>(1, tasks.size()));
        if (!tasks.empty())
            hcopy(p->buf.p, tasks.data(), sizeof(BrickTask) * tasks.size(), hipMemcpyHostToDevice);
        p->launches = launches;
        p->targets.alloc(sizeof(double2 *) * std::max<size_t>(1, targets.size()));
        if (!targets.empty())
            hcopy(p->targets.p, targets.data(), sizeof(double2 *) * targets.size(),
                                hipMemcpyHostToDevice);
        p->nherm = (int)herm.size();
        p->herm_ptrs.alloc(sizeof(double2 *) * std::max<size_t>(1, herm.size()));
        p->herm_n.alloc(sizeof(int) * std::max<size_t>(1, herm.size()));
        if (!herm.empty()) {
            hcopy(p->herm_ptrs.p, herm.data(), sizeof(double2 *) * herm.size(), hipMemcpyHostToDevice);
            hcopy(p->herm_n.p, herm_n.data(), sizeof(int) * herm.size(), hipMemcpyHostToDevice);
        }
        return p;
    }
};

void launch_plan_entry(mg_ctx *c, const Plan &p, const Launch &l, int nmax) {
    const BrickTask *base = p.buf.as<BrickTask>();
    if (l.ngram > 0) {
        const size_t need = sizeof(double2) * 64 * (size_t)l.ngram * (size_t)l.gram_blocks;
        if (c->dgpart.bytes < need) {
            HIP_CHECK(hipStreamSynchronize(c->stream));
            c->dgpart.alloc(need);
            c->graph_epoch++;          // graphs captured against the old buffer are stale
        }
        launch_brick_tasks(base + l.off, l.n, l.maxblocks, nmax, l.brick, l.nbuf, l.nmat, p.prec, c->stream, l.ngram,
                           c->dgpart.as<double2>(), l.gram_blocks, nullptr, l.gram_task_max);
        launch_gram_slots_final(c->dgpart.as<double2>(), l.ngram, l.gram_blocks,
                                p.targets.as<double2 *>() + l.slot_off, c->nspfmax, c->stream);
    } else {
        launch_brick_tasks(base + l.off, l.n, l.maxblocks, nmax, l.brick, l.nbuf, l.nmat, p.prec, c->stream, 0,
                           nullptr, 0, c->plan_skip);
    }
}

void run_plan(mg_ctx *c, const Plan &p) {
    const BrickTask *base = p.buf.as<BrickTask>();
    const int nmax = nmax_template(c);
    const size_t es = p.prec ? sizeof(float2) : sizeof(double2);
    for (auto &l : p.launches)
        if (es * ((size_t)l.nmat * nmax * nmax + (size_t)l.nbuf * group_capacity(l.brick)) > 160 * 1024)
            throw HipError{"brick program exceeds the 160 KB of LDS"};
    (void)base;
    static const bool verbose_all = std::getenv("MG_VERBOSE") != nullptr;
    if (!c->prof && !verbose_all) {
        // a plan graph holds its launches' arguments: plans enqueued with a
        // skip flag or into a captured attempt launch their kernels directly
        if (c->use_graphs && p.runs >= 1 && !c->capturing && !c->plan_skip) {
            if (p.gexec && p.epoch != c->graph_epoch) {
                HIP_CHECK(hipGraphExecDestroy(p.gexec));
                p.gexec = nullptr;
            }
            if (!p.gexec) {
                hipGraph_t g = nullptr;
                HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                for (auto &l : p.launches) launch_plan_entry(c, p, l, nmax);
                if (p.nherm)
                    launch_herm_batch(p.herm_ptrs.as<double2 *>(), p.herm_n.as<int>(), p.nherm, c->nspfmax, c->stream);
                HIP_CHECK(hipStreamEndCapture(c->stream, &g));
                HIP_CHECK(hipGraphInstantiate(&p.gexec, g, nullptr, nullptr, 0));
                HIP_CHECK(hipGraphDestroy(g));
                p.epoch = c->graph_epoch;
            }
            HIP_CHECK(hipGraphLaunch(p.gexec, c->stream));
            return;
        }
        p.runs++;
        for (auto &l : p.launches) launch_plan_entry(c, p, l, nmax);
        if (p.nherm) launch_herm_batch(p.herm_ptrs.as<double2 *>(), p.herm_n.as<int>(), p.nherm, c->nspfmax, c->stream);
        return;
    }
    while (c->prof_ev.size() < 2 * p.launches.size()) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        c->prof_ev.push_back(e);
    }
    for (size_t i = 0; i < p.launches.size(); i++) {
        const Launch &l = p.launches[i];
        HIP_CHECK(hipEventRecord(c->prof_ev[2 * i], c->stream));
        launch_plan_entry(c, p, l, nmax);
        HIP_CHECK(hipEventRecord(c->prof_ev[2 * i + 1], c->stream));
    }
    if (p.nherm) launch_herm_batch(p.herm_ptrs.as<double2 *>(), p.herm_n.as<int>(), p.nherm, c->nspfmax, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    static const bool verbose = std::getenv("MG_VERBOSE") != nullptr;
    for (size_t i = 0; i < p.launches.size(); i++) {
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, c->prof_ev[2 * i], c->prof_ev[2 * i + 1]));
        c->prof_ms += ms;
        c->prof_launches++;
        if (verbose)
            fprintf(stderr, "  launch %zu: tasks=%d blocks=%lld brick=%d nbuf=%d nmat=%d ngram=%d  %.3f ms\n", i,
                    p.launches[i].n, (long long)p.launches[i].maxblocks, p.launches[i].brick, p.launches[i].nbuf,
                    p.launches[i].nmat, p.launches[i].ngram, ms);
    }
}

int64_t prod_range(const std::vector<int64_t> &sz, int a, int b) {
    int64_t p = 1;
    for (int m = a; m < b; m++) p *= sz[m];
    return p;
}

// Calc_DHD_B (cmf2_derivatives.f90:1-80): HB_s = sum_k (h_k on mode k) B_s
// (the diagonal sums Hsum) + sum_chains ae (x_k O_k) B_rhs.  Schedule over
// the mode groups g = 0..G-1:
//   first  (g = 0, G > 1), one task per input state b: load B_b once, write
//          the group-0 diagonal sum to HB_b and every chain out of b through
//          its group-0 ops to its temporary;
//   middle (0 < g < G-1): diagonal tasks (HB_s += group-g sum of B_s) and
//          chain tasks (temporaries through their group-g ops, in place when
//          square);
//   final  (g = G-1), one task per output state s: HB_s partial + last-group
//          sum of B_s + every chain into s through its last-group ops (ae
//          folded into its last matrix), written once.
// With one group only the final kind runs, reading the chains' inputs.
// gbasis: H in the Gaussian (non-orthogonal) basis -- chains_g, h1mode on
// every Gaussian mode, no diagonal sums (VMF with Gaussian modes).
// f64: complex FP64 storage and matrices in a mixed-form context (the first
// Krylov product of the mixed Lanczos, whose input is the FP64 start vector).
std::unique_ptr<Plan> build_apply_h(mg_ctx *c, const void *B, void *HB, bool gbasis = false, bool f64 = false) {
    PlanBuilder pb;
    const std::vector<Chain> &chains = gbasis ? c->chains_g : c->chains;
    const bool wide = f64 && c->prec;
    const int sprec = wide ? 0 : c->prec;
    const size_t esz = sprec ? sizeof(float2) : sizeof(double2);
    auto opm = [&](int m, int op) -> const void * {
        if (wide) {
            if (gbasis || !c->ngwp) return c->h1(m, op);
            return c->dhA.as<double2>() + (size_t)c->nspfmax * c->nspfmax * (op + (size_t)c->nopmax * m);
        }
        return gbasis ? c->h1op(m, op) : c->haop(m, op);
    };
    auto hsumop = [&](int s, int m) -> const void * { return wide ? (const void *)c->hsum(s, m) : c->hsumop(s, m); };
    auto at = [&](const void *p, int64_t n) { return static_cast<void *>(static_cast<char *>(const_cast<void *>(p)) + (size_t)n * esz); };
    const int G = (int)c->groups.size();
    const int nm = c->nm;
    const size_t mixed = (size_t)c->mixed;
    const size_t K = chains.size();
    if (wide && c->dchain64.bytes < sizeof(double2) * 2 * std::max<size_t>(1, K) * mixed)
        c->dchain64.alloc(sizeof(double2) * 2 * std::max<size_t>(1, K) * mixed);
    void *chainbuf = wide ? c->dchain64.p : c->dchain.p;
    auto Tbuf = [&](size_t k, int i) { return at(chainbuf, (int64_t)(mixed * (2 * k + i))); };
    auto state_sizes = [&](int s) {
        std::vector<int64_t> sz(nm);
        for (int m = 0; m < nm; m++) sz[m] = c->NS(m, s);
        return sz;
    };
    auto diag_ops = [&](int s, int g) {
        std::vector<int> ms;
        for (int m = c->groups[g].first; m < c->groups[g].second; m++) {
            const int row = s * nm + m;
            if (!gbasis && c->hs_rowptr[row + 1] != c->hs_rowptr[row]) ms.push_back(m);
        }
        return ms;
    };
    auto chain_ops = [&](const Chain &ch, int g) {
        std::vector<ChainOp> v;
        for (auto &o : ch.ops)
            if (o.mode >= c->groups[g].first && o.mode < c->groups[g].second) v.push_back(o);
        return v;
    };
    std::vector<const void *> cur(K);
    std::vector<int> curT(K, -1);
    std::vector<std::vector<int64_t>> csz(K);
    for (size_t k = 0; k < K; k++) {
        cur[k] = at(B, c->off[chains[k].rhs]);
        csz[k] = state_sizes(chains[k].rhs);
    }
    // Apply chain k's ops of group g to the data in buffer `in`; the result is
    // left in the returned buffer (in place when square, else ping-pong with
    // `spare`).  last_to: if >= 0, the last op accumulates into that buffer
    // with `scale` folded into its matrix.
    auto chain_group_ops = [&](TaskBuilder &tb, size_t k, int g, int in, int spare, int last_to, cplx scale) {
        const Chain &ch = chains[k];
        std::vector<ChainOp> ops = chain_ops(ch, g);
        int curb = in;
        for (size_t i = 0; i < ops.size(); i++) {
            const ChainOp &o = ops[i];
            const int nout = (int)c->NS(o.mode, ch.lhs);
            const bool last = (i + 1 == ops.size());
            if (last && last_to >= 0) {
                tb.op(curb, last_to, o.mode, opm(o.mode, o.op), 0, nout, 1, scale);
                curb = last_to;
            } else {
                const int ni = (int)tb.shape[curb][o.mode - tb.m0];
                const int dst = (ni == nout) ? curb : (curb == in ? spare : in);
                tb.op(curb, dst, o.mode, opm(o.mode, o.op), 0, nout, 0);
                curb = dst;
            }
            csz[k][o.mode] = nout;
        }
        return curb;
    };

    if (G >= 2) {
        // first pass, per input state; the chains out of a state continue in
        // further programs of the same launch (each re-reading B_b) when
        // they outgrow one program's matrix or instruction table (multi-set
        // inter-state terms carry an overlap matrix on every mode)
        std::vector<TaskBuilder> first;
        for (int b = 0; b < c->ns; b++) {
            const std::vector<int64_t> sz = state_sizes(b);
            const int64_t Rb = prod_range(sz, c->groups[0].second, nm);
            TaskBuilder tb(c, 0, 1, Rb);
            tb.load(0, at(B, c->off[b]), sz);
            const std::vector<int> dm = diag_ops(b, 0);
            if (dm.empty()) {
                tb.zero(1, sz);
            } else {
                for (size_t i = 0; i < dm.size(); i++)
                    tb.op(0, 1, dm[i], hsumop(b, dm[i]), 0, (int)sz[dm[i]], i > 0 ? 1 : 0);
            }
            tb.store(1, at(HB, c->off[b]), 0);
            for (size_t k = 0; k < K; k++) {
                const Chain &ch = chains[k];
                if (ch.rhs != b || chain_ops(ch, 0).empty()) continue;
                // first op reads the untouched input brick (buffer 0)
                const std::vector<ChainOp> ops = chain_ops(ch, 0);
                if (tb.t.nmat + (int)ops.size() > kMaxBrickMats || tb.t.ninstr + (int)ops.size() + 1 > kMaxBrickInstr) {
                    first.push_back(tb);
                    tb = TaskBuilder(c, 0, 1, Rb);
                    tb.load(0, at(B, c->off[b]), sz);
                }
                const ChainOp &o = ops[0];
                tb.op(0, 1, o.mode, opm(o.mode, o.op), 0, (int)c->NS(o.mode, ch.lhs), 0);
                csz[k][o.mode] = c->NS(o.mode, ch.lhs);
                int curb = 1;
                for (size_t i = 1; i < ops.size(); i++) {
                    const int nout = (int)c->NS(ops[i].mode, ch.lhs);
                    const int ni = (int)tb.shape[curb][ops[i].mode - tb.m0];
                    const int dst = (ni == nout) ? curb : (curb == 1 ? 2 : 1);
                    tb.op(curb, dst, ops[i].mode, opm(ops[i].mode, ops[i].op), 0, nout, 0);
                    csz[k][ops[i].mode] = nout;
                    curb = dst;
                }
                tb.store(curb, Tbuf(k, 0), 0);
                cur[k] = Tbuf(k, 0);
                curT[k] = 0;
            }
            first.push_back(tb);
        }
        pb.phase(first);
        // middle passes
        for (int g = 1; g + 1 < G; g++) {
            std::vector<TaskBuilder> diag, chn;
            for (int s = 0; s < c->ns; s++) {
                const std::vector<int> dm = diag_ops(s, g);
                if (dm.empty()) continue;
                const std::vector<int64_t> sz = state_sizes(s);
                TaskBuilder tb(c, g, prod_range(sz, 0, c->groups[g].first), prod_range(sz, c->groups[g].second, nm));
                tb.load(0, at(B, c->off[s]), sz);
                for (size_t i = 0; i < dm.size(); i++)
                    tb.op(0, 1, dm[i], hsumop(s, dm[i]), 0, (int)sz[dm[i]], i > 0 ? 1 : 0);
                tb.store(1, at(HB, c->off[s]), 1);
                diag.push_back(tb);
            }
            for (size_t k = 0; k < K; k++) {
                const Chain &ch = chains[k];
                if (chain_ops(ch, g).empty()) continue;
                TaskBuilder tb(c, g, prod_range(csz[k], 0, c->groups[g].first),
                               prod_range(csz[k], c->groups[g].second, nm));
                tb.load(0, cur[k], csz[k]);
                bool square = true;
                for (auto &o : chain_ops(ch, g)) square = square && (csz[k][o.mode] == c->NS(o.mode, ch.lhs));
                const int resb = chain_group_ops(tb, k, g, 0, 1, -1, 1.0);
                // in place (same tensor) only when the tensor's layout is unchanged
                const int next = (curT[k] >= 0 && square) ? curT[k] : (curT[k] == 0 ? 1 : 0);
                tb.store(resb, Tbuf(k, next), 0);
                cur[k] = Tbuf(k, next);
                curT[k] = next;
                chn.push_back(tb);
            }
            pb.phase(diag);
            pb.phase(chn);
        }
    }
    // final pass, per output state
    {
        const int g = G - 1;
        std::vector<std::vector<std::vector<size_t>>> fin_batches(c->ns);
        for (int s = 0; s < c->ns; s++) {
            const std::vector<int64_t> sz = state_sizes(s);
            const int64_t L = prod_range(sz, 0, c->groups[g].first);
            const std::vector<int> dm = diag_ops(s, g);
            // chains into s in batches that fit one brick program's matrix
            // and instruction tables; batch b > 0 continues from HB in a
            // later phase
            std::vector<std::vector<size_t>> batch(1);
            int nmat = (int)dm.size(), ninstr = 3 + (int)dm.size();
            for (size_t k = 0; k < K; k++) {
                const Chain &ch = chains[k];
                if (ch.lhs != s) continue;
                if (prod_range(csz[k], 0, c->groups[g].first) != L)
                    throw HipError{"chain temporary does not match its output state's layout"};
                const int km = (int)chain_ops(ch, g).size(), ki = 1 + std::max(1, km);
                if (nmat + km > kMaxBrickMats - 2 || ninstr + ki > kMaxBrickInstr - 2) {
                    batch.emplace_back();
                    nmat = 0;
                    ninstr = 2;
                }
                batch.back().push_back(k);
                nmat += km;
                ninstr += ki;
            }
            fin_batches[s] = batch;
            (void)L;
        }
        size_t nb = 1;
        for (int s = 0; s < c->ns; s++) nb = std::max(nb, fin_batches[s].size());
        for (size_t b = 0; b < nb; b++) {
            std::vector<TaskBuilder> fin;
            for (int s = 0; s < c->ns; s++) {
                if (b >= fin_batches[s].size()) continue;
                const std::vector<int64_t> sz = state_sizes(s);
                const int64_t L = prod_range(sz, 0, c->groups[g].first);
                TaskBuilder tb(c, g, L, 1);
                if (b == 0) {
                    if (G >= 2) tb.load(1, at(HB, c->off[s]), sz);
                    else tb.zero(1, sz);
                    const std::vector<int> dm = diag_ops(s, g);
                    if (!dm.empty()) {
                        tb.load(0, at(B, c->off[s]), sz);
                        for (int m : dm) tb.op(0, 1, m, hsumop(s, m), 0, (int)sz[m], 1);
                    }
                } else {
                    tb.load(1, at(HB, c->off[s]), sz);
                }
                for (size_t k : fin_batches[s][b]) {
                    const Chain &ch = chains[k];
                    tb.load(0, cur[k], csz[k]);
                    if (chain_ops(ch, g).empty()) tb.axpy(0, 1, ch.ae);
                    else chain_group_ops(tb, k, g, 0, 2, 1, ch.ae);
                }
                tb.store(1, at(HB, c->off[s]), 0);
                fin.push_back(tb);
            }
            pb.phase(fin);
        }
    }
    return pb.finish(sprec);
}

const Plan &get_plan(mg_ctx *c, const PlanKey &key, std::unique_ptr<Plan> (*build)(mg_ctx *, const PlanKey &)) {
    auto it = c->plans.find(key);
    if (it != c->plans.end()) return *it->second;
    std::unique_ptr<Plan> p;
    try {
        p = build(c, key);
    } catch (const HipError &e) {
        throw HipError{"building launch plan " + std::to_string((uintptr_t)key[0]) + ": " + e.msg};
    }
    const Plan &ref = *p;
    c->plans[key] = std::move(p);
    return ref;
}

void apply_h(mg_ctx *c, const void *B, void *HB) {
    PlanKey key{(const void *)1, B, HB, nullptr};
    const Plan &p = get_plan(c, key, [](mg_ctx *cc, const PlanKey &k) {
        return build_apply_h(cc, k[1], (void *)k[2]);
    });
    run_plan(c, p);
    check_launch();
}

// H B with complex FP64 storage and matrices in a mixed-form context.
void apply_h64(mg_ctx *c, const double2 *B, double2 *HB) {
    PlanKey key{(const void *)10, B, HB, nullptr};
    const Plan &p = get_plan(c, key, [](mg_ctx *cc, const PlanKey &k) {
        return build_apply_h(cc, k[1], (void *)k[2], false, true);
    });
    run_plan(c, p);
    check_launch();
}

// H A in the Gaussian basis (conj(Adot_tmp) / der_factor of Calc_MeanField,
// gmctdh_objects.f90:550-578), work type.
void apply_hg(mg_ctx *c, const void *B, void *HB) {
    PlanKey key{(const void *)9, B, HB, nullptr};
    const Plan &p = get_plan(c, key, [](mg_ctx *cc, const PlanKey &k) {
        return build_apply_h(cc, k[1], (void *)k[2], true);
    });
    run_plan(c, p);
    check_launch();
}

GwpParams gwp_params(mg_ctx *c, double dt = 0.0) {
    GwpParams P;
    P.gwpa = c->dgwpa.as<double>();
    P.maxdof = c->maxdof;
    P.rho = c->drho.as<double2>();
    P.mf = c->dmf.as<double2>();
    P.nspfmax = c->nspfmax;
    P.nopmax = c->nopmax;
    P.nm = c->nm;
    P.ns = c->ns;
    P.eps_psi = c->eps_psi;
    P.eps_C = c->eps_C;
    P.eps_ovl = c->eps_ovl;
    P.min_stepsize = c->min_stepsize;
    P.dt = dt;
    P.dormand_prince = (c->int_type_psi == 2);
    P.der_factor = make_double2(0.0, -1.0);
    P.max_attempts = 200000;
    P.dt_src = nullptr;
    P.dt_scale = 1.0;
    return P;
}

// Y_s = (x_m M_(m,s)) X_s for every state over the modes in `mask` (bit m;
// norm_wf / WF_Ovl chains over all modes, A_to_B / B_to_A and the overlap
// weights of the density matrices over Gaussian modes), FP64.
std::unique_ptr<Plan> build_apply_kron(mg_ctx *c, const double2 *X, double2 *Y, const double2 *mats,
                                       unsigned mask) {
    PlanBuilder pb;
    const int G = (int)c->groups.size();
    const size_t nn = (size_t)c->nspfmax * c->nspfmax;
    auto in_mask = [&](int m) { return m >= 32 || ((mask >> m) & 1u); };
    int last = -1;
    for (int g = 0; g < G; g++)
        for (int m = c->groups[g].first; m < c->groups[g].second; m++)
            if (in_mask(m)) last = g;
    if (last < 0) throw HipError{"mode product chain without modes"};
    double2 *tmp = c->dkron.as<double2>();
    std::vector<const double2 *> cur(c->ns);
    for (int s = 0; s < c->ns; s++) cur[s] = X + c->off[s];
    for (int g = 0; g <= last; g++) {
        bool any = false;
        for (int m = c->groups[g].first; m < c->groups[g].second; m++) any = any || in_mask(m);
        if (!any) continue;
        std::vector<TaskBuilder> tasks;
        for (int s = 0; s < c->ns; s++) {
            std::vector<int64_t> sz(c->nm);
            for (int m = 0; m < c->nm; m++) sz[m] = c->NS(m, s);
            TaskBuilder tb(c, g, prod_range(sz, 0, c->groups[g].first), prod_range(sz, c->groups[g].second, c->nm));
            tb.load(0, cur[s], sz);
            for (int m = c->groups[g].first; m < c->groups[g].second; m++)
                if (in_mask(m)) tb.op(0, 0, m, mats + nn * (m + (size_t)c->nm * s), 0, (int)sz[m], 0);
            // the temporary is updated in place from the second group on
            double2 *out = (g == last) ? Y + c->off[s] : tmp + c->off[s];
            tb.store(0, out, 0);
            cur[s] = out;
            tasks.push_back(tb);
        }
        pb.phase(tasks);
    }
    return pb.finish(0);
}

void apply_kron(mg_ctx *c, const double2 *X, double2 *Y, const double2 *mats, unsigned mask = ~0u) {
    PlanKey key{(const void *)(uintptr_t)(mask == ~0u ? 2 : (0x100000000ull | mask)), X, Y, mats};
    const Plan &p = get_plan(c, key, [](mg_ctx *cc, const PlanKey &k) {
        const uintptr_t t = (uintptr_t)k[0];
        return build_apply_kron(cc, (const double2 *)k[1], (double2 *)k[2], (const double2 *)k[3],
                                t == 2 ? ~0u : (unsigned)(t & 0xffffffffu));
    });
    run_plan(c, p);
    check_launch();
}

// DeltaPsi's three overlaps <A1|(x_m S_m) A1> (S, S3, S2: the SPF overlap
// sets of delta_psi_overlaps) in one plan.  Groups before the last carry the
// three chains in temporaries (dkron twice, dHtmp); the last group's program
// loads A1 once, finishes each chain in a second buffer and accumulates the
// Gram A1^H T over the group's last mode, whose trace is <A1|T>: no chain is
// written out whole and no separate dot pass reads it back.
std::unique_ptr<Plan> build_delta_psi(mg_ctx *c, const double2 *A1) {
    PlanBuilder pb;
    const int G = (int)c->groups.size(), nm = c->nm, ns = c->ns;
    const size_t nn = (size_t)c->nspfmax * c->nspfmax;
    const double2 *mats[3] = {c->dS.as<double2>(), c->dS3.as<double2>(), c->dS2.as<double2>()};
    double2 *tmp[3] = {c->dkron.as<double2>(), c->dkron.as<double2>() + c->nconf, c->dHtmp.as<double2>()};
    if (c->ddpsi.bytes < sizeof(double2) * 3 * ns * nn) c->ddpsi.alloc(sizeof(double2) * 3 * ns * nn);
    double2 *slots = c->ddpsi.as<double2>();
    for (int g = 0; g < G; g++) {
        const int m0 = c->groups[g].first, m1 = c->groups[g].second;
        std::vector<TaskBuilder> tasks;
        for (int s = 0; s < ns; s++) {
            std::vector<int64_t> sz(nm);
            for (int m = 0; m < nm; m++) sz[m] = c->NS(m, s);
            const int64_t L = prod_range(sz, 0, m0), R = prod_range(sz, m1, nm);
            auto ops = [&](TaskBuilder &tb, int buf, int k) {
                for (int m = m0; m < m1; m++) tb.op(buf, buf, m, mats[k] + nn * (m + (size_t)nm * s), 0, (int)sz[m], 0);
            };
            if (g < G - 1) {
                for (int k = 0; k < 3; k++) {
                    TaskBuilder tb(c, g, L, R);
                    tb.load(0, (g == 0 ? A1 : tmp[k]) + c->off[s], sz);
                    ops(tb, 0, k);
                    tb.store(0, tmp[k] + c->off[s], 0);
                    tasks.push_back(tb);
                }
            } else {
                TaskBuilder tb(c, g, L, R);
                tb.load(0, A1 + c->off[s], sz);
                for (int k = 0; k < 3; k++) {
                    tb.load(1, (G == 1 ? A1 : tmp[k]) + c->off[s], sz);
                    ops(tb, 1, k);
                    tb.gram(0, 1, m1 - 1, slots + nn * ((size_t)k * ns + s));
                }
                tasks.push_back(tb);
            }
        }
        pb.phase(tasks);
    }
    return pb.finish(0);
}

// A_to_B / B_to_A (cmf2_base.f90:264-345): B = D^H S A and A = D B over the
// Gaussian modes, with the intra-state overlaps of the current SPFs (gwp_s00)
// and the D matrices dd.
void a_to_b(mg_ctx *c, const double2 *A, double2 *B, const double2 *dd) {
    launch_gwp_ab_mats(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, c->dSd.as<double2>(), dd, gwp_params(c),
                       c->dABmats.as<double2>(), c->stream);
    apply_kron(c, A, B, c->dABmats.as<double2>(), c->gwp_mask);
}

void b_to_a(mg_ctx *c, const double2 *B, double2 *A, const double2 *dd) {
    launch_gwp_ab_mats(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, c->dSd.as<double2>(), dd, gwp_params(c),
                       c->dABmats.as<double2>(), c->stream);
    const size_t half = (size_t)c->nspfmax * c->nspfmax * c->nm * c->ns;
    apply_kron(c, B, A, c->dABmats.as<double2>() + half, c->gwp_mask);
}

double2 dot(mg_ctx *c, const void *x, const void *y, int64_t n, int prec = 0) {
    double *part = c->dpartials.as<double>();
    double2 *out = reinterpret_cast<double2 *>(c->dscal.as<double>() + 8);
    launch_dot(x, y, n, part, c->nblocks_red, out, prec, c->stream);
    double2 h;
    HIP_CHECK(hipMemcpyAsync(&h, out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return h;
}

double norm2(mg_ctx *c, const void *x, int64_t n, int prec = 0) {
    double *part = c->dpartials.as<double>();
    double *out = c->dscal.as<double>() + 12;
    launch_norm2(x, n, part, c->nblocks_red, out, prec, c->stream);
    double h;
    HIP_CHECK(hipMemcpyAsync(&h, out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return h;
}

// ----------------------------------------------------------------------------
// SPF matrices: h1mode (hamilton.f90:433-485) and SPF overlaps.
void run_spf_tasks(mg_ctx *c, const PlanKey &key, std::vector<SpfTask> (*build)(mg_ctx *, const PlanKey &)) {
    auto &cache = c->spf_plans;
    auto it = cache.find(key);
    if (it == cache.end()) {
        std::vector<SpfTask> tasks = build(c, key);
        const size_t base = spf_scratch_base(c);
        int64_t maxng = 0;
        for (int m = 0; m < c->nm; m++)
            if (!c->gwp[m]) maxng = std::max(maxng, c->ngpm[m]);
        for (size_t i = 0; i < tasks.size(); i++)
            tasks[i].scratch = c->dspfscr.as<double2>() + base + i * 2 * (size_t)maxng * c->nspfmax;
        auto p = std::make_unique<SpfPlan>();
        p->n = (int)tasks.size();
        p->buf.alloc(sizeof(SpfTask) * std::max<size_t>(1, tasks.size()));
        if (!tasks.empty())
            hcopy(p->buf.p, tasks.data(), sizeof(SpfTask) * tasks.size(), hipMemcpyHostToDevice);
        it = cache.emplace(key, std::move(p)).first;
    }
    launch_spf_tasks(it->second->buf.as<SpfTask>(), it->second->n, dof_tables(c), 0, c->stream);
    check_launch();
}

std::vector<SpfTask> build_h1mode(mg_ctx *c, const PlanKey &key) {
    const double2 *psi = (const double2 *)key[1];
    std::vector<SpfTask> tasks;
    for (int m = 0; m < c->nm; m++)
        for (int i = 0; i < c->nOper[m]; i++) {
            if (c->gwp[m] || c->isIdent[m + (size_t)c->nm * i]) continue;
            const Oper &v = c->opers[(size_t)m * c->nopmax + i];
            const int s1 = v.lhs - 1, s2 = v.rhs - 1;
            SpfTask t;
            std::memset(&t, 0, sizeof t);
            t.phiA = psi + c->IP(m, s1);
            t.phiB = psi + c->IP(m, s2);
            t.nG = (int)c->ngpm[m];
            t.na = (int)c->NS(m, s1);
            t.nb = (int)c->NS(m, s2);
            t.mode = m;
            t.ncode = (int)c->nmodeDOF[m];
            for (int k = 0; k < t.ncode; k++) t.codes[k] = v.codes[k];
            t.sym = (s1 == s2);
            t.out = c->h1(m, i);
            t.ld = c->nspfmax;
            tasks.push_back(t);
        }
    return tasks;
}

// Calc_S00 of every Gaussian mode (gmctdh_objects.f90:323-381).
void gwp_s00(mg_ctx *c, const double2 *psi) {
    launch_gwp_s00(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, psi, gwp_params(c), c->dS00.as<double2>(),
                   c->dSd.as<double2>(), c->stream);
}

// Calc_h1mode of every mode (hamilton.f90:416-531): grid modes through the
// SPF block products, Gaussian modes from S00 and the Gaussian integrals.
// Without Gaussian modes the H.A diagonal sums follow at once; with them
// they wait for D^H h D (prepare_ha).
void calc_h1mode(mg_ctx *c, const double2 *psi) {
    run_spf_tasks(c, PlanKey{(const void *)3, psi, nullptr, nullptr}, build_h1mode);
    if (c->ngwp) {
        gwp_s00(c, psi);
        launch_gwp_h1(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, psi, gwp_params(c), c->dS00.as<double2>(),
                      c->dh1.as<double2>(), c->stream);
        if (c->prec)
            launch_convert(c->dh1f.p, 1, c->dh1.p, 0, (int64_t)(c->dh1.bytes / sizeof(double2)), c->stream);
        check_launch();
        return;
    }
    launch_hsum(c->dhsum.as<double2>(), c->dh1.as<double2>(), c->dhs_rowptr.as<int>(), c->dhs_src.as<int>(),
                c->dhs_coef.as<double2>(), c->ns * c->nm, c->nspfmax, c->stream);
    if (c->prec) {
        launch_convert(c->dh1f.p, 1, c->dh1.p, 0, (int64_t)(c->dh1.bytes / sizeof(double2)), c->stream);
        launch_convert(c->dhsumf.p, 1, c->dhsum.p, 0, (int64_t)(c->dhsum.bytes / sizeof(double2)), c->stream);
    }
}

// H.A operator table with Gaussian modes (Calc_DHD, hamilton.f90:535-574):
// h1mode on grid modes, D^H h D on Gaussian modes, then the diagonal sums.
void prepare_ha(mg_ctx *c, const double2 *dd) {
    HIP_CHECK(hipMemcpyAsync(c->dhA.p, c->dh1.p, c->dhA.bytes, hipMemcpyDeviceToDevice, c->stream));
    launch_gwp_dhd(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, c->dh1.as<double2>(), dd, gwp_params(c),
                   c->dhA.as<double2>(), c->stream);
    launch_hsum(c->dhsum.as<double2>(), c->dhA.as<double2>(), c->dhs_rowptr.as<int>(), c->dhs_src.as<int>(),
                c->dhs_coef.as<double2>(), c->ns * c->nm, c->nspfmax, c->stream);
    if (c->prec) {
        launch_convert(c->dhAf.p, 1, c->dhA.p, 0, (int64_t)(c->dhA.bytes / sizeof(double2)), c->stream);
        launch_convert(c->dhsumf.p, 1, c->dhsum.p, 0, (int64_t)(c->dhsum.bytes / sizeof(double2)), c->stream);
    }
    check_launch();
}

// D0 = InverseSquareRootU(S00) per Gaussian mode and set (cmf2_base.f90:72-80);
// needs gwp_s00 of the SPFs.
void gwp_dd_init(mg_ctx *c, double2 *dd) {
    launch_gwp_dd_init(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, c->dSd.as<double2>(), gwp_params(c), dd,
                       c->dstatus.as<int>(), c->stream);
    check_launch();
}

// out(m, s) = <phi1_(m,s) | phi2_(m,s)>; sym for phi1 == phi2 (zherk).
std::vector<SpfTask> build_overlaps(mg_ctx *c, const PlanKey &key) {
    const double2 *psi1 = (const double2 *)key[1], *psi2 = (const double2 *)key[2];
    double2 *out = (double2 *)key[3];
    const int sym = (psi1 == psi2) ? 1 : 0;
    std::vector<SpfTask> tasks;
    const size_t nn = (size_t)c->nspfmax * c->nspfmax;
    for (int s = 0; s < c->ns; s++)
        for (int m = 0; m < c->nm; m++) {
            if (c->gwp[m]) continue;
            SpfTask t;
            std::memset(&t, 0, sizeof t);
            t.phiA = psi1 + c->IP(m, s);
            t.phiB = psi2 + c->IP(m, s);
            t.nG = (int)c->ngpm[m];
            t.na = t.nb = (int)c->NS(m, s);
            t.mode = m;
            t.ncode = 0;
            t.sym = sym;
            t.out = out + nn * (m + (size_t)c->nm * s);
            t.ld = c->nspfmax;
            tasks.push_back(t);
        }
    return tasks;
}

void calc_overlaps(mg_ctx *c, const double2 *psi1, const double2 *psi2, double2 *out) {
    run_spf_tasks(c, PlanKey{(const void *)4, psi1, psi2, out}, build_overlaps);
    if (c->ngwp) {
        launch_gwp_overlaps(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, psi1, psi2, psi1 == psi2 ? 1 : 0, gwp_params(c),
                            out, c->stream);
        check_launch();
    }
}

// norm_wf (gmctdh_objects.f90:21-97)
double norm_wf(mg_ctx *c, const double2 *A, const double2 *psi) {
    calc_overlaps(c, psi, psi, c->dS.as<double2>());
    apply_kron(c, A, c->dHtmp.as<double2>(), c->dS.as<double2>());
    return dot(c, c->dHtmp.as<double2>(), A, c->nconf).x;
}

// WF_Ovl (gmctdh_objects.f90:101-180): <Psi1|Psi2>
cplx wf_ovl(mg_ctx *c, const double2 *A1, const double2 *psi1, const double2 *A2, const double2 *psi2) {
    calc_overlaps(c, psi1, psi2, c->dS2.as<double2>());
    apply_kron(c, A2, c->dHtmp.as<double2>(), c->dS2.as<double2>());
    double2 v = dot(c, A1, c->dHtmp.as<double2>(), c->nconf);
    return cplx(v.x, v.y);
}

// DeltaPsi (cmf2_base.f90:146-153 via norm_wf / WF_Ovl) in two halves: the
// SPF overlap matrices (which read p1, p2), then the A-vector contractions
// (which do not), so p2 may be updated on another stream in between.
void delta_psi_overlaps(mg_ctx *c, const double2 *p1, const double2 *p2) {
    calc_overlaps(c, p1, p1, c->dS.as<double2>());
    calc_overlaps(c, p2, p2, c->dS3.as<double2>());
    calc_overlaps(c, p1, p2, c->dS2.as<double2>());
}

double delta_psi_contract(mg_ctx *c, const double2 *A1, const double2 *A2) {
    apply_kron(c, A1, c->dHtmp.as<double2>(), c->dS.as<double2>());
    const double o11 = dot(c, c->dHtmp.as<double2>(), A1, c->nconf).x;
    apply_kron(c, A2, c->dHtmp.as<double2>(), c->dS3.as<double2>());
    const double o22 = dot(c, c->dHtmp.as<double2>(), A2, c->nconf).x;
    apply_kron(c, A2, c->dHtmp.as<double2>(), c->dS2.as<double2>());
    const double2 v = dot(c, A1, c->dHtmp.as<double2>(), c->nconf);
    const cplx o12(v.x, v.y);
    return std::sqrt(std::fabs(2.0 * (1.0 - o12.real() / std::sqrt(o11 * o22))));
}

// ----------------------------------------------------------------------------
// Density matrices and mean fields.
// The Gram contractions read the FP64 A (mixed form included, unless rho64
// is switched off: then an FP32 copy of A, left in dAf for the mean-field
// chains that follow) and accumulate in FP64; exact = true always reads the
// FP64 A (the reported populations and norm).
void calc_rho(mg_ctx *c, const double2 *A64, bool exact = false) {
    const void *A = A64;
    const int gp = (exact || c->rho64) ? 0 : c->prec;
    if (gp) {
        launch_convert(c->dAf.p, 1, A64, 0, c->nconf, c->stream);
        A = c->dAf.p;
    }
    if (c->ngwp) {
        // Overlap-weighted density matrices (gmctdh_objects.f90:439-457):
        // rho(m) = Gram over the other modes of (A, X_m), X_m = (x_{q != m}
        // S00_q) A over the Gaussian modes q; FP64.  X is shared by all grid
        // modes (every Gaussian mode weighted).
        unsigned have = 0xffffffffu;
        double2 *X = c->dHtmp.as<double2>();
        for (int m = 0; m < c->nm; m++) {
            const unsigned mask = c->gwp_mask & ~(1u << m);
            const double2 *Xm = A64;
            if (mask) {
                if (mask != have) apply_kron(c, A64, X, c->dSd.as<double2>(), mask);
                have = mask;
                Xm = X;
            }
            for (int st = 0; st < c->ns; st++) {
                GramTask g;
                std::memset(&g, 0, sizeof g);
                int64_t ll = 1, nl = 1;
                for (int q = 0; q < m; q++) ll *= c->NS(q, st);
                for (int q = m + 1; q < c->nm; q++) nl *= c->NS(q, st);
                g.X = A64 + c->off[st];
                g.Y = Xm + c->off[st];
                g.ll = ll;
                g.nl = nl;
                g.n1 = g.n2 = (int)c->NS(m, st);
                g.G = c->rho(m, st);
                g.ld = c->nspfmax;
                g.alpha = make_double2(1, 0);
                g.accumulate = 0;
                g.hermitian = 1;
                launch_gram(g, c->dpartials.as<double2>(), c->nblocks_red, c->nspfmax, 0, c->stream);
            }
        }
    } else if (c->nspfmax <= 8) {
        // one brick pass per mode group and state: all the group's Grams at once
        const void *Ain = gp ? A : (const void *)A64;
        PlanKey key{(const void *)(uintptr_t)(gp ? 5 : 6), Ain, nullptr, nullptr};
        auto it = c->plans.find(key);
        if (it == c->plans.end()) {
            const size_t esz = gp ? sizeof(float2) : sizeof(double2);
            PlanBuilder pb;
            const int G = (int)c->groups.size();
            for (int g = 0; g < G; g++) {
                std::vector<TaskBuilder> tasks;
                for (int s = 0; s < c->ns; s++) {
                    std::vector<int64_t> sz(c->nm);
                    for (int m = 0; m < c->nm; m++) sz[m] = c->NS(m, s);
                    TaskBuilder tb(c, g, prod_range(sz, 0, c->groups[g].first),
                                   prod_range(sz, c->groups[g].second, c->nm));
                    tb.load(0, static_cast<const char *>(Ain) + esz * (size_t)c->off[s], sz);
                    for (int m = c->groups[g].first; m < c->groups[g].second; m++) tb.gram(0, 0, m, c->rho(m, s));
                    tasks.push_back(tb);
                }
                pb.phase(tasks);
            }
            for (int s = 0; s < c->ns; s++)
                for (int m = 0; m < c->nm; m++) pb.hermitian(c->rho(m, s), (int)c->NS(m, s));
            it = c->plans.emplace(key, pb.finish(gp)).first;
        }
        run_plan(c, *it->second);
    } else
    for (int s = 0; s < c->ns; s++)
        for (int m = 0; m < c->nm; m++) {
            GramTask g;
            std::memset(&g, 0, sizeof g);
            int64_t ll = 1, nl = 1;
            for (int q = 0; q < m; q++) ll *= c->NS(q, s);
            for (int q = m + 1; q < c->nm; q++) nl *= c->NS(q, s);
            g.X = g.Y = gp ? c->at(A, c->off[s]) : (const void *)(A64 + c->off[s]);
            g.ll = ll;
            g.nl = nl;
            g.n1 = g.n2 = (int)c->NS(m, s);
            g.G = c->rho(m, s);
            g.ld = c->nspfmax;
            g.alpha = make_double2(1, 0);
            g.accumulate = 0;
            g.hermitian = 1;
            launch_gram(g, c->dpartials.as<double2>(), c->nblocks_red, c->nspfmax, gp, c->stream);
        }
    // rho_ave for single-set modes (gmctdh_objects.f90:464-474)
    std::vector<int> rowptr(1, 0), src;
    std::vector<double2> coef;
    std::vector<int> targets;
    for (int m = 0; m < c->nm; m++) {
        if (c->MultiSet[m]) continue;
        targets.push_back(m);
        for (int s = 0; s < c->ns; s++) { src.push_back(m + c->nm * s); coef.push_back(make_double2(1, 0)); }
        rowptr.push_back((int)src.size());
    }
    if (!targets.empty()) {
        // single-set modes are rare; do it with tiny axpys
        const size_t nn = (size_t)c->nspfmax * c->nspfmax;
        for (int m : targets) {
            launch_fill(c->rho_ave(m), (int64_t)nn, make_double2(0, 0), c->stream);
            for (int s = 0; s < c->ns; s++)
                launch_mat_axpy(c->rho_ave(m), c->rho(m, s), make_double2(1, 0), c->nspfmax, c->nspfmax, c->stream);
        }
    }
    check_launch();
}

// Mean fields through brick programs.  For chain term t (lhs a, rhs b, op O_m
// on every non-identity mode m) the hole contraction over mode m is
// Gram(L_m, R_m) with L_m = (x_{q<m} O_q^H) A_a and R_m = (x_{q>m} O_q) A_b.
// Over the mode groups: R passes (right to left) store R_g = ops of groups > g
// applied to A_b; Gram pass g loads L_g and R_g, and for each mode m of the
// group (ascending) forms R_m from R_g in a work buffer, accumulates the
// Gram, then moves L on by O_m^H; L_{g+1} is stored in place for pass g+1.
// The per-(term, mode) Grams and ae * rho for one-mode terms are then summed
// into the mean-field matrices (mirror terms as ae* G^H) by k_mat_combine.
// f64: complex FP64 storage and matrices in a mixed-form context.
std::unique_ptr<Plan> build_meanfield(mg_ctx *c, const void *A, bool f64 = false) {
    PlanBuilder pb;
    const int G = (int)c->groups.size(), nm = c->nm;
    const size_t mixed = (size_t)c->mixed, nn = (size_t)c->nspfmax * c->nspfmax;
    const bool wide = f64 && c->prec;
    const int sprec = wide ? 0 : c->prec;
    const size_t esz = sprec ? sizeof(float2) : sizeof(double2);
    auto at = [&](const void *p, int64_t n) { return static_cast<void *>(static_cast<char *>(const_cast<void *>(p)) + (size_t)n * esz); };
    auto h1op = [&](int m, int op) -> const void * { return wide ? (const void *)c->h1(m, op) : c->h1op(m, op); };
    std::vector<int> terms;
    for (size_t ti = 0; ti < c->mf_chain_terms.size(); ti++)
        if (c->mf_partner[ti] != -2) terms.push_back((int)ti);
    if (wide && c->dmfbuf64.bytes < esz * mixed * (size_t)G * terms.size())
        c->dmfbuf64.alloc(esz * mixed * (size_t)G * std::max<size_t>(1, terms.size()));
    DevBuf &mfbuf = wide ? c->dmfbuf64 : c->dmfbuf;
    if (mfbuf.bytes < esz * mixed * (size_t)G * terms.size() ||
        c->dgslots.bytes < sizeof(double2) * nn * nm * terms.size())
        throw HipError{"mean-field work buffers too small"};
    auto tensor = [&](size_t k, int g) { return at(mfbuf.p, (int64_t)(mixed * (k * G + g))); };
    auto is_op = [&](int t, int m) {
        const int op = c->iOper[m + (size_t)nm * t];
        return !c->isIdent[m + (size_t)nm * op];
    };
    // Gram slot per (term, mode): dgslots[k * nm + m]
    auto slot = [&](size_t k, int m) { return c->dgslots.as<double2>() + nn * (k * nm + m); };
    std::vector<std::vector<const void *>> R(terms.size(), std::vector<const void *>(G));
    // R passes
    for (size_t k = 0; k < terms.size(); k++)
        R[k][G - 1] = at(A, c->off[(int)c->rhs[c->mf_chain_terms[terms[k]]] - 1]);
    for (int g = G - 2; g >= 0; g--) {
        std::vector<TaskBuilder> tasks;
        for (size_t k = 0; k < terms.size(); k++) {
            const int t = c->mf_chain_terms[terms[k]];
            const int a = (int)c->lhs[t] - 1, b = (int)c->rhs[t] - 1;
            const int g1 = g + 1;
            std::vector<int> ms;
            for (int m = c->groups[g1].first; m < c->groups[g1].second; m++)
                if (is_op(t, m)) ms.push_back(m);
            if (ms.empty()) {
                R[k][g] = R[k][g1];
                continue;
            }
            std::vector<int64_t> sz(nm);
            for (int m = 0; m < nm; m++) sz[m] = c->NS(m, b);
            TaskBuilder tb(c, g1, prod_range(sz, 0, c->groups[g1].first), prod_range(sz, c->groups[g1].second, nm));
            tb.load(0, R[k][g1], sz);
            for (int m : ms) {
                const int op = c->iOper[m + (size_t)nm * t];
                tb.op(0, 0, m, h1op(m, op), 0, (int)c->NS(m, a), 0);
            }
            tb.store(0, tensor(k, g), 0);
            R[k][g] = tensor(k, g);
            tasks.push_back(tb);
        }
        pb.phase(tasks);
    }
    // Gram passes
    std::vector<const void *> Lcur(terms.size());
    for (size_t k = 0; k < terms.size(); k++)
        Lcur[k] = at(A, c->off[(int)c->lhs[c->mf_chain_terms[terms[k]]] - 1]);
    int last_mode_with_op = -1;
    for (int g = 0; g < G; g++) {
        std::vector<TaskBuilder> tasks;
        for (size_t k = 0; k < terms.size(); k++) {
            const int t = c->mf_chain_terms[terms[k]];
            const int a = (int)c->lhs[t] - 1, b = (int)c->rhs[t] - 1;
            (void)a;
            std::vector<int> ms;
            for (int m = c->groups[g].first; m < c->groups[g].second; m++)
                if (is_op(t, m)) ms.push_back(m);
            if (ms.empty()) continue;
            last_mode_with_op = -1;
            for (int m = nm - 1; m >= 0; m--)
                if (is_op(t, m)) { last_mode_with_op = m; break; }
            std::vector<int64_t> sz(nm);
            for (int m = 0; m < nm; m++) sz[m] = c->NS(m, b);
            TaskBuilder tb(c, g, prod_range(sz, 0, c->groups[g].first), prod_range(sz, c->groups[g].second, nm));
            tb.load(0, Lcur[k], sz);
            tb.load(1, R[k][g], sz);
            for (size_t i = 0; i < ms.size(); i++) {
                const int m = ms[i];
                int rsrc = 1;
                for (size_t j = i + 1; j < ms.size(); j++) {
                    const int op = c->iOper[ms[j] + (size_t)nm * t];
                    tb.op(rsrc, 2, ms[j], h1op(ms[j], op), 0, (int)sz[ms[j]], 0);
                    rsrc = 2;
                }
                tb.gram(0, rsrc, m, slot(k, m));
                if (m != last_mode_with_op) {
                    const int op = c->iOper[m + (size_t)nm * t];
                    tb.op(0, 0, m, h1op(m, op), 1, (int)sz[m], 0);
                }
            }
            if (g + 1 < G) {
                // L moves on only if a later group still has Grams for this term
                void *Lout = tensor(k, G - 1);
                tb.store(0, Lout, 0);
                Lcur[k] = Lout;
            }
            tasks.push_back(tb);
        }
        pb.phase(tasks);
    }
    return pb.finish(sprec);
}

// Combination table: MF(m, op) = sum ae * rho (one-mode terms) + sum ae G_t(m)
// + sum conj(ae) G_t(m)^H (mirror terms), for every (m, op).
void build_mf_combine(mg_ctx *c) {
    const int nm = c->nm;
    const size_t nn = (size_t)c->nspfmax * c->nspfmax;
    std::vector<int> terms;
    for (size_t ti = 0; ti < c->mf_chain_terms.size(); ti++)
        if (c->mf_partner[ti] != -2) terms.push_back((int)ti);
    const int nrow = nm * c->nopmax;
    std::vector<std::vector<std::tuple<const double2 *, double2, int>>> rows(nrow);
    {   // identity matrix for ae * 1 contributions
        std::vector<double2> I(nn, make_double2(0, 0));
        for (int j = 0; j < c->nspfmax; j++) I[j + (size_t)c->nspfmax * j] = make_double2(1, 0);
        c->dident.alloc(sizeof(double2) * nn);
        hcopy(c->dident.p, I.data(), sizeof(double2) * nn, hipMemcpyHostToDevice);
    }
    for (int r = 0; r < nrow; r++)
        for (int q = c->mf1_rowptr[r]; q < c->mf1_rowptr[r + 1]; q++) {
            const double2 *src = c->mf1_src[q] >= 0 ? c->drho.as<double2>() + nn * c->mf1_src[q] : c->dident.as<double2>();
            rows[r].push_back({src, d2(c->mf1_coef[q]), 0});
        }
    for (size_t k = 0; k < terms.size(); k++) {
        const int ti = terms[k];
        const int t = c->mf_chain_terms[ti];
        const int partner = c->mf_partner[ti];
        for (int m = 0; m < nm; m++) {
            const int op = c->iOper[m + (size_t)nm * t];
            if (c->isIdent[m + (size_t)nm * op]) continue;
            const double2 *G = c->dgslots.as<double2>() + nn * (k * nm + m);
            rows[m * c->nopmax + op].push_back({G, d2(c->ae[t]), 0});
            if (partner >= 0) {
                const int op2 = c->iOper[m + (size_t)nm * partner];
                rows[m * c->nopmax + op2].push_back({G, d2(std::conj(c->ae[t])), 1});
            }
        }
    }
    std::vector<double2 *> dst(nrow);
    std::vector<int> rowptr(1, 0), trans;
    std::vector<const double2 *> src;
    std::vector<double2> coef;
    for (int r = 0; r < nrow; r++) {
        dst[r] = c->dmf.as<double2>() + nn * r;
        for (auto &e : rows[r]) {
            src.push_back(std::get<0>(e));
            coef.push_back(std::get<1>(e));
            trans.push_back(std::get<2>(e));
        }
        rowptr.push_back((int)src.size());
    }
    auto up = [&](DevBuf &b, const void *h, size_t bytes) {
        b.alloc(std::max<size_t>(bytes, 8));
        if (bytes) hcopy(b.p, h, bytes, hipMemcpyHostToDevice);
    };
    up(c->dmf_dst, dst.data(), sizeof(double2 *) * dst.size());
    up(c->dmf_rowptr, rowptr.data(), sizeof(int) * rowptr.size());
    up(c->dmf_src, src.data(), sizeof(double2 *) * src.size());
    up(c->dmf_coef, coef.data(), sizeof(double2) * coef.size());
    up(c->dmf_trans, trans.data(), sizeof(int) * trans.size());
}

void calc_meanfield_bricks(mg_ctx *c, const double2 *A64, bool af_ready) {
    const void *A = A64;
    if (c->prec && !c->mf64) {
        if (!af_ready || c->rho64) launch_convert(c->dAf.p, 1, A64, 0, c->nconf, c->stream);
        A = c->dAf.p;
    }
    PlanKey key{(const void *)(uintptr_t)(c->mf64 ? 11 : 7), A, nullptr, nullptr};
    auto it = c->plans.find(key);
    if (it == c->plans.end()) {
        it = c->plans.emplace(key, build_meanfield(c, A, c->mf64)).first;
        if (!c->dmf_rowptr.p) build_mf_combine(c);
    }
    if (!c->mf_chain_terms.empty()) run_plan(c, *it->second);
    launch_mat_combine(c->dmf_dst.as<double2 *>(), c->dmf_rowptr.as<int>(), c->dmf_src.as<const double2 *>(),
                       c->dmf_coef.as<double2>(), c->dmf_trans.as<int>(), c->nm * c->nopmax, c->nspfmax, c->stream);
    launch_rmf(c->drho.as<double2>(), c->dmf.as<double2>(), c->drmf.as<double2>(), c->drmf_jobs.as<int>(),
               c->n_rmf_jobs, c->nspfmax, c->nopmax, c->eps_rho, c->stream, c->drmf_minw.as<double>());
    check_launch();
}

// Calc_MeanField, Hamiltonian part (gmctdh_objects.f90:518-605) + Calc_rhom1MF
// (:1147-1194).  Requires calc_rho(A) and calc_h1mode(psi).
// af_ready: the mixed form's dAf already holds A (calc_rho(A) just ran).
void calc_meanfield(mg_ctx *c, const double2 *A64, bool af_ready = false) {
    if (c->mf_bricks) {
        calc_meanfield_bricks(c, A64, af_ready);
        return;
    }
    // one-mode same-state terms: MF += ae * rho  (all other modes are identity)
    launch_hsum(c->dmf.as<double2>(), c->drho.as<double2>(), c->dmf1_rowptr.as<int>(), c->dmf1_src.as<int>(),
                c->dmf1_coef.as<double2>(), c->nm * c->nopmax, c->nspfmax, c->stream);
    const size_t mixed = (size_t)c->mixed;
    // the chains run in the work-tensor type: the mixed form multiplies an
    // FP32 copy of A; the Gram contractions accumulate in FP64 either way
    const void *A = A64;
    if (c->prec && !c->mf_chain_terms.empty()) {
        if (!af_ready || c->rho64) launch_convert(c->dAf.p, 1, A64, 0, c->nconf, c->stream);
        A = c->dAf.p;
    }
    void *AL = c->dmfbuf.p;
    void *AR = c->at(AL, (int64_t)(mixed * c->nm));
    for (size_t ti = 0; ti < c->mf_chain_terms.size(); ti++) {
        const int t = c->mf_chain_terms[ti];
        const int partner = c->mf_partner[ti];
        if (partner == -2) continue;   // derived from its mirror term
        const int a = (int)c->lhs[t] - 1, b = (int)c->rhs[t] - 1;
        const cplx ae = c->ae[t];
        // right chain R_i (modes f-1 .. 0), pointers alias on identity modes
        std::vector<const void *> R(c->nm);
        std::vector<int64_t> sz(c->nm);
        for (int m = 0; m < c->nm; m++) sz[m] = c->NS(m, b);
        R[0] = c->at(A, c->off[b]);
        for (int i = 1; i < c->nm; i++) {
            const int jm = c->nm - i;
            const int op = c->iOper[jm + (size_t)c->nm * t];
            if (c->mfskip[jm + (size_t)c->nm * op]) { R[i] = R[i - 1]; continue; }
            ModeTask mt;
            int64_t ll = 1, nl = 1;
            for (int q = 0; q < jm; q++) ll *= sz[q];
            for (int q = jm + 1; q < c->nm; q++) nl *= sz[q];
            mt.X = R[i - 1];
            mt.Y = c->at(AR, (int64_t)(mixed * i));
            mt.ll = ll;
            mt.nl = nl;
            mt.nin = (int)sz[jm];
            mt.nout = (int)c->NS(jm, a);
            mt.M = c->h1op(jm, op);
            mt.ld = c->nspfmax;
            mt.adj = 0;
            launch_mode_product(mt, c->nspfmax, c->prec, c->stream);
            sz[jm] = mt.nout;
            R[i] = mt.Y;
        }
        // left chain L_i (modes 0 .. f-2) with h^H
        std::vector<const void *> Lc(c->nm);
        for (int m = 0; m < c->nm; m++) sz[m] = c->NS(m, a);
        Lc[0] = c->at(A, c->off[a]);
        for (int i = 1; i < c->nm; i++) {
            const int m = i - 1;
            const int op = c->iOper[m + (size_t)c->nm * t];
            if (c->mfskip[m + (size_t)c->nm * op]) { Lc[i] = Lc[i - 1]; continue; }
            ModeTask mt;
            int64_t ll = 1, nl = 1;
            for (int q = 0; q < m; q++) ll *= sz[q];
            for (int q = m + 1; q < c->nm; q++) nl *= sz[q];
            mt.X = Lc[i - 1];
            mt.Y = c->at(AL, (int64_t)(mixed * i));
            mt.ll = ll;
            mt.nl = nl;
            mt.nin = (int)sz[m];
            mt.nout = (int)c->NS(m, b);
            mt.M = c->h1op(m, op);
            mt.ld = c->nspfmax;
            mt.adj = 1;
            launch_mode_product(mt, c->nspfmax, c->prec, c->stream);
            sz[m] = mt.nout;
            Lc[i] = mt.Y;
        }
        for (int m = 0; m < c->nm; m++) {
            const int op = c->iOper[m + (size_t)c->nm * t];
            if (c->isIdent[m + (size_t)c->nm * op]) continue;
            int64_t ll = 1, nl = 1;
            for (int q = 0; q < m; q++) ll *= c->NS(q, b);
            for (int q = m + 1; q < c->nm; q++) nl *= c->NS(q, a);
            GramTask g;
            std::memset(&g, 0, sizeof g);
            g.X = Lc[m];
            g.Y = R[c->nm - 1 - m];
            g.ll = ll;
            g.nl = nl;
            g.n1 = (int)c->NS(m, a);
            g.n2 = (int)c->NS(m, b);
            g.ld = c->nspfmax;
            g.alpha = d2(ae);
            g.hermitian = 0;
            if (partner < 0) {
                g.G = c->mf(m, op);
                g.accumulate = 1;
                launch_gram(g, c->dpartials.as<double2>(), c->nblocks_red, c->nspfmax, c->prec, c->stream);
            } else {
                double2 *G = c->dGtmp.as<double2>();
                g.G = G;
                g.accumulate = 0;
                launch_gram(g, c->dpartials.as<double2>(), c->nblocks_red, c->nspfmax, c->prec, c->stream);
                const int op2 = c->iOper[m + (size_t)c->nm * partner];
                launch_mf_accum(c->mf(m, op), c->mf(m, op2), G, g.n1, g.n2, c->nspfmax, c->stream);
            }
        }
    }
    launch_rmf(c->drho.as<double2>(), c->dmf.as<double2>(), c->drmf.as<double2>(), c->drmf_jobs.as<int>(),
               c->n_rmf_jobs, c->nspfmax, c->nopmax, c->eps_rho, c->stream, c->drmf_minw.as<double>());
    check_launch();
}

// ----------------------------------------------------------------------------
// propagate_SPF for all DVR modes (cmf2_base.f90:210-260), one workgroup each.
SpfParams spf_params(mg_ctx *c, double dt) {
    SpfParams P;
    P.rmf = c->drmf.as<double2>();
    P.rho = c->drho.as<double2>();
    P.nspfmax = c->nspfmax;
    P.nopmax = c->nopmax;
    P.eps_psi = c->eps_psi;
    P.min_stepsize = c->min_stepsize;
    P.dt = dt;
    P.dormand_prince = (c->int_type_psi == 2);
    P.der_factor = make_double2(0.0, -1.0);
    P.max_attempts = 200000;
    P.dt_src = nullptr;
    P.dt_scale = 1.0;
    return P;
}

void spf_check(mg_ctx *c) {
    const int st = *c->spf_status_host;
    if (st == 1) throw HipError{"SPF overlap matrix not positive definite"};
    if (st == 2) throw HipError{"SPF Runge-Kutta did not converge"};
    if (st == 3) throw HipError{"Gaussian overlap matrix not positive definite"};
    if (st == 4) throw HipError{"operator not supported on a Gaussian mode"};
}

// Device-to-device copy of n complex FP64 elements on the engine stream.
void copy_d2d(mg_ctx *c, double2 *dst, const double2 *src, int64_t n) {
    HIP_CHECK(hipMemcpyAsync(dst, src, sizeof(double2) * n, hipMemcpyDeviceToDevice, c->stream));
}

// ----------------------------------------------------------------------------
// Learned step controller (no counterpart in the reference).  After an
// accepted CMF2 step the 12 features below describe the step history and the
// wavefunction; a 12-64-64-1 perceptron (mg_mlp.hip) maps them to
// y = log(h_next / h).  The CMF2 error estimate still accepts or rejects
// every attempt unchanged and rejections use the classical rule, so the
// model can only change how often attempts are rejected, never the accuracy
// criterion.
void note_attempt(mg_ctx *c, double h, double error) {
    c->hist_h.insert(c->hist_h.begin(), h);
    c->hist_e.insert(c->hist_e.begin(), error);
    if (c->hist_h.size() > 4) { c->hist_h.pop_back(); c->hist_e.pop_back(); }
}

// f0..f3 log h of the last 4 attempts (most recent first, padded with the
// oldest), f4..f7 log10 of their error ratios err / eps_int, f8 log10 of the
// relative change of the mean fields since the previous accepted step, f9
// log10 of the smallest eigenvalue of the grid modes' density matrices, f10
// log10 |tr rho(mode 1) - 1| (norm drift), f11 the most populated state
// / (nstates - 1).  rho and the mean fields are those of the step's half
// step, the last ones the step evaluated.
void step_features(mg_ctx *c, float *f) {
    const size_t nh = c->hist_h.size();
    for (int i = 0; i < 4; i++) {
        const size_t k = std::min<size_t>(i, nh - 1);
        f[i] = (float)std::log(c->hist_h[k]);
        f[4 + i] = (float)std::log10(std::max(c->hist_e[k], 1e-12));
    }
    const int64_t nmf = (int64_t)(c->dmf.bytes / sizeof(double2));
    double rel = 0.0;
    if (c->mf_prev_valid) {
        HIP_CHECK(hipMemcpyAsync(c->dmf_tmp.p, c->dmf.p, c->dmf.bytes, hipMemcpyDeviceToDevice, c->stream));
        launch_axpby(c->dmf_tmp.p, c->dmf_prev.p, -1.0, 1.0, nmf, 0, c->stream);
        const double d = norm2(c, c->dmf_tmp.p, nmf), t = norm2(c, c->dmf.p, nmf);
        rel = t > 0 ? std::sqrt(d / t) : 0.0;
    }
    HIP_CHECK(hipMemcpyAsync(c->dmf_prev.p, c->dmf.p, c->dmf.bytes, hipMemcpyDeviceToDevice, c->stream));
    c->mf_prev_valid = true;
    f[8] = (float)std::log10(rel + 1e-12);
    double mw = 1.0;
    if (c->n_rmf_jobs > 0) {
        std::vector<double> w(c->n_rmf_jobs);
        HIP_CHECK(hipMemcpyAsync(w.data(), c->drmf_minw.p, sizeof(double) * w.size(), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        mw = *std::min_element(w.begin(), w.end());
    }
    f[9] = (float)std::log10(std::max(mw, 1e-30));
    const size_t nn = (size_t)c->nspfmax * c->nspfmax;
    std::vector<double2> r(nn * c->nm * c->ns);
    HIP_CHECK(hipMemcpyAsync(r.data(), c->drho.p, sizeof(double2) * r.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    double tot = 0.0, best = -1.0;
    int arg = 0;
    for (int s = 0; s < c->ns; s++) {
        double p = 0.0;
        for (int64_t j = 0; j < c->NS(0, s); j++) p += r[nn * ((size_t)c->nm * s) + j + (size_t)c->nspfmax * j].x;
        tot += p;
        if (p > best) { best = p; arg = s; }
    }
    f[10] = (float)std::log10(std::fabs(tot - 1.0) + 1e-16);
    f[11] = c->ns > 1 ? (float)arg / (float)(c->ns - 1) : 0.0f;
}

// Next step after an accepted attempt (hdid, error ratio): the classical
// rule min(h / sqrt(err), 1.5 h) or the model's proposal h exp(y) clamped to
// [0.2 h, 2 h]; both capped at the output interval.
double propose_next(mg_ctx *c, double hdid, double error) {
    const double classical = std::min(hdid / std::sqrt(error), hdid * 1.5);
    double hn = classical;
    if (c->mlp_on || c->log_features) {
        float f[kMlpNFeat];
        step_features(c, f);
        if (c->log_features) {
            c->feat_log.insert(c->feat_log.end(), f, f + kMlpNFeat);
            c->feat_h.push_back(hdid);
        }
        if (c->mlp_on) {
            HIP_CHECK(hipMemcpyAsync(c->dmlp_x.p, f, sizeof f, hipMemcpyHostToDevice, c->stream));
            float *w = c->dmlp_work.as<float>();
            MlpWork mw{w, w + kMlpHidden, w + 2 * kMlpHidden, w + 3 * kMlpHidden, w + 4 * kMlpHidden,
                       w + 5 * kMlpHidden, w + 6 * kMlpHidden, w + 6 * kMlpHidden + 1};
            launch_mlp_forward(c->dmlp_x.as<float>(), c->dmlp_p.as<float>(), c->dmlp_wb.as<uint16_t>(), mw,
                               c->dmlp_y.as<float>(), nullptr, 1, c->stream);
            check_launch();
            float y = 0.f;
            HIP_CHECK(hipMemcpyAsync(&y, c->dmlp_y.p, sizeof y, hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
            const double cand = hdid * std::exp((double)y);
            if (std::isfinite(cand) && cand > 0.0) {
                hn = std::min(std::max(cand, 0.2 * hdid), 2.0 * hdid);
                c->mlp_proposals++;
            } else {
                c->mlp_fallbacks++;
            }
        }
    }
    return std::min(hn, c->writetime);
}


// Phase buckets of MG_PHASES: 0 rho + h1mode, 1 mean fields, 2 SPF
// propagation (the synchronous one), 3 Lanczos, 4 DeltaPsi, 5 copies,
// errors and the controller.
const char *const kPhaseNames[6] = {"rho+h1mode", "meanfield", "spf(sync)", "lanczos", "deltapsi", "other"};
void phase(mg_ctx *c, int bucket) {
    if (!c->phases) return;
    HIP_CHECK(hipStreamSynchronize(c->stream));
    const auto t = std::chrono::steady_clock::now();
    c->phase_ms[bucket] += std::chrono::duration<double, std::milli>(t - c->phase_t0).count();
    c->phase_t0 = t;
}

// ----------------------------------------------------------------------------
// Device-controlled CMF2 attempt (cmf2_base.f90:76-203), grid and Gaussian
// modes.  Every decision of an attempt is taken on the device: the Lanczos
// stopping test and propagator coefficients (k_lan_ctrl, k_lan_coef),
// DeltaPsi and err_A, the accept/reject decision and the copies that commit
// an accepted step (k_attempt_decide, k_copy_if); the attempted step h comes
// from device memory.  The host writes h (and a replayed decision) before,
// and reads the errors, Krylov sizes and the decision after an attempt: the
// attempt is one stream-ordered sequence with no host synchronisation inside,
// and with MG_ATTEMPT_GRAPH=1 it is captured once into a HIP graph and
// replayed for every later attempt.  A Krylov build enqueues kmax
// iterations; those past convergence return at once on a device flag.

AttemptDev *att_dev(mg_ctx *c) { return c->datt.as<AttemptDev>(); }
LanczosDev *lan_dev(mg_ctx *c, int i) {
    return reinterpret_cast<LanczosDev *>(c->datt.as<char>() + sizeof(AttemptDev) + 64 + (size_t)i * sizeof(LanczosDev));
}

// Lanczos (cmf2_integration.f90:30-103) from B0 (FP64) over dtscale * h; BT
// FP64 and normalised.  build = false reuses L's Krylov space.  b0f: an FP32
// copy of B0 already in memory (the mixed form's mean-field input), used as
// the work-type u_0 instead of converting B0 again.  normalise = false
// leaves BT unnormalised with its squared norm in dscal[6], for a consumer
// that normalises as it reads.
void lanczos_dev(mg_ctx *c, const double2 *B0, double2 *BT, LanczosDev *L, double dtscale, bool build, int kmax,
                 const void *b0f = nullptr, bool normalise = true) {
    const int64_t N = c->nconf;
    const int pr = c->prec;
    void *BK = c->dBK.p;
    double *dsc = c->dscal.as<double>();
    double *nsq = dsc + 24;          // n_j^2, j <= kNKrylovMax
    double *part = c->dpartials.as<double>();
    double2 *dot_d = reinterpret_cast<double2 *>(dsc + 2);
    const AttemptDev *at = att_dev(c);
    if (build) {
        launch_lan_begin(L, c->stream);
        launch_norm2(B0, N, part, c->nblocks_red, nsq + 0, 0, c->stream);
        if (pr && !b0f) launch_convert(BK, 1, B0, 0, N, c->stream);
        const void *u0 = pr ? (b0f ? b0f : BK) : (const void *)B0;
        for (int j = 1; j <= kmax; j++) {
            const void *uj = (j == 1) ? u0 : c->at(BK, N * (j - 1));
            void *w = c->at(BK, N * j);
            const void *up = (j >= 2) ? ((j == 2) ? u0 : c->at(BK, N * (j - 2))) : nullptr;
            const int *skip = (j >= 2) ? &L->done : nullptr;
            if (j == 1 && pr && c->lan64) {
                double2 *w64 = c->dW64.as<double2>();
                apply_h64(c, B0, w64);
                launch_dot(B0, w64, N, part, c->nblocks_red, dot_d, 0, c->stream);
                launch_lanczos_step(w64, B0, nullptr, dsc + 2, nsq, 0, N, part, c->nblocks_red, nsq + 1, 0, c->stream);
                launch_convert(w, 1, w64, 0, N, c->stream);
            } else {
                struct SkipScope {
                    mg_ctx *c;
                    ~SkipScope() { c->plan_skip = nullptr; }
                } scope{c};
                c->plan_skip = skip;
                apply_h(c, uj, w);
                launch_dot(uj, w, N, part, c->nblocks_red, dot_d, pr, c->stream, skip);
                launch_lanczos_step(w, uj, up, dsc + 2, nsq, j - 1, N, part, c->nblocks_red, nsq + j, pr, c->stream,
                                    skip);
            }
            launch_lan_ctrl(L, at, dot_d, nsq, j, dtscale, kmax, c->stream);
        }
    }
    double2 *dcoef = reinterpret_cast<double2 *>(dsc + 48);
    launch_lan_coef(L, at, dtscale, build ? 1 : 0, dcoef, c->stream);
    // B0 is u_0 of the space (also when it is reused)
    launch_combine_v0(BT, B0, BK, N, dcoef, 0, N, pr, c->stream, &L->nk);
    launch_norm2(BT, N, part, c->nblocks_red, dsc + 6, 0, c->stream);
    if (normalise) launch_scale(BT, dsc + 6, N, 0, c->stream);
    check_launch();
}

// DeltaPsi (gmctdh_objects.f90:204-227) of (A1, psi1) and (A1, psi2) from
// the SPF overlaps queued by delta_psi_overlaps; the result in at->err_psi.
void delta_psi_contract_dev(mg_ctx *c, const double2 *A1) {
    double *part = c->dpartials.as<double>();
    double2 *o = reinterpret_cast<double2 *>(c->dscal.as<double>() + 88);   // o11, o22, o12
    if (c->nspfmax <= 8 && c->ngwp == 0 && c->dps_fused) {
        // real parts only enter err_psi, and Re <T|A1> = Re <A1|T>
        PlanKey key{(const void *)12, A1, nullptr, nullptr};
        const Plan &p = get_plan(c, key, [](mg_ctx *cc, const PlanKey &k) {
            return build_delta_psi(cc, (const double2 *)k[1]);
        });
        run_plan(c, p);
        launch_dpsi_traces(c->ddpsi.as<double2>(), c->ns, c->nspfmax, o, c->stream);
        launch_attempt_psi(att_dev(c), o + 0, o + 1, o + 2, c->stream);
        check_launch();
        return;
    }
    double2 *T = c->dHtmp.as<double2>();
    apply_kron(c, A1, T, c->dS.as<double2>());
    launch_dot(T, A1, c->nconf, part, c->nblocks_red, o + 0, 0, c->stream);
    apply_kron(c, A1, T, c->dS3.as<double2>());
    launch_dot(T, A1, c->nconf, part, c->nblocks_red, o + 1, 0, c->stream);
    apply_kron(c, A1, T, c->dS2.as<double2>());
    launch_dot(A1, T, c->nconf, part, c->nblocks_red, o + 2, 0, c->stream);
    launch_attempt_psi(att_dev(c), o + 0, o + 1, o + 2, c->stream);
}

// SPF half step over dtscale * h (h on the device) on stream st; Gaussian
// modes carry their D matrices dd (cmf2_base.f90:110-124).
void propagate_spf_dev(mg_ctx *c, double2 *psi, double dtscale, hipStream_t st, double2 *dd = nullptr) {
    if (c->ndvr_desc > 0) {
        SpfParams P = spf_params(c, 0.0);
        P.dt_src = &att_dev(c)->h;
        P.dt_scale = dtscale;
        launch_spf_propagate(c->dspfmodes.as<SpfModeDesc>(), c->ndvr_desc, psi, c->dspfscr.as<double2>(), P,
                             dof_tables(c), c->dstatus.as<int>(), st);
    }
    if (c->ngwp) {
        if (!dd) throw HipError{"Gaussian-mode propagation without D matrices"};
        GwpParams G = gwp_params(c, 0.0);
        G.dt_src = &att_dev(c)->h;
        G.dt_scale = dtscale;
        launch_gwp_propagate(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, psi, dd, c->dgwpscr.as<double2>(), G,
                             c->dstatus.as<int>(), st);
    }
}

// Second stream for an SPF half step that overlaps the A-vector work after
// everything queued so far (the stream itself when MG_ATTEMPT_FORK=0), and
// the join back.
hipStream_t attempt_fork(mg_ctx *c) {
    if (!c->att_fork) return c->stream;
    HIP_CHECK(hipEventRecord(c->ev_fork, c->stream));
    HIP_CHECK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    return c->stream2;
}
void attempt_join(mg_ctx *c) {
    if (!c->att_fork) return;
    HIP_CHECK(hipEventRecord(c->ev_join, c->stream2));
    HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
}

// Everything of one attempt, in stream order (graph-safe: no host sync, no
// allocation, no host-side decision).
void cmf2_attempt_enqueue(mg_ctx *c, int kmax) {
    const int64_t N = c->nconf, D = c->dimpsi;
    double2 *A = c->dA.as<double2>(), *psi = c->dpsi.as<double2>();
    double2 *B0 = c->dB0.as<double2>(), *BH = c->dBH.as<double2>(), *B02 = c->dB02.as<double2>();
    double2 *pH = c->dpH.as<double2>(), *pH2 = c->dpH2.as<double2>();
    AttemptDev *at = att_dev(c);
    LanczosDev *L1 = lan_dev(c, 0), *L2 = lan_dev(c, 1);
    launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
    calc_rho(c, A);
    calc_h1mode(c, psi);
    phase(c, 0);
    calc_meanfield(c, A, true);
    phase(c, 1);
    copy_d2d(c, B0, A, N);
    copy_d2d(c, pH, psi, D);
    // the SPF half step needs only rho(A) and the mean fields: stream2, beside
    // the Lanczos propagation of A
    // the mixed form's mean fields left an FP32 copy of A (= B0) in dAf
    const void *a_f32 = (c->prec && c->mf_bricks && !c->mf64) ? c->dAf.p : nullptr;
    propagate_spf_dev(c, pH, 0.5, attempt_fork(c));
    lanczos_dev(c, B0, BH, L1, 0.5, true, kmax, a_f32);
    phase(c, 3);
    attempt_join(c);
    phase(c, 2);
    calc_rho(c, BH);
    calc_h1mode(c, pH);
    phase(c, 0);
    calc_meanfield(c, BH, true);
    phase(c, 1);
    copy_d2d(c, pH2, psi, D);
    propagate_spf_dev(c, pH2, 0.5, c->stream);
    phase(c, 2);
    // DeltaPsi reads pH2 only through its SPF overlaps: once those are
    // queued, the second SPF half step overlaps the A-vector contractions
    delta_psi_overlaps(c, pH, pH2);
    propagate_spf_dev(c, pH2, 0.5, attempt_fork(c));
    delta_psi_contract_dev(c, BH);
    phase(c, 4);
    attempt_join(c);
    phase(c, 2);
    calc_h1mode(c, pH2);
    phase(c, 0);
    // dAf still holds the FP32 copy of BH from the second mean fields; B02
    // and the accepted B0 stay unnormalised (squared norm in dscal[6]) and
    // are normalised by their one reader, as the scaling pass would
    double *bt_n2 = c->dscal.as<double>() + 6;
    lanczos_dev(c, BH, B02, L2, -0.5, true, kmax, a_f32, false);
    phase(c, 3);
    double *n2 = c->dscal.as<double>() + 12;
    launch_diff_norm2(B02, B0, N, c->dpartials.as<double>(), c->nblocks_red, n2, c->stream, bt_n2);
    launch_attempt_decide(at, L1, L2, n2, c->eps_int, c->dstatus.as<int>(), c->stream);
    // the accepted step: A from the half step with the same Krylov space
    // (computed either way; committed only when accepted)
    lanczos_dev(c, BH, B0, L2, 0.5, false, kmax, nullptr, false);
    phase(c, 3);
    launch_copy_if(A, B0, N, &at->accept, c->stream, bt_n2);
    launch_copy_if(psi, pH2, D, &at->accept, c->stream);
    HIP_CHECK(hipMemcpyAsync(c->att_host, at, sizeof(AttemptDev), hipMemcpyDeviceToHost, c->stream));
    check_launch();
    phase(c, 5);
}

// The attempt with Gaussian modes (cmf2_base.f90:76-203): the coefficients
// are propagated as B = D^H S A in the basis orthonormalised by D (D0 from
// the step's start, carried by the SPF propagation), A = D B at the half
// step and, committed only when accepted, at the end.  Same device-side
// decisions as the grid attempt.
void cmf2_attempt_enqueue_gwp(mg_ctx *c, int kmax) {
    const int64_t N = c->nconf, D = c->dimpsi;
    const int64_t nD = (int64_t)c->nspfmax * c->nspfmax * c->nm * c->ns;
    double2 *A = c->dA.as<double2>(), *psi = c->dpsi.as<double2>(), *AH = c->dAH.as<double2>();
    double2 *B0 = c->dB0.as<double2>(), *BH = c->dBH.as<double2>(), *B02 = c->dB02.as<double2>();
    double2 *pH = c->dpH.as<double2>(), *pH2 = c->dpH2.as<double2>();
    double2 *DD = c->dDD.as<double2>(), *DH = c->dDH.as<double2>(), *DH2 = c->dDH2.as<double2>();
    AttemptDev *at = att_dev(c);
    LanczosDev *L1 = lan_dev(c, 0), *L2 = lan_dev(c, 1);
    launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
    calc_h1mode(c, psi);                 // S00(psi) and h1mode
    calc_rho(c, A);
    calc_meanfield(c, A, true);
    prepare_ha(c, DD);
    a_to_b(c, A, B0, DD);
    copy_d2d(c, pH, psi, D);
    copy_d2d(c, DH, DD, nD);
    propagate_spf_dev(c, pH, 0.5, attempt_fork(c), DH);
    lanczos_dev(c, B0, BH, L1, 0.5, true, kmax);
    attempt_join(c);
    b_to_a(c, BH, AH, DH);
    calc_h1mode(c, pH);
    calc_rho(c, AH);
    calc_meanfield(c, AH, true);
    copy_d2d(c, pH2, psi, D);
    copy_d2d(c, DH2, DD, nD);
    propagate_spf_dev(c, pH2, 0.5, c->stream, DH2);
    delta_psi_overlaps(c, pH, pH2);
    delta_psi_contract_dev(c, AH);
    propagate_spf_dev(c, pH2, 0.5, c->stream, DH2);
    calc_h1mode(c, pH2);
    prepare_ha(c, DH2);
    lanczos_dev(c, BH, B02, L2, -0.5, true, kmax);
    double *n2 = c->dscal.as<double>() + 12;
    launch_diff_norm2(B02, B0, N, c->dpartials.as<double>(), c->nblocks_red, n2, c->stream);
    launch_attempt_decide(at, L1, L2, n2, c->eps_int, c->dstatus.as<int>(), c->stream);
    lanczos_dev(c, BH, B0, L2, 0.5, false, kmax);
    b_to_a(c, B0, AH, DH2);              // A of the accepted step, committed below
    launch_copy_if(A, AH, N, &at->accept, c->stream);
    launch_copy_if(psi, pH2, D, &at->accept, c->stream);
    launch_copy_if(DD, DH2, nD, &at->accept, c->stream);
    HIP_CHECK(hipMemcpyAsync(c->att_host, at, sizeof(AttemptDev), hipMemcpyDeviceToHost, c->stream));
    check_launch();
}

// One attempt: eagerly on the first attempts (they build every launch plan
// and size the Gram partial buffer), then through its captured graph.
void cmf2_attempt_run(mg_ctx *c, int kmax) {
    const bool graphs = c->use_graphs && c->att_graphs && !c->phases && !c->prof;
    if (c->att_exec && (c->att_epoch != c->graph_epoch || c->att_kmax != kmax || !graphs)) {
        HIP_CHECK(hipGraphExecDestroy(c->att_exec));
        c->att_exec = nullptr;
    }
    auto enqueue = [&] {
        if (c->ngwp) cmf2_attempt_enqueue_gwp(c, kmax);
        else cmf2_attempt_enqueue(c, kmax);
    };
    if (!graphs || c->att_runs < 1) {
        enqueue();
        c->att_runs++;
        return;
    }
    if (!c->att_exec) {
        const int epoch = c->graph_epoch;
        hipGraph_t g = nullptr;
        c->capturing = true;
        HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue();
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(c->stream, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            c->capturing = false;
            throw;
        }
        c->capturing = false;
        HIP_CHECK(hipStreamEndCapture(c->stream, &g));
        if (c->graph_epoch != epoch) {   // a buffer was reallocated while enqueuing
            (void)hipGraphDestroy(g);
            throw HipError{"attempt graph: launch plans changed during capture"};
        }
        HIP_CHECK(hipGraphInstantiate(&c->att_exec, g, nullptr, nullptr, 0));
        HIP_CHECK(hipGraphDestroy(g));
        c->att_epoch = epoch;
        c->att_kmax = kmax;
    }
    HIP_CHECK(hipGraphLaunch(c->att_exec, c->stream));
}

void cmf2_step(mg_ctx *c, double tt, double htry, double *hdid_out, double *hnext_out, std::vector<mg_attempt> *log) {
    if (c->ngwp) {
        // the step's D0 from the Cholesky factor of S00 (cmf2_base.f90:61-66)
        launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
        gwp_s00(c, c->dpsi.as<double2>());
        gwp_dd_init(c, c->dDD.as<double2>());
        HIP_CHECK(hipMemcpyAsync(c->spf_status_host, c->dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        spf_check(c);
    }
    if (!c->datt.p) c->datt.alloc(sizeof(AttemptDev) + 64 + 2 * sizeof(LanczosDev));
    const double safe = 0.95;
    double hdid = htry;
    bool full = false;
    for (;;) {
        int forced = -1;
        const size_t rpos = c->replay_pos;
        if (c->replay_pos < c->replay.size()) {
            const mg_attempt &r = c->replay[c->replay_pos++];
            hdid = r.h;
            forced = (int)r.accepted;
        }
        // iterations enqueued per Krylov build: the reference's limit on the
        // first (eager) attempt and after a build that needed more, else two
        // beyond the largest Krylov size seen (cmf2_integration.f90:57-64)
        const int kmax = (c->att_runs == 0 || full)
                             ? kNKrylovMax - 1
                             : std::min({kNKrylovMax - 1, c->nk_seen + 2, c->att_kmax_cap});
        full = false;
        if (c->phases) phase(c, 5);
        c->att_host->h = hdid;
        c->att_host->forced = forced;
        HIP_CHECK(hipMemcpyAsync(att_dev(c), c->att_host, 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        cmf2_attempt_run(c, kmax);
        HIP_CHECK(hipStreamSynchronize(c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream2));
        const AttemptDev r = *c->att_host;
        if (r.spf_status) {
            *c->spf_status_host = r.spf_status;
            spf_check(c);
        }
        if (r.overflow == 1) throw HipError{"Maximum number of Krylov vectors reached"};
        if (r.overflow == 2) {
            // the Krylov space outgrew the enqueued iterations: the attempt
            // committed nothing; repeat it with the reference's limit
            full = true;
            c->replay_pos = rpos;
            continue;
        }
        c->nk_seen = std::max(c->nk_seen, std::max(r.nk1, r.nk2));
        const bool reject = !r.accept;
        if (log) log->push_back({tt, hdid, r.err_A, r.err_psi, (int64_t)(!reject), r.nk1, r.nk2});
        note_attempt(c, hdid, r.error);
        if (reject) {
            hdid = hdid / std::sqrt(r.error) * safe;
            if (hdid < c->min_stepsize) throw HipError{"CMF2 step size fell below minstep"};
            continue;
        }
        *hnext_out = propose_next(c, hdid, r.error);
        if (c->phases) phase(c, 5);
        break;
    }
    *hdid_out = hdid;
    HIP_CHECK(hipStreamSynchronize(c->stream));
}

// Energy with Gaussian modes: <A|H|A> = <B|D^H h D|B> with D0 from the
// Cholesky factor of the current overlaps and B = D0^H S A (D0 D0^H = S^-1).
double energy_gwp(mg_ctx *c, const double2 *A) {
    double2 *Dobs = c->dDobs.as<double2>(), *Bo = c->dIn.as<double2>(), *HB = c->dOut.as<double2>();
    gwp_dd_init(c, Dobs);
    prepare_ha(c, Dobs);
    a_to_b(c, A, Bo, Dobs);
    if (c->prec) {
        launch_convert(c->dXf.p, 1, Bo, 0, c->nconf, c->stream);
        apply_h(c, c->dXf.p, c->dYf.p);
        return dot(c, c->dXf.p, c->dYf.p, c->nconf, 1).x;
    }
    apply_h(c, Bo, HB);
    return dot(c, Bo, HB, c->nconf).x;
}

// WriteOutput observables (output.f90:33-87, 143-150).
// A wavefunction that is no longer finite ends the run with an error at the
// next output time (the reference stops when LAPACK reports info != 0 on
// such data, maths.f90:37-46).
void check_finite(const mg_output &o, double t) {
    if (!std::isfinite(o.norm) || !std::isfinite(o.energy) || !std::isfinite(o.auto_re) ||
        !std::isfinite(o.auto_im))
        throw HipError{"the wavefunction is not finite at t = " + std::to_string(t) + " au"};
}

void observables(mg_ctx *c, double t, mg_output *out, double *pops) {
    double2 *A = c->dA.as<double2>(), *psi = c->dpsi.as<double2>();
    if (c->ngwp) gwp_s00(c, psi);
    calc_rho(c, A, true);
    calc_h1mode(c, psi);
    double energy;
    if (c->ngwp) {
        energy = energy_gwp(c, A);
    } else if (c->prec) {
        launch_convert(c->dAf.p, 1, A, 0, c->nconf, c->stream);
        apply_h(c, c->dAf.p, c->dYf.p);
        energy = dot(c, c->dAf.p, c->dYf.p, c->nconf, 1).x;
    } else {
        apply_h(c, A, c->dOut.p);
        energy = dot(c, A, c->dOut.p, c->nconf).x;
    }
    const size_t nn = (size_t)c->nspfmax * c->nspfmax;
    std::vector<double2> r(nn * c->nm * c->ns);
    HIP_CHECK(hipMemcpyAsync(r.data(), c->drho.p, sizeof(double2) * r.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    std::vector<double2> sd;
    if (c->ngwp && c->gwp[0]) {
        sd.resize(nn * c->nm * c->ns);
        HIP_CHECK(hipMemcpyAsync(sd.data(), c->dSd.p, sizeof(double2) * sd.size(), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    double tot = 0.0;
    const size_t ld = (size_t)c->nspfmax;
    for (int s = 0; s < c->ns; s++) {
        const int jel = c->MultiSet[0] ? s : 0;
        const int64_t n = c->NS(0, jel);
        double pop = 0.0;
        if (!sd.empty()) {
            // Gaussian first mode: Re sum rho o S00 (output.f90:64-74)
            const double2 *rr = r.data() + nn * (0 + (size_t)c->nm * s), *S = sd.data() + nn * (0 + (size_t)c->nm * s);
            for (int64_t j2 = 0; j2 < n; j2++) {
                for (int64_t j1 = 0; j1 < j2; j1++) pop += 2.0 * cmul(rr[j1 + ld * j2], S[j1 + ld * j2]).x;
                pop += cmul(rr[j2 + ld * j2], S[j2 + ld * j2]).x;
            }
        } else
        for (int64_t i = 0; i < n; i++) pop += r[nn * (0 + (size_t)c->nm * s) + i + (size_t)c->nspfmax * i].x;
        if (pops) pops[s] = pop;
        tot += pop;
    }
    const cplx ct = wf_ovl(c, c->dAstart.as<double2>(), c->dpsistart.as<double2>(), A, psi);
    out->time = t;
    out->auto_re = ct.real();
    out->auto_im = ct.imag();
    out->norm = tot;
    out->energy = energy / tot;
}

// ----------------------------------------------------------------------------
// Variable mean field (VMF) integration, DVR modes (derivatives.f90:39-243,
// integrator.f90:36-221, propagator.f90:108-140).
// Cash-Karp (1990) and Dormand-Prince (1980) pairs, as in integrator.f90:108-194.
const double kCkB[6][5] = {{0, 0, 0, 0, 0}, {0.2, 0, 0, 0, 0}, {0.075, 0.225, 0, 0, 0}, {0.3, -0.9, 1.2, 0, 0},
                           {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0},
                           {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
const double kCkC[7] = {37.0 / 378.0, 0, 250.0 / 621.0, 125.0 / 594.0, 0, 512.0 / 1771.0, 0};
const double kCkD[7] = {2825.0 / 27648.0, 0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 0.25, 0};
const double kDpB[6][5] = {{0, 0, 0, 0, 0}, {0.2, 0, 0, 0, 0}, {0.075, 0.225, 0, 0, 0},
                           {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0},
                           {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0},
                           {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0}};
const double kDpC[7] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0};
const double kDpD[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0,
                        0.025};

void vmf_alloc(mg_ctx *c) {
    if (c->vmf_ready) return;
    const size_t N = (size_t)c->nconf, D = (size_t)c->dimpsi, z = sizeof(double2);
    for (int q = 0; q < 7; q++) {
        c->dvA[q].alloc(c->es * N);
        c->dvP[q].alloc(z * D);
    }
    c->dvAt.alloc(z * N);
    c->dvAe.alloc(z * N);
    c->dvPt.alloc(z * D);
    c->dvPe.alloc(z * D);
    c->vmf_ready = true;
}

// Calc_Derivs at (A, psi): kA <- H A (work type; dA = der_factor kA, applied
// in the stage combinations), kp <- dpsi of every mode.  Failures of the SPF
// projector solve accumulate in dstatus.
// Calc_Derivs with Gaussian modes (derivatives.f90:61-89, 155-238).  kA
// receives dA / der_factor = S^-1 H_g A - i sum_m tau_m A (vmf_combine
// multiplies by der_factor dt): H_g is H in the Gaussian basis (S00 for the
// identity on Gaussian modes), S^-1 the Tikhonov S00^-1 on every Gaussian
// mode, tau_m of Calc_Der_SPF; the Gaussian-parameter derivatives come from
// k_gwp_der with D = 1 (its dD output is then -tau).
void calc_derivs_gwp(mg_ctx *c, const double2 *A, const double2 *psi, void *kA, double2 *kp) {
    const int64_t N = c->nconf;
    calc_h1mode(c, psi);
    calc_rho(c, A);
    calc_meanfield(c, A, true);
    launch_fill(kp, c->dimpsi, make_double2(0, 0), c->stream);
    SpfParams P = spf_params(c, 0.0);
    if (c->ndvr_desc > 0)
        launch_der_spf(c->dspfmodes.as<SpfModeDesc>(), -c->ndvr_desc, psi, kp, c->dspfscr.as<double2>(), P,
                       dof_tables(c), c->dstatus.as<int>(), c->stream);
    const GwpParams G = gwp_params(c);
    for (int i = 0; i < c->ngwp_desc; i++)
        launch_gwp_der(c->dgwpdesc.as<GwpDesc>(), i, psi, c->dDid.as<double2>(), kp, c->dTauSet.as<double2>(),
                       c->dgwpscr.as<double2>(), G, c->dstatus.as<int>(), c->stream);
    launch_gwp_state_mats(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, c->dTauSet.as<double2>(), -1.0, G,
                          c->dTauS.as<double2>(), c->stream);
    launch_gwp_s00m1(c->dgwpdesc.as<GwpDesc>(), c->ngwp_desc, c->dSd.as<double2>(), G, c->dSm1.as<double2>(),
                     c->stream);
    double2 *HgA = c->dB0.as<double2>(), *X = c->dBH.as<double2>(), *Z = c->dB02.as<double2>();
    apply_hg(c, c->prec ? c->dAf.p : (const void *)A, kA);
    if (c->prec) launch_convert(HgA, 0, kA, 1, N, c->stream);
    else copy_d2d(c, HgA, static_cast<const double2 *>(kA), N);
    apply_kron(c, HgA, X, c->dSm1.as<double2>(), c->gwp_mask);
    for (int m = 0; m < c->nm; m++) {
        if (!c->gwp[m]) continue;
        apply_kron(c, A, Z, c->dTauS.as<double2>(), 1u << m);
        launch_cscale(Z, Z, make_double2(0.0, -1.0), N, c->stream);
        launch_axpby(X, Z, 1.0, 1.0, N, 0, c->stream);
    }
    if (c->prec) launch_convert(kA, 1, X, 0, N, c->stream);
    else copy_d2d(c, static_cast<double2 *>(kA), X, N);
    check_launch();
}

void calc_derivs(mg_ctx *c, const double2 *A, const double2 *psi, void *kA, double2 *kp) {
    if (c->ngwp) {
        calc_derivs_gwp(c, A, psi, kA, kp);
        return;
    }
    calc_rho(c, A);
    calc_h1mode(c, psi);
    calc_meanfield(c, A, true);
    apply_h(c, c->prec ? c->dAf.p : (const void *)A, kA);
    SpfParams P = spf_params(c, 0.0);
    launch_der_spf(c->dspfmodes.as<SpfModeDesc>(), -c->nm, psi, kp, c->dspfscr.as<double2>(), P, dof_tables(c),
                   c->dstatus.as<int>(), c->stream);
    check_launch();
}

// y_out = y0 + sum_q scale * w[q] k[q] for the A part (scale = dt der_factor)
// and the psi part (scale = dt).
void vmf_combine(mg_ctx *c, const double *w, int nq, double dt, double2 *Aout, double2 *Pout) {
    const void *kA[7], *kp[7];
    double2 wA[7], wP[7];
    for (int q = 0; q < nq; q++) {
        kA[q] = c->dvA[q].p;
        kp[q] = c->dvP[q].p;
        wA[q] = make_double2(0.0, -dt * w[q]);   // dt w der_factor, der_factor = -i
        wP[q] = make_double2(dt * w[q], 0.0);
    }
    launch_rk_combine(Aout, c->dA.as<double2>(), kA, wA, nq, c->nconf, c->prec, c->stream);
    launch_rk_combine(Pout, c->dpsi.as<double2>(), kp, wP, nq, c->dimpsi, 0, c->stream);
}

// rkck / rkdp (integrator.f90:95-221) from the resident (A, psi); stage 0's
// derivatives are in dvA[0], dvP[0]; Dormand-Prince leaves those of the
// 5th-order point in dvA[6], dvP[6].
void vmf_rk(mg_ctx *c, bool dp, double dt) {
    double2 *At = c->dvAt.as<double2>(), *Pt = c->dvPt.as<double2>();
    for (int st = 1; st < 6; st++) {
        vmf_combine(c, dp ? kDpB[st] : kCkB[st], st, dt, At, Pt);
        calc_derivs(c, At, Pt, c->dvA[st].p, c->dvP[st].as<double2>());
    }
    vmf_combine(c, dp ? kDpC : kCkC, 6, dt, At, Pt);
    if (dp) {
        calc_derivs(c, At, Pt, c->dvA[6].p, c->dvP[6].as<double2>());
        vmf_combine(c, kDpD, 7, dt, c->dvAe.as<double2>(), c->dvPe.as<double2>());
    } else {
        vmf_combine(c, kCkD, 6, dt, c->dvAe.as<double2>(), c->dvPe.as<double2>());
    }
}

void spf_status_check(mg_ctx *c) {
    int st = 0;
    HIP_CHECK(hipMemcpyAsync(&st, c->dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (st == 1) throw HipError{"SPF overlap matrix not positive definite"};
}

// rkqs (integrator.f90:36-91): error = DeltaPsi(5th, 4th order) / eps_int.
void vmf_rkqs(mg_ctx *c, bool dp, double htry, double tt, double *hdid, double *hnext,
              std::vector<mg_attempt> *log) {
    const double safety = 0.9, pgrow = -0.2, pshrnk = -0.25, minscal = 0.1, maxscal = 5.0;
    const double errcon = std::pow(5.0 / safety, 1.0 / pgrow);
    const double minstep = c->min_stepsize;
    double dt = htry, errmax = 0.0;
    for (;;) {
        int forced = -1;
        if (c->replay_pos < c->replay.size()) {
            const mg_attempt &r = c->replay[c->replay_pos++];
            dt = r.h;
            forced = (int)r.accepted;
        }
        launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
        vmf_rk(c, dp, dt);
        delta_psi_overlaps(c, c->dvPt.as<double2>(), c->dvPe.as<double2>());
        errmax = delta_psi_contract(c, c->dvAt.as<double2>(), c->dvAe.as<double2>()) / c->eps_int;
        spf_status_check(c);
        const bool accept = (forced >= 0) ? (forced == 1) : (errmax < 1.0 || dt <= minstep);
        if (log) log->push_back({tt, dt, errmax * c->eps_int, 0.0, (int64_t)accept, 0, 0});
        if (accept) break;
        dt = std::max(minscal * dt, safety * dt * std::pow(errmax, pshrnk));
        dt = std::max(dt, minstep);
    }
    double hn = (errmax > errcon) ? safety * dt * std::pow(errmax, pgrow) : dt * maxscal;
    hn = std::max(hn, minstep);
    hn = std::min(hn, c->writetime);
    *hnext = hn;
    *hdid = dt;
    copy_d2d(c, c->dA.as<double2>(), c->dvAt.as<double2>(), c->nconf);
    copy_d2d(c, c->dpsi.as<double2>(), c->dvPt.as<double2>(), c->dimpsi);
    if (dp) {   // first-same-as-last: the new point's derivatives become stage 0
        std::swap(c->dvA[0].p, c->dvA[6].p);
        std::swap(c->dvP[0].p, c->dvP[6].p);
        // cached H.A plans are keyed by buffer address and stay valid for both
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
}

// Renormalisation of propagate_wf (propagator.f90:129-144, 288-305):
// A /= sqrt(Trace_wf(A, psi)), the norm with the SPF overlaps
// (gmctdh_objects.f90:184-202, one packet).  Mode 3 after every step that
// does not end an output interval; every nonzero mode at the end of an
// interval, mode 1 only once.
void renormalise(mg_ctx *c) {
    const double n = norm_wf(c, c->dA.as<double2>(), c->dpsi.as<double2>());
    if (!(n > 0.0) || !std::isfinite(n)) throw HipError{"cannot renormalise: the wavefunction norm is not positive"};
    launch_axpby(c->dA.p, c->dA.p, 1.0 / std::sqrt(n), 0.0, c->nconf, 0, c->stream);
    check_launch();
}
void renormalise_interval(mg_ctx *c) {
    if (c->inorm_psi == 0) return;
    renormalise(c);
    if (c->inorm_psi == 1) c->inorm_psi = 0;
}

float *noalloc_scratch();   // the no-allocation entry point's device array (below)

template <class F>
int guarded(mg_ctx *c, F &&f) {
    if (!c) return -1;
    struct Scope {
        mg_ctx *prev;
        explicit Scope(mg_ctx *c) : prev(tl_ctx) { tl_ctx = c; }
        ~Scope() { tl_ctx = prev; }
    } scope(c);
    // On an error, wait for both streams before returning: a step can throw
    // (Krylov overflow, minimum step) while the SPF propagation it forked is
    // still running on stream2, and a retry on the same context must not
    // overlap it.
    auto drain = [c] {
        if (c->stream2) (void)hipStreamSynchronize(c->stream2);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
    };
    try {
        f();
        return 0;
    } catch (const HipError &e) {
        c->err = e.msg;
        drain();
        return 1;
    } catch (const std::exception &e) {
        c->err = e.what();
        drain();
        return 2;
    }
}

}  // namespace

// ============================================================================
// C-ABI
extern "C" {

int mg_create(const mg_problem *prob, int64_t precision, mg_ctx **out) {
    if (!out || !prob) return -1;
    *out = nullptr;
    mg_ctx *c = new mg_ctx();
    c->prec = (int)precision;
    c->es = precision ? sizeof(float2) : sizeof(double2);
    int rc = guarded(c, [&] {
        if (precision != 0 && precision != 1)
            throw HipError{"precision must be 0 (complex FP64) or 1 (mixed complex FP32/FP64)"};
        // host tables and the descriptor's limits first (setup makes no HIP
        // call): an unsupported problem is rejected before the device is used
        setup(c, prob);
        int ndev = 0;
        HIP_CHECK(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw HipError{"no HIP device"};
        HIP_CHECK(hipStreamCreate(&c->stream));
        HIP_CHECK(hipStreamCreate(&c->stream2));
        (void)noalloc_scratch();   // the device address of the no-allocation scratch, once per device
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        HIP_CHECK(hipHostMalloc((void **)&c->spf_status_host, sizeof(int), hipHostMallocDefault));
        *c->spf_status_host = 0;
        if (const char *e = std::getenv("MG_GRAPHS")) c->use_graphs = std::atoi(e) != 0;
        if (const char *e = std::getenv("MG_PHASES")) c->phases = std::atoi(e) != 0;
        if (const char *e = std::getenv("MG_ATTEMPT_GRAPH")) c->att_graphs = std::atoi(e) != 0;
        if (const char *e = std::getenv("MG_ATTEMPT_KMAX")) c->att_kmax_cap = std::max(2, std::atoi(e));
        if (const char *e = std::getenv("MG_ATTEMPT_FORK")) c->att_fork = std::atoi(e) != 0;
        if (const char *e = std::getenv("MG_DPSI_FUSED")) c->dps_fused = std::atoi(e) != 0;
        HIP_CHECK(hipHostMalloc((void **)&c->att_host, sizeof(AttemptDev), hipHostMallocDefault));
        std::memset(c->att_host, 0, sizeof(AttemptDev));
        c->phase_t0 = std::chrono::steady_clock::now();
        allocate(c);
        c->stepsize = c->writetime * c->eps_int;  // propagator.f90:103
    });
    if (rc != 0) {
        fprintf(stderr, "mg_create: %s\n", c->err.c_str());
        *out = c;  // caller can read the error, then destroy
        return rc;
    }
    *out = c;
    return 0;
}

int mg_destroy(mg_ctx *c) {
    if (!c) return 0;
    if (c->phases) {
        fprintf(stderr, "[mg] CMF2 phases (ms):");
        for (int i = 0; i < 6; i++) fprintf(stderr, " %s %.1f", kPhaseNames[i], c->phase_ms[i]);
        fprintf(stderr, "\n");
    }
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (auto e : c->prof_ev) (void)hipEventDestroy(e);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->spf_status_host) (void)hipHostFree(c->spf_status_host);
    if (c->att_exec) (void)hipGraphExecDestroy(c->att_exec);
    if (c->att_host) (void)hipHostFree(c->att_host);
    hipStream_t s = c->stream, s2 = c->stream2;
    delete c;
    if (s) (void)hipStreamDestroy(s);
    if (s2) (void)hipStreamDestroy(s2);
    return 0;
}

const char *mg_last_error(const mg_ctx *c) { return c ? c->err.c_str() : "null context"; }

int mg_dims(const mg_ctx *c, int64_t *nspfmax, int64_t *nopmax) {
    if (!c) return -1;
    if (nspfmax) *nspfmax = c->nspfmax;
    if (nopmax) *nopmax = c->nopmax;
    return 0;
}

int mg_set_replay(mg_ctx *c, const mg_attempt *log, int64_t n) {
    if (!c || (n > 0 && !log)) return -1;
    c->replay.assign(log, log + (n > 0 ? n : 0));
    c->replay_pos = 0;
    return 0;
}

int mg_stepsize(const mg_ctx *c, double *h) {
    if (!c || !h) return -1;
    *h = c->stepsize;
    return 0;
}

int mg_set_stepsize(mg_ctx *c, double h) {
    return guarded(c, [&] {
        if (!(std::isfinite(h) && h > 0.0)) throw HipError{"step size must be finite and > 0"};
        c->stepsize = h;
    });
}

int mg_apply_h_profile(mg_ctx *c, const void *dB, void *dHB, double *kernel_ms, int64_t *nlaunch) {
    return guarded(c, [&] {
        c->prof = true;
        c->prof_ms = 0.0;
        c->prof_launches = 0;
        const void *in = dB ? dB : (c->prec ? c->dXf.p : c->dIn.p);
        void *out = dHB ? dHB : (c->prec ? c->dYf.p : c->dOut.p);
        try {
            apply_h(c, in, out);
        } catch (...) {
            c->prof = false;
            throw;
        }
        c->prof = false;
        if (kernel_ms) *kernel_ms = c->prof_ms;
        if (nlaunch) *nlaunch = c->prof_launches;
    });
}
int mg_profile_copy(mg_ctx *c, int64_t n, double *kernel_ms) {
    return guarded(c, [&] {
        if (n < 0 || n > c->nconf) throw HipError{"copy size out of range"};
        void *src = c->prec ? c->dXf.p : c->dIn.p;
        void *dst = c->prec ? c->dYf.p : c->dOut.p;
        HIP_CHECK(hipEventRecord(c->ev0, c->stream));
        launch_copy(dst, src, n, c->prec, c->stream);
        HIP_CHECK(hipEventRecord(c->ev1, c->stream));
        HIP_CHECK(hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (kernel_ms) *kernel_ms = ms;
    });
}
int64_t mg_nconf(const mg_ctx *c) { return c ? c->nconf : -1; }
int64_t mg_dimpsi(const mg_ctx *c) { return c ? c->dimpsi : -1; }

int mg_set_state(mg_ctx *c, const double *A, const double *psi) {
    return guarded(c, [&] {
        hcopy(c->dA.p, A, sizeof(double2) * c->nconf, hipMemcpyHostToDevice);
        hcopy(c->dpsi.p, psi, sizeof(double2) * c->dimpsi, hipMemcpyHostToDevice);
    });
}

int mg_get_state(mg_ctx *c, double *A, double *psi) {
    return guarded(c, [&] {
        HIP_CHECK(hipStreamSynchronize(c->stream));
        if (A) hcopy(A, c->dA.p, sizeof(double2) * c->nconf, hipMemcpyDeviceToHost);
        if (psi) hcopy(psi, c->dpsi.p, sizeof(double2) * c->dimpsi, hipMemcpyDeviceToHost);
    });
}

int mg_set_start(mg_ctx *c, const double *A, const double *psi) {
    return guarded(c, [&] {
        hcopy(c->dAstart.p, A, sizeof(double2) * c->nconf, hipMemcpyHostToDevice);
        hcopy(c->dpsistart.p, psi, sizeof(double2) * c->dimpsi, hipMemcpyHostToDevice);
    });
}

int mg_calc_h1mode(mg_ctx *c, double *h1_out) {
    return guarded(c, [&] {
        calc_h1mode(c, c->dpsi.as<double2>());
        if (h1_out) {
            HIP_CHECK(hipMemcpyAsync(h1_out, c->dh1.p, c->dh1.bytes, hipMemcpyDeviceToHost, c->stream));
        }
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int mg_calc_rho(mg_ctx *c, double *rho_out) {
    return guarded(c, [&] {
        if (c->ngwp) gwp_s00(c, c->dpsi.as<double2>());
        calc_rho(c, c->dA.as<double2>());
        if (rho_out) {
            const size_t nn = (size_t)c->nspfmax * c->nspfmax;
            HIP_CHECK(hipMemcpyAsync(rho_out, c->drho.p, sizeof(double2) * nn * c->nm * c->ns, hipMemcpyDeviceToHost,
                                     c->stream));
        }
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int mg_calc_meanfield(mg_ctx *c, double *mf_out, double *rmf_out) {
    return guarded(c, [&] {
        calc_meanfield(c, c->dA.as<double2>());
        if (mf_out) HIP_CHECK(hipMemcpyAsync(mf_out, c->dmf.p, c->dmf.bytes, hipMemcpyDeviceToHost, c->stream));
        if (rmf_out) HIP_CHECK(hipMemcpyAsync(rmf_out, c->drmf.p, c->drmf.bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int mg_apply_h(mg_ctx *c, const double *B, double *HB) {
    return guarded(c, [&] {
        HIP_CHECK(hipMemcpyAsync(c->dIn.p, B, sizeof(double2) * c->nconf, hipMemcpyHostToDevice, c->stream));
        if (c->prec) {
            launch_convert(c->dXf.p, 1, c->dIn.p, 0, c->nconf, c->stream);
            apply_h(c, c->dXf.p, c->dYf.p);
            launch_convert(c->dOut.p, 0, c->dYf.p, 1, c->nconf, c->stream);
        } else {
            apply_h(c, c->dIn.p, c->dOut.p);
        }
        HIP_CHECK(hipMemcpyAsync(HB, c->dOut.p, sizeof(double2) * c->nconf, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int mg_prepare_gwp(mg_ctx *c) {
    return guarded(c, [&] {
        if (!c->ngwp) return;
        launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
        calc_h1mode(c, c->dpsi.as<double2>());
        gwp_dd_init(c, c->dDD.as<double2>());
        prepare_ha(c, c->dDD.as<double2>());
        HIP_CHECK(hipMemcpyAsync(c->spf_status_host, c->dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        spf_check(c);
    });
}

int mg_gwp_overlaps(mg_ctx *c, double *s00_out, double *dd_out) {
    return guarded(c, [&] {
        if (!c->ngwp) throw HipError{"no Gaussian modes"};
        launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
        gwp_s00(c, c->dpsi.as<double2>());
        gwp_dd_init(c, c->dDD.as<double2>());
        if (s00_out) HIP_CHECK(hipMemcpyAsync(s00_out, c->dS00.p, c->dS00.bytes, hipMemcpyDeviceToHost, c->stream));
        if (dd_out) HIP_CHECK(hipMemcpyAsync(dd_out, c->dDD.p, c->dDD.bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipMemcpyAsync(c->spf_status_host, c->dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        spf_check(c);
    });
}

int mg_der_spf_gwp(mg_ctx *c, int64_t im, const double *psi_in, const double *dd_in, double *dpsi_out,
                   double *ddd_out) {
    return guarded(c, [&] {
        if (im < 1 || im > c->nm || !c->gwp[im - 1]) throw HipError{"mode is not a Gaussian mode"};
        int idesc = 0;
        for (int m = 0; m < im - 1; m++) idesc += c->gwp[m];
        double2 *A = c->dA.as<double2>(), *psi = c->dpsi.as<double2>();
        calc_h1mode(c, psi);
        calc_rho(c, A, true);
        calc_meanfield(c, A);
        double2 *pin = c->dpH.as<double2>(), *pout = c->dpH2.as<double2>();
        double2 *din = c->dDH.as<double2>(), *dout = c->dDH2.as<double2>();
        HIP_CHECK(hipMemcpyAsync(pin, psi_in, sizeof(double2) * c->dimpsi, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(din, dd_in, c->dDH.bytes, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemsetAsync(pout, 0, sizeof(double2) * c->dimpsi, c->stream));
        HIP_CHECK(hipMemsetAsync(dout, 0, c->dDH2.bytes, c->stream));
        launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
        launch_gwp_der(c->dgwpdesc.as<GwpDesc>(), idesc, pin, din, pout, dout, c->dgwpscr.as<double2>(),
                       gwp_params(c), c->dstatus.as<int>(), c->stream);
        check_launch();
        HIP_CHECK(hipMemcpyAsync(dpsi_out, pout, sizeof(double2) * c->dimpsi, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipMemcpyAsync(ddd_out, dout, c->dDH2.bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipMemcpyAsync(c->spf_status_host, c->dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        spf_check(c);
    });
}

int mg_a_to_b(mg_ctx *c, const double *A, const double *dd, double *B) {
    return guarded(c, [&] {
        if (!c->ngwp) throw HipError{"no Gaussian modes"};
        double2 *din = c->dDH.as<double2>();
        HIP_CHECK(hipMemcpyAsync(c->dIn.p, A, sizeof(double2) * c->nconf, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(din, dd, c->dDH.bytes, hipMemcpyHostToDevice, c->stream));
        gwp_s00(c, c->dpsi.as<double2>());
        a_to_b(c, c->dIn.as<double2>(), c->dOut.as<double2>(), din);
        HIP_CHECK(hipMemcpyAsync(B, c->dOut.p, sizeof(double2) * c->nconf, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

// ---- learned step controller ----
static_assert(kMlpNParam == MG_MLP_NPARAM && kMlpNFeat == MG_MLP_NFEAT, "MLP parameter layout");
namespace {
struct MlpDev {
    DevBuf p, wb, X, t, y, work, grad, m, v;
    int cap = 0;
    MlpWork w;
    void reserve(int B) {
        if (B <= cap) return;
        cap = B;
        X.alloc(sizeof(float) * (size_t)kMlpNFeat * B);
        t.alloc(sizeof(float) * (size_t)B);
        y.alloc(sizeof(float) * (size_t)B);
        work.alloc(sizeof(float) * ((size_t)6 * kMlpHidden * B + 2 * (size_t)B));
        float *f = work.as<float>();
        const size_t hb = (size_t)kMlpHidden * B;
        w = MlpWork{f, f + hb, f + 2 * hb, f + 3 * hb, f + 4 * hb, f + 5 * hb, f + 6 * hb, f + 6 * hb + B};
    }
    void init_params(const float *host) {
        p.alloc(sizeof(float) * kMlpNParam);
        wb.alloc(sizeof(uint16_t) * kMlpNWeights);
        grad.alloc(sizeof(float) * kMlpNParam);
        m.alloc(sizeof(float) * kMlpNParam);
        v.alloc(sizeof(float) * kMlpNParam);
        hcopy(p.p, host, sizeof(float) * kMlpNParam, hipMemcpyHostToDevice);
        hzero(m.p, m.bytes);
        hzero(v.p, v.bytes);
    }
};
}  // namespace

int mg_mlp_forward(mg_ctx *c, const float *params, const float *X, int64_t n, float *y) {
    return guarded(c, [&] {
        MlpDev d;
        d.init_params(params);
        const int B = (int)std::min<int64_t>(std::max<int64_t>(n, 1), 4096);
        d.reserve(B);
        launch_mlp_bf16(d.p.as<float>(), d.wb.as<uint16_t>(), c->stream);
        for (int64_t o = 0; o < n; o += B) {
            const int b = (int)std::min<int64_t>(B, n - o);
            HIP_CHECK(hipMemcpyAsync(d.X.p, X + kMlpNFeat * o, sizeof(float) * kMlpNFeat * b, hipMemcpyHostToDevice,
                                     c->stream));
            launch_mlp_forward(d.X.as<float>(), d.p.as<float>(), d.wb.as<uint16_t>(), d.w, d.y.as<float>(), nullptr, b,
                               c->stream);
            HIP_CHECK(hipMemcpyAsync(y + o, d.y.p, sizeof(float) * b, hipMemcpyDeviceToHost, c->stream));
        }
        check_launch();
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

// ---------------------------------------------------------------------------
// mg_mlp_forward_noalloc: mg_mlp_forward without any device allocation.  Its
// device memory (parameters, BF16 weights, one batch of inputs, outputs and
// activations) is the file-scope array `scratch`, MG_MLP_NOALLOC_MAX_ROWS
// rows at most.  Host code never passes `scratch` itself to a kernel or a
// copy (in host code the name is a stand-in, not a device address): the
// device address comes from hipGetSymbolAddress, once per device, at the
// first mg_create on that device, and is kept per device.  Every call on a
// device shares that device's array, so calls that overlap on one GPU race.
namespace {
constexpr size_t kNaRows = MG_MLP_NOALLOC_MAX_ROWS;
constexpr size_t kNaAlign = 64;   // floats (256 bytes)
constexpr size_t na_round(size_t n) { return (n + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr size_t kNaParams = 0;
constexpr size_t kNaWb = kNaParams + na_round(kMlpNParam);
constexpr size_t kNaX = kNaWb + na_round((kMlpNWeights + 1) / 2);
constexpr size_t kNaY = kNaX + na_round((size_t)kMlpNFeat * kNaRows);
constexpr size_t kNaWork = kNaY + na_round(kNaRows);
constexpr size_t kNaTotal = kNaWork + na_round((size_t)6 * kMlpHidden * kNaRows + 2 * kNaRows);
}  // namespace

static __device__ __align__(256) float scratch[kNaTotal];

namespace {
constexpr int kNaMaxDevices = 64;
std::mutex na_mutex;
float *na_ptr[kNaMaxDevices] = {nullptr};

// Device address of `scratch` on the current device (looked up once per
// device; the lookup is made by mg_create, so a context's device has it).
float *noalloc_scratch() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) throw HipError{"device index out of range for the no-allocation scratch"};
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipError{"hipGetSymbolAddress returned no address for the no-allocation scratch"};
        na_ptr[dev] = static_cast<float *>(p);
    }
    return na_ptr[dev];
}
}  // namespace

int mg_mlp_forward_noalloc(mg_ctx *c, const float *params, const float *X, int64_t n, float *y) {
    return guarded(c, [&] {
        if (n < 1 || n > (int64_t)kNaRows)
            throw HipError{"mg_mlp_forward_noalloc: n = " + std::to_string(n) + " rows; between 1 and " +
                           std::to_string(kNaRows) + " rows (MG_MLP_NOALLOC_MAX_ROWS)"};
        if (!params || !X || !y) throw HipError{"mg_mlp_forward_noalloc: null pointer"};
        float *base = noalloc_scratch();
        float *dp = base + kNaParams, *dx = base + kNaX, *dy = base + kNaY, *f = base + kNaWork;
        uint16_t *dwb = reinterpret_cast<uint16_t *>(base + kNaWb);
        const int B = (int)n;
        const size_t hb = (size_t)kMlpHidden * B;
        const MlpWork w{f, f + hb, f + 2 * hb, f + 3 * hb, f + 4 * hb, f + 5 * hb, f + 6 * hb, f + 6 * hb + B};
        HIP_CHECK(hipMemcpyAsync(dp, params, sizeof(float) * kMlpNParam, hipMemcpyHostToDevice, c->stream));
        launch_mlp_bf16(dp, dwb, c->stream);
        HIP_CHECK(hipMemcpyAsync(dx, X, sizeof(float) * kMlpNFeat * B, hipMemcpyHostToDevice, c->stream));
        launch_mlp_forward(dx, dp, dwb, w, dy, nullptr, B, c->stream);
        check_launch();
        HIP_CHECK(hipMemcpyAsync(y, dy, sizeof(float) * B, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int mg_mlp_grad(mg_ctx *c, const float *params, const float *X, const float *t, int64_t n, float *grad,
                double *loss) {
    return guarded(c, [&] {
        if (n < 1 || n > 65536) throw HipError{"mg_mlp_grad: 1 <= n <= 65536"};
        MlpDev d;
        d.init_params(params);
        d.reserve((int)n);
        HIP_CHECK(hipMemcpyAsync(d.X.p, X, sizeof(float) * kMlpNFeat * n, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(d.t.p, t, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        launch_mlp_bf16(d.p.as<float>(), d.wb.as<uint16_t>(), c->stream);
        launch_mlp_forward(d.X.as<float>(), d.p.as<float>(), d.wb.as<uint16_t>(), d.w, d.y.as<float>(), d.t.as<float>(),
                           (int)n, c->stream);
        launch_mlp_backward_grad(d.X.as<float>(), d.p.as<float>(), d.wb.as<uint16_t>(), d.w, d.grad.as<float>(),
                                 (int)n, c->stream);
        check_launch();
        std::vector<float> e(n);
        const int off = 2 * kMlpNFeat;
        for (int i = 0; i < off; i++) grad[i] = 0.0f;
        HIP_CHECK(hipMemcpyAsync(grad + off, d.grad.p, sizeof(float) * (kMlpNParam - off), hipMemcpyDeviceToHost,
                                 c->stream));
        HIP_CHECK(hipMemcpyAsync(e.data(), d.w.err2, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        double sum = 0.0;
        for (float v : e) sum += v;
        if (loss) *loss = sum / (double)n;
    });
}

int mg_mlp_train(mg_ctx *c, const float *X, const float *t, int64_t n, int64_t epochs, int64_t batch, double lr,
                 uint64_t seed, float *params, double *loss_out) {
    return guarded(c, [&] {
        if (n < 2 || epochs < 1 || batch < 1 || batch > 65536) throw HipError{"mg_mlp_train: bad sizes"};
        const int F = kMlpNFeat, Hd = kMlpHidden;
        std::vector<float> P(kMlpNParam, 0.0f);
        // standardisation
        for (int i = 0; i < F; i++) {
            double mu = 0.0, sq = 0.0;
            for (int64_t r = 0; r < n; r++) mu += X[F * r + i];
            mu /= (double)n;
            for (int64_t r = 0; r < n; r++) sq += (X[F * r + i] - mu) * (X[F * r + i] - mu);
            P[i] = (float)mu;
            P[F + i] = (float)std::max(std::sqrt(sq / (double)n), 1e-3);
        }
        // Glorot-uniform weights, zero biases, output bias at the target mean
        std::mt19937_64 rng(seed);
        auto uni = [&](double a) { return (float)(a * (2.0 * ((double)(rng() >> 11) / 9007199254740992.0) - 1.0)); };
        int o = 2 * F;
        for (int i = 0; i < F * Hd; i++) P[o + i] = uni(std::sqrt(6.0 / (F + Hd)));
        o += F * Hd + Hd;
        for (int i = 0; i < Hd * Hd; i++) P[o + i] = uni(std::sqrt(6.0 / (Hd + Hd)));
        o += Hd * Hd + Hd;
        for (int i = 0; i < Hd; i++) P[o + i] = uni(std::sqrt(6.0 / (Hd + 1)));
        double tm = 0.0;
        for (int64_t r = 0; r < n; r++) tm += t[r];
        P[kMlpNParam - 1] = (float)(tm / (double)n);
        MlpDev d;
        d.init_params(P.data());
        const int B = (int)std::min<int64_t>(batch, n);
        d.reserve(B);
        DevBuf dXall, dTall;
        dXall.alloc(sizeof(float) * F * n);
        dTall.alloc(sizeof(float) * n);
        std::vector<int64_t> perm(n);
        for (int64_t r = 0; r < n; r++) perm[r] = r;
        std::vector<float> Xp((size_t)F * n), Tp(n);
        int step = 0;
        for (int64_t ep = 0; ep < epochs; ep++) {
            for (int64_t r = n - 1; r > 0; r--) std::swap(perm[r], perm[(int64_t)(rng() % (uint64_t)(r + 1))]);
            for (int64_t r = 0; r < n; r++) {
                std::memcpy(&Xp[(size_t)F * r], X + (size_t)F * perm[r], sizeof(float) * F);
                Tp[r] = t[perm[r]];
            }
            HIP_CHECK(hipMemcpyAsync(dXall.p, Xp.data(), sizeof(float) * F * n, hipMemcpyHostToDevice, c->stream));
            HIP_CHECK(hipMemcpyAsync(dTall.p, Tp.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
            for (int64_t b0 = 0; b0 < n; b0 += B) {
                const int b = (int)std::min<int64_t>(B, n - b0);
                const float *Xb = dXall.as<float>() + (size_t)F * b0, *Tb = dTall.as<float>() + b0;
                launch_mlp_bf16(d.p.as<float>(), d.wb.as<uint16_t>(), c->stream);
                launch_mlp_forward(Xb, d.p.as<float>(), d.wb.as<uint16_t>(), d.w, d.y.as<float>(), Tb, b, c->stream);
                launch_mlp_backward_grad(Xb, d.p.as<float>(), d.wb.as<uint16_t>(), d.w, d.grad.as<float>(), b,
                                         c->stream);
                launch_mlp_adam(d.p.as<float>(), d.grad.as<float>(), d.m.as<float>(), d.v.as<float>(), (float)lr,
                                ++step, c->stream);
            }
            check_launch();
            HIP_CHECK(hipStreamSynchronize(c->stream));
        }
        // final mean squared error over the training set
        launch_mlp_bf16(d.p.as<float>(), d.wb.as<uint16_t>(), c->stream);
        double sum = 0.0;
        std::vector<float> e(B);
        for (int64_t b0 = 0; b0 < n; b0 += B) {
            const int b = (int)std::min<int64_t>(B, n - b0);
            launch_mlp_forward(dXall.as<float>() + (size_t)F * b0, d.p.as<float>(), d.wb.as<uint16_t>(), d.w,
                               d.y.as<float>(), dTall.as<float>() + b0, b, c->stream);
            HIP_CHECK(hipMemcpyAsync(e.data(), d.w.err2, sizeof(float) * b, hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
            for (int i = 0; i < b; i++) sum += e[i];
        }
        if (loss_out) *loss_out = sum / (double)n;
        hcopy(params, d.p.p, sizeof(float) * kMlpNParam, hipMemcpyDeviceToHost);
    });
}

int mg_set_step_model(mg_ctx *c, const float *params) {
    return guarded(c, [&] {
        if (!params) {
            c->mlp_on = false;
            return;
        }
        c->dmlp_p.alloc(sizeof(float) * kMlpNParam);
        c->dmlp_wb.alloc(sizeof(uint16_t) * kMlpNWeights);
        c->dmlp_x.alloc(sizeof(float) * kMlpNFeat);
        c->dmlp_y.alloc(sizeof(float) * 4);
        c->dmlp_work.alloc(sizeof(float) * (6 * kMlpHidden + 2));
        hcopy(c->dmlp_p.p, params, sizeof(float) * kMlpNParam, hipMemcpyHostToDevice);
        launch_mlp_bf16(c->dmlp_p.as<float>(), c->dmlp_wb.as<uint16_t>(), c->stream);
        check_launch();
        HIP_CHECK(hipStreamSynchronize(c->stream));
        c->mlp_on = true;
        c->mlp_proposals = c->mlp_fallbacks = 0;
    });
}

// Ensemble of independent wavepackets.  Every member is a context of its own
// (buffers, two streams, launch plans and graphs), created, propagated and
// destroyed by one of `nthreads` host threads, so the device interleaves the
// kernels of up to nthreads members.  Members start largest first; a member
// is admitted only while the estimated device bytes of the members in flight
// stay within mem_fraction of the free device memory (one member is always
// admitted, so an oversized member runs alone).
namespace {
double member_bytes(const mg_problem *p, int64_t precision) {
    double nconf = 0.0;
    for (int64_t s = 0; s < p->nstates; s++) {
        double v = 1.0;
        for (int64_t m = 0; m < p->nmodes; m++) {
            const int64_t sm = p->MultiSet && p->MultiSet[m] ? s : 0;
            v *= (double)p->nSPF[m + p->nmodes * sm];
        }
        nconf += v;
    }
    // A-vectors: state, start, Krylov space (up to 21), H.A, partial sums
    // and mean-field scratch; 40 copies bounds what allocate() takes.
    return nconf * (precision ? 8.0 : 16.0) * 40.0 + 64.0 * 1024 * 1024;
}
}  // namespace

int mg_ensemble_propagate(const mg_problem *const *probs, const double *const *A0, const double *const *psi0,
                          int64_t n, int64_t precision, int64_t nthreads, double mem_fraction, int64_t flags,
                          double t0, double tfinal, mg_output *outs, double *pops, int64_t max_out, int64_t *n_out,
                          int64_t *status, double *member_wall, char *errbuf, int64_t errlen) {
    if (!probs || !A0 || !psi0 || n < 0 || !outs || !n_out || !status || max_out < 1) return -1;
    if (errbuf && errlen > 0) errbuf[0] = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return -1;
    const double budget = std::max(0.05, std::min(0.95, mem_fraction)) * (double)free_b;
    std::vector<int64_t> order((size_t)n);
    std::vector<double> est((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        order[(size_t)i] = i;
        est[(size_t)i] = member_bytes(probs[i], precision);
        status[i] = -1;
        n_out[i] = 0;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return est[a] > est[b]; });
    std::mutex mu;
    std::condition_variable cv;
    double in_flight = 0.0;
    int64_t active = 0, next = 0, done = 0;
    std::string first_err;
    auto worker = [&]() {
        if (hipSetDevice(dev) != hipSuccess) return;
        for (;;) {
            int64_t i;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] {
                    return next >= n || active == 0 || in_flight + est[(size_t)order[(size_t)next]] <= budget;
                });
                if (next >= n) return;
                i = order[(size_t)next];
                next++;
                active++;
                in_flight += est[(size_t)i];
            }
            const auto w0 = std::chrono::steady_clock::now();
            mg_ctx *c = nullptr;
            int rc = mg_create(probs[i], precision, &c);
            if (rc == 0) rc = mg_set_state(c, A0[i], psi0[i]);
            if (rc == 0) rc = mg_set_start(c, A0[i], psi0[i]);
            if (rc == 0 && (flags & 1)) {
                int64_t b = 0;
                double ms = 0.0;
                rc = mg_autotune(c, 1, &b, &ms);
            }
            std::string msg;
            try {
                std::vector<double> pp((size_t)max_out * (size_t)std::max<int64_t>(1, probs[i]->nstates));
                if (rc == 0)
                    rc = mg_propagate(c, t0, tfinal, outs + i * max_out, pp.data(), max_out, &n_out[i], nullptr, 0,
                                      nullptr);
                if (pops)
                    for (int64_t k = 0; k < n_out[i]; k++)
                        for (int64_t st = 0; st < probs[i]->nstates && st < 8; st++)
                            pops[((size_t)i * max_out + k) * 8 + st] = pp[(size_t)k * probs[i]->nstates + st];
                if (rc != 0 && c) msg = c->err;
            } catch (const std::exception &e) {   // host allocation failure
                rc = 2;
                msg = e.what();
            }
            if (c) mg_destroy(c);
            if (member_wall)
                member_wall[i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
            std::lock_guard<std::mutex> lk(mu);
            status[i] = rc;
            if (rc != 0 && first_err.empty())
                first_err = "member " + std::to_string(i) + ": " + (msg.empty() ? "creation failed" : msg);
            done++;
            if (flags & 2) {
                std::fprintf(stderr, "[ensemble] member %lld done (%lld of %lld), status %d, %.2f s\n", (long long)i,
                             (long long)done, (long long)n, rc,
                             std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count());
                std::fflush(stderr);
            }
            active--;
            in_flight -= est[(size_t)i];
            cv.notify_all();
        }
    };
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(nthreads, std::max<int64_t>(n, 1)));
    std::vector<std::thread> pool;
    for (int64_t k = 1; k < nt; k++) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    if (!first_err.empty() && errbuf && errlen > 0) {
        std::strncpy(errbuf, first_err.c_str(), (size_t)errlen - 1);
        errbuf[errlen - 1] = 0;
    }
    for (int64_t i = 0; i < n; i++)
        if (status[i] != 0) return 1;
    return 0;
}

int mg_set_graphs(mg_ctx *c, int64_t enable) {
    return guarded(c, [&] {
        HIP_CHECK(hipStreamSynchronize(c->stream));
        c->use_graphs = enable != 0;
    });
}

int mg_log_step_features(mg_ctx *c, int64_t enable) {
    return guarded(c, [&] {
        c->log_features = enable != 0;
        c->feat_log.clear();
        c->feat_h.clear();
    });
}

int mg_step_features(mg_ctx *c, float *feat_out, double *h_out, int64_t max, int64_t *n) {
    return guarded(c, [&] {
        const int64_t k = std::min<int64_t>(max, (int64_t)c->feat_h.size());
        if (feat_out) std::memcpy(feat_out, c->feat_log.data(), sizeof(float) * kMlpNFeat * k);
        if (h_out) std::memcpy(h_out, c->feat_h.data(), sizeof(double) * k);
        *n = k;
    });
}

int mg_controller_stats(mg_ctx *c, int64_t *proposals, int64_t *fallbacks) {
    return guarded(c, [&] {
        *proposals = c->mlp_proposals;
        *fallbacks = c->mlp_fallbacks;
    });
}

int mg_apply_h_device(mg_ctx *c, const void *dB, void *dHB) {
    return guarded(c, [&] {
        HIP_CHECK(hipEventRecord(c->ev0, c->stream));
        apply_h(c, dB, dHB);
        HIP_CHECK(hipEventRecord(c->ev1, c->stream));
        HIP_CHECK(hipEventSynchronize(c->ev1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
    });
}

int mg_apply_h_stats(mg_ctx *c, double *bytes_alg, double *flops_alg) {
    return guarded(c, [&] {
        // bytes: read B once, write HB once; flops: 8 per complex MAC over all
        // mode products of the diagonal sums and the chains.
        double fl = 0.0;
        for (int s = 0; s < c->ns; s++)
            for (int m = 0; m < c->nm; m++) {
                const int row = s * c->nm + m;
                if (c->hs_rowptr[row + 1] == c->hs_rowptr[row]) continue;
                fl += 8.0 * c->NS(m, s) * (double)c->nAConf[s];
            }
        for (auto &ch : c->chains) {
            std::vector<int64_t> sz(c->nm);
            for (int m = 0; m < c->nm; m++) sz[m] = c->NS(m, ch.rhs);
            for (auto &o : ch.ops) {
                double tot = 1.0;
                for (int m = 0; m < c->nm; m++) tot *= (double)sz[m];
                const double nout = (double)c->NS(o.mode, ch.lhs);
                fl += 8.0 * nout * tot;  // tot/nin fibers * nin*nout MACs
                sz[o.mode] = (int64_t)nout;
            }
        }
        const double by = 2.0 * (double)c->es * (double)c->nconf;
        if (bytes_alg) *bytes_alg = by;
        if (flops_alg) *flops_alg = fl;
        c->bytes_alg = by;
        c->flops_alg = fl;
    });
}

int mg_der_spf(mg_ctx *c, int64_t im, const double *psi_in, double *dpsi_out) {
    return guarded(c, [&] {
        if (im < 1 || im > c->nm) throw HipError{"mode index out of range"};
        const int desc = c->dvr_desc[im - 1];
        if (desc < 0) throw HipError{"mode is a Gaussian mode (use mg_der_spf_gwp)"};
        HIP_CHECK(hipMemcpyAsync(c->dpH.p, psi_in, sizeof(double2) * c->dimpsi, hipMemcpyHostToDevice, c->stream));
        launch_fill(c->dpH2.as<double2>(), c->dimpsi, make_double2(0, 0), c->stream);
        launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
        SpfParams P = spf_params(c, 0.0);
        launch_der_spf(c->dspfmodes.as<SpfModeDesc>(), desc, c->dpH.as<double2>(), c->dpH2.as<double2>(),
                       c->dspfscr.as<double2>(), P, dof_tables(c), c->dstatus.as<int>(), c->stream);
        HIP_CHECK(hipMemcpyAsync(dpsi_out, c->dpH2.p, sizeof(double2) * c->dimpsi, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int mg_autotune(mg_ctx *c, int64_t reps, int64_t *best_brick, double *best_ms) {
    return guarded(c, [&] {
        // candidates around the default that fit the LDS; score = H.A time x 8
        // + mean-field time x 2 (their weight in one CMF2 attempt)
        const int def = c->prec ? 2048 : 1024;
        // candidates: LDS brick sizes around the default that fit, each with
        // strided-group rows of 32 and 64 bytes when they group the modes
        // differently
        std::vector<std::pair<int, int>> cand;
        for (int b : {def / 4, def / 2, def, def * 2, def * 4}) {
            if (!brick_fits_lds(c, b)) continue;
            std::vector<std::pair<int, int>> seen;
            for (int rb : {32, 64}) {
                c->brick = b;
                c->run_bytes = rb;
                make_groups(c);
                bool dup = false;
                for (auto &q : cand)
                    if (q.first == b) {
                        c->brick = b;
                        c->run_bytes = q.second;
                        const auto g0 = c->groups;
                        c->run_bytes = rb;
                        make_groups(c);
                        dup = dup || (g0 == c->groups);
                    }
                if (!dup) cand.push_back({b, rb});
            }
        }
        const int R = reps > 0 ? (int)reps : 2;
        // input: the resident A (an FP32 copy in the mixed form)
        if (c->prec) launch_convert(c->dAf.p, 1, c->dA.p, 0, c->nconf, c->stream);
        const void *in = c->prec ? c->dAf.p : c->dA.p;
        void *out = c->prec ? c->dYf.p : c->dOut.p;
        double best = 1e300;
        int bestb = c->brick, bestr = c->run_bytes;
        for (auto [b, rb] : cand) {
            HIP_CHECK(hipStreamSynchronize(c->stream));
            c->plans.clear();
            c->graph_epoch++;          // an attempt graph holds the cleared plans' task tables
            c->brick = b;
            c->run_bytes = rb;
            make_groups(c);
            alloc_mf_tensors(c);
            calc_rho(c, c->dA.as<double2>());
            calc_h1mode(c, c->dpsi.as<double2>());
            apply_h(c, in, out);                        // plan build and warm-up
            calc_meanfield(c, c->dA.as<double2>(), true);
            double tha = 1e300, tmf = 1e300;
            for (int r = 0; r < R; r++) {
                HIP_CHECK(hipEventRecord(c->ev0, c->stream));
                apply_h(c, in, out);
                HIP_CHECK(hipEventRecord(c->ev1, c->stream));
                HIP_CHECK(hipEventSynchronize(c->ev1));
                float ms = 0.f;
                HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
                tha = std::min(tha, (double)ms);
                calc_rho(c, c->dA.as<double2>());
                HIP_CHECK(hipEventRecord(c->ev0, c->stream));
                calc_meanfield(c, c->dA.as<double2>(), true);
                HIP_CHECK(hipEventRecord(c->ev1, c->stream));
                HIP_CHECK(hipEventSynchronize(c->ev1));
                HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
                tmf = std::min(tmf, (double)ms);
            }
            const double score = 8.0 * tha + 2.0 * tmf;
            if (std::getenv("MG_VERBOSE"))
                fprintf(stderr, "autotune: brick %d rows %d B (%zu groups)  H.A %.3f ms  mean fields %.3f ms  score %.3f\n",
                        b, rb, c->groups.size(), tha, tmf, score);
            if (score < best) {
                best = score;
                bestb = b;
                bestr = rb;
            }
        }
        HIP_CHECK(hipStreamSynchronize(c->stream));
        c->plans.clear();
        c->graph_epoch++;
        c->brick = bestb;
        c->run_bytes = bestr;
        make_groups(c);
        alloc_mf_tensors(c);
        if (best_brick) *best_brick = bestb;
        if (best_ms) *best_ms = best;
    });
}

int mg_calc_derivs(mg_ctx *c, double *dA_out, double *dpsi_out) {
    return guarded(c, [&] {
        vmf_alloc(c);
        launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
        calc_derivs(c, c->dA.as<double2>(), c->dpsi.as<double2>(), c->dvA[0].p, c->dvP[0].as<double2>());
        spf_status_check(c);
        // dA = der_factor * H A, as complex FP64
        launch_fill(c->dOut.as<double2>(), c->nconf, make_double2(0, 0), c->stream);
        const void *k[1] = {c->dvA[0].p};
        const double2 w[1] = {make_double2(0.0, -1.0)};
        launch_rk_combine(c->dIn.as<double2>(), c->dOut.as<double2>(), k, w, 1, c->nconf, c->prec, c->stream);
        if (dA_out)
            HIP_CHECK(hipMemcpyAsync(dA_out, c->dIn.p, sizeof(double2) * c->nconf, hipMemcpyDeviceToHost, c->stream));
        if (dpsi_out)
            HIP_CHECK(hipMemcpyAsync(dpsi_out, c->dvP[0].p, sizeof(double2) * c->dimpsi, hipMemcpyDeviceToHost,
                                     c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int mg_propagate_vmf(mg_ctx *c, int64_t method, double t0, double tfinal, mg_output *outs, double *pops,
                     int64_t max_out, int64_t *n_out, mg_attempt *atts, int64_t max_att, int64_t *n_att) {
    return guarded(c, [&] {
        if (method != 1 && method != 2) throw HipError{"VMF method must be 1 (Cash-Karp) or 2 (Dormand-Prince)"};
        if (!n_out) throw HipError{"n_out is required"};
        const bool dp = (method == 2);
        vmf_alloc(c);
        const double wt = c->writetime, minstep = c->min_stepsize;
        const double maxtime = tfinal + wt * 0.1;
        double time = t0, stepsize = c->stepsize;
        *n_out = 0;
        if (n_att) *n_att = 0;
        std::vector<mg_attempt> log;
        if (dp) {   // the reference leaves these unfilled for its first step (propagator.f90:97-113)
            launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
            calc_derivs(c, c->dA.as<double2>(), c->dpsi.as<double2>(), c->dvA[0].p, c->dvP[0].as<double2>());
            spf_status_check(c);
        }
        while (time < maxtime) {
            if (*n_out < max_out) {
                observables(c, time, &outs[*n_out], pops ? pops + c->ns * *n_out : nullptr);
                check_finite(outs[*n_out], time);
                (*n_out)++;
            }
            const double tnext = time + wt;
            for (;;) {
                double done = 0.0, next = 0.0;
                if (!dp) {
                    launch_fill(reinterpret_cast<double2 *>(c->dstatus.p), 1, make_double2(0, 0), c->stream);
                    calc_derivs(c, c->dA.as<double2>(), c->dpsi.as<double2>(), c->dvA[0].p,
                                c->dvP[0].as<double2>());
                }
                log.clear();
                vmf_rkqs(c, dp, stepsize, time, &done, &next, atts ? &log : nullptr);
                if (atts && n_att)
                    for (auto &a : log)
                        if (*n_att < max_att) atts[(*n_att)++] = a;
                time = time + done;
                if (time >= tnext - minstep) break;
                const int64_t k = (int64_t)((tnext - time) / next);
                stepsize = (k == 0) ? tnext - time : (tnext - time) / (double)(k + 1);
                if (c->inorm_psi == 3) renormalise(c);
            }
            renormalise_interval(c);
        }
        c->stepsize = stepsize;
    });
}

int mg_cmf2_step(mg_ctx *c, double t, double htry, double *hdid, double *hnext) {
    return guarded(c, [&] { cmf2_step(c, t, htry, hdid, hnext, nullptr); });
}

int mg_observables(mg_ctx *c, double t, mg_output *out, double *pops) {
    return guarded(c, [&] { observables(c, t, out, pops); });
}

int mg_set_normalize(mg_ctx *c, int64_t mode) {
    return guarded(c, [&] {
        if (mode < 0 || mode > 3) throw HipError{"normalize must be 0, 1, 2 or 3"};
        c->inorm_psi = (int)mode;
    });
}

int mg_propagate(mg_ctx *c, double t0, double tfinal, mg_output *outs, double *pops, int64_t max_out, int64_t *n_out,
                 mg_attempt *atts, int64_t max_att, int64_t *n_att) {
    return guarded(c, [&] {
        const double wt = c->writetime, minstep = c->min_stepsize;
        const double maxtime = tfinal + wt * 0.1;
        double time = t0, stepsize = c->stepsize;
        *n_out = 0;
        if (n_att) *n_att = 0;
        std::vector<mg_attempt> log;
        while (time < maxtime) {
            if (*n_out < max_out) {
                observables(c, time, &outs[*n_out], pops ? pops + c->ns * *n_out : nullptr);
                check_finite(outs[*n_out], time);
                (*n_out)++;
            }
            const double tnext = time + wt;
            for (;;) {
                double done = 0.0, next = 0.0;
                log.clear();
                cmf2_step(c, time, stepsize, &done, &next, atts ? &log : nullptr);
                if (atts && n_att)
                    for (auto &a : log)
                        if (*n_att < max_att) atts[(*n_att)++] = a;
                time = time + done;
                if (time >= tnext - minstep) break;
                const int64_t k = (int64_t)((tnext - time) / next);
                stepsize = (k == 0) ? tnext - time : (tnext - time) / (double)(k + 1);
                c->stepsize = stepsize;
                if (c->inorm_psi == 3) renormalise(c);
            }
            renormalise_interval(c);
        }
        c->stepsize = stepsize;
    });
}

}  // extern "C"
