// Gaussian-SPF (G-MCTDH) kernels of the MI355X engine (gfx950, wave64).
//
// A Gaussian mode carries n Gaussians g_j = exp(A_jk x_k^2 + xi_jk x_k) per
// DOF k with fixed widths A < 0; the SPF parameters are the complex xi.  All
// matrices here are small (n <= 16 Gaussians per mode and set), so every
// kernel works on one mode per workgroup in FP64:
//   k_gwp_s00       overlaps S00 (Calc_S00, gmctdh_objects.f90:323-381)
//   k_gwp_h1        h1mode of the Gaussian modes (Calc_h1mode, hamilton.f90:486-525)
//   k_gwp_overlaps  SPF overlap matrices for norm_wf / WF_Ovl / DeltaPsi
//   k_gwp_dd_init   D0 = InverseSquareRootU(S00) (maths.f90:621-644, cmf2_base.f90:72-80)
//   k_gwp_dhd       D^H h D (Calc_DHD, hamilton.f90:535-574)
//   k_gwp_ab_mats   the per-mode matrices of A_to_B / B_to_A (cmf2_base.f90:264-345)
//   k_gwp_propagate propagate_SPF of the Gaussian modes with the C/Y equations
//                   of motion (Calc_Der_SPF + Calc_CY, cmf2_derivatives.f90:175-232,
//                   gmctdh_objects.f90:1198-1333) and rk_SPF_ck / rk_SPF_dp
//                   carrying D along (cmf2_integration.f90:107-412).
#include "mg_internal.h"

#include <math.h>

namespace mg {

namespace {

__device__ __forceinline__ double2 z0() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 cexpd(double2 z) {
    double sn, cs;
    sincos(z.y, &sn, &cs);
    const double e = exp(z.x);
    return make_double2(e * cs, e * sn);
}
__device__ __forceinline__ double2 cdivr(double2 a, double r) { return make_double2(a.x / r, a.y / r); }
__device__ __forceinline__ double2 csq(double2 a) { return cmul(a, a); }
__device__ __forceinline__ double2 cmi(double2 a) { return make_double2(a.y, -a.x); }   // -i a
__device__ __forceinline__ double cabs2(double2 a) { return hypot(a.x, a.y); }

// <g1|g2> of normalised nmd-dimensional Gaussians (GauOvlMultiD,
// gmctdh_objects.f90:1497-1519, equal-width branch :1509-1513).
__device__ double2 gau_ovl(const double *a1, const double *a2, const double2 *x1, const double2 *x2, int nmd) {
    bool equal = true;
    for (int k = 0; k < nmd; k++)
        if (!(fabs(a1[k] - a2[k]) < 2.220446049250313e-16)) equal = false;
    double2 c = z0();
    if (equal) {
        for (int k = 0; k < nmd; k++) {
            const double2 s = cadd(cconj(x1[k]), x2[k]);
            const double2 t = cadd(make_double2(x1[k].x * x1[k].x + x2[k].x * x2[k].x, 0.0), cscale(-0.5, csq(s)));
            c = cadd(c, cdivr(t, a1[k]));
        }
        return cexpd(cscale(0.25, c));
    }
    double pr = 1.0;
    for (int k = 0; k < nmd; k++) {
        const double A1 = a1[k], A2 = a2[k];
        const double2 s = cadd(cconj(x1[k]), x2[k]);
        double2 t = make_double2(x1[k].x * x1[k].x / A1 + x2[k].x * x2[k].x / A2, 0.0);
        t = csub(t, cdivr(csq(s), A1 + A2));
        c = cadd(c, t);
        pr *= 4.0 * A1 * A2 / ((A1 + A2) * (A1 + A2));
    }
    return cscale(sqrt(sqrt(pr)), cexpd(cscale(0.25, c)));
}

// binomial(n, 2k) (2k-1)!!
__device__ double moment_coeff(int k, int n) {
    double b = 1.0;
    for (int i = 0; i < 2 * k; i++) b = b * (double)(n - i) / (double)(i + 1);
    double df = 1.0;
    for (int i = 2 * k - 1; i > 1; i -= 2) df *= (double)i;
    return b * df;
}

__device__ double2 cpow_int(double2 z, int n) {
    double2 r = make_double2(1.0, 0.0);
    for (int i = 0; i < n; i++) r = cmul(r, z);
    return r;
}

// <g1|O|g2>/<g1|g2> (GauInt1D, hamilton.f90:745-849): 0 identity, -1 p,
// -2 d^2/dq^2, -3 qp + pq, 1..20 q^n; anything else sets *bad.
__device__ double2 gau_int(double A1, double A2, double2 xi1, double2 xi2, int op, int *bad) {
    const double2 x1c = cconj(xi1);
    const double s = A1 + A2;
    if (op == 0) return make_double2(1.0, 0.0);
    if (op == -1) return cmi(cdivr(csub(cscale(A1, xi2), cscale(A2, x1c)), s));
    if (op == -2) {
        const double2 t = cdivr(csub(cscale(A2, x1c), cscale(A1, xi2)), s);
        return cadd(csq(t), make_double2(2.0 * A1 * A2 / s, 0.0));
    }
    if (op == -3) {
        const double2 g =
            cadd(make_double2(A1 * A1 - A2 * A2, 0.0), cmul(csub(cscale(A2, x1c), cscale(A1, xi2)), cadd(x1c, xi2)));
        return cmi(cdivr(g, s * s));
    }
    if (op == 1) return cdivr(cadd(x1c, xi2), -2.0 * s);
    if (op == 2) {
        const double s2 = 2.0 * s;
        const double2 t = cdivr(cadd(x1c, xi2), s2);
        return csub(csq(t), make_double2(1.0 / s2, 0.0));
    }
    if (op >= 3 && op <= 20) {
        const double2 c = cadd(x1c, xi2);
        const double a = -2.0 * s;
        double2 g = z0();
        if (op % 2) {
            for (int k = 0; k <= op / 2; k++)
                g = cadd(g, cscale(moment_coeff(k, op) / pow(a, (double)(op - k)), cpow_int(c, op - 2 * k)));
        } else {
            const double2 c2 = cdivr(csq(c), a);
            double2 c1 = make_double2(1.0, 0.0);
            for (int k = op / 2; k >= 0; k--) {
                g = cadd(g, cscale(moment_coeff(k, op), c1));
                c1 = cmul(c1, c2);
            }
            g = cdivr(g, pow(a, (double)(op / 2)));
        }
        return g;
    }
    *bad = 1;
    return z0();
}

// <dg1/dxi1|O|g2>/<g1|g2> (Gau10Int1D, hamilton.f90:853-902).
__device__ double2 gau10_int(double A1, double A2, double2 xi1, double2 xi2, int op, int *bad) {
    if (op == -1)
        return csub(cmi(cmul(xi2, gau_int(A1, A2, xi1, xi2, 1, bad))),
                    cmul(make_double2(0.0, 2.0 * A2), gau_int(A1, A2, xi1, xi2, 2, bad)));
    if (op == -2)
        return cadd(cadd(cmul(cadd(make_double2(A2 + A2, 0.0), csq(xi2)), gau_int(A1, A2, xi1, xi2, 1, bad)),
                         cmul(cscale(4.0 * A2, xi2), gau_int(A1, A2, xi1, xi2, 2, bad))),
                    cscale(4.0 * A2 * A2, gau_int(A1, A2, xi1, xi2, 3, bad)));
    if (op == -3)
        return cmi(cadd(cadd(gau_int(A1, A2, xi1, xi2, 1, bad), cmul(cscale(2.0, xi2), gau_int(A1, A2, xi1, xi2, 2, bad))),
                        cscale(4.0 * A2, gau_int(A1, A2, xi1, xi2, 3, bad))));
    if (op >= 0 && op <= 19) return gau_int(A1, A2, xi1, xi2, op + 1, bad);
    *bad = 1;
    return z0();
}

__device__ __forceinline__ const double *gwpa_ptr(const GwpParams &P, int m, int j) {
    return P.gwpa + (size_t)P.maxdof * (j + (size_t)P.nspfmax * m);
}

// Overlap block of sets (s1, s2) of one mode from xi blocks x1 (set s1) and
// x2 (set s2): intra-set (same xi, sym) -> 1 on the diagonal, computed upper
// triangle, mirrored lower; otherwise every element.
__device__ double2 ovl_elem(const GwpParams &P, int m, int nmd, const double2 *x1, const double2 *x2, int j1, int j2,
                            bool sym) {
    if (sym) {
        if (j1 == j2) return make_double2(1.0, 0.0);
        if (j1 > j2)
            return cconj(gau_ovl(gwpa_ptr(P, m, j2), gwpa_ptr(P, m, j1), x1 + nmd * j2, x1 + nmd * j1, nmd));
        return gau_ovl(gwpa_ptr(P, m, j1), gwpa_ptr(P, m, j2), x1 + nmd * j1, x1 + nmd * j2, nmd);
    }
    return gau_ovl(gwpa_ptr(P, m, j1), gwpa_ptr(P, m, j2), x1 + nmd * j1, x2 + nmd * j2, nmd);
}

// S (nsets x nsets blocks of nmax^2) of a mode from local xi blocks at loff.
// Blocks (s1 > s2) are the conjugate transposes of (s2, s1).
__device__ void calc_s_local(const GwpDesc &d, const GwpParams &P, const double2 *xi, const int *loff, double2 *S) {
    const int nmax = P.nspfmax, nn = nmax * nmax, nb = d.nsets * d.nsets;
    for (int e = threadIdx.x; e < nb * nn; e += blockDim.x) {
        const int b = e / nn, r = e % nn;
        const int s1 = b % d.nsets, s2 = b / d.nsets, j1 = r % nmax, j2 = r / nmax;
        double2 v = z0();
        if (j1 < d.nspf[s1] && j2 < d.nspf[s2]) {
            if (s1 == s2) v = ovl_elem(P, d.mode, d.nmd, xi + loff[s1], xi + loff[s1], j1, j2, true);
            else if (s1 < s2) v = ovl_elem(P, d.mode, d.nmd, xi + loff[s1], xi + loff[s2], j1, j2, false);
            else v = cconj(ovl_elem(P, d.mode, d.nmd, xi + loff[s2], xi + loff[s1], j2, j1, false));
        }
        S[e] = v;
    }
}

// h1mode element of operator q (sets s1, s2) from the overlaps S.
__device__ double2 h1_elem(const GwpDesc &d, const GwpParams &P, const double2 *xi, const int *loff, const double2 *S,
                           int q, int j1, int j2, int *bad) {
    const int nmax = P.nspfmax, nn = nmax * nmax;
    const int s1 = d.op_lhs[q], s2 = d.op_rhs[q];
    const double2 *Sb = S + (size_t)nn * (s1 + d.nsets * s2);
    if (d.op_ident[q]) return Sb[j1 + nmax * j2];
    const bool same = (s1 == s2);
    if (same && j1 > j2) return cconj(h1_elem(d, P, xi, loff, S, q, j2, j1, bad));
    double2 v = Sb[j1 + nmax * j2];
    const double2 *x1 = xi + loff[s1] + d.nmd * j1, *x2 = xi + loff[s2] + d.nmd * j2;
    double2 pr = make_double2(1.0, 0.0);
    for (int k = 0; k < d.nmd; k++)
        pr = cmul(pr, gau_int(gwpa_ptr(P, d.mode, j1)[k], gwpa_ptr(P, d.mode, j2)[k], x1[k], x2[k], d.op_codes[q][k], bad));
    return cmul(v, pr);
}

// Regularised inverse w / (w^2 + eps^2) of the Hermitian n x n matrix in
// `in` (ld nmax) into `out` (StdRegInv 'tikh', maths.f90:8-83).
__device__ void tikh_inverse(int n, const double2 *in, int nmax, double eps, double2 *out, double2 *A, double2 *U,
                             double *w) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double2 v = (i <= j) ? in[i + nmax * j] : cconj(in[j + nmax * i]);
            if (i == j) v.y = 0.0;
            A[i + kSmall * j] = v;
        }
    dev_herm_eig(n, A, w, U);
    for (int k = 0; k < n; k++) w[k] = w[k] / (w[k] * w[k] + eps * eps);
    for (int j2 = 0; j2 < n; j2++)
        for (int j1 = 0; j1 < n; j1++) {
            double2 acc = z0();
            for (int k = 0; k < n; k++) acc = cadd(acc, cscale(w[k], cmul(U[j1 + kSmall * k], cconj(U[j2 + kSmall * k]))));
            out[j1 + nmax * j2] = acc;
        }
    for (int j = 0; j < n; j++) out[j + nmax * j].y = 0.0;
    for (int j2 = 0; j2 < n; j2++)
        for (int j1 = j2 + 1; j1 < n; j1++) out[j1 + nmax * j2] = cconj(out[j2 + nmax * j1]);
}

// Per-mode scratch of the propagation (complex elements).
struct GwpScratch {
    double2 *y0, *k[7], *tmp, *perr, *D0, *kD[7], *tmpD, *S, *Sm, *h, *work;
};

constexpr int kSetWork = 4 * kSmall * kSmall + 4 * kSmall * 16 + 3 * kSmall + kMaxDofPerMode * 256;

__device__ GwpScratch gwp_carve(double2 *base, const GwpDesc &d, int nspfmax) {
    const size_t nn = (size_t)nspfmax * nspfmax, nd = nn * d.nsets;
    GwpScratch s;
    double2 *p = base + d.scr_off;
    s.y0 = p; p += d.blk;
    for (int i = 0; i < 7; i++) { s.k[i] = p; p += d.blk; }
    s.tmp = p; p += d.blk;
    s.perr = p; p += d.blk;
    s.D0 = p; p += nd;
    for (int i = 0; i < 7; i++) { s.kD[i] = p; p += nd; }
    s.tmpD = p; p += nd;
    s.S = p; p += nn * d.nsets * d.nsets;
    s.Sm = p; p += nd;
    s.h = p; p += nn * d.nop;
    s.work = p;
    return s;
}

// Calc_Der_SPF, Gaussian branch, with Calc_CY (no DD0 form): dxi = der_factor
// * StdRegSol_tikh(C, Y, eps_C), dD = -tau D.  S00, S00^-1 and h1mode of the
// mode are evaluated at xi; rho and the mean fields are the caller's.
__device__ void dev_der_gwp(const GwpDesc &d, const GwpParams &P, const double2 *xi, const double2 *D,
                            double2 *dxi, double2 *dD, GwpScratch &W, int *fail) {
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nmax = P.nspfmax, nn = nmax * nmax, nmd = d.nmd, m = d.mode;
    int loff[kMaxSets];
    {
        int o = 0;
        for (int s = 0; s < d.nsets; s++) { loff[s] = o; o += nmd * d.nspf[s]; }
    }
    calc_s_local(d, P, xi, loff, W.S);
    __syncthreads();
    int bad = 0;
    for (int e = tid; e < d.nop * nn; e += nth) {
        const int q = e / nn, r = e % nn, j1 = r % nmax, j2 = r / nmax;
        const int n1 = d.nspf[d.op_lhs[q]], n2 = d.nspf[d.op_rhs[q]];
        W.h[e] = (j1 < n1 && j2 < n2) ? h1_elem(d, P, xi, loff, W.S, q, j1, j2, &bad) : z0();
    }
    if (tid < d.nsets) {
        const int s = tid, n = d.nspf[s];
        double2 *wk = W.work + (size_t)kSetWork * s;
        tikh_inverse(n, W.S + (size_t)nn * (s + d.nsets * s), nmax, P.eps_ovl, W.Sm + (size_t)nn * s, wk,
                     wk + kSmall * kSmall, reinterpret_cast<double *>(wk + 2 * kSmall * kSmall));
    }
    __syncthreads();
    if (tid < d.nsets) {
        const int s = tid, nS = d.nspf[s], nG = nS * nmd;
        double2 *wk = W.work + (size_t)kSetWork * s;
        double2 *Sab = wk;                                  // kSmall^2 (C in place)
        double2 *U = wk + kSmall * kSmall;                  // kSmall^2
        double *w = reinterpret_cast<double *>(wk + 2 * kSmall * kSmall);
        double2 *E = wk + 3 * kSmall * kSmall;              // kSmall^2
        double2 *Sa0 = wk + 4 * kSmall * kSmall;            // kSmall x 16
        double2 *SaSm = Sa0 + kSmall * 16;
        double2 *Ha0 = SaSm + kSmall * 16;
        double2 *tau = Ha0 + kSmall * 16;                   // kSmall x 16 (uses nmax^2)
        double2 *Y = tau + kSmall * 16;
        double2 *gx = Y + 3 * kSmall;                       // nmd x nmax x nmax
        const double2 *S = W.S + (size_t)nn * (s + d.nsets * s);
        const double2 *Sm = W.Sm + (size_t)nn * s;
        const double2 *x = xi + loff[s];
#define GX(k, a, b) gx[(k) + nmd * ((a) + nmax * (b))]
        for (int j2 = 0; j2 < nS; j2++)
            for (int j1 = 0; j1 <= j2; j1++)
                for (int k = 0; k < nmd; k++)
                    GX(k, j1, j2) = gau_int(gwpa_ptr(P, m, j1)[k], gwpa_ptr(P, m, j2)[k], x[nmd * j1 + k],
                                            x[nmd * j2 + k], 1, &bad);
        for (int j2 = 0; j2 < nS; j2++)
            for (int j1 = 0; j1 < nS; j1++)
                for (int k = 0; k < nmd; k++)
                    Sa0[(j1 * nmd + k) + kSmall * j2] = (j1 <= j2) ? cmul(S[j1 + nmax * j2], GX(k, j1, j2))
                                                                   : cconj(cmul(S[j2 + nmax * j1], GX(k, j2, j1)));
        for (int c = 0; c < nS; c++)
            for (int r = 0; r < nG; r++) {
                double2 acc = z0();
                for (int q = 0; q < nS; q++) acc = cadd(acc, cmul(Sa0[r + kSmall * q], Sm[q + nmax * c]));
                SaSm[r + kSmall * c] = acc;
            }
        for (int i = 0; i < kSmall * kSmall; i++) Sab[i] = z0();
        for (int j2 = 0; j2 < nS; j2++)
            for (int k2 = 0; k2 < nmd; k2++)
                for (int j1 = 0; j1 <= j2; j1++)
                    for (int k1 = 0; k1 < nmd; k1++) {
                        double2 v;
                        if (k1 == k2)
                            v = gau_int(gwpa_ptr(P, m, j1)[k1], gwpa_ptr(P, m, j2)[k2], x[nmd * j1 + k1], x[nmd * j2 + k2],
                                        2, &bad);
                        else
                            v = cmul(GX(k1, j1, j2), GX(k2, j1, j2));
                        Sab[(j1 * nmd + k1) + kSmall * (j2 * nmd + k2)] = cmul(S[j1 + nmax * j2], v);
                    }
        for (int c = 0; c < nG; c++)
            for (int r = 0; r <= c; r++) {
                double2 acc = z0();
                for (int q = 0; q < nS; q++)
                    acc = cadd(acc, cadd(cmul(SaSm[r + kSmall * q], cconj(Sa0[c + kSmall * q])),
                                         cmul(Sa0[r + kSmall * q], cconj(SaSm[c + kSmall * q]))));
                double2 v = csub(Sab[r + kSmall * c], cscale(0.5, acc));
                if (r == c) v.y = 0.0;
                Sab[r + kSmall * c] = v;
            }
        // C = rho o Sab on the upper triangle, as a full Hermitian matrix in E
        const double2 *rh = P.rho + (size_t)nn * d.rho_index[s];
        for (int j2 = 0; j2 < nS; j2++)
            for (int k2 = 0; k2 < nmd; k2++)
                for (int j1 = 0; j1 <= j2; j1++)
                    for (int k1 = 0; k1 < nmd; k1++) {
                        const int r = j1 * nmd + k1, c = j2 * nmd + k2;
                        if (r > c) continue;
                        double2 v = cmul(rh[j1 + nmax * j2], Sab[r + kSmall * c]);
                        if (r == c) v.y = 0.0;
                        E[r + kSmall * c] = v;
                        E[c + kSmall * r] = cconj(v);
                    }
        // Y = sum_op MF (H_a0 - Sa0 S^-1 h)
        for (int i = 0; i < nG; i++) Y[i] = z0();
        for (int q = 0; q < d.nop; q++) {
            if (d.op_ident[q] || d.op_lhs[q] != s) continue;
            const int s2 = d.op_rhs[q], n2 = d.nspf[s2];
            const double2 *S12 = W.S + (size_t)nn * (s + d.nsets * s2);
            const double2 *x2b = xi + loff[s2];
            for (int j2 = 0; j2 < n2; j2++)
                for (int j1 = 0; j1 < nS; j1++) {
                    double2 gI[kMaxDofPerMode];
                    for (int k = 0; k < nmd; k++)
                        gI[k] = gau_int(gwpa_ptr(P, m, j1)[k], gwpa_ptr(P, m, j2)[k], x[nmd * j1 + k], x2b[nmd * j2 + k],
                                        d.op_codes[q][k], &bad);
                    for (int k = 0; k < nmd; k++) {
                        double2 pr = S12[j1 + nmax * j2];
                        for (int kk = 0; kk < nmd; kk++)
                            if (kk != k) pr = cmul(pr, gI[kk]);
                        Ha0[(j1 * nmd + k) + kSmall * j2] =
                            cmul(pr, gau10_int(gwpa_ptr(P, m, j1)[k], gwpa_ptr(P, m, j2)[k], x[nmd * j1 + k],
                                               x2b[nmd * j2 + k], d.op_codes[q][k], &bad));
                    }
                }
            const double2 *h = W.h + (size_t)nn * q;
            for (int j2 = 0; j2 < n2; j2++)
                for (int r = 0; r < nG; r++) {
                    double2 acc = z0();
                    for (int qq = 0; qq < nS; qq++) acc = cadd(acc, cmul(SaSm[r + kSmall * qq], h[qq + nmax * j2]));
                    Ha0[r + kSmall * j2] = csub(Ha0[r + kSmall * j2], acc);
                }
            const double2 *mf = P.mf + (size_t)nn * d.op_index[q];
            for (int j2 = 0; j2 < n2; j2++)
                for (int j1 = 0; j1 < nS; j1++)
                    for (int k = 0; k < nmd; k++)
                        Y[j1 * nmd + k] = cadd(Y[j1 * nmd + k], cmul(mf[j1 + nmax * j2], Ha0[(j1 * nmd + k) + kSmall * j2]));
        }
        // dxi = der_factor * U diag(w / (w^2 + eps_C^2)) U^H Y (StdRegSol 'tikh')
        dev_herm_eig(nG, E, w, U);
        double2 uy[kSmall];
        for (int k = 0; k < nG; k++) {
            double2 acc = z0();
            for (int i = 0; i < nG; i++) acc = cadd(acc, cmul(cconj(U[i + kSmall * k]), Y[i]));
            uy[k] = cscale(w[k] / (w[k] * w[k] + P.eps_C * P.eps_C), acc);
        }
        double2 *dx = dxi + loff[s];
        for (int i = 0; i < nG; i++) {
            double2 acc = z0();
            for (int k = 0; k < nG; k++) acc = cadd(acc, cmul(U[i + kSmall * k], uy[k]));
            dx[i] = cmul(P.der_factor, acc);
        }
        // tau, then dD = -tau D
        for (int j2 = 0; j2 < nS; j2++)
            for (int j1 = 0; j1 < nS; j1++) {
                double2 acc = z0();
                for (int k = 0; k < nmd; k++) acc = cadd(acc, cmul(cconj(SaSm[(j2 * nmd + k) + kSmall * j1]), dx[j2 * nmd + k]));
                tau[j1 + nmax * j2] = acc;
            }
        for (int j = 0; j < nS; j++) {
            double dd = 0.0;
            for (int k = 0; k < nmd; k++) dd += x[nmd * j + k].x * dx[nmd * j + k].x / gwpa_ptr(P, m, j)[k];
            tau[j + nmax * j].x += dd * 0.5;
        }
        const double2 *D0 = D + (size_t)nn * s;
        double2 *dDs = dD + (size_t)nn * s;
        for (int i = 0; i < nn; i++) dDs[i] = z0();
        for (int b = 0; b < nS; b++)
            for (int a = 0; a < nS; a++) {
                double2 acc = z0();
                for (int q = 0; q < nS; q++) acc = cadd(acc, cmul(tau[a + nmax * q], D0[q + nmax * b]));
                dDs[a + nmax * b] = make_double2(-acc.x, -acc.y);
            }
#undef GX
    }
    if (bad) atomicOr(fail, 2);
    __syncthreads();
}

// Calc_Err_SPF, Gaussian branch (cmf2_derivatives.f90:293-334).
__device__ double dev_err_gwp(const GwpDesc &d, const GwpParams &P, const double2 *p1, const double2 *p2, double *part) {
    const int tid = threadIdx.x, nmax = P.nspfmax, nmd = d.nmd, m = d.mode;
    int loff[kMaxSets];
    {
        int o = 0;
        for (int s = 0; s < d.nsets; s++) { loff[s] = o; o += nmd * d.nspf[s]; }
    }
    if (tid < d.nsets) {
        const int s = tid, nS = d.nspf[s];
        const double2 *a = p1 + loff[s], *b = p2 + loff[s];
        const double2 *r = P.rho + (size_t)nmax * nmax * d.rho_index[s];
        double err = 0.0;
        for (int j2 = 0; j2 < nS; j2++) {
            for (int j1 = 0; j1 < j2; j1++) {
                const double *a1 = gwpa_ptr(P, m, j1), *a2 = gwpa_ptr(P, m, j2);
                double2 o = cadd(gau_ovl(a1, a2, a + nmd * j1, a + nmd * j2, nmd), gau_ovl(a1, a2, b + nmd * j1, b + nmd * j2, nmd));
                o = csub(o, gau_ovl(a1, a2, a + nmd * j1, b + nmd * j2, nmd));
                o = csub(o, gau_ovl(a1, a2, b + nmd * j1, a + nmd * j2, nmd));
                err += 2.0 * cmul(r[j1 + nmax * j2], cconj(o)).x;
            }
            const double2 g = gau_ovl(gwpa_ptr(P, m, j2), gwpa_ptr(P, m, j2), a + nmd * j2, b + nmd * j2, nmd);
            err += cmul(r[j2 + nmax * j2], make_double2(2.0 - 2.0 * g.x, 0.0)).x;
        }
        part[s] = err;
    }
    __syncthreads();
    double e = 0.0;
    for (int s = 0; s < d.nsets; s++) e += part[s];
    __syncthreads();
    return sqrt(fabs(e));
}

__constant__ double g_ckB[6][5] = {
    {0, 0, 0, 0, 0}, {0.2, 0, 0, 0, 0}, {0.075, 0.225, 0, 0, 0}, {0.3, -0.9, 1.2, 0, 0},
    {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
__constant__ double g_ckC[6] = {37.0 / 378.0, 0, 250.0 / 621.0, 125.0 / 594.0, 0, 512.0 / 1771.0};
__constant__ double g_ckD[6] = {2825.0 / 27648.0, 0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 0.25};
__constant__ double g_dpB[6][5] = {
    {0, 0, 0, 0, 0}, {0.2, 0, 0, 0, 0}, {0.075, 0.225, 0, 0, 0}, {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0}};
__constant__ double g_dpC[6] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0};
__constant__ double g_dpD[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0,
                                187.0 / 2100.0, 0.025};

// out = y0 + h sum_q w[q] k[q] (st == 1: y0 + w[0] h k[0]) over n elements.
__device__ void stage(double2 *out, const double2 *y0, double2 *const *k, const double *w, int nq, double h, int n,
                      bool first) {
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        if (first) {
            out[e] = cadd(y0[e], cscale(w[0] * h, k[0][e]));
            continue;
        }
        double2 acc = z0();
        for (int q = 0; q < nq; q++)
            if (w[q] != 0.0) acc = cadd(acc, cscale(w[q], k[q][e]));
        out[e] = cadd(y0[e], cscale(h, acc));
    }
}

__device__ void load_mode(const GwpDesc &d, int nmax, int nm, const double2 *psi, const double2 *dd, double2 *y,
                          double2 *D) {
    const int nn = nmax * nmax;
    int o = 0;
    for (int s = 0; s < d.nsets; s++) {
        const int n = d.nmd * d.nspf[s];
        for (int e = threadIdx.x; e < n; e += blockDim.x) y[o + e] = psi[d.off[s] + e];
        o += n;
        if (D)
            for (int e = threadIdx.x; e < nn; e += blockDim.x)
                D[(size_t)nn * s + e] = dd[(size_t)nn * (d.mode + (size_t)nm * s) + e];
    }
}

__device__ void store_mode(const GwpDesc &d, int nmax, int nm, const double2 *y, const double2 *D, double2 *psi,
                           double2 *dd) {
    const int nn = nmax * nmax;
    int o = 0;
    for (int s = 0; s < d.nsets; s++) {
        const int n = d.nmd * d.nspf[s];
        for (int e = threadIdx.x; e < n; e += blockDim.x) psi[d.off[s] + e] = y[o + e];
        o += n;
        if (dd)
            for (int e = threadIdx.x; e < nn; e += blockDim.x)
                dd[(size_t)nn * (d.mode + (size_t)nm * s) + e] = D[(size_t)nn * s + e];
    }
}

// ---------------------------------------------------------------------------
// Kernels.

__global__ void __launch_bounds__(64) k_gwp_s00(const GwpDesc *__restrict__ descs, const double2 *psi, GwpParams P,
                                                double2 *s00, double2 *sd) {
    const GwpDesc &d = descs[blockIdx.x];
    const int nmax = P.nspfmax, nn = nmax * nmax, m = d.mode;
    // set pairs (intra- and, for multi-set modes, inter-state)
    for (int e = threadIdx.x; e < d.nsets * d.nsets * nn; e += blockDim.x) {
        const int b = e / nn, r = e % nn;
        const int s1 = b % d.nsets, s2 = b / d.nsets, j1 = r % nmax, j2 = r / nmax;
        double2 v = z0();
        if (j1 < d.nspf[s1] && j2 < d.nspf[s2]) {
            const double2 *x1 = psi + d.off[s1], *x2 = psi + d.off[s2];
            if (s1 == s2) v = ovl_elem(P, m, d.nmd, x1, x1, j1, j2, true);
            else if (s1 < s2) v = ovl_elem(P, m, d.nmd, x1, x2, j1, j2, false);
            else v = cconj(ovl_elem(P, m, d.nmd, x2, x1, j2, j1, false));
        }
        s00[(size_t)nn * (m + (size_t)P.nm * (s1 + (size_t)P.ns * s2)) + r] = v;
    }
    // intra-state overlap per electronic state (the set's for single-set modes)
    for (int e = threadIdx.x; e < P.ns * nn; e += blockDim.x) {
        const int st = e / nn, r = e % nn, j1 = r % nmax, j2 = r / nmax;
        const int set = d.nsets == 1 ? 0 : st;
        double2 v = z0();
        if (j1 < d.nspf[set] && j2 < d.nspf[set])
            v = ovl_elem(P, m, d.nmd, psi + d.off[set], psi + d.off[set], j1, j2, true);
        sd[(size_t)nn * (m + (size_t)P.nm * st) + r] = v;
    }
}

__global__ void __launch_bounds__(64) k_gwp_h1(const GwpDesc *__restrict__ descs, const double2 *psi, GwpParams P,
                                               const double2 *s00, double2 *h1) {
    const GwpDesc &d = descs[blockIdx.x];
    const int nmax = P.nspfmax, nn = nmax * nmax, m = d.mode;
    int bad = 0;
    for (int e = threadIdx.x; e < d.nop * nn; e += blockDim.x) {
        const int q = e / nn, r = e % nn, j1 = r % nmax, j2 = r / nmax;
        const int s1 = d.op_lhs[q], s2 = d.op_rhs[q];
        double2 v = z0();
        if (j1 < d.nspf[s1] && j2 < d.nspf[s2]) {
            const double2 *S = s00 + (size_t)nn * (m + (size_t)P.nm * (s1 + (size_t)P.ns * s2));
            if (d.op_ident[q]) {
                v = S[j1 + nmax * j2];
            } else {
                const bool flip = (s1 == s2) && j1 > j2;
                const int a = flip ? j2 : j1, b = flip ? j1 : j2;
                const double2 *x1 = psi + d.off[s1] + d.nmd * a, *x2 = psi + d.off[s2] + d.nmd * b;
                double2 pr = make_double2(1.0, 0.0);
                for (int k = 0; k < d.nmd; k++)
                    pr = cmul(pr, gau_int(gwpa_ptr(P, m, a)[k], gwpa_ptr(P, m, b)[k], x1[k], x2[k], d.op_codes[q][k], &bad));
                v = cmul(S[a + nmax * b], pr);
                if (flip) v = cconj(v);
            }
        }
        h1[(size_t)nn * d.op_index[q] + r] = v;
    }
}

__global__ void __launch_bounds__(64) k_gwp_overlaps(const GwpDesc *__restrict__ descs, const double2 *psi1,
                                                     const double2 *psi2, int sym, GwpParams P, double2 *out) {
    const GwpDesc &d = descs[blockIdx.x];
    const int nmax = P.nspfmax, nn = nmax * nmax, m = d.mode;
    for (int e = threadIdx.x; e < P.ns * nn; e += blockDim.x) {
        const int st = e / nn, r = e % nn, j1 = r % nmax, j2 = r / nmax;
        const int set = d.nsets == 1 ? 0 : st;
        double2 v = z0();
        if (j1 < d.nspf[set] && j2 < d.nspf[set])
            v = ovl_elem(P, m, d.nmd, psi1 + d.off[set], psi2 + d.off[set], j1, j2, sym != 0);
        out[(size_t)nn * (m + (size_t)P.nm * st) + r] = v;
    }
}

// D = U^-1 with S = U^H U, one thread per (mode, set).
__global__ void k_gwp_dd_init(const GwpDesc *__restrict__ descs, const double2 *sd, GwpParams P, double2 *dd,
                              int *status) {
    const GwpDesc &d = descs[blockIdx.x];
    const int s = threadIdx.x;
    if (s >= d.nsets) return;
    const int nmax = P.nspfmax, nn = nmax * nmax, n = d.nspf[s];
    const double2 *S = sd + (size_t)nn * (d.mode + (size_t)P.nm * s);
    double2 *D = dd + (size_t)nn * (d.mode + (size_t)P.nm * s);
    double2 U[kSmall * kSmall];
    for (int j = 0; j < n; j++) {
        double dj = S[j + nmax * j].x;
        for (int k = 0; k < j; k++) dj -= cmul(cconj(U[k + kSmall * j]), U[k + kSmall * j]).x;
        if (!(dj > 0.0)) {
            atomicMax(status, 3);
            return;
        }
        const double ujj = sqrt(dj);
        U[j + kSmall * j] = make_double2(ujj, 0.0);
        for (int i = j + 1; i < n; i++) {
            double2 v = S[j + nmax * i];
            for (int k = 0; k < j; k++) v = csub(v, cmul(cconj(U[k + kSmall * j]), U[k + kSmall * i]));
            U[j + kSmall * i] = cdivr(v, ujj);
        }
    }
    for (int i = 0; i < nn; i++) D[i] = z0();
    for (int c = 0; c < n; c++) {
        D[c + nmax * c] = make_double2(1.0 / U[c + kSmall * c].x, 0.0);
        for (int r = c - 1; r >= 0; r--) {
            double2 acc = z0();
            for (int k = r + 1; k <= c; k++) acc = cadd(acc, cmul(U[r + kSmall * k], D[k + nmax * c]));
            D[r + nmax * c] = cdivr(make_double2(-acc.x, -acc.y), U[r + kSmall * r].x);
        }
    }
}

// D^H h D per operator, one thread per operator of the mode.
__global__ void k_gwp_dhd(const GwpDesc *__restrict__ descs, const double2 *h1, const double2 *dd, GwpParams P,
                          double2 *dhd) {
    const GwpDesc &d = descs[blockIdx.x];
    const int q = threadIdx.x;
    if (q >= d.nop) return;
    const int nmax = P.nspfmax, nn = nmax * nmax;
    const int s1 = d.op_lhs[q], s2 = d.op_rhs[q], n1 = d.nspf[s1], n2 = d.nspf[s2];
    double2 *R = dhd + (size_t)nn * d.op_index[q];
    for (int i = 0; i < nn; i++) R[i] = z0();
    if (d.op_ident[q]) {
        for (int j = 0; j < (n1 < n2 ? n1 : n2); j++) R[j + nmax * j] = make_double2(1.0, 0.0);
        return;
    }
    const double2 *h = h1 + (size_t)nn * d.op_index[q];
    const double2 *D1 = dd + (size_t)nn * (d.mode + (size_t)P.nm * s1);
    const double2 *D2 = dd + (size_t)nn * (d.mode + (size_t)P.nm * s2);
    double2 aux[kSmall * kSmall];
    if (s1 == s2) {
        for (int b = 0; b < n2; b++)
            for (int a = 0; a < n1; a++) {
                double2 acc = z0();
                for (int k = 0; k < n1; k++) acc = cadd(acc, cmul(h[a + nmax * k], D2[k + nmax * b]));
                aux[a + kSmall * b] = acc;
            }
        for (int b = 0; b < n1; b++)
            for (int a = 0; a <= b; a++) {
                double2 x = z0(), y = z0();
                for (int k = 0; k < n1; k++) {
                    x = cadd(x, cmul(cconj(D1[k + nmax * a]), aux[k + kSmall * b]));
                    y = cadd(y, cmul(cconj(aux[k + kSmall * a]), D1[k + nmax * b]));
                }
                double2 v = cadd(cscale(0.5, x), cscale(0.5, y));
                if (a == b) v.y = 0.0;
                R[a + nmax * b] = v;
            }
    } else {
        for (int b = 0; b < n2; b++)
            for (int a = 0; a < n1; a++) {
                double2 acc = z0();
                for (int k = 0; k < n1; k++) acc = cadd(acc, cmul(cconj(D1[k + nmax * a]), h[k + nmax * b]));
                aux[a + kSmall * b] = acc;
            }
        for (int b = 0; b < n2; b++)
            for (int a = 0; a < n1; a++) {
                double2 acc = z0();
                for (int k = 0; k < n2; k++) acc = cadd(acc, cmul(aux[a + kSmall * k], D2[k + nmax * b]));
                R[a + nmax * b] = acc;
            }
    }
    const double eps = 2.220446049250313e-16;
    for (int b = 0; b < n2; b++)
        for (int a = 0; a < (s1 == s2 ? b + 1 : n1); a++)
            if (!(cabs2(R[a + nmax * b]) > eps)) R[a + nmax * b] = z0();
    if (s1 == s2)
        for (int b = 0; b < n1; b++)
            for (int a = b + 1; a < n1; a++) R[a + nmax * b] = cconj(R[b + nmax * a]);
}

// A_to_B matrices (S D)^H and B_to_A matrices D per (mode, state): mats[0..]
// and mats[nn * nm * ns ..].
__global__ void k_gwp_ab_mats(const GwpDesc *__restrict__ descs, const double2 *sd, const double2 *dd, GwpParams P,
                              double2 *mats) {
    const GwpDesc &d = descs[blockIdx.x];
    const int nmax = P.nspfmax, nn = nmax * nmax;
    const size_t half = (size_t)nn * P.nm * P.ns;
    for (int e = threadIdx.x; e < P.ns * nn; e += blockDim.x) {
        const int st = e / nn, r = e % nn, a = r % nmax, b = r / nmax;
        const int set = d.nsets == 1 ? 0 : st, n = d.nspf[set];
        const double2 *S = sd + (size_t)nn * (d.mode + (size_t)P.nm * st);
        const double2 *D = dd + (size_t)nn * (d.mode + (size_t)P.nm * set);
        double2 v = z0(), w = z0();
        if (a < n && b < n) {
            double2 acc = z0();  // (S D)[b, a]
            for (int k = 0; k < n; k++) acc = cadd(acc, cmul(S[b + nmax * k], D[k + nmax * a]));
            v = cconj(acc);
            w = D[a + nmax * b];
        }
        mats[(size_t)nn * (d.mode + (size_t)P.nm * st) + r] = v;
        mats[half + (size_t)nn * (d.mode + (size_t)P.nm * st) + r] = w;
    }
}

// propagate_SPF of one Gaussian mode per workgroup over P.dt; psi and dd are
// updated in place.  status: 1 non-finite error, 2 too many attempts,
// 4 unsupported operator.
__global__ void __launch_bounds__(64) k_gwp_propagate(const GwpDesc *__restrict__ descs, double2 *psi, double2 *dd,
                                                      double2 *scratch, GwpParams P, int *status) {
    const GwpDesc &d = descs[blockIdx.x];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nmax = P.nspfmax, nn = nmax * nmax, nD = nn * d.nsets;
    GwpScratch S = gwp_carve(scratch, d, nmax);
    __shared__ int fail;
    __shared__ double part[kMaxSets];
    if (tid == 0) fail = 0;
    load_mode(d, nmax, P.nm, psi, dd, S.y0, S.D0);
    __syncthreads();
    const double safety = 0.9, pgrow = -0.2, pshrnk = -0.25, minscal = 0.1, maxscal = 5.0;
    const double errcon = pow(5.0 / safety, 1.0 / pgrow);
    const bool dp = P.dormand_prince != 0;
    const double DT = P.dt_src ? P.dt_src[0] * P.dt_scale : P.dt;
    double dt_done = 0.0, stepsize = DT;
    int attempts = 0;
    for (;;) {
        double h = stepsize, error = 0.0;
        for (;;) {
            dev_der_gwp(d, P, S.y0, S.D0, S.k[0], S.kD[0], S, &fail);
            for (int st = 1; st < 6; st++) {
                const double *b = dp ? g_dpB[st] : g_ckB[st];
                stage(S.tmp, S.y0, S.k, b, st, h, d.blk, st == 1);
                stage(S.tmpD, S.D0, S.kD, b, st, h, nD, st == 1);
                __syncthreads();
                dev_der_gwp(d, P, S.tmp, S.tmpD, S.k[st], S.kD[st], S, &fail);
            }
            stage(S.tmp, S.y0, S.k, dp ? g_dpC : g_ckC, 6, h, d.blk, false);
            stage(S.tmpD, S.D0, S.kD, dp ? g_dpC : g_ckC, 6, h, nD, false);
            if (!dp) stage(S.perr, S.y0, S.k, g_ckD, 6, h, d.blk, false);
            __syncthreads();
            if (dp) {
                dev_der_gwp(d, P, S.tmp, S.tmpD, S.k[6], S.kD[6], S, &fail);
                stage(S.perr, S.y0, S.k, g_dpD, 7, h, d.blk, false);
                __syncthreads();
            }
            error = dev_err_gwp(d, P, S.tmp, S.perr, part) / P.eps_psi;
            attempts++;
            if (error < 1.0) break;
            if (fail || attempts > P.max_attempts || !(error == error)) break;
            const double a = minscal * h, b2 = safety * h * pow(error, pshrnk);
            h = a > b2 ? a : b2;
        }
        if (fail || attempts > P.max_attempts || !(error == error)) {
            if (tid == 0 && !(error == error)) atomicOr(&fail, 1);
            break;
        }
        const double hnext = (error > errcon) ? safety * h * pow(error, pgrow) : h * maxscal;
        for (int e = tid; e < d.blk; e += nth) S.y0[e] = S.tmp[e];
        for (int e = tid; e < nD; e += nth) S.D0[e] = S.tmpD[e];
        __syncthreads();
        dt_done = dt_done + h;
        if (dt_done >= DT) break;
        double hn = hnext;
        if (hn < P.min_stepsize * 2) hn = P.min_stepsize * 2;
        stepsize = (DT - dt_done) / (double)((long long)((DT - dt_done) / hn) + 1);
        if (stepsize > DT - dt_done) stepsize = DT - dt_done;
    }
    __syncthreads();
    if (tid == 0) {
        int code = 0;
        if (fail & 2) code = 4;
        else if (fail & 1) code = 1;
        else if (attempts > P.max_attempts) code = 2;
        if (code) atomicMax(status, code);
    }
    store_mode(d, nmax, P.nm, S.y0, S.D0, psi, dd);
}

// One derivative evaluation of Gaussian mode idesc at (psi, dd).
__global__ void __launch_bounds__(64) k_gwp_der(const GwpDesc *__restrict__ descs, int idesc, const double2 *psi,
                                                const double2 *dd, double2 *dpsi, double2 *ddd, double2 *scratch,
                                                GwpParams P, int *status) {
    const GwpDesc &d = descs[idesc];
    const int nmax = P.nspfmax;
    GwpScratch S = gwp_carve(scratch, d, nmax);
    __shared__ int fail;
    if (threadIdx.x == 0) fail = 0;
    load_mode(d, nmax, P.nm, psi, dd, S.y0, S.D0);
    __syncthreads();
    dev_der_gwp(d, P, S.y0, S.D0, S.k[0], S.kD[0], S, &fail);
    store_mode(d, nmax, P.nm, S.k[0], S.kD[0], dpsi, ddd);
    if (threadIdx.x == 0 && fail) atomicMax(status, (fail & 2) ? 4 : 1);
}

// Tikhonov inverse of the intra-state overlap per (mode, state) (Calc_S00m1,
// gmctdh_objects.f90:385-404), one thread per state.
__global__ void k_gwp_s00m1(const GwpDesc *__restrict__ descs, const double2 *sd, GwpParams P, double2 *out) {
    const GwpDesc &d = descs[blockIdx.x];
    const int st = threadIdx.x;
    if (st >= P.ns) return;
    const int nmax = P.nspfmax, nn = nmax * nmax, set = d.nsets == 1 ? 0 : st, n = d.nspf[set];
    double2 A[kSmall * kSmall], U[kSmall * kSmall];
    double w[kSmall];
    double2 *o = out + (size_t)nn * (d.mode + (size_t)P.nm * st);
    for (int i = 0; i < nn; i++) o[i] = z0();
    tikh_inverse(n, sd + (size_t)nn * (d.mode + (size_t)P.nm * st), nmax, P.eps_ovl, o, A, U, w);
}

// dst(m, state) = scale * src(m, set of the state), per Gaussian mode.
__global__ void k_gwp_state_mats(const GwpDesc *__restrict__ descs, const double2 *src, double scale, GwpParams P,
                                 double2 *dst) {
    const GwpDesc &d = descs[blockIdx.x];
    const int nn = P.nspfmax * P.nspfmax;
    for (int e = threadIdx.x; e < P.ns * nn; e += blockDim.x) {
        const int st = e / nn, r = e % nn, set = d.nsets == 1 ? 0 : st;
        dst[(size_t)nn * (d.mode + (size_t)P.nm * st) + r] =
            cscale(scale, src[(size_t)nn * (d.mode + (size_t)P.nm * set) + r]);
    }
}

__global__ void k_cscale(double2 *out, const double2 *x, double2 alpha, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = cmul(alpha, x[i]);
}

}  // namespace

size_t gwp_scratch_elems(const GwpDesc &d, int nspfmax) {
    const size_t nn = (size_t)nspfmax * nspfmax, nd = nn * d.nsets;
    // y0, k[7], tmp, perr; D0, kD[7], tmpD, Sm; S; h; per-set work
    return (size_t)10 * d.blk + 10 * nd + nn * d.nsets * d.nsets + nn * d.nop + (size_t)kSetWork * d.nsets;
}

void launch_gwp_s00(const GwpDesc *d_descs, int ndesc, const double2 *psi, GwpParams p, double2 *s00, double2 *sd,
                    hipStream_t s) {
    if (ndesc > 0) hipLaunchKernelGGL(k_gwp_s00, dim3(ndesc), dim3(64), 0, s, d_descs, psi, p, s00, sd);
}

void launch_gwp_h1(const GwpDesc *d_descs, int ndesc, const double2 *psi, GwpParams p, const double2 *s00,
                   double2 *h1, hipStream_t s) {
    if (ndesc > 0) hipLaunchKernelGGL(k_gwp_h1, dim3(ndesc), dim3(64), 0, s, d_descs, psi, p, s00, h1);
}

void launch_gwp_overlaps(const GwpDesc *d_descs, int ndesc, const double2 *psi1, const double2 *psi2, int sym,
                         GwpParams p, double2 *out, hipStream_t s) {
    if (ndesc > 0) hipLaunchKernelGGL(k_gwp_overlaps, dim3(ndesc), dim3(64), 0, s, d_descs, psi1, psi2, sym, p, out);
}

void launch_gwp_dd_init(const GwpDesc *d_descs, int ndesc, const double2 *sd, GwpParams p, double2 *dd, int *status,
                        hipStream_t s) {
    if (ndesc > 0) hipLaunchKernelGGL(k_gwp_dd_init, dim3(ndesc), dim3(kMaxSets), 0, s, d_descs, sd, p, dd, status);
}

void launch_gwp_dhd(const GwpDesc *d_descs, int ndesc, const double2 *h1, const double2 *dd, GwpParams p,
                    double2 *dhd, hipStream_t s) {
    if (ndesc > 0) hipLaunchKernelGGL(k_gwp_dhd, dim3(ndesc), dim3(kMaxGwpOps), 0, s, d_descs, h1, dd, p, dhd);
}

void launch_gwp_ab_mats(const GwpDesc *d_descs, int ndesc, const double2 *sd, const double2 *dd, GwpParams p,
                        double2 *mats, hipStream_t s) {
    if (ndesc > 0) hipLaunchKernelGGL(k_gwp_ab_mats, dim3(ndesc), dim3(64), 0, s, d_descs, sd, dd, p, mats);
}

void launch_gwp_propagate(const GwpDesc *d_descs, int ndesc, double2 *psi, double2 *dd, double2 *scratch,
                          GwpParams p, int *status, hipStream_t s) {
    if (ndesc > 0)
        hipLaunchKernelGGL(k_gwp_propagate, dim3(ndesc), dim3(64), 0, s, d_descs, psi, dd, scratch, p, status);
}

void launch_gwp_der(const GwpDesc *d_descs, int idesc, const double2 *psi, const double2 *dd, double2 *dpsi,
                    double2 *ddd, double2 *scratch, GwpParams p, int *status, hipStream_t s) {
    hipLaunchKernelGGL(k_gwp_der, dim3(1), dim3(64), 0, s, d_descs, idesc, psi, dd, dpsi, ddd, scratch, p, status);
}

void launch_gwp_s00m1(const GwpDesc *d_descs, int ndesc, const double2 *sd, GwpParams p, double2 *out,
                      hipStream_t s) {
    if (ndesc > 0) hipLaunchKernelGGL(k_gwp_s00m1, dim3(ndesc), dim3(kMaxSets), 0, s, d_descs, sd, p, out);
}

void launch_gwp_state_mats(const GwpDesc *d_descs, int ndesc, const double2 *src, double scale, GwpParams p,
                           double2 *dst, hipStream_t s) {
    if (ndesc > 0) hipLaunchKernelGGL(k_gwp_state_mats, dim3(ndesc), dim3(64), 0, s, d_descs, src, scale, p, dst);
}

void launch_cscale(double2 *out, const double2 *x, double2 alpha, int64_t n, hipStream_t s) {
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_cscale, dim3((unsigned)g), dim3(256), 0, s, out, x, alpha, n);
}

}  // namespace mg
