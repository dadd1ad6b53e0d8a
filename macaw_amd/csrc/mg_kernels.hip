// Device kernels of the MI355X G-MCTDH engine (gfx950, wave64).
//
// Hot path: the sum-of-products Hamiltonian applied to the coefficient tensor
// (Calc_DHD_B, cmf2_derivatives.f90:1-80) is expressed as passes over groups
// of consecutive modes; each workgroup stages an Lt x M brick of the tensor in
// LDS and applies all mode products of the group there (k_group), so a chain
// of f mode products costs one HBM read + write per group instead of per mode.
// Hole contractions (density matrices, mean fields; gmctdh_objects.f90:408-605)
// are split-K reductions with a deterministic second stage (k_gram*).
// The SPF equations of motion and their adaptive Runge-Kutta integration
// (cmf2_integration.f90:107-412, cmf2_base.f90:210-260) run entirely on the
// device, one workgroup per logical mode (k_spf_propagate).
#include "mg_internal.h"

#include <math.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>

// 1: FP32 ops along a unit-stride mode use mfma_op8_unit (build with
// -DMG_UNIT_STRIDE_OP=0 for the general lane layout, for comparisons)
#ifndef MG_UNIT_STRIDE_OP
#define MG_UNIT_STRIDE_OP 1
#endif

namespace mg {

__device__ __forceinline__ double2 zero2() { return make_double2(0.0, 0.0); }

__device__ __forceinline__ void cfma(double2 &acc, double2 m, double2 x) {
    acc.x = fma(m.x, x.x, acc.x);
    acc.x = fma(-m.y, x.y, acc.x);
    acc.y = fma(m.x, x.y, acc.y);
    acc.y = fma(m.y, x.x, acc.y);
}

// Complex FP32: four FMAs per multiply-add (on the VALU the 3-multiply form
// needs more instructions than this).
__device__ __forceinline__ void cfma(float2 &acc, float2 m, float2 x) {
    acc.x = fmaf(m.x, x.x, acc.x);
    acc.x = fmaf(-m.y, x.y, acc.x);
    acc.y = fmaf(m.x, x.y, acc.y);
    acc.y = fmaf(m.y, x.x, acc.y);
}

// Storage-type helpers: complex FP64 (double2) or complex FP32 (float2).
template <class T> __device__ __forceinline__ T zeroT();
template <> __device__ __forceinline__ double2 zeroT<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 zeroT<float2>() { return make_float2(0.f, 0.f); }
__device__ __forceinline__ double2 to_d2(double2 a) { return a; }
__device__ __forceinline__ double2 to_d2(float2 a) { return make_double2((double)a.x, (double)a.y); }
template <class T> __device__ __forceinline__ T from_d2(double2 a);
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }
__device__ __forceinline__ float2 cmulT(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulT(double2 a, double2 b) { return cmul(a, b); }
__device__ __forceinline__ float2 caddT(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 caddT(double2 a, double2 b) { return cadd(a, b); }
__device__ __forceinline__ float2 cconjT(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ double2 cconjT(double2 a) { return cconj(a); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Global-memory loads/stores through address space 1: the task tables hand
// the kernels generic pointers, which would otherwise become flat accesses
// that also count against the LDS wait counter.
typedef unsigned long long mg_u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 gload(const float2 *p) {
    const unsigned long long v = *(const __attribute__((address_space(1))) unsigned long long *)p;
    return __builtin_bit_cast(float2, v);
}
__device__ __forceinline__ double2 gload(const double2 *p) {
    const mg_u64x2 v = *(const __attribute__((address_space(1))) mg_u64x2 *)p;
    return __builtin_bit_cast(double2, v);
}
__device__ __forceinline__ void gstore(float2 *p, float2 v) {
    *(__attribute__((address_space(1))) unsigned long long *)p = __builtin_bit_cast(unsigned long long, v);
}
__device__ __forceinline__ void gstore(double2 *p, double2 v) {
    *(__attribute__((address_space(1))) mg_u64x2 *)p = __builtin_bit_cast(mg_u64x2, v);
}
constexpr int kLoadUnroll = 8;

// ---------------------------------------------------------------------------
// 8x8 complex mode product on the matrix cores.  The complex product y = M x
// is the real 16x16 product Yr = Mr Xr, done for 16 fibers at once with four
// 16x16x4 MFMAs (K = 16 in chunks of 4).  Lane l = j + 16 q (j: fiber of the
// group, q = l / 16) holds:
//   B (chunk c): Xr[4c + q][j];  we order Xr's rows so that lane q's four
//     values are x_q.re, x_q.im, x_{q+4}.re, x_{q+4}.im (two complex loads):
//     row k -> x = 4 (k / 8) + k % 4, part = (k / 4) % 2;
//   A (chunk c): Mr[i = j][k = 4c + q], built once per op;
//   D: FP32 16x16x4 gives rows 4q + v (v = 0..3) of column j, FP64 gives rows
//     4v + q.  Output rows are ordered so that both land as two complex
//     values per lane: FP32 row i -> y = i / 2, part = i % 2 (y_{2q}, y_{2q+1});
//     FP64 row i -> y = 4 (i / 8) + i % 4, part = (i / 4) % 2 (y_q, y_{q+4}).
// Lane layouts of both builtins were checked on gfx950 with a one-wave probe.
typedef float mg_floatx4 __attribute__((ext_vector_type(4)));
typedef double mg_doublex4 __attribute__((ext_vector_type(4)));

template <class T> struct MfmaTraits;
template <> struct MfmaTraits<float2> {
    typedef float S;
    typedef mg_floatx4 V;
    static __device__ __forceinline__ V mma(S a, S b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int out_y(int i) { return i >> 1; }
    static __device__ __forceinline__ int out_part(int i) { return i & 1; }
    // complex outputs held by lane group q: (regs 0,1) and (regs 2,3)
    static __device__ __forceinline__ int y0(int q) { return 2 * q; }
    static __device__ __forceinline__ int y1(int q) { return 2 * q + 1; }
};
template <> struct MfmaTraits<double2> {
    typedef double S;
    typedef mg_doublex4 V;
    static __device__ __forceinline__ V mma(S a, S b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int out_y(int i) { return 4 * (i >> 3) + (i & 3); }
    static __device__ __forceinline__ int out_part(int i) { return (i >> 2) & 1; }
    static __device__ __forceinline__ int y0(int q) { return q; }
    static __device__ __forceinline__ int y1(int q) { return q + 4; }
};

// One op over all fibers of a brick: dst(fiber) = M src(fiber)  (acc == 0) or
// dst(fiber) += M src(fiber)  (acc == 1, dst holds the running sum).  M is the
// 8x8 op matrix in LDS, element (y, x) at M[y + 8 x].  Every lane of a wave
// takes part (MFMA needs full EXEC); lanes past nfib load zeros and skip the
// store.  In place (src == dst) is safe: a wave reads its 16 fibers before
// it writes them, and no other wave touches them.
// FP32 op along a unit-stride mode (the first mode of the leading group):
// lane group q takes components 2q and 2q+1 of its fiber in one 16-byte LDS
// access (a wave's 64 lanes read 1 KiB contiguously, no bank conflicts), the
// real embedding of the op matrix is ordered to match, and the outputs
// y_2q, y_2q+1 of the FP32 layout are contiguous too.  Same products as
// mfma_op8, summed in another order.
__device__ __forceinline__ void mfma_op8_unit(const float2 *src, float2 *dst, const float2 *M, int nfib, int acc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int j = lane & 15, q = lane >> 4;
    float a[4];
    {
        const int y = j >> 1, po = j & 1;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int x = 2 * q + (c >> 1), pi = c & 1;
            const float2 m = M[y + 8 * x];
            a[c] = (po == 0) ? (pi == 0 ? m.x : -m.y) : (pi == 0 ? m.y : m.x);
        }
    }
    for (int g = wave; g * 16 < nfib; g += nw) {
        const int f = g * 16 + j;
        const bool valid = f < nfib;
        const int base = (valid ? f : 0) * 8 + 2 * q;
        float4 xv = *reinterpret_cast<const float4 *>(src + base);
        if (!valid) xv = make_float4(0.f, 0.f, 0.f, 0.f);
        mg_floatx4 d;
        if (acc) {
            const float4 cv = *reinterpret_cast<const float4 *>(dst + base);
            d[0] = cv.x; d[1] = cv.y; d[2] = cv.z; d[3] = cv.w;
        } else {
            d[0] = 0.f; d[1] = 0.f; d[2] = 0.f; d[3] = 0.f;
        }
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], xv.x, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], xv.y, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], xv.z, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], xv.w, d, 0, 0, 0);
        if (valid) *reinterpret_cast<float4 *>(dst + base) = make_float4(d[0], d[1], d[2], d[3]);
    }
}

// UNIT: take the unit-stride path where it applies (the streaming brick
// programs; the Gram programs keep one code path: the second one costs them
// registers and occupancy, measured 18% slower mean fields)
template <class T, bool UNIT>
__device__ __forceinline__ void mfma_op8(const T *src, T *dst, const T *M, int nfib, int stride, FastDiv fds,
                                         int acc) {
    if constexpr (UNIT && std::is_same<T, float2>::value) {
        if (MG_UNIT_STRIDE_OP && stride == 1) {
            mfma_op8_unit(src, dst, M, nfib, acc);
            return;
        }
    }
    typedef MfmaTraits<T> Tr;
    typedef typename Tr::S S;
    typedef typename Tr::V V;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int j = lane & 15, q = lane >> 4;
    S a[4];
    {
        const int i = j;
        const int y = Tr::out_y(i), po = Tr::out_part(i);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int k = 4 * c + q;
            const int x = 4 * (k >> 3) + (k & 3), pi = (k >> 2) & 1;
            const T m = M[y + 8 * x];
            a[c] = (po == 0) ? (pi == 0 ? m.x : -m.y) : (pi == 0 ? m.y : m.x);
        }
    }
    const int ya = Tr::y0(q), yb = Tr::y1(q);
    for (int g = wave; g * 16 < nfib; g += nw) {
        const int f = g * 16 + j;
        const bool valid = f < nfib;
        const int ff = valid ? f : 0;
        const int outer = (int)fdiv((uint32_t)ff, fds), inner = ff - outer * stride;
        const int base = inner + outer * stride * 8;
        T x0 = src[base + q * stride];
        T x1 = src[base + (q + 4) * stride];
        if (!valid) { x0 = zeroT<T>(); x1 = zeroT<T>(); }
        V d;
        if (acc) {
            const T c0 = dst[base + ya * stride], c1 = dst[base + yb * stride];
            d[0] = c0.x; d[1] = c0.y; d[2] = c1.x; d[3] = c1.y;
        } else {
            d[0] = 0; d[1] = 0; d[2] = 0; d[3] = 0;
        }
        d = Tr::mma(a[0], x0.x, d);
        d = Tr::mma(a[1], x0.y, d);
        d = Tr::mma(a[2], x1.x, d);
        d = Tr::mma(a[3], x1.y, d);
        if (valid) {
            T o0, o1;
            o0.x = d[0]; o0.y = d[1]; o1.x = d[2]; o1.y = d[3];
            dst[base + ya * stride] = o0;
            dst[base + yb * stride] = o1;
        }
    }
}

// Bricks are stored in LDS unpadded, element i in slot i.  Measured at
// 4 x 8^8 against one padding slot every 8 values: mean fields 10.2 -> 7.2 ms
// (mixed) and 19.6 -> 13.4 ms (FP64), H.A 4.46 -> 4.07 ms; a multi-level
// padding and an XOR swizzle that a 32-bank model makes conflict-free for
// every access pattern of these kernels were both slower than no padding.
int group_capacity(int brick) { return brick; }

// ---------------------------------------------------------------------------
// Brick programs (see BrickTask).  LDS: [nmat op matrices][nbuf buffers].
template <class T>
__device__ __forceinline__ void brick_load(T *buf, const T *X, int64_t l0, int64_t r0, int64_t L, int Lt, int Rt,
                                           FastDiv fd_Lt, int M, FastDiv fd_M, int add, T alpha) {
    const int nth = blockDim.x, tid = threadIdx.x;
    const int tot = Lt * M * Rt;
    for (int base = 0; base < tot; base += nth * kLoadUnroll) {
        T v[kLoadUnroll];
        const bool full = base + nth * kLoadUnroll <= tot;
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++) {
            const int idx = base + u * nth + tid;
            const int ci = (full || idx < tot) ? idx : 0;
            const int rest = (int)fdiv((uint32_t)ci, fd_Lt), lt = ci - rest * Lt;
            const int rt = (int)fdiv((uint32_t)rest, fd_M), g = rest - rt * M;
            v[u] = gload(X + (l0 + lt + L * ((int64_t)g + (int64_t)M * (r0 + rt))));
        }
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++) {
            const int idx = base + u * nth + tid;
            if (full || idx < tot) {
                if (add) buf[idx] = caddT(buf[idx], cmulT(alpha, v[u]));
                else buf[idx] = v[u];
            }
        }
    }
}

template <class T>
__device__ __forceinline__ void brick_store(const T *buf, T *Y, int64_t l0, int64_t r0, int64_t L, int Lt, int Rt,
                                            FastDiv fd_Lt, int M, FastDiv fd_M, int acc, T alpha) {
    const int nth = blockDim.x, tid = threadIdx.x;
    const int tot = Lt * M * Rt;
    for (int base = 0; base < tot; base += nth * kLoadUnroll) {
        int64_t a[kLoadUnroll];
        T yv[kLoadUnroll];
        const bool full = base + nth * kLoadUnroll <= tot;
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++) {
            const int idx = base + u * nth + tid;
            const int ci = (full || idx < tot) ? idx : 0;
            const int rest = (int)fdiv((uint32_t)ci, fd_Lt), lt = ci - rest * Lt;
            const int rt = (int)fdiv((uint32_t)rest, fd_M), g = rest - rt * M;
            a[u] = l0 + lt + L * ((int64_t)g + (int64_t)M * (r0 + rt));
            if (acc) yv[u] = gload(Y + a[u]);
        }
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++) {
            const int idx = base + u * nth + tid;
            if (full || idx < tot) {
                T v = cmulT(alpha, buf[idx]);
                if (acc) v = caddT(yv[u], v);
                gstore(Y + a[u], v);
            }
        }
    }
}

// Generic mode product over the fibers of one mode (any sizes <= NMAX).
template <int NMAX, class T>
__device__ __forceinline__ void brick_op_generic(const T *src, T *dst, const T *Mq, int ni, int no, int nfib,
                                                 int stride, FastDiv fds, int acc) {
    const int nth = blockDim.x, tid = threadIdx.x;
    for (int f = tid; f < nfib; f += nth) {
        const int outer = (int)fdiv((uint32_t)f, fds), inner = f - outer * stride;
        const int bi = inner + outer * stride * ni;
        const int bo = inner + outer * stride * no;
        T x[NMAX];
#pragma unroll
        for (int j = 0; j < NMAX; j++) x[j] = (j < ni) ? src[bi + j * stride] : zeroT<T>();
#pragma unroll
        for (int i = 0; i < NMAX; i++) {
            if (i < no) {
                T r = acc ? dst[bo + i * stride] : zeroT<T>();
#pragma unroll
                for (int j = 0; j < NMAX; j++)
                    if (j < ni) cfma(r, Mq[i + NMAX * j], x[j]);
                dst[bo + i * stride] = r;
            }
        }
    }
}

// Gram accumulation on the matrix cores: G = X^H Y over the fibers of one
// mode is the real 16x16 product [Xr Xi]^T [Yr Yi] (K = fibers, 4 per MFMA):
// Re G = D[a][b] + D[a+8][b+8], Im G = D[a][b+8] - D[a+8][b].  Lane l holds
// the real part (l % 16 < 8) or imaginary part of element l % 8 of fiber
// 4 s + l / 16 for both operands.  FP32 MFMA within a brick (<= 256 fibers),
// FP64 across bricks.
template <class T> struct GramMfma;
template <> struct GramMfma<float2> {
    // X == Y (density matrices): one LDS load per step instead of two
    static __device__ __forceinline__ void run(const float2 *X, const float2 *Y, int n1, int n2, int nfib,
                                               int stride, FastDiv fds, double acc[4]) {
        if (X == Y) run_t<true>(X, Y, n1, n2, nfib, stride, fds, acc);
        else run_t<false>(X, Y, n1, n2, nfib, stride, fds, acc);
    }
    template <bool SAME>
    static __device__ __forceinline__ void run_t(const float2 *X, const float2 *Y, int n1, int n2, int nfib,
                                                 int stride, FastDiv fds, double acc[4]) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
        const int i = lane & 15, q = lane >> 4, e = i & 7, im = i >> 3;
        mg_floatx4 d = {0.f, 0.f, 0.f, 0.f};
        const int nsteps = (nfib + 3) >> 2;
        // four steps' loads in flight before their MFMAs (same order of
        // accumulation; steps past the end add exact zeros)
        constexpr int U = 4;
        for (int st0 = wave; st0 < nsteps; st0 += U * nw) {
            float av[U], bv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int st = st0 + u * nw;
                const int f = 4 * st + q;
                const bool ok = st < nsteps && f < nfib;
                const int ff = ok ? f : 0;
                const int outer = (int)fdiv((uint32_t)ff, fds), inner = ff - outer * stride;
                const float2 xv = X[inner + outer * stride * n1 + (e < n1 ? e : 0) * stride];
                const float2 yv = SAME ? xv : Y[inner + outer * stride * n2 + (e < n2 ? e : 0) * stride];
                av[u] = (ok && e < n1) ? (im ? xv.y : xv.x) : 0.f;
                bv[u] = (ok && e < n2) ? (im ? yv.y : yv.x) : 0.f;
            }
#pragma unroll
            for (int u = 0; u < U; u++) d = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], bv[u], d, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) acc[v] += (double)d[v];
    }
    // D row held in register v by lane group q
    static __device__ __forceinline__ int row(int q, int v) { return 4 * q + v; }
};
template <> struct GramMfma<double2> {
    // X == Y (density matrices): one LDS load per step instead of two
    static __device__ __forceinline__ void run(const double2 *X, const double2 *Y, int n1, int n2, int nfib,
                                               int stride, FastDiv fds, double acc[4]) {
        if (X == Y) run_t<true>(X, Y, n1, n2, nfib, stride, fds, acc);
        else run_t<false>(X, Y, n1, n2, nfib, stride, fds, acc);
    }
    template <bool SAME>
    static __device__ __forceinline__ void run_t(const double2 *X, const double2 *Y, int n1, int n2, int nfib,
                                                 int stride, FastDiv fds, double acc[4]) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
        const int i = lane & 15, q = lane >> 4, e = i & 7, im = i >> 3;
        mg_doublex4 d = {0.0, 0.0, 0.0, 0.0};
        const int nsteps = (nfib + 3) >> 2;
        // four steps' loads in flight before their MFMAs (same order of
        // accumulation; steps past the end add exact zeros)
        constexpr int U = 4;
        for (int st0 = wave; st0 < nsteps; st0 += U * nw) {
            double av[U], bv[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int st = st0 + u * nw;
                const int f = 4 * st + q;
                const bool ok = st < nsteps && f < nfib;
                const int ff = ok ? f : 0;
                const int outer = (int)fdiv((uint32_t)ff, fds), inner = ff - outer * stride;
                const double2 xv = X[inner + outer * stride * n1 + (e < n1 ? e : 0) * stride];
                const double2 yv = SAME ? xv : Y[inner + outer * stride * n2 + (e < n2 ? e : 0) * stride];
                av[u] = (ok && e < n1) ? (im ? xv.y : xv.x) : 0.0;
                bv[u] = (ok && e < n2) ? (im ? yv.y : yv.x) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; u++) d = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], d, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) acc[v] += d[v];
    }
    static __device__ __forceinline__ int row(int q, int v) { return 4 * v + q; }
};

// Workgroups are dispatched round-robin over the 8 XCDs (linear workgroup id
// modulo 8), and each XCD has its own L2.  With a grid width that is a
// multiple of 8, block x runs on XCD x % 8 whatever its task row; the remap
// gives the blocks of one XCD consecutive logical indices, so the bricks a
// XCD works on at the same time are neighbours in memory and short runs
// (32-byte rows of a strided brick) complete their 128-byte lines in that
// XCD's L2 instead of in several.  A bijection on [0, G); identity otherwise.
__device__ __forceinline__ int64_t xcd_block(int b, int G) {
    if (G & 7) return b;
    return (int64_t)(b & 7) * (G >> 3) + (b >> 3);
}

// GRAM: the launch carries kBGram instructions; the grid is persistent over
// the bricks and every workgroup ends by writing its Gram partials.
// NG: Gram slots the program's accumulators hold (kMaxGrams, or 4 when no
// task of the launch has more: fewer registers, more waves per SIMD)
template <int NMAX, class T, bool GRAM, int NG = kMaxGrams>
__global__ void __launch_bounds__(GRAM ? 256 : 1024) k_brick(const BrickTask *__restrict__ tasks, int capacity, int nmat_lds,
                                                double2 *__restrict__ partials, const int *skip) {
    if (!GRAM && skip && *skip) return;
    const BrickTask &t = tasks[blockIdx.y];
    const int64_t L = t.L, R = t.R;
    const int Lt = (int)t.Lt, Rt = (int)t.Rt;
    const int64_t nlb = L / Lt, nrb = R / Rt;
    const int64_t nbricks = nlb * nrb;
    if (!GRAM && xcd_block(blockIdx.x, gridDim.x) >= nbricks) return;
    const FastDiv fd_Lt = t.fd_Lt;
    extern __shared__ char lds_raw[];
    T *mats = reinterpret_cast<T *>(lds_raw);
    T *bufs = mats + (size_t)nmat_lds * NMAX * NMAX;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nmat = t.nmat, ld = t.ld;
    for (int e = tid; e < nmat * NMAX * NMAX; e += nth) {
        const int q = e / (NMAX * NMAX), r = e - q * NMAX * NMAX;
        const int i = r % NMAX, j = r / NMAX;
        T v = zeroT<T>();
        if (i < t.matnout[q] && j < t.matnin[q]) {
            const T *Mg = static_cast<const T *>(t.mat[q]);
            v = t.matadj[q] ? cconjT(Mg[j + ld * i]) : Mg[i + ld * j];
            v = cmulT(from_d2<T>(t.matscale[q]), v);
        }
        mats[e] = v;
    }
    __syncthreads();
    double gacc[GRAM ? NG : 1][4];
#pragma unroll
    for (int g = 0; g < (GRAM ? NG : 1); g++)
#pragma unroll
        for (int v = 0; v < 4; v++) gacc[g][v] = 0.0;
    const int nins = t.ninstr;
    // a workgroup walks its task's bricks with the grid stride, so the op
    // matrices staged in LDS above serve every brick it takes
    for (int64_t bi = xcd_block(blockIdx.x, gridDim.x); bi < nbricks; bi += gridDim.x) {
        const int64_t l0 = (bi % nlb) * Lt, r0 = (bi / nlb) * Rt;
        for (int k = 0; k < nins; k++) {
            const BrickInstr &I = t.ins[k];
            const int code = I.code;
            T *A = bufs + (size_t)I.a * capacity;
            T *B = bufs + (size_t)I.b * capacity;
            if (code == kBLoad || code == kBLoadAdd) {
                brick_load(A, static_cast<const T *>(I.ptr), l0, r0, L, Lt, Rt, fd_Lt, I.M, I.fd_M,
                           code == kBLoadAdd, from_d2<T>(I.alpha));
            } else if (code == kBOp) {
                const T *Mq = mats + I.mat * NMAX * NMAX;
                if (NMAX == 8 && I.ni == 8 && I.no == 8)
                    mfma_op8<T, !GRAM>(A, B, Mq, I.nfib, I.stride, I.fd_stride, I.acc);
                else
                    brick_op_generic<NMAX, T>(A, B, Mq, I.ni, I.no, I.nfib, I.stride, I.fd_stride, I.acc);
            } else if (code == kBAxpy) {
                const int tot = Lt * I.M * Rt;
                const T al = from_d2<T>(I.alpha);
                for (int idx = tid; idx < tot; idx += nth)
                    B[idx] = caddT(B[idx], cmulT(al, A[idx]));
            } else if (code == kBStore) {
                brick_store(A, static_cast<T *>(const_cast<void *>(I.ptr)), l0, r0, L, Lt, Rt, fd_Lt, I.M, I.fd_M,
                            I.acc, from_d2<T>(I.alpha));
            } else if (code == kBZero) {
                const int tot = Lt * I.M * Rt;
                for (int idx = tid; idx < tot; idx += nth) A[idx] = zeroT<T>();
            } else if (GRAM && code == kBGram) {
                double part[4] = {0.0, 0.0, 0.0, 0.0};
                GramMfma<T>::run(A, B, I.ni, I.no, I.nfib, I.stride, I.fd_stride, part);
#pragma unroll
                for (int g = 0; g < (GRAM ? NG : 1); g++)
                    if (g == I.mat)
#pragma unroll
                        for (int v = 0; v < 4; v++) gacc[g][v] += part[v];
            }
            __syncthreads();
        }
    }
    if (GRAM) {
        // per slot: waves' 16x16 partial D -> LDS -> 8x8 complex partial
        double *red = reinterpret_cast<double *>(lds_raw);
        const int lane = tid & 63, wave = tid >> 6, nw = nth >> 6;
        const int q = lane >> 4, col = lane & 15;
#pragma unroll
        for (int g = 0; g < (GRAM ? NG : 1); g++) {
            if (g >= t.ngram) break;
#pragma unroll
            for (int v = 0; v < 4; v++) red[wave * 256 + GramMfma<T>::row(q, v) * 16 + col] = gacc[g][v];
            __syncthreads();
            if (tid < 64) {
                const int a = tid & 7, b = tid >> 3;
                double d00 = 0.0, d11 = 0.0, d01 = 0.0, d10 = 0.0;
                for (int w = 0; w < nw; w++) {
                    const double *D = red + w * 256;
                    d00 += D[a * 16 + b];
                    d11 += D[(a + 8) * 16 + b + 8];
                    d01 += D[a * 16 + b + 8];
                    d10 += D[(a + 8) * 16 + b];
                }
                partials[((int64_t)(t.gram_base + g) * 64 + tid) * gridDim.x + blockIdx.x] =
                    make_double2(d00 + d11, d01 - d10);
            }
            __syncthreads();
        }
    }
}

template <int NMAX, class T, bool GRAM, int NG>
static void launch_brick_nmax(const BrickTask *d_tasks, int ntasks, int64_t max_blocks, int capacity, int nbuf,
                              int nmat, hipStream_t s, double2 *partials, const int *skip) {
    static const int nthreads = [] {
        const char *e = std::getenv("MG_THREADS");
        const int v = e ? std::atoi(e) : 256;
        return (v == 128 || v == 256 || v == 512 || v == 1024) ? v : 256;
    }();
    size_t shm = sizeof(T) * ((size_t)nmat * NMAX * NMAX + (size_t)nbuf * capacity);
    if (GRAM) shm = std::max(shm, sizeof(double) * 256 * 4);
    // Thread-safe one-time initialisation (ensemble members launch from several host threads).
    static const bool attr_set = [] {
        (void)hipFuncSetAttribute((const void *)k_brick<NMAX, T, GRAM, NG>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        return true;
    }();
    (void)attr_set;
    hipLaunchKernelGGL((k_brick<NMAX, T, GRAM, NG>), dim3((unsigned)max_blocks, (unsigned)ntasks), dim3(GRAM ? 256 : nthreads), shm,
                       s, d_tasks, capacity, nmat, partials, skip);
}

template <class T, bool GRAM, int NG = kMaxGrams>
static void launch_brick_T(const BrickTask *d_tasks, int ntasks, int64_t mb, int nmax, int capacity, int nbuf,
                           int nmat, hipStream_t s, double2 *partials, const int *skip) {
    if (nmax <= 4) launch_brick_nmax<4, T, GRAM, NG>(d_tasks, ntasks, mb, capacity, nbuf, nmat, s, partials, skip);
    else if (nmax <= 8) launch_brick_nmax<8, T, GRAM, NG>(d_tasks, ntasks, mb, capacity, nbuf, nmat, s, partials, skip);
    else launch_brick_nmax<16, T, GRAM, NG>(d_tasks, ntasks, mb, capacity, nbuf, nmat, s, partials, skip);
}

void launch_brick_tasks(const BrickTask *d_tasks, int ntasks, int64_t max_blocks_per_task, int nmax, int brick,
                        int nbuf, int nmat, int prec, hipStream_t s, int ngram_total, double2 *partials,
                        int gram_blocks, const int *skip, int gram_task_max) {
    if (ntasks <= 0 || max_blocks_per_task <= 0) return;
    const int capacity = group_capacity(brick);
    if (ngram_total > 0) {
        // Gram launches always run (their partials feed a final reduction)
        const bool ng4 = gram_task_max <= 4;
        if (prec == 0) {
            if (ng4)
                launch_brick_T<double2, true, 4>(d_tasks, ntasks, gram_blocks, nmax, capacity, nbuf, nmat, s,
                                                 partials, nullptr);
            else
                launch_brick_T<double2, true>(d_tasks, ntasks, gram_blocks, nmax, capacity, nbuf, nmat, s, partials,
                                              nullptr);
        } else {
            if (ng4)
                launch_brick_T<float2, true, 4>(d_tasks, ntasks, gram_blocks, nmax, capacity, nbuf, nmat, s,
                                                partials, nullptr);
            else
                launch_brick_T<float2, true>(d_tasks, ntasks, gram_blocks, nmax, capacity, nbuf, nmat, s, partials,
                                             nullptr);
        }
        return;
    }
    // workgroups per task: at most 4096 (complex FP32) / 1024 (complex FP64),
    // each walking several bricks so that its staged op matrices serve them
    // all (measured on the 4 x 8^8 H.A: -2.5% FP32, -8% FP64 against one
    // workgroup per brick); MG_BRICK_GRID overrides, 0 = one per brick
    static const int64_t grid_env = [] {
        const char *e = std::getenv("MG_BRICK_GRID");
        return e ? std::max<int64_t>(0, std::atoll(e)) : (int64_t)-1;
    }();
    const int64_t grid_cap = grid_env >= 0 ? grid_env : (prec ? 4096 : 1024);
    const int64_t gx = grid_cap > 0 ? std::min(max_blocks_per_task, grid_cap) : max_blocks_per_task;
    if (prec == 0)
        launch_brick_T<double2, false>(d_tasks, ntasks, gx, nmax, capacity, nbuf, nmat, s, nullptr, skip);
    else
        launch_brick_T<float2, false>(d_tasks, ntasks, gx, nmax, capacity, nbuf, nmat, s, nullptr, skip);
}

// One wave per (slot, entry): the slot's block partials summed in a fixed
// order into targets[slot][a + ld b] (a, b < min(8, ld)).
__global__ void __launch_bounds__(256) k_gram_slots_final(const double2 *partials, int nslots, int nblocks,
                                                          double2 *const *targets, int ld) {
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= nslots * 64) return;
    const int slot = w >> 6, e = w & 63, a = e & 7, b = e >> 3;
    const double2 *p = partials + (int64_t)w * nblocks;
    double sx = 0.0, sy = 0.0;
    for (int k = lane; k < nblocks; k += 64) {
        sx += p[k].x;
        sy += p[k].y;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0 && a < ld && b < ld) targets[slot][a + ld * b] = make_double2(sx, sy);
}

void launch_gram_slots_final(const double2 *partials, int nslots, int nblocks, double2 *const *targets, int ld,
                             hipStream_t s) {
    if (nslots <= 0) return;
    const int waves = nslots * 64;
    hipLaunchKernelGGL(k_gram_slots_final, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, partials, nslots,
                       nblocks, targets, ld);
}

// Exact Hermitian symmetry for a batch of n x n matrices (upper triangle wins).
__global__ void k_herm_batch(double2 *const *targets, const int *nv, int ld) {
    double2 *G = targets[blockIdx.x];
    const int n = nv[blockIdx.x];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int a = e % n, b = e / n;
        if (a == b) G[a + ld * a].y = 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int a = e % n, b = e / n;
        if (a > b) G[a + ld * b] = cconj(G[b + ld * a]);
    }
}

void launch_herm_batch(double2 *const *targets, const int *n, int count, int ld, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_herm_batch, dim3(count), dim3(256), 0, s, targets, n, ld);
}

// dst[t] = sum_{c in row t} coef[c] * (trans[c] ? src[c]^H : src[c]), ld x ld
// matrices, overwriting dst[t]; one thread per (t, entry), fixed order.
__global__ void k_mat_combine(double2 *const *dst, const int *rowptr, const double2 *const *src, const double2 *coef,
                              const int *trans, int ndst, int ld) {
    const int nn = ld * ld;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= (int64_t)ndst * nn) return;
    const int t = (int)(id / nn), e = (int)(id % nn);
    const int a = e % ld, b = e / ld;
    double2 acc = zero2();
    for (int c = rowptr[t]; c < rowptr[t + 1]; c++) {
        const double2 v = trans[c] ? cconj(src[c][b + ld * a]) : src[c][e];
        acc = cadd(acc, cmul(coef[c], v));
    }
    dst[t][e] = acc;
}

void launch_mat_combine(double2 *const *dst, const int *rowptr, const double2 *const *src, const double2 *coef,
                        const int *trans, int ndst, int ld, hipStream_t s) {
    if (ndst <= 0) return;
    const int64_t tot = (int64_t)ndst * ld * ld;
    hipLaunchKernelGGL(k_mat_combine, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, dst, rowptr, src, coef,
                       trans, ndst, ld);
}

// ---------------------------------------------------------------------------
// Single-mode product straight from global memory (one thread per fiber).
template <int NMAX, class T>
__global__ void __launch_bounds__(256) k_mode_product(ModeTask t) {
    __shared__ T M[NMAX * NMAX];
    const T *Mg = static_cast<const T *>(t.M);
    for (int e = threadIdx.x; e < NMAX * NMAX; e += blockDim.x) {
        int i = e % NMAX, j = e / NMAX;
        T v = zeroT<T>();
        if (i < t.nout && j < t.nin) v = t.adj ? cconjT(Mg[j + t.ld * i]) : Mg[i + t.ld * j];
        M[e] = v;
    }
    __syncthreads();
    const T *X = static_cast<const T *>(t.X);
    T *Y = static_cast<T *>(t.Y);
    const uint32_t nfib = (uint32_t)(t.ll * t.nl);
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nfib; f += gridDim.x * blockDim.x) {
        const uint32_t r = fdiv(f, t.fd_ll), l = f - r * (uint32_t)t.ll;
        const T *x0 = X + l + t.ll * (int64_t)t.nin * r;
        T *y0 = Y + l + t.ll * (int64_t)t.nout * r;
        T x[NMAX], y[NMAX];
#pragma unroll
        for (int j = 0; j < NMAX; j++) x[j] = (j < t.nin) ? gload(x0 + t.ll * j) : zeroT<T>();
#pragma unroll
        for (int i = 0; i < NMAX; i++) {
            T acc = zeroT<T>();
#pragma unroll
            for (int j = 0; j < NMAX; j++)
                if (j < t.nin) cfma(acc, M[i + NMAX * j], x[j]);
            y[i] = acc;
        }
        // all outputs first, then the stores (no store-data register reuse)
#pragma unroll
        for (int i = 0; i < NMAX; i++)
            if (i < t.nout) gstore(y0 + t.ll * i, y[i]);
    }
}

void launch_mode_product(const ModeTask &t0, int nmax, int prec, hipStream_t s) {
    ModeTask t = t0;
    t.fd_ll = make_fastdiv((uint32_t)t.ll);
    const int64_t nfib = t.ll * t.nl;
    int64_t blocks = (nfib + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    const dim3 g((unsigned)blocks), b(256);
    if (prec == 0) {
        if (nmax <= 4) hipLaunchKernelGGL((k_mode_product<4, double2>), g, b, 0, s, t);
        else if (nmax <= 8) hipLaunchKernelGGL((k_mode_product<8, double2>), g, b, 0, s, t);
        else hipLaunchKernelGGL((k_mode_product<16, double2>), g, b, 0, s, t);
    } else {
        if (nmax <= 4) hipLaunchKernelGGL((k_mode_product<4, float2>), g, b, 0, s, t);
        else if (nmax <= 8) hipLaunchKernelGGL((k_mode_product<8, float2>), g, b, 0, s, t);
        else hipLaunchKernelGGL((k_mode_product<16, float2>), g, b, 0, s, t);
    }
}

// ---------------------------------------------------------------------------
// Hole contraction, stage 1: per-block partial n1 x n2 sums (chunked to 8x8),
// accumulated in FP64 whatever the storage type.

// The positions (l, r) of the contraction are cut into tiles of Lt x Rt
// (Lt | ll, and Rt | nl when Lt == ll) so that each tile is a set of
// contiguous runs in memory; a tile of X (n1c planes) and of Y (n2c planes) is
// staged in LDS as [plane][position], converted to FP64 on use.  Wave w owns
// the 4x4 sub-block (4*(w&1), 4*(w>>1)) of the 8x8 chunk and keeps 16 FP64
// accumulators per lane; lanes stride over the tile's positions.
struct GramTile {
    int Lt, Rt;
    int64_t ntl, ntiles;         // tiles along l, total
    FastDiv fd_Lt, fd_n1c, fd_n2c;
};

template <class T>
__device__ __forceinline__ void gram_load_tile(T *dst, const T *src, int64_t l0, int64_t r0, int64_t ll, int n,
                                               int jo, int nc, const GramTile &g, FastDiv fd_nc) {
    const int TP = g.Lt * g.Rt;
    const int tot = TP * nc;
    const int nth = blockDim.x;
    for (int base = 0; base < tot; base += nth * kLoadUnroll) {
        T v[kLoadUnroll];
        int d[kLoadUnroll];
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++) {
            const int idx = base + u * nth + threadIdx.x;
            const int ci = idx < tot ? idx : 0;
            const int rest = (int)fdiv((uint32_t)ci, g.fd_Lt), lt = ci - rest * g.Lt;
            const int rt = (int)fdiv((uint32_t)rest, fd_nc), a = rest - rt * nc;
            d[u] = idx < tot ? a * TP + lt + g.Lt * rt : -1;
            v[u] = gload(src + (l0 + lt + ll * ((int64_t)(jo + a) + (int64_t)n * (r0 + rt))));
        }
#pragma unroll
        for (int u = 0; u < kLoadUnroll; u++)
            if (d[u] >= 0) dst[d[u]] = v[u];
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_gram_partial(GramTask t, int j1o, int n1c, int j2o, int n2c, GramTile g,
                                                      double2 *partials) {
    extern __shared__ char lds_raw[];
    const int TP = g.Lt * g.Rt;
    T *xs = reinterpret_cast<T *>(lds_raw);
    const bool same = (t.X == t.Y) && (j1o == j2o);
    T *ys = same ? xs : xs + n1c * TP;
    const T *X = static_cast<const T *>(t.X);
    const T *Yv = static_cast<const T *>(t.Y);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int A = (wave & 1) * 4, B = (wave >> 1) * 4;
    const bool active = (A < n1c) && (B < n2c);
    double2 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = zero2();
    for (int64_t tile = blockIdx.x; tile < g.ntiles; tile += gridDim.x) {
        const int64_t lb = tile % g.ntl, rb = tile / g.ntl;
        const int64_t l0 = lb * g.Lt, r0 = rb * g.Rt;
        gram_load_tile(xs, X, l0, r0, t.ll, t.n1, j1o, n1c, g, g.fd_n1c);
        if (!same) gram_load_tile(ys, Yv, l0, r0, t.ll, t.n2, j2o, n2c, g, g.fd_n2c);
        __syncthreads();
        if (active) {
            for (int q = lane; q < TP; q += 64) {
                double2 x[4], y[4];
#pragma unroll
                for (int i = 0; i < 4; i++) x[i] = (A + i < n1c) ? to_d2(xs[(A + i) * TP + q]) : zero2();
#pragma unroll
                for (int j = 0; j < 4; j++) y[j] = (B + j < n2c) ? to_d2(ys[(B + j) * TP + q]) : zero2();
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        acc[i][j].x = fma(x[i].x, y[j].x, acc[i][j].x);
                        acc[i][j].x = fma(x[i].y, y[j].y, acc[i][j].x);
                        acc[i][j].y = fma(x[i].x, y[j].y, acc[i][j].y);
                        acc[i][j].y = fma(-x[i].y, y[j].x, acc[i][j].y);
                    }
            }
        }
        __syncthreads();
    }
    if (!active) return;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double sx = wave_sum(acc[i][j].x), sy = wave_sum(acc[i][j].y);
            if (lane == 0 && A + i < n1c && B + j < n2c)
                partials[(int64_t)((A + i) + 8 * (B + j)) * gridDim.x + blockIdx.x] = make_double2(sx, sy);
        }
}

// Stage 2: one wave per matrix entry sums that entry's block partials in a
// fixed order (deterministic), then applies alpha / accumulate.
__global__ void __launch_bounds__(256) k_gram_final(GramTask t, int j1o, int n1c, int j2o, int n2c,
                                                    const double2 *partials, int nblocks) {
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (e >= 64) return;
    const int a = e % 8, b = e / 8;
    if (a >= n1c || b >= n2c) return;
    const double2 *p = partials + (int64_t)e * nblocks;
    double sx = 0.0, sy = 0.0;
    for (int k = lane; k < nblocks; k += 64) {
        sx += p[k].x;
        sy += p[k].y;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane != 0) return;
    double2 v = cmul(t.alpha, make_double2(sx, sy));
    const int j1 = j1o + a, j2 = j2o + b;
    double2 *g = t.G + j1 + (int64_t)t.ld * j2;
    if (t.accumulate) v = cadd(*g, v);
    *g = v;
}

__global__ void k_gram_hermitize(double2 *G, int n, int ld) {
    const int e = threadIdx.x;
    const bool in = e < n * n;
    const int a = in ? e % n : 0, b = in ? e / n : 0;
    if (in && a == b) G[a + ld * a].y = 0.0;
    __syncthreads();
    if (in && a > b) G[a + ld * b] = cconj(G[b + ld * a]);
}

static int64_t largest_divisor_le(int64_t L, int64_t cap) {
    if (cap >= L) return L;
    for (int64_t d = cap; d >= 1; d--)
        if (L % d == 0) return d;
    return 1;
}

void launch_gram(const GramTask &t0, double2 *partials, int nblocks, int nmax, int prec, hipStream_t s) {
    (void)nmax;
    GramTask t = t0;
    t.fd_ll = make_fastdiv((uint32_t)t.ll);
    // tile: ~256 positions (FP32) or 128 (FP64): 2 x 8 planes stay <= 32 KB
    const int64_t target = prec ? 256 : 128;
    GramTile g;
    g.Lt = (int)largest_divisor_le(t.ll, target);
    g.Rt = (g.Lt == t.ll) ? (int)largest_divisor_le(t.nl, std::max<int64_t>(1, target / t.ll)) : 1;
    g.ntl = t.ll / g.Lt;
    g.ntiles = g.ntl * (t.nl / g.Rt);
    g.fd_Lt = make_fastdiv((uint32_t)g.Lt);
    const int TP = g.Lt * g.Rt;
    int nb = nblocks;
    if (g.ntiles < nb) nb = (int)g.ntiles;
    if (nb < 1) nb = 1;
    for (int j1o = 0; j1o < t.n1; j1o += 8) {
        const int n1c = (t.n1 - j1o) < 8 ? (t.n1 - j1o) : 8;
        for (int j2o = 0; j2o < t.n2; j2o += 8) {
            const int n2c = (t.n2 - j2o) < 8 ? (t.n2 - j2o) : 8;
            g.fd_n1c = make_fastdiv((uint32_t)n1c);
            g.fd_n2c = make_fastdiv((uint32_t)n2c);
            const bool same = (t.X == t.Y) && (j1o == j2o);
            const size_t planes = (size_t)n1c + (same ? 0 : (size_t)n2c);
            if (prec == 0)
                hipLaunchKernelGGL(k_gram_partial<double2>, dim3(nb), dim3(256), sizeof(double2) * planes * TP, s, t,
                                   j1o, n1c, j2o, n2c, g, partials);
            else
                hipLaunchKernelGGL(k_gram_partial<float2>, dim3(nb), dim3(256), sizeof(float2) * planes * TP, s, t,
                                   j1o, n1c, j2o, n2c, g, partials);
            hipLaunchKernelGGL(k_gram_final, dim3(16), dim3(256), 0, s, t, j1o, n1c, j2o, n2c, partials, nb);
        }
    }
    if (t.hermitian) hipLaunchKernelGGL(k_gram_hermitize, dim3(1), dim3(256), 0, s, t.G, t.n1, t.ld);
}

// ---------------------------------------------------------------------------
// SPF block products (Calc_h1mode DVR branch, hamilton.f90:433-485, and the
// SPF overlap matrices of norm_wf / WF_Ovl).  Always FP64.
__device__ void apply_prim_block(const DofTables &dt, int mode, int k, int code, int nS, int nG,
                                 double2 *blk, double2 *tmp) {
    const int dof = dt.mode_dofs[mode * kMaxDofPerMode + k];
    int ll = 1, nl = 1;
    for (int q = 0; q < k; q++) ll *= dt.ngp[dt.mode_dofs[mode * kMaxDofPerMode + q]];
    for (int q = k + 1; q < dt.mode_ndof[mode]; q++) nl *= dt.ngp[dt.mode_dofs[mode * kMaxDofPerMode + q]];
    const int ng = dt.ngp[dof];
    const int total = nG * nS;
    const int tid = threadIdx.x, nth = blockDim.x;
    if (code > 200) {
        // parametrised diagonal operator: its values on this DOF's grid
        // (MultTensDiag with gridOper, hamilton.f90:467-468)
        const double *f = dt.oper + (int64_t)dt.ngpmax * ((code - 201) + (int64_t)dt.nparam * dof);
        for (int e = tid; e < total; e += nth) blk[e] = cscale(f[(e / ll) % ng], blk[e]);
        __syncthreads();
        return;
    }
    if (code <= -200) {
        // projector phi phi^T along this DOF (MultTensProj, hamilton.f90:632-661)
        const double *phi = dt.oper + (int64_t)dt.ngpmax * ((-code - 201) + (int64_t)dt.nparam * dof);
        for (int e = tid; e < total; e += nth) {
            const int l = e % ll;
            const int g2 = (e / ll) % ng;
            const int rest = e / (ll * ng);
            const double2 *c = blk + (int64_t)rest * ng * ll + l;
            double2 acc = zero2();
            for (int g = 0; g < ng; g++) {
                acc.x = fma(phi[g], c[g * ll].x, acc.x);
                acc.y = fma(phi[g], c[g * ll].y, acc.y);
            }
            tmp[e] = cscale(phi[g2], acc);
        }
        __syncthreads();
        for (int e = tid; e < total; e += nth) blk[e] = tmp[e];
        __syncthreads();
        return;
    }
    if (code >= 1) {
        const double *x = dt.grid + (int64_t)dt.ngpmax * dof;
        for (int e = tid; e < total; e += nth) {
            const int g = (e / ll) % ng;
            double v = x[g];
            for (int q = 1; q < code; q++) v *= x[g];
            blk[e] = cscale(v, blk[e]);
        }
        __syncthreads();
        return;
    }
    const double *Mat = (code == -1 ? dt.d1 : (code == -2 ? dt.d2 : dt.xd1)) +
                        (int64_t)dt.ngpmax * dt.ngpmax * dof;
    const bool imag_fac = (code == -1 || code == -3);
    (void)nl;
    // tmp(l, g2, rest) = sum_g blk(l, g, rest) * Mat(g, g2)
    for (int e = tid; e < total; e += nth) {
        const int l = e % ll;
        const int g2 = (e / ll) % ng;
        const int rest = e / (ll * ng);
        const double2 *c = blk + (int64_t)rest * ng * ll + l;
        double2 acc = zero2();
        for (int g = 0; g < ng; g++) {
            const double m = Mat[g + (int64_t)dt.ngpmax * g2];
            acc.x = fma(m, c[g * ll].x, acc.x);
            acc.y = fma(m, c[g * ll].y, acc.y);
        }
        tmp[e] = imag_fac ? make_double2(-acc.y, acc.x) : acc;
    }
    __syncthreads();
    for (int e = tid; e < total; e += nth) blk[e] = tmp[e];
    __syncthreads();
}

__global__ void __launch_bounds__(256) k_spf_products(const SpfTask *__restrict__ tasks, DofTables dt) {
    const SpfTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, nth = blockDim.x;
    double2 *sect = t.scratch, *tmp = t.scratch + (int64_t)t.nG * t.nb;
    const int total = t.nG * t.nb;
    for (int e = tid; e < total; e += nth) sect[e] = t.phiB[e];
    __syncthreads();
    for (int k = 0; k < t.ncode; k++)
        if (t.codes[k] != 0) apply_prim_block(dt, t.mode, k, t.codes[k], t.nb, t.nG, sect, tmp);
    for (int e = tid; e < t.na * t.nb; e += nth) {
        const int a = e % t.na, b = e / t.na;
        if (t.sym && a > b) continue;
        double2 x = zero2(), y = zero2();
        for (int g = 0; g < t.nG; g++) {
            x = cadd(x, cconjmul(t.phiA[g + t.nG * a], sect[g + t.nG * b]));
            if (t.sym) y = cadd(y, cconjmul(sect[g + t.nG * a], t.phiB[g + t.nG * b]));
        }
        if (t.sym) {
            double2 v = make_double2(0.5 * x.x + 0.5 * y.x, 0.5 * x.y + 0.5 * y.y);
            if (a == b) v.y = 0.0;
            t.out[a + t.ld * b] = v;
            t.out[b + t.ld * a] = cconj(v);
        } else {
            t.out[a + t.ld * b] = x;
        }
    }
}

void launch_spf_tasks(const SpfTask *d_tasks, int ntasks, DofTables dt, int max_ng, hipStream_t s) {
    (void)max_ng;
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(k_spf_products, dim3(ntasks), dim3(256), 0, s, d_tasks, dt);
}

// ---------------------------------------------------------------------------
// Vector kernels (Lanczos, cmf2_integration.f90:30-103): storage T, FP64
// arithmetic and FP64 reductions.
static int64_t grid_for(int64_t n) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    return blocks;
}

template <class T>
__global__ void k_fill(T *p, int64_t n, double2 v) {
    const T vv = from_d2<T>(v);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = vv;
}

void launch_fill(double2 *p, int64_t n, double2 v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill<double2>, dim3((unsigned)grid_for(n)), dim3(256), 0, s, p, n, v);
}

template <class T>
__global__ void k_axpby(T *y, const T *x, double a, double b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 xv = to_d2(x[i]), yv = to_d2(y[i]);
        y[i] = from_d2<T>(make_double2(a * xv.x + b * yv.x, a * xv.y + b * yv.y));
    }
}

void launch_axpby(void *y, const void *x, double a, double b, int64_t n, int prec, hipStream_t s) {
    const dim3 g((unsigned)grid_for(n)), bl(256);
    if (prec == 0) hipLaunchKernelGGL(k_axpby<double2>, g, bl, 0, s, (double2 *)y, (const double2 *)x, a, b, n);
    else hipLaunchKernelGGL(k_axpby<float2>, g, bl, 0, s, (float2 *)y, (const float2 *)x, a, b, n);
}

template <class T>
__global__ void __launch_bounds__(256) k_dot_partial(const T *x, const T *y, int64_t n, double *partials,
                                                     const int *skip) {
    if (skip && *skip) return;
    double sx = 0.0, sy = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d2(x[i]), b = to_d2(y[i]);
        sx = fma(a.x, b.x, sx);
        sx = fma(a.y, b.y, sx);
        sy = fma(a.x, b.y, sy);
        sy = fma(-a.y, b.x, sy);
    }
    __shared__ double red[2][4];
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { red[0][wave] = sx; red[1][wave] = sy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ax = 0.0, ay = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) { ax += red[0][w]; ay += red[1][w]; }
        partials[2 * blockIdx.x] = ax;
        partials[2 * blockIdx.x + 1] = ay;
    }
}

// One wave sums the block partials in a fixed order (deterministic).
__global__ void __launch_bounds__(64) k_dot_final(const double *partials, int nblocks, double2 *out,
                                                  const int *skip) {
    if (skip && *skip) return;
    double ax = 0.0, ay = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += 64) { ax += partials[2 * k]; ay += partials[2 * k + 1]; }
    ax = wave_sum(ax);
    ay = wave_sum(ay);
    if (threadIdx.x == 0) *out = make_double2(ax, ay);
}

void launch_dot(const void *x, const void *y, int64_t n, double *partials, int nblocks, double2 *out, int prec,
                hipStream_t s, const int *skip) {
    if (prec == 0)
        hipLaunchKernelGGL(k_dot_partial<double2>, dim3(nblocks), dim3(256), 0, s, (const double2 *)x,
                           (const double2 *)y, n, partials, skip);
    else
        hipLaunchKernelGGL(k_dot_partial<float2>, dim3(nblocks), dim3(256), 0, s, (const float2 *)x,
                           (const float2 *)y, n, partials, skip);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(64), 0, s, partials, nblocks, out, skip);
}

template <class T>
__global__ void __launch_bounds__(256) k_norm2_partial(const T *x, int64_t n, double *partials) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d2(x[i]);
        acc = fma(a.x, a.x, acc);
        acc = fma(a.y, a.y, acc);
    }
    __shared__ double red[4];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) a += red[w];
        partials[blockIdx.x] = a;
    }
}

// |y - x|^2 partials (complex FP64), the same accumulation as k_norm2_partial
// of the stored difference: err_A of a CMF2 attempt without writing B02 - B0.
__global__ void __launch_bounds__(256) k_diff_norm2_partial(const double2 *y, const double2 *x, int64_t n,
                                                            double *partials, const double *ynorm2) {
    // ynorm2: y is the unnormalised vector of squared norm ynorm2[0]; its
    // normalised values are formed as k_scale forms them
    const double d = ynorm2 ? sqrt(ynorm2[0]) : 1.0;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 yi = y[i];
        if (ynorm2) yi = make_double2(yi.x / d, yi.y / d);
        const double2 a = make_double2(yi.x - x[i].x, yi.y - x[i].y);
        acc = fma(a.x, a.x, acc);
        acc = fma(a.y, a.y, acc);
    }
    __shared__ double red[4];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) a += red[w];
        partials[blockIdx.x] = a;
    }
}

__global__ void __launch_bounds__(64) k_norm2_final(const double *partials, int nblocks, double *out,
                                                    const int *skip = nullptr) {
    if (skip && *skip) return;
    double a = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += 64) a += partials[k];
    a = wave_sum(a);
    if (threadIdx.x == 0) *out = a;
}

void launch_norm2(const void *x, int64_t n, double *partials, int nblocks, double *out, int prec, hipStream_t s) {
    if (prec == 0)
        hipLaunchKernelGGL(k_norm2_partial<double2>, dim3(nblocks), dim3(256), 0, s, (const double2 *)x, n,
                           partials);
    else
        hipLaunchKernelGGL(k_norm2_partial<float2>, dim3(nblocks), dim3(256), 0, s, (const float2 *)x, n,
                           partials);
    hipLaunchKernelGGL(k_norm2_final, dim3(1), dim3(64), 0, s, partials, nblocks, out, (const int *)nullptr);
}

void launch_diff_norm2(const double2 *y, const double2 *x, int64_t n, double *partials, int nblocks, double *out,
                       hipStream_t s, const double *ynorm2) {
    hipLaunchKernelGGL(k_diff_norm2_partial, dim3(nblocks), dim3(256), 0, s, y, x, n, partials, ynorm2);
    hipLaunchKernelGGL(k_norm2_final, dim3(1), dim3(64), 0, s, partials, nblocks, out, (const int *)nullptr);
}

template <class T>
__global__ void k_scale(T *x, const double *norm2, int64_t n) {
    const double d = sqrt(norm2[0]);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d2(x[i]);
        x[i] = from_d2<T>(make_double2(a.x / d, a.y / d));
    }
}

void launch_scale(void *x, const double *norm2, int64_t n, int prec, hipStream_t s) {
    const dim3 g((unsigned)grid_for(n)), b(256);
    if (prec == 0) hipLaunchKernelGGL(k_scale<double2>, g, b, 0, s, (double2 *)x, norm2, n);
    else hipLaunchKernelGGL(k_scale<float2>, g, b, 0, s, (float2 *)x, norm2, n);
}

template <class T>
__global__ void k_combine(T *out, const T *basis, int64_t stride, const double2 *coef, int k, int64_t n) {
    __shared__ double2 c[kNKrylovMax];
    if (threadIdx.x < k) c[threadIdx.x] = coef[threadIdx.x];
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = zero2();
        for (int j = 0; j < k; j++) cfma(acc, c[j], to_d2(basis[j * stride + i]));
        out[i] = from_d2<T>(acc);
    }
}

void launch_combine(void *out, const void *basis, int64_t stride, const double2 *coef, int k, int64_t n, int prec,
                    hipStream_t s) {
    const dim3 g((unsigned)grid_for(n)), b(256);
    if (prec == 0)
        hipLaunchKernelGGL(k_combine<double2>, g, b, 0, s, (double2 *)out, (const double2 *)basis, stride, coef, k, n);
    else
        hipLaunchKernelGGL(k_combine<float2>, g, b, 0, s, (float2 *)out, (const float2 *)basis, stride, coef, k, n);
}

// Lanczos step on unnormalised Krylov vectors u_j = n_j v_j (n_j^2 in
// nsq[j] on the device): with h = H u_j and r = Re<u_j|h>,
//   u_{j+1} = (h - (r / n_j^2) u_j) / n_j - (n_j / n_{j-1}) u_{j-1}
// (= H v_j - alpha_j v_j - beta_{j-1} v_{j-1}, beta_{j-1} = n_j), written over
// h; accumulates |u_{j+1}|^2 partials.  No normalisation pass is needed.
template <class T>
__global__ void __launch_bounds__(256) k_lanczos_step(T *w, const T *uj, const T *uprev, const double *dot_re,
                                                      const double *nsq, int j, int64_t n, double *partials,
                                                      const int *skip) {
    if (skip && *skip) return;
    const double nj = sqrt(nsq[j]);
    const double a = 1.0 / nj, b = dot_re[0] / (nsq[j] * nj);
    const double cp = (j >= 1 && uprev) ? sqrt(nsq[j] / nsq[j - 1]) : 0.0;
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 h = to_d2(w[i]), u = to_d2(uj[i]);
        double2 r = make_double2(a * h.x - b * u.x, a * h.y - b * u.y);
        if (cp != 0.0) {
            const double2 p = to_d2(uprev[i]);
            r.x -= cp * p.x;
            r.y -= cp * p.y;
        }
        const T st = from_d2<T>(r);
        w[i] = st;
        const double2 q = to_d2(st);
        acc = fma(q.x, q.x, acc);
        acc = fma(q.y, q.y, acc);
    }
    __shared__ double red[4];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q];
        partials[blockIdx.x] = t;
    }
}

void launch_lanczos_step(void *w, const void *uj, const void *uprev, const double *dot_re, const double *nsq, int j,
                         int64_t n, double *partials, int nblocks, double *nsq_next, int prec, hipStream_t s,
                         const int *skip) {
    if (prec == 0)
        hipLaunchKernelGGL(k_lanczos_step<double2>, dim3(nblocks), dim3(256), 0, s, (double2 *)w,
                           (const double2 *)uj, (const double2 *)uprev, dot_re, nsq, j, n, partials, skip);
    else
        hipLaunchKernelGGL(k_lanczos_step<float2>, dim3(nblocks), dim3(256), 0, s, (float2 *)w, (const float2 *)uj,
                           (const float2 *)uprev, dot_re, nsq, j, n, partials, skip);
    hipLaunchKernelGGL(k_norm2_final, dim3(1), dim3(64), 0, s, partials, nblocks, nsq_next, skip);
}

// out = coef[0] v0 (FP64) + sum_{j>=1} coef[j] basis_j (storage type T, the
// j-th vector at basis + j * stride).
template <class T>
__global__ void k_combine_v0(double2 *out, const double2 *v0, const T *basis, int64_t stride, const double2 *coef,
                             int k, int64_t n, const int *kdev) {
    if (kdev) k = *kdev;
    __shared__ double2 c[kNKrylovMax];
    if (threadIdx.x < k) c[threadIdx.x] = coef[threadIdx.x];
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = cmul(c[0], v0[i]);
        for (int j = 1; j < k; j++) cfma(acc, c[j], to_d2(basis[j * stride + i]));
        out[i] = acc;
    }
}

void launch_combine_v0(double2 *out, const double2 *v0, const void *basis, int64_t stride, const double2 *coef, int k,
                       int64_t n, int prec, hipStream_t s, const int *kdev) {
    const dim3 g((unsigned)grid_for(n)), b(256);
    if (prec == 0)
        hipLaunchKernelGGL(k_combine_v0<double2>, g, b, 0, s, out, v0, (const double2 *)basis, stride, coef, k, n,
                           kdev);
    else
        hipLaunchKernelGGL(k_combine_v0<float2>, g, b, 0, s, out, v0, (const float2 *)basis, stride, coef, k, n,
                           kdev);
}

// Precision conversion of complex vectors / matrices.
template <class TI, class TO>
__global__ void k_convert(TO *dst, const TI *src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = from_d2<TO>(to_d2(src[i]));
}

void launch_convert(void *dst, int pdst, const void *src, int psrc, int64_t n, hipStream_t s) {
    const dim3 g((unsigned)grid_for(n)), b(256);
    if (psrc == 0 && pdst == 1)
        hipLaunchKernelGGL((k_convert<double2, float2>), g, b, 0, s, (float2 *)dst, (const double2 *)src, n);
    else if (psrc == 1 && pdst == 0)
        hipLaunchKernelGGL((k_convert<float2, double2>), g, b, 0, s, (double2 *)dst, (const float2 *)src, n);
    else if (psrc == 0)
        hipLaunchKernelGGL((k_convert<double2, double2>), g, b, 0, s, (double2 *)dst, (const double2 *)src, n);
    else
        hipLaunchKernelGGL((k_convert<float2, float2>), g, b, 0, s, (float2 *)dst, (const float2 *)src, n);
}

// Runge-Kutta stage combination: out = y0 + sum_q w[q] k[q] (complex
// weights, zero weights skipped), y0 / out FP64, k in the storage type T.
struct RkTerms {
    const void *k[7];
    double2 w[7];
    int nq;
};

template <class T>
__global__ void k_rk_combine(double2 *out, const double2 *y0, RkTerms t, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = y0[i];
        for (int q = 0; q < t.nq; q++) {
            if (t.w[q].x == 0.0 && t.w[q].y == 0.0) continue;
            cfma(acc, t.w[q], to_d2(static_cast<const T *>(t.k[q])[i]));
        }
        out[i] = acc;
    }
}

void launch_rk_combine(double2 *out, const double2 *y0, const void *const *k, const double2 *w, int nq, int64_t n,
                       int prec, hipStream_t s) {
    RkTerms t;
    std::memset(&t, 0, sizeof t);
    t.nq = nq;
    for (int q = 0; q < nq; q++) {
        t.k[q] = k[q];
        t.w[q] = w[q];
    }
    const dim3 g((unsigned)grid_for(n)), b(256);
    if (prec == 0) hipLaunchKernelGGL(k_rk_combine<double2>, g, b, 0, s, out, y0, t, n);
    else hipLaunchKernelGGL(k_rk_combine<float2>, g, b, 0, s, out, y0, t, n);
}

// Plain copy (profiling reference for traffic counters).
template <class T>
__global__ void __launch_bounds__(256) k_copy(T *__restrict__ dst, const T *__restrict__ src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy(void *dst, const void *src, int64_t n, int prec, hipStream_t s) {
    const dim3 g((unsigned)grid_for(n)), b(256);
    if (prec == 0) hipLaunchKernelGGL(k_copy<double2>, g, b, 0, s, (double2 *)dst, (const double2 *)src, n);
    else hipLaunchKernelGGL(k_copy<float2>, g, b, 0, s, (float2 *)dst, (const float2 *)src, n);
}

// hsum[t][e] = sum_{c in row t} coef[c] * (src[c] >= 0 ? h1[src[c]][e] : I[e])
__global__ void k_hsum(double2 *hsum, const double2 *h1, const int *rowptr, const int *src, const double2 *coef,
                       int nmat, int nmax) {
    const int nn = nmax * nmax;
    const int64_t id = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (id >= (int64_t)nmat * nn) return;
    const int t = (int)(id / nn), e = (int)(id % nn);
    double2 acc = zero2();
    for (int c = rowptr[t]; c < rowptr[t + 1]; c++) {
        double2 v;
        if (src[c] >= 0) v = h1[(int64_t)src[c] * nn + e];
        else v = make_double2((e % nmax) == (e / nmax) ? 1.0 : 0.0, 0.0);
        acc = cadd(acc, cmul(coef[c], v));
    }
    hsum[id] = acc;
}

void launch_hsum(double2 *hsum, const double2 *h1, const int *rowptr, const int *src, const double2 *coef,
                 int nmat, int nmax, hipStream_t s) {
    if (nmat <= 0) return;
    const int64_t tot = (int64_t)nmat * nmax * nmax;
    hipLaunchKernelGGL(k_hsum, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, hsum, h1, rowptr, src, coef,
                       nmat, nmax);
}

// Mean-field accumulation of a term and of its mirror term:
// dst1 += G (n1 x n2), dst2 += G^H (n2 x n1).
__global__ void k_mf_accum(double2 *dst1, double2 *dst2, const double2 *G, int n1, int n2, int ld) {
    const int e = threadIdx.x;
    if (e >= n1 * n2) return;
    const int a = e % n1, b = e / n1;
    const double2 g = G[a + ld * b];
    if (dst2 == dst1) {   // same operator (single-set mode): dst += G + G^H, n1 == n2
        dst1[a + ld * b] = cadd(dst1[a + ld * b], cadd(g, cconj(G[b + ld * a])));
        return;
    }
    dst1[a + ld * b] = cadd(dst1[a + ld * b], g);
    if (dst2) dst2[b + ld * a] = cadd(dst2[b + ld * a], cconj(g));
}

void launch_mf_accum(double2 *dst1, double2 *dst2, const double2 *G, int n1, int n2, int ld, hipStream_t s) {
    hipLaunchKernelGGL(k_mf_accum, dim3(1), dim3(256), 0, s, dst1, dst2, G, n1, n2, ld);
}

__global__ void k_mat_axpy(double2 *dst, const double2 *src, double2 alpha, int n, int ld) {
    const int e = threadIdx.x;
    if (e >= n * n) return;
    const int a = e % n, b = e / n;
    dst[a + ld * b] = cadd(dst[a + ld * b], cmul(alpha, src[a + ld * b]));
}

void launch_mat_axpy(double2 *dst, const double2 *src, double2 alpha, int n, int ld, hipStream_t s) {
    hipLaunchKernelGGL(k_mat_axpy, dim3(1), dim3(256), 0, s, dst, src, alpha, n, ld);
}

// Calc_rhom1MF (gmctdh_objects.f90:1147-1194): per job (mode, set) the
// 'poly'-regularised inverse of rho (maths.f90:8-83) times each mean field.
// job layout: [mode, rho_index, n, nops, (op, n2) x nops], stride 36.
// minw (optional): the smallest eigenvalue of each job's rho.
// rho^-1 MF per (mode, state) job, one wavefront each: the cyclic Jacobi of
// dev_herm_eig with every rotation's column, row and eigenvector updates
// spread over the lanes (lane k updates element k; each element's arithmetic
// and the sweep test are those of the one-thread routine), then the
// regularised inverse and the products with the mean-field matrices, one
// element per lane.
__global__ void __launch_bounds__(64) k_rmf(const double2 *rho, const double2 *mf, double2 *rmf, const int *jobs,
                                            int njobs, int nspfmax, int nopmax, double eps, double *minw) {
    const int jb = blockIdx.x;
    if (jb >= njobs) return;
    const int lane = threadIdx.x;
    __shared__ double2 A[kSmall * kSmall], U[kSmall * kSmall];
    __shared__ double w[kSmall];
    __shared__ int converged;
    const int *J = jobs + 36 * jb;
    const int m = J[0], ri = J[1], n = J[2], nops = J[3];
    const int nn = nspfmax * nspfmax;
    const double2 *r = rho + (int64_t)ri * nn;
    for (int e = lane; e < n * n; e += 64) {
        const int i = e % n, j = e / n;
        A[i + kSmall * j] = r[i + nspfmax * j];
        U[i + kSmall * j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    for (int sweep = 0; sweep < 100; sweep++) {
        if (lane == 0) {
            double off = 0.0, diag = 0.0;
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    const double2 a = A[i + kSmall * j];
                    const double a2 = a.x * a.x + a.y * a.y;
                    if (i != j) off += a2; else diag += a2;
                }
            converged = (off <= 1e-32 * (diag + 1e-300) || off == 0.0) ? 1 : 0;
        }
        __syncthreads();
        if (converged) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double2 apq = A[p + kSmall * q];
                const double aabs = hypot(apq.x, apq.y);
                if (aabs == 0.0) continue;
                const double2 ph = make_double2(apq.x / aabs, apq.y / aabs);
                const double app = A[p + kSmall * p].x, aqq = A[q + kSmall * q].x;
                const double theta = (aqq - app) / (2.0 * aabs);
                const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(tt * tt + 1.0), sn = tt * c;
                const double2 jpp = make_double2(c, 0.0), jqq = cscale(c, cconj(ph)), jpq = make_double2(sn, 0.0),
                              jqp = cscale(-sn, cconj(ph));
                // every lane has read the rotation's inputs before any update
                __syncthreads();
                if (lane < n) {
                    const int k = lane;
                    const double2 akp = A[k + kSmall * p], akq = A[k + kSmall * q];
                    A[k + kSmall * p] = cadd(cmul(akp, jpp), cmul(akq, jqp));
                    A[k + kSmall * q] = cadd(cmul(akp, jpq), cmul(akq, jqq));
                    const double2 ukp = U[k + kSmall * p], ukq = U[k + kSmall * q];
                    U[k + kSmall * p] = cadd(cmul(ukp, jpp), cmul(ukq, jqp));
                    U[k + kSmall * q] = cadd(cmul(ukp, jpq), cmul(ukq, jqq));
                }
                __syncthreads();
                if (lane < n) {
                    const int k = lane;
                    const double2 apk = A[p + kSmall * k], aqk = A[q + kSmall * k];
                    A[p + kSmall * k] = cadd(cmul(cconj(jpp), apk), cmul(cconj(jqp), aqk));
                    A[q + kSmall * k] = cadd(cmul(cconj(jpq), apk), cmul(cconj(jqq), aqk));
                }
                __syncthreads();
                if (lane == 0) {
                    A[p + kSmall * q] = make_double2(0.0, 0.0);
                    A[q + kSmall * p] = make_double2(0.0, 0.0);
                    A[p + kSmall * p].y = 0.0;
                    A[q + kSmall * q].y = 0.0;
                }
                __syncthreads();
            }
    }
    if (lane == 0) {
        for (int i = 0; i < n; i++) w[i] = A[i + kSmall * i].x;
        if (minw) {
            double mn = w[0];
            for (int j = 1; j < n; j++) mn = fmin(mn, w[j]);
            minw[jb] = mn;
        }
        for (int j = 0; j < n; j++) {
            double v = w[j];
            if (v < eps) v = 0.5 * (eps + v * v / eps);
            w[j] = 1.0 / v;
        }
    }
    __syncthreads();
    // rinv (Hermitian: the upper triangle, mirrored) into A
    for (int e = lane; e < n * n; e += 64) {
        const int j1 = e % n, j2 = e / n;
        if (j1 > j2) continue;
        double2 acc = zero2();
        for (int k = 0; k < n; k++) acc = cadd(acc, cscale(w[k], cmul(U[j1 + kSmall * k], cconj(U[j2 + kSmall * k]))));
        if (j1 == j2) acc.y = 0.0;
        A[j1 + kSmall * j2] = acc;
        A[j2 + kSmall * j1] = cconj(acc);
    }
    __syncthreads();
    for (int q = 0; q < nops; q++) {
        const int op = J[4 + 2 * q], n2 = J[5 + 2 * q];
        const double2 *src = mf + (int64_t)nn * (op + (int64_t)nopmax * m);
        double2 *dst = rmf + (int64_t)nn * (op + (int64_t)nopmax * m);
        for (int e = lane; e < n * n2; e += 64) {
            const int a = e % n, b = e / n;
            double2 acc = zero2();
            for (int k = 0; k < n; k++) acc = cadd(acc, cmul(A[a + kSmall * k], src[k + nspfmax * b]));
            dst[a + nspfmax * b] = acc;
        }
    }
}

void launch_rmf(const double2 *rho, const double2 *mf, double2 *rmf, const int *jobs, int njobs, int nspfmax,
                int nopmax, double eps, hipStream_t s, double *minw) {
    // nspfmax <= kSmall: the engine refuses more SPFs per mode at set-up
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_rmf, dim3(njobs), dim3(64), 0, s, rho, mf, rmf, jobs, njobs, nspfmax, nopmax, eps, minw);
}

// ---------------------------------------------------------------------------
// SPF propagation: one workgroup per logical mode.
// k: the seven Runge-Kutta stage derivatives, stage q at k + q * blk
struct SpfScratch {
    double2 *y0, *k, *tmp, *perr, *sect, *sect2, *ovl, *od, *prod;
    int blk;
    __device__ double2 *stage(int q) const { return k + (size_t)q * blk; }
};

__device__ SpfScratch carve(double2 *base, const SpfModeDesc &d, int nspfmax) {
    SpfScratch s;
    double2 *p = base + d.scr_off;
    s.y0 = p; p += d.blk;
    s.k = p; p += (size_t)7 * d.blk;
    s.blk = d.blk;
    s.tmp = p; p += d.blk;
    s.perr = p; p += d.blk;
    s.sect = p; p += d.nG * nspfmax;
    s.sect2 = p; p += d.nG * nspfmax;
    s.ovl = p; p += (size_t)kMaxSets * nspfmax * nspfmax;     // one n x n block per set
    s.od = p; p += (size_t)kMaxSets * nspfmax * nspfmax;      // kMaxSets blocks as well
    s.prod = p;                                               // d.nuq blocks of nG x nspfmax
    return s;
}

// The projector's and the error norm's small matrices (read serially by one
// lane per set) in LDS rather than in the global scratch pool; the launch
// provides spf_lds_bytes(nspfmax) of dynamic LDS.
__device__ __forceinline__ void spf_lds_overlaps(SpfScratch &S, int nspfmax) {
    extern __shared__ double2 spf_lds[];
    S.ovl = spf_lds;
    S.od = spf_lds + (size_t)kMaxSets * nspfmax * nspfmax;
}
static size_t spf_lds_bytes(int nspfmax) { return sizeof(double2) * 2 * kMaxSets * (size_t)nspfmax * nspfmax; }
// 16 SPFs and 8 sets take 64 KiB plus the kernels' static LDS: allow more
// than the default dynamic limit, once per kernel (thread-safe statics)
template <class K> static void spf_allow_lds(K kernel) {
    (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
}

// Calc_Der_SPF, DVR branch, on mode-local blocks (cmf2_derivatives.f90:118-173).
__device__ void dev_der_spf(const SpfModeDesc &d, const SpfParams &P, const DofTables &dt, const double2 *in,
                            double2 *out, SpfScratch &S, int *fail) {
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nG = d.nG, nmax = P.nspfmax;
    const int *loff = d.loc_off;
    // the primitive products h phi of the distinct (codes, rhs set) pairs,
    // each formed once (several operators share one: the same one-mode
    // operator in several mean-field terms).  One-DOF modes with powers of x
    // or DVR matrices form them all in one pass, element by element, with
    // apply_prim_block's arithmetic; other modes apply the DOFs in turn.
    if (d.prod_direct) {
        const int dof = dt.mode_dofs[d.mode * kMaxDofPerMode];
        const double *x = dt.grid + (int64_t)dt.ngpmax * dof;
        for (int i = tid; i < d.prod_off[d.nprod]; i += nth) {
            int k = 0;
            while (k + 1 < d.nprod && i >= d.prod_off[k + 1]) k++;
            const int u = d.prod_uq[k], e = i - d.prod_off[k];
            const int s2 = d.uq_set[u], code = d.uq_codes[u][0];
            const double2 *c0 = in + loff[s2];
            const int g2 = e % nG, j = e / nG;
            double2 v;
            if (code >= 1) {
                double f = x[g2];
                for (int q = 1; q < code; q++) f *= x[g2];
                v = cscale(f, c0[e]);
            } else {
                const double *Mat = (code == -1 ? dt.d1 : (code == -2 ? dt.d2 : dt.xd1)) +
                                    (int64_t)dt.ngpmax * dt.ngpmax * dof;
                const double2 *c = c0 + (int64_t)j * nG;
                double2 acc = zero2();
                for (int g = 0; g < nG; g++) {
                    const double m = Mat[g + (int64_t)dt.ngpmax * g2];
                    acc.x = fma(m, c[g].x, acc.x);
                    acc.y = fma(m, c[g].y, acc.y);
                }
                v = (code == -1 || code == -3) ? make_double2(-acc.y, acc.x) : acc;
            }
            S.prod[(size_t)u * nG * nmax + e] = v;
        }
    } else {
        for (int u = 0; u < d.nuq; u++) {
            if (d.uq_ident[u]) continue;
            const int s2 = d.uq_set[u], n2 = d.nspf[s2];
            double2 *Pu = S.prod + (size_t)u * nG * nmax;
            for (int e = tid; e < nG * n2; e += nth) Pu[e] = in[loff[s2] + e];
            __syncthreads();
            for (int k = 0; k < dt.mode_ndof[d.mode]; k++)
                if (d.uq_codes[u][k] != 0) apply_prim_block(dt, d.mode, k, d.uq_codes[u][k], n2, nG, Pu, S.sect2);
        }
    }
    __syncthreads();
    // out(g, j1) of set s1 = sum over the operators q into s1, in operator
    // order, of der_factor * (rmf_q h_q phi_s2)(g, j1)
    for (int e = tid; e < d.blk; e += nth) {
        int s1 = 0;
        while (s1 + 1 < d.nsets && e >= loff[s1 + 1]) s1++;
        const int loc = e - loff[s1], g = loc % nG, j1 = loc / nG;
        double2 o = zero2();
        for (int q = 0; q < d.nop; q++) {
            if (d.op_lhs_set[q] != s1) continue;
            const int s2 = d.op_rhs_set[q], n2 = d.nspf[s2], u = d.op_uq[q];
            const double2 *src = d.uq_ident[u] ? in + loff[s2] : S.prod + (size_t)u * nG * nmax;
            const double2 *r = P.rmf + (int64_t)nmax * nmax * d.op_index[q];
            double2 acc = zero2();
            for (int j2 = 0; j2 < n2; j2++) cfma(acc, r[j1 + nmax * j2], src[g + nG * j2]);
            o = cadd(o, cmul(P.der_factor, acc));
        }
        out[e] = o;
    }
    __syncthreads();
    // projector: d <- d - phi (phi^H phi)^-1 phi^H d, every set at once:
    // the sets' overlaps in parallel, each set's Cholesky solve on lane 0 of
    // its own wave, then the updates (each set's arithmetic is unchanged)
    const int nn = nmax * nmax;
    for (int s = 0; s < d.nsets; s++) {
        const int n = d.nspf[s];
        const double2 *phi = in + loff[s];
        const double2 *dd = out + loff[s];
        double2 *ovl = S.ovl + nn * s, *od = S.od + nn * s;
        for (int e = tid; e < n * n; e += nth) {
            const int a = e % n, b = e / n;
            double2 x = zero2(), y = zero2();
            for (int g = 0; g < nG; g++) {
                x = cadd(x, cconjmul(phi[g + nG * a], phi[g + nG * b]));
                y = cadd(y, cconjmul(phi[g + nG * a], dd[g + nG * b]));
            }
            ovl[a + n * b] = x;
            od[a + n * b] = y;
        }
    }
    __syncthreads();
    {
        const int lane = tid & 63, wave = tid >> 6, nw = (nth + 63) >> 6;
        for (int s = wave; s < d.nsets; s += nw) {
            if (lane != 0) continue;
            const int n = d.nspf[s];
            // Cholesky S = U^H U (upper), then solve for each column of od
            double2 *U = S.ovl + nn * s, *od = S.od + nn * s;
            for (int j = 0; j < n; j++) {
                double dg = U[j + n * j].x;
                for (int k = 0; k < j; k++) dg -= U[k + n * j].x * U[k + n * j].x + U[k + n * j].y * U[k + n * j].y;
                if (!(dg > 0.0)) { *fail = 1; dg = 1.0; }
                const double ujj = sqrt(dg);
                U[j + n * j] = make_double2(ujj, 0.0);
                for (int i = j + 1; i < n; i++) {
                    double2 v = U[j + n * i];
                    for (int k = 0; k < j; k++) v = csub(v, cconjmul(U[k + n * j], U[k + n * i]));
                    U[j + n * i] = make_double2(v.x / ujj, v.y / ujj);
                }
            }
            for (int c = 0; c < n; c++) {
                double2 *y = od + n * c;
                for (int i = 0; i < n; i++) {
                    double2 v = y[i];
                    for (int k = 0; k < i; k++) v = csub(v, cconjmul(U[k + n * i], y[k]));
                    const double u = U[i + n * i].x;
                    y[i] = make_double2(v.x / u, v.y / u);
                }
                for (int i = n - 1; i >= 0; i--) {
                    double2 v = y[i];
                    for (int k = i + 1; k < n; k++) v = csub(v, cmul(U[i + n * k], y[k]));
                    const double u = U[i + n * i].x;
                    y[i] = make_double2(v.x / u, v.y / u);
                }
            }
        }
    }
    __syncthreads();
    for (int s = 0; s < d.nsets; s++) {
        const int n = d.nspf[s];
        const double2 *phi = in + loff[s];
        double2 *dd = out + loff[s];
        const double2 *od = S.od + nn * s;
        for (int e = tid; e < nG * n; e += nth) {
            const int g = e % nG, b = e / nG;
            double2 acc = zero2();
            for (int k = 0; k < n; k++) cfma(acc, od[k + n * b], phi[g + nG * k]);
            dd[e] = csub(dd[e], acc);
        }
    }
    __syncthreads();
}

// Calc_Err_SPF, DVR branch (cmf2_derivatives.f90:253-291).
__device__ double dev_err_spf(const SpfModeDesc &d, const SpfParams &P, const double2 *p1, const double2 *p2,
                              SpfScratch &S) {
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nG = d.nG, nmax = P.nspfmax;
    __shared__ double err_sh;
    if (tid == 0) err_sh = 0.0;
    int o = 0;
    for (int s = 0; s < d.nsets; s++) {
        const int n = d.nspf[s];
        for (int e = tid; e < n * n; e += nth) {
            const int a = e % n, b = e / n;
            double2 x = zero2();
            if (a <= b)
                for (int g = 0; g < nG; g++) {
                    double2 da = csub(p1[o + g + nG * a], p2[o + g + nG * a]);
                    double2 db = csub(p1[o + g + nG * b], p2[o + g + nG * b]);
                    x = cadd(x, cconjmul(da, db));
                }
            S.ovl[a + n * b] = x;
        }
        __syncthreads();
        if (tid == 0) {
            const double2 *r = P.rho + (int64_t)nmax * nmax * d.rho_index[s];
            double err = err_sh;
            for (int j2 = 0; j2 < n; j2++) {
                for (int j1 = 0; j1 < j2; j1++) {
                    double2 prod = cmul(r[j1 + nmax * j2], cconj(S.ovl[j1 + n * j2]));
                    err += 2.0 * prod.x;
                }
                double2 prod = cmul(r[j2 + nmax * j2], make_double2(S.ovl[j2 + n * j2].x, 0.0));
                err += prod.x;
            }
            err_sh = err;
        }
        __syncthreads();
        o += nG * n;
    }
    const double e = sqrt(fabs(err_sh));
    __syncthreads();
    return e;
}

__constant__ double c_ckB[6][5] = {
    {0, 0, 0, 0, 0}, {0.2, 0, 0, 0, 0}, {0.075, 0.225, 0, 0, 0}, {0.3, -0.9, 1.2, 0, 0},
    {-11.0 / 54.0, 2.5, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
__constant__ double c_ckC[6] = {37.0 / 378.0, 0, 250.0 / 621.0, 125.0 / 594.0, 0, 512.0 / 1771.0};
__constant__ double c_ckD[6] = {2825.0 / 27648.0, 0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 0.25};
__constant__ double c_dpB[6][5] = {
    {0, 0, 0, 0, 0}, {0.2, 0, 0, 0, 0}, {0.075, 0.225, 0, 0, 0}, {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0}};
__constant__ double c_dpC[6] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0};
__constant__ double c_dpD[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0,
                                187.0 / 2100.0, 0.025};

// propagate_SPF over P.dt (cmf2_base.f90:210-260) with rk_SPF_ck / rk_SPF_dp
// (cmf2_integration.f90:107-412).  status: 1 = Cholesky failure, 2 = too many
// attempts.
__global__ void __launch_bounds__(256) k_spf_propagate(const SpfModeDesc *__restrict__ modes, double2 *psi,
                                                       double2 *scratch, SpfParams P, DofTables dt, int *status) {
    const SpfModeDesc &d = modes[blockIdx.x];
    const int tid = threadIdx.x, nth = blockDim.x;
    SpfScratch S = carve(scratch, d, P.nspfmax);
    spf_lds_overlaps(S, P.nspfmax);
    __shared__ int fail;
    if (tid == 0) fail = 0;
    // gather the mode's set blocks
    {
        int o = 0;
        for (int s = 0; s < d.nsets; s++) {
            const int n = d.nG * d.nspf[s];
            for (int e = tid; e < n; e += nth) S.y0[o + e] = psi[d.off[s] + e];
            o += n;
        }
    }
    __syncthreads();
    const double safety = 0.9, pgrow = -0.2, pshrnk = -0.25, minscal = 0.1, maxscal = 5.0;
    const double errcon = pow(5.0 / safety, 1.0 / pgrow);
    const bool dp = P.dormand_prince != 0;
    const double DT = P.dt_src ? P.dt_src[0] * P.dt_scale : P.dt;
    double dt_done = 0.0, stepsize = DT;
    int attempts = 0;
    for (;;) {
        // ---- one adaptive RK step (rk_SPF_ck / rk_SPF_dp) ----
        double h = stepsize, error = 0.0;
        for (;;) {
            dev_der_spf(d, P, dt, S.y0, S.stage(0), S, &fail);
            for (int st = 1; st < 6; st++) {
                for (int e = tid; e < d.blk; e += nth) {
                    double2 v;
                    if (st == 1) {
                        const double c = (dp ? c_dpB[1][0] : c_ckB[1][0]) * h;
                        v = cadd(S.y0[e], cscale(c, S.stage(0)[e]));
                    } else {
                        double2 acc = zero2();
                        for (int q = 0; q < st; q++) {
                            const double b = dp ? c_dpB[st][q] : c_ckB[st][q];
                            acc = cadd(acc, cscale(b, S.stage(q)[e]));
                        }
                        v = cadd(S.y0[e], cscale(h, acc));
                    }
                    S.tmp[e] = v;
                }
                __syncthreads();
                dev_der_spf(d, P, dt, S.tmp, S.stage(st), S, &fail);
            }
            for (int e = tid; e < d.blk; e += nth) {
                double2 acc = zero2();
                for (int q = 0; q < 6; q++) {
                    const double c = dp ? c_dpC[q] : c_ckC[q];
                    if (c != 0.0) acc = cadd(acc, cscale(c, S.stage(q)[e]));
                }
                S.tmp[e] = cadd(S.y0[e], cscale(h, acc));
                if (!dp) {
                    double2 ae = zero2();
                    for (int q = 0; q < 6; q++)
                        if (c_ckD[q] != 0.0) ae = cadd(ae, cscale(c_ckD[q], S.stage(q)[e]));
                    S.perr[e] = cadd(S.y0[e], cscale(h, ae));
                }
            }
            __syncthreads();
            if (dp) {
                dev_der_spf(d, P, dt, S.tmp, S.stage(6), S, &fail);
                for (int e = tid; e < d.blk; e += nth) {
                    double2 ae = zero2();
                    for (int q = 0; q < 7; q++)
                        if (c_dpD[q] != 0.0) ae = cadd(ae, cscale(c_dpD[q], S.stage(q)[e]));
                    S.perr[e] = cadd(S.y0[e], cscale(h, ae));
                }
                __syncthreads();
            }
            error = dev_err_spf(d, P, S.tmp, S.perr, S) / P.eps_psi;
            attempts++;
            if (error < 1.0) break;
            if (fail || attempts > P.max_attempts || !(error == error)) break;
            const double a = minscal * h, b = safety * h * pow(error, pshrnk);
            h = a > b ? a : b;
        }
        if (fail || attempts > P.max_attempts || !(error == error)) break;
        double hnext = (error > errcon) ? safety * h * pow(error, pgrow) : h * maxscal;
        for (int e = tid; e < d.blk; e += nth) S.y0[e] = S.tmp[e];
        __syncthreads();
        // ---- step landing (propagate_SPF) ----
        dt_done = dt_done + h;
        if (dt_done >= DT) break;
        if (hnext < P.min_stepsize * 2) hnext = P.min_stepsize * 2;
        stepsize = (DT - dt_done) / (double)((long long)((DT - dt_done) / hnext) + 1);
        if (stepsize > DT - dt_done) stepsize = DT - dt_done;
    }
    __syncthreads();
    if (tid == 0 && (fail || attempts > P.max_attempts)) atomicMax(status, fail ? 1 : 2);
    {
        int o = 0;
        for (int s = 0; s < d.nsets; s++) {
            const int n = d.nG * d.nspf[s];
            for (int e = tid; e < n; e += nth) psi[d.off[s] + e] = S.y0[o + e];
            o += n;
        }
    }
}

// Calc_Der_SPF for one mode (mode >= 0) or for every mode (mode < 0, one
// workgroup each) on a full psi vector; dpsi receives those modes' blocks.
__global__ void __launch_bounds__(256) k_der_spf(const SpfModeDesc *__restrict__ modes, int mode, const double2 *psi,
                                                 double2 *dpsi, double2 *scratch, SpfParams P, DofTables dt,
                                                 int *status) {
    const SpfModeDesc &d = modes[mode >= 0 ? mode : blockIdx.x];
    const int tid = threadIdx.x, nth = blockDim.x;
    SpfScratch S = carve(scratch, d, P.nspfmax);
    spf_lds_overlaps(S, P.nspfmax);
    __shared__ int fail;
    if (tid == 0) fail = 0;
    int o = 0;
    for (int s = 0; s < d.nsets; s++) {
        const int n = d.nG * d.nspf[s];
        for (int e = tid; e < n; e += nth) S.y0[o + e] = psi[d.off[s] + e];
        o += n;
    }
    __syncthreads();
    dev_der_spf(d, P, dt, S.y0, S.stage(0), S, &fail);
    o = 0;
    for (int s = 0; s < d.nsets; s++) {
        const int n = d.nG * d.nspf[s];
        for (int e = tid; e < n; e += nth) dpsi[d.off[s] + e] = S.stage(0)[o + e];
        o += n;
    }
    if (tid == 0 && fail) atomicMax(status, 1);
}

void launch_der_spf(const SpfModeDesc *d_modes, int mode, const double2 *psi, double2 *dpsi, double2 *scratch,
                    SpfParams p, DofTables dt, int *d_status, hipStream_t s) {
    static const bool lds_set = (spf_allow_lds(k_der_spf), true);
    (void)lds_set;
    hipLaunchKernelGGL(k_der_spf, dim3(mode >= 0 ? 1 : -mode), dim3(256), spf_lds_bytes(p.nspfmax), s, d_modes, mode,
                       psi, dpsi, scratch, p,
                       dt, d_status);
}

void launch_spf_propagate(const SpfModeDesc *d_modes, int nmodes, double2 *psi, double2 *scratch, SpfParams p,
                          DofTables dt, int *d_status, hipStream_t s) {
    static const bool lds_set = (spf_allow_lds(k_spf_propagate), true);
    (void)lds_set;
    hipLaunchKernelGGL(k_spf_propagate, dim3(nmodes), dim3(256), spf_lds_bytes(p.nspfmax), s, d_modes, psi, scratch,
                       p, dt, d_status);
}


// ---------------------------------------------------------------------------
// Device-controlled CMF2 attempt: the Lanczos stopping test, the propagator
// coefficients and the accept/reject decision on the device (one thread
// each), so that a whole attempt is one HIP graph.
__global__ void k_lan_begin(LanczosDev *L) {
    L->done = 0;
    L->nk = 0;
    L->overflow = 0;
    L->error = 0.0;
}

void launch_lan_begin(LanczosDev *L, hipStream_t s) { hipLaunchKernelGGL(k_lan_begin, dim3(1), dim3(1), 0, s, L); }

// After the H-image of basis vector j - 1 (cmf2_integration.f90:56-84):
// alpha_{j-1} = <u|Hu> / n_{j-1}^2, beta_{j-1} = n_j (the vectors are kept
// unnormalised, u_j = n_j v_j), error *= beta_{j-1} / j |dt| from j = 2 on;
// the build stops at the first j >= 2 with error < eps_B = 1e-8.  kmax is
// the last iteration enqueued: reaching it unconverged sets overflow (1 if
// it is the reference's own limit kNKrylovMax - 1, else 2).
__global__ void k_lan_ctrl(LanczosDev *L, const AttemptDev *at, const double2 *dot, const double *nsq, int j,
                           double dtscale, int kmax) {
    if (L->done) return;
    const double dt = fabs(at->h * dtscale);
    L->nrm[j - 1] = nsq[j - 1];
    L->nrm[j] = nsq[j];
    L->alpha[j - 1] = dot->x / nsq[j - 1];
    L->beta[j - 1] = sqrt(nsq[j]);
    if (j == 1) {
        L->error = L->beta[0] * dt;
        return;
    }
    L->error = L->error * L->beta[j - 1] / (double)j * dt;
    if (L->error < 1e-8) {
        L->done = 1;
        L->nk = j;
    } else if (j >= kmax) {
        L->done = 1;
        L->nk = j;
        L->overflow = (kmax >= kNKrylovMax - 1) ? 1 : 2;
    }
}

void launch_lan_ctrl(LanczosDev *L, const AttemptDev *at, const double2 *dot, const double *nsq, int j,
                     double dtscale, int kmax, hipStream_t s) {
    hipLaunchKernelGGL(k_lan_ctrl, dim3(1), dim3(1), 0, s, L, at, dot, nsq, j, dtscale, kmax);
}

// Eigen-decomposition of the Krylov tridiagonal (cyclic Jacobi, the host
// routine's rotations) when build is set, then the coefficients of
// B(t + dt) = sum_i coef_i u_i: coef_i = sum_r U_ir U_0r exp(-i w_r dt) / n_i
// (cmf2_integration.f90:86-97).
__global__ void k_lan_coef(LanczosDev *L, const AttemptDev *at, double dtscale, int build, double2 *coef) {
    const int n = L->nk;
    if (build) {
        // one thread: the tridiagonal in LDS rather than in private memory
        __shared__ double A[kNKrylovMax * kNKrylovMax];
        double *U = L->U;
        for (int i = 0; i < n * n; i++) A[i] = 0.0;
        for (int i = 0; i < n; i++) {
            A[i + n * i] = L->alpha[i];
            if (i + 1 < n) { A[i + n * (i + 1)] = L->beta[i]; A[(i + 1) + n * i] = L->beta[i]; }
        }
        for (int i = 0; i < n; i++)
            for (int jj = 0; jj < n; jj++) U[i + n * jj] = (i == jj) ? 1.0 : 0.0;
        for (int sweep = 0; sweep < 100; sweep++) {
            double off = 0.0, dg = 0.0;
            for (int jj = 0; jj < n; jj++)
                for (int i = 0; i < n; i++) {
                    const double v = A[i + n * jj] * A[i + n * jj];
                    if (i != jj) off += v; else dg += v;
                }
            if (off <= 1e-32 * (dg + 1e-300) || off == 0.0) break;
            for (int p = 0; p < n - 1; p++)
                for (int q = p + 1; q < n; q++) {
                    const double apq = A[p + n * q];
                    if (apq == 0.0) continue;
                    const double theta = (A[q + n * q] - A[p + n * p]) / (2.0 * apq);
                    const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    const double cc = 1.0 / sqrt(t * t + 1.0), ss = t * cc;
                    for (int k = 0; k < n; k++) {
                        const double akp = A[k + n * p], akq = A[k + n * q];
                        A[k + n * p] = cc * akp - ss * akq;
                        A[k + n * q] = ss * akp + cc * akq;
                    }
                    for (int k = 0; k < n; k++) {
                        const double apk = A[p + n * k], aqk = A[q + n * k];
                        A[p + n * k] = cc * apk - ss * aqk;
                        A[q + n * k] = ss * apk + cc * aqk;
                    }
                    A[p + n * q] = A[q + n * p] = 0.0;
                    for (int k = 0; k < n; k++) {
                        const double ukp = U[k + n * p], ukq = U[k + n * q];
                        U[k + n * p] = cc * ukp - ss * ukq;
                        U[k + n * q] = ss * ukp + cc * ukq;
                    }
                }
        }
        for (int i = 0; i < n; i++) {
            L->w[i] = A[i + n * i];
            L->ln[i] = sqrt(L->nrm[i]);
        }
    }
    const double dt = at->h * dtscale;
    for (int i = 0; i < n; i++) {
        double re = 0.0, im = 0.0;
        for (int r = 0; r < n; r++) {
            double sn, cs;
            sincos(-L->w[r] * dt, &sn, &cs);
            const double u = L->U[i + n * r] * L->U[0 + n * r];
            re += u * cs;
            im += u * sn;
        }
        coef[i] = make_double2(re / L->ln[i], im / L->ln[i]);
    }
}

void launch_lan_coef(LanczosDev *L, const AttemptDev *at, double dtscale, int build, double2 *coef, hipStream_t s) {
    hipLaunchKernelGGL(k_lan_coef, dim3(1), dim3(1), 0, s, L, at, dtscale, build, coef);
}

// DeltaPsi (gmctdh_objects.f90:204-227) from the three overlaps.
// o[k] = sum over states of the trace of DeltaPsi's overlap Gram (set k,
// state s), in a fixed order
__global__ void k_dpsi_traces(const double2 *G, int ns, int ld, double2 *o) {
    const int k = threadIdx.x;
    if (k >= 3) return;
    double2 acc = make_double2(0.0, 0.0);
    for (int s = 0; s < ns; s++) {
        const double2 *g = G + (size_t)ld * ld * ((size_t)k * ns + s);
        for (int i = 0; i < ld; i++) acc = make_double2(acc.x + g[i + ld * i].x, acc.y + g[i + ld * i].y);
    }
    o[k] = acc;
}

void launch_dpsi_traces(const double2 *G, int ns, int ld, double2 *o, hipStream_t s) {
    hipLaunchKernelGGL(k_dpsi_traces, dim3(1), dim3(64), 0, s, G, ns, ld, o);
}

__global__ void k_attempt_psi(AttemptDev *at, const double2 *o11, const double2 *o22, const double2 *o12) {
    at->err_psi = sqrt(fabs(2.0 * (1.0 - o12->x / sqrt(o11->x * o22->x))));
}

void launch_attempt_psi(AttemptDev *at, const double2 *o11, const double2 *o22, const double2 *o12, hipStream_t s) {
    hipLaunchKernelGGL(k_attempt_psi, dim3(1), dim3(1), 0, s, at, o11, o22, o12);
}

// err_A = |B02 - B0| / 2, error = max(err_A, err_psi) / eps_int, rejected
// when error > 1 (cmf2_base.f90:177-184) or as replayed; an unfinished
// Krylov build always rejects (the host then reports or repeats it).
__global__ void k_attempt_decide(AttemptDev *at, const LanczosDev *L1, const LanczosDev *L2, const double *n2,
                                 double eps_int, const int *spf_status) {
    at->err_A = sqrt(n2[0]) * 0.5;
    at->error = fmax(at->err_A, at->err_psi) / eps_int;
    bool reject = (at->forced >= 0) ? (at->forced == 0) : (at->error > 1.0);
    at->overflow = max(L1->overflow, L2->overflow);
    at->spf_status = spf_status[0];
    if (at->overflow || at->spf_status) reject = true;
    at->accept = reject ? 0 : 1;
    at->nk1 = L1->nk;
    at->nk2 = L2->nk;
}

void launch_attempt_decide(AttemptDev *at, const LanczosDev *L1, const LanczosDev *L2, const double *n2,
                           double eps_int, const int *spf_status, hipStream_t s) {
    hipLaunchKernelGGL(k_attempt_decide, dim3(1), dim3(1), 0, s, at, L1, L2, n2, eps_int, spf_status);
}

// srcnorm2: src is unnormalised with squared norm srcnorm2[0]; dst receives
// it normalised as k_scale would
__global__ void k_copy_if(double2 *dst, const double2 *src, int64_t n, const int *flag, const double *srcnorm2) {
    if (!*flag) return;
    const double d = srcnorm2 ? sqrt(srcnorm2[0]) : 1.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = src[i];
        dst[i] = srcnorm2 ? make_double2(a.x / d, a.y / d) : a;
    }
}

void launch_copy_if(double2 *dst, const double2 *src, int64_t n, const int *flag, hipStream_t s,
                    const double *srcnorm2) {
    hipLaunchKernelGGL(k_copy_if, dim3((unsigned)grid_for(n)), dim3(256), 0, s, dst, src, n, flag, srcnorm2);
}

}  // namespace mg
