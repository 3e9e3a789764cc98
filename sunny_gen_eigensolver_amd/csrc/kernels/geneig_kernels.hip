// Batched generalized Hermitian-definite eigensolve H u = lambda S u for
// gfx950 (CDNA4), one matrix order N per compiled module.
//
// This file is compiled at run time by hipRTC, once per (N, workgroup size),
// with -DGENEIG_N=<N> -DGENEIG_NT=<threads>; the library embeds its text.
// hipcc also compiles it ahead of time as a build check.
//
// Pipeline per matrix (the reference's fused kernel order,
// src/unified/kernel_source.hpp:238-276):
//   potrf(S) = L L^H  ->  M = L^-1 H L^-H  ->  Householder tridiagonalisation
//   Q^H M Q = T  ->  implicit QL on T with rotations accumulated into Z
//   ->  sort ascending  ->  V = Q Z  ->  U = L^-H V.
//
// Layout: column-major, element (i,j) of matrix b at b*N*N + j*N + i, complex
// numbers as (re, im) real pairs. In LDS the two N x N complex work
// matrices use an odd leading dimension so that a wave reading one matrix
// row (lane = column) touches 64 distinct 16-byte bank groups.
//
// Thread roles: a workgroup is NW = GENEIG_NT / 64 waves. Phases whose work is
// independent per column (or per row) give each wave a fixed set of columns
// held in registers (lane = row): the triangular solves run without any
// workgroup barrier, the pivot entry travelling by v_readlane. Phases that mix
// columns (Cholesky, tridiagonalisation, reflector application) synchronise
// with workgroup barriers. The QL iteration is a scalar recurrence: one wave
// runs it, every lane redundantly, and each lane applies the rotations to its
// own row of Z.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#ifndef GENEIG_N
#define GENEIG_N 54
#endif
#ifndef GENEIG_NT
#define GENEIG_NT 256
#endif

namespace gk {

constexpr int N = GENEIG_N;
constexpr int NT = GENEIG_NT;
constexpr int NW = NT / 64;
constexpr int LDA = N | 1;                 // complex leading dimension in LDS
constexpr int LDZ = N | 1;                 // real leading dimension of Z in LDS
constexpr int CPW = (N + NW - 1) / NW;     // columns (or rows) owned per wave
constexpr int QL_MAXIT = 30;               // sweeps per eigenvalue before giving up

static_assert(NT % 64 == 0 && NW >= 1, "workgroup must be whole waves");
static_assert(N >= 1 && N <= 192, "orders above 64 run the staged kernels, up to 192");

// Module precision: FP64 by default; -DGENEIG_FP32 builds the FP32 variant
// of the split-mode stages used by the mixed-precision path.
#ifdef GENEIG_FP32
typedef float real;
#else
typedef double real;
#endif

struct __attribute__((aligned(2 * sizeof(real)))) cx {
    real x, y;
};

__device__ __forceinline__ float sqrt_r(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrt_r(double x) { return __builtin_sqrt(x); }

#ifdef GENEIG_FP32
constexpr real REAL_HALF_ULP = 0x1p-24f;   // unit roundoff
constexpr real REAL_TINY = 0x1p-126f;      // smallest normal
#else
constexpr real REAL_HALF_ULP = 0x1p-53;
constexpr real REAL_TINY = 0x1p-1022;
#endif

__device__ __forceinline__ cx cm(real a, real b) { cx r; r.x = a; r.y = b; return r; }
__device__ __forceinline__ cx operator+(cx a, cx b) { return cm(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cx operator-(cx a, cx b) { return cm(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cx operator*(cx a, real s) { return cm(a.x * s, a.y * s); }
__device__ __forceinline__ cx cmul(cx a, cx b) {
    return cm(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ cx cmulc(cx a, cx b) {
    return cm(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a) * b
__device__ __forceinline__ cx cconjmul(cx a, cx b) {
    return cm(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ cx cconj(cx a) { return cm(a.x, -a.y); }
// acc + a * b and m - a * conj(b) as four FMAs each.
__device__ __forceinline__ cx cfma(cx a, cx b, cx acc) {
    return cm(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}
__device__ __forceinline__ cx cfms_conj(cx a, cx b, cx m) {
    return cm(fma(-a.x, b.x, fma(-a.y, b.y, m.x)), fma(-a.y, b.x, fma(a.x, b.y, m.y)));
}

// Complex numbers as stored in global memory by the mixed-precision path:
// complex128 inputs and outputs, complex64 hand-offs between its stages.
struct __attribute__((aligned(16))) cd {
    double x, y;
};
struct __attribute__((aligned(8))) cf {
    float x, y;
};
__device__ __forceinline__ cx to_cx(const cd& a) { return cm((real)a.x, (real)a.y); }
__device__ __forceinline__ cx to_cx(const cf& a) { return cm((real)a.x, (real)a.y); }
__device__ __forceinline__ cx to_cx(const cx& a) { return a; }
__device__ __forceinline__ real cabs2(cx a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// Value of `v` in lane `src` (src wave-uniform), via v_readlane.
__device__ __forceinline__ double bcast(double v, int src) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bcast(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ cx bcast(cx v, int src) { return cm(bcast(v.x, src), bcast(v.y, src)); }

// DPP move of a value between lanes (each 32-bit half with the same control).
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Sum over the 64 lanes, returned in every lane: a butterfly inside each
// 16-lane row with DPP (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror), then the four row sums via v_readlane.
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
    v += dpp_mov<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);   // row_half_mirror
    v += dpp_mov<0x140>(v);   // row_mirror
    return (bcast(v, 0) + bcast(v, 16)) + (bcast(v, 32) + bcast(v, 48));
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
    v = fmax(v, dpp_mov<0xB1>(v));
    v = fmax(v, dpp_mov<0x4E>(v));
    v = fmax(v, dpp_mov<0x141>(v));
    v = fmax(v, dpp_mov<0x140>(v));
    return fmax(fmax(bcast(v, 0), bcast(v, 16)), fmax(bcast(v, 32), bcast(v, 48)));
}

// Lane `target` takes x, the others keep v.
__device__ __forceinline__ real sel_lane(real v, real x, int target) {
    return (lane_id() == target) ? x : v;
}

// 1/sqrt(t) for the QL rotations.
#ifdef GENEIG_FP32
__device__ __forceinline__ float rsqrt_nr(float t) {
    // v_rsq_f32 is accurate to about one ulp: no refinement needed in FP32.
    return __builtin_amdgcn_rsqf(t);
}
#else
__device__ __forceinline__ double rsqrt_nr(double t) {
    // v_rsq_f64 is good to about 2^-24 (measured on gfx950), so one
    // third-order step y (1 + e/2 + 3 e^2 / 8), e = 1 - t y^2 (the binomial
    // series of (1 - e)^(-1/2)), leaves an error of order e^3 (below 2^-70)
    // plus the final rounding: four dependent operations instead of the six
    // of two Newton steps.
    const double y = __builtin_amdgcn_rsq(t);
    const double e = fma(-(t * y), y, 1.0);
    const double q = fma(e, 0.375, 0.5);
    return fma(y * e, q, y);
}
#endif

// 1/x for the QL shift and the refinement's gap divisions: v_rcp_f32 (about
// one ulp); in FP64 v_rcp_f64 and one Newton step (relative error of order
// 2^-48, far below what either use needs).
#ifdef GENEIG_FP32
__device__ __forceinline__ float rcp_nr(float x) { return __builtin_amdgcn_rcpf(x); }
#else
__device__ __forceinline__ double rcp_nr(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, y, 1.0), y, y);
}
#endif

// One rotation of an implicit QL sweep (Handbook II/3, tql2's inner loop)
// on the pair (d_i, e_i) with the sweep's running (s, c, g, p) and d_up (the
// original d[i+1]); returns the new e[i+1] and d[i+1] through e_next and
// d_next. 2 c b is formed as (2 g b) y, so that only s, c and that product
// wait for the reciprocal square root. t = f^2 + g^2 > 0 inside the window
// (f = s e_i != 0 there); the smallest normal added to g^2 keeps t, and so
// the rotation, finite for inputs scaled near underflow.
__device__ __forceinline__ void ql_rotation(real e_i, real d_i, real& s, real& c, real& g, real& p, real& d_up,
                                            real& e_next, real& d_next) {
    const real f = s * e_i;
    const real b = c * e_i;
    const real t = fma(f, f, fma(g, g, REAL_TINY));
    const real y = rsqrt_nr(t);
    const real gb2 = real(2.0) * g * b;
    e_next = t * y;
    s = f * y;
    c = g * y;
    g = d_up - p;
    const real r2 = fma(d_i - g, s, gb2 * y);
    p = s * r2;
    d_next = g + p;
    g = fma(c, r2, -b);
    d_up = d_i;
}

// Orders this wave's LDS writes before its later LDS reads.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 2 x 2 Hermitian-definite pencil (T, G) (G close to I): eigenvalues
// mu1 <= mu2 and G-orthonormal eigenvectors X = [x1 x2], by a Cholesky of G
// and the closed-form eigensystem of C = L^-1 T L^-H. Arguments wave-uniform.
__device__ __forceinline__ void eig2_pencil(real t11, real t22, cx t12, real g11, real g22, cx g12,
                                            real& mu1, real& mu2, cx (&X)[2][2]) {
    const real l11 = sqrt_r(g11);
    const cx l21 = cm(g12.x / l11, -g12.y / l11);            // conj(g12) / l11
    const real l22 = sqrt_r(g22 - cabs2(l21));
    const cx m11 = cm(t11 / l11, 0.0), m12 = cm(t12.x / l11, t12.y / l11);
    const cx m21 = (cm(t12.x, -t12.y) - cmul(l21, m11)) * (real(1.0) / l22);
    const cx m22 = (cm(t22, 0.0) - cmul(l21, m12)) * (real(1.0) / l22);
    const cx c21 = m21 * (real(1.0) / l11);
    const real a = t11 / (l11 * l11);
    const real c = (m22 - cmulc(c21, l21)).x / l22;
    const cx bb = cm(c21.x, -c21.y);                          // C(1, 2)
    const real d = real(0.5) * (c - a);
    const real r = sqrt_r(d * d + cabs2(bb));
    mu1 = real(0.5) * (a + c) - r;
    mu2 = real(0.5) * (a + c) + r;
    cx y1a, y1b;                                              // eigenvector of mu1
    if (d >= real(0.0)) {
        y1a = cm(d + r, 0.0);
        y1b = cm(-bb.x, bb.y);
    } else {
        y1a = bb;
        y1b = cm(d - r, 0.0);
    }
    const real nrm = sqrt_r(cabs2(y1a) + cabs2(y1b));
    if (nrm > real(0.0)) {
        y1a = y1a * (real(1.0) / nrm);
        y1b = y1b * (real(1.0) / nrm);
    } else {
        y1a = cm(1.0, 0.0);
        y1b = cm(0.0, 0.0);
    }
    const cx y2a = cm(-y1b.x, y1b.y), y2b = cm(y1a.x, -y1a.y);
    const cx cl21 = cm(l21.x, -l21.y);
    X[1][0] = y1b * (real(1.0) / l22);
    X[0][0] = (y1a - cmul(cl21, X[1][0])) * (real(1.0) / l11);
    X[1][1] = y2b * (real(1.0) / l22);
    X[0][1] = (y2a - cmul(cl21, X[1][1])) * (real(1.0) / l11);
}

// A 16 x 16 complex tile on the matrix cores in the module precision:
// v_mfma_f64_16x16x4f64, or v_mfma_f32_16x16x4f32 in the FP32 build. Both take
// A(i, k) from lane i + 16 k and B(k, j) from lane j + 16 k (i, j < 16,
// k < 4) and leave D(i, j) in lane j + 16 g, element e: FP64 with g = i % 4,
// e = i / 4; FP32 with g = i / 4, e = i % 4 (acc_row gives i). Element e of
// a result is therefore the B operand of the k-step over rows acc_row(e, g).
// A complex product is four real MFMAs into a (re, im) pair of accumulators.
typedef real r4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ r4 mfma16(real a, real b, r4 c) {
#ifdef GENEIG_FP32
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
#else
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
#endif
}
__device__ __forceinline__ constexpr int acc_row(int e, int g) {
#ifdef GENEIG_FP32
    return 4 * g + e;
#else
    return 4 * e + g;
#endif
}
struct ctile {
    r4 re, im;
};
__device__ __forceinline__ void cmfma(ctile& c, cx a, cx b) {
    c.re = mfma16(a.x, b.x, c.re);
    c.re = mfma16(-a.y, b.y, c.re);
    c.im = mfma16(a.x, b.y, c.im);
    c.im = mfma16(a.y, b.x, c.im);
}
__device__ __forceinline__ cx tile_get(const ctile& t, int v) { return cm(t.re[v], t.im[v]); }

// The same product in three real MFMAs (Gauss's 3M form): p += ar br,
// q += ai bi, s += (ar + ai)(br + bi); then re = p - q, im = s - p - q. The
// rounding error of im is bounded by eps (|ar| + |ai|)(|br| + |bi|) per term,
// at most twice the four-product form's. bs = br + bi, formed once per B.
struct ctile3 {
    r4 p, q, s;
};
__device__ __forceinline__ void c3_zero(ctile3& c) {
    c.p = r4{real(0.0), real(0.0), real(0.0), real(0.0)};
    c.q = c.p;
    c.s = c.p;
}
__device__ __forceinline__ void cmfma3(ctile3& c, cx a, cx b, real bs) {
    c.p = mfma16(a.x, b.x, c.p);
    c.q = mfma16(a.y, b.y, c.q);
    c.s = mfma16(a.x + a.y, bs, c.s);
}
__device__ __forceinline__ ctile c3_final(const ctile3& c) {
    ctile t;
    t.re = c.p - c.q;
    t.im = c.s - c.p - c.q;
    return t;
}

// The 16 x 16 triangular system L(ib, ib)^H x = acc of block row ib, in
// place in the FP64 tile layout (row r of the tile: element r / 4 of lane
// group r % 4), rows from the last: u_r from its owner lane of the same
// column, then the rows above it updated. Rows at or beyond N are left alone.
template <class FL>
__device__ __forceinline__ void backsub_diag(ctile& acc, FL fL, int ib) {
    const int l = lane_id();
    const int c16 = l & 15, g4 = l >> 4;
#pragma unroll
    for (int r = 15; r >= 0; --r) {
        const int rg = ib * 16 + r;
        if (rg >= N) continue;
        const int own = c16 + 16 * (r % 4);
        const real ur = __shfl(acc.re[r / 4], own, 64);
        const real ui = __shfl(acc.im[r / 4], own, 64);
        const real inv = real(1.0) / fL(rg, rg).x;
        const cx u = cm(ur * inv, ui * inv);
        if (g4 == r % 4) {
            acc.re[r / 4] = u.x;
            acc.im[r / 4] = u.y;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 4 * v + g4;              // row of the tile in this element
            if (4 * v < r) {
                const bool upd = i < r;
                const cx lr = fL(rg, ib * 16 + (upd ? i : 0));
                const cx t = cconjmul(lr, u);
                if (upd) {
                    acc.re[v] -= t.x;
                    acc.im[v] -= t.y;
                }
            }
        }
    }
}

// U(:, this wave's 16-column block cb) = L^-H V, L lower triangular, by
// block back substitution from the last 16-row block: row tile ib is
// V(ib) - sum_{jb > ib} L(jb, ib)^H U(jb) on the matrix cores (the finished
// tiles U(jb) are the B operand, straight from registers), then the 16 x 16
// triangular system L(ib, ib)^H U(ib) = that, solved row by row in the tile
// layout: row r of the tile is element r / 4 of lane group r % 4, so each
// step takes u_r from its owner lane of the same column and updates the rows
// above it. No explicit inverse of L is formed (with an ill-conditioned S a
// product with W = L^-1 doubles the residual that the substitution leaves,
// as zhegv's back-transform does). NBLK: 16-blocks per dimension; KE: the
// k-steps that touch rows below N. fL(i, k): L(i, k) for i >= k, k < N.
// fV(i, j): V(i, j), zero outside N.
template <int NBLK, int KE, class FL, class FV>
__device__ __forceinline__ void backsub_lh(ctile (&Ut)[NBLK], FL fL, FV fV, int cb) {
    static_assert(sizeof(real) == 8, "the FP64 tile layout: row 4 e + g of element e, lane group g");
    const int l = lane_id();
    const int c16 = l & 15, g4 = l >> 4;
    const int j = cb * 16 + c16;
#pragma unroll
    for (int ib = NBLK - 1; ib >= 0; --ib) {
        ctile acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const cx x = fV(ib * 16 + 4 * v + g4, j);
            acc.re[v] = x.x;
            acc.im[v] = x.y;
        }
        if (ib < NBLK - 1) {
            // sum_{k in blocks jb > ib} conj(L(k, 16 ib + i)) U(k, j).
            ctile3 a3;
            c3_zero(a3);
#pragma unroll
            for (int q = ib + 1; q < NBLK; ++q) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if (16 * q + 4 * v >= KE) continue;
                    const cx b = tile_get(Ut[q], v);
                    const real bs = b.x + b.y;
                    const int k = 16 * q + 4 * v + g4;
                    const int i = ib * 16 + c16;
                    const cx a = (k < N && i < N) ? cconj(fL(k, i)) : cm(0.0, 0.0);
                    cmfma3(a3, a, b, bs);
                }
            }
            const ctile sm = c3_final(a3);
            acc.re = acc.re - sm.re;
            acc.im = acc.im - sm.im;
        }
        backsub_diag(acc, fL, ib);
        Ut[ib] = acc;
    }
}

// A per-lane set of N bits in registers (words selected by unrolled
// compares, so no index is dynamic).
struct LaneBits {
    static constexpr int NW = (N + 63) / 64;
    unsigned long long w[NW];
    __device__ __forceinline__ void assign(int k, bool v) {
#pragma unroll
        for (int q = 0; q < NW; ++q)
            if (q == (k >> 6)) {
                const unsigned long long bit = 1ull << (k & 63);
                w[q] = v ? (w[q] | bit) : (w[q] & ~bit);
            }
    }
    // Lowest k >= from whose bit equals v; NW * 64 if there is none.
    __device__ __forceinline__ int find(int from, bool v) const {
        int r = NW * 64;
#pragma unroll
        for (int q = NW - 1; q >= 0; --q) {
            unsigned long long x = v ? w[q] : ~w[q];
            const int sh = from - 64 * q;
            if (sh >= 64) x = 0;
            else if (sh > 0) x &= ~0ull << sh;
            if (x) r = 64 * q + __builtin_ctzll(x);
        }
        return r;
    }
};

// Implicit QL recurrence for LANES matrices at once, one per lane of the
// calling wave (lanes LANES and up pass active = false), with the rotations
// logged (the first half of the split QL; method: Bowdler, Martin, Reinsch &
// Wilkinson, Handbook II/3). d and e of lane l's matrix sit in LDS at
// ds[i * LANES + l] and es[i * LANES + l]: two lanes at any indices never
// share a bank (LANES a power of two up to 64); d is left holding the
// eigenvalues. The recurrence is one dependent chain per lane, so the wave's
// speed is its latency; fewer matrices per wave (more waves, each with less
// LDS) let the SIMDs interleave waves. The lanes advance as per-lane state machines, every lane by
// one rotation per iteration; a lane whose sweep ends picks its next window
// in the same iteration from a bit mask of negligible off-diagonals (no scan
// of the off-diagonals). Sweep k's window (l, m) goes to sw[2k], sw[2k + 1]
// and its rotations (c, s) to rot[2 (r + j)], rot[2 (r + j) + 1] (r = the
// rotations of the sweeps before it, j = i - l). Returns through nsw the
// sweeps, through unconverged the off-diagonals left after QL_MAXIT sweeps at
// one l (LAPACK's INFO), and through overflow whether the log's capacity
// (cap sweeps, rot_cap rotations) was reached; lanes with active false do
// nothing but take part in the wave's loop.
template <int LANES>
__device__ __forceinline__ void ql_lanes_logged(real* ds, real* es, bool active, int cap, int rot_cap, int* sw,
                                                real* rot, int& nsw, int& unconverged, int& overflow) {
    static_assert(LANES >= 1 && LANES <= 64 && (LANES & (LANES - 1)) == 0, "lanes: a power of two up to 64");
    const int lane = lane_id() & (LANES - 1);
    const real eps = REAL_HALF_ULP;
    // Bit k of small: off-diagonal k is negligible against its diagonals
    // (|e_k| <= eps (|d_k| + |d_k+1|)); bits N-1 and up are set. A sweep over
    // [l, m] changes only e_l .. e_m and d_l .. d_m, and each rotation leaves
    // its e_(i+1), d_(i+1) and d_(i+2) final for the sweep, so the rotations
    // re-test exactly the bits the sweep can change and the choice of the
    // next window needs no scan: l moves to the lowest clear bit at or above
    // it (the deflations), m is the lowest set bit above that.
    LaneBits small;
#pragma unroll
    for (int q = 0; q < LaneBits::NW; ++q) small.w[q] = 0;
    for (int k = N - 1; k < 64 * LaneBits::NW; ++k) small.assign(k, true);
    for (int k = 0; k < N - 1; ++k)
        small.assign(k, fabs(es[k * LANES + lane]) <= eps * (fabs(ds[k * LANES + lane]) + fabs(ds[(k + 1) * LANES + lane])));
    enum { START = 0, SWEEP = 1, DONE = 2 };
    int phase = active ? START : DONE;
    int l = 0, m = 0, i = 0, iter = 0, nrot = 0;
    nsw = 0;
    unconverged = 0;
    overflow = 0;
    real s = real(1.0), c = real(1.0), p = real(0.0), g = real(0.0), d_up = real(0.0), d_l = real(0.0);
    real d_prev = real(0.0);                            // d_(i+2) of the current rotation
    real e_cur = real(0.0), d_cur = real(0.0);          // (e_i, d_i), fetched ahead
    int step = 0;
    while (__ballot(phase != DONE)) {
        // Lanes between sweeps are served together, on every fourth
        // iteration (or at once when no lane is sweeping): the wave pays for
        // the window choice and shift a quarter as often, a lane waits at
        // most three iterations per sweep.
        const bool serve = (step++ & 7) == 0 || __ballot(phase == SWEEP) == 0;
        if (phase == SWEEP) {
            real e_next, d_next, d_i = d_up;
            ql_rotation(e_cur, d_cur, s, c, g, p, d_i, e_next, d_next);
            es[(i + 1) * LANES + lane] = e_next;
            ds[(i + 1) * LANES + lane] = d_next;
            rot[2 * (nrot + i - l)] = c;
            rot[2 * (nrot + i - l) + 1] = s;
            if (i + 1 < m) small.assign(i + 1, fabs(e_next) <= eps * (fabs(d_next) + fabs(d_prev)));
            d_prev = d_next;
            // The next rotation's operands, fetched now so that their latency
            // passes during the rest of the iteration: e_(i-1), d_(i-1) and
            // d_up = the original d_i (this sweep writes d_i only in its
            // next rotation). Unconditional, and after this rotation's last
            // use of the current ones (the scheduling barrier), so the loads
            // land in e_cur / d_cur / d_up themselves.
            __builtin_amdgcn_sched_barrier(0);
            const int ip = i > 0 ? i - 1 : 0;
            e_cur = es[ip * LANES + lane];
            d_cur = ds[ip * LANES + lane];
            d_up = ds[i * LANES + lane];
            if (--i < l) {
                const real dl_new = d_l - p;
                ds[l * LANES + lane] = dl_new;
                es[l * LANES + lane] = g;
                es[m * LANES + lane] = real(0.0);
                small.assign(l, fabs(g) <= eps * (fabs(dl_new) + fabs(d_prev)));
                small.assign(m, true);
                nrot += m - l;
                phase = START;
            }
        }
        if (phase == START && serve) {
            // Deflate: l moves past every negligible off-diagonal at l.
            const int l_open = small.find(l, false);
            const int l_new = l_open < N ? l_open : N;
            if (l_new != l) {
                l = l_new;
                iter = 0;
            }
            if (l >= N) {
                phase = DONE;
            } else {
                m = small.find(l, true);
                if (iter == QL_MAXIT) {
                    for (int k = l; k < N - 1; ++k) unconverged += (es[k * LANES + lane] != real(0.0)) ? 1 : 0;
                    phase = DONE;
                } else if (nsw >= cap || nrot + (m - l) > rot_cap) {
                    overflow = 1;
                    phase = DONE;
                } else {
                    ++iter;
                    d_l = ds[l * LANES + lane];
                    const real e_l = es[l * LANES + lane];
                    // The shift (Wilkinson's, from the leading 2 x 2 block)
                    // through reciprocals instead of correctly rounded
                    // divisions and square root: it only steers convergence,
                    // and every rotation stays exactly orthogonal whatever
                    // its value. e_l is not negligible here, so |g| < 1 / (2
                    // eps) and g^2 + 1 stays finite.
                    g = (ds[(l + 1) * LANES + lane] - d_l) * rcp_nr(real(2.0) * e_l);
                    const real t = fma(g, g, real(1.0));
                    const real r = t * rsqrt_nr(t);
                    d_up = ds[m * LANES + lane];
                    e_cur = es[(m - 1) * LANES + lane];
                    d_cur = ds[(m - 1) * LANES + lane];
                    d_prev = d_up;
                    g = d_up - d_l + e_l * rcp_nr(g + copysign(r, g));
                    s = real(1.0);
                    c = real(1.0);
                    p = real(0.0);
                    i = m - 1;
                    sw[2 * nsw] = l;
                    sw[2 * nsw + 1] = m;
                    ++nsw;
                    phase = SWEEP;
                }
            }
        }
    }
}

}  // namespace gk

// The staged kernels (orders above 64; their Cholesky for every order).
#include "geneig_large.hip"

#if GENEIG_N <= 64

namespace gk {

// Shared-memory work area of one matrix.
struct Work {
    cx A[N * LDA];        // S -> L (lower); Householder vectors (strict upper, row k = v_k)
    cx B[N * LDA];        // H -> M -> T; then Z (real); then V
    cx tau[N];
    cx vbuf[N];
    cx wbuf[N];
    cx vb[2][N];          // Householder vector / Cholesky pivot column, by step parity
    cx wb[NW][N];         // per-wave copy of the rank-2 vector w
    real invL[N];        // 1 / L(k,k)
    real d[N];
    real e[N];
    cx part[2][NW * N];   // reflector partial sums, real-buffered
    int flag[2];
};

// ---------------------------------------------------------------------------
// Stage 1: Cholesky S = L L^H (lower, right-looking). Returns 0 or k+1 for the
// first pivot k that is not positive (zpotrf's INFO). Each step updates the
// trailing lower triangle with the unscaled pivot column and scales the
// previous pivot column, so no thread reads a value another thread writes in
// the same step; one barrier per step.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int potrf(Work& w) {
    // Column j belongs to wave j % NW. The owner of column k finishes it
    // (pivot check, scaling) right after its own trailing updates of step k-1
    // and publishes it in vb[k & 1]; all waves then update their own columns.
    // One barrier per step.
    const int lane = lane_id(), wave = wave_id();
    auto finish_column = [&](int k) {
        const cx a = (lane >= k && lane < N) ? w.A[k * LDA + lane] : cm(0.0, 0.0);
        const real dkk = bcast(a.x, k);
        if (!(dkk > real(0.0))) {
            if (lane == 0) w.flag[1] = k + 1;
            return;
        }
        const real rinv = real(1.0) / sqrt_r(dkk);
        const cx l = a * rinv;
        if (lane >= k && lane < N) w.A[k * LDA + lane] = l;
        if (lane < N) w.vb[k & 1][lane] = l;
        if (lane == 0) w.invL[k] = rinv;
    };
    if (threadIdx.x == 0) w.flag[1] = 0;
    __syncthreads();
    if (wave_id() == 0) finish_column(0);
    __syncthreads();
    for (int k = 0; k < N; ++k) {
        const int fail = w.flag[1];
        if (fail) return fail;
        const int p = k & 1;
        const int lr = lane < N ? lane : N - 1;
        const cx li = w.vb[p][lr];
        // All loads first (clamped addresses), then masked stores: no branch
        // separates the LDS reads, so their latencies overlap.
        cx a[CPW], lj[CPW];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int j = (wave + NW * c < N) ? wave + NW * c : N - 1;
            a[c] = w.A[j * LDA + lr];
            lj[c] = w.vb[p][j];
        }
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int j = wave + NW * c;
            if (j > k && j < N && lane >= j && lane < N) w.A[j * LDA + lane] = a[c] - cmulc(li, lj[c]);
        }
        if (k + 1 < N && wave == (k + 1) % NW) finish_column(k + 1);
        __syncthreads();
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Stage 2: M = L^-1 H L^-H. Left solve column by column (each wave owns
// columns, lane = row), then right solve row by row (each wave owns rows,
// lane = column). Pivot values travel by readlane: no barriers inside.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void trsm_left(Work& w) {
    // X = L^-1 B, column j owned by wave j % NW (lane = row). Step k reads the
    // pivot entry B(k, j) with a broadcast LDS read; a wave's own writes are
    // visible to its later reads, so no barrier is needed between steps.
    const int lane = lane_id(), wave = wave_id();
    const int lr = lane < N ? lane : N - 1;
    for (int k = 0; k < N; ++k) {
        const real inv = w.invL[k];
        const cx lik = w.A[k * LDA + lr];
        cx pv[CPW], bv[CPW];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int j = (wave + NW * c < N) ? wave + NW * c : N - 1;
            pv[c] = w.B[j * LDA + k];
            bv[c] = w.B[j * LDA + lr];
        }
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int j = wave + NW * c;
            const cx x = pv[c] * inv;
            const cx nb = (lane == k) ? x : bv[c] - cmul(lik, x);
            if (j < N && lane >= k && lane < N) w.B[j * LDA + lane] = nb;
        }
    }
}

__device__ __forceinline__ void trsm_right(Work& w) {
    // Y L^H = X row by row, row i owned by wave i % NW (lane = column):
    // Y(i,k) = X(i,k) / L(k,k), then X(i,j) -= Y(i,k) conj(L(j,k)), j > k.
    const int lane = lane_id(), wave = wave_id();
    const int lr = lane < N ? lane : N - 1;
    for (int k = 0; k < N; ++k) {
        const real inv = w.invL[k];
        const cx ljk = w.A[k * LDA + lr];
        cx pv[CPW], bv[CPW];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int i = (wave + NW * c < N) ? wave + NW * c : N - 1;
            pv[c] = w.B[k * LDA + i];
            bv[c] = w.B[lr * LDA + i];
        }
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int i = wave + NW * c;
            const cx y = pv[c] * inv;
            const cx nb = (lane == k) ? y : bv[c] - cmulc(y, ljk);
            if (i < N && lane >= k && lane < N) w.B[lane * LDA + i] = nb;
        }
    }
}

// ---------------------------------------------------------------------------
// Stage 3: Householder tridiagonalisation of the Hermitian M (full storage).
// Reflector k: H_k = I - tau_k v v^H with v(k+1) = 1, v(k+2..) kept in
// B(k+2.., k); reflector generation follows LAPACK zlarfg (beta real,
// beta = -sign(Re alpha) * ||(alpha, x)||). Three barriers per step.
// ---------------------------------------------------------------------------
// Householder vector for column k, computed by the wave that owns column k:
// v -> vb[k & 1] (zeros above row k+1) and B(k+2.., k); e[k], d[k], tau[k].
__device__ __forceinline__ void house(Work& w, int k) {
    const int lane = lane_id();
    const int m0 = k + 1;
    const cx x = (lane >= m0 && lane < N) ? w.B[k * LDA + lane] : cm(0.0, 0.0);
    const cx alpha = bcast(x, m0);
    const real xn2 = wave_sum(lane > m0 ? cabs2(x) : real(0.0));
    cx t = cm(0.0, 0.0);
    real beta = alpha.x;
    cx v = cm(0.0, 0.0);
    if (xn2 == real(0.0) && alpha.y == real(0.0)) {
        if (lane == m0) v = cm(1.0, 0.0);
    } else {
        const real nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        beta = (alpha.x >= real(0.0)) ? -nrm : nrm;
        t = cm((beta - alpha.x) / beta, -alpha.y / beta);
        const cx den = cm(alpha.x - beta, alpha.y);
        const real dd = real(1.0) / cabs2(den);
        const cx scal = cm(den.x * dd, -den.y * dd);   // 1 / (alpha - beta)
        if (lane == m0) v = cm(1.0, 0.0);
        else if (lane > m0) v = cmul(x, scal);
    }
    if (lane > m0 && lane < N) w.B[k * LDA + lane] = v;
    if (lane < N) w.vb[k & 1][lane] = v;
    if (lane == 0) {
        w.e[k] = beta;
        w.d[k] = w.B[k * LDA + k].x;
        w.tau[k] = t;
    }
}

__device__ __forceinline__ void hetrd(Work& w) {
    // Column j belongs to wave j % NW. Step k: every wave forms partial sums
    // of p = M v over its own columns (barrier), combines them into w in its
    // own lanes, applies the rank-2 update to its own columns, and the owner of
    // column k+1 then builds the next Householder vector (barrier).
    const int lane = lane_id(), wave = wave_id();
    if (wave == 0) house(w, 0);
    __syncthreads();
    for (int k = 0; k < N - 1; ++k) {
        const int m0 = k + 1;
        const int p = k & 1;
        const cx t = w.tau[k];
        if (t.x != real(0.0) || t.y != real(0.0)) {
            const int lr = lane < N ? lane : N - 1;
            const cx vl = (lane < N) ? w.vb[p][lr] : cm(0.0, 0.0);
            cx a[CPW], vj[CPW];
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int j = (wave + NW * c < N) ? wave + NW * c : N - 1;
                a[c] = w.B[j * LDA + lr];
                vj[c] = w.vb[p][j];
            }
            cx acc = cm(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int j = wave + NW * c;
                const cx prod = cmul(a[c], vj[c]);
                if (j >= m0 && j < N) acc = acc + prod;
            }
            if (lane < N) w.part[0][wave * N + lane] = acc;
            __syncthreads();
            cx pi = cm(0.0, 0.0);
            if (lane < N) {
#pragma unroll
                for (int q = 0; q < NW; ++q) pi = pi + w.part[0][q * N + lane];
            }
            pi = cmul(t, pi);
            const cx pvl = cconjmul(pi, vl);
            const cx pv = cm(wave_sum(pvl.x), wave_sum(pvl.y));
            const cx cc = cmul(t, pv) * -0.5;
            const cx wi = pi + cmul(cc, vl);
            if (lane < N) w.wb[wave][lane] = wi;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            cx wj[CPW];
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int j = (wave + NW * c < N) ? wave + NW * c : N - 1;
                wj[c] = w.wb[wave][j];
            }
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int j = wave + NW * c;
                if (j >= m0 && j < N && lane >= m0 && lane < N)
                    w.B[j * LDA + lane] = a[c] - cmulc(vl, wj[c]) - cmulc(wi, vj[c]);
            }
        }
        if (k + 1 < N - 1 && wave == (k + 1) % NW) house(w, k + 1);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        w.d[N - 1] = w.B[(N - 1) * LDA + N - 1].x;
        w.e[N - 1] = 0.0;
        w.tau[N - 1] = cm(0.0, 0.0);
    }
    __syncthreads();
}

// Move the Householder vectors from B's lower part to A's strict upper
// triangle (A(k, i) = v_k(i), i >= k+2) so that B can hold Z and V.
__device__ __forceinline__ void stash_reflectors(Work& w) {
    const int lane = lane_id(), wave = wave_id();
    for (int k = wave; k < N - 2; k += NW)
        if (lane >= k + 2 && lane < N) w.A[lane * LDA + k] = w.B[k * LDA + lane];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Stage 4: implicit QL with Wilkinson-type shift on the real symmetric
// tridiagonal (d, e), one wave per matrix. Lane j holds d[j] and e[j]; the
// scalar recurrence runs redundantly in every lane, reading entries with
// v_readlane and writing them back with a lane select. Lane r holds row r of
// Z in registers: the sweep is unrolled over the static column index, each
// step predicated on the active window [l, m), so Z never leaves the VGPRs.
// Rotation lengths use a reciprocal square root refined by two Newton steps
// instead of a correctly rounded sqrt and division (c^2 + s^2 = 1 to a few
// ulp). Convergence scan: every lane tests its own off-diagonal, one ballot
// finds the first negligible one. Method: Bowdler, Martin, Reinsch &
// Wilkinson, "The QR and QL algorithms for symmetric matrices" (Handbook
// II/3). Returns zheev-style INFO (off-diagonals left unconverged, 0 = ok).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int ql_wave(real& dreg, real& ereg, real (&z)[N]) {
    const int lane = lane_id();
    const real eps = REAL_HALF_ULP;
    int l = 0, iter = 0;
    while (l < N) {
        const real dnext = __shfl_down(dreg, 1, 64);
        const bool small = fabs(ereg) <= eps * (fabs(dreg) + fabs(dnext));
        unsigned long long mask = __ballot(small && lane >= l && lane < N - 1);
        mask |= 1ull << (N - 1);
        mask &= ~((1ull << l) - 1ull);
        const int m = __ffsll((long long)mask) - 1;
        if (m == l) {
            ++l;
            iter = 0;
            continue;
        }
        if (iter++ == QL_MAXIT) {
            const unsigned long long nz = __ballot(ereg != real(0.0) && lane >= l && lane < N - 1);
            return __popcll(nz);
        }
        const real d_l = bcast(dreg, l);
        const real e_l = bcast(ereg, l);
        real g = (bcast(dreg, l + 1) - d_l) / (real(2.0) * e_l);
        real r = sqrt(g * g + real(1.0));
        g = bcast(dreg, m) - d_l + e_l / (g + copysign(r, g));
        real s = 1.0, c = 1.0, p = 0.0;
        real d_up = bcast(dreg, m);      // original d[i+1] for the current i
        // Step i reads lane i before any step of this sweep writes it (step i
        // writes lane i + 1), so the reads come from copies taken here: they
        // then wait for nothing written during the sweep.
        const real d0 = dreg, e0 = ereg;
#pragma unroll
        for (int i = N - 2; i >= 0; --i) {
            if (i < m && i >= l) {
                real e_next, d_next;
                ql_rotation(bcast(e0, i), bcast(d0, i), s, c, g, p, d_up, e_next, d_next);
                ereg = sel_lane(ereg, e_next, i + 1);
                dreg = sel_lane(dreg, d_next, i + 1);
                const real zi = z[i], zi1 = z[i + 1];
                z[i + 1] = fma(s, zi, c * zi1);
                z[i] = fma(c, zi, -(s * zi1));
            }
        }
        dreg = sel_lane(dreg, d_l - p, l);
        ereg = sel_lane(ereg, g, l);
        ereg = sel_lane(ereg, 0.0, m);
    }
    return 0;
}

// Ascending position of lane j's eigenvalue (ties by index).
__device__ __forceinline__ int ascending_rank(real dreg) {
    const int lane = lane_id();
    int rank = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const real dk = bcast(dreg, k);
        rank += (dk < dreg || (dk == dreg && k < lane)) ? 1 : 0;
    }
    return rank;
}

// Wave-level eigen-solve of the tridiagonal held in dsrc/esrc (LDS): writes the
// ascending eigenvalues to W_out and leaves Z (columns in ascending order) in
// `z` (lane r = row r) plus its rank permutation in `rank`.
__device__ __forceinline__ int tridiag_eig_wave(const real* dsrc, const real* esrc, real* __restrict__ W_out,
                                real (&z)[N], int& rank) {
    const int lane = lane_id();
    real dreg = (lane < N) ? dsrc[lane] : 0.0;
    real ereg = (lane < N - 1) ? esrc[lane] : 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) z[j] = (lane == j) ? 1.0 : 0.0;
    const int info = ql_wave(dreg, ereg, z);
    rank = ascending_rank(lane < N ? dreg : __builtin_inf());
    if (lane < N) W_out[rank] = dreg;
    return info;
}

// Write this wave's Z rows into Zs (real, leading dimension LDZ) with columns
// permuted to ascending eigenvalue order.
__device__ __forceinline__ void store_z_sorted(real* Zs, const real (&z)[N], int rank) {
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int pos = __builtin_amdgcn_readlane(rank, j);
        if (lane < N) Zs[pos * LDZ + lane] = z[j];
    }
}

// ---------------------------------------------------------------------------
// Stage 5: V = Q Z, Q = H_0 H_1 ... H_{N-2}, applied last first. Each wave
// owns whole rows of V in registers (lane = column). For reflector k the row
// vector y = tau_k v_k^H V is a sum over rows: each wave adds its own rows,
// the NW partial sums meet in LDS behind one barrier, then every wave updates
// its rows V(i,:) -= v_k(i) y. Partial-sum buffers alternate between applied
// reflectors so a fast wave never overwrites sums a slow wave still reads.
// Z is read from B's storage as real; V is written back to B (complex).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void apply_q(Work& w) {
    const int lane = lane_id(), wave = wave_id();
    const real* Z = reinterpret_cast<const real*>(w.B);
    cx b[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int i = wave + NW * c;
        b[c] = cm((i < N && lane < N) ? Z[lane * LDZ + i] : 0.0, 0.0);
    }
    int applied = 0;
    for (int k = N - 2; k >= 0; --k) {
        const cx t = w.tau[k];
        if (t.x == real(0.0) && t.y == real(0.0)) continue;
        const int m0 = k + 1;
        cx v[CPW];
        cx acc = cm(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int i = wave + NW * c;
            v[c] = cm(0.0, 0.0);
            if (i == m0) v[c] = cm(1.0, 0.0);
            else if (i > m0 && i < N) v[c] = w.A[i * LDA + k];
            acc = acc + cconjmul(v[c], b[c]);
        }
        cx* part = w.part[applied & 1];
        ++applied;
        if (lane < N) part[wave * N + lane] = acc;
        __syncthreads();
        cx y = cm(0.0, 0.0);
        if (lane < N) {
#pragma unroll
            for (int q = 0; q < NW; ++q) y = y + part[q * N + lane];
        }
        y = cmul(t, y);
#pragma unroll
        for (int c = 0; c < CPW; ++c) b[c] = b[c] - cmul(v[c], y);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int i = wave + NW * c;
        if (i < N && lane < N) w.B[lane * LDA + i] = b[c];
    }
}

// Stage 6: U = L^-H V column by column (lane = row), written straight to HBM.
__device__ __forceinline__ void backtrsm_store(Work& w, cx* __restrict__ Ub) {
    const int lane = lane_id(), wave = wave_id();
    cx b[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int j = wave + NW * c;
        b[c] = (j < N && lane < N) ? w.B[j * LDA + lane] : cm(0.0, 0.0);
    }
#pragma unroll 2
    for (int k = N - 1; k >= 0; --k) {
        // L(k, i) for lane i < k: column i, row k.
        const cx lki = (lane < k) ? w.A[lane * LDA + k] : cm(0.0, 0.0);
        const real inv = w.invL[k];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const cx uk = bcast(b[c], k) * inv;
            if (lane == k) b[c] = uk;
            else if (lane < k) b[c] = b[c] - cconjmul(lki, uk);
        }
    }
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int j = wave + NW * c;
        if (j < N && lane < N) Ub[j * N + lane] = b[c];
    }
}

__device__ __forceinline__ void load_pair(Work& w, const cx* __restrict__ Hb, const cx* __restrict__ Sb) {
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        w.A[j * LDA + i] = Sb[e];
        w.B[j * LDA + i] = Hb[e];
    }
    __syncthreads();
}

__device__ __forceinline__ void fail_store(int code, real* __restrict__ Wb, cx* __restrict__ Ub,
                           int* __restrict__ info_b) {
    for (int e = threadIdx.x; e < N * N; e += NT) Ub[e] = cm(0.0, 0.0);
    for (int e = threadIdx.x; e < N; e += NT) Wb[e] = __builtin_nan("");
    if (threadIdx.x == 0) *info_b = code;
}

// ---------------------------------------------------------------------------
// Stash of one matrix's factor data in its own U slot (global memory) while
// the workgroup moves on to the next matrix, in the layout of A after
// stash_reflectors: L on and below the diagonal with 1/L(k,k) in the
// diagonal's imaginary part, tau_k on the first superdiagonal (k, k+1), and
// reflector entries v_k(i) at (k, i), i >= k+2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stash_factors(const Work& w, cx* __restrict__ Ub) {
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        cx v;
        if (i > j) v = w.A[j * LDA + i];
        else if (i == j) v = cm(w.A[j * LDA + j].x, w.invL[j]);
        else if (i == j - 1) v = w.tau[i];
        else v = w.B[i * LDA + j];
        Ub[e] = v;
    }
}

__device__ __forceinline__ void unstash_factors(Work& w, const cx* __restrict__ Ub) {
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        const cx v = Ub[e];
        if (i == j) {
            w.A[j * LDA + j] = cm(v.x, 0.0);
            w.invL[j] = v.y;
        } else {
            w.A[j * LDA + i] = v;
            if (i == j - 1) w.tau[i] = v;
        }
    }
    if (threadIdx.x == 0) w.tau[N - 1] = cm(0.0, 0.0);
}

// ---------------------------------------------------------------------------
// Back half for one matrix, run by a single wave with lane j holding column j
// of V in registers (rows statically indexed; every step of the unrolled row
// loops is predicated on a wave-uniform condition):
//   Z (ascending columns) -> V = Q Z (reflectors applied last first: per
//   reflector a column-local dot product and update, no cross-lane traffic)
//   -> U = L^-H V (backward substitution, pivots local to each lane).
// Reflectors, tau, L and 1/L(k,k) are read from the stash layout in global
// memory (wave-uniform addresses). Z arrives in `zbuf` (this wave's LDS slice,
// N x N doubles, column-major) written by store_z_slice.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_z_slice(real* zbuf, const real (&z)[N], int rank) {
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int pos = __builtin_amdgcn_readlane(rank, j);
        if (lane < N) zbuf[pos * N + lane] = z[j];
    }
}

__device__ __forceinline__ void load_z_columns(const real* zbuf, cx (&b)[N]) {
    const int lane = lane_id();
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] = cm((lane < N) ? zbuf[lane * N + i] : 0.0, 0.0);
}

// Packed per-wave copies of the factors, staged in the wave's LDS slice:
//   reflectors: v_k(i), i = k+2..N-1, at RPK(k) + (i - k - 2); tau_k after them;
//   L by rows:  L(k, i), i <= k, at k(k+1)/2 + i, the diagonal slot holding
//               (1/L(k,k), 0).
constexpr int NREFL = (N - 1) * (N - 2) / 2;
constexpr int NLPK = N * (N + 1) / 2;
__device__ __forceinline__ int rpk(int k) { return k * (2 * N - k - 3) / 2; }

constexpr int RB = 8;                           // rows per branch-free block
constexpr int NRB = (N + RB - 1) / RB;

// Stage reflectors + tau (from the stash in global memory) into `sl`.
__device__ __forceinline__ void stage_reflectors(cx* sl, const cx* __restrict__ Ub) {
    const int lane = lane_id();
    for (int e = lane; e < N * N; e += 64) {
        const int i = e % N, j = e / N;           // stash element (row i, col j)
        if (i < j - 1) sl[rpk(i) + (j - i - 2)] = Ub[e];
        else if (i == j - 1) sl[NREFL + i] = Ub[e];
    }
}

// Stage L (rows packed, 1/L(k,k) on the diagonal) into `sl`.
__device__ __forceinline__ void stage_lower(cx* sl, const cx* __restrict__ Ub) {
    const int lane = lane_id();
    for (int e = lane; e < N * N; e += 64) {
        const int i = e % N, j = e / N;
        if (i > j) sl[i * (i + 1) / 2 + j] = Ub[e];
        else if (i == j) sl[i * (i + 1) / 2 + i] = cm(Ub[e].y, 0.0);
    }
}


// V <- Q V for this lane's column; reflectors from the staged slice.
__device__ __forceinline__ void apply_q_wave(const cx* sl, cx (&b)[N]) {
    for (int k = N - 2; k >= 0; --k) {
        const cx t = sl[NREFL + k];
        if (t.x == real(0.0) && t.y == real(0.0)) continue;
        const int base = rpk(k) - k - 2;
        cx acc = cm(0.0, 0.0);
#pragma unroll
        for (int blk = 0; blk < NRB; ++blk) {
            if (blk * RB + RB - 1 < k + 1) continue;
            cx vv[RB];
#pragma unroll
            for (int ii = 0; ii < RB; ++ii) {
                const int i = blk * RB + ii;
                if (i < N) {
                    const cx r = sl[(i > k + 1) ? base + i : 0];
                    vv[ii] = (i > k + 1) ? r : cm((i == k + 1) ? 1.0 : 0.0, 0.0);
                }
            }
#pragma unroll
            for (int ii = 0; ii < RB; ++ii) {
                const int i = blk * RB + ii;
                if (i < N) acc = acc + cconjmul(vv[ii], b[i]);
            }
        }
        const cx y = cmul(t, acc);
#pragma unroll
        for (int blk = 0; blk < NRB; ++blk) {
            if (blk * RB + RB - 1 < k + 1) continue;
            cx vv[RB];
#pragma unroll
            for (int ii = 0; ii < RB; ++ii) {
                const int i = blk * RB + ii;
                if (i < N) {
                    const cx r = sl[(i > k + 1) ? base + i : 0];
                    vv[ii] = (i > k + 1) ? r : cm((i == k + 1) ? 1.0 : 0.0, 0.0);
                }
            }
#pragma unroll
            for (int ii = 0; ii < RB; ++ii) {
                const int i = blk * RB + ii;
                if (i < N) b[i] = b[i] - cmul(vv[ii], y);
            }
        }
    }
}

// Value barrier: the compiler may not see through it, so a select between
// statically indexed registers is never rewritten as a dynamically indexed
// (scratch) array access.
__device__ __forceinline__ real opq(real x) {
    asm volatile("" : "+v"(x));
    return x;
}
__device__ __forceinline__ cx opq(cx a) { return cm(opq(a.x), opq(a.y)); }

// U = L^-H V for this lane's column; L from the staged slice (rows packed).
// Row blocks entirely above the pivot are untouched, blocks entirely below it
// take a plain update, and only the pivot's block uses selects.
__device__ __forceinline__ void backtrsm_wave(const cx* sl, cx (&b)[N]) {
    for (int k = N - 1; k >= 0; --k) {
        const int row = k * (k + 1) / 2;
        const real inv = sl[row + k].x;
        const int kb = k / RB;
        cx bk = cm(0.0, 0.0);
#pragma unroll
        for (int blk = 0; blk < NRB; ++blk) {
            if (blk != kb) continue;
#pragma unroll
            for (int ii = 0; ii < RB; ++ii) {
                const int i = blk * RB + ii;
                if (i < N) bk = (i == k) ? opq(b[i]) : bk;
            }
        }
        const cx u = bk * inv;
#pragma unroll
        for (int blk = 0; blk < NRB; ++blk) {
            if (blk > kb) continue;
            cx ll[RB];
#pragma unroll
            for (int ii = 0; ii < RB; ++ii) {
                const int i = blk * RB + ii;
                if (i < N) ll[ii] = sl[row + ((i < k) ? i : 0)];
            }
            if (blk < kb) {
#pragma unroll
                for (int ii = 0; ii < RB; ++ii) {
                    const int i = blk * RB + ii;
                    if (i < N) b[i] = b[i] - cconjmul(ll[ii], u);
                }
            } else {
#pragma unroll
                for (int ii = 0; ii < RB; ++ii) {
                    const int i = blk * RB + ii;
                    if (i < N) {
                        const cx nb = b[i] - cconjmul(ll[ii], u);
                        b[i] = (i < k) ? nb : ((i == k) ? opq(u) : b[i]);
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ void store_columns(cx* __restrict__ Ub, const cx (&b)[N]) {
    const int lane = lane_id();
    if (lane < N) {
#pragma unroll
        for (int i = 0; i < N; ++i) Ub[lane * N + i] = b[i];
    }
}

// Per-wave LDS slice: a Z transpose buffer, then the staged reflectors, then L.
constexpr size_t SLICE_BYTES =
    ((((size_t)N * N * 8 > (size_t)NLPK * 16 ? (size_t)N * N * 8 : (size_t)NLPK * 16) >
      (size_t)(NREFL + N) * 16
          ? ((size_t)N * N * 8 > (size_t)NLPK * 16 ? (size_t)N * N * 8 : (size_t)NLPK * 16)
          : (size_t)(NREFL + N) * 16) + 15) / 16 * 16;
constexpr int ZSLICES = (int)(sizeof(Work) / SLICE_BYTES) < NW ? (int)(sizeof(Work) / SLICE_BYTES) : NW;
static_assert(ZSLICES >= 1, "work area must hold one slice");

// Back half of every wave's matrix: Z -> V = Q Z (-> U = L^-H V when `with_l`)
// -> Ub. Waves whose slices do not fit at once take turns; every wave passes
// the same barriers.

__device__ __forceinline__ void back_half_waves(Work& w, bool active, const real (&z)[N], int rank,
                                                cx* __restrict__ Ub, const cx* __restrict__ Lb,
                                                bool with_l) {
    const int wave = wave_id();
    char* base = reinterpret_cast<char*>(&w);
    for (int r = 0; r < (NW + ZSLICES - 1) / ZSLICES; ++r) {
        if (active && wave / ZSLICES == r) {
            cx* sl = reinterpret_cast<cx*>(base + (size_t)(wave % ZSLICES) * SLICE_BYTES);
            real* zbuf = reinterpret_cast<real*>(sl);
            cx b[N];
            store_z_slice(zbuf, z, rank);
            wave_lds_sync();
            load_z_columns(zbuf, b);
            wave_lds_sync();
            stage_reflectors(sl, Ub);
            wave_lds_sync();
            apply_q_wave(sl, b);
            if (with_l) {
                wave_lds_sync();
                stage_lower(sl, Lb);
                wave_lds_sync();
                backtrsm_wave(sl, b);
            }
            store_columns(Ub, b);
        }
        __syncthreads();
    }
}

// Hermitian and triangular N x N matrices packed by columns: element (i, j),
// i >= j, of the lower triangle at pk(j) + i - j.
constexpr int NPK = N * (N + 1) / 2;
__device__ __forceinline__ int pk(int j) { return j * N - (j * (j - 1)) / 2; }

// ===========================================================================
// Register-resident tridiagonalisation (geneig_mx_front): each wave keeps its
// own columns (j = wave + NW c) in registers, lane = row, so the
// matrix-vector products and rank-2 updates never read the matrix from LDS;
// LDS carries only the published v / w and the partial sums.
// ===========================================================================
// Householder vector for column k held in registers (slot c) of its owner.
template <class WK>
__device__ __forceinline__ void house_reg(WK& w, int k, cx& col) {
    const int lane = lane_id();
    const int m0 = k + 1;
    const cx x = (lane >= m0 && lane < N) ? col : cm(0.0, 0.0);
    const cx alpha = bcast(x, m0);
    const real xn2 = wave_sum(lane > m0 ? cabs2(x) : real(0.0));
    cx t = cm(0.0, 0.0);
    real beta = alpha.x;
    cx v = cm(0.0, 0.0);
    if (xn2 == real(0.0) && alpha.y == real(0.0)) {
        if (lane == m0) v = cm(1.0, 0.0);
    } else {
        const real nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        beta = (alpha.x >= real(0.0)) ? -nrm : nrm;
        t = cm((beta - alpha.x) / beta, -alpha.y / beta);
        const cx den = cm(alpha.x - beta, alpha.y);
        const real dd = real(1.0) / cabs2(den);
        const cx scal = cm(den.x * dd, -den.y * dd);   // 1 / (alpha - beta)
        if (lane == m0) v = cm(1.0, 0.0);
        else if (lane > m0) v = cmul(x, scal);
    }
    const real dk = bcast(col.x, k);
    if (lane > m0 && lane < N) col = v;                 // reflector kept in the column
    if (lane < N) w.u.h.vb[k & 1][lane] = v;
    if (lane == 0) {
        w.e[k] = beta;
        w.d[k] = dk;
        w.tau[k] = t;
    }
}

// Tridiagonalisation with the full Hermitian M in the owning waves'
// registers (schedule as in hetrd: two barriers per step).
template <class WK>
__device__ __forceinline__ void hetrd_reg(WK& w, cx (&m)[CPW]) {
    const int lane = lane_id(), wave = wave_id();
    const int lr = lane < N ? lane : N - 1;
    if (wave == 0) house_reg(w, 0, m[0]);
    __syncthreads();
    for (int k = 0; k < N - 1; ++k) {
        const int m0 = k + 1;
        const int p = k & 1;
        const cx t = w.tau[k];
        if (t.x != real(0.0) || t.y != real(0.0)) {
            // v is zero above row m0 and outside N, so the rows above the
            // active block add nothing and (with w zeroed there) take no
            // update; finished and padding columns are skipped by wave-uniform
            // branches. v_j is read from LDS (a broadcast) where it is used,
            // here and in the update below, instead of being held for the
            // whole step (vb[p] stays valid until step k + 2).
            const cx vl = (lane < N) ? w.u.h.vb[p][lr] : cm(0.0, 0.0);
            cx acc = cm(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int j = wave + NW * c;
                if (j >= m0 && j < N) acc = cfma(m[c], w.u.h.vb[p][j], acc);
            }
            if (lane < N) w.u.h.part[wave * N + lane] = acc;
            __syncthreads();
            cx pi = cm(0.0, 0.0);
            if (lane < N) {
#pragma unroll
                for (int q = 0; q < NW; ++q) pi = pi + w.u.h.part[q * N + lane];
            }
            pi = cmul(t, pi);
            const cx pvl = cconjmul(pi, vl);
            const cx pv = cm(wave_sum(pvl.x), wave_sum(pvl.y));
            const cx cc = cmul(t, pv) * -0.5;
            const cx wall = pi + cmul(cc, vl);
            if (lane < N) w.u.h.wb[wave][lane] = wall;
            const cx wi = (lane >= m0 && lane < N) ? wall : cm(0.0, 0.0);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int j = wave + NW * c;
                if (j >= m0 && j < N)
                    m[c] = cfms_conj(wi, w.u.h.vb[p][j], cfms_conj(vl, w.u.h.wb[wave][j], m[c]));
            }
        }
        if (k + 1 < N - 1 && wave == (k + 1) % NW) {
            // The owned column k + 1 through one copy of house_reg (selected
            // and written back by wave-uniform compares; the value barriers
            // keep the selects from becoming a scratch-backed array index).
            const int cs = (k + 1) / NW;
            cx col = m[0];
#pragma unroll
            for (int c = 1; c < CPW; ++c) col = (c == cs) ? opq(m[c]) : col;
            house_reg(w, k + 1, col);
#pragma unroll
            for (int c = 0; c < CPW; ++c) m[c] = (c == cs) ? opq(col) : m[c];
        }
        __syncthreads();
    }
    if (wave == (N - 1) % NW) {
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            if (wave + NW * c == N - 1) {
                const real dl = bcast(m[c].x, N - 1);
                if (lane == 0) {
                    w.d[N - 1] = dl;
                    w.e[N - 1] = 0.0;
                    w.tau[N - 1] = cm(0.0, 0.0);
                }
            }
        }
    }
    __syncthreads();
}

// Stage data kept per matrix of a workgroup between the front and back halves.
struct Staged {
    real d[NW][N];
    real e[NW][N];
    int status[NW];        // 0 = ok, otherwise the matrix is finished (failed)
    int qinfo[NW];
};

}  // namespace gk

using gk::cx;
using gk::real;


// ---------------------------------------------------------------------------
// Fused solve. A workgroup of NW waves takes NW consecutive matrices: it runs
// the dense front half (Cholesky, two-sided solve, tridiagonalisation) for
// each in turn with all waves, stashing the factors in the matrix's own U
// slot; then every wave runs the QL of one matrix at the same time (Z in its
// registers); then the dense back half (V = Q Z, U = L^-H V) for each in turn.
// ---------------------------------------------------------------------------
#ifndef GENEIG_FP32
extern "C" __global__ void __launch_bounds__(GENEIG_NT)
geneig_fused(const cx* __restrict__ H, const cx* __restrict__ S, real* __restrict__ W,
             cx* __restrict__ U, int* __restrict__ info, int batch) {
    using namespace gk;
    __shared__ Work w;
    __shared__ Staged st;
    const int lane = lane_id(), wave = wave_id();
    const int b0 = blockIdx.x * NW;
    for (int q = 0; q < NW; ++q) {
        const int b = b0 + q;
        if (b >= batch) break;
        const size_t off = (size_t)b * N * N;
        load_pair(w, H + off, S + off);
        const int pinfo = potrf(w);
        if (pinfo) {
            fail_store(N + pinfo, W + (size_t)b * N, U + off, info + b);
            if (threadIdx.x == 0) st.status[q] = 1;
            __syncthreads();
            continue;
        }
        trsm_left(w);
        __syncthreads();
        trsm_right(w);
        __syncthreads();
        hetrd(w);
        stash_factors(w, U + off);
        if (threadIdx.x < N) {
            st.d[q][threadIdx.x] = w.d[threadIdx.x];
            st.e[q][threadIdx.x] = w.e[threadIdx.x];
        }
        if (threadIdx.x == 0) st.status[q] = 0;
        __syncthreads();
    }
    real z[N];
    int rank = 0;
    {
        const int b = b0 + wave;
        if (b < batch && st.status[wave] == 0) {
            const int qi = tridiag_eig_wave(st.d[wave], st.e[wave], W + (size_t)b * N, z, rank);
            if (lane == 0) st.qinfo[wave] = qi;
        }
    }
    __syncthreads();
    {
        const int b = b0 + wave;
        const bool active = b < batch && st.status[wave] == 0;
        const size_t off = active ? (size_t)b * N * N : 0;
        back_half_waves(w, active, z, rank, U + off, U + off, true);
        if (active && lane == 0) info[b] = st.qinfo[wave];
    }
}



#endif  // !GENEIG_FP32

// ---------------------------------------------------------------------------
// QL in two kernels. The scalar recurrence of a sweep is the same for every
// row of Z, so instead of every lane of a wave evaluating it for one matrix,
// geneig_ql_record runs it for 64 matrices at once, one per lane (its d and e
// in LDS, element i of lane l at [i][l], so a lane always hits its own bank
// whatever index it is at), as a per-lane state machine that advances every
// lane by one rotation per iteration; a lane whose sweep ends picks its next
// window from a bit mask of negligible off-diagonals in the same iteration
// (no scan of the off-diagonals). Each matrix's sweeps (l, m), rotations, ascending ranks and
// eigenvalues go to its log; geneig_split_ql then rotates the register rows
// of Z with them. A matrix whose QL outgrows its log (or the caller's sweep
// cap, at most the log's) is redone by geneig_ql_full instead.
// ---------------------------------------------------------------------------
namespace gk {
constexpr int QS_CAP = 8 * N;                       // sweeps per matrix
// Recorder matrices per wave: 64. At these orders a large batch keeps every
// SIMD busy with one such wave; 16 per wave (four waves per SIMD) was slower
// (N = 54: FP32 recorder 1.66 -> 1.94 ms): each wave issues the whole
// per-iteration instruction stream for a quarter of the matrices.
constexpr int QL_REC_LANES = 64;
constexpr int QR_CAP = 2 * N * N;                   // rotations per matrix
constexpr size_t QG_SW = 16;                        // int[2 QS_CAP]: (l, m) per sweep
constexpr size_t QG_RANK = QG_SW + 8 * (size_t)QS_CAP;                     // int[N]
constexpr size_t QG_LAM = (QG_RANK + 4 * (size_t)N + 15) / 16 * 16;        // real[N]
constexpr size_t QG_ROT = (QG_LAM + 8 * (size_t)N + 15) / 16 * 16;         // (c, s) pairs
constexpr size_t QG_BYTES = (QG_ROT + 16 * (size_t)QR_CAP + 255) / 256 * 256;
}  // namespace gk

// Bytes of one matrix's QL log, for the host's consistency check.
extern "C" __device__ const unsigned long long geneig_ql_log_bytes = gk::QG_BYTES;

// Matrices per wave of the recorder (one per lane; the other lanes idle).
extern "C" __device__ const int geneig_ql_record_lanes = gk::QL_REC_LANES;

extern "C" __global__ void __launch_bounds__(64)
geneig_ql_record(const real* __restrict__ W, const real* __restrict__ E, const int* __restrict__ info,
                 char* __restrict__ qlog, long long log_stride, int batch, int sweep_cap) {
    using namespace gk;
    constexpr int LN = QL_REC_LANES;
    const int cap = sweep_cap < QS_CAP ? sweep_cap : QS_CAP;
    __shared__ real ds[N][LN];
    __shared__ real es[N][LN];
    const int lane = lane_id();
    const int b = blockIdx.x * LN + lane;
    const bool mine = lane < LN && b < batch && info[b] == 0;
    if (lane < LN) {
        for (int i = 0; i < N; ++i) {
            ds[i][lane] = mine ? W[(size_t)b * N + i] : real(0.0);
            es[i][lane] = (mine && i < N - 1) ? E[(size_t)b * N + i] : real(0.0);
        }
    }
    char* lg = qlog + (size_t)(mine ? b : 0) * (size_t)log_stride;
    int* hdr = reinterpret_cast<int*>(lg);
    int* sw = reinterpret_cast<int*>(lg + QG_SW);
    int* rk = reinterpret_cast<int*>(lg + QG_RANK);
    real* lam = reinterpret_cast<real*>(lg + QG_LAM);
    real* rot = reinterpret_cast<real*>(lg + QG_ROT);
    int nsw, unconverged, overflow;
    ql_lanes_logged<LN>(&ds[0][0], &es[0][0], mine, cap, QR_CAP, sw, rot, nsw, unconverged, overflow);
    if (!mine) return;
    if (!overflow) {
        // Ascending ranks (ties by index) and the eigenvalues.
        for (int k = 0; k < N; ++k) {
            const real dk = ds[k][lane];
            int rank = 0;
            for (int j = 0; j < N; ++j) {
                const real dj = ds[j][lane];
                rank += (dj < dk || (dj == dk && j < k)) ? 1 : 0;
            }
            rk[k] = rank;
            lam[rank] = dk;
        }
    }
    hdr[0] = nsw;
    hdr[1] = overflow;
    hdr[2] = unconverged;
}

// QL, second half: one wave per matrix (lane r = row r of Z, in registers)
// applies the logged rotations sweep by sweep (each sweep's rotations fetched
// by the lanes ahead of time, then broadcast by v_readlane), writes the sorted
// eigenvalues to W and Z with its columns in ascending order to Zw, and info
// (the number of unconverged off-diagonals, LAPACK's convention). Matrices
// whose log overflowed are left to geneig_ql_full (info stays 0 for them).
namespace gk {
// Z (columns in ascending order of rank) to Zb and info, one wave per matrix.
__device__ __forceinline__ void store_z(real* Zb, const real (&z)[N], int rank, int* info, int b, int qi) {
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int pos = __builtin_amdgcn_readlane(rank, j);
        if (lane < N) Zb[pos * N + lane] = z[j];
    }
    if (lane == 0) info[b] = qi;
}
}  // namespace gk

extern "C" __global__ void __launch_bounds__(GENEIG_NT)
geneig_split_ql(real* __restrict__ W, const real* __restrict__ E, real* __restrict__ Zw,
                int* __restrict__ info, const char* __restrict__ qlog, long long log_stride, int batch) {
    using namespace gk;
    __shared__ real rcs[NW][2][2][N];             // [wave][sweep parity][c or s][column]
    const int wave = wave_id();
    const int b = blockIdx.x * NW + wave;
    if (b >= batch) return;
    if (info[b] != 0) return;
    const int lane = lane_id();
    real* Wb = W + (size_t)b * N;
    const char* lg = qlog + (size_t)b * (size_t)log_stride;
    const int* hdr = reinterpret_cast<const int*>(lg);
    if (hdr[1]) return;
    real z[N];
#pragma unroll
    for (int j = 0; j < N; ++j) z[j] = (lane == j) ? 1.0 : 0.0;
    const int* sw = reinterpret_cast<const int*>(lg + QG_SW);
    const real* rot = reinterpret_cast<const real*>(lg + QG_ROT);
    const int nsw = hdr[0];
    int off = 0;
    int l = nsw > 0 ? sw[0] : 0, m = nsw > 0 ? sw[1] : 0;
    // Lane p holds the rotation of column p (identity outside the window).
    real cr = (lane >= l && lane < m) ? rot[2 * (lane - l)] : real(1.0);
    real sr = (lane >= l && lane < m) ? rot[2 * (lane - l) + 1] : real(0.0);
    for (int t = 0; t < nsw; ++t) {
        // The sweep's rotations at their column index in this wave's LDS
        // slice: the columns then run unrolled at compile-time addresses,
        // chunks of QZ_CH that miss the window skipped, the rest without
        // branches (an identity rotation changes nothing, exactly).
        real* rc = rcs[wave][t & 1][0];
        real* rs = rcs[wave][t & 1][1];
        if (lane < N) {
            rc[lane] = cr;
            rs[lane] = sr;
        }
        wave_lds_sync();
        const int off_n = off + (m - l);
        const bool last = t + 1 == nsw;
        const int ln = last ? 0 : sw[2 * t + 2], mn = last ? 0 : sw[2 * t + 3];
        cr = (lane >= ln && lane < mn) ? rot[2 * (off_n + lane - ln)] : real(1.0);
        sr = (lane >= ln && lane < mn) ? rot[2 * (off_n + lane - ln) + 1] : real(0.0);
#pragma unroll
        for (int c0 = (N - 2) / QZ_CH * QZ_CH; c0 >= 0; c0 -= QZ_CH) {
            if (c0 + QZ_CH - 1 < l || c0 >= m) continue;
#pragma unroll
            for (int i = c0 + QZ_CH - 1; i >= c0; --i) {
                if (i > N - 2) continue;
                const real c = rc[i], s = rs[i];
                const real zi = z[i], zi1 = z[i + 1];
                z[i + 1] = fma(s, zi, c * zi1);
                z[i] = fma(c, zi, -(s * zi1));
            }
        }
        off = off_n;
        l = ln;
        m = mn;
    }
    const int* rk = reinterpret_cast<const int*>(lg + QG_RANK);
    const real* lam = reinterpret_cast<const real*>(lg + QG_LAM);
    const int rank = (lane < N) ? rk[lane] : N;
    if (lane < N) Wb[lane] = lam[lane];
    store_z(Zw + (size_t)b * N * N, z, rank, info, b, hdr[2]);
}

// QL without a log: the recurrence inline, every lane redundantly, with the
// rows of Z in registers (a kernel of its own, so that the log-driven kernel
// above keeps its small register footprint). With all == 0 it solves the
// matrices whose log overflowed; with all != 0 every matrix (the host's choice
// for small batches, where the recorder's one-wave latency would dominate;
// the log is then not read).
extern "C" __global__ void __launch_bounds__(GENEIG_NT)
geneig_ql_full(real* __restrict__ W, const real* __restrict__ E, real* __restrict__ Zw,
               int* __restrict__ info, const char* __restrict__ qlog, long long log_stride, int batch, int all) {
    using namespace gk;
    const int b = blockIdx.x * NW + wave_id();
    if (b >= batch) return;
    if (info[b] != 0) return;
    if (!all && !reinterpret_cast<const int*>(qlog + (size_t)b * (size_t)log_stride)[1]) return;
    const int lane = lane_id();
    real* Wb = W + (size_t)b * N;
    const real* Eb = E + (size_t)b * N;
    real z[N];
#pragma unroll
    for (int j = 0; j < N; ++j) z[j] = (lane == j) ? 1.0 : 0.0;
    real dreg = (lane < N) ? Wb[lane] : 0.0;
    real ereg = (lane < N - 1) ? Eb[lane] : 0.0;
    const int qi = ql_wave(dreg, ereg, z);
    const int rank = ascending_rank(lane < N ? dreg : __builtin_inf());
    if (lane < N) Wb[rank] = dreg;
    store_z(Zw + (size_t)b * N * N, z, rank, info, b, qi);
}


// ---------------------------------------------------------------------------
// Mixed precision (FP32 tridiagonalisation and eigenvectors, FP64 Cholesky
// and refinement). Per matrix:
//   geneig_lg_chol    (FP64 module): S = L L^H and W = L^-1 in FP64 by 16-column
//                     panels with the products on the matrix cores; W to Wd
//                     (its lower triangle packed by columns, or in full with
//                     the upper zero where it did not fit LDS). A non-positive
//                     pivot ends the matrix here with zhegv's INFO, as in the
//                     other modes.
//   geneig_mx_formm   (FP64 module): M = W H W^H on the matrix cores; its
//                     lower triangle packed to Mp (complex128).
//   geneig_mx_front   (FP32 module): Householder tridiagonalisation of M (Mp
//                     rounded to complex64, Hermitian from its lower
//                     triangle); reflectors and tau to U32, T to W32 / E32.
//   geneig_split_ql   (FP32 module): QL with FP32 unit roundoff, Z to Z32.
//   geneig_mx_back    (FP32 module): V = Q Z into U32.
//   geneig_mx_refine  (FP64 module): Ogita-Aishima refinement of (M, V) in
//                     the Cholesky frame on the matrix cores (M from Mp),
//                     Rayleigh-Ritz on clusters, then U = L^-H V by back
//                     substitution with L (packed); W, U out.
// The FP64 split mode runs the same stages in the FP64 module, without the
// refinement: geneig_lg_chol, geneig_mx_formm, geneig_mx_front (T to W and E,
// the stash in U), geneig_ql_record + geneig_split_ql (or geneig_ql_full),
// geneig_mx_back (V = Q Z in U), then geneig_mx_backtri (U = L^-H V in place).
// ---------------------------------------------------------------------------

// Built in both precisions: FP32 for the mixed mode, FP64 for the split mode.
namespace gk {
// Shared memory of the front half: the tridiagonalisation's buffers only
// (M arrives formed, straight into the owning waves' registers).
struct WorkF {
    union {
        struct {
            cx vb[2][N];
            cx wb[NW][N];
            cx part[NW * N];
        } h;
    } u;
    cx tau[N];
    real d[N];
    real e[N];
};

// Stash for the back half: tau on the superdiagonal and the reflector
// entries from the owning waves' registers (column k, rows >= k+2, go to
// (k, i)); the lower triangle is left alone (nothing reads it).
__device__ __forceinline__ void stash_refl_reg(const WorkF& w, const cx (&m)[CPW], cx* __restrict__ Ub) {
    const int lane = lane_id(), wave = wave_id();
    for (int i = threadIdx.x; i < N - 1; i += NT) Ub[(size_t)(i + 1) * N + i] = w.tau[i];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int k = wave + NW * c;
        if (k < N && lane >= k + 2 && lane < N) Ub[(size_t)lane * N + k] = m[c];
    }
}
}  // namespace gk

// M's packed lower triangle in LDS (rounded to complex64 in the FP32 build),
// only until the owning waves hold their columns, so the area is shared with
// the tridiagonalisation's buffers.
namespace gk {
union FrontLds {
    WorkF w;
    cx ms[NPK];
};
// Waves per SIMD the front half is compiled for: five in FP32 (67 VGPRs at
// N = 54; with the packed, shared LDS area (12 KB) the registers, not the
// LDS, bound the workgroups per CU), four in FP64 (114 VGPRs; before the
// FMA-form updates and column skips it needed 154, three waves per SIMD).
constexpr int FRONT_WAVES_PER_EU = sizeof(real) == 4 ? 5 : 4;
}  // namespace gk

extern "C" __global__ void __launch_bounds__(GENEIG_NT) __attribute__((amdgpu_waves_per_eu(gk::FRONT_WAVES_PER_EU)))
geneig_mx_front(const gk::cd* __restrict__ Mp, real* __restrict__ W32, cx* __restrict__ U32,
                const int* __restrict__ info, real* __restrict__ E32, int batch) {
    using namespace gk;
    __shared__ FrontLds lds;
    WorkF& w = lds.w;
    cx* ms = lds.ms;
    const int b = blockIdx.x;
    if (b >= batch || info[b] != 0) return;
    const size_t off = (size_t)b * N * N;
    const int lane = lane_id(), wave = wave_id();
    // M's lower triangle (complex128, packed by columns; one coalesced column
    // per wave and step) into LDS; each owned column is then Hermitian from
    // it: M(i, j) below the diagonal, conj(M(j, i)) above, a real diagonal.
    {
        const cd* Mb = Mp + (size_t)b * NPK;
        for (int j = wave; j < N; j += NW) {
            const int i = j + lane;
            if (i < N) {
                const cx v = to_cx(Mb[pk(j) + lane]);
                ms[pk(j) + lane] = (i == j) ? cm(v.x, real(0.0)) : v;
            }
        }
        __syncthreads();
    }
    cx m[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int j = wave + NW * c;
        const bool in = lane < N && j < N;
        const bool lo = lane >= j;
        const cx v = ms[in ? (lo ? pk(j) + lane - j : pk(lane) + j - lane) : 0];
        m[c] = !in ? cm(real(0.0), real(0.0)) : (lo ? v : cconj(v));
    }
    __syncthreads();    // ms is read; its area now holds w
    hetrd_reg(w, m);
    stash_refl_reg(w, m, U32 + off);
    if (threadIdx.x < N) {
        W32[(size_t)b * N + threadIdx.x] = w.d[threadIdx.x];
        E32[(size_t)b * N + threadIdx.x] = w.e[threadIdx.x];
    }
}

// V = Q Z on the matrix cores (back_wy; the back half without the W^H
// factor, which the FP64 refinement, or geneig_mx_backtri in the FP64 split
// mode, applies), one workgroup of one wave per 16-column panel per matrix:
// reflectors from the stash in U32 (v_k(i) at row k, column i; tau_k at row
// k, column k + 1), V over it. Four waves per SIMD in FP32 (the compiler
// would otherwise spend registers on three), three in FP64 (whose tiles take
// twice the registers).
namespace gk {
constexpr int BACK_WAVES_PER_EU = sizeof(real) == 4 ? 4 : 3;
}
extern "C" __global__ void __launch_bounds__(gk::BNT) __attribute__((amdgpu_waves_per_eu(gk::BACK_WAVES_PER_EU)))
geneig_mx_back(cx* __restrict__ U32, const real* __restrict__ Zw, const int* __restrict__ info, int batch) {
    using namespace gk;
    __shared__ BackWork bw;
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    const size_t off = (size_t)b * N * N;
    cx* Ub = U32 + off;
    const real* Zb = Zw + off;
    back_wy<true>(
        bw, [&](int k, int row) { return Ub[(size_t)row * N + k]; }, [&](int k) { return Ub[(size_t)(k + 1) * N + k]; },
        [&](int row, int col) { return Zb[(size_t)col * N + row]; },
        [&](int row, int col, cx v) { Ub[(size_t)col * N + row] = v; });
}


#ifndef GENEIG_FP32

namespace gk {
// Complex FP64 GEMMs of the refinement on the matrix cores:
// v_mfma_f64_16x16x4f64 takes A(i, k) from lane i + 16 k and B(k, j) from lane
// j + 16 k (i, j < 16, k < 4) and leaves D(i, j) in lane j + 16 (i % 4),
// element i / 4. A 16 x 16 complex tile is a pair of such accumulators; a wave
// owns one 16-column block of the result (all row tiles), so reductions over
// a column stay inside the wave.
constexpr int NB = (N + 15) / 16;          // 16-blocks per dimension
constexpr int REF_NT = 64 * NB;            // one wave per column block
constexpr int KEND = (N + 3) / 4 * 4;      // k-steps that touch rows/columns < N
constexpr int LDR = N | 1;
// Row tiles per accumulation group: four, so that each B operand is loaded
// once per k-step for all of them (two: M = W H W^H 2.87 ms at N = 54, four
// 2.73 ms).
constexpr int RGR = NB < 4 ? NB : 4;

// Element (i, k) of an N x N column-major matrix (leading dimension ld), zero
// outside; the address is clamped so the load is branch-free.
__device__ __forceinline__ cx elem(const cx* M, int ld, int i, int k) {
    const bool in = i < N && k < N;
    const cx v = M[(in ? k : 0) * ld + (in ? i : 0)];
    return in ? v : cm(0.0, 0.0);
}

// acc[rt] (row tile rt, column block cb) = A B, A(i, k) = fa(i, k),
// B(k, j) = fb(k, j), k < ke (the fetchers return zero outside N). With
// UPPER_A, A is upper triangular: row tile rt only takes k >= 16 rt. Only
// the groups of row tiles that meet [rlo, rhi] (wave-uniform) are formed;
// the other tiles are left as they are.
template <bool UPPER_A = false, class FA, class FB>
__device__ __forceinline__ void cgemm(ctile (&acc)[NB], FA fa, FB fb, int cb, int ke = KEND, int rlo = 0,
                                      int rhi = NB - 1) {
    const int l = lane_id();
    const int c16 = l & 15, k4 = l >> 4;
    // Row tiles in groups of RGR, so that only RGR 3M accumulators are live.
#pragma unroll
    for (int g0 = 0; g0 < NB; g0 += RGR) {
        if (g0 + RGR - 1 < rlo || g0 > rhi) continue;
        ctile3 a3[RGR];
#pragma unroll
        for (int g = 0; g < RGR; ++g) c3_zero(a3[g]);
#pragma unroll 2
        for (int k0 = 0; k0 < ke; k0 += 4) {
            const cx b = fb(k0 + k4, cb * 16 + c16);
            const real bs = b.x + b.y;
#pragma unroll
            for (int g = 0; g < RGR; ++g) {
                const int rt = g0 + g;
                if (rt < NB && (!UPPER_A || k0 >= 16 * rt)) cmfma3(a3[g], fa(rt * 16 + c16, k0 + k4), b, bs);
            }
        }
#pragma unroll
        for (int g = 0; g < RGR; ++g)
            if (g0 + g < NB) acc[g0 + g] = c3_final(a3[g]);
    }
}

// acc[rt] = A Bt, where the right operand Bt is this wave's own column block
// held in registers in the result layout: the result layout of row tile q is
// the B-operand layout of the k-steps 16 q .. 16 q + 15 (element v of tile q
// = rows 16 q + 4 v .. 16 q + 4 v + 3), so no LDS round trip is needed.
// With LOWER_A, A is lower triangular (row tile rt only takes k-tiles q <= rt)
// and only the row tiles rt >= rt_min are wanted.
template <bool LOWER_A = false, class FA>
__device__ __forceinline__ void cgemm_reg(ctile (&acc)[NB], FA fa, const ctile (&Bt)[NB], int rt_min = 0) {
    const int l = lane_id();
    const int c16 = l & 15, k4 = l >> 4;
#pragma unroll
    for (int g0 = 0; g0 < NB; g0 += RGR) {
        ctile3 a3[RGR];
#pragma unroll
        for (int g = 0; g < RGR; ++g) c3_zero(a3[g]);
#pragma unroll
        for (int q = 0; q < NB; ++q) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (16 * q + 4 * v >= KEND) continue;    // k-steps entirely beyond N
                const cx b = tile_get(Bt[q], v);
                const real bs = b.x + b.y;
                const int k = 16 * q + 4 * v + k4;
#pragma unroll
                for (int g = 0; g < RGR; ++g) {
                    const int rt = g0 + g;
                    if (rt >= NB || (LOWER_A && (q > rt || rt < rt_min))) continue;
                    cmfma3(a3[g], fa(rt * 16 + c16, k), b, bs);
                }
            }
            __builtin_amdgcn_sched_barrier(0);      // bounds the operand loads in flight
        }
#pragma unroll
        for (int g = 0; g < RGR; ++g)
            if (g0 + g < NB) acc[g0 + g] = c3_final(a3[g]);
    }
}

// The refinement's two products whose results are corrections (Y = V^H R
// with the residual R = M V - V diag(lambda), and V E) on the FP32 matrix
// cores (v_mfma_f32_16x16x4f32, twice the FP64 rate), operands rounded to
// FP32 and sums in FP32: R and E are small, so the rounding is relative to
// them (about 6e-8 of a correction of 1e-6 in the first step and far less
// in the second), below what the step's FP64 parts leave. M V, the
// Rayleigh quotients, the residual itself, the Gram entries and V + V E are
// FP64. FP32 result layout: D(i, j) in lane j + 16 (i / 4), element i % 4.
template <bool B>
struct BoolC {
    static constexpr bool value = B;
};
typedef float f4 __attribute__((ext_vector_type(4)));
struct ftile {
    f4 re, im;
};
struct ftile3 {
    f4 p, q, s;
};
__device__ __forceinline__ void cmfma3f(ftile3& c, float ax, float ay, float bx, float by, float bs) {
    c.p = __builtin_amdgcn_mfma_f32_16x16x4f32(ax, bx, c.p, 0, 0, 0);
    c.q = __builtin_amdgcn_mfma_f32_16x16x4f32(ay, by, c.q, 0, 0, 0);
    c.s = __builtin_amdgcn_mfma_f32_16x16x4f32(ax + ay, bs, c.s, 0, 0, 0);
}
// Row of element e of an FP32 result tile held by lane group g.
__device__ __forceinline__ constexpr int frow(int e, int g) { return 4 * g + e; }

// acc[rt] = A Bt in FP32, Bt this wave's column block in registers: an FP64
// result (B64, element v of lane group g = row 16 q + 4 v + g) or an FP32 one
// (element v of lane group g = row 16 q + 4 g + v). Each (q, v) is one k-step
// over the four rows that lane groups 0..3 hold as element v; A is fetched
// for the same rows.
template <bool B64, class FA, class BT>
__device__ __forceinline__ void cgemm_reg_f32(ftile (&acc)[NB], FA fa, const BT (&Bt)[NB]) {
    const int l = lane_id();
    const int c16 = l & 15, k4 = l >> 4;
#pragma unroll
    for (int g0 = 0; g0 < NB; g0 += RGR) {
        ftile3 a3[RGR];
#pragma unroll
        for (int g = 0; g < RGR; ++g) {
            a3[g].p = f4{0.f, 0.f, 0.f, 0.f};
            a3[g].q = a3[g].p;
            a3[g].s = a3[g].p;
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (B64 && 16 * q + 4 * v >= KEND) continue;     // k-steps entirely beyond N
                if (!B64 && 16 * q >= N) continue;
                const int k = 16 * q + (B64 ? 4 * v + k4 : 4 * k4 + v);
                const float bx = (float)Bt[q].re[v], by = (float)Bt[q].im[v];
                const float bs = bx + by;
#pragma unroll
                for (int g = 0; g < RGR; ++g) {
                    const int rt = g0 + g;
                    if (rt >= NB) continue;
                    const cx a = fa(rt * 16 + c16, k);
                    cmfma3f(a3[g], (float)a.x, (float)a.y, bx, by, bs);
                }
            }
            __builtin_amdgcn_sched_barrier(0);      // bounds the operand loads in flight
        }
#pragma unroll
        for (int g = 0; g < RGR; ++g)
            if (g0 + g < NB) {
                acc[g0 + g].re = a3[g].p - a3[g].q;
                acc[g0 + g].im = a3[g].s - a3[g].p - a3[g].q;
            }
    }
}

// Sum over the four lanes that hold one column (lanes c, c + 16, c + 32, c + 48).
__device__ __forceinline__ real col_sum(real v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

struct RefWork {
    cx U[N * LDR];        // eigenvector estimate V
    cx M[NPK];            // M = W H W^H, lower triangle packed by columns
    real lam[N];
    real emax[NB];        // per wave: largest entry of the last step's correction
};

// M(i, k) from its packed lower triangle (Hermitian), zero outside.
__device__ __forceinline__ cx melem(const cx* P, int i, int k) {
    const bool in = i < N && k < N;
    const bool lo = i >= k;
    const int ii = in ? (lo ? i : k) : 0, kk = in ? (lo ? k : i) : 0;
    const cx v = P[pk(kk) + ii - kk];
    return in ? (lo ? v : cm(v.x, -v.y)) : cm(0.0, 0.0);
}

// L(i, k) of a lower-triangular L from its packed lower triangle, zero above
// the diagonal and outside.
__device__ __forceinline__ cx lelem(const cx* P, int i, int k) {
    const bool in = i < N && k < N && i >= k;
    const cx v = P[in ? pk(k) + i - k : 0];
    return in ? v : cm(0.0, 0.0);
}

// Lower triangle of the column-major N x N matrix A (global) into P, packed
// by columns (cooperative, coalesced).
__device__ __forceinline__ void stage_lower_packed(cx* P, const cx* A, int tid, int nt) {
    for (int e = tid; e < N * N; e += nt) {
        const int i = e % N, j = e / N;
        if (i >= j) P[pk(j) + i - j] = A[e];
    }
}

// W = L^-1 of matrix b as geneig_lg_chol leaves it for the mixed mode: its
// lower triangle packed by columns (W_PACKED_OUT), or in full (upper zero).
__device__ __forceinline__ const cx* w_of(const cx* Wd, int b) {
    return Wd + (size_t)b * (W_PACKED_OUT ? (size_t)NPK : (size_t)N * N);
}
// W's lower triangle into P, packed by columns.
__device__ __forceinline__ void stage_w(cx* P, const cx* Wb, int tid, int nt) {
    if constexpr (W_PACKED_OUT) {
        for (int e = tid; e < NPK; e += nt) P[e] = Wb[e];
    } else {
        stage_lower_packed(P, Wb, tid, nt);
    }
}
// W(i, k) straight from Wb (global), zero above the diagonal and outside.
__device__ __forceinline__ cx w_elem(const cx* Wb, int i, int k) {
    if constexpr (W_PACKED_OUT) return lelem(Wb, i, k);
    else return elem(Wb, N, i, k);
}

// Rayleigh-Ritz on columns (a, b) of V for the Hermitian M (held in r.M), by
// one wave: the 2 x 2 projected pencil (V^H M V, V^H V) of the pair, then
// V(:, [a b]) <- V(:, [a b]) X and lam[a], lam[b] <- the Ritz values.
__device__ __forceinline__ void ritz_pair(RefWork& r, int a, int b) {
    const int lane = lane_id();
    const int lr = lane < N ? lane : N - 1;
    const bool in = lane < N;
    cx ha = cm(0.0, 0.0), hb = cm(0.0, 0.0);
#pragma unroll 2
    for (int k = 0; k < N; ++k) {
        const cx mk = melem(r.M, lr, k);
        ha = ha + cmul(mk, r.U[a * LDR + k]);
        hb = hb + cmul(mk, r.U[b * LDR + k]);
    }
    const cx ua = in ? r.U[a * LDR + lr] : cm(0.0, 0.0);
    const cx ub = in ? r.U[b * LDR + lr] : cm(0.0, 0.0);
    const real t11 = wave_sum(in ? cconjmul(ua, ha).x : real(0.0));
    const real t22 = wave_sum(in ? cconjmul(ub, hb).x : real(0.0));
    const cx t12v = cconjmul(ua, hb);
    const cx t12 = cm(wave_sum(in ? t12v.x : real(0.0)), wave_sum(in ? t12v.y : real(0.0)));
    const real g11 = wave_sum(in ? cabs2(ua) : real(0.0));
    const real g22 = wave_sum(in ? cabs2(ub) : real(0.0));
    const cx g12v = cconjmul(ua, ub);
    const cx g12 = cm(wave_sum(in ? g12v.x : real(0.0)), wave_sum(in ? g12v.y : real(0.0)));
    real mu1, mu2;
    cx X[2][2];
    eig2_pencil(t11, t22, t12, g11, g22, g12, mu1, mu2, X);
    if (in) {
        r.U[a * LDR + lane] = cmul(ua, X[0][0]) + cmul(ub, X[1][0]);
        r.U[b * LDR + lane] = cmul(ua, X[0][1]) + cmul(ub, X[1][1]);
    }
    if (lane == 0) {
        r.lam[a] = mu1;
        r.lam[b] = mu2;
    }
    wave_lds_sync();
}
}  // namespace gk

namespace gk {
// geneig_mx_formm stages H and W in LDS when two workgroups still fit a CU.
constexpr bool FORMM_LDS = ((size_t)N * LDR + NPK) * sizeof(cx) <= 80 * 1024;

// The lower triangle of this wave's column block of M (tiles in the MFMA
// result layout, column j in this lane) to Mb, packed by columns.
__device__ __forceinline__ void store_m_packed(cx* __restrict__ Mb, const ctile (&M)[NB], int j, bool jin, int g4) {
#pragma unroll
    for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = rt * 16 + g4 + 4 * v;
            if (i < N && jin && i >= j) Mb[pk(j) + i - j] = tile_get(M[rt], v);
        }
    }
}
}  // namespace gk

// M = W H W^H of the mixed mode, once, on the matrix cores (W = L^-1 lower
// triangular: Z(:, j) = H W^H(:, j) needs k <= j, M(i, j) = W(i, :) Z(:, j)
// needs k <= i). Its lower triangle goes packed by columns to Mp (complex128),
// which both the FP32 front half and the refinement read.
extern "C" __global__ void __launch_bounds__(gk::REF_NT, 2)
geneig_mx_formm(const cx* __restrict__ H, const cx* __restrict__ Wd, const int* __restrict__ info,
                cx* __restrict__ Mp, int batch) {
    using namespace gk;
    const int b = blockIdx.x;
    if (b >= batch || info[b] != 0) return;
    const int l = lane_id(), cb = wave_id();
    const int g4 = l >> 4;
    const int j = cb * 16 + (l & 15);          // this lane's column
    const bool jin = j < N;
    const size_t off = (size_t)b * N * N;
    ctile Z[NB], M[NB];
    const int kz = 16 * (cb + 1) < KEND ? 16 * (cb + 1) : KEND;
    if constexpr (FORMM_LDS) {
        // H and W's lower triangle staged in LDS (coalesced loads; the
        // products then read LDS only).
        __shared__ cx Hs[FORMM_LDS ? N * LDR : 1];
        __shared__ cx Ws[FORMM_LDS ? NPK : 1];
        const cx* Hb = H + off;
        for (int e = threadIdx.x; e < N * N; e += REF_NT) Hs[(e / N) * LDR + e % N] = Hb[e];
        stage_w(Ws, w_of(Wd, b), threadIdx.x, REF_NT);
        __syncthreads();
        auto fW = [&](int i, int k) { return lelem(Ws, i, k); };
        cgemm(Z, [&](int i, int k) { return elem(Hs, LDR, i, k); },
              [&](int k, int jj) { return cconj(fW(jj, k)); }, cb, kz);
        cgemm_reg<true>(M, fW, Z, cb);
    } else {
        // Operands straight from global memory (no LDS, so more workgroups
        // per CU hide the latency).
        const cx* Hb = H + off;
        const cx* Wb = w_of(Wd, b);
        cgemm(Z, [&](int i, int k) { return elem(Hb, N, i, k); },
              [&](int k, int jj) { return cconj(w_elem(Wb, jj, k)); }, cb, kz);
        cgemm_reg<true>(M, [&](int i, int k) { return w_elem(Wb, i, k); }, Z, cb);
    }
    store_m_packed(Mp + (size_t)b * NPK, M, j, jin, g4);
}

// M = W diag(h) W^H for a diagonal H (Sunny's paraunitary metric H = Sigma =
// diag(I, -I), or any real diagonal): one product on the matrix cores,
// M(i, j) = sum_{k <= min(i, j)} W(i, k) h_k conj(W(j, k)), with the right
// operand formed from W's packed lower triangle and h as it is read; H itself
// is never loaded. h is row b * hd_stride of Hd (stride 0: one diagonal for
// the whole batch). Outputs as geneig_mx_formm.
extern "C" __global__ void __launch_bounds__(gk::REF_NT, 2)
geneig_mx_formm_diag(const real* __restrict__ Hd, long long hd_stride, const cx* __restrict__ Wd,
                     const int* __restrict__ info, cx* __restrict__ Mp, int batch) {
    using namespace gk;
    const int b = blockIdx.x;
    if (b >= batch || info[b] != 0) return;
    const int l = lane_id(), cb = wave_id();
    const int g4 = l >> 4;
    const int j = cb * 16 + (l & 15);          // this lane's column
    const bool jin = j < N;
    const size_t off = (size_t)b * N * N;
    __shared__ cx Ws[NPK];
    __shared__ real hs[KEND];
    stage_w(Ws, w_of(Wd, b), threadIdx.x, REF_NT);
    if (threadIdx.x < KEND) hs[threadIdx.x] = threadIdx.x < N ? Hd[(size_t)b * hd_stride + threadIdx.x] : real(0.0);
    __syncthreads();
    // Column block cb needs k <= j < 16 (cb + 1).
    const int kz = 16 * (cb + 1) < KEND ? 16 * (cb + 1) : KEND;
    ctile M[NB];
    cgemm(M, [&](int i, int k) { return lelem(Ws, i, k); },
          [&](int k, int jj) { return cconj(lelem(Ws, jj, k)) * hs[k]; }, cb, kz);
    store_m_packed(Mp + (size_t)b * NPK, M, j, jin, g4);
}

// U = L^-H V of the FP64 split mode, in place (V arrives in U): V and L's
// packed lower triangle staged in LDS, then each wave's 16-column block by
// block back substitution (backsub_lh). Matrices whose Cholesky failed keep
// the outputs it wrote.
extern "C" __global__ void __launch_bounds__(gk::REF_NT, 2)
geneig_mx_backtri(const cx* __restrict__ Lp, cx* __restrict__ U, const int* __restrict__ info, int batch) {
    using namespace gk;
    __shared__ cx Vs[N * LDR];
    __shared__ cx Ls[NPK];
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    const int l = lane_id(), cb = wave_id();
    const int g4 = l >> 4;
    const int j = cb * 16 + (l & 15);          // this lane's column
    cx* Ub = U + (size_t)b * N * N;
    const cx* Lb = Lp + (size_t)b * NPK;
    for (int e = threadIdx.x; e < N * N; e += REF_NT) Vs[(e / N) * LDR + e % N] = Ub[e];
    for (int e = threadIdx.x; e < NPK; e += REF_NT) Ls[e] = Lb[e];
    __syncthreads();
    ctile Uo[NB];
    backsub_lh<NB, KEND>(Uo, [&](int i, int k) { return Ls[pk(k) + i - k]; },
               [&](int i, int jj) { return elem(Vs, LDR, i, jj); }, cb);
#pragma unroll
    for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = rt * 16 + g4 + 4 * v;
            if (i < N && j < N) Ub[(size_t)j * N + i] = tile_get(Uo[rt], v);
        }
    }
}

// Refinement in the Cholesky frame: with W = L^-1 (FP64), M = W H W^H and the
// FP32 eigenvectors V of M, iters Ogita-Aishima steps on the matrix cores
// (T = V^H M V, G = V^H V, lambda_j = T_jj / G_jj, V <- V + V E with E as
// described at its computation), Rayleigh-Ritz on every cluster (a run of
// eigenvalues with adjacent gaps below delta: one exact 2 x 2 step for a
// pair, Jacobi sweeps over all its pairs otherwise), then U = L^-H V by
// back substitution with the packed L (Lp). In step it (from 0),
// delta = 2^it delta_rel max |lambda|. U^H S U = V^H V up to the accuracy of
// the substitution, as in the FP64 modes. Convergence test: the error left
// after a step is of the order of the square of its correction, so a last
// step whose correction E has an entry above `tol` marks the matrix
// info = N (a value zhegv's INFO never takes), with W and U as computed.
extern "C" __global__ void __launch_bounds__(gk::REF_NT, 2)
geneig_mx_refine(const cx* __restrict__ Mp, const cx* __restrict__ Lp, real* __restrict__ W,
                 cx* __restrict__ U, int* __restrict__ info, const gk::cf* __restrict__ V32,
                 const float* __restrict__ W32, int batch, int iters, real delta_rel, real tol) {
    using namespace gk;
    __shared__ RefWork r;
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;     // Cholesky failures keep their outputs
    const int l = lane_id(), cb = wave_id();
    const int g4 = l >> 4;
    const int j = cb * 16 + (l & 15);          // this lane's column
    const bool jin = j < N;
    const size_t off = (size_t)b * N * N;
    const cx* Lb = Lp + (size_t)b * NPK;            // L, packed by columns
    auto fM = [&](int i, int k) { return melem(r.M, i, k); };
    auto fV = [&](int i, int k) { return elem(r.U, LDR, i, k); };
    auto fVh = [&](int i, int k) { return cconj(elem(r.U, LDR, k, i)); };
    {
        const cx* Mb = Mp + (size_t)b * NPK;
        for (int e = threadIdx.x; e < NPK; e += REF_NT) r.M[e] = Mb[e];
        for (int e = threadIdx.x; e < N * N; e += REF_NT) r.U[(e / N) * LDR + e % N] = to_cx(V32[off + e]);
        if (threadIdx.x < N) r.lam[threadIdx.x] = (real)W32[(size_t)b * N + threadIdx.x];
        __syncthreads();
    }
    real delta = real(0.0);
    real emax = real(0.0);                       // this lane's largest |E| entry, last step
    // One step; the last (LAST) and the earlier ones are separate
    // instances, so that each is register-allocated on its own.
    auto step = [&](int it, auto last_c) {
        constexpr bool LAST = decltype(last_c)::value;
        ctile X[NB];
        cgemm(X, fM, fV, cb);                    // X = M V
        // Rayleigh quotients of this wave's columns, lambda_j = v_j^H x_j /
        // v_j^H v_j (the four lanes of column j hold a quarter of its rows each).
        real tj = real(0.0), gj = real(0.0);
#pragma unroll
        for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = rt * 16 + g4 + 4 * v;
                const cx vv = elem(r.U, LDR, i, j);
                tj += cconjmul(vv, tile_get(X[rt], v)).x;
                gj += cabs2(vv);
            }
        }
        tj = col_sum(tj);
        gj = col_sum(gj);
        const real lj = jin ? tj / gj : real(0.0);
        if (jin && g4 == 0) r.lam[j] = lj;
        __syncthreads();
        // The threshold doubles every step, so a pair treated as close once
        // stays close (its vectors are then mixed) and ends in Rayleigh-Ritz.
        delta = delta_rel * (real)(1 << it) * wave_max(l < N ? fabs(r.lam[l < N ? l : 0]) : real(0.0));
        // Residual R = M V - V diag(lambda) in place of X, then Y = V^H R.
#pragma unroll
        for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = rt * 16 + g4 + 4 * v;
                const cx rr = tile_get(X[rt], v) - elem(r.U, LDR, i, j) * lj;
                X[rt].re[v] = rr.x;
                X[rt].im[v] = rr.y;
            }
        }
        // Ogita-Aishima's correction E_ij = (T_ij - lambda_j G_ij) /
        // (lambda_j - lambda_i) with T = V^H M V and G = V^H V is
        // Y_ij / (lambda_j - lambda_i): for separated pairs only Y is needed.
        // The diagonal takes E_jj = (1 - G_jj) / 2 and pairs closer than
        // delta E_ij = -G_ij / 2, each G_ij of those a dot product of two
        // columns (the pairs are few: found per column against every
        // eigenvalue, whatever their order). Yt holds Y on entry and E on
        // return, element v of lane group g at row rowf(v, g) of its tile.
        unsigned long long pm = 0;               // close partners of column j (lanes g4 == 0)
        if (jin && g4 == 0) {
            for (int i = 0; i < N; ++i)
                if (i != j && fabs(r.lam[i] - lj) < delta) pm |= 1ull << i;
        }
        auto correction = [&](auto& Yt, auto rowf) {
#pragma unroll
            for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = rt * 16 + rowf(v, g4);
                    const real li = r.lam[i < N ? i : 0];
                    const real gap = lj - li;
                    cx e = cm((real)Yt[rt].re[v], (real)Yt[rt].im[v]) * rcp_nr(gap);
                    if (i == j) e = cm(real(0.5) * (real(1.0) - gj), 0.0);
                    else if (fabs(gap) < delta) e = cm(0.0, 0.0);    // set below
                    if (i >= N || !jin) e = cm(0.0, 0.0);
                    Yt[rt].re[v] = e.x;
                    Yt[rt].im[v] = e.y;
                }
            }
            for (;;) {
                const unsigned long long any = __ballot(pm != 0);
                if (any == 0) break;
                const int src = (int)__builtin_ctzll(any);
                const int jj = __builtin_amdgcn_readlane(j, src);
                const unsigned lo = __builtin_amdgcn_readlane((unsigned)pm, src);
                const unsigned hi = __builtin_amdgcn_readlane((unsigned)(pm >> 32), src);
                const int ii = lo ? (int)__builtin_ctz(lo) : 32 + (int)__builtin_ctz(hi);
                if (l == src) pm &= pm - 1;
                cx g = (l < N) ? cconjmul(r.U[ii * LDR + l], r.U[jj * LDR + l]) : cm(0.0, 0.0);
                g = wave_sum_cx(g) * real(-0.5);
#pragma unroll
                for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        if (j == jj && rt * 16 + rowf(v, g4) == ii) {
                            Yt[rt].re[v] = g.x;
                            Yt[rt].im[v] = g.y;
                        }
                    }
                }
            }
        };
        // The earlier steps take Y and V E on the FP32 matrix cores (their
        // rounding is removed by the next step); the last one is FP64, since
        // near the closeness threshold the FP32 rounding of Y over a gap of
        // delta would stay in V (S-orthogonality 14.6 -> 251 N eps at N = 54).
        if constexpr (!LAST) {
            ftile Y[NB];
            cgemm_reg_f32<true>(Y, fVh, X);
            correction(Y, [](int v, int g) { return frow(v, g); });
            ftile D[NB];
            cgemm_reg_f32<false>(D, fV, Y);      // V E
#pragma unroll
            for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = rt * 16 + frow(v, g4);
                    const cx u = elem(r.U, LDR, i, j) + cm((real)D[rt].re[v], (real)D[rt].im[v]);
                    X[rt].re[v] = u.x;
                    X[rt].im[v] = u.y;
                }
            }
        } else {
            ctile Y[NB];
            cgemm_reg(Y, fVh, X);
            correction(Y, [](int v, int g) { return 4 * v + g; });
#pragma unroll
            for (int rt = 0; rt < NB; ++rt)
#pragma unroll
                for (int v = 0; v < 4; ++v) emax = fmax(emax, fmax(fabs(Y[rt].re[v]), fabs(Y[rt].im[v])));
            cgemm_reg(X, fV, Y);                 // V E
#pragma unroll
            for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = rt * 16 + g4 + 4 * v;
                    const cx u = elem(r.U, LDR, i, j) + tile_get(X[rt], v);
                    X[rt].re[v] = u.x;
                    X[rt].im[v] = u.y;
                }
            }
        }
        __syncthreads();     // every wave is done reading V
        constexpr bool f32l = !LAST;             // X in the FP32 result layout
#pragma unroll
        for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = rt * 16 + (f32l ? frow(v, g4) : 4 * v + g4);
                if (i < N && jin) r.U[j * LDR + i] = tile_get(X[rt], v);
            }
        }
        __syncthreads();
    };
    for (int it = 0; it + 1 < iters; ++it) step(it, BoolC<false>{});
    if (iters > 0) step(iters - 1, BoolC<true>{});
    // Clusters: runs [c0, c1] with adjacent gaps below delta. Their bounds are
    // fixed first (bit i: gap i to i + 1 is close; every wave computes the same
    // mask before any wave changes lam), then cluster number nc goes to wave
    // nc mod NB (clusters are disjoint).
    unsigned long long close = 0;
    if (iters > 0) close = __ballot(l < N - 1 && fabs(r.lam[l < N - 1 ? l + 1 : 0] - r.lam[l < N - 1 ? l : 0]) < delta);
    __syncthreads();
    if (iters > 0) {
        int c0 = 0, nc = 0;
        while (c0 < N - 1) {
            int c1 = c0;
            while (c1 + 1 < N && ((close >> c1) & 1ull)) ++c1;
            if (c1 > c0 && (nc++ % NB) == cb) {
                const int sweeps = (c1 == c0 + 1) ? 1 : 4;
                for (int sw = 0; sw < sweeps; ++sw)
                    for (int a = c0; a < c1; ++a)
                        for (int bb = a + 1; bb <= c1; ++bb) ritz_pair(r, a, bb);
                // Keep the run ascending.
                for (int a = c0 + 1; a <= c1; ++a)
                    for (int q = a; q > c0 && r.lam[q] < r.lam[q - 1]; --q) {
                        const real tl = r.lam[q];
                        if (l < N) {
                            const cx t0 = r.U[q * LDR + l];
                            r.U[q * LDR + l] = r.U[(q - 1) * LDR + l];
                            r.U[(q - 1) * LDR + l] = t0;
                        }
                        wave_lds_sync();
                        if (l == 0) {
                            r.lam[q] = r.lam[q - 1];
                            r.lam[q - 1] = tl;
                        }
                        wave_lds_sync();
                    }
            }
            c0 = c1 + 1;
        }
    }
    emax = wave_max(emax);
    if (l == 0) r.emax[cb] = emax;
    __syncthreads();
    if (threadIdx.x == 0 && iters > 0 && info[b] == 0) {
        real m = r.emax[0];
#pragma unroll
        for (int q = 1; q < NB; ++q) m = fmax(m, r.emax[q]);
        if (!(m <= tol)) info[b] = N;
    }
    // U = L^-H V by block back substitution, L's packed lower triangle staged
    // over M (no longer needed).
    for (int e = threadIdx.x; e < NPK; e += REF_NT) r.M[e] = Lb[e];
    __syncthreads();
    ctile Uo[NB];
    backsub_lh<NB, KEND>(Uo, [&](int i, int k) { return r.M[pk(k) + i - k]; }, fV, cb);
    cx* Ub = U + off;
#pragma unroll
    for (int rt = 0; rt < NB; ++rt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = rt * 16 + g4 + 4 * v;
            if (i < N && jin) Ub[(size_t)j * N + i] = tile_get(Uo[rt], v);
        }
    }
    if (threadIdx.x < N) W[(size_t)b * N + threadIdx.x] = r.lam[threadIdx.x];
}
#endif  // !GENEIG_FP32

#endif  // GENEIG_N <= 64

#ifndef GENEIG_FP32
// A diagonal H written out in full (complex128, column-major) from its real
// diagonal, row b * hd_stride of Hd, for the paths that read H itself.
extern "C" __global__ void __launch_bounds__(256)
geneig_expand_hdiag(const double* __restrict__ Hd, long long hd_stride, gk::cx* __restrict__ H, int batch) {
    using namespace gk;
    const int b = blockIdx.x;
    if (b >= batch) return;
    cx* Hb = H + (size_t)b * N * N;
    const double* h = Hd + (size_t)b * hd_stride;
    for (int e = threadIdx.x; e < N * N; e += 256) {
        const int i = e % N, j = e / N;
        Hb[e] = cm(i == j ? h[i] : 0.0, 0.0);
    }
}
#endif  // !GENEIG_FP32
