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
// numbers as (re, im) double pairs. In LDS the two N x N complex work
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
static_assert(N >= 1 && N <= 64, "LDS-resident path holds one row per lane");

struct __attribute__((aligned(16))) c64 {
    double x, y;
};

__device__ __forceinline__ c64 cm(double a, double b) { c64 r; r.x = a; r.y = b; return r; }
__device__ __forceinline__ c64 operator+(c64 a, c64 b) { return cm(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ c64 operator-(c64 a, c64 b) { return cm(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ c64 operator*(c64 a, double s) { return cm(a.x * s, a.y * s); }
__device__ __forceinline__ c64 cmul(c64 a, c64 b) {
    return cm(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ c64 cmulc(c64 a, c64 b) {
    return cm(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a) * b
__device__ __forceinline__ c64 cconjmul(c64 a, c64 b) {
    return cm(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ c64 cconj(c64 a) { return cm(a.x, -a.y); }
__device__ __forceinline__ double cabs2(c64 a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// Value of `v` in lane `src` (src wave-uniform), via v_readlane.
__device__ __forceinline__ double bcast(double v, int src) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ c64 bcast(c64 v, int src) { return cm(bcast(v.x, src), bcast(v.y, src)); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Shared-memory work area of one matrix.
struct Work {
    c64 A[N * LDA];        // S -> L (lower); Householder vectors (strict upper, row k = v_k)
    c64 B[N * LDA];        // H -> M -> T; then Z (real); then V
    c64 tau[N];
    c64 vbuf[N];
    c64 wbuf[N];
    double invL[N];        // 1 / L(k,k)
    double d[N];
    double e[N];
    c64 part[2][NW * N];   // reflector partial sums, double-buffered
    int flag[2];
};

// ---------------------------------------------------------------------------
// Stage 1: Cholesky S = L L^H (lower, right-looking). Returns 0 or k+1 for the
// first pivot k that is not positive (zpotrf's INFO). Each step updates the
// trailing lower triangle with the unscaled pivot column and scales the
// previous pivot column, so no thread reads a value another thread writes in
// the same step; one barrier per step.
// ---------------------------------------------------------------------------
__device__ int potrf(Work& w) {
    const int lane = lane_id(), wave = wave_id();
    double prev_rinv = 0.0;
    for (int k = 0; k < N; ++k) {
        const double dkk = w.A[k * LDA + k].x;
        if (!(dkk > 0.0)) return k + 1;
        const double rinv = 1.0 / sqrt(dkk);
        const double rinv2 = rinv * rinv;
        const c64 aik = (lane < N) ? w.A[k * LDA + lane] : cm(0.0, 0.0);
        if (k > 0 && wave == (k - 1) % NW && lane >= k - 1 && lane < N)
            w.A[(k - 1) * LDA + lane] = w.A[(k - 1) * LDA + lane] * prev_rinv;
        for (int j = k + 1 + wave; j < N; j += NW) {
            const c64 ajk = bcast(aik, j);
            if (lane >= j && lane < N) {
                c64 t = w.A[j * LDA + lane];
                t = t - cmulc(aik, ajk) * rinv2;
                w.A[j * LDA + lane] = t;
            }
        }
        if (threadIdx.x == 0) w.invL[k] = rinv;
        prev_rinv = rinv;
        __syncthreads();
    }
    if (wave == (N - 1) % NW && lane == N - 1)
        w.A[(N - 1) * LDA + N - 1] = w.A[(N - 1) * LDA + N - 1] * prev_rinv;
    __syncthreads();
    return 0;
}

// ---------------------------------------------------------------------------
// Stage 2: M = L^-1 H L^-H. Left solve column by column (each wave owns
// columns, lane = row), then right solve row by row (each wave owns rows,
// lane = column). Pivot values travel by readlane: no barriers inside.
// ---------------------------------------------------------------------------
__device__ void trsm_left(Work& w) {
    const int lane = lane_id(), wave = wave_id();
    c64 b[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int j = wave + NW * c;
        b[c] = (j < N && lane < N) ? w.B[j * LDA + lane] : cm(0.0, 0.0);
    }
#pragma unroll 2
    for (int k = 0; k < N; ++k) {
        const c64 lik = (lane < N) ? w.A[k * LDA + lane] : cm(0.0, 0.0);
        const double inv = w.invL[k];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const c64 xk = bcast(b[c], k) * inv;
            if (lane == k) b[c] = xk;
            else if (lane > k) b[c] = b[c] - cmul(lik, xk);
        }
    }
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int j = wave + NW * c;
        if (j < N && lane < N) w.B[j * LDA + lane] = b[c];
    }
}

__device__ void trsm_right(Work& w) {
    const int lane = lane_id(), wave = wave_id();
    c64 b[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int i = wave + NW * c;
        b[c] = (i < N && lane < N) ? w.B[lane * LDA + i] : cm(0.0, 0.0);
    }
#pragma unroll 2
    for (int k = 0; k < N; ++k) {
        const c64 ljk = (lane < N) ? w.A[k * LDA + lane] : cm(0.0, 0.0);
        const double inv = w.invL[k];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const c64 yk = bcast(b[c], k) * inv;
            if (lane == k) b[c] = yk;
            else if (lane > k) b[c] = b[c] - cmulc(yk, ljk);
        }
    }
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int i = wave + NW * c;
        if (i < N && lane < N) w.B[lane * LDA + i] = b[c];
    }
}

// ---------------------------------------------------------------------------
// Stage 3: Householder tridiagonalisation of the Hermitian M (full storage).
// Reflector k: H_k = I - tau_k v v^H with v(k+1) = 1, v(k+2..) kept in
// B(k+2.., k); reflector generation follows LAPACK zlarfg (beta real,
// beta = -sign(Re alpha) * ||(alpha, x)||). Three barriers per step.
// ---------------------------------------------------------------------------
__device__ void hetrd(Work& w) {
    const int lane = lane_id(), wave = wave_id();
    for (int k = 0; k < N - 1; ++k) {
        const int m0 = k + 1;
        if (wave == 0) {
            const c64 x = (lane >= m0 && lane < N) ? w.B[k * LDA + lane] : cm(0.0, 0.0);
            const c64 alpha = bcast(x, m0);
            const double xn2 = wave_sum(lane > m0 ? cabs2(x) : 0.0);
            c64 t = cm(0.0, 0.0);
            double beta = alpha.x;
            c64 v = cm(0.0, 0.0);
            if (xn2 == 0.0 && alpha.y == 0.0) {
                if (lane == m0) v = cm(1.0, 0.0);
            } else {
                const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
                beta = (alpha.x >= 0.0) ? -nrm : nrm;
                t = cm((beta - alpha.x) / beta, -alpha.y / beta);
                const c64 den = cm(alpha.x - beta, alpha.y);
                const double dd = 1.0 / cabs2(den);
                const c64 scal = cm(den.x * dd, -den.y * dd);   // 1 / (alpha - beta)
                if (lane == m0) v = cm(1.0, 0.0);
                else if (lane > m0) v = cmul(x, scal);
            }
            if (lane > m0 && lane < N) w.B[k * LDA + lane] = v;
            if (lane < N) w.vbuf[lane] = v;
            if (lane == 0) {
                w.e[k] = beta;
                w.d[k] = w.B[k * LDA + k].x;
                w.tau[k] = t;
            }
        }
        __syncthreads();
        const c64 t = w.tau[k];
        if (t.x == 0.0 && t.y == 0.0) continue;
        if (wave == 0) {
            // p_i = t * sum_j M(i,j) v_j, reading column i (M Hermitian).
            c64 acc = cm(0.0, 0.0);
            if (lane >= m0 && lane < N) {
                for (int j = m0; j < N; ++j) {
                    const c64 mji = w.B[lane * LDA + j];
                    acc = acc + cconjmul(mji, w.vbuf[j]);
                }
            }
            const c64 p = cmul(t, acc);
            const c64 v = (lane < N) ? w.vbuf[lane] : cm(0.0, 0.0);
            const c64 pvl = cconjmul(p, v);
            const c64 pv = cm(wave_sum(pvl.x), wave_sum(pvl.y));
            const c64 cc = cmul(t, pv) * -0.5;
            const c64 wv = p + cmul(cc, v);
            if (lane < N) w.wbuf[lane] = wv;
        }
        __syncthreads();
        {
            const c64 vi = (lane < N) ? w.vbuf[lane] : cm(0.0, 0.0);
            const c64 wi = (lane < N) ? w.wbuf[lane] : cm(0.0, 0.0);
            for (int j = m0 + wave; j < N; j += NW) {
                const c64 vj = w.vbuf[j];
                const c64 wj = w.wbuf[j];
                if (lane >= m0 && lane < N) {
                    c64 a = w.B[j * LDA + lane];
                    a = a - cmulc(vi, wj) - cmulc(wi, vj);
                    w.B[j * LDA + lane] = a;
                }
            }
        }
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
__device__ void stash_reflectors(Work& w) {
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
__device__ __forceinline__ double sel_lane(double v, double x, int target) {
    return (lane_id() == target) ? x : v;
}

__device__ __forceinline__ double rsqrt_nr(double t) {
    double y = __builtin_amdgcn_rsq(t);
    const double h = 0.5 * t;
    double r = fma(-(h * y), y, 0.5);
    y = fma(y, r, y);
    r = fma(-(h * y), y, 0.5);
    y = fma(y, r, y);
    return y;
}

__device__ int ql_wave(double& dreg, double& ereg, double (&z)[N]) {
    const int lane = lane_id();
    const double eps = 0x1p-53;
    int l = 0, iter = 0;
    while (l < N) {
        const double dnext = __shfl_down(dreg, 1, 64);
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
            const unsigned long long nz = __ballot(ereg != 0.0 && lane >= l && lane < N - 1);
            return __popcll(nz);
        }
        const double d_l = bcast(dreg, l);
        const double e_l = bcast(ereg, l);
        double g = (bcast(dreg, l + 1) - d_l) / (2.0 * e_l);
        double r = sqrt(g * g + 1.0);
        g = bcast(dreg, m) - d_l + e_l / (g + copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        double d_up = bcast(dreg, m);      // original d[i+1] for the current i
        bool underflow = false;
#pragma unroll
        for (int i = N - 2; i >= 0; --i) {
            if (i < m && i >= l && !underflow) {
                const double e_i = bcast(ereg, i);
                const double d_i = bcast(dreg, i);
                const double f = s * e_i;
                const double b = c * e_i;
                const double t = fma(f, f, g * g);
                if (t == 0.0) {
                    ereg = sel_lane(ereg, 0.0, i + 1);
                    dreg = sel_lane(dreg, d_up - p, i + 1);
                    ereg = sel_lane(ereg, 0.0, m);
                    underflow = true;
                } else {
                    const double y = rsqrt_nr(t);
                    ereg = sel_lane(ereg, t * y, i + 1);
                    s = f * y;
                    c = g * y;
                    g = d_up - p;
                    const double r2 = fma(d_i - g, s, 2.0 * c * b);
                    p = s * r2;
                    dreg = sel_lane(dreg, g + p, i + 1);
                    g = fma(c, r2, -b);
                    const double zi = z[i], zi1 = z[i + 1];
                    z[i + 1] = fma(s, zi, c * zi1);
                    z[i] = fma(c, zi, -(s * zi1));
                    d_up = d_i;
                }
            }
        }
        if (underflow) continue;
        dreg = sel_lane(dreg, d_l - p, l);
        ereg = sel_lane(ereg, g, l);
        ereg = sel_lane(ereg, 0.0, m);
    }
    return 0;
}

// Ascending position of lane j's eigenvalue (ties by index).
__device__ int ascending_rank(double dreg) {
    const int lane = lane_id();
    int rank = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double dk = bcast(dreg, k);
        rank += (dk < dreg || (dk == dreg && k < lane)) ? 1 : 0;
    }
    return rank;
}

// Wave-level eigen-solve of the tridiagonal held in dsrc/esrc (LDS): writes the
// ascending eigenvalues to W_out and leaves Z (columns in ascending order) in
// `z` (lane r = row r) plus its rank permutation in `rank`.
__device__ int tridiag_eig_wave(const double* dsrc, const double* esrc, double* __restrict__ W_out,
                                double (&z)[N], int& rank) {
    const int lane = lane_id();
    double dreg = (lane < N) ? dsrc[lane] : 0.0;
    double ereg = (lane < N - 1) ? esrc[lane] : 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) z[j] = (lane == j) ? 1.0 : 0.0;
    const int info = ql_wave(dreg, ereg, z);
    rank = ascending_rank(lane < N ? dreg : __builtin_inf());
    if (lane < N) W_out[rank] = dreg;
    return info;
}

// Write this wave's Z rows into Zs (real, leading dimension LDZ) with columns
// permuted to ascending eigenvalue order.
__device__ void store_z_sorted(double* Zs, const double (&z)[N], int rank) {
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
__device__ void apply_q(Work& w) {
    const int lane = lane_id(), wave = wave_id();
    const double* Z = reinterpret_cast<const double*>(w.B);
    c64 b[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int i = wave + NW * c;
        b[c] = cm((i < N && lane < N) ? Z[lane * LDZ + i] : 0.0, 0.0);
    }
    int applied = 0;
    for (int k = N - 2; k >= 0; --k) {
        const c64 t = w.tau[k];
        if (t.x == 0.0 && t.y == 0.0) continue;
        const int m0 = k + 1;
        c64 v[CPW];
        c64 acc = cm(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int i = wave + NW * c;
            v[c] = cm(0.0, 0.0);
            if (i == m0) v[c] = cm(1.0, 0.0);
            else if (i > m0 && i < N) v[c] = w.A[i * LDA + k];
            acc = acc + cconjmul(v[c], b[c]);
        }
        c64* part = w.part[applied & 1];
        ++applied;
        if (lane < N) part[wave * N + lane] = acc;
        __syncthreads();
        c64 y = cm(0.0, 0.0);
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
__device__ void backtrsm_store(Work& w, c64* __restrict__ Ub) {
    const int lane = lane_id(), wave = wave_id();
    c64 b[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
        const int j = wave + NW * c;
        b[c] = (j < N && lane < N) ? w.B[j * LDA + lane] : cm(0.0, 0.0);
    }
#pragma unroll 2
    for (int k = N - 1; k >= 0; --k) {
        // L(k, i) for lane i < k: column i, row k.
        const c64 lki = (lane < k) ? w.A[lane * LDA + k] : cm(0.0, 0.0);
        const double inv = w.invL[k];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const c64 uk = bcast(b[c], k) * inv;
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

__device__ void load_pair(Work& w, const c64* __restrict__ Hb, const c64* __restrict__ Sb) {
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        w.A[j * LDA + i] = Sb[e];
        w.B[j * LDA + i] = Hb[e];
    }
    __syncthreads();
}

__device__ void fail_store(int code, double* __restrict__ Wb, c64* __restrict__ Ub,
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
__device__ void stash_factors(const Work& w, c64* __restrict__ Ub) {
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        c64 v;
        if (i > j) v = w.A[j * LDA + i];
        else if (i == j) v = cm(w.A[j * LDA + j].x, w.invL[j]);
        else if (i == j - 1) v = w.tau[i];
        else v = w.B[i * LDA + j];
        Ub[e] = v;
    }
}

__device__ void unstash_factors(Work& w, const c64* __restrict__ Ub) {
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        const c64 v = Ub[e];
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

// Stage data kept per matrix of a workgroup between the front and back halves.
struct Staged {
    double d[NW][N];
    double e[NW][N];
    int status[NW];        // 0 = ok, otherwise the matrix is finished (failed)
    int qinfo[NW];
};

}  // namespace gk

using gk::c64;

// ---------------------------------------------------------------------------
// Fused solve. A workgroup of NW waves takes NW consecutive matrices: it runs
// the dense front half (Cholesky, two-sided solve, tridiagonalisation) for
// each in turn with all waves, stashing the factors in the matrix's own U
// slot; then every wave runs the QL of one matrix at the same time (Z in its
// registers); then the dense back half (V = Q Z, U = L^-H V) for each in turn.
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(GENEIG_NT)
geneig_fused(const c64* __restrict__ H, const c64* __restrict__ S, double* __restrict__ W,
             c64* __restrict__ U, int* __restrict__ info, int batch) {
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
    double z[N];
    int rank = 0;
    {
        const int b = b0 + wave;
        if (b < batch && st.status[wave] == 0) {
            const int qi = tridiag_eig_wave(st.d[wave], st.e[wave], W + (size_t)b * N, z, rank);
            if (lane == 0) st.qinfo[wave] = qi;
        }
    }
    __syncthreads();
    for (int q = 0; q < NW; ++q) {
        const int b = b0 + q;
        if (b >= batch) break;
        if (st.status[q] != 0) continue;
        const size_t off = (size_t)b * N * N;
        if (wave == q) store_z_sorted(reinterpret_cast<double*>(w.B), z, rank);
        unstash_factors(w, U + off);
        __syncthreads();
        apply_q(w);
        __syncthreads();
        backtrsm_store(w, U + off);
        if (threadIdx.x == 0) info[b] = st.qinfo[q];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Split mode (reference: src/hybrid/hybrid_solver.cpp:295-393), three kernels
// over the same device stages:
//   geneig_split_reduce : potrf + two-sided solve; L -> Lw, M -> U (scratch)
//   geneig_split_heev   : tridiagonalise M, QL, sort, V = Q Z -> U, W
//   geneig_split_back   : U = L^-H V
// Lw holds L's lower triangle with 1/L(k,k) in the imaginary part of the
// diagonal, so the back substitution uses the same reciprocal pivots as the
// fused kernel. A matrix whose Cholesky failed carries info > N and is skipped
// by the later kernels. The reduce and back kernels take one matrix per
// workgroup; the heev kernel takes NW, like the fused kernel.
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(GENEIG_NT)
geneig_split_reduce(const c64* __restrict__ H, const c64* __restrict__ S, double* __restrict__ W,
                    c64* __restrict__ U, int* __restrict__ info, c64* __restrict__ Lw, int batch) {
    using namespace gk;
    __shared__ Work w;
    const int b = blockIdx.x;
    if (b >= batch) return;
    const size_t off = (size_t)b * N * N;
    load_pair(w, H + off, S + off);
    const int pinfo = potrf(w);
    if (pinfo) {
        fail_store(N + pinfo, W + (size_t)b * N, U + off, info + b);
        return;
    }
    trsm_left(w);
    __syncthreads();
    trsm_right(w);
    __syncthreads();
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        c64 l = (i >= j) ? w.A[j * LDA + i] : cm(0.0, 0.0);
        if (i == j) l.y = w.invL[j];
        Lw[off + e] = l;
        U[off + e] = w.B[j * LDA + i];
    }
    if (threadIdx.x == 0) info[b] = 0;
}

extern "C" __global__ void __launch_bounds__(GENEIG_NT)
geneig_split_heev(double* __restrict__ W, c64* __restrict__ U, int* __restrict__ info, int batch) {
    using namespace gk;
    __shared__ Work w;
    __shared__ Staged st;
    const int lane = lane_id(), wave = wave_id();
    const int b0 = blockIdx.x * NW;
    for (int q = 0; q < NW; ++q) {
        const int b = b0 + q;
        if (b >= batch) break;
        if (info[b] != 0) {
            if (threadIdx.x == 0) st.status[q] = 1;
            continue;
        }
        const size_t off = (size_t)b * N * N;
        for (int e = threadIdx.x; e < N * N; e += NT) w.B[(e / N) * LDA + (e % N)] = U[off + e];
        __syncthreads();
        hetrd(w);
        // Reflectors and tau into U's strict upper part (the lower part is unused).
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int i = e % N, j = e / N;
            if (i == j - 1) U[off + e] = w.tau[i];
            else if (i < j - 1) U[off + e] = w.B[i * LDA + j];
        }
        if (threadIdx.x < N) {
            st.d[q][threadIdx.x] = w.d[threadIdx.x];
            st.e[q][threadIdx.x] = w.e[threadIdx.x];
        }
        if (threadIdx.x == 0) st.status[q] = 0;
        __syncthreads();
    }
    double z[N];
    int rank = 0;
    {
        const int b = b0 + wave;
        if (b < batch && st.status[wave] == 0) {
            const int qi = tridiag_eig_wave(st.d[wave], st.e[wave], W + (size_t)b * N, z, rank);
            if (lane == 0) st.qinfo[wave] = qi;
        }
    }
    __syncthreads();
    for (int q = 0; q < NW; ++q) {
        const int b = b0 + q;
        if (b >= batch) break;
        if (st.status[q] != 0) continue;
        const size_t off = (size_t)b * N * N;
        if (wave == q) store_z_sorted(reinterpret_cast<double*>(w.B), z, rank);
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int i = e % N, j = e / N;
            if (i < j - 1) w.A[j * LDA + i] = U[off + e];
            else if (i == j - 1) w.tau[i] = U[off + e];
        }
        if (threadIdx.x == 0) w.tau[N - 1] = cm(0.0, 0.0);
        __syncthreads();
        apply_q(w);
        __syncthreads();
        for (int e = threadIdx.x; e < N * N; e += NT) U[off + e] = w.B[(e / N) * LDA + (e % N)];
        if (threadIdx.x == 0) info[b] = st.qinfo[q];
        __syncthreads();
    }
}

extern "C" __global__ void __launch_bounds__(GENEIG_NT)
geneig_split_back(c64* __restrict__ U, const int* __restrict__ info, const c64* __restrict__ Lw,
                  int batch) {
    using namespace gk;
    __shared__ Work w;
    const int b = blockIdx.x;
    if (b >= batch) return;
    if (info[b] > N) return;
    const size_t off = (size_t)b * N * N;
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e % N, j = e / N;
        const c64 l = Lw[off + e];
        if (i == j) {
            w.A[j * LDA + i] = cm(l.x, 0.0);
            w.invL[j] = l.y;
        } else {
            w.A[j * LDA + i] = l;
        }
        w.B[j * LDA + i] = U[off + e];
    }
    __syncthreads();
    backtrsm_store(w, U + off);
}
