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
// Stage 4: implicit QL with Wilkinson-type shift on (d, e), rotations applied
// to Z (real, in B's storage). Run by wave 0: every lane computes the same
// scalar recurrence, lane r owns row r of Z and carries the column value that
// the next rotation of the sweep touches. Then selection sort ascending.
// Method: Bowdler, Martin, Reinsch & Wilkinson, "The QR and QL algorithms for
// symmetric matrices" (Handbook II/3). Returns zheev-style INFO (number of
// off-diagonals left unconverged) and writes sorted eigenvalues to W_out.
// ---------------------------------------------------------------------------
__device__ int tql(Work& w, double* __restrict__ W_out) {
    const int lane = lane_id();
    double* Z = reinterpret_cast<double*>(w.B);
    double* d = w.d;
    double* e = w.e;
    const bool own = lane < N;
    const double eps = 0x1p-53;
    int info = 0;
    for (int l = 0; l < N && info == 0; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < N - 1; ++m) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (iter++ == QL_MAXIT) {
                int cnt = 0;
                for (int i = l; i < N - 1; ++i) cnt += (e[i] != 0.0);
                info = cnt;
                break;
            }
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = sqrt(g * g + 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            double zc = own ? Z[m * LDZ + lane] : 0.0;
            bool underflow = false;
            double dn = d[m - 1], en = e[m - 1];
            for (int i = m - 1; i >= l; --i) {
                const double di = dn, ei = en;
                if (i > l) { dn = d[i - 1]; en = e[i - 1]; }
                const double f = s * ei;
                const double b = c * ei;
                r = sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    if (own) Z[(i + 1) * LDZ + lane] = zc;
                    underflow = true;
                    break;
                }
                const double rinv = 1.0 / r;
                s = f * rinv;
                c = g * rinv;
                g = d[i + 1] - p;
                r = (di - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                const double zi = own ? Z[i * LDZ + lane] : 0.0;
                if (own) Z[(i + 1) * LDZ + lane] = s * zi + c * zc;
                zc = c * zi - s * zc;
            }
            if (underflow) continue;
            if (own) Z[l * LDZ + lane] = zc;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    // Selection sort, ascending, first minimum wins (lane j holds d[j]).
    double dj = own ? d[lane] : __builtin_inf();
    for (int i = 0; i < N - 1; ++i) {
        double best = (lane >= i) ? dj : __builtin_inf();
        int bi = lane;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ob < best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        const int k = bi;
        if (k != i) {
            const double di = bcast(dj, i);
            if (lane == i) dj = best;
            else if (lane == k) dj = di;
            if (own) {
                const double zi = Z[i * LDZ + lane];
                const double zk = Z[k * LDZ + lane];
                Z[i * LDZ + lane] = zk;
                Z[k * LDZ + lane] = zi;
            }
        }
    }
    if (own) W_out[lane] = dj;
    return info;
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

}  // namespace gk

using gk::c64;

// Fused solve, one workgroup per matrix.
extern "C" __global__ void __launch_bounds__(GENEIG_NT)
geneig_fused(const c64* __restrict__ H, const c64* __restrict__ S, double* __restrict__ W,
             c64* __restrict__ U, int* __restrict__ info, int batch) {
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
    hetrd(w);
    stash_reflectors(w);
    {
        double* Z = reinterpret_cast<double*>(w.B);
        for (int e = threadIdx.x; e < N * LDZ; e += NT) Z[e] = ((e % LDZ) == (e / LDZ)) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (wave_id() == 0) {
        const int qinfo = tql(w, W + (size_t)b * N);
        if (lane_id() == 0) w.flag[0] = qinfo;
    }
    __syncthreads();
    apply_q(w);
    __syncthreads();
    backtrsm_store(w, U + off);
    if (threadIdx.x == 0) info[b] = w.flag[0];
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
// by the later kernels.
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
    const int b = blockIdx.x;
    if (b >= batch) return;
    if (info[b] != 0) return;
    const size_t off = (size_t)b * N * N;
    for (int e = threadIdx.x; e < N * N; e += NT) w.B[(e / N) * LDA + (e % N)] = U[off + e];
    __syncthreads();
    hetrd(w);
    stash_reflectors(w);
    {
        double* Z = reinterpret_cast<double*>(w.B);
        for (int e = threadIdx.x; e < N * LDZ; e += NT) Z[e] = ((e % LDZ) == (e / LDZ)) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (wave_id() == 0) {
        const int qinfo = tql(w, W + (size_t)b * N);
        if (lane_id() == 0) w.flag[0] = qinfo;
    }
    __syncthreads();
    apply_q(w);
    __syncthreads();
    for (int e = threadIdx.x; e < N * N; e += NT) U[off + e] = w.B[(e / N) * LDA + (e % N)];
    if (threadIdx.x == 0) info[b] = w.flag[0];
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
