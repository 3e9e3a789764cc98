// Orders 64 < N <= 192: the HBM-staged pipeline. Included by
// geneig_kernels.hip for every order: the LDS-resident kernels (one row per
// lane, a whole pair of matrices in LDS) stop at N = 64, so above that every
// kernel below is built; at N <= 64 only geneig_lg_chol is, for the mixed
// mode's FP64 Cholesky and W = L^-1.
//
// Above that an N x N complex128 matrix alone is 66 KB (N = 64) to 590 KB
// (N = 192), so every stage is its own kernel, one workgroup of 256 threads
// per matrix, and the stages hand over through a per-matrix workspace in
// global memory (layout below):
//   geneig_lg_chol    FP64  S = L L^H by 16-column panels (the update of a
//                           panel by the columns left of it on the matrix
//                           cores), then W = L^-1 in place by 16-row blocks.
//   geneig_lg_formm   FP64  M = W H W^H on the matrix cores, stored as a full
//                           (exactly Hermitian) matrix.
//   geneig_lg_hetrd   P     Householder tridiagonalisation with the lower
//                           triangle of M packed in LDS. In FP64 above N = 131
//                           the triangle does not fit: the first K0 columns are
//                           reduced in place in global memory, then the
//                           trailing triangle moves to LDS.
//   geneig_lg_ql      P     implicit QL: each wave runs the scalar recurrence
//                           (d, e spread over its lanes), records the rotations
//                           of a sweep in LDS and applies them to its own rows
//                           of Z (in LDS, or in global memory for the rows that
//                           do not fit), then sorts.
//   geneig_lg_back    P     V = Q Z, blocks of 8 columns per wave in registers.
//   geneig_lg_backtri FP64  U = L^-H V by block back substitution (the
//                           off-diagonal blocks on the matrix cores).
// P is the module precision: FP64 for the FP64 modes; FP32 in the mixed mode,
// which then refines (M, V) in FP64 (geneig_lg_refine) instead of
// geneig_lg_backtri.

namespace gk {

// The whole staged pipeline is built above order 64, and at any order when
// GENEIG_STAGED is defined (the HBM-staged alternative the autotuner weighs
// against the LDS-resident kernels); otherwise only geneig_lg_chol is, for the
// mixed mode's FP64 Cholesky and W = L^-1.
#if GENEIG_N > 64 || defined(GENEIG_STAGED)
#define GENEIG_LG_PIPELINE 1
#else
#define GENEIG_LG_PIPELINE 0
#endif

// Workgroup size of the staged kernels: 256; where only the Cholesky is
// built (orders up to 64) one wave per 16-row tile, so that small orders do
// not carry idle waves.
constexpr int LNT = (N <= 64 && !GENEIG_LG_PIPELINE) ? 64 * ((N + 15) / 16) : 256;
constexpr int LNW = LNT / 64;
constexpr int ND = (N + 63) / 64;             // rows per lane when lanes index rows
constexpr int LNB = (N + 15) / 16;             // 16-blocks per dimension
constexpr int LKEND = (N + 3) / 4 * 4;         // k-steps that touch rows/columns < N
constexpr int QZ_CH = 8;                       // columns per branch-free chunk of the QL appliers
constexpr size_t NN = (size_t)N * N;
constexpr size_t LDS_CAP = 160 * 1024;

// Per-matrix workspace (byte offsets). The same layout in the FP64 and FP32
// modules: regions are sized for FP64 data.
constexpr size_t WS_WI = 0;                   // L, then W = L^-1 (complex128, upper zero)
constexpr size_t WS_M = 16 * NN;              // M = W H W^H (complex128, full)
constexpr size_t WS_X = 32 * NN;              // scratch: W H; Z rows beyond LDS; refinement products
constexpr size_t WS_V = 48 * NN;              // V (complex, module precision; FP64 in refinement)
constexpr size_t WS_R = 64 * NN;              // reflectors: v_k(i) at (i, k), i >= k+2; tau_k at (k+1, k)
constexpr size_t WS_Z = 80 * NN;              // Z, columns in ascending eigenvalue order (real)
constexpr size_t WS_D = 88 * NN;              // d[N], e[N], ascending eigenvalues[N] (real)
constexpr size_t WS_L = (88 * NN + 24 * (size_t)N + 255) / 256 * 256;   // L packed by columns (complex128)
constexpr size_t WS_Q = (WS_L + 8 * (size_t)N * (N + 1) + 255) / 256 * 256;   // QL rotation log
// The log: header (sweeps, overflow flag), the window (l, m) of every sweep,
// the ascending rank of every eigenvalue, then the rotations (c_i, s_i) of all
// sweeps back to back. A matrix whose QL needs more sweeps or rotations than
// the log holds is redone by the applier with the rotations applied inline.
constexpr int QL_SW_CAP = 8 * N;
constexpr int QL_ROT_CAP = 2 * N * N;
constexpr size_t Q_SW = 16;
constexpr size_t Q_RANK = Q_SW + 8 * (size_t)QL_SW_CAP;
constexpr size_t Q_ROT = (Q_RANK + 4 * (size_t)N + 15) / 16 * 16;
constexpr size_t Q_BYTES = Q_ROT + 16 * (size_t)QL_ROT_CAP;
constexpr size_t WS_BYTES = (WS_Q + Q_BYTES + 255) / 256 * 256;

__device__ __forceinline__ char* ws_of(char* ws, long long wss, int b) { return ws + (size_t)b * (size_t)wss; }

// Element of a distributed array (element i at lane i % 64, register i / 64)
// at a wave-uniform, run-time index.
__device__ __forceinline__ real dget(const real (&a)[ND], int idx) {
    real v = a[0];
#pragma unroll
    for (int q = 1; q < ND; ++q)
        if ((idx >> 6) == q) v = a[q];
    return bcast(v, idx & 63);
}
__device__ __forceinline__ void dset(real (&a)[ND], int idx, real x) {
#pragma unroll
    for (int q = 0; q < ND; ++q)
        if ((idx >> 6) == q) a[q] = sel_lane(a[q], x, idx & 63);
}

__device__ __forceinline__ cx wave_sum_cx(cx v) { return cm(wave_sum(v.x), wave_sum(v.y)); }

// V = Q Z on the matrix cores, for the staged pipeline (geneig_lg_back) and
// the mixed mode's FP32 back half up to order 64 (geneig_mx_back), which keep
// the reflectors in different places. The reflectors are taken 16 at a
// time: H_k0 H_k0+1 ... H_k0+15 = I - Y T Y^H with Y the block's reflector
// columns (v_k(k+1) = 1, zero above) and T upper triangular (T_jj = tau_j,
// T(0:j, j) = -tau_j T(0:j, 0:j) Y(:, 0:j)^H v_j, the forward columnwise form
// of LAPACK's zlarft). Blocks are applied last first: X <- X - Y (T (Y^H X)).
// One wave per 16-column panel of X, held in registers as row tiles in the
// accumulator layout; a block touches only the row tiles at or below its
// first reflector's row.
//   prologue: wave w forms T for blocks w, w + BRT, ... (Gram Y^H Y on the
//             matrix cores, the recurrence on 16 lanes) into LDS;
//   per block: Y staged in LDS (coalesced, masked), then per wave
//             W1 = Y^H X, C2 = T W1, X -= Y C2, three complex products in the
//             3M form.
// Every read of the reflectors happens before the barrier that follows the
// last staging, so V may overwrite them.
constexpr int BRT = (N + 15) / 16;          // row tiles = column panels = waves
constexpr int BNT = 64 * BRT;
constexpr int NBK = (N - 1 + 15) / 16;      // blocks of the N - 1 reflectors
constexpr int BYLD = 16 * BRT + 1;          // staged column stride (odd: no bank conflicts)
constexpr int TLD = 17;                     // row stride of T in LDS
struct BackWork {
    cx T[NBK > 0 ? NBK : 1][16 * TLD];
    cx Y[16 * BYLD];                          // Y(row, m) at [m * BYLD + row]
};
static_assert(BRT * 256 <= 16 * BYLD, "the Gram scratch of every wave fits the staged block");
static_assert(16 * 16 * BRT <= 4 * BNT, "a block is at most four elements per thread");

// vget(k, row): the stored v_k(row), k <= N-2, k+2 <= row < N; tget(k): tau_k;
// zget(row, col): Z; vput(row, col, v): V(row, col) = v. KFAST: consecutive
// reflectors of one row are adjacent in memory (else consecutive rows of one
// reflector), which sets the order in which threads stage a block. Each
// thread holds at most 4 elements of a block; the next block's are loaded
// while the current one is applied.
constexpr int BSPT = 4;                     // staged elements per thread: 16 x 16 BRT / BNT
template <bool KFAST, class FV, class FT, class FZ, class FP>
__device__ __forceinline__ void back_wy(BackWork& bw, FV vget, FT tget, FZ zget, FP vput) {
    const int lane = lane_id(), wave = wave_id();
    const int c16 = lane & 15, k4 = lane >> 4;
    // Element (row, m) of reflector block k0: v_{k0+m}(row).
    auto refl = [&](int k0, int m, int row) -> cx {
        const int k = k0 + m;
        if (k > N - 2 || row <= k || row >= N) return cm(0.0, 0.0);
        if (row == k + 1) return cm(1.0, 0.0);
        return vget(k, row);
    };

    // T of every block. The Gram matrix goes through this wave's slice of the
    // (not yet used) staging buffer; lane i < 16 then forms row i of T.
    cx* G = bw.Y + wave * 256;
    for (int bi = wave; bi < NBK; bi += BRT) {
        const int k0 = 16 * bi;
        ctile3 a3;
        c3_zero(a3);
#pragma unroll
        for (int rt = 0; rt < BRT; ++rt) {
            if (rt < bi) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const cx y = refl(k0, c16, 16 * rt + acc_row(e, k4));
                cmfma3(a3, cconj(y), y, y.x + y.y);
            }
        }
        const ctile g = c3_final(a3);
#pragma unroll
        for (int e = 0; e < 4; ++e) G[acc_row(e, k4) * 16 + c16] = tile_get(g, e);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < 16) {
            const int i = lane;
            cx t[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const cx tau = (k0 + j <= N - 2) ? tget(k0 + j) : cm(0.0, 0.0);
                cx s = cm(0.0, 0.0);
#pragma unroll
                for (int m = 0; m < j; ++m)
                    if (m >= i) s = s + cmul(t[m], G[m * 16 + j]);
                t[j] = (i < j) ? cm(0.0, 0.0) - cmul(tau, s) : (i == j ? tau : cm(0.0, 0.0));
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) bw.T[bi][i * TLD + j] = t[j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // This wave's panel: columns 16 wave .. 16 wave + 15 of Z.
    const int col = 16 * wave + c16;
    ctile x[BRT];
#pragma unroll
    for (int rt = 0; rt < BRT; ++rt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = 16 * rt + acc_row(e, k4);
            x[rt].re[e] = (row < N && col < N) ? zget(row, col) : real(0.0);
            x[rt].im[e] = real(0.0);
        }

    // Element s of this thread's share of block bi: (m, row), or m < 0 past its end.
    auto slot = [&](int bi, int sidx, int& m, int& row) {
        const int r0 = 16 * bi, nr = 16 * BRT - r0;
        const int idx = (int)threadIdx.x + BNT * sidx;
        m = -1;
        row = 0;
        if (idx < 16 * nr) {
            m = KFAST ? idx % 16 : idx / nr;
            row = r0 + (KFAST ? idx / 16 : idx % nr);
        }
    };
    cx pre[BSPT];
    auto fetch = [&](int bi) {
#pragma unroll
        for (int u = 0; u < BSPT; ++u) {
            int m, row;
            slot(bi, u, m, row);
            pre[u] = (m >= 0) ? refl(16 * bi, m, row) : cm(0.0, 0.0);
        }
    };
    if (NBK > 0) fetch(NBK - 1);
    for (int bi = NBK - 1; bi >= 0; --bi) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < BSPT; ++u) {
            int m, row;
            slot(bi, u, m, row);
            if (m >= 0) bw.Y[m * BYLD + row] = pre[u];
        }
        __syncthreads();
        if (bi > 0) fetch(bi - 1);
        // W1 = Y^H X (16 x 16): A(m, row) = conj(Y(row, m)), B = the X tiles.
        ctile3 a3;
        c3_zero(a3);
#pragma unroll
        for (int rt = 0; rt < BRT; ++rt) {
            if (rt < bi) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const cx y = bw.Y[c16 * BYLD + 16 * rt + acc_row(e, k4)];
                const cx xe = tile_get(x[rt], e);
                cmfma3(a3, cconj(y), xe, xe.x + xe.y);
            }
        }
        const ctile w1 = c3_final(a3);
        // C2 = T W1.
        c3_zero(a3);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const cx t = bw.T[bi][c16 * TLD + acc_row(e, k4)];
            const cx we = tile_get(w1, e);
            cmfma3(a3, t, we, we.x + we.y);
        }
        const ctile c2 = c3_final(a3);
        // X(row tile) -= Y(row tile, :) C2.
        real c2s[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) c2s[e] = c2.re[e] + c2.im[e];
#pragma unroll
        for (int rt = 0; rt < BRT; ++rt) {
            if (rt < bi) continue;
            c3_zero(a3);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const cx y = bw.Y[acc_row(e, k4) * BYLD + 16 * rt + c16];
                cmfma3(a3, y, tile_get(c2, e), c2s[e]);
            }
            const ctile d = c3_final(a3);
            x[rt].re = x[rt].re - d.re;
            x[rt].im = x[rt].im - d.im;
        }
    }
#pragma unroll
    for (int rt = 0; rt < BRT; ++rt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int row = 16 * rt + acc_row(e, k4);
            if (row < N && col < N) vput(row, col, tile_get(x[rt], e));
        }
}

// ---------------------------------------------------------------------------
// Tridiagonalisation (module precision).
// ---------------------------------------------------------------------------
// Waves of the tridiagonalisation's workgroup: eight, twice the other staged
// kernels', so that each step's column loops are half as long per wave and
// a CU holds twice the waves to hide their latency; four where the FP32
// triangle and eight waves' partial sums would not fit LDS together (N = 192).
constexpr int HT_NW = 8;
constexpr int HT_NT = 64 * HT_NW;
// Bytes beside the packed triangle: v (two steps), w, per-wave partial sums
// of A v and the column sums of its upper part.
constexpr size_t HT_EXTRA = (size_t)(2 + 1 + HT_NW + 1) * N * sizeof(cx) + 64;
// LDS budget of the tridiagonalisation: in FP64 half the CU's LDS, so that
// two workgroups share a CU; the columns whose packed triangle does not fit
// are reduced in place in global memory first (K0 below). Measured at batch
// 16384 / 2048: N = 96 46.8 -> 26.6 ms (K0 = 5), N = 192 39.0 -> 25.3 ms
// (K0 = 94) against the whole LDS at one workgroup per CU; a third of the
// LDS (three workgroups) was slower (29.5 ms at N = 96). The FP32 build
// reduces its first K0 columns in an FP32 copy of M in the V slot (free
// until V = QZ), since the refinement still reads the FP64 M.
constexpr size_t HT_CAP = LDS_CAP / 2;
constexpr int ht_fit() {
    int nl = N;
    while (nl > 1 && (size_t)nl * (nl + 1) / 2 * sizeof(cx) + HT_EXTRA > HT_CAP) --nl;
    return nl;
}
constexpr int NL = ht_fit();                  // order of the LDS-resident trailing triangle
constexpr int K0 = N - NL;                    // columns reduced in global memory first

struct HtWork {
    cx P[NL * (NL + 1) / 2];                  // lower triangle of rows/columns K0.., packed by columns
    cx vb[2][N];
    cx wb[N];
    cx part[HT_NW][N];
    cx pc[N];
    cx tau[2];
};

__device__ __forceinline__ int pkl(int c) { return c * NL - (c * (c - 1)) / 2; }

// The matrix-vector product by row sums with mirrored reads (orders whose
// rows fit two per lane; the LDS-resident steps only): it replaces a wave
// reduction per column by one more LDS read per row (N = 96 FP64: 57.1 ->
// 53.0 ms per batch). At N = 192 it was slower (94.5 -> 105.1 ms FP64,
// where the first 94 columns are reduced in global memory, and 64.7 -> 65.7
// ms FP32).
constexpr bool ROW_SUMS = ND <= 2;

struct HtLds {
    static constexpr bool in_lds = true;
    cx* P;
    __device__ __forceinline__ cx& at(int i, int j) const { return P[pkl(j - K0) + i - j]; }
};
struct HtGlb {
    static constexpr bool in_lds = false;
    cx* G;
    __device__ __forceinline__ cx& at(int i, int j) const { return G[(size_t)j * N + i]; }
};

// Householder vector of column k, by the wave that owns column k (k % HT_NW):
// v(k+1) = 1, tau and the real beta as in LAPACK's zlarfg; v to LDS and to
// the reflector store, d_k, e_k = beta and tau to global memory.
template <class ST>
__device__ __forceinline__ void ht_house(HtWork& w, const ST& A, int k, cx* __restrict__ R,
                                         real* __restrict__ dg, real* __restrict__ eg) {
    const int lane = lane_id();
    const int m0 = k + 1;
    cx x[ND];
    real xn2 = real(0.0);
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const int r = 64 * q + lane;
        x[q] = (r > m0 && r < N) ? A.at(r, k) : cm(0.0, 0.0);
        xn2 += cabs2(x[q]);
    }
    xn2 = wave_sum(xn2);
    const cx alpha = A.at(m0, k);
    cx t = cm(0.0, 0.0), scal = cm(0.0, 0.0);
    real beta = alpha.x;
    if (!(xn2 == real(0.0) && alpha.y == real(0.0))) {
        const real nrm = sqrt_r(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        beta = (alpha.x >= real(0.0)) ? -nrm : nrm;
        t = cm((beta - alpha.x) / beta, -alpha.y / beta);
        const cx den = cm(alpha.x - beta, alpha.y);
        const real dd = real(1.0) / cabs2(den);
        scal = cm(den.x * dd, -den.y * dd);    // 1 / (alpha - beta)
    }
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const int r = 64 * q + lane;
        const cx v = (r == m0) ? cm(1.0, 0.0) : cmul(x[q], scal);
        if (r < N) w.vb[k & 1][r] = (r >= m0) ? v : cm(0.0, 0.0);
        if (r > m0 && r < N) R[(size_t)k * N + r] = v;
    }
    if (lane == 0) {
        R[(size_t)k * N + m0] = t;
        dg[k] = A.at(k, k).x;
        eg[k] = beta;
        w.tau[k & 1] = t;
    }
}

// Step k: A' -= v w^H + w v^H on the trailing rows/columns m0 = k+1.. with
// w = p - (tau/2)(p^H v) v, p = tau A' v. Column j belongs to wave j % HT_NW
// (lane = row, ND rows per lane). The product reads the lower triangle once:
// A(r, j) v_j goes to row r (per-lane sums, combined across waves in LDS)
// and conj(A(r, j)) v_r, r > j, to row j (a wave sum). The owner of column
// k+1 forms the next reflector right after its updates.
template <class ST>
__device__ __forceinline__ void ht_step(HtWork& w, const ST& A, int k, cx* __restrict__ R,
                                        real* __restrict__ dg, real* __restrict__ eg) {
    const int lane = lane_id(), wave = wave_id();
    const int m0 = k + 1, p = k & 1;
    const cx t = w.tau[p];
    const int jfirst = m0 + (((wave - m0) % HT_NW) + HT_NW) % HT_NW;
    if (t.x != real(0.0) || t.y != real(0.0)) {
        cx vr[ND], acc[ND];
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            const int r = 64 * q + lane;
            vr[q] = (r < N) ? w.vb[p][r] : cm(0.0, 0.0);
            acc[q] = cm(0.0, 0.0);
        }
        if constexpr (ROW_SUMS && ST::in_lds) {
            // p = M v over the active block by row sums only: column j
            // (owned by this wave) adds M(r, j) v_j to every active row r,
            // M(r, j) read from the stored lower triangle below the diagonal
            // and as the conjugate of M(j, r) above it (no reduction across
            // lanes per column). Unrolled so that several columns' loads
            // are in flight at once.
#pragma unroll 4
            for (int j = jfirst; j < N; j += HT_NW) {
                const cx vj = w.vb[p][j];
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    if (r >= m0 && r < N) {
                        const cx a = (r >= j) ? A.at(r, j) : cconj(A.at(j, r));
                        acc[q] = cfma(a, vj, acc[q]);
                    }
                }
            }
        } else {
            // Each stored element once: row sums below the diagonal per
            // lane, and the column sums of its mirror by a wave reduction
            // per column (w.pc).
            for (int j = jfirst; j < N; j += HT_NW) {
                const cx vj = w.vb[p][j];
                cx s = cm(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    if (r >= j && r < N) {
                        const cx a = A.at(r, j);
                        acc[q] = cfma(a, vj, acc[q]);
                        if (r > j) s = s + cconjmul(a, vr[q]);
                    }
                }
                s = wave_sum_cx(s);
                if (lane == 0) w.pc[j] = s;
            }
        }
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            const int r = 64 * q + lane;
            if (r < N) w.part[wave][r] = acc[q];
        }
        __syncthreads();
        cx wr[ND];
        cx pv = cm(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            const int r = 64 * q + lane;
            cx pr = cm(0.0, 0.0);
            if (r >= m0 && r < N) {
                if constexpr (!(ROW_SUMS && ST::in_lds)) pr = w.pc[r];
#pragma unroll
                for (int u = 0; u < HT_NW; ++u) pr = pr + w.part[u][r];
                pr = cmul(t, pr);
            }
            pv = pv + cconjmul(pr, vr[q]);
            wr[q] = pr;
        }
        pv = wave_sum_cx(pv);
        const cx cc = cmul(t, pv) * real(-0.5);
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            wr[q] = wr[q] + cmul(cc, vr[q]);
            const int r = 64 * q + lane;
            if (r < N && (r % HT_NW) == wave) w.wb[r] = wr[q];   // entries of this wave's columns
        }
        wave_lds_sync();
#pragma unroll 4
        for (int j = jfirst; j < N; j += HT_NW) {
            const cx vj = w.vb[p][j], wj = w.wb[j];
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                const int r = 64 * q + lane;
                if (r >= j && r < N) A.at(r, j) = cfms_conj(wr[q], vj, cfms_conj(vr[q], wj, A.at(r, j)));
            }
        }
    }
    if (k + 1 < N - 1 && wave == (k + 1) % HT_NW) ht_house(w, A, k + 1, R, dg, eg);
    __syncthreads();
}

// ---------------------------------------------------------------------------
// QL (module precision).
// ---------------------------------------------------------------------------
constexpr int QNT = 64 * ND;                  // one row of Z per thread
constexpr size_t QL_FIXED = (size_t)ND * 2 * N * sizeof(real) + (size_t)N * sizeof(int) + 64;
constexpr int ql_lds_waves() {
    int wv = ND;
    while (wv > 0 && (size_t)(64 * wv < N ? 64 * wv : N) * N * sizeof(real) + QL_FIXED > LDS_CAP) --wv;
    return wv;
}
constexpr int QW_LDS = ql_lds_waves();        // waves whose rows of Z live in LDS
constexpr int ZR_LDS = (64 * QW_LDS < N) ? 64 * QW_LDS : N;   // rows of Z in LDS (= its leading dimension)

// Apply the rotations (c_i, s_i), i = m-1 down to l, of one sweep to a row of
// Z whose element c sits at z[c * LD]. Loads run four steps ahead.
template <int LD, class ZP>
__device__ __forceinline__ void ql_apply(ZP z, const real* rc, const real* rs, int l, int m) {
    real zc = z[m * LD];
    int i = m - 1;
    for (; i - 3 >= l; i -= 4) {
        const real z0 = z[i * LD], z1 = z[(i - 1) * LD], z2 = z[(i - 2) * LD], z3 = z[(i - 3) * LD];
        real c = rc[i], s = rs[i];
        z[(i + 1) * LD] = fma(s, z0, c * zc);
        zc = fma(c, z0, -(s * zc));
        c = rc[i - 1];
        s = rs[i - 1];
        z[i * LD] = fma(s, z1, c * zc);
        zc = fma(c, z1, -(s * zc));
        c = rc[i - 2];
        s = rs[i - 2];
        z[(i - 1) * LD] = fma(s, z2, c * zc);
        zc = fma(c, z2, -(s * zc));
        c = rc[i - 3];
        s = rs[i - 3];
        z[(i - 2) * LD] = fma(s, z3, c * zc);
        zc = fma(c, z3, -(s * zc));
    }
    for (; i >= l; --i) {
        const real zi = z[i * LD];
        const real c = rc[i], s = rs[i];
        z[(i + 1) * LD] = fma(s, zi, c * zc);
        zc = fma(c, zi, -(s * zc));
    }
    z[l * LD] = zc;
}

// Rotation pair of the log.
struct rpair {
    real c, s;
};

// First m >= l with a negligible off-diagonal e_m (m = N - 1 if none).
__device__ __forceinline__ int ql_find_m(const real (&dr)[ND], const real (&er)[ND], int l) {
    const int lane = lane_id();
    int m = N - 1;
#pragma unroll
    for (int q = ND - 1; q >= 0; --q) {
        real dn = __shfl_down(dr[q], 1, 64);
        if (q + 1 < ND) {
            const real nx = bcast(dr[q + 1 < ND ? q + 1 : q], 0);
            if (lane == 63) dn = nx;
        }
        const int i = 64 * q + lane;
        const bool small = fabs(er[q]) <= REAL_HALF_ULP * (fabs(dr[q]) + fabs(dn));
        const unsigned long long mask = __ballot(small && i >= l && i < N - 1);
        if (mask) m = 64 * q + __ffsll((long long)mask) - 1;
    }
    return m;
}

__device__ __forceinline__ int ql_unconverged(const real (&er)[ND], int l) {
    const int lane = lane_id();
    int n = 0;
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const int i = 64 * q + lane;
        n += __popcll(__ballot(er[q] != real(0.0) && i >= l && i < N - 1));
    }
    return n;
}

// One implicit QL sweep on the window [l, m) with a Wilkinson shift
// (Bowdler, Martin, Reinsch and Wilkinson, Handbook II/3): updates (d, e)
// and leaves the sweep's rotation i in element i of (cr, sr).
__device__ __forceinline__ void ql_sweep(real (&dr)[ND], real (&er)[ND], int l, int m, real (&cr)[ND],
                                         real (&sr)[ND]) {
    const real d_l = dget(dr, l);
    const real e_l = dget(er, l);
    real g = (dget(dr, l + 1) - d_l) / (real(2.0) * e_l);
    const real rr = sqrt_r(g * g + real(1.0));
    const real d_m = dget(dr, m);
    g = d_m - d_l + e_l / (g + copysign(rr, g));
    real s = real(1.0), c = real(1.0), p = real(0.0);
    real d_up = d_m;
    // Step i reads element i before any step of this sweep writes it (step i
    // writes i + 1): copies taken here keep those reads off the sweep's chain.
    real d0[ND], e0[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        cr[q] = real(1.0);
        sr[q] = real(0.0);
        d0[q] = dr[q];
        e0[q] = er[q];
    }
#pragma unroll
    for (int i = N - 2; i >= 0; --i) {
        if (i < m && i >= l) {
            real e_next, d_next;
            ql_rotation(bcast(e0[i >> 6], i & 63), bcast(d0[i >> 6], i & 63), s, c, g, p, d_up, e_next, d_next);
            er[(i + 1) >> 6] = sel_lane(er[(i + 1) >> 6], e_next, (i + 1) & 63);
            dr[(i + 1) >> 6] = sel_lane(dr[(i + 1) >> 6], d_next, (i + 1) & 63);
            cr[i >> 6] = sel_lane(cr[i >> 6], c, i & 63);
            sr[i >> 6] = sel_lane(sr[i >> 6], s, i & 63);
        }
    }
    dset(dr, l, d_l - p);
    dset(er, l, g);
    dset(er, m, real(0.0));
}

// Ascending rank of eigenvalue cidx (ties by index).
__device__ __forceinline__ int ql_rank(const real (&dr)[ND], real dc, int cidx) {
    int rank = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const real dk = bcast(dr[k >> 6], k & 63);
        rank += (dk < dc || (dk == dc && k < cidx)) ? 1 : 0;
    }
    return rank;
}

// ---------------------------------------------------------------------------
// Complex FP64 products on the matrix cores (FP64 module only), in 16 x 16
// complex tiles (ctile / cmfma).
// ---------------------------------------------------------------------------
#ifndef GENEIG_FP32

// Element (i, k) of an N x N column-major matrix in global memory, zero
// outside; the address is clamped so the load is branch-free.
__device__ __forceinline__ cx gel(const cx* __restrict__ M, int i, int k) {
    const bool in = i < N && k < N;
    const cx v = M[(size_t)(in ? k : 0) * N + (in ? i : 0)];
    return in ? v : cm(0.0, 0.0);
}

// acc[g] = sum over k in [kb, ke) of A(i, k) B(k, j) for row tile rt0 + g and
// column tile ct; fa / fb return zero outside N. kb, ke multiples of 4.
// Accumulated in the 3M form (three real MFMAs per complex product).
template <int RG, class FA, class FB>
__device__ __forceinline__ void gemm_rg(ctile (&acc)[RG], FA fa, FB fb, int rt0, int ct, int kb, int ke) {
    const int l = lane_id();
    const int c16 = l & 15, k4 = l >> 4;
    ctile3 a3[RG];
#pragma unroll
    for (int g = 0; g < RG; ++g) c3_zero(a3[g]);
#pragma unroll 2
    for (int k0 = kb; k0 < ke; k0 += 4) {
        const cx b = fb(k0 + k4, ct * 16 + c16);
        const real bs = b.x + b.y;
#pragma unroll
        for (int g = 0; g < RG; ++g) cmfma3(a3[g], fa((rt0 + g) * 16 + c16, k0 + k4), b, bs);
    }
#pragma unroll
    for (int g = 0; g < RG; ++g) acc[g] = c3_final(a3[g]);
}
// Row of element v of a result tile held by this lane (column: ct * 16 + lane % 16).
__device__ __forceinline__ int tile_row(int rt, int v) { return rt * 16 + 4 * v + (lane_id() >> 4); }

// Panel column stride (odd, so the 16 columns of a tile row fall in distinct
// LDS banks) and the stride of one column of a block row in the inverse's
// staging (17: consecutive threads' columns in distinct banks).
constexpr int PLD = N | 1;
constexpr int YLD = 17;
struct CholWork {
    cx pan[(16 * PLD > YLD * 16 * LNB) ? 16 * PLD : YLD * 16 * LNB];   // panel; trtri: Y, then X
    cx lii[256];              // trtri: diagonal block of L
    real liinv[16];           // trtri: 1 / L(r, r) of that block
    int fail;                 // pivot failure of the panel (orders up to 64)
};
#endif  // !GENEIG_FP32

}  // namespace gk

using gk::cx;
using gk::real;

#ifndef GENEIG_FP32
// S = L L^H and W = L^-1 in the workspace (upper triangle zero), and L
// packed by columns: in the staged pipeline to the workspace (WS_L), else to
// Lout when it is not NULL. On a
// non-positive pivot k (0-based) of a leading minor: info = N + k + 1, W NaN,
// U zero (LAPACK zhegv's convention), and the later stages skip the matrix.
// For small orders (while three workgroups still fit a CU), L and then W live
// packed in LDS (the products read LDS), and W goes out at the end (packed
// where W_PACKED_OUT below holds, in full otherwise); otherwise they live in
// the workspace.
namespace gk {
constexpr bool CHOL_LDS = N <= 64 && sizeof(CholWork) + (size_t)N * (N + 1) / 2 * sizeof(cx) <= 52 * 1024;
constexpr int LPK = N * (N + 1) / 2;
__device__ __forceinline__ int lpk(int j) { return j * N - (j * (j - 1)) / 2; }
// Where only the Cholesky is built (the mixed mode's module, orders up to 64)
// and W was formed in LDS, W goes out as its lower triangle packed by columns
// (element (i, j), i >= j, at lpk(j) + i - j; matrix b at b LPK): half the
// bytes of the full layout, which the staged pipeline's consumers read.
constexpr bool W_PACKED_OUT = CHOL_LDS && !GENEIG_LG_PIPELINE;
}  // namespace gk

// Whether geneig_lg_chol writes W packed (the host sets its stride by it).
extern "C" __device__ const int geneig_w_packed = gk::W_PACKED_OUT ? 1 : 0;

extern "C" __global__ void __launch_bounds__(gk::LNT, 3)
geneig_lg_chol(const cx* __restrict__ S, real* __restrict__ W, cx* __restrict__ U, int* __restrict__ info,
               char* __restrict__ ws, long long wss, cx* __restrict__ Lout, int batch) {
    using namespace gk;
    __shared__ CholWork cw;
    __shared__ cx Lp[CHOL_LDS ? LPK : 1];
    const int b = blockIdx.x;
    if (b >= batch) return;
    const int tid = threadIdx.x, wave = wave_id(), lane = lane_id();
    cx* Lg = reinterpret_cast<cx*>(ws_of(ws, wss, b) + WS_WI);
    const cx* Sb = S + (size_t)b * NN;
    if constexpr (!CHOL_LDS) {
        for (int e = tid; e < (int)NN; e += LNT)
            if (e % N < e / N) Lg[e] = cm(0.0, 0.0);
        __syncthreads();
    }
    // L(i, k) (zero above the diagonal and outside), and its store (i >= k).
    auto fL = [&](int i, int k) -> cx {
        if constexpr (CHOL_LDS) {
            const bool in = i < N && k < N && i >= k;
            const cx v = Lp[in ? lpk(k) + i - k : 0];
            return in ? v : cm(0.0, 0.0);
        } else {
            return gel(Lg, i, k);
        }
    };
    auto setL = [&](int i, int k, cx v) {
        if constexpr (CHOL_LDS) Lp[lpk(k) + i - k] = v;
        else Lg[(size_t)k * N + i] = v;
    };
    for (int p = 0; p < LNB; ++p) {
        const int j0 = 16 * p;
        const int jw = (N - j0 < 16) ? N - j0 : 16;
        // Panel rows j0.. of columns j0 .. j0+jw: S minus the columns left of it.
        for (int rt = p + wave; rt < LNB; rt += LNW) {
            ctile acc[1];
            gemm_rg<1>(acc, fL, [&](int k, int c) { return cconj(c < jw ? fL(j0 + c, k) : cm(0.0, 0.0)); },
                       rt, 0, 0, j0);
            const int c = lane & 15;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = tile_row(rt, v);
                if (i < N && c < jw) cw.pan[c * PLD + (i - j0)] = Sb[(size_t)(j0 + c) * N + i] - tile_get(acc[0], v);
            }
        }
        __syncthreads();
        int fail = 0;
        if constexpr (N <= 64) {
            // Every panel row is a lane of wave 0: the unblocked Cholesky of the
            // panel runs there with row t in registers, L(c2, c) broadcast from
            // lane c2 by v_readlane, no barrier per column.
            if (wave == 0) {
                const int t = lane;
                const bool tin = t < N - j0;
                cx r[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) r[c] = (tin && c < jw) ? cw.pan[c * PLD + t] : cm(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    if (c < jw && !fail) {
                        const real piv = bcast(r[c].x, c);
                        if (!(piv > real(0.0))) {
                            fail = j0 + c + 1;
                        } else {
                            // 1 / sqrt(piv) to about an ulp (v_rsq_f64 and a
                            // third-order step), sqrt(piv) = piv / sqrt(piv).
                            const real ri = rsqrt_nr(piv);
                            const real sq = piv * ri;
                            const cx lc = (t == c) ? cm(sq, 0.0) : r[c] * ri;
                            r[c] = lc;
#pragma unroll
                            for (int c2 = c + 1; c2 < 16; ++c2)
                                if (c2 < jw) r[c2] = r[c2] - cmulc(lc, bcast(lc, c2));
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < 16; ++c)
                    if (c < jw && tin && t >= c) setL(j0 + t, j0 + c, r[c]);
                if (lane == 0) cw.fail = fail;
            }
            __syncthreads();
            fail = cw.fail;
        } else {
            // Unblocked Cholesky of the panel in LDS, thread t = row j0 + t, one
            // barrier per column: every thread derives the pivot of column c from
            // LDS, reads column c unscaled (scaling each entry itself) and writes
            // its scaled entry of column c straight to L in global memory, so
            // column c is never rewritten in LDS while others read it.
            const int t = tid;
            const bool tin = t < N - j0;
            for (int c = 0; c < jw; ++c) {
                const real piv = cw.pan[c * PLD + c].x;
                if (!(piv > real(0.0))) {
                    fail = j0 + c + 1;
                    break;
                }
                const real sq = sqrt_r(piv);
                const real ri = real(1.0) / sq;
                if (t >= c && tin) {
                    const cx ltc = (t == c) ? cm(sq, 0.0) : cw.pan[c * PLD + t] * ri;
                    setL(j0 + t, j0 + c, ltc);
                    for (int c2 = c + 1; c2 < jw && c2 <= t; ++c2)
                        cw.pan[c2 * PLD + t] = cw.pan[c2 * PLD + t] - cmulc(ltc, cw.pan[c * PLD + c2] * ri);
                }
                __syncthreads();
            }
        }
        if (fail) {
            real* Wb = W + (size_t)b * N;
            cx* Ub = U + (size_t)b * NN;
            for (int e = tid; e < (int)NN; e += LNT) Ub[e] = cm(0.0, 0.0);
            for (int e = tid; e < N; e += LNT) Wb[e] = __builtin_nan("");
            if (tid == 0) info[b] = N + fail;
            return;
        }
        __syncthreads();
    }
    // L itself, packed by columns, for a back-transform by substitution,
    // before W overwrites it: to the workspace in the staged pipeline, else
    // to Lout when the caller asks for it.
    cx* Lo = nullptr;
    if constexpr (GENEIG_LG_PIPELINE) Lo = reinterpret_cast<cx*>(ws_of(ws, wss, b) + WS_L);
    else if (Lout) Lo = Lout + (size_t)b * LPK;
    if (Lo) {
        if constexpr (CHOL_LDS) {
            for (int e = tid; e < LPK; e += LNT) Lo[e] = Lp[e];
        } else {
            for (int e = tid; e < (int)NN; e += LNT) {
                const int i = e % N, j = e / N;
                if (i >= j) Lo[lpk(j) + i - j] = Lg[e];
            }
        }
        __syncthreads();
    }
    // W = L^-1 in place, block row ib after block row ib-1: with X = W,
    // X(ib, j) = L_ii^-1 (delta - sum_{k < r0} L(ib, k) X(k, j)).
    for (int ib = 0; ib < LNB; ++ib) {
        const int r0 = 16 * ib;
        const int rw = (N - r0 < 16) ? N - r0 : 16;
        for (int e = tid; e < 256; e += LNT) {
            const int rr = e & 15, c = e >> 4;
            cw.lii[e] = (rr < rw && c < rw) ? fL(r0 + rr, r0 + c) : cm(0.0, 0.0);
        }
        if (tid < rw) cw.liinv[tid] = real(1.0) / fL(r0 + tid, r0 + tid).x;
        for (int ct = wave; ct < ib; ct += LNW) {
            ctile acc[1];
            gemm_rg<1>(acc, fL, fL, ib, ct, 16 * ct, r0);
            const int j = ct * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = tile_row(ib, v);
                if (i < N) cw.pan[j * YLD + (i - r0)] = tile_get(acc[0], v);
            }
        }
        __syncthreads();
        // Forward substitution with L_ii, thread j = column j, its 16 entries
        // in registers (unrolled; the term of the newest entry comes last, so
        // each entry waits on its predecessor for one multiply-add only).
        const int j = tid;
        if (j < r0 + rw) {
            const cx* xj = cw.pan + j * YLD;
            cx x[16];
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) x[rr] = (rr < rw && j < r0) ? xj[rr] : cm(0.0, 0.0);
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) {
                if (rr < rw) {
                    cx s = (j < r0) ? cm(0.0, 0.0) - x[rr] : cm((rr == j - r0) ? 1.0 : 0.0, 0.0);
#pragma unroll
                    for (int ss = 0; ss < rr; ++ss) s = s - cmul(cw.lii[ss * 16 + rr], x[ss]);
                    x[rr] = s * cw.liinv[rr];
                }
            }
#pragma unroll
            for (int rr = 0; rr < 16; ++rr)
                if (rr < rw && (CHOL_LDS ? r0 + rr >= j : true)) setL(r0 + rr, j, x[rr]);
        }
        __syncthreads();
    }
    if constexpr (W_PACKED_OUT) {
        for (int e = tid; e < LPK; e += LNT) Lg[e] = Lp[e];
    } else if constexpr (CHOL_LDS) {
        for (int e = tid; e < (int)NN; e += LNT) {
            const int i = e % N, j = e / N;
            Lg[e] = (i >= j) ? Lp[lpk(j) + i - j] : cm(0.0, 0.0);
        }
    }
    if (tid == 0) info[b] = 0;
}

#if GENEIG_LG_PIPELINE
// M = W H W^H: X = W H (W lower triangular: k <= row), then the lower tiles
// of X W^H, mirrored so that M is exactly Hermitian.
extern "C" __global__ void __launch_bounds__(gk::LNT)
geneig_lg_formm(const cx* __restrict__ H, const int* __restrict__ info, char* __restrict__ ws, long long wss,
                int batch) {
    using namespace gk;
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    const int wave = wave_id(), lane = lane_id();
    char* wsb = ws_of(ws, wss, b);
    const cx* Wi = reinterpret_cast<const cx*>(wsb + WS_WI);
    cx* X = reinterpret_cast<cx*>(wsb + WS_X);
    cx* M = reinterpret_cast<cx*>(wsb + WS_M);
    const cx* Hb = H + (size_t)b * NN;
    constexpr int NRG = (LNB + 3) / 4;
    for (int t = wave; t < NRG * LNB; t += LNW) {
        const int rt0 = 4 * (t / LNB), ct = t % LNB;
        const int ke = (16 * (rt0 + 4) < LKEND) ? 16 * (rt0 + 4) : LKEND;
        ctile acc[4];
        gemm_rg<4>(acc, [&](int i, int k) { return gel(Wi, i, k); }, [&](int k, int j) { return gel(Hb, k, j); },
                   rt0, ct, 0, ke);
        const int j = ct * 16 + (lane & 15);
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = tile_row(rt0 + g, v);
                if (i < N && j < N) X[(size_t)j * N + i] = tile_get(acc[g], v);
            }
    }
    __syncthreads();
    int t = 0;
    for (int ct = 0; ct < LNB; ++ct) {
        const int ke = (16 * (ct + 1) < LKEND) ? 16 * (ct + 1) : LKEND;
        for (int rt0 = ct; rt0 < LNB; rt0 += 4, ++t) {
            if (t % LNW != wave) continue;
            ctile acc[4];
            gemm_rg<4>(acc, [&](int i, int k) { return gel(X, i, k); },
                       [&](int k, int jj) { return cconj(gel(Wi, jj, k)); }, rt0, ct, 0, ke);
            const int j = ct * 16 + (lane & 15);
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = tile_row(rt0 + g, v);
                    if (i < N && j < N && i >= j) {
                        const cx m = tile_get(acc[g], v);
                        if (i == j) {
                            M[(size_t)j * N + i] = cm(m.x, 0.0);
                        } else {
                            M[(size_t)j * N + i] = m;
                            M[(size_t)i * N + j] = cconj(m);
                        }
                    }
                }
        }
    }
}

// U = L^-H V (V in FP64 from the workspace) by block back substitution with
// L's packed lower triangle from the workspace (no explicit inverse in the
// back-transform): a wave takes one 16-column block of U at a time, row tile
// by row tile from the last, keeping the finished tiles in its LDS slice
// (N x 16) as the B operand of the blocks above, and storing each tile to U
// as it completes. BS_SLICES waves work (the column blocks spread evenly over
// them): as many as leave two workgroups per CU, or, where fewer than two
// slices would then fit, as many as fit one workgroup.
namespace gk {
constexpr int BS_SLICE_BYTES = N * 16 * 16;
constexpr int BS_FIT2 = (78 * 1024) / BS_SLICE_BYTES;
constexpr int BS_FIT1 = (150 * 1024) / BS_SLICE_BYTES;
constexpr int BS_SLICES = (BS_FIT2 >= 2 ? BS_FIT2 : BS_FIT1) < LNW ? (BS_FIT2 >= 2 ? BS_FIT2 : BS_FIT1) : LNW;
static_assert(BS_SLICES >= 1, "one back-substitution slice must fit");
}  // namespace gk

extern "C" __global__ void __launch_bounds__(gk::LNT)
geneig_lg_backtri(cx* __restrict__ U, const int* __restrict__ info, char* __restrict__ ws, long long wss,
                  int batch) {
    using namespace gk;
    __shared__ cx bsl[BS_SLICES][N * 16];
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    const int wave = wave_id(), lane = lane_id();
    const int c16 = lane & 15, g4 = lane >> 4;
    char* wsb = ws_of(ws, wss, b);
    const cx* Lb = reinterpret_cast<const cx*>(wsb + WS_L);
    const cx* V = reinterpret_cast<const cx*>(wsb + WS_V);
    cx* Ub = U + (size_t)b * NN;
    auto fL = [&](int i, int k) { return Lb[lpk(k) + i - k]; };
    {
        if (wave < BS_SLICES) {
            cx* sl = bsl[wave];
            for (int ct = wave; ct < LNB; ct += BS_SLICES) {
                const int j = ct * 16 + c16;
                for (int ib = LNB - 1; ib >= 0; --ib) {
                    ctile acc;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const cx x = gel(V, ib * 16 + 4 * v + g4, j);
                        acc.re[v] = x.x;
                        acc.im[v] = x.y;
                    }
                    if (ib < LNB - 1) {
                        // sum_{k in blocks above} conj(L(k, 16 ib + i)) U(k, j).
                        ctile3 a3;
                        c3_zero(a3);
                        const int i = ib * 16 + c16;
                        for (int k0 = 16 * (ib + 1); k0 < LKEND; k0 += 4) {
                            const int k = k0 + g4;
                            const cx bu = (k < N) ? sl[c16 * N + k] : cm(0.0, 0.0);
                            const cx a = (k < N && i < N) ? cconj(fL(k, i)) : cm(0.0, 0.0);
                            cmfma3(a3, a, bu, bu.x + bu.y);
                        }
                        const ctile sm = c3_final(a3);
                        acc.re = acc.re - sm.re;
                        acc.im = acc.im - sm.im;
                    }
                    backsub_diag(acc, fL, ib);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = ib * 16 + 4 * v + g4;
                        if (row < N) {
                            const cx u = cm(acc.re[v], acc.im[v]);
                            sl[c16 * N + row] = u;
                            if (j < N) Ub[(size_t)j * N + row] = u;
                        }
                    }
                    wave_lds_sync();
                }
            }
        }
    }
}
#endif  // GENEIG_LG_PIPELINE
#endif  // !GENEIG_FP32

#if GENEIG_LG_PIPELINE
// Bytes of one matrix's workspace, for the host's consistency check.
extern "C" __device__ const unsigned long long geneig_lg_ws_bytes = gk::WS_BYTES;

// Workgroup size of geneig_lg_hetrd, read by the host before launching it.
extern "C" __device__ const int geneig_lg_hetrd_threads = gk::HT_NT;

namespace gk {
// Columns 0 .. K0-1 of the tridiagonalisation, whose trailing triangle
// does not fit LDS yet, in panels of up to HT_NBP columns with the updates
// deferred (the panel step of LAPACK's blocked zhetrd): within a panel the
// matrix in global memory keeps its value from the panel's start, the
// current column and the matrix-vector product add the panel's rank-2
// corrections -sum_q (v_q w_q^H + w_q v_q^H) from the v and w vectors kept in
// LDS (in the area of the trailing triangle, idle until it is loaded), and
// the trailing lower triangle is updated in global memory once per panel.
// Against updating it every step, this reads the active block once per step
// instead of twice and writes it once per panel instead of once per step
// (N = 192, batch 2048: 17.8 -> 14.3 ms; N = 150, batch 4096: 20.5 ->
// 19.2 ms). Where K0 is smaller the steps update in place, since each
// panel step's serial column and reflector, which the other waves wait for,
// outweigh the traffic saved (N = 128, batch 8192, K0 = 44: 19.7 ms in place,
// 20.9 ms in panels; N = 96, K0 = 8: 16.9 and 17.4 ms).
constexpr int HT_NBP_FIT = (NL * (NL + 1) / 2) / (2 * N);
constexpr int HT_NBP = HT_NBP_FIT < 16 ? HT_NBP_FIT : 16;
constexpr bool HT_PANELS = K0 >= 64;

__device__ __forceinline__ void ht_global_panels(HtWork& w, cx* __restrict__ Mg, cx* __restrict__ R,
                                                 real* __restrict__ dg, real* __restrict__ eg) {
    static_assert(HT_NBP >= 1, "the panel vectors must fit the triangle's area");
    const int lane = lane_id(), wave = wave_id();
    cx* Vp = w.P;                    // [HT_NBP][N]
    cx* Wp = w.P + HT_NBP * N;       // [HT_NBP][N]
    const HtGlb G{Mg};
    for (int k0 = 0; k0 < K0; k0 += HT_NBP) {
        const int np = (K0 - k0 < HT_NBP) ? K0 - k0 : HT_NBP;
        for (int t = 0; t < np; ++t) {
            const int k = k0 + t, m0 = k + 1, p = k & 1;
            if (wave == 0) {
                // Column k as it stands after the panel's earlier steps, rows >= k.
                cx x[ND];
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    cx a = (r >= k && r < N) ? G.at(r, k) : cm(0.0, 0.0);
                    if (r >= k && r < N) {
                        for (int tt = 0; tt < t; ++tt)
                            a = cfms_conj(Wp[tt * N + r], Vp[tt * N + k], cfms_conj(Vp[tt * N + r], Wp[tt * N + k], a));
                    }
                    x[q] = a;
                }
                // Householder vector (LAPACK's zlarfg conventions, as ht_house).
                real xn2 = real(0.0);
                cx alpha = cm(0.0, 0.0), dk = cm(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    if (r > m0 && r < N) xn2 += cabs2(x[q]);
                    if ((m0 >> 6) == q) alpha = bcast(x[q], m0 & 63);
                    if ((k >> 6) == q) dk = bcast(x[q], k & 63);
                }
                xn2 = wave_sum(xn2);
                cx tau = cm(0.0, 0.0), scal = cm(0.0, 0.0);
                real beta = alpha.x;
                if (!(xn2 == real(0.0) && alpha.y == real(0.0))) {
                    const real nrm = sqrt_r(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
                    beta = (alpha.x >= real(0.0)) ? -nrm : nrm;
                    tau = cm((beta - alpha.x) / beta, -alpha.y / beta);
                    const cx den = cm(alpha.x - beta, alpha.y);
                    const real dd = real(1.0) / cabs2(den);
                    scal = cm(den.x * dd, -den.y * dd);    // 1 / (alpha - beta)
                }
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    const cx v = (r == m0) ? cm(1.0, 0.0) : cmul(x[q], scal);
                    if (r < N) w.vb[p][r] = (r >= m0) ? v : cm(0.0, 0.0);
                    if (r > m0 && r < N) R[(size_t)k * N + r] = v;
                }
                if (lane == 0) {
                    R[(size_t)k * N + m0] = tau;
                    dg[k] = dk.x;
                    eg[k] = beta;
                    w.tau[p] = tau;
                }
            }
            __syncthreads();
            const cx tk = w.tau[p];
            const int jfirst = m0 + (((wave - m0) % HT_NW) + HT_NW) % HT_NW;
            cx vr[ND], wr[ND];
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                const int r = 64 * q + lane;
                vr[q] = (r < N) ? w.vb[p][r] : cm(0.0, 0.0);
                wr[q] = cm(0.0, 0.0);
            }
            if (tk.x != real(0.0) || tk.y != real(0.0)) {
                // A v over the active block from the panel-start matrix: row
                // sums below the diagonal per lane, the mirror's column sums by
                // a wave reduction per column.
                cx acc[ND];
#pragma unroll
                for (int q = 0; q < ND; ++q) acc[q] = cm(0.0, 0.0);
                for (int j = jfirst; j < N; j += HT_NW) {
                    const cx vj = w.vb[p][j];
                    cx sc = cm(0.0, 0.0);
#pragma unroll
                    for (int q = 0; q < ND; ++q) {
                        const int r = 64 * q + lane;
                        if (r >= j && r < N) {
                            const cx a = G.at(r, j);
                            acc[q] = cfma(a, vj, acc[q]);
                            if (r > j) sc = sc + cconjmul(a, vr[q]);
                        }
                    }
                    sc = wave_sum_cx(sc);
                    if (lane == 0) w.pc[j] = sc;
                }
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    if (r < N) w.part[wave][r] = acc[q];
                }
                __syncthreads();
                // The panel's corrections: -(v_q (w_q^H v) + w_q (v_q^H v)).
                cx pr[ND];
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    pr[q] = cm(0.0, 0.0);
                    if (r >= m0 && r < N) {
                        cx s = w.pc[r];
#pragma unroll
                        for (int u = 0; u < HT_NW; ++u) s = s + w.part[u][r];
                        pr[q] = s;
                    }
                }
                for (int tt = 0; tt < t; ++tt) {
                    cx al = cm(0.0, 0.0), be = cm(0.0, 0.0);
#pragma unroll
                    for (int q = 0; q < ND; ++q) {
                        const int r = 64 * q + lane;
                        if (r < N) {
                            al = al + cconjmul(Wp[tt * N + r], vr[q]);
                            be = be + cconjmul(Vp[tt * N + r], vr[q]);
                        }
                    }
                    al = wave_sum_cx(al);
                    be = wave_sum_cx(be);
#pragma unroll
                    for (int q = 0; q < ND; ++q) {
                        const int r = 64 * q + lane;
                        if (r >= m0 && r < N)
                            pr[q] = pr[q] - cmul(Vp[tt * N + r], al) - cmul(Wp[tt * N + r], be);
                    }
                }
                cx pv = cm(0.0, 0.0);
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    pr[q] = cmul(tk, pr[q]);
                    pv = pv + cconjmul(pr[q], vr[q]);
                }
                pv = wave_sum_cx(pv);
                const cx cc = cmul(tk, pv) * real(-0.5);
#pragma unroll
                for (int q = 0; q < ND; ++q) wr[q] = pr[q] + cmul(cc, vr[q]);
            }
            // v_t and w_t of the panel (zero w, and so no correction, for an
            // identity reflector).
            if (wave == 0) {
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    if (r < N) {
                        Vp[t * N + r] = vr[q];
                        Wp[t * N + r] = wr[q];
                    }
                }
            }
            __syncthreads();
        }
        // The panel's rank-2 updates into the trailing lower triangle, once.
        const int jt = k0 + np;
        const int jf = jt + (((wave - jt) % HT_NW) + HT_NW) % HT_NW;
        for (int j = jf; j < N; j += HT_NW) {
            cx a[ND];
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                const int r = 64 * q + lane;
                a[q] = (r >= j && r < N) ? G.at(r, j) : cm(0.0, 0.0);
            }
            for (int tt = 0; tt < np; ++tt) {
                const cx vj = Vp[tt * N + j], wj = Wp[tt * N + j];
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    const int r = 64 * q + lane;
                    if (r >= j && r < N)
                        a[q] = cfms_conj(Wp[tt * N + r], vj, cfms_conj(Vp[tt * N + r], wj, a[q]));
                }
            }
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                const int r = 64 * q + lane;
                if (r >= j && r < N) G.at(r, j) = a[q];
            }
        }
        __syncthreads();
    }
}
}  // namespace gk

// Tridiagonalisation of M (read as FP64, reduced in the module precision):
// d, e and the reflectors to the workspace.
extern "C" __global__ void __launch_bounds__(gk::HT_NT)
geneig_lg_hetrd(const int* __restrict__ info, char* __restrict__ ws, long long wss, int batch) {
    using namespace gk;
    __shared__ HtWork w;
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    const int tid = threadIdx.x;
    char* wsb = ws_of(ws, wss, b);
    const cd* Mg = reinterpret_cast<const cd*>(wsb + WS_M);
    cx* R = reinterpret_cast<cx*>(wsb + WS_R);
    real* dg = reinterpret_cast<real*>(wsb + WS_D);
    real* eg = dg + N;
    const HtLds L{w.P};
    auto load_trailing = [&](const auto* src) {
        for (int e = tid; e < (int)NN; e += HT_NT) {
            const int i = e % N, j = e / N;
            if (j >= K0 && i >= j) w.P[pkl(j - K0) + i - j] = to_cx(src[e]);
        }
    };
    if constexpr (K0 > 0) {
        // Columns 0 .. K0-1 in global memory, then the trailing triangle to
        // LDS. FP64: in place (the FP64 modes do not need M after this
        // stage); FP32: in an FP32 copy of M's lower triangle in the V slot.
#ifdef GENEIG_FP32
        cx* Gm = reinterpret_cast<cx*>(wsb + WS_V);
        for (int e = tid; e < (int)NN; e += HT_NT) {
            const int i = e % N, j = e / N;
            if (i >= j) Gm[e] = to_cx(Mg[e]);
        }
        __syncthreads();
#else
        cx* Gm = reinterpret_cast<cx*>(wsb + WS_M);
#endif
        if constexpr (HT_PANELS) {
            ht_global_panels(w, Gm, R, dg, eg);
            load_trailing(Gm);
            __syncthreads();
            if (wave_id() == 0) ht_house(w, L, K0, R, dg, eg);
            __syncthreads();
        } else {
            const HtGlb G{Gm};
            if (wave_id() == 0) ht_house(w, G, 0, R, dg, eg);
            __syncthreads();
            for (int k = 0; k < K0; ++k) ht_step(w, G, k, R, dg, eg);
            load_trailing(Gm);
            __syncthreads();
        }
    } else {
        load_trailing(Mg);
        __syncthreads();
        if (wave_id() == 0) ht_house(w, L, 0, R, dg, eg);
        __syncthreads();
    }
    for (int k = K0; k < N - 1; ++k) ht_step(w, L, k, R, dg, eg);
    if (tid == 0) {
        dg[N - 1] = L.at(N - 1, N - 1).x;
        eg[N - 1] = real(0.0);
    }
}

// Implicit QL on (d, e) with the rotations accumulated into Z (one row per
// thread, QNT threads), then ascending order: eigenvalues to wout (matrix b
// at wout + b * wout_stride) and Z with its columns in that order to the
// workspace. Every wave runs the scalar recurrence itself and applies each
// sweep's rotations to its own rows at once. A sweep that has not converged
// after QL_MAXIT iterations sets info to the number of off-diagonals left
// (LAPACK's convention). The inline form, for matrices whose rotations
// overflow the log.
namespace gk {
__device__ __forceinline__ void ql_inline(int b, real* __restrict__ wout, long long wout_stride,
                                          int* __restrict__ info, char* wsb, real* zs, real (*rot)[2][N],
                                          int* rank_s) {
    const int lane = lane_id(), wave = wave_id();
    const int r = wave * 64 + lane;
    const bool rin = r < N;
    const bool in_lds = wave < QW_LDS;
    const real* dsrc = reinterpret_cast<const real*>(wsb + WS_D);
    const real* esrc = dsrc + N;
    real* zg = reinterpret_cast<real*>(wsb + WS_X);      // rows of Z beyond the LDS
    real* zout = reinterpret_cast<real*>(wsb + WS_Z);
    real dr[ND], er[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const int i = 64 * q + lane;
        dr[q] = (i < N) ? dsrc[i] : real(0.0);
        er[q] = (i < N - 1) ? esrc[i] : real(0.0);
    }
    if (rin) {
        if (in_lds) {
            for (int c = 0; c < N; ++c) zs[c * ZR_LDS + r] = (c == r) ? real(1.0) : real(0.0);
        } else {
            for (int c = 0; c < N; ++c) zg[(size_t)c * N + r] = (c == r) ? real(1.0) : real(0.0);
        }
    }
    real* rc = rot[wave][0];
    real* rs = rot[wave][1];
    int l = 0, iter = 0, unconverged = 0;
    while (l < N) {
        const int m = ql_find_m(dr, er, l);
        if (m == l) {
            ++l;
            iter = 0;
            continue;
        }
        if (iter++ == QL_MAXIT) {
            unconverged = ql_unconverged(er, l);
            break;
        }
        real cr[ND], sr[ND];
        ql_sweep(dr, er, l, m, cr, sr);
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            const int i = 64 * q + lane;
            if (i < N) {
                rc[i] = cr[q];
                rs[i] = sr[q];
            }
        }
        wave_lds_sync();
        if (rin) {
            if (in_lds) ql_apply<ZR_LDS>(zs + r, rc, rs, l, m);
            else ql_apply<N>(zg + r, rc, rs, l, m);
        }
        wave_lds_sync();
    }
    // Wave q ranks eigenvalues 64 q + lane.
    {
        const int cidx = 64 * wave + lane;
        real dc = dr[0];
#pragma unroll
        for (int q = 1; q < ND; ++q)
            if (wave == q) dc = dr[q];
        const int rank = ql_rank(dr, dc, cidx);
        if (cidx < N) {
            rank_s[cidx] = rank;
            wout[(size_t)b * (size_t)wout_stride + rank] = dc;
        }
    }
    if (threadIdx.x == 0 && unconverged) info[b] = unconverged;
    __syncthreads();
    if (rin) {
        if (in_lds) {
            for (int c = 0; c < N; ++c) zout[(size_t)rank_s[c] * N + r] = zs[c * ZR_LDS + r];
        } else {
            for (int c = 0; c < N; ++c) zout[(size_t)rank_s[c] * N + r] = zg[(size_t)c * N + r];
        }
    }
}
}  // namespace gk

// QL, first half: one wave per matrix runs the scalar recurrence alone
// (registers only, so many matrices share a CU and hide each other's
// latency) and logs every sweep's window and rotations, then the ranks, and
// writes the ascending eigenvalues to wout and info as above. More sweeps than
// the log holds, or than the caller's sweep_cap, mark the log as overflowed.
extern "C" __global__ void __launch_bounds__(64)
geneig_lg_qlr(real* __restrict__ wout, long long wout_stride, int* __restrict__ info, char* __restrict__ ws,
              long long wss, int batch, int sweep_cap) {
    using namespace gk;
    const int cap = sweep_cap < QL_SW_CAP ? sweep_cap : QL_SW_CAP;
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    const int lane = lane_id();
    char* wsb = ws_of(ws, wss, b);
    const real* dsrc = reinterpret_cast<const real*>(wsb + WS_D);
    const real* esrc = dsrc + N;
    char* qb = wsb + WS_Q;
    int* hdr = reinterpret_cast<int*>(qb);
    int* sw = reinterpret_cast<int*>(qb + Q_SW);
    int* rk = reinterpret_cast<int*>(qb + Q_RANK);
    rpair* rot = reinterpret_cast<rpair*>(qb + Q_ROT);
    real dr[ND], er[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const int i = 64 * q + lane;
        dr[q] = (i < N) ? dsrc[i] : real(0.0);
        er[q] = (i < N - 1) ? esrc[i] : real(0.0);
    }
    int l = 0, iter = 0, nsw = 0, nrot = 0, unconverged = 0;
    bool overflow = false;
    while (l < N) {
        const int m = ql_find_m(dr, er, l);
        if (m == l) {
            ++l;
            iter = 0;
            continue;
        }
        if (iter++ == QL_MAXIT) {
            unconverged = ql_unconverged(er, l);
            break;
        }
        if (nsw >= cap || nrot + (m - l) > QL_ROT_CAP) {
            overflow = true;
            break;
        }
        real cr[ND], sr[ND];
        ql_sweep(dr, er, l, m, cr, sr);
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            const int i = 64 * q + lane;
            if (i >= l && i < m) rot[nrot + i - l] = rpair{cr[q], sr[q]};
        }
        if (lane == 0) {
            sw[2 * nsw] = l;
            sw[2 * nsw + 1] = m;
        }
        ++nsw;
        nrot += m - l;
    }
    if (overflow) {
        if (lane == 0) {
            hdr[0] = 0;
            hdr[1] = 1;
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const int cidx = 64 * q + lane;
        const int rank = ql_rank(dr, dr[q], cidx);
        if (cidx < N) {
            rk[cidx] = rank;
            wout[(size_t)b * (size_t)wout_stride + rank] = dr[q];
        }
    }
    if (lane == 0) {
        hdr[0] = nsw;
        hdr[1] = 0;
        if (unconverged) info[b] = unconverged;
    }
}

// QL, first half, for large batches: 64 matrices per wave, one per lane
// (ql_lanes_logged), instead of one wave per matrix. The one-wave form spends
// the whole wave on one scalar recurrence, so its time grows with the batch;
// here every lane runs its own and the time is about one lane's recurrence,
// whatever the batch up to the waves the CUs hold at once (measured at
// N = 96: 5.3 ms in FP64, 3.7 ms in FP32, against 12.2 and 6.0 ms for the
// one-wave form at 16,384 matrices, and 3.9 / 3.1 ms against 4.2 / 2.1 ms at
// 4,096). Same inputs, log, outputs and info as geneig_lg_qlr. d and e of
// the wave's matrices sit in LDS (2 N LG_QLR_LANES reals), orders up to 128.
#if GENEIG_N <= 128
namespace gk {
// Matrices per wave (one per lane; the other lanes idle): 16, so that a
// batch of a few thousand matrices still spreads over every SIMD (N = 96,
// 1,024 matrices: 4.1 ms against 6.0 ms with 64 per wave). Above order 128
// the batches are too small for either form to beat one wave per matrix
// (N = 192 FP32, 2,048 matrices: 17.4 ms with 4 per wave against 5.9 ms).
constexpr int LG_QLR_LANES = 16;
}  // namespace gk
extern "C" __device__ const int geneig_lg_qlr_lanes_count = gk::LG_QLR_LANES;

extern "C" __global__ void __launch_bounds__(64)
geneig_lg_qlr_lanes(real* __restrict__ wout, long long wout_stride, int* __restrict__ info, char* __restrict__ ws,
                    long long wss, int batch, int sweep_cap) {
    using namespace gk;
    const int cap = sweep_cap < QL_SW_CAP ? sweep_cap : QL_SW_CAP;
    constexpr int LN = LG_QLR_LANES;
    __shared__ real ds[N * LN];
    __shared__ real es[N * LN];
    const int lane = lane_id();
    const int b = blockIdx.x * LN + lane;
    const bool mine = lane < LN && b < batch && info[b] <= N;
    char* wsb = ws_of(ws, wss, mine ? b : 0);
    const real* dsrc = reinterpret_cast<const real*>(wsb + WS_D);
    const real* esrc = dsrc + N;
    if (lane < LN) {
        for (int i = 0; i < N; ++i) {
            ds[i * LN + lane] = mine ? dsrc[i] : real(0.0);
            es[i * LN + lane] = (mine && i < N - 1) ? esrc[i] : real(0.0);
        }
    }
    char* qb = wsb + WS_Q;
    int* hdr = reinterpret_cast<int*>(qb);
    int* sw = reinterpret_cast<int*>(qb + Q_SW);
    int* rk = reinterpret_cast<int*>(qb + Q_RANK);
    real* rot = reinterpret_cast<real*>(qb + Q_ROT);
    int nsw, unconverged, overflow;
    ql_lanes_logged<LN>(ds, es, mine, cap, QL_ROT_CAP, sw, rot, nsw, unconverged, overflow);
    if (!mine) return;
    if (overflow) {
        hdr[0] = 0;
        hdr[1] = 1;
        return;
    }
    // Ascending ranks (ties by index) and the eigenvalues.
    for (int k = 0; k < N; ++k) {
        const real dk = ds[k * LN + lane];
        int rank = 0;
        for (int j = 0; j < N; ++j) {
            const real dj = ds[j * LN + lane];
            rank += (dj < dk || (dj == dk && j < k)) ? 1 : 0;
        }
        rk[k] = rank;
        wout[(size_t)b * (size_t)wout_stride + rank] = dk;
    }
    hdr[0] = nsw;
    hdr[1] = 0;
    if (unconverged) info[b] = unconverged;
}
#endif

// Rows of Z in registers at every staged order (a row is N registers in
// FP32, 2N in FP64; above order 104 the compiler keeps part of it in the
// accumulation registers, one wave per SIMD; at N = 192 FP64, 256 VGPRs and
// 172 AGPRs, no spills). It replaced, above order 128 in FP64, a form with
// the rows in LDS or global memory (N = 192: 31.0 -> 9 ms).
// QL, second half: Z = product of the logged rotations, one row of Z per
// thread held in registers, then Z with its columns in ascending order to the
// workspace. Every sweep's rotations are staged by the wave in its own LDS
// slice at their column index, identity (1, 0) outside the sweep's window, so
// the columns run unrolled at compile-time addresses: chunks of QZ_CH columns
// that miss the window are skipped (a uniform branch), the rest are applied
// without branches (an identity rotation leaves both entries unchanged,
// exactly). The next sweep's rotations are fetched while the current one is
// applied. Matrices whose log overflowed are left to geneig_lg_qlz_inline.
extern "C" __global__ void __launch_bounds__(gk::QNT)
geneig_lg_qlz(real* __restrict__ wout, long long wout_stride, int* __restrict__ info, char* __restrict__ ws,
              long long wss, int batch) {
    using namespace gk;
    __shared__ real rcs[ND][2][2][N];             // [wave][sweep parity][c or s][column]
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    char* wsb = ws_of(ws, wss, b);
    const char* qb = wsb + WS_Q;
    const int* hdr = reinterpret_cast<const int*>(qb);
    if (hdr[1]) return;
    const int* sw = reinterpret_cast<const int*>(qb + Q_SW);
    const int* rk = reinterpret_cast<const int*>(qb + Q_RANK);
    const rpair* rot = reinterpret_cast<const rpair*>(qb + Q_ROT);
    const int lane = lane_id(), wave = wave_id();
    const int r = wave * 64 + lane;
    real* zout = reinterpret_cast<real*>(wsb + WS_Z);
    real z[N];
#pragma unroll
    for (int c = 0; c < N; ++c) z[c] = (c == r) ? real(1.0) : real(0.0);
    // Rotation of column p (lane p % 64, register p / 64) of the sweep whose
    // rotations start at base (rotation l + j at base[j]).
    auto fetch = [&](rpair (&f)[ND], const rpair* base, int l, int m) {
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            const int p = 64 * q + lane;
            f[q] = (p >= l && p < m) ? base[p - l] : rpair{real(1.0), real(0.0)};
        }
    };
    const int nsw = hdr[0];
    int off = 0;
    int l = nsw > 0 ? sw[0] : 0, m = nsw > 0 ? sw[1] : 0;
    rpair cur[ND];
    fetch(cur, rot, l, m);
    for (int t = 0; t < nsw; ++t) {
        real* rc = rcs[wave][t & 1][0];
        real* rs = rcs[wave][t & 1][1];
#pragma unroll
        for (int q = 0; q < ND; ++q) {
            const int p = 64 * q + lane;
            if (p < N) {
                rc[p] = cur[q].c;
                rs[p] = cur[q].s;
            }
        }
        wave_lds_sync();
        const int off_n = off + (m - l);
        const bool last = t + 1 == nsw;
        const int ln = last ? 0 : sw[2 * t + 2], mn = last ? 0 : sw[2 * t + 3];
        fetch(cur, rot + off_n, ln, mn);
        // Columns i = m-1 down to l, rotation i acting on (i, i+1).
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
    if (r < N) {
#pragma unroll
        for (int c = 0; c < N; ++c) zout[(size_t)rk[c] * N + r] = z[c];
    }
}

// Matrices whose QL rotation log overflowed: QL solved inline (ql_inline),
// eigenvalues to wout, Z to the workspace; every other matrix returns at once.
extern "C" __global__ void __launch_bounds__(gk::QNT)
geneig_lg_qlz_inline(real* __restrict__ wout, long long wout_stride, int* __restrict__ info,
                     char* __restrict__ ws, long long wss, int batch) {
    using namespace gk;
    __shared__ real zs[ZR_LDS > 0 ? ZR_LDS * N : 1];
    __shared__ real rotb[ND][2][N];
    __shared__ int rank_s[N];
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    char* wsb = ws_of(ws, wss, b);
    const int* hdr = reinterpret_cast<const int*>(wsb + WS_Q);
    if (!hdr[1]) return;
    gk::ql_inline(b, wout, wout_stride, info, wsb, zs, rotb, rank_s);
}

#if defined(GENEIG_FP32) || GENEIG_N <= 128
// V = Q Z on the matrix cores (back_wy): the FP32 build at every staged
// order, the FP64 build up to order 128, where its register tiles fit.
// Workgroup size of geneig_lg_back, read by the host before launching it.
extern "C" __device__ const int geneig_lg_back_threads = gk::BNT;

extern "C" __global__ void __launch_bounds__(gk::BNT)
geneig_lg_back(const int* __restrict__ info, char* __restrict__ ws, long long wss, int batch) {
    using namespace gk;
    __shared__ BackWork bw;
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    char* wsb = ws_of(ws, wss, b);
    const cx* R = reinterpret_cast<const cx*>(wsb + WS_R);
    const real* Z = reinterpret_cast<const real*>(wsb + WS_Z);
    cx* V = reinterpret_cast<cx*>(wsb + WS_V);
    back_wy<false>(
        bw, [&](int k, int row) { return R[(size_t)k * N + row]; }, [&](int k) { return R[(size_t)k * N + k + 1]; },
        [&](int row, int col) { return Z[(size_t)col * N + row]; },
        [&](int row, int col, cx v) { V[(size_t)col * N + row] = v; });
}
#else
// V = Q Z = H_0 H_1 ... H_{N-2} Z (reflectors applied last first), eight
// columns per wave at a time in registers (lane = row, ND rows per lane):
// per reflector a wave sum per column and a rank-1 update.
extern "C" __device__ const int geneig_lg_back_threads = gk::LNT;

extern "C" __global__ void __launch_bounds__(gk::LNT)
geneig_lg_back(const int* __restrict__ info, char* __restrict__ ws, long long wss, int batch) {
    using namespace gk;
    constexpr int CB = 8;
    constexpr int NCB = (N + CB - 1) / CB;
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    const int lane = lane_id(), wave = wave_id();
    char* wsb = ws_of(ws, wss, b);
    const cx* R = reinterpret_cast<const cx*>(wsb + WS_R);
    const real* Z = reinterpret_cast<const real*>(wsb + WS_Z);
    cx* V = reinterpret_cast<cx*>(wsb + WS_V);
    for (int cbk = wave; cbk < NCB; cbk += LNW) {
        const int c0 = cbk * CB;
        cx x[ND][CB];
#pragma unroll
        for (int q = 0; q < ND; ++q)
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                const int r = 64 * q + lane, col = c0 + c;
                x[q][c] = cm((r < N && col < N) ? Z[(size_t)col * N + r] : real(0.0), 0.0);
            }
        // Reflector k is fetched while reflector k+1 is applied.
        auto fetch = [&](int k, cx (&v)[ND], cx& tau) {
            tau = R[(size_t)k * N + k + 1];
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                const int r = 64 * q + lane;
                v[q] = (r == k + 1) ? cm(1.0, 0.0) : ((r > k + 1 && r < N) ? R[(size_t)k * N + r] : cm(0.0, 0.0));
            }
        };
        cx v[ND], tau;
        fetch(N - 2, v, tau);
        for (int k = N - 2; k >= 0; --k) {
            cx vn[ND], taun = cm(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < ND; ++q) vn[q] = cm(0.0, 0.0);
            if (k > 0) fetch(k - 1, vn, taun);
            if (tau.x != real(0.0) || tau.y != real(0.0)) {
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    cx s = cm(0.0, 0.0);
#pragma unroll
                    for (int q = 0; q < ND; ++q) s = s + cconjmul(v[q], x[q][c]);
                    s = cmul(tau, wave_sum_cx(s));
#pragma unroll
                    for (int q = 0; q < ND; ++q) x[q][c] = x[q][c] - cmul(v[q], s);
                }
            }
#pragma unroll
            for (int q = 0; q < ND; ++q) v[q] = vn[q];
            tau = taun;
        }
#pragma unroll
        for (int q = 0; q < ND; ++q)
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                const int r = 64 * q + lane, col = c0 + c;
                if (r < N && col < N) V[(size_t)col * N + r] = x[q][c];
            }
    }
}

#endif  // MFMA V = Q Z

#ifndef GENEIG_FP32
namespace gk {
// Row tiles per task in the refinement's products (two keep its registers
// low enough for two workgroups per CU).
constexpr int RRG = 2;

// acc1[g] = A B1 and acc2[g] = A B2 over k in [kb, ke) (shared A loads),
// in the 3M form.
template <int RG, class FA, class FB1, class FB2>
__device__ __forceinline__ void gemm2_rg(ctile (&acc1)[RG], ctile (&acc2)[RG], FA fa, FB1 fb1, FB2 fb2, int rt0,
                                         int ct, int kb, int ke) {
    const int l = lane_id();
    const int c16 = l & 15, k4 = l >> 4;
    ctile3 a1[RG], a2[RG];
#pragma unroll
    for (int g = 0; g < RG; ++g) {
        c3_zero(a1[g]);
        c3_zero(a2[g]);
    }
#pragma unroll 2
    for (int k0 = kb; k0 < ke; k0 += 4) {
        const cx b1 = fb1(k0 + k4, ct * 16 + c16);
        const cx b2 = fb2(k0 + k4, ct * 16 + c16);
        const real bs1 = b1.x + b1.y, bs2 = b2.x + b2.y;
#pragma unroll
        for (int g = 0; g < RG; ++g) {
            const cx a = fa((rt0 + g) * 16 + c16, k0 + k4);
            cmfma3(a1[g], a, b1, bs1);
            cmfma3(a2[g], a, b2, bs2);
        }
    }
#pragma unroll
    for (int g = 0; g < RG; ++g) {
        acc1[g] = c3_final(a1[g]);
        acc2[g] = c3_final(a2[g]);
    }
}

// Rayleigh-Ritz on columns (a, b) of V (global memory) for the Hermitian M
// (global), by one wave (lane = row): the 2 x 2 projected pencil
// (V^H M V, V^H V) of the pair, then V(:, [a b]) <- V(:, [a b]) X and
// lam[a], lam[b] <- the Ritz values. The columns' entries travel between
// lanes by v_readlane, never as wave-uniform loads of memory this kernel
// writes.
__device__ __forceinline__ void ritz_pair_g(const cx* __restrict__ M, cx* V, real* lam, int a, int b) {
    const int lane = lane_id();
    cx ua[ND], ub[ND], ha[ND], hb[ND];
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const int r = 64 * q + lane;
        ua[q] = (r < N) ? V[(size_t)a * N + r] : cm(0.0, 0.0);
        ub[q] = (r < N) ? V[(size_t)b * N + r] : cm(0.0, 0.0);
        ha[q] = cm(0.0, 0.0);
        hb[q] = cm(0.0, 0.0);
    }
#pragma unroll
    for (int qk = 0; qk < ND; ++qk) {
#pragma unroll 1
        for (int kk = 0; kk < 64 && 64 * qk + kk < N; ++kk) {
            const int k = 64 * qk + kk;
            const cx va = bcast(ua[qk], kk), vb = bcast(ub[qk], kk);
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                const int r = 64 * q + lane;
                const cx m = (r < N) ? M[(size_t)k * N + r] : cm(0.0, 0.0);
                ha[q] = ha[q] + cmul(m, va);
                hb[q] = hb[q] + cmul(m, vb);
            }
        }
    }
    real t11 = 0.0, t22 = 0.0, g11 = 0.0, g22 = 0.0;
    cx t12 = cm(0.0, 0.0), g12 = cm(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        t11 += cconjmul(ua[q], ha[q]).x;
        t22 += cconjmul(ub[q], hb[q]).x;
        t12 = t12 + cconjmul(ua[q], hb[q]);
        g11 += cabs2(ua[q]);
        g22 += cabs2(ub[q]);
        g12 = g12 + cconjmul(ua[q], ub[q]);
    }
    t11 = wave_sum(t11);
    t22 = wave_sum(t22);
    t12 = wave_sum_cx(t12);
    g11 = wave_sum(g11);
    g22 = wave_sum(g22);
    g12 = wave_sum_cx(g12);
    real mu1, mu2;
    cx X[2][2];
    eig2_pencil(t11, t22, t12, g11, g22, g12, mu1, mu2, X);
#pragma unroll
    for (int q = 0; q < ND; ++q) {
        const int r = 64 * q + lane;
        if (r < N) {
            V[(size_t)a * N + r] = cmul(ua[q], X[0][0]) + cmul(ub[q], X[1][0]);
            V[(size_t)b * N + r] = cmul(ua[q], X[0][1]) + cmul(ub[q], X[1][1]);
        }
    }
    if (lane == 0) {
        lam[a] = mu1;
        lam[b] = mu2;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
}  // namespace gk

// Refinement of the FP32 eigenpairs in the Cholesky frame (mixed mode): with
// W = L^-1, M = W H W^H (both FP64, in the workspace) and V the FP32
// eigenvectors of M, `iters` Ogita-Aishima steps on the matrix cores
// (P = M V, lambda_j = v_j^H p_j / v_j^H v_j, T = V^H P, G = V^H V,
// E = -(G - I)/2 + A with A_ij = (T_ij - (lambda_i + lambda_j) G_ij / 2) /
// (lambda_j - lambda_i) for pairs whose gap is at least delta and 0
// otherwise, V <- V + V E), Rayleigh-Ritz on every cluster (a run of
// eigenvalues with adjacent gaps below delta: one exact 2 x 2 step for a
// pair, Jacobi sweeps over its pairs otherwise), then U = W^H V. In step it
// (from 0), delta = 2^it delta_rel max |lambda|. Buffers: V in the scratch
// and V slots alternately, P in the reflector slot, E in the caller's U.
// Convergence test as in geneig_mx_refine: a last correction with an entry
// above `tol` marks the matrix info = N.
namespace gk {
// LDS of geneig_lg_refine, padded to 56 KB so that at most two workgroups
// share a CU: its products stream their operands through L2, and more
// matrices per CU evict each other's. N = 96, batch 16384: 27.5 ms at four
// workgroups per CU (126 VGPRs allow it), 24.4 ms at two, 38.7 ms at one.
constexpr int REF_GNR = (56 * 1024) / (int)sizeof(real);
}  // namespace gk
extern "C" __global__ void __launch_bounds__(gk::LNT, 2)
geneig_lg_refine(real* __restrict__ W, cx* __restrict__ U, int* __restrict__ info, char* __restrict__ ws,
                 long long wss, int batch, int iters, real delta_rel, real tol) {
    using namespace gk;
    __shared__ real lam[N];
    __shared__ real gnr[REF_GNR > N ? REF_GNR : N];    // v_j^H v_j (padded, see REF_GNR)
    __shared__ real wmax[LNW];
    const int b = blockIdx.x;
    if (b >= batch || info[b] > N) return;
    const int tid = threadIdx.x, wave = wave_id(), lane = lane_id();
    char* wsb = ws_of(ws, wss, b);
    const cx* Wi = reinterpret_cast<const cx*>(wsb + WS_WI);
    const cx* M = reinterpret_cast<const cx*>(wsb + WS_M);
    cx* Va = reinterpret_cast<cx*>(wsb + WS_X);
    cx* Vb = reinterpret_cast<cx*>(wsb + WS_V);
    cx* P = reinterpret_cast<cx*>(wsb + WS_R);
    cx* Ub = U + (size_t)b * NN;
    cx* E = Ub;
    const cf* V32 = reinterpret_cast<const cf*>(wsb + WS_V);
    const float* lam32 = reinterpret_cast<const float*>(wsb + WS_D) + 2 * N;
    for (int e = tid; e < (int)NN; e += LNT) Va[e] = to_cx(V32[e]);
    for (int e = tid; e < N; e += LNT) lam[e] = (real)lam32[e];
    __syncthreads();
    constexpr int NRG = (LNB + RRG - 1) / RRG;
    real delta = 0.0;
    real emax = 0.0;                             // this lane's largest |E| entry, last step
    for (int it = 0; it < iters; ++it) {
        const cx* V = Va;
        // P = M V.
        for (int t = wave; t < NRG * LNB; t += LNW) {
            const int rt0 = RRG * (t / LNB), ct = t % LNB;
            ctile acc[RRG];
            gemm_rg<RRG>(acc, [&](int i, int k) { return gel(M, i, k); }, [&](int k, int j) { return gel(V, k, j); },
                       rt0, ct, 0, LKEND);
            const int j = ct * 16 + (lane & 15);
#pragma unroll
            for (int g = 0; g < RRG; ++g)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = tile_row(rt0 + g, v);
                    if (i < N && j < N) P[(size_t)j * N + i] = tile_get(acc[g], v);
                }
        }
        __syncthreads();
        // lambda_j = v_j^H p_j / v_j^H v_j, column j by wave j % LNW.
        real mx = 0.0;
        for (int j = wave; j < N; j += LNW) {
            real tj = 0.0, gj = 0.0;
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                const int r = 64 * q + lane;
                if (r < N) {
                    const cx v = V[(size_t)j * N + r];
                    tj += cconjmul(v, P[(size_t)j * N + r]).x;
                    gj += cabs2(v);
                }
            }
            const real gs = wave_sum(gj);
            const real lj = wave_sum(tj) / gs;
            if (lane == 0) {
                lam[j] = lj;
                gnr[j] = gs;
            }
            mx = fmax(mx, fabs(lj));
        }
        if (lane == 0) wmax[wave] = mx;
        __syncthreads();
        mx = wmax[0];
#pragma unroll
        for (int u = 1; u < LNW; ++u) mx = fmax(mx, wmax[u]);
        // The threshold doubles every step, so a pair treated as close once
        // stays close and ends in Rayleigh-Ritz.
        delta = delta_rel * (real)(1 << it) * mx;
        // E = (T - G diag(lambda)) / gap with T = V^H M V, G = V^H V, that is
        // Y / gap with Y = V^H R and the residual R = P - V diag(lambda)
        // (formed as the product's right operand): one product instead of
        // T and G. The diagonal takes (1 - G_jj) / 2 and pairs closer than
        // delta -G_ij / 2: those entries are written after the tiles, each
        // G_ij a dot product of two columns.
        for (int t = wave; t < NRG * LNB; t += LNW) {
            const int rt0 = RRG * (t / LNB), ct = t % LNB;
            const int j = ct * 16 + (lane & 15);
            const real lj = lam[j < N ? j : 0];
            ctile Y[RRG];
            gemm_rg<RRG>(Y, [&](int i, int k) { return cconj(gel(V, k, i)); },
                         [&](int k, int jj) { return gel(P, k, jj) - gel(V, k, jj) * lj; }, rt0, ct, 0, LKEND);
#pragma unroll
            for (int g = 0; g < RRG; ++g)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = tile_row(rt0 + g, v);
                    if (i < N && j < N) {
                        const real li = lam[i];
                        const real gap = lj - li;
                        cx e = tile_get(Y[g], v) * rcp_nr(gap);
                        if (i == j) e = cm(real(0.5) * (real(1.0) - gnr[j]), 0.0);
                        else if (fabs(gap) < delta) e = cm(0.0, 0.0);
                        if (it == iters - 1) emax = fmax(emax, fmax(fabs(e.x), fabs(e.y)));
                        E[(size_t)j * N + i] = e;
                    }
                }
        }
        __syncthreads();
        // Close pairs (i, j), column j by wave j % LNW: E_ij = -v_i^H v_j / 2.
        for (int j = wave; j < N; j += LNW) {
            const real lj = lam[j];
            for (int i0 = 0; i0 < N; i0 += 64) {
                const int ir = i0 + lane;
                unsigned long long m = __ballot(ir < N && ir != j && fabs(lam[ir < N ? ir : 0] - lj) < delta);
                while (m) {
                    const int ii = i0 + (int)__builtin_ctzll(m);
                    m &= m - 1;
                    cx g = cm(0.0, 0.0);
#pragma unroll
                    for (int q = 0; q < ND; ++q) {
                        const int r = 64 * q + lane;
                        if (r < N) g = g + cconjmul(V[(size_t)ii * N + r], V[(size_t)j * N + r]);
                    }
                    g = wave_sum_cx(g) * real(-0.5);
                    if (lane == 0) E[(size_t)j * N + ii] = g;
                    if (it == iters - 1) emax = fmax(emax, fmax(fabs(g.x), fabs(g.y)));
                }
            }
        }
        __syncthreads();
        // V <- V + V E, into the other buffer.
        for (int t = wave; t < NRG * LNB; t += LNW) {
            const int rt0 = RRG * (t / LNB), ct = t % LNB;
            ctile acc[RRG];
            gemm_rg<RRG>(acc, [&](int i, int k) { return gel(V, i, k); }, [&](int k, int j) { return gel(E, k, j); },
                       rt0, ct, 0, LKEND);
            const int j = ct * 16 + (lane & 15);
#pragma unroll
            for (int g = 0; g < RRG; ++g)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = tile_row(rt0 + g, v);
                    if (i < N && j < N) Vb[(size_t)j * N + i] = gel(V, i, j) + tile_get(acc[g], v);
                }
        }
        __syncthreads();
        cx* tmp = Va;
        Va = Vb;
        Vb = tmp;
    }
    // Clusters: runs [c0, c1] with adjacent gaps below delta (wave 0).
    if (wave == 0 && iters > 0) {
        int c0 = 0;
        while (c0 < N - 1) {
            int c1 = c0;
            while (c1 + 1 < N && fabs(lam[c1 + 1] - lam[c1]) < delta) ++c1;
            if (c1 > c0) {
                const int sweeps = (c1 == c0 + 1) ? 1 : 4;
                for (int sw = 0; sw < sweeps; ++sw)
                    for (int a = c0; a < c1; ++a)
                        for (int bb = a + 1; bb <= c1; ++bb) ritz_pair_g(M, Va, lam, a, bb);
                // Keep the run ascending.
                for (int a = c0 + 1; a <= c1; ++a)
                    for (int q = a; q > c0 && lam[q] < lam[q - 1]; --q) {
                        const real tl = lam[q];
#pragma unroll
                        for (int qq = 0; qq < ND; ++qq) {
                            const int r = 64 * qq + lane;
                            if (r < N) {
                                const cx t0 = Va[(size_t)q * N + r];
                                Va[(size_t)q * N + r] = Va[(size_t)(q - 1) * N + r];
                                Va[(size_t)(q - 1) * N + r] = t0;
                            }
                        }
                        wave_lds_sync();
                        if (lane == 0) {
                            lam[q] = lam[q - 1];
                            lam[q - 1] = tl;
                        }
                        wave_lds_sync();
                    }
            }
            c0 = c1 + 1;
        }
    }
    __syncthreads();
    emax = wave_max(emax);
    if (lane == 0) wmax[wave] = emax;
    __syncthreads();
    if (tid == 0 && iters > 0 && info[b] == 0) {
        real m = wmax[0];
#pragma unroll
        for (int u = 1; u < LNW; ++u) m = fmax(m, wmax[u]);
        if (!(m <= tol)) info[b] = N;
    }
    // U = W^H V.
    const cx* V = Va;
    for (int t = wave; t < NRG * LNB; t += LNW) {
        const int rt0 = RRG * (t / LNB), ct = t % LNB;
        ctile acc[RRG];
        gemm_rg<RRG>(acc, [&](int i, int k) { return cconj(gel(Wi, k, i)); }, [&](int k, int j) { return gel(V, k, j); },
                   rt0, ct, 16 * rt0, LKEND);
        const int j = ct * 16 + (lane & 15);
#pragma unroll
        for (int g = 0; g < RRG; ++g)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = tile_row(rt0 + g, v);
                if (i < N && j < N) Ub[(size_t)j * N + i] = tile_get(acc[g], v);
            }
    }
    for (int e = tid; e < N; e += LNT) W[(size_t)b * N + e] = lam[e];
}
#endif  // !GENEIG_FP32

#endif  // GENEIG_LG_PIPELINE

#ifndef GENEIG_FP32
// Uneven batches (every bucket, in every FP64 module): the bucket's matrices
// (work list ids, packed back to back at element offsets moff, orders nn) are
// embedded as (H (+) pad I, S (+) I) into a contiguous batch of order N for
// the bucket's FP64 split solver; pad is far above any eigenvalue of the
// pencils, so the padded problem decouples exactly (the Cholesky, the
// reflectors and the QL rotations never mix the two blocks) and its n
// smallest eigenpairs are the original ones. The scatter returns the leading
// n eigenvalues, the leading n x n block of U and zhegv's info for order n (a
// failing leading minor k lies in the original block: N + k becomes n + k).
// One workgroup of 256 threads per matrix.
namespace gk {
constexpr double LG_PAD = 0x1p600;
}

extern "C" __global__ void __launch_bounds__(256)
geneig_lg_gather(const cx* __restrict__ H, const cx* __restrict__ S, const int* __restrict__ ids,
                 const long long* __restrict__ moff, const int* __restrict__ nn, int count, cx* __restrict__ Hp,
                 cx* __restrict__ Sp) {
    using namespace gk;
    const int b = blockIdx.x;
    if (b >= count) return;
    const int id = ids[b];
    const int n = nn[id];
    const long long mo = moff[id];
    cx* hb = Hp + (size_t)b * NN;
    cx* sb = Sp + (size_t)b * NN;
    for (int e = threadIdx.x; e < (int)NN; e += (int)blockDim.x) {
        const int i = e % N, j = e / N;
        const bool in = i < n && j < n;
        sb[e] = in ? S[mo + (long long)j * n + i] : cm(i == j ? 1.0 : 0.0, 0.0);
        hb[e] = in ? H[mo + (long long)j * n + i] : cm(i == j ? LG_PAD : 0.0, 0.0);
    }
}

extern "C" __global__ void __launch_bounds__(256)
geneig_lg_scatter(const real* __restrict__ Wp, const cx* __restrict__ Up, const int* __restrict__ Ip,
                  const int* __restrict__ ids, const long long* __restrict__ moff,
                  const long long* __restrict__ woff, const int* __restrict__ nn, int count, real* __restrict__ W,
                  cx* __restrict__ U, int* __restrict__ info) {
    using namespace gk;
    const int b = blockIdx.x;
    if (b >= count) return;
    const int id = ids[b];
    const int n = nn[id];
    const long long mo = moff[id], wo = woff[id];
    const cx* ub = Up + (size_t)b * NN;
    for (int e = threadIdx.x; e < n * n; e += (int)blockDim.x) U[mo + e] = ub[(size_t)(e / n) * N + e % n];
    for (int e = threadIdx.x; e < n; e += (int)blockDim.x) W[wo + e] = Wp[(size_t)b * N + e];
    if (threadIdx.x == 0) {
        const int ip = Ip[b];
        info[id] = (ip > N) ? n + (ip - N) : ip;
    }
}
#endif  // !GENEIG_FP32
