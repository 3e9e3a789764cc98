// CPU oracle for the batched generalized Hermitian eigensolve (test
// infrastructure only; see geneig_oracle.h). Built with g++ + libstdc++ so the
// generators reproduce the reference's fixtures bit for bit.
#include "geneig_oracle.h"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

using cd = std::complex<double>;

namespace {

inline cd* C(go_c128* p) { return reinterpret_cast<cd*>(p); }
inline const cd* C(const go_c128* p) { return reinterpret_cast<const cd*>(p); }
inline size_t idx(int n, int i, int j) { return (size_t)j * n + i; }

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------

// Hermitian H = A + A^H, A complex Gaussian; one fresh normal_distribution per
// call, column-major.
// Follows tests/common/lapack_reference.cpp:86-101.
void gen_hermitian(int n, std::mt19937_64& rng, cd* H) {
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<cd> A((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double im = g(rng);  // g++ evaluates cdouble(g(rng), g(rng))'s
            double re = g(rng);  // arguments right to left: imaginary part first
            A[idx(n, i, j)] = cd(re, im);
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            H[idx(n, i, j)] = A[idx(n, i, j)] + std::conj(A[idx(n, j, i)]);
}

// Random HPD S = A A^H + N I. Follows tests/common/lapack_reference.cpp:105-127
// (same summation order over k so the result is bit-identical).
void gen_hpd_random(int n, std::mt19937_64& rng, cd* S) {
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<cd> A((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double im = g(rng);  // g++ evaluates cdouble(g(rng), g(rng))'s
            double re = g(rng);  // arguments right to left: imaginary part first
            A[idx(n, i, j)] = cd(re, im);
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            cd acc(0.0, 0.0);
            for (int k = 0; k < n; ++k) acc += A[idx(n, i, k)] * std::conj(A[idx(n, j, k)]);
            S[idx(n, i, j)] = acc;
        }
    for (int i = 0; i < n; ++i) S[idx(n, i, i)] += cd((double)n, 0.0);
}

// Near-identity S = (1 + N/2) I + 0.1 (B + B^H).
// Follows tests/common/lapack_reference.cpp:132-150.
void gen_hpd_near_identity(int n, std::mt19937_64& rng, cd* S) {
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<cd> B((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double im = g(rng);  // g++ evaluates cdouble(g(rng), g(rng))'s
            double re = g(rng);  // arguments right to left: imaginary part first
            B[idx(n, i, j)] = cd(re, im);
        }
    const double shift = 1.0 + 0.5 * (double)n;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            cd v = 0.1 * (B[idx(n, i, j)] + std::conj(B[idx(n, j, i)]));
            if (i == j) v += cd(shift, 0.0);
            S[idx(n, i, j)] = v;
        }
}

// Unitary Q from the Householder QR of a complex Gaussian matrix (Q's columns
// are fixed up so R has a positive real diagonal, i.e. Haar-distributed Q).
void gen_unitary(int n, std::mt19937_64& rng, cd* Q) {
    std::normal_distribution<double> g(0.0, 1.0);
    std::vector<cd> A((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double im = g(rng);  // g++ evaluates cdouble(g(rng), g(rng))'s
            double re = g(rng);  // arguments right to left: imaginary part first
            A[idx(n, i, j)] = cd(re, im);
        }
    // Householder QR, reflectors kept in vs / taus.
    std::vector<cd> vs((size_t)n * n, cd(0.0));
    std::vector<cd> taus(n, cd(0.0));
    std::vector<cd> rdiag(n);
    for (int k = 0; k < n; ++k) {
        double xn2 = 0.0;
        for (int i = k; i < n; ++i) xn2 += std::norm(A[idx(n, i, k)]);
        cd alpha = A[idx(n, k, k)];
        double xn = std::sqrt(xn2);
        cd* v = &vs[(size_t)k * n];
        if (xn == 0.0) { rdiag[k] = 0.0; continue; }
        cd ph = (std::abs(alpha) > 0.0) ? alpha / std::abs(alpha) : cd(1.0);
        cd beta = -ph * xn;                    // reflect x onto beta e_k
        for (int i = k; i < n; ++i) v[i] = A[idx(n, i, k)];
        v[k] -= beta;
        double vn2 = 0.0;
        for (int i = k; i < n; ++i) vn2 += std::norm(v[i]);
        taus[k] = (vn2 > 0.0) ? cd(2.0 / vn2) : cd(0.0);
        for (int j = k; j < n; ++j) {
            cd w(0.0);
            for (int i = k; i < n; ++i) w += std::conj(v[i]) * A[idx(n, i, j)];
            w *= taus[k];
            for (int i = k; i < n; ++i) A[idx(n, i, j)] -= v[i] * w;
        }
        rdiag[k] = A[idx(n, k, k)];
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Q[idx(n, i, j)] = (i == j) ? cd(1.0) : cd(0.0);
    for (int k = n - 1; k >= 0; --k) {
        const cd* v = &vs[(size_t)k * n];
        if (taus[k] == cd(0.0)) continue;
        for (int j = 0; j < n; ++j) {
            cd w(0.0);
            for (int i = k; i < n; ++i) w += std::conj(v[i]) * Q[idx(n, i, j)];
            w *= taus[k];
            for (int i = k; i < n; ++i) Q[idx(n, i, j)] -= v[i] * w;
        }
    }
    for (int j = 0; j < n; ++j) {
        cd ph = (std::abs(rdiag[j]) > 0.0) ? rdiag[j] / std::abs(rdiag[j]) : cd(1.0);
        for (int i = 0; i < n; ++i) Q[idx(n, i, j)] *= ph;
    }
}

// X = Q diag(w) Q^H, Hermitian-symmetrised.
void congruence(int n, const cd* Q, const double* w, cd* X) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            cd acc(0.0);
            for (int k = 0; k < n; ++k) acc += Q[idx(n, i, k)] * w[k] * std::conj(Q[idx(n, j, k)]);
            X[idx(n, i, j)] = acc;
        }
    for (int j = 0; j < n; ++j) {
        X[idx(n, j, j)] = cd(X[idx(n, j, j)].real(), 0.0);
        for (int i = j + 1; i < n; ++i) {
            cd a = 0.5 * (X[idx(n, i, j)] + std::conj(X[idx(n, j, i)]));
            X[idx(n, i, j)] = a;
            X[idx(n, j, i)] = std::conj(a);
        }
    }
}

// ---------------------------------------------------------------------------
// LAPACK binding (Fortran ABI, hidden string lengths passed explicitly)
// ---------------------------------------------------------------------------
typedef void (*zhegv_t)(const int*, const char*, const char*, const int*, cd*, const int*,
                        cd*, const int*, double*, cd*, const int*, double*, int*, size_t, size_t);
typedef void (*zpotrf_t)(const char*, const int*, cd*, const int*, int*, size_t);
typedef void (*zhegst_t)(const int*, const char*, const int*, cd*, const int*, const cd*,
                         const int*, int*, size_t);
typedef void (*zheev_t)(const char*, const char*, const int*, cd*, const int*, double*, cd*,
                        const int*, double*, int*, size_t, size_t);
typedef void (*dsteqr_t)(const char*, const int*, double*, double*, double*, const int*,
                         double*, int*, size_t);
typedef void (*setthreads_t)(int);

struct Lapack {
    void* handle = nullptr;
    std::string provider;
    zhegv_t zhegv = nullptr;
    zpotrf_t zpotrf = nullptr;
    zhegst_t zhegst = nullptr;
    zheev_t zheev = nullptr;
    dsteqr_t dsteqr = nullptr;
};
Lapack g_lapack;

void* sym(void* h, const char* prefix, const char* name) {
    std::string s = std::string(prefix) + name;
    return dlsym(h, s.c_str());
}

bool try_load(const char* path) {
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return false;
    for (const char* prefix : {"", "scipy_"}) {
        auto zhegv = (zhegv_t)sym(h, prefix, "zhegv_");
        if (!zhegv) continue;
        g_lapack.handle = h;
        g_lapack.provider = std::string(path) + (prefix[0] ? " (scipy_ prefix)" : "");
        g_lapack.zhegv = zhegv;
        g_lapack.zpotrf = (zpotrf_t)sym(h, prefix, "zpotrf_");
        g_lapack.zhegst = (zhegst_t)sym(h, prefix, "zhegst_");
        g_lapack.zheev = (zheev_t)sym(h, prefix, "zheev_");
        g_lapack.dsteqr = (dsteqr_t)sym(h, prefix, "dsteqr_");
        // Pin the provider to one thread: batch parallelism is ours.
        for (const char* nm : {"scipy_openblas_set_num_threads", "openblas_set_num_threads"}) {
            auto st = (setthreads_t)dlsym(h, nm);
            if (st) { st(1); break; }
        }
        return true;
    }
    dlclose(h);
    return false;
}

bool lapack_ready() { return g_lapack.zhegv != nullptr; }

// ---------------------------------------------------------------------------
// Restatement of the GPU algorithm (FP64). The GPU kernels implement these
// same steps in the same order; see each function's note.
// ---------------------------------------------------------------------------

// Right-looking lower Cholesky S = L L^H, in place; strict upper zeroed.
// Returns 0 or k+1 for the first non-positive pivot k (0-based), matching
// zpotrf's INFO (reference call site: src/unified/kernel_source.hpp:244).
int rs_potrf(int n, cd* A) {
    for (int k = 0; k < n; ++k) {
        double dkk = A[idx(n, k, k)].real();
        if (!(dkk > 0.0)) return k + 1;
        double lkk = std::sqrt(dkk);
        double inv = 1.0 / lkk;
        A[idx(n, k, k)] = cd(lkk, 0.0);
        for (int i = k + 1; i < n; ++i) A[idx(n, i, k)] *= inv;
        for (int j = k + 1; j < n; ++j) {
            cd ljk = std::conj(A[idx(n, j, k)]);
            for (int i = j; i < n; ++i) A[idx(n, i, j)] -= A[idx(n, i, k)] * ljk;
        }
    }
    for (int j = 1; j < n; ++j)
        for (int i = 0; i < j; ++i) A[idx(n, i, j)] = cd(0.0);
    return 0;
}

// A <- L^{-1} A L^{-H}: forward substitution from the left, then from the
// right (kernel_source.hpp:246-248 order: TrsmLeft then TrsmRight).
void rs_reduce(int n, const cd* L, cd* A) {
    // Left: rows.
    for (int k = 0; k < n; ++k) {
        double inv = 1.0 / L[idx(n, k, k)].real();
        for (int j = 0; j < n; ++j) A[idx(n, k, j)] *= inv;
        for (int j = 0; j < n; ++j) {
            cd xkj = A[idx(n, k, j)];
            for (int i = k + 1; i < n; ++i) A[idx(n, i, j)] -= L[idx(n, i, k)] * xkj;
        }
    }
    // Right: columns. Solve Y L^H = A, column k of Y in increasing k.
    for (int k = 0; k < n; ++k) {
        double inv = 1.0 / L[idx(n, k, k)].real();
        for (int i = 0; i < n; ++i) A[idx(n, i, k)] *= inv;
        for (int j = k + 1; j < n; ++j) {
            cd ljk = std::conj(L[idx(n, j, k)]);
            for (int i = 0; i < n; ++i) A[idx(n, i, j)] -= A[idx(n, i, k)] * ljk;
        }
    }
}

// Householder tridiagonalisation of a Hermitian matrix held in full storage:
// Q^H A Q = T with real d, e. Reflector k (0-based) is H_k = I - tau_k v v^H,
// v[k+1] = 1, v[k+2..] stored in A(k+2.., k). Reflector generation follows the
// zlarfg convention (beta real, beta = -sign(Re alpha) * norm). The
// matrix-vector product and the rank-2 update act on the full matrix.
void rs_hetrd(int n, cd* A, double* d, double* e, cd* tau) {
    std::vector<cd> v(n), p(n);
    for (int k = 0; k + 1 < n; ++k) {
        const int m0 = k + 1;
        cd alpha = A[idx(n, m0, k)];
        double xn2 = 0.0;
        for (int i = m0 + 1; i < n; ++i) xn2 += std::norm(A[idx(n, i, k)]);
        cd t(0.0);
        double beta = alpha.real();
        if (xn2 == 0.0 && alpha.imag() == 0.0) {
            t = cd(0.0);
        } else {
            double nrm = std::sqrt(alpha.real() * alpha.real() + alpha.imag() * alpha.imag() + xn2);
            beta = (alpha.real() >= 0.0) ? -nrm : nrm;
            t = cd((beta - alpha.real()) / beta, -alpha.imag() / beta);
            cd scal = 1.0 / (alpha - beta);
            for (int i = m0 + 1; i < n; ++i) A[idx(n, i, k)] *= scal;
        }
        e[k] = beta;
        tau[k] = t;
        d[k] = A[idx(n, k, k)].real();
        for (int i = 0; i < n; ++i) v[i] = cd(0.0);
        v[m0] = cd(1.0);
        for (int i = m0 + 1; i < n; ++i) v[i] = A[idx(n, i, k)];
        if (t != cd(0.0)) {
            // p = tau * A22 v
            for (int i = m0; i < n; ++i) {
                cd acc(0.0);
                for (int j = m0; j < n; ++j) acc += A[idx(n, i, j)] * v[j];
                p[i] = t * acc;
            }
            // w = p - (tau/2) (p^H v) v
            cd pv(0.0);
            for (int i = m0; i < n; ++i) pv += std::conj(p[i]) * v[i];
            cd c = -0.5 * t * pv;
            for (int i = m0; i < n; ++i) p[i] += c * v[i];
            // A22 -= v w^H + w v^H
            for (int j = m0; j < n; ++j)
                for (int i = m0; i < n; ++i)
                    A[idx(n, i, j)] -= v[i] * std::conj(p[j]) + p[i] * std::conj(v[j]);
        }
    }
    d[n - 1] = A[idx(n, n - 1, n - 1)].real();
    if (n >= 1) tau[n - 1] = cd(0.0);
}

// Implicit QL with Wilkinson-type shift on the real symmetric tridiagonal
// (d, e), accumulating the rotations into Z (n x n, column-major, starts as I).
// Eigenvalues sorted ascending with Z's columns. Textbook method: Bowdler,
// Martin, Reinsch & Wilkinson, "The QR and QL algorithms for symmetric
// matrices" (Handbook for Automatic Computation II/3). Returns 0, or the
// number of off-diagonals that failed to converge within 30 sweeps per
// eigenvalue (zheev's INFO meaning). The GPU computes rotation lengths with a
// Newton-refined reciprocal square root instead of hypot and a division;
// results agree to rounding.
int rs_tql(int n, double* d, double* e, double* Z) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Z[idx(n, i, j)] = (i == j) ? 1.0 : 0.0;
    if (n <= 1) return 0;
    e[n - 1] = 0.0;
    const double eps = 0x1p-53;
    int info = 0;
    for (int l = 0; l < n && info == 0; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (iter++ == 30) {
                for (int i = l; i < n - 1; ++i) info += (e[i] != 0.0);
                break;
            }
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool underflow = false;
            for (; i >= l; --i) {
                double f = s * e[i];
                double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; ++k) {
                    double zf = Z[idx(n, k, i + 1)];
                    double zi = Z[idx(n, k, i)];
                    Z[idx(n, k, i + 1)] = s * zi + c * zf;
                    Z[idx(n, k, i)] = c * zi - s * zf;
                }
            }
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    // ascending order (also after a convergence failure, as the GPU does)
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < p) { k = j; p = d[j]; }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int r = 0; r < n; ++r) std::swap(Z[idx(n, r, i)], Z[idx(n, r, k)]);
        }
    }
    return info;
}

// V <- Q V with Q = H_0 H_1 ... H_{n-2} from rs_hetrd (applied last-first).
void rs_apply_q(int n, const cd* A, const cd* tau, cd* V) {
    std::vector<cd> v(n);
    for (int k = n - 2; k >= 0; --k) {
        cd t = tau[k];
        if (t == cd(0.0)) continue;
        const int m0 = k + 1;
        v[m0] = cd(1.0);
        for (int i = m0 + 1; i < n; ++i) v[i] = A[idx(n, i, k)];
        for (int j = 0; j < n; ++j) {
            cd w(0.0);
            for (int i = m0; i < n; ++i) w += std::conj(v[i]) * V[idx(n, i, j)];
            w *= t;
            for (int i = m0; i < n; ++i) V[idx(n, i, j)] -= v[i] * w;
        }
    }
}

// V <- L^{-H} V (backward substitution), kernel_source.hpp:252.
void rs_backtrsm(int n, const cd* L, cd* V) {
    for (int k = n - 1; k >= 0; --k) {
        double inv = 1.0 / L[idx(n, k, k)].real();
        for (int j = 0; j < n; ++j) V[idx(n, k, j)] *= inv;
        for (int j = 0; j < n; ++j) {
            cd ukj = V[idx(n, k, j)];
            for (int i = 0; i < k; ++i) V[idx(n, i, j)] -= std::conj(L[idx(n, k, i)]) * ukj;
        }
    }
}

int rs_zhegv(int n, const cd* H, const cd* S, double* W, cd* U) {
    const size_t nn = (size_t)n * n;
    std::vector<cd> L(S, S + nn), M(H, H + nn), tau(n);
    std::vector<double> e(n), Z(nn);
    int info = rs_potrf(n, L.data());
    if (info) {
        for (int i = 0; i < n; ++i) W[i] = NAN;
        for (size_t i = 0; i < nn; ++i) U[i] = cd(0.0);
        return n + info;
    }
    rs_reduce(n, L.data(), M.data());
    rs_hetrd(n, M.data(), W, e.data(), tau.data());
    int qinfo = rs_tql(n, W, e.data(), Z.data());
    for (size_t i = 0; i < nn; ++i) U[i] = cd(Z[i], 0.0);
    rs_apply_q(n, M.data(), tau.data(), U);
    rs_backtrsm(n, L.data(), U);
    return qinfo;
}

// ---------------------------------------------------------------------------
// Metrics
// ---------------------------------------------------------------------------
void matvec(int n, const cd* A, const cd* x, cd* y) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int j = 0; j < n; ++j) {
        cd xj = x[j];
        for (int i = 0; i < n; ++i) y[i] += A[idx(n, i, j)] * xj;
    }
}

double norm1(int n, const cd* A) {
    double best = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += std::abs(A[idx(n, i, j)]);
        best = std::max(best, s);
    }
    return best;
}

}  // namespace

extern "C" {

void go_gen_reference(int n, int batch, uint64_t seed, int s_mode, go_c128* H, go_c128* S) {
    // One mt19937_64 stream, H then S per matrix (lapack_reference.cpp:165-175).
    std::mt19937_64 rng(seed);
    const size_t nn = (size_t)n * n;
    for (int b = 0; b < batch; ++b) {
        gen_hermitian(n, rng, C(H) + nn * b);
        if (s_mode == 0) gen_hpd_random(n, rng, C(S) + nn * b);
        else             gen_hpd_near_identity(n, rng, C(S) + nn * b);
    }
}

void go_gen_spectrum(int n, int batch, uint64_t seed, int n_pairs, double pair_gap,
                     double cond_s, go_c128* H, go_c128* S) {
    std::mt19937_64 rng(seed);
    const size_t nn = (size_t)n * n;
    std::vector<cd> Q(nn), B(nn);
    std::vector<double> lam(n), sig(n);
    for (int b = 0; b < batch; ++b) {
        cd* Hb = C(H) + nn * b;
        cd* Sb = C(S) + nn * b;
        gen_unitary(n, rng, Q.data());
        std::uniform_real_distribution<double> u(-10.0, 10.0);
        for (int i = 0; i < n; ++i) lam[i] = u(rng);
        for (int k = 0; k < n_pairs && 2 * k + 1 < n; ++k) lam[2 * k + 1] = lam[2 * k] + pair_gap;
        congruence(n, Q.data(), lam.data(), Hb);
        if (cond_s > 0.0) {
            gen_unitary(n, rng, Q.data());
            for (int i = 0; i < n; ++i)
                sig[i] = (n > 1) ? std::pow(cond_s, (double)i / (double)(n - 1)) : 1.0;
            congruence(n, Q.data(), sig.data(), Sb);
        } else {
            std::normal_distribution<double> g(0.0, 1.0);
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double re = g(rng);
                    double im = g(rng);
                    B[idx(n, i, j)] = cd(re, im);
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    cd acc(0.0);
                    for (int k = 0; k < n; ++k) acc += B[idx(n, i, k)] * std::conj(B[idx(n, j, k)]);
                    Sb[idx(n, i, j)] = acc;
                }
            for (int j = 0; j < n; ++j) {
                Sb[idx(n, j, j)] = cd(Sb[idx(n, j, j)].real() + 0.05 * n, 0.0);
                for (int i = j + 1; i < n; ++i) Sb[idx(n, j, i)] = std::conj(Sb[idx(n, i, j)]);
            }
        }
    }
}

// GEIG v1: int32[6] header {magic 0x47454947, version 1, N, B, seed, 0}, then
// H, S (complex128), W (float64), U (complex128), all column-major.
// Follows tests/common/lapack_reference.cpp:257-279.
int go_geig_read_header(const char* path, int32_t hdr[6]) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return -1;
    f.read(reinterpret_cast<char*>(hdr), 6 * sizeof(int32_t));
    if (!f) return -2;
    if (hdr[0] != 0x47454947) return -3;
    if (hdr[1] != 1) return -4;
    return 0;
}

int go_geig_read(const char* path, go_c128* H, go_c128* S, double* W, go_c128* U) {
    int32_t hdr[6];
    int rc = go_geig_read_header(path, hdr);
    if (rc) return rc;
    std::ifstream f(path, std::ios::binary);
    f.seekg(6 * sizeof(int32_t));
    const size_t n = hdr[2], b = hdr[3];
    const size_t mat = n * n * b;
    f.read(reinterpret_cast<char*>(H), 16 * mat);
    f.read(reinterpret_cast<char*>(S), 16 * mat);
    f.read(reinterpret_cast<char*>(W), 8 * n * b);
    f.read(reinterpret_cast<char*>(U), 16 * mat);
    return f ? 0 : -5;
}

int go_geig_write(const char* path, int n, int batch, int seed, const go_c128* H,
                  const go_c128* S, const double* W, const go_c128* U) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return -1;
    int32_t hdr[6] = {0x47454947, 1, n, batch, seed, 0};
    f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    const size_t mat = (size_t)n * n * batch;
    f.write(reinterpret_cast<const char*>(H), 16 * mat);
    f.write(reinterpret_cast<const char*>(S), 16 * mat);
    f.write(reinterpret_cast<const char*>(W), 8 * (size_t)n * batch);
    f.write(reinterpret_cast<const char*>(U), 16 * mat);
    return f ? 0 : -2;
}

int go_lapack_load(const char* path) {
    if (lapack_ready()) return 0;
    if (path && path[0]) return try_load(path) ? 0 : -1;
    const char* env = std::getenv("GENEIG_LAPACK_LIB");
    if (env && env[0] && try_load(env)) return 0;
    for (const char* p : {"liblapack.so.3", "liblapack.so", "libopenblas.so.0", "libopenblas.so"})
        if (try_load(p)) return 0;
    return -1;
}

const char* go_lapack_provider(void) { return g_lapack.provider.c_str(); }

// zhegv(itype=1, jobz='V', uplo='L'): the reference's live oracle call,
// tests/static/test_full_batched_n54.cpp:243-247.
int go_zhegv(int n, const go_c128* H, const go_c128* S, double* W, go_c128* U) {
    if (!lapack_ready()) return -1000;
    const size_t nn = (size_t)n * n;
    std::vector<cd> B(C(S), C(S) + nn);
    std::memcpy(U, H, 16 * nn);
    int itype = 1, lda = n, info = 0, lwork = -1;
    cd wq;
    std::vector<double> rwork(std::max(1, 3 * n - 2));
    g_lapack.zhegv(&itype, "V", "L", &n, C(U), &lda, B.data(), &lda, W, &wq, &lwork,
                   rwork.data(), &info, 1, 1);
    lwork = std::max(1, (int)wq.real());
    std::vector<cd> work(lwork);
    g_lapack.zhegv(&itype, "V", "L", &n, C(U), &lda, B.data(), &lda, W, work.data(), &lwork,
                   rwork.data(), &info, 1, 1);
    return info;
}

int go_zhegv_batch(int n, int batch, const go_c128* H, const go_c128* S, double* W,
                   go_c128* U, int* info, int threads) {
    if (!lapack_ready()) return -1000;
    const size_t nn = (size_t)n * n;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
    for (int b = 0; b < batch; ++b)
        info[b] = go_zhegv(n, H + nn * b, S + nn * b, W + (size_t)n * b, U + nn * b);
    return 0;
}

int go_zpotrf(int n, go_c128* A) {
    if (!lapack_ready()) return -1000;
    int info = 0;
    g_lapack.zpotrf("L", &n, C(A), &n, &info, 1);
    for (int j = 1; j < n; ++j)
        for (int i = 0; i < j; ++i) C(A)[idx(n, i, j)] = 0.0;
    return info;
}

int go_zhegst(int n, go_c128* A, const go_c128* L) {
    if (!lapack_ready()) return -1000;
    int itype = 1, info = 0;
    g_lapack.zhegst(&itype, "L", &n, C(A), &n, C(L), &n, &info, 1);
    return info;
}

int go_zheev(int n, go_c128* A, double* W) {
    if (!lapack_ready()) return -1000;
    int info = 0, lwork = -1;
    cd wq;
    std::vector<double> rwork(std::max(1, 3 * n - 2));
    g_lapack.zheev("V", "L", &n, C(A), &n, W, &wq, &lwork, rwork.data(), &info, 1, 1);
    lwork = std::max(1, (int)wq.real());
    std::vector<cd> work(lwork);
    g_lapack.zheev("V", "L", &n, C(A), &n, W, work.data(), &lwork, rwork.data(), &info, 1, 1);
    return info;
}

int go_dsteqr(int n, double* d, double* e, double* Z) {
    if (!lapack_ready() || !g_lapack.dsteqr) return -1000;
    int info = 0;
    std::vector<double> work(std::max(1, 2 * n - 2));
    g_lapack.dsteqr("I", &n, d, e, Z, &n, work.data(), &info, 1);
    return info;
}

int go_rs_potrf(int n, go_c128* A) { return rs_potrf(n, C(A)); }
void go_rs_reduce(int n, const go_c128* L, go_c128* A) { rs_reduce(n, C(L), C(A)); }
void go_rs_hetrd(int n, go_c128* A, double* d, double* e, go_c128* tau) {
    rs_hetrd(n, C(A), d, e, C(tau));
}
int go_rs_tql(int n, double* d, double* e, double* Z) { return rs_tql(n, d, e, Z); }
void go_rs_apply_q(int n, const go_c128* A, const go_c128* tau, go_c128* V) {
    rs_apply_q(n, C(A), C(tau), C(V));
}
void go_rs_backtrsm(int n, const go_c128* L, go_c128* V) { rs_backtrsm(n, C(L), C(V)); }
int go_rs_zhegv(int n, const go_c128* H, const go_c128* S, double* W, go_c128* U) {
    return rs_zhegv(n, C(H), C(S), W, C(U));
}
void go_rs_zhegv_batch(int n, int batch, const go_c128* H, const go_c128* S, double* W,
                       go_c128* U, int* info, int threads) {
    const size_t nn = (size_t)n * n;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4)
    for (int b = 0; b < batch; ++b)
        info[b] = rs_zhegv(n, C(H) + nn * b, C(S) + nn * b, W + (size_t)n * b, C(U) + nn * b);
}

// eig_rel and phase_err as in tests/static/test_full_batched_n54.cpp:99-124
// (S-inner-product phase, columns with gap < degen_gap excluded).
void go_compare(int n, const go_c128* S, const double* W_ref, const go_c128* U_ref,
                const double* W, const go_c128* U, double degen_gap, double out[3]) {
    double eig_rel = 0.0, phase = 0.0;
    int excluded = 0;
    for (int i = 0; i < n; ++i) {
        double scale = std::max(std::fabs(W_ref[i]), 1.0);
        double rel = std::fabs(W[i] - W_ref[i]) / scale;
        if (!(rel == rel)) rel = INFINITY;
        eig_rel = std::max(eig_rel, rel);
    }
    std::vector<cd> Sv(n);
    for (int i = 0; i < n; ++i) {
        matvec(n, C(S), C(U_ref) + (size_t)i * n, Sv.data());
        cd inner(0.0);
        const cd* ug = C(U) + (size_t)i * n;
        for (int k = 0; k < n; ++k) inner += std::conj(ug[k]) * Sv[k];
        double pe = 1.0 - std::abs(inner);
        if (!(pe == pe)) pe = INFINITY;
        double gl = (i > 0) ? std::fabs(W_ref[i] - W_ref[i - 1]) : INFINITY;
        double gr = (i < n - 1) ? std::fabs(W_ref[i + 1] - W_ref[i]) : INFINITY;
        if (std::min(gl, gr) < degen_gap) ++excluded;
        else phase = std::max(phase, pe);
    }
    out[0] = eig_rel;
    out[1] = phase;
    out[2] = excluded;
}

double go_residual(int n, const go_c128* H, const go_c128* S, const double* W, const go_c128* U) {
    const size_t nn = (size_t)n * n;
    std::vector<cd> R(nn), hu(n), su(n);
    for (int j = 0; j < n; ++j) {
        const cd* u = C(U) + (size_t)j * n;
        matvec(n, C(H), u, hu.data());
        matvec(n, C(S), u, su.data());
        for (int i = 0; i < n; ++i) R[idx(n, i, j)] = hu[i] - W[j] * su[i];
    }
    double den = norm1(n, C(H)) * norm1(n, C(U));
    double r = norm1(n, R.data());
    double out = den > 0 ? r / den : r;
    return (out == out) ? out : INFINITY;
}

double go_s_orth(int n, const go_c128* S, const go_c128* U) {
    const size_t nn = (size_t)n * n;
    std::vector<cd> SU(nn), G(nn);
    for (int j = 0; j < n; ++j) matvec(n, C(S), C(U) + (size_t)j * n, SU.data() + (size_t)j * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            cd acc(0.0);
            const cd* ui = C(U) + (size_t)i * n;
            const cd* suj = SU.data() + (size_t)j * n;
            for (int k = 0; k < n; ++k) acc += std::conj(ui[k]) * suj[k];
            G[idx(n, i, j)] = acc - ((i == j) ? cd(1.0) : cd(0.0));
        }
    double r = norm1(n, G.data());
    return (r == r) ? r : INFINITY;
}

void go_check_batch(int n, int batch, const go_c128* H, const go_c128* S, const double* W_ref,
                    const go_c128* U_ref, const double* W, const go_c128* U, double degen_gap,
                    int threads, double* out) {
    const size_t nn = (size_t)n * n;
    if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 2)
    for (int b = 0; b < batch; ++b) {
        double* o = out + 5 * (size_t)b;
        if (W_ref && U_ref) {
            go_compare(n, S + nn * b, W_ref + (size_t)n * b, U_ref + nn * b, W + (size_t)n * b,
                       U + nn * b, degen_gap, o);
        } else {
            o[0] = o[1] = o[2] = 0.0;
        }
        o[3] = go_residual(n, H + nn * b, S + nn * b, W + (size_t)n * b, U + nn * b);
        o[4] = go_s_orth(n, S + nn * b, U + nn * b);
    }
}

}  // extern "C"
