/* CPU oracle for the batched generalized Hermitian eigensolve H u = lambda S u.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing on the product path (libsunny_geneig.so,
 * the Python package) links, loads or calls this library. It is the checker:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Contents (each function cites the reference file:line it follows):
 *   - input generators (bit-exact with the reference's fixture generator),
 *   - GEIG v1 reader / writer,
 *   - an FP64 LAPACK binding (zhegv, zpotrf, zhegst, zheev, dsteqr) loaded by
 *     dlopen from any provider (system liblapack / OpenBLAS, or scipy's bundled
 *     OpenBLAS with its scipy_ symbol prefix),
 *   - a plain-C++ restatement of the GPU algorithm, stage by stage,
 *   - the comparison metrics used by the reference's tests plus the residual
 *     and S-orthogonality norms.
 *
 * Matrices are column-major complex128 (std::complex<double> / two doubles),
 * element (i,j) of matrix b at b*N*N + j*N + i, leading dimension N.
 */
#ifndef GENEIG_ORACLE_H
#define GENEIG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } go_c128;

/* ---- generators ---------------------------------------------------------- */
/* s_mode: 0 = random HPD (A A^H + N I), 1 = near identity. */
void go_gen_reference(int n, int batch, uint64_t seed, int s_mode,
                      go_c128* H, go_c128* S);
/* Spectrum-controlled generator: H = Q diag(lambda) Q^H, lambda ~ U[-10,10]
 * with `n_pairs` near-degenerate pairs split by `pair_gap`; S = B B^H + 0.05 N I
 * (cond_s <= 0) or S = Q_S diag(sigma) Q_S^H with sigma geometric in
 * [1, cond_s] (cond_s > 0). */
void go_gen_spectrum(int n, int batch, uint64_t seed, int n_pairs,
                     double pair_gap, double cond_s, go_c128* H, go_c128* S);

/* ---- GEIG v1 file format ------------------------------------------------ */
int go_geig_read_header(const char* path, int32_t hdr[6]);
int go_geig_read(const char* path, go_c128* H, go_c128* S, double* W, go_c128* U);
int go_geig_write(const char* path, int n, int batch, int seed,
                  const go_c128* H, const go_c128* S, const double* W,
                  const go_c128* U);

/* ---- LAPACK ----------------------------------------------------------------
 * go_lapack_load(path): path == NULL or "" tries the default providers.
 * Returns 0 on success. go_lapack_provider() names what was loaded. */
int         go_lapack_load(const char* path);
const char* go_lapack_provider(void);
int go_zhegv(int n, const go_c128* H, const go_c128* S, double* W, go_c128* U);
int go_zhegv_batch(int n, int batch, const go_c128* H, const go_c128* S,
                   double* W, go_c128* U, int* info, int threads);
int go_zpotrf(int n, go_c128* A);                       /* lower, in place  */
int go_zhegst(int n, go_c128* A, const go_c128* L);     /* itype 1, lower   */
int go_zheev(int n, go_c128* A, double* W);             /* jobz V, lower    */
int go_dsteqr(int n, double* d, double* e, double* Z);  /* compz I          */

/* ---- restatement of the GPU algorithm ------------------------------------ */
int  go_rs_potrf(int n, go_c128* A);
void go_rs_reduce(int n, const go_c128* L, go_c128* A);
void go_rs_hetrd(int n, go_c128* A, double* d, double* e, go_c128* tau);
int  go_rs_tql(int n, double* d, double* e, double* Z);
void go_rs_apply_q(int n, const go_c128* A, const go_c128* tau, go_c128* V);
void go_rs_backtrsm(int n, const go_c128* L, go_c128* V);
int  go_rs_zhegv(int n, const go_c128* H, const go_c128* S, double* W, go_c128* U);
void go_rs_zhegv_batch(int n, int batch, const go_c128* H, const go_c128* S,
                       double* W, go_c128* U, int* info, int threads);

/* ---- metrics --------------------------------------------------------------
 * out[0] eig_rel   = max_i |W-W_ref| / max(|W_ref|,1)
 * out[1] phase_err = max_i (1 - |u_i^H S u_ref_i|) over non-degenerate i
 * out[2] number of columns excluded (gap < degen_gap)
 * out[3] residual  = ||H U - S U diag(W)||_1 / (||H||_1 ||U||_1)
 * out[4] s_orth    = ||U^H S U - I||_1                                        */
void   go_compare(int n, const go_c128* S, const double* W_ref, const go_c128* U_ref,
                  const double* W, const go_c128* U, double degen_gap, double out[3]);
double go_residual(int n, const go_c128* H, const go_c128* S, const double* W,
                   const go_c128* U);
double go_s_orth(int n, const go_c128* S, const go_c128* U);
void   go_check_batch(int n, int batch, const go_c128* H, const go_c128* S,
                      const double* W_ref, const go_c128* U_ref,
                      const double* W, const go_c128* U, double degen_gap,
                      int threads, double* out /* batch x 5 */);

#ifdef __cplusplus
}
#endif
#endif
