/*
 * oracle.h -- CPU restatement of the reference LibAccInt CPU path for the
 * Rys ERI + Fock (J/K) build.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker that the HIP path is compared against (tests/, the smoke check and
 * the bench's CPU-baseline leg).  No product code links or calls it.
 *
 * Every function cites the reference file:line it follows.  Plain C99 +
 * OpenMP, FP64 throughout.
 */
#ifndef LAI_ORACLE_H
#define LAI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_L 4          /* shells up to g */
#define ORACLE_MAX_ROOTS 16     /* MAX_RYS_ROOTS, rys_quadrature.hpp:35 */

/* A basis: shells in input order, coefficients normalised as in
 * src/shared/basis/shell.cpp:32-113.  Function offsets are assigned in shell
 * order (src/shared/basis/basis_set.cpp:161-185). */
typedef struct oracle_basis {
    int n_shells;
    int nbf;
    int *L;          /* [n_shells] */
    int *nprim;      /* [n_shells] (zero-coefficient primitives kept) */
    int *prim_off;   /* [n_shells] */
    int *func_off;   /* [n_shells] */
    double *center;  /* [3*n_shells] */
    double *exps;    /* [sum nprim] */
    double *coefs;   /* [sum nprim], normalised */
} oracle_basis;

/* ---- math ------------------------------------------------------------- */
void oracle_boys_array(int n_max, double T, double *out);
void oracle_rys(int n_roots, double T, double *roots, double *weights);

/* ---- basis ------------------------------------------------------------ */
oracle_basis *oracle_basis_create(int n_shells, const int *L, const double *centers,
                                  const int *nprim, const double *exps,
                                  const double *raw_coefs);
/* As oracle_basis_create with the coefficients taken as given (already
 * normalised), so that a zero-exponent unit s function can stand in as the
 * fourth shell of a three- or two-centre integral: (ab|P u) = (ab|P), the
 * integral of src/host/kernels/three_center_eri_kernel.cpp and
 * two_center_coulomb_kernel.cpp (density fitting). */
oracle_basis *oracle_basis_create_normalized(int n_shells, const int *L, const double *centers,
                                             const int *nprim, const double *exps,
                                             const double *coefs);
void oracle_basis_free(oracle_basis *b);
int oracle_basis_nbf(const oracle_basis *b);
/* copy out normalised coefficients and function offsets */
void oracle_basis_export(const oracle_basis *b, double *coefs, int *func_off);

/* ---- integrals -------------------------------------------------------- */
/* (ab|cd) block for shells a,b,c,d, row-major [na][nb][nc][nd] */
void oracle_eri_shell_quartet(const oracle_basis *b, int sa, int sb, int sc, int sd,
                              double *out);
/* full nbf^4 ERI tensor (small bases only) */
void oracle_eri_tensor(const oracle_basis *b, double *out);
/* one-electron matrices nbf x nbf; charges as (Z, x, y, z) arrays */
void oracle_overlap(const oracle_basis *b, double *S);
void oracle_kinetic(const oracle_basis *b, double *T);
void oracle_nuclear(const oracle_basis *b, int n_charges, const double *Z,
                    const double *xyz, double *V);

/* ---- screening -------------------------------------------------------- */
/* Q_ij = sqrt(max |(ij|ij)|), packed i<=j at j*(j+1)/2 + i... as reference:
 * index = packed_index(i,j) with i<=j (schwarz_bounds.hpp), see oracle.c */
void oracle_schwarz(const oracle_basis *b, double *Q_packed, int n_threads);
/* Number of canonical quartets passing Q_ij*Q_kl >= thr (thr <= 0: all), in
 * the reference's canonical loop (include/libaccint/engine/engine.hpp:1815-1819),
 * and the list checksum: sum over quartets of splitmix64 of the packed
 * (i, j, k, l) (16 bits each), modulo 2^64. */
int64_t oracle_screened_list(const oracle_basis *b, const double *Q_packed, double thr,
                             uint64_t *hash_out);

/* ---- Fock build ------------------------------------------------------- */
/* Canonical screened J/K (engine.hpp:1805-1861 loop + fock_builder.cpp:123-222
 * accumulate_symmetric).  thr <= 0 disables screening (Q may be NULL).
 * Parallelised over shell pairs with thread-local J/K (harness-level). */
void oracle_jk_canonical(const oracle_basis *b, const double *D, const double *Q_packed,
                         double thr, double *J, double *K, int n_threads);
/* Reference default path: all n_shells^4 quartets with accumulate()
 * (cpu_engine.hpp:553-622, fock_builder.cpp:64-121), dynamic,8 scheduling,
 * thread-local J/K.  Optional bounds restrict the first-shell range
 * [i_begin, i_end) so a bounded sample can be timed. */
void oracle_jk_full(const oracle_basis *b, const double *D, double *J, double *K,
                    int n_threads, int i_begin, int i_end);

/* Canonical screened J/K over every stride-th bra pair starting at offset;
 * returns the quartet and ERI counts (a bounded sample for CPU timing). */
void oracle_jk_canonical_sample(const oracle_basis *b, const double *D, const double *Q_packed,
                                double thr, int stride, int offset, double *J, double *K,
                                int n_threads, long long *n_quartets, long long *n_eri);

#ifdef __cplusplus
}
#endif
#endif
