/*
 * oracle.c -- CPU restatement of the reference LibAccInt CPU algorithm for the
 * Rys ERI + J/K path.  TEST INFRASTRUCTURE ONLY (the checker); see oracle.h.
 *
 * Sources restated (paths relative to the reference repository):
 *   Boys:     src/shared/math/boys_function.cpp:72-128,168-192,232-303
 *   Rys:      src/shared/math/rys_quadrature.cpp:24-69,75-163,169-225,
 *             238-365,455-513
 *   Gaussian product: src/shared/math/gaussian_product.cpp:13-30
 *   Normalisation: src/shared/basis/shell.cpp:32-113,
 *             include/libaccint/kernels/norm_correction.hpp:28-34
 *   Cartesian order: include/libaccint/math/cartesian_indices.hpp:26-35
 *   ERI:      src/host/kernels/eri_kernel.cpp:104-381
 *   S/T/V:    src/host/kernels/overlap_kernel.cpp:38-170,
 *             kinetic_kernel.cpp:92-215, nuclear_kernel.cpp:51-240
 *   Fock:     src/shared/consumers/fock_builder.cpp:64-222
 *   Schwarz:  src/shared/screening/schwarz_bounds.cpp:20-59
 *   Loops:    include/libaccint/engine/engine.hpp:1815-1861,
 *             include/libaccint/engine/cpu_engine.hpp:553-622
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "boys_cheb_table.h"

#define PI_ 3.14159265358979323846
#define SQRT_PI_ 1.77245385090551602729

/* ======================================================================== */
/* Boys function                                                              */
/* ======================================================================== */

static double dfact_table(int n) /* (2n-1)!!, boys_function.cpp:27-45 */
{
    double r = 1.0;
    for (int k = 1; k <= n; ++k) r *= (double)(2 * k - 1);
    return r;
}

/* boys_function.cpp:72-91 */
static double boys_asymptotic(int n, double T)
{
    const double df = dfact_table(n);
    const double p2 = pow(2.0, n + 1);
    const double tp = pow(T, n + 0.5);
    return df / p2 * SQRT_PI_ / tp;
}

/* boys_function.cpp:150-170, Clenshaw with halved c0 */
static double clenshaw(const double *c, int N, double x)
{
    double d1 = 0.0, d2 = 0.0;
    for (int j = N - 1; j >= 1; --j) {
        double dk = 2.0 * x * d1 - d2 + c[j];
        d2 = d1;
        d1 = dk;
    }
    return c[0] + x * d1 - d2;
}

static double boys_cheb(int n, double T) /* boys_function.cpp:168-192 */
{
    int iv = (int)(T / 1.0);
    if (iv >= ORACLE_BOYS_N_INTERVALS) iv = ORACLE_BOYS_N_INTERVALS - 1;
    const double a = iv * 1.0, b = a + 1.0;
    const double x = (2.0 * T - a - b) / (b - a);
    return clenshaw(ORACLE_BOYS_CHEB[iv][n], ORACLE_BOYS_N_TERMS, x);
}

static double boys_taylor(int n, double T) /* boys_function.cpp:232-249 */
{
    double result = 1.0 / (2 * n + 1);
    double term = 1.0;
    for (int k = 1; k <= 20; ++k) {
        term *= -T / k;
        double contribution = term / (2 * n + 2 * k + 1);
        result += contribution;
        if (fabs(contribution) < 1e-16 * fabs(result)) break;
    }
    return result;
}

/* boys_function.cpp:253-303 */
void oracle_boys_array(int n_max, double T, double *out)
{
    if (T == 0.0) {
        for (int n = 0; n <= n_max; ++n) out[n] = 1.0 / (2 * n + 1);
        return;
    }
    if (T < 1e-14) {
        for (int n = 0; n <= n_max; ++n) out[n] = boys_taylor(n, T);
        return;
    }
    if (T >= 36.0) { /* boys_asymptotic_array, :95-128 */
        out[0] = boys_asymptotic(0, T);
        if (n_max == 0) return;
        const double e = exp(-T);
        const double inv2T = 1.0 / (2.0 * T);
        for (int n = 1; n <= n_max; ++n) out[n] = ((2 * n - 1) * out[n - 1] - e) * inv2T;
        return;
    }
    out[n_max] = boys_cheb(n_max, T);
    if (n_max == 0) return;
    const double e = exp(-T);
    const double twoT = 2.0 * T;
    for (int n = n_max; n >= 1; --n) out[n - 1] = (twoT * out[n] + e) / (2 * n - 1);
}

/* ======================================================================== */
/* Rys roots and weights                                                      */
/* ======================================================================== */

/* modified Chebyshev, rys_quadrature.cpp:24-69 */
static void rys_chebyshev(int n, const double *mom, double *alpha, double *beta)
{
    const int n2 = 2 * n;
    double sigma[3][2 * ORACLE_MAX_ROOTS];
    memset(sigma, 0, sizeof(sigma));
    int rm2 = 0, rm1 = 1, rc = 2;
    for (int k = 0; k < n2; ++k) sigma[rm1][k] = mom[k];
    alpha[0] = mom[1] / mom[0];
    beta[0] = mom[0];
    for (int l = 1; l < n; ++l) {
        for (int k = l; k < n2 - l; ++k)
            sigma[rc][k] = sigma[rm1][k + 1] - alpha[l - 1] * sigma[rm1][k] - beta[l - 1] * sigma[rm2][k];
        alpha[l] = sigma[rc][l + 1] / sigma[rc][l] - sigma[rm1][l] / sigma[rm1][l - 1];
        beta[l] = sigma[rc][l] / sigma[rm1][l - 1];
        int t = rm2;
        rm2 = rm1;
        rm1 = rc;
        rc = t;
    }
}

/* implicit QL with first-row eigenvector accumulation, rys_quadrature.cpp:75-163 */
static void tridiag_ql(int n, double *d, double *e, double *z)
{
    if (n == 1) return;
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            while (m < n - 1) {
                double tst = fabs(d[m]) + fabs(d[m + 1]);
                if (tst + fabs(e[m]) == tst) break;
                ++m;
            }
            if (m == l) break;
            if (++iter > 50) break;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            for (i = m - 1; i >= l; --i) {
                double f = s * e[i];
                double bb = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r < 1e-300) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * bb;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - bb;
                f = z[i + 1];
                z[i + 1] = s * z[i] + c * f;
                z[i] = c * z[i] - s * f;
            }
            if (r < 1e-300 && i >= l) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    for (int i = 0; i < n - 1; ++i) { /* selection sort ascending, :142-162 */
        int k = i;
        double pk = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < pk) {
                k = j;
                pk = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = pk;
            double t = z[k];
            z[k] = z[i];
            z[i] = t;
        }
    }
}

/* rys_quadrature.cpp:169-197 */
static void rys_poly_eval(int n, double u, const double *alpha, const double *beta,
                          double *value, double *deriv)
{
    double pp = 1.0, pc = u - alpha[0], dpp = 0.0, dpc = 1.0;
    if (n == 1) {
        *value = pc;
        *deriv = dpc;
        return;
    }
    for (int k = 1; k < n; ++k) {
        double um = u - alpha[k];
        double pn = um * pc - beta[k] * pp;
        double dpn = pc + um * dpc - beta[k] * dpp;
        pp = pc;
        pc = pn;
        dpp = dpc;
        dpc = dpn;
    }
    *value = pc;
    *deriv = dpc;
}

/* rys_quadrature.cpp:203-225 */
static void rys_newton_polish(int n, const double *alpha, const double *beta, double *roots,
                              int max_iter)
{
    for (int i = 0; i < n; ++i) {
        double u = roots[i];
        for (int it = 0; it < max_iter; ++it) {
            double v, dv;
            rys_poly_eval(n, u, alpha, beta, &v, &dv);
            if (fabs(dv) < 1e-300) break;
            double delta = v / dv;
            u -= delta;
            if (fabs(delta) <= fabs(u) * 2.0 * 2.220446049250313e-16) break;
        }
        roots[i] = u;
    }
}

/* Gauss-Legendre on [0,1] via Golub-Welsch, rys_quadrature.cpp:238-265.
 * Only M = 100 is reachable for n <= 16 (M = max(100, 6n)); computed once. */
#define GL_M 100
static double gl_nodes[GL_M], gl_wts[GL_M];
static void gl_init(void) __attribute__((constructor));
static void gl_init(void)
{
    double d[GL_M], e[GL_M], z[GL_M];
    for (int i = 0; i < GL_M; ++i) {
        d[i] = 0.0;
        e[i] = 0.0;
        z[i] = (i == 0) ? 1.0 : 0.0;
    }
    for (int i = 0; i < GL_M - 1; ++i) {
        int k = i + 1;
        double bk = (double)(k * k) / (4.0 * k * k - 1.0);
        e[i] = sqrt(bk);
    }
    tridiag_ql(GL_M, d, e, z);
    for (int i = 0; i < GL_M; ++i) {
        gl_nodes[i] = (d[i] + 1.0) / 2.0;
        gl_wts[i] = z[i] * z[i];
    }
}

/* discretised Stieltjes, rys_quadrature.cpp:299-365 */
static void rys_stieltjes(int n, double T, double *alpha, double *beta)
{
    const int M = GL_M;
    double u[GL_M], w[GL_M], pprev[GL_M], pcur[GL_M], pnext[GL_M];
    for (int j = 0; j < M; ++j) {
        double t = gl_nodes[j];
        u[j] = t * t;
        w[j] = gl_wts[j] * exp(-T * u[j]);
        pprev[j] = 0.0;
        pcur[j] = 1.0;
    }
    double norm_prev = 0.0;
    for (int j = 0; j < M; ++j) norm_prev += w[j];
    beta[0] = norm_prev;
    double up = 0.0;
    for (int j = 0; j < M; ++j) up += w[j] * u[j];
    alpha[0] = up / norm_prev;
    for (int k = 1; k < n; ++k) {
        for (int j = 0; j < M; ++j) {
            pnext[j] = (u[j] - alpha[k - 1]) * pcur[j];
            if (k >= 2) pnext[j] -= beta[k - 1] * pprev[j];
        }
        double nc = 0.0;
        for (int j = 0; j < M; ++j) nc += w[j] * pnext[j] * pnext[j];
        beta[k] = nc / norm_prev;
        up = 0.0;
        for (int j = 0; j < M; ++j) up += w[j] * u[j] * pnext[j] * pnext[j];
        alpha[k] = up / nc;
        norm_prev = nc;
        memcpy(pprev, pcur, sizeof(double) * M);
        memcpy(pcur, pnext, sizeof(double) * M);
    }
}

/* rys_quadrature.cpp:412-437 */
static void rys_from_coeffs(int n, const double *alpha, const double *beta, double *roots,
                            double *weights)
{
    double d[ORACLE_MAX_ROOTS], e[ORACLE_MAX_ROOTS], z[ORACLE_MAX_ROOTS];
    for (int i = 0; i < n; ++i) {
        d[i] = alpha[i];
        z[i] = (i == 0) ? 1.0 : 0.0;
    }
    for (int i = 0; i < n - 1; ++i) e[i] = sqrt(fabs(beta[i + 1]));
    e[n - 1] = 0.0;
    tridiag_ql(n, d, e, z);
    for (int i = 0; i < n; ++i) {
        roots[i] = d[i];
        weights[i] = beta[0] * z[i] * z[i];
    }
}

/* rys_quadrature.cpp:455-513 */
void oracle_rys(int n, double T, double *roots, double *weights)
{
    if (n == 1) {
        double m[2];
        oracle_boys_array(1, T, m);
        roots[0] = m[1] / m[0];
        weights[0] = m[0];
        return;
    }
    double mom[2 * ORACLE_MAX_ROOTS], alpha[ORACLE_MAX_ROOTS], beta[ORACLE_MAX_ROOTS];
    oracle_boys_array(2 * n - 1, T, mom);
    rys_chebyshev(n, mom, alpha, beta);
    int ok = 1;
    for (int i = 1; i < n; ++i)
        if (beta[i] <= 0.0) {
            ok = 0;
            break;
        }
    if (ok) {
        rys_from_coeffs(n, alpha, beta, roots, weights);
        int valid = 1;
        for (int i = 0; i < n; ++i)
            if (roots[i] <= 0.0 || roots[i] >= 1.0) valid = 0;
        if (valid) {
            rys_newton_polish(n, alpha, beta, roots, 10);
            for (int i = 0; i < n; ++i) {
                roots[i] = fmax(1e-14, fmin(roots[i], 1.0 - 1e-14));
                weights[i] = fmax(0.0, weights[i]);
            }
            return;
        }
    }
    rys_stieltjes(n, T, alpha, beta);
    rys_from_coeffs(n, alpha, beta, roots, weights);
    rys_newton_polish(n, alpha, beta, roots, 10);
    for (int i = 0; i < n; ++i) {
        roots[i] = fmax(1e-14, fmin(roots[i], 1.0 - 1e-14));
        weights[i] = fmax(0.0, weights[i]);
    }
}

/* ======================================================================== */
/* Basis                                                                      */
/* ======================================================================== */

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static int dfo(int n) /* double_factorial_odd, normalization.hpp:33-42 */
{
    if (n <= 0) return 1;
    int r = 1;
    for (int i = 1; i <= n; ++i) r *= (2 * i - 1);
    return r;
}

/* cartesian_indices.hpp:26-35: lx descending, then ly descending */
static void cart_indices(int L, int (*idx)[3])
{
    int c = 0;
    for (int lx = L; lx >= 0; --lx)
        for (int ly = L - lx; ly >= 0; --ly) {
            idx[c][0] = lx;
            idx[c][1] = ly;
            idx[c][2] = L - lx - ly;
            ++c;
        }
}

static double norm_corr(int lx, int ly, int lz) /* norm_correction.hpp:28-34 */
{
    const double den = (double)(dfo(lx) * dfo(ly) * dfo(lz));
    return 1.0 / sqrt(den);
}

oracle_basis *oracle_basis_create(int n_shells, const int *L, const double *centers,
                                  const int *nprim, const double *exps, const double *raw)
{
    oracle_basis *b = (oracle_basis *)calloc(1, sizeof(oracle_basis));
    b->n_shells = n_shells;
    b->L = (int *)malloc(sizeof(int) * n_shells);
    b->nprim = (int *)malloc(sizeof(int) * n_shells);
    b->prim_off = (int *)malloc(sizeof(int) * n_shells);
    b->func_off = (int *)malloc(sizeof(int) * n_shells);
    b->center = (double *)malloc(sizeof(double) * 3 * n_shells);
    int np = 0, nbf = 0;
    for (int s = 0; s < n_shells; ++s) {
        b->L[s] = L[s];
        b->nprim[s] = nprim[s];
        b->prim_off[s] = np;
        b->func_off[s] = nbf;
        np += nprim[s];
        nbf += ncart(L[s]);
        for (int k = 0; k < 3; ++k) b->center[3 * s + k] = centers[3 * s + k];
    }
    b->nbf = nbf;
    b->exps = (double *)malloc(sizeof(double) * (np > 0 ? np : 1));
    b->coefs = (double *)malloc(sizeof(double) * (np > 0 ? np : 1));
    memcpy(b->exps, exps, sizeof(double) * np);
    /* shell.cpp:32-79: primitive normalisation then contracted self-overlap */
    for (int s = 0; s < n_shells; ++s) {
        const int l = L[s], K = nprim[s], o = b->prim_off[s];
        double *c = b->coefs + o;
        const double *a = b->exps + o;
        for (int i = 0; i < K; ++i) {
            const double t1 = pow(2.0 * a[i] / PI_, 0.75);
            const double t2 = pow(4.0 * a[i], l * 0.5);
            const double t3 = sqrt((double)dfo(l));
            c[i] = raw[o + i] * (t1 * t2 / t3);
        }
        double so = 0.0;
        for (int i = 0; i < K; ++i)
            for (int j = 0; j < K; ++j) {
                const double p = a[i] + a[j];
                double ov = pow(PI_ / p, 1.5);
                for (int k = 0; k < l; ++k) ov *= 0.5 / p;
                so += c[i] * c[j] * ov;
            }
        const double scale = 1.0 / sqrt(so);
        for (int i = 0; i < K; ++i) c[i] *= scale;
    }
    return b;
}

oracle_basis *oracle_basis_create_normalized(int n_shells, const int *L, const double *centers,
                                             const int *nprim, const double *exps,
                                             const double *coefs)
{
    oracle_basis *b = oracle_basis_create(n_shells, L, centers, nprim, exps, coefs);
    int np = 0;
    for (int s = 0; s < n_shells; ++s) np += nprim[s];
    memcpy(b->coefs, coefs, sizeof(double) * np);
    return b;
}

void oracle_basis_free(oracle_basis *b)
{
    if (!b) return;
    free(b->L);
    free(b->nprim);
    free(b->prim_off);
    free(b->func_off);
    free(b->center);
    free(b->exps);
    free(b->coefs);
    free(b);
}

int oracle_basis_nbf(const oracle_basis *b) { return b->nbf; }

void oracle_basis_export(const oracle_basis *b, double *coefs, int *func_off)
{
    int np = 0;
    for (int s = 0; s < b->n_shells; ++s) np += b->nprim[s];
    if (coefs) memcpy(coefs, b->coefs, sizeof(double) * np);
    if (func_off) memcpy(func_off, b->func_off, sizeof(int) * b->n_shells);
}

/* ======================================================================== */
/* ERI (eri_kernel.cpp:104-381)                                               */
/* ======================================================================== */

#define DA_MAX (2 * ORACLE_MAX_L + 1)
#define DB_MAX (ORACLE_MAX_L + 1)
#define TAB_MAX (DA_MAX * DB_MAX * DA_MAX * DB_MAX)

/* eri_kernel.cpp:104-164 */
static void build_2d(int La, int Lb, int Lc, int Ld, double PA, double QC, double AB, double CD,
                     double B10, double B01, double B00, double *I, int da, int db, int dc, int dd)
{
    const int sa = db * dc * dd, sb = dc * dd, sc = dd;
#define IX(a, b, c, d) I[(a) * sa + (b) * sb + (c) * sc + (d)]
    memset(I, 0, sizeof(double) * da * db * dc * dd);
    IX(0, 0, 0, 0) = 1.0;
    for (int a = 0; a < La + Lb; ++a) {
        IX(a + 1, 0, 0, 0) = PA * IX(a, 0, 0, 0);
        if (a > 0) IX(a + 1, 0, 0, 0) += (double)a * B10 * IX(a - 1, 0, 0, 0);
    }
    for (int c = 0; c < Lc + Ld; ++c)
        for (int a = 0; a <= La + Lb; ++a) {
            IX(a, 0, c + 1, 0) = QC * IX(a, 0, c, 0);
            if (c > 0) IX(a, 0, c + 1, 0) += (double)c * B01 * IX(a, 0, c - 1, 0);
            if (a > 0) IX(a, 0, c + 1, 0) += (double)a * B00 * IX(a - 1, 0, c, 0);
        }
    for (int b = 0; b < Lb; ++b)
        for (int a = 0; a <= La + Lb - b - 1; ++a)
            for (int c = 0; c <= Lc + Ld; ++c)
                for (int d = 0; d < dd; ++d) IX(a, b + 1, c, d) = IX(a + 1, b, c, d) + AB * IX(a, b, c, d);
    for (int d = 0; d < Ld; ++d)
        for (int a = 0; a <= La; ++a)
            for (int b = 0; b <= Lb; ++b)
                for (int c = 0; c <= Lc + Ld - d - 1; ++c)
                    IX(a, b, c, d + 1) = IX(a, b, c + 1, d) + CD * IX(a, b, c, d);
#undef IX
}

typedef struct {
    double x, y, z;
} pt3;

void oracle_eri_shell_quartet(const oracle_basis *bs, int s0, int s1, int s2, int s3, double *out)
{
    const int La = bs->L[s0], Lb = bs->L[s1], Lc = bs->L[s2], Ld = bs->L[s3];
    const int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
    const int nout = na * nb * nc * nd;
    for (int i = 0; i < nout; ++i) out[i] = 0.0;
    const pt3 A = {bs->center[3 * s0], bs->center[3 * s0 + 1], bs->center[3 * s0 + 2]};
    const pt3 B = {bs->center[3 * s1], bs->center[3 * s1 + 1], bs->center[3 * s1 + 2]};
    const pt3 C = {bs->center[3 * s2], bs->center[3 * s2 + 1], bs->center[3 * s2 + 2]};
    const pt3 D = {bs->center[3 * s3], bs->center[3 * s3 + 1], bs->center[3 * s3 + 2]};
    const double ABx = A.x - B.x, ABy = A.y - B.y, ABz = A.z - B.z;
    const double CDx = C.x - D.x, CDy = C.y - D.y, CDz = C.z - D.z;
    int ia[15][3], ib[15][3], ic[15][3], id[15][3];
    cart_indices(La, ia);
    cart_indices(Lb, ib);
    cart_indices(Lc, ic);
    cart_indices(Ld, id);
    double ca_[15], cb_[15], cc_[15], cd_[15];
    for (int i = 0; i < na; ++i) ca_[i] = norm_corr(ia[i][0], ia[i][1], ia[i][2]);
    for (int i = 0; i < nb; ++i) cb_[i] = norm_corr(ib[i][0], ib[i][1], ib[i][2]);
    for (int i = 0; i < nc; ++i) cc_[i] = norm_corr(ic[i][0], ic[i][1], ic[i][2]);
    for (int i = 0; i < nd; ++i) cd_[i] = norm_corr(id[i][0], id[i][1], id[i][2]);
    const int nr = (La + Lb + Lc + Ld) / 2 + 1;
    double roots[ORACLE_MAX_ROOTS], weights[ORACLE_MAX_ROOTS];
    const int da = La + Lb + 1, db = Lb + 1, dc = Lc + Ld + 1, dd = Ld + 1;
    const int sa = db * dc * dd, sb = dc * dd, sc = dd;
    double Ix[TAB_MAX], Iy[TAB_MAX], Iz[TAB_MAX];
    const double *ea = bs->exps + bs->prim_off[s0], *eb = bs->exps + bs->prim_off[s1];
    const double *ec = bs->exps + bs->prim_off[s2], *ed = bs->exps + bs->prim_off[s3];
    const double *ka = bs->coefs + bs->prim_off[s0], *kb = bs->coefs + bs->prim_off[s1];
    const double *kc = bs->coefs + bs->prim_off[s2], *kd = bs->coefs + bs->prim_off[s3];
    const double AB2 = ABx * ABx + ABy * ABy + ABz * ABz;
    const double CD2 = CDx * CDx + CDy * CDy + CDz * CDz;
    for (int p = 0; p < bs->nprim[s0]; ++p) {
        const double alpha = ea[p], cA = ka[p];
        for (int q = 0; q < bs->nprim[s1]; ++q) {
            const double beta = eb[q], cB = kb[q];
            /* gaussian_product.cpp:13-30 */
            const double zeta = alpha + beta;
            const double izeta = 1.0 / zeta;
            const double mu = alpha * beta * izeta;
            const pt3 P = {(alpha * A.x + beta * B.x) * izeta, (alpha * A.y + beta * B.y) * izeta,
                           (alpha * A.z + beta * B.z) * izeta};
            const double KAB = exp(-mu * AB2);
            for (int r = 0; r < bs->nprim[s2]; ++r) {
                const double gamma = ec[r], cC = kc[r];
                for (int s = 0; s < bs->nprim[s3]; ++s) {
                    const double delta = ed[s], cD = kd[s];
                    const double eta = gamma + delta;
                    const double ieta = 1.0 / eta;
                    const double nu = gamma * delta * ieta;
                    const pt3 Q = {(gamma * C.x + delta * D.x) * ieta, (gamma * C.y + delta * D.y) * ieta,
                                   (gamma * C.z + delta * D.z) * ieta};
                    const double KCD = exp(-nu * CD2);
                    const double rho = zeta * eta / (zeta + eta);
                    const double PQx = P.x - Q.x, PQy = P.y - Q.y, PQz = P.z - Q.z;
                    const double PQ2 = PQx * PQx + PQy * PQy + PQz * PQz;
                    const double T = rho * PQ2;
                    const double pi52 = PI_ * PI_ * sqrt(PI_);
                    const double pref = 2.0 * pi52 / (zeta * eta * sqrt(zeta + eta)) * KAB * KCD;
                    const double coeff = cA * cB * cC * cD * pref;
                    oracle_rys(nr, T, roots, weights);
                    for (int rt = 0; rt < nr; ++rt) {
                        const double u = roots[rt], w = weights[rt];
                        const double roz = rho / zeta, roe = rho / eta;
                        const double B10 = 0.5 / zeta * (1.0 - roz * u);
                        const double B01 = 0.5 / eta * (1.0 - roe * u);
                        const double B00 = 0.5 / (zeta + eta) * u;
                        const double PAx = (P.x - A.x) - roz * u * PQx;
                        const double PAy = (P.y - A.y) - roz * u * PQy;
                        const double PAz = (P.z - A.z) - roz * u * PQz;
                        const double QCx = (Q.x - C.x) + roe * u * PQx;
                        const double QCy = (Q.y - C.y) + roe * u * PQy;
                        const double QCz = (Q.z - C.z) + roe * u * PQz;
                        build_2d(La, Lb, Lc, Ld, PAx, QCx, ABx, CDx, B10, B01, B00, Ix, da, db, dc, dd);
                        build_2d(La, Lb, Lc, Ld, PAy, QCy, ABy, CDy, B10, B01, B00, Iy, da, db, dc, dd);
                        build_2d(La, Lb, Lc, Ld, PAz, QCz, ABz, CDz, B10, B01, B00, Iz, da, db, dc, dd);
                        const double wc = coeff * w;
                        int o = 0;
                        for (int a = 0; a < na; ++a)
                            for (int b = 0; b < nb; ++b)
                                for (int c = 0; c < nc; ++c)
                                    for (int d = 0; d < nd; ++d, ++o) {
                                        const double v =
                                            Ix[ia[a][0] * sa + ib[b][0] * sb + ic[c][0] * sc + id[d][0]] *
                                            Iy[ia[a][1] * sa + ib[b][1] * sb + ic[c][1] * sc + id[d][1]] *
                                            Iz[ia[a][2] * sa + ib[b][2] * sb + ic[c][2] * sc + id[d][2]];
                                        out[o] += wc * v;
                                    }
                    }
                }
            }
        }
    }
    int o = 0;
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nc; ++c)
                for (int d = 0; d < nd; ++d, ++o) out[o] *= ca_[a] * cb_[b] * cc_[c] * cd_[d];
}

void oracle_eri_tensor(const oracle_basis *b, double *out)
{
    const int n = b->nbf;
    double buf[15 * 15 * 15 * 15];
    for (int i = 0; i < b->n_shells; ++i)
        for (int j = 0; j < b->n_shells; ++j)
            for (int k = 0; k < b->n_shells; ++k)
                for (int l = 0; l < b->n_shells; ++l) {
                    oracle_eri_shell_quartet(b, i, j, k, l, buf);
                    const int na = ncart(b->L[i]), nb = ncart(b->L[j]), nc = ncart(b->L[k]),
                              nd = ncart(b->L[l]);
                    int o = 0;
                    for (int a = 0; a < na; ++a)
                        for (int bb = 0; bb < nb; ++bb)
                            for (int c = 0; c < nc; ++c)
                                for (int d = 0; d < nd; ++d, ++o) {
                                    const size_t mu = b->func_off[i] + a, nu = b->func_off[j] + bb;
                                    const size_t la = b->func_off[k] + c, si = b->func_off[l] + d;
                                    out[((mu * n + nu) * n + la) * n + si] = buf[o];
                                }
                }
}

/* ======================================================================== */
/* One-electron integrals                                                     */
/* ======================================================================== */

#define D1 (ORACLE_MAX_L + 3)

/* overlap_kernel.cpp:30-60 */
static void build_1d_overlap(int La, int Lb, double XPA, double XPB, double o2z, double I[D1][D1])
{
    for (int i = 0; i <= La; ++i)
        for (int j = 0; j <= Lb; ++j) I[i][j] = 0.0;
    I[0][0] = 1.0;
    for (int i = 0; i < La; ++i) {
        I[i + 1][0] = XPA * I[i][0];
        if (i > 0) I[i + 1][0] += (double)i * o2z * I[i - 1][0];
    }
    for (int j = 0; j < Lb; ++j)
        for (int i = 0; i <= La; ++i) {
            double v = XPB * I[i][j];
            if (i > 0) v += (double)i * o2z * I[i - 1][j];
            if (j > 0) v += (double)j * o2z * I[i][j - 1];
            I[i][j + 1] = v;
        }
}

/* kinetic_kernel.cpp:80-92 */
static double kin1d(double S[D1][D1], int i, int j, double beta)
{
    double v = 4.0 * beta * beta * S[i][j + 2];
    v -= 2.0 * beta * (double)(2 * j + 1) * S[i][j];
    if (j >= 2) v += (double)(j * (j - 1)) * S[i][j - 2];
    return -0.5 * v;
}

/* nuclear_kernel.cpp:30-55 */
static void build_1d_rys(int La, int Lb, double PA, double PB, double B00, double I[D1][D1])
{
    for (int i = 0; i <= La; ++i)
        for (int j = 0; j <= Lb; ++j) I[i][j] = 0.0;
    I[0][0] = 1.0;
    for (int i = 0; i < La; ++i) {
        I[i + 1][0] = PA * I[i][0];
        if (i > 0) I[i + 1][0] += (double)i * B00 * I[i - 1][0];
    }
    for (int j = 0; j < Lb; ++j)
        for (int i = 0; i <= La; ++i) {
            double v = PB * I[i][j];
            if (i > 0) v += (double)i * B00 * I[i - 1][j];
            if (j > 0) v += (double)j * B00 * I[i][j - 1];
            I[i][j + 1] = v;
        }
}

/* kind 0 = S, 1 = T, 2 = V */
static void one_e_pair(const oracle_basis *bs, int s0, int s1, int kind, int nch, const double *Zc,
                       const double *xyz, double *buf)
{
    const int La = bs->L[s0], Lb = bs->L[s1];
    const int na = ncart(La), nb = ncart(Lb);
    for (int i = 0; i < na * nb; ++i) buf[i] = 0.0;
    const pt3 A = {bs->center[3 * s0], bs->center[3 * s0 + 1], bs->center[3 * s0 + 2]};
    const pt3 B = {bs->center[3 * s1], bs->center[3 * s1 + 1], bs->center[3 * s1 + 2]};
    int ia[15][3], ib[15][3];
    cart_indices(La, ia);
    cart_indices(Lb, ib);
    const double *ea = bs->exps + bs->prim_off[s0], *eb = bs->exps + bs->prim_off[s1];
    const double *ka = bs->coefs + bs->prim_off[s0], *kb = bs->coefs + bs->prim_off[s1];
    const double dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
    const double AB2 = dx * dx + dy * dy + dz * dz;
    double Ix[D1][D1], Iy[D1][D1], Iz[D1][D1];
    const int nr = (La + Lb) / 2 + 1;
    double roots[ORACLE_MAX_ROOTS], weights[ORACLE_MAX_ROOTS];
    for (int p = 0; p < bs->nprim[s0]; ++p) {
        const double alpha = ea[p], cA = ka[p];
        for (int q = 0; q < bs->nprim[s1]; ++q) {
            const double beta = eb[q], cB = kb[q];
            const double zeta = alpha + beta, iz = 1.0 / zeta, mu = alpha * beta * iz;
            const pt3 P = {(alpha * A.x + beta * B.x) * iz, (alpha * A.y + beta * B.y) * iz,
                           (alpha * A.z + beta * B.z) * iz};
            const double KAB = exp(-mu * AB2);
            const double PAx = P.x - A.x, PAy = P.y - A.y, PAz = P.z - A.z;
            const double PBx = P.x - B.x, PBy = P.y - B.y, PBz = P.z - B.z;
            if (kind == 0 || kind == 1) {
                const double o2z = 0.5 / zeta;
                const double pref = pow(PI_ / zeta, 1.5) * KAB;
                const int lbx = (kind == 1) ? Lb + 2 : Lb;
                build_1d_overlap(La, lbx, PAx, PBx, o2z, Ix);
                build_1d_overlap(La, lbx, PAy, PBy, o2z, Iy);
                build_1d_overlap(La, lbx, PAz, PBz, o2z, Iz);
                const double pc = cA * cB * pref;
                for (int a = 0; a < na; ++a)
                    for (int b = 0; b < nb; ++b) {
                        const int ax = ia[a][0], ay = ia[a][1], az = ia[a][2];
                        const int bx = ib[b][0], by = ib[b][1], bz = ib[b][2];
                        double v;
                        if (kind == 0) {
                            v = Ix[ax][bx] * Iy[ay][by] * Iz[az][bz];
                        } else {
                            const double Tx = kin1d(Ix, ax, bx, beta), Ty = kin1d(Iy, ay, by, beta),
                                         Tz = kin1d(Iz, az, bz, beta);
                            const double Sx = Ix[ax][bx], Sy = Iy[ay][by], Sz = Iz[az][bz];
                            v = Tx * Sy * Sz + Sx * Ty * Sz + Sx * Sy * Tz;
                        }
                        buf[a * nb + b] += pc * v;
                    }
            } else {
                const double pref = (2.0 * PI_ / zeta) * KAB;
                const double B00b = 0.5 / zeta;
                for (int c = 0; c < nch; ++c) {
                    const double Z = Zc[c];
                    if (Z == 0.0) continue;
                    const double PCx = P.x - xyz[3 * c], PCy = P.y - xyz[3 * c + 1], PCz = P.z - xyz[3 * c + 2];
                    const double T = zeta * (PCx * PCx + PCy * PCy + PCz * PCz);
                    oracle_rys(nr, T, roots, weights);
                    for (int r = 0; r < nr; ++r) {
                        const double u = roots[r], w = weights[r];
                        const double B00 = B00b * (1.0 - u);
                        build_1d_rys(La, Lb, PAx - u * PCx, PBx - u * PCx, B00, Ix);
                        build_1d_rys(La, Lb, PAy - u * PCy, PBy - u * PCy, B00, Iy);
                        build_1d_rys(La, Lb, PAz - u * PCz, PBz - u * PCz, B00, Iz);
                        const double co = -Z * cA * cB * pref * w;
                        for (int a = 0; a < na; ++a)
                            for (int b = 0; b < nb; ++b)
                                buf[a * nb + b] +=
                                    co * (Ix[ia[a][0]][ib[b][0]] * Iy[ia[a][1]][ib[b][1]] * Iz[ia[a][2]][ib[b][2]]);
                    }
                }
            }
        }
    }
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b)
            buf[a * nb + b] *= norm_corr(ia[a][0], ia[a][1], ia[a][2]) * norm_corr(ib[b][0], ib[b][1], ib[b][2]);
}

/* full matrix over the upper-triangle shell-pair list, mirrored
 * (src/host/engine/cpu_engine.cpp:482-593) */
static void one_e_matrix(const oracle_basis *bs, int kind, int nch, const double *Z, const double *xyz,
                         double *M)
{
    const int n = bs->nbf;
    double buf[15 * 15];
    for (int i = 0; i < bs->n_shells; ++i)
        for (int j = i; j < bs->n_shells; ++j) {
            one_e_pair(bs, i, j, kind, nch, Z, xyz, buf);
            const int na = ncart(bs->L[i]), nb = ncart(bs->L[j]);
            for (int a = 0; a < na; ++a)
                for (int b = 0; b < nb; ++b) {
                    const int mu = bs->func_off[i] + a, nu = bs->func_off[j] + b;
                    M[mu * n + nu] = buf[a * nb + b];
                    M[nu * n + mu] = buf[a * nb + b];
                }
        }
}

void oracle_overlap(const oracle_basis *b, double *S) { one_e_matrix(b, 0, 0, 0, 0, S); }
void oracle_kinetic(const oracle_basis *b, double *T) { one_e_matrix(b, 1, 0, 0, 0, T); }
void oracle_nuclear(const oracle_basis *b, int nch, const double *Z, const double *xyz, double *V)
{
    one_e_matrix(b, 2, nch, Z, xyz, V);
}

/* ======================================================================== */
/* Schwarz + screened list                                                    */
/* ======================================================================== */

static size_t packed(size_t i, size_t j) /* schwarz_bounds.hpp:110-113 */
{
    if (i > j) {
        size_t t = i;
        i = j;
        j = t;
    }
    return j * (j + 1) / 2 + i;
}

void oracle_schwarz(const oracle_basis *b, double *Q, int n_threads)
{
    const int n = b->n_shells;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
#endif
    for (int i = 0; i < n; ++i) {
        double *buf = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
        for (int j = i; j < n; ++j) {
            oracle_eri_shell_quartet(b, i, j, i, j, buf);
            const int ni = ncart(b->L[i]), nj = ncart(b->L[j]);
            double mx = 0.0;
            for (int a = 0; a < ni; ++a)
                for (int bb = 0; bb < nj; ++bb) {
                    const double v = fabs(buf[((a * nj + bb) * ni + a) * nj + bb]);
                    mx = fmax(mx, v);
                }
            Q[packed(i, j)] = sqrt(mx);
        }
        free(buf);
    }
    (void)n_threads;
}

/* order-independent list checksum: sum over quartets of splitmix64 of the
 * packed (i, j, k, l), modulo 2^64 */
static uint64_t quartet_mix(uint32_t i, uint32_t j, uint32_t k, uint32_t l)
{
    uint64_t x = ((uint64_t)(i & 0xffffu) << 48) | ((uint64_t)(j & 0xffffu) << 32) |
                 ((uint64_t)(k & 0xffffu) << 16) | (uint64_t)(l & 0xffffu);
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

int64_t oracle_screened_list(const oracle_basis *b, const double *Q, double thr, uint64_t *hash_out)
{
    const size_t n = (size_t)b->n_shells;
    int64_t count = 0;
    uint64_t h = 0;
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i; j < n; ++j)
            for (size_t k = i; k < n; ++k)
                for (size_t l = (k == i) ? j : k; l < n; ++l) {
                    if (thr > 0.0 && !(Q[packed(i, j)] * Q[packed(k, l)] >= thr)) continue;
                    ++count;
                    h += quartet_mix((uint32_t)i, (uint32_t)j, (uint32_t)k, (uint32_t)l);
                }
    if (hash_out) *hash_out = h;
    return count;
}

/* ======================================================================== */
/* Fock build                                                                 */
/* ======================================================================== */

/* fock_builder.cpp:123-222 (accumulate_symmetric) */
static void acc_symmetric(const double *g, const double *D, double *J, double *K, size_t n, int fa,
                          int fb, int fc, int fd, int na, int nb, int nc, int nd, int ij, int kl, int bk)
{
    int o = 0;
    for (int a = 0; a < na; ++a) {
        const size_t mu = fa + a;
        for (int b = 0; b < nb; ++b) {
            const size_t nu = fb + b;
            for (int c = 0; c < nc; ++c) {
                const size_t lam = fc + c;
                for (int d = 0; d < nd; ++d, ++o) {
                    const size_t sig = fd + d;
                    const double v = g[o];
                    const double Dls = D[lam * n + sig], Dsl = D[sig * n + lam];
                    const double Dmn = D[mu * n + nu], Dnm = D[nu * n + mu];
                    const double Dns = D[nu * n + sig], Dms = D[mu * n + sig];
                    const double Dnl = D[nu * n + lam], Dml = D[mu * n + lam];
                    J[mu * n + nu] += v * Dls;
                    K[mu * n + lam] += v * Dns;
                    if (!kl) {
                        J[mu * n + nu] += v * Dsl;
                        K[mu * n + sig] += v * Dnl;
                    }
                    if (!ij) {
                        J[nu * n + mu] += v * Dls;
                        K[nu * n + lam] += v * Dms;
                    }
                    if (!ij && !kl) {
                        J[nu * n + mu] += v * Dsl;
                        K[nu * n + sig] += v * Dml;
                    }
                    if (!bk) {
                        J[lam * n + sig] += v * Dmn;
                        K[lam * n + mu] += v * Dns;
                        if (!ij) {
                            J[lam * n + sig] += v * Dnm;
                            K[lam * n + nu] += v * Dms;
                        }
                        if (!kl) {
                            J[sig * n + lam] += v * Dmn;
                            K[sig * n + mu] += v * Dnl;
                        }
                        if (!ij && !kl) {
                            J[sig * n + lam] += v * Dnm;
                            K[sig * n + nu] += v * Dml;
                        }
                    }
                }
            }
        }
    }
}

void oracle_jk_canonical(const oracle_basis *b, const double *D, const double *Q, double thr,
                         double *J, double *K, int n_threads)
{
    const int n = b->n_shells;
    const size_t nbf = (size_t)b->nbf;
    memset(J, 0, sizeof(double) * nbf * nbf);
    memset(K, 0, sizeof(double) * nbf * nbf);
    const int use_q = (Q != NULL && thr > 0.0);
    /* flatten the (i,j) pair list so threads take whole bra pairs */
    const int npair = n * (n + 1) / 2;
    int *pi = (int *)malloc(sizeof(int) * npair), *pj = (int *)malloc(sizeof(int) * npair);
    int c = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            pi[c] = i;
            pj[c] = j;
            ++c;
        }
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
    {
        double *Jl = (double *)calloc(nbf * nbf, sizeof(double));
        double *Kl = (double *)calloc(nbf * nbf, sizeof(double));
        double *buf = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int p = 0; p < npair; ++p) {
            const int i = pi[p], j = pj[p];
            for (int k = i; k < n; ++k)
                for (int l = (k == i) ? j : k; l < n; ++l) {
                    if (use_q && !(Q[packed(i, j)] * Q[packed(k, l)] >= thr)) continue;
                    oracle_eri_shell_quartet(b, i, j, k, l, buf);
                    acc_symmetric(buf, D, Jl, Kl, nbf, b->func_off[i], b->func_off[j], b->func_off[k],
                                  b->func_off[l], ncart(b->L[i]), ncart(b->L[j]), ncart(b->L[k]),
                                  ncart(b->L[l]), i == j, k == l, i == k && j == l);
                }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (size_t x = 0; x < nbf * nbf; ++x) {
                J[x] += Jl[x];
                K[x] += Kl[x];
            }
        }
        free(Jl);
        free(Kl);
        free(buf);
    }
    free(pi);
    free(pj);
    (void)n_threads;
}

/* fock_builder.cpp:64-121 (accumulate), cpu_engine.hpp:553-622 loop */
void oracle_jk_full(const oracle_basis *b, const double *D, double *J, double *K, int n_threads,
                    int i_begin, int i_end)
{
    const int n = b->n_shells;
    const size_t nbf = (size_t)b->nbf;
    if (i_begin < 0) i_begin = 0;
    if (i_end <= 0 || i_end > n) i_end = n;
    memset(J, 0, sizeof(double) * nbf * nbf);
    memset(K, 0, sizeof(double) * nbf * nbf);
    const long long total = (long long)(i_end - i_begin) * n * n * n;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
#endif
    {
        double *Jl = (double *)calloc(nbf * nbf, sizeof(double));
        double *Kl = (double *)calloc(nbf * nbf, sizeof(double));
        double *buf = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
        for (long long t = 0; t < total; ++t) {
            const int l = (int)(t % n);
            const int k = (int)((t / n) % n);
            const int j = (int)((t / ((long long)n * n)) % n);
            const int i = i_begin + (int)(t / ((long long)n * n * n));
            oracle_eri_shell_quartet(b, i, j, k, l, buf);
            const int na = ncart(b->L[i]), nb = ncart(b->L[j]), nc = ncart(b->L[k]), nd = ncart(b->L[l]);
            const int fa = b->func_off[i], fb = b->func_off[j], fc = b->func_off[k], fd = b->func_off[l];
            int o = 0;
            for (int a = 0; a < na; ++a)
                for (int bb = 0; bb < nb; ++bb)
                    for (int c = 0; c < nc; ++c)
                        for (int d = 0; d < nd; ++d, ++o) {
                            const size_t mu = fa + a, nu = fb + bb, lam = fc + c, sig = fd + d;
                            Jl[mu * nbf + nu] += buf[o] * D[lam * nbf + sig];
                            Kl[mu * nbf + lam] += buf[o] * D[nu * nbf + sig];
                        }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (size_t x = 0; x < nbf * nbf; ++x) {
                J[x] += Jl[x];
                K[x] += Kl[x];
            }
        }
        free(Jl);
        free(Kl);
        free(buf);
    }
    (void)n_threads;
}

/* canonical screened J/K over every stride-th bra pair (offset) of the
 * reference order, counting quartets and ERIs (bounded CPU-baseline sample) */
void oracle_jk_canonical_sample(const oracle_basis *b, const double *D, const double *Q, double thr, int stride,
                                int offset, double *J, double *K, int n_threads, long long *n_quartets,
                                long long *n_eri)
{
    const int n = b->n_shells;
    const size_t nbf = (size_t)b->nbf;
    memset(J, 0, sizeof(double) * nbf * nbf);
    memset(K, 0, sizeof(double) * nbf * nbf);
    const int npair = n * (n + 1) / 2;
    int *pi = (int *)malloc(sizeof(int) * npair), *pj = (int *)malloc(sizeof(int) * npair);
    int c = 0, p = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j, ++p)
            if (p % stride == offset) {
                pi[c] = i;
                pj[c] = j;
                ++c;
            }
    long long nq = 0, ne = 0;
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads) reduction(+ : nq, ne)
#endif
    {
        double *Jl = (double *)calloc(nbf * nbf, sizeof(double));
        double *Kl = (double *)calloc(nbf * nbf, sizeof(double));
        double *buf = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int t = 0; t < c; ++t) {
            const int i = pi[t], j = pj[t];
            for (int k = i; k < n; ++k)
                for (int l = (k == i) ? j : k; l < n; ++l) {
                    if (!(Q[packed(i, j)] * Q[packed(k, l)] >= thr)) continue;
                    oracle_eri_shell_quartet(b, i, j, k, l, buf);
                    const int na = ncart(b->L[i]), nb = ncart(b->L[j]), nc = ncart(b->L[k]), nd = ncart(b->L[l]);
                    acc_symmetric(buf, D, Jl, Kl, nbf, b->func_off[i], b->func_off[j], b->func_off[k],
                                  b->func_off[l], na, nb, nc, nd, i == j, k == l, i == k && j == l);
                    ++nq;
                    ne += (long long)na * nb * nc * nd;
                }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (size_t x = 0; x < nbf * nbf; ++x) {
                J[x] += Jl[x];
                K[x] += Kl[x];
            }
        }
        free(Jl);
        free(Kl);
        free(buf);
    }
    free(pi);
    free(pj);
    *n_quartets = nq;
    *n_eri = ne;
    (void)n_threads;
}
