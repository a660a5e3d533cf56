/* Rys nodes and weights for the CPU oracle (test infrastructure only).
 *
 * This is deliberately a different route from the product's mpmath-generated tables:
 *  - the weight exp(-T t^2) on t in [0, b], b = min(1, 12/sqrt(T)), is discretised with
 *    a composite Gauss-Legendre rule (6 panels x 24 points);
 *  - the discretised Stieltjes procedure (W. Gautschi, "Orthogonal Polynomials:
 *    Computation and Approximation", Oxford 2004, sec. 2.2.3.1) gives the recurrence
 *    coefficients of the orthogonal polynomials in x = t^2;
 *  - the nodes are the eigenvalues of the Jacobi matrix (implicit QL on the symmetric
 *    tridiagonal matrix) and w_i = mu_0 v_{0i}^2 (Golub & Welsch 1969).
 * For speed `oracle_rys_roots` interpolates Chebyshev fits (width 0.5, degree 15) of
 * the direct values, built on first use; above T = 40 + 5N it uses the half-range
 * Gauss-Hermite nodes (the roots of H_2N from the Hermite Jacobi matrix): x = s^2/T,
 * w = W/sqrt(T).
 * The reference obtains the same nodes from libcint-derived fits
 * (cuobc/lib/gint/rys_roots.cu:1314-1777); it returns u = x/(1-x).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include "oracle.h"

#define NMAX_ROOTS 10
#define NPANEL 6
#define NGL 24
#define NQ (NPANEL * NGL)

static double gl_x[NGL], gl_w[NGL];

/* eigenvalues/vectors of a symmetric tridiagonal matrix, implicit QL.
 * d: diagonal (n), e: off-diagonal (e[1..n-1] used), z: n x n, starts as identity. */
static void tqli(double *d, double *e, int n, double *z)
{
    for (int i = 1; i < n; i++) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= 1e-17 * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) break;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i], b = c * e[i];
                    e[i + 1] = (r = hypot(f, g));
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                    for (int k = 0; k < n; k++) {
                        f = z[k * n + i + 1];
                        z[k * n + i + 1] = s * z[k * n + i] + c * f;
                        z[k * n + i] = c * z[k * n + i] - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
}

static void jacobi_to_rule(int n, const double *alpha, const double *beta, double mu0,
                           double *x, double *w)
{
    double d[NMAX_ROOTS * 2], e[NMAX_ROOTS * 2], z[NMAX_ROOTS * NMAX_ROOTS * 4];
    for (int i = 0; i < n; i++) {
        d[i] = alpha[i];
        e[i] = i ? sqrt(beta[i]) : 0.0;
    }
    memset(z, 0, sizeof(double) * n * n);
    for (int i = 0; i < n; i++) z[i * n + i] = 1.0;
    tqli(d, e, n, z);
    /* sort ascending */
    int idx[NMAX_ROOTS * 2];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (d[idx[j]] < d[idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int i = 0; i < n; i++) {
        x[i] = d[idx[i]];
        double v = z[0 * n + idx[i]];
        w[i] = mu0 * v * v;
    }
}

static void init_gl(void)
{
    /* Gauss-Legendre nodes on [-1,1] by Newton on P_n */
    int n = NGL;
    for (int i = 0; i < n; i++) {
        double x = cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dp = n * (x * p1 - p0) / (x * x - 1.0);
        gl_x[i] = x;
        gl_w[i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
}

void oracle_rys_roots_direct(int n, double T, double *x, double *w)
{
    static pthread_once_t once = PTHREAD_ONCE_INIT;
    pthread_once(&once, init_gl);
    double b = 1.0;
    if (T > 144.0) b = 12.0 / sqrt(T);
    double xs[NQ], ws[NQ];
    double h = b / NPANEL;
    int q = 0;
    for (int p = 0; p < NPANEL; p++) {
        double a0 = p * h;
        for (int i = 0; i < NGL; i++, q++) {
            double t = a0 + (gl_x[i] + 1.0) * 0.5 * h;
            xs[q] = t * t;
            ws[q] = gl_w[i] * 0.5 * h * exp(-T * t * t);
        }
    }
    /* discretised Stieltjes procedure in x */
    double pprev[NQ], pcur[NQ], pnext[NQ];
    double alpha[NMAX_ROOTS], beta[NMAX_ROOTS];
    double norm_prev = 0.0, norm = 0.0, mom = 0.0;
    for (int j = 0; j < NQ; j++) {
        pprev[j] = 0.0;
        pcur[j] = 1.0;
        norm += ws[j];
        mom += ws[j] * xs[j];
    }
    double mu0 = norm;
    alpha[0] = mom / norm;
    beta[0] = mu0;
    for (int k = 1; k < n; k++) {
        double nn = 0.0, mm = 0.0;
        for (int j = 0; j < NQ; j++) {
            pnext[j] = (xs[j] - alpha[k - 1]) * pcur[j] - (k > 1 ? beta[k - 1] : 0.0) * pprev[j];
            nn += ws[j] * pnext[j] * pnext[j];
            mm += ws[j] * xs[j] * pnext[j] * pnext[j];
        }
        norm_prev = norm;
        norm = nn;
        alpha[k] = mm / nn;
        beta[k] = nn / norm_prev;
        memcpy(pprev, pcur, sizeof(pcur));
        memcpy(pcur, pnext, sizeof(pnext));
    }
    jacobi_to_rule(n, alpha, beta, mu0, x, w);
}

/* ---- cached Chebyshev fits -------------------------------------------------- */
#define FIT_H 0.5
#define FIT_DEG 15
static double *fit_tab[NMAX_ROOTS + 1];
static double herm_s2[NMAX_ROOTS + 1][NMAX_ROOTS], herm_w[NMAX_ROOTS + 1][NMAX_ROOTS];
static double fit_tmax(int n) { return 40.0 + 5.0 * n; }

static void build_fits(void)
{
    for (int n = 1; n <= NMAX_ROOTS; n++) {
        int nint = (int)ceil(fit_tmax(n) / FIT_H);
        int m = FIT_DEG + 1;
        double *tab = (double *)malloc(sizeof(double) * nint * 2 * n * m);
        for (int it = 0; it < nint; it++) {
            double vals[2 * NMAX_ROOTS][FIT_DEG + 1];
            for (int k = 0; k < m; k++) {
                double s = cos(M_PI * (k + 0.5) / m);
                double T = it * FIT_H + (s + 1.0) * 0.5 * FIT_H;
                double x[NMAX_ROOTS], w[NMAX_ROOTS];
                oracle_rys_roots_direct(n, T, x, w);
                for (int r = 0; r < n; r++) {
                    vals[r][k] = x[r];
                    vals[n + r][k] = w[r];
                }
            }
            for (int r = 0; r < 2 * n; r++) {
                for (int j = 0; j < m; j++) {
                    double c = 0.0;
                    for (int k = 0; k < m; k++)
                        c += vals[r][k] * cos(M_PI * j * (k + 0.5) / m);
                    c *= 2.0 / m;
                    if (j == 0) c *= 0.5;
                    tab[((size_t)it * 2 * n + r) * m + j] = c;
                }
            }
        }
        fit_tab[n] = tab;
        /* half-range Hermite rule: Jacobi matrix of H_2n has alpha = 0, beta_k = k/2 */
        int n2 = 2 * n;
        double alpha[2 * NMAX_ROOTS], beta[2 * NMAX_ROOTS], xs[2 * NMAX_ROOTS], ws[2 * NMAX_ROOTS];
        for (int k = 0; k < n2; k++) {
            alpha[k] = 0.0;
            beta[k] = 0.5 * k;
        }
        jacobi_to_rule(n2, alpha, beta, sqrt(M_PI), xs, ws);
        for (int r = 0; r < n; r++) {
            double s = xs[n + r]; /* positive roots, ascending */
            herm_s2[n][r] = s * s;
            herm_w[n][r] = ws[n + r];
        }
    }
}

void oracle_rys_roots(int n, double T, double *x, double *w)
{
    static pthread_once_t once = PTHREAD_ONCE_INIT;
    pthread_once(&once, build_fits);
    if (T >= fit_tmax(n)) {
        double it = 1.0 / T, sit = sqrt(it);
        for (int r = 0; r < n; r++) {
            x[r] = herm_s2[n][r] * it;
            w[r] = herm_w[n][r] * sit;
        }
        return;
    }
    int i = (int)(T / FIT_H);
    double s = 2.0 * (T - i * FIT_H) / FIT_H - 1.0;
    const double *tab = fit_tab[n] + (size_t)i * 2 * n * (FIT_DEG + 1);
    for (int r = 0; r < 2 * n; r++) {
        const double *c = tab + r * (FIT_DEG + 1);
        double b1 = 0.0, b2 = 0.0;
        for (int j = FIT_DEG; j >= 1; j--) {
            double b0 = 2.0 * s * b1 - b2 + c[j];
            b2 = b1;
            b1 = b0;
        }
        double v = s * b1 - b2 + c[0];
        if (r < n) x[r] = v; else w[r - n] = v;
    }
}
