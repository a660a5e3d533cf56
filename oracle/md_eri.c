/* McMurchie-Davidson electron-repulsion integrals for the CPU oracle (test
 * infrastructure only).  An algorithm independent of Rys quadrature:
 *
 *   (ab|cd) = 2 pi^(5/2) / (p q sqrt(p+q)) sum_tuv E^ab_tuv
 *             sum_tau,nu,phi (-1)^(tau+nu+phi) E^cd_tau,nu,phi R_(t+tau)(u+nu)(v+phi)
 *
 * with Hermite expansion coefficients E (Helgaker, Jorgensen, Olsen, "Molecular
 * Electronic-Structure Theory", eqs. 9.5.6-9.5.8) and Hermite Coulomb integrals
 * R^n_tuv from R^n_000 = (-2 alpha)^n F_n(alpha R_PQ^2) (eqs. 9.9.18-9.9.20).
 * Cartesian functions are c_env * fac_sp(l) * x^i y^j z^k exp(-a r^2), where
 * fac_sp(0) = 1/(2 sqrt(pi)), fac_sp(1) = sqrt(3/(4 pi)), fac_sp(l>1) = 1
 * (cuobc/lib/gint/pair_data.c:27-28).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define ATM_SLOTS 6
#define BAS_SLOTS 8
#define LMAX 4
#define LTOT (4 * LMAX)

static double fac_sp(int l)
{
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

static void cart_powers(int l, int *px, int *py, int *pz)
{
    int n = 0;
    for (int lx = l; lx >= 0; lx--)
        for (int ly = l - lx; ly >= 0; ly--, n++) {
            px[n] = lx;
            py[n] = ly;
            pz[n] = l - lx - ly;
        }
}

/* E[i][j][t] for i <= li, j <= lj, t <= i + j in one Cartesian direction. */
#define EDIM (LMAX + 1)
#define ETDIM (2 * LMAX + 1)
static void hermite_e(int li, int lj, double a, double b, double xa, double xb,
                      double E[EDIM][EDIM][ETDIM])
{
    double p = a + b, mu = a * b / p, xab = xa - xb;
    double xp = (a * xa + b * xb) / p;
    double xpa = xp - xa, xpb = xp - xb, oo2p = 0.5 / p;
    memset(E, 0, sizeof(double) * EDIM * EDIM * ETDIM);
    E[0][0][0] = exp(-mu * xab * xab);
    for (int i = 0; i <= li; i++) {
        for (int j = 0; j <= lj; j++) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; t++) {
                double v;
                if (i > 0) {
                    v = xpa * E[i - 1][j][t];
                    if (t > 0) v += oo2p * E[i - 1][j][t - 1];
                    if (t + 1 <= i - 1 + j) v += (t + 1) * E[i - 1][j][t + 1];
                } else {
                    v = xpb * E[i][j - 1][t];
                    if (t > 0) v += oo2p * E[i][j - 1][t - 1];
                    if (t + 1 <= i + j - 1) v += (t + 1) * E[i][j - 1][t + 1];
                }
                E[i][j][t] = v;
            }
        }
    }
}

#define RD (LTOT + 1)
static inline int ridx(int t, int u, int v) { return (t * RD + u) * RD + v; }

/* R0[t][u][v] = R^0_tuv for t+u+v <= L.  work: (L+1) * RD^3 doubles. */
static void hermite_r(int L, double alpha, double X, double Y, double Z, double *R0, double *work)
{
    double F[LTOT + 1];
    oracle_boys(L, alpha * (X * X + Y * Y + Z * Z), F);
    size_t cube = (size_t)RD * RD * RD;
    double m2a = -2.0 * alpha, pw = 1.0;
    for (int n = 0; n <= L; n++) {
        work[n * cube + ridx(0, 0, 0)] = pw * F[n];
        pw *= m2a;
    }
    for (int N = 1; N <= L; N++) {
        for (int n = 0; n <= L - N; n++) {
            double *Rn = work + n * cube, *Rn1 = work + (n + 1) * cube;
            for (int t = 0; t <= N; t++)
                for (int u = 0; u <= N - t; u++) {
                    int v = N - t - u;
                    double val;
                    if (t > 0) {
                        val = X * Rn1[ridx(t - 1, u, v)];
                        if (t > 1) val += (t - 1) * Rn1[ridx(t - 2, u, v)];
                    } else if (u > 0) {
                        val = Y * Rn1[ridx(t, u - 1, v)];
                        if (u > 1) val += (u - 1) * Rn1[ridx(t, u - 2, v)];
                    } else {
                        val = Z * Rn1[ridx(t, u, v - 1)];
                        if (v > 1) val += (v - 1) * Rn1[ridx(t, u, v - 2)];
                    }
                    Rn[ridx(t, u, v)] = val;
                }
        }
    }
    memcpy(R0, work, sizeof(double) * cube);
}

typedef struct {
    int l, nprim;
    const double *exp, *coef, *r;
} shell_t;

static shell_t get_shell(const int32_t *atm, const int32_t *bas, const double *env, int ish)
{
    const int32_t *b = bas + ish * BAS_SLOTS;
    shell_t s;
    s.l = b[1];
    s.nprim = b[2];
    s.exp = env + b[5];
    s.coef = env + b[6];
    s.r = env + atm[b[0] * ATM_SLOTS + 1];
    return s;
}

int oracle_md_shell_quartet(const int32_t *atm, const int32_t *bas, const double *env,
                            int ish, int jsh, int ksh, int lsh, double *out)
{
    shell_t A = get_shell(atm, bas, env, ish), B = get_shell(atm, bas, env, jsh);
    shell_t C = get_shell(atm, bas, env, ksh), D = get_shell(atm, bas, env, lsh);
    if (A.l > LMAX || B.l > LMAX || C.l > LMAX || D.l > LMAX) return 1;
    int nfa = ncart(A.l), nfb = ncart(B.l), nfc = ncart(C.l), nfd = ncart(D.l);
    int ax[15], ay[15], az[15], bx[15], by[15], bz[15];
    int cx[15], cy[15], cz[15], dx[15], dy[15], dz[15];
    cart_powers(A.l, ax, ay, az);
    cart_powers(B.l, bx, by, bz);
    cart_powers(C.l, cx, cy, cz);
    cart_powers(D.l, dx, dy, dz);
    int Lab = A.l + B.l, Lcd = C.l + D.l, L = Lab + Lcd;
    size_t nout = (size_t)nfa * nfb * nfc * nfd;
    memset(out, 0, sizeof(double) * nout);
    size_t cube = (size_t)RD * RD * RD;
    double *R0 = (double *)malloc(sizeof(double) * cube);
    double *work = (double *)malloc(sizeof(double) * cube * (L + 1));
    /* G[cd component][t][u][v] = sum_tau,nu,phi (-1)^.. E^cd R_(t+tau)(u+nu)(v+phi) */
    int nh = (Lab + 1) * (Lab + 1) * (Lab + 1);
    double *G = (double *)malloc(sizeof(double) * nfc * nfd * nh);
    static const double pi52x2 = 34.98683665524972; /* 2 pi^(5/2) */
    double norm = fac_sp(A.l) * fac_sp(B.l) * fac_sp(C.l) * fac_sp(D.l);
    double Eab[3][EDIM][EDIM][ETDIM], Ecd[3][EDIM][EDIM][ETDIM];
    for (int ip = 0; ip < A.nprim; ip++)
    for (int jp = 0; jp < B.nprim; jp++) {
        double a = A.exp[ip], b = B.exp[jp], p = a + b;
        double P[3];
        for (int d = 0; d < 3; d++) {
            P[d] = (a * A.r[d] + b * B.r[d]) / p;
            hermite_e(A.l, B.l, a, b, A.r[d], B.r[d], Eab[d]);
        }
        double cab = A.coef[ip] * B.coef[jp];
        for (int kp = 0; kp < C.nprim; kp++)
        for (int lp = 0; lp < D.nprim; lp++) {
            double c = C.exp[kp], dd = D.exp[lp], q = c + dd;
            double Q[3];
            for (int d = 0; d < 3; d++) {
                Q[d] = (c * C.r[d] + dd * D.r[d]) / q;
                hermite_e(C.l, D.l, c, dd, C.r[d], D.r[d], Ecd[d]);
            }
            double alpha = p * q / (p + q);
            hermite_r(L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], R0, work);
            double pref = pi52x2 / (p * q * sqrt(p + q)) * cab * C.coef[kp] * D.coef[lp] * norm;
            for (int kc = 0; kc < nfc; kc++)
            for (int lc = 0; lc < nfd; lc++) {
                double *g = G + (size_t)(kc * nfd + lc) * nh;
                int mx = cx[kc] + dx[lc], my = cy[kc] + dy[lc], mz = cz[kc] + dz[lc];
                for (int t = 0; t <= Lab; t++)
                for (int u = 0; u <= Lab - t; u++)
                for (int v = 0; v <= Lab - t - u; v++) {
                    double s = 0.0;
                    for (int tau = 0; tau <= mx; tau++) {
                        double ex = Ecd[0][cx[kc]][dx[lc]][tau];
                        if (ex == 0.0) continue;
                        for (int nu = 0; nu <= my; nu++) {
                            double ey = Ecd[1][cy[kc]][dy[lc]][nu];
                            if (ey == 0.0) continue;
                            for (int phi = 0; phi <= mz; phi++) {
                                double ez = Ecd[2][cz[kc]][dz[lc]][phi];
                                double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                                s += sg * ex * ey * ez * R0[ridx(t + tau, u + nu, v + phi)];
                            }
                        }
                    }
                    g[(t * (Lab + 1) + u) * (Lab + 1) + v] = s;
                }
            }
            for (int ia = 0; ia < nfa; ia++)
            for (int jb = 0; jb < nfb; jb++) {
                int nx = ax[ia] + bx[jb], ny = ay[ia] + by[jb], nz = az[ia] + bz[jb];
                for (int kc = 0; kc < nfc; kc++)
                for (int lc = 0; lc < nfd; lc++) {
                    const double *g = G + (size_t)(kc * nfd + lc) * nh;
                    double s = 0.0;
                    for (int t = 0; t <= nx; t++) {
                        double ex = Eab[0][ax[ia]][bx[jb]][t];
                        for (int u = 0; u <= ny; u++) {
                            double ey = Eab[1][ay[ia]][by[jb]][u];
                            for (int v = 0; v <= nz; v++)
                                s += ex * ey * Eab[2][az[ia]][bz[jb]][v] *
                                     g[(t * (Lab + 1) + u) * (Lab + 1) + v];
                        }
                    }
                    out[((size_t)(ia * nfb + jb) * nfc + kc) * nfd + lc] += pref * s;
                }
            }
        }
    }
    free(G);
    free(R0);
    free(work);
    return 0;
}

int oracle_md_eri(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                  const double *env, double *eri)
{
    (void)natm;
    int *ao_loc = (int *)malloc(sizeof(int) * (nbas + 1));
    ao_loc[0] = 0;
    for (int i = 0; i < nbas; i++) ao_loc[i + 1] = ao_loc[i] + ncart(bas[i * BAS_SLOTS + 1]);
    size_t n = ao_loc[nbas];
    int err = 0;
    #pragma omp parallel for schedule(dynamic) collapse(2)
    for (int i = 0; i < nbas; i++)
    for (int j = 0; j < nbas; j++) {
        double *buf = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
        for (int k = 0; k < nbas; k++)
        for (int l = 0; l < nbas; l++) {
            if (oracle_md_shell_quartet(atm, bas, env, i, j, k, l, buf)) { err = 1; continue; }
            int ni = ao_loc[i + 1] - ao_loc[i], nj = ao_loc[j + 1] - ao_loc[j];
            int nk = ao_loc[k + 1] - ao_loc[k], nl = ao_loc[l + 1] - ao_loc[l];
            for (int a = 0; a < ni; a++)
            for (int b = 0; b < nj; b++)
            for (int c = 0; c < nk; c++)
            for (int d = 0; d < nl; d++)
                eri[(((ao_loc[i] + a) * n + ao_loc[j] + b) * n + ao_loc[k] + c) * n + ao_loc[l] + d] =
                    buf[((a * nj + b) * nk + c) * nl + d];
        }
        free(buf);
    }
    free(ao_loc);
    return err;
}

/* One-electron integrals by McMurchie-Davidson, for pinning the product's integrals.
 * S from E^ab_0, V from R_tuv(p, P - C); the kinetic energy uses the symmetric form
 * T_ab = 1/2 sum_d <d_d a | d_d b> with d/dx (x^i e^(-a x^2)) = i x^(i-1) - 2a x^(i+1),
 * a different route from the -1/2 <a|lap b> form.  Cartesian basis, nao x nao each. */
int oracle_md_int1e(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                    const double *env, double *S, double *T, double *V)
{
    int *ao_loc = (int *)malloc(sizeof(int) * (nbas + 1));
    ao_loc[0] = 0;
    for (int i = 0; i < nbas; i++) ao_loc[i + 1] = ao_loc[i] + ncart(bas[i * BAS_SLOTS + 1]);
    int n = ao_loc[nbas];
    size_t cube = (size_t)RD * RD * RD;
    for (int i = 0; i < nbas; i++)
    for (int j = 0; j < nbas; j++) {
        shell_t A = get_shell(atm, bas, env, i), B = get_shell(atm, bas, env, j);
        if (A.l > 3 || B.l > 3) { free(ao_loc); return 1; }
        int ax[15], ay[15], az[15], bx[15], by[15], bz[15];
        cart_powers(A.l, ax, ay, az);
        cart_powers(B.l, bx, by, bz);
        int nfa = ncart(A.l), nfb = ncart(B.l);
        double *R0 = (double *)malloc(sizeof(double) * cube);
        double *work = (double *)malloc(sizeof(double) * cube * (A.l + B.l + 1));
        double norm = fac_sp(A.l) * fac_sp(B.l);
        for (int fa = 0; fa < nfa; fa++)
        for (int fb = 0; fb < nfb; fb++) {
            int I = ao_loc[i] + fa, J = ao_loc[j] + fb;
            S[I * n + J] = T[I * n + J] = V[I * n + J] = 0.0;
        }
        for (int ip = 0; ip < A.nprim; ip++)
        for (int jp = 0; jp < B.nprim; jp++) {
            double a = A.exp[ip], b = B.exp[jp], p = a + b;
            double cc = A.coef[ip] * B.coef[jp] * norm;
            double E[3][EDIM][EDIM][ETDIM], P[3];
            for (int d = 0; d < 3; d++) {
                P[d] = (a * A.r[d] + b * B.r[d]) / p;
                hermite_e(A.l + 1, B.l + 1, a, b, A.r[d], B.r[d], E[d]);
            }
            double sq = sqrt(M_PI / p);
            #define S1(d, ii, jj) (((ii) < 0 || (jj) < 0) ? 0.0 : E[d][ii][jj][0] * sq)
            for (int fa = 0; fa < nfa; fa++)
            for (int fb = 0; fb < nfb; fb++) {
                int pa[3] = {ax[fa], ay[fa], az[fa]}, pb[3] = {bx[fb], by[fb], bz[fb]};
                double s[3], k[3];
                for (int d = 0; d < 3; d++) {
                    int ii = pa[d], jj = pb[d];
                    s[d] = S1(d, ii, jj);
                    k[d] = 0.5 * (ii * jj * S1(d, ii - 1, jj - 1) - 2.0 * b * ii * S1(d, ii - 1, jj + 1)
                                  - 2.0 * a * jj * S1(d, ii + 1, jj - 1) + 4.0 * a * b * S1(d, ii + 1, jj + 1));
                }
                int I = ao_loc[i] + fa, J = ao_loc[j] + fb;
                S[I * n + J] += cc * s[0] * s[1] * s[2];
                T[I * n + J] += cc * (k[0] * s[1] * s[2] + s[0] * k[1] * s[2] + s[0] * s[1] * k[2]);
            }
            #undef S1
            for (int c = 0; c < natm; c++) {
                const double *rc = env + atm[c * ATM_SLOTS + 1];
                hermite_r(A.l + B.l, p, P[0] - rc[0], P[1] - rc[1], P[2] - rc[2], R0, work);
                double f = -atm[c * ATM_SLOTS] * 2.0 * M_PI / p * cc;
                for (int fa = 0; fa < nfa; fa++)
                for (int fb = 0; fb < nfb; fb++) {
                    double s = 0.0;
                    for (int t = 0; t <= ax[fa] + bx[fb]; t++)
                    for (int u = 0; u <= ay[fa] + by[fb]; u++)
                    for (int v = 0; v <= az[fa] + bz[fb]; v++)
                        s += E[0][ax[fa]][bx[fb]][t] * E[1][ay[fa]][by[fb]][u] * E[2][az[fa]][bz[fb]][v]
                             * R0[ridx(t, u, v)];
                    V[(ao_loc[i] + fa) * n + ao_loc[j] + fb] += f * s;
                }
            }
        }
        free(R0);
        free(work);
    }
    free(ao_loc);
    return 0;
}
