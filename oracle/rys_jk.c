/* Rys-quadrature J/K for the CPU oracle (test infrastructure only).
 *
 * A plain-C restatement of the reference's direct J/K algorithm:
 *  - primitive pair data a12 = ai + aj, e12 = fac_sp(li) fac_sp(lj) ci cj
 *    exp(-ai aj |Ri - Rj|^2 / a12), P = (ai Ri + aj Rj)/a12, e12 halved when the two
 *    shells are the same (cuobc/lib/gint/pair_data.c:83-141, diag_fac = 0.5 from
 *    cuobc/scf/hf.py:398-402);
 *  - unique quartets (ij|kl) with i >= j, k >= l, pair(ij) >= pair(kl); the 2 pi^(5/2)
 *    norm is halved when the two pairs coincide (cuobc/lib/gvhf/g2e.cu:1022-1029,
 *    cuobc/lib/gint/g2e.c:41);
 *  - N = (li+lj+lk+ll)/2 + 1 Rys roots (g2e.c:40) in the u = t^2/(1-t^2) variable;
 *  - 2D integrals per root: u2 = a0 u, tmp4 = 1/(2 (u2 (aij+akl) + aij akl)),
 *    b00 = u2 tmp4, b10 = b00 + tmp4 akl, b01 = b00 + tmp4 aij,
 *    c00 = (P - A) - 2 b00 akl (P - Q), c0p = (Q - C) + 2 b00 aij (P - Q), vertical
 *    recursion in (n, m) then horizontal transfers to j and l; Ix[0] = norm,
 *    Iy[0] = eij ekl / (sqrt(aij + akl) aij akl), Iz[0] = w (g2e.cu:32-490);
 *  - digestion Vj_ij += s D_kl, Vj_kl += s D_ij, Vk_ik += s D_jl, Vk_il += s D_jk,
 *    Vk_jk += s D_il, Vk_jl += s D_ik (cuobc/lib/gvhf/contract_jk.cu:219-316);
 *  - J = 2 (Vj + Vj^T), K = Vk + Vk^T (cuobc/scf/hf.py:170-181).
 * Quartets are kept when q_ij q_kl > cutoff (Schwarz, hf.py:357-389); the reference
 * approximates the same test with log-q bins (nr_jk_driver.cu:252-263).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "oracle.h"

#define ATM_SLOTS 6
#define BAS_SLOTS 8
#define LMAX 4
#define NR_MAX 10
#define G1 (2 * LMAX + 1)

static double fac_sp(int l)
{
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}
static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

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

/* 2D integrals for one direction and root: g[i][j][k][l] (dims (li+1)(lj+1)(lk+1)(ll+1)). */
static void g2d(int li, int lj, int lk, int ll, double g0, double c00, double c0p,
                double b00, double b10, double b01, double ab, double cd, double *g)
{
    int nmax = li + lj, mmax = lk + ll;
    double v[G1][G1];
    v[0][0] = g0;
    for (int n = 0; n < nmax; n++)
        v[n + 1][0] = c00 * v[n][0] + (n ? n * b10 * v[n - 1][0] : 0.0);
    for (int m = 0; m < mmax; m++)
        for (int n = 0; n <= nmax; n++) {
            double t = c0p * v[n][m];
            if (m) t += m * b01 * v[n][m - 1];
            if (n) t += n * b00 * v[n - 1][m];
            v[n][m + 1] = t;
        }
    /* horizontal transfer to j: h[i][j][m], i <= nmax - j */
    double h[G1][LMAX + 1][G1];
    for (int i = 0; i <= nmax; i++)
        for (int m = 0; m <= mmax; m++) h[i][0][m] = v[i][m];
    for (int j = 1; j <= lj; j++)
        for (int i = 0; i <= nmax - j; i++)
            for (int m = 0; m <= mmax; m++)
                h[i][j][m] = h[i + 1][j - 1][m] + ab * h[i][j - 1][m];
    /* horizontal transfer to l */
    double o[LMAX + 1][LMAX + 1][G1][LMAX + 1];
    for (int i = 0; i <= li; i++)
        for (int j = 0; j <= lj; j++) {
            for (int k = 0; k <= mmax; k++) o[i][j][k][0] = h[i][j][k];
            for (int l = 1; l <= ll; l++)
                for (int k = 0; k <= mmax - l; k++)
                    o[i][j][k][l] = o[i][j][k + 1][l - 1] + cd * o[i][j][k][l - 1];
        }
    int n1 = lj + 1, n2 = lk + 1, n3 = ll + 1;
    for (int i = 0; i <= li; i++)
        for (int j = 0; j <= lj; j++)
            for (int k = 0; k <= lk; k++)
                for (int l = 0; l <= ll; l++)
                    g[((i * n1 + j) * n2 + k) * n3 + l] = o[i][j][k][l];
}

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

/* Contracted Cartesian quartet out[nfi][nfj][nfk][nfl]; pair factors diag_ij / diag_kl
 * (1 or 0.5) multiply e12 and norm_fac multiplies 2 pi^(5/2). */
static int rys_quartet(const int32_t *atm, const int32_t *bas, const double *env,
                       int ish, int jsh, int ksh, int lsh, double diag_ij, double diag_kl,
                       double norm_fac, double *out)
{
    shell_t I = get_shell(atm, bas, env, ish), J = get_shell(atm, bas, env, jsh);
    shell_t K = get_shell(atm, bas, env, ksh), L = get_shell(atm, bas, env, lsh);
    if (I.l > LMAX || J.l > LMAX || K.l > LMAX || L.l > LMAX) return 1;
    int nroots = (I.l + J.l + K.l + L.l) / 2 + 1;
    int nfi = ncart(I.l), nfj = ncart(J.l), nfk = ncart(K.l), nfl = ncart(L.l);
    int ix[15], iy[15], iz[15], jx[15], jy[15], jz[15], kx[15], ky[15], kz[15], lx[15], ly[15], lz[15];
    cart_powers(I.l, ix, iy, iz);
    cart_powers(J.l, jx, jy, jz);
    cart_powers(K.l, kx, ky, kz);
    cart_powers(L.l, lx, ly, lz);
    int d1 = J.l + 1, d2 = K.l + 1, d3 = L.l + 1;
    int gsz = (I.l + 1) * d1 * d2 * d3;
    size_t nf = (size_t)nfi * nfj * nfk * nfl;
    memset(out, 0, sizeof(double) * nf);
    double gx[625], gy[625], gz[625];
    double norm = 2.0 * pow(M_PI, 2.5) * norm_fac;
    double fsp = fac_sp(I.l) * fac_sp(J.l) * fac_sp(K.l) * fac_sp(L.l);
    double rij2 = 0, rkl2 = 0;
    for (int d = 0; d < 3; d++) {
        rij2 += (I.r[d] - J.r[d]) * (I.r[d] - J.r[d]);
        rkl2 += (K.r[d] - L.r[d]) * (K.r[d] - L.r[d]);
    }
    for (int ip = 0; ip < I.nprim; ip++)
    for (int jp = 0; jp < J.nprim; jp++) {
        double ai = I.exp[ip], aj = J.exp[jp], aij = ai + aj;
        double eij = I.coef[ip] * J.coef[jp] * exp(-rij2 * ai * aj / aij) * diag_ij;
        double P[3];
        for (int d = 0; d < 3; d++) P[d] = (ai * I.r[d] + aj * J.r[d]) / aij;
        for (int kp = 0; kp < K.nprim; kp++)
        for (int lp = 0; lp < L.nprim; lp++) {
            double ak = K.exp[kp], al = L.exp[lp], akl = ak + al;
            double ekl = K.coef[kp] * L.coef[lp] * exp(-rkl2 * ak * al / akl) * diag_kl;
            double Q[3];
            for (int d = 0; d < 3; d++) Q[d] = (ak * K.r[d] + al * L.r[d]) / akl;
            double aijkl = aij + akl, a1 = aij * akl, a0 = a1 / aijkl;
            double pq[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
            double T = a0 * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
            double xr[NR_MAX], wr[NR_MAX];
            oracle_rys_roots(nroots, T, xr, wr);
            double fac = eij * ekl / (sqrt(aijkl) * a1) * fsp;
            for (int r = 0; r < nroots; r++) {
                double u = xr[r] / (1.0 - xr[r]);
                double u2 = a0 * u;
                double tmp4 = 0.5 / (u2 * aijkl + a1);
                double b00 = u2 * tmp4, b10 = b00 + tmp4 * akl, b01 = b00 + tmp4 * aij;
                double g0[3] = {norm, fac, wr[r]};
                double *gs[3] = {gx, gy, gz};
                for (int d = 0; d < 3; d++) {
                    double c00 = (P[d] - I.r[d]) - 2.0 * b00 * akl * pq[d];
                    double c0p = (Q[d] - K.r[d]) + 2.0 * b00 * aij * pq[d];
                    g2d(I.l, J.l, K.l, L.l, g0[d], c00, c0p, b00, b10, b01,
                        I.r[d] - J.r[d], K.r[d] - L.r[d], gs[d]);
                }
                (void)gsz;
                size_t n = 0;
                for (int a = 0; a < nfi; a++)
                for (int b = 0; b < nfj; b++)
                for (int c = 0; c < nfk; c++)
                for (int e = 0; e < nfl; e++, n++) {
                    int ox = ((ix[a] * d1 + jx[b]) * d2 + kx[c]) * d3 + lx[e];
                    int oy = ((iy[a] * d1 + jy[b]) * d2 + ky[c]) * d3 + ly[e];
                    int oz = ((iz[a] * d1 + jz[b]) * d2 + kz[c]) * d3 + lz[e];
                    out[n] += gx[ox] * gy[oy] * gz[oz];
                }
            }
        }
    }
    return 0;
}

int oracle_rys_shell_quartet(const int32_t *atm, const int32_t *bas, const double *env,
                             int ish, int jsh, int ksh, int lsh, double *out)
{
    return rys_quartet(atm, bas, env, ish, jsh, ksh, lsh, 1.0, 1.0, 1.0, out);
}

static int *make_ao_loc(const int32_t *bas, int nbas)
{
    int *ao_loc = (int *)malloc(sizeof(int) * (nbas + 1));
    ao_loc[0] = 0;
    for (int i = 0; i < nbas; i++) ao_loc[i + 1] = ao_loc[i] + ncart(bas[i * BAS_SLOTS + 1]);
    return ao_loc;
}

int oracle_q_cond(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                  const double *env, double *q, int nthreads)
{
    (void)natm;
    int err = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    #pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
        #pragma omp for schedule(dynamic)
        for (int i = 0; i < nbas; i++)
            for (int j = 0; j <= i; j++) {
                if (rys_quartet(atm, bas, env, i, j, i, j, 1.0, 1.0, 1.0, buf)) { err = 1; continue; }
                int ni = ncart(bas[i * BAS_SLOTS + 1]), nj = ncart(bas[j * BAS_SLOTS + 1]);
                double m = 0.0;
                for (int a = 0; a < ni; a++)
                    for (int b = 0; b < nj; b++) {
                        double v = fabs(buf[((a * nj + b) * ni + a) * nj + b]);
                        if (v > m) m = v;
                    }
                q[i * nbas + j] = q[j * nbas + i] = sqrt(m);
            }
        free(buf);
    }
    return err;
}

int oracle_rys_jk(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                  const double *env, const double *q, double cutoff,
                  const double *dm, int n_dm, int nao, double *vj, double *vk,
                  int nthreads)
{
    return oracle_rys_jk_rows(atm, natm, bas, nbas, env, q, cutoff, dm, n_dm, nao, vj, vk,
                              nthreads, 1, 0);
}

int oracle_rys_jk_rows(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                       const double *env, const double *q, double cutoff,
                       const double *dm, int n_dm, int nao, double *vj, double *vk,
                       int nthreads, int stride, int offset)
{
    (void)natm;
    int *ao_loc = make_ao_loc(bas, nbas);
    if (ao_loc[nbas] != nao) { free(ao_loc); return 2; }
    size_t nn = (size_t)nao * nao;
    int npair = nbas * (nbas + 1) / 2;
    int *pi = (int *)malloc(sizeof(int) * npair), *pj = (int *)malloc(sizeof(int) * npair);
    for (int i = 0, p = 0; i < nbas; i++)
        for (int j = 0; j <= i; j++, p++) { pi[p] = i; pj[p] = j; }
    if (nthreads > 0) omp_set_num_threads(nthreads);
    int nt = omp_get_max_threads();
    double *accj = vj ? (double *)calloc((size_t)nt * n_dm * nn, sizeof(double)) : NULL;
    double *acck = vk ? (double *)calloc((size_t)nt * n_dm * nn, sizeof(double)) : NULL;
    int err = 0;
    #pragma omp parallel
    {
        int tid = omp_get_thread_num();
        double *buf = (double *)malloc(sizeof(double) * 15 * 15 * 15 * 15);
        double *myj = accj ? accj + (size_t)tid * n_dm * nn : NULL;
        double *myk = acck ? acck + (size_t)tid * n_dm * nn : NULL;
        #pragma omp for schedule(dynamic, 1)
        for (int pij = npair - 1; pij >= 0; pij--) {
            if (stride > 1 && pij % stride != offset) continue;
            int i = pi[pij], j = pj[pij];
            for (int pkl = 0; pkl <= pij; pkl++) {
                int k = pi[pkl], l = pj[pkl];
                if (q && q[i * nbas + j] * q[k * nbas + l] <= cutoff) continue;
                double dij = (i == j) ? 0.5 : 1.0, dkl = (k == l) ? 0.5 : 1.0;
                double nf = (pij == pkl) ? 0.5 : 1.0;
                if (rys_quartet(atm, bas, env, i, j, k, l, dij, dkl, nf, buf)) { err = 1; continue; }
                int i0 = ao_loc[i], j0 = ao_loc[j], k0 = ao_loc[k], l0 = ao_loc[l];
                int ni = ao_loc[i + 1] - i0, nj = ao_loc[j + 1] - j0;
                int nk = ao_loc[k + 1] - k0, nl = ao_loc[l + 1] - l0;
                for (int n = 0; n < n_dm; n++) {
                    const double *D = dm + n * nn;
                    double *Vj = myj ? myj + n * nn : NULL, *Vk = myk ? myk + n * nn : NULL;
                    size_t c = 0;
                    for (int a = i0; a < i0 + ni; a++)
                    for (int b = j0; b < j0 + nj; b++)
                    for (int cc = k0; cc < k0 + nk; cc++)
                    for (int d = l0; d < l0 + nl; d++, c++) {
                        double s = buf[c];
                        if (Vj) {
                            Vj[a * nao + b] += s * D[cc * nao + d];
                            Vj[cc * nao + d] += s * D[a * nao + b];
                        }
                        if (Vk) {
                            Vk[a * nao + cc] += s * D[b * nao + d];
                            Vk[a * nao + d] += s * D[b * nao + cc];
                            Vk[b * nao + cc] += s * D[a * nao + d];
                            Vk[b * nao + d] += s * D[a * nao + cc];
                        }
                    }
                }
            }
        }
        free(buf);
    }
    for (int n = 0; n < n_dm; n++) {
        for (size_t x = 0; x < nn; x++) {
            double sj = 0.0, sk = 0.0;
            for (int t = 0; t < nt; t++) {
                if (accj) sj += accj[((size_t)t * n_dm + n) * nn + x];
                if (acck) sk += acck[((size_t)t * n_dm + n) * nn + x];
            }
            if (vj) vj[n * nn + x] = sj;
            if (vk) vk[n * nn + x] = sk;
        }
        for (int a = 0; a < nao; a++)
            for (int b = 0; b <= a; b++) {
                if (vj) {
                    double s = 2.0 * (vj[n * nn + a * nao + b] + vj[n * nn + b * nao + a]);
                    vj[n * nn + a * nao + b] = vj[n * nn + b * nao + a] = s;
                }
                if (vk) {
                    double s = vk[n * nn + a * nao + b] + vk[n * nn + b * nao + a];
                    vk[n * nn + a * nao + b] = vk[n * nn + b * nao + a] = s;
                }
            }
    }
    free(accj);
    free(acck);
    free(pi);
    free(pj);
    free(ao_loc);
    return err;
}
