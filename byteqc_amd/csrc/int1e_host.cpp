// One-electron integrals (overlap, kinetic, nuclear attraction) for the SCF driver,
// host C++ with std::thread.  McMurchie-Davidson form:
//   S_ab = prod_d E^ab_0(d) (pi/p)^(3/2)
//   T_ab from 1D overlaps: T_ij = -1/2 [j(j-1) S_i,j-2 - 2b(2j+1) S_ij + 4b^2 S_i,j+2]
//   V_ab = -sum_C Z_C 2 pi / p sum_tuv E^ab_t E^ab_u E^ab_v R_tuv(p, P - C)
// Cartesian functions as in the J/K kernels: c_env * fac_sp(l) * x^i y^j z^k e^(-a r^2).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/mi355jk.h"

namespace {

constexpr int LMAX = 3;
constexpr int EI = LMAX + 1, EJ = LMAX + 3, ET = 2 * LMAX + 3;
constexpr int LR = 2 * LMAX;  // max t+u+v for V

double fac_sp(int l) { return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0); }

// Boys F_0..F_n(T): series at the top order plus downward recursion for T < 35,
// F_0 = sqrt(pi/T) erf(sqrt T)/2 and upward recursion above.
void boys(int n, double T, double *f)
{
    const double e = exp(-T);
    if (T < 35.0) {
        double term = 1.0 / (2 * n + 1), sum = term;
        for (int k = 1; k < 300; ++k) {
            term *= 2.0 * T / (2 * n + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        f[n] = e * sum;
        for (int m = n - 1; m >= 0; --m) f[m] = (2.0 * T * f[m + 1] + e) / (2 * m + 1);
    } else {
        f[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
        for (int m = 0; m < n; ++m) f[m + 1] = ((2 * m + 1) * f[m] - e) / (2.0 * T);
    }
}

// E[i][j][t], i <= li, j <= lj, in one direction; E[0][0][0] = exp(-mu X_AB^2)
void hermite_coeffs(int li, int lj, double a, double b, double xa, double xb, double E[EI][EJ][ET])
{
    const double p = a + b, xp = (a * xa + b * xb) / p;
    const double pa = xp - xa, pb = xp - xb, o2p = 0.5 / p;
    for (int i = 0; i <= li; ++i)
        for (int j = 0; j <= lj; ++j)
            for (int t = 0; t < ET; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = exp(-a * b / p * (xa - xb) * (xa - xb));
    for (int j = 1; j <= lj; ++j)
        for (int t = 0; t <= j; ++t)
            E[0][j][t] = (t ? o2p * E[0][j - 1][t - 1] : 0.0) + pb * E[0][j - 1][t] +
                         (t + 1 <= j - 1 ? (t + 1) * E[0][j - 1][t + 1] : 0.0);
    for (int i = 1; i <= li; ++i)
        for (int j = 0; j <= lj; ++j)
            for (int t = 0; t <= i + j; ++t)
                E[i][j][t] = (t ? o2p * E[i - 1][j][t - 1] : 0.0) + pa * E[i - 1][j][t] +
                             (t + 1 <= i - 1 + j ? (t + 1) * E[i - 1][j][t + 1] : 0.0);
}

struct Shell {
    int l, nprim, atom;
    const double *exp, *coef;
    double r[3];
};

void cart_powers(int l, int *px, int *py, int *pz)
{
    int n = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly, ++n) {
            px[n] = lx;
            py[n] = ly;
            pz[n] = l - lx - ly;
        }
}

// R^0_tuv for t+u+v <= L (cube of side LR+1), R^n_000 = (-2p)^n F_n(p |PC|^2)
void hermite_r(int L, double p, double X, double Y, double Z, double *R0)
{
    constexpr int D = LR + 1;
    static thread_local std::vector<double> work;
    work.assign((size_t)(L + 1) * D * D * D, 0.0);
    auto idx = [&](int n, int t, int u, int v) { return ((size_t)(n * D + t) * D + u) * D + v; };
    double F[LR + 1];
    boys(L, p * (X * X + Y * Y + Z * Z), F);
    double pw = 1.0;
    for (int n = 0; n <= L; ++n) {
        work[idx(n, 0, 0, 0)] = pw * F[n];
        pw *= -2.0 * p;
    }
    for (int N = 1; N <= L; ++N)
        for (int n = 0; n <= L - N; ++n)
            for (int t = 0; t <= N; ++t)
                for (int u = 0; u <= N - t; ++u) {
                    const int v = N - t - u;
                    double val;
                    if (t > 0) val = X * work[idx(n + 1, t - 1, u, v)] + (t > 1 ? (t - 1) * work[idx(n + 1, t - 2, u, v)] : 0.0);
                    else if (u > 0) val = Y * work[idx(n + 1, t, u - 1, v)] + (u > 1 ? (u - 1) * work[idx(n + 1, t, u - 2, v)] : 0.0);
                    else val = Z * work[idx(n + 1, t, u, v - 1)] + (v > 1 ? (v - 1) * work[idx(n + 1, t, u, v - 2)] : 0.0);
                    work[idx(n, t, u, v)] = val;
                }
    memcpy(R0, work.data(), sizeof(double) * D * D * D);
}

void pair_block(const Shell &A, const Shell &B, int natm, const double *zc, const double *rc,
                double *S, double *T, double *V, double *buf_r)
{
    constexpr int D = LR + 1;
    const int nfa = (A.l + 1) * (A.l + 2) / 2, nfb = (B.l + 1) * (B.l + 2) / 2;
    int ax[10], ay[10], az[10], bx[10], by[10], bz[10];
    cart_powers(A.l, ax, ay, az);
    cart_powers(B.l, bx, by, bz);
    for (int k = 0; k < nfa * nfb; ++k) S[k] = T[k] = V[k] = 0.0;
    const double norm = fac_sp(A.l) * fac_sp(B.l);
    double E[3][EI][EJ][ET];
    const double rr = (A.r[0] - B.r[0]) * (A.r[0] - B.r[0]) + (A.r[1] - B.r[1]) * (A.r[1] - B.r[1]) +
                      (A.r[2] - B.r[2]) * (A.r[2] - B.r[2]);
    for (int ip = 0; ip < A.nprim; ++ip)
        for (int jp = 0; jp < B.nprim; ++jp) {
            const double a = A.exp[ip], b = B.exp[jp], p = a + b;
            const double cc = A.coef[ip] * B.coef[jp] * norm;
            if (fabs(cc) * exp(-a * b / p * rr) < 1e-22) continue;
            double P[3];
            for (int d = 0; d < 3; ++d) {
                P[d] = (a * A.r[d] + b * B.r[d]) / p;
                hermite_coeffs(A.l, B.l + 2, a, b, A.r[d], B.r[d], E[d]);
            }
            const double sq = sqrt(M_PI / p);
            // 1D overlaps and kinetic pieces
            auto s1 = [&](int d, int i, int j) { return j < 0 ? 0.0 : E[d][i][j][0] * sq; };
            auto t1 = [&](int d, int i, int j) {
                return -0.5 * (j * (j - 1) * s1(d, i, j - 2) - 2.0 * b * (2 * j + 1) * s1(d, i, j) +
                               4.0 * b * b * s1(d, i, j + 2));
            };
            for (int fa = 0; fa < nfa; ++fa)
                for (int fb = 0; fb < nfb; ++fb) {
                    const double sx = s1(0, ax[fa], bx[fb]), sy = s1(1, ay[fa], by[fb]), sz = s1(2, az[fa], bz[fb]);
                    S[fa * nfb + fb] += cc * sx * sy * sz;
                    T[fa * nfb + fb] += cc * (t1(0, ax[fa], bx[fb]) * sy * sz + sx * t1(1, ay[fa], by[fb]) * sz +
                                              sx * sy * t1(2, az[fa], bz[fb]));
                }
            const int L = A.l + B.l;
            for (int c = 0; c < natm; ++c) {
                hermite_r(L, p, P[0] - rc[3 * c], P[1] - rc[3 * c + 1], P[2] - rc[3 * c + 2], buf_r);
                const double f = -zc[c] * 2.0 * M_PI / p * cc;
                for (int fa = 0; fa < nfa; ++fa)
                    for (int fb = 0; fb < nfb; ++fb) {
                        double s = 0.0;
                        for (int t = 0; t <= ax[fa] + bx[fb]; ++t)
                            for (int u = 0; u <= ay[fa] + by[fb]; ++u)
                                for (int v = 0; v <= az[fa] + bz[fb]; ++v)
                                    s += E[0][ax[fa]][bx[fb]][t] * E[1][ay[fa]][by[fb]][u] * E[2][az[fa]][bz[fb]][v] *
                                         buf_r[(t * D + u) * D + v];
                        V[fa * nfb + fb] += f * s;
                    }
            }
        }
}

}  // namespace

extern "C" int mjk_int1e(const int32_t *atm, int natm, const int32_t *bas, int nbas, const double *env,
                         double *s, double *t, double *v, int nthreads)
{
    if (!atm || !bas || !env || !s || !t || !v || nbas <= 0) return MJK_ERR_ARG;
    std::vector<Shell> sh(nbas);
    std::vector<int> ao(nbas + 1, 0);
    for (int i = 0; i < nbas; ++i) {
        const int32_t *b = bas + 8 * i;
        if (b[1] > LMAX) return MJK_ERR_UNSUPPORTED;
        if (b[3] != 1) return MJK_ERR_ARG;
        sh[i].l = b[1];
        sh[i].nprim = b[2];
        sh[i].atom = b[0];
        sh[i].exp = env + b[5];
        sh[i].coef = env + b[6];
        const double *r = env + atm[6 * b[0] + 1];
        for (int d = 0; d < 3; ++d) sh[i].r[d] = r[d];
        ao[i + 1] = ao[i] + (b[1] + 1) * (b[1] + 2) / 2;
    }
    std::vector<double> zc(natm), rc(3 * natm);
    for (int c = 0; c < natm; ++c) {
        zc[c] = atm[6 * c];
        for (int d = 0; d < 3; ++d) rc[3 * c + d] = env[atm[6 * c + 1] + d];
    }
    const size_t n = ao[nbas];
    const long npair = (long)nbas * (nbas + 1) / 2;
    std::atomic<long> next(0);
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    auto work = [&]() {
        constexpr int D = LR + 1;
        std::vector<double> rbuf(D * D * D), bs(100), bt(100), bv(100);
        for (;;) {
            const long k = next.fetch_add(16);
            if (k >= npair) break;
            for (long kk = k; kk < std::min(npair, k + 16); ++kk) {
                int i = (int)((sqrt(8.0 * kk + 1.0) - 1.0) / 2.0);
                while ((long)i * (i + 1) / 2 > kk) --i;
                while ((long)(i + 1) * (i + 2) / 2 <= kk) ++i;
                const int j = (int)(kk - (long)i * (i + 1) / 2);
                pair_block(sh[i], sh[j], natm, zc.data(), rc.data(), bs.data(), bt.data(), bv.data(), rbuf.data());
                const int ni = ao[i + 1] - ao[i], nj = ao[j + 1] - ao[j];
                for (int a = 0; a < ni; ++a)
                    for (int b = 0; b < nj; ++b) {
                        const size_t x = (ao[i] + a) * n + ao[j] + b, y = (ao[j] + b) * n + ao[i] + a;
                        s[x] = s[y] = bs[a * nj + b];
                        t[x] = t[y] = bt[a * nj + b];
                        v[x] = v[y] = bv[a * nj + b];
                    }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int k = 0; k < nthreads; ++k) pool.emplace_back(work);
    for (auto &th : pool) th.join();
    return MJK_OK;
}
