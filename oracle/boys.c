/* Boys function for the CPU oracle (test infrastructure only).
 *
 * F_n(T) = int_0^1 t^(2n) exp(-T t^2) dt.
 *   T < 30: the series F_n(T) = exp(-T) sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))
 *           at the top order, then the stable downward recursion
 *           F_n = (2T F_{n+1} + exp(-T)) / (2n+1);
 *   T >= 30: F_0 = sqrt(pi/T) erf(sqrt(T)) / 2 and the upward recursion
 *           F_{n+1} = ((2n+1) F_n - exp(-T)) / (2T), stable for large T.
 * The reference evaluates F_0 the same way for (ss|ss) (cuobc/lib/gvhf/g2e.cu:1418).
 */
#include <math.h>
#include "oracle.h"

void oracle_boys(int nmax, double T, double *f)
{
    double e = exp(-T);
    if (T < 30.0) {
        double term = 1.0 / (2 * nmax + 1);
        double sum = term;
        for (int k = 1; k < 400; k++) {
            term *= 2.0 * T / (2 * nmax + 2 * k + 1);
            sum += term;
            if (term < 1e-18 * sum) break;
        }
        f[nmax] = e * sum;
        for (int n = nmax - 1; n >= 0; n--)
            f[n] = (2.0 * T * f[n + 1] + e) / (2 * n + 1);
    } else {
        double st = sqrt(T);
        f[0] = 0.5 * sqrt(M_PI) / st * erf(st);
        for (int n = 0; n < nmax; n++)
            f[n + 1] = ((2 * n + 1) * f[n] - e) / (2.0 * T);
    }
}
