/* CPU oracle for the direct-SCF J/K path.  TEST INFRASTRUCTURE ONLY: nothing in
 * byteqc_amd/ links, loads or calls this code; tests/, the smoke check and the
 * bench's CPU-baseline leg use it as the checker.
 *
 * Integral tables use the PySCF layout (_atm: 6 ints per atom, _bas: 8 ints per
 * shell, _env: doubles) of a Cartesian, segmented (nctr == 1) basis.  AO order is
 * shell order, Cartesian components with x power descending
 * (cuobc/lib/gvhf/contract_jk.cu:228-260).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boys function F_0..F_nmax(T) into f[0..nmax]. */
void oracle_boys(int nmax, double T, double *f);

/* Rys nodes x_i = t_i^2 (ascending) and weights of exp(-T t^2) on [0,1]. */
void oracle_rys_roots(int n, double T, double *x, double *w);
/* Same, straight from the discretised Stieltjes procedure (no table). */
void oracle_rys_roots_direct(int n, double T, double *x, double *w);

/* Cartesian overlap, kinetic and nuclear-attraction matrices (nao x nao each). */
int oracle_md_int1e(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                    const double *env, double *S, double *T, double *V);

/* Full Cartesian ERI tensor (nao^4, row-major (ij|kl)) by McMurchie-Davidson. */
int oracle_md_eri(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                  const double *env, double *eri);

/* One contracted Cartesian shell quartet by McMurchie-Davidson:
 * out[nfi][nfj][nfk][nfl]. */
int oracle_md_shell_quartet(const int32_t *atm, const int32_t *bas, const double *env,
                            int ish, int jsh, int ksh, int lsh, double *out);

/* Same quartet by Rys quadrature, following the reference's 2D recursion. */
int oracle_rys_shell_quartet(const int32_t *atm, const int32_t *bas, const double *env,
                             int ish, int jsh, int ksh, int lsh, double *out);

/* Schwarz factors q[i*nbas+j] = sqrt(max |(ij|ij)|) by Rys quadrature. */
int oracle_q_cond(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                  const double *env, double *q, int nthreads);

/* Direct J/K over unique shell quartets with Schwarz screening (q may be NULL:
 * no screening).  dm, vj, vk: n_dm x nao x nao row-major; vj/vk may be NULL. */
int oracle_rys_jk(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                  const double *env, const double *q, double cutoff,
                  const double *dm, int n_dm, int nao, double *vj, double *vk,
                  int nthreads);

/* Same, restricted to bra pair rows pij with pij % stride == offset (pairs enumerated
 * i >= j in shell order), for timing a systematic sample of the quartet work. */
int oracle_rys_jk_rows(const int32_t *atm, int natm, const int32_t *bas, int nbas,
                       const double *env, const double *q, double cutoff,
                       const double *dm, int n_dm, int nao, double *vj, double *vk,
                       int nthreads, int stride, int offset);

#ifdef __cplusplus
}
#endif
