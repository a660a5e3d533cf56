/* mi355jk: MI355X-native direct-SCF Coulomb/exchange (J/K) builder, C ABI.
 *
 * All entry points are extern "C", take plain pointers, sizes and scalars, return an
 * int status (MJK_OK on success) and never exit the process.  The text of the last
 * error on the calling thread is returned by mjk_last_error().  Device pointers are
 * HIP device allocations on the current device; `stream` is a hipStream_t passed as
 * void* (NULL = the null stream).  Launches are asynchronous on that stream.
 *
 * The entry points replace ByteQC's ctypes interface to libgint/libgvhf:
 *   mjk_q_cond              <- PySCF VHFOpt.get_q_cond (CVHFsetnr_direct_scf), used at
 *                              cuobc/scf/hf.py:323-341
 *   mjk_init_basis_prod     <- GINTinit_basis_prod_cpu + GINTinit_basis_prod_gpu
 *                              (cuobc/lib/gint/nr_fill_ao_ints.cu:452-491)
 *   mjk_del_basis_prod      <- GINTdel_basis_prod (nr_fill_ao_ints.cu:394-410)
 *   mjk_build_jk            <- GINTbuild_jk (cuobc/lib/gvhf/nr_jk_driver.cu:227-283)
 *   mjk_build_jk_all        <- the class-pair loop of get_jk (cuobc/scf/hf.py:111-166)
 *                              done natively, one call per J/K build
 */
#ifndef MI355JK_H
#define MI355JK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MJK_VERSION 1

enum {
    MJK_OK = 0,
    MJK_ERR_ARG = 1,          /* bad argument (null pointer, size, class index) */
    MJK_ERR_UNSUPPORTED = 2,  /* angular momentum above MJK_LMAX */
    MJK_ERR_HIP = 3,          /* a HIP runtime call or kernel launch failed */
    MJK_ERR_OOM = 4           /* device or host allocation failed */
};

#define MJK_LMAX 3

/* Opaque device-resident basis-product cache (shells, shell pairs, primitive pairs). */
typedef struct mjk_basis_prod mjk_basis_prod;

int mjk_version(void);
const char *mjk_last_error(void);

/* Schwarz factors q[i * nbas + j] = sqrt(max_{a in i, b in j} |(ab|ab)|) for a
 * Cartesian, segmented basis in PySCF _atm/_bas/_env layout.  q is a HOST array of
 * nbas * nbas doubles.  Runs on the current device and synchronises. */
int mjk_q_cond(const int32_t *atm, int natm, const int32_t *bas, int nbas,
               const double *env, double *q);

/* Build the device cache from host tables (copied; nothing is borrowed).
 *   diag_fac        factor on e12 when the two shells of a pair coincide (0.5 for the
 *                   lower-triangle pair lists, as in cuobc/scf/hf.py:398-402)
 *   bas_pair2shls   int32[2][n_pairs]: bra shells then ket shells of every pair
 *   bas_pairs_locs  int32[ncptype + 1]: pair-class offsets into bas_pair2shls; all
 *                   pairs of a class share (l_i, nprim_i, l_j, nprim_j) with l_i >= l_j
 *   pair_q          double[n_pairs]: Schwarz factor of every pair.  Pairs may come in
 *                   any order; ordering them by (i, j) within a class (shells sorted in
 *                   space) lets the kernels reuse exchange partial sums across kets
 *   atm/bas/env     Cartesian, segmented basis (PySCF layout); env coefficients carry
 *                   the radial normalisation.
 */
int mjk_init_basis_prod(mjk_basis_prod **pbp, double diag_fac,
                        const int32_t *bas_pair2shls, const int32_t *bas_pairs_locs,
                        int ncptype, const double *pair_q,
                        const int32_t *atm, int natm, const int32_t *bas, int nbas,
                        const double *env);

/* Free the cache; *pbp is set to NULL.  Safe on NULL / already freed. */
int mjk_del_basis_prod(mjk_basis_prod **pbp);

int mjk_nao(const mjk_basis_prod *bp);

/* Per-build screening and precision controls. */
typedef struct {
    double direct_scf_tol;    /* quartet kept when q_ij q_kl dmax > tol (1e-13) */
    double precision_cutoff;  /* quartets with q_ij q_kl dmax below this run in FP32;
                                 0 = everything FP64 */
    int density_screen;       /* 1: dmax is the max |D| over the 6 shell blocks the quartet
                                 touches (recomputed from dm each build); 0: dmax = 1;
                                 2 (mjk_build_jk only): reuse the block maxima computed by
                                 the previous call on this cache */
    int reserved[3];
} mjk_screen;

/* Accumulate one class pair (cp_ij >= cp_kl) into vj/vk (device, n_dm x nao x nao,
 * row-major, NOT zeroed here).  vj or vk may be NULL.  What is accumulated is the
 * unsymmetrised V of the reference: J = 2 (V_j + V_j^T), K = V_k + V_k^T. */
int mjk_build_jk(mjk_basis_prod *bp, double *vj, double *vk, const double *dm,
                 int nao, int n_dm, int cp_ij, int cp_kl, const mjk_screen *screen,
                 void *stream);

/* Zero vj/vk, run every class pair, and symmetrise in place:
 * vj <- 2 (V_j + V_j^T), vk <- V_k + V_k^T.  vj or vk may be NULL. */
int mjk_build_jk_all(mjk_basis_prod *bp, double *vj, double *vk, const double *dm,
                     int nao, int n_dm, const mjk_screen *screen, void *stream);

/* Counters of the last mjk_build_jk_all on this cache (host values, synchronises):
 * out[0] quartets computed in FP64, out[1] in FP32, out[2] kernel launches, out[3] quartet
 * slots the register and group kernels executed (lanes or lane groups times ket
 * iterations), so out[0..1] / out[3] is their lane utilisation. */
int mjk_last_stats(mjk_basis_prod *bp, double *out);

/* Overlap, kinetic and nuclear-attraction matrices (host arrays, nao x nao each) of a
 * Cartesian segmented basis, on nthreads host threads (<= 0: all cores).  The SCF
 * driver's one-electron part; it replaces PySCF's int1e_ovlp / int1e_kin / int1e_nuc
 * that the reference's RHF inherits (cuobc/scf/hf.py:272-274). */
int mjk_int1e(const int32_t *atm, int natm, const int32_t *bas, int nbas, const double *env,
              double *s, double *t, double *v, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
