/* mi355jk: MI355X-native direct-SCF Coulomb/exchange (J/K) builder, C ABI.
 *
 * All entry points are extern "C", take plain pointers, sizes and scalars, return an
 * int status (MJK_OK on success) and never exit the process.  The text of the last
 * error on the calling thread is returned by mjk_last_error().  Device pointers are
 * HIP device allocations on the current device; `stream` is a hipStream_t passed as
 * void* (NULL = the null stream).  Launches are asynchronous on that stream.  A cache
 * (mjk_basis_prod) is used by one host thread at a time: its builds share the cache's
 * side streams, counters and density-maxima workspace; separate caches are independent.
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

/* Streams one mjk_build_jk_all spreads its class pairs over: the caller's stream plus
 * lanes - 1 side streams of the cache (default priority), joined back into the caller's
 * stream before the call's last kernels.  1 runs everything on the caller's stream;
 * default 4.  No counterpart in the reference, whose build launches on one stream. */
#define MJK_MAX_LANES 8
int mjk_set_streams(mjk_basis_prod *bp, int lanes);

/* Longest chunk of ket pairs one workgroup of the register and group kernels walks
 * (0: the kernels' defaults, 128 and 32).  Shorter chunks mean more, shorter tiles.
 * No counterpart in the reference. */
int mjk_set_ket_chunk(mjk_basis_prod *bp, int cap);

/* Test hook for failure handling: the next n device allocations of the library's caches
 * succeed and every later one fails with MJK_ERR_OOM (no device memory is touched for it),
 * until called with n < 0.  Process-wide.  No counterpart in the reference, whose allocation
 * helpers exit() on a device error (cuobc/lib/gint/cuda_alloc.cuh:22-31). */
int mjk_fail_alloc_after(int n);

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

/* Add the Cartesian ERI block (ij|kl) of every pair of bra class cp_ij and ket class cp_kl
 * into eri (device), element (a, b, c, d) of shells (i, j, k, l) at
 *   sum over (ao_i + a - ao_offsets[0]) * strides[0], (ao_j + b - ao_offsets[1]) * strides[1],
 *            (ao_k + c - ao_offsets[2]) * strides[2], (ao_l + d - ao_offsets[3]) * strides[3].
 * Replaces GINTfill_int2e (cuobc/lib/gint/nr_fill_ao_ints.cu:493-551) without its bins: every
 * listed pair of both classes is computed.  Build the cache with diag_fac = 1 (full pair
 * lists on diagonal groups, as get_int4c's diag_block_with_triu) for plain ERIs.  Strides are
 * 64-bit (the reference's are int). */
int mjk_fill_int2e(mjk_basis_prod *bp, double *eri, int nao, const int64_t *strides,
                   const int32_t *ao_offsets, int cp_ij, int cp_kl, void *stream);

/* ---- Batched builds: many molecules, one launch per class pair ------------------------
 * No counterpart in the reference, whose get_jk handles one molecule per call and whose
 * multi-GPU map hands whole class pairs to devices (cuobc/lib/multigpu.py:264-291).
 *
 * The caller builds ONE basis-product cache over all molecules (shells and atoms
 * concatenated; only intra-molecule shell pairs listed; inside each class the pairs of
 * molecule 0 first, then molecule 1, ...) and hands it to mjk_batch_create, which takes
 * ownership (*pbp is set to NULL on success).  mol_pair_off is [ncptype][n_mol + 1]: the
 * pairs of molecule m in class c are [off[c][m], off[c][m+1]) of that class.  Molecule m's
 * D, J and K are nao[m] x nao[m] row-major at element dm_off[m] of the batch arrays, and
 * its shells' Cartesian AO indices start at ao_base[m] of the concatenated basis.  Every
 * quartet class launches once for all molecules over a host-built table of tiles. */
typedef struct mjk_batch mjk_batch;
int mjk_batch_create(mjk_batch **pb, mjk_basis_prod **pbp, int n_mol, const int32_t *mol_pair_off,
                     const int64_t *dm_off, const int32_t *nao, const int32_t *ao_base);
int mjk_batch_destroy(mjk_batch **pb);
/* J = 2(V + V^T), K = V + V^T per molecule (vj or vk may be NULL); Schwarz screening only
 * (screen->density_screen must be 0).  use_graph != 0 records the whole build into a
 * hipGraph on the first call and replays it while (vj, vk, dm, tolerances) stay the same;
 * it needs a non-default stream.  The graph holds no memset nodes. */
int mjk_batch_build(mjk_batch *b, double *vj, double *vk, const double *dm, const mjk_screen *screen,
                    int use_graph, void *stream);
int mjk_batch_num_class_pairs(const mjk_batch *b);
/* Per class pair, 10 int32 each: bra class, ket class, l_a, l_b, l_c, l_d, ket chunk, tiles,
 * kernel family (0 register, 1 group, 2 block-LDS) and bra pairs per tile. */
int mjk_batch_tiles(const mjk_batch *b, int32_t *out, int n);
/* Set the ket chunk (ket pairs per tile) of every class pair, in mjk_batch_tiles order. */
int mjk_batch_set_kc(mjk_batch *b, const int32_t *kc, int n);
/* Average time (ms, hipEvents) of class pair i over reps launches after one warm-up; the
 * launches accumulate into vj/vk, so the caller rebuilds afterwards.  The autotuner's probe. */
int mjk_batch_time_class_pair(mjk_batch *b, int i, double *vj, double *vk, const double *dm,
                              const mjk_screen *screen, int reps, void *stream, float *ms);
/* Counters of the last build, as mjk_last_stats. */
int mjk_batch_stats(mjk_batch *b, double *out);

/* ---- Learned initial guess: perceptron kernels ------------------------------------
 * No counterpart in the reference, whose SCF starts from PySCF's minao/SAD guess
 * (PySCF RHF.get_init_guess, inherited at cuobc/scf/hf.py:272-274).  All pointers are
 * device pointers; bf16 values are uint16_t bit patterns (round to nearest even); every
 * call is asynchronous on `stream`.  Matrices are row-major.
 *
 * c[m][n] = op(a) op(b) + beta c, bf16 operands, fp32 accumulate on MFMA:
 *   layout 0: a is [m][k] (lda), b is [k][n] (ldb)
 *   layout 1: a is [m][k],       b is [n][k]        (c = a b^T)
 *   layout 2: a is [k][m],       b is [k][n]        (c = a^T b) */
int mjk_gemm_bf16(int layout, int m, int n, int k, const uint16_t *a, int lda,
                  const uint16_t *b, int ldb, float *c, int ldc, float beta, void *stream);
/* z[m][n] += bias[n] (bias may be NULL); a_out = bf16(act(z)) if a_out is not NULL;
 * act 0 = identity, 1 = SiLU z / (1 + exp(-z)). */
int mjk_bias_act(int m, int n, float *z, const float *bias, uint16_t *a_out, int act,
                 void *stream);
/* dz = bf16(da * act'(z)); dbias[n] = sum_m da act'(z) in fp32 when dbias is not NULL. */
int mjk_act_backward(int m, int n, const float *da, const float *z, uint16_t *dz, float *dbias,
                     int act, void *stream);
/* dy = scale * mask * (y - target) over n elements (mask may be NULL = 1); if loss is not
 * NULL, *loss += sum mask (y - target)^2 (the caller zeroes it). */
int mjk_mse_grad(size_t n, const float *y, const float *target, const float *mask, float scale,
                 float *dy, double *loss, void *stream);
/* One Adam step (step >= 1, bias-corrected) on n fp32 weights; w_bf16 (may be NULL)
 * receives the rounded copy the GEMMs read. */
int mjk_adam(size_t n, float *w, const float *g, float *m1, float *m2, uint16_t *w_bf16,
             float lr, float beta1, float beta2, float eps, int step, void *stream);
int mjk_to_bf16(size_t n, const float *x, uint16_t *y, void *stream);

#ifdef __cplusplus
}
#endif

#endif
