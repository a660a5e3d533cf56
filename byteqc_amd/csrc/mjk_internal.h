// Internal declarations shared by the host dispatcher and the helper kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include "jk_device.cuh"

namespace mjk {

// Shells of a Cartesian, segmented basis on the device.
struct ShellTab {
    const double *x, *y, *z;     // centre of each shell
    const int *nprim, *poff;     // primitive count and offset into exps / coefs
    const double *exps, *coefs;  // coefficients carry the radial normalisation
    int nbas;
};

hipError_t launch_qcond(int la, int lb, const ShellTab &s, const int *pi, const int *pj, int n,
                        double *q, hipStream_t st);
hipError_t launch_dm_cond(const double *dm, int n_dm, int nao, const int *ao_loc, int nbas,
                          double *dc, unsigned long long *dmax, hipStream_t st);
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long *p, size_t n, hipStream_t st);
hipError_t launch_symmetrize(double *v, int n_dm, int nao, double f, hipStream_t st);
// Records msg as the calling thread's mjk_last_error() text and returns code.
int report_error(int code, const std::string &msg);

hipError_t launch_symmetrize_batch(double *v, const MolRec *mols, int n_mol, int max_nao, double f, hipStream_t st);
// Kernel family of a canonical class (0 register, 1 group, 2 block-LDS; -1 unknown) and
// the bra pairs one of its tiles covers.
int class_kernel_info(int la, int lb, int lc, int ld, int *bra_per_tile);

// One class pair of an ERI-tensor fill (fill_int2e.hip).
struct FillArgs {
    PairSet bra, ket;
    double *eri;
    long sb0, sb1, sk0, sk1;  // element strides of the bra's i, j and the ket's k, l
    int ob0, ob1, ok0, ok1;   // AO offsets subtracted from them
    long base;                // first quartet of this launch
};
hipError_t launch_fill_class(int la, int lb, int lc, int ld, FillArgs f, hipStream_t st);

hipError_t launch_class(int la, int lb, int lc, int ld, bool fp32, const JKArgs &a, hipStream_t st, long *nt);

}  // namespace mjk
