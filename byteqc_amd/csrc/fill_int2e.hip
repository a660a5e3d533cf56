// Explicit ERI blocks: the restatement of GINTfill_int2e (cuobc/lib/gint/nr_fill_ao_ints.cu:
// 493-551, kernels in cuobc/lib/gint/fill_ints.cu:28-198).  Every quartet of a bra and a ket
// pair class adds its Cartesian block (ij|kl) into a strided tensor.  One 256-thread
// workgroup per quartet: three threads build the x, y and z 2D tables of one root in LDS
// (the same recursion as the J/K kernels), then every thread accumulates its share of the
// block, whose three table offsets it packed into one integer at the start.  Each quartet's
// block belongs to one workgroup, so the tensor is updated with plain read-add-writes.
// A setup and test path (ERI tensors of small systems), not a hot path.
#include <stdint.h>

#include "../../include/mi355jk.h"
#include "jk_launch.cuh"
#include "mjk_internal.h"

namespace mjk {

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) fill_kernel(const FillArgs f)
{
    constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    constexpr int NFA = ncart(LA), NFB = ncart(LB), NFC = ncart(LC), NFD = ncart(LD);
    constexpr int NFAB = NFA * NFB, NFCD = NFC * NFD, NE = NFAB * NFCD;
    constexpr int G2 = (LC + 1) * (LD + 1);
    constexpr int GS = (LA + 1) * (LB + 1) * G2;
    constexpr int PER = (NE + 255) / 256;
    static_assert(GS <= 256, "table offsets are packed in 8 bits");
    __shared__ double g[3][GS];

    const long quart = f.base + (long)blockIdx.x;
    const int p = (int)(quart / f.ket.npairs), q = (int)(quart % f.ket.npairs);
    const int t = threadIdx.x;

    // packed (x, y, z) table offsets of this thread's elements
    int offs[PER];
    double acc[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        acc[k] = 0.0;
        const int e = t + 256 * k;
        offs[k] = 0;
        if (e < NE) {
            const int row = e / NFCD, c = e % NFCD;
            const int ia = row / NFB, jb = row % NFB, kc = c / NFD, ld = c % NFD;
            const int ox = (cart_x(LA, ia) * (LB + 1) + cart_x(LB, jb)) * G2 + cart_x(LC, kc) * (LD + 1) + cart_x(LD, ld);
            const int oy = (cart_y(LA, ia) * (LB + 1) + cart_y(LB, jb)) * G2 + cart_y(LC, kc) * (LD + 1) + cart_y(LD, ld);
            const int oz = (cart_z(LA, ia) * (LB + 1) + cart_z(LB, jb)) * G2 + cart_z(LC, kc) * (LD + 1) + cart_z(LD, ld);
            offs[k] = ox | (oy << 8) | (oz << 16);
        }
    }
    const double *bd = f.bra.pdbl + 8 * (size_t)p;
    const double *kd = f.ket.pdbl + 8 * (size_t)q;
    const double A[3] = {bd[1], bd[2], bd[3]}, AB[3] = {bd[4], bd[5], bd[6]};
    const double Cc[3] = {kd[1], kd[2], kd[3]}, CD[3] = {kd[4], kd[5], kd[6]};
    const int npb = f.bra.nprim, npk = f.ket.nprim;
    for (int ip = 0; ip < npb; ++ip) {
        const double *bp_ = f.bra.prim + 6 * ((size_t)p * npb + ip);
        const double aij = bp_[0], eij = bp_[1], iaij = bp_[5];
        const double P[3] = {bp_[2], bp_[3], bp_[4]};
        for (int kq = 0; kq < npk; ++kq) {
            const double *kp_ = f.ket.prim + 6 * ((size_t)q * npk + kq);
            const double akl = kp_[0], ekl = kp_[1], iakl = kp_[5];
            const double Q[3] = {kp_[2], kp_[3], kp_[4]};
            const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
            const double iaijkl = 1.0 / (aij + akl);
            const double T = aij * akl * iaijkl * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = PI52X2 * eij * ekl * iaij * iakl * sqrt(iaijkl);
            for (int r = 0; r < NR; ++r) {
                double x, w;
                rys_root<double, NR>(T, r, x, w);
                if (t < 3) {
                    const double b00 = x * 0.5 * iaijkl;
                    const double b10 = (0.5 - b00 * akl) * iaij;
                    const double b01 = (0.5 - b00 * aij) * iakl;
                    const double tb = 2.0 * b00;
                    const double c00 = (P[t] - A[t]) - tb * akl * PQ[t];
                    const double c0p = (Q[t] - Cc[t]) + tb * aij * PQ[t];
                    rys2d<double, LA, LB, LC, LD>(t == 2 ? w * pref : 1.0, c00, c0p, b00, b10, b01, AB[t], CD[t],
                                                  g[t]);
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < PER; ++k)
                    if (t + 256 * k < NE) {
                        const int o = offs[k];
                        acc[k] += g[0][o & 255] * g[1][(o >> 8) & 255] * g[2][(o >> 16) & 255];
                    }
                __syncthreads();
            }
        }
    }
    const int4 b4 = f.bra.pint[p], k4 = f.ket.pint[q];
    const long i0 = b4.z - f.ob0, j0 = b4.w - f.ob1, k0 = k4.z - f.ok0, l0 = k4.w - f.ok1;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = t + 256 * k;
        if (e < NE) {
            const int row = e / NFCD, c = e % NFCD;
            const int ia = row / NFB, jb = row % NFB, kc = c / NFD, ld = c % NFD;
            f.eri[(i0 + ia) * f.sb0 + (j0 + jb) * f.sb1 + (k0 + kc) * f.sk0 + (l0 + ld) * f.sk1] += acc[k];
        }
    }
}

constexpr long FILL_MAX_BLOCKS = 1L << 30;

hipError_t launch_fill_class(int la, int lb, int lc, int ld, FillArgs f, hipStream_t st)
{
    const long nq = (long)f.bra.npairs * f.ket.npairs;
#define MJK_FILL(A, B, C, D)                                                                      \
    if constexpr (canon(A, B, C, D)) {                                                            \
        if (la == A && lb == B && lc == C && ld == D) {                                           \
            for (long base = 0; base < nq; base += FILL_MAX_BLOCKS) {                             \
                f.base = base;                                                                    \
                const long cnt = nq - base < FILL_MAX_BLOCKS ? nq - base : FILL_MAX_BLOCKS;       \
                hipLaunchKernelGGL((fill_kernel<A, B, C, D>), dim3((unsigned)cnt), dim3(256), 0, st, f); \
                hipError_t e = hipGetLastError();                                                 \
                if (e != hipSuccess) return e;                                                    \
            }                                                                                     \
            return hipSuccess;                                                                    \
        }                                                                                         \
    }
#define MJK_FILL_D(A, B, C) MJK_FILL(A, B, C, 0) MJK_FILL(A, B, C, 1) MJK_FILL(A, B, C, 2) MJK_FILL(A, B, C, 3)
#define MJK_FILL_C(A, B) MJK_FILL_D(A, B, 0) MJK_FILL_D(A, B, 1) MJK_FILL_D(A, B, 2) MJK_FILL_D(A, B, 3)
    MJK_FILL_C(0, 0) MJK_FILL_C(1, 0) MJK_FILL_C(1, 1) MJK_FILL_C(2, 0) MJK_FILL_C(2, 1)
    MJK_FILL_C(2, 2) MJK_FILL_C(3, 0) MJK_FILL_C(3, 1) MJK_FILL_C(3, 2) MJK_FILL_C(3, 3)
#undef MJK_FILL_C
#undef MJK_FILL_D
#undef MJK_FILL
    return hipErrorInvalidValue;
}

}  // namespace mjk
