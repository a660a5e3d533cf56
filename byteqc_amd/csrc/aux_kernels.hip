// Small kernels around the J/K build: Schwarz factors, density block maxima,
// zero fill and the final symmetrisation.
#include "jk_launch.cuh"
#include "mjk_internal.h"

namespace mjk {

__host__ __device__ constexpr double fac_sp(int l)
{
    return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0);
}

// q[i*nbas+j] = sqrt(max_(a in i, b in j) |(ab|ab)|) for pairs of class (LA, LB).
template <int LA, int LB>
__global__ void qcond_kernel(ShellTab s, const int *pi, const int *pj, int npairs, double *q)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    constexpr int NR = LA + LB + 1;
    constexpr int NFA = ncart(LA), NFB = ncart(LB);
    constexpr int GS = (LA + 1) * (LB + 1) * (LA + 1) * (LB + 1);
    const int i = pi[t], j = pj[t];
    const double A[3] = {s.x[i], s.y[i], s.z[i]}, B[3] = {s.x[j], s.y[j], s.z[j]};
    const double AB[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    const double rr = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    const int npi = s.nprim[i], npj = s.nprim[j];
    const double *ei = s.exps + s.poff[i], *ej = s.exps + s.poff[j];
    const double *ci = s.coefs + s.poff[i], *cj = s.coefs + s.poff[j];
    double diag[NFA * NFB];
#pragma unroll
    for (int k = 0; k < NFA * NFB; ++k) diag[k] = 0.0;
    const double fsp = fac_sp(LA) * fac_sp(LB);
    for (int p1 = 0; p1 < npi * npj; ++p1) {
        const double ai = ei[p1 / npj], aj = ej[p1 % npj], aij = ai + aj;
        const double eij = fsp * ci[p1 / npj] * cj[p1 % npj] * exp(-rr * ai * aj / aij);
        const double P[3] = {(ai * A[0] + aj * B[0]) / aij, (ai * A[1] + aj * B[1]) / aij,
                             (ai * A[2] + aj * B[2]) / aij};
        for (int p2 = 0; p2 < npi * npj; ++p2) {
            const double ak = ei[p2 / npj], al = ej[p2 % npj], akl = ak + al;
            const double ekl = fsp * ci[p2 / npj] * cj[p2 % npj] * exp(-rr * ak * al / akl);
            const double Q[3] = {(ak * A[0] + al * B[0]) / akl, (ak * A[1] + al * B[1]) / akl,
                                 (ak * A[2] + al * B[2]) / akl};
            const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
            const double aijkl = aij + akl, a0 = aij * akl / aijkl;
            const double T = a0 * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = PI52X2 * eij * ekl / (aij * akl * sqrt(aijkl));
            for (int r = 0; r < NR; ++r) {
                double x, w;
                rys_root64<NR>(T, r, x, w);
                const double b00 = x * 0.5 / aijkl;
                const double b10 = (0.5 - b00 * akl) / aij, b01 = (0.5 - b00 * aij) / akl;
                double g[3][GS];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double c00 = (P[d] - A[d]) - 2.0 * b00 * akl * PQ[d];
                    const double c0p = (Q[d] - A[d]) + 2.0 * b00 * aij * PQ[d];
                    rys2d<double, LA, LB, LA, LB>(d == 2 ? w * pref : 1.0, c00, c0p, b00, b10, b01,
                                                  AB[d], AB[d], g[d]);
                }
#pragma unroll
                for (int ia = 0; ia < NFA; ++ia)
#pragma unroll
                    for (int jb = 0; jb < NFB; ++jb) {
                        const int ix = cart_x(LA, ia), iy = cart_y(LA, ia), iz = cart_z(LA, ia);
                        const int jx = cart_x(LB, jb), jy = cart_y(LB, jb), jz = cart_z(LB, jb);
                        const int ox = ((ix * (LB + 1) + jx) * (LA + 1) + ix) * (LB + 1) + jx;
                        const int oy = ((iy * (LB + 1) + jy) * (LA + 1) + iy) * (LB + 1) + jy;
                        const int oz = ((iz * (LB + 1) + jz) * (LA + 1) + iz) * (LB + 1) + jz;
                        diag[ia * NFB + jb] += g[0][ox] * g[1][oy] * g[2][oz];
                    }
            }
        }
    }
    double m = 0.0;
#pragma unroll
    for (int k = 0; k < NFA * NFB; ++k) m = fmax(m, fabs(diag[k]));
    const double v = sqrt(m);
    q[(size_t)i * s.nbas + j] = v;
    q[(size_t)j * s.nbas + i] = v;
}

hipError_t launch_qcond(int la, int lb, const ShellTab &s, const int *pi, const int *pj, int n,
                        double *q, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const dim3 block(64), grid((n + 63) / 64);
#define MJK_Q(A, B)                                                                   \
    if (la == A && lb == B) {                                                         \
        hipLaunchKernelGGL((qcond_kernel<A, B>), grid, block, 0, st, s, pi, pj, n, q); \
        return hipGetLastError();                                                     \
    }
    MJK_Q(0, 0) MJK_Q(1, 0) MJK_Q(1, 1) MJK_Q(2, 0) MJK_Q(2, 1) MJK_Q(2, 2)
    MJK_Q(3, 0) MJK_Q(3, 1) MJK_Q(3, 2) MJK_Q(3, 3)
#undef MJK_Q
    return hipErrorInvalidValue;
}

// dc[a*nbas+b] = max over densities and the (a, b) AO block of |D|; *dmax = global max.
__global__ void dm_cond_kernel(const double *dm, int n_dm, int nao, const int *ao_loc, int nbas,
                               double *dc, unsigned long long *dmax)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nbas * nbas) return;
    const int a = (int)(t / nbas), b = (int)(t % nbas);
    const int a0 = ao_loc[a], a1 = ao_loc[a + 1], b0 = ao_loc[b], b1 = ao_loc[b + 1];
    double m = 0.0;
    for (int n = 0; n < n_dm; ++n) {
        const double *D = dm + (size_t)n * nao * nao;
        for (int x = a0; x < a1; ++x)
            for (int y = b0; y < b1; ++y) m = fmax(m, fabs(D[(size_t)x * nao + y]));
    }
    dc[t] = m;
    atomicMax(dmax, (unsigned long long)__double_as_longlong(m));
}

__global__ void fill_kernel(double *p, size_t n, double v)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void fill_u64_kernel(unsigned long long *p, size_t n)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

// V <- f (V + V^T) for each of n_dm square matrices.
__global__ void symmetrize_kernel(double *v, int n_dm, int nao, double f)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nn = (size_t)nao * nao;
    if (t >= nn * n_dm) return;
    const int n = (int)(t / nn);
    const int a = (int)((t % nn) / nao), b = (int)(t % nao);
    if (b > a) return;
    double *V = v + n * nn;
    const double s = f * (V[(size_t)a * nao + b] + V[(size_t)b * nao + a]);
    V[(size_t)a * nao + b] = s;
    V[(size_t)b * nao + a] = s;
}

hipError_t launch_dm_cond(const double *dm, int n_dm, int nao, const int *ao_loc, int nbas,
                          double *dc, unsigned long long *dmax, hipStream_t st)
{
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, st, dmax, (size_t)1);
    const long n = (long)nbas * nbas;
    hipLaunchKernelGGL(dm_cond_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dm,
                       n_dm, nao, ao_loc, nbas, dc, dmax);
    return hipGetLastError();
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_symmetrize(double *v, int n_dm, int nao, double f, hipStream_t st)
{
    const size_t n = (size_t)nao * nao * n_dm;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, v,
                       n_dm, nao, f);
    return hipGetLastError();
}

// J = f (V + V^T) per molecule of a batch; blockIdx.y is the molecule
__global__ void symmetrize_batch_kernel(double *v, const MolRec *mols, double f)
{
    const MolRec m = mols[blockIdx.y];
    const size_t nn = (size_t)m.nao * m.nao;
    double *V = v + m.off;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x) {
        const int a = (int)(t / m.nao), b = (int)(t % m.nao);
        if (b > a) continue;
        const double s = f * (V[(size_t)a * m.nao + b] + V[(size_t)b * m.nao + a]);
        V[(size_t)a * m.nao + b] = s;
        V[(size_t)b * m.nao + a] = s;
    }
}

hipError_t launch_symmetrize_batch(double *v, const MolRec *mols, int n_mol, int max_nao, double f, hipStream_t st)
{
    if (n_mol <= 0) return hipSuccess;
    const size_t nn = (size_t)max_nao * max_nao;
    unsigned gx = (unsigned)((nn + 255) / 256);
    if (gx > 1024) gx = 1024;
    if (gx == 0) gx = 1;
    hipLaunchKernelGGL(symmetrize_batch_kernel, dim3(gx, (unsigned)n_mol), dim3(256), 0, st, v, mols, f);
    return hipGetLastError();
}

int class_kernel_info(int la, int lb, int lc, int ld, int *bra_per_tile)
{
#define MJK_INFO(A, B, C, D)                                                                     \
    if constexpr (canon(A, B, C, D)) {                                                           \
        if (la == A && lb == B && lc == C && ld == D) {                                          \
            if constexpr (Cfg<A, B, C, D>::REG) {                                                \
                *bra_per_tile = REG_BRA;                                                         \
                return 0;                                                                        \
            } else if constexpr (GCfg<A, B, C, D>::OK) {                                         \
                *bra_per_tile = GCfg<A, B, C, D>::W;                                             \
                return 1;                                                                        \
            } else {                                                                             \
                *bra_per_tile = Cfg<A, B, C, D>::BP;                                             \
                return 2;                                                                        \
            }                                                                                    \
        }                                                                                        \
    }
#define MJK_INFO_D(A, B, C) MJK_INFO(A, B, C, 0) MJK_INFO(A, B, C, 1) MJK_INFO(A, B, C, 2) MJK_INFO(A, B, C, 3)
#define MJK_INFO_C(A, B) MJK_INFO_D(A, B, 0) MJK_INFO_D(A, B, 1) MJK_INFO_D(A, B, 2) MJK_INFO_D(A, B, 3)
    MJK_INFO_C(0, 0) MJK_INFO_C(1, 0) MJK_INFO_C(1, 1) MJK_INFO_C(2, 0) MJK_INFO_C(2, 1)
    MJK_INFO_C(2, 2) MJK_INFO_C(3, 0) MJK_INFO_C(3, 1) MJK_INFO_C(3, 2) MJK_INFO_C(3, 3)
#undef MJK_INFO_C
#undef MJK_INFO_D
#undef MJK_INFO
    return -1;
}

hipError_t launch_class(int la, int lb, int lc, int ld, bool fp32, const JKArgs &a, hipStream_t st,
                        long *nt)
{
#define MJK_LAB(A, B) \
    if (la == A && lb == B) return launch_lab<A, B>(lc, ld, fp32, a, st, nt);
    MJK_LAB(0, 0) MJK_LAB(1, 0) MJK_LAB(1, 1) MJK_LAB(2, 0) MJK_LAB(2, 1)
    MJK_LAB(2, 2) MJK_LAB(3, 0) MJK_LAB(3, 1) MJK_LAB(3, 2) MJK_LAB(3, 3)
#undef MJK_LAB
    return hipErrorInvalidValue;
}

}  // namespace mjk
