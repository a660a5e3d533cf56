// Launch wrapper and per-class dispatch for the J/K kernels.
#pragma once
#include "jk_device.cuh"

namespace mjk {

// Canonical orientation: l_i >= l_j, l_k >= l_l and the bra pair has at least as many
// Cartesian products as the ket pair (every thread owns whole bra rows).
__host__ __device__ constexpr bool canon(int la, int lb, int lc, int ld)
{
    return la >= lb && lc >= ld && ncart(la) * ncart(lb) >= ncart(lc) * ncart(ld);
}

constexpr long MAX_TILES_PER_LAUNCH = 1L << 23;

template <typename R, int LA, int LB, int LC, int LD>
hipError_t launch(JKArgs a, hipStream_t st, long *ntiles_out)
{
    using C = Cfg<LA, LB, LC, LD>;
    a.nb_bra = (a.bra.npairs + C::BP - 1) / C::BP;
    a.nb_ket = (a.ket.npairs + C::KP - 1) / C::KP;
    long ntiles;
    if (!a.same) ntiles = (long)a.nb_bra * a.nb_ket;
    else if (C::BP == C::KP) ntiles = (long)a.nb_bra * (a.nb_bra + 1) / 2;
    else ntiles = tri_tiles_bp1(a.bra.npairs, C::KP);
    if (ntiles_out) *ntiles_out += ntiles;
    for (long base = 0; base < ntiles; base += MAX_TILES_PER_LAUNCH) {
        const long n = ntiles - base < MAX_TILES_PER_LAUNCH ? ntiles - base : MAX_TILES_PER_LAUNCH;
        a.tile_base = base;
        hipLaunchKernelGGL((jk_kernel<R, LA, LB, LC, LD>), dim3((unsigned)n), dim3(C::NT), 0, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int LA, int LB>
hipError_t launch_lab(int lc, int ld, bool fp32, const JKArgs &a, hipStream_t st, long *nt)
{
#define MJK_TRY(LC, LD)                                                                        \
    if constexpr (canon(LA, LB, LC, LD)) {                                                     \
        if (lc == LC && ld == LD)                                                              \
            return fp32 ? launch<float, LA, LB, LC, LD>(a, st, nt)                             \
                        : launch<double, LA, LB, LC, LD>(a, st, nt);                           \
    }
#define MJK_TRY_D(LC) MJK_TRY(LC, 0) MJK_TRY(LC, 1) MJK_TRY(LC, 2) MJK_TRY(LC, 3)
    MJK_TRY_D(0) MJK_TRY_D(1) MJK_TRY_D(2) MJK_TRY_D(3)
#undef MJK_TRY_D
#undef MJK_TRY
    return hipErrorInvalidValue;
}

#define MJK_LAB_DECL(A, B) \
    extern template hipError_t launch_lab<A, B>(int, int, bool, const JKArgs &, hipStream_t, long *);
MJK_LAB_DECL(0, 0) MJK_LAB_DECL(1, 0) MJK_LAB_DECL(1, 1) MJK_LAB_DECL(2, 0) MJK_LAB_DECL(2, 1)
MJK_LAB_DECL(2, 2) MJK_LAB_DECL(3, 0) MJK_LAB_DECL(3, 1) MJK_LAB_DECL(3, 2) MJK_LAB_DECL(3, 3)
#undef MJK_LAB_DECL

hipError_t launch_class(int la, int lb, int lc, int ld, bool fp32, const JKArgs &a, hipStream_t st, long *nt);

}  // namespace mjk
