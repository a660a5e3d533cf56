// Launch wrapper and per-class dispatch for the J/K kernels.
#pragma once
#include "jk_device.cuh"
#include "jk_reg.cuh"
#include "jk_grp.cuh"

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
    long ntiles;
    // ket chunk: long chunks amortise the per-lane bra state, short ones give small classes
    // enough workgroups (about 2048 or more per launch)
    auto pick_kc = [&](long nb_bra, int kmax) {
        int kc = a.kc_cap > 0 && a.kc_cap < kmax ? a.kc_cap : kmax;
        while (kc > 1 && nb_bra * ((a.ket.npairs + kc - 1) / kc) < 2048) kc >>= 1;
        return kc;
    };
    if (a.tiles) {
        // batched build: the host enumerated the tiles (register and group kernels only)
        if constexpr (!(C::REG || GCfg<LA, LB, LC, LD>::OK)) return hipErrorNotSupported;
        ntiles = a.ndesc;
    } else if constexpr (C::REG) {
        a.nb_bra = (a.bra.npairs + REG_BRA - 1) / REG_BRA;
        a.kc = pick_kc(a.nb_bra, REG_KC);
        a.nb_ket = (a.ket.npairs + a.kc - 1) / a.kc;
        ntiles = a.same ? reg_tri_tiles(a.nb_bra, a.nb_ket, a.kc) : (long)a.nb_bra * a.nb_ket;
    } else if constexpr (GCfg<LA, LB, LC, LD>::OK) {
        constexpr int W = GCfg<LA, LB, LC, LD>::W;
        a.nb_bra = (a.bra.npairs + W - 1) / W;
        a.kc = pick_kc(a.nb_bra, GRP_KC);
        a.nb_ket = (a.ket.npairs + a.kc - 1) / a.kc;
        ntiles = (long)a.nb_bra * a.nb_ket;  // same-class tiles above the triangle exit at once
    } else {
        a.nb_bra = (a.bra.npairs + C::BP - 1) / C::BP;
        a.nb_ket = (a.ket.npairs + C::KP - 1) / C::KP;
        if (!a.same) ntiles = (long)a.nb_bra * a.nb_ket;
        else if (C::BP == C::KP) ntiles = (long)a.nb_bra * (a.nb_bra + 1) / 2;
        else ntiles = tri_tiles_bp1(a.bra.npairs, C::KP);
    }
    if (ntiles_out) *ntiles_out += ntiles;
    // the register and group kernels digest one density per launch
    constexpr bool ONE_DM = C::REG || GCfg<LA, LB, LC, LD>::OK;
    const int ndm_launch = ONE_DM ? a.n_dm : 1;
    const size_t nn = (size_t)a.nao * a.nao;
    for (int n = 0; n < ndm_launch; ++n) {
        JKArgs b = a;
        if (ONE_DM) {
            b.n_dm = 1;
            b.dm = a.dm + n * nn;
            b.vj = a.vj ? a.vj + n * nn : nullptr;
            b.vk = a.vk ? a.vk + n * nn : nullptr;
            if (n > 0) b.counter = nullptr;
        }
        for (long base = 0; base < ntiles; base += MAX_TILES_PER_LAUNCH) {
            const long cnt = ntiles - base < MAX_TILES_PER_LAUNCH ? ntiles - base : MAX_TILES_PER_LAUNCH;
            b.tile_base = base;
            if constexpr (C::REG)
                hipLaunchKernelGGL((jk_reg_kernel<R, LA, LB, LC, LD>), dim3((unsigned)cnt), dim3(REG_BRA), 0, st, b);
            else if constexpr (GCfg<LA, LB, LC, LD>::OK)
                hipLaunchKernelGGL((jk_grp_kernel<R, LA, LB, LC, LD>), dim3((unsigned)cnt), dim3(64), 0, st, b);
            else
                hipLaunchKernelGGL((jk_kernel<R, LA, LB, LC, LD>), dim3((unsigned)cnt), dim3(C::NT), 0, st, b);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
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
