// Register-resident J/K kernel for the small angular-momentum classes.
//
// One lane owns one bra shell pair for the whole launch tile; all lanes of the
// workgroup walk the same sequence of ket pairs (a chunk of the ket class), so every
// ket quantity is uniform across the wave:
//   * the quartet's Cartesian integrals live in registers (2D tables of one root at a
//     time, contracted over primitives and roots);
//   * J_ij accumulates per lane over the whole chunk and is flushed once;
//   * J_kl is uniform per ket: a wave-wide sum, then one atomic per element per wave;
//   * K_ik and K_jk accumulate per lane while consecutive kets share shell k (ket pairs
//     are ordered by (k, l) inside a Schwarz decade) and are flushed when k changes;
//   * K_il is summed over the run of lanes whose bra pairs share shell i, then written;
//   * K_jl is written per quartet.
// Every D element a quartet needs is loaded at the top of its iteration, so the load
// latency hides behind the integral evaluation.
// The six updates are those of cuobc/lib/gvhf/contract_jk.cu:219-316.
#pragma once
#include "jk_device.cuh"

namespace mjk {

constexpr int REG_BRA = 256;  // bra pairs (lanes) per tile
// Workgroups per CU each class is compiled for.  Two halve the registers a lane may hold:
// faster where the extra wave hides more latency than the spills cost, measured per class
// on (H2O)64/def2-TZVP ((fp|ss) 92 -> 70 ms, (dp|ss) 68 -> 60, while
// (dp|ps) went 80 -> 112 and stays at one).
constexpr int reg_min_blocks(int la, int lb, int lc, int ld)
{
    const int key = ((la * 4 + lb) * 4 + lc) * 4 + ld;
    switch (key) {
    case ((3 * 4 + 1) * 4 + 0) * 4 + 0:  // (fp|ss)
    case ((3 * 4 + 0) * 4 + 1) * 4 + 0:  // (fs|ps)
    case ((2 * 4 + 1) * 4 + 0) * 4 + 0:  // (dp|ss)
    case ((2 * 4 + 0) * 4 + 1) * 4 + 0:  // (ds|ps)
    case ((1 * 4 + 1) * 4 + 1) * 4 + 0:  // (pp|ps)
        return 2;
    default:
        return 1;
    }
}
constexpr int REG_KC = 128;   // longest ket chunk per tile

// Tiles of a same-class launch with ket chunks of kc pairs (kc divides REG_BRA): bra block
// b needs ket chunks 0 .. min(nb_ket, (b+1) r) - 1, r = REG_BRA / kc.  Returns the number of
// tiles before bra block bi.
__host__ __device__ inline long reg_tri_tiles(long bi, long nb_ket, long kc)
{
    const long r = REG_BRA / kc;
    const long B = nb_ket / r;
    const long m = bi < B ? bi : B;
    return r * m * (m + 1) / 2 + (bi > B ? (bi - B) * nb_ket : 0);
}

// Inclusive segmented sum over lanes; seg_head is the first lane of this lane's segment
// (segments are runs of consecutive lanes).
__device__ inline double seg_scan(double v, int lane, int seg_head)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double u = __shfl_up(v, off, 64);
        if (lane - off >= seg_head) v += u;
    }
    return v;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename R, int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(REG_BRA, reg_min_blocks(LA, LB, LC, LD))
jk_reg_kernel(const JKArgs a)
{
    constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    constexpr int NFA = ncart(LA), NFB = ncart(LB), NFC = ncart(LC), NFD = ncart(LD);
    constexpr int NFAB = NFA * NFB, NFCD = NFC * NFD;
    constexpr int G2 = (LC + 1) * (LD + 1);
    constexpr int GS = (LA + 1) * (LB + 1) * G2;
    constexpr bool D_EARLY = (NFA + NFB) * (NFC + NFD) + NFCD <= 16;

    // ---- tile
    const long tile = a.tile_base + (long)blockIdx.x;
    const int t = threadIdx.x;
    const int lane = t & 63;
    int p, q_begin, q_end, p_first = 0;
    bool pvalid;
    const double *__restrict__ D = a.dm;
    double *VJ = a.vj, *VK = a.vk;
    int nao = a.nao, ao_base = 0;
    if (a.tiles) {
        const int4 d = a.tiles[tile];
        p_first = d.x;
        p = d.x + t;
        pvalid = p < d.y;
        q_begin = d.z;
        q_end = d.w;
        const MolRec m = a.mols[a.tile_mol[tile]];
        D += m.off;
        VJ = VJ ? VJ + m.off : nullptr;
        VK = VK ? VK + m.off : nullptr;
        nao = m.nao;
        ao_base = m.ao_base;
    } else {
        int bi, ki;
        if (!a.same) {
            bi = (int)(tile / a.nb_ket);
            ki = (int)(tile % a.nb_ket);
        } else {
            long lo = 0, hi = a.nb_bra;
            while (hi - lo > 1) {
                const long mid = (lo + hi) / 2;
                if (reg_tri_tiles(mid, a.nb_ket, a.kc) <= tile) lo = mid; else hi = mid;
            }
            bi = (int)lo;
            ki = (int)(tile - reg_tri_tiles(lo, a.nb_ket, a.kc));
        }
        p = bi * REG_BRA + t;
        pvalid = p < a.bra.npairs;
        q_begin = ki * a.kc;
        q_end = q_begin + a.kc;
        if (q_end > a.ket.npairs) q_end = a.ket.npairs;
    }

    // lanes without a bra pair borrow the tile's first one, which lies in the same molecule,
    // so every index they form stays inside this molecule's D block
    const int pp = pvalid ? p : p_first;
    const int4 b4 = a.bra.pint[pp];
    const double *bd = a.bra.pdbl + 8 * pp;
    const int ish = b4.x, jsh = b4.y;
    const double qp = pvalid ? bd[0] : 0.0;
    const double dmax = a.dm_max ? *a.dm_max : 1.0;
    // cheap tile test: the largest bound any quartet of the tile can reach
    {
        double qk = 0.0;
        for (int q = q_begin + lane; q < q_end; q += 64) qk = fmax(qk, a.ket.pdbl[8 * q]);
        double qb = qp;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            qk = fmax(qk, __shfl_xor(qk, off, 64));
            qb = fmax(qb, __shfl_xor(qb, off, 64));
        }
        // the precision split is per tile: this launch takes the tiles whose largest bound m
        // satisfies lo <= m < hi, and computes every quartet of such a tile
        const double m = qb * qk * dmax;
        if (!__syncthreads_or(m > a.tol)) return;
        if (!__syncthreads_or(m >= a.lo)) return;
        if (__syncthreads_or(m >= a.hi)) return;
    }
    // lanes holding bra pairs with the same shell i form runs: K_il is summed over a run
    int seg_head, seg_tail;
    {
        const int key = pvalid ? ish : -1 - lane;
        const int prev = __shfl_up(key, 1, 64), next = __shfl_down(key, 1, 64);
        const bool head = lane == 0 || prev != key;
        const bool tail = lane == 63 || next != key;
        const unsigned long long hm = __builtin_amdgcn_ballot_w64(head);
        const unsigned long long below = hm & ((lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1));
        seg_head = 63 - __builtin_clzll(below);
        seg_tail = tail;
    }
    // lanes of a wave holding the same shell j share one LDS slot for K_jl (the slot of the
    // first such lane, which alone issues the global atomics): one atomic per distinct j
    // instead of one per lane
    constexpr int NJL = NFB * NFD;
    __shared__ double kjl_slot[REG_BRA / 64][64][NJL];
    const int wv = t >> 6;
    int jlead = lane;
    {
        const int key = pvalid ? jsh : -1 - lane;
        for (int u = 63; u >= 0; --u)
            if (__shfl(key, u, 64) == key) jlead = u;
#pragma unroll
        for (int e = 0; e < NJL; ++e) kjl_slot[wv][lane][e] = 0.0;
    }
    const double A[3] = {bd[1], bd[2], bd[3]};
    const double AB[3] = {bd[4], bd[5], bd[6]};
    const int i0 = b4.z - ao_base, j0 = b4.w - ao_base;
    const bool doj = VJ != nullptr, dok = VK != nullptr;
    const int ri = i0 * nao, rj = j0 * nao;  // row offsets (nao^2 < 2^31)

    double dij[NFAB];
#pragma unroll
    for (int r = 0; r < NFAB; ++r) dij[r] = (doj && pvalid) ? D[ri + (r / NFB) * nao + j0 + r % NFB] : 0.0;
    double jij[NFAB];
    double kik[NFA][NFC], kjk[NFB][NFC];
#pragma unroll
    for (int r = 0; r < NFAB; ++r) jij[r] = 0.0;
#pragma unroll
    for (int x = 0; x < NFA; ++x)
#pragma unroll
        for (int c = 0; c < NFC; ++c) kik[x][c] = 0.0;
#pragma unroll
    for (int x = 0; x < NFB; ++x)
#pragma unroll
        for (int c = 0; c < NFC; ++c) kjk[x][c] = 0.0;
    int cur_k = -1, cur_k0 = 0;
    bool kdirty = false;
    unsigned long long nquart = 0, nslots = 0;
    const int npb = a.bra.nprim, npk = a.ket.nprim;

    // the next ket's records are loaded one iteration ahead (software pipelining), so their
    // latency overlaps the current quartet
    int4 k4n = make_int4(0, 0, 0, 0);
    double kdn[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (q_begin < q_end) {
        k4n = a.ket.pint[q_begin];
#pragma unroll
        for (int e = 0; e < 7; ++e) kdn[e] = a.ket.pdbl[8 * q_begin + e];
    }
    for (int q = q_begin; q < q_end; ++q) {
        // ---- ket q: all uniform
        const int4 k4 = k4n;
        double kd[7];
#pragma unroll
        for (int e = 0; e < 7; ++e) kd[e] = kdn[e];
        if (q + 1 < q_end) {
            k4n = a.ket.pint[q + 1];
#pragma unroll
            for (int e = 0; e < 7; ++e) kdn[e] = a.ket.pdbl[8 * (q + 1) + e];
        }
        const int ksh = __builtin_amdgcn_readfirstlane(k4.x);
        const int lsh = __builtin_amdgcn_readfirstlane(k4.y);
        const int k0 = __builtin_amdgcn_readfirstlane(k4.z) - ao_base;
        const int l0 = __builtin_amdgcn_readfirstlane(k4.w) - ao_base;
        const double qq = kd[0];
        const double Cc[3] = {kd[1], kd[2], kd[3]};
        const double CD[3] = {kd[4], kd[5], kd[6]};
        // ---- D elements of this quartet: issued before the integrals when they are few,
        // else next to the digestion (the 2D tables are dead there)
        double dkl[NFCD], dik[NFA][NFC], djk[NFB][NFC], dil[NFA][NFD], djl[NFB][NFD];
        auto load_d = [&]() {
#pragma unroll
            for (int c = 0; c < NFCD; ++c) dkl[c] = doj ? D[(k0 + c / NFD) * nao + l0 + c % NFD] : 0.0;
#pragma unroll
            for (int x = 0; x < NFA; ++x) {
#pragma unroll
                for (int c = 0; c < NFC; ++c) dik[x][c] = (dok && pvalid) ? D[ri + x * nao + k0 + c] : 0.0;
#pragma unroll
                for (int d = 0; d < NFD; ++d) dil[x][d] = (dok && pvalid) ? D[ri + x * nao + l0 + d] : 0.0;
            }
#pragma unroll
            for (int x = 0; x < NFB; ++x) {
#pragma unroll
                for (int c = 0; c < NFC; ++c) djk[x][c] = (dok && pvalid) ? D[rj + x * nao + k0 + c] : 0.0;
#pragma unroll
                for (int d = 0; d < NFD; ++d) djl[x][d] = (dok && pvalid) ? D[rj + x * nao + l0 + d] : 0.0;
            }
        };
        if constexpr (D_EARLY) load_d();
        if (ksh != cur_k) {
            if (dok && kdirty) {
#pragma unroll
                for (int x = 0; x < NFA; ++x)
#pragma unroll
                    for (int c = 0; c < NFC; ++c) {
                        unsafeAtomicAdd(VK + ri + x * nao + cur_k0 + c, kik[x][c]);
                        kik[x][c] = 0.0;
                    }
#pragma unroll
                for (int x = 0; x < NFB; ++x)
#pragma unroll
                    for (int c = 0; c < NFC; ++c) {
                        unsafeAtomicAdd(VK + rj + x * nao + cur_k0 + c, kjk[x][c]);
                        kjk[x][c] = 0.0;
                    }
            }
            kdirty = false;
            cur_k = ksh;
            cur_k0 = k0;
        }
        bool active = pvalid && (!a.same || q <= p);
        if (active) {
            double bound = qp * qq;
            if (a.dm_cond) {
                const double *dc = a.dm_cond;
                const int nb = a.nbas;
                const double d = fmax(fmax(dc[ish * nb + jsh], dc[ksh * nb + lsh]),
                                      fmax(fmax(dc[ish * nb + ksh], dc[ish * nb + lsh]),
                                           fmax(dc[jsh * nb + ksh], dc[jsh * nb + lsh])));
                bound *= d;
            }
            active = bound > a.tol;
        }
        if (!__builtin_amdgcn_ballot_w64(active)) continue;  // wave-uniform skip
        ++nslots;
        R acc[NFAB][NFCD];
#pragma unroll
        for (int r = 0; r < NFAB; ++r)
#pragma unroll
            for (int c = 0; c < NFCD; ++c) acc[r][c] = (R)0;
        if (active) {
            ++nquart;
            const double norm = (a.same && p == q) ? 0.5 * PI52X2 : PI52X2;
            for (int ip = 0; ip < npb; ++ip) {
                const double *bp_ = a.bra.prim + 6 * (p * npb + ip);
                const double aij = bp_[0], eij = bp_[1], iaij = bp_[5];
                const double P[3] = {bp_[2], bp_[3], bp_[4]};
                for (int kq = 0; kq < npk; ++kq) {
                    const double *kp_ = a.ket.prim + 6 * (q * npk + kq);
                    const double akl = kp_[0], ekl = kp_[1], iakl = kp_[5];
                    const double Q[3] = {kp_[2], kp_[3], kp_[4]};
                    const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                    const double iaijkl = 1.0 / (aij + akl);
                    const double T = aij * akl * iaijkl * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                    const double pref = norm * eij * ekl * iaij * iakl * sqrt(iaijkl);
                    if (fabs(pref) < a.prim_tol) continue;
                    const double inv2 = 0.5 * iaijkl;
#pragma unroll 1
                    for (int r = 0; r < NR; ++r) {
                        R x, w;
                        rys_root<R, NR>(T, r, x, w);
                        const R b00 = x * (R)inv2;
                        const R b10 = ((R)0.5 - b00 * (R)akl) * (R)iaij;
                        const R b01 = ((R)0.5 - b00 * (R)aij) * (R)iakl;
                        const R tb = (R)2 * b00;
                        R gx[GS], gy[GS], gz[GS];
                        rys2d<R, LA, LB, LC, LD>((R)1, (R)(P[0] - A[0]) - tb * (R)akl * (R)PQ[0],
                                                 (R)(Q[0] - Cc[0]) + tb * (R)aij * (R)PQ[0],
                                                 b00, b10, b01, (R)AB[0], (R)CD[0], gx);
                        rys2d<R, LA, LB, LC, LD>((R)1, (R)(P[1] - A[1]) - tb * (R)akl * (R)PQ[1],
                                                 (R)(Q[1] - Cc[1]) + tb * (R)aij * (R)PQ[1],
                                                 b00, b10, b01, (R)AB[1], (R)CD[1], gy);
                        rys2d<R, LA, LB, LC, LD>(w * (R)pref, (R)(P[2] - A[2]) - tb * (R)akl * (R)PQ[2],
                                                 (R)(Q[2] - Cc[2]) + tb * (R)aij * (R)PQ[2],
                                                 b00, b10, b01, (R)AB[2], (R)CD[2], gz);
#pragma unroll
                        for (int row = 0; row < NFAB; ++row) {
                            const int ia = row / NFB, jb = row % NFB;
                            const int ox = (cart_x(LA, ia) * (LB + 1) + cart_x(LB, jb)) * G2;
                            const int oy = (cart_y(LA, ia) * (LB + 1) + cart_y(LB, jb)) * G2;
                            const int oz = (cart_z(LA, ia) * (LB + 1) + cart_z(LB, jb)) * G2;
#pragma unroll
                            for (int c = 0; c < NFCD; ++c) {
                                const int kc = c / NFD, ld = c % NFD;
                                const int cx = cart_x(LC, kc) * (LD + 1) + cart_x(LD, ld);
                                const int cy = cart_y(LC, kc) * (LD + 1) + cart_y(LD, ld);
                                const int cz = cart_z(LC, kc) * (LD + 1) + cart_z(LD, ld);
                                acc[row][c] += gx[ox + cx] * gy[oy + cy] * gz[oz + cz];
                            }
                        }
                    }
                }
            }
        }
        if constexpr (!D_EARLY) load_d();
        // ---- digestion (inactive lanes hold zeros)
        if (doj) {
            double jkl[NFCD];
#pragma unroll
            for (int c = 0; c < NFCD; ++c) jkl[c] = 0.0;
#pragma unroll
            for (int r = 0; r < NFAB; ++r)
#pragma unroll
                for (int c = 0; c < NFCD; ++c) {
                    jij[r] += (double)acc[r][c] * dkl[c];
                    jkl[c] += (double)acc[r][c] * dij[r];
                }
#pragma unroll
            for (int c = 0; c < NFCD; ++c) {
                const double s = wave_sum(jkl[c]);
                if (lane == 0) unsafeAtomicAdd(VJ + (k0 + c / NFD) * nao + l0 + c % NFD, s);
            }
        }
        if (dok) {
            kdirty |= active;
            double kil[NFA][NFD], kjl[NFB][NFD];
#pragma unroll
            for (int x = 0; x < NFA; ++x)
#pragma unroll
                for (int d = 0; d < NFD; ++d) kil[x][d] = 0.0;
#pragma unroll
            for (int x = 0; x < NFB; ++x)
#pragma unroll
                for (int d = 0; d < NFD; ++d) kjl[x][d] = 0.0;
#pragma unroll
            for (int ia = 0; ia < NFA; ++ia)
#pragma unroll
                for (int jb = 0; jb < NFB; ++jb)
#pragma unroll
                    for (int kc = 0; kc < NFC; ++kc)
#pragma unroll
                        for (int ld = 0; ld < NFD; ++ld) {
                            const double v = (double)acc[ia * NFB + jb][kc * NFD + ld];
                            kik[ia][kc] += v * djl[jb][ld];
                            kjk[jb][kc] += v * dil[ia][ld];
                            kil[ia][ld] += v * djk[jb][kc];
                            kjl[jb][ld] += v * dik[ia][kc];
                        }
#pragma unroll
            for (int x = 0; x < NFA; ++x)
#pragma unroll
                for (int d = 0; d < NFD; ++d) {
                    const double v = seg_scan(kil[x][d], lane, seg_head);
                    if (seg_tail && pvalid && v != 0.0) unsafeAtomicAdd(VK + ri + x * nao + l0 + d, v);
                }
            if (active) {
#pragma unroll
                for (int x = 0; x < NFB; ++x)
#pragma unroll
                    for (int d = 0; d < NFD; ++d) atomicAdd(&kjl_slot[wv][jlead][x * NFD + d], kjl[x][d]);
            }
            // LDS instructions of one wave execute in issue order, so the leaders' reads below
            // see every lane's add; only the compiler must not reorder them (a memory fence here
            // would also wait for the wave's outstanding global atomics every ket)
            __builtin_amdgcn_wave_barrier();
            if (jlead == lane && pvalid) {
#pragma unroll
                for (int x = 0; x < NFB; ++x)
#pragma unroll
                    for (int d = 0; d < NFD; ++d) {
                        const double v = kjl_slot[wv][lane][x * NFD + d];
                        if (v != 0.0) {
                            unsafeAtomicAdd(VK + rj + x * nao + l0 + d, v);
                            kjl_slot[wv][lane][x * NFD + d] = 0.0;
                        }
                    }
            }
        }
    }
    // ---- final flushes
    if (dok && kdirty) {
#pragma unroll
        for (int x = 0; x < NFA; ++x)
#pragma unroll
            for (int c = 0; c < NFC; ++c) unsafeAtomicAdd(VK + ri + x * nao + cur_k0 + c, kik[x][c]);
#pragma unroll
        for (int x = 0; x < NFB; ++x)
#pragma unroll
            for (int c = 0; c < NFC; ++c) unsafeAtomicAdd(VK + rj + x * nao + cur_k0 + c, kjk[x][c]);
    }
    if (doj && pvalid) {
#pragma unroll
        for (int r = 0; r < NFAB; ++r)
            if (jij[r] != 0.0) unsafeAtomicAdd(VJ + ri + (r / NFB) * nao + j0 + r % NFB, jij[r]);
    }
    if (a.counter) {
        double cnt = wave_sum((double)nquart);
        if (lane == 0 && cnt > 0) atomicAdd(a.counter, (unsigned long long)cnt);
        if (lane == 0 && a.slots) atomicAdd(a.slots, 64ull * nslots);
    }
}

}  // namespace mjk
