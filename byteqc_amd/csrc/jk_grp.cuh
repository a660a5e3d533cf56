// LDS-staged J/K kernel for the larger angular-momentum classes.
//
// A workgroup is one wavefront (64 lanes) split into W groups of S lanes; group g owns
// bra shell pair p = tile * W + g and all groups walk the same chunk of ket pairs.
// For each ket and primitive quartet the group's lanes evaluate the 3 N_roots 2D
// tables (one root and one Cartesian direction per lane, looping when S < 3 N_roots)
// into the group's LDS slice; then each lane assembles the integrals of its RPT bra
// rows and every ket column in registers.  Digestion:
//   * J_ij per lane row, kept in registers over the chunk;
//   * J_kl is uniform over the wave: wave-wide sum, one atomic per element;
//   * K_ik, K_jk are summed in LDS while consecutive kets share shell k, K_il, K_jl per
//     quartet, then written with one FP64 atomic per element.
// The six updates are those of cuobc/lib/gvhf/contract_jk.cu:219-316.
#pragma once
#include "jk_device.cuh"
#include "jk_reg.cuh"

namespace mjk {

// longest ket chunk per tile: shorter chunks give more tiles and shorter per-tile chains of
// ket iterations, which is what bounds a sparse (density-screened) launch of the large classes
constexpr int GRP_KC = 32;

template <int LA, int LB, int LC, int LD>
struct GCfg {
    static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    static constexpr int NFA = ncart(LA), NFB = ncart(LB), NFC = ncart(LC), NFD = ncart(LD);
    static constexpr int NFAB = NFA * NFB, NFCD = NFC * NFD;
    static constexpr int G2 = (LC + 1) * (LD + 1);
    static constexpr int GS = (LA + 1) * (LB + 1) * G2;
    static constexpr int TAB = NR * 3 * GS;                      // 2D tables per group
    static constexpr int KIK = 0, KJK = NFA * NFC, KIL = KJK + NFB * NFC, KJL = KIL + NFA * NFD;
    static constexpr int KBUF = KJL + NFB * NFD;                 // exchange partial sums per group
    static constexpr int ACC = 72;                               // accumulators per lane
    static constexpr int LDS_BYTES = 24 * 1024;

    static constexpr int rpt_of(int s) { return (NFAB + s - 1) / s; }
    static constexpr int w_of(int s) { return 64 / s; }
    static constexpr bool fits(int s)
    {
        return s <= 64 && s <= NFAB && rpt_of(s) * NFCD <= (NFCD > ACC ? NFCD : ACC) &&
               (w_of(s) == 1 || w_of(s) * (TAB * 8 + KBUF * 8) <= LDS_BYTES);
    }
    // lanes per quartet: best use of the 64 lanes' row slots (W NFAB / (64 RPT)), then
    // more groups per wave
    static constexpr int pick()
    {
        int best = 0, bnum = 0, bden = 1, bw = 0;
        for (int s = 1; s <= 64; ++s) {
            if (!fits(s)) continue;
            const int num = w_of(s) * NFAB, den = 64 * rpt_of(s);
            if (best == 0 || num * bden > bnum * den || (num * bden == bnum * den && w_of(s) > bw)) {
                best = s; bnum = num; bden = den; bw = w_of(s);
            }
        }
        return best;
    }
    static constexpr bool OK = pick() > 0;
    static constexpr int S = OK ? pick() : 1;
    static constexpr int RPT = rpt_of(S);
    static constexpr int W = w_of(S);
};

// Classes whose group kernel runs faster compiled for two waves per SIMD (a few scratch
// spills against twice the latency hiding), measured per class on (H2O)64/def2-TZVP:
// (fp|fp) 107 -> 67 ms, (fd|fp) 53 -> 40, (fp|ds) 59 -> 47, (dd|ss) 22 -> 19; where the spills are larger it
// loses ((fp|dd) 64 -> 133) and the class keeps the compiler's default.
constexpr int grp_waves_per_eu(int la, int lb, int lc, int ld)
{
    constexpr int t[][4] = {{3, 1, 3, 1}, {3, 2, 3, 1}, {3, 1, 2, 0}, {3, 1, 1, 0}, {3, 2, 3, 0},
                            {3, 1, 1, 1}, {2, 2, 2, 0}, {3, 3, 3, 0}, {3, 2, 2, 0}, {3, 2, 2, 2},
                            {3, 3, 2, 0}, {2, 2, 2, 1}, {3, 2, 1, 1}, {2, 2, 0, 0}, {3, 3, 1, 1},
                            {3, 2, 2, 1}};
    for (const auto &c : t)
        if (c[0] == la && c[1] == lb && c[2] == lc && c[3] == ld) return 2;
    return 1;
}

template <typename R, int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(grp_waves_per_eu(LA, LB, LC, LD))))
jk_grp_kernel(const JKArgs a)
{
    using C = GCfg<LA, LB, LC, LD>;
    constexpr int NR = C::NR, NFA = C::NFA, NFB = C::NFB, NFC = C::NFC, NFD = C::NFD;
    constexpr int NFAB = C::NFAB, NFCD = C::NFCD, GS = C::GS, G2 = C::G2;
    constexpr int S = C::S, RPT = C::RPT, W = C::W;
    __shared__ R tab[W][C::TAB];
    __shared__ double kbuf[W][C::KBUF];

    const long tile = a.tile_base + (long)blockIdx.x;
    const int lane = threadIdx.x;
    const int g = lane / S, sub = lane % S;
    const bool lane_ok = g < W;
    const int gg = lane_ok ? g : 0;
    int p, q_begin, q_end, p_first = 0;
    bool pvalid;
    const double *D = a.dm;
    double *VJ = a.vj, *VK = a.vk;
    int nao = a.nao, ao_base = 0;
    if (a.tiles) {
        const int4 d = a.tiles[tile];
        p_first = d.x;
        p = d.x + g;
        pvalid = lane_ok && p < d.y;
        q_begin = d.z;
        q_end = d.w;
        const MolRec m = a.mols[a.tile_mol[tile]];
        D += m.off;
        VJ = VJ ? VJ + m.off : nullptr;
        VK = VK ? VK + m.off : nullptr;
        nao = m.nao;
        ao_base = m.ao_base;
    } else {
        const int bi = (int)(tile / a.nb_ket), ki = (int)(tile % a.nb_ket);
        p = bi * W + g;
        q_begin = ki * a.kc;
        if (a.same && q_begin > bi * W + W - 1) return;  // the tile lies above the triangle
        q_end = q_begin + a.kc;
        if (q_end > a.ket.npairs) q_end = a.ket.npairs;
        pvalid = lane_ok && p < a.bra.npairs;
    }
    // lanes without a bra pair borrow the tile's first one, which lies in the same molecule,
    // so every index they form stays inside this molecule's D block
    const int pp = pvalid ? p : p_first;
    const int4 b4 = a.bra.pint[pp];
    const double *bd = a.bra.pdbl + 8 * pp;
    const int ish = b4.x, jsh = b4.y;
    const double qp = pvalid ? bd[0] : 0.0;
    const double dmax = a.dm_max ? *a.dm_max : 1.0;
    {
        double qk = 0.0;
        for (int q = q_begin + lane; q < q_end; q += 64) qk = fmax(qk, a.ket.pdbl[8 * q]);
        double qb = qp;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            qk = fmax(qk, __shfl_xor(qk, off, 64));
            qb = fmax(qb, __shfl_xor(qb, off, 64));
        }
        // per-tile precision split: lo <= (largest bound of the tile) < hi
        const double m = qb * qk * dmax;
        if (!(m > a.tol && m >= a.lo && m < a.hi)) return;  // wave-uniform
    }
    const double A[3] = {bd[1], bd[2], bd[3]};
    const double AB[3] = {bd[4], bd[5], bd[6]};
    const int i0 = b4.z - ao_base, j0 = b4.w - ao_base;
    const bool doj = VJ != nullptr, dok = VK != nullptr;

    // this lane's bra rows
    int ox[RPT], oy[RPT], oz[RPT], rowia[RPT], rowjb[RPT];
    bool rowok[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int row = sub + k * S;
        rowok[k] = lane_ok && row < NFAB;
        const int rr = rowok[k] ? row : 0;
        const int ia = rr / NFB, jb = rr % NFB;
        rowia[k] = ia;
        rowjb[k] = jb;
        ox[k] = (cart_x(LA, ia) * (LB + 1) + cart_x(LB, jb)) * G2;
        oy[k] = (cart_y(LA, ia) * (LB + 1) + cart_y(LB, jb)) * G2;
        oz[k] = (cart_z(LA, ia) * (LB + 1) + cart_z(LB, jb)) * G2;
    }
    double jij[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) jij[k] = 0.0;
    for (int x = sub; x < C::KBUF; x += S)
        if (lane_ok) kbuf[gg][x] = 0.0;
    int cur_k = -1, cur_k0 = 0;
    unsigned long long nquart = 0, nslots = 0;
    const int npb = a.bra.nprim, npk = a.ket.nprim;
    __syncthreads();

    for (int q = q_begin; q < q_end; ++q) {
        const int4 k4 = a.ket.pint[q];
        const int ksh = __builtin_amdgcn_readfirstlane(k4.x);
        const int lsh = __builtin_amdgcn_readfirstlane(k4.y);
        const int k0 = __builtin_amdgcn_readfirstlane(k4.z) - ao_base;
        const int l0 = __builtin_amdgcn_readfirstlane(k4.w) - ao_base;
        const double *kd = a.ket.pdbl + 8 * q;
        if (ksh != cur_k) {
            if (dok && cur_k >= 0 && pvalid) {
                for (int x = sub; x < C::KIL; x += S) {
                    const double v = kbuf[gg][x];
                    if (v != 0.0) {
                        const size_t row = x < C::KJK ? (size_t)(i0 + x / NFC) : (size_t)(j0 + (x - C::KJK) / NFC);
                        unsafeAtomicAdd(VK + row * nao + cur_k0 + x % NFC, v);
                        kbuf[gg][x] = 0.0;
                    }
                }
            }
            cur_k = ksh;
            cur_k0 = k0;
            __syncthreads();
        }
        bool active = pvalid && (!a.same || q <= p);
        if (active) {
            double bound = qp * kd[0];
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
        if (!__builtin_amdgcn_ballot_w64(active)) continue;  // wave-uniform
        if (active && sub == 0) ++nquart;
        ++nslots;
        R acc[RPT][NFCD];
#pragma unroll
        for (int k = 0; k < RPT; ++k)
#pragma unroll
            for (int c = 0; c < NFCD; ++c) acc[k][c] = (R)0;
        const double Cc[3] = {kd[1], kd[2], kd[3]};
        const double CD[3] = {kd[4], kd[5], kd[6]};
        const double norm = (a.same && p == q) ? 0.5 * PI52X2 : PI52X2;
        for (int ip = 0; ip < npb; ++ip) {
            for (int kq = 0; kq < npk; ++kq) {
                const double *bp_ = a.bra.prim + 6 * (pp * npb + ip);
                const double *kp_ = a.ket.prim + 6 * (q * npk + kq);
                const double aij = bp_[0], akl = kp_[0], iaij = bp_[5], iakl = kp_[5];
                const double iaijkl = 1.0 / (aij + akl);
                const double pref = active ? norm * bp_[1] * kp_[1] * iaij * iakl * sqrt(iaijkl) : 0.0;
                const bool pactive = active && fabs(pref) >= a.prim_tol;
                if (!__builtin_amdgcn_ballot_w64(pactive)) continue;  // wave-uniform: barriers stay paired
                if (pactive) {
                    const double P[3] = {bp_[2], bp_[3], bp_[4]};
                    const double Q[3] = {kp_[2], kp_[3], kp_[4]};
                    const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                    const double T = aij * akl * iaijkl * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                    const double inv2 = 0.5 * iaijkl;
                    for (int it = sub; it < 3 * NR; it += S) {
                        const int r = it / 3, d = it % 3;
                        R x, w;
                        rys_root<R, NR>(T, r, x, w);
                        const R b00 = x * (R)inv2;
                        const R b10 = ((R)0.5 - b00 * (R)akl) * (R)iaij;
                        const R b01 = ((R)0.5 - b00 * (R)aij) * (R)iakl;
                        const R tb = (R)2 * b00;
                        const double Pd = d == 0 ? P[0] : (d == 1 ? P[1] : P[2]);
                        const double Qd = d == 0 ? Q[0] : (d == 1 ? Q[1] : Q[2]);
                        const double PQd = d == 0 ? PQ[0] : (d == 1 ? PQ[1] : PQ[2]);
                        const double Ad = d == 0 ? A[0] : (d == 1 ? A[1] : A[2]);
                        const double Cd = d == 0 ? Cc[0] : (d == 1 ? Cc[1] : Cc[2]);
                        const double ABd = d == 0 ? AB[0] : (d == 1 ? AB[1] : AB[2]);
                        const double CDd = d == 0 ? CD[0] : (d == 1 ? CD[1] : CD[2]);
                        R g0 = (R)1;
                        if (d == 2) g0 = w * (R)pref;
                        rys2d<R, LA, LB, LC, LD>(g0, (R)(Pd - Ad) - tb * (R)akl * (R)PQd,
                                                 (R)(Qd - Cd) + tb * (R)aij * (R)PQd, b00, b10, b01,
                                                 (R)ABd, (R)CDd, &tab[gg][(r * 3 + d) * GS]);
                    }
                }
                __syncthreads();
                if (pactive) {
#pragma unroll 1
                    for (int r = 0; r < NR; ++r) {
                        const R *gx = &tab[gg][(r * 3) * GS], *gy = gx + GS, *gz = gx + 2 * GS;
#pragma unroll
                        for (int k = 0; k < RPT; ++k) {
                            R xs[G2], ys[G2], zs[G2];
#pragma unroll
                            for (int e = 0; e < G2; ++e) {
                                xs[e] = gx[ox[k] + e];
                                ys[e] = gy[oy[k] + e];
                                zs[e] = gz[oz[k] + e];
                            }
#pragma unroll
                            for (int c = 0; c < NFCD; ++c) {
                                const int kc = c / NFD, ld = c % NFD;
                                const int cx = cart_x(LC, kc) * (LD + 1) + cart_x(LD, ld);
                                const int cy = cart_y(LC, kc) * (LD + 1) + cart_y(LD, ld);
                                const int cz = cart_z(LC, kc) * (LD + 1) + cart_z(LD, ld);
                                acc[k][c] += xs[cx] * ys[cy] * zs[cz];
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        // ---- digestion
        if (doj) {
            double jkl[NFCD];
#pragma unroll
            for (int c = 0; c < NFCD; ++c) jkl[c] = 0.0;
            if (active) {
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    if (!rowok[k]) continue;
                    const double dij = D[(size_t)(i0 + rowia[k]) * nao + j0 + rowjb[k]];
#pragma unroll
                    for (int c = 0; c < NFCD; ++c) {
                        const double v = (double)acc[k][c];
                        jij[k] += v * D[(size_t)(k0 + c / NFD) * nao + l0 + c % NFD];
                        jkl[c] += v * dij;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < NFCD; ++c) {
                const double s = wave_sum(jkl[c]);
                if (lane == 0) unsafeAtomicAdd(VJ + (size_t)(k0 + c / NFD) * nao + l0 + c % NFD, s);
            }
        }
        if (dok) {
            if (active) {
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    if (!rowok[k]) continue;
                    const int ia = rowia[k], jb = rowjb[k];
                    const size_t ri = (size_t)(i0 + ia) * nao, rj = (size_t)(j0 + jb) * nao;
                    double pik[NFC], pjk[NFC], pil[NFD], pjl[NFD];
#pragma unroll
                    for (int x = 0; x < NFC; ++x) { pik[x] = 0.0; pjk[x] = 0.0; }
#pragma unroll
                    for (int x = 0; x < NFD; ++x) { pil[x] = 0.0; pjl[x] = 0.0; }
#pragma unroll
                    for (int kc = 0; kc < NFC; ++kc) {
                        const double djk = D[rj + k0 + kc], dik = D[ri + k0 + kc];
#pragma unroll
                        for (int ld = 0; ld < NFD; ++ld) {
                            const double v = (double)acc[k][kc * NFD + ld];
                            pik[kc] += v * D[rj + l0 + ld];
                            pjk[kc] += v * D[ri + l0 + ld];
                            pil[ld] += v * djk;
                            pjl[ld] += v * dik;
                        }
                    }
#pragma unroll
                    for (int x = 0; x < NFC; ++x) {
                        atomicAdd(&kbuf[gg][C::KIK + ia * NFC + x], pik[x]);
                        atomicAdd(&kbuf[gg][C::KJK + jb * NFC + x], pjk[x]);
                    }
#pragma unroll
                    for (int x = 0; x < NFD; ++x) {
                        atomicAdd(&kbuf[gg][C::KIL + ia * NFD + x], pil[x]);
                        atomicAdd(&kbuf[gg][C::KJL + jb * NFD + x], pjl[x]);
                    }
                }
            }
            __syncthreads();
            if (active) {
                for (int x = C::KIL + sub; x < C::KBUF; x += S) {
                    const size_t row = x < C::KJL ? (size_t)(i0 + (x - C::KIL) / NFD) : (size_t)(j0 + (x - C::KJL) / NFD);
                    unsafeAtomicAdd(VK + row * nao + l0 + (x - C::KIL) % NFD, kbuf[gg][x]);
                    kbuf[gg][x] = 0.0;
                }
            }
            __syncthreads();
        }
    }
    if (dok && cur_k >= 0 && pvalid) {
        for (int x = sub; x < C::KIL; x += S) {
            const double v = kbuf[gg][x];
            if (v != 0.0) {
                const size_t row = x < C::KJK ? (size_t)(i0 + x / NFC) : (size_t)(j0 + (x - C::KJK) / NFC);
                unsafeAtomicAdd(VK + row * nao + cur_k0 + x % NFC, v);
            }
        }
    }
    if (doj && pvalid) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (rowok[k] && jij[k] != 0.0)
                unsafeAtomicAdd(VJ + (size_t)(i0 + rowia[k]) * nao + j0 + rowjb[k], jij[k]);
    }
    if (a.counter) {
        const double cnt = wave_sum((double)nquart);
        if (lane == 0 && cnt > 0) atomicAdd(a.counter, (unsigned long long)cnt);
        if (lane == 0 && a.slots) atomicAdd(a.slots, (unsigned long long)W * nslots);
    }
}

}  // namespace mjk
