// Device code of the direct J/K builder: Rys nodes, 2D recursion, per-class kernels.
//
// Work decomposition (one launch per class pair, i.e. per (l_i l_j | l_k l_l) and
// primitive-count combination):
//   * bra and ket shell-pair lists are sorted by descending Schwarz factor q;
//   * the (bra, ket) plane is cut into tiles of BP bra pairs x KP ket pairs; one
//     workgroup owns one tile (the upper triangle is not launched when bra and ket are
//     the same class);
//   * S threads share one quartet.  Thread `sub` of a quartet owns the bra Cartesian
//     rows sub, sub + S, ... and every ket column, accumulating the contracted
//     integrals in registers over the primitive and root loops;
//   * REG classes (S == 1, small tables) keep each root's 2D integrals in registers;
//     LDS classes stage the 2D tables of all roots of the tile's quartets in LDS, the
//     recursions spread over consecutive threads, the assembly reading them back;
//   * digestion: J_ij, J_kl, K_ik, K_il, K_jk, K_jl partial sums per thread, then FP64
//     atomics into the unsymmetrised V matrices (cuobc/lib/gvhf/contract_jk.cu:219-316
//     defines the six updates).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RYS_STORAGE static __device__ const
#include "rys_tables.inc"

namespace mjk {

constexpr double PI52X2 = 34.98683665524972497;  // 2 pi^(5/2)

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian component c of angular momentum l: powers in libcint order
// (x power descending, then y power descending).
__host__ __device__ constexpr int cart_x(int l, int c)
{
    int lx = l, n = c;
    while (n > l - lx) { n -= l - lx + 1; --lx; }
    return lx;
}
__host__ __device__ constexpr int cart_y(int l, int c)
{
    int lx = l, n = c;
    while (n > l - lx) { n -= l - lx + 1; --lx; }
    return l - lx - n;
}
__host__ __device__ constexpr int cart_z(int l, int c) { return l - cart_x(l, c) - cart_y(l, c); }

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

// One shell-pair class, pointers offset to the class start.
struct PairSet {
    const int *ish, *jsh;
    const double *q;
    const double *a, *e, *px, *py, *pz;  // [npairs * nprim]
    // per pair: AO offsets of both shells, centre of shell i and R_i - R_j, so that a
    // pair's geometry is one level of independent loads
    const int *aoi, *aoj;
    const double *cx, *cy, *cz, *abx, *aby, *abz;
    // packed records used by the register and group kernels (three base pointers keep
    // the kernel arguments in scalar registers):
    //   pint[p] = {ish, jsh, ao_i, ao_j}
    //   pdbl[8p + 0..6] = {q, x_i, y_i, z_i, (R_i - R_j)_x, _y, _z}
    //   prim[6 (p nprim + k) + 0..5] = {a, e, P_x, P_y, P_z, 1/a}
    const int4 *pint;
    const double *pdbl;
    const double *prim;
    int npairs, nprim;
};

// One molecule of a batched build: its D, J and K start at element `off` of the batch
// arrays (nao x nao each) and its shells' AO indices start at ao_base.
struct MolRec {
    long off;
    int nao, ao_base;
};

struct JKArgs {
    PairSet bra, ket;
    // batched builds: tile t covers bra pairs [x, y) and ket pairs [z, w) of tiles[t], all
    // of molecule tile_mol[t]; null for a single molecule (tiles from the class sizes)
    const int4 *tiles;
    const int *tile_mol;
    const MolRec *mols;
    long ndesc;             // number of descriptors in tiles
    const double *bx, *by, *bz;
    const int *ao_loc;
    const double *dm;
    double *vj, *vk;
    const double *dm_cond;  // nbas x nbas shell-block max |D| or null
    unsigned long long *counter;  // quartets computed (optional)
    unsigned long long *slots;    // lane slots of executed ket iterations (optional)
    double tol;             // keep quartets with bound > tol
    double lo, hi;          // precision split: the register and group kernels take the tiles whose
                            // largest bound m has lo <= m < hi; the block-LDS kernel takes the
                            // quartets with lo <= bound < hi
    int nbas, nao, n_dm;
    const double *dm_max;   // device scalar: global max |D| (null: 1)
    long tile_base;         // first tile of this launch
    int kc;                 // ket pairs per tile (register and group kernels)
    int kc_cap;             // upper bound on kc set by the caller (0: the kernel's default)
    double prim_tol;        // primitive quartets with |prefactor| below this are skipped
    int same;               // bra and ket are the same class: p >= q only
    int nb_bra, nb_ket;     // tiles along bra / ket
};

// ---- Rys nodes (x = t^2) and weights --------------------------------------------
template <int N>
__device__ inline void rys_root64(double T, int r, double &x, double &w)
{
    if (T >= rys_tmax[N]) {
        const int o = N * (N - 1) / 2 + r;
        const double it = 1.0 / T;
        x = rys_herm_s2[o] * it;
        w = rys_herm_w[o] * sqrt(it);
        return;
    }
    const int i = (int)T;  // RYS_H64 == 1
    const double s = 2.0 * (T - (double)i) - 1.0;
    const double *cx = rys_tab64 + rys_off64[N] + ((size_t)i * 2 * N + r) * (RYS_DEG64 + 1);
    const double *cw = cx + (size_t)N * (RYS_DEG64 + 1);
    double b1x = 0.0, b2x = 0.0, b1w = 0.0, b2w = 0.0;
#pragma unroll
    for (int j = RYS_DEG64; j >= 1; --j) {
        const double tx = 2.0 * s * b1x - b2x + cx[j];
        const double tw = 2.0 * s * b1w - b2w + cw[j];
        b2x = b1x; b1x = tx;
        b2w = b1w; b1w = tw;
    }
    x = s * b1x - b2x + cx[0];
    w = s * b1w - b2w + cw[0];
}

// FP32 nodes; large T (the asymptotic branch) is evaluated in FP64.
template <int N>
__device__ inline void rys_root32(double T, int r, float &x, float &w)
{
    if (T >= rys_tmax[N]) {
        double xd, wd;
        const int o = N * (N - 1) / 2 + r;
        const double it = 1.0 / T;
        xd = rys_herm_s2[o] * it;
        wd = rys_herm_w[o] * sqrt(it);
        x = (float)xd;
        w = (float)wd;
        return;
    }
    const int i = (int)(T * 0.5);  // RYS_H32 == 2
    const float s = (float)(T - 2.0 * i) - 1.0f;
    const float *cx = rys_tab32 + rys_off32[N] + ((size_t)i * 2 * N + r) * (RYS_DEG32 + 1);
    const float *cw = cx + (size_t)N * (RYS_DEG32 + 1);
    float b1x = 0.f, b2x = 0.f, b1w = 0.f, b2w = 0.f;
#pragma unroll
    for (int j = RYS_DEG32; j >= 1; --j) {
        const float tx = 2.f * s * b1x - b2x + cx[j];
        const float tw = 2.f * s * b1w - b2w + cw[j];
        b2x = b1x; b1x = tx;
        b2w = b1w; b1w = tw;
    }
    x = s * b1x - b2x + cx[0];
    w = s * b1w - b2w + cw[0];
}

// One-root rule in closed form: w = F0(T), x = F1(T)/F0(T), with
// F0 = sqrt(pi/T) erf(sqrt T)/2, F1 = (F0 - exp(-T))/(2T), and the Taylor series
// F_n = sum_k (-T)^k / (k! (2n + 2k + 1)) below T = 0.1 where the difference cancels.
__device__ inline void rys_root1_closed(double T, double &x, double &w)
{
    double f0, f1;
    if (T < 0.1) {
        double t = 1.0, s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int k = 0; k < 10; ++k) {
            s0 += t / (2 * k + 1);
            s1 += t / (2 * k + 3);
            t *= -T / (k + 1);
        }
        f0 = s0;
        f1 = s1;
    } else {
        const double st = sqrt(T);
        f0 = 0.886226925452758014 * erf(st) / st;  // sqrt(pi)/2
        f1 = (f0 - exp(-T)) * 0.5 / T;
    }
    w = f0;
    x = f1 / f0;
}

// One-root rule from the Boys grid: F_n(T_i - d) = sum_k F_(n+k)(T_i) d^k / k!, ten terms,
// |d| <= h/2 = 0.05 (truncation below 1e-17); above the grid F0 = sqrt(pi/T)/2 and
// F1/F0 = 1/(2T) (exp(-T) < 5e-18 there).
__device__ inline void rys_root1_grid(double T, double &x, double &w)
{
    if (T >= RYS_BOYS_TMAX) {
        w = 0.886226925452758014 / sqrt(T);
        x = 0.5 / T;
        return;
    }
    constexpr double inv_fact[10] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720,
                                     1.0 / 5040, 1.0 / 40320, 1.0 / 362880};
    const int i = (int)(T * (1.0 / RYS_BOYS_H) + 0.5);
    const double d = i * RYS_BOYS_H - T;
    const double *g = rys_boys_grid + i * (RYS_BOYS_K + 1);
    double f0 = g[9] * inv_fact[9], f1 = g[10] * inv_fact[9];
#pragma unroll
    for (int k = 8; k >= 0; --k) {
        f0 = f0 * d + g[k] * inv_fact[k];
        f1 = f1 * d + g[k + 1] * inv_fact[k];
    }
    w = f0;
    x = f1 / f0;
}

template <typename R, int N>
__device__ inline void rys_root(double T, int r, R &x, R &w);
template <> __device__ inline void rys_root<double, 1>(double T, int r, double &x, double &w) { rys_root1_grid(T, x, w); }
template <> __device__ inline void rys_root<double, 2>(double T, int r, double &x, double &w) { rys_root64<2>(T, r, x, w); }
template <> __device__ inline void rys_root<double, 3>(double T, int r, double &x, double &w) { rys_root64<3>(T, r, x, w); }
template <> __device__ inline void rys_root<double, 4>(double T, int r, double &x, double &w) { rys_root64<4>(T, r, x, w); }
template <> __device__ inline void rys_root<double, 5>(double T, int r, double &x, double &w) { rys_root64<5>(T, r, x, w); }
template <> __device__ inline void rys_root<double, 6>(double T, int r, double &x, double &w) { rys_root64<6>(T, r, x, w); }
template <> __device__ inline void rys_root<double, 7>(double T, int r, double &x, double &w) { rys_root64<7>(T, r, x, w); }
template <> __device__ inline void rys_root<float, 1>(double T, int r, float &x, float &w)
{
    double xd, wd;
    rys_root1_grid(T, xd, wd);
    x = (float)xd;
    w = (float)wd;
}
template <> __device__ inline void rys_root<float, 2>(double T, int r, float &x, float &w) { rys_root32<2>(T, r, x, w); }
template <> __device__ inline void rys_root<float, 3>(double T, int r, float &x, float &w) { rys_root32<3>(T, r, x, w); }
template <> __device__ inline void rys_root<float, 4>(double T, int r, float &x, float &w) { rys_root32<4>(T, r, x, w); }
template <> __device__ inline void rys_root<float, 5>(double T, int r, float &x, float &w) { rys_root32<5>(T, r, x, w); }
template <> __device__ inline void rys_root<float, 6>(double T, int r, float &x, float &w) { rys_root32<6>(T, r, x, w); }
template <> __device__ inline void rys_root<float, 7>(double T, int r, float &x, float &w) { rys_root32<7>(T, r, x, w); }

// ---- 2D integrals of one root and one Cartesian direction ------------------------
// Vertical recursion in (n, m) = (i + j, k + l) from g(0,0) = g0:
//   g(n+1, 0) = c00 g(n, 0) + n b10 g(n-1, 0)
//   g(n, m+1) = c0p g(n, m) + m b01 g(n, m-1) + n b00 g(n-1, m)
// then horizontal transfers g(i, j+1) = g(i+1, j) + AB g(i, j) and likewise with CD.
// out[((i (LB+1) + j) (LC+1) + k) (LD+1) + l], i <= LA, j <= LB, k <= LC, l <= LD.
template <typename R, int LA, int LB, int LC, int LD, typename OUT>
__device__ inline void rys2d(R g0, R c00, R c0p, R b00, R b10, R b01, R ab, R cd, OUT out)
{
    constexpr int NAB = LA + LB, NCD = LC + LD;
    R v[NAB + 1][NCD + 1];
    v[0][0] = g0;
#pragma unroll
    for (int n = 0; n < NAB; ++n) {
        R t = c00 * v[n][0];
        if (n > 0) t += (R)n * b10 * v[n - 1][0];
        v[n + 1][0] = t;
    }
#pragma unroll
    for (int m = 0; m < NCD; ++m) {
#pragma unroll
        for (int n = 0; n <= NAB; ++n) {
            R t = c0p * v[n][m];
            if (m > 0) t += (R)m * b01 * v[n][m - 1];
            if (n > 0) t += (R)n * b00 * v[n - 1][m];
            v[n][m + 1] = t;
        }
    }
    R h[NAB + 1][LB + 1][NCD + 1];
#pragma unroll
    for (int i = 0; i <= NAB; ++i)
#pragma unroll
        for (int m = 0; m <= NCD; ++m) h[i][0][m] = v[i][m];
#pragma unroll
    for (int j = 1; j <= LB; ++j)
#pragma unroll
        for (int i = 0; i <= NAB - j; ++i)
#pragma unroll
            for (int m = 0; m <= NCD; ++m) h[i][j][m] = h[i + 1][j - 1][m] + ab * h[i][j - 1][m];
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j) {
            R o[NCD + 1][LD + 1];
#pragma unroll
            for (int k = 0; k <= NCD; ++k) o[k][0] = h[i][j][k];
#pragma unroll
            for (int l = 1; l <= LD; ++l)
#pragma unroll
                for (int k = 0; k <= NCD - l; ++k) o[k][l] = o[k + 1][l - 1] + cd * o[k][l - 1];
#pragma unroll
            for (int k = 0; k <= LC; ++k)
#pragma unroll
                for (int l = 0; l <= LD; ++l)
                    out[((i * (LB + 1) + j) * (LC + 1) + k) * (LD + 1) + l] = o[k][l];
        }
}

// ---- per-class launch configuration ----------------------------------------------
// Classes (canonical orientation) under the register limit that ran faster on the
// wave-group kernel, measured per class on (H2O)64/def2-TZVP: (pp|ds) 93 -> 56 ms,
// (ds|ds) 47 -> 35, (dd|ss) 31 -> 19 ((dp|ps) went 77 -> 93 and stays on the register kernel).
__host__ __device__ constexpr bool prefer_group(int la, int lb, int lc, int ld)
{
    return (la == 1 && lb == 1 && lc == 2 && ld == 0) || (la == 2 && lb == 0 && lc == 2 && ld == 0) ||
           (la == 2 && lb == 2 && lc == 0 && ld == 0);
}

#ifndef REG_LIMIT
#define REG_LIMIT 150
#endif
template <int LA, int LB, int LC, int LD>
struct Cfg {
    static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    static constexpr int NFA = ncart(LA), NFB = ncart(LB), NFC = ncart(LC), NFD = ncart(LD);
    static constexpr int NFAB = NFA * NFB, NFCD = NFC * NFD;
    static constexpr int G2 = (LC + 1) * (LD + 1);
    static constexpr int GS = (LA + 1) * (LB + 1) * G2;
    // register-resident classes: one root's 2D tables, the quartet's integrals and the
    // per-lane exchange/Coulomb partial sums fit the 512-entry VGPR+AGPR file
    static constexpr bool REG = 3 * GS + NFAB * NFCD + NFAB + (NFA + NFB) * NFC <= REG_LIMIT &&
                                !prefer_group(LA, LB, LC, LD);
    static constexpr int ACC_MAX = 64;
    static constexpr int RPT0 = REG ? NFAB : cmax(1, cmin(NFAB, ACC_MAX / NFCD));
    static constexpr int S = REG ? 1 : (NFAB + RPT0 - 1) / RPT0;
    static constexpr int RPT = (NFAB + S - 1) / S;
    static constexpr int LDS_Q = NR * 3 * GS;  // doubles of 2D tables per quartet
    static constexpr int QB = REG ? 256 : cmax(1, cmin(256 / S, (48 * 1024 / 8) / LDS_Q));
    static constexpr int BP = REG ? 16 : 1;
    static constexpr int KP = REG ? 16 : QB;
    static constexpr int NT = QB * S;
    static constexpr int LDS_G = REG ? 1 : QB * LDS_Q;
};

// ---- tiles -----------------------------------------------------------------------
// Same-class launches only visit tiles holding some p >= q.  Square tiles (BP == KP)
// form a triangle; BP == 1 tiles give bra pair p the ket blocks 0 .. p / KP.
__host__ __device__ inline long tri_tiles_bp1(long p, int kp)
{
    // sum_{p' < p} (p' / kp + 1)
    const long m = p / kp, r = p % kp;
    return p + kp * m * (m - 1) / 2 + m * r;
}

template <int BP, int KP>
__device__ inline void tile_coords(const JKArgs &a, long tile, int &bi, int &ki)
{
    if (!a.same) {
        bi = (int)(tile / a.nb_ket);
        ki = (int)(tile % a.nb_ket);
    } else if constexpr (BP == KP) {
        long b = (long)((sqrt(8.0 * (double)tile + 1.0) - 1.0) * 0.5);
        while (b * (b + 1) / 2 > tile) --b;
        while ((b + 1) * (b + 2) / 2 <= tile) ++b;
        bi = (int)b;
        ki = (int)(tile - b * (b + 1) / 2);
    } else {
        static_assert(BP == 1, "same-class tiles need BP == KP or BP == 1");
        long lo = 0, hi = a.nb_bra;  // largest p with tri_tiles_bp1(p) <= tile
        while (hi - lo > 1) {
            const long mid = (lo + hi) / 2;
            if (tri_tiles_bp1(mid, KP) <= tile) lo = mid; else hi = mid;
        }
        bi = (int)lo;
        ki = (int)(tile - tri_tiles_bp1(lo, KP));
    }
}

// ---- the kernel ------------------------------------------------------------------
template <typename R, int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__((Cfg<LA, LB, LC, LD>::NT))
jk_kernel(const JKArgs a)
{
    using C = Cfg<LA, LB, LC, LD>;
    constexpr int NR = C::NR, NFB = C::NFB, NFC = C::NFC, NFD = C::NFD;
    constexpr int NFAB = C::NFAB, NFCD = C::NFCD, GS = C::GS, G2 = C::G2;
    constexpr int S = C::S, RPT = C::RPT, QB = C::QB, KP = C::KP, BP = C::BP;
    __shared__ R g_lds[C::LDS_G];
    // per-slot geometry of LDS classes: A(3) AB(3) C(3) CD(3) norm, p, q, active
    __shared__ double slot[C::REG ? 1 : QB][16];

    // ---- tile coordinates
    int bi, ki;
    tile_coords<BP, KP>(a, a.tile_base + (long)blockIdx.x, bi, ki);
    const int p0 = bi * BP, q0 = ki * KP;

    const int t = threadIdx.x;
    const int qs = t / S, sub = t % S;
    bool active = false;
    int p = 0, q = 0, ish = 0, jsh = 0, ksh = 0, lsh = 0;
    if (qs < QB) {
        p = p0 + qs / KP;
        q = q0 + qs % KP;
        if (p < a.bra.npairs && q < a.ket.npairs && (!a.same || q <= p)) {
            ish = a.bra.ish[p]; jsh = a.bra.jsh[p];
            ksh = a.ket.ish[q]; lsh = a.ket.jsh[q];
            double bound = a.bra.q[p] * a.ket.q[q];
            if (a.dm_cond) {
                const double *dc = a.dm_cond;
                const int nb = a.nbas;
                double d = fmax(fmax(dc[ish * nb + jsh], dc[ksh * nb + lsh]),
                                fmax(fmax(dc[ish * nb + ksh], dc[ish * nb + lsh]),
                                     fmax(dc[jsh * nb + ksh], dc[jsh * nb + lsh])));
                bound *= d;
            }
            active = bound > a.tol && bound >= a.lo && bound < a.hi;
        }
    }
    {
        const int nact = __syncthreads_count(active && sub == 0);
        if (nact == 0) return;
        if (a.counter && t == 0) atomicAdd(a.counter, (unsigned long long)nact);
    }

    if (!C::REG && qs < QB && sub == 0) {
        double *sl = slot[qs];
        if (active) {
            sl[0] = a.bx[ish]; sl[1] = a.by[ish]; sl[2] = a.bz[ish];
            sl[3] = a.bx[ish] - a.bx[jsh]; sl[4] = a.by[ish] - a.by[jsh]; sl[5] = a.bz[ish] - a.bz[jsh];
            sl[6] = a.bx[ksh]; sl[7] = a.by[ksh]; sl[8] = a.bz[ksh];
            sl[9] = a.bx[ksh] - a.bx[lsh]; sl[10] = a.by[ksh] - a.by[lsh]; sl[11] = a.bz[ksh] - a.bz[lsh];
            sl[12] = (a.same && p == q) ? 0.5 * PI52X2 : PI52X2;
        }
        sl[13] = (double)p; sl[14] = (double)q; sl[15] = active ? 1.0 : 0.0;
    }
    __syncthreads();

    const int npb = a.bra.nprim, npk = a.ket.nprim;
    R acc[RPT][NFCD];
#pragma unroll
    for (int k = 0; k < RPT; ++k)
#pragma unroll
        for (int c = 0; c < NFCD; ++c) acc[k][c] = (R)0;

    if constexpr (C::REG) {
        if (active) {
            const double A[3] = {a.bx[ish], a.by[ish], a.bz[ish]};
            const double AB[3] = {A[0] - a.bx[jsh], A[1] - a.by[jsh], A[2] - a.bz[jsh]};
            const double Cc[3] = {a.bx[ksh], a.by[ksh], a.bz[ksh]};
            const double CD[3] = {Cc[0] - a.bx[lsh], Cc[1] - a.by[lsh], Cc[2] - a.bz[lsh]};
            const double norm = (a.same && p == q) ? 0.5 * PI52X2 : PI52X2;
            for (int ip = 0; ip < npb; ++ip) {
                const int pi_ = p * npb + ip;
                const double aij = a.bra.a[pi_], eij = a.bra.e[pi_];
                const double P[3] = {a.bra.px[pi_], a.bra.py[pi_], a.bra.pz[pi_]};
                for (int kq = 0; kq < npk; ++kq) {
                    const int qi_ = q * npk + kq;
                    const double akl = a.ket.a[qi_], ekl = a.ket.e[qi_];
                    const double Q[3] = {a.ket.px[qi_], a.ket.py[qi_], a.ket.pz[qi_]};
                    const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                    const double aijkl = aij + akl;
                    const double a0 = aij * akl / aijkl;
                    const double T = a0 * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                    const double pref = norm * eij * ekl / (aij * akl * sqrt(aijkl));
                    const double inv2 = 0.5 / aijkl;
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
                        R x, w;
                        rys_root<R, NR>(T, r, x, w);
                        const R b00 = x * (R)inv2;
                        const R b10 = ((R)0.5 - b00 * (R)akl) / (R)aij;
                        const R b01 = ((R)0.5 - b00 * (R)aij) / (R)akl;
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
    } else {
        // row offsets of this thread
        int ox[RPT], oy[RPT], oz[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            int row = sub + k * S;
            if (row >= NFAB) row = 0;
            const int ia = row / NFB, jb = row % NFB;
            ox[k] = (cart_x(LA, ia) * (LB + 1) + cart_x(LB, jb)) * G2;
            oy[k] = (cart_y(LA, ia) * (LB + 1) + cart_y(LB, jb)) * G2;
            oz[k] = (cart_z(LA, ia) * (LB + 1) + cart_z(LB, jb)) * G2;
        }
        constexpr int NITEM = QB * NR * 3;
        constexpr int NT = C::NT;
        for (int ip = 0; ip < npb; ++ip) {
            for (int kq = 0; kq < npk; ++kq) {
                __syncthreads();
                for (int it = t; it < NITEM; it += NT) {
                    const int qs2 = it / (NR * 3);
                    const int rd = it % (NR * 3);
                    const int r = rd / 3, d = rd % 3;
                    const double *sl = slot[qs2];
                    if (sl[15] == 0.0) continue;
                    const int p2 = (int)sl[13], q2 = (int)sl[14];
                    const int pi_ = p2 * npb + ip, qi_ = q2 * npk + kq;
                    const double aij = a.bra.a[pi_], akl = a.ket.a[qi_];
                    const double Px = a.bra.px[pi_], Py = a.bra.py[pi_], Pz = a.bra.pz[pi_];
                    const double Qx = a.ket.px[qi_], Qy = a.ket.py[qi_], Qz = a.ket.pz[qi_];
                    const double PQx = Px - Qx, PQy = Py - Qy, PQz = Pz - Qz;
                    const double aijkl = aij + akl;
                    const double a0 = aij * akl / aijkl;
                    const double T = a0 * (PQx * PQx + PQy * PQy + PQz * PQz);
                    R x, w;
                    rys_root<R, NR>(T, r, x, w);
                    const double inv2 = 0.5 / aijkl;
                    const R b00 = x * (R)inv2;
                    const R b10 = ((R)0.5 - b00 * (R)akl) / (R)aij;
                    const R b01 = ((R)0.5 - b00 * (R)aij) / (R)akl;
                    const R tb = (R)2 * b00;
                    const double Pd = d == 0 ? Px : (d == 1 ? Py : Pz);
                    const double Qd = d == 0 ? Qx : (d == 1 ? Qy : Qz);
                    const double PQd = d == 0 ? PQx : (d == 1 ? PQy : PQz);
                    const R c00 = (R)(Pd - sl[d]) - tb * (R)akl * (R)PQd;
                    const R c0p = (R)(Qd - sl[6 + d]) + tb * (R)aij * (R)PQd;
                    R g0 = (R)1;
                    if (d == 2) {
                        const double pref = sl[12] * a.bra.e[pi_] * a.ket.e[qi_] / (aij * akl * sqrt(aijkl));
                        g0 = w * (R)pref;
                    }
                    R *out = g_lds + ((size_t)(qs2 * NR + r) * 3 + d) * GS;
                    rys2d<R, LA, LB, LC, LD>(g0, c00, c0p, b00, b10, b01, (R)sl[3 + d], (R)sl[9 + d], out);
                }
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
                        const R *gx = g_lds + ((size_t)(qs * NR + r) * 3) * GS;
                        const R *gy = gx + GS, *gz = gx + 2 * GS;
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
            }
        }
    }

    if (!active) return;

    // ---- digestion
    const int nao = a.nao;
    const int i0 = a.ao_loc[ish], j0 = a.ao_loc[jsh], k0 = a.ao_loc[ksh], l0 = a.ao_loc[lsh];
    const size_t nn = (size_t)nao * nao;
    for (int n = 0; n < a.n_dm; ++n) {
        const double *D = a.dm + n * nn;
        if (a.vj) {
            double *V = a.vj + n * nn;
            double dkl[NFCD], jkl[NFCD];
#pragma unroll
            for (int c = 0; c < NFCD; ++c) {
                dkl[c] = D[(size_t)(k0 + c / NFD) * nao + l0 + c % NFD];
                jkl[c] = 0.0;
            }
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int row = sub + k * S;
                if (row < NFAB) {
                    const int ia = row / NFB, jb = row % NFB;
                    const double dij = D[(size_t)(i0 + ia) * nao + j0 + jb];
                    double sij = 0.0;
#pragma unroll
                    for (int c = 0; c < NFCD; ++c) {
                        const double v = (double)acc[k][c];
                        sij += v * dkl[c];
                        jkl[c] += v * dij;
                    }
                    unsafeAtomicAdd(V + (size_t)(i0 + ia) * nao + j0 + jb, sij);
                }
            }
#pragma unroll
            for (int c = 0; c < NFCD; ++c)
                unsafeAtomicAdd(V + (size_t)(k0 + c / NFD) * nao + l0 + c % NFD, jkl[c]);
        }
        if (a.vk) {
            double *V = a.vk + n * nn;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int row = sub + k * S;
                if (row >= NFAB) continue;
                const int ia = row / NFB, jb = row % NFB;
                const size_t ri = (size_t)(i0 + ia) * nao, rj = (size_t)(j0 + jb) * nao;
                double dik[NFC], dil[NFD], djk[NFC], djl[NFD];
#pragma unroll
                for (int x = 0; x < NFC; ++x) { dik[x] = D[ri + k0 + x]; djk[x] = D[rj + k0 + x]; }
#pragma unroll
                for (int x = 0; x < NFD; ++x) { dil[x] = D[ri + l0 + x]; djl[x] = D[rj + l0 + x]; }
                double vik[NFC], vjk[NFC], vil[NFD], vjl[NFD];
#pragma unroll
                for (int x = 0; x < NFC; ++x) { vik[x] = 0.0; vjk[x] = 0.0; }
#pragma unroll
                for (int x = 0; x < NFD; ++x) { vil[x] = 0.0; vjl[x] = 0.0; }
#pragma unroll
                for (int c = 0; c < NFCD; ++c) {
                    const int kc = c / NFD, ld = c % NFD;
                    const double v = (double)acc[k][c];
                    vik[kc] += v * djl[ld];
                    vil[ld] += v * djk[kc];
                    vjk[kc] += v * dil[ld];
                    vjl[ld] += v * dik[kc];
                }
#pragma unroll
                for (int x = 0; x < NFC; ++x) {
                    unsafeAtomicAdd(V + ri + k0 + x, vik[x]);
                    unsafeAtomicAdd(V + rj + k0 + x, vjk[x]);
                }
#pragma unroll
                for (int x = 0; x < NFD; ++x) {
                    unsafeAtomicAdd(V + ri + l0 + x, vil[x]);
                    unsafeAtomicAdd(V + rj + l0 + x, vjl[x]);
                }
            }
        }
    }
}

}  // namespace mjk
