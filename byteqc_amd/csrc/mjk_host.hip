// Host side of libmi355jk.so: the C ABI declared in include/mi355jk.h.
//
// Ownership: every device buffer a cache needs is allocated here and freed by
// mjk_del_basis_prod; input host tables are copied, nothing is borrowed.  Errors are
// reported by return code plus a thread-local message; nothing here exits the process.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <vector>

#include "../../include/mi355jk.h"
#include "mjk_internal.h"

using namespace mjk;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *where)
{
    return fail(e == hipErrorOutOfMemory ? MJK_ERR_OOM : MJK_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

}  // namespace

int mjk::report_error(int code, const std::string &msg) { return fail(code, msg); }

namespace {

#define HIP_TRY(call)                                           \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return hip_fail(e_, #call);       \
    } while (0)

constexpr int ATM_SLOTS = 6, BAS_SLOTS = 8;
constexpr int ATOM_OF = 0, ANG_OF = 1, NPRIM_OF = 2, NCTR_OF = 3, PTR_EXP = 5, PTR_COEFF = 6;
constexpr int PTR_COORD = 1;

double fac_sp_host(int l) { return l == 0 ? 0.282094791773878143 : (l == 1 ? 0.488602511902919921 : 1.0); }

// A list of device allocations released together.
// Allocations left before an injected failure (mjk_fail_alloc_after; -1: none injected).
std::atomic<long> g_alloc_budget{-1};

struct DevPool {
    std::vector<void *> ptrs;
    template <typename T>
    hipError_t alloc(T **p, size_t n)
    {
        *p = nullptr;
        if (n == 0) n = 1;
        for (long b = g_alloc_budget.load(); b >= 0;) {
            if (b == 0) return hipErrorOutOfMemory;
            if (g_alloc_budget.compare_exchange_weak(b, b - 1)) break;
        }
        hipError_t e = hipMalloc((void **)p, n * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
    template <typename T>
    hipError_t upload(T **p, const T *h, size_t n)
    {
        hipError_t e = alloc(p, n);
        if (e != hipSuccess || n == 0) return e;
        return hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice);
    }
    void release()
    {
        for (void *p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

struct ClassType {
    int l_bra, l_ket, nprim_bra, nprim_ket, nprim12, npairs;
    int pair_off, prim_off;
};

}  // namespace

struct mjk_basis_prod {
    int device = 0;
    int nbas = 0, nao = 0, ncptype = 0, n_pairs = 0, n_prim_pairs = 0;
    std::vector<ClassType> cp;
    DevPool pool;
    double *bx = nullptr, *by = nullptr, *bz = nullptr;
    int *ao_loc = nullptr;
    int *pair_i = nullptr, *pair_j = nullptr;
    double *pair_q = nullptr;
    double *a12 = nullptr, *e12 = nullptr, *x12 = nullptr, *y12 = nullptr, *z12 = nullptr;
    int *pair_aoi = nullptr, *pair_aoj = nullptr;
    double *pair_geo = nullptr;  // [6][n_pairs]: x_i, y_i, z_i, (R_i - R_j)_x,y,z
    int4 *pair_int = nullptr;    // [n_pairs] {ish, jsh, ao_i, ao_j}
    double *pair_dbl = nullptr;  // [n_pairs][8] {q, x_i, y_i, z_i, AB_x, AB_y, AB_z, 0}
    double *prim_rec = nullptr;  // [n_prim_pairs][6] {a, e, Px, Py, Pz, 1/a}
    double *dm_cond = nullptr;            // nbas * nbas workspace
    unsigned long long *dm_max = nullptr;  // scalar (double bits)
    unsigned long long *counters = nullptr;  // [0] fp64 quartets, [1] fp32 quartets, [2] quartet slots
    long last_launches = 0;
    // class pairs of one build run on `lanes` streams: the caller's and lanes - 1 of these
    int lanes = 4;
    int kc_cap = 0;  // longest ket chunk per tile (0: each kernel's default)
    hipStream_t side[MJK_MAX_LANES - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[MJK_MAX_LANES - 1] = {};
    int n_side = 0;  // side streams created so far
};

extern "C" {

int mjk_version(void) { return MJK_VERSION; }

const char *mjk_last_error(void) { return g_last_error.c_str(); }

int mjk_nao(const mjk_basis_prod *bp) { return bp ? bp->nao : -1; }

int mjk_q_cond(const int32_t *atm, int natm, const int32_t *bas, int nbas, const double *env,
               double *q)
{
    if (!atm || !bas || !env || !q || nbas <= 0 || natm <= 0) return fail(MJK_ERR_ARG, "mjk_q_cond: bad argument");
    std::vector<double> x(nbas), y(nbas), z(nbas), exps, coefs;
    std::vector<int> nprim(nbas), poff(nbas), ls(nbas);
    for (int i = 0; i < nbas; ++i) {
        const int32_t *b = bas + i * BAS_SLOTS;
        if (b[NCTR_OF] != 1) return fail(MJK_ERR_ARG, "mjk_q_cond: basis must be segmented (nctr == 1)");
        if (b[ANG_OF] > MJK_LMAX) return fail(MJK_ERR_UNSUPPORTED, "mjk_q_cond: l > MJK_LMAX");
        if (b[ATOM_OF] < 0 || b[ATOM_OF] >= natm) return fail(MJK_ERR_ARG, "mjk_q_cond: bad atom index");
        const double *r = env + atm[b[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        x[i] = r[0]; y[i] = r[1]; z[i] = r[2];
        ls[i] = b[ANG_OF];
        nprim[i] = b[NPRIM_OF];
        poff[i] = (int)exps.size();
        for (int p = 0; p < b[NPRIM_OF]; ++p) {
            exps.push_back(env[b[PTR_EXP] + p]);
            coefs.push_back(env[b[PTR_COEFF] + p]);
        }
    }
    // pairs grouped by canonical (l_big, l_small)
    std::vector<int> pi[4][4], pj[4][4];
    for (int i = 0; i < nbas; ++i)
        for (int j = 0; j <= i; ++j) {
            int a = i, b = j;
            if (ls[a] < ls[b]) std::swap(a, b);
            pi[ls[a]][ls[b]].push_back(a);
            pj[ls[a]][ls[b]].push_back(b);
        }
    DevPool pool;
    ShellTab s{};
    double *dx, *dy, *dz, *dexp, *dcoef, *dq;
    int *dnp, *dpo;
    hipError_t e;
#define UP(ptr, vec) if ((e = pool.upload(&ptr, vec.data(), vec.size())) != hipSuccess) { pool.release(); return hip_fail(e, "mjk_q_cond upload"); }
    UP(dx, x) UP(dy, y) UP(dz, z) UP(dexp, exps) UP(dcoef, coefs) UP(dnp, nprim) UP(dpo, poff)
    if ((e = pool.alloc(&dq, (size_t)nbas * nbas)) != hipSuccess) { pool.release(); return hip_fail(e, "mjk_q_cond alloc"); }
    s.x = dx; s.y = dy; s.z = dz; s.nprim = dnp; s.poff = dpo; s.exps = dexp; s.coefs = dcoef; s.nbas = nbas;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b <= a; ++b) {
            int n = (int)pi[a][b].size();
            if (!n) continue;
            int *dpi, *dpj;
            UP(dpi, pi[a][b]) UP(dpj, pj[a][b])
            if ((e = launch_qcond(a, b, s, dpi, dpj, n, dq, 0)) != hipSuccess) { pool.release(); return hip_fail(e, "mjk_q_cond launch"); }
        }
#undef UP
    e = hipMemcpy(q, dq, sizeof(double) * nbas * nbas, hipMemcpyDeviceToHost);
    pool.release();
    if (e != hipSuccess) return hip_fail(e, "mjk_q_cond copy");
    return MJK_OK;
}

int mjk_init_basis_prod(mjk_basis_prod **pbp, double diag_fac, const int32_t *bas_pair2shls,
                        const int32_t *bas_pairs_locs, int ncptype, const double *pair_q,
                        const int32_t *atm, int natm, const int32_t *bas, int nbas,
                        const double *env)
{
    if (!pbp || !bas_pair2shls || !bas_pairs_locs || !pair_q || !atm || !bas || !env || ncptype < 0 || nbas <= 0)
        return fail(MJK_ERR_ARG, "mjk_init_basis_prod: bad argument");
    *pbp = nullptr;
    const int n_pairs = bas_pairs_locs[ncptype];
    const int32_t *p2i = bas_pair2shls, *p2j = bas_pair2shls + n_pairs;
    std::vector<int> ao_loc(nbas + 1, 0);
    std::vector<double> x(nbas), y(nbas), z(nbas);
    for (int i = 0; i < nbas; ++i) {
        const int32_t *b = bas + i * BAS_SLOTS;
        if (b[NCTR_OF] != 1) return fail(MJK_ERR_ARG, "mjk_init_basis_prod: basis must be segmented (nctr == 1)");
        if (b[ANG_OF] > MJK_LMAX || b[ANG_OF] < 0)
            return fail(MJK_ERR_UNSUPPORTED, "mjk_init_basis_prod: angular momentum above MJK_LMAX");
        if (b[ATOM_OF] < 0 || b[ATOM_OF] >= natm) return fail(MJK_ERR_ARG, "mjk_init_basis_prod: bad atom index");
        const double *r = env + atm[b[ATOM_OF] * ATM_SLOTS + PTR_COORD];
        x[i] = r[0]; y[i] = r[1]; z[i] = r[2];
        const int l = b[ANG_OF];
        ao_loc[i + 1] = ao_loc[i] + (l + 1) * (l + 2) / 2;
    }
    auto *bp = new (std::nothrow) mjk_basis_prod();
    if (!bp) return fail(MJK_ERR_OOM, "mjk_init_basis_prod: host allocation failed");
    (void)hipGetDevice(&bp->device);
    bp->nbas = nbas;
    bp->nao = ao_loc[nbas];
    bp->ncptype = ncptype;
    bp->n_pairs = n_pairs;
    std::vector<double> a12, e12, x12, y12, z12, geo(6 * (size_t)n_pairs);
    std::vector<int> pi(n_pairs), pj(n_pairs), paoi(n_pairs), paoj(n_pairs);
    std::vector<int4> pint(n_pairs);
    std::vector<double> pdbl(8 * (size_t)n_pairs, 0.0), prec;
    int prim_total = 0;
    for (int c = 0; c < ncptype; ++c) {
        ClassType ct{};
        ct.pair_off = bas_pairs_locs[c];
        ct.npairs = bas_pairs_locs[c + 1] - bas_pairs_locs[c];
        ct.prim_off = prim_total;
        if (ct.npairs > 0) {
            const int i0 = p2i[ct.pair_off], j0 = p2j[ct.pair_off];
            ct.l_bra = bas[i0 * BAS_SLOTS + ANG_OF];
            ct.l_ket = bas[j0 * BAS_SLOTS + ANG_OF];
            ct.nprim_bra = bas[i0 * BAS_SLOTS + NPRIM_OF];
            ct.nprim_ket = bas[j0 * BAS_SLOTS + NPRIM_OF];
            if (ct.l_bra < ct.l_ket) {
                delete bp;
                return fail(MJK_ERR_ARG, "mjk_init_basis_prod: pair classes need l_i >= l_j");
            }
        }
        ct.nprim12 = ct.nprim_bra * ct.nprim_ket;
        for (int pr = ct.pair_off; pr < ct.pair_off + ct.npairs; ++pr) {
            const int i = p2i[pr], j = p2j[pr];
            if (i < 0 || i >= nbas || j < 0 || j >= nbas) {
                delete bp;
                return fail(MJK_ERR_ARG, "mjk_init_basis_prod: shell index out of range");
            }
            const int32_t *bi = bas + i * BAS_SLOTS, *bj = bas + j * BAS_SLOTS;
            if (bi[ANG_OF] != ct.l_bra || bj[ANG_OF] != ct.l_ket || bi[NPRIM_OF] != ct.nprim_bra ||
                bj[NPRIM_OF] != ct.nprim_ket) {
                delete bp;
                return fail(MJK_ERR_ARG, "mjk_init_basis_prod: pair class mixes (l, nprim) patterns");
            }
            pi[pr] = i;
            pj[pr] = j;
            paoi[pr] = ao_loc[i];
            paoj[pr] = ao_loc[j];
            geo[0 * (size_t)n_pairs + pr] = x[i];
            geo[1 * (size_t)n_pairs + pr] = y[i];
            geo[2 * (size_t)n_pairs + pr] = z[i];
            geo[3 * (size_t)n_pairs + pr] = x[i] - x[j];
            geo[4 * (size_t)n_pairs + pr] = y[i] - y[j];
            geo[5 * (size_t)n_pairs + pr] = z[i] - z[j];
            pint[pr] = make_int4(i, j, ao_loc[i], ao_loc[j]);
            double *pd = &pdbl[8 * (size_t)pr];
            pd[0] = pair_q[pr];
            pd[1] = x[i]; pd[2] = y[i]; pd[3] = z[i];
            pd[4] = x[i] - x[j]; pd[5] = y[i] - y[j]; pd[6] = z[i] - z[j];
            const double *ai = env + bi[PTR_EXP], *aj = env + bj[PTR_EXP];
            const double *ci = env + bi[PTR_COEFF], *cj = env + bj[PTR_COEFF];
            const double rx = x[i] - x[j], ry = y[i] - y[j], rz = z[i] - z[j];
            const double rr = rx * rx + ry * ry + rz * rz;
            const double norm = fac_sp_host(ct.l_bra) * fac_sp_host(ct.l_ket) * (i == j ? diag_fac : 1.0);
            for (int ip = 0; ip < ct.nprim_bra; ++ip)
                for (int jp = 0; jp < ct.nprim_ket; ++jp) {
                    const double aij = ai[ip] + aj[jp];
                    a12.push_back(aij);
                    e12.push_back(norm * ci[ip] * cj[jp] * exp(-rr * ai[ip] * aj[jp] / aij));
                    x12.push_back((ai[ip] * x[i] + aj[jp] * x[j]) / aij);
                    y12.push_back((ai[ip] * y[i] + aj[jp] * y[j]) / aij);
                    z12.push_back((ai[ip] * z[i] + aj[jp] * z[j]) / aij);
                    prec.insert(prec.end(), {aij, e12.back(), x12.back(), y12.back(), z12.back(), 1.0 / aij});
                }
        }
        prim_total += ct.npairs * ct.nprim12;
        bp->cp.push_back(ct);
    }
    bp->n_prim_pairs = prim_total;
    hipError_t e;
#define UP(ptr, vec) if ((e = bp->pool.upload(&bp->ptr, vec.data(), vec.size())) != hipSuccess) goto oom;
    UP(bx, x) UP(by, y) UP(bz, z) UP(ao_loc, ao_loc) UP(pair_i, pi) UP(pair_j, pj)
    UP(a12, a12) UP(e12, e12) UP(x12, x12) UP(y12, y12) UP(z12, z12)
    UP(pair_aoi, paoi) UP(pair_aoj, paoj) UP(pair_geo, geo)
    UP(pair_int, pint) UP(pair_dbl, pdbl) UP(prim_rec, prec)
#undef UP
    if ((e = bp->pool.upload(&bp->pair_q, pair_q, (size_t)n_pairs)) != hipSuccess) goto oom;
    if ((e = bp->pool.alloc(&bp->dm_max, 1)) != hipSuccess) goto oom;
    if ((e = bp->pool.alloc(&bp->counters, 4)) != hipSuccess) goto oom;
    if ((e = hipMemset(bp->counters, 0, 4 * sizeof(unsigned long long))) != hipSuccess) goto oom;
    *pbp = bp;
    return MJK_OK;
oom:
    bp->pool.release();
    delete bp;
    return fail(e == hipErrorOutOfMemory ? MJK_ERR_OOM : MJK_ERR_HIP,
                std::string("mjk_init_basis_prod: ") + hipGetErrorString(e));
}

int mjk_del_basis_prod(mjk_basis_prod **pbp)
{
    if (!pbp || !*pbp) return MJK_OK;
    mjk_basis_prod *bp = *pbp;
    for (int i = 0; i < bp->n_side; ++i) {
        (void)hipStreamDestroy(bp->side[i]);
        (void)hipEventDestroy(bp->ev_join[i]);
    }
    if (bp->ev_fork) (void)hipEventDestroy(bp->ev_fork);
    (*pbp)->pool.release();
    delete *pbp;
    *pbp = nullptr;
    return MJK_OK;
}

}  // extern "C"

namespace {

PairSet pair_set(const mjk_basis_prod *bp, const ClassType &ct)
{
    PairSet s{};
    s.ish = bp->pair_i + ct.pair_off;
    s.jsh = bp->pair_j + ct.pair_off;
    s.q = bp->pair_q + ct.pair_off;
    s.a = bp->a12 + ct.prim_off;
    s.e = bp->e12 + ct.prim_off;
    s.px = bp->x12 + ct.prim_off;
    s.py = bp->y12 + ct.prim_off;
    s.pz = bp->z12 + ct.prim_off;
    s.aoi = bp->pair_aoi + ct.pair_off;
    s.aoj = bp->pair_aoj + ct.pair_off;
    const size_t np = bp->n_pairs;
    s.cx = bp->pair_geo + 0 * np + ct.pair_off;
    s.cy = bp->pair_geo + 1 * np + ct.pair_off;
    s.cz = bp->pair_geo + 2 * np + ct.pair_off;
    s.abx = bp->pair_geo + 3 * np + ct.pair_off;
    s.aby = bp->pair_geo + 4 * np + ct.pair_off;
    s.abz = bp->pair_geo + 5 * np + ct.pair_off;
    s.pint = bp->pair_int + ct.pair_off;
    s.pdbl = bp->pair_dbl + 8 * (size_t)ct.pair_off;
    s.prim = bp->prim_rec + 6 * (size_t)ct.prim_off;
    s.npairs = ct.npairs;
    s.nprim = ct.nprim12;
    return s;
}

int ncart_h(int l) { return (l + 1) * (l + 2) / 2; }

// Classes (canonical orientation) whose FP32 tiles ran faster than their FP64 tiles on
// MI355X ((H2O)64/def2-TZVP, per-tile split at a 1e-7 bound).  Without packed math, CDNA4
// issues FP32 VALU work at the FP64 rate, so FP32 only pays where halved registers or LDS
// raise occupancy; elsewhere the conversions cost more and the class stays FP64, which is
// also the more accurate choice.
bool fp32_pays(int la, int lb, int lc, int ld)
{
    static const int table[][4] = {{3, 2, 1, 1}, {3, 1, 2, 0}, {3, 2, 2, 2}, {3, 1, 1, 1}, {2, 2, 2, 0},
                                   {3, 1, 1, 0}, {3, 0, 0, 0}, {3, 2, 2, 0}, {2, 2, 2, 2}, {2, 1, 1, 1},
                                   {3, 2, 3, 0}};
    for (const auto &c : table)
        if (c[0] == la && c[1] == lb && c[2] == lc && c[3] == ld) return true;
    return false;
}

int run_class_pair(mjk_basis_prod *bp, double *vj, double *vk, const double *dm, int n_dm,
                   int cp_ij, int cp_kl, const mjk_screen &sc, bool dens, hipStream_t st)
{
    const ClassType &c1 = bp->cp[cp_ij], &c2 = bp->cp[cp_kl];
    if (c1.npairs == 0 || c2.npairs == 0) return MJK_OK;
    const ClassType *bra = &c1, *ket = &c2;
    if (ncart_h(c1.l_bra) * ncart_h(c1.l_ket) < ncart_h(c2.l_bra) * ncart_h(c2.l_ket)) std::swap(bra, ket);
    JKArgs a{};
    a.bra = pair_set(bp, *bra);
    a.ket = pair_set(bp, *ket);
    a.bx = bp->bx; a.by = bp->by; a.bz = bp->bz;
    a.ao_loc = bp->ao_loc;
    a.dm = dm; a.vj = vj; a.vk = vk;
    a.dm_cond = dens ? bp->dm_cond : nullptr;
    a.dm_max = dens ? (const double *)bp->dm_max : nullptr;
    a.tol = sc.direct_scf_tol;
    // a primitive quartet whose prefactor is 1e-6 below the quartet cutoff cannot matter
    a.prim_tol = sc.direct_scf_tol * 1e-6;
    a.nbas = bp->nbas; a.nao = bp->nao; a.n_dm = n_dm;
    a.same = cp_ij == cp_kl;
    a.kc_cap = bp->kc_cap;
    const double pc = fp32_pays(bra->l_bra, bra->l_ket, ket->l_bra, ket->l_ket) ? sc.precision_cutoff : 0.0;
    long nt = 0;
    // FP64 part: bound >= precision_cutoff (everything when the cutoff is 0)
    a.lo = pc > 0 ? pc : 0.0;
    a.hi = INFINITY;
    a.counter = bp->counters;
    a.slots = bp->counters + 2;
    hipError_t e = launch_class(bra->l_bra, bra->l_ket, ket->l_bra, ket->l_ket, false, a, st, &nt);
    if (e != hipSuccess)
        return fail(MJK_ERR_HIP, std::string("mjk_build_jk: launch of (") + "spdf"[bra->l_bra] + "spdf"[bra->l_ket] + "|" +
                                 "spdf"[ket->l_bra] + "spdf"[ket->l_ket] + ") failed: " + hipGetErrorString(e));
    if (pc > 0) {
        a.lo = 0.0;
        a.hi = pc;
        a.counter = bp->counters + 1;
        e = launch_class(bra->l_bra, bra->l_ket, ket->l_bra, ket->l_ket, true, a, st, &nt);
        if (e != hipSuccess) return hip_fail(e, "mjk_build_jk: FP32 launch");
    }
    bp->last_launches += nt > 0 ? 1 + (pc > 0) : 0;
    return MJK_OK;
}

int check_common(mjk_basis_prod *bp, const double *dm, int nao, int n_dm, const mjk_screen *screen)
{
    if (!bp) return fail(MJK_ERR_ARG, "null basis-product cache");
    if (!dm) return fail(MJK_ERR_ARG, "null density");
    if (!screen) return fail(MJK_ERR_ARG, "null screen options");
    if (nao != bp->nao) return fail(MJK_ERR_ARG, "nao does not match the cache");
    if (n_dm <= 0) return fail(MJK_ERR_ARG, "n_dm must be positive");
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != bp->device) return fail(MJK_ERR_ARG, "current device differs from the cache's device");
    return MJK_OK;
}

// The nbas x nbas block maxima are allocated on the first density-screened build (a
// concatenated batch basis never needs them).
int ensure_dm_cond(mjk_basis_prod *bp)
{
    if (bp->dm_cond) return MJK_OK;
    hipError_t e = bp->pool.alloc(&bp->dm_cond, (size_t)bp->nbas * bp->nbas);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? MJK_ERR_OOM : MJK_ERR_HIP,
                    std::string("density screen workspace: ") + hipGetErrorString(e));
    return MJK_OK;
}

// Side streams (default priority, not synchronising with the null stream) and their
// events, created on the cache's device the first time a build asks for them.
int ensure_lanes(mjk_basis_prod *bp, int lanes)
{
    while (bp->n_side < lanes - 1) {
        const int i = bp->n_side;
        hipError_t e = hipStreamCreateWithFlags(&bp->side[i], hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(e, "side stream");
        e = hipEventCreateWithFlags(&bp->ev_join[i], hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipStreamDestroy(bp->side[i]);
            return hip_fail(e, "side stream event");
        }
        bp->n_side = i + 1;
    }
    if (lanes > 1 && !bp->ev_fork) {
        hipError_t e = hipEventCreateWithFlags(&bp->ev_fork, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "fork event");
    }
    return MJK_OK;
}

}  // namespace

extern "C" {

int mjk_fail_alloc_after(int n)
{
    g_alloc_budget.store(n < 0 ? -1 : (long)n);
    return MJK_OK;
}

int mjk_set_streams(mjk_basis_prod *bp, int lanes)
{
    if (!bp) return fail(MJK_ERR_ARG, "mjk_set_streams: null basis-product cache");
    if (lanes < 1 || lanes > MJK_MAX_LANES) return fail(MJK_ERR_ARG, "mjk_set_streams: streams out of range");
    bp->lanes = lanes;
    return MJK_OK;
}

int mjk_set_ket_chunk(mjk_basis_prod *bp, int cap)
{
    if (!bp) return fail(MJK_ERR_ARG, "mjk_set_ket_chunk: null basis-product cache");
    if (cap < 0 || cap > 4096) return fail(MJK_ERR_ARG, "mjk_set_ket_chunk: chunk out of range");
    bp->kc_cap = cap;
    return MJK_OK;
}

int mjk_build_jk(mjk_basis_prod *bp, double *vj, double *vk, const double *dm, int nao, int n_dm,
                 int cp_ij, int cp_kl, const mjk_screen *screen, void *stream)
{
    int rc = check_common(bp, dm, nao, n_dm, screen);
    if (rc) return rc;
    if (cp_ij < 0 || cp_kl < 0 || cp_ij >= bp->ncptype || cp_kl >= bp->ncptype || cp_kl > cp_ij)
        return fail(MJK_ERR_ARG, "mjk_build_jk: class indices must satisfy 0 <= cp_kl <= cp_ij < ncptype");
    if (!vj && !vk) return MJK_OK;
    hipStream_t st = (hipStream_t)stream;
    const bool dens = screen->density_screen != 0;
    if (screen->density_screen == 2 && !bp->dm_cond)
        return fail(MJK_ERR_ARG, "mjk_build_jk: density_screen 2 needs an earlier density_screen 1 call");
    if (screen->density_screen == 1 && (rc = ensure_dm_cond(bp)) != MJK_OK) return rc;
    if (screen->density_screen == 1) {
        hipError_t e = launch_dm_cond(dm, n_dm, nao, bp->ao_loc, bp->nbas, bp->dm_cond, bp->dm_max, st);
        if (e != hipSuccess) return hip_fail(e, "mjk_build_jk: density screen");
    }
    return run_class_pair(bp, vj, vk, dm, n_dm, cp_ij, cp_kl, *screen, dens, st);
}

int mjk_build_jk_all(mjk_basis_prod *bp, double *vj, double *vk, const double *dm, int nao, int n_dm,
                     const mjk_screen *screen, void *stream)
{
    int rc = check_common(bp, dm, nao, n_dm, screen);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)n_dm * nao * nao;
    hipError_t e;
    if (vj && (e = launch_fill(vj, n, 0.0, st)) != hipSuccess) return hip_fail(e, "mjk_build_jk_all: zero vj");
    if (vk && (e = launch_fill(vk, n, 0.0, st)) != hipSuccess) return hip_fail(e, "mjk_build_jk_all: zero vk");
    if ((e = launch_fill_u64(bp->counters, 4, st)) != hipSuccess) return hip_fail(e, "mjk_build_jk_all: counters");
    bp->last_launches = 0;
    if (!vj && !vk) return MJK_OK;
    const bool dens = screen->density_screen != 0;
    if (dens && (rc = ensure_dm_cond(bp)) != MJK_OK) return rc;
    if (dens && (e = launch_dm_cond(dm, n_dm, nao, bp->ao_loc, bp->nbas, bp->dm_cond, bp->dm_max, st)) != hipSuccess)
        return hip_fail(e, "mjk_build_jk_all: density screen");
    // heavy classes first
    std::vector<std::pair<double, std::pair<int, int>>> order;
    for (int i = 0; i < bp->ncptype; ++i)
        for (int j = 0; j <= i; ++j) {
            const ClassType &a = bp->cp[i], &b = bp->cp[j];
            const double w = (double)a.npairs * b.npairs * a.nprim12 * b.nprim12 *
                             ncart_h(a.l_bra) * ncart_h(a.l_ket) * ncart_h(b.l_bra) * ncart_h(b.l_ket);
            order.push_back({-w, {i, j}});
        }
    std::sort(order.begin(), order.end());
    // The class pairs are spread over several streams.  A launch lasts as long as its
    // longest chain of ket iterations; when density screening leaves few quartets (the
    // incremental builds of an SCF) that tail, not the work, sets its duration, and
    // concurrent launches overlap the tails.  Each class pair goes to the stream with the
    // least modelled work so far; the caller's stream waits for all of them.
    const int lanes = bp->lanes;
    if ((rc = ensure_lanes(bp, lanes)) != MJK_OK) return rc;
    hipStream_t ls[MJK_MAX_LANES];
    double load[MJK_MAX_LANES];
    ls[0] = st;
    load[0] = 0.0;
    for (int k = 1; k < lanes; ++k) {
        ls[k] = bp->side[k - 1];
        load[k] = 0.0;
    }
    if (lanes > 1) {
        if ((e = hipEventRecord(bp->ev_fork, st)) != hipSuccess) return hip_fail(e, "mjk_build_jk_all: fork");
        for (int k = 1; k < lanes; ++k)
            if ((e = hipStreamWaitEvent(ls[k], bp->ev_fork, 0)) != hipSuccess) return hip_fail(e, "mjk_build_jk_all: fork");
    }
    for (auto &o : order) {
        const int k = (int)(std::min_element(load, load + lanes) - load);
        load[k] -= o.first;
        rc = run_class_pair(bp, vj, vk, dm, n_dm, o.second.first, o.second.second, *screen, dens, ls[k]);
        if (rc) break;
    }
    // join every lane, also after a failed launch, so no work is left behind the caller's stream
    for (int k = 1; k < lanes; ++k) {
        if ((e = hipEventRecord(bp->ev_join[k - 1], ls[k])) != hipSuccess) return hip_fail(e, "mjk_build_jk_all: join");
        if ((e = hipStreamWaitEvent(st, bp->ev_join[k - 1], 0)) != hipSuccess) return hip_fail(e, "mjk_build_jk_all: join");
    }
    if (rc) return rc;
    if (vj && (e = launch_symmetrize(vj, n_dm, nao, 2.0, st)) != hipSuccess) return hip_fail(e, "symmetrize vj");
    if (vk && (e = launch_symmetrize(vk, n_dm, nao, 1.0, st)) != hipSuccess) return hip_fail(e, "symmetrize vk");
    return MJK_OK;
}

// ---- batched builds ---------------------------------------------------------------------

}  // extern "C"

namespace {

struct BatchCP {
    int cb = 0, ck = 0;      // bra and ket class (bra has at least as many Cartesian products)
    int l[4] = {0, 0, 0, 0};
    int kind = 0;            // 0 register, 1 group, 2 block-LDS (single molecule only)
    int bra_tile = 1, kc = 0;
    bool same = false;
    double weight = 0.0;     // modelled cost, heavy class pairs launch first
    int4 *d_tiles = nullptr;
    int *d_tmol = nullptr;
    long ntiles = 0;
};

}  // namespace

struct mjk_batch {
    mjk_basis_prod *bp = nullptr;
    int n_mol = 0, max_nao = 0;
    size_t total = 0;
    std::vector<MolRec> mols;
    MolRec *d_mols = nullptr;
    std::vector<int> mol_off;  // [ncptype][n_mol + 1]
    std::vector<BatchCP> cps;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *g_key[3] = {nullptr, nullptr, nullptr};
    double g_tol = 0.0, g_pc = 0.0;
};

namespace {

void batch_drop_graph(mjk_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int default_kc(const BatchCP &c) { return c.kind == 0 ? 64 : 16; }

// Tiles of one class pair over every molecule: bra blocks of c.bra_tile pairs times ket
// chunks of c.kc pairs, both inside one molecule; same-class pairs keep q <= p.
int batch_make_tiles(mjk_batch *b, BatchCP &c)
{
    if (c.d_tiles) (void)hipFree(c.d_tiles);
    if (c.d_tmol) (void)hipFree(c.d_tmol);
    c.d_tiles = nullptr;
    c.d_tmol = nullptr;
    c.ntiles = 0;
    if (c.kind == 2) return MJK_OK;  // computed tiles (single molecule)
    std::vector<int4> t;
    std::vector<int> m;
    const int nm1 = b->n_mol + 1;
    const int *ob = &b->mol_off[(size_t)c.cb * nm1], *ok = &b->mol_off[(size_t)c.ck * nm1];
    for (int mol = 0; mol < b->n_mol; ++mol) {
        const int pb0 = ob[mol], pb1 = ob[mol + 1], pk0 = ok[mol], pk1 = ok[mol + 1];
        if (pb0 == pb1 || pk0 == pk1) continue;
        for (int p0 = pb0; p0 < pb1; p0 += c.bra_tile) {
            const int p1 = std::min(p0 + c.bra_tile, pb1);
            for (int q0 = pk0; q0 < pk1; q0 += c.kc) {
                if (c.same && q0 > p1 - 1) break;
                t.push_back(make_int4(p0, p1, q0, std::min(q0 + c.kc, pk1)));
                m.push_back(mol);
            }
        }
    }
    c.ntiles = (long)t.size();
    if (t.empty()) return MJK_OK;
    hipError_t e;
    if ((e = hipMalloc((void **)&c.d_tiles, t.size() * sizeof(int4))) != hipSuccess ||
        (e = hipMalloc((void **)&c.d_tmol, m.size() * sizeof(int))) != hipSuccess ||
        (e = hipMemcpy(c.d_tiles, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c.d_tmol, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? MJK_ERR_OOM : MJK_ERR_HIP,
                    std::string("batch tiles: ") + hipGetErrorString(e));
    return MJK_OK;
}

int batch_launch_cp(mjk_batch *b, const BatchCP &c, double *vj, double *vk, const double *dm,
                    const mjk_screen &sc, hipStream_t st)
{
    mjk_basis_prod *bp = b->bp;
    const ClassType &bra = bp->cp[c.cb], &ket = bp->cp[c.ck];
    if (bra.npairs == 0 || ket.npairs == 0) return MJK_OK;
    if (c.kind != 2 && c.ntiles == 0) return MJK_OK;
    JKArgs a{};
    a.bra = pair_set(bp, bra);
    a.ket = pair_set(bp, ket);
    a.bx = bp->bx; a.by = bp->by; a.bz = bp->bz;
    a.ao_loc = bp->ao_loc;
    a.dm = dm; a.vj = vj; a.vk = vk;
    a.tol = sc.direct_scf_tol;
    a.prim_tol = sc.direct_scf_tol * 1e-6;
    a.nbas = bp->nbas;
    a.nao = b->mols[0].nao;
    a.n_dm = 1;
    a.same = c.same;
    a.counter = bp->counters;
    a.slots = bp->counters + 2;
    if (c.kind != 2) {
        a.tiles = c.d_tiles;
        a.tile_mol = c.d_tmol;
        a.mols = b->d_mols;
        a.ndesc = c.ntiles;
    }
    const double pc = fp32_pays(c.l[0], c.l[1], c.l[2], c.l[3]) ? sc.precision_cutoff : 0.0;
    a.lo = pc > 0 ? pc : 0.0;
    a.hi = INFINITY;
    long nt = 0;
    hipError_t e = launch_class(c.l[0], c.l[1], c.l[2], c.l[3], false, a, st, &nt);
    if (e != hipSuccess) return hip_fail(e, "mjk_batch_build: class launch");
    if (pc > 0) {
        a.lo = 0.0;
        a.hi = pc;
        a.counter = bp->counters + 1;
        e = launch_class(c.l[0], c.l[1], c.l[2], c.l[3], true, a, st, &nt);
        if (e != hipSuccess) return hip_fail(e, "mjk_batch_build: FP32 class launch");
    }
    return MJK_OK;
}

int batch_enqueue(mjk_batch *b, double *vj, double *vk, const double *dm, const mjk_screen &sc, hipStream_t st)
{
    hipError_t e;
    if (vj && (e = launch_fill(vj, b->total, 0.0, st)) != hipSuccess) return hip_fail(e, "mjk_batch_build: zero vj");
    if (vk && (e = launch_fill(vk, b->total, 0.0, st)) != hipSuccess) return hip_fail(e, "mjk_batch_build: zero vk");
    if ((e = launch_fill_u64(b->bp->counters, 4, st)) != hipSuccess) return hip_fail(e, "mjk_batch_build: counters");
    for (const BatchCP &c : b->cps) {
        int rc = batch_launch_cp(b, c, vj, vk, dm, sc, st);
        if (rc) return rc;
    }
    if (vj && (e = launch_symmetrize_batch(vj, b->d_mols, b->n_mol, b->max_nao, 2.0, st)) != hipSuccess)
        return hip_fail(e, "mjk_batch_build: symmetrize vj");
    if (vk && (e = launch_symmetrize_batch(vk, b->d_mols, b->n_mol, b->max_nao, 1.0, st)) != hipSuccess)
        return hip_fail(e, "mjk_batch_build: symmetrize vk");
    return MJK_OK;
}

}  // namespace

extern "C" {

int mjk_batch_create(mjk_batch **pb, mjk_basis_prod **pbp, int n_mol, const int32_t *mol_pair_off,
                     const int64_t *dm_off, const int32_t *nao, const int32_t *ao_base)
{
    if (!pb || !pbp || !*pbp || n_mol <= 0 || !mol_pair_off || !dm_off || !nao || !ao_base)
        return fail(MJK_ERR_ARG, "mjk_batch_create: bad argument");
    *pb = nullptr;
    mjk_basis_prod *bp = *pbp;
    const int nct = bp->ncptype, nm1 = n_mol + 1;
    for (int c = 0; c < nct; ++c) {
        const int32_t *o = mol_pair_off + (size_t)c * nm1;
        if (o[0] != 0 || o[n_mol] != bp->cp[c].npairs)
            return fail(MJK_ERR_ARG, "mjk_batch_create: molecule pair offsets do not cover class " + std::to_string(c));
        for (int m = 0; m < n_mol; ++m)
            if (o[m + 1] < o[m]) return fail(MJK_ERR_ARG, "mjk_batch_create: molecule pair offsets decrease");
    }
    mjk_batch *b = new mjk_batch();
    b->n_mol = n_mol;
    b->mol_off.assign(mol_pair_off, mol_pair_off + (size_t)nct * nm1);
    for (int m = 0; m < n_mol; ++m) {
        if (nao[m] <= 0 || dm_off[m] < 0 || ao_base[m] < 0) {
            delete b;
            return fail(MJK_ERR_ARG, "mjk_batch_create: bad molecule record " + std::to_string(m));
        }
        b->mols.push_back(MolRec{(long)dm_off[m], nao[m], ao_base[m]});
        b->max_nao = std::max(b->max_nao, (int)nao[m]);
        b->total = std::max(b->total, (size_t)dm_off[m] + (size_t)nao[m] * nao[m]);
    }
    hipError_t e;
    if ((e = hipMalloc((void **)&b->d_mols, n_mol * sizeof(MolRec))) != hipSuccess ||
        (e = hipMemcpy(b->d_mols, b->mols.data(), n_mol * sizeof(MolRec), hipMemcpyHostToDevice)) != hipSuccess) {
        if (b->d_mols) (void)hipFree(b->d_mols);
        delete b;
        return hip_fail(e, "mjk_batch_create");
    }
    b->bp = bp;
    for (int i = 0; i < nct; ++i)
        for (int j = 0; j <= i; ++j) {
            BatchCP c;
            const ClassType *c1 = &bp->cp[i], *c2 = &bp->cp[j];
            c.cb = i;
            c.ck = j;
            if (ncart_h(c1->l_bra) * ncart_h(c1->l_ket) < ncart_h(c2->l_bra) * ncart_h(c2->l_ket)) {
                std::swap(c1, c2);
                std::swap(c.cb, c.ck);
            }
            c.l[0] = c1->l_bra; c.l[1] = c1->l_ket; c.l[2] = c2->l_bra; c.l[3] = c2->l_ket;
            c.same = i == j;
            c.kind = class_kernel_info(c.l[0], c.l[1], c.l[2], c.l[3], &c.bra_tile);
            if (c.kind < 0 || (c.kind == 2 && n_mol > 1)) {
                *pbp = bp;  // ownership stays with the caller
                b->bp = nullptr;
                mjk_batch_destroy(&b);
                return fail(MJK_ERR_UNSUPPORTED, std::string("mjk_batch_create: class (") + "spdf"[c.l[0]] +
                                                     "spdf"[c.l[1]] + "|" + "spdf"[c.l[2]] + "spdf"[c.l[3]] +
                                                     ") has no batched kernel");
            }
            c.kc = default_kc(c);
            c.weight = (double)c1->npairs * c2->npairs * c1->nprim12 * c2->nprim12 * ncart_h(c.l[0]) *
                       ncart_h(c.l[1]) * ncart_h(c.l[2]) * ncart_h(c.l[3]);
            b->cps.push_back(c);
        }
    std::sort(b->cps.begin(), b->cps.end(), [](const BatchCP &x, const BatchCP &y) { return x.weight > y.weight; });
    for (BatchCP &c : b->cps) {
        int rc = batch_make_tiles(b, c);
        if (rc) {
            *pbp = bp;
            b->bp = nullptr;
            mjk_batch_destroy(&b);
            return rc;
        }
    }
    *pbp = nullptr;  // the batch owns the basis-product cache from here on
    *pb = b;
    return MJK_OK;
}

int mjk_batch_destroy(mjk_batch **pb)
{
    if (!pb || !*pb) return MJK_OK;
    mjk_batch *b = *pb;
    batch_drop_graph(b);
    for (BatchCP &c : b->cps) {
        if (c.d_tiles) (void)hipFree(c.d_tiles);
        if (c.d_tmol) (void)hipFree(c.d_tmol);
    }
    if (b->d_mols) (void)hipFree(b->d_mols);
    if (b->bp) mjk_del_basis_prod(&b->bp);
    delete b;
    *pb = nullptr;
    return MJK_OK;
}

int mjk_batch_build(mjk_batch *b, double *vj, double *vk, const double *dm, const mjk_screen *screen,
                    int use_graph, void *stream)
{
    if (!b || !dm || !screen) return fail(MJK_ERR_ARG, "mjk_batch_build: bad argument");
    if (screen->density_screen) return fail(MJK_ERR_ARG, "mjk_batch_build: batched builds screen by Schwarz only");
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != b->bp->device) return fail(MJK_ERR_ARG, "mjk_batch_build: current device differs from the batch's");
    if (!vj && !vk) return MJK_OK;
    hipStream_t st = (hipStream_t)stream;
    if (!use_graph) return batch_enqueue(b, vj, vk, dm, *screen, st);
    if (!st) return fail(MJK_ERR_ARG, "mjk_batch_build: graph capture needs a non-default stream");
    const bool same_key = b->exec && b->g_key[0] == vj && b->g_key[1] == vk && b->g_key[2] == dm &&
                          b->g_tol == screen->direct_scf_tol && b->g_pc == screen->precision_cutoff;
    if (!same_key) {
        batch_drop_graph(b);
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(e, "mjk_batch_build: begin capture");
        int rc = batch_enqueue(b, vj, vk, dm, *screen, st);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "mjk_batch_build: end capture");
        b->graph = g;
        if ((e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0)) != hipSuccess) {
            batch_drop_graph(b);
            return hip_fail(e, "mjk_batch_build: instantiate");
        }
        b->g_key[0] = vj;
        b->g_key[1] = vk;
        b->g_key[2] = dm;
        b->g_tol = screen->direct_scf_tol;
        b->g_pc = screen->precision_cutoff;
    }
    hipError_t e = hipGraphLaunch(b->exec, st);
    if (e != hipSuccess) return hip_fail(e, "mjk_batch_build: graph launch");
    return MJK_OK;
}

int mjk_batch_num_class_pairs(const mjk_batch *b) { return b ? (int)b->cps.size() : 0; }

int mjk_batch_tiles(const mjk_batch *b, int32_t *out, int n)
{
    if (!b || !out || n < (int)b->cps.size()) return fail(MJK_ERR_ARG, "mjk_batch_tiles: bad argument");
    for (size_t i = 0; i < b->cps.size(); ++i) {
        const BatchCP &c = b->cps[i];
        int32_t *o = out + 10 * i;
        o[0] = c.cb; o[1] = c.ck;
        o[2] = c.l[0]; o[3] = c.l[1]; o[4] = c.l[2]; o[5] = c.l[3];
        o[6] = c.kc;
        o[7] = (int32_t)std::min<long>(c.ntiles, 0x7fffffff);
        o[8] = c.kind;
        o[9] = c.bra_tile;
    }
    return MJK_OK;
}

int mjk_batch_set_kc(mjk_batch *b, const int32_t *kc, int n)
{
    if (!b || !kc || n != (int)b->cps.size()) return fail(MJK_ERR_ARG, "mjk_batch_set_kc: bad argument");
    for (int i = 0; i < n; ++i)
        if (kc[i] < 1 || kc[i] > 4096) return fail(MJK_ERR_ARG, "mjk_batch_set_kc: ket chunk out of range");
    batch_drop_graph(b);
    for (int i = 0; i < n; ++i) {
        if (b->cps[i].kc == kc[i]) continue;
        b->cps[i].kc = kc[i];
        int rc = batch_make_tiles(b, b->cps[i]);
        if (rc) return rc;
    }
    return MJK_OK;
}

int mjk_batch_time_class_pair(mjk_batch *b, int i, double *vj, double *vk, const double *dm,
                              const mjk_screen *screen, int reps, void *stream, float *ms)
{
    if (!b || i < 0 || i >= (int)b->cps.size() || !dm || !screen || reps < 1 || !ms)
        return fail(MJK_ERR_ARG, "mjk_batch_time_class_pair: bad argument");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    hipError_t e;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(e, "event");
    }
    int rc = batch_launch_cp(b, b->cps[i], vj, vk, dm, *screen, st);  // warm-up
    if (!rc) (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps && !rc; ++r) rc = batch_launch_cp(b, b->cps[i], vj, vk, dm, *screen, st);
    if (!rc) {
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        *ms = t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int mjk_fill_int2e(mjk_basis_prod *bp, double *eri, int nao, const int64_t *strides, const int32_t *ao_offsets,
                   int cp_ij, int cp_kl, void *stream)
{
    if (!bp || !eri || !strides || !ao_offsets) return fail(MJK_ERR_ARG, "mjk_fill_int2e: null argument");
    if (nao != bp->nao) return fail(MJK_ERR_ARG, "mjk_fill_int2e: nao does not match the cache");
    if (cp_ij < 0 || cp_kl < 0 || cp_ij >= bp->ncptype || cp_kl >= bp->ncptype)
        return fail(MJK_ERR_ARG, "mjk_fill_int2e: class index out of range");
    int dev = -1;
    (void)hipGetDevice(&dev);
    if (dev != bp->device) return fail(MJK_ERR_ARG, "mjk_fill_int2e: current device differs from the cache's device");
    const ClassType &c1 = bp->cp[cp_ij], &c2 = bp->cp[cp_kl];
    if (c1.npairs == 0 || c2.npairs == 0) return MJK_OK;
    // the kernels take the class with more Cartesian products as the bra; the strides follow
    const bool swap = ncart_h(c1.l_bra) * ncart_h(c1.l_ket) < ncart_h(c2.l_bra) * ncart_h(c2.l_ket);
    const ClassType &b = swap ? c2 : c1, &k = swap ? c1 : c2;
    FillArgs f{};
    f.bra = pair_set(bp, b);
    f.ket = pair_set(bp, k);
    f.eri = eri;
    const int ib = swap ? 2 : 0, ik = swap ? 0 : 2;
    f.sb0 = (long)strides[ib];
    f.sb1 = (long)strides[ib + 1];
    f.sk0 = (long)strides[ik];
    f.sk1 = (long)strides[ik + 1];
    f.ob0 = ao_offsets[ib];
    f.ob1 = ao_offsets[ib + 1];
    f.ok0 = ao_offsets[ik];
    f.ok1 = ao_offsets[ik + 1];
    hipError_t e = launch_fill_class(b.l_bra, b.l_ket, k.l_bra, k.l_ket, f, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "mjk_fill_int2e: launch");
    return MJK_OK;
}

int mjk_batch_stats(mjk_batch *b, double *out)
{
    if (!b) return fail(MJK_ERR_ARG, "mjk_batch_stats: null batch");
    return mjk_last_stats(b->bp, out);
}

int mjk_last_stats(mjk_basis_prod *bp, double *out)
{
    if (!bp || !out) return fail(MJK_ERR_ARG, "mjk_last_stats: bad argument");
    unsigned long long c[3];
    hipError_t e = hipMemcpy(c, bp->counters, sizeof(c), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "mjk_last_stats");
    out[0] = (double)c[0];
    out[1] = (double)c[1];
    out[2] = (double)bp->last_launches;
    out[3] = (double)c[2];
    return MJK_OK;
}

}  // extern "C"
