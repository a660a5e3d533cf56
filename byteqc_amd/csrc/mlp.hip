// Kernels of the learned initial density guess: a perceptron with bf16 weights and
// activations on MFMA (fp32 accumulate), its backward pass, a masked squared-error loss and
// Adam.  The GEMM tile is 64 x 64 per 256-thread workgroup; each wave owns a 32 x 32
// quarter made of 2 x 2 v_mfma_f32_16x16x16_bf16 blocks, with both operands staged in LDS
// as [row][k] so a lane's MFMA fragment (four consecutive k) is one 8-byte LDS read.
// The reference has no counterpart: its SCF starts from PySCF's minao/SAD guess.
#include <stdint.h>

#include "../../include/mi355jk.h"
#include "mjk_internal.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (inputs are finite)
__device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }

__device__ inline float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;  // LDS row pitch (bf16 elements)

// C[M][N] (row-major, ldc) = op(A) op(B) + beta C, bf16 operands, fp32 accumulate.
//   LAYOUT 0 (NN): A is [M][K] (lda), B is [K][N] (ldb)
//   LAYOUT 1 (NT): A is [M][K],       B is [N][K]        (C = A B^T)
//   LAYOUT 2 (TN): A is [K][M],       B is [K][N]        (C = A^T B)
template <int LAYOUT>
__global__ void __launch_bounds__(256)
gemm_bf16_kernel(int M, int N, int K, const uint16_t *__restrict__ A, int lda,
                 const uint16_t *__restrict__ B, int ldb, float *__restrict__ C, int ldc, float beta)
{
    __shared__ __align__(16) uint16_t sa[BM][LDK];
    __shared__ __align__(16) uint16_t sb[BN][LDK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += BK) {
        // stage sa[m][k] = A(m0 + m, k0 + k); each thread moves 8 elements
        if (LAYOUT == 2) {  // m contiguous in memory: consecutive threads take consecutive m
            const int m = t & 63, kb = (t >> 6) * 8, gm = m0 + m;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int gk = k0 + kb + j;
                sa[m][kb + j] = (gm < M && gk < K) ? A[(size_t)gk * lda + gm] : (uint16_t)0;
            }
        } else {            // k contiguous
            const int m = t >> 2, kb = (t & 3) * 8, gm = m0 + m;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int gk = k0 + kb + j;
                sa[m][kb + j] = (gm < M && gk < K) ? A[(size_t)gm * lda + gk] : (uint16_t)0;
            }
        }
        // stage sb[n][k] = B(k0 + k, n0 + n)
        if (LAYOUT == 1) {  // k contiguous
            const int n = t >> 2, kb = (t & 3) * 8, gn = n0 + n;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int gk = k0 + kb + j;
                sb[n][kb + j] = (gn < N && gk < K) ? B[(size_t)gn * ldb + gk] : (uint16_t)0;
            }
        } else {            // n contiguous
            const int n = t & 63, kb = (t >> 6) * 8, gn = n0 + n;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int gk = k0 + kb + j;
                sb[n][kb + j] = (gn < N && gk < K) ? B[(size_t)gk * ldb + gn] : (uint16_t)0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            // lane holds rows (lane & 15) and k = kk + 4 (lane >> 4) .. + 3 of each operand
            v4s fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                fa[i] = *(const v4s *)&sa[wm + i * 16 + (lane & 15)][kk + (lane >> 4) * 4];
                fb[i] = *(const v4s *)&sb[wn + i * 16 + (lane & 15)][kk + (lane >> 4) * 4];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator element r of a 16 x 16 block: row 4 (lane >> 4) + r, column lane & 15
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    float v = acc[i][j][r];
                    if (beta != 0.f) v += beta * C[(size_t)gm * ldc + gn];
                    C[(size_t)gm * ldc + gn] = v;
                }
            }
}

// z += bias (per column); a_out = bf16(act(z)) when a_out is given
__global__ void bias_act_kernel(int M, int N, float *z, const float *bias, uint16_t *a_out, int act)
{
    const size_t n = (size_t)M * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float v = z[i] + (bias ? bias[i % N] : 0.f);
        z[i] = v;
        if (a_out) a_out[i] = f2bf(act == 1 ? silu(v) : v);
    }
}

// dz = da * act'(z), stored as bf16 for the next GEMMs
__global__ void act_backward_kernel(int M, int N, const float *da, const float *z, uint16_t *dz, int act)
{
    const size_t n = (size_t)M * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float g = act == 1 ? da[i] * silu_grad(z[i]) : da[i];
        dz[i] = f2bf(g);
    }
}

// dbias[n] = sum_m da[m][n] act'(z[m][n]) in fp32 (one thread per column, 64-row slabs
// summed by a second level of atomics)
__global__ void bias_grad_kernel(int M, int N, const float *da, const float *z, float *dbias, int act)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int r0 = blockIdx.y * 256, r1 = min(M, r0 + 256);
    float s = 0.f;
    for (int m = r0; m < r1; ++m) {
        const size_t i = (size_t)m * N + n;
        s += act == 1 ? da[i] * silu_grad(z[i]) : da[i];
    }
    atomicAdd(dbias + n, s);
}

// dy = scale * mask * (y - t); loss += sum mask (y - t)^2
__global__ void mse_grad_kernel(size_t n, const float *y, const float *tg, const float *mask, float scale,
                                float *dy, double *loss)
{
    __shared__ double red[256];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float mk = mask ? mask[i] : 1.f;
        const float d = y[i] - tg[i];
        dy[i] = scale * mk * d;
        s += (double)mk * d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

__global__ void adam_kernel(size_t n, float *w, const float *g, float *m1, float *m2, uint16_t *wb, float lr,
                            float b1, float b2, float eps, float c1, float c2)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        const float wi = w[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

__global__ void to_bf16_kernel(size_t n, const float *x, uint16_t *y)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

__global__ void fill_f32_kernel(size_t n, float *x)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = 0.f;
}

unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? (b ? b : 1) : 4096);
}

int launched(const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mjk::report_error(MJK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return MJK_OK;
}

}  // namespace

extern "C" {

int mjk_gemm_bf16(int layout, int m, int n, int k, const uint16_t *a, int lda, const uint16_t *b, int ldb,
                  float *c, int ldc, float beta, void *stream)
{
    if (m < 0 || n < 0 || k < 0 || !c || (k > 0 && (!a || !b)))
        return mjk::report_error(MJK_ERR_ARG, "mjk_gemm_bf16: bad sizes or null operand");
    const int need_a = layout == 2 ? m : k, need_b = layout == 1 ? k : n;
    if (lda < need_a || ldb < need_b || ldc < n)
        return mjk::report_error(MJK_ERR_ARG, "mjk_gemm_bf16: leading dimension too small");
    if (m == 0 || n == 0) return MJK_OK;
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((n + BN - 1) / BN, (m + BM - 1) / BM);
    if (grid.y > 65535) return mjk::report_error(MJK_ERR_ARG, "mjk_gemm_bf16: m too large");
    switch (layout) {
    case 0: gemm_bf16_kernel<0><<<grid, 256, 0, st>>>(m, n, k, a, lda, b, ldb, c, ldc, beta); break;
    case 1: gemm_bf16_kernel<1><<<grid, 256, 0, st>>>(m, n, k, a, lda, b, ldb, c, ldc, beta); break;
    case 2: gemm_bf16_kernel<2><<<grid, 256, 0, st>>>(m, n, k, a, lda, b, ldb, c, ldc, beta); break;
    default: return mjk::report_error(MJK_ERR_ARG, "mjk_gemm_bf16: layout must be 0 (NN), 1 (NT) or 2 (TN)");
    }
    return launched("mjk_gemm_bf16");
}

int mjk_bias_act(int m, int n, float *z, const float *bias, uint16_t *a_out, int act, void *stream)
{
    if (m < 0 || n <= 0 || !z || act < 0 || act > 1) return mjk::report_error(MJK_ERR_ARG, "mjk_bias_act: bad argument");
    const size_t cnt = (size_t)m * n;
    if (!cnt) return MJK_OK;
    bias_act_kernel<<<grid_for(cnt), 256, 0, (hipStream_t)stream>>>(m, n, z, bias, a_out, act);
    return launched("mjk_bias_act");
}

int mjk_act_backward(int m, int n, const float *da, const float *z, uint16_t *dz, float *dbias, int act, void *stream)
{
    if (m < 0 || n <= 0 || !da || !dz || act < 0 || act > 1 || (act == 1 && !z))
        return mjk::report_error(MJK_ERR_ARG, "mjk_act_backward: bad argument");
    const size_t cnt = (size_t)m * n;
    hipStream_t st = (hipStream_t)stream;
    if (dbias) {
        fill_f32_kernel<<<grid_for(n), 256, 0, st>>>((size_t)n, dbias);
        if (m > 0) {
            const dim3 grid((n + 255) / 256, (m + 255) / 256);
            if (grid.y > 65535) return mjk::report_error(MJK_ERR_ARG, "mjk_act_backward: m too large");
            bias_grad_kernel<<<grid, 256, 0, st>>>(m, n, da, z, dbias, act);
        }
    }
    if (cnt) act_backward_kernel<<<grid_for(cnt), 256, 0, st>>>(m, n, da, z, dz, act);
    return launched("mjk_act_backward");
}

int mjk_mse_grad(size_t n, const float *y, const float *target, const float *mask, float scale, float *dy,
                 double *loss, void *stream)
{
    if (!y || !target || !dy) return mjk::report_error(MJK_ERR_ARG, "mjk_mse_grad: null argument");
    if (!n) return MJK_OK;
    mse_grad_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(n, y, target, mask, scale, dy, loss);
    return launched("mjk_mse_grad");
}

int mjk_adam(size_t n, float *w, const float *g, float *m1, float *m2, uint16_t *w_bf16, float lr, float beta1,
             float beta2, float eps, int step, void *stream)
{
    if (!w || !g || !m1 || !m2 || step < 1) return mjk::report_error(MJK_ERR_ARG, "mjk_adam: bad argument");
    if (!n) return MJK_OK;
    const float c1 = 1.f / (1.f - powf(beta1, (float)step));
    const float c2 = 1.f / (1.f - powf(beta2, (float)step));
    adam_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(n, w, g, m1, m2, w_bf16, lr, beta1, beta2, eps, c1, c2);
    return launched("mjk_adam");
}

int mjk_to_bf16(size_t n, const float *x, uint16_t *y, void *stream)
{
    if (!x || !y) return mjk::report_error(MJK_ERR_ARG, "mjk_to_bf16: null argument");
    if (!n) return MJK_OK;
    to_bf16_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(n, x, y);
    return launched("mjk_to_bf16");
}

}  // extern "C"
