// Kernels of the learned label proposal: a 3-layer perceptron over the score
// step's rank features (doubled centred ranks / G, in BF16),
//     h1 = gelu(x W1 + b1), h2 = gelu(h1 W2 + b2), logits = h2 W3 + b3,
// trained with softmax cross-entropy and Adam (FP32 master weights, BF16
// copies for the contractions).  All contractions run on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csg_mfma.h"

namespace csg {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float  v4f_m __attribute__((ext_vector_type(4)));

// erfc(a) for a >= 0 by Abramowitz & Stegun 7.1.26 (absolute error below
// 1.5e-7): t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-a^2), t = 1 / (1 + p a),
// on the hardware reciprocal and base-2 exponential.
__device__ inline float erfc_as(float a)
{
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, a, 1.f));
    const float p =
        t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f),
                         -0.284496736f),
                 0.254829592f);
    return p * __builtin_amdgcn_exp2f(-(a * a) * 1.4426950408889634f);
}

// Phi(z) = 0.5 erfc(-z / sqrt 2): from erfc of |z| / sqrt 2 on either side, so
// the lower tail keeps its small values (absolute error below 1e-6).
__device__ inline float gelu_cdf(float z)
{
    const float ec = erfc_as(fabsf(z) * 0.70710678118654752f);
    return z >= 0.f ? 1.f - 0.5f * ec : 0.5f * ec;
}

__device__ inline float gelu_f(float z) { return z * gelu_cdf(z); }

__device__ inline float gelu_grad(float z)
{
    const float pdf = 0.39894228040143268f * __builtin_amdgcn_exp2f(-0.5f * z * z * 1.4426950408889634f);
    return gelu_cdf(z) + z * pdf;
}

// C[M x N] (FP32, row-major, ldc) = sum_k A(m, k) * B(k, n) with
//   A(m, k) = TA ? A[k * lda + m] : A[m * lda + k]
//   B(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n]
// (BF16 operands).  Block tile 64 x 64 (4 waves of 32 x 32), K staged 32 at a
// time through LDS in k-contiguous rows; out-of-range rows / k read as zero.
// A k-contiguous operand needs K % 32 == 0 and a leading dimension that is a
// multiple of 8.  beta = 0 overwrites C, beta = 1 accumulates into it.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_bgemm(const __bf16* __restrict__ A, int lda,
                                               const __bf16* __restrict__ B, int ldb,
                                               float* __restrict__ C, int ldc, int M, int N,
                                               int K, float beta)
{
    constexpr int BM = 64, BN = 64, BK = 32, LD = 40;  // LDS row: 32 k + 8 pad
    __shared__ __attribute__((aligned(16))) __bf16 sa[BM * LD];
    __shared__ __attribute__((aligned(16))) __bf16 sb[BN * LD];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    v4f_m     acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            acc[i][j] = v4f_m{ 0.f, 0.f, 0.f, 0.f };
    const __bf16 zero = (__bf16)0.f;
    for (int k0 = 0; k0 < K; k0 += BK)
    {
        // 64 x 32 values per operand.  A k-contiguous operand (A untransposed,
        // B transposed) is staged with one 16-byte load per thread (K is a
        // multiple of 32 and the leading dimension of 8); the other layout
        // with 8 coalesced 2-byte loads per thread.
        if constexpr (!TA)
        {
            const int r = tid >> 2, k = (tid & 3) * 8, gm = m0 + r;
            int4      v = int4{ 0, 0, 0, 0 };
            if (gm < M)
                v = *(const int4*)(A + (size_t)gm * lda + k0 + k);
            *(int4*)(&sa[r * LD + k]) = v;
        }
        else
        {
#pragma unroll
            for (int u = 0; u < 8; ++u)
            {
                const int e = tid + u * 256;
                const int k = e >> 6, r = e & 63;  // consecutive threads read consecutive m
                const int gm = m0 + r, gk = k0 + k;
                __bf16    v  = zero;
                if (gm < M && gk < K)
                    v = A[(size_t)gk * lda + gm];
                sa[r * LD + k] = v;
            }
        }
        if constexpr (TB)
        {
            const int c = tid >> 2, k = (tid & 3) * 8, gn = n0 + c;
            int4      v = int4{ 0, 0, 0, 0 };
            if (gn < N)
                v = *(const int4*)(B + (size_t)gn * ldb + k0 + k);
            *(int4*)(&sb[c * LD + k]) = v;
        }
        else
        {
#pragma unroll
            for (int u = 0; u < 8; ++u)
            {
                const int e = tid + u * 256;
                const int k = e >> 6, c = e & 63;
                const int gn = n0 + c, gk = k0 + k;
                __bf16    v  = zero;
                if (gn < N && gk < K)
                    v = B[(size_t)gk * ldb + gn];
                sb[c * LD + k] = v;
            }
        }
        __syncthreads();
        const int r16 = lane & 15, kofs = (lane >> 4) * 8;
        v8bf      a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = *(const v8bf*)(&sa[(wm * 32 + i * 16 + r16) * LD + kofs]);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            b[j] = *(const v8bf*)(&sb[(wn * 32 + j * 16 + r16) * LD + kofs]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + e;
            if (m >= M)
                continue;
#pragma unroll
            for (int j = 0; j < 2; ++j)
            {
                const int n = n0 + wn * 32 + j * 16 + (lane & 15);
                if (n < N)
                {
                    float* c = C + (size_t)m * ldc + n;
                    *c       = beta != 0.f ? *c + acc[i][j][e] : acc[i][j][e];
                }
            }
        }
}

// Forward layer product C = A B^T with both operands k-contiguous:
// A [M x K] (lda), B [N x K] (ldb, the layer's transposed BF16 weights),
// K % 32 == 0, lda and ldb multiples of 8.  Block tile 128 x 128, four waves
// of 64 x 64 (4 x 4 MFMA tiles each), K staged 32 at a time through two LDS
// buffers.  Epilogue (row-major, leading dimension ldc):
//   MODE 0: C = acc (FP32);
//   MODE 1: v = acc + bias[n]; z = v (FP32, when given); H = gelu(v) (BF16).
template <int MODE>
__global__ __launch_bounds__(256) void k_bgemm_nt(const __bf16* __restrict__ A, int lda,
                                                  const __bf16* __restrict__ B, int ldb, int M,
                                                  int N, int K, int ldc, float* __restrict__ C,
                                                  const float* __restrict__ bias,
                                                  float* __restrict__ z, __bf16* __restrict__ H)
{
    constexpr int T = 128, BK = 32, LD = 40;
    __shared__ __attribute__((aligned(16))) __bf16 sa[2][T * LD];
    __shared__ __attribute__((aligned(16))) __bf16 sb[2][T * LD];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int m0 = blockIdx.y * T, n0 = blockIdx.x * T;
    const int r0 = tid >> 2, kc = (tid & 3) * 8;  // rows r0 and r0 + 64, 8 values at kc
    int4      va[2], vb[2];
    auto      load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            const int gm = m0 + r0 + 64 * u, gn = n0 + r0 + 64 * u;
            va[u] = gm < M ? *(const int4*)(A + (size_t)gm * lda + k0 + kc) : int4{ 0, 0, 0, 0 };
            vb[u] = gn < N ? *(const int4*)(B + (size_t)gn * ldb + k0 + kc) : int4{ 0, 0, 0, 0 };
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            *(int4*)(&sa[buf][(r0 + 64 * u) * LD + kc]) = va[u];
            *(int4*)(&sb[buf][(r0 + 64 * u) * LD + kc]) = vb[u];
        }
    };
    v4f_m acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            acc[i][j] = v4f_m{ 0.f, 0.f, 0.f, 0.f };
    load(0);
    store(0);
    __syncthreads();
    const int r16 = lane & 15, kofs = (lane >> 4) * 8;
    int       buf = 0;
    for (int k0 = 0; k0 < K; k0 += BK)
    {
        const bool more = k0 + BK < K;
        if (more)
            load(k0 + BK);
        v8bf a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            a[i] = *(const v8bf*)(&sa[buf][(wm * 64 + i * 16 + r16) * LD + kofs]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            b[j] = *(const v8bf*)(&sb[buf][(wn * 64 + j * 16 + r16) * LD + kofs]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (more)
        {
            // buffer buf ^ 1 was last read before the previous barrier
            store(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }
    // the lane's four column biases, loaded once (a column past N reads
    // column N - 1 and stores nothing)
    float bj[4] = { 0.f, 0.f, 0.f, 0.f };
    if constexpr (MODE != 0)
    {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            bj[j] = bias[min(n0 + wn * 64 + j * 16 + r16, N - 1)];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + e;
            if (m >= M)
                continue;
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int n = n0 + wn * 64 + j * 16 + r16;
                if (n >= N)
                    continue;
                const size_t o = (size_t)m * ldc + n;
                if constexpr (MODE == 0)
                    C[o] = acc[i][j][e];
                else
                {
                    const float v = acc[i][j][e] + bj[j];
                    if (z)
                        z[o] = v;
                    H[o] = (__bf16)gelu_f(v);
                }
            }
        }
}

// The hidden layers' product for wide layers (N a multiple of 256, K of 32),
// with MODE 1's epilogue: block tile 256 x 256, eight waves of 128 x 64, each
// in 4 x 2 tiles of v_mfma_f32_32x32x16_bf16 (the 32 x 32 MFMAs issue closer
// to their peak rate than the 16 x 16 ones).  K is staged 32 at a time
// through four LDS stages filled by direct-to-LDS loads, three in flight; a
// stage holds [A 256 rows][B 256 rows] of 32 BF16 (64 bytes), the 16-byte
// slot of row r holding K chunk c being c ^ bw_swz(r).  Epilogue: v = acc +
// bias[n]; z = v (FP32, when given); the BF16 gelu(v) of a wave's 128 x 64
// outputs is staged in LDS and written by 16-byte row stores.  Tiles run in
// XCD order, the column tiles of a row tile next to each other.
constexpr int BW_BM = 256, BW_BN = 256, BW_BK = 32, BW_NS = 4;
constexpr int BW_STAGE = (BW_BM + BW_BN) * BW_BK * 2;  // 32 KB
constexpr int BW_LDS   = BW_NS * BW_STAGE;             // also the epilogue's 8 x 16 KB

__device__ __forceinline__ int bw_swz(int r) { return (r >> 2) & 3; }

typedef float v16f_m __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(512) void k_bgemm_nt_w(const __bf16* __restrict__ A, int lda,
                                                    const __bf16* __restrict__ B, int ldb, int M,
                                                    int N, int K, int ldh,
                                                    const float* __restrict__ bias,
                                                    float* __restrict__ z, __bf16* __restrict__ H)
{
    extern __shared__ __attribute__((aligned(16))) int8_t bw_sm[];
    const int      tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int      wm = wid >> 2, wn = wid & 3;  // 128-row half, 64-column quarter
    const uint32_t nct = (uint32_t)(N / BW_BN);
    const uint32_t t   = xcd_order(blockIdx.x, gridDim.x);
    const int      m0 = (int)(t / nct) * BW_BM, n0 = (int)(t % nct) * BW_BN;
    // this lane's source (element offset, swizzled chunk) for its four row
    // groups g = wid + 8 u of 16 rows: groups 0-15 A rows, 16-31 B rows; a
    // wave's load fills one group, lane l row l / 4, slot l % 4
    size_t src[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        const int g = wid + 8 * u;
        const int r = (g < 16 ? g : g - 16) * 16 + (lane >> 2);
        const int c = (lane & 3) ^ bw_swz(r);
        src[u]      = g < 16 ? (size_t)min(m0 + r, M - 1) * lda + c * 8
                             : (size_t)min(n0 + r, N - 1) * ldb + c * 8;
    }
    auto issue = [&](int8_t* stage, int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int     g    = wid + 8 * u;
            const __bf16* base = g < 16 ? A : B;
            __builtin_amdgcn_global_load_lds(
                base + src[u] + k0, (__attribute__((address_space(3))) void*)(stage + g * 1024), 16,
                0, 0);
        }
    };
    v16f_m acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e)
                acc[i][j][e] = 0.f;
    const int T = K / BW_BK;
    issue(bw_sm, 0);
    if (T > 1)
        issue(bw_sm + BW_STAGE, BW_BK);
    if (T > 2)
        issue(bw_sm + 2 * BW_STAGE, 2 * BW_BK);
    // 32 x 32 x 16 operands: lane l holds row (l % 32), K elements 8 (l / 32) .. + 7
    const int r32 = lane & 31, kh = lane >> 5;
    for (int s = 0; s < T; ++s)
    {
        // stage s is complete when at most the loads of stages s + 1 and s + 2
        // (four each) are outstanding; the barrier publishes every wave's
        // loads and ends every wave's reads of stage s - 1, whose buffer
        // stage s + 3 reuses
        if (s + 2 < T)
            __builtin_amdgcn_s_waitcnt(0x0F78);  // vmcnt <= 8
        else if (s + 1 < T)
            __builtin_amdgcn_s_waitcnt(0x0F74);  // vmcnt <= 4
        else
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt 0
        asm volatile("s_barrier" ::: "memory");
        if (s + 3 < T)
            issue(bw_sm + ((s + 3) % BW_NS) * BW_STAGE, (s + 3) * BW_BK);
        const int8_t* st = bw_sm + (s % BW_NS) * BW_STAGE;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
        {
            const int ch = ks * 2 + kh;
            v8bf      a[4], b[2];
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
                const int row = wm * 128 + i * 32 + r32;
                a[i] = *(const v8bf*)(st + row * 64 + ((ch ^ bw_swz(row)) << 4));
            }
#pragma unroll
            for (int j = 0; j < 2; ++j)
            {
                const int row = wn * 64 + j * 32 + r32;
                b[j] = *(const v8bf*)(st + (BW_BM + row) * 64 + ((ch ^ bw_swz(row)) << 4));
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    // the lane's two column biases (N is a multiple of BW_BN)
    float bj[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
        bj[j] = bias[n0 + wn * 64 + j * 32 + r32];
    __syncthreads();  // every wave is done with the stages
    // 32 x 32 result: acc[e] of lane l is row 8 (e / 4) + 4 (l / 32) + e % 4,
    // column l % 32.  The wave's 128 x 64 BF16 outputs go to its 16 KB of
    // LDS (16-byte chunk c of row r at slot c ^ (r & 7)), then out by rows.
    __bf16* w = (__bf16*)bw_sm + wid * (128 * 64);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 16; ++e)
#pragma unroll
            for (int j = 0; j < 2; ++j)
            {
                const float v   = acc[i][j][e] + bj[j];
                const int   row = i * 32 + (e >> 2) * 8 + kh * 4 + (e & 3);
                const int   col = j * 32 + r32;
                if (z)
                {
                    const int m = m0 + wm * 128 + row;
                    if (m < M)
                        z[(size_t)m * ldh + n0 + wn * 64 + col] = v;
                }
                w[row * 64 + (((col >> 3) ^ (row & 7)) << 3) + (col & 7)] = (__bf16)gelu_f(v);
            }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // this wave's rows only
#pragma unroll
    for (int q = 0; q < 16; ++q)
    {
        const int r = q * 8 + (lane >> 3), c = lane & 7;
        const int m = m0 + wm * 128 + r;
        const int4 v = *(const int4*)(w + r * 64 + ((c ^ (r & 7)) << 3));
        if (m < M)
            *(int4*)(H + (size_t)m * ldh + n0 + wn * 64 + c * 8) = v;
    }
}

// dst [C x R] (BF16) = transpose of src [R x C] (FP32).
__global__ void k_transpose_bf16(const float* __restrict__ src, int R, int Cn,
                                 __bf16* __restrict__ dst)
{
    __shared__ float tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    for (int y = ty; y < 32; y += 8)
    {
        const int r = r0 + y, c = c0 + tx;
        tile[y][tx] = r < R && c < Cn ? src[(size_t)r * Cn + c] : 0.f;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8)
    {
        const int c = c0 + y, r = r0 + tx;
        if (c < Cn && r < R)
            dst[(size_t)c * R + r] = (__bf16)tile[tx][y];
    }
}

// dz = dh * gelu'(z), FP32 for the bias gradient and BF16 for the next product.
__global__ void k_gelu_back(const float* __restrict__ dh, const float* __restrict__ z, size_t n,
                            float* __restrict__ dz, __bf16* __restrict__ dzb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float v = dh[i] * gelu_grad(z[i]);
    dz[i]         = v;
    dzb[i]        = (__bf16)v;
}

// Softmax cross-entropy over the first `L` of `ldl` logit columns (+ bias):
// dlogits = (softmax - onehot) / rows (FP32 and BF16); loss summed atomically.
__global__ void k_softmax_xent(const float* __restrict__ acc, const float* __restrict__ bias,
                               const int32_t* __restrict__ labels, int rows, int L, int ldl,
                               float* __restrict__ dlog, __bf16* __restrict__ dlogb,
                               float* __restrict__ loss)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows)
        return;
    const float* a  = acc + (size_t)r * ldl;
    float        mx = -3.0e38f;
    for (int j = 0; j < L; ++j)
        mx = fmaxf(mx, a[j] + bias[j]);
    float s = 0.f;
    for (int j = 0; j < L; ++j)
        s += expf(a[j] + bias[j] - mx);
    const int   y   = labels[r];
    const float inv = 1.f / (float)rows;
    for (int j = 0; j < ldl; ++j)
    {
        float g = 0.f;
        if (j < L)
            g = (expf(a[j] + bias[j] - mx) / s - (j == y ? 1.f : 0.f)) * inv;
        dlog[(size_t)r * ldl + j]  = g;
        dlogb[(size_t)r * ldl + j] = (__bf16)g;
    }
    atomicAdd(loss, (logf(s) + mx - (a[y] + bias[y])) * inv);
}

// Column sums of a [rows x cols] FP32 matrix (bias gradients).
__global__ void k_colsum(const float* __restrict__ x, int rows, int cols, float* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols)
        return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r)
        s += x[(size_t)r * cols + c];
    out[c] = s;
}

// Adam on FP32 master weights; refreshes the BF16 copy (when given).
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, __bf16* __restrict__ wb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i]           = mi;
    v[i]           = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i]           = wi;
    if (wb)
        wb[i] = (__bf16)wi;
}

__global__ void k_to_bf16(const float* __restrict__ x, size_t n, __bf16* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = (__bf16)x[i];
}

// Rank features: X / G in BF16 for `rows` rows of a [rows x Kp] int16 matrix.
__global__ void k_rank_features(const int16_t* __restrict__ X, size_t n, float inv_g,
                                __bf16* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = (__bf16)((float)X[i] * inv_g);
}

// Per row: the K largest logits (+ bias) in descending order (ties keep the
// lower label first) and the softmax margin p(top1) - p(top2); one wave per
// row, L <= 256 logits (four per lane).  A row with a NaN logit gets margin
// NaN and labels -1.
// Largest (value, label) pair of the wave -- larger value first, lower
// label among equal values -- returned to every lane: rows of 16 by DPP
// row shifts (a lane whose source lies outside its row compares with
// itself), then the four rows' winners.  The order is total on the pairs
// taken here (no NaN values, distinct labels), so the winner does not
// depend on the reduction order.  Every lane of the wave must be active.
__device__ inline void wave_argmax(float& bv, int& bj)
{
    auto take = [&](float ov, int oj) {
        if (ov > bv || (ov == bv && oj < bj))
        {
            bv = ov;
            bj = oj;
        }
    };
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
        const int b = __builtin_bit_cast(int, bv);
        int       tv, tj;
        switch (s)
        {
        case 0: tv = CSG_DPP_ROW_SHR_KEEP(b, 1); tj = CSG_DPP_ROW_SHR_KEEP(bj, 1); break;
        case 1: tv = CSG_DPP_ROW_SHR_KEEP(b, 2); tj = CSG_DPP_ROW_SHR_KEEP(bj, 2); break;
        case 2: tv = CSG_DPP_ROW_SHR_KEEP(b, 4); tj = CSG_DPP_ROW_SHR_KEEP(bj, 4); break;
        default: tv = CSG_DPP_ROW_SHR_KEEP(b, 8); tj = CSG_DPP_ROW_SHR_KEEP(bj, 8); break;
        }
        take(__builtin_bit_cast(float, tv), tj);
    }
    const int   b  = __builtin_bit_cast(int, bv);
    const int   jj = bj;
    const float v0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 15));
    const int   j0 = __builtin_amdgcn_readlane(jj, 15);
    const float v1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 31));
    const int   j1 = __builtin_amdgcn_readlane(jj, 31);
    const float v2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 47));
    const int   j2 = __builtin_amdgcn_readlane(jj, 47);
    const float v3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 63));
    const int   j3 = __builtin_amdgcn_readlane(jj, 63);
    bv = v0;
    bj = j0;
    take(v1, j1);
    take(v2, j2);
    take(v3, j3);
}

template <int K>
__global__ __launch_bounds__(256) void k_topk(const float* __restrict__ acc,
                                              const float* __restrict__ bias, int rows, int L,
                                              int ldl, int32_t* __restrict__ top,
                                              float* __restrict__ margin)
{
    constexpr int KK   = K < 2 ? 2 : K;  // the margin needs the runner-up
    const int     lane = threadIdx.x & 63;
    const int     r    = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows)
        return;
    const float* a = acc + (size_t)r * ldl;
    float        v[4];
    bool         nan = false;
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        const int j = lane + 64 * u;
        v[u]        = j < L ? a[j] + bias[j] : -__builtin_inff();
        nan |= __builtin_isnan(v[u]);
    }
    if (__ballot(nan))
    {
        if (lane < K)
            top[(size_t)r * K + lane] = -1;
        if (lane == 0)
            margin[r] = __builtin_nanf("");
        return;
    }
    const float mx = wave_max(fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3])));
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        s += lane + 64 * u < L ? expf(v[u] - mx) : 0.f;
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o);
    float b0 = 0.f, b1 = 0.f;
    for (int t = 0; t < KK; ++t)
    {
        // the largest remaining value, lowest label among equals
        float bv = -__builtin_inff();
        int   bj = 1 << 30;
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int j = lane + 64 * u;
            if (j < L && (v[u] > bv || (v[u] == bv && j < bj)))
            {
                bv = v[u];
                bj = j;
            }
        }
        wave_argmax(bv, bj);
        if (bj >= L)
            bj = -1;  // fewer than KK labels
        if (t < K && lane == 0)
            top[(size_t)r * K + t] = bj;
        if (t == 0)
            b0 = bv;
        if (t == 1)
            b1 = bv;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (lane + 64 * u == bj)
                v[u] = -__builtin_inff() * 0.f;  // taken: NaN never wins a comparison
    }
    if (lane == 0)
        margin[r] = L > 1 ? (expf(b0 - mx) - expf(b1 - mx)) / s : 1.f;
}

// The same for L <= 64 with one thread per row, bit for bit: 128 rows per
// block staged through LDS (row stride 65, so a wave's rows fall in distinct
// banks), each row's values in registers.  The selection scans the labels in
// ascending order with k_topk's predicate, and the softmax sum adds the
// per-label terms in the tree order of k_topk's butterfly (lane 0's sum:
// pairs (j, j + o) for o = 32, 16, ..., 1).  The wave kernel's reductions are
// chains of cross-lane shuffles, about a microsecond per row.
template <int K>
__global__ __launch_bounds__(128) void k_topk_rows(const float* __restrict__ acc,
                                                   const float* __restrict__ bias, int rows,
                                                   int L, int ldl, int32_t* __restrict__ top,
                                                   float* __restrict__ margin)
{
    constexpr int KK = K < 2 ? 2 : K;
    __shared__ float tile[128 * 65];
    const int        r0 = blockIdx.x * 128;
    const int        nr = min(128, rows - r0);
    for (int idx = threadIdx.x; idx < nr * L; idx += 128)
    {
        const int rr = idx / L, j = idx - rr * L;
        tile[rr * 65 + j] = acc[(size_t)(r0 + rr) * ldl + j] + bias[j];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= nr)
        return;
    const int r = r0 + t;
    float     v[64];
    bool      nan = false;
#pragma unroll
    for (int j = 0; j < 64; ++j)
    {
        v[j] = j < L ? tile[t * 65 + j] : -__builtin_inff();
        nan |= __builtin_isnan(v[j]);
    }
    if (nan)
    {
        for (int u = 0; u < K; ++u)
            top[(size_t)r * K + u] = -1;
        margin[r] = __builtin_nanf("");
        return;
    }
    float mx = v[0];
#pragma unroll
    for (int j = 1; j < 64; ++j)
        mx = fmaxf(mx, v[j]);
    float e[64];
#pragma unroll
    for (int j = 0; j < 64; ++j)
        e[j] = j < L ? expf(v[j] - mx) : 0.f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < o; ++j)
            e[j] = e[j] + e[j + o];
    const float s     = e[0];
    uint64_t    taken = 0;
    float       b0 = 0.f, b1 = 0.f;
#pragma unroll
    for (int tt = 0; tt < KK; ++tt)
    {
        float bv = -__builtin_inff();
        int   bj = 1 << 30;
#pragma unroll
        for (int j = 0; j < 64; ++j)
            if (!((taken >> j) & 1) && j < L && (v[j] > bv || (v[j] == bv && j < bj)))
            {
                bv = v[j];
                bj = j;
            }
        if (bj >= L)
            bj = -1;  // fewer than KK labels
        else
            taken |= 1ull << bj;
        if (tt < K)
            top[(size_t)r * K + tt] = bj;
        if (tt == 0)
            b0 = bv;
        if (tt == 1)
            b1 = bv;
    }
    margin[r] = L > 1 ? (expf(b0 - mx) - expf(b1 - mx)) / s : 1.f;
}

// Rows perm[0, rows) of a [* x cols] BF16 matrix (cols a multiple of 8) and
// their labels, gathered into a minibatch.
__global__ void k_gather_batch(const __bf16* __restrict__ feat, const int32_t* __restrict__ labels,
                               const uint32_t* __restrict__ perm, int rows, int cols,
                               __bf16* __restrict__ xb, int32_t* __restrict__ yb)
{
    const int per = cols / 8;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * per)
        return;
    const int r = (int)(i / per), c = (int)(i % per) * 8;
    const uint32_t src = perm[r];
    *(int4*)(xb + (size_t)r * cols + c) = *(const int4*)(feat + (size_t)src * cols + c);
    if (c == 0)
        yb[r] = labels[src];
}

// [rows x g] FP32 features into a zero-padded [rows x cols] BF16 matrix.
__global__ void k_pad_bf16(const float* __restrict__ x, int rows, int g, int cols,
                           __bf16* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols)
        return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    out[i]      = (__bf16)(c < g ? x[(size_t)r * g + c] : 0.f);
}

}  // namespace csg
