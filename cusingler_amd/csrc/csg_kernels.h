// Device kernels of the SingleR annotation engine (gfx950 / MI355X).
//
// Value convention.  Ranks are kept as doubled, centred integers
//     X = 2 * rank - (G + 1)
// in int16, where rank is the 1-based tie-averaged rank of the value among the
// G genes of the active gene set (src/cusingler.cu:745-801 rankdata_bin3,
// :159-179 rankdata and :601-625 rankdata_batch all produce it).  X is exact;
// the reference's centred float rank x = rank - (G+1)/2 equals X/2.  Every
// FP32 sum below is formed from X instead of x: scaling all inputs by a power
// of two scales every rounded sum, product, square root and quotient exactly,
// so the resulting correlation has the same bits as the reference's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csg_stdalgo.h"

namespace csg {

constexpr int kWave = 64;

// ------------------------------------------------------------- utilities ---

// DPP lane moves within rows of 16 lanes (no LDS traffic): row_shr:n gives
// lane l the value of lane l - n, row_shl:n that of lane l + n; a source
// outside the row reads 0.
#define CSG_DPP_ROW_SHR(v, n) __builtin_amdgcn_update_dpp(0, (v), 0x110 + (n), 0xF, 0xF, true)
#define CSG_DPP_ROW_SHL(v, n) __builtin_amdgcn_update_dpp(0, (v), 0x100 + (n), 0xF, 0xF, true)
// row_shr:n where a lane whose source is outside the row keeps its own value
#define CSG_DPP_ROW_SHR_KEEP(v, n) \
    __builtin_amdgcn_update_dpp((v), (v), 0x110 + (n), 0xF, 0xF, false)

// Inclusive prefix sum over the wave by DPP: Hillis-Steele steps 1, 2, 4, 8
// inside each row of 16, then the totals of the lower rows (lanes 15, 31,
// 47) added.  Every lane of the wave must be active.
__device__ inline int wave_incl_scan(int v)
{
    v += CSG_DPP_ROW_SHR(v, 1);
    v += CSG_DPP_ROW_SHR(v, 2);
    v += CSG_DPP_ROW_SHR(v, 4);
    v += CSG_DPP_ROW_SHR(v, 8);
    const int r0  = __builtin_amdgcn_readlane(v, 15);
    const int r1  = __builtin_amdgcn_readlane(v, 31);
    const int r2  = __builtin_amdgcn_readlane(v, 47);
    const int row = (threadIdx.x & (kWave - 1)) >> 4;
    return v + (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
}

// Lane 0's result of the shuffle-down tree `for (o = 1; o < 64; o *= 2)
// v += shfl_down(v, o)`, returned to every lane: lane l + o's partial is
// added to lane l's at each level, so lane 0 ends with (T0 + T1) + (T2 + T3)
// where Tk is the same pairwise tree over row k (lanes 16k .. 16k + 15).
// The rows' trees are built by DPP row shifts toward lane 0, the four row
// totals read from lanes 0, 16, 32 and 48: the same additions of the same
// operands, so the same bits.  Every lane of the wave must be active.
__device__ inline float wave_tree_sum(float v)
{
    int b = __builtin_bit_cast(int, v);
    v += __builtin_bit_cast(float, CSG_DPP_ROW_SHL(b, 1));
    b = __builtin_bit_cast(int, v);
    v += __builtin_bit_cast(float, CSG_DPP_ROW_SHL(b, 2));
    b = __builtin_bit_cast(int, v);
    v += __builtin_bit_cast(float, CSG_DPP_ROW_SHL(b, 4));
    b = __builtin_bit_cast(int, v);
    v += __builtin_bit_cast(float, CSG_DPP_ROW_SHL(b, 8));
    b              = __builtin_bit_cast(int, v);
    const float t0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 0));
    const float t1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 16));
    const float t2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 32));
    const float t3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 48));
    return (t0 + t1) + (t2 + t3);
}

// Wave sum, in every lane: row totals by DPP, then the four rows' totals
// (lanes 15, 31, 47, 63).  Every lane of the wave must be active.
__device__ inline int wave_sum(int v)
{
    v += CSG_DPP_ROW_SHR(v, 1);
    v += CSG_DPP_ROW_SHR(v, 2);
    v += CSG_DPP_ROW_SHR(v, 4);
    v += CSG_DPP_ROW_SHR(v, 8);
    return __builtin_amdgcn_readlane(v, 15) + __builtin_amdgcn_readlane(v, 31) +
           __builtin_amdgcn_readlane(v, 47) + __builtin_amdgcn_readlane(v, 63);
}

__device__ inline long long wave_sum64(long long v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1)
        v += __shfl_xor(v, d, kWave);
    return v;
}

// Wave minimum / maximum, in every lane: row results by DPP (a lane whose
// source lies outside its row compares with itself), then the four rows'
// results.  Every lane of the wave must be active.
__device__ inline float wave_min(float v)
{
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
        const int b = __builtin_bit_cast(int, v);
        int       t;
        switch (s)
        {
        case 0: t = CSG_DPP_ROW_SHR_KEEP(b, 1); break;
        case 1: t = CSG_DPP_ROW_SHR_KEEP(b, 2); break;
        case 2: t = CSG_DPP_ROW_SHR_KEEP(b, 4); break;
        default: t = CSG_DPP_ROW_SHR_KEEP(b, 8); break;
        }
        v = fminf(v, __builtin_bit_cast(float, t));
    }
    const int b = __builtin_bit_cast(int, v);
    return fminf(fminf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 15)),
                       __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 31))),
                 fminf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 47)),
                       __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 63))));
}

__device__ inline float wave_max(float v)
{
#pragma unroll
    for (int s = 0; s < 4; ++s)
    {
        const int b = __builtin_bit_cast(int, v);
        int       t;
        switch (s)
        {
        case 0: t = CSG_DPP_ROW_SHR_KEEP(b, 1); break;
        case 1: t = CSG_DPP_ROW_SHR_KEEP(b, 2); break;
        case 2: t = CSG_DPP_ROW_SHR_KEEP(b, 4); break;
        default: t = CSG_DPP_ROW_SHR_KEEP(b, 8); break;
        }
        v = fmaxf(v, __builtin_bit_cast(float, t));
    }
    const int b = __builtin_bit_cast(int, v);
    return fmaxf(fmaxf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 15)),
                       __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 31))),
                 fmaxf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 47)),
                       __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 63))));
}

// Correlation from the three FP32 sums, in the reference's operation order:
// divisor = sqrt(sxx * syy); divisor != 0 ? sxy / divisor : NaN
// (src/cusingler.cu:849-854, :940-944).  sqrtf and '/' are correctly rounded
// in HIP device code (IEEE single, as the reference's nvcc defaults).
__device__ inline float corr_from_sums(float sxy, float sxx, float syy)
{
    const float prod = sxx * syy;
    const float div  = sqrtf(prod);
    return div != 0.f ? sxy / div : __builtin_nanf("");
}

// Block-wide exclusive scan of `v` for a block of NT threads (NT a multiple of
// 64, at most 1024); `tmp` holds NT/64 ints of LDS.
template <int NT> __device__ inline int block_excl_scan(int v, int* tmp, int* total)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wid  = threadIdx.x / kWave;
    int       inc  = wave_incl_scan(v);
    if (lane == kWave - 1)
        tmp[wid] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w)
    {
        const int t = tmp[w];
        if (w < wid)
            off += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// Turns per-code counts in bins[0..nbins) into doubled centred ranks in place:
// X(c) = 2*start(c) + cnt(c) + 1 - (G + 1), start(c) = sum of counts below c.
// `base` is added to every start (the implicit-zero count when code 0 is not
// in the bins).
template <int NT> __device__ inline void bins_to_ranks(int* bins, int nbins, int base, int G, int* tmp)
{
    const int per = (nbins + NT - 1) / NT;
    const int b0  = threadIdx.x * per;
    const int b1  = min(b0 + per, nbins);
    int       s   = 0;
    for (int i = b0; i < b1; ++i)
        s += bins[i];
    int total;
    int start = block_excl_scan<NT>(s, tmp, &total) + base;
    for (int i = b0; i < b1; ++i)
    {
        const int c = bins[i];
        bins[i]     = 2 * start + c - G;
        start += c;
    }
    __syncthreads();
}

// ------------------------------------------------- score-step rank rows ---

// 1/sqrt(s) of an exact sum of squares s < 2^53: the conversion is exact and
// the root and the division are correctly rounded IEEE double operations, so
// host and device give the same bits (+inf for s = 0).
__host__ __device__ inline double inv_root(long long s)
{
    return 1.0 / sqrt((double)s);
}

// The reference's FP32 sum of squares of one rank row: `sumxx += x*x` in gene
// order (src/cusingler.cu:931-938), contracted to fma, by one thread.
__device__ inline float seq_sum_sq(const int16_t* row, uint32_t G)
{
    float    acc = 0.f;
    uint32_t g   = 0;
    for (; g + 8 <= G; g += 8)
    {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            v[u] = (float)row[g + u];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            acc = __builtin_fmaf(v[u], v[u], acc);
    }
    for (; g < G; ++g)
    {
        const float v = (float)row[g];
        acc           = __builtin_fmaf(v, v, acc);
    }
    return acc;
}

// One block per row of a coded CSR matrix over G genes.  Writes the dense
// doubled centred ranks X[row, 0..Kp) (zero past G), the inverse root of the
// exact int64 sum of X^2, and where asked (non-null) the FP32 sum of X^2 of
// seq_sum_sq and the network features BF16(X * inv_g).
template <int NT>
__global__ __launch_bounds__(NT) void k_rank_csr(const uint16_t* __restrict__ data,
                                                 const int* __restrict__ indices,
                                                 const int* __restrict__ indptr, uint32_t row0,
                                                 uint32_t nrows, uint32_t G, uint32_t Kp,
                                                 int nbins, int16_t* __restrict__ X,
                                                 float* __restrict__ sq_f32,
                                                 double* __restrict__ ir_exact,
                                                 __bf16* __restrict__ feat, float inv_g)
{
    extern __shared__ int smem[];
    int*     bins = smem;                          // nbins
    int16_t* row  = (int16_t*)(smem + ((nbins + 3) & ~3));  // Kp
    __shared__ int       tmp[NT / kWave];
    __shared__ long long red[NT / kWave];

    const uint32_t r = blockIdx.x;
    if (r >= nrows)
        return;
    const int beg = indptr[row0 + r];
    const int end = indptr[row0 + r + 1];
    const int nnz = end - beg;
    for (int i = threadIdx.x; i < nbins; i += NT)
        bins[i] = 0;
    __syncthreads();
    for (int i = beg + threadIdx.x; i < end; i += NT)
        atomicAdd(&bins[data[i]], 1);
    __syncthreads();
    // code 0 (implicit zeros) is not stored; its count is G - nnz
    if (threadIdx.x == 0)
        bins[0] = 0;
    __syncthreads();
    bins_to_ranks<NT>(bins, nbins, (int)G - nnz, (int)G, tmp);
    const int16_t x0 = (int16_t)(-nnz);  // 2*0 + (G-nnz) - G
    for (uint32_t g = threadIdx.x; g < Kp; g += NT)
        row[g] = g < G ? x0 : (int16_t)0;
    __syncthreads();
    for (int i = beg + threadIdx.x; i < end; i += NT)
        row[indices[i]] = (int16_t)bins[data[i]];
    __syncthreads();
    int16_t* out = X + (size_t)r * Kp;
    for (uint32_t g = threadIdx.x * 8; g < Kp; g += NT * 8)
        *(int4*)(out + g) = *(const int4*)(row + g);
    if (feat)
        for (uint32_t g = threadIdx.x; g < Kp; g += NT)
            feat[(size_t)r * Kp + g] = (__bf16)((float)row[g] * inv_g);
    long long s = 0;
    for (uint32_t g = threadIdx.x; g < G; g += NT)
    {
        const long long v = row[g];
        s += v * v;
    }
    s = wave_sum64(s);
    if ((threadIdx.x & (kWave - 1)) == 0)
        red[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        long long t = 0;
        for (int w = 0; w < NT / kWave; ++w)
            t += red[w];
        ir_exact[r] = inv_root(t);
        if (sq_f32)
            sq_f32[r] = seq_sum_sq(row, G);
    }
}

// The same outputs with one wave per row (four rows per 256-thread block,
// rows strided over the grid, no block barrier): per-wave LDS of `wave_ints`
// ints holds the code bins and the row.
__global__ __launch_bounds__(256) void k_rank_csr_w(
    const uint16_t* __restrict__ data, const int* __restrict__ indices,
    const int* __restrict__ indptr, uint32_t row0, uint32_t nrows, uint32_t G, uint32_t Kp,
    int nbins, int wave_ints, int16_t* __restrict__ X, float* __restrict__ sq_f32,
    double* __restrict__ ir_exact, __bf16* __restrict__ feat, float inv_g)
{
    extern __shared__ int smem[];
    const int wid  = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    int*      bins = smem + wid * wave_ints;
    int16_t*  row  = (int16_t*)(bins + ((nbins + 3) & ~3));
    for (uint32_t r = blockIdx.x * 4 + wid; r < nrows; r += gridDim.x * 4)
    {
        const int beg = indptr[row0 + r];
        const int end = indptr[row0 + r + 1];
        const int nnz = end - beg;
        for (int i = lane; i < nbins; i += 64)
            bins[i] = 0;
        __builtin_amdgcn_wave_barrier();
        for (int i = beg + lane; i < end; i += 64)
            atomicAdd(&bins[data[i]], 1);
        __builtin_amdgcn_wave_barrier();
        if (lane == 0)
            bins[0] = 0;  // code 0 (implicit zeros) is not stored; its count is G - nnz
        __builtin_amdgcn_wave_barrier();
        {
            const int per = (nbins + 63) / 64;
            const int b0 = min(lane * per, nbins), b1 = min(b0 + per, nbins);
            int       sc = 0;
            for (int i = b0; i < b1; ++i)
                sc += bins[i];
            int start = wave_incl_scan(sc) - sc + ((int)G - nnz);
            for (int i = b0; i < b1; ++i)
            {
                const int c = bins[i];
                bins[i]     = 2 * start + c - (int)G;
                start += c;
            }
        }
        const int16_t x0 = (int16_t)(-nnz);  // 2*0 + (G-nnz) - G
        for (uint32_t g = lane; g < Kp; g += 64)
            row[g] = g < G ? x0 : (int16_t)0;
        __builtin_amdgcn_wave_barrier();
        for (int i = beg + lane; i < end; i += 64)
            row[indices[i]] = (int16_t)bins[data[i]];
        __builtin_amdgcn_wave_barrier();
        int16_t* out = X + (size_t)r * Kp;
        for (uint32_t g = lane * 8; g < Kp; g += 64 * 8)
            *(int4*)(out + g) = *(const int4*)(row + g);
        if (feat)
            for (uint32_t g = lane * 8; g < Kp; g += 64 * 8)
            {
                const int4     v = *(const int4*)(row + g);
                const int16_t* x = (const int16_t*)&v;
                __bf16         h[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    h[u] = (__bf16)((float)x[u] * inv_g);
                *(int4*)(feat + (size_t)r * Kp + g) = *(const int4*)h;
            }
        long long s = 0;
        for (uint32_t g = lane; g < G; g += 64)
        {
            const long long v = row[g];
            s += v * v;
        }
        s = wave_sum64(s);
        if (lane == 0)
        {
            ir_exact[r] = inv_root(s);
            if (sq_f32)
                sq_f32[r] = seq_sum_sq(row, G);
        }
        __builtin_amdgcn_wave_barrier();  // LDS reads done before the next row
    }
}

// The reference's FP32 sums of squares of the listed rows of a rank matrix.
__global__ void k_sq_seq(const int16_t* __restrict__ X, uint32_t Kp, uint32_t G,
                         const uint32_t* __restrict__ rows, const uint32_t* __restrict__ n_rows,
                         float* __restrict__ sq_f32)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_rows[0])
        sq_f32[rows[i]] = seq_sum_sq(X + (size_t)rows[i] * Kp, G);
}

// ------------------------------------------- plain FP32 score contraction ---

// S[m, n] = corr(Xq[m], Xr[n]) with sum_k Xq[m,k]*Xr[n,k] accumulated by fma in
// k order 0..Kp-1 (zero padding past G adds exact zeros), i.e. the reference's
// sequential loop (src/cusingler.cu:931-938).  128x128 tile per 256-thread
// block, 8x8 outputs per thread, K staged through LDS 16 at a time.  The row
// map gives the operand rows of the tile (a_row / b_row, always inside the
// operand buffers), which results are stored (row_ok / col_ok), the FP32 sums
// of squares and the output rows.
constexpr int PG_BM = 128, PG_BN = 128, PG_BK = 16;

struct DensePMap
{
    int          m0, n0, nq, nr;
    const float* sq_q;
    const float* sq_r;
    float*       S;
    size_t       ldS;
    __device__ size_t a_row(int r) const { return (size_t)(m0 + r); }
    __device__ size_t b_row(int r) const { return (size_t)(n0 + r); }
    __device__ bool   row_ok(int m) const { return m0 + m < nq; }
    __device__ float  row_sq(int m) const { return sq_q[m0 + m]; }
    __device__ float* row_out(int m) const { return S + (size_t)(m0 + m) * ldS + n0; }
    __device__ bool   col_ok(int n) const { return n0 + n < nr; }
    __device__ float  col_sq(int n) const { return sq_r[n0 + n]; }
};

// One label of a grouped contraction: the query rows listed in cells[0, mcount)
// against the reference rows b0 .. b0 + ncount - 1; row m of the result is
// stored contiguously at out + m * ld.  Rows past the list repeat its last
// row, reference rows past b_max read row b_max.
struct GroupPMap
{
    const uint32_t* cells;
    int             mcount, ncount;
    uint32_t        b0, b_max;
    const float*    sq_q;
    const float*    sq_r;
    float*          out;
    size_t          ld;
    __device__ size_t a_row(int r) const { return cells[r < mcount ? r : mcount - 1]; }
    __device__ size_t b_row(int r) const { return b0 + r < b_max ? b0 + r : b_max; }
    __device__ bool   row_ok(int m) const { return m < mcount; }
    __device__ float  row_sq(int m) const { return sq_q[cells[m]]; }
    __device__ float* row_out(int m) const { return out + (size_t)m * ld; }
    __device__ bool   col_ok(int n) const { return n < ncount; }
    __device__ float  col_sq(int n) const { return sq_r[b0 + n]; }
};

// BM = 128 or 64 query rows per tile (MI = BM / 16 rows per thread).
template <class Map, int BM = PG_BM>
__device__ inline void plain_block(const int16_t* __restrict__ Xq, const int16_t* __restrict__ Xr,
                                   int Kp, const Map& map)
{
    constexpr int MI = BM / 16;
    __shared__ float As[2][PG_BK][BM + 4];
    __shared__ float Bs[2][PG_BK][PG_BN + 4];
    const int tid = threadIdx.x;
    const int tx  = tid & 15;
    const int ty  = tid >> 4;

    float acc[MI][8];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j)
            acc[i][j] = 0.f;

    const int  lr   = tid >> 1;        // tile row loaded by this thread
    const int  lk   = (tid & 1) * 8;   // first of 8 k values
    const bool a_ld = lr < BM;         // the thread also stages a query row
    const int16_t* qa = Xq + map.a_row(a_ld ? lr : 0) * (size_t)Kp + lk;
    const int16_t* rb = Xr + map.b_row(lr) * (size_t)Kp + lk;

    auto stage = [&](int buf, int4 av, int4 bv) {
        const int16_t* a = (const int16_t*)&av;
        const int16_t* b = (const int16_t*)&bv;
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            if (a_ld)
                As[buf][lk + u][lr] = (float)a[u];
            Bs[buf][lk + u][lr] = (float)b[u];
        }
    };

    int4 av = a_ld ? *(const int4*)qa : int4{ 0, 0, 0, 0 };
    int4 bv = *(const int4*)rb;
    stage(0, av, bv);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < Kp; k0 += PG_BK)
    {
        const bool more = k0 + PG_BK < Kp;
        if (more)
        {
            if (a_ld)
                av = *(const int4*)(qa + k0 + PG_BK);
            bv = *(const int4*)(rb + k0 + PG_BK);
        }
#pragma unroll 4
        for (int kk = 0; kk < PG_BK; ++kk)
        {
            float a[MI];
            {
                const float4 a0 = *(const float4*)&As[buf][kk][ty * 4];
                a[0] = a0.x, a[1] = a0.y, a[2] = a0.z, a[3] = a0.w;
                if constexpr (MI == 8)
                {
                    const float4 a1 = *(const float4*)&As[buf][kk][64 + ty * 4];
                    a[4] = a1.x, a[5] = a1.y, a[6] = a1.z, a[7] = a1.w;
                }
            }
            const float4 b0 = *(const float4*)&Bs[buf][kk][tx * 4];
            const float4 b1 = *(const float4*)&Bs[buf][kk][64 + tx * 4];
            const float  b[8] = { b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w };
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    acc[i][j] = __builtin_fmaf(a[i], b[j], acc[i][j]);
        }
        if (more)
        {
            stage(buf ^ 1, av, bv);
            __syncthreads();
            buf ^= 1;
        }
    }

    // the thread's row and column sums of squares, all loaded before any is
    // used (a row / column outside the tile reads row / column 0)
    float rsq[MI], csq[8];
#pragma unroll
    for (int i = 0; i < MI; ++i)
    {
        const int m = i < 4 ? ty * 4 + i : 64 + ty * 4 + (i - 4);
        rsq[i]      = map.row_sq(map.row_ok(m) ? m : 0);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        const int n = j < 4 ? tx * 4 + j : 64 + tx * 4 + (j - 4);
        csq[j]      = map.col_sq(map.col_ok(n) ? n : 0);
    }
#pragma unroll
    for (int i = 0; i < MI; ++i)
    {
        const int m = i < 4 ? ty * 4 + i : 64 + ty * 4 + (i - 4);
        if (!map.row_ok(m))
            continue;
        float* out = map.row_out(m);
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            const int n = j < 4 ? tx * 4 + j : 64 + tx * 4 + (j - 4);
            if (map.col_ok(n))
                out[n] = corr_from_sums(acc[i][j], rsq[i], csq[j]);
        }
    }
}

__global__ __launch_bounds__(256) void k_gemm_plain(const int16_t* __restrict__ Xq,
                                                    const int16_t* __restrict__ Xr, int Kp,
                                                    const float* __restrict__ sq_q,
                                                    const float* __restrict__ sq_r, int nq, int nr,
                                                    float* __restrict__ S, size_t ldS)
{
    const DensePMap map{ (int)blockIdx.y * PG_BM, (int)blockIdx.x * PG_BN, nq, nr, sq_q, sq_r, S,
                         ldS };
    plain_block(Xq, Xr, Kp, map);
}

// Query rows per tile of the grouped FP32 contraction (its lists are short).
constexpr int PG_GBM = 64;

// Tile (label, first cell, first reference row) of a grouped contraction.
struct GroupTile
{
    uint32_t label, m0, n0, pad_;
};

// Grouped FP32 contraction in the reference's order (PG_GBM x PG_BN tiles):
// block b takes tile tiles[b] = (label l, m0, n0), the cells cells[a_base[l] + m0 ..] (a_cnt[l]
// in all) against reference rows lstart[l] + n0 .. of label l (llen[l] in
// all); each cell's scores against the label's rows are stored contiguously,
// label l's block starting at S + c_off[l].
__global__ __launch_bounds__(256) void k_gemm_plain_group(
    const GroupTile* __restrict__ tiles, const uint32_t* __restrict__ cells,
    const uint32_t* __restrict__ a_base, const uint32_t* __restrict__ a_cnt,
    const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ llen,
    const uint64_t* __restrict__ c_off, const int16_t* __restrict__ Xq,
    const int16_t* __restrict__ Xr, int Kp, uint32_t b_max, const float* __restrict__ sq_q,
    const float* __restrict__ sq_r, float* __restrict__ S, const uint32_t* __restrict__ ntiles)
{
    // any grid over the device-built tile list (one tile per block when the
    // grid is large enough)
    const uint32_t T = ntiles[0];
    for (uint32_t k = blockIdx.x; k < T; k += gridDim.x)
    {
        const GroupTile t  = tiles[k];
        const uint32_t  nl = llen[t.label];
        const GroupPMap map{ cells + a_base[t.label] + t.m0,
                             (int)(a_cnt[t.label] - t.m0),
                             (int)(nl - t.n0),
                             lstart[t.label] + t.n0,
                             b_max,
                             sq_q,
                             sq_r,
                             S + c_off[t.label] + (size_t)t.m0 * nl + t.n0,
                             nl };
        plain_block<GroupPMap, PG_GBM>(Xq, Xr, Kp, map);
        __syncthreads();  // the next tile reuses the LDS buffers
    }
}

// ------------------------------------------------ per-label quantile select ---

// A segment of scores whose per-label quantile is wanted.
struct Segment
{
    uint64_t off;    // first score
    uint32_t n;      // number of scores
    uint32_t param;  // index into the PctParam table
};

// Segment source for the score step: item = q * L + l reads
// S[q * ldS + lstart[l] .. + llen[l]].
struct ScoreSegments
{
    const uint32_t* lstart;
    const uint32_t* llen;
    uint32_t        L;
    uint64_t        ldS;
    __device__ Segment get(uint64_t item) const
    {
        const uint32_t it = (uint32_t)item;  // a chunk holds fewer than 2^32 items
        const uint32_t q  = it / L;
        const uint32_t l  = it - q * L;
        return Segment{ (uint64_t)q * ldS + lstart[l], llen[l], l };
    }
};

struct TableSegments
{
    const Segment* seg;
    __device__ Segment get(uint64_t item) const { return seg[item]; }
};

// Order statistics a_(k0) and a_(k1) (0-based, ascending, k0 <= k1) of a NaN
// free segment, found by one wave: min/max, then a 256-bucket histogram whose
// bucket index is monotone in the value, then exact ranking of the values in
// the buckets that hold the two ranks.  While those buckets hold more than
// kCand values the search narrows to their value range and repeats; the last
// level ranks the remaining candidates against the whole segment.
constexpr int kBuckets  = 256;
constexpr int kCand     = 256;
constexpr int kMaxLevel = 5;

struct WaveSelectSmem
{
    int   hist[kBuckets];
    float cand[kCand];
    int   ncand;
};

__device__ inline int bucket_of(float x, float lo, double dscale)
{
    int b = (int)(((double)x - (double)lo) * dscale);
    return b < 0 ? 0 : (b >= kBuckets ? kBuckets - 1 : b);
}

__device__ inline void wave_select2(const float* __restrict__ v, uint32_t n, uint32_t k0,
                                    uint32_t k1, WaveSelectSmem& sm, float* out0, float* out1)
{
    const int lane = threadIdx.x & (kWave - 1);
    float     lo   = __builtin_inff();
    float     hi   = -__builtin_inff();
    for (uint32_t i = lane; i < n; i += kWave)
    {
        const float x = v[i];
        lo            = fminf(lo, x);
        hi            = fmaxf(hi, x);
    }
    lo             = wave_min(lo);
    hi             = wave_max(hi);
    uint32_t below = 0;  // number of segment values below lo
    float    r0 = lo, r1 = lo;
    for (int level = 0;; ++level)
    {
        if (lo == hi)
        {
            r0 = r1 = lo;
            break;
        }
        const double dscale = (double)kBuckets / ((double)hi - (double)lo);
        for (int b = lane; b < kBuckets; b += kWave)
            sm.hist[b] = 0;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = lane; i < n; i += kWave)
        {
            const float x = v[i];
            if (x >= lo && x <= hi)
                atomicAdd(&sm.hist[bucket_of(x, lo, dscale)], 1);
        }
        __builtin_amdgcn_wave_barrier();
        int c[4], s = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            c[u] = sm.hist[lane * 4 + u];
            s += c[u];
        }
        int ex = wave_incl_scan(s) - s + (int)below;
        int b0 = -1, b1 = -1;
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            if ((uint32_t)ex <= k0 && k0 < (uint32_t)(ex + c[u]))
                b0 = lane * 4 + u;
            if ((uint32_t)ex <= k1 && k1 < (uint32_t)(ex + c[u]))
                b1 = lane * 4 + u;
            ex += c[u];
        }
        const unsigned long long m0 = __ballot(b0 >= 0);
        const unsigned long long m1 = __ballot(b1 >= 0);
        const int blo = __shfl(b0, m0 ? __builtin_ctzll(m0) : 0, kWave);
        const int bhi = __shfl(b1, m1 ? __builtin_ctzll(m1) : 0, kWave);
        int cb = 0, cc = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int b = lane * 4 + u;
            if (b < blo)
                cb += c[u];
            else if (b <= bhi)
                cc += c[u];
        }
        cb = wave_sum(cb);
        cc = wave_sum(cc);
        const bool last = cc <= kCand || level >= kMaxLevel;
        if (last)
        {
            bool  hit0 = false, hit1 = false;
            float v0 = 0.f, v1 = 0.f;
            if (cc <= kCand)
            {
                if (lane == 0)
                    sm.ncand = 0;
                __builtin_amdgcn_wave_barrier();
                for (uint32_t i = lane; i < n; i += kWave)
                {
                    const float x = v[i];
                    if (x >= lo && x <= hi)
                    {
                        const int b = bucket_of(x, lo, dscale);
                        if (b >= blo && b <= bhi)
                            sm.cand[atomicAdd(&sm.ncand, 1)] = x;
                    }
                }
                __builtin_amdgcn_wave_barrier();
                const uint32_t base = below + (uint32_t)cb;
                for (int i = lane; i < cc; i += kWave)
                {
                    const float x    = sm.cand[i];
                    uint32_t    less = 0, eq = 0;
                    for (int j = 0; j < cc; ++j)
                    {
                        const float y = sm.cand[j];
                        less += y < x;
                        eq += y == x;
                    }
                    const uint32_t p0 = base + less, p1 = p0 + eq;
                    if (p0 <= k0 && k0 < p1)
                    {
                        hit0 = true;
                        v0   = x;
                    }
                    if (p0 <= k1 && k1 < p1)
                    {
                        hit1 = true;
                        v1   = x;
                    }
                }
            }
            else
            {
                // many candidates after kMaxLevel narrowings: rank each one
                // against the whole segment
                for (uint32_t i = lane; i < n; i += kWave)
                {
                    const float x = v[i];
                    if (!(x >= lo && x <= hi))
                        continue;
                    const int b = bucket_of(x, lo, dscale);
                    if (b < blo || b > bhi)
                        continue;
                    uint32_t less = 0, eq = 0;
                    for (uint32_t j = 0; j < n; ++j)
                    {
                        const float y = v[j];
                        less += y < x;
                        eq += y == x;
                    }
                    const uint32_t p0 = less, p1 = less + eq;
                    if (p0 <= k0 && k0 < p1)
                    {
                        hit0 = true;
                        v0   = x;
                    }
                    if (p0 <= k1 && k1 < p1)
                    {
                        hit1 = true;
                        v1   = x;
                    }
                }
            }
            const unsigned long long f0 = __ballot(hit0);
            const unsigned long long f1 = __ballot(hit1);
            r0 = __shfl(v0, f0 ? __builtin_ctzll(f0) : 0, kWave);
            r1 = __shfl(v1, f1 ? __builtin_ctzll(f1) : 0, kWave);
            break;
        }
        float nlo = __builtin_inff(), nhi = -__builtin_inff();
        for (uint32_t i = lane; i < n; i += kWave)
        {
            const float x = v[i];
            if (x >= lo && x <= hi)
            {
                const int b = bucket_of(x, lo, dscale);
                if (b >= blo && b <= bhi)
                {
                    nlo = fminf(nlo, x);
                    nhi = fmaxf(nhi, x);
                }
            }
        }
        below += (uint32_t)cb;
        lo = wave_min(nlo);
        hi = wave_max(nhi);
        __builtin_amdgcn_wave_barrier();
    }
    *out0 = r0;
    *out1 = r1;
}

// Register-resident variant for segments of at most 64 * R values: every
// lane holds its values in registers (slot u is element lane + 64 u), the
// range [lo, hi] is cut into 64 buckets (one per lane, monotone in value),
// the two target ranks are located by a prefix over the bucket counts, and the
// values of the buckets that hold them are ranked exactly across the lanes.  While those
// buckets hold more than 64 values the range narrows to them and repeats; a
// last level ranks the candidates against the whole segment.
template <int R>
__device__ inline void wave_select2_reg(const float (&x)[R], const float* __restrict__ v,
                                        uint32_t n, float lo, float hi, uint32_t k0, uint32_t k1,
                                        WaveSelectSmem& sm, float* out0, float* out1)
{
    const int lane  = threadIdx.x & (kWave - 1);
    uint32_t  below = 0;
    float    r0 = lo, r1 = lo;
    for (int level = 0;; ++level)
    {
        if (lo == hi)
        {
            r0 = r1 = lo;
            break;
        }
        const double dscale = 64.0 / ((double)hi - (double)lo);
        sm.hist[lane]       = 0;
        __builtin_amdgcn_wave_barrier();
        int b[R];
#pragma unroll
        for (int u = 0; u < R; ++u)
        {
            const uint32_t i = lane + 64 * u;
            b[u]             = -1;
            if (i < n && x[u] >= lo && x[u] <= hi)
            {
                int t = (int)(((double)x[u] - (double)lo) * dscale);
                b[u]  = t < 0 ? 0 : (t > 63 ? 63 : t);
                atomicAdd(&sm.hist[b[u]], 1);
            }
        }
        __builtin_amdgcn_wave_barrier();
        const int c    = sm.hist[lane];
        const int incl = wave_incl_scan(c);
        const int ex   = incl - c + (int)below;
        const unsigned long long m0 = __ballot((uint32_t)ex <= k0 && k0 < (uint32_t)(ex + c));
        const unsigned long long m1 = __ballot((uint32_t)ex <= k1 && k1 < (uint32_t)(ex + c));
        const int blo = m0 ? __builtin_ctzll(m0) : 0;
        const int bhi = m1 ? __builtin_ctzll(m1) : 63;
        // values below bucket blo, and in buckets blo .. bhi
        const int cb = __builtin_amdgcn_readlane(ex, blo) - (int)below;
        const int cc = __builtin_amdgcn_readlane(incl, bhi) - __builtin_amdgcn_readlane(incl - c, blo);
        const bool last = cc <= kWave || level >= kMaxLevel;
        if (last)
        {
            bool  hit0 = false, hit1 = false;
            float v0 = 0.f, v1 = 0.f;
            if (cc <= kWave)
            {
                // the candidates compacted into lanes 0 .. cc - 1 (ballot
                // prefix counts give the slots), then each ranked against
                // all of them by reading the other lanes' values
                uint32_t pos = 0;
#pragma unroll
                for (int u = 0; u < R; ++u)
                {
                    const bool               f  = b[u] >= blo && b[u] <= bhi;
                    const unsigned long long mk = __ballot(f);
                    if (f)
                        sm.cand[pos + __builtin_amdgcn_mbcnt_hi(
                                          (uint32_t)(mk >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u))] = x[u];
                    pos += (uint32_t)__builtin_popcountll(mk);
                }
                __builtin_amdgcn_wave_barrier();
                const uint32_t base = below + (uint32_t)cb;
                const float    xv   = lane < cc ? sm.cand[lane] : 0.f;
                const int      xb   = __builtin_bit_cast(int, xv);
                uint32_t       less = 0, eq = 0;
                for (int j = 0; j < cc; ++j)
                {
                    const float y = __builtin_bit_cast(float, __builtin_amdgcn_readlane(xb, j));
                    less += y < xv;
                    eq += y == xv;
                }
                const uint32_t p0 = base + less, p1 = p0 + eq;
                if (lane < cc && p0 <= k0 && k0 < p1)
                {
                    hit0 = true;
                    v0   = xv;
                }
                if (lane < cc && p0 <= k1 && k1 < p1)
                {
                    hit1 = true;
                    v1   = xv;
                }
            }
            else
            {
#pragma unroll
                for (int u = 0; u < R; ++u)
                {
                    if (!(b[u] >= blo && b[u] <= bhi))
                        continue;
                    uint32_t less = 0, eq = 0;
                    for (uint32_t j = 0; j < n; ++j)
                    {
                        const float y = v[j];
                        less += y < x[u];
                        eq += y == x[u];
                    }
                    if (less <= k0 && k0 < less + eq)
                    {
                        hit0 = true;
                        v0   = x[u];
                    }
                    if (less <= k1 && k1 < less + eq)
                    {
                        hit1 = true;
                        v1   = x[u];
                    }
                }
            }
            const unsigned long long f0 = __ballot(hit0);
            const unsigned long long f1 = __ballot(hit1);
            r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v0),
                                                                     f0 ? __builtin_ctzll(f0) : 0));
            r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v1),
                                                                     f1 ? __builtin_ctzll(f1) : 0));
            break;
        }
        float nlo = __builtin_inff(), nhi = -__builtin_inff();
#pragma unroll
        for (int u = 0; u < R; ++u)
            if (b[u] >= blo && b[u] <= bhi)
            {
                nlo = fminf(nlo, x[u]);
                nhi = fmaxf(nhi, x[u]);
            }
        below += (uint32_t)cb;
        lo = wave_min(nlo);
        hi = wave_max(nhi);
        __builtin_amdgcn_wave_barrier();
    }
    *out0 = r0;
    *out1 = r1;
}

// Reference percentile of a NaN-free segment from its order statistics.
__device__ inline float pct_from_stats(const PctParam& pp, float p, float a_km1, float a_k,
                                       float vmin, float vmax)
{
    if (pp.kind == 1)
        return vmin;
    if (pp.kind == 2)
        return vmax;
    return pct_interp(p, pp.prev_index, pp.step, a_km1, a_k);
}

// A segment of at most 64 R scores: load into R register slots, send a
// segment with a NaN to the NaN list, otherwise select its order statistics.
template <int R>
__device__ inline void select_segment_reg(const Segment& sg, const float* __restrict__ v,
                                          const PctParam& pp, float p, uint64_t item,
                                          float* __restrict__ out, uint64_t* __restrict__ nan_list,
                                          unsigned int* __restrict__ nan_count,
                                          WaveSelectSmem& wsm, int lane)
{
    float x[R];
    bool  nan = false;
    float lo = __builtin_inff(), hi = -__builtin_inff();
#pragma unroll
    for (int u = 0; u < R; ++u)
    {
        const uint32_t i = lane + 64 * u;
        x[u]             = i < sg.n ? v[i] : 0.f;
        if (i < sg.n)
        {
            nan |= __builtin_isnan(x[u]);
            lo = fminf(lo, x[u]);
            hi = fmaxf(hi, x[u]);
        }
    }
    if (pp.kind == 0)
    {
        // len <= 1: the reference returns arr.front() unsorted
        if (lane == 0)
            out[item] = sg.n > 0 ? x[0] : __builtin_nanf("");
        return;
    }
    if (__ballot(nan))
    {
        if (lane == 0)
        {
            const unsigned int s = atomicAdd(nan_count, 1u);
            nan_list[s]          = item;
            out[item]            = __builtin_nanf("");
        }
        return;
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    float a0 = 0.f, a1 = 0.f;
    if (pp.kind == 3)
        wave_select2_reg<R>(x, v, sg.n, lo, hi, (uint32_t)(pp.k - 1), (uint32_t)pp.k, wsm, &a0,
                            &a1);
    if (lane == 0)
        out[item] = pct_from_stats(pp, p, a0, a1, lo, hi);
}

// One wave per segment: the reference's percentile (src/cusingler.cu:947-978)
// of the segment's scores.  Segments holding a NaN are appended to `nan_list`
// (the libstdc++ sort emulation runs on them in k_select_nan).
template <class Src>
__device__ inline void select_item(const Src& src, uint64_t item,
                                   const float* __restrict__ scores,
                                   const PctParam* __restrict__ params, float p,
                                   float* __restrict__ out, uint64_t* __restrict__ nan_list,
                                   unsigned int* __restrict__ nan_count, WaveSelectSmem& wsm,
                                   int lane)
{
    const Segment  sg = src.get(item);
    const float*   v  = scores + sg.off;
    const PctParam pp = params[sg.param];
    if (sg.n <= 512)
    {
        // the segment stays in registers for every pass, in the fewest slots
        // (a power of two) that hold it; they are loaded before the
        // parameters are looked at, so both loads are in flight together.
        // (Longer segments take the streaming path below: a register path
        // for them would cost every wave of the kernel the VGPRs of 16
        // slots, and occupancy with them.)
        if (sg.n <= 64)
            select_segment_reg<1>(sg, v, pp, p, item, out, nan_list, nan_count, wsm, lane);
        else if (sg.n <= 128)
            select_segment_reg<2>(sg, v, pp, p, item, out, nan_list, nan_count, wsm, lane);
        else if (sg.n <= 256)
            select_segment_reg<4>(sg, v, pp, p, item, out, nan_list, nan_count, wsm, lane);
        else
            select_segment_reg<8>(sg, v, pp, p, item, out, nan_list, nan_count, wsm, lane);
        return;
    }
    if (pp.kind == 0)
    {
        if (lane == 0)
            out[item] = v[0];
        return;
    }
    bool nan = false;
    for (uint32_t i = lane; i < sg.n; i += kWave)
        nan |= __builtin_isnan(v[i]);
    if (__ballot(nan))
    {
        if (lane == 0)
        {
            const unsigned int s = atomicAdd(nan_count, 1u);
            nan_list[s]          = item;
            out[item]            = __builtin_nanf("");
        }
        return;
    }
    float a0 = 0.f, a1 = 0.f, vmin = 0.f, vmax = 0.f;
    if (pp.kind == 3)
        wave_select2(v, sg.n, (uint32_t)(pp.k - 1), (uint32_t)pp.k, wsm, &a0, &a1);
    else
    {
        float lo = __builtin_inff(), hi = -__builtin_inff();
        for (uint32_t i = lane; i < sg.n; i += kWave)
        {
            lo = fminf(lo, v[i]);
            hi = fmaxf(hi, v[i]);
        }
        vmin = wave_min(lo);
        vmax = wave_max(hi);
    }
    if (lane == 0)
        out[item] = pct_from_stats(pp, p, a0, a1, vmin, vmax);
}

template <class Src>
__global__ __launch_bounds__(256, 8) void k_select(Src src, uint64_t n_items,
                                                const float* __restrict__ scores,
                                                const PctParam* __restrict__ params, float p,
                                                float* __restrict__ out,
                                                uint64_t* __restrict__ nan_list,
                                                unsigned int* __restrict__ nan_count)
{
    __shared__ WaveSelectSmem sm[256 / kWave];
    const int      wid  = threadIdx.x / kWave;
    const int      lane = threadIdx.x & (kWave - 1);
    const uint64_t item = (uint64_t)blockIdx.x * (256 / kWave) + wid;
    if (item >= n_items)
        return;
    select_item(src, item, scores, params, p, out, nan_list, nan_count, sm[wid], lane);
}

// The same over *n_dev items counted on the device: a grid-stride loop, one
// wave per item at a time.
template <class Src>
__global__ __launch_bounds__(256, 8) void k_select_n(Src src, const uint32_t* __restrict__ n_dev,
                                                  const float* __restrict__ scores,
                                                  const PctParam* __restrict__ params, float p,
                                                  float* __restrict__ out,
                                                  uint64_t* __restrict__ nan_list,
                                                  unsigned int* __restrict__ nan_count)
{
    __shared__ WaveSelectSmem sm[256 / kWave];
    // (the wave index is uniform: in scalar registers, so is the item)
    const int      wid  = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int      lane = threadIdx.x & (kWave - 1);
    const uint32_t n    = *n_dev;
    for (uint32_t item = blockIdx.x * (256 / kWave) + wid; item < n;
         item += gridDim.x * (256 / kWave))
        select_item(src, item, scores, params, p, out, nan_list, nan_count, sm[wid], lane);
}

// NaN-carrying segments listed by k_select / k_select_n, sorted in place (the
// segments' scores are not read again), each by one lane with the libstdc++
// introsort emulation; the reference's positions are read after it.  A
// segment whose sort would read out of bounds in the reference sets *status.
template <class Src>
__global__ void k_select_nan_inplace(Src src, const uint64_t* __restrict__ nan_list,
                                     const unsigned int* __restrict__ count,
                                     float* __restrict__ scores,
                                     const PctParam* __restrict__ params, float p,
                                     float* __restrict__ out, int* __restrict__ status)
{
    const unsigned int n = *count;
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint64_t item = nan_list[i];
        const Segment  sg   = src.get(item);
        const PctParam pp   = params[sg.param];
        float*         a    = scores + sg.off;
        if (!std_sort(a, (long)sg.n))
        {
            atomicExch(status, 1);
            out[item] = __builtin_nanf("");
            continue;
        }
        float r;
        if (pp.kind == 1)
            r = a[0];
        else if (pp.kind == 2)
            r = a[sg.n - 1];
        else
            r = pct_interp(p, pp.prev_index, pp.step, a[pp.k - 1], a[pp.k]);
        out[item] = r;
    }
}

// NaN-carrying segments: one lane copies the segment to `work` and runs the
// libstdc++ introsort emulation, then reads the reference's positions.
template <class Src>
__global__ void k_select_nan(Src src, const uint64_t* __restrict__ nan_list, unsigned int count,
                             const float* __restrict__ scores, const PctParam* __restrict__ params,
                             float p, float* __restrict__ out, float* __restrict__ work,
                             const uint64_t* __restrict__ work_off, int* __restrict__ status)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const uint64_t item = nan_list[i];
    const Segment  sg   = src.get(item);
    const PctParam pp   = params[sg.param];
    float*         a    = work + work_off[i];
    for (uint32_t t = 0; t < sg.n; ++t)
        a[t] = scores[sg.off + t];
    if (!std_sort(a, (long)sg.n))
    {
        atomicExch(status, 1);
        out[item] = __builtin_nanf("");
        return;
    }
    float r;
    if (pp.kind == 1)
        r = a[0];
    else if (pp.kind == 2)
        r = a[sg.n - 1];
    else
        r = pct_interp(p, pp.prev_index, pp.step, a[pp.k - 1], a[pp.k]);
    out[item] = r;
}

// First label (last maximum of libstdc++ minmax_element) and candidate mask
// (score >= max - thre, in float) per query cell (src/cusingler.cu:1119-1135).
__global__ void k_label(const float* __restrict__ P, uint32_t nq, uint32_t L, float thre,
                        int32_t* __restrict__ first, float* __restrict__ mask)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    const float* s = P + (size_t)q * L;
    int          imin, imax;
    std_minmax(s, (int)L, &imin, &imax);
    first[q]       = imax;
    const float th = s[imax] - thre;
    for (uint32_t j = 0; j < L; ++j)
        mask[(size_t)q * L + j] = s[j] >= th ? 1.f : 0.f;
}

}  // namespace csg
