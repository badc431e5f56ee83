// Reference training on the device: trainData (src/io.cpp:141-262).
//
// 1. Per label and gene, the median over all of the label's cells with
//    implicit zeros (src/io.cpp:146-163: even count -> the mean, in float, of
//    elements n/2 and n/2 - 1 of the sorted column; odd -> element n/2).  The
//    reference's stored values are regrouped into (label, gene) segments
//    (count, scan, scatter); a segment needs sorting only when a wanted
//    order statistic falls outside the run of zeros.
// 2. Per ordered label pair (i, j), i != j: d = int(round((m_i - m_j) *
//    1e6)) per gene, genes ordered by d descending then gene ascending
//    (:228-236).  The positive genes in that order are the pair's window
//    (ref_train_values); its first 333 enter thre_genes; the first k_pair
//    genes of the whole order (positives, then d = 0 genes by ascending index,
//    then negatives) enter common_genes (:239-257).
// A gene's sort key is ((2^31 - 1 - d) << 32) | gene, ascending.
#pragma once

#include "csg_code.h"

namespace csg {

constexpr int TRAIN_NT = 256;

__device__ inline uint64_t train_key(int d, uint32_t g)
{
    return ((uint64_t)(uint32_t)(0x7FFFFFFF - d) << 32) | g;
}

// d of gene g for the pair (mi, mj): the reference's int(round((v1 - v2) * 1e6))
// -- a float difference, a double product, round half away from zero.
__device__ inline int train_diff(float a, float b)
{
    const float  x = a - b;
    const double y = (double)x * 1e6;
    return (int)__builtin_round(y);
}

__global__ void k_train_fill(uint32_t* __restrict__ p, size_t n, uint32_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Entries per (label, gene) segment; `seg` receives each entry's segment (or
// ~0 for a dropped column).
__global__ __launch_bounds__(TRAIN_NT) void k_train_count(
    const int32_t* __restrict__ indices, const int32_t* __restrict__ indptr, uint32_t rows,
    const uint8_t* __restrict__ label, const int32_t* __restrict__ colmap, uint32_t width,
    uint32_t W, uint32_t* __restrict__ cnt, uint32_t* __restrict__ seg, uint32_t* __restrict__ bad)
{
    const int      lane = threadIdx.x & 63;
    const uint32_t nw   = gridDim.x * (TRAIN_NT / 64);
    for (uint32_t r = blockIdx.x * (TRAIN_NT / 64) + (threadIdx.x >> 6); r < rows; r += nw)
    {
        const uint32_t l = label[r];
        for (int i = indptr[r] + lane; i < indptr[r + 1]; i += 64)
        {
            const int32_t c = indices[i];
            int32_t       m = -1;
            if (c < 0 || (uint32_t)c >= width)
                atomicOr(bad, 1u);
            else
                m = colmap ? colmap[c] : c;
            if (m < 0 || (uint32_t)m >= W)
            {
                seg[i] = ~0u;
                continue;
            }
            const uint32_t s = l * W + (uint32_t)m;
            seg[i]           = s;
            atomicAdd(&cnt[s], 1u);
        }
    }
}

// Tile totals of an exclusive scan (tiles of TRAIN_TILE entries).
constexpr int TRAIN_TILE = 4096;

__global__ __launch_bounds__(1024) void k_train_tile_sum(const uint32_t* __restrict__ v, size_t n,
                                                        uint64_t* __restrict__ tot)
{
    __shared__ uint64_t red[16];
    const size_t t0 = (size_t)blockIdx.x * TRAIN_TILE;
    uint64_t     s  = 0;
    for (int k = 0; k < TRAIN_TILE / 1024; ++k)
    {
        const size_t i = t0 + (size_t)k * 1024 + threadIdx.x;
        s += i < n ? v[i] : 0u;
    }
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint64_t t = 0;
        for (int w = 0; w < 16; ++w)
            t += red[w];
        tot[blockIdx.x] = t;
    }
}

// Exclusive scan of v into out (uint64), each tile from its offset tile_off.
__global__ __launch_bounds__(1024) void k_train_tile_scan(const uint32_t* __restrict__ v, size_t n,
                                                         const uint64_t* __restrict__ tile_off,
                                                         uint64_t* __restrict__ out)
{
    __shared__ uint64_t part[1024];
    const size_t t0  = (size_t)blockIdx.x * TRAIN_TILE;
    const int    per = TRAIN_TILE / 1024;
    const size_t i0  = t0 + (size_t)threadIdx.x * per;
    uint64_t     s   = 0;
    for (int k = 0; k < per; ++k)
        s += i0 + k < n ? v[i0 + k] : 0u;
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1)
    {
        const uint64_t a = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0ull;
        __syncthreads();
        part[threadIdx.x] += a;
        __syncthreads();
    }
    uint64_t run = tile_off[blockIdx.x] + part[threadIdx.x] - s;
    for (int k = 0; k < per; ++k)
        if (i0 + k < n)
        {
            out[i0 + k] = run;
            run += v[i0 + k];
        }
}

// Values into their segments (order inside a segment is free: it is sorted
// where it matters).
__global__ __launch_bounds__(TRAIN_NT) void k_train_scatter(
    const float* __restrict__ data, const uint32_t* __restrict__ seg, size_t nnz,
    const uint64_t* __restrict__ off, uint32_t* __restrict__ fill, float* __restrict__ vals)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
         i += (size_t)gridDim.x * blockDim.x)
    {
        const uint32_t s = seg[i];
        if (s == ~0u)
            continue;
        vals[off[s] + atomicAdd(&fill[s], 1u)] = data[i];
    }
}

__device__ inline float train_key_value(uint64_t k)
{
    const uint32_t u = (uint32_t)(k >> 32);
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Element k of the label's sorted column: stored values sorted in s[0, c)
// (negatives, stored zeros, positives), nz implicit zeros.
__device__ inline float train_order(const uint64_t* s, int c, int neg, int zs, int nz, int k)
{
    if (k < neg)
        return train_key_value(s[k]);
    if (k < neg + zs + nz)
        return 0.f;
    return train_key_value(s[k - nz]);
}

// The median of a (label, gene) segment from the counts and, where needed,
// its sorted stored values `s`.
__device__ inline float train_median(const uint64_t* s, int c, int neg, int zs, int n)
{
    const int nz = n - c;
    if (n % 2 == 0)
    {
        const float hi = train_order(s, c, neg, zs, nz, n / 2);
        const float lo = train_order(s, c, neg, zs, nz, n / 2 - 1);
        return (hi + lo) / 2;
    }
    return train_order(s, c, neg, zs, nz, n / 2);
}

// Whether a segment's median needs its stored values sorted: a wanted order
// statistic lies among the stored negatives or positives.
__device__ inline bool train_needs_sort(int c, int neg, int zs, int n)
{
    const int nz = n - c, z0 = neg, z1 = neg + zs + nz;  // zeros at [z0, z1)
    const int k1 = n / 2, k0 = n % 2 == 0 ? n / 2 - 1 : n / 2;
    return !(k0 >= z0 && k1 < z1);
}

constexpr int TRAIN_WAVE_CAP = 1024;  // stored values a wave sorts in its LDS

// One wave per (label, gene) segment: medians of segments of at most
// TRAIN_WAVE_CAP stored values; longer ones that need sorting are listed in
// `big` for k_train_median_big.
__global__ __launch_bounds__(TRAIN_NT) void k_train_median(
    const uint32_t* __restrict__ cnt, const uint64_t* __restrict__ off,
    const float* __restrict__ vals, const uint32_t* __restrict__ nlab, uint32_t L, uint32_t W,
    float* __restrict__ med, uint32_t* __restrict__ big, uint32_t* __restrict__ n_big)
{
    __shared__ uint64_t skeys[TRAIN_NT / 64][TRAIN_WAVE_CAP];
    const int      lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t*      keys = skeys[wid];
    const uint64_t nseg = (uint64_t)L * W;
    const uint64_t nw   = (uint64_t)gridDim.x * (TRAIN_NT / 64);
    for (uint64_t sg = (uint64_t)blockIdx.x * (TRAIN_NT / 64) + wid; sg < nseg; sg += nw)
    {
        const int      c  = (int)cnt[sg];
        const int      n  = (int)nlab[sg / W];
        const float*   v  = vals + off[sg];
        if (c == 0)
        {
            if (lane == 0)
                med[sg] = 0.f;
            continue;
        }
        int neg = 0, zs = 0;
        for (int i = lane; i < c; i += 64)
        {
            const float x = v[i];
            neg += x < 0.f;
            zs += !(x < 0.f) && !(x > 0.f);
        }
        neg = wave_sum(neg);
        zs  = wave_sum(zs);
        if (!train_needs_sort(c, neg, zs, n))
        {
            if (lane == 0)
                med[sg] = 0.f;
            continue;
        }
        if (c > TRAIN_WAVE_CAP)
        {
            if (lane == 0)
                big[atomicAdd(n_big, 1u)] = (uint32_t)sg;
            continue;
        }
        int n2 = 64;
        while (n2 < c)
            n2 <<= 1;
        for (int i = lane; i < n2; i += 64)
            keys[i] = i < c ? (uint64_t)code_key(v[i]) << 32 : ~0ull;
        __builtin_amdgcn_wave_barrier();
        code_bitonic(keys, n2, lane, 64, [] { __builtin_amdgcn_wave_barrier(); });
        if (lane == 0)
            med[sg] = train_median(keys, c, neg, zs, n);
        __builtin_amdgcn_wave_barrier();  // LDS reads done before the next segment
    }
}

// One block per listed segment, its values sorted in the block's slice of
// global scratch (`slice` keys per block).
__global__ __launch_bounds__(TRAIN_NT) void k_train_median_big(
    const uint32_t* __restrict__ cnt, const uint64_t* __restrict__ off,
    const float* __restrict__ vals, const uint32_t* __restrict__ nlab, uint32_t W,
    const uint32_t* __restrict__ big, const uint32_t* __restrict__ n_big,
    uint64_t* __restrict__ scratch, size_t slice, float* __restrict__ med)
{
    __shared__ int red[2][TRAIN_NT / 64];
    uint64_t*      keys = scratch + (size_t)blockIdx.x * slice;
    const uint32_t nb   = *n_big;
    for (uint32_t it = blockIdx.x; it < nb; it += gridDim.x)
    {
        const uint32_t sg = big[it];
        const int      c  = (int)cnt[sg];
        const int      n  = (int)nlab[sg / W];
        const float*   v  = vals + off[sg];
        int n2 = 64;
        while (n2 < c)
            n2 <<= 1;
        __syncthreads();  // the previous segment's keys are no longer read
        int neg = 0, zs = 0;
        for (int i = threadIdx.x; i < n2; i += TRAIN_NT)
        {
            const float x = i < c ? v[i] : 0.f;
            if (i < c)
            {
                neg += x < 0.f;
                zs += !(x < 0.f) && !(x > 0.f);
            }
            keys[i] = i < c ? (uint64_t)code_key(x) << 32 : ~0ull;
        }
        neg = wave_sum(neg);
        zs  = wave_sum(zs);
        if ((threadIdx.x & 63) == 0)
        {
            red[0][threadIdx.x >> 6] = neg;
            red[1][threadIdx.x >> 6] = zs;
        }
        __syncthreads();
        neg = zs = 0;
        for (int w = 0; w < TRAIN_NT / 64; ++w)
        {
            neg += red[0][w];
            zs += red[1][w];
        }
        code_bitonic(keys, n2, threadIdx.x, TRAIN_NT, [] { __syncthreads(); });
        if (threadIdx.x == 0)
            med[sg] = train_median(keys, c, neg, zs, n);
    }
}

// Positive and zero differences per ordered pair (one block per pair).
__global__ __launch_bounds__(TRAIN_NT) void k_train_pair_count(const float* __restrict__ med,
                                                               uint32_t L, uint32_t W,
                                                               uint32_t* __restrict__ npos,
                                                               uint32_t* __restrict__ nzero)
{
    __shared__ int red[2][TRAIN_NT / 64];
    for (uint32_t pr = blockIdx.x; pr < L * L; pr += gridDim.x)
    {
        const uint32_t i = pr / L, j = pr % L;
        int            p = 0, z = 0;
        if (i != j)
            for (uint32_t g = threadIdx.x; g < W; g += TRAIN_NT)
            {
                const int d = train_diff(med[(size_t)i * W + g], med[(size_t)j * W + g]);
                p += d > 0;
                z += d == 0;
            }
        p = wave_sum(p);
        z = wave_sum(z);
        __syncthreads();  // red of the previous pair has been read
        if ((threadIdx.x & 63) == 0)
        {
            red[0][threadIdx.x >> 6] = p;
            red[1][threadIdx.x >> 6] = z;
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            int tp = 0, tz = 0;
            for (int w = 0; w < TRAIN_NT / 64; ++w)
            {
                tp += red[0][w];
                tz += red[1][w];
            }
            npos[pr]  = (uint32_t)tp;
            nzero[pr] = (uint32_t)tz;
        }
    }
}

// Block-wide exclusive scan of one int per thread; *total receives the sum.
__device__ inline int train_block_scan(int v, int* red, int* total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int incl = wave_incl_scan(v);
    __syncthreads();
    if (lane == 63)
        red[wid] = incl;
    __syncthreads();
    int add = 0, tot = 0;
    for (int w = 0; w < TRAIN_NT / 64; ++w)
    {
        add += w < wid ? red[w] : 0;
        tot += red[w];
    }
    *total = tot;
    return incl - v + add;
}

// Per ordered pair: the sorted positive genes into train_values at
// off[pair], their first thre_n into the thre bitmap, the first k_pair of
// the whole order into the common bitmap.  Keys are sorted in the block's
// scratch slice (`slice` keys).
__global__ __launch_bounds__(TRAIN_NT) void k_train_pair_sort(
    const float* __restrict__ med, uint32_t L, uint32_t W, const uint32_t* __restrict__ npos,
    const uint32_t* __restrict__ nzero, const uint64_t* __restrict__ voff, uint32_t thre_n,
    uint32_t k_pair, uint64_t* __restrict__ scratch, size_t slice,
    uint32_t* __restrict__ train_values, uint32_t* __restrict__ thre_bits,
    uint32_t* __restrict__ common_bits)
{
    __shared__ int red[TRAIN_NT / 64];
    __shared__ int sbase;
    uint64_t*      keys = scratch + (size_t)blockIdx.x * slice;
    for (uint32_t pr = blockIdx.x; pr < L * L; pr += gridDim.x)
    {
        const uint32_t i = pr / L, j = pr % L;
        if (i == j)
            continue;  // uniform per block
        const float* mi = med + (size_t)i * W;
        const float* mj = med + (size_t)j * W;
        const int    P  = (int)npos[pr], Z = (int)nzero[pr];
        // gather the keys of one class (0: positive, 1: negative) in any order
        auto gather = [&](int cls) {
            if (threadIdx.x == 0)
                sbase = 0;
            __syncthreads();
            for (uint32_t g0 = 0; g0 < W; g0 += TRAIN_NT)
            {
                const uint32_t g  = g0 + threadIdx.x;
                int            d  = 0;
                bool           in = false;
                if (g < W)
                {
                    d  = train_diff(mi[g], mj[g]);
                    in = cls == 0 ? d > 0 : d < 0;
                }
                int       tot;
                const int at = train_block_scan(in ? 1 : 0, red, &tot);
                if (in)
                    keys[sbase + at] = train_key(d, g);
                __syncthreads();
                if (threadIdx.x == 0)
                    sbase += tot;
                __syncthreads();
            }
        };
        auto sort = [&](int n) {
            int n2 = 64;
            while (n2 < n)
                n2 <<= 1;
            for (int k = n + (int)threadIdx.x; k < n2; k += TRAIN_NT)
                keys[k] = ~0ull;
            __syncthreads();
            code_bitonic(keys, n2, threadIdx.x, TRAIN_NT, [] { __syncthreads(); });
        };
        if (P > 0)
        {
            gather(0);
            sort(P);
            for (int t = threadIdx.x; t < P; t += TRAIN_NT)
            {
                const uint32_t g = (uint32_t)keys[t];
                train_values[voff[pr] + t] = g;
                if ((uint32_t)t < thre_n)
                    atomicOr(&thre_bits[g >> 5], 1u << (g & 31));
                if ((uint32_t)t < k_pair)
                    atomicOr(&common_bits[g >> 5], 1u << (g & 31));
            }
        }
        if ((uint32_t)P >= k_pair)
        {
            __syncthreads();
            continue;
        }
        // the rest of the first k_pair: d = 0 genes by ascending index, then
        // the best negatives
        const int need = (int)k_pair - P;
        if (threadIdx.x == 0)
            sbase = 0;
        __syncthreads();
        for (uint32_t g0 = 0; g0 < W && sbase < need; g0 += TRAIN_NT)
        {
            const uint32_t g  = g0 + threadIdx.x;
            const bool     in = g < W && train_diff(mi[g], mj[g]) == 0;
            int            tot;
            const int      at = train_block_scan(in ? 1 : 0, red, &tot);
            if (in && sbase + at < need)
                atomicOr(&common_bits[g >> 5], 1u << (g & 31));
            __syncthreads();
            if (threadIdx.x == 0)
                sbase += tot;
            __syncthreads();
        }
        __syncthreads();  // every thread has left the loop (its test reads sbase)
        if (Z < need)
        {
            const int nn = (int)W - P - Z;
            gather(1);
            sort(nn);
            for (int t = threadIdx.x; t < need - Z && t < nn; t += TRAIN_NT)
            {
                const uint32_t g = (uint32_t)keys[t];
                atomicOr(&common_bits[g >> 5], 1u << (g & 31));
            }
        }
        __syncthreads();
    }
}

}  // namespace csg
