// Value coding of a float CSR on the device: csrFilter (src/io.cpp:470-523)
// for the query rows, for both gene sets of the path at once.
//
// Per row and gene set, the entries whose column the set holds are kept in
// row order, re-indexed to the column's position in the (ascending) set, and
// their value replaced by its 1-based rank among the row's distinct kept
// values; the set's max_uniq_gene is the largest (#distinct + 1) over rows.
// -0.0 and +0.0 are one value (they compare equal, as the reference's
// std::set<float> and unordered_map<float> treat them).
//
// Three passes over a slice of rows: k_code_count (kept entries per row and
// set, one wave per row), k_code_scan (the output row pointers, one block),
// k_code_rows (one wave per row: the row's kept (key, slot) pairs sorted in
// the wave's LDS by a bitonic network, dense ranks from a wave scan of the
// distinct-key flags, codes scattered back to their slots).  A row with more
// kept entries than a wave's LDS holds is listed and coded by
// k_code_rows_big (one block per row, the same steps over the block's LDS).
#pragma once

#include "csg_kernels.h"

namespace csg {

// Order-preserving 32-bit key of a float (ascending keys = ascending values
// for every non-NaN value; -0.0 is mapped to +0.0 first).
__device__ inline uint32_t code_key(float v)
{
    const uint32_t u = __float_as_uint(v == 0.f ? 0.f : v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Kept entries per set a wave codes in its LDS (8 bytes each, both sets of a
// row held at once); longer rows go to the block kernel.
constexpr int CODE_WAVE_CAP = 512;
constexpr int CODE_NT       = 256;  // threads of the coding kernels' blocks
constexpr int CODE_UNROLL   = 4;    // groups of 64 entries a wave loads at once

// Flags of a coding pass (bit 0: a NaN among the kept values, bit 1: a column
// outside the query width, bit 2: a row with more distinct values than
// uint16 codes hold).
constexpr uint32_t CODE_NAN = 1, CODE_BADCOL = 2, CODE_OVERFLOW = 4;

struct CodeSet
{
    const int32_t* map;      // query column -> column in the gene set, or -1 (width entries)
    const int32_t* ptr;      // output row pointers (k_code_scan)
    uint16_t*      data;     // output codes
    int32_t*       indices;  // output columns
    uint32_t*      count;    // kept entries per row (k_code_count)
};

// Kept entries per row and set.  Rows [0, nrows) of the CSR given by
// (data, indices, indptr) with indptr relative to `base` (the first entry of
// row 0 sits at data[0]).  Column indices are int32, or uint16 for batches
// uploaded in that narrower form (query widths up to 65536).
template <class IT>
__global__ __launch_bounds__(CODE_NT) void k_code_count(
    const IT* __restrict__ indices,
    const int32_t* __restrict__ indptr, int64_t base, uint32_t nrows, uint32_t width,
    const int32_t* __restrict__ map_a, const int32_t* __restrict__ map_b,
    uint32_t* __restrict__ cnt_a, uint32_t* __restrict__ cnt_b, uint32_t* __restrict__ flags)
{
    const int lane = threadIdx.x & 63;
    const uint32_t nw = gridDim.x * (CODE_NT / 64);
    for (uint32_t r = blockIdx.x * (CODE_NT / 64) + (threadIdx.x >> 6); r < nrows; r += nw)
    {
        const int64_t beg = indptr[r] - base, end = indptr[r + 1] - base;
        int           ca = 0, cb = 0;
        uint32_t      fl = 0;
        for (int64_t i0 = beg; i0 < end; i0 += 64 * CODE_UNROLL)
        {
            int32_t c[CODE_UNROLL];
#pragma unroll
            for (int u = 0; u < CODE_UNROLL; ++u)
            {
                const int64_t i = i0 + u * 64 + lane;
                c[u]            = i < end ? (int32_t)indices[i] : 0;
                if (i < end && (c[u] < 0 || (uint32_t)c[u] >= width))
                {
                    fl |= CODE_BADCOL;
                    c[u] = -1;
                }
                else if (i >= end)
                    c[u] = -1;
            }
#pragma unroll
            for (int u = 0; u < CODE_UNROLL; ++u)
                if (c[u] >= 0)
                {
                    ca += map_a[c[u]] >= 0;
                    cb += map_b[c[u]] >= 0;
                }
        }
        ca = wave_sum(ca);
        cb = wave_sum(cb);
        if (fl)
            atomicOr(flags, fl);
        if (lane == 0)
        {
            cnt_a[r] = (uint32_t)ca;
            cnt_b[r] = (uint32_t)cb;
        }
    }
}

// Exclusive prefix sums of two count arrays into row pointers, from the
// running totals base[0] and base[1] (CSRs that grow slice by slice; both
// are advanced by the slice's totals).  One block.
__global__ __launch_bounds__(1024) void k_code_scan(const uint32_t* __restrict__ cnt_a,
                                                    const uint32_t* __restrict__ cnt_b,
                                                    uint32_t nrows, int32_t* __restrict__ ptr_a,
                                                    int32_t* __restrict__ ptr_b,
                                                    int64_t* __restrict__ base)
{
    __shared__ int64_t part[2][1024];
    const uint32_t t   = threadIdx.x;
    const uint32_t per = (nrows + 1023) / 1024;
    const uint32_t r0 = min(t * per, nrows), r1 = min(r0 + per, nrows);
    int64_t        sa = 0, sb = 0;
    for (uint32_t r = r0; r < r1; ++r)
    {
        sa += cnt_a[r];
        sb += cnt_b[r];
    }
    part[0][t] = sa;
    part[1][t] = sb;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1)
    {
        const int64_t va = t >= off ? part[0][t - off] : 0, vb = t >= off ? part[1][t - off] : 0;
        __syncthreads();
        part[0][t] += va;
        part[1][t] += vb;
        __syncthreads();
    }
    const int64_t ba = base[0], bb = base[1];
    int64_t       ra = ba + part[0][t] - sa, rb = bb + part[1][t] - sb;
    for (uint32_t r = r0; r < r1; ++r)
    {
        ptr_a[r] = (int32_t)ra;
        ptr_b[r] = (int32_t)rb;
        ra += cnt_a[r];
        rb += cnt_b[r];
    }
    __syncthreads();  // every thread has read the bases
    if (t == 1023)
    {
        ptr_a[nrows] = (int32_t)(ba + part[0][1023]);
        ptr_b[nrows] = (int32_t)(bb + part[1][1023]);
        base[0]      = ba + part[0][1023];
        base[1]      = bb + part[1][1023];
    }
}

// Bitonic sort of keys[0, n2) ascending (n2 a power of two) by `nt` threads
// starting at thread `t`; `sync` separates the passes.
template <class Sync>
__device__ inline void code_bitonic(uint64_t* keys, int n2, int t, int nt, Sync sync)
{
    for (int k = 2; k <= n2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1)
        {
            for (int p = t; p < n2 / 2; p += nt)
            {
                const int      i  = (p / j) * 2 * j + (p % j);
                const int      q  = i + j;
                const uint64_t a  = keys[i], b = keys[q];
                const bool     up = (i & k) == 0;
                if ((a > b) == up)
                {
                    keys[i] = b;
                    keys[q] = a;
                }
            }
            sync();
        }
}

// Codes of one row and set from its kept entries sorted in keys[0, n) (key in
// the high word, output slot in the low word): the rank of a key is the
// number of distinct keys up to it.  Positions are split into contiguous
// runs of `per` per thread; `scan(v)` returns the inclusive prefix sum of v
// over the threads in order and `total` the count of distinct keys.
template <class Scan>
__device__ inline uint32_t code_ranks(const uint64_t* keys, int n, int t, int nt, uint16_t* out,
                                      Scan scan)
{
    const int per = (n + nt - 1) / nt;
    const int p0 = min(t * per, n), p1 = min(p0 + per, n);
    int       flags = 0;
    for (int p = p0; p < p1; ++p)
        flags += p == 0 || (keys[p] >> 32) != (keys[p - 1] >> 32);
    const int incl = scan(flags);
    int       rank = incl - flags;
    for (int p = p0; p < p1; ++p)
    {
        rank += p == 0 || (keys[p] >> 32) != (keys[p - 1] >> 32);
        out[(uint32_t)keys[p]] = (uint16_t)rank;
    }
    return (uint32_t)incl;  // the last thread's value is the distinct count
}

// Sorts keys[0, n) of one row and set in the wave's LDS and writes the
// codes: scattered by slot into the wave's LDS row `codes`, then stored to
// `out` in order (coalesced); returns the distinct count (every lane).
__device__ inline uint32_t code_wave_row(uint64_t* keys, uint16_t* codes, int n, int lane,
                                         uint16_t* out)
{
    int n2 = 64;
    while (n2 < n)
        n2 <<= 1;
    for (int k = n + lane; k < n2; k += 64)
        keys[k] = ~0ull;
    __builtin_amdgcn_wave_barrier();
    code_bitonic(keys, n2, lane, 64, [] { __builtin_amdgcn_wave_barrier(); });
    const uint32_t d = code_ranks(keys, n, lane, 64, codes,
                                  [](int v) { return wave_incl_scan(v); });
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < n; k += 64)
        out[k] = codes[k];
    __builtin_amdgcn_wave_barrier();  // LDS reads done before keys and codes are rewritten
    return __builtin_amdgcn_readlane(d, 63);
}

// One wave per row: both sets' codes of rows [0, nrows), the row's entries
// read once for both.  `over` lists the (row, set) pairs -- set in bit 31 --
// whose kept entries exceed the wave's LDS; a NaN among kept values raises
// CODE_NAN (the reference orders only the kept values).
template <class IT>
__global__ __launch_bounds__(CODE_NT) void k_code_rows(
    const float* __restrict__ data, const IT* __restrict__ indices,
    const int32_t* __restrict__ indptr, int64_t base, uint32_t nrows, uint32_t width, CodeSet sa,
    CodeSet sb, uint32_t* __restrict__ max_uniq, uint32_t* __restrict__ over,
    uint32_t* __restrict__ n_over, uint32_t* __restrict__ flags)
{
    __shared__ uint64_t skeys[CODE_NT / 64][2][CODE_WAVE_CAP];
    __shared__ uint16_t scodes[CODE_NT / 64][CODE_WAVE_CAP];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t* ka = skeys[wid][0];
    uint64_t* kb = skeys[wid][1];
    uint16_t* cw = scodes[wid];
    const uint32_t nw = gridDim.x * (CODE_NT / 64);
    const uint64_t below = (1ull << lane) - 1;
    uint32_t       mu[2] = { 1u, 1u };
    uint32_t       fl    = 0;
    for (uint32_t r = blockIdx.x * (CODE_NT / 64) + wid; r < nrows; r += nw)
    {
        const int  na = (int)sa.count[r], nb = (int)sb.count[r];
        const bool wa = na > 0 && na <= CODE_WAVE_CAP, wb = nb > 0 && nb <= CODE_WAVE_CAP;
        if (lane == 0)
        {
            if (na > CODE_WAVE_CAP)
                over[atomicAdd(n_over, 1u)] = r;
            if (nb > CODE_WAVE_CAP)
                over[atomicAdd(n_over, 1u)] = r | 0x80000000u;
        }
        if (!wa && !wb)
            continue;  // uniform per wave
        const int64_t beg = indptr[r] - base, end = indptr[r + 1] - base;
        const int64_t oa = sa.ptr[r], ob = sb.ptr[r];
        // kept entries in row order: columns out, (key, slot) pairs to LDS;
        // CODE_UNROLL groups of 64 entries per pass, their column and map
        // loads issued before any is used
        int slot_a = 0, slot_b = 0;
        for (int64_t i0 = beg; i0 < end; i0 += 64 * CODE_UNROLL)
        {
            int32_t c[CODE_UNROLL], ma[CODE_UNROLL], mb[CODE_UNROLL];
#pragma unroll
            for (int u = 0; u < CODE_UNROLL; ++u)
            {
                const int64_t i = i0 + u * 64 + lane;
                c[u]            = i < end ? (int32_t)indices[i] : -1;
            }
#pragma unroll
            for (int u = 0; u < CODE_UNROLL; ++u)
            {
                const bool ok = c[u] >= 0 && (uint32_t)c[u] < width;
                ma[u]         = ok && wa ? sa.map[c[u]] : -1;
                mb[u]         = ok && wb ? sb.map[c[u]] : -1;
            }
#pragma unroll
            for (int u = 0; u < CODE_UNROLL; ++u)
            {
                const int64_t  i    = i0 + u * 64 + lane;
                const uint64_t bala = __ballot(ma[u] >= 0), balb = __ballot(mb[u] >= 0);
                if (ma[u] >= 0 || mb[u] >= 0)
                {
                    const float    v   = data[i];
                    const uint64_t key = (uint64_t)code_key(v) << 32;
                    if (v != v)
                        fl |= CODE_NAN;
                    if (ma[u] >= 0)
                    {
                        const int k        = slot_a + __popcll(bala & below);
                        sa.indices[oa + k] = ma[u];
                        ka[k]              = key | (uint32_t)k;
                    }
                    if (mb[u] >= 0)
                    {
                        const int k        = slot_b + __popcll(balb & below);
                        sb.indices[ob + k] = mb[u];
                        kb[k]              = key | (uint32_t)k;
                    }
                }
                slot_a += __popcll(bala);
                slot_b += __popcll(balb);
            }
        }
        if (wa)
        {
            const uint32_t d = code_wave_row(ka, cw, na, lane, sa.data + oa);
            mu[0]            = max(mu[0], d + 1);
            if (d > 65535u)
                fl |= CODE_OVERFLOW;
        }
        if (wb)
        {
            const uint32_t d = code_wave_row(kb, cw, nb, lane, sb.data + ob);
            mu[1]            = max(mu[1], d + 1);
            if (d > 65535u)
                fl |= CODE_OVERFLOW;
        }
    }
    if (fl)
        atomicOr(flags, fl);
    // one address takes every wave's maximum: read it first, so that waves
    // whose maximum is not larger (nearly all, once it has grown) skip the
    // serialising atomic
    if (lane == 0)
    {
        if (mu[0] > __atomic_load_n(&max_uniq[0], __ATOMIC_RELAXED))
            atomicMax(&max_uniq[0], mu[0]);
        if (mu[1] > __atomic_load_n(&max_uniq[1], __ATOMIC_RELAXED))
            atomicMax(&max_uniq[1], mu[1]);
    }
}

// Block-wide inclusive scan of one int per thread (CODE_NT threads).
__device__ inline int code_block_scan(int v, int* red)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int incl = wave_incl_scan(v);
    if (lane == 63)
        red[wid] = incl;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wid; ++w)
        add += red[w];
    __syncthreads();
    return incl + add;
}

// One block per listed row (set in bit 31) with more kept entries than a
// wave holds; `cap` keys of dynamic LDS.
template <class IT>
__global__ __launch_bounds__(CODE_NT) void k_code_rows_big(
    const float* __restrict__ data, const IT* __restrict__ indices,
    const int32_t* __restrict__ indptr, int64_t base, uint32_t width, CodeSet sa, CodeSet sb,
    const uint32_t* __restrict__ over, const uint32_t* __restrict__ n_over, int cap,
    uint32_t* __restrict__ max_uniq, uint32_t* __restrict__ flags)
{
    extern __shared__ uint64_t bkeys[];
    __shared__ int red[CODE_NT / 64];
    __shared__ int sslot;
    const uint32_t no = *n_over;
    for (uint32_t it = blockIdx.x; it < no; it += gridDim.x)
    {
        const uint32_t r = over[it] & 0x7FFFFFFFu;
        const int      s = (int)(over[it] >> 31);
        const CodeSet& cs = s ? sb : sa;
        const int      n  = (int)cs.count[r];
        int            n2 = 64;  // the sort network's width
        while (n2 < n)
            n2 <<= 1;
        __syncthreads();  // the previous row's LDS reads are done
        if (n2 > cap)
        {
            if (threadIdx.x == 0)
                atomicOr(flags, CODE_OVERFLOW);
            continue;  // uniform per block
        }
        const int64_t beg = indptr[r] - base, end = indptr[r + 1] - base;
        const int64_t o   = cs.ptr[r];
        if (threadIdx.x == 0)
            sslot = 0;
        __syncthreads();
        for (int64_t i0 = beg; i0 < end; i0 += CODE_NT)
        {
            const int64_t i = i0 + threadIdx.x;
            int32_t       m = -1;
            if (i < end)
            {
                const int32_t c = (int32_t)indices[i];
                m               = c >= 0 && (uint32_t)c < width ? cs.map[c] : -1;
            }
            const int kept = m >= 0;
            const int incl = code_block_scan(kept, red);
            const int at   = sslot;
            if (kept)
            {
                const int   k     = at + incl - 1;
                const float v     = data[i];
                cs.indices[o + k] = m;
                bkeys[k]          = ((uint64_t)code_key(v) << 32) | (uint32_t)k;
                if (v != v)
                    atomicOr(flags, CODE_NAN);
            }
            __syncthreads();
            if (threadIdx.x == CODE_NT - 1)
                sslot = at + incl;
            __syncthreads();
        }
        for (int k = n + threadIdx.x; k < n2; k += CODE_NT)
            bkeys[k] = ~0ull;
        __syncthreads();
        code_bitonic(bkeys, n2, threadIdx.x, CODE_NT, [] { __syncthreads(); });
        const uint32_t d = code_ranks(bkeys, n, threadIdx.x, CODE_NT, cs.data + o,
                                      [&](int v) { return code_block_scan(v, red); });
        if (threadIdx.x == CODE_NT - 1)
        {
            atomicMax(&max_uniq[s], d + 1);
            if (d > 65535u)
                atomicOr(flags, CODE_OVERFLOW);
        }
    }
}

// Clears the coding passes' counters: misc = [flags, n_over, max_uniq[2]
// (u32), bases[2] (i64)], 32 bytes.
__global__ void k_code_init(uint32_t* __restrict__ misc)
{
    if (threadIdx.x < 8)
        misc[threadIdx.x] = 0u;
}

__global__ void k_code_zero(uint32_t* __restrict__ v)
{
    if (threadIdx.x == 0)
        *v = 0u;
}

// Copies the flags and both max_uniq values out of misc (status[3] = 0).
__global__ void k_code_status(const uint32_t* __restrict__ misc, uint32_t* __restrict__ status)
{
    if (threadIdx.x == 0)
    {
        status[0] = misc[0];
        status[1] = misc[2];
        status[2] = misc[3];
        status[3] = 0u;
    }
}

}  // namespace csg
