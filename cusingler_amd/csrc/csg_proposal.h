// Work lists of the proposal-restricted score step.
//
// A proposal ranks, per query cell, the KP labels a network scores highest
// and gives a softmax margin.  Cells whose margin reaches the threshold are
// scored exactly against the reference rows of their first K labels; the
// others (fallback cells) against every label.  Work is grouped by label:
// label l's list holds the cells scoring it in this stage (in atomic slot
// order) followed by the cells scored on every label, and one grouped
// contraction scores each list against the label's reference rows.  Each (cell, label) item then owns a
// contiguous run of llen[l] scores (in reference row order, as in the full
// score matrix), so the quantile select and the label logic see exactly the
// values the full path produces for that pair.
#pragma once

#include "csg_exact.h"
#include "csg_kernels.h"

namespace csg {

__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// A cell is confident when its margin reaches `thr` and its KP ranked
// labels are valid.
__device__ inline bool prop_confident(const int32_t* top, const float* margin, uint32_t c, int KP,
                                      uint32_t L, float thr)
{
    bool ok = margin[c] >= thr;  // NaN margins fall back
#pragma unroll 8
    for (int t = 0; t < KP; ++t)
    {
        const int32_t l = top[(size_t)c * KP + t];
        ok &= l >= 0 && (uint32_t)l < L;
    }
    return ok;
}

// First stage: every confident cell takes a slot in the lists of its first
// K ranked labels; the other cells are appended to the fallback list.
// cnt[0, L) counts per label, cnt[L] the fallback cells (L <= 256).  Slots
// are taken in LDS first and each block adds its per-label totals to the
// global counters once.
__global__ __launch_bounds__(256) void k_prop_count(const int32_t* __restrict__ top,
                                                    const float* __restrict__ margin, uint32_t nq,
                                                    int KP, int K, uint32_t L, float thr,
                                                    uint32_t* __restrict__ cnt,
                                                    uint32_t* __restrict__ slot,
                                                    uint32_t* __restrict__ fb)
{
    __shared__ uint32_t hist[257], base[257];
    for (uint32_t i = threadIdx.x; i <= L; i += blockDim.x)
        hist[i] = 0;
    __syncthreads();
    const uint32_t c    = blockIdx.x * blockDim.x + threadIdx.x;
    const bool     live = c < nq;
    const bool     conf = live && prop_confident(top, margin, c, KP, L, thr);
    uint32_t       ls[8];
    int            lab[8];
    if (conf)
        for (int t = 0; t < K; ++t)
        {
            lab[t] = top[(size_t)c * KP + t];
            ls[t]  = atomicAdd(&hist[lab[t]], 1u);
        }
    else if (live)
        ls[0] = atomicAdd(&hist[L], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i <= L; i += blockDim.x)
        if (hist[i])
            base[i] = atomicAdd(&cnt[i], hist[i]);
    __syncthreads();
    if (conf)
        for (int t = 0; t < K; ++t)
            slot[(size_t)c * KP + t] = base[lab[t]] + ls[t];
    else if (live)
        fb[base[L] + ls[0]] = c;
}

// Work tables of the cells scored on a window [t0, t1) of their ranked
// labels: `list` names them (NULL: every confident cell of the chunk, with
// n = nq); each takes the slot its counting pass assigned.
// (n_dev, when given: the list's length on the device, at most n.)
__global__ void k_prop_scatter_win(const int32_t* __restrict__ top,
                                   const float* __restrict__ margin, float thr,
                                   const uint32_t* __restrict__ list, uint32_t n,
                                   const uint32_t* __restrict__ n_dev, int KP, int t0, int t1,
                                   uint32_t L, const uint32_t* __restrict__ slot,
                                   const uint32_t* __restrict__ a_base,
                                   const uint64_t* __restrict__ c_off,
                                   const uint32_t* __restrict__ llen, uint32_t* __restrict__ cells,
                                   RecheckItem* __restrict__ items, Segment* __restrict__ segs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n_dev ? min(n, *n_dev) : n))
        return;
    const uint32_t c = list ? list[i] : i;
    if (!list && !prop_confident(top, margin, c, KP, L, thr))
        return;
    for (int t = t0; t < t1; ++t)
    {
        const uint32_t l   = (uint32_t)top[(size_t)c * KP + t];
        const uint32_t s   = slot[(size_t)c * KP + t];
        const uint32_t pos = a_base[l] + s;
        cells[pos]         = c;
        items[pos]         = RecheckItem{ c, l };
        segs[pos]          = Segment{ c_off[l] + (uint64_t)s * llen[l], llen[l], l };
    }
}

// Work tables of the cells scored on every label: cell list[f] takes slot
// base[l] + f in label l's list (base NULL: 0); the list holds *n_dev cells
// (at most n).
__global__ void k_prop_scatter_all(const uint32_t* __restrict__ list, uint32_t n,
                                   const uint32_t* __restrict__ n_dev, uint32_t L,
                                   const uint32_t* __restrict__ base,
                                   const uint32_t* __restrict__ a_base,
                                   const uint64_t* __restrict__ c_off,
                                   const uint32_t* __restrict__ llen, uint32_t* __restrict__ cells,
                                   RecheckItem* __restrict__ items, Segment* __restrict__ segs)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= min(n, *n_dev))
        return;
    const uint32_t c = list[f];
    for (uint32_t l = 0; l < L; ++l)
    {
        const uint32_t s   = (base ? base[l] : 0u) + f;
        const uint32_t pos = a_base[l] + s;
        cells[pos]         = c;
        items[pos]         = RecheckItem{ c, l };
        segs[pos]          = Segment{ c_off[l] + (uint64_t)s * llen[l], llen[l], l };
    }
}

// The round's *n_dev items (planned on the device), grid-stride.
__global__ void k_prop_scatter_p(const RecheckItem* __restrict__ items,
                                 const uint32_t* __restrict__ n_dev,
                                 const float* __restrict__ val, float* __restrict__ P, uint32_t L)
{
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        P[(size_t)items[i].q * L + items[i].l] = val[i];
}

// The proposal's tallies of a chunk, added to the device counters: [16]
// cells decided on proposed labels, [17] cells below the margin, [18] cells
// the edge tests sent to scoring on every label, [20] cells widened.
__global__ void k_prop_tally(uint32_t nq, const uint32_t* __restrict__ nfb,
                             const uint32_t* __restrict__ n1, const uint32_t* __restrict__ n_full,
                             int widen, unsigned long long* __restrict__ dcnt)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        const uint32_t f = *nfb, full = *n_full;
        dcnt[16] += nq - f - full;
        dcnt[17] += f;
        dcnt[18] += full;
        if (widen)
            dcnt[20] += *n1;
    }
}

// Plan of one grouped scoring round, on the device: label l's list holds
// na_l = (cw ? cw[l] : 0) + (n_all ? *n_all : 0) cells; a_base and c_off are
// the exclusive scans of na_l and of na_l * llen[l] (the label's first item
// and first score), a_cnt = na_l; the tiles (label, m0, n0) of BM x BN cover
// every label's na_l x llen[l] block, label by label, m0 outer.  meta[0]
// receives the tile count, meta[1] the item count; dcnt[19] and dcnt[21]
// accumulate the items and scores.  Blocks of 256 threads, L <= 256: every
// block scans the labels in LDS, block 0 writes the label tables and counts,
// the blocks share the tile list.
__global__ __launch_bounds__(256) void k_prop_plan(
    const uint32_t* __restrict__ cw, const uint32_t* __restrict__ n_all,
    const uint32_t* __restrict__ llen, uint32_t L, uint32_t BM, uint32_t BN,
    uint32_t* __restrict__ a_base, uint32_t* __restrict__ a_cnt, uint64_t* __restrict__ c_off,
    GroupTile* __restrict__ tiles, uint32_t* __restrict__ meta,
    unsigned long long* __restrict__ dcnt)
{
    __shared__ uint32_t s_na[256], s_tb[257], s_nb[256];
    __shared__ uint64_t s_sc[256];
    const uint32_t      t  = threadIdx.x;
    const uint32_t      fb = n_all ? *n_all : 0u;
    if (t < L)
    {
        const uint32_t na = (cw ? cw[t] : 0u) + fb, nl = llen[t];
        s_na[t] = na;
        s_nb[t] = (nl + BN - 1) / BN;
        s_tb[t] = ((na + BM - 1) / BM) * s_nb[t];
        s_sc[t] = (uint64_t)na * nl;
    }
    __syncthreads();
    if (t == 0)
    {
        const bool first = blockIdx.x == 0;
        uint32_t   ia = 0, it = 0;
        uint64_t   isc = 0;
        for (uint32_t l = 0; l < L; ++l)
        {
            if (first)
            {
                a_base[l] = ia;
                a_cnt[l]  = s_na[l];
                c_off[l]  = isc;
            }
            const uint32_t tl = s_tb[l];
            s_tb[l]           = it;
            ia += s_na[l];
            isc += s_sc[l];
            it += tl;
        }
        s_tb[L] = it;
        if (first)
        {
            meta[0] = it;
            meta[1] = ia;
            atomicAdd(&dcnt[19], (unsigned long long)ia);
            atomicAdd(&dcnt[21], (unsigned long long)isc);
        }
    }
    __syncthreads();
    const uint32_t nt = s_tb[L];
    for (uint32_t k = blockIdx.x * blockDim.x + t; k < nt; k += gridDim.x * blockDim.x)
    {
        // label of tile k: the last l with s_tb[l] <= k (binary search)
        uint32_t lo = 0, hi = L - 1;
        while (lo < hi)
        {
            const uint32_t mid = (lo + hi + 1) / 2;
            if (s_tb[mid] <= k)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint32_t r = k - s_tb[lo], nb = s_nb[lo];
        tiles[k]         = GroupTile{ lo, (r / nb) * BM, (r % nb) * BN, 0u };
    }
}

// Edge test on the exactly scored labels of a cell (P holds -inf for the
// labels not scored yet).  Labels outside the scored part of the ranking
// matter only if one would reach the candidate mask (score >= max - thre) or
// the maximum itself; the network ranks them below its label at position
// `edge`, so the cell is listed for more scoring when that label may itself
// lie within `thre` of the exact maximum (wp: the exact path's half-width)
// -- in particular whenever the exact winner is that weakest scored label,
// i.e. when the winner could lie outside the scored part.  `in` names the
// cells to test (NULL: every confident cell, n = nq).  With `widen`, a
// listed cell also takes slots (counted in cnt) in the lists of its ranked
// labels [edge + 1, KP).  Cells with a NaN score are left to the
// certification, which re-scores all of their labels.
// One thread per cell.  A listed cell's list position and label slots are
// taken in LDS first and each block adds its totals to the global counters
// once (per-thread atomics on the few counters, which share a cache line,
// had serialised in one L2 channel).
__global__ __launch_bounds__(256) void k_prop_check(
    const float* __restrict__ P, uint32_t n, const uint32_t* __restrict__ n_dev, uint32_t L,
    const int32_t* __restrict__ top,
    const float* __restrict__ margin, int KP, int edge, float thr, float thre, double wp,
    const uint32_t* __restrict__ in, int widen, uint32_t* __restrict__ cnt,
    uint32_t* __restrict__ slot, uint32_t* __restrict__ list, uint32_t* __restrict__ n_list)
{
    __shared__ uint32_t hist[257], base[257];
    for (uint32_t j = threadIdx.x; j <= L; j += blockDim.x)
        hist[j] = 0;
    __syncthreads();
    const uint32_t i    = blockIdx.x * blockDim.x + threadIdx.x;
    bool           pass = false;
    uint32_t       c    = 0;
    if (i < (n_dev ? min(n, *n_dev) : n))
    {
        c = in ? in[i] : i;
        if (in || prop_confident(top, margin, c, KP, L, thr))
        {
            const float* s   = P + (size_t)c * L;
            float        mx  = -__builtin_inff();
            bool         nan = false;
            // (no early exit at a NaN, so that the row's loads are in flight together)
#pragma unroll 8
            for (uint32_t j = 0; j < L; ++j)
            {
                const float v = s[j];
                nan |= __builtin_isnan(v);
                mx = fmaxf(mx, v);
            }
            if (!nan)
            {
                const float last = s[top[(size_t)c * KP + edge]];
                pass = !((double)last + 2.0 * wp < (double)mx - (double)thre);
            }
        }
    }
    uint32_t lpos = 0, ls[8];
    int      lab[8];
    if (pass)
    {
        lpos = atomicAdd(&hist[L], 1u);
        if (widen)
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t > edge && t < KP)
                {
                    lab[t] = top[(size_t)c * KP + t];
                    ls[t]  = atomicAdd(&hist[lab[t]], 1u);
                }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j <= L; j += blockDim.x)
        if (hist[j])
            base[j] = atomicAdd(j < L ? &cnt[j] : n_list, hist[j]);
    __syncthreads();
    if (pass)
    {
        list[base[L] + lpos] = c;
        if (widen)
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t > edge && t < KP)
                    slot[(size_t)c * KP + t] = base[lab[t]] + ls[t];
    }
}

}  // namespace csg
