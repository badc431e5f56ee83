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
__global__ void k_prop_scatter_win(const int32_t* __restrict__ top,
                                   const float* __restrict__ margin, float thr,
                                   const uint32_t* __restrict__ list, uint32_t n, int KP, int t0,
                                   int t1, uint32_t L, const uint32_t* __restrict__ slot,
                                   const uint32_t* __restrict__ a_base,
                                   const uint64_t* __restrict__ c_off,
                                   const uint32_t* __restrict__ llen, uint32_t* __restrict__ cells,
                                   RecheckItem* __restrict__ items, Segment* __restrict__ segs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
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
// base[l] + f in label l's list (base NULL: 0).
__global__ void k_prop_scatter_all(const uint32_t* __restrict__ list, uint32_t n, uint32_t L,
                                   const uint32_t* __restrict__ base,
                                   const uint32_t* __restrict__ a_base,
                                   const uint64_t* __restrict__ c_off,
                                   const uint32_t* __restrict__ llen, uint32_t* __restrict__ cells,
                                   RecheckItem* __restrict__ items, Segment* __restrict__ segs)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n)
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

__global__ void k_prop_scatter_p(const RecheckItem* __restrict__ items, uint32_t n,
                                 const float* __restrict__ val, float* __restrict__ P, uint32_t L)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        P[(size_t)items[i].q * L + items[i].l] = val[i];
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
    const float* __restrict__ P, uint32_t n, uint32_t L, const int32_t* __restrict__ top,
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
    if (i < n)
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
