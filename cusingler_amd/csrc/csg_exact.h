// Certified labels for the exact score path.
//
// The exact path knows every rank dot product exactly (S4) and the exact sums
// of squares Sxx, Syy of both rows, and returns (S4 / sqrt(Sxx)) / sqrt(Syy)
// as S4 times the two correctly rounded inverse roots in double (relative
// error below 2^-50), rounded once to float: the correlation to within half a
// float ulp plus 1e-15.  The
// reference computes the same quantity in FP32: its sequential sum of
// products differs from S4 by at most gamma_{G-1} * sum |p_k| <= gamma_{G-1} *
// sqrt(Sxx Syy) (gamma_n = n u / (1 - n u), u = 2^-24), and its divisor
// sqrtf(sxx_f * syy_f) differs from sqrt(Sxx Syy) by a relative e_D <=
// gamma_{G-1} + 2u (1 + gamma_{G-1}) (two sequential sums of exact squares,
// each within a factor 1 +- gamma, one product, one root), so each of its scores lies within W = (gamma + e_D) / (1 - e_D) + 2u
// of the exact path's (host-computed).  Order statistics move by at most W
// and the quantile formula is a convex combination of two of them, so every
// per-label quantile of the reference lies within W_P = W + 16u of the exact
// path's.  A cell's first label and mask are decided from those intervals
// when they are unambiguous; the (cell, label) pairs that are not are scored
// again in the reference's exact FP32 order and the decision is redone with
// their exact values.
#pragma once

#include "csg_kernels.h"

namespace csg {

struct RecheckItem
{
    uint32_t q;  // query row within the chunk
    uint32_t l;  // label
};

// Neighbouring floats of a finite x.
__device__ inline float next_up(float x)
{
    if (x == 0.f)
        return __uint_as_float(1u);
    const unsigned int b = __float_as_uint(x);
    return __uint_as_float(x > 0.f ? b + 1u : b - 1u);
}
__device__ inline float next_down(float x) { return -next_up(-x); }

// Per cell: decide first label and mask from the exact-path quantiles P
// (half-width wp), or list the labels that need the reference's FP32 value.
// status[q]: 0 decided, 1 partial recheck, 2 full recheck (a NaN quantile:
// the libstdc++ minmax semantics then need every exact value).
// One wave per cell, the lanes over its labels (four per lane: L <= 256),
// so that a row's quantiles are read and its mask written coalesced.  The
// bounds of the reference's maximum, max_j (v_j -+ wp) in double, are
// (max_j v_j) -+ wp rounded once (rounding is monotone).
__global__ __launch_bounds__(256) void k_label_certify(
    const float* __restrict__ P, uint32_t nq, uint32_t L, double wp, float thre,
    int32_t* __restrict__ first, float* __restrict__ mask, uint8_t* __restrict__ status,
    RecheckItem* __restrict__ items, unsigned int* __restrict__ n_items)
{
    // n_items[0]: (cell, label) pairs listed; n_items[1]: cells left uncertified
    const uint32_t q    = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int      lane = threadIdx.x & 63;
    if (q >= nq)
        return;  // (the whole wave)
    const float* s  = P + (size_t)q * L;
    float*       mk = mask + (size_t)q * L;
    float        v[4];
    bool         nan = false;
    float        vmax = -__builtin_inff();
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        const uint32_t j = lane + 64 * u;
        v[u]             = j < L ? s[j] : 0.f;
        if (j < L)
        {
            nan |= __builtin_isnan(v[u]);
            vmax = fmaxf(vmax, v[u]);
        }
    }
    // lane 0 reserves `count` list slots; every lane gets the base
    auto reserve = [&](unsigned int count) {
        unsigned int base = 0;
        if (lane == 0)
            base = atomicAdd(n_items, count);
        return (unsigned int)__builtin_amdgcn_readfirstlane((int)base);
    };
    if (__ballot(nan))
    {
        if (lane == 0)
        {
            status[q] = 2;
            atomicAdd(n_items + 1, 1u);
        }
        const unsigned int base = reserve(L);
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const uint32_t j = lane + 64 * u;
            if (j < L)
                items[base + j] = RecheckItem{ q, j };
        }
        return;
    }
    vmax              = wave_max(vmax);
    const double mlo  = (double)vmax - wp;
    const double mhi  = (double)vmax + wp;
    // float bounds of the reference's maximum and threshold
    float flo = (float)mlo;
    if ((double)flo > mlo)
        flo = next_down(flo);
    float fhi = (float)mhi;
    if ((double)fhi < mhi)
        fhi = next_up(fhi);
    const double tlo = (double)(flo - thre);
    const double thi = (double)(fhi - thre);
    uint32_t ncand = 0;
    int      cand  = -1;
    bool     unsure = false;
    bool     list[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        const uint32_t j   = lane + 64 * u;
        const double   dv  = (double)v[u];
        const bool     ok  = j < L;
        const bool     inC = ok && dv + wp >= mlo;
        const bool     hi1 = dv - wp >= thi, lo0 = dv + wp < tlo;
        if (ok && hi1)
            mk[j] = 1.f;
        else if (ok && lo0)
            mk[j] = 0.f;
        const bool unk = ok && !hi1 && !lo0;
        unsure |= unk;
        list[u] = inC || unk;
        const unsigned long long m = __ballot(inC);
        ncand += (uint32_t)__builtin_popcountll(m);
        if (m)
            cand = 64 * u + 63 - __builtin_clzll(m);  // the last candidate label
    }
    if (ncand == 1 && !__ballot(unsure))
    {
        if (lane == 0)
        {
            first[q]  = cand;
            status[q] = 0;
        }
        return;
    }
    if (lane == 0)
    {
        status[q] = 1;
        atomicAdd(n_items + 1, 1u);
    }
    unsigned long long m[4];
    unsigned int       total = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        m[u] = __ballot(list[u]);
        total += (unsigned int)__builtin_popcountll(m[u]);
    }
    unsigned int at = reserve(total);
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        if (list[u])
            items[at + __builtin_amdgcn_mbcnt_hi((uint32_t)(m[u] >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m[u], 0u))] =
                RecheckItem{ q, lane + 64u * u };
        at += (unsigned int)__builtin_popcountll(m[u]);
    }
}

// ---- re-scoring lists grouped by label, built on the device

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Pairs per label.
// (Grid-stride over the *n_dev items the certification listed.)
__global__ void k_rc_count(const RecheckItem* __restrict__ items,
                           const uint32_t* __restrict__ n_dev, uint32_t* __restrict__ cnt)
{
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&cnt[items[i].l], 1u);
}

// One thread: label l's cells start at abase[l] of the grouped list, its
// scores at coff[l] (acnt[l] cells x llen[l] rows), its tiles at tbase[l]
// (TM-cell x TN-row tiles); the fill counters are cleared; ntiles[0] = tiles.
__global__ void k_rc_scan(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ llen,
                          uint32_t L, uint32_t TM, uint32_t TN, uint32_t* __restrict__ abase,
                          uint64_t* __restrict__ coff, uint32_t* __restrict__ tbase,
                          uint32_t* __restrict__ fill, uint32_t* __restrict__ ntiles)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    uint32_t a = 0, t = 0;
    uint64_t c = 0;
    for (uint32_t l = 0; l < L; ++l)
    {
        abase[l] = a;
        coff[l]  = c;
        tbase[l] = t;
        fill[l]  = 0;
        a += cnt[l];
        c += (uint64_t)cnt[l] * llen[l];
        t += ((cnt[l] + TM - 1) / TM) * ((llen[l] + TN - 1) / TN);
    }
    ntiles[0] = t;
}

// Each pair takes the next place of its label: the label's cell list, the
// pair in label order and its score segment (contiguous within the label's
// block).  Places within a label follow atomics; every pair's values do not
// depend on its place.
__global__ void k_rc_scatter(const RecheckItem* __restrict__ items,
                             const uint32_t* __restrict__ n_dev,
                             const uint32_t* __restrict__ abase, const uint64_t* __restrict__ coff,
                             const uint32_t* __restrict__ llen, uint32_t* __restrict__ fill,
                             uint32_t* __restrict__ cells, RecheckItem* __restrict__ sorted,
                             Segment* __restrict__ segs)
{
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const RecheckItem it  = items[i];
        const uint32_t    len = llen[it.l];
        const uint32_t    p   = atomicAdd(&fill[it.l], 1u);
        const uint32_t    pos = abase[it.l] + p;
        cells[pos]            = it.q;
        sorted[pos]           = it;
        segs[pos]             = Segment{ coff[it.l] + (uint64_t)p * len, len, it.l };
    }
}

// Tile list of the grouped contraction, one thread per label.
__global__ void k_rc_tiles(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ llen,
                           const uint32_t* __restrict__ tbase, uint32_t L, uint32_t TM,
                           uint32_t TN, GroupTile* __restrict__ tiles)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L)
        return;
    uint32_t t = tbase[l];
    for (uint32_t m0 = 0; m0 < cnt[l]; m0 += TM)
        for (uint32_t n0 = 0; n0 < llen[l]; n0 += TN)
            tiles[t++] = GroupTile{ l, m0, n0, 0u };
}

// The uncertified cells (status 1 or 2), listed once each.
__global__ void k_rc_cells(const uint8_t* __restrict__ status, uint32_t nq,
                           uint32_t* __restrict__ list, uint32_t* __restrict__ n_list)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq && status[q] != 0)
        list[atomicAdd(n_list, 1u)] = q;
}

// Writes the re-scored quantiles into P and redoes the decision of every cell
// that needed them (status 1 or 2).
__global__ void k_apply_recheck(const RecheckItem* __restrict__ items,
                                const uint32_t* __restrict__ n_dev,
                                const float* __restrict__ Pref, float* __restrict__ P,
                                uint32_t L, uint8_t* __restrict__ exact)
{
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const RecheckItem it           = items[i];
        P[(size_t)it.q * L + it.l]     = Pref[i];
        exact[(size_t)it.q * L + it.l] = 1;
    }
}

// out[0] = *a, out[1] = *b (two device counts gathered for one read-back).
__global__ void k_gather_u32(const unsigned int* __restrict__ a, const unsigned int* __restrict__ b,
                             uint32_t* __restrict__ out)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        out[0] = *a;
        out[1] = *b;
    }
}

// The certification's tallies of a chunk, added to the device counters:
// [3] re-scored pairs, [4] uncertified cells, [5] NaN-carrying segments.
__global__ void k_cert_tally(const unsigned int* __restrict__ nitems,
                             const unsigned int* __restrict__ nan_count,
                             unsigned long long* __restrict__ dcnt)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        dcnt[3] += nitems[0];
        dcnt[4] += nitems[1];
        dcnt[5] += *nan_count;
    }
}

__global__ void k_label_final(const float* __restrict__ P, const uint8_t* __restrict__ status,
                              const uint8_t* __restrict__ exact, uint32_t nq, uint32_t L,
                              float thre, int32_t* __restrict__ first, float* __restrict__ mask)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq || status[q] == 0)
        return;
    const float*   s  = P + (size_t)q * L;
    const uint8_t* ex = exact + (size_t)q * L;
    float*         mk = mask + (size_t)q * L;
    if (status[q] == 2)
    {
        // every value is the reference's: its own minmax / threshold logic
        int imin, imax;
        std_minmax(s, (int)L, &imin, &imax);
        first[q]       = imax;
        const float th = s[imax] - thre;
        for (uint32_t j = 0; j < L; ++j)
            mk[j] = s[j] >= th ? 1.f : 0.f;
        return;
    }
    // the maximum lies among the re-scored labels; the last one holding it
    int   jm = -1;
    float m  = 0.f;
    for (uint32_t j = 0; j < L; ++j)
    {
        if (!ex[j])
            continue;
        if (jm < 0 || !(s[j] < m))
        {
            m  = s[j];
            jm = (int)j;
        }
    }
    first[q]       = jm;
    const float th = m - thre;
    for (uint32_t j = 0; j < L; ++j)
        if (ex[j])
            mk[j] = s[j] >= th ? 1.f : 0.f;
}

}  // namespace csg
