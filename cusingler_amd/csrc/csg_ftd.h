// Device-resident fine-tune rounds (cufinetune / finetune_round,
// src/cusingler.cu:1138-1333).
//
// Every round runs on the device with its work counts kept in device memory,
// so that a round is a fixed sequence of launches (captured in two HIP
// graphs by the host) and the host reads back one small counter block per
// round:
//   1. group the active cells by candidate set: a 64-bit hash of the set's
//      label bitset is inserted into an open-addressing table (the first
//      inserter owns the slot), slots receive group ids, and every cell
//      checks its full set against its group owner's (a mismatch -- two sets
//      under one hash -- is reported and the host falls back);
//   2. per group: the ascending label list, G_sub = the union of every
//      ordered pair's first min(len, k_t) window genes (k_t =
//      round(500 (2/3)^log2 k)), built as a bitset in LDS and scanned into an
//      ascending gene list;
//   3. one block prefix-sums the groups' work sizes into offsets;
//   4. work-queue ("persistent") kernels -- grids sized to the machine, each
//      block pulling items until the device-side count is exhausted -- rank
//      the reference and query rows over G_sub, correlate (exact INT8 MFMA
//      tiles or FP32 chunk-tree tiles), select the per-label quantiles and
//      filter every cell's candidate list in place; cells the certified
//      filter cannot decide are redone in FP32 order;
//   5. the cells that still have several candidates form the next round.
// Work placement depends on atomics, values do not: every (cell, reference
// row) sum and every per-segment quantile is computed exactly as in the
// host-grouped path, so labels and round counts are identical.
#pragma once

#include "csg_ft_kernels.h"
#include "csg_stdalgo.h"

namespace csg {

struct FtdCounts
{
    uint32_t n_active;   // cells active in this round
    uint32_t n_next;     // cells active in the next round
    uint32_t n_groups;
    uint32_t n_cells;    // cells with work this round (groups of >= 20 genes)
    uint32_t n_tiles;    // wide correlation tiles (exact 16 x 256 or FP32 32 x 32)
    uint32_t n_thin;     // FP32 thin tiles
    uint32_t n_redo;     // cells the certified filter left undecided
    uint32_t n_rthin;    // thin tiles of their FP32 redo
    uint32_t n_nan;      // NaN segments of the select
    uint32_t n_nan2;     // NaN segments of the redo select
    uint32_t collision;  // two candidate sets met under one hash
    uint32_t undefined;  // the libstdc++ sort emulation would leave its array
    uint32_t max_stride;
    uint32_t round;
    uint32_t tot_redo;  // redone cells over all rounds
    uint32_t g_lo, g_hi;  // group range of the current work batch
    uint32_t n_fused;     // groups of the fused FP32 form
    uint32_t n_etiles;    // exact 16-cell tiles (of n_tiles)
    uint32_t n_btiles;    // exact FTD_BIG-cell tiles (of n_tiles)
    uint32_t hT;          // hash slots in use this round (a power of two, >= 2 n_active)
    uint64_t n_ref_items;
    uint64_t ref_elems;
    uint64_t cell_elems;
    uint64_t n_scores;
    uint64_t n_segs;
    uint64_t n_redo_segs;
    uint64_t rthin_bound;  // tiles a redo of every exact-form cell would take
    uint64_t pair_genes;   // sum over rounds of cells x reference rows x |G_sub| (work measure)
    // Chained rounds (several rounds launched without a host read-back): a
    // round whose work exceeds the tables the chain was launched with, or
    // whose grouping met a hash collision, sets `halt` in its scan and leaves
    // the state at its start (its compaction done); every later kernel of the
    // chain then returns at once.  Sticky until the host clears it.
    uint32_t halt;
    uint32_t max_count;   // largest candidate list of the round's active cells
    uint64_t tot_active;  // active cells summed over rounds
    uint64_t tot_groups;
    uint64_t tot_scores;
    uint64_t tot_fused;
    uint32_t tot_rounds;  // rounds with active cells (all rounds of the call)
    uint32_t pad_;
};

// Capacities of the work tables a chain of rounds was launched with (in the
// units of the matching FtdCounts totals); a round needing more halts.
struct FtdCaps
{
    uint64_t ref_items, cells, ref_elems, cell_elems, scores, segs, tiles, thin, rthin;
};

// Launch mask of a round: the work-queue kernel families whose work lists
// are not empty.  A round's part B launches the families of its mask only.
constexpr uint32_t FTDM_RANK = 1, FTDM_MFMA = 2, FTDM_MFMA64 = 4, FTDM_CORR = 8, FTDM_THIN = 16,
                   FTDM_FUSED = 32, FTDM_REDO = 64, FTDM_ALL = 127;

__host__ __device__ inline uint32_t ftd_mask_need(uint64_t ref_elems, uint32_t n_tiles,
                                                  uint32_t n_etiles, uint32_t n_btiles,
                                                  uint32_t n_thin, uint32_t n_fused)
{
    const uint32_t ftiles = n_tiles - n_etiles - n_btiles;
    return (ref_elems ? FTDM_RANK : 0u) | (n_etiles ? FTDM_MFMA : 0u) |
           (n_btiles ? FTDM_MFMA64 : 0u) | (ftiles ? FTDM_CORR : 0u) | (n_thin ? FTDM_THIN : 0u) |
           (n_fused ? FTDM_FUSED : 0u) | (n_etiles || n_btiles ? FTDM_REDO : 0u);
}

// Height of the big exact fine-tune tile (cells).
constexpr int FTD_BIG = 32;

// Per-group flags (bit 0: fewer than 20 genes, bit 1: exact form, bit 2:
// fused FP32 form -- a group of at most FTD_FUSE_MAX cells whose reference
// rows are ranked and correlated in one pass without storing their ranks).
constexpr uint8_t FTD_SHORT = 1, FTD_EXACT = 2, FTD_FUSED = 4;
constexpr int     FTD_FUSE_MAX = 4;

__device__ inline void ftd_key(const uint8_t* __restrict__ list, uint32_t k, uint64_t key[4])
{
    key[0] = key[1] = key[2] = key[3] = 0ull;
    for (uint32_t t = 0; t < k; ++t)
    {
        const uint32_t l = list[t];
        key[l >> 6] |= 1ull << (l & 63);
    }
}

__device__ inline uint64_t ftd_mix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ inline uint64_t ftd_hash(const uint64_t key[4])
{
    uint64_t h = ftd_mix(key[0]);
    h          = ftd_mix(h ^ key[1]);
    h          = ftd_mix(h ^ key[2]);
    h          = ftd_mix(h ^ key[3]);
    return h ? h : 1ull;  // 0 marks an empty slot
}

// Initial candidate lists: ascending labels of each cell's mask row
// (src/cusingler.cu:1307-1312); cells with several candidates (and rounds
// left) are listed as the first round's active cells.
__global__ void k_ftd_init(const float* __restrict__ mask, uint32_t Nq, uint32_t L, int times,
                           uint8_t* __restrict__ lists, uint32_t* __restrict__ counts,
                           int32_t* __restrict__ rounds, uint32_t* __restrict__ act_next,
                           FtdCounts* __restrict__ cnt)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= Nq)
        return;
    const float* m = mask + (size_t)c * L;
    uint8_t*     o = lists + (size_t)c * L;
    uint32_t     k = 0;
    // eight mask values loaded together, then listed (the list stores would
    // otherwise wait behind each load)
    for (uint32_t l0 = 0; l0 < L; l0 += 8)
    {
        float mv[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u)
            mv[u] = m[min(l0 + u, L - 1)];  // (past the row: its last value, not used)
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u)
            if (l0 + u < L && mv[u] != 0.f)
                o[k++] = (uint8_t)(l0 + u);
    }
    counts[c] = k;
    rounds[c] = 0;
    if (k > 1 && (times == 0 || 0 < times))
        act_next[atomicAdd(&cnt->n_next, 1u)] = c;
}

// Round start, part 2 (after ftd_counters_body): the previous round's next
// list becomes this round's active list; the group member counters and the
// hash slots this round uses are cleared.  (i0, stride: this thread's first
// index and the threads' count.)
__device__ inline void ftd_begin_body(const FtdCounts* __restrict__ cnt,
                                      uint32_t* __restrict__ act,
                                      const uint32_t* __restrict__ act_next,
                                      unsigned long long* __restrict__ htab,
                                      uint32_t* __restrict__ gcount, uint32_t i0, uint32_t stride)
{
    const uint32_t n = cnt->n_active;
    const uint32_t T = cnt->hT;
    for (uint32_t i = i0; i < n; i += stride)
    {
        act[i]    = act_next[i];
        gcount[i] = 0;
    }
    for (uint32_t s = i0; s < T; s += stride)
        htab[s] = 0ull;
}

__global__ void k_ftd_begin(const FtdCounts* __restrict__ cnt, uint32_t* __restrict__ act,
                            const uint32_t* __restrict__ act_next,
                            unsigned long long* __restrict__ htab,
                            uint32_t* __restrict__ gcount)
{
    if (cnt->halt)
        return;
    ftd_begin_body(cnt, act, act_next, htab, gcount, blockIdx.x * blockDim.x + threadIdx.x,
                   gridDim.x * blockDim.x);
}

// Round start, part 1 (one thread): counters, and the hash slots of the
// round: the smallest power of two of at least 2 n_active (and 1024), at most
// T, so that an insert always finds a free slot.
__device__ inline void ftd_counters_body(FtdCounts* __restrict__ cnt, uint32_t T)
{
    const uint32_t na = cnt->n_next;
    uint32_t       hT = 1024;
    while (hT < 2 * na && hT < T)
        hT <<= 1;
    cnt->hT          = hT;
    cnt->n_active    = na;
    cnt->n_next      = 0;
    cnt->n_groups    = 0;
    cnt->n_cells     = 0;
    cnt->n_tiles     = 0;
    cnt->n_thin      = 0;
    cnt->n_redo      = 0;
    cnt->n_rthin     = 0;
    cnt->n_nan       = 0;
    cnt->n_nan2      = 0;
    cnt->collision   = 0;
    cnt->max_count   = 0;
    cnt->max_stride  = 0;
    cnt->round      += 1;
    cnt->g_lo        = 0u;
    cnt->g_hi        = 0xFFFFFFFFu;
    cnt->n_fused     = 0;
    cnt->n_ref_items = 0;
    cnt->ref_elems   = 0;
    cnt->cell_elems  = 0;
    cnt->n_scores    = 0;
    cnt->n_segs      = 0;
    cnt->n_redo_segs = 0;
    cnt->rthin_bound = 0;
}

__global__ void k_ftd_counters(FtdCounts* __restrict__ cnt, uint32_t T)
{
    if (blockIdx.x != 0 || threadIdx.x != 0 || cnt->halt)
        return;
    ftd_counters_body(cnt, T);
}

__device__ inline void ftd_hash_body(const FtdCounts* __restrict__ cnt,
                                     const uint32_t* __restrict__ act,
                                     const uint8_t* __restrict__ lists,
                                     const uint32_t* __restrict__ counts, uint32_t L,
                                     unsigned long long* __restrict__ htab,
                                     uint32_t* __restrict__ howner, uint32_t* __restrict__ cslot,
                                     uint32_t i0, uint32_t stride)
{
    const uint32_t n = cnt->n_active;
    const uint32_t T = cnt->hT;
    for (uint32_t i = i0; i < n; i += stride)
    {
        const uint32_t c = act[i];
        uint64_t       key[4];
        ftd_key(lists + (size_t)c * L, counts[c], key);
        const unsigned long long h = ftd_hash(key);
        uint32_t                 s = (uint32_t)h & (T - 1);
        // T >= 2 n: a free slot is always reached
        for (;;)
        {
            const unsigned long long prev = atomicCAS(&htab[s], 0ull, h);
            if (prev == 0ull)
            {
                howner[s] = c;
                break;
            }
            if (prev == h)
                break;
            s = (s + 1) & (T - 1);
        }
        cslot[i] = s;
    }
}

__global__ void k_ftd_hash(const FtdCounts* __restrict__ cnt, const uint32_t* __restrict__ act,
                           const uint8_t* __restrict__ lists, const uint32_t* __restrict__ counts,
                           uint32_t L, unsigned long long* __restrict__ htab,
                           uint32_t* __restrict__ howner, uint32_t* __restrict__ cslot)
{
    if (cnt->halt)
        return;
    ftd_hash_body(cnt, act, lists, counts, L, htab, howner, cslot,
                  blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__device__ inline void ftd_gid_body(FtdCounts* __restrict__ cnt,
                                    const unsigned long long* __restrict__ htab,
                                    const uint32_t* __restrict__ howner,
                                    uint32_t* __restrict__ hgid, uint32_t* __restrict__ gowner,
                                    uint32_t i0, uint32_t stride)
{
    const uint32_t T = cnt->hT;
    for (uint32_t s = i0; s < T; s += stride)
        if (htab[s])
        {
            const uint32_t g = atomicAdd(&cnt->n_groups, 1u);
            hgid[s]          = g;
            gowner[g]        = howner[s];
        }
}

__global__ void k_ftd_gid(FtdCounts* __restrict__ cnt, const unsigned long long* __restrict__ htab,
                          const uint32_t* __restrict__ howner, uint32_t* __restrict__ hgid,
                          uint32_t* __restrict__ gowner)
{
    if (cnt->halt)
        return;
    ftd_gid_body(cnt, htab, howner, hgid, gowner, blockIdx.x * blockDim.x + threadIdx.x,
                 gridDim.x * blockDim.x);
}

__device__ inline void ftd_member_body(FtdCounts* __restrict__ cnt,
                                       const uint32_t* __restrict__ act,
                                       const uint8_t* __restrict__ lists,
                                       const uint32_t* __restrict__ counts, uint32_t L,
                                       const uint32_t* __restrict__ cslot,
                                       const uint32_t* __restrict__ hgid,
                                       const uint32_t* __restrict__ gowner,
                                       uint32_t* __restrict__ gcount, uint32_t* __restrict__ cgid,
                                       uint32_t* __restrict__ cpos, uint32_t i0, uint32_t stride)
{
    const uint32_t n  = cnt->n_active;
    uint32_t       mc = 0;
    for (uint32_t i = i0; i < n; i += stride)
    {
        const uint32_t c  = act[i];
        const uint32_t g  = hgid[cslot[i]];
        const uint32_t o  = gowner[g];
        const uint32_t kc = counts[c];
        mc                = max(mc, kc);
        if (o != c)
        {
            uint64_t a[4], b[4];
            ftd_key(lists + (size_t)c * L, kc, a);
            ftd_key(lists + (size_t)o * L, counts[o], b);
            if (a[0] != b[0] || a[1] != b[1] || a[2] != b[2] || a[3] != b[3])
                atomicOr(&cnt->collision, 1u);
        }
        cgid[i] = g;
        cpos[i] = atomicAdd(&gcount[g], 1u);
    }
    // (every lane of the wave reaches this point; counts are below 2^24)
    const float wm = wave_max((float)mc);
    if ((threadIdx.x & 63) == 0 && wm > 0.f)
        atomicMax(&cnt->max_count, (uint32_t)wm);
}

__global__ void k_ftd_member(FtdCounts* __restrict__ cnt, const uint32_t* __restrict__ act,
                             const uint8_t* __restrict__ lists,
                             const uint32_t* __restrict__ counts, uint32_t L,
                             const uint32_t* __restrict__ cslot, const uint32_t* __restrict__ hgid,
                             const uint32_t* __restrict__ gowner, uint32_t* __restrict__ gcount,
                             uint32_t* __restrict__ cgid, uint32_t* __restrict__ cpos)
{
    if (cnt->halt)
        return;
    ftd_member_body(cnt, act, lists, counts, L, cslot, hgid, gowner, gcount, cgid, cpos,
                    blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// The grouping of a round of at most FTD_SMALL_NT active cells in one block
// (the five launches above as barrier-separated phases of one launch, the
// same bodies): a round of a few cells pays one launch instead of five.
constexpr int FTD_SMALL_NT = 1024;

__global__ __launch_bounds__(FTD_SMALL_NT) void k_ftd_group_small(
    FtdCounts* __restrict__ cnt, uint32_t T, uint32_t* __restrict__ act,
    const uint32_t* __restrict__ act_next, unsigned long long* __restrict__ htab,
    uint32_t* __restrict__ gcount, const uint8_t* __restrict__ lists,
    const uint32_t* __restrict__ counts, uint32_t L, uint32_t* __restrict__ howner,
    uint32_t* __restrict__ cslot, uint32_t* __restrict__ hgid, uint32_t* __restrict__ gowner,
    uint32_t* __restrict__ cgid, uint32_t* __restrict__ cpos)
{
    if (cnt->halt)
        return;
    const uint32_t t = threadIdx.x;
    if (t == 0)
        ftd_counters_body(cnt, T);
    __syncthreads();
    ftd_begin_body(cnt, act, act_next, htab, gcount, t, FTD_SMALL_NT);
    __syncthreads();
    ftd_hash_body(cnt, act, lists, counts, L, htab, howner, cslot, t, FTD_SMALL_NT);
    __syncthreads();
    ftd_gid_body(cnt, htab, howner, hgid, gowner, t, FTD_SMALL_NT);
    __syncthreads();
    ftd_member_body(cnt, act, lists, counts, L, cslot, hgid, gowner, gcount, cgid, cpos, t,
                    FTD_SMALL_NT);
}

// One block per group: ascending labels, their row offsets, G_sub.
// `exact_ok`: the exact form is wanted (groups of at most 1919 genes use it);
// groups of at most `fuse_max` cells take the fused FP32 form instead.
constexpr int FTD_SETUP_NT = 1024;

// Ordered pair windows staged in LDS at a time by ftd_flag_windows.
constexpr int FTD_PAIRS = 1024;

// The gene flags of a candidate set's G_sub: for every ordered pair (i, j),
// i != j, of lst[0..k), flag[g] = 1 for the first min(len_ij, cut) genes g
// of the pair's window (src/cusingler.cu:1144-1156).  A whole block of NT
// threads: the windows' (pos, len) are staged in `pw` (FTD_PAIRS entries of
// LDS) by parallel loads, then every thread flags genes over the flattened
// (pair, gene) range with eight independent loads in flight -- a set of k
// labels has k (k - 1) windows, and walking them one wave per window left
// the loads of a large set serialised.  Plain stores of the same value; the
// caller clears `flag` before and synchronises after.
template <int NT, class LT>
__device__ inline void ftd_flag_windows(const LT* lst, uint32_t k, uint32_t L, uint32_t cut,
                                        const uint32_t* __restrict__ ctdidx,
                                        const uint32_t* __restrict__ ctdiff, uint8_t* flag,
                                        uint2* pw)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t np  = k * k;
    for (uint32_t p0 = 0; p0 < np; p0 += FTD_PAIRS)
    {
        const uint32_t pn = min(np - p0, (uint32_t)FTD_PAIRS);
        __syncthreads();  // (flags cleared; the previous batch's windows read)
        for (uint32_t q = tid; q < pn; q += NT)
        {
            const uint32_t pi = p0 + q;
            const uint32_t i = lst[pi / k], j = lst[pi % k];
            uint2          w{ 0u, 0u };
            if (i != j)
            {
                w.x = ctdidx[((size_t)i * L + j) * 2];
                w.y = min(ctdidx[((size_t)i * L + j) * 2 + 1], cut);
            }
            pw[q] = w;
        }
        __syncthreads();
        // (a window holds at most `cut` genes: pair q, gene t at flattened
        // index q * cut + t)
        const uint32_t tot = pn * cut;
        for (uint32_t e0 = tid; e0 < tot; e0 += 8 * NT)
        {
            uint32_t g[8];
            bool     v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
            {
                const uint32_t e = e0 + u * NT;
                const uint32_t q = e / cut, t = e - q * cut;
                v[u]             = e < tot && t < pw[min(q, pn - 1)].y;
                g[u]             = v[u] ? ctdiff[pw[q].x + t] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (v[u])
                    flag[g[u]] = 1;
        }
    }
}

__global__ __launch_bounds__(FTD_SETUP_NT) void k_ftd_setup(
    FtdCounts* __restrict__ cnt, const uint32_t* __restrict__ gowner,
    const uint32_t* __restrict__ gcount, const uint8_t* __restrict__ lists,
    const uint32_t* __restrict__ counts, uint32_t L, const uint32_t* __restrict__ ctdidx,
    const uint32_t* __restrict__ ctdiff, const uint32_t* __restrict__ cut_tab,
    const uint32_t* __restrict__ llen, uint32_t Gf, uint32_t gstride, int exact_ok,
    uint32_t fuse_max, FtGroup* __restrict__ groups, uint8_t* __restrict__ gflag,
    uint32_t* __restrict__ glabels,
    uint32_t* __restrict__ glpos, uint32_t* __restrict__ ggenes)
{
    // dynamic LDS: ceil(Gf / 32) bit words, one flag byte per gene (whole
    // words), FTD_PAIRS staged pair windows
    extern __shared__ uint32_t sbits[];
    __shared__ uint32_t slab[256];
    __shared__ uint32_t spart[FTD_SETUP_NT];
    __shared__ uint8_t  sflag[256];
    __shared__ uint32_t sk, snref;
    if (cnt->halt)
        return;
    const uint32_t ng  = cnt->n_groups;
    const uint32_t W   = (Gf + 31) / 32;
    const uint32_t tid = threadIdx.x;
    for (uint32_t g = blockIdx.x; g < ng; g += gridDim.x)
    {
        __syncthreads();  // the previous group's LDS is no longer read
        if (tid < 64)
        {
            // the owner's candidate labels in ascending order, by the first
            // wave: membership flags, one ballot per 64 labels, ranks from
            // the ballot's lower bits, row offsets from a wave scan of the
            // labels' sizes (the loads of list entries and sizes in parallel)
            const uint32_t o   = gowner[g];
            const uint32_t c   = counts[o];
            const uint8_t* lst = lists + (size_t)o * L;
#pragma unroll
            for (uint32_t w = 0; w < 4; ++w)
                sflag[w * 64 + tid] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (uint32_t t = tid; t < c; t += 64)
                sflag[lst[t]] = 1;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const uint64_t below = (1ull << tid) - 1ull;
            uint32_t       k = 0, nref = 0;
#pragma unroll
            for (uint32_t w = 0; w < 4; ++w)
            {
                const uint32_t l  = w * 64 + tid;
                const bool     in = l < L && sflag[l];
                const uint64_t m  = __ballot(in);
                const uint32_t n  = in ? llen[l] : 0u;
                const int      ex = wave_incl_scan((int)n) - (int)n;
                if (in)
                {
                    const uint32_t r            = k + (uint32_t)__builtin_popcountll(m & below);
                    slab[r]                     = l;
                    glabels[(size_t)g * L + r]  = l;
                    glpos[(size_t)g * L + r]    = nref + (uint32_t)ex;
                }
                k += (uint32_t)__builtin_popcountll(m);
                nref += (uint32_t)__builtin_amdgcn_readlane(ex + (int)n, 63);
            }
            if (tid == 0)
            {
                sk    = k;
                snref = nref;
            }
        }
        uint8_t* gflags = (uint8_t*)(sbits + W);
        for (uint32_t w = tid; w < W * 8; w += FTD_SETUP_NT)
            ((uint32_t*)gflags)[w] = 0u;
        __syncthreads();
        const uint32_t k   = sk;
        const uint32_t cut = k ? cut_tab[k] : 0u;
        // a flag byte per gene of every ordered pair's window (plain stores
        // of the same value, no atomics on the few words a small gene set has)
        if (k > 1)
            ftd_flag_windows<FTD_SETUP_NT>(slab, k, L, cut, ctdidx, ctdiff, gflags,
                                           (uint2*)(gflags + (size_t)W * 32));
        __syncthreads();
        for (uint32_t w = tid; w < W; w += FTD_SETUP_NT)
        {
            const uint32_t* f4 = (const uint32_t*)(gflags + w * 32);
            uint32_t        v  = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q)
            {
                const uint32_t f = f4[q];  // flags of genes 32 w + 4 q .. + 3, one per byte
                v |= ((f & 1u) | ((f >> 7) & 2u) | ((f >> 14) & 4u) | ((f >> 21) & 8u)) << (4 * q);
            }
            sbits[w] = v;
        }
        __syncthreads();
        // ascending gene list: contiguous word ranges per thread, block scan
        const uint32_t per = (W + FTD_SETUP_NT - 1) / FTD_SETUP_NT;
        const uint32_t w0 = min(tid * per, W), w1 = min(w0 + per, W);
        uint32_t       own = 0;
        for (uint32_t w = w0; w < w1; ++w)
            own += __builtin_popcount(sbits[w]);
        spart[tid] = own;
        __syncthreads();
        for (uint32_t off = 1; off < FTD_SETUP_NT; off <<= 1)
        {
            const uint32_t v = tid >= off ? spart[tid - off] : 0u;
            __syncthreads();
            spart[tid] += v;
            __syncthreads();
        }
        uint32_t       at  = spart[tid] - own;
        const uint32_t gs  = spart[FTD_SETUP_NT - 1];
        uint32_t*      dst = ggenes + (size_t)g * gstride;
        for (uint32_t w = w0; w < w1; ++w)
        {
            uint32_t v = sbits[w];
            while (v)
            {
                dst[at++] = w * 32 + (uint32_t)__builtin_ctz(v);
                v &= v - 1;
            }
        }
        if (tid == 0)
        {
            FtGroup G{};
            G.gene_off  = (uint32_t)((size_t)g * gstride);
            G.gs        = gs;
            G.stride    = (gs + 7) & ~7u;
            G.nref      = snref;
            G.ncell     = gcount[g];
            G.label_off = (uint32_t)((size_t)g * L);
            G.nlabel    = k;
            groups[g]   = G;
            // fewer than 20 genes: the round returns {front}
            // (src/cusingler.cu:1158-1159)
            const bool fused = G.ncell <= fuse_max;
            gflag[g]         = (uint8_t)((gs < 20 ? FTD_SHORT : 0) | (fused ? FTD_FUSED : 0) |
                                 (!fused && exact_ok && gs <= 1919 ? FTD_EXACT : 0));
            if (fused && gs >= 20)
                atomicAdd(&cnt->n_fused, 1u);
        }
    }
}

// Exclusive prefix sums of the groups' work sizes (one block of FTD_SCAN_NT
// threads): row, element, score and segment offsets, tile offsets, totals.
// FP32-form groups of at most `thin_max` cells use thin tiles.
constexpr int FTD_SCAN_NT = 512;

struct FtdSizes
{
    uint64_t ref_items, cell_items, ref_elems, cell_elems, scores, tiles, thin, segs, rthin;
    uint64_t etiles, btiles;  // exact 16-cell and FTD_BIG-cell tiles (of tiles)
};

// Exact-form groups of at least t64 cells use FTD_BIG-cell tiles, the others
// 16-cell tiles.
__host__ __device__ inline FtdSizes ftd_sizes(const FtGroup& G, uint8_t flag, uint32_t thin_max,
                                              uint32_t t64)
{
    FtdSizes z{};
    if (flag & FTD_SHORT)
        return z;
    const uint64_t nc = G.ncell, nr = G.nref;
    z.ref_items       = nr;
    z.cell_items      = nc;
    z.ref_elems       = nr * G.stride;
    z.cell_elems      = nc * G.stride;
    z.scores          = nc * nr;
    z.segs            = nc * G.nlabel;
    if (flag & FTD_FUSED)
        z.ref_elems = 0;  // the fused pass stores no reference ranks and no tiles
    else if (flag & FTD_EXACT)
    {
        const uint64_t tc = nc >= t64 ? FTD_BIG : 16;
        z.tiles           = ((nc + tc - 1) / tc) * ((nr + 255) / 256);
        z.rthin           = nc * ((nr + 255) / 256);
        (tc == 16 ? z.etiles : z.btiles) = z.tiles;
    }
    else if (nc <= thin_max)
        z.thin = (nr + 255) / 256;
    else
        z.tiles = ((nc + FT_TC - 1) / FT_TC) * ((nr + FT_TR - 1) / FT_TR);
    return z;
}

//
// `acc` (the round's own scan, not a batch range's): the round's totals are
// added to the call's tot_* counters and pair_genes.  `chain` (a round of a
// chain launched without a host read-back): a round whose totals exceed
// `caps`, which needs a kernel family outside `chain_mask` (the families the
// chain launches), or whose grouping met a collision, halts the chain
// instead -- its active list goes back to n_next and its work counts to zero,
// so that the rest of the round does nothing and the host can run it again
// from part A.
__global__ __launch_bounds__(FTD_SCAN_NT) void k_ftd_scan(FtdCounts* __restrict__ cnt,
                                                          FtGroup* __restrict__ groups,
                                                          const uint8_t* __restrict__ gflag,
                                                          uint32_t thin_max, uint32_t t64,
                                                          uint32_t* __restrict__ gtile0,
                                                          uint32_t* __restrict__ gthin0,
                                                          uint64_t* __restrict__ gseg0,
                                                          int acc, int chain,
                                                          uint32_t chain_mask, FtdCaps caps)
{
    constexpr int NQ = 12;
    __shared__ uint64_t part[NQ][FTD_SCAN_NT];
    if (cnt->halt)
        return;
    // groups [lo, hi) of the current batch (all groups unless the host set a
    // range); the batch's work counters start at zero
    const uint32_t      lo  = cnt->g_lo;
    const uint32_t      hi  = min(cnt->g_hi, cnt->n_groups);
    const uint32_t      ng  = hi > lo ? hi - lo : 0u;
    const uint32_t      tid = threadIdx.x;
    const uint32_t      per = (ng + FTD_SCAN_NT - 1) / FTD_SCAN_NT;
    const uint32_t      g0 = lo + min(tid * per, ng), g1 = lo + min(tid * per + per, ng);
    uint64_t            sum[NQ] = {};
    uint32_t            smax    = 0;
    for (uint32_t g = g0; g < g1; ++g)
    {
        const FtdSizes z = ftd_sizes(groups[g], gflag[g], thin_max, t64);
        sum[0] += z.ref_items;
        sum[1] += z.cell_items;
        sum[2] += z.ref_elems;
        sum[3] += z.cell_elems;
        sum[4] += z.scores;
        sum[5] += z.tiles;
        sum[6] += z.thin;
        sum[7] += z.segs;
        sum[8] += z.rthin;
        sum[9] += z.etiles;
        sum[10] += z.btiles;
        if (!(gflag[g] & FTD_SHORT))
        {
            const FtGroup& G = groups[g];
            smax = max(smax, G.stride);
            sum[11] += (uint64_t)G.ncell * G.nref * G.gs;
        }
    }
    for (int q = 0; q < NQ; ++q)
        part[q][tid] = sum[q];
    __syncthreads();
    for (uint32_t off = 1; off < FTD_SCAN_NT; off <<= 1)
    {
        uint64_t v[NQ];
        for (int q = 0; q < NQ; ++q)
            v[q] = tid >= off ? part[q][tid - off] : 0ull;
        __syncthreads();
        for (int q = 0; q < NQ; ++q)
            part[q][tid] += v[q];
        __syncthreads();
    }
    uint64_t run[NQ];
    for (int q = 0; q < NQ; ++q)
        run[q] = part[q][tid] - sum[q];
    for (uint32_t g = g0; g < g1; ++g)
    {
        FtGroup&       G = groups[g];
        const FtdSizes z = ftd_sizes(G, gflag[g], thin_max, t64);
        G.ref_item0      = (uint32_t)run[0];
        G.cell_item0     = (uint32_t)run[1];
        G.ref_dst0       = run[2];
        G.cell_dst0      = run[3];
        G.score0         = run[4];
        gtile0[g]        = (uint32_t)run[5];
        gthin0[g]        = (uint32_t)run[6];
        gseg0[g]         = run[7];
        run[0] += z.ref_items;
        run[1] += z.cell_items;
        run[2] += z.ref_elems;
        run[3] += z.cell_elems;
        run[4] += z.scores;
        run[5] += z.tiles;
        run[6] += z.thin;
        run[7] += z.segs;
        run[8] += z.rthin;
        run[9] += z.etiles;
        run[10] += z.btiles;
    }
    if (smax)
        atomicMax(&cnt->max_stride, smax);
    if (tid == 0)
    {
        cnt->n_redo      = 0;
        cnt->n_rthin     = 0;
        cnt->n_nan       = 0;
        cnt->n_nan2      = 0;
        cnt->n_redo_segs = 0;
    }
    if (tid == FTD_SCAN_NT - 1)
    {
        uint64_t t[NQ];
        for (int q = 0; q < NQ; ++q)
            t[q] = part[q][tid];
        // (every table of a count is at least as large as the cap given)
        const bool over = t[0] > caps.ref_items || t[1] > caps.cells || t[2] > caps.ref_elems ||
                          t[3] > caps.cell_elems || t[4] > caps.scores || t[7] > caps.segs ||
                          t[5] > caps.tiles || t[6] > caps.thin || t[8] > caps.rthin;
        const uint32_t need = ftd_mask_need(t[2], (uint32_t)t[5], (uint32_t)t[9],
                                            (uint32_t)t[10], (uint32_t)t[6], cnt->n_fused);
        if (chain && (over || (need & ~chain_mask) || cnt->collision))
        {
            // the round is run again from its part A by the host: its
            // active list (left in act_next by its compaction) is listed
            // again and nothing of the round's work runs
            cnt->halt     = 1;
            cnt->n_next   = cnt->n_active;
            cnt->round   -= 1;
            cnt->g_lo     = 1;
            cnt->g_hi     = 0;
            for (int q = 0; q < NQ; ++q)
                t[q] = 0;
        }
        else if (acc && !cnt->collision)
        {
            cnt->tot_rounds += cnt->n_active ? 1u : 0u;
            cnt->tot_active += cnt->n_active;
            cnt->tot_groups += cnt->n_groups;
            cnt->tot_scores += t[4];
            cnt->tot_fused += cnt->n_fused;
            cnt->pair_genes += t[11];
        }
        cnt->n_ref_items = t[0];
        cnt->n_cells     = (uint32_t)t[1];
        cnt->ref_elems   = t[2];
        cnt->cell_elems  = t[3];
        cnt->n_scores    = t[4];
        cnt->n_tiles     = (uint32_t)t[5];
        cnt->n_thin      = (uint32_t)t[6];
        cnt->n_segs      = t[7];
        cnt->rthin_bound = t[8];
        cnt->n_etiles    = (uint32_t)t[9];
        cnt->n_btiles    = (uint32_t)t[10];
    }
}

// Cells with work get their rank item and FtCell (ordered by group, then
// position, so that their segment offsets increase); cells of groups with
// fewer than 20 genes end with {front}.  Each segment records its cell
// (segcell), each reference rank item its group (k_ftd_refmap), so that the
// work-queue kernels find them without a search.  Idempotent (the round is
// counted by k_ftd_compact), so the autotuner may run it ahead of the
// round's graph.
__device__ inline void ftd_cells_body(const FtdCounts* __restrict__ cnt, const uint32_t* __restrict__ act,
                            const uint32_t* __restrict__ cgid, const uint32_t* __restrict__ cpos,
                            const FtGroup* __restrict__ groups, const uint8_t* __restrict__ gflag,
                            const uint64_t* __restrict__ gseg0, uint32_t* __restrict__ counts,
                            FtRankItem* __restrict__ items,
                            FtCell* __restrict__ cells, uint32_t* __restrict__ segcell)
{
    const uint32_t n = cnt->n_active;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint32_t c = act[i];
        const uint32_t g = cgid[i];
        if (g < cnt->g_lo || g >= cnt->g_hi)
            continue;  // another batch
        if (gflag[g] & FTD_SHORT)
        {
            counts[c] = 1;
            continue;
        }
        const FtGroup  G   = groups[g];
        const uint32_t pos = cpos[i];
        const uint32_t idx = G.cell_item0 + pos;
        const uint64_t s0  = gseg0[g] + (uint64_t)pos * G.nlabel;
        items[idx]         = FtRankItem{ c, g, G.cell_dst0 + (uint64_t)pos * G.stride };
        cells[idx]         = FtCell{ c, g, pos, 0u, s0 };
        for (uint32_t j = 0; j < G.nlabel; ++j)
            segcell[s0 + j] = idx;
    }
}

// Group and source row of every reference rank item (one block per group;
// the group's rows are its labels' rows in ascending label order).
__device__ inline void ftd_refmap_body(const FtdCounts* __restrict__ cnt, const FtGroup* __restrict__ groups,
                             const uint8_t* __restrict__ gflag, const uint32_t* __restrict__ labels,
                             const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ llen,
                             uint32_t* __restrict__ refgroup, uint32_t* __restrict__ refrow)
{
    const uint32_t hi = min(cnt->g_hi, cnt->n_groups);
    for (uint32_t g = cnt->g_lo + blockIdx.x; g < hi; g += gridDim.x)
    {
        if (gflag[g] & FTD_SHORT)
            continue;
        const FtGroup G   = groups[g];
        uint32_t      off = G.ref_item0;
        for (uint32_t t = 0; t < G.nlabel; ++t)
        {
            const uint32_t l = labels[G.label_off + t], n = llen[l], s0 = lstart[l];
            for (uint32_t r = threadIdx.x; r < n; r += blockDim.x)
            {
                refgroup[off + r] = g;
                refrow[off + r]   = s0 + r;
            }
            off += n;
        }
    }
}

// Correlation tile lists of every group with work.
__device__ inline void ftd_tiles_body(const FtdCounts* __restrict__ cnt, const FtGroup* __restrict__ groups,
                            const uint8_t* __restrict__ gflag, uint32_t thin_max, uint32_t t64,
                            const uint32_t* __restrict__ gtile0, const uint32_t* __restrict__ gthin0,
                            FtCorrTile* __restrict__ tiles, FtCorrTile* __restrict__ thin)
{
    const uint32_t hi = min(cnt->g_hi, cnt->n_groups);
    for (uint32_t g = cnt->g_lo + blockIdx.x * blockDim.x + threadIdx.x; g < hi;
         g += gridDim.x * blockDim.x)
    {
        const uint8_t f = gflag[g];
        if (f & FTD_SHORT)
            continue;
        const FtGroup G = groups[g];
        if (G.ncell == 0 || (f & FTD_FUSED))
            continue;
        if (f & FTD_EXACT)
        {
            // lim marks the tile height: 0 for 16 cells, FTD_BIG for the big tile
            const uint32_t tc = G.ncell >= t64 ? (uint32_t)FTD_BIG : 16u;
            uint32_t       t  = gtile0[g];
            for (uint32_t c0 = 0; c0 < G.ncell; c0 += tc)
                for (uint32_t r0 = 0; r0 < G.nref; r0 += 256)
                    tiles[t++] = FtCorrTile{ g, c0, r0, tc == 16 ? 0u : tc };
        }
        else if (G.ncell <= thin_max)
        {
            uint32_t t = gthin0[g];
            for (uint32_t r0 = 0; r0 < G.nref; r0 += 256)
                thin[t++] = FtCorrTile{ g, 0u, r0, 0u };
        }
        else
        {
            uint32_t t = gtile0[g];
            for (uint32_t c0 = 0; c0 < G.ncell; c0 += FT_TC)
                for (uint32_t r0 = 0; r0 < G.nref; r0 += FT_TR)
                    tiles[t++] = FtCorrTile{ g, c0, r0, 0u };
        }
    }
}

__global__ void k_ftd_cells(const FtdCounts* __restrict__ cnt, const uint32_t* __restrict__ act,
                            const uint32_t* __restrict__ cgid, const uint32_t* __restrict__ cpos,
                            const FtGroup* __restrict__ groups, const uint8_t* __restrict__ gflag,
                            const uint64_t* __restrict__ gseg0, uint32_t* __restrict__ counts,
                            FtRankItem* __restrict__ items, FtCell* __restrict__ cells,
                            uint32_t* __restrict__ segcell)
{
    ftd_cells_body(cnt, act, cgid, cpos, groups, gflag, gseg0, counts, items, cells, segcell);
}

__global__ void k_ftd_refmap(const FtdCounts* __restrict__ cnt, const FtGroup* __restrict__ groups,
                             const uint8_t* __restrict__ gflag, const uint32_t* __restrict__ labels,
                             const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ llen,
                             uint32_t* __restrict__ refgroup, uint32_t* __restrict__ refrow)
{
    ftd_refmap_body(cnt, groups, gflag, labels, lstart, llen, refgroup, refrow);
}

__global__ void k_ftd_tiles(const FtdCounts* __restrict__ cnt, const FtGroup* __restrict__ groups,
                            const uint8_t* __restrict__ gflag, uint32_t thin_max, uint32_t t64,
                            const uint32_t* __restrict__ gtile0, const uint32_t* __restrict__ gthin0,
                            FtCorrTile* __restrict__ tiles, FtCorrTile* __restrict__ thin)
{
    ftd_tiles_body(cnt, groups, gflag, thin_max, t64, gtile0, gthin0, tiles, thin);
}

// The three set-up passes of part B in one launch (they only read the scan's
// results and write disjoint tables): cells, reference maps, tile lists.
__global__ void k_ftd_prep(const FtdCounts* __restrict__ cnt, const uint32_t* __restrict__ act,
                           const uint32_t* __restrict__ cgid, const uint32_t* __restrict__ cpos,
                           const FtGroup* __restrict__ groups, const uint8_t* __restrict__ gflag,
                           const uint64_t* __restrict__ gseg0, uint32_t* __restrict__ counts,
                           FtRankItem* __restrict__ items, FtCell* __restrict__ cells,
                           uint32_t* __restrict__ segcell, const uint32_t* __restrict__ labels,
                           const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ llen,
                           uint32_t* __restrict__ refgroup, uint32_t* __restrict__ refrow,
                           uint32_t thin_max, uint32_t t64, const uint32_t* __restrict__ gtile0,
                           const uint32_t* __restrict__ gthin0, FtCorrTile* __restrict__ tiles,
                           FtCorrTile* __restrict__ thin)
{
    ftd_cells_body(cnt, act, cgid, cpos, groups, gflag, gseg0, counts, items, cells, segcell);
    ftd_refmap_body(cnt, groups, gflag, labels, lstart, llen, refgroup, refrow);
    ftd_tiles_body(cnt, groups, gflag, thin_max, t64, gtile0, gthin0, tiles, thin);
}

// ---- work-queue kernels: the grid covers the machine, items come from the
// ---- device-side counts

// The last step of a wave's rank table: bins[0] = the zero count (the zero
// code is counted by ballots, never by the LDS atomics), then the code
// counts become bins[c] = 2 (#less) + (#equal) - gs by wave scans: up to
// 512 bins window by window (lane l takes bin 64 w + l, so the LDS accesses
// are free of bank conflicts, a running carry joins the windows), beyond
// that one scan over contiguous runs of ceil(nbins / 64) bins per lane.
__device__ inline void ftd_bins_to_ranks(int* bins, int nbins, int gs, int zeros, int lane)
{
    if (lane == 0)
        bins[0] = zeros;
    __builtin_amdgcn_wave_barrier();
    if (nbins <= 512)
    {
        int carry = 0;
        for (int b0 = 0; b0 < nbins; b0 += 64)
        {
            const int i    = b0 + lane;
            const int c    = i < nbins ? bins[i] : 0;
            const int incl = wave_incl_scan(c);
            if (i < nbins)
                bins[i] = 2 * (carry + incl - c) + c - gs;
            carry += __builtin_amdgcn_readlane(incl, 63);
        }
        __builtin_amdgcn_wave_barrier();
        return;
    }
    const int per = (nbins + 63) / 64;
    const int b0 = min(lane * per, nbins), b1 = min(b0 + per, nbins);
    int       s  = 0;
    for (int i = b0; i < b1; ++i)
        s += bins[i];
    int start = wave_incl_scan(s) - s;
    for (int i = b0; i < b1; ++i)
    {
        const int c = bins[i];
        bins[i]     = 2 * start + c - gs;
        start += c;
    }
    __builtin_amdgcn_wave_barrier();
}

// Doubled centred rank table of one reference row over G_sub, built by one
// wave in LDS (codes stored as CT: 8 bits when every code is below 256, else
// 16): codes[i] = the row's code at gene gl[i] (all gathered first,
// so that the loads overlap), the code histogram (the zero code -- most of a
// row -- counted by ballots rather than LDS atomics), and
// bins[c] = 2 (#less) + (#equal) - gs from a wave scan over the bins.
template <class CT>
__device__ inline void ftd_row_table(const CT* __restrict__ src, const uint32_t* __restrict__ gl,
                                     int gs, int nbins, int* bins, uint16_t* codes, int lane)
{
    for (int i = lane; i < nbins; i += 64)
        bins[i] = 0;
#pragma unroll 4
    for (int i = lane; i < gs; i += 64)
        codes[i] = src[gl[i]];
    __builtin_amdgcn_wave_barrier();
    int zeros = 0;
    for (int i0 = 0; i0 < gs; i0 += 64)
    {
        const int      i = i0 + lane;
        const uint16_t c = i < gs ? codes[i] : (uint16_t)0;
        zeros += __builtin_popcountll(__ballot(i < gs && c == 0));
        if (i < gs && c != 0)
            atomicAdd(&bins[c], 1);
    }
    ftd_bins_to_ranks(bins, nbins, gs, zeros, lane);
}

// Reference rows, one wave per row (4 rows per block, no block barrier): the
// rank table of ftd_row_table (per-wave LDS of `nbins` code bins and the
// row's codes, `wave_ints` ints), the ranks stored, and the 128-chunk +
// stride-tree sum of squares with lane l holding chunks 2l and 2l+1 (their
// sum is the tree's first level) and wave_tree_sum for the remaining six
// levels.
template <class CT>
__device__ inline void ft_rank_ref_wave(uint32_t item, const FtGroup* __restrict__ groups,
                                        const uint32_t* __restrict__ refgroup,
                                        const uint32_t* __restrict__ refrow,
                                        const uint32_t* __restrict__ genes,
                                        const CT* __restrict__ ref_codes, uint32_t Gf,
                                        int nbins, int16_t* __restrict__ out,
                                        float* __restrict__ out_sq, int* bins, uint16_t* codes,
                                        int lane)
{
    const FtGroup  g     = groups[refgroup[item]];
    const uint32_t local = item - g.ref_item0;
    const int      gs    = (int)g.gs;
    ftd_row_table(ref_codes + (size_t)refrow[item] * Gf, genes + g.gene_off, gs, nbins, bins,
                  codes, lane);
    int16_t* dst = out + g.ref_dst0 + (uint64_t)local * g.stride;
    for (int i = lane; i < (int)g.stride; i += 64)
        dst[i] = i < gs ? (int16_t)bins[codes[i]] : (int16_t)0;
    const int step = (gs + 127) / 128;
    float     a0 = 0.f, a1 = 0.f;
    {
        const int lo = 2 * lane * step, hi = min(lo + step, gs);
        for (int i = lo; i < hi; ++i)
        {
            const float v = (float)bins[codes[i]];
            a0            = __builtin_fmaf(v, v, a0);
        }
        const int lo1 = lo + step, hi1 = min(lo1 + step, gs);
        for (int i = lo1; i < hi1; ++i)
        {
            const float v = (float)bins[codes[i]];
            a1            = __builtin_fmaf(v, v, a1);
        }
    }
    const float p = wave_tree_sum(a0 + a1);
    if (lane == 0)
        out_sq[item] = p;
    __builtin_amdgcn_wave_barrier();  // LDS reads done before the next row
}

// Reference rows of every group except the fused ones whose chunk length
// ceil(gs / 128) exceeds step_lo (the others are k_ftd_fused_r's).
template <class CT>
__global__ __launch_bounds__(256) void k_ftd_rank_ref_w(
    const FtdCounts* __restrict__ cnt, const FtGroup* __restrict__ groups,
    const uint8_t* __restrict__ gflag, const uint32_t* __restrict__ refgroup,
    const uint32_t* __restrict__ refrow, const uint32_t* __restrict__ genes,
    const CT* __restrict__ ref_codes, uint32_t Gf, int nbins, int wave_ints,
    int16_t* __restrict__ out, float* __restrict__ out_sq, int step_lo)
{
    extern __shared__ int smem[];
    const int      wid   = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int*           bins  = smem + wid * wave_ints;
    uint16_t*      codes = (uint16_t*)(bins + ((nbins + 3) & ~3));
    const uint32_t n     = (uint32_t)cnt->n_ref_items;
    for (uint32_t item = blockIdx.x * 4 + wid; item < n; item += gridDim.x * 4)
    {
        const uint32_t gi = refgroup[item];
        if ((gflag[gi] & FTD_FUSED) || ((int)groups[gi].gs + 127) / 128 <= step_lo)
            continue;  // uniform per wave
        ft_rank_ref_wave(item, groups, refgroup, refrow, genes, ref_codes, Gf, nbins, out,
                         out_sq, bins, codes, lane);
    }
}

// Fused FP32 form, one wave per reference row of a group of at most
// FTD_FUSE_MAX cells: the row is ranked over G_sub as in ft_rank_ref_wave
// (LDS code histogram, wave scan), and instead of storing its ranks the wave
// forms, in the same pass and the same 128-chunk + stride-tree FP32 order as
// the thin tile (lane l sums chunks 2l and 2l + 1 sequentially by fma,
// wave_tree_sum completes the tree), the row's sum of squares and its dot
// product with each of the group's cell rank rows, and stores the group's
// scores corr_from_sums(sxy, sxx, syy) -- the same values the thin tile
// computes from stored ranks.  Takes the fused groups whose chunk length
// ceil(gs / 128) exceeds step_lo (the others are k_ftd_fused_r's).
template <class CT>
__global__ __launch_bounds__(256) void k_ftd_fused(
    const FtdCounts* __restrict__ cnt, const FtGroup* __restrict__ groups,
    const uint8_t* __restrict__ gflag, const uint32_t* __restrict__ refgroup,
    const uint32_t* __restrict__ refrow, const uint32_t* __restrict__ genes,
    const CT* __restrict__ ref_codes, uint32_t Gf, int nbins, int wave_ints,
    const int16_t* __restrict__ cell_rank, const float* __restrict__ cell_sq,
    float* __restrict__ scores, int step_lo)
{
    extern __shared__ int smem[];
    const int      wid   = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int*           bins  = smem + wid * wave_ints;
    uint16_t*      codes = (uint16_t*)(bins + ((nbins + 3) & ~3));
    const uint32_t n     = (uint32_t)cnt->n_ref_items;
    for (uint32_t item = blockIdx.x * 4 + wid; item < n; item += gridDim.x * 4)
    {
        const uint32_t gi = refgroup[item];
        if (!(gflag[gi] & FTD_FUSED))
            continue;  // uniform per wave
        const FtGroup  g     = groups[gi];
        const uint32_t local = item - g.ref_item0;
        const int      gs    = (int)g.gs;
        if ((gs + 127) / 128 <= step_lo)
            continue;  // uniform per wave
        ftd_row_table(ref_codes + (size_t)refrow[item] * Gf, genes + g.gene_off, gs, nbins, bins,
                      codes, lane);
        const int      nc   = (int)min(g.ncell, (uint32_t)FTD_FUSE_MAX);
        const int      st   = (int)g.stride;
        const int16_t* xc   = cell_rank + g.cell_dst0;
        const int      step = (gs + 127) / 128;
        float          q[2] = { 0.f, 0.f };
        float          d[2][FTD_FUSE_MAX];
#pragma unroll
        for (int h = 0; h < 2; ++h)
        {
#pragma unroll
            for (int c = 0; c < FTD_FUSE_MAX; ++c)
                d[h][c] = 0.f;
            const int lo = (2 * lane + h) * step, hi = min(lo + step, gs);
            for (int i = lo; i < hi; ++i)
            {
                const float y = (float)bins[codes[i]];
                q[h]          = __builtin_fmaf(y, y, q[h]);
#pragma unroll
                for (int c = 0; c < FTD_FUSE_MAX; ++c)
                    if (c < nc)
                        d[h][c] = __builtin_fmaf((float)xc[c * st + i], y, d[h][c]);
            }
        }
        const float syy = wave_tree_sum(q[0] + q[1]);
#pragma unroll
        for (int c = 0; c < FTD_FUSE_MAX; ++c)
        {
            if (c >= nc)
                break;
            const float sxy = wave_tree_sum(d[0][c] + d[1][c]);
            if (lane == 0)
                scores[g.score0 + (uint64_t)c * g.nref + local] =
                    corr_from_sums(sxy, cell_sq[g.cell_item0 + c], syy);
        }
        __builtin_amdgcn_wave_barrier();  // LDS reads done before the next row
    }
}

// Most consecutive reference rows one wave of k_ftd_fused_r takes at a time.
constexpr int FTD_FUSED_RUN = 64;

// Two cells' rank values of one element, in half precision: a doubled
// centred rank 2 r - (gs + 1) over gs <= 2048 genes is an integer with
// |x| <= gs - 1 <= 2047, which FP16 (11 significant bits) holds exactly.
typedef _Float16 ftd_h2 __attribute__((ext_vector_type(2)));

// Chunk sums of one row for a group of NC cells (lane l: chunks 2l + h of
// `step` elements, element (2l + h) step + k in slot [h][k], bit h STEP + k
// of `ok` set for the slots inside G_sub).  A slot outside has y = 0 (and
// some finite rank value x): its fmas add +-0, which leaves the sums' bits
// as they are (no sum here is ever -0), so the sums are those of the
// sequential loops over the valid elements.
template <int STEP, int NC>
__device__ inline void ftd_fused_row(const int* bins, const uint32_t (&code)[2][STEP],
                                     const ftd_h2 (&x)[2][STEP][FTD_FUSE_MAX / 2], uint32_t ok,
                                     float& syy, float (&sxy)[FTD_FUSE_MAX])
{
    float q[2] = { 0.f, 0.f };
    float d[2][NC];
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
#pragma unroll
        for (int c = 0; c < NC; ++c)
            d[h][c] = 0.f;
#pragma unroll
        for (int k = 0; k < STEP; ++k)
        {
            const int   v = bins[code[h][k]];
            const float y = (ok >> (h * STEP + k)) & 1u ? (float)v : 0.f;
            q[h]          = __builtin_fmaf(y, y, q[h]);
#pragma unroll
            for (int c = 0; c < NC; ++c)
                d[h][c] = __builtin_fmaf((float)x[h][k][c >> 1][c & 1], y, d[h][c]);
        }
    }
    syy = q[0] + q[1];
#pragma unroll
    for (int c = 0; c < NC; ++c)
        sxy[c] = d[0][c] + d[1][c];
}

// Fused FP32 form with the cells held in registers, for byte codes (at most
// 256 code bins): the scores of k_ftd_fused (the same rank tables, chunk
// sums and tree, so the same bits) for the fused groups whose chunk length
// ceil(gs / 128) lies in (STEP_LO, STEP].  A wave takes runs of consecutive
// reference items (one vector load of their groups and flags, a ballot of
// the fused ones): ceil(items / waves) of them, at most FTD_FUSED_RUN, so
// that a wave's rows are consecutive rows of one group as far as possible
// (blocks outnumber the resident slots, so uneven runs even out).
// Consecutive rows of a group share its gene list and cells, so a lane
// loads its chunk elements' genes and the cells' rank values once per group
// and keeps them in registers (the ranks as FP16 pairs, exact up to 2048
// genes, chunk length 16).  Per row: the row's codes gathered at the
// lane's chunk elements, counted into the wave's 256 LDS bins (ballots for
// code 0, LDS atomics for the rest), the counts turned into rank values by
// a DPP scan (lane l holds bins 4l .. 4l + 3: one 16-byte read and write),
// the chunk sums accumulated from the codes the lane still holds, and the
// trees completed by DPP (wave_tree_sum) -- no LDS traffic beyond the bins.
//
// The (4, 8] instance is held to 128 VGPRs (4 waves per SIMD instead of 3,
// two values spilled outside the row loop).
//
// RANKS = true: the same rows of the groups that are not fused, whose ranks
// are stored for the tiles (k_ftd_rank_ref_w's output): the rank values go
// through an LDS row (chunk layout in, lane-strided out) so that the row is
// stored coalesced, zero-padded to the group's stride, with the row's
// chunk-tree sum of squares.
template <int STEP, int STEP_LO, bool RANKS = false>
__global__ __launch_bounds__(256, STEP == 8 && !RANKS ? 4 : 1) void k_ftd_fused_r(
    const FtdCounts* __restrict__ cnt, const FtGroup* __restrict__ groups,
    const uint8_t* __restrict__ gflag, const uint32_t* __restrict__ refgroup,
    const uint32_t* __restrict__ refrow, const uint32_t* __restrict__ genes,
    const uint8_t* __restrict__ ref_codes, uint32_t Gf, const int16_t* __restrict__ cell_rank,
    const float* __restrict__ cell_sq, float* __restrict__ scores,
    int16_t* __restrict__ out_rank = nullptr, float* __restrict__ out_sq = nullptr)
{
    static_assert(STEP <= 16, "the slot mask holds 2 STEP bits; FP16 ranks are exact to 2048 genes");
    // two copies of the 256 bins per wave (lanes 0-31 and 32-63 count into
    // their own, halving the same-address conflicts of the atomics; the
    // second starts 4 banks on), summed into the first for the rank values
    __shared__ __attribute__((aligned(16))) int sbins[4][2 * 256 + 4];
    __shared__ __attribute__((aligned(16))) int16_t srow[4][RANKS ? 128 * STEP : 2];
    const int      wid   = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int*           bins  = sbins[wid];
    int4*          bins4 = (int4*)bins;
    int*           mybin = bins + (lane >> 5) * (256 + 4);
    const uint64_t n     = cnt->n_ref_items;
    uint32_t       cur   = 0xFFFFFFFFu;
    bool           mine  = false;
    FtGroup        g{};
    int            gs = 0, nc = 0;
    uint32_t       ok = 0;
    uint32_t       gl[2][STEP];
    ftd_h2         x[2][STEP][FTD_FUSE_MAX / 2];
    const uint64_t waves = (uint64_t)gridDim.x * 4;
    const int      rl    = (int)min((uint64_t)FTD_FUSED_RUN,
                                    max((uint64_t)1, (n + waves - 1) / waves));
    for (uint64_t run = (uint64_t)blockIdx.x * 4 + wid; run * rl < n; run += waves)
    {
        const uint64_t base = run * rl;
        uint32_t       rg = 0, rr = 0;
        bool           fz = false;
        if (lane < rl && base + lane < n)
        {
            rg = refgroup[base + lane];
            rr = refrow[base + lane];
            fz = ((gflag[rg] & FTD_FUSED) != 0) != RANKS;
        }
        uint64_t m = __ballot(fz);
        // the next row's codes, gathered while this row is ranked (only
        // when it is a row of the same group: the gene registers hold)
        bool     pf = false;
        uint32_t pcode[2][STEP];
        while (m)
        {
            const int k = __builtin_ctzll(m);
            m &= m - 1;
            const uint32_t gi   = __builtin_amdgcn_readlane(rg, k);
            const uint64_t item = base + (uint64_t)k;
            if (gi != cur)
            {
                pf = false;
                cur            = gi;
                g              = groups[gi];
                gs             = (int)g.gs;
                const int step = (gs + 127) / 128;
                nc             = (int)min(g.ncell, (uint32_t)FTD_FUSE_MAX);
                mine           = step > STEP_LO && step <= STEP;
                if (mine)
                {
                    // every load unconditional, so that they are all in
                    // flight together: a slot outside G_sub takes element 0
                    // (its code is gathered but never counted, its rank
                    // value multiplies y = 0), a cell past nc takes cell 0
                    // (never read)
                    const int16_t* xc = cell_rank + g.cell_dst0;
                    const int      st = (int)g.stride;
                    ok                = 0;
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int j = 0; j < STEP; ++j)
                        {
                            const int  i  = (2 * lane + h) * step + j;
                            const bool v  = j < step && i < gs;
                            const int  ic = v ? i : 0;
                            ok |= (uint32_t)v << (h * STEP + j);
                            gl[h][j] = genes[g.gene_off + ic];
                            if constexpr (!RANKS)
                            {
#pragma unroll
                                for (int c = 0; c < FTD_FUSE_MAX; ++c)
                                    x[h][j][c >> 1][c & 1] =
                                        (_Float16)xc[(c < nc ? c : 0) * st + ic];
                            }
                        }
                }
            }
            if (!mine)
                continue;
            uint32_t code[2][STEP];
            if (pf)
            {
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int j = 0; j < STEP; ++j)
                        code[h][j] = pcode[h][j];
            }
            else
            {
                const uint8_t* src =
                    ref_codes + (size_t)__builtin_amdgcn_readlane(rr, k) * Gf;
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int j = 0; j < STEP; ++j)
                        code[h][j] = src[gl[h][j]];
            }
            pf = false;
            if (m)
            {
                const int k2 = __builtin_ctzll(m);
                if ((uint32_t)__builtin_amdgcn_readlane(rg, k2) == cur)
                {
                    const uint8_t* src2 =
                        ref_codes + (size_t)__builtin_amdgcn_readlane(rr, k2) * Gf;
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int j = 0; j < STEP; ++j)
                            pcode[h][j] = src2[gl[h][j]];
                    pf = true;
                }
            }
            bins4[lane]      = make_int4(0, 0, 0, 0);
            bins4[65 + lane] = make_int4(0, 0, 0, 0);
            __builtin_amdgcn_wave_barrier();
            int zeros = 0;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int j = 0; j < STEP; ++j)
                {
                    const bool v = (ok >> (h * STEP + j)) & 1u;
                    zeros += __builtin_popcountll(__ballot(v && code[h][j] == 0));
                    if (v && code[h][j] != 0)
                        atomicAdd(&mybin[code[h][j]], 1);
                }
            __builtin_amdgcn_wave_barrier();
            int4       b  = bins4[lane];
            const int4 b2 = bins4[65 + lane];
            b.x += b2.x;
            b.y += b2.y;
            b.z += b2.z;
            b.w += b2.w;
            if (lane == 0)
                b.x = zeros;  // no atomic touched bin 0
            const int s  = b.x + b.y + b.z + b.w;
            int       ex = wave_incl_scan(s) - s;
            int4      r;
            r.x = 2 * ex + b.x - gs;
            ex += b.x;
            r.y = 2 * ex + b.y - gs;
            ex += b.y;
            r.z = 2 * ex + b.z - gs;
            ex += b.z;
            r.w         = 2 * ex + b.w - gs;
            bins4[lane] = r;
            __builtin_amdgcn_wave_barrier();
            const uint64_t local = item - g.ref_item0;
            if constexpr (RANKS)
            {
                const int step = (gs + 127) / 128;
                int16_t*  row  = srow[wid];
                float     q[2] = { 0.f, 0.f };
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int j = 0; j < STEP; ++j)
                    {
                        if (j >= step)
                            break;
                        // a slot past G_sub stores the padding zero
                        const int y = (ok >> (h * STEP + j)) & 1u ? bins[code[h][j]] : 0;
                        row[(2 * lane + h) * step + j] = (int16_t)y;
                        const float yf = (float)y;
                        q[h]           = __builtin_fmaf(yf, yf, q[h]);
                    }
                __builtin_amdgcn_wave_barrier();
                int16_t* dst = out_rank + g.ref_dst0 + local * g.stride;
                for (int i = lane; i < (int)g.stride; i += 64)
                    dst[i] = row[i];
                const float p = wave_tree_sum(q[0] + q[1]);
                if (lane == 0)
                    out_sq[item] = p;
                __builtin_amdgcn_wave_barrier();  // LDS reads done before the next row
                continue;
            }
            float syy, sxy[FTD_FUSE_MAX];
            switch (nc)
            {
            case 1: ftd_fused_row<STEP, 1>(bins, code, x, ok, syy, sxy); break;
            case 2: ftd_fused_row<STEP, 2>(bins, code, x, ok, syy, sxy); break;
            case 3: ftd_fused_row<STEP, 3>(bins, code, x, ok, syy, sxy); break;
            default: ftd_fused_row<STEP, 4>(bins, code, x, ok, syy, sxy); break;
            }
            syy = wave_tree_sum(syy);
#pragma unroll
            for (int c = 0; c < FTD_FUSE_MAX; ++c)
            {
                if (c >= nc)
                    break;
                const float v = wave_tree_sum(sxy[c]);
                if (lane == 0)
                    scores[g.score0 + (uint64_t)c * g.nref + local] =
                        corr_from_sums(v, cell_sq[g.cell_item0 + c], syy);
            }
            __builtin_amdgcn_wave_barrier();  // LDS reads done before the next row
        }
    }
}

__global__ __launch_bounds__(128) void k_ftd_rank_ref(
    const FtdCounts* __restrict__ cnt, const FtGroup* __restrict__ groups,
    const uint32_t* __restrict__ labels, const uint32_t* __restrict__ lstart,
    const uint32_t* __restrict__ llen, const uint32_t* __restrict__ genes,
    const uint16_t* __restrict__ ref_codes, uint32_t Gf, int nbins, int16_t* __restrict__ out,
    float* __restrict__ out_sq)
{
    extern __shared__ int smem[];
    __shared__ float part[128];
    __shared__ int   tmp[2];
    const uint32_t   n  = (uint32_t)cnt->n_ref_items;
    const uint32_t   ng = cnt->n_groups;
    for (uint32_t item = blockIdx.x; item < n; item += gridDim.x)
    {
        __syncthreads();
        ft_rank_ref_item(item, groups, ng, labels, lstart, llen, genes, ref_codes, Gf, nbins, out,
                         out_sq, smem, part, tmp);
    }
}

// Query rows, one block per cell item; items of groups whose chunk length
// ceil(gs / 128) is at most step_lo are k_ftd_rank_qry_w's.
__global__ __launch_bounds__(128) void k_ftd_rank_qry(
    const FtdCounts* __restrict__ cnt, const FtRankItem* __restrict__ items,
    const FtGroup* __restrict__ groups, const uint32_t* __restrict__ genes,
    const uint16_t* __restrict__ qdata, const int* __restrict__ qind,
    const int* __restrict__ qptr, uint32_t Gf, int nbins, int16_t* __restrict__ out,
    float* __restrict__ out_sq, int step_lo)
{
    extern __shared__ int smem[];
    __shared__ float part[128];
    __shared__ int   tmp[2];
    const uint32_t   n = cnt->n_cells;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x)
    {
        const FtRankItem it = items[i];
        if (((int)groups[it.group].gs + 127) / 128 <= step_lo)
            continue;  // uniform per block
        __syncthreads();
        ft_rank_qry_item(it, i, groups, genes, qdata, qind, qptr, Gf, nbins, out, out_sq, smem,
                         part, tmp);
    }
}

// Per-wave LDS (ints) of k_ftd_rank_qry_w: 256 code bins, the staged rank
// row of 128 STEP int16, the dense row of Gf 16-bit codes (16-byte rounded).
__host__ __device__ constexpr int ftd_qry_wave_ints(int step, uint32_t Gf)
{
    return 256 + 64 * step + (int)(((Gf + 1) / 2 + 3) & ~3u);
}

// Query rows of the fine-tune, one wave per cell item (four per block), for
// codes below 256 and G_sub of at most 128 STEP genes: the cell's CSR row
// densified over the fine-tune genes in the wave's LDS, gathered at G_sub in
// the chunk layout of k_ftd_fused_r (the gene indices kept in registers
// across a wave's run of consecutive items, which mostly share a group),
// counted into 256 LDS bins (ballots for code 0, LDS atomics), turned into
// rank values by a DPP scan over 16-byte bin quads, stored through an LDS
// row (coalesced, zero-padded to the group's stride), with the chunk-tree
// sum of squares (wave_tree_sum: the same additions as the block tree of
// ft_rank_qry_item).
template <int STEP>
__global__ __launch_bounds__(256) void k_ftd_rank_qry_w(
    const FtdCounts* __restrict__ cnt, const FtRankItem* __restrict__ items,
    const FtGroup* __restrict__ groups, const uint32_t* __restrict__ genes,
    const uint16_t* __restrict__ qdata, const int* __restrict__ qind,
    const int* __restrict__ qptr, uint32_t Gf, int16_t* __restrict__ out,
    float* __restrict__ out_sq)
{
    extern __shared__ __attribute__((aligned(16))) int qsm[];
    const int      wid   = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int      wints = ftd_qry_wave_ints(STEP, Gf);
    const int      dints = wints - 256 - 64 * STEP;
    int*           bins  = qsm + wid * wints;
    int4*          bins4 = (int4*)bins;
    int16_t*       row   = (int16_t*)(bins + 256);
    int*           dw    = bins + 256 + 64 * STEP;
    uint16_t*      dense = (uint16_t*)dw;
    const uint32_t n     = cnt->n_cells;
    const uint32_t waves = gridDim.x * 4;
    const uint32_t rl    = max(1u, (n + waves - 1) / waves);
    uint32_t       cur   = 0xFFFFFFFFu;
    bool           mine  = false;
    int            gs = 0, step = 0, stride = 0;
    uint32_t       ok = 0;
    uint32_t       gl[2][STEP];
    for (uint64_t run = (uint64_t)blockIdx.x * 4 + wid; run * rl < n; run += waves)
    {
        const uint32_t i1 = (uint32_t)min((uint64_t)n, run * rl + rl);
        for (uint32_t idx = (uint32_t)(run * rl); idx < i1; ++idx)
        {
            const FtRankItem it = items[idx];
            if (it.group != cur)
            {
                cur               = it.group;
                const FtGroup g   = groups[cur];
                gs                = (int)g.gs;
                step              = (gs + 127) / 128;
                stride            = (int)g.stride;
                mine              = step <= STEP;
                ok                = 0;
                if (mine)
                {
#pragma unroll
                    for (int h = 0; h < 2; ++h)
#pragma unroll
                        for (int j = 0; j < STEP; ++j)
                        {
                            const int  i = (2 * lane + h) * step + j;
                            const bool v = j < step && i < gs;
                            ok |= (uint32_t)v << (h * STEP + j);
                            gl[h][j] = genes[g.gene_off + (v ? i : 0)];
                        }
                }
            }
            if (!mine)
                continue;
            for (int w = lane; w < dints; w += 64)
                dw[w] = 0;
            bins4[lane] = make_int4(0, 0, 0, 0);
            __builtin_amdgcn_wave_barrier();
            const int beg = qptr[it.src_row], end = qptr[it.src_row + 1];
            for (int t = beg + lane; t < end; t += 64)
                dense[qind[t]] = qdata[t];
            __builtin_amdgcn_wave_barrier();
            uint32_t code[2][STEP];
            int      zeros = 0;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int j = 0; j < STEP; ++j)
                {
                    code[h][j]   = dense[gl[h][j]];
                    const bool v = (ok >> (h * STEP + j)) & 1u;
                    zeros += __builtin_popcountll(__ballot(v && code[h][j] == 0));
                    if (v && code[h][j] != 0)
                        atomicAdd(&bins[code[h][j]], 1);
                }
            __builtin_amdgcn_wave_barrier();
            int4 b = bins4[lane];
            if (lane == 0)
                b.x = zeros;  // no atomic touched bin 0
            const int s  = b.x + b.y + b.z + b.w;
            int       ex = wave_incl_scan(s) - s;
            int4      r;
            r.x = 2 * ex + b.x - gs;
            ex += b.x;
            r.y = 2 * ex + b.y - gs;
            ex += b.y;
            r.z = 2 * ex + b.z - gs;
            ex += b.z;
            r.w         = 2 * ex + b.w - gs;
            bins4[lane] = r;
            __builtin_amdgcn_wave_barrier();
            float q[2] = { 0.f, 0.f };
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int j = 0; j < STEP; ++j)
                {
                    if (j >= step)
                        break;
                    // a slot past G_sub stores the padding zero
                    const int y = (ok >> (h * STEP + j)) & 1u ? bins[code[h][j]] : 0;
                    row[(2 * lane + h) * step + j] = (int16_t)y;
                    const float yf = (float)y;
                    q[h]           = __builtin_fmaf(yf, yf, q[h]);
                }
            __builtin_amdgcn_wave_barrier();
            int16_t* dst = out + it.dst;
            for (int i = lane; i < stride; i += 64)
                dst[i] = row[i];
            const float p = wave_tree_sum(q[0] + q[1]);
            if (lane == 0)
                out_sq[idx] = p;
            __builtin_amdgcn_wave_barrier();  // LDS reads done before the next item
        }
    }
}

// Exact tiles (16 cells x 256 rows, INT8 limbs on MFMA).
__global__ __launch_bounds__(256) void k_ftd_mfma(
    const FtdCounts* __restrict__ cnt, const FtCorrTile* __restrict__ tiles,
    const FtGroup* __restrict__ groups, const uint8_t* __restrict__ gflag,
    const int16_t* __restrict__ cell_rank, const float* __restrict__ cell_sq,
    const int16_t* __restrict__ ref_rank, const float* __restrict__ ref_sq,
    float* __restrict__ scores)
{
    __shared__ __attribute__((aligned(16))) int8_t sa[2][16 * FT_MFMA_LD];
    __shared__ __attribute__((aligned(16))) int8_t sb[2][256 * FT_MFMA_LD];
    const uint32_t n = cnt->n_tiles;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x)
    {
        const FtCorrTile tl = tiles[t];
        if (!(gflag[tl.group] & FTD_EXACT) || tl.lim != 0)
            continue;  // uniform per block: the whole tile is skipped
        __syncthreads();
        ft_mfma_tile(tl, groups, cell_rank, cell_sq, ref_rank, ref_sq, scores, sa, sb);
    }
}

// Exact tiles of FTD_BIG cells x 256 rows (groups of at least t64 cells).
__global__ __launch_bounds__(256, 2) void k_ftd_mfma64(
    const FtdCounts* __restrict__ cnt, const FtCorrTile* __restrict__ tiles,
    const FtGroup* __restrict__ groups, const uint8_t* __restrict__ gflag,
    const int16_t* __restrict__ cell_rank, const float* __restrict__ cell_sq,
    const int16_t* __restrict__ ref_rank, const float* __restrict__ ref_sq,
    float* __restrict__ scores)
{
    __shared__ __attribute__((aligned(16))) int8_t sa[2][FTD_BIG * FT_MFMA_LD];
    __shared__ __attribute__((aligned(16))) int8_t sb[2][256 * FT_MFMA_LD];
    const uint32_t n = cnt->n_tiles;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x)
    {
        const FtCorrTile tl = tiles[t];
        if (!(gflag[tl.group] & FTD_EXACT) || tl.lim != (uint32_t)FTD_BIG)
            continue;
        __syncthreads();
        ft_mfma_tile_t<FTD_BIG>(tl, groups, cell_rank, cell_sq, ref_rank, ref_sq, scores, sa, sb);
    }
}

// FP32 wide tiles (FT_TC x FT_TR) of the FP32-form groups.
__global__ __launch_bounds__(FT_NT) void k_ftd_corr(
    const FtdCounts* __restrict__ cnt, const FtCorrTile* __restrict__ tiles,
    const FtGroup* __restrict__ groups, const uint8_t* __restrict__ gflag,
    const int16_t* __restrict__ cell_rank, const float* __restrict__ cell_sq,
    const int16_t* __restrict__ ref_rank, const float* __restrict__ ref_sq,
    float* __restrict__ scores, int kt_max)
{
    extern __shared__ int srow32[];
    const uint32_t n = cnt->n_tiles;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x)
    {
        const FtCorrTile tl = tiles[t];
        if (gflag[tl.group] & FTD_EXACT)
            continue;
        __syncthreads();
        ft_corr_tile<FT_TC, FT_TR, FT_RC, FT_RR>(tl, groups, cell_rank, cell_sq, ref_rank, ref_sq,
                                                 scores, srow32, kt_max);
    }
}

// FP32 thin tiles; `redo` selects the redo tile list and its count.
__global__ __launch_bounds__(256) void k_ftd_thin(
    const FtdCounts* __restrict__ cnt, int redo, const FtCorrTile* __restrict__ tiles,
    const FtGroup* __restrict__ groups, const int16_t* __restrict__ cell_rank,
    const float* __restrict__ cell_sq, const int16_t* __restrict__ ref_rank,
    const float* __restrict__ ref_sq, float* __restrict__ scores, int kt_max)
{
    extern __shared__ int16_t cthin[];
    const uint32_t n = redo ? cnt->n_rthin : cnt->n_thin;
    for (uint32_t t = blockIdx.x; t < n; t += gridDim.x)
    {
        __syncthreads();
        ft_corr_thin_tile<FT_THIN>(tiles[t], groups, cell_rank, cell_sq, ref_rank, ref_sq, scores,
                                   cthin, kt_max);
    }
}

// Segments of the device-resident select: the FtSegments rule with the cell
// of each segment looked up instead of searched.
struct FtdSegments
{
    const FtCell*   cells;
    const uint32_t* segcell;  // cell item of every segment
    const FtGroup*  groups;
    const uint32_t* labels;
    const uint32_t* label_pos;
    const uint32_t* llen;
    __device__ Segment get(uint64_t s) const
    {
        const FtCell   fc = cells[segcell[s]];
        const FtGroup  g  = groups[fc.group];
        const uint32_t j  = (uint32_t)(s - fc.seg0);
        const uint32_t l  = labels[g.label_off + j];
        return Segment{ g.score0 + (uint64_t)fc.pos * g.nref + label_pos[g.label_off + j], llen[l],
                        l };
    }
};

// The redo select's items are the global segments listed in `list`.
struct FtdListSegments
{
    FtdSegments     inner;
    const uint64_t* list;
    __device__ Segment get(uint64_t i) const { return inner.get(list[i]); }
};

// One wave per segment, waves pulling segments until the count is reached.
// redo = 0: all segments (count n_segs, NaN list counted in n_nan);
// redo = 1: the listed segments (n_redo_segs, n_nan2).
template <class Src>
__global__ __launch_bounds__(256, 8) void k_ftd_select(FtdCounts* __restrict__ cnt, int redo, Src src,
                                                    const float* __restrict__ scores,
                                                    const PctParam* __restrict__ params, float p,
                                                    float* __restrict__ out,
                                                    uint64_t* __restrict__ nan_list)
{
    __shared__ WaveSelectSmem sm[256 / kWave];
    const int      wid  = threadIdx.x / kWave;
    const int      lane = threadIdx.x & (kWave - 1);
    const uint64_t n    = redo ? cnt->n_redo_segs : cnt->n_segs;
    unsigned int*  nanc = redo ? &cnt->n_nan2 : &cnt->n_nan;
    for (uint64_t item = (uint64_t)blockIdx.x * (256 / kWave) + wid; item < n;
         item += (uint64_t)gridDim.x * (256 / kWave))
        select_item(src, item, scores, params, p, out, nan_list, nanc, sm[wid], lane);
}

// NaN-carrying segments: the libstdc++ sort emulation, in place in the
// score buffer (the segment is not read again this round).
// Exact-form segments of the first select keep their NaN (gflag given): the
// certified filter sends their cells to the FP32 redo.
template <class Src>
__global__ void k_ftd_select_nan(FtdCounts* __restrict__ cnt, int redo, Src src,
                                 const uint64_t* __restrict__ nan_list, float* __restrict__ scores,
                                 const PctParam* __restrict__ params, float p,
                                 float* __restrict__ out, const FtCell* __restrict__ cells,
                                 const uint32_t* __restrict__ segcell,
                                 const uint8_t* __restrict__ gflag)
{
    const uint32_t n = redo ? cnt->n_nan2 : cnt->n_nan;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint64_t item = nan_list[i];
        if (!redo && (gflag[cells[segcell[item]].group] & FTD_EXACT))
            continue;
        const Segment  sg   = src.get(item);
        const PctParam pp   = params[sg.param];
        float*         a    = scores + sg.off;
        if (!std_sort(a, (long)sg.n))
        {
            atomicOr(&cnt->undefined, 1u);
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

// A cell's candidate list and its quantiles in list order, staged in LDS
// (lists of up to FTD_FAST labels).  The list holds exactly the group's label
// set, so label l's position in the group's ascending label list is the
// number of listed labels below l (a popcount of the list's bitset).
constexpr int FTD_FAST = 64, FTD_FILTER_NT = 128;

__device__ inline void ftd_stage_cell(const uint8_t* __restrict__ gin, uint32_t k,
                                      const float* __restrict__ P, uint64_t seg0, uint8_t* sin,
                                      float* ssc)
{
    uint64_t m[4] = { 0ull, 0ull, 0ull, 0ull };
    for (uint32_t t = 0; t < k; ++t)
    {
        const uint32_t l = gin[t];
        sin[t]           = (uint8_t)l;
        m[l >> 6] |= 1ull << (l & 63);
    }
    for (uint32_t t = 0; t < k; ++t)
    {
        const uint32_t l = sin[t], w = l >> 6;
        uint32_t       j = __builtin_popcountll(m[w] & ((1ull << (l & 63)) - 1ull));
        for (uint32_t v = 0; v < w; ++v)
            j += __builtin_popcountll(m[v]);
        ssc[t] = P[seg0 + j];
    }
}

// Filter of every cell with work, candidate lists updated in place (the new
// list is formed in `stage` first).  Exact-form cells use the certified
// filter; an undecided one is listed for the FP32 redo, with its tiles and
// segments.
__global__ __launch_bounds__(FTD_FILTER_NT) void k_ftd_filter(
    FtdCounts* __restrict__ cnt, const FtCell* __restrict__ cells,
    const FtGroup* __restrict__ groups, const uint8_t* __restrict__ gflag,
    const uint32_t* __restrict__ labels, const float* __restrict__ P, uint8_t* __restrict__ lists,
    uint32_t* __restrict__ counts, uint32_t L, float thre, uint8_t* __restrict__ stage,
    uint32_t* __restrict__ redo, FtCorrTile* __restrict__ rthin, uint64_t* __restrict__ rseg)
{
    __shared__ uint8_t sin_[FTD_FILTER_NT][FTD_FAST];
    __shared__ float   ssc_[FTD_FILTER_NT][FTD_FAST];
    const uint32_t     n = cnt->n_cells;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const FtCell   fc  = cells[i];
        const FtGroup  G   = groups[fc.group];
        uint8_t*       gin = lists + (size_t)fc.cell * L;
        uint8_t*       out = stage + (size_t)i * L;
        const uint32_t k   = counts[fc.cell];
        const bool     ex  = gflag[fc.group] & FTD_EXACT;
        uint32_t       m   = 0;
        int            st  = 0;
        if (k <= (uint32_t)FTD_FAST)
        {
            uint8_t* sin = sin_[threadIdx.x];
            float*   ssc = ssc_[threadIdx.x];
            ftd_stage_cell(gin, k, P, fc.seg0, sin, ssc);
            auto score_of = [&](uint32_t t) { return ssc[t]; };
            if (ex)
                st = ft_filter_exact_list(score_of, G.gs, sin, k, thre, out, &m);
            else
                m = ft_filter_list(score_of, sin, k, thre, out);
        }
        else if (ex)
            st = ft_filter_exact_cell(fc, G, labels + G.label_off, P, gin, k, thre, out, &m);
        else
            m = ft_filter_cell(fc, G, labels + G.label_off, P, gin, k, thre, out);
        if (st)
        {
            uint32_t t = atomicAdd(&cnt->n_rthin, (G.nref + 255) / 256);
            for (uint32_t r0 = 0; r0 < G.nref; r0 += 256)
                rthin[t++] = FtCorrTile{ fc.group, fc.pos, r0, 1u };
            const uint64_t s0 = atomicAdd((unsigned long long*)&cnt->n_redo_segs,
                                          (unsigned long long)G.nlabel);
            for (uint32_t j = 0; j < G.nlabel; ++j)
                rseg[s0 + j] = fc.seg0 + j;
            redo[atomicAdd(&cnt->n_redo, 1u)] = i;
            atomicAdd(&cnt->tot_redo, 1u);
            continue;
        }
        for (uint32_t t = 0; t < m; ++t)
            gin[t] = out[t];
        counts[fc.cell] = m;
    }
}

// Quantiles of the redo segments back into the per-segment array.
__global__ void k_ftd_redo_scatter(const FtdCounts* __restrict__ cnt,
                                   const uint64_t* __restrict__ rseg,
                                   const float* __restrict__ pred, float* __restrict__ P)
{
    const uint64_t n = cnt->n_redo_segs;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        P[rseg[i]] = pred[i];
}

// FP32 filter of the redo cells (their certified filter left the lists
// untouched).
__global__ void k_ftd_filter_redo(const FtdCounts* __restrict__ cnt,
                                  const uint32_t* __restrict__ redo,
                                  const FtCell* __restrict__ cells,
                                  const FtGroup* __restrict__ groups,
                                  const uint32_t* __restrict__ labels, const float* __restrict__ P,
                                  uint8_t* __restrict__ lists, uint32_t* __restrict__ counts,
                                  uint32_t L, float thre, uint8_t* __restrict__ stage)
{
    const uint32_t n = cnt->n_redo;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
    {
        const uint32_t i  = redo[r];
        const FtCell   fc = cells[i];
        const FtGroup  G  = groups[fc.group];
        uint8_t*       in = lists + (size_t)fc.cell * L;
        uint8_t*       out = stage + (size_t)i * L;
        const uint32_t m  = ft_filter_cell(fc, G, labels + G.label_off, P, in, counts[fc.cell],
                                           thre, out);
        for (uint32_t t = 0; t < m; ++t)
            in[t] = out[t];
        counts[fc.cell] = m;
    }
}

// Every active cell counts the round; the next round takes the cells with
// several candidates and rounds left.
__global__ void k_ftd_compact(FtdCounts* __restrict__ cnt, const uint32_t* __restrict__ act,
                              const uint32_t* __restrict__ counts, int32_t* __restrict__ rounds,
                              int times, uint32_t* __restrict__ act_next)
{
    if (cnt->halt)
        return;
    const uint32_t n = cnt->n_active;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint32_t c = act[i];
        const int32_t  r = rounds[c] + 1;
        rounds[c]        = r;
        if (counts[c] > 1 && (times == 0 || r < times))
            act_next[atomicAdd(&cnt->n_next, 1u)] = c;
    }
}

// A halted chain's round is resumed from its part A by the host.
__global__ void k_ftd_resume(FtdCounts* __restrict__ cnt)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        cnt->halt = 0;
}

// Counter reset for the autotuner's repeated select launches.
__global__ void k_ftd_reset_nan(FtdCounts* __restrict__ cnt)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        cnt->n_nan = 0;
}

__global__ void k_ftd_final(const uint8_t* __restrict__ lists, const uint32_t* __restrict__ counts,
                            uint32_t Nq, uint32_t L, uint32_t* __restrict__ final_labels)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < Nq)
        final_labels[c] = counts[c] == 0 ? 0u : (uint32_t)lists[(size_t)c * L];
}

}  // namespace csg
