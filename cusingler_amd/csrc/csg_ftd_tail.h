// The last fine-tune rounds as one persistent kernel (cufinetune /
// finetune_round, src/cusingler.cu:1138-1333).
//
// Once a round has few active cells, the per-round launches of the
// device-resident schedule (grouping, set-up, ranks, correlations, select,
// filter, compaction: about a dozen kernels and a counter read-back per
// round) cost far more than the round's work.  Here one cooperative launch
// runs every remaining round of those cells, each cell a group of its own
// (the reference refines one cell at a time; sharing a candidate set across
// cells only saves work, it never changes a value):
//   * begin (the cell's owner block): G_sub of its candidate list (the union
//     of every ordered pair's first min(len, k_t) window genes), the cell's
//     doubled centred ranks over G_sub with their chunk-tree sum of squares,
//     and its reference rows split into work chunks;
//   * grid barrier;
//   * rows (every wave): a chunk's reference rows ranked over G_sub and
//     correlated with the cell in the reference's 128-chunk + stride-tree
//     FP32 order -- the fused form's arithmetic, so the same bits;
//   * grid barrier;
//   * end (the owner block): each candidate label's quantile (the wave
//     select, the libstdc++ sort emulation for NaN-carrying segments), the
//     filter (first min, last max, front insertion of the maximum), the
//     round counted; a cell still undecided begins its next round at once.
// Two grid barriers per round (the launch is cooperative, so every block is
// resident); all blocks leave the loop together when no cell begins a round.  Blocks only communicate through global memory across
// barriers; within a phase a block reads back nothing it wrote to global
// memory (its own results are kept in LDS).
#pragma once

#include "csg_ftd.h"

namespace csg {

constexpr int FTT_NT = 512;  // threads per block (8 waves)
constexpr int FTT_NW = FTT_NT / 64;
constexpr int FTT_CH = 64;   // reference rows per work chunk (one row per wave item)

// One cell of the tail (a slot for the whole launch).
struct FtdTailSlot
{
    uint32_t cell;
    uint32_t gs;     // |G_sub| of the slot's current round
    uint32_t nrows;  // reference rows of the current round's candidate labels
    uint32_t pad_;
    uint64_t row0;   // first score of the slot's region (its rows at the start)
    float    csq;    // the cell's chunk-tree sum of squares over G_sub
    uint32_t pad2_;
};

struct FtdTailChunk
{
    uint32_t slot, r0;  // rows r0 .. r0 + FTT_CH - 1 of the slot's round
};
// (the row phase's items are the chunks' rows: item i is row i % FTT_CH of
// chunk i / FTT_CH, skipped past the slot's row count)

struct FtdTailState
{
    uint32_t nchunk[2];  // work chunks of the rounds of parity p
    uint32_t nact[2];    // cells taking part in the round of parity p
    uint64_t rows;       // score rows reserved by the slots
    uint32_t rounds;     // rounds with at least one cell
    uint32_t bar0;       // the first barrier: blocks arrived
    uint32_t bar;        // later barriers: XCD groups arrived, over all barriers
    uint32_t timeout;    // a barrier wait gave up (the results are void)
    uint32_t pad_;
    uint32_t xn[16];     // blocks per XCD (counted at the start)
    uint32_t xarr[16];   // blocks of each XCD arrived, over all later barriers
    uint64_t tot_active, tot_scores, pair_genes;
    // device clock (s_memrealtime, 100 MHz) when block 0 left barrier g, for
    // g < FTT_STAMPS: per-phase durations for the host's phase report
    uint64_t stamp[64];
};
constexpr int FTT_STAMPS = 64;

struct FtdTailArgs
{
    FtdCounts*       cnt;  // `undefined` is raised here
    FtdTailState*    st;   // zeroed by the host
    FtdTailSlot*     slots;
    FtdTailChunk*    chunks;
    const uint32_t*  act;  // the first round's active cells
    uint32_t         n;
    uint8_t*         lists;
    uint32_t*        counts;
    int32_t*         rounds;
    uint32_t         L;
    int              times;
    const uint32_t*  ctdidx;
    const uint32_t*  ctdiff;
    const uint32_t*  cut_tab;
    const uint32_t*  lstart;
    const uint32_t*  llen;
    uint32_t         Gf;
    uint32_t         gst;          // row stride of genes / crank (Gf rounded up to 8)
    int              nbins;
    int              wave_ints;    // per-wave LDS of the row phase (bins, then codes)
    const uint16_t*  qdata;
    const int*       qind;
    const int*       qptr;
    uint32_t*        genes;   // [n x gst] each slot's G_sub, ascending
    int16_t*         crank;   // [n x gst] each slot's cell ranks over G_sub
    float*           scores;  // the slots' score regions
    float*           P;       // [n x L] quantiles in list order
    uint64_t*        nanl;    // [n x L] NaN-carrying list positions
    const PctParam*  pct;
    float            quantile;
    float            thre;
};

// Per-block LDS of the tail kernel: the largest of the begin phase's, the
// end phase's and the row phase's per-wave tables.
__host__ __device__ inline size_t ftd_tail_lds(uint32_t Gf, int nbins, int wave_ints)
{
    const size_t gw    = ((size_t)Gf + 3) / 4;  // Gf bytes as words
    const size_t nb    = ((size_t)nbins + 3) & ~(size_t)3;
    // begin: gene flags (bytes), gene list (u16), dense query row (u16),
    // codes (u16), bins, then the staged pair windows
    const size_t begin = 4 * (gw + 3 * ((Gf + 1) / 2 + 1) + nb + 2 * FTD_PAIRS) + 64;
    // end: list offsets, the new list, a wave select table per wave
    const size_t endp  = 4 * 256 + 256 + FTT_NW * sizeof(WaveSelectSmem) + 64;
    const size_t rows  = (size_t)FTT_NW * wave_ints * 4;
    return std::max(begin, std::max(endp, rows));
}

// The XCD (accelerator complex die) a wave runs on: blocks of one XCD share
// its L2, the device's coherence point for them.
__device__ inline uint32_t ftd_xcc_id()
{
    // hardware register XCC_ID (20), bits 3:0
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;
}

// Bounded wait until *ctr >= target (about a second at most): false after a
// timeout, which is recorded for every block to see.
__device__ inline bool ftd_tail_wait(FtdTailState* st, const uint32_t* ctr, uint32_t target)
{
    for (uint32_t spin = 0;; ++spin)
    {
        if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target)
            return true;
        if (spin > (1u << 24))
        {
            __hip_atomic_fetch_or(&st->timeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Grid barriers of the cooperative launch (which makes every block
// co-resident).  A block's writes reach its XCD's L2 once its waves have
// waited for their stores; an agent-scope release (an L2 write-back) then
// publishes everything that L2 holds, and an acquire (an L1/L2
// invalidation) takes in what the other XCDs published.
//   * The first barrier: every block releases, arrives on `bar0` and waits
//     for all (it also counts the blocks of each XCD, `xn`).
//   * Later barriers (`gen` = 0, 1, ...): a block waits for its stores and
//     arrives on its XCD's counter; the last block of an XCD to arrive
//     releases that XCD's L2 once for all of them and arrives on `bar`; every
//     block waits for all XCDs and acquires.  One write-back per XCD per
//     barrier instead of one per block (each write-back walks the L2).
// A wait that does not end within about a second sets `timeout`; every block
// that sees it returns, so all of them end and the host reports the launch
// as failed.  Both return false (in every thread) after a timeout.
__device__ inline bool ftd_tail_sync_first(FtdTailState* st)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&st->bar0, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ok = ftd_tail_wait(st, &st->bar0, gridDim.x);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = s_ok && __hip_atomic_load(&st->timeout, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) == 0;
        if (blockIdx.x == 0)
            st->stamp[0] = __builtin_amdgcn_s_memrealtime();
    }
    __syncthreads();
    return s_ok != 0;
}

// `x` / `gx`: this block's XCD and its number of blocks; `nx`: XCDs with
// blocks.  Barrier `gen` stamps stamp[gen + 1] (block 0).
__device__ inline bool ftd_tail_sync(FtdTailState* st, uint32_t gen, uint32_t x, uint32_t gx,
                                     uint32_t nx)
{
    __shared__ int s_ok;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are in the L2
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const uint32_t prev =
            __hip_atomic_fetch_add(&st->xarr[x], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prev + 1 == (gen + 1) * gx)
        {
            // the XCD's last block: every block of it has arrived
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(&st->bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_ok = ftd_tail_wait(st, &st->bar, (gen + 1) * nx);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = s_ok && __hip_atomic_load(&st->timeout, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) == 0;
        if (blockIdx.x == 0 && gen + 1 < (uint32_t)FTT_STAMPS)
            st->stamp[gen + 1] = __builtin_amdgcn_s_memrealtime();
    }
    __syncthreads();
    return s_ok != 0;
}

// Begin a round of slot s (whole block): G_sub, the cell ranks, the chunks.
// `lst` / `k`: the cell's current list (in LDS).  Returns through the slot
// record and the global tables; a cell whose G_sub has fewer than 20 genes
// ends here with {front} (src/cusingler.cu:1158-1159).
__device__ inline void ftd_tail_begin(const FtdTailArgs& a, uint32_t s, const uint8_t* lst,
                                      uint32_t k, int p, bool count, uint8_t* smem)
{
    FtdTailState* st   = a.st;
    FtdTailSlot&  sl   = a.slots[s];
    const uint32_t c   = sl.cell;
    const int     tid  = threadIdx.x;
    const int     lane = tid & 63;
    const uint32_t Gf  = a.Gf;
    const size_t  gw   = ((size_t)Gf + 3) / 4;
    const size_t  hw   = (Gf + 1) / 2 + 1;
    uint8_t*      flag  = smem;                                // Gf bytes
    uint16_t*     glist = (uint16_t*)(smem + 4 * gw);          // G_sub, ascending
    uint16_t*     dense = glist + 2 * hw;                      // the cell's codes over Gf
    uint16_t*     codes = dense + 2 * hw;                      // codes at G_sub
    int*          bins  = (int*)(codes + 2 * hw);              // nbins
    uint2*        pw    = (uint2*)(((uintptr_t)(bins + (((size_t)a.nbins + 3) & ~(size_t)3)) + 7) &
                              ~(uintptr_t)7);  // FTD_PAIRS windows, 8-byte aligned
    __shared__ int s_tmp[FTT_NW];
    __shared__ int s_gs;
    for (size_t w = tid; w < gw; w += FTT_NT)
        ((uint32_t*)flag)[w] = 0u;
    // the ordered pairs' windows (k > 1 for a cell taking part in a round)
    const uint32_t cut = a.cut_tab[k];
    ftd_flag_windows<FTT_NT>(lst, k, a.L, cut, a.ctdidx, a.ctdiff, flag, pw);
    __syncthreads();
    // ascending gene list: contiguous gene ranges per thread, block scan
    const uint32_t per = (Gf + FTT_NT - 1) / FTT_NT;
    const uint32_t g0 = min((uint32_t)tid * per, Gf), g1 = min(g0 + per, Gf);
    int            own = 0;
    for (uint32_t g = g0; g < g1; ++g)
        own += flag[g];
    int       total;
    int       at  = block_excl_scan<FTT_NT>(own, s_tmp, &total);
    uint32_t* dst = a.genes + (size_t)s * a.gst;
    for (uint32_t g = g0; g < g1; ++g)
        if (flag[g])
        {
            glist[at] = (uint16_t)g;
            dst[at]   = g;
            ++at;
        }
    const int gs = total;
    if (tid == 0)
        s_gs = gs;
    __syncthreads();
    if (gs < 20)
    {
        // the round returns {front}: the list keeps its first label
        if (tid == 0)
        {
            a.counts[c] = 1;
            a.rounds[c] += 1;
            sl.gs    = 0;
            sl.nrows = 0;
        }
        return;
    }
    // the cell's ranks over G_sub by wave 0: its CSR row densified in LDS,
    // the codes gathered, the code histogram (ballots for code 0), the rank
    // table by a wave scan; then its chunk-tree sum of squares
    if (tid < 64)
    {
        for (uint32_t w = lane; w < hw; w += 64)
            ((uint32_t*)dense)[w] = 0u;
        for (int i = lane; i < a.nbins; i += 64)
            bins[i] = 0;
        __builtin_amdgcn_wave_barrier();
        const int beg = a.qptr[c], end = a.qptr[c + 1];
        for (int t = beg + lane; t < end; t += 64)
            dense[a.qind[t]] = a.qdata[t];
        __builtin_amdgcn_wave_barrier();
        for (int i = lane; i < gs; i += 64)
            codes[i] = dense[glist[i]];
        __builtin_amdgcn_wave_barrier();
        int zeros = 0;
        for (int i0 = 0; i0 < gs; i0 += 64)
        {
            const int      i  = i0 + lane;
            const uint16_t cc = i < gs ? codes[i] : (uint16_t)0;
            zeros += __builtin_popcountll(__ballot(i < gs && cc == 0));
            if (i < gs && cc != 0)
                atomicAdd(&bins[cc], 1);
        }
        ftd_bins_to_ranks(bins, a.nbins, gs, zeros, lane);
        int16_t* xr = a.crank + (size_t)s * a.gst;
        for (int i = lane; i < gs; i += 64)
            xr[i] = (int16_t)bins[codes[i]];
        const int step = (gs + 127) / 128;
        float     q[2] = { 0.f, 0.f };
#pragma unroll
        for (int h = 0; h < 2; ++h)
        {
            const int lo = (2 * lane + h) * step, hi = min(lo + step, gs);
            for (int i = lo; i < hi; ++i)
            {
                const float v = (float)bins[codes[i]];
                q[h]          = __builtin_fmaf(v, v, q[h]);
            }
        }
        const float csq = wave_tree_sum(q[0] + q[1]);
        // the candidate labels' rows, split into chunks
        uint32_t nrows = 0;
        for (uint32_t j = lane; j < k; j += 64)
            nrows += a.llen[lst[j]];
        nrows = (uint32_t)wave_sum((int)nrows);
        const uint32_t nch = (nrows + FTT_CH - 1) / FTT_CH;
        uint32_t       b0  = 0;
        if (lane == 0)
        {
            sl.gs    = (uint32_t)gs;
            sl.nrows = nrows;
            sl.csq   = csq;
            b0       = atomicAdd(&st->nchunk[p], nch);
            if (count)
            {
                atomicAdd((unsigned long long*)&st->tot_scores, (unsigned long long)nrows);
                atomicAdd((unsigned long long*)&st->pair_genes, (unsigned long long)nrows * gs);
            }
        }
        b0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)b0);
        for (uint32_t i = lane; i < nch; i += 64)
            a.chunks[b0 + i] = FtdTailChunk{ s, i * FTT_CH };
    }
}

// The row phase: every wave takes chunks; per reference row the rank table
// over G_sub (ftd_row_table) and the fused form's chunk sums with the cell's
// ranks, completed by wave_tree_sum: the scores of k_ftd_fused.
template <class CT>
__device__ inline void ftd_tail_rows(const FtdTailArgs& a, const CT* __restrict__ ref_codes,
                                     int p, int* smem)
{
    const int      wid   = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int*           bins  = smem + wid * a.wave_ints;
    uint16_t*      codes = (uint16_t*)(bins + ((a.nbins + 3) & ~3));
    const uint32_t nit   = a.st->nchunk[p] * (uint32_t)FTT_CH;
    const uint32_t waves = gridDim.x * FTT_NW;
    for (uint32_t it = blockIdx.x * FTT_NW + wid; it < nit; it += waves)
    {
        const FtdTailChunk ch = a.chunks[it / FTT_CH];
        const FtdTailSlot  sl = a.slots[ch.slot];
        const uint32_t     r  = ch.r0 + it % FTT_CH;
        if (r >= sl.nrows)
            continue;  // (uniform per wave)
        const uint8_t*  lst  = a.lists + (size_t)sl.cell * a.L;
        const int       gs   = (int)sl.gs;
        const uint32_t* gl   = a.genes + (size_t)ch.slot * a.gst;
        const int16_t*  xc   = a.crank + (size_t)ch.slot * a.gst;
        const int       step = (gs + 127) / 128;
        // the row's list position and its row within the label
        uint32_t j = 0, cum = 0, l = lst[0], ln = a.llen[l];
        while (cum + ln <= r)
        {
            cum += ln;
            l  = lst[++j];
            ln = a.llen[l];
        }
        {
            const uint32_t row = a.lstart[l] + (r - cum);
            ftd_row_table(ref_codes + (size_t)row * a.Gf, gl, gs, a.nbins, bins, codes, lane);
            float q[2] = { 0.f, 0.f }, d[2] = { 0.f, 0.f };
#pragma unroll
            for (int h = 0; h < 2; ++h)
            {
                const int lo = (2 * lane + h) * step, hi = min(lo + step, gs);
                for (int i = lo; i < hi; ++i)
                {
                    const float y = (float)bins[codes[i]];
                    q[h]          = __builtin_fmaf(y, y, q[h]);
                    d[h]          = __builtin_fmaf((float)xc[i], y, d[h]);
                }
            }
            const float syy = wave_tree_sum(q[0] + q[1]);
            const float sxy = wave_tree_sum(d[0] + d[1]);
            if (lane == 0)
                a.scores[sl.row0 + r] = corr_from_sums(sxy, sl.csq, syy);
            __builtin_amdgcn_wave_barrier();  // LDS reads done before the next row
        }
    }
}

// Segments of one slot's round: list position j of the slot.
struct FtdTailSegments
{
    const uint8_t*  lst;
    const uint32_t* loff;  // row offset of each list position (LDS)
    const uint32_t* llen;
    uint64_t        row0;
    __device__ Segment get(uint64_t j) const
    {
        const uint32_t l = lst[j];
        return Segment{ row0 + loff[j], llen[l], l };
    }
};

// End the round of slot s (whole block): quantiles, filter, round count;
// returns (to every thread) whether the cell takes part in the next round,
// with its new list in `lst` / *k.
__device__ inline bool ftd_tail_end(const FtdTailArgs& a, uint32_t s, uint8_t* lst, uint32_t* kp,
                                    uint8_t* smem)
{
    const FtdTailSlot sl   = a.slots[s];
    const uint32_t    c    = sl.cell;
    const int         tid  = threadIdx.x;
    const int         wid  = tid >> 6, lane = tid & 63;
    const uint32_t    k    = *kp;
    uint32_t*         loff = (uint32_t*)smem;          // 64 words (k <= L <= 256: 256)
    uint8_t*          nout = (uint8_t*)(loff + 256);   // 256 bytes
    WaveSelectSmem*   wsm  = (WaveSelectSmem*)(nout + 256);
    __shared__ unsigned int s_nan;
    __shared__ int          s_next;
    __shared__ uint32_t     s_m;
    __shared__ StdSortRange s_stack[kStdSortStack];  // (thread 0's sort: no scratch)
    if (tid == 0)
    {
        uint32_t off = 0;
        for (uint32_t t = 0; t < k; ++t)
        {
            loff[t] = off;
            off += a.llen[lst[t]];
        }
        s_nan = 0;
    }
    __syncthreads();
    const FtdTailSegments src{ lst, loff, a.llen, sl.row0 };
    float*                 P    = a.P + (size_t)s * a.L;
    uint64_t*              nanl = a.nanl + (size_t)s * a.L;
    for (uint32_t j = wid; j < k; j += FTT_NW)
        select_item(src, j, a.scores, a.pct, a.quantile, P, nanl, &s_nan, wsm[wid], lane);
    __syncthreads();
    if (tid == 0)
    {
        // NaN-carrying segments: the libstdc++ sort emulation in place
        for (unsigned int t = 0; t < s_nan; ++t)
        {
            const uint64_t j  = nanl[t];
            const Segment  sg = src.get(j);
            const PctParam pp = a.pct[sg.param];
            float*         v  = a.scores + sg.off;
            if (!std_sort(v, (long)sg.n, s_stack))
            {
                atomicOr(&a.cnt->undefined, 1u);
                P[j] = __builtin_nanf("");
                continue;
            }
            P[j] = pp.kind == 1 ? v[0]
                 : pp.kind == 2 ? v[sg.n - 1]
                                : pct_interp(a.quantile, pp.prev_index, pp.step, v[pp.k - 1],
                                             v[pp.k]);
        }
        auto           score_of = [&](uint32_t t) { return P[t]; };
        const uint32_t m        = ft_filter_list(score_of, lst, k, a.thre, nout);
        uint8_t*       gl       = a.lists + (size_t)c * a.L;
        for (uint32_t t = 0; t < m; ++t)
        {
            lst[t] = nout[t];
            gl[t]  = nout[t];
        }
        a.counts[c]     = m;
        const int32_t r = a.rounds[c] + 1;
        a.rounds[c]     = r;
        s_m             = m;
        s_next          = m > 1 && (a.times == 0 || r < a.times);
    }
    __syncthreads();
    *kp = s_m;  // (every thread: the next round's list length)
    return s_next != 0;
}

// One cooperative launch: every round of the given active cells.  Slot s is
// owned by block s % gridDim.x (its begin and end phases).
template <class CT>
__global__ __launch_bounds__(FTT_NT) void k_ftd_tail(FtdTailArgs a, const CT* __restrict__ ref_codes)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t tsm[];
    __shared__ uint8_t  s_list[256];
    __shared__ uint32_t s_k;
    FtdTailState*       st = a.st;
    __shared__ uint32_t s_x, s_gx, s_nx;
    if (threadIdx.x == 0)
    {
        s_x = ftd_xcc_id();
        __hip_atomic_fetch_add(&st->xn[s_x], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // first round: every slot's cell, its score region, its begin
    for (uint32_t s = blockIdx.x; s < a.n; s += gridDim.x)
    {
        if (threadIdx.x == 0)
        {
            const uint32_t c = a.act[s];
            const uint32_t k = a.counts[c];
            uint32_t       nrows = 0;
            for (uint32_t t = 0; t < k; ++t)
            {
                s_list[t] = a.lists[(size_t)c * a.L + t];
                nrows += a.llen[s_list[t]];
            }
            s_k                = k;
            a.slots[s].cell    = c;
            a.slots[s].row0    = atomicAdd((unsigned long long*)&st->rows, (unsigned long long)nrows);
            atomicAdd(&st->nact[0], 1u);
        }
        __syncthreads();
        ftd_tail_begin(a, s, s_list, s_k, 0, false, tsm);
        __syncthreads();
    }
    if (!ftd_tail_sync_first(st))
        return;
    if (threadIdx.x == 0)
    {
        uint32_t nx = 0;
        for (int i = 0; i < 16; ++i)
            nx += st->xn[i] ? 1u : 0u;
        s_gx = st->xn[s_x];
        s_nx = nx;
    }
    __syncthreads();
    const uint32_t x = s_x, gx = s_gx, nx = s_nx;
    uint32_t       gen = 0;
    // (the first round was counted by the part A that handed it over: the
    // totals cover the rounds after it)
    for (int p = 0, round = 0;; p ^= 1, ++round)
    {
        const uint32_t nact = st->nact[p];
        if (nact == 0)
            break;  // (every block reads the same count after the barrier)
        if (blockIdx.x == 0 && threadIdx.x == 0 && round > 0)
        {
            st->rounds += 1;
            st->tot_active += nact;
        }
        ftd_tail_rows<CT>(a, ref_codes, p, (int*)tsm);
        if (!ftd_tail_sync(st, gen++, x, gx, nx))
            return;
        if (blockIdx.x == 0 && threadIdx.x == 0)
        {
            // this parity's counters are read again only after the next
            // barrier, by the round after next
            st->nchunk[p] = 0;
            st->nact[p]   = 0;
        }
        for (uint32_t s = blockIdx.x; s < a.n; s += gridDim.x)
        {
            __syncthreads();
            if (a.slots[s].gs == 0)
                continue;  // not in this round (uniform per block)
            if (threadIdx.x == 0)
            {
                const uint32_t c = a.slots[s].cell;
                const uint32_t k = a.counts[c];
                for (uint32_t t = 0; t < k; ++t)
                    s_list[t] = a.lists[(size_t)c * a.L + t];
                s_k = k;
            }
            __syncthreads();
            uint32_t k = s_k;
            if (ftd_tail_end(a, s, s_list, &k, tsm))
            {
                if (threadIdx.x == 0)
                    atomicAdd(&st->nact[p ^ 1], 1u);
                ftd_tail_begin(a, s, s_list, k, p ^ 1, true, tsm);
            }
            else if (threadIdx.x == 0)
                a.slots[s].gs = 0;
        }
        if (!ftd_tail_sync(st, gen++, x, gx, nx))
            return;
    }
}

}  // namespace csg
