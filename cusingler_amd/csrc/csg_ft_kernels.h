// Device kernels of the fine-tune step (src/cusingler.cu:1138-1333).
//
// The reference refines one query cell at a time: each round it builds the
// gene subset G_sub of the cell's candidate labels, ranks the cell and every
// reference cell of those labels over G_sub, correlates them with a
// 128-chunk + tree FP32 reduction (spearman_reduce, :804-856), takes each
// label's quantile and filters the labels.  Here a round processes all cells
// at once: cells with the same candidate set share G_sub and the reference
// ranks, so the host groups them and these kernels run over flat work tables.
#pragma once

#include "csg_kernels.h"

namespace csg {

// FP32 chunk + tree sum of spearman_reduce: thread t of 128 sums genes
// [t*step, (t+1)*step) of `f(i)` by fma in order, then partial[t] +=
// partial[t + stride] for stride 1, 2, ..., 64 (src/cusingler.cu:823-845).
// Called by all 128 threads of a block; `part` is 128 floats of LDS.
template <class F> __device__ inline float chunk_tree_sum128(int n, F f, float* part)
{
    const int t    = threadIdx.x;
    const int step = (n + 127) / 128;
    float     acc  = 0.f;
    const int lo   = t * step;
    const int hi   = min(lo + step, n);
    for (int i = lo; i < hi; ++i)
    {
        const float v = f(i);
        acc           = __builtin_fmaf(v, v, acc);
    }
    part[t] = acc;
    __syncthreads();
    for (int stride = 1; stride < 128; stride *= 2)
    {
        if ((t % (2 * stride)) == 0)
            part[t] += part[t + stride];
        __syncthreads();
    }
    const float r = part[0];
    __syncthreads();
    return r;
}

// FP32 correlation tiles: TC cells x TR reference rows per block, RC x RR
// pairs per thread; groups of at most FT_THIN cells use the thin tile
// (FT_THIN cells x 256 rows).
constexpr int FT_TC = 32, FT_TR = 32, FT_RC = 2, FT_RR = 2;
constexpr int FT_NT   = (FT_TC / FT_RC) * (FT_TR / FT_RR);
constexpr int FT_THIN = 8;
// LDS stage width (genes) of the thin tiles: FT_THIN cell rows of int16 in
// the 64 KiB a block gets by default
constexpr int FT_THIN_KT = (65536 / (FT_THIN * 2)) & ~7;

// A row to rank over a group's gene subset.
struct FtRankItem
{
    uint32_t src_row;  // reference row (dense codes) or query row (CSR)
    uint32_t group;
    uint64_t dst;      // element offset of the output row in the rank buffer
};

struct FtGroup
{
    uint32_t gene_off;    // into the gene table
    uint32_t gs;          // |G_sub|
    uint32_t stride;      // row stride of the rank buffers (gs rounded up to 8)
    uint32_t nref;        // reference rows of the group's labels
    uint32_t ncell;       // cells of the group
    uint32_t ref_item0;   // first reference rank item (row) of the group
    uint32_t cell_item0;  // first cell item of the group
    uint32_t label_off;   // into the label table (ascending labels)
    uint32_t nlabel;
    uint32_t pad_;
    uint64_t ref_dst0;    // element offset of the first reference rank row
    uint64_t cell_dst0;   // element offset of the first cell rank row
    uint64_t score0;      // first score of the group (cell-major, nref per cell)
};

// Doubled centred ranks of codes[0..gs) (code 0 = implicit zero, counted like
// any other value) into row[], over nbins code bins held in LDS.
template <int NT>
__device__ inline void rank_codes(const uint16_t* codes, int gs, int nbins, int* bins, int16_t* row,
                                  int* tmp)
{
    for (int i = threadIdx.x; i < nbins; i += NT)
        bins[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < gs; i += NT)
        atomicAdd(&bins[codes[i]], 1);
    __syncthreads();
    bins_to_ranks<NT>(bins, nbins, 0, gs, tmp);
    for (int i = threadIdx.x; i < gs; i += NT)
        row[i] = (int16_t)bins[codes[i]];
    __syncthreads();
}

// Group holding batch-wide item `i` given the groups' first items (sorted).
__device__ inline uint32_t find_group(const FtGroup* __restrict__ groups, uint32_t ngroups,
                                      uint32_t i, bool by_ref)
{
    uint32_t lo = 0, hi = ngroups;  // last group with first item <= i
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) / 2;
        const uint32_t f   = by_ref ? groups[mid].ref_item0 : groups[mid].cell_item0;
        if (f <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Reference rows: one block per reference row of every group (the group and
// the row follow from the block index and the group's label list); gather the
// dense fine-tune codes at G_sub (get_device_ref_lines, src/cusingler.cu:
// 566-577), rank them (rankdata_bin3, :745-801) and form the chunk-tree sum of
// squares.
__device__ inline void ft_rank_ref_item(uint32_t item, const FtGroup* __restrict__ groups,
                                        uint32_t ngroups, const uint32_t* __restrict__ labels,
                                        const uint32_t* __restrict__ lstart,
                                        const uint32_t* __restrict__ llen,
                                        const uint32_t* __restrict__ genes,
                                        const uint16_t* __restrict__ ref_codes, uint32_t Gf,
                                        int nbins, int16_t* __restrict__ out,
                                        float* __restrict__ out_sq, int* smem, float* part, int* tmp)
{
    const uint32_t gi      = find_group(groups, ngroups, item, true);
    const FtGroup  g       = groups[gi];
    uint32_t       j       = item - g.ref_item0;
    uint32_t       src_row = 0;
    for (uint32_t t = 0; t < g.nlabel; ++t)
    {
        const uint32_t l = labels[g.label_off + t];
        if (j < llen[l])
        {
            src_row = lstart[l] + j;
            break;
        }
        j -= llen[l];
    }
    const uint32_t  local = item - g.ref_item0;
    int*            bins  = smem;
    uint16_t*       codes = (uint16_t*)(smem + ((nbins + 3) & ~3));
    int16_t*        row   = (int16_t*)(codes + ((g.gs + 7) & ~7u));
    const uint16_t* src   = ref_codes + (size_t)src_row * Gf;
    const uint32_t* gl    = genes + g.gene_off;
    for (uint32_t i = threadIdx.x; i < g.gs; i += 128)
        codes[i] = src[gl[i]];
    __syncthreads();
    rank_codes<128>(codes, (int)g.gs, nbins, bins, row, tmp);
    int16_t* dst = out + g.ref_dst0 + (uint64_t)local * g.stride;
    for (uint32_t i = threadIdx.x; i < g.stride; i += 128)
        dst[i] = i < g.gs ? row[i] : (int16_t)0;
    const float sq = chunk_tree_sum128((int)g.gs, [&](int i) { return (float)row[i]; }, part);
    if (threadIdx.x == 0)
        out_sq[item] = sq;
}

__global__ __launch_bounds__(128) void k_ft_rank_ref(const FtGroup* __restrict__ groups,
                                                     uint32_t ngroups,
                                                     const uint32_t* __restrict__ labels,
                                                     const uint32_t* __restrict__ lstart,
                                                     const uint32_t* __restrict__ llen,
                                                     const uint32_t* __restrict__ genes,
                                                     const uint16_t* __restrict__ ref_codes,
                                                     uint32_t Gf, int nbins,
                                                     int16_t* __restrict__ out,
                                                     float* __restrict__ out_sq)
{
    extern __shared__ int smem[];
    __shared__ float part[128];
    __shared__ int   tmp[2];
    ft_rank_ref_item(blockIdx.x, groups, ngroups, labels, lstart, llen, genes, ref_codes, Gf,
                     nbins, out, out_sq, smem, part, tmp);
}

// Query rows: densify the CSR row over the fine-tune genes, gather G_sub
// (csr2dense single-row form, src/cusingler.cu:241-264), rank (rankdata,
// :159-179) and form the chunk-tree sum of squares.
__device__ inline void ft_rank_qry_item(const FtRankItem& it, uint32_t idx,
                                        const FtGroup* __restrict__ groups,
                                        const uint32_t* __restrict__ genes,
                                        const uint16_t* __restrict__ qdata,
                                        const int* __restrict__ qind, const int* __restrict__ qptr,
                                        uint32_t Gf, int nbins, int16_t* __restrict__ out,
                                        float* __restrict__ out_sq, int* smem, float* part, int* tmp)
{
    const FtGroup g     = groups[it.group];
    // LDS: code bins | the dense row over the fine-tune genes | the codes at
    // G_sub; the rank row reuses the dense row's space (dead after the
    // gather, and G_sub is a subset of the fine-tune genes)
    int*          bins  = smem;
    uint16_t*     dense = (uint16_t*)(smem + ((nbins + 3) & ~3));
    uint16_t*     codes = dense + ((Gf + 7) & ~7u);
    int16_t*      row   = (int16_t*)dense;
    for (uint32_t i = threadIdx.x; i < Gf; i += 128)
        dense[i] = 0;
    __syncthreads();
    const int beg = qptr[it.src_row], end = qptr[it.src_row + 1];
    for (int i = beg + (int)threadIdx.x; i < end; i += 128)
        dense[qind[i]] = qdata[i];
    __syncthreads();
    const uint32_t* gl = genes + g.gene_off;
    for (uint32_t i = threadIdx.x; i < g.gs; i += 128)
        codes[i] = dense[gl[i]];
    __syncthreads();
    rank_codes<128>(codes, (int)g.gs, nbins, bins, row, tmp);
    int16_t* dst = out + it.dst;
    for (uint32_t i = threadIdx.x; i < g.stride; i += 128)
        dst[i] = i < g.gs ? row[i] : (int16_t)0;
    const float sq = chunk_tree_sum128((int)g.gs, [&](int i) { return (float)row[i]; }, part);
    if (threadIdx.x == 0)
        out_sq[idx] = sq;
}

__global__ __launch_bounds__(128) void k_ft_rank_qry(const FtRankItem* __restrict__ items,
                                                     const FtGroup* __restrict__ groups,
                                                     const uint32_t* __restrict__ genes,
                                                     const uint16_t* __restrict__ qdata,
                                                     const int* __restrict__ qind,
                                                     const int* __restrict__ qptr, uint32_t Gf,
                                                     int nbins, int16_t* __restrict__ out,
                                                     float* __restrict__ out_sq)
{
    extern __shared__ int smem[];
    __shared__ float part[128];
    __shared__ int   tmp[2];
    ft_rank_qry_item(items[blockIdx.x], blockIdx.x, groups, genes, qdata, qind, qptr, Gf, nbins,
                     out, out_sq, smem, part, tmp);
}

// Correlation tile of one group: TC cells x TR reference rows, one pair per
// thread, rows staged in LDS as int16.  sxy follows spearman_reduce: 128
// chunks of ceil(gs/128) genes summed by fma in order, folded by the
// stride-1..64 tree, which for chunk sums c_0..c_127 is the balanced binary
// tree ((c0+c1)+(c2+c3))+...; here the tree is built left to right with a
// 7-level stack (a completed node is added to its left sibling).
struct FtCorrTile
{
    uint32_t group;
    uint32_t cell0;  // first cell of the tile within the group
    uint32_t ref0;   // first reference row of the tile within the group
    uint32_t lim;    // at most this many cells from cell0 (0: no limit)
};

// Merges chunk sum `v` of chunk `ch` into the left-to-right tree: while the
// low bits of ch are ones the node closes a pair with its left sibling; the
// last chunk (127) closes the whole tree into `full`.
__device__ inline void ft_tree_push(float (&stk)[7], float& full, float v, int ch)
{
    int  cc   = ch;
    bool open = true;
#pragma unroll
    for (int l = 0; l < 7; ++l)
    {
        if (open)
        {
            if (cc & 1)
            {
                v = stk[l] + v;
                cc >>= 1;
            }
            else
            {
                stk[l] = v;
                open   = false;
            }
        }
    }
    if (open)
        full = v;
}

// Root of the tree after the nz chunks that hold genes: the remaining
// chunks are zero and adding a zero subtree leaves a sum unchanged, so the
// root is the pending left subtrees (set bits of nz) folded from the lowest
// level up.
__device__ inline float ft_tree_root(const float (&stk)[7], float full, int nz)
{
    float v    = 0.f;
    bool  have = false;
#pragma unroll
    for (int l = 0; l < 7; ++l)
        if ((nz >> l) & 1)
        {
            v    = have ? stk[l] + v : stk[l];
            have = true;
        }
    return nz == 128 ? full : v;
}

// LDS stage width (genes, a multiple of 8) of the FP32 correlation tiles for
// rows of stride st when a block may use `lds` bytes: the whole row when it
// fits, else the widest stage that does (the tile then walks K in stages).
__host__ __device__ inline int ft_corr_stage(int st, size_t lds, int rows)
{
    const int words = (int)(lds / (4 * (size_t)rows)) - 1;  // pitch = kt/2 + 1 words
    const int kt    = (words * 2) & ~7;
    return st <= kt ? st : kt;
}

template <int TC, int TR, int RC, int RR>
__device__ inline void ft_corr_tile(const FtCorrTile& tl, const FtGroup* __restrict__ groups,
                                    const int16_t* __restrict__ cell_rank,
                                    const float* __restrict__ cell_sq,
                                    const int16_t* __restrict__ ref_rank,
                                    const float* __restrict__ ref_sq, float* __restrict__ scores,
                                    int* srow32, int kt_max)
{
    constexpr int    NT    = (TC / RC) * (TR / RR);
    const FtGroup    g     = groups[tl.group];
    const int        gs    = (int)g.gs;
    const int        st    = (int)g.stride;  // multiple of 8
    // genes are staged in LDS kt at a time (all of them when the row fits);
    // the LDS row pitch kt + 2 int16 is an odd number of 32-bit words, so
    // lanes reading one gene of different rows hit different banks
    const int        kt    = st < kt_max ? st : kt_max;  // multiple of 8
    const int        pw    = kt / 2 + 1;     // pitch in words
    const int        ncell = min(TC, (int)g.ncell - (int)tl.cell0);
    const int        nref  = min(TR, (int)g.nref - (int)tl.ref0);
    int*             cs32  = srow32;             // TC rows
    int*             rs32  = srow32 + TC * pw;   // TR rows
    const int16_t*   cbase = cell_rank + g.cell_dst0 + (uint64_t)tl.cell0 * st;
    const int16_t*   rbase = ref_rank + g.ref_dst0 + (uint64_t)tl.ref0 * st;
    const int16_t*   cs    = (const int16_t*)cs32;
    const int16_t*   rs    = (const int16_t*)rs32;
    const int        p16   = 2 * pw;  // pitch in int16
    // thread -> RC cells x RR refs; the RR refs of a thread are TR/RR apart
    const int tc = threadIdx.x / (TR / RR);
    const int tr = threadIdx.x % (TR / RR);
    int       ci[RC], ri[RR];
#pragma unroll
    for (int a = 0; a < RC; ++a)
        ci[a] = min(tc + a * (TC / RC), TC - 1);
#pragma unroll
    for (int b = 0; b < RR; ++b)
        ri[b] = min(tr + b * (TR / RR), TR - 1);
    const int step = (gs + 127) / 128;
    const int nz   = (gs + step - 1) / step;  // chunks holding genes; later ones are 0
    float     stk[RC][RR][7];
    float     full[RC][RR];
    float     acc[RC][RR];
#pragma unroll
    for (int a = 0; a < RC; ++a)
#pragma unroll
        for (int b = 0; b < RR; ++b)
        {
            full[a][b] = 0.f;
            acc[a][b]  = 0.f;
#pragma unroll
            for (int l = 0; l < 7; ++l)
                stk[a][b][l] = 0.f;
        }
    int ch     = 0;                  // chunk being summed
    int ch_end = step < gs ? step : gs;
    for (int k0 = 0; k0 < gs; k0 += kt)
    {
        const int w4 = min(kt, st - k0) / 8;  // int4 per staged row
        if (k0)
            __syncthreads();  // the previous stage is no longer read
        for (int v = threadIdx.x; v < ncell * w4; v += NT)
        {
            const int  row = v / w4, c = v % w4;
            const int4 val = *(const int4*)(cbase + (uint64_t)row * st + k0 + c * 8);
            int*       d   = cs32 + row * pw + c * 4;
            d[0] = val.x;
            d[1] = val.y;
            d[2] = val.z;
            d[3] = val.w;
        }
        for (int v = threadIdx.x; v < nref * w4; v += NT)
        {
            const int  row = v / w4, c = v % w4;
            const int4 val = *(const int4*)(rbase + (uint64_t)row * st + k0 + c * 8);
            int*       d   = rs32 + row * pw + c * 4;
            d[0] = val.x;
            d[1] = val.y;
            d[2] = val.z;
            d[3] = val.w;
        }
        __syncthreads();
        const int k1 = min(k0 + kt, gs);
        int       i  = k0;
        while (i < k1)
        {
            const int e = min(ch_end, k1);
            for (; i < e; ++i)
            {
                float x[RC], y[RR];
#pragma unroll
                for (int a = 0; a < RC; ++a)
                    x[a] = (float)cs[ci[a] * p16 + i - k0];
#pragma unroll
                for (int b = 0; b < RR; ++b)
                    y[b] = (float)rs[ri[b] * p16 + i - k0];
#pragma unroll
                for (int a = 0; a < RC; ++a)
#pragma unroll
                    for (int b = 0; b < RR; ++b)
                        acc[a][b] = __builtin_fmaf(x[a], y[b], acc[a][b]);
            }
            if (i == ch_end)
            {
#pragma unroll
                for (int a = 0; a < RC; ++a)
#pragma unroll
                    for (int b = 0; b < RR; ++b)
                    {
                        ft_tree_push(stk[a][b], full[a][b], acc[a][b], ch);
                        acc[a][b] = 0.f;
                    }
                ++ch;
                ch_end = min(ch_end + step, gs);
            }
        }
    }
#pragma unroll
    for (int a = 0; a < RC; ++a)
    {
        const int c = tc + a * (TC / RC);
        if (c >= ncell)
            continue;
        const float sq  = cell_sq[g.cell_item0 + tl.cell0 + c];
        float*      out = scores + g.score0 + (uint64_t)(tl.cell0 + c) * g.nref + tl.ref0;
#pragma unroll
        for (int b = 0; b < RR; ++b)
        {
            const int r = tr + b * (TR / RR);
            if (r < nref)
                out[r] = corr_from_sums(ft_tree_root(stk[a][b], full[a][b], nz), sq,
                                        ref_sq[g.ref_item0 + tl.ref0 + r]);
        }
    }
}

template <int TC, int TR, int RC, int RR>
__global__ __launch_bounds__((TC / RC) * (TR / RR)) void k_ft_corr(
    const FtCorrTile* __restrict__ tiles, const FtGroup* __restrict__ groups,
    const int16_t* __restrict__ cell_rank, const float* __restrict__ cell_sq,
    const int16_t* __restrict__ ref_rank, const float* __restrict__ ref_sq,
    float* __restrict__ scores, int kt_max)
{
    extern __shared__ int srow32[];
    ft_corr_tile<TC, TR, RC, RR>(tiles[blockIdx.x], groups, cell_rank, cell_sq, ref_rank, ref_sq,
                                 scores, srow32, kt_max);
}

// Thin tile for groups with few cells: up to TC cells (rows in LDS) x 256
// reference rows, one reference row per thread read straight from memory
// (each row is used by so few cells that staging it gains nothing).  Same
// 128-chunk + tree order as k_ft_corr.
template <int TC>
__device__ inline void ft_corr_thin_tile(const FtCorrTile& tl, const FtGroup* __restrict__ groups,
                                         const int16_t* __restrict__ cell_rank,
                                         const float* __restrict__ cell_sq,
                                         const int16_t* __restrict__ ref_rank,
                                         const float* __restrict__ ref_sq,
                                         float* __restrict__ scores, int16_t* cthin, int kt_max)
{
    const FtGroup    g     = groups[tl.group];
    const int        gs    = (int)g.gs;
    const int        st    = (int)g.stride;
    // cell rows staged kt genes at a time (all of them when they fit)
    const int        kt    = st < kt_max ? st : kt_max;  // multiple of 8
    const int        ncell = min(tl.lim ? (int)tl.lim : TC, min(TC, (int)g.ncell - (int)tl.cell0));
    const int        nref  = min(256, (int)g.nref - (int)tl.ref0);
    const int16_t*   cbase = cell_rank + g.cell_dst0 + (uint64_t)tl.cell0 * st;
    const int        r     = threadIdx.x;
    const bool       live  = r < nref;  // idle lanes still take part in the stage barriers
    const int16_t*   y     = ref_rank + g.ref_dst0 + (uint64_t)(tl.ref0 + (live ? r : 0)) * st;
    const int        step  = (gs + 127) / 128;
    const int        nz    = (gs + step - 1) / step;
    float            stk[TC][7];
    float            full[TC];
    float            acc[TC];
#pragma unroll
    for (int a = 0; a < TC; ++a)
    {
        full[a] = 0.f;
        acc[a]  = 0.f;
#pragma unroll
        for (int l = 0; l < 7; ++l)
            stk[a][l] = 0.f;
    }
    int ch     = 0;
    int ch_end = step < gs ? step : gs;
    for (int k0 = 0; k0 < gs; k0 += kt)
    {
        const int w8 = min(kt, st - k0) / 8;  // int4 per staged row
        if (k0)
            __syncthreads();  // the previous stage is no longer read
        for (int v = threadIdx.x; v < ncell * w8; v += 256)
        {
            const int row = v / w8, c = v % w8;
            *(int4*)(cthin + row * kt + c * 8) =
                *(const int4*)(cbase + (uint64_t)row * st + k0 + c * 8);
        }
        __syncthreads();
        if (!live)
            continue;
        const int k1 = min(k0 + kt, gs);
        int       i  = k0;
        while (i < k1)
        {
            const int e = min(ch_end, k1);
            for (; i < e; ++i)
            {
                const float yv = (float)y[i];
#pragma unroll
                for (int a = 0; a < TC; ++a)
                    if (a < ncell)
                        acc[a] = __builtin_fmaf((float)cthin[a * kt + i - k0], yv, acc[a]);
            }
            if (i == ch_end)
            {
#pragma unroll
                for (int a = 0; a < TC; ++a)
                {
                    ft_tree_push(stk[a], full[a], acc[a], ch);
                    acc[a] = 0.f;
                }
                ++ch;
                ch_end = min(ch_end + step, gs);
            }
        }
    }
    if (!live)
        return;  // no barrier follows in this tile
    const float syy = ref_sq[g.ref_item0 + tl.ref0 + r];
#pragma unroll
    for (int a = 0; a < TC; ++a)
    {
        if (a >= ncell)
            continue;
        scores[g.score0 + (uint64_t)(tl.cell0 + a) * g.nref + tl.ref0 + r] =
            corr_from_sums(ft_tree_root(stk[a], full[a], nz), cell_sq[g.cell_item0 + tl.cell0 + a],
                           syy);
    }
}

template <int TC>
__global__ __launch_bounds__(256) void k_ft_corr_thin(const FtCorrTile* __restrict__ tiles,
                                                      const FtGroup* __restrict__ groups,
                                                      const int16_t* __restrict__ cell_rank,
                                                      const float* __restrict__ cell_sq,
                                                      const int16_t* __restrict__ ref_rank,
                                                      const float* __restrict__ ref_sq,
                                                      float* __restrict__ scores,
                                                      int kt_max)
{
    extern __shared__ int16_t cthin[];
    ft_corr_thin_tile<TC>(tiles[blockIdx.x], groups, cell_rank, cell_sq, ref_rank, ref_sq, scores,
                          cthin, kt_max);
}

// New candidate list of one cell from its per-label scores, following
// finetune_round's tail (src/cusingler.cu:1270-1295): NaN maximum -> {front};
// keep labels with score > min and score >= max - thre, labels equal to the
// maximum moved to the front (each inserted before the previous ones);
// nothing kept -> {front}.
struct FtCell
{
    uint32_t cell;   // query row
    uint32_t group;
    uint32_t pos;    // cell position within the group's rows
    uint32_t pad_;
    uint64_t seg0;   // first per-label quantile (ascending label order)
};

// Segments of the fine-tune select: item s belongs to the cell whose seg0 is
// the last one <= s; its label is the group's (s - seg0)-th, its scores the
// cell's row of the group's score block at that label's offset.
struct FtSegments
{
    const FtCell*   cells;
    uint32_t        ncells;
    const FtGroup*  groups;
    const uint32_t* labels;
    const uint32_t* label_pos;  // offset of each group label within the group's rows
    const uint32_t* llen;
    __device__ Segment get(uint64_t s) const
    {
        uint32_t lo = 0, hi = ncells;
        while (hi - lo > 1)
        {
            const uint32_t mid = (lo + hi) / 2;
            if (cells[mid].seg0 <= s)
                lo = mid;
            else
                hi = mid;
        }
        const FtCell   fc = cells[lo];
        const FtGroup  g  = groups[fc.group];
        const uint32_t j  = (uint32_t)(s - fc.seg0);
        const uint32_t l  = labels[g.label_off + j];
        const uint32_t c  = fc.pos;
        return Segment{ g.score0 + (uint64_t)c * g.nref + label_pos[g.label_off + j], llen[l], l };
    }
};

// One cell of the FP32 filter: `in` holds its k candidates, the new list goes
// to `out`; returns its length.
template <class SF>
__device__ inline uint32_t ft_filter_list(const SF& score_of, const uint8_t* in, uint32_t k,
                                          float thre, uint8_t* out)
{
    // minmax_element over the scores in list order
    int mn = 0, mx = 0;
    if (k >= 2)
    {
        float s0 = score_of(0), s1 = score_of(1);
        if (s1 < s0)
        {
            mn = 1;
            mx = 0;
        }
        else
        {
            mn = 0;
            mx = 1;
        }
        float smn = s1 < s0 ? s1 : s0, smx = s1 < s0 ? s0 : s1;
        uint32_t t = 2;
        while (t < k)
        {
            const float a = score_of(t);
            if (t + 1 == k)
            {
                if (a < smn)
                {
                    mn  = t;
                    smn = a;
                }
                else if (!(a < smx))
                {
                    mx  = t;
                    smx = a;
                }
                break;
            }
            const float b = score_of(t + 1);
            if (b < a)
            {
                if (b < smn)
                {
                    mn  = t + 1;
                    smn = b;
                }
                if (!(a < smx))
                {
                    mx  = t;
                    smx = a;
                }
            }
            else
            {
                if (a < smn)
                {
                    mn  = t;
                    smn = a;
                }
                if (!(b < smx))
                {
                    mx  = t + 1;
                    smx = b;
                }
            }
            t += 2;
        }
    }
    const float vmin = score_of(mn);
    const float vmax = score_of(mx);
    uint32_t    n    = 0;
    if (__builtin_isnan(vmax))
    {
        out[0] = in[0];
        return 1;
    }
    const float th = vmax - thre;
    // labels equal to the maximum, latest first
    for (int t = (int)k - 1; t >= 0; --t)
    {
        const float s = score_of(t);
        if (s <= vmin || s < th)
            continue;
        if (s == vmax)
            out[n++] = in[t];
    }
    for (uint32_t t = 0; t < k; ++t)
    {
        const float s = score_of(t);
        if (s <= vmin || s < th)
            continue;
        if (!(s == vmax))
            out[n++] = in[t];
    }
    if (n == 0)
        out[n++] = in[0];
    return n;
}

__device__ inline uint32_t ft_filter_cell(const FtCell& fc, const FtGroup& g,
                                          const uint32_t* __restrict__ gl,
                                          const float* __restrict__ P, const uint8_t* in,
                                          uint32_t k, float thre, uint8_t* out)
{
    auto score_of = [&](uint32_t pos) {
        const uint32_t lab = in[pos];
        uint32_t       j   = 0;
        while (j < g.nlabel && gl[j] != lab)
            ++j;
        return j < g.nlabel ? P[fc.seg0 + j] : __builtin_nanf("");
    };
    return ft_filter_list(score_of, in, k, thre, out);
}

__global__ void k_ft_filter(const FtCell* __restrict__ cells, uint32_t ncells,
                            const FtGroup* __restrict__ groups,
                            const uint32_t* __restrict__ labels, const float* __restrict__ P,
                            const uint8_t* __restrict__ in_lists,
                            const uint32_t* __restrict__ in_counts, uint32_t L, float thre,
                            uint8_t* __restrict__ out_lists, uint32_t* __restrict__ out_counts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells)
        return;
    const FtCell  fc = cells[i];
    const FtGroup g  = groups[fc.group];
    out_counts[i]    = ft_filter_cell(fc, g, labels + g.label_off, P, in_lists + (size_t)i * L,
                                      in_counts[i], thre, out_lists + (size_t)i * L);
}

// Exact fine-tune contraction (INT8 limbs on MFMA) for gene subsets of at most
// 1919 genes: S4 = sum_i X_cell[i] * X_ref[i] exactly, divided by the
// reference's own chunk-tree FP32 divisor sqrt(sxx * syy).  Tile: 16 cells x
// 256 reference rows per 256-thread block (4 waves of 16 x 64), K staged 64
// at a time; the ranks are split into limbs while staging (X = 256 h + l) and
// the hi*hi product is folded into the cross partial as (16h)(16h).
typedef int v4i_ft __attribute__((ext_vector_type(4)));

constexpr int FT_MFMA_LD = 80;  // bytes per LDS row of the exact fine-tune tile

// TC cells (16 or 64) x 256 reference rows; wave w takes reference rows
// [64 w, 64 w + 64) of all TC cells (TC / 16 x 4 MFMA tiles).
template <int TC>
__device__ inline void ft_mfma_tile_t(const FtCorrTile& tl, const FtGroup* __restrict__ groups,
                                      const int16_t* __restrict__ cell_rank,
                                      const float* __restrict__ cell_sq,
                                      const int16_t* __restrict__ ref_rank,
                                      const float* __restrict__ ref_sq,
                                      float* __restrict__ scores, int8_t (*sa)[TC * FT_MFMA_LD],
                                      int8_t (*sb)[256 * FT_MFMA_LD])
{
    constexpr int LD = FT_MFMA_LD;
    constexpr int NI = TC / 16;
    const FtGroup  g     = groups[tl.group];
    const int      st    = (int)g.stride;
    const int      ncell = min(TC, (int)g.ncell - (int)tl.cell0);
    const int      nref  = min(256, (int)g.nref - (int)tl.ref0);
    const int16_t* cbase = cell_rank + g.cell_dst0 + (uint64_t)tl.cell0 * st;
    const int16_t* rbase = ref_rank + g.ref_dst0 + (uint64_t)tl.ref0 * st;
    const int      tid   = threadIdx.x;
    const int      lane  = tid & 63;
    const int      wid   = tid >> 6;
    const int      r16   = lane & 15;
    const int      kofs  = (lane >> 4) * 16;
    v4i_ft         acc_x[NI][4], acc_ll[NI][4];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            acc_x[i][j]  = v4i_ft{ 0, 0, 0, 0 };
            acc_ll[i][j] = v4i_ft{ 0, 0, 0, 0 };
        }
    // X = 256 h + l, l the signed low byte and h = (X + 128) >> 8 (|X| <=
    // 1919 here): two int16 per packed add and arithmetic shift, the bytes
    // gathered by permutes (12 instructions for 8 values)
    auto split8 = [](int4 v, int2& hi, int2& lo) {
        typedef short v2s __attribute__((ext_vector_type(2)));
        auto h2 = [](int d) {
            v2s t = __builtin_bit_cast(v2s, d);
            t     = (t + (v2s){ 128, 128 }) >> (v2s){ 8, 8 };
            return __builtin_bit_cast(uint32_t, t);
        };
        // bytes 0 and 2 of each of two dwords (low dword first)
        constexpr uint32_t even = 0x06040200u;
        lo = int2{ (int)__builtin_amdgcn_perm((uint32_t)v.y, (uint32_t)v.x, even),
                   (int)__builtin_amdgcn_perm((uint32_t)v.w, (uint32_t)v.z, even) };
        hi = int2{ (int)__builtin_amdgcn_perm(h2(v.y), h2(v.x), even),
                   (int)__builtin_amdgcn_perm(h2(v.w), h2(v.z), even) };
    };
    // the K loop keeps the next step's rank chunks in registers while the
    // current step's MFMAs run (loads of step k0 + 64 overlap them)
    constexpr int NA = (TC * 8 + 255) / 256;
    int4          pa[NA], pb[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NA; ++u)
        {
            const int c = tid + u * 256, row = c >> 3, ch = c & 7, k = k0 + ch * 8;
            pa[u]       = int4{ 0, 0, 0, 0 };
            if (c < TC * 8 && row < ncell && k < st)
                pa[u] = *(const int4*)(cbase + (uint64_t)row * st + k);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const int c = tid + u * 256, row = c >> 3, ch = c & 7, k = k0 + ch * 8;
            pb[u]       = int4{ 0, 0, 0, 0 };
            if (row < nref && k < st)
                pb[u] = *(const int4*)(rbase + (uint64_t)row * st + k);
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < st; k0 += 64)
    {
        __syncthreads();
        // A: TC rows x 8 chunks of 8 values
#pragma unroll
        for (int u = 0; u < NA; ++u)
        {
            const int c = tid + u * 256;
            if (c >= TC * 8)
                break;
            const int row = c >> 3, ch = c & 7;
            int2      hi, lo;
            split8(pa[u], hi, lo);
            *(int2*)(&sa[0][row * LD + ch * 8]) = hi;
            *(int2*)(&sa[1][row * LD + ch * 8]) = lo;
        }
        // B: 256 rows x 8 chunks = 2048 chunks, 8 per thread
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const int c = tid + u * 256, row = c >> 3, ch = c & 7;
            int2      hi, lo;
            split8(pb[u], hi, lo);
            *(int2*)(&sb[0][row * LD + ch * 8]) = hi;
            *(int2*)(&sb[1][row * LD + ch * 8]) = lo;
        }
        __syncthreads();
        if (k0 + 64 < st)
            fetch(k0 + 64);
        v4i_ft ah[NI], al[NI], ah16[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i)
        {
            ah[i] = *(const v4i_ft*)(&sa[0][(i * 16 + r16) * LD + kofs]);
            al[i] = *(const v4i_ft*)(&sa[1][(i * 16 + r16) * LD + kofs]);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                ah16[i][q] = (int)(((unsigned)ah[i][q] << 4) & 0xF0F0F0F0u);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int    row = wid * 64 + j * 16 + r16;
            const v4i_ft bh  = *(const v4i_ft*)(&sb[0][row * LD + kofs]);
            const v4i_ft bl  = *(const v4i_ft*)(&sb[1][row * LD + kofs]);
            v4i_ft       bh16;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                bh16[q] = (int)(((unsigned)bh[q] << 4) & 0xF0F0F0F0u);
#pragma unroll
            for (int i = 0; i < NI; ++i)
            {
                acc_x[i][j]  = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah[i], bl, acc_x[i][j], 0, 0, 0);
                acc_x[i][j]  = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bh, acc_x[i][j], 0, 0, 0);
                acc_x[i][j]  = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah16[i], bh16, acc_x[i][j], 0, 0, 0);
                acc_ll[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bl, acc_ll[i][j], 0, 0, 0);
            }
        }
    }
    // the epilogue's sums of squares (rows and columns outside the tile read
    // row / column 0), all loaded before any is used
    float csq[NI][4], rsq[4];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = i * 16 + (lane >> 4) * 4 + e;
            csq[i][e]   = cell_sq[g.cell_item0 + tl.cell0 + (m < ncell ? m : 0)];
        }
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        const int n = wid * 64 + j * 16 + r16;
        rsq[j]      = ref_sq[g.ref_item0 + tl.ref0 + (n < nref ? n : 0)];
    }
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = i * 16 + (lane >> 4) * 4 + e;
            if (m >= ncell)
                continue;
            float* out = scores + g.score0 + (uint64_t)(tl.cell0 + m) * g.nref + tl.ref0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int n = wid * 64 + j * 16 + r16;
                if (n >= nref)
                    continue;
                const long long s4   = ((long long)acc_x[i][j][e] << 8) + (long long)acc_ll[i][j][e];
                const float     prod = csq[i][e] * rsq[j];
                const float     d    = sqrtf(prod);
                out[n] = d != 0.f ? (float)((double)s4 / (double)d) : __builtin_nanf("");
            }
        }
}

__device__ inline void ft_mfma_tile(const FtCorrTile& tl, const FtGroup* __restrict__ groups,
                                    const int16_t* __restrict__ cell_rank,
                                    const float* __restrict__ cell_sq,
                                    const int16_t* __restrict__ ref_rank,
                                    const float* __restrict__ ref_sq, float* __restrict__ scores,
                                    int8_t (*sa)[16 * FT_MFMA_LD], int8_t (*sb)[256 * FT_MFMA_LD])
{
    ft_mfma_tile_t<16>(tl, groups, cell_rank, cell_sq, ref_rank, ref_sq, scores, sa, sb);
}

__global__ __launch_bounds__(256) void k_ft_mfma(const FtCorrTile* __restrict__ tiles,
                                                 const FtGroup* __restrict__ groups,
                                                 const int16_t* __restrict__ cell_rank,
                                                 const float* __restrict__ cell_sq,
                                                 const int16_t* __restrict__ ref_rank,
                                                 const float* __restrict__ ref_sq,
                                                 float* __restrict__ scores)
{
    __shared__ __attribute__((aligned(16))) int8_t sa[2][16 * FT_MFMA_LD];
    __shared__ __attribute__((aligned(16))) int8_t sb[2][256 * FT_MFMA_LD];
    ft_mfma_tile(tiles[blockIdx.x], groups, cell_rank, cell_sq, ref_rank, ref_sq, scores, sa, sb);
}

// Certified filter for the exact fine-tune: the reference's chunk-tree FP32
// sum differs from the exact one by at most gamma_{step+6} * sum|p| (chunk
// sums of `step` exact products, then a 7-level tree), so every per-label
// quantile of the reference lies within `wp` of the exact one.  When the
// maximum and the minimum are unique and every threshold decision is clear,
// the new list is [max label] + [kept labels in list order]; otherwise (or
// with a NaN) status = 1 and the cell is redone in the reference's FP32 order.
__device__ inline float ft_next_up(float x)
{
    if (x == 0.f)
        return __uint_as_float(1u);
    const unsigned int b = __float_as_uint(x);
    return __uint_as_float(x > 0.f ? b + 1u : b - 1u);
}

// One cell of the certified filter: returns 0 and writes the new list
// (length in *cnt_out) when certified, else 1 (list untouched).
template <class SF>
__device__ inline int ft_filter_exact_list(const SF& score_of, uint32_t gs, const uint8_t* in,
                                           uint32_t k, float thre, uint8_t* out,
                                           uint32_t* cnt_out)
{
    const double u     = 5.9604644775390625e-08;
    const int    step  = ((int)gs + 127) / 128;
    const double n     = (double)(step + 6);
    const double gamma = n * u / (1.0 - n * u);
    const double wp    = (gamma / (1.0 - gamma - 2 * u) + 3 * u) * (1.0 + 1e-6) + 16 * u;
    double       mlo = -1e300, mhi = -1e300, nlo = 1e300, nhi = 1e300;
    for (uint32_t t = 0; t < k; ++t)
    {
        const float v = score_of(t);
        if (__builtin_isnan(v))
            return 1;
        mlo = fmax(mlo, (double)v - wp);
        mhi = fmax(mhi, (double)v + wp);
        nlo = fmin(nlo, (double)v - wp);
        nhi = fmin(nhi, (double)v + wp);
    }
    int nmax = 0, nmin = 0, imax = -1, imin = -1;
    for (uint32_t t = 0; t < k; ++t)
    {
        const double v = (double)score_of(t);
        if (v + wp >= mlo)
        {
            ++nmax;
            imax = (int)t;
        }
        if (v - wp <= nhi)
        {
            ++nmin;
            imin = (int)t;
        }
    }
    if (nmax != 1 || nmin != 1 || imax == imin)
        return 1;
    // float bounds of the reference's maximum, then of max - thre
    float flo = (float)mlo;
    if ((double)flo > mlo)
        flo = -ft_next_up(-flo);
    float fhi = (float)mhi;
    if ((double)fhi < mhi)
        fhi = ft_next_up(fhi);
    const double tlo = (double)(flo - thre);
    const double thi = (double)(fhi - thre);
    // decide every label before writing (out may be read by nobody else, but
    // an uncertified cell must leave it untouched)
    for (uint32_t t = 0; t < k; ++t)
    {
        if ((int)t == imax || (int)t == imin)
            continue;
        const double v = (double)score_of(t);
        if (!(v - wp >= thi) && !(v + wp < tlo))
            return 1;
    }
    uint32_t cnt = 0;
    out[cnt++]   = in[imax];
    for (uint32_t t = 0; t < k; ++t)
    {
        if ((int)t == imax || (int)t == imin)
            continue;
        if ((double)score_of(t) - wp >= thi)
            out[cnt++] = in[t];
    }
    *cnt_out = cnt;
    return 0;
}

__device__ inline int ft_filter_exact_cell(const FtCell& fc, const FtGroup& g,
                                           const uint32_t* __restrict__ gl,
                                           const float* __restrict__ P, const uint8_t* in,
                                           uint32_t k, float thre, uint8_t* out,
                                           uint32_t* cnt_out)
{
    auto score_of = [&](uint32_t pos) {
        const uint32_t lab = in[pos];
        uint32_t       j   = 0;
        while (j < g.nlabel && gl[j] != lab)
            ++j;
        return j < g.nlabel ? P[fc.seg0 + j] : __builtin_nanf("");
    };
    return ft_filter_exact_list(score_of, g.gs, in, k, thre, out, cnt_out);
}

__global__ void k_ft_filter_exact(const FtCell* __restrict__ cells, uint32_t ncells,
                                  const FtGroup* __restrict__ groups,
                                  const uint32_t* __restrict__ labels, const float* __restrict__ P,
                                  const uint8_t* __restrict__ in_lists,
                                  const uint32_t* __restrict__ in_counts, uint32_t L, float thre,
                                  uint8_t* __restrict__ out_lists, uint32_t* __restrict__ out_counts,
                                  uint8_t* __restrict__ status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells)
        return;
    const FtCell  fc = cells[i];
    const FtGroup g  = groups[fc.group];
    status[i] = (uint8_t)ft_filter_exact_cell(fc, g, labels + g.label_off, P,
                                              in_lists + (size_t)i * L, in_counts[i], thre,
                                              out_lists + (size_t)i * L, out_counts + i);
}

}  // namespace csg
