// MFMA score contractions (gfx950): the exact INT8-limb path and the two
// approximate paths (FP16 in / FP32 accumulate, FP32 in / FP32 accumulate).
//
// Operand layout.  A 16x16 MFMA tile takes, in lane l of the wave, row
// (l % 16) of A and column (l % 16) of B together with one slice of the K
// block chosen by the lane group l / 16; the same slice index is used for A
// and B, so the products are summed over the whole K block whatever order the
// hardware walks the slice in.  The 16x16 result sits in lane l as rows
// 4*(l/16) + i, i = 0..3, of column l % 16.
#pragma once

#include "csg_kernels.h"

namespace csg {

typedef int      v4i __attribute__((ext_vector_type(4)));
typedef float    v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// Correlation of the exact path: the exact rank dot product s4 (an integer
// below 2^53, held exactly in a double) times the two rows' inverse roots
// 1/sqrt(sum X^2) (inv_root, csg_kernels.h), in double, rounded
// once to float.  A row without variance has s4 = 0 and an infinite inverse
// root, so the score is NaN, as the reference's 0 divisor makes it.
__device__ inline float exact_corr(double s4, double ir_row, double ir_col)
{
    return (float)((s4 * ir_row) * ir_col);
}

// ------------------------------------------------------------ limb split ---

// X = 256*h + l with l = signed low byte, h = (X - l) / 256; both in int8 for
// |X| <= 32639.  Writes the two planes of a [rows x Kp] int16 matrix.
__global__ void k_split_limbs(const int16_t* __restrict__ X, size_t n, int8_t* __restrict__ hi,
                              int8_t* __restrict__ lo)
{
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (i >= n)
        return;
    const int4 v = *(const int4*)(X + i);
    const int16_t* x = (const int16_t*)&v;
    int8_t hh[8], ll[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
    {
        const int l = (int)(int8_t)(x[u] & 0xff);
        ll[u]       = (int8_t)l;
        hh[u]       = (int8_t)((x[u] - l) >> 8);
    }
    *(int2*)(hi + i) = *(const int2*)hh;
    *(int2*)(lo + i) = *(const int2*)ll;
}

// Three-product limb form, for |X| <= 1007 (gene sets of up to 1008 genes):
// X = 65 h + l with h = floor((X + 32) / 65) in [-15, 15] and l in [-32, 32].
// Then X Y = 65 [(h + l)(h' + l') + (8h)(8h')] - 64 l l', since
// 65 (h l' + l h') + 65 l l' + (65 + 65 * 64) h h' - 64 l l' expands to
// 65^2 h h' + 65 (h l' + l h') + l l'.  The planes hold u = 8 (h + 16) in
// [8, 248] (bytes whose low three bits are zero) and v = l + 32 in [0, 64];
// a 32-bit word of four such bytes gives the three MFMA operands with no
// carry between bytes:
//     (u >> 3) + v     = (h + l) + 48     (the sum of two bytes below 96)
//     u ^ 0x80808080   = 8h                (a signed byte)
//     v                = l + 32
// and the offsets are taken out per row: with s_p = sum_k ((h + l) + 48) and
// s_v = sum_k (l + 32) over the row's Kp columns,
//     X.Y = 65 A1 - 64 A2 - d(X) - d(Y),   d = 3120 s_p - 2048 s_v - 42112 Kp
//         = 3120 sum h + 1072 sum l + 42112 Kp,
// where A1 = sum (u>>3 + v)(u'>>3 + v') + (8h)(8h') and A2 = sum v v' are
// the two int32 partials.  One wave per row; `d` is exact in a double.
__global__ __launch_bounds__(256) void k_split_kf(const int16_t* __restrict__ X, size_t rows,
                                                  int Kp, int8_t* __restrict__ hi,
                                                  int8_t* __restrict__ lo, double* __restrict__ d)
{
    const size_t row = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int    lane = threadIdx.x & 63;
    if (row >= rows)
        return;
    const int16_t* x  = X + row * Kp;
    int8_t*        uh = hi + row * Kp;
    int8_t*        vl = lo + row * Kp;
    int            sh = 0, sl = 0;
    for (int k = lane * 8; k < Kp; k += 512)
    {
        const int4     v  = *(const int4*)(x + k);
        const int16_t* xv = (const int16_t*)&v;
        uint8_t        uu[8], ww[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
        {
            const int h = (xv[u] + 32 + 65 * 16) / 65 - 16;  // numerator > 0
            const int l = xv[u] - 65 * h;
            sh += h;
            sl += l;
            uu[u] = (uint8_t)(8 * (h + 16));
            ww[u] = (uint8_t)(l + 32);
        }
        *(int2*)(uh + k) = *(const int2*)uu;
        *(int2*)(vl + k) = *(const int2*)ww;
    }
    sh = wave_sum(sh);
    sl = wave_sum(sl);
    if (lane == 0)
        d[row] = 3120.0 * sh + 1072.0 * sl + 42112.0 * Kp;
}

// FP16 copy of a [rows x Kp] int16 matrix (exact for |X| <= 2048).
__global__ void k_to_half(const int16_t* __restrict__ X, size_t n, _Float16* __restrict__ out)
{
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (i >= n)
        return;
    const int4 v = *(const int4*)(X + i);
    const int16_t* x = (const int16_t*)&v;
    v8h h;
#pragma unroll
    for (int u = 0; u < 8; ++u)
        h[u] = (_Float16)(float)x[u];
    *(v8h*)(out + i) = h;
}

// ------------------------------------------------- exact INT8-limb GEMM ---

// Block tile 64 query rows x 128 reference rows, 4 waves in a 2 x 2 grid of
// 32 x 64 wave tiles, K staged 64 at a time through double-buffered LDS
// (rows padded to 80 bytes).  Every output keeps three int32 partials
// (hi*hi, hi*lo + lo*hi, lo*lo) recombined exactly in double, so
//     S4 = sum_k Xq[m,k] * Xr[n,k]
// is exact, and the score is exact_corr(S4, ir_q[m], ir_r[n]).
constexpr int I8_BM = 64, I8_BN = 128, I8_BK = 64, I8_LD = 80;

// Row / column maps of one 64 x 128 block of k_gemm_int8 (the same roles as
// the wide kernel's maps below): a_row / b_row give the operand rows staged
// for block rows / columns r (inside the operand buffers); row_ok / col_ok
// which results are stored; row_out the output row of block row m (column 0
// of the block); row_ir / col_ir the inverse roots of block row m / column n.
struct DenseI8Map
{
    int           m0, n0, nq, nr;
    const double* ir_q;
    const double* ir_r;
    float*        S;
    size_t        ldS;
    __device__ size_t a_row(int r) const { return (size_t)(m0 + r); }
    __device__ size_t b_row(int r) const { return (size_t)(n0 + r); }
    __device__ bool   row_ok(int m) const { return m0 + m < nq; }
    __device__ double row_ir(int m) const { return ir_q[m0 + m]; }
    __device__ float* row_out(int m) const { return S + (size_t)(m0 + m) * ldS + n0; }
    __device__ bool   col_ok(int n) const { return n0 + n < nr; }
    __device__ double col_ir(int n) const { return ir_r[n0 + n]; }
};

// One label of a grouped contraction in 64 x 128 blocks: the query rows
// listed in cells[0, mcount) against reference rows b0 .. b0 + ncount - 1;
// row m of the result is stored contiguously at out + m * ld.  Rows past the
// list repeat its last row, reference rows past b_max read row b_max.
struct GroupI8Map
{
    const uint32_t* cells;
    int             mcount, ncount;
    uint32_t        b0, b_max;
    const double*   ir_q;
    const double*   ir_r;
    float*          out;
    size_t          ld;
    __device__ size_t a_row(int r) const { return cells[r < mcount ? r : mcount - 1]; }
    __device__ size_t b_row(int r) const { return b0 + r < b_max ? b0 + r : b_max; }
    __device__ bool   row_ok(int m) const { return m < mcount; }
    __device__ double row_ir(int m) const { return ir_q[a_row(m)]; }
    __device__ float* row_out(int m) const { return out + (size_t)m * ld; }
    __device__ bool   col_ok(int n) const { return n < ncount; }
    __device__ double col_ir(int n) const { return ir_r[b_row(n)]; }
};

template <class Map>
__device__ inline void gemm_int8_block(const int8_t* __restrict__ qh,
                                       const int8_t* __restrict__ ql,
                                       const int8_t* __restrict__ rh,
                                       const int8_t* __restrict__ rl, int Kp, const Map& map)
{
    __shared__ __attribute__((aligned(16))) int8_t sm[2][(2 * I8_BM + 2 * I8_BN) * I8_LD];
    const int tid  = threadIdx.x;
    const int lane = tid & 63;
    const int wid  = tid >> 6;
    const int wm   = wid >> 1;  // 0..1: 32-row half
    const int wn   = wid & 1;   // 0..1: 64-col half

    // global -> LDS staging: A planes 64 rows x 64 B (2 x 256 x 16 B), B planes
    // 128 rows x 64 B (2 x 512 x 16 B); thread t moves 16-byte chunks
    size_t asrc[2], bsrc[4];
#pragma unroll
    for (int u = 0; u < 2; ++u)
        asrc[u] = map.a_row(((tid + u * 256) >> 2) & 63) * (size_t)Kp;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        bsrc[u] = map.b_row(((tid + u * 256) >> 2) & 127) * (size_t)Kp;
    auto load_stage = [&](int k0, int4 (&ra)[2], int4 (&rb)[4]) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            const int c   = tid + u * 256;   // 0..511 over (plane, row, chunk)
            const int pl  = c >> 8;          // 0: hi, 1: lo
            const int ch  = c & 3;
            ra[u] = *(const int4*)((pl ? ql : qh) + asrc[u] + k0 + ch * 16);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int c   = tid + u * 256;   // 0..1023
            const int pl  = c >> 9;
            const int ch  = c & 3;
            rb[u] = *(const int4*)((pl ? rl : rh) + bsrc[u] + k0 + ch * 16);
        }
    };
    auto store_stage = [&](int buf, const int4 (&ra)[2], const int4 (&rb)[4]) {
        int8_t* base = sm[buf];
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            const int c = tid + u * 256, pl = c >> 8, row = (c >> 2) & 63, ch = c & 3;
            *(int4*)(base + (pl * I8_BM + row) * I8_LD + ch * 16) = ra[u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int c = tid + u * 1024 / 4, pl = c >> 9, row = (c >> 2) & 127, ch = c & 3;
            *(int4*)(base + (2 * I8_BM + pl * I8_BN + row) * I8_LD + ch * 16) = rb[u];
        }
    };

    v4i acc_hh[2][4], acc_x[2][4], acc_ll[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            acc_hh[i][j] = v4i{ 0, 0, 0, 0 };
            acc_x[i][j]  = v4i{ 0, 0, 0, 0 };
            acc_ll[i][j] = v4i{ 0, 0, 0, 0 };
        }

    int4 ra[2], rb[4];
    load_stage(0, ra, rb);
    store_stage(0, ra, rb);
    __syncthreads();
    const int r16  = lane & 15;
    const int kofs = (lane >> 4) * 16;
    int       buf  = 0;
    for (int k0 = 0; k0 < Kp; k0 += I8_BK)
    {
        const bool more = k0 + I8_BK < Kp;
        if (more)
            load_stage(k0 + I8_BK, ra, rb);
        const int8_t* base = sm[buf];
        v4i ah[2], al[2], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
        {
            const int row = wm * 32 + i * 16 + r16;
            ah[i] = *(const v4i*)(base + (row)*I8_LD + kofs);
            al[i] = *(const v4i*)(base + (I8_BM + row) * I8_LD + kofs);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int row = wn * 64 + j * 16 + r16;
            bh[j] = *(const v4i*)(base + (2 * I8_BM + row) * I8_LD + kofs);
            bl[j] = *(const v4i*)(base + (2 * I8_BM + I8_BN + row) * I8_LD + kofs);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                acc_hh[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah[i], bh[j], acc_hh[i][j], 0, 0, 0);
                acc_x[i][j]  = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah[i], bl[j], acc_x[i][j], 0, 0, 0);
                acc_x[i][j]  = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bh[j], acc_x[i][j], 0, 0, 0);
                acc_ll[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bl[j], acc_ll[i][j], 0, 0, 0);
            }
        if (more)
        {
            __syncthreads();  // every wave is done reading `buf`
            store_stage(buf ^ 1, ra, rb);
            __syncthreads();
            buf ^= 1;
        }
    }

    // epilogue
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = wm * 32 + i * 16 + (lane >> 4) * 4 + e;
            if (!map.row_ok(m))
                continue;
            const double ir  = map.row_ir(m);
            float*       out = map.row_out(m);
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int n = wn * 64 + j * 16 + r16;
                if (!map.col_ok(n))
                    continue;
                const double s4 = (double)acc_hh[i][j][e] * 65536.0 +
                                  (double)acc_x[i][j][e] * 256.0 + (double)acc_ll[i][j][e];
                out[n] = exact_corr(s4, ir, map.col_ir(n));
            }
        }
}

__global__ __launch_bounds__(256) void k_gemm_int8(
    const int8_t* __restrict__ qh, const int8_t* __restrict__ ql, const int8_t* __restrict__ rh,
    const int8_t* __restrict__ rl, int Kp, const double* __restrict__ ir_q,
    const double* __restrict__ ir_r, int nq, int nr, float* __restrict__ S, size_t ldS)
{
    const DenseI8Map map{ (int)blockIdx.y * I8_BM, (int)blockIdx.x * I8_BN, nq, nr, ir_q, ir_r, S,
                          ldS };
    gemm_int8_block(qh, ql, rh, rl, Kp, map);
}

// Wide-tile variant for |X| <= 1919 (gene sets of at most 1920 genes).  Then
// the high limb h lies in [-8, 7], 16*h is still an int8 and
//     X*Y = 256 * (h*l' + l*h' + (16h)*(16h')) + l*l'
// so two int32 partials suffice: acc_x takes three products, acc_ll one.  16*h
// is formed in registers from the packed high limbs ((w << 4) & 0xF0F0F0F0 per
// dword: the low nibble of each byte is h in two's complement).  With half the
// accumulators a wave holds a 64 x 64 tile and the block 128 x 256.
//
// Operand staging.  K advances 64 at a time through W8_NS stages of LDS, each
// [A hi 128 rows][A lo 128][B hi 256][B lo 256] of 64 bytes, filled by
// direct-to-LDS loads (global_load_lds_dwordx4: lane i of a wave writes 16
// bytes at the wave's LDS address + 16 i) so that two stages are in flight
// while a third is read, without staging registers.  A wave instruction fills
// one group of 16 rows; the 16-byte slot of row r holding K chunk c is
// c ^ w8_swz(r), which keeps the 16-row ds_read_b128 of the MFMA operands free
// of bank conflicts.
constexpr int W8_BM = 128, W8_BN = 256, W8_BK = 64, W8_NS = 3;
constexpr int W8_STAGE = (2 * W8_BM + 2 * W8_BN) * W8_BK;  // 48 KB
// the tile's row and column inverse roots (doubles) follow the stages
constexpr int W8_IR_OFF = W8_NS * W8_STAGE;
constexpr int W8_LDS    = W8_IR_OFF + (W8_BM + W8_BN) * 8;  // dynamic LDS per block
// the three-product form also stages the rows' offset corrections d
constexpr int W8_LDS_KF = W8_LDS + (W8_BM + W8_BN) * 8;

__device__ inline v4i times16(v4i h)
{
    v4i r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        r[i] = (int)(((unsigned)h[i] << 4) & 0xF0F0F0F0u);
    return r;
}

__device__ __forceinline__ int w8_swz(int r) { return ((r >> 1) + (r >> 3)) & 3; }

// Dispatch order -> tile order: workgroup b runs on XCD b % 8 (round-robin
// dispatch), and XCD x takes the contiguous range [start(x), start(x + 1)) of
// the n tiles, so tiles that share operands (neighbours in the tile order)
// share one XCD's L2.
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t n)
{
    const uint32_t x = b % 8, local = b / 8, q = n / 8, r = n % 8;
    return x * q + min(x, r) + local;
}

// Grids over a tile list built on the device (`T` tiles, not known when the
// grid is sized): the blocks of XCD x (block b runs on XCD b % 8; the grid
// is a multiple of 8) take the contiguous range [x T / 8, (x + 1) T / 8) in
// dispatch order, so tiles that share operands (neighbours in the list)
// share one XCD's L2.  A grid of at least T blocks gives each block one
// tile (the blocks past the range return), a smaller one loops.
struct XcdRange
{
    uint32_t k, end, step;
    // skip: the range's first tiles that an earlier launch took
    __device__ XcdRange(uint32_t T, uint32_t skip = 0)
    {
        const uint32_t x = blockIdx.x % 8;
        k                = (uint32_t)((uint64_t)T * x / 8) + skip + blockIdx.x / 8;
        end              = (uint32_t)((uint64_t)T * (x + 1) / 8);
        step             = gridDim.x / 8;
    }
};

// Row / column maps of one 128 x 256 block of the wide kernel.  a_row(r) and
// b_row(r) give the operand rows staged for block rows / columns r (always
// inside the operand buffers); row_ok / col_ok say which results are stored;
// row_ir_at / col_ir_at give the inverse root 1/sqrt(sum X^2) of an A / B row
// (indexed like a_row / b_row).
struct DenseW8Map
{
    int           m0, n0, nq, nr;
    const double* ir_q;
    const double* ir_r;
    float*        S;
    size_t        ldS;
    const double* d_q = nullptr;  // three-product form: row corrections
    const double* d_r = nullptr;
    __device__ size_t a_row(int r) const { return (size_t)(m0 + r); }
    __device__ size_t b_row(int r) const { return (size_t)(n0 + r); }
    __device__ bool   row_ok(int m) const { return m0 + m < nq; }
    __device__ double row_ir_at(size_t a) const { return ir_q[a]; }
    __device__ float* row_out(int m) const { return S + (size_t)(m0 + m) * ldS + n0; }
    __device__ bool   col_ok(int n) const { return n0 + n < nr; }
    __device__ double col_ir_at(size_t b) const { return ir_r[b]; }
};

// One label of a grouped contraction: the query rows listed in cells[0, mcount)
// against the reference rows b0 .. b0 + ncount - 1; row m of the result is
// stored contiguously at out + m * ld.  Rows past the list repeat its last
// row, reference rows past b_max read row b_max.
struct GroupW8Map
{
    const uint32_t* cells;
    int             mcount, ncount;
    uint32_t        b0, b_max;
    const double*   ir_q;
    const double*   ir_r;
    float*          out;
    size_t          ld;
    const double*   d_q = nullptr;  // three-product form: row corrections
    const double*   d_r = nullptr;
    __device__ size_t a_row(int r) const { return cells[r < mcount ? r : mcount - 1]; }
    __device__ size_t b_row(int r) const { return b0 + r < b_max ? b0 + r : b_max; }
    __device__ bool   row_ok(int m) const { return m < mcount; }
    __device__ double row_ir_at(size_t a) const { return ir_q[a]; }
    __device__ float* row_out(int m) const { return out + (size_t)m * ld; }
    __device__ bool   col_ok(int n) const { return n < ncount; }
    __device__ double col_ir_at(size_t b) const { return ir_r[b]; }
};

// Wave `wid`'s six direct-to-LDS loads of one stage: row group g = wid + 8 u
// (groups 0-7 A hi, 8-15 A lo, 16-31 B hi, 32-47 B lo); src[u] is this lane's
// source row offset plus its swizzled 16-byte chunk.
__device__ __forceinline__ void w8_issue(const int8_t* __restrict__ qh,
                                         const int8_t* __restrict__ ql,
                                         const int8_t* __restrict__ rh,
                                         const int8_t* __restrict__ rl, const uint32_t (&src)[6],
                                         int8_t* stage, int wid, int k0)
{
#pragma unroll
    for (int u = 0; u < 6; ++u)
    {
        const int     g    = wid + 8 * u;
        const int8_t* base = g < 8 ? qh : g < 16 ? ql : g < 32 ? rh : rl;
        __builtin_amdgcn_global_load_lds(
            base + (size_t)src[u] + k0, (__attribute__((address_space(3))) void*)(stage + g * 1024),
            16, 0, 0);
    }
}

// `sm` is the block's W8_LDS (W8_LDS_KF) bytes of dynamic LDS; Kp is a
// multiple of 64.  KF: the planes hold the three-product form (k_split_kf)
// and map.d_q / map.d_r the rows' corrections.
template <bool KF, class Map>
__device__ inline void gemm_int8w_block(const int8_t* __restrict__ qh,
                                        const int8_t* __restrict__ ql,
                                        const int8_t* __restrict__ rh,
                                        const int8_t* __restrict__ rl, int Kp, const Map& map,
                                        int8_t* sm)
{
    const int tid  = threadIdx.x;
    const int lane = tid & 63;
    const int wid  = tid >> 6;
    const int wm   = wid >> 2;  // 0..1: 64-row half of the block
    const int wn   = wid & 3;   // 0..3: 64-column quarter

    // A rows: both planes take row rq; B rows: rq and rq + 128 (element
    // offsets below 2^32: a plane holds fewer than 2^32 bytes)
    const int    rq   = wid * 16 + (lane >> 2);
    const size_t arow = map.a_row(rq);
    uint32_t     src[6];
#pragma unroll
    for (int u = 0; u < 6; ++u)
    {
        const int g = wid + 8 * u;
        const int r = (g < 16 ? (g & 7) : ((g - 16) & 15)) * 16 + (lane >> 2);  // row in its plane
        src[u]      = (uint32_t)((g < 16 ? arow : map.b_row(r)) * (size_t)Kp) +
                 (uint32_t)(((lane & 3) ^ w8_swz(r)) * 16);
    }
    v4i acc_x[4][4], acc_ll[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            acc_x[i][j]  = v4i{ 0, 0, 0, 0 };
            acc_ll[i][j] = v4i{ 0, 0, 0, 0 };
        }
    // inverse roots of the rows this lane's loads fetch (A row rq, B rows rq
    // and rq + 128, the same clamped row indices), loaded with them -- the
    // pipeline's counted waits only ever cover more loads than their own --
    // and stored to LDS after the K loop for the epilogue
    const int T = Kp / W8_BK;
    w8_issue(qh, ql, rh, rl, src, sm, wid, 0);
    // (between the stages: complete by the first stage's wait)
    asm volatile("" ::: "memory");
    const double ir_a  = map.row_ir_at(arow);
    const double ir_b0 = map.col_ir_at(map.b_row(rq));
    const double ir_b1 = map.col_ir_at(map.b_row(rq + 128));
    double       d_a = 0.0, d_b0 = 0.0, d_b1 = 0.0;
    if constexpr (KF)
    {
        d_a  = map.d_q[arow];
        d_b0 = map.d_r[map.b_row(rq)];
        d_b1 = map.d_r[map.b_row(rq + 128)];
    }
    asm volatile("" ::: "memory");
    if (T > 1)
        w8_issue(qh, ql, rh, rl, src, sm + W8_STAGE, wid, W8_BK);
    const int r16 = lane & 15;
    const int kc  = lane >> 4;
    for (int t = 0; t < T; ++t)
    {
        // this wave's loads of stage t are done when at most the six of stage
        // t + 1 are outstanding; the barrier then makes every wave's loads of
        // stage t visible and ends every wave's reads of stage t - 1, whose
        // buffer the loads of stage t + 2 overwrite
        if (t + 1 < T)
            __builtin_amdgcn_s_waitcnt(0x0F76);  // vmcnt <= 6
        else
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt 0
        asm volatile("s_barrier" ::: "memory");
        if (t + 2 < T)
            w8_issue(qh, ql, rh, rl, src, sm + ((t + 2) % W8_NS) * W8_STAGE, wid,
                     (t + 2) * W8_BK);
        const int8_t* s = sm + (t % W8_NS) * W8_STAGE;
        v4i ah[4], al[4], ah16[4], bh[4], bl[4], bh16[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            const int row = wm * 64 + i * 16 + r16;
            const int o   = row * W8_BK + ((kc ^ w8_swz(row)) << 4);
            ah[i]         = *(const v4i*)(s + o);
            al[i]         = *(const v4i*)(s + W8_BM * W8_BK + o);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int row = wn * 64 + j * 16 + r16;
            const int o   = row * W8_BK + ((kc ^ w8_swz(row)) << 4);
            bh[j]         = *(const v4i*)(s + 2 * W8_BM * W8_BK + o);
            bl[j]         = *(const v4i*)(s + (2 * W8_BM + W8_BN) * W8_BK + o);
        }
        if constexpr (KF)
        {
            // ah / bh: u = 8 (h + 16), al / bl: v = l + 32 -> operands
            // (u >> 3) + v and u ^ 0x80808080 (no carry between bytes)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                {
                    ah16[i][c] = (int)((unsigned)ah[i][c] ^ 0x80808080u);
                    ah[i][c]   = (int)(((unsigned)ah[i][c] >> 3) + (unsigned)al[i][c]);
                    bh16[i][c] = (int)((unsigned)bh[i][c] ^ 0x80808080u);
                    bh[i][c]   = (int)(((unsigned)bh[i][c] >> 3) + (unsigned)bl[i][c]);
                }
            // acc_x: A1 (two products), acc_ll: A2; an accumulator's
            // successive MFMAs are at least 16 issues apart
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                    {
                        if (p == 0)
                            acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah[i], bh[j], acc_x[i][j], 0, 0, 0);
                        else if (p == 1)
                            acc_ll[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bl[j], acc_ll[i][j], 0, 0, 0);
                        else
                            acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah16[i], bh16[j], acc_x[i][j], 0, 0, 0);
                    }
            continue;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
        {
            ah16[i] = times16(ah[i]);
            bh16[i] = times16(bh[i]);
        }
        // product by product over all 16 output tiles: an accumulator's
        // successive MFMAs are 16 issues apart, so none waits for the last
        // (integer sums: the order does not change the result)
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                {
                    if (p == 0)
                        acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah[i], bl[j], acc_x[i][j], 0, 0, 0);
                    else if (p == 1)
                        acc_ll[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bl[j], acc_ll[i][j], 0, 0, 0);
                    else if (p == 2)
                        acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bh[j], acc_x[i][j], 0, 0, 0);
                    else
                        acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah16[i], bh16[j], acc_x[i][j], 0, 0, 0);
                }
    }
    double* s_ir = (double*)(sm + W8_IR_OFF);
    double* s_d  = s_ir + W8_BM + W8_BN;
    if ((lane & 3) == 0)
    {
        s_ir[rq]               = ir_a;
        s_ir[W8_BM + rq]       = ir_b0;
        s_ir[W8_BM + 128 + rq] = ir_b1;
        if constexpr (KF)
        {
            s_d[rq]               = d_a;
            s_d[W8_BM + rq]       = d_b0;
            s_d[W8_BM + 128 + rq] = d_b1;
        }
    }
    __syncthreads();
    double rir[4][4], cir[4], rd[4][4], cd[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        cir[j] = s_ir[W8_BM + wn * 64 + j * 16 + r16];
        cd[j]  = KF ? s_d[W8_BM + wn * 64 + j * 16 + r16] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            rir[i][e] = s_ir[wm * 64 + i * 16 + (lane >> 4) * 4 + e];
            rd[i][e]  = KF ? s_d[wm * 64 + i * 16 + (lane >> 4) * 4 + e] : 0.0;
        }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = wm * 64 + i * 16 + (lane >> 4) * 4 + e;
            if (!map.row_ok(m))
                continue;
            float* out = map.row_out(m);
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int n = wn * 64 + j * 16 + r16;
                if (!map.col_ok(n))
                    continue;
                // integers below 2^53: exact in any order
                const double s4 =
                    KF ? ((double)acc_x[i][j][e] * 65.0 - (double)acc_ll[i][j][e] * 64.0) -
                             (rd[i][e] + cd[j])
                       : (double)acc_x[i][j][e] * 256.0 + (double)acc_ll[i][j][e];
                __builtin_nontemporal_store(exact_corr(s4, rir[i][e], cir[j]), out + n);
            }
        }
}

template <bool KF>
__global__ __launch_bounds__(512) void k_gemm_int8w(
    const int8_t* __restrict__ qh, const int8_t* __restrict__ ql, const int8_t* __restrict__ rh,
    const int8_t* __restrict__ rl, int Kp, const double* __restrict__ ir_q,
    const double* __restrict__ ir_r, int nq, int nr, float* __restrict__ S, size_t ldS,
    const double* __restrict__ d_q, const double* __restrict__ d_r)
{
    extern __shared__ __attribute__((aligned(16))) int8_t w8_sm[];
    const uint32_t gx = gridDim.x;
    const uint32_t t  = xcd_order(blockIdx.x + gx * blockIdx.y, gx * gridDim.y);
    // XCD ranges are walked in groups of 8 row tiles, column-major inside a
    // group, so a group's query rows stay in L2 while its columns stream
    const uint32_t gsz = 8 * gx, grp = t / gsz, in = t % gsz;
    const uint32_t rows = min(8u, gridDim.y - grp * 8);
    const DenseW8Map map{ (int)(grp * 8 + in % rows) * W8_BM, (int)(in / rows) * W8_BN, nq, nr,
                          ir_q, ir_r, S, ldS, d_q, d_r };
    gemm_int8w_block<KF>(qh, ql, rh, rl, Kp, map, w8_sm);
}

// Grouped contraction over the *ntiles tiles of a device-built list
// (XcdRange): without LOOP each block takes at most one tile, the `skip`
// first of each XCD range being another launch's; with LOOP the blocks loop
// over the rest of each range (a loop over tiles costs the tile body
// registers, so the one-tile form, with a grid sized to the expected tile
// count, does the work and the looping form takes what outgrew that grid).
// Tile (label l, m0, n0) takes the cells
// cells[a_base[l] + m0 ..] (a_cnt[l] in all) against reference rows
// lstart[l] + n0 .. of label l (llen[l] in all); each cell's scores against
// the label's rows are stored contiguously, label l's block starting at
// S + c_off[l].  Tiles are listed label by label, so a label's tiles run on
// one XCD.
template <bool KF, bool LOOP>
__global__ __launch_bounds__(512) void k_gemm_int8w_group(
    const GroupTile* __restrict__ tiles, const uint32_t* __restrict__ cells,
    const uint32_t* __restrict__ a_base, const uint32_t* __restrict__ a_cnt,
    const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ llen,
    const uint64_t* __restrict__ c_off, const int8_t* __restrict__ qh,
    const int8_t* __restrict__ ql, const int8_t* __restrict__ rh, const int8_t* __restrict__ rl,
    int Kp, uint32_t b_max, const double* __restrict__ ir_q, const double* __restrict__ ir_r,
    float* __restrict__ S, const double* __restrict__ d_q, const double* __restrict__ d_r,
    const uint32_t* __restrict__ ntiles, uint32_t skip)
{
    extern __shared__ __attribute__((aligned(16))) int8_t w8_sm[];
    for (XcdRange r(*ntiles, skip); r.k < r.end; r.k += r.step)
    {
        const GroupTile t  = tiles[r.k];
        const uint32_t  nl = llen[t.label];
        const GroupW8Map map{ cells + a_base[t.label] + t.m0,
                              (int)(a_cnt[t.label] - t.m0),
                              (int)(nl - t.n0),
                              lstart[t.label] + t.n0,
                              b_max,
                              ir_q,
                              ir_r,
                              S + c_off[t.label] + (size_t)t.m0 * nl + t.n0,
                              nl,
                              d_q,
                              d_r };
        gemm_int8w_block<KF>(qh, ql, rh, rl, Kp, map, w8_sm);
        if constexpr (!LOOP)
            break;
        __syncthreads();  // the next tile's loads overwrite the stages
    }
}

// ------------------------------------------------ three-partial wide tile ---
//
// For score gene sets beyond the folded kernel's 1920 genes (|X| up to 32639,
// h in [-128, 127]): X*Y = 65536 h*h' + 256 (h*l' + l*h') + l*l' in three int32
// partials.  A 128 x 128 block of 8 waves of 32 x 64 (2 x 4 MFMA tiles, 96
// accumulator registers); K advances 64 at a time through four 32 KB LDS
// stages [A hi 128 rows][A lo][B hi][B lo] of 64 bytes, filled by
// direct-to-LDS loads with three stages in flight, the same swizzled 16-byte
// slots as the folded kernel.  A wave instruction fills one group of 16 rows;
// lane i of every wave loads row (wave * 16 + i / 4) of each plane.
constexpr int X3_BM = 128, X3_BN = 128, X3_NS = 4;
constexpr int X3_STAGE  = (2 * X3_BM + 2 * X3_BN) * W8_BK;  // 32 KB
constexpr int X3_IR_OFF = X3_NS * X3_STAGE;
constexpr int X3_LDS    = X3_IR_OFF + (X3_BM + X3_BN) * 8;  // dynamic LDS per block

// Wave `wid`'s four direct-to-LDS loads of one stage: group wid of each plane.
__device__ __forceinline__ void x3_issue(const int8_t* __restrict__ qh,
                                         const int8_t* __restrict__ ql,
                                         const int8_t* __restrict__ rh,
                                         const int8_t* __restrict__ rl, size_t asrc, size_t bsrc,
                                         int8_t* stage, int wid, int k0)
{
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
        const int8_t* base = u == 0 ? qh : u == 1 ? ql : u == 2 ? rh : rl;
        __builtin_amdgcn_global_load_lds(
            base + (u < 2 ? asrc : bsrc) + k0,
            (__attribute__((address_space(3))) void*)(stage + (wid + 8 * u) * 1024), 16, 0, 0);
    }
}

// `sm` is the block's X3_LDS bytes of dynamic LDS; Kp is a multiple of 64.  The
// maps are the folded kernel's (block columns 0 .. 127).
template <class Map>
__device__ inline void gemm_int8x_block(const int8_t* __restrict__ qh,
                                        const int8_t* __restrict__ ql,
                                        const int8_t* __restrict__ rh,
                                        const int8_t* __restrict__ rl, int Kp, const Map& map,
                                        int8_t* sm)
{
    const int tid  = threadIdx.x;
    const int lane = tid & 63;
    const int wid  = tid >> 6;
    const int wm   = wid >> 1;  // 0..3: 32-row quarter of the block
    const int wn   = wid & 1;   // 0..1: 64-column half
    const int rq   = wid * 16 + (lane >> 2);  // this lane's staged row of every plane
    const size_t arow  = map.a_row(rq), brow = map.b_row(rq);
    const size_t chunk = (size_t)(((lane & 3) ^ w8_swz(rq)) * 16);
    const size_t asrc = arow * (size_t)Kp + chunk, bsrc = brow * (size_t)Kp + chunk;
    v4i acc_hh[2][4], acc_x[2][4], acc_ll[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            acc_hh[i][j] = v4i{ 0, 0, 0, 0 };
            acc_x[i][j]  = v4i{ 0, 0, 0, 0 };
            acc_ll[i][j] = v4i{ 0, 0, 0, 0 };
        }
    const int T = Kp / W8_BK;
    x3_issue(qh, ql, rh, rl, asrc, bsrc, sm, wid, 0);
    // the inverse roots of the lane's rows, between the first two stages
    // (complete by the first stage's wait, which leaves only later stages
    // outstanding)
    asm volatile("" ::: "memory");
    const double ir_a = map.row_ir_at(arow);
    const double ir_b = map.col_ir_at(brow);
    asm volatile("" ::: "memory");
    if (T > 1)
        x3_issue(qh, ql, rh, rl, asrc, bsrc, sm + X3_STAGE, wid, W8_BK);
    if (T > 2)
        x3_issue(qh, ql, rh, rl, asrc, bsrc, sm + 2 * X3_STAGE, wid, 2 * W8_BK);
    const int r16 = lane & 15;
    const int kc  = lane >> 4;
    for (int t = 0; t < T; ++t)
    {
        // stage t is complete when at most the loads of the stages issued
        // after it (t + 1, t + 2: four each) are outstanding; the barrier
        // then makes every wave's loads visible and ends every wave's reads of
        // stage t - 1, whose buffer stage t + 3 fills
        if (t + 2 < T)
            __builtin_amdgcn_s_waitcnt(0x0F78);  // vmcnt <= 8
        else if (t + 1 < T)
            __builtin_amdgcn_s_waitcnt(0x0F74);  // vmcnt <= 4
        else
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt 0
        asm volatile("s_barrier" ::: "memory");
        if (t + 3 < T)
            x3_issue(qh, ql, rh, rl, asrc, bsrc, sm + ((t + 3) % X3_NS) * X3_STAGE, wid,
                     (t + 3) * W8_BK);
        const int8_t* st = sm + (t % X3_NS) * X3_STAGE;
        v4i ah[2], al[2], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
        {
            const int row = wm * 32 + i * 16 + r16;
            const int o   = row * W8_BK + ((kc ^ w8_swz(row)) << 4);
            ah[i]         = *(const v4i*)(st + o);
            al[i]         = *(const v4i*)(st + X3_BM * W8_BK + o);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int row = wn * 64 + j * 16 + r16;
            const int o   = row * W8_BK + ((kc ^ w8_swz(row)) << 4);
            bh[j]         = *(const v4i*)(st + 2 * X3_BM * W8_BK + o);
            bl[j]         = *(const v4i*)(st + (2 * X3_BM + X3_BN) * W8_BK + o);
        }
        // product by product over the wave's 8 output tiles (integer sums:
        // the order does not change the result)
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < 2; ++i)
                {
                    if (p == 0)
                        acc_hh[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah[i], bh[j], acc_hh[i][j], 0, 0, 0);
                    else if (p == 1)
                        acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah[i], bl[j], acc_x[i][j], 0, 0, 0);
                    else if (p == 2)
                        acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bh[j], acc_x[i][j], 0, 0, 0);
                    else
                        acc_ll[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(al[i], bl[j], acc_ll[i][j], 0, 0, 0);
                }
    }
    double* s_ir = (double*)(sm + X3_IR_OFF);
    if ((lane & 3) == 0)
    {
        s_ir[rq]         = ir_a;
        s_ir[X3_BM + rq] = ir_b;
    }
    __syncthreads();
    double rir[2][4], cir[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        cir[j] = s_ir[X3_BM + wn * 64 + j * 16 + r16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            rir[i][e] = s_ir[wm * 32 + i * 16 + (lane >> 4) * 4 + e];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = wm * 32 + i * 16 + (lane >> 4) * 4 + e;
            if (!map.row_ok(m))
                continue;
            float* out = map.row_out(m);
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int n = wn * 64 + j * 16 + r16;
                if (!map.col_ok(n))
                    continue;
                const double s4 = (double)acc_hh[i][j][e] * 65536.0 +
                                  (double)acc_x[i][j][e] * 256.0 + (double)acc_ll[i][j][e];
                __builtin_nontemporal_store(exact_corr(s4, rir[i][e], cir[j]), out + n);
            }
        }
}

__global__ __launch_bounds__(512) void k_gemm_int8x(
    const int8_t* __restrict__ qh, const int8_t* __restrict__ ql, const int8_t* __restrict__ rh,
    const int8_t* __restrict__ rl, int Kp, const double* __restrict__ ir_q,
    const double* __restrict__ ir_r, int nq, int nr, float* __restrict__ S, size_t ldS)
{
    extern __shared__ __attribute__((aligned(16))) int8_t x3_sm[];
    const uint32_t gx = gridDim.x;
    const uint32_t t  = xcd_order(blockIdx.x + gx * blockIdx.y, gx * gridDim.y);
    // XCD ranges walked in groups of 8 row tiles, column-major inside a group
    const uint32_t gsz = 8 * gx, grp = t / gsz, in = t % gsz;
    const uint32_t rows = min(8u, gridDim.y - grp * 8);
    const DenseW8Map map{ (int)(grp * 8 + in % rows) * X3_BM, (int)(in / rows) * X3_BN, nq, nr,
                          ir_q, ir_r, S, ldS };
    gemm_int8x_block(qh, ql, rh, rl, Kp, map, x3_sm);
}

// The grouped contraction in these blocks: the tables and output layout of
// k_gemm_int8w_group, tiles of X3_BM cells x X3_BN reference rows.
__global__ __launch_bounds__(512) void k_gemm_int8x_group(
    const GroupTile* __restrict__ tiles, const uint32_t* __restrict__ cells,
    const uint32_t* __restrict__ a_base, const uint32_t* __restrict__ a_cnt,
    const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ llen,
    const uint64_t* __restrict__ c_off, const int8_t* __restrict__ qh,
    const int8_t* __restrict__ ql, const int8_t* __restrict__ rh, const int8_t* __restrict__ rl,
    int Kp, uint32_t b_max, const double* __restrict__ ir_q, const double* __restrict__ ir_r,
    float* __restrict__ S, const uint32_t* __restrict__ ntiles)
{
    extern __shared__ __attribute__((aligned(16))) int8_t x3_sm[];
    for (XcdRange r(*ntiles); r.k < r.end; r.k += r.step)
    {
        const GroupTile t  = tiles[r.k];
        const uint32_t  nl = llen[t.label];
        const GroupW8Map map{ cells + a_base[t.label] + t.m0,
                              (int)(a_cnt[t.label] - t.m0),
                              (int)(nl - t.n0),
                              lstart[t.label] + t.n0,
                              b_max,
                              ir_q,
                              ir_r,
                              S + c_off[t.label] + (size_t)t.m0 * nl + t.n0,
                              nl };
        gemm_int8x_block(qh, ql, rh, rl, Kp, map, x3_sm);
        __syncthreads();  // the next tile's loads overwrite the stages
    }
}

// The grouped contraction in 64 x 128 blocks with three partials, for any
// score gene set (the wide kernel's folded limbs need |X| <= 1919): tiles
// (label l, m0, n0) with m0 a multiple of 64 and n0 of 128, the same tables,
// output layout and grid as k_gemm_int8w_group.
__global__ __launch_bounds__(256) void k_gemm_int8_group(
    const GroupTile* __restrict__ tiles, const uint32_t* __restrict__ cells,
    const uint32_t* __restrict__ a_base, const uint32_t* __restrict__ a_cnt,
    const uint32_t* __restrict__ lstart, const uint32_t* __restrict__ llen,
    const uint64_t* __restrict__ c_off, const int8_t* __restrict__ qh,
    const int8_t* __restrict__ ql, const int8_t* __restrict__ rh, const int8_t* __restrict__ rl,
    int Kp, uint32_t b_max, const double* __restrict__ ir_q, const double* __restrict__ ir_r,
    float* __restrict__ S, const uint32_t* __restrict__ ntiles)
{
    for (XcdRange r(*ntiles); r.k < r.end; r.k += r.step)
    {
        const GroupTile t  = tiles[r.k];
        const uint32_t  nl = llen[t.label];
        const GroupI8Map map{ cells + a_base[t.label] + t.m0,
                              (int)(a_cnt[t.label] - t.m0),
                              (int)(nl - t.n0),
                              lstart[t.label] + t.n0,
                              b_max,
                              ir_q,
                              ir_r,
                              S + c_off[t.label] + (size_t)t.m0 * nl + t.n0,
                              nl };
        gemm_int8_block(qh, ql, rh, rl, Kp, map);
        __syncthreads();  // the next tile reuses the LDS buffers
    }
}

// ------------------------------------- approximate MFMA contractions ---

// FP16 operands, FP32 accumulation (v_mfma_f32_16x16x32_f16).  Same tiling as
// the INT8 kernel with one plane of 2-byte values: LDS rows of 64 halves.
constexpr int H_BM = 64, H_BN = 128, H_BK = 64, H_LD = 72;  // halves per LDS row

__global__ __launch_bounds__(256) void k_gemm_f16(const _Float16* __restrict__ qa,
                                                  const _Float16* __restrict__ rb_, int Kp,
                                                  const float* __restrict__ sq_q,
                                                  const float* __restrict__ sq_r, int nq, int nr,
                                                  float* __restrict__ S, size_t ldS)
{
    __shared__ __attribute__((aligned(16))) _Float16 sm[2][(H_BM + H_BN) * H_LD];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int m0 = blockIdx.y * H_BM, n0 = blockIdx.x * H_BN;
    auto load = [&](int k0, int4 (&ra)[2], int4 (&rb)[4]) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            const int c = tid + u * 256, row = c >> 3, ch = c & 7;  // 64 rows x 8 chunks
            ra[u] = *(const int4*)(qa + (size_t)(m0 + row) * Kp + k0 + ch * 8);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int c = tid + u * 256, row = c >> 3, ch = c & 7;  // 128 rows
            rb[u] = *(const int4*)(rb_ + (size_t)(n0 + row) * Kp + k0 + ch * 8);
        }
    };
    auto store = [&](int buf, const int4 (&ra)[2], const int4 (&rb)[4]) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            const int c = tid + u * 256, row = c >> 3, ch = c & 7;
            *(int4*)(&sm[buf][row * H_LD + ch * 8]) = ra[u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int c = tid + u * 256, row = c >> 3, ch = c & 7;
            *(int4*)(&sm[buf][(H_BM + row) * H_LD + ch * 8]) = rb[u];
        }
    };
    v4f acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            acc[i][j] = v4f{ 0.f, 0.f, 0.f, 0.f };
    int4 ra[2], rb[4];
    load(0, ra, rb);
    store(0, ra, rb);
    __syncthreads();
    const int r16 = lane & 15, kofs = (lane >> 4) * 8;
    int       buf = 0;
    for (int k0 = 0; k0 < Kp; k0 += H_BK)
    {
        const bool more = k0 + H_BK < Kp;
        if (more)
            load(k0 + H_BK, ra, rb);
#pragma unroll
        for (int kk = 0; kk < H_BK; kk += 32)
        {
            v8h a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *(const v8h*)(&sm[buf][(wm * 32 + i * 16 + r16) * H_LD + kk + kofs]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                b[j] = *(const v8h*)(&sm[buf][(H_BM + wn * 64 + j * 16 + r16) * H_LD + kk + kofs]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more)
        {
            __syncthreads();
            store(buf ^ 1, ra, rb);
            __syncthreads();
            buf ^= 1;
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + e;
            if (m >= nq)
                continue;
            const float sq = sq_q[m];
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int n = n0 + wn * 64 + j * 16 + r16;
                if (n < nr)
                    S[(size_t)m * ldS + n] = corr_from_sums(acc[i][j][e], sq, sq_r[n]);
            }
        }
}

// FP32 operands, FP32 accumulation (v_mfma_f32_16x16x4_f32): lane l holds
// one float of row (l % 16) at k = k0 + l / 16.  Operands are converted from
// the int16 ranks while staging.
constexpr int F_BM = 64, F_BN = 128, F_BK = 32, F_LD = 36;  // floats per LDS row

__global__ __launch_bounds__(256) void k_gemm_f32mfma(const int16_t* __restrict__ Xq,
                                                      const int16_t* __restrict__ Xr, int Kp,
                                                      const float* __restrict__ sq_q,
                                                      const float* __restrict__ sq_r, int nq,
                                                      int nr, float* __restrict__ S, size_t ldS)
{
    __shared__ __attribute__((aligned(16))) float sm[(F_BM + F_BN) * F_LD];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int m0 = blockIdx.y * F_BM, n0 = blockIdx.x * F_BN;
    v4f acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            acc[i][j] = v4f{ 0.f, 0.f, 0.f, 0.f };
    const int r16 = lane & 15, kq = lane >> 4;
    for (int k0 = 0; k0 < Kp; k0 += F_BK)
    {
        // 192 rows x 32 k int16 = 6144 values: 256 threads x 24 (3 x 8)
#pragma unroll
        for (int u = 0; u < 3; ++u)
        {
            const int c = tid + u * 256, row = c >> 2, ch = c & 3;  // 192 rows x 4 chunks of 8
            const int16_t* src = row < F_BM ? Xq + (size_t)(m0 + row) * Kp
                                            : Xr + (size_t)(n0 + row - F_BM) * Kp;
            const int4 v = *(const int4*)(src + k0 + ch * 8);
            const int16_t* x = (const int16_t*)&v;
#pragma unroll
            for (int t = 0; t < 8; ++t)
                sm[row * F_LD + ch * 8 + t] = (float)x[t];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < F_BK; kk += 4)
        {
            float a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = sm[(wm * 32 + i * 16 + r16) * F_LD + kk + kq];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                b[j] = sm[(F_BM + wn * 64 + j * 16 + r16) * F_LD + kk + kq];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int e = 0; e < 4; ++e)
        {
            const int m = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + e;
            if (m >= nq)
                continue;
            const float sq = sq_q[m];
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                const int n = n0 + wn * 64 + j * 16 + r16;
                if (n < nr)
                    S[(size_t)m * ldS + n] = corr_from_sums(acc[i][j][e], sq, sq_r[n]);
            }
        }
}

}  // namespace csg
