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
__global__ void k_label_certify(const float* __restrict__ P, uint32_t nq, uint32_t L, double wp,
                                float thre, int32_t* __restrict__ first, float* __restrict__ mask,
                                uint8_t* __restrict__ status, RecheckItem* __restrict__ items,
                                unsigned int* __restrict__ n_items)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    const float* s  = P + (size_t)q * L;
    float*       mk = mask + (size_t)q * L;
    bool         nan = false;
    double       mlo = -1e300, mhi = -1e300;
    for (uint32_t j = 0; j < L; ++j)
    {
        const float v = s[j];
        if (__builtin_isnan(v))
        {
            nan = true;
            break;
        }
        mlo = fmax(mlo, (double)v - wp);
        mhi = fmax(mhi, (double)v + wp);
    }
    if (nan)
    {
        status[q] = 2;
        for (uint32_t j = 0; j < L; ++j)
        {
            const unsigned int t = atomicAdd(n_items, 1u);
            items[t]             = RecheckItem{ q, j };
        }
        return;
    }
    // float bounds of the reference's maximum and threshold
    float flo = (float)mlo;
    if ((double)flo > mlo)
        flo = next_down(flo);
    float fhi = (float)mhi;
    if ((double)fhi < mhi)
        fhi = next_up(fhi);
    const double tlo = (double)(flo - thre);
    const double thi = (double)(fhi - thre);
    int          ncand = 0, cand = -1;
    bool         unsure = false;
    for (uint32_t j = 0; j < L; ++j)
    {
        const double v = (double)s[j];
        if (v + wp >= mlo)
        {
            ++ncand;
            cand = (int)j;
        }
        if (v - wp >= thi)
            mk[j] = 1.f;
        else if (v + wp < tlo)
            mk[j] = 0.f;
        else
            unsure = true;
    }
    if (ncand == 1 && !unsure)
    {
        first[q]  = cand;
        status[q] = 0;
        return;
    }
    status[q] = 1;
    for (uint32_t j = 0; j < L; ++j)
    {
        const double v   = (double)s[j];
        const bool   inC = v + wp >= mlo;
        const bool   unk = !(v - wp >= thi) && !(v + wp < tlo);
        if (inC || unk)
        {
            const unsigned int t = atomicAdd(n_items, 1u);
            items[t]             = RecheckItem{ q, j };
        }
    }
}

// FP32 scores of one query row against every reference row of one label in
// the reference's sequential order (the plain contraction restricted to a
// (cell, label) pair).  One wave per (item, 64 reference rows), one row per
// lane; each 64 x 64 tile of ranks is staged with 16-byte loads into LDS
// rows of 33 words (an odd pitch, so the lanes' column reads do not collide).
__global__ __launch_bounds__(64) void k_rescore(const RecheckItem* __restrict__ items,
                                                const uint64_t* __restrict__ out_off,
                                                const int16_t* __restrict__ Xq,
                                                const int16_t* __restrict__ Xr, int Kp,
                                                const float* __restrict__ sq_q,
                                                const float* __restrict__ sq_r,
                                                const uint32_t* __restrict__ lstart,
                                                const uint32_t* __restrict__ llen,
                                                float* __restrict__ out)
{
    __shared__ int     tile[64 * 33];
    __shared__ int16_t xs[64];
    const RecheckItem it   = items[blockIdx.x];
    const int         lane = threadIdx.x;
    const uint32_t    r0   = lstart[it.l];
    const uint32_t    n    = llen[it.l];
    const int16_t*    xq   = Xq + (size_t)it.q * Kp;
    const float       sq   = sq_q[it.q];
    float*            o    = out + out_off[blockIdx.x];
    const int16_t*    t16  = (const int16_t*)tile;
    // block (item, y) takes reference rows [64 y, 64 y + 64) of the label
    const uint32_t    base = blockIdx.y * 64;
    if (base >= n)
        return;  // uniform over the block
    {
        const uint32_t nrow = min(64u, n - base);
        float          acc  = 0.f;
        for (int k0 = 0; k0 < Kp; k0 += 64)
        {
            __syncthreads();
#pragma unroll
            for (int t = 0; t < 8; ++t)
            {
                const int c = lane + 64 * t, row = c >> 3, ch = c & 7;
                int4      v = int4{ 0, 0, 0, 0 };
                if ((uint32_t)row < nrow)
                    v = *(const int4*)(Xr + (size_t)(r0 + base + row) * Kp + k0 + ch * 8);
                int* d = tile + row * 33 + ch * 4;
                d[0]   = v.x;
                d[1]   = v.y;
                d[2]   = v.z;
                d[3]   = v.w;
            }
            xs[lane] = xq[k0 + lane];
            __syncthreads();
            if ((uint32_t)lane < nrow)
            {
#pragma unroll 16
                for (int k = 0; k < 64; ++k)
                    acc = __builtin_fmaf((float)xs[k], (float)t16[lane * 66 + k], acc);
            }
        }
        if ((uint32_t)lane < nrow)
            o[base + lane] = corr_from_sums(acc, sq, sq_r[r0 + base + lane]);
    }
}

// Writes the re-scored quantiles into P and redoes the decision of every cell
// that needed them (status 1 or 2).
__global__ void k_apply_recheck(const RecheckItem* __restrict__ items, unsigned int n,
                                const float* __restrict__ Pref, float* __restrict__ P,
                                uint32_t L, uint8_t* __restrict__ exact)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const RecheckItem it                = items[i];
    P[(size_t)it.q * L + it.l]          = Pref[i];
    exact[(size_t)it.q * L + it.l]      = 1;
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
