// Host/device restatement of the two libstdc++ algorithms whose exact
// behaviour decides labels in the reference: std::sort on float arrays
// (introsort: median-of-3 pivot, unguarded partition, depth limit
// 2*floor(log2 n) then heapsort, final insertion sort with a 16-element
// threshold) and std::minmax_element (first minimum, last maximum).
//
// The reference calls std::sort inside `percentile` (src/cusingler.cu:947-978)
// and std::minmax_element in get_label / finetune_round
// (src/cusingler.cu:1122, :1270).  When every value is an ordinary float the
// result of the sort is simply the sorted multiset and any selection method
// gives the same order statistics; when NaN values are present (a reference or
// query row that is constant over the gene subset gives sxx*syy == 0 and a NaN
// score, src/cusingler.cu:849-854) the array is no longer totally ordered and
// the positions the introsort leaves values in are what the reference reads.
// These functions reproduce those positions step for step so the device can
// return the reference's result on such inputs too.  They are compiled for
// both the host (checked against the real std::sort in the CPU tests) and the
// device (run by a single lane on the rare NaN-carrying segments).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define CSG_HD __host__ __device__
#else
#define CSG_HD
#endif

namespace csg {

CSG_HD inline bool flt_less(float a, float b) { return a < b; }

CSG_HD inline void swap_f(float* a, float* b)
{
    float t = *a;
    *a      = *b;
    *b      = t;
}

CSG_HD inline int floor_log2(long n)
{
    int k = 0;
    while (n > 1)
    {
        n >>= 1;
        ++k;
    }
    return k;
}

// Moves the median of *a, *b, *c into *result.
CSG_HD inline void move_median_to_first(float* result, float* a, float* b, float* c)
{
    if (flt_less(*a, *b))
    {
        if (flt_less(*b, *c))
            swap_f(result, b);
        else if (flt_less(*a, *c))
            swap_f(result, c);
        else
            swap_f(result, a);
    }
    else if (flt_less(*a, *c))
        swap_f(result, a);
    else if (flt_less(*b, *c))
        swap_f(result, c);
    else
        swap_f(result, b);
}

// Hoare partition of [first, last) around *pivot with no bounds checks in the
// scans, as the library does.  `lo`/`hi` bound the whole array; a scan that
// would leave them sets *oob (the library would read out of bounds there).
CSG_HD inline float* unguarded_partition(float* first, float* last, const float* pivot,
                                         const float* lo, const float* hi, bool* oob)
{
    while (true)
    {
        while (flt_less(*first, *pivot))
        {
            ++first;
            if (first >= hi)
            {
                *oob = true;
                return first - 1;
            }
        }
        --last;
        while (flt_less(*pivot, *last))
        {
            --last;
            if (last < lo)
            {
                *oob = true;
                return first;
            }
        }
        if (!(first < last))
            return first;
        swap_f(first, last);
        ++first;
    }
}

CSG_HD inline void push_heap(float* first, long hole, long top, float value)
{
    long parent = (hole - 1) / 2;
    while (hole > top && flt_less(first[parent], value))
    {
        first[hole] = first[parent];
        hole        = parent;
        parent      = (hole - 1) / 2;
    }
    first[hole] = value;
}

CSG_HD inline void adjust_heap(float* first, long hole, long len, float value)
{
    const long top    = hole;
    long       second = hole;
    while (second < (len - 1) / 2)
    {
        second = 2 * (second + 1);
        if (flt_less(first[second], first[second - 1]))
            second--;
        first[hole] = first[second];
        hole        = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2)
    {
        second      = 2 * (second + 1);
        first[hole] = first[second - 1];
        hole        = second - 1;
    }
    push_heap(first, hole, top, value);
}

CSG_HD inline void heap_sort(float* first, float* last)
{
    const long len = last - first;
    if (len >= 2)
    {
        long parent = (len - 2) / 2;
        while (true)
        {
            float v = first[parent];
            adjust_heap(first, parent, len, v);
            if (parent == 0)
                break;
            parent--;
        }
    }
    while (last - first > 1)
    {
        --last;
        float v = *last;
        *last   = *first;
        adjust_heap(first, 0, last - first, v);
    }
}

CSG_HD inline void insertion_sort(float* first, float* last)
{
    if (first == last)
        return;
    for (float* i = first + 1; i != last; ++i)
    {
        if (flt_less(*i, *first))
        {
            float v = *i;
            for (float* p = i; p != first; --p)
                *p = *(p - 1);
            *first = v;
        }
        else
        {
            float  v    = *i;
            float* hole = i;
            float* next = i - 1;
            while (flt_less(v, *next))
            {
                *hole = *next;
                hole  = next;
                --next;
            }
            *hole = v;
        }
    }
}

// Unguarded insertion of [first, last): values move left while smaller than
// their neighbour, relying on a smaller-or-incomparable value somewhere to the
// left.  Walking below `lo` sets *oob.
CSG_HD inline void unguarded_insertion_sort(float* first, float* last, const float* lo, bool* oob)
{
    for (float* i = first; i != last; ++i)
    {
        float  v    = *i;
        float* hole = i;
        float* next = i - 1;
        while (true)
        {
            if (next < lo)
            {
                *oob = true;
                *hole = v;
                return;
            }
            if (!flt_less(v, *next))
                break;
            *hole = *next;
            hole  = next;
            --next;
        }
        *hole = v;
    }
}

// Pending ranges of std_sort (the library's recursion, made explicit): at
// most 2 floor(log2 n) + 1 are open at once.
struct StdSortRange
{
    float* first;
    float* last;
    int    depth;
};
constexpr int kStdSortStack = 130;

// std::sort(a, a + n) with operator<, its pending ranges in `stack`
// (kStdSortStack entries: LDS for a device caller that must not spill them
// to scratch).  Returns false if the library's unguarded scans would have
// left the array (then the reference itself has undefined behaviour and no
// result is claimed).
CSG_HD inline bool std_sort(float* a, long n, StdSortRange* stack)
{
    if (n <= 1)
        return true;
    bool        oob = false;
    const float* lo = a;
    const float* hi = a + n;
    using Range = StdSortRange;
    int   sp   = 0;
    stack[sp++] = Range{ a, a + n, 2 * floor_log2(n) };
    while (sp > 0 && !oob)
    {
        Range r = stack[--sp];
        float* first = r.first;
        float* last  = r.last;
        int    depth = r.depth;
        while (last - first > 16)
        {
            if (depth == 0)
            {
                heap_sort(first, last);
                break;
            }
            --depth;
            float* mid = first + (last - first) / 2;
            move_median_to_first(first, first + 1, mid, last - 1);
            float* cut = unguarded_partition(first + 1, last, first, lo, hi, &oob);
            if (oob)
                return false;
            // The library recurses into [cut, last) and loops on [first, cut);
            // the two ranges are disjoint, so the order they run in does not
            // change the result.
            stack[sp++] = Range{ cut, last, depth };
            last        = cut;
        }
    }
    if (oob)
        return false;
    if (n > 16)
    {
        insertion_sort(a, a + 16);
        unguarded_insertion_sort(a + 16, a + n, lo, &oob);
    }
    else
        insertion_sort(a, a + n);
    return !oob;
}

// std::sort(a, a + n) with its pending ranges on the caller's stack.
CSG_HD inline bool std_sort(float* a, long n)
{
    StdSortRange stack[kStdSortStack];
    return std_sort(a, n, stack);
}

// std::minmax_element(a, a + n) with operator<: writes the indices of the
// first minimum and of the last maximum (libstdc++ pairs the elements).
CSG_HD inline void std_minmax(const float* a, int n, int* imin, int* imax)
{
    if (n <= 1)
    {
        *imin = 0;
        *imax = 0;
        return;
    }
    int mn, mx;
    if (flt_less(a[1], a[0]))
    {
        mn = 1;
        mx = 0;
    }
    else
    {
        mn = 0;
        mx = 1;
    }
    int i = 2;
    while (i < n)
    {
        if (i + 1 == n)
        {
            if (flt_less(a[i], a[mn]))
                mn = i;
            else if (!flt_less(a[i], a[mx]))
                mx = i;
            break;
        }
        const int j = i + 1;
        if (flt_less(a[j], a[i]))
        {
            if (flt_less(a[j], a[mn]))
                mn = j;
            if (!flt_less(a[i], a[mx]))
                mx = i;
        }
        else
        {
            if (flt_less(a[i], a[mn]))
                mn = i;
            if (!flt_less(a[j], a[mx]))
                mx = j;
        }
        i += 2;
    }
    *imin = mn;
    *imax = mx;
}

// Parameters of the reference's `percentile(arr, len, p)` that depend only on
// (len, p) (src/cusingler.cu:947-978): which branch is taken and, for the
// interpolating branch, the position `k` of lower_bound(index, p) together
// with index[k-1] and the float step.
struct PctParam
{
    int32_t kind;        // 0: len <= 1 (front), 1: p <= 0 (arr[0]), 2: p >= last (arr[len-1]), 3: interpolate
    int32_t k;           // kind 3: arr[k-1] and arr[k] are read
    float   prev_index;  // index[k-1]
    float   step;        // (float)(1.0 / (len - 1))
};

// Host-side construction; the float arithmetic follows the reference's host
// code (index[i] = i * step in float, lower_bound on that vector).
inline PctParam make_pct_param(long len, float p)
{
    PctParam r{ 0, 0, 0.f, 0.f };
    if (len <= 1)
        return r;
    const float step = (float)(1.0 / (double)(len - 1));
    r.step           = step;
    const float first = 0 * step;
    const float back  = (float)(len - 1) * step;
    if (p <= first)
    {
        r.kind = 1;
        return r;
    }
    if (back <= p)
    {
        r.kind = 2;
        return r;
    }
    // lower_bound: first i with !(index[i] < p)
    long lo = 0, hi = len;
    while (lo < hi)
    {
        long  mid = lo + (hi - lo) / 2;
        float v   = (float)mid * step;
        if (v < p)
            lo = mid + 1;
        else
            hi = mid;
    }
    r.kind       = 3;
    r.k          = (int32_t)lo;
    r.prev_index = (float)(lo - 1) * step;
    return r;
}

// (p - prevIndex) * (nextValue - prevValue) / step + prevValue, every
// operation rounded to float on its own (the reference evaluates it in host
// float arithmetic without fused operations).
CSG_HD inline float pct_interp(float p, float prev_index, float step, float prev, float next)
{
#if defined(__HIPCC__)
#pragma clang fp contract(off)
#endif
    const float t0 = p - prev_index;
    const float t1 = next - prev;
    const float t2 = t0 * t1;
    const float t3 = (float)((double)t2 / (double)step);
    return t3 + prev;
}

}  // namespace csg
