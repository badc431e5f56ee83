// Host preprocessing of the SingleR annotation path: the reference's
// DataParser (src/io.cpp) restated for many cores.  Every result is
// bit-identical to the reference's; only the way it is computed differs:
//   - per-label gene medians come from order statistics of the stored values
//     plus the implicit-zero count instead of sorting zero-padded columns;
//   - each ordered label pair sorts only its positive differences and selects
//     the few further genes the score gene set needs, instead of sorting all
//     genes;
//   - rows are value-coded in parallel with a sort + unique per row instead of
//     a std::set per row.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cusingler_amd.h"
#include "../../include/cusingler_prep.h"

namespace {

struct Csr
{
    std::vector<float> data;
    std::vector<int>   indices;
    std::vector<int>   indptr;
};

struct Coded
{
    std::vector<uint16_t> data;
    std::vector<int32_t>  indices;
    std::vector<int32_t>  indptr;
};

template <class F> void parallel_for(int threads, size_t n, F f)
{
    threads = std::max(1, std::min<int>(threads, (int)std::min<size_t>(n, 256)));
    if (threads <= 1 || n < 2)
    {
        for (size_t i = 0; i < n; ++i)
            f(i, 0);
        return;
    }
    std::vector<std::thread> ts;
    std::atomic_size_t       next{ 0 };
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&, t]() {
            while (true)
            {
                size_t i = next.fetch_add(1);
                if (i >= n)
                    break;
                f(i, t);
            }
        });
    for (auto& th : ts)
        th.join();
}

}  // namespace

struct csg_prep
{
    int      cores = 1;
    uint32_t ref_height = 0, ref_width = 0, qry_height = 0, qry_width = 0;
    uint32_t L = 0;
    Csr      ref, qry;
    std::vector<uint32_t> ctidx;         // [start, len] per label
    std::vector<uint32_t> train_idxs;    // [pos, len] per ordered pair
    std::vector<uint32_t> train_values;  // positive-difference genes per pair
    std::vector<uint32_t> common_genes, thre_genes;  // ascending
    bool filtered = false;
    std::vector<int> qcol;  // query column -> shared column (removeCols), -1: dropped
    std::vector<int> rcol;  // the same for the reference's columns
    // step results
    Coded                 cref, cqry;
    std::vector<uint32_t> ctdidx, ctdiff;
    uint64_t              max_uniq = 0;
    std::vector<uint32_t> step_genes;
};

namespace {

int fail(char* buf, size_t len, int code, const std::string& m)
{
    if (buf && len)
    {
        std::snprintf(buf, len, "%s", m.c_str());
    }
    return code;
}

// removeCols (src/io.cpp:687-711): keep mapped columns, renumbered.
void remove_cols(Csr& a, const std::vector<int>& map_of)
{
    Csr o;
    o.indptr.reserve(a.indptr.size());
    o.indptr.push_back(0);
    for (size_t r = 0; r + 1 < a.indptr.size(); ++r)
    {
        for (int i = a.indptr[r]; i < a.indptr[r + 1]; ++i)
        {
            const int m = map_of[a.indices[i]];
            if (m < 0)
                continue;
            o.indices.push_back(m);
            o.data.push_back(a.data[i]);
        }
        o.indptr.push_back((int)o.data.size());
    }
    a = std::move(o);
}

// k-th smallest (0-based) of the multiset {vals} + {nzero zeros}.
float order_stat(std::vector<float>& vals, size_t nzero, size_t k)
{
    // partition: negatives | zeros (stored) | positives
    auto mid1 = std::partition(vals.begin(), vals.end(), [](float v) { return v < 0.f; });
    auto mid2 = std::partition(mid1, vals.end(), [](float v) { return !(v > 0.f); });
    const size_t nneg = (size_t)(mid1 - vals.begin());
    const size_t nz   = (size_t)(mid2 - mid1) + nzero;
    if (k < nneg)
    {
        std::nth_element(vals.begin(), vals.begin() + k, mid1);
        return vals[k];
    }
    if (k < nneg + nz)
        return 0.f;
    const size_t j = k - nneg - nz;
    std::nth_element(mid2, mid2 + j, vals.end());
    return *(mid2 + j);
}

// Median over `len` cells with implicit zeros (src/io.cpp:146-163): odd len ->
// element len/2; even -> (a[len/2] + a[len/2-1]) / 2 in float.
float median_with_zeros(std::vector<float>& vals, size_t len)
{
    const size_t nzero = len - vals.size();
    if (len % 2 == 0)
    {
        const float hi = order_stat(vals, nzero, len / 2);
        const float lo = order_stat(vals, nzero, len / 2 - 1);
        return (hi + lo) / 2;
    }
    return order_stat(vals, nzero, len / 2);
}

// Value coding of one matrix over a gene set (csrFilter, src/io.cpp:470-523).
void code_matrix(const Csr& a, const std::vector<int>& col_map, int threads, Coded& out,
                 uint64_t& max_uniq, bool& overflow)
{
    const size_t rows = a.indptr.size() - 1;
    std::vector<uint32_t> counts(rows + 1, 0);
    parallel_for(threads, rows, [&](size_t r, int) {
        uint32_t n = 0;
        for (int i = a.indptr[r]; i < a.indptr[r + 1]; ++i)
            n += col_map[a.indices[i]] >= 0;
        counts[r + 1] = n;
    });
    out.indptr.assign(rows + 1, 0);
    for (size_t r = 0; r < rows; ++r)
        out.indptr[r + 1] = out.indptr[r] + (int32_t)counts[r + 1];
    out.data.assign((size_t)out.indptr[rows], 0);
    out.indices.assign((size_t)out.indptr[rows], 0);
    std::vector<uint64_t> tmax(std::max(1, threads), 0);
    std::atomic_bool      ovf{ false };
    parallel_for(threads, rows, [&](size_t r, int t) {
        thread_local std::vector<float> u;
        u.clear();
        for (int i = a.indptr[r]; i < a.indptr[r + 1]; ++i)
            if (col_map[a.indices[i]] >= 0)
                u.push_back(a.data[i]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end(), [](float x, float y) { return !(x < y) && !(y < x); }),
                u.end());
        const uint64_t nu = u.size() + 1;
        if (nu > tmax[t % tmax.size()])
            tmax[t % tmax.size()] = nu;
        if (u.size() > 65535)
            ovf = true;
        size_t o = (size_t)out.indptr[r];
        for (int i = a.indptr[r]; i < a.indptr[r + 1]; ++i)
        {
            const int m = col_map[a.indices[i]];
            if (m < 0)
                continue;
            const size_t c = (size_t)(std::lower_bound(u.begin(), u.end(), a.data[i]) - u.begin());
            out.indices[o] = m;
            out.data[o]    = (uint16_t)(c + 1);
            ++o;
        }
    });
    for (auto v : tmax)
        max_uniq = std::max(max_uniq, v);
    overflow = overflow || ovf;
}

}  // namespace

extern "C" {

int csg_prep_train(int cores, uint32_t ref_height, uint32_t ref_width, const float* ref_data,
                   const int32_t* ref_indices, const int32_t* ref_indptr, uint32_t qry_height,
                   uint32_t qry_width, const float* qry_data, const int32_t* qry_indices,
                   const int32_t* qry_indptr, uint32_t n_codes, const int32_t* celltype,
                   const char* const* ref_genes, uint32_t n_ref_genes,
                   const char* const* qry_genes, uint32_t n_qry_genes, csg_prep** out,
                   char* errbuf, size_t errlen)
{
    if (!out)
        return fail(errbuf, errlen, CSG_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (n_codes == 0 || n_codes > 256)
        return fail(errbuf, errlen, CSG_ERR_ARGUMENT,
                    "the number of label codes must be in [1, 256] (labels are uint8)");
    if (ref_height == 0 || qry_height == 0)
        return fail(errbuf, errlen, CSG_ERR_ARGUMENT, "empty reference or query matrix");
    auto check_csr = [&](const char* what, uint32_t rows, uint32_t width, const float* d,
                         const int32_t* ind, const int32_t* ptr) -> int {
        if (ptr[0] != 0)
            return fail(errbuf, errlen, CSG_ERR_ARGUMENT, std::string(what) + ": indptr[0] != 0");
        for (uint32_t r = 0; r < rows; ++r)
            if (ptr[r + 1] < ptr[r])
                return fail(errbuf, errlen, CSG_ERR_ARGUMENT,
                            std::string(what) + ": indptr not monotone");
        for (int64_t i = 0; i < ptr[rows]; ++i)
        {
            if (ind[i] < 0 || (uint32_t)ind[i] >= width)
                return fail(errbuf, errlen, CSG_ERR_ARGUMENT,
                            std::string(what) + ": column index out of range");
            if (std::isnan(d[i]))
                return fail(errbuf, errlen, CSG_ERR_UNDEFINED,
                            std::string(what) + ": NaN value (undefined in the reference)");
        }
        return 0;
    };
    int rc = check_csr("reference", ref_height, ref_width, ref_data, ref_indices, ref_indptr);
    if (rc)
        return rc;
    if ((rc = check_csr("query", qry_height, qry_width, qry_data, qry_indices, qry_indptr)))
        return rc;
    if (n_ref_genes != ref_width || n_qry_genes != qry_width)
        return fail(errbuf, errlen, CSG_ERR_ARGUMENT, "gene name lists must match the widths");

    csg_prep* p   = new csg_prep();
    p->cores      = std::max(1, cores);
    p->ref_height = ref_height;
    p->ref_width  = ref_width;
    p->qry_height = qry_height;
    p->qry_width  = qry_width;
    p->L          = n_codes;
    p->ref.data.assign(ref_data, ref_data + ref_indptr[ref_height]);
    p->ref.indices.assign(ref_indices, ref_indices + ref_indptr[ref_height]);
    p->ref.indptr.assign(ref_indptr, ref_indptr + ref_height + 1);
    p->qry.data.assign(qry_data, qry_data + qry_indptr[qry_height]);
    p->qry.indices.assign(qry_indices, qry_indices + qry_indptr[qry_height]);
    p->qry.indptr.assign(qry_indptr, qry_indptr + qry_height + 1);

    // ---- findIntersectionGenes (src/io.cpp:306-349)
    std::set<std::string> rs(ref_genes, ref_genes + n_ref_genes);
    std::set<std::string> qs(qry_genes, qry_genes + n_qry_genes);
    std::set<std::string> cs;
    std::set_intersection(rs.begin(), rs.end(), qs.begin(), qs.end(),
                          std::inserter(cs, cs.begin()));
    std::vector<int> rmap, qmap;
    if (rs.size() == qs.size() && rs.size() == cs.size())
        p->filtered = false;
    else
    {
        if (cs.empty())
        {
            delete p;
            return fail(errbuf, errlen, CSG_ERR_ARGUMENT,
                        "No intersection genes in two file, please check input data!");
        }
        p->filtered = true;
        rmap.assign(n_ref_genes, -1);
        qmap.assign(n_qry_genes, -1);
        int k = 0;
        for (uint32_t i = 0; i < n_ref_genes; ++i)
            if (cs.count(ref_genes[i]))
                rmap[i] = k++;
        k = 0;
        for (uint32_t i = 0; i < n_qry_genes; ++i)
            if (cs.count(qry_genes[i]))
                qmap[i] = k++;
    }
    if (p->filtered)
    {
        p->qcol = qmap;
        p->rcol = rmap;
    }
    else
    {
        p->qcol.resize(qry_width);
        for (uint32_t i = 0; i < qry_width; ++i)
            p->qcol[i] = (int)i;
        p->rcol.resize(ref_width);
        for (uint32_t i = 0; i < ref_width; ++i)
            p->rcol[i] = (int)i;
    }

    // ---- loadRefData: groupbyCelltypes + resort (src/io.cpp:616-684)
    std::vector<std::vector<uint32_t>> aux(n_codes);
    for (uint32_t i = 0; i < ref_height; ++i)
    {
        const uint8_t c = (uint8_t)celltype[i];  // src/io.cpp:813
        if (c >= n_codes)
        {
            delete p;
            return fail(errbuf, errlen, CSG_ERR_ARGUMENT,
                        "cell type code " + std::to_string(c) + " has no name in codes");
        }
        aux[c].push_back(i);
    }
    {
        Csr      o;
        uint32_t start = 0;
        o.indptr.push_back(0);
        for (uint32_t l = 0; l < n_codes; ++l)
        {
            p->ctidx.push_back(start);
            p->ctidx.push_back((uint32_t)aux[l].size());
            start += (uint32_t)aux[l].size();
            for (uint32_t r : aux[l])
            {
                const int s = p->ref.indptr[r], e = p->ref.indptr[r + 1];
                o.indices.insert(o.indices.end(), p->ref.indices.begin() + s,
                                 p->ref.indices.begin() + e);
                o.data.insert(o.data.end(), p->ref.data.begin() + s, p->ref.data.begin() + e);
                o.indptr.push_back((int)o.data.size());
            }
        }
        p->ref = std::move(o);
    }
    for (uint32_t l = 0; l < n_codes; ++l)
        if (p->ctidx[2 * l + 1] == 0)
        {
            delete p;
            return fail(errbuf, errlen, CSG_ERR_UNDEFINED,
                        "label " + std::to_string(l) +
                            " has no reference cell (its median is undefined in the reference)");
        }
    if (p->filtered)
        remove_cols(p->ref, rmap);

    // ---- trainData (src/io.cpp:141-262).  The width is still the reference's
    // pre-intersection width: genes past the shared ones have median 0.
    const uint32_t W = ref_width;
    std::vector<std::vector<float>> med(n_codes, std::vector<float>(W, 0.f));
    {
        // gather each label's values per gene (column lists)
        parallel_for(p->cores, n_codes, [&](size_t l, int) {
            const uint32_t pos = p->ctidx[2 * l], len = p->ctidx[2 * l + 1];
            std::vector<uint32_t> cnt(W + 1, 0);
            for (uint32_t r = pos; r < pos + len; ++r)
                for (int i = p->ref.indptr[r]; i < p->ref.indptr[r + 1]; ++i)
                    cnt[p->ref.indices[i] + 1]++;
            for (uint32_t g = 0; g < W; ++g)
                cnt[g + 1] += cnt[g];
            std::vector<float>    vals(cnt[W]);
            std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
            for (uint32_t r = pos; r < pos + len; ++r)
                for (int i = p->ref.indptr[r]; i < p->ref.indptr[r + 1]; ++i)
                    vals[fill[p->ref.indices[i]]++] = p->ref.data[i];
            std::vector<float> col;
            for (uint32_t g = 0; g < W; ++g)
            {
                if (cnt[g + 1] == cnt[g])
                {
                    med[l][g] = 0.f;  // all zeros
                    continue;
                }
                col.assign(vals.begin() + cnt[g], vals.begin() + cnt[g + 1]);
                med[l][g] = median_with_zeros(col, len);
            }
        });
    }
    const size_t common_n = (size_t)std::round(500 * std::pow((2 / 3.0), std::log2((double)n_codes)));
    const size_t thre_n   = (size_t)std::round(500 * std::pow((2 / 3.0), std::log2(2)));
    if (common_n > W)
    {
        delete p;
        return fail(errbuf, errlen, CSG_ERR_UNDEFINED,
                    "fewer reference genes than the score gene count per label pair");
    }
    // per ordered pair (a, b), a != b: genes by descending int difference,
    // ties to the lower gene index (src/io.cpp:228-236)
    const size_t                       npair = (size_t)n_codes * n_codes;
    std::vector<std::vector<uint32_t>> pos_genes(npair), top_genes(npair);
    parallel_for(p->cores, npair, [&](size_t pr, int) {
        const uint32_t a = (uint32_t)(pr / n_codes), b = (uint32_t)(pr % n_codes);
        if (a == b)
            return;
        const std::vector<float>& v1 = med[a];
        const std::vector<float>& v2 = med[b];
        std::vector<std::pair<int, uint32_t>> pos, rest;
        for (uint32_t i = 0; i < W; ++i)
        {
            const int d = int(std::round((v1[i] - v2[i]) * 1e6));
            if (d > 0)
                pos.push_back({ d, i });
            else
                rest.push_back({ d, i });
        }
        auto before = [](const std::pair<int, uint32_t>& x, const std::pair<int, uint32_t>& y) {
            return x.first != y.first ? x.first > y.first : x.second < y.second;
        };
        std::sort(pos.begin(), pos.end(), before);
        auto& pg = pos_genes[pr];
        for (auto& e : pos)
            pg.push_back(e.second);
        auto& tg = top_genes[pr];
        for (size_t k = 0; k < common_n && k < pos.size(); ++k)
            tg.push_back(pos[k].second);
        if (common_n > pos.size())
        {
            const size_t need = common_n - pos.size();
            std::partial_sort(rest.begin(), rest.begin() + need, rest.end(), before);
            for (size_t k = 0; k < need; ++k)
                tg.push_back(rest[k].second);
        }
    });
    {
        std::vector<char> in_thre(W, 0), in_common(W, 0);
        uint32_t          idx_start = 0;
        for (size_t pr = 0; pr < npair; ++pr)
        {
            const uint32_t a = (uint32_t)(pr / n_codes), b = (uint32_t)(pr % n_codes);
            if (a == b)
            {
                p->train_idxs.push_back(0);
                p->train_idxs.push_back(0);
                continue;
            }
            const auto& pg = pos_genes[pr];
            for (size_t k = 0; k < pg.size(); ++k)
            {
                p->train_values.push_back(pg[k]);
                if (k < thre_n)
                    in_thre[pg[k]] = 1;
            }
            p->train_idxs.push_back(idx_start);
            p->train_idxs.push_back((uint32_t)pg.size());
            idx_start += (uint32_t)pg.size();
            for (uint32_t g : top_genes[pr])
                in_common[g] = 1;
        }
        for (uint32_t g = 0; g < W; ++g)
        {
            if (in_common[g])
                p->common_genes.push_back(g);
            if (in_thre[g])
                p->thre_genes.push_back(g);
        }
    }
    // ---- loadQryData (src/io.cpp:525-538)
    if (p->filtered)
        remove_cols(p->qry, qmap);
    *out = p;
    return CSG_OK;
}

int csg_prep_generate(csg_prep* p, int step, csg_coded* out)
{
    if (!p || !out || (step != 0 && step != 1))
        return CSG_ERR_ARGUMENT;
    const std::vector<uint32_t>& gs = step == 0 ? p->common_genes : p->thre_genes;
    const uint32_t               W  = p->ref_width;
    std::vector<int>             col_map(std::max<uint32_t>(W, 1), -1);
    for (size_t k = 0; k < gs.size(); ++k)
        col_map[gs[k]] = (int)k;
    p->ctdidx.clear();
    p->ctdiff.clear();
    if (step == 1)
    {
        // pair windows remapped into the fine-tune gene space (src/io.cpp:722-756)
        size_t start = 0;
        for (size_t i = 0; i < p->train_idxs.size(); i += 2)
        {
            const uint32_t pos = p->train_idxs[i], len = p->train_idxs[i + 1];
            if (pos == 0 && len == 0)
            {
                p->ctdidx.push_back(0);
                p->ctdidx.push_back(0);
                continue;
            }
            for (uint32_t j = pos; j < pos + len; ++j)
            {
                const int m = col_map[p->train_values[j]];
                if (m >= 0)
                    p->ctdiff.push_back((uint32_t)m);
            }
            p->ctdidx.push_back((uint32_t)start);
            p->ctdidx.push_back((uint32_t)(p->ctdiff.size() - start));
            start = p->ctdiff.size();
        }
    }
    p->max_uniq   = 0;
    bool overflow = false;
    // query columns live in the same (shared) index space as the reference
    std::vector<int> qmap(std::max<uint32_t>(p->qry_width, W) + 1, -1);
    for (size_t k = 0; k < gs.size(); ++k)
        if (gs[k] < qmap.size())
            qmap[gs[k]] = (int)k;
    code_matrix(p->ref, col_map, p->cores, p->cref, p->max_uniq, overflow);
    code_matrix(p->qry, qmap, p->cores, p->cqry, p->max_uniq, overflow);
    if (overflow)
        return CSG_ERR_UNDEFINED;  // > 65535 distinct values in a row overflow uint16 codes
    p->step_genes        = gs;
    out->ref_height      = p->ref_height;
    out->qry_height      = p->qry_height;
    out->width           = (uint32_t)gs.size();
    out->ct_num          = p->L;
    out->ref_data        = p->cref.data.data();
    out->ref_indices     = p->cref.indices.data();
    out->ref_indptr      = p->cref.indptr.data();
    out->qry_data        = p->cqry.data.data();
    out->qry_indices     = p->cqry.indices.data();
    out->qry_indptr      = p->cqry.indptr.data();
    out->max_uniq_gene   = p->max_uniq;
    out->ctidx           = p->ctidx.data();
    out->ctdidx          = step == 1 ? p->ctdidx.data() : p->train_idxs.data();
    out->ctdiff          = step == 1 ? p->ctdiff.data() : p->train_values.data();
    out->n_ctdiff        = step == 1 ? p->ctdiff.size() : p->train_values.size();
    out->genes           = p->step_genes.data();
    return CSG_OK;
}

uint64_t csg_prep_train_table(const csg_prep* p, int which, uint32_t* out)
{
    if (!p)
        return 0;
    const std::vector<uint32_t>* v = nullptr;
    switch (which)
    {
    case 0: v = &p->train_idxs; break;
    case 1: v = &p->train_values; break;
    case 2: v = &p->common_genes; break;
    case 3: v = &p->thre_genes; break;
    default: return 0;
    }
    if (out)
        std::memcpy(out, v->data(), v->size() * sizeof(uint32_t));
    return v->size();
}

int csg_prep_filtered(const csg_prep* p) { return p && p->filtered ? 1 : 0; }

int csg_prep_ref_map(const csg_prep* p, int32_t* out)
{
    if (!p || !out)
        return CSG_ERR_ARGUMENT;
    for (uint32_t c = 0; c < p->ref_width; ++c)
        out[c] = c < p->rcol.size() ? p->rcol[c] : -1;
    return CSG_OK;
}

int csg_prep_query_map(const csg_prep* p, int step, int32_t* out)
{
    if (!p || !out || (step != 0 && step != 1))
        return CSG_ERR_ARGUMENT;
    const std::vector<uint32_t>& gs = step == 0 ? p->common_genes : p->thre_genes;
    std::vector<int> pos(std::max<size_t>(p->ref_width, p->qry_width) + 1, -1);
    for (size_t k = 0; k < gs.size(); ++k)
        if (gs[k] < pos.size())
            pos[gs[k]] = (int)k;
    for (uint32_t c = 0; c < p->qry_width; ++c)
    {
        const int s = c < p->qcol.size() ? p->qcol[c] : -1;
        out[c]      = s >= 0 && (size_t)s < pos.size() ? pos[s] : -1;
    }
    return CSG_OK;
}

void csg_prep_release(csg_prep* p) { delete p; }

}  // extern "C"
