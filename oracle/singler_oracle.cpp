// CPU oracle for the SingleR annotation path — TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU, the arithmetic of the reference's
// preprocessing (src/io.cpp) and of its device engine (src/cusingler.cu) so
// that the HIP engine can be checked against it.  It is loaded only by the
// tests, by the smoke check and by the CPU-baseline leg of the benchmark; the
// product never links or calls it.
//
// Variants of the correlation:
//   VAR_REF   (0): FP32 in the reference's summation order.  The score step
//                  sums genes i = 0..G-1 sequentially (src/cusingler.cu:929-938);
//                  the fine-tune step sums 128 contiguous chunks and folds them
//                  with a stride-1,2,4..64 tree (src/cusingler.cu:810-845).
//                  `s += x*y` is written as fmaf(x, y, s): nvcc contracts it by
//                  default (--fmad=true).  For G <= 4096 every product x*y of
//                  half-integer centred ranks is exact, so fmaf and mul+add
//                  agree; beyond that the contraction choice is parity unpinned.
//   VAR_EXACT (1): exact integer sums of doubled centred ranks
//                  X = 2*rank - (G+1) in int64, normalised in double and
//                  rounded once to float.  This is also the FP64 oracle for the
//                  correlation (it is exact up to that final rounding).
// The per-label quantile, the label choice and the fine-tune filter always use
// the reference's float arithmetic (src/cusingler.cu:947-978,1119-1135,1256-1295).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include <omp.h>

typedef unsigned char  uint8;
typedef unsigned short uint16;
typedef unsigned int   uint32;
typedef unsigned long  uint64;

enum
{
    VAR_REF   = 0,
    VAR_EXACT = 1
};

namespace {

struct CodedCsr
{
    std::vector<uint16> data;
    std::vector<int>    indptr, indices;
    uint32              height = 0, width = 0;
};

struct Model
{
    int cores = 1;
    // raw inputs (after prepareData)
    std::vector<float>       ref_data, qry_data;
    std::vector<int>         ref_indices, ref_indptr, qry_indices, qry_indptr;
    uint32                   ref_height = 0, ref_width = 0, qry_height = 0, qry_width = 0;
    std::vector<uint8>       celltype_codes;
    size_t                   n_labels = 0;
    std::vector<std::string> ref_genes, qry_genes;
    bool                     filter_genes = false;
    std::map<uint32, uint32> ref_gene_index, qry_gene_index;
    // train results
    std::vector<uint32> ctidx;        // [start, len] per label
    std::vector<uint32> train_idxs;   // [pos, len] per ordered pair (gene space)
    std::vector<uint32> train_values; // positive-diff genes per pair
    std::set<uint32>    common_genes, thre_genes;
    // coded matrices
    CodedCsr ref0, qry0, ref1, qry1;
    uint64   max_uniq0 = 0, max_uniq1 = 0;
    std::vector<uint32> ctdidx1, ctdiff1;  // pair windows in thre-gene space
    std::string error;
};

// ---------------------------------------------------------------- prepare ---

// src/io.cpp:306-349 (findIntersectionGenes)
bool find_intersection(Model& m)
{
    std::set<std::string> r(m.ref_genes.begin(), m.ref_genes.end());
    std::set<std::string> q(m.qry_genes.begin(), m.qry_genes.end());
    std::set<std::string> c;
    std::set_intersection(r.begin(), r.end(), q.begin(), q.end(), std::inserter(c, c.begin()));
    if (r.size() == q.size() && r.size() == c.size())
    {
        m.filter_genes = false;
        return true;
    }
    if (c.empty())
    {
        m.error = "No intersection genes in two file, please check input data!";
        return false;
    }
    m.filter_genes = true;
    int k = 0;
    for (uint32 i = 0; i < m.ref_genes.size(); ++i)
        if (c.count(m.ref_genes[i]))
            m.ref_gene_index.insert({ i, (uint32)k++ });
    k = 0;
    for (uint32 i = 0; i < m.qry_genes.size(); ++i)
        if (c.count(m.qry_genes[i]))
            m.qry_gene_index.insert({ i, (uint32)k++ });
    return true;
}

// src/io.cpp:687-711 (removeCols): keep mapped columns, renumber by position.
void remove_cols(std::vector<float>& data, std::vector<int>& indptr, std::vector<int>& indices,
                 const std::map<uint32, uint32>& cols)
{
    std::vector<int>   ni, np{ 0 };
    std::vector<float> nd;
    for (size_t j = 0; j + 1 < indptr.size(); ++j)
    {
        for (int i = indptr[j]; i < indptr[j + 1]; ++i)
        {
            auto it = cols.find((uint32)indices[i]);
            if (it == cols.end())
                continue;
            ni.push_back((int)it->second);
            nd.push_back(data[i]);
        }
        np.push_back((int)nd.size());
    }
    indices.swap(ni);
    indptr.swap(np);
    data.swap(nd);
}

// src/io.cpp:616-636 (groupbyCelltypes) + 657-684 (resort): stable regroup of
// the reference rows by label code; ctidx = [start, len] per label.
bool group_and_resort(Model& m)
{
    std::vector<std::vector<uint32>> aux(m.n_labels);
    for (size_t i = 0; i < m.celltype_codes.size(); ++i)
    {
        if (m.celltype_codes[i] >= m.n_labels)
        {
            m.error = "celltype code out of range of codes";
            return false;
        }
        aux[m.celltype_codes[i]].push_back((uint32)i);
    }
    std::vector<uint32> ids;
    uint32              start = 0;
    for (auto& v : aux)
    {
        m.ctidx.push_back(start);
        m.ctidx.push_back((uint32)v.size());
        start += (uint32)v.size();
        ids.insert(ids.end(), v.begin(), v.end());
    }
    std::vector<int>   ni, np{ 0 };
    std::vector<float> nd;
    for (auto line : ids)
    {
        int s = m.ref_indptr[line], e = m.ref_indptr[line + 1];
        np.push_back(np.back() + e - s);
        ni.insert(ni.end(), m.ref_indices.begin() + s, m.ref_indices.begin() + e);
        nd.insert(nd.end(), m.ref_data.begin() + s, m.ref_data.begin() + e);
    }
    m.ref_indices.swap(ni);
    m.ref_indptr.swap(np);
    m.ref_data.swap(nd);
    return true;
}

// src/io.cpp:146-163: median over all len cells with implicit zeros.
float median_of(std::vector<float> cols, size_t len)
{
    cols.resize(len, 0);
    if (cols.size() % 2 == 0)
    {
        std::sort(cols.begin(), cols.end());
        return (cols[cols.size() / 2] + cols[cols.size() / 2 - 1]) / 2;
    }
    std::nth_element(cols.begin(), cols.begin() + cols.size() / 2, cols.end());
    return cols[cols.size() / 2];
}

// src/io.cpp:141-262 (trainData).  ref_width is still the pre-intersection
// width here, so trailing genes beyond the shared ones have median 0.
bool train(Model& m)
{
    const uint32                    W = m.ref_width;
    std::vector<std::vector<float>> medians(m.n_labels);
    for (size_t ct = 0; ct < m.n_labels; ++ct)
    {
        uint32                          pos = m.ctidx[2 * ct], len = m.ctidx[2 * ct + 1];
        std::vector<std::vector<float>> sub(W);
        for (uint32 r = pos; r < pos + len; ++r)
            for (int i = m.ref_indptr[r]; i < m.ref_indptr[r + 1]; ++i)
                sub[m.ref_indices[i]].push_back(m.ref_data[i]);
        medians[ct].assign(W, 0.f);
#pragma omp parallel for schedule(dynamic, 256) num_threads(m.cores)
        for (long g = 0; g < (long)W; ++g)
            medians[ct][g] = median_of(sub[g], len);
    }
    const size_t common_n = (size_t)round(500 * pow((2 / 3.0), log2((double)m.n_labels)));
    const size_t thre_n   = (size_t)round(500 * pow((2 / 3.0), log2(2)));
    uint32       idx_start = 0;
    // src/io.cpp:170,213-217: median_map is a std::map keyed by uint8 label.
    std::map<uint8, size_t> order;
    for (size_t ct = 0; ct < m.n_labels; ++ct)
        order.insert({ (uint8)ct, ct });
    for (auto& a : order)
    {
        for (auto& b : order)
        {
            if (a.first == b.first)
            {
                m.train_idxs.push_back(0);
                m.train_idxs.push_back(0);
                continue;
            }
            const auto&                         v1 = medians[a.second];
            const auto&                         v2 = medians[b.second];
            std::vector<std::pair<int, uint32>> diff;
            diff.reserve(W);
            for (uint32 i = 0; i < W; ++i)
                diff.push_back({ int(round((v1[i] - v2[i]) * 1e6)), W - i - 1 });
            std::sort(diff.begin(), diff.end(), std::greater<std::pair<int, uint32>>());
            uint32 i = 0;
            for (; i < diff.size(); ++i)
            {
                if (diff[i].first <= 0)
                    break;
                m.train_values.push_back(W - diff[i].second - 1);
                if (i < thre_n)
                    m.thre_genes.insert(W - diff[i].second - 1);
            }
            m.train_idxs.push_back(idx_start);
            m.train_idxs.push_back(i);
            idx_start += i;
            for (uint32 k = 0; k < common_n && k < diff.size(); ++k)
                m.common_genes.insert(W - diff[k].second - 1);
        }
    }
    return true;
}

// src/io.cpp:470-523 (csrFilter): keep gene-set columns renumbered in set
// order; each stored value becomes its 1-based rank among the row's distinct
// kept values; max_uniq = max(#distinct + 1).
void csr_filter(const std::vector<float>& data, const std::vector<int>& indptr,
                const std::vector<int>& indices, const std::set<uint32>& cols, CodedCsr& out,
                uint64& max_uniq)
{
    std::unordered_map<uint32, uint32> im;
    uint32                             k = 0;
    for (auto c : cols)
        im[c] = k++;
    out.data.clear();
    out.indices.clear();
    out.indptr.assign(1, 0);
    uint64 cur = 0;
    for (size_t j = 0; j + 1 < indptr.size(); ++j)
    {
        std::set<float> uniq;
        for (int i = indptr[j]; i < indptr[j + 1]; ++i)
            if (im.count((uint32)indices[i]))
                uniq.insert(data[i]);
        cur = std::max<uint64>(uniq.size() + 1, cur);
        std::unordered_map<float, uint16> code;
        uint16                            c = 1;
        for (auto v : uniq)
            code[v] = c++;
        for (int i = indptr[j]; i < indptr[j + 1]; ++i)
        {
            auto it = im.find((uint32)indices[i]);
            if (it == im.end())
                continue;
            out.indices.push_back((int)it->second);
            out.data.push_back(code[data[i]]);
        }
        out.indptr.push_back((int)out.data.size());
    }
    max_uniq = std::max(max_uniq, cur);
    out.width = (uint32)cols.size();
}

// src/io.cpp:714-784 (generateDenseMatrix).  Step 1 also remaps the pair
// windows into the thre-gene index space (:722-756).
void generate(Model& m, int step)
{
    if (step == 0)
    {
        csr_filter(m.ref_data, m.ref_indptr, m.ref_indices, m.common_genes, m.ref0, m.max_uniq0);
        csr_filter(m.qry_data, m.qry_indptr, m.qry_indices, m.common_genes, m.qry0, m.max_uniq0);
        m.ref0.height = m.ref_height;
        m.qry0.height = m.qry_height;
        return;
    }
    std::map<uint32, uint32> im;
    uint32                   k = 0;
    for (auto c : m.thre_genes)
        im[c] = k++;
    m.ctdidx1.clear();
    m.ctdiff1.clear();
    size_t start = 0;
    for (size_t i = 0; i < m.train_idxs.size(); i += 2)
    {
        uint32 pos = m.train_idxs[i], len = m.train_idxs[i + 1];
        if (pos == 0 && len == 0)
        {
            m.ctdidx1.push_back(0);
            m.ctdidx1.push_back(0);
            continue;
        }
        for (uint32 j = pos; j < pos + len; ++j)
        {
            auto it = im.find(m.train_values[j]);
            if (it == im.end())
                continue;
            m.ctdiff1.push_back(it->second);
        }
        m.ctdidx1.push_back((uint32)start);
        m.ctdidx1.push_back((uint32)(m.ctdiff1.size() - start));
        start = m.ctdiff1.size();
    }
    csr_filter(m.ref_data, m.ref_indptr, m.ref_indices, m.thre_genes, m.ref1, m.max_uniq1);
    csr_filter(m.qry_data, m.qry_indptr, m.qry_indices, m.thre_genes, m.qry1, m.max_uniq1);
    m.ref1.height = m.ref_height;
    m.qry1.height = m.qry_height;
}

// ------------------------------------------------------------ arithmetic ---

// Average (tie-mean) 1-based rank of each code in a dense code row:
// rank = #less + (#equal + 1) / 2 — src/cusingler.cu:745-801 (rankdata_bin3,
// ranks[c] = start + (bins[c]+1)*0.5) and :159-179 (rankdata, less+(eq-1)/2
// with less starting at 1); both give the same exact half-integers.
void rank_row(const uint16* codes, uint32 n, float* out)
{
    uint32 maxc = 0;
    for (uint32 i = 0; i < n; ++i)
        maxc = std::max<uint32>(maxc, codes[i]);
    std::vector<int> bins(maxc + 1, 0);
    for (uint32 i = 0; i < n; ++i)
        bins[codes[i]]++;
    std::vector<float> rk(maxc + 1);
    int                start = 0;
    for (uint32 c = 0; c <= maxc; ++c)
    {
        rk[c] = (float)(start + (bins[c] + 1) * 0.5);
        start += bins[c];
    }
    for (uint32 i = 0; i < n; ++i)
        out[i] = rk[codes[i]];
}

void dense_row(const CodedCsr& a, uint32 row, std::vector<uint16>& out)
{
    out.assign(a.width, 0);
    for (int i = a.indptr[row]; i < a.indptr[row + 1]; ++i)
        out[a.indices[i]] = a.data[i];
}

// src/cusingler.cu:918-945 (spearman): sequential FP32 over genes.
float corr_seq(const float* q, const float* r, uint32 G)
{
    float mean = (G + 1) / 2.0;
    float sxy = 0, sxx = 0, syy = 0;
    for (uint32 i = 0; i < G; ++i)
    {
        float x = q[i] - mean;
        float y = r[i] - mean;
        sxy     = fmaf(x, y, sxy);
        sxx     = fmaf(x, x, sxx);
        syy     = fmaf(y, y, syy);
    }
    float divisor = sqrtf(sxx * syy);
    return divisor != 0 ? sxy / divisor : NAN;
}

// src/cusingler.cu:804-856 (spearman_reduce): 128 contiguous chunks of
// ceil(G/128) genes each summed sequentially, then a stride 1..64 tree.
float corr_tree(const float* q, const float* r, uint32 G)
{
    float mean = (G + 1) / 2.0;
    int   step = G / 128;
    if (G % 128 != 0)
        step++;
    float sxy[128], sxx[128], syy[128];
    for (int t = 0; t < 128; ++t)
    {
        sxy[t] = sxx[t] = syy[t] = 0;
        for (int i = t * step; i < (t + 1) * step; ++i)
        {
            if ((uint32)i < G)
            {
                float x = q[i] - mean;
                float y = r[i] - mean;
                sxy[t]  = fmaf(x, y, sxy[t]);
                sxx[t]  = fmaf(x, x, sxx[t]);
                syy[t]  = fmaf(y, y, syy[t]);
            }
        }
    }
    for (int stride = 1; stride < 128; stride *= 2)
        for (int t = 0; t < 128; t += 2 * stride)
        {
            sxy[t] += sxy[t + stride];
            sxx[t] += sxx[t + stride];
            syy[t] += syy[t + stride];
        }
    float divisor = sqrtf(sxx[0] * syy[0]);
    return divisor != 0 ? sxy[0] / divisor : NAN;
}

// Exact correlation of doubled centred ranks, normalised in double.
float corr_exact(const float* q, const float* r, uint32 G)
{
    int64_t sxy = 0, sxx = 0, syy = 0;
    for (uint32 i = 0; i < G; ++i)
    {
        int64_t x = (int64_t)(2.0f * q[i]) - (int64_t)(G + 1);
        int64_t y = (int64_t)(2.0f * r[i]) - (int64_t)(G + 1);
        sxy += x * y;
        sxx += x * x;
        syy += y * y;
    }
    double d = std::sqrt((double)sxx * (double)syy);
    return d != 0 ? (float)((double)sxy / d) : NAN;
}

// Same, returned in double (the FP64 oracle value of the correlation).
double corr_f64(const float* q, const float* r, uint32 G)
{
    int64_t sxy = 0, sxx = 0, syy = 0;
    for (uint32 i = 0; i < G; ++i)
    {
        int64_t x = (int64_t)(2.0f * q[i]) - (int64_t)(G + 1);
        int64_t y = (int64_t)(2.0f * r[i]) - (int64_t)(G + 1);
        sxy += x * y;
        sxx += x * x;
        syy += y * y;
    }
    double d = std::sqrt((double)sxx * (double)syy);
    return d != 0 ? (double)sxy / d : NAN;
}

// src/cusingler.cu:947-978 (percentile), float arithmetic throughout.
float percentile(std::vector<float> arr, int len, float p)
{
    if (len <= 1)
        return arr.front();
    float res;
    std::sort(arr.begin(), arr.begin() + len);
    std::vector<float> index;
    float              step = 1.0 / (len - 1);
    for (int i = 0; i < len; ++i)
        index.push_back(i * step);
    if (p <= index.front())
        res = arr[0];
    else if (index.back() <= p)
        res = arr[len - 1];
    else
    {
        auto  it        = std::lower_bound(index.begin(), index.end(), p);
        float prevIndex = *(it - 1);
        float prevValue = arr.at(it - index.begin() - 1);
        float nextValue = arr.at(it - index.begin());
        res             = (p - prevIndex) * (nextValue - prevValue) / step + prevValue;
    }
    return res;
}

// One round of src/cusingler.cu:1138-1296 (finetune_round) for query row q.
std::vector<uint32> finetune_round(const Model& m, uint32 q, const std::vector<uint32>& top,
                                   float quantile, float thre, int variant,
                                   std::vector<float>* round_scores)
{
    const uint32  L = (uint32)m.n_labels;
    std::set<uint32> uniq;
    int gene_thre = round(500 * pow((2 / 3.0), log2((double)top.size())));
    for (auto i : top)
        for (auto j : top)
        {
            if (i == j)
                continue;
            int pos = m.ctdidx1[(i * L + j) * 2];
            int len = m.ctdidx1[(i * L + j) * 2 + 1];
            if (len > gene_thre)
                len = gene_thre;
            uniq.insert(m.ctdiff1.begin() + pos, m.ctdiff1.begin() + pos + len);
        }
    if (uniq.size() < 20)
        return { top.front() };
    std::vector<uint32> genes(uniq.begin(), uniq.end());
    const uint32        G = (uint32)genes.size();

    std::vector<uint16> dense, sub(G);
    std::vector<float>  qr(G), rr(G);
    dense_row(m.qry1, q, dense);
    for (uint32 i = 0; i < G; ++i)
        sub[i] = dense[genes[i]];
    rank_row(sub.data(), G, qr.data());

    std::vector<float> scores;
    for (auto label : top)
    {
        uint32             pos = m.ctidx[label * 2], len = m.ctidx[label * 2 + 1];
        std::vector<float> s(len);
        for (uint32 k = 0; k < len; ++k)
        {
            dense_row(m.ref1, pos + k, dense);
            for (uint32 i = 0; i < G; ++i)
                sub[i] = dense[genes[i]];
            rank_row(sub.data(), G, rr.data());
            s[k] = variant == VAR_REF ? corr_tree(qr.data(), rr.data(), G)
                                      : corr_exact(qr.data(), rr.data(), G);
        }
        scores.push_back(percentile(s, (int)len, quantile));
    }
    if (round_scores)
        *round_scores = scores;
    auto  ele = std::minmax_element(scores.begin(), scores.end());
    float th  = *ele.second - thre;
    if (std::isnan(*ele.second))
        return { top.front() };
    std::vector<uint32> res;
    for (uint32 i = 0; i < scores.size(); ++i)
    {
        if (scores[i] <= *ele.first || scores[i] < th)
            continue;
        if (scores[i] == *ele.second)
            res.insert(res.begin(), top[i]);
        else
            res.push_back(top[i]);
    }
    if (res.empty())
        res.push_back(top.front());
    return res;
}

}  // namespace

// ================================================================= C-ABI ===
extern "C" {

static std::vector<std::string> split_names(const char* blob, int n)
{
    std::vector<std::string> out;
    out.reserve(n);
    const char* p = blob;
    for (int i = 0; i < n; ++i)
    {
        out.emplace_back(p);
        p += out.back().size() + 1;
    }
    return out;
}

// Runs prepareData, findIntersectionGenes, loadRefData, trainData,
// loadQryData, preprocess and both generateDenseMatrix steps
// (src/pipeline.cpp:128-156, src/io.cpp:792-821).  Returns NULL on error; the
// message is available from or_last_error().
static std::string g_err;

void* or_train(int cores, uint32 ref_h, uint32 ref_w, const float* ref_data,
               const int* ref_indices, const int* ref_indptr, uint32 qry_h, uint32 qry_w,
               const float* qry_data, const int* qry_indices, const int* qry_indptr, int n_codes,
               const int* celltype, const char* ref_genes, int n_ref_genes, const char* qry_genes,
               int n_qry_genes)
{
    Model* m        = new Model();
    m->cores        = cores > 0 ? cores : 1;
    m->ref_height   = ref_h;
    m->ref_width    = ref_w;
    m->qry_height   = qry_h;
    m->qry_width    = qry_w;
    m->ref_indptr.assign(ref_indptr, ref_indptr + ref_h + 1);
    m->ref_indices.assign(ref_indices, ref_indices + ref_indptr[ref_h]);
    m->ref_data.assign(ref_data, ref_data + ref_indptr[ref_h]);
    m->qry_indptr.assign(qry_indptr, qry_indptr + qry_h + 1);
    m->qry_indices.assign(qry_indices, qry_indices + qry_indptr[qry_h]);
    m->qry_data.assign(qry_data, qry_data + qry_indptr[qry_h]);
    m->n_labels = (size_t)n_codes;
    // src/io.cpp:813: int codes are narrowed to uint8
    for (uint32 i = 0; i < ref_h; ++i)
        m->celltype_codes.push_back((uint8)celltype[i]);
    m->ref_genes = split_names(ref_genes, n_ref_genes);
    m->qry_genes = split_names(qry_genes, n_qry_genes);

    if (!find_intersection(*m) || !group_and_resort(*m))
    {
        g_err = m->error;
        delete m;
        return nullptr;
    }
    if (m->filter_genes)
    {
        remove_cols(m->ref_data, m->ref_indptr, m->ref_indices, m->ref_gene_index);
        remove_cols(m->qry_data, m->qry_indptr, m->qry_indices, m->qry_gene_index);
    }
    train(*m);
    generate(*m, 0);
    generate(*m, 1);
    return m;
}

const char* or_last_error() { return g_err.c_str(); }

void or_free(void* h) { delete (Model*)h; }

// Sizes: [n_labels, G_s, G_f, max_uniq0, max_uniq1, |train_values|,
//         |ctdiff1|, ref_nnz0, qry_nnz0, ref_nnz1, qry_nnz1, ref_height, qry_height]
void or_sizes(void* h, int64_t* out)
{
    Model* m = (Model*)h;
    out[0]   = (int64_t)m->n_labels;
    out[1]   = m->ref0.width;
    out[2]   = m->ref1.width;
    out[3]   = (int64_t)m->max_uniq0;
    out[4]   = (int64_t)m->max_uniq1;
    out[5]   = (int64_t)m->train_values.size();
    out[6]   = (int64_t)m->ctdiff1.size();
    out[7]   = (int64_t)m->ref0.data.size();
    out[8]   = (int64_t)m->qry0.data.size();
    out[9]   = (int64_t)m->ref1.data.size();
    out[10]  = (int64_t)m->qry1.data.size();
    out[11]  = m->ref_height;
    out[12]  = m->qry_height;
}

// Copies one array out.  which: 0 ctidx (u32 2L), 1 train_idxs (u32 2L^2),
// 2 train_values (u32), 3 common_genes (u32 G_s), 4 thre_genes (u32 G_f),
// 5 ctdidx1 (u32 2L^2), 6 ctdiff1 (u32),
// 10/11/12 ref0 data(u16)/indices(i32)/indptr(i32), 13/14/15 qry0,
// 20/21/22 ref1, 23/24/25 qry1.
void or_get(void* h, int which, void* out)
{
    Model* m    = (Model*)h;
    auto   cp32 = [&](const std::vector<uint32>& v) { memcpy(out, v.data(), v.size() * 4); };
    auto   cpset = [&](const std::set<uint32>& s) {
        uint32* o = (uint32*)out;
        for (auto v : s)
            *o++ = v;
    };
    auto csr = [&](const CodedCsr& a, int part) {
        if (part == 0)
            memcpy(out, a.data.data(), a.data.size() * 2);
        else if (part == 1)
            memcpy(out, a.indices.data(), a.indices.size() * 4);
        else
            memcpy(out, a.indptr.data(), a.indptr.size() * 4);
    };
    switch (which)
    {
    case 0: cp32(m->ctidx); break;
    case 1: cp32(m->train_idxs); break;
    case 2: cp32(m->train_values); break;
    case 3: cpset(m->common_genes); break;
    case 4: cpset(m->thre_genes); break;
    case 5: cp32(m->ctdidx1); break;
    case 6: cp32(m->ctdiff1); break;
    case 10: case 11: case 12: csr(m->ref0, which - 10); break;
    case 13: case 14: case 15: csr(m->qry0, which - 13); break;
    case 20: case 21: case 22: csr(m->ref1, which - 20); break;
    case 23: case 24: case 25: csr(m->qry1, which - 23); break;
    default: break;
    }
}

// libstdc++ std::minmax_element (first min, last max) as the reference calls it
// (src/cusingler.cu:1122, :1270).
void or_minmax(const float* a, int n, int* imin, int* imax)
{
    auto e = std::minmax_element(a, a + n);
    *imin  = (int)(e.first - a);
    *imax  = (int)(e.second - a);
}

// Rank a dense code row (A6).
void or_rank(const uint16* codes, uint32 n, float* out) { rank_row(codes, n, out); }

// Reference percentile (A8).
float or_percentile(const float* arr, int len, float p)
{
    return percentile(std::vector<float>(arr, arr + len), len, p);
}

// Correlation of two rank rows: variant 0 sequential FP32, 2 chunk+tree FP32,
// 1 exact (float), 3 FP64 value returned through *out64.
float or_corr(const float* q, const float* r, uint32 G, int variant, double* out64)
{
    if (out64)
        *out64 = corr_f64(q, r, G);
    if (variant == 0)
        return corr_seq(q, r, G);
    if (variant == 2)
        return corr_tree(q, r, G);
    return corr_exact(q, r, G);
}

// Score-step correlation matrix over G_s for query rows [q0, q1):
// out[(q - q0) * Nr + r] (reference-grouped ref row order).
void or_score_matrix(void* h, uint32 q0, uint32 q1, int variant, float* out, double* out64)
{
    Model*       m  = (Model*)h;
    const uint32 G  = m->ref0.width;
    const uint32 Nr = m->ref_height;
    std::vector<float> rr((size_t)Nr * G);
    {
        std::vector<uint16> d;
        for (uint32 r = 0; r < Nr; ++r)
        {
            dense_row(m->ref0, r, d);
            rank_row(d.data(), G, rr.data() + (size_t)r * G);
        }
    }
#pragma omp parallel for schedule(dynamic, 4) num_threads(m->cores)
    for (long q = q0; q < (long)q1; ++q)
    {
        std::vector<uint16> d;
        std::vector<float>  qr(G);
        dense_row(m->qry0, (uint32)q, d);
        rank_row(d.data(), G, qr.data());
        for (uint32 r = 0; r < Nr; ++r)
        {
            const float* y = rr.data() + (size_t)r * G;
            size_t       o = (size_t)(q - q0) * Nr + r;
            if (out)
                out[o] = variant == VAR_REF ? corr_seq(qr.data(), y, G) : corr_exact(qr.data(), y, G);
            if (out64)
                out64[o] = corr_f64(qr.data(), y, G);
        }
    }
}

// src/cusingler.cu:986-1136 (get_label) for query rows [q0, q1): per-label
// percentile scores (q1-q0) x L, first label (last maximum, libstdc++
// minmax_element), and the 0/1 candidate mask (score >= max - thre).
void or_score(void* h, uint32 q0, uint32 q1, float quantile, float thre, int variant,
              int32_t* first, float* mask, float* scores)
{
    Model*       m  = (Model*)h;
    const uint32 G  = m->ref0.width;
    const uint32 Nr = m->ref_height;
    const uint32 L  = (uint32)m->n_labels;
    std::vector<float> rr((size_t)Nr * G);
    {
        std::vector<uint16> d;
        for (uint32 r = 0; r < Nr; ++r)
        {
            dense_row(m->ref0, r, d);
            rank_row(d.data(), G, rr.data() + (size_t)r * G);
        }
    }
#pragma omp parallel for schedule(dynamic, 4) num_threads(m->cores)
    for (long q = q0; q < (long)q1; ++q)
    {
        std::vector<uint16> d;
        std::vector<float>  qr(G), row(Nr), ls(L);
        dense_row(m->qry0, (uint32)q, d);
        rank_row(d.data(), G, qr.data());
        for (uint32 r = 0; r < Nr; ++r)
        {
            const float* y = rr.data() + (size_t)r * G;
            row[r] = variant == VAR_REF ? corr_seq(qr.data(), y, G) : corr_exact(qr.data(), y, G);
        }
        for (uint32 l = 0; l < L; ++l)
        {
            uint32 pos = m->ctidx[2 * l], len = m->ctidx[2 * l + 1];
            ls[l] = percentile(std::vector<float>(row.begin() + pos, row.begin() + pos + len), (int)len,
                               quantile);
        }
        const size_t o   = (size_t)(q - q0);
        auto         ele = std::minmax_element(ls.begin(), ls.end());
        first[o]         = (int32_t)(ele.second - ls.begin());
        float th         = *ele.second - thre;
        for (uint32 l = 0; l < L; ++l)
        {
            mask[o * L + l]   = ls[l] >= th ? 1.f : 0.f;
            if (scores)
                scores[o * L + l] = ls[l];
        }
    }
}

// src/cusingler.cu:1297-1333 (cufinetune) for query rows [q0, q1) given the
// candidate mask rows of those cells ((q1-q0) x L).  rounds_out (optional)
// receives the number of rounds each cell ran.
void or_finetune(void* h, uint32 q0, uint32 q1, const float* mask, float quantile, float thre,
                 int times_limit, int variant, uint32* final_labels, int32_t* rounds_out)
{
    Model*       m = (Model*)h;
    const uint32 L = (uint32)m->n_labels;
#pragma omp parallel for schedule(dynamic, 1) num_threads(m->cores)
    for (long q = q0; q < (long)q1; ++q)
    {
        const size_t        o = (size_t)(q - q0);
        std::vector<uint32> top;
        for (uint32 l = 0; l < L; ++l)
            if (mask[o * L + l] != 0)
                top.push_back(l);
        int times = 0, rounds = 0;
        while (top.size() > 1 && (times_limit == 0 || times++ < times_limit))
        {
            top = finetune_round(*m, (uint32)q, top, quantile, thre, variant, nullptr);
            rounds++;
        }
        final_labels[o] = top.empty() ? 0 : top.front();
        if (rounds_out)
            rounds_out[o] = rounds;
    }
}

// One fine-tune round for one cell (for step-by-step checks): writes the new
// candidate list (returns its length) and the per-label scores of the round.
int or_finetune_round(void* h, uint32 q, const uint32* top, int n_top, float quantile, float thre,
                      int variant, uint32* out_top, float* out_scores)
{
    Model*              m = (Model*)h;
    std::vector<float>  sc;
    std::vector<uint32> t(top, top + n_top);
    auto                res = finetune_round(*m, q, t, quantile, thre, variant, &sc);
    for (size_t i = 0; i < res.size(); ++i)
        out_top[i] = res[i];
    for (size_t i = 0; i < sc.size(); ++i)
        out_scores[i] = sc[i];
    return (int)res.size();
}

}  // extern "C"
