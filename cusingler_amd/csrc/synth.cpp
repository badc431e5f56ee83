// Seeded synthetic single-cell count sampler (host C++, OpenMP).
//
// Samples counts ~ NegativeBinomial(mean mu, size r) for every (cell, gene) of a
// cell x gene matrix and stores log2(1 + count) for counts >= 1 as float32 CSR
// (data / int32 indices / int64 row offsets), the layout the annotation entry
// point takes for its reference and query matrices.
//
// mu(cell, gene) = mu_label[label(cell), gene] * lib(cell).  Each entry draws one
// counter-based uniform u = hash(seed, stream, row, gene) and inverts the NB CDF
// from k = 0 upward, so the result depends only on (seed, stream, row, gene) and
// never on the thread count.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline double uniform01(uint64_t seed, uint64_t stream, uint64_t row, uint64_t gene)
{
    uint64_t h = splitmix64(seed ^ splitmix64(stream * 0x632BE59BD9B4E019ull + row));
    h          = splitmix64(h ^ (gene * 0xD6E8FEB86659FD93ull));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
}

// Inverse-CDF NB sample with mean mu and size r (variance mu + mu^2 / r).
inline int nb_sample(double mu, double r, double u)
{
    if (mu <= 0.0)
        return 0;
    const double q  = mu / (r + mu);
    double       pk = std::exp(r * std::log(r / (r + mu)));  // P(0)
    double       cdf = pk;
    if (u < cdf)
        return 0;
    int k = 0;
    while (u >= cdf && k < 1000000)
    {
        ++k;
        pk *= (k - 1 + r) / k * q;
        cdf += pk;
        if (pk < 1e-300)
            break;
    }
    return k;
}

}  // namespace

extern "C" {

// Pass 1+2 in one: per-row vectors filled in parallel, then concatenated.
// mu_label: [n_labels x n_genes] row-major float.  labels: [n_cells] int32.
// lib: [n_cells] float.  Returns nnz; *out_data / *out_indices are malloc'ed
// and must be released with csg_synth_free.  out_indptr has n_cells + 1 slots.
int64_t csg_synth_sample(uint64_t seed, uint64_t stream, int64_t n_cells, int32_t n_genes,
                         const float* mu_label, const int32_t* labels, const float* lib,
                         double nb_size, int32_t force_min_nnz, int64_t* out_indptr,
                         float** out_data, int32_t** out_indices)
{
    std::vector<std::vector<int32_t>> idx(n_cells);
    std::vector<std::vector<float>>   val(n_cells);

#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < n_cells; ++c)
    {
        const float* mu_row = mu_label + (size_t)labels[c] * n_genes;
        const double l      = lib[c];
        auto&        vi     = idx[c];
        auto&        vv     = val[c];
        vi.reserve(n_genes / 8 + 8);
        vv.reserve(n_genes / 8 + 8);
        for (int32_t g = 0; g < n_genes; ++g)
        {
            double mu = (double)mu_row[g] * l;
            double u  = uniform01(seed, stream, (uint64_t)c, (uint64_t)g);
            int    k  = nb_sample(mu, nb_size, u);
            if (k > 0)
            {
                vi.push_back(g);
                vv.push_back((float)std::log2(1.0 + (double)k));
            }
        }
        // Guarantee a minimum number of expressed genes: take the genes with
        // the largest label mean that are still zero and give them count 1.
        if ((int)vi.size() < force_min_nnz)
        {
            std::vector<char> taken(n_genes, 0);
            for (auto g : vi)
                taken[g] = 1;
            while ((int)vi.size() < force_min_nnz)
            {
                int   best = -1;
                float bm   = -1.f;
                for (int32_t g = 0; g < n_genes; ++g)
                    if (!taken[g] && mu_row[g] > bm)
                    {
                        bm   = mu_row[g];
                        best = g;
                    }
                if (best < 0)
                    break;
                taken[best] = 1;
                // keep indices sorted
                size_t pos = 0;
                while (pos < vi.size() && vi[pos] < best)
                    ++pos;
                vi.insert(vi.begin() + pos, best);
                vv.insert(vv.begin() + pos, 1.0f);  // log2(1 + 1)
            }
        }
    }

    out_indptr[0] = 0;
    for (int64_t c = 0; c < n_cells; ++c)
        out_indptr[c + 1] = out_indptr[c] + (int64_t)idx[c].size();
    const int64_t nnz = out_indptr[n_cells];
    float*   data     = (float*)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    int32_t* indices  = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
    {
        const int64_t o = out_indptr[c];
        if (!idx[c].empty())
        {
            std::memcpy(indices + o, idx[c].data(), sizeof(int32_t) * idx[c].size());
            std::memcpy(data + o, val[c].data(), sizeof(float) * val[c].size());
        }
    }
    *out_data    = data;
    *out_indices = indices;
    return nnz;
}

void csg_synth_free(void* p) { free(p); }

// Expected nonzero fraction for a given scale s (used to calibrate s):
// mean over (cell sample, gene) of 1 - P(0) with mean s * mu_label[label, g] * lib.
double csg_synth_nonzero_fraction(double s, int64_t n_cells, int32_t n_genes,
                                  const float* mu_label, const int32_t* labels,
                                  const float* lib, double nb_size)
{
    double acc = 0.0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (int64_t c = 0; c < n_cells; ++c)
    {
        const float* mu_row = mu_label + (size_t)labels[c] * n_genes;
        double       a      = 0.0;
        for (int32_t g = 0; g < n_genes; ++g)
        {
            double mu = s * (double)mu_row[g] * lib[c];
            a += 1.0 - std::exp(nb_size * std::log(nb_size / (nb_size + mu)));
        }
        acc += a;
    }
    return acc / ((double)n_cells * (double)n_genes);
}

}  // extern "C"
