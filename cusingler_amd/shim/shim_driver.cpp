// The reference's pipeline order (src/pipeline.cpp:47-111, 128-156) run through
// the shim of its device API: train (this library's host preprocessing, the
// reference's DataParser restated), generateDenseMatrix(0) -> copyin_score ->
// get_label -> destroy_score, generateDenseMatrix(1) -> copyin -> cufinetune
// -> destroy.  Reads a dataset directory, writes first.i32 and final.u32.
//
//   shim_driver <dataset dir> <output dir> [path]
//
// Dataset files: meta.txt "ref_height ref_width qry_height qry_width n_codes
// cores", ref_/qry_ data.f32, indices.i32, indptr.i32, celltype.i32,
// ref_genes.txt, qry_genes.txt (one name per line).

#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cusingler_amd.h"
#include "../../include/cusingler_prep.h"
#include "cusingler_shim.h"

namespace {

template <class T> std::vector<T> read_bin(const std::string& path)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f)
        throw std::runtime_error("cannot read " + path);
    const std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<T> v((size_t)n / sizeof(T));
    f.read((char*)v.data(), n);
    return v;
}

std::vector<std::string> read_lines(const std::string& path)
{
    std::ifstream            f(path);
    std::vector<std::string> out;
    std::string              s;
    while (std::getline(f, s))
        out.push_back(s);
    return out;
}

template <class T> void write_bin(const std::string& path, const std::vector<T>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write((const char*)v.data(), (std::streamsize)(v.size() * sizeof(T)));
}

// generateDenseMatrix(step)'s outputs into InputData (the reference's
// DataParser fills the same fields, src/io.cpp:714-784)
void fill(InputData& d, const csg_coded& c)
{
    d.ref_height = c.ref_height;
    d.ref_width  = c.width;
    d.qry_height = c.qry_height;
    d.qry_width  = c.width;
    d.ct_num     = (int)c.ct_num;
    d.ref_data.assign(c.ref_data, c.ref_data + c.ref_indptr[c.ref_height]);
    d.ref_indices.assign(c.ref_indices, c.ref_indices + c.ref_indptr[c.ref_height]);
    d.ref_indptr.assign(c.ref_indptr, c.ref_indptr + c.ref_height + 1);
    d.qry_data.assign(c.qry_data, c.qry_data + c.qry_indptr[c.qry_height]);
    d.qry_indices.assign(c.qry_indices, c.qry_indices + c.qry_indptr[c.qry_height]);
    d.qry_indptr.assign(c.qry_indptr, c.qry_indptr + c.qry_height + 1);
    d.ctidx.assign(c.ctidx, c.ctidx + 2 * c.ct_num);
    d.ctdidx.assign(c.ctdidx, c.ctdidx + 2 * (size_t)c.ct_num * c.ct_num);
    d.ctdiff.assign(c.ctdiff, c.ctdiff + c.n_ctdiff);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3)
    {
        std::fprintf(stderr, "usage: shim_driver <dataset dir> <output dir> [path]\n");
        return 2;
    }
    const std::string in = argv[1], out = argv[2];
    unsigned rh, rw, qh, qw, nc;
    int      cores;
    {
        std::ifstream m(in + "/meta.txt");
        m >> rh >> rw >> qh >> qw >> nc >> cores;
    }
    auto rd = read_bin<float>(in + "/ref_data.f32");
    auto ri = read_bin<int32_t>(in + "/ref_indices.i32");
    auto rp = read_bin<int32_t>(in + "/ref_indptr.i32");
    auto qd = read_bin<float>(in + "/qry_data.f32");
    auto qi = read_bin<int32_t>(in + "/qry_indices.i32");
    auto qp = read_bin<int32_t>(in + "/qry_indptr.i32");
    auto ct = read_bin<int32_t>(in + "/celltype.i32");
    auto rg = read_lines(in + "/ref_genes.txt");
    auto qg = read_lines(in + "/qry_genes.txt");
    std::vector<const char*> rgp, qgp;
    for (auto& s : rg)
        rgp.push_back(s.c_str());
    for (auto& s : qg)
        qgp.push_back(s.c_str());
    csg_prep* prep = nullptr;
    char      err[512];
    if (csg_prep_train(cores, rh, rw, rd.data(), ri.data(), rp.data(), qh, qw, qd.data(),
                       qi.data(), qp.data(), nc, ct.data(), rgp.data(), (uint32_t)rgp.size(),
                       qgp.data(), (uint32_t)qgp.size(), &prep, err, sizeof(err)) != 0)
    {
        std::fprintf(stderr, "train: %s\n", err);
        return 1;
    }
    try
    {
        if (!initGPU(0))
            return 1;
        if (argc > 3)
            csg_shim_set_path(std::stoi(argv[3]));
        InputData d;
        csg_coded c;
        // score step (Pipeline::score)
        if (csg_prep_generate(prep, 0, &c) != 0)
            return 1;
        fill(d, c);
        const uint64 mu0 = c.max_uniq_gene;
        if (!copyin_score(d))
            return 1;
        std::vector<int> first = get_label(d, mu0, cores, 0.8f, 0.05f);
        if (!destroy_score())
            return 1;
        // fine-tune (Pipeline::finetune); InputData::labels holds the mask
        if (csg_prep_generate(prep, 1, &c) != 0)
            return 1;
        std::vector<float> mask = d.labels;
        fill(d, c);
        d.labels = mask;
        const uint64 mu1 = c.max_uniq_gene;
        if (!copyin(d))
            return 1;
        std::vector<uint32> final_labels = cufinetune(mu1, 0.8f, 0.05f, 0);
        if (!destroy())
            return 1;
        write_bin(out + "/first.i32", first);
        write_bin(out + "/final.u32", final_labels);
    }
    catch (const std::exception& e)
    {
        std::fprintf(stderr, "%s\n", e.what());
        csg_prep_release(prep);
        return 1;
    }
    csg_prep_release(prep);
    return 0;
}
