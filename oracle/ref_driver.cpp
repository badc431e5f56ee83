// Driver for the reference's own host preprocessing (src/io.cpp), built from
// the reference sources where they lie (see oracle/Makefile) into
// oracle/_ref/ref_driver.  TEST INFRASTRUCTURE ONLY: it exists to pin the
// oracle's restatement of src/io.cpp and is never part of the product.
//
// Usage: ref_driver <in_dir> <out_dir>
//   in_dir : meta.txt ("ref_h ref_w qry_h qry_w n_codes cores"),
//            ref_data.f32 ref_indices.i32 ref_indptr.i32,
//            qry_data.f32 qry_indices.i32 qry_indptr.i32, celltype.i32,
//            ref_genes.txt qry_genes.txt codes.txt cellnames.txt (one per line)
//   out_dir: step0_* and step1_* coded CSR arrays, gene sets, pair windows.
//
// The sequence of calls is src/pipeline.cpp:144-152 followed by
// generateDenseMatrix(0) (src/pipeline.cpp:54) and generateDenseMatrix(1)
// (src/pipeline.cpp:94).

#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "io.h"

template <typename T> static std::vector<T> read_bin(const std::string& p)
{
    std::ifstream f(p, std::ios::binary | std::ios::ate);
    size_t        n = (size_t)f.tellg();
    f.seekg(0);
    std::vector<T> v(n / sizeof(T));
    f.read((char*)v.data(), (std::streamsize)n);
    return v;
}

static std::vector<std::string> read_lines(const std::string& p)
{
    std::ifstream            f(p);
    std::vector<std::string> v;
    std::string              s;
    while (std::getline(f, s))
        v.push_back(s);
    return v;
}

template <typename T> static void write_bin(const std::string& p, const T* d, size_t n)
{
    std::ofstream f(p, std::ios::binary);
    f.write((const char*)d, (std::streamsize)(n * sizeof(T)));
}

template <typename T> static void write_vec(const std::string& p, const std::vector<T>& v)
{
    write_bin(p, v.data(), v.size());
}

static void write_set(const std::string& p, const std::set<uint32>& s)
{
    std::vector<uint32> v(s.begin(), s.end());
    write_vec(p, v);
}

int main(int argc, char** argv)
{
    if (argc != 3)
    {
        std::cerr << "usage: ref_driver <in_dir> <out_dir>\n";
        return 2;
    }
    std::string in = argv[1], out = argv[2];
    uint32      ref_h, ref_w, qry_h, qry_w;
    int         n_codes, cores;
    {
        std::ifstream m(in + "/meta.txt");
        m >> ref_h >> ref_w >> qry_h >> qry_w >> n_codes >> cores;
    }
    auto ref_data    = read_bin<float>(in + "/ref_data.f32");
    auto ref_indices = read_bin<int>(in + "/ref_indices.i32");
    auto ref_indptr  = read_bin<int>(in + "/ref_indptr.i32");
    auto qry_data    = read_bin<float>(in + "/qry_data.f32");
    auto qry_indices = read_bin<int>(in + "/qry_indices.i32");
    auto qry_indptr  = read_bin<int>(in + "/qry_indptr.i32");
    auto celltype    = read_bin<int>(in + "/celltype.i32");
    auto ref_genes   = read_lines(in + "/ref_genes.txt");
    auto qry_genes   = read_lines(in + "/qry_genes.txt");
    auto codes       = read_lines(in + "/codes.txt");
    auto cellnames   = read_lines(in + "/cellnames.txt");
    codes.resize(n_codes);

    PyDataParser p(cores);
    p.prepareData(ref_h, ref_w, ref_data, ref_indices, ref_indptr, qry_h, qry_w, qry_data,
                  qry_indices, qry_indptr, codes, celltype, cellnames, ref_genes, qry_genes);
    p.findIntersectionGenes();
    p.loadRefData();
    p.trainData();
    p.loadQryData();
    p.preprocess();

    uint64 mu0 = 0;
    p.generateDenseMatrix(0, mu0);
    auto& rd = p.raw_data;
    write_vec(out + "/step0_ref_data.u16", rd.ref_data);
    write_vec(out + "/step0_ref_indices.i32", rd.ref_indices);
    write_vec(out + "/step0_ref_indptr.i32", rd.ref_indptr);
    write_vec(out + "/step0_qry_data.u16", rd.qry_data);
    write_vec(out + "/step0_qry_indices.i32", rd.qry_indices);
    write_vec(out + "/step0_qry_indptr.i32", rd.qry_indptr);
    write_vec(out + "/ctidx.u32", rd.ctidx);
    write_vec(out + "/train_idxs.u32", p.ref_train_idxs);
    write_vec(out + "/train_values.u32", p.ref_train_values);
    write_set(out + "/common_genes.u32", p.common_genes);
    write_set(out + "/thre_genes.u32", p.thre_genes);

    uint64 mu1 = 0;
    p.generateDenseMatrix(1, mu1);
    write_vec(out + "/step1_ref_data.u16", rd.ref_data);
    write_vec(out + "/step1_ref_indices.i32", rd.ref_indices);
    write_vec(out + "/step1_ref_indptr.i32", rd.ref_indptr);
    write_vec(out + "/step1_qry_data.u16", rd.qry_data);
    write_vec(out + "/step1_qry_indices.i32", rd.qry_indices);
    write_vec(out + "/step1_qry_indptr.i32", rd.qry_indptr);
    write_vec(out + "/ctdidx1.u32", rd.ctdidx);
    write_vec(out + "/ctdiff1.u32", rd.ctdiff);
    {
        std::ofstream m(out + "/meta.txt");
        m << mu0 << " " << mu1 << " " << rd.ct_num << "\n";
    }
    return 0;
}
