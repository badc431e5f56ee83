// See csg_pipeline.h.
#include "csg_pipeline.h"

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>

namespace csg {

namespace {
double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

Pipeline::Pipeline(int cores_, int gpuid_) : cores(cores_), gpuid(gpuid_) {}

Pipeline::~Pipeline()
{
    if (eng_)
        csg_release(eng_);
    if (prep_)
        csg_prep_release(prep_);
}

void Pipeline::check(int rc, const char* what) const
{
    if (rc == CSG_OK)
        return;
    std::string msg = std::string(what) + " failed (status " + std::to_string(rc) + ")";
    if (eng_ && csg_last_error(eng_)[0])
        msg += ": " + std::string(csg_last_error(eng_));
    throw std::runtime_error(msg);
}

bool Pipeline::train(uint32_t ref_height, uint32_t ref_width, const std::vector<float>& ref_data,
                     const std::vector<int>& ref_indices, const std::vector<int>& ref_indptr,
                     uint32_t qry_height, uint32_t qry_width, const std::vector<float>& qry_data,
                     const std::vector<int>& qry_indices, const std::vector<int>& qry_indptr,
                     const std::vector<std::string>& codes, const std::vector<int>& celltypes_,
                     const std::vector<std::string>& cellnames,
                     const std::vector<std::string>& ref_geneidx,
                     const std::vector<std::string>& qry_geneidx)
{
    // initGPU (src/pipeline.cpp:131-136)
    if (!eng_)
    {
        int rc = csg_init_gpu(gpuid, &eng_);
        if (rc != CSG_OK)
            throw std::runtime_error("Fail to initlize gpu with id: " + std::to_string(gpuid));
    }
    if (ref_indptr.size() != (size_t)ref_height + 1 || qry_indptr.size() != (size_t)qry_height + 1)
        throw std::runtime_error("indptr length must be height + 1");
    if ((size_t)ref_indptr.back() > ref_data.size() || (size_t)ref_indptr.back() > ref_indices.size() ||
        (size_t)qry_indptr.back() > qry_data.size() || (size_t)qry_indptr.back() > qry_indices.size())
        throw std::runtime_error("indptr points past data / indices");
    if (celltypes_.size() != ref_height)
        throw std::runtime_error("celltype must hold one code per reference cell");
    if (cellnames.size() != qry_height)
        throw std::runtime_error("cellnames must hold one name per query cell");
    const double t0 = now_s();
    std::vector<const char*> rg, qg;
    for (auto& s : ref_geneidx)
        rg.push_back(s.c_str());
    for (auto& s : qry_geneidx)
        qg.push_back(s.c_str());
    char err[512] = { 0 };
    if (prep_)
    {
        csg_prep_release(prep_);
        prep_ = nullptr;
    }
    int rc = csg_prep_train(cores, ref_height, ref_width, ref_data.data(), ref_indices.data(),
                            ref_indptr.data(), qry_height, qry_width, qry_data.data(),
                            qry_indices.data(), qry_indptr.data(), (uint32_t)codes.size(),
                            celltypes_.data(), rg.data(), (uint32_t)rg.size(), qg.data(),
                            (uint32_t)qg.size(), &prep_, err, sizeof(err));
    if (rc != CSG_OK)
        throw std::runtime_error(std::string("train failed: ") + err);
    celltypes = codes;
    cells     = cellnames;
    t_train   = now_s() - t0;
    if (verbose)
        std::cout << "train data cost time(s): " << t_train << std::endl;
    return true;
}

bool Pipeline::score(float quantile, float finetune_thre, int path)
{
    if (!prep_)
        throw std::runtime_error("score before train");
    const double t0 = now_s();
    csg_coded    c{};
    if (csg_prep_generate(prep_, 0, &c) != CSG_OK)
        throw std::runtime_error("generateDenseMatrix(0) failed");
    check(csg_copyin_score(eng_, c.ref_height, c.width, c.ref_data, c.ref_indices, c.ref_indptr,
                           c.qry_height, c.qry_data, c.qry_indices, c.qry_indptr, c.ct_num,
                           c.ctidx, c.max_uniq_gene),
          "copyin_score");
    first_idx.assign(c.qry_height, 0);
    scores.assign((size_t)c.qry_height * c.ct_num, 0.f);
    check(csg_get_label(eng_, quantile, finetune_thre, path, first_idx.data(), nullptr,
                        scores.data()),
          "get_label");
    first_labels.clear();
    for (auto i : first_idx)
        first_labels.push_back(celltypes.at((size_t)i));
    check(csg_destroy_score(eng_), "destroy_score");
    t_score = now_s() - t0;
    if (verbose)
        std::cout << "score data cost time(s): " << t_score << std::endl;
    return true;
}

bool Pipeline::finetune(float quantile, float finetune_thre, int finetune_times)
{
    if (!prep_)
        throw std::runtime_error("finetune before train");
    const double t0 = now_s();
    csg_coded    c{};
    if (csg_prep_generate(prep_, 1, &c) != CSG_OK)
        throw std::runtime_error("generateDenseMatrix(1) failed");
    check(csg_copyin(eng_, c.ref_height, c.width, c.ref_data, c.ref_indices, c.ref_indptr,
                     c.qry_height, c.qry_data, c.qry_indices, c.qry_indptr, c.ct_num, c.ctidx,
                     c.ctdidx, c.ctdiff, c.n_ctdiff, c.max_uniq_gene, nullptr),
          "copyin");
    final_idx.assign(c.qry_height, 0);
    check(csg_finetune(eng_, quantile, finetune_thre, finetune_times, final_idx.data(), nullptr),
          "finetune");
    final_labels.clear();
    for (auto i : final_idx)
        final_labels.push_back(celltypes.at(i));
    check(csg_destroy(eng_), "destroy");
    t_finetune = now_s() - t0;
    if (verbose)
        std::cout << "finetune cost time(s): " << t_finetune << std::endl;
    return true;
}

std::vector<std::vector<std::string>> Pipeline::dump() const
{
    std::vector<std::vector<std::string>> res;
    for (size_t i = 0; i < final_labels.size(); ++i)
        res.push_back({ cells[i], first_labels[i], final_labels[i] });
    return res;
}

bool Pipeline::dump(const std::string& stat_file) const
{
    std::ofstream ofs(stat_file);
    ofs << "cell\tfirstLabel\tfinalLabel\n";
    for (size_t i = 0; i < final_labels.size(); ++i)
        ofs << cells[i] << "\t" << first_labels[i] << "\t" << final_labels[i] << "\n";
    return true;
}

}  // namespace csg
