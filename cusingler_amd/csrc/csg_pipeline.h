// The annotation pipeline over the host preprocessing and the device engine:
// the counterpart of the reference's Pipeline / PyPipeline
// (src/pipeline.h:17-58, src/pipeline.cpp:19-164).  Same steps and meaning:
// train -> score -> finetune -> dump; failures throw std::runtime_error with
// the engine's message instead of calling exit().
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/cusingler_amd.h"
#include "../../include/cusingler_prep.h"

namespace csg {

class Pipeline
{
public:
    Pipeline(int cores, int gpuid);
    ~Pipeline();
    Pipeline(const Pipeline&) = delete;
    Pipeline& operator=(const Pipeline&) = delete;

    // PyPipeline::train (src/pipeline.cpp:128-156)
    bool train(uint32_t ref_height, uint32_t ref_width, const std::vector<float>& ref_data,
               const std::vector<int>& ref_indices, const std::vector<int>& ref_indptr,
               uint32_t qry_height, uint32_t qry_width, const std::vector<float>& qry_data,
               const std::vector<int>& qry_indices, const std::vector<int>& qry_indptr,
               const std::vector<std::string>& codes, const std::vector<int>& celltypes,
               const std::vector<std::string>& cellnames,
               const std::vector<std::string>& ref_geneidx,
               const std::vector<std::string>& qry_geneidx);
    // Pipeline::score (src/pipeline.cpp:47-86); `path` picks the correlation form
    bool score(float quantile, float finetune_thre, int path = CSG_PATH_PLAIN);
    // Pipeline::finetune (src/pipeline.cpp:88-111)
    bool finetune(float quantile, float finetune_thre, int finetune_times);
    // PyPipeline::dump (src/pipeline.cpp:158-164)
    std::vector<std::vector<std::string>> dump() const;
    // Pipeline::dump (src/pipeline.cpp:113-121): TSV with a header line
    bool dump(const std::string& stat_file) const;

    int                      cores;
    int                      gpuid;
    std::vector<std::string> cells, first_labels, final_labels;
    std::vector<std::string> celltypes;
    std::vector<float>       scores;   // query x label quantile scores of the score step
    std::vector<int32_t>     first_idx;
    std::vector<uint32_t>    final_idx;
    double                   t_train = 0, t_score = 0, t_finetune = 0;
    bool                     verbose = true;

private:
    csg_prep*   prep_ = nullptr;
    csg_engine* eng_  = nullptr;
    void        check(int rc, const char* what) const;
};

}  // namespace csg
