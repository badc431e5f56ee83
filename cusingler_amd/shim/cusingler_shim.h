// The reference's device API (src/cusingler.cuh:13-22) implemented over this
// library's C-ABI (include/cusingler_amd.h): a drop-in for the CUDA engine of
// the reference's src/pipeline.cpp.
//
// Inside the reference's source tree compile with -DCSG_SHIM_REFERENCE_TYPES
// so that InputData comes from its own types.h; standalone (tests, other
// hosts) the definition below is used: the same fields with the same names
// and types (src/types.h:21-42), the interface the API passes.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#ifdef CSG_SHIM_REFERENCE_TYPES
#include "types.h"
#else
typedef unsigned char  uint8;
typedef unsigned short uint16;
typedef unsigned int   uint32;
typedef unsigned long  uint64;

struct InputData
{
    std::vector<uint16>      ref;
    std::vector<uint16>      ref_data;
    std::vector<int>         ref_indptr, ref_indices;
    uint32                   ref_height = 0, ref_width = 0;
    std::vector<uint16>      qry;
    std::vector<uint16>      qry_data;
    std::vector<int>         qry_indptr, qry_indices;
    uint32                   qry_height = 0, qry_width = 0;
    std::vector<std::string> qry_cellnames;
    std::vector<float>       labels;     // the candidate mask, cell x label
    std::vector<std::string> celltypes;
    int                      ct_num = 0;
    std::vector<uint32>      ctdiff;
    std::vector<uint32>      ctdidx;
    std::vector<uint32>      ctidx;
};
#endif

struct csg_engine;

bool                initGPU(const int gpuid);
bool                copyin_score(InputData& rawdata);
bool                destroy_score();
std::vector<int>    get_label(InputData& rawdata, const uint64 max_uniq_gene, int cores,
                              float quantile, float finetune_thre);
bool                copyin(InputData& rawdata);
bool                destroy();
std::vector<uint32> cufinetune(const uint64 max_uniq_gene, float quantile, float finetune_thre,
                               int finetune_times);

// The engine behind the functions above (after initGPU), for the options the
// reference does not have: the score path (csg_get_label's `path`, plain by
// default), csg_set_finetune_path, the learned proposal.
csg_engine* csg_shim_engine();
void        csg_shim_set_path(int path);
