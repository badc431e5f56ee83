// The reference's device API over the C-ABI (see cusingler_shim.h).
//
// The reference hands max_uniq_gene to get_label and cufinetune, while this
// library's copy-in calls take it (it sizes the rank kernels' code bins).  So
// copyin_score and copyin keep a reference to the caller's InputData and the
// upload happens in get_label / cufinetune, with the value those receive --
// the same data, uploaded once, at the point the reference first needs it.
// Failures print the library's message to stderr: the bool functions return
// false and the vector functions throw std::runtime_error (the reference
// calls exit(), src/cusingler.cu:48-57).

#include "cusingler_shim.h"

#include <cstdio>
#include <stdexcept>

#include "../../include/cusingler_amd.h"

namespace {

csg_engine* g_eng   = nullptr;
InputData*  g_score = nullptr;  // copyin_score's data, uploaded by get_label
InputData*  g_ft    = nullptr;  // copyin's data, uploaded by cufinetune
int         g_path  = CSG_PATH_PLAIN;

bool ok(int rc, const char* what)
{
    if (rc == CSG_OK)
        return true;
    std::fprintf(stderr, "%s: %s\n", what, g_eng ? csg_last_error(g_eng) : "no engine");
    return false;
}

}  // namespace

bool initGPU(const int gpuid)
{
    if (g_eng)
    {
        csg_release(g_eng);
        g_eng = nullptr;
    }
    return ok(csg_init_gpu(gpuid, &g_eng), "initGPU");
}

bool copyin_score(InputData& rawdata)
{
    if (!g_eng)
        return ok(CSG_ERR_STATE, "copyin_score before initGPU");
    g_score = &rawdata;
    return true;
}

std::vector<int> get_label(InputData& d, const uint64 max_uniq_gene, int cores, float quantile,
                           float finetune_thre)
{
    (void)cores;  // the reference sizes its host percentile threads with it
    if (!g_eng || g_score != &d)
        throw std::runtime_error("get_label without copyin_score of the same data");
    if (!ok(csg_copyin_score(g_eng, d.ref_height, d.ref_width, d.ref_data.data(),
                             d.ref_indices.data(), d.ref_indptr.data(), d.qry_height,
                             d.qry_data.data(), d.qry_indices.data(), d.qry_indptr.data(),
                             (uint32_t)d.ct_num, d.ctidx.data(), max_uniq_gene),
            "copyin_score"))
        throw std::runtime_error("copyin_score failed");
    std::vector<int> first(d.qry_height);
    d.labels.assign((size_t)d.qry_height * (size_t)d.ct_num, 0.f);
    if (!ok(csg_get_label(g_eng, quantile, finetune_thre, g_path, first.data(), d.labels.data(),
                          nullptr),
            "get_label"))
        throw std::runtime_error("get_label failed");
    return first;
}

bool destroy_score()
{
    g_score = nullptr;
    return g_eng && ok(csg_destroy_score(g_eng), "destroy_score");
}

bool copyin(InputData& rawdata)
{
    if (!g_eng)
        return ok(CSG_ERR_STATE, "copyin before initGPU");
    g_ft = &rawdata;
    return true;
}

std::vector<uint32> cufinetune(const uint64 max_uniq_gene, float quantile, float finetune_thre,
                               int finetune_times)
{
    if (!g_eng || !g_ft)
        throw std::runtime_error("cufinetune without copyin");
    InputData& d = *g_ft;
    // the candidate mask get_label wrote into InputData::labels
    const float* mask = d.labels.size() == (size_t)d.qry_height * (size_t)d.ct_num
                            ? d.labels.data()
                            : nullptr;
    if (!ok(csg_copyin(g_eng, d.ref_height, d.ref_width, d.ref_data.data(), d.ref_indices.data(),
                       d.ref_indptr.data(), d.qry_height, d.qry_data.data(),
                       d.qry_indices.data(), d.qry_indptr.data(), (uint32_t)d.ct_num,
                       d.ctidx.data(), d.ctdidx.data(), d.ctdiff.data(), d.ctdiff.size(),
                       max_uniq_gene, mask),
            "copyin"))
        throw std::runtime_error("copyin failed");
    std::vector<uint32> final_labels(d.qry_height);
    if (!ok(csg_finetune(g_eng, quantile, finetune_thre, finetune_times, final_labels.data(),
                         nullptr),
            "cufinetune"))
        throw std::runtime_error("cufinetune failed");
    return final_labels;
}

bool destroy()
{
    g_ft = nullptr;
    return g_eng && ok(csg_destroy(g_eng), "destroy");
}

csg_engine* csg_shim_engine() { return g_eng; }

void csg_shim_set_path(int path) { g_path = path; }
