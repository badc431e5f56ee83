// pybind11 module `cusingler`: the reference's Python entry point
// (src/main.cpp:32-95) over the MI355X engine.  `run` keeps the reference's
// signature and defaults; `quantile` is in percent and is divided by 100 in
// float (src/main.cpp:47-48).  Extra keyword arguments (defaults reproduce
// the reference): `path` selects the score-step correlation form and
// `return_scores` appends the query x label quantile scores.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <string>
#include <vector>

#include "csg_pipeline.h"

namespace py = pybind11;
using namespace pybind11::literals;

namespace {

int path_code(const std::string& p)
{
    if (p == "plain")
        return CSG_PATH_PLAIN;
    if (p == "exact_int8")
        return CSG_PATH_EXACT_INT8;
    if (p == "fp32_mfma")
        return CSG_PATH_FP32_MFMA;
    if (p == "fp16")
        return CSG_PATH_FP16_MFMA;
    throw std::invalid_argument("path must be one of plain, exact_int8, fp32_mfma, fp16");
}

py::object run(int cores, int gpuid, uint32_t ref_height, uint32_t ref_width,
               std::vector<float>& ref_data, std::vector<int>& ref_indices,
               std::vector<int>& ref_indptr, uint32_t qry_height, uint32_t qry_width,
               std::vector<float>& qry_data, std::vector<int>& qry_indices,
               std::vector<int>& qry_indptr, std::vector<std::string>& codes,
               std::vector<int>& celltypes, std::vector<std::string>& cellnames,
               std::vector<std::string>& ref_geneidx, std::vector<std::string>& qry_geneidx,
               float quantile, float finetune_thre, int finetune_times, const std::string& path,
               bool return_scores, bool verbose)
{
    const int          pc = path_code(path);
    csg::Pipeline      pipeline(cores, gpuid);
    pipeline.verbose = verbose;
    {
        py::gil_scoped_release nogil;
        pipeline.train(ref_height, ref_width, ref_data, ref_indices, ref_indptr, qry_height,
                       qry_width, qry_data, qry_indices, qry_indptr, codes, celltypes, cellnames,
                       ref_geneidx, qry_geneidx);
        pipeline.score(quantile / 100, finetune_thre, pc);
        pipeline.finetune(quantile / 100, finetune_thre, finetune_times);
    }
    auto res = pipeline.dump();
    if (!return_scores)
        return py::cast(res);
    py::array_t<float> sc({ (py::ssize_t)qry_height, (py::ssize_t)codes.size() });
    std::copy(pipeline.scores.begin(), pipeline.scores.end(), sc.mutable_data());
    return py::make_tuple(res, sc);
}

}  // namespace

PYBIND11_MODULE(cusingler, m)
{
    m.doc() = "cell annotation accelerated by GPU (MI355X)";
    m.def("run", &run,
          "The main entry of cusingler: reference CSR + labels and query CSR in, "
          "[[cell, firstLabel, finalLabel], ...] out.",
          "cores"_a, "gpuid"_a, "ref_height"_a, "ref_width"_a, "ref_data"_a, "ref_indices"_a,
          "ref_indptr"_a, "qry_height"_a, "qry_width"_a, "qry_data"_a, "qry_indices"_a,
          "qry_indptr"_a, "codes"_a, "celltype"_a, "cellnames"_a, "ref_geneidx"_a,
          "qry_geneidx"_a, "quantile"_a = 80, "finetune_thre"_a = 0.05, "finetune_times"_a = 0,
          py::kw_only(), "path"_a = "plain", "return_scores"_a = false, "verbose"_a = true);
#ifdef VERSION_INFO
    m.attr("__version__") = VERSION_INFO;
#else
    m.attr("__version__") = "dev";
#endif
}
