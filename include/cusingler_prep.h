/* C-ABI of the host preprocessing of the SingleR annotation path
 * (libcusingler_amd.so), the counterpart of the reference's DataParser
 * (src/io.h:16-122, src/io.cpp:141-821).
 *
 * csg_prep_train runs, in the reference's order (src/pipeline.cpp:128-152):
 * prepareData, findIntersectionGenes, loadRefData (group by label, resort,
 * removeCols), trainData (per-label medians and ordered-pair marker genes),
 * loadQryData and preprocess.  csg_prep_generate(step) is
 * generateDenseMatrix(step): the value-coded CSR matrices over the score gene
 * set (step 0, `common_genes`) or the fine-tune gene set (step 1,
 * `thre_genes`, with the pair gene windows remapped into that set).
 * Results are bit-identical to the reference's; the work is spread over
 * `cores` threads.
 */
#ifndef CUSINGLER_PREP_H
#define CUSINGLER_PREP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csg_prep csg_prep;

/* Value-coded CSR matrices and label tables of one step (views into the
 * csg_prep object, valid until the next csg_prep_generate or release). */
typedef struct csg_coded
{
    uint32_t        ref_height, qry_height, width, ct_num;
    const uint16_t* ref_data;
    const int32_t*  ref_indices;
    const int32_t*  ref_indptr;
    const uint16_t* qry_data;
    const int32_t*  qry_indices;
    const int32_t*  qry_indptr;
    uint64_t        max_uniq_gene; /* max over rows of (#distinct values + 1) */
    const uint32_t* ctidx;         /* [start, len] per label (2*ct_num) */
    const uint32_t* ctdidx;        /* [pos, len] per ordered pair (2*ct_num^2) */
    const uint32_t* ctdiff;        /* pair gene windows */
    uint64_t        n_ctdiff;
    const uint32_t* genes;         /* the step's gene set (ascending, train space) */
} csg_coded;

/* PyPipeline::train (src/pipeline.cpp:128-156).  celltype[i] is the label
 * code of reference cell i (narrowed to uint8 as in src/io.cpp:813); codes
 * 0..n_codes-1 name the labels.  Gene names are NUL-terminated strings.
 * Returns 0 or a csg_status code; errbuf receives the message. */
int csg_prep_train(int cores, uint32_t ref_height, uint32_t ref_width, const float* ref_data,
                   const int32_t* ref_indices, const int32_t* ref_indptr, uint32_t qry_height,
                   uint32_t qry_width, const float* qry_data, const int32_t* qry_indices,
                   const int32_t* qry_indptr, uint32_t n_codes, const int32_t* celltype,
                   const char* const* ref_genes, uint32_t n_ref_genes,
                   const char* const* qry_genes, uint32_t n_qry_genes, csg_prep** out,
                   char* errbuf, size_t errlen);

/* generateDenseMatrix(step) (src/io.cpp:714-784). */
int csg_prep_generate(csg_prep* p, int step, csg_coded* out);

/* Training results for inspection: which = 0 ref_train_idxs ([pos, len] per
 * ordered pair in train space), 1 ref_train_values, 2 common_genes,
 * 3 thre_genes.  Returns the element count; copies when out != NULL. */
uint64_t csg_prep_train_table(const csg_prep* p, int which, uint32_t* out);

/* 1 if the gene lists differ and columns were remapped (filter_genes). */
int csg_prep_filtered(const csg_prep* p);

/* Column map of the query for value coding on the device
 * (csg_copyin_query_raw): for every column c of the query as given to
 * csg_prep_train (qry_width entries), the column of step `step`'s gene set
 * (0: common_genes, 1: thre_genes) that csrFilter (src/io.cpp:470-523)
 * writes for it, or -1 where the column is dropped (not shared with the
 * reference, src/io.cpp:306-349,525-538, or outside the gene set).  The
 * composition of removeCols' renumbering and generateDenseMatrix's index map
 * (src/io.cpp:714-721); it depends only on the gene lists, so it codes any
 * query batch over the same genes.  Returns 0 or CSG_ERR_ARGUMENT. */
int csg_prep_query_map(const csg_prep* p, int step, int32_t* out);

/* Column map of the reference for training on the device (csg_train): for
 * every reference column (ref_width entries) its shared column after
 * removeCols (src/io.cpp:306-349,687-711), or -1 where the gene is not
 * shared; the identity when both gene lists are equal.  Returns 0 or
 * CSG_ERR_ARGUMENT. */
int csg_prep_ref_map(const csg_prep* p, int32_t* out);

void csg_prep_release(csg_prep* p);

#ifdef __cplusplus
}
#endif

#endif /* CUSINGLER_PREP_H */
