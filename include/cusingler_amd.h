/* C-ABI of the MI355X SingleR annotation engine (libcusingler_amd.so).
 *
 * Each entry point replaces one function of the reference's device API
 * (src/cusingler.cuh:13-22) and keeps its meaning: the same inputs (the coded
 * CSR matrices, label ranges and pair gene windows the host preprocessing
 * produces, src/types.h:21-42), the same outputs (first-label indices, the
 * 0/1 candidate mask, final-label indices).  Differences at the boundary:
 *   - state lives in an opaque engine handle instead of file-scope globals
 *     (src/cusingler.cu:26-46), so several engines may coexist;
 *   - every function returns a status code (0 = success) instead of calling
 *     exit() (src/cusingler.cu:48-57); csg_last_error() gives the message;
 *   - plain pointers and sizes only, no C++ types.
 *
 * Index conventions: reference rows are grouped by label, label l owning rows
 * [ctidx[2l], ctidx[2l] + ctidx[2l+1]) (src/io.cpp:616-684); value codes are
 * the per-row dense ranks of distinct stored values, starting at 1, with 0
 * for implicit zeros (src/io.cpp:470-523).
 */
#ifndef CUSINGLER_AMD_H
#define CUSINGLER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csg_engine csg_engine;

enum csg_status
{
    CSG_OK            = 0,
    CSG_ERR_ARGUMENT  = 1, /* invalid sizes, codes or indices */
    CSG_ERR_DEVICE    = 2, /* HIP runtime error (the message names the call) */
    CSG_ERR_MEMORY    = 3, /* not enough device memory for the requested plan */
    CSG_ERR_STATE     = 4, /* entry points called out of order */
    CSG_ERR_UNDEFINED = 5  /* the reference's own behaviour is undefined here */
};

/* Correlation forms of the score step (csg_get_label `path`). */
enum csg_path
{
    CSG_PATH_PLAIN      = 0, /* FP32, reference summation order: bit-exact scores */
    CSG_PATH_EXACT_INT8 = 1, /* exact rank dot products on INT8-limb MFMA, FP64
                                normalisation, certified labels with FP32 recheck */
    CSG_PATH_FP32_MFMA  = 2, /* FP32-in/FP32-accumulate MFMA (approximate) */
    CSG_PATH_FP16_MFMA  = 3  /* FP16-in/FP32-accumulate MFMA (approximate) */
};

/* initGPU (src/cusingler.cuh:15, src/cusingler.cu:74-90): selects device
 * `gpuid` and creates an engine bound to it. */
int csg_init_gpu(int gpuid, csg_engine** out);

/* Message of the last failing call on this engine ("" if none). */
const char* csg_last_error(const csg_engine* eng);

/* Releases every device buffer and the engine itself. */
void csg_release(csg_engine* eng);

/* copyin_score (src/cusingler.cuh:18, src/cusingler.cu:283-413): uploads the
 * score-step inputs, coded over the score gene set (width = |common_genes|).
 * max_uniq_gene is the preprocessing's bound on codes + 1
 * (src/io.cpp:502,520).  ctidx holds 2*ct_num entries.  With the three query
 * pointers NULL only the reference is uploaded and room is reserved for
 * query batches of up to qry_height rows, which csg_copyin_query_raw
 * supplies (max_uniq_gene then bounds the reference's codes). */
int csg_copyin_score(csg_engine* eng, uint32_t ref_height, uint32_t width,
                     const uint16_t* ref_data, const int32_t* ref_indices,
                     const int32_t* ref_indptr, uint32_t qry_height, const uint16_t* qry_data,
                     const int32_t* qry_indices, const int32_t* qry_indptr, uint32_t ct_num,
                     const uint32_t* ctidx, uint64_t max_uniq_gene);

/* get_label (src/cusingler.cuh:19, src/cusingler.cu:986-1136): Spearman
 * correlation of every query cell with every reference cell, the per-label
 * `quantile` (a fraction, 0.8 by default) of those scores, the first label
 * (last maximum) and the candidate mask (score >= max - finetune_thre).
 * Outputs may be NULL; results stay on the device for csg_finetune either way.
 * With all three NULL the call returns once the step is queued on the
 * engine's stream; its counters and the CSG_ERR_UNDEFINED of a NaN-carrying
 * score segment are then reported by the next csg_finetune, csg_counters,
 * csg_phase_times or csg_get_label.
 *   first_labels : int32 [qry_height]
 *   labels_mask  : float [qry_height * ct_num] (0/1, the reference's `labels`)
 *   scores       : float [qry_height * ct_num] per-label quantile scores */
int csg_get_label(csg_engine* eng, float quantile, float finetune_thre, int path,
                  int32_t* first_labels, float* labels_mask, float* scores);

/* destroy_score (src/cusingler.cuh:20): frees the score-step buffers. */
int csg_destroy_score(csg_engine* eng);

/* copyin (src/cusingler.cuh:24, src/cusingler.cu:415-521): uploads the
 * fine-tune inputs, coded over the fine-tune gene set (width = |thre_genes|),
 * and the per ordered label pair gene windows: ctdidx holds [pos, len] at
 * (i*ct_num + j)*2 into ctdiff (gene indices in the fine-tune space).
 * labels_mask (qry_height*ct_num floats) may be NULL to use the mask the last
 * csg_get_label left on the device.  With the three query pointers NULL
 * (and no mask) room is reserved for query batches of up to qry_height rows,
 * as in csg_copyin_score. */
int csg_copyin(csg_engine* eng, uint32_t ref_height, uint32_t width, const uint16_t* ref_data,
               const int32_t* ref_indices, const int32_t* ref_indptr, uint32_t qry_height,
               const uint16_t* qry_data, const int32_t* qry_indices, const int32_t* qry_indptr,
               uint32_t ct_num, const uint32_t* ctidx, const uint32_t* ctdidx,
               const uint32_t* ctdiff, uint64_t n_ctdiff, uint64_t max_uniq_gene,
               const float* labels_mask);

/* cufinetune (src/cusingler.cuh:26, src/cusingler.cu:1297-1333): iterative
 * refinement of every cell's candidate labels; finetune_times == 0 runs until
 * one label is left.  final_labels: uint32 [qry_height]; rounds (optional):
 * int32 [qry_height] rounds run per cell. */
int csg_finetune(csg_engine* eng, float quantile, float finetune_thre, int finetune_times,
                 uint32_t* final_labels, int32_t* rounds);

/* Correlation form of the fine-tune step: CSG_PATH_EXACT_INT8 contracts the
 * ranks exactly on INT8-limb MFMA and certifies every cell's filter decisions
 * against the reference's FP32 chunk-tree order, redoing the cells it cannot
 * certify in that order (same results as the plain form); any other value
 * uses the plain FP32 chunk-tree kernels.  csg_get_label sets it to the path
 * it ran; this call overrides that.  No counterpart in the reference. */
int csg_set_finetune_path(csg_engine* eng, int path);

/* Fine-tune schedule: 0 (default) runs every round on the device -- cells
 * grouped by candidate set through a device hash table, per-group gene
 * subsets built from the pair windows, work-queue kernels over device-side
 * counts, each round replayed from two captured HIP graphs with one counter
 * read-back; 1 groups on the host and runs the device work in batches.
 * Labels and round counts are the same.  No counterpart in the reference. */
int csg_set_finetune_schedule(csg_engine* eng, int schedule);

/* Largest fine-tune group (in cells) that the device-resident schedule
 * correlates in the fused FP32 form: one pass per reference row ranks the row
 * over the round's gene subset and forms its scores against the group's cells
 * in the reference's 128-chunk + stride-tree order, without storing the
 * row's ranks (default 4; 0 turns the fused form off; at most 4).  Labels
 * and round counts do not depend on it.  No counterpart in the reference. */
int csg_set_finetune_fuse(csg_engine* eng, int max_cells);

/* Longest chain of fine-tune rounds that the device-resident schedule
 * launches together without reading anything back (default 0: chains off;
 * at most 64).  Once a round has at most 1,024 active cells, the
 * following rounds run as chains of captured rounds: a round removes at least
 * one label from every cell, so a chain covers the rounds the longest
 * candidate list can still take.  Each chained round checks on the device
 * that its work fits the tables and the launch mask the chain was captured
 * with, and that its correlations (cells x reference rows) are at most
 * max_scores (0: no limit); one that does not halts the chain with its state
 * intact, and the host runs it again unchained.  Labels and round counts do
 * not depend on these settings.  No counterpart in the reference. */
int csg_set_finetune_chain(csg_engine* eng, int max_rounds, uint64_t max_scores);

/* The device-resident schedule's last rounds in one cooperative launch: once
 * a round has at most max_cells active cells (default 64; 0 never), that
 * round and every later one run in a persistent kernel of blocks_per_cu
 * blocks of 512 threads per CU (default 4; 1 to 16, fewer if they cannot be
 * co-resident).
 * Each cell is a group of its own; per round the cells' owner blocks build
 * G_sub and rank the cell, every wave ranks and correlates reference rows,
 * and the owner blocks select the quantiles and filter, with two grid
 * barriers per round and no host involvement.  The scores are the fused
 * FP32 form's, so labels and round counts do not depend on these settings.
 * No counterpart in the reference. */
int csg_set_finetune_tail(csg_engine* eng, uint32_t max_cells, int blocks_per_cu);

/* Phase durations of the last fine-tune's tail kernel, in microseconds of
 * the device clock between successive grid barriers: [0] the first round's
 * begin, then per round the rows phase and the end + next begin phase.
 * *count receives how many were written (0 when the tail did not run, or
 * minus the reason it was last refused: 1 its LDS, 2 no cooperative launch,
 * 3 no co-resident block, 4 the launch refused). */
int csg_finetune_tail_profile(csg_engine* eng, double* us, int max, int* count);

/* Autotuner of the device-resident fine-tune: on the first round with enough
 * work after the fine-tune inputs change, the block counts of the reference
 * rank, correlation and select work-queue kernels and the exact correlation
 * tile shape (16-cell tiles, or 32-cell tiles for groups of at least t cells)
 * are chosen by timing each candidate on that round's work.  autotune = 1 / 0
 * turns it on / off (and asks for a new tuning), -1 leaves it; grids[0..5]
 * receive (tuned, rank, correlation, select block counts, t or -1 for 16-cell
 * tiles only, fused-form block count) and best_ms[0..3] the tuned kernels'
 * times (rank, correlation, select, fused; either array may be NULL).  No
 * counterpart in the reference. */
int csg_finetune_tuning(csg_engine* eng, int autotune, int32_t* grids, float* best_ms);

/* Device memory budget (bytes) of one device-resident fine-tune work batch:
 * a round whose work tables would exceed it is processed in group ranges
 * that fit (default 8 GiB; 0 leaves it).  *batches (optional) receives the
 * number of such range batches run so far.  No counterpart in the
 * reference. */
int csg_set_finetune_budget(csg_engine* eng, uint64_t bytes, int64_t* batches);

/* Loads a batch of query cells as the reference's pipeline receives them --
 * float CSR values over the query's own gene list (after log-normalising,
 * src/pipeline.cpp:128-152) -- in place of the loaded query, for both the
 * score step and the fine-tune, and value-codes it on the device: csrFilter
 * (src/io.cpp:470-523) over the score gene set and over the fine-tune gene
 * set at once, bit-identical to the host's codes.  The values and columns
 * travel in slices on a second stream (pinned host memory makes that a
 * direct DMA), each slice coded while the next one is in flight.
 *   qry_indptr    : qry_height + 1 row pointers into qry_data / qry_indices
 *   map_score     : qry_width entries, csg_prep_query_map(step 0)
 *   map_finetune  : qry_width entries, csg_prep_query_map(step 1)
 * Needs csg_copyin_score and csg_copyin loaded (with data, or reserving
 * qry_height rows or more).  The batch's codes raise max_uniq_gene of each
 * step above the reference's where a cell needs it.  Then csg_get_label and
 * csg_finetune annotate the batch; the candidate mask comes from
 * csg_get_label.  CSG_ERR_UNDEFINED for a NaN among a row's kept values or
 * more than 65535 distinct values in a row (the reference's uint16 codes).  The host arrays
 * are read before the call returns.  No counterpart in the reference: its
 * DataParser codes every query row on the host before copyin_score. */
int csg_copyin_query_raw(csg_engine* eng, uint32_t qry_height, uint32_t qry_width,
                         const float* qry_data, const int32_t* qry_indices,
                         const int32_t* qry_indptr, const int32_t* map_score,
                         const int32_t* map_finetune);

/* Value coding of a float CSR that is already on the device (csrFilter,
 * src/io.cpp:470-523), for two gene sets a and b at once -- the device half of
 * csg_copyin_query_raw, for callers whose cells arrive on the GPU.  All
 * pointers but max_uniq are device pointers:
 *   data, indices, indptr : rows x width float CSR (row r's entries at
 *                           [indptr[r], indptr[r + 1]))
 *   map_a, map_b          : width entries each, column -> gene-set column or
 *                           -1 (csg_prep_query_map)
 *   data_x, indices_x     : the coded CSR of set x (room for indptr[rows]
 *                           entries); indptr_x : rows + 1 row pointers
 *   max_uniq (host, 2)    : max over rows of (#distinct kept values + 1) per
 *                           set (may be NULL)
 * Runs on the engine's stream and returns when the codes are written;
 * CSG_ERR_UNDEFINED for a NaN among a row's kept values or a row of more than 65535 distinct
 * values.  Work buffers (2 x rows counts and row lists) are the engine's,
 * grown on demand.  No counterpart in the reference (its coding runs on the
 * host, src/io.cpp:714-784). */
int csg_code_csr(csg_engine* eng, uint32_t rows, uint32_t width, const float* data,
                 const int32_t* indices, const int32_t* indptr, const int32_t* map_a,
                 const int32_t* map_b, uint16_t* data_a, int32_t* indices_a, int32_t* indptr_a,
                 uint16_t* data_b, int32_t* indices_b, int32_t* indptr_b, uint32_t* max_uniq);

/* csg_code_csr for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same inputs and outputs as csg_code_csr
 * (device pointers), but the work memory (row counts and row lists) lives in
 * a fixed per-device array of the library, 16 MiB + 256 B, held from the
 * first csg_init_gpu on that device until the process ends; the call
 * allocates nothing and does not synchronise.  At most
 * CSG_CODE_NOALLOC_MAX_ROWS rows per call (CSG_ERR_ARGUMENT beyond).  Every
 * call on one device shares that array, so calls that overlap on one GPU
 * (two engines, or two host threads, on the same device) race; calls on one
 * engine are ordered by its stream.  Results are valid once the engine's
 * stream has passed the call: `status` (device, 4 words) then holds
 * [flags, max_uniq of set a, max_uniq of set b, 0], flags bit 0 a NaN kept value,
 * bit 1 a column outside `width`, bit 2 a row whose distinct values exceed
 * uint16 codes or the block coder's LDS (codes are then not valid). */
#define CSG_CODE_NOALLOC_MAX_ROWS (1 << 20)

/* Waits until the engine's stream has passed every call made on it (for the
 * entry points that do not synchronise).  No counterpart in the reference. */
int csg_synchronize(csg_engine* eng);

int csg_code_csr_noalloc(csg_engine* eng, uint32_t rows, uint32_t width, const float* data,
                         const int32_t* indices, const int32_t* indptr, const int32_t* map_a,
                         const int32_t* map_b, uint16_t* data_a, int32_t* indices_a,
                         int32_t* indptr_a, uint16_t* data_b, int32_t* indices_b,
                         int32_t* indptr_b, uint32_t* status);

/* Diagnostic: copies the query's value-coded CSR of step 0 (score gene set)
 * or 1 (fine-tune gene set) from the device -- what csg_copyin_query_raw
 * produced, or what csg_copyin_score / csg_copyin received.  indptr receives
 * rows + 1 row pointers; data / indices (may be NULL) indptr[rows] -
 * indptr[0] entries.  No counterpart in the reference. */
int csg_query_coded(csg_engine* eng, int step, uint16_t* data, int32_t* indices, int32_t* indptr);

/* Page-locked host memory from the library's HIP runtime, for loaders that
 * read query batches straight into the buffers csg_prefetch_query_raw /
 * csg_copyin_query_raw upload (DMA without a staging copy).  csg_host_free
 * takes only pointers csg_host_alloc returned.  No counterpart in the
 * reference. */
int csg_host_alloc(uint64_t bytes, void** ptr);
int csg_host_free(void* ptr);

/* Entries per upload slice of csg_copyin_query_raw (default 2^26; whole rows
 * per slice, a longer row is a slice of its own).  No counterpart in the
 * reference. */
int csg_set_upload_slice(csg_engine* eng, uint64_t entries);

/* trainData (src/io.cpp:141-262) on the device: per label and gene the
 * median with implicit zeros, per ordered label pair the genes by descending
 * rounded median difference (ties to the lower gene), the pair's positive
 * genes (ref_train_idxs / ref_train_values), the top 333 positives of every
 * pair (thre_genes) and the top k_pair(L) genes of every pair regardless of
 * sign (common_genes) -- bit-identical to the reference's.  Input: the
 * reference's float CSR as given to the pipeline (rows in any label order:
 * the medians do not depend on it), its per-row label codes (narrowed to
 * uint8), and ref_map (width entries: the column after removeCols, or -1;
 * csg_prep_ref_map; NULL = identity).  ref_width is the reference's
 * pre-intersection width, as in the reference (genes past the shared ones
 * have median 0).  sizes (optional) receives [|ref_train_values|,
 * |common_genes|, |thre_genes|]; csg_train_table copies them out. */
int csg_train(csg_engine* eng, uint32_t ref_height, uint32_t ref_width, const float* ref_data,
              const int32_t* ref_indices, const int32_t* ref_indptr, const int32_t* ref_map,
              uint32_t width, uint32_t ct_num, const int32_t* celltype, uint64_t* sizes);

/* Results of the last csg_train: which = 0 ref_train_idxs ([pos, len] per
 * ordered pair, 2 ct_num^2), 1 ref_train_values, 2 common_genes, 3
 * thre_genes (ascending).  Returns the element count; copies when out !=
 * NULL. */
uint64_t csg_train_table(const csg_engine* eng, int which, uint32_t* out);

/* Starts the upload of the next query batch's values and columns on the
 * engine's second stream and returns at once, so that the upload overlaps the
 * score step and fine-tune of the batch before it (the reference slices and
 * uploads synchronously, src/cusingler.cu:335-409).  The
 * next csg_copyin_query_raw with the same three pointers and row count codes
 * that batch from its device copy instead of uploading it; the host arrays
 * must stay valid and unchanged until that call.  Two batches can be in
 * flight (two device copies, consumed oldest first), so the upload of batch
 * i + 1 can be started before batch i is coded; a third prefetch drops the
 * older unconsumed one.  No counterpart in the reference. */
int csg_prefetch_query_raw(csg_engine* eng, uint32_t qry_height, const float* qry_data,
                           const int32_t* qry_indices, const int32_t* qry_indptr);

/* csg_prefetch_query_raw and csg_copyin_query_raw with 16-bit column indices
 * (query widths up to 65536; CSG_ERR_ARGUMENT beyond): 6 instead of 8 bytes
 * per stored value cross the bus, for loaders that read the columns in that
 * form (cusingler_amd/h5ad.py does when the width allows).  A prefetched
 * batch is consumed only by the copy-in of the same index width.  Results are
 * those of the 32-bit forms.  No counterpart in the reference. */
int csg_prefetch_query_raw_u16(csg_engine* eng, uint32_t qry_height, const float* qry_data,
                               const uint16_t* qry_indices, const int32_t* qry_indptr);
int csg_copyin_query_raw_u16(csg_engine* eng, uint32_t qry_height, uint32_t qry_width,
                             const float* qry_data, const uint16_t* qry_indices,
                             const int32_t* qry_indptr, const int32_t* map_score,
                             const int32_t* map_finetune);

/* Caps the query chunk of the score step (src/cusingler.cu:335-409 slices the
 * query to the free device memory; here the score buffer takes at most 1/16
 * of the free memory, a chunk of the learned proposal path fills at most that
 * buffer, and a chunk of a dense score matrix -- the paths without the
 * proposal, csg_score_matrix -- holds at most 8 GiB of scores).  rows: a
 * multiple of 128, 0 removes the cap; it applies from the next
 * csg_copyin_score.  *chunk (optional) receives the proposal path's chunk of
 * the current inputs.  Lets a caller bound the score buffer, and tests run
 * several chunks on small inputs. */
int csg_set_query_chunk(csg_engine* eng, uint32_t rows, uint32_t* chunk);

/* destroy (src/cusingler.cuh:25): frees the fine-tune buffers. */
int csg_destroy(csg_engine* eng);

/* Per-phase device time of the calls since the last reset, from HIP events on
 * the engine's stream, in milliseconds: [0] rank, [1] correlation,
 * [2] quantile select, [3] label/mask, [4] certification recheck,
 * [5] fine-tune total, [6] fine-tune correlation, [7] fine-tune rank,
 * [8] fine-tune select+filter, [9] learned proposal (network inference and
 * work lists), [10] query upload and value coding (csg_copyin_query_raw).
 * counts[i] is the number of launches timed in phase i.  Either pointer may
 * be NULL. */
#define CSG_N_PHASES 11
int csg_phase_times(csg_engine* eng, double* ms, int64_t* counts, int reset);

/* Which phases csg_phase_times times (no counterpart in the reference: an
 * instrumentation switch of this engine).  Every timed phase records two HIP
 * events on the stream, which costs about 1% of a proposal-path step at the
 * atlas size.  mode 0: none; 1: the score-step correlation phase only (the
 * contraction's live launch time, as bench.py's roofline uses it); 2: every
 * phase (the default); 3: every phase, and the device-resident fine-tune
 * launches each round's work kernels eagerly instead of replaying their
 * captured graph, so that its rank [7], correlation [6] and select+filter [8]
 * families are timed too.  CSG_ERR_ARGUMENT for any other mode. */
int csg_set_phase_timing(csg_engine* eng, int mode);

/* Counters of the last csg_get_label / csg_finetune call:
 * [0] query cells, [1] score-step gene count G_s, [2] reference cells,
 * [3] (cell, label) pairs re-scored in FP32 by the exact path,
 * [4] cells the exact path could not certify, [5] NaN-carrying segments,
 * [6] fine-tune rounds, [7] fine-tune (cell, round) items,
 * [8] fine-tune candidate-set groups, [9] fine-tune correlation pairs,
 * [10] fine-tune sum over items of gene-subset size * reference rows,
 * [11] fine-tune cells the exact form redid in FP32 order,
 * [12..15] fine-tune host microseconds: candidate lists from the mask,
 * grouping by candidate set, work-table building, batch round trips,
 * [16] cells scored on proposed labels only, [17] cells below the proposal
 * margin (scored on every label), [18] confident cells the edge tests sent
 * to scoring on every label, [19] (cell, label) pairs the proposal path
 * scored, [20] cells widened to their second group of ranked labels,
 * [21] correlations the proposal path computed (sum over its pairs of the
 * label's reference cells), [22] fine-tune groups of the device-resident
 * schedule that took the fused FP32 form (csg_set_finetune_fuse), [23]
 * fine-tune rounds launched in chains without a host read-back (including
 * rounds past the last one, which do nothing), [24] chains halted by a
 * round that needed larger work tables, another launch mask or the host
 * schedule (csg_set_finetune_chain), [25] fine-tune rounds run by the tail
 * kernel (csg_set_finetune_tail). */
#define CSG_N_COUNTERS 26
int csg_counters(csg_engine* eng, int64_t* out);

/* ---- learned label proposal (no counterpart in the reference) ----
 *
 * A perceptron over the score step's rank features (doubled centred ranks
 * over the score gene set, divided by G_s, in BF16): G_s -> H -> H -> L with
 * GELU, trained with softmax cross-entropy and Adam (FP32 master weights,
 * BF16 operands on MFMA, FP32 accumulation).  Parameters form one FP32 array
 * [W1 (Kp x H) | b1 (H) | W2 (H x H) | b2 (H) | W3 (H x Lp) | b3 (Lp)],
 * row-major, Kp = G_s rounded up to 64, Lp = L rounded up to 64; input rows
 * >= G_s and output columns >= L are inert.  The network belongs to the
 * engine and survives csg_copyin_score calls with the same G_s and L. */

/* Random initialisation (seeded) for the loaded score data's G_s and L;
 * hidden = H, a multiple of 64 in [64, 4096]. */
int csg_mlp_init(csg_engine* eng, int hidden, uint64_t seed);

/* dims[0..5] = G_s, Kp, H, L, Lp, number of parameters. */
int csg_mlp_dims(const csg_engine* eng, int64_t* dims);

/* Copy the FP32 parameters out / in (setting them clears the Adam state). */
int csg_mlp_get_params(csg_engine* eng, float* out);
int csg_mlp_set_params(csg_engine* eng, const float* in);

/* Loss and gradient of the mean cross-entropy over n rows of features x
 * ([n x G_s] FP32, rounded to BF16 on the device) with the given labels, at
 * the current parameters; grad receives one value per parameter.  For
 * checking the training kernels; the parameters are not changed. */
int csg_mlp_grad(csg_engine* eng, uint32_t n, const float* x, const int32_t* labels, float* grad,
                 float* loss);

/* Train on the loaded query cells (csg_copyin_score) with one label per cell
 * (e.g. the final labels of the plain method): `epochs` shuffled passes of
 * minibatches of `batch` rows, Adam at learning rate lr with cosine decay to
 * lr / 20.  epoch_loss (optional) receives the mean loss of each epoch. */
int csg_mlp_train(csg_engine* eng, const int32_t* labels, int epochs, int batch, float lr,
                  uint64_t seed, float* epoch_loss);

/* Top-k labels ([Nq x k], descending logits) and softmax margin p1 - p2 of
 * every loaded query cell, 1 <= k <= min(8, L). */
int csg_mlp_predict(csg_engine* eng, int k, int32_t* top, float* margin);

/* Proposal-restricted scoring for csg_get_label on the exact INT8 path
 * (k = 0 turns it off).  The network ranks kp = min(2k, 8, L) labels per
 * cell.  Cells with margin >= min_margin are scored on their first k labels
 * only and the rest on every label.  A proposal-scored cell whose k-th label
 * may lie within finetune_thre of its exact winner (so that the candidate
 * set, or the winner, could extend past the proposal) is scored on its labels
 * k .. kp-1 too, and on every label if its kp-th label is still that close.
 * Unscored labels get quantile -inf, so they stay out of the mask; the labels
 * then differ from the full path's only where a label the network ranked
 * below the last scored one would have entered the mask although that one
 * did not.  Any score gene set (up to 1920 genes the grouped contraction
 * folds the high limbs into two partials, beyond it keeps three); at most
 * 256 labels. */
int csg_set_proposal(csg_engine* eng, int k, float min_margin);

/* Diagnostic: the per-pair correlation matrix [nq x ref_height] of query rows
 * [q0, q0 + nq) of the loaded score data (nq at most one dense chunk) by the
 * given path's contraction (the exact path's values before certification),
 * for stating each path's error against FP64.  No counterpart in the
 * reference (the reference keeps this matrix internal, src/cusingler.cu:
 * 1071-1090). */
int csg_score_matrix(csg_engine* eng, int path, uint32_t q0, uint32_t nq, float* out);

/* Limb form of the exact INT8 contraction.  form 0 (default): for gene sets
 * of up to 1008 genes (|X| <= 1007) the three-product form, X = 65h + l with
 * X Y = 65 [(h + l)(h' + l') + (8h)(8h')] - 64 l l' (three INT8 products per
 * score, two int32 partials, per-row offset corrections); wider sets keep the
 * folded four-product form.  form 1: the folded form for every gene set.
 * Scores are exact in both, so results do not depend on this setting.  *prev
 * (may be null) receives the previous form.  No counterpart in the
 * reference. */
int csg_set_limb_form(csg_engine* eng, int form, int* prev);

/* Diagnostic of the contraction kernels on seeded random ranks (|X| <= 1500,
 * 256 x 256 rows, K = 192): which = 0 INT8 limbs, 1 FP16, 2 FP32 MFMA,
 * 3 plain FP32, 4 INT8 limbs wide-tile kernel (gene sets up to 1920), 5 INT8
 * limbs three-partial wide-tile kernel (|X| <= 32000: every high-limb value),
 * 6 the wide-tile kernel's three-product form (|X| <= 1007).
 * *max_err receives 0 when the device results equal the host's
 * exact (0, 4, 5, 6) / sequential-fma (3) values, else the largest difference (+1); for
 * 1 and 2 the largest error of the dot product relative to sqrt(sxx * syy),
 * i.e. in correlation units.  which = 7 compares the proposal's two top-k
 * kernels (one thread per row for up to 64 labels, one wave per row) on
 * seeded logits with ties, -inf and NaN rows; *max_err receives the number
 * of differing labels and margins (bits).  No counterpart in the reference:
 * it checks the MFMA operand layouts and kernel equivalences. */
int csg_selftest(csg_engine* eng, int which, double* max_err);

/* Diagnostic of the fine-tune correlation kernels (wide and thin tiles): 40 x
 * 70 pairs of seeded random ranks over `gs` genes (1..16384), compared bit
 * for bit with the 128-chunk + tree FP32 order of spearman_reduce
 * (src/cusingler.cu:804-856) computed on the host; each kernel runs with its
 * widest LDS stage and with 64-gene stages; *mismatches receives the number
 * of differing sums over the four runs.  No counterpart in the reference. */
int csg_selftest_ft(csg_engine* eng, int gs, int64_t* mismatches);

/* Host-side evaluation of the reference's `percentile` (src/cusingler.cu:
 * 947-978) with the same libstdc++ sort emulation the device uses; returns
 * CSG_ERR_UNDEFINED if the library's unguarded scans would leave the array. */
int csg_host_percentile(const float* arr, int64_t len, float quantile, float* out);

/* Host-side libstdc++ minmax_element emulation (first min, last max). */
void csg_host_minmax(const float* arr, int n, int* imin, int* imax);

#ifdef __cplusplus
}
#endif

#endif /* CUSINGLER_AMD_H */
