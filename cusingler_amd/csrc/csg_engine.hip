// Host side of the MI355X SingleR annotation engine: the C-ABI declared in
// include/cusingler_amd.h, device memory plans and kernel launches.
//
// Score step (get_label, src/cusingler.cu:986-1136): the reference rows are
// ranked once; query rows are processed in chunks of Qc rows: rank, contract
// against every reference row (or, with the learned proposal, against the
// proposed labels' rows), per-label quantile select, then one label pass over
// all cells.  Fine-tune (cufinetune, :1297-1333): level-synchronous over all
// cells; by default each round groups the active cells by candidate set on the
// device and ranks, correlates and filters every group in work-queue kernels
// replayed from captured graphs; the host-grouped schedule does the grouping
// on the host.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cusingler_amd.h"
#include "csg_exact.h"
#include "csg_code.h"
#include "csg_train.h"
#include "csg_ft_kernels.h"
#include "csg_ftd.h"
#include "csg_ftd_tail.h"
#include "csg_kernels.h"
#include "csg_mfma.h"
#include "csg_mlp.h"
#include "csg_proposal.h"

using namespace csg;

namespace {

struct PhaseEvents
{
    int        phase;
    hipEvent_t a, b;
    int        launches;
};

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

int64_t now_us()
{
    return std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// fine-tune correlation tile: 32 cells x 32 reference rows, 2 x 2 per thread

// Work memory of csg_code_csr_noalloc, one copy per device for the life of
// the process: [misc 256 B | row counts 2 x CSG_CODE_NOALLOC_MAX_ROWS u32 |
// row lists 2 x CSG_CODE_NOALLOC_MAX_ROWS u32].  Host code never uses this
// name as an address: each engine resolves the device address once at
// csg_init_gpu (hipGetSymbolAddress) and passes that.
constexpr size_t NOALLOC_MISC  = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_MISC + (size_t)4 * 4 * CSG_CODE_NOALLOC_MAX_ROWS;
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];

}  // namespace

struct csg_engine
{
    int             device = 0;
    hipStream_t     stream = nullptr;
    size_t          lds_limit = 65536;  // dynamic LDS a block may use
    std::string     err;
    hipDeviceProp_t prop{};

    // ---- score step ----
    bool      score_loaded = false;
    uint32_t  Nr = 0, Nq = 0, G = 0, Kp = 0, L = 0;
    bool      w8_ok = false;  // the wide INT8 kernels' LDS could be granted
    bool      kf_ok = false;  // ... and the three-product form's
    int       limb_form = 0;  // 0: three-product form where it applies, 1: folded four products
    bool      x3_ok = false;  // the three-partial wide INT8 kernels' LDS could be granted
    bool      bw_ok = false;  // the wide BF16 network kernel's LDS could be granted
    int       nbins = 0;
    std::vector<uint32_t> ctidx;
    uint16_t *d_rdata = nullptr, *d_qdata = nullptr;
    int *d_rind = nullptr, *d_rptr = nullptr, *d_qind = nullptr, *d_qptr = nullptr;
    uint32_t* d_lstart = nullptr;
    uint32_t* d_llen   = nullptr;
    PctParam* d_pct    = nullptr;
    float     pct_quantile = -1.f;
    int16_t*  d_Xr = nullptr;
    float*    d_sqr_f = nullptr;
    double*    d_irr   = nullptr;  // inverse roots of the exact sums of X^2 (reference rows)
    bool      ref_ranked = false;
    uint32_t  Qc = 0;        // query rows per chunk (the proposal path)
    uint32_t  Qc_dense = 0;  // query rows per chunk of the dense score matrices
    uint32_t  qc_cap = 0;  // caller's cap on the query chunk (rows), 0: none
    // query batches value-coded on the device (csg_copyin_query_raw)
    uint32_t    q_cap = 0, fq_cap = 0;          // query rows the loaded inputs have room for
    bool        q_ready = false, fq_ready = false;  // query data present
    int         nbins_ref = 0, fnbins_ref = 0;  // the callers' max_uniq_gene (reference side)
    size_t      q_nnz_cap = 0, fq_nnz_cap = 0;  // entries d_qdata / d_fqdata hold
    hipStream_t copy_stream = nullptr;
    hipEvent_t  code_copied[2] = { nullptr, nullptr }, code_done[2] = { nullptr, nullptr };
    float*      d_raw_data[2] = { nullptr, nullptr };
    int32_t*    d_raw_idx[2]  = { nullptr, nullptr };
    size_t      raw_cap = 0;  // entries of each slice buffer
    int32_t*    d_raw_ptr = nullptr;
    int32_t *   d_map_s = nullptr, *d_map_f = nullptr;
    size_t      map_cap = 0;
    uint32_t*   d_code_cnt = nullptr;   // kept entries per row, both sets
    uint32_t*   d_code_over = nullptr;  // rows for the block coding kernel
    size_t      code_rows_cap = 0;
    uint8_t*    d_code_misc = nullptr;  // flags, n_over, max_uniq[2] (u32), bases[2] (i64)
    uint8_t*    h_code_misc = nullptr;  // pinned
    uint64_t    code_slice = (uint64_t)1 << 26;  // entries per upload slice
    uint8_t*    noalloc_scratch = nullptr;  // device address of `scratch` on this engine's device
    // query batches uploaded ahead (csg_prefetch_query_raw): two slots,
    // consumed oldest first
    struct PreSlot
    {
        float*      d_data = nullptr;
        int32_t*    d_idx  = nullptr;
        size_t      cap    = 0;
        hipEvent_t  copied = nullptr;
        bool        valid  = false;
        uint64_t    seq    = 0;
        const void *h_data = nullptr, *h_idx = nullptr, *h_ptr = nullptr;
        uint32_t    rows   = 0;
        int         idx_bytes = 4;  // column index width of the staged batch
    } pre[2];
    uint64_t pre_seq = 0;
    // results of the last csg_train
    std::vector<uint32_t> tr_idxs, tr_values, tr_common, tr_thre;
    int16_t*  d_Xq = nullptr;
    float*    d_sqq_f = nullptr;
    double*    d_irq   = nullptr;  // the same for the query chunk
    float*    d_S = nullptr;
    float*    d_P = nullptr;
    int32_t*  d_first = nullptr;
    float*    d_mask = nullptr;
    bool      mask_valid = false;
    uint64_t* d_nan_list = nullptr;
    unsigned int* d_nan_count = nullptr;
    int*      d_status = nullptr;
    // counters the score step's kernels add on the device (the indices of
    // `counters`), [32] the undefined-sort flag of its NaN segments, [36, 42)
    // the proposal rounds' plans (d_pplan: per round tiles, items, 2 unused,
    // as uint32), [42] the certification's pairs and tiles (uint32); read
    // back in one copy with the step's closing synchronisation
    unsigned long long* d_dcnt = nullptr;
    unsigned long long* h_dcnt = nullptr;  // pinned, 48 entries
    // csg_get_label without outputs leaves its read-back in flight: folded
    // (counters, grid estimates, phase events, the NaN-sort flag) by the next
    // call that synchronises (fold_score_step)
    bool dcnt_pending = false, dcnt_prop = false, dcnt_exact = false;
    // exact / approximate MFMA paths
    int8_t *  d_rh = nullptr, *d_rl = nullptr, *d_qh = nullptr, *d_ql = nullptr;
    double *  d_dr = nullptr, *d_dq = nullptr;  // rows' corrections, three-product form
    _Float16 *d_rf16 = nullptr, *d_qf16 = nullptr;
    int       ref_prepared_path = -1;
    bool      ref_kf = false;  // the prepared limb planes hold the three-product form
    uint8_t*  d_cstatus = nullptr;
    uint8_t*  d_exact = nullptr;
    RecheckItem* d_items = nullptr;
    unsigned int* d_nitems = nullptr;
    float*    d_Pref = nullptr;
    Segment*  d_item_seg = nullptr;
    // re-scored pairs grouped by label for the grouped FP32 contraction
    uint32_t * d_rc_cells = nullptr, *d_rc_abase = nullptr, *d_rc_ucells = nullptr;
    RecheckItem* d_rc_sorted = nullptr;
    uint64_t*  d_rc_coff  = nullptr;
    GroupTile* d_rc_tiles = nullptr;
    size_t     rc_cap_seg = 0, rc_cap_ucells = 0, rc_cap_sorted = 0, rc_cap_cells = 0, rc_cap_l = 0,
           rc_cap_coff = 0, rc_cap_tiles = 0;

    // ---- fine-tune ----
    bool      ft_loaded = false;
    uint32_t  fNr = 0, fNq = 0, Gf = 0, fL = 0;
    int       fnbins = 0;
    std::vector<uint32_t> fctidx, ctdidx, ctdiff;
    std::vector<float>    host_mask;  // empty: use the device mask
    uint16_t* d_fref_codes = nullptr;
    uint8_t*  d_fref_codes8 = nullptr;  // the same codes in bytes, when all are below 256
    uint32_t *d_flstart = nullptr, *d_fllen = nullptr;
    uint16_t *d_fqdata = nullptr;
    int *d_fqind = nullptr, *d_fqptr = nullptr;
    PctParam* d_fpct = nullptr;
    float     fpct_quantile = -1.f;
    // fine-tune work buffers (grow-only)
    uint8_t*  d_ft_tables = nullptr;
    size_t    d_ft_tables_cap = 0;
    int16_t * d_ft_rr = nullptr, *d_ft_cr = nullptr;
    size_t    d_ft_rr_cap = 0, d_ft_cr_cap = 0;
    float *   d_ft_rsq = nullptr, *d_ft_csq = nullptr, *d_ft_scores = nullptr, *d_ft_P = nullptr;
    size_t    d_ft_rsq_cap = 0, d_ft_csq_cap = 0, d_ft_scores_cap = 0, d_ft_P_cap = 0;
    uint64_t* d_ft_nan = nullptr;
    size_t    d_ft_nan_cap = 0;
    uint8_t*  d_ft_out = nullptr;
    size_t    d_ft_out_cap = 0;
    uint8_t * h_ft_stage = nullptr, *h_ft_out = nullptr;
    size_t    h_ft_stage_cap = 0, h_ft_out_cap = 0;
    uint8_t * d_ft_redo = nullptr, *h_ft_redo = nullptr;
    size_t    d_ft_redo_cap = 0, h_ft_redo_cap = 0;
    bool      ft_exact = false;  // exact (certified) fine-tune contraction
    std::unordered_map<uint64_t, std::vector<uint64_t>> window_cache;  // pair gene windows

    // device-resident fine-tune rounds (schedule 0)
    int       ft_schedule = 0;  // 0: device-resident rounds, 1: host-grouped batches
    bool      ftd_static  = false;
    uint8_t*  d_fl_lists = nullptr;
    uint32_t *d_fl_counts = nullptr, *d_fl_act = nullptr, *d_fl_next = nullptr;
    int32_t*  d_fl_rounds = nullptr;
    uint32_t *d_fl_cslot = nullptr, *d_fl_cgid = nullptr, *d_fl_cpos = nullptr;
    uint64_t* d_fl_htab = nullptr;
    uint32_t *d_fl_howner = nullptr, *d_fl_hgid = nullptr, *d_fl_gowner = nullptr,
             *d_fl_gcount = nullptr;
    FtGroup*  d_fl_groups = nullptr;
    uint8_t*  d_fl_gflag  = nullptr;
    uint32_t *d_fl_glabels = nullptr, *d_fl_glpos = nullptr, *d_fl_ggenes = nullptr,
             *d_fl_gtile0 = nullptr, *d_fl_gthin0 = nullptr;
    uint64_t*  d_fl_gseg0 = nullptr;
    FtdCounts* d_fl_cnt   = nullptr;
    FtdCounts* h_fl_cnt   = nullptr;  // pinned
    uint32_t * d_fl_cut = nullptr, *d_fl_ctdidx = nullptr, *d_fl_ctdiff = nullptr,
             *d_fl_final = nullptr;
    float*      d_fl_mask  = nullptr;
    FtRankItem* d_fl_items = nullptr;
    FtCell*     d_fl_cells = nullptr;
    FtCorrTile *d_fl_tiles = nullptr, *d_fl_thin = nullptr, *d_fl_rthin = nullptr;
    uint8_t*    d_fl_stage = nullptr;
    uint32_t*   d_fl_redo  = nullptr;
    uint64_t *  d_fl_rseg = nullptr, *d_fl_nan1 = nullptr, *d_fl_nan2 = nullptr;
    uint32_t *  d_fl_segcell = nullptr, *d_fl_refgroup = nullptr, *d_fl_refrow = nullptr;
    float*      d_fl_pred = nullptr;
    struct
    {
        size_t items, cells, stage, redo, tiles, thin, rthin, rseg, pred, nan1, nan2, segcell,
            refgroup, refrow;
    } fl_cap{};
    uint32_t               ftd_gcap = 0, ftd_T = 0;
    size_t                 ftd_gstride = 0;
    hipGraphExec_t         ftd_gA = nullptr;   // part A, many cells
    std::vector<uintptr_t> ftd_sigA;
    hipGraphExec_t         ftd_gAs = nullptr;  // part A, a round of few cells
    std::vector<uintptr_t> ftd_sigAs;
    // part A of the rounds after the first, led by the previous round's
    // compaction ([0] many cells, [1] few)
    hipGraphExec_t         ftd_gAc[2] = { nullptr, nullptr };
    std::vector<uintptr_t> ftd_sigAc[2];
    bool                   ftd_small_a = false;  // the next part A is the few-cell one
    // part B graphs, one per launch mask (which work-queue kernels a round
    // needs, from its counts)
    struct FtdGraphB
    {
        uint32_t               mask = 0;
        hipGraphExec_t         exec = nullptr;
        std::vector<uintptr_t> sig;
    };
    std::vector<FtdGraphB> ftd_gB;
    uint32_t               ftd_mask = 0xFFFFFFFFu;  // launch mask of the current round
    // chains of rounds launched together without a host read-back (few-cell
    // rounds; one graph per chain length, launch mask and tables)
    struct FtdChainGraph
    {
        int                    rounds = 0;
        uint32_t               mask   = 0;
        hipGraphExec_t         exec   = nullptr;
        std::vector<uintptr_t> sig;
    };
    std::vector<FtdChainGraph> ftd_gC;
    int                        ftd_chain_max = 0;  // longest chain (rounds); 0: no chains
    // the last rounds in one cooperative launch (csg_ftd_tail.h)
    uint32_t      ftd_tail_max = 64;  // most active cells handed to it (0: never)
    bool          tt_ran = false;       // the last fine-tune ended in it
    int           tt_refused = 0;  // why it was last refused: 1 LDS, 2 no cooperative
                                   // launch, 3 no resident block, 4 launch refused
    int           tt_blocks_per_cu = 4;
    size_t        tt_lds_limit = 65536;
    FtdTailState* d_tt_state = nullptr;
    FtdTailState* h_tt_state = nullptr;  // pinned
    FtdTailSlot*  d_tt_slots = nullptr;
    FtdTailChunk* d_tt_chunks = nullptr;
    uint32_t*     d_tt_genes = nullptr;
    int16_t*      d_tt_crank = nullptr;
    float *       d_tt_scores = nullptr, *d_tt_P = nullptr;
    uint64_t*     d_tt_nanl = nullptr;
    size_t tt_cap_state = 0, tt_cap_slots = 0, tt_cap_chunks = 0, tt_cap_genes = 0,
           tt_cap_crank = 0, tt_cap_scores = 0, tt_cap_P = 0, tt_cap_nanl = 0;
    uint64_t                   ftd_chain_scores = 0;  // most scores of a chained round (0: any)
    // launch settings of the work-queue kernels (blocks), tunable
    int   ftd_grid_threads = 1024, ftd_grid_setup = 512, ftd_grid_rank = 2048,
        ftd_grid_corr = 512, ftd_grid_select = 2048, ftd_grid_fused = 4096, ftd_thin_max = 8,
        ftd_t64 = 48;
    int ftd_fuse_max = FTD_FUSE_MAX;  // groups of at most this many cells take the fused form
    bool ftd_eager = false;  // part B is being launched eagerly (not captured): phase events allowed
    bool     ftd_autotune = true, ftd_tuned = false;
    uint64_t ftd_budget   = (uint64_t)8 << 30;  // device bytes of one work batch
    int64_t  ftd_batches  = 0;                   // work batches of oversized rounds
    float ftd_tune_ms[4] = { 0.f, 0.f, 0.f, 0.f };  // best times of the tuned kernels

    // ---- learned proposal ----
    // Network G_s -> H -> H -> L (GELU), parameters in one FP32 array
    // [W1 (Kp x H) | b1 (H) | W2 (H x H) | b2 (H) | W3 (H x Lp) | b3 (Lp)],
    // row-major, input rows >= G_s and output columns >= L inert (zero).
    bool     mlp_ready = false;
    uint32_t mG = 0, mKp = 0, mH = 0, mL = 0, mLp = 0;
    size_t   m_off[6] = {}, m_n = 0;
    float *  d_mw = nullptr, *d_mg = nullptr, *d_mm = nullptr, *d_mv = nullptr;
    __bf16*  d_mwb = nullptr;   // BF16 copy of every parameter
    __bf16*  d_mwbT = nullptr;  // BF16 transposes of W1, W2, W3 (same offsets)
    int64_t  m_step = 0;       // Adam steps taken
    // activations (grow-only, in rows)
    size_t   m_rows_cap = 0, m_train_cap = 0;
    __bf16 * d_mx = nullptr, *d_mh1 = nullptr, *d_mh2 = nullptr, *d_mdlb = nullptr, *d_mdzb = nullptr;
    float *  d_macc = nullptr, *d_mz1 = nullptr, *d_mz2 = nullptr, *d_mdl = nullptr, *d_mdh = nullptr,
          *d_mdz = nullptr, *d_mloss = nullptr;
    int32_t* d_my = nullptr;
    // proposal work lists (allocated with the exact path's buffers)
    int        prop_k = 0, prop_kp = 0;  // proposal size, labels ranked per cell
    float      prop_margin = 0.2f;
    int32_t*   d_ptop = nullptr;
    float*     d_pmargin = nullptr;
    uint32_t * d_pcnt = nullptr, *d_pslot = nullptr, *d_pfb = nullptr, *d_pcells = nullptr;
    uint32_t*  d_plist = nullptr;  // cells the edge tests list (two lists of Qc)
    // per-round label tables of the grouped contraction, built on the device
    // by k_prop_plan, views into d_pmeta: [c_off L (u64) | a_base L | a_cnt L
    // | 48 bytes unused | tiles (bound: every label's list at Qc cells)]; the
    // round plans (d_pplan) are part of d_dcnt
    uint32_t * d_pabase = nullptr, *d_pacnt = nullptr, *d_pplan = nullptr;
    // the previous score step's device counts, read back with its closing
    // synchronisation, size the next one's grids (a kernel loops where a
    // count outgrew its grid): per proposal round the tiles and items, the
    // certification's pairs and tiles
    uint32_t   est_tiles[3] = {}, est_items[3] = {}, est_pairs = 0, est_rc_tiles = 0;
    uint64_t*  d_pcoff = nullptr;
    uint8_t*   d_pmeta = nullptr;
    size_t     pmeta_cap = 0;
    RecheckItem* d_pitems = nullptr;
    Segment*   d_pseg = nullptr;
    GroupTile* d_ptiles = nullptr;

    // ---- timing / counters ----
    double  phase_ms[CSG_N_PHASES]  = {};
    int64_t phase_cnt[CSG_N_PHASES] = {};
    // 0 no phase events, 1 the correlation phase only, 2 all, 3 all and the
    // device fine-tune's kernel families (its part B launched eagerly)
    int     phase_mode = 2;
    std::vector<PhaseEvents> pending;
    std::vector<hipEvent_t>  event_pool;
    int64_t counters[CSG_N_COUNTERS] = {};

    int fail(int code, const std::string& m)
    {
        err = m;
        return code;
    }
};

#define HIPCHK(call)                                                                           \
    do                                                                                         \
    {                                                                                          \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return eng->fail(CSG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK()                                                                            \
    do                                                                                         \
    {                                                                                          \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess)                                                                  \
            return eng->fail(CSG_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

namespace {

template <class T> void dfree(T*& p)
{
    if (p)
        (void)hipFree((void*)p);
    p = nullptr;
}

template <class T> int dalloc(csg_engine* eng, T** p, size_t count)
{
    dfree(*p);
    if (count == 0)
        count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess)
    {
        *p = nullptr;
        return eng->fail(CSG_ERR_MEMORY, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                             " bytes: " + hipGetErrorString(e));
    }
    return CSG_OK;
}

template <class T> int upload(csg_engine* eng, T** p, const T* src, size_t count)
{
    int rc = dalloc(eng, p, count);
    if (rc)
        return rc;
    if (count)
        HIPCHK(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, eng->stream));
    return CSG_OK;
}

// Grow-only device allocation.
// Grid of a persistent or grid-stride kernel over device-counted work:
// per_cu blocks per CU, a multiple of 8 (XcdRange).
unsigned persistent_grid(const csg_engine* eng, int per_cu)
{
    const unsigned cu = (unsigned)std::max(1, eng->prop.multiProcessorCount);
    return (cu * per_cu + 7) / 8 * 8;
}

// Grid for `est` units of work (the previous step's count), `per_block`
// units per block, at least the persistent grid of per_cu blocks per CU; a
// multiple of 8 (XcdRange), with some room for growth.
unsigned est_grid(const csg_engine* eng, uint32_t est, uint32_t per_block, int per_cu)
{
    const uint64_t need = ((uint64_t)est * 33 / 32 + per_block - 1) / per_block + 8;
    return std::max<unsigned>(persistent_grid(eng, per_cu),
                              (unsigned)std::min<uint64_t>(need / 8 * 8 + 8, 1u << 30));
}

template <class T> int ensure(csg_engine* eng, T** p, size_t* cap, size_t n)
{
    if (n <= *cap && *p)
        return CSG_OK;
    const size_t want = std::max(n, *cap + *cap / 2);
    int          rc   = dalloc(eng, p, want);
    *cap              = rc ? 0 : want;
    return rc;
}

// Grow-only upload: `count` elements into *p (capacity *cap).
template <class T> int upload_keep(csg_engine* eng, T** p, size_t* cap, const T* src, size_t count)
{
    int rc = ensure(eng, p, cap, std::max<size_t>(count, 1));
    if (rc)
        return rc;
    if (count)
        HIPCHK(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, eng->stream));
    return CSG_OK;
}

void free_proposal(csg_engine* eng);
void free_mlp(csg_engine* eng);
int ftd_free(csg_engine* eng);

hipEvent_t take_event(csg_engine* eng)
{
    if (!eng->event_pool.empty())
    {
        hipEvent_t e = eng->event_pool.back();
        eng->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct PhaseScope
{
    csg_engine* eng;
    PhaseEvents pe;
    bool        on;
    PhaseScope(csg_engine* e, int phase, int launches)
        : eng(e), on(e->phase_mode >= 2 || (e->phase_mode == 1 && phase == 1))
    {
        if (!on)
            return;
        pe.phase    = phase;
        pe.a        = take_event(e);
        pe.b        = take_event(e);
        pe.launches = launches;
        (void)hipEventRecord(pe.a, e->stream);
    }
    ~PhaseScope()
    {
        if (!on)
            return;
        (void)hipEventRecord(pe.b, eng->stream);
        eng->pending.push_back(pe);
    }
};

// Folds finished phase events into the totals (after a stream sync).
void collect_phases(csg_engine* eng)
{
    for (auto& pe : eng->pending)
    {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.a, pe.b) == hipSuccess)
        {
            eng->phase_ms[pe.phase] += ms;
            eng->phase_cnt[pe.phase] += pe.launches;
        }
        eng->event_pool.push_back(pe.a);
        eng->event_pool.push_back(pe.b);
    }
    eng->pending.clear();
}

int validate_csr(csg_engine* eng, const char* what, uint32_t rows, uint32_t width,
                 const uint16_t* data, const int32_t* ind, const int32_t* ptr, uint64_t nbins)
{
    if (ptr[0] != 0)
        return eng->fail(CSG_ERR_ARGUMENT, std::string(what) + ": indptr[0] != 0");
    for (uint32_t r = 0; r < rows; ++r)
        if (ptr[r + 1] < ptr[r])
            return eng->fail(CSG_ERR_ARGUMENT, std::string(what) + ": indptr not monotone");
    const int64_t nnz = ptr[rows];
    for (int64_t i = 0; i < nnz; ++i)
    {
        if (ind[i] < 0 || (uint32_t)ind[i] >= width)
            return eng->fail(CSG_ERR_ARGUMENT, std::string(what) + ": column index out of range");
        if (data[i] == 0 || data[i] >= nbins)
            return eng->fail(CSG_ERR_ARGUMENT,
                             std::string(what) + ": value code outside [1, max_uniq_gene)");
    }
    return CSG_OK;
}

int check_ctidx(csg_engine* eng, uint32_t ref_height, uint32_t L, const uint32_t* ctidx)
{
    if (L == 0 || L > 256)
        return eng->fail(CSG_ERR_ARGUMENT, "ct_num must be in [1, 256] (labels are uint8 codes)");
    uint64_t pos = 0;
    for (uint32_t l = 0; l < L; ++l)
    {
        if (ctidx[2 * l] != pos)
            return eng->fail(CSG_ERR_ARGUMENT, "ctidx ranges are not contiguous");
        if (ctidx[2 * l + 1] == 0)
            return eng->fail(CSG_ERR_UNDEFINED,
                             "label " + std::to_string(l) +
                                 " has no reference cell (undefined in the reference)");
        pos += ctidx[2 * l + 1];
    }
    if (pos != ref_height)
        return eng->fail(CSG_ERR_ARGUMENT, "ctidx does not cover the reference rows");
    return CSG_OK;
}

int set_pct(csg_engine* eng, PctParam** d, float* cached, const std::vector<uint32_t>& ctidx,
            uint32_t L, float q)
{
    if (*d && *cached == q)
        return CSG_OK;
    std::vector<PctParam> pp(L);
    for (uint32_t l = 0; l < L; ++l)
        pp[l] = make_pct_param((long)ctidx[2 * l + 1], q);
    int rc = upload(eng, d, pp.data(), pp.size());
    if (rc)
        return rc;
    *cached = q;
    return CSG_OK;
}

size_t rank_smem(int nbins, uint32_t Kp) { return (size_t)((nbins + 3) & ~3) * 4 + (size_t)Kp * 2; }

// Lengths of listed segments (for sizing the NaN work buffer).
template <class Src>
__global__ void k_seg_len(Src src, const uint64_t* __restrict__ items, unsigned int n,
                          uint32_t* __restrict__ out)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = src.get(items[i]).n;
}

// Runs the libstdc++ sort emulation on the NaN-carrying segments found by a
// k_select launch (listed in eng->d_nan_list, counted in eng->d_nan_count).
template <class Src>
int resolve_nan_segments(csg_engine* eng, Src src, const float* scores, const PctParam* params,
                         float p, float* out)
{
    unsigned int cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, eng->d_nan_count, sizeof(cnt), hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    if (cnt == 0)
        return CSG_OK;
    eng->counters[5] += cnt;
    uint32_t* d_len = nullptr;
    int       rc    = dalloc(eng, &d_len, cnt);
    if (rc)
        return rc;
    k_seg_len<Src><<<(cnt + 255) / 256, 256, 0, eng->stream>>>(src, eng->d_nan_list, cnt, d_len);
    std::vector<uint32_t> len(cnt);
    hipError_t e1 = hipGetLastError();
    hipError_t e2 = hipMemcpyAsync(len.data(), d_len, cnt * 4, hipMemcpyDeviceToHost, eng->stream);
    hipError_t e3 = hipStreamSynchronize(eng->stream);
    dfree(d_len);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
        return eng->fail(CSG_ERR_DEVICE, "NaN segment sizing failed");
    std::vector<uint64_t> off(cnt);
    uint64_t              total = 0;
    for (unsigned int i = 0; i < cnt; ++i)
    {
        off[i] = total;
        total += len[i];
    }
    float*    d_work = nullptr;
    uint64_t* d_off  = nullptr;
    if ((rc = dalloc(eng, &d_work, total)) || (rc = upload(eng, &d_off, off.data(), off.size())))
    {
        dfree(d_work);
        return rc;
    }
    HIPCHK(hipMemsetAsync(eng->d_status, 0, sizeof(int), eng->stream));
    k_select_nan<Src><<<(cnt + 63) / 64, 64, 0, eng->stream>>>(src, eng->d_nan_list, cnt, scores,
                                                               params, p, out, d_work, d_off,
                                                               eng->d_status);
    LAUNCHCHK();
    int status = 0;
    HIPCHK(hipMemcpyAsync(&status, eng->d_status, sizeof(int), hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    dfree(d_work);
    dfree(d_off);
    if (status)
        return eng->fail(CSG_ERR_UNDEFINED,
                         "a NaN-carrying score segment makes the reference's std::sort read out "
                         "of bounds; its percentile is undefined");
    return CSG_OK;
}

}  // namespace

// ===================================================================== C-ABI

extern "C" {

int csg_init_gpu(int gpuid, csg_engine** out)
{
    if (!out)
        return CSG_ERR_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || gpuid < 0 || gpuid >= n)
        return CSG_ERR_DEVICE;
    csg_engine* eng = new csg_engine();
    eng->device     = gpuid;
    if (hipSetDevice(gpuid) != hipSuccess || hipGetDeviceProperties(&eng->prop, gpuid) != hipSuccess ||
        hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking) != hipSuccess)
    {
        delete eng;
        return CSG_ERR_DEVICE;
    }
    // this device's copy of the allocation-free entry point's work memory
    if (hipGetSymbolAddress((void**)&eng->noalloc_scratch, HIP_SYMBOL(scratch)) != hipSuccess ||
        !eng->noalloc_scratch)
    {
        (void)hipGetLastError();
        (void)hipStreamDestroy(eng->stream);
        delete eng;
        return CSG_ERR_DEVICE;
    }
    // larger dynamic LDS for the fine-tune kernels where the device allows it
    // (static LDS of the kernels stays below 2 KiB)
    eng->lds_limit = 65536;
    if ((size_t)eng->prop.sharedMemPerBlock > 65536 + 2048)
    {
        const int  want = (int)eng->prop.sharedMemPerBlock - 2048;
        const bool ok1  = hipFuncSetAttribute((const void*)k_ft_corr<FT_TC, FT_TR, FT_RC, FT_RR>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              want) == hipSuccess;
        const bool ok2  = hipFuncSetAttribute((const void*)k_ft_rank_qry,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              want) == hipSuccess;
        bool       ok3  = true;
        for (const void* f : { (const void*)k_ftd_rank_ref, (const void*)k_ftd_rank_qry,
                               (const void*)k_ftd_corr, (const void*)k_ft_rank_ref,
                               (const void*)k_code_rows_big<int32_t>,
                               (const void*)k_code_rows_big<uint16_t>, (const void*)k_ftd_setup })
            ok3 = ok3 && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             want) == hipSuccess;
        if (ok1 && ok2 && ok3)
            eng->lds_limit = (size_t)want;
        (void)hipGetLastError();  // a refused attribute must not surface at the next launch
    }
    // the fine-tune tail kernel: as much dynamic LDS as the device gives a
    // block beside the kernel's static LDS (its sort stack: about 3.4 KB)
    eng->tt_lds_limit = 65536;
    {
        hipFuncAttributes fa8{}, fa16{};
        if (hipFuncGetAttributes(&fa8, (const void*)k_ftd_tail<uint8_t>) == hipSuccess &&
            hipFuncGetAttributes(&fa16, (const void*)k_ftd_tail<uint16_t>) == hipSuccess)
        {
            const size_t st   = std::max(fa8.sharedSizeBytes, fa16.sharedSizeBytes);
            const size_t have = (size_t)eng->prop.sharedMemPerBlock;
            if (have > 65536 + st)
            {
                const int want = (int)(have - st);
                if (hipFuncSetAttribute((const void*)k_ftd_tail<uint8_t>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        want) == hipSuccess &&
                    hipFuncSetAttribute((const void*)k_ftd_tail<uint16_t>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        want) == hipSuccess)
                    eng->tt_lds_limit = (size_t)want;
            }
        }
        (void)hipGetLastError();
    }
    // the wide INT8 kernels stage three K steps of operands in dynamic LDS
    eng->w8_ok = (size_t)eng->prop.sharedMemPerBlock >= (size_t)W8_LDS &&
                 hipFuncSetAttribute((const void*)k_gemm_int8w<false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     W8_LDS) == hipSuccess &&
                 hipFuncSetAttribute((const void*)k_gemm_int8w_group<false, false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     W8_LDS) == hipSuccess &&
                 hipFuncSetAttribute((const void*)k_gemm_int8w_group<false, true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     W8_LDS) == hipSuccess;
    eng->kf_ok = eng->w8_ok && (size_t)eng->prop.sharedMemPerBlock >= (size_t)W8_LDS_KF &&
                 hipFuncSetAttribute((const void*)k_gemm_int8w<true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     W8_LDS_KF) == hipSuccess &&
                 hipFuncSetAttribute((const void*)k_gemm_int8w_group<true, false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     W8_LDS_KF) == hipSuccess &&
                 hipFuncSetAttribute((const void*)k_gemm_int8w_group<true, true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     W8_LDS_KF) == hipSuccess;
    eng->x3_ok = (size_t)eng->prop.sharedMemPerBlock >= (size_t)X3_LDS &&
                 hipFuncSetAttribute((const void*)k_gemm_int8x,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     X3_LDS) == hipSuccess &&
                 hipFuncSetAttribute((const void*)k_gemm_int8x_group,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     X3_LDS) == hipSuccess;
    eng->bw_ok = (size_t)eng->prop.sharedMemPerBlock >= (size_t)BW_LDS &&
                 hipFuncSetAttribute((const void*)k_bgemm_nt_w,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     BW_LDS) == hipSuccess;
    (void)hipGetLastError();
    {
        const int cu          = std::max(1, eng->prop.multiProcessorCount);
        eng->ftd_grid_threads = cu * 4;
        eng->ftd_grid_setup   = cu * 4;
        eng->ftd_grid_rank    = cu * 8;
        eng->ftd_grid_fused   = cu * 16;
        eng->ftd_grid_corr    = cu * 2;
        eng->ftd_grid_select  = cu * 8;
    }
    if (dalloc(eng, &eng->d_nan_count, 1) || dalloc(eng, &eng->d_status, 1) ||
        dalloc(eng, &eng->d_dcnt, 48) ||
        hipHostMalloc((void**)&eng->h_dcnt, 48 * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess)
    {
        delete eng;
        return CSG_ERR_MEMORY;
    }
    *out = eng;
    return CSG_OK;
}

const char* csg_last_error(const csg_engine* eng) { return eng ? eng->err.c_str() : "no engine"; }

int csg_destroy_score(csg_engine* eng)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    (void)hipSetDevice(eng->device);
    (void)hipStreamSynchronize(eng->stream);
    dfree(eng->d_rdata);
    dfree(eng->d_qdata);
    dfree(eng->d_rind);
    dfree(eng->d_rptr);
    dfree(eng->d_qind);
    dfree(eng->d_qptr);
    dfree(eng->d_lstart);
    dfree(eng->d_llen);
    dfree(eng->d_pct);
    dfree(eng->d_Xr);
    dfree(eng->d_sqr_f);
    dfree(eng->d_irr);
    dfree(eng->d_Xq);
    dfree(eng->d_sqq_f);
    dfree(eng->d_irq);
    dfree(eng->d_S);
    dfree(eng->d_nan_list);
    dfree(eng->d_rh);
    dfree(eng->d_rl);
    dfree(eng->d_qh);
    dfree(eng->d_ql);
    dfree(eng->d_dr);
    dfree(eng->d_dq);
    dfree(eng->d_rf16);
    dfree(eng->d_qf16);
    dfree(eng->d_cstatus);
    dfree(eng->d_exact);
    dfree(eng->d_items);
    dfree(eng->d_nitems);
    dfree(eng->d_Pref);
    dfree(eng->d_item_seg);
    dfree(eng->d_rc_cells);
    dfree(eng->d_rc_abase);
    dfree(eng->d_rc_sorted);
    dfree(eng->d_rc_coff);
    dfree(eng->d_rc_tiles);
    dfree(eng->d_rc_ucells);
    eng->rc_cap_ucells = eng->rc_cap_sorted = 0;
    eng->rc_cap_seg = eng->rc_cap_cells = eng->rc_cap_l =
        eng->rc_cap_coff = eng->rc_cap_tiles = 0;
    free_proposal(eng);
    eng->ref_prepared_path = -1;
    eng->pct_quantile = -1.f;
    eng->score_loaded = false;
    eng->ref_ranked   = false;
    return CSG_OK;
}

int csg_destroy(csg_engine* eng)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    (void)hipSetDevice(eng->device);
    (void)hipStreamSynchronize(eng->stream);
    dfree(eng->d_fref_codes);
    dfree(eng->d_fref_codes8);
    dfree(eng->d_flstart);
    dfree(eng->d_fllen);
    dfree(eng->d_fqdata);
    dfree(eng->d_fqind);
    dfree(eng->d_fqptr);
    dfree(eng->d_fpct);
    dfree(eng->d_ft_tables);
    dfree(eng->d_ft_rr);
    dfree(eng->d_ft_cr);
    dfree(eng->d_ft_rsq);
    dfree(eng->d_ft_csq);
    dfree(eng->d_ft_scores);
    dfree(eng->d_ft_P);
    dfree(eng->d_ft_nan);
    dfree(eng->d_ft_out);
    dfree(eng->d_ft_redo);
    eng->d_ft_redo_cap = 0;
    if (eng->h_ft_redo)
        (void)hipHostFree(eng->h_ft_redo);
    eng->h_ft_redo     = nullptr;
    eng->h_ft_redo_cap = 0;
    eng->d_ft_tables_cap = eng->d_ft_rr_cap = eng->d_ft_cr_cap = eng->d_ft_rsq_cap = 0;
    eng->d_ft_csq_cap = eng->d_ft_scores_cap = eng->d_ft_P_cap = eng->d_ft_nan_cap = 0;
    eng->d_ft_out_cap = 0;
    if (eng->h_ft_stage)
        (void)hipHostFree(eng->h_ft_stage);
    if (eng->h_ft_out)
        (void)hipHostFree(eng->h_ft_out);
    eng->h_ft_stage = eng->h_ft_out = nullptr;
    eng->h_ft_stage_cap = eng->h_ft_out_cap = 0;
    ftd_free(eng);
    eng->fpct_quantile = -1.f;
    eng->ft_loaded     = false;
    eng->window_cache.clear();
    eng->host_mask.clear();
    return CSG_OK;
}

void csg_release(csg_engine* eng)
{
    if (!eng)
        return;
    csg_destroy_score(eng);
    csg_destroy(eng);
    free_mlp(eng);
    if (eng->copy_stream)
        (void)hipStreamSynchronize(eng->copy_stream);
    for (int b = 0; b < 2; ++b)
    {
        dfree(eng->d_raw_data[b]);
        dfree(eng->d_raw_idx[b]);
        if (eng->code_copied[b])
            (void)hipEventDestroy(eng->code_copied[b]);
        if (eng->code_done[b])
            (void)hipEventDestroy(eng->code_done[b]);
    }
    dfree(eng->d_raw_ptr);
    for (auto& ps : eng->pre)
    {
        dfree(ps.d_data);
        dfree(ps.d_idx);
        if (ps.copied)
            (void)hipEventDestroy(ps.copied);
    }
    dfree(eng->d_map_s);
    dfree(eng->d_map_f);
    dfree(eng->d_code_cnt);
    dfree(eng->d_code_over);
    dfree(eng->d_code_misc);
    if (eng->h_code_misc)
        (void)hipHostFree(eng->h_code_misc);
    if (eng->copy_stream)
        (void)hipStreamDestroy(eng->copy_stream);
    dfree(eng->d_P);
    dfree(eng->d_first);
    dfree(eng->d_mask);
    dfree(eng->d_nan_count);
    dfree(eng->d_dcnt);
    if (eng->h_dcnt)
        (void)hipHostFree(eng->h_dcnt);
    eng->h_dcnt = nullptr;
    dfree(eng->d_status);
    collect_phases(eng);
    for (auto e : eng->event_pool)
        (void)hipEventDestroy(e);
    if (eng->stream)
        (void)hipStreamDestroy(eng->stream);
    delete eng;
}

int csg_copyin_score(csg_engine* eng, uint32_t ref_height, uint32_t width, const uint16_t* ref_data,
                     const int32_t* ref_indices, const int32_t* ref_indptr, uint32_t qry_height,
                     const uint16_t* qry_data, const int32_t* qry_indices,
                     const int32_t* qry_indptr, uint32_t ct_num, const uint32_t* ctidx,
                     uint64_t max_uniq_gene)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    HIPCHK(hipSetDevice(eng->device));
    csg_destroy_score(eng);
    if (width == 0 || width > 32767)
        return eng->fail(CSG_ERR_ARGUMENT, "score gene set size must be in [1, 32767]");
    if (ref_height == 0 || qry_height == 0)
        return eng->fail(CSG_ERR_ARGUMENT, "empty reference or query matrix");
    if (max_uniq_gene < 1 || max_uniq_gene > 16384)
        return eng->fail(CSG_ERR_ARGUMENT, "max_uniq_gene must be in [1, 16384]");
    int rc = check_ctidx(eng, ref_height, ct_num, ctidx);
    if (rc)
        return rc;
    // no query data: reserve room for query batches of up to qry_height rows
    // (csg_copyin_query_raw supplies them)
    const bool reserve = !qry_data && !qry_indices && !qry_indptr;
    if ((rc = validate_csr(eng, "reference", ref_height, width, ref_data, ref_indices, ref_indptr,
                           max_uniq_gene)) ||
        (!reserve && (rc = validate_csr(eng, "query", qry_height, width, qry_data, qry_indices,
                                        qry_indptr, max_uniq_gene))))
        return rc;
    eng->Nr    = ref_height;
    eng->Nq    = qry_height;
    eng->G     = width;
    eng->Kp    = (uint32_t)round_up(width, 64);
    eng->L     = ct_num;
    eng->nbins = (int)max_uniq_gene;
    if (rank_smem(eng->nbins, eng->Kp) > 65536)
        return eng->fail(CSG_ERR_ARGUMENT, "gene set and code range exceed the rank kernel's LDS");
    eng->ctidx.assign(ctidx, ctidx + 2 * ct_num);
    eng->nbins_ref = (int)max_uniq_gene;
    eng->q_cap     = qry_height;
    eng->q_ready   = !reserve;
    const size_t rnnz = (size_t)ref_indptr[ref_height];
    const size_t qnnz = reserve ? 0 : (size_t)qry_indptr[qry_height];
    eng->q_nnz_cap    = qnnz;
    if ((rc = upload(eng, &eng->d_rdata, ref_data, rnnz)) ||
        (rc = upload(eng, &eng->d_rind, ref_indices, rnnz)) ||
        (rc = upload(eng, &eng->d_rptr, ref_indptr, (size_t)ref_height + 1)) ||
        (!reserve && ((rc = upload(eng, &eng->d_qdata, qry_data, qnnz)) ||
                      (rc = upload(eng, &eng->d_qind, qry_indices, qnnz)) ||
                      (rc = upload(eng, &eng->d_qptr, qry_indptr, (size_t)qry_height + 1)))))
        return rc;
    std::vector<uint32_t> ls(ct_num), ll(ct_num);
    for (uint32_t l = 0; l < ct_num; ++l)
    {
        ls[l] = ctidx[2 * l];
        ll[l] = ctidx[2 * l + 1];
    }
    if ((rc = upload(eng, &eng->d_lstart, ls.data(), ls.size())) ||
        (rc = upload(eng, &eng->d_llen, ll.data(), ll.size())))
        return rc;
    // reference ranks (rows padded to the 128-row tile)
    const size_t nr_pad = round_up(ref_height, 256);
    if ((rc = dalloc(eng, &eng->d_Xr, nr_pad * eng->Kp)) ||
        (rc = dalloc(eng, &eng->d_sqr_f, nr_pad)) || (rc = dalloc(eng, &eng->d_irr, nr_pad)))
        return rc;
    HIPCHK(hipMemsetAsync(eng->d_Xr, 0, nr_pad * eng->Kp * sizeof(int16_t), eng->stream));
    // query chunks: the score buffer takes at most a quarter of the free
    // device memory.  The proposal path's scores are addressed by 64-bit
    // segment offsets, so its chunk (Qc rows, every score of a chunk fitting
    // the buffer) is bounded by memory only; a dense score matrix chunk
    // (Qc_dense rows) also holds at most 2^31 - 2^20 scores, so that every
    // chunk-local dense index stays below 2^31.
    size_t budget = (size_t)1 << 36;
    {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b / 16 < budget)
            budget = std::max<size_t>(free_b / 16, (size_t)1 << 26);
    }
    const size_t dense_budget = std::min(budget, ((size_t)1 << 31) - ((size_t)1 << 20));
    auto rows_for = [&](size_t scores) {
        size_t q = scores / (size_t)ref_height;
        q        = std::max<size_t>(128, q / 128 * 128);
        q        = std::min<size_t>(q, round_up(qry_height, 128));
        if (eng->qc_cap)
            q = std::min<size_t>(q, eng->qc_cap);
        return q;
    };
    // (and a chunk's INT8 operand plane, Qc x Kp bytes, below 2^32: the wide
    // kernels' 32-bit row offsets)
    const size_t qc     = std::min(rows_for(budget),
                                   std::max<size_t>(128, (((size_t)1 << 32) - 1) / eng->Kp / 128 * 128));
    eng->Qc             = (uint32_t)qc;
    eng->Qc_dense       = (uint32_t)std::min(qc, rows_for(dense_budget));
    if ((rc = dalloc(eng, &eng->d_Xq, qc * eng->Kp)) || (rc = dalloc(eng, &eng->d_sqq_f, qc)) ||
        (rc = dalloc(eng, &eng->d_irq, qc)) || (rc = dalloc(eng, &eng->d_S, qc * ref_height)) ||
        (rc = dalloc(eng, &eng->d_nan_list, qc * ct_num)) ||
        (rc = dalloc(eng, &eng->d_P, (size_t)qry_height * ct_num)) ||
        (rc = dalloc(eng, &eng->d_first, qry_height)) ||
        (rc = dalloc(eng, &eng->d_mask, (size_t)qry_height * ct_num)))
        return rc;
    HIPCHK(hipMemsetAsync(eng->d_Xq, 0, qc * eng->Kp * sizeof(int16_t), eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    eng->mask_valid   = false;
    eng->score_loaded = true;
    return CSG_OK;
}

namespace {

// Score-step ranks of rows [row0, row0 + nrows) (k_rank_csr_w, one wave per
// row, when its per-wave LDS fits 16 KiB; else one block per row).  sq_f
// (the FP32 sums of squares) and feat (the network features) may be null.
int rank_rows(csg_engine* eng, const uint16_t* data, const int* ind, const int* ptr, uint32_t row0,
              uint32_t nrows, int16_t* X, float* sq_f, double* ir, __bf16* feat = nullptr)
{
    if (nrows == 0)
        return CSG_OK;
    const float inv_g     = 1.f / (float)std::max<uint32_t>(eng->G, 1);
    const int   wave_ints = ((eng->nbins + 3) & ~3) + (int)eng->Kp / 2;
    if ((size_t)wave_ints * 4 * 4 <= 65536)
    {
        const unsigned grid = (unsigned)std::min<size_t>((nrows + 3) / 4, 16384);
        k_rank_csr_w<<<grid, 256, (size_t)wave_ints * 4 * 4, eng->stream>>>(
            data, ind, ptr, row0, nrows, eng->G, eng->Kp, eng->nbins, wave_ints, X, sq_f, ir,
            feat, inv_g);
    }
    else
    {
        const size_t smem = rank_smem(eng->nbins, eng->Kp);
        k_rank_csr<256><<<nrows, 256, smem, eng->stream>>>(data, ind, ptr, row0, nrows, eng->G,
                                                           eng->Kp, eng->nbins, X, sq_f, ir, feat,
                                                           inv_g);
    }
    LAUNCHCHK();
    return CSG_OK;
}

int score_chunk(csg_engine* eng, int path, uint32_t nq)
{
    if (path == CSG_PATH_PLAIN)
    {
        dim3 grid((eng->Nr + PG_BN - 1) / PG_BN, (nq + PG_BM - 1) / PG_BM);
        k_gemm_plain<<<grid, 256, 0, eng->stream>>>(eng->d_Xq, eng->d_Xr, (int)eng->Kp,
                                                    eng->d_sqq_f, eng->d_sqr_f, (int)nq,
                                                    (int)eng->Nr, eng->d_S, (size_t)eng->Nr);
    }
    else if (path == CSG_PATH_EXACT_INT8 && eng->ref_kf)
    {
        dim3 grid((eng->Nr + W8_BN - 1) / W8_BN, (nq + W8_BM - 1) / W8_BM);
        k_gemm_int8w<true><<<grid, 512, W8_LDS_KF, eng->stream>>>(
            eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl, (int)eng->Kp, eng->d_irq, eng->d_irr,
            (int)nq, (int)eng->Nr, eng->d_S, (size_t)eng->Nr, eng->d_dq, eng->d_dr);
    }
    else if (path == CSG_PATH_EXACT_INT8 && eng->G <= 1920 && eng->w8_ok)
    {
        dim3 grid((eng->Nr + W8_BN - 1) / W8_BN, (nq + W8_BM - 1) / W8_BM);
        k_gemm_int8w<false><<<grid, 512, W8_LDS, eng->stream>>>(
            eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl, (int)eng->Kp, eng->d_irq, eng->d_irr,
            (int)nq, (int)eng->Nr, eng->d_S, (size_t)eng->Nr, nullptr, nullptr);
    }
    else if (path == CSG_PATH_EXACT_INT8 && eng->x3_ok)
    {
        // gene sets beyond the folded limbs: three partials, 128 x 128 blocks
        dim3 grid((eng->Nr + X3_BN - 1) / X3_BN, (nq + X3_BM - 1) / X3_BM);
        k_gemm_int8x<<<grid, 512, X3_LDS, eng->stream>>>(eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl,
                                                    (int)eng->Kp, eng->d_irq, eng->d_irr,
                                                    (int)nq, (int)eng->Nr, eng->d_S,
                                                    (size_t)eng->Nr);
    }
    else if (path == CSG_PATH_EXACT_INT8)
    {
        dim3 grid((eng->Nr + I8_BN - 1) / I8_BN, (nq + I8_BM - 1) / I8_BM);
        k_gemm_int8<<<grid, 256, 0, eng->stream>>>(eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl,
                                                   (int)eng->Kp, eng->d_irq, eng->d_irr,
                                                   (int)nq, (int)eng->Nr, eng->d_S,
                                                   (size_t)eng->Nr);
    }
    else if (path == CSG_PATH_FP16_MFMA)
    {
        dim3 grid((eng->Nr + H_BN - 1) / H_BN, (nq + H_BM - 1) / H_BM);
        k_gemm_f16<<<grid, 256, 0, eng->stream>>>(eng->d_qf16, eng->d_rf16, (int)eng->Kp,
                                                  eng->d_sqq_f, eng->d_sqr_f, (int)nq,
                                                  (int)eng->Nr, eng->d_S, (size_t)eng->Nr);
    }
    else
    {
        dim3 grid((eng->Nr + F_BN - 1) / F_BN, (nq + F_BM - 1) / F_BM);
        k_gemm_f32mfma<<<grid, 256, 0, eng->stream>>>(eng->d_Xq, eng->d_Xr, (int)eng->Kp,
                                                      eng->d_sqq_f, eng->d_sqr_f, (int)nq,
                                                      (int)eng->Nr, eng->d_S, (size_t)eng->Nr);
    }
    LAUNCHCHK();
    return CSG_OK;
}

// The exact path's limb form for the loaded gene set: the three-product form
// (k_split_kf) needs |X| <= 1007, i.e. G <= 1008, and the wide kernels.
bool use_kf(const csg_engine* eng)
{
    return eng->limb_form == 0 && eng->kf_ok && eng->G <= 1008;
}

// Operand planes of `rows` rows of X for the MFMA paths; the exact path's in
// the three-product form (with the rows' corrections in d) when kf.
int prepare_operands(csg_engine* eng, int path, const int16_t* X, size_t rows, int8_t* h,
                     int8_t* l, _Float16* f, bool kf = false, double* d = nullptr)
{
    const size_t n = rows * eng->Kp;
    const unsigned blocks = (unsigned)((n / 8 + 255) / 256);
    if (path == CSG_PATH_EXACT_INT8 && kf)
        k_split_kf<<<(unsigned)((rows + 3) / 4), 256, 0, eng->stream>>>(X, rows, (int)eng->Kp, h,
                                                                       l, d);
    else if (path == CSG_PATH_EXACT_INT8)
        k_split_limbs<<<blocks, 256, 0, eng->stream>>>(X, n, h, l);
    else if (path == CSG_PATH_FP16_MFMA)
        k_to_half<<<blocks, 256, 0, eng->stream>>>(X, n, f);
    else
        return CSG_OK;
    LAUNCHCHK();
    return CSG_OK;
}

// Half-width of the reference's per-label quantile around the exact path's
// (see csg_exact.h): per score, gamma_{G-1} / (1 - e_D) for the reference's
// FP32 sum of products, e_D / (1 - e_D) for its FP32 divisor (relative error
// e_D = gamma_{G-1} + 2u (1 + gamma_{G-1})) and 2u for the two roundings;
// plus 16u for the two
// float evaluations of the quantile formula.
double certify_width(uint32_t G)
{
    const double u     = 5.9604644775390625e-08;
    const double n     = G > 1 ? (double)(G - 1) : 0.0;
    const double gamma = n * u / (1.0 - n * u);
    const double ed    = gamma + 2 * u * (1.0 + gamma);
    const double wpair = (gamma + ed) / (1.0 - ed) + 2 * u;
    return wpair * (1.0 + 1e-6) + 16 * u;
}

int certify_chunk(csg_engine* eng, uint32_t q0, uint32_t nq, float quantile, float thre)
{
    const uint32_t L  = eng->L;
    float*         P  = eng->d_P + (size_t)q0 * L;
    int32_t*       F  = eng->d_first + q0;
    float*         M  = eng->d_mask + (size_t)q0 * L;
    const double   wp = certify_width(eng->G);
    // the re-score tables, for every (cell, label) pair of a chunk
    uint32_t max_len = 1;
    for (uint32_t l = 0; l < L; ++l)
        max_len = std::max(max_len, eng->ctidx[2 * l + 1]);
    const size_t pairs = (size_t)eng->Qc * L;
    const size_t tile_bound = (pairs / PG_GBM + L) * ((max_len + PG_BN - 1) / PG_BN);
    int rc;
    if ((rc = ensure(eng, &eng->d_item_seg, &eng->rc_cap_seg, pairs)) ||
        (rc = ensure(eng, &eng->d_rc_cells, &eng->rc_cap_cells, pairs)) ||
        (rc = ensure(eng, &eng->d_rc_sorted, &eng->rc_cap_sorted, pairs)) ||
        (rc = ensure(eng, &eng->d_rc_abase, &eng->rc_cap_l, 4 * (size_t)L + 2)) ||
        (rc = ensure(eng, &eng->d_rc_coff, &eng->rc_cap_coff, L)) ||
        (rc = ensure(eng, &eng->d_rc_tiles, &eng->rc_cap_tiles, tile_bound)) ||
        (rc = ensure(eng, &eng->d_rc_ucells, &eng->rc_cap_ucells, eng->Qc)))
        return rc;
    HIPCHK(hipMemsetAsync(eng->d_nitems, 0, 2 * sizeof(unsigned int), eng->stream));
    HIPCHK(hipMemsetAsync(eng->d_exact, 0, (size_t)nq * L, eng->stream));
    {
        PhaseScope ps(eng, 3, 1);
        k_label_certify<<<(nq + 3) / 4, 256, 0, eng->stream>>>(
            P, nq, L, wp, thre, F, M, eng->d_cstatus, eng->d_items, eng->d_nitems);
        LAUNCHCHK();
    }
    // The pairs to re-score (d_nitems[0], in d_items) grouped by label on the
    // device: label l's cells listed from abase[l], a cell's scores against
    // the label's rows contiguous from coff[l] on, so that every pair's
    // scores are one contiguous segment.  The counts stay on the device: the
    // tables are sized for every pair of the chunk (csg_get_label), the
    // kernels loop over the counted items and tiles.  A cell's pairs take at
    // most Nr score slots (one per reference row of its labels), so the
    // grouped scores fit the score buffer.
    uint32_t* n_pairs = eng->d_nitems;
    // one uint32 table: [cnt L | abase L | tbase L | fill L | ntiles | nu]
    uint32_t* cnt    = eng->d_rc_abase;
    uint32_t* abase  = cnt + L;
    uint32_t* tbase  = abase + L;
    uint32_t* fill   = tbase + L;
    uint32_t* ntiles = fill + L;
    uint32_t* nu     = ntiles + 1;
    const unsigned gs = est_grid(eng, eng->est_pairs, 256, 4);
    {
        PhaseScope ps(eng, 4, 3);
        k_fill_u32<<<1, 256, 0, eng->stream>>>(cnt, (size_t)4 * L + 2, 0u);
        k_rc_count<<<gs, 256, 0, eng->stream>>>(eng->d_items, n_pairs, cnt);
        k_rc_scan<<<1, 64, 0, eng->stream>>>(cnt, eng->d_llen, L, PG_GBM, PG_BN, abase,
                                             eng->d_rc_coff, tbase, fill, ntiles);
        k_rc_scatter<<<gs, 256, 0, eng->stream>>>(eng->d_items, n_pairs, abase, eng->d_rc_coff,
                                                  eng->d_llen, fill, eng->d_rc_cells,
                                                  eng->d_rc_sorted, eng->d_item_seg);
        k_rc_tiles<<<(L + 63) / 64, 64, 0, eng->stream>>>(cnt, eng->d_llen, tbase, L, PG_GBM,
                                                          PG_BN, eng->d_rc_tiles);
        k_rc_cells<<<(nq + 255) / 256, 256, 0, eng->stream>>>(eng->d_cstatus, nq,
                                                              eng->d_rc_ucells, nu);
        // the reference's FP32 sums of squares of the re-scored cells (the
        // exact path's rank pass leaves them out)
        k_sq_seq<<<(nq + 255) / 256, 256, 0, eng->stream>>>(eng->d_Xq, eng->Kp, eng->G,
                                                            eng->d_rc_ucells, nu, eng->d_sqq_f);
        k_gemm_plain_group<<<est_grid(eng, eng->est_rc_tiles, 1, 4), 256, 0, eng->stream>>>(
            eng->d_rc_tiles, eng->d_rc_cells, abase, cnt, eng->d_lstart, eng->d_llen,
            eng->d_rc_coff, eng->d_Xq, eng->d_Xr, (int)eng->Kp,
            (uint32_t)round_up(eng->Nr, 256) - 1, eng->d_sqq_f, eng->d_sqr_f, eng->d_S, ntiles);
        LAUNCHCHK();
        HIPCHK(hipMemsetAsync(eng->d_nan_count, 0, sizeof(unsigned int), eng->stream));
        TableSegments src{ eng->d_item_seg };
        k_select_n<TableSegments><<<est_grid(eng, eng->est_pairs, 4, 8), 256, 0, eng->stream>>>(
            src, n_pairs, eng->d_S, eng->d_pct, quantile, eng->d_Pref, eng->d_nan_list,
            eng->d_nan_count);
        LAUNCHCHK();
        // NaN-carrying segments: the libstdc++ sort emulation in place (an
        // undefined sort sets the flag the step's closing read-back checks)
        k_select_nan_inplace<TableSegments><<<64, 64, 0, eng->stream>>>(
            src, eng->d_nan_list, eng->d_nan_count, eng->d_S, eng->d_pct, quantile, eng->d_Pref,
            (int*)(eng->d_dcnt + 32));
        LAUNCHCHK();
    }
    {
        PhaseScope ps(eng, 4, 2);
        k_apply_recheck<<<gs, 256, 0, eng->stream>>>(eng->d_rc_sorted, n_pairs, eng->d_Pref, P, L,
                                                     eng->d_exact);
        LAUNCHCHK();
        k_label_final<<<(nq + 255) / 256, 256, 0, eng->stream>>>(P, eng->d_cstatus, eng->d_exact,
                                                                 nq, L, thre, F, M);
        LAUNCHCHK();
        k_cert_tally<<<1, 64, 0, eng->stream>>>(eng->d_nitems, eng->d_nan_count, eng->d_dcnt);
        LAUNCHCHK();
    }
    return CSG_OK;
}


// ------------------------------------------------------- learned proposal

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

extern "C++" {
template <bool TA, bool TB>
int bgemm(csg_engine* eng, const __bf16* A, int lda, const __bf16* B, int ldb, float* C, int ldc,
          int M, int N, int K)
{
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    k_bgemm<TA, TB><<<grid, 256, 0, eng->stream>>>(A, lda, B, ldb, C, ldc, M, N, K, 0.f);
    LAUNCHCHK();
    return CSG_OK;
}
}

// Activation buffers for `rows` rows; with `train` also those of the
// backward pass.
int mlp_work(csg_engine* eng, size_t rows, bool train)
{
    int          rc;
    const size_t H = eng->mH, Kp = eng->mKp, Lp = eng->mLp, W = std::max(H, Lp);
    if (rows > eng->m_rows_cap || !eng->d_mx)
    {
        eng->m_rows_cap = 0;
        if ((rc = dalloc(eng, &eng->d_mx, rows * Kp)) || (rc = dalloc(eng, &eng->d_macc, rows * W)) ||
            (rc = dalloc(eng, &eng->d_mh1, rows * H)) || (rc = dalloc(eng, &eng->d_mh2, rows * H)))
            return rc;
        eng->m_rows_cap = rows;
    }
    if (train && (rows > eng->m_train_cap || !eng->d_mz1))
    {
        eng->m_train_cap = 0;
        if ((rc = dalloc(eng, &eng->d_mz1, rows * H)) || (rc = dalloc(eng, &eng->d_mz2, rows * H)) ||
            (rc = dalloc(eng, &eng->d_mdl, rows * Lp)) || (rc = dalloc(eng, &eng->d_mdlb, rows * Lp)) ||
            (rc = dalloc(eng, &eng->d_mdh, rows * H)) || (rc = dalloc(eng, &eng->d_mdz, rows * H)) ||
            (rc = dalloc(eng, &eng->d_mdzb, rows * H)) || (rc = dalloc(eng, &eng->d_my, rows)) ||
            (rc = dalloc(eng, &eng->d_mloss, 1)))
            return rc;
        eng->m_train_cap = rows;
    }
    return CSG_OK;
}

// Forward pass over the BF16 rows in d_mx: logits without the output bias
// land in d_macc [rows x Lp]; the pre-activations z1, z2 are kept for the
// backward pass when `keep`.
int mlp_forward(csg_engine* eng, int rows, bool keep)
{
    const int     H = (int)eng->mH, Kp = (int)eng->mKp, Lp = (int)eng->mLp;
    const __bf16* wt = eng->d_mwbT;
    const float*  w  = eng->d_mw;
    const dim3    gh((H + 127) / 128, (rows + 127) / 128), gl((Lp + 127) / 128, (rows + 127) / 128);
    // the hidden layers in the pipelined wide kernel where its shapes and LDS allow
    const bool     wide = eng->bw_ok && H % BW_BN == 0 && Kp % BW_BK == 0 && H % BW_BK == 0;
    const unsigned gw   = (unsigned)(((rows + BW_BM - 1) / BW_BM) * (H / BW_BN));
    if (wide)
        k_bgemm_nt_w<<<gw, 512, BW_LDS, eng->stream>>>(eng->d_mx, Kp, wt + eng->m_off[0], Kp, rows,
                                                       H, Kp, H, w + eng->m_off[1],
                                                       keep ? eng->d_mz1 : nullptr, eng->d_mh1);
    else
        k_bgemm_nt<1><<<gh, 256, 0, eng->stream>>>(eng->d_mx, Kp, wt + eng->m_off[0], Kp, rows, H,
                                                    Kp, H, nullptr, w + eng->m_off[1],
                                                    keep ? eng->d_mz1 : nullptr, eng->d_mh1);
    LAUNCHCHK();
    if (wide)
        k_bgemm_nt_w<<<gw, 512, BW_LDS, eng->stream>>>(eng->d_mh1, H, wt + eng->m_off[2], H, rows,
                                                       H, H, H, w + eng->m_off[3],
                                                       keep ? eng->d_mz2 : nullptr, eng->d_mh2);
    else
        k_bgemm_nt<1><<<gh, 256, 0, eng->stream>>>(eng->d_mh1, H, wt + eng->m_off[2], H, rows, H,
                                                    H, H, nullptr, w + eng->m_off[3],
                                                    keep ? eng->d_mz2 : nullptr, eng->d_mh2);
    LAUNCHCHK();
    k_bgemm_nt<0><<<gl, 256, 0, eng->stream>>>(eng->d_mh2, H, wt + eng->m_off[4], H, rows, Lp, H,
                                                Lp, eng->d_macc, nullptr, nullptr, nullptr);
    LAUNCHCHK();
    return CSG_OK;
}

// BF16 transposes of the three weight matrices (the forward operands).
int mlp_transpose(csg_engine* eng)
{
    const int    Kp = (int)eng->mKp, H = (int)eng->mH, Lp = (int)eng->mLp;
    const int    R[3] = { Kp, H, H }, Cn[3] = { H, H, Lp };
    const size_t off[3] = { eng->m_off[0], eng->m_off[2], eng->m_off[4] };
    for (int i = 0; i < 3; ++i)
    {
        const dim3 g((Cn[i] + 31) / 32, (R[i] + 31) / 32);
        k_transpose_bf16<<<g, 256, 0, eng->stream>>>(eng->d_mw + off[i], R[i], Cn[i],
                                                     eng->d_mwbT + off[i]);
        LAUNCHCHK();
    }
    return CSG_OK;
}

// Gradients of the mean cross-entropy into d_mg, from the output gradient
// the softmax kernel left in d_mdl / d_mdlb and the forward pass's state.
int mlp_backward(csg_engine* eng, int rows)
{
    const int     H = (int)eng->mH, Kp = (int)eng->mKp, Lp = (int)eng->mLp;
    const __bf16* wb = eng->d_mwb;
    float*        g  = eng->d_mg;
    const size_t  n  = (size_t)rows * H;
    int           rc;
    if ((rc = bgemm<true, false>(eng, eng->d_mh2, H, eng->d_mdlb, Lp, g + eng->m_off[4], Lp, H, Lp,
                                 rows)))
        return rc;
    k_colsum<<<nblk(Lp), 256, 0, eng->stream>>>(eng->d_mdl, rows, Lp, g + eng->m_off[5]);
    LAUNCHCHK();
    if ((rc = bgemm<false, true>(eng, eng->d_mdlb, Lp, wb + eng->m_off[4], Lp, eng->d_mdh, H, rows,
                                 H, Lp)))
        return rc;
    k_gelu_back<<<nblk(n), 256, 0, eng->stream>>>(eng->d_mdh, eng->d_mz2, n, eng->d_mdz, eng->d_mdzb);
    LAUNCHCHK();
    if ((rc = bgemm<true, false>(eng, eng->d_mh1, H, eng->d_mdzb, H, g + eng->m_off[2], H, H, H,
                                 rows)))
        return rc;
    k_colsum<<<nblk(H), 256, 0, eng->stream>>>(eng->d_mdz, rows, H, g + eng->m_off[3]);
    LAUNCHCHK();
    if ((rc = bgemm<false, true>(eng, eng->d_mdzb, H, wb + eng->m_off[2], H, eng->d_mdh, H, rows, H,
                                 H)))
        return rc;
    k_gelu_back<<<nblk(n), 256, 0, eng->stream>>>(eng->d_mdh, eng->d_mz1, n, eng->d_mdz, eng->d_mdzb);
    LAUNCHCHK();
    if ((rc = bgemm<true, false>(eng, eng->d_mx, Kp, eng->d_mdzb, H, g + eng->m_off[0], H, Kp, H,
                                 rows)))
        return rc;
    k_colsum<<<nblk(H), 256, 0, eng->stream>>>(eng->d_mdz, rows, H, g + eng->m_off[1]);
    LAUNCHCHK();
    return CSG_OK;
}

// Loss and output gradient of the rows in d_mx with labels in d_my (the
// batch mean, added to d_mloss).
int mlp_loss(csg_engine* eng, int rows)
{
    k_softmax_xent<<<nblk(rows), 256, 0, eng->stream>>>(eng->d_macc, eng->d_mw + eng->m_off[5],
                                                       eng->d_my, rows, (int)eng->mL,
                                                       (int)eng->mLp, eng->d_mdl, eng->d_mdlb,
                                                       eng->d_mloss);
    LAUNCHCHK();
    return CSG_OK;
}

int mlp_adam(csg_engine* eng, float lr)
{
    eng->m_step++;
    const float c1 = (float)(1.0 - std::pow(0.9, (double)eng->m_step));
    const float c2 = (float)(1.0 - std::pow(0.999, (double)eng->m_step));
    k_adam<<<nblk(eng->m_n), 256, 0, eng->stream>>>(eng->d_mw, eng->d_mg, eng->d_mm, eng->d_mv,
                                                    eng->m_n, lr, 0.9f, 0.999f, 1e-8f, c1, c2,
                                                    eng->d_mwb);
    LAUNCHCHK();
    return mlp_transpose(eng);
}

int mlp_check(csg_engine* eng)
{
    if (!eng->mlp_ready)
        return eng->fail(CSG_ERR_STATE, "no proposal network: call csg_mlp_init first");
    if (!eng->score_loaded || eng->mG != eng->G || eng->mL != eng->L)
        return eng->fail(CSG_ERR_STATE,
                         "the proposal network's gene set / label count differ from the loaded data");
    return CSG_OK;
}

// Top-K labels and softmax margin of the `nq` ranked query rows in d_Xq.
// Proposals of nq query rows whose network features are already in d_mx
// (rank_rows writes them).
int mlp_propose(csg_engine* eng, uint32_t nq, int K, int32_t* top, float* margin)
{
    int rc;
    if ((rc = mlp_work(eng, nq, false)))
        return rc;
    if ((rc = mlp_forward(eng, (int)nq, false)))
        return rc;
    const float* b3 = eng->d_mw + eng->m_off[5];
    const int    L = (int)eng->mL, Lp = (int)eng->mLp;
    switch (K)
    {
#define CSG_TOPK(k)                                                                                \
    case k:                                                                                        \
        if (L <= 64)                                                                               \
            k_topk_rows<k><<<(nq + 127) / 128, 128, 0, eng->stream>>>(eng->d_macc, b3, (int)nq, L, \
                                                                     Lp, top, margin);             \
        else                                                                                       \
            k_topk<k><<<(nq + 3) / 4, 256, 0, eng->stream>>>(eng->d_macc, b3, (int)nq, L, Lp, top, \
                                                             margin);                              \
        break;
        CSG_TOPK(1)
        CSG_TOPK(2)
        CSG_TOPK(3)
        CSG_TOPK(4)
        CSG_TOPK(5)
        CSG_TOPK(6)
        CSG_TOPK(7)
        CSG_TOPK(8)
#undef CSG_TOPK
    default:
        return eng->fail(CSG_ERR_ARGUMENT, "proposal size must be in [1, 8]");
    }
    LAUNCHCHK();
    return CSG_OK;
}

// One grouped scoring round of a chunk, planned on the device (no host
// read-back).  Label l's list holds cw[l] cells (device counts; NULL: none)
// scored on the window [t0, t1) of their ranked labels (`win` names them, or
// NULL for every confident cell of the chunk; n_win bounds their number,
// *n_win_dev counts them when given), followed by the *n_all cells of `all`
// (NULL: none), scored on every label (`base`: cw, their slot offset).  Every
// (cell, label) item is contracted, its quantile selected and stored in the
// chunk's quantile rows P.  A label's list holds at most nq cells (a cell is
// in each list once), so the items and scores fit the tables and the score
// buffer, and the tiles the planned bound.
struct PropRound
{
    const uint32_t* cw        = nullptr;
    const uint32_t* win       = nullptr;
    uint32_t        n_win     = 0;
    const uint32_t* n_win_dev = nullptr;
    int             t0 = 0, t1 = 0;
    const uint32_t* all   = nullptr;
    const uint32_t* n_all = nullptr;
    const uint32_t* base  = nullptr;
};

// Bound on a round's tiles: every label's list at Qc cells, in the smallest
// tiles of the three grouped kernels.
size_t prop_tile_bound(const csg_engine* eng)
{
    size_t t = 0;
    for (uint32_t l = 0; l < eng->L; ++l)
        t += ((size_t)eng->ctidx[2 * l + 1] + I8_BN - 1) / I8_BN;
    return t * (((size_t)eng->Qc + I8_BM - 1) / I8_BM);
}

int prop_round(csg_engine* eng, const PropRound& r, int round, uint32_t nq, float quantile,
               float* P)
{
    const uint32_t L = eng->L;
    // wide 128 x 256 blocks with folded limbs up to 1920 genes, else the
    // three-partial 128 x 128 blocks (64 x 128 where their LDS is refused)
    const bool wide = eng->G <= 1920 && eng->w8_ok;
    const bool x3   = !wide && eng->x3_ok;
    const uint32_t tbm = wide ? W8_BM : x3 ? X3_BM : I8_BM, tbn = wide ? W8_BN : x3 ? X3_BN : I8_BN;
    uint32_t* ntiles = eng->d_pplan + 4 * round;
    uint32_t* nitems = ntiles + 1;
    const unsigned g_tiles = est_grid(eng, eng->est_tiles[round], 1, 1);
    const unsigned g_items = est_grid(eng, eng->est_items[round], 4, 8);
    k_prop_plan<<<32, 256, 0, eng->stream>>>(r.cw, r.n_all, eng->d_llen, L, tbm, tbn,
                                            eng->d_pabase, eng->d_pacnt, eng->d_pcoff,
                                            eng->d_ptiles, ntiles, eng->d_dcnt);
    LAUNCHCHK();
    {
        PhaseScope ps(eng, 9, 1);
        if (r.t1 > r.t0 && r.n_win)
        {
            k_prop_scatter_win<<<nblk(r.n_win), 256, 0, eng->stream>>>(
                eng->d_ptop, eng->d_pmargin, eng->prop_margin, r.win, r.n_win, r.n_win_dev,
                eng->prop_kp, r.t0, r.t1, L, eng->d_pslot, eng->d_pabase, eng->d_pcoff,
                eng->d_llen, eng->d_pcells, eng->d_pitems, eng->d_pseg);
            LAUNCHCHK();
        }
        if (r.all)
        {
            k_prop_scatter_all<<<nblk(nq), 256, 0, eng->stream>>>(
                r.all, nq, r.n_all, L, r.base, eng->d_pabase, eng->d_pcoff, eng->d_llen,
                eng->d_pcells, eng->d_pitems, eng->d_pseg);
            LAUNCHCHK();
        }
    }
    {
        PhaseScope ps(eng, 1, 1);
        const uint32_t b_max = (uint32_t)round_up(eng->Nr, 256) - 1;
        if (wide)
        {
            // one tile per block on the expected count's grid, then the tiles
            // past its capacity (none unless the count grew) in a looping form
            const bool    kf  = eng->ref_kf;
            const size_t  lds = kf ? W8_LDS_KF : W8_LDS;
            const double* dq = kf ? eng->d_dq : nullptr;
            const double* dr = kf ? eng->d_dr : nullptr;
            const unsigned gp = persistent_grid(eng, 1);
            if (kf)
            {
                k_gemm_int8w_group<true, false><<<g_tiles, 512, lds, eng->stream>>>(
                    eng->d_ptiles, eng->d_pcells, eng->d_pabase, eng->d_pacnt, eng->d_lstart,
                    eng->d_llen, eng->d_pcoff, eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl,
                    (int)eng->Kp, b_max, eng->d_irq, eng->d_irr, eng->d_S, dq, dr, ntiles, 0u);
                k_gemm_int8w_group<true, true><<<gp, 512, lds, eng->stream>>>(
                    eng->d_ptiles, eng->d_pcells, eng->d_pabase, eng->d_pacnt, eng->d_lstart,
                    eng->d_llen, eng->d_pcoff, eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl,
                    (int)eng->Kp, b_max, eng->d_irq, eng->d_irr, eng->d_S, dq, dr, ntiles,
                    g_tiles / 8);
            }
            else
            {
                k_gemm_int8w_group<false, false><<<g_tiles, 512, lds, eng->stream>>>(
                    eng->d_ptiles, eng->d_pcells, eng->d_pabase, eng->d_pacnt, eng->d_lstart,
                    eng->d_llen, eng->d_pcoff, eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl,
                    (int)eng->Kp, b_max, eng->d_irq, eng->d_irr, eng->d_S, dq, dr, ntiles, 0u);
                k_gemm_int8w_group<false, true><<<gp, 512, lds, eng->stream>>>(
                    eng->d_ptiles, eng->d_pcells, eng->d_pabase, eng->d_pacnt, eng->d_lstart,
                    eng->d_llen, eng->d_pcoff, eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl,
                    (int)eng->Kp, b_max, eng->d_irq, eng->d_irr, eng->d_S, dq, dr, ntiles,
                    g_tiles / 8);
            }
        }
        else if (x3)
            k_gemm_int8x_group<<<g_tiles, 512, X3_LDS, eng->stream>>>(
                eng->d_ptiles, eng->d_pcells, eng->d_pabase, eng->d_pacnt, eng->d_lstart,
                eng->d_llen, eng->d_pcoff, eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl,
                (int)eng->Kp, b_max, eng->d_irq, eng->d_irr, eng->d_S, ntiles);
        else
            k_gemm_int8_group<<<std::max(g_tiles, persistent_grid(eng, 4)), 256, 0, eng->stream>>>(
                eng->d_ptiles, eng->d_pcells, eng->d_pabase, eng->d_pacnt, eng->d_lstart,
                eng->d_llen, eng->d_pcoff, eng->d_qh, eng->d_ql, eng->d_rh, eng->d_rl,
                (int)eng->Kp, b_max, eng->d_irq, eng->d_irr, eng->d_S, ntiles);
        LAUNCHCHK();
    }
    {
        PhaseScope ps(eng, 2, 2);
        HIPCHK(hipMemsetAsync(eng->d_nan_count, 0, sizeof(unsigned int), eng->stream));
        // NaN quantiles stay NaN: the certification re-scores those cells
        k_select_n<TableSegments><<<g_items, 256, 0, eng->stream>>>(
            TableSegments{ eng->d_pseg }, nitems, eng->d_S, eng->d_pct, quantile, eng->d_Pref,
            eng->d_nan_list, eng->d_nan_count);
        LAUNCHCHK();
        k_prop_scatter_p<<<est_grid(eng, eng->est_items[round], 256, 4), 256, 0, eng->stream>>>(
            eng->d_pitems, nitems, eng->d_Pref, P, L);
        LAUNCHCHK();
    }
    return CSG_OK;
}

// Proposal-restricted exact quantiles of one chunk (ranked, limbs split).
// The network ranks KP = min(2K, 8, L) labels per cell.  Stage 1 scores the
// confident cells on their first K labels and the rest on every label;
// cells whose K-th label is at the mask's edge (k_prop_check) are scored on
// their labels K .. KP-1 in stage 2, and those whose KP-th label is still at
// the edge on every label in stage 3.  Unscored labels keep -inf, which the
// certification treats as outside every mask.
int propose_chunk(csg_engine* eng, uint32_t q0, uint32_t nq, float quantile, float thre)
{
    const uint32_t L  = eng->L;
    const int      K  = eng->prop_k;
    const int      KP = std::min<int>(std::min(2 * K, 8), (int)L);
    eng->prop_kp      = KP;
    const double   wp = certify_width(eng->G);
    uint32_t*      cA = eng->d_pcnt;
    uint32_t*      cB = eng->d_pcnt + (L + 1);
    uint32_t*      cC = eng->d_pcnt + 2 * (L + 1);
    int            rc;
    {
        PhaseScope ps(eng, 9, 5);
        if ((rc = mlp_propose(eng, nq, KP, eng->d_ptop, eng->d_pmargin)))
            return rc;
        HIPCHK(hipMemsetAsync(eng->d_pcnt, 0, 3 * (L + 1) * sizeof(uint32_t), eng->stream));
        k_prop_count<<<nblk(nq), 256, 0, eng->stream>>>(eng->d_ptop, eng->d_pmargin, nq, KP, K, L,
                                                       eng->prop_margin, cA, eng->d_pslot,
                                                       eng->d_pfb);
        LAUNCHCHK();
    }
    float* P = eng->d_P + (size_t)q0 * L;
    k_fill_f32<<<nblk((size_t)nq * L), 256, 0, eng->stream>>>(P, (size_t)nq * L, -__builtin_inff());
    LAUNCHCHK();
    // stage 1: the confident cells on their first K labels, the cells below
    // the margin (cA[L] of them, listed in d_pfb) on every label
    PropRound r1;
    r1.cw    = cA;
    r1.n_win = nq;
    r1.t1    = K;
    r1.all   = eng->d_pfb;
    r1.n_all = cA + L;
    r1.base  = cA;
    if ((rc = prop_round(eng, r1, 0, nq, quantile, P)))
        return rc;
    // stage 2: widen the edge cells (cB[L], listed in d_plist) to their next
    // ranked labels
    const bool widen = KP > K;
    k_prop_check<<<nblk(nq), 256, 0, eng->stream>>>(P, nq, nullptr, L, eng->d_ptop,
                                                   eng->d_pmargin, KP, K - 1, eng->prop_margin,
                                                   thre, wp, nullptr, widen ? 1 : 0, cB,
                                                   eng->d_pslot, eng->d_plist, cB + L);
    LAUNCHCHK();
    const uint32_t* full   = eng->d_plist;
    const uint32_t* n_full = cB + L;
    if (widen)
    {
        PropRound r2;
        r2.cw        = cB;
        r2.win       = eng->d_plist;
        r2.n_win     = nq;
        r2.n_win_dev = cB + L;
        r2.t0        = K;
        r2.t1        = KP;
        if ((rc = prop_round(eng, r2, 1, nq, quantile, P)))
            return rc;
        k_prop_check<<<nblk(nq), 256, 0, eng->stream>>>(P, nq, cB + L, L, eng->d_ptop,
                                                       eng->d_pmargin, KP, KP - 1,
                                                       eng->prop_margin, thre, wp, eng->d_plist, 0,
                                                       nullptr, nullptr, eng->d_plist + eng->Qc,
                                                       cC + L);
        LAUNCHCHK();
        full   = eng->d_plist + eng->Qc;
        n_full = cC + L;
    }
    // stage 3: the cells whose last ranked label is still at the edge, on
    // every label
    PropRound r3;
    r3.all   = full;
    r3.n_all = n_full;
    if ((rc = prop_round(eng, r3, 2, nq, quantile, P)))
        return rc;
    k_prop_tally<<<1, 64, 0, eng->stream>>>(nq, cA + L, cB + L, n_full, widen ? 1 : 0,
                                            eng->d_dcnt);
    LAUNCHCHK();
    return CSG_OK;
}

int alloc_proposal(csg_engine* eng)
{
    const size_t qc = eng->Qc, L = eng->L;
    int          rc;
    if ((rc = dalloc(eng, &eng->d_ptop, qc * 8)) || (rc = dalloc(eng, &eng->d_pmargin, qc)) ||
        (rc = dalloc(eng, &eng->d_pcnt, 3 * (L + 1))) || (rc = dalloc(eng, &eng->d_pslot, qc * 8)) ||
        (rc = dalloc(eng, &eng->d_pfb, qc)) || (rc = dalloc(eng, &eng->d_plist, 2 * qc)) || (rc = dalloc(eng, &eng->d_pcells, qc * L)) ||
        (rc = dalloc(eng, &eng->d_pitems, qc * L)) || (rc = dalloc(eng, &eng->d_pseg, qc * L)) ||
        (rc = ensure(eng, &eng->d_pmeta, &eng->pmeta_cap,
                     (size_t)16 * L + 48 + prop_tile_bound(eng) * sizeof(GroupTile))))
        return rc;
    eng->d_pcoff  = (uint64_t*)eng->d_pmeta;
    eng->d_pabase = (uint32_t*)(eng->d_pmeta + (size_t)8 * L);
    eng->d_pacnt  = eng->d_pabase + L;
    eng->d_pplan  = (uint32_t*)(eng->d_dcnt + 36);
    eng->d_ptiles = (GroupTile*)(eng->d_pmeta + (size_t)16 * L + 48);
    return CSG_OK;
}

void free_proposal(csg_engine* eng)
{
    dfree(eng->d_ptop);
    dfree(eng->d_pmargin);
    dfree(eng->d_pcnt);
    dfree(eng->d_pslot);
    dfree(eng->d_pfb);
    dfree(eng->d_plist);
    dfree(eng->d_pcells);
    dfree(eng->d_pitems);
    dfree(eng->d_pseg);
    dfree(eng->d_pmeta);
    eng->pmeta_cap = 0;
    eng->d_pabase = eng->d_pacnt = eng->d_pplan = nullptr;
    eng->d_pcoff  = nullptr;
    eng->d_ptiles = nullptr;
}

// The score step's read-back, once the stream has reached it: the counters the
// kernels kept on the device, the work counts that size the next step's
// grids, the phase events, and the undefined-sort flag of its NaN segments.
int fold_score_step(csg_engine* eng)
{
    if (!eng->dcnt_pending)
        return CSG_OK;
    HIPCHK(hipStreamSynchronize(eng->stream));
    eng->dcnt_pending = false;
    collect_phases(eng);
    for (int i = 0; i < CSG_N_COUNTERS; ++i)
        eng->counters[i] += (int64_t)eng->h_dcnt[i];
    const uint32_t* plan = (const uint32_t*)(eng->h_dcnt + 36);
    if (eng->dcnt_prop)
        for (int r = 0; r < 3; ++r)
        {
            eng->est_tiles[r] = plan[4 * r];
            eng->est_items[r] = plan[4 * r + 1];
        }
    if (eng->dcnt_exact)
    {
        eng->est_pairs    = plan[12];
        eng->est_rc_tiles = plan[13];
    }
    if ((int)(eng->h_dcnt[32] & 0xFFFFFFFFull))
    {
        eng->mask_valid = false;
        return eng->fail(CSG_ERR_UNDEFINED,
                         "a NaN-carrying score segment makes the reference's std::sort read out "
                         "of bounds; its percentile is undefined");
    }
    return CSG_OK;
}

}  // namespace

int csg_get_label(csg_engine* eng, float quantile, float finetune_thre, int path,
                  int32_t* first_labels, float* labels_mask, float* scores)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (!eng->score_loaded)
        return eng->fail(CSG_ERR_STATE, "csg_get_label before csg_copyin_score");
    if (!eng->q_ready)
        return eng->fail(CSG_ERR_STATE, "no query data: csg_copyin_score reserved room only; "
                                        "load a batch with csg_copyin_query_raw");
    if (path < CSG_PATH_PLAIN || path > CSG_PATH_FP16_MFMA)
        return eng->fail(CSG_ERR_ARGUMENT, "unknown correlation path");
    const bool prop = eng->prop_k > 0;
    if (prop && path != CSG_PATH_EXACT_INT8)
        return eng->fail(CSG_ERR_ARGUMENT, "the learned proposal runs on the exact INT8 path");
    HIPCHK(hipSetDevice(eng->device));
    int rc;
    if ((rc = fold_score_step(eng)))  // a previous step's deferred read-back
        return rc;
    if (prop && (rc = mlp_check(eng)))
        return rc;
    if (prop && ((uint32_t)eng->prop_k > eng->L || eng->L > 256))
        return eng->fail(CSG_ERR_ARGUMENT, "proposal size exceeds the label count, or more than "
                                           "256 labels");
    if ((rc = set_pct(eng, &eng->d_pct, &eng->pct_quantile, eng->ctidx, eng->L, quantile)))
        return rc;
    if (prop && !eng->d_pcells && (rc = alloc_proposal(eng)))
        return rc;
    std::memset(eng->counters, 0, sizeof(eng->counters));
    HIPCHK(hipMemsetAsync(eng->d_dcnt, 0, 48 * sizeof(unsigned long long), eng->stream));
    eng->ft_exact    = path == CSG_PATH_EXACT_INT8;
    eng->counters[0] = eng->Nq;
    eng->counters[1] = eng->G;
    eng->counters[2] = eng->Nr;
    const size_t nr_pad = round_up(eng->Nr, 256);
    const size_t qc     = eng->Qc;
    if (path == CSG_PATH_EXACT_INT8 && !eng->d_rh)
    {
        if ((rc = dalloc(eng, &eng->d_rh, nr_pad * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_rl, nr_pad * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_qh, qc * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_ql, qc * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_dr, nr_pad)) || (rc = dalloc(eng, &eng->d_dq, qc)) ||
            (rc = dalloc(eng, &eng->d_cstatus, qc)) ||
            (rc = dalloc(eng, &eng->d_exact, qc * eng->L)) ||
            (rc = dalloc(eng, &eng->d_items, qc * eng->L)) ||
            (rc = dalloc(eng, &eng->d_nitems, 2)) || (rc = dalloc(eng, &eng->d_Pref, qc * eng->L)))
            return rc;
    }
    if (path == CSG_PATH_FP16_MFMA && !eng->d_rf16)
    {
        if ((rc = dalloc(eng, &eng->d_rf16, nr_pad * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_qf16, qc * eng->Kp)))
            return rc;
    }
    if (!eng->ref_ranked)
    {
        PhaseScope ps(eng, 0, 1);
        if ((rc = rank_rows(eng, eng->d_rdata, eng->d_rind, eng->d_rptr, 0, eng->Nr, eng->d_Xr,
                            eng->d_sqr_f, eng->d_irr)))
            return rc;
        eng->ref_ranked        = true;
        eng->ref_prepared_path = -1;
    }
    if ((path == CSG_PATH_EXACT_INT8 || path == CSG_PATH_FP16_MFMA) &&
        (eng->ref_prepared_path != path ||
         (path == CSG_PATH_EXACT_INT8 && eng->ref_kf != use_kf(eng))))
    {
        PhaseScope ps(eng, 0, 1);
        const bool kf = path == CSG_PATH_EXACT_INT8 && use_kf(eng);
        if ((rc = prepare_operands(eng, path, eng->d_Xr, nr_pad, eng->d_rh, eng->d_rl, eng->d_rf16,
                                   kf, eng->d_dr)))
            return rc;
        eng->ref_prepared_path = path;
        eng->ref_kf            = kf;
    }
    const uint32_t chunk = prop ? eng->Qc : eng->Qc_dense;
    for (uint32_t q0 = 0; q0 < eng->Nq; q0 += chunk)
    {
        const uint32_t nq = std::min(chunk, eng->Nq - q0);
        {
            PhaseScope ps(eng, 0, path == CSG_PATH_PLAIN || path == CSG_PATH_FP32_MFMA ? 1 : 2);
            // the exact path needs the FP32 sums only for re-scored cells
            // (certify_chunk), the proposal the network features
            if (prop && (rc = mlp_work(eng, nq, false)))
                return rc;
            if ((rc = rank_rows(eng, eng->d_qdata, eng->d_qind, eng->d_qptr, q0, nq, eng->d_Xq,
                                path == CSG_PATH_EXACT_INT8 ? nullptr : eng->d_sqq_f, eng->d_irq,
                                prop ? eng->d_mx : nullptr)))
                return rc;
            if ((rc = prepare_operands(eng, path, eng->d_Xq, round_up(nq, 128), eng->d_qh,
                                       eng->d_ql, eng->d_qf16,
                                       path == CSG_PATH_EXACT_INT8 && eng->ref_kf, eng->d_dq)))
                return rc;
        }
        if (prop)
        {
            if ((rc = propose_chunk(eng, q0, nq, quantile, finetune_thre)) ||
                (rc = certify_chunk(eng, q0, nq, quantile, finetune_thre)))
                return rc;
            continue;
        }
        {
            PhaseScope ps(eng, 1, 1);
            if ((rc = score_chunk(eng, path, nq)))
                return rc;
        }
        HIPCHK(hipMemsetAsync(eng->d_nan_count, 0, sizeof(unsigned int), eng->stream));
        ScoreSegments src{ eng->d_lstart, eng->d_llen, eng->L, (uint64_t)eng->Nr };
        const uint64_t n_items = (uint64_t)nq * eng->L;
        {
            PhaseScope ps(eng, 2, 1);
            k_select<ScoreSegments><<<(unsigned)((n_items + 3) / 4), 256, 0, eng->stream>>>(
                src, n_items, eng->d_S, eng->d_pct, quantile, eng->d_P + (size_t)q0 * eng->L,
                eng->d_nan_list, eng->d_nan_count);
            LAUNCHCHK();
        }
        if (path == CSG_PATH_EXACT_INT8)
        {
            // NaN quantiles stay NaN here; the certification re-scores those cells
            if ((rc = certify_chunk(eng, q0, nq, quantile, finetune_thre)))
                return rc;
            continue;
        }
        if ((rc = resolve_nan_segments(eng, src, eng->d_S, eng->d_pct, quantile,
                                       eng->d_P + (size_t)q0 * eng->L)))
            return rc;
    }
    if (path != CSG_PATH_EXACT_INT8)
    {
        PhaseScope ps(eng, 3, 1);
        k_label<<<(eng->Nq + 255) / 256, 256, 0, eng->stream>>>(eng->d_P, eng->Nq, eng->L,
                                                                finetune_thre, eng->d_first,
                                                                eng->d_mask);
        LAUNCHCHK();
    }
    const size_t nql = (size_t)eng->Nq * eng->L;
    if (first_labels)
        HIPCHK(hipMemcpyAsync(first_labels, eng->d_first, eng->Nq * sizeof(int32_t),
                              hipMemcpyDeviceToHost, eng->stream));
    if (labels_mask)
        HIPCHK(hipMemcpyAsync(labels_mask, eng->d_mask, nql * sizeof(float), hipMemcpyDeviceToHost,
                              eng->stream));
    if (scores)
        HIPCHK(hipMemcpyAsync(scores, eng->d_P, nql * sizeof(float), hipMemcpyDeviceToHost,
                              eng->stream));
    // the counters the kernels kept on the device, and the work counts that
    // size the next step's grids, in one copy with the closing read-back
    if (path == CSG_PATH_EXACT_INT8)
    {
        k_gather_u32<<<1, 64, 0, eng->stream>>>(eng->d_nitems, eng->d_rc_abase + 4 * eng->L,
                                                (uint32_t*)(eng->d_dcnt + 42));
        LAUNCHCHK();
    }
    HIPCHK(hipMemcpyAsync(eng->h_dcnt, eng->d_dcnt, 48 * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, eng->stream));
    eng->dcnt_pending = true;
    eng->dcnt_prop    = prop;
    eng->dcnt_exact   = path == CSG_PATH_EXACT_INT8;
    eng->mask_valid   = true;
    // with outputs the call returns them synchronised; without, the stream
    // runs on (a following csg_finetune queues behind it) and the read-back
    // is folded by the next call that synchronises
    if (first_labels || labels_mask || scores)
        return fold_score_step(eng);
    return CSG_OK;
}

// ================================================================ fine-tune

namespace {

extern "C++" {
template <class CT>
__global__ void k_densify(const uint16_t* __restrict__ data, const int* __restrict__ ind,
                          const int* __restrict__ ptr, uint32_t nrows, uint32_t width,
                          CT* __restrict__ out)
{
    const uint32_t r = blockIdx.x;
    if (r >= nrows)
        return;
    CT* o = out + (size_t)r * width;
    for (uint32_t g = threadIdx.x; g < width; g += blockDim.x)
        o[g] = 0;
    __syncthreads();
    for (int i = ptr[r] + (int)threadIdx.x; i < ptr[r + 1]; i += blockDim.x)
        o[ind[i]] = (CT)data[i];
}
}

}  // namespace

int csg_copyin(csg_engine* eng, uint32_t ref_height, uint32_t width, const uint16_t* ref_data,
               const int32_t* ref_indices, const int32_t* ref_indptr, uint32_t qry_height,
               const uint16_t* qry_data, const int32_t* qry_indices, const int32_t* qry_indptr,
               uint32_t ct_num, const uint32_t* ctidx, const uint32_t* ctdidx,
               const uint32_t* ctdiff, uint64_t n_ctdiff, uint64_t max_uniq_gene,
               const float* labels_mask)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    HIPCHK(hipSetDevice(eng->device));
    csg_destroy(eng);
    if (width > 32767)
        return eng->fail(CSG_ERR_ARGUMENT, "fine-tune gene set size must be at most 32767");
    if (ref_height == 0 || qry_height == 0)
        return eng->fail(CSG_ERR_ARGUMENT, "empty reference or query matrix");
    if (max_uniq_gene < 1 || max_uniq_gene > 16384)
        return eng->fail(CSG_ERR_ARGUMENT, "max_uniq_gene must be in [1, 16384]");
    int rc = check_ctidx(eng, ref_height, ct_num, ctidx);
    if (rc)
        return rc;
    const uint32_t w = std::max<uint32_t>(width, 1);
    // no query data: reserve room for query batches of up to qry_height rows
    // (csg_copyin_query_raw supplies them)
    const bool reserve = !qry_data && !qry_indices && !qry_indptr;
    if (reserve && labels_mask)
        return eng->fail(CSG_ERR_ARGUMENT, "a candidate mask needs the query it belongs to");
    if ((rc = validate_csr(eng, "reference", ref_height, w, ref_data, ref_indices, ref_indptr,
                           max_uniq_gene)) ||
        (!reserve && (rc = validate_csr(eng, "query", qry_height, w, qry_data, qry_indices,
                                        qry_indptr, max_uniq_gene))))
        return rc;
    for (uint64_t i = 0; i < (uint64_t)ct_num * ct_num; ++i)
        if ((uint64_t)ctdidx[2 * i] + ctdidx[2 * i + 1] > n_ctdiff)
            return eng->fail(CSG_ERR_ARGUMENT, "ctdidx window outside ctdiff");
    for (uint64_t i = 0; i < n_ctdiff; ++i)
        if (ctdiff[i] >= w)
            return eng->fail(CSG_ERR_ARGUMENT, "ctdiff gene index outside the fine-tune gene set");
    eng->fNr        = ref_height;
    eng->fNq        = qry_height;
    eng->fq_cap     = qry_height;
    eng->fq_ready   = !reserve;
    eng->Gf         = width;
    eng->fL         = ct_num;
    eng->fnbins     = (int)max_uniq_gene;
    eng->fnbins_ref = (int)max_uniq_gene;
    eng->fctidx.assign(ctidx, ctidx + 2 * ct_num);
    eng->ctdidx.assign(ctdidx, ctdidx + 2 * (size_t)ct_num * ct_num);
    eng->ctdiff.assign(ctdiff, ctdiff + n_ctdiff);
    if (labels_mask)
        eng->host_mask.assign(labels_mask, labels_mask + (size_t)qry_height * ct_num);
    {
        std::vector<uint32_t> ls(ct_num), ll(ct_num);
        for (uint32_t l = 0; l < ct_num; ++l)
        {
            ls[l] = ctidx[2 * l];
            ll[l] = ctidx[2 * l + 1];
        }
        if ((rc = upload(eng, &eng->d_flstart, ls.data(), ls.size())) ||
            (rc = upload(eng, &eng->d_fllen, ll.data(), ll.size())))
            return rc;
    }
    const size_t rnnz = (size_t)ref_indptr[ref_height];
    const size_t qnnz = reserve ? 0 : (size_t)qry_indptr[qry_height];
    eng->fq_nnz_cap   = qnnz;
    uint16_t* d_rd = nullptr;
    int *     d_ri = nullptr, *d_rp = nullptr;
    if ((rc = upload(eng, &d_rd, ref_data, rnnz)) || (rc = upload(eng, &d_ri, ref_indices, rnnz)) ||
        (rc = upload(eng, &d_rp, ref_indptr, (size_t)ref_height + 1)) ||
        (rc = dalloc(eng, &eng->d_fref_codes, (size_t)ref_height * w)) ||
        (!reserve && ((rc = upload(eng, &eng->d_fqdata, qry_data, qnnz)) ||
                      (rc = upload(eng, &eng->d_fqind, qry_indices, qnnz)) ||
                      (rc = upload(eng, &eng->d_fqptr, qry_indptr, (size_t)qry_height + 1)))))
    {
        dfree(d_rd);
        dfree(d_ri);
        dfree(d_rp);
        return rc;
    }
    k_densify<uint16_t><<<ref_height, 256, 0, eng->stream>>>(d_rd, d_ri, d_rp, ref_height, w,
                                                             eng->d_fref_codes);
    dfree(eng->d_fref_codes8);
    if (eng->fnbins <= 256)
    {
        if ((rc = dalloc(eng, &eng->d_fref_codes8, (size_t)ref_height * w)))
        {
            dfree(d_rd);
            dfree(d_ri);
            dfree(d_rp);
            return rc;
        }
        k_densify<uint8_t><<<ref_height, 256, 0, eng->stream>>>(d_rd, d_ri, d_rp, ref_height, w,
                                                                eng->d_fref_codes8);
    }
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(eng->stream);
    dfree(d_rd);
    dfree(d_ri);
    dfree(d_rp);
    if (le != hipSuccess || se != hipSuccess)
        return eng->fail(CSG_ERR_DEVICE, std::string("densify: ") +
                                             hipGetErrorString(le != hipSuccess ? le : se));
    eng->ft_loaded = true;
    return CSG_OK;
}

namespace {

struct FtBatch
{
    std::vector<FtGroup>    groups;
    std::vector<uint32_t>   genes;
    std::vector<uint32_t>   labels;
    std::vector<FtRankItem> cell_items;
    std::vector<uint32_t>   label_pos;  // per group label: offset within the group's rows
    uint64_t                n_ref_items = 0;
    std::vector<FtCorrTile> tiles, thin_tiles;
    std::vector<FtCell>     cells;
    uint64_t                n_segs = 0;
    std::vector<uint8_t>    in_lists;
    std::vector<uint32_t>   in_counts;
    std::vector<uint32_t>   cell_ids;  // query row per cell item
    uint64_t ref_elems = 0, cell_elems = 0, n_scores = 0;
    void clear()
    {
        groups.clear();
        genes.clear();
        labels.clear();
        cell_items.clear();
        label_pos.clear();
        n_ref_items = 0;
        tiles.clear();
        thin_tiles.clear();
        cells.clear();
        n_segs = 0;
        in_lists.clear();
        in_counts.clear();
        cell_ids.clear();
        ref_elems = cell_elems = n_scores = 0;
    }
};



// Grow-only pinned host allocation.
int ensure_host(csg_engine* eng, uint8_t** p, size_t* cap, size_t n)
{
    if (n <= *cap && *p)
        return CSG_OK;
    if (*p)
        (void)hipHostFree(*p);
    *p              = nullptr;
    const size_t want = std::max(n, *cap + *cap / 2);
    HIPCHK(hipHostMalloc((void**)p, want, hipHostMallocDefault));
    *cap = want;
    return CSG_OK;
}

size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// Cells of an exact fine-tune batch whose decisions could not be certified:
// their scores are recomputed in the reference's FP32 chunk-tree order (thin
// tiles limited to the one cell), their quantiles re-selected (libstdc++ sort
// emulation for NaN segments) and the reference's filter applied.  The
// group tables of the batch are still on the device.
int redo_ft_cells(csg_engine* eng, const FtBatch& b, const std::vector<uint32_t>& redo,
                  const FtGroup* d_groups, const uint32_t* d_labels, const uint32_t* d_lpos,
                  float quantile, float thre, std::vector<uint8_t>& lists,
                  std::vector<uint32_t>& counts)
{
    const uint32_t          L = eng->fL;
    const uint32_t          n = (uint32_t)redo.size();
    std::vector<FtCell>     cells(n);
    std::vector<FtCorrTile> tiles;
    std::vector<uint8_t>    in_lists((size_t)n * L);
    std::vector<uint32_t>   in_counts(n);
    uint64_t                nseg = 0;
    for (uint32_t t = 0; t < n; ++t)
    {
        const FtCell&  fc = b.cells[redo[t]];
        const FtGroup& g  = b.groups[fc.group];
        cells[t]          = FtCell{ fc.cell, fc.group, fc.pos, 0, nseg };
        nseg += g.nlabel;
        for (uint32_t r0 = 0; r0 < g.nref; r0 += 256)
            tiles.push_back(FtCorrTile{ fc.group, fc.pos, r0, 1 });
        std::memcpy(&in_lists[(size_t)t * L], &b.in_lists[(size_t)redo[t] * L], L);
        in_counts[t] = b.in_counts[redo[t]];
    }
    size_t off   = 0;
    auto   place = [&](size_t bytes) {
        const size_t o = off;
        off            = align16(off + bytes);
        return o;
    };
    const size_t o_cells = place(n * sizeof(FtCell));
    const size_t o_tiles = place(tiles.size() * sizeof(FtCorrTile));
    const size_t o_inl   = place(in_lists.size());
    const size_t o_inc   = place(n * 4);
    const size_t o_out   = place((size_t)n * (L + 4));
    const size_t o_P     = place(nseg * 4);
    int          rc;
    if ((rc = ensure(eng, &eng->d_ft_redo, &eng->d_ft_redo_cap, off)) ||
        (rc = ensure_host(eng, &eng->h_ft_redo, &eng->h_ft_redo_cap, off)))
        return rc;
    uint8_t* h = eng->h_ft_redo;
    std::memcpy(h + o_cells, cells.data(), n * sizeof(FtCell));
    std::memcpy(h + o_tiles, tiles.data(), tiles.size() * sizeof(FtCorrTile));
    std::memcpy(h + o_inl, in_lists.data(), in_lists.size());
    std::memcpy(h + o_inc, in_counts.data(), n * 4);
    HIPCHK(hipMemcpyAsync(eng->d_ft_redo, h, o_out, hipMemcpyHostToDevice, eng->stream));
    uint8_t*      d       = eng->d_ft_redo;
    const FtCell* d_cells = (const FtCell*)(d + o_cells);
    uint32_t      max_st  = 8;
    for (auto& g : b.groups)
        max_st = std::max(max_st, g.stride);
    {
        PhaseScope ps(eng, 6, 1);
        const int  ktt = std::min<int>((int)max_st, FT_THIN_KT);
        k_ft_corr_thin<FT_THIN><<<(unsigned)tiles.size(), 256, (size_t)FT_THIN * ktt * 2,
                                  eng->stream>>>((const FtCorrTile*)(d + o_tiles), d_groups,
                                                 eng->d_ft_cr, eng->d_ft_csq, eng->d_ft_rr,
                                                 eng->d_ft_rsq, eng->d_ft_scores, ktt);
        LAUNCHCHK();
    }
    FtSegments src{ d_cells, n, d_groups, d_labels, d_lpos, eng->d_fllen };
    float*     d_P = (float*)(d + o_P);
    {
        PhaseScope ps(eng, 8, 1);
        HIPCHK(hipMemsetAsync(eng->d_nan_count, 0, sizeof(unsigned int), eng->stream));
        k_select<FtSegments><<<(unsigned)((nseg + 3) / 4), 256, 0, eng->stream>>>(
            src, nseg, eng->d_ft_scores, eng->d_fpct, quantile, d_P, eng->d_ft_nan,
            eng->d_nan_count);
        LAUNCHCHK();
    }
    std::swap(eng->d_nan_list, eng->d_ft_nan);
    rc = resolve_nan_segments(eng, src, eng->d_ft_scores, eng->d_fpct, quantile, d_P);
    std::swap(eng->d_nan_list, eng->d_ft_nan);
    if (rc)
        return rc;
    uint8_t*  d_ol = d + o_out;
    uint32_t* d_oc = (uint32_t*)(d + o_out + (size_t)n * L);
    {
        PhaseScope ps(eng, 8, 1);
        k_ft_filter<<<(n + 127) / 128, 128, 0, eng->stream>>>(
            d_cells, n, d_groups, d_labels, d_P, (const uint8_t*)(d + o_inl),
            (const uint32_t*)(d + o_inc), L, thre, d_ol, d_oc);
        LAUNCHCHK();
    }
    HIPCHK(hipMemcpyAsync(h + o_out, d + o_out, (size_t)n * (L + 4), hipMemcpyDeviceToHost,
                          eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    const uint8_t*  ol = h + o_out;
    const uint32_t* oc = (const uint32_t*)(h + o_out + (size_t)n * L);
    for (uint32_t t = 0; t < n; ++t)
    {
        const uint32_t c = b.cell_ids[redo[t]];
        counts[c]        = oc[t];
        std::memcpy(&lists[(size_t)c * L], &ol[(size_t)t * L], L);
    }
    return CSG_OK;
}

// Uploads a batch (one packed copy), runs rank / correlate / select / filter,
// and writes the new candidate lists of the batch's cells into `lists` /
// `counts`.
int run_ft_batch(csg_engine* eng, FtBatch& b, float quantile, float thre,
                 std::vector<uint8_t>& lists, std::vector<uint32_t>& counts)
{
    const uint32_t L     = eng->fL;
    const uint32_t ncell = (uint32_t)b.cells.size();
    uint32_t       max_gs = 0;
    for (auto& g : b.groups)
        max_gs = std::max(max_gs, g.gs);
    // the exact form needs |X| <= 1919 for the folded limb products
    const bool exact = eng->ft_exact && max_gs <= 1919;
    b.tiles.clear();
    b.thin_tiles.clear();
    for (uint32_t gi = 0; gi < (uint32_t)b.groups.size(); ++gi)
    {
        const FtGroup& g = b.groups[gi];
        if (exact)
        {
            for (uint32_t c0 = 0; c0 < g.ncell; c0 += 16)
                for (uint32_t r0 = 0; r0 < g.nref; r0 += 256)
                    b.tiles.push_back(FtCorrTile{ gi, c0, r0, 0 });
        }
        else if (g.ncell <= (uint32_t)FT_THIN)
        {
            for (uint32_t r0 = 0; r0 < g.nref; r0 += 256)
                b.thin_tiles.push_back(FtCorrTile{ gi, 0, r0, 0 });
        }
        else
        {
            for (uint32_t c0 = 0; c0 < g.ncell; c0 += FT_TC)
                for (uint32_t r0 = 0; r0 < g.nref; r0 += FT_TR)
                    b.tiles.push_back(FtCorrTile{ gi, c0, r0, 0 });
        }
    }
    // table layout in the packed upload
    size_t off = 0;
    auto   place = [&](size_t bytes) {
        const size_t o = off;
        off            = align16(off + bytes);
        return o;
    };
    const size_t o_groups = place(b.groups.size() * sizeof(FtGroup));
    const size_t o_genes  = place(b.genes.size() * 4);
    const size_t o_labels = place(b.labels.size() * 4);
    const size_t o_lpos   = place(b.label_pos.size() * 4);
    const size_t o_citems = place(b.cell_items.size() * sizeof(FtRankItem));
    const size_t o_tiles  = place(b.tiles.size() * sizeof(FtCorrTile));
    const size_t o_thin   = place(b.thin_tiles.size() * sizeof(FtCorrTile));
    const size_t o_cells  = place(b.cells.size() * sizeof(FtCell));
    const size_t o_inl    = place(b.in_lists.size());
    const size_t o_inc    = place(b.in_counts.size() * 4);
    const size_t total    = off;
    int          rc;
    if ((rc = ensure_host(eng, &eng->h_ft_stage, &eng->h_ft_stage_cap, total)) ||
        (rc = ensure(eng, &eng->d_ft_tables, &eng->d_ft_tables_cap, total)))
        return rc;
    uint8_t* h = eng->h_ft_stage;
    auto     put = [&](size_t o, const void* src, size_t bytes) {
        if (bytes)
            std::memcpy(h + o, src, bytes);
    };
    put(o_groups, b.groups.data(), b.groups.size() * sizeof(FtGroup));
    put(o_genes, b.genes.data(), b.genes.size() * 4);
    put(o_labels, b.labels.data(), b.labels.size() * 4);
    put(o_lpos, b.label_pos.data(), b.label_pos.size() * 4);
    put(o_citems, b.cell_items.data(), b.cell_items.size() * sizeof(FtRankItem));
    put(o_tiles, b.tiles.data(), b.tiles.size() * sizeof(FtCorrTile));
    put(o_thin, b.thin_tiles.data(), b.thin_tiles.size() * sizeof(FtCorrTile));
    put(o_cells, b.cells.data(), b.cells.size() * sizeof(FtCell));
    put(o_inl, b.in_lists.data(), b.in_lists.size());
    put(o_inc, b.in_counts.data(), b.in_counts.size() * 4);
    HIPCHK(hipMemcpyAsync(eng->d_ft_tables, h, total, hipMemcpyHostToDevice, eng->stream));
    uint8_t*          d          = eng->d_ft_tables;
    const FtGroup*    d_groups   = (const FtGroup*)(d + o_groups);
    const uint32_t*   d_genes    = (const uint32_t*)(d + o_genes);
    const uint32_t*   d_labels   = (const uint32_t*)(d + o_labels);
    const uint32_t*   d_lpos     = (const uint32_t*)(d + o_lpos);
    const FtRankItem* d_citems   = (const FtRankItem*)(d + o_citems);
    const FtCorrTile* d_tiles    = (const FtCorrTile*)(d + o_tiles);
    const FtCorrTile* d_thin     = (const FtCorrTile*)(d + o_thin);
    const FtCell*     d_cells    = (const FtCell*)(d + o_cells);
    const uint8_t*    d_in_lists = (const uint8_t*)(d + o_inl);
    const uint32_t*   d_in_cnt   = (const uint32_t*)(d + o_inc);
    // work buffers
    const size_t nseg = b.n_segs;
    if ((rc = ensure(eng, &eng->d_ft_rr, &eng->d_ft_rr_cap, b.ref_elems)) ||
        (rc = ensure(eng, &eng->d_ft_cr, &eng->d_ft_cr_cap, b.cell_elems)) ||
        (rc = ensure(eng, &eng->d_ft_rsq, &eng->d_ft_rsq_cap, b.n_ref_items)) ||
        (rc = ensure(eng, &eng->d_ft_csq, &eng->d_ft_csq_cap, b.cell_items.size())) ||
        (rc = ensure(eng, &eng->d_ft_scores, &eng->d_ft_scores_cap, b.n_scores)) ||
        (rc = ensure(eng, &eng->d_ft_P, &eng->d_ft_P_cap, nseg)) ||
        (rc = ensure(eng, &eng->d_ft_nan, &eng->d_ft_nan_cap, nseg)) ||
        (rc = ensure(eng, &eng->d_ft_out, &eng->d_ft_out_cap, (size_t)ncell * (L + 5))) ||
        (rc = ensure_host(eng, &eng->h_ft_out, &eng->h_ft_out_cap, (size_t)ncell * (L + 5))))
        return rc;
    uint8_t*  d_out_lists  = eng->d_ft_out;
    uint32_t* d_out_counts = (uint32_t*)(eng->d_ft_out + (size_t)ncell * L);
    uint8_t*  d_status     = eng->d_ft_out + (size_t)ncell * (L + 4);

    const size_t stride_max = round_up(max_gs, 8);
    const size_t lds_rank_q = (size_t)((eng->fnbins + 3) & ~3) * 4 +
                              round_up(std::max<uint32_t>(eng->Gf, 1), 8) * 2 + stride_max * 2;
    const size_t lds_rank_r = (size_t)((eng->fnbins + 3) & ~3) * 4 + stride_max * 4;
    // the correlation tiles walk K in LDS stages, so any gene subset fits
    const int    kt_corr    = ft_corr_stage((int)stride_max, eng->lds_limit, FT_TC + FT_TR);
    const int    kt_thin    = std::min<int>((int)stride_max, FT_THIN_KT);
    const size_t lds_corr   = (size_t)(FT_TC + FT_TR) * (kt_corr / 2 + 1) * 4;
    const size_t lds_thin   = (size_t)FT_THIN * kt_thin * 2;
    if (lds_rank_q > eng->lds_limit || lds_rank_r > eng->lds_limit ||
        (b.tiles.empty() && b.thin_tiles.empty()) || b.n_ref_items == 0 || b.cell_items.empty())
        return eng->fail(CSG_ERR_ARGUMENT, "fine-tune gene subset of " + std::to_string(max_gs) +
                                               " genes with " + std::to_string(eng->fnbins) +
                                               " value codes exceeds the rank kernels' LDS");
    {
        PhaseScope   ps(eng, 7, 2);
        k_ft_rank_ref<<<(unsigned)b.n_ref_items, 128, lds_rank_r, eng->stream>>>(
            d_groups, (uint32_t)b.groups.size(), d_labels, eng->d_flstart, eng->d_fllen, d_genes,
            eng->d_fref_codes, std::max<uint32_t>(eng->Gf, 1), eng->fnbins, eng->d_ft_rr,
            eng->d_ft_rsq);
        LAUNCHCHK();
        k_ft_rank_qry<<<(unsigned)b.cell_items.size(), 128, lds_rank_q, eng->stream>>>(
            d_citems, d_groups, d_genes, eng->d_fqdata, eng->d_fqind, eng->d_fqptr,
            std::max<uint32_t>(eng->Gf, 1), eng->fnbins, eng->d_ft_cr, eng->d_ft_csq);
        LAUNCHCHK();
    }
    if (exact)
    {
        PhaseScope ps(eng, 6, 1);
        k_ft_mfma<<<(unsigned)b.tiles.size(), 256, 0, eng->stream>>>(
            d_tiles, d_groups, eng->d_ft_cr, eng->d_ft_csq, eng->d_ft_rr, eng->d_ft_rsq,
            eng->d_ft_scores);
        LAUNCHCHK();
    }
    else
    {
        PhaseScope ps(eng, 6, (int)!b.tiles.empty() + (int)!b.thin_tiles.empty());
        if (!b.tiles.empty())
        {
            k_ft_corr<FT_TC, FT_TR, FT_RC, FT_RR><<<(unsigned)b.tiles.size(), FT_NT, lds_corr,
                                                    eng->stream>>>(
                d_tiles, d_groups, eng->d_ft_cr, eng->d_ft_csq, eng->d_ft_rr, eng->d_ft_rsq,
                eng->d_ft_scores, kt_corr);
            LAUNCHCHK();
        }
        if (!b.thin_tiles.empty())
        {
            k_ft_corr_thin<FT_THIN><<<(unsigned)b.thin_tiles.size(), 256, lds_thin, eng->stream>>>(
                d_thin, d_groups, eng->d_ft_cr, eng->d_ft_csq, eng->d_ft_rr, eng->d_ft_rsq,
                eng->d_ft_scores, kt_thin);
            LAUNCHCHK();
        }
    }
    {
        PhaseScope ps(eng, 8, 1);
        HIPCHK(hipMemsetAsync(eng->d_nan_count, 0, sizeof(unsigned int), eng->stream));
        FtSegments src{ d_cells, ncell, d_groups, d_labels, d_lpos, eng->d_fllen };
        k_select<FtSegments><<<(unsigned)((nseg + 3) / 4), 256, 0, eng->stream>>>(
            src, nseg, eng->d_ft_scores, eng->d_fpct, quantile, eng->d_ft_P, eng->d_ft_nan,
            eng->d_nan_count);
        LAUNCHCHK();
    }
    if (exact)
    {
        // certified filter; NaN quantiles are left NaN and send their cells
        // to the FP32 redo below
        {
            PhaseScope ps(eng, 8, 1);
            k_ft_filter_exact<<<(ncell + 127) / 128, 128, 0, eng->stream>>>(
                d_cells, ncell, d_groups, d_labels, eng->d_ft_P, d_in_lists, d_in_cnt, L, thre,
                d_out_lists, d_out_counts, d_status);
            LAUNCHCHK();
        }
        HIPCHK(hipMemcpyAsync(eng->h_ft_out, eng->d_ft_out, (size_t)ncell * (L + 5),
                              hipMemcpyDeviceToHost, eng->stream));
        HIPCHK(hipStreamSynchronize(eng->stream));
        const uint8_t*        ol = eng->h_ft_out;
        const uint32_t*       oc = (const uint32_t*)(eng->h_ft_out + (size_t)ncell * L);
        const uint8_t*        st = eng->h_ft_out + (size_t)ncell * (L + 4);
        std::vector<uint32_t> redo;
        for (uint32_t i = 0; i < ncell; ++i)
        {
            if (st[i])
            {
                redo.push_back(i);
                continue;
            }
            const uint32_t c = b.cell_ids[i];
            counts[c]        = oc[i];
            std::memcpy(&lists[(size_t)c * L], &ol[(size_t)i * L], L);
        }
        if (redo.empty())
            return CSG_OK;
        eng->counters[11] += (int64_t)redo.size();
        return redo_ft_cells(eng, b, redo, d_groups, d_labels, d_lpos, quantile, thre, lists,
                             counts);
    }
    std::swap(eng->d_nan_list, eng->d_ft_nan);
    rc = resolve_nan_segments(eng, FtSegments{ d_cells, ncell, d_groups, d_labels, d_lpos, eng->d_fllen },
                              eng->d_ft_scores, eng->d_fpct, quantile, eng->d_ft_P);
    std::swap(eng->d_nan_list, eng->d_ft_nan);
    if (rc)
        return rc;
    {
        PhaseScope ps(eng, 8, 1);
        k_ft_filter<<<(ncell + 127) / 128, 128, 0, eng->stream>>>(
            d_cells, ncell, d_groups, d_labels, eng->d_ft_P, d_in_lists, d_in_cnt, L, thre,
            d_out_lists, d_out_counts);
        LAUNCHCHK();
    }
    HIPCHK(hipMemcpyAsync(eng->h_ft_out, eng->d_ft_out, (size_t)ncell * (L + 4),
                          hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    const uint8_t*  ol = eng->h_ft_out;
    const uint32_t* oc = (const uint32_t*)(eng->h_ft_out + (size_t)ncell * L);
    for (uint32_t i = 0; i < ncell; ++i)
    {
        const uint32_t c = b.cell_ids[i];
        counts[c]        = oc[i];
        std::memcpy(&lists[(size_t)c * L], &ol[(size_t)i * L], L);
    }
    return CSG_OK;
}

// Bitset over the fine-tune genes of the first min(len, cut) genes of the gene
// window of ordered label pair (i, j); built on first use and kept until the
// fine-tune inputs change.
const uint64_t* window_bits(csg_engine* eng, uint32_t i, uint32_t j, int cut)
{
    const uint64_t key = ((uint64_t)i << 40) | ((uint64_t)j << 24) | (uint64_t)(uint32_t)cut;
    auto           it  = eng->window_cache.find(key);
    if (it != eng->window_cache.end())
        return it->second.data();
    const size_t          words = (std::max<uint32_t>(eng->Gf, 1) + 63) / 64;
    std::vector<uint64_t> bits(words, 0ull);
    const uint32_t        L   = eng->fL;
    const uint32_t        pos = eng->ctdidx[((size_t)i * L + j) * 2];
    int                   len = (int)eng->ctdidx[((size_t)i * L + j) * 2 + 1];
    if (len > cut)
        len = cut;
    for (int t = 0; t < len; ++t)
    {
        const uint32_t g = eng->ctdiff[pos + t];
        bits[g / 64] |= 1ull << (g % 64);
    }
    return eng->window_cache.emplace(key, std::move(bits)).first->second.data();
}

// Mask rows packed into 64-bit words (bit l of word l/64), so the host reads
// Nq * ceil(L/64) words instead of Nq * L floats.
__global__ void k_pack_mask(const float* __restrict__ mask, uint32_t nq, uint32_t L,
                            uint64_t* __restrict__ out)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq)
        return;
    const uint32_t W = (L + 63) / 64;
    for (uint32_t w = 0; w < W; ++w)
    {
        uint64_t bits = 0;
        for (uint32_t j = w * 64; j < min(L, w * 64 + 64); ++j)
            if (mask[(size_t)q * L + j] != 0.f)
                bits |= 1ull << (j - w * 64);
        out[(size_t)q * W + w] = bits;
    }
}

struct SetKey
{
    uint64_t w[4];
    bool     operator==(const SetKey& o) const
    {
        return w[0] == o.w[0] && w[1] == o.w[1] && w[2] == o.w[2] && w[3] == o.w[3];
    }
};
struct SetKeyHash
{
    size_t operator()(const SetKey& k) const
    {
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (int i = 0; i < 4; ++i)
        {
            h ^= k.w[i] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
            h *= 0xBF58476D1CE4E5B9ull;
        }
        return (size_t)(h ^ (h >> 31));
    }
};

}  // namespace

}  // extern "C"

namespace {

// ------------------------------------------------- device-resident rounds

constexpr int FTD_TPB = 256;  // threads of the thread-level work-queue kernels

void ftd_free_chains(csg_engine* eng)
{
    for (auto& c : eng->ftd_gC)
        if (c.exec)
            (void)hipGraphExecDestroy(c.exec);
    eng->ftd_gC.clear();
}

int ftd_free(csg_engine* eng)
{
    if (eng->ftd_gAs)
        (void)hipGraphExecDestroy(eng->ftd_gAs);
    eng->ftd_gAs = nullptr;
    eng->ftd_sigAs.clear();
    for (int i = 0; i < 2; ++i)
    {
        if (eng->ftd_gAc[i])
            (void)hipGraphExecDestroy(eng->ftd_gAc[i]);
        eng->ftd_gAc[i] = nullptr;
        eng->ftd_sigAc[i].clear();
    }
    if (eng->ftd_gA)
        (void)hipGraphExecDestroy(eng->ftd_gA);
    for (auto& b : eng->ftd_gB)
        if (b.exec)
            (void)hipGraphExecDestroy(b.exec);
    eng->ftd_gB.clear();
    ftd_free_chains(eng);
    eng->ftd_gA = nullptr;
    eng->ftd_sigA.clear();
    dfree(eng->d_fl_lists);
    dfree(eng->d_fl_counts);
    dfree(eng->d_fl_rounds);
    dfree(eng->d_fl_act);
    dfree(eng->d_fl_next);
    dfree(eng->d_fl_cslot);
    dfree(eng->d_fl_cgid);
    dfree(eng->d_fl_cpos);
    dfree(eng->d_fl_htab);
    dfree(eng->d_fl_howner);
    dfree(eng->d_fl_hgid);
    dfree(eng->d_fl_gowner);
    dfree(eng->d_fl_gcount);
    dfree(eng->d_fl_groups);
    dfree(eng->d_fl_gflag);
    dfree(eng->d_fl_glabels);
    dfree(eng->d_fl_glpos);
    dfree(eng->d_fl_ggenes);
    dfree(eng->d_fl_gtile0);
    dfree(eng->d_fl_gthin0);
    dfree(eng->d_fl_gseg0);
    dfree(eng->d_fl_cnt);
    dfree(eng->d_fl_cut);
    dfree(eng->d_fl_ctdidx);
    dfree(eng->d_fl_ctdiff);
    dfree(eng->d_fl_final);
    dfree(eng->d_fl_mask);
    dfree(eng->d_fl_items);
    dfree(eng->d_fl_cells);
    dfree(eng->d_fl_tiles);
    dfree(eng->d_fl_thin);
    dfree(eng->d_fl_rthin);
    dfree(eng->d_fl_stage);
    dfree(eng->d_fl_redo);
    dfree(eng->d_fl_rseg);
    dfree(eng->d_fl_pred);
    dfree(eng->d_fl_nan1);
    dfree(eng->d_fl_nan2);
    dfree(eng->d_fl_segcell);
    dfree(eng->d_fl_refgroup);
    dfree(eng->d_fl_refrow);
    eng->fl_cap = {};
    dfree(eng->d_tt_state);
    dfree(eng->d_tt_slots);
    dfree(eng->d_tt_chunks);
    dfree(eng->d_tt_genes);
    dfree(eng->d_tt_crank);
    dfree(eng->d_tt_scores);
    dfree(eng->d_tt_P);
    dfree(eng->d_tt_nanl);
    eng->tt_cap_state = eng->tt_cap_slots = eng->tt_cap_chunks = eng->tt_cap_genes = 0;
    eng->tt_cap_crank = eng->tt_cap_scores = eng->tt_cap_P = eng->tt_cap_nanl = 0;
    if (eng->h_tt_state)
        (void)hipHostFree(eng->h_tt_state);
    eng->h_tt_state = nullptr;
    eng->ftd_gcap = 0;
    eng->ftd_T    = 0;
    if (eng->h_fl_cnt)
        (void)hipHostFree(eng->h_fl_cnt);
    eng->h_fl_cnt = nullptr;
    eng->ftd_static = false;
    eng->ftd_tuned  = false;
    return CSG_OK;
}

// Per-cell state, the pair windows and the k -> k_t table (sized by the
// fine-tune inputs; kept until they change).
int ftd_static(csg_engine* eng)
{
    if (eng->ftd_static)
        return CSG_OK;
    // per-cell state for the largest batch the inputs have room for
    const uint32_t Nq = std::max(eng->fNq, eng->fq_cap), L = eng->fL;
    int            rc;
    std::vector<uint32_t> cut(L + 1, 0u);
    for (uint32_t k = 1; k <= L; ++k)
        cut[k] = (uint32_t)std::round(500 * std::pow((2 / 3.0), std::log2((double)k)));
    if ((rc = dalloc(eng, &eng->d_fl_lists, (size_t)Nq * L)) ||
        (rc = dalloc(eng, &eng->d_fl_counts, Nq)) || (rc = dalloc(eng, &eng->d_fl_rounds, Nq)) ||
        (rc = dalloc(eng, &eng->d_fl_act, Nq)) || (rc = dalloc(eng, &eng->d_fl_next, Nq)) ||
        (rc = dalloc(eng, &eng->d_fl_cslot, Nq)) || (rc = dalloc(eng, &eng->d_fl_cgid, Nq)) ||
        (rc = dalloc(eng, &eng->d_fl_cpos, Nq)) || (rc = dalloc(eng, &eng->d_fl_cnt, 1)) ||
        (rc = dalloc(eng, &eng->d_fl_final, Nq)) ||
        (rc = upload(eng, &eng->d_fl_cut, cut.data(), cut.size())) ||
        (rc = upload(eng, &eng->d_fl_ctdidx, eng->ctdidx.data(), eng->ctdidx.size())) ||
        (rc = upload(eng, &eng->d_fl_ctdiff, eng->ctdiff.data(), std::max<size_t>(eng->ctdiff.size(), 1))))
        return rc;
    if (!eng->h_fl_cnt)
        HIPCHK(hipHostMalloc((void**)&eng->h_fl_cnt, sizeof(FtdCounts), hipHostMallocDefault));
    eng->ftd_gstride = round_up(std::max<uint32_t>(eng->Gf, 1), 4);
    eng->ftd_static  = true;
    return CSG_OK;
}

// Group tables and the hash table for up to `n` active cells.
int ftd_groups(csg_engine* eng, uint32_t n)
{
    if (n <= eng->ftd_gcap && eng->d_fl_groups)
        return CSG_OK;
    // the captured rounds name the tables about to move
    if (eng->ftd_gAs)
        (void)hipGraphExecDestroy(eng->ftd_gAs);
    eng->ftd_gAs = nullptr;
    eng->ftd_sigAs.clear();
    for (int i = 0; i < 2; ++i)
    {
        if (eng->ftd_gAc[i])
            (void)hipGraphExecDestroy(eng->ftd_gAc[i]);
        eng->ftd_gAc[i] = nullptr;
        eng->ftd_sigAc[i].clear();
    }
    if (eng->ftd_gA)
        (void)hipGraphExecDestroy(eng->ftd_gA);
    eng->ftd_gA = nullptr;
    eng->ftd_sigA.clear();
    for (auto& b : eng->ftd_gB)
        if (b.exec)
            (void)hipGraphExecDestroy(b.exec);
    eng->ftd_gB.clear();
    ftd_free_chains(eng);
    const uint32_t cap = std::max(n, eng->ftd_gcap + eng->ftd_gcap / 2);
    uint32_t       T   = 1024;
    while (T < 2 * cap)
        T <<= 1;
    const size_t L = eng->fL, gst = eng->ftd_gstride;
    int          rc;
    eng->ftd_gcap = 0;
    if ((rc = dalloc(eng, &eng->d_fl_htab, T)) || (rc = dalloc(eng, &eng->d_fl_howner, T)) ||
        (rc = dalloc(eng, &eng->d_fl_hgid, T)) || (rc = dalloc(eng, &eng->d_fl_gowner, cap)) ||
        (rc = dalloc(eng, &eng->d_fl_gcount, cap)) || (rc = dalloc(eng, &eng->d_fl_groups, cap)) ||
        (rc = dalloc(eng, &eng->d_fl_gflag, cap)) ||
        (rc = dalloc(eng, &eng->d_fl_glabels, (size_t)cap * L)) ||
        (rc = dalloc(eng, &eng->d_fl_glpos, (size_t)cap * L)) ||
        (rc = dalloc(eng, &eng->d_fl_ggenes, (size_t)cap * gst)) ||
        (rc = dalloc(eng, &eng->d_fl_gtile0, cap)) || (rc = dalloc(eng, &eng->d_fl_gthin0, cap)) ||
        (rc = dalloc(eng, &eng->d_fl_gseg0, cap)))
        return rc;
    eng->ftd_gcap = cap;
    eng->ftd_T    = T;
    return CSG_OK;
}

// Work tables for one round's counts (grow-only).
int ftd_work(csg_engine* eng, const FtdCounts& h)
{
    auto& c  = eng->fl_cap;
    const size_t L = eng->fL;
    int   rc;
    if ((rc = ensure(eng, &eng->d_ft_rr, &eng->d_ft_rr_cap, std::max<uint64_t>(h.ref_elems, 8))) ||
        (rc = ensure(eng, &eng->d_ft_rsq, &eng->d_ft_rsq_cap, std::max<uint64_t>(h.n_ref_items, 1))) ||
        (rc = ensure(eng, &eng->d_ft_cr, &eng->d_ft_cr_cap, std::max<uint64_t>(h.cell_elems, 8))) ||
        (rc = ensure(eng, &eng->d_ft_csq, &eng->d_ft_csq_cap, std::max<uint64_t>(h.n_cells, 1))) ||
        (rc = ensure(eng, &eng->d_ft_scores, &eng->d_ft_scores_cap, std::max<uint64_t>(h.n_scores, 1))) ||
        (rc = ensure(eng, &eng->d_ft_P, &eng->d_ft_P_cap, std::max<uint64_t>(h.n_segs, 1))) ||
        (rc = ensure(eng, &eng->d_fl_items, &c.items, std::max<uint64_t>(h.n_cells, 1))) ||
        (rc = ensure(eng, &eng->d_fl_cells, &c.cells, std::max<uint64_t>(h.n_cells, 1))) ||
        (rc = ensure(eng, &eng->d_fl_stage, &c.stage, std::max<uint64_t>((uint64_t)h.n_cells * L, 1))) ||
        (rc = ensure(eng, &eng->d_fl_redo, &c.redo, std::max<uint64_t>(h.n_cells, 1))) ||
        (rc = ensure(eng, &eng->d_fl_tiles, &c.tiles, std::max<uint64_t>(h.n_tiles, 1))) ||
        (rc = ensure(eng, &eng->d_fl_thin, &c.thin, std::max<uint64_t>(h.n_thin, 1))) ||
        (rc = ensure(eng, &eng->d_fl_rthin, &c.rthin, std::max<uint64_t>(h.rthin_bound, 1))) ||
        (rc = ensure(eng, &eng->d_fl_rseg, &c.rseg, std::max<uint64_t>(h.n_segs, 1))) ||
        (rc = ensure(eng, &eng->d_fl_pred, &c.pred, std::max<uint64_t>(h.n_segs, 1))) ||
        (rc = ensure(eng, &eng->d_fl_nan1, &c.nan1, std::max<uint64_t>(h.n_segs, 1))) ||
        (rc = ensure(eng, &eng->d_fl_nan2, &c.nan2, std::max<uint64_t>(h.n_segs, 1))) ||
        (rc = ensure(eng, &eng->d_fl_segcell, &c.segcell, std::max<uint64_t>(h.n_segs, 1))) ||
        (rc = ensure(eng, &eng->d_fl_refgroup, &c.refgroup, std::max<uint64_t>(h.n_ref_items, 1))) ||
        (rc = ensure(eng, &eng->d_fl_refrow, &c.refrow, std::max<uint64_t>(h.n_ref_items, 1))))
        return rc;
    return CSG_OK;
}

// Dynamic LDS of the work-queue kernels for gene subsets of up to Gf genes.
struct FtdLds
{
    size_t setup, rank_ref, rank_qry, corr, thin;
    int    kt_corr, kt_thin;  // LDS stage widths (genes) of the FP32 correlation tiles
    int    wave_ints;  // per-wave LDS of the wave-per-row reference rank kernel
    size_t rank_ref_w;
};

FtdLds ftd_lds(const csg_engine* eng)
{
    const size_t st = round_up(std::max<uint32_t>(eng->Gf, 1), 8);
    const size_t nb = (size_t)((eng->fnbins + 3) & ~3) * 4;
    FtdLds       d;
    // bit words and a flag byte per gene (the flags padded to whole words)
    d.setup    = ((size_t)std::max<uint32_t>(eng->Gf, 1) + 31) / 32 * (4 + 32) +
              (size_t)FTD_PAIRS * 8;
    d.rank_ref = nb + st * 4;
    d.rank_qry = nb + st * 2 + st * 2;
    // the correlation tiles walk K in LDS stages, so any gene subset fits
    d.kt_corr  = ft_corr_stage((int)st, eng->lds_limit, FT_TC + FT_TR);
    d.corr     = (size_t)(FT_TC + FT_TR) * (d.kt_corr / 2 + 1) * 4;
    d.kt_thin  = std::min<int>((int)st, FT_THIN_KT);
    d.thin     = (size_t)FT_THIN * d.kt_thin * 2;
    d.wave_ints  = (int)(((eng->fnbins + 3) & ~3) + st / 2);
    d.rank_ref_w = (size_t)4 * d.wave_ints * 4;
    return d;
}

// Largest group (cells) of the fused form: the engine's setting when the
// per-wave LDS of the wave rank kernel fits, else 0 (no fused groups).
uint32_t ftd_fuse(const csg_engine* eng)
{
    return ftd_lds(eng).rank_ref_w <= 65536 ? (uint32_t)eng->ftd_fuse_max : 0u;
}

// Launch mask of a round: the work-queue kernels whose work lists can be
// non-empty, from the round's totals (a superset for every group batch of
// the round).  Kernels outside the mask are not launched; each mask has its
// own part B graph.
uint32_t ftd_mask_of(const csg_engine* eng, const FtdCounts& h)
{
    return ftd_mask_need(h.ref_elems, h.n_tiles, h.n_etiles, h.n_btiles, h.n_thin,
                         ftd_fuse(eng) ? h.n_fused : 0u);
}

// Capacities of the current work tables, in the units of the round totals
// they hold (a chained round whose totals exceed them halts).
FtdCaps ftd_caps(const csg_engine* eng)
{
    const auto&    c = eng->fl_cap;
    const uint64_t L = std::max<uint32_t>(eng->fL, 1);
    FtdCaps        k;
    k.ref_items  = std::min({ (uint64_t)eng->d_ft_rsq_cap, (uint64_t)c.refgroup, (uint64_t)c.refrow });
    k.cells      = std::min({ (uint64_t)eng->d_ft_csq_cap, (uint64_t)c.items, (uint64_t)c.cells,
                              (uint64_t)c.redo, (uint64_t)c.stage / L });
    k.ref_elems  = eng->d_ft_rr_cap;
    k.cell_elems = eng->d_ft_cr_cap;
    k.scores     = eng->d_ft_scores_cap;
    if (eng->ftd_chain_scores)
        k.scores = std::min<uint64_t>(k.scores, eng->ftd_chain_scores);
    k.segs       = std::min({ (uint64_t)eng->d_ft_P_cap, (uint64_t)c.rseg, (uint64_t)c.pred,
                              (uint64_t)c.nan1, (uint64_t)c.nan2, (uint64_t)c.segcell });
    k.tiles      = c.tiles;
    k.thin       = c.thin;
    k.rthin      = c.rthin;
    return k;
}

// Round part A: grouping, group set-up and the prefix sums.  `chain`: the
// round is one of a chain launched without a host read-back, with the work
// tables of these capacities and part B's families of `chain_mask` (a round
// needing more halts the chain, see k_ftd_scan).
int ftd_launch_a(csg_engine* eng, bool small, int compact_times, const FtdCaps* chain = nullptr,
                 uint32_t chain_mask = 0)
{
    const uint32_t L = eng->fL, T = eng->ftd_T;
    const unsigned g = (unsigned)eng->ftd_grid_threads;
    const FtdLds   d = ftd_lds(eng);
    FtdCounts*     cnt = eng->d_fl_cnt;
    if (compact_times >= 0)
        // the previous round's compaction into this round's active list
        k_ftd_compact<<<g, FTD_TPB, 0, eng->stream>>>(cnt, eng->d_fl_act, eng->d_fl_counts,
                                                      eng->d_fl_rounds, compact_times,
                                                      eng->d_fl_next);
    if (small)
        // at most FTD_SMALL_NT active cells (this round's can only be fewer
        // than the last round's): the grouping in one block
        k_ftd_group_small<<<1, FTD_SMALL_NT, 0, eng->stream>>>(
            cnt, T, eng->d_fl_act, eng->d_fl_next, (unsigned long long*)eng->d_fl_htab,
            eng->d_fl_gcount, eng->d_fl_lists, eng->d_fl_counts, L, eng->d_fl_howner,
            eng->d_fl_cslot, eng->d_fl_hgid, eng->d_fl_gowner, eng->d_fl_cgid, eng->d_fl_cpos);
    else
    {
        k_ftd_counters<<<1, 64, 0, eng->stream>>>(cnt, T);
        k_ftd_begin<<<g, FTD_TPB, 0, eng->stream>>>(cnt, eng->d_fl_act, eng->d_fl_next,
                                                    (unsigned long long*)eng->d_fl_htab,
                                                    eng->d_fl_gcount);
        k_ftd_hash<<<g, FTD_TPB, 0, eng->stream>>>(cnt, eng->d_fl_act, eng->d_fl_lists,
                                                   eng->d_fl_counts, L,
                                                   (unsigned long long*)eng->d_fl_htab,
                                                   eng->d_fl_howner, eng->d_fl_cslot);
        k_ftd_gid<<<g, FTD_TPB, 0, eng->stream>>>(cnt,
                                                  (const unsigned long long*)eng->d_fl_htab,
                                                  eng->d_fl_howner, eng->d_fl_hgid,
                                                  eng->d_fl_gowner);
        k_ftd_member<<<g, FTD_TPB, 0, eng->stream>>>(
            cnt, eng->d_fl_act, eng->d_fl_lists, eng->d_fl_counts, L, eng->d_fl_cslot,
            eng->d_fl_hgid, eng->d_fl_gowner, eng->d_fl_gcount, eng->d_fl_cgid, eng->d_fl_cpos);
    }
    k_ftd_setup<<<(unsigned)eng->ftd_grid_setup, FTD_SETUP_NT, d.setup, eng->stream>>>(
        cnt, eng->d_fl_gowner, eng->d_fl_gcount, eng->d_fl_lists, eng->d_fl_counts, L,
        eng->d_fl_ctdidx, eng->d_fl_ctdiff, eng->d_fl_cut, eng->d_fllen,
        std::max<uint32_t>(eng->Gf, 1), (uint32_t)eng->ftd_gstride, eng->ft_exact ? 1 : 0,
        ftd_fuse(eng), eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_glabels, eng->d_fl_glpos,
        eng->d_fl_ggenes);
    k_ftd_scan<<<1, FTD_SCAN_NT, 0, eng->stream>>>(
        cnt, eng->d_fl_groups, eng->d_fl_gflag, (uint32_t)eng->ftd_thin_max,
        (uint32_t)eng->ftd_t64, eng->d_fl_gtile0, eng->d_fl_gthin0, eng->d_fl_gseg0, 1,
        chain ? 1 : 0, chain_mask, chain ? *chain : FtdCaps{});
    LAUNCHCHK();
    return CSG_OK;
}

// Reference-row ranks of the groups that are not fused: with byte codes the
// register-resident kernel (k_ftd_fused_r in its rank mode) for G_sub up to
// 1280 genes; longer gene sets and 16-bit codes one wave per row when the
// per-wave LDS fits, else one block per row.
int ftd_launch_rank_ref(csg_engine* eng, int grid)
{
    const FtdLds   d   = ftd_lds(eng);
    const uint32_t Gf  = std::max<uint32_t>(eng->Gf, 1);
    FtdCounts*     cnt = eng->d_fl_cnt;
    int            lo  = 0;  // rows of chunk length <= lo already ranked
    if (eng->d_fref_codes8)
    {
        // byte codes: G_sub up to 1280 genes in the register-resident
        // kernel, by chunk length classes (0, 4], (4, 8], (8, 10]
#define CSG_FTD_RANK_R(S, LO)                                                               \
    k_ftd_fused_r<S, LO, true><<<(unsigned)grid, 256, 0, eng->stream>>>(                    \
        cnt, eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_refgroup, eng->d_fl_refrow,       \
        eng->d_fl_ggenes, eng->d_fref_codes8, Gf, nullptr, nullptr, nullptr, eng->d_ft_rr, \
        eng->d_ft_rsq)
        CSG_FTD_RANK_R(4, 0);
        lo = 4;
        if (Gf > 512)
        {
            CSG_FTD_RANK_R(8, 4);
            lo = 8;
        }
        if (Gf > 1024)
        {
            CSG_FTD_RANK_R(10, 8);
            lo = 10;
        }
#undef CSG_FTD_RANK_R
        if (Gf <= 1280)
        {
            LAUNCHCHK();
            return CSG_OK;
        }
    }
    if (d.rank_ref_w <= 65536)
    {
        if (eng->d_fref_codes8)
            k_ftd_rank_ref_w<uint8_t><<<(unsigned)grid, 256, d.rank_ref_w, eng->stream>>>(
                cnt, eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_refgroup, eng->d_fl_refrow,
                eng->d_fl_ggenes, eng->d_fref_codes8, Gf, eng->fnbins, d.wave_ints, eng->d_ft_rr,
                eng->d_ft_rsq, lo);
        else
            k_ftd_rank_ref_w<uint16_t><<<(unsigned)grid, 256, d.rank_ref_w, eng->stream>>>(
                cnt, eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_refgroup, eng->d_fl_refrow,
                eng->d_fl_ggenes, eng->d_fref_codes, Gf, eng->fnbins, d.wave_ints, eng->d_ft_rr,
                eng->d_ft_rsq, lo);
    }
    else
        // (rows already ranked above are ranked again to the same values)
        k_ftd_rank_ref<<<(unsigned)grid, 128, d.rank_ref, eng->stream>>>(
            cnt, eng->d_fl_groups, eng->d_fl_glabels, eng->d_flstart, eng->d_fllen,
            eng->d_fl_ggenes, eng->d_fref_codes, Gf, eng->fnbins, eng->d_ft_rr, eng->d_ft_rsq);
    LAUNCHCHK();
    return CSG_OK;
}

// Query-cell ranks: with codes below 256 one wave per cell for G_sub up to
// 1024 genes (when its LDS fits), one block per cell for the rest.
int ftd_launch_rank_qry(csg_engine* eng, int grid)
{
    const FtdLds   d   = ftd_lds(eng);
    const uint32_t Gf  = std::max<uint32_t>(eng->Gf, 1);
    FtdCounts*     cnt = eng->d_fl_cnt;
    const size_t   lw  = (size_t)4 * ftd_qry_wave_ints(8, Gf) * 4;
    int            lo  = 0;
    if (eng->d_fref_codes8 && lw <= 65536)
    {
        k_ftd_rank_qry_w<8><<<(unsigned)grid, 256, lw, eng->stream>>>(
            cnt, eng->d_fl_items, eng->d_fl_groups, eng->d_fl_ggenes, eng->d_fqdata,
            eng->d_fqind, eng->d_fqptr, Gf, eng->d_ft_cr, eng->d_ft_csq);
        lo = 8;
    }
    if (lo == 0 || Gf > 1024)
        k_ftd_rank_qry<<<(unsigned)grid, 128, d.rank_qry, eng->stream>>>(
            cnt, eng->d_fl_items, eng->d_fl_groups, eng->d_fl_ggenes, eng->d_fqdata,
            eng->d_fqind, eng->d_fqptr, Gf, eng->fnbins, eng->d_ft_cr, eng->d_ft_csq, lo);
    LAUNCHCHK();
    return CSG_OK;
}

int ftd_launch_corr(csg_engine* eng, int grid)
{
    const FtdLds   d   = ftd_lds(eng);
    FtdCounts*     cnt = eng->d_fl_cnt;
    const uint32_t m   = eng->ftd_mask;
    if (m & FTDM_MFMA)
        k_ftd_mfma<<<(unsigned)grid, 256, 0, eng->stream>>>(
            cnt, eng->d_fl_tiles, eng->d_fl_groups, eng->d_fl_gflag, eng->d_ft_cr, eng->d_ft_csq,
            eng->d_ft_rr, eng->d_ft_rsq, eng->d_ft_scores);
    if (m & FTDM_MFMA64)
        k_ftd_mfma64<<<(unsigned)grid, 256, 0, eng->stream>>>(
            cnt, eng->d_fl_tiles, eng->d_fl_groups, eng->d_fl_gflag, eng->d_ft_cr, eng->d_ft_csq,
            eng->d_ft_rr, eng->d_ft_rsq, eng->d_ft_scores);
    if (m & FTDM_CORR)
        k_ftd_corr<<<(unsigned)grid, FT_NT, d.corr, eng->stream>>>(
            cnt, eng->d_fl_tiles, eng->d_fl_groups, eng->d_fl_gflag, eng->d_ft_cr, eng->d_ft_csq,
            eng->d_ft_rr, eng->d_ft_rsq, eng->d_ft_scores, d.kt_corr);
    if (m & FTDM_THIN)
        k_ftd_thin<<<(unsigned)grid, 256, d.thin, eng->stream>>>(
            cnt, 0, eng->d_fl_thin, eng->d_fl_groups, eng->d_ft_cr, eng->d_ft_csq, eng->d_ft_rr,
            eng->d_ft_rsq, eng->d_ft_scores, d.kt_thin);
    LAUNCHCHK();
    return CSG_OK;
}

// Fused-form groups (k_ftd_fused, one wave per reference row).
int ftd_launch_fused(csg_engine* eng, int grid)
{
    if (!(eng->ftd_mask & FTDM_FUSED) || !ftd_fuse(eng))
        return CSG_OK;
    const FtdLds   d   = ftd_lds(eng);
    FtdCounts*     cnt = eng->d_fl_cnt;
    const uint32_t Gf  = std::max<uint32_t>(eng->Gf, 1);
    if (eng->d_fref_codes8)
    {
        // byte codes: chunk lengths (0, 4], (4, 8] and (8, 10] with the
        // cells in registers, longer ones (G_sub above 1280 genes, beyond
        // the FP32 tiles' LDS plan today) through the LDS-table kernel
#define CSG_FTD_FUSED_R(S, LO)                                                          \
    k_ftd_fused_r<S, LO><<<(unsigned)grid, 256, 0, eng->stream>>>(                      \
        cnt, eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_refgroup, eng->d_fl_refrow,   \
        eng->d_fl_ggenes, eng->d_fref_codes8, Gf, eng->d_ft_cr, eng->d_ft_csq,          \
        eng->d_ft_scores)
        CSG_FTD_FUSED_R(4, 0);
        if (Gf > 512)
            CSG_FTD_FUSED_R(8, 4);
        if (Gf > 1024)
            CSG_FTD_FUSED_R(10, 8);
#undef CSG_FTD_FUSED_R
        if (Gf > 1280)
            k_ftd_fused<uint8_t><<<(unsigned)grid, 256, d.rank_ref_w, eng->stream>>>(
                cnt, eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_refgroup, eng->d_fl_refrow,
                eng->d_fl_ggenes, eng->d_fref_codes8, Gf, eng->fnbins, d.wave_ints, eng->d_ft_cr,
                eng->d_ft_csq, eng->d_ft_scores, 10);
    }
    else
        k_ftd_fused<uint16_t><<<(unsigned)grid, 256, d.rank_ref_w, eng->stream>>>(
            cnt, eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_refgroup, eng->d_fl_refrow,
            eng->d_fl_ggenes, eng->d_fref_codes, Gf, eng->fnbins, d.wave_ints, eng->d_ft_cr,
            eng->d_ft_csq, eng->d_ft_scores, 0);
    LAUNCHCHK();
    return CSG_OK;
}

int ftd_launch_select(csg_engine* eng, int grid, float quantile)
{
    FtdCounts*        cnt = eng->d_fl_cnt;
    const FtdSegments seg{ eng->d_fl_cells, eng->d_fl_segcell, eng->d_fl_groups,
                           eng->d_fl_glabels, eng->d_fl_glpos, eng->d_fllen };
    k_ftd_select<FtdSegments><<<(unsigned)grid, 256, 0, eng->stream>>>(
        cnt, 0, seg, eng->d_ft_scores, eng->d_fpct, quantile, eng->d_ft_P, eng->d_fl_nan1);
    LAUNCHCHK();
    return CSG_OK;
}

// Round part B: ranks, correlations, quantiles, filters, the redo of the
// undecided exact-form cells and the next round's active list.
int ftd_launch_b(csg_engine* eng, float quantile, float thre, int times)
{
    const uint32_t   L   = eng->fL;
    const uint32_t   Gf  = std::max<uint32_t>(eng->Gf, 1);
    const unsigned   g   = (unsigned)eng->ftd_grid_threads;
    const FtdLds     d   = ftd_lds(eng);
    FtdCounts*       cnt = eng->d_fl_cnt;
    const FtdSegments seg{ eng->d_fl_cells, eng->d_fl_segcell, eng->d_fl_groups,
                           eng->d_fl_glabels, eng->d_fl_glpos, eng->d_fllen };
    const FtdListSegments rseg{ seg, eng->d_fl_rseg };
    k_ftd_prep<<<std::max<unsigned>(g, (unsigned)eng->ftd_grid_setup), FTD_TPB, 0, eng->stream>>>(
        cnt, eng->d_fl_act, eng->d_fl_cgid, eng->d_fl_cpos, eng->d_fl_groups, eng->d_fl_gflag,
        eng->d_fl_gseg0, eng->d_fl_counts, eng->d_fl_items, eng->d_fl_cells, eng->d_fl_segcell,
        eng->d_fl_glabels, eng->d_flstart, eng->d_fllen, eng->d_fl_refgroup, eng->d_fl_refrow,
        (uint32_t)eng->ftd_thin_max, (uint32_t)eng->ftd_t64, eng->d_fl_gtile0, eng->d_fl_gthin0,
        eng->d_fl_tiles, eng->d_fl_thin);
    // kernel families timed only when launched eagerly (phase events cannot
    // be recorded into a captured graph)
    auto fam = [&](int phase) {
        return std::unique_ptr<PhaseScope>(eng->ftd_eager ? new PhaseScope(eng, phase, 1) : nullptr);
    };
    {
        auto ps = fam(7);
        if ((eng->ftd_mask & FTDM_RANK) && ftd_launch_rank_ref(eng, eng->ftd_grid_rank))
            return CSG_ERR_DEVICE;
        if (ftd_launch_rank_qry(eng, eng->ftd_grid_rank))
            return CSG_ERR_DEVICE;
    }
    {
        auto ps = fam(6);
        if (ftd_launch_corr(eng, eng->ftd_grid_corr) || ftd_launch_fused(eng, eng->ftd_grid_fused))
            return CSG_ERR_DEVICE;
    }
    auto ps_sel = fam(8);
    if (ftd_launch_select(eng, eng->ftd_grid_select, quantile))
        return CSG_ERR_DEVICE;
    k_ftd_select_nan<FtdSegments><<<g, FTD_TPB, 0, eng->stream>>>(
        cnt, 0, seg, eng->d_fl_nan1, eng->d_ft_scores, eng->d_fpct, quantile, eng->d_ft_P,
        eng->d_fl_cells, eng->d_fl_segcell, eng->d_fl_gflag);
    k_ftd_filter<<<g * (FTD_TPB / FTD_FILTER_NT), FTD_FILTER_NT, 0, eng->stream>>>(
        cnt, eng->d_fl_cells, eng->d_fl_groups,
                                                 eng->d_fl_gflag, eng->d_fl_glabels, eng->d_ft_P,
                                                 eng->d_fl_lists, eng->d_fl_counts, L, thre,
                                                 eng->d_fl_stage, eng->d_fl_redo,
                                                 eng->d_fl_rthin, eng->d_fl_rseg);
    LAUNCHCHK();
    if (!(eng->ftd_mask & FTDM_REDO))
        return CSG_OK;
    // FP32 redo of the undecided exact-form cells
    k_ftd_thin<<<(unsigned)eng->ftd_grid_corr, 256, d.thin, eng->stream>>>(
        cnt, 1, eng->d_fl_rthin, eng->d_fl_groups, eng->d_ft_cr, eng->d_ft_csq, eng->d_ft_rr,
        eng->d_ft_rsq, eng->d_ft_scores, d.kt_thin);
    k_ftd_select<FtdListSegments><<<(unsigned)eng->ftd_grid_select, 256, 0, eng->stream>>>(
        cnt, 1, rseg, eng->d_ft_scores, eng->d_fpct, quantile, eng->d_fl_pred, eng->d_fl_nan2);
    k_ftd_select_nan<FtdListSegments><<<g, FTD_TPB, 0, eng->stream>>>(
        cnt, 1, rseg, eng->d_fl_nan2, eng->d_ft_scores, eng->d_fpct, quantile, eng->d_fl_pred,
        eng->d_fl_cells, eng->d_fl_segcell, eng->d_fl_gflag);
    k_ftd_redo_scatter<<<g, FTD_TPB, 0, eng->stream>>>(cnt, eng->d_fl_rseg, eng->d_fl_pred,
                                                       eng->d_ft_P);
    k_ftd_filter_redo<<<g, FTD_TPB, 0, eng->stream>>>(cnt, eng->d_fl_redo, eng->d_fl_cells,
                                                      eng->d_fl_groups, eng->d_fl_glabels,
                                                      eng->d_ft_P, eng->d_fl_lists,
                                                      eng->d_fl_counts, L, thre,
                                                      eng->d_fl_stage);
    LAUNCHCHK();
    return CSG_OK;
}

// Round end: the round is counted and the next round's active list formed.
// Device bytes of one round's work tables for the given totals.
uint64_t ftd_bytes(const FtdCounts& h, uint32_t L)
{
    return 2 * h.ref_elems + 2 * h.cell_elems + 4 * h.n_scores + 4 * h.n_ref_items +
           (uint64_t)h.n_cells * (sizeof(FtRankItem) + sizeof(FtCell) + L + 12) +
           (uint64_t)h.n_segs * 36 + ((uint64_t)h.n_tiles + h.n_thin + h.rthin_bound) * 16;
}

// Autotuner of the work-queue launch settings.  On the first device round
// after the fine-tune inputs change (with enough work to time), each
// candidate block count of the reference-rank, correlation and select
// kernels is timed on that round's own work -- the kernels are pure
// functions of their inputs, so repeating them changes nothing -- and the
// fastest count is kept for the graphs.  The round's cell, query-rank and
// tile kernels run first (they are idempotent too); the graph then runs the
// whole round.
float ftd_time(csg_engine* eng, const std::function<int()>& launch, int reps)
{
    hipEvent_t a = take_event(eng), b = take_event(eng);
    float      best = 1e30f;
    for (int r = 0; r < reps; ++r)
    {
        (void)hipEventRecord(a, eng->stream);
        if (launch())
        {
            best = 1e30f;
            break;
        }
        (void)hipEventRecord(b, eng->stream);
        float ms = 0.f;
        if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess)
            best = std::min(best, ms);
    }
    eng->event_pool.push_back(a);
    eng->event_pool.push_back(b);
    return best;
}

int ftd_tune(csg_engine* eng, float quantile)
{
    int rc;
    const int      cu  = std::max(1, eng->prop.multiProcessorCount);
    const unsigned g   = (unsigned)eng->ftd_grid_threads;
    const uint32_t Gf  = std::max<uint32_t>(eng->Gf, 1);
    const FtdLds   d   = ftd_lds(eng);
    FtdCounts*     cnt = eng->d_fl_cnt;
    k_ftd_cells<<<g, FTD_TPB, 0, eng->stream>>>(cnt, eng->d_fl_act, eng->d_fl_cgid, eng->d_fl_cpos,
                                                eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_gseg0,
                                                eng->d_fl_counts, eng->d_fl_items,
                                                eng->d_fl_cells, eng->d_fl_segcell);
    k_ftd_refmap<<<(unsigned)eng->ftd_grid_setup, 256, 0, eng->stream>>>(
        cnt, eng->d_fl_groups, eng->d_fl_gflag, eng->d_fl_glabels, eng->d_flstart, eng->d_fllen,
        eng->d_fl_refgroup, eng->d_fl_refrow);
    if (ftd_launch_rank_qry(eng, eng->ftd_grid_rank))
        return CSG_ERR_DEVICE;
    k_ftd_tiles<<<g, FTD_TPB, 0, eng->stream>>>(cnt, eng->d_fl_groups, eng->d_fl_gflag,
                                                (uint32_t)eng->ftd_thin_max, (uint32_t)eng->ftd_t64, eng->d_fl_gtile0,
                                                eng->d_fl_gthin0, eng->d_fl_tiles, eng->d_fl_thin);
    LAUNCHCHK();
    auto pick = [&](const std::vector<int>& mult, const std::function<int(int)>& run, int* grid,
                    float* best_ms) {
        float best = 1e30f;
        int   bg   = *grid;
        for (int m : mult)
        {
            const float t = ftd_time(eng, [&]() { return run(m * cu); }, 2);
            if (t < best)
            {
                best = t;
                bg   = m * cu;
            }
        }
        *grid    = bg;
        *best_ms = best;
    };
    pick({ 2, 4, 8, 16, 32 }, [&](int gr) { return ftd_launch_rank_ref(eng, gr); },
         &eng->ftd_grid_rank, &eng->ftd_tune_ms[0]);
    (void)ftd_launch_rank_ref(eng, eng->ftd_grid_rank);
    // big-tile threshold x correlation grid: the scan and tile kernels are
    // re-run for each threshold (only the tile offsets depend on it)
    auto retile = [&](uint32_t t64) -> int {
        eng->ftd_t64 = (int)t64;
        k_ftd_scan<<<1, FTD_SCAN_NT, 0, eng->stream>>>(cnt, eng->d_fl_groups, eng->d_fl_gflag,
                                                       (uint32_t)eng->ftd_thin_max, t64,
                                                       eng->d_fl_gtile0, eng->d_fl_gthin0,
                                                       eng->d_fl_gseg0, 0, 0, 0u, FtdCaps{});
        LAUNCHCHK();
        FtdCounts hc;
        HIPCHK(hipMemcpyAsync(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, eng->stream));
        HIPCHK(hipStreamSynchronize(eng->stream));
        // the tile kinds (16-cell / big exact, FP32) follow the threshold:
        // the launch mask must, or a candidate would time skipped tiles
        eng->ftd_mask = ftd_mask_of(eng, hc);
        int rc2;
        if ((rc2 = ensure(eng, &eng->d_fl_tiles, &eng->fl_cap.tiles, std::max<uint64_t>(hc.n_tiles, 1))))
            return rc2;
        k_ftd_tiles<<<g, FTD_TPB, 0, eng->stream>>>(cnt, eng->d_fl_groups, eng->d_fl_gflag,
                                                    (uint32_t)eng->ftd_thin_max, t64,
                                                    eng->d_fl_gtile0, eng->d_fl_gthin0,
                                                    eng->d_fl_tiles, eng->d_fl_thin);
        LAUNCHCHK();
        return CSG_OK;
    };
    float    best_corr = 1e30f;
    uint32_t best_t64  = (uint32_t)eng->ftd_t64;
    int      best_grid = eng->ftd_grid_corr;
    for (uint32_t t64 : { 0xFFFFFFFFu, 32u, 64u, 128u })
    {
        int rc2;
        if ((rc2 = retile(t64)))
            return rc2;
        int   grid = eng->ftd_grid_corr;
        float ms   = 0.f;
        pick({ 1, 2, 3, 4, 6 }, [&](int gr) { return ftd_launch_corr(eng, gr); }, &grid, &ms);
        if (ms < best_corr)
        {
            best_corr = ms;
            best_t64  = t64;
            best_grid = grid;
        }
    }
    eng->ftd_grid_corr  = best_grid;
    eng->ftd_tune_ms[1] = best_corr;
    if ((rc = retile(best_t64)))
        return rc;
    if ((eng->ftd_mask & FTDM_FUSED) && ftd_fuse(eng))
        pick({ 4, 8, 16, 32 }, [&](int gr) { return ftd_launch_fused(eng, gr); },
             &eng->ftd_grid_fused, &eng->ftd_tune_ms[3]);
    pick({ 2, 4, 8, 16 },
         [&](int gr) {
             k_ftd_reset_nan<<<1, 64, 0, eng->stream>>>(cnt);
             return ftd_launch_select(eng, gr, quantile);
         },
         &eng->ftd_grid_select, &eng->ftd_tune_ms[2]);
    k_ftd_reset_nan<<<1, 64, 0, eng->stream>>>(cnt);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(eng->stream));
    eng->ftd_tuned = true;
    return CSG_OK;
}

// Captures `launch` into a graph (re-captured when any pointer or launch
// setting in `sig` changed).
template <class F>
int ftd_graph(csg_engine* eng, hipGraphExec_t* exec, std::vector<uintptr_t>* have,
              const std::vector<uintptr_t>& sig, F launch)
{
    if (*exec && *have == sig)
        return CSG_OK;
    if (*exec)
        (void)hipGraphExecDestroy(*exec);
    *exec = nullptr;
    have->clear();
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(eng->stream, hipStreamCaptureModeThreadLocal));
    const int  rc = launch();
    hipError_t e  = hipStreamEndCapture(eng->stream, &graph);
    if (rc)
    {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess)
        return eng->fail(CSG_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
    {
        *exec = nullptr;
        return eng->fail(CSG_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    *have = sig;
    return CSG_OK;
}

std::vector<uintptr_t> ftd_sig_a(const csg_engine* eng)
{
    return { (uintptr_t)eng->d_fl_htab, (uintptr_t)eng->d_fl_groups, (uintptr_t)eng->d_fl_ggenes,
             (uintptr_t)eng->d_fl_glabels, (uintptr_t)eng->d_fl_gseg0, (uintptr_t)eng->ftd_T,
             (uintptr_t)eng->ftd_grid_setup, (uintptr_t)eng->ftd_thin_max,
             (uintptr_t)eng->ftd_t64, (uintptr_t)ftd_fuse(eng),
             (uintptr_t)eng->ft_exact, (uintptr_t)eng->ftd_grid_threads };
}

std::vector<uintptr_t> ftd_sig_b(const csg_engine* eng, float quantile, float thre, int times)
{
    uint32_t q, t;
    std::memcpy(&q, &quantile, 4);
    std::memcpy(&t, &thre, 4);
    return { (uintptr_t)eng->d_ft_rr, (uintptr_t)eng->d_ft_rsq, (uintptr_t)eng->d_ft_cr,
             (uintptr_t)eng->d_ft_csq, (uintptr_t)eng->d_ft_scores, (uintptr_t)eng->d_ft_P,
             (uintptr_t)eng->d_fl_items, (uintptr_t)eng->d_fl_cells, (uintptr_t)eng->d_fl_stage,
             (uintptr_t)eng->d_fl_redo, (uintptr_t)eng->d_fl_tiles, (uintptr_t)eng->d_fl_thin,
             (uintptr_t)eng->d_fl_rthin, (uintptr_t)eng->d_fl_rseg, (uintptr_t)eng->d_fl_pred,
             (uintptr_t)eng->d_fl_nan1, (uintptr_t)eng->d_fl_nan2, (uintptr_t)eng->d_fpct,
             (uintptr_t)eng->d_fl_segcell, (uintptr_t)eng->d_fl_refgroup,
             (uintptr_t)eng->d_fl_refrow, (uintptr_t)eng->d_fref_codes8,
             (uintptr_t)eng->d_fref_codes, (uintptr_t)eng->d_fqdata, (uintptr_t)eng->d_fqind,
             (uintptr_t)eng->d_fqptr, (uintptr_t)eng->fnbins,
             (uintptr_t)eng->d_fl_groups, (uintptr_t)eng->d_fl_ggenes, (uintptr_t)ftd_fuse(eng),
             (uintptr_t)eng->ftd_mask, (uintptr_t)q,
             (uintptr_t)t, (uintptr_t)(int64_t)times, (uintptr_t)eng->ftd_grid_rank,
             (uintptr_t)eng->ftd_grid_corr, (uintptr_t)eng->ftd_grid_select,
             (uintptr_t)eng->ftd_grid_fused,
             (uintptr_t)eng->ftd_thin_max, (uintptr_t)eng->ftd_t64,
             (uintptr_t)eng->ftd_grid_threads };
}

// The part B graph of the current launch mask (captured on first use and
// re-captured when a pointer or setting in its signature changed).
int ftd_graph_b(csg_engine* eng, float quantile, float thre, int times, hipGraphExec_t* out)
{
    csg_engine::FtdGraphB* b = nullptr;
    for (auto& e : eng->ftd_gB)
        if (e.mask == eng->ftd_mask)
            b = &e;
    if (!b)
    {
        eng->ftd_gB.emplace_back();
        b       = &eng->ftd_gB.back();
        b->mask = eng->ftd_mask;
    }
    const int rc = ftd_graph(eng, &b->exec, &b->sig, ftd_sig_b(eng, quantile, thre, times),
                             [&]() { return ftd_launch_b(eng, quantile, thre, times); });
    *out = b->exec;
    return rc;
}

// A chain of `rounds` few-cell rounds in one graph: each round's part A (led
// by the previous round's compaction) and its part B with the launch mask
// `mask`, replayed without a host read-back between the rounds.  A round
// that needs larger work tables or a family outside the mask halts the chain
// (k_ftd_scan).  One graph per chain length and mask, re-captured when a
// table or setting in its signature changes.
int ftd_graph_chain(csg_engine* eng, int rounds, uint32_t mask, float quantile, float thre,
                    int times, hipGraphExec_t* out)
{
    const FtdCaps  caps = ftd_caps(eng);
    const uint32_t keep = eng->ftd_mask;
    eng->ftd_mask       = mask;
    std::vector<uintptr_t>       sig = ftd_sig_a(eng);
    const std::vector<uintptr_t> sb  = ftd_sig_b(eng, quantile, thre, times);
    sig.insert(sig.end(), sb.begin(), sb.end());
    for (uint64_t v : { caps.ref_items, caps.cells, caps.ref_elems, caps.cell_elems, caps.scores,
                        caps.segs, caps.tiles, caps.thin, caps.rthin })
        sig.push_back((uintptr_t)v);
    sig.push_back((uintptr_t)rounds);
    csg_engine::FtdChainGraph* c = nullptr;
    for (auto& e : eng->ftd_gC)
        if (e.rounds == rounds && e.mask == mask)
            c = &e;
    if (!c)
    {
        if (eng->ftd_gC.size() >= 8)
        {
            if (eng->ftd_gC.front().exec)
                (void)hipGraphExecDestroy(eng->ftd_gC.front().exec);
            eng->ftd_gC.erase(eng->ftd_gC.begin());
        }
        eng->ftd_gC.emplace_back();
        c         = &eng->ftd_gC.back();
        c->rounds = rounds;
        c->mask   = mask;
    }
    const int rc = ftd_graph(eng, &c->exec, &c->sig, sig, [&]() -> int {
        for (int r = 0; r < rounds; ++r)
        {
            int e;
            if ((e = ftd_launch_a(eng, true, times, &caps, mask)) ||
                (e = ftd_launch_b(eng, quantile, thre, times)))
                return e;
        }
        return CSG_OK;
    });
    eng->ftd_mask = keep;
    *out          = rc ? nullptr : c->exec;
    return rc;
}

// A round whose work tables exceed the budget: its groups are processed in
// ranges whose tables fit (each range re-scanned on the device), every range
// through the same round graph.
int ftd_batches(csg_engine* eng, const FtdCounts& h, float quantile, float thre, int times)
{
    const uint32_t        ng = h.n_groups, L = eng->fL;
    std::vector<FtGroup>  groups(ng);
    std::vector<uint8_t>  flags(ng);
    HIPCHK(hipMemcpy(groups.data(), eng->d_fl_groups, ng * sizeof(FtGroup), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(flags.data(), eng->d_fl_gflag, ng, hipMemcpyDeviceToHost));
    // greedy ranges under the budget; the totals of the largest range size the tables
    std::vector<std::pair<uint32_t, uint32_t>> ranges;
    FtdCounts                                  cap{}, cur{};
    uint32_t                                   g0 = 0;
    auto add = [](FtdCounts& t, const FtdSizes& z) {
        t.n_ref_items += z.ref_items;
        t.n_cells += (uint32_t)z.cell_items;
        t.ref_elems += z.ref_elems;
        t.cell_elems += z.cell_elems;
        t.n_scores += z.scores;
        t.n_tiles += (uint32_t)z.tiles;
        t.n_thin += (uint32_t)z.thin;
        t.n_segs += z.segs;
        t.rthin_bound += z.rthin;
    };
    auto keep_max = [&](const FtdCounts& t) {
        cap.n_ref_items = std::max(cap.n_ref_items, t.n_ref_items);
        cap.n_cells     = std::max(cap.n_cells, t.n_cells);
        cap.ref_elems   = std::max(cap.ref_elems, t.ref_elems);
        cap.cell_elems  = std::max(cap.cell_elems, t.cell_elems);
        cap.n_scores    = std::max(cap.n_scores, t.n_scores);
        cap.n_tiles     = std::max(cap.n_tiles, t.n_tiles);
        cap.n_thin      = std::max(cap.n_thin, t.n_thin);
        cap.n_segs      = std::max(cap.n_segs, t.n_segs);
        cap.rthin_bound = std::max(cap.rthin_bound, t.rthin_bound);
    };
    for (uint32_t g = 0; g < ng; ++g)
    {
        const FtdSizes z = ftd_sizes(groups[g], flags[g], (uint32_t)eng->ftd_thin_max,
                                     (uint32_t)eng->ftd_t64);
        FtdCounts      t = cur;
        add(t, z);
        if (g > g0 && ftd_bytes(t, L) > eng->ftd_budget)
        {
            ranges.emplace_back(g0, g);
            keep_max(cur);
            cur = FtdCounts{};
            g0  = g;
            add(cur, z);
        }
        else
            cur = t;
    }
    ranges.emplace_back(g0, ng);
    keep_max(cur);
    int rc;
    if ((rc = ftd_work(eng, cap)))
        return rc;
    hipGraphExec_t gB = nullptr;
    if ((rc = ftd_graph_b(eng, quantile, thre, times, &gB)))
        return rc;
    for (auto& r : ranges)
    {
        const uint32_t lohi[2] = { r.first, r.second };
        HIPCHK(hipMemcpyAsync(&eng->d_fl_cnt->g_lo, lohi, sizeof(lohi), hipMemcpyHostToDevice,
                              eng->stream));
        k_ftd_scan<<<1, FTD_SCAN_NT, 0, eng->stream>>>(eng->d_fl_cnt, eng->d_fl_groups,
                                                       eng->d_fl_gflag,
                                                       (uint32_t)eng->ftd_thin_max,
                                                       (uint32_t)eng->ftd_t64,
                                                       eng->d_fl_gtile0, eng->d_fl_gthin0,
                                                       eng->d_fl_gseg0, 0, 0, 0u, FtdCaps{});
        LAUNCHCHK();
        HIPCHK(hipGraphLaunch(gB, eng->stream));
        // the range values above are read before the next copy replaces them
        HIPCHK(hipStreamSynchronize(eng->stream));
    }
    eng->ftd_batches += (int64_t)ranges.size();
    return CSG_OK;
}

// The rounds from the current one on, for its n active cells (part A's
// list, counts and lists at the round's start), as one cooperative launch of
// k_ftd_tail.  *ran stays false (nothing launched) when the kernel's LDS or a
// co-resident grid does not fit.
int ftd_tail(csg_engine* eng, uint32_t n, uint32_t max_count, float quantile, float thre,
             int times, bool* ran)
{
    *ran               = false;
    const uint32_t Gf  = std::max<uint32_t>(eng->Gf, 1), L = eng->fL;
    const FtdLds   d   = ftd_lds(eng);
    const size_t   lds = ftd_tail_lds(Gf, eng->fnbins, d.wave_ints);
    eng->tt_refused    = lds > eng->tt_lds_limit ? 1 : !eng->prop.cooperativeLaunch ? 2 : 0;
    if (eng->tt_refused || n == 0)
        return CSG_OK;
    const void* fn = eng->d_fref_codes8 ? (const void*)k_ftd_tail<uint8_t>
                                        : (const void*)k_ftd_tail<uint16_t>;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, FTT_NT, lds) != hipSuccess ||
        per < 1)
    {
        (void)hipGetLastError();
        eng->tt_refused = 3;
        return CSG_OK;
    }
    const int cu   = std::max(1, eng->prop.multiProcessorCount);
    const int grid = std::min(per, std::max(1, eng->tt_blocks_per_cu)) * cu;
    // score rows: a cell's labels are at most the max_count largest ones
    std::vector<uint32_t> sz(L);
    for (uint32_t l = 0; l < L; ++l)
        sz[l] = eng->fctidx[2 * l + 1];
    std::sort(sz.begin(), sz.end(), std::greater<uint32_t>());
    uint64_t top = 0;
    for (uint32_t i = 0; i < std::min(max_count, L); ++i)
        top += sz[i];
    const uint64_t rows   = std::max<uint64_t>((uint64_t)n * top, 1);
    const uint64_t chunks = rows / FTT_CH + n;
    const uint32_t gst    = round_up(Gf, 8);
    int            rc;
    if ((rc = ensure(eng, &eng->d_tt_state, &eng->tt_cap_state, 1)) ||
        (rc = ensure(eng, &eng->d_tt_slots, &eng->tt_cap_slots, n)) ||
        (rc = ensure(eng, &eng->d_tt_chunks, &eng->tt_cap_chunks, chunks)) ||
        (rc = ensure(eng, &eng->d_tt_genes, &eng->tt_cap_genes, (size_t)n * gst)) ||
        (rc = ensure(eng, &eng->d_tt_crank, &eng->tt_cap_crank, (size_t)n * gst)) ||
        (rc = ensure(eng, &eng->d_tt_scores, &eng->tt_cap_scores, rows)) ||
        (rc = ensure(eng, &eng->d_tt_P, &eng->tt_cap_P, (size_t)n * L)) ||
        (rc = ensure(eng, &eng->d_tt_nanl, &eng->tt_cap_nanl, (size_t)n * L)))
        return rc;
    if (!eng->h_tt_state)
        HIPCHK(hipHostMalloc((void**)&eng->h_tt_state, sizeof(FtdTailState), hipHostMallocDefault));
    HIPCHK(hipMemsetAsync(eng->d_tt_state, 0, sizeof(FtdTailState), eng->stream));
    FtdTailArgs a{};
    a.cnt       = eng->d_fl_cnt;
    a.st        = eng->d_tt_state;
    a.slots     = eng->d_tt_slots;
    a.chunks    = eng->d_tt_chunks;
    a.act       = eng->d_fl_act;
    a.n         = n;
    a.lists     = eng->d_fl_lists;
    a.counts    = eng->d_fl_counts;
    a.rounds    = eng->d_fl_rounds;
    a.L         = L;
    a.times     = times;
    a.ctdidx    = eng->d_fl_ctdidx;
    a.ctdiff    = eng->d_fl_ctdiff;
    a.cut_tab   = eng->d_fl_cut;
    a.lstart    = eng->d_flstart;
    a.llen      = eng->d_fllen;
    a.Gf        = Gf;
    a.gst       = gst;
    a.nbins     = eng->fnbins;
    a.wave_ints = d.wave_ints;
    a.qdata     = eng->d_fqdata;
    a.qind      = eng->d_fqind;
    a.qptr      = eng->d_fqptr;
    a.genes     = eng->d_tt_genes;
    a.crank     = eng->d_tt_crank;
    a.scores    = eng->d_tt_scores;
    a.P         = eng->d_tt_P;
    a.nanl      = eng->d_tt_nanl;
    a.pct       = eng->d_fpct;
    a.quantile  = quantile;
    a.thre      = thre;
    const void* codes = eng->d_fref_codes8 ? (const void*)eng->d_fref_codes8
                                           : (const void*)eng->d_fref_codes;
    void* params[] = { (void*)&a, (void*)&codes };
    if (hipLaunchCooperativeKernel(fn, dim3((unsigned)grid), dim3(FTT_NT), params,
                                   (unsigned)lds, eng->stream) != hipSuccess)
    {
        // (refused: the rounds go on one launch sequence at a time)
        (void)hipGetLastError();
        eng->tt_refused = 4;
        return CSG_OK;
    }
    HIPCHK(hipMemcpyAsync(eng->h_tt_state, eng->d_tt_state, sizeof(FtdTailState),
                          hipMemcpyDeviceToHost, eng->stream));
    *ran = true;
    return CSG_OK;
}

// The whole fine-tune on the device.  Returns CSG_OK with *fallback set when
// a hash collision was met: lists / counts / rounds then hold the state at
// the start of that round for the host-grouped schedule to continue from.
int finetune_device(csg_engine* eng, float quantile, float thre, int times,
                    uint32_t* final_labels, int32_t* rounds_out, bool* fallback,
                    std::vector<uint8_t>& lists, std::vector<uint32_t>& counts,
                    std::vector<int32_t>& nrounds)
{
    *fallback         = false;
    const uint32_t Nq = eng->fNq, L = eng->fL;
    int            rc;
    const FtdLds   d = ftd_lds(eng);
    // (the set-up kernel's static LDS is about 5.5 KB)
    if (d.rank_qry > eng->lds_limit || d.setup + 8192 > std::max<size_t>(eng->lds_limit, 65536) ||
        d.rank_ref > eng->lds_limit)
        return eng->fail(CSG_ERR_ARGUMENT, "fine-tune gene set of " + std::to_string(eng->Gf) +
                                               " genes with " + std::to_string(eng->fnbins) +
                                               " value codes exceeds the rank kernels' LDS");
    if ((rc = ftd_static(eng)))
        return rc;
    const float* mask = nullptr;
    if (!eng->host_mask.empty())
    {
        if ((rc = upload(eng, &eng->d_fl_mask, eng->host_mask.data(), eng->host_mask.size())))
            return rc;
        mask = eng->d_fl_mask;
    }
    else
    {
        if (!eng->mask_valid || eng->Nq != Nq || eng->L != L)
            return eng->fail(CSG_ERR_STATE, "no candidate mask: pass one to csg_copyin or run "
                                            "csg_get_label on the same cells first");
        mask = eng->d_mask;
    }
    HIPCHK(hipMemsetAsync(eng->d_fl_cnt, 0, sizeof(FtdCounts), eng->stream));
    k_ftd_init<<<nblk(Nq), 256, 0, eng->stream>>>(mask, Nq, L, times, eng->d_fl_lists,
                                                 eng->d_fl_counts, eng->d_fl_rounds,
                                                 eng->d_fl_next, eng->d_fl_cnt);
    LAUNCHCHK();
    FtdCounts& h = *eng->h_fl_cnt;
    HIPCHK(hipMemcpyAsync(&h, eng->d_fl_cnt, sizeof(FtdCounts), hipMemcpyDeviceToHost,
                          eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    if ((rc = ftd_groups(eng, std::max<uint32_t>(h.n_next, 1))))
        return rc;
    // the grouping of a round of few cells runs in one block; a round's
    // active cells are at most the previous round's
    eng->ftd_small_a = h.n_next <= (uint32_t)FTD_SMALL_NT;
    // a round removes at least one label from every active cell (its
    // lowest-scoring ones), so rounds after one whose largest list held m
    // labels number at most m - 2, and the compaction of the last of them
    // opens round m - 1 (which then has no active cells)
    uint32_t max_count = L;
    bool     comp      = false;  // part A opens with the previous round's compaction
    bool     tail      = false;  // the last rounds ran in the tail kernel
    eng->tt_ran        = false;
    auto     state_to_host = [&]() -> int {
        // state at the start of the current round, for the host schedule
        *fallback = true;
        lists.resize((size_t)Nq * L);
        counts.resize(Nq);
        nrounds.resize(Nq);
        HIPCHK(hipMemcpy(lists.data(), eng->d_fl_lists, (size_t)Nq * L, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(counts.data(), eng->d_fl_counts, (size_t)Nq * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(nrounds.data(), eng->d_fl_rounds, (size_t)Nq * 4, hipMemcpyDeviceToHost));
        return CSG_OK;
    };
    auto add_totals = [&]() {
        eng->counters[6] += h.tot_rounds;
        eng->counters[7] += (int64_t)h.tot_active;
        eng->counters[8] += (int64_t)h.tot_groups;
        eng->counters[9] += (int64_t)h.tot_scores;
        eng->counters[22] += (int64_t)h.tot_fused;
        eng->counters[10] += (int64_t)h.pair_genes;
        if (tail)
        {
            // rounds after the tail's first (which part A counted); each
            // cell a group of its own
            const FtdTailState& t = *eng->h_tt_state;
            eng->counters[6] += t.rounds;
            eng->counters[7] += (int64_t)t.tot_active;
            eng->counters[8] += (int64_t)t.tot_active;
            eng->counters[9] += (int64_t)t.tot_scores;
            eng->counters[10] += (int64_t)t.pair_genes;
            eng->counters[25] += t.rounds + 1;
        }
    };
    for (;;)
    {
        const int64_t t0 = now_us();
        // few-cell rounds run in chains: the rounds a list of max_count
        // labels can still take, launched together with the last round's
        // launch mask and work tables, one read-back per chain
        const int chain = comp && eng->ftd_small_a && eng->ftd_chain_max > 0 &&
                                  eng->phase_mode != 3
                              ? (int)std::min<uint32_t>(max_count > 2 ? max_count - 1 : 1u,
                                                        (uint32_t)eng->ftd_chain_max)
                              : 0;
        if (chain > 0)
        {
            hipGraphExec_t gC = nullptr;
            if ((rc = ftd_graph_chain(eng, chain, eng->ftd_mask, quantile, thre, times, &gC)))
                return rc;
            HIPCHK(hipGraphLaunch(gC, eng->stream));
            HIPCHK(hipMemcpyAsync(&h, eng->d_fl_cnt, sizeof(FtdCounts), hipMemcpyDeviceToHost,
                                  eng->stream));
            HIPCHK(hipStreamSynchronize(eng->stream));
            eng->counters[15] += now_us() - t0;
            eng->counters[23] += chain;
            if (h.undefined)
                return eng->fail(CSG_ERR_UNDEFINED,
                                 "a NaN-carrying fine-tune segment makes the reference's sort "
                                 "leave its array (undefined behaviour)");
            if (!h.halt)
            {
                if (h.n_active == 0)
                    break;  // the chain ran past the last round
                max_count = h.max_count;
                continue;   // (the next part A opens with this chain's last compaction)
            }
            // a halted round: its compaction ran, the rest runs below from
            // its part A (tables grown, its own launch mask)
            eng->counters[24]++;
            k_ftd_resume<<<1, 64, 0, eng->stream>>>(eng->d_fl_cnt);
            LAUNCHCHK();
            comp = false;
        }
        const bool      small = eng->ftd_small_a;
        hipGraphExec_t*         gA   = comp ? &eng->ftd_gAc[small] : small ? &eng->ftd_gAs : &eng->ftd_gA;
        std::vector<uintptr_t>* sA   = comp ? &eng->ftd_sigAc[small] : small ? &eng->ftd_sigAs : &eng->ftd_sigA;
        std::vector<uintptr_t>  sig  = ftd_sig_a(eng);
        if (comp)
            sig.push_back((uintptr_t)(uint32_t)times);
        if ((rc = ftd_graph(eng, gA, sA, sig,
                            [&]() { return ftd_launch_a(eng, small, comp ? times : -1); })))
            return rc;
        HIPCHK(hipGraphLaunch(*gA, eng->stream));
        HIPCHK(hipMemcpyAsync(&h, eng->d_fl_cnt, sizeof(FtdCounts), hipMemcpyDeviceToHost,
                              eng->stream));
        HIPCHK(hipStreamSynchronize(eng->stream));
        if (h.undefined)
            return eng->fail(CSG_ERR_UNDEFINED,
                             "a NaN-carrying fine-tune segment makes the reference's sort leave "
                             "its array (undefined behaviour)");
        if (h.n_active == 0)
            break;
        if (h.collision)
        {
            add_totals();
            return state_to_host();
        }
        max_count = h.max_count;
        if (eng->ftd_tail_max && h.n_active <= eng->ftd_tail_max)
        {
            // this round and the rest in one launch (part A's grouping unused)
            bool ran = false;
            if ((rc = ftd_tail(eng, h.n_active, h.max_count, quantile, thre, times, &ran)))
                return rc;
            if (ran)
            {
                tail = eng->tt_ran = true;
                eng->counters[15] += now_us() - t0;
                break;
            }
        }
        eng->ftd_small_a = h.n_active <= (uint32_t)FTD_SMALL_NT;
        eng->ftd_mask    = ftd_mask_of(eng, h);
        if (ftd_bytes(h, L) <= eng->ftd_budget)
        {
            if ((rc = ftd_work(eng, h)))
                return rc;
            if (!eng->ftd_tuned && eng->ftd_autotune && h.n_cells >= 4096 &&
                (rc = ftd_tune(eng, quantile)))
                return rc;
            if (eng->phase_mode == 3)
            {
                // part B launched eagerly, its kernel families timed
                eng->ftd_eager = true;
                rc             = ftd_launch_b(eng, quantile, thre, times);
                eng->ftd_eager = false;
                if (rc)
                    return rc;
            }
            else
            {
                hipGraphExec_t gB = nullptr;
                if ((rc = ftd_graph_b(eng, quantile, thre, times, &gB)))
                    return rc;
                HIPCHK(hipGraphLaunch(gB, eng->stream));
            }
        }
        else if ((rc = ftd_batches(eng, h, quantile, thre, times)))
            return rc;
        // (the compaction into the next round's active list leads that
        // round's part A)
        comp = true;
        eng->counters[15] += now_us() - t0;
    }
    k_ftd_final<<<nblk(Nq), 256, 0, eng->stream>>>(eng->d_fl_lists, eng->d_fl_counts, Nq, L,
                                                  eng->d_fl_final);
    LAUNCHCHK();
    // the labels and round counts come back through page-locked memory (one
    // DMA each) and are copied out on the host: a copy into the caller's
    // pageable memory is staged by the runtime piece by piece
    const size_t nb = (size_t)Nq * 4;
    if ((rc = ensure_host(eng, &eng->h_ft_out, &eng->h_ft_out_cap, 2 * nb)))
        return rc;
    HIPCHK(hipMemcpyAsync(eng->h_ft_out, eng->d_fl_final, nb, hipMemcpyDeviceToHost,
                          eng->stream));
    if (rounds_out)
        HIPCHK(hipMemcpyAsync(eng->h_ft_out + nb, eng->d_fl_rounds, nb, hipMemcpyDeviceToHost,
                              eng->stream));
    HIPCHK(hipMemcpyAsync(&h, eng->d_fl_cnt, sizeof(FtdCounts), hipMemcpyDeviceToHost,
                          eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    std::memcpy(final_labels, eng->h_ft_out, nb);
    if (rounds_out)
        std::memcpy(rounds_out, eng->h_ft_out + nb, nb);
    eng->counters[11] += h.tot_redo;
    add_totals();
    if (tail && eng->h_tt_state->timeout)
        return eng->fail(CSG_ERR_DEVICE, "the fine-tune tail kernel's grid barrier timed out");
    if (h.undefined)
        return eng->fail(CSG_ERR_UNDEFINED,
                         "a NaN-carrying fine-tune segment makes the reference's sort leave its "
                         "array (undefined behaviour)");
    return CSG_OK;
}

// Host-grouped schedule: rounds grouped by candidate set on the host, the
// device ranking, correlating and filtering every group in batches.  Starts
// from the given lists / counts / round counts.
int finetune_host(csg_engine* eng, float quantile, float finetune_thre, int finetune_times,
                  std::vector<uint8_t>& lists, std::vector<uint32_t>& counts,
                  std::vector<int32_t>& nrounds, int64_t t_start)
{
    const uint32_t L  = eng->fL;
    const uint32_t Nq = eng->fNq;
    int            rc = CSG_OK;
    const size_t          gwords = (std::max<uint32_t>(eng->Gf, 1) + 63) / 64;
    std::vector<uint64_t> gbits(gwords);
    std::vector<uint32_t> gsub;
    const uint64_t        budget_elems = (uint64_t)1 << 28;  // per batch (scores + ranks)
    FtBatch               b;
    std::vector<uint32_t> active;
    active.reserve(Nq);
    for (uint32_t c = 0; c < Nq; ++c)
        if (counts[c] > 1 && (finetune_times == 0 || nrounds[c] < finetune_times))
            active.push_back(c);

    int64_t t_mark = now_us();
    eng->counters[12] += t_mark - t_start;  // mask -> candidate lists
    std::unordered_map<SetKey, uint32_t, SetKeyHash> gid;
    std::vector<std::vector<uint32_t>>              gcells;
    std::vector<SetKey>                             gkey;
    while (!active.empty())
    {
        // active cells grouped by candidate set
        gid.clear();
        gcells.clear();
        gkey.clear();
        gid.reserve(active.size() / 4 + 16);
        for (uint32_t c : active)
        {
            SetKey key{ { 0, 0, 0, 0 } };
            for (uint32_t t = 0; t < counts[c]; ++t)
            {
                const uint32_t l = lists[(size_t)c * L + t];
                key.w[l / 64] |= 1ull << (l % 64);
            }
            auto it = gid.find(key);
            if (it == gid.end())
            {
                it = gid.emplace(key, (uint32_t)gcells.size()).first;
                gcells.emplace_back();
                gkey.push_back(key);
            }
            gcells[it->second].push_back(c);
        }
        eng->counters[6]++;
        {
            const int64_t t = now_us();
            eng->counters[13] += t - t_mark;  // grouping
            t_mark = t;
        }
        b.clear();
        auto flush = [&]() -> int {
            if (b.groups.empty())
                return CSG_OK;
            const int64_t t0 = now_us();
            eng->counters[14] += t0 - t_mark;  // table building
            int r = run_ft_batch(eng, b, quantile, finetune_thre, lists, counts);
            b.clear();
            t_mark = now_us();
            eng->counters[15] += t_mark - t0;  // batch upload + kernels + download
            return r;
        };
        for (size_t gi = 0; gi < gcells.size(); ++gi)
        {
            std::vector<uint32_t> labs;
            for (uint32_t l = 0; l < L; ++l)
                if (gkey[gi].w[l / 64] >> (l % 64) & 1)
                    labs.push_back(l);
            const int k = (int)labs.size();
            // G_sub: union of each ordered pair's first min(len, gene_thre)
            // genes (src/cusingler.cu:1140-1156), as a union of cached window
            // bitsets; scanning the bits yields the genes in ascending order
            const int gene_thre = (int)std::round(500 * std::pow((2 / 3.0), std::log2((double)k)));
            std::fill(gbits.begin(), gbits.end(), 0ull);
            for (uint32_t i : labs)
                for (uint32_t j : labs)
                {
                    if (i == j)
                        continue;
                    const uint64_t* w = window_bits(eng, i, j, gene_thre);
                    for (size_t t = 0; t < gwords; ++t)
                        gbits[t] |= w[t];
                }
            gsub.clear();
            for (size_t t = 0; t < gwords; ++t)
            {
                uint64_t v = gbits[t];
                while (v)
                {
                    gsub.push_back((uint32_t)(t * 64 + __builtin_ctzll(v)));
                    v &= v - 1;
                }
            }
            auto& cells = gcells[gi];
            eng->counters[7] += cells.size();
            eng->counters[8] += 1;
            for (uint32_t c : cells)
                nrounds[c]++;
            if (gsub.size() < 20)
            {
                // fewer than 20 genes: the round returns {front} (:1158-1159)
                for (uint32_t c : cells)
                    counts[c] = 1;
                continue;
            }
            uint32_t nref = 0;
            for (uint32_t l : labs)
                nref += eng->fctidx[2 * l + 1];
            const uint32_t stride = (uint32_t)round_up(gsub.size(), 8);
            // split very large groups by cells so that one batch stays bounded
            size_t cpos = 0;
            while (cpos < cells.size())
            {
                const uint64_t fixed = (uint64_t)nref * stride;
                const uint64_t per   = (uint64_t)stride + nref;
                const uint64_t used  = b.ref_elems + b.cell_elems + b.n_scores + fixed;
                const uint64_t room  = budget_elems > used ? budget_elems - used : 0;
                size_t take = (size_t)std::min<uint64_t>(cells.size() - cpos, room / per);
                if (take == 0)
                {
                    if (!b.groups.empty())
                    {
                        if ((rc = flush()))
                            return rc;
                        continue;
                    }
                    take = 1;
                }
                FtGroup g{};
                g.gene_off   = (uint32_t)b.genes.size();
                g.gs         = (uint32_t)gsub.size();
                g.stride     = stride;
                g.nref       = nref;
                g.ncell      = (uint32_t)take;
                g.ref_item0  = (uint32_t)b.n_ref_items;
                g.cell_item0 = (uint32_t)b.cell_items.size();
                g.label_off  = (uint32_t)b.labels.size();
                g.nlabel     = (uint32_t)labs.size();
                g.ref_dst0   = b.ref_elems;
                g.cell_dst0  = b.cell_elems;
                g.score0     = b.n_scores;
                const uint32_t gidx = (uint32_t)b.groups.size();
                b.genes.insert(b.genes.end(), gsub.begin(), gsub.end());
                b.labels.insert(b.labels.end(), labs.begin(), labs.end());
                {
                    uint32_t o = 0;
                    for (uint32_t l : labs)
                    {
                        b.label_pos.push_back(o);
                        o += eng->fctidx[2 * l + 1];
                    }
                }
                b.n_ref_items += nref;
                b.ref_elems += (uint64_t)nref * stride;
                for (size_t t = 0; t < take; ++t)
                {
                    const uint32_t c = cells[cpos + t];
                    b.cell_items.push_back(FtRankItem{ c, gidx, b.cell_elems });
                    b.cell_elems += stride;
                    b.cells.push_back(FtCell{ c, gidx, (uint32_t)t, 0, b.n_segs });
                    b.cell_ids.push_back(c);
                    b.n_segs += labs.size();
                    b.in_lists.insert(b.in_lists.end(), &lists[(size_t)c * L],
                                      &lists[(size_t)c * L] + L);
                    b.in_counts.push_back(counts[c]);
                }
                eng->counters[9] += (int64_t)take * nref;
                eng->counters[10] += (int64_t)take * nref * gsub.size();
                b.n_scores += (uint64_t)take * nref;
                b.groups.push_back(g);
                cpos += take;
                if (b.ref_elems + b.cell_elems + b.n_scores >= budget_elems)
                    if ((rc = flush()))
                        return rc;
            }
        }
        if ((rc = flush()))
            return rc;
        const int64_t t_keep = now_us();
        eng->counters[14] += t_keep - t_mark;
        t_mark = t_keep;
        size_t keep = 0;
        for (uint32_t c : active)
            if (counts[c] > 1 && (finetune_times == 0 || nrounds[c] < finetune_times))
                active[keep++] = c;
        active.resize(keep);
    }
    return CSG_OK;
}

}  // namespace

extern "C" {

int finetune_entry(csg_engine* eng, float quantile, float finetune_thre, int finetune_times,
                   uint32_t* final_labels, int32_t* rounds_out);

int csg_finetune(csg_engine* eng, float quantile, float finetune_thre, int finetune_times,
                 uint32_t* final_labels, int32_t* rounds_out)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    // a score step left in flight runs ahead of this call's work on the
    // stream; its read-back is folded once the fine-tune has synchronised
    int       rc  = finetune_entry(eng, quantile, finetune_thre, finetune_times, final_labels,
                                   rounds_out);
    const int rc2 = fold_score_step(eng);
    return rc ? rc : rc2;
}

int finetune_entry(csg_engine* eng, float quantile, float finetune_thre, int finetune_times,
                   uint32_t* final_labels, int32_t* rounds_out)
{
    if (!eng->ft_loaded)
        return eng->fail(CSG_ERR_STATE, "csg_finetune before csg_copyin");
    if (!eng->fq_ready)
        return eng->fail(CSG_ERR_STATE, "no query data: csg_copyin reserved room only; load a "
                                        "batch with csg_copyin_query_raw");
    HIPCHK(hipSetDevice(eng->device));
    const uint32_t L  = eng->fL;
    const uint32_t Nq = eng->fNq;
    int            rc = set_pct(eng, &eng->d_fpct, &eng->fpct_quantile, eng->fctidx, L, quantile);
    if (rc)
        return rc;
    for (int i = 6; i < 16; ++i)  // the score step's own counters stay
        eng->counters[i] = 0;
    for (int i : { 22, 23, 24, 25 })
        eng->counters[i] = 0;
    PhaseScope    total(eng, 5, 0);
    const int64_t t_start = now_us();
    // candidate lists, counts and round counts on the host: filled by the
    // device schedule only when it hands over to the host-grouped one (sized
    // there), so that a device fine-tune allocates none of them
    std::vector<uint8_t>  lists;
    std::vector<uint32_t> counts;
    std::vector<int32_t>  nrounds;
    if (eng->ft_schedule == 0 && L <= 256)
    {
        bool fallback = false;
        rc = finetune_device(eng, quantile, finetune_thre, finetune_times, final_labels, rounds_out,
                             &fallback, lists, counts, nrounds);
        if (rc || !fallback)
            return rc;
        if ((rc = finetune_host(eng, quantile, finetune_thre, finetune_times, lists, counts,
                                nrounds, t_start)))
            return rc;
        for (uint32_t c = 0; c < Nq; ++c)
        {
            final_labels[c] = counts[c] == 0 ? 0u : (uint32_t)lists[(size_t)c * L];
            if (rounds_out)
                rounds_out[c] = nrounds[c];
        }
        return CSG_OK;
    }
    lists.assign((size_t)Nq * L, 0);
    counts.assign(Nq, 0);
    nrounds.assign(Nq, 0);
    // candidate lists from the mask, ascending label order (src/cusingler.cu:1307-1312)
    if (!eng->host_mask.empty())
    {
        const float* mask = eng->host_mask.data();
        for (uint32_t c = 0; c < Nq; ++c)
            for (uint32_t l = 0; l < L; ++l)
                if (mask[(size_t)c * L + l] != 0)
                    lists[(size_t)c * L + counts[c]++] = (uint8_t)l;
    }
    else
    {
        if (!eng->mask_valid || eng->Nq != Nq || eng->L != L)
            return eng->fail(CSG_ERR_STATE, "no candidate mask: pass one to csg_copyin or run "
                                            "csg_get_label on the same cells first");
        const uint32_t W = (L + 63) / 64;
        uint64_t*      d_bits = nullptr;
        if ((rc = dalloc(eng, &d_bits, (size_t)Nq * W)))
            return rc;
        std::vector<uint64_t> bits((size_t)Nq * W);
        k_pack_mask<<<(Nq + 255) / 256, 256, 0, eng->stream>>>(eng->d_mask, Nq, L, d_bits);
        hipError_t e1 = hipGetLastError();
        hipError_t e2 = hipMemcpyAsync(bits.data(), d_bits, bits.size() * 8, hipMemcpyDeviceToHost,
                                       eng->stream);
        hipError_t e3 = hipStreamSynchronize(eng->stream);
        dfree(d_bits);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
            return eng->fail(CSG_ERR_DEVICE, "mask copy failed");
        for (uint32_t c = 0; c < Nq; ++c)
            for (uint32_t w = 0; w < W; ++w)
            {
                uint64_t v = bits[(size_t)c * W + w];
                while (v)
                {
                    const int t = __builtin_ctzll(v);
                    v &= v - 1;
                    lists[(size_t)c * L + counts[c]++] = (uint8_t)(w * 64 + t);
                }
            }
    }

    if ((rc = finetune_host(eng, quantile, finetune_thre, finetune_times, lists, counts, nrounds,
                            t_start)))
        return rc;
    for (uint32_t c = 0; c < Nq; ++c)
    {
        final_labels[c] = counts[c] == 0 ? 0u : (uint32_t)lists[(size_t)c * L];
        if (rounds_out)
            rounds_out[c] = nrounds[c];
    }
    return CSG_OK;
}

int csg_score_matrix(csg_engine* eng, int path, uint32_t q0, uint32_t nq, float* out)
{
    if (!eng || !out)
        return CSG_ERR_ARGUMENT;
    if (!eng->score_loaded)
        return eng->fail(CSG_ERR_STATE, "csg_score_matrix before csg_copyin_score");
    if (path < CSG_PATH_PLAIN || path > CSG_PATH_FP16_MFMA)
        return eng->fail(CSG_ERR_ARGUMENT, "unknown correlation path");
    if (nq == 0 || nq > eng->Qc_dense || q0 > eng->Nq || nq > eng->Nq - q0)
        return eng->fail(CSG_ERR_ARGUMENT, "query rows outside the loaded cells or above one chunk");
    HIPCHK(hipSetDevice(eng->device));
    const size_t nr_pad = round_up(eng->Nr, 256);
    const size_t qc     = eng->Qc;
    int          rc;
    if (path == CSG_PATH_EXACT_INT8 && !eng->d_rh)
    {
        if ((rc = dalloc(eng, &eng->d_rh, nr_pad * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_rl, nr_pad * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_qh, qc * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_ql, qc * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_dr, nr_pad)) || (rc = dalloc(eng, &eng->d_dq, qc)) ||
            (rc = dalloc(eng, &eng->d_cstatus, qc)) ||
            (rc = dalloc(eng, &eng->d_exact, qc * eng->L)) ||
            (rc = dalloc(eng, &eng->d_items, qc * eng->L)) ||
            (rc = dalloc(eng, &eng->d_nitems, 2)) || (rc = dalloc(eng, &eng->d_Pref, qc * eng->L)))
            return rc;
    }
    if (path == CSG_PATH_FP16_MFMA && !eng->d_rf16)
    {
        if ((rc = dalloc(eng, &eng->d_rf16, nr_pad * eng->Kp)) ||
            (rc = dalloc(eng, &eng->d_qf16, qc * eng->Kp)))
            return rc;
    }
    if (!eng->ref_ranked)
    {
        if ((rc = rank_rows(eng, eng->d_rdata, eng->d_rind, eng->d_rptr, 0, eng->Nr, eng->d_Xr,
                            eng->d_sqr_f, eng->d_irr)))
            return rc;
        eng->ref_ranked        = true;
        eng->ref_prepared_path = -1;
    }
    if ((path == CSG_PATH_EXACT_INT8 || path == CSG_PATH_FP16_MFMA) &&
        (eng->ref_prepared_path != path ||
         (path == CSG_PATH_EXACT_INT8 && eng->ref_kf != use_kf(eng))))
    {
        const bool kf = path == CSG_PATH_EXACT_INT8 && use_kf(eng);
        if ((rc = prepare_operands(eng, path, eng->d_Xr, nr_pad, eng->d_rh, eng->d_rl, eng->d_rf16,
                                   kf, eng->d_dr)))
            return rc;
        eng->ref_prepared_path = path;
        eng->ref_kf            = kf;
    }
    if ((rc = rank_rows(eng, eng->d_qdata, eng->d_qind, eng->d_qptr, q0, nq, eng->d_Xq,
                        path == CSG_PATH_EXACT_INT8 ? nullptr : eng->d_sqq_f, eng->d_irq)) ||
        (rc = prepare_operands(eng, path, eng->d_Xq, round_up(nq, 128), eng->d_qh, eng->d_ql,
                               eng->d_qf16, path == CSG_PATH_EXACT_INT8 && eng->ref_kf,
                               eng->d_dq)) ||
        (rc = score_chunk(eng, path, nq)))
        return rc;
    HIPCHK(hipMemcpyAsync(out, eng->d_S, (size_t)nq * eng->Nr * sizeof(float),
                          hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return CSG_OK;
}

int csg_set_finetune_schedule(csg_engine* eng, int schedule)
{
    if (!eng || schedule < 0 || schedule > 1)
        return CSG_ERR_ARGUMENT;
    eng->ft_schedule = schedule;
    return CSG_OK;
}

int csg_set_finetune_budget(csg_engine* eng, uint64_t bytes, int64_t* batches)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (bytes)
        eng->ftd_budget = bytes;
    if (batches)
        *batches = eng->ftd_batches;
    return CSG_OK;
}

namespace {

// The second stream of the query uploads, its events and the coding counters
// (created on first use).
int ensure_copy_stream(csg_engine* eng)
{
    int rc;
    if (!eng->d_code_misc && (rc = dalloc(eng, &eng->d_code_misc, 32)))
        return rc;
    if (eng->copy_stream)
        return CSG_OK;
    HIPCHK(hipStreamCreateWithFlags(&eng->copy_stream, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b)
    {
        HIPCHK(hipEventCreateWithFlags(&eng->code_copied[b], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&eng->code_done[b], hipEventDisableTiming));
    }
    for (auto& ps : eng->pre)
        HIPCHK(hipEventCreateWithFlags(&ps.copied, hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void**)&eng->h_code_misc, 32, hipHostMallocDefault));
    return CSG_OK;
}

// The coding passes of rows [0, nr) of a device CSR (row r's entries at
// data/idx [ptr[r] - base, ptr[r + 1] - base)) into sets a and b, on `st`.
// Scratch: cnt_a / cnt_b (nr each), over (2 nr), misc (32 bytes, cleared by
// the caller; its running bases continue from earlier slices).  No
// allocation, no synchronisation.
extern "C++" {
template <class IT>
void launch_coding(const csg_engine* eng, hipStream_t st, const float* data, const IT* idx,
                   const int32_t* ptr, int64_t base, uint32_t nr, uint32_t width,
                   const CodeSet& sa, const CodeSet& sb, uint32_t* over, uint8_t* misc)
{
    uint32_t* d_flags = (uint32_t*)misc;
    uint32_t* d_nover = d_flags + 1;
    uint32_t* d_maxu  = d_flags + 2;
    int64_t*  d_base  = (int64_t*)(misc + 16);
    const int big_cap = (int)std::min<size_t>(eng->lds_limit / 8, (size_t)1 << 20);
    // one wave per row (rows strided over the grid beyond 2^20 blocks)
    const unsigned g  = (unsigned)std::max<uint32_t>(1, std::min<uint32_t>((nr + 3) / 4, 1u << 20));
    k_code_count<IT><<<g, CODE_NT, 0, st>>>(idx, ptr, base, nr, width, sa.map, sb.map, sa.count,
                                        sb.count, d_flags);
    k_code_scan<<<1, 1024, 0, st>>>(sa.count, sb.count, nr, (int32_t*)sa.ptr, (int32_t*)sb.ptr,
                                    d_base);
    k_code_zero<<<1, 64, 0, st>>>(d_nover);
    // rows strided over a grid of 8 blocks per CU: each wave codes many rows
    // and reports its maximum once
    const unsigned gr = (unsigned)std::max<uint32_t>(
        1, std::min<uint32_t>((nr + 3) / 4, (uint32_t)std::max(1, eng->prop.multiProcessorCount) * 8));
    k_code_rows<IT><<<gr, CODE_NT, 0, st>>>(data, idx, ptr, base, nr, width, sa, sb, d_maxu, over,
                                        d_nover, d_flags);
    k_code_rows_big<IT><<<256, CODE_NT, (size_t)big_cap * 8, st>>>(data, idx, ptr, base, width, sa, sb,
                                                               over, d_nover, big_cap, d_maxu,
                                                               d_flags);
}
}  // extern "C++"

// Status of a finished coding (misc read back): CSG_OK or the error.
int coding_status(csg_engine* eng, const uint32_t* misc)
{
    if (misc[0] & CODE_BADCOL)
        return eng->fail(CSG_ERR_ARGUMENT, "a query column index lies outside the query width");
    if (misc[0] & CODE_NAN)
        return eng->fail(CSG_ERR_UNDEFINED, "NaN query value (the reference's std::set<float> "
                                            "ordering is undefined for it)");
    if (misc[0] & CODE_OVERFLOW)
        return eng->fail(CSG_ERR_UNDEFINED, "a query row holds more distinct values than uint16 "
                                            "codes (65535) or the block coder's LDS");
    return CSG_OK;
}

}  // namespace

int csg_train(csg_engine* eng, uint32_t ref_height, uint32_t ref_width, const float* ref_data,
              const int32_t* ref_indices, const int32_t* ref_indptr, const int32_t* ref_map,
              uint32_t width, uint32_t ct_num, const int32_t* celltype, uint64_t* sizes)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (ref_height == 0 || ref_width == 0 || width == 0 || !ref_indptr || !celltype)
        return eng->fail(CSG_ERR_ARGUMENT, "empty reference matrix");
    if (ct_num == 0 || ct_num > 256)
        return eng->fail(CSG_ERR_ARGUMENT, "1 to 256 labels (the reference's uint8 codes)");
    HIPCHK(hipSetDevice(eng->device));
    const uint32_t L = ct_num, W = ref_width;
    // label codes narrowed to uint8 (src/io.cpp:813); each label needs a cell
    std::vector<uint8_t>  lab(ref_height);
    std::vector<uint32_t> nlab(L, 0);
    for (uint32_t r = 0; r < ref_height; ++r)
    {
        lab[r] = (uint8_t)celltype[r];
        if (lab[r] >= L)
            return eng->fail(CSG_ERR_ARGUMENT, "cell type code without a label name");
        nlab[lab[r]]++;
        if (ref_indptr[r + 1] < ref_indptr[r])
            return eng->fail(CSG_ERR_ARGUMENT, "reference row pointers decrease");
    }
    for (uint32_t l = 0; l < L; ++l)
        if (nlab[l] == 0)
            return eng->fail(CSG_ERR_UNDEFINED, "label " + std::to_string(l) +
                                                    " has no reference cell (its median is "
                                                    "undefined in the reference)");
    const size_t k_pair = (size_t)std::round(500 * std::pow((2 / 3.0), std::log2((double)L)));
    const size_t thre_n = (size_t)std::round(500 * std::pow((2 / 3.0), std::log2(2)));
    if (k_pair > W)
        return eng->fail(CSG_ERR_UNDEFINED, "fewer reference genes than the score gene count per "
                                            "label pair");
    const size_t nnz  = (size_t)(ref_indptr[ref_height] - ref_indptr[0]);
    const size_t nseg = (size_t)L * W;
    if (ref_indptr[0] != 0)
        return eng->fail(CSG_ERR_ARGUMENT, "reference row pointers must start at 0");
    // device work memory, freed on return
    float *    d_data = nullptr, *d_vals = nullptr, *d_med = nullptr;
    int32_t *  d_ind = nullptr, *d_ptr = nullptr, *d_map = nullptr;
    uint8_t*   d_lab = nullptr;
    uint32_t * d_cnt = nullptr, *d_fill = nullptr, *d_seg = nullptr, *d_big = nullptr,
             *d_misc = nullptr, *d_nlab = nullptr, *d_npos = nullptr, *d_nzero = nullptr,
             *d_vout = nullptr, *d_bits = nullptr;
    uint64_t * d_off = nullptr, *d_tiles = nullptr, *d_scr = nullptr, *d_voff = nullptr;
    auto cleanup = [&]() {
        dfree(d_data); dfree(d_vals); dfree(d_med); dfree(d_ind); dfree(d_ptr); dfree(d_map);
        dfree(d_lab); dfree(d_cnt); dfree(d_fill); dfree(d_seg); dfree(d_big); dfree(d_misc);
        dfree(d_nlab); dfree(d_npos); dfree(d_nzero); dfree(d_vout); dfree(d_bits); dfree(d_off);
        dfree(d_tiles); dfree(d_scr); dfree(d_voff);
    };
    int rc = CSG_OK;
    auto run = [&]() -> int {
        const size_t ntile = (nseg + TRAIN_TILE - 1) / TRAIN_TILE;
        if ((rc = upload(eng, &d_data, ref_data, std::max<size_t>(nnz, 1))) ||
            (rc = upload(eng, &d_ind, ref_indices, std::max<size_t>(nnz, 1))) ||
            (rc = upload(eng, &d_ptr, ref_indptr, (size_t)ref_height + 1)) ||
            (rc = upload(eng, &d_lab, lab.data(), lab.size())) ||
            (rc = upload(eng, &d_nlab, nlab.data(), nlab.size())) ||
            (ref_map && (rc = upload(eng, &d_map, ref_map, width))) ||
            (rc = dalloc(eng, &d_cnt, nseg)) || (rc = dalloc(eng, &d_fill, nseg)) ||
            (rc = dalloc(eng, &d_off, nseg)) || (rc = dalloc(eng, &d_seg, std::max<size_t>(nnz, 1))) ||
            (rc = dalloc(eng, &d_vals, std::max<size_t>(nnz, 1))) || (rc = dalloc(eng, &d_med, nseg)) ||
            (rc = dalloc(eng, &d_big, nseg)) || (rc = dalloc(eng, &d_misc, 4)) ||
            (rc = dalloc(eng, &d_tiles, ntile)))
            return rc;
        const unsigned fg = 2048;
        k_train_fill<<<fg, 256, 0, eng->stream>>>(d_cnt, nseg, 0u);
        k_train_fill<<<fg, 256, 0, eng->stream>>>(d_fill, nseg, 0u);
        k_train_fill<<<1, 64, 0, eng->stream>>>(d_misc, 4, 0u);
        const unsigned rg = (unsigned)std::min<uint32_t>((ref_height + 3) / 4, 1u << 20);
        k_train_count<<<rg, TRAIN_NT, 0, eng->stream>>>(d_ind, d_ptr, ref_height, d_lab, d_map,
                                                        width, W, d_cnt, d_seg, d_misc);
        k_train_tile_sum<<<(unsigned)ntile, 1024, 0, eng->stream>>>(d_cnt, nseg, d_tiles);
        LAUNCHCHK();
        std::vector<uint64_t> tiles(ntile);
        HIPCHK(hipMemcpyAsync(tiles.data(), d_tiles, ntile * 8, hipMemcpyDeviceToHost, eng->stream));
        uint32_t misc[4];
        HIPCHK(hipMemcpyAsync(misc, d_misc, 16, hipMemcpyDeviceToHost, eng->stream));
        HIPCHK(hipStreamSynchronize(eng->stream));
        if (misc[0])
            return eng->fail(CSG_ERR_ARGUMENT, "a reference column index lies outside its width");
        uint64_t run_off = 0;
        for (auto& t : tiles)
        {
            const uint64_t v = t;
            t                = run_off;
            run_off += v;
        }
        HIPCHK(hipMemcpyAsync(d_tiles, tiles.data(), ntile * 8, hipMemcpyHostToDevice, eng->stream));
        k_train_tile_scan<<<(unsigned)ntile, 1024, 0, eng->stream>>>(d_cnt, nseg, d_tiles, d_off);
        k_train_scatter<<<4096, 256, 0, eng->stream>>>(d_data, d_seg, nnz, d_off, d_fill, d_vals);
        k_train_median<<<8192, TRAIN_NT, 0, eng->stream>>>(d_cnt, d_off, d_vals, d_nlab, L, W,
                                                          d_med, d_big, d_misc + 1);
        LAUNCHCHK();
        HIPCHK(hipMemcpyAsync(misc, d_misc, 16, hipMemcpyDeviceToHost, eng->stream));
        HIPCHK(hipStreamSynchronize(eng->stream));
        if (misc[1])
        {
            // segments longer than a wave's LDS: sorted in global scratch
            size_t slice = 64;
            const uint32_t nmax = *std::max_element(nlab.begin(), nlab.end());
            while (slice < nmax)
                slice <<= 1;
            const unsigned bg = std::min<uint32_t>(misc[1], 256u);
            if ((rc = dalloc(eng, &d_scr, (size_t)bg * slice)))
                return rc;
            k_train_median_big<<<bg, TRAIN_NT, 0, eng->stream>>>(d_cnt, d_off, d_vals, d_nlab, W,
                                                                d_big, d_misc + 1, d_scr, slice,
                                                                d_med);
            LAUNCHCHK();
            dfree(d_scr);
        }
        // positive / zero differences per ordered pair, then their windows
        const size_t npair = (size_t)L * L;
        if ((rc = dalloc(eng, &d_npos, npair)) || (rc = dalloc(eng, &d_nzero, npair)))
            return rc;
        const unsigned pg = (unsigned)std::min<size_t>(npair, 4096);
        k_train_pair_count<<<pg, TRAIN_NT, 0, eng->stream>>>(d_med, L, W, d_npos, d_nzero);
        LAUNCHCHK();
        std::vector<uint32_t> npos(npair), nzero(npair);
        HIPCHK(hipMemcpyAsync(npos.data(), d_npos, npair * 4, hipMemcpyDeviceToHost, eng->stream));
        HIPCHK(hipMemcpyAsync(nzero.data(), d_nzero, npair * 4, hipMemcpyDeviceToHost, eng->stream));
        HIPCHK(hipStreamSynchronize(eng->stream));
        std::vector<uint64_t> voff(npair, 0);
        eng->tr_idxs.assign(2 * npair, 0);
        uint64_t total = 0;
        size_t   kmax  = 64;
        for (size_t pr = 0; pr < npair; ++pr)
        {
            if (pr / L == pr % L)
                continue;  // (0, 0) for a label with itself
            const uint64_t P = npos[pr], Z = nzero[pr];
            voff[pr]              = total;
            eng->tr_idxs[2 * pr]     = (uint32_t)total;
            eng->tr_idxs[2 * pr + 1] = (uint32_t)P;
            total += P;
            size_t need = P;
            if (P < k_pair && Z < k_pair - P)
                need = std::max<size_t>(need, W - P - Z);
            while (kmax < need)
                kmax <<= 1;
        }
        const unsigned sg    = (unsigned)std::min<size_t>(npair, 512);
        const size_t   words = ((size_t)W + 31) / 32;
        if ((rc = upload(eng, &d_voff, voff.data(), npair)) ||
            (rc = dalloc(eng, &d_vout, std::max<uint64_t>(total, 1))) ||
            (rc = dalloc(eng, &d_bits, 2 * words)) || (rc = dalloc(eng, &d_scr, (size_t)sg * kmax)))
            return rc;
        k_train_fill<<<64, 256, 0, eng->stream>>>(d_bits, 2 * words, 0u);
        k_train_pair_sort<<<sg, TRAIN_NT, 0, eng->stream>>>(
            d_med, L, W, d_npos, d_nzero, d_voff, (uint32_t)thre_n, (uint32_t)k_pair, d_scr, kmax,
            d_vout, d_bits, d_bits + words);
        LAUNCHCHK();
        eng->tr_values.resize(total);
        std::vector<uint32_t> bits(2 * words);
        if (total)
            HIPCHK(hipMemcpyAsync(eng->tr_values.data(), d_vout, total * 4, hipMemcpyDeviceToHost,
                                  eng->stream));
        HIPCHK(hipMemcpyAsync(bits.data(), d_bits, 2 * words * 4, hipMemcpyDeviceToHost,
                              eng->stream));
        HIPCHK(hipStreamSynchronize(eng->stream));
        eng->tr_thre.clear();
        eng->tr_common.clear();
        for (uint32_t g = 0; g < W; ++g)
        {
            if ((bits[g >> 5] >> (g & 31)) & 1u)
                eng->tr_thre.push_back(g);
            if ((bits[words + (g >> 5)] >> (g & 31)) & 1u)
                eng->tr_common.push_back(g);
        }
        return CSG_OK;
    };
    rc = run();
    cleanup();
    if (rc)
        return rc;
    if (sizes)
    {
        sizes[0] = eng->tr_values.size();
        sizes[1] = eng->tr_common.size();
        sizes[2] = eng->tr_thre.size();
    }
    return CSG_OK;
}

uint64_t csg_train_table(const csg_engine* eng, int which, uint32_t* out)
{
    if (!eng)
        return 0;
    const std::vector<uint32_t>* v = nullptr;
    switch (which)
    {
    case 0: v = &eng->tr_idxs; break;
    case 1: v = &eng->tr_values; break;
    case 2: v = &eng->tr_common; break;
    case 3: v = &eng->tr_thre; break;
    default: return 0;
    }
    if (out && !v->empty())
        std::memcpy(out, v->data(), v->size() * sizeof(uint32_t));
    return v->size();
}

extern "C++" {
template <class IT>
int prefetch_query_raw(csg_engine* eng, uint32_t qry_height, const float* qry_data,
                       const IT* qry_indices, const int32_t* qry_indptr)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (qry_height == 0 || !qry_indptr)
        return eng->fail(CSG_ERR_ARGUMENT, "empty query batch");
    const size_t nnz = (size_t)(qry_indptr[qry_height] - qry_indptr[0]);
    if (nnz && (!qry_data || !qry_indices))
        return eng->fail(CSG_ERR_ARGUMENT, "missing query values or columns");
    HIPCHK(hipSetDevice(eng->device));
    int rc;
    if ((rc = ensure_copy_stream(eng)))
        return rc;
    // a free slot, else the older one (whose batch is then dropped)
    csg_engine::PreSlot* ps = !eng->pre[0].valid ? &eng->pre[0]
                              : !eng->pre[1].valid ? &eng->pre[1]
                              : eng->pre[0].seq < eng->pre[1].seq ? &eng->pre[0] : &eng->pre[1];
    ps->valid = false;
    if (ps->cap < nnz || !ps->d_data)
    {
        // an upload into the old buffers may still be in flight
        HIPCHK(hipStreamSynchronize(eng->copy_stream));
        dfree(ps->d_data);
        dfree(ps->d_idx);
        ps->cap = 0;
        const size_t n = std::max<size_t>(nnz, 1);
        if ((rc = dalloc(eng, &ps->d_data, n)) || (rc = dalloc(eng, &ps->d_idx, n)))
            return rc;
        ps->cap = n;
    }
    // the copy stream is in order: this upload follows the slot's previous
    // upload, and the slot's previous batch has been coded (csg_copyin_query_raw
    // returns after its coding)
    if (nnz)
    {
        HIPCHK(hipMemcpyAsync(ps->d_data, qry_data, nnz * 4, hipMemcpyHostToDevice,
                              eng->copy_stream));
        HIPCHK(hipMemcpyAsync(ps->d_idx, qry_indices, nnz * sizeof(IT), hipMemcpyHostToDevice,
                              eng->copy_stream));
    }
    HIPCHK(hipEventRecord(ps->copied, eng->copy_stream));
    ps->idx_bytes = (int)sizeof(IT);
    ps->valid  = true;
    ps->seq    = ++eng->pre_seq;
    ps->h_data = qry_data;
    ps->h_idx  = qry_indices;
    ps->h_ptr  = qry_indptr;
    ps->rows   = qry_height;
    return CSG_OK;
}
}  // extern "C++"

extern "C++" {
template <class IT>
int copyin_query_raw(csg_engine* eng, uint32_t qry_height, uint32_t qry_width,
                     const float* qry_data, const IT* qry_indices, const int32_t* qry_indptr,
                     const int32_t* map_score, const int32_t* map_finetune)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (!eng->score_loaded || !eng->ft_loaded)
        return eng->fail(CSG_ERR_STATE, "csg_copyin_query_raw needs the score and fine-tune "
                                        "inputs (csg_copyin_score, csg_copyin) loaded first");
    if (qry_height == 0 || qry_width == 0 || !qry_indptr || !map_score || !map_finetune)
        return eng->fail(CSG_ERR_ARGUMENT, "empty query batch or missing column maps");
    if (qry_height > eng->q_cap || qry_height > eng->fq_cap)
        return eng->fail(CSG_ERR_ARGUMENT, "query batch of " + std::to_string(qry_height) +
                                               " rows exceeds the " +
                                               std::to_string(std::min(eng->q_cap, eng->fq_cap)) +
                                               " rows the loaded inputs have room for");
    HIPCHK(hipSetDevice(eng->device));
    // the row pointers and the maps, checked on the host
    const int64_t base0 = qry_indptr[0];
    for (uint32_t r = 0; r < qry_height; ++r)
        if (qry_indptr[r + 1] < qry_indptr[r])
            return eng->fail(CSG_ERR_ARGUMENT, "query row pointers decrease");
    const size_t nnz = (size_t)(qry_indptr[qry_height] - base0);
    if (nnz && (!qry_data || !qry_indices))
        return eng->fail(CSG_ERR_ARGUMENT, "missing query values or columns");
    for (uint32_t c = 0; c < qry_width; ++c)
        if (map_score[c] >= (int32_t)eng->G || map_finetune[c] >= (int32_t)eng->Gf ||
            map_score[c] < -1 || map_finetune[c] < -1)
            return eng->fail(CSG_ERR_ARGUMENT, "column map entry outside its gene set");
    int rc;
    if ((rc = ensure_copy_stream(eng)))
        return rc;
    // a batch uploaded ahead by csg_prefetch_query_raw from these arrays (the
    // oldest such) is coded from its device copy in one slice
    csg_engine::PreSlot* slot = nullptr;
    for (auto& ps : eng->pre)
        if (ps.valid && ps.h_data == (const void*)qry_data && ps.h_idx == (const void*)qry_indices &&
            ps.h_ptr == (const void*)qry_indptr && ps.rows == qry_height &&
            ps.idx_bytes == (int)sizeof(IT) &&
            (!slot || ps.seq < slot->seq))
            slot = &ps;
    const bool staged = slot != nullptr;
    if (slot)
        slot->valid = false;
    // slices of whole rows of at most code_slice entries (a row longer than
    // that is a slice of its own)
    std::vector<uint32_t> cut{ 0 };
    size_t                biggest = 0;
    if (staged)
        cut.push_back(qry_height);
    for (uint32_t r = 0; r < qry_height && !staged;)
    {
        uint32_t e = r + 1;
        while (e < qry_height &&
               (uint64_t)(qry_indptr[e + 1] - qry_indptr[r]) <= eng->code_slice)
            ++e;
        biggest = std::max<size_t>(biggest, (size_t)(qry_indptr[e] - qry_indptr[r]));
        cut.push_back(e);
        r = e;
    }
    // work buffers (grow-only; a buffer pair shares its capacity)
    if (eng->map_cap < qry_width)
    {
        dfree(eng->d_map_s);
        dfree(eng->d_map_f);
        eng->map_cap = 0;
        if ((rc = dalloc(eng, &eng->d_map_s, qry_width)) ||
            (rc = dalloc(eng, &eng->d_map_f, qry_width)))
            return rc;
        eng->map_cap = qry_width;
    }
    if (eng->code_rows_cap < qry_height)
    {
        dfree(eng->d_raw_ptr);
        dfree(eng->d_code_cnt);
        dfree(eng->d_code_over);
        eng->code_rows_cap = 0;
        if ((rc = dalloc(eng, &eng->d_raw_ptr, (size_t)qry_height + 1)) ||
            (rc = dalloc(eng, &eng->d_code_cnt, 2 * (size_t)qry_height)) ||
            (rc = dalloc(eng, &eng->d_code_over, 2 * (size_t)qry_height)))
            return rc;
        eng->code_rows_cap = qry_height;
    }
    if (eng->raw_cap < biggest || !eng->d_raw_data[0])
    {
        for (int b = 0; b < 2; ++b)
        {
            dfree(eng->d_raw_data[b]);
            dfree(eng->d_raw_idx[b]);
        }
        eng->raw_cap = 0;
        const size_t n = std::max<size_t>(biggest, 1);
        for (int b = 0; b < 2; ++b)
            if ((rc = dalloc(eng, &eng->d_raw_data[b], n)) ||
                (rc = dalloc(eng, &eng->d_raw_idx[b], n)))
                return rc;
        eng->raw_cap = n;
    }
    // coded outputs: at most every entry kept; row pointers for the largest
    // batch the inputs have room for
    if (eng->q_nnz_cap < nnz || !eng->d_qdata || !eng->d_qptr)
    {
        dfree(eng->d_qdata);
        dfree(eng->d_qind);
        dfree(eng->d_qptr);
        eng->q_nnz_cap = 0;
        const size_t n = std::max<size_t>(nnz, 1);
        if ((rc = dalloc(eng, &eng->d_qdata, n)) || (rc = dalloc(eng, &eng->d_qind, n)) ||
            (rc = dalloc(eng, &eng->d_qptr, (size_t)eng->q_cap + 1)))
            return rc;
        eng->q_nnz_cap = n;
    }
    if (eng->fq_nnz_cap < nnz || !eng->d_fqdata || !eng->d_fqptr)
    {
        dfree(eng->d_fqdata);
        dfree(eng->d_fqind);
        dfree(eng->d_fqptr);
        eng->fq_nnz_cap = 0;
        const size_t n = std::max<size_t>(nnz, 1);
        if ((rc = dalloc(eng, &eng->d_fqdata, n)) || (rc = dalloc(eng, &eng->d_fqind, n)) ||
            (rc = dalloc(eng, &eng->d_fqptr, (size_t)eng->fq_cap + 1)))
            return rc;
        eng->fq_nnz_cap = n;
    }
    HIPCHK(hipMemcpyAsync(eng->d_map_s, map_score, (size_t)qry_width * 4, hipMemcpyHostToDevice,
                          eng->stream));
    HIPCHK(hipMemcpyAsync(eng->d_map_f, map_finetune, (size_t)qry_width * 4,
                          hipMemcpyHostToDevice, eng->stream));
    HIPCHK(hipMemcpyAsync(eng->d_raw_ptr, qry_indptr, ((size_t)qry_height + 1) * 4,
                          hipMemcpyHostToDevice, eng->stream));
    k_code_init<<<1, 64, 0, eng->stream>>>((uint32_t*)eng->d_code_misc);
    uint32_t* cnt_s   = eng->d_code_cnt;
    uint32_t* cnt_f   = eng->d_code_cnt + qry_height;
    // the copy stream waits for the maps and pointers, then the slices
    // alternate between two buffers: slice i's upload waits for slice i - 2's
    // coding, its coding for its upload
    hipEvent_t start = take_event(eng);
    HIPCHK(hipEventRecord(start, eng->stream));
    HIPCHK(hipStreamWaitEvent(eng->copy_stream, start, 0));
    eng->event_pool.push_back(start);
    {
        PhaseScope ps(eng, 10, (int)(cut.size() - 1) * 5);
        for (size_t i = 0; i + 1 < cut.size(); ++i)
        {
            const int      b    = (int)(i & 1);
            const uint32_t r0   = cut[i], r1 = cut[i + 1], nr = r1 - r0;
            const int64_t  e0   = qry_indptr[r0];
            const size_t   n    = (size_t)(qry_indptr[r1] - e0);
            const float* src_data = eng->d_raw_data[b];
            const IT*    src_idx  = (const IT*)eng->d_raw_idx[b];
            if (staged)
            {
                HIPCHK(hipStreamWaitEvent(eng->stream, slot->copied, 0));
                src_data = slot->d_data;
                src_idx  = (const IT*)slot->d_idx;
            }
            else
            {
                if (i >= 2)
                    HIPCHK(hipStreamWaitEvent(eng->copy_stream, eng->code_done[b], 0));
                if (n)
                {
                    HIPCHK(hipMemcpyAsync(eng->d_raw_data[b], qry_data + (e0 - base0), n * 4,
                                          hipMemcpyHostToDevice, eng->copy_stream));
                    HIPCHK(hipMemcpyAsync(eng->d_raw_idx[b], qry_indices + (e0 - base0),
                                          n * sizeof(IT), hipMemcpyHostToDevice,
                                          eng->copy_stream));
                }
                HIPCHK(hipEventRecord(eng->code_copied[b], eng->copy_stream));
                HIPCHK(hipStreamWaitEvent(eng->stream, eng->code_copied[b], 0));
            }
            const CodeSet ss{ eng->d_map_s, eng->d_qptr + r0, eng->d_qdata, eng->d_qind, cnt_s + r0 };
            const CodeSet sf{ eng->d_map_f, eng->d_fqptr + r0, eng->d_fqdata, eng->d_fqind,
                              cnt_f + r0 };
            launch_coding(eng, eng->stream, src_data, src_idx, eng->d_raw_ptr + r0,
                          staged ? base0 : e0, nr, qry_width, ss, sf, eng->d_code_over,
                          eng->d_code_misc);
            LAUNCHCHK();
            HIPCHK(hipEventRecord(eng->code_done[b], eng->stream));
        }
    }
    HIPCHK(hipMemcpyAsync(eng->h_code_misc, eng->d_code_misc, 32, hipMemcpyDeviceToHost,
                          eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    collect_phases(eng);
    uint32_t misc[4];
    std::memcpy(misc, eng->h_code_misc, 16);
    // the query is replaced whatever follows: a failed batch leaves none
    eng->q_ready = eng->fq_ready = false;
    if ((rc = coding_status(eng, misc)))
        return rc;
    const int nb = std::max(eng->nbins_ref, (int)misc[2]), fnb = std::max(eng->fnbins_ref, (int)misc[3]);
    if (nb > 16384 || fnb > 16384 || rank_smem(nb, eng->Kp) > 65536)
        return eng->fail(CSG_ERR_ARGUMENT, "a query row holds more distinct values than the rank "
                                           "kernels' code bins");
    eng->Nq     = qry_height;
    eng->fNq    = qry_height;
    eng->nbins  = nb;
    eng->fnbins = fnb;
    if (fnb > 256 && eng->d_fref_codes8)
        dfree(eng->d_fref_codes8);  // byte-wide fine-tune codes no longer hold every code
    eng->mask_valid = false;
    eng->host_mask.clear();
    eng->q_ready = eng->fq_ready = true;
    return CSG_OK;
}
}  // extern "C++"

int csg_prefetch_query_raw(csg_engine* eng, uint32_t qry_height, const float* qry_data,
                           const int32_t* qry_indices, const int32_t* qry_indptr)
{
    return prefetch_query_raw(eng, qry_height, qry_data, qry_indices, qry_indptr);
}

int csg_prefetch_query_raw_u16(csg_engine* eng, uint32_t qry_height, const float* qry_data,
                               const uint16_t* qry_indices, const int32_t* qry_indptr)
{
    return prefetch_query_raw(eng, qry_height, qry_data, qry_indices, qry_indptr);
}

int csg_copyin_query_raw(csg_engine* eng, uint32_t qry_height, uint32_t qry_width,
                         const float* qry_data, const int32_t* qry_indices,
                         const int32_t* qry_indptr, const int32_t* map_score,
                         const int32_t* map_finetune)
{
    return copyin_query_raw(eng, qry_height, qry_width, qry_data, qry_indices, qry_indptr,
                            map_score, map_finetune);
}

int csg_copyin_query_raw_u16(csg_engine* eng, uint32_t qry_height, uint32_t qry_width,
                             const float* qry_data, const uint16_t* qry_indices,
                             const int32_t* qry_indptr, const int32_t* map_score,
                             const int32_t* map_finetune)
{
    if (eng && qry_width > 65536)
        return eng->fail(CSG_ERR_ARGUMENT, "16-bit column indices need a query width of at "
                                           "most 65536");
    return copyin_query_raw(eng, qry_height, qry_width, qry_data, qry_indices, qry_indptr,
                            map_score, map_finetune);
}

int csg_code_csr(csg_engine* eng, uint32_t rows, uint32_t width, const float* data,
                 const int32_t* indices, const int32_t* indptr, const int32_t* map_a,
                 const int32_t* map_b, uint16_t* data_a, int32_t* indices_a, int32_t* indptr_a,
                 uint16_t* data_b, int32_t* indices_b, int32_t* indptr_b, uint32_t* max_uniq)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (rows == 0 || width == 0 || !data || !indices || !indptr || !map_a || !map_b || !data_a ||
        !indices_a || !indptr_a || !data_b || !indices_b || !indptr_b)
        return eng->fail(CSG_ERR_ARGUMENT, "empty CSR or missing device pointer");
    HIPCHK(hipSetDevice(eng->device));
    int rc;
    if (eng->code_rows_cap < rows || !eng->d_code_cnt)
    {
        dfree(eng->d_raw_ptr);
        dfree(eng->d_code_cnt);
        dfree(eng->d_code_over);
        eng->code_rows_cap = 0;
        if ((rc = dalloc(eng, &eng->d_raw_ptr, (size_t)rows + 1)) ||
            (rc = dalloc(eng, &eng->d_code_cnt, 2 * (size_t)rows)) ||
            (rc = dalloc(eng, &eng->d_code_over, 2 * (size_t)rows)))
            return rc;
        eng->code_rows_cap = rows;
    }
    if ((rc = ensure_copy_stream(eng)))
        return rc;
    const CodeSet sa{ map_a, indptr_a, data_a, indices_a, eng->d_code_cnt };
    const CodeSet sb{ map_b, indptr_b, data_b, indices_b, eng->d_code_cnt + rows };
    k_code_init<<<1, 64, 0, eng->stream>>>((uint32_t*)eng->d_code_misc);
    launch_coding(eng, eng->stream, data, indices, indptr, 0, rows, width, sa, sb,
                  eng->d_code_over, eng->d_code_misc);
    LAUNCHCHK();
    uint32_t misc[4];
    HIPCHK(hipMemcpyAsync(misc, eng->d_code_misc, 16, hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    if ((rc = coding_status(eng, misc)))
        return rc;
    if (max_uniq)
    {
        max_uniq[0] = misc[2];
        max_uniq[1] = misc[3];
    }
    return CSG_OK;
}

int csg_code_csr_noalloc(csg_engine* eng, uint32_t rows, uint32_t width, const float* data,
                         const int32_t* indices, const int32_t* indptr, const int32_t* map_a,
                         const int32_t* map_b, uint16_t* data_a, int32_t* indices_a,
                         int32_t* indptr_a, uint16_t* data_b, int32_t* indices_b,
                         int32_t* indptr_b, uint32_t* status)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (rows == 0 || width == 0 || !data || !indices || !indptr || !map_a || !map_b || !data_a ||
        !indices_a || !indptr_a || !data_b || !indices_b || !indptr_b || !status)
        return eng->fail(CSG_ERR_ARGUMENT, "empty CSR or missing device pointer");
    if (rows > (uint32_t)CSG_CODE_NOALLOC_MAX_ROWS)
        return eng->fail(CSG_ERR_ARGUMENT,
                         "csg_code_csr_noalloc codes at most " +
                             std::to_string(CSG_CODE_NOALLOC_MAX_ROWS) + " rows per call (asked " +
                             std::to_string(rows) + "); use csg_code_csr or split the rows");
    HIPCHK(hipSetDevice(eng->device));
    // the work memory's device address (never the array's name in host code)
    uint8_t*      misc  = eng->noalloc_scratch;
    uint32_t*     cnt   = (uint32_t*)(misc + NOALLOC_MISC);
    uint32_t*     over  = cnt + 2 * (size_t)CSG_CODE_NOALLOC_MAX_ROWS;
    const CodeSet sa{ map_a, indptr_a, data_a, indices_a, cnt };
    const CodeSet sb{ map_b, indptr_b, data_b, indices_b, cnt + rows };
    k_code_init<<<1, 64, 0, eng->stream>>>((uint32_t*)misc);
    launch_coding(eng, eng->stream, data, indices, indptr, 0, rows, width, sa, sb, over, misc);
    k_code_status<<<1, 64, 0, eng->stream>>>((const uint32_t*)misc, status);
    LAUNCHCHK();
    return CSG_OK;
}

int csg_synchronize(csg_engine* eng)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    HIPCHK(hipSetDevice(eng->device));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return CSG_OK;
}

int csg_query_coded(csg_engine* eng, int step, uint16_t* data, int32_t* indices, int32_t* indptr)
{
    if (!eng || (step != 0 && step != 1) || !indptr)
        return CSG_ERR_ARGUMENT;
    const bool ok = step == 0 ? eng->score_loaded && eng->q_ready : eng->ft_loaded && eng->fq_ready;
    if (!ok)
        return eng->fail(CSG_ERR_STATE, "no query data of that step on the device");
    HIPCHK(hipSetDevice(eng->device));
    const uint32_t n   = step == 0 ? eng->Nq : eng->fNq;
    const int32_t* ptr = step == 0 ? eng->d_qptr : eng->d_fqptr;
    HIPCHK(hipMemcpy(indptr, ptr, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
    const size_t nnz = (size_t)(indptr[n] - indptr[0]);
    if (data && nnz)
        HIPCHK(hipMemcpy(data, (step == 0 ? eng->d_qdata : eng->d_fqdata) + indptr[0], nnz * 2,
                         hipMemcpyDeviceToHost));
    if (indices && nnz)
        HIPCHK(hipMemcpy(indices, (step == 0 ? eng->d_qind : eng->d_fqind) + indptr[0], nnz * 4,
                         hipMemcpyDeviceToHost));
    return CSG_OK;
}

int csg_host_alloc(uint64_t bytes, void** ptr)
{
    if (!ptr)
        return CSG_ERR_ARGUMENT;
    *ptr = nullptr;
    if (hipHostMalloc(ptr, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess)
    {
        (void)hipGetLastError();
        *ptr = nullptr;
        return CSG_ERR_MEMORY;
    }
    return CSG_OK;
}

int csg_host_free(void* ptr)
{
    if (!ptr)
        return CSG_OK;
    return hipHostFree(ptr) == hipSuccess ? CSG_OK : CSG_ERR_DEVICE;
}

int csg_set_upload_slice(csg_engine* eng, uint64_t entries)
{
    if (!eng || entries == 0)
        return CSG_ERR_ARGUMENT;
    eng->code_slice = entries;
    return CSG_OK;
}

int csg_set_query_chunk(csg_engine* eng, uint32_t rows, uint32_t* chunk)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (rows % 128)
        return eng->fail(CSG_ERR_ARGUMENT, "the query chunk cap must be a multiple of 128 rows");
    eng->qc_cap = rows;
    if (chunk)
        *chunk = eng->Qc;
    return CSG_OK;
}

int csg_set_finetune_fuse(csg_engine* eng, int max_cells)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (max_cells < 0 || max_cells > FTD_FUSE_MAX)
        return eng->fail(CSG_ERR_ARGUMENT, "fused fine-tune groups hold 0 to 4 cells");
    eng->ftd_fuse_max = max_cells;
    return CSG_OK;
}

int csg_finetune_tail_profile(csg_engine* eng, double* us, int max, int* count)
{
    if (!eng || !us || !count || max < 0)
        return CSG_ERR_ARGUMENT;
    *count = -eng->tt_refused;
    if (!eng->h_tt_state || !eng->tt_ran)
        return CSG_OK;
    const FtdTailState& t = *eng->h_tt_state;
    // barriers passed: the first, then two per round with cells and the
    // last one (the round that found no cell)
    const int nb = std::min<int>(FTT_STAMPS, 1 + 2 * ((int)t.rounds + 1));
    int       n  = 0;
    for (int g = 1; g < nb && n < max; ++g, ++n)
        us[n] = (double)(t.stamp[g] - t.stamp[g - 1]) / 100.0;  // 100 MHz clock
    *count = n;
    return CSG_OK;
}

int csg_set_finetune_tail(csg_engine* eng, uint32_t max_cells, int blocks_per_cu)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (blocks_per_cu < 1 || blocks_per_cu > 16)
        return eng->fail(CSG_ERR_ARGUMENT, "the tail kernel takes 1 to 16 blocks per CU");
    eng->ftd_tail_max     = max_cells;
    eng->tt_blocks_per_cu = blocks_per_cu;
    return CSG_OK;
}

int csg_set_limb_form(csg_engine* eng, int form, int* prev)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (form < 0 || form > 1)
        return eng->fail(CSG_ERR_ARGUMENT, "limb form 0 (automatic) or 1 (folded)");
    if (prev)
        *prev = eng->limb_form;
    eng->limb_form = form;
    return CSG_OK;
}

int csg_set_finetune_chain(csg_engine* eng, int max_rounds, uint64_t max_scores)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (max_rounds < 0 || max_rounds > 64)
        return eng->fail(CSG_ERR_ARGUMENT, "fine-tune chains hold 0 to 64 rounds");
    eng->ftd_chain_max    = max_rounds;
    eng->ftd_chain_scores = max_scores;
    return CSG_OK;
}

int csg_finetune_tuning(csg_engine* eng, int autotune, int32_t* grids, float* best_ms)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (autotune >= 0)
    {
        eng->ftd_autotune = autotune != 0;
        eng->ftd_tuned    = false;
    }
    if (grids)
    {
        grids[0] = eng->ftd_tuned ? 1 : 0;
        grids[1] = eng->ftd_grid_rank;
        grids[2] = eng->ftd_grid_corr;
        grids[3] = eng->ftd_grid_select;
        grids[4] = eng->ftd_t64 == (int)0xFFFFFFFF ? -1 : eng->ftd_t64;
        grids[5] = eng->ftd_grid_fused;
    }
    if (best_ms)
        for (int i = 0; i < 4; ++i)
            best_ms[i] = eng->ftd_tune_ms[i];
    return CSG_OK;
}

int csg_set_finetune_path(csg_engine* eng, int path)
{
    if (!eng || path < CSG_PATH_PLAIN || path > CSG_PATH_FP16_MFMA)
        return CSG_ERR_ARGUMENT;
    eng->ft_exact = path == CSG_PATH_EXACT_INT8;
    return CSG_OK;
}

int csg_set_phase_timing(csg_engine* eng, int mode)
{
    if (!eng || mode < 0 || mode > 3)
        return CSG_ERR_ARGUMENT;
    eng->phase_mode = mode;
    return CSG_OK;
}

int csg_phase_times(csg_engine* eng, double* ms, int64_t* counts, int reset)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    (void)hipStreamSynchronize(eng->stream);
    int rc = fold_score_step(eng);
    collect_phases(eng);
    for (int i = 0; i < CSG_N_PHASES; ++i)
    {
        if (ms)
            ms[i] = eng->phase_ms[i];
        if (counts)
            counts[i] = eng->phase_cnt[i];
        if (reset)
        {
            eng->phase_ms[i]  = 0;
            eng->phase_cnt[i] = 0;
        }
    }
    return rc;
}

int csg_counters(csg_engine* eng, int64_t* out)
{
    if (!eng || !out)
        return CSG_ERR_ARGUMENT;
    const int rc = fold_score_step(eng);
    std::memcpy(out, eng->counters, sizeof(eng->counters));
    return rc;
}

// The top-k kernels: k_topk_rows (one thread per row, L <= 64) against
// k_topk (one wave per row) on seeded logits with ties, -inf entries and NaN
// rows, for L = 1, 2, 17, 50, 64 and K = 1, 4, 8; returns the number of
// differing outputs (labels, or margins by their bits).
int selftest_topk(csg_engine* eng, double* mismatches)
{
    const int Lp = 64, rows = 1000;
    std::vector<float> acc((size_t)rows * Lp), bias(Lp);
    uint64_t           st = 0x2545F4914F6CDD1Dull;
    auto               rnd = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    for (int j = 0; j < Lp; ++j)
        bias[j] = (float)((int)(rnd() % 9) - 4) * 0.125f;
    for (int r = 0; r < rows; ++r)
        for (int j = 0; j < Lp; ++j)
        {
            const uint64_t u = rnd();
            // coarse values (many ties), some -inf, a NaN row every 97 rows
            float v = (float)((int)(u % 33) - 16) * 0.25f;
            if (u % 61 == 0)
                v = -__builtin_inff();
            if (r % 97 == 5 && j == (int)(u % 50))
                v = __builtin_nanf("");
            acc[(size_t)r * Lp + j] = v;
        }
    float *d_acc = nullptr, *d_bias = nullptr, *d_m = nullptr;
    int32_t* d_t = nullptr;
    int      rc  = CSG_OK;
    if ((rc = upload(eng, &d_acc, acc.data(), acc.size())) ||
        (rc = upload(eng, &d_bias, bias.data(), bias.size())) ||
        (rc = dalloc(eng, &d_t, (size_t)2 * rows * 8)) || (rc = dalloc(eng, &d_m, (size_t)2 * rows)))
    {
        dfree(d_acc); dfree(d_bias); dfree(d_t); dfree(d_m);
        return rc;
    }
    double bad = 0;
    for (int L : { 1, 2, 17, 50, 64 })
        for (int K : { 1, 4, 8 })
        {
            if (K > L)
                continue;
            int32_t* t1 = d_t + (size_t)rows * 8;
            float*   m1 = d_m + rows;
#define CSG_TOPK_PAIR(k)                                                                           \
    case k:                                                                                        \
        k_topk<k><<<(rows + 3) / 4, 256, 0, eng->stream>>>(d_acc, d_bias, rows, L, Lp, d_t, d_m);  \
        k_topk_rows<k><<<(rows + 127) / 128, 128, 0, eng->stream>>>(d_acc, d_bias, rows, L, Lp, t1, \
                                                                    m1);                           \
        break;
            switch (K)
            {
                CSG_TOPK_PAIR(1)
                CSG_TOPK_PAIR(4)
                CSG_TOPK_PAIR(8)
            }
#undef CSG_TOPK_PAIR
            std::vector<int32_t>  ta((size_t)rows * K), tb((size_t)rows * K);
            std::vector<uint32_t> ma(rows), mb(rows);
            hipError_t e0 = hipGetLastError();
            hipError_t e1 = hipMemcpyAsync(ta.data(), d_t, ta.size() * 4, hipMemcpyDeviceToHost, eng->stream);
            hipError_t e2 = hipMemcpyAsync(tb.data(), t1, tb.size() * 4, hipMemcpyDeviceToHost, eng->stream);
            hipError_t e3 = hipMemcpyAsync(ma.data(), d_m, rows * 4, hipMemcpyDeviceToHost, eng->stream);
            hipError_t e4 = hipMemcpyAsync(mb.data(), m1, rows * 4, hipMemcpyDeviceToHost, eng->stream);
            hipError_t e5 = hipStreamSynchronize(eng->stream);
            if (e0 != hipSuccess || e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess ||
                e4 != hipSuccess || e5 != hipSuccess)
            {
                dfree(d_acc); dfree(d_bias); dfree(d_t); dfree(d_m);
                return eng->fail(CSG_ERR_DEVICE, "top-k selftest failed");
            }
            for (size_t i = 0; i < ta.size(); ++i)
                bad += ta[i] != tb[i];
            for (int r = 0; r < rows; ++r)
            {
                const bool na = (ma[r] & 0x7F800000u) == 0x7F800000u && (ma[r] & 0x7FFFFFu);
                const bool nb = (mb[r] & 0x7F800000u) == 0x7F800000u && (mb[r] & 0x7FFFFFu);
                bad += na != nb || (!na && ma[r] != mb[r]);
            }
        }
    dfree(d_acc); dfree(d_bias); dfree(d_t); dfree(d_m);
    *mismatches = bad;
    return CSG_OK;
}

int csg_selftest(csg_engine* eng, int which, double* max_err)
{
    if (!eng || !max_err || which < 0 || which > 7)
        return CSG_ERR_ARGUMENT;
    if (which == 7)
    {
        HIPCHK(hipSetDevice(eng->device));
        return selftest_topk(eng, max_err);
    }
    if (which == 4 && !eng->w8_ok)
        return eng->fail(CSG_ERR_DEVICE, "the device does not grant the wide kernel's LDS");
    if (which == 6 && !eng->kf_ok)
        return eng->fail(CSG_ERR_DEVICE,
                         "the device does not grant the three-product wide kernel's LDS");
    if (which == 5 && !eng->x3_ok)
        return eng->fail(CSG_ERR_DEVICE, "the device does not grant the three-partial kernel's LDS");
    HIPCHK(hipSetDevice(eng->device));
    // the three-partial kernel over the whole int8 range of the high limb
    // the three-product form over its whole range, |X| <= 1007
    const int nq = 256, nr = 256, Kp = 192, G = which == 5 ? 32000 : which == 6 ? 1007 : 1500;
    std::vector<int16_t> xq((size_t)nq * Kp), xr((size_t)nr * Kp);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return (int)(st % (2 * G + 1)) - G;
    };
    for (auto& v : xq)
        v = (int16_t)rnd();
    for (auto& v : xr)
        v = (int16_t)rnd();
    std::vector<float>     ones(std::max(nq, nr), 1.f);
    std::vector<double>    ones64(std::max(nq, nr), 1.0);
    int16_t *dq = nullptr, *dr = nullptr;
    int8_t *qh = nullptr, *ql = nullptr, *rh = nullptr, *rl = nullptr;
    _Float16 *qf = nullptr, *rf = nullptr;
    float *d1 = nullptr, *dS = nullptr;
    double *d1i = nullptr, *cq = nullptr, *cr = nullptr;
    int    rc = CSG_OK;
    auto   cleanup = [&]() {
        dfree(dq); dfree(dr); dfree(qh); dfree(ql); dfree(rh); dfree(rl); dfree(qf); dfree(rf);
        dfree(d1); dfree(dS); dfree(d1i); dfree(cq); dfree(cr);
    };
    if ((rc = upload(eng, &dq, xq.data(), xq.size())) || (rc = upload(eng, &dr, xr.data(), xr.size())) ||
        (rc = upload(eng, &d1, ones.data(), ones.size())) ||
        (rc = upload(eng, &d1i, ones64.data(), ones64.size())) || (rc = dalloc(eng, &dS, (size_t)nq * nr)) ||
        (rc = dalloc(eng, &qh, xq.size())) || (rc = dalloc(eng, &ql, xq.size())) ||
        (rc = dalloc(eng, &rh, xr.size())) || (rc = dalloc(eng, &rl, xr.size())) ||
        (rc = dalloc(eng, &qf, xq.size())) || (rc = dalloc(eng, &rf, xr.size())) ||
        (rc = dalloc(eng, &cq, (size_t)nq)) || (rc = dalloc(eng, &cr, (size_t)nr)))
    {
        cleanup();
        return rc;
    }
    const unsigned bq = (unsigned)((xq.size() / 8 + 255) / 256), br = (unsigned)((xr.size() / 8 + 255) / 256);
    if (which == 0)
    {
        k_split_limbs<<<bq, 256, 0, eng->stream>>>(dq, xq.size(), qh, ql);
        k_split_limbs<<<br, 256, 0, eng->stream>>>(dr, xr.size(), rh, rl);
        k_gemm_int8<<<dim3(nr / I8_BN, nq / I8_BM), 256, 0, eng->stream>>>(qh, ql, rh, rl, Kp, d1i,
                                                                          d1i, nq, nr, dS, nr);
    }
    else if (which == 1)
    {
        k_to_half<<<bq, 256, 0, eng->stream>>>(dq, xq.size(), qf);
        k_to_half<<<br, 256, 0, eng->stream>>>(dr, xr.size(), rf);
        k_gemm_f16<<<dim3(nr / H_BN, nq / H_BM), 256, 0, eng->stream>>>(qf, rf, Kp, d1, d1, nq, nr,
                                                                        dS, nr);
    }
    else if (which == 4)
    {
        k_split_limbs<<<bq, 256, 0, eng->stream>>>(dq, xq.size(), qh, ql);
        k_split_limbs<<<br, 256, 0, eng->stream>>>(dr, xr.size(), rh, rl);
        k_gemm_int8w<false><<<dim3(nr / W8_BN, nq / W8_BM), 512, W8_LDS, eng->stream>>>(
            qh, ql, rh, rl, Kp, d1i, d1i, nq, nr, dS, nr, nullptr, nullptr);
    }
    else if (which == 6)
    {
        k_split_kf<<<(unsigned)(nq / 4), 256, 0, eng->stream>>>(dq, (size_t)nq, Kp, qh, ql, cq);
        k_split_kf<<<(unsigned)(nr / 4), 256, 0, eng->stream>>>(dr, (size_t)nr, Kp, rh, rl, cr);
        k_gemm_int8w<true><<<dim3(nr / W8_BN, nq / W8_BM), 512, W8_LDS_KF, eng->stream>>>(
            qh, ql, rh, rl, Kp, d1i, d1i, nq, nr, dS, nr, cq, cr);
    }
    else if (which == 5)
    {
        k_split_limbs<<<bq, 256, 0, eng->stream>>>(dq, xq.size(), qh, ql);
        k_split_limbs<<<br, 256, 0, eng->stream>>>(dr, xr.size(), rh, rl);
        k_gemm_int8x<<<dim3(nr / X3_BN, nq / X3_BM), 512, X3_LDS, eng->stream>>>(
            qh, ql, rh, rl, Kp, d1i, d1i, nq, nr, dS, nr);
    }
    else if (which == 2)
        k_gemm_f32mfma<<<dim3(nr / F_BN, nq / F_BM), 256, 0, eng->stream>>>(dq, dr, Kp, d1, d1, nq,
                                                                            nr, dS, nr);
    else
        k_gemm_plain<<<dim3(nr / PG_BN, nq / PG_BM), 256, 0, eng->stream>>>(dq, dr, Kp, d1, d1, nq,
                                                                            nr, dS, nr);
    hipError_t le = hipGetLastError();
    std::vector<float> out((size_t)nq * nr);
    hipError_t ce = hipMemcpyAsync(out.data(), dS, out.size() * 4, hipMemcpyDeviceToHost, eng->stream);
    hipError_t se = hipStreamSynchronize(eng->stream);
    cleanup();
    if (le != hipSuccess || ce != hipSuccess || se != hipSuccess)
        return eng->fail(CSG_ERR_DEVICE, "selftest kernel failed");
    double worst = 0;
    for (int m = 0; m < nq; ++m)
        for (int n = 0; n < nr; ++n)
        {
            long long s4  = 0, sxx = 0, syy = 0;
            float     seq = 0.f;
            for (int k = 0; k < Kp; ++k)
            {
                const long long a = xq[(size_t)m * Kp + k], b = xr[(size_t)n * Kp + k];
                s4 += a * b;
                sxx += a * a;
                syy += b * b;
                seq = std::fma((float)a, (float)b, seq);
            }
            const float  dev = out[(size_t)m * nr + n];
            double       err;
            if (which == 0 || which == 4 || which == 5 || which == 6)
                err = dev == (float)(double)s4 ? 0.0 : std::fabs((double)dev - (double)s4) + 1.0;
            else if (which == 3)
                err = dev == seq ? 0.0 : std::fabs((double)dev - (double)seq) + 1.0;
            else
                err = std::fabs((double)dev - (double)s4) / std::sqrt((double)sxx * (double)syy);
            worst = std::max(worst, err);
        }
    *max_err = worst;
    return CSG_OK;
}

int csg_selftest_ft(csg_engine* eng, int gs, int64_t* mismatches)
{
    if (!eng || !mismatches || gs < 1 || gs > 16384)
        return CSG_ERR_ARGUMENT;
    HIPCHK(hipSetDevice(eng->device));
    const int      ncell = 40, nref = 70;
    const uint32_t st    = (uint32_t)round_up(gs, 8);
    std::vector<int16_t> cr((size_t)ncell * st, 0), rr((size_t)nref * st, 0);
    uint64_t             sd = 0x2545F4914F6CDD1Dull ^ (uint64_t)gs;
    auto rnd = [&]() {
        sd ^= sd << 13;
        sd ^= sd >> 7;
        sd ^= sd << 17;
        return (int)(sd % (2 * gs + 1)) - gs;
    };
    for (int c = 0; c < ncell; ++c)
        for (int i = 0; i < gs; ++i)
            cr[(size_t)c * st + i] = (int16_t)rnd();
    for (int r = 0; r < nref; ++r)
        for (int i = 0; i < gs; ++i)
            rr[(size_t)r * st + i] = (int16_t)rnd();
    FtGroup g{};
    g.gs = (uint32_t)gs;
    g.stride = st;
    g.nref   = nref;
    g.ncell  = ncell;
    std::vector<FtCorrTile> tiles;
    for (uint32_t c0 = 0; c0 < (uint32_t)ncell; c0 += FT_TC)
        for (uint32_t r0 = 0; r0 < (uint32_t)nref; r0 += FT_TR)
            tiles.push_back(FtCorrTile{ 0, c0, r0, 0 });
    std::vector<float> ones(std::max(ncell, nref), 1.f);
    FtGroup*    dg = nullptr;
    FtCorrTile* dt = nullptr;
    int16_t *   dc = nullptr, *dr = nullptr;
    float *     d1 = nullptr, *ds = nullptr;
    int         rc = CSG_OK;
    auto        cleanup = [&]() {
        dfree(dg); dfree(dt); dfree(dc); dfree(dr); dfree(d1); dfree(ds);
    };
    if ((rc = upload(eng, &dg, &g, 1)) || (rc = upload(eng, &dt, tiles.data(), tiles.size())) ||
        (rc = upload(eng, &dc, cr.data(), cr.size())) || (rc = upload(eng, &dr, rr.data(), rr.size())) ||
        (rc = upload(eng, &d1, ones.data(), ones.size())) ||
        (rc = dalloc(eng, &ds, (size_t)ncell * nref)))
    {
        cleanup();
        return rc;
    }
    // both tile kernels, each with its widest LDS stage and with 64-gene
    // stages (several stages per row beyond 64 genes)
    std::vector<FtCorrTile> thin;
    for (uint32_t c0 = 0; c0 < (uint32_t)ncell; c0 += FT_THIN)
        for (uint32_t r0 = 0; r0 < (uint32_t)nref; r0 += 256)
            thin.push_back(FtCorrTile{ 0, c0, r0, 0 });
    FtCorrTile* dth = nullptr;
    if ((rc = upload(eng, &dth, thin.data(), thin.size())))
    {
        cleanup();
        return rc;
    }
    std::vector<std::vector<float>> outs;
    hipError_t                      err = hipSuccess;
    for (int v = 0; v < 4 && err == hipSuccess; ++v)
    {
        const bool wide = v < 2;
        const int  kt   = wide ? ft_corr_stage((int)st, eng->lds_limit, FT_TC + FT_TR) : 64;
        const int  ktt  = wide ? std::min<int>((int)st, FT_THIN_KT) : 64;
        if (v % 2 == 0)
            k_ft_corr<FT_TC, FT_TR, FT_RC, FT_RR><<<(unsigned)tiles.size(), FT_NT,
                                                    (size_t)(FT_TC + FT_TR) * (kt / 2 + 1) * 4,
                                                    eng->stream>>>(dt, dg, dc, d1, dr, d1, ds, kt);
        else
            k_ft_corr_thin<FT_THIN><<<(unsigned)thin.size(), 256, (size_t)FT_THIN * ktt * 2,
                                      eng->stream>>>(dth, dg, dc, d1, dr, d1, ds, ktt);
        err = hipGetLastError();
        outs.emplace_back((size_t)ncell * nref);
        if (err == hipSuccess)
            err = hipMemcpyAsync(outs.back().data(), ds, outs.back().size() * 4,
                                 hipMemcpyDeviceToHost, eng->stream);
        if (err == hipSuccess)
            err = hipStreamSynchronize(eng->stream);
    }
    dfree(dth);
    cleanup();
    if (err != hipSuccess)
        return eng->fail(CSG_ERR_DEVICE, std::string("ft selftest kernel failed: ") +
                                             hipGetErrorString(err));
    // host: 128 chunks of ceil(gs/128) genes summed by fma, stride 1..64 tree
    int64_t   bad  = 0;
    const int step = (gs + 127) / 128;
    for (int c = 0; c < ncell; ++c)
        for (int r = 0; r < nref; ++r)
        {
            float part[128];
            for (int t = 0; t < 128; ++t)
            {
                float acc = 0.f;
                for (int i = t * step; i < std::min((t + 1) * step, gs); ++i)
                    acc = std::fma((float)cr[(size_t)c * st + i], (float)rr[(size_t)r * st + i], acc);
                part[t] = acc;
            }
            for (int stride = 1; stride < 128; stride *= 2)
                for (int t = 0; t < 128; t += 2 * stride)
                    part[t] = part[t] + part[t + stride];
            for (auto& o : outs)
                if (!(o[(size_t)c * nref + r] == part[0]))
                    ++bad;
        }
    *mismatches = bad;
    return CSG_OK;
}

int csg_host_percentile(const float* arr, int64_t len, float quantile, float* out)
{
    if (len <= 0 || !arr || !out)
        return CSG_ERR_ARGUMENT;
    const PctParam     pp = make_pct_param((long)len, quantile);
    std::vector<float> a(arr, arr + len);
    if (pp.kind == 0)
    {
        *out = a[0];
        return CSG_OK;
    }
    if (!std_sort(a.data(), (long)len))
        return CSG_ERR_UNDEFINED;
    if (pp.kind == 1)
        *out = a[0];
    else if (pp.kind == 2)
        *out = a[len - 1];
    else
        *out = pct_interp(quantile, pp.prev_index, pp.step, a[pp.k - 1], a[pp.k]);
    return CSG_OK;
}

void csg_host_minmax(const float* arr, int n, int* imin, int* imax) { std_minmax(arr, n, imin, imax); }


}  // extern "C"

// ================================================================ learned proposal

namespace {

void free_mlp(csg_engine* eng)
{
    dfree(eng->d_mw);
    dfree(eng->d_mg);
    dfree(eng->d_mm);
    dfree(eng->d_mv);
    dfree(eng->d_mwb);
    dfree(eng->d_mwbT);
    dfree(eng->d_mx);
    dfree(eng->d_macc);
    dfree(eng->d_mh1);
    dfree(eng->d_mh2);
    dfree(eng->d_mz1);
    dfree(eng->d_mz2);
    dfree(eng->d_mdl);
    dfree(eng->d_mdlb);
    dfree(eng->d_mdh);
    dfree(eng->d_mdz);
    dfree(eng->d_mdzb);
    dfree(eng->d_my);
    dfree(eng->d_mloss);
    eng->m_rows_cap = eng->m_train_cap = 0;
    eng->mlp_ready = false;
}

// Sets the FP32 parameters (host array of m_n values), refreshes the BF16
// copy and clears the optimiser state.
int mlp_load(csg_engine* eng, const float* w)
{
    HIPCHK(hipMemcpyAsync(eng->d_mw, w, eng->m_n * sizeof(float), hipMemcpyHostToDevice,
                          eng->stream));
    HIPCHK(hipMemsetAsync(eng->d_mm, 0, eng->m_n * sizeof(float), eng->stream));
    HIPCHK(hipMemsetAsync(eng->d_mv, 0, eng->m_n * sizeof(float), eng->stream));
    HIPCHK(hipMemsetAsync(eng->d_mg, 0, eng->m_n * sizeof(float), eng->stream));
    k_to_bf16<<<nblk(eng->m_n), 256, 0, eng->stream>>>(eng->d_mw, eng->m_n, eng->d_mwb);
    LAUNCHCHK();
    int rc = mlp_transpose(eng);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(eng->stream));
    eng->m_step = 0;
    return CSG_OK;
}

}  // namespace

extern "C" {

int csg_mlp_init(csg_engine* eng, int hidden, uint64_t seed)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (!eng->score_loaded)
        return eng->fail(CSG_ERR_STATE, "csg_mlp_init before csg_copyin_score");
    if (hidden < 64 || hidden > 4096 || hidden % 64)
        return eng->fail(CSG_ERR_ARGUMENT, "hidden width must be a multiple of 64 in [64, 4096]");
    HIPCHK(hipSetDevice(eng->device));
    free_mlp(eng);
    eng->mG  = eng->G;
    eng->mKp = eng->Kp;
    eng->mH  = (uint32_t)hidden;
    eng->mL  = eng->L;
    eng->mLp = (uint32_t)round_up(eng->L, 64);
    const size_t Kp = eng->mKp, H = eng->mH, Lp = eng->mLp;
    const size_t sizes[6] = { Kp * H, H, H * H, H, H * Lp, Lp };
    size_t       off      = 0;
    for (int i = 0; i < 6; ++i)
    {
        eng->m_off[i] = off;
        off += sizes[i];
    }
    eng->m_n = off;
    int rc;
    if ((rc = dalloc(eng, &eng->d_mw, off)) || (rc = dalloc(eng, &eng->d_mg, off)) ||
        (rc = dalloc(eng, &eng->d_mm, off)) || (rc = dalloc(eng, &eng->d_mv, off)) ||
        (rc = dalloc(eng, &eng->d_mwb, off)) || (rc = dalloc(eng, &eng->d_mwbT, off)))
        return rc;
    // He-normal hidden layers, variance 1 / fan_in at the output layer, zero
    // biases; padded input rows and output columns stay zero
    std::vector<float>              w(off, 0.f);
    std::mt19937_64                 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    const float s1 = std::sqrt(2.f / (float)eng->mG), s2 = std::sqrt(2.f / (float)H),
                s3 = std::sqrt(1.f / (float)H);
    for (size_t i = 0; i < eng->mG; ++i)
        for (size_t o = 0; o < H; ++o)
            w[eng->m_off[0] + i * H + o] = s1 * nd(rng);
    for (size_t i = 0; i < H * H; ++i)
        w[eng->m_off[2] + i] = s2 * nd(rng);
    for (size_t i = 0; i < H; ++i)
        for (size_t j = 0; j < eng->mL; ++j)
            w[eng->m_off[4] + i * Lp + j] = s3 * nd(rng);
    if ((rc = mlp_load(eng, w.data())))
        return rc;
    eng->mlp_ready = true;
    return CSG_OK;
}

int csg_mlp_dims(const csg_engine* eng, int64_t* dims)
{
    if (!eng || !dims)
        return CSG_ERR_ARGUMENT;
    if (!eng->mlp_ready)
        return CSG_ERR_STATE;
    dims[0] = eng->mG;
    dims[1] = eng->mKp;
    dims[2] = eng->mH;
    dims[3] = eng->mL;
    dims[4] = eng->mLp;
    dims[5] = (int64_t)eng->m_n;
    return CSG_OK;
}

int csg_mlp_get_params(csg_engine* eng, float* out)
{
    if (!eng || !out)
        return CSG_ERR_ARGUMENT;
    if (!eng->mlp_ready)
        return eng->fail(CSG_ERR_STATE, "no proposal network");
    HIPCHK(hipSetDevice(eng->device));
    HIPCHK(hipMemcpyAsync(out, eng->d_mw, eng->m_n * sizeof(float), hipMemcpyDeviceToHost,
                          eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return CSG_OK;
}

int csg_mlp_set_params(csg_engine* eng, const float* in)
{
    if (!eng || !in)
        return CSG_ERR_ARGUMENT;
    if (!eng->mlp_ready)
        return eng->fail(CSG_ERR_STATE, "no proposal network: call csg_mlp_init first");
    for (size_t i = 0; i < eng->m_n; ++i)
        if (!std::isfinite(in[i]))
            return eng->fail(CSG_ERR_ARGUMENT, "non-finite network parameter");
    HIPCHK(hipSetDevice(eng->device));
    return mlp_load(eng, in);
}

int csg_mlp_grad(csg_engine* eng, uint32_t n, const float* x, const int32_t* labels, float* grad,
                 float* loss)
{
    if (!eng || !x || !labels || !grad)
        return CSG_ERR_ARGUMENT;
    if (!eng->mlp_ready)
        return eng->fail(CSG_ERR_STATE, "no proposal network: call csg_mlp_init first");
    if (n == 0 || n > 65536)
        return eng->fail(CSG_ERR_ARGUMENT, "row count must be in [1, 65536]");
    for (uint32_t i = 0; i < n; ++i)
        if (labels[i] < 0 || (uint32_t)labels[i] >= eng->mL)
            return eng->fail(CSG_ERR_ARGUMENT, "label out of range");
    HIPCHK(hipSetDevice(eng->device));
    int rc;
    if ((rc = mlp_work(eng, n, true)))
        return rc;
    float* dx = nullptr;
    if ((rc = upload(eng, &dx, x, (size_t)n * eng->mG)))
        return rc;
    k_pad_bf16<<<nblk((size_t)n * eng->mKp), 256, 0, eng->stream>>>(dx, (int)n, (int)eng->mG,
                                                                  (int)eng->mKp, eng->d_mx);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(eng->d_my, labels, n * sizeof(int32_t), hipMemcpyHostToDevice,
                           eng->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(eng->d_mloss, 0, sizeof(float), eng->stream);
    if (e != hipSuccess)
    {
        (void)hipStreamSynchronize(eng->stream);
        dfree(dx);
        return eng->fail(CSG_ERR_DEVICE, std::string("csg_mlp_grad: ") + hipGetErrorString(e));
    }
    rc = mlp_forward(eng, (int)n, true);
    if (!rc)
        rc = mlp_loss(eng, (int)n);
    if (!rc)
        rc = mlp_backward(eng, (int)n);
    (void)hipStreamSynchronize(eng->stream);
    dfree(dx);
    if (rc)
        return rc;
    HIPCHK(hipMemcpy(grad, eng->d_mg, eng->m_n * sizeof(float), hipMemcpyDeviceToHost));
    if (loss)
        HIPCHK(hipMemcpy(loss, eng->d_mloss, sizeof(float), hipMemcpyDeviceToHost));
    return CSG_OK;
}

int csg_mlp_train(csg_engine* eng, const int32_t* labels, int epochs, int batch, float lr,
                  uint64_t seed, float* epoch_loss)
{
    if (!eng || !labels)
        return CSG_ERR_ARGUMENT;
    int rc;
    if ((rc = mlp_check(eng)))
        return rc;
    if (epochs < 1 || batch < 1 || batch > 65536 || !(lr > 0.f) || !std::isfinite(lr))
        return eng->fail(CSG_ERR_ARGUMENT, "epochs >= 1, batch in [1, 65536] and lr > 0 required");
    const uint32_t N = eng->Nq;
    for (uint32_t i = 0; i < N; ++i)
        if (labels[i] < 0 || (uint32_t)labels[i] >= eng->mL)
            return eng->fail(CSG_ERR_ARGUMENT, "label out of range");
    HIPCHK(hipSetDevice(eng->device));
    const size_t Kp = eng->Kp;
    __bf16*      feat = nullptr;
    int32_t*     lab  = nullptr;
    uint32_t*    perm = nullptr;
    auto         cleanup = [&]() {
        (void)hipStreamSynchronize(eng->stream);
        dfree(feat);
        dfree(lab);
        dfree(perm);
    };
    if ((rc = dalloc(eng, &feat, (size_t)N * Kp)) || (rc = upload(eng, &lab, labels, N)) ||
        (rc = dalloc(eng, &perm, N)) || (rc = mlp_work(eng, (size_t)batch, true)))
    {
        cleanup();
        return rc;
    }
    // features of every loaded query cell: ranks over the score gene set / G
    for (uint32_t q0 = 0; q0 < N && !rc; q0 += eng->Qc)
    {
        const uint32_t nq = std::min(eng->Qc, N - q0);
        rc = rank_rows(eng, eng->d_qdata, eng->d_qind, eng->d_qptr, q0, nq, eng->d_Xq, nullptr,
                       eng->d_irq, feat + (size_t)q0 * Kp);
    }
    std::vector<uint32_t> order(N);
    std::mt19937_64       rng(seed ^ 0x9E3779B97F4A7C15ull);
    const int64_t         per_epoch = (N + batch - 1) / batch;
    const int64_t         total     = per_epoch * epochs;
    int64_t               it        = 0;
    for (int e = 0; e < epochs && !rc; ++e)
    {
        for (uint32_t i = 0; i < N; ++i)
            order[i] = i;
        std::shuffle(order.begin(), order.end(), rng);
        hipError_t he = hipMemcpyAsync(perm, order.data(), N * sizeof(uint32_t),
                                       hipMemcpyHostToDevice, eng->stream);
        if (he == hipSuccess)
            he = hipMemsetAsync(eng->d_mloss, 0, sizeof(float), eng->stream);
        if (he != hipSuccess)
        {
            rc = eng->fail(CSG_ERR_DEVICE, std::string("csg_mlp_train: ") + hipGetErrorString(he));
            break;
        }
        for (uint32_t b0 = 0; b0 < N && !rc; b0 += (uint32_t)batch, ++it)
        {
            const int rows = (int)std::min<uint32_t>((uint32_t)batch, N - b0);
            k_gather_batch<<<nblk((size_t)rows * (Kp / 8)), 256, 0, eng->stream>>>(
                feat, lab, perm + b0, rows, (int)Kp, eng->d_mx, eng->d_my);
            if (hipGetLastError() != hipSuccess)
            {
                rc = eng->fail(CSG_ERR_DEVICE, "batch gather launch failed");
                break;
            }
            // cosine decay from lr to lr / 20 over all steps
            const double f    = total > 1 ? (double)it / (double)(total - 1) : 0.0;
            const float  lr_t = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.14159265358979323846 * f))));
            if ((rc = mlp_forward(eng, rows, true)) || (rc = mlp_loss(eng, rows)) ||
                (rc = mlp_backward(eng, rows)) || (rc = mlp_adam(eng, lr_t)))
                break;
        }
        if (!rc)
        {
            float sum = 0.f;
            he        = hipMemcpyAsync(&sum, eng->d_mloss, sizeof(float), hipMemcpyDeviceToHost,
                                       eng->stream);
            if (he == hipSuccess)
                he = hipStreamSynchronize(eng->stream);
            if (he != hipSuccess)
                rc = eng->fail(CSG_ERR_DEVICE, std::string("csg_mlp_train: ") + hipGetErrorString(he));
            else if (epoch_loss)
                epoch_loss[e] = sum / (float)per_epoch;
        }
    }
    cleanup();
    return rc;
}

int csg_mlp_predict(csg_engine* eng, int k, int32_t* top, float* margin)
{
    if (!eng || !top || !margin)
        return CSG_ERR_ARGUMENT;
    int rc;
    if ((rc = mlp_check(eng)))
        return rc;
    if (k < 1 || k > 8 || (uint32_t)k > eng->L)
        return eng->fail(CSG_ERR_ARGUMENT, "k must be in [1, min(8, labels)]");
    HIPCHK(hipSetDevice(eng->device));
    int32_t* dtop = nullptr;
    float*   dmar = nullptr;
    if ((rc = dalloc(eng, &dtop, (size_t)eng->Qc * k)) || (rc = dalloc(eng, &dmar, eng->Qc)))
    {
        dfree(dtop);
        return rc;
    }
    for (uint32_t q0 = 0; q0 < eng->Nq && !rc; q0 += eng->Qc)
    {
        const uint32_t nq = std::min(eng->Qc, eng->Nq - q0);
        if ((rc = mlp_work(eng, nq, false)) ||
            (rc = rank_rows(eng, eng->d_qdata, eng->d_qind, eng->d_qptr, q0, nq, eng->d_Xq,
                            nullptr, eng->d_irq, eng->d_mx)) ||
            (rc = mlp_propose(eng, nq, k, dtop, dmar)))
            break;
        hipError_t he = hipMemcpyAsync(top + (size_t)q0 * k, dtop, (size_t)nq * k * sizeof(int32_t),
                                       hipMemcpyDeviceToHost, eng->stream);
        if (he == hipSuccess)
            he = hipMemcpyAsync(margin + q0, dmar, nq * sizeof(float), hipMemcpyDeviceToHost,
                                eng->stream);
        if (he == hipSuccess)
            he = hipStreamSynchronize(eng->stream);
        if (he != hipSuccess)
            rc = eng->fail(CSG_ERR_DEVICE, std::string("csg_mlp_predict: ") + hipGetErrorString(he));
    }
    (void)hipStreamSynchronize(eng->stream);
    dfree(dtop);
    dfree(dmar);
    return rc;
}

int csg_set_proposal(csg_engine* eng, int k, float min_margin)
{
    if (!eng)
        return CSG_ERR_ARGUMENT;
    if (k < 0 || k > 8)
        return eng->fail(CSG_ERR_ARGUMENT, "proposal size must be in [0, 8]");
    if (!(min_margin >= 0.f && min_margin <= 1.f))
        return eng->fail(CSG_ERR_ARGUMENT, "minimum margin must be in [0, 1]");
    eng->prop_k      = k;
    eng->prop_margin = min_margin;
    return CSG_OK;
}

}  // extern "C"
