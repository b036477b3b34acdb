// Internal runtime of the propagation library: the context, its buffer
// and table caches, the problem descriptors and the batch builder shared by
// the C entry points (api.cpp, inverse.cpp, observe.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/optics_gpu.h"
#include "host_guards.hpp"
#include "prop_kernels.hpp"

namespace ocl {
hipError_t launch_pass_c64(int pass, const ocd::ProbDev* d, int np, int total, int block, size_t lds, int grid,
                           hipStream_t st);
hipError_t launch_pass_c128(int pass, const ocd::ProbDev* d, int np, int total, int block, size_t lds, int grid,
                            hipStream_t st);
hipError_t set_lds_limits();
hipError_t launch_mask_apply_c64(const void* in, void* out, const ocd::MaskDev& m, int nx, int ny, int* flag,
                                 hipStream_t st);
hipError_t launch_mask_trans(bool dbl, void* out, const ocd::MaskDev& m, int nx, int ny, hipStream_t st);
hipError_t launch_mask_apply_c128(const void* in, void* out, const ocd::MaskDev& m, int nx, int ny, int* flag,
                                  hipStream_t st);
hipError_t launch_check_finite(bool dbl, const void* in, size_t n, int* flag, hipStream_t st);
hipError_t launch_copy(bool dbl, const void* in, void* out, size_t n, hipStream_t st);
hipError_t launch_fill_int(int* p, int v, int n, hipStream_t st);
hipError_t fast_launch_c64(const ocd::FastSel& s, const ocd::ProbDev* d, int np, int total, int grid, hipStream_t st);
hipError_t fast_launch_c128(const ocd::FastSel& s, const ocd::ProbDev* d, int np, int total, int grid, hipStream_t st);
hipError_t fast_launch_chain_c64(int which, int n, int v, const ocd::ProbDev* a, const ocd::ProbDev* f, int np,
                                 int total, int grid, hipStream_t st);
hipError_t fast_launch_chain_c128(int which, int n, int v, const ocd::ProbDev* a, const ocd::ProbDev* f, int np,
                                  int total, int grid, hipStream_t st);
hipError_t fast_launch_link_c64(int pk, int n, int v, const ocd::ProbDev* a, const ocd::ProbDev* b, int np,
                                int total, int grid, hipStream_t st);
hipError_t fast_launch_link_c128(int pk, int n, int v, const ocd::ProbDev* a, const ocd::ProbDev* b, int np,
                                 int total, int grid, hipStream_t st);
void fast_layout_c64(int n, int v, int* G, int* B, int* CB);
void fast_layout_c128(int n, int v, int* G, int* B, int* CB);
hipError_t fast_set_lds_limits_c64();
hipError_t fast_set_lds_limits_c128();

// Number of layout variants of the power-of-two kernels at length n (0: none).
inline int fast_variants(int n) {
    switch (n) {
        case 64: case 128: case 256: case 512: return 1;
        case 1024: case 2048: case 4096: case 8192: return 4;
        default: return 0;
    }
}

// true when n has a specialised kernel; fills its workgroup layout.
inline bool fast_supported(bool dbl, int n, int v, int* G, int* B, int* CB) {
    if (fast_variants(n) == 0) return false;
    if (dbl) fast_layout_c128(n, v, G, B, CB);
    else fast_layout_c64(n, v, G, B, CB);
    return *G > 0;
}
hipError_t launch_seed(bool dbl, const void* u, const double* dmdi, void* out, size_t n, hipStream_t st);
hipError_t launch_adj_elem(bool dbl, void* co, const void* ckpt, double* grad, const ocd::MaskDev& m, int nx, int ny,
                           hipStream_t st);
hipError_t launch_block_reorder(bool dbl, const void* src, void* dst, int nx, int ny, int B, hipStream_t st);
// sums: 8 + ocd::kSeedParts doubles; sums[0] receives the merit sum.
hipError_t launch_seed_mse(bool dbl, const void* u, const double* target, void* out, size_t n, double* sums,
                           hipStream_t st);
// MseToTarget seeds of the [frame][n] fields (frame = b * ns + s) for the
// band/scene selector (-1 = all); sums as launch_seed_mse (sums[0] = the sum
// of the squared differences of the selected frames' summed intensity).
hipError_t launch_seed_frames(bool dbl, const void* u, const double* target, void* out, size_t n, int nb, int ns,
                              int sel_b, int sel_s, double* sums, hipStream_t st);
// out[i] = sum over frames of slab[f][i], in frame order (m per frame).
hipError_t launch_frame_sum(const double* slab, double* out, size_t m, int frames, hipStream_t st);
// gs::run bookkeeping between the passes of an iteration, one workgroup
// per restart (parts: pstride doubles per restart, nparts[r] entries of 4).
hipError_t launch_gs_after_fwd(ocd::GsState* st, const double* parts, const int* nparts, size_t pstride, size_t n,
                               ocd::GsRec* hist, size_t hstride, int restarts, hipStream_t stream);
hipError_t launch_gs_after_bwd(ocd::GsState* st, const double* parts, const int* nparts, size_t pstride, size_t n,
                               ocd::GsRec* hist, size_t hstride, int restarts, hipStream_t stream);
hipError_t launch_gs_keep_best(const ocd::GsState* st, const double* phase, double* best, size_t n, int restarts,
                               hipStream_t stream);
hipError_t big_set_lds_limits();
hipError_t big_line(bool dbl, int dir, const void* in, void* out, const ocd::LineSpec& s, const ocd::FftPlanDev& pl,
                    const void* twl, size_t lds, int max_grid, hipStream_t st);
hipError_t big_embed(bool dbl, const ocd::ProbDev& P, void* a, hipStream_t st);
hipError_t big_kernel(bool dbl, const ocd::ProbDev& P, void* k, hipStream_t st);
hipError_t big_transfer(bool dbl, const ocd::ProbDev& P, void* a, const void* kspec, hipStream_t st);
hipError_t big_crop(bool dbl, const ocd::ProbDev& P, const void* a, hipStream_t st);
hipError_t launch_render_obs(bool dbl, const void* u, double* obs, size_t n, int frames, int nr, int cx, int cy,
                             int cl, double hs, double hl, int* flag, hipStream_t st);
hipError_t launch_mse(bool dbl, const void* u, const double* target, size_t n, int nr, double* part, double* merit,
                      hipStream_t st);
hipError_t launch_mtf(const double* psf, size_t nx, size_t ny, size_t count, double* sums, const void* tw,
                      double* values, int* flag, hipStream_t st);
hipError_t launch_adam(double* x, const double* g, double* m, double* v, const double* lo, const double* hi, size_t n,
                       double* part, double* scale, int has_clip, double cn, double lr, double b1, double b2,
                       double eps, double bc1, double bc2, hipStream_t st);
}  // namespace ocl

namespace ocr {
using namespace ocd;


constexpr size_t kLdsBudget = 140 * 1024;  // per workgroup (gfx950: 160 KB per CU)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct TableEntry {
    void* dev = nullptr;
    size_t bytes = 0;
};

}  // namespace ocr

struct oc_context {
    int device = 0;
    int num_cu = 256;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, ocr::DevBuf> bufs;
    std::map<std::pair<int, int>, void*> twiddles;  // (n, double?) -> table
    std::map<std::string, ocr::TableEntry> tables;       // per-parameter axis tables
    size_t tables_bytes = 0;
    // host staging for descriptor uploads: a ring of slots, each with the
    // event recorded after the last copy out of it, so that building a
    // batch waits only on the copy kPinSlots uploads back
    static constexpr int kPinSlots = 8;
    void* pinned[kPinSlots] = {};
    size_t pinned_cap[kPinSlots] = {};
    hipEvent_t pinned_done[kPinSlots] = {};
    int pin_next = 0;
    // pinned host buffer for small read-backs at the end of a call (a
    // pageable destination makes the runtime stage and wait per copy)
    void* readback = nullptr;
    size_t readback_cap = 0;
    // Device copies of recent transient descriptor arrays.  A batch whose
    // descriptors are byte-identical to a kept copy (the same problems on
    // the same buffers, as in an inverse-design loop) launches on that copy
    // instead of uploading again.  Slots are reused least recently used
    // first; an overwrite is queued on the stream after every kernel that
    // reads the old contents.
    static constexpr int kDescSlots = 64;
    static constexpr size_t kDescSlotBytes = 16384;
    void* desc_pool = nullptr;
    std::vector<unsigned char> desc_host[kDescSlots];
    uint64_t desc_hash[kDescSlots] = {};
    uint64_t desc_used[kDescSlots] = {};
    uint64_t desc_clock = 0;
    // per-pass timing
    bool timing = false;
    uint32_t timing_mask = 0xFFFFFFFFu;  // pass kinds that get timing brackets
    struct Pending {
        int pass;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double pass_ms[OC_NUM_PASS_KINDS] = {0};
    uint64_t pass_n[OC_NUM_PASS_KINDS] = {0};
    bool force_generic = false;  // route every batch through the runtime-plan kernels
    void* queued = nullptr;      // batches submitted without waiting (QueuedSet, created on first use)
    std::recursive_mutex mu;     // held by every entry point for the duration of its call
    // No-allocation mode (oc_propagate_batch_noalloc_dev_c64): while `arena`
    // is set, every device buffer, table and descriptor of the call is carved
    // from the device's scratch array [scratch, scratch + scratch_cap) and
    // filled from host copies kept in arena_keep until the call has
    // synchronised; arena_tables holds this call's tables by key.
    void* scratch = nullptr;  // this device's scratch array (hipGetSymbolAddress at creation)
    size_t scratch_cap = 0;
    bool arena = false;
    size_t arena_used = 0;
    std::vector<std::vector<unsigned char>> arena_keep;
    std::map<std::string, const void*> arena_tables;
    std::map<std::pair<int, int>, int> fft_variant;  // (n, double?) -> fast-kernel layout variant
    // batch schedules by (path pick?, length, double?, size class)
    std::map<std::tuple<int, uint64_t, int, uint64_t>, oc_schedule> sched;
    std::map<int, std::pair<int, int>> sep_variant;  // n -> (column, row) layouts of the two-pass Fresnel
    // The autotuned layout if there is one, else the layout that timed
    // fastest for a masked-ASM + Fresnel round on MI355X (oc_autotune_fft
    // over 16 problems, both precisions; 1024/2048 complex128 are within 1%
    // across layouts).
    int variant_for(int n, bool dbl) const {
        auto it = fft_variant.find(std::make_pair(n, dbl ? 1 : 0));
        if (it != fft_variant.end()) return it->second;
        switch (n) {
            case 1024: case 2048: return 1;
            case 4096: case 8192: return dbl ? 3 : 2;
            default: return 0;
        }
    }
    // Layouts of the two passes of the separable complex64 Fresnel, which
    // need only agree on the blocked width B of their intermediate: the
    // tuned or forced pair if there is one, else (for 4096) the pair that
    // timed fastest on MI355X (columns with V0's 4-column workgroups, rows
    // with V3's 2-row ones), else the propagation layout for both.
    std::pair<int, int> sep_variants_for(int n) const {
        auto it = sep_variant.find(n);
        if (it != sep_variant.end()) return it->second;
        if (n == 4096 && fft_variant.find(std::make_pair(n, 0)) == fft_variant.end()) return {0, 3};
        const int v = variant_for(n, false);
        return {v, v};
    }
};

// Serialises the calls made on one context (its caches, buffers and stream
// order); calls on different contexts proceed concurrently.  Recursive, so
// an entry point may call another.
#define OC_CONTEXT_LOCK(ctx)                                                                           \
    std::unique_lock<std::recursive_mutex> oc_context_lock_ =                                        \
        (ctx) ? std::unique_lock<std::recursive_mutex>((ctx)->mu) : std::unique_lock<std::recursive_mutex>()

namespace ocr {

int fail(oc_context* ctx, int code, const std::string& msg);
hipEvent_t pool_event(oc_context* ctx);
// Records a timing bracket around `launch` when timing is enabled.
template <typename F>
hipError_t timed(oc_context* ctx, int pass, F&& launch) {
    if (!ctx->timing || !((ctx->timing_mask >> pass) & 1u)) return launch();
    hipEvent_t a = pool_event(ctx), b = pool_event(ctx);
    if (!a || !b) return launch();
    (void)hipEventRecord(a, ctx->stream);
    hipError_t e = launch();
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({pass, a, b});
    return e;
}

void collect_timing(oc_context* ctx);
#define HIP_TRY(ctx, expr)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(ctx, e_ == hipErrorOutOfMemory ? OC_ERR_OUT_OF_MEMORY : OC_ERR_DEVICE,       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                         \
    } while (0)

int ensure_buf(oc_context* ctx, const std::string& name, size_t bytes, void** out);
// No-allocation mode: `bytes` (256-byte aligned) of the scratch array, or
// OC_ERR_INVALID_ARGUMENT when it is exhausted.
int arena_alloc(oc_context* ctx, size_t bytes, void** out);
// No-allocation mode: a scratch copy of host data (queued on the stream; the
// host copy is kept until the call ends).
int arena_upload(oc_context* ctx, const void* host, size_t bytes, const void** out);
// Scratch bytes a no-allocation batch of these (unprepared) problems needs at
// most; every allocation of the call fits in it.
size_t arena_need(const std::vector<struct Prob>& probs, bool dbl);
// Looks up / stores a device table by key: the context's cache, or this
// call's scratch copies in the no-allocation mode.
const void* find_table(oc_context* ctx, const std::string& key);
int store_table(oc_context* ctx, const std::string& key, const void* host, size_t bytes, const void** out);
// Problems of a batch call from its descriptors (engine and mask checks).
int probs_from_descs(const oc_prop_desc* d, size_t count, std::vector<struct Prob>& v);
// Attaches this device's scratch array to a new context (noalloc.hip).
int noalloc_attach(oc_context* ctx);
// Staging slot `slot` with room for `bytes` (the caller has waited for the
// slot's previous copy).
int ensure_pinned(oc_context* ctx, int slot, size_t bytes);
int ensure_readback(oc_context* ctx, size_t bytes, void** out);
int get_twiddles(oc_context* ctx, int n, bool dbl, const void** out);
bool make_radices(int n, std::vector<int>& rad, bool& generic);
// Plan of length n: the default radices, or rad[0..nrad) when given.
int make_plan(oc_context* ctx, int n, bool dbl, FftPlanDev* pl, const int* rad = nullptr, int nrad = 0);
// A radix plan the in-place LDS stages run: radices 2, 3, 4, 5, 7, 8, 16
// with product n, at most kMaxStages of them.
bool valid_radices(int n, const int* rad, int nrad);
// Batch schedules (oc_schedule_*): size class of a count, the context's
// pick for (pad?, length, precision, class) or the nearest class's, and
// their application to a call's prepared problems.
uint64_t size_class(size_t n);
const oc_schedule* find_sched(const oc_context* ctx, bool pad, uint64_t len, bool dbl, uint64_t cls);
void apply_schedules(oc_context* ctx, bool dbl, std::vector<struct Prob>& probs);
size_t plan_lds_elems(const FftPlanDev& pl);
// Internal description of one propagation problem.
struct Prob {
    oc_band_geom g;
    double z;
    int engine;
    int band_limited, strict;
    double padding, max_padding, medium;
    const oc_phase_mask* mask;
    const void* in;
    void* out;
    int status = OC_OK;
    bool launch = false;     // goes through the FFT passes
    bool check_only = false; // finiteness check first (z == 0 copy / guard failure)
    int pending = OC_OK;     // status to report if the input is finite
    size_t nxp = 0, nyp = 0;
    size_t ref_nxp = 0, ref_nyp = 0;  // the reference's padded size (nxp/nyp may run longer)
    // schedule of the run length: -1 default, 0 LDS kernels with rad[0..nrad)
    // (nrad 0: the default plan), 1 register-resident kernels in `variant`
    int path = -1;
    int variant = -1;
    int nrad = 0;
    int rad[kMaxStages] = {};
    oc_asm_report asm_rep{};
    oc_fresnel_report fr_rep{};
    oc_rs_report rs_rep{};
    oc_fraunhofer_config fh_cfg{};
    oc_fraunhofer_report fh_rep{};
    OpsDev ops{};
};

std::string fmtkey(const char* tag, std::initializer_list<double> vals);
int fresnel_factors(oc_context* ctx, const Prob& p, EngineDev* e);
int engine_tables(oc_context* ctx, const Prob& p, bool dbl, EngineDev* e);
int mask_footprint(const oc_band_geom& g, const oc_phase_mask& m);
int mask_tables(oc_context* ctx, const oc_band_geom& g, const oc_phase_mask& m, const double* values_dev,
                MaskDev* md);
int pick_pow2_divisor(size_t n, int cap);
void prepare(Prob& p);
struct PassCfg {
    int block = 64;
    size_t lds = 0;
    int total = 0;
};

// Workgroup shape of one problem: transforms per row-pass workgroup (cnt) and
// the column block width B of the blocked intermediate.
struct Layout {
    int cnt = 1, B = 1;
    int ccnt = 1;  // columns per column-pass workgroup (divides B)
    size_t row_bytes = 0, col_bytes = 0;
    bool ok = false;
};

Layout choose_layout(const FftPlanDev& prow, const FftPlanDev& pcol, size_t esz, size_t ny, size_t nyp, bool rs);
struct Batch {
    bool dbl = false;
    bool fast = false;  // power-of-two kernels (all problems share P_x = P_y = fast_n)
    int fast_n = 0;
    int fast_v = 0;
    int fast_v_row = 0;      // layout of the separable Fresnel's row pass (pass 2)
    int post_kind = -2;      // post-op shared by every problem (-1: mixed)
    int col_kind = -2;       // engine shared by every problem (-1: mixed)
    int pre_kind = -2;       // pre-op shared by every problem (-1: mixed)
    bool sep = false;        // separable complex64 Fresnel: two passes
    bool half = false;       // power-of-two batch whose problems are all N = P/2 (the chain's specialised kernels)
    bool rs = false;
    std::vector<ProbDev> pd;
    std::vector<int> idx;  // index into the problem vector of each descriptor
    void* dprobs = nullptr;
    PassCfg pc[4];
    std::vector<void*> owned;  // allocations of a persistent batch
};

void free_batch(oc_context* ctx, Batch& b);
// Batches of oc_propagate_batch_submit_dev_* not yet waited for.
struct Queued {
    std::vector<Prob> probs;
    int* dflags = nullptr;
    oc_prop_desc* descs = nullptr;
    // a submitted chain: probs holds the first stages then the second
    // stages; the statuses go to descs (first) and second (a failed first
    // stage fails the chain)
    oc_prop_desc* second = nullptr;
};

struct QueuedSet {
    std::vector<Queued> batches;
};

QueuedSet& queued_set(oc_context* ctx);
void release_queued(oc_context* ctx);
int batch_alloc(oc_context* ctx, Batch& b, bool persistent, const std::string& name, size_t bytes, void** out);
// force_v >= 0: every pass of the batch uses layout variant force_v (the
// separable Fresnel's two passes included) instead of the context's choice.
int build_batch(oc_context* ctx, bool dbl, std::vector<Prob>& probs, int* dflags, Batch& b, bool persistent,
                const std::vector<int>* subset = nullptr, const std::string& tag = "", int force_v = -1);
constexpr unsigned kAllPasses = 0xF;
// pass 4 (power-of-two batches with a GS phase post-op): one iteration's
// inverse row pass fused with the next one's forward row pass
constexpr int PASS_GS_ROW = 4;
// pass 5: the same with beta = 1 and 2 pi wide bounds (amplitude replacement;
// the phases are left to the last iteration's inverse pass)
constexpr int PASS_GS_AMP = 5;
// timing slots of the separable Fresnel passes (see OC_NUM_PASS_KINDS)
constexpr int PASS_T_SEP_COL = 6, PASS_T_SEP_ROW = 7;
int launch_batch(oc_context* ctx, Batch& b, unsigned mask = kAllPasses);
void evict_tables(oc_context* ctx);
// zero_flags = false: the caller has zeroed dflags[0..probs.size()) on the stream
int launch_problems(oc_context* ctx, bool dbl, std::vector<Prob>& probs, int* dflags, bool zero_flags = true);
int finish_problems(oc_context* ctx, std::vector<Prob>& probs, const int* dflags);
// Consecutive propagation steps linked through their rows (f_row_link);
// 0 = launched, 1 = the run does not qualify (nothing launched), < 0 = -error.
int launch_linked(oc_context* ctx, bool dbl, std::vector<std::vector<Prob>>& run, int* dflags,
                  const std::vector<int>& off, std::vector<std::vector<Prob>>& queued, bool keep_mid);
int run_batch(oc_context* ctx, bool dbl, std::vector<Prob>& probs);
// true when the batched passes hold problem p's row and column transforms on
// chip (else it runs through launch_big)
bool fits_on_chip(const Prob& p, bool dbl);
// One prepared problem through the multi-pass path (padded transforms beyond
// the on-chip size) with its finiteness flag at *flag; 0 when queued.
int launch_big(oc_context* ctx, bool dbl, Prob& p, int* flag);
Prob make_prob(const oc_band_geom* g, double z, int engine, int band_limited, int strict, double padding,
               double max_padding, double medium, const oc_phase_mask* mask, const void* in, void* out);
bool geom_ok(const oc_band_geom* g);
// Host-memory wrapper: H2D, run, D2H.
int run_host(oc_context* ctx, Prob p, const double* in, double* out, const oc_phase_mask* host_mask, Prob* done);

// Device allocation released at scope exit (after a stream sync).
struct Scoped {
    oc_context* ctx;
    std::vector<void*> ptrs;
    explicit Scoped(oc_context* c) : ctx(c) {}
    ~Scoped() {
        if (!ptrs.empty()) (void)hipStreamSynchronize(ctx->stream);
        for (void* p : ptrs) (void)hipFree(p);
    }
    int alloc(size_t bytes, void** out) {
        hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 256));
        if (e != hipSuccess) return fail(ctx, e == hipErrorOutOfMemory ? OC_ERR_OUT_OF_MEMORY : OC_ERR_DEVICE,
                                         std::string("hipMalloc: ") + hipGetErrorString(e));
        ptrs.push_back(*out);
        return OC_OK;
    }
};

}  // namespace ocr
