// mw_batch.h -- batched evaluation over several resident wedges: the
// planner and runner shared by the batch C ABI and the request server.
#pragma once

#include <vector>

#include "mw_context.h"

namespace mw {

// Evaluation constants of one parameter vector of a context (the vector is
// laid out as set_astronomy_parameters reads it; NULL keeps the context's
// template values).
int slot_constants(mw_ctx* c, const double* vec, EvalConst* k);

// One parameter vector of one context inside a batch.
struct Slot {
    mw_ctx* ctx;
    EvalConst k;
};

// Plans and runs one batch on one stream with its own buffers:
//   H2D constants -> reset counters -> persistent integral launch(es) ->
//   reduction -> I = main - cuts on device -> star pass -> reduction -> D2H.
// Work items are ordered by cost per block, largest first, so the
// persistent grid's tail is short; results are per logical block and
// reduced in fixed order, so every slot's result is bitwise independent of
// what else is in the batch.
class BatchRunner {
public:
    BatchRunner() = default;
    ~BatchRunner();
    BatchRunner(const BatchRunner&) = delete;
    BatchRunner& operator=(const BatchRunner&) = delete;

    int init(int device, int flags);
    // Builds the work items for these slots (same device as init).  Must be
    // called again when the slots, their contexts or their background modes
    // change; keeps constants in the pinned staging buffer.
    // precision: MW_PREC_FP64, MW_PREC_MIXED or MW_PREC_RESAMPLED.
    int plan(const std::vector<Slot>& slots, int precision);
    // Rewrites the staged constants of the planned slots (same modes).
    void set_constants(const std::vector<Slot>& slots);
    bool same_modes(const std::vector<Slot>& slots) const;
    int table_choice(const RegionDev& rd, const EvalConst& k) const;
    int star_level(const mw_ctx& c, const EvalConst& k) const;
    // Enqueues the whole sequence on the runner's stream (capturable).  With
    // kernel_events, events bracket the integral launches (not while a
    // graph is being captured).
    int enqueue(bool kernel_events = false);
    double integral_ms();            // between the kernel events of the last enqueue
    int record_start();
    int record_end();
    // Waits, then fills one mw_result per slot.
    int finish(mw_result* results);
    bool query_done();               // non-blocking completion test
    double elapsed_ms();             // between record_start and record_end
    void set_variant(int variant, int blocks_per_cu) { variant_ = variant; blocks_per_cu_ = blocks_per_cu; }
    // Largest per-work-item scratch (register spill) size of the integral
    // kernels the planned batch would launch at this variant.
    int variant_scratch(int variant, size_t* bytes) const;
    int n_slots() const { return (int)slot_n_regions_.size(); }
    size_t scratch_bytes() const;
    hipStream_t stream() const { return stream_; }
    long integral_blocks() const { return total_iblocks_; }
    long integral_points() const { return points_; }
    long integral_points_recurrence() const { return points_rec_; }
    long integral_points_background() const { return points_bg_; }
    long integral_stream_terms() const { return terms_; }
    long star_blocks() const { return total_sblocks_; }

    static size_t scratch_bytes_for(const std::vector<Slot>& slots);

private:
    struct Group {
        int bg;
        int first_item, n_items, first_block, n_blocks;
        bool split = true;   // every item in the split layout (integral groups)
        bool exact = true;   // no item on a distance-grid recurrence (integral groups)
    };
    int device_ = 0, flags_ = 0, n_cu_ = 256;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[4] = {};
    bool mixed_ = false, resample_ = false;
    int variant_ = 0, blocks_per_cu_ = 2;
    std::vector<int> levels_;        // quadrature table (level, layout) per (slot, region), slot-major
    long points_ = 0;                // integrand evaluations of the planned integral pass
    long points_rec_ = 0;            // of those, on distance grids by recurrence
    long points_bg_ = 0;             // background evaluations (strided on distance grids)
    long terms_ = 0;                 // stream terms by recurrence
    std::vector<int> slot_bg_, slot_n_regions_, slot_first_row_, slot_S_;
    std::vector<Group> igroups_, sgroups_;
    int n_ritems_ = 0, total_iblocks_ = 0, total_sblocks_ = 0;
    std::vector<long> slot_stars_;
    DevBuf<EvalConst> consts_;
    DevBuf<RegionLaunch> ritems_;
    DevBuf<StarLaunch> sitems_;
    DevBuf<ReduceItem> rred_, sred_;
    DevBuf<double> ipart_, spart_, irows_, srows_;
    DevBuf<int> first_, count_;
    DevBuf<unsigned> counters_;
    HostBuf<EvalConst> h_consts_;
    HostBuf<double> h_irows_, h_srows_;
    HostBuf<unsigned char> h_plan_;   // staging of the planned work items
};

}  // namespace mw
