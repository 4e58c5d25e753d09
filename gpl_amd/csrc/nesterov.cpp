// The Nesterov placement loop on the GPU: the host side of NesterovPlace / NesterovBase
// (nesterov.rs:2512-3168, 4115-4610) driving the HIP kernels of kernels.hip.
//
// The per-cell, per-pin, per-net and per-bin work runs on the device; the scalar control (step
// length acceptance, the penalty and WA-coefficient schedule, divergence, convergence) runs
// here on values the device reduces in FP64 and hands back once per back-tracking trial (the
// sums straight into pinned memory).  Slot rotation (prev <- cur <- next) is a swap of device
// pointers.  In launch modes 4-6 the same decisions are taken on the device by the trial's last
// kernel (device-driven trials): the host then steps in only where the device halts, on the
// events that need it (a zero or non-finite sum, a divergence, a new snapshot, a convergence, a
// re-sort, the iteration limit), and the rows are those of the host-driven loop, bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "host.h"
#include "kernels.h"

namespace {

using gpl::strfmt;
using Clock = std::chrono::steady_clock;

enum { S_CUR_SLP, S_CUR_SLP_WL, S_CUR_SLP_DEN, S_CUR_SLP_SUM, S_NEXT_SLP, S_NEXT_SLP_WL, S_NEXT_SLP_DEN, S_NEXT_SLP_SUM,
       S_PREV_SLP, S_PREV_SLP_WL, S_PREV_SLP_DEN, S_PREV_SLP_SUM, S_CUR, S_NEXT, S_SNAP, S_SNAP_SLP, S_SNAP_SLP_SUM,
       S_SNAP_PREV_SLP_SUM, S_COUNT };
enum Slot { SLOT_PREV, SLOT_CUR, SLOT_NEXT };

struct HipError {
    hipError_t e;
    const char* what;
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError{e, what};
}

struct Region {
    gpl_region R{};
    int32_t n = 0;
    gplk::RegionDev dev{};
    gplk::SolveBufs sb{};
    float2* v[S_COUNT] = {};
    double* partials = nullptr;
    bool has_locked = false;
    int64_t overflow = 0, overflow_unscaled = 0;
    float snapshot_density_penalty = 0, snapshot_step_length = 0, step_length = 0, density_penalty = 0;
    float wl_grad_sum = 0, density_grad_sum = 0, sum_overflow = 0, sum_overflow_unscaled = 0, base_wl_coef_nb = 0;
    int64_t prev_hpwl = 0, prev_reported_hpwl = 0;
    float prev_reported_overflow_unscaled = 0, min_sum_overflow = 1e30f, hpwl_with_min_sum_overflow = 1e30f;
    float coordi_distance = 0, peak_coordi_distance = 0, max_phi_coef = 0;
    bool max_phi_coef_changed = false, is_converged = false, is_diverged = false, reprint_header = false;
    int64_t nia = 0;  // nesterov_insts_area
    int32_t edges_rounded = 0;  // dev.edges_rounded as found at construction
    int32_t narrow_ok = 0;      // 32-bit scatter sums are exact for this region's bins
};

// The autotuner's pick for one region (kernel variants with bit-identical results: the scatter's
// chunk of cells per block and its LDS sum width, the gradient's LDS staging of the pin terms) and
// where it came from: 0 the defaults or a caller's choice, 1 measured here, 2 a cached measurement
// of the same shape.
struct Tune {
    int32_t chunk = gplk::kPlaceChunk, narrow = 0, stage = 1, source = 0;
    double scatter_us = 0.0, gradient_us = 0.0;
    bool decayed = false;  // the scatter measured on a decayed storage order (just before a re-sort)
};

// Measured picks by problem shape (cells, bins, pins of the region), shared by every context of the
// process so that a shape is measured once.
struct TuneKey {
    int32_t n, cnt_x, cnt_y, pins;
    bool operator<(const TuneKey& o) const {
        return std::tie(n, cnt_x, cnt_y, pins) < std::tie(o.n, o.cnt_x, o.cnt_y, o.pins);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, Tune> g_tune_cache;

// Device time per kernel family (profiling mode).
enum Fam { F_UPDATE, F_SCATTER, F_BINS, F_DCT, F_WA, F_GRAD, F_REDUCE, F_COUNT };
const char* kFamName[F_COUNT] = {"update", "scatter", "bin_density", "dct_solve", "wa_forward", "gradient", "reduce"};

}  // namespace

struct gpl_ctx {
    hipStream_t stream = nullptr;
    // A second stream (default priority) for the bins and the Poisson solve of a trial, which run
    // beside the WA kernels; the gradient waits on ev_join.  Off while profiling, so that each
    // kernel family's events bracket that family alone.
    // the objective entries' buffers (allocated on first use)
    float2* obj_xy = nullptr;
    float2* obj_grad = nullptr;
    double* obj_out = nullptr;
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool overlap = true, join_pending = false;  // overlap: launch_mode < 2
    // Launch mode: 0 a back-tracking trial as one HIP graph (side streams inside), 1 eager launches
    // on the side streams, 2 eager launches on one stream.  Profiling runs as 2.
    int launch_mode = 0;
    hipGraph_t tgraph = nullptr;
    hipGraphExec_t texec = nullptr;
    uint64_t tgraph_key = ~0ull;
    std::vector<gplk::TrialNode> tnodes;
    // Grouped trials (launch mode 7): the trial's kernels recorded, then launched grouped over the
    // regions (one launch per kernel variant and step), the WA kernels on the side stream; the
    // arguments' pinned staging and its device copy.
    bool grouped_trials = false;
    gplk::TrialRecorder grec;
    unsigned char* gst_h = nullptr;
    unsigned char* gst_d = nullptr;
    size_t gst_cap = 0;
    // Device-driven trials (launch mode 4): the trial graph ends in k_control, which takes the
    // loop's decisions on the device from the control record d_ctl, so that trials run back to
    // back (ctl_depth graphs in flight) and the host steps in only when the device halts on an
    // event it handles.  h_ctl is the record's pinned staging copy.
    bool device_trials = false, ctl_failed = false;
    gplk::RudyBufs rudy{};  // routability's congestion measurement (grown on demand)
    uint8_t* d_cap_red = nullptr;
    size_t cap_red_cap = 0;
    gplk::CtlGlobal* d_ctl = nullptr;
    gplk::CtlGlobal* h_ctl = nullptr;
    hipGraph_t cgraph = nullptr;
    hipGraphExec_t cexec = nullptr;
    int32_t ctl_depth = 2;        // trials in flight
    int32_t ctl_launch = 0;       // 0 graphs, 1 kernels one by one on the side streams, 2 on one stream
    bool ctl_ready = false;
    std::vector<gplk::RegionDev> ctl_dv;
    gplk::NetsDev ctl_nn{};
    int64_t ctl_graphs = 0, ctl_halts = 0;
    double ctl_host_ms = 0.0;
    // a third stream for the high-fanout nets' WA kernel, beside the small nets' one
    hipStream_t aux2 = nullptr;
    hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
    bool join2_pending = false;
    std::vector<void*> allocs;
    int32_t ncells = 0, npins = 0, nnets = 0;
    gplk::CellsDev cells{};
    gplk::NetsDev nets{};
    std::vector<Region> rg;
    gplk::Sums* d_sums = nullptr;  // one per region
    gplk::Sums* h_sums = nullptr;  // pinned host copy
    // the trial graph's completion signal: its last k_finish bumps d_seq and writes the value to
    // the pinned h_seq after the statistics; the host spins on h_seq instead of sleeping in a
    // stream synchronisation
    unsigned long long* d_seq = nullptr;
    unsigned long long* h_seq = nullptr;
    unsigned long long seq_want = 0;
    int signal_region = -1;  // the region whose gradient closes the trial graph
    unsigned long long* d_hpwl = nullptr;
    unsigned long long* d_global_adds = nullptr;  // the scatter's atomic adds into the HBM bins
    uint8_t* d_flags = nullptr;
    std::vector<uint8_t> h_flags;
    std::vector<int32_t> net_sorted_of;  // original net -> sorted index
    std::vector<int32_t> net_orig_of;    // sorted index -> original net
    int32_t* d_net_orig = nullptr;       // the same on the device (net re-sorts permute it)
    bool net_map_stale = false;          // the host maps lag the device's order
    int32_t* d_tile = nullptr;           // the WA tiles' first nets (nets.tile)
    gplk::NetResortBufs nrb{};
    std::vector<int32_t> cell_inst, inst_box;
    std::vector<int32_t> cell_old_of, cell_new_of;  // device cell order <-> problem order
    std::vector<int32_t> cell_inst_prob;            // placer instance of each problem cell
    bool host_map_stale = false;
    // spatial re-sort of the device cells
    gplk::CellsDev cells_alt{};
    gplk::ResortBufs rb{};
    gplk::RegionTable rtab{};
    std::vector<std::vector<float2*>> region_scratch;  // per region, kGatherMax vectors
    int32_t net_pin_count = 0;
    bool resort_enabled = true;
    uint64_t last_global_adds = 0;
    double adds_base = 0.0;  // global adds per trial in the first iteration after the last sort
    int32_t last_resort_iter = -1000, resorts = 0;
    int32_t resort_gap = 50;       // the re-sort policy (gpl_set_resort_policy)
    float resort_growth = 1.5f;
    int32_t site_x = 0, dbu = 1000;
    gpl_opts o{};
    // NesterovPlace
    float wl_coef_x = 0, wl_coef_y = 0, base_wl_coef = 0;
    int32_t recursion_wl = 0, recursion_init_slp = 0;
    float initial_prev_coordi_update_coef = 100.0f, reference_hpwl = 0;
    int32_t max_iter = 0;
    float average_overflow = 0, average_overflow_unscaled = 0;
    int64_t min_hpwl = INT64_MAX;
    bool is_min_hpwl = false;
    int32_t diverge_snapshot_iter = 0;
    float diverge_snapshot_overflow = 0;
    int32_t num_region_diverged = 0, diverge_code = 0;
    std::string diverge_msg;
    // LoopSnapshots (nesterov.rs:4428), routability off
    bool diverge_saved = false;
    float diverge_coef_x = 0, diverge_coef_y = 0;
    int64_t diverge_hpwl = 0;
    int32_t diverge_reverts = 0;
    // Routability-driven placement (opts.routability): RouteBase (route_base.rs:279-326) and the
    // Nesterov placer's routability state.  Cells routability takes out of a region (fillers) are
    // parked after the region's live cells by a re-sort (kCellParked), their per-cell slots kept in
    // the region's store, so that every kernel covers the live cells alone.
    struct Route {
        bool on = false;
        gplk::RudyGrid grid{};
        gplk::RouteVarsDev vars{};
        int32_t revert_count = 0;  // RouteBase::revert_count: the measurements so far
        float min_rc = 1e30f;
        int32_t min_rc_violated = 0;
        std::vector<float> min_rc_target_density;
        std::vector<int64_t> inflated_delta, accumulated, min_rc_inflated;
        int64_t original_movable_area = 0;
        std::vector<float> inflation_ratio;  // per tile (y-major): the last measurement's ratios
        std::vector<float> inflated_ratio;   // per tile: the pass's, trimmed to the budget
        std::vector<float> net_weight;       // per problem net: RouteBase::net_weights (0: not in the map)
        std::vector<float> min_rc_net_weight;
        int32_t weighted = 0;
        std::vector<int32_t> size;           // (dx, dy) per problem cell: the cell boxes now
        std::vector<int32_t> min_rc_size;    // per problem cell at the minimum congestion ((0, 0) until kept)
        bool need = true;                    // is_routability_need
        int32_t settle_start = -1;           // routability_settle_wait_start_iter
        // LoopSnapshots' routability half (nesterov.rs:4428-4437)
        bool saved = false;
        float coef_x = 0.0f, coef_y = 0.0f, a = 1.0f;
        int32_t attempts = 0;
        int32_t route_iters = 0, reverts = 0;
        float final_rc = 0.0f;
        int32_t last_overflowed = 0, last_actionable = 0;
        std::vector<gpl_route_pass> passes;
        // per region: its fillers in the reference's cell order (taken out from the back), the
        // fillers taken out since the last minimum (in removal order), the filler area, the
        // parked cells now on the device, the store, the first filler's problem index
        std::vector<std::vector<int32_t>> fillers, removed;
        std::vector<std::vector<uint8_t>> parked;  // per region, per filler: parked on the device
        std::vector<int64_t> total_filler_area;
        std::vector<float2*> store;
        std::vector<int32_t> filler_base, num_fillers, n_alloc;
        std::vector<long long*> np_raw, np_union;
        int2* d_sizes = nullptr;     // N: a region's cell sizes in device order
        int32_t* d_list = nullptr;   // 3 N: index lists of the parking kernels
        float* d_ratio = nullptr;    // tiles: the inflated ratios
        float* d_demand = nullptr;   // E: the nets' demand in congested tiles
        unsigned long long* d_delta = nullptr;  // the inflation's added area
        std::vector<float> demand;   // per problem net: its demand in the measured congested tiles
    } rt;
    std::vector<uint8_t> cell_flags_prob;  // GPL_CELL_* per problem cell, as constructed
    // bookkeeping
    int64_t cur_hpwl = 0;
    int32_t trials = 0, total_trials = 0, last_iter = -1;
    bool converged = false, initialized = false;
    double t_init = 0, t_loop = 0, t_converge = 0;
    std::vector<gpl_iter_row> hist;
    gpl::Log log;
    std::string err;
    int32_t inject_kind = 0, inject_after = -1, iters_run = 0;
    // host-forced loop-control cases (gpl_debug_inject kinds 5 and 6): the WA gradient sum read as
    // zero once in iteration inject_zero_wl_at (the coefficient-halving recursion), a divergence
    // reported in iterations inject_diverge_at / inject_diverge_at2 (the min-HPWL snapshot revert);
    // -1: off.  iter_now is the iteration being run.
    int32_t inject_zero_wl_at = -1, inject_diverge_at = -1, inject_diverge_at2 = -1, iter_now = -1;
    // the autotuner: 1 measure at the loop's init and again after every re-sort, 0 keep the current
    // choices (the defaults, or those gpl_set_tuning forced)
    int autotune = 1;
    std::vector<Tune> tune;
    double tune_ms = 0.0;
    int32_t tune_runs = 0;
    // profiling
    bool profiling = false;
    hipEvent_t ev[64];
    int nev = 0;
    int ev_fam[32];
    double fam_ms[F_COUNT] = {};
    int fam_n[F_COUNT] = {};
    // host time of the graph-launched trials (ms): node arguments, the launch call, the wait, and
    // the host's work between one trial's return and the next's start
    enum { H_PARAMS, H_LAUNCH, H_WAIT, H_BETWEEN, H_COUNT };
    bool finish_to_host = false;  // k_finish also writes the pinned host statistics (graph capture)
    double host_ms[H_COUNT] = {};
    int host_n = 0;
    std::chrono::steady_clock::time_point host_last{};
    bool host_have_last = false;

    template <typename T>
    T* dalloc(size_t n) {
        void* p = nullptr;
        if (n == 0) n = 1;
        check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
        allocs.push_back(p);
        return (T*)p;
    }
    template <typename T>
    // a device copy of n host values; waits, so the source may go out of scope on return
    T* upload(const T* h, size_t n) {
        T* d = dalloc<T>(std::max<size_t>(n, 1));
        if (n) {
            check(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, stream), "upload");
            check(hipStreamSynchronize(stream), "sync");
        }
        return d;
    }
    ~gpl_ctx() {
        if (stream) hipStreamSynchronize(stream);
        if (aux) hipStreamSynchronize(aux);
        for (void* p : allocs) hipFree(p);
        if (gst_h) hipHostFree(gst_h);
        if (gst_d) hipFree(gst_d);
        if (h_sums) hipHostFree(h_sums);
        if (h_seq) hipHostFree(h_seq);
        if (h_ctl) hipHostFree(h_ctl);
        rudy.release();
        if (d_cap_red) hipFree(d_cap_red);
        if (cexec) hipGraphExecDestroy(cexec);
        if (cgraph) hipGraphDestroy(cgraph);
        if (profiling)
            for (int i = 0; i < 64; ++i) hipEventDestroy(ev[i]);
        if (aux) hipStreamSynchronize(aux);
        if (aux2) hipStreamSynchronize(aux2);
        if (texec) hipGraphExecDestroy(texec);
        if (tgraph) hipGraphDestroy(tgraph);
        if (ev_fork) hipEventDestroy(ev_fork);
        if (ev_join) hipEventDestroy(ev_join);
        if (ev_fork2) hipEventDestroy(ev_fork2);
        if (ev_join2) hipEventDestroy(ev_join2);
        if (aux) hipStreamDestroy(aux);
        if (aux2) hipStreamDestroy(aux2);
        if (stream) hipStreamDestroy(stream);
    }

    // ---- profiling: events bracket each kernel family; times summed at each sync ----
    void mark(int fam) {
        if (!profiling || nev >= 62) return;
        check(hipEventRecord(ev[nev], stream), "event");
        ev_fam[nev / 2] = fam;
        ++nev;
    }
    void mark_end() {
        if (!profiling || nev % 2 == 0) return;
        check(hipEventRecord(ev[nev], stream), "event");
        ++nev;
    }
    void collect() {
        if (!profiling) return;
        for (int i = 0; i + 1 < nev; i += 2) {
            float ms = 0;
            hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            fam_ms[ev_fam[i / 2]] += ms;
            fam_n[ev_fam[i / 2]] += 1;
        }
        nev = 0;
    }

    void sync_stats() {
        join();
        check(hipMemcpyAsync(h_sums, d_sums, sizeof(gplk::Sums) * (rg.size() + 1), hipMemcpyDeviceToHost, stream), "stats copy");
        check(hipStreamSynchronize(stream), "sync");
        collect();
    }

    // ---- the reference's steps ------------------------------------------------------------
    int64_t nesterov_insts_area(const Region& r) const {
        return r.R.std_insts_area + (int64_t)std::round(float(r.R.macro_insts_area) * r.R.target_density);
    }

    // a kernel, not a memset, so that a captured trial graph holds no memset node
    void zero_hpwl() {
        join2();  // no high-fanout WA kernel may still be adding into the counter
        gplk::launch_set_u64(stream, d_hpwl, 0ull);
    }

    // getHpwl (nesterov.rs:2297), read back.
    int64_t hpwl_now() {
        zero_hpwl();
        gplk::launch_hpwl(stream, nets, cells, d_hpwl);
        unsigned long long h = 0;
        check(hipMemcpyAsync(&h, d_hpwl, sizeof h, hipMemcpyDeviceToHost, stream), "hpwl copy");
        check(hipStreamSynchronize(stream), "sync");
        return (int64_t)h;
    }

    // the bin densities and their overflow sums (finished here, or with the gradient's sums when
    // `finish` is false: a trial reads both only at its end)
    void bins_and_finish(Region& r, int b, bool finish = true) {
        mark(F_BINS);
        gplk::launch_bin_density(stream, r.dev, r.partials);
        if (finish) gplk::launch_finish(stream, r.partials, gplk::block_partials_bins(r.dev), 0, d_sums + b, nullptr);
        mark_end();
    }

    // the bins' per-bin pass and the solve (the overflow sums finished here, or with the
    // gradient's sums when `finish` is false)
    void bins_and_solve(Region& r, int b, bool finish = true, bool want_phi = false) {
        bins_and_finish(r, b, finish);
        solve(r, want_phi);
    }

    void solve(Region& r, bool want_phi = false) {
        mark(F_DCT);
        gplk::launch_solve(stream, r.dev, r.sb, want_phi);
        mark_end();
    }

    // updateWireLengthForceWA with the pins' gradient terms (read by the gradient kernel); the
    // net records only for the kernel-level entry.
    void wa_forward(bool write_rec = false, bool zeroed = false) {
        mark(F_WA);
        if (zeroed) join2();  // (zero_hpwl joins too)
        else zero_hpwl();
        const bool ovl = overlap && !profiling && nets.num_small < nnets;
        if (ovl) {
            check(hipEventRecord(ev_fork2, stream), "event");
            check(hipStreamWaitEvent(aux2, ev_fork2, 0), "wait");
        }
        gplk::launch_wa_forward(stream, ovl ? aux2 : stream, nets, cells, wl_coef_x, wl_coef_y, d_hpwl, write_rec);
        if (ovl) {
            check(hipEventRecord(ev_join2, aux2), "event");
            join2_pending = true;
        }
        mark_end();
    }

    // the main stream waits for the high-fanout nets' WA kernel
    void join2() {
        if (!join2_pending) return;
        check(hipStreamWaitEvent(stream, ev_join2, 0), "wait");
        join2_pending = false;
    }

    // The gradient of region b into `slot`; with `trial`, the distances of the step length too, and
    // the bins' overflow sums finished together with the gradient's (field(false) left them).  The
    // WA and density parts are stored apart only for the kernel-level entries (`parts`): the loop
    // reads the preconditioned sum alone.
    void launch_gradient(int b, int slot, bool trial, int nan_inject = 0, bool parts = false) {
        Region& r = rg[b];
        int base = slot == SLOT_PREV ? S_PREV_SLP : slot == SLOT_CUR ? S_CUR_SLP : S_NEXT_SLP;
        mark(F_GRAD);
        gplk::launch_gradient(stream, r.dev, cells, nets, wl_coef_x, wl_coef_y, r.density_penalty, parts ? r.v[base + 1] : nullptr,
                              parts ? r.v[base + 2] : nullptr,
                              r.v[base + 3], trial ? r.v[S_CUR_SLP] : nullptr, trial ? r.v[S_NEXT_SLP] : nullptr,
                              trial ? r.v[S_CUR_SLP_SUM] : nullptr, r.partials, nan_inject);
        const bool sig = finish_to_host && b == signal_region;
        gplk::launch_finish(stream, r.partials, trial ? gplk::block_partials_bins(r.dev) : 0, gplk::block_partials_cells(r.dev), d_sums + b,
                            d_hpwl, finish_to_host ? h_sums + b : nullptr, finish_to_host ? d_sums + rg.size() : nullptr,
                            finish_to_host ? h_sums + rg.size() : nullptr, sig ? d_seq : nullptr, sig ? h_seq : nullptr);
        mark_end();
    }

    bool inject_now() const { return inject_kind == 1 && inject_after >= 0 && iters_run >= inject_after; }

    // updateGradients' sums read back; npUpdate*Gradient's zero-sum recursion and NaN check
    // (nesterov.rs:2512-2528).  Regions from `first` are (re)evaluated in order.
    void np_update_gradients(int first, int last, int slot, bool trial) {
        join();
        launch_gradients(first, last, slot, trial);
        sync_stats();
        gradients_read(first, last, slot, trial);
    }

    void launch_gradients(int first, int last, int slot, bool trial) {
        for (int b = first; b < last; ++b) {
            if (rg[b].is_converged) continue;
            launch_gradient(b, slot, trial, slot == SLOT_NEXT && inject_now());
        }
    }

    // the host side of np_update_gradients once the statistics are back
    void gradients_read(int first, int last, int slot, bool trial) {
        for (int b = first; b < last; ++b) {
            Region& r = rg[b];
            if (!r.is_converged) {
                r.wl_grad_sum = (float)h_sums[b].wl_abs;
                r.density_grad_sum = (float)h_sums[b].den_abs;
                if (inject_zero_wl_at >= 0 && iter_now == inject_zero_wl_at && slot == SLOT_NEXT) {
                    r.wl_grad_sum = 0.0f;
                    inject_zero_wl_at = -1;
                }
            }
            if (r.wl_grad_sum == 0.0f && recursion_wl < 10) {
                wl_coef_x = (float)((double)wl_coef_x * 0.5);
                wl_coef_y = (float)((double)wl_coef_y * 0.5);
                base_wl_coef = (float)((double)base_wl_coef * 0.5);
                wa_forward();
                recursion_wl += 1;
                np_update_gradients(b, last, slot, trial);
                return;
            }
            auto bad = [](float v) { return std::isnan(v) || std::isinf(v); };
            if (bad(r.wl_grad_sum) || bad(r.density_grad_sum)) {
                diverge_code = 306;
                diverge_msg = "RePlAce diverged at wire/density gradient Sum. An internal value is NaN or Inf.";
                num_region_diverged = 1;
            }
        }
    }

    void update_wirelength_coef(float overflow) {
        float c;
        if (overflow > 1.0f)
            c = (float)0.1;
        else if ((double)overflow < 0.1)
            c = 10.0f;
        else
            c = (float)(1.0 / std::pow(10.0, ((double)overflow - 0.1) * 20.0 / 9.0 - 1.0));
        wl_coef_x = c * base_wl_coef;
        wl_coef_y = c * base_wl_coef;
    }

    // initDensity1 (nesterov.rs:2546).
    void init_density1(int b) {
        Region& r = rg[b];
        mark(F_SCATTER);
        gplk::launch_center_init(stream, r.dev, cells, r.v[S_CUR_SLP], r.v[S_PREV_SLP], r.v[S_CUR], d_global_adds);
        mark_end();
        bins_and_solve(r, b);
        r.prev_hpwl = hpwl_now();
        sync_stats();
        r.overflow = (int64_t)h_sums[b].overflow;
        r.overflow_unscaled = (int64_t)h_sums[b].overflow_unscaled;
        r.base_wl_coef_nb = (float)((double)o.init_wirelength_coef / ((double)(float)(r.R.size_x + r.R.size_y) * 0.5));
        float nia = (float)r.nia;
        r.sum_overflow = (float)r.overflow / nia;
        r.sum_overflow_unscaled = (float)r.overflow_unscaled / nia;
    }

    // getDistance of two vector pairs (nesterov.rs:319), FP64 sums.
    std::pair<float, float> distances(int b, const float2* a, const float2* bb, const float2* c, const float2* d) {
        Region& r = rg[b];
        mark(F_REDUCE);
        gplk::launch_distance(stream, r.n, a, bb, c, d, r.partials);
        gplk::launch_finish(stream, r.partials, 0, gplk::block_partials_cells(r.dev), d_sums + b, nullptr);
        mark_end();
        sync_stats();
        if (r.n == 0) return {0.0f, 0.0f};
        float x = (float)std::sqrt(h_sums[b].coord_sq / (2.0 * (double)r.n));
        float y = (float)std::sqrt(h_sums[b].grad_sq / (2.0 * (double)r.n));
        return {x, y};
    }

    // initDensity2 (nesterov.rs:2600).
    float init_density2(int b) {
        Region& r = rg[b];
        if (r.wl_grad_sum == 0.0f) {
            r.density_penalty = o.init_density_penalty;
            launch_gradient(b, SLOT_PREV, false);
            sync_stats();
            r.wl_grad_sum = (float)h_sums[b].wl_abs;
            r.density_grad_sum = (float)h_sums[b].den_abs;
        }
        if (r.wl_grad_sum != 0.0f) r.density_penalty = (r.wl_grad_sum / r.density_grad_sum) * o.init_density_penalty;
        float nia = (float)r.nia;
        r.sum_overflow = (float)r.overflow / nia;
        r.sum_overflow_unscaled = (float)r.overflow_unscaled / nia;
        auto dd = distances(b, r.v[S_PREV_SLP], r.v[S_CUR_SLP], r.v[S_PREV_SLP_SUM], r.v[S_CUR_SLP_SUM]);
        r.coordi_distance = dd.first;
        r.step_length = r.coordi_distance / dd.second;
        return r.step_length;
    }

    // NesterovPlace::init (nesterov.rs:2674).
    int32_t np_init() {
        float total_sum_overflow = 0.0f, total_base = 0.0f;
        for (size_t b = 0; b < rg.size(); ++b) {
            init_density1(int(b));
            total_sum_overflow += rg[b].sum_overflow;
            total_base += rg[b].base_wl_coef_nb;
        }
        float n = (float)rg.size();
        average_overflow = total_sum_overflow / n;
        base_wl_coef = total_base / n;
        update_wirelength_coef(average_overflow);
        wa_forward();
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            np_update_gradients(int(b), int(b) + 1, SLOT_CUR, false);
            // only region b's gradient is wanted here; later regions are recomputed below
            mark(F_SCATTER);
            gplk::launch_initial_prev(stream, r.dev, cells, r.v[S_CUR_SLP], r.v[S_CUR_SLP_SUM], initial_prev_coordi_update_coef, r.v[S_PREV_SLP]);
            gplk::launch_center_scatter(stream, r.dev, cells, r.v[S_PREV_SLP], d_global_adds);
            mark_end();
            bins_and_solve(r, int(b));
            sync_stats();
            r.overflow = (int64_t)h_sums[b].overflow;
            r.overflow_unscaled = (int64_t)h_sums[b].overflow_unscaled;
        }
        wa_forward();
        np_update_gradients(0, int(rg.size()), SLOT_PREV, false);
        for (size_t b = 0; b < rg.size(); ++b) {
            float step = init_density2(int(b));
            if ((std::isnan(step) || std::isinf(step)) && recursion_init_slp < 10) {
                initial_prev_coordi_update_coef *= 10.0f;
                recursion_init_slp += 1;
                return np_init();
            }
            if (std::isnan(step) || std::isinf(step)) {
                log.error(304, strfmt("RePlAce diverged at initial iteration with steplength being %g. Re-run with a smaller init_density_penalty value.", step));
                return 304;
            }
        }
        return GPL_OK;
    }

    // The first live region's scatter clears the HPWL counter for the trial's WA kernels.
    unsigned long long* zero_for(size_t b) const {
        for (size_t q = 0; q < rg.size(); ++q)
            if (!rg[q].is_converged) return q == b ? d_hpwl : nullptr;
        return nullptr;
    }

    // nesterovUpdateCoordinates for every region (new points, boxes and centres); field() then
    // scatters the boxes and solves on the side stream while wa_forward(false, true) runs the WA
    // kernels on the centres (the HPWL counter cleared here).
    void update_coordinates(float coeff) {
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            if (r.is_converged) continue;
            mark(F_UPDATE);
            gplk::launch_update(stream, r.dev, cells, r.v[S_CUR_SLP], r.v[S_CUR_SLP_SUM], r.v[S_CUR], r.v[S_NEXT], r.v[S_NEXT_SLP],
                                r.step_length, coeff, zero_for(b));
            mark_end();
        }
        if (overlap && !profiling) check(hipEventRecord(ev_fork, stream), "event");
    }

    // The bins and the Poisson solve of every region after update_coordinates: on the second
    // stream when overlapping (beside the WA kernels), joined before the gradient.
    void field() {
        const bool ovl = overlap && !profiling;
        if (ovl) {
            check(hipStreamWaitEvent(aux, ev_fork, 0), "wait");
            std::swap(stream, aux);
        }
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            if (r.is_converged) continue;
            mark(F_SCATTER);
            gplk::launch_density_scatter(stream, r.dev, cells, d_global_adds);
            mark_end();
            // the overflow sums finished with the trial's gradient sums; while profiling the bins
            // pass and the solve run apart (one family each), else fused (the same bits)
            if (profiling) bins_and_solve(r, int(b), false);
            else gplk::launch_bins_solve(stream, r.dev, r.sb, r.partials);
        }
        if (ovl) {
            std::swap(stream, aux);
            check(hipEventRecord(ev_join, aux), "event");
            join_pending = true;
        }
    }

    // the main stream waits for the second stream's bins and solve
    void join() {
        join2();
        if (!join_pending) return;
        check(hipStreamWaitEvent(stream, ev_join, 0), "wait");
        join_pending = false;
    }

    // nesterovUpdateStepLength (nesterov.rs:2750) from the trial's FP64 distance sums.
    bool update_step_length(int b) {
        Region& r = rg[b];
        if (r.is_converged) return true;
        double n2 = 2.0 * (double)r.n;
        r.coordi_distance = r.n ? (float)std::sqrt(h_sums[b].coord_sq / n2) : 0.0f;
        float grad = r.n ? (float)std::sqrt(h_sums[b].grad_sq / n2) : 0.0f;
        float nw = r.coordi_distance / grad;
        if (inject_kind == 2 && inject_after >= 0 && iters_run >= inject_after) nw = NAN;
        if (std::isnan(nw) || std::isinf(nw)) {
            r.is_diverged = true;
            return false;
        }
        if ((double)nw > (double)r.step_length * 0.95) {
            r.step_length = nw;
            return false;
        }
        if ((double)nw < 0.01) {
            r.step_length = (float)0.01;
            return false;
        }
        r.step_length = nw;
        return true;
    }

    // doBackTracking (nesterov.rs:2776).
    // ---- one back-tracking trial as a HIP graph ----------------------------------------------
    // The trial (scatter; bins and solve on the second stream beside the WA kernels; gradients,
    // finishing reductions, the statistics copy) is captured once per set of live regions; before
    // each launch the nodes whose arguments change (step, momentum, coefficients, penalty, the
    // rotated slots, the cell arrays a re-sort swapped) get their new arguments.  One launch and
    // one wait per trial instead of ~20 API calls.
    // Waits for the trial graph's completion signal (spinning: the statistics are then complete in
    // pinned memory, and the stream's later work queues behind the graph as usual); after 2 s
    // without it, a stream synchronisation reports whatever went wrong.
    void wait_trial_signal() { wait_seq(++seq_want); }

    void wait_seq(unsigned long long want) {
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t k = 0;; ++k) {
            if (*reinterpret_cast<volatile unsigned long long*>(h_seq) >= want) return;
            if ((k & 1023u) == 1023u && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) break;
        }
        check(hipStreamSynchronize(stream), "sync");
        if (*reinterpret_cast<volatile unsigned long long*>(h_seq) < want) throw HipError{hipErrorUnknown, "trial signal"};
    }

    // ---- device-driven trials ----------------------------------------------------------------
    // Whether the next iteration's trials may run on the device: launch mode 4, no profiling, no
    // live region with locked cells (their next slots are fixed by a host-launched kernel) and few
    // enough regions for the control record.
    bool ctl_usable() const {
        if (!device_trials || ctl_failed || launch_mode != 0 || profiling || rt.on) return false;
        if (rg.size() > size_t(gplk::kCtlMaxRegions)) return false;
        for (auto& r : rg)
            if (!r.is_converged && r.has_locked) return false;
        return true;
    }

    // The device-driven trial: every region's update, scatter, bins and solve (second stream),
    // the WA kernels (third stream for the high-fanout nets), gradients and finishing reductions,
    // the last of which runs the control.  The kernels read their step, penalty, coefficients and
    // slots from the control record, so the same launches serve every trial: captured once as a
    // graph (launch mode 4), or launched one by one on the side streams (5) or on one stream (6).  Their arguments (ctl_dv, ctl_nn) are taken again after a re-sort or
    // a new autotuner pick (the cell and net arrays, the kernel variants).
    void ctl_prepare() {
        ctl_nn = nets;
        ctl_nn.cg = &d_ctl->h;
        ctl_dv.resize(rg.size());
        for (size_t b = 0; b < rg.size(); ++b) {
            ctl_dv[b] = rg[b].dev;
            ctl_dv[b].ctl = &d_ctl->h.reg[b];
            ctl_dv[b].cg = &d_ctl->h;
        }
        ctl_ready = true;
    }

    void enqueue_ctl_trial(bool side) {
        const size_t R = rg.size();
        hipStream_t s1 = side ? aux : stream;
        for (size_t b = 0; b < R; ++b) {
            Region& r = rg[b];
            gplk::launch_update(stream, ctl_dv[b], cells, r.v[S_CUR_SLP], r.v[S_CUR_SLP_SUM], r.v[S_CUR], r.v[S_NEXT], r.v[S_NEXT_SLP],
                                0.0f, 0.0f, nullptr);
        }
        if (side) {
            check(hipEventRecord(ev_fork, stream), "event");
            check(hipStreamWaitEvent(aux, ev_fork, 0), "wait");
        }
        for (size_t b = 0; b < R; ++b) {
            Region& r = rg[b];
            gplk::launch_density_scatter(s1, ctl_dv[b], cells, d_global_adds);
            gplk::launch_bins_solve(s1, ctl_dv[b], r.sb, r.partials, &d_ctl->h.reg[b].hot.flags);
        }
        if (side) check(hipEventRecord(ev_join, aux), "event");
        const bool big = side && nets.num_small < nnets;
        if (big) {
            check(hipEventRecord(ev_fork2, stream), "event");
            check(hipStreamWaitEvent(aux2, ev_fork2, 0), "wait");
        }
        gplk::launch_wa_forward(stream, big ? aux2 : stream, ctl_nn, cells, 0.0f, 0.0f, d_hpwl, false);
        if (big) {
            check(hipEventRecord(ev_join2, aux2), "event");
            check(hipStreamWaitEvent(stream, ev_join2, 0), "wait");
        }
        if (side) check(hipStreamWaitEvent(stream, ev_join, 0), "wait");
        gplk::CtlArgs ca{d_ctl, d_sums, d_sums + R, d_seq, h_seq, int32_t(R)};
        for (size_t b = 0; b < R; ++b) {
            Region& r = rg[b];
            gplk::launch_gradient(stream, ctl_dv[b], cells, ctl_nn, 0.0f, 0.0f, 0.0f, nullptr, nullptr, r.v[S_NEXT_SLP_SUM],
                                  r.v[S_CUR_SLP], r.v[S_NEXT_SLP], r.v[S_CUR_SLP_SUM], r.partials, 0);
            gplk::launch_finish(stream, r.partials, gplk::block_partials_bins(r.dev), gplk::block_partials_cells(r.dev), d_sums + b, d_hpwl,
                                nullptr, nullptr, nullptr, nullptr, nullptr, &d_ctl->h.reg[b].hot.flags, b + 1 == R ? ca : gplk::CtlArgs{});
        }
    }

    bool build_ctl_graph() {
        drop_ctl_graph();
        join();
        check(hipStreamSynchronize(stream), "sync");
        ctl_prepare();
        check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
        hipError_t err = hipSuccess;
        try {
            enqueue_ctl_trial(true);
        } catch (const HipError& e) {
            err = e.e;
        }
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(stream, &g);
        if (err != hipSuccess || ec != hipSuccess || !g) {
            if (g) hipGraphDestroy(g);
            return false;
        }
        cgraph = g;
        if (hipGraphInstantiate(&cexec, cgraph, nullptr, nullptr, 0) != hipSuccess) {
            drop_ctl_graph();
            return false;
        }
        return true;
    }

    // The host's loop state into the control record (pinned staging copy, then an asynchronous
    // copy on the stream): the point sets as rotation 0, the iteration `ls.iter` about to start with
    // momentum coefficient `coeff`, at most up to iteration `stop_at`.
    void ctl_upload(float coeff, int32_t stop_at) {
        gplk::CtlHead& g = h_ctl->h;
        g.halt = 0;
        g.stage = 0;
        g.nreg = int32_t(rg.size());
        g.iter = ls.iter;
        g.iter_stop = stop_at;
        g.max_iter = max_iter;
        g.trials = 0;
        g.total_trials = total_trials;
        g.recursion_wl = recursion_wl;
        g.disable_revert = o.disable_revert_if_diverge ? 1 : 0;
        g.wl_coef_x = wl_coef_x;
        g.wl_coef_y = wl_coef_y;
        g.base_wl_coef = base_wl_coef;
        g.cur_a = ls.cur_a;
        g.coeff = coeff;
        g.reference_hpwl = reference_hpwl;
        g.average_overflow = average_overflow;
        g.average_overflow_unscaled = average_overflow_unscaled;
        g.overflow_target = o.overflow;
        g.min_phi = o.min_phi_coef;
        g.diverge_snapshot_overflow = diverge_snapshot_overflow;
        g.is_min_hpwl = is_min_hpwl ? 1 : 0;
        g.diverge_snapshot_iter = diverge_snapshot_iter;
        g.resort_enabled = resort_enabled ? 1 : 0;
        g.last_resort_iter = last_resort_iter;
        g.iters_run = iters_run;
        g.min_hpwl = min_hpwl;
        g.cur_hpwl = cur_hpwl;
        g.last_global_adds = last_global_adds;
        g.adds_base = adds_base;
        g.nhist = 0;
        g.nlog = 0;
        const bool after = inject_after >= 0;
        g.nan_from = inject_kind == 1 && after ? inject_after : -1;
        g.step_nan_from = inject_kind == 2 && after ? inject_after : -1;
        g.resort_from = inject_kind == 3 && after ? inject_after : -1;
        g.zero_wl_at = inject_zero_wl_at;
        g.resort_gap = resort_gap;
        g.resort_growth = resort_growth;
        g.diverge_at = inject_diverge_at;
        g.diverge_at2 = inject_diverge_at2;
        for (size_t b = 0; b < rg.size(); ++b) {
            const Region& r = rg[b];
            gplk::CtlRegion& q = g.reg[b];
            for (int k = 0; k < 4; ++k) {
                q.sets[0][k] = r.v[S_CUR_SLP + k];
                q.sets[1][k] = r.v[S_NEXT_SLP + k];
                q.sets[2][k] = r.v[S_PREV_SLP + k];
            }
            q.cn[0] = r.v[S_CUR];
            q.cn[1] = r.v[S_NEXT];
            q.rot = 0;
            q.cnp = 0;
            q.is_converged = r.is_converged ? 1 : 0;
            q.max_phi_coef_changed = r.max_phi_coef_changed ? 1 : 0;
            q.n = r.n;
            q.group = r.R.group;
            q.step_length = r.step_length;
            q.density_penalty = r.density_penalty;
            q.max_phi_coef = r.max_phi_coef;
            q.sum_overflow = r.sum_overflow;
            q.sum_overflow_unscaled = r.sum_overflow_unscaled;
            q.min_sum_overflow = r.min_sum_overflow;
            q.hpwl_with_min_sum_overflow = r.hpwl_with_min_sum_overflow;
            q.coordi_distance = r.coordi_distance;
            q.peak_coordi_distance = r.peak_coordi_distance;
            q.prev_reported_overflow_unscaled = r.prev_reported_overflow_unscaled;
            q.wl_grad_sum = r.wl_grad_sum;
            q.density_grad_sum = r.density_grad_sum;
            q.prev_hpwl = r.prev_hpwl;
            q.prev_reported_hpwl = r.prev_reported_hpwl;
            q.overflow = r.overflow;
            q.overflow_unscaled = r.overflow_unscaled;
            q.nia = r.nia;
            q.non_place_insts_area = r.R.non_place_insts_area;
        }
        gplk::ctl_refresh(g);
        check(hipMemcpyAsync(d_ctl, h_ctl, sizeof(gplk::CtlHead), hipMemcpyHostToDevice, stream), "control upload");
    }

    // The control record back into the host's state after a halt: the progress rows logged and the
    // history rows kept in order, every scalar, and each region's point sets in their rotation.
    void ctl_apply() {
        const gplk::CtlHead& g = h_ctl->h;
        for (int i = 0; i < g.nlog; ++i) {
            const gplk::CtlLog& L = h_ctl->log[i];
            report_row(rg[L.region], L.iter, L.overflow, L.hpwl, L.pct, L.penalty);
        }
        for (int i = 0; i < g.nhist; ++i) {
            const gplk::CtlHist& H = h_ctl->hist[i];
            gpl_iter_row row{};
            row.iter = H.iter;
            row.trials = H.trials;
            row.overflow = H.overflow;
            row.density_penalty = H.density_penalty;
            row.step_length = H.step_length;
            row.wl_coef = H.wl_coef;
            row.hpwl = H.hpwl;
            hist.push_back(row);
        }
        if (g.nhist > 0) last_iter = h_ctl->hist[g.nhist - 1].iter;
        wl_coef_x = g.wl_coef_x;
        wl_coef_y = g.wl_coef_y;
        average_overflow = g.average_overflow;
        average_overflow_unscaled = g.average_overflow_unscaled;
        diverge_snapshot_overflow = g.diverge_snapshot_overflow;
        is_min_hpwl = g.is_min_hpwl != 0;
        diverge_snapshot_iter = g.diverge_snapshot_iter;
        iters_run = g.iters_run;
        total_trials = g.total_trials;
        min_hpwl = g.min_hpwl;
        cur_hpwl = g.cur_hpwl;
        last_global_adds = g.last_global_adds;
        adds_base = g.adds_base;
        ls.iter = g.iter;
        ls.cur_a = g.cur_a;
        trials = g.stage == 1 ? g.trials : (g.nhist > 0 ? h_ctl->hist[g.nhist - 1].trials : trials);
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            const gplk::CtlRegion& q = g.reg[b];
            const int ns = q.rot == 2 ? 0 : q.rot + 1, ps = ns == 2 ? 0 : ns + 1;
            for (int k = 0; k < 4; ++k) {
                r.v[S_CUR_SLP + k] = q.sets[q.rot][k];
                r.v[S_NEXT_SLP + k] = q.sets[ns][k];
                r.v[S_PREV_SLP + k] = q.sets[ps][k];
            }
            r.v[S_CUR] = q.cn[q.cnp];
            r.v[S_NEXT] = q.cn[q.cnp ^ 1];
            r.max_phi_coef_changed = q.max_phi_coef_changed != 0;
            r.step_length = q.step_length;
            r.density_penalty = q.density_penalty;
            r.max_phi_coef = q.max_phi_coef;
            r.sum_overflow = q.sum_overflow;
            r.sum_overflow_unscaled = q.sum_overflow_unscaled;
            r.min_sum_overflow = q.min_sum_overflow;
            r.hpwl_with_min_sum_overflow = q.hpwl_with_min_sum_overflow;
            r.coordi_distance = q.coordi_distance;
            r.peak_coordi_distance = q.peak_coordi_distance;
            r.prev_reported_overflow_unscaled = q.prev_reported_overflow_unscaled;
            r.wl_grad_sum = q.wl_grad_sum;
            r.density_grad_sum = q.density_grad_sum;
            r.prev_hpwl = q.prev_hpwl;
            r.prev_reported_hpwl = q.prev_reported_hpwl;
            r.overflow = q.overflow;
            r.overflow_unscaled = q.overflow_unscaled;
        }
    }

    // Runs trials on the device from iteration ls.iter (its momentum `coeff` taken) until the device
    // halts, with ctl_depth graphs in flight; the host's state then follows the device's.  Returns
    // the halt's stage (1: the host finishes the halted trial and the iteration; 2: the host takes
    // the iteration from isDiverged; 3: the iteration ended), or 0 when the graph cannot be built
    // (the host runs the iteration).
    int run_device_trials(float coeff, int32_t stop_at) {
        if (!d_ctl) {
            d_ctl = dalloc<gplk::CtlGlobal>(1);
            check(hipHostMalloc((void**)&h_ctl, sizeof(gplk::CtlGlobal), hipHostMallocDefault), "pinned");
        }
        if (ctl_launch == 0 && !cexec && !build_ctl_graph()) {
            ctl_failed = true;
            return 0;
        }
        if (ctl_launch != 0 && !ctl_ready) ctl_prepare();
        const auto t0 = std::chrono::steady_clock::now();
        join();
        ctl_upload(coeff, stop_at);
        gplk::launch_set_u64(stream, d_hpwl, 0ull);  // the first trial's HPWL counter
        const unsigned long long base = seq_want;
        unsigned long long launched = 0, done = 0;
        bool halted = false;
        for (;;) {
            while (!halted && launched - done < (unsigned long long)ctl_depth) {
                if (ctl_launch == 0) check(hipGraphLaunch(cexec, stream), "graph launch");
                else enqueue_ctl_trial(ctl_launch == 1);
                ++launched;
            }
            if (done == launched) break;
            wait_seq(base + done + 1);
            ++done;
            if (!halted && *reinterpret_cast<volatile unsigned long long*>(h_seq + 1) != 0ull) halted = true;
        }
        seq_want = base + launched;
        ctl_graphs += (int64_t)launched;
        ++ctl_halts;
        check(hipMemcpyAsync(h_ctl, d_ctl, sizeof(gplk::CtlGlobal), hipMemcpyDeviceToHost, stream), "control download");
        sync_stats();
        ctl_apply();
        ctl_host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return h_ctl->h.stage;
    }


    void drop_graph() {
        if (texec) hipGraphExecDestroy(texec);
        if (tgraph) hipGraphDestroy(tgraph);
        texec = nullptr;
        tgraph = nullptr;
        tnodes.clear();
        tgraph_key = ~0ull;
        drop_ctl_graph();
    }

    void drop_ctl_graph() {
        if (cexec) hipGraphExecDestroy(cexec);
        if (cgraph) hipGraphDestroy(cgraph);
        cexec = nullptr;
        cgraph = nullptr;
        ctl_ready = false;
    }

    uint64_t live_regions_key() const {
        uint64_t k = 0;
        for (size_t b = 0; b < rg.size(); ++b)
            if (!rg[b].is_converged) k |= 1ull << b;
        return k;
    }

    // Captured with the side streams: the bins and the solve form a branch beside the WA kernels.
    bool build_graph() {
        drop_graph();
        check(hipStreamSynchronize(stream), "sync");
        check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "capture");
        hipError_t err = hipSuccess;
        try {
            update_coordinates(0.0f);
            field();
            wa_forward(false, true);
            join();
            // the trial's statistics go straight to the pinned host copy: the graph holds no copy node
            finish_to_host = true;
            signal_region = -1;
            for (size_t b = 0; b < rg.size(); ++b)
                if (!rg[b].is_converged) signal_region = int(b);
            launch_gradients(0, int(rg.size()), SLOT_NEXT, true);
            finish_to_host = false;
        } catch (const HipError& e) {
            err = e.e;
        }
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(stream, &g);
        join_pending = join2_pending = false;
        if (err != hipSuccess || ec != hipSuccess || !g) {
            if (g) hipGraphDestroy(g);
            return false;
        }
        tgraph = g;
        if (hipGraphInstantiate(&texec, tgraph, nullptr, nullptr, 0) != hipSuccess) {
            drop_graph();
            return false;
        }
        tnodes.resize(8 + 3 * rg.size());
        int m = gplk::find_trial_nodes(tgraph, tnodes.data(), int(tnodes.size()));
        int live = 0;
        for (auto& r : rg) live += r.is_converged ? 0 : 1;
        if (m != 3 * live + (nets.num_small > 0 ? 1 : 0) + (nets.num_small < nnets ? 1 : 0)) {
            drop_graph();
            return false;
        }
        tnodes.resize(m);
        tgraph_key = live_regions_key();
        return true;
    }

    Region* region_of(int cell_begin) {
        for (auto& r : rg)
            if (r.R.cell_begin == cell_begin) return &r;
        return nullptr;
    }

    // one trial through the graph; false when graphs are off (the caller launches eagerly)
    bool trial_by_graph(float coeff) {
        if (launch_mode != 0 || profiling) return false;
        if (!texec || tgraph_key != live_regions_key()) {
            if (!build_graph()) {
                launch_mode = 1;
                return false;
            }
        }
        using clk = std::chrono::steady_clock;
        auto ms_since = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        const clk::time_point t0 = clk::now();
        if (host_have_last) host_ms[H_BETWEEN] += ms_since(host_last, t0);
        const int nan = inject_now() ? 1 : 0;
        for (auto& n : tnodes) {
            hipError_t e = hipSuccess;
            if (n.kind == gplk::TK_UPDATE) {
                Region& r = *region_of(n.cell_begin);
                e = gplk::set_update_node(texec, n, r.dev, cells, r.v[S_CUR_SLP], r.v[S_CUR_SLP_SUM], r.v[S_CUR], r.v[S_NEXT],
                                          r.v[S_NEXT_SLP], r.step_length, coeff, zero_for(size_t(&r - rg.data())));
            } else if (n.kind == gplk::TK_SCATTER) {
                Region& r = *region_of(n.cell_begin);
                e = gplk::set_scatter_node(texec, n, r.dev, cells, d_global_adds);
            } else if (n.kind == gplk::TK_GRADIENT) {
                Region& r = *region_of(n.cell_begin);
                e = gplk::set_gradient_node(texec, n, r.dev, cells, nets, wl_coef_x, wl_coef_y, r.density_penalty, nullptr,
                                            nullptr, r.v[S_NEXT_SLP + 3], r.v[S_CUR_SLP], r.v[S_NEXT_SLP],
                                            r.v[S_CUR_SLP_SUM], r.partials, nan);
            } else {
                e = gplk::set_wa_node(texec, n, nets, cells, wl_coef_x, wl_coef_y, d_hpwl);
            }
            check(e, "graph node");
        }
        const clk::time_point t1 = clk::now();
        check(hipGraphLaunch(texec, stream), "graph launch");
        const clk::time_point t2 = clk::now();
        wait_trial_signal();
        const clk::time_point t3 = clk::now();
        gradients_read(0, int(rg.size()), SLOT_NEXT, true);
        host_ms[H_PARAMS] += ms_since(t0, t1);
        host_ms[H_LAUNCH] += ms_since(t1, t2);
        host_ms[H_WAIT] += ms_since(t2, t3);
        ++host_n;
        host_last = clk::now();
        host_have_last = true;
        return true;
    }

    // one trial recorded and launched grouped (launch mode 7); false when off or when a kernel of
    // the trial has no grouped form (the caller then launches it as usual)
    bool trial_grouped(float coeff) {
        if (!grouped_trials || profiling) return false;
        grec.recs.clear();
        grec.failed = false;
        gplk::set_recorder(&grec);
        try {
            record_trial(coeff);
        } catch (...) {
            gplk::set_recorder(nullptr);
            throw;
        }
        gplk::set_recorder(nullptr);
        if (grec.failed) return false;
        const gplk::TrialRecorder* one = &grec;
        const size_t need = gplk::grouped_arg_bytes(&one, 1);
        if (need > gst_cap) {
            check(hipStreamSynchronize(stream), "sync");
            if (gst_h) hipHostFree(gst_h);
            if (gst_d) hipFree(gst_d);
            gst_h = nullptr;
            gst_d = nullptr;
            gst_cap = 0;
            check(hipHostMalloc((void**)&gst_h, need, hipHostMallocDefault), "hipHostMalloc");
            check(hipMalloc((void**)&gst_d, need), "hipMalloc");
            gst_cap = need;
        }
        check(gplk::launch_grouped(stream, &one, 1, gst_h, gst_d, aux, ev_fork, ev_join), "grouped trial");
        wait_trial_signal();
        gradients_read(0, int(rg.size()), SLOT_NEXT, true);
        return true;
    }

    // `resume` > 0: the iteration's first resume - 1 trials ran on the device and trial `resume`
    // did too, up to its gradient sums (in d_sums, and with stats_on_host in h_sums already): the
    // host takes it from there.
    void do_back_tracking(float coeff, int resume = 0, bool stats_on_host = false) {
        trials = resume > 0 ? resume - 1 : 0;
        for (int nbt = trials; nbt < 10; ++nbt) {
            ++trials;
            num_region_diverged = 0;
            if (resume > 0 && nbt == resume - 1) {
                if (!stats_on_host) sync_stats();
                gradients_read(0, int(rg.size()), SLOT_NEXT, true);
            } else if (!trial_by_graph(coeff) && !trial_grouped(coeff)) {
                update_coordinates(coeff);
                field();
                wa_forward(false, true);
                np_update_gradients(0, int(rg.size()), SLOT_NEXT, true);
            }
            for (size_t b = 0; b < rg.size(); ++b) {
                Region& r = rg[b];
                if (!r.is_converged) {
                    r.overflow = (int64_t)h_sums[b].overflow;
                    r.overflow_unscaled = (int64_t)h_sums[b].overflow_unscaled;
                }
                num_region_diverged += r.is_diverged ? 1 : 0;
            }
            cur_hpwl = (int64_t)h_sums[rg.size()].hpwl;
            if (num_region_diverged > 0) {
                diverge_code = 306;
                diverge_msg = "RePlAce diverged at wire/density gradient Sum.";
                break;
            }
            size_t ok = 0;
            num_region_diverged = 0;
            for (size_t b = 0; b < rg.size(); ++b) {
                ok += update_step_length(int(b)) ? 1 : 0;
                num_region_diverged += rg[b].is_diverged ? 1 : 0;
            }
            if (num_region_diverged > 0) {
                diverge_code = 305;
                diverge_msg = "RePlAce diverged during gradient descent calculation, resulting in an invalid step length (Inf or NaN). This is often caused by numerical instability or high placement density. Consider reducing placement density to potentially resolve the issue.";
                break;
            }
            if (ok != rg.size()) break;
        }
    }

    void adjust_phi(Region& r) {
        if (r.is_converged) return;
        if (!r.max_phi_coef_changed && r.sum_overflow_unscaled < 0.35f) {
            r.max_phi_coef_changed = true;
            r.max_phi_coef = (float)((double)r.max_phi_coef * 0.99);
        }
        if (r.max_phi_coef <= 1.0f) r.max_phi_coef = 1.01f;
    }

    float phi_coef(const Region& r, float scaled) const {
        float m = r.max_phi_coef;
        float v = scaled < 0.0f ? m : (float)((double)m * std::pow((double)m, -(double)scaled));
        return std::max(o.min_phi_coef, v);
    }

    void report_header() {
        log.report(strfmt("%9s | %8s | %13s | %8s | %9s | %5s", "Iteration", "Overflow", "HPWL (um)", "HPWL(%)", "Penalty", "Group"));
        log.report("---------------------------------------------------------------");
    }

    // A region's progress row (every tenth iteration), after the table's header at the first.
    void report_row(Region& r, int iter, float overflow, int64_t hpwl, float pct, float penalty) {
        std::string group = r.R.group >= 0 ? strfmt(" (group%d)", r.R.group) : std::string();
        if ((iter == 0 || r.reprint_header) && r.R.group < 0) {
            if (iter == 0) log.info(31, "HPWL: Half-Perimeter Wirelength");
            report_header();
            r.reprint_header = false;
        }
        log.report(strfmt("%9d | %8.4f | %13.6e | %+7.2f%% | %9.2e | %5s", iter, overflow, (double)hpwl / (double)dbu, pct, (double)penalty,
                          group.c_str()));
    }

    // NesterovBase::updateNextIter (nesterov.rs:2837).
    void nb_update_next_iter(int b, int iter) {
        Region& r = rg[b];
        if (r.is_converged) return;
        for (int q = 0; q < 4; ++q) std::swap(r.v[S_PREV_SLP + q], r.v[S_CUR_SLP + q]);
        if (r.has_locked) {
            float2* nv[5] = {r.v[S_NEXT_SLP], r.v[S_NEXT_SLP_WL], r.v[S_NEXT_SLP_DEN], r.v[S_NEXT_SLP_SUM], r.v[S_NEXT]};
            const float2* cv[5] = {r.v[S_CUR_SLP], r.v[S_CUR_SLP_WL], r.v[S_CUR_SLP_DEN], r.v[S_CUR_SLP_SUM], r.v[S_CUR]};
            gplk::launch_lock_fix(stream, r.dev, cells, nv, cv, 5);
        }
        for (int q = 0; q < 4; ++q) std::swap(r.v[S_CUR_SLP + q], r.v[S_NEXT_SLP + q]);
        std::swap(r.v[S_CUR], r.v[S_NEXT]);
        float fraction = (float)iter / (float)max_iter;
        float denom = std::max((float)r.nia, fraction * (float)r.R.non_place_insts_area * 0.05f);
        r.sum_overflow = (float)r.overflow / denom;
        r.sum_overflow_unscaled = (float)r.overflow_unscaled / denom;
        int64_t hpwl = cur_hpwl;
        if (iter == 0 || iter % 10 == 0) {
            float pct = 0.0f;
            if (r.prev_reported_hpwl != 0) pct = (float)((double)(hpwl - r.prev_reported_hpwl) / (double)r.prev_reported_hpwl * 100.0);
            r.prev_reported_hpwl = hpwl;
            r.prev_reported_overflow_unscaled = r.sum_overflow_unscaled;
            report_row(r, iter, r.sum_overflow_unscaled, hpwl, pct, r.density_penalty);
        }
        float scaled = (float)(hpwl - r.prev_hpwl) / reference_hpwl;
        float phi = phi_coef(r, scaled);
        r.density_penalty *= phi;
        r.prev_hpwl = hpwl;
        r.peak_coordi_distance = std::max(r.peak_coordi_distance, r.coordi_distance);
        if (iter > 50 && r.min_sum_overflow > r.sum_overflow_unscaled) {
            r.min_sum_overflow = r.sum_overflow_unscaled;
            r.hpwl_with_min_sum_overflow = (float)r.prev_hpwl;
        }
    }

    // NesterovPlace::updateNextIter (nesterov.rs:2907).
    void np_update_next_iter(int iter) {
        float total = 0.0f, total_unscaled = 0.0f;
        for (size_t b = 0; b < rg.size(); ++b) {
            nb_update_next_iter(int(b), iter);
            total += rg[b].sum_overflow;
            total_unscaled += rg[b].sum_overflow_unscaled;
        }
        float n = (float)rg.size();
        average_overflow = total / n;
        average_overflow_unscaled = total_unscaled / n;
        update_wirelength_coef(average_overflow);
        if (!o.disable_revert_if_diverge) {
            int64_t hpwl = cur_hpwl;
            if (hpwl < min_hpwl && (double)average_overflow_unscaled <= 0.25) {
                min_hpwl = hpwl;
                diverge_snapshot_overflow = average_overflow_unscaled;
                diverge_snapshot_iter = iter + 1;
                is_min_hpwl = true;
            } else {
                is_min_hpwl = false;
            }
        }
    }

    // The problem's index of each device net, refreshed after net re-sorts.
    void ensure_net_map() {
        if (!net_map_stale) return;
        check(hipMemcpyAsync(net_orig_of.data(), d_net_orig, sizeof(int32_t) * net_orig_of.size(), hipMemcpyDeviceToHost, stream), "copy");
        check(hipStreamSynchronize(stream), "sync");
        for (size_t i = 0; i < net_orig_of.size(); ++i) net_sorted_of[net_orig_of[i]] = int32_t(i);
        net_map_stale = false;
    }

    // The problem's index of each device cell, refreshed after re-sorts.
    void ensure_host_map() {
        if (!host_map_stale) return;
        check(hipMemcpyAsync(cell_old_of.data(), cells.old_of, sizeof(int32_t) * ncells, hipMemcpyDeviceToHost, stream), "copy");
        check(hipStreamSynchronize(stream), "sync");
        for (int32_t i = 0; i < ncells; ++i) {
            cell_new_of[cell_old_of[i]] = i;
            cell_inst[i] = cell_inst_prob[cell_old_of[i]];
        }
        host_map_stale = false;
    }

    // Re-sort every region's cells by the Morton code of their current bin: the cell arrays, the
    // pin CSR and every per-cell vector are permuted and the net pins' cells renamed.  Only the
    // device's storage order changes; the integer scatter is exact in any order and the FP64 sums
    // stay deterministic.
    void resort() {
        const int32_t N = ncells;
        mark(F_REDUCE);
        gplk::launch_resort_sort(stream, N, cells.ctr, cells.flags, rtab, rb);
        gplk::launch_resort_cells(stream, N, cells, cells_alt, rb, nets.cpi);
        std::swap(cells, cells_alt);
        d_flags = cells.flags;
        // the vectors the loop reads again: the coordinates, the preconditioned gradients and, while
        // a snapshot can still be restored, the snapshots (the wl / den vectors and the next-point
        // slots are rewritten before any read); all of a region's in one launch
        static const int live[] = {S_CUR_SLP, S_CUR_SLP_SUM, S_PREV_SLP, S_PREV_SLP_SUM, S_CUR, S_SNAP, S_SNAP_SLP, S_SNAP_SLP_SUM,
                                   S_SNAP_PREV_SLP_SUM};
        const int nlive = (diverge_saved || diverge_reverts > 0 || rt.saved) ? 9 : 5;
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            gplk::GatherSet gs{};
            gs.count = nlive;
            for (int i = 0; i < nlive; ++i) {
                gs.in[i] = r.v[live[i]];
                gs.out[i] = region_scratch[b][i];
            }
            gplk::launch_gather_f2_set(stream, r.n, r.R.cell_begin, rb.perm, gs);
            for (int i = 0; i < nlive; ++i) std::swap(r.v[live[i]], region_scratch[b][i]);
        }
        gplk::launch_remap_pcell(stream, net_pin_count, nets.pcell, rb.inv);
        // the nets follow the cells: the small nets ordered again by their first cell's window
        gplk::launch_net_resort(stream, nets, net_pin_count, ncells, d_net_orig, cells, nrb);
        {
            int32_t* t = const_cast<int32_t*>(nets.start);
            nets.start = nrb.start;
            nrb.start = t;
            std::swap(nets.pcell, nrb.pcell);
            std::swap(nets.cpi, nrb.cpi);
            int2* o = const_cast<int2*>(nets.poff);
            nets.poff = nrb.poff;
            nrb.poff = o;
            o = const_cast<int2*>(nets.pxy0);
            nets.pxy0 = nrb.pxy0;
            nrb.pxy0 = o;
            float* w = const_cast<float*>(nets.weight);
            nets.weight = nrb.weight;
            nrb.weight = w;
            std::swap(d_net_orig, nrb.orig);
        }
        gplk::launch_tile_bounds(stream, nets, d_tile);
        net_map_stale = true;
        mark_end();
        host_map_stale = true;
        ++resorts;
        drop_ctl_graph();  // it holds the old cell and net arrays
    }

    // The scatter's cost follows its count of global atomic adds: one per distinct bin a chunk of
    // cells touches.  A re-sort when that count per trial has grown by half over its level in the
    // first iteration after the last sort, at most every 50 iterations: a re-sort also moves each
    // net's cells away from the nets' own order (the WA kernels then gather from farther apart),
    // so frequent re-sorts cost more than the scatter gains (at 1M cells: every >= 8 iterations at
    // +25 %: 40 re-sorts, 4340 it/s; every >= 50 at +50 %: 8 re-sorts, 4850 it/s; none: 4680).
    void maybe_resort(int32_t iter) {
        if (!resort_enabled) return;
        if (inject_kind == 3 && inject_after >= 0 && iter >= inject_after) {
            resort();
            last_resort_iter = iter;
            return;
        }
        uint64_t ga = 0;
        std::memcpy(&ga, &h_sums[rg.size()].pad[0], sizeof ga);
        const double rate = double(ga - last_global_adds) / double(std::max(trials, 1));
        last_global_adds = ga;
        if (iter == last_resort_iter + 1 || adds_base == 0.0) adds_base = rate;
        if (rate > (double)resort_growth * adds_base && iter - last_resort_iter >= resort_gap) {
            autotune_before_resort();
            resort();
            last_resort_iter = iter;
        }
    }

    // Per-cell float pairs of region b between the problem's order (host) and the device's.
    void upload_region_vec(int b, float2* dev, const float* host) {
        ensure_host_map();
        Region& r = rg[b];
        std::vector<float2> tmp(r.n);
        for (int k = 0; k < r.n; ++k) {
            int old = cell_old_of[r.R.cell_begin + k] - r.R.cell_begin;
            tmp[k] = make_float2(host[2 * old], host[2 * old + 1]);
        }
        check(hipMemcpyAsync(dev, tmp.data(), sizeof(float2) * r.n, hipMemcpyHostToDevice, stream), "copy");
        check(hipStreamSynchronize(stream), "sync");
    }
    void download_region_vec(int b, const float2* dev, float* host) {
        ensure_host_map();
        Region& r = rg[b];
        std::vector<float2> tmp(r.n);
        check(hipMemcpyAsync(tmp.data(), dev, sizeof(float2) * r.n, hipMemcpyDeviceToHost, stream), "copy");
        check(hipStreamSynchronize(stream), "sync");
        for (int k = 0; k < r.n; ++k) {
            int old = cell_old_of[r.R.cell_begin + k] - r.R.cell_begin;
            host[2 * old] = tmp[k].x;
            host[2 * old + 1] = tmp[k].y;
        }
    }

    void copy_vec(float2* dst, const float2* src, int n) {
        if (n > 0) check(hipMemcpyAsync(dst, src, sizeof(float2) * n, hipMemcpyDeviceToDevice, stream), "snapshot copy");
    }

    void save_snapshot(Region& r) {
        if (r.is_converged) return;
        copy_vec(r.v[S_SNAP], r.v[S_CUR], r.n);
        copy_vec(r.v[S_SNAP_SLP], r.v[S_CUR_SLP], r.n);
        copy_vec(r.v[S_SNAP_SLP_SUM], r.v[S_CUR_SLP_SUM], r.n);
        copy_vec(r.v[S_SNAP_PREV_SLP_SUM], r.v[S_PREV_SLP_SUM], r.n);
        r.snapshot_density_penalty = r.density_penalty;
        r.snapshot_step_length = r.step_length;
    }

    void revert_to_snapshot(int b) {
        Region& r = rg[b];
        if (r.is_converged) return;
        copy_vec(r.v[S_CUR], r.v[S_SNAP], r.n);
        copy_vec(r.v[S_CUR_SLP], r.v[S_SNAP_SLP], r.n);
        copy_vec(r.v[S_CUR_SLP_SUM], r.v[S_SNAP_SLP_SUM], r.n);
        copy_vec(r.v[S_PREV_SLP_SUM], r.v[S_SNAP_PREV_SLP_SUM], r.n);
        r.density_penalty = r.snapshot_density_penalty;
        r.step_length = r.snapshot_step_length;
        mark(F_SCATTER);
        gplk::launch_center_scatter(stream, r.dev, cells, r.v[S_CUR], d_global_adds);
        mark_end();
        bins_and_solve(r, b);
        sync_stats();
        r.overflow = (int64_t)h_sums[b].overflow;
        r.overflow_unscaled = (int64_t)h_sums[b].overflow_unscaled;
        r.is_diverged = false;
        r.prev_reported_hpwl = 0;
        r.prev_reported_overflow_unscaled = 0.0f;
    }

    bool check_divergence(Region& r) {
        if (r.sum_overflow_unscaled < 0.2f && r.sum_overflow_unscaled - r.min_sum_overflow >= 0.02f &&
            r.hpwl_with_min_sum_overflow * 1.2f < (float)r.prev_hpwl) {
            r.is_diverged = true;
            log.warn(323, "Divergence detected between consecutive iterations");
        }
        if (r.prev_reported_overflow_unscaled > 0.0f && r.prev_reported_hpwl > 0) {
            float change = r.sum_overflow_unscaled - r.prev_reported_overflow_unscaled;
            float increase = (float)(r.prev_hpwl - r.prev_reported_hpwl) / (float)r.prev_reported_hpwl;
            if (change >= 0.05f && increase >= 0.25f) {
                r.is_diverged = true;
                log.warn(324, strfmt("Divergence detected between reported values. Overflow change: %g, HPWL increase: %g%%.", change, increase * 100.0f));
            }
        }
        if (iter_now >= 0 && (iter_now == inject_diverge_at || iter_now == inject_diverge_at2)) r.is_diverged = true;
        return r.is_diverged;
    }

    void revert_to_diverge_snapshot() {
        log.warn(999, strfmt("Revert to iter: %4d overflow: %.3f HPWL: %lld", diverge_snapshot_iter, diverge_snapshot_overflow, (long long)diverge_hpwl));
        wl_coef_x = diverge_coef_x;
        wl_coef_y = diverge_coef_y;
        wa_forward();
        for (size_t b = 0; b < rg.size(); ++b) {
            revert_to_snapshot(int(b));
            rg[b].min_sum_overflow = 1e30f;
            rg[b].hpwl_with_min_sum_overflow = 1e30f;
        }
        sync_stats();
        cur_hpwl = hpwl_now();
    }

    // isDiverged (nesterov.rs:3009): 0 none, 1 resume, 2 stop.  The min-HPWL snapshot only once
    // routability is off (it shares the routability snapshot's buffer); a divergence tries the
    // routability snapshot first.
    int np_is_diverged(float* cur_a) {
        num_region_diverged = 0;
        for (auto& r : rg) num_region_diverged += check_divergence(r) ? 1 : 0;
        const bool routability_off = !rt.on || !rt.need;
        if (num_region_diverged == 0 && is_min_hpwl && !o.disable_revert_if_diverge && routability_off) {
            diverge_coef_x = wl_coef_x;
            diverge_coef_y = wl_coef_y;
            diverge_hpwl = min_hpwl;
            for (auto& r : rg) save_snapshot(r);
            diverge_saved = true;
        }
        if (num_region_diverged == 0) return 0;
        log.report(strfmt("Divergence occured in %d regions.", num_region_diverged));
        bool disable = o.disable_revert_if_diverge != 0;
        if (!disable && rt.on && try_routability_diverge_recovery(cur_a)) {
            num_region_diverged = 0;
            return 1;
        }
        if (!disable && diverge_saved && diverge_reverts < 3) {
            diverge_reverts += 1;
            log.warn(998, strfmt("Divergence detected, reverting to snapshot with min hpwl and resuming (%d of 3 reverts allowed).", diverge_reverts));
            revert_to_diverge_snapshot();
            *cur_a = 1.0f;
            diverge_saved = false;
            num_region_diverged = 0;
            return 1;
        }
        if (!disable && diverge_reverts > 0) {
            log.warn(997, "Divergence detected again, reverting to snapshot with min hpwl and stopping.");
            revert_to_diverge_snapshot();
            num_region_diverged = 0;
            return 2;
        }
        diverge_code = 307;
        diverge_msg = "RePlAce divergence detected: Current overflow is low and increasing relative to minimum,and the HPWL has significantly worsened. Consider re-running with a smaller max_phi_cof value.";
        return 2;
    }

    void lock_region(int b);

    // ---- the autotuner ---------------------------------------------------------------------------
    void apply_tune(int b) {
        Region& r = rg[b];
        const Tune& t = tune[b];
        r.dev.chunk = t.chunk;
        r.dev.narrow = t.narrow && r.narrow_ok ? 1 : 0;
        r.dev.grad_stage = t.stage ? gplk::kGradStage : 0;
    }

    // The shortest of `reps` device timings of f (HIP events on the stream), in microseconds; `after`
    // runs untimed after each (the scatter's bins pass, which clears the accumulators).
    template <typename F, typename A>
    double time_us(F f, A after, int reps) {
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "event");
        check(hipEventCreate(&e1), "event");
        double best = 1e30;
        for (int k = 0; k < reps; ++k) {
            check(hipEventRecord(e0, stream), "event");
            f();
            check(hipEventRecord(e1, stream), "event");
            after();
            check(hipEventSynchronize(e1), "sync");
            float ms = 0.0f;
            hipEventElapsedTime(&ms, e0, e1);
            best = std::min(best, 1000.0 * (double)ms);
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return best;
    }

    // The scatter's variants timed on the boxes the cells hold now, then the bins pass (clearing the
    // accumulators; the density it recomputes is that of the same boxes): the fastest is kept.
    void tune_scatter(int b, int reps) {
        Region& r = rg[b];
        if (r.n <= 0) return;
        Tune& t = tune[b];
        const Tune before = t;
        double best = 1e30;
        for (int chunk : {gplk::kPlaceChunk, 256})
            for (int nar : {1, 0}) {
                if (nar && !r.narrow_ok) continue;
                r.dev.chunk = chunk;
                r.dev.narrow = nar;
                const double us = time_us([&] { gplk::launch_density_scatter(stream, r.dev, cells, nullptr); },
                                          [&] { gplk::launch_bin_density(stream, r.dev, r.partials); }, reps);
                if (us < best) {
                    best = us;
                    t.chunk = chunk;
                    t.narrow = nar;
                }
            }
        t.scatter_us = best;
        t.source = 1;
        apply_tune(b);
        if (t.chunk != before.chunk || t.narrow != before.narrow) drop_graph();  // the graph holds the old variant
    }

    // The gradient with and without its LDS staging of the pin terms, into the next slot (free
    // before np_init and between trials).
    void tune_gradient(int b) {
        Region& r = rg[b];
        if (r.n <= 0) return;
        Tune& t = tune[b];
        double best = 1e30;
        for (int stage : {1, 0}) {
            r.dev.grad_stage = stage ? gplk::kGradStage : 0;
            const double us = time_us(
                [&] {
                    gplk::launch_gradient(stream, r.dev, cells, nets, wl_coef_x, wl_coef_y, r.density_penalty, nullptr, nullptr,
                                          r.v[S_NEXT_SLP_SUM], nullptr, nullptr, nullptr, r.partials, 0);
                },
                [] {}, 3);
            if (us < best) {
                best = us;
                t.stage = stage;
            }
        }
        t.gradient_us = best;
        apply_tune(b);
    }

    // Before np_init: every region measured on the initial boxes (or its shape's cached pick taken).
    void autotune_init() {
        if (!autotune) return;
        const auto t0 = Clock::now();
        join();
        bool any = false;
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            const TuneKey key{r.n, r.R.cnt_x, r.R.cnt_y, net_pin_count};
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune_cache.find(key);
            if (it != g_tune_cache.end()) {
                tune[b] = it->second;
                tune[b].source = 2;
                apply_tune(int(b));
                continue;
            }
            any = true;
            tune_scatter(int(b), 3);
            // a field and pin terms for the gradient's timing
            gplk::launch_density_scatter(stream, r.dev, cells, nullptr);
            gplk::launch_bin_density(stream, r.dev, r.partials);
            gplk::launch_solve(stream, r.dev, r.sb, false);
            wa_forward();
            join();
            tune_gradient(int(b));
            g_tune_cache[key] = tune[b];
        }
        check(hipStreamSynchronize(stream), "sync");
        if (any) ++tune_runs;
        tune_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        drop_graph();
    }

    // Just before a shape's first re-sort the storage order is at its most decayed, where the
    // scatter's variants differ most: they are measured there once and that pick is kept (and
    // cached for the shape, so later contexts start with it, and contexts that reach the point
    // after another of the same shape take its pick).
    void autotune_before_resort() {
        if (!autotune) return;
        bool need = false;
        for (size_t b = 0; b < rg.size(); ++b) need = need || (!rg[b].is_converged && !tune[b].decayed);
        if (!need) return;
        const auto t0 = Clock::now();
        join();
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            if (r.is_converged || tune[b].decayed) continue;
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                auto it = g_tune_cache.find(TuneKey{r.n, r.R.cnt_x, r.R.cnt_y, net_pin_count});
                if (it != g_tune_cache.end() && it->second.decayed) {
                    const Tune before = tune[b];
                    tune[b] = it->second;
                    tune[b].source = 2;
                    apply_tune(int(b));
                    if (tune[b].chunk != before.chunk || tune[b].narrow != before.narrow || tune[b].stage != before.stage) drop_graph();
                    continue;
                }
            }
            tune_scatter(int(b), 2);
            tune[b].decayed = true;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune_cache[TuneKey{r.n, r.R.cnt_x, r.R.cnt_y, net_pin_count}] = tune[b];
        }
        check(hipStreamSynchronize(stream), "sync");
        ++tune_runs;
        tune_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    }

    // checkConvergence (nesterov.rs:3107).
    // ---- routability-driven placement ---------------------------------------------------------
    bool is_std(int32_t g) const { return (cell_flags_prob[g] & GPL_CELL_INSTANCE) && !(cell_flags_prob[g] & GPL_CELL_MACRO); }
    double dbu_area(double a) const { return a / (double(dbu) * double(dbu)); }

    // The routability state at construction: RouteBase::new + route_base_init (route_base.rs:
    // 298-339, nesterov.rs:3776-3782), the grid the router would supply, the parking stores.
    void route_setup(const gpl_problem* p) {
        rt.on = true;
        // the die (region 0's box when the problem names none)
        const gpl_region& R0 = p->regions[0];
        int32_t die[4] = {p->die[0], p->die[1], p->die[2], p->die[3]};
        if (die[2] <= die[0] || die[3] <= die[1]) {
            die[0] = R0.lx;
            die[1] = R0.ly;
            die[2] = R0.ux;
            die[3] = R0.uy;
        }
        const int32_t gcells = o.route_gcells > 0 ? o.route_gcells : 64;
        const int32_t tile = std::max(1, (die[2] - die[0]) / gcells);
        const int32_t pitch = std::max(1, tile / (o.route_tracks > 0 ? o.route_tracks : 8));
        const int32_t pitches[2] = {pitch, pitch};
        rt.grid = gplk::RudyGrid{die[0], die[1], die[2], die[3], gcells, std::max(1, (die[3] - die[1]) / tile), tile,
                                 gpl_rudy_wire_width(pitches, 2)};
        rt.vars = gplk::RouteVarsDev{o.routability_min_congestion, o.routability_inflation_ratio_coef, o.routability_max_inflation_ratio,
                                     0.8f, {o.routability_rc_k[0], o.routability_rc_k[1], o.routability_rc_k[2], o.routability_rc_k[3]}};
        const size_t nb = rg.size();
        rt.min_rc_target_density.assign(nb, 0.0f);
        rt.inflated_delta.assign(nb, 0);
        rt.accumulated.assign(nb, 0);
        rt.min_rc_inflated.assign(nb, 0);
        rt.original_movable_area = 0;
        for (auto& r : rg) rt.original_movable_area += nesterov_insts_area(r);
        rt.size.resize(2 * size_t(ncells));
        for (int32_t g = 0; g < ncells; ++g) {
            rt.size[2 * g] = p->cell_box[4 * g + 2] - p->cell_box[4 * g];
            rt.size[2 * g + 1] = p->cell_box[4 * g + 3] - p->cell_box[4 * g + 1];
        }
        rt.min_rc_size.assign(2 * size_t(ncells), 0);
        rt.net_weight.assign(size_t(nnets), 0.0f);
        rt.min_rc_net_weight.assign(size_t(nnets), 0.0f);
        rt.fillers.resize(nb);
        rt.removed.assign(nb, {});
        rt.parked.resize(nb);
        rt.total_filler_area.resize(nb);
        for (size_t b = 0; b < nb; ++b) {
            const gpl_region& R = p->regions[b];
            rt.filler_base.push_back(R.cell_end - R.num_fillers);
            rt.num_fillers.push_back(R.num_fillers);
            rt.n_alloc.push_back(R.cell_end - R.cell_begin);
            for (int32_t k = 0; k < R.num_fillers; ++k) rt.fillers[b].push_back(R.cell_end - R.num_fillers + k);
            rt.parked[b].assign(size_t(R.num_fillers), 0);
            rt.total_filler_area[b] = R.total_filler_area;
            rt.store.push_back(dalloc<float2>(size_t(gplk::kGatherMax) * std::max(1, R.num_fillers)));
            const size_t M = size_t(R.cnt_x) * R.cnt_y;
            std::vector<int64_t> none(M, 0), neg(M, -1);
            rt.np_raw.push_back((long long*)upload(R.non_place_raw ? R.non_place_raw : none.data(), M));
            rt.np_union.push_back((long long*)upload(R.non_place_union ? R.non_place_union : neg.data(), M));
        }
        rt.d_sizes = dalloc<int2>(size_t(ncells));
        rt.d_list = dalloc<int32_t>(3 * size_t(ncells));
        rt.d_ratio = dalloc<float>(size_t(rt.grid.cnt_x) * rt.grid.cnt_y);
        rt.d_demand = dalloc<float>(size_t(std::max(nnets, 1)));
        rt.d_delta = dalloc<unsigned long long>(1);
    }

    // setTargetDensity (nesterov.rs:3793-3801): the region's and its bins' target density, the
    // bins' fixed area again at it.
    void set_target_density(int b, float td) {
        Region& r = rg[b];
        r.R.target_density = td;
        r.dev.target_density = td;
        check(gplk::launch_non_place(stream, r.R.cnt_x * r.R.cnt_y, rt.np_raw[b], rt.np_union[b], td, const_cast<long long*>(r.dev.non_place),
                                     const_cast<float2*>(r.dev.non_place_f)),
              "non place");
        r.nia = nesterov_insts_area(r);
        drop_graph();
    }

    // updateAreas (nesterov.rs:3804-3822) from the cell sizes now.
    void update_areas(int b) {
        Region& r = rg[b];
        int64_t st = 0, mac = 0;
        for (int32_t g = r.R.cell_begin; g < rt.filler_base[b]; ++g) {
            if (!(cell_flags_prob[g] & GPL_CELL_INSTANCE)) continue;
            const int64_t a = int64_t(rt.size[2 * g]) * int64_t(rt.size[2 * g + 1]);
            if (cell_flags_prob[g] & GPL_CELL_MACRO)
                mac += a;
            else
                st += a;
        }
        r.R.std_insts_area = st;
        r.R.macro_insts_area = mac;
        r.nia = nesterov_insts_area(r);
    }

    // the region's live cells' sizes in device order
    void upload_sizes(int b) {
        ensure_host_map();
        Region& r = rg[b];
        std::vector<int2> sz((size_t)std::max(r.n, 1));
        for (int k = 0; k < r.n; ++k) {
            const int32_t g = cell_old_of[r.R.cell_begin + k];
            sz[k] = make_int2(rt.size[2 * size_t(g)], rt.size[2 * size_t(g) + 1]);
        }
        check(hipMemcpyAsync(rt.d_sizes, sz.data(), sizeof(int2) * sz.size(), hipMemcpyHostToDevice, stream), "upload");
        check(hipStreamSynchronize(stream), "sync");
    }

    // updateDensitySize of the region's live cells at their sizes now (nesterov.rs:2022-2036).
    void update_density_size(int b);

    // The device's parked cells made those routability took out: newly taken fillers saved and
    // parked, restored ones back among the live cells with their slots; one re-sort.
    void sync_parking() {
        ensure_host_map();
        bool change = false;
        std::vector<std::vector<int32_t>> to_park(rg.size()), to_back(rg.size());
        for (size_t b = 0; b < rg.size(); ++b) {
            std::vector<uint8_t> live(size_t(rt.num_fillers[b]), 0);
            for (int32_t g : rt.fillers[b]) live[size_t(g - rt.filler_base[b])] = 1;
            for (int32_t k = 0; k < rt.num_fillers[b]; ++k) {
                if (!live[k] && !rt.parked[b][k]) to_park[b].push_back(rt.filler_base[b] + k);
                if (live[k] && rt.parked[b][k]) to_back[b].push_back(rt.filler_base[b] + k);
            }
            change = change || !to_park[b].empty() || !to_back[b].empty();
        }
        if (!change) return;
        join();
        auto slots = [&](Region& r) {
            static const int live[] = {S_CUR_SLP, S_CUR_SLP_SUM, S_PREV_SLP, S_PREV_SLP_SUM, S_CUR, S_SNAP, S_SNAP_SLP, S_SNAP_SLP_SUM,
                                       S_SNAP_PREV_SLP_SUM};
            gplk::ParkSlots ps{};
            ps.count = gplk::kGatherMax;
            for (int i = 0; i < gplk::kGatherMax; ++i) ps.v[i] = r.v[live[i]];
            return ps;
        };
        // d_list: the cells' device indices, their indices within the region (the slots'), their
        // store indices
        auto lists = [&](const std::vector<int32_t>& cells_, int b) {
            const size_t k = cells_.size();
            std::vector<int32_t> h(3 * k);
            for (size_t i = 0; i < k; ++i) {
                h[i] = cell_new_of[cells_[i]];
                h[k + i] = h[i] - rg[b].R.cell_begin;
                h[2 * k + i] = cells_[i] - rt.filler_base[b];
            }
            if (!h.empty()) check(hipMemcpyAsync(rt.d_list, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice, stream), "upload");
            check(hipStreamSynchronize(stream), "sync");
        };
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            const int kp = int(to_park[b].size()), kb = int(to_back[b].size());
            if (kp) {
                lists(to_park[b], int(b));
                check(gplk::launch_park_save(stream, kp, rt.d_list + kp, rt.d_list + 2 * kp, slots(r), rt.store[b], rt.num_fillers[b]), "park");
                check(gplk::launch_flag_list(stream, kp, rt.d_list, cells.flags, gplk::kCellParked, 0), "flags");
            }
            if (kb) {
                lists(to_back[b], int(b));
                check(gplk::launch_flag_list(stream, kb, rt.d_list, cells.flags, 0, gplk::kCellParked), "flags");
            }
            for (int32_t g : to_park[b]) rt.parked[b][size_t(g - rt.filler_base[b])] = 1;
            for (int32_t g : to_back[b]) rt.parked[b][size_t(g - rt.filler_base[b])] = 0;
            int32_t np_ = 0;
            for (uint8_t q : rt.parked[b]) np_ += q;
            r.n = rt.n_alloc[b] - np_;
            r.dev.cell_end = r.R.cell_begin + r.n;
        }
        resort();
        ensure_host_map();
        for (size_t b = 0; b < rg.size(); ++b) {
            const int kb = int(to_back[b].size());
            if (!kb) continue;
            lists(to_back[b], int(b));
            check(gplk::launch_park_restore(stream, kb, rt.d_list + kb, rt.d_list + 2 * kb, slots(rg[b]), rt.store[b], rt.num_fillers[b]),
                  "unpark");
        }
        check(hipStreamSynchronize(stream), "sync");
        last_resort_iter = iter_now;
        drop_graph();
    }

    // cutFillerCells (nesterov.rs:3826-3870): whole fillers from the back of the region's cells
    // for the inflation's area, the target density raised for what they cannot absorb.
    void cut_filler_cells(int b, int64_t inflation_area) {
        Region& r = rg[b];
        if (inflation_area < 0) {
            log.warn(313, strfmt("Negative area provided to remove fillers: %g. Expected positive value, ignoring.", dbu_area(double(inflation_area))));
            return;
        }
        const int64_t single = int64_t(r.R.filler_dx) * int64_t(r.R.filler_dy);
        auto& fl = rt.fillers[b];
        const int64_t max_remove = single > 0 ? std::min<int64_t>(inflation_area / single, int64_t(fl.size())) : 0;
        int64_t available = single * int64_t(fl.size()), removed = 0;
        while (!fl.empty() && removed < max_remove) {
            rt.removed[b].push_back(fl.back());
            fl.pop_back();
            available -= single;
            ++removed;
        }
        rt.total_filler_area[b] = available;
        if (single * int64_t(fl.size()) != rt.total_filler_area[b])
            log.warn(312, strfmt("Unexpected filler area! The value %g, should be equal to totalFillerArea_ %g.", dbu_area(double(single * int64_t(fl.size()))),
                                 dbu_area(double(rt.total_filler_area[b]))));
        const int64_t removed_area = single * removed, remaining = inflation_area - removed_area;
        if (remaining > single) {
            const int64_t total = nesterov_insts_area(r) + removed_area + rt.total_filler_area[b] + remaining;
            set_target_density(b, (float)total / (float)r.R.white_space_area);
            r.R.movable_area = (int64_t)((float)r.R.white_space_area * r.R.target_density);
        }
    }

    // restoreRemovedFillers (nesterov.rs:3873-3891): put back at the end in the order taken.
    void restore_removed_fillers(int b) {
        const int64_t single = int64_t(rg[b].R.filler_dx) * int64_t(rg[b].R.filler_dy);
        for (int32_t g : rt.removed[b]) {
            rt.fillers[b].push_back(g);
            rt.total_filler_area[b] += single;
        }
        rt.removed[b].clear();
    }

    // applyNetWeights (nesterov.rs:3922-3937): each net's weight from the map (at least 1), every
    // other net 1; uploaded in the device's net order.
    int32_t apply_net_weights() {
        ensure_net_map();
        int32_t weighted = 0;
        std::vector<float> w((size_t)nnets);
        for (int32_t i = 0; i < nnets; ++i) {
            const float x = rt.net_weight[size_t(net_orig_of[i])];
            w[i] = std::max(1.0f, x);
            if (w[i] > 1.0f) ++weighted;
        }
        check(hipMemcpyAsync(const_cast<float*>(nets.weight), w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, stream), "upload");
        check(hipStreamSynchronize(stream), "sync");
        rt.weighted = weighted;
        return weighted;
    }

    // revertToMinCongestion (nesterov.rs:3896-3918).
    void revert_to_min_congestion() {
        if (rt.revert_count < 1) return;
        log.info(55, strfmt("Reverting inflation and target density to the pass with the minimum observed routing congestion (%.4f), target density %.4f.",
                            (double)rt.min_rc, (double)(rt.min_rc_target_density.empty() ? 0.0f : rt.min_rc_target_density[0])));
        for (int32_t g = 0; g < ncells; ++g)
            if (is_std(g)) {
                rt.size[2 * size_t(g)] = rt.min_rc_size[2 * size_t(g)];
                rt.size[2 * size_t(g) + 1] = rt.min_rc_size[2 * size_t(g) + 1];
            }
        for (size_t b = 0; b < rg.size(); ++b) {
            set_target_density(int(b), rt.min_rc_target_density[b]);
            restore_removed_fillers(int(b));
        }
        sync_parking();
        for (size_t b = 0; b < rg.size(); ++b) {
            update_density_size(int(b));
            rt.accumulated[b] = rt.min_rc_inflated[b];
        }
        rt.net_weight = rt.min_rc_net_weight;
        apply_net_weights();
        std::fill(rt.inflated_delta.begin(), rt.inflated_delta.end(), 0);
    }

    // The tile under a point (getTile, route_base.rs:428-437): origin (0, 0), the last column and
    // row clamped; -1 left of or below the grid.
    int64_t tile_at(int32_t x, int32_t y) const {
        const int32_t ix = std::min(x / rt.grid.tile, rt.grid.cnt_x - 1), iy = std::min(y / rt.grid.tile, rt.grid.cnt_y - 1);
        if (ix < 0 || iy < 0) return -1;
        return int64_t(iy) * rt.grid.cnt_x + ix;
    }

    // The density centres (d_cx, d_cy) of every live cell, the problem's order.
    std::vector<int32_t> density_centres() {
        ensure_host_map();
        std::vector<int4> db((size_t)ncells);
        check(hipMemcpyAsync(db.data(), cells.dbox, sizeof(int4) * ncells, hipMemcpyDeviceToHost, stream), "copy");
        check(hipStreamSynchronize(stream), "sync");
        std::vector<int32_t> c(2 * size_t(ncells));
        for (int32_t i = 0; i < ncells; ++i) {
            const int32_t g = cell_old_of[i];
            c[2 * size_t(g)] = (db[i].x + db[i].z) / 2;
            c[2 * size_t(g) + 1] = (db[i].y + db[i].w) / 2;
        }
        return c;
    }

    // scaleInflationToBudget (nesterov.rs:3946-3978): whether the budget is spent.
    bool scale_inflation_to_budget() {
        if (rt.original_movable_area <= 0) return false;
        const double budget = double(rt.original_movable_area) * double(o.routability_max_inflation_total);
        int64_t total = 0;
        for (int64_t a : rt.accumulated) total += a;
        const double spent = double(total);
        const double allowed = std::min(budget / 6.0, budget - spent);
        const std::vector<int32_t> ctr = density_centres();
        double wanted = 0.0;
        for (size_t b = 0; b < rg.size(); ++b)
            for (int32_t g = rg[b].R.cell_begin; g < rt.filler_base[b]; ++g) {
                if (!is_std(g)) continue;
                const int64_t t = tile_at(ctr[2 * size_t(g)], ctr[2 * size_t(g) + 1]);
                if (t < 0 || t >= int64_t(rt.inflated_ratio.size())) continue;
                const float ratio = rt.inflated_ratio[size_t(t)];
                if ((double)ratio <= 1.0) continue;
                const double area = double(rt.size[2 * size_t(g)]) * double(rt.size[2 * size_t(g) + 1]);
                wanted += area * ((double)ratio - 1.0);
            }
        if (wanted <= 0.0 || wanted <= allowed) return spent + wanted >= budget;
        const double scale = std::max(0.0, allowed / wanted);
        for (float& r : rt.inflated_ratio)
            if (r > 1.0f) r = (float)(1.0 + scale * ((double)r - 1.0));
        return spent + allowed >= budget;
    }

    // RUDY and the tile grid at the current placement (calculateRudyTiles + updateRudyAverage,
    // route_base.rs:443-536): the weighted congestion; the tiles' inflation ratios kept.
    float measure_rudy() {
        join();
        wa_forward(true);  // the nets' boxes at the cells' centres, into the net records
        join();
        check(gplk::launch_rudy(stream, nets, d_net_orig, rt.grid, nullptr, rudy), "rudy");
        check(gplk::launch_rudy_tiles(stream, cells, ncells, rt.grid, rt.vars, rudy), "rudy tiles");
        double st[8];
        const size_t tiles = size_t(rt.grid.cnt_x) * rt.grid.cnt_y;
        rt.inflation_ratio.resize(tiles);
        check(hipMemcpyAsync(st, rudy.stats, sizeof st, hipMemcpyDeviceToHost, stream), "copy");
        check(hipMemcpyAsync(rt.inflation_ratio.data(), rudy.ratio, sizeof(float) * tiles, hipMemcpyDeviceToHost, stream), "copy");
        check(hipStreamSynchronize(stream), "sync");
        rt.last_overflowed = int32_t(st[1]);
        rt.last_actionable = int32_t(st[2]);
        return float(st[7]);
    }

    // getNetDemandInTiles over the measured congested tiles (inflation ratio above 1), per problem
    // net, while the net records hold the measured boxes.
    void congested_net_demand() {
        rt.demand.clear();
        if (o.routability_net_weight_max <= 1.0f || o.routability_congested_nets_percentage <= 0.0f || nnets == 0) return;
        check(gplk::launch_net_congested_demand(stream, nets, d_net_orig, rt.grid, rudy.ratio, rt.d_demand, rudy), "net demand");
        rt.demand.resize(size_t(nnets));
        check(hipMemcpyAsync(rt.demand.data(), rt.d_demand, sizeof(float) * rt.demand.size(), hipMemcpyDeviceToHost, stream), "copy");
        check(hipStreamSynchronize(stream), "sync");
    }

    // weightCongestedNets (route_base.rs:381-426): the nets depositing the most demand in the
    // tiles about to be inflated weighted; -1 when none were (the map cleared either way).
    int32_t weight_congested_nets() {
        std::fill(rt.net_weight.begin(), rt.net_weight.end(), 0.0f);
        const float wmax = o.routability_net_weight_max, pct = o.routability_congested_nets_percentage;
        if (wmax <= 1.0f || pct <= 0.0f || nnets == 0 || rt.demand.size() != size_t(nnets)) return -1;
        const std::vector<float>& dem = rt.demand;
        std::vector<std::pair<int32_t, float>> d;
        for (int32_t e = 0; e < nnets; ++e)
            if (dem[e] > 0.0f) d.push_back({e, dem[e]});
        if (d.empty()) return -1;
        // by demand, largest first (ties in the nets' order: a net tied at the cutoff gets weight 1
        // whichever way the tie is broken, unless every chosen net ties with the worst)
        std::stable_sort(d.begin(), d.end(), [](const std::pair<int32_t, float>& a, const std::pair<int32_t, float>& b) { return a.second > b.second; });
        const size_t n = d.size();
        const size_t share = (size_t)std::ceil((double)((float)n * pct) / 100.0);
        const size_t count = std::min(n, std::max<size_t>(share, 1));
        const float worst = d[0].second, cutoff = d[count - 1].second;
        for (size_t i = 0; i < count; ++i) {
            const float w = worst > cutoff ? 1.0f + (wmax - 1.0f) * (d[i].second - cutoff) / (worst - cutoff) : wmax;
            rt.net_weight[size_t(d[i].first)] = w;
        }
        return int32_t(n);
    }

    // RouteBase::routability (nesterov.rs:3985-4093): (is_routability_need, is_revert_init_needed).
    std::pair<bool, bool> routability(int32_t revert_count) {
        rt.revert_count += 1;
        const float cur_rc = measure_rudy();
        congested_net_demand();
        {
            const int32_t over = rt.last_overflowed, act = rt.last_actionable;
            const double pct = act > 0 ? double(over) / double(act) * 100.0 : 0.0;
            log.info(40, strfmt("Routability pass %d: weighted congestion %.4f (target %.4f), %d of %d actionable tiles overflowed (%.1f%%).",
                                rt.revert_count, (double)cur_rc, (double)o.routability_target_rc, over, act, pct));
        }
        gpl_route_pass row{};
        row.iter = iter_now;
        row.rc = cur_rc;
        row.overflowed_tiles = rt.last_overflowed;
        row.actionable_tiles = rt.last_actionable;
        auto finish = [&](bool need, bool revert) {
            int64_t add = 0;
            for (int64_t v : rt.inflated_delta) add += v;
            row.inflation = add;
            row.target_density = rg[0].R.target_density;
            row.fillers = int32_t(rt.fillers[0].size());
            row.weighted_nets = rt.weighted;
            row.need = need;
            row.revert = revert;
            rt.passes.push_back(row);
            return std::make_pair(need, revert);
        };
        if (cur_rc < o.routability_target_rc) {
            log.info(50, strfmt("Weighted routing congestion is lower than target routing congestion(%.4f), end routability optimization.",
                                (double)o.routability_target_rc));
            auto res = finish(false, false);
            std::fill(rt.inflated_delta.begin(), rt.inflated_delta.end(), 0);
            return res;
        }
        if ((rt.min_rc - cur_rc) > rt.min_rc * 0.005f) {
            log.info(48, strfmt("Congestion improved on previous minimum (%.4g); updating minimum.", (double)rt.min_rc));
            rt.min_rc = cur_rc;
            rt.min_rc_violated = 0;
            for (size_t b = 0; b < rg.size(); ++b) {
                rt.min_rc_target_density[b] = rg[b].R.target_density;
                rt.removed[b].clear();
            }
            for (int32_t g = 0; g < ncells; ++g)
                if (is_std(g)) {
                    rt.min_rc_size[2 * size_t(g)] = rt.size[2 * size_t(g)];
                    rt.min_rc_size[2 * size_t(g) + 1] = rt.size[2 * size_t(g) + 1];
                }
            rt.min_rc_inflated = rt.accumulated;
            rt.min_rc_net_weight = rt.net_weight;
        } else {
            rt.min_rc_violated += 1;
            log.info(49, strfmt("No improvement on minimum (%.4f); consecutive non-improvement count: %d.", (double)rt.min_rc, rt.min_rc_violated));
        }
        rt.inflated_ratio.resize(rt.inflation_ratio.size());
        for (size_t t = 0; t < rt.inflation_ratio.size(); ++t) rt.inflated_ratio[t] = rt.inflation_ratio[t] > 1.0f ? rt.inflation_ratio[t] : 1.0f;
        const bool budget_spent = scale_inflation_to_budget();
        check(hipMemcpyAsync(rt.d_ratio, rt.inflated_ratio.data(), sizeof(float) * rt.inflated_ratio.size(), hipMemcpyHostToDevice, stream),
              "upload");
        for (size_t b = 0; b < rg.size(); ++b) {
            Region& r = rg[b];
            rt.inflated_delta[b] = 0;
            {
                // the standard cells in inflated tiles grown (device), their sizes back in the problem's order
                upload_sizes(int(b));
                gplk::launch_set_u64(stream, rt.d_delta, 0ull);
                check(gplk::launch_route_inflate(stream, r.R.cell_begin, r.n, cells, rt.grid, rt.d_ratio, rt.d_sizes, rt.d_delta), "inflate");
                std::vector<int2> sz((size_t)std::max(r.n, 1));
                unsigned long long dl = 0;
                check(hipMemcpyAsync(sz.data(), rt.d_sizes, sizeof(int2) * sz.size(), hipMemcpyDeviceToHost, stream), "copy");
                check(hipMemcpyAsync(&dl, rt.d_delta, sizeof dl, hipMemcpyDeviceToHost, stream), "copy");
                check(hipStreamSynchronize(stream), "sync");
                for (int k = 0; k < r.n; ++k) {
                    const int32_t g = cell_old_of[r.R.cell_begin + k];
                    rt.size[2 * size_t(g)] = sz[k].x;
                    rt.size[2 * size_t(g) + 1] = sz[k].y;
                }
                rt.inflated_delta[b] = (int64_t)dl;
            }
            rt.accumulated[b] += rt.inflated_delta[b];
            const int64_t delta = rt.inflated_delta[b];
            const float pct = ((float)delta / (float)nesterov_insts_area(r)) * 100.0f;
            const float delta_um2 = (float)dbu_area(double(delta));
            cut_filler_cells(int(b), delta);
            {
                int64_t total = 0;
                for (int64_t a : rt.accumulated) total += a;
                const double spent = rt.original_movable_area > 0 ? 100.0 * double(total) / double(rt.original_movable_area) : 0.0;
                const std::string group = r.R.group >= 0 ? strfmt(" in group group%d", r.R.group) : std::string();
                log.info(86, strfmt("Inflation %+.3f um^2 (%+.2f%% of movable), %.2f%% of the %.2f%% budget spent; target density %.4f%s.",
                                    (double)delta_um2, (double)pct, spent, 100.0 * (double)o.routability_max_inflation_total,
                                    (double)r.R.target_density, group.c_str()));
            }
            if (r.R.target_density > o.routability_max_density) {
                const std::string group = r.R.group >= 0 ? strfmt(" in group group%d", r.R.group) : std::string();
                log.info(53, strfmt("Target density %.4f exceeds the maximum allowed %.4f%s.", (double)r.R.target_density,
                                    (double)o.routability_max_density, group.c_str()));
                revert_to_min_congestion();
                return finish(false, true);
            }
        }
        if (revert_count >= o.routability_max_passes) {
            log.info(91, strfmt("Routability mode reached the maximum allowed reverts %d", revert_count));
            revert_to_min_congestion();
            return finish(false, true);
        }
        if (rt.min_rc_violated >= 3) {
            log.info(54, strfmt("No improvement in routing congestion for %d consecutive iterations (limit is %d).", rt.min_rc_violated, 3));
            revert_to_min_congestion();
            return finish(false, true);
        }
        sync_parking();
        for (size_t b = 0; b < rg.size(); ++b) {
            update_areas(int(b));
            update_density_size(int(b));
        }
        const int32_t depositors = weight_congested_nets();
        const int32_t weighted = apply_net_weights();
        if (depositors >= 0)
            log.info(93, strfmt("Weighted %d of %d nets feeding congested tiles, up to %.2fx wirelength weight.", weighted, depositors,
                                (double)o.routability_net_weight_max));
        auto res = finish(!budget_spent, true);
        std::fill(rt.inflated_delta.begin(), rt.inflated_delta.end(), 0);
        if (budget_spent) {
            int64_t total = 0;
            for (int64_t a : rt.accumulated) total += a;
            log.info(97, strfmt("Total routability inflation %.2f%% of the movable area spent the %.2f%% budget, ending routability.",
                                100.0 * double(total) / double(rt.original_movable_area), 100.0 * (double)o.routability_max_inflation_total));
            revert_to_min_congestion();
        }
        return res;
    }

    // isPlacementSettled / getWorstSettleRatio (nesterov.rs:3172-3181).
    bool is_placement_settled() const {
        for (auto& r : rg)
            if (!(r.peak_coordi_distance > 0.0f && r.coordi_distance <= 0.6f * r.peak_coordi_distance)) return false;
        return true;
    }
    float worst_settle_ratio() const {
        float w = 0.0f;
        for (auto& r : rg) w = std::max(w, r.peak_coordi_distance <= 0.0f ? 1.0f : r.coordi_distance / r.peak_coordi_distance);
        return w;
    }

    // tryRoutabilityDivergeRecovery (nesterov.rs:3062-3090): true when it reverted.
    bool try_routability_diverge_recovery(float* cur_a) {
        if (!rt.saved || rt.attempts >= 2) return false;
        if (!rt.need && (diverge_saved || diverge_reverts > 0)) return false;
        rt.attempts = rt.need ? rt.attempts + 1 : 2;
        const bool last = rt.attempts == 2;
        if (rt.need) revert_to_min_congestion();
        if (last) rt.need = false;
        log.warn(112, strfmt("Divergence detected, reverting to the routability snapshot (attempt %d of %d), further inflation %s.", rt.attempts, 2,
                             last ? "disabled" : "still allowed"));
        wl_coef_x = rt.coef_x;
        wl_coef_y = rt.coef_y;
        wa_forward();
        for (size_t b = 0; b < rg.size(); ++b) {
            revert_to_snapshot(int(b));
            rg[b].min_sum_overflow = 1e30f;
            rg[b].hpwl_with_min_sum_overflow = 1e30f;
        }
        sync_stats();
        *cur_a = 1.0f;
        rt.settle_start = -1;
        min_hpwl = INT64_MAX;
        is_min_hpwl = false;
        return true;
    }

    // The loop's routability step after the divergence test (nesterov.rs:4521-4567): the
    // routability snapshot, then a routability pass once the overflow reached its end and the
    // cells settled.
    void routability_step(int32_t iter, float* cur_a) {
        if (!rt.saved && o.routability_snapshot_overflow >= average_overflow_unscaled) {
            rt.coef_x = wl_coef_x;
            rt.coef_y = wl_coef_y;
            rt.a = *cur_a;
            rt.saved = true;
            for (auto& r : rg) save_snapshot(r);
            log.info(38, strfmt("Routability snapshot saved at iter = %d", iter + 1));
            log.report(strfmt("%9d | %8.4f | %13.6e | %8s | %9s | %5s", iter, (double)average_overflow_unscaled, (double)cur_hpwl / (double)dbu, " ", " ", " "));
        }
        const bool active = rt.need && average_overflow_unscaled <= o.routability_end_overflow;
        const bool settled = active && is_placement_settled();
        if (active && !settled && rt.settle_start == -1) {
            rt.settle_start = iter;
            log.info(99, strfmt("Routability end extended at iter = %d: overflow %.4f reached %.4f, but cells are still moving at %.0f%% of peak displacement (waiting for %.0f%%).",
                                iter + 1, (double)average_overflow_unscaled, (double)o.routability_end_overflow,
                                (double)(worst_settle_ratio() * 100.0f), (double)(0.6f * 100.0f)));
        } else if (settled && rt.settle_start != -1) {
            log.info(103, strfmt("Routability end extension finished at iter = %d after %d extra iterations: overflow %.4f, cells moving at %.0f%% of peak displacement.",
                                 iter + 1, iter - rt.settle_start, (double)average_overflow_unscaled, (double)(worst_settle_ratio() * 100.0f)));
            rt.settle_start = -1;
        }
        if (!settled) return;
        rg[0].reprint_header = true;
        rt.reverts += 1;
        const std::pair<bool, bool> res = routability(rt.reverts);
        const bool need = res.first, revert_init = res.second;
        rt.need = need;
        min_hpwl = INT64_MAX;
        is_min_hpwl = false;
        if (need || revert_init) {
            *cur_a = rt.a;
            wl_coef_x = rt.coef_x;
            wl_coef_y = rt.coef_y;
            wa_forward();
            for (size_t b = 0; b < rg.size(); ++b) {
                revert_to_snapshot(int(b));
                rg[b].min_sum_overflow = 1e30f;
                rg[b].hpwl_with_min_sum_overflow = 1e30f;
            }
            sync_stats();
        }
        if (need && revert_init) log.info(87, "Routability end iteration: increase inflation and revert back to snapshot.");
        if (!need && revert_init) log.info(89, "Routability finished. Reverting to minimal observed routing congestion, could not reach target.");
        if (!need && !revert_init) log.info(90, "Routability finished. Target routing congestion achieved succesfully.");
    }

    bool check_convergence(int b, int iter) {
        Region& r = rg[b];
        if (r.is_converged) return true;
        if (r.sum_overflow_unscaled > o.overflow) return false;
        int64_t hpwl = cur_hpwl;
        std::string group = r.R.group >= 0 ? strfmt("group%d", r.R.group) : std::string();
        log.report(strfmt("%9d | %8.4f | %13.6e | %8s | %9.2e | %5s", iter, r.sum_overflow_unscaled, (double)hpwl / (double)dbu, "",
                          (double)r.density_penalty, group.c_str()));
        log.report("---------------------------------------------------------------");
        if (r.R.group >= 0)
            log.info(1016, strfmt("Region '%s' placement finished at iteration %d", group.c_str(), iter));
        else {
            log.info(1001, strfmt("Global placement finished at iteration %d", iter));
            if (rt.on) log.info(1017, strfmt("Routability mode iteration count: %d", rt.route_iters));
        }
        if (rt.on) {
            rt.final_rc = measure_rudy();
            log.info(1005, strfmt("Routability final weighted congestion: %.4f", (double)rt.final_rc));
        }
        double dbu2 = double(dbu) * double(dbu);
        double w = double(r.R.white_space_area);
        log.info(1002, strfmt("Placed Cell Area            %10.4f", double(r.nia) / dbu2));
        log.info(1003, strfmt("Available Free Area         %10.4f", w / dbu2));
        log.info(1004, strfmt("Minimum Feasible Density        %.4f (cell_area / free_area)", r.R.uniform_target_density));
        log.info(1006, "  Suggested Target Densities:");
        log.info(1007, strfmt("    - For 90%% usage of free space: %.4f", double(r.nia) / (w * 0.90)));
        log.info(1008, strfmt("    - For 80%% usage of free space: %.4f", double(r.nia) / (w * 0.80)));
        if (double(r.nia) / (w * 0.50) <= 1.0) log.info(1009, strfmt("    - For 50%% usage of free space: %.4f", double(r.nia) / (w * 0.50)));
        if (r.R.uniform_target_density > 0.95f)
            log.warn(1015, strfmt("High uniform density (>%.2f) may cause congestion or legalization issues.", r.R.uniform_target_density));
        lock_region(b);
        r.is_converged = true;
        return true;
    }

    void push_history(int iter) {
        gpl_iter_row row{};
        row.iter = iter;
        row.trials = trials;
        row.overflow = average_overflow_unscaled;
        row.density_penalty = rg[0].density_penalty;
        row.step_length = rg[0].step_length;
        row.wl_coef = wl_coef_x;
        row.hpwl = rg[0].prev_hpwl;
        hist.push_back(row);
    }

    // resume_nesterov_place (nesterov.rs:4440), routability and timing off, in three pieces so
    // a caller can run it in slices: begin (the per-call resets), step (iterations), end (the
    // max-iteration snapshot rule and the closing report).
    struct LoopState {
        float cur_a = 1.0f;
        int32_t iter = 0;
        Clock::time_point t0;
        int64_t original_area = 0;
        bool active = false, done = false;
    } ls;

    int32_t loop_begin(int32_t start_iter) {
        ls = LoopState{};
        ls.t0 = Clock::now();
        log.info(84, "---- Execute Nesterov Global Placement.");
        if (num_region_diverged > 0) {
            log.error(diverge_code, diverge_msg);
            ls.done = true;
            return diverge_code;
        }
        for (auto& r : rg) {
            r.max_phi_coef_changed = false;
            r.min_sum_overflow = 1e30f;
            r.hpwl_with_min_sum_overflow = 1e30f;
            r.is_converged = false;
            ls.original_area += r.nia;
        }
        diverge_saved = false;
        diverge_reverts = 0;
        diverge_hpwl = 0;
        rt.saved = false;
        rt.attempts = 0;
        rt.route_iters = 0;
        rt.reverts = 0;
        converged = false;
        ls.iter = start_iter;
        ls.active = true;
        return GPL_OK;
    }

    // At most n iterations of the loop body (nesterov.rs:4478-4575).  Each pass is one iteration
    // (ls.iter + 1 at its end, or the loop ends); with device-driven trials the device runs whole
    // iterations until it halts, and the host finishes the one it halted in.
    void loop_step(int32_t n) {
        const int32_t stop_at = n >= INT32_MAX - ls.iter ? INT32_MAX : ls.iter + n;
        while (ls.active && !ls.done && ls.iter < stop_at) {
            float coeff = 0.0f;
            if (!iter_begin(coeff)) break;
            int resume = 0;
            bool booked = false;  // the iteration's bookkeeping (up to the history row) is done
            if (ctl_usable()) {
                const int stage = run_device_trials(coeff, stop_at);
                iter_now = ls.iter;
                if (stage == 3) continue;
                if (stage == 2) booked = true;
                if (stage == 1) {
                    resume = h_ctl->h.trials + 1;
                    coeff = h_ctl->h.coeff;
                }
            }
            if (!iter_rest(coeff, resume, booked, false)) break;
        }
    }

    // An iteration's start: the iteration limit, the momentum (a_k) and its coefficient; false
    // when the loop ends here.
    bool iter_begin(float& coeff) {
        if (ls.iter >= max_iter) {
            ls.done = true;
            return false;
        }
        float prev_a = ls.cur_a;
        ls.cur_a = (float)((1.0 + std::sqrt(4.0 * (double)prev_a * (double)prev_a + 1.0)) * 0.5);
        coeff = (float)(((double)prev_a - 1.0) / (double)ls.cur_a);
        iter_now = ls.iter;
        return true;
    }

    // The rest of the iteration: the back-tracking (from trial `resume` when the first trials ran
    // elsewhere; stats_on_host: that trial's statistics are already in the pinned copy), the
    // next-iteration bookkeeping unless `booked`, divergence, routability, convergence, the
    // re-sort; false when the loop ends here.
    bool iter_rest(float coeff, int resume, bool booked, bool stats_on_host) {
        if (!booked) {
            do_back_tracking(coeff, resume, stats_on_host);
            total_trials += trials;
            for (auto& r : rg) adjust_phi(r);
            if (num_region_diverged > 0) {
                ls.done = true;
                return false;
            }
            np_update_next_iter(ls.iter);
            push_history(ls.iter);
            last_iter = ls.iter;
            ++iters_run;
        }
        // is_routability_gpl_iter: each such iteration extends the limit by one
        if (rt.on && rt.saved && average_overflow_unscaled > o.routability_end_overflow) {
            rt.route_iters += 1;
            max_iter += 1;
        }
        int act = np_is_diverged(&ls.cur_a);
        if (act == 2) {
            ls.done = true;
            return false;
        }
        if (act == 1) {
            ++ls.iter;
            return true;
        }
        if (rt.on) routability_step(ls.iter, &ls.cur_a);
        size_t conv = 0;
        for (size_t b = 0; b < rg.size(); ++b) conv += check_convergence(int(b), ls.iter) ? 1 : 0;
        if (conv == rg.size()) {
            converged = true;
            t_converge = t_init + std::chrono::duration<double>(Clock::now() - ls.t0).count();
            ls.done = true;
            return false;
        }
        maybe_resort(ls.iter);
        ++ls.iter;
        return true;
    }

    // One trial's kernels (the trial graph's sequence: the update, the scatter, the bins and the
    // solve, the WA kernels, the gradients with the statistics and the completion signal to the
    // host) into the thread's recorder, for a grouped launch with other placements'.
    void record_trial(float coeff) {
        finish_to_host = true;
        signal_region = -1;
        for (size_t b = 0; b < rg.size(); ++b)
            if (!rg[b].is_converged) signal_region = int(b);
        update_coordinates(coeff);
        field();
        wa_forward(false, true);
        join();
        launch_gradients(0, int(rg.size()), SLOT_NEXT, true);
        finish_to_host = false;
    }

    int32_t loop_end(int32_t* iter_out) {
        if (!ls.active) {
            if (iter_out) *iter_out = ls.iter;
            return num_region_diverged > 0 ? diverge_code : GPL_OK;
        }
        ls.active = false;
        int32_t iter = ls.iter;
        if (iter >= max_iter && !o.disable_revert_if_diverge && (diverge_saved || diverge_reverts > 0) &&
            diverge_snapshot_overflow <= average_overflow_unscaled && diverge_hpwl < cur_hpwl) {
            log.warn(121, "Reached the maximum number of iterations, reverting to snapshot with min hpwl.");
            revert_to_diverge_snapshot();
        }
        if (iter >= max_iter)
            log.warn(1010, strfmt("GPL reached the maximum number of iterations for nesterov %d. Placement may have failed to converge.", max_iter));
        int64_t new_area = 0;
        for (auto& r : rg) new_area += r.nia;
        double dbu2 = double(dbu) * double(dbu);
        if (rt.on) {
            log.info(1011, strfmt("Original area (um^2): %.2f", double(ls.original_area) / dbu2));
            int64_t inflation = 0;
            for (int64_t a : rt.accumulated) inflation += a;
            const float pct = (float)(100.0 * double(inflation) / double(ls.original_area));
            log.info(1012, strfmt("Total routability artificial inflation: %.2f (%+.2f%%)", double(inflation) / dbu2, (double)pct));
        }
        float diff = (float)(100.0 * double(new_area - ls.original_area) / double(ls.original_area));
        log.info(1014, strfmt("Final placement area: %.2f (%+.2f%%)", double(new_area) / dbu2, diff));
        check(hipStreamSynchronize(stream), "sync");
        t_loop = std::chrono::duration<double>(Clock::now() - ls.t0).count();
        if (iter_out) *iter_out = iter;
        if (num_region_diverged > 0) {
            log.error(diverge_code, diverge_msg);
            return diverge_code;
        }
        return GPL_OK;
    }

    int32_t resume(int32_t start_iter, int32_t* iter_out) {
        int32_t rc = loop_begin(start_iter);
        if (rc != GPL_OK) {
            if (iter_out) *iter_out = start_iter;
            ls.active = false;
            return rc;
        }
        loop_step(INT32_MAX);
        return loop_end(iter_out);
    }
};

extern "C" __global__ void gpl_k_lock_region(gplk::RegionDev r, uint8_t* flags) {
    int n = r.cell_end - r.cell_begin;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint8_t f = flags[r.cell_begin + k];
        if (f & 1u) flags[r.cell_begin + k] = f | 4u;
    }
}

void gpl_ctx::lock_region(int b) {
    Region& r = rg[b];
    if (r.n > 0) hipLaunchKernelGGL(gpl_k_lock_region, dim3((r.n + 255) / 256), dim3(256), 0, stream, r.dev, d_flags);
    r.has_locked = true;
}

namespace {

// Twiddle tables of the length-L transforms (h = L/2), computed in double and rounded once.
void make_tables(int L, std::vector<float2>& tw, std::vector<float2>& th, std::vector<float2>& dw) {
    const double pi = 3.14159265358979323846;
    const int h = L / 2;
    tw.resize(h);
    th.resize(size_t(h) + 1);
    dw.resize(L);
    for (int m = 0; m < h; ++m) tw[m] = make_float2((float)std::cos(-2.0 * pi * m / h), (float)std::sin(-2.0 * pi * m / h));
    for (int k = 0; k <= h; ++k) th[k] = make_float2((float)std::cos(-2.0 * pi * k / L), (float)std::sin(-2.0 * pi * k / L));
    for (int k = 0; k < L; ++k) dw[k] = make_float2((float)std::cos(pi * k / (2.0 * L)), (float)-std::sin(pi * k / (2.0 * L)));
}

bool pow2(int v) { return v >= 2 && (v & (v - 1)) == 0; }
// a bin count the solve transforms: 2^k (2..2048), or 3 * 2^k (24..1536: the row FFT's half
// length then takes one radix-3 stage)
bool fft_len_ok(int v) { return v <= 2048 && (pow2(v) || (v % 3 == 0 && pow2(v / 3) && v / 3 >= 8)); }

int32_t fail(gpl_ctx* c, int32_t code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

}  // namespace

extern "C" {

int32_t gpl_ctx_create(const gpl_problem* p, gpl_ctx** out) {
    if (!p || !out) return GPL_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        fprintf(stderr, "gpl_ctx_create: no HIP device\n");
        return GPL_ERR_HIP;
    }
    for (int b = 0; b < p->num_regions; ++b) {
        const gpl_region& R = p->regions[b];
        if (!fft_len_ok(R.cnt_x) || !fft_len_ok(R.cnt_y)) {
            fprintf(stderr, "gpl_ctx_create: bin grid %d x %d: each side must be 2^k (2..2048) or 3 * 2^k (24..1536)\n", R.cnt_x,
                    R.cnt_y);
            return GPL_ERR_REFUSED;
        }
    }
    gpl_ctx* c = new gpl_ctx();
    try {
        check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream");
        check(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking), "stream");
        check(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "event");
        check(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "event");
        check(hipStreamCreateWithFlags(&c->aux2, hipStreamNonBlocking), "stream");
        check(hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming), "event");
        check(hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming), "event");
        c->o = p->opts;
        c->max_iter = p->opts.max_iter;
        c->dbu = p->dbu;
        c->site_x = p->site_x;
        c->ncells = p->num_cells;
        c->npins = p->num_pins;
        c->nnets = p->num_nets;
        const int32_t N = p->num_cells, E = p->num_nets;
        // Cells in Morton order of their initial density-box centre's bin, within each region:
        // a block of consecutive cells then covers a compact window of bins, which the scatter
        // accumulates in LDS.  new_of[old] / old_of[new]; the C ABI's per-cell vectors stay in
        // the problem's order (permuted at the boundary).
        c->cell_old_of.resize(N);
        c->cell_new_of.resize(N);
        std::iota(c->cell_old_of.begin(), c->cell_old_of.end(), 0);
        for (int32_t b = 0; b < p->num_regions; ++b) {
            const gpl_region& R = p->regions[b];
            std::vector<uint64_t> key(size_t(R.cell_end - R.cell_begin));
            for (int32_t g = R.cell_begin; g < R.cell_end; ++g) {
                const int32_t* d = &p->cell_dbox[4 * g];
                double cx = 0.5 * (double(d[0]) + double(d[2])), cy = 0.5 * (double(d[1]) + double(d[3]));
                uint32_t bx = uint32_t(std::min(std::max((cx - R.lx) / R.size_x, 0.0), double(R.cnt_x - 1)));
                uint32_t by = uint32_t(std::min(std::max((cy - R.ly) / R.size_y, 0.0), double(R.cnt_y - 1)));
                uint64_t m = 0;
                for (int bit = 0; bit < 16; ++bit) m |= (uint64_t((bx >> bit) & 1u) << (2 * bit)) | (uint64_t((by >> bit) & 1u) << (2 * bit + 1));
                key[g - R.cell_begin] = m;
            }
            std::stable_sort(c->cell_old_of.begin() + R.cell_begin, c->cell_old_of.begin() + R.cell_end,
                             [&](int32_t a, int32_t bb) { return key[a - R.cell_begin] < key[bb - R.cell_begin]; });
        }
        for (int32_t i = 0; i < N; ++i) c->cell_new_of[c->cell_old_of[i]] = i;
        auto newc = [&](int32_t old) { return old >= 0 ? c->cell_new_of[old] : -1; };
        // nets sorted by windows of 4096 cells of their first pin's cell, then by degree (so that
        // a wave's lanes loop equally), then by the first cell: the nets in flight at once touch
        // one window's cells, whose centres and gradient terms stay in L2; nets above 128 pins
        // last, one block each.  The small nets are then cut into 8 contiguous slices of equal pin
        // counts, one per XCD.
        std::vector<int32_t> order(E);
        std::iota(order.begin(), order.end(), 0);
        auto deg = [&](int32_t e) { return p->net_pin_start[e + 1] - p->net_pin_start[e]; };
        auto first_cell = [&](int32_t e) {
            int32_t best = INT32_MAX;
            for (int32_t k = p->net_pin_start[e]; k < p->net_pin_start[e + 1]; ++k) {
                int32_t q = p->net_pins[k];
                if (p->pin_cell[q] >= 0) best = std::min(best, newc(p->pin_cell[q]));
            }
            return best;
        };
        std::vector<int32_t> fc(E);
        for (int32_t e = 0; e < E; ++e) fc[e] = first_cell(e);
        const int32_t kBig = 128;
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
            const bool ba = deg(a) > kBig, bb = deg(b) > kBig;
            if (ba != bb) return bb;
            if ((fc[a] >> 12) != (fc[b] >> 12)) return fc[a] < fc[b];
            return deg(a) != deg(b) ? deg(a) < deg(b) : fc[a] < fc[b];
        });
        c->net_orig_of = order;
        c->net_sorted_of.assign(E, 0);
        for (int32_t i = 0; i < E; ++i) c->net_sorted_of[order[i]] = i;
        int32_t num_small = 0;
        while (num_small < E && deg(order[num_small]) <= kBig) ++num_small;
        {
            // the WA tiles: ranges of kTileSpan small-net pins (bounds found on the device, where a
            // net re-sort recomputes them), XCD x taking a contiguous eighth
            int64_t small_pins = 0;
            for (int32_t i = 0; i < num_small; ++i) small_pins += deg(order[i]);
            c->nets.num_tiles = int32_t((small_pins + gplk::kTileSpan - 1) / gplk::kTileSpan);
            c->nets.tile_blocks = std::max(1, (c->nets.num_tiles + 7) / 8);
            c->d_tile = c->dalloc<int32_t>(size_t(c->nets.num_tiles) + 1);
            c->nets.tile = c->d_tile;
        }
        std::vector<int32_t> nstart(E + 1, 0), pcell, np_net;
        std::vector<int2> poff, pxy0;
        std::vector<float> nweight(E);
        std::vector<int32_t> npin_of(size_t(p->num_pins), -1);  // problem pin -> net-order pin index
        for (int32_t i = 0; i < E; ++i) {
            int32_t e = order[i];
            nweight[i] = p->net_weight[e];
            for (int32_t k = p->net_pin_start[e]; k < p->net_pin_start[e + 1]; ++k) {
                int32_t q = p->net_pins[k];
                npin_of[q] = int32_t(pcell.size());
                pcell.push_back(newc(p->pin_cell[q]));
                pxy0.push_back(make_int2(p->pin_xy[2 * q], p->pin_xy[2 * q + 1]));
                // a pin on no movable cell keeps its fixed position in the offset slot
                poff.push_back(pcell.back() >= 0 ? make_int2(p->pin_offset[2 * q], p->pin_offset[2 * q + 1]) : pxy0.back());
            }
            nstart[i + 1] = int32_t(pcell.size());
        }
        c->nets.num_nets = E;
        c->nets.num_small = num_small;
        c->nets.start = c->upload(nstart.data(), nstart.size());
        c->nets.pcell = c->upload(pcell.data(), pcell.size());
        c->nets.poff = c->upload(poff.data(), poff.size());
        c->nets.pxy0 = c->upload(pxy0.data(), pxy0.size());
        c->nets.weight = c->upload(nweight.data(), nweight.size());
        {
            std::vector<gplk::NetRec> rec(E);
            for (int32_t i = 0; i < E; ++i) {
                std::memset(&rec[i], 0, sizeof rec[i]);
                rec[i].w = nweight[i];
            }
            c->nets.rec = c->upload(rec.data(), rec.size());
        }
        // cells, in the new order
        std::vector<int4> dbox(N);
        std::vector<float2> pre(N);
        std::vector<float> dscale(N);
        std::vector<uint8_t> flags(N);
        std::vector<int32_t> pin_start(size_t(N) + 1, 0), cell_inst(N);
        std::vector<int32_t> cp_npin(p->cell_pin_start[N]);
        std::vector<int32_t> cpi(pcell.size(), -1);
        for (int32_t i = 0; i < N; ++i) {
            int32_t g = c->cell_old_of[i];
            const int32_t* d = &p->cell_dbox[4 * g];
            dbox[i] = make_int4(d[0], d[1], d[2], d[3]);
            const int32_t* bx = &p->cell_box[4 * g];
            pre[i] = make_float2((float)(p->cell_pin_start[g + 1] - p->cell_pin_start[g]), (float)(bx[2] - bx[0]) * (float)(bx[3] - bx[1]));
            dscale[i] = p->density_scale[g];
            flags[i] = p->cell_flags[g];
            cell_inst[i] = p->cell_inst[g];
            int32_t at = pin_start[i];
            for (int32_t k = p->cell_pin_start[g]; k < p->cell_pin_start[g + 1]; ++k, ++at) {
                int32_t q = p->cell_pins[k];
                cp_npin[at] = npin_of[q];
                cpi[npin_of[q]] = at;
            }
            pin_start[i + 1] = at;
        }
        c->h_flags = flags;
        c->cells.dbox = c->upload(dbox.data(), N);
        c->cells.dscale = c->upload(dscale.data(), N);
        c->d_flags = c->upload(flags.data(), N);
        c->cells.flags = c->d_flags;
        {
            std::vector<int2> ctr(N, make_int2(INT_MIN, INT_MIN));
            c->cells.ctr = c->upload(ctr.data(), N);
        }

        c->cells.pin_start = c->upload(pin_start.data(), size_t(N) + 1);
        c->cells.cp_npin = c->upload(cp_npin.data(), cp_npin.size());

        c->nets.cpi = c->upload(cpi.data(), cpi.size());
        c->nets.pgrad = c->dalloc<float2>(std::max<size_t>(cp_npin.size(), 1));
        gplk::launch_tile_bounds(c->stream, c->nets, c->d_tile);
        // the net re-sort's buffers and second copy of every net array
        {
            gplk::NetResortBufs& b = c->nrb;
            const int32_t Es = std::max(1, num_small);
            const size_t P = std::max<size_t>(pcell.size(), 1);
            b.keys_in = c->dalloc<unsigned long long>(Es);
            b.keys_out = c->dalloc<unsigned long long>(Es);
            b.idx_in = c->dalloc<int32_t>(Es);
            b.perm = c->dalloc<int32_t>(Es);
            b.deg = c->dalloc<int32_t>(size_t(Es) + 1);
            b.temp_bytes = gplk::net_resort_temp_bytes(Es);
            b.temp = c->dalloc<uint8_t>(b.temp_bytes);
            int32_t wbits = 1;
            while ((int64_t(1) << wbits) <= (int64_t(N) >> 12) + 1) ++wbits;
            b.end_bit = 8 + wbits;
            b.start = c->dalloc<int32_t>(size_t(E) + 1);
            b.pcell = c->dalloc<int32_t>(P);
            b.poff = c->dalloc<int2>(P);
            b.pxy0 = c->dalloc<int2>(P);
            b.cpi = c->dalloc<int32_t>(P);
            b.weight = c->dalloc<float>(std::max(1, E));
            b.orig = c->dalloc<int32_t>(std::max(1, E));
            c->d_net_orig = c->upload(order.data(), order.size());
        }
        c->cells.pre = c->upload(pre.data(), N);
        c->cell_inst = cell_inst;
        c->cell_inst_prob.assign(p->cell_inst, p->cell_inst + N);
        c->inst_box.assign(p->inst_box, p->inst_box + 4 * size_t(p->num_insts));
        c->cells.old_of = c->upload(c->cell_old_of.data(), N);
        // the re-sort's second copy of every cell array and its work buffers
        {
            gplk::CellsDev& a = c->cells_alt;
            a.dbox = c->dalloc<int4>(N);
            a.ctr = c->dalloc<int2>(N);
            a.dscale = c->dalloc<float>(N);
            a.flags = c->dalloc<uint8_t>(N);
            a.pin_start = c->dalloc<int32_t>(size_t(N) + 1);
            a.cp_npin = c->dalloc<int32_t>(cp_npin.size());

            a.pre = c->dalloc<float2>(N);
            a.old_of = c->dalloc<int32_t>(N);
            c->rb.keys_in = c->dalloc<uint32_t>(N);
            c->rb.keys_out = c->dalloc<uint32_t>(N);
            c->rb.idx_in = c->dalloc<int32_t>(N);
            c->rb.perm = c->dalloc<int32_t>(N);
            c->rb.inv = c->dalloc<int32_t>(N);
            c->rb.cnt = c->dalloc<int32_t>(size_t(N) + 1);
            c->rb.temp_bytes = gplk::resort_temp_bytes(N);
            c->rb.temp = c->dalloc<uint8_t>(c->rb.temp_bytes);
            c->net_pin_count = int32_t(pcell.size());
        }
        // regions
        c->rg.resize(p->num_regions);
        for (int32_t b = 0; b < p->num_regions; ++b) {
            Region& r = c->rg[b];
            r.R = p->regions[b];
            r.n = r.R.cell_end - r.R.cell_begin;
            r.max_phi_coef = c->o.max_phi_coef;
            r.nia = c->nesterov_insts_area(r);
            for (int32_t g = r.R.cell_begin; g < r.R.cell_end; ++g)
                if (p->cell_flags[g] & GPL_CELL_LOCKED) r.has_locked = true;
            gplk::RegionDev& d = r.dev;
            d.lx = r.R.lx;
            d.ly = r.R.ly;
            d.ux = r.R.ux;
            d.uy = r.R.uy;
            d.cnt_x = r.R.cnt_x;
            d.cnt_y = r.R.cnt_y;
            d.size_x = r.R.size_x;
            d.size_y = r.R.size_y;
            d.target_density = r.R.target_density;
            d.cell_begin = r.R.cell_begin;
            d.cell_end = r.R.cell_end;
            {
                // a chunk's sum in one bin is at most kPlaceChunk x the bin's area
                int64_t wmax = 0, hmax = 0;
                for (int32_t x = 0; x < r.R.cnt_x; ++x) wmax = std::max<int64_t>(wmax, int64_t(r.R.bin_x[x + 1]) - r.R.bin_x[x]);
                for (int32_t y = 0; y < r.R.cnt_y; ++y) hmax = std::max<int64_t>(hmax, int64_t(r.R.bin_y[y + 1]) - r.R.bin_y[y]);
                d.narrow = (double)gplk::kPlaceChunk * (double)wmax * (double)hmax < 4294967296.0 ? 1 : 0;
                r.narrow_ok = d.narrow;
                d.chunk = gplk::kPlaceChunk;
            }
            {
                // the kernels compute the edges themselves when the tables are init_bins' own
                bool rounded = true;
                for (int32_t x = 0; x <= r.R.cnt_x; ++x)
                    rounded = rounded && r.R.bin_x[x] == d.lx + (int32_t)std::round(double(x) * d.size_x);
                for (int32_t y = 0; y <= r.R.cnt_y; ++y)
                    rounded = rounded && r.R.bin_y[y] == d.ly + (int32_t)std::round(double(y) * d.size_y);
                d.edges_rounded = rounded ? 1 : 0;
                r.edges_rounded = d.edges_rounded;
                d.grad_stage = gplk::kGradStage;
            }
            size_t M = size_t(r.R.cnt_x) * r.R.cnt_y;
            d.bin_x = c->upload(r.R.bin_x, size_t(r.R.cnt_x) + 1);
            d.bin_y = c->upload(r.R.bin_y, size_t(r.R.cnt_y) + 1);
            d.inst = c->dalloc<unsigned long long>(M);
            d.filler = c->dalloc<unsigned long long>(M);
            d.non_place = (const long long*)c->upload(r.R.non_place, M);
            {
                // int64 -> float rounds to nearest even on the host as on the device
                std::vector<float2> npf(M);
                for (size_t k = 0; k < M; ++k)
                    npf[k] = make_float2((float)(long long)r.R.non_place[k], (float)(long long)r.R.non_place_unscaled[k]);
                d.non_place_f = c->upload(npf.data(), M);
            }
            d.density = c->dalloc<float>(M);
            d.field = c->dalloc<float2>(M);
            d.phi = c->dalloc<float>(M);
            r.sb.work0 = c->dalloc<float>(2 * M);  // the solve's (Ex', Ey') pairs
            r.sb.work1 = c->dalloc<float>(M);
            r.sb.work2 = c->dalloc<float>(M);
            r.sb.work3 = c->dalloc<float>(M);
            r.sb.wx = c->upload(r.R.wx, r.R.cnt_x);
            r.sb.wy = c->upload(r.R.wy, r.R.cnt_y);
            std::vector<float2> tw, th, dw;
            make_tables(r.R.cnt_x, tw, th, dw);
            r.sb.tw_x = c->upload(tw.data(), tw.size());
            r.sb.th_x = c->upload(th.data(), th.size());
            r.sb.dw_x = c->upload(dw.data(), dw.size());
            make_tables(r.R.cnt_y, tw, th, dw);
            r.sb.tw_y = c->upload(tw.data(), tw.size());
            r.sb.th_y = c->upload(th.data(), th.size());
            r.sb.dw_y = c->upload(dw.data(), dw.size());
            for (int s = 0; s < S_COUNT; ++s) {
                r.v[s] = c->dalloc<float2>(r.n);
                check(hipMemsetAsync(r.v[s], 0, sizeof(float2) * std::max(1, r.n), c->stream), "memset");
            }
            r.partials = c->dalloc<double>(6 * gplk::kPartials);
            check(hipMemsetAsync(r.partials, 0, sizeof(double) * 6 * gplk::kPartials, c->stream), "memset");
            gplk::launch_zero_bins(c->stream, d);
        }
        c->tune.assign(p->num_regions, Tune{});
        for (int32_t b = 0; b < p->num_regions; ++b) c->tune[b].narrow = c->rg[b].narrow_ok;
        if (p->num_regions > 1) {
            // region groups: grouped trials (launch mode 7) by default -- every region's small
            // kernels in one launch per step instead of a graph node each (9 regions: 612 -> 328 us
            // per iteration; one region keeps the trial graph, 5260 against 4490 it/s at 1M cells)
            c->grouped_trials = true;
            c->launch_mode = 2;
            c->overlap = false;
        }
        // one Sums per region, then one whose hpwl field is the HPWL counter the WA kernels add into
        c->d_sums = c->dalloc<gplk::Sums>(size_t(p->num_regions) + 1);
        check(hipMemsetAsync(c->d_sums, 0, sizeof(gplk::Sums) * (size_t(p->num_regions) + 1), c->stream), "memset");
        check(hipHostMalloc((void**)&c->h_sums, sizeof(gplk::Sums) * (size_t(p->num_regions) + 1), hipHostMallocDefault), "pinned");
        c->d_hpwl = &c->d_sums[p->num_regions].hpwl;
        c->d_seq = c->dalloc<unsigned long long>(1);
        check(hipMemsetAsync(c->d_seq, 0, sizeof(unsigned long long), c->stream), "memset");
        // [0] the sequence number, [1] the device-driven trials' halt event
        check(hipHostMalloc((void**)&c->h_seq, 2 * sizeof(unsigned long long), hipHostMallocDefault), "pinned");
        c->h_seq[0] = c->h_seq[1] = 0ull;
        c->d_global_adds = (unsigned long long*)&c->d_sums[p->num_regions].pad[0];

        // the region table of the re-sort; per-region scratch vectors for its gathers
        c->resort_enabled = p->num_regions <= gplk::kMaxRegions;
        c->rtab.n = std::min(p->num_regions, gplk::kMaxRegions);
        for (int32_t b = 0; b < c->rtab.n; ++b) {
            const gpl_region& R = p->regions[b];
            c->rtab.begin[b] = R.cell_begin;
            c->rtab.begin[b + 1] = R.cell_end;
            c->rtab.lx[b] = R.lx;
            c->rtab.ly[b] = R.ly;
            c->rtab.cnt_x[b] = R.cnt_x;
            c->rtab.cnt_y[b] = R.cnt_y;
            c->rtab.size_x[b] = R.size_x;
            c->rtab.size_y[b] = R.size_y;
        }
        {
            // the sort key: the region index above the Morton code of the cell's 4 x 4-bin square
            int32_t maxc = 1;
            for (int32_t b = 0; b < c->rtab.n; ++b) maxc = std::max({maxc, c->rtab.cnt_x[b], c->rtab.cnt_y[b]});
            int32_t side = (maxc + (1 << gplk::kResortShift) - 1) >> gplk::kResortShift, bits = 1;
            while ((1 << bits) < side) ++bits;
            int32_t rbits = 0;
            while ((1 << rbits) < c->rtab.n) ++rbits;
            c->rtab.morton_bits = 2 * bits;
            c->rtab.end_bit = 2 * bits + 1 + rbits;
            if (c->rtab.end_bit > 32 || bits > 11) c->resort_enabled = false;
        }
        for (int32_t b = 0; b < p->num_regions; ++b) {
            c->region_scratch.emplace_back();
            for (int i = 0; i < gplk::kGatherMax; ++i) c->region_scratch.back().push_back(c->dalloc<float2>(std::max(1, c->rg[b].n)));
        }
        c->cell_flags_prob.assign(p->cell_flags, p->cell_flags + N);
        if (p->opts.routability) {
            // the cells routability takes out are parked by re-sorts
            if (!c->resort_enabled) {
                fprintf(stderr, "gpl_ctx_create: routability needs the device re-sort (at most %d regions)\n", gplk::kMaxRegions);
                delete c;
                return GPL_ERR_REFUSED;
            }
            c->route_setup(p);
        }
        check(hipStreamSynchronize(c->stream), "sync");
    } catch (const HipError& e) {
        fprintf(stderr, "gpl_ctx_create: %s: %s\n", e.what, hipGetErrorString(e.e));
        delete c;
        return GPL_ERR_HIP;
    }
    *out = c;
    return GPL_OK;
}

void gpl_ctx_destroy(gpl_ctx* c) { delete c; }

const char* gpl_last_error(const gpl_ctx* c) { return c ? c->err.c_str() : "null context"; }

int32_t gpl_init_nesterov_loop(gpl_ctx* c) {
    if (!c) return GPL_ERR_ARG;
    try {
        c->autotune_init();  // kernel variants picked per region (untimed: part of the set-up)
        auto t0 = Clock::now();
        int64_t design_hpwl = c->hpwl_now();
        float user = c->o.reference_hpwl;
        c->reference_hpwl = user > 0.0f ? user : std::max(446000000.0f, 0.01f * (float)design_hpwl);
        int32_t rc = c->np_init();
        c->initialized = true;
        check(hipStreamSynchronize(c->stream), "sync");
        c->t_init = std::chrono::duration<double>(Clock::now() - t0).count();
        if (rc != GPL_OK) {
            c->num_region_diverged = 1;
            c->diverge_code = rc;
        }
        return rc;
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
}

int32_t gpl_run_nesterov_loop(gpl_ctx* c, int32_t start_iter, int32_t* iter_out) {
    if (!c) return GPL_ERR_ARG;
    if (!c->initialized) return fail(c, GPL_ERR_STATE, "gpl_init_nesterov_loop first");
    try {
        return c->resume(start_iter, iter_out);
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
}

int32_t gpl_loop_begin(gpl_ctx* c, int32_t start_iter) {
    if (!c) return GPL_ERR_ARG;
    if (!c->initialized) return fail(c, GPL_ERR_STATE, "gpl_init_nesterov_loop first");
    try {
        return c->loop_begin(start_iter);
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
}

int32_t gpl_loop_step(gpl_ctx* c, int32_t n, int32_t* iter_out, int32_t* done) {
    if (!c) return GPL_ERR_ARG;
    if (!c->ls.active) return fail(c, GPL_ERR_STATE, "gpl_loop_begin first");
    try {
        c->loop_step(n);
        check(hipStreamSynchronize(c->stream), "sync");
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    if (iter_out) *iter_out = c->ls.iter;
    if (done) *done = c->ls.done ? 1 : 0;
    return GPL_OK;
}

int32_t gpl_loop_end(gpl_ctx* c, int32_t* iter_out) {
    if (!c) return GPL_ERR_ARG;
    try {
        return c->loop_end(iter_out);
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
}

int32_t gpl_do_nesterov_place(gpl_ctx* c, int32_t* iter_out) {
    int32_t rc = gpl_init_nesterov_loop(c);
    if (rc < 0) return rc;
    return gpl_run_nesterov_loop(c, 0, iter_out);
}

// Independent placements in lockstep: each iteration's first trial of every context as one
// launch per kernel (the contexts' records grouped by sequence), the rest of each iteration by its
// own context; the contexts share one stream meanwhile.  Each context's result is its own
// gpl_do_nesterov_place's.
int32_t gpl_place_batch(gpl_ctx* const* cs, int32_t n, int32_t* iters) {
    if (!cs || n <= 0) return GPL_ERR_ARG;
    for (int32_t i = 0; i < n; ++i) {
        if (!cs[i]) return GPL_ERR_ARG;
        for (int32_t j = 0; j < i; ++j)
            if (cs[j] == cs[i]) return GPL_ERR_ARG;
    }
    bool plain = true;  // every context can take part (else each runs on its own)
    for (int32_t i = 0; i < n; ++i) plain = plain && !cs[i]->rt.on && !cs[i]->profiling;
    if (!plain) {
        int32_t rc = GPL_OK;
        for (int32_t i = 0; i < n; ++i) {
            int32_t it = 0;
            const int32_t r = gpl_do_nesterov_place(cs[i], &it);
            if (iters) iters[i] = it;
            if (r < 0 && rc == GPL_OK) rc = r;
        }
        return rc;
    }
    for (int32_t i = 0; i < n; ++i) {
        const int32_t rc = gpl_init_nesterov_loop(cs[i]);
        if (rc < 0) return rc;
    }
    struct Saved {
        hipStream_t stream, aux, aux2;
        int launch_mode;
        bool overlap, device_trials;
    };
    std::vector<Saved> saved(n);
    hipStream_t bs = nullptr, side = nullptr;  // the contexts' stream; the grouped WA kernels'
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // The contexts' round: the live ones and those whose first trial ran grouped; the grouped
    // launches' arguments (pinned staging) and their device copy.
    struct Round {
        std::vector<int32_t> members, live, grouped;
        unsigned char* hstage = nullptr;
        unsigned char* dstage = nullptr;
        size_t cap = 0;
        bool pending = false;  // a launched trial not yet finished
    } L;
    int32_t rc = GPL_OK;
    gpl_ctx* failing = cs[0];
    try {
        check(hipStreamCreate(&bs), "stream");
        check(hipStreamCreate(&side), "stream");
        check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "event");
        check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "event");
        for (int32_t i = 0; i < n; ++i) {
            gpl_ctx* c = cs[i];
            c->join();
            check(hipStreamSynchronize(c->stream), "sync");
            c->drop_graph();
            saved[i] = Saved{c->stream, c->aux, c->aux2, c->launch_mode, c->overlap, c->device_trials};
            c->stream = c->aux = c->aux2 = bs;
            c->launch_mode = 2;
            c->overlap = false;
            c->device_trials = false;
            c->loop_begin(0);
        }
        // (two rounds in turn -- one on the GPU while the host finishes the other -- measured
        // slower: each round's launches cost the host as much as the whole batch's)
        for (int32_t i = 0; i < n; ++i) L.members.push_back(i);
        std::vector<gplk::TrialRecorder> rec(n);
        std::vector<float> coeff(n, 0.0f);
        std::vector<const gplk::TrialRecorder*> grp;
        // an iteration's start for the live contexts: the first trials recorded, then launched
        // grouped (one launch per kernel variant and step over every context and region)
        auto launch_round = [&]() {
            L.live.clear();
            L.grouped.clear();
            for (int32_t i : L.members) {
                gpl_ctx* c = cs[i];
                if (c->ls.active && !c->ls.done && c->iter_begin(coeff[i])) L.live.push_back(i);
            }
            L.pending = !L.live.empty();
            for (int32_t i : L.live) {
                failing = cs[i];
                rec[i].recs.clear();
                rec[i].failed = false;
                gplk::set_recorder(&rec[i]);
                try {
                    cs[i]->record_trial(coeff[i]);
                } catch (...) {
                    gplk::set_recorder(nullptr);
                    throw;
                }
                gplk::set_recorder(nullptr);
            }
            grp.clear();
            for (int32_t i : L.live)
                if (!rec[i].failed) {
                    grp.push_back(&rec[i]);
                    L.grouped.push_back(i);
                }
            if (grp.empty()) return;
            const size_t need = gplk::grouped_arg_bytes(grp.data(), int(grp.size()));
            if (need > L.cap) {
                check(hipStreamSynchronize(bs), "sync");
                if (L.hstage) hipHostFree(L.hstage);
                if (L.dstage) hipFree(L.dstage);
                L.hstage = nullptr;
                L.dstage = nullptr;
                L.cap = 0;
                check(hipHostMalloc((void**)&L.hstage, need, hipHostMallocDefault), "hipHostMalloc");
                check(hipMalloc((void**)&L.dstage, need), "hipMalloc");
                L.cap = need;
            }
            failing = cs[L.grouped[0]];
            check(gplk::launch_grouped(bs, grp.data(), int(grp.size()), L.hstage, L.dstage, side, ev_fork, ev_join), "grouped trial");
        };
        // the rest of each iteration by its own context (a context whose record failed: the whole
        // iteration there)
        auto finish_round = [&]() {
            for (int32_t i : L.grouped) cs[i]->wait_trial_signal();
            for (int32_t i : L.live) {
                failing = cs[i];
                const bool done_here = std::find(L.grouped.begin(), L.grouped.end(), i) != L.grouped.end();
                cs[i]->iter_rest(coeff[i], done_here ? 1 : 0, false, done_here);
            }
            L.pending = false;
        };
        for (launch_round(); L.pending; launch_round()) finish_round();
    } catch (const HipError& e) {
        gplk::set_recorder(nullptr);
        rc = fail(failing, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    if (bs) hipStreamSynchronize(bs);
    for (int32_t i = 0; i < n; ++i) {
        gpl_ctx* c = cs[i];
        if (!saved[i].stream) continue;
        c->stream = saved[i].stream;
        c->aux = saved[i].aux;
        c->aux2 = saved[i].aux2;
        c->launch_mode = saved[i].launch_mode;
        c->overlap = saved[i].overlap;
        c->device_trials = saved[i].device_trials;
        if (rc == GPL_OK) {
            int32_t it = 0;
            try {
                const int32_t r = c->loop_end(&it);
                if (r < 0 && rc == GPL_OK) rc = r;
            } catch (const HipError& e) {
                rc = fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
            }
            if (iters) iters[i] = it;
        }
    }
    if (side) hipStreamSynchronize(side);
    if (L.hstage) hipHostFree(L.hstage);
    if (L.dstage) hipFree(L.dstage);
    if (ev_fork) hipEventDestroy(ev_fork);
    if (ev_join) hipEventDestroy(ev_join);
    if (side) hipStreamDestroy(side);
    if (bs) hipStreamDestroy(bs);
    return rc;
}

int32_t gpl_continue_nesterov_place(gpl_ctx* c, float overflow, int32_t max_iter, int32_t start_iter, int32_t* iter_out) {
    if (!c) return GPL_ERR_ARG;
    c->o.overflow = overflow;
    c->max_iter = max_iter;
    return gpl_run_nesterov_loop(c, start_iter, iter_out);
}

int32_t gpl_get_stats(gpl_ctx* c, gpl_stats* s) {
    if (!c || !s) return GPL_ERR_ARG;
    std::memset(s, 0, sizeof *s);
    s->iter = c->last_iter;
    s->converged = c->converged ? 1 : 0;
    s->status = c->num_region_diverged > 0 ? c->diverge_code : 0;
    s->total_trials = c->total_trials;
    s->hpwl = c->cur_hpwl;
    s->overflow = c->average_overflow_unscaled;
    s->overflow_scaled = c->average_overflow;
    s->density_penalty = c->rg[0].density_penalty;
    s->step_length = c->rg[0].step_length;
    s->wl_coef = c->wl_coef_x;
    s->base_wl_coef = c->base_wl_coef;
    s->reference_hpwl = c->reference_hpwl;
    s->wl_grad_sum = c->rg[0].wl_grad_sum;
    s->density_grad_sum = c->rg[0].density_grad_sum;
    s->num_diverge_reverts = c->diverge_reverts;
    s->init_seconds = c->t_init;
    s->loop_seconds = c->t_loop;
    s->first_below_target_seconds = c->t_converge;
    return GPL_OK;
}

int32_t gpl_get_history(gpl_ctx* c, gpl_iter_row* rows, int32_t cap) {
    if (!c) return GPL_ERR_ARG;
    int32_t n = std::min<int32_t>(cap, int32_t(c->hist.size()));
    if (rows && n > 0) std::memcpy(rows, c->hist.data(), sizeof(gpl_iter_row) * n);
    return int32_t(c->hist.size());
}

int32_t gpl_get_log(gpl_ctx* c, char* buf, int32_t cap) {
    if (!c) return GPL_ERR_ARG;
    std::string s = c->log.joined();
    if (buf && cap > 0) {
        size_t n = std::min<size_t>(size_t(cap) - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return int32_t(s.size());
}

int32_t gpl_placed_locations(gpl_ctx* c, int32_t* out, int32_t cap) {
    if (!c) return GPL_ERR_ARG;
    try {
        c->ensure_host_map();
        std::vector<int4> dbox(c->ncells);
        check(hipMemcpyAsync(dbox.data(), c->cells.dbox, sizeof(int4) * c->ncells, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipStreamSynchronize(c->stream), "sync");
        int32_t n = 0;
        for (int32_t g_old = 0; g_old < c->ncells; ++g_old) {
            int32_t g = c->cell_new_of[g_old];
            int32_t k = c->cell_inst[g];
            if (k < 0) continue;
            if (out && n < cap) {
                const int32_t* b = &c->inst_box[4 * size_t(k)];
                int32_t dcx = (dbox[g].x + dbox[g].z) / 2, dcy = (dbox[g].y + dbox[g].w) / 2;
                out[3 * n] = k;
                out[3 * n + 1] = dcx - (b[2] - b[0]) / 2 + c->site_x * c->o.pad_left;
                out[3 * n + 2] = dcy - (b[3] - b[1]) / 2;
            }
            ++n;
        }
        return n;
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
}

// ---- kernel-level entries ---------------------------------------------------------------------

static float2* slot_ptr(gpl_ctx* c, int region, int slot) {
    Region& r = c->rg[region];
    switch (slot) {
    case GPL_SLOT_CUR_SLP: return r.v[S_CUR_SLP];
    case GPL_SLOT_NEXT_SLP: return r.v[S_NEXT_SLP];
    case GPL_SLOT_PREV_SLP: return r.v[S_PREV_SLP];
    case GPL_SLOT_CUR: return r.v[S_CUR];
    case GPL_SLOT_NEXT: return r.v[S_NEXT];
    default: return nullptr;
    }
}

static bool bad_region(gpl_ctx* c, int32_t region) { return !c || region < 0 || region >= (int32_t)c->rg.size(); }

int32_t gpl_set_coordinates(gpl_ctx* c, int32_t region, int32_t slot, const float* xy) {
    if (bad_region(c, region) || !xy) return GPL_ERR_ARG;
    float2* d = slot_ptr(c, region, slot);
    if (!d) return GPL_ERR_ARG;
    try {
        c->upload_region_vec(region, d, xy);
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_get_coordinates(gpl_ctx* c, int32_t region, int32_t slot, float* xy) {
    if (bad_region(c, region) || !xy) return GPL_ERR_ARG;
    float2* d = slot_ptr(c, region, slot);
    if (!d) return GPL_ERR_ARG;
    try {
        c->download_region_vec(region, d, xy);
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_density_scatter(gpl_ctx* c, int32_t region, int32_t slot, int64_t* inst_area, int64_t* filler_area, float* density,
                            double* overflow_f64) {
    if (bad_region(c, region)) return GPL_ERR_ARG;
    float2* xy = slot_ptr(c, region, slot);
    if (!xy) return GPL_ERR_ARG;
    try {
        Region& r = c->rg[region];
        size_t M = size_t(r.R.cnt_x) * r.R.cnt_y;
        gplk::launch_center_scatter(c->stream, r.dev, c->cells, xy, c->d_global_adds);
        if (inst_area) check(hipMemcpyAsync(inst_area, r.dev.inst, 8 * M, hipMemcpyDeviceToHost, c->stream), "copy");
        if (filler_area) check(hipMemcpyAsync(filler_area, r.dev.filler, 8 * M, hipMemcpyDeviceToHost, c->stream), "copy");
        c->bins_and_finish(r, region);
        if (density) check(hipMemcpyAsync(density, r.dev.density, 4 * M, hipMemcpyDeviceToHost, c->stream), "copy");
        c->sync_stats();
        r.overflow = (int64_t)c->h_sums[region].overflow;
        r.overflow_unscaled = (int64_t)c->h_sums[region].overflow_unscaled;
        if (overflow_f64) {
            overflow_f64[0] = c->h_sums[region].overflow;
            overflow_f64[1] = c->h_sums[region].overflow_unscaled;
        }
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_set_density(gpl_ctx* c, int32_t region, const float* density) {
    if (bad_region(c, region) || !density) return GPL_ERR_ARG;
    try {
        Region& r = c->rg[region];
        size_t M = size_t(r.R.cnt_x) * r.R.cnt_y;
        check(hipMemcpyAsync(r.dev.density, density, 4 * M, hipMemcpyHostToDevice, c->stream), "copy");
        check(hipStreamSynchronize(c->stream), "sync");
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_density_solve(gpl_ctx* c, int32_t region, float* phi, float* field_x, float* field_y) {
    if (bad_region(c, region)) return GPL_ERR_ARG;
    try {
        Region& r = c->rg[region];
        size_t M = size_t(r.R.cnt_x) * r.R.cnt_y;
        c->solve(r, phi != nullptr);
        if (phi) check(hipMemcpyAsync(phi, r.dev.phi, 4 * M, hipMemcpyDeviceToHost, c->stream), "copy");
        std::vector<float2> f;
        if (field_x || field_y) {
            f.resize(M);
            check(hipMemcpyAsync(f.data(), r.dev.field, sizeof(float2) * M, hipMemcpyDeviceToHost, c->stream), "copy");
        }
        check(hipStreamSynchronize(c->stream), "sync");
        for (size_t k = 0; k < f.size(); ++k) {
            if (field_x) field_x[k] = f[k].x;
            if (field_y) field_y[k] = f[k].y;
        }
        c->collect();
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_wa_forward(gpl_ctx* c, float wl_coef_x, float wl_coef_y, float* net_sums, int32_t* net_box, int64_t* hpwl) {
    if (!c) return GPL_ERR_ARG;
    try {
        c->wl_coef_x = wl_coef_x;
        c->wl_coef_y = wl_coef_y;
        c->wa_forward(true);
        c->join();
        const int32_t E = c->nnets;
        std::vector<gplk::NetRec> rec(E);
        unsigned long long h = 0;
        check(hipMemcpyAsync(rec.data(), c->nets.rec, sizeof(gplk::NetRec) * E, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipMemcpyAsync(&h, c->d_hpwl, sizeof h, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipStreamSynchronize(c->stream), "sync");
        c->collect();
        c->ensure_net_map();
        for (int32_t e = 0; e < E; ++e) {
            const gplk::NetRec& R = rec[c->net_sorted_of[e]];
            if (net_sums) {
                float* q = &net_sums[8 * size_t(e)];
                q[0] = R.sx.x;
                q[1] = R.sx.y;
                q[2] = R.sx.z;
                q[3] = R.sx.w;
                q[4] = R.sy.x;
                q[5] = R.sy.y;
                q[6] = R.sy.z;
                q[7] = R.sy.w;
            }
            if (net_box) {
                int32_t* b = &net_box[4 * size_t(e)];
                b[0] = R.box.x;
                b[1] = R.box.y;
                b[2] = R.box.z;
                b[3] = R.box.w;
            }
        }
        if (hpwl) *hpwl = (int64_t)h;
        c->cur_hpwl = (int64_t)h;
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

// The gradient kernel into scratch slots (the next slp slots), read back.
static int32_t gradients_out(gpl_ctx* c, int32_t region, float penalty, float* wl, float* den, float* sum, double* abs_sums, int slot) {
    if (bad_region(c, region)) return GPL_ERR_ARG;
    try {
        Region& r = c->rg[region];
        float saved = r.density_penalty;
        r.density_penalty = penalty;
        c->launch_gradient(region, slot, false, 0, true);
        r.density_penalty = saved;
        int base = slot == SLOT_PREV ? S_PREV_SLP : slot == SLOT_CUR ? S_CUR_SLP : S_NEXT_SLP;
        c->sync_stats();
        if (wl) c->download_region_vec(region, r.v[base + 1], wl);
        if (den) c->download_region_vec(region, r.v[base + 2], den);
        if (sum) c->download_region_vec(region, r.v[base + 3], sum);
        if (abs_sums) {
            abs_sums[0] = c->h_sums[region].wl_abs;
            abs_sums[1] = c->h_sums[region].den_abs;
        }
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_wa_backward(gpl_ctx* c, int32_t region, float* wl_grad) { return gradients_out(c, region, 0.0f, wl_grad, nullptr, nullptr, nullptr, SLOT_NEXT); }

int32_t gpl_density_gather(gpl_ctx* c, int32_t region, float* den_grad) {
    return gradients_out(c, region, 0.0f, nullptr, den_grad, nullptr, nullptr, SLOT_NEXT);
}

int32_t gpl_update_gradients(gpl_ctx* c, int32_t region, int32_t slot, float density_penalty, float* wl, float* den, float* sum,
                             double* abs_sums) {
    int s = slot == GPL_SLOT_PREV_SLP ? SLOT_PREV : slot == GPL_SLOT_CUR_SLP ? SLOT_CUR : SLOT_NEXT;
    return gradients_out(c, region, density_penalty, wl, den, sum, abs_sums, s);
}

int32_t gpl_nesterov_step(gpl_ctx* c, int32_t region, float step, float coeff, float* next, float* next_slp) {
    if (bad_region(c, region)) return GPL_ERR_ARG;
    try {
        Region& r = c->rg[region];
        gplk::launch_update(c->stream, r.dev, c->cells, r.v[S_CUR_SLP], r.v[S_CUR_SLP_SUM], r.v[S_CUR], r.v[S_NEXT], r.v[S_NEXT_SLP], step,
                            coeff, nullptr);
        gplk::launch_density_scatter(c->stream, r.dev, c->cells, c->d_global_adds);
        c->bins_and_finish(r, region);
        c->sync_stats();
        if (next) c->download_region_vec(region, r.v[S_NEXT], next);
        if (next_slp) c->download_region_vec(region, r.v[S_NEXT_SLP], next_slp);
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_debug_inject(gpl_ctx* c, int32_t kind, int32_t after_iter) {
    if (!c || kind < 0 || kind > 6) return GPL_ERR_ARG;
    if (kind == 5) {
        c->inject_zero_wl_at = after_iter;
        return GPL_OK;
    }
    if (kind == 6) {
        if (c->inject_diverge_at < 0) c->inject_diverge_at = after_iter;
        else c->inject_diverge_at2 = after_iter;
        return GPL_OK;
    }
    if (kind == 0) c->inject_zero_wl_at = c->inject_diverge_at = c->inject_diverge_at2 = -1;
    if (kind == 0 || kind == 4) {
        // kind 4: the kernels' general paths (bin edges from the tables, the gradient's per-cell
        // gathers of pin terms); they give the same results as the fast ones.  The autotuner then
        // keeps its hands off the gradient's staging.
        for (size_t b = 0; b < c->rg.size(); ++b) {
            Region& r = c->rg[b];
            r.dev.edges_rounded = kind == 4 ? 0 : r.edges_rounded;
            if (kind == 4) c->tune[b].stage = 0;
            c->apply_tune(int(b));
        }
        c->drop_graph();  // the graphs hold the kernels' old paths
        if (kind == 4) {
            c->autotune = 0;
            return GPL_OK;
        }
    }
    c->inject_kind = kind;
    c->inject_after = kind == 0 ? -1 : after_iter;
    return GPL_OK;
}

int32_t gpl_set_tuning(gpl_ctx* c, int32_t region, int32_t chunk, int32_t table_bits, int32_t stage) {
    if (!c) return GPL_ERR_ARG;
    if (region < 0) {
        c->autotune = chunk != 0 ? 1 : 0;  // region -1: the tuner on (chunk 1) or off (chunk 0)
        return GPL_OK;
    }
    if (bad_region(c, region) || (chunk != 256 && chunk != gplk::kPlaceChunk) || (table_bits != 32 && table_bits != 64) ||
        (stage != 0 && stage != 1))
        return GPL_ERR_ARG;
    try {
        c->join();
        check(hipStreamSynchronize(c->stream), "sync");
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    Tune& t = c->tune[region];
    t.chunk = chunk;
    t.narrow = table_bits == 32 ? 1 : 0;
    t.stage = stage;
    t.source = 0;
    c->apply_tune(region);
    c->autotune = 0;
    c->drop_graph();
    return GPL_OK;
}

int32_t gpl_get_tuning(gpl_ctx* c, int32_t* out, int32_t cap) {
    if (!c) return GPL_ERR_ARG;
    const int32_t n = int32_t(c->rg.size());
    for (int32_t b = 0; b < n && out && b < cap; ++b) {
        const Region& r = c->rg[b];
        out[6 * b] = r.dev.chunk;
        out[6 * b + 1] = r.dev.narrow ? 32 : 64;
        out[6 * b + 2] = r.dev.grad_stage ? 1 : 0;
        out[6 * b + 3] = c->tune[b].source;
        out[6 * b + 4] = int32_t(std::lround(100.0 * c->tune[b].scatter_us));
        out[6 * b + 5] = int32_t(std::lround(100.0 * c->tune[b].gradient_us));
    }
    return n;
}

int32_t gpl_set_launch_mode(gpl_ctx* c, int32_t mode) {
    if (!c || mode < 0 || mode > 7) return GPL_ERR_ARG;
    try {
        c->join();
        check(hipStreamSynchronize(c->stream), "sync");
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    c->drop_graph();
    c->grouped_trials = mode == 7;
    if (mode == 7) mode = 2;  // the trials grouped; every other launch one by one on one stream
    c->launch_mode = mode >= 3 ? 0 : mode;
    c->overlap = mode < 2 || mode >= 4;
    c->device_trials = mode >= 4;
    c->ctl_launch = mode >= 4 ? mode - 4 : 0;
    return GPL_OK;
}

int32_t gpl_set_resort_policy(gpl_ctx* c, int32_t min_gap, float growth) {
    if (!c || min_gap < 1 || !(growth >= 1.0f)) return GPL_ERR_ARG;
    c->resort_gap = min_gap;
    c->resort_growth = growth;
    return GPL_OK;
}

int32_t gpl_set_profiling(gpl_ctx* c, int32_t on) {
    if (!c) return GPL_ERR_ARG;
    if (on && !c->profiling) {
        for (int i = 0; i < 64; ++i) hipEventCreateWithFlags(&c->ev[i], hipEventDisableSystemFence);
    }
    if (!on && c->profiling) {
        hipStreamSynchronize(c->stream);
        for (int i = 0; i < 64; ++i) hipEventDestroy(c->ev[i]);
    }
    c->profiling = on != 0;
    c->nev = 0;
    for (int f = 0; f < F_COUNT; ++f) {
        c->fam_ms[f] = 0;
        c->fam_n[f] = 0;
    }
    return GPL_OK;
}

int32_t gpl_kernel_times(gpl_ctx* c, gpl_kernel_time* out, int32_t cap) {
    if (!c) return GPL_ERR_ARG;
    int32_t n = 0;
    for (int f = 0; f < F_COUNT && n < cap; ++f, ++n) {
        std::memset(&out[n], 0, sizeof out[n]);
        std::snprintf(out[n].name, sizeof out[n].name, "%s", kFamName[f]);
        out[n].launches = c->fam_n[f];
        out[n].ms = c->fam_ms[f];
    }
    // the graph-launched trials' host time (after the HBM atomic count: callers that take a fixed
    // number of entries see the kernel families first)
    static const char* kHostName[] = {"host_params", "host_launch", "host_wait", "host_between"};
    if (n < cap) {
        // the scatter's atomic adds into the HBM bins since the context began (the count is in
        // `ms`), and the re-sorts (in `launches`)
        unsigned long long ga = 0;
        if (hipMemcpy(&ga, c->d_global_adds, sizeof ga, hipMemcpyDeviceToHost) == hipSuccess) {
            std::memset(&out[n], 0, sizeof out[n]);
            std::snprintf(out[n].name, sizeof out[n].name, "scatter_global_adds");
            out[n].launches = c->resorts;
            out[n].ms = double(ga);
            ++n;
        }
    }
    if (n < cap) {
        // the autotuner's measurements: runs in `launches`, their host time in `ms`
        std::memset(&out[n], 0, sizeof out[n]);
        std::snprintf(out[n].name, sizeof out[n].name, "autotune");
        out[n].launches = c->tune_runs;
        out[n].ms = c->tune_ms;
        ++n;
    }
    if (n < cap && c->ctl_halts > 0) {
        // device-driven trials: the device's halts in `launches`, the host's time from each entry
        // to the halt's state taken back in `ms`; then the trial graphs launched
        std::memset(&out[n], 0, sizeof out[n]);
        std::snprintf(out[n].name, sizeof out[n].name, "device_halts");
        out[n].launches = int32_t(c->ctl_halts);
        out[n].ms = c->ctl_host_ms;
        ++n;
        if (n < cap) {
            std::memset(&out[n], 0, sizeof out[n]);
            std::snprintf(out[n].name, sizeof out[n].name, "device_launched");
            out[n].launches = int32_t(c->ctl_graphs);
            ++n;
        }
    }
    for (int h = 0; h < gpl_ctx::H_COUNT && n < cap && c->host_n > 0; ++h, ++n) {
        std::memset(&out[n], 0, sizeof out[n]);
        std::snprintf(out[n].name, sizeof out[n].name, "%s", kHostName[h]);
        out[n].launches = c->host_n;
        out[n].ms = c->host_ms[h];
    }
    return n;
}

// The objective's common first step: region 0's points uploaded (device order) and its density
// and potential computed from the continuous boxes.  The region's bin arrays are reused, so the
// loop's fields are overwritten (every trial recomputes them).
static void objective_density(gpl_ctx* c, const float* xy) {
    Region& r = c->rg[0];
    c->join();
    if (!c->obj_xy) {
        c->obj_xy = c->dalloc<float2>(std::max(r.n, 1));
        c->obj_grad = c->dalloc<float2>(std::max(r.n, 1));
        c->obj_out = c->dalloc<double>(2);
    }
    c->upload_region_vec(0, c->obj_xy, xy);
    gplk::launch_objective_density(c->stream, r.dev, r.sb, c->cells, c->obj_xy);
}

int32_t gpl_objective_forward(gpl_ctx* c, const float* xy, float wl_coef, float lambda, double* value, double* wa_part,
                              double* density_part) {
    if (!c || !xy || c->rg.empty()) return GPL_ERR_ARG;
    try {
        objective_density(c, xy);
        Region& r = c->rg[0];
        gplk::launch_objective_value(c->stream, r.dev, c->nets, c->obj_xy, wl_coef, r.partials, c->obj_out);
        double h[2] = {0.0, 0.0};
        check(hipMemcpyAsync(h, c->obj_out, sizeof h, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipStreamSynchronize(c->stream), "sync");
        const double d = 0.5 * h[1];
        if (wa_part) *wa_part = h[0];
        if (density_part) *density_part = d;
        if (value) *value = h[0] + (double)lambda * d;
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_objective_backward(gpl_ctx* c, const float* xy, float wl_coef, float lambda, float* grad) {
    if (!c || !xy || !grad || c->rg.empty()) return GPL_ERR_ARG;
    try {
        objective_density(c, xy);
        Region& r = c->rg[0];
        gplk::launch_objective_gradient(c->stream, r.dev, c->cells, c->nets, c->obj_xy, wl_coef, lambda, r.partials, c->obj_grad);
        c->download_region_vec(0, c->obj_grad, grad);
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

// ---- routability's congestion measurement ------------------------------------------------------
int32_t gpl_rudy_wire_width(const int32_t* pitches, int32_t n) {
    double terms = 0.0;
    for (int32_t i = 0; i < n; ++i) terms += 1.0 / double(pitches[i]);
    return terms != 0.0 ? (int32_t)((1.0 / terms) * 1.25) : 100;
}

static bool bad_grid(const gpl_rudy_grid* g) {
    return !g || g->cnt_x <= 0 || g->cnt_y <= 0 || g->tile_size <= 0 || (int64_t)g->cnt_x * g->cnt_y > ((int64_t)1 << 26);
}

static gplk::RudyGrid rudy_grid(const gpl_rudy_grid* g) {
    return gplk::RudyGrid{g->lx, g->ly, g->ux, g->uy, g->cnt_x, g->cnt_y, g->tile_size, g->wire_width};
}

int32_t gpl_rudy(gpl_ctx* c, const gpl_rudy_grid* g, const uint8_t* cap_red, float* demand, float* blockage) {
    if (!c || bad_grid(g)) return GPL_ERR_ARG;
    try {
        const size_t tiles = size_t(g->cnt_x) * size_t(g->cnt_y);
        c->join();
        // the nets' boxes at the cells' current centres, into the net records
        c->wa_forward(true);
        c->join();
        const uint8_t* cr = nullptr;
        if (cap_red) {
            if (c->cap_red_cap < 2 * tiles) {
                if (c->d_cap_red) check(hipFree(c->d_cap_red), "free");
                c->d_cap_red = nullptr;
                check(hipMalloc(&c->d_cap_red, 2 * tiles), "hipMalloc");
                c->cap_red_cap = 2 * tiles;
            }
            check(hipMemcpyAsync(c->d_cap_red, cap_red, 2 * tiles, hipMemcpyHostToDevice, c->stream), "upload");
            cr = c->d_cap_red;
        }
        check(gplk::launch_rudy(c->stream, c->nets, c->d_net_orig, rudy_grid(g), cr, c->rudy), "rudy");
        if (demand) check(hipMemcpyAsync(demand, c->rudy.demand, sizeof(float) * tiles, hipMemcpyDeviceToHost, c->stream), "copy");
        if (blockage) check(hipMemcpyAsync(blockage, c->rudy.blockage, sizeof(float) * tiles, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipStreamSynchronize(c->stream), "sync");
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

int32_t gpl_rudy_tiles(gpl_ctx* c, const gpl_rudy_grid* g, const gpl_route_vars* v, float* inflation_ratio, uint8_t* actionable,
                       gpl_rudy_stats* stats) {
    if (!c || bad_grid(g) || !v || !c->rudy.demand || size_t(g->cnt_x) * size_t(g->cnt_y) > c->rudy.cap_tiles) return GPL_ERR_ARG;
    try {
        const size_t tiles = size_t(g->cnt_x) * size_t(g->cnt_y);
        gplk::RouteVarsDev rv{v->min_congestion_for_inflation, v->inflation_ratio_coef, v->max_inflation_ratio, v->ignore_edge_ratio,
                              {v->rc_k[0], v->rc_k[1], v->rc_k[2], v->rc_k[3]}};
        c->join();
        check(gplk::launch_rudy_tiles(c->stream, c->cells, c->ncells, rudy_grid(g), rv, c->rudy), "rudy tiles");
        double st[8];
        check(hipMemcpyAsync(st, c->rudy.stats, sizeof st, hipMemcpyDeviceToHost, c->stream), "copy");
        if (inflation_ratio)
            check(hipMemcpyAsync(inflation_ratio, c->rudy.ratio, sizeof(float) * tiles, hipMemcpyDeviceToHost, c->stream), "copy");
        if (actionable) check(hipMemcpyAsync(actionable, c->rudy.act, tiles, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipStreamSynchronize(c->stream), "sync");
        if (stats) {
            stats->total_overflow = st[0];
            stats->overflowed_tiles = int32_t(st[1]);
            stats->actionable_tiles = int32_t(st[2]);
            for (int k = 0; k < 4; ++k) stats->top[k] = st[3 + k];
            stats->rc = float(st[7]);
        }
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

// The cells of a region between the problem's order (host arrays) and the device's.
static void region_perm(gpl_ctx* c, int32_t b, std::vector<int32_t>& dev_to_prob) {
    c->ensure_host_map();
    const Region& r = c->rg[b];
    dev_to_prob.resize(r.n);
    for (int k = 0; k < r.n; ++k) dev_to_prob[k] = c->cell_old_of[r.R.cell_begin + k] - r.R.cell_begin;
}

// updateDensitySize's per-region constants in the reference's extended precision (nesterov.rs:2022-2036).
static gplk::DensitySizeDev density_size_consts(const gpl_region& R) {
    const long double s2 = 1.414213562373095048801L;
    const long double tx = s2 * (long double)R.size_x, ty = s2 * (long double)R.size_y;
    gplk::DensitySizeDev d{};
    const long double fx = floorl(tx), fy = floorl(ty);
    d.max_x = (int32_t)(fx == tx ? fx - 1 : fx);  // dx < tx for an integer dx
    d.max_y = (int32_t)(fy == ty ? fy - 1 : fy);
    d.thr_x = (float)tx;
    d.thr_y = (float)ty;
    d.size_x = (float)(s2 * (long double)(float)R.size_x);
    d.size_y = (float)(s2 * (long double)(float)R.size_y);
    return d;
}

void gpl_ctx::update_density_size(int b) {
    upload_sizes(b);
    Region& r = rg[b];
    check(gplk::launch_density_size(stream, r.R.cell_begin, r.n, cells, density_size_consts(r.R), rt.d_sizes), "density size");
    check(hipStreamSynchronize(stream), "sync");
    drop_graph();
}

int32_t gpl_route_inflate(gpl_ctx* c, int32_t region, const gpl_rudy_grid* g, const float* inflated_ratio, int32_t* sizes, int64_t* delta) {
    if (bad_region(c, region) || bad_grid(g) || !inflated_ratio || !sizes) return GPL_ERR_ARG;
    void* d_ratio = nullptr;
    void* d_sizes = nullptr;
    void* d_delta = nullptr;
    int32_t rc = GPL_OK;
    try {
        Region& r = c->rg[region];
        const size_t tiles = size_t(g->cnt_x) * size_t(g->cnt_y);
        std::vector<int32_t> perm;
        region_perm(c, region, perm);
        std::vector<int2> sz(std::max(r.n, 1));
        for (int k = 0; k < r.n; ++k) sz[k] = make_int2(sizes[2 * perm[k]], sizes[2 * perm[k] + 1]);
        c->join();
        check(hipMalloc(&d_ratio, sizeof(float) * tiles), "hipMalloc");
        check(hipMalloc(&d_sizes, sizeof(int2) * sz.size()), "hipMalloc");
        check(hipMalloc(&d_delta, sizeof(unsigned long long)), "hipMalloc");
        check(hipMemcpyAsync(d_ratio, inflated_ratio, sizeof(float) * tiles, hipMemcpyHostToDevice, c->stream), "upload");
        check(hipMemcpyAsync(d_sizes, sz.data(), sizeof(int2) * sz.size(), hipMemcpyHostToDevice, c->stream), "upload");
        gplk::launch_set_u64(c->stream, (unsigned long long*)d_delta, 0ull);
        check(gplk::launch_route_inflate(c->stream, r.R.cell_begin, r.n, c->cells, rudy_grid(g), (const float*)d_ratio, (int2*)d_sizes,
                                         (unsigned long long*)d_delta),
              "inflate");
        unsigned long long dl = 0;
        check(hipMemcpyAsync(sz.data(), d_sizes, sizeof(int2) * sz.size(), hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipMemcpyAsync(&dl, d_delta, sizeof dl, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipStreamSynchronize(c->stream), "sync");
        for (int k = 0; k < r.n; ++k) {
            sizes[2 * perm[k]] = sz[k].x;
            sizes[2 * perm[k] + 1] = sz[k].y;
        }
        if (delta) *delta = (int64_t)dl;
    } catch (const HipError& e) {
        rc = fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    for (void* p : {d_ratio, d_sizes, d_delta})
        if (p) hipFree(p);
    return rc;
}

int32_t gpl_update_density_size(gpl_ctx* c, int32_t region, const int32_t* sizes) {
    if (bad_region(c, region) || !sizes) return GPL_ERR_ARG;
    void* d_sizes = nullptr;
    int32_t rc = GPL_OK;
    try {
        Region& r = c->rg[region];
        std::vector<int32_t> perm;
        region_perm(c, region, perm);
        std::vector<int2> sz(std::max(r.n, 1));
        for (int k = 0; k < r.n; ++k) sz[k] = make_int2(sizes[2 * perm[k]], sizes[2 * perm[k] + 1]);
        c->join();
        check(hipMalloc(&d_sizes, sizeof(int2) * sz.size()), "hipMalloc");
        check(hipMemcpyAsync(d_sizes, sz.data(), sizeof(int2) * sz.size(), hipMemcpyHostToDevice, c->stream), "upload");
        check(gplk::launch_density_size(c->stream, r.R.cell_begin, r.n, c->cells, density_size_consts(r.R), (const int2*)d_sizes), "density size");
        check(hipStreamSynchronize(c->stream), "sync");
    } catch (const HipError& e) {
        rc = fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    if (d_sizes) hipFree(d_sizes);
    return rc;
}

int32_t gpl_get_route_stats(gpl_ctx* c, gpl_route_stats* s, gpl_route_pass* passes, int32_t cap) {
    if (!c) return GPL_ERR_ARG;
    if (s) {
        std::memset(s, 0, sizeof *s);
        s->passes = int32_t(c->rt.passes.size());
        s->route_iters = c->rt.route_iters;
        s->final_rc = c->rt.final_rc;
        s->min_rc = c->rt.min_rc;
        for (int64_t a : c->rt.accumulated) s->total_inflation += a;
        s->original_area = c->rt.original_movable_area;
    }
    const int32_t n = int32_t(c->rt.passes.size());
    for (int32_t i = 0; passes && i < n && i < cap; ++i) passes[i] = c->rt.passes[size_t(i)];
    return n;
}

int32_t gpl_get_density_boxes(gpl_ctx* c, int32_t region, int32_t* dbox, float* density_scale) {
    if (bad_region(c, region)) return GPL_ERR_ARG;
    try {
        Region& r = c->rg[region];
        std::vector<int32_t> perm;
        region_perm(c, region, perm);
        std::vector<int4> db(std::max(r.n, 1));
        std::vector<float> sc(std::max(r.n, 1));
        c->join();
        check(hipMemcpyAsync(db.data(), c->cells.dbox + r.R.cell_begin, sizeof(int4) * r.n, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipMemcpyAsync(sc.data(), c->cells.dscale + r.R.cell_begin, sizeof(float) * r.n, hipMemcpyDeviceToHost, c->stream), "copy");
        check(hipStreamSynchronize(c->stream), "sync");
        for (int k = 0; k < r.n; ++k) {
            if (dbox) {
                int32_t* o = dbox + 4 * size_t(perm[k]);
                o[0] = db[k].x;
                o[1] = db[k].y;
                o[2] = db[k].z;
                o[3] = db[k].w;
            }
            if (density_scale) density_scale[perm[k]] = sc[k];
        }
    } catch (const HipError& e) {
        return fail(c, GPL_ERR_HIP, strfmt("%s: %s", e.what, hipGetErrorString(e.e)));
    }
    return GPL_OK;
}

}  // extern "C"
