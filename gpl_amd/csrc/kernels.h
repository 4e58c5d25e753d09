// Device kernels of the Nesterov engine (kernels.hip) and their launchers, called by the host
// controller (nesterov.cpp).  Layout in HBM, per region r of the problem:
//   cells   struct-of-arrays, each region's cells contiguous [cell_begin, cell_end), in Morton
//           order of their bins (re-sorted on the device as cells move): dbox int4 (the density
//           box), centre int2, density_scale f32, flags u8, the coordinate slots float2 (cur_slp,
//           next_slp, prev_slp, cur, next, snapshots), the gradient slots float2 (wl, den, sum per
//           slp slot), pin CSR (start int32; per cell-pin: net-pin index int32, the pin's WA
//           gradient float2), preconditioner base (pins, area) float2.
//   pins    in net order (net CSR): cell int32, offset int2, initial position int2, cell-pin
//           index int32.
//   nets    sorted by degree, then spatially (the host keeps the permutation): start int32,
//           weight f32; a 64-byte record (box, the WA sums) for the kernel-level entry.
//   bins    y-major (k = y*cnt_x + x): inst/filler int64 accumulators, non_place
//           int64 and (with the unscaled one) as f32 pairs, density f32, the field as (x, y) f32 pairs, scratch for the
//           transforms.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace gplk {

struct CtlRegion;
struct CtlHead;

struct RegionDev {
    int32_t lx, ly, ux, uy;
    int32_t cnt_x, cnt_y;
    double size_x, size_y;
    float target_density;
    int32_t cell_begin, cell_end;
    int32_t narrow;  // 1: the scatter's LDS tables may hold 32-bit sums (kPlaceChunk x the largest bin area < 2^32)
    int32_t edges_rounded;  // 1: bin_x[i] = lx + round(i size_x) and bin_y alike (init_bins' edges)
    int32_t grad_stage;     // the gradient stages a block's pin terms in LDS up to this many (at most kGradStage)
    int32_t chunk;          // cells a scatter block takes at a time: 256 or kPlaceChunk (0: kPlaceChunk); the autotuner's pick
    // bins
    const int32_t* bin_x;  // cnt_x + 1
    const int32_t* bin_y;  // cnt_y + 1
    unsigned long long* inst;    // int64 accumulators (two's complement in u64)
    unsigned long long* filler;
    const long long* non_place;
    const float2* non_place_f;  // per bin ((float)non_place, (float)non_place_unscaled), as the bins pass uses them
    float* density;
    float2* field;  // the electric field per bin as (x, y) pairs
    float* phi;
    // device-driven trials (null otherwise): the region's control record and the global one; the
    // kernels then read their step, penalty, coefficients and slot pointers from them and do
    // nothing while ctl->skip is set
    const CtlRegion* ctl;
    const CtlHead* cg;
};

// The Poisson solve's buffers and tables of a region (host side: the row kernels get them through
// their own arguments, so the kernels that take a RegionDev by value do not carry them).
struct SolveBufs {
    float* work0;  // transform scratch: 2 cnt_x*cnt_y (the three-pass solve's field pairs), the others cnt_x*cnt_y
    float* work1;
    float* work2;
    float* work3;
    const float* wx;
    const float* wy;
    // transform tables for a length L (h = L/2), as float2: the half-length FFT's twiddles
    // e^{-2 pi i m / h} (m < h), the real-transform split e^{-2 pi i k / L} (k <= h) and the DCT
    // twiddles (cos, -sin)(pi k / 2L) (k < L)
    const float2* tw_x;
    const float2* th_x;
    const float2* dw_x;
    const float2* tw_y;
    const float2* th_y;
    const float2* dw_y;
};

struct CellsDev {
    int4* dbox;
    int2* ctr;                  // the density box's centre; (INT_MIN, INT_MIN) until first centred
    float* dscale;
    uint8_t* flags;             // GPL_CELL_* (locked is updated on convergence)
    int32_t* pin_start;         // num_cells + 1
    int32_t* cp_npin;           // per cell-pin: its index in the net-order pin arrays
    float2* pre;                // (pin count, dx*dy) per cell, as float
    int32_t* old_of;            // the problem's index of each device cell
};

// One net's state in one 64-byte record (one cache line for the gradient's per-pin gather):
// box, the WA sums along x (min_x, x_min_x, max_x, x_max_x) and y, and the net weight.
struct __align__(16) NetRec {
    int4 box;
    float4 sx;
    float4 sy;
    float w;
    int32_t pad[3];
};

struct NetsDev {
    int32_t num_nets;
    const int32_t* start;   // num_nets + 1, into the net-order pin arrays
    int32_t* pcell;         // per net-pin: cell or -1 (renamed by a re-sort)
    const int2* poff;       // per net-pin: offset from the cell centre; a pin on no movable cell: its position
    const int2* pxy0;       // per net-pin: position as init left it
    const float* weight;    // per net
    int32_t* cpi;           // per net-pin: its cell-pin index, or -1 (renumbered by a re-sort)
    float2* pgrad;          // per cell-pin: the pin's weighted WA gradient, written by the WA kernel
    NetRec* rec;            // box and sums per net, written when asked for (the kernel-level entry)
    int32_t num_small;      // nets [0, num_small) go by tiles; the rest take a block each
    // Tile t holds the small nets whose first pin lies in [t kTileSpan, (t+1) kTileSpan), nets
    // [tile[t], tile[t+1]) (at most kTileSpan + 127 < kTilePins pins); XCD x takes the tiles
    // [x tile_blocks, (x+1) tile_blocks)
    const int32_t* tile;    // num_tiles + 1
    int32_t num_tiles;
    int32_t tile_blocks;
    const CtlHead* cg;      // device-driven trials: the WA coefficients from here, nothing while halted
};

// Per-trial FP64 reductions of one region, finished into this struct by k_finish.
struct Sums {
    double overflow, overflow_unscaled;   // sum of the per-bin overflow terms
    double wl_abs, den_abs;               // sum |wl| and |den| over x and y
    double coord_sq, grad_sq;             // squared distances (step length)
    unsigned long long hpwl;              // pin HPWL (integer, exact)
    int32_t pad[2];
};

// ---- device-driven trials ----------------------------------------------------------------------
// The loop's control record in device memory.  The trial graph's last kernel (k_control) takes the
// host's decisions of nesterov.cpp on the device after every back-tracking trial -- step length and
// its acceptance, another trial or the iteration's end (phi, penalty, WA coefficient, the divergence
// watermarks, the momentum) -- so that trials follow each other with no host round trip; it halts
// (every kernel of later trials doing nothing) on the rare events the host handles: a zero or
// non-finite gradient sum or step, a divergence, a new min-HPWL snapshot, a region's convergence,
// a re-sort, the iteration limit.
constexpr int kCtlMaxRegions = 16;
constexpr int kCtlRing = 1024;  // history and progress rows kept between two halts

// The arguments a region's trial kernels read (one 64-byte line, loaded at once): set by the host's
// upload and refreshed by the control after every trial.
struct alignas(64) CtlHot {
    // bit 0: the region's kernels do nothing (converged, or the device halted); bit 1: the gradient
    // puts a NaN into its sums (a host-forced case)
    int32_t flags;
    float step_length, density_penalty, coeff;
    float2* cur_slp;
    float2* cur_sum;
    float2* cur;
    float2* next;
    float2* next_slp;
    float2* next_sum;
};

struct alignas(64) CtlRegion {
    CtlHot hot;
    float2* sets[3][4];  // the three slp point sets (coordinates, wl, den, preconditioned sum)
    float2* cn[2];       // cur / next
    int32_t rot, cnp;    // cur_slp = sets[rot], next_slp = sets[(rot+1)%3], prev_slp = sets[(rot+2)%3]; cur = cn[cnp]
    int32_t is_converged, max_phi_coef_changed, n, group;
    float step_length, density_penalty, max_phi_coef, sum_overflow, sum_overflow_unscaled, min_sum_overflow;
    float hpwl_with_min_sum_overflow, coordi_distance, peak_coordi_distance, prev_reported_overflow_unscaled, wl_grad_sum;
    float density_grad_sum;
    long long prev_hpwl, prev_reported_hpwl, overflow, overflow_unscaled, nia, non_place_insts_area;
};

struct CtlHist {  // one iteration's history row (gpl_iter_row)
    int32_t iter, trials;
    float overflow, density_penalty, step_length, wl_coef;
    long long hpwl;
};

struct CtlLog {  // one region's progress row (every tenth iteration)
    int32_t iter, region;
    float overflow, pct, penalty;
    int32_t pad;
    long long hpwl;
};

enum CtlEvent {
    CTL_RUN = 0,
    CTL_ZERO_WL = 1,   // stage 1: a zero WA gradient sum (the coefficient-halving recursion)
    CTL_NAN = 2,       // stage 1: a non-finite gradient sum (GPL-0306)
    CTL_STEP = 3,      // stage 1: a non-finite step length (GPL-0305)
    CTL_DIVERGE = 4,   // stage 2: a region's divergence rule fired
    CTL_SNAPSHOT = 5,  // stage 2: a new min-HPWL snapshot is due
    CTL_CONVERGE = 6,  // stage 2: a region reached the target overflow
    CTL_RESORT = 7,    // stage 2: the re-sort trigger fired
    CTL_STOP = 8,      // stage 3: the iteration limit, the caller's stop, or the rows are full
};

// The loop's scalars and the regions' records (the part the control stages in LDS).
struct alignas(64) CtlHead {
    // read by the trial kernels: the halt event, the momentum coefficient, the WA coefficients
    int32_t halt;
    float coeff, wl_coef_x, wl_coef_y;
    int32_t stage;  // where the device halted: 1 after a trial's gradient sums, 2 after an
                    // iteration's bookkeeping, 3 before the next iteration
    int32_t nreg, iter, iter_stop, max_iter, trials, total_trials, recursion_wl, disable_revert;
    float base_wl_coef, cur_a, reference_hpwl, average_overflow, average_overflow_unscaled;
    float overflow_target, min_phi, diverge_snapshot_overflow;
    int32_t is_min_hpwl, diverge_snapshot_iter, resort_enabled, last_resort_iter, iters_run;
    long long min_hpwl, cur_hpwl;
    unsigned long long last_global_adds;
    double adds_base;
    int32_t nhist, nlog;
    // host-forced cases (-1 off): a NaN in the gradient sums from iters_run >= nan_from, a NaN step
    // length from iters_run >= step_nan_from, a re-sort at every iteration >= resort_from, the WA sum
    // read as zero in iteration zero_wl_at, a divergence in iterations diverge_at / diverge_at2
    int32_t nan_from, step_nan_from, resort_from, zero_wl_at, diverge_at, diverge_at2;
    // the re-sort policy: a re-sort when the scatter's global adds per trial grew by resort_growth
    // over their level after the last one, at least resort_gap iterations after it
    int32_t resort_gap;
    float resort_growth;
    CtlRegion reg[kCtlMaxRegions];
};

struct CtlGlobal {
    CtlHead h;
    // rows since the last halt (the host reads them there)
    CtlHist hist[kCtlRing];
    CtlLog log[kCtlRing];
};

// The control's arguments to the trial's last finishing reduction (g null: no control).
struct CtlArgs {
    CtlGlobal* g;
    const Sums* sums;  // every region's, then the tail (HPWL and global-add counters)
    Sums* tail;
    unsigned long long* dev_seq;
    unsigned long long* host_seq;  // pinned: [0] sequence number, [1] halt event
    int32_t nreg;
};

// the hot fields of every region from its canonical ones (skip: halted or converged)
__host__ __device__ inline void ctl_refresh(CtlHead& g) {
    for (int b = 0; b < g.nreg; ++b) {
        CtlRegion& q = g.reg[b];
        const int ns = q.rot == 2 ? 0 : q.rot + 1;
        q.hot.flags = ((g.halt != 0 || q.is_converged) ? 1 : 0) | (g.nan_from >= 0 && g.iters_run >= g.nan_from ? 2 : 0);
        q.hot.step_length = q.step_length;
        q.hot.density_penalty = q.density_penalty;
        q.hot.coeff = g.coeff;
        q.hot.cur_slp = q.sets[q.rot][0];
        q.hot.cur_sum = q.sets[q.rot][3];
        q.hot.cur = q.cn[q.cnp];
        q.hot.next = q.cn[q.cnp ^ 1];
        q.hot.next_slp = q.sets[ns][0];
        q.hot.next_sum = q.sets[ns][3];
    }
}

// ---- routability's congestion measurement (route.hip) ------------------------------------------
// The RUDY tile grid (RudyInputs / makeGrid): the die, the tiles per side, the tile size and the
// wire width (Rudy::new).
struct RudyGrid {
    int32_t lx, ly, ux, uy;
    int32_t cnt_x, cnt_y, tile, wire_width;
};
// RouteBaseVars' inflation fields: minCongestionForInflation, inflationRatioCoef,
// maxInflationRatio, ignoreEdgeRatio, the top-share weights rcK.
struct RouteVarsDev {
    float min_congestion, coef, max_ratio, ignore_edge_ratio;
    float rc_k[4];
};
// Device buffers of the measurement (grown on demand, kept for the context): the pairs and their
// sort; per tile the demand and blockage (x-major), the actionable flags, inflation ratios and
// congestions (y-major), the congestions sorted; the statistics (total overflow, overflowed,
// actionable, the four top averages, the weighted congestion).
struct RudyBufs {
    unsigned long long* count = nullptr;
    unsigned long long* offset = nullptr;
    unsigned long long* keys = nullptr;
    unsigned long long* keys2 = nullptr;
    float* vals = nullptr;
    float* vals2 = nullptr;
    unsigned char* temp = nullptr;  // the library sorts' and scans' scratch
    float* demand = nullptr;
    float* blockage = nullptr;
    uint8_t* act = nullptr;
    float* ratio = nullptr;
    float* cong = nullptr;
    float* sorted = nullptr;
    double* stats = nullptr;
    float* over = nullptr;          // the overflowed tiles' congestions in the tiles' order
    int* nover = nullptr;           // their count
    int* col_start = nullptr;       // per column (cnt_x + 1): its congested tiles in col_y
    int* col_y = nullptr;           // the congested tiles' rows, column by column, ascending
    size_t cap_nets = 0, cap_off = 0, cap_pairs = 0, cap_pairs2 = 0, cap_pairs3 = 0, cap_pairs4 = 0, cap_temp = 0;
    size_t cap_tiles = 0, cap_tiles2 = 0, cap_act = 0, cap_ratio = 0, cap_cong = 0, cap_sorted = 0, cap_stats = 0;
    size_t cap_over = 0, cap_nover = 0, cap_col = 0, cap_coly = 0;
    void release();
};
// RUDY demand and blockage of the grid from the nets' boxes in n.rec (a WA forward with the net
// records written), (capacity, reduction) pairs per tile (x-major, device; null: none); the nets'
// problem indices in net_orig order the float sums.  Waits for the stream (the pair count).
hipError_t launch_rudy(hipStream_t s, const NetsDev& n, const int32_t* net_orig, const RudyGrid& g, const uint8_t* cap_red_dev, RudyBufs& b);
// The tiles of the last launch_rudy: actionable (standard instance cells' density centres among
// the ncells device cells), inflation ratios, the averages into b.stats.
hipError_t launch_rudy_tiles(hipStream_t s, const CellsDev& c, int ncells, const RudyGrid& g, const RouteVarsDev& v, RudyBufs& b);

// updateDensitySize's per-region constants (the reference's extended-precision sqrt(2) bin sizes):
// max_x the largest integer below sqrt(2) size_x, thr_x that threshold rounded to float, size_x
// sqrt(2) (float) size_x rounded to float; y alike.
struct DensitySizeDev {
    int32_t max_x, max_y;
    float thr_x, thr_y, size_x, size_y;
};
// routability's cell inflation of cells [begin, begin + n) (sizes: their cell-box sizes, device
// order, updated; *delta += the area added) at the tiles' inflated ratios (y-major)
hipError_t launch_route_inflate(hipStream_t s, int begin, int n, const CellsDev& c, const RudyGrid& g, const float* ratio, int2* sizes,
                                unsigned long long* delta);
// updateDensitySize of cells [begin, begin + n) at cell-box sizes (device order): density boxes,
// density scales, the preconditioner's areas
hipError_t launch_density_size(hipStream_t s, int begin, int n, const CellsDev& c, const DensitySizeDev& d, const int2* sizes);



constexpr int kPartials = 2048;  // per-block partial slots per quantity
constexpr int kGradStage = 2048;  // pin terms of a gradient block's cells staged in LDS, at most
constexpr int kPlaceChunk = 512;  // cells per block iteration of the scatter family
constexpr int kTilePins = 2048;   // pins a WA tile stages in LDS
constexpr int kTileSpan = kTilePins - 128;  // pin range whose nets form a tile (a small net has <= 128 pins)
static_assert(kTilePins >= 256 && kTilePins % 256 == 0, "a WA tile holds any small net, staged by 256 threads");
static_assert(sizeof(CtlHot) == 64 && sizeof(Sums) == 64, "one 64-byte line each");
static_assert(sizeof(CtlHead) % 8 == 0 && sizeof(CtlHead) / 8 <= 1024 && (kCtlMaxRegions + 1) * sizeof(Sums) / 8 <= 1024,
              "the control stages one word per thread of the 1024-thread finishing block");

// ---- launchers (stream-ordered, asynchronous) -----------------------------------------------
void launch_zero_bins(hipStream_t s, const RegionDev& r);
// init_density1's layout-inside + centring: cur_slp = prev_slp = cur = density centre; scatter.
void launch_center_init(hipStream_t s, const RegionDev& r, const CellsDev& c, float2* cur_slp, float2* prev_slp, float2* cur,
                        unsigned long long* global_adds);
// (global_adds, when not null, counts the 64-bit atomic adds the scatter sent to the bins in HBM)
// centre the region's cells at `xy` and scatter them into the bins.
void launch_center_scatter(hipStream_t s, const RegionDev& r, const CellsDev& c, const float2* xy, unsigned long long* global_adds);
// nesterovUpdateCoordinates + centring: next, next_slp, the density boxes and the centres.
// (zero, when not null, is cleared first: the trial's HPWL counter)
void launch_update(hipStream_t s, const RegionDev& r, const CellsDev& c, const float2* cur_slp, const float2* cur_sum, const float2* cur,
                   float2* next, float2* next_slp, float step, float coeff, unsigned long long* zero);
// the scatter of the density boxes launch_update left into the bins.
void launch_density_scatter(hipStream_t s, const RegionDev& r, const CellsDev& c, unsigned long long* global_adds);
// per bin: density, overflow partials, the accumulators cleared (unless `clear` is false).
void launch_bin_density(hipStream_t s, const RegionDev& r, double* partials, bool clear = true);
// b[i] = a[i], i < n <= 64 (device to device, one small kernel: no copy node in a captured graph)
void launch_copy_f64(hipStream_t s, const double* a, double* b, int n);
// the Poisson solve of r.density into r.field (and r.phi when want_phi); nothing while *skip (when
// given) is set.
void launch_solve(hipStream_t s, const RegionDev& r, const SolveBufs& sb, bool want_phi, const int* skip = nullptr);
// launch_bin_density (clearing, partials into `partials`) then launch_solve without phi, the bins
// pass fused into the solve's first row pass where the grid allows: the same densities, partials
// and field bit for bit.
void launch_bins_solve(hipStream_t s, const RegionDev& r, const SolveBufs& sb, double* partials, const int* skip = nullptr);
// WA forward over every net, the HPWL, and each pin's weighted WA gradient into n.pgrad (at its
// cell-pin index); the net records too when write_rec.  The high-fanout nets go to s_big (which
// may be s; the caller orders s_big after the zeroed HPWL counter and joins it).
void launch_wa_forward(hipStream_t s, hipStream_t s_big, const NetsDev& n, const CellsDev& c, float coef_x, float coef_y,
                       unsigned long long* hpwl, bool write_rec);
// per cell of a region: WA gradient (its pins' n.pgrad terms summed in pin order), density
// gradient, preconditioned sum; partial sums of |wl|,
// |den| and (when cmp_slp != nullptr) the squared distances to cmp_slp/cmp_sum.
void launch_gradient(hipStream_t s, const RegionDev& r, const CellsDev& c, const NetsDev& n, float coef_x, float coef_y,
                     float penalty, float2* wl, float2* den, float2* sum, const float2* a_slp, const float2* b_slp,
                     const float2* a_sum, double* partials, int nan_inject);
// finish: sums of the block partials into *out (fixed order).
// the FP64 partials finished into *out (and *hpwl copied); with host_out (pinned host memory) the
// same values also written there, and with host_tail the HPWL counter and the global-add count of
// dev_tail (the Sums record that holds them) copied to host_tail; with host_seq, after all of
// that is visible to the host, *dev_seq is incremented and its new value written to *host_seq
// With skip, nothing while *skip is set -- except, with ctl.g, the device-driven trial's control,
// which then runs after the sums (k_finish of the trial's last region): it stages the control
// record and the sums in LDS, takes the loop's decisions after the trial (the host's expressions of
// nesterov.cpp), writes the record back, clears the HPWL counter for the next trial (unless it
// halted) and signals the host (host_seq[1] = the halt event, then host_seq[0] = the sequence
// number, a system-scope release store).
void launch_finish(hipStream_t s, const double* partials, int nparts_bins, int nparts_cells, Sums* out, const unsigned long long* hpwl,
                   Sums* host_out = nullptr, const Sums* dev_tail = nullptr, Sums* host_tail = nullptr,
                   unsigned long long* dev_seq = nullptr, unsigned long long* host_seq = nullptr, const int* skip = nullptr,
                   const CtlArgs& ctl = CtlArgs{});
// squared distance between two coordinate sets (prev/cur at init).
void launch_distance(hipStream_t s, int n, const float2* a, const float2* b, const float2* c, const float2* d, double* partials);
// initial prev_slp: cur_slp - c * cur_sum, kept inside.
void launch_initial_prev(hipStream_t s, const RegionDev& r, const CellsDev& c, const float2* cur_slp, const float2* cur_sum,
                         float coef, float2* prev_slp);
// lock the instance cells of a converged region; rotate locked cells' slots.
void launch_lock_fix(hipStream_t s, const RegionDev& r, const CellsDev& c, float2* const* next_vec, const float2* const* cur_vec, int nvec);
void launch_hpwl(hipStream_t s, const NetsDev& n, const CellsDev& c, unsigned long long* hpwl);
// The differentiable objective over region r's cells at continuous centres xy (device order):
// density: rho into r.density from the continuous boxes, then the solve with phi;
// value: out2[0] = sum_e w_e WA_e (exact exponential), out2[1] = sum_b rho_b phi_b (FP64);
// gradient: per cell d/dxy of WA + lambda * 1/2 sum rho phi (needs the density step first).
void launch_objective_density(hipStream_t s, const RegionDev& r, const SolveBufs& sb, const CellsDev& c, const float2* xy);
void launch_objective_value(hipStream_t s, const RegionDev& r, const NetsDev& n, const float2* xy, float coef, double* partials,
                            double* out2);
void launch_objective_gradient(hipStream_t s, const RegionDev& r, const CellsDev& c, const NetsDev& n, const float2* xy, float coef,
                               float lambda, double* partials, float2* grad);

// ---- spatial re-sort of the cells (device order only; results do not depend on it) ---------
constexpr int kMaxRegions = 32;
constexpr int kResortShift = 2;  // the sort key's Morton code runs over 4 x 4-bin squares
struct RegionTable {
    int32_t n;
    int32_t begin[kMaxRegions + 1];
    int32_t lx[kMaxRegions], ly[kMaxRegions], cnt_x[kMaxRegions], cnt_y[kMaxRegions];
    double size_x[kMaxRegions], size_y[kMaxRegions];
    int32_t morton_bits;  // bits of the Morton part of a key (a parked-cell bit, then the region index above them)
    int32_t end_bit;      // key bits the radix sort looks at
};
struct ResortBufs {
    uint32_t* keys_in;
    uint32_t* keys_out;
    int32_t* idx_in;
    int32_t* perm;   // new -> old
    int32_t* inv;    // old -> new
    int32_t* cnt;    // pins per new cell, then exclusive-scanned into the new pin starts (N + 1)
    void* temp;
    size_t temp_bytes;
};
size_t resort_temp_bytes(int n_cells);
// keys (region, Morton code of the current bin) and the stable sort: perm and inv
void launch_resort_sort(hipStream_t s, int n_cells, const int2* ctr, const uint8_t* flags, const RegionTable& t, ResortBufs& b);
// every per-cell array gathered into `out` (pin CSR rebuilt, the net pins' cell-pin indices
// `cpi` renumbered), old_of composed
void launch_resort_cells(hipStream_t s, int n_cells, const CellsDev& in, const CellsDev& out, ResortBufs& b, int32_t* cpi);
// one region's per-cell float pairs gathered: out[k] = in[perm[begin + k] - begin]
void launch_gather_f2(hipStream_t s, int n, int begin, const int32_t* perm, const float2* in, float2* out);
// up to kGatherMax such vectors of one region in one launch
constexpr int kGatherMax = 9;
struct GatherSet {
    const float2* in[kGatherMax];
    float2* out[kGatherMax];
    int32_t count;
};
void launch_gather_f2_set(hipStream_t s, int n, int begin, const int32_t* perm, const GatherSet& g);
// net pins' cells renamed: pcell[k] = inv[pcell[k]]
void launch_remap_pcell(hipStream_t s, int n_pins, int32_t* pcell, const int32_t* inv);

// ---- net re-sort (after a cell re-sort): the small nets ordered again by (window of 4096 cells of
// their first cell, degree), stably, as at construction, so that a WA tile's nets gather the
// centres of nearby cells; the high-fanout nets keep their places at the end
struct NetResortBufs {
    unsigned long long* keys_in;
    unsigned long long* keys_out;
    int32_t* idx_in;
    int32_t* perm;     // new -> old (small nets)
    int32_t* deg;      // degrees in the new order, scanned into the new starts
    void* temp;
    size_t temp_bytes;
    int32_t end_bit;   // key bits the sort looks at
    // the alternate net arrays the permutation writes (swapped with NetsDev's after)
    int32_t* start;
    int32_t* pcell;
    int2* poff;
    int2* pxy0;
    int32_t* cpi;
    float* weight;
    int32_t* orig;     // the problem's index of each net (device order)
};
size_t net_resort_temp_bytes(int num_small);
// writes the permuted arrays into b (start, pcell, poff, pxy0, cpi, weight, orig) and the cells'
// cell-pin -> net-pin map; the caller swaps b's arrays with the NetsDev's and then calls
// launch_tile_bounds
void launch_net_resort(hipStream_t s, const NetsDev& n, int num_pins, int num_cells, const int32_t* orig, CellsDev& c, NetResortBufs& b);
// tile[t] = the first small net whose first pin is at or past t kTileSpan, t <= num_tiles
void launch_tile_bounds(hipStream_t s, const NetsDev& n, int32_t* tile);

void launch_set_u64(hipStream_t s, unsigned long long* p, unsigned long long v);

// ---- a back-tracking trial as a HIP graph: the kernel nodes whose arguments change between
// trials (the update, the density scatter and the gradient per region, the two WA kernels),
// found by kernel and region in a captured graph and given new arguments before each launch ---
enum TrialKind { TK_SCATTER = 0, TK_WA_SMALL = 1, TK_WA_BIG = 2, TK_GRADIENT = 3, TK_UPDATE = 4 };
struct TrialNode {
    hipGraphNode_t node;
    hipKernelNodeParams p;  // as captured (function, grid, block, LDS)
    int kind;
    int cell_begin;         // the node's region (update, scatter, gradient), by its first cell
};
int find_trial_nodes(hipGraph_t g, TrialNode* out, int cap);  // count, or -1
hipError_t set_update_node(hipGraphExec_t e, const TrialNode& n, const RegionDev& r, const CellsDev& c, const float2* cur_slp,
                           const float2* cur_sum, const float2* cur, float2* next, float2* next_slp, float step, float coeff,
                           unsigned long long* zero);
hipError_t set_scatter_node(hipGraphExec_t e, const TrialNode& n, const RegionDev& r, const CellsDev& c,
                            unsigned long long* global_adds);
// ---- trials recorded and launched grouped: one launch per kernel variant ----------------------
// While a recorder is set on the calling thread, the trial's launchers (the update, the density
// scatter, the solve's row passes with the bins pass fused, the WA kernels, the gradient, the
// finishing reduction) record their kernel variants, grids and arguments instead of launching; a
// launch with no grouped form (any other launcher, or a variant without one) marks the record
// failed and launches nothing.  launch_grouped runs the records of one or more recorders (one
// trial of each of several placements, each with one or more regions) as one launch per kernel
// variant and step: every region's chain (update, scatter, row passes, gradient, finishing
// reduction) in its order, the WA kernels after every update, the finishing reductions in each
// recorder's order; a grouped launch's blocks take each record's arguments in turn (record g:
// its own grid's worth of blocks), after one copy of the arguments to the device.
constexpr int kRecArgBytes = 768;
struct LaunchRec {
    int kind, grid, block;
    unsigned lds;
    int bytes;
    alignas(16) unsigned char args[kRecArgBytes];
};
struct TrialRecorder {
    std::vector<LaunchRec> recs;
    bool failed = false;
};
void set_recorder(TrialRecorder* r);
size_t grouped_arg_bytes(const TrialRecorder* const* recs, int n);  // the staging bytes launch_grouped needs
// recs[0 .. n): recorders none of which failed; host_stage (pinned) and dev_stage hold
// grouped_arg_bytes(recs, n) bytes; the host stage must not change until the copy is done.
// With a side stream (and two events), the WA kernels run there beside the scatter, bins and
// solve, from the updates on, and the gradients wait for them (as in the trial graph).
hipError_t launch_grouped(hipStream_t s, const TrialRecorder* const* recs, int n, unsigned char* host_stage, unsigned char* dev_stage,
                          hipStream_t side = nullptr, hipEvent_t ev_fork = nullptr, hipEvent_t ev_join = nullptr);

hipError_t set_wa_node(hipGraphExec_t e, const TrialNode& n, const NetsDev& nd, const CellsDev& c, float cx, float cy,
                       unsigned long long* hpwl);
hipError_t set_gradient_node(hipGraphExec_t e, const TrialNode& n, const RegionDev& r, const CellsDev& c, const NetsDev& nd,
                             float cx, float cy, float penalty, float2* wl, float2* den, float2* sum, const float2* a_slp,
                             const float2* b_slp, const float2* a_sum, double* partials, int nan_inject);

int block_partials_bins(const RegionDev& r);
int block_partials_cells(const RegionDev& r);

// ---- routability passes (route.hip) ------------------------------------------------------------
// A cell's flag while routability has taken it out of its region (a removed filler): the re-sort
// puts such cells after the region's live cells, which the kernels alone cover.
constexpr uint8_t kCellParked = 8u;
// The per-cell slots a removed filler keeps (the re-sort's live set): saved to and restored from a
// region's store, stride cells per slot.
struct ParkSlots {
    float2* v[kGatherMax];
    int32_t count;
};
// store[s * stride + idx[i]] = v[s][pos[i]] (save) or the reverse (restore), i < k
hipError_t launch_park_save(hipStream_t s, int k, const int32_t* pos, const int32_t* idx, const ParkSlots& v, float2* store, int stride);
hipError_t launch_park_restore(hipStream_t s, int k, const int32_t* pos, const int32_t* idx, const ParkSlots& v, const float2* store,
                               int stride);
// flags[pos[i]] = (flags[pos[i]] & ~clear) | set, i < k
hipError_t launch_flag_list(hipStream_t s, int k, const int32_t* pos, uint8_t* flags, uint8_t set, uint8_t clear);
// updateBinsNonPlaceArea at target density td from its target-independent halves (nesterov.rs:
// 1956-2018): per bin non_place = min((raw * td) as i64, (union * td * 1.10) as i64) (the cap where
// a fixed instance touches the bin, union >= 0; 0 where raw is 0), non_place_unscaled =
// (union * td) as i64; written as int64 and as the bins pass's float pairs.
hipError_t launch_non_place(hipStream_t s, int m, const long long* raw, const long long* uni, float td, long long* non_place,
                            float2* non_place_f);
// getNetDemandInTiles (route_base.rs:200-218): each net's demand summed in float over the tiles of
// its box (x outer, y inner) whose inflation ratio (y-major, the last launch_rudy_tiles) is above 1,
// written at the net's problem index (net_orig) into out.  The congested tiles are first listed per
// column (rows ascending), so a net visits only those of its box, in the same order.
hipError_t launch_net_congested_demand(hipStream_t s, const NetsDev& n, const int32_t* net_orig, const RudyGrid& g, const float* ratio,
                                       float* out, RudyBufs& b);

}  // namespace gplk
